# SeisWEMAMD.jl -- Julia bindings of libseiswem.so (MI355X shot-profile WEM) for Seismic.jl users.
#
# Two ways in:
#  1. unchanged: Seismic.jl's ShotProfileWEM (src/Imaging/ShotProfileWEM.jl:334-349) calls
#     ccall((:main,"shotwem"), Int32, (Int32, Ptr{Ptr{UInt8}}), length(argv), argv); put the
#     directory holding this build's shotwem.so first on LD_LIBRARY_PATH and that call lands in
#     swem_shot_main with the same argv keys and .seis files.
#  2. in memory, without files: the functions below bind wem() (src/Imaging/c/wem.c:3-11) and
#     the multi-shot form of ShotProfileWEM with nangx = nangy = 1.
#
# Arrays follow the reference's alloc2float layouts seen from Julia (column-major): d is
# (nt, nm), m and vel are (nz, nm), trace ix = imx*nmy + imy.
module SeisWEMAMD

const libseiswem = "libseiswem"

const SWEM_PSPI_COMPAT = Cint(0)
const SWEM_PSPI_LINEAR = Cint(1)
const SWEM_PREC_KZ64 = Cint(0)
const SWEM_PREC_FP32 = Cint(1)

struct Geometry
    nt::Cint; ot::Cfloat; dt::Cfloat
    nmx::Cint; omx::Cfloat; dmx::Cfloat
    nmy::Cint; omy::Cfloat; dmy::Cfloat
    nz::Cint; oz::Cfloat; dz::Cfloat
    gz::Cfloat; sz::Cfloat
    nref::Cint; fmin::Cfloat; fmax::Cfloat
    padt::Cint; padx::Cint
    pspi::Cint
end

mutable struct Options
    pspi_mode::Cint
    precision::Cint
    device::Cint
    freqs_per_block::Cint
    verbose::Cint
    scratch_bytes::Csize_t
    variant::Cint
    table_bytes::Csize_t
end
Options() = Options(SWEM_PSPI_COMPAT, SWEM_PREC_KZ64, 0, 0, 0, 0, 0, 0)

function check(rc::Cint)
    if rc != 0
        msg = unsafe_string(ccall((:swem_last_error, libseiswem), Cstring, ()))
        error("seiswem error $rc: $msg")
    end
    nothing
end

"""
    wem!(d, m, wav, vel; nt, ot, dt, nmx, omx, dmx, nmy, omy, dmy, sx, sy, nz, oz, dz, gz, sz,
         nref, fmin, fmax, padt, padx, adj, pspi, verbose=false)

wem() of src/Imaging/c/wem.c:3-11 on Julia arrays: adj=true fills m (nz, nm) from d (nt, nm),
adj=false fills d from m.
"""
function wem!(d::Matrix{Float32}, m::Matrix{Float32}, wav::Vector{Float32}, vel::Matrix{Float32};
              nt, ot=0f0, dt, nmx, omx=0f0, dmx, nmy=1, omy=0f0, dmy=dmx, sx, sy=0f0, nz, oz=0f0, dz,
              gz=0f0, sz=0f0, nref=5, fmin=0f0, fmax=80f0, padt=1, padx=1, adj=true, pspi=true,
              verbose=false)
    check(ccall((:swem_wem, libseiswem), Cint,
                (Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat},
                 Cint, Cfloat, Cfloat, Cint, Cfloat, Cfloat, Cint, Cfloat, Cfloat,
                 Cfloat, Cfloat, Cint, Cfloat, Cfloat, Cfloat, Cfloat,
                 Ptr{Cfloat}, Cint, Cfloat, Cfloat, Cint, Cint, Cint, Cint, Cint),
                d, m, wav, nt, ot, dt, nmx, omx, dmx, nmy, omy, dmy, sx, sy, nz, oz, dz, gz, sz,
                vel, nref, fmin, fmax, padt, padx, adj, pspi, verbose))
    adj ? m : d
end

"""
    shots!(d, m, wav, vel, geom, sx, sy; adj=true, opt=Options())

ShotProfileWEM over several shots (src/Imaging/ShotProfileWEM.jl:39-310, nangx = nangy = 1):
adj=true stacks the migrated images of d (nt, nm, nshot) into m (nz, nm); adj=false models
every shot of d from m.
"""
function shots!(d::Array{Float32,3}, m::Matrix{Float32}, wav::Vector{Float32}, vel::Matrix{Float32},
                geom::Geometry, sx::Vector{Float32}, sy::Vector{Float32}; adj::Bool=true,
                opt::Options=Options())
    g = Ref(geom)
    check(ccall((:swem_shots, libseiswem), Cint,
                (Ref{Geometry}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Cfloat}, Cint,
                 Ptr{Cfloat}, Ptr{Cfloat}, Cint, Ref{Options}),
                g, vel, wav, sx, sy, length(sx), d, m, adj, opt))
    adj ? m : d
end

end # module
