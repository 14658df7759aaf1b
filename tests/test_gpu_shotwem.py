"""GPU test of the file-based drop-in: shotwem.so's `main` with the reference's argv protocol
and .seis files (shotwem.c:12-140), migrating and modelling one shot, checked against the FP64
oracle and the reference's header rewrites (shotwem.c:86-98, 113-117)."""
import ctypes

import numpy as np
import pytest

import oracle as O
import seiswem
from seiswem import _lib, synth
from conftest import rel_l2
from test_capi import header_array, write_seis

pytestmark = pytest.mark.gpu


def call_shotwem(args):
    so = ctypes.CDLL(seiswem.SHOTWEM_PATH)
    so.main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    so.main.restype = ctypes.c_int
    argv = [b"0"] + [a.encode() for a in args]
    return so.main(len(argv), (ctypes.c_char_p * len(argv))(*argv))


def read_seis(name, n1, ntr):
    lib = seiswem.load()
    fh = _lib.SeisExtent()
    data = np.zeros((ntr, n1), np.float32)
    h = header_array(ntr)
    _lib.check(lib.swem_seis_read(name.encode(), data.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), h.ctypes.data,
                                  ctypes.byref(fh)))
    return data, h, fh


@pytest.mark.parametrize("pspi", [False, True])
def test_shotwem_main_migrates_and_models_like_the_oracle(gpu, tmp_path, monkeypatch, pspi):
    monkeypatch.chdir(tmp_path)
    monkeypatch.delenv("DATAPATH", raising=False)
    lib = seiswem.load()
    nx, nz, nt, dx, dz, dt, ox = 64, 40, 256, 10.0, 10.0, 0.004, 0.0
    rng = np.random.default_rng(2)
    vel = synth.blob_velocity(nx, nz, dx, dz, rng, nblobs=3, v0=2000.0)
    wav = synth.wavelet_trace(nt, dt, 15.0)
    d = rng.standard_normal((nx, nt)).astype(np.float32)
    mx = ox + dx * np.arange(nx)
    write_seis(lib, "vel.seis", vel, header_array(nx, tracenum=np.arange(1, nx + 1), n1=nz, mx=mx, imx=np.arange(nx)),
               nz, nx, o1=0.0, d1=dz, o2=ox, d2=dx)
    write_seis(lib, "wav.seis", wav[None], header_array(1, n1=nt), nt, 1, o1=0.0, d1=dt)
    write_seis(lib, "d.seis", d, header_array(nx, tracenum=np.arange(1, nx + 1), n1=nt, gx=mx, sx=200.0), nt, nx,
               o1=0.0, d1=dt, o2=ox, d2=dx)
    common = [f"pspi={'y' if pspi else 'n'}", "vel=vel.seis", "wav=wav.seis", "sx=200", "fmax=40", "nref=4",
              "padt=2", "padx=2"]
    assert call_shotwem(["adj=y", "d=d.seis", "m=m.seis"] + common) == 0
    m, hm, fh = read_seis("m.seis", nz, nx)
    assert (fh.n1, fh.n2) == (nz, nx) and abs(fh.d1 - dz) < 1e-6
    assert np.allclose(hm["mx"], mx) and np.array_equal(hm["imx"], np.arange(nx))  # shotwem.c:113-117
    p = O.WemParams(nt=nt, dt=dt, nmx=nx, dmx=dx, nz=nz, dz=dz, fmin=0.0, fmax=40.0, nref=4, padt=2, padx=2,
                    pspi=pspi, pspi_mode=0)
    m_ref = O.wem(p, vel, wav, 200.0, d=d, adj=True)
    assert rel_l2(m, m_ref) <= 1e-4
    # modelling of the image just written
    assert call_shotwem(["adj=n", "d=dm.seis", "m=m.seis"] + common) == 0
    dm, hd, fh = read_seis("dm.seis", nt, nx)
    assert (fh.n1, fh.n2) == (nt, nx) and abs(fh.d1 - dt) < 1e-7
    assert np.allclose(hd["gx"], mx) and np.allclose(hd["hx"], mx - 200.0)  # shotwem.c:86-98
    d_ref = O.wem(p, vel, wav, 200.0, m=m, adj=False)
    assert rel_l2(dm, d_ref) <= 1e-4


def _grid_files(lib, nx, nz, nt, dx, dz, dt, rng):
    vel = synth.blob_velocity(nx, nz, dx, dz, rng, nblobs=3, v0=2000.0)
    wav = synth.wavelet_trace(nt, dt, 15.0)
    mx = dx * np.arange(nx)
    write_seis(lib, "vel.seis", vel, header_array(nx, tracenum=np.arange(1, nx + 1), n1=nz, d1=dz, mx=mx,
                                                  imx=np.arange(nx)), nz, nx, o1=0.0, d1=dz, o2=0.0, d2=dx)
    write_seis(lib, "wav.seis", wav[None], header_array(1, n1=nt, d1=dt), nt, 1, o1=0.0, d1=dt)
    return vel, wav, mx


def test_shot_profile_wem_files_stacks_and_concatenates_like_the_reference(gpu, tmp_path, monkeypatch):
    """ShotProfileWEM(m, d, adj) on .seis files (ShotProfileWEM.jl:39-310): per-shot windows by
    sx/sy, images stacked at imx(gx) * ny + imy, modelled shots concatenated with their headers."""
    monkeypatch.chdir(tmp_path)
    monkeypatch.delenv("DATAPATH", raising=False)
    lib = seiswem.load()
    nx, nz, nt, dx, dz, dt = 64, 40, 256, 10.0, 10.0, 0.004
    rng = np.random.default_rng(5)
    vel, wav, mx = _grid_files(lib, nx, nz, nt, dx, dz, dt, rng)
    sx = np.array([150.0, 320.0, 480.0], np.float32)
    d = rng.standard_normal((3, nx, nt)).astype(np.float32)
    # shot 1's traces are stored with their receivers in reverse order: wem images trace ix at grid
    # point ix and the stack then places it at the receiver's imx (shotwem.c:113-117).
    gx = np.stack([mx, mx[::-1], mx])
    order = np.array([0, 2, 1])  # shots interleaved in the file, windowed by their sx header
    h = np.concatenate([header_array(nx, n1=nt, d1=dt, sx=sx[k], gx=gx[k]) for k in order])
    write_seis(lib, "d.seis", d[order].reshape(-1, nt), h, nt, 3 * nx, o1=0.0, d1=dt)
    common = dict(vel="vel.seis", wav="wav.seis", fmax=40.0, nref=4, padt=2, padx=2, pspi=True)
    m = seiswem.ShotProfileWEMFiles("m.seis", "d.seis", True, sx=sx, sy=np.zeros(3, np.float32), **common)
    p = O.WemParams(nt=nt, dt=dt, nmx=nx, dmx=dx, nz=nz, dz=dz, fmin=0.0, fmax=40.0, nref=4, padt=2, padx=2,
                    pspi=True, pspi_mode=0)
    m_ref = np.zeros((nx, nz))
    for k in range(3):
        img = O.wem(p, vel, wav, float(sx[k]), d=d[k], adj=True)
        m_ref[np.trunc(gx[k] / dx).astype(int)] += img
    assert rel_l2(m, m_ref) <= 1e-4
    mr, hm, fh = seiswem.seis.read("m.seis")
    assert np.array_equal(mr, m)
    assert (fh.n1, fh.n2, fh.n3, fh.n4, fh.n5) == (nz, nx, 1, 1, 1)
    assert [fh.label1, fh.label2, fh.label4] == [b"Depth", b"mx", b"angx"] and fh.unit4 == b"degrees"
    assert np.array_equal(hm["tracenum"], np.arange(1, nx + 1))  # SeisWrite.jl:36-39
    assert np.array_equal(hm["imx"], np.arange(nx)) and np.array_equal(hm["mx"], mx.astype(np.float32))
    assert np.all(hm["n1"] == nz) and np.all(hm["d1"] == np.float32(dz))

    # demigration of the stacked image, two shots concatenated (ShotProfileWEM.jl:203-309)
    dm = seiswem.ShotProfileWEMFiles("m.seis", "dm.seis", False, sx=sx[:2], sy=np.zeros(2, np.float32), **common)
    dr, hd, fh = seiswem.seis.read("dm.seis")
    assert dr.shape == (2 * nx, nt) and np.array_equal(dr.reshape(2, nx, nt), dm)
    assert (fh.n1, fh.n2, fh.n3, fh.n4, fh.n5) == (nt, nx, 1, 2, 2)
    assert [fh.label1, fh.label2, fh.label4, fh.label5] == [b"Time", b"gx", b"sx", b"sy"]
    assert abs(fh.o4 - sx[0]) < 1e-3 and abs(fh.d4 - (sx[1] - sx[0])) < 1e-3 and fh.d5 == 1.0
    assert np.array_equal(hd["tracenum"], np.arange(1, 2 * nx + 1))
    for k in range(2):
        hk = hd[k * nx:(k + 1) * nx]
        assert np.all(hk["sx"] == sx[k]) and np.array_equal(hk["gx"], mx.astype(np.float32))
        assert np.array_equal(hk["hx"], mx.astype(np.float32) - sx[k]) and np.all(hk["n1"] == nt)
        d_ref = O.wem(p, vel, wav, float(sx[k]), m=m, adj=False)
        assert rel_l2(dm[k], d_ref) <= 1e-4


def test_compute_angles_main_writes_the_angle_files(gpu, tmp_path, monkeypatch):
    """compute_angles.so's main (compute_angles.c:12-118) from .seis files: the angles equal the
    in-memory swem_compute_angles of the same grid, the headers are the velocity's with sx set,
    and dip_flag=y adds the dip files."""
    monkeypatch.chdir(tmp_path)
    monkeypatch.delenv("DATAPATH", raising=False)
    lib = seiswem.load()
    nx, nz, nt, dx, dz, dt = 128, 60, 256, 10.0, 10.0, 0.004
    rng = np.random.default_rng(5)
    vel = synth.blob_velocity(nx, nz, dx, dz, rng, nblobs=2, v0=2000.0)
    wav = synth.wavelet_trace(nt, dt, 20.0)
    mx = dx * np.arange(nx)
    hv = header_array(nx, tracenum=np.arange(1, nx + 1), n1=nz, o1=0.0, d1=dz, mx=mx, imx=np.arange(nx))
    write_seis(lib, "vel.seis", vel, hv, nz, nx, o1=0.0, d1=dz, d2=dx)
    write_seis(lib, "wav.seis", wav[None], header_array(1, n1=nt, o1=0.0, d1=dt), nt, 1, o1=0.0, d1=dt)
    dip = np.full((nx, nz), 2.5, np.float32)
    write_seis(lib, "dipx.seis", dip, hv, nz, nx, o1=0.0, d1=dz, d2=dx)
    write_seis(lib, "dipy.seis", dip, hv, nz, nx, o1=0.0, d1=dz, d2=dx)
    so = ctypes.CDLL(seiswem.ANGLES_PATH)
    so.main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    so.main.restype = ctypes.c_int

    def run(args):
        argv = [b"0"] + [a.encode() for a in args]
        return so.main(len(argv), (ctypes.c_char_p * len(argv))(*argv))

    common = ["vel=vel.seis", "wav=wav.seis", "sx=640", "fmin=5", "fmax=40", "dipx=dipx.seis", "dipy=dipy.seis"]
    assert run(["angx=ax.seis", "angy=ay.seis"] + common) == 0
    ax, hx, fh = read_seis("ax.seis", nz, nx)
    ay, _, _ = read_seis("ay.seis", nz, nx)
    g = seiswem.WemGeometry(nt=nt, dt=dt, nmx=nx, dmx=dx, nz=nz, dz=dz, fmin=5.0, fmax=40.0, pspi=False)
    rx, ry = seiswem.compute_angles(vel, wav, g, 640.0)
    assert np.array_equal(ax, rx) and np.array_equal(ay, ry)
    assert np.all(hx["sx"] == 640.0) and np.array_equal(hx["imx"], np.arange(nx)) and (fh.n1, fh.n2) == (nz, nx)
    assert run(["angx=bx.seis", "angy=by.seis", "dip_flag=y"] + common) == 0
    bx, _, _ = read_seis("bx.seis", nz, nx)
    by, _, _ = read_seis("by.seis", nz, nx)
    assert np.array_equal(bx, rx + np.float32(2.5)) and np.array_equal(by, ry + np.float32(2.5))


def test_shot_profile_wem_files_angle_gathers(gpu, tmp_path, monkeypatch):
    """ShotProfileWEM with nangx > 1 on files (ShotProfileWEM.jl:93-201, 240-275): angle files
    from compute_angles.so for two shots concatenated as ComputeAngles.jl writes them; the gathers
    and their headers, and the modelled shots, against ShotProfileWEMAngles on arrays."""
    monkeypatch.chdir(tmp_path)
    monkeypatch.delenv("DATAPATH", raising=False)
    lib = seiswem.load()
    nx, nz, nt, dx, dz, dt = 64, 40, 256, 10.0, 10.0, 0.004
    rng = np.random.default_rng(8)
    vel, wav, mx = _grid_files(lib, nx, nz, nt, dx, dz, dt, rng)
    sx = np.array([200.0, 440.0], np.float32)
    so = ctypes.CDLL(seiswem.ANGLES_PATH)
    so.main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    so.main.restype = ctypes.c_int
    parts = {"x": [], "y": []}
    for k, s in enumerate(sx):
        argv = [b"0"] + [a.encode() for a in (f"angx=ax{k}.seis", f"angy=ay{k}.seis", "vel=vel.seis", "wav=wav.seis",
                                              f"sx={s}", "fmax=40")]
        assert so.main(len(argv), (ctypes.c_char_p * len(argv))(*argv)) == 0
        for a in "xy":
            parts[a].append(seiswem.seis.read(f"a{a}{k}.seis"))
    for a in "xy":  # ComputeAngles.jl:89-99: shots appended to one file
        data = np.concatenate([p[0] for p in parts[a]])
        hh = np.concatenate([p[1] for p in parts[a]])
        hh["tracenum"] = np.arange(1, hh.size + 1)
        seiswem.seis.write(f"ang{a}.seis", data, hh, parts[a][0][2])
    d = rng.standard_normal((2, nx, nt)).astype(np.float32)
    h = np.concatenate([header_array(nx, n1=nt, d1=dt, sx=sx[k], gx=mx) for k in range(2)])
    write_seis(lib, "d.seis", d.reshape(-1, nt), h, nt, 2 * nx, o1=0.0, d1=dt)
    akw = dict(nangx=6, oangx=-45.0, dangx=15.0)
    common = dict(vel="vel.seis", wav="wav.seis", angx="angx.seis", angy="angy.seis", fmax=40.0, pspi=False,
                  padt=2, padx=2, **akw)
    gath = seiswem.ShotProfileWEMFiles("m.seis", "d.seis", True, sx=sx, sy=np.zeros(2, np.float32), **common)
    g = seiswem.WemGeometry(nt=nt, dt=dt, nmx=nx, dmx=dx, nz=nz, dz=dz, fmax=40.0, pspi=False)
    angx = np.stack([parts["x"][k][0] for k in range(2)])
    angy = np.stack([parts["y"][k][0] for k in range(2)])
    ref = seiswem.ShotProfileWEMAngles(d, True, vel=vel, wav=wav, geom=g, sx=sx, angx=angx, angy=angy, **akw)
    assert rel_l2(gath, ref) <= 1e-6
    mr, hm, fh = seiswem.seis.read("m.seis")
    assert (fh.n1, fh.n2, fh.n3, fh.n4, fh.n5) == (nz, nx, 1, 6, 1) and abs(fh.o4 + 45.0) < 1e-4
    assert np.array_equal(mr, gath.reshape(nx * 6, nz))
    assert np.array_equal(hm["iang"], np.tile(np.arange(6), nx)) and np.array_equal(hm["imx"], np.repeat(np.arange(nx), 6))
    assert np.allclose(hm["ang"], np.tile(-45.0 + 15.0 * np.arange(6), nx))
    dm = seiswem.ShotProfileWEMFiles("m.seis", "dm.seis", False, sx=sx, sy=np.zeros(2, np.float32), **common)
    dref = seiswem.ShotProfileWEMAngles(gath, False, vel=vel, wav=wav, geom=g, sx=sx, angx=angx, angy=angy, **akw)
    assert rel_l2(dm, dref) <= 1e-6


def test_shotewem_main_migrates_and_models_through_files(gpu, tmp_path, monkeypatch):
    """shotewem.so's main (shotewem.c:11-223): the three component files in, the three image files
    out (trid 1 / 2, mx from the data's gx), equal to swem_ewem on the same arrays; then modelling
    back to component files with the gx / hx rewrites."""
    monkeypatch.chdir(tmp_path)
    monkeypatch.delenv("DATAPATH", raising=False)
    lib = seiswem.load()
    nx, nz, nt, dx, dz, dt = 64, 30, 256, 10.0, 10.0, 0.004
    rng = np.random.default_rng(11)
    vp, wav, mx = _grid_files(lib, nx, nz, nt, dx, dz, dt, rng)
    vs = (vp / 1.8).astype(np.float32)
    write_seis(lib, "vs.seis", vs, header_array(nx, tracenum=np.arange(1, nx + 1), n1=nz, d1=dz, mx=mx,
                                                imx=np.arange(nx)), nz, nx, o1=0.0, d1=dz, o2=0.0, d2=dx)
    comps = rng.standard_normal((3, nx, nt)).astype(np.float32)
    for k, name in enumerate(("ux", "uy", "uz")):
        write_seis(lib, f"{name}.seis", comps[k], header_array(nx, n1=nt, d1=dt, sx=250.0, gx=mx), nt, nx, o1=0.0,
                   d1=dt, o2=0.0, d2=dx)
    so = ctypes.CDLL(seiswem.SHOTEWEM_PATH)
    so.main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    so.main.restype = ctypes.c_int

    def run(args):
        argv = [b"0"] + [a.encode() for a in args]
        return so.main(len(argv), (ctypes.c_char_p * len(argv))(*argv))

    files = ["ux=ux.seis", "uy=uy.seis", "uz=uz.seis", "mpp=mpp.seis", "mps1=mps1.seis", "mps2=mps2.seis",
             "vp=vel.seis", "vs=vs.seis", "wav=wav.seis", "sx=250", "fmax=30", "padt=2", "padx=2"]
    assert run(["adj=y"] + files) == 0
    ref = seiswem.ewem(comps[0], comps[1], comps[2], None, None, None, wav, nt, 0.0, dt, nx, 0.0, dx, 1, 0.0, dx,
                       250.0, 0.0, nz, 0.0, dz, 0.0, 0.0, vp, vs, 0.0, 30.0, 2, 2, True)
    for k, name in enumerate(("mpp", "mps1", "mps2")):
        img, h, fh = read_seis(f"{name}.seis", nz, nx)
        assert np.array_equal(img, ref[k]) and (fh.n1, fh.n2) == (nz, nx)
        assert np.all(h["trid"] == (1 if k == 0 else 2)) and np.array_equal(h["imx"], np.arange(nx))
    assert np.linalg.norm(ref[0]) > 0 and np.all(ref[1] == 0)
    assert run(["adj=n"] + [f.replace("ux.seis", "mx.seis").replace("uy.seis", "my.seis").replace("uz.seis", "mz.seis")
                            for f in files]) == 0
    dref = seiswem.ewem(None, None, None, ref[0], ref[1], ref[2], wav, nt, 0.0, dt, nx, 0.0, dx, 1, 0.0, dx, 250.0, 0.0,
                        nz, 0.0, dz, 0.0, 0.0, vp, vs, 0.0, 30.0, 2, 2, False)
    for k, name in enumerate(("mx", "my", "mz")):
        d_k, h, fh = read_seis(f"{name}.seis", nt, nx)
        assert np.array_equal(d_k, dref[k]) and (fh.n1, fh.n2) == (nt, nx)
        assert np.array_equal(h["gx"], mx.astype(np.float32)) and np.array_equal(h["hx"], mx.astype(np.float32) - 250.0)
