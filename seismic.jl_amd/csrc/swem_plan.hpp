// swem_plan.hpp -- the plan object (device-resident tables and work buffers) shared by the C ABI
// in swem_host.cpp and the serving loop in swem_server.cpp.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "swem_internal.hpp"

namespace swem {

// Device memory source of DevBuf.  A thread that installs one (the serving loop's worker, for
// the requests it admits) gets every plan allocation from it; otherwise hipMalloc / hipFree.
struct DevAllocator {
    virtual void* alloc(size_t bytes) = 0;
    virtual void release(void* p, size_t bytes) = 0;
    virtual ~DevAllocator() = default;
};
// Host -> device copies of plan tables.  Installed by the serving loop: copies go through
// pinned staging on the request's stream.  Otherwise synchronous hipMemcpy.
struct HostStager {
    virtual int put(void* dev, const void* host, size_t bytes) = 0;
    virtual ~HostStager() = default;
};
extern thread_local DevAllocator* tl_dev_alloc;
extern thread_local HostStager* tl_stager;

int upload_bytes(void* dev, const void* host, size_t bytes);

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevAllocator* from = nullptr;
    int alloc(size_t count)
    {
        if (count <= n && p) return SWEM_OK;
        release();
        if (count == 0) return SWEM_OK;
        if (tl_dev_alloc) {
            p = static_cast<T*>(tl_dev_alloc->alloc(count * sizeof(T)));
            if (!p) return set_error(SWEM_ENOMEM, "request memory block exhausted (%zu bytes more)", count * sizeof(T));
            from = tl_dev_alloc;
        } else {
            SWEM_HIP(hipMalloc((void**)&p, count * sizeof(T)));
        }
        n = count;
        return SWEM_OK;
    }
    int upload(const T* h, size_t count)
    {
        SWEM_TRY(alloc(count));
        return upload_bytes(p, h, count * sizeof(T));
    }
    void release()
    {
        if (p) {
            if (from)
                from->release(p, n * sizeof(T));
            else
                (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
        from = nullptr;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Fourier bases of one band (time <-> frequency transforms), shared by the plans of the
// serving loop that have the same nt, dt, ntfft and band.
struct Bases {
    DevBuf<float> adj, fwd, src;
};
struct BasisCache {
    virtual const Bases* find_or_build(const swem_geometry& g, const Sampling& s, int nf) = 0;
    virtual ~BasisCache() = default;
};

}  // namespace swem

struct swem_plan {
    swem_geometry g{};
    swem_options opt{};
    swem::Sampling s{};
    int nm = 0, nf = 0, nkx = 0, T = 0, F = 1, Ffwd = 1, mode = 0;
    int narr = 0;  // x transform array length: nkx, or the Bluestein length for other nkx
    int narr_y = 0;  // 3-D y transform array length: nky, or its Bluestein length
    // frequency chunks: [c * nfc, min(nf, (c + 1) * nfc)); the phase tables hold one chunk
    int nfc = 0, nchunks = 1, resident = -1;
    bool three_d = false;
    // band rule: 0 the sampling of wem.c:79-89, 1 that of ewem.c:43-52 (elastic plans)
    int band_rule = 0;
    int F3 = 8, F3f = 8;  // frequencies per 3-D y-pass workgroup (migration, modelling)
    // 3-D step state of one launch (frequencies per launch); the autotuner may change it
    size_t state3d_bytes = size_t(128) << 20;
    int tune_nz = 0;  // autotuning only: march the first tune_nz depths (0 = all)
    // 3-D migration of one shot as a captured graph: the launches of every frequency chunk and
    // depth read the shot's values from shot3d; the graph is captured once per buffer layout
    // (graph3d_key) and replayed for every shot.  use_graphs = 0 launches them one by one.
    int use_graphs = 1;
    swem::DevBuf<swem::Shot3D> shot3d;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph3d = nullptr;
    hipGraphExec_t graph3d_exec = nullptr;
    std::vector<uintptr_t> graph3d_key;
    std::vector<float> vel;  // host copy (source of every table)
    // construction context: stream of the table builds, optional shared bases
    hipStream_t build_stream = nullptr;
    bool served = false;  // built by the serving loop: no events, no device-wide waits
    swem::BasisCache* basis_cache = nullptr;
    // tables
    swem::DevBuf<float2> tw, twy, spec_s, ktab, stab;
    swem::DevBuf<float2> chirp, bchirp;  // x Bluestein: [nkx] chirp, [2][narr] kernel spectra
    swem::DevBuf<float2> chirp_y, bchirp_y;  // 3-D y Bluestein: [nky] chirp, [2][narr_y]
    swem::MixedRadix mrx{}, mry{};           // mixed-radix plans of nkx, nky (n = 0: not mixed radix)
    swem::DevBuf<float> taper, taper_y, p2f, pdf, w2, wavd;
    swem::Bases own_bases;
    const float* basis_adj = nullptr;  // [nt][2nf] (own_bases or a shared cache entry)
    const float* basis_fwd = nullptr;  // [2nf][nt]
    swem::DevBuf<double> p2, pd;
    swem::DevBuf<uint8_t> sel, sel2;
    swem::DevBuf<uint32_t> used;
    swem::DevBuf<int> cref;  // [nz][kCrefStride] uncut points of a depth (PSPI compat)
    bool packed_ok = false;  // PSPI compat migration may run the packed kernel
    // work
    swem::DevBuf<float2> spec_r, smig, dgw, W, R;  // R: 3-D PSPI per-vref slices
    swem::DevBuf<float> partial, mT;
    swem::DevBuf<int> igx;
    std::vector<int> igx_host;
    int* igx_pinned = nullptr;
    int igx_cap = 0;
    size_t budget = 0;
    // timing of the last call
    std::vector<hipEvent_t> events;
    int nev = 0;
    std::vector<int> ev_kind;  // stage of the interval that ends at event i (0 spectra, 1 march, 2 reduce)
    int last_launches = 0;
    double last_flops = 0;
    hipStream_t last_stream = nullptr;
};

namespace swem {

// Host-side set-up of a plan (validation, launch shapes, chunk sizes); no device work.
// table_cap = 0: the default cap from the free device memory.
int plan_setup(swem_plan* p, const swem_geometry* g, const float* vel, const swem_options* opt);
// Device tables of a set-up plan, enqueued on p->build_stream (synchronised unless served).
int plan_build(swem_plan* p, const float* wav);
// Device bytes a set-up plan allocates for its tables plus one migration of nshot shots
// (records and image excluded); an upper bound including allocation padding.
size_t plan_migrate_bytes(const swem_plan* p, int nshot);
// Host bytes the plan stages for its uploads (served plans).
size_t plan_upload_bytes(const swem_plan* p, int nshot);
int plan_migrate(swem_plan* p, const float* d_dev, const float* sx, const float* sy, int nshot, float* m_dev,
                 int accumulate, int per_shot, hipStream_t st);
void plan_free(swem_plan* p);

}  // namespace swem
