// swem_internal.hpp -- declarations shared by the host code and the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

#include "../../include/seiswem.h"

namespace swem {

// Error state (thread-local message, see swem_host.cpp).
int set_error(int code, const char* fmt, ...);
int hip_error(hipError_t e, const char* what, const char* file, int line);

#define SWEM_HIP(call)                                                        \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) return ::swem::hip_error(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define SWEM_TRY(expr)            \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != SWEM_OK) return rc_; \
    } while (0)

// Derived sampling of one wem() call, with the reference's casts (wem.c:79-89, 305-308).
struct Sampling {
    int ntfft, nw, nkx, nky, ifmin, ifmax, lmx, lmy, fop_n;
    float dkx, dky, dw;
};

int compute_sampling(const swem_geometry& g, Sampling& s);
int source_index(const swem_geometry& g, float sx, float sy, int& igx, int& igy);
int pspi_tables(const swem_geometry& g, const float* vel, float* vref, int* iref1, int* iref2, uint8_t* cut);
void taper_f32(int n, int l, float* t);
void slowness_f32(const swem_geometry& g, const float* vel, float* po, float* pd);

// Arguments of the 2-D depth-march kernels (one subgroup of N/P threads per frequency,
// F subgroups per workgroup, grid = (ngroups, nshot)).
struct March2DArgs {
    int nmx, nz, nf, nref, ngroups, ifmin, nshot;
    float oz, dz, sz, gz;
    double dw;        // 2 pi / (ntfft dt)          (FP64 frequency step)
    float dwf;        // reference float dw        (wem.c:85)
    double dkx;       // 2 pi / (nkx dmx)
    float dkxf;       // reference float dkx       (wem.c:83)
    double dz_pi;     // dz / pi
    float evan;       // -0.2 |dz|                  (wem.c:343)
    float inv_nk;     // 1 / nkx                    (the two 1/sqrt(nk) of wem.c:344,354)
    const float2* tw;       // [nkx] exp(-2 pi i m / nkx)
    const float2* spec_s;   // [nf] source spectrum / sqrt(ntfft)
    const int* igx;         // [nshot]
    const float2* spec_r;   // [nshot][nmx][nf]  receiver spectra w^2 D / sqrt(ntfft)
    const float* pdpi;      // [nz][nmx] pd * dz / pi (SSF)
    const float* taper;     // [nmx]
    const double* p2;       // SSF: po^2 [nz];  PSPI: 1/vref^2 [nz][nref]
    const float* p2f;       // float counterparts (naive FP32 precision)
    const uint8_t* sel;     // [nz][nmx] iref1 | cut << 7 (PSPI)
    const uint8_t* sel2;    // [nz][nmx] iref2 (PSPI linear)
    const float* w2;        // [nz][nmx] weight of iref2 (PSPI linear)
    const uint32_t* used;   // [nz] bit ir set when some x selects vref ir at iz
    float* partial;         // [nshot][ngroups][nz][nmx] migration partial images
    const float* mT;        // [nz][nmx] image (modelling input, depth-major)
    float2* smig;           // [nshot][nf][nz][nmx] source wavefield scratch (modelling)
    float2* dgw;            // [nshot][nmx][nf] modelled spectra (modelling output)
};

// mode: 0 = SSF, 1 = PSPI compat, 2 = PSPI linear.  prec: SWEM_PREC_*.
int launch_march2d(const March2DArgs& a, int nkx, int F, int mode, int prec, int adj, hipStream_t st);
// Threads per frequency subgroup for a given nkx (0 if unsupported).
int march2d_threads(int nkx);
bool march2d_supported(int nkx);
int march2d_max_block();
size_t march2d_lds_bytes(int nkx, int F);

// C[M x Nc] = A[M x K] * B[K x Nc], row-major with leading dimensions lda, ldb, ldc.
int launch_sgemm(int M, int Nc, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
                 hipStream_t st);

// Image reduction: out[(s)][x][iz] (+)= sum_{shot in group, g} partial[shot][g][iz][x], Kahan,
// fixed order.  per_shot = 1 writes one image per shot (no stacking).
int launch_reduce_image(const float* partial, int nshot, int ngroups, int nz, int nmx, float* out, int accumulate,
                        int per_shot, hipStream_t st);
// [nm][nz] -> [nz][nm]
int launch_transpose(const float* in, int rows, int cols, float* out, hipStream_t st);

}  // namespace swem
