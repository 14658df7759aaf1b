// swem_march2d.hip -- fused 2-D depth-march kernels (split-step Fourier and PSPI).
//
// One subgroup of T = N/P threads owns one (shot, frequency) pair for the whole depth march;
// F subgroups (different frequencies of the same shot) form a workgroup.  The source and
// receiver wavefields never leave registers: each depth step runs the x-FFTs through the
// subgroup's LDS buffers, applies the kz phase (built from FP64 s = w^2 p^2 - kx^2), the
// split-step correction and the taper, and the cross-correlation image of the F frequencies is
// summed in LDS in a fixed order and written once per step to a partial image.  The partial
// images are reduced over frequency groups and shots by reduce_image with compensated sums,
// so a shot's image does not depend on what else was launched with it.
//
// Reference semantics (src/Imaging/c/wem.c):
//   extrap1f adjoint   wem.c:229-250  source marched with -dz, receiver with +dz, image
//                                     m[x][iz] += factor Re(ur conj(us)), factor 2 (1 at iw=0)
//   extrap1f forward   wem.c:252-281  source stored per depth, receiver marched upward from
//                                     nz-1 with injection ur += m us before each step
//   ssop               wem.c:289-378  adj: FFT, kz phase, IFFT, split-step phase, taper;
//                                     forward branch: taper, split-step phase, FFT, kz, IFFT
//   pspiop             wem.c:380-471  FFT, per vref (phase, IFFT), per-x selection with the
//                                     real-part cut of linear_interp (wem.c:473-478)
#include "swem_fft.hpp"
#include "swem_internal.hpp"

namespace swem {

template <int N>
struct Pts {
    static constexpr int value = N >= 512 ? 8 : (N >= 256 ? 4 : 2);
};

constexpr int kModeSSF = 0, kModePspiCompat = 1, kModePspiLinear = 2;
// At most 512 threads per workgroup: up to 256 VGPRs per lane, which the PSPI step needs.
constexpr int kMaxBlock = 512;

template <int N>
__device__ __forceinline__ double kx2_d(int k, double dkx)
{
    double kx = (double)(k <= N / 2 ? k : k - N) * dkx;
    return kx * kx;
}

// Reference float wavenumber (wem.c:336-337).
template <int N>
__device__ __forceinline__ float kx2_f(int k, float dkx)
{
    float kx = k <= N / 2 ? dkx * (float)k : -(dkx * (float)N - dkx * (float)k);
    return kx * kx;
}

// exp(i sqrt(s) dz) for s >= 0, exp(-0.2 sqrt(-s) |dz|) otherwise (wem.c:342-343), with s
// computed in FP64 and the phase reduced to [-pi, pi] in FP64 before the FP32 sincospi.
__device__ __forceinline__ float2 kz_phase64(double s, double dz_pi, float evan)
{
    double sq = sqrt(fabs(s));
    if (s >= 0.0) {
        double a = sq * dz_pi;
        a -= 2.0 * rint(0.5 * a);
        float sn, cs;
        sincospif((float)a, &sn, &cs);
        return make_float2(cs, sn);
    }
    return make_float2(__expf(evan * (float)sq), 0.f);
}

// Naive FP32: s, sqrtf and the phase in float, as wem.c:341-343 writes them.
__device__ __forceinline__ float2 kz_phase32(float s, float dz, float evan)
{
    if (s >= 0.f) {
        float sn, cs;
        sincosf(sqrtf(s) * dz, &sn, &cs);
        return make_float2(cs, sn);
    }
    return make_float2(expf(evan * sqrtf(fabsf(s))), 0.f);
}

// Phase of wavenumber k for p2 (= po^2 or 1/vref^2), scaled by 1/nk; +dz orientation.
template <int N, int PREC>
__device__ __forceinline__ float2 kphase(const March2DArgs& a, int k, double w2p2, float w2p2f)
{
    float2 L;
    if constexpr (PREC == SWEM_PREC_KZ64)
        L = kz_phase64(w2p2 - kx2_d<N>(k, a.dkx), a.dz_pi, a.evan);
    else
        L = kz_phase32(w2p2f - kx2_f<N>(k, a.dkxf), a.dz, a.evan);
    return cscale(L, a.inv_nk);
}

template <bool CONJ>
__device__ __forceinline__ float2 cmul_c(float2 x, float2 L)
{
    return CONJ ? cmulc(x, L) : cmul(x, L);
}

struct StepCtx {
    int t, iz;
    double w;     // angular frequency (FP64)
    float wf;     // angular frequency used for the split-step phase
};

// ssop adjoint branch (wem.c:313-320, 334-359) on NW wavefields; CONJ_a = step of -dz.
template <int N, int P, int PREC, int NW, bool C0, bool C1>
__device__ __forceinline__ void ssf_adj_step(float2 (&v)[NW][P], float2* lds, const March2DArgs& a, const StepCtx& c,
                                             const float (&tap)[P])
{
    constexpr int T = N / P;
    fft<N, P, NW, -1>(v, lds, a.tw, c.t);
    const double w2p2 = c.w * c.w * a.p2[c.iz];
    const float w2p2f = (c.wf * c.wf) * a.p2f[c.iz];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const float2 L = kphase<N, PREC>(a, c.t + q * T, w2p2, w2p2f);
        v[0][q] = cmul_c<C0>(v[0][q], L);
        if constexpr (NW > 1) v[1][q] = cmul_c<C1>(v[1][q], L);
    }
    fft<N, P, NW, +1>(v, lds, a.tw, c.t);
    const float* pdrow = a.pdpi + (size_t)c.iz * a.nmx;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = c.t + q * T;
        if (x < a.nmx) {
            float sn, cs;
            sincospif(c.wf * pdrow[x], &sn, &cs);
            const float2 S = make_float2(cs, sn);
            v[0][q] = cscale(cmul_c<C0>(v[0][q], S), tap[q]);
            if constexpr (NW > 1) v[1][q] = cscale(cmul_c<C1>(v[1][q], S), tap[q]);
        } else {
            v[0][q] = make_float2(0.f, 0.f);
            if constexpr (NW > 1) v[1][q] = make_float2(0.f, 0.f);
        }
    }
}

// ssop forward branch with dz < 0 (wem.c:321-332, 334-348, 362-369): taper, split-step
// phase, FFT, kz phase, IFFT.
template <int N, int P, int PREC>
__device__ __forceinline__ void ssf_fwd_step(float2 (&v)[1][P], float2* lds, const March2DArgs& a, const StepCtx& c,
                                             const float (&tap)[P])
{
    constexpr int T = N / P;
    const float* pdrow = a.pdpi + (size_t)c.iz * a.nmx;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = c.t + q * T;
        if (x < a.nmx) {
            float sn, cs;
            sincospif(c.wf * pdrow[x], &sn, &cs);
            v[0][q] = cmulc(cscale(v[0][q], tap[q]), make_float2(cs, sn));
        } else
            v[0][q] = make_float2(0.f, 0.f);
    }
    fft<N, P, 1, -1>(v, lds, a.tw, c.t);
    const double w2p2 = c.w * c.w * a.p2[c.iz];
    const float w2p2f = (c.wf * c.wf) * a.p2f[c.iz];
#pragma unroll
    for (int q = 0; q < P; ++q) v[0][q] = cmulc(v[0][q], kphase<N, PREC>(a, c.t + q * T, w2p2, w2p2f));
    fft<N, P, 1, +1>(v, lds, a.tw, c.t);
#pragma unroll
    for (int q = 0; q < P; ++q)
        if (c.t + q * T >= a.nmx) v[0][q] = make_float2(0.f, 0.f);
}

// Weight of reference velocity ir at a point with selection byte s1 (iref1 | cut<<7),
// iref2 s2 and interpolation weight w2.
template <int MODE>
__device__ __forceinline__ float pspi_weight(int ir, uint8_t s1, uint8_t s2, float w2)
{
    const int i1 = s1 & 0x7f;
    const bool cut = (s1 >> 7) != 0;
    if constexpr (MODE == kModePspiCompat) {
        return i1 == ir ? 1.f : 0.f;
    } else {
        if (!cut) return i1 == ir ? 1.f : 0.f;
        float wt = 0.f;
        if (i1 == ir) wt += 1.f - w2;
        if ((int)s2 == ir) wt += w2;
        return wt;
    }
}

// pspiop with the reference structure (wem.c:409-460): FFT, for each vref phase + IFFT, then
// per-x selection.  Compat: u = Re(dref[iref1]) when cut else dref[iref1]; linear: weights.
template <int N, int P, int PREC, int NW, bool C0, bool C1, int MODE>
__device__ __forceinline__ void pspi_select_step(float2 (&v)[NW][P], float2* lds, const March2DArgs& a,
                                                 const StepCtx& c)
{
    constexpr int T = N / P;
    fft<N, P, NW, -1>(v, lds, a.tw, c.t);
    float2 U[NW][P];
    uint8_t s1[P], s2[P];
    float w2[P];
    const size_t row = (size_t)c.iz * a.nmx;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = c.t + q * T;
        const bool in = x < a.nmx;
        s1[q] = in ? a.sel[row + x] : (uint8_t)0x7f;
        if constexpr (MODE == kModePspiLinear) {
            s2[q] = in ? a.sel2[row + x] : (uint8_t)0x7f;
            w2[q] = in ? a.w2[row + x] : 0.f;
        } else {
            s2[q] = 0;
            w2[q] = 0.f;
        }
#pragma unroll
        for (int b = 0; b < NW; ++b) U[b][q] = make_float2(0.f, 0.f);
    }
    const uint32_t used = a.used[c.iz];
    for (int ir = 0; ir < a.nref; ++ir) {
        if (!((used >> ir) & 1u)) continue;
        const double w2p2 = c.w * c.w * a.p2[(size_t)c.iz * a.nref + ir];
        const float w2p2f = (c.wf * c.wf) * a.p2f[(size_t)c.iz * a.nref + ir];
        float2 Bv[NW][P];
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const float2 L = kphase<N, PREC>(a, c.t + q * T, w2p2, w2p2f);
            Bv[0][q] = cmul_c<C0>(v[0][q], L);
            if constexpr (NW > 1) Bv[1][q] = cmul_c<C1>(v[1][q], L);
        }
        fft<N, P, NW, +1>(Bv, lds, a.tw, c.t);
#pragma unroll
        for (int q = 0; q < P; ++q) {
            if constexpr (MODE == kModePspiCompat) {
                if ((s1[q] & 0x7f) == ir) {
                    const bool cut = (s1[q] >> 7) != 0;
#pragma unroll
                    for (int b = 0; b < NW; ++b) U[b][q] = cut ? make_float2(Bv[b][q].x, 0.f) : Bv[b][q];
                }
            } else {
                const float wt = pspi_weight<MODE>(ir, s1[q], s2[q], w2[q]);
#pragma unroll
                for (int b = 0; b < NW; ++b) U[b][q] = cadd(U[b][q], cscale(Bv[b][q], wt));
            }
        }
    }
#pragma unroll
    for (int b = 0; b < NW; ++b)
#pragma unroll
        for (int q = 0; q < P; ++q) v[b][q] = U[b][q];
}

// Exact adjoint of the linear select-after step, for the forward receiver (dz < 0):
// u <- IFFT( sum_ir conj(L_ir) FFT(W_ir u) ).
template <int N, int P, int PREC>
__device__ __forceinline__ void pspi_weight_step(float2 (&v)[1][P], float2* lds, const March2DArgs& a,
                                                 const StepCtx& c)
{
    constexpr int T = N / P;
    float2 acc[1][P];
    uint8_t s1[P], s2[P];
    float w2[P];
    const size_t row = (size_t)c.iz * a.nmx;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = c.t + q * T;
        const bool in = x < a.nmx;
        s1[q] = in ? a.sel[row + x] : (uint8_t)0x7f;
        s2[q] = in ? a.sel2[row + x] : (uint8_t)0x7f;
        w2[q] = in ? a.w2[row + x] : 0.f;
        acc[0][q] = make_float2(0.f, 0.f);
    }
    const uint32_t used = a.used[c.iz];
    for (int ir = 0; ir < a.nref; ++ir) {
        if (!((used >> ir) & 1u)) continue;
        float2 Bv[1][P];
#pragma unroll
        for (int q = 0; q < P; ++q) Bv[0][q] = cscale(v[0][q], pspi_weight<kModePspiLinear>(ir, s1[q], s2[q], w2[q]));
        fft<N, P, 1, -1>(Bv, lds, a.tw, c.t);
        const double w2p2 = c.w * c.w * a.p2[(size_t)c.iz * a.nref + ir];
        const float w2p2f = (c.wf * c.wf) * a.p2f[(size_t)c.iz * a.nref + ir];
#pragma unroll
        for (int q = 0; q < P; ++q) acc[0][q] = cadd(acc[0][q], cmulc(Bv[0][q], kphase<N, PREC>(a, c.t + q * T, w2p2, w2p2f)));
    }
    fft<N, P, 1, +1>(acc, lds, a.tw, c.t);
#pragma unroll
    for (int q = 0; q < P; ++q) v[0][q] = (c.t + q * T < a.nmx) ? acc[0][q] : make_float2(0.f, 0.f);
}

// One step of the extrapolator used for the source / adjoint receiver (select-after for PSPI).
template <int N, int P, int PREC, int MODE, int NW, bool C0, bool C1>
__device__ __forceinline__ void adj_step(float2 (&v)[NW][P], float2* lds, const March2DArgs& a, const StepCtx& c,
                                         const float (&tap)[P])
{
    if constexpr (MODE == kModeSSF)
        ssf_adj_step<N, P, PREC, NW, C0, C1>(v, lds, a, c, tap);
    else
        pspi_select_step<N, P, PREC, NW, C0, C1, MODE>(v, lds, a, c);
}

template <int N, int P, int MODE>
__device__ __forceinline__ void load_taper(float (&tap)[P], const March2DArgs& a, int t)
{
    constexpr int T = N / P;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = t + q * T;
        tap[q] = (MODE == kModeSSF && x < a.nmx) ? a.taper[x] : 1.f;
    }
}

// Migration (adjoint): grid (ngroups, nshot), block F*T.
template <int N, int MODE, int PREC>
__global__ __launch_bounds__(kMaxBlock) void march2d_adj(March2DArgs a)
{
    constexpr int P = Pts<N>::value;
    constexpr int T = N / P;
    constexpr int L = LdsLen<N>::value;
    extern __shared__ float2 smem[];
    const int F = blockDim.x / T;
    const int tid = threadIdx.x, f = tid / T, t = tid - f * T;
    const int g = blockIdx.x, shot = blockIdx.y;
    const int fl = g * F + f;
    const bool active = fl < a.nf;
    const int iw = a.ifmin + fl;
    StepCtx c;
    c.t = t;
    c.w = (double)iw * a.dw;
    c.wf = PREC == SWEM_PREC_KZ64 ? (float)c.w : (float)iw * a.dwf;
    const float factor = iw == 0 ? 1.f : 2.f;
    float2* lds = smem + (size_t)f * 2 * L;

    float2 v[2][P];
    float tap[P];
    load_taper<N, P, MODE>(tap, a, t);
    const int igx = a.igx[shot];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = t + q * T;
        v[0][q] = (active && x == igx) ? a.spec_s[fl] : make_float2(0.f, 0.f);
        v[1][q] = (active && x < a.nmx) ? a.spec_r[((size_t)shot * a.nmx + x) * a.nf + fl] : make_float2(0.f, 0.f);
    }
    float* part = a.partial + ((size_t)shot * a.ngroups + g) * (size_t)a.nz * a.nmx;
    float* red = reinterpret_cast<float*>(smem);

    for (int iz = 0; iz < a.nz; ++iz) {
        c.iz = iz;
        const float z = a.oz + a.dz * (float)iz;
        const bool ds = z >= a.sz, dr = z >= a.gz;
        if (ds && dr) {
            adj_step<N, P, PREC, MODE, 2, true, false>(v, lds, a, c, tap);
        } else if (ds) {
            float2 u[1][P];
#pragma unroll
            for (int q = 0; q < P; ++q) u[0][q] = v[0][q];
            adj_step<N, P, PREC, MODE, 1, true, true>(u, lds, a, c, tap);
#pragma unroll
            for (int q = 0; q < P; ++q) v[0][q] = u[0][q];
        } else if (dr) {
            float2 u[1][P];
#pragma unroll
            for (int q = 0; q < P; ++q) u[0][q] = v[1][q];
            adj_step<N, P, PREC, MODE, 1, false, false>(u, lds, a, c, tap);
#pragma unroll
            for (int q = 0; q < P; ++q) v[1][q] = u[0][q];
        }
        // imaging condition (wem.c:243-247), summed over the F frequencies in fixed order
        __syncthreads();
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int x = t + q * T;
            if (x < a.nmx)
                red[f * N + x] = dr ? factor * (v[1][q].x * v[0][q].x + v[1][q].y * v[0][q].y) : 0.f;
        }
        __syncthreads();
        for (int x = tid; x < a.nmx; x += blockDim.x) {
            float s = red[x];
            for (int ff = 1; ff < F; ++ff) s += red[ff * N + x];
            part[(size_t)iz * a.nmx + x] = s;
        }
        __syncthreads();
    }
}

// Modelling (forward): source march stored to smig, then receiver march upward with injection.
template <int N, int MODE, int PREC>
__global__ __launch_bounds__(kMaxBlock) void march2d_fwd(March2DArgs a)
{
    constexpr int P = Pts<N>::value;
    constexpr int T = N / P;
    constexpr int L = LdsLen<N>::value;
    extern __shared__ float2 smem[];
    const int F = blockDim.x / T;
    const int tid = threadIdx.x, f = tid / T, t = tid - f * T;
    const int g = blockIdx.x, shot = blockIdx.y;
    const int fl = g * F + f;
    const bool active = fl < a.nf;
    const int iw = a.ifmin + fl;
    StepCtx c;
    c.t = t;
    c.w = (double)iw * a.dw;
    c.wf = PREC == SWEM_PREC_KZ64 ? (float)c.w : (float)iw * a.dwf;
    float2* lds = smem + (size_t)f * 2 * L;
    float tap[P];
    load_taper<N, P, MODE>(tap, a, t);
    const int igx = a.igx[shot];
    float2* sm = a.smig + ((size_t)shot * a.nf + (active ? fl : 0)) * (size_t)a.nz * a.nmx;

    float2 u[1][P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = t + q * T;
        u[0][q] = (active && x == igx) ? a.spec_s[fl] : make_float2(0.f, 0.f);
    }
    // source march, stored per depth (wem.c:254-265)
    for (int iz = 0; iz < a.nz; ++iz) {
        c.iz = iz;
        const float z = a.oz + a.dz * (float)iz;
        const bool ds = z >= a.sz;
        if (ds) adj_step<N, P, PREC, MODE, 1, true, true>(u, lds, a, c, tap);
        if (active) {
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int x = t + q * T;
                if (x < a.nmx) sm[(size_t)iz * a.nmx + x] = ds ? u[0][q] : make_float2(0.f, 0.f);
            }
        }
    }
    // receiver march upward with injection (wem.c:266-276)
#pragma unroll
    for (int q = 0; q < P; ++q) u[0][q] = make_float2(0.f, 0.f);
    for (int iz = a.nz - 1; iz >= 0; --iz) {
        c.iz = iz;
        const float z = a.oz + a.dz * (float)iz;
        if (!(z >= a.gz)) continue;
        if (active) {
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int x = t + q * T;
                if (x < a.nmx) {
                    const float mm = a.mT[(size_t)iz * a.nmx + x];
                    const float2 s = sm[(size_t)iz * a.nmx + x];
                    u[0][q] = cadd(u[0][q], cscale(s, mm));
                }
            }
        }
        if constexpr (MODE == kModeSSF)
            ssf_fwd_step<N, P, PREC>(u, lds, a, c, tap);
        else if constexpr (MODE == kModePspiCompat)
            pspi_select_step<N, P, PREC, 1, true, true, MODE>(u, lds, a, c);
        else
            pspi_weight_step<N, P, PREC>(u, lds, a, c);
    }
    if (active) {
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int x = t + q * T;
            if (x < a.nmx) a.dgw[((size_t)shot * a.nmx + x) * a.nf + fl] = u[0][q];
        }
    }
}

// ------------------------------------------------------------------------ dispatch

int march2d_threads(int nkx)
{
    switch (nkx) {
    case 128: return 128 / Pts<128>::value;
    case 256: return 256 / Pts<256>::value;
    case 512: return 512 / Pts<512>::value;
    case 1024: return 1024 / Pts<1024>::value;
    case 2048: return 2048 / Pts<2048>::value;
    case 4096: return 4096 / Pts<4096>::value;
    default: return 0;
    }
}

bool march2d_supported(int nkx) { return march2d_threads(nkx) > 0; }

int march2d_max_block() { return kMaxBlock; }

size_t march2d_lds_bytes(int nkx, int F)
{
    return (size_t)F * 2 * (size_t)(nkx + (nkx >> 3)) * sizeof(float2);
}

template <int N, int MODE, int PREC>
static int launch_one(const March2DArgs& a, int F, int adj, hipStream_t st)
{
    constexpr int T = N / Pts<N>::value;
    dim3 grid(a.ngroups, a.nshot), block(T * F);
    size_t lds = march2d_lds_bytes(N, F);
    if (adj) {
        SWEM_HIP(hipFuncSetAttribute((const void*)march2d_adj<N, MODE, PREC>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((march2d_adj<N, MODE, PREC>), grid, block, lds, st, a);
    } else {
        SWEM_HIP(hipFuncSetAttribute((const void*)march2d_fwd<N, MODE, PREC>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((march2d_fwd<N, MODE, PREC>), grid, block, lds, st, a);
    }
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

template <int N>
static int launch_n(const March2DArgs& a, int F, int mode, int prec, int adj, hipStream_t st)
{
    if (mode == kModeSSF) {
        if (prec == SWEM_PREC_FP32) return launch_one<N, kModeSSF, SWEM_PREC_FP32>(a, F, adj, st);
        return launch_one<N, kModeSSF, SWEM_PREC_KZ64>(a, F, adj, st);
    }
    if (prec != SWEM_PREC_KZ64)
        return set_error(SWEM_EUNSUPPORTED, "naive FP32 precision is built for split-step Fourier only");
    if (mode == kModePspiCompat) return launch_one<N, kModePspiCompat, SWEM_PREC_KZ64>(a, F, adj, st);
    return launch_one<N, kModePspiLinear, SWEM_PREC_KZ64>(a, F, adj, st);
}

int launch_march2d(const March2DArgs& a, int nkx, int F, int mode, int prec, int adj, hipStream_t st)
{
    const int T = march2d_threads(nkx);
    if (T <= 0) return set_error(SWEM_EUNSUPPORTED, "2-D march: nkx=%d is not one of 128..4096 (powers of two)", nkx);
    if (F < 1 || T * F > kMaxBlock) return set_error(SWEM_EINVAL, "2-D march: %d frequencies of %d threads exceed %d", F, T, kMaxBlock);
    if (a.nmx > nkx || a.nref > 31 || a.ngroups < 1 || a.nshot < 1)
        return set_error(SWEM_EINVAL, "2-D march: inconsistent launch (nmx=%d nkx=%d nref=%d)", a.nmx, nkx, a.nref);
    if (march2d_lds_bytes(nkx, F) > 160 * 1024)
        return set_error(SWEM_EINVAL, "2-D march: LDS for F=%d at nkx=%d exceeds 160 KiB", F, nkx);
    switch (nkx) {
    case 128: return launch_n<128>(a, F, mode, prec, adj, st);
    case 256: return launch_n<256>(a, F, mode, prec, adj, st);
    case 512: return launch_n<512>(a, F, mode, prec, adj, st);
    case 1024: return launch_n<1024>(a, F, mode, prec, adj, st);
    case 2048: return launch_n<2048>(a, F, mode, prec, adj, st);
    case 4096: return launch_n<4096>(a, F, mode, prec, adj, st);
    }
    return set_error(SWEM_EUNSUPPORTED, "2-D march: nkx=%d", nkx);
}

}  // namespace swem
