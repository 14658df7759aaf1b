// swem_march3d.hip -- 3-D split-step Fourier depth march (nmy > 1).
//
// A 3-D slice (nkx x nky complex, 2 MiB at 512 x 512) does not fit in LDS, so one depth step
// is two passes over HBM per wavefield, each a read and a write of the live part of the slice:
//
//   state between steps: W[item][wf][x][ky], x < nmx only (rows x >= nmx of the padded slice are
//   zero on input and discarded on output, so they are never stored), y transformed to ky.
//
//   march3d_cols  (x-pass)  for each ky column: FFT over x (input rows x < nmx, zeros above),
//                           times the kz phase exp(+-i kz dz)/nk built from FP64
//                           s = w^2 po^2 - kx^2 - ky^2, inverse FFT, keep x < nmx.
//   march3d_rows  (y-pass)  for each row x: inverse FFT over y, keep y < nmy, then the
//                           split-step phase and taper (adjoint, wem.c:350-359), the imaging
//                           condition of the F frequencies of the workgroup summed in a fixed
//                           order, then the forward y-FFT that starts the next step.
//
// so a step of the adjoint (wem.c:313-320, 334-359) FFT2 -> L -> IFFT2 -> SS -> taper is carried
// out as  [FFTy] FFTx L IFFTx | IFFTy SS taper image FFTy ...  and the forward step (wem.c:321-
// 332, 362-369) taper -> SS -> FFT2 -> L -> IFFT2 as  IFFTy inject taper SS FFTy | FFTx L IFFTx.
// An item is one frequency of one shot; a launch covers every item of the frequency chunk.
// Transforms are wave-level (one 64-lane wave per row / column, N/64 points per lane, no
// workgroup barrier inside a transform).  An axis whose length n is not a power of two 128..1024
// is transformed through Bluestein's identity on an array of N >= 2n - 1 points (line_fft).
#include "swem_esep.hpp"
#include "swem_fft.hpp"
#include "swem_internal.hpp"

namespace swem {

namespace {

constexpr double kTwoPi = 6.283185307179586;

// exp(i 2 pi r) for r already reduced to [-1/2, 1/2] (hardware sine / cosine take revolutions).
__device__ __forceinline__ float2 cis_rev(double r)
{
    const float rf = (float)r;
    return make_float2(__builtin_amdgcn_cosf(rf), __builtin_amdgcn_sinf(rf));
}

__device__ __forceinline__ double reduce_rev(double phase)
{
    double r = phase * (1.0 / kTwoPi);
    return r - rint(r);
}

// FP64 square root of s >= 0 by the hardware reciprocal square root (5e-8 relative) and one
// Newton step: 4e-15 relative at most (measured over 1e-20 .. 1), against about 17 instructions
// of the library's correctly rounded expansion; 0 up to 1e-300 (the estimate is not defined at
// 0); a NaN stays NaN.
__device__ __forceinline__ double sqrt_nr(double s)
{
    const double r = __builtin_amdgcn_rsq(s);
    double y = s * r;
    const double h = 0.5 * r;
    y = fma(y, fma(-y, h, 0.5), y);
    return s <= 1e-300 ? 0.0 : y;
}

// kz phase from s = w^2 p^2 - kx^2 - ky^2 evaluated in FP64 (the cancellation near the
// evanescent cutoff): square root, phase and its reduction to revolutions in FP64, rounded to
// float only for the hardware sine / cosine of the reduced angle; the evanescent factor
// exp(-0.2 sqrt(-s) |dz|) from the FP64 root (wem.c:341-343).  1/nk folded in.
__device__ __forceinline__ float2 kz_phase_s64(const March3DArgs& a, double s)
{
    const double q = sqrt_nr(fabs(s));
    if (s >= 0.0) {
        const float2 c = cis_rev(reduce_rev(q * a.dzd));
        return make_float2(c.x * a.inv_nk, c.y * a.inv_nk);
    }
    return make_float2(expf((float)(-0.2 * q * fabs(a.dzd))) * a.inv_nk, 0.f);
}

// kz phase of one (kx, ky) point, +dz orientation, with 1/nk folded in (wem.c:336-344).
__device__ __forceinline__ float2 kz_phase(const March3DArgs& a, double w, double p2, float p2f, int ikx, int iky)
{
    if (a.prec == SWEM_PREC_KZ64) {
        const int kxi = ikx <= a.nkx / 2 ? ikx : a.nkx - ikx;
        const int kyi = iky <= a.nky / 2 ? iky : a.nky - iky;
        const double kx = kxi * a.dkx, ky = kyi * a.dky;
        return kz_phase_s64(a, w * w * p2 - kx * kx - ky * ky);
    }
    {
        // the reference's float expressions (wem.c:336-343), mirrored wavenumbers included, with
        // no contraction into fused multiply-adds
#pragma clang fp contract(off)
        const float wf = (float)w;
        const float kx = ikx <= a.nkx / 2 ? a.dkxf * (float)ikx : -(a.dkxf * (float)a.nkx - a.dkxf * (float)ikx);
        const float ky = iky <= a.nky / 2 ? a.dkyf * (float)iky : -(a.dkyf * (float)a.nky - a.dkyf * (float)iky);
        const float s = (wf * wf) * p2f - kx * kx - ky * ky;
        if (s >= 0.f) {
            float sn, cs;
            sincosf(sqrtf(s) * (float)a.dzd, &sn, &cs);
            return make_float2(cs * a.inv_nk, sn * a.inv_nk);
        }
        return make_float2(expf(-0.2f * sqrtf(fabsf(s)) * fabsf((float)a.dzd)) * a.inv_nk, 0.f);
    }
}

// split-step phase exp(i w pd dz) (+dz orientation, wem.c:326,353)
__device__ __forceinline__ float2 ss_phase(const March3DArgs& a, double w, size_t ixy)
{
    if (a.prec == SWEM_PREC_KZ64) return cis_rev(reduce_rev(w * a.pd[(size_t)a.iz * a.nm + ixy] * a.dzd));
    {
#pragma clang fp contract(off)
        float sn, cs;
        sincosf((float)w * a.pdf[(size_t)a.iz * a.nm + ixy] * (float)a.dzd, &sn, &cs);
        return make_float2(cs, sn);
    }
}

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float2* state(const March3DArgs& a, int item, int wf)
{
    return a.W + ((size_t)item * 2 + wf) * (size_t)a.nmx * a.nky;
}

// Point i of a y line is stored (lines hold nky points; a Bluestein array is longer).
template <int BY>
__device__ __forceinline__ bool yin(const March3DArgs& a, int i)
{
    return BY == kLenPow2 || i < a.nky;
}

// Transform of one line of n points held by a wave as N array points (lane t, slot q: point
// t + 64 q).  KIND kLenPow2: n = N, the wave FFT.  kLenMixed: n = 2^a 3^b 5^c <= N, the
// mixed-radix transform in the wave's LDS buffer (tw: [n] of n).  kLenBluestein: Bluestein's
// chirp-z identity on the power-of-two array length N >= 2n - 1, X_k = c_k IFFT_N(FFT_N(c x) B)_k
// with the chirp c_j = exp(SIGN i pi j^2 / n) and B the spectrum of the wrapped conjugate chirp
// over N, built in FP64 on the host (c = 0 for j >= n, so points beyond n enter and leave as zero).
template <int N, int SIGN, int KIND>
__device__ __forceinline__ void line_fft(float2 (&v)[N / 64], float2* buf, const float2* __restrict__ tw,
                                         const float2* __restrict__ chirp, const float2* __restrict__ bch,
                                         const MixedRadix& mr, int n, int t)
{
    constexpr int P = N / 64;
    if constexpr (KIND == kLenPow2) {
        wfft<N, SIGN>(v, buf, tw, t);
    } else if constexpr (KIND == kLenMixed) {
        int pp = 0;
        mixed_fft<N, P, SIGN, false>(v, buf, buf, tw, mr, t, pp, []() { wave_sync_lds(); });
    } else {
        auto chirp_mul = [&]() {
            float2 c[P];
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int j = t + 64 * q;
                const float2 cj = j < n ? chirp[j] : make_float2(0.f, 0.f);  // exp(-i pi j^2 / n)
                c[q] = SIGN < 0 ? cj : make_float2(cj.x, -cj.y);
            }
#pragma unroll
            for (int q = 0; q < P; ++q) v[q] = cmul(v[q], c[q]);
        };
        chirp_mul();
        wfft<N, -1>(v, buf, tw, t);
        const float2* B = bch + (SIGN < 0 ? 0 : N);
        float2 bq[P];
#pragma unroll
        for (int q = 0; q < P; ++q) bq[q] = B[t + 64 * q];
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = cmul(v[q], bq[q]);
        wave_sync_lds();
        wfft<N, +1>(v, buf, tw, t);
        chirp_mul();
    }
}

template <int N, int SIGN, int BX>
__device__ __forceinline__ void xfft(float2 (&v)[N / 64], float2* buf, const March3DArgs& a, int t)
{
    line_fft<N, SIGN, BX>(v, buf, a.twx, a.chx, a.bchx, a.mrx, a.nkx, t);
}

template <int N, int SIGN, int BY>
__device__ __forceinline__ void yfft(float2 (&v)[N / 64], float2* buf, const March3DArgs& a, int t)
{
    line_fft<N, SIGN, BY>(v, buf, a.twy, a.chy, a.bchy, a.mry, a.nky, t);
}

}  // namespace

// Initial wavefields of the chunk (adjoint: source delta at the shot position and receiver
// spectra; forward: source delta, receiver zero), forward y-FFT, stored as the step state.
// Workgroup: WAVES rows (one wave each); grid: nitem * 2 * nmx rows.
template <int NY, int BY>
__global__ __launch_bounds__(256) void march3d_init(March3DArgs a)
{
    constexpr int P = NY / 64;
    constexpr int L = LdsLen<NY>::value;
    __shared__ float2 lds[4 * L];
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const size_t row = (size_t)blockIdx.x * 4 + w;
    const size_t nrow = (size_t)a.nitem * 2 * a.nmx;
    if (row >= nrow) return;  // whole waves only: the transform below has no workgroup barrier
    const int x = (int)(row % a.nmx);
    const int wf = (int)((row / a.nmx) & 1);
    const int item = (int)(row / ((size_t)2 * a.nmx));
    const int fl = a.f0 + item;  // frequency index in the band (one shot per launch)
    float2 v[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int y = t + 64 * q;
        float2 u = make_float2(0.f, 0.f);
        if (y < a.nmy) {
            const size_t ixy = (size_t)x * a.nmy + y;
            if (wf == 0)
                u = (int)ixy == (a.shot ? a.shot->igx : a.igx) ? a.spec_s[fl] : u;
            else if (a.adj)
                u = a.spec_r[ixy * a.nf + fl];
        }
        v[q] = u;
    }
    yfft<NY, -1, BY>(v, lds + w * L, a, t);
    float2* dst = state(a, item, wf) + (size_t)x * a.nky;
#pragma unroll
    for (int q = 0; q < P; ++q) if (yin<BY>(a, t + 64 * q)) dst[t + 64 * q] = v[q];
}

// kz phase of the KZ64 path with the column's part hoisted: c0 = w^2 po^2 - ky^2 (FP64, once per
// column), s = c0 - kxi^2 dkx^2 per point (kxi^2 exact in FP64).
__device__ __forceinline__ float2 kz_phase64(const March3DArgs& a, double c0, int ikx)
{
    const int kxi = ikx <= a.nkx / 2 ? ikx : a.nkx - ikx;
    return kz_phase_s64(a, fma(-(double)(kxi * kxi), a.dkx2, c0));
}

// kz phases of a lane's points ikx = t + 64 q of one ky column on a power-of-two nkx = N (KZ64):
// q <= P/2 directly, the others from their mirror nkx - ikx (the same |kx|, so the same phase bit
// for bit), which lane 64 - t holds at slot P - 1 - q (lane 0: its own slot P - q).
template <int N>
__device__ __forceinline__ void kz_phases64_col(const March3DArgs& a, double c0, int t, float2 (&ph)[N / 64])
{
    constexpr int P = N / 64;
#pragma unroll
    for (int q = 0; q <= P / 2 && q < P; ++q) ph[q] = kz_phase64(a, c0, t + 64 * q);
    const int src = (64 - t) & 63;
#pragma unroll
    for (int q = P / 2 + 1; q < P; ++q) {
        const float2 m = ph[P - 1 - q];
        const float x = __shfl(m.x, src, 64), y = __shfl(m.y, src, 64);
        ph[q] = t == 0 ? ph[P - q] : make_float2(x, y);
    }
}

// x-pass.  Workgroup: C waves, wave w owns column ky0 + w of every stepped wavefield (NWF = 2:
// source and receiver, which share the kz phase; NWF = 1: the one wmask names).  The tile of C
// consecutive ky columns is staged through LDS so the global reads and writes are row segments
// of C complex.  conj_src / conj_rec: use conj(L) (-dz).  Workgroups are renumbered so that each
// XCD (consecutive workgroup ids go round robin over the 8 XCDs) takes a contiguous run of tiles:
// neighbouring tiles share 128-byte lines, which then meet in one L2.
template <int NX, int NWF, int C, int BX>
__global__ __launch_bounds__(64 * C) __attribute__((amdgpu_waves_per_eu(NX >= 1024 ? 2 : 4))) void march3d_cols(March3DArgs a)
{
    constexpr int P = NX / 64;
    constexpr int L = LdsLen<NX>::value;
    extern __shared__ float2 smem[];
    const int tiles = (a.nky + C - 1) / C;  // the last tile is partial when C does not divide nky
    const int nblk = a.nitem * tiles;
    const int full = nblk & ~7;
    int b = blockIdx.x;
    if (b < full) b = (b & 7) * (full >> 3) + (b >> 3);
    const int item = b / tiles;
    const int ky0 = (b - item * tiles) * C;
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int wf1 = (a.wmask & 1) ? 0 : 1;  // the wavefield of an NWF = 1 launch
    auto wf_of = [&](int k) { return NWF == 2 ? k : wf1; };
    // stage: buffer (c * NWF + k) holds column ky0 + c of the k-th stepped wavefield.  The loads of
    // one wavefield are all issued before its LDS writes (rows x0 + 64 j, j < P cover x < nmx).
    const int c = threadIdx.x % C, x0 = threadIdx.x / C;
    const bool kin = ky0 + c < a.nky;
#pragma unroll
    for (int k = 0; k < NWF; ++k) {
        const float2* src = state(a, item, wf_of(k)) + ky0 + c;
        float2* dst = smem + (c * NWF + k) * L;
        float2 st[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int x = x0 + 64 * j;
            st[j] = x < a.nmx && kin ? src[(size_t)x * a.nky] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int x = x0 + 64 * j;
            if (x < a.nmx) dst[lds_pad(x)] = st[j];
        }
    }
    __syncthreads();
    float2 v[NWF][P];
#pragma unroll
    for (int k = 0; k < NWF; ++k) {
        float2* buf = smem + (w * NWF + k) * L;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int x = t + 64 * q;
            v[k][q] = x < a.nmx ? buf[lds_pad(x)] : make_float2(0.f, 0.f);
        }
        wave_sync_lds();
        xfft<NX, -1, BX>(v[k], buf, a, t);
    }
    const int iw = a.ifmin + a.f0 + item;
    const double wd = (double)iw * a.dw;
    const double p2 = a.p2[a.iz];
    const int iky = ky0 + w;
    const int kyi = iky <= a.nky / 2 ? iky : a.nky - iky;
    const double ky = kyi * a.dky;
    const double c0 = wd * wd * p2 - ky * ky;
    bool cj[NWF];
#pragma unroll
    for (int k = 0; k < NWF; ++k) cj[k] = wf_of(k) == 0 ? a.conj_src : a.conj_rec;
    float2 phm[P];
    const bool mirror = BX == kLenPow2 && a.prec == SWEM_PREC_KZ64;
    if constexpr (BX == kLenPow2)
        if (mirror) kz_phases64_col<NX>(a, c0, t, phm);
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const float2 ph = mirror ? phm[q]
                          : a.prec == SWEM_PREC_KZ64 ? kz_phase64(a, c0, t + 64 * q)
                                                     : kz_phase(a, wd, p2, a.p2f[a.iz], t + 64 * q, iky);
#pragma unroll
        for (int k = 0; k < NWF; ++k) v[k][q] = cmul(v[k][q], make_float2(ph.x, cj[k] ? -ph.y : ph.y));
    }
#pragma unroll
    for (int k = 0; k < NWF; ++k) {
        float2* buf = smem + (w * NWF + k) * L;
        wave_sync_lds();
        xfft<NX, +1, BX>(v[k], buf, a, t);
        wave_sync_lds();
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int x = t + 64 * q;
            if (x < a.nmx) buf[lds_pad(x)] = v[k][q];
        }
    }
    __syncthreads();
    // recomputed rather than kept live through the transforms (register pressure)
    int c1 = c, x1 = x0;
    asm volatile("" : "+v"(c1), "+v"(x1));
#pragma unroll
    for (int k = 0; k < NWF; ++k) {
        float2* dst = state(a, item, wf_of(k)) + ky0 + c1;
        const float2* src = smem + (c1 * NWF + k) * L;
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int x = x1 + 64 * j;
            if (x < a.nmx && ky0 + c1 < a.nky) dst[(size_t)x * a.nky] = src[lds_pad(x)];
        }
    }
}

// ------------------------------------------------------------------------ PSPI (compat)
// pspiop (wem.c:380-471) in the same two passes: the x-pass transforms each ky column once and,
// for every vref the depth uses, multiplies by its kz phase and transforms back, writing one
// slice per vref; the y-pass transforms the rows of those slices back and keeps, per point, the
// slice of its vref iref1 -- its real part where vref2 - vref1 > 10 (linear_interp returns a
// float, wem.c:450-458, 473-478).  No taper and no split-step phase (wem.c:418, 461-463).

__device__ __forceinline__ float2* rslice(const March3DArgs& a, int item, int wf, int ir)
{
    return a.R + (((size_t)item * 2 + wf) * a.nref + ir) * (size_t)a.nmx * a.nky;
}

// x-pass: as march3d_cols up to the forward x-FFT, then per vref in use: kz phase of
// vref[ir][iz] (conj for a -dz step), inverse x-FFT, slice tile written through LDS.
template <int NX, int NWF, int C, int BX>
__global__ __launch_bounds__(64 * C) __attribute__((amdgpu_waves_per_eu(NX >= 512 ? 2 : 4))) void march3d_cols_pspi(
    March3DArgs a)
{
    constexpr int P = NX / 64;
    constexpr int L = LdsLen<NX>::value;
    extern __shared__ float2 smem[];
    const int tiles = (a.nky + C - 1) / C;  // the last tile is partial when C does not divide nky
    const int nblk = a.nitem * tiles;
    const int full = nblk & ~7;
    int b = blockIdx.x;
    if (b < full) b = (b & 7) * (full >> 3) + (b >> 3);
    const int item = b / tiles;
    const int ky0 = (b - item * tiles) * C;
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int wf1 = (a.wmask & 1) ? 0 : 1;
    auto wf_of = [&](int k) { return NWF == 2 ? k : wf1; };
    const int c = threadIdx.x % C, x0 = threadIdx.x / C;
    const bool kin = ky0 + c < a.nky;
#pragma unroll
    for (int k = 0; k < NWF; ++k) {
        const float2* src = state(a, item, wf_of(k)) + ky0 + c;
        float2* dst = smem + (c * NWF + k) * L;
        float2 st[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int x = x0 + 64 * j;
            st[j] = x < a.nmx && kin ? src[(size_t)x * a.nky] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int x = x0 + 64 * j;
            if (x < a.nmx) dst[lds_pad(x)] = st[j];
        }
    }
    __syncthreads();
    float2 v[NWF][P];
#pragma unroll
    for (int k = 0; k < NWF; ++k) {
        float2* buf = smem + (w * NWF + k) * L;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int x = t + 64 * q;
            v[k][q] = x < a.nmx ? buf[lds_pad(x)] : make_float2(0.f, 0.f);
        }
        wave_sync_lds();
        xfft<NX, -1, BX>(v[k], buf, a, t);
    }
    const int iw = a.ifmin + a.f0 + item;
    const double wd = (double)iw * a.dw;
    const int iky = ky0 + w;
    const int kyi = iky <= a.nky / 2 ? iky : a.nky - iky;
    const double ky = kyi * a.dky;
    bool cj[NWF];
#pragma unroll
    for (int k = 0; k < NWF; ++k) cj[k] = wf_of(k) == 0 ? a.conj_src : a.conj_rec;
    const uint32_t used = a.used[a.iz];
    int c1 = c, x1 = x0;
    asm volatile("" : "+v"(c1), "+v"(x1));
    for (int ir = 0; ir < a.nref; ++ir) {
        if (!((used >> ir) & 1u)) continue;
        const double p2 = a.p2[(size_t)a.iz * a.nref + ir];
        const float p2f = a.p2f[(size_t)a.iz * a.nref + ir];
        const double c0 = wd * wd * p2 - ky * ky;
        // the two wavefields of a migration share each phase (the one-wavefield shapes keep
        // computing it per point: the mirrored set would spill there)
        float2 phm[P];
        const bool mirror = NWF == 2 && BX == kLenPow2 && a.prec == SWEM_PREC_KZ64;
        if constexpr (NWF == 2 && BX == kLenPow2)
            if (mirror) kz_phases64_col<NX>(a, c0, t, phm);
#pragma unroll
        for (int k = 0; k < NWF; ++k) {
            float2 bq[P];
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const float2 ph = mirror ? phm[q]
                                  : a.prec == SWEM_PREC_KZ64 ? kz_phase64(a, c0, t + 64 * q)
                                                             : kz_phase(a, wd, p2, p2f, t + 64 * q, iky);
                bq[q] = cmul(v[k][q], make_float2(ph.x, cj[k] ? -ph.y : ph.y));
            }
            float2* buf = smem + (w * NWF + k) * L;
            wave_sync_lds();
            xfft<NX, +1, BX>(bq, buf, a, t);
            wave_sync_lds();
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int x = t + 64 * q;
                if (x < a.nmx) buf[lds_pad(x)] = bq[q];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NWF; ++k) {
            float2* dst = rslice(a, item, wf_of(k), ir) + ky0 + c1;
            const float2* src = smem + (c1 * NWF + k) * L;
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const int x = x1 + 64 * j;
                if (x < a.nmx && ky0 + c1 < a.nky) dst[(size_t)x * a.nky] = src[lds_pad(x)];
            }
        }
        __syncthreads();  // the tile is rewritten by the next vref
    }
}

// x-pass of the linear-mode modelling receiver (the transpose of march3d_cols_pspi): for every
// vref in use at the depth, the slice R[ir] (the weighted wavefield, y transformed) is x
// transformed and multiplied by its kz phase (conj for the -dz step), the products are summed,
// and one inverse x-FFT gives the stepped receiver (y transformed) in the state.
template <int NX, int C, int BX>
__global__ __launch_bounds__(64 * C) __attribute__((amdgpu_waves_per_eu(NX >= 512 ? 2 : 4))) void march3d_cols_gather(
    March3DArgs a)
{
    constexpr int P = NX / 64;
    constexpr int L = LdsLen<NX>::value;
    extern __shared__ float2 smem[];
    const int tiles = (a.nky + C - 1) / C;
    const int nblk = a.nitem * tiles;
    const int full = nblk & ~7;
    int b = blockIdx.x;
    if (b < full) b = (b & 7) * (full >> 3) + (b >> 3);
    const int item = b / tiles;
    const int ky0 = (b - item * tiles) * C;
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int c = threadIdx.x % C, x0 = threadIdx.x / C;
    const bool kin = ky0 + c < a.nky;
    const int iw = a.ifmin + a.f0 + item;
    const double wd = (double)iw * a.dw;
    const int iky = ky0 + w;
    const int kyi = iky <= a.nky / 2 ? iky : a.nky - iky;
    const double ky = kyi * a.dky;
    const uint32_t used = a.used[a.iz];
    float2* buf = smem + w * L;
    float2 acc[P];
#pragma unroll
    for (int q = 0; q < P; ++q) acc[q] = make_float2(0.f, 0.f);
    for (int ir = 0; ir < a.nref; ++ir) {
        if (!((used >> ir) & 1u)) continue;
        // stage the tile's columns of slice ir (row segments of C complex)
        {
            const float2* src = rslice(a, item, 1, ir) + ky0 + c;
            float2* dst = smem + c * L;
            float2 st[P];
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const int x = x0 + 64 * j;
                st[j] = x < a.nmx && kin ? src[(size_t)x * a.nky] : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const int x = x0 + 64 * j;
                if (x < a.nmx) dst[lds_pad(x)] = st[j];
            }
        }
        __syncthreads();
        float2 v[P];
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int x = t + 64 * q;
            v[q] = x < a.nmx ? buf[lds_pad(x)] : make_float2(0.f, 0.f);
        }
        wave_sync_lds();
        xfft<NX, -1, BX>(v, buf, a, t);
        const double p2 = a.p2[(size_t)a.iz * a.nref + ir];
        const float p2f = a.p2f[(size_t)a.iz * a.nref + ir];
        const double c0 = wd * wd * p2 - ky * ky;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const float2 ph = a.prec == SWEM_PREC_KZ64 ? kz_phase64(a, c0, t + 64 * q)
                                                       : kz_phase(a, wd, p2, p2f, t + 64 * q, iky);
            acc[q] = cadd(acc[q], cmul(v[q], make_float2(ph.x, a.conj_rec ? -ph.y : ph.y)));
        }
        __syncthreads();  // the tile is restaged for the next vref
    }
    xfft<NX, +1, BX>(acc, buf, a, t);
    wave_sync_lds();
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = t + 64 * q;
        if (x < a.nmx) buf[lds_pad(x)] = acc[q];
    }
    __syncthreads();
    int c1 = c, x1 = x0;
    asm volatile("" : "+v"(c1), "+v"(x1));
    float2* dst = state(a, item, 1) + ky0 + c1;
    const float2* src = smem + c1 * L;
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const int x = x1 + 64 * j;
        if (x < a.nmx && ky0 + c1 < a.nky) dst[(size_t)x * a.nky] = src[lds_pad(x)];
    }
}

// u = the pspiop selection of row x of wavefield wf from the slices of depth iz_sel: inverse
// y-FFT of each slice in use, per point the slice of its iref1, real part where cut.
template <int NY, int BY>
__device__ __forceinline__ void pspi_select_row(const March3DArgs& a, int item, int wf, int x, float2 (&u)[NY / 64],
                                                float2* buf, int t)
{
    constexpr int P = NY / 64;
    const uint32_t used = a.used[a.iz_sel];
    const uint8_t* sr = a.sel + (size_t)a.iz_sel * a.nm + (size_t)x * a.nmy;
    uint8_t s1[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int y = t + 64 * q;
        s1[q] = y < a.nmy ? sr[y] : (uint8_t)0x7f;
        u[q] = make_float2(0.f, 0.f);
    }
    // linear mode: the weights of wem.c:476's interpolation, (1 - w2) for iref1 and w2 for iref2
    // at the cut points, 1 for iref1 elsewhere
    uint8_t s2[P];
    float wt2[P];
    if (a.linear) {
        const uint8_t* sr2 = a.sel2 + (size_t)a.iz_sel * a.nm + (size_t)x * a.nmy;
        const float* wr = a.w2 + (size_t)a.iz_sel * a.nm + (size_t)x * a.nmy;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int y = t + 64 * q;
            s2[q] = y < a.nmy ? sr2[y] : (uint8_t)0x7f;
            wt2[q] = y < a.nmy ? wr[y] : 0.f;
        }
    }
    for (int ir = 0; ir < a.nref; ++ir) {
        if (!((used >> ir) & 1u)) continue;
        const float2* row = rslice(a, item, wf, ir) + (size_t)x * a.nky;
        float2 bq[P];
#pragma unroll
        for (int q = 0; q < P; ++q) bq[q] = yin<BY>(a, t + 64 * q) ? row[t + 64 * q] : make_float2(0.f, 0.f);
        yfft<NY, +1, BY>(bq, buf, a, t);
        wave_sync_lds();
        if (a.linear) {
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const bool cut = s1[q] >> 7;
                float wt = (s1[q] & 0x7f) == ir ? (cut ? 1.f - wt2[q] : 1.f) : 0.f;
                if (cut && s2[q] == ir) wt += wt2[q];
                u[q] = cadd(u[q], cscale(bq[q], wt));
            }
        } else {
#pragma unroll
            for (int q = 0; q < P; ++q)
                if ((s1[q] & 0x7f) == ir) u[q] = (s1[q] >> 7) ? make_float2(bq[q].x, 0.f) : bq[q];
        }
    }
}

// Weight of vref ir at a point in the linear mode: (1 - w2) for iref1 and w2 for iref2 where cut,
// 1 for iref1 elsewhere (wem.c:476).
__device__ __forceinline__ float pspi3_weight(uint8_t s1, uint8_t s2, float w2, int ir)
{
    const bool cut = s1 >> 7;
    float wt = (s1 & 0x7f) == ir ? (cut ? 1.f - w2 : 1.f) : 0.f;
    if (cut && s2 == ir) wt += w2;
    return wt;
}

// y-pass of PSPI, F waves (frequencies g F + f) of one row x.
//   phase 0 migration: select the stepped wavefields (the source is re-read from its state when
//           only the receiver steps), image (summed over the F frequencies in a fixed order),
//           forward y-FFT of what steps next
//   phase 1 modelling source: select, store the depth's source in smig (zero where z < sz),
//           forward y-FFT for the next step
//   phase 2 modelling receiver: select the previous step's result (zero at the first), inject
//           m smig of this depth, forward y-FFT
//   phase 3 modelling output: select the last step's result (zero when no step ran), spectra out
template <int NY, int BY>
__global__ __launch_bounds__(512) void march3d_rows_pspi(March3DArgs a, int F, int phase)
{
    constexpr int P = NY / 64;
    constexpr int L = LdsLen<NY>::value;
    extern __shared__ float2 smem[];
    const int ngroups = (a.nitem + F - 1) / F;
    const int g = blockIdx.x % ngroups, x = blockIdx.x / ngroups;
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int item = g * F + w;
    const bool active = item < a.nitem;
    float2* buf = smem + w * L;
    if (phase >= 4) {
        // linear-mode modelling receiver.  phase 4 at depth iz: u = the last step's result (the
        // inverse y-FFT of the state; zero before the first step) plus the injection m smig of
        // this depth, then per vref in use u weighted, forward y-FFT, slice R[ir].  phase 5: the
        // last step's result out as the modelled spectra.
        if (!active) return;  // no workgroup barriers below
        const int fl = a.f0 + item;
        float2 u[P];
        if (a.first) {
#pragma unroll
            for (int q = 0; q < P; ++q) u[q] = make_float2(0.f, 0.f);
        } else {
            const float2* src = state(a, item, 1) + (size_t)x * a.nky;
#pragma unroll
            for (int q = 0; q < P; ++q) u[q] = yin<BY>(a, t + 64 * q) ? src[t + 64 * q] : make_float2(0.f, 0.f);
            yfft<NY, +1, BY>(u, buf, a, t);
            wave_sync_lds();
        }
        if (phase == 5) {
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int y = t + 64 * q;
                if (y < a.nmy) a.dgw[((size_t)x * a.nmy + y) * a.nf + fl] = u[q];
            }
            return;
        }
        const float2* sm = a.smig + ((size_t)item * a.nz + a.iz) * a.nm + (size_t)x * a.nmy;
        const float* mt = a.mT + (size_t)a.iz * a.nm + (size_t)x * a.nmy;
        const size_t o = (size_t)a.iz * a.nm + (size_t)x * a.nmy;
        uint8_t s1[P], s2[P];
        float wt2[P];
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int y = t + 64 * q;
            if (y < a.nmy) {
                u[q] = cadd(u[q], cscale(sm[y], mt[y]));
                s1[q] = a.sel[o + y];
                s2[q] = a.sel2[o + y];
                wt2[q] = a.w2[o + y];
            } else {
                u[q] = make_float2(0.f, 0.f);
                s1[q] = s2[q] = 0x7f;
                wt2[q] = 0.f;
            }
        }
        const uint32_t used = a.used[a.iz];
        for (int ir = 0; ir < a.nref; ++ir) {
            if (!((used >> ir) & 1u)) continue;
            float2 v[P];
#pragma unroll
            for (int q = 0; q < P; ++q) v[q] = cscale(u[q], pspi3_weight(s1[q], s2[q], wt2[q], ir));
            wave_sync_lds();
            yfft<NY, -1, BY>(v, buf, a, t);
            float2* dst = rslice(a, item, 1, ir) + (size_t)x * a.nky;
#pragma unroll
            for (int q = 0; q < P; ++q) if (yin<BY>(a, t + 64 * q)) dst[t + 64 * q] = v[q];
        }
        return;
    }
    if (phase != 0) {
        if (!active) return;  // no workgroup barriers below
        const int fl = a.f0 + item;
        float2 u[P];
        if (phase == 1) {
            if (a.ds) pspi_select_row<NY, BY>(a, item, 0, x, u, buf, t);
            float2* sm = a.smig + ((size_t)item * a.nz + a.iz) * a.nm + (size_t)x * a.nmy;
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int y = t + 64 * q;
                if (y < a.nmy) sm[y] = a.ds ? u[q] : make_float2(0.f, 0.f);
            }
            if (!(a.ds && a.next_s)) return;
        } else {
            if (a.first) {
#pragma unroll
                for (int q = 0; q < P; ++q) u[q] = make_float2(0.f, 0.f);
            } else {
                pspi_select_row<NY, BY>(a, item, 1, x, u, buf, t);
            }
            if (phase == 3) {
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    const int y = t + 64 * q;
                    if (y < a.nmy) a.dgw[((size_t)x * a.nmy + y) * a.nf + fl] = u[q];
                }
                return;
            }
            const float2* sm = a.smig + ((size_t)item * a.nz + a.iz) * a.nm + (size_t)x * a.nmy;
            const float* mt = a.mT + (size_t)a.iz * a.nm + (size_t)x * a.nmy;
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int y = t + 64 * q;
                if (y < a.nmy) u[q] = cadd(u[q], cscale(sm[y], mt[y]));
            }
        }
        wave_sync_lds();
        yfft<NY, -1, BY>(u, buf, a, t);
        float2* dst = state(a, item, phase == 1 ? 0 : 1) + (size_t)x * a.nky;
#pragma unroll
        for (int q = 0; q < P; ++q) if (yin<BY>(a, t + 64 * q)) dst[t + 64 * q] = u[q];
        return;
    }
    // phase 0: migration
    const int iw = a.ifmin + a.f0 + (active ? item : 0);
    float2 vs[P], vr[P];
#pragma unroll
    for (int q = 0; q < P; ++q) vs[q] = vr[q] = make_float2(0.f, 0.f);
    if (active && a.ds) {
        pspi_select_row<NY, BY>(a, item, 0, x, vs, buf, t);
    } else if (active && a.dr) {
        // an unstepped source: its state is FFT_y(u), so the inverse transform returns nky u
        const float2* src = state(a, item, 0) + (size_t)x * a.nky;
#pragma unroll
        for (int q = 0; q < P; ++q) vs[q] = yin<BY>(a, t + 64 * q) ? src[t + 64 * q] : make_float2(0.f, 0.f);
        yfft<NY, +1, BY>(vs, buf, a, t);
        wave_sync_lds();
        const float rn = 1.f / (float)a.nky;  // exact when nky is a power of two
#pragma unroll
        for (int q = 0; q < P; ++q) vs[q] = cscale(vs[q], rn);
    }
    if (active && a.dr) pspi_select_row<NY, BY>(a, item, 1, x, vr, buf, t);
    float cc[P];
    const float factor = iw == 0 ? 1.f : 2.f;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        if (t + 64 * q >= a.nmy) vs[q] = vr[q] = make_float2(0.f, 0.f);
        cc[q] = (active && a.dr) ? factor * (vr[q].x * vs[q].x + vr[q].y * vs[q].y) : 0.f;
    }
    wave_sync_lds();
    float* red = reinterpret_cast<float*>(buf);
#pragma unroll
    for (int q = 0; q < P; ++q) red[t + 64 * q] = cc[q];
    __syncthreads();
    float* part = a.partial + (size_t)g * a.nm + (size_t)x * a.nmy;
    for (int y = threadIdx.x; y < a.nmy; y += blockDim.x) {
        float sum = 0.f;
        for (int ff = 0; ff < F; ++ff) sum += reinterpret_cast<const float*>(smem + ff * L)[y];
        part[y] = sum;
    }
    __syncthreads();
    if (active && a.ds && a.next_s) {
        yfft<NY, -1, BY>(vs, buf, a, t);
        float2* dst = state(a, item, 0) + (size_t)x * a.nky;
#pragma unroll
        for (int q = 0; q < P; ++q) if (yin<BY>(a, t + 64 * q)) dst[t + 64 * q] = vs[q];
        wave_sync_lds();
    }
    if (active && a.dr && a.next_r) {
        yfft<NY, -1, BY>(vr, buf, a, t);
        float2* dst = state(a, item, 1) + (size_t)x * a.nky;
#pragma unroll
        for (int q = 0; q < P; ++q) if (yin<BY>(a, t + 64 * q)) dst[t + 64 * q] = vr[q];
    }
}

// y-pass of the migration (adjoint): F waves, wave f owns row x of the source and the receiver
// of frequency item g F + f.  ds / dr: the source / receiver was stepped at this depth; next_s /
// next_r: it is stepped at the next depth (its forward y-FFT is needed); the imaging runs when dr.
// The split-step phase is computed once and applied conjugated to the source.
template <int NY, int BY>
__global__ __launch_bounds__(BY ? 512 : 1024) void march3d_rows_adj(March3DArgs a, int F)
{
    constexpr int P = NY / 64;
    constexpr int L = LdsLen<NY>::value;
    extern __shared__ float2 smem[];
    const int ngroups = (a.nitem + F - 1) / F;
    const int g = blockIdx.x % ngroups, x = blockIdx.x / ngroups;
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int item = g * F + w;
    const bool active = item < a.nitem;
    const int iw = a.ifmin + a.f0 + (active ? item : 0);
    const double wd = (double)iw * a.dw;
    float2* buf = smem + w * L;
    const bool need_s = active && (a.ds || a.dr), need_r = active && a.dr;
    float2 vs[P], vr[P];
#pragma unroll
    for (int q = 0; q < P; ++q) vs[q] = vr[q] = make_float2(0.f, 0.f);
    // both rows are loaded before the first transform (their latencies overlap)
    if (need_s) {
        const float2* src = state(a, item, 0) + (size_t)x * a.nky;
#pragma unroll
        for (int q = 0; q < P; ++q) vs[q] = yin<BY>(a, t + 64 * q) ? src[t + 64 * q] : make_float2(0.f, 0.f);
    }
    if (need_r) {
        const float2* src = state(a, item, 1) + (size_t)x * a.nky;
#pragma unroll
        for (int q = 0; q < P; ++q) vr[q] = yin<BY>(a, t + 64 * q) ? src[t + 64 * q] : make_float2(0.f, 0.f);
    }
    if (need_s) {
        yfft<NY, +1, BY>(vs, buf, a, t);
        wave_sync_lds();
        if (!a.ds) {  // unstepped source: its state is FFT_y(u), the inverse returns nky u
            const float rn = 1.f / (float)a.nky;  // exact when nky is a power of two
#pragma unroll
            for (int q = 0; q < P; ++q) vs[q] = cscale(vs[q], rn);
        }
    }
    if (need_r) yfft<NY, +1, BY>(vr, buf, a, t);
    float cc[P];
    const float factor = iw == 0 ? 1.f : 2.f;
    const float tx = a.taper_x[x];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int y = t + 64 * q;
        if (y < a.nmy) {
            if (a.ds || a.dr) {
                const float2 sp = ss_phase(a, wd, (size_t)x * a.nmy + y);
                const float tp = tx * a.taper_y[y];
                if (a.ds) vs[q] = cscale(cmulc(vs[q], sp), tp);
                if (a.dr) vr[q] = cscale(cmul(vr[q], sp), tp);
            }
        } else {
            vs[q] = vr[q] = make_float2(0.f, 0.f);
        }
        cc[q] = need_r ? factor * (vr[q].x * vs[q].x + vr[q].y * vs[q].y) : 0.f;
    }
    // image of the workgroup's frequencies, summed in a fixed order (wem.c:243-247)
    wave_sync_lds();
    float* red = reinterpret_cast<float*>(buf);
#pragma unroll
    for (int q = 0; q < P; ++q) red[t + 64 * q] = cc[q];
    __syncthreads();
    float* part = a.partial + (size_t)g * a.nm + (size_t)x * a.nmy;
    for (int y = threadIdx.x; y < a.nmy; y += blockDim.x) {
        float sum = 0.f;
        for (int ff = 0; ff < F; ++ff) sum += reinterpret_cast<const float*>(smem + ff * L)[y];
        part[y] = sum;
    }
    __syncthreads();
    // forward y-FFTs that start the next step
    if (active && a.ds && a.next_s) {
        yfft<NY, -1, BY>(vs, buf, a, t);
        float2* dst = state(a, item, 0) + (size_t)x * a.nky;
#pragma unroll
        for (int q = 0; q < P; ++q) if (yin<BY>(a, t + 64 * q)) dst[t + 64 * q] = vs[q];
        wave_sync_lds();
    }
    if (active && a.dr && a.next_r) {
        yfft<NY, -1, BY>(vr, buf, a, t);
        float2* dst = state(a, item, 1) + (size_t)x * a.nky;
#pragma unroll
        for (int q = 0; q < P; ++q) if (yin<BY>(a, t + 64 * q)) dst[t + 64 * q] = vr[q];
    }
}

// y-pass of modelling (forward), one wavefield per wave, F frequencies of one row x.
//   phase 0 (source march): IFFTy, SS (conj) and taper as the adjoint source, store the
//           spatial source of this depth in smig (zero where z < sz), FFTy for the next step.
//   phase 1 (receiver march, upward): IFFTy of the previous step's result (skipped at the first
//           step, whose state is zero), inject m * smig, taper, conj(SS), FFTy.
//   phase 2 (output): IFFTy of the last step's result, write the spectra of the record.
//   phase 3 (incidence angles, compute_angles wem.c:712-719): the source marched with the
//           receiver's step; at the first call the state is the source trace, later IFFTy of
//           the previous step's result is stored in smig at depth iz_sel; with ds the next step
//           starts (taper, conj(SS), FFTy).
template <int NY, int BY>
__global__ __launch_bounds__(BY ? 512 : 1024) void march3d_rows_fwd(March3DArgs a, int F, int phase)
{
    constexpr int P = NY / 64;
    constexpr int L = LdsLen<NY>::value;
    extern __shared__ float2 smem[];
    const int ngroups = (a.nitem + F - 1) / F;
    const int g = blockIdx.x % ngroups, x = blockIdx.x / ngroups;
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int item = g * F + w;
    if (item >= a.nitem) return;  // no workgroup barriers in this kernel
    const int fl = a.f0 + item;
    const double wd = (double)(a.ifmin + fl) * a.dw;
    float2* buf = smem + w * L;
    const int wf = phase == 0 ? 0 : 1;
    float2* row = state(a, item, wf) + (size_t)x * a.nky;
    float2 v[P];
    if (phase == 3) {
        const float tx = a.taper_x[x];
        if (a.first) {
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int y = t + 64 * q;
                v[q] = (y < a.nmy && x * a.nmy + y == a.igx) ? a.spec_s[fl] : make_float2(0.f, 0.f);
            }
        } else {
#pragma unroll
            for (int q = 0; q < P; ++q) v[q] = yin<BY>(a, t + 64 * q) ? row[t + 64 * q] : make_float2(0.f, 0.f);
            yfft<NY, +1, BY>(v, buf, a, t);
            float2* sm = a.smig + ((size_t)item * a.nz + a.iz_sel) * a.nm + (size_t)x * a.nmy;
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int y = t + 64 * q;
                if (y < a.nmy) sm[y] = v[q];
            }
        }
        if (!a.ds) return;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int y = t + 64 * q;
            v[q] = y < a.nmy ? cmulc(cscale(v[q], tx * a.taper_y[y]), ss_phase(a, wd, (size_t)x * a.nmy + y))
                             : make_float2(0.f, 0.f);
        }
        wave_sync_lds();
        yfft<NY, -1, BY>(v, buf, a, t);
#pragma unroll
        for (int q = 0; q < P; ++q) if (yin<BY>(a, t + 64 * q)) row[t + 64 * q] = v[q];
        return;
    }
    if (phase == 1 && a.first) {
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = make_float2(0.f, 0.f);
    } else {
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = yin<BY>(a, t + 64 * q) ? row[t + 64 * q] : make_float2(0.f, 0.f);
        yfft<NY, +1, BY>(v, buf, a, t);
    }
    const float tx = a.taper_x[x];
    float2* sm = a.smig + ((size_t)item * a.nz + a.iz) * a.nm + (size_t)x * a.nmy;
    const float* mt = a.mT + (size_t)a.iz * a.nm + (size_t)x * a.nmy;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int y = t + 64 * q;
        if (y >= a.nmy) {
            v[q] = make_float2(0.f, 0.f);
            continue;
        }
        if (phase == 0) {
            if (a.ds) v[q] = cscale(cmulc(v[q], ss_phase(a, wd, (size_t)x * a.nmy + y)), tx * a.taper_y[y]);
            sm[y] = a.ds ? v[q] : make_float2(0.f, 0.f);
        } else if (phase == 1) {
            const float2 u = cadd(v[q], cscale(sm[y], mt[y]));
            v[q] = cmulc(cscale(u, tx * a.taper_y[y]), ss_phase(a, wd, (size_t)x * a.nmy + y));
        }
    }
    if (phase == 2) {
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int y = t + 64 * q;
            if (y < a.nmy) a.dgw[((size_t)x * a.nmy + y) * a.nf + fl] = v[q];
        }
        return;
    }
    if (phase == 0 && !(a.ds && a.next_s)) return;  // state unchanged / no further step
    wave_sync_lds();
    yfft<NY, -1, BY>(v, buf, a, t);
#pragma unroll
    for (int q = 0; q < P; ++q) if (yin<BY>(a, t + 64 * q)) row[t + 64 * q] = v[q];
}

// ------------------------------------------------------------------------ dispatch

// Array length of a 3-D line transform of n points: n for powers of two 128..1024, else the
// Bluestein length (the power of two >= max(128, 2n - 1), up to 1024: n <= 512; n = 1 is the
// nkx of a single inline, nmx = 1); 0 if neither.
int march3d_kind(int n)
{
    if (n >= 128 && !(n & (n - 1))) return kLenPow2;
    return mixed_radix_plan(n, nullptr) ? kLenMixed : kLenBluestein;
}

int march3d_len(int n)
{
    if (n < 1) return 0;
    const int kind = march3d_kind(n);
    if (kind == kLenPow2) return n <= 1024 ? n : 0;
    int m = 128;
    if (kind == kLenMixed) {
        while (m < n) m <<= 1;
        return m <= 1024 ? m : 0;
    }
    if (n > 512) return 0;
    while (m < 2 * n - 1) m <<= 1;
    return m;
}

bool march3d_supported(int nkx, int nky) { return march3d_len(nkx) > 0 && march3d_len(nky) > 0; }

// ------------------------------------------------------------------------ elastic WEM (3-D)
// ewem (ewem.c:3-366) with nmy > 1 in the two passes of the 3-D march.  A depth step of the
// receiver pair is: separation (x-pass: FFT_x, the 2 x 2 map of kx per point, IFFT_x), truncation
// to the nmx x nmy grid (y-pass), the split-step steps of the potentials with vp and vs (x-pass kz
// phase; y-pass split-step phase and taper), the imaging against the source, the combination
// (x-pass) and truncation (y-pass).  Every x-pass applies one 1/(nkx nky): the state it writes, y
// transformed, returns the step's result in space under the inverse y-FFT.

__device__ __forceinline__ float2* eslot(const March3DArgs& a, const Ewem3DArgs& e, int item, int s)
{
    return e.W + ((size_t)item * 3 + s) * (size_t)a.nmx * a.nky;
}

// Initial state: the source delta at igx (slot 0) and the receiver spectra (slots 1 and 2, zero
// when modelling), forward y-FFT.  One wave per row; grid: nitem * 3 * nmx rows.
template <int NY, int BY>
__global__ __launch_bounds__(256) void ewem3d_init(March3DArgs a, Ewem3DArgs e)
{
    constexpr int P = NY / 64;
    constexpr int L = LdsLen<NY>::value;
    __shared__ float2 lds[4 * L];
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const size_t row = (size_t)blockIdx.x * 4 + w;
    if (row >= (size_t)a.nitem * 3 * a.nmx) return;  // whole waves only
    const int x = (int)(row % a.nmx);
    const int s = (int)((row / a.nmx) % 3);
    const int item = (int)(row / ((size_t)3 * a.nmx));
    const int fl = a.f0 + item;
    float2 v[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int y = t + 64 * q;
        float2 u = make_float2(0.f, 0.f);
        if (y < a.nmy) {
            const size_t ixy = (size_t)x * a.nmy + y;
            if (s == 0) u = (int)ixy == e.igx ? a.spec_s[fl] : u;
            else if (e.sep_adj) u = (s == 1 ? e.spx : e.spz)[ixy * a.nf + fl];
        }
        v[q] = u;
    }
    yfft<NY, -1, BY>(v, lds + w * L, a, t);
    float2* dst = eslot(a, e, item, s) + (size_t)x * a.nky;
#pragma unroll
    for (int q = 0; q < P; ++q) if (yin<BY>(a, t + 64 * q)) dst[t + 64 * q] = v[q];
}

// x-pass.  NWF = 1 (e.mode 3): the source (slot 0), kz phase of vp (conj_src).  NWF = 2: the
// receiver pair (slots 1, 2): MAPS, the 2 x 2 maps of kx (ewem.c:593-630, 680-710): e.mode 0
// separation, 1 combination, 4 combination of depth izc then separation of depth iz with the
// truncation to x < nmx between them; else e.mode 2, the kz phases of vp and vs (conj_rec).
// Columns staged through LDS in tiles of C.
template <int NX, int NWF, int C, int BX, bool MAPS>
__global__ __launch_bounds__(64 * C) __attribute__((amdgpu_waves_per_eu(NX >= 1024 ? 2 : (MAPS && NX >= 512 ? 3 : 4)))) void ewem3d_cols(
    March3DArgs ap, March3DArgs as, Ewem3DArgs e)
{
    constexpr int P = NX / 64;
    constexpr int L = LdsLen<NX>::value;
    extern __shared__ float2 smem[];
    const March3DArgs& a = ap;
    const int tiles = (a.nky + C - 1) / C;
    const int nblk = a.nitem * tiles;
    const int full = nblk & ~7;
    int b = blockIdx.x;
    if (b < full) b = (b & 7) * (full >> 3) + (b >> 3);
    const int item = b / tiles;
    const int ky0 = (b - item * tiles) * C;
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    auto slot_of = [&](int k) { return NWF == 1 ? 0 : 1 + k; };
    const int c = threadIdx.x % C, x0 = threadIdx.x / C;
    const bool kin = ky0 + c < a.nky;
#pragma unroll
    for (int k = 0; k < NWF; ++k) {
        const float2* src = eslot(a, e, item, slot_of(k)) + ky0 + c;
        float2* dst = smem + (c * NWF + k) * L;
        float2 st[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int x = x0 + 64 * j;
            st[j] = x < a.nmx && kin ? src[(size_t)x * a.nky] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int x = x0 + 64 * j;
            if (x < a.nmx) dst[lds_pad(x)] = st[j];
        }
    }
    __syncthreads();
    const int iw = a.ifmin + a.f0 + item;
    const int iky = ky0 + w;
    // a column of wavefield k: from the wave's buffer into registers, forward x-FFT
    auto fetch = [&](int k, float2 (&u)[P]) {
        float2* buf = smem + (w * NWF + k) * L;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int x = t + 64 * q;
            u[q] = x < a.nmx ? buf[lds_pad(x)] : make_float2(0.f, 0.f);
        }
        wave_sync_lds();
        xfft<NX, -1, BX>(u, buf, a, t);
    };
    // inverse x-FFT and back into the buffer (x < nmx)
    auto put = [&](int k, float2 (&u)[P]) {
        float2* buf = smem + (w * NWF + k) * L;
        wave_sync_lds();
        xfft<NX, +1, BX>(u, buf, a, t);
        wave_sync_lds();
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int x = t + 64 * q;
            if (x < a.nmx) buf[lds_pad(x)] = u[q];
        }
    };
    if constexpr (MAPS) {
        static_assert(NWF == 2, "the 2 x 2 maps act on the receiver pair");
        float2 v[2][P];
        fetch(0, v[0]);
        fetch(1, v[1]);
        const float wf = (float)iw * e.dwf;
        const int n = a.nkx;
        // the 2 x 2 map of kx (separation or combination) with the layer velocities of depth iz
        auto map = [&](bool sep, int iz, float sc) {
            const float vp = e.vp0[iz], vs = e.vs0[iz];
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int k = t + 64 * q;
                const float kx = (float)k < (float)n / 2.f ? e.dkxf * (float)k : -(e.dkxf * (float)n - e.dkxf * (float)k);
                float2 r1, r2;
                esep_k(kx, wf, vp, vs, sep, e.sep_adj != 0, v[0][q], v[1][q], r1, r2);
                v[0][q] = k < n ? cscale(r1, sc) : make_float2(0.f, 0.f);
                v[1][q] = k < n ? cscale(r2, sc) : make_float2(0.f, 0.f);
            }
        };
        if (e.mode == 4) {
            // combination of depth izc, back to x (truncated to x < nmx), separation of depth iz
            map(false, e.izc, e.s1);
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                float2* buf = smem + (w * NWF + k) * L;
                wave_sync_lds();
                xfft<NX, +1, BX>(v[k], buf, a, t);
#pragma unroll
                for (int q = 0; q < P; ++q)
                    if (t + 64 * q >= a.nmx) v[k][q] = make_float2(0.f, 0.f);
                wave_sync_lds();
                xfft<NX, -1, BX>(v[k], buf, a, t);
            }
            map(true, a.iz, e.s2);
        } else {
            map(e.mode == 0, a.iz, e.s1);
        }
        put(0, v[0]);
        put(1, v[1]);
    } else {
        // kz phases, one wavefield at a time (the source with vp; the pair with vp and vs)
        const double wd = (double)iw * a.dw;
        const int kyi = iky <= a.nky / 2 ? iky : a.nky - iky;
        const double ky = kyi * a.dky;
        const bool cj = NWF == 1 ? e.conj_src : e.conj_rec;
#pragma unroll
        for (int k = 0; k < NWF; ++k) {
            const March3DArgs& pl = (NWF == 2 && k == 1) ? as : ap;  // us2 steps with vs
            float2 v[P];
            fetch(k, v);
            const double p2 = pl.p2[a.iz];
            const double c0 = wd * wd * p2 - ky * ky;
            float2 phm[P];
            const bool mirror = BX == kLenPow2 && a.prec == SWEM_PREC_KZ64;
            if constexpr (BX == kLenPow2)
                if (mirror) kz_phases64_col<NX>(pl, c0, t, phm);
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const float2 ph = mirror ? phm[q]
                                  : a.prec == SWEM_PREC_KZ64 ? kz_phase64(pl, c0, t + 64 * q)
                                                             : kz_phase(pl, wd, p2, pl.p2f[a.iz], t + 64 * q, iky);
                v[q] = cmul(v[q], make_float2(ph.x, cj ? -ph.y : ph.y));
            }
            put(k, v);
        }
    }
    __syncthreads();
    int c1 = c, x1 = x0;
    asm volatile("" : "+v"(c1), "+v"(x1));
#pragma unroll
    for (int k = 0; k < NWF; ++k) {
        float2* dst = eslot(a, e, item, slot_of(k)) + ky0 + c1;
        const float2* src = smem + (c1 * NWF + k) * L;
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int x = x1 + 64 * j;
            if (x < a.nmx && ky0 + c1 < a.nky) dst[(size_t)x * a.nky] = src[lds_pad(x)];
        }
    }
}

// y-pass, F waves (frequencies g F + f) of one row x.
//   phase 0 migration: the stepped source (inverse y-FFT, conj split-step phase of vp, taper) and
//           potentials (split-step phases of vp and vs, taper); images factor Re(conj(s / maxs) up)
//           and factor Re(conj(s / maxs) us2) summed over the F frequencies (zero where the source is
//           not stepped: ewem.c's smig is zero there); forward y-FFTs of what steps next
//   phase 1 truncation of the receiver pair to y < nmy (between a map and a step)
//   phase 2 modelling source: stepped source stored (divided by max|wav|) in smig, zero where
//           z < sz; forward y-FFT for the next step
//   phase 3 modelling receiver: truncation, injection up += smig mpp, us2 += smig mps2, taper,
//           conj split-step phases (the forward branch of ssop, -dz), forward y-FFT
//   phase 4 modelling output: the receiver pair's spectra
template <int NY, int BY>
__global__ __launch_bounds__(512) void ewem3d_rows(March3DArgs ap, March3DArgs as, Ewem3DArgs e, int F, int phase)
{
    constexpr int P = NY / 64;
    constexpr int L = LdsLen<NY>::value;
    extern __shared__ float2 smem[];
    const March3DArgs& a = ap;
    const int ngroups = (a.nitem + F - 1) / F;
    const int g = blockIdx.x % ngroups, x = blockIdx.x / ngroups;
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int item = g * F + w;
    const bool active = item < a.nitem;
    const int fl = a.f0 + (active ? item : 0);
    const double wd = (double)(a.ifmin + fl) * a.dw;
    float2* buf = smem + w * L;
    const float tx = a.taper_x[x];
    auto load_row = [&](int s, float2 (&v)[P]) {
        const float2* src = eslot(a, e, item, s) + (size_t)x * a.nky;
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = yin<BY>(a, t + 64 * q) ? src[t + 64 * q] : make_float2(0.f, 0.f);
    };
    auto store_row = [&](int s, float2 (&v)[P]) {
        wave_sync_lds();
        yfft<NY, -1, BY>(v, buf, a, t);
        float2* dst = eslot(a, e, item, s) + (size_t)x * a.nky;
#pragma unroll
        for (int q = 0; q < P; ++q) if (yin<BY>(a, t + 64 * q)) dst[t + 64 * q] = v[q];
    };
    if (phase == 0) {
        const bool do_s = active && e.ds, do_r = active && e.dr;
        float2 vs[P], v1[P], v2[P];
#pragma unroll
        for (int q = 0; q < P; ++q) vs[q] = v1[q] = v2[q] = make_float2(0.f, 0.f);
        if (do_s) load_row(0, vs);
        if (do_r) {
            load_row(1, v1);
            load_row(2, v2);
        }
        if (do_s) {
            yfft<NY, +1, BY>(vs, buf, a, t);
            wave_sync_lds();
        }
        if (do_r) {
            yfft<NY, +1, BY>(v1, buf, a, t);
            wave_sync_lds();
            yfft<NY, +1, BY>(v2, buf, a, t);
        }
        float cp[P], cs[P];
        const float factor = a.ifmin + fl == 0 ? 1.f : 2.f;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int y = t + 64 * q;
            cp[q] = cs[q] = 0.f;
            if (y >= a.nmy) {
                vs[q] = v1[q] = v2[q] = make_float2(0.f, 0.f);
                continue;
            }
            const size_t ixy = (size_t)x * a.nmy + y;
            const float tp = tx * a.taper_y[y];
            if (do_s) vs[q] = cscale(cmulc(vs[q], ss_phase(ap, wd, ixy)), tp);
            if (do_r) {
                v1[q] = cscale(cmul(v1[q], ss_phase(ap, wd, ixy)), tp);
                v2[q] = cscale(cmul(v2[q], ss_phase(as, wd, ixy)), tp);
                if (do_s) {
                    const float2 sm = cscale(vs[q], e.inv_maxs);
                    cp[q] = factor * (sm.x * v1[q].x + sm.y * v1[q].y);
                    cs[q] = factor * (sm.x * v2[q].x + sm.y * v2[q].y);
                }
            }
        }
        // the two images of the F frequencies, each summed in a fixed order
        for (int im = 0; im < 2; ++im) {
            wave_sync_lds();
            float* red = reinterpret_cast<float*>(buf);
#pragma unroll
            for (int q = 0; q < P; ++q) red[t + 64 * q] = im == 0 ? cp[q] : cs[q];
            __syncthreads();
            float* part = (im == 0 ? e.ppart : e.spart) + (size_t)g * a.nm + (size_t)x * a.nmy;
            for (int y = threadIdx.x; y < a.nmy; y += blockDim.x) {
                float sum = 0.f;
                for (int ff = 0; ff < F; ++ff) sum += reinterpret_cast<const float*>(smem + ff * L)[y];
                part[y] = sum;
            }
            __syncthreads();
        }
        if (do_s && e.next_s) store_row(0, vs);
        if (do_r) {
            store_row(1, v1);
            store_row(2, v2);
        }
        return;
    }
    if (!active) return;  // phases 1-4 have no workgroup barriers
    if (phase == 2) {
        float2 vs[P];
        if (e.ds) {
            load_row(0, vs);
            yfft<NY, +1, BY>(vs, buf, a, t);
        }
        float2* sm = e.smig + ((size_t)item * a.nz + a.iz) * a.nm + (size_t)x * a.nmy;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int y = t + 64 * q;
            if (y >= a.nmy) {
                vs[q] = make_float2(0.f, 0.f);
                continue;
            }
            if (e.ds) vs[q] = cscale(cmulc(vs[q], ss_phase(ap, wd, (size_t)x * a.nmy + y)), tx * a.taper_y[y]);
            sm[y] = e.ds ? cscale(vs[q], e.inv_maxs) : make_float2(0.f, 0.f);
        }
        if (e.ds && e.next_s) store_row(0, vs);
        return;
    }
    float2 v1[P], v2[P];
    load_row(1, v1);
    load_row(2, v2);
    yfft<NY, +1, BY>(v1, buf, a, t);
    wave_sync_lds();
    yfft<NY, +1, BY>(v2, buf, a, t);
    if (phase == 4) {
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int y = t + 64 * q;
            if (y < a.nmy) {
                const size_t o = ((size_t)x * a.nmy + y) * a.nf + fl;
                e.dgx[o] = v1[q];
                e.dgz[o] = v2[q];
            }
        }
        return;
    }
    const float2* sm = e.smig + ((size_t)item * a.nz + a.iz) * a.nm + (size_t)x * a.nmy;
    const size_t mo = (size_t)a.iz * a.nm + (size_t)x * a.nmy;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int y = t + 64 * q;
        if (y >= a.nmy) {
            v1[q] = v2[q] = make_float2(0.f, 0.f);
            continue;
        }
        if (phase == 3) {
            const size_t ixy = (size_t)x * a.nmy + y;
            const float tp = tx * a.taper_y[y];
            const float2 u1 = cadd(v1[q], cscale(sm[y], e.mppT[mo + y]));
            const float2 u2 = cadd(v2[q], cscale(sm[y], e.mpsT[mo + y]));
            v1[q] = cmulc(cscale(u1, tp), ss_phase(ap, wd, ixy));
            v2[q] = cmulc(cscale(u2, tp), ss_phase(as, wd, ixy));
        }
    }
    store_row(1, v1);
    store_row(2, v2);
}

template <int N, int BL>
static int launch_ewem3d_n(int which, const March3DArgs& ap, const March3DArgs& as, const Ewem3DArgs& e, int F,
                           int phase, hipStream_t st)
{
    constexpr int L = LdsLen<N>::value;
    if (which == 0) {
        const size_t rows = (size_t)ap.nitem * 3 * ap.nmx;
        hipLaunchKernelGGL((ewem3d_init<N, BL>), dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, st, ap, e);
        SWEM_HIP(hipGetLastError());
        return SWEM_OK;
    }
    if (which == 1) {
        const bool src = e.mode == 3, maps = e.mode != 2 && !src;  // the source's kz; the 2 x 2 maps
        // the maps' kernel holds both wavefields through two transforms each: at N = 512 it runs
        // 4 columns per workgroup at 3 waves per SIMD rather than spill at 4
        constexpr int C2 = N >= 1024 ? 4 : 8, C1 = 2 * C2, CM = N >= 512 ? 4 : 8;
        const int C = src ? C1 : (maps ? CM : C2), nwf = src ? 1 : 2;
        const size_t lds = (size_t)C * nwf * L * sizeof(float2);
        const dim3 grid(ap.nitem * ((ap.nky + C - 1) / C)), block(64 * C);
        auto k = src ? ewem3d_cols<N, 1, C1, BL, false>
                     : (maps ? ewem3d_cols<N, 2, CM, BL, true> : ewem3d_cols<N, 2, C2, BL, false>);
        if (lds > 48 * 1024) SWEM_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k, grid, block, lds, st, ap, as, e);
        SWEM_HIP(hipGetLastError());
        return SWEM_OK;
    }
    const int ngroups = (ap.nitem + F - 1) / F;
    const size_t lds = (size_t)F * L * sizeof(float2);
    if (lds > 48 * 1024)
        SWEM_HIP(hipFuncSetAttribute((const void*)ewem3d_rows<N, BL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((ewem3d_rows<N, BL>), dim3(ngroups * ap.nmx), dim3(64 * F), lds, st, ap, as, e, F, phase);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

// Columns (waves) per x-pass workgroup: 16 buffers of staging per workgroup (8 columns of two
// wavefields or 16 of one) below nkx = 1024, half that at 1024.
static int march3d_cols_tile(int nwf, int nkx) { return (nkx >= 1024 ? 8 : 16) / nwf; }

template <typename K, typename... X>
static int launch3(K kernel, dim3 grid, dim3 block, size_t lds, hipStream_t st, const March3DArgs& a, X... extra)
{
    if (lds > 48 * 1024)
        SWEM_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kernel, grid, block, lds, st, a, extra...);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

template <int N, int BL>
static int launch3_n(int which, const March3DArgs& a, int F, int phase, hipStream_t st)
{
    constexpr int L = LdsLen<N>::value;
    if (which == 0) {  // init
        const size_t rows = (size_t)a.nitem * 2 * a.nmx;
        hipLaunchKernelGGL((march3d_init<N, BL>), dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, st, a);
        SWEM_HIP(hipGetLastError());
        return SWEM_OK;
    }
    if (which == 1) {  // cols
        const int nwf = (a.wmask & 1) + ((a.wmask >> 1) & 1);
        const int C = march3d_cols_tile(nwf, N);
        const size_t lds = (size_t)C * nwf * L * sizeof(float2);
        const dim3 grid(a.nitem * ((a.nky + C - 1) / C)), block(64 * C);
        constexpr int C2 = N >= 1024 ? 4 : 8, C1 = 2 * C2;
        return nwf == 2 ? launch3(march3d_cols<N, 2, C2, BL>, grid, block, lds, st, a)
                        : launch3(march3d_cols<N, 1, C1, BL>, grid, block, lds, st, a);
    }
    if (which == 4) {  // PSPI x-pass
        const int nwf = (a.wmask & 1) + ((a.wmask >> 1) & 1);
        const int C = march3d_cols_tile(nwf, N);
        const size_t lds = (size_t)C * nwf * L * sizeof(float2);
        const dim3 grid(a.nitem * ((a.nky + C - 1) / C)), block(64 * C);
        constexpr int C2 = N >= 1024 ? 4 : 8, C1 = 2 * C2;
        return nwf == 2 ? launch3(march3d_cols_pspi<N, 2, C2, BL>, grid, block, lds, st, a)
                        : launch3(march3d_cols_pspi<N, 1, C1, BL>, grid, block, lds, st, a);
    }
    if (which == 6) {  // linear-mode modelling receiver x-pass (gather over vrefs)
        constexpr int CG = N >= 1024 ? 8 : 16;
        const size_t lds = (size_t)CG * L * sizeof(float2);
        const dim3 grid(a.nitem * ((a.nky + CG - 1) / CG)), block(64 * CG);
        return launch3(march3d_cols_gather<N, CG, BL>, grid, block, lds, st, a);
    }
    const int ngroups = (a.nitem + F - 1) / F;
    if (which == 5) {
        const size_t lds = (size_t)F * L * sizeof(float2);
        return launch3(march3d_rows_pspi<N, BL>, dim3(ngroups * a.nmx), dim3(64 * F), lds, st, a, F, phase);
    }
    if (which == 2) {
        const size_t lds = (size_t)F * L * sizeof(float2);
        return launch3(march3d_rows_adj<N, BL>, dim3(ngroups * a.nmx), dim3(64 * F), lds, st, a, F);
    }
    const size_t lds = (size_t)F * L * sizeof(float2);
    SWEM_HIP(hipFuncSetAttribute((const void*)march3d_rows_fwd<N, BL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds));
    hipLaunchKernelGGL((march3d_rows_fwd<N, BL>), dim3(ngroups * a.nmx), dim3(64 * F), lds, st, a, F, phase);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

// which: 0 init (y-FFT of the initial wavefields), 1 x-pass, 2 adjoint y-pass, 3 forward y-pass
// (phase 0 source, 1 receiver, 2 output).  F: frequencies per y-pass workgroup.
int launch_ewem3d(int which, const March3DArgs& ap, const March3DArgs& as, const Ewem3DArgs& e, int F, int phase,
                  hipStream_t st)
{
    const bool xpass = which == 1;
    const int n = xpass ? ap.nkx : ap.nky, m = march3d_len(n);
    if (!march3d_supported(ap.nkx, ap.nky) || ap.nmx > ap.nkx || ap.nmy > ap.nky || ap.nitem < 1 || !e.W)
        return set_error(SWEM_EINVAL, "3-D ewem: inconsistent launch (nkx=%d nky=%d)", ap.nkx, ap.nky);
    if (which == 2 && (F < 1 || F > 8))  // the y-pass holds three rows per wave: 512 threads
        return set_error(SWEM_EINVAL, "3-D ewem: %d frequencies per workgroup (1..8)", F);
    const int kind = march3d_kind(n);
    if ((kind == kLenBluestein && (xpass ? !(ap.chx && ap.bchx) : !(ap.chy && ap.bchy))) ||
        (kind == kLenMixed && (xpass ? ap.mrx.n : ap.mry.n) != n))
        return set_error(SWEM_EINVAL, "3-D ewem: transform tables missing (n=%d)", n);
    switch (kind * 16 + m / 128) {
#define SWEM_E3(M, K) \
    case K * 16 + M / 128: return launch_ewem3d_n<M, K>(which, ap, as, e, F, phase, st);
        SWEM_E3(128, 0) SWEM_E3(256, 0) SWEM_E3(512, 0) SWEM_E3(1024, 0)
        SWEM_E3(128, 1) SWEM_E3(256, 1) SWEM_E3(512, 1) SWEM_E3(1024, 1)
        SWEM_E3(128, 2) SWEM_E3(256, 2) SWEM_E3(512, 2) SWEM_E3(1024, 2)
#undef SWEM_E3
    }
    return set_error(SWEM_EUNSUPPORTED, "3-D ewem: n=%d", n);
}

int launch_march3d(int which, const March3DArgs& a, int F, int phase, hipStream_t st)
{
    const bool xpass = which == 1 || which == 4 || which == 6;
    const int n = xpass ? a.nkx : a.nky, m = march3d_len(n);
    if (!march3d_supported(a.nkx, a.nky) || a.nmx > a.nkx || a.nmy > a.nky || a.nitem < 1)
        return set_error(SWEM_EINVAL, "3-D march: inconsistent launch (nkx=%d nky=%d)", a.nkx, a.nky);
    const int kind = march3d_kind(n);
    if (kind == kLenBluestein && (xpass ? !(a.chx && a.bchx) : !(a.chy && a.bchy)))
        return set_error(SWEM_EINVAL, "3-D march: Bluestein tables missing (n=%d)", n);
    if (kind == kLenMixed && (xpass ? a.mrx.n : a.mry.n) != n)
        return set_error(SWEM_EINVAL, "3-D march: mixed-radix plan missing (n=%d)", n);
    // y-passes: up to 16 frequencies per workgroup (8 on a Bluestein y axis, whose kernels are
    // compiled for 512 threads to keep their longer transforms in registers; 8 for PSPI)
    const int fmax = which == 5 ? 8 : (!xpass && kind != kLenPow2 ? 8 : 16);
    if (((which == 2 || which == 3 || which == 5) && F > fmax) || F < 1)
        return set_error(SWEM_EINVAL, "3-D march: %d frequencies per workgroup (at most %d here)", F, fmax);
    if ((which == 4 || which == 5 || which == 6) && (!a.R || !a.sel || !a.used || a.nref < 1 || a.nref > 31))
        return set_error(SWEM_EINVAL, "3-D PSPI march: tables missing");
    if ((a.linear || which == 6) && (!a.sel2 || !a.w2))
        return set_error(SWEM_EINVAL, "3-D PSPI march: linear-mode tables missing");
    switch (kind * 16 + m / 128) {
#define SWEM_M3(M, K) \
    case K * 16 + M / 128: return launch3_n<M, K>(which, a, F, phase, st);
        SWEM_M3(128, 0) SWEM_M3(256, 0) SWEM_M3(512, 0) SWEM_M3(1024, 0)
        SWEM_M3(128, 1) SWEM_M3(256, 1) SWEM_M3(512, 1) SWEM_M3(1024, 1)
        SWEM_M3(128, 2) SWEM_M3(256, 2) SWEM_M3(512, 2) SWEM_M3(1024, 2)
#undef SWEM_M3
    }
    return set_error(SWEM_EUNSUPPORTED, "3-D march: n=%d", n);
}

}  // namespace swem
