// swem_march2d.hip -- fused 2-D depth-march kernels (split-step Fourier and PSPI).
//
// Migration (adjoint): a workgroup holds F frequencies of one shot; each frequency has a source
// and a receiver wavefield, each owned by its own subgroup (a 64-lane wave holding N/64 points
// per lane, or N/P threads holding P points), kept in registers for the whole depth march.  Each
// depth step runs the x-FFTs through LDS, multiplies by the kz phase and the split-step phase
// read from the FP64-built phase tables (swem_phase.hip), applies the taper (SSF) or the per-vref
// propagation and selection (PSPI).  The source subgroup then hands its wavefield to the receiver
// subgroup through LDS for the cross-correlation image; the image of the F frequencies is summed
// in a fixed order and written once per depth step to a partial image, which reduce_image sums
// over frequency groups and shots with compensated sums -- so a shot's image does not depend on
// what else was launched with it.
//
// The default migration kernels keep both wavefields of a frequency in one subgroup instead:
// PAIR steps them in turn (image register-local), and for PSPI compat PACKED steps both with one
// transform per vref (their real parts packed in one complex array, see adj_body_packed).
//
// Modelling (forward): each subgroup owns one frequency, marches the source down storing it per
// depth in HBM, then marches the receiver up injecting m * source before each step.
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
// A step of -dz uses the complex conjugate of the +dz phases (the evanescent factor is real).
#include <type_traits>

#include "swem_esep.hpp"
#include "swem_fft.hpp"
#include "swem_internal.hpp"

namespace swem {

template <int N>
struct Pts {
    static constexpr int value = N >= 512 ? 8 : (N >= 256 ? 4 : 2);
};

constexpr int kModeSSF = 0, kModePspiCompat = 1, kModePspiLinear = 2;
// Subgroup kernels: 8 points per thread, up to 1024 threads (128 VGPRs, four waves per SIMD),
// or 16 points per thread, up to 512 threads (256 VGPRs): radix-16 passes halve the LDS round
// trips.  Wave kernels: 256 threads, two waves per SIMD.
constexpr int kMaxBlock = 1024;
constexpr int kWaveBlock = 256;

// Block index -> (frequency group, shot).  Workgroups are dispatched to the 8 XCDs round
// robin, so block b runs on XCD b % 8; consecutive blocks of one XCD take the shots of one
// frequency group, which then read the same phase-table rows from that XCD's L2.
__device__ __forceinline__ void decode_block(int ngroups, int nshot, int& g, int& shot)
{
    const int b = blockIdx.x, x = b & 7, m = b >> 3;
    shot = m % nshot;
    g = (m / nshot) * 8 + x;
    if (g >= ngroups) g = -1;
}

// Raw buffer access: a resource of `bytes` bytes at `base` (wave-uniform); loads beyond it
// return zero and stores beyond it are dropped by the hardware, so edge lanes need no branch.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float2 buf_load2(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
__device__ __forceinline__ float buf_load1(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ __forceinline__ void buf_store1(__amdgpu_buffer_rsrc_t r, uint32_t off, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, 0);
}
// Partial-image rows are written once and read once by the reduction: stored non-temporal (cache
// policy NT), so they stream through L2 instead of evicting the phase-table rows that the shots
// of one frequency share there.
constexpr int kCachePolicyNT = 2;
__device__ __forceinline__ void buf_store1_nt(__amdgpu_buffer_rsrc_t r, uint32_t off, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, kCachePolicyNT);
}
__device__ __forceinline__ void buf_store2(__amdgpu_buffer_rsrc_t r, uint32_t off, float2 v)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r,
                                          (int)off, 0, 0);
}

// kz phase of the point k = t + q T of the strided mapping from a table row (L is even in kx,
// the row holds kk = 0..N/2): kk = k for q < P/2, N - k = (P - q) T - t above.
template <int N, int P>
__device__ __forceinline__ float2 kload(const float2* __restrict__ row, int t, int q)
{
    constexpr int T = N / P;
    return row[q < P / 2 ? t + q * T : (P - q) * T - t];
}

__device__ __forceinline__ float2 cmul_c(float2 x, float2 L, bool conj) { return conj ? cmulc(x, L) : cmul(x, L); }

// A table row of the strided mapping into registers with every load issued before the first
// use (under register pressure the scheduler otherwise pairs each load with its use and
// serialises the latencies).
template <int N, int P>
__device__ __forceinline__ void kload_row(float2 (&L)[P], const float2* __restrict__ row, int t)
{
#pragma unroll
    for (int q = 0; q < P; ++q) L[q] = kload<N, P>(row, t, q);
    __builtin_amdgcn_sched_barrier(0);
}

// One x-transform of the wavefield: a wave-level transform (WAVE, P = N/64, no barriers) or a
// subgroup transform through ping-pong buffers.
template <int N, int P, int SIGN, bool WAVE, bool PRE = false, bool HIN = false>
__device__ __forceinline__ void xfft(float2 (&v)[P], float2* lds, const float2* __restrict__ tw, int t, int& pp)
{
    if constexpr (WAVE) {
        static_assert(P * 64 == N, "wave transforms hold N/64 points per lane");
        wfft<N, SIGN>(v, lds, tw, t);
    } else
        fft<N, P, SIGN, PRE, HIN>(v, lds, tw, t, pp);
}

// Transform policies of the step functions.  XPow2: the transform length is the register array
// length N (a power of two); PRE: every twiddle seed of a transform loaded at its start, two
// registers per seed held through the first pass (measured at nkx 2048, 64 shots: 6-11 % faster
// than per-pass seeds in every 2-D mode, migration and modelling).  XBlue: a transform of runtime length n = nkx (any length) through
// two power-of-two transforms of the array length M >= 2 n - 1 (Bluestein): with
// w_j = exp(SIGN i pi j^2 / n),  X_k = w_k sum_j (x_j w_j) conj(w_(k-j)),  the sum a cyclic
// convolution of length M taken as IFFT_M(FFT_M(x w) B) with B = FFT_M(conj w, wrapped) / M
// built in FP64.  Entries j, k >= n of the array are zero on input and on output.  The PSPI
// selection tables keep the array length as their row stride (0x7f padding beyond nmx) in both.
template <int N, int P, bool PRE = true>
struct XPow2 {
    // HIN: slots q >= P/2 are zero on input (see fft)
    template <int SIGN, bool WAVE, bool HIN = false>
    __device__ __forceinline__ static void fft(float2 (&v)[P], float2* lds, const March2DArgs& a, int t, int& pp)
    {
        xfft<N, P, SIGN, WAVE, PRE, HIN>(v, lds, a.tw, t, pp);
    }
    __device__ __forceinline__ static void kload_row(float2 (&L)[P], const float2* __restrict__ row, const March2DArgs&, int t)
    {
        swem::kload_row<N, P>(L, row, t);
    }
};

template <int M, int P>
struct XBlue {
    static constexpr int T = M / P;
    template <int SIGN, bool WAVE, bool HIN = false>
    __device__ __forceinline__ static void fft(float2 (&v)[P], float2* lds, const March2DArgs& a, int t, int& pp)
    {
        static_assert(!WAVE, "Bluestein transforms use the subgroup FFT");
        chirp_mul<SIGN>(v, a, t);
        swem::fft<M, P, -1>(v, lds, a.tw, t, pp);
        const float2* B = a.bchirp + (SIGN < 0 ? 0 : M);
        {
            float2 bq[P];
#pragma unroll
            for (int q = 0; q < P; ++q) bq[q] = B[t + q * T];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < P; ++q) v[q] = cmul(v[q], bq[q]);
        }
        swem::fft<M, P, +1>(v, lds, a.tw, t, pp);
        chirp_mul<SIGN>(v, a, t);
    }
    // v_j <- v_j w_j for j < n, 0 beyond (the chirp is reloaded rather than kept in registers)
    template <int SIGN>
    __device__ __forceinline__ static void chirp_mul(float2 (&v)[P], const March2DArgs& a, int t)
    {
        const int n = a.nkx;
        float2 w[P];
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int j = t + q * T;
            const float2 c = j < n ? a.chirp[j] : make_float2(0.f, 0.f);  // exp(-i pi j^2 / n)
            w[q] = SIGN < 0 ? c : make_float2(c.x, -c.y);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = cmul(v[q], w[q]);
    }
    __device__ __forceinline__ static void kload_row(float2 (&L)[P], const float2* __restrict__ row, const March2DArgs& a, int t)
    {
        const int n = a.nkx;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int k = t + q * T;
            L[q] = row[k < n ? (k <= n / 2 ? k : n - k) : 0];
        }
        __builtin_amdgcn_sched_barrier(0);
    }
};

// XMixed: a mixed-radix length n = nkx (2^a 3^b 5^c) in the power-of-two array of M >= n points,
// transformed in the subgroup's two LDS buffers (mixed_fft); entries j >= n stay zero.
template <int M, int P>
struct XMixed {
    template <int SIGN, bool WAVE, bool HIN = false>
    __device__ __forceinline__ static void fft(float2 (&v)[P], float2* lds, const March2DArgs& a, int t, int& pp)
    {
        static_assert(!WAVE, "mixed-radix transforms use the subgroup buffers");
        mixed_fft<M, P, SIGN, true>(v, lds, lds + LdsLen<M>::value, a.tw, a.mr, t, pp, []() { __syncthreads(); });
    }
    __device__ __forceinline__ static void kload_row(float2 (&L)[P], const float2* __restrict__ row, const March2DArgs& a, int t)
    {
        XBlue<M, P>::kload_row(L, row, a, t);
    }
};

struct StepCtx {
    int t, iz, fc;  // thread within the subgroup, depth index, frequency index within the chunk
                    // (fc and iz are wave-uniform: table rows are addressed from scalar registers)
};

__device__ __forceinline__ const float2* krow(const March2DArgs& a, const StepCtx& c, int ir)
{
    return a.ktab + (((size_t)c.fc * a.nz + c.iz) * a.nref + ir) * a.nkh;
}

// Split-step phase row of (fc, iz) as a buffer of nmx entries: lanes x >= nmx read zero.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t ss_row(const March2DArgs& a, const StepCtx& c)
{
    return buf_rsrc(a.stab + ((size_t)c.fc * a.nz + c.iz) * a.nmx, (uint32_t)a.nmx * 8u);
}

// ssop adjoint branch (wem.c:313-320, 334-359) on one wavefield.  ACT = false turns the step
// into FFT / IFFT with unit phases (a numerical identity) so every subgroup of a workgroup runs
// the same barriers when only one of the two wavefields is stepped at a depth.
template <int N, int P, bool WAVE, bool ACT, class X = XPow2<N, P>>
__device__ __forceinline__ void ssf_adj_step(float2 (&u)[P], float2* lds, int& pp, const March2DArgs& a,
                                             const StepCtx& c, bool conj, const float (&tap)[P])
{
    constexpr int T = N / P;
    X::template fft<-1, WAVE>(u, lds, a, c.t, pp);
    if constexpr (ACT) {
        float2 L[P];
        X::kload_row(L, krow(a, c, 0), a, c.t);
#pragma unroll
        for (int q = 0; q < P; ++q) u[q] = cmul_c(u[q], L[q], conj);
    } else {
#pragma unroll
        for (int q = 0; q < P; ++q) u[q] = cscale(u[q], a.inv_nk);
    }
    X::template fft<+1, WAVE>(u, lds, a, c.t, pp);
    if constexpr (ACT) {
        const auto sr = ss_row(a, c);
        float2 sp[P];
#pragma unroll
        for (int q = 0; q < P; ++q) sp[q] = buf_load2(sr, (uint32_t)(c.t + q * T) * 8u);
        __builtin_amdgcn_sched_barrier(0);  // all loads in flight before the first use
#pragma unroll
        for (int q = 0; q < P; ++q)  // zero beyond nmx: the split-step row reads zero there
            u[q] = cscale(cmul_c(u[q], sp[q], conj), tap[q]);
    } else {
#pragma unroll
        for (int q = 0; q < P; ++q)
            if (c.t + q * T >= a.nmx) u[q] = make_float2(0.f, 0.f);
    }
}

// ssop forward branch with dz < 0 (wem.c:321-332, 334-348, 362-369): taper, split-step
// phase, FFT, kz phase, IFFT.
template <int N, int P, bool WAVE, class X = XPow2<N, P>>
__device__ __forceinline__ void ssf_fwd_step(float2 (&u)[P], float2* lds, int& pp, const March2DArgs& a,
                                             const StepCtx& c, const float (&tap)[P])
{
    constexpr int T = N / P;
    const auto sr = ss_row(a, c);
    {
        float2 sp[P];
#pragma unroll
        for (int q = 0; q < P; ++q) sp[q] = buf_load2(sr, (uint32_t)(c.t + q * T) * 8u);
        __builtin_amdgcn_sched_barrier(0);  // all loads in flight before the first use
#pragma unroll
        for (int q = 0; q < P; ++q) u[q] = cmulc(cscale(u[q], tap[q]), sp[q]);
    }
    X::template fft<-1, WAVE>(u, lds, a, c.t, pp);
    {
        float2 L[P];
        X::kload_row(L, krow(a, c, 0), a, c.t);
#pragma unroll
        for (int q = 0; q < P; ++q) u[q] = cmulc(u[q], L[q]);
    }
    X::template fft<+1, WAVE>(u, lds, a, c.t, pp);
#pragma unroll
    for (int q = 0; q < P; ++q)
        if (c.t + q * T >= a.nmx) u[q] = make_float2(0.f, 0.f);
}

// Weight of reference velocity ir at a point with selection byte s1 (iref1 | cut<<7), iref2
// s2 and interpolation weight w2 (linear mode: the formula at wem.c:476).
__device__ __forceinline__ float pspi_linear_weight(int ir, uint8_t s1, uint8_t s2, float w2)
{
    const int i1 = s1 & 0x7f;
    if (!(s1 >> 7)) return i1 == ir ? 1.f : 0.f;
    float wt = 0.f;
    if (i1 == ir) wt += 1.f - w2;
    if ((int)s2 == ir) wt += w2;
    return wt;
}

// pspiop with the reference structure (wem.c:409-460): FFT, for each vref phase + IFFT, then
// per-x selection.  Compat: u = Re(dref[iref1]) when cut else dref[iref1]; linear: weights.
// The selection rows are padded to N columns (0x7f beyond nmx: no vref selected).
template <int N, int P, int MODE, bool WAVE, bool ACT, class X = XPow2<N, P>>
__device__ __forceinline__ void pspi_select_step(float2 (&u)[P], float2* lds, int& pp, const March2DArgs& a,
                                                 const StepCtx& c, bool conj)
{
    constexpr int T = N / P;
    X::template fft<-1, WAVE>(u, lds, a, c.t, pp);
    float2 U[P];
    uint8_t s1[P], s2[P];
    float w2[P];
    const size_t rowx = (size_t)c.iz * N;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = c.t + q * T;
        s1[q] = a.sel[rowx + x];
        if constexpr (MODE == kModePspiLinear) {
            s2[q] = a.sel2[rowx + x];
            w2[q] = a.w2[rowx + x];
        }
        U[q] = make_float2(0.f, 0.f);
    }
    const uint32_t used = a.used[c.iz];
    const int first = __builtin_ctz(used | 0x80000000u);
    for (int ir = 0; ir < a.nref; ++ir) {
        if (!((used >> ir) & 1u)) continue;
        float2 B[P];
        if constexpr (ACT) {
            X::kload_row(B, krow(a, c, ir), a, c.t);
#pragma unroll
            for (int q = 0; q < P; ++q) B[q] = cmul_c(u[q], B[q], conj);
        } else {
#pragma unroll
            for (int q = 0; q < P; ++q) B[q] = cscale(u[q], a.inv_nk);
        }
        X::template fft<+1, WAVE>(B, lds, a, c.t, pp);
#pragma unroll
        for (int q = 0; q < P; ++q) {
            if constexpr (!ACT) {
                if (ir == first) U[q] = B[q];
            } else if constexpr (MODE == kModePspiCompat) {
                if ((s1[q] & 0x7f) == ir) U[q] = (s1[q] >> 7) ? make_float2(B[q].x, 0.f) : B[q];
            } else {
                U[q] = cadd(U[q], cscale(B[q], pspi_linear_weight(ir, s1[q], s2[q], w2[q])));
            }
        }
    }
#pragma unroll
    for (int q = 0; q < P; ++q) u[q] = (c.t + q * T < a.nmx) ? U[q] : make_float2(0.f, 0.f);
}

// Exact adjoint of the linear select-after step, for the forward receiver (dz < 0):
// u <- IFFT( sum_ir conj(L_ir) FFT(W_ir u) ).
template <int N, int P, bool WAVE, class X = XPow2<N, P>>
__device__ __forceinline__ void pspi_weight_step(float2 (&u)[P], float2* lds, int& pp, const March2DArgs& a,
                                                 const StepCtx& c)
{
    constexpr int T = N / P;
    float2 acc[P];
    uint8_t s1[P], s2[P];
    float w2[P];
    const size_t rowx = (size_t)c.iz * N;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = c.t + q * T;
        s1[q] = a.sel[rowx + x];
        s2[q] = a.sel2[rowx + x];
        w2[q] = a.w2[rowx + x];
        acc[q] = make_float2(0.f, 0.f);
    }
    const uint32_t used = a.used[c.iz];
    for (int ir = 0; ir < a.nref; ++ir) {
        if (!((used >> ir) & 1u)) continue;
        float2 B[P];
#pragma unroll
        for (int q = 0; q < P; ++q) B[q] = cscale(u[q], pspi_linear_weight(ir, s1[q], s2[q], w2[q]));
        X::template fft<-1, WAVE>(B, lds, a, c.t, pp);
        const float2* row = krow(a, c, ir);
        float2 L[P];
        X::kload_row(L, row, a, c.t);
#pragma unroll
        for (int q = 0; q < P; ++q) acc[q] = cadd(acc[q], cmulc(B[q], L[q]));
    }
    X::template fft<+1, WAVE>(acc, lds, a, c.t, pp);
#pragma unroll
    for (int q = 0; q < P; ++q) u[q] = (c.t + q * T < a.nmx) ? acc[q] : make_float2(0.f, 0.f);
}

// Source / adjoint-receiver step of either method (select-after for PSPI).  act is uniform
// over the subgroup: the branch is taken once per step, not per element.
template <int N, int P, int MODE, bool WAVE, class X = XPow2<N, P>>
__device__ __forceinline__ void adj_step(float2 (&u)[P], float2* lds, int& pp, const March2DArgs& a, const StepCtx& c,
                                         bool act, bool conj, const float (&tap)[P])
{
    if constexpr (MODE == kModeSSF) {
        if (act)
            ssf_adj_step<N, P, WAVE, true, X>(u, lds, pp, a, c, conj, tap);
        else
            ssf_adj_step<N, P, WAVE, false, X>(u, lds, pp, a, c, conj, tap);
    } else {
        if (act)
            pspi_select_step<N, P, MODE, WAVE, true, X>(u, lds, pp, a, c, conj);
        else
            pspi_select_step<N, P, MODE, WAVE, false, X>(u, lds, pp, a, c, conj);
    }
}

// Forward-receiver step (dz < 0).
template <int N, int P, int MODE, bool WAVE, class X = XPow2<N, P>>
__device__ __forceinline__ void fwd_step(float2 (&u)[P], float2* lds, int& pp, const March2DArgs& a, const StepCtx& c,
                                         const float (&tap)[P])
{
    if constexpr (MODE == kModeSSF)
        ssf_fwd_step<N, P, WAVE, X>(u, lds, pp, a, c, tap);
    else if constexpr (MODE == kModePspiCompat)
        pspi_select_step<N, P, MODE, WAVE, true, X>(u, lds, pp, a, c, true);
    else
        pspi_weight_step<N, P, WAVE, X>(u, lds, pp, a, c);
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

// Migration march shared by both kernel families.  The workgroup has 2F subgroups of T
// threads; subgroup s = 2f is the source and 2f + 1 the receiver of frequency f.  BUFS LDS
// buffers of LdsLen<N> complex per subgroup (1 for wave transforms, 2 for ping-pong).
template <int N, int P, int MODE, bool WAVE>
__device__ __forceinline__ void adj_body(const March2DArgs& a)
{
    constexpr int T = N / P;
    constexpr int L = LdsLen<N>::value;
    constexpr int BUFS = WAVE ? 1 : 2;
    extern __shared__ float2 smem[];
    int g, shot;
    decode_block(a.ngroups, a.nshot, g, shot);
    if (g < 0) return;  // padding block of the XCD-aware grid: exits before any barrier
    const int F = blockDim.x / (2 * T);
    const int tid = threadIdx.x, s = tid / T, t = tid - s * T;
    const int f = s >> 1;
    const bool src = (s & 1) == 0;
    const int fc = g * F + f;  // frequency within the chunk
    const bool active = fc < a.nfc;
    const int fl = a.f0 + fc;  // frequency within the band
    const int iw = a.ifmin + fl;
    StepCtx c;
    c.t = t;
    c.fc = __builtin_amdgcn_readfirstlane(active ? fc : 0);
    const float factor = iw == 0 ? 1.f : 2.f;
    float2* lds = smem + (size_t)s * BUFS * L;
    const float2* xsrc = smem + (size_t)(2 * f) * BUFS * L;
    int pp = 0;

    float2 u[P];
    float tap[P];
    load_taper<N, P, MODE>(tap, a, t);
    const int igx = a.igx[shot];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = t + q * T;
        if (src)
            u[q] = (active && x == igx) ? a.spec_s[fl] : make_float2(0.f, 0.f);
        else
            u[q] = (active && x < a.nmx) ? a.spec_r[((size_t)shot * a.nmx + x) * a.nf + fl] : make_float2(0.f, 0.f);
    }
    float* part = a.partial + ((size_t)shot * a.ngroups + g) * (size_t)a.nz * a.nmx;

    for (int iz = 0; iz < a.nz; ++iz) {
        c.iz = iz;
        const float z = a.oz + a.dz * (float)iz;
        const bool ds = z >= a.sz, dr = z >= a.gz;
        if (ds || dr) adj_step<N, P, MODE, WAVE>(u, lds, pp, a, c, src ? ds : dr, src, tap);
        // imaging condition (wem.c:243-247): the source subgroup hands its wavefield over in LDS
        const int xb = pp * L;
        if (src) {
#pragma unroll
            for (int q = 0; q < P; ++q) lds[xb + lds_pad(t + q * T)] = u[q];
        }
        __syncthreads();
        float cc[P];
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const float2 us = xsrc[xb + lds_pad(t + q * T)];
            cc[q] = dr ? factor * (u[q].x * us.x + u[q].y * us.y) : 0.f;
        }
        if constexpr (!WAVE) pp ^= 1;
        if (F == 1) {
            if (!src) {
                const auto pr = buf_rsrc(part + (size_t)iz * a.nmx, (uint32_t)a.nmx * 4u);
#pragma unroll
                for (int q = 0; q < P; ++q) buf_store1_nt(pr, (uint32_t)(t + q * T) * 4u, cc[q]);
            }
        } else {
            if constexpr (WAVE) __syncthreads();  // single buffer: the source copy is read
            // the receiver subgroup parks its image row in its own buffer; all threads then sum
            // over the F frequencies in a fixed order
            const int rb = pp * L;
            if (!src) {
                float* red = reinterpret_cast<float*>(lds + rb);
#pragma unroll
                for (int q = 0; q < P; ++q) red[t + q * T] = cc[q];
            }
            __syncthreads();
            for (int x = tid; x < a.nmx; x += blockDim.x) {
                float sum = 0.f;
                for (int ff = 0; ff < F; ++ff)
                    sum += reinterpret_cast<const float*>(smem + (size_t)(2 * ff + 1) * BUFS * L + rb)[x];
                part[(size_t)iz * a.nmx + x] = sum;
            }
            if constexpr (!WAVE) pp ^= 1;
        }
        if constexpr (WAVE) __syncthreads();  // readers of the single buffers are done
    }
}

// Modelling march: NSUB subgroups, one frequency each, no interaction between subgroups.
template <int N, int P, int MODE, bool WAVE, class X = XPow2<N, P>>
__device__ __forceinline__ void fwd_body(const March2DArgs& a)
{
    constexpr int T = N / P;
    constexpr int L = LdsLen<N>::value;
    constexpr int BUFS = WAVE ? 1 : 2;
    extern __shared__ float2 smem[];
    int g, shot;
    decode_block(a.ngroups, a.nshot, g, shot);
    if (g < 0) return;
    const int NSUB = blockDim.x / T;
    const int tid = threadIdx.x, s = tid / T, t = tid - s * T;
    const int fc = g * NSUB + s;
    const bool active = fc < a.nfc;
    const int fl = a.f0 + fc;
    StepCtx c;
    c.t = t;
    c.fc = __builtin_amdgcn_readfirstlane(active ? fc : 0);
    float2* lds = smem + (size_t)s * BUFS * L;
    int pp = 0;
    float tap[P];
    load_taper<N, P, MODE>(tap, a, t);
    const int igx = a.igx[shot];
    float2* sm = a.smig + ((size_t)shot * a.nfc + (active ? fc : 0)) * (size_t)a.nz * a.nmx;

    float2 u[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = t + q * T;
        u[q] = (active && x == igx) ? a.spec_s[fl] : make_float2(0.f, 0.f);
    }
    // source march, stored per depth (wem.c:254-265)
    for (int iz = 0; iz < a.nz; ++iz) {
        c.iz = iz;
        const float z = a.oz + a.dz * (float)iz;
        const bool ds = z >= a.sz;
        if (ds) adj_step<N, P, MODE, WAVE, X>(u, lds, pp, a, c, true, true, tap);
        if (active) {
            const auto sr = buf_rsrc(sm + (size_t)iz * a.nmx, (uint32_t)a.nmx * 8u);
#pragma unroll
            for (int q = 0; q < P; ++q)
                buf_store2(sr, (uint32_t)(t + q * T) * 8u, ds ? u[q] : make_float2(0.f, 0.f));
        }
    }
    // receiver march upward with injection (wem.c:266-276)
#pragma unroll
    for (int q = 0; q < P; ++q) u[q] = make_float2(0.f, 0.f);
    for (int iz = a.nz - 1; iz >= 0; --iz) {
        c.iz = iz;
        const float z = a.oz + a.dz * (float)iz;
        if (!(z >= a.gz)) continue;
        if (active) {
            const auto sr = buf_rsrc(sm + (size_t)iz * a.nmx, (uint32_t)a.nmx * 8u);
            const auto mr = buf_rsrc(a.mT + (size_t)iz * a.nmx, (uint32_t)a.nmx * 4u);
            float2 sv[P];
            float mv[P];
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const uint32_t x = (uint32_t)(t + q * T);
                sv[q] = buf_load2(sr, x * 8u);
                mv[q] = buf_load1(mr, x * 4u);
            }
            __builtin_amdgcn_sched_barrier(0);  // all loads in flight before the first use
#pragma unroll
            for (int q = 0; q < P; ++q) u[q] = cadd(u[q], cscale(sv[q], mv[q]));
        }
        fwd_step<N, P, MODE, WAVE, X>(u, lds, pp, a, c, tap);
    }
    if (active) {
        const auto dr = buf_rsrc(a.dgw + (size_t)shot * a.nmx * a.nf, (uint32_t)a.nmx * a.nf * 8u);
#pragma unroll
        for (int q = 0; q < P; ++q) buf_store2(dr, ((uint32_t)(t + q * T) * (uint32_t)a.nf + (uint32_t)fl) * 8u, u[q]);
    }
}

// Migration march with one subgroup per frequency that steps the source, then the receiver,
// through the same LDS buffers (variant PAIR).  One frequency per workgroup; the imaging
// condition is register-local.
template <int N, int P, int MODE, bool WAVE, class X = XPow2<N, P>>
__device__ __forceinline__ void adj_body_pair(const March2DArgs& a)
{
    constexpr int T = N / P;
    extern __shared__ float2 smem[];
    int g, shot;
    decode_block(a.ngroups, a.nshot, g, shot);
    if (g < 0) return;  // padding block of the XCD-aware grid
    const int t = threadIdx.x;
    const int fc = g;  // one frequency per workgroup
    const bool active = fc < a.nfc;
    const int fl = a.f0 + fc;
    const int iw = a.ifmin + fl;
    StepCtx c;
    c.t = t;
    c.fc = __builtin_amdgcn_readfirstlane(active ? fc : 0);
    const float factor = iw == 0 ? 1.f : 2.f;
    int pp = 0;
    float tap[P];
    load_taper<N, P, MODE>(tap, a, t);
    const int igx = a.igx[shot];
    float2 us[P], ur[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = t + q * T;
        us[q] = (active && x == igx) ? a.spec_s[fl] : make_float2(0.f, 0.f);
        ur[q] = (active && x < a.nmx) ? a.spec_r[((size_t)shot * a.nmx + x) * a.nf + fl] : make_float2(0.f, 0.f);
    }
    float* part = a.partial + ((size_t)shot * a.ngroups + g) * (size_t)a.nz * a.nmx;
    for (int iz = 0; iz < a.nz; ++iz) {
        c.iz = iz;
        const float z = a.oz + a.dz * (float)iz;
        const bool ds = z >= a.sz, dr = z >= a.gz;
        if (ds) adj_step<N, P, MODE, WAVE, X>(us, smem, pp, a, c, true, true, tap);
        if (dr) adj_step<N, P, MODE, WAVE, X>(ur, smem, pp, a, c, true, false, tap);
        // imaging condition (wem.c:243-247)
        const auto pr = buf_rsrc(part + (size_t)iz * a.nmx, (uint32_t)a.nmx * 4u);
#pragma unroll
        for (int q = 0; q < P; ++q)
            buf_store1_nt(pr, (uint32_t)(t + q * T) * 4u, dr ? factor * (ur[q].x * us[q].x + ur[q].y * us[q].y) : 0.f);
    }
}

// ---------------------------------------------------------------- packed PSPI compat migration
// In compat mode pspiop keeps only the real part of its result wherever vref2 - vref1 > 10
// (wem.c:450-458 with linear_interp, wem.c:473-478), so after a step both wavefields are real
// except at the few points of each depth with iref1 == iref2 (v at the depth's maximum, or every
// point of a depth with vmax <= vmin + 10).  With s = s_R + i s_I, r = r_R + i r_I the state is
//   z = s_R + i a r_R        e = s_I - i a r_I          (a: balances the two magnitudes)
// and, because L_ir(k) is even in kx and FFTs of real sequences are Hermitian,
//   Re P_src(s) + i a Re P_rec(r) = IFFT( Re L Z + Im L E )                        (all x)
//   Im P_src(s) - i a Im P_rec(r) = IFFT( Re L E - Im L Z )                        (all x)
// with Z = FFT(z), E = FFT(e), P_src using conj(L) (-dz) and P_rec using L.  One inverse
// transform per vref thus steps both wavefields; one more, with the vref of the uncut points,
// gives their new e: 2 + nref + 1 transforms per depth instead of 2 (1 + nref).  The image is
// factor Re(r conj s) = factor (Re z Im z - Re e Im e) / a.

// Sum over the T threads of a subgroup in a fixed order (red: T/64 floats of LDS scratch).
template <int T>
__device__ __forceinline__ float subgroup_sum(float v, float* red)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if constexpr (T == 64) {
        return v;
    } else {
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        float s = 0.f;
#pragma unroll
        for (int i = 0; i < T / 64; ++i) s += red[i];
        __syncthreads();
        return s;
    }
}


template <int T>
__device__ __forceinline__ void subgroup_sync()
{
    if constexpr (T == 64) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// HALF: nmx <= N / 2, so the points q >= P / 2 (x >= N / 2) are zero in the state and never selected:
// those constants let the first forward pass and the last inverse pass drop their work.
// X: the transform policy (XPow2: nkx = N; XMixed: a mixed-radix nkx <= N, whose twiddle table
// holds nkx entries and whose spectra are zero at k >= nkx).
template <int N, int P, bool WAVE, bool HALF, class X = XPow2<N, P, true>>
__device__ __forceinline__ void adj_body_packed(const March2DArgs& a)
{
    constexpr int T = N / P;
    constexpr int K = kUncutSparse;
    constexpr int NW = T / 64;  // waves of the subgroup
    constexpr bool POW2 = std::is_same<X, XPow2<N, P, true>>::value;
    // Bluestein lengths: the table tw is of the array length, not of nkx, so no direct sums (the
    // uncut points' e and output always come from transforms)
    constexpr bool BLUE = std::is_same<X, XBlue<N, P>>::value;
    // exp(-2 pi i m / nkx) for the direct sums (m >= 0)
    auto twid = [&](int m) { return POW2 ? a.tw[m & (N - 1)] : a.tw[m % a.nkx]; };
    static_assert(4 + K + 2 * NW * K <= kPackedScratch, "packed scratch");
    extern __shared__ float2 smem[];
    int g, shot;
    decode_block(a.ngroups, a.nshot, g, shot);
    if (g < 0) return;  // padding block of the XCD-aware grid
    const int t = threadIdx.x;
    const int fc = g;  // one frequency per workgroup
    const bool active = fc < a.nfc;
    const int fl = a.f0 + fc;
    const int iw = a.ifmin + fl;
    StepCtx c;
    c.t = t;
    c.fc = __builtin_amdgcn_readfirstlane(active ? fc : 0);
    const float factor = iw == 0 ? 1.f : 2.f;
    int pp = 0;
    const int igx = a.igx[shot];
    const float2 ss = active ? a.spec_s[fl] : make_float2(0.f, 0.f);
    // scratch after the transform buffers: reduction slots, e at the previous depth's uncut
    // points, per-wave partial sums of the new e at this depth's uncut points
    float2* scr = smem + (WAVE ? 1 : 2) * LdsLen<N>::value;
    float* red = reinterpret_cast<float*>(scr);
    float2* scr_e = scr + 4;
    float2* scr_v = scr + 4 + K;
    float2 z[P], e[P];
    float nr = 0.f;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = t + q * T;
        e[q] = (active && x < a.nmx) ? a.spec_r[((size_t)shot * a.nmx + x) * a.nf + fl] : make_float2(0.f, 0.f);
        nr += e[q].x * e[q].x + e[q].y * e[q].y;
    }
    nr = subgroup_sum<T>(nr, red);
    const float ns = sqrtf(ss.x * ss.x + ss.y * ss.y);
    float* part = a.partial + ((size_t)shot * a.ngroups + g) * (size_t)a.nz * a.nmx;
    if (nr == 0.f || ns == 0.f) {
        // a zero receiver or source spectrum images to zero at every depth (the march is linear
        // in each): the packed pair would leave rounding noise of the other wavefield instead
        // (a NaN in either takes the march and reaches the image)
        for (int iz = 0; iz < a.nz; ++iz) {
            const auto pr = buf_rsrc(part + (size_t)iz * a.nmx, (uint32_t)a.nmx * 4u);
#pragma unroll
            for (int q = 0; q < P; ++q) buf_store1_nt(pr, (uint32_t)(t + q * T) * 4u, 0.f);
        }
        return;
    }
    float alpha = ns / sqrtf(nr);
    float inv_alpha = 1.f / alpha;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = t + q * T;
        const float2 us = x == igx ? ss : make_float2(0.f, 0.f);
        z[q] = make_float2(us.x, alpha * e[q].x);
        e[q] = make_float2(us.y, -alpha * e[q].y);
        if (HALF && q >= P / 2) z[q] = e[q] = make_float2(0.f, 0.f);
    }
    // e before the step: dense (the initial imaginary parts), sparse (nonzero only at the
    // listed uncut points of the previous depth: prev_cnt <= K) or zero
    int prev_cnt = -1;  // -1: dense
    const int* prev_row = a.cref;
    for (int iz = 0; iz < a.nz; ++iz) {
        c.iz = iz;
        const int* crow = a.cref + (size_t)iz * kCrefStride;
        const int cr = crow[0], cnt = crow[1];
        // HALF: slots q >= P/2 of z and e are zero (x >= N/2 >= nmx is never selected)
        X::template fft<-1, WAVE, HALF>(z, smem, a, t, pp);
        if (prev_cnt < 0 || prev_cnt > (BLUE ? 0 : K)) {
            X::template fft<-1, WAVE, HALF>(e, smem, a, t, pp);
        } else if (prev_cnt > 0) {
            // E(k) = sum_j e(x_j) w^(k x_j), w = exp(-2 pi i / N): the owners publish e(x_j)
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if (j >= prev_cnt) break;
                const int xj = prev_row[3 + j];
#pragma unroll
                for (int q = 0; q < P; ++q)
                    if (t + q * T == xj) scr_e[j] = e[q];
            }
            subgroup_sync<T>();
#pragma unroll
            for (int q = 0; q < P; ++q) e[q] = make_float2(0.f, 0.f);
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if (j >= prev_cnt) break;
                const int xj = prev_row[3 + j];
                const float2 ev = scr_e[j];
                float2 w = twid(t * xj);
                const float2 st = twid(T * xj);
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    e[q] = cadd(e[q], cmul(ev, w));
                    w = cmul(w, st);
                }
            }
            if constexpr (!POW2) {  // the spectrum of a length nkx < N ends at nkx
#pragma unroll
                for (int q = 0; q < P; ++q)
                    if (t + q * T >= a.nkx) e[q] = make_float2(0.f, 0.f);
            }
        }  // prev_cnt == 0: e == 0, E == 0
        // selection bytes packed four to a register: (sel(q) = s1w[q / 4] >> 8 (q % 4)) & 0xff
        uint32_t s1w[(P + 3) / 4];
        float2 out[P];
#pragma unroll
        for (int q = 0; q < (P + 3) / 4; ++q) s1w[q] = 0u;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            s1w[q >> 2] |= (uint32_t)a.sel[(size_t)iz * N + t + q * T] << (8 * (q & 3));
            out[q] = make_float2(0.f, 0.f);
        }
        auto s1 = [&](int q) { return (s1w[q >> 2] >> (8 * (q & 3))) & 0xffu; };
        const uint32_t used = a.used[iz];
        // Few uncut points, and no cut point selecting their vref cr: their output under cr, like
        // their new e, comes from direct sums instead of an inverse transform of cr (HALF kernels:
        // the others have no registers left for the second set of sums).
        const bool direct_cr = !BLUE && HALF && cnt > 0 && cnt <= K && crow[2] != 0;
        for (int ir = 0; ir < a.nref; ++ir) {
            if (!((used >> ir) & 1u) || (direct_cr && ir == cr)) continue;
            const float2* row = krow(a, c, ir);
            float2 w[P];
            X::kload_row(w, row, a, c.t);
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const float2 L = w[q];
                w[q] = make_float2(z[q].x * L.x + e[q].x * L.y, z[q].y * L.x + e[q].y * L.y);
            }
            X::template fft<+1, WAVE>(w, smem, a, t, pp);
#pragma unroll
            for (int q = 0; q < (HALF ? P / 2 : P); ++q)
                if ((int)(s1(q) & 0x7fu) == ir) out[q] = w[q];
        }
        if (cnt > 0) {
            // imaginary parts at the uncut points of this depth (all select vref cr):
            // e_new = IFFT(Re L E - Im L Z), the conjugate of IFFT(V) by L(-k) = L(k)
            const float2* row = krow(a, c, cr);
            float2 lv[P];
            X::kload_row(lv, row, a, c.t);
            if (BLUE || cnt > K) {
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    const float2 L = lv[q];
                    e[q] = make_float2(L.x * e[q].x - L.y * z[q].x, L.x * e[q].y - L.y * z[q].y);
                }
                X::template fft<+1, WAVE>(e, smem, a, t, pp);
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    const int x = t + q * T;
                    if ((HALF && q >= P / 2) || !(x < a.nmx && !(s1(q) >> 7))) e[q] = make_float2(0.f, 0.f);
                }
            } else {
                // e_new(x_j) = sum_k U(k) conj(w^(k x_j)) at the few uncut points and, when direct_cr,
                // out(x_j) = sum_k W(k) conj(w^(k x_j)) with W = Re L Z + Im L E: partial sums per
                // thread, then over the subgroup in a fixed order
                // U into e and W into lv, in place (the spectra z are not needed after them)
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    const float2 L = lv[q];
                    const float2 u = make_float2(L.x * e[q].x - L.y * z[q].x, L.x * e[q].y - L.y * z[q].y);
                    lv[q] = make_float2(z[q].x * L.x + e[q].x * L.y, z[q].y * L.x + e[q].y * L.y);
                    e[q] = u;
                }
                float2 acc[K], accw[K];
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    acc[j] = accw[j] = make_float2(0.f, 0.f);
                    if (j >= cnt) continue;
                    const int xj = crow[3 + j];
                    float2 w = twid(t * xj);
                    const float2 st = twid(T * xj);
#pragma unroll
                    for (int q = 0; q < P; ++q) {
                        acc[j] = cadd(acc[j], cmulc(e[q], w));
                        if (direct_cr) accw[j] = cadd(accw[j], cmulc(lv[q], w));
                        w = cmul(w, st);
                    }
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) {
                        acc[j].x += __shfl_xor(acc[j].x, o, 64);
                        acc[j].y += __shfl_xor(acc[j].y, o, 64);
                        if (direct_cr) {
                            accw[j].x += __shfl_xor(accw[j].x, o, 64);
                            accw[j].y += __shfl_xor(accw[j].y, o, 64);
                        }
                    }
                    if constexpr (NW > 1)
                        if ((t & 63) == 0) {
                            scr_v[(t >> 6) * K + j] = acc[j];
                            scr_v[(NW + (t >> 6)) * K + j] = accw[j];
                        }
                }
                if constexpr (NW > 1) {
                    subgroup_sync<T>();
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        if (j >= cnt) break;
                        acc[j] = scr_v[j];
                        accw[j] = scr_v[NW * K + j];
#pragma unroll
                        for (int wv = 1; wv < NW; ++wv) {
                            acc[j] = cadd(acc[j], scr_v[wv * K + j]);
                            accw[j] = cadd(accw[j], scr_v[(NW + wv) * K + j]);
                        }
                    }
                }
#pragma unroll
                for (int q = 0; q < P; ++q) e[q] = make_float2(0.f, 0.f);
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    if (j >= cnt) break;
                    const int xj = crow[3 + j];
#pragma unroll
                    for (int q = 0; q < (HALF ? P / 2 : P); ++q)
                        if (t + q * T == xj) {
                            e[q] = acc[j];
                            if (direct_cr) out[q] = accw[j];
                        }
                }
            }
        }
        prev_cnt = cnt;
        prev_row = crow;
        // imaging condition (wem.c:243-247): factor Re(ur conj(us)).  Without uncut points e is
        // zeroed, so the arithmetic above needs no branch.
        const auto pr = buf_rsrc(part + (size_t)iz * a.nmx, (uint32_t)a.nmx * 4u);
        float n2s = 0.f, n2r = 0.f;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            z[q] = out[q];
            if (cnt == 0 || (HALF && q >= P / 2)) e[q] = make_float2(0.f, 0.f);
            if (HALF && q >= P / 2) {
                z[q] = make_float2(0.f, 0.f);
                continue;  // x >= nmx: nothing to image
            }
            buf_store1_nt(pr, (uint32_t)(t + q * T) * 4u, factor * (z[q].x * z[q].y - e[q].x * e[q].y) * inv_alpha);
        }
        // keep the two packed wavefields of comparable size (the rounding of each transform is
        // relative to their sum): every 8 depths, rescale the receiver part when the ratio of
        // their energies leaves [1/4, 4]
        if ((iz & 7) != 7) continue;
#pragma unroll
        for (int q = 0; q < (HALF ? P / 2 : P); ++q) {
            n2s += z[q].x * z[q].x + e[q].x * e[q].x;
            n2r += z[q].y * z[q].y + e[q].y * e[q].y;
        }
        n2s = subgroup_sum<T>(n2s, red);
        n2r = subgroup_sum<T>(n2r, red);
        if (n2s > 0.f && n2r > 0.f && (n2s > 4.f * n2r || n2r > 4.f * n2s)) {
            const float beta = sqrtf(n2s / n2r);
#pragma unroll
            for (int q = 0; q < P; ++q) {
                z[q].y *= beta;
                e[q].y *= beta;
            }
            alpha *= beta;
            inv_alpha = 1.f / alpha;
        }
    }
}

template <int N, int P, bool HALF>
__global__ __launch_bounds__(N / P) __attribute__((amdgpu_waves_per_eu(2))) void march2d_adj_packed(March2DArgs a)
{
    adj_body_packed<N, P, false, HALF>(a);
}

template <int N, bool HALF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void march2d_adj_packed_w(March2DArgs a)
{
    adj_body_packed<N, N / 64, true, HALF>(a);
}

// Packed compat migration of a mixed-radix or Bluestein nkx on the array length M (the subgroup
// shape of the other non-power-of-two kernels: M / 8 threads from M = 1024, one wave below).
template <int M, int P, bool HALF>
__global__ __launch_bounds__(M / P) __attribute__((amdgpu_waves_per_eu(2))) void march2d_adj_packed_mixed(March2DArgs a)
{
    adj_body_packed<M, P, false, HALF, XMixed<M, P>>(a);
}
template <int M, int P, bool HALF>
__global__ __launch_bounds__(M / P) __attribute__((amdgpu_waves_per_eu(2))) void march2d_adj_packed_blue(March2DArgs a)
{
    adj_body_packed<M, P, false, HALF, XBlue<M, P>>(a);
}

#define SWEM_SUBGROUP_BOUNDS(P) __launch_bounds__((P) == 16 ? 512 : kMaxBlock)

template <int N, int MODE, int P>
__global__ SWEM_SUBGROUP_BOUNDS(P) void march2d_adj(March2DArgs a) { adj_body<N, P, MODE, false>(a); }

template <int N, int MODE, int P>
__global__ SWEM_SUBGROUP_BOUNDS(P) void march2d_fwd(March2DArgs a) { fwd_body<N, P, MODE, false>(a); }

template <int N, int MODE>
__global__ __launch_bounds__(kWaveBlock) __attribute__((amdgpu_waves_per_eu(2))) void march2d_adj_w(March2DArgs a)
{
    adj_body<N, N / 64, MODE, true>(a);
}

template <int N, int MODE>
__global__ __launch_bounds__(kWaveBlock) __attribute__((amdgpu_waves_per_eu(2))) void march2d_fwd_w(March2DArgs a)
{
    fwd_body<N, N / 64, MODE, true>(a);
}

// PAIR: 16 points per thread from nkx = 1024 (nkx/16 threads, radix-16 passes), one wave with
// nkx/64 points per lane below it.
template <int N, int MODE, int P>
__global__ __launch_bounds__(N / P) __attribute__((amdgpu_waves_per_eu(2))) void march2d_adj_pair(March2DArgs a)
{
    adj_body_pair<N, P, MODE, false>(a);
}

template <int N, int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void march2d_adj_pair_w(March2DArgs a)
{
    adj_body_pair<N, N / 64, MODE, true>(a);
}

// Transform lengths that are not powers of two: Bluestein through the array length M (one
// subgroup of M/P threads marches both wavefields of a frequency in migration, as PAIR does;
// modelling takes F frequencies per workgroup of at most 512 threads).  P = 8 from M = 1024:
// the chirp products add live registers, and 16 points per thread would spill.
template <int M, int MODE, int P>
__global__ __launch_bounds__(M / P) __attribute__((amdgpu_waves_per_eu(2))) void march2d_adj_blue(March2DArgs a)
{
    adj_body_pair<M, P, MODE, false, XBlue<M, P>>(a);
}

template <int M, int MODE, int P>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void march2d_fwd_blue(March2DArgs a)
{
    fwd_body<M, P, MODE, false, XBlue<M, P>>(a);
}

template <int M, int MODE, int P>
__global__ __launch_bounds__(M / P) __attribute__((amdgpu_waves_per_eu(2))) void march2d_adj_mixed(March2DArgs a)
{
    adj_body_pair<M, P, MODE, false, XMixed<M, P>>(a);
}

template <int M, int MODE, int P>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void march2d_fwd_mixed(March2DArgs a)
{
    fwd_body<M, P, MODE, false, XMixed<M, P>>(a);
}

// ------------------------------------------------------------------------ dispatch
// Kernel shape of a variant: threads per wavefield subgroup T, subgroups per frequency
// (2 = source and receiver in separate subgroups, 1 = one subgroup marches both), the largest
// workgroup, LDS buffers per subgroup.
//   WAVE         one 64-lane wave per wavefield, no barrier inside a transform (nkx <= 2048)
//   SUBGROUP     nkx/8 threads per wavefield (nkx/4, nkx/2 below 512), ping-pong LDS
//   SUBGROUP16   nkx/16 threads per wavefield, radix-16 passes (nkx >= 1024)
//   PAIR         (migration) one subgroup per frequency marches source and receiver in turn:
//                the imaging needs no hand-off and a barrier holds only that subgroup's waves;
//                a wave (nkx <= 1024) or nkx/16 threads.  Modelling uses the AUTO shape.
//   PACKED       PAIR's shape; PSPI compat migration with both wavefields in one transform
//                (the host selects it only where every depth steps both wavefields and the
//                uncut points of a depth share one vref; other modes run PAIR).
//   Bluestein    any other transform length up to 2048 (nkx = 1 too: one trace, nmx = 1, whose
//                transform is the identity): the PAIR body (migration) or the
//                subgroup body (modelling) over the array length M = blue (P = 8 from M = 1024,
//                one wave below); the variant option does not apply.
int march2d_kind(int nkx)
{
    if (nkx >= 128 && !(nkx & (nkx - 1))) return kLenPow2;
    return mixed_radix_plan(nkx, nullptr) ? kLenMixed : kLenBluestein;
}

int march2d_blue_len(int nkx)
{
    if (nkx < 1 || (nkx >= 128 && !(nkx & (nkx - 1)))) return 0;
    int m = 128;
    if (march2d_kind(nkx) == kLenMixed) {
        while (m < nkx) m <<= 1;
        return m <= 4096 ? m : 0;
    }
    if (nkx > 2048) return 0;
    while (m < 2 * nkx - 1) m <<= 1;
    return m;
}

March2DShape march2d_shape(int nkx, int variant, int adj)
{
    March2DShape sh{};
    if (const int m = march2d_blue_len(nkx)) {
        sh.blue = m;
        sh.mixed = march2d_kind(nkx) == kLenMixed;
        sh.variant = adj ? SWEM_VARIANT_PAIR : SWEM_VARIANT_SUBGROUP;
        if (adj && variant == SWEM_VARIANT_PACKED) sh.variant = SWEM_VARIANT_PACKED;
        sh.T = m >= 1024 ? m / 8 : 64;
        sh.per_freq = 1;
        sh.max_block = adj ? sh.T : 512;
        sh.bufs = 2;
        return sh;
    }
    if (nkx < 128 || nkx > 4096 || (nkx & (nkx - 1))) return sh;
    int v = variant;
    if ((v == SWEM_VARIANT_PAIR || v == SWEM_VARIANT_PAIR8 || v == SWEM_VARIANT_PACKED) && !adj) v = SWEM_VARIANT_AUTO;
    if (v == SWEM_VARIANT_PAIR8 && nkx < 512) v = SWEM_VARIANT_PAIR;
    if (v == SWEM_VARIANT_AUTO) v = adj ? SWEM_VARIANT_PAIR : (nkx <= 512 ? SWEM_VARIANT_WAVE : SWEM_VARIANT_SUBGROUP16);
    if (v == SWEM_VARIANT_WAVE && nkx > 2048) v = SWEM_VARIANT_SUBGROUP16;
    if (v == SWEM_VARIANT_SUBGROUP16 && nkx < 1024) v = SWEM_VARIANT_SUBGROUP;
    sh.variant = v;
    sh.per_freq = adj ? 2 : 1;
    switch (v) {
    case SWEM_VARIANT_WAVE:
        sh.T = 64, sh.max_block = kWaveBlock, sh.bufs = 1;
        break;
    case SWEM_VARIANT_SUBGROUP16:
        sh.T = nkx / 16, sh.max_block = 512, sh.bufs = 2;
        break;
    case SWEM_VARIANT_PAIR:
    case SWEM_VARIANT_PACKED:
        sh.T = nkx >= 1024 ? nkx / 16 : 64, sh.per_freq = 1, sh.max_block = sh.T, sh.bufs = nkx >= 1024 ? 2 : 1;
        break;
    case SWEM_VARIANT_PAIR8:
        sh.T = nkx / 8, sh.per_freq = 1, sh.max_block = sh.T, sh.bufs = 2;
        break;
    default:
        sh.variant = SWEM_VARIANT_SUBGROUP;
        sh.T = nkx / (nkx >= 512 ? 8 : (nkx >= 256 ? 4 : 2)), sh.max_block = kMaxBlock, sh.bufs = 2;
    }
    return sh;
}

bool march2d_supported(int nkx) { return march2d_shape(nkx, SWEM_VARIANT_AUTO, 1).T > 0; }

size_t march2d_lds_bytes(int nkx, const March2DShape& sh, int F)
{
    if (sh.blue) nkx = sh.blue;
    const size_t len = (size_t)(((nkx + nkx / 8 + 31) & ~31) + 4);  // LdsLen<nkx>
    return (size_t)F * sh.per_freq * sh.bufs * len * sizeof(float2);
}

int march2d_grid_blocks(int ngroups, int nshot) { return ((ngroups + 7) / 8) * 8 * nshot; }

template <typename K>
static int launch_kernel(K kernel, const March2DArgs& a, dim3 block, size_t lds, hipStream_t st)
{
    SWEM_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    dim3 grid(march2d_grid_blocks(a.ngroups, a.nshot));
    hipLaunchKernelGGL(kernel, grid, block, lds, st, a);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

template <int N, int MODE>
static int launch_one(const March2DArgs& a, const March2DShape& sh, int F, int adj, hipStream_t st)
{
    const size_t lds = march2d_lds_bytes(N, sh, F);
    dim3 block(sh.T * sh.per_freq * F);
    if (sh.variant == SWEM_VARIANT_PACKED && MODE == kModePspiCompat && adj) {
        const size_t lds_p = lds + kPackedScratch * sizeof(float2);
        const bool half = 2 * a.nmx <= N;
        if constexpr (N < 1024)
            return half ? launch_kernel(march2d_adj_packed_w<N, true>, a, block, lds_p, st)
                        : launch_kernel(march2d_adj_packed_w<N, false>, a, block, lds_p, st);
        else
            return half ? launch_kernel(march2d_adj_packed<N, 16, true>, a, block, lds_p, st)
                        : launch_kernel(march2d_adj_packed<N, 16, false>, a, block, lds_p, st);
    }
    if (sh.variant == SWEM_VARIANT_PAIR || sh.variant == SWEM_VARIANT_PACKED) {
        if constexpr (N < 1024)
            return launch_kernel(march2d_adj_pair_w<N, MODE>, a, block, lds, st);
        else
            return launch_kernel(march2d_adj_pair<N, MODE, 16>, a, block, lds, st);
    }
    if (sh.variant == SWEM_VARIANT_PAIR8) {
        if constexpr (N >= 512)
            return launch_kernel(march2d_adj_pair<N, MODE, 8>, a, block, lds, st);
    }
    if constexpr (N <= 2048) {
        if (sh.variant == SWEM_VARIANT_WAVE)
            return adj ? launch_kernel(march2d_adj_w<N, MODE>, a, block, lds, st)
                       : launch_kernel(march2d_fwd_w<N, MODE>, a, block, lds, st);
    }
    if constexpr (N >= 1024) {
        if (sh.variant == SWEM_VARIANT_SUBGROUP16)
            return adj ? launch_kernel(march2d_adj<N, MODE, 16>, a, block, lds, st)
                       : launch_kernel(march2d_fwd<N, MODE, 16>, a, block, lds, st);
    }
    constexpr int P = Pts<N>::value;
    return adj ? launch_kernel(march2d_adj<N, MODE, P>, a, block, lds, st)
               : launch_kernel(march2d_fwd<N, MODE, P>, a, block, lds, st);
}

template <int M, int MODE>
static int launch_blue_mode(const March2DArgs& a, const March2DShape& sh, int F, int adj, hipStream_t st)
{
    constexpr int P = M >= 1024 ? 8 : M / 64;
    const size_t lds = march2d_lds_bytes(a.nkx, sh, F);
    dim3 block(sh.T * sh.per_freq * F);
    if constexpr (MODE == kModePspiCompat) {
        if (adj && sh.variant == SWEM_VARIANT_PACKED) {
            const size_t lds_p = lds + kPackedScratch * sizeof(float2);
            if (sh.mixed)
                return 2 * a.nmx <= M ? launch_kernel(march2d_adj_packed_mixed<M, P, true>, a, block, lds_p, st)
                                      : launch_kernel(march2d_adj_packed_mixed<M, P, false>, a, block, lds_p, st);
            // Bluestein: M >= 2 nkx - 1 >= 2 nmx - 1, and M even, so the upper half is always zero
            return launch_kernel(march2d_adj_packed_blue<M, P, true>, a, block, lds_p, st);
        }
    }
    if (sh.mixed)
        return adj ? launch_kernel(march2d_adj_mixed<M, MODE, P>, a, block, lds, st)
                   : launch_kernel(march2d_fwd_mixed<M, MODE, P>, a, block, lds, st);
    return adj ? launch_kernel(march2d_adj_blue<M, MODE, P>, a, block, lds, st)
               : launch_kernel(march2d_fwd_blue<M, MODE, P>, a, block, lds, st);
}

template <int M>
static int launch_blue(const March2DArgs& a, const March2DShape& sh, int F, int mode, int adj, hipStream_t st)
{
    if (mode == kModeSSF) return launch_blue_mode<M, kModeSSF>(a, sh, F, adj, st);
    if (mode == kModePspiCompat) return launch_blue_mode<M, kModePspiCompat>(a, sh, F, adj, st);
    return launch_blue_mode<M, kModePspiLinear>(a, sh, F, adj, st);
}

template <int N>
static int launch_n(const March2DArgs& a, const March2DShape& sh, int F, int mode, int adj, hipStream_t st)
{
    if (mode == kModeSSF) return launch_one<N, kModeSSF>(a, sh, F, adj, st);
    if (mode == kModePspiCompat) return launch_one<N, kModePspiCompat>(a, sh, F, adj, st);
    return launch_one<N, kModePspiLinear>(a, sh, F, adj, st);
}

// F = frequencies per workgroup (migration: 2F subgroups, or F for PAIR; modelling: F subgroups).
int launch_march2d(const March2DArgs& a, int nkx, int F, int mode, int adj, int variant, hipStream_t st)
{
    const March2DShape sh = march2d_shape(nkx, variant, adj);
    if (sh.T <= 0)
        return set_error(SWEM_EUNSUPPORTED,
                         "2-D march: nkx=%d is neither a power of two 128..4096, 2^a 3^b 5^c up to 4096 nor 1..2048", nkx);
    if (F < 1 || sh.T * sh.per_freq * F > sh.max_block)
        return set_error(SWEM_EINVAL, "2-D march: %d frequencies of %d subgroups of %d threads exceed %d", F,
                         sh.per_freq, sh.T, sh.max_block);
    if (a.nmx > nkx || a.nref < 1 || a.nref > 31 || a.ngroups < 1 || a.nshot < 1 || a.nkh != nkx / 2 + 1 ||
        a.f0 < 0 || a.nfc < 1 || a.f0 + a.nfc > a.nf)
        return set_error(SWEM_EINVAL, "2-D march: inconsistent launch (nmx=%d nkx=%d nref=%d)", a.nmx, nkx, a.nref);
    if (march2d_lds_bytes(nkx, sh, F) > 160 * 1024)
        return set_error(SWEM_EINVAL, "2-D march: LDS for %d frequencies at nkx=%d exceeds 160 KiB", F, nkx);
    if (sh.blue) {
        if (a.nkx != nkx || (sh.mixed ? a.mr.n != nkx : (!a.chirp || !a.bchirp)))
            return set_error(SWEM_EINVAL, "2-D march: %s tables missing", sh.mixed ? "mixed-radix" : "Bluestein");
        switch (sh.blue) {
        case 128: return launch_blue<128>(a, sh, F, mode, adj, st);
        case 256: return launch_blue<256>(a, sh, F, mode, adj, st);
        case 512: return launch_blue<512>(a, sh, F, mode, adj, st);
        case 1024: return launch_blue<1024>(a, sh, F, mode, adj, st);
        case 2048: return launch_blue<2048>(a, sh, F, mode, adj, st);
        case 4096: return launch_blue<4096>(a, sh, F, mode, adj, st);
        }
    }
    switch (nkx) {
    case 128: return launch_n<128>(a, sh, F, mode, adj, st);
    case 256: return launch_n<256>(a, sh, F, mode, adj, st);
    case 512: return launch_n<512>(a, sh, F, mode, adj, st);
    case 1024: return launch_n<1024>(a, sh, F, mode, adj, st);
    case 2048: return launch_n<2048>(a, sh, F, mode, adj, st);
    case 4096: return launch_n<4096>(a, sh, F, mode, adj, st);
    }
    return set_error(SWEM_EUNSUPPORTED, "2-D march: nkx=%d", nkx);
}


// ------------------------------------------------------------------ source incidence angles
// compute_angles (wem.c:566-776) for 2-D.  The source march is the modelling receiver's step
// (ssop with adj = false and -dz, which is ssop_source, wem.c:778-826) stored at every depth.
// The derivatives of calculate_derivatives (wem.c:828-981) are separable: the inverse 3-D
// transform of i kx FFT3(u), divided by nkx nky nkz, equals IFFT_x(i kx FFT_x(u)) / nkx, so the
// x parts run along rows and the z part along columns of the stored volume.
template <int N, int P, class X>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void angle_march(March2DArgs a)
{
    constexpr int T = N / P;
    constexpr int L = LdsLen<N>::value;
    extern __shared__ float2 smem[];
    int g, shot;
    decode_block(a.ngroups, a.nshot, g, shot);
    if (g < 0) return;
    const int NSUB = blockDim.x / T;
    const int tid = threadIdx.x, s = tid / T, t = tid - s * T;
    const int fc = g * NSUB + s;
    const bool active = fc < a.nfc;
    const int fl = a.f0 + fc;
    StepCtx c;
    c.t = t;
    c.fc = __builtin_amdgcn_readfirstlane(active ? fc : 0);
    float2* lds = smem + (size_t)s * 2 * L;
    int pp = 0;
    float tap[P];
    load_taper<N, P, kModeSSF>(tap, a, t);
    const int igx = a.igx[shot];
    float2* vol = a.smig + (size_t)(active ? fc : 0) * a.nz * a.nmx;
    float2 u[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = t + q * T;
        u[q] = (active && x == igx) ? a.spec_s[fl] : make_float2(0.f, 0.f);
    }
    for (int iz = 0; iz < a.nz; ++iz) {  // extrapolate_source, wem.c:712-719
        c.iz = iz;
        const float z = a.oz + a.dz * (float)iz;
        const bool ds = z >= a.sz;
        if (ds) ssf_fwd_step<N, P, false, X>(u, lds, pp, a, c, tap);
        if (active) {
            const auto sr = buf_rsrc(vol + (size_t)iz * a.nmx, (uint32_t)a.nmx * 8u);
#pragma unroll
            for (int q = 0; q < P; ++q) buf_store2(sr, (uint32_t)(t + q * T) * 8u, ds ? u[q] : make_float2(0.f, 0.f));
        }
    }
}

__device__ __forceinline__ float kbin(int k, int n, float dk)
{
    return k <= n / 2 ? dk * (float)k : -(dk * (float)n - dk * (float)k);
}

__device__ __forceinline__ float2 times_i(float2 v, float k) { return make_float2(-v.y * k, v.x * k); }

// z pass (wem.c:888-899): |IFFT_z(i kz FFT_z(u))| / nz along the z line of one (frequency, point).
template <int N, int P, class X>
__global__ __launch_bounds__(N / P) void angle_cols(March2DArgs az, AngleArgs b)
{
    constexpr int T = N / P;
    extern __shared__ float2 smem[];
    const int l = blockIdx.x;
    const size_t base = (size_t)(l / b.inner) * b.outer + (size_t)(l % b.inner) * b.inner_stride;
    const int t = threadIdx.x;
    int pp = 0;
    const float2* col = b.vol + base;
    float2 v[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int iz = t + q * T;
        v[q] = iz < b.nz ? col[(size_t)iz * b.elem] : make_float2(0.f, 0.f);
    }
    __builtin_amdgcn_sched_barrier(0);
    X::template fft<-1, false>(v, smem, az, t, pp);
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int k = t + q * T;
        v[q] = times_i(v[q], k < b.nz ? kbin(k, b.nz, b.dkz) : 0.f);
    }
    X::template fft<+1, false>(v, smem, az, t, pp);
    float* out = b.dzmag + base;
    const float inv = 1.f / (float)b.nz;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int iz = t + q * T;
        if (iz < b.nz) out[(size_t)iz * b.elem] = sqrtf(v[q].x * v[q].x + v[q].y * v[q].y) * inv;
    }
}

// Lateral pass along one x (or y) line of one depth: per frequency in order, the derivative along
// the line (wem.c:858-887) against the z pass's |u_z| (wem.c:747-748) and the left / right split
// (wem.c:901-970, wem.c:739-740), added to the sums of earlier chunks.
template <int N, int P, class X>
__global__ __launch_bounds__(N / P) void angle_rows(March2DArgs ax, AngleArgs b)
{
    constexpr int T = N / P;
    extern __shared__ float2 smem[];
    const int l = blockIdx.x;
    const size_t base = (size_t)(l / b.inner) * b.outer + (size_t)(l % b.inner) * b.inner_stride;
    const int t = threadIdx.x;
    int pp = 0;
    const int n = ax.nkx;
    const float deg = (float)(180.0 / 3.141592653589793);
    const float inv = 1.f / (float)n;
    const float inv_nw = 1.f / (float)b.nw;
    float acc[P], sgn[P], kx[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = t + q * T;
        const size_t i = base + (size_t)x * b.elem;
        acc[q] = x < b.len ? b.acc[i] : 0.f;
        sgn[q] = x < b.len ? b.sgn[i] : 0.f;
        kx[q] = x < n ? kbin(x, n, b.dk) : 0.f;
    }
    for (int f = 0; f < b.nfc; ++f) {
        const float2* ur = b.vol + (size_t)f * b.item + base;
        const float* zr = b.dzmag + (size_t)f * b.item + base;
        float2 A[P];
        float mu[P], mz[P];
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int x = t + q * T;
            A[q] = x < b.len ? ur[(size_t)x * b.elem] : make_float2(0.f, 0.f);
            mz[q] = x < b.len ? zr[(size_t)x * b.elem] : 0.f;
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < P; ++q) mu[q] = sqrtf(A[q].x * A[q].x + A[q].y * A[q].y);
        X::template fft<-1, false>(A, smem, ax, t, pp);
        float mx[P], ml[P];
        float2 B[P];
#pragma unroll
        for (int q = 0; q < P; ++q) B[q] = times_i(A[q], kx[q]);
        X::template fft<+1, false>(B, smem, ax, t, pp);
#pragma unroll
        for (int q = 0; q < P; ++q) mx[q] = sqrtf(B[q].x * B[q].x + B[q].y * B[q].y) * inv;
        // negative wavenumbers only (k > n/2), then positive only (k <= n/2)
#pragma unroll
        for (int q = 0; q < P; ++q) B[q] = times_i(A[q], (t + q * T) <= n / 2 ? 0.f : kx[q]);
        X::template fft<+1, false>(B, smem, ax, t, pp);
#pragma unroll
        for (int q = 0; q < P; ++q) ml[q] = sqrtf(B[q].x * B[q].x + B[q].y * B[q].y);
#pragma unroll
        for (int q = 0; q < P; ++q) B[q] = times_i(A[q], (t + q * T) <= n / 2 ? kx[q] : 0.f);
        X::template fft<+1, false>(B, smem, ax, t, pp);
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const float mr = sqrtf(B[q].x * B[q].x + B[q].y * B[q].y);
            sgn[q] += (ml[q] >= mr ? mu[q] : -mu[q]) * inv_nw;
            acc[q] += deg * atanf(mx[q] / (mz[q] + 1e-10f));
        }
    }
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = t + q * T;
        if (x < b.len) {
            const size_t i = base + (size_t)x * b.elem;
            b.acc[i] = acc[q];
            b.sgn[i] = sgn[q];
        }
    }
}

template <int N, class X>
static int launch_angles_n(int which, const March2DArgs& ax, const March2DArgs& az, const AngleArgs& b, int F,
                           hipStream_t st)
{
    constexpr int P = N >= 1024 ? 8 : N / 64, T = N / P;
    const size_t len = (size_t)LdsLen<N>::value * sizeof(float2);
    if (which == 0) {
        if (F < 1 || F * T > 512 || F * 2 * len > 160 * 1024) return set_error(SWEM_EINVAL, "angles: %d frequencies per block", F);
        return launch_kernel(angle_march<N, P, X>, ax, dim3(T * F), F * 2 * len, st);
    }
    auto k = which == 1 ? angle_cols<N, P, X> : angle_rows<N, P, X>;
    const March2DArgs& a = which == 1 ? az : ax;
    const unsigned grid = (unsigned)b.lines;
    SWEM_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(2 * len)));
    hipLaunchKernelGGL(k, dim3(grid), dim3(T), 2 * len, st, a, b);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

template <int N>
static int launch_angles_len(int which, int kind, const March2DArgs& ax, const March2DArgs& az, const AngleArgs& b,
                             int F, hipStream_t st)
{
    constexpr int P = N >= 1024 ? 8 : N / 64;
    if (kind == kLenMixed) return launch_angles_n<N, XMixed<N, P>>(which, ax, az, b, F, st);
    return kind == kLenBluestein ? launch_angles_n<N, XBlue<N, P>>(which, ax, az, b, F, st)
                                 : launch_angles_n<N, XPow2<N, P>>(which, ax, az, b, F, st);
}

int angles_supported(int n) { return n >= 2 && (march2d_blue_len(n) || (n >= 128 && n <= 4096 && !(n & (n - 1)))); }

// The tables a launch needs for its transform length are present.
static bool len_tables(const March2DArgs& a, int kind)
{
    return kind == kLenPow2 || (kind == kLenMixed ? a.mr.n == a.nkx : a.chirp && a.bchirp);
}

int launch_angles(int which, const March2DArgs& ax, const March2DArgs& az, const AngleArgs& b, int F, hipStream_t st)
{
    const int n = which == 1 ? az.nkx : ax.nkx;
    const int m = march2d_blue_len(n);
    const int len = m ? m : n, kind = m ? march2d_kind(n) : kLenPow2;
    if (!angles_supported(n) || !len_tables(which == 1 ? az : ax, kind))
        return set_error(SWEM_EUNSUPPORTED, "angles: transform length %d", n);
    if (b.len < 1 || b.nz < 1 || b.nfc < 1 || (which > 0 && (b.inner < 1 || b.lines < 1)) ||
        (which == 2 && b.len != n) || (which == 1 && b.nz != n))
        return set_error(SWEM_EINVAL, "angles: inconsistent launch (len=%d nz=%d n=%d)", b.len, b.nz, n);
    switch (len) {
    case 128: return launch_angles_len<128>(which, kind, ax, az, b, F, st);
    case 256: return launch_angles_len<256>(which, kind, ax, az, b, F, st);
    case 512: return launch_angles_len<512>(which, kind, ax, az, b, F, st);
    case 1024: return launch_angles_len<1024>(which, kind, ax, az, b, F, st);
    case 2048: return launch_angles_len<2048>(which, kind, ax, az, b, F, st);
    case 4096: return launch_angles_len<4096>(which, kind, ax, az, b, F, st);
    }
    return set_error(SWEM_EUNSUPPORTED, "angles: transform length %d", n);
}


// ------------------------------------------------------------------------ elastic WEM (2-D)
// ewem (ewem.c:3-366).  Per frequency: the P source (vp, -dz, adjoint order) normalised by the
// largest source sample, and the receiver components (ux, uz).  Per depth the components are
// separated into the P and second S potentials (elastic_separate_3d, ewem.c:553-741; the first
// S potential is never formed), each propagated with its own velocity's split-step step, and
// recombined; the P and S images correlate the potentials with the source.  The separation
// matrix of wavenumber kx (kx = dkx ikx for ikx < nkx/2, else -(dkx nkx - dkx ikx)) uses the
// depth's layer velocities vp0 = 1/po_p, vs0 = 1/po_s and is computed in place.
// (i1, i2) -> (o1, o2) through the wavenumber domain; o1, o2 are zero at x >= nmx.
template <int N, int P, class X>
__device__ __forceinline__ void esep(const float2 (&i1)[P], const float2 (&i2)[P], float2 (&o1)[P], float2 (&o2)[P],
                                     float2* lds, int& pp, const March2DArgs& a, const EwemArgs& e, int t, float w,
                                     float vp, float vs, bool sep, bool adj)
{
    constexpr int T = N / P;
    const int n = a.nkx;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        o1[q] = i1[q];
        o2[q] = i2[q];
    }
    X::template fft<-1, false>(o1, lds, a, t, pp);
    X::template fft<-1, false>(o2, lds, a, t, pp);
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int k = t + q * T;
        const float kx = k < n ? ((float)k < (float)n / 2.f ? e.dkx * (float)k : -(e.dkx * (float)n - e.dkx * (float)k))
                               : 0.f;
        float2 r1, r2;
        esep_k(kx, w, vp, vs, sep, adj, o1[q], o2[q], r1, r2);
        o1[q] = k < n ? cscale(r1, a.inv_nk) : make_float2(0.f, 0.f);
        o2[q] = k < n ? cscale(r2, a.inv_nk) : make_float2(0.f, 0.f);
    }
    X::template fft<+1, false>(o1, lds, a, t, pp);
    X::template fft<+1, false>(o2, lds, a, t, pp);
#pragma unroll
    for (int q = 0; q < P; ++q)
        if (t + q * T >= a.nmx) o1[q] = o2[q] = make_float2(0.f, 0.f);
}

template <int N, int P, class X, bool ADJ>
__global__ __launch_bounds__(N / P) __attribute__((amdgpu_waves_per_eu(2))) void ewem_march(March2DArgs ap,
                                                                                             March2DArgs as, EwemArgs e)
{
    constexpr int T = N / P;
    extern __shared__ float2 smem[];
    int g, shot;
    decode_block(ap.ngroups, ap.nshot, g, shot);
    if (g < 0) return;
    const int t = threadIdx.x;
    const int fc = g;
    const bool active = fc < ap.nfc;
    const int fl = ap.f0 + fc;
    const int iw = ap.ifmin + fl;
    const float w = (float)iw * e.dwf;
    StepCtx c;
    c.t = t;
    c.fc = __builtin_amdgcn_readfirstlane(active ? fc : 0);
    int pp = 0;
    float tap[P];
    load_taper<N, P, kModeSSF>(tap, ap, t);
    const int igx = ap.igx[shot];
    const float factor = iw == 0 ? 1.f : 2.f;
    const size_t sp = (size_t)shot * ap.nmx * e.nf;  // this shot's spectra
    float2 us[P], rx[P], rz[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int x = t + q * T;
        us[q] = (active && x == igx) ? ap.spec_s[fl] : make_float2(0.f, 0.f);
        rx[q] = (ADJ && active && x < ap.nmx) ? e.spx[sp + (size_t)x * e.nf + fl] : make_float2(0.f, 0.f);
        rz[q] = (ADJ && active && x < ap.nmx) ? e.spz[sp + (size_t)x * e.nf + fl] : make_float2(0.f, 0.f);
    }
    // partial images / stored source of (shot, frequency)
    const size_t vol = ((size_t)shot * ap.ngroups + (active ? fc : 0)) * ap.nz * ap.nmx;
    if constexpr (ADJ) {
        for (int iz = 0; iz < ap.nz; ++iz) {  // elastic_extrap1f, ewem.c:302-327
            c.iz = iz;
            const float z = ap.oz + ap.dz * (float)iz;
            const bool ds = z >= ap.sz, dr = z >= ap.gz;
            if (ds) ssf_adj_step<N, P, false, true, X>(us, smem, pp, ap, c, true, tap);
            float2 up[P], u2[P];
            if (dr) {
                const float vp = e.vp0[iz], vs = e.vs0[iz];
                esep<N, P, X>(rx, rz, up, u2, smem, pp, ap, e, t, w, vp, vs, true, true);
                ssf_adj_step<N, P, false, true, X>(up, smem, pp, ap, c, false, tap);
                ssf_adj_step<N, P, false, true, X>(u2, smem, pp, as, c, false, tap);
                esep<N, P, X>(up, u2, rx, rz, smem, pp, ap, e, t, w, vp, vs, false, true);
            }
            if (active) {
                const auto rp = buf_rsrc(e.ppart + vol + (size_t)iz * ap.nmx, (uint32_t)ap.nmx * 4u);
                const auto rs = buf_rsrc(e.spart + vol + (size_t)iz * ap.nmx, (uint32_t)ap.nmx * 4u);
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    const float2 sm = ds ? cscale(us[q], e.inv_maxs) : make_float2(0.f, 0.f);
                    const uint32_t off = (uint32_t)(t + q * T) * 4u;
                    buf_store1(rp, off, dr ? factor * (sm.x * up[q].x + sm.y * up[q].y) : 0.f);
                    buf_store1(rs, off, dr ? factor * (sm.x * u2[q].x + sm.y * u2[q].y) : 0.f);
                }
            }
        }
    } else {
        for (int iz = 0; iz < ap.nz; ++iz) {  // source, stored per depth (ewem.c:302-310)
            c.iz = iz;
            const float z = ap.oz + ap.dz * (float)iz;
            const bool ds = z >= ap.sz;
            if (ds) ssf_adj_step<N, P, false, true, X>(us, smem, pp, ap, c, true, tap);
            if (active) {
                const auto sr = buf_rsrc(e.smig + vol + (size_t)iz * ap.nmx, (uint32_t)ap.nmx * 8u);
#pragma unroll
                for (int q = 0; q < P; ++q)
                    buf_store2(sr, (uint32_t)(t + q * T) * 8u, ds ? cscale(us[q], e.inv_maxs) : make_float2(0.f, 0.f));
            }
        }
        for (int iz = ap.nz - 1; iz >= 0; --iz) {  // receiver upward (ewem.c:328-346)
            c.iz = iz;
            const float z = ap.oz + ap.dz * (float)iz;
            if (!(z >= ap.gz)) continue;
            const float vp = e.vp0[iz], vs = e.vs0[iz];
            float2 up[P], u2[P];
            esep<N, P, X>(rx, rz, up, u2, smem, pp, ap, e, t, w, vp, vs, true, false);
            if (active) {
                const auto sr = buf_rsrc(e.smig + vol + (size_t)iz * ap.nmx, (uint32_t)ap.nmx * 8u);
                const auto mp = buf_rsrc(e.mppT + (size_t)iz * ap.nmx, (uint32_t)ap.nmx * 4u);
                const auto ms = buf_rsrc(e.mpsT + (size_t)iz * ap.nmx, (uint32_t)ap.nmx * 4u);
                float2 sv[P];
                float mpv[P], msv[P];
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    const uint32_t x = (uint32_t)(t + q * T);
                    sv[q] = buf_load2(sr, x * 8u);
                    mpv[q] = buf_load1(mp, x * 4u);
                    msv[q] = buf_load1(ms, x * 4u);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    up[q] = cadd(up[q], cscale(sv[q], mpv[q]));
                    u2[q] = cadd(u2[q], cscale(sv[q], msv[q]));
                }
            }
            ssf_fwd_step<N, P, false, X>(up, smem, pp, ap, c, tap);
            ssf_fwd_step<N, P, false, X>(u2, smem, pp, as, c, tap);
            esep<N, P, X>(up, u2, rx, rz, smem, pp, ap, e, t, w, vp, vs, false, false);
        }
        if (active) {
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int x = t + q * T;
                if (x < ap.nmx) {
                    e.dgx[sp + (size_t)x * e.nf + fl] = rx[q];
                    e.dgz[sp + (size_t)x * e.nf + fl] = rz[q];
                }
            }
        }
    }
}

template <int N, class X>
static int launch_ewem_n(const March2DArgs& ap, const March2DArgs& as, const EwemArgs& e, int adj, hipStream_t st)
{
    constexpr int P = N >= 1024 ? 8 : N / 64, T = N / P;
    const size_t lds = (size_t)LdsLen<N>::value * 2 * sizeof(float2);
    auto k = adj ? ewem_march<N, P, X, true> : ewem_march<N, P, X, false>;
    SWEM_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k, dim3(march2d_grid_blocks(ap.ngroups, ap.nshot)), dim3(T), lds, st, ap, as, e);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

template <int N>
static int launch_ewem_len(int kind, const March2DArgs& ap, const March2DArgs& as, const EwemArgs& e, int adj,
                           hipStream_t st)
{
    constexpr int P = N >= 1024 ? 8 : N / 64;
    if (kind == kLenMixed) return launch_ewem_n<N, XMixed<N, P>>(ap, as, e, adj, st);
    return kind == kLenBluestein ? launch_ewem_n<N, XBlue<N, P>>(ap, as, e, adj, st)
                                 : launch_ewem_n<N, XPow2<N, P>>(ap, as, e, adj, st);
}

int launch_ewem(const March2DArgs& ap, const March2DArgs& as, const EwemArgs& e, int adj, hipStream_t st)
{
    const int n = ap.nkx, m = march2d_blue_len(n), len = m ? m : n, kind = m ? march2d_kind(n) : kLenPow2;
    if (!angles_supported(n) || as.nkx != n || !len_tables(ap, kind) || !len_tables(as, kind) || ap.nshot < 1 ||
        ap.ngroups != ap.nfc || ap.nmx > n)
        return set_error(SWEM_EINVAL, "ewem: inconsistent launch (nkx=%d nmx=%d)", n, ap.nmx);
    switch (len) {
    case 128: return launch_ewem_len<128>(kind, ap, as, e, adj, st);
    case 256: return launch_ewem_len<256>(kind, ap, as, e, adj, st);
    case 512: return launch_ewem_len<512>(kind, ap, as, e, adj, st);
    case 1024: return launch_ewem_len<1024>(kind, ap, as, e, adj, st);
    case 2048: return launch_ewem_len<2048>(kind, ap, as, e, adj, st);
    case 4096: return launch_ewem_len<4096>(kind, ap, as, e, adj, st);
    }
    return set_error(SWEM_EUNSUPPORTED, "ewem: transform length %d", n);
}

}  // namespace swem
