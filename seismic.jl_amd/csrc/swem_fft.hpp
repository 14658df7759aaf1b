// swem_fft.hpp -- register/LDS Stockham FFT for gfx950 and complex helpers.
//
// A subgroup of T = N/P threads transforms N points.  Thread t holds its points in the
// "strided mapping" v[q] = x[t + q*T], q = 0..P-1.  Each pass of radix R (R | P) does P/R
// butterflies per thread; butterfly j = t + b*T reads x[j + r*N/R] -- exactly register slot
// b + r*P/R of the strided mapping -- twiddles them, runs an R-point DFT in registers and
// writes the self-sorting Stockham position (j/Ns)*Ns*R + j%Ns + r*Ns to LDS.  The last pass
// (Ns*R == N) writes back to the same register slots, so both the input and the output of a
// transform are in the strided mapping and the output is in natural order: the forward FFT
// leaves X[k = t + q*T] in v[q].  A transform costs (passes - 1) LDS round trips.
//
// LDS: each subgroup owns two buffers of LdsLen<N> complex and alternates between them
// (parity `pp`, carried across calls), so one barrier per round trip suffices: a buffer written
// in round trip k is rewritten in round trip k+2, after every thread has passed barrier k+1 and
// hence finished reading it.  Indices are padded per pass (PassPad) to spread the stores and
// loads over the banks.
// Sign convention as FFTW: SIGN = -1 forward exp(-i...), SIGN = +1 backward, unnormalised.
#pragma once
#include <hip/hip_runtime.h>

namespace swem {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by SIGN*i
template <int SIGN>
__device__ __forceinline__ float2 mul_si(float2 a)
{
    return SIGN < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int R, int SIGN>
struct Dft;

template <int SIGN>
struct Dft<2, SIGN> {
    static __device__ __forceinline__ void run(float2* v)
    {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int SIGN>
struct Dft<4, SIGN> {
    static __device__ __forceinline__ void run(float2* v)
    {
        float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        float2 t2 = cadd(v[1], v[3]), t3 = mul_si<SIGN>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <int SIGN>
struct Dft<8, SIGN> {
    static __device__ __forceinline__ void run(float2* v)
    {
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, SIGN>::run(e);
        Dft<4, SIGN>::run(o);
        const float r = 0.70710678118654752440f;
        // W8^1 o = r (o + SIGN i o), W8^2 o = SIGN i o, W8^3 o = r (SIGN i o - o)
        float2 o1 = cscale(cadd(o[1], mul_si<SIGN>(o[1])), r);
        float2 o2 = mul_si<SIGN>(o[2]);
        float2 o3 = cscale(csub(mul_si<SIGN>(o[3]), o[3]), r);
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

template <int SIGN>
struct Dft<16, SIGN> {
    static __device__ __forceinline__ void run(float2* v)
    {
        float2 e[8], o[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            e[i] = v[2 * i];
            o[i] = v[2 * i + 1];
        }
        Dft<8, SIGN>::run(e);
        Dft<8, SIGN>::run(o);
        // X_k = E_k + W16^k O_k, X_{k+8} = E_k - W16^k O_k, W16^k = cos(pi k/8) + SIGN i sin(pi k/8)
        const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
        const float r = 0.70710678118654752440f;
        float2 t[8];
        t[0] = o[0];
        t[1] = cmul(o[1], make_float2(c1, SIGN * s1));
        t[2] = cscale(cadd(o[2], mul_si<SIGN>(o[2])), r);
        t[3] = cmul(o[3], make_float2(s1, SIGN * c1));
        t[4] = mul_si<SIGN>(o[4]);
        t[5] = cmul(o[5], make_float2(-s1, SIGN * c1));
        t[6] = cscale(csub(mul_si<SIGN>(o[6]), o[6]), r);
        t[7] = cmul(o[7], make_float2(-c1, SIGN * s1));
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            v[k] = cadd(e[k], t[k]);
            v[k + 8] = csub(e[k], t[k]);
        }
    }
};

template <int SIGN>
struct Dft<3, SIGN> {
    // X1,2 = x0 - (x1 + x2) / 2 +- SIGN i (sqrt(3) / 2) (x1 - x2)
    static __device__ __forceinline__ void run(float2* v)
    {
        const float h = 0.86602540378443864676f;
        const float2 a = cadd(v[1], v[2]), b = cscale(mul_si<SIGN>(csub(v[1], v[2])), h);
        const float2 c = csub(v[0], cscale(a, 0.5f));
        v[0] = cadd(v[0], a);
        v[1] = cadd(c, b);
        v[2] = csub(c, b);
    }
};

template <int SIGN>
struct Dft<5, SIGN> {
    // with a1 = x1 + x4, a2 = x2 + x3, b1 = x1 - x4, b2 = x2 - x3 and w = exp(SIGN 2 pi i / 5):
    // X1,4 = x0 + c1 a1 + c2 a2 +- SIGN i (s1 b1 + s2 b2), X2,3 = x0 + c2 a1 + c1 a2 +- SIGN i (s2 b1 - s1 b2)
    static __device__ __forceinline__ void run(float2* v)
    {
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
        const float2 a1 = cadd(v[1], v[4]), a2 = cadd(v[2], v[3]);
        const float2 b1 = csub(v[1], v[4]), b2 = csub(v[2], v[3]);
        const float2 x0 = v[0];
        const float2 p1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
        const float2 p2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
        const float2 q1 = mul_si<SIGN>(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
        const float2 q2 = mul_si<SIGN>(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
        v[0] = cadd(x0, cadd(a1, a2));
        v[1] = cadd(p1, q1);
        v[4] = csub(p1, q1);
        v[2] = cadd(p2, q2);
        v[3] = csub(p2, q2);
    }
};

// x * exp(SIGN i pi M / 8) for a compile-time M in 0..7 (the twiddles W_R^n of R <= 16).
template <int M, int SIGN>
__device__ __forceinline__ float2 cmul_pi8(float2 x)
{
    constexpr float C[5] = {1.f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f, 0.f};
    if constexpr (M == 0) return x;
    else if constexpr (M == 4) return mul_si<SIGN>(x);
    else if constexpr (M == 2) return cscale(cadd(x, mul_si<SIGN>(x)), C[2]);
    else if constexpr (M == 6) return cscale(csub(mul_si<SIGN>(x), x), C[2]);
    else {
        // cos and sin of M pi / 8 (M odd): cos = +-C[M or 8 - M], sin = C[4 - M or M - 4]
        constexpr float c = M < 4 ? C[M] : -C[8 - M];
        constexpr float sn = M < 4 ? C[4 - M] : C[M - 4];
        return cmul(x, make_float2(c, SIGN * sn));
    }
}

// R-point DFT whose inputs r >= R/2 are zero: X_2m = DFT_{R/2}(x_n)_m, X_2m+1 = DFT_{R/2}(x_n W_R^n)_m
// over n < R/2 (W_R = exp(SIGN 2 pi i / R)); v[r] for r >= R/2 is not read.
template <int R, int SIGN>
__device__ __forceinline__ void dft_half_in(float2* v)
{
    constexpr int H = R / 2;
    if constexpr (R == 2) {
        v[1] = v[0];
    } else {
        float2 e[H], o[H];
#pragma unroll
        for (int n = 0; n < H; ++n) e[n] = v[n];
        o[0] = v[0];
        if constexpr (H > 1) o[1] = cmul_pi8<16 / R, SIGN>(v[1]);
        if constexpr (H > 2) {
            o[2] = cmul_pi8<32 / R, SIGN>(v[2]);
            o[3] = cmul_pi8<48 / R, SIGN>(v[3]);
        }
        if constexpr (H > 4) {
            o[4] = cmul_pi8<64 / R, SIGN>(v[4]);
            o[5] = cmul_pi8<80 / R, SIGN>(v[5]);
            o[6] = cmul_pi8<96 / R, SIGN>(v[6]);
            o[7] = cmul_pi8<112 / R, SIGN>(v[7]);
        }
        Dft<H, SIGN>::run(e);
        Dft<H, SIGN>::run(o);
#pragma unroll
        for (int m = 0; m < H; ++m) {
            v[2 * m] = e[m];
            v[2 * m + 1] = o[m];
        }
    }
}

// Bank-conflict padding: one complex inserted every 2^SH.  The LDS has 64 banks of 4 bytes (32
// complex per row).  A Stockham pass writes at lane stride R (first pass) or in groups of NS
// lanes, then reads back at unit lane stride; SH is chosen per pass so both the writes and the
// reads of a 64-lane instruction touch every bank exactly twice (the minimum for 8-byte
// accesses), except the radix-8 NS = 8 and radix-4 NS > 1 passes, which keep one 3-way access.
template <int SH>
__device__ __forceinline__ int padsh(int i)
{
    return i + (i >> SH);
}
// Padding of the strided-mapping accesses x[t + q T] outside a transform (hand-offs).
__device__ __forceinline__ int lds_pad(int i) { return padsh<5>(i); }
template <int NS, int R>
struct PassPad {
    static constexpr int value = (R == 8)   ? (NS == 8 ? 3 : 5)
                                 : (R == 4) ? (NS == 1 ? 5 : 3)
                                 : (R == 2) ? ((NS == 8 || NS == 16) ? 3 : 5)
                                            : 5;
};
// Complex elements per LDS buffer: the largest padded index (shift 3), rounded to a multiple of
// 32 plus 4, so consecutive buffers start 8 banks apart (tiles staged across buffers spread
// over all banks).
template <int N>
struct LdsLen {
    static constexpr int value = ((N + N / 8 + 31) & ~31) + 4;
};

// Number of radix passes of fft<N,P> (for flop accounting and tests).
template <int N, int P>
constexpr int fft_passes()
{
    int n = 1, p = 0;
    while (n < N) {
        int rem = N / n;
        n *= rem >= P ? P : rem;
        ++p;
    }
    return p;
}

// Twiddle seeds of fft<N,P>: for each pass after the first (NS > 1, radix R) and each butterfly
// b of a thread, the powers w^e, w = exp(-2 pi i k / (NS R)), k = (t + b T) mod NS, for the
// seeded exponents e (1, 2, 3 and the multiples of 4 below R), read from the table at the start
// of a transform.  The other powers are one product of two seeds, w^(4a + c) = w^(4a) w^c, so
// a twiddle carries at most one rounding beyond the table's (powers by longer chains of
// products drift: their systematic error, repeated over every transform of a march, grows the
// image error tenfold).  Radix 16: 6 loads per butterfly instead of 15, radix 8: 4 instead of 7.
template <int R>
constexpr int seed_count()
{
    return R <= 4 ? R - 1 : 3 + (R - 1) / 4;
}
// the i-th seeded exponent: 1, 2, 3, then 4, 8, 12, ...
constexpr int seed_exp(int i) { return i < 3 ? i + 1 : 4 * (i - 2); }

template <int N, int P>
constexpr int fft_seeds_before(int ns_end)
{
    int n = 1, c = 0;
    while (n < N && n < ns_end) {
        const int rem = N / n, r = rem >= P ? P : rem;
        if (n > 1) c += (P / r) * (r <= 4 ? r - 1 : 3 + (r - 1) / 4);
        n *= r;
    }
    return c;
}
template <int N, int P>
struct FftSeeds {
    static constexpr int count = fft_seeds_before<N, P>(N);
    float2 w[count > 0 ? count : 1];
};

// The seeds of pass NS (all passes from NS on when ALL).
template <int N, int P, int NS, bool ALL>
__device__ __forceinline__ void fft_seeds_load(FftSeeds<N, P>& s, const float2* __restrict__ tw, int t)
{
    constexpr int T = N / P;
    constexpr int REM = N / NS;
    constexpr int R = REM >= P ? P : REM;
    if constexpr (NS > 1) {
        constexpr int off = fft_seeds_before<N, P>(NS);
        constexpr int S = seed_count<R>();
#pragma unroll
        for (int b = 0; b < P / R; ++b) {
            // opaque: the loads stay in the transform instead of pinning registers across the
            // caller's loops
            int ti = ((t + b * T) & (NS - 1)) * (N / (NS * R));
            asm volatile("" : "+v"(ti));
#pragma unroll
            for (int i = 0; i < S; ++i) s.w[off + b * S + i] = tw[seed_exp(i) * ti];
        }
    }
    if constexpr (ALL && NS * R < N) fft_seeds_load<N, P, NS * R, true>(s, tw, t);
}

// pw[r] = w^r, r = 1 .. R-1, from the seeds sd[i] = w^seed_exp(i) (conjugated for SIGN > 0)
template <int R, int SIGN>
__device__ __forceinline__ void twiddle_powers(const float2* sd, float2 (&pw)[R])
{
    constexpr int S = seed_count<R>();
#pragma unroll
    for (int i = 0; i < S; ++i) pw[seed_exp(i)] = SIGN > 0 ? cconj(sd[i]) : sd[i];
#pragma unroll
    for (int r = 5; r < R; ++r)
        if (r % 4 != 0) pw[r] = cmul(pw[r & ~3], pw[r & 3]);
}

// One Stockham pass and the passes after it on one wavefield v in the strided mapping.
// lds: the subgroup's two buffers (2 * LdsLen<N> complex); pp: buffer parity, toggled per
// round trip; s: the twiddle seeds, loaded at the start of the transform (PRE) or of each pass
// (every load of a pass in flight before the first use: under register pressure the scheduler
// would otherwise pair each load with its use and serialise the latencies).
template <int N, int P, int SIGN, int NS, bool PRE, bool HIN>
__device__ __forceinline__ void fft_pass(float2 (&v)[P], float2* lds, const float2* __restrict__ tw,
                                         FftSeeds<N, P>& s, int t, int& pp)
{
    constexpr int T = N / P;
    constexpr int REM = N / NS;
    constexpr int R = REM >= P ? P : REM;
    constexpr int B = P / R;
    constexpr bool LAST = (NS * R == N);
    constexpr int L = LdsLen<N>::value;
    constexpr int SH = PassPad<NS, R>::value;
    float2* buf = lds + pp * L;
    if constexpr (NS > 1 && !PRE) {
        fft_seeds_load<N, P, NS, false>(s, tw, t);
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = t + b * T;
        const int k = j & (NS - 1);
        float2 x[R];
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] = v[b + r * B];
        if constexpr (NS > 1) {
            float2 pw[R];
            twiddle_powers<R, SIGN>(&s.w[fft_seeds_before<N, P>(NS) + b * seed_count<R>()], pw);
#pragma unroll
            for (int r = 1; r < R; ++r) x[r] = cmul(x[r], pw[r]);
        }
        if constexpr (HIN && NS == 1 && B == 1) dft_half_in<R, SIGN>(x);
        else Dft<R, SIGN>::run(x);
        if constexpr (LAST) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b + r * B] = x[r];
        } else {
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[padsh<SH>(base + r * NS)] = x[r];
        }
    }
    if constexpr (!LAST) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = buf[padsh<SH>(t + q * T)];
        pp ^= 1;
        fft_pass<N, P, SIGN, NS * R, PRE, false>(v, lds, tw, s, t, pp);
    }
}

// The transform of one wavefield v in the strided mapping.  tw[m] = exp(-2 pi i m/N).  PRE:
// every seed loaded at the start (their latency overlaps the first pass; 2 registers per seed
// held through it), else per pass.  HIN: the points N/2 .. N-1 (slots q >= P/2) are zero on
// input and are not read (the first pass runs half-input DFTs).
template <int N, int P, int SIGN, bool PRE = false, bool HIN = false>
__device__ __forceinline__ void fft(float2 (&v)[P], float2* lds, const float2* __restrict__ tw, int t, int& pp)
{
    FftSeeds<N, P> s;
    if constexpr (PRE) fft_seeds_load<N, P, 1, true>(s, tw, t);
    fft_pass<N, P, SIGN, 1, PRE, HIN>(v, lds, tw, s, t, pp);
}


// Wave-level variant: one 64-lane wave transforms N = 64 P points held as v[q] = x[t + 64 q].
// All exchanges stay inside the wave, whose LDS instructions execute in program order, so a
// read issued after a write observes it and one buffer (LdsLen<N> complex, private to the wave)
// suffices with no s_barrier; wave_barrier only stops the compiler from reordering the LDS
// accesses across the exchange.
template <int N, int NS>
struct WPass {
    static constexpr int P = N / 64;
    static constexpr int REM = N / NS;
    static constexpr int RMAX = P < 8 ? P : 8;
    static constexpr int R = REM >= RMAX ? RMAX : REM;
    static constexpr int B = P / R;
    static constexpr bool LAST = (NS * R == N);
    static constexpr int S = seed_count<R>();
    static constexpr int NTW = NS > 1 ? B * S : 1;  // twiddle seeds per lane (1: unused slot)
};

// The twiddle seeds of pass NS of wfft<N> (as in fft): w[b S + i] = tw[seed_exp(i) k N / (NS R)].
template <int N, int NS>
__device__ __forceinline__ void wfft_twiddles(const float2* __restrict__ tw, int t, float2 (&w)[WPass<N, NS>::NTW])
{
    using S = WPass<N, NS>;
    if constexpr (NS > 1) {
#pragma unroll
        for (int b = 0; b < S::B; ++b) {
            const int k = (t + b * 64) & (NS - 1);
            int ti = k * (N / (NS * S::R));
            asm volatile("" : "+v"(ti));  // keep the loads in the transform (see fft)
#pragma unroll
            for (int i = 0; i < S::S; ++i) w[b * S::S + i] = tw[seed_exp(i) * ti];
        }
    }
}

// One pass and the passes after it; w holds this pass's twiddles.  The next pass's twiddles are
// loaded at the start of this one, so their latency overlaps this pass's butterflies and exchange.
template <int N, int SIGN, int NS>
__device__ __forceinline__ void wfft_pass(float2 (&v)[N / 64], float2* buf, const float2* __restrict__ tw, int t,
                                          const float2 (&w)[WPass<N, NS>::NTW])
{
    using S = WPass<N, NS>;
    constexpr int P = S::P, R = S::R, B = S::B;
    constexpr int T = 64;
    constexpr int SH = PassPad<NS, R>::value;
    static_assert(P % R == 0, "radix must divide the points per lane");
    constexpr int NSN = S::LAST ? NS : NS * R;
    // prefetch only where it costs at most 7 complex registers; larger transforms load their
    // seeds in the pass that uses them
    constexpr bool PRE = WPass<N, NSN>::NTW <= 7;
    float2 wn[WPass<N, NSN>::NTW];
    if constexpr (!S::LAST && PRE) wfft_twiddles<N, NSN>(tw, t, wn);
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = t + b * T;
        const int k = j & (NS - 1);
        float2 x[R];
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] = v[b + r * B];
        if constexpr (NS > 1) {
            float2 pw[R];
            twiddle_powers<R, SIGN>(&w[b * S::S], pw);
#pragma unroll
            for (int r = 1; r < R; ++r) x[r] = cmul(x[r], pw[r]);
        }
        Dft<R, SIGN>::run(x);
        if constexpr (S::LAST) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b + r * B] = x[r];
        } else {
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[padsh<SH>(base + r * NS)] = x[r];
        }
    }
    if constexpr (!S::LAST) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = buf[padsh<SH>(t + q * T)];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if constexpr (!PRE) wfft_twiddles<N, NSN>(tw, t, wn);
        wfft_pass<N, SIGN, NSN>(v, buf, tw, t, wn);
    }
}

template <int N, int SIGN>
__device__ __forceinline__ void wfft(float2 (&v)[N / 64], float2* buf, const float2* __restrict__ tw, int t)
{
    const float2 none[1] = {make_float2(0.f, 0.f)};
    wfft_pass<N, SIGN, 1>(v, buf, tw, t, none);
}

// ------------------------------------------------------------------ mixed-radix transforms
// A length n = 2^a 3^b 5^c held in an array of M >= n points (M a power of two; T = M / P
// threads, point j = t + q T in v[q]; entries j >= n are zero on input and on output).  The
// transform runs in LDS: the points are stored in natural order; every pass of radix R (in the
// plan's order) reads butterfly b < n / R at b + r n / R, multiplies input r by
// exp(SIGN 2 pi i r k / (NS R)) with k = b mod NS, runs the R-point DFT and writes the
// self-sorting Stockham position (b - k) R + k + r NS, NS being the product of the radices before
// the pass; the points are then loaded back.  Each thread holds at most IT = ceil(P / R)
// butterflies of a pass.  PP: two buffers alternated per pass (one barrier per pass), parity `pp`
// carried as in fft; otherwise one buffer with a barrier between a pass's loads and its stores.
// Twiddles: w = tw[k n / (NS R)] from the table tw[j] = exp(-2 pi i j / n), its powers by
// products, w^r = w^h w^(r-h) with h the largest power of two below r (one load per butterfly:
// seeds as in fft, four loads per radix-8 butterfly, measured slower here, 818 -> 943 ms for a
// 250 x 250 3-D shot).  LDS indices are unpadded: the loads and the later passes' stores run at unit lane
// stride, the first pass's stores at stride R.

template <int R, int SIGN, int IT, bool PP, class Sync>
__device__ __forceinline__ void mixed_pass(const float2* src, float2* dst, const float2* __restrict__ tw, int n,
                                           int ns, int tws, uint32_t magic, int t, int T, Sync sync)
{
    const int nb = n / R;
    float2 x[IT][R];
    int base[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int b = t + i * T;
        if (i == 0 || b < nb) {
            const int k = ns > 1 ? b - (int)__umulhi((uint32_t)b, magic) * ns : 0;
            base[i] = (b - k) * R + k;
            float2 w1 = make_float2(1.f, 0.f);
            if (ns > 1 && b < nb) w1 = tw[k * tws];
#pragma unroll
            for (int r = 0; r < R; ++r) x[i][r] = b < nb ? src[b + r * nb] : make_float2(0.f, 0.f);
            if (ns > 1) {
                float2 pw[R];
                pw[1] = SIGN > 0 ? cconj(w1) : w1;
#pragma unroll
                for (int r = 2; r < R; ++r) {
                    int h = 1;
                    while (2 * h < r) h *= 2;
                    pw[r] = cmul(pw[h], pw[r - h]);
                }
#pragma unroll
                for (int r = 1; r < R; ++r) x[i][r] = cmul(x[i][r], pw[r]);
            }
            Dft<R, SIGN>::run(x[i]);
        }
    }
    if constexpr (!PP) sync();
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int b = t + i * T;
        if (b < nb) {
#pragma unroll
            for (int r = 0; r < R; ++r) dst[base[i] + r * ns] = x[i][r];
        }
    }
    sync();
}

template <int M, int P, int SIGN, bool PP, class Plan, class Sync>
__device__ __forceinline__ void mixed_fft(float2 (&v)[P], float2* buf0, float2* buf1, const float2* __restrict__ tw,
                                          const Plan& mr, int t, int& pp, Sync sync)
{
    constexpr int T = M / P;
    const int n = mr.n;
    float2* cur = (PP && pp) ? buf1 : buf0;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int j = t + q * T;
        if (j < n) cur[j] = v[q];
    }
    sync();
    for (int p = 0; p < mr.np; ++p) {
        float2* nxt = PP ? (cur == buf0 ? buf1 : buf0) : cur;
        const int ns = mr.ns[p], tws = mr.tws[p];
        const uint32_t mg = mr.magic[p];
        switch (mr.r[p]) {
        case 8: mixed_pass<8, SIGN, (P + 7) / 8, PP>(cur, nxt, tw, n, ns, tws, mg, t, T, sync); break;
        case 4: mixed_pass<4, SIGN, (P + 3) / 4, PP>(cur, nxt, tw, n, ns, tws, mg, t, T, sync); break;
        case 2: mixed_pass<2, SIGN, (P + 1) / 2, PP>(cur, nxt, tw, n, ns, tws, mg, t, T, sync); break;
        case 3: mixed_pass<3, SIGN, (P + 2) / 3, PP>(cur, nxt, tw, n, ns, tws, mg, t, T, sync); break;
        default: mixed_pass<5, SIGN, (P + 4) / 5, PP>(cur, nxt, tw, n, ns, tws, mg, t, T, sync); break;
        }
        cur = nxt;
    }
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int j = t + q * T;
        v[q] = j < n ? cur[j] : make_float2(0.f, 0.f);
    }
    if constexpr (PP) pp = cur == buf0 ? 1 : 0;  // the next transform writes the other buffer
    else sync();
}

}  // namespace swem
