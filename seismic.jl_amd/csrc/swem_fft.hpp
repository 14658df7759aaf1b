// swem_fft.hpp -- register/LDS Stockham FFT for gfx950 and complex helpers.
//
// A subgroup of T = N/P threads transforms N points.  Thread t holds its points in the
// "strided mapping" v[q] = x[t + q*T], q = 0..P-1.  Each pass of radix R (R | P) does P/R
// butterflies per thread; butterfly j = t + b*T reads x[j + r*N/R] -- exactly register slot
// b + r*P/R of the strided mapping -- twiddles them, runs an R-point DFT in registers and
// writes the self-sorting Stockham position (j/Ns)*Ns*R + j%Ns + r*Ns to LDS.  The last pass
// (Ns*R == N) writes back to the same register slots, so both the input and the output of a
// transform are in the strided mapping and the output is in natural order: forward FFT
// leaves X[k = t + q*T] in v[q].  A transform therefore costs (passes - 1) LDS round trips.
// LDS indices are padded by one complex every 8 to spread the stride-R stores over banks.
// NW wavefields are transformed together (shared twiddles, shared barriers).
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
struct Dft<1, SIGN> {
    static __device__ __forceinline__ void run(float2*) {}
};

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
        // W16^k = cos(pi k/8) + SIGN i sin(pi k/8)
        const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
        const float c2 = 0.70710678118654752440f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            float cr, ci;
            switch (k) {
            case 0: cr = 1.f; ci = 0.f; break;
            case 1: cr = c1; ci = s1; break;
            case 2: cr = c2; ci = c2; break;
            case 3: cr = s1; ci = c1; break;
            case 4: cr = 0.f; ci = 1.f; break;
            case 5: cr = -s1; ci = c1; break;
            case 6: cr = -c2; ci = c2; break;
            default: cr = -c1; ci = s1; break;
            }
            float2 w = make_float2(cr, SIGN < 0 ? -ci : ci);
            float2 t = k == 0 ? o[0] : cmul(o[k], w);
            v[k] = cadd(e[k], t);
            v[k + 8] = csub(e[k], t);
        }
    }
};

__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 3); }
template <int N>
struct LdsLen {
    static constexpr int value = N + (N >> 3);
};

// One Stockham pass and the passes after it.  v: NW wavefields in the strided mapping.
// lds: NW buffers of LdsLen<N> complex each, private to the subgroup.  tw[m] = exp(-2 pi i m/N).
template <int N, int P, int NW, int SIGN, int NS = 1>
__device__ __forceinline__ void fft(float2 (&v)[NW][P], float2* lds, const float2* __restrict__ tw, int t)
{
    constexpr int T = N / P;
    constexpr int REM = N / NS;
    constexpr int R = REM >= P ? P : REM;
    constexpr int B = P / R;
    constexpr bool LAST = (NS * R == N);
    constexpr int L = LdsLen<N>::value;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = t + b * T;
        const int k = j & (NS - 1);
        float2 w[R];
        if constexpr (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 x = tw[r * k * (N / (NS * R))];
                w[r] = SIGN > 0 ? make_float2(x.x, -x.y) : x;
            }
        }
#pragma unroll
        for (int a = 0; a < NW; ++a) {
            float2 x[R];
#pragma unroll
            for (int r = 0; r < R; ++r) x[r] = v[a][b + r * B];
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) x[r] = cmul(x[r], w[r]);
            }
            Dft<R, SIGN>::run(x);
            if constexpr (LAST) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[a][b + r * B] = x[r];
            } else {
                const int base = (j - k) * R + k;
#pragma unroll
                for (int r = 0; r < R; ++r) lds[a * L + lds_pad(base + r * NS)] = x[r];
            }
        }
    }
    if constexpr (!LAST) {
        __syncthreads();
#pragma unroll
        for (int a = 0; a < NW; ++a)
#pragma unroll
            for (int q = 0; q < P; ++q) v[a][q] = lds[a * L + lds_pad(t + q * T)];
        __syncthreads();
        fft<N, P, NW, SIGN, NS * R>(v, lds, tw, t);
    }
}

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

}  // namespace swem
