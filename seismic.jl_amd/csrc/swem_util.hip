// swem_util.hip -- band-limited time<->frequency transforms (as a tiled SGEMM against an FP64-built
// Fourier basis), the compensated image reduction and a transpose.
//
// The reference transforms every trace with an r2c / c2r FFT of length (nw-1)*2 and keeps the
// band [ifmin, ifmax) (wem.c:121-140, 177-188, f_op wem.c:480-514).  Only nf of nw bins are used
// and the transform length need not be a power of two (2000, 3000, 4000, 8000 in the
// workloads), so the band is computed directly: spec[x][f] = sum_t d[x][t] B[t][f] with
// B[t][2f] = c_f cos(2 pi iw t / n), B[t][2f+1] = -c_f sin(2 pi iw t / n), the angle reduced
// exactly in integers and evaluated in FP64 on the host.
#include <algorithm>

#include "swem_internal.hpp"

namespace swem {

// 64x64 output tile per 256-thread workgroup, 4x4 per thread, K tile 16.
__global__ __launch_bounds__(256) void sgemm_nn(int M, int Nc, int K, const float* __restrict__ A, int lda,
                                                const float* __restrict__ B, int ldb, float* __restrict__ C, int ldc)
{
    __shared__ float As[16][64 + 4];
    __shared__ float Bs[16][64 + 4];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
    for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int idx = tid + i * 256;
            const int mm = idx >> 4, kk = idx & 15;
            As[kk][mm] = (m0 + mm < M && k0 + kk < K) ? A[(size_t)(m0 + mm) * lda + k0 + kk] : 0.f;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int idx = tid + i * 256;
            const int kk = idx >> 6, nn = idx & 63;
            Bs[kk][nn] = (k0 + kk < K && n0 + nn < Nc) ? B[(size_t)(k0 + kk) * ldb + n0 + nn] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) av[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(av[i], bv[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = m0 + ty + 16 * i;
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + tx + 16 * j;
            if (n < Nc) C[(size_t)m * ldc + n] = acc[i][j];
        }
    }
}

int launch_sgemm(int M, int Nc, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
                 hipStream_t st)
{
    if (M <= 0 || Nc <= 0 || K <= 0) return SWEM_OK;
    if ((Nc + 63) / 64 > 65535 * 16 || (M + 63) / 64 > 65535) return set_error(SWEM_EINVAL, "sgemm: %d x %d too large", M, Nc);
    dim3 grid((Nc + 63) / 64, (M + 63) / 64), block(256);
    hipLaunchKernelGGL(sgemm_nn, grid, block, 0, st, M, Nc, K, A, lda, B, ldb, C, ldc);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

// out (+)= sum over (shot, group) of partial[shot][g][iz][x], Kahan-compensated, fixed order.
__global__ __launch_bounds__(256) void reduce_image(const float* __restrict__ partial, int nshot, int ngroups, int nz,
                                                    int nmx, float* __restrict__ out, int accumulate, int per_shot)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    const int s0 = per_shot ? (int)blockIdx.z : 0;
    const int s1 = per_shot ? s0 + 1 : nshot;
    if (x >= nmx) return;
    float sum = 0.f, comp = 0.f;
    auto add = [&](float v) {
        const float y = v - comp;
        const float t = sum + y;
        comp = (t - sum) - y;
        sum = t;
    };
    const size_t gs = (size_t)nz * nmx;
    for (int s = s0; s < s1; ++s) {
        const float* p = partial + ((size_t)s * ngroups * nz + iz) * nmx + x;
        int g = 0;
        // eight loads in flight ahead of the (serial, fixed-order) compensated sum
        for (; g + 8 <= ngroups; g += 8) {
            float v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = p[(size_t)(g + k) * gs];
#pragma unroll
            for (int k = 0; k < 8; ++k) add(v[k]);
        }
        for (; g < ngroups; ++g) add(p[(size_t)g * gs]);
    }
    const size_t o = per_shot ? ((size_t)s0 * nmx + x) * nz + iz : (size_t)x * nz + iz;
    out[o] = accumulate ? out[o] + sum : sum;
}

int launch_reduce_image(const float* partial, int nshot, int ngroups, int nz, int nmx, float* out, int accumulate,
                        int per_shot, hipStream_t st)
{
    dim3 grid((nmx + 255) / 256, nz, per_shot ? nshot : 1), block(256);
    hipLaunchKernelGGL(reduce_image, grid, block, 0, st, partial, nshot, ngroups, nz, nmx, out, accumulate, per_shot);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

// ----------------------------------------------------------- solver vector operations
// m1 = a m1 + b m2 (CGStep.jl:1-4)
__global__ __launch_bounds__(256) void vec_cgstep_k(float a, float* __restrict__ m1, float b,
                                                    const float* __restrict__ m2, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m1[i] = a * m1[i] + b * m2[i];
}

// y = x w elementwise (a diagonal operator: WeightingOp.jl:1-5; y may be x)
__global__ __launch_bounds__(256) void vec_mul_k(const float* x, const float* __restrict__ w, float* y, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = x[i] * w[i];
}

int launch_vec_mul(const float* x, const float* w, float* y, size_t n, hipStream_t st)
{
    if (!n) return SWEM_OK;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(vec_mul_k, dim3(blocks), dim3(256), 0, st, x, w, y, n);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

int launch_vec_cgstep(float a, float* m1, float b, const float* m2, size_t n, hipStream_t st)
{
    if (!n) return SWEM_OK;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(vec_cgstep_k, dim3(blocks), dim3(256), 0, st, a, m1, b, m2, n);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

// <x, y> in double: each of kDotBlocks blocks sums a fixed grid-stride subset and reduces it in a
// fixed tree; one block then sums the block results in order.  The result depends only on n.
constexpr int kDotBlocks = 512;

__global__ __launch_bounds__(256) void vec_dot_k(const float* __restrict__ x, const float* __restrict__ y, size_t n,
                                                 double* __restrict__ part)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        acc += (double)x[i] * (double)y[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void vec_dot_final_k(const double* __restrict__ part, int nb, double* __restrict__ out)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += part[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// part: kDotBlocks + 1 doubles of device scratch; the result lands in part[kDotBlocks].
int launch_vec_dot(const float* x, const float* y, size_t n, double* part, hipStream_t st)
{
    hipLaunchKernelGGL(vec_dot_k, dim3(kDotBlocks), dim3(256), 0, st, x, y, n, part);
    SWEM_HIP(hipGetLastError());
    hipLaunchKernelGGL(vec_dot_final_k, dim3(1), dim3(256), 0, st, part, kDotBlocks, part + kDotBlocks);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

int vec_dot_scratch_doubles() { return kDotBlocks + 1; }

// One depth of an image: out[x * ldo] (+)= sum_g partial[g][x], compensated, fixed order.
__global__ __launch_bounds__(256) void reduce_depth(const float* __restrict__ partial, int ngroups, int n,
                                                    float* __restrict__ out, int ldo, int accumulate)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    float sum = 0.f, comp = 0.f;
    for (int g = 0; g < ngroups; ++g) {
        const float y = partial[(size_t)g * n + x] - comp;
        const float t = sum + y;
        comp = (t - sum) - y;
        sum = t;
    }
    float* o = out + (size_t)x * ldo;
    *o = accumulate ? *o + sum : sum;
}

int launch_reduce_depth(const float* partial, int ngroups, int n, float* out, int ldo, int accumulate, hipStream_t st)
{
    hipLaunchKernelGGL(reduce_depth, dim3((n + 255) / 256), dim3(256), 0, st, partial, ngroups, n, out, ldo, accumulate);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

__global__ __launch_bounds__(256) void reduce_depth_shot(const float* __restrict__ partial, int ngroups, int n,
                                                         const Shot3D* __restrict__ shot, int iz, int ldo,
                                                         int acc_chunk)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    float sum = 0.f, comp = 0.f;
    for (int g = 0; g < ngroups; ++g) {
        const float y = partial[(size_t)g * n + x] - comp;
        const float t = sum + y;
        comp = (t - sum) - y;
        sum = t;
    }
    float* o = shot->out + (size_t)x * ldo + iz;
    *o = (acc_chunk || shot->acc) ? *o + sum : sum;
}

int launch_reduce_depth_shot(const float* partial, int ngroups, int n, const Shot3D* shot, int iz, int ldo,
                             int acc_chunk, hipStream_t st)
{
    hipLaunchKernelGGL(reduce_depth_shot, dim3((n + 255) / 256), dim3(256), 0, st, partial, ngroups, n, shot, iz, ldo,
                       acc_chunk);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

__global__ __launch_bounds__(64) void set_shot3d(Shot3D* dst, Shot3D v)
{
    if (threadIdx.x == 0) *dst = v;
}

int launch_set_shot3d(Shot3D* dst, const Shot3D& v, hipStream_t st)
{
    hipLaunchKernelGGL(set_shot3d, dim3(1), dim3(64), 0, st, dst, v);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

__global__ __launch_bounds__(256) void transpose_k(const float* __restrict__ in, int rows, int cols,
                                                   float* __restrict__ out)
{
    __shared__ float tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int i = ty; i < 32; i += 8) {
        const int r = r0 + i, c = c0 + tx;
        tile[i][tx] = (r < rows && c < cols) ? in[(size_t)r * cols + c] : 0.f;
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        const int c = c0 + i, r = r0 + tx;
        if (c < cols && r < rows) out[(size_t)c * rows + r] = tile[tx][i];
    }
}

int launch_transpose(const float* in, int rows, int cols, float* out, hipStream_t st)
{
    dim3 grid((cols + 31) / 32, (rows + 31) / 32), block(256);
    hipLaunchKernelGGL(transpose_k, grid, block, 0, st, in, rows, cols, out);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

// Entry (t, f): residue e = iw t mod n (exact in 64-bit integers), then the products of the host
// builder in the same order, so every float matches it.
__global__ __launch_bounds__(256) void build_bases_k(int nt, int nf, int n, int ifmin, double rnt,
                                                     const double* __restrict__ tab, float* badj, float* bsrc,
                                                     float* bfwd)
{
    const long long total = (long long)nt * nf;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int t = (int)(i / nf), f = (int)(i - (long long)t * nf);
        const int iw = ifmin + f;
        const int e = (int)(((long long)iw * t) % n);
        const double c = tab[e], sn = tab[n + e], w2s = tab[2 * n + f];
        const double cf = (iw == 0 || 2 * iw == n) ? 1.0 : 2.0;
        const size_t o = (size_t)t * 2 * nf + 2 * f;
        badj[o] = (float)(w2s * c);
        badj[o + 1] = (float)(-w2s * sn);
        bsrc[o] = (float)(rnt * c);
        bsrc[o + 1] = (float)(-rnt * sn);
        bfwd[(size_t)(2 * f) * nt + t] = (float)(cf * w2s * c);
        bfwd[(size_t)(2 * f + 1) * nt + t] = (float)(-cf * w2s * sn);
    }
}

int launch_build_bases(int nt, int nf, int n, int ifmin, double rnt, const double* tab, float* badj, float* bsrc,
                       float* bfwd, hipStream_t st)
{
    if (nt < 1 || nf < 1 || n < nt || ifmin < 0 || ifmin + nf > n)
        return set_error(SWEM_EINVAL, "bases: nt=%d nf=%d n=%d ifmin=%d", nt, nf, n, ifmin);
    const long long total = (long long)nt * nf;
    hipLaunchKernelGGL(build_bases_k, dim3((unsigned)std::min<long long>((total + 255) / 256, 1 << 16)), dim3(256),
                       0, st, nt, nf, n, ifmin, rnt, tab, badj, bsrc, bfwd);
    SWEM_HIP(hipGetLastError());
    return SWEM_OK;
}

// Integers written to device memory from a kernel argument (the values travel by value with
// the launch, so the host array may be reused as soon as the launch returns).
struct IntBlock {
    int v[kIntBlock];
};

__global__ __launch_bounds__(kIntBlock) void fill_ints(int* dst, int n, IntBlock b)
{
    if ((int)threadIdx.x < n) dst[threadIdx.x] = b.v[threadIdx.x];
}

int launch_fill_ints(int* dst, const int* vals, int n, hipStream_t st)
{
    for (int i0 = 0; i0 < n; i0 += kIntBlock) {
        IntBlock b{};
        const int m = n - i0 < kIntBlock ? n - i0 : kIntBlock;
        for (int i = 0; i < m; i++) b.v[i] = vals[i0 + i];
        hipLaunchKernelGGL(fill_ints, dim3(1), dim3(kIntBlock), 0, st, dst + i0, m, b);
        SWEM_HIP(hipGetLastError());
    }
    return SWEM_OK;
}

}  // namespace swem
