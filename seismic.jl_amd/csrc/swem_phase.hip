// swem_phase.hip -- FP64 phase tables of the 2-D depth march, built on the GPU once per plan
// (or per frequency chunk) and shared by every shot migrated or modelled with the plan.
//
// kz table  L[f][iz][ir][kk] = exp(i sqrt(s) dz) / nk      (s >= 0)
//                              exp(-0.2 sqrt(-s) |dz|) / nk  (s < 0)
//           s = w^2 p2[iz][ir] - kx^2, kx = kk dkx, kk = 0..nkx/2 (L is even in kx)
//           (ssop wem.c:341-344 with p2 = po^2; pspiop wem.c:434-437 with p2 = 1/vref^2;
//           the two 1/sqrt(nk) of wem.c:344,354 folded in)
// SS table  S[f][iz][x] = exp(i w pd[x][iz] dz)            (ssop wem.c:326,353)
// Tables hold the +dz orientation; a -dz step uses their complex conjugate.
// KZ64: every quantity in double (s, sqrt, sincos, exp), rounded to float at the end.
// FP32: the reference's float expressions (s = (w*w)*p2 - kx*kx in float, sqrtf, sincosf).
#include <algorithm>

#include "swem_internal.hpp"

namespace swem {

__global__ __launch_bounds__(256) void build_kz_table(PhaseTableArgs a)
{
    const size_t total = (size_t)a.nfc * a.nz * a.nref * a.nkh;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int kk = (int)(i % a.nkh);
        size_t r = i / a.nkh;
        const int ir = (int)(r % a.nref);
        r /= a.nref;
        const int iz = (int)(r % a.nz);
        const int f = (int)(r / a.nz);
        const int iw = a.ifmin + a.f0 + f;
        float2 L;
        if (a.prec == SWEM_PREC_KZ64) {
            const double w = (double)iw * a.dw;
            const double kx = (double)kk * a.dkx;
            const double s = w * w * a.p2[(size_t)iz * a.nref + ir] - kx * kx;
            double re, im;
            if (s >= 0.0) {
                sincos(sqrt(s) * a.dz, &im, &re);
            } else {
                re = exp(-0.2 * sqrt(-s) * fabs(a.dz));
                im = 0.0;
            }
            L = make_float2((float)(re * a.inv_nk), (float)(im * a.inv_nk));
        } else {
            // wem.c:336-343 in float: kx = (float)dkx*ikx, s = (w*w)*p2 - kx*kx
            const float w = (float)iw * a.dwf;
            const float kx = a.dkxf * (float)kk;
            const float s = (w * w) * a.p2f[(size_t)iz * a.nref + ir] - kx * kx;
            float re, im;
            if (s >= 0.f) {
                sincosf(sqrtf(s) * (float)a.dz, &im, &re);
            } else {
                re = expf(-0.2f * sqrtf(fabsf(s)) * fabsf((float)a.dz));
                im = 0.f;
            }
            L = make_float2(re * (float)a.inv_nk, im * (float)a.inv_nk);
        }
        a.ktab[i] = L;
    }
}

__global__ __launch_bounds__(256) void build_ss_table(PhaseTableArgs a)
{
    const size_t total = (size_t)a.nfc * a.nz * a.nmx;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % a.nmx);
        size_t r = i / a.nmx;
        const int iz = (int)(r % a.nz);
        const int f = (int)(r / a.nz);
        const int iw = a.ifmin + a.f0 + f;
        float2 S;
        if (a.prec == SWEM_PREC_KZ64) {
            double sn, cs;
            sincos((double)iw * a.dw * a.pd[(size_t)iz * a.nmx + x] * a.dz, &sn, &cs);
            S = make_float2((float)cs, (float)sn);
        } else {
            float sn, cs;
            sincosf((float)iw * a.dwf * a.pdf[(size_t)iz * a.nmx + x] * (float)a.dz, &sn, &cs);
            S = make_float2(cs, sn);
        }
        a.stab[i] = S;
    }
}

int launch_phase_tables(const PhaseTableArgs& a, bool ss, hipStream_t st)
{
    const size_t nk = (size_t)a.nfc * a.nz * a.nref * a.nkh;
    const int blocks = (int)std::min<size_t>((nk + 255) / 256, 65536);
    hipLaunchKernelGGL(build_kz_table, dim3(blocks), dim3(256), 0, st, a);
    SWEM_HIP(hipGetLastError());
    if (ss) {
        const size_t ns = (size_t)a.nfc * a.nz * a.nmx;
        const int b2 = (int)std::min<size_t>((ns + 255) / 256, 65536);
        hipLaunchKernelGGL(build_ss_table, dim3(b2), dim3(256), 0, st, a);
        SWEM_HIP(hipGetLastError());
    }
    return SWEM_OK;
}

}  // namespace swem
