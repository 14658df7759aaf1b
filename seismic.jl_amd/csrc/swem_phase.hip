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
#include <cstdint>

#include "swem_internal.hpp"

namespace swem {

// One workgroup per table row (f, iz, ir) at a time: the row's w and p2 are loaded once and the
// entries need no index division.
__global__ __launch_bounds__(256) void build_kz_table(PhaseTableArgs a)
{
    const int rows = a.nfc * a.nz * a.nref;
    for (int r = blockIdx.x; r < rows; r += gridDim.x) {
        const int ir = r % a.nref;
        const int iz = (r / a.nref) % a.nz;
        const int f = r / (a.nref * a.nz);
        const int iw = a.ifmin + a.f0 + f;
        float2* row = a.ktab + (size_t)r * a.nkh;  // = ((f nz + iz) nref + ir) nkh
        if (a.prec == SWEM_PREC_KZ64) {
            const double w = (double)iw * a.dw;
            const double p2 = a.p2[(size_t)iz * a.nref + ir];
            for (int kk = threadIdx.x; kk < a.nkh; kk += blockDim.x) {
                const double kx = (double)kk * a.dkx;
                const double s = w * w * p2 - kx * kx;
                double re, im;
                if (s >= 0.0) {
                    sincos(sqrt(s) * a.dz, &im, &re);
                } else {
                    re = exp(-0.2 * sqrt(-s) * fabs(a.dz));
                    im = 0.0;
                }
                row[kk] = make_float2((float)(re * a.inv_nk), (float)(im * a.inv_nk));
            }
        } else {
            // wem.c:336-343 in float: kx = (float)dkx*ikx, s = (w*w)*p2 - kx*kx, no contraction.
            // The row holds kk = 0..nkx/2 and serves ikx = nkx - kk by symmetry, so the mirrored
            // wavenumbers are (float)dkx*kk rather than the reference's -((float)dkx*nkx - dkx*ikx)
            // (they differ in the last bit of kx).
#pragma clang fp contract(off)
            const float w = (float)iw * a.dwf;
            const float p2 = a.p2f[(size_t)iz * a.nref + ir];
            for (int kk = threadIdx.x; kk < a.nkh; kk += blockDim.x) {
                const float kx = a.dkxf * (float)kk;
                const float s = (w * w) * p2 - kx * kx;
                float re, im;
                if (s >= 0.f) {
                    sincosf(sqrtf(s) * (float)a.dz, &im, &re);
                } else {
                    re = expf(-0.2f * sqrtf(fabsf(s)) * fabsf((float)a.dz));
                    im = 0.f;
                }
                row[kk] = make_float2(re * (float)a.inv_nk, im * (float)a.inv_nk);
            }
        }
    }
}

// One workgroup per row (f, iz) at a time.
__global__ __launch_bounds__(256) void build_ss_table(PhaseTableArgs a)
{
    const int rows = a.nfc * a.nz;
    for (int r = blockIdx.x; r < rows; r += gridDim.x) {
        const int iz = r % a.nz;
        const int f = r / a.nz;
        const int iw = a.ifmin + a.f0 + f;
        float2* row = a.stab + (size_t)r * a.nmx;  // = (f nz + iz) nmx
        for (int x = threadIdx.x; x < a.nmx; x += blockDim.x) {
            float2 S;
            if (a.prec == SWEM_PREC_KZ64) {
                double sn, cs;
                sincos((double)iw * a.dw * a.pd[(size_t)iz * a.nmx + x] * a.dz, &sn, &cs);
                S = make_float2((float)cs, (float)sn);
            } else {
#pragma clang fp contract(off)
                float sn, cs;
                sincosf((float)iw * a.dwf * a.pdf[(size_t)iz * a.nmx + x] * (float)a.dz, &sn, &cs);
                S = make_float2(cs, sn);
            }
            row[x] = S;
        }
    }
}

int launch_phase_tables(const PhaseTableArgs& a, bool ss, hipStream_t st)
{
    const long long rows = (long long)a.nfc * a.nz * a.nref;
    if (rows < 1 || rows > INT32_MAX || a.nkh < 1 || a.nmx < 1)
        return set_error(SWEM_EINVAL, "phase tables: %lld rows", rows);
    hipLaunchKernelGGL(build_kz_table, dim3((unsigned)std::min<long long>(rows, 1 << 20)), dim3(256), 0, st, a);
    SWEM_HIP(hipGetLastError());
    if (ss) {
        const int rows_s = a.nfc * a.nz;
        hipLaunchKernelGGL(build_ss_table, dim3((unsigned)std::min(rows_s, 1 << 20)), dim3(256), 0, st, a);
        SWEM_HIP(hipGetLastError());
    }
    return SWEM_OK;
}

}  // namespace swem
