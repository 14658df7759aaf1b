// swem_esep.hpp -- the elastic separation / combination map of one wavenumber
// (elastic_separate_3d, ewem.c:553-741), shared by the 2-D and 3-D elastic kernels.
#pragma once
#include "swem_fft.hpp"

namespace swem {

// (a1, a2) -> (o1, o2): sep maps the components (ux, uz) to the potentials (up, us2), !sep maps the
// potentials back; adj selects the transposed matrices.  kx is the wavenumber as ewem.c:593 forms it
// (the matrix does not depend on ky), vp and vs the depth's layer velocities 1/po_p, 1/po_s; below
// w = 20 rad/s or where the P and S polarisations are nearly orthogonal the components pass through
// (up = uz, us2 = ux and back).
__device__ __forceinline__ void esep_k(float kx, float w, float vp, float vs, bool sep, bool adj, float2 a1,
                                       float2 a2, float2& o1, float2& o2)
{
    const float sp = w * w / (vp * vp) - kx * kx, ss = w * w / (vs * vs) - kx * kx;
    float kzp = sp > 0.f ? sqrtf(sp) : 0.f, kzs = ss > 0.f ? sqrtf(ss) : 0.f;
    const float np = sqrtf(kx * kx + kzp * kzp), ns = sqrtf(kx * kx + kzs * kzs);
    const float kxp = kx / np, kxs = kx / ns;
    kzp = kzp / np;
    kzs = kzs / ns;
    const float norm = kxp * kxs + kzp * kzs;
    if (norm >= 0.1f && w >= 20.f) {
        float m11, m12, m21, m22;
        if (sep) {
            if (!adj) m11 = kxs / norm, m12 = kzs / norm, m21 = -kzp / norm, m22 = kxp / norm;
            else m11 = kxp, m12 = kzp, m21 = -kzs, m22 = kxs;
        } else {
            if (!adj) m11 = kxp, m12 = -kzs, m21 = kzp, m22 = kxs;
            else m11 = kxs / norm, m12 = -kzp / norm, m21 = kzs / norm, m22 = kxp / norm;
        }
        o1 = cadd(cscale(a1, m11), cscale(a2, m12));
        o2 = cadd(cscale(a1, m21), cscale(a2, m22));
    } else {
        o1 = a2;
        o2 = a1;
    }
}

}  // namespace swem
