// swem_tables.cpp -- discrete decisions of wem() with the reference's float/int casts.
// Every expression below is evaluated in float exactly where src/Imaging/c/wem.c does, one
// operation per statement so no contraction or excess precision can change a branch.
// Built with -ffp-contract=off (host compiler), so each float statement rounds as written.
#include <cmath>
#include <cstring>

#include <vector>

#include "swem_internal.hpp"

namespace swem {

static const double kPi = 3.141592653589793;  // PI of seismic.h:18

// wem.c:79-89 and the taper widths of ssop (wem.c:305-308).
int compute_sampling(const swem_geometry& g, Sampling& s)
{
    if (g.nt < 1 || g.nmx < 1 || g.nmy < 1 || g.nz < 1 || g.padt < 1 || g.padx < 1)
        return set_error(SWEM_EINVAL, "sizes must be positive (nt=%d nmx=%d nmy=%d nz=%d padt=%d padx=%d)", g.nt, g.nmx,
                         g.nmy, g.nz, g.padt, g.padx);
    if (!(g.dt > 0.f) || !(g.dmx > 0.f)) return set_error(SWEM_EINVAL, "dt and dmx must be positive");
    if (g.nmy > 1 && !(g.dmy > 0.f)) return set_error(SWEM_EINVAL, "dmy must be positive when nmy > 1");
    // ntfft = (int) 2*padt*((float) nt)/2
    float fnt = (float)(2 * g.padt) * (float)g.nt;
    fnt = fnt / 2;
    s.ntfft = (int)fnt;
    s.nw = s.ntfft / 2 + 1;
    s.fop_n = (s.nw - 1) * 2;  // f_op's own length (wem.c:487)
    if (s.fop_n < g.nt || s.fop_n < 2)
        return set_error(SWEM_EINVAL, "padt*nt=%d is too short for nt=%d (the reference f_op would overrun)", s.ntfft, g.nt);
    s.nkx = g.nmx > 1 ? g.padx * g.nmx : 1;
    s.nky = g.nmy > 1 ? g.padx * g.nmy : 1;
    s.dkx = (float)(2 * kPi / ((double)(float)s.nkx) / (double)g.dmx);
    s.dky = (float)(2 * kPi / ((double)(float)s.nky) / (double)(g.dmy > 0.f ? g.dmy : g.dmx));
    s.dw = (float)(2 * kPi / ((double)(float)s.ntfft) / (double)g.dt);
    float t = g.fmax * g.dt;
    t = t * (float)s.ntfft;
    t = t + 1;
    if (t < (float)s.nw) {
        float u = (float)(int)g.fmax * g.dt;
        u = u * (float)s.ntfft;
        u = u + 1;
        s.ifmax = (int)u;
    } else
        s.ifmax = s.nw;
    t = g.fmin * g.dt;
    t = t * (float)s.ntfft;
    t = t + 1;
    if (t < (float)s.ifmax) {
        float u = (float)(int)g.fmin * g.dt;
        u = u * (float)s.ntfft;
        s.ifmin = (int)u;
    } else
        s.ifmin = 0;
    s.lmx = g.nmx > 100 ? 30 : 0;
    s.lmy = g.nmy > 100 ? 30 : 0;
    return SWEM_OK;
}

// igx = (int) (sx - omx)/dmx (wem.c:112-113): the cast binds to (sx - omx).
int source_index(const swem_geometry& g, float sx, float sy, int& igx, int& igy)
{
    float q = (float)(int)(sx - g.omx) / g.dmx;
    igx = (int)q;
    const float dmy = g.dmy > 0.f ? g.dmy : g.dmx;
    q = (float)(int)(sy - g.omy) / dmy;
    igy = (int)q;
    if (igx < 0 || igx >= g.nmx || igy < 0 || igy >= g.nmy)
        return set_error(SWEM_EINVAL, "source (sx=%g, sy=%g) maps to (igx=%d, igy=%d) outside the %dx%d grid", sx, sy,
                         igx, igy, g.nmx, g.nmy);
    return SWEM_OK;
}

// wem.c:27-36 in float (naive-FP32 tables): float sum, (float)nmx*nmy, 1./po in double.  The
// traces are read in storage order with one accumulator per depth, so each depth's sum runs over
// ix = 0, 1, ... in the reference's order.
void slowness_f32(const swem_geometry& g, const float* vel, float* po, float* pd)
{
    const int nm = g.nmx * g.nmy, nz = g.nz;
    std::vector<float> sum(nz, 0.f);
    for (int ix = 0; ix < nm; ix++) {
        const float* v = vel + (size_t)ix * nz;
        for (int iz = 0; iz < nz; iz++) sum[iz] += v[iz];
    }
    for (int iz = 0; iz < nz; iz++) {
        float p = sum[iz];
        p /= (float)g.nmx * g.nmy;
        po[iz] = (float)(1. / (double)p);
    }
    for (int ix = 0; ix < nm; ix++) {
        const float* v = vel + (size_t)ix * nz;
        float* d = pd + (size_t)ix * nz;
        for (int iz = 0; iz < nz; iz++) d[iz] = (float)(1.0 / (double)v[iz] - (double)po[iz]);
    }
}

// wem.c:40-70 and the branch test of wem.c:454.  A velocity below vmin (3-D only: vmin is
// searched over the first nmx traces, wem.c:49) would index vref[-1] in the reference; the
// index is clamped to 0.
int pspi_tables(const swem_geometry& g, const float* vel, float* vref, int* iref1, int* iref2, uint8_t* cut)
{
    const int nmx = g.nmx, nm = g.nmx * g.nmy, nz = g.nz, nref = g.nref;
    if (nref < 2) return set_error(SWEM_EINVAL, "PSPI needs nref >= 2 (nref=%d divides by zero at wem.c:52)", nref);
    if (nref > 31) return set_error(SWEM_EUNSUPPORTED, "nref=%d > 31", nref);
    // per depth: vmin over the first nmx traces, vmax over all (traces in storage order)
    std::vector<float> vmin(vel, vel + nz), vmax(vel + (size_t)(nmx - 1) * nz, vel + (size_t)nmx * nz);
    for (int ix = 0; ix < nmx; ix++) {
        const float* v = vel + (size_t)ix * nz;
        for (int iz = 0; iz < nz; iz++)
            if (v[iz] < vmin[iz]) vmin[iz] = v[iz];
    }
    for (int ix = 0; ix < nm; ix++) {
        const float* v = vel + (size_t)ix * nz;
        for (int iz = 0; iz < nz; iz++)
            if (v[iz] > vmax[iz]) vmax[iz] = v[iz];
    }
    for (int iz = 0; iz < nz; iz++)
        for (int ir = 0; ir < nref; ir++) {
            float q = (float)ir * (vmax[iz] - vmin[iz]);
            q = q / ((float)nref - 1);
            vref[(size_t)ir * nz + iz] = vmin[iz] + q;
        }
    for (int ix = 0; ix < nm; ix++)
        for (int iz = 0; iz < nz; iz++) {
            const size_t i = (size_t)ix * nz + iz;
            const float v = vel[i], lo = vmin[iz], hi = vmax[iz];
            int i1, i2;
            if (hi > lo + 10) {
                float q = (float)(nref - 1) * (v - lo);
                q = q / (hi - lo);
                int ir = (int)q;
                if (ir < 0) ir = 0;
                i1 = ir;
                i2 = ir + 1;
                if (ir > nref - 2) {
                    i1 = nref - 1;
                    i2 = nref - 1;
                }
            } else {
                i1 = 0;
                i2 = 0;
            }
            if (iref1) iref1[i] = i1;
            if (iref2) iref2[i] = i2;
            const float dv = vref[(size_t)i2 * nz + iz] - vref[(size_t)i1 * nz + iz];
            if (cut) cut[i] = ((double)dv > 10.0) ? 1 : 0;
        }
    return SWEM_OK;
}

// boundary_condition along one axis (wem.c:551-558).
void taper_f32(int n, int l, float* t)
{
    float tv = 1.f;
    for (int i = 0; i < n; i++) {
        if (i >= 0 && i < l) {
            const float a = (float)(0.015 * (double)((float)l - (float)i));
            tv = expf(-powf(a, 2.f));
        }
        if (i >= l && i <= n - l) tv = 1.f;
        if (i > n - l && i < n) {
            const float a = (float)(0.015 * (double)((float)i - (float)n + (float)l));
            tv = expf(-powf(a, 2.f));
        }
        t[i] = tv;
    }
}

// Passes of the mixed-radix transform of n: radix 5 while 5 divides the rest, then 3, 8, 4, 2
// (the odd radices first: their passes store at an odd lane stride, which spreads over the LDS
// banks; measured 3-5 % faster than 8, 4, 2, 5, 3 at n = 500, 2000 and 3000).
bool mixed_radix_plan(int n, MixedRadix* mr)
{
    if (n < 1 || n > 4096) return false;
    MixedRadix m{};
    m.n = n;
    int rest = n, ns = 1;
    static const int radices[] = {5, 3, 8, 4, 2};
    for (int r : radices)
        while (rest % r == 0) {
            if (m.np == kMixedMaxPass) return false;
            m.r[m.np] = (uint8_t)r;
            m.ns[m.np] = (uint16_t)ns;
            m.tws[m.np] = (uint16_t)(n / (ns * r));
            m.magic[m.np] = ns > 1 ? (uint32_t)((1ull << 32) / (unsigned)ns + 1) : 0u;
            ns *= r;
            rest /= r;
            m.np++;
        }
    if (rest != 1) return false;
    if (mr) *mr = m;
    return true;
}

}  // namespace swem
