/*
 * wem_oracle.c -- CPU oracle for shot-profile WEM (TEST INFRASTRUCTURE ONLY; never shipped,
 * never called by the product path).  Discrete decisions shared by both precisions live here;
 * the precision-templated body is wem_oracle_t.h, included once for double and once for float.
 */
#define _GNU_SOURCE
#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "wem_oracle.h"

#define ORC_PI 3.141592653589793 /* PI of seismic.h:18 */

/* Sampling with the reference's casts (wem.c:79-89; taper widths wem.c:305-308). */
int orc_sampling_compute(const orc_geom* g, orc_sampling* s)
{
    if (!g || !s) return ORC_EINVAL;
    if (g->nt < 1 || g->nmx < 1 || g->nmy < 1 || g->nz < 1 || g->padt < 1 || g->padx < 1) return ORC_EINVAL;
    if (!(g->dt > 0) || !(g->dmx > 0) || (g->nmy > 1 && !(g->dmy > 0))) return ORC_EINVAL;
    /* ntfft = (int) 2*padt*((float) nt)/2 */
    float fnt = (float)(2 * g->padt) * (float)g->nt;
    fnt = fnt / 2;
    s->ntfft = (int)fnt;
    s->nw = s->ntfft / 2 + 1;
    s->fop_n = (s->nw - 1) * 2;
    if (s->fop_n < g->nt || s->fop_n < 2) return ORC_EINVAL; /* f_op would overrun (wem.c:493) */
    s->nkx = g->nmx > 1 ? g->padx * g->nmx : 1;
    s->nky = g->nmy > 1 ? g->padx * g->nmy : 1;
    s->dkx = (float)(2 * ORC_PI / ((double)(float)s->nkx) / (double)g->dmx);
    s->dky = (float)(2 * ORC_PI / ((double)(float)s->nky) / (double)g->dmy);
    s->dw = (float)(2 * ORC_PI / ((double)(float)s->ntfft) / (double)g->dt);
    /* if(fmax*dt*ntfft+1<nw) ifmax = (int) fmax*dt*ntfft + 1; else ifmax = nw; */
    float t = g->fmax * g->dt;
    t = t * (float)s->ntfft;
    t = t + 1;
    if (t < (float)s->nw) {
        float u = (float)(int)g->fmax * g->dt;
        u = u * (float)s->ntfft;
        u = u + 1;
        s->ifmax = (int)u;
    } else s->ifmax = s->nw;
    /* if(fmin*dt*ntfft+1<ifmax) ifmin = (int) fmin*dt*ntfft; else ifmin = 0; */
    t = g->fmin * g->dt;
    t = t * (float)s->ntfft;
    t = t + 1;
    if (t < (float)s->ifmax) {
        float u = (float)(int)g->fmin * g->dt;
        u = u * (float)s->ntfft;
        s->ifmin = (int)u;
    } else s->ifmin = 0;
    s->lmx = g->nmx > 100 ? 30 : 0;
    s->lmy = g->nmy > 100 ? 30 : 0;
    return ORC_OK;
}

/* igx = (int) (sx - omx)/dmx (wem.c:112-113): the cast binds to (sx - omx). */
int orc_source_index(const orc_geom* g, float sx, float sy, int* igx, int* igy)
{
    float q = (float)(int)(sx - g->omx) / g->dmx;
    *igx = (int)q;
    if (g->nmy > 1) {
        q = (float)(int)(sy - g->omy) / g->dmy;
        *igy = (int)q;
    } else {
        float dmy = g->dmy > 0 ? g->dmy : g->dmx;
        q = (float)(int)(sy - g->omy) / dmy;
        *igy = (int)q;
    }
    if (*igx < 0 || *igx >= g->nmx || *igy < 0 || *igy >= g->nmy) return ORC_EINVAL;
    return ORC_OK;
}

/* wem.c:27-36 in the reference's arithmetic: float sum, (float)nmx*nmy, 1./po in double. */
int orc_slowness_f32(const orc_geom* g, const float* vel, float* po, float* pd)
{
    int nm = g->nmx * g->nmy, nz = g->nz;
    for (int iz = 0; iz < nz; iz++) {
        float p = 0.f;
        for (int ix = 0; ix < nm; ix++) p += vel[(size_t)ix * nz + iz];
        p /= (float)g->nmx * g->nmy;
        p = (float)(1. / (double)p);
        po[iz] = p;
        for (int ix = 0; ix < nm; ix++) pd[(size_t)ix * nz + iz] = (float)(1.0 / (double)vel[(size_t)ix * nz + iz] - (double)p);
    }
    return ORC_OK;
}

/* wem.c:40-70 (reference-velocity ladder and indices) plus the branch test of wem.c:454.
 * A velocity below vmin (possible in 3-D, since vmin is searched over ix < nmx only,
 * wem.c:49) would give iref < 0 and an out-of-bounds vref read in the reference; it is
 * clamped to 0 here. */
int orc_pspi_tables(const orc_geom* g, const float* vel, float* vref, int* iref1, int* iref2, unsigned char* cut)
{
    int nmx = g->nmx, nm = g->nmx * g->nmy, nz = g->nz, nref = g->nref;
    if (nref < 2) return ORC_EINVAL; /* (float)nref-1 == 0 divides by zero at wem.c:52 */
    for (int iz = 0; iz < nz; iz++) {
        float vmin = vel[iz];
        for (int ix = 0; ix < nmx; ix++) if (vel[(size_t)ix * nz + iz] < vmin) vmin = vel[(size_t)ix * nz + iz];
        float vmax = vel[(size_t)(nmx - 1) * nz + iz];
        for (int ix = 0; ix < nm; ix++) if (vel[(size_t)ix * nz + iz] > vmax) vmax = vel[(size_t)ix * nz + iz];
        for (int ir = 0; ir < nref; ir++) {
            float q = (float)ir * (vmax - vmin);
            q = q / ((float)nref - 1);
            vref[(size_t)ir * nz + iz] = vmin + q;
        }
        for (int ix = 0; ix < nm; ix++) {
            size_t i = (size_t)ix * nz + iz;
            float v = vel[i];
            int i1, i2;
            if (vmax > vmin + 10) {
                float q = (float)(nref - 1) * (v - vmin);
                q = q / (vmax - vmin);
                int ir = (int)q;
                if (ir < 0) ir = 0;
                i1 = ir;
                i2 = ir + 1;
                if (ir > nref - 2) { i1 = nref - 1; i2 = nref - 1; }
            } else { i1 = 0; i2 = 0; }
            iref1[i] = i1;
            iref2[i] = i2;
            float v1 = vref[(size_t)i1 * nz + iz], v2 = vref[(size_t)i2 * nz + iz];
            float dv = v2 - v1;
            cut[i] = ((double)dv > 10.0) ? 1 : 0;
        }
    }
    return ORC_OK;
}

/* boundary_condition one axis (wem.c:551-558): expf(-powf(0.015*((float) l - i),2.)). */
void orc_taper_f32(int n, int l, float* t)
{
    float tv = 1.f;
    for (int i = 0; i < n; i++) {
        if (i >= 0 && i < l) tv = expf(-powf((float)(0.015 * (double)((float)l - i)), 2.f));
        if (i >= l && i <= n - l) tv = 1.f;
        if (i > n - l && i < n) tv = expf(-powf((float)(0.015 * (double)((float)i - n + l)), 2.f));
        t[i] = tv;
    }
}

void orc_taper_f64(int n, int l, double* t)
{
    for (int i = 0; i < n; i++) {
        double tv = 1.0;
        if (i < l) tv = exp(-(0.015 * (l - i)) * (0.015 * (l - i)));
        else if (i > n - l) tv = exp(-(0.015 * (i - n + l)) * (0.015 * (i - n + l)));
        t[i] = tv;
    }
}

/* ---------------------------------------------------------------- double instantiation */
#define RT double
#define SFX(x) orc64_##x
#define ORC_IS_F64 1
#define MSQRT sqrt
#define MEXP exp
#define MCEXP cexp
#define MCREAL creal
#define MCONJ conj
#define MFABS fabs
#include "wem_oracle_t.h"
#undef RT
#undef SFX
#undef ORC_IS_F64
#undef MSQRT
#undef MEXP
#undef MCEXP
#undef MCREAL
#undef MCONJ
#undef MFABS

/* ----------------------------------------------------------------- float instantiation */
#define RT float
#define SFX(x) orc32_##x
#define ORC_IS_F64 0
#define MSQRT sqrtf
#define MEXP expf
#define MCEXP cexpf
#define MCREAL crealf
#define MCONJ conjf
#define MFABS fabsf
#include "wem_oracle_t.h"
