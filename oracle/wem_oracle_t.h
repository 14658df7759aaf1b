/*
 * wem_oracle_t.h -- precision-templated body of the CPU oracle (test infrastructure only).
 * Included twice by wem_oracle.c: RT=double (SFX -> orc64_) and RT=float (SFX -> orc32_).
 * Each function cites the part of src/Imaging/c/wem.c it restates.
 */
#define CT RT _Complex

typedef struct SFX(ctx) {
    const orc_geom* g;
    orc_sampling s;
    int nm, nk;
    RT* po;      /* [nz]      layer slowness (wem.c:30-34) */
    RT* pd;      /* [nm][nz]  slowness perturbation (wem.c:35) */
    RT* tx;      /* [nmx]     taper along x (wem.c:552-554) */
    RT* ty;      /* [nmy]     taper along y (wem.c:556-558) */
    RT* kx2;     /* [nkx]     kx*kx (wem.c:336-337) */
    RT* ky2;     /* [nky]     ky*ky (wem.c:339-340) */
    RT* vref;    /* [nref][nz] reference velocities (wem.c:52) */
    RT* w2;      /* [nm][nz]  weight of iref2 for pspi_mode 1 */
    int* iref1;  /* [nm][nz]  (wem.c:57-61) */
    int* iref2;
    unsigned char* cut; /* [nm][nz] vref2 - vref1 > 10 (wem.c:454) */
    CT* twx;     /* twiddles exp(-2 pi i j / nkx) */
    CT* twy;
} SFX(ctx);

typedef struct SFX(work) {
    CT* a;       /* [nk] */
    CT* b;       /* [nk] */
    CT* acc;     /* [nk] */
    CT* col;     /* [max(nkx,nky)] column gather */
    CT* col2;
    CT* dref;    /* [nm][nref] (wem.c:398,407) */
} SFX(work);

/* ---------------------------------------------------------------- FFT (own, mixed radix) */

static void SFX(twiddles)(CT* tw, int n)
{
    for (int j = 0; j < n; j++) {
        double ang = -2.0 * 3.14159265358979323846 * (double)j / (double)n;
        tw[j] = (RT)cos(ang) + (RT)sin(ang) * (RT _Complex)_Complex_I;
    }
}

static int SFX(small_factor)(int n)
{
    if (n % 4 == 0) return 4;
    for (int p = 2; p * p <= n; p++) if (n % p == 0) return p;
    return n;
}

/* out[k] = sum_j in[j*stride] * exp(sign*2*pi*i*j*k/n); tw holds exp(-2 pi i e/N), n | N.
 * Recursive decimation in time: X[k + r m] = sum_q W_n^{qk} W_p^{qr} Y_q[k]. */
static void SFX(fft_rec)(const CT* in, int stride, CT* out, int n, const CT* tw, int N, int sign)
{
    if (n == 1) { out[0] = in[0]; return; }
    int p = SFX(small_factor)(n);
    int m = n / p;
    for (int q = 0; q < p; q++) SFX(fft_rec)(in + (size_t)q * stride, stride * p, out + (size_t)q * m, m, tw, N, sign);
    int ts = N / n;
    CT t[p];
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < p; q++) {
            CT w = tw[(size_t)((q * k) % n) * ts];
            if (sign > 0) w = MCONJ(w);
            t[q] = out[(size_t)q * m + k] * w;
        }
        for (int r = 0; r < p; r++) {
            CT acc = 0;
            for (int q = 0; q < p; q++) {
                CT w = tw[(size_t)(((q * r) % p) * m) * ts];
                if (sign > 0) w = MCONJ(w);
                acc += t[q] * w;
            }
            out[(size_t)r * m + k] = acc;
        }
    }
}

/* In-place 1-D transform of length n with a scratch copy. */
static void SFX(fft1)(CT* x, int n, const CT* tw, int sign, CT* scratch)
{
    if (n == 1) return;
    memcpy(scratch, x, sizeof(CT) * (size_t)n);
    SFX(fft_rec)(scratch, 1, x, n, tw, n, sign);
}

/* 2-D transform of a[ikx*nky + iky], the layout of fftwf_plan_dft(2, {nkx,nky}) (wem.c:100-104). */
static void SFX(fft2)(const SFX(ctx)* c, CT* a, int sign, SFX(work)* w)
{
    int nkx = c->s.nkx, nky = c->s.nky;
    if (nky > 1)
        for (int ix = 0; ix < nkx; ix++) SFX(fft1)(a + (size_t)ix * nky, nky, c->twy, sign, w->col2);
    if (nkx > 1) {
        if (nky == 1) { SFX(fft1)(a, nkx, c->twx, sign, w->col2); return; }
        for (int iy = 0; iy < nky; iy++) {
            for (int ix = 0; ix < nkx; ix++) w->col[ix] = a[(size_t)ix * nky + iy];
            SFX(fft1)(w->col, nkx, c->twx, sign, w->col2);
            for (int ix = 0; ix < nkx; ix++) a[(size_t)ix * nky + iy] = w->col[ix];
        }
    }
}

int SFX(fft)(RT* x, int n, int sign)
{
    if (n < 1) return ORC_EINVAL;
    CT* tw = (CT*)malloc(sizeof(CT) * (size_t)n);
    CT* sc = (CT*)malloc(sizeof(CT) * (size_t)n);
    if (!tw || !sc) { free(tw); free(sc); return ORC_ENOMEM; }
    SFX(twiddles)(tw, n);
    SFX(fft1)((CT*)x, n, tw, sign, sc);
    free(tw); free(sc);
    return ORC_OK;
}

/* f_op (wem.c:480-514): length (nw-1)*2 r2c / c2r, unnormalised.  c2r uses Re of the DC and
 * Nyquist bins only, the Hermitian reconstruction FFTW's c2r performs. */
int SFX(fop)(RT* mi, RT* d, int nw, int nt, int adj)
{
    int n = (nw - 1) * 2;
    if (n < 2 || nt > n) return ORC_EINVAL;
    CT* m = (CT*)mi;
    CT* tw = (CT*)malloc(sizeof(CT) * (size_t)n);
    CT* x = (CT*)malloc(sizeof(CT) * (size_t)n);
    CT* sc = (CT*)malloc(sizeof(CT) * (size_t)n);
    if (!tw || !x || !sc) { free(tw); free(x); free(sc); return ORC_ENOMEM; }
    SFX(twiddles)(tw, n);
    if (adj) {
        for (int it = 0; it < n; it++) x[it] = it < nt ? d[it] : 0;
        SFX(fft1)(x, n, tw, -1, sc);
        for (int iw = 0; iw < nw; iw++) m[iw] = x[iw];
    } else {
        x[0] = MCREAL(m[0]);
        for (int iw = 1; iw < nw - 1; iw++) { x[iw] = m[iw]; x[n - iw] = MCONJ(m[iw]); }
        x[n / 2] = MCREAL(m[nw - 1]);
        SFX(fft1)(x, n, tw, +1, sc);
        for (int it = 0; it < nt; it++) d[it] = MCREAL(x[it]);
    }
    free(tw); free(x); free(sc);
    return ORC_OK;
}

/* --------------------------------------------------------------- tables of one wem() call */

static void SFX(ctx_free)(SFX(ctx)* c)
{
    free(c->po); free(c->pd); free(c->tx); free(c->ty); free(c->kx2); free(c->ky2);
    free(c->vref); free(c->w2); free(c->iref1); free(c->iref2); free(c->cut);
    free(c->twx); free(c->twy);
}

static int SFX(ctx_init)(SFX(ctx)* c, const orc_geom* g, const float* vel)
{
    memset(c, 0, sizeof(*c));
    c->g = g;
    int rc = orc_sampling_compute(g, &c->s);
    if (rc) return rc;
    int nmx = g->nmx, nmy = g->nmy, nz = g->nz, nm = nmx * nmy;
    c->nm = nm;
    c->nk = c->s.nkx * c->s.nky;
    c->po = (RT*)malloc(sizeof(RT) * nz);
    c->pd = (RT*)malloc(sizeof(RT) * (size_t)nm * nz);
    c->tx = (RT*)malloc(sizeof(RT) * nmx);
    c->ty = (RT*)malloc(sizeof(RT) * nmy);
    c->kx2 = (RT*)malloc(sizeof(RT) * c->s.nkx);
    c->ky2 = (RT*)malloc(sizeof(RT) * c->s.nky);
    c->twx = (CT*)malloc(sizeof(CT) * c->s.nkx);
    c->twy = (CT*)malloc(sizeof(CT) * c->s.nky);
    int nref = g->nref > 0 ? g->nref : 1;
    c->vref = (RT*)malloc(sizeof(RT) * (size_t)nref * nz);
    c->w2 = (RT*)calloc((size_t)nm * nz, sizeof(RT));
    c->iref1 = (int*)calloc((size_t)nm * nz, sizeof(int));
    c->iref2 = (int*)calloc((size_t)nm * nz, sizeof(int));
    c->cut = (unsigned char*)calloc((size_t)nm * nz, 1);
    if (!c->po || !c->pd || !c->tx || !c->ty || !c->kx2 || !c->ky2 || !c->twx || !c->twy ||
        !c->vref || !c->w2 || !c->iref1 || !c->iref2 || !c->cut) { SFX(ctx_free)(c); return ORC_ENOMEM; }

    /* slowness split, wem.c:27-36 */
#if ORC_IS_F64
    for (int iz = 0; iz < nz; iz++) {
        double sum = 0.0;
        for (int ix = 0; ix < nm; ix++) sum += (double)vel[(size_t)ix * nz + iz];
        double po = 1.0 / (sum / (double)nm);
        c->po[iz] = po;
        for (int ix = 0; ix < nm; ix++) c->pd[(size_t)ix * nz + iz] = 1.0 / (double)vel[(size_t)ix * nz + iz] - po;
    }
    orc_taper_f64(nmx, c->s.lmx, c->tx);
    orc_taper_f64(nmy, c->s.lmy, c->ty);
    /* kx = 2 pi ikx/(nkx dmx) for ikx <= nkx/2, else -2 pi (nkx-ikx)/(nkx dmx) (wem.c:336-337) */
    for (int ik = 0; ik < c->s.nkx; ik++) {
        double kx = ik <= c->s.nkx / 2 ? 2.0 * M_PI * ik / ((double)c->s.nkx * g->dmx)
                                      : -2.0 * M_PI * (c->s.nkx - ik) / ((double)c->s.nkx * g->dmx);
        c->kx2[ik] = kx * kx;
    }
    for (int ik = 0; ik < c->s.nky; ik++) {
        double ky = ik <= c->s.nky / 2 ? 2.0 * M_PI * ik / ((double)c->s.nky * g->dmy)
                                      : -2.0 * M_PI * (c->s.nky - ik) / ((double)c->s.nky * g->dmy);
        c->ky2[ik] = ky * ky;
    }
#else
    {
        float* pof = (float*)malloc(sizeof(float) * nz);
        float* pdf = (float*)malloc(sizeof(float) * (size_t)nm * nz);
        if (!pof || !pdf) { free(pof); free(pdf); SFX(ctx_free)(c); return ORC_ENOMEM; }
        orc_slowness_f32(g, vel, pof, pdf);
        for (int iz = 0; iz < nz; iz++) c->po[iz] = pof[iz];
        for (size_t i = 0; i < (size_t)nm * nz; i++) c->pd[i] = pdf[i];
        free(pof); free(pdf);
    }
    orc_taper_f32(nmx, c->s.lmx, c->tx);
    orc_taper_f32(nmy, c->s.lmy, c->ty);
    for (int ik = 0; ik < c->s.nkx; ik++) {
        float kx;
        if (ik <= (int)c->s.nkx / 2) kx = (float)c->s.dkx * ik;
        else kx = -((float)c->s.dkx * c->s.nkx - c->s.dkx * ik);
        c->kx2[ik] = kx * kx;
    }
    for (int ik = 0; ik < c->s.nky; ik++) {
        float ky;
        if (ik <= (int)c->s.nky / 2) ky = (float)c->s.dky * ik;
        else ky = -((float)c->s.dky * c->s.nky - c->s.dky * ik);
        c->ky2[ik] = ky * ky;
    }
#endif
    SFX(twiddles)(c->twx, c->s.nkx);
    SFX(twiddles)(c->twy, c->s.nky);

    if (g->pspi) {
        if (g->nref < 2) { SFX(ctx_free)(c); return ORC_EINVAL; }
        float* vreff = (float*)malloc(sizeof(float) * (size_t)g->nref * nz);
        if (!vreff) { SFX(ctx_free)(c); return ORC_ENOMEM; }
        rc = orc_pspi_tables(g, vel, vreff, c->iref1, c->iref2, c->cut);
        if (rc) { free(vreff); SFX(ctx_free)(c); return rc; }
        for (int iz = 0; iz < nz; iz++) {
#if ORC_IS_F64
            /* vmin, vmax as the reference finds them (float); the ladder in double */
            double vmin = vreff[0 * nz + iz], vmax = vreff[(size_t)(g->nref - 1) * nz + iz];
            for (int ir = 0; ir < g->nref; ir++)
                c->vref[(size_t)ir * nz + iz] = vmin + (double)ir * (vmax - vmin) / ((double)g->nref - 1.0);
#else
            for (int ir = 0; ir < g->nref; ir++) c->vref[(size_t)ir * nz + iz] = vreff[(size_t)ir * nz + iz];
#endif
        }
        /* linear-interpolation weight of iref2 (the formula at wem.c:476) */
        for (int ix = 0; ix < nm; ix++)
            for (int iz = 0; iz < nz; iz++) {
                size_t i = (size_t)ix * nz + iz;
                if (c->cut[i]) {
                    RT v1 = c->vref[(size_t)c->iref1[i] * nz + iz], v2 = c->vref[(size_t)c->iref2[i] * nz + iz];
                    c->w2[i] = ((RT)vel[i] - v1) / (v2 - v1);
                }
            }
        free(vreff);
    }
    return ORC_OK;
}

static int SFX(work_init)(SFX(work)* w, const SFX(ctx)* c)
{
    int nkmax = c->s.nkx > c->s.nky ? c->s.nkx : c->s.nky;
    int nref = c->g->nref > 0 ? c->g->nref : 1;
    w->a = (CT*)malloc(sizeof(CT) * c->nk);
    w->b = (CT*)malloc(sizeof(CT) * c->nk);
    w->acc = (CT*)malloc(sizeof(CT) * c->nk);
    w->col = (CT*)malloc(sizeof(CT) * nkmax);
    w->col2 = (CT*)malloc(sizeof(CT) * nkmax);
    w->dref = (CT*)malloc(sizeof(CT) * (size_t)c->nm * nref);
    if (!w->a || !w->b || !w->acc || !w->col || !w->col2 || !w->dref) return ORC_ENOMEM;
    return ORC_OK;
}

static void SFX(work_free)(SFX(work)* w)
{
    free(w->a); free(w->b); free(w->acc); free(w->col); free(w->col2); free(w->dref);
}

/* ------------------------------------------------------------------- one depth step */

/* Angular frequency of bin iw: w = iw*dw (wem.c:228). */
static RT SFX(omega)(const SFX(ctx)* c, int iw)
{
#if ORC_IS_F64
    return 2.0 * M_PI * (double)iw / ((double)c->s.ntfft * (double)c->g->dt);
#else
    return iw * c->s.dw;
#endif
}

/* boundary_condition (wem.c:546-564) */
static void SFX(taper)(const SFX(ctx)* c, CT* u)
{
    int nmx = c->g->nmx, nmy = c->g->nmy;
    for (int ix = 0; ix < nmx; ix++)
        for (int iy = 0; iy < nmy; iy++) u[(size_t)ix * nmy + iy] *= c->tx[ix] * c->ty[iy];
}

/* kz phase shift of one wavenumber (wem.c:341-343 for SSF with p2 = po^2, wem.c:434-436
 * for PSPI with p2 = 1/vref^2). */
static CT SFX(kzphase)(RT w, RT p2, RT k2x, RT k2y, RT dzs)
{
    RT s = (w * w) * p2 - k2x - k2y;
    if (s >= 0) return MCEXP(MSQRT(s) * dzs * (RT _Complex)_Complex_I);
    return (CT)MEXP((RT)(-0.2) * MSQRT(MFABS(s)) * MFABS(dzs));
}

/* ssop (wem.c:289-378) */
static void SFX(ssop)(const SFX(ctx)* c, CT* u, RT w, int iz, RT dzs, int adj, SFX(work)* wk)
{
    int nmx = c->g->nmx, nmy = c->g->nmy, nkx = c->s.nkx, nky = c->s.nky, nz = c->g->nz;
    CT* a = wk->a;
#if ORC_IS_F64
    RT rnk = 1.0 / sqrt((double)c->nk);
#else
    RT rnk = 1.0f / sqrtf((float)nkx * nky);
#endif
    if (!adj) SFX(taper)(c, u);
    for (int ix = 0; ix < nkx; ix++)
        for (int iy = 0; iy < nky; iy++) {
            size_t k = (size_t)ix * nky + iy;
            if (ix < nmx && iy < nmy) {
                size_t x = (size_t)ix * nmy + iy;
                if (adj) a[k] = u[x];
                else a[k] = u[x] * MCEXP(w * c->pd[x * nz + iz] * dzs * (RT _Complex)_Complex_I);
            } else a[k] = 0;
        }
    SFX(fft2)(c, a, -1, wk);
    RT p2 = c->po[iz] * c->po[iz];
    for (int ix = 0; ix < nkx; ix++)
        for (int iy = 0; iy < nky; iy++) {
            size_t k = (size_t)ix * nky + iy;
            a[k] = a[k] * SFX(kzphase)(w, p2, c->kx2[ix], c->ky2[iy], dzs) * rnk;
        }
    SFX(fft2)(c, a, +1, wk);
    for (int ix = 0; ix < nmx; ix++)
        for (int iy = 0; iy < nmy; iy++) {
            size_t k = (size_t)ix * nky + iy, x = (size_t)ix * nmy + iy;
            if (adj) u[x] = a[k] * MCEXP(w * c->pd[x * nz + iz] * dzs * (RT _Complex)_Complex_I) * rnk;
            else u[x] = a[k] * rnk;
        }
    if (adj) SFX(taper)(c, u);
}

/* pspiop (wem.c:380-471) with linear_interp (wem.c:473-478).  select_after=1 is the
 * reference's structure (propagate with every vref, then pick per x).  select_after=0 is
 * its exact adjoint for pspi_mode 1 (weight per x, propagate, sum). */
static void SFX(pspiop)(const SFX(ctx)* c, CT* u, RT w, int iz, RT dzs, int select_after, SFX(work)* wk)
{
    const orc_geom* g = c->g;
    int nmx = g->nmx, nmy = g->nmy, nkx = c->s.nkx, nky = c->s.nky, nz = g->nz, nref = g->nref;
    int nm = c->nm;
    CT* a = wk->a;
    CT* b = wk->b;
#if ORC_IS_F64
    RT rnk = 1.0 / sqrt((double)c->nk);
#else
    RT rnk = 1.0f / sqrtf((float)nkx * nky);
#endif
    if (select_after) {
        for (int ix = 0; ix < nkx; ix++)
            for (int iy = 0; iy < nky; iy++) {
                size_t k = (size_t)ix * nky + iy;
                a[k] = (ix < nmx && iy < nmy) ? u[(size_t)ix * nmy + iy] : 0;
            }
        SFX(fft2)(c, a, -1, wk);
        for (int ir = 0; ir < nref; ir++) {
            RT vr = c->vref[(size_t)ir * nz + iz];
#if ORC_IS_F64
            RT p2 = 1.0 / (vr * vr);
#else
            RT p2 = 1 / (vr * vr);
#endif
            for (int ix = 0; ix < nkx; ix++)
                for (int iy = 0; iy < nky; iy++) {
                    size_t k = (size_t)ix * nky + iy;
                    b[k] = a[k] * SFX(kzphase)(w, p2, c->kx2[ix], c->ky2[iy], dzs) * rnk;
                }
            SFX(fft2)(c, b, +1, wk);
            for (int ix = 0; ix < nmx; ix++)
                for (int iy = 0; iy < nmy; iy++)
                    wk->dref[((size_t)ix * nmy + iy) * nref + ir] = b[(size_t)ix * nky + iy] * rnk;
        }
        for (int x = 0; x < nm; x++) {
            size_t i = (size_t)x * nz + iz;
            CT y1 = wk->dref[(size_t)x * nref + c->iref1[i]];
            if (c->cut[i]) {
                if (g->pspi_mode == 0) u[x] = MCREAL(y1);   /* float linear_interp returns y1 (wem.c:477) */
                else {
                    CT y2 = wk->dref[(size_t)x * nref + c->iref2[i]];
                    u[x] = ((RT)1 - c->w2[i]) * y1 + c->w2[i] * y2;
                }
            } else u[x] = y1;
        }
    } else {
        CT* acc = wk->acc;
        for (int k = 0; k < c->nk; k++) acc[k] = 0;
        for (int ir = 0; ir < nref; ir++) {
            int any = 0;
            for (int ix = 0; ix < nkx; ix++)
                for (int iy = 0; iy < nky; iy++) {
                    size_t k = (size_t)ix * nky + iy;
                    CT v = 0;
                    if (ix < nmx && iy < nmy) {
                        size_t x = (size_t)ix * nmy + iy, i = x * nz + iz;
                        RT wt = 0;
                        if (c->cut[i]) {
                            if (c->iref1[i] == ir) wt += (RT)1 - c->w2[i];
                            if (c->iref2[i] == ir) wt += c->w2[i];
                        } else if (c->iref1[i] == ir) wt = 1;
                        v = u[x] * wt;
                        if (wt != 0) any = 1;
                    }
                    a[k] = v;
                }
            if (!any) continue;
            SFX(fft2)(c, a, -1, wk);
            RT vr = c->vref[(size_t)ir * nz + iz];
            RT p2 = (RT)1 / (vr * vr);
            for (int ix = 0; ix < nkx; ix++)
                for (int iy = 0; iy < nky; iy++) {
                    size_t k = (size_t)ix * nky + iy;
                    acc[k] += a[k] * SFX(kzphase)(w, p2, c->kx2[ix], c->ky2[iy], dzs) * rnk;
                }
        }
        SFX(fft2)(c, acc, +1, wk);
        for (int ix = 0; ix < nmx; ix++)
            for (int iy = 0; iy < nmy; iy++) u[(size_t)ix * nmy + iy] = acc[(size_t)ix * nky + iy] * rnk;
    }
}

/* One step of either extrapolator, as extrap1f calls them (wem.c:237-242, 258-259, 273-274). */
static void SFX(step_ctx)(const SFX(ctx)* c, CT* u, RT w, int iz, RT dzs, int adj, SFX(work)* wk)
{
    if (c->g->pspi) {
        int select_after = (c->g->pspi_mode == 0) ? 1 : adj;
        SFX(pspiop)(c, u, w, iz, dzs, select_after, wk);
    } else SFX(ssop)(c, u, w, iz, dzs, adj, wk);
}

int SFX(step)(RT* u, const float* vel, const orc_geom* g, int iw, int iz, float dzs, int adj)
{
    SFX(ctx) c;
    int rc = SFX(ctx_init)(&c, g, vel);
    if (rc) return rc;
    SFX(work) wk;
    memset(&wk, 0, sizeof(wk));
    rc = SFX(work_init)(&wk, &c);
    if (!rc) SFX(step_ctx)(&c, (CT*)u, SFX(omega)(&c, iw), iz, (RT)dzs, adj, &wk);
    SFX(work_free)(&wk);
    SFX(ctx_free)(&c);
    return rc;
}

/* ------------------------------------------------------------------------------ wem() */

int SFX(wem)(RT* d, RT* m, const float* wav, const float* vel, const orc_geom* g,
             float sx, float sy, int adj)
{
    SFX(ctx) c;
    int rc = SFX(ctx_init)(&c, g, vel);
    if (rc) return rc;
    int igx, igy;
    rc = orc_source_index(g, sx, sy, &igx, &igy);
    if (rc) { SFX(ctx_free)(&c); return rc; }
    const orc_sampling* s = &c.s;
    int nm = c.nm, nz = g->nz, nt = g->nt, nw = s->nw;
    int ifmin = s->ifmin, ifmax = s->ifmax, nf = ifmax - ifmin;
    if (adj) memset(m, 0, sizeof(RT) * (size_t)nm * nz);
    else memset(d, 0, sizeof(RT) * (size_t)nm * nt);
    if (nf <= 0) { SFX(ctx_free)(&c); return ORC_OK; }

    /* spectra (wem.c:112-140) */
    CT* ws = (CT*)malloc(sizeof(CT) * nw);
    CT* dg = (CT*)calloc((size_t)nm * nf, sizeof(CT));
    RT* tr = (RT*)malloc(sizeof(RT) * nt);
    CT* dw = (CT*)malloc(sizeof(CT) * nw);
    if (!ws || !dg || !tr || !dw) { free(ws); free(dg); free(tr); free(dw); SFX(ctx_free)(&c); return ORC_ENOMEM; }
    for (int it = 0; it < nt; it++) tr[it] = wav[it];
    SFX(fop)((RT*)ws, tr, nw, nt, 1);
    if (adj) {
        for (int ix = 0; ix < nm; ix++) {
            for (int it = 0; it < nt; it++) tr[it] = d[(size_t)ix * nt + it];
            SFX(fop)((RT*)dw, tr, nw, nt, 1);
            for (int iw = ifmin; iw < ifmax; iw++) {
                RT w = SFX(omega)(&c, iw);
                dg[(size_t)ix * nf + (iw - ifmin)] = w * w * dw[iw];
            }
        }
    }
#if ORC_IS_F64
    RT rnt = 1.0 / sqrt((double)s->ntfft);
#else
    RT rnt = 1.0f / sqrtf((float)s->ntfft);
#endif
    int nthr = omp_get_max_threads();
    RT* mthr = NULL;
    if (adj) {
        mthr = (RT*)calloc((size_t)nthr * nm * nz, sizeof(RT));
        if (!mthr) { free(ws); free(dg); free(tr); free(dw); SFX(ctx_free)(&c); return ORC_ENOMEM; }
    }
    int err = 0;
    /* frequency loop (wem.c:166-171) and extrap1f (wem.c:206-287) */
#pragma omp parallel
    {
        int t = omp_get_thread_num();
        SFX(work) wk;
        memset(&wk, 0, sizeof(wk));
        CT* us = (CT*)malloc(sizeof(CT) * nm);
        CT* ur = (CT*)malloc(sizeof(CT) * nm);
        CT* smig = adj ? NULL : (CT*)malloc(sizeof(CT) * (size_t)nm * nz);
        int ok = SFX(work_init)(&wk, &c) == ORC_OK && us && ur && (adj || smig);
        if (!ok) {
#pragma omp atomic write
            err = 1;
        }
#pragma omp for schedule(dynamic, 1)
        for (int iw = ifmin; iw < ifmax; iw++) {
            if (!ok) continue;
            RT w = SFX(omega)(&c, iw);
            RT factor = iw == 0 ? 1 : 2;
            for (int x = 0; x < nm; x++) us[x] = 0;
            us[(size_t)igx * g->nmy + igy] = ws[iw] * rnt;
            if (adj) {
                RT* mt = mthr + (size_t)t * nm * nz;
                for (int x = 0; x < nm; x++) ur[x] = dg[(size_t)x * nf + (iw - ifmin)] * rnt;
                for (int iz = 0; iz < nz; iz++) {
                    float z = g->oz + g->dz * iz;
                    if (z >= g->sz) SFX(step_ctx)(&c, us, w, iz, (RT)(-g->dz), 1, &wk);
                    if (z >= g->gz) {
                        SFX(step_ctx)(&c, ur, w, iz, (RT)g->dz, 1, &wk);
                        for (int x = 0; x < nm; x++)
                            mt[(size_t)x * nz + iz] += factor * MCREAL(ur[x] * MCONJ(us[x]));
                    }
                }
            } else {
                for (int iz = 0; iz < nz; iz++) {
                    float z = g->oz + g->dz * iz;
                    if (z >= g->sz) {
                        SFX(step_ctx)(&c, us, w, iz, (RT)(-g->dz), 1, &wk);
                        for (int x = 0; x < nm; x++) smig[(size_t)x * nz + iz] = us[x];
                    } else
                        for (int x = 0; x < nm; x++) smig[(size_t)x * nz + iz] = 0;
                }
                for (int x = 0; x < nm; x++) ur[x] = 0;
                for (int iz = nz - 1; iz >= 0; iz--) {
                    float z = g->oz + g->dz * iz;
                    if (z >= g->gz) {
                        for (int x = 0; x < nm; x++) ur[x] += m[(size_t)x * nz + iz] * smig[(size_t)x * nz + iz];
                        SFX(step_ctx)(&c, ur, w, iz, (RT)(-g->dz), 0, &wk);
                    }
                }
                for (int x = 0; x < nm; x++) dg[(size_t)x * nf + (iw - ifmin)] = ur[x] * rnt;
            }
        }
        SFX(work_free)(&wk);
        free(us); free(ur); free(smig);
    }
    if (!err) {
        if (adj) {
            /* reduction over threads in thread order (wem.c:175) */
            for (int x = 0; x < nm; x++)
                for (int t = 0; t < nthr; t++)
                    for (int iz = 0; iz < nz; iz++) m[(size_t)x * nz + iz] += mthr[((size_t)t * nm + x) * nz + iz];
        } else {
            /* omega^2 weighting and c2r (wem.c:177-188) */
            for (int ix = 0; ix < nm; ix++) {
                for (int iw = 0; iw < nw; iw++) dw[iw] = 0;
                for (int iw = ifmin; iw < ifmax; iw++) {
                    RT w = SFX(omega)(&c, iw);
                    dw[iw] = w * w * dg[(size_t)ix * nf + (iw - ifmin)];
                }
                SFX(fop)((RT*)dw, tr, nw, nt, 0);
                for (int it = 0; it < nt; it++) d[(size_t)ix * nt + it] = tr[it];
            }
        }
    }
    free(mthr); free(ws); free(dg); free(tr); free(dw);
    SFX(ctx_free)(&c);
    return err ? ORC_ENOMEM : ORC_OK;
}

int SFX(shots)(RT* d, RT* m, const float* wav, const float* vel, const orc_geom* g,
               const float* sx, const float* sy, int nshot, int adj)
{
    size_t nm = (size_t)g->nmx * g->nmy, nmz = nm * g->nz, nmt = nm * g->nt;
    if (adj) {
        RT* ms = (RT*)malloc(sizeof(RT) * nmz);
        if (!ms) return ORC_ENOMEM;
        memset(m, 0, sizeof(RT) * nmz);
        for (int is = 0; is < nshot; is++) {
            int rc = SFX(wem)(d + (size_t)is * nmt, ms, wav, vel, g, sx[is], sy ? sy[is] : 0.f, 1);
            if (rc) { free(ms); return rc; }
            for (size_t i = 0; i < nmz; i++) m[i] += ms[i];
        }
        free(ms);
    } else {
        for (int is = 0; is < nshot; is++) {
            int rc = SFX(wem)(d + (size_t)is * nmt, m, wav, vel, g, sx[is], sy ? sy[is] : 0.f, 0);
            if (rc) return rc;
        }
    }
    return ORC_OK;
}


/* ------------------------------------------------------------------ compute_angles() */

/* 3-D transform of a[ix*nky*nkz + iy*nkz + iz] (fftwf_plan_dft(3, {nkx,nky,nkz}), wem.c:631-633):
 * 1-D transforms along z, y and x.  tw* hold exp(-2 pi i j / n) of each length. */
static void SFX(fft3)(CT* a, int nkx, int nky, int nkz, const CT* twx, const CT* twy, const CT* twz, int sign,
                      CT* line, CT* sc)
{
    for (size_t l = 0; l < (size_t)nkx * nky; l++) SFX(fft1)(a + l * nkz, nkz, twz, sign, sc);
    if (nky > 1)
        for (int ix = 0; ix < nkx; ix++)
            for (int iz = 0; iz < nkz; iz++) {
                CT* base = a + (size_t)ix * nky * nkz + iz;
                for (int iy = 0; iy < nky; iy++) line[iy] = base[(size_t)iy * nkz];
                SFX(fft1)(line, nky, twy, sign, sc);
                for (int iy = 0; iy < nky; iy++) base[(size_t)iy * nkz] = line[iy];
            }
    if (nkx > 1)
        for (size_t l = 0; l < (size_t)nky * nkz; l++) {
            CT* base = a + l;
            for (int ix = 0; ix < nkx; ix++) line[ix] = base[(size_t)ix * nky * nkz];
            SFX(fft1)(line, nkx, twx, sign, sc);
            for (int ix = 0; ix < nkx; ix++) base[(size_t)ix * nky * nkz] = line[ix];
        }
}

/* Signed wavenumber of bin i of n with spacing dk (wem.c:799-804, 862-863, 886-887, 907-908). */
static RT SFX(kbin)(int i, int n, RT dk)
{
#if ORC_IS_F64
    return i <= n / 2 ? dk * i : -(dk * n - dk * i);
#else
    return i <= (int)n / 2 ? (float)dk * i : -((float)dk * n - dk * i);
#endif
}

/* compute_angles (wem.c:566-692), extrapolate_source (wem.c:694-776), ssop_source
 * (wem.c:778-826) and calculate_derivatives (wem.c:828-981): source incidence angles in
 * degrees, angx / angy [nm][nz].  The source march is ssop with adj = false and -dz
 * (ssop_source: taper, split-step phase, FFT, kz phase, IFFT; ntfft = nt, nkx = nmx, so the
 * geometry is taken with padt = padx = 1); per frequency the derivatives of the stored source
 * volume come from one forward 3-D transform (nkx, nky, nkz = nz) and inverse transforms of
 * i kx, i ky, i kz times it, normalised by 1/(nkx nky nkz), and the left / right signs compare
 * the inverse transforms of the negative-only and positive-only kx (ky) parts.  Frequencies are
 * summed in order (the reference sums them from OpenMP threads without synchronisation,
 * wem.c:657-665), then angx *= sign(angx_sign) / (ifmax - ifmin + 1) (wem.c:671-672). */
int SFX(angles)(RT* angx, RT* angy, const float* wav, const float* vel, const orc_geom* g0, float sx, float sy)
{
    orc_geom ga = *g0;
    ga.padt = 1;
    ga.padx = 1;
    ga.pspi = 0;
    const orc_geom* g = &ga;
    SFX(ctx) c;
    int rc = SFX(ctx_init)(&c, g, vel);
    if (rc) return rc;
    int igx, igy;
    rc = orc_source_index(g, sx, sy, &igx, &igy);
    if (rc) { SFX(ctx_free)(&c); return rc; }
    const orc_sampling* s = &c.s;
    const int nmy = g->nmy, nz = g->nz, nt = g->nt, nw = s->nw, nm = c.nm;
    const int nkx = s->nkx, nky = s->nky, nkz = nz;
    const size_t nvol = (size_t)nkx * nky * nkz;
    const int ifmin = s->ifmin, ifmax = s->ifmax;
    memset(angx, 0, sizeof(RT) * (size_t)nm * nz);
    memset(angy, 0, sizeof(RT) * (size_t)nm * nz);
    RT* sgx = (RT*)calloc((size_t)nm * nz, sizeof(RT));
    RT* sgy = (RT*)calloc((size_t)nm * nz, sizeof(RT));
    CT* ws = (CT*)malloc(sizeof(CT) * nw);
    RT* tr = (RT*)malloc(sizeof(RT) * nt);
    CT* u = (CT*)malloc(sizeof(CT) * nm);
    CT* us = (CT*)malloc(sizeof(CT) * (size_t)nm * nz);
    CT* a = (CT*)malloc(sizeof(CT) * nvol);
    CT* b = (CT*)malloc(sizeof(CT) * nvol);
    CT* bl = (CT*)malloc(sizeof(CT) * nvol);
    CT* br = (CT*)malloc(sizeof(CT) * nvol);
    RT* dxa = (RT*)malloc(sizeof(RT) * (size_t)nm * nz);
    RT* dya = (RT*)malloc(sizeof(RT) * (size_t)nm * nz);
    int nl = nkx > nky ? nkx : nky;
    if (nkz > nl) nl = nkz;
    CT* line = (CT*)malloc(sizeof(CT) * nl);
    CT* sc = (CT*)malloc(sizeof(CT) * nl);
    CT* twz = (CT*)malloc(sizeof(CT) * nkz);
    SFX(work) wk;
    memset(&wk, 0, sizeof(wk));
    int ok = sgx && sgy && ws && tr && u && us && a && b && bl && br && dxa && dya && line && sc && twz &&
             SFX(work_init)(&wk, &c) == ORC_OK;
    if (ok && ifmax > ifmin) {
        SFX(twiddles)(twz, nkz);
        for (int it = 0; it < nt; it++) tr[it] = wav[it];
        SFX(fop)((RT*)ws, tr, nw, nt, 1);
#if ORC_IS_F64
        const RT rnt = 1.0 / sqrt((double)s->ntfft), rvol = (double)nvol;
        const RT dkx = 2.0 * M_PI / ((double)nkx * g->dmx), dky = 2.0 * M_PI / ((double)nky * g->dmy);
        const RT dkz = 2.0 * M_PI / ((double)nkz * g->dz), deg = 180.0 / M_PI, eps = 1e-10;
#else
        const RT rnt = 1.0f / sqrtf((float)s->ntfft), rvol = (float)nkx * nky * nkz;
        const RT dkx = s->dkx, dky = s->dky;
        const RT dkz = (float)(2 * 3.141592653589793 / ((double)(float)nkz) / (double)g->dz);
        const RT deg = (float)(180 / 3.141592653589793), eps = 1e-10f;
#endif
        for (int iw = ifmin; iw < ifmax; iw++) {
            RT w = SFX(omega)(&c, iw);
            for (int x = 0; x < nm; x++) u[x] = 0;
            u[(size_t)igx * nmy + igy] = ws[iw] * rnt;
            for (int iz = 0; iz < nz; iz++) {
                float z = g->oz + g->dz * iz;
                if (z >= g->sz) {
                    SFX(ssop)(&c, u, w, iz, (RT)(-g->dz), 0, &wk);
                    for (int x = 0; x < nm; x++) us[(size_t)x * nz + iz] = u[x];
                } else
                    for (int x = 0; x < nm; x++) us[(size_t)x * nz + iz] = 0;
            }
            /* calculate_derivatives (wem.c:828-981); nkx = nmx, nky = nmy, nkz = nz */
            memcpy(a, us, sizeof(CT) * nvol);
            SFX(fft3)(a, nkx, nky, nkz, c.twx, c.twy, twz, -1, line, sc);
            for (int pass = 0; pass < 3; pass++) {  /* x, y, z derivative magnitudes */
                for (int ix = 0; ix < nkx; ix++)
                    for (int iy = 0; iy < nky; iy++)
                        for (int iz = 0; iz < nkz; iz++) {
                            size_t k = ((size_t)ix * nky + iy) * nkz + iz;
                            RT kk = pass == 0 ? SFX(kbin)(ix, nkx, dkx)
                                  : pass == 1 ? SFX(kbin)(iy, nky, dky) : SFX(kbin)(iz, nkz, dkz);
                            b[k] = (RT _Complex)_Complex_I * kk * a[k];
                        }
                SFX(fft3)(b, nkx, nky, nkz, c.twx, c.twy, twz, +1, line, sc);
                for (size_t k = 0; k < nvol; k++) {
                    CT v = b[k] / rvol;
#if ORC_IS_F64
                    RT mag = cabs(v);
#else
                    RT mag = cabsf(v);
#endif
                    if (pass == 0) dxa[k] = mag;
                    else if (pass == 1) dya[k] = mag;
                    else {
#if ORC_IS_F64
                        angx[k] += deg * atan(dxa[k] / (mag + eps));
                        angy[k] += deg * atan(dya[k] / (mag + eps));
#else
                        angx[k] += deg * atanf(dxa[k] / (mag + eps));
                        angy[k] += deg * atanf(dya[k] / (mag + eps));
#endif
                    }
                }
            }
            for (int axis = 0; axis < 2; axis++) {  /* left / right signs along x, then y */
                for (int ix = 0; ix < nkx; ix++)
                    for (int iy = 0; iy < nky; iy++)
                        for (int iz = 0; iz < nkz; iz++) {
                            size_t k = ((size_t)ix * nky + iy) * nkz + iz;
                            int i = axis == 0 ? ix : iy, n = axis == 0 ? nkx : nky;
                            RT kk = SFX(kbin)(i, n, axis == 0 ? dkx : dky);
                            RT kl = i <= n / 2 ? 0 : kk, kr = i <= n / 2 ? kk : 0;
                            bl[k] = (RT _Complex)_Complex_I * kl * a[k];
                            br[k] = (RT _Complex)_Complex_I * kr * a[k];
                        }
                SFX(fft3)(bl, nkx, nky, nkz, c.twx, c.twy, twz, +1, line, sc);
                SFX(fft3)(br, nkx, nky, nkz, c.twx, c.twy, twz, +1, line, sc);
                RT* sg = axis == 0 ? sgx : sgy;
                for (size_t k = 0; k < nvol; k++) {
#if ORC_IS_F64
                    RT mu = cabs(us[k]), ml = cabs(bl[k]), mr = cabs(br[k]);
#else
                    RT mu = cabsf(us[k]), ml = cabsf(bl[k]), mr = cabsf(br[k]);
#endif
                    sg[k] += (ml >= mr ? mu : -mu) / (RT)nw;
                }
            }
        }
        const RT nrm = (RT)1 / (RT)(ifmax - ifmin + 1);
        for (size_t k = 0; k < (size_t)nm * nz; k++) {
            angx[k] *= (sgx[k] >= 0 ? (RT)1 : (RT)-1) * nrm;
            angy[k] *= (sgy[k] >= 0 ? (RT)1 : (RT)-1) * nrm;
        }
    }
    SFX(work_free)(&wk);
    free(sgx); free(sgy); free(ws); free(tr); free(u); free(us); free(a); free(b); free(bl); free(br);
    free(dxa); free(dya); free(line); free(sc); free(twz);
    SFX(ctx_free)(&c);
    return ok ? ORC_OK : ORC_ENOMEM;
}


/* ------------------------------------------------------------------------- ewem() (elastic) */

/* elastic_separate_3d (ewem.c:553-741) at one wavenumber: sep maps (ux, uz) to the potentials
 * (up, us2), !sep maps (up, us2) back; adj selects the transposed matrices. */
void SFX(esep_k)(RT kx, RT w, RT vp, RT vs, int sep, int adj, const RT* in1, const RT* in2, RT* o1, RT* o2)
{
    CT a1 = in1[0] + in1[1] * (RT _Complex)_Complex_I, a2 = in2[0] + in2[1] * (RT _Complex)_Complex_I, r1, r2;
    RT sp = w * w / (vp * vp) - kx * kx, ss = w * w / (vs * vs) - kx * kx;
    RT kzp = sp > 0 ? MSQRT(sp) : 0, kzs = ss > 0 ? MSQRT(ss) : 0;
    RT np = MSQRT(kx * kx + kzp * kzp), ns = MSQRT(kx * kx + kzs * kzs);
    RT kxp = kx / np, kxs = kx / ns;
    kzp = kzp / np;
    kzs = kzs / ns;
    RT norm = kxp * kxs + kzp * kzs;
    if (norm >= (RT)0.1 && w >= (RT)20) {
        if (sep) {
            if (!adj) { r1 = (kxs * a1 + kzs * a2) / norm; r2 = (-kzp * a1 + kxp * a2) / norm; }
            else { r1 = kxp * a1 + kzp * a2; r2 = -kzs * a1 + kxs * a2; }
        } else {
            if (!adj) { r1 = kxp * a1 - kzs * a2; r2 = kzp * a1 + kxs * a2; }
            else { r1 = (kxs * a1 - kzp * a2) / norm; r2 = (kzs * a1 + kxp * a2) / norm; }
        }
    } else {  /* up = uz, us2 = ux (sep); ux = us2, uz = up (!sep) */
        r1 = a2;
        r2 = a1;
    }
    o1[0] = MCREAL(r1); o1[1] = MCREAL(-r1 * (RT _Complex)_Complex_I);
    o2[0] = MCREAL(r2); o2[1] = MCREAL(-r2 * (RT _Complex)_Complex_I);
}

/* Separation / combination of a wavefield pair (elastic_separate_3d, ewem.c:553-741): pad both
 * to the nkx x nky grid, transform (2-D when nmy > 1, unitary), apply esep_k per wavenumber with
 * the matrix of kx alone (kx = dkx ikx for ikx < nkx/2, else -(dkx nkx - dkx ikx), ewem.c:593; the
 * reference's matrix does not depend on ky), transform back and truncate to nmx x nmy. */
static void SFX(esep)(const SFX(ctx)* c, CT* i1, CT* i2, CT* o1, CT* o2, RT w, RT vp, RT vs, int sep, int adj,
                      SFX(work)* wk, CT* b1, CT* b2)
{
    int nmx = c->g->nmx, nmy = c->g->nmy, nkx = c->s.nkx, nky = c->s.nky;
#if ORC_IS_F64
    RT rnk = 1.0 / sqrt((double)nkx * nky), dkx = 2.0 * M_PI / ((double)nkx * c->g->dmx);
#else
    RT rnk = 1.0f / sqrtf((float)nkx * nky), dkx = c->s.dkx;
#endif
    for (int ix = 0; ix < nkx; ix++)
        for (int iy = 0; iy < nky; iy++) {
            size_t k = (size_t)ix * nky + iy, x = (size_t)ix * nmy + iy;
            b1[k] = (ix < nmx && iy < nmy) ? i1[x] : 0;
            b2[k] = (ix < nmx && iy < nmy) ? i2[x] : 0;
        }
    SFX(fft2)(c, b1, -1, wk);
    SFX(fft2)(c, b2, -1, wk);
    for (int ik = 0; ik < nkx; ik++) {
        RT kx = ik < nkx / 2.0 ? dkx * ik : -(dkx * nkx - dkx * ik);
        for (int iy = 0; iy < nky; iy++) {
            size_t k = (size_t)ik * nky + iy;
            CT a1 = b1[k] * rnk, a2 = b2[k] * rnk;
            RT r1[2], r2[2];
            SFX(esep_k)(kx, w, vp, vs, sep, adj, (RT*)&a1, (RT*)&a2, r1, r2);
            b1[k] = r1[0] + r1[1] * (RT _Complex)_Complex_I;
            b2[k] = r2[0] + r2[1] * (RT _Complex)_Complex_I;
        }
    }
    SFX(fft2)(c, b1, +1, wk);
    SFX(fft2)(c, b2, +1, wk);
    for (int ix = 0; ix < nmx; ix++)
        for (int iy = 0; iy < nmy; iy++) {
            size_t k = (size_t)ix * nky + iy, x = (size_t)ix * nmy + iy;
            o1[x] = b1[k] * rnk;
            o2[x] = b2[k] * rnk;
        }
}

/* The sampling of ewem (ewem.c:46-57): ntfft = 2 trunc(padt nt / 2), ifmax = trunc(fmax dt ntfft)
 * + 1, ifmin = trunc(fmin dt ntfft) (no integer cast of fmax / fmin as in wem.c). */
static void SFX(ewem_sampling)(const orc_geom* g, orc_sampling* s)
{
    float h = (float)g->padt * (float)g->nt / 2.f;
    s->ntfft = (int)(2.f * truncf(h));
    s->nw = s->ntfft / 2 + 1;
    s->fop_n = (s->nw - 1) * 2;
    s->dw = (float)(2 * 3.141592653589793 / ((double)(float)s->ntfft) / (double)g->dt);
    float t = g->fmax * g->dt * (float)s->ntfft;
    s->ifmax = t + 1 < (float)s->nw ? (int)trunc(t) + 1 : s->nw;
    t = g->fmin * g->dt * (float)s->ntfft;
    s->ifmin = t + 1 < (float)s->ifmax ? (int)trunc(t) : 0;
}

/* ewem (ewem.c:3-253) and elastic_extrap1f (ewem.c:255-366), 2-D and 3-D: adj maps the
 * components ux, uz [nm][nt] to the images mpp, mps2 [nm][nz] (mps1 stays zero: the reference
 * never forms the first S potential, and uy is not used); !adj maps mpp, mps2 to ux, uz (uy stays
 * zero).  The source (vp) is injected at igx nmy + igy (ewem.c:99-108) and normalised by
 * max|wav| / sqrt(ntfft) (ewem.c:146); the receiver spectra carry no w^2. */
int SFX(ewem)(RT* ux, RT* uz, RT* mpp, RT* mps2, const float* wav, const float* vp, const float* vs,
              const orc_geom* g0, float sx, float sy, int adj)
{
    orc_geom ga = *g0;
    ga.pspi = 0;
    const orc_geom* g = &ga;
    SFX(ctx) cp, cs;
    int rc = SFX(ctx_init)(&cp, g, vp);
    if (rc) return rc;
    rc = SFX(ctx_init)(&cs, g, vs);
    if (rc) { SFX(ctx_free)(&cp); return rc; }
    SFX(ewem_sampling)(g, &cp.s);
    SFX(ewem_sampling)(g, &cs.s);
    const orc_sampling* s = &cp.s;
    const int nmx = g->nmx, nmy = g->nmy, nm = nmx * nmy, nz = g->nz, nt = g->nt, nw = s->nw, nk = s->nkx * s->nky;
    const int ifmin = s->ifmin, ifmax = s->ifmax;
    const int igx = (int)truncf((sx - g->omx) / g->dmx);
    const int igy = nmy > 1 ? (int)truncf((sy - g->omy) / g->dmy) : 0;  /* ewem.c:99-100 */
    if (igx < 0 || igx >= nmx || igy < 0 || igy >= nmy || s->fop_n < nt) { SFX(ctx_free)(&cp); SFX(ctx_free)(&cs); return ORC_EINVAL; }
    if (adj) { memset(mpp, 0, sizeof(RT) * (size_t)nm * nz); memset(mps2, 0, sizeof(RT) * (size_t)nm * nz); }
    else { memset(ux, 0, sizeof(RT) * (size_t)nm * nt); memset(uz, 0, sizeof(RT) * (size_t)nm * nt); }
#if ORC_IS_F64
    const RT rnt = 1.0 / sqrt((double)s->ntfft);
#else
    const RT rnt = 1.0f / sqrtf((float)s->ntfft);
#endif
    RT maxs = 0;
    for (int it = 0; it < nt; it++) {
        RT v = MFABS((RT)wav[it]) * rnt;
        if (v > maxs) maxs = v;
    }
    const int nf = ifmax > ifmin ? ifmax - ifmin : 0;
    CT* ws = (CT*)malloc(sizeof(CT) * nw);
    CT* dw = (CT*)malloc(sizeof(CT) * nw);
    RT* tr = (RT*)malloc(sizeof(RT) * nt);
    CT* gx = (CT*)calloc((size_t)nm * (nf ? nf : 1), sizeof(CT));
    CT* gz = (CT*)calloc((size_t)nm * (nf ? nf : 1), sizeof(CT));
    CT* us = (CT*)malloc(sizeof(CT) * nm);
    CT* rx = (CT*)malloc(sizeof(CT) * nm);
    CT* rz = (CT*)malloc(sizeof(CT) * nm);
    CT* up = (CT*)malloc(sizeof(CT) * nm);
    CT* u2 = (CT*)malloc(sizeof(CT) * nm);
    CT* smig = (CT*)malloc(sizeof(CT) * (size_t)nm * nz);
    CT* b1 = (CT*)malloc(sizeof(CT) * nk);
    CT* b2 = (CT*)malloc(sizeof(CT) * nk);
    RT* vp0 = (RT*)malloc(sizeof(RT) * nz);
    RT* vs0 = (RT*)malloc(sizeof(RT) * nz);
    SFX(work) wk;
    memset(&wk, 0, sizeof(wk));
    int ok = ws && dw && tr && gx && gz && us && rx && rz && up && u2 && smig && b1 && b2 && vp0 && vs0 &&
             SFX(work_init)(&wk, &cp) == ORC_OK;
    if (ok && nf > 0) {
        for (int iz = 0; iz < nz; iz++) { vp0[iz] = (RT)1 / cp.po[iz]; vs0[iz] = (RT)1 / cs.po[iz]; }
        for (int it = 0; it < nt; it++) tr[it] = wav[it];
        SFX(fop)((RT*)ws, tr, nw, nt, 1);
        if (adj)
            for (int ix = 0; ix < nm; ix++) {
                for (int it = 0; it < nt; it++) tr[it] = ux[(size_t)ix * nt + it];
                SFX(fop)((RT*)dw, tr, nw, nt, 1);
                for (int iw = ifmin; iw < ifmax; iw++) gx[(size_t)ix * nf + iw - ifmin] = dw[iw];
                for (int it = 0; it < nt; it++) tr[it] = uz[(size_t)ix * nt + it];
                SFX(fop)((RT*)dw, tr, nw, nt, 1);
                for (int iw = ifmin; iw < ifmax; iw++) gz[(size_t)ix * nf + iw - ifmin] = dw[iw];
            }
        for (int iw = ifmin; iw < ifmax; iw++) {
            const RT w = SFX(omega)(&cp, iw), factor = iw == 0 ? 1 : 2;
            for (int x = 0; x < nm; x++) us[x] = 0;
            us[(size_t)igx * nmy + igy] = ws[iw] * rnt;
            for (int iz = 0; iz < nz; iz++) {  /* source, ewem.c:302-310 */
                float z = g->oz + g->dz * iz;
                if (z >= g->sz) {
                    SFX(ssop)(&cp, us, w, iz, (RT)(-g->dz), 1, &wk);
                    for (int x = 0; x < nm; x++) smig[(size_t)x * nz + iz] = us[x] / maxs;
                } else
                    for (int x = 0; x < nm; x++) smig[(size_t)x * nz + iz] = 0;
            }
            if (adj) {  /* ewem.c:311-327 */
                for (int x = 0; x < nm; x++) {
                    rx[x] = gx[(size_t)x * nf + iw - ifmin] * rnt;
                    rz[x] = gz[(size_t)x * nf + iw - ifmin] * rnt;
                }
                for (int iz = 0; iz < nz; iz++) {
                    float z = g->oz + g->dz * iz;
                    if (!(z >= g->gz)) continue;
                    SFX(esep)(&cp, rx, rz, up, u2, w, vp0[iz], vs0[iz], 1, 1, &wk, b1, b2);
                    SFX(ssop)(&cp, up, w, iz, (RT)g->dz, 1, &wk);
                    SFX(ssop)(&cs, u2, w, iz, (RT)g->dz, 1, &wk);
                    SFX(esep)(&cp, up, u2, rx, rz, w, vp0[iz], vs0[iz], 0, 1, &wk, b1, b2);
                    for (int x = 0; x < nm; x++) {
                        CT sm = MCONJ(smig[(size_t)x * nz + iz]);
                        mpp[(size_t)x * nz + iz] += factor * MCREAL(sm * up[x]);
                        mps2[(size_t)x * nz + iz] += factor * MCREAL(sm * u2[x]);
                    }
                }
            } else {  /* ewem.c:328-346 */
                for (int x = 0; x < nm; x++) rx[x] = rz[x] = 0;
                for (int iz = nz - 1; iz >= 0; iz--) {
                    float z = g->oz + g->dz * iz;
                    if (!(z >= g->gz)) continue;
                    SFX(esep)(&cp, rx, rz, up, u2, w, vp0[iz], vs0[iz], 1, 0, &wk, b1, b2);
                    for (int x = 0; x < nm; x++) {
                        up[x] += smig[(size_t)x * nz + iz] * mpp[(size_t)x * nz + iz];
                        u2[x] += smig[(size_t)x * nz + iz] * mps2[(size_t)x * nz + iz];
                    }
                    SFX(ssop)(&cp, up, w, iz, (RT)(-g->dz), 0, &wk);
                    SFX(ssop)(&cs, u2, w, iz, (RT)(-g->dz), 0, &wk);
                    SFX(esep)(&cp, up, u2, rx, rz, w, vp0[iz], vs0[iz], 0, 0, &wk, b1, b2);
                }
                for (int x = 0; x < nm; x++) {
                    gx[(size_t)x * nf + iw - ifmin] = rx[x] * rnt;
                    gz[(size_t)x * nf + iw - ifmin] = rz[x] * rnt;
                }
            }
        }
        if (!adj)  /* ewem.c:219-241: c2r of the band, no w^2 */
            for (int ix = 0; ix < nm; ix++)
                for (int k = 0; k < 2; k++) {
                    CT* src = k == 0 ? gx : gz;
                    RT* out = k == 0 ? ux : uz;
                    for (int iw = 0; iw < nw; iw++) dw[iw] = 0;
                    for (int iw = ifmin; iw < ifmax; iw++) dw[iw] = src[(size_t)ix * nf + iw - ifmin];
                    SFX(fop)((RT*)dw, tr, nw, nt, 0);
                    for (int it = 0; it < nt; it++) out[(size_t)ix * nt + it] = tr[it];
                }
    }
    SFX(work_free)(&wk);
    free(ws); free(dw); free(tr); free(gx); free(gz); free(us); free(rx); free(rz); free(up); free(u2);
    free(smig); free(b1); free(b2); free(vp0); free(vs0);
    SFX(ctx_free)(&cp);
    SFX(ctx_free)(&cs);
    return ok ? ORC_OK : ORC_ENOMEM;
}

#undef CT
