// swem_host.cpp -- plans (device-resident tables and work buffers), the in-memory C ABI
// (swem_wem, swem_shots, swem_plan_*) and the status helpers.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "swem_plan.hpp"

namespace swem {

static thread_local std::string g_err;
thread_local DevAllocator* tl_dev_alloc = nullptr;
thread_local HostStager* tl_stager = nullptr;

int upload_bytes(void* dev, const void* host, size_t bytes)
{
    if (!bytes) return SWEM_OK;
    if (tl_stager) return tl_stager->put(dev, host, bytes);
    SWEM_HIP(hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice));
    return SWEM_OK;
}

int set_error(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_error(hipError_t e, const char* what, const char* file, int line)
{
    return set_error(e == hipErrorOutOfMemory ? SWEM_ENOMEM : SWEM_EDEVICE, "%s failed: %s (%s:%d)", what,
                     hipGetErrorString(e), file, line);
}

static const double kPi = 3.141592653589793;

static int check_device_uncached(int dev);

// A device that passed the check once is current and usable; only hipSetDevice is repeated.
static int check_device(int dev)
{
    static std::atomic<unsigned long long> ok_mask{0};
    if (dev >= 0 && dev < 64 && (ok_mask.load() >> dev) & 1ull) {
        SWEM_HIP(hipSetDevice(dev));
        return SWEM_OK;
    }
    SWEM_TRY(check_device_uncached(dev));
    if (dev >= 0 && dev < 64) ok_mask.fetch_or(1ull << dev);
    return SWEM_OK;
}

static int check_device_uncached(int dev)
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count < 1)
        return set_error(SWEM_EDEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    if (dev < 0 || dev >= count) return set_error(SWEM_EINVAL, "device %d out of range (%d devices)", dev, count);
    SWEM_HIP(hipSetDevice(dev));
    hipDeviceProp_t prop;
    SWEM_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(SWEM_EDEVICE, "device %d is %s; this library is built for gfx950 (MI355X)", dev, prop.gcnArchName);
    return SWEM_OK;
}

}  // namespace swem

using namespace swem;

namespace swem {

// Algorithmic FP32 flops of one (frequency, depth) step: 2 wavefields + imaging.
static double step_flops(const swem_plan* p)
{
    const double nk = p->s.nkx * (double)p->s.nky, nm = p->nm, lg = std::log2(nk);
    double fw;
    if (p->g.pspi)
        fw = 5.0 * nk * lg * (1.0 + p->g.nref) + 6.0 * p->g.nref * nk + 8.0 * nm;
    else
        fw = 10.0 * nk * lg + 6.0 * nk + 8.0 * nm;
    return 2.0 * fw + 4.0 * nm;
}

// Frequencies per migration workgroup (2F subgroups of T threads) or subgroups per modelling
// workgroup (one frequency each).  Preferred: at most 512 threads and 80 KiB of LDS so two
// workgroups share a CU; fewer when a single shot would leave CUs idle.
static int choose_group(const swem_plan* p, int adj, int forced)
{
    if (forced > 0) return forced;
    const March2DShape sh = march2d_shape(p->nkx, p->opt.variant, adj);
    int F = std::max(1, std::min(512, sh.max_block) / (sh.T * sh.per_freq));
    while (F > 1 && march2d_lds_bytes(p->nkx, sh, F) > 80 * 1024) F--;
    while (F > 1 && (p->nf + F - 1) / F < 512) F--;
    return F;
}

// Fourier bases of the band (f_op, wem.c:480-514, with the w^2 of wem.c:129,182 and the
// 1/sqrt(ntfft) of wem.c:231-232,278), built in FP64.  Angle index (iw*t) mod n exact in integers:
// the angles are 2 pi e / n for the residues e, one cos / sin per residue ([ct | st], n each),
// and w^2 / sqrt(ntfft) per frequency (w2s, nf).
static void basis_angles(const swem_geometry& g, const Sampling& sm, int nf, std::vector<double>& tab)
{
    const int n = sm.fop_n;
    const double rnt = 1.0 / std::sqrt((double)sm.ntfft);
    const double dw = 2.0 * kPi / ((double)sm.ntfft * (double)g.dt);
    tab.resize((size_t)2 * n + nf);
    for (int e = 0; e < n; e++) {
        const double ang = 2.0 * kPi * (double)e / (double)n;
        tab[e] = std::cos(ang);
        tab[(size_t)n + e] = std::sin(ang);
    }
    for (int f = 0; f < nf; f++) {
        const double w = (sm.ifmin + f) * dw;
        tab[(size_t)2 * n + f] = w * w * rnt;
    }
}

int build_bases(const swem_geometry& g, const Sampling& sm, int nf, Bases& b)
{
    const int nt = g.nt, n = sm.fop_n;
    const double rnt = 1.0 / std::sqrt((double)sm.ntfft);
    std::vector<double> tab;
    basis_angles(g, sm, nf, tab);
    const double *ct = tab.data(), *st = ct + n, *w2s = st + n;
    std::vector<float> badj((size_t)nt * 2 * nf), bsrc((size_t)nt * 2 * nf), bfwd((size_t)2 * nf * nt);
    // [t][f] layouts, rows in order (iw < n and t < n: the residue advances by one subtraction)
    for (int t = 0; t < nt; t++) {
        int e = (int)(((long long)sm.ifmin * t) % n);
        float* ra = badj.data() + (size_t)t * 2 * nf;
        float* rs = bsrc.data() + (size_t)t * 2 * nf;
        for (int f = 0; f < nf; f++) {
            const double c = ct[e], sn = st[e];
            ra[2 * f] = (float)(w2s[f] * c);
            ra[2 * f + 1] = (float)(-w2s[f] * sn);
            rs[2 * f] = (float)(rnt * c);
            rs[2 * f + 1] = (float)(-rnt * sn);
            e += t;
            if (e >= n) e -= n;
        }
    }
    // [f][t] layout
    for (int f = 0; f < nf; f++) {
        const int iw = sm.ifmin + f;
        const double cf = (iw == 0 || 2 * iw == n) ? 1.0 : 2.0;
        float* rc = bfwd.data() + (size_t)(2 * f) * nt;
        float* rn = rc + nt;
        int e = 0;
        for (int t = 0; t < nt; t++) {
            rc[t] = (float)(cf * w2s[f] * ct[e]);
            rn[t] = (float)(-cf * w2s[f] * st[e]);
            e += iw;
            if (e >= n) e -= n;
        }
    }
    SWEM_TRY(b.adj.upload(badj.data(), badj.size()));
    SWEM_TRY(b.fwd.upload(bfwd.data(), bfwd.size()));
    SWEM_TRY(b.src.upload(bsrc.data(), bsrc.size()));
    return SWEM_OK;
}

// The same bases filled on the device from the uploaded angle table (the same FP64 products,
// rounded to float): 2 n + nf doubles cross PCIe instead of the three bases.
static int build_bases_device(const swem_geometry& g, const Sampling& sm, int nf, Bases& b, hipStream_t st)
{
    const int nt = g.nt, n = sm.fop_n;
    std::vector<double> tab;
    basis_angles(g, sm, nf, tab);
    DevBuf<double> dtab;
    SWEM_TRY(dtab.upload(tab.data(), tab.size()));
    SWEM_TRY(b.adj.alloc((size_t)nt * 2 * nf));
    SWEM_TRY(b.fwd.alloc((size_t)2 * nf * nt));
    SWEM_TRY(b.src.alloc((size_t)nt * 2 * nf));
    SWEM_TRY(launch_build_bases(nt, nf, n, sm.ifmin, 1.0 / std::sqrt((double)sm.ntfft), dtab.p, b.adj.p, b.src.p,
                                b.fwd.p, st));
    SWEM_HIP(hipStreamSynchronize(st));  // dtab is freed on return
    return SWEM_OK;
}

// In-place radix-2 transform of a power-of-two length, exp(sign 2 pi i jk / n), FP64.
static void fft_pow2_f64(std::vector<std::complex<double>>& x, int sign)
{
    const size_t n = x.size();
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(x[i], x[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        for (size_t k = 0; k < len / 2; k++) {
            const double a = sign * 2.0 * kPi * (double)k / (double)len;
            const std::complex<double> w(std::cos(a), std::sin(a));
            for (size_t i = 0; i < n; i += len) {
                const std::complex<double> u = x[i + k], v = x[i + k + len / 2] * w;
                x[i + k] = u + v;
                x[i + k + len / 2] = u - v;
            }
        }
    }
}

// Bluestein tables of a length-n transform through length m (m >= 2n - 1, a power of two):
// chirp_j = exp(-i pi j^2 / n) (j^2 reduced mod 2n, exact in integers), and for the two signs s
// of the transform the spectra B_s = FFT_m(b_s) / m of the wrapped even kernel
// b_s(j) = b_s(m - j) = exp(-s i pi j^2 / n), j < n (zero between).
static void bluestein_tables(int n, int m, std::vector<float2>& chirp, std::vector<float2>& bchirp)
{
    std::vector<std::complex<double>> c(n);
    for (int j = 0; j < n; j++) {
        const long long e = ((long long)j * j) % (2LL * n);
        const double a = -kPi * (double)e / (double)n;
        c[j] = std::complex<double>(std::cos(a), std::sin(a));
    }
    chirp.resize(n);
    for (int j = 0; j < n; j++) chirp[j] = make_float2((float)c[j].real(), (float)c[j].imag());
    bchirp.assign((size_t)2 * m, make_float2(0.f, 0.f));
    for (int h = 0; h < 2; h++) {  // h = 0: sign -1 (kernel conj(chirp)), h = 1: sign +1 (chirp)
        std::vector<std::complex<double>> b(m, 0.0);
        for (int j = 0; j < n; j++) {
            const std::complex<double> v = h == 0 ? std::conj(c[j]) : c[j];
            b[j] = v;
            if (j) b[m - j] = v;
        }
        fft_pow2_f64(b, -1);
        for (int k = 0; k < m; k++)
            bchirp[(size_t)h * m + k] = make_float2((float)(b[k].real() / m), (float)(b[k].imag() / m));
    }
}

// Tables of a transform of n points on an array of m: the twiddles exp(-2 pi i j / len) in FP64
// (len = m for power-of-two and Bluestein arrays, n for a mixed-radix length), Bluestein's chirps,
// and the mixed-radix plan (n = 0 in mr for the other kinds).
static int length_tables(int n, int m, int kind, DevBuf<float2>& tw, DevBuf<float2>& chirp, DevBuf<float2>& bchirp,
                         MixedRadix& mr)
{
    const int len = kind == kLenMixed ? n : m;
    std::vector<float2> t(len);
    for (int j = 0; j < len; j++) {
        const double a = -2.0 * kPi * (double)j / (double)len;
        t[j] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    SWEM_TRY(tw.upload(t.data(), t.size()));
    mr = MixedRadix{};
    if (kind == kLenMixed && !mixed_radix_plan(n, &mr))
        return set_error(SWEM_EDEVICE, "no mixed-radix plan for n=%d", n);
    if (kind == kLenBluestein) {
        std::vector<float2> c, bc;
        bluestein_tables(n, m, c, bc);
        SWEM_TRY(chirp.upload(c.data(), c.size()));
        SWEM_TRY(bchirp.upload(bc.data(), bc.size()));
    }
    return SWEM_OK;
}

static int build_tables(swem_plan* p, const float* wav)
{
    const swem_geometry& g = p->g;
    const float* bsrc = nullptr;
    const int nmx = g.nmx, nz = g.nz, nm = p->nm, nf = p->nf, nt = g.nt, nkx = p->nkx, narr = p->narr;
    // x transform tables: twiddles in FP64 of the array length (power of two, Bluestein) or of
    // nkx itself (mixed radix), and the chirps of Bluestein's identity
    SWEM_TRY(length_tables(nkx, narr, p->three_d ? march3d_kind(nkx) : march2d_kind(nkx), p->tw, p->chirp, p->bchirp,
                           p->mrx));
    // taper (wem.c:546-564)
    std::vector<float> tap(nmx);
    taper_f32(nmx, p->s.lmx, tap.data());
    SWEM_TRY(p->taper.upload(tap.data(), tap.size()));
    if (p->three_d) {
        const int nky = p->s.nky, my = p->narr_y;
        SWEM_TRY(length_tables(nky, my, march3d_kind(nky), p->twy, p->chirp_y, p->bchirp_y, p->mry));
        std::vector<float> tapy(g.nmy);
        taper_f32(g.nmy, p->s.lmy, tapy.data());
        SWEM_TRY(p->taper_y.upload(tapy.data(), tapy.size()));
    }
    // slowness split (wem.c:27-36): FP64 for the propagation, float copy for naive FP32;
    // depth-major [nz][nmx] for the split-step table
    std::vector<double> po(nz);
    std::vector<float> pof(nz), pdf_xz((size_t)nm * nz);
    slowness_f32(g, p->vel.data(), pof.data(), pdf_xz.data());
    if (!g.pspi) {
        std::vector<double> pd((size_t)nz * nm);
        std::vector<float> pdf((size_t)nz * nm);
        for (int iz = 0; iz < nz; iz++) {
            double sum = 0;
            for (int ix = 0; ix < nm; ix++) sum += p->vel[(size_t)ix * nz + iz];
            po[iz] = 1.0 / (sum / (double)nm);
            for (int ix = 0; ix < nm; ix++) {
                pd[(size_t)iz * nm + ix] = 1.0 / (double)p->vel[(size_t)ix * nz + iz] - po[iz];
                pdf[(size_t)iz * nm + ix] = pdf_xz[(size_t)ix * nz + iz];
            }
        }
        SWEM_TRY(p->pd.upload(pd.data(), pd.size()));
        SWEM_TRY(p->pdf.upload(pdf.data(), pdf.size()));

        std::vector<double> p2(nz);
        std::vector<float> p2f(nz);
        for (int iz = 0; iz < nz; iz++) {
            p2[iz] = po[iz] * po[iz];
            p2f[iz] = pof[iz] * pof[iz];
        }
        SWEM_TRY(p->p2.upload(p2.data(), p2.size()));
        SWEM_TRY(p->p2f.upload(p2f.data(), p2f.size()));
        std::vector<uint32_t> used(nz, 1u);
        SWEM_TRY(p->used.upload(used.data(), used.size()));
        std::vector<int> cref((size_t)nz * kCrefStride, 0);
        SWEM_TRY(p->cref.upload(cref.data(), cref.size()));
        p->packed_ok = false;
    } else {
        const int nref = g.nref;
        std::vector<float> vref((size_t)nref * nz);
        std::vector<int> i1((size_t)nm * nz), i2((size_t)nm * nz);
        std::vector<uint8_t> cut((size_t)nm * nz);
        SWEM_TRY(pspi_tables(g, p->vel.data(), vref.data(), i1.data(), i2.data(), cut.data()));
        std::vector<double> p2((size_t)nz * nref);
        std::vector<float> p2f((size_t)nz * nref);
        std::vector<double> vd((size_t)nz * nref);
        for (int iz = 0; iz < nz; iz++) {
            const double vmin = vref[iz], vmax = vref[(size_t)(nref - 1) * nz + iz];
            for (int ir = 0; ir < nref; ir++) {
                const double v = vmin + (double)ir * (vmax - vmin) / ((double)nref - 1.0);
                vd[(size_t)iz * nref + ir] = v;
                p2[(size_t)iz * nref + ir] = 1.0 / (v * v);
                const float vf = vref[(size_t)ir * nz + iz];
                p2f[(size_t)iz * nref + ir] = 1 / (vf * vf);
            }
        }
        // 2-D: [nz][narr], rows padded to the transform array length (0x7f: no vref at x >= nmx);
        // 3-D: [nz][nm], trace ix = imx * nmy + imy
        const size_t srow = p->three_d ? (size_t)nm : (size_t)narr;
        std::vector<uint8_t> sel((size_t)nz * srow, 0x7f), sel2((size_t)nz * srow, 0x7f);
        std::vector<float> w2((size_t)nz * srow, 0.f);
        std::vector<uint32_t> used(nz, 0u);
        for (int ix = 0; ix < nm; ix++)
            for (int iz = 0; iz < nz; iz++) {
                const size_t i = (size_t)ix * nz + iz, o = (size_t)iz * srow + ix;
                sel[o] = (uint8_t)(i1[i] | (cut[i] ? 0x80 : 0));
                sel2[o] = (uint8_t)i2[i];
                used[iz] |= 1u << i1[i];
                if (cut[i]) {
                    const double v1 = vd[(size_t)iz * nref + i1[i]], v2 = vd[(size_t)iz * nref + i2[i]];
                    w2[o] = (float)(((double)p->vel[i] - v1) / (v2 - v1));
                    if (p->opt.pspi_mode == SWEM_PSPI_LINEAR) used[iz] |= 1u << i2[i];
                }
            }
        SWEM_TRY(p->p2.upload(p2.data(), p2.size()));
        SWEM_TRY(p->p2f.upload(p2f.data(), p2f.size()));
        SWEM_TRY(p->sel.upload(sel.data(), sel.size()));
        SWEM_TRY(p->sel2.upload(sel2.data(), sel2.size()));
        SWEM_TRY(p->w2.upload(w2.data(), w2.size()));
        SWEM_TRY(p->used.upload(used.data(), used.size()));
        // Packed compat migration: every depth steps both wavefields (z >= sz and z >= gz,
        // wem.c:235-241) and the uncut points of a depth share one vref.
        // Per depth: vref of the uncut points (-1: none), their count, 1 when no cut point selects
        // that vref, and the first kUncutSparse positions.
        std::vector<int> cref((size_t)nz * kCrefStride, 0);
        bool packed = p->opt.pspi_mode == SWEM_PSPI_COMPAT && !p->three_d;  // any 2-D length
        for (int iz = 0; iz < nz; iz++) {
            const float z = g.oz + g.dz * (float)iz;
            if (!(z >= g.sz && z >= g.gz)) packed = false;
            cref[(size_t)iz * kCrefStride] = -1;
        }
        for (int ix = 0; ix < nm; ix++)  // traces in storage order; per depth ix ascends
            for (int iz = 0; iz < nz; iz++) {
                const size_t i = (size_t)ix * nz + iz;
                if (cut[i]) continue;
                int* row = cref.data() + (size_t)iz * kCrefStride;
                if (row[0] < 0) row[0] = i1[i];
                else if (row[0] != i1[i]) packed = false;
                if (row[1] < kUncutSparse) row[3 + row[1]] = ix;
                row[1]++;
            }
        for (int iz = 0; iz < nz; iz++) {  // does a cut point select the uncut points' vref?
            int* row = cref.data() + (size_t)iz * kCrefStride;
            row[2] = row[0] >= 0;
            for (int ix = 0; ix < nm && row[2]; ix++) {
                const size_t i = (size_t)ix * nz + iz;
                if (cut[i] && i1[i] == row[0]) row[2] = 0;
            }
        }
        SWEM_TRY(p->cref.upload(cref.data(), cref.size()));
        p->packed_ok = packed;
    }
    if (p->basis_cache) {
        const Bases* b = p->basis_cache->find_or_build(g, p->s, nf);
        if (!b) return SWEM_ENOMEM;
        p->basis_adj = b->adj.p;
        p->basis_fwd = b->fwd.p;
        bsrc = b->src.p;
    } else {
        if (p->served)
            SWEM_TRY(build_bases(g, p->s, nf, p->own_bases));
        else
            SWEM_TRY(build_bases_device(g, p->s, nf, p->own_bases, p->build_stream));
        p->basis_adj = p->own_bases.adj.p;
        p->basis_fwd = p->own_bases.fwd.p;
        bsrc = p->own_bases.src.p;
    }
    // source spectrum (wem.c:115-119) through the same transform on the device
    SWEM_TRY(p->wavd.upload(wav, nt));
    SWEM_TRY(p->spec_s.alloc(nf));
    SWEM_TRY(launch_sgemm(1, 2 * nf, nt, p->wavd.p, nt, bsrc, 2 * nf, reinterpret_cast<float*>(p->spec_s.p), 2 * nf,
                          p->build_stream));
    return SWEM_OK;
}

// Bytes of the phase tables of one frequency.
static size_t table_bytes_per_freq(const swem_plan* p)
{
    const size_t nref = p->g.pspi ? p->g.nref : 1, nkh = p->nkx / 2 + 1;
    return sizeof(float2) * (size_t)p->g.nz * (nref * nkh + (p->g.pspi ? 0 : (size_t)p->g.nmx));
}

// Builds the phase tables of frequency chunk c (no-op when they are resident).
static int ensure_chunk(swem_plan* p, int c, hipStream_t st)
{
    if (p->resident == c) return SWEM_OK;
    PhaseTableArgs t{};
    t.f0 = c * p->nfc;
    t.nfc = std::min(p->nfc, p->nf - t.f0);
    t.ifmin = p->s.ifmin;
    t.nz = p->g.nz;
    t.nref = p->g.pspi ? p->g.nref : 1;
    t.nkh = p->nkx / 2 + 1;
    t.nmx = p->g.nmx;
    t.prec = p->opt.precision;
    t.dw = 2.0 * kPi / ((double)p->s.ntfft * (double)p->g.dt);
    t.dkx = 2.0 * kPi / ((double)p->s.nkx * (double)p->g.dmx);
    t.dz = p->g.dz;
    t.inv_nk = 1.0 / (double)p->s.nkx;
    t.dwf = p->s.dw;
    t.dkxf = p->s.dkx;
    t.p2 = p->p2.p;
    t.p2f = p->p2f.p;
    t.pd = p->pd.p;
    t.pdf = p->pdf.p;
    t.ktab = p->ktab.p;
    t.stab = p->stab.p;
    p->resident = -1;
    SWEM_TRY(launch_phase_tables(t, !p->g.pspi, st));
    p->resident = c;
    return SWEM_OK;
}

// Frequencies per phase-table chunk of a 2-D plan under its table cap.
static void choose_chunks(swem_plan* p)
{
    const size_t per = table_bytes_per_freq(p);
    size_t cap = p->opt.table_bytes;
    if (!cap) {
        size_t fr = 0, tot = 0;
        cap = hipMemGetInfo(&fr, &tot) == hipSuccess ? std::min((size_t)32 << 30, fr / 4) : (size_t)8 << 30;
    }
    int nfc = (int)std::min<size_t>(p->nf, cap / per);
    if (nfc < p->nf) nfc = std::max(p->F, std::max(p->Ffwd, (nfc / (p->F * p->Ffwd)) * p->F * p->Ffwd));
    if (nfc < 1) nfc = 1;
    p->nfc = std::min(nfc, p->nf);
    p->nchunks = (p->nf + p->nfc - 1) / p->nfc;
}

static int alloc_tables(swem_plan* p)
{
    const size_t nref = p->g.pspi ? p->g.nref : 1, nkh = p->nkx / 2 + 1;
    SWEM_TRY(p->ktab.alloc((size_t)p->nfc * p->g.nz * nref * nkh));
    if (!p->g.pspi) SWEM_TRY(p->stab.alloc((size_t)p->nfc * p->g.nz * p->g.nmx));
    return ensure_chunk(p, 0, p->build_stream);
}

static size_t default_budget()
{
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return (size_t)8 << 30;
    return std::min((size_t)96 << 30, fr / 2);
}

static int plan_events(swem_plan* p, int n)
{
    while ((int)p->events.size() < n) {
        hipEvent_t e;
        SWEM_HIP(hipEventCreate(&e));
        p->events.push_back(e);
        p->ev_kind.push_back(0);
    }
    return SWEM_OK;
}

static int record(swem_plan* p, int kind, hipStream_t st)
{
    if (p->served) return SWEM_OK;
    SWEM_TRY(plan_events(p, p->nev + 1));
    p->ev_kind[p->nev] = kind;
    SWEM_HIP(hipEventRecord(p->events[p->nev], st));
    p->nev++;
    return SWEM_OK;
}

static int set_igx(swem_plan* p, const float* sx, const float* sy, int nshot, hipStream_t st)
{
    p->igx_host.resize(nshot);
    for (int i = 0; i < nshot; i++) {
        int ix, iy;
        SWEM_TRY(source_index(p->g, sx[i], sy ? sy[i] : 0.f, ix, iy));
        p->igx_host[i] = ix * p->g.nmy + iy;
    }
    SWEM_TRY(p->igx.alloc(nshot));
    if (p->served) return upload_bytes(p->igx.p, p->igx_host.data(), sizeof(int) * nshot);
    if (p->igx_cap < nshot) {
        if (p->igx_pinned) (void)hipHostFree(p->igx_pinned);
        p->igx_pinned = nullptr;
        SWEM_HIP(hipHostMalloc((void**)&p->igx_pinned, sizeof(int) * nshot, hipHostMallocDefault));
        p->igx_cap = nshot;
    }
    // the previous call's copy out of the pinned buffer must have finished
    if (p->last_stream || p->nev) SWEM_HIP(hipStreamSynchronize(p->last_stream));
    std::copy(p->igx_host.begin(), p->igx_host.end(), p->igx_pinned);
    SWEM_HIP(hipMemcpyAsync(p->igx.p, p->igx_pinned, sizeof(int) * nshot, hipMemcpyHostToDevice, st));
    return SWEM_OK;
}

// Arguments of frequency chunk c; F = frequencies (or subgroups) per workgroup.
static March2DArgs base_args(const swem_plan* p, int c, int F)
{
    March2DArgs a{};
    a.nmx = p->g.nmx;
    a.nz = p->g.nz;
    a.nf = p->nf;
    a.nref = p->g.pspi ? p->g.nref : 1;
    a.f0 = c * p->nfc;
    a.nfc = std::min(p->nfc, p->nf - a.f0);
    a.ngroups = (a.nfc + F - 1) / F;
    a.nkh = p->nkx / 2 + 1;
    a.ifmin = p->s.ifmin;
    a.oz = p->g.oz;
    a.dz = p->g.dz;
    a.sz = p->g.sz;
    a.gz = p->g.gz;
    a.inv_nk = 1.0f / (float)p->s.nkx;
    a.nkx = p->nkx;
    a.tw = p->tw.p;
    a.mr = p->mrx;
    a.chirp = p->chirp.p;
    a.bchirp = p->bchirp.p;
    a.ktab = p->ktab.p;
    a.stab = p->stab.p;
    a.spec_s = p->spec_s.p;
    a.igx = p->igx.p;
    a.taper = p->taper.p;
    a.sel = p->sel.p;
    a.sel2 = p->sel2.p;
    a.w2 = p->w2.p;
    a.used = p->used.p;
    a.cref = p->cref.p;
    return a;
}

// Migration kernel variant: AUTO and PACKED run the packed PSPI compat kernel where the plan
// allows it; PACKED falls back to PAIR elsewhere.
static int adj_variant(const swem_plan* p)
{
    const int v = p->opt.variant;
    if (v == SWEM_VARIANT_AUTO || v == SWEM_VARIANT_PACKED) return p->packed_ok ? SWEM_VARIANT_PACKED : SWEM_VARIANT_AUTO;
    return v;
}

static int mode_of(const swem_plan* p)
{
    if (!p->g.pspi) return 0;
    return p->opt.pspi_mode == SWEM_PSPI_LINEAR ? 2 : 1;
}

static int migrate3d_impl(swem_plan* p, const float* d_dev, const float* sx, const float* sy, int nshot, float* m_dev,
                          int accumulate, int per_shot, hipStream_t st);
static int model3d_impl(swem_plan* p, const float* m_dev, const float* sx, const float* sy, int nshot, float* d_dev,
                        hipStream_t st);

static int migrate_impl(swem_plan* p, const float* d_dev, const float* sx, const float* sy, int nshot, float* m_dev,
                        int accumulate, int per_shot, hipStream_t st)
{
    if (!p || !d_dev || !sx || !m_dev || nshot < 1) return set_error(SWEM_EINVAL, "migrate: null argument or nshot < 1");
    SWEM_HIP(hipSetDevice(p->opt.device));
    if (p->three_d) return migrate3d_impl(p, d_dev, sx, sy, nshot, m_dev, accumulate, per_shot, st);
    const size_t nm = p->nm, nz = p->g.nz, nt = p->g.nt, nf = p->nf;
    const size_t gmax = (size_t)(p->nfc + p->F - 1) / p->F;
    const size_t per = nm * nf * sizeof(float2) + gmax * nz * p->g.nmx * sizeof(float);
    int chunk = (int)std::max<size_t>(1, std::min<size_t>(nshot, p->budget / std::max<size_t>(per, 1)));
    SWEM_TRY(p->spec_r.alloc((size_t)chunk * nm * nf));
    SWEM_TRY(p->partial.alloc((size_t)chunk * gmax * nz * p->g.nmx));
    SWEM_TRY(set_igx(p, sx, sy, nshot, st));
    p->nev = 0;
    p->last_launches = 0;
    p->last_flops = 0;
    p->last_stream = st;
    SWEM_TRY(record(p, -1, st));
    for (int s0 = 0; s0 < nshot; s0 += chunk) {
        const int ns = std::min(chunk, nshot - s0);
        SWEM_TRY(launch_sgemm(ns * (int)nm, 2 * (int)nf, (int)nt, d_dev + (size_t)s0 * nm * nt, (int)nt, p->basis_adj,
                              2 * (int)nf, reinterpret_cast<float*>(p->spec_r.p), 2 * (int)nf, st));
        SWEM_TRY(record(p, 0, st));
        for (int c = 0; c < p->nchunks; c++) {
            if (p->resident != c) {
                SWEM_TRY(ensure_chunk(p, c, st));
                SWEM_TRY(record(p, 0, st));
            }
            March2DArgs a = base_args(p, c, p->F);
            a.nshot = ns;
            a.igx = p->igx.p + s0;
            a.spec_r = p->spec_r.p;
            a.partial = p->partial.p;
            SWEM_TRY(launch_march2d(a, p->nkx, p->F, mode_of(p), 1, adj_variant(p), st));
            SWEM_TRY(record(p, 1, st));
            p->last_launches++;
            p->last_flops += step_flops(p) * (double)a.nfc * nz * ns;
            float* out = per_shot ? m_dev + (size_t)s0 * nm * nz : m_dev;
            const int acc = c > 0 || (!per_shot && (accumulate || s0 > 0));
            SWEM_TRY(launch_reduce_image(p->partial.p, ns, a.ngroups, (int)nz, p->g.nmx, out, acc, per_shot, st));
            SWEM_TRY(record(p, 2, st));
        }
    }
    return SWEM_OK;
}

static March3DArgs base3(const swem_plan* p)
{
    March3DArgs a{};
    a.nmx = p->g.nmx;
    a.nmy = p->g.nmy;
    a.nm = p->nm;
    a.nkx = p->s.nkx;
    a.nky = p->s.nky;
    a.nz = p->g.nz;
    a.nf = p->nf;
    a.ifmin = p->s.ifmin;
    a.prec = p->opt.precision;
    a.dw = 2.0 * kPi / ((double)p->s.ntfft * (double)p->g.dt);
    a.dkx = 2.0 * kPi / ((double)p->s.nkx * (double)p->g.dmx);
    a.dky = 2.0 * kPi / ((double)p->s.nky * (double)p->g.dmy);
    a.dzd = p->g.dz;
    a.dkx2 = a.dkx * a.dkx;
    a.dkxf = p->s.dkx;
    a.dkyf = p->s.dky;
    a.inv_nk = (float)(1.0 / ((double)p->s.nkx * (double)p->s.nky));
    a.p2 = p->p2.p;
    a.p2f = p->p2f.p;
    a.pd = p->pd.p;
    a.pdf = p->pdf.p;
    a.taper_x = p->taper.p;
    a.taper_y = p->taper_y.p;
    a.twx = p->tw.p;
    a.mrx = p->mrx;
    a.mry = p->mry;
    a.twy = p->twy.p;
    a.chx = p->chirp.p;
    a.bchx = p->bchirp.p;
    a.chy = p->chirp_y.p;
    a.bchy = p->bchirp_y.p;
    a.spec_s = p->spec_s.p;
    a.nref = p->g.pspi ? p->g.nref : 1;
    a.sel = p->sel.p;
    a.used = p->used.p;
    a.R = p->R.p;
    a.linear = p->g.pspi && p->opt.pspi_mode == SWEM_PSPI_LINEAR;
    a.sel2 = p->sel2.p;
    a.w2 = p->w2.p;
    return a;
}

// Bytes of the per-vref x-transformed slices of one 3-D PSPI frequency (both wavefields).
static size_t pspi3_slice_bytes(const swem_plan* p)
{
    return p->g.pspi ? 2 * (size_t)p->g.nref * p->g.nmx * p->s.nky * sizeof(float2) : 0;
}

static bool stepped_at(const swem_plan* p, int iz, float level)
{
    const float z = p->g.oz + p->g.dz * (float)iz;  // wem.c:235 in float
    return iz >= 0 && iz < p->g.nz && z >= level;
}

// Step state of one 3-D launch: at most p->state3d_bytes (default 128 MiB), so that a launch's
// wavefields stay in the 256 MB last-level (Infinity) cache between the x-pass and the y-pass of
// a depth step.
static int chunk3(const swem_plan* p, size_t per_item, int F)
{
    const size_t state_item = 2 * (size_t)p->g.nmx * p->s.nky * sizeof(float2);
    size_t n = std::max<size_t>(1, p->budget / std::max<size_t>(per_item, 1));
    n = std::min(n, std::max<size_t>(1, p->state3d_bytes / state_item));
    if (n >= (size_t)p->nf) return p->nf;
    n = std::max<size_t>(F, (n / F) * F);
    return (int)std::min<size_t>(n, p->nf);
}

// The launches of one shot's 3-D migration: every frequency chunk of nfc frequencies, the
// initial wavefields, and per depth the x-pass, the y-pass (imaging) and the depth's image summed
// over the chunk's frequency groups.  The shot's source index, output image and accumulate flag
// are read from p->shot3d, so the sequence is the same for every shot.
static int enqueue_shot3d(swem_plan* p, int nfc, int nz_run, hipStream_t st)
{
    const int nf = p->nf, nm = p->nm, nz = p->g.nz, F = p->F3;
    for (int f0 = 0; f0 < nf; f0 += nfc) {
        March3DArgs a = base3(p);
        a.f0 = f0;
        a.nitem = std::min(nfc, nf - f0);
        a.igx = -1;
        a.shot = p->shot3d.p;
        a.adj = 1;
        a.spec_r = p->spec_r.p;
        a.W = p->W.p;
        a.partial = p->partial.p;
        a.R = p->R.p;
        a.conj_src = 1;
        a.conj_rec = 0;
        SWEM_TRY(launch_march3d(0, a, F, 0, st));
        for (int iz = 0; iz < nz_run; iz++) {
            a.iz = iz;
            a.ds = stepped_at(p, iz, p->g.sz);
            a.dr = stepped_at(p, iz, p->g.gz);
            a.next_s = stepped_at(p, iz + 1, p->g.sz);
            a.next_r = stepped_at(p, iz + 1, p->g.gz);
            a.wmask = (a.ds ? 1 : 0) | (a.dr ? 2 : 0);
            a.iz_sel = iz;
            if (p->g.pspi) {
                if (a.wmask) SWEM_TRY(launch_march3d(4, a, F, 0, st));
                SWEM_TRY(launch_march3d(5, a, F, 0, st));
            } else {
                if (a.wmask) SWEM_TRY(launch_march3d(1, a, F, 0, st));
                SWEM_TRY(launch_march3d(2, a, F, 0, st));
            }
            // this depth's image, summed over the chunk's frequency groups
            SWEM_TRY(launch_reduce_depth_shot(p->partial.p, (a.nitem + F - 1) / F, nm, p->shot3d.p, iz, nz, f0 > 0, st));
        }
    }
    return SWEM_OK;
}

// Captures enqueue_shot3d as a graph of kernel nodes on the plan's capture stream and
// instantiates it (replacing an earlier one).
static int capture_shot3d(swem_plan* p, int nfc, int nz_run)
{
    if (p->graph3d_exec) (void)hipGraphExecDestroy(p->graph3d_exec);
    if (p->graph3d) (void)hipGraphDestroy(p->graph3d);
    p->graph3d_exec = nullptr;
    p->graph3d = nullptr;
    p->graph3d_key.clear();
    if (!p->cap_stream) SWEM_HIP(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
    SWEM_HIP(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeRelaxed));
    const int rc = enqueue_shot3d(p, nfc, nz_run, p->cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture", __FILE__, __LINE__);
    const hipError_t ei = hipGraphInstantiate(&p->graph3d_exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        p->graph3d_exec = nullptr;
        return hip_error(ei, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    p->graph3d = g;
    return SWEM_OK;
}

static int migrate3d_impl(swem_plan* p, const float* d_dev, const float* sx, const float* sy, int nshot, float* m_dev,
                          int accumulate, int per_shot, hipStream_t st)
{
    const size_t nm = p->nm, nz = p->g.nz, nt = p->g.nt, nf = p->nf;
    const int F = p->F3;
    const int nz_run = p->tune_nz > 0 ? std::min(p->tune_nz, (int)nz) : (int)nz;
    const size_t per_item =
        2 * (size_t)p->g.nmx * p->s.nky * sizeof(float2) + nm * sizeof(float) / F + 1 + pspi3_slice_bytes(p);
    const int nfc = chunk3(p, per_item, F);
    const size_t gmax = (nfc + F - 1) / F;
    SWEM_TRY(p->W.alloc((size_t)nfc * 2 * p->g.nmx * p->s.nky));
    SWEM_TRY(p->partial.alloc(gmax * nm));
    if (p->g.pspi) SWEM_TRY(p->R.alloc((size_t)nfc * pspi3_slice_bytes(p) / sizeof(float2)));
    SWEM_TRY(p->spec_r.alloc(nm * nf));
    SWEM_TRY(p->shot3d.alloc(1));
    SWEM_TRY(set_igx(p, sx, sy, nshot, st));
    p->nev = 0;
    p->last_launches = 0;
    p->last_flops = 0;
    p->last_stream = st;
    SWEM_TRY(record(p, -1, st));
    const uintptr_t key[] = {(uintptr_t)p->W.p, (uintptr_t)p->partial.p, (uintptr_t)p->R.p, (uintptr_t)p->spec_r.p,
                             (uintptr_t)p->shot3d.p, (uintptr_t)nfc, (uintptr_t)F, (uintptr_t)nz_run, (uintptr_t)nf};
    const bool graphs = p->use_graphs && !p->served && p->tune_nz == 0;
    if (graphs && p->graph3d_key != std::vector<uintptr_t>(std::begin(key), std::end(key))) {
        SWEM_TRY(capture_shot3d(p, nfc, nz_run));
        p->graph3d_key.assign(std::begin(key), std::end(key));
    }
    const int nchunks = ((int)nf + nfc - 1) / nfc;
    for (int s = 0; s < nshot; s++) {
        SWEM_TRY(launch_sgemm((int)nm, 2 * (int)nf, (int)nt, d_dev + (size_t)s * nm * nt, (int)nt, p->basis_adj,
                              2 * (int)nf, reinterpret_cast<float*>(p->spec_r.p), 2 * (int)nf, st));
        SWEM_TRY(record(p, 0, st));
        Shot3D v;
        v.igx = p->igx_host[s];
        v.acc = !per_shot && (accumulate || s > 0);
        v.out = per_shot ? m_dev + (size_t)s * nm * nz : m_dev;
        SWEM_TRY(launch_set_shot3d(p->shot3d.p, v, st));
        if (graphs)
            SWEM_HIP(hipGraphLaunch(p->graph3d_exec, st));
        else
            SWEM_TRY(enqueue_shot3d(p, nfc, nz_run, st));
        SWEM_TRY(record(p, 1, st));
        p->last_launches += nchunks;
        p->last_flops += step_flops(p) * (double)nf * nz;
    }
    return SWEM_OK;
}

// 3-D PSPI modelling march of one frequency chunk (compat: select after every step, wem.c:252-281
// with pspiop).  Source: x-pass then y-pass per depth, the depth's source stored.  Receiver,
// upward: y-pass (select the previous step's slices, inject m * source, forward y-FFT), x-pass.
static int model3d_pspi(swem_plan* p, March3DArgs a, int F, hipStream_t st)
{
    const int nz = p->g.nz;
    for (int iz = 0; iz < nz; iz++) {
        a.iz = iz;
        a.iz_sel = iz;
        a.ds = stepped_at(p, iz, p->g.sz);
        a.next_s = stepped_at(p, iz + 1, p->g.sz);
        a.wmask = 1;
        if (a.ds) SWEM_TRY(launch_march3d(4, a, F, 0, st));
        SWEM_TRY(launch_march3d(5, a, F, 1, st));
    }
    a.first = 1;
    if (a.linear) {
        // the exact adjoint of the migration's selection (wem.c:476 weights before propagation):
        // per depth, upward, the y-pass adds the injection and writes the weighted wavefield of
        // every vref in use, and the x-pass sums their propagated spectra into the state
        for (int iz = nz - 1; iz >= 0; iz--) {
            if (!stepped_at(p, iz, p->g.gz)) continue;
            a.iz = iz;
            a.iz_sel = iz;
            a.wmask = 2;
            SWEM_TRY(launch_march3d(5, a, F, 4, st));
            SWEM_TRY(launch_march3d(6, a, F, 0, st));
            a.first = 0;
        }
        return launch_march3d(5, a, F, 5, st);
    }
    int last = -1;  // depth of the last receiver x-pass
    for (int iz = nz - 1; iz >= 0; iz--) {
        if (!stepped_at(p, iz, p->g.gz)) continue;
        a.iz = iz;
        a.iz_sel = last < 0 ? iz : last;
        a.wmask = 2;
        SWEM_TRY(launch_march3d(5, a, F, 2, st));
        SWEM_TRY(launch_march3d(4, a, F, 0, st));
        a.first = 0;
        last = iz;
    }
    a.iz_sel = last < 0 ? 0 : last;
    a.first = last < 0 ? 1 : 0;
    return launch_march3d(5, a, F, 3, st);
}

static int model3d_impl(swem_plan* p, const float* m_dev, const float* sx, const float* sy, int nshot, float* d_dev,
                        hipStream_t st)
{
    const size_t nm = p->nm, nz = p->g.nz, nt = p->g.nt, nf = p->nf;
    const int F = p->F3f;
    const size_t per_item =
        2 * (size_t)p->g.nmx * p->s.nky * sizeof(float2) + nz * nm * sizeof(float2) + pspi3_slice_bytes(p);
    const int nfc = chunk3(p, per_item, F);
    SWEM_TRY(p->W.alloc((size_t)nfc * 2 * p->g.nmx * p->s.nky));
    if (p->g.pspi) SWEM_TRY(p->R.alloc((size_t)nfc * pspi3_slice_bytes(p) / sizeof(float2)));
    SWEM_TRY(p->smig.alloc((size_t)nfc * nz * nm));
    SWEM_TRY(p->dgw.alloc(nm * nf));
    SWEM_TRY(p->mT.alloc(nm * nz));
    SWEM_TRY(set_igx(p, sx, sy, nshot, st));
    p->nev = 0;
    p->last_launches = 0;
    p->last_flops = 0;
    p->last_stream = st;
    SWEM_TRY(record(p, -1, st));
    SWEM_TRY(launch_transpose(m_dev, (int)nm, (int)nz, p->mT.p, st));
    SWEM_TRY(record(p, 0, st));
    for (int s = 0; s < nshot; s++) {
        for (int f0 = 0; f0 < (int)nf; f0 += nfc) {
            March3DArgs a = base3(p);
            a.f0 = f0;
            a.nitem = std::min(nfc, (int)nf - f0);
            a.igx = p->igx_host[s];
            a.adj = 0;
            a.W = p->W.p;
            a.smig = p->smig.p;
            a.mT = p->mT.p;
            a.dgw = p->dgw.p;
            a.conj_src = 1;
            a.conj_rec = 1;
            a.R = p->R.p;
            SWEM_TRY(launch_march3d(0, a, F, 0, st));
            if (p->g.pspi) {
                SWEM_TRY(model3d_pspi(p, a, F, st));
                SWEM_TRY(record(p, 1, st));
                p->last_launches++;
                p->last_flops += step_flops(p) * (double)a.nitem * nz;
                continue;
            }
            // source march, stored per depth (wem.c:254-265)
            for (int iz = 0; iz < (int)nz; iz++) {
                a.iz = iz;
                a.ds = stepped_at(p, iz, p->g.sz);
                a.next_s = stepped_at(p, iz + 1, p->g.sz);
                a.wmask = 1;
                if (a.ds) SWEM_TRY(launch_march3d(1, a, F, 0, st));
                SWEM_TRY(launch_march3d(3, a, F, 0, st));
            }
            // receiver march upward with injection (wem.c:266-276)
            a.first = 1;
            for (int iz = (int)nz - 1; iz >= 0; iz--) {
                if (!stepped_at(p, iz, p->g.gz)) continue;
                a.iz = iz;
                a.wmask = 2;
                SWEM_TRY(launch_march3d(3, a, F, 1, st));
                SWEM_TRY(launch_march3d(1, a, F, 0, st));
                a.first = 0;
            }
            SWEM_TRY(launch_march3d(3, a, F, 2, st));
            SWEM_TRY(record(p, 1, st));
            p->last_launches++;
            p->last_flops += step_flops(p) * (double)a.nitem * nz;
        }
        SWEM_TRY(launch_sgemm((int)nm, (int)nt, 2 * (int)nf, reinterpret_cast<const float*>(p->dgw.p), 2 * (int)nf,
                              p->basis_fwd, (int)nt, d_dev + (size_t)s * nm * nt, (int)nt, st));
        SWEM_TRY(record(p, 2, st));
    }
    return SWEM_OK;
}

static int model_impl(swem_plan* p, const float* m_dev, const float* sx, const float* sy, int nshot, float* d_dev,
                      hipStream_t st)
{
    if (!p || !d_dev || !sx || !m_dev || nshot < 1) return set_error(SWEM_EINVAL, "model: null argument or nshot < 1");
    SWEM_HIP(hipSetDevice(p->opt.device));
    if (p->three_d) return model3d_impl(p, m_dev, sx, sy, nshot, d_dev, st);
    const size_t nm = p->nm, nz = p->g.nz, nt = p->g.nt, nf = p->nf;
    const size_t per = nm * nf * sizeof(float2) + (size_t)p->nfc * nz * p->g.nmx * sizeof(float2);
    int chunk = (int)std::max<size_t>(1, std::min<size_t>(nshot, p->budget / std::max<size_t>(per, 1)));
    SWEM_TRY(p->dgw.alloc((size_t)chunk * nm * nf));
    SWEM_TRY(p->smig.alloc((size_t)chunk * p->nfc * nz * p->g.nmx));
    SWEM_TRY(p->mT.alloc(nm * nz));
    SWEM_TRY(set_igx(p, sx, sy, nshot, st));
    p->nev = 0;
    p->last_launches = 0;
    p->last_flops = 0;
    p->last_stream = st;
    SWEM_TRY(record(p, -1, st));
    SWEM_TRY(launch_transpose(m_dev, (int)nm, (int)nz, p->mT.p, st));
    SWEM_TRY(record(p, 0, st));
    for (int s0 = 0; s0 < nshot; s0 += chunk) {
        const int ns = std::min(chunk, nshot - s0);
        for (int c = 0; c < p->nchunks; c++) {
            if (p->resident != c) {
                SWEM_TRY(ensure_chunk(p, c, st));
                SWEM_TRY(record(p, 0, st));
            }
            March2DArgs a = base_args(p, c, p->Ffwd);
            a.nshot = ns;
            a.igx = p->igx.p + s0;
            a.mT = p->mT.p;
            a.smig = p->smig.p;
            a.dgw = p->dgw.p;
            SWEM_TRY(launch_march2d(a, p->nkx, p->Ffwd, mode_of(p), 0, p->opt.variant, st));
            SWEM_TRY(record(p, 1, st));
            p->last_launches++;
            p->last_flops += step_flops(p) * (double)a.nfc * nz * ns;
        }
        SWEM_TRY(launch_sgemm(ns * (int)nm, (int)nt, 2 * (int)nf, reinterpret_cast<const float*>(p->dgw.p), 2 * (int)nf,
                              p->basis_fwd, (int)nt, d_dev + (size_t)s0 * nm * nt, (int)nt, st));
        SWEM_TRY(record(p, 2, st));
    }
    return SWEM_OK;
}

}  // namespace swem

// ============================================================================ C ABI

extern "C" {

int swem_version(void) { return SWEM_VERSION; }

const char* swem_last_error(void) { return g_err.c_str(); }

void swem_options_init(swem_options* opt)
{
    if (!opt) return;
    std::memset(opt, 0, sizeof(*opt));
    opt->pspi_mode = SWEM_PSPI_COMPAT;
    opt->precision = SWEM_PREC_KZ64;
}

int swem_device_count(int* count)
{
    if (!count) return set_error(SWEM_EINVAL, "null count");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return set_error(SWEM_EDEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    return SWEM_OK;
}

}  // extern "C"

namespace swem {

// The sampling of ewem (ewem.c:43-52): ntfft = 2 truncf(padt nt / 2) (even), ifmax =
// trunc(fmax dt ntfft) + 1, ifmin = trunc(fmin dt ntfft), without wem.c's integer casts of fmax
// and fmin (wem.c:79-89).
static void ewem_sampling(const swem_geometry& g, int& ntfft, int& ifmin, int& ifmax)
{
    ntfft = (int)(2.f * std::trunc((float)g.padt * (float)g.nt / 2.f));
    const int nw = ntfft / 2 + 1;
    float t = g.fmax * g.dt * (float)ntfft;
    ifmax = t + 1 < (float)nw ? (int)std::trunc(t) + 1 : nw;
    t = g.fmin * g.dt * (float)ntfft;
    ifmin = t + 1 < (float)ifmax ? (int)std::trunc(t) : 0;
}

int plan_setup(swem_plan* p, const swem_geometry* g, const float* vel, const swem_options* opt)
{
    swem_options o;
    swem_options_init(&o);
    if (opt) o = *opt;
    if (o.pspi_mode != SWEM_PSPI_COMPAT && o.pspi_mode != SWEM_PSPI_LINEAR)
        return set_error(SWEM_EINVAL, "pspi_mode %d", o.pspi_mode);
    if (o.precision != SWEM_PREC_KZ64 && o.precision != SWEM_PREC_FP32)
        return set_error(SWEM_EINVAL, "precision %d", o.precision);
    if (o.variant < SWEM_VARIANT_AUTO || o.variant > SWEM_VARIANT_PACKED) return set_error(SWEM_EINVAL, "variant %d", o.variant);
    p->g = *g;
    p->opt = o;
    SWEM_TRY(compute_sampling(p->g, p->s));
    if (p->band_rule == 1) {
        ewem_sampling(p->g, p->s.ntfft, p->s.ifmin, p->s.ifmax);
        p->s.nw = p->s.ntfft / 2 + 1;
        p->s.fop_n = (p->s.nw - 1) * 2;
        if (p->s.fop_n < g->nt || p->s.fop_n < 2)
            return set_error(SWEM_EINVAL, "ewem: ntfft=%d is too short for nt=%d", p->s.ntfft, g->nt);
        p->s.dw = (float)(2 * kPi / ((double)(float)p->s.ntfft) / (double)g->dt);  // ewem.c:48
    }
    p->nm = g->nmx * g->nmy;
    p->nf = p->s.ifmax - p->s.ifmin;
    p->nkx = p->s.nkx;
    p->three_d = g->nmy > 1;
    p->narr = p->nkx;
    if (!p->three_d && march2d_blue_len(p->nkx)) p->narr = march2d_blue_len(p->nkx);
    if (p->three_d) {
        if (!march3d_supported(p->s.nkx, p->s.nky))
            return set_error(SWEM_EUNSUPPORTED,
                             "3-D: nkx=%d, nky=%d: this build handles 1..512, 2^a 3^b 5^c and powers of two to 1024",
                             p->s.nkx, p->s.nky);
        p->narr = march3d_len(p->s.nkx);
        p->narr_y = march3d_len(p->s.nky);
    } else if (!march2d_supported(p->nkx)) {
        return set_error(SWEM_EUNSUPPORTED,
                         "nkx = padx*nmx = %d: this build handles 1..2048, 2^a 3^b 5^c and powers of two to 4096", p->nkx);
    }
    if (p->nf < 1) return set_error(SWEM_EINVAL, "empty frequency band [ifmin=%d, ifmax=%d)", p->s.ifmin, p->s.ifmax);
    if (g->pspi && (g->nref < 2 || g->nref > 31)) return set_error(SWEM_EINVAL, "PSPI needs 2 <= nref <= 31 (nref=%d)", g->nref);
    if (!(g->dz != 0.f)) return set_error(SWEM_EINVAL, "dz must be non-zero");
    SWEM_TRY(check_device(o.device));
    if (p->three_d) {
        // y-pass workgroups: one wave per frequency (both wavefields in migration)
        if (o.freqs_per_block > 0) p->F3 = p->F3f = o.freqs_per_block;
        const int fmax = (g->pspi || p->narr_y != p->s.nky) ? 8 : 16;  // Bluestein y-passes: 512 threads
        if (p->F3 < 1 || p->F3 > fmax || p->F3f > fmax)
            return set_error(SWEM_EINVAL, "3-D: freqs_per_block=%d (1..%d)", o.freqs_per_block, fmax);
    } else {
        const March2DShape sa = march2d_shape(p->nkx, o.variant, 1), sf = march2d_shape(p->nkx, o.variant, 0);
        p->T = sa.T;
        p->F = choose_group(p, 1, o.freqs_per_block);
        p->Ffwd = choose_group(p, 0, o.freqs_per_block);
        if (sa.per_freq * p->F * sa.T > sa.max_block || march2d_lds_bytes(p->nkx, sa, p->F) > 160 * 1024 ||
            p->Ffwd * sf.T > sf.max_block || march2d_lds_bytes(p->nkx, sf, p->Ffwd) > 160 * 1024)
            return set_error(SWEM_EINVAL, "freqs_per_block=%d does not fit a workgroup at nkx=%d", o.freqs_per_block,
                             p->s.nkx);
        choose_chunks(p);
    }
    p->mode = mode_of(p);
    p->vel.assign(vel, vel + (size_t)p->nm * g->nz);
    p->budget = o.scratch_bytes ? o.scratch_bytes : default_budget();
    return SWEM_OK;
}

int plan_build(swem_plan* p, const float* wav)
{
    SWEM_TRY(build_tables(p, wav));
    if (!p->three_d) SWEM_TRY(alloc_tables(p));
    if (!p->served) SWEM_HIP(hipDeviceSynchronize());
    return SWEM_OK;
}

void plan_free(swem_plan* p)
{
    if (!p) return;
    (void)hipSetDevice(p->opt.device);
    if (!p->served && (p->last_stream || p->nev)) (void)hipStreamSynchronize(p->last_stream);
    for (hipEvent_t e : p->events) (void)hipEventDestroy(e);
    if (p->igx_pinned) (void)hipHostFree(p->igx_pinned);
    p->twy.release(); p->taper_y.release(); p->W.release(); p->wavd.release();
    p->tw.release(); p->chirp.release(); p->bchirp.release(); p->chirp_y.release(); p->bchirp_y.release(); p->spec_s.release(); p->ktab.release(); p->stab.release(); p->pd.release(); p->pdf.release();
    p->taper.release(); p->p2f.release(); p->w2.release();
    p->own_bases.adj.release(); p->own_bases.fwd.release(); p->own_bases.src.release();
    p->p2.release(); p->sel.release(); p->sel2.release();
    p->used.release(); p->cref.release(); p->R.release(); p->spec_r.release(); p->smig.release(); p->dgw.release(); p->partial.release();
    p->mT.release(); p->igx.release(); p->shot3d.release();
    if (p->graph3d_exec) (void)hipGraphExecDestroy(p->graph3d_exec);
    if (p->graph3d) (void)hipGraphDestroy(p->graph3d);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    delete p;
}

int plan_migrate(swem_plan* p, const float* d_dev, const float* sx, const float* sy, int nshot, float* m_dev,
                 int accumulate, int per_shot, hipStream_t st)
{
    return migrate_impl(p, d_dev, sx, sy, nshot, m_dev, accumulate, per_shot, st);
}

static size_t pad256(size_t b) { return (b + 511) & ~(size_t)255; }

size_t plan_upload_bytes(const swem_plan* p, int nshot)
{
    const size_t nz = p->g.nz, nm = p->nm, nmx = p->g.nmx, nref = p->g.pspi ? p->g.nref : 1;
    size_t b = pad256(p->narr * 8) + pad256(p->s.nky * 8) + pad256(nmx * 4) + pad256(p->g.nmy * 4) + pad256(p->g.nt * 4);
    b += pad256(nz * nref * 8) + pad256(nz * nref * 4) + pad256(nz * 4);
    if (p->g.pspi) {
        const size_t srow = p->three_d ? nm : (size_t)p->narr;
        b += 2 * pad256(nz * srow) + pad256(nz * srow * 4) + pad256(nz * kCrefStride * 4) + pad256(nz * 4);
    }
    else
        b += pad256(nz * nm * 8) + pad256(nz * nm * 4);
    const int kx = p->three_d ? march3d_kind(p->nkx) : march2d_kind(p->nkx);
    if (kx == kLenBluestein) b += pad256(p->nkx * 8) + pad256(p->narr * 16);
    if (p->three_d && march3d_kind(p->s.nky) == kLenBluestein)
        b += pad256(p->narr_y * 8) + pad256(p->s.nky * 8) + pad256(p->narr_y * 16);
    return b + pad256((size_t)nshot * 4);
}

size_t plan_migrate_bytes(const swem_plan* p, int nshot)
{
    const size_t nz = p->g.nz, nm = p->nm, nmx = p->g.nmx, nf = p->nf, nref = p->g.pspi ? p->g.nref : 1;
    size_t b = plan_upload_bytes(p, nshot) + pad256(nf * 8);
    if (!p->basis_cache) b += 3 * pad256((size_t)p->g.nt * 2 * nf * 4);
    if (p->three_d) {
        const int F = p->F3;
        const size_t per_item = 2 * nmx * p->s.nky * sizeof(float2) + nm * sizeof(float) / F + 1 + pspi3_slice_bytes(p);
        const size_t nfc = chunk3(p, per_item, F), gmax = (nfc + F - 1) / F;
        b += pad256(nfc * 2 * nmx * p->s.nky * 8) + pad256(gmax * nm * 4) + pad256(nm * nf * 8);
        if (p->g.pspi) b += pad256(nfc * pspi3_slice_bytes(p));
        return b;
    }
    const size_t nkh = p->nkx / 2 + 1;
    b += pad256((size_t)p->nfc * nz * nref * nkh * 8);
    if (!p->g.pspi) b += pad256((size_t)p->nfc * nz * nmx * 8);
    const size_t gmax = (size_t)(p->nfc + p->F - 1) / p->F;
    const size_t per = nm * nf * sizeof(float2) + gmax * nz * nmx * sizeof(float);
    const size_t chunk = std::max<size_t>(1, std::min<size_t>(nshot, p->budget / std::max<size_t>(per, 1)));
    b += pad256(chunk * nm * nf * 8) + pad256(chunk * gmax * nz * nmx * 4);
    return b;
}

}  // namespace swem

extern "C" {

static int plan_create_rule(const swem_geometry* g, const float* vel, const float* wav, const swem_options* opt,
                            int band_rule, swem_plan** out)
{
    if (!g || !vel || !wav || !out) return set_error(SWEM_EINVAL, "plan_create: null argument");
    *out = nullptr;
    swem_plan* p = new swem_plan();
    p->band_rule = band_rule;
    int rc = plan_setup(p, g, vel, opt);
    if (!rc) rc = plan_build(p, wav);
    if (rc) {
        plan_free(p);
        return rc;
    }
    *out = p;
    return SWEM_OK;
}

int swem_plan_create(const swem_geometry* g, const float* vel, const float* wav, const swem_options* opt,
                     swem_plan** out)
{
    if (!g || !vel || !wav || !out) return set_error(SWEM_EINVAL, "plan_create: null argument");
    *out = nullptr;
    swem_plan* p = new swem_plan();
    int rc = plan_setup(p, g, vel, opt);
    if (!rc) rc = plan_build(p, wav);
    if (rc) {
        plan_free(p);
        return rc;
    }
    *out = p;
    return SWEM_OK;
}

int swem_plan_destroy(swem_plan* p)
{
    plan_free(p);
    return SWEM_OK;
}

int swem_plan_migrate(swem_plan* p, const float* d_dev, const float* sx, const float* sy, int nshot, float* m_dev,
                      int accumulate, void* stream)
{
    return migrate_impl(p, d_dev, sx, sy, nshot, m_dev, accumulate, 0, (hipStream_t)stream);
}

int swem_plan_migrate_each(swem_plan* p, const float* d_dev, const float* sx, const float* sy, int nshot, float* m_dev,
                           void* stream)
{
    return migrate_impl(p, d_dev, sx, sy, nshot, m_dev, 0, 1, (hipStream_t)stream);
}

// The 2-D migration of migrate_impl with every work buffer carved out of the fixed device
// scratch (swem_noalloc.hip) instead of allocated: source indices, receiver spectra and partial
// images.  Shots run in chunks and the frequencies of a table chunk in sub-chunks sized so that
// the spectra of a shot chunk and the partial images of a sub-chunk fit the scratch; each
// sub-chunk's partial images are reduced into the image before the next launch reuses them.
// The kernels and the reduction are migrate_impl's, so the image equals swem_plan_migrate's up
// to the order of the frequency-chunk sums.  No timing events are recorded.
int swem_plan_migrate_noalloc(swem_plan* p, const float* d_dev, const float* sx, const float* sy, int nshot,
                              float* m_dev, int accumulate, void* stream)
{
    if (!p || !d_dev || !sx || !m_dev || nshot < 1)
        return set_error(SWEM_EINVAL, "migrate_noalloc: null argument or nshot < 1");
    if (p->three_d) return set_error(SWEM_EUNSUPPORTED, "migrate_noalloc: 2-D plans only (this plan is 3-D)");
    SWEM_HIP(hipSetDevice(p->opt.device));
    const hipStream_t st = (hipStream_t)stream;
    void* base = nullptr;
    size_t cap = 0;
    SWEM_TRY(noalloc_scratch(p->opt.device, &base, &cap));
    const size_t nm = p->nm, nz = p->g.nz, nt = p->g.nt, nf = p->nf, nmx = p->g.nmx;
    const int F = p->F;
    constexpr size_t kIgxBytes = size_t(64) << 10;  // source indices of one shot chunk
    constexpr int kIgxMax = (int)(kIgxBytes / sizeof(int));
    const auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t spec_shot = up(nm * nf * sizeof(float2));
    const size_t part_group = nz * nmx * sizeof(float);
    const size_t avail = cap - kIgxBytes;
    if (spec_shot + part_group > avail)
        return set_error(SWEM_ENOMEM,
                         "migrate_noalloc: one shot needs %zu bytes of work memory (spectra %zu, one partial image "
                         "%zu), above the fixed scratch of %zu bytes",
                         spec_shot + part_group, spec_shot, part_group, avail);
    // shots per chunk and groups per launch: the most workgroups per launch the scratch allows
    const int gtot = (p->nfc + F - 1) / F;
    int ns = 1, ngr = 1;
    long long best = 0;
    const int ns_max = (int)std::min<size_t>({(size_t)nshot, (size_t)kIgxMax, avail / (spec_shot + part_group)});
    for (int n = 1; n <= ns_max; n++) {
        const int g = (int)std::min<size_t>(gtot, (avail / n - spec_shot) / part_group);
        if (g >= 1 && (long long)n * g >= best) best = (long long)n * g, ns = n, ngr = g;
    }
    char* sc = static_cast<char*>(base);
    int* igx_dev = reinterpret_cast<int*>(sc);
    float2* spec = reinterpret_cast<float2*>(sc + kIgxBytes);
    float* part = reinterpret_cast<float*>(sc + kIgxBytes + (size_t)ns * spec_shot);
    int host_igx[kIntBlock];
    for (int s0 = 0; s0 < nshot; s0 += ns) {
        const int nsc = std::min(ns, nshot - s0);
        // source indices, written by a kernel whose arguments carry them (stream-ordered after the
        // previous chunk's kernels, which read the same scratch)
        for (int i0 = 0; i0 < nsc; i0 += kIntBlock) {
            const int n = std::min(kIntBlock, nsc - i0);
            for (int i = 0; i < n; i++) {
                int ix, iy;
                SWEM_TRY(source_index(p->g, sx[s0 + i0 + i], sy ? sy[s0 + i0 + i] : 0.f, ix, iy));
                host_igx[i] = ix * p->g.nmy + iy;
            }
            SWEM_TRY(launch_fill_ints(igx_dev + i0, host_igx, n, st));
        }
        SWEM_TRY(launch_sgemm(nsc * (int)nm, 2 * (int)nf, (int)nt, d_dev + (size_t)s0 * nm * nt, (int)nt, p->basis_adj,
                              2 * (int)nf, reinterpret_cast<float*>(spec), 2 * (int)nf, st));
        for (int c = 0; c < p->nchunks; c++) {
            if (p->resident != c) SWEM_TRY(ensure_chunk(p, c, st));
            const March2DArgs a0 = base_args(p, c, F);
            for (int j0 = 0; j0 < a0.nfc; j0 += ngr * F) {
                March2DArgs a = a0;
                a.nfc = std::min(ngr * F, a0.nfc - j0);
                a.f0 = a0.f0 + j0;
                a.ngroups = (a.nfc + F - 1) / F;
                a.ktab = a0.ktab + (size_t)j0 * nz * a.nref * a.nkh;
                if (a0.stab) a.stab = a0.stab + (size_t)j0 * nz * nmx;
                a.nshot = nsc;
                a.igx = igx_dev;
                a.spec_r = spec;
                a.partial = part;
                SWEM_TRY(launch_march2d(a, p->nkx, F, mode_of(p), 1, adj_variant(p), st));
                const int acc = accumulate || s0 > 0 || c > 0 || j0 > 0;
                SWEM_TRY(launch_reduce_image(part, nsc, a.ngroups, (int)nz, (int)nmx, m_dev, acc, 0, st));
            }
        }
    }
    return SWEM_OK;
}

int swem_plan_model(swem_plan* p, const float* m_dev, const float* sx, const float* sy, int nshot, float* d_dev,
                    void* stream)
{
    return model_impl(p, m_dev, sx, sy, nshot, d_dev, (hipStream_t)stream);
}

int swem_plan_timing(swem_plan* p, swem_timing* out)
{
    if (!p || !out) return set_error(SWEM_EINVAL, "plan_timing: null argument");
    std::memset(out, 0, sizeof(*out));
    if (p->nev < 2) return SWEM_OK;
    SWEM_HIP(hipSetDevice(p->opt.device));
    SWEM_HIP(hipEventSynchronize(p->events[p->nev - 1]));
    float ms = 0.f;
    SWEM_HIP(hipEventElapsedTime(&ms, p->events[0], p->events[p->nev - 1]));
    out->total_ms = ms;
    for (int i = 1; i < p->nev; i++) {
        SWEM_HIP(hipEventElapsedTime(&ms, p->events[i - 1], p->events[i]));
        if (p->ev_kind[i] == 0) out->spectra_ms += ms;
        else if (p->ev_kind[i] == 1) out->march_ms += ms;
        else out->reduce_ms += ms;
    }
    out->march_launches = p->last_launches;
    out->march_flops = p->last_flops;
    return SWEM_OK;
}

// Measured choice of the 2-D migration kernel: every variant whose shape fits the plan's nkx is
// timed on nshot records of pseudo-random values and the fastest becomes the plan's variant.
// times[SWEM_VARIANT_PACKED + 1] receives each variant's march time in ms (0: not a distinct
// candidate).
// Records of nshot shots spread over the line and an image buffer, for timing candidate launch
// shapes.  The march does the same work for any records whose spectra are nonzero; 2-D records
// are pseudo-random, since the packed kernel writes a zero image without marching a frequency
// whose receiver spectrum is zero; 3-D ones are zero (no 3-D kernel skips work on them).
struct TuneBuffers {
    float* d = nullptr;
    float* m = nullptr;
    std::vector<float> sx, sy;
    ~TuneBuffers()
    {
        (void)hipDeviceSynchronize();
        if (d) (void)hipFree(d);
        if (m) (void)hipFree(m);
    }
    int init(const swem_plan* p, int nshot)
    {
        const size_t nm = p->nm, nt = p->g.nt, nz = p->g.nz;
        SWEM_HIP(hipMalloc((void**)&d, (size_t)nshot * nm * nt * sizeof(float)));
        SWEM_HIP(hipMalloc((void**)&m, nm * nz * sizeof(float)));
        if (p->three_d) {
            SWEM_HIP(hipMemset(d, 0, (size_t)nshot * nm * nt * sizeof(float)));
        } else {
            std::vector<float> h((size_t)nshot * nm * nt);
            uint32_t x = 2463534242u;  // xorshift32
            for (float& v : h) {
                x ^= x << 13;
                x ^= x >> 17;
                x ^= x << 5;
                v = (float)(int32_t)x * (1.f / 2147483648.f);
            }
            SWEM_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
        }
        sx.resize(nshot);
        sy.assign(nshot, p->g.omy + p->g.dmy * (float)(p->g.nmy / 2));
        for (int i = 0; i < nshot; i++)
            sx[i] = p->g.omx + p->g.dmx * (float)((i * (p->g.nmx - 1)) / std::max(1, nshot - 1));
        return SWEM_OK;
    }
};

// Time one candidate: a warm-up migration, then the timed one (march kernels only).
static int tune_time(swem_plan* p, TuneBuffers& b, int nshot, float& ms)
{
    swem_timing t{};
    for (int rep = 0; rep < 2; rep++) {
        SWEM_TRY(migrate_impl(p, b.d, b.sx.data(), b.sy.data(), nshot, b.m, 0, 0, nullptr));
        SWEM_TRY(swem_plan_timing(p, &t));
    }
    ms = (float)t.march_ms;
    return SWEM_OK;
}

static void tune_add(swem_tune_result* r, int variant, int fpb, size_t chunk, float ms)
{
    if (r->ncand >= SWEM_TUNE_MAX) return;
    r->cand_variant[r->ncand] = variant;
    r->cand_fpb[r->ncand] = fpb;
    r->cand_chunk_bytes[r->ncand] = chunk;
    r->cand_ms[r->ncand] = ms;
    r->ncand++;
}

// 2-D: every migration kernel variant that fits nkx, each with every frequency group size that
// fits its workgroup and LDS (1, 2, 4, 8, 16 frequencies per workgroup).
static int autotune2d(swem_plan* p, int nshot, swem_tune_result* r)
{
    TuneBuffers b;
    SWEM_TRY(b.init(p, nshot));
    const swem_options keep = p->opt;
    const int keepF = p->F;
    int rc = SWEM_OK;
    int seen[SWEM_VARIANT_PACKED + 1] = {};
    for (int v = SWEM_VARIANT_SUBGROUP; v <= SWEM_VARIANT_PACKED && !rc; v++) {
        if (v == SWEM_VARIANT_PACKED && !p->packed_ok) continue;
        const March2DShape sh = march2d_shape(p->nkx, v, 1);
        if (sh.T <= 0 || seen[sh.variant]) continue;  // not available, or the same kernel as an earlier one
        seen[sh.variant] = 1;
        p->opt.variant = v;
        for (int F = 1; F <= 16 && !rc; F *= 2) {
            if (v == SWEM_VARIANT_PACKED && F > 1) break;  // one frequency per packed workgroup
            if (keep.freqs_per_block > 0 && F != keep.freqs_per_block) continue;
            if (sh.per_freq * F * sh.T > sh.max_block || march2d_lds_bytes(p->nkx, sh, F) > 160 * 1024) break;
            if (F > 1 && (p->nfc + F - 1) / F < 1) break;
            p->F = F;
            float ms = 0.f;
            rc = tune_time(p, b, nshot, ms);
            if (!rc) tune_add(r, v, F, 0, ms);
        }
    }
    p->opt = keep;
    p->F = keepF;
    return rc;
}

// 3-D: frequencies per y-pass workgroup (2, 4, 8 and 16 where the y-pass kernel allows it) and the
// step state of one launch (64, 128, 256 MiB: the frequencies marched together, against the
// 256 MB last-level cache), timed on the first 16 depths of the plan (every depth is the same
// work).
static int autotune3d(swem_plan* p, int nshot, swem_tune_result* r)
{
    TuneBuffers b;
    SWEM_TRY(b.init(p, nshot));
    const int keepF = p->F3;
    const size_t keepS = p->state3d_bytes;
    const int fmax = (p->g.pspi || p->narr_y != p->s.nky) ? 8 : 16;
    p->tune_nz = std::min(16, p->g.nz);
    int rc = SWEM_OK;
    for (int F = 2; F <= fmax && !rc; F *= 2) {
        if (p->opt.freqs_per_block > 0 && F != p->opt.freqs_per_block) continue;
        for (size_t S = size_t(64) << 20; S <= (size_t(256) << 20) && !rc; S *= 2) {
            p->F3 = F;
            p->state3d_bytes = S;
            float ms = 0.f;
            rc = tune_time(p, b, nshot, ms);
            if (!rc) tune_add(r, SWEM_VARIANT_AUTO, F, S, ms);
        }
    }
    p->tune_nz = 0;
    p->F3 = keepF;
    p->state3d_bytes = keepS;
    return rc;
}

int swem_plan_autotune_ex(swem_plan* p, int nshot, swem_tune_result* out)
{
    if (!p || nshot < 1 || !out) return set_error(SWEM_EINVAL, "autotune: null plan, null result or nshot < 1");
    SWEM_HIP(hipSetDevice(p->opt.device));
    std::memset(out, 0, sizeof(*out));
    SWEM_TRY(p->three_d ? autotune3d(p, nshot, out) : autotune2d(p, nshot, out));
    int best = -1;
    for (int i = 0; i < out->ncand; i++)
        if (best < 0 || out->cand_ms[i] < out->cand_ms[best]) best = i;
    if (best < 0) return set_error(SWEM_EUNSUPPORTED, "autotune: no kernel shape fits this plan (nkx=%d)", p->nkx);
    out->variant = out->cand_variant[best];
    out->freqs_per_block = out->cand_fpb[best];
    out->chunk_bytes = out->cand_chunk_bytes[best];
    out->best_ms = out->cand_ms[best];
    if (p->three_d) {
        p->F3 = out->freqs_per_block;
        p->state3d_bytes = out->chunk_bytes;
    } else {
        p->opt.variant = out->variant;
        p->F = out->freqs_per_block;
    }
    return SWEM_OK;
}

int swem_plan_autotune(swem_plan* p, int nshot, int* best, float* times)
{
    if (!p || nshot < 1) return set_error(SWEM_EINVAL, "autotune: null plan or nshot < 1");
    swem_tune_result r;
    SWEM_TRY(swem_plan_autotune_ex(p, nshot, &r));
    if (times) {
        for (int v = 0; v <= SWEM_VARIANT_PACKED; v++) times[v] = 0.f;
        for (int i = 0; i < r.ncand; i++) {  // each variant's best group size
            float& t = times[r.cand_variant[i]];
            if (t == 0.f || r.cand_ms[i] < t) t = r.cand_ms[i];
        }
    }
    if (best) *best = r.variant;
    return SWEM_OK;
}

int swem_plan_set_graphs(swem_plan* p, int enable)
{
    if (!p) return set_error(SWEM_EINVAL, "plan_set_graphs: null plan");
    p->use_graphs = enable ? 1 : 0;
    return SWEM_OK;
}

int swem_plan_bytes(const swem_plan* p, size_t* table_bytes, size_t* work_bytes)
{
    if (!p) return set_error(SWEM_EINVAL, "plan_bytes: null plan");
    if (table_bytes)
        *table_bytes = p->tw.bytes() + p->spec_s.bytes() + p->ktab.bytes() + p->stab.bytes() + p->pd.bytes() +
                       p->pdf.bytes() + p->taper.bytes() + p->p2f.bytes() +
                       p->w2.bytes() + p->own_bases.adj.bytes() + p->own_bases.fwd.bytes() + p->own_bases.src.bytes() + p->wavd.bytes() + p->p2.bytes() + p->sel.bytes() +
                       p->sel2.bytes() + p->used.bytes() + p->chirp.bytes() + p->bchirp.bytes() +
                       p->twy.bytes() + p->chirp_y.bytes() + p->bchirp_y.bytes() + p->taper_y.bytes();
    if (work_bytes)
        *work_bytes = p->spec_r.bytes() + p->smig.bytes() + p->dgw.bytes() + p->partial.bytes() + p->mT.bytes() +
                      p->igx.bytes();
    return SWEM_OK;
}

static void geom_from_args(swem_geometry* g, int nt, float ot, float dt, int nmx, float omx, float dmx, int nmy,
                           float omy, float dmy, int nz, float oz, float dz, float gz, float sz, int nref, float fmin,
                           float fmax, int padt, int padx, int pspi)
{
    g->nt = nt; g->ot = ot; g->dt = dt;
    g->nmx = nmx; g->omx = omx; g->dmx = dmx;
    g->nmy = nmy; g->omy = omy; g->dmy = dmy;
    g->nz = nz; g->oz = oz; g->dz = dz;
    g->gz = gz; g->sz = sz;
    g->nref = nref; g->fmin = fmin; g->fmax = fmax;
    g->padt = padt; g->padx = padx; g->pspi = pspi;
}

int swem_shots(const swem_geometry* g, const float* vel, const float* wav, const float* sx, const float* sy, int nshot,
               float* d, float* m, int adj, const swem_options* opt)
{
    if (!g || !vel || !wav || !sx || !d || !m || nshot < 1) return set_error(SWEM_EINVAL, "shots: null argument");
    swem_plan* p = nullptr;
    SWEM_TRY(swem_plan_create(g, vel, wav, opt, &p));
    const size_t nm = (size_t)g->nmx * g->nmy, dbytes = (size_t)nshot * nm * g->nt * sizeof(float),
                 mbytes = nm * g->nz * sizeof(float);
    float *dd = nullptr, *dm = nullptr;
    int rc = SWEM_OK;
    hipError_t e = hipMalloc((void**)&dd, dbytes);
    if (e == hipSuccess) e = hipMalloc((void**)&dm, mbytes);
    if (e != hipSuccess) rc = hip_error(e, "hipMalloc", __FILE__, __LINE__);
    if (!rc) {
        if (adj) {
            e = hipMemcpy(dd, d, dbytes, hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = hip_error(e, "hipMemcpy", __FILE__, __LINE__);
            if (!rc) rc = swem_plan_migrate(p, dd, sx, sy, nshot, dm, 0, nullptr);
            if (!rc) {
                e = hipMemcpy(m, dm, mbytes, hipMemcpyDeviceToHost);
                if (e != hipSuccess) rc = hip_error(e, "hipMemcpy", __FILE__, __LINE__);
            }
        } else {
            e = hipMemcpy(dm, m, mbytes, hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = hip_error(e, "hipMemcpy", __FILE__, __LINE__);
            if (!rc) rc = swem_plan_model(p, dm, sx, sy, nshot, dd, nullptr);
            if (!rc) {
                e = hipMemcpy(d, dd, dbytes, hipMemcpyDeviceToHost);
                if (e != hipSuccess) rc = hip_error(e, "hipMemcpy", __FILE__, __LINE__);
            }
        }
    }
    if (dd) (void)hipFree(dd);
    if (dm) (void)hipFree(dm);
    swem_plan_destroy(p);
    return rc;
}

int swem_wem_ex(float* d, float* m, const float* wav, const swem_geometry* g, float sx, float sy, const float* vel,
                int adj, const swem_options* opt)
{
    return swem_shots(g, vel, wav, &sx, &sy, 1, d, m, adj, opt);
}

int swem_wem(float* d, float* m, const float* wav, int nt, float ot, float dt, int nmx, float omx, float dmx, int nmy,
             float omy, float dmy, float sx, float sy, int nz, float oz, float dz, float gz, float sz, const float* vel,
             int nref, float fmin, float fmax, int padt, int padx, int adj, int pspi, int verbose)
{
    swem_geometry g;
    geom_from_args(&g, nt, ot, dt, nmx, omx, dmx, nmy, omy, dmy, nz, oz, dz, gz, sz, nref, fmin, fmax, padt, padx, pspi);
    swem_options o;
    swem_options_init(&o);
    o.verbose = verbose;
    return swem_wem_ex(d, m, wav, &g, sx, sy, vel, adj, &o);
}

// compute_angles (wem.c:566-692) on the device, 2-D: the plan is built with padt = padx = 1 and
// split-step (ntfft = nt, nkx = nmx, wem.c:588-594); per frequency chunk the source march stores
// the volume, the column pass its z derivative and the row pass sums the angles and signs.
static int angles3d_impl(swem_plan* p, float sx, float sy, float* angx, float* angy);
static int angle_z_tables(int nz, DevBuf<float2>& twz, DevBuf<float2>& chz, DevBuf<float2>& bchz, March2DArgs& az);

static int angles_impl(swem_plan* p, float sx, float sy, float* angx, float* angy)
{
    if (p->three_d) return angles3d_impl(p, sx, sy, angx, angy);
    hipStream_t st = nullptr;
    const int nmx = p->g.nmx, nz = p->g.nz;
    const size_t nmz = (size_t)nmx * nz;
    for (size_t i = 0; i < nmz; i++) angx[i] = angy[i] = 0.f;
    if (p->nf < 1) return SWEM_OK;
    DevBuf<float2> twz, chz, bchz, vol;
    DevBuf<float> dzmag, acc, sgn;
    March2DArgs az;
    SWEM_TRY(angle_z_tables(nz, twz, chz, bchz, az));
    // frequencies per pass: the volume (complex) and |u_z| (float) of each within half the budget
    const int nsub = (int)std::max<size_t>(1, std::min<size_t>(p->nfc, p->budget / 2 / (nmz * 12)));
    SWEM_TRY(vol.alloc((size_t)nsub * nmz));
    SWEM_TRY(dzmag.alloc((size_t)nsub * nmz));
    SWEM_TRY(acc.alloc(nmz));
    SWEM_TRY(sgn.alloc(nmz));
    SWEM_HIP(hipMemsetAsync(acc.p, 0, nmz * sizeof(float), st));
    SWEM_HIP(hipMemsetAsync(sgn.p, 0, nmz * sizeof(float), st));
    SWEM_TRY(set_igx(p, &sx, &sy, 1, st));
    const int T = p->narr >= 1024 ? p->narr / 8 : 64;
    const size_t lds = ((size_t)(p->narr + p->narr / 8 + 31) / 32 * 32 + 4) * 2 * sizeof(float2);
    int F = std::max(1, 512 / T);
    while (F > 1 && F * lds > 80 * 1024) F--;
    AngleArgs b{};
    b.nz = nz;
    b.nw = p->s.nw;
    b.dkz = (float)(2.0 * kPi / (double)(float)nz / (double)p->g.dz);
    b.item = (long long)nmz;
    b.len = nmx;
    b.dk = p->s.dkx;
    b.vol = vol.p;
    b.dzmag = dzmag.p;
    b.acc = acc.p;
    b.sgn = sgn.p;
    for (int c = 0; c < p->nchunks; c++) {
        SWEM_TRY(ensure_chunk(p, c, st));
        const March2DArgs a0 = base_args(p, c, F);
        for (int f1 = 0; f1 < a0.nfc; f1 += nsub) {  // a pass over frequencies f1 .. of the chunk's tables
            March2DArgs ax = a0;
            ax.f0 = a0.f0 + f1;
            ax.nfc = std::min(nsub, a0.nfc - f1);
            ax.ngroups = (ax.nfc + F - 1) / F;
            ax.ktab = a0.ktab + (size_t)f1 * nz * a0.nref * a0.nkh;
            ax.stab = a0.stab + (size_t)f1 * nmz;
            ax.nshot = 1;
            ax.igx = p->igx.p;
            ax.smig = vol.p;
            b.nfc = ax.nfc;
            SWEM_TRY(launch_angles(0, ax, az, b, F, st));
            // z pass: the z line of every (frequency, x)
            b.inner = nmx, b.lines = ax.nfc * nmx, b.outer = (long long)nmz, b.inner_stride = 1, b.elem = nmx;
            SWEM_TRY(launch_angles(1, ax, az, b, F, st));
            // x pass: the x line of every depth
            b.inner = 1, b.lines = nz, b.outer = nmx, b.inner_stride = 0, b.elem = 1;
            SWEM_TRY(launch_angles(2, ax, az, b, F, st));
        }
    }
    std::vector<float> ha(nmz), hs(nmz);
    SWEM_HIP(hipMemcpyAsync(ha.data(), acc.p, nmz * sizeof(float), hipMemcpyDeviceToHost, st));
    SWEM_HIP(hipMemcpyAsync(hs.data(), sgn.p, nmz * sizeof(float), hipMemcpyDeviceToHost, st));
    SWEM_HIP(hipStreamSynchronize(st));
    // angx *= signf1(angx_sign) / (ifmax - ifmin + 1) (wem.c:671); angy stays 0 in 2-D (ky = 0)
    for (int ix = 0; ix < nmx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            const size_t i = (size_t)iz * nmx + ix;
            angx[(size_t)ix * nz + iz] = ha[i] * ((hs[i] >= 0.f ? 1.f : -1.f) / (float)(p->nf + 1));
        }
    return SWEM_OK;
}

// z tables of the angle passes: the transform tables of nz on its 2-D array length.
static int angle_z_tables(int nz, DevBuf<float2>& twz, DevBuf<float2>& chz, DevBuf<float2>& bchz, March2DArgs& az)
{
    const int mz = march2d_blue_len(nz), nza = mz ? mz : nz;
    az = March2DArgs{};
    SWEM_TRY(length_tables(nz, nza, mz ? march2d_kind(nz) : kLenPow2, twz, chz, bchz, az.mr));
    az.nkx = nz;
    az.tw = twz.p;
    az.chirp = chz.p;
    az.bchirp = bchz.p;
    return SWEM_OK;
}

// compute_angles in 3-D (nkx = nmx, nky = nmy; wem.c:588-594): per chunk of frequencies the
// source march of the 3-D modelling receiver (x-pass with conj(L), y-pass phase 3) stores every
// depth slice, then the z pass and the x and y passes add the angles of both axes.
static int angles3d_impl(swem_plan* p, float sx, float sy, float* angx, float* angy)
{
    hipStream_t st = nullptr;
    const int nmx = p->g.nmx, nmy = p->g.nmy, nz = p->g.nz;
    const size_t nm = p->nm, nmz = nm * nz;
    for (size_t i = 0; i < nmz; i++) angx[i] = angy[i] = 0.f;
    if (p->nf < 1) return SWEM_OK;
    DevBuf<float2> twz, chz, bchz, vol, W;
    DevBuf<float> dzmag, accx, sgx, accy, sgy;
    March2DArgs az;
    SWEM_TRY(angle_z_tables(nz, twz, chz, bchz, az));
    const size_t per_item = nmz * (sizeof(float2) + sizeof(float)) + 2 * (size_t)nmx * p->s.nky * sizeof(float2);
    const size_t cap = std::min<size_t>(p->budget / 2, (size_t)16 << 30);
    const int nfc = (int)std::max<size_t>(1, std::min<size_t>(p->nf, cap / per_item));
    SWEM_TRY(vol.alloc((size_t)nfc * nmz));
    SWEM_TRY(dzmag.alloc((size_t)nfc * nmz));
    SWEM_TRY(W.alloc((size_t)nfc * 2 * nmx * p->s.nky));
    for (DevBuf<float>* b : {&accx, &sgx, &accy, &sgy}) {
        SWEM_TRY(b->alloc(nmz));
        SWEM_HIP(hipMemsetAsync(b->p, 0, nmz * sizeof(float), st));
    }
    SWEM_TRY(set_igx(p, &sx, &sy, 1, st));
    const int F = p->F3f;
    March2DArgs ax{}, ay{};
    ax.nkx = nmx;  // nkx = nmx, nky = nmy (padx = 1); Bluestein tables where the march has them
    ax.tw = p->tw.p;
    ax.mr = p->mrx;
    ax.chirp = p->chirp.p;
    ax.bchirp = p->bchirp.p;
    ay.nkx = nmy;
    ay.tw = p->twy.p;
    ay.mr = p->mry;
    ay.chirp = p->chirp_y.p;
    ay.bchirp = p->bchirp_y.p;
    AngleArgs b{};
    b.nz = nz;
    b.nw = p->s.nw;
    b.dkz = (float)(2.0 * kPi / (double)(float)nz / (double)p->g.dz);
    b.item = (long long)nmz;
    b.vol = vol.p;
    b.dzmag = dzmag.p;
    for (int f0 = 0; f0 < p->nf; f0 += nfc) {
        March3DArgs a = base3(p);
        a.f0 = f0;
        a.nitem = std::min(nfc, p->nf - f0);
        a.igx = p->igx_host[0];
        a.adj = 0;
        a.W = W.p;
        a.smig = vol.p;
        a.conj_src = 1;
        a.conj_rec = 1;
        a.wmask = 2;
        SWEM_HIP(hipMemsetAsync(vol.p, 0, (size_t)a.nitem * nmz * sizeof(float2), st));
        // extrapolate_source (wem.c:712-719): the depths z >= sz step; the others stay zero
        int prev = -1;
        for (int iz = 0; iz < nz; iz++) {
            if (!stepped_at(p, iz, p->g.sz)) continue;
            a.iz = iz;
            a.iz_sel = prev;
            a.first = prev < 0;
            a.ds = 1;
            SWEM_TRY(launch_march3d(3, a, F, 3, st));
            SWEM_TRY(launch_march3d(1, a, F, 0, st));
            prev = iz;
        }
        if (prev >= 0) {
            a.iz_sel = prev;
            a.first = 0;
            a.ds = 0;
            SWEM_TRY(launch_march3d(3, a, F, 3, st));
        }
        b.nfc = a.nitem;
        // z pass over every (frequency, trace)
        b.len = nmx, b.dk = p->s.dkx;
        b.inner = (int)nm, b.lines = a.nitem * (int)nm, b.outer = (long long)nmz, b.inner_stride = 1, b.elem = (long long)nm;
        SWEM_TRY(launch_angles(1, ax, az, b, 1, st));
        // x lines: (iz, y), points nmy apart
        b.acc = accx.p, b.sgn = sgx.p;
        b.len = nmx, b.dk = p->s.dkx;
        b.inner = nmy, b.lines = nz * nmy, b.outer = (long long)nm, b.inner_stride = 1, b.elem = nmy;
        SWEM_TRY(launch_angles(2, ax, az, b, 1, st));
        // y lines: (iz, x), contiguous
        b.acc = accy.p, b.sgn = sgy.p;
        b.len = nmy, b.dk = p->s.dky;
        b.inner = nmx, b.lines = nz * nmx, b.outer = (long long)nm, b.inner_stride = nmy, b.elem = 1;
        SWEM_TRY(launch_angles(2, ay, az, b, 1, st));
    }
    std::vector<float> hax(nmz), hsx(nmz), hay(nmz), hsy(nmz);
    SWEM_HIP(hipMemcpyAsync(hax.data(), accx.p, nmz * sizeof(float), hipMemcpyDeviceToHost, st));
    SWEM_HIP(hipMemcpyAsync(hsx.data(), sgx.p, nmz * sizeof(float), hipMemcpyDeviceToHost, st));
    SWEM_HIP(hipMemcpyAsync(hay.data(), accy.p, nmz * sizeof(float), hipMemcpyDeviceToHost, st));
    SWEM_HIP(hipMemcpyAsync(hsy.data(), sgy.p, nmz * sizeof(float), hipMemcpyDeviceToHost, st));
    SWEM_HIP(hipStreamSynchronize(st));
    const float inv = 1.f / (float)(p->nf + 1);
    for (size_t i = 0; i < nm; i++)
        for (int iz = 0; iz < nz; iz++) {
            const size_t k = (size_t)iz * nm + i;
            angx[i * nz + iz] = hax[k] * ((hsx[k] >= 0.f ? 1.f : -1.f) * inv);
            angy[i * nz + iz] = hay[k] * ((hsy[k] >= 0.f ? 1.f : -1.f) * inv);
        }
    return SWEM_OK;
}

int swem_compute_angles_ex(float* angx, float* angy, const float* wav, const swem_geometry* g0, float sx, float sy,
                           const float* vel, const swem_options* opt)
{
    if (!angx || !angy || !wav || !vel || !g0) return set_error(SWEM_EINVAL, "compute_angles: null argument");
    const swem_geometry& g = *g0;
    if (g.nmy > 1 && !march3d_supported(g.nmx, g.nmy))
        return set_error(SWEM_EUNSUPPORTED, "compute_angles: 3-D takes nmx and nmy in 2..512 or powers of two to 1024 "
                         "(%d x %d)", g.nmx, g.nmy);
    if (g.nmx < 2 || !angles_supported(g.nmx) || !angles_supported(g.nz))
        return set_error(SWEM_EUNSUPPORTED, "compute_angles: nmx = %d, nz = %d: this build handles 2..2048 and powers of "
                         "two to 4096 (nmx >= 2)", g.nmx, g.nz);
    swem_geometry ga = g;
    ga.padt = ga.padx = 1;
    ga.pspi = 0;
    ga.nref = 1;
    ga.gz = 0.f;
    swem_options o;
    if (opt) o = *opt;
    else swem_options_init(&o);
    swem_plan* p = nullptr;
    SWEM_TRY(swem_plan_create(&ga, vel, wav, &o, &p));
    const int rc = angles_impl(p, sx, sy, angx, angy);
    swem_plan_destroy(p);
    return rc;
}

int swem_compute_angles(float* angx, float* angy, const float* wav, int nt, float ot, float dt, int nmx, float omx,
                        float dmx, int nmy, float omy, float dmy, float sx, float sy, int nz, float oz, float dz,
                        float sz, const float* vel, float fmin, float fmax, int verbose)
{
    swem_geometry g;
    geom_from_args(&g, nt, ot, dt, nmx, omx, dmx, nmy, omy, dmy, nz, oz, dz, 0.f, sz, 1, fmin, fmax, 1, 1, 0);
    swem_options o;
    swem_options_init(&o);
    o.verbose = verbose;
    return swem_compute_angles_ex(angx, angy, wav, &g, sx, sy, vel, &o);
}

// ------------------------------------------------------------------------ elastic WEM (2-D)
// The sampling of ewem (ewem.c:46-57), which truncates fmax dt ntfft where wem.c casts fmax.

// nshot shots of elastic migration (records [nshot][nm][nt] -> images [nm][nz] summed over shots)
// or modelling (images -> records of every shot), all shots of a chunk in one launch.
// Host tables of elastic shots, 2-D and 3-D: the Fourier bases without w^2 (ewem.c:116-140,
// 219-241: R2C / sqrt(ntfft) and the c2r of the band), the layer velocities of the separation in
// float as ewem.c:64-81 forms them (1 / po, po the inverse of the mean velocity over the nmx nmy
// traces), and max|wav| / sqrt(ntfft) (ewem.c:146).
struct EwemTables {
    std::vector<float> badj, bfwd, vp0, vs0;
    float maxs = 0.f;
};

static int ewem_tables(const swem_plan* pp, const swem_plan* ps, const float* wav, EwemTables& t)
{
    const int nt = pp->g.nt, nf = pp->nf, nz = pp->g.nz, nm = pp->g.nmx * pp->g.nmy;
    const int n = pp->s.fop_n;
    const double rnt = 1.0 / std::sqrt((double)pp->s.ntfft);
    t.badj.assign((size_t)nt * 2 * nf, 0.f);
    t.bfwd.assign((size_t)2 * nf * nt, 0.f);
    for (int f = 0; f < nf; f++) {
        const int iw = pp->s.ifmin + f;
        const double cf = (iw == 0 || 2 * iw == n) ? 1.0 : 2.0;
        for (int it = 0; it < nt; it++) {
            const long long e = ((long long)iw * it) % n;
            const double ang = 2.0 * kPi * (double)e / (double)n, c = std::cos(ang), sn = std::sin(ang);
            t.badj[(size_t)it * 2 * nf + 2 * f] = (float)(rnt * c);
            t.badj[(size_t)it * 2 * nf + 2 * f + 1] = (float)(-rnt * sn);
            t.bfwd[(size_t)(2 * f) * nt + it] = (float)(cf * rnt * c);
            t.bfwd[(size_t)(2 * f + 1) * nt + it] = (float)(-cf * rnt * sn);
        }
    }
    t.vp0.assign(nz, 0.f);
    t.vs0.assign(nz, 0.f);
    for (int k = 0; k < 2; k++) {
        const std::vector<float>& v = k == 0 ? pp->vel : ps->vel;
        for (int iz = 0; iz < nz; iz++) {
            float sum = 0.f;
            for (int ix = 0; ix < nm; ix++) sum += v[(size_t)ix * nz + iz];
            const float po = 1.f / (sum / (float)nm);
            (k == 0 ? t.vp0 : t.vs0)[iz] = (float)(1.0 / (double)po);
        }
    }
    const float rntf = std::sqrt((float)pp->s.ntfft);
    t.maxs = 0.f;
    for (int it = 0; it < nt; it++) t.maxs = std::max(t.maxs, std::fabs(wav[it]) / rntf);
    if (!(t.maxs > 0.f)) return set_error(SWEM_EINVAL, "ewem: the wavelet is zero");
    return SWEM_OK;
}

static int eshots_impl(swem_plan* pp, swem_plan* ps, int nshot, const int* igx, float* ux, float* uy, float* uz,
                       float* mpp, float* mps1, float* mps2, const float* wav, int adj)
{
    hipStream_t st = nullptr;
    const int nmx = pp->g.nmx, nz = pp->g.nz, nt = pp->g.nt, nf = pp->nf;
    const size_t nmz = (size_t)nmx * nz, nmt = (size_t)nmx * nt;
    if (adj) {
        std::fill(mpp, mpp + nmz, 0.f);
        std::fill(mps1, mps1 + nmz, 0.f);
        std::fill(mps2, mps2 + nmz, 0.f);
    } else {
        std::fill(ux, ux + nshot * nmt, 0.f);
        std::fill(uy, uy + nshot * nmt, 0.f);
        std::fill(uz, uz + nshot * nmt, 0.f);
    }
    if (nf < 1) return SWEM_OK;
    if (pp->nchunks != ps->nchunks || pp->nfc != ps->nfc)
        return set_error(SWEM_EINVAL, "ewem: the P and S plans chunk the band differently");
    EwemTables tb;
    SWEM_TRY(ewem_tables(pp, ps, wav, tb));
    const std::vector<float>&badj = tb.badj, &bfwd = tb.bfwd, &vp0 = tb.vp0, &vs0 = tb.vs0;
    const float maxs = tb.maxs;
    DevBuf<float> dbadj, dbfwd, dvp0, dvs0, drx, drz, ppart, spart, dmpp, dmps, mppT, mpsT;
    DevBuf<float2> spx, spz, smig, dgx, dgz;
    DevBuf<int> dig;
    SWEM_TRY(dvp0.upload(vp0.data(), nz));
    SWEM_TRY(dvs0.upload(vs0.data(), nz));
    SWEM_TRY(dig.upload(igx, nshot));
    const size_t nfc = (size_t)pp->nfc;
    // shots per launch: records, spectra and the per-frequency partial images or stored sources
    const size_t per_shot = 2 * nmt * sizeof(float) + 2 * (size_t)nmx * nf * sizeof(float2) +
                            (adj ? 2 * nfc * nmz * sizeof(float) : nfc * nmz * sizeof(float2));
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(nshot, pp->budget / std::max<size_t>(per_shot, 1)));
    SWEM_TRY(drx.alloc((size_t)chunk * nmt));
    SWEM_TRY(drz.alloc((size_t)chunk * nmt));
    if (adj) {
        SWEM_TRY(dbadj.upload(badj.data(), badj.size()));
        SWEM_TRY(spx.alloc((size_t)chunk * nmx * nf));
        SWEM_TRY(spz.alloc((size_t)chunk * nmx * nf));
        SWEM_TRY(ppart.alloc((size_t)chunk * nfc * nmz));
        SWEM_TRY(spart.alloc((size_t)chunk * nfc * nmz));
        SWEM_TRY(dmpp.alloc(nmz));
        SWEM_TRY(dmps.alloc(nmz));
    } else {
        SWEM_TRY(dbfwd.upload(bfwd.data(), bfwd.size()));
        std::vector<float> t1(nmz), t2(nmz);
        for (int ix = 0; ix < nmx; ix++)
            for (int iz = 0; iz < nz; iz++) {
                t1[(size_t)iz * nmx + ix] = mpp[(size_t)ix * nz + iz];
                t2[(size_t)iz * nmx + ix] = mps2[(size_t)ix * nz + iz];
            }
        SWEM_TRY(mppT.upload(t1.data(), nmz));
        SWEM_TRY(mpsT.upload(t2.data(), nmz));
        SWEM_TRY(smig.alloc((size_t)chunk * nfc * nmz));
        SWEM_TRY(dgx.alloc((size_t)chunk * nmx * nf));
        SWEM_TRY(dgz.alloc((size_t)chunk * nmx * nf));
    }
    bool first_image = true;
    for (int s0 = 0; s0 < nshot; s0 += chunk) {
        const int ns = std::min(chunk, nshot - s0);
        if (adj) {
            SWEM_HIP(hipMemcpyAsync(drx.p, ux + s0 * nmt, ns * nmt * sizeof(float), hipMemcpyHostToDevice, st));
            SWEM_HIP(hipMemcpyAsync(drz.p, uz + s0 * nmt, ns * nmt * sizeof(float), hipMemcpyHostToDevice, st));
            SWEM_TRY(launch_sgemm(ns * nmx, 2 * nf, nt, drx.p, nt, dbadj.p, 2 * nf, reinterpret_cast<float*>(spx.p), 2 * nf,
                                  st));
            SWEM_TRY(launch_sgemm(ns * nmx, 2 * nf, nt, drz.p, nt, dbadj.p, 2 * nf, reinterpret_cast<float*>(spz.p), 2 * nf,
                                  st));
        }
        for (int c = 0; c < pp->nchunks; c++) {
            SWEM_TRY(ensure_chunk(pp, c, st));
            SWEM_TRY(ensure_chunk(ps, c, st));
            March2DArgs ap = base_args(pp, c, 1), as = base_args(ps, c, 1);
            ap.nshot = as.nshot = ns;
            ap.igx = as.igx = dig.p + s0;
            EwemArgs e{};
            e.nf = nf;
            e.dwf = pp->s.dw;
            e.dkx = pp->s.dkx;
            e.inv_maxs = 1.f / maxs;
            e.vp0 = dvp0.p;
            e.vs0 = dvs0.p;
            e.spx = spx.p;
            e.spz = spz.p;
            e.ppart = ppart.p;
            e.spart = spart.p;
            e.smig = smig.p;
            e.mppT = mppT.p;
            e.mpsT = mpsT.p;
            e.dgx = dgx.p;
            e.dgz = dgz.p;
            SWEM_TRY(launch_ewem(ap, as, e, adj, st));
            if (adj) {
                SWEM_TRY(launch_reduce_image(ppart.p, ns, ap.ngroups, nz, nmx, dmpp.p, !first_image, 0, st));
                SWEM_TRY(launch_reduce_image(spart.p, ns, ap.ngroups, nz, nmx, dmps.p, !first_image, 0, st));
                first_image = false;
            }
        }
        if (!adj) {
            SWEM_TRY(launch_sgemm(ns * nmx, nt, 2 * nf, reinterpret_cast<const float*>(dgx.p), 2 * nf, dbfwd.p, nt, drx.p,
                                  nt, st));
            SWEM_TRY(launch_sgemm(ns * nmx, nt, 2 * nf, reinterpret_cast<const float*>(dgz.p), 2 * nf, dbfwd.p, nt, drz.p,
                                  nt, st));
            SWEM_HIP(hipMemcpyAsync(ux + s0 * nmt, drx.p, ns * nmt * sizeof(float), hipMemcpyDeviceToHost, st));
            SWEM_HIP(hipMemcpyAsync(uz + s0 * nmt, drz.p, ns * nmt * sizeof(float), hipMemcpyDeviceToHost, st));
            SWEM_HIP(hipStreamSynchronize(st));  // the record buffers are reused by the next chunk
        }
    }
    if (adj) {
        SWEM_HIP(hipMemcpyAsync(mpp, dmpp.p, nmz * sizeof(float), hipMemcpyDeviceToHost, st));
        SWEM_HIP(hipMemcpyAsync(mps2, dmps.p, nmz * sizeof(float), hipMemcpyDeviceToHost, st));
    }
    SWEM_HIP(hipStreamSynchronize(st));
    return SWEM_OK;
}

// 3-D elastic shots (ewem.c:3-366 with nmy > 1): per shot and frequency chunk, the P source and
// the receiver pair marched through the two-pass depth step of the 3-D kernels (launch_ewem3d):
// migration per depth: source kz x-pass, separation, truncation, kz x-pass of the potentials (vp,
// vs), imaging y-pass, combination, truncation, and the depth's two images summed over the chunk's
// frequency groups; modelling: the source over all depths first (stored), then the receiver upward:
// separation, injection y-pass, kz x-pass, truncation, combination, truncation (or the output).
static int eshots3d_impl(swem_plan* pp, swem_plan* ps, int nshot, const float* sx, const float* sy, float* ux,
                         float* uy, float* uz, float* mpp, float* mps1, float* mps2, const float* wav, int adj)
{
    hipStream_t st = nullptr;
    const swem_geometry& g = pp->g;
    const int nmx = g.nmx, nmy = g.nmy, nm = nmx * nmy, nz = g.nz, nt = g.nt, nf = pp->nf;
    const size_t nmz = (size_t)nm * nz, nmt = (size_t)nm * nt;
    if (adj) {
        std::fill(mpp, mpp + nmz, 0.f);
        std::fill(mps1, mps1 + nmz, 0.f);
        std::fill(mps2, mps2 + nmz, 0.f);
    } else {
        std::fill(ux, ux + (size_t)nshot * nmt, 0.f);
        std::fill(uy, uy + (size_t)nshot * nmt, 0.f);
        std::fill(uz, uz + (size_t)nshot * nmt, 0.f);
    }
    std::vector<int> isrc(nshot);
    for (int i = 0; i < nshot; i++) {
        const float fx = std::trunc((sx[i] - g.omx) / g.dmx);  // ewem.c:99-100
        const float fy = std::trunc(((sy ? sy[i] : 0.f) - g.omy) / g.dmy);
        if (!(fx >= 0.f && fx < (float)nmx && fy >= 0.f && fy < (float)nmy))
            return set_error(SWEM_EINVAL, "ewem: source (%g, %g) outside the grid", (double)sx[i], sy ? (double)sy[i] : 0.0);
        isrc[i] = (int)fx * nmy + (int)fy;
    }
    if (nf < 1) return SWEM_OK;
    EwemTables tb;
    SWEM_TRY(ewem_tables(pp, ps, wav, tb));
    const std::vector<float>&badj = tb.badj, &bfwd = tb.bfwd, &vp0 = tb.vp0, &vs0 = tb.vs0;
    const float maxs = tb.maxs;
    // frequencies per launch: three slots of step state per frequency (plus the stored source when
    // modelling) within the work budget, and the step state within the last-level cache bound
    const int F = std::min(8, pp->F3);
    const size_t slot = (size_t)nmx * pp->s.nky * sizeof(float2);
    const size_t per_item = 3 * slot + (adj ? 2 * (size_t)nm * sizeof(float) / F + 1 : nmz * sizeof(float2));
    size_t nfc = std::max<size_t>(1, pp->budget / per_item);
    nfc = std::min(nfc, std::max<size_t>(1, pp->state3d_bytes / (3 * slot)));
    if (nfc < (size_t)nf) nfc = std::max<size_t>(F, nfc / F * F);
    nfc = std::min<size_t>(nfc, nf);
    const size_t gmax = (nfc + F - 1) / F;
    DevBuf<float> dbadj, dbfwd, dvp0, dvs0, drx, drz, ppart, spart, dmpp, dmps, mppT, mpsT;
    DevBuf<float2> W, spx, spz, smig, dgx, dgz;
    SWEM_TRY(dvp0.upload(vp0.data(), nz));
    SWEM_TRY(dvs0.upload(vs0.data(), nz));
    SWEM_TRY(W.alloc(nfc * 3 * (size_t)nmx * pp->s.nky));
    SWEM_TRY(drx.alloc(nmt));
    SWEM_TRY(drz.alloc(nmt));
    if (adj) {
        SWEM_TRY(dbadj.upload(badj.data(), badj.size()));
        SWEM_TRY(spx.alloc((size_t)nm * nf));
        SWEM_TRY(spz.alloc((size_t)nm * nf));
        SWEM_TRY(ppart.alloc(gmax * nm));
        SWEM_TRY(spart.alloc(gmax * nm));
        SWEM_TRY(dmpp.alloc(nmz));
        SWEM_TRY(dmps.alloc(nmz));
    } else {
        SWEM_TRY(dbfwd.upload(bfwd.data(), bfwd.size()));
        std::vector<float> t1(nmz), t2(nmz);
        for (int ix = 0; ix < nm; ix++)
            for (int iz = 0; iz < nz; iz++) {
                t1[(size_t)iz * nm + ix] = mpp[(size_t)ix * nz + iz];
                t2[(size_t)iz * nm + ix] = mps2[(size_t)ix * nz + iz];
            }
        SWEM_TRY(mppT.upload(t1.data(), nmz));
        SWEM_TRY(mpsT.upload(t2.data(), nmz));
        SWEM_TRY(smig.alloc(nfc * nmz));
        SWEM_TRY(dgx.alloc((size_t)nm * nf));
        SWEM_TRY(dgz.alloc((size_t)nm * nf));
    }
    bool first_image = true;
    for (int s = 0; s < nshot; s++) {
        if (adj) {
            SWEM_HIP(hipMemcpyAsync(drx.p, ux + (size_t)s * nmt, nmt * sizeof(float), hipMemcpyHostToDevice, st));
            SWEM_HIP(hipMemcpyAsync(drz.p, uz + (size_t)s * nmt, nmt * sizeof(float), hipMemcpyHostToDevice, st));
            SWEM_TRY(launch_sgemm(nm, 2 * nf, nt, drx.p, nt, dbadj.p, 2 * nf, reinterpret_cast<float*>(spx.p), 2 * nf, st));
            SWEM_TRY(launch_sgemm(nm, 2 * nf, nt, drz.p, nt, dbadj.p, 2 * nf, reinterpret_cast<float*>(spz.p), 2 * nf, st));
        }
        for (int f0 = 0; f0 < nf; f0 += (int)nfc) {
            March3DArgs ap = base3(pp), as = base3(ps);
            ap.f0 = as.f0 = f0;
            ap.nitem = as.nitem = std::min((int)nfc, nf - f0);
            Ewem3DArgs e{};
            e.W = W.p;
            e.igx = isrc[s];
            e.sep_adj = adj;
            e.dwf = pp->s.dw;
            e.dkxf = pp->s.dkx;
            e.inv_maxs = 1.f / maxs;
            e.vp0 = dvp0.p;
            e.vs0 = dvs0.p;
            e.spx = spx.p;
            e.spz = spz.p;
            e.ppart = ppart.p;
            e.spart = spart.p;
            e.smig = smig.p;
            e.mppT = mppT.p;
            e.mpsT = mpsT.p;
            e.dgx = dgx.p;
            e.dgz = dgz.p;
            SWEM_TRY(launch_ewem3d(0, ap, as, e, F, 0, st));
            auto at = [&](int iz) {
                ap.iz = as.iz = iz;
                e.ds = stepped_at(pp, iz, g.sz);
                e.dr = stepped_at(pp, iz, g.gz);
                e.next_s = stepped_at(pp, iz + 1, g.sz);
            };
            auto xpass = [&](int mode) {
                e.mode = mode;
                return launch_ewem3d(1, ap, as, e, F, 0, st);
            };
            const int ngr = (ap.nitem + F - 1) / F;
            // The truncation to y < nmy commutes with the separation and the combination (maps of
            // kx alone), so the combination of one depth and the separation of the next share one
            // x-pass and one truncation; an x-pass on a state another x-pass wrote scales by 1 / nkx.
            const float inv_nk = ap.inv_nk, inv_nkx = (float)(1.0 / (double)ap.nkx);
            if (adj) {
                e.conj_src = 1;
                e.conj_rec = 0;
                bool comb_due = false;  // the receiver pair of the previous depth awaits its combination
                for (int iz = 0; iz < nz; iz++) {
                    at(iz);
                    if (e.ds) SWEM_TRY(xpass(3));
                    if (e.dr) {
                        e.s1 = inv_nk;
                        e.s2 = inv_nkx;
                        e.izc = iz - 1;
                        SWEM_TRY(xpass(comb_due ? 4 : 0));
                        SWEM_TRY(launch_ewem3d(2, ap, as, e, F, 1, st));
                        SWEM_TRY(xpass(2));
                    }
                    SWEM_TRY(launch_ewem3d(2, ap, as, e, F, 0, st));
                    comb_due = e.dr;  // the last depth's combination is never used
                    const int acc = !first_image || f0 > 0;
                    SWEM_TRY(launch_reduce_depth(ppart.p, ngr, nm, dmpp.p + iz, nz, acc, st));
                    SWEM_TRY(launch_reduce_depth(spart.p, ngr, nm, dmps.p + iz, nz, acc, st));
                }
            } else {
                e.conj_src = 1;
                e.conj_rec = 1;
                for (int iz = 0; iz < nz; iz++) {  // the source, stored per depth
                    at(iz);
                    if (e.ds) SWEM_TRY(xpass(3));
                    SWEM_TRY(launch_ewem3d(2, ap, as, e, F, 2, st));
                }
                int comb_at = -1;  // depth whose combination is due
                for (int iz = nz - 1; iz >= 0; iz--) {
                    at(iz);
                    if (!e.dr) continue;
                    if (comb_at >= 0) {
                        e.izc = comb_at;
                        e.s1 = e.s2 = inv_nkx;
                        SWEM_TRY(xpass(4));
                    } else {
                        e.s1 = inv_nk;  // the initial (zero) state
                        SWEM_TRY(xpass(0));
                    }
                    SWEM_TRY(launch_ewem3d(2, ap, as, e, F, 3, st));
                    SWEM_TRY(xpass(2));
                    comb_at = iz;
                }
                if (comb_at >= 0) {  // the uppermost combination
                    at(comb_at);
                    e.s1 = inv_nkx;
                    SWEM_TRY(xpass(1));
                }
                // the modelled spectra: the last combination's state (or the zero initial state)
                SWEM_TRY(launch_ewem3d(2, ap, as, e, F, 4, st));
            }
        }
        if (adj) first_image = false;
        if (!adj) {
            SWEM_TRY(launch_sgemm(nm, nt, 2 * nf, reinterpret_cast<const float*>(dgx.p), 2 * nf, dbfwd.p, nt, drx.p, nt, st));
            SWEM_TRY(launch_sgemm(nm, nt, 2 * nf, reinterpret_cast<const float*>(dgz.p), 2 * nf, dbfwd.p, nt, drz.p, nt, st));
            SWEM_HIP(hipMemcpyAsync(ux + (size_t)s * nmt, drx.p, nmt * sizeof(float), hipMemcpyDeviceToHost, st));
            SWEM_HIP(hipMemcpyAsync(uz + (size_t)s * nmt, drz.p, nmt * sizeof(float), hipMemcpyDeviceToHost, st));
        }
        SWEM_HIP(hipStreamSynchronize(st));  // the record buffers are reused by the next shot
    }
    if (adj) {
        SWEM_HIP(hipMemcpyAsync(mpp, dmpp.p, nmz * sizeof(float), hipMemcpyDeviceToHost, st));
        SWEM_HIP(hipMemcpyAsync(mps2, dmps.p, nmz * sizeof(float), hipMemcpyDeviceToHost, st));
    }
    SWEM_HIP(hipStreamSynchronize(st));
    return SWEM_OK;
}

int swem_eshots(const swem_geometry* g, const float* vp, const float* vs, const float* wav, const float* sx,
                const float* sy, int nshot, float* ux, float* uy, float* uz, float* mpp, float* mps1, float* mps2,
                int adj, const swem_options* opt)
{
    if (!g || !vp || !vs || !wav || !sx || !ux || !uy || !uz || !mpp || !mps1 || !mps2 || nshot < 1)
        return set_error(SWEM_EINVAL, "ewem: null argument or nshot < 1");
    swem_geometry ge = *g;
    ge.pspi = 0;
    int ntfft, ifmin, ifmax;
    ewem_sampling(ge, ntfft, ifmin, ifmax);
    const bool three_d = ge.nmy > 1;
    if (!(ge.dmx > 0.f)) return set_error(SWEM_EINVAL, "ewem: dmx must be positive");
    std::vector<int> igx(nshot);
    for (int i = 0; i < nshot && !three_d; i++) {
        const float fx = std::trunc((sx[i] - ge.omx) / ge.dmx);  // ewem.c:99
        if (!(fx >= 0.f && fx < (float)ge.nmx)) return set_error(SWEM_EINVAL, "ewem: source x %g outside the grid", (double)sx[i]);
        igx[i] = (int)fx;
    }
    swem_options o;
    if (opt) o = *opt;
    else swem_options_init(&o);
    swem_plan *pp = nullptr, *ps = nullptr;
    // both plans sample the band as ewem.c does (ewem.c:43-52)
    SWEM_TRY(plan_create_rule(&ge, vp, wav, &o, 1, &pp));
    int rc = plan_create_rule(&ge, vs, wav, &o, 1, &ps);
    if (!rc && (pp->s.ntfft != ntfft || pp->s.ifmin != ifmin || pp->s.ifmax != ifmax))
        rc = set_error(SWEM_EDEVICE, "ewem: internal band mismatch (ntfft %d, [%d, %d)) against (%d, [%d, %d))", ntfft,
                       ifmin, ifmax, pp->s.ntfft, pp->s.ifmin, pp->s.ifmax);
    if (!rc)
        rc = three_d ? eshots3d_impl(pp, ps, nshot, sx, sy, ux, uy, uz, mpp, mps1, mps2, wav, adj)
                     : eshots_impl(pp, ps, nshot, igx.data(), ux, uy, uz, mpp, mps1, mps2, wav, adj);
    if (ps) swem_plan_destroy(ps);
    swem_plan_destroy(pp);
    return rc;
}

int swem_ewem(float* ux, float* uy, float* uz, float* mpp, float* mps1, float* mps2, const float* wav, int nt,
              float ot, float dt, int nmx, float omx, float dmx, int nmy, float omy, float dmy, float sx, float sy,
              int nz, float oz, float dz, float gz, float sz, const float* vp, const float* vs, float fmin, float fmax,
              int padt, int padx, int adj, int verbose)
{
    if (!ux || !uy || !uz || !mpp || !mps1 || !mps2 || !wav || !vp || !vs)
        return set_error(SWEM_EINVAL, "ewem: null argument");
    swem_geometry g;
    geom_from_args(&g, nt, ot, dt, nmx, omx, dmx, nmy, omy, dmy, nz, oz, dz, gz, sz, 1, fmin, fmax, padt, padx, 0);
    swem_options o;
    swem_options_init(&o);
    o.verbose = verbose;
    return swem_eshots(&g, vp, vs, wav, &sx, &sy, 1, ux, uy, uz, mpp, mps1, mps2, adj, &o);
}

int swem_malloc(void** ptr, size_t bytes)
{
    if (!ptr) return set_error(SWEM_EINVAL, "null pointer");
    SWEM_HIP(hipMalloc(ptr, bytes));
    return SWEM_OK;
}

int swem_free(void* ptr)
{
    if (ptr) SWEM_HIP(hipFree(ptr));
    return SWEM_OK;
}

int swem_memcpy_htod(void* dst, const void* src, size_t bytes)
{
    SWEM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SWEM_OK;
}

int swem_memcpy_dtoh(void* dst, const void* src, size_t bytes)
{
    SWEM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return SWEM_OK;
}

int swem_memcpy_dtod(void* dst, const void* src, size_t bytes)
{
    SWEM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return SWEM_OK;
}

int swem_vec_cgstep(float a, float* m1, float b, const float* m2, size_t n, void* stream)
{
    if ((!m1 || !m2) && n) return set_error(SWEM_EINVAL, "vec_cgstep: null vector");
    return launch_vec_cgstep(a, m1, b, m2, n, (hipStream_t)stream);
}

int swem_vec_mul(const float* x, const float* w, float* y, size_t n, void* stream)
{
    if ((!x || !w || !y) && n) return set_error(SWEM_EINVAL, "vec_mul: null vector");
    return launch_vec_mul(x, w, y, n, (hipStream_t)stream);
}

int swem_vec_dot(const float* x, const float* y, size_t n, double* out, void* stream)
{
    if (!out || ((!x || !y) && n)) return set_error(SWEM_EINVAL, "vec_dot: null argument");
    // scratch and pinned result slot per device, shared by the calls of the process (serialised)
    static std::mutex mu;
    static double* scratch[64] = {};
    static double* host[64] = {};
    std::lock_guard<std::mutex> lk(mu);
    int dev = 0;
    SWEM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return set_error(SWEM_EDEVICE, "vec_dot: device %d", dev);
    if (!scratch[dev]) {
        SWEM_HIP(hipMalloc((void**)&scratch[dev], sizeof(double) * vec_dot_scratch_doubles()));
        SWEM_HIP(hipHostMalloc((void**)&host[dev], sizeof(double), hipHostMallocDefault));
    }
    hipStream_t st = (hipStream_t)stream;
    SWEM_TRY(launch_vec_dot(x, y, n, scratch[dev], st));
    SWEM_HIP(hipMemcpyAsync(host[dev], scratch[dev] + vec_dot_scratch_doubles() - 1, sizeof(double),
                            hipMemcpyDeviceToHost, st));
    SWEM_HIP(hipStreamSynchronize(st));
    *out = host[dev][0];
    return SWEM_OK;
}

int swem_memset(void* ptr, int value, size_t bytes)
{
    SWEM_HIP(hipMemset(ptr, value, bytes));
    return SWEM_OK;
}

int swem_synchronize(void)
{
    SWEM_HIP(hipDeviceSynchronize());
    return SWEM_OK;
}

int swem_sampling(const swem_geometry* g, int* out)
{
    if (!g || !out) return set_error(SWEM_EINVAL, "sampling: null argument");
    Sampling s;
    SWEM_TRY(compute_sampling(*g, s));
    out[0] = s.ntfft; out[1] = s.nw; out[2] = s.nkx; out[3] = s.nky;
    out[4] = s.ifmin; out[5] = s.ifmax; out[6] = s.lmx; out[7] = s.lmy;
    return SWEM_OK;
}

int swem_source_index(const swem_geometry* g, float sx, float sy, int* igx, int* igy)
{
    if (!g || !igx || !igy) return set_error(SWEM_EINVAL, "source_index: null argument");
    return source_index(*g, sx, sy, *igx, *igy);
}

int swem_pspi_tables(const swem_geometry* g, const float* vel, float* vref, int* iref1, int* iref2, unsigned char* cut)
{
    if (!g || !vel || !vref) return set_error(SWEM_EINVAL, "pspi_tables: null argument");
    if (g->nmx < 1 || g->nmy < 1 || g->nz < 1) return set_error(SWEM_EINVAL, "pspi_tables: bad sizes");
    return pspi_tables(*g, vel, vref, iref1, iref2, cut);
}

int swem_taper(int n, int l, float* t)
{
    if (!t || n < 0 || l < 0) return set_error(SWEM_EINVAL, "taper: bad argument");
    taper_f32(n, l, t);
    return SWEM_OK;
}

}  // extern "C"
