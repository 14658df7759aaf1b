// swem_seis.cpp -- the .seis file format, the key=value argv protocol and swem_shot_main, the
// drop-in for shotwem.so's main (shotwem.c:12-140).
//
// Format (seismic.c:270-461, seismic.h:20-86, Header.jl:1-33): a text header of "\tkey=value"
// lines; a data file of native floats, n1 fastest, named by in="..."; a header file of
// 31 x 4-byte words per trace named by headers="...".  SeisWrite names the files
// <DATAPATH><name>@data@ and <DATAPATH><name>@headers@.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <algorithm>
#include <vector>

#include "swem_internal.hpp"

namespace swem {

// struct SeisHeader (seismic.h:20-53): 31 four-byte fields.
struct SeisHeader {
    int tracenum;
    float o1;
    int n1;
    float d1;
    float sx, sy, gx, gy, mx, my, hx, hy, h, az, ang;
    int isx, isy, igx, igy, imx, imy, ihx, ihy, ih, iaz, iang;
    float selev, gelev, sstat, gstat;
    int trid;
};
static_assert(sizeof(SeisHeader) == 4 * SWEM_SEIS_HEADER_WORDS, "SeisHeader must be 124 bytes");

// par_read_* (seismic.c:2-71): the name is the text before the first '=', compared exactly;
// the first match wins; a bool is true iff its value starts with 'y'.
static const char* par_value(int argc, char** argv, const char* key)
{
    for (int i = 1; i < argc; i++) {
        const char* a = argv[i];
        if (!a) continue;
        const char* eq = std::strchr(a, '=');
        size_t n = eq ? (size_t)(eq - a) : std::strlen(a);
        if (n == std::strlen(key) && std::strncmp(a, key, n) == 0) return eq ? eq + 1 : a + n;
    }
    return nullptr;
}
static bool par_string(int argc, char** argv, const char* key, std::string& v)
{
    const char* s = par_value(argc, argv, key);
    if (!s) return false;
    v = s;
    return true;
}
static bool par_int(int argc, char** argv, const char* key, int& v)
{
    const char* s = par_value(argc, argv, key);
    if (!s) return false;
    v = std::atoi(s);
    return true;
}
static bool par_float(int argc, char** argv, const char* key, float& v)
{
    const char* s = par_value(argc, argv, key);
    if (!s) return false;
    v = (float)std::atof(s);
    return true;
}
static bool par_bool(int argc, char** argv, const char* key, bool& v)
{
    const char* s = par_value(argc, argv, key);
    if (!s) return false;
    v = s[0] == 'y';
    return true;
}

static void copy_str(char* dst, const std::string& s)
{
    std::strncpy(dst, s.c_str(), 99);
    dst[99] = '\0';
}

}  // namespace swem

using namespace swem;

extern "C" {

// InitFileHeader (seismic.c:346-378).
void swem_seis_extent_init(swem_seis_extent* fh)
{
    std::memset(fh, 0, sizeof(*fh));
    fh->n1 = fh->n2 = fh->n3 = fh->n4 = fh->n5 = 1;
    fh->o1 = 0;
    fh->o2 = fh->o3 = fh->o4 = fh->o5 = 1;
    fh->d1 = 0;
    fh->d2 = fh->d3 = fh->d4 = fh->d5 = 1;
    std::strcpy(fh->data_format, "native_float");
    fh->esize = 4;
    std::strcpy(fh->dname, "NULL");
    std::strcpy(fh->hname, "NULL");
}

// ReadFileHeader (seismic.c:380-421): the last occurrence of each key wins.
int swem_seis_read_header(const char* name, swem_seis_extent* fh)
{
    if (!name || !fh) return set_error(SWEM_EINVAL, "seis_read_header: null argument");
    FILE* fp = std::fopen(name, "r");
    if (!fp) return set_error(SWEM_EIO, "cannot open seis header %s: %s", name, std::strerror(errno));
    char line[4096];
    while (std::fgets(line, sizeof line, fp)) {
        char* s = line;
        while (*s == '\t' || *s == ' ') s++;
        char* eq = std::strchr(s, '=');
        if (!eq) continue;
        std::string key(s, eq - s);
        std::string val(eq + 1);
        while (!val.empty() && (val.back() == '\n' || val.back() == '\r')) val.pop_back();
        std::string sval = val;
        if (!sval.empty() && sval[0] == '"') {
            size_t e = sval.find('"', 1);
            sval = sval.substr(1, e == std::string::npos ? std::string::npos : e - 1);
        }
        const char* v = val.c_str();
        if (key == "in") copy_str(fh->dname, sval);
        else if (key == "headers") copy_str(fh->hname, sval);
        else if (key == "title") copy_str(fh->title, sval);
        else if (key == "data_format") copy_str(fh->data_format, sval);
        else if (key == "esize") fh->esize = std::atoi(v);
        else if (key == "n1") fh->n1 = std::atoi(v);
        else if (key == "n2") fh->n2 = std::atoi(v);
        else if (key == "n3") fh->n3 = std::atoi(v);
        else if (key == "n4") fh->n4 = std::atoi(v);
        else if (key == "n5") fh->n5 = std::atoi(v);
        else if (key == "o1") fh->o1 = (float)std::atof(v);
        else if (key == "o2") fh->o2 = (float)std::atof(v);
        else if (key == "o3") fh->o3 = (float)std::atof(v);
        else if (key == "o4") fh->o4 = (float)std::atof(v);
        else if (key == "o5") fh->o5 = (float)std::atof(v);
        else if (key == "d1") fh->d1 = (float)std::atof(v);
        else if (key == "d2") fh->d2 = (float)std::atof(v);
        else if (key == "d3") fh->d3 = (float)std::atof(v);
        else if (key == "d4") fh->d4 = (float)std::atof(v);
        else if (key == "d5") fh->d5 = (float)std::atof(v);
        else if (key == "label1") copy_str(fh->label1, sval);
        else if (key == "label2") copy_str(fh->label2, sval);
        else if (key == "label3") copy_str(fh->label3, sval);
        else if (key == "label4") copy_str(fh->label4, sval);
        else if (key == "label5") copy_str(fh->label5, sval);
        else if (key == "unit1") copy_str(fh->unit1, sval);
        else if (key == "unit2") copy_str(fh->unit2, sval);
        else if (key == "unit3") copy_str(fh->unit3, sval);
        else if (key == "unit4") copy_str(fh->unit4, sval);
        else if (key == "unit5") copy_str(fh->unit5, sval);
    }
    std::fclose(fp);
    // Every extent must be at least 1 (callers size buffers and index the first and last trace
    // from them), and a file of more than 2^36 samples is taken as a corrupt header.
    const int ns[5] = {fh->n1, fh->n2, fh->n3, fh->n4, fh->n5};
    double total = 1.0;
    for (int i = 0; i < 5; i++) {
        if (ns[i] < 1) return set_error(SWEM_EINVAL, "seis header %s: n%d=%d (extents must be >= 1)", name, i + 1, ns[i]);
        total *= ns[i];
    }
    if (total > 68719476736.0)
        return set_error(SWEM_EINVAL, "seis header %s: %g samples (n1..n5 product above 2^36)", name, total);
    return SWEM_OK;
}

// SeisRead (seismic.c:270-297).
int swem_seis_read(const char* name, float* data, void* headers, swem_seis_extent* fh)
{
    if (!name || !fh) return set_error(SWEM_EINVAL, "seis_read: null argument");
    swem_seis_extent_init(fh);
    SWEM_TRY(swem_seis_read_header(name, fh));
    const size_t ntr = (size_t)fh->n2 * fh->n3 * fh->n4 * fh->n5;
    if (headers) {
        FILE* fp = std::fopen(fh->hname, "rb");
        if (!fp) return set_error(SWEM_EIO, "SeisRead: Unable to open %s", fh->hname);
        size_t got = std::fread(headers, 4 * SWEM_SEIS_HEADER_WORDS, ntr, fp);
        std::fclose(fp);
        if (got != ntr) return set_error(SWEM_EIO, "SeisRead: %s holds %zu of %zu trace headers", fh->hname, got, ntr);
    }
    if (data) {
        FILE* fp = std::fopen(fh->dname, "rb");
        if (!fp) return set_error(SWEM_EIO, "SeisRead: Unable to open %s", fh->dname);
        size_t got = std::fread(data, sizeof(float), ntr * fh->n1, fp);
        std::fclose(fp);
        if (got != ntr * fh->n1) return set_error(SWEM_EIO, "SeisRead: %s holds %zu of %zu samples", fh->dname, got, ntr * fh->n1);
    }
    return SWEM_OK;
}

// SeisWrite + WriteFileHeader (seismic.c:300-344, 423-461), including the repeated unit5 line.
int swem_seis_write(const char* name, const float* data, const void* headers, swem_seis_extent* fh)
{
    if (!name || !fh) return set_error(SWEM_EINVAL, "seis_write: null argument");
    const char* dp = std::getenv("DATAPATH");
    std::string base = dp ? std::string(dp) + name : std::string(name);
    std::string dn = base + "@data@", hn = base + "@headers@";
    // hn ("@headers@") is the longer of the two names; both fields hold 100 bytes
    if (dn.size() >= sizeof fh->dname || hn.size() >= sizeof fh->hname)
        return set_error(SWEM_EINVAL, "seis file name too long: %s", hn.c_str());
    std::strcpy(fh->dname, dn.c_str());
    std::strcpy(fh->hname, hn.c_str());
    FILE* fp = std::fopen(name, "w");
    if (!fp) return set_error(SWEM_EIO, "cannot write %s: %s", name, std::strerror(errno));
    std::fprintf(fp, "\tn1=%d\n\tn2=%d\n\tn3=%d\n\tn4=%d\n\tn5=%d\n", fh->n1, fh->n2, fh->n3, fh->n4, fh->n5);
    std::fprintf(fp, "\to1=%f\n\to2=%f\n\to3=%f\n\to4=%f\n\to5=%f\n", fh->o1, fh->o2, fh->o3, fh->o4, fh->o5);
    std::fprintf(fp, "\td1=%f\n\td2=%f\n\td3=%f\n\td4=%f\n\td5=%f\n", fh->d1, fh->d2, fh->d3, fh->d4, fh->d5);
    std::fprintf(fp, "\tlabel1=\"%s\"\n\tlabel2=\"%s\"\n\tlabel3=\"%s\"\n\tlabel4=\"%s\"\n\tlabel5=\"%s\"\n", fh->label1,
                 fh->label2, fh->label3, fh->label4, fh->label5);
    std::fprintf(fp, "\tunit1=\"%s\"\n\tunit2=\"%s\"\n\tunit3=\"%s\"\n\tunit4=\"%s\"\n\tunit5=\"%s\"\n\tunit5=\"%s\"\n",
                 fh->unit1, fh->unit2, fh->unit3, fh->unit4, fh->unit5, fh->unit5);
    std::fprintf(fp, "\ttitle=\"%s\"\n\tdata_format=\"%s\"\n\tesize=%d\n\tin=\"%s\"\n\theaders=\"%s\"\n", fh->title,
                 fh->data_format, fh->esize, fh->dname, fh->hname);
    std::fclose(fp);
    const size_t ntr = (size_t)fh->n2 * fh->n3 * fh->n4 * fh->n5;
    fp = std::fopen(fh->hname, "wb");
    if (!fp) return set_error(SWEM_EIO, "SeisWrite: Unable to open %s", fh->hname);
    if (headers) std::fwrite(headers, 4 * SWEM_SEIS_HEADER_WORDS, ntr, fp);
    std::fclose(fp);
    fp = std::fopen(fh->dname, "wb");
    if (!fp) return set_error(SWEM_EIO, "SeisWrite: Unable to open %s", fh->dname);
    if (data) std::fwrite(data, sizeof(float), ntr * fh->n1, fp);
    std::fclose(fp);
    return SWEM_OK;
}

// shotwem.c:12-140.
int swem_shot_main(int argc, char** argv)
{
    bool adj = true, pspi = true, verbose = false;
    std::string d_name, m_name, vel_name, wav_name;
    float sz = 0, gz = 0, sx = 0, sy = 0, fmin = 0, fmax = 80;
    int padt = 1, padx = 1, nref = 5;
    par_bool(argc, argv, "adj", adj);
    par_bool(argc, argv, "pspi", pspi);
    par_bool(argc, argv, "verbose", verbose);
    if (!par_string(argc, argv, "d", d_name) || !par_string(argc, argv, "m", m_name) ||
        !par_string(argc, argv, "vel", vel_name) || !par_string(argc, argv, "wav", wav_name)) {
        std::fprintf(stderr, "shotwem: one-shot wave-equation migration (adj=y) or modelling (adj=n) of .seis files;"
                             " required: d= m= vel= wav=\n");
        return set_error(SWEM_EINVAL, "shotwem: d=, m=, vel= and wav= are required");
    }
    par_float(argc, argv, "sz", sz);
    par_float(argc, argv, "gz", gz);
    par_float(argc, argv, "sx", sx);
    par_float(argc, argv, "sy", sy);
    par_float(argc, argv, "fmin", fmin);
    par_float(argc, argv, "fmax", fmax);
    par_int(argc, argv, "padt", padt);
    par_int(argc, argv, "padx", padx);
    par_int(argc, argv, "nref", nref);

    swem_seis_extent fh;
    swem_seis_extent_init(&fh);
    SWEM_TRY(swem_seis_read_header(wav_name.c_str(), &fh));
    const int nt = fh.n1;
    const size_t nwtr = (size_t)fh.n2 * fh.n3 * fh.n4 * fh.n5;
    std::vector<float> wav((size_t)nt * std::max<size_t>(nwtr, 1));
    std::vector<SeisHeader> hw(std::max<size_t>(nwtr, 1));
    SWEM_TRY(swem_seis_read(wav_name.c_str(), wav.data(), hw.data(), &fh));
    const float ot = fh.o1, dt = fh.d1;

    swem_seis_extent_init(&fh);
    SWEM_TRY(swem_seis_read_header(vel_name.c_str(), &fh));
    const int nz = fh.n1;
    const size_t ntraces = (size_t)fh.n2 * fh.n3 * fh.n4 * fh.n5;
    std::vector<float> vel((size_t)nz * ntraces);
    std::vector<SeisHeader> hv(ntraces);
    SWEM_TRY(swem_seis_read(vel_name.c_str(), vel.data(), hv.data(), &fh));
    const int nx = hv[ntraces - 1].imx - hv[0].imx + 1;
    const int ny = hv[ntraces - 1].imy - hv[0].imy + 1;
    const float oz = fh.o1, dz = fh.d1;
    const float ox = hv[0].mx;
    const float dx = ntraces > 1 ? hv[1].mx - hv[0].mx : fh.d2;
    const float oy = hv[0].my;
    const float dy = dx;  // shotwem.c:68
    if (nx < 1 || ny < 1 || (size_t)nx * ny != ntraces)
        return set_error(SWEM_EINVAL, "shotwem: velocity headers give nx=%d ny=%d for %zu traces", nx, ny, ntraces);
    const size_t nm = (size_t)nx * ny;

    std::vector<float> d(nm * nt, 0.f), m(nm * nz, 0.f);
    std::vector<SeisHeader> hd(nm), hm(nm);
    if (adj) {
        swem_seis_extent_init(&fh);
        SWEM_TRY(swem_seis_read_header(d_name.c_str(), &fh));
        if ((size_t)fh.n2 * fh.n3 * fh.n4 * fh.n5 != nm || fh.n1 != nt)
            return set_error(SWEM_EINVAL, "shotwem: data %s is %d x %d, expected %d x %zu", d_name.c_str(), fh.n1,
                             fh.n2 * fh.n3 * fh.n4 * fh.n5, nt, nm);
        SWEM_TRY(swem_seis_read(d_name.c_str(), d.data(), hd.data(), &fh));
        for (size_t ix = 0; ix < nm; ix++) hm[ix] = hv[ix];
    } else {
        swem_seis_extent_init(&fh);
        SWEM_TRY(swem_seis_read_header(m_name.c_str(), &fh));
        if ((size_t)fh.n2 * fh.n3 * fh.n4 * fh.n5 != nm || fh.n1 != nz)
            return set_error(SWEM_EINVAL, "shotwem: image %s is %d x %d, expected %d x %zu", m_name.c_str(), fh.n1,
                             fh.n2 * fh.n3 * fh.n4 * fh.n5, nz, nm);
        SWEM_TRY(swem_seis_read(m_name.c_str(), m.data(), hm.data(), &fh));
        for (size_t ix = 0; ix < nm; ix++) {
            hd[ix] = hm[ix];
            hd[ix].n1 = nt;
            hd[ix].d1 = dt;
            hd[ix].o1 = ot;
            hd[ix].gx = hm[ix].mx;
            hd[ix].gy = hm[ix].my;
            hd[ix].hx = hm[ix].mx - sx;
            hd[ix].hy = hm[ix].my - sy;
        }
    }
    swem_geometry g;
    std::memset(&g, 0, sizeof g);
    g.nt = nt; g.ot = ot; g.dt = dt;
    g.nmx = nx; g.omx = ox; g.dmx = dx;
    g.nmy = ny; g.omy = oy; g.dmy = dy;
    g.nz = nz; g.oz = oz; g.dz = dz;
    g.gz = gz; g.sz = sz;
    g.nref = nref; g.fmin = fmin; g.fmax = fmax;
    g.padt = padt; g.padx = padx; g.pspi = pspi ? 1 : 0;
    swem_options o;
    swem_options_init(&o);
    o.verbose = verbose ? 1 : 0;
    SWEM_TRY(swem_wem_ex(d.data(), m.data(), wav.data(), &g, sx, sy, vel.data(), adj ? 1 : 0, &o));
    swem_seis_extent_init(&fh);
    if (adj) {
        for (size_t ix = 0; ix < nm; ix++) {
            hm[ix].mx = hd[ix].gx;
            hm[ix].my = hd[ix].gy;
            hm[ix].imx = (int)std::trunc((hm[ix].mx - ox) / dx);
            hm[ix].imy = (int)std::trunc((hm[ix].my - oy) / dy);
        }
        fh.n1 = nz; fh.o1 = oz; fh.d1 = dz;
        fh.n2 = nx; fh.o2 = ox; fh.d2 = dx;
        fh.n3 = ny; fh.o3 = oy; fh.d3 = dy;
        SWEM_TRY(swem_seis_write(m_name.c_str(), m.data(), hm.data(), &fh));
    } else {
        fh.n1 = nt; fh.o1 = ot; fh.d1 = dt;
        fh.n2 = nx; fh.o2 = ox; fh.d2 = dx;
        fh.n3 = ny; fh.o3 = oy; fh.d3 = dy;
        SWEM_TRY(swem_seis_write(d_name.c_str(), d.data(), hd.data(), &fh));
    }
    return SWEM_OK;
}

}  // extern "C"

// Drop-in for compute_angles.so's main (compute_angles.c:12-118), the library ComputeAngles.jl
// calls through ccall((:main, "compute_angles")): the angles of one shot from the velocity and
// wavelet files, written as angx / angy with the velocity's headers and sx set.  Geometry comes
// from the trace headers as the reference reads it (nt = n1 of the wavelet file, ot / dt from its
// first trace header; nz = n1 and oz / dz from the first velocity trace header, nx / ny from the
// imx / imy range, ox = mx of the first trace, dx from the first two, dy as compute_angles.c:63).
// dipx / dipy are read only with dip_flag=y (the reference requires the names either way).
int swem_angles_main(int argc, char** argv)
{
    bool verbose = false, dip_flag = false;
    std::string angx_name, angy_name, vel_name, wav_name, dipx_name, dipy_name;
    float sz = 0, sx = 0, sy = 0, fmin = 0, fmax = 80;
    par_bool(argc, argv, "verbose", verbose);
    par_bool(argc, argv, "dip_flag", dip_flag);
    if (!par_string(argc, argv, "angx", angx_name) || !par_string(argc, argv, "angy", angy_name) ||
        !par_string(argc, argv, "vel", vel_name) || !par_string(argc, argv, "wav", wav_name) ||
        (dip_flag && (!par_string(argc, argv, "dipx", dipx_name) || !par_string(argc, argv, "dipy", dipy_name)))) {
        std::fprintf(stderr, "compute_angles: source incidence angles of one shot from .seis files;"
                             " required: angx= angy= vel= wav= (dipx= dipy= with dip_flag=y)\n");
        return set_error(SWEM_EINVAL, "compute_angles: angx=, angy=, vel= and wav= are required");
    }
    par_float(argc, argv, "sz", sz);
    par_float(argc, argv, "sx", sx);
    par_float(argc, argv, "sy", sy);
    par_float(argc, argv, "fmin", fmin);
    par_float(argc, argv, "fmax", fmax);

    swem_seis_extent fh;
    swem_seis_extent_init(&fh);
    SWEM_TRY(swem_seis_read_header(wav_name.c_str(), &fh));
    const int nt = fh.n1;
    const size_t nwtr = std::max<size_t>((size_t)fh.n2 * fh.n3 * fh.n4 * fh.n5, 1);
    std::vector<float> wav((size_t)nt * nwtr);
    std::vector<SeisHeader> hw(nwtr);
    SWEM_TRY(swem_seis_read(wav_name.c_str(), wav.data(), hw.data(), &fh));
    const float ot = hw[0].o1, dt = hw[0].d1;

    swem_seis_extent_init(&fh);
    SWEM_TRY(swem_seis_read_header(vel_name.c_str(), &fh));
    const int nz = fh.n1;
    const size_t ntraces = (size_t)fh.n2 * fh.n3 * fh.n4 * fh.n5;
    if (ntraces < 2) return set_error(SWEM_EINVAL, "compute_angles: velocity %s has %zu traces", vel_name.c_str(), ntraces);
    std::vector<float> vel((size_t)nz * ntraces);
    std::vector<SeisHeader> hv(ntraces);
    SWEM_TRY(swem_seis_read(vel_name.c_str(), vel.data(), hv.data(), &fh));
    const int nx = hv[ntraces - 1].imx - hv[0].imx + 1;
    const int ny = hv[ntraces - 1].imy - hv[0].imy + 1;
    if (nx < 1 || ny < 1 || (size_t)nx * ny != ntraces)
        return set_error(SWEM_EINVAL, "compute_angles: velocity headers give nx=%d ny=%d for %zu traces", nx, ny, ntraces);
    const float oz = hv[0].o1, dz = hv[0].d1;
    const float ox = hv[0].mx, dx = hv[1].mx - hv[0].mx;
    const float oy = hv[0].my;
    const float dy = ntraces > (size_t)nx ? hv[nx].my - hv[nx - 1].my : dx;
    const size_t nm = (size_t)nx * ny;
    std::vector<float> angx(nm * nz), angy(nm * nz);
    SWEM_TRY(swem_compute_angles(angx.data(), angy.data(), wav.data(), nt, ot, dt, nx, ox, dx, ny, oy, dy, sx, sy, nz,
                                 oz, dz, sz, vel.data(), fmin, fmax, verbose ? 1 : 0));
    if (dip_flag) {  // compute_angles.c:84-100: angles relative to the reflector normal
        for (int k = 0; k < 2; k++) {
            const std::string& name = k == 0 ? dipx_name : dipy_name;
            std::vector<float>& ang = k == 0 ? angx : angy;
            swem_seis_extent_init(&fh);
            SWEM_TRY(swem_seis_read_header(name.c_str(), &fh));
            if (fh.n1 != nz || (size_t)fh.n2 * fh.n3 * fh.n4 * fh.n5 != ntraces)
                return set_error(SWEM_EINVAL, "compute_angles: %s does not match the velocity grid", name.c_str());
            std::vector<float> dip((size_t)nz * ntraces);
            std::vector<SeisHeader> hd(ntraces);
            SWEM_TRY(swem_seis_read(name.c_str(), dip.data(), hd.data(), &fh));
            for (size_t i = 0; i < nm * nz; i++) ang[i] += dip[i];
        }
    }
    std::vector<SeisHeader> ha(hv.begin(), hv.begin() + nm);
    for (auto& h : ha) h.sx = sx;
    swem_seis_extent_init(&fh);
    fh.n1 = nz; fh.o1 = oz; fh.d1 = dz;
    fh.n2 = nx; fh.o2 = ox; fh.d2 = dx;
    fh.n3 = ny; fh.o3 = oy; fh.d3 = dy;
    SWEM_TRY(swem_seis_write(angx_name.c_str(), angx.data(), ha.data(), &fh));
    SWEM_TRY(swem_seis_write(angy_name.c_str(), angy.data(), ha.data(), &fh));
    return SWEM_OK;
}

// Drop-in for shotewem.so's main (shotewem.c:11-223), the library ShotProfileEWEM.jl calls: one
// shot of elastic migration (ux, uy, uz -> mpp, mps1, mps2) or modelling, .seis files in and out,
// with the reference's header rewrites (shotewem.c:117-151, 171-190).
int swem_eshot_main(int argc, char** argv)
{
    bool adj = true, verbose = false;
    std::string nm_[9];
    const char* keys[9] = {"ux", "uy", "uz", "mpp", "mps1", "mps2", "vp", "vs", "wav"};
    float sz = 0, gz = 0, sx = 0, sy = 0, fmin = 0, fmax = 80;
    int padt = 1, padx = 1;
    par_bool(argc, argv, "adj", adj);
    par_bool(argc, argv, "verbose", verbose);
    for (int k = 0; k < 9; k++)
        if (!par_string(argc, argv, keys[k], nm_[k])) {
            std::fprintf(stderr, "shotewem: one-shot elastic wave-equation migration (adj=y) or modelling (adj=n) of"
                                 " .seis files; required: ux= uy= uz= mpp= mps1= mps2= vp= vs= wav=\n");
            return set_error(SWEM_EINVAL, "shotewem: ux=, uy=, uz=, mpp=, mps1=, mps2=, vp=, vs= and wav= are required");
        }
    par_float(argc, argv, "sz", sz);
    par_float(argc, argv, "gz", gz);
    par_float(argc, argv, "sx", sx);
    par_float(argc, argv, "sy", sy);
    par_float(argc, argv, "fmin", fmin);
    par_float(argc, argv, "fmax", fmax);
    par_int(argc, argv, "padt", padt);
    par_int(argc, argv, "padx", padx);

    swem_seis_extent fh;
    swem_seis_extent_init(&fh);
    SWEM_TRY(swem_seis_read_header(nm_[8].c_str(), &fh));
    const int nt = fh.n1;
    const size_t nwtr = std::max<size_t>((size_t)fh.n2 * fh.n3 * fh.n4 * fh.n5, 1);
    std::vector<float> wav((size_t)nt * nwtr);
    std::vector<SeisHeader> hw(nwtr);
    SWEM_TRY(swem_seis_read(nm_[8].c_str(), wav.data(), hw.data(), &fh));
    const float ot = fh.o1, dt = fh.d1;
    swem_seis_extent_init(&fh);
    SWEM_TRY(swem_seis_read_header(nm_[6].c_str(), &fh));
    const int nz = fh.n1;
    const size_t ntraces = (size_t)fh.n2 * fh.n3 * fh.n4 * fh.n5;
    if (ntraces < 2) return set_error(SWEM_EINVAL, "shotewem: velocity %s has %zu traces", nm_[6].c_str(), ntraces);
    std::vector<float> vp((size_t)nz * ntraces), vs((size_t)nz * ntraces);
    std::vector<SeisHeader> hv(ntraces), hvs(ntraces);
    SWEM_TRY(swem_seis_read(nm_[6].c_str(), vp.data(), hv.data(), &fh));
    SWEM_TRY(swem_seis_read(nm_[7].c_str(), vs.data(), hvs.data(), &fh));
    const int nx = hv[ntraces - 1].imx - hv[0].imx + 1;
    const int ny = hv[ntraces - 1].imy - hv[0].imy + 1;
    if (nx < 1 || ny < 1 || (size_t)nx * ny != ntraces)
        return set_error(SWEM_EINVAL, "shotewem: velocity headers give nx=%d ny=%d for %zu traces", nx, ny, ntraces);
    const float oz = fh.o1, dz = fh.d1;
    const float ox = hv[0].mx, dx = hv[1].mx - hv[0].mx, oy = hv[0].my, dy = dx;
    const size_t nm = ntraces;
    std::vector<float> d[3], m[3];
    std::vector<SeisHeader> hd[3], hm[3];
    for (int k = 0; k < 3; k++) {
        d[k].assign(nm * nt, 0.f);
        m[k].assign(nm * nz, 0.f);
        hd[k].resize(nm);
        hm[k].resize(nm);
        const bool in_d = adj;
        const std::string& name = in_d ? nm_[k] : nm_[3 + k];
        const int n1 = in_d ? nt : nz;
        swem_seis_extent_init(&fh);
        SWEM_TRY(swem_seis_read_header(name.c_str(), &fh));
        if ((size_t)fh.n2 * fh.n3 * fh.n4 * fh.n5 != nm || fh.n1 != n1)
            return set_error(SWEM_EINVAL, "shotewem: %s is %d x %d, expected %d x %zu", name.c_str(), fh.n1,
                             fh.n2 * fh.n3 * fh.n4 * fh.n5, n1, nm);
        SWEM_TRY(swem_seis_read(name.c_str(), in_d ? d[k].data() : m[k].data(), in_d ? hd[k].data() : hm[k].data(), &fh));
    }
    SWEM_TRY(swem_ewem(d[0].data(), d[1].data(), d[2].data(), m[0].data(), m[1].data(), m[2].data(), wav.data(), nt, ot,
                       dt, nx, ox, dx, ny, oy, dy, sx, sy, nz, oz, dz, gz, sz, vp.data(), vs.data(), fmin, fmax, padt,
                       padx, adj ? 1 : 0, verbose ? 1 : 0));
    swem_seis_extent_init(&fh);
    if (adj) {
        fh.n1 = nz; fh.o1 = oz; fh.d1 = dz;
        fh.n2 = nx; fh.o2 = ox; fh.d2 = dx;
        fh.n3 = ny; fh.o3 = oy; fh.d3 = dy;
        for (int k = 0; k < 3; k++) {
            for (size_t ix = 0; ix < nm; ix++) {
                SeisHeader h = hv[ix];
                h.trid = k == 0 ? 1 : 2;
                h.mx = hd[0][ix].gx;
                h.my = hd[0][ix].gy;
                h.imx = (int)std::trunc((h.mx - ox) / dx);
                h.imy = (int)std::trunc((h.my - oy) / dy);
                hm[k][ix] = h;
            }
            SWEM_TRY(swem_seis_write(nm_[3 + k].c_str(), m[k].data(), hm[k].data(), &fh));
        }
    } else {
        fh.n1 = nt; fh.o1 = ot; fh.d1 = dt;
        fh.n2 = nx; fh.o2 = ox; fh.d2 = dx;
        fh.n3 = ny; fh.o3 = oy; fh.d3 = dy;
        for (int k = 0; k < 3; k++) {
            for (size_t ix = 0; ix < nm; ix++) {
                SeisHeader h = hm[0][ix];
                h.n1 = nt;
                h.d1 = dt;
                h.o1 = ot;
                h.gx = hm[0][ix].mx;
                h.gy = hm[0][ix].my;
                h.hx = hm[0][ix].mx - sx;
                h.hy = hm[0][ix].my - sy;
                hd[k][ix] = h;
            }
            SWEM_TRY(swem_seis_write(nm_[k].c_str(), d[k].data(), hd[k].data(), &fh));
        }
    }
    return SWEM_OK;
}
