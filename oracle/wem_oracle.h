/*
 * wem_oracle.h -- CPU oracle for shot-profile wave-equation migration (test infrastructure).
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker.  The product path is the
 * HIP library under seismic.jl_amd/csrc, which never links or calls anything here.
 *
 * It restates the reference's C path (src/Imaging/c/wem.c:3-564) in plain C, compiled twice:
 *   orc64_*  -- every continuous quantity in double precision: the FP64 oracle.
 *   orc32_*  -- every continuous quantity in float, with the reference's float expressions
 *               (s = w*w*po*po - kx*kx - ky*ky in float, sqrtf, expf...): "naive FP32".
 * Discrete decisions (ntfft, nw, ifmin/ifmax, igx/igy, iref1/iref2, taper widths, the PSPI
 * "vref2-vref1>10" test) are always evaluated with the reference's float/int casts, once, in
 * code shared by both variants, so both variants and the GPU library take identical branches.
 *
 * The reference cannot be compiled here (it needs FFTW3, which the image does not have), so
 * the FFT is this file's own mixed-radix transform with FFTW's documented sign conventions:
 * FORWARD = exp(-i...), BACKWARD = exp(+i...), both unnormalised.
 *
 * Parity of every function is pinned by tests/test_oracle_*.py against closed forms
 * (brute-force DFT, plane-wave phase shift, Gazdag v(z) phase shift, PSPI==SSF for v(z),
 * f_op round trip, taper values, discrete decisions at the configuration values, the
 * adjoint dot test, 1-D kinematics).
 */
#ifndef SEISWEM_ORACLE_H
#define SEISWEM_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Geometry and parameters of one wem() call (wem.c:3-11, shotwem.c:27-42). */
typedef struct orc_geom {
    int nt;   float ot, dt;
    int nmx;  float omx, dmx;
    int nmy;  float omy, dmy;
    int nz;   float oz, dz;
    float gz, sz;
    int nref; float fmin, fmax;
    int padt, padx;
    int pspi;       /* 0 = split-step Fourier (ssop), 1 = PSPI (pspiop) */
    int pspi_mode;  /* 0 = compat (reference: real-part cut, select after propagation, both
                       directions); 1 = linear interpolation, forward = exact adjoint */
} orc_geom;

/* Derived sampling (wem.c:79-89, ssop wem.c:305-308). */
typedef struct orc_sampling {
    int ntfft, nw, nkx, nky, ifmin, ifmax, lmx, lmy;
    int fop_n;          /* f_op transform length (nw-1)*2, wem.c:487 */
    float dkx, dky, dw; /* float values exactly as the reference computes them */
} orc_sampling;

/* Error codes (0 = ok). */
#define ORC_OK 0
#define ORC_EINVAL 1
#define ORC_ENOMEM 2

int orc_sampling_compute(const orc_geom* g, orc_sampling* s);
/* igx = (int)(sx-omx)/dmx (wem.c:112-113). Returns ORC_EINVAL when outside the grid. */
int orc_source_index(const orc_geom* g, float sx, float sy, int* igx, int* igy);
/* PSPI tables with the reference's float arithmetic (wem.c:40-70, 450-459).
 * vref [nref][nz] float, iref1/iref2 [nm][nz], cut [nm][nz] = (vref2 - vref1 > 10.0). */
int orc_pspi_tables(const orc_geom* g, const float* vel, float* vref, int* iref1, int* iref2,
                    unsigned char* cut);
/* Slowness split with the reference's float arithmetic (wem.c:27-36): po [nz], pd [nm][nz]. */
int orc_slowness_f32(const orc_geom* g, const float* vel, float* po, float* pd);
/* Edge taper of boundary_condition (wem.c:546-564) for one axis, float expressions. */
void orc_taper_f32(int n, int l, float* t);
void orc_taper_f64(int n, int l, double* t);

/* Unnormalised in-place complex FFT, interleaved (re,im); sign -1 = FFTW_FORWARD. */
int orc64_fft(double* x, int n, int sign);
int orc32_fft(float* x, int n, int sign);

/* f_op (wem.c:480-514): adj=1 r2c of d[0..nt) zero-padded to (nw-1)*2 -> m[0..nw);
 * adj=0 c2r of m[0..nw) -> d[0..nt). */
int orc64_fop(double* m, double* d, int nw, int nt, int adj);
int orc32_fop(float* m, float* d, int nw, int nt, int adj);

/* One extrapolation step on a wavefield u [nm] (interleaved complex), exactly ssop/pspiop
 * (wem.c:289-471) at frequency index iw and depth index iz with signed step dzs.
 * adj selects the ssop branch (and, for pspi_mode 1, select-after (1) vs weight-before (0)). */
int orc64_step(double* u, const float* vel, const orc_geom* g, int iw, int iz, float dzs, int adj);
int orc32_step(float* u, const float* vel, const orc_geom* g, int iw, int iz, float dzs, int adj);

/* One shot: wem() (wem.c:3-204).  adj=1: d [nm][nt] in, m [nm][nz] out (overwritten).
 * adj=0: m in, d out (overwritten).  Returns an error code. */
int orc64_wem(double* d, double* m, const float* wav, const float* vel, const orc_geom* g,
              float sx, float sy, int adj);
int orc32_wem(float* d, float* m, const float* wav, const float* vel, const orc_geom* g,
              float sx, float sy, int adj);

/* Several shots over one velocity grid, as ShotProfileWEM stacks them with nangx=nangy=1
 * (ShotProfileWEM.jl:104-201, 288-308): adj=1 m = sum_shots wem_adj(d[shot]);
 * adj=0 d[shot] = wem_fwd(m).  d [nshot][nm][nt], m [nm][nz]. */
int orc64_shots(double* d, double* m, const float* wav, const float* vel, const orc_geom* g,
                const float* sx, const float* sy, int nshot, int adj);
int orc32_shots(float* d, float* m, const float* wav, const float* vel, const orc_geom* g,
                const float* sx, const float* sy, int nshot, int adj);
/* compute_angles (wem.c:566-992): source incidence angles (degrees) angx, angy [nm][nz] of one
 * shot; the geometry's padt / padx / pspi are ignored (ntfft = nt, nkx = nmx, split-step). */
int orc64_angles(double* angx, double* angy, const float* wav, const float* vel, const orc_geom* g, float sx, float sy);
int orc32_angles(float* angx, float* angy, const float* wav, const float* vel, const orc_geom* g, float sx, float sy);
/* ewem (ewem.c:3-366), 2-D: adj: ux, uz [nm][nt] -> mpp, mps2 [nm][nz]; else the reverse. */
int orc64_ewem(double* ux, double* uz, double* mpp, double* mps2, const float* wav, const float* vp, const float* vs,
               const orc_geom* g, float sx, float sy, int adj);
int orc32_ewem(float* ux, float* uz, float* mpp, float* mps2, const float* wav, const float* vp, const float* vs,
               const orc_geom* g, float sx, float sy, int adj);
/* elastic_separate_3d's matrix at one wavenumber (ewem.c:600-630, 680-710): complex in / out as
 * (re, im) pairs. */
void orc64_esep_k(double kx, double w, double vp, double vs, int sep, int adj, const double* in1, const double* in2,
                  double* o1, double* o2);
void orc32_esep_k(float kx, float w, float vp, float vs, int sep, int adj, const float* in1, const float* in2,
                  float* o1, float* o2);

#ifdef __cplusplus
}
#endif
#endif
