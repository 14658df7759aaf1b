/*
 * seiswem.h -- C ABI of libseiswem.so, the MI355X (gfx950) shot-profile wave-equation
 * migration / demigration library (split-step Fourier and PSPI, 2-D and 3-D).
 *
 * Every entry point returns an int status (SWEM_OK == 0) and never calls exit(); the text of
 * the last error on the calling thread is available from swem_last_error().  Plain pointers
 * and sizes only.  Array layouts follow the reference's alloc2float convention
 * (src/API/c/seismic.c:92-106): a [n2][n1] array is contiguous with n1 fastest, so
 *   d   [nm][nt]  traces of one shot, time fastest      (wem.c:3, d[ix][it])
 *   m   [nm][nz]  image, depth fastest                  (wem.c:3, m[ix][iz])
 *   vel [nm][nz]  velocity, depth fastest               (wem.c:9, vel[ix][iz])
 *   wav [nt]      source wavelet                         (wem.c:3, wav[0][it])
 * with trace index ix = imx*nmy + imy (y fastest) and nm = nmx*nmy.
 */
#ifndef SEISWEM_H
#define SEISWEM_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SWEM_VERSION 100

/* Status codes. */
#define SWEM_OK 0
#define SWEM_EINVAL 1        /* invalid argument (missing file name, bad size, nref < 2, ...) */
#define SWEM_ENOMEM 2        /* host or device allocation failed, or budget exceeded */
#define SWEM_EDEVICE 3       /* HIP runtime error, or no usable gfx950 device */
#define SWEM_EIO 4           /* .seis file could not be opened / read / written */
#define SWEM_EUNSUPPORTED 5  /* valid input this build does not handle yet (named in the message) */
#define SWEM_EBUSY 6         /* serving queue full */

/* pspi_mode */
#define SWEM_PSPI_COMPAT 0   /* the reference: real-part cut + nearest lower vref, both directions
                                (wem.c:450-459, linear_interp wem.c:473-478); not an adjoint pair */
#define SWEM_PSPI_LINEAR 1   /* linear interpolation between vref1 and vref2 (the formula at
                                wem.c:476); the forward operator is the exact adjoint */

/* precision */
#define SWEM_PREC_KZ64 0     /* FP32 propagation, kz phase built from FP64 s = w^2 p^2 - k^2,
                                compensated image sums (default) */
#define SWEM_PREC_FP32 1     /* naive FP32: s, sqrt and phase in float like wem.c:341-343 */

/* variant of the 2-D depth-march kernels */
#define SWEM_VARIANT_AUTO 0
#define SWEM_VARIANT_SUBGROUP 1  /* nkx/8 threads per wavefield, ping-pong LDS, barrier per pass */
#define SWEM_VARIANT_WAVE 2      /* one 64-lane wave per wavefield, no barrier inside a transform */
#define SWEM_VARIANT_SUBGROUP16 3 /* nkx/16 threads per wavefield, radix-16 passes (nkx >= 1024) */
#define SWEM_VARIANT_PAIR 4      /* migration: one subgroup per frequency marches source and receiver */
#define SWEM_VARIANT_PAIR8 5     /* PAIR with nkx/8 threads (8 points each), nkx >= 512 */
#define SWEM_VARIANT_PACKED 6    /* PSPI compat migration: both real wavefields in one transform (AUTO where valid) */

/* Geometry of one migration problem: the scalar arguments of wem() (wem.c:3-11). */
typedef struct swem_geometry {
    int nt;   float ot, dt;
    int nmx;  float omx, dmx;
    int nmy;  float omy, dmy;
    int nz;   float oz, dz;
    float gz, sz;
    int nref; float fmin, fmax;
    int padt, padx;
    int pspi;               /* 0 = SSF (ssop, wem.c:289), 1 = PSPI (pspiop, wem.c:380) */
} swem_geometry;

typedef struct swem_options {
    int pspi_mode;          /* SWEM_PSPI_* (default COMPAT) */
    int precision;          /* SWEM_PREC_* (default KZ64) */
    int device;             /* HIP device ordinal (default 0) */
    int freqs_per_block;    /* frequencies marched together by one workgroup; 0 = autotuned */
    int verbose;            /* progress to stderr like progress_msg (wem.c:516-520) */
    size_t scratch_bytes;   /* device work memory of one call (spectra, partial images, source
                               wavefields); 0 = min(96 GiB, free/2).  Shots beyond it run in
                               several launches. */
    int variant;            /* SWEM_VARIANT_* (default AUTO) */
    size_t table_bytes;     /* device memory for the phase tables; 0 = min(32 GiB, free/4).
                               A band whose tables exceed it is marched in frequency chunks,
                               the tables of each chunk rebuilt per call. */
} swem_options;

/* Per-call timing of the last plan call (HIP events on the plan's stream). */
typedef struct swem_timing {
    double total_ms;        /* first event to last event */
    double spectra_ms;      /* time -> frequency transforms */
    double march_ms;        /* sum of depth-march kernel durations */
    double reduce_ms;       /* image reduction / frequency -> time */
    int march_launches;     /* number of depth-march kernel launches */
    double march_flops;     /* algorithmic FP32 flops of the march launches (5 N log2 N per FFT) */
} swem_timing;

/* ---------------------------------------------------------------- library status */
int swem_version(void);
const char* swem_last_error(void);
void swem_options_init(swem_options* opt);
int swem_device_count(int* count);

/* ------------------------------------------------------------ reference drop-ins */

/* Replaces shotwem.so's `int main(int argc, char** argv)` (shotwem.c:12-140): the same
 * key=value argv protocol (seismic.c:2-71; keys adj, pspi, verbose, d, m, vel, wav, sz, gz,
 * sx, sy, fmin, fmax, padt, padx, nref), the same .seis files in and out (seismic.c:270-461),
 * and the same header rewrites (shotwem.c:86-98, 113-117).  Unlike the reference it returns
 * SWEM_EINVAL instead of calling exit(1) when a required file name is missing.  The library
 * shotwem.so built beside libseiswem.so exports this function under the name `main`, so the
 * Julia ccall of ShotProfileWEM.jl:347 binds to it unchanged. */
int swem_shot_main(int argc, char** argv);

/* Replaces compute_angles.so's `int main(int argc, char** argv)` (compute_angles.c:12-118): keys
 * angx, angy, vel, wav (required), dipx, dipy (required with dip_flag=y), dip_flag, sx, sy, sz,
 * fmin, fmax, verbose; writes the angle files angx and angy.  compute_angles.so built beside
 * libseiswem.so exports it as `main` for ComputeAngles.jl's ccall.  Returns SWEM_EINVAL instead
 * of calling exit(1) when a required name is missing. */
int swem_angles_main(int argc, char** argv);

/* Replaces shotewem.so's `int main(int argc, char** argv)` (shotewem.c:11-223): keys ux, uy, uz,
 * mpp, mps1, mps2, vp, vs, wav (required), adj, verbose, sx, sy, sz, gz, fmin, fmax, padt, padx;
 * the header rewrites of shotewem.c:117-151 and 171-190.  shotewem.so exports it as `main`. */
int swem_eshot_main(int argc, char** argv);

/* Replaces wem() (wem.c:3-11) with host arrays: adj=1 migrates d [nm][nt] into m [nm][nz]
 * (overwritten), adj=0 models m into d (overwritten).  pspi selects PSPI (compat mode) or SSF. */
int swem_wem(float* d, float* m, const float* wav,
             int nt, float ot, float dt,
             int nmx, float omx, float dmx,
             int nmy, float omy, float dmy,
             float sx, float sy,
             int nz, float oz, float dz, float gz, float sz,
             const float* vel, int nref, float fmin, float fmax,
             int padt, int padx,
             int adj, int pspi, int verbose);

/* Replaces compute_angles() (wem.c:566-574, the routine behind compute_angles.c:12-118):
 * source-side incidence angles in degrees, angx and angy [nm][nz] (host arrays, overwritten),
 * of one shot at (sx, sy), from a split-step source march with ntfft = nt and nkx = nmx and the
 * spatial derivatives of the source wavefield per frequency (wem.c:694-981).  2-D: nmx and nz up
 * to 2048, or powers of two to 4096; 3-D: nmx and nmy in 2..512 or powers of two to 1024 (others
 * return SWEM_EUNSUPPORTED).  angy is zero in 2-D, as in the reference. */
int swem_compute_angles(float* angx, float* angy, const float* wav,
                        int nt, float ot, float dt,
                        int nmx, float omx, float dmx,
                        int nmy, float omy, float dmy,
                        float sx, float sy,
                        int nz, float oz, float dz, float sz,
                        const float* vel, float fmin, float fmax, int verbose);

/* compute_angles with a geometry and explicit options (device, work budget scratch_bytes, table
 * cap table_bytes); padt, padx, pspi, nref and gz of the geometry are not used. */
int swem_compute_angles_ex(float* angx, float* angy, const float* wav, const swem_geometry* g,
                           float sx, float sy, const float* vel, const swem_options* opt);

/* Replaces ewem() (ewem.c:3-15), elastic shot-profile migration (adj=1: components ux, uy, uz
 * [nm][nt] -> images mpp, mps1, mps2 [nm][nz]) and modelling (adj=0: the reverse), with the P
 * source, the P / S separation of the receiver components at every depth and split-step
 * propagation with vp and vs (ewem.c:255-741).  2-D (nmy = 1) and 3-D (nmy > 1: the source at
 * igx nmy + igy, the separation matrix of kx alone applied on the 2-D transform, as
 * elastic_separate_3d does, ewem.c:553-741); as in the reference, uy is not used (migration) and
 * returned zero (modelling), and mps1 is returned zero and not used.
 * The band is sampled as ewem.c samples it (ewem.c:43-52: an even ntfft = 2 truncf(padt nt / 2),
 * ifmax = trunc(fmax dt ntfft) + 1), which differs from wem.c's rule for odd padt nt or a
 * fractional fmax / fmin.  3-D takes the 3-D march's transform lengths (nkx, nky up to 512, powers
 * of two up to 1024); a source outside the grid returns SWEM_EINVAL. */
int swem_ewem(float* ux, float* uy, float* uz, float* mpp, float* mps1, float* mps2, const float* wav,
              int nt, float ot, float dt,
              int nmx, float omx, float dmx,
              int nmy, float omy, float dmy,
              float sx, float sy,
              int nz, float oz, float dz, float gz, float sz,
              const float* vp, const float* vs, float fmin, float fmax,
              int padt, int padx, int adj, int verbose);

/* ShotProfileEWEM (ShotProfileEWEM.jl) in memory, 2-D: nshot shots over one pair of velocity
 * grids, all shots of a chunk in one launch.  adj=1: mpp, mps1, mps2 [nm][nz] = sums over shots of
 * ewem's images of ux, uz [nshot][nm][nt] (uy unused); adj=0: ux, uy, uz [nshot][nm][nt] from the
 * images (uy zero).  sy may be NULL. */
int swem_eshots(const swem_geometry* g, const float* vp, const float* vs, const float* wav,
                const float* sx, const float* sy, int nshot,
                float* ux, float* uy, float* uz, float* mpp, float* mps1, float* mps2,
                int adj, const swem_options* opt);

/* wem() with explicit options (pspi_mode, precision, device...). */
int swem_wem_ex(float* d, float* m, const float* wav, const swem_geometry* g,
                float sx, float sy, const float* vel, int adj, const swem_options* opt);

/* ShotProfileWEM (ShotProfileWEM.jl:39-310) in memory with nangx = nangy = 1: several shots
 * over one velocity grid.  adj=1: m [nm][nz] = sum over shots of wem_adj(d[shot]) (the
 * stacking of ShotProfileWEM.jl:140-201); adj=0: d[shot] = wem_fwd(m) for every shot
 * (ShotProfileWEM.jl:288-308).  d [nshot][nm][nt]; sx, sy [nshot] (sy may be NULL). */
int swem_shots(const swem_geometry* g, const float* vel, const float* wav,
               const float* sx, const float* sy, int nshot,
               float* d, float* m, int adj, const swem_options* opt);

/* ---------------------------------------------------------- device-resident plans */
typedef struct swem_plan swem_plan;

/* Uploads velocity-derived tables and the source spectrum once; vel and wav are host arrays. */
int swem_plan_create(const swem_geometry* g, const float* vel, const float* wav,
                     const swem_options* opt, swem_plan** out);
int swem_plan_destroy(swem_plan* plan);

/* Migration of nshot shots whose records d_dev [nshot][nm][nt] are already in device memory.
 * Writes the stacked image m_dev [nm][nz] (device), adding to it when accumulate != 0.
 * sx, sy are host arrays [nshot] (sy may be NULL).  stream is a hipStream_t (NULL = default).
 * Returns after the work is enqueued; synchronise the stream before reading m_dev. */
int swem_plan_migrate(swem_plan* plan, const float* d_dev, const float* sx, const float* sy,
                      int nshot, float* m_dev, int accumulate, void* stream);
/* As swem_plan_migrate but one image per shot: m_dev [nshot][nm][nz], no stacking. */
int swem_plan_migrate_each(swem_plan* plan, const float* d_dev, const float* sx, const float* sy,
                           int nshot, float* m_dev, void* stream);
/* No counterpart in the reference: swem_plan_migrate for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and result as swem_plan_migrate (within FP32 rounding of the frequency-chunk
 * sums), 2-D plans only; calls no hipMalloc / hipFree.  Its work memory (source indices,
 * receiver spectra, partial images) is a fixed 128 MiB scratch array in device memory, one per
 * GPU, held until the process ends; shots and frequencies are launched in chunks that fit it.
 * SWEM_ENOMEM when one shot's spectra plus one partial image exceed it (nm * nf * 8 +
 * nz * nmx * 4 bytes above 128 MiB - 64 KiB), SWEM_EUNSUPPORTED for a 3-D plan.  Every call on
 * a GPU shares that GPU's scratch: calls that overlap on one GPU (two host threads, or two
 * streams in flight at once) race and must be serialised by the caller. */
int swem_plan_migrate_noalloc(swem_plan* plan, const float* d_dev, const float* sx, const float* sy,
                              int nshot, float* m_dev, int accumulate, void* stream);
/* Demigration (modelling) of the image m_dev [nm][nz] into d_dev [nshot][nm][nt]. */
int swem_plan_model(swem_plan* plan, const float* m_dev, const float* sx, const float* sy,
                    int nshot, float* d_dev, void* stream);
/* Timing of the last migrate/model call on this plan (synchronises its stream). */
int swem_plan_timing(swem_plan* plan, swem_timing* out);

/* Measured launch shapes of a plan (no reference counterpart; the reference plans FFTW with
 * FFTW_ESTIMATE, wem.c:103-104).  2-D: every migration kernel variant that fits the plan's nkx,
 * each with every frequency group size (frequencies per workgroup, 1..16) its workgroup and LDS
 * allow, is timed on nshot shots of pseudo-random records (the work is the same for any records
 * with nonzero spectra).  3-D:
 * frequencies per y-pass workgroup (2..16, 8 at most for PSPI or a Bluestein y axis) times the step
 * state of one launch (64, 128, 256 MiB, i.e. frequencies marched together against the 256 MB
 * last-level cache), timed on the plan's first 16 depths.  The fastest candidate becomes the
 * plan's shape for later migrations (a freqs_per_block given at plan creation is kept).  Each
 * candidate takes two migrations; synchronous. */
#define SWEM_TUNE_MAX 48
typedef struct swem_tune_result {
    int variant;            /* 2-D: chosen SWEM_VARIANT_*; 3-D: SWEM_VARIANT_AUTO */
    int freqs_per_block;    /* chosen frequencies per workgroup (3-D: per y-pass workgroup) */
    size_t chunk_bytes;     /* 3-D: chosen step state of one launch; 2-D: 0 */
    float best_ms;          /* march time of the choice (3-D: of the first 16 depths) */
    int ncand;              /* candidates timed */
    int cand_variant[SWEM_TUNE_MAX], cand_fpb[SWEM_TUNE_MAX];
    size_t cand_chunk_bytes[SWEM_TUNE_MAX];
    float cand_ms[SWEM_TUNE_MAX];
} swem_tune_result;
int swem_plan_autotune_ex(swem_plan* plan, int nshot, swem_tune_result* out);
/* The 2-D variant choice of swem_plan_autotune_ex: *best receives the chosen variant and
 * times[0..SWEM_VARIANT_PACKED] each variant's march time in ms at its best group size (0: not a
 * distinct candidate; may be NULL).  A 3-D plan is tuned as by swem_plan_autotune_ex, *best =
 * SWEM_VARIANT_AUTO and times all 0. */
int swem_plan_autotune(swem_plan* plan, int nshot, int* best, float* times);
/* No reference counterpart.  enable != 0 (the default): a 3-D migration launches each shot as one
 * captured graph (every frequency chunk and depth: x-pass, y-pass, depth image), captured on the
 * plan's first 3-D migration and again whenever its work buffers or launch shapes change, replayed
 * for every later shot; the shot's source index and output are read from device memory.
 * enable = 0 launches the same kernels one by one.  2-D plans ignore it. */
int swem_plan_set_graphs(swem_plan* plan, int enable);
/* Device bytes held by the plan's tables and its largest work chunk. */
int swem_plan_bytes(const swem_plan* plan, size_t* table_bytes, size_t* work_bytes);

/* ------------------------------------------------ serving loop (continuous batching) */
/* No counterpart in the reference, whose driver runs one shotwem process per shot
 * (ShotProfileWEM.jl:104-128, 288): these are the batched / served entry points of the
 * drop-in.  A server owns one device arena, one pinned staging arena, a few streams and a worker
 * thread.  Migration requests of any shape are submitted from any thread (inputs are copied);
 * the worker admits them as arena memory and in-flight slots free up, runs each with its own
 * plan on one of the streams (requests overlap on the device) and keeps the Fourier bases of
 * each band resident.  Each result is bit-identical to the same request served alone. */
#define SWEM_REQ_PENDING 0
#define SWEM_REQ_RUNNING 1
#define SWEM_REQ_DONE 2
#define SWEM_REQ_FAILED 3

typedef struct swem_server swem_server;
typedef struct swem_server_config {
    int device;
    size_t device_bytes;    /* device arena; 0 = min(64 GiB, free/2) */
    size_t pinned_bytes;    /* pinned staging arena; 0 = 4 GiB */
    int streams;            /* HIP streams requests are spread over (default 4) */
    int max_inflight;       /* requests on the device at once (default 8) */
    int max_pending;        /* queued requests before submit returns SWEM_EBUSY (default 4096) */
    double skip_window_ms;  /* a head-of-queue request that does not fit lets later ones pass
                               for this long after its submission (default 500) */
    int max_fuse;           /* single-shot requests with the same geometry, options, velocity and
                               wavelet that are queued together run as the shots of one launch on
                               one plan, at most this many (default 32; 1 = never); each keeps its
                               own image, bit-identical to the request served alone */
} swem_server_config;

typedef struct swem_request {   /* one migration request: the arguments of swem_shots, adj=1 */
    swem_geometry g;
    int pspi_mode, precision;
    const float* vel;       /* [nm][nz] */
    const float* wav;       /* [nt] */
    const float* d;         /* [nshot][nm][nt] */
    const float* sx;        /* [nshot] */
    const float* sy;        /* [nshot] or NULL */
    int nshot;
} swem_request;

typedef struct swem_request_status {
    int state;              /* SWEM_REQ_* */
    int error;              /* status code of a failed request */
    double submit_ms, admit_ms, done_ms;   /* server clock (ms since creation) */
    char message[256];
} swem_request_status;

typedef struct swem_server_stats {
    long long submitted, completed, failed;
    int pending, inflight;
    size_t device_bytes, device_bytes_used, device_bytes_peak, basis_bytes;
    double busy_ms;         /* time with at least one request on the device */
    long long fused_launches, fused_requests;  /* launches that carried several requests, and those requests */
} swem_server_stats;

void swem_server_config_init(swem_server_config* cfg);
int swem_server_create(const swem_server_config* cfg, swem_server** out);
/* Finishes the admitted and queued requests, then frees everything. */
int swem_server_destroy(swem_server* server);
/* Queues a request; *id names it in poll / wait.  SWEM_EBUSY when the queue is full. */
int swem_batch_submit(swem_server* server, const swem_request* req, long long* id);
int swem_batch_poll(swem_server* server, long long id, swem_request_status* status);
/* Blocks until the request is done, copies its stacked image [nm][nz] into m (if not NULL),
 * fills status and forgets the request.  Returns the request's error code if it failed.  One
 * thread may wait for a given id: a second concurrent wait on it returns SWEM_EINVAL. */
int swem_batch_wait(swem_server* server, long long id, float* m, swem_request_status* status);
int swem_server_get_stats(swem_server* server, swem_server_stats* out);

/* ---------------------------------------------------------- device memory helpers */
int swem_malloc(void** ptr, size_t bytes);
int swem_free(void* ptr);
int swem_memcpy_htod(void* dst, const void* src, size_t bytes);
int swem_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int swem_memcpy_dtod(void* dst, const void* src, size_t bytes);
int swem_memset(void* ptr, int value, size_t bytes);
int swem_synchronize(void);

/* Device vector operations of the least-squares solver, for vectors resident in device memory
 * (models [nm][nz], records [nshot][nm][nt]):
 *   swem_vec_cgstep  m1 = a*m1 + b*m2                       (CGStep, src/Solvers/CGStep.jl:1-4)
 *   swem_vec_mul     y = x .* w, y may be x                  (WeightingOp, src/Operators/WeightingOp.jl:1-5:
 *                    the data weights and mute of ShotProfileLSWEM, src/Imaging/ShotProfileLSWEM.jl:12-15)
 *   swem_vec_dot     *out = sum x*y accumulated in double in an order that depends only on n
 *                    (InnerProduct, src/Tools/InnerProduct.jl:1-5, whose in-memory form rounds to
 *                    Float32); waits for the stream. */
int swem_vec_cgstep(float a, float* m1, float b, const float* m2, size_t n, void* stream);
int swem_vec_dot(const float* x, const float* y, size_t n, double* out, void* stream);
int swem_vec_mul(const float* x, const float* w, float* y, size_t n, void* stream);

/* ------------------------------------------- discrete decisions (host only, no GPU) */

/* ntfft, nw, nkx, nky, ifmin, ifmax, lmx, lmy with the reference's casts (wem.c:79-89,
 * 305-308): out[8] in that order. */
int swem_sampling(const swem_geometry* g, int* out);
/* igx = (int)(sx-omx)/dmx (wem.c:112-113); SWEM_EINVAL outside the grid. */
int swem_source_index(const swem_geometry* g, float sx, float sy, int* igx, int* igy);
/* Reference velocities vref [nref][nz], indices iref1/iref2 [nm][nz] and the PSPI branch
 * flag cut [nm][nz] = (vref2 - vref1 > 10) with the reference's float arithmetic
 * (wem.c:40-70, 454). */
int swem_pspi_tables(const swem_geometry* g, const float* vel, float* vref, int* iref1,
                     int* iref2, unsigned char* cut);
/* Edge taper of boundary_condition (wem.c:546-564) for one axis of n points, width l. */
int swem_taper(int n, int l, float* t);

/* .seis I/O with the reference's layout (seismic.c:270-461, seismic.h:20-86): a text header
 * file, a data file of native floats (n1 fastest) and 31 x 4-byte trace headers per trace. */
#define SWEM_SEIS_HEADER_WORDS 31
typedef struct swem_seis_extent {
    int n1, n2, n3, n4, n5;
    float o1, o2, o3, o4, o5;
    float d1, d2, d3, d4, d5;
    char label1[100], label2[100], label3[100], label4[100], label5[100];
    char unit1[100], unit2[100], unit3[100], unit4[100], unit5[100];
    char title[100];
    char data_format[100];
    int esize;
    char dname[100];
    char hname[100];
} swem_seis_extent;
void swem_seis_extent_init(swem_seis_extent* fh);
int swem_seis_read_header(const char* name, swem_seis_extent* fh);
/* data [ntrace][n1] floats, headers [ntrace][31] 4-byte words (either may be NULL). */
int swem_seis_read(const char* name, float* data, void* headers, swem_seis_extent* fh);
int swem_seis_write(const char* name, const float* data, const void* headers, swem_seis_extent* fh);

#ifdef __cplusplus
}
#endif
#endif
