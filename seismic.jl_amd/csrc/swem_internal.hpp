// swem_internal.hpp -- declarations shared by the host code and the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

#include "../../include/seiswem.h"

namespace swem {

// Error state (thread-local message, see swem_host.cpp).
int set_error(int code, const char* fmt, ...);
int hip_error(hipError_t e, const char* what, const char* file, int line);

#define SWEM_HIP(call)                                                        \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) return ::swem::hip_error(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define SWEM_TRY(expr)            \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != SWEM_OK) return rc_; \
    } while (0)

// Derived sampling of one wem() call, with the reference's casts (wem.c:79-89, 305-308).
struct Sampling {
    int ntfft, nw, nkx, nky, ifmin, ifmax, lmx, lmy, fop_n;
    float dkx, dky, dw;
};

int compute_sampling(const swem_geometry& g, Sampling& s);
int source_index(const swem_geometry& g, float sx, float sy, int& igx, int& igy);
int pspi_tables(const swem_geometry& g, const float* vel, float* vref, int* iref1, int* iref2, uint8_t* cut);
void taper_f32(int n, int l, float* t);
void slowness_f32(const swem_geometry& g, const float* vel, float* po, float* pd);

// Mixed-radix transform plan of a length n = 2^a 3^b 5^c <= 4096 (mixed_fft in swem_fft.hpp):
// the radix of every pass in order (8 while 8 divides the rest, then 4, 2, 5, 3), the product of
// the radices before it, its twiddle step n / (ns r) and the multiplier that divides by ns
// (b / ns = umulhi(b, magic) for b ns < 2^32; 0 where ns = 1).
constexpr int kMixedMaxPass = 12;
struct MixedRadix {
    int n, np;  // length; passes (np = 0 for n = 1: the identity)
    uint16_t ns[kMixedMaxPass];
    uint16_t tws[kMixedMaxPass];
    uint32_t magic[kMixedMaxPass];
    uint8_t r[kMixedMaxPass];
};
// The plan of n (mr may be null); false when n is not 2^a 3^b 5^c or exceeds 4096.
bool mixed_radix_plan(int n, MixedRadix* mr);
// How a transform of length n runs: a power-of-two kernel on n itself, Bluestein's identity on a
// longer power-of-two array, or the mixed-radix transform on the next power of two.
enum { kLenPow2 = 0, kLenBluestein = 1, kLenMixed = 2 };

// Arguments of the FP64 phase-table builders (swem_phase.hip).
struct PhaseTableArgs {
    int nfc, f0, ifmin, nz, nref, nkh, nmx, prec;
    double dw, dkx, dz, inv_nk;
    float dwf, dkxf;
    const double* p2;   // [nz][nref] po^2 (SSF, nref = 1) or 1/vref^2 (PSPI)
    const float* p2f;   // float counterparts (naive FP32)
    const double* pd;   // [nz][nmx] slowness perturbation (SSF)
    const float* pdf;   // float counterpart (naive FP32)
    float2* ktab;       // [nfc][nz][nref][nkh]
    float2* stab;       // [nfc][nz][nmx]
};
int launch_phase_tables(const PhaseTableArgs& a, bool ss, hipStream_t st);

// Arguments of the 2-D depth-march kernels.  A launch covers the frequency chunk
// [f0, f0 + nfc) of the band (indices relative to ifmin) for nshot shots; the block index is
// decoded XCD-aware into (frequency group g, shot) so the shots that share a frequency run on
// the same XCD and share its L2 copy of the phase tables.
// Uncut-point table of the packed PSPI compat kernel: depths with at most kUncutSparse uncut points
// use direct sums over them instead of two transforms.
constexpr int kUncutSparse = 4;
constexpr int kCrefStride = 3 + kUncutSparse;
// Complex LDS scratch after the transform buffers of a packed-kernel workgroup.
constexpr int kPackedScratch = 80;

struct March2DArgs {
    int nmx, nz, nf, nref, ngroups, ifmin, nshot;
    int f0, nfc, nkh;
    float oz, dz, sz, gz;
    float inv_nk;           // 1 / nkx (identity steps)
    int nkx;                // transform length
    const float2* tw;       // [nkx] exp(-2 pi i m / nkx) (power-of-two and mixed-radix lengths);
                            // Bluestein: [M] of the array length M
    MixedRadix mr;          // mixed-radix lengths: the plan of nkx
    const float2* chirp;    // Bluestein: [nkx] exp(-i pi j^2 / nkx)
    const float2* bchirp;   // Bluestein: [2][M] FFT_M of the wrapped conj chirp / M (SIGN -1, +1)
    const float2* ktab;     // [nfc][nz][nref][nkh] kz phase / nk, +dz
    const float2* stab;     // [nfc][nz][nmx] split-step phase, +dz (SSF)
    const float2* spec_s;   // [nf] source spectrum / sqrt(ntfft)
    const int* igx;         // [nshot]
    const float2* spec_r;   // [nshot][nmx][nf]  receiver spectra w^2 D / sqrt(ntfft)
    const float* taper;     // [nmx]
    const uint8_t* sel;     // [nz][nkx] iref1 | cut << 7 (PSPI; 0x7f beyond nmx)
    const uint8_t* sel2;    // [nz][nkx] iref2 (PSPI linear)
    const float* w2;        // [nz][nkx] weight of iref2 (PSPI linear)
    const uint32_t* used;   // [nz] bit ir set when some x selects vref ir at iz
    const int* cref;        // [nz][kCrefStride] uncut points of depth iz (PSPI compat): vref (-1: none),
                            // count, 1 when no cut point selects that vref, the first kUncutSparse
                            // positions x
    float* partial;         // [nshot][ngroups][nz][nmx] migration partial images
    const float* mT;        // [nz][nmx] image (modelling input, depth-major)
    float2* smig;           // [nshot][nfc][nz][nmx] source wavefield scratch (modelling)
    float2* dgw;            // [nshot][nmx][nf] modelled spectra (modelling output)
};

// Source incidence angles (compute_angles, wem.c:566-992).  The march writes every depth of the
// source wavefield to vol [nfc][nz][nm] (trace ix * nmy + iy); the z pass writes |u_z| to dzmag;
// the lateral pass, along x or y lines, adds that axis's angles and left / right signs of every
// frequency, in order, to acc / sgn [nz][nm].  A line is `len` points `elem` apart starting at
// (l / inner) * outer + (l % inner) * inner_stride for line l (z pass: l over nm points of one
// frequency; lateral pass: l over the nz * (nm / len) lines of a depth).  ax: lateral transforms
// (length len, tables of the plan or of y); az: z transforms (length nz).
struct AngleArgs {
    int len, inner, lines, nfc, nw, nz;  // lines: workgroups of the pass
    long long outer, inner_stride, elem, item;  // item: distance between frequencies in vol / dzmag
    float dk, dkz;          // wavenumber spacing of the line's axis and of z (float, as the reference)
    const float2* vol;      // [nfc][nz][nm]
    float* dzmag;           // [nfc][nz][nm]
    float* acc;             // [nz][nm] sum of angles (degrees)
    float* sgn;             // [nz][nm] sum of signed |u| / nw
};
// which: 0 = march (F frequencies per workgroup), 1 = column pass, 2 = row pass.
int launch_angles(int which, const March2DArgs& ax, const March2DArgs& az, const AngleArgs& b, int F, hipStream_t st);
int angles_supported(int n);

// Elastic WEM (ewem.c), 2-D.  One workgroup per (frequency of the chunk, shot); ap / as carry
// the P and S plans' tables (kz phases of po_p / po_s, split-step phases of pd_p / pd_s), ap also
// the source spectrum, igx [nshot], nshot and the geometry.
struct EwemArgs {
    int nf;                 // band size (stride of the spectra)
    float dwf, dkx;         // float dw and dkx, as the reference computes w and kx
    float inv_maxs;         // 1 / (max|wav| / sqrt(ntfft)) (ewem.c:146)
    const float* vp0;       // [nz] 1 / po_p (separation)
    const float* vs0;       // [nz] 1 / po_s
    const float2* spx;      // [nshot][nmx][nf] R2C(ux) / sqrt(ntfft) (migration)
    const float2* spz;      // [nshot][nmx][nf]
    float* ppart;           // [nshot][nfc][nz][nmx] partial P image (migration)
    float* spart;           // [nshot][nfc][nz][nmx] partial S image
    float2* smig;           // [nshot][nfc][nz][nmx] stored source (modelling)
    const float* mppT;      // [nz][nmx] images (modelling input)
    const float* mpsT;
    float2* dgx;            // [nshot][nmx][nf] modelled spectra (modelling output)
    float2* dgz;
};
int launch_ewem(const March2DArgs& ap, const March2DArgs& as, const EwemArgs& e, int adj, hipStream_t st);

// mode: 0 = SSF, 1 = PSPI compat, 2 = PSPI linear.  variant: SWEM_VARIANT_*.  F: frequencies
// per workgroup.
int launch_march2d(const March2DArgs& a, int nkx, int F, int mode, int adj, int variant, hipStream_t st);
// Number of blocks of the XCD-aware 1-D grid of a launch.
int march2d_grid_blocks(int ngroups, int nshot);
// Shape of the 2-D kernel a variant selects (T = 0: nkx unsupported).
struct March2DShape {
    int variant;    // resolved SWEM_VARIANT_*
    int T;          // threads per wavefield subgroup
    int per_freq;   // subgroups per frequency
    int max_block;  // largest workgroup
    int bufs;       // LDS buffers per subgroup
    int blue;       // array length of a length that is not a power of two 128..4096 (0: nkx is one)
    int mixed;      // that length runs the mixed-radix transform (else Bluestein's)
};
March2DShape march2d_shape(int nkx, int variant, int adj);
bool march2d_supported(int nkx);
// Array length of a 2-D transform length that is not a power of two 128..4096: the next power
// of two (at least 128) for mixed-radix lengths up to 4096, Bluestein's 2 nkx - 1 rounded up for
// the others up to 2048 (0: a power-of-two kernel takes nkx, or nkx is out of range).
int march2d_blue_len(int nkx);
// kLenPow2, kLenBluestein or kLenMixed for a supported 2-D length.
int march2d_kind(int nkx);
size_t march2d_lds_bytes(int nkx, const March2DShape& sh, int F);

// Arguments of the 3-D split-step kernels (swem_march3d.hip).  A launch covers the items
// (frequencies f0 .. f0 + nitem of one shot) of a frequency chunk at depth iz.
// Per-shot values of a 3-D migration read by the kernels from device memory, so that the launch
// sequence of a shot does not depend on the shot (it is captured once as a graph and replayed).
struct Shot3D {
    int igx;     // source trace ix * nmy + iy
    int acc;     // add this shot's image to out (else overwrite)
    float* out;  // image [nm][nz] the shot's depths are reduced into
};

struct March3DArgs {
    int nmx, nmy, nm, nkx, nky, nz, nf, ifmin, f0, nitem, iz, prec;
    int igx;                    // source trace ix * nmy + iy
    int adj;                    // init: load receiver spectra (migration) or zeros (modelling)
    int wmask;                  // x-pass: bit 0 source, bit 1 receiver
    int conj_src, conj_rec;     // x-pass: conj(L) for a -dz step
    int ds, dr, next_s, next_r; // stepped at this depth / at the next depth
    int first;                  // modelling receiver: first step (zero state)
    double dw, dkx, dky, dzd, dkx2;  // dkx2 = dkx^2
    float dkxf, dkyf, inv_nk;
    const double* p2;           // [nz] po^2
    const float* p2f;           // [nz] float po^2 (naive FP32)
    const double* pd;           // [nz][nm] slowness perturbation
    const float* pdf;           // [nz][nm] float counterpart
    const float* taper_x;       // [nmx]
    const float* taper_y;       // [nmy]
    const float2* twx;          // [march3d_len(nkx)] twiddles of the x array length; [nkx] of nkx
                                // for a mixed-radix length
    const float2* twy;          // [march3d_len(nky)] (mixed radix: [nky])
    MixedRadix mrx, mry;        // mixed-radix plans of nkx, nky (kLenMixed axes)
    // Bluestein tables of an axis whose array length differs from its nk: [nk] chirp
    // exp(-i pi j^2 / nk), [2][array length] kernel spectra (transform sign -1, +1)
    const float2 *chx, *bchx, *chy, *bchy;
    const float2* spec_s;       // [nf]
    const float2* spec_r;       // [nm][nf] this shot's receiver spectra
    float2* W;                  // [nitem][2][nmx][nky] step state (y transformed)
    float* partial;             // [ngroups][nm] partial image of this depth
    const float* mT;            // [nz][nm] image (modelling input)
    float2* smig;               // [nitem][nz][nm] source wavefields (modelling)
    float2* dgw;                // [nm][nf] modelled spectra of this shot
    // PSPI (compat): reference velocities, per-point selection iref1 | cut << 7 ([nz][nm]),
    // vrefs in use per depth, per-vref x-transformed slices [nitem][2][nref][nmx][nky], and the
    // depth whose step produced the slices a y-pass selects from
    int nref, iz_sel;
    const uint8_t* sel;
    const uint32_t* used;
    float2* R;
    const Shot3D* shot;         // when set, igx is read from it (migration)
    // PSPI linear mode: iref2 per point ([nz][nm]) and the interpolation weight of iref2
    int linear;
    const uint8_t* sel2;
    const float* w2;
};
// 3-D elastic WEM (ewem.c:3-366 with nmy > 1): three wavefield slots per frequency item,
// W [nitem][3][nmx][nky] (y transformed, x in space): 0 the P source, 1 and 2 the receiver pair
// (ux, uz before the separation and after the combination; up, us2 between).
struct Ewem3DArgs {
    float2* W;
    int igx;                // source trace igx nmy + igy (ewem.c:99-108)
    int sep_adj;            // matrix branch of the separation / combination (1: migration)
    int ds, dr, next_s;     // source / receiver stepped at this depth; source stepped at the next
    int conj_src, conj_rec; // x-pass: conj(L) for the -dz steps
    int mode;               // x-pass: 0 separation, 1 combination, 2 kz phases of the pair, 3 of the source,
                            // 4 combination at depth izc then separation at iz (one x-pass, the
                            // truncation to x < nmx between them)
    int izc;                // mode 4: depth of the combination
    float s1, s2;           // scales of the x-pass's first and second 2 x 2 maps: 1 / (nkx nky) on a
                            // state a y-pass wrote, 1 / nkx on one an x-pass wrote
    float dwf, dkxf;        // float dw and dkx, as ewem.c forms w and the separation's kx
    float inv_maxs;         // 1 / (max|wav| / sqrt(ntfft)) (ewem.c:146)
    const float* vp0;       // [nz] 1 / po_p (separation)
    const float* vs0;       // [nz] 1 / po_s
    const float2* spx;      // [nm][nf] R2C(ux) / sqrt(ntfft) (migration)
    const float2* spz;      // [nm][nf]
    float* ppart;           // [ngroups][nm] partial P image of one depth (migration)
    float* spart;           // [ngroups][nm] partial S image
    float2* smig;           // [nitem][nz][nm] stored source / max|wav| (modelling)
    const float* mppT;      // [nz][nm] images (modelling input)
    const float* mpsT;
    float2* dgx;            // [nm][nf] modelled spectra (modelling output)
    float2* dgz;
};
// which: 0 initial wavefields, 1 x-pass (e.mode), 2 y-pass (phase 0 migration imaging, 1 receiver
// truncation to the nmx x nmy grid, 2 modelling source, 3 modelling receiver injection, 4 modelling
// output).  ap / as: the P and S plans' 3-D arguments (f0, nitem, iz set by the caller).
int launch_ewem3d(int which, const March3DArgs& ap, const March3DArgs& as, const Ewem3DArgs& e, int F, int phase,
                  hipStream_t st);
// Array length of a 3-D line of n points: n (powers of two 128..1024), the next power of two (at
// least 128) for mixed-radix n up to 1024, Bluestein's 2 n - 1 rounded up for other n up to 512;
// 0 when unsupported.
int march3d_len(int n);
int march3d_kind(int n);
bool march3d_supported(int nkx, int nky);
// which: 0 init, 1 x-pass, 2 migration y-pass, 3 modelling y-pass (phase 0 source,
// 1 receiver, 2 output), 4 PSPI x-pass, 5 PSPI y-pass (phase 0 migration, 1 modelling source,
// 2 modelling receiver, 3 modelling output, 4 linear-mode modelling receiver, 5 its output),
// 6 linear-mode modelling receiver x-pass.  F: frequencies per y-pass workgroup.
int launch_march3d(int which, const March3DArgs& a, int F, int phase, hipStream_t st);

// C[M x Nc] = A[M x K] * B[K x Nc], row-major with leading dimensions lda, ldb, ldc.
int launch_sgemm(int M, int Nc, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
                 hipStream_t st);

// Image reduction: out[(s)][x][iz] (+)= sum_{shot in group, g} partial[shot][g][iz][x], Kahan,
// fixed order.  per_shot = 1 writes one image per shot (no stacking).
// Fourier bases from tab = [cos | sin] (n each) of 2 pi e / n and w^2 / sqrt(ntfft) (nf):
// badj, bsrc [nt][2 nf], bfwd [2 nf][nt] (f_op with the w^2 and 1/sqrt(ntfft) factors).
int launch_build_bases(int nt, int nf, int n, int ifmin, double rnt, const double* tab, float* badj, float* bsrc,
                       float* bfwd, hipStream_t st);
// Fixed device scratch of the allocation-free migration (swem_noalloc.hip): one array per device,
// kNoallocScratchBytes long, whose device address for `device` is returned in *ptr (taken once
// per device with hipGetSymbolAddress).
constexpr size_t kNoallocScratchBytes = size_t(128) << 20;
constexpr int kMaxDevices = 64;
int noalloc_scratch(int device, void** ptr, size_t* bytes);
// dst[0..n) = vals[0..n) on stream st through kernel arguments (kIntBlock values per launch).
constexpr int kIntBlock = 256;
int launch_fill_ints(int* dst, const int* vals, int n, hipStream_t st);
// Image of depth iz of a 3-D chunk: shot->out[x * ldo + iz] (+)= compensated sum over the chunk's
// frequency groups, adding when acc_chunk or shot->acc is set.
int launch_reduce_depth_shot(const float* partial, int ngroups, int n, const Shot3D* shot, int iz, int ldo,
                             int acc_chunk, hipStream_t st);
// *dst = v on stream st (the value travels as a kernel argument).
int launch_set_shot3d(Shot3D* dst, const Shot3D& v, hipStream_t st);
int launch_reduce_image(const float* partial, int nshot, int ngroups, int nz, int nmx, float* out, int accumulate,
                        int per_shot, hipStream_t st);
// Solver vector operations: m1 = a m1 + b m2; <x, y> in double into part[kDotBlocks] (part has
// vec_dot_scratch_doubles() entries).
int launch_vec_cgstep(float a, float* m1, float b, const float* m2, size_t n, hipStream_t st);
int launch_vec_mul(const float* x, const float* w, float* y, size_t n, hipStream_t st);
int launch_vec_dot(const float* x, const float* y, size_t n, double* part, hipStream_t st);
int vec_dot_scratch_doubles();
// One depth of an image: out[x * ldo] (+)= sum_g partial[g][x] (compensated, fixed order).
int launch_reduce_depth(const float* partial, int ngroups, int n, float* out, int ldo, int accumulate, hipStream_t st);
// [nm][nz] -> [nz][nm]
int launch_transpose(const float* in, int rows, int cols, float* out, hipStream_t st);

}  // namespace swem
