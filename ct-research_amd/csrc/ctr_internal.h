// ctr_internal.h - kernel parameter blocks and launcher declarations shared by
// the C ABI (ctr_api.cpp) and the HIP kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "ctr/ctr.h"

namespace ctr {

constexpr float kInf = __builtin_huge_valf();

// Ray-driven Siddon projector (cone; the fan beam is the nz = n_xi = 1 case).
// Box [0,bt] x [0,bs] x [0,bz] mm, voxel pitch (rt, rs, rz).
struct RayParams {
    int nt, ns, nz;
    int np, nx, nb;
    double rt, rs, rz;
    double bt, bs, bz;
    double D;
    double sz;              // source z (mm, volume frame)
    const double *view;     // [nb][4]: sin b, cos b, source t, source s
    const double *p;        // [np] column positions at the isocentre
    const double *xi;       // [nx] row positions at the isocentre
    const float *volT;      // optional [nz][nt][ns] copy of the volume, read by views whose
                            // central ray runs mostly along t (neighbouring rays then differ in s)
    int parallel;           // parallel beam: ray (view, p) is the line through centre + p (cos, sin)
                            // along (-sin, cos); view[4b+2..3] unused
    int residual_norm;      // the SIRT residual pass stores r times the ray norm (the cone
                            // adjoint's input form); the fan adjoint reads r itself
    int skip_empty;         // a slab's second and third voxels are read only by lanes whose
                            // segment in them is not empty (pays once the volume outgrows the
                            // Infinity Cache; costs instructions below that)
};

// Per-view constants of the voxel-driven backprojectors, float:
//   dect_t(i,j) = at*i + as*j + a0, dect_s(i,j) = bt*i + bs*j + b0
//   (i, j voxel indices; the voxel centre frame of Backprojection3D.cu:31-37),
//   wv = view weight (dbeta for FDK), st/ss = source position (volume frame).
struct BpView {
    float at, as, a0, bt, bs, b0, wv, pad0;
    float st, ss, sb, cb, pad1, pad2, pad3, pad4;
};

struct BpParams {
    int nt, ns, nz, np, nx, nb;
    int tiles_t, tiles_s, tiles_z;
    float rt, rs, rz;       // voxel pitch
    float zc0;              // (0.5 rz - Cz): z of slice-0 centre relative to the source plane
    float D;
    float u_scale, u_off;   // detector column coordinate u = P*u_scale + u_off (sample index units)
    float v_scale, v_off;   // detector row coordinate v = xi*v_scale + v_off
    float umax, vmax;       // valid iff 0 <= u < umax and 0 <= v < vmax
    float ht, hs, hz;       // half voxel sizes (chord of BP_REF)
    const BpView *views;
    const float *pd;        // [np] detector columns (float, device)
    const float *xd;        // [nx] detector rows (float, device)
    const float2 *zq;       // [nz + 32]: (zs^2, hz/|zs|) per slice, zs = slice centre relative to the source
    int accumulate;
    int quirks;
    int rows_in;            // FDK: tiles that leave the detector only in u take the clamped interior loop
                            // (0: the masked loop, for validation; CTR_BP_ROWSIN=0)
    // exact-adjoint tables (geometry only, built with the cone tables):
    const float *adj_inva;  // [nb][np] 1/(p cos b + D sin b), 0 where that is 0
    const float *adj_invb;  // [nb][np] 1/(p sin b - D cos b), 0 where that is 0
    const float *adj_invxi; // [nx] 1/xi, 0 at xi = 0
    const float *adj_norm;  // [nx][np] ray norm sqrt(p^2 + xi^2 + D^2); the adjoint reads its
                            // input multiplied by it (this table itself for A'1)
    const double *view_d;   // [nb][4] sin, cos, source t, source s in FP64 (the projector's rays)
    const double *pdd;      // [np] detector columns in FP64
    double rt_d, rs_d;      // voxel pitch in FP64
};

// The ray norm of detector sample (p, xi), in the float operations that the
// adjoint's input scaling uses everywhere (tables, residual, pre-pass).
__host__ __device__ inline float ray_norm(float p, float xi, float D)
{
    return sqrtf(fmaf(p, p, fmaf(xi, xi, D * D)));
}

enum BpKind { kBpFdk = 0, kBpRef = 1 };

// Exact adjoint: SIRT update fused into the epilogue when sirt != 0.
struct SirtUpdate {
    const float *vinv;      // [nvox] 1/(A'1 + 1e-9)
    float *x;               // [nvox] iterate, updated in place
    float lambda, lo, hi;
};

// Row filter (pre-weight + convolution along p).
struct FilterParams {
    int np, nx, rows;       // detector columns, rows per view, total rows
    int mode;               // 0: w = D/sqrt(D^2+p^2+xi^2); 1: w = colw[p]; 2: w = 1;
                            // 3: curved short scan, w = (D/P) cos(g P) Parker(b P, start P, g P, fan P)
                            //    with P = rowp[row % nx], b = betad[row / nx], g = gamd[p]
    float D;
    double Dd;              // mode 0: D in FP64 (the pre-weight is formed in FP64)
    const double *pw;       // mode 0: [nx][np] FP64 weights D / sqrt(D^2 + p^2 + xi^2) (launch_cone_weights),
                            //         or null: the same weights formed per sample
    const float *pd;        // [np] column positions
    const float *xd;        // [nx] row positions
    const float *colw;      // [np] column weights (mode 1)
    const double *taps;     // [2np-1] two-sided taps times the pitch: h2[k] = d*h[|k-(np-1)|]
    const double *spec;     // [fft_len] real DFT of the circular taps divided by fft_len, or null (direct path)
    int fft_len;
    int spec_stride;        // nonzero: spec holds one spectrum per detector row (row % nx) this many
                            // doubles apart, and the FFT path pairs rows of equal row index
    const double *rowp;     // mode 3: [nx] row scale Dsd / sqrt(Dsd^2 + xi^2) (FP64: the window's
                            //         sine argument (pi + P (fan - b)) / (P (fan/2 + g)) is
                            //         ill-conditioned in P near the end of the scan)
    const double *gamd;     // mode 3: [np] channel angles
    const double *betad;    // mode 3: [rows / nx] view angles
    double start, fan;      // mode 3: Parker window start and full fan angle
};

// Parker short-scan window (ParkFunction.m:6-16): region tests in FP64,
// the sine in FP32.  A ramp region of zero width (a channel at +-fan/2)
// gives weight 0.
__device__ __forceinline__ float parker_weight(double b, double st, double g, double fan)
{
    constexpr double pi = 3.141592653589793;
    if (b >= st && b < st + fan - 2.0 * g) {
        const double den = 0.5 * fan - g;
        if (!(den > 0.0)) return 0.0f;
        const float v = sinf((float)(0.25 * pi * b / den));
        return v * v;
    }
    if (b >= st + fan - 2.0 * g && b <= st + pi - 2.0 * g) return 1.0f;
    if (b > st + pi - 2.0 * g && b <= st + pi + fan) {
        const double den = 0.5 * fan + g;
        if (!(den > 0.0)) return 0.0f;
        const float v = sinf((float)(0.25 * pi * (pi + fan - b) / den));
        return v * v;
    }
    return 0.0f;
}

// Curved (cylindrical, equiangular) cone-beam FDK backprojection.
struct CurvedBpParams {
    int nt, ns, nz, ng, nx, nb;
    float rt, rs, rz, ct, cs, cz;
    float off;                    // voxel centre (i + off) r - c: 0.5, or -0.5 for the reference's offset
    float D, Dsd;
    float g_scale, g_off, gmax;   // channel coordinate gamma*g_scale + g_off, valid in [0, gmax)
    float x_scale, x_off, xmax;   // row coordinate xi*x_scale + x_off, valid in [0, xmax)
    float dbeta;
    const float *view;            // [nb][2] sin, cos
    int accumulate;
    // tiles of 16 x 16 columns x tz slices and their per-view detector
    // footprints (curved_plan, launch_curved_footprints)
    int tz, tiles_t, tiles_s, tiles_z;
    const int4 *foot;             // [tiles][nb]: (channel0, row0, W | H << 16, flags)
};
// Tile plan of the curved backprojection: fills P.tz and P.tiles_*; returns
// the number of footprint entries (tiles x views).
size_t curved_plan(CurvedBpParams &P);
hipError_t launch_curved_footprints(const CurvedBpParams &P, int4 *foot, hipStream_t st);
hipError_t launch_curved_bp(const CurvedBpParams &P, const float *filt, float *vol, hipStream_t st);

// Fan-beam pixel-driven kernels.
struct FanParams {
    int nt, ns, np, nb;
    float rt, rs;
    float ct, cs;           // rotation centre
    float D;
    float u_scale, u_off, umax;
    float dbeta;
    float g0, g_scale, gmax;   // equiangular: gamma index = (gamma - g0)*g_scale, valid in [0,gmax)
    double rt_d, rs_d, ct_d, cs_d, u_scale_d, u_off_d;   // FP64 copies for the reference-quirk path
    const double *view;     // [nb][4]: sin, cos, source t, source s
    const float *pf;        // [np] bin positions (float)
    const double *pdbl;     // [np] bin positions (FP64, as the forward projector's rays)
    const float *gam;       // [np] bin angles (float, equiangular)
    int accumulate;
    int quirks;
    int equiangular;
};

// ---- launchers (return hipError_t of the launch) --------------------------
hipError_t launch_ray_fp(const RayParams &g, const float *vol, float *proj, hipStream_t st);
// Siddon system matrix: counts[ray] = entries of the ray's row; then the rows
// at row_ptr[ray] (column = voxel index, value = segment length).
hipError_t launch_sysmat_count(const RayParams &g, long long *counts, hipStream_t st);
hipError_t launch_sysmat_fill(const RayParams &g, const long long *row_ptr, int *col, float *val, hipStream_t st);
// Inclusive prefix sum of n counts (hipCUB); tmp == nullptr queries tmp_bytes.
hipError_t scan_counts(const long long *counts, long long *out, int n, void *tmp, size_t &tmp_bytes, hipStream_t st);
// SIRT residual: r = (b - A x) / (rowlen + 1e-9); partials[block] = (sum e^2, sum |e|)
hipError_t launch_ray_fp_residual(const RayParams &g, const float *vol, const float *b, float *r,
                                  double *partials, int *nblocks, hipStream_t st);
int ray_fp_blocks(const RayParams &g);

// Tiling of the voxel-driven backprojection and its per-(tile, view) detector
// footprint table (int4 per entry, nfoot entries), computed once per geometry.
// Flags of a (tile, view) footprint (the .w of its int4).
enum { kFootEmpty = 1, kFootInterior = 2, kFootFits = 4, kFootRowsIn = 8 };

struct BpPlan {
    int faces;              // footprints of voxel boxes (exact adjoint) instead of voxel centres
    int tz, wp, capf, hmax;
    int cm;                 // exact adjoint: candidate detector columns per chunk
    int fit_all;            // exact adjoint: every footprint fits LDS (set once the table is built)
    int tiles_t, tiles_s, tiles_z;
    size_t nfoot;
};
BpPlan bp_plan(const BpParams &P, int pitch);        // pitch: 64, 72 or 80 floats per staged row
BpPlan adjoint_plan(const BpParams &P, int cm);      // cm: 3, 4 or 6 candidate columns per chunk
hipError_t launch_bp_footprints(const BpParams &P, const BpPlan &L, int4 *foot, hipStream_t st);
hipError_t launch_bp_tiled(int kind, const BpParams &P, const BpPlan &L, const int4 *foot, const float *proj,
                           float *vol, hipStream_t st);
// proj is the adjoint's input already multiplied by the ray norm (P.adj_norm):
// the SIRT residual kernels write it so, other callers scale with
// launch_scale_by_norm first.  ones_input computes A'1 from P.adj_norm.
hipError_t launch_adjoint(const BpParams &P, const BpPlan &L, const int4 *foot, const float *proj, float *vol,
                          const SirtUpdate *upd, int ones_input, hipStream_t st);
// Fills P.adj_inva, adj_invb, adj_invxi and adj_norm (the pointers name the
// destinations; P's views, pd and xd must be on the device).
hipError_t launch_adjoint_tables(const BpParams &P, hipStream_t st);
// out[0 .. kFootCountBlocks) sum to the number of footprints that are neither
// empty nor fit LDS.
constexpr int kFootCountBlocks = 256;
hipError_t launch_foot_unfit(const int4 *foot, size_t n, int *out, hipStream_t st);
// out[b][r][c] = in[b][r][c] * P.adj_norm[r][c] over all views.
hipError_t launch_scale_by_norm(const BpParams &P, const float *in, float *out, hipStream_t st);

// Batched exact adjoint (independent problems in one launch): job
// blockIdx.y, its tile blockIdx.x (blocks past the job's tiles leave at
// once).  P.tiles_* are the problem's plan; jobs live in device memory.
struct AdjointJob {
    BpParams P;
    const int4 *foot;
    const float *proj;      // weighted residual times the ray norm (ignored when ones_input)
    float *vol;             // output (when not a SIRT update)
    SirtUpdate U;
    int block0, nblocks;    // the job's first block in the batch's block count, its tiles
};
// fit_all: every job's plan has fit_all set (the variant without the
// global-memory path for footprints too large for LDS); max_blocks: the
// largest job's tile count.
hipError_t launch_adjoint_batched(int wp, int cm, bool fit_all, const AdjointJob *jobs, int njobs, int max_blocks,
                                  int sirt, int ones_input, hipStream_t st);

// Batched SIRT residual pass: r = (b - A x)/(A1 + 1e-9) and per-block (sum e^2, sum |e|) per job.
struct FpJob {
    RayParams g;
    const float *vol;
    const float *b;
    float *r;
    double *partials;
    int block0, gx, gy;     // first block and the job's (columns, rows) block grid; views = z
};
int fp_job_blocks(const RayParams &g, int *gx, int *gy);
// job blockIdx.y, its block blockIdx.x; max_blocks: the largest job's block count
hipError_t launch_ray_fp_residual_batched(const FpJob *jobs, int njobs, int max_blocks, hipStream_t st);
// sums[(*iter) * njobs + p] = sum of job p's partial pairs (one block per job)
struct ReduceJob {
    const double *partials;
    int n;
};
hipError_t launch_reduce_jobs(const ReduceJob *jobs, int njobs, const int *iter, double *sums, hipStream_t st);
hipError_t launch_iter_inc(int *iter, hipStream_t st);

hipError_t launch_row_filter(const FilterParams &P, const float *in, float *out, hipStream_t st);
// pw[x][q] = D / sqrt(D^2 + pd[q]^2 + xd[x]^2) in FP64 (the mode-0 weight table).
hipError_t launch_cone_weights(double D, const float *pd, const float *xd, int np, int nx, double *pw, hipStream_t st);
// transform length of the FFT filter path for np columns (0: direct convolution)
int fft_filter_length(int np);
// Restricted Chebyshev emulation of SIRT: mask = v > thr; out = a * m; the
// masked plain step w0 = mask lambda V g with acc = c0 w0; one recurrence
// term w_{j+1} = mask (2 (2M - I) w_j - w_{j-1}) (first: mask (2M - I) w_0)
// from g = A'(-W A w_j), written over w_{j-1}, with acc += c w_{j+1};
// out = clamp(a + b).
hipError_t launch_threshold_mask(const float *v, float thr, size_t n, float *mask, hipStream_t st);
hipError_t launch_mul(const float *a, const float *m, size_t n, float *out, hipStream_t st);
hipError_t launch_cheb_start(const float *g, const float *vinv, const float *mask, float lambda, float c0, size_t n,
                             float *w0, float *acc, hipStream_t st);
hipError_t launch_cheb_step(const float *g, const float *vinv, const float *mask, const float *cur, float *prev,
                            float *acc, float c, int first, size_t n, hipStream_t st);
hipError_t launch_add_clamp(const float *a, const float *b, float lo, float hi, size_t n, float *out, hipStream_t st);
// Copy a wx x wy x wz box from src (row sx, plane sx*sy) at (sx0, sy0, sz0) to
// dst (row dx, plane dx*dy) at (dx0, dy0, dz0).
hipError_t launch_box_copy(const float *src, int sx, int sy, int sx0, int sy0, int sz0, float *dst, int dx, int dy,
                           int dx0, int dy0, int dz0, int wx, int wy, int wz, hipStream_t st);

// Allocation-free ramp filter: the library's fixed device array (one per
// GPU, file-scope in k_filter.hip) holds, per call, the detector column and
// row positions, the two-sided taps and the FFT spectrum.
constexpr size_t kNoAllocScratchBytes = 128 * 1024;
constexpr int kNoAllocMaxCols = 2048;     // taps: 2 * 2048 - 1 doubles
constexpr int kNoAllocMaxRows = 8192;
constexpr size_t kNoAllocOffXf = 8192;            // [kNoAllocMaxRows] float row positions
constexpr size_t kNoAllocOffTaps = 40960;         // [2 * kNoAllocMaxCols - 1] double taps
constexpr size_t kNoAllocOffSpec = 73728;         // [<= 2048] double spectrum
static_assert(kNoAllocOffXf >= sizeof(float) * kNoAllocMaxCols, "scratch layout");
static_assert(kNoAllocOffTaps - kNoAllocOffXf >= sizeof(float) * kNoAllocMaxRows, "scratch layout");
static_assert(kNoAllocOffSpec - kNoAllocOffTaps >= sizeof(double) * (2 * kNoAllocMaxCols - 1), "scratch layout");
static_assert(kNoAllocOffSpec + sizeof(double) * 2048 <= kNoAllocScratchBytes, "scratch layout");
// Device address of the array on `device` (which must be current), looked up
// once per device and kept.
hipError_t noalloc_scratch(int device, void **ptr);
// Taps h2[k] = d h[|k - (n-1)|] of the half SL / RL flat-detector filter and,
// when L > 0, the spectrum of its circular kernel divided by L.
hipError_t launch_filter_tables(int filter, double d, int n, int L, double *taps, double *spec, hipStream_t st);

hipError_t launch_fan_bp_ref(const FanParams &P, const float *proj, float *img, hipStream_t st);
hipError_t launch_fan_adjoint(const FanParams &P, const float *proj, float *img, const SirtUpdate *upd,
                              int ones_input, hipStream_t st);
hipError_t launch_fan_fbp_bp(const FanParams &P, const float *filt, float *img, hipStream_t st);

// Parallel-beam pixel-driven kernels: the FBP backprojection (t = x cos + y sin
// per pixel and view, linear interpolation on the uniform detector grid, times
// dtheta) and the exact adjoint of the Siddon parallel projector.
struct ParBpParams {
    int nt, ns, np, nb;
    float rt, rs, ct, cs;
    double rt_d, rs_d, ct_d, cs_d;
    float u_scale, u_off, umax;   // u = t*u_scale + u_off in sample units, valid in [0, umax)
    float dtheta;
    const float *view;            // FBP: [nb][2] sin, cos of the single-precision view angles
    const double *view_d;         // adjoint: [nb][4] sin, cos (the projector's view table)
    const double *t_d;            // adjoint: [np] ray offsets (the projector's)
    int accumulate;
};
hipError_t launch_par_fbp_bp(const ParBpParams &P, const float *filt, float *img, hipStream_t st);
hipError_t launch_par_adjoint(const ParBpParams &P, const float *proj, float *img, const SirtUpdate *U,
                              int ones_input, hipStream_t st);

// Iterative-FDK error slice (IterativeFDKerrCompute.cu).
struct FdkErrParams {
    int nt, ns, nz, nb, nu, np, ng, t;
    double r0, r1, r2, ct, cs, cz, D, pint;
    const float *cosb, *sinb;   // [nb] float(cos(double beta)), float(sin(...))
    const float *u;             // [nu] integration rows U
    const float *tab_t;         // [nu][np] float(P (D - U) / D)
    const double *p;            // [np] detector samples P
    const double *g;            // [ng] ramp taps
    const double *cp;           // [np] sqrt(1 + (P/D)^2) D / sqrt(D^2 + P^2) dU PInt dbeta
};
// xy: scratch of launch_fdk_error_slice_scratch(P) bytes: the voxel map
// [nb][nu][np] (int), the per-(view, U) sample intervals (int2), then a copy
// of the volume with z fastest (float [ns][nt][nz])
hipError_t launch_fdk_error_slice(const FdkErrParams &P, const float *pic, int *xy, float *out, hipStream_t st);
constexpr int kFdkErrRowPad = 8;   // ints readable past the intervals (the kernel's unrolled map reads)
inline size_t fdk_error_slice_zcopy_offset(const FdkErrParams &P)   // in ints, a multiple of 64 (256 bytes)
{
    const size_t n = (size_t)P.nb * P.nu * P.np;
    return (((n + 1) & ~(size_t)1) + 2 * (size_t)P.nb * P.nu + kFdkErrRowPad + 63) & ~(size_t)63;
}
inline size_t launch_fdk_error_slice_scratch(const FdkErrParams &P)
{
    return (fdk_error_slice_zcopy_offset(P) + (size_t)P.nt * P.ns * P.nz) * sizeof(int);
}

// Ellipsoid of the phantom rasterizer: amplitude, semi-axes, rotated centre
// (sx, sy) = R(phi) (x0, y0), z0, cos/sin of the rotation about z.
struct PhantomEll {
    double A, a, b, c, sx, sy, z0, cp, sp;
};
hipError_t launch_phantom(const PhantomEll *ell, int n_ell, int nt, int ns, int nz, float *vol, hipStream_t st);
// per-block (sum (x - ref)^2, sum ref^2)
hipError_t launch_diff_sums(const float *x, const float *ref, size_t n, double *partials, int *nblocks, hipStream_t st);

// Utilities.
hipError_t launch_fill(float *dst, float value, size_t n, hipStream_t st);
hipError_t launch_transpose_st(const float *in, float *out, int nt, int ns, int nz, hipStream_t st);   // [z][t][s] <- [z][s][t]
hipError_t launch_reciprocal_eps(float *v, size_t n, float eps, hipStream_t st);   // v <- 1/(v+eps)
// Deterministic sum of n (a, b) pairs into out2[0..1].  Above kReduceDirect
// pairs it runs in two stages and writes up to kReduceSlack stage pairs after
// the n pairs: a partial buffer holds 2 (n + kReduceSlack) doubles.
constexpr int kReduceDirect = 4096, kReduceSlack = 256;
hipError_t launch_reduce_partials(double *partials, int n, double *out2, hipStream_t st);
hipError_t launch_sum_sq_sum(const float *x, size_t n, double *partials, int *nblocks, hipStream_t st);  // (sum x^2, sum x)

}  // namespace ctr
