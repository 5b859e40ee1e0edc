/*
 * ctr.h - C ABI of the MI355X (gfx950) CT projector library.
 *
 * The library provides the cone-beam and fan-beam projector pair of
 * CT-Research (Siddon ray-driven forward projection, voxel-driven
 * backprojection), the FDK / fan-beam FBP ramp filter and backprojection,
 * the exact matched adjoint of the Siddon projector and the SIRT loop
 * built on the pair.  Every computation runs in hand-written HIP kernels
 * on one GPU; there is no CPU fallback.
 *
 * Conventions shared with the reference MEX operators
 *   - volume axes (t, s, z), t fastest: index k*Nt*Ns + j*Nt + i
 *   - 2-D image: index j*Nt + i
 *   - the volume occupies [0,size[0]] x [0,size[1]] x [0,size[2]] mm and
 *     rotates about its centre size/2
 *   - source S(beta) = (Ct - D sin(beta), Cs + D cos(beta), Cz), beta in radians
 *   - virtual flat detector through the isocentre: ray (p, xi) runs from S
 *     through C + p (cos beta, sin beta, 0) + xi (0, 0, 1)
 *   - projections: 3-D beta*(n_xi*n_p) + xi*n_p + p, 2-D beta*n_p + p
 *
 * Pointer convention
 *   Geometry arrays (beta, p, xi) are always host memory.  Data pointers
 *   (volumes, images, projections) are host memory unless CTR_DEVICE_PTRS
 *   is set.  Host-pointer calls are synchronous.  Device-pointer calls are
 *   ordered on the context's stream and return before the GPU finishes;
 *   call ctr_ctx_synchronize() before reading the results from the host.
 *
 * Errors
 *   Every entry point returns a ctr_status.  No entry point aborts.  The
 *   message of the last failure on a context is ctr_last_error(ctx); for a
 *   failure with no usable context it is ctr_last_error(NULL) (per thread).
 *
 * Threading
 *   A context is used by one host thread at a time.  Use one context per
 *   thread for concurrent callers.
 */
#ifndef CTR_CTR_H_
#define CTR_CTR_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CTR_VERSION_STRING "ctr 0.1.0 (gfx950)"

typedef enum ctr_status {
    CTR_OK = 0,
    CTR_EINVAL = 1,          /* invalid argument (null pointer, bad size, bad geometry) */
    CTR_ENOMEM = 2,          /* device or pinned-host allocation failed */
    CTR_EHIP = 3,            /* a HIP runtime call or a kernel launch failed */
    CTR_ETOOBIG = 4,         /* problem exceeds an index or workspace limit */
    CTR_EFAULT_INJECTED = 5, /* failure forced by the CTR_FAULT environment hook */
    CTR_ENODEV = 6           /* no usable HIP device */
} ctr_status;

/* Call flags (bitwise OR). */
#define CTR_DEVICE_PTRS 0x1u /* data pointers are device memory, call is stream-ordered */
#define CTR_ACCUMULATE 0x2u  /* backprojections add into the output instead of overwriting it */
#define CTR_REF_QUIRKS 0x4u  /* reproduce the reference's documented arithmetic quirks
                                (fan BP entry-face test, FDK stale sample above the detector) */
#define CTR_AUTOTUNE 0x8u    /* cone beam: run ctr_autotune_cone for this geometry on its first use */

/* Ramp filters (half-scaled: the 1/2 redundancy weight of a 360-degree scan is folded in). */
#define CTR_FILTER_SL_HALF 0 /* Shepp-Logan  h[n] = -1/(pi^2 d^2 (4n^2-1)) */
#define CTR_FILTER_RL_HALF 1 /* Ram-Lak      h[0] = 1/(8 d^2), h[odd n] = -1/(2 pi^2 n^2 d^2), h[even n] = 0 */

/* Fan-beam detector kinds for filtered backprojection. */
#define CTR_FAN_FLAT 0        /* p[] are flat-detector positions at the isocentre */
#define CTR_FAN_EQUIANGULAR 1 /* p[] are D*tan(gamma_i) with gamma_i uniformly spaced */

/* Cone-beam circular-orbit geometry (the argument set of ProjectionCone_3D / Backprojection / FDK). */
typedef struct ctr_cone_geom {
    int nt, ns, nz;         /* voxels along t, s, z */
    double size[3];         /* physical extent of the volume (mm) */
    double distance;        /* source-to-isocentre distance D (mm) */
    int n_beta, n_p, n_xi;  /* views, detector columns, detector rows */
    const float *beta;      /* [n_beta] view angles (rad) */
    const float *p;         /* [n_p] column positions at the isocentre (mm), monotonic */
    const float *xi;        /* [n_xi] row positions at the isocentre (mm), monotonic */
} ctr_cone_geom;

/* Fan-beam geometry (the argument set of ProjectionFan_2D / BackprojectionFan2D / FBPfan). */
typedef struct ctr_fan_geom {
    int nt, ns;             /* pixels along t, s */
    double size[2];         /* physical extent of the image (mm) */
    double distance;        /* source-to-isocentre distance (mm) */
    int n_beta, n_p;        /* views, detector bins */
    const double *beta;     /* [n_beta] view angles (rad) */
    const double *p;        /* [n_p] bin positions at the isocentre (mm), monotonic */
} ctr_fan_geom;

/* Parallel-beam geometry (the argument set of ProjectionParallel_2D / FBPparallel).
 * Ray (theta, t) is the line (x - size[0]/2) cos(theta) + (y - size[1]/2) sin(theta) = t,
 * x along t pixels, y along s pixels (ProjectionParallel_2D.cu:22-28). */
typedef struct ctr_par_geom {
    int nt, ns;             /* pixels along t (x) and s (y) */
    double size[2];         /* physical extent of the image along t and s (mm) */
    int n_theta, n_t;       /* views, detector bins */
    const double *theta;    /* [n_theta] angles of the ray normals (rad) */
    const double *t;        /* [n_t] signed ray offsets from the image centre (mm), monotonic */
} ctr_par_geom;

/* Curved (cylindrical, equiangular) cone-beam short-scan geometry (the argument
 * set of ImplementUI: My_toolbox/Implement_UnitedImaging_Data/ImplementUI.cpp:7-24).
 * Projections [n_beta][n_xi][n_gamma], gamma fastest. */
typedef struct ctr_curved_geom {
    int nt, ns, nz;         /* voxels along t, s, z */
    double size[3];         /* physical extent of the volume (mm) */
    double distance;        /* source-to-isocentre distance D (mm) */
    double distance_s2d;    /* source-to-detector distance (mm): radius of the detector cylinder */
    int n_beta, n_gamma, n_xi;
    const double *beta;     /* [n_beta] view angles (rad), uniform */
    const double *gamma;    /* [n_gamma] channel angles (rad), uniform, monotonic */
    const double *xi;       /* [n_xi] row positions on the detector (mm), uniform, monotonic */
    double beta_start;      /* short-scan start angle of the Parker window (rad) */
    double fan_angle;       /* full fan angle of the Parker window (rad) */
} ctr_curved_geom;

/* Arguments of the iterative-FDK error slice (the argument set of
 * IterativeFDKerrCompute: My_novel_algorithm/IterativeFDK/IterativeFDKerrCompute.cpp:5-28). */
typedef struct ctr_fdkerr_geom {
    int nt, ns, nz;         /* voxels along t, s, z */
    double size[3];         /* physical extent of the volume (mm) */
    double distance;        /* source-to-isocentre distance (mm) */
    int n_beta, n_u, n_p, n_g;
    const float *beta;      /* [n_beta] view angles (rad), uniform */
    const float *u;         /* [n_u] integration rows U (mm) along the source axis, uniform */
    const double *p;        /* [n_p] detector samples (mm), uniform */
    const double *g;        /* [n_g] ramp taps, centre at index n_p - 1 (IterativeFDK.m:303-312); n_g <= 32767 */
    int t;                  /* voxel slice along t, 0-based */
} ctr_fdkerr_geom;

/* SIRT options (Iterative_algorithm/SIRT3D.m, SIRT.m). */
typedef struct ctr_sirt_opts {
    int n_iter;             /* iterations (SIRT3D.m "Times") */
    float lambda;           /* relaxation factor (SIRT3D.m "Lamda") */
    float min_val;          /* lower clamp of every iterate (SIRT3D.m: 0) */
    float max_val;          /* upper clamp (SIRT.m: 1); +INFINITY for none */
    int use_x0;             /* nonzero: x holds the starting iterate, else start from 0 */
    double *residual_l2;    /* optional host [n_iter+1]: ||b - A x_k||_2 / ||b||_2, k = 0..n_iter */
    double *residual_l1;    /* optional host [n_iter+1]: sum|b - A x_k| / (sum(b) * n_voxels) */
} ctr_sirt_opts;

typedef struct ctr_ctx ctr_ctx;

/* ---- library and context --------------------------------------------- */

const char *ctr_version(void);
const char *ctr_status_string(int status);
const char *ctr_last_error(const ctr_ctx *ctx);

/* Create a context on HIP device `device`: a stream, cached device workspaces
 * and geometry tables.  Replaces the per-call cudaSetDevice(0) / cudaMalloc /
 * cudaDeviceReset of every reference launcher (e.g.
 * GPU_accelerated_Operator/ProjectionCone_3D/ProjectionCone_3D.cu:433-519,
 * ProjectionCone_3D.cpp:53). */
ctr_status ctr_ctx_create(int device, ctr_ctx **out);
void ctr_ctx_destroy(ctr_ctx *ctx);
/* Order the context's work on hip_stream (a hipStream_t; NULL is HIP's null
 * stream, e.g. torch's default stream).  The context's own stream is the
 * value ctr_ctx_stream() returns right after creation; pass it back to
 * restore it. */
ctr_status ctr_ctx_set_stream(ctr_ctx *ctx, void *hip_stream);
void *ctr_ctx_stream(const ctr_ctx *ctx);
ctr_status ctr_ctx_synchronize(ctr_ctx *ctx);

/* Device memory helpers for callers without another allocator. */
ctr_status ctr_device_alloc(ctr_ctx *ctx, size_t bytes, void **out);
ctr_status ctr_device_free(ctr_ctx *ctx, void *ptr);
ctr_status ctr_copy_to_device(ctr_ctx *ctx, void *dst, const void *src, size_t bytes);
ctr_status ctr_copy_to_host(ctr_ctx *ctx, void *dst, const void *src, size_t bytes);

/* Argument checks only (no device work). */
ctr_status ctr_cone_geom_check(const ctr_cone_geom *g, int need_uniform_pitch);
ctr_status ctr_fan_geom_check(const ctr_fan_geom *g, int need_uniform_pitch);
ctr_status ctr_par_geom_check(const ctr_par_geom *g, int need_uniform_pitch);
ctr_status ctr_curved_geom_check(const ctr_curved_geom *g);

/* ---- cone beam --------------------------------------------------------- */

/* Siddon line integrals of vol [nz*ns*nt] into proj [n_beta*n_xi*n_p].
 * The volume must be finite: a non-finite voxel makes the rays through it
 * non-finite, and a ray that only grazes it (zero segment length) may or
 * may not be, depending on the voxel-read path (masked above 256 MiB of
 * volume); rays away from it are unaffected.
 * Replaces ProjectionCone_3D() (GPU_accelerated_Operator/ProjectionCone_3D/ProjectionCone_3D.h:32-35,
 * kernel ProjectionCone_3D.cu:8-379). */
ctr_status ctr_cone_fp(ctr_ctx *ctx, const ctr_cone_geom *g, const float *vol, float *proj, unsigned flags);

/* Voxel-driven backprojection: bilinear detector sample x chord of the
 * source-to-voxel-centre line through the voxel.  Replaces BackPro()
 * (GPU_accelerated_Operator/BackprojectionCone_3D/Backprojection3D.h:31-33, kernel Backprojection3D.cu:8-291). */
ctr_status ctr_cone_bp_ref(ctr_ctx *ctx, const ctr_cone_geom *g, const float *proj, float *vol, unsigned flags);

/* Exact transpose of ctr_cone_fp (voxel-driven gather of Siddon segment
 * lengths).  No reference counterpart: SIRT3D.m:81 uses SysMatrix' of the
 * same Siddon system matrix. */
ctr_status ctr_cone_bp_adjoint(ctr_ctx *ctx, const ctr_cone_geom *g, const float *proj, float *vol, unsigned flags);

/* FDK pre-weight D/sqrt(D^2+p^2+xi^2) and ramp filtering along p, scaled by
 * the column pitch.  Replaces GFunction + PreWeightFilter
 * (Analytic_Algorithm/FBP_type_analytic_algorithm_GPU_cpp_sourcecode/FDK_GPU/FDK.cu:6-48). */
ctr_status ctr_cone_filter(ctr_ctx *ctx, const ctr_cone_geom *g, int filter, const float *proj, float *filtered,
                           unsigned flags);

/* ctr_cone_filter without any allocation, for callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Data pointers must be device memory
 * (CTR_DEVICE_PTRS is required).  The detector positions, the taps and the
 * FFT spectrum go to a fixed 128 KiB device array of the library, one per
 * GPU, held until the process ends: at most 2048 detector columns and 8192
 * rows (CTR_ETOOBIG beyond).  Every call on a GPU rewrites that array, so
 * calls that overlap on one GPU (other streams, threads or contexts) race.
 * The outputs agree with ctr_cone_filter's to FP64 rounding of the spectrum. */
ctr_status ctr_cone_filter_noalloc(ctr_ctx *ctx, const ctr_cone_geom *g, int filter, const float *proj,
                                   float *filtered, unsigned flags);

/* FDK backprojection of filtered projections: bilinear sample x mag^2 x dbeta.
 * Replaces the BackProjection kernel of FDK_GPU/FDK.cu:50-126. */
ctr_status ctr_fdk_bp(ctr_ctx *ctx, const ctr_cone_geom *g, const float *filtered, float *vol, unsigned flags);

/* Filter + backprojection.  Replaces FDKpro() (FDK_GPU/FDK.h:29-31). */
ctr_status ctr_fdk(ctr_ctx *ctx, const ctr_cone_geom *g, int filter, const float *proj, float *vol, unsigned flags);

/* ---- fan beam ---------------------------------------------------------- */

/* Replaces ProjectionFan_2D() (GPU_accelerated_Operator/ProjectionFan_2D/ProjectionFan_2D.h:32-34). */
ctr_status ctr_fan_fp(ctr_ctx *ctx, const ctr_fan_geom *g, const float *img, float *proj, unsigned flags);

/* Linear detector sample x pixel chord.  Replaces FBPfan() of
 * GPU_accelerated_Operator/BackprojectionFan_2D/BackprojectionFan2D.cu:119 (kernel :7-116). */
ctr_status ctr_fan_bp_ref(ctr_ctx *ctx, const ctr_fan_geom *g, const float *proj, float *img, unsigned flags);

/* Exact transpose of ctr_fan_fp, for any monotone detector (uniform or not,
 * e.g. equiangular p = D tan(gamma)). */
ctr_status ctr_fan_bp_adjoint(ctr_ctx *ctx, const ctr_fan_geom *g, const float *proj, float *img, unsigned flags);

/* Fan pre-weight and ramp filter (flat: FBP_fan_GPU/FBPfan.cu:21-43; equiangular:
 * FBP_type_analytic_algorithm/FanProjection.m:229-293). */
ctr_status ctr_fan_filter(ctr_ctx *ctx, const ctr_fan_geom *g, int detector, int filter, const float *proj,
                          float *filtered, unsigned flags);

/* Filtered backprojection.  Flat replaces FBPfan() (FBP_fan_GPU/FBPfan.h:31-32);
 * equiangular follows FanProjection.m:295-324 with weight 1/2 for a full scan. */
ctr_status ctr_fan_fbp(ctr_ctx *ctx, const ctr_fan_geom *g, int detector, int filter, const float *proj, float *img,
                       unsigned flags);

/* ---- parallel beam ----------------------------------------------------- */

/* Siddon line integrals of img [ns*nt] into proj [n_theta*n_t].
 * Replaces ProjectionParallel_2D() (GPU_accelerated_Operator/ProjectionParallel_2D/ProjectionParallel_2D.h:33,
 * kernel ProjectionParallel_2D.cu:9-253). */
ctr_status ctr_par_fp(ctr_ctx *ctx, const ctr_par_geom *g, const float *img, float *proj, unsigned flags);

/* Exact transpose of ctr_par_fp.  No reference counterpart: ART.m and
 * SIRT.m use the transpose of the same Siddon system matrix. */
ctr_status ctr_par_bp_adjoint(ctr_ctx *ctx, const ctr_par_geom *g, const float *proj, float *img, unsigned flags);

/* Ramp filter along t (no pre-weight), scaled by the bin pitch of the
 * single-precision t positions.  Replaces GFunction + Filter
 * (Analytic_Algorithm/FBP_type_analytic_algorithm_GPU_cpp_sourcecode/FBP_parallel_GPU/FBPparallel.cu:6-39). */
ctr_status ctr_par_filter(ctr_ctx *ctx, const ctr_par_geom *g, int filter, const float *proj, float *filtered,
                          unsigned flags);

/* Filtered backprojection: linear detector interpolation x dtheta with theta
 * and t rounded to single precision, as FBPparallel receives them.
 * Replaces FBPpara() (FBP_parallel_GPU/FBPparallel.h:31, kernel FBPparallel.cu:41-89). */
ctr_status ctr_par_fbp(ctr_ctx *ctx, const ctr_par_geom *g, int filter, const float *proj, float *img,
                       unsigned flags);

/* ---- curved detector, short scan --------------------------------------- */

/* Parker-weighted pre-weight (D/P) cos(gamma P) and per-row equiangular
 * Shepp-Logan filter (full scale), P = Dsd / sqrt(Dsd^2 + xi^2).  Replaces
 * GFunction + PreWeightFiltration (ImplementUI.cu:6-78).  At most 1024
 * channels. */
ctr_status ctr_curved_filter(ctr_ctx *ctx, const ctr_curved_geom *g, const float *proj, float *filtered,
                             unsigned flags);

/* Backprojection of filtered projections: bilinear sample at
 * (atan(u/(D-s)), Dsd z/rho) times dbeta rho / L2^1.5.  Replaces the
 * BackProjection kernel (ImplementUI.cu:80-153).  CTR_REF_QUIRKS places
 * voxel centres at (i - 0.5) pitch instead of (i + 0.5) (ImplementUI.cu:110). */
ctr_status ctr_curved_bp(ctr_ctx *ctx, const ctr_curved_geom *g, const float *filtered, float *vol, unsigned flags);

/* Filter + backprojection.  Replaces FDKUI() (ImplementUI.h:31-33). */
ctr_status ctr_curved_fdk(ctr_ctx *ctx, const ctr_curved_geom *g, const float *proj, float *vol, unsigned flags);

/* ---- iterative FDK ------------------------------------------------------- */

/* Error slice of the iterative FDK correction for voxel slice t of vol
 * [nz*ns*nt]: err [n_beta*nz*ns] (s fastest), summed over U and P as
 * IterativeFDK.m:330-333 then sums it over the views.  Replaces
 * IterativeFDKerrCompute() (IterativeFDKerrCompute.cu:77-332, kernel :12-75).
 * Accumulates in FP64 (the reference in FP32). */
ctr_status ctr_fdk_error_slice(ctr_ctx *ctx, const ctr_fdkerr_geom *a, const float *vol, float *err, unsigned flags);

/* ---- phantoms and metrics ------------------------------------------------ */

/* Sum of ellipsoid indicators (My_toolbox/My_phantom/phantom3d.m:71-110) at the
 * voxel centres of a grid spanning [-1, 1] on each axis (end points
 * included), written as vol [nz*ns*nt]: t is phantom3d's y and s its x (the
 * layout a plain reshape of its array gives, FDK_GPU.m:74).  ell [n_ell][8]:
 * A, a, b, c, x0, y0, z0, phi (degrees), as phantom3d.m's tables. */
ctr_status ctr_phantom_ellipsoids(ctr_ctx *ctx, const double *ell, int n_ell, int nt, int ns, int nz, float *vol,
                                  unsigned flags);

/* rRMSE (My_toolbox/Evalution_parameter/rRMSE.m:2):
 * sqrt(mean((x-ref)^2)) / (sqrt(mean(ref^2)) + eps), FP64 on the device. */
ctr_status ctr_rrmse(ctr_ctx *ctx, const float *x, const float *ref, size_t n, double *out, unsigned flags);

/* LocalError (My_toolbox/Evalution_parameter/LocalError.m:2): ||cur - prev|| / ||prev||. */
ctr_status ctr_local_error(ctr_ctx *ctx, const float *cur, const float *prev, size_t n, double *out, unsigned flags);

/* ---- Siddon system matrix ------------------------------------------------
 * The matrix A of the Siddon projector in CSR form: row r is ray r in the
 * projection layout, column j is voxel j in the volume layout, the value is
 * the ray's segment length in the voxel (mm); A x equals the forward
 * projection.  Each row lists its voxels in the order the ray visits them,
 * without repeats.  Replaces the MATLAB builders GenSysMatCircular3D.m:8-100,
 * GenSysMatFan.m:6-85 and GenSysMatParal.m (the SysMatrix of SIRT.m:79,
 * MLOSEM.m:30).
 *
 * row_ptr: [n_rays + 1] int64; col: int32 voxel indices; val: float.
 * Call with col = val = NULL to fill row_ptr and *nnz.  Then call with col
 * and val of capacity *nnz (in: the capacity, out: the entries written) to
 * fill all three.  With CTR_DEVICE_PTRS the arrays are device memory; nnz is
 * always a host pointer. */
ctr_status ctr_cone_sysmat(ctr_ctx *ctx, const ctr_cone_geom *g, int64_t *row_ptr, int32_t *col, float *val,
                           int64_t *nnz, unsigned flags);
ctr_status ctr_fan_sysmat(ctr_ctx *ctx, const ctr_fan_geom *g, int64_t *row_ptr, int32_t *col, float *val,
                          int64_t *nnz, unsigned flags);
ctr_status ctr_par_sysmat(ctr_ctx *ctx, const ctr_par_geom *g, int64_t *row_ptr, int32_t *col, float *val,
                          int64_t *nnz, unsigned flags);

/* ---- SIRT -------------------------------------------------------------- */

/* x <- clamp(x + lambda * V .* A'(W .* (b - A x))), W = 1/(A 1 + 1e-9), V = 1/(A' 1 + 1e-9),
 * with A the Siddon projector and A' its exact transpose (Iterative_algorithm/SIRT3D.m:78-92).
 * b: [n_beta*n_xi*n_p], x: [nz*ns*nt] (in/out). */
ctr_status ctr_sirt3d(ctr_ctx *ctx, const ctr_cone_geom *g, const float *b, float *x, const ctr_sirt_opts *opts,
                      unsigned flags);

/* 2-D fan-beam SIRT (Iterative_algorithm/SIRT.m:126-154), any monotone detector. */
ctr_status ctr_sirt2d(ctr_ctx *ctx, const ctr_fan_geom *g, const float *b, float *x, const ctr_sirt_opts *opts,
                      unsigned flags);

/* ---- autotuner ----------------------------------------------------------
 * Times the tile and layout variants of the cone-beam kernels on the given
 * geometry (synthetic data, HIP events on the context's stream) and keeps the
 * fastest for every later call on this geometry in this context.  No
 * counterpart in the reference. */
typedef struct ctr_tune_result {
    int bp_pitch;           /* chosen row pitch (floats) of the FDK / BP_REF footprint buffers */
    int adj_cm;             /* chosen candidate-column chunk of the exact adjoint */
    int fp_transpose;       /* 1: views along t read a [z][t][s] copy of the volume */
    double bp_ms[3];        /* FDK backprojection, pitch 64 / 72 / 80 */
    double adj_ms[3];       /* exact adjoint, chunk 3 / 4 / 6 */
    double fp_ms[2];        /* forward projection, without / with the transposed copy */
} ctr_tune_result;

ctr_status ctr_autotune_cone(ctr_ctx *ctx, const ctr_cone_geom *g, ctr_tune_result *out);

/* How the cone-beam backprojection plan of this geometry classifies its
 * (tile, view) pairs: counts[0] all pairs, [1] empty (the tile projects off
 * the detector), [2] interior (every sample on the detector), [3] rows-in
 * (only u leaves the detector; FDK takes its clamped interior loop unless
 * the environment sets CTR_BP_ROWSIN=0), [4] other edge pairs staged in LDS,
 * [5] footprints too large for LDS (read from global memory).  adjoint != 0
 * reports the exact adjoint's plan ([3] is then 0).  A diagnostic with no
 * counterpart in the reference. */
ctr_status ctr_cone_bp_plan_stats(ctr_ctx *ctx, const ctr_cone_geom *g, int adjoint, long long counts[6]);

/* Many independent SIRT3D problems (uneven volumes, views and detectors)
 * advanced together: each iteration is one batched residual launch, one
 * reduction and one batched update launch per LDS layout, captured once in
 * a HIP graph and replayed n_iter times.  opts->residual_l2/l1 are ignored;
 * each problem reports through its own residual_l2. */
typedef struct ctr_sirt_problem {
    ctr_cone_geom geom;
    const float *b;         /* [n_beta*n_xi*n_p] projections */
    float *x;               /* [nz*ns*nt] iterate (in when opts->use_x0, out) */
    double *residual_l2;    /* optional host [n_iter+1]: ||b - A x_k||_2 / ||b||_2 */
} ctr_sirt_problem;

ctr_status ctr_sirt3d_batched(ctr_ctx *ctx, const ctr_sirt_problem *problems, int n, const ctr_sirt_opts *opts,
                              unsigned flags);

/* ---- learned SIRT update (no counterpart in the reference) -------------
 *
 * A 3-D CNN maps (x_k, s_k / rms(s_k)) to a correction d_k, where s_k is the
 * plain SIRT step lambda V .* A'(W .* (b - A x_k)).  Network: conv3d
 * 2->32->32->32->1, 3x3x3 kernels, zero padding, ReLU between layers, BF16
 * weights and activations, FP32 accumulation.  Parameters are FP32 in
 * PyTorch conv3d order: W1[32][2][3][3][3], b1[32], W2[32][32][3][3][3], b2[32],
 * W3[32][32][3][3][3], b3[32], W4[1][32][3][3][3], b4[1]. */
#define CTR_CNN_NUM_PARAMS 57985

typedef struct ctr_cnn ctr_cnn;

ctr_status ctr_cnn_create(ctr_ctx *ctx, ctr_cnn **out);   /* parameters start at zero */
void ctr_cnn_destroy(ctr_cnn *net);
ctr_status ctr_cnn_set_params(ctr_cnn *net, const float *params, size_t n);   /* host array, n = CTR_CNN_NUM_PARAMS */
ctr_status ctr_cnn_get_params(ctr_cnn *net, float *params, size_t n);

/* out = CNN(x, g * g_scale) over an nx x ny x nz volume (x fastest). */
ctr_status ctr_cnn_forward(ctr_cnn *net, int nx, int ny, int nz, const float *x, const float *g, float g_scale,
                           float *out, unsigned flags);

/* The same output as ctr_cnn_forward, computed tile by tile: tiles of tile^3
 * output voxels whose input boxes reach halo voxels past them (cropped at the
 * volume's faces).  halo >= 4 (the network's receptive radius) makes the
 * result identical to the whole-volume pass; the workspaces hold one tile, so
 * volumes whose activations would not fit whole (512^3 and up) still run.
 * No counterpart in the reference. */
ctr_status ctr_cnn_forward_tiled(ctr_cnn *net, int nx, int ny, int nz, const float *x, const float *g, float g_scale,
                                 float *out, int tile, int halo, unsigned flags);

/* Mean squared error of CNN(x_i, g_i * g_scale[i]) against target_i over the
 * voxels at least `border` from every face, averaged over n_patch patches of
 * nx x ny x nz (x, g, target: [n_patch][nz][ny][nx]; g_scale: host [n_patch]).
 * Returns the loss and, when grad is not NULL, dloss/dparams (host
 * [CTR_CNN_NUM_PARAMS]).  With step_lr > 0 it also takes one Adam step
 * (beta1 0.9, beta2 0.999, eps 1e-8) with that learning rate. */
ctr_status ctr_cnn_train_step(ctr_cnn *net, int n_patch, int nx, int ny, int nz, const float *x, const float *g,
                              const float *g_scale, const float *target, int border, float step_lr, double *loss,
                              float *grad, unsigned flags);

typedef struct ctr_learned_opts {
    int n_steps;            /* learned steps */
    float lambda;           /* relaxation of the plain step */
    float min_val, max_val; /* clamp of every iterate */
    int use_x0;             /* nonzero: x holds the starting iterate, else start from 0 */
    double *residual_l2;    /* optional host [n_steps+1]: ||b - A x_k||_2 / ||b||_2 */
    int *accepted;          /* optional host [n_steps]: 1 learned step kept, 0 fell back to the plain step */
    int judge_vs_plain;     /* nonzero: keep the learned step only if its residual is below the plain
                               step's (one more forward projection per step) */
    int cnn_tile;           /* 0: the network runs over the whole volume; > 0: in tiles of cnn_tile^3
                               voxels with 8-voxel halos (same result, workspaces of one tile) */
} ctr_learned_opts;

/* The plain SIRT step at x: s = lambda * V .* A'(W .* (b - A x)) (the
 * network's second input; ctr_sirt3d applies clamp(x + s) each iteration). */
ctr_status ctr_sirt3d_step(ctr_ctx *ctx, const ctr_cone_geom *g, const float *b, const float *x, float *s,
                           float lambda, unsigned flags);

/* Learned SIRT: x_{k+1} = clamp(x_k + s_k + rms(s_k) d_k) when that lowers
 * the data residual ||b - A x|| (below ||b - A x_k||, or below the plain
 * step's residual with judge_vs_plain), else the plain step clamp(x_k + s_k). */
ctr_status ctr_learned_sirt3d(ctr_ctx *ctx, const ctr_cone_geom *g, const float *b, float *x, ctr_cnn *net,
                              const ctr_learned_opts *opts, unsigned flags);

/* ---- SIRT emulated by a restricted Chebyshev polynomial (no counterpart
 * in the reference) ------------------------------------------------------
 *
 * Plain clamped SIRT from zero (SIRT3D.m:67-92) for n_start iterations, then
 * the remaining n_target - n_start iterations emulated in degree operator
 * applications: on a support S, x = x_s + p(lambda M_S) s_S, where M = V A'W A,
 * M_S its restriction to S, s_S the masked plain step at the masked start
 * x_s, and p the Chebyshev interpolant of p_N(s) = (1 - (1 - s)^N)/s on the
 * spectrum [0, lambda].  S is {net(x, s / rms(s)) > support_threshold} with a
 * support network (its output is a probability that the voxel is positive
 * after n_target plain iterations), else {x > min_val}.  With support_degree
 * > 0 a first emulation of that degree runs on S, and the main one on the
 * support of its clamped result.  The plain projector judges the clamped
 * result: it replaces x_{n_start} only when its data residual is lower, else
 * the plain iterate stays.  Then n_polish plain clamped iterations.  One
 * operator application costs about one SIRT iteration, so the call costs
 * about n_start + support_degree + degree + n_polish + 2 of them. */
typedef struct ctr_emulated_opts {
    int n_start;            /* plain iterations before the emulation */
    int n_target;           /* plain iterations the result stands for, counted from zero */
    int degree;             /* Chebyshev degree: operator applications of the emulation (1..1024) */
    int n_polish;           /* plain iterations after it */
    float lambda;           /* relaxation, in (0, 2) */
    float min_val, max_val; /* clamp of every iterate */
    float support_threshold;/* support from the network's output above this (with a network) */
    int cnn_tile;           /* network tiles as in ctr_learned_opts */
    int support_degree;     /* > 0: a first emulation of this degree refines the support */
    double *residual_l2;    /* optional host [3]: ||b - A x|| / ||b|| after the start, of the emulated
                               iterate, at the end */
    int *fell_back;         /* optional host: 1 when the emulated iterate did not lower the residual */
} ctr_emulated_opts;

ctr_status ctr_sirt3d_emulated(ctr_ctx *ctx, const ctr_cone_geom *g, const float *b, float *x, ctr_cnn *support_net,
                               const ctr_emulated_opts *opts, unsigned flags);

#ifdef __cplusplus
}
#endif

#endif /* CTR_CTR_H_ */
