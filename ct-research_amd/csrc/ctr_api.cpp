// ctr_api.cpp - the C ABI of libctr: contexts, argument checks, geometry
// tables, workspaces and dispatch to the HIP kernels.  There is no CPU
// compute path: every operator runs on the GPU or returns an error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <vector>

#include "ctr_cnn.h"
#include "ctr_internal.h"

using namespace ctr;

namespace {

thread_local std::string g_last_error;
// Failures on this thread, counted: a context's message is the thread's
// latest only while no later failure (e.g. a geometry check that runs
// before the context is looked at) happened on the thread.
thread_local unsigned long long g_error_seq = 0;

enum Slot { kWsIn = 0, kWsOut, kWsTmp, kWsR, kWsVinv, kWsPart, kWsSums, kWsG, kWsS, kWsD, kWsX2, kWsR2, kWsT, kWsYn,
            kWsMask, kWsXs, kWsCount };

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

struct ConeTables {
    uint64_t key = 0;
    bool valid = false;
    std::vector<double> h_view;     // [nb][4]
    std::vector<double> h_pd, h_xd; // positions (double)
    std::vector<BpView> h_bp;
    std::vector<float> h_pf, h_xf;
    std::vector<float> h_zq;        // [nz + 32][2]: zs^2, hz/|zs| per slice (zs = slice centre - source z)
    DevBuf d_view, d_p, d_xi, d_bp, d_pf, d_xf, d_zq;
    DevBuf d_ainv, d_binv, d_xinv, d_norm;   // exact-adjoint tables, built on the device
    DevBuf d_fw;                              // FDK pre-weights [nx][np] (FP64), built on first filter use
    uint64_t fw_key = 0;
};

struct FanTables {
    uint64_t key = 0;
    bool valid = false;
    std::vector<double> h_view, h_p, h_zero;
    std::vector<float> h_pf, h_gam, h_colw, h_xf;
    DevBuf d_view, d_p, d_zero, d_pf, d_gam, d_colw, d_xf;
};

struct ParTables {
    uint64_t key = 0;
    bool valid = false;
    std::vector<double> h_view, h_t, h_zero;   // projector: [nb][4] sin, cos of theta; ray offsets
    std::vector<float> h_vf;                   // FBP: [nb][2] sin, cos of the single-precision theta
    DevBuf d_view, d_t, d_zero, d_vf;
};

struct CurvedTables {
    uint64_t key = 0;
    bool valid = false;
    std::vector<float> h_view;             // [nb][2] sin, cos
    std::vector<double> h_rowp, h_beta, h_gam;   // [nx] row scale; view and channel angles
    int fft_len = 0;
    std::vector<double> h_spec;            // [nx][fft_len] per-row spectrum / fft_len
    DevBuf d_view, d_rowp, d_beta, d_gam, d_spec;
    uint64_t foot_key = 0;                 // backprojection tile footprints (per geometry and voxel offset)
    DevBuf d_foot;
};

struct FdkErrTables {
    DevBuf d_cos, d_sin, d_u, d_t, d_p, d_g, d_cp;
};

struct PhantomTable {
    std::vector<PhantomEll> h;
    DevBuf d;
};

struct FootCache {
    DevBuf buf;
    DevBuf unfit;           // one int: footprints too large for the adjoint's LDS
    uint64_t key = 0;
    BpPlan plan{};
};

struct TapTable {
    uint64_t key = 0;
    bool valid = false;
    std::vector<double> h;
    DevBuf d;
    int fft_len = 0;                // FFT path length (0: direct convolution)
    std::vector<double> spec;       // [fft_len] H[k] / fft_len
    DevBuf d_spec;
};

// Tile/layout choices of the cone-beam kernels for one geometry (defaults
// until ctr_autotune_cone measures them).
struct Tuning {
    int bp_pitch = 80;      // row pitch of the FDK / BP_REF footprint buffers
    int adj_cm = 3;         // candidate-column chunk of the exact adjoint
    int fp_transpose = 1;   // forward projector reads the [z][t][s] copy for views along t
};

}  // namespace

struct ctr_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::string last_error;
    unsigned long long error_seq = 0;   // g_error_seq of the failure that set last_error
    DevBuf ws[kWsCount];
    ConeTables cone;
    FanTables fan;
    ParTables par;
    CurvedTables curved;
    FdkErrTables fdkerr;
    PhantomTable phantom;
    TapTable taps;
    FootCache foot_bp, foot_adj;   // footprint tables of the interpolating BP and of the exact adjoint
    // SIRT's V = 1/(A'1 + 1e-9) depends on the geometry only: ws[kWsVinv] keeps
    // it for the operator with this key; reserve_vinv clears the key whenever
    // the buffer may be reallocated, and sirt_norm sets it once V is written
    uint64_t vinv_key = 0;
    int fault = 0;   // CTR_FAULT test hook: 1 alloc, 2 launch, 3 copy
    // device address of the library's fixed scratch array on this device
    // (the allocation-free filter's tables), looked up at creation
    void *scratch = nullptr;
    hipError_t scratch_err = hipSuccess;
    std::vector<ctr_cnn *> nets;   // networks created on this context (released with it)
    std::map<uint64_t, Tuning> tune;   // autotuned choices per cone geometry
};

namespace {
void cnn_release(ctr_cnn *net);   // frees a network's device buffers and detaches it from its context
}

namespace {

ctr_status fail(ctr_ctx *ctx, ctr_status st, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ++g_error_seq;
    if (ctx) {
        ctx->last_error = buf;
        ctx->error_seq = g_error_seq;
    }
    g_last_error = buf;
    return st;
}

#define HIP_TRY(ctx, expr)                                                                              \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return fail((ctx), CTR_EHIP, "%s: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

int parse_fault()
{
    const char *f = getenv("CTR_FAULT");
    if (!f) return 0;
    if (!strcmp(f, "alloc")) return 1;
    if (!strcmp(f, "launch")) return 2;
    if (!strcmp(f, "copy")) return 3;
    return 0;
}

ctr_status buf_reserve(ctr_ctx *ctx, DevBuf &b, size_t bytes)
{
    if (ctx->fault == 1) return fail(ctx, CTR_EFAULT_INJECTED, "CTR_FAULT=alloc: device allocation of %zu bytes", bytes);
    if (bytes <= b.bytes && b.ptr) return CTR_OK;
    if (b.ptr) {
        // the old buffer may still be read by queued work
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, CTR_ENOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    b.ptr = p;
    b.bytes = bytes;
    return CTR_OK;
}

// The workspace of SIRT's V: a reallocation (even one that fails, or that
// returns the old address) invalidates the V kept in it.
ctr_status reserve_vinv(ctr_ctx *ctx, size_t bytes)
{
    DevBuf &b = ctx->ws[kWsVinv];
    if (!(b.ptr && bytes <= b.bytes)) ctx->vinv_key = 0;
    return buf_reserve(ctx, b, bytes);
}

void buf_free(DevBuf &b)
{
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

ctr_status upload(ctr_ctx *ctx, DevBuf &b, const void *src, size_t bytes)
{
    ctr_status s = buf_reserve(ctx, b, bytes);
    if (s != CTR_OK) return s;
    if (ctx->fault == 3) return fail(ctx, CTR_EFAULT_INJECTED, "CTR_FAULT=copy: host-to-device copy");
    HIP_TRY(ctx, hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return CTR_OK;
}

ctr_status check_launch(ctr_ctx *ctx, hipError_t e, const char *what)
{
    if (ctx->fault == 2) return fail(ctx, CTR_EFAULT_INJECTED, "CTR_FAULT=launch: %s", what);
    if (e != hipSuccess) return fail(ctx, CTR_EHIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    return CTR_OK;
}

uint64_t fnv(uint64_t h, const void *data, size_t n)
{
    const unsigned char *p = static_cast<const unsigned char *>(data);
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

bool finite_pos(double x) { return std::isfinite(x) && x > 0.0; }

template <typename T>
ctr_status check_axis(const T *a, int n, bool need_uniform, const char *name)
{
    if (!a) return fail(nullptr, CTR_EINVAL, "%s is NULL", name);
    for (int i = 0; i < n; ++i)
        if (!std::isfinite((double)a[i])) return fail(nullptr, CTR_EINVAL, "%s[%d] is not finite", name, i);
    if (n < 2) return CTR_OK;
    const double d0 = (double)a[1] - (double)a[0];
    if (d0 == 0.0) return fail(nullptr, CTR_EINVAL, "%s is not strictly monotonic", name);
    for (int i = 1; i < n; ++i) {
        const double d = (double)a[i] - (double)a[i - 1];
        if ((d > 0) != (d0 > 0) || d == 0.0) return fail(nullptr, CTR_EINVAL, "%s is not strictly monotonic", name);
        if (need_uniform && std::fabs(d - d0) > 1e-3 * std::fabs(d0))
            return fail(nullptr, CTR_EINVAL, "%s has a non-uniform pitch (step %d is %g, first step %g)", name, i, d, d0);
    }
    return CTR_OK;
}

ctr_status check_cone(const ctr_cone_geom *g, int need_uniform)
{
    if (!g) return fail(nullptr, CTR_EINVAL, "geometry is NULL");
    if (g->nt < 1 || g->ns < 1 || g->nz < 1) return fail(nullptr, CTR_EINVAL, "volume dimensions must be positive");
    // the projector reads the volume through one buffer descriptor with 32-bit byte offsets
    if ((double)g->nt * g->ns * g->nz > (double)(1 << 30))
        return fail(nullptr, CTR_ETOOBIG, "volume has more than 2^30 voxels");
    for (int a = 0; a < 3; ++a)
        if (!finite_pos(g->size[a])) return fail(nullptr, CTR_EINVAL, "size[%d] must be positive and finite", a);
    if (!finite_pos(g->distance)) return fail(nullptr, CTR_EINVAL, "distance must be positive and finite");
    if (g->n_beta < 1 || g->n_p < 1 || g->n_xi < 1) return fail(nullptr, CTR_EINVAL, "detector/view counts must be positive");
    if (need_uniform && (g->n_p < 2 || g->n_xi < 2))
        return fail(nullptr, CTR_EINVAL, "backprojection needs at least 2 detector columns and rows");
    if ((double)g->n_beta * g->n_p * g->n_xi > (double)std::numeric_limits<int>::max())
        return fail(nullptr, CTR_ETOOBIG, "more than 2^31-1 detector samples");
    // a view is addressed through one buffer descriptor with 32-bit byte offsets
    if ((double)g->n_p * g->n_xi > (double)(1 << 28))
        return fail(nullptr, CTR_ETOOBIG, "a view has more than 2^28 detector samples");
    const double half = 0.5 * std::sqrt(g->size[0] * g->size[0] + g->size[1] * g->size[1]);
    if (!(g->distance > half * (1.0 + 1e-6)))
        return fail(nullptr, CTR_EINVAL, "source orbit (distance %g) intersects the volume (half diagonal %g)",
                    g->distance, half);
    // view angles need not be monotonic: only finite
    if (!g->beta) return fail(nullptr, CTR_EINVAL, "beta is NULL");
    for (int i = 0; i < g->n_beta; ++i)
        if (!std::isfinite(g->beta[i])) return fail(nullptr, CTR_EINVAL, "beta[%d] is not finite", i);
    ctr_status s;
    if ((s = check_axis(g->p, g->n_p, need_uniform != 0, "p")) != CTR_OK) return s;
    if ((s = check_axis(g->xi, g->n_xi, need_uniform != 0, "xi")) != CTR_OK) return s;
    return CTR_OK;
}

ctr_status check_fan(const ctr_fan_geom *g, int need_uniform)
{
    if (!g) return fail(nullptr, CTR_EINVAL, "geometry is NULL");
    if (g->nt < 1 || g->ns < 1) return fail(nullptr, CTR_EINVAL, "image dimensions must be positive");
    if ((double)g->nt * g->ns > (double)(1 << 30))
        return fail(nullptr, CTR_ETOOBIG, "image has more than 2^30 pixels");
    for (int a = 0; a < 2; ++a)
        if (!finite_pos(g->size[a])) return fail(nullptr, CTR_EINVAL, "size[%d] must be positive and finite", a);
    if (!finite_pos(g->distance)) return fail(nullptr, CTR_EINVAL, "distance must be positive and finite");
    if (g->n_beta < 1 || g->n_p < 1) return fail(nullptr, CTR_EINVAL, "detector/view counts must be positive");
    if (need_uniform && g->n_p < 2) return fail(nullptr, CTR_EINVAL, "backprojection needs at least 2 detector bins");
    const double half = 0.5 * std::sqrt(g->size[0] * g->size[0] + g->size[1] * g->size[1]);
    if (!(g->distance > half * (1.0 + 1e-6)))
        return fail(nullptr, CTR_EINVAL, "source orbit (distance %g) intersects the image (half diagonal %g)",
                    g->distance, half);
    if (!g->beta) return fail(nullptr, CTR_EINVAL, "beta is NULL");
    for (int i = 0; i < g->n_beta; ++i)
        if (!std::isfinite(g->beta[i])) return fail(nullptr, CTR_EINVAL, "beta[%d] is not finite", i);
    return check_axis(g->p, g->n_p, need_uniform != 0, "p");
}

ctr_status check_par(const ctr_par_geom *g, int need_uniform)
{
    if (!g) return fail(nullptr, CTR_EINVAL, "geometry is NULL");
    if (g->nt < 1 || g->ns < 1) return fail(nullptr, CTR_EINVAL, "image dimensions must be positive");
    if ((double)g->nt * g->ns > (double)(1 << 30))
        return fail(nullptr, CTR_ETOOBIG, "image has more than 2^30 pixels");
    for (int a = 0; a < 2; ++a)
        if (!finite_pos(g->size[a])) return fail(nullptr, CTR_EINVAL, "size[%d] must be positive and finite", a);
    if (g->n_theta < 1 || g->n_t < 1) return fail(nullptr, CTR_EINVAL, "detector/view counts must be positive");
    if (g->n_theta > 65535) return fail(nullptr, CTR_ETOOBIG, "at most 65535 views");
    if (need_uniform && g->n_t < 2) return fail(nullptr, CTR_EINVAL, "backprojection needs at least 2 detector bins");
    if (!g->theta) return fail(nullptr, CTR_EINVAL, "theta is NULL");
    for (int i = 0; i < g->n_theta; ++i)
        if (!std::isfinite(g->theta[i])) return fail(nullptr, CTR_EINVAL, "theta[%d] is not finite", i);
    return check_axis(g->t, g->n_t, need_uniform != 0, "t");
}

ctr_status check_curved(const ctr_curved_geom *g)
{
    if (!g) return fail(nullptr, CTR_EINVAL, "geometry is NULL");
    if (g->nt < 1 || g->ns < 1 || g->nz < 1) return fail(nullptr, CTR_EINVAL, "volume dimensions must be positive");
    if ((double)g->nt * g->ns * g->nz > (double)std::numeric_limits<int>::max())
        return fail(nullptr, CTR_ETOOBIG, "volume has more than 2^31-1 voxels");
    for (int a = 0; a < 3; ++a)
        if (!finite_pos(g->size[a])) return fail(nullptr, CTR_EINVAL, "size[%d] must be positive and finite", a);
    if (!finite_pos(g->distance) || !finite_pos(g->distance_s2d))
        return fail(nullptr, CTR_EINVAL, "distances must be positive and finite");
    const double half = 0.5 * std::sqrt(g->size[0] * g->size[0] + g->size[1] * g->size[1]);
    if (!(g->distance > half * (1.0 + 1e-6)))
        return fail(nullptr, CTR_EINVAL, "source orbit (distance %g) intersects the volume (half diagonal %g)",
                    g->distance, half);
    if (g->n_beta < 2 || g->n_gamma < 2 || g->n_xi < 2)
        return fail(nullptr, CTR_EINVAL, "need at least 2 views, 2 channels and 2 rows");
    if (g->n_gamma > 1024) return fail(nullptr, CTR_ETOOBIG, "the curved-detector filter supports at most 1024 channels");
    // the backprojection reads the filtered projections with 32-bit byte offsets
    if ((double)g->n_beta * g->n_gamma * g->n_xi > (double)(1 << 30))
        return fail(nullptr, CTR_ETOOBIG, "more than 2^30 detector samples");
    if (g->n_beta > 65535) return fail(nullptr, CTR_ETOOBIG, "at most 65535 views");
    if (!std::isfinite(g->beta_start) || !finite_pos(g->fan_angle) || g->fan_angle >= 3.141592653589793)
        return fail(nullptr, CTR_EINVAL, "beta_start must be finite and fan_angle in (0, pi)");
    ctr_status s;
    if ((s = check_axis(g->beta, g->n_beta, true, "beta")) != CTR_OK) return s;
    if ((s = check_axis(g->gamma, g->n_gamma, true, "gamma")) != CTR_OK) return s;
    if ((s = check_axis(g->xi, g->n_xi, true, "xi")) != CTR_OK) return s;
    for (int i = 0; i < g->n_gamma; ++i)
        if (!(std::fabs(g->gamma[i]) < 0.5 * 3.141592653589793))
            return fail(nullptr, CTR_EINVAL, "gamma[%d] is not inside (-pi/2, pi/2)", i);
    return CTR_OK;
}

ctr_status need_ctx(ctr_ctx *ctx)
{
    if (!ctx) return fail(nullptr, CTR_EINVAL, "context is NULL");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return CTR_OK;
}

// ---- geometry tables -------------------------------------------------------

uint64_t cone_key(const ctr_cone_geom *g)
{
    uint64_t h = 1469598103934665603ull;
    h = fnv(h, &g->nt, sizeof(int) * 3);
    h = fnv(h, g->size, sizeof(double) * 3);
    h = fnv(h, &g->distance, sizeof(double));
    h = fnv(h, &g->n_beta, sizeof(int) * 3);
    h = fnv(h, g->beta, sizeof(float) * g->n_beta);
    h = fnv(h, g->p, sizeof(float) * g->n_p);
    h = fnv(h, g->xi, sizeof(float) * g->n_xi);
    return h;
}

BpParams cone_bp_params_from(const ConeTables &T, const ctr_cone_geom *g, unsigned flags);

ctr_status cone_tables_into(ctr_ctx *ctx, ConeTables &T, const ctr_cone_geom *g)
{
    const uint64_t key = cone_key(g);
    if (T.valid && T.key == key) return CTR_OK;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));   // queued work may still read the old tables
    T.valid = false;
    const int nb = g->n_beta;
    const double D = g->distance, Ct = g->size[0] / 2, Cs = g->size[1] / 2;
    const double rt = g->size[0] / g->nt, rs = g->size[1] / g->ns;
    const float dbeta = nb >= 2 ? std::fabs(g->beta[1] - g->beta[0]) : 0.0f;
    T.h_view.resize(4 * (size_t)nb);
    T.h_bp.resize(nb);
    for (int b = 0; b < nb; ++b) {
        const double beta = (double)g->beta[b], sb = std::sin(beta), cb = std::cos(beta);
        T.h_view[4 * b + 0] = sb;
        T.h_view[4 * b + 1] = cb;
        T.h_view[4 * b + 2] = Ct - D * sb;
        T.h_view[4 * b + 3] = Cs + D * cb;
        const double x0 = 0.5 * rt - Ct, y0 = 0.5 * rs - Cs;
        BpView v{};
        v.at = (float)(rt * cb);
        v.as = (float)(rs * sb);
        v.a0 = (float)(x0 * cb + y0 * sb);
        v.bt = (float)(-rt * sb);
        v.bs = (float)(rs * cb);
        v.b0 = (float)(-x0 * sb + y0 * cb);
        v.wv = dbeta;
        v.st = (float)(Ct - D * sb);
        v.ss = (float)(Cs + D * cb);
        v.sb = (float)sb;
        v.cb = (float)cb;
        T.h_bp[b] = v;
    }
    // per-slice chord factors, padded by 32 slices so that a partial z tile reads in bounds
    const double rz = g->size[2] / g->nz, Cz = g->size[2] / 2;
    T.h_zq.resize(2 * ((size_t)g->nz + 32));
    for (int z = 0; z < g->nz + 32; ++z) {
        const double zs = (z + 0.5) * rz - Cz;
        T.h_zq[2 * z] = (float)(zs * zs);
        T.h_zq[2 * z + 1] = (float)(0.5 * rz / std::fabs(zs));
    }
    T.h_pd.assign(g->p, g->p + g->n_p);
    T.h_xd.assign(g->xi, g->xi + g->n_xi);
    T.h_pf.assign(g->p, g->p + g->n_p);
    T.h_xf.assign(g->xi, g->xi + g->n_xi);
    ctr_status s;
    if ((s = upload(ctx, T.d_view, T.h_view.data(), sizeof(double) * T.h_view.size())) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_p, T.h_pd.data(), sizeof(double) * T.h_pd.size())) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_xi, T.h_xd.data(), sizeof(double) * T.h_xd.size())) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_bp, T.h_bp.data(), sizeof(BpView) * T.h_bp.size())) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_pf, T.h_pf.data(), sizeof(float) * T.h_pf.size())) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_xf, T.h_xf.data(), sizeof(float) * T.h_xf.size())) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_zq, T.h_zq.data(), sizeof(float) * T.h_zq.size())) != CTR_OK) return s;
    const size_t nvc = (size_t)nb * g->n_p, nrc = (size_t)g->n_xi * g->n_p;
    if ((s = buf_reserve(ctx, T.d_ainv, nvc * sizeof(float))) != CTR_OK) return s;
    if ((s = buf_reserve(ctx, T.d_binv, nvc * sizeof(float))) != CTR_OK) return s;
    if ((s = buf_reserve(ctx, T.d_xinv, (size_t)g->n_xi * sizeof(float))) != CTR_OK) return s;
    if ((s = buf_reserve(ctx, T.d_norm, nrc * sizeof(float))) != CTR_OK) return s;
    const BpParams P = cone_bp_params_from(T, g, 0);
    if ((s = check_launch(ctx, launch_adjoint_tables(P, ctx->stream), "adjoint tables")) != CTR_OK) return s;
    // later users may run on other streams (batched SIRT)
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    T.key = key;
    T.valid = true;
    return CTR_OK;
}

RayParams cone_ray_params_from(const ConeTables &T, const ctr_cone_geom *g)
{
    RayParams r{};
    r.nt = g->nt;
    r.ns = g->ns;
    r.nz = g->nz;
    r.np = g->n_p;
    r.nx = g->n_xi;
    r.nb = g->n_beta;
    r.rt = g->size[0] / g->nt;
    r.rs = g->size[1] / g->ns;
    r.rz = g->size[2] / g->nz;
    r.bt = g->size[0];
    r.bs = g->size[1];
    r.bz = g->size[2];
    r.D = g->distance;
    r.sz = g->size[2] / 2;
    r.view = (const double *)T.d_view.ptr;
    r.p = (const double *)T.d_p.ptr;
    r.xi = (const double *)T.d_xi.ptr;
    r.residual_norm = 1;
    // masked second/third voxel reads once the volume exceeds the 256 MiB
    // Infinity Cache (measured: 512^3 301 -> 270 ms, 256^3 19.3 -> 20.4 ms);
    // CTR_FP_SKIP=0/1 overrides
    r.skip_empty = (size_t)g->nt * g->ns * g->nz * sizeof(float) > ((size_t)256 << 20);
    if (const char *e = std::getenv("CTR_FP_SKIP")) r.skip_empty = std::atoi(e) != 0;
    return r;
}

BpParams cone_bp_params_from(const ConeTables &T, const ctr_cone_geom *g, unsigned flags)
{
    BpParams P{};
    P.nt = g->nt;
    P.ns = g->ns;
    P.nz = g->nz;
    P.np = g->n_p;
    P.nx = g->n_xi;
    P.nb = g->n_beta;
    P.rt = (float)(g->size[0] / g->nt);
    P.rs = (float)(g->size[1] / g->ns);
    P.rz = (float)(g->size[2] / g->nz);
    P.rt_d = g->size[0] / g->nt;
    P.rs_d = g->size[1] / g->ns;
    P.zc0 = (float)(0.5 * g->size[2] / g->nz - g->size[2] / 2);
    P.D = (float)g->distance;
    // uniform grid through the end samples: a pitch taken from the first two
    // float samples (as Backprojection3D.cu:304-305 does) drifts by up to
    // (n-1)*ulp/pitch samples across the detector when the pitch is not exact in float
    const double dp = ((double)g->p[g->n_p - 1] - (double)g->p[0]) / (g->n_p - 1);
    const double dx = ((double)g->xi[g->n_xi - 1] - (double)g->xi[0]) / (g->n_xi - 1);
    P.u_scale = (float)(1.0 / dp);
    P.u_off = (float)(-(double)g->p[0] / dp);
    P.v_scale = (float)(1.0 / dx);
    P.v_off = (float)(-(double)g->xi[0] / dx);
    P.umax = (float)(g->n_p - 1);
    P.vmax = (float)(g->n_xi - 1);
    P.ht = 0.5f * P.rt;
    P.hs = 0.5f * P.rs;
    P.hz = 0.5f * P.rz;
    P.views = (const BpView *)T.d_bp.ptr;
    P.view_d = (const double *)T.d_view.ptr;
    P.pdd = (const double *)T.d_p.ptr;
    P.pd = (const float *)T.d_pf.ptr;
    P.xd = (const float *)T.d_xf.ptr;
    P.zq = (const float2 *)T.d_zq.ptr;
    {
        const char *e = getenv("CTR_BP_ROWSIN");
        P.rows_in = (e && !strcmp(e, "0")) ? 0 : 1;
    }
    P.accumulate = (flags & CTR_ACCUMULATE) ? 1 : 0;
    P.quirks = (flags & CTR_REF_QUIRKS) ? 1 : 0;
    P.adj_inva = (const float *)T.d_ainv.ptr;
    P.adj_invb = (const float *)T.d_binv.ptr;
    P.adj_invxi = (const float *)T.d_xinv.ptr;
    P.adj_norm = (const float *)T.d_norm.ptr;
    return P;
}

// Footprint table of the tiled backprojection for the current cone geometry
// (computed on the device once per geometry, reused by every later call).
Tuning tuning_for(const ctr_ctx *ctx, uint64_t key)
{
    const auto it = ctx->tune.find(key);
    return it == ctx->tune.end() ? Tuning{} : it->second;
}

ctr_status ensure_footprints_key(ctr_ctx *ctx, const BpParams &P, FootCache &fc, bool adjoint, uint64_t base_key)
{
    const Tuning tu = tuning_for(ctx, base_key);
    BpPlan L = adjoint ? adjoint_plan(P, tu.adj_cm) : bp_plan(P, tu.bp_pitch);
    uint64_t key = base_key;
    key = fnv(key, &L, sizeof L);
    if (fc.buf.ptr && fc.key == key) return CTR_OK;
    ctr_status s = buf_reserve(ctx, fc.buf, L.nfoot * sizeof(int4));
    if (s != CTR_OK) return s;
    s = check_launch(ctx, launch_bp_footprints(P, L, (int4 *)fc.buf.ptr, ctx->stream), "footprint table");
    if (s != CTR_OK) return s;
    // the exact adjoint has a leaner variant for plans whose footprints all
    // fit LDS: counted once per geometry (not while a graph is captured)
    L.fit_all = 0;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (adjoint && hipStreamIsCapturing(ctx->stream, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
        s = buf_reserve(ctx, fc.unfit, sizeof(int) * kFootCountBlocks);
        if (s != CTR_OK) return s;
        s = check_launch(ctx, launch_foot_unfit((const int4 *)fc.buf.ptr, L.nfoot, (int *)fc.unfit.ptr, ctx->stream),
                         "footprint count");
        if (s != CTR_OK) return s;
        int counts[kFootCountBlocks];
        HIP_TRY(ctx, hipMemcpyAsync(counts, fc.unfit.ptr, sizeof counts, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        long long unfit = 0;
        for (int c : counts) unfit += c;
        L.fit_all = unfit == 0 ? 1 : 0;
    }
    fc.key = key;
    fc.plan = L;
    return CTR_OK;
}

ctr_status cone_tables(ctr_ctx *ctx, const ctr_cone_geom *g) { return cone_tables_into(ctx, ctx->cone, g); }
RayParams cone_ray_params(ctr_ctx *ctx, const ctr_cone_geom *g) { return cone_ray_params_from(ctx->cone, g); }
BpParams cone_bp_params(ctr_ctx *ctx, const ctr_cone_geom *g, unsigned flags)
{
    return cone_bp_params_from(ctx->cone, g, flags);
}
ctr_status ensure_footprints(ctr_ctx *ctx, const BpParams &P, FootCache &fc, bool adjoint)
{
    return ensure_footprints_key(ctx, P, fc, adjoint, ctx->cone.key);
}

// The [z][t][s] copy of a volume for the forward projector (its views along t
// read it), unless the geometry's tuning turned that layout off.
ctr_status fp_transposed(ctr_ctx *ctx, RayParams &R, const float *vol, DevBuf &buf, hipStream_t st,
                         bool enable = true)
{
    R.volT = nullptr;
    if (!enable) return CTR_OK;
    ctr_status s = buf_reserve(ctx, buf, (size_t)R.nt * R.ns * R.nz * sizeof(float));
    if (s != CTR_OK) return s;
    s = check_launch(ctx, launch_transpose_st(vol, (float *)buf.ptr, R.nt, R.ns, R.nz, st), "volume transpose");
    if (s != CTR_OK) return s;
    R.volT = (const float *)buf.ptr;
    return CTR_OK;
}

uint64_t fan_key(const ctr_fan_geom *g)
{
    uint64_t h = 1469598103934665603ull;
    h = fnv(h, &g->nt, sizeof(int) * 2);
    h = fnv(h, g->size, sizeof(double) * 2);
    h = fnv(h, &g->distance, sizeof(double));
    h = fnv(h, &g->n_beta, sizeof(int) * 2);
    h = fnv(h, g->beta, sizeof(double) * g->n_beta);
    h = fnv(h, g->p, sizeof(double) * g->n_p);
    return h;
}

ctr_status fan_tables(ctr_ctx *ctx, const ctr_fan_geom *g)
{
    FanTables &T = ctx->fan;
    const uint64_t key = fan_key(g);
    if (T.valid && T.key == key) return CTR_OK;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    T.valid = false;
    const int nb = g->n_beta, np = g->n_p;
    const double D = g->distance, Ct = g->size[0] / 2, Cs = g->size[1] / 2;
    T.h_view.resize(4 * (size_t)nb);
    for (int b = 0; b < nb; ++b) {
        const double sb = std::sin(g->beta[b]), cb = std::cos(g->beta[b]);
        T.h_view[4 * b + 0] = sb;
        T.h_view[4 * b + 1] = cb;
        T.h_view[4 * b + 2] = Ct - D * sb;
        T.h_view[4 * b + 3] = Cs + D * cb;
    }
    T.h_p.assign(g->p, g->p + np);
    T.h_zero.assign(1, 0.0);
    T.h_pf.resize(np);
    T.h_gam.resize(np);
    T.h_colw.resize(np);
    for (int i = 0; i < np; ++i) {
        T.h_pf[i] = (float)g->p[i];
        const double gam = std::atan(g->p[i] / D);
        T.h_gam[i] = (float)gam;
        T.h_colw[i] = (float)(D * std::cos(gam));
    }
    T.h_xf.assign(1, 0.0f);
    ctr_status s;
    if ((s = upload(ctx, T.d_view, T.h_view.data(), sizeof(double) * T.h_view.size())) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_p, T.h_p.data(), sizeof(double) * np)) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_zero, T.h_zero.data(), sizeof(double))) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_pf, T.h_pf.data(), sizeof(float) * np)) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_gam, T.h_gam.data(), sizeof(float) * np)) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_colw, T.h_colw.data(), sizeof(float) * np)) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_xf, T.h_xf.data(), sizeof(float))) != CTR_OK) return s;
    T.key = key;
    T.valid = true;
    return CTR_OK;
}

RayParams fan_ray_params(ctr_ctx *ctx, const ctr_fan_geom *g)
{
    const FanTables &T = ctx->fan;
    RayParams r{};
    r.nt = g->nt;
    r.ns = g->ns;
    r.nz = 1;
    r.np = g->n_p;
    r.nx = 1;
    r.nb = g->n_beta;
    r.rt = g->size[0] / g->nt;
    r.rs = g->size[1] / g->ns;
    r.rz = 1.0;
    r.bt = g->size[0];
    r.bs = g->size[1];
    r.bz = 1.0;
    r.D = g->distance;
    r.sz = 0.5;
    r.view = (const double *)T.d_view.ptr;
    r.p = (const double *)T.d_p.ptr;
    r.xi = (const double *)T.d_zero.ptr;
    return r;
}

FanParams fan_params(ctr_ctx *ctx, const ctr_fan_geom *g, unsigned flags)
{
    const FanTables &T = ctx->fan;
    FanParams P{};
    P.nt = g->nt;
    P.ns = g->ns;
    P.np = g->n_p;
    P.nb = g->n_beta;
    P.rt = (float)(g->size[0] / g->nt);
    P.rs = (float)(g->size[1] / g->ns);
    P.ct = (float)(g->size[0] / 2);
    P.cs = (float)(g->size[1] / 2);
    P.rt_d = g->size[0] / g->nt;
    P.rs_d = g->size[1] / g->ns;
    P.ct_d = g->size[0] / 2;
    P.cs_d = g->size[1] / 2;
    P.D = (float)g->distance;
    if (g->n_p >= 2) {
        // uniform grids through the end samples (see cone_bp_params)
        const int n = g->n_p;
        const double p0 = (double)T.h_pf[0], dp = ((double)T.h_pf[n - 1] - p0) / (n - 1);
        P.u_scale = (float)(1.0 / dp);
        P.u_off = (float)(-p0 / dp);
        P.u_scale_d = 1.0 / dp;
        P.u_off_d = -p0 / dp;
        const double g0 = std::atan(g->p[0] / g->distance);
        const double dg = (std::atan(g->p[n - 1] / g->distance) - g0) / (n - 1);
        P.g0 = (float)g0;
        P.g_scale = (float)(1.0 / dg);
    }
    P.umax = (float)(g->n_p - 1);
    P.gmax = (float)(g->n_p - 1);
    P.dbeta = g->n_beta >= 2 ? (float)std::fabs(g->beta[1] - g->beta[0]) : 0.0f;
    P.view = (const double *)T.d_view.ptr;
    P.pf = (const float *)T.d_pf.ptr;
    P.pdbl = (const double *)T.d_p.ptr;
    P.gam = (const float *)T.d_gam.ptr;
    P.accumulate = (flags & CTR_ACCUMULATE) ? 1 : 0;
    P.quirks = (flags & CTR_REF_QUIRKS) ? 1 : 0;
    return P;
}

uint64_t par_key(const ctr_par_geom *g)
{
    uint64_t h = 1469598103934665603ull;
    h = fnv(h, &g->nt, sizeof(int) * 2);
    h = fnv(h, g->size, sizeof(double) * 2);
    h = fnv(h, &g->n_theta, sizeof(int) * 2);
    h = fnv(h, g->theta, sizeof(double) * g->n_theta);
    h = fnv(h, g->t, sizeof(double) * g->n_t);
    return h;
}

ctr_status par_tables(ctr_ctx *ctx, const ctr_par_geom *g)
{
    ParTables &T = ctx->par;
    const uint64_t key = par_key(g);
    if (T.valid && T.key == key) return CTR_OK;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    T.valid = false;
    const int nb = g->n_theta, n = g->n_t;
    T.h_view.assign(4 * (size_t)nb, 0.0);
    T.h_vf.resize(2 * (size_t)nb);
    for (int b = 0; b < nb; ++b) {
        T.h_view[4 * b + 0] = std::sin(g->theta[b]);
        T.h_view[4 * b + 1] = std::cos(g->theta[b]);
        const double th = (double)(float)g->theta[b];   // FBPparallel receives single angles
        T.h_vf[2 * b + 0] = (float)std::sin(th);
        T.h_vf[2 * b + 1] = (float)std::cos(th);
    }
    T.h_t.assign(g->t, g->t + n);
    T.h_zero.assign(1, 0.0);
    ctr_status s;
    if ((s = upload(ctx, T.d_view, T.h_view.data(), sizeof(double) * T.h_view.size())) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_t, T.h_t.data(), sizeof(double) * n)) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_zero, T.h_zero.data(), sizeof(double))) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_vf, T.h_vf.data(), sizeof(float) * T.h_vf.size())) != CTR_OK) return s;
    T.key = key;
    T.valid = true;
    return CTR_OK;
}

// The parallel projector runs the shared Siddon kernel: one detector row,
// rays through the foot points centre + t (cos, sin) along (-sin, cos).
RayParams par_ray_params(ctr_ctx *ctx, const ctr_par_geom *g)
{
    const ParTables &T = ctx->par;
    RayParams r{};
    r.nt = g->nt;
    r.ns = g->ns;
    r.nz = 1;
    r.np = g->n_t;
    r.nx = 1;
    r.nb = g->n_theta;
    r.rt = g->size[0] / g->nt;
    r.rs = g->size[1] / g->ns;
    r.rz = 1.0;
    r.bt = g->size[0];
    r.bs = g->size[1];
    r.bz = 1.0;
    r.sz = 0.5;
    r.parallel = 1;
    r.view = (const double *)T.d_view.ptr;
    r.p = (const double *)T.d_t.ptr;
    r.xi = (const double *)T.d_zero.ptr;
    return r;
}

// Pixel-driven parameters.  single_axis: the detector grid of the float
// t positions (FBPparallel receives single arrays), else of the double ones.
ParBpParams par_bp_params(ctr_ctx *ctx, const ctr_par_geom *g, bool single_axis, unsigned flags)
{
    const ParTables &T = ctx->par;
    ParBpParams P{};
    P.nt = g->nt;
    P.ns = g->ns;
    P.np = g->n_t;
    P.nb = g->n_theta;
    P.rt_d = g->size[0] / g->nt;
    P.rs_d = g->size[1] / g->ns;
    P.ct_d = g->size[0] / 2;
    P.cs_d = g->size[1] / 2;
    P.rt = (float)P.rt_d;
    P.rs = (float)P.rs_d;
    P.ct = (float)P.ct_d;
    P.cs = (float)P.cs_d;
    const int n = g->n_t;
    if (n >= 2) {
        // uniform grid through the end samples
        const double t0 = single_axis ? (double)(float)g->t[0] : g->t[0];
        const double t1 = single_axis ? (double)(float)g->t[n - 1] : g->t[n - 1];
        const double dt = (t1 - t0) / (n - 1);
        P.u_scale = (float)(1.0 / dt);
        P.u_off = (float)(-t0 / dt);
    }
    P.umax = (float)(n - 1);
    P.dtheta = g->n_theta >= 2 ? (float)std::fabs((double)(float)g->theta[1] - (double)(float)g->theta[0]) : 0.0f;
    P.view = (const float *)T.d_vf.ptr;
    P.view_d = (const double *)T.d_view.ptr;
    P.t_d = (const double *)T.d_t.ptr;
    P.accumulate = (flags & CTR_ACCUMULATE) ? 1 : 0;
    return P;
}

uint64_t curved_key(const ctr_curved_geom *g)
{
    uint64_t h = 1469598103934665603ull;
    h = fnv(h, &g->nt, sizeof(int) * 3);
    h = fnv(h, g->size, sizeof(double) * 3);
    h = fnv(h, &g->distance, sizeof(double) * 2);
    h = fnv(h, &g->n_beta, sizeof(int) * 3);
    h = fnv(h, g->beta, sizeof(double) * g->n_beta);
    h = fnv(h, g->gamma, sizeof(double) * g->n_gamma);
    h = fnv(h, g->xi, sizeof(double) * g->n_xi);
    h = fnv(h, &g->beta_start, sizeof(double) * 2);
    return h;
}

// View table, row scales and the per-row spectra of the equiangular
// Shepp-Logan filter (ImplementUI.cu:6-22): taps dgP G[n] with
// G[0] = 2/(pi dgP)^2, G[n] = -2 n^2 / (pi^2 (4n^2-1) sin^2(n dgP)), dgP = dgamma P.
ctr_status curved_tables(ctr_ctx *ctx, const ctr_curved_geom *g)
{
    CurvedTables &T = ctx->curved;
    const uint64_t key = curved_key(g);
    if (T.valid && T.key == key) return CTR_OK;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    T.valid = false;
    const double pi = 3.141592653589793;
    const int nb = g->n_beta, ng = g->n_gamma, nx = g->n_xi;
    T.h_view.resize(2 * (size_t)nb);
    for (int b = 0; b < nb; ++b) {
        T.h_view[2 * b] = (float)std::sin(g->beta[b]);
        T.h_view[2 * b + 1] = (float)std::cos(g->beta[b]);
    }
    T.h_beta.assign(g->beta, g->beta + nb);
    T.h_gam.assign(g->gamma, g->gamma + ng);
    T.h_rowp.resize(nx);
    const int L = fft_filter_length(ng);
    if (!L) return fail(ctx, CTR_ETOOBIG, "no FFT length for %d channels", ng);
    T.fft_len = L;
    T.h_spec.assign((size_t)nx * L, 0.0);
    std::vector<double> cosL(L), taps(ng);
    for (int i = 0; i < L; ++i) cosL[i] = std::cos(2.0 * pi * i / L);
    const double dg = std::fabs(g->gamma[1] - g->gamma[0]);
    for (int x = 0; x < nx; ++x) {
        const double P = g->distance_s2d / std::sqrt(g->distance_s2d * g->distance_s2d + g->xi[x] * g->xi[x]);
        T.h_rowp[x] = P;
        const double dgp = dg * P;
        taps[0] = dgp * 2.0 / ((pi * dgp) * (pi * dgp));
        for (int n = 1; n < ng; ++n) {
            const double sn = std::sin(n * dgp);
            taps[n] = dgp * (-2.0 * n * n / (pi * pi) / (4.0 * n * n - 1.0) / (sn * sn));
        }
        double *H = T.h_spec.data() + (size_t)x * L;
        for (int k = 0; k < L; ++k) {
            double acc = taps[0];
            for (int m = 1; m < ng; ++m) acc += 2.0 * taps[m] * cosL[((long)k * m) % L];
            H[k] = acc / L;
        }
    }
    ctr_status s;
    if ((s = upload(ctx, T.d_view, T.h_view.data(), sizeof(float) * T.h_view.size())) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_rowp, T.h_rowp.data(), sizeof(double) * nx)) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_beta, T.h_beta.data(), sizeof(double) * nb)) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_gam, T.h_gam.data(), sizeof(double) * ng)) != CTR_OK) return s;
    if ((s = upload(ctx, T.d_spec, T.h_spec.data(), sizeof(double) * T.h_spec.size())) != CTR_OK) return s;
    T.key = key;
    T.valid = true;
    return CTR_OK;
}

CurvedBpParams curved_bp_params(ctr_ctx *ctx, const ctr_curved_geom *g, unsigned flags)
{
    CurvedBpParams P{};
    P.nt = g->nt;
    P.ns = g->ns;
    P.nz = g->nz;
    P.ng = g->n_gamma;
    P.nx = g->n_xi;
    P.nb = g->n_beta;
    P.rt = (float)(g->size[0] / g->nt);
    P.rs = (float)(g->size[1] / g->ns);
    P.rz = (float)(g->size[2] / g->nz);
    P.ct = (float)(g->size[0] / 2);
    P.cs = (float)(g->size[1] / 2);
    P.cz = (float)(g->size[2] / 2);
    P.off = (flags & CTR_REF_QUIRKS) ? -0.5f : 0.5f;
    P.D = (float)g->distance;
    P.Dsd = (float)g->distance_s2d;
    // uniform grids through the end samples
    const double dgs = (g->gamma[g->n_gamma - 1] - g->gamma[0]) / (g->n_gamma - 1);
    P.g_scale = (float)(1.0 / dgs);
    P.g_off = (float)(-g->gamma[0] / dgs);
    P.gmax = (float)(g->n_gamma - 1);
    const double dxs = (g->xi[g->n_xi - 1] - g->xi[0]) / (g->n_xi - 1);
    P.x_scale = (float)(1.0 / dxs);
    P.x_off = (float)(-g->xi[0] / dxs);
    P.xmax = (float)(g->n_xi - 1);
    P.dbeta = (float)std::fabs(g->beta[1] - g->beta[0]);
    P.view = (const float *)ctx->curved.d_view.ptr;
    P.accumulate = (flags & CTR_ACCUMULATE) ? 1 : 0;
    return P;
}

// Tile plan and per-(tile, view) footprints of the curved backprojection,
// computed on the device once per geometry and voxel-centre convention.
ctr_status curved_footprints(ctr_ctx *ctx, CurvedBpParams &P)
{
    CurvedTables &T = ctx->curved;
    const size_t n = curved_plan(P);
    uint64_t key = fnv(T.key, &P.off, sizeof P.off);
    key = fnv(key, &P.tz, sizeof P.tz);
    if (!(T.d_foot.ptr && T.foot_key == key)) {
        ctr_status s = buf_reserve(ctx, T.d_foot, n * sizeof(int4));
        if (s != CTR_OK) return s;
        s = check_launch(ctx, launch_curved_footprints(P, (int4 *)T.d_foot.ptr, ctx->stream), "curved footprints");
        if (s != CTR_OK) return s;
        T.foot_key = key;
    }
    P.foot = (const int4 *)T.d_foot.ptr;
    return CTR_OK;
}

// Half-scaled ramp taps times the pitch, two-sided: h2[k] = d * h[|k-(n-1)|].
ctr_status tap_table(ctr_ctx *ctx, int filter, int equiangular, double d, int n)
{
    TapTable &T = ctx->taps;
    uint64_t key = 1469598103934665603ull;
    key = fnv(key, &filter, sizeof filter);
    key = fnv(key, &equiangular, sizeof equiangular);
    key = fnv(key, &d, sizeof d);
    key = fnv(key, &n, sizeof n);
    if (T.valid && T.key == key) return CTR_OK;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    T.valid = false;
    const double pi = 3.141592653589793;
    std::vector<double> h(n);
    for (int k = 0; k < n; ++k) {
        double v;
        if (filter == CTR_FILTER_SL_HALF) {
            v = -1.0 / ((d * pi) * (d * pi)) / (4.0 * k * k - 1.0);
            if (!equiangular) v = (double)(float)v;   // dev_G is float (FDK.cu:6-19)
        } else if (k == 0) {
            v = 1.0 / (8.0 * d * d);
        } else if (k % 2 == 0) {
            v = 0.0;
        } else if (equiangular) {
            const double s = std::sin(k * d);
            v = -1.0 / (2.0 * pi * pi * s * s);
        } else {
            v = -1.0 / (2.0 * pi * pi * (double)k * k * d * d);
        }
        h[k] = v * d;
    }
    T.h.resize(2 * (size_t)n - 1);
    for (int k = 0; k < 2 * n - 1; ++k) T.h[k] = h[std::abs(k - (n - 1))];
    ctr_status s = upload(ctx, T.d, T.h.data(), sizeof(double) * T.h.size());
    if (s != CTR_OK) return s;
    // spectrum of the circular kernel c[m] = c[L-m] = h[m] (m < n) for the FFT
    // path: H[k] = h[0] + 2 sum_m h[m] cos(2 pi k m / L), real because c is even
    T.fft_len = fft_filter_length(n);
    if (T.fft_len) {
        const int L = T.fft_len;
        std::vector<double> cosL(L);
        for (int i = 0; i < L; ++i) cosL[i] = std::cos(2.0 * pi * i / L);
        T.spec.assign(L, 0.0);
        for (int k = 0; k < L; ++k) {
            long double acc = h[0];
            for (int m = 1; m < n; ++m) acc += 2.0L * h[m] * cosL[((long)k * m) % L];
            T.spec[k] = (double)(acc / L);
        }
        s = upload(ctx, T.d_spec, T.spec.data(), sizeof(double) * L);
        if (s != CTR_OK) return s;
    }
    T.key = key;
    T.valid = true;
    return CTR_OK;
}

// Host/device pointer plumbing: in host mode inputs are staged into
// workspaces and results copied back; in device mode pointers pass through.
struct Io {
    ctr_ctx *ctx;
    bool dev;
};

ctr_status stage_in(ctr_ctx *ctx, bool dev, int slot, const float *src, size_t n, const float **out)
{
    if (!src) return fail(ctx, CTR_EINVAL, "input data pointer is NULL");
    if (dev) {
        *out = src;
        return CTR_OK;
    }
    ctr_status s = buf_reserve(ctx, ctx->ws[slot], n * sizeof(float));
    if (s != CTR_OK) return s;
    if (ctx->fault == 3) return fail(ctx, CTR_EFAULT_INJECTED, "CTR_FAULT=copy: host-to-device copy");
    HIP_TRY(ctx, hipMemcpyAsync(ctx->ws[slot].ptr, src, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    *out = (const float *)ctx->ws[slot].ptr;
    return CTR_OK;
}

// Output buffer; with accumulate in host mode the current contents are uploaded.
ctr_status stage_out(ctr_ctx *ctx, bool dev, bool load, int slot, float *dst, size_t n, float **out)
{
    if (!dst) return fail(ctx, CTR_EINVAL, "output data pointer is NULL");
    if (dev) {
        *out = dst;
        return CTR_OK;
    }
    ctr_status s = buf_reserve(ctx, ctx->ws[slot], n * sizeof(float));
    if (s != CTR_OK) return s;
    if (load) {
        if (ctx->fault == 3) return fail(ctx, CTR_EFAULT_INJECTED, "CTR_FAULT=copy: host-to-device copy");
        HIP_TRY(ctx, hipMemcpyAsync(ctx->ws[slot].ptr, dst, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    }
    *out = (float *)ctx->ws[slot].ptr;
    return CTR_OK;
}

ctr_status finish_out(ctr_ctx *ctx, bool dev, const float *dbuf, float *dst, size_t n)
{
    if (dev) return CTR_OK;
    if (ctx->fault == 3) return fail(ctx, CTR_EFAULT_INJECTED, "CTR_FAULT=copy: device-to-host copy");
    HIP_TRY(ctx, hipMemcpyAsync(dst, dbuf, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return CTR_OK;
}

size_t nvox3(const ctr_cone_geom *g) { return (size_t)g->nt * g->ns * g->nz; }
size_t nray3(const ctr_cone_geom *g) { return (size_t)g->n_beta * g->n_p * g->n_xi; }
size_t nvox2(const ctr_fan_geom *g) { return (size_t)g->nt * g->ns; }
size_t nray2(const ctr_fan_geom *g) { return (size_t)g->n_beta * g->n_p; }
size_t npix(const ctr_par_geom *g) { return (size_t)g->nt * g->ns; }
size_t nrayp(const ctr_par_geom *g) { return (size_t)g->n_theta * g->n_t; }
size_t nvoxc(const ctr_curved_geom *g) { return (size_t)g->nt * g->ns * g->nz; }
size_t nrayc(const ctr_curved_geom *g) { return (size_t)g->n_beta * g->n_gamma * g->n_xi; }

#define TRY(expr)                          \
    do {                                   \
        ctr_status s_ = (expr);            \
        if (s_ != CTR_OK) return s_;       \
    } while (0)

ctr_status cone_filter_impl(ctr_ctx *ctx, const ctr_cone_geom *g, int filter, const float *in, float *out)
{
    if (g->n_p > 2048) return fail(ctx, CTR_ETOOBIG, "ramp filter supports at most 2048 detector columns");
    const double d = (double)std::fabs(g->p[1] - g->p[0]);
    TRY(tap_table(ctx, filter, 0, (double)(float)d, g->n_p));
    FilterParams F{};
    F.np = g->n_p;
    F.nx = g->n_xi;
    F.rows = g->n_beta * g->n_xi;
    F.mode = 0;
    F.D = (float)g->distance;
    F.Dd = g->distance;
    F.pd = (const float *)ctx->cone.d_pf.ptr;
    F.xd = (const float *)ctx->cone.d_xf.ptr;
    F.taps = (const double *)ctx->taps.d.ptr;
    F.spec = ctx->taps.fft_len ? (const double *)ctx->taps.d_spec.ptr : nullptr;
    F.fft_len = ctx->taps.fft_len;
    // the FP64 pre-weights once per geometry (at most 32 MiB; beyond, the
    // filter forms the same weights per sample)
    const size_t nw = (size_t)g->n_p * g->n_xi;
    if (nw <= ((size_t)1 << 22)) {
        ConeTables &T = ctx->cone;
        if (T.fw_key != T.key || !T.d_fw.ptr) {
            T.fw_key = 0;
            TRY(buf_reserve(ctx, T.d_fw, nw * sizeof(double)));
            TRY(check_launch(ctx, launch_cone_weights(g->distance, F.pd, F.xd, g->n_p, g->n_xi, (double *)T.d_fw.ptr,
                                                      ctx->stream), "pre-weight table"));
            T.fw_key = T.key;
        }
        F.pw = (const double *)T.d_fw.ptr;
    }
    return check_launch(ctx, launch_row_filter(F, in, out, ctx->stream), "ramp filter");
}

ctr_status fan_filter_impl(ctr_ctx *ctx, const ctr_fan_geom *g, int detector, int filter, const float *in, float *out)
{
    if (g->n_p > 2048) return fail(ctx, CTR_ETOOBIG, "ramp filter supports at most 2048 detector bins");
    double d;
    if (detector == CTR_FAN_EQUIANGULAR) {
        d = std::fabs(std::atan(g->p[1] / g->distance) - std::atan(g->p[0] / g->distance));
    } else {
        d = (double)std::fabs((float)g->p[1] - (float)g->p[0]);
    }
    TRY(tap_table(ctx, filter, detector == CTR_FAN_EQUIANGULAR, d, g->n_p));
    FilterParams F{};
    F.np = g->n_p;
    F.nx = 1;
    F.rows = g->n_beta;
    F.mode = detector == CTR_FAN_EQUIANGULAR ? 1 : 0;
    F.D = (float)g->distance;
    F.Dd = g->distance;
    F.pd = (const float *)ctx->fan.d_pf.ptr;
    F.xd = (const float *)ctx->fan.d_xf.ptr;
    F.colw = (const float *)ctx->fan.d_colw.ptr;
    F.taps = (const double *)ctx->taps.d.ptr;
    F.spec = ctx->taps.fft_len ? (const double *)ctx->taps.d_spec.ptr : nullptr;
    F.fft_len = ctx->taps.fft_len;
    return check_launch(ctx, launch_row_filter(F, in, out, ctx->stream), "fan ramp filter");
}

ctr_status check_filter_args(ctr_ctx *ctx, int filter)
{
    if (filter != CTR_FILTER_SL_HALF && filter != CTR_FILTER_RL_HALF)
        return fail(ctx, CTR_EINVAL, "unknown filter %d", filter);
    return CTR_OK;
}

// CTR_AUTOTUNE: tune a cone geometry on its first use in this context.
ctr_status maybe_tune(ctr_ctx *ctx, const ctr_cone_geom *g, unsigned flags)
{
    if (!(flags & CTR_AUTOTUNE)) return CTR_OK;
    TRY(cone_tables(ctx, g));
    if (ctx->tune.count(ctx->cone.key)) return CTR_OK;
    return ctr_autotune_cone(ctx, g, nullptr);
}

// FBPparallel.cu:6-39 (half-scaled taps, pitch |t[1]-t[0]| of the single t)
ctr_status par_filter_impl(ctr_ctx *ctx, const ctr_par_geom *g, int filter, const float *in, float *out)
{
    if (g->n_t > 2048) return fail(ctx, CTR_ETOOBIG, "ramp filter supports at most 2048 detector bins");
    const double d = (double)std::fabs((float)g->t[1] - (float)g->t[0]);
    TRY(tap_table(ctx, filter, 0, d, g->n_t));
    FilterParams F{};
    F.np = g->n_t;
    F.nx = 1;
    F.rows = g->n_theta;
    F.mode = 2;
    F.taps = (const double *)ctx->taps.d.ptr;
    F.spec = ctx->taps.fft_len ? (const double *)ctx->taps.d_spec.ptr : nullptr;
    F.fft_len = ctx->taps.fft_len;
    return check_launch(ctx, launch_row_filter(F, in, out, ctx->stream), "parallel ramp filter");
}

// ImplementUI.cu:24-78 on the FFT path with one spectrum per detector row
ctr_status curved_filter_impl(ctr_ctx *ctx, const ctr_curved_geom *g, const float *in, float *out)
{
    const CurvedTables &T = ctx->curved;
    FilterParams F{};
    F.np = g->n_gamma;
    F.nx = g->n_xi;
    F.rows = g->n_beta * g->n_xi;
    F.mode = 3;
    F.D = (float)g->distance;
    F.Dd = g->distance;
    F.spec = (const double *)T.d_spec.ptr;
    F.fft_len = T.fft_len;
    F.spec_stride = T.fft_len;
    F.rowp = (const double *)T.d_rowp.ptr;
    F.gamd = (const double *)T.d_gam.ptr;
    F.betad = (const double *)T.d_beta.ptr;
    F.start = g->beta_start;
    F.fan = g->fan_angle;
    return check_launch(ctx, launch_row_filter(F, in, out, ctx->stream), "curved-detector filter");
}

// Shared body of the system-matrix entry points: per-ray counts, their
// prefix sum into row_ptr, then (with col/val) the rows.
ctr_status sysmat_impl(ctr_ctx *ctx, const RayParams &R, size_t nray, int64_t *row_ptr, int32_t *col, float *val,
                       int64_t *nnz, unsigned flags)
{
    if (!nnz || !row_ptr) return fail(ctx, CTR_EINVAL, "row_ptr and nnz must not be NULL");
    if ((col == nullptr) != (val == nullptr)) return fail(ctx, CTR_EINVAL, "col and val must both be given or both NULL");
    if (nray > (size_t)std::numeric_limits<int>::max()) return fail(ctx, CTR_ETOOBIG, "more than 2^31-1 rays");
    const bool dev = flags & CTR_DEVICE_PTRS;
    const int n = (int)nray;
    TRY(buf_reserve(ctx, ctx->ws[kWsG], nray * sizeof(long long)));
    long long *counts = (long long *)ctx->ws[kWsG].ptr;
    long long *rp = (long long *)row_ptr;
    if (!dev) {
        TRY(buf_reserve(ctx, ctx->ws[kWsS], (nray + 1) * sizeof(long long)));
        rp = (long long *)ctx->ws[kWsS].ptr;
    }
    TRY(check_launch(ctx, launch_sysmat_count(R, counts, ctx->stream), "system-matrix row counts"));
    HIP_TRY(ctx, hipMemsetAsync(rp, 0, sizeof(long long), ctx->stream));
    size_t tmp_bytes = 0;
    HIP_TRY(ctx, scan_counts(counts, rp + 1, n, nullptr, tmp_bytes, ctx->stream));
    TRY(buf_reserve(ctx, ctx->ws[kWsD], tmp_bytes > 0 ? tmp_bytes : 1));
    HIP_TRY(ctx, scan_counts(counts, rp + 1, n, ctx->ws[kWsD].ptr, tmp_bytes, ctx->stream));
    long long total = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&total, rp + nray, sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (col) {
        if (*nnz < total)
            return fail(ctx, CTR_EINVAL, "col/val capacity %lld is below the %lld entries", (long long)*nnz, total);
        int *dc = (int *)col;
        float *dv = val;
        if (!dev) {
            TRY(buf_reserve(ctx, ctx->ws[kWsIn], (size_t)(total > 0 ? total : 1) * sizeof(int)));
            TRY(buf_reserve(ctx, ctx->ws[kWsOut], (size_t)(total > 0 ? total : 1) * sizeof(float)));
            dc = (int *)ctx->ws[kWsIn].ptr;
            dv = (float *)ctx->ws[kWsOut].ptr;
        }
        TRY(check_launch(ctx, launch_sysmat_fill(R, rp, dc, dv, ctx->stream), "system-matrix rows"));
        if (!dev && total > 0) {
            HIP_TRY(ctx, hipMemcpyAsync(col, dc, (size_t)total * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(ctx, hipMemcpyAsync(val, dv, (size_t)total * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        }
    }
    if (!dev)
        HIP_TRY(ctx, hipMemcpyAsync(row_ptr, rp, (nray + 1) * sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    *nnz = total;
    return CTR_OK;
}

ctr_status check_equiangular(ctr_ctx *ctx, const ctr_fan_geom *g)
{
    const double g0 = std::atan(g->p[0] / g->distance), d0 = std::atan(g->p[1] / g->distance) - g0;
    for (int i = 1; i < g->n_p; ++i) {
        const double d = std::atan(g->p[i] / g->distance) - std::atan(g->p[i - 1] / g->distance);
        if (std::fabs(d - d0) > 1e-3 * std::fabs(d0))
            return fail(ctx, CTR_EINVAL, "equiangular detector needs uniformly spaced atan(p/D)");
    }
    return CTR_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *ctr_version(void) { return CTR_VERSION_STRING; }

const char *ctr_status_string(int status)
{
    switch (status) {
    case CTR_OK: return "ok";
    case CTR_EINVAL: return "invalid argument";
    case CTR_ENOMEM: return "out of memory";
    case CTR_EHIP: return "HIP runtime error";
    case CTR_ETOOBIG: return "problem too large";
    case CTR_EFAULT_INJECTED: return "injected fault";
    case CTR_ENODEV: return "no HIP device";
    default: return "unknown status";
    }
}

const char *ctr_last_error(const ctr_ctx *ctx)
{
    // a failure raised before the call reached its context (argument and
    // geometry checks) is newer than the context's own last message
    if (!ctx || g_error_seq > ctx->error_seq) return g_last_error.c_str();
    return ctx->last_error.c_str();
}

ctr_status ctr_ctx_create(int device, ctr_ctx **out)
{
    if (!out) return fail(nullptr, CTR_EINVAL, "out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return fail(nullptr, CTR_ENODEV, "no HIP device available (%s)", hipGetErrorString(e));
    }
    if (device < 0 || device >= n) return fail(nullptr, CTR_EINVAL, "device %d out of range (0..%d)", device, n - 1);
    e = hipSetDevice(device);
    if (e != hipSuccess) return fail(nullptr, CTR_EHIP, "hipSetDevice(%d): %s", device, hipGetErrorString(e));
    ctr_ctx *ctx = new ctr_ctx();
    ctx->device = device;
    ctx->scratch_err = noalloc_scratch(device, &ctx->scratch);
    ctx->fault = parse_fault();
    e = hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(nullptr, CTR_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    ctx->stream = ctx->own;
    *out = ctx;
    return CTR_OK;
}

void ctr_ctx_destroy(ctr_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    // networks still alive become detached handles: ctr_cnn_destroy only frees the struct
    while (!ctx->nets.empty()) cnn_release(ctx->nets.back());
    for (auto &b : ctx->ws) buf_free(b);
    for (DevBuf *b : {&ctx->cone.d_view, &ctx->cone.d_p, &ctx->cone.d_xi, &ctx->cone.d_bp, &ctx->cone.d_pf,
                      &ctx->cone.d_xf, &ctx->cone.d_zq, &ctx->fan.d_view, &ctx->fan.d_p, &ctx->fan.d_zero, &ctx->fan.d_pf,
                      &ctx->fan.d_gam, &ctx->fan.d_colw, &ctx->fan.d_xf, &ctx->taps.d, &ctx->taps.d_spec, &ctx->foot_bp.buf, &ctx->foot_adj.buf})
        buf_free(*b);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
}

ctr_status ctr_ctx_set_stream(ctr_ctx *ctx, void *hip_stream)
{
    TRY(need_ctx(ctx));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->stream = (hipStream_t)hip_stream;
    return CTR_OK;
}

void *ctr_ctx_stream(const ctr_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

ctr_status ctr_ctx_synchronize(ctr_ctx *ctx)
{
    TRY(need_ctx(ctx));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return CTR_OK;
}

ctr_status ctr_device_alloc(ctr_ctx *ctx, size_t bytes, void **out)
{
    TRY(need_ctx(ctx));
    if (!out) return fail(ctx, CTR_EINVAL, "out is NULL");
    if (ctx->fault == 1) return fail(ctx, CTR_EFAULT_INJECTED, "CTR_FAULT=alloc");
    hipError_t e = hipMalloc(out, bytes ? bytes : 16);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, CTR_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return CTR_OK;
}

ctr_status ctr_device_free(ctr_ctx *ctx, void *ptr)
{
    TRY(need_ctx(ctx));
    if (ptr) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        HIP_TRY(ctx, hipFree(ptr));
    }
    return CTR_OK;
}

ctr_status ctr_copy_to_device(ctr_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    TRY(need_ctx(ctx));
    if (!dst || !src) return fail(ctx, CTR_EINVAL, "NULL pointer");
    if (ctx->fault == 3) return fail(ctx, CTR_EFAULT_INJECTED, "CTR_FAULT=copy");
    HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return CTR_OK;
}

ctr_status ctr_copy_to_host(ctr_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    TRY(need_ctx(ctx));
    if (!dst || !src) return fail(ctx, CTR_EINVAL, "NULL pointer");
    if (ctx->fault == 3) return fail(ctx, CTR_EFAULT_INJECTED, "CTR_FAULT=copy");
    HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return CTR_OK;
}

ctr_status ctr_cone_geom_check(const ctr_cone_geom *g, int need_uniform_pitch) { return check_cone(g, need_uniform_pitch); }

ctr_status ctr_fan_geom_check(const ctr_fan_geom *g, int need_uniform_pitch) { return check_fan(g, need_uniform_pitch); }
ctr_status ctr_par_geom_check(const ctr_par_geom *g, int need_uniform_pitch) { return check_par(g, need_uniform_pitch); }
ctr_status ctr_curved_geom_check(const ctr_curved_geom *g) { return check_curved(g); }

// ---- cone ------------------------------------------------------------------

ctr_status ctr_cone_fp(ctr_ctx *ctx, const ctr_cone_geom *g, const float *vol, float *proj, unsigned flags)
{
    TRY(check_cone(g, 0));
    TRY(need_ctx(ctx));
    TRY(maybe_tune(ctx, g, flags));
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(cone_tables(ctx, g));
    const float *dv;
    float *dp;
    TRY(stage_in(ctx, dev, kWsIn, vol, nvox3(g), &dv));
    TRY(stage_out(ctx, dev, false, kWsOut, proj, nray3(g), &dp));
    RayParams R = cone_ray_params(ctx, g);
    TRY(fp_transposed(ctx, R, dv, ctx->ws[kWsT], ctx->stream, tuning_for(ctx, ctx->cone.key).fp_transpose));
    TRY(check_launch(ctx, launch_ray_fp(R, dv, dp, ctx->stream), "cone forward projection"));
    return finish_out(ctx, dev, dp, proj, nray3(g));
}

ctr_status ctr_cone_bp_ref(ctr_ctx *ctx, const ctr_cone_geom *g, const float *proj, float *vol, unsigned flags)
{
    TRY(check_cone(g, 1));
    TRY(need_ctx(ctx));
    TRY(maybe_tune(ctx, g, flags));
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(cone_tables(ctx, g));
    const float *dp;
    float *dv;
    TRY(stage_in(ctx, dev, kWsIn, proj, nray3(g), &dp));
    TRY(stage_out(ctx, dev, (flags & CTR_ACCUMULATE) != 0, kWsOut, vol, nvox3(g), &dv));
    BpParams P = cone_bp_params(ctx, g, flags);
    TRY(ensure_footprints(ctx, P, ctx->foot_bp, false));
    TRY(check_launch(ctx, launch_bp_tiled(kBpRef, P, ctx->foot_bp.plan, (const int4 *)ctx->foot_bp.buf.ptr, dp, dv, ctx->stream),
                     "cone reference backprojection"));
    return finish_out(ctx, dev, dv, vol, nvox3(g));
}

ctr_status ctr_cone_bp_adjoint(ctr_ctx *ctx, const ctr_cone_geom *g, const float *proj, float *vol, unsigned flags)
{
    TRY(check_cone(g, 1));
    TRY(need_ctx(ctx));
    TRY(maybe_tune(ctx, g, flags));
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(cone_tables(ctx, g));
    const float *dp;
    float *dv;
    TRY(stage_in(ctx, dev, kWsIn, proj, nray3(g), &dp));
    TRY(stage_out(ctx, dev, (flags & CTR_ACCUMULATE) != 0, kWsOut, vol, nvox3(g), &dv));
    BpParams P = cone_bp_params(ctx, g, flags);
    TRY(ensure_footprints(ctx, P, ctx->foot_adj, true));
    // the kernel reads its input times the ray norm
    TRY(buf_reserve(ctx, ctx->ws[kWsYn], nray3(g) * sizeof(float)));
    float *yn = (float *)ctx->ws[kWsYn].ptr;
    TRY(check_launch(ctx, launch_scale_by_norm(P, dp, yn, ctx->stream), "adjoint input scaling"));
    TRY(check_launch(ctx,
                     launch_adjoint(P, ctx->foot_adj.plan, (const int4 *)ctx->foot_adj.buf.ptr, yn, dv, nullptr, 0,
                                    ctx->stream),
                     "cone adjoint"));
    return finish_out(ctx, dev, dv, vol, nvox3(g));
}

ctr_status ctr_cone_filter(ctr_ctx *ctx, const ctr_cone_geom *g, int filter, const float *proj, float *filtered,
                           unsigned flags)
{
    TRY(check_cone(g, 1));
    TRY(need_ctx(ctx));
    TRY(check_filter_args(ctx, filter));
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(cone_tables(ctx, g));
    const float *dp;
    float *df;
    TRY(stage_in(ctx, dev, kWsIn, proj, nray3(g), &dp));
    TRY(stage_out(ctx, dev, false, kWsTmp, filtered, nray3(g), &df));
    TRY(cone_filter_impl(ctx, g, filter, dp, df));
    return finish_out(ctx, dev, df, filtered, nray3(g));
}

ctr_status ctr_cone_filter_noalloc(ctr_ctx *ctx, const ctr_cone_geom *g, int filter, const float *proj,
                                   float *filtered, unsigned flags)
{
    TRY(check_cone(g, 1));
    TRY(need_ctx(ctx));
    TRY(check_filter_args(ctx, filter));
    if (!(flags & CTR_DEVICE_PTRS))
        return fail(ctx, CTR_EINVAL, "the allocation-free filter takes device pointers (CTR_DEVICE_PTRS)");
    if (!proj || !filtered) return fail(ctx, CTR_EINVAL, "data pointer is NULL");
    if (g->n_p > kNoAllocMaxCols || g->n_xi > kNoAllocMaxRows)
        return fail(ctx, CTR_ETOOBIG, "the allocation-free filter holds at most %d columns and %d rows (got %d x %d)",
                    kNoAllocMaxCols, kNoAllocMaxRows, g->n_p, g->n_xi);
    if (!ctx->scratch)
        return fail(ctx, CTR_EHIP, "no scratch array on device %d: %s", ctx->device, hipGetErrorString(ctx->scratch_err));
    const int n = g->n_p, L = fft_filter_length(n);
    const double d = (double)std::fabs(g->p[1] - g->p[0]);
    unsigned char *base = (unsigned char *)ctx->scratch;
    float *pf = (float *)base, *xf = (float *)(base + kNoAllocOffXf);
    double *taps = (double *)(base + kNoAllocOffTaps), *spec = (double *)(base + kNoAllocOffSpec);
    if (ctx->fault == 3) return fail(ctx, CTR_EFAULT_INJECTED, "CTR_FAULT=copy: host-to-device copy");
    HIP_TRY(ctx, hipMemcpyAsync(pf, g->p, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(xf, g->xi, sizeof(float) * g->n_xi, hipMemcpyHostToDevice, ctx->stream));
    TRY(check_launch(ctx, launch_filter_tables(filter, (double)(float)d, n, L, taps, spec, ctx->stream),
                     "filter tables"));
    FilterParams F{};
    F.np = n;
    F.nx = g->n_xi;
    F.rows = g->n_beta * g->n_xi;
    F.mode = 0;
    F.D = (float)g->distance;
    F.Dd = g->distance;
    F.pd = pf;
    F.xd = xf;
    F.taps = taps;
    F.spec = L ? spec : nullptr;
    F.fft_len = L;
    return check_launch(ctx, launch_row_filter(F, proj, filtered, ctx->stream), "ramp filter");
}

ctr_status ctr_fdk_bp(ctr_ctx *ctx, const ctr_cone_geom *g, const float *filtered, float *vol, unsigned flags)
{
    TRY(check_cone(g, 1));
    TRY(need_ctx(ctx));
    TRY(maybe_tune(ctx, g, flags));
    if (g->n_beta < 2) return fail(ctx, CTR_EINVAL, "FDK needs at least 2 views (dbeta = |beta[1]-beta[0]|)");
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(cone_tables(ctx, g));
    const float *dp;
    float *dv;
    TRY(stage_in(ctx, dev, kWsIn, filtered, nray3(g), &dp));
    TRY(stage_out(ctx, dev, (flags & CTR_ACCUMULATE) != 0, kWsOut, vol, nvox3(g), &dv));
    BpParams P = cone_bp_params(ctx, g, flags);
    TRY(ensure_footprints(ctx, P, ctx->foot_bp, false));
    TRY(check_launch(ctx, launch_bp_tiled(kBpFdk, P, ctx->foot_bp.plan, (const int4 *)ctx->foot_bp.buf.ptr, dp, dv, ctx->stream),
                     "FDK backprojection"));
    return finish_out(ctx, dev, dv, vol, nvox3(g));
}

ctr_status ctr_fdk(ctr_ctx *ctx, const ctr_cone_geom *g, int filter, const float *proj, float *vol, unsigned flags)
{
    TRY(check_cone(g, 1));
    TRY(need_ctx(ctx));
    TRY(maybe_tune(ctx, g, flags));
    TRY(check_filter_args(ctx, filter));
    if (g->n_beta < 2) return fail(ctx, CTR_EINVAL, "FDK needs at least 2 views (dbeta = |beta[1]-beta[0]|)");
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(cone_tables(ctx, g));
    const float *dp;
    float *dv;
    TRY(stage_in(ctx, dev, kWsIn, proj, nray3(g), &dp));
    TRY(stage_out(ctx, dev, (flags & CTR_ACCUMULATE) != 0, kWsOut, vol, nvox3(g), &dv));
    TRY(buf_reserve(ctx, ctx->ws[kWsTmp], nray3(g) * sizeof(float)));
    float *filt = (float *)ctx->ws[kWsTmp].ptr;
    TRY(cone_filter_impl(ctx, g, filter, dp, filt));
    BpParams P = cone_bp_params(ctx, g, flags);
    TRY(ensure_footprints(ctx, P, ctx->foot_bp, false));
    TRY(check_launch(ctx, launch_bp_tiled(kBpFdk, P, ctx->foot_bp.plan, (const int4 *)ctx->foot_bp.buf.ptr, filt, dv, ctx->stream),
                     "FDK backprojection"));
    if (dev) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));   // the filtered workspace is reused by the next call
    return finish_out(ctx, dev, dv, vol, nvox3(g));
}

// ---- fan -------------------------------------------------------------------

ctr_status ctr_fan_fp(ctr_ctx *ctx, const ctr_fan_geom *g, const float *img, float *proj, unsigned flags)
{
    TRY(check_fan(g, 0));
    TRY(need_ctx(ctx));
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(fan_tables(ctx, g));
    const float *di;
    float *dp;
    TRY(stage_in(ctx, dev, kWsIn, img, nvox2(g), &di));
    TRY(stage_out(ctx, dev, false, kWsOut, proj, nray2(g), &dp));
    RayParams R = fan_ray_params(ctx, g);
    TRY(fp_transposed(ctx, R, di, ctx->ws[kWsT], ctx->stream));
    TRY(check_launch(ctx, launch_ray_fp(R, di, dp, ctx->stream), "fan forward projection"));
    return finish_out(ctx, dev, dp, proj, nray2(g));
}

ctr_status ctr_fan_bp_ref(ctr_ctx *ctx, const ctr_fan_geom *g, const float *proj, float *img, unsigned flags)
{
    TRY(check_fan(g, 1));
    TRY(need_ctx(ctx));
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(fan_tables(ctx, g));
    const float *dp;
    float *di;
    TRY(stage_in(ctx, dev, kWsIn, proj, nray2(g), &dp));
    TRY(stage_out(ctx, dev, (flags & CTR_ACCUMULATE) != 0, kWsOut, img, nvox2(g), &di));
    FanParams P = fan_params(ctx, g, flags);
    TRY(check_launch(ctx, launch_fan_bp_ref(P, dp, di, ctx->stream), "fan reference backprojection"));
    return finish_out(ctx, dev, di, img, nvox2(g));
}

ctr_status ctr_fan_bp_adjoint(ctr_ctx *ctx, const ctr_fan_geom *g, const float *proj, float *img, unsigned flags)
{
    TRY(check_fan(g, 0));   // any monotone detector
    TRY(need_ctx(ctx));
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(fan_tables(ctx, g));
    const float *dp;
    float *di;
    TRY(stage_in(ctx, dev, kWsIn, proj, nray2(g), &dp));
    TRY(stage_out(ctx, dev, (flags & CTR_ACCUMULATE) != 0, kWsOut, img, nvox2(g), &di));
    FanParams P = fan_params(ctx, g, flags);
    TRY(check_launch(ctx, launch_fan_adjoint(P, dp, di, nullptr, 0, ctx->stream), "fan adjoint"));
    return finish_out(ctx, dev, di, img, nvox2(g));
}

ctr_status ctr_fan_filter(ctr_ctx *ctx, const ctr_fan_geom *g, int detector, int filter, const float *proj,
                          float *filtered, unsigned flags)
{
    // an equiangular detector is uniform in atan(p/D), not in p
    TRY(check_fan(g, detector == CTR_FAN_FLAT));
    if (g->n_p < 2) return fail(ctx, CTR_EINVAL, "filtering needs at least 2 detector bins");
    TRY(need_ctx(ctx));
    TRY(check_filter_args(ctx, filter));
    if (detector != CTR_FAN_FLAT && detector != CTR_FAN_EQUIANGULAR)
        return fail(ctx, CTR_EINVAL, "unknown detector kind %d", detector);
    if (detector == CTR_FAN_EQUIANGULAR) TRY(check_equiangular(ctx, g));
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(fan_tables(ctx, g));
    const float *dp;
    float *df;
    TRY(stage_in(ctx, dev, kWsIn, proj, nray2(g), &dp));
    TRY(stage_out(ctx, dev, false, kWsTmp, filtered, nray2(g), &df));
    TRY(fan_filter_impl(ctx, g, detector, filter, dp, df));
    return finish_out(ctx, dev, df, filtered, nray2(g));
}

ctr_status ctr_fan_fbp(ctr_ctx *ctx, const ctr_fan_geom *g, int detector, int filter, const float *proj, float *img,
                       unsigned flags)
{
    TRY(check_fan(g, detector == CTR_FAN_FLAT));
    if (g->n_p < 2) return fail(ctx, CTR_EINVAL, "filtered backprojection needs at least 2 detector bins");
    TRY(need_ctx(ctx));
    TRY(check_filter_args(ctx, filter));
    if (detector != CTR_FAN_FLAT && detector != CTR_FAN_EQUIANGULAR)
        return fail(ctx, CTR_EINVAL, "unknown detector kind %d", detector);
    if (detector == CTR_FAN_EQUIANGULAR) TRY(check_equiangular(ctx, g));
    if (g->n_beta < 2) return fail(ctx, CTR_EINVAL, "FBP needs at least 2 views (dbeta = |beta[1]-beta[0]|)");
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(fan_tables(ctx, g));
    const float *dp;
    float *di;
    TRY(stage_in(ctx, dev, kWsIn, proj, nray2(g), &dp));
    TRY(stage_out(ctx, dev, (flags & CTR_ACCUMULATE) != 0, kWsOut, img, nvox2(g), &di));
    TRY(buf_reserve(ctx, ctx->ws[kWsTmp], nray2(g) * sizeof(float)));
    float *filt = (float *)ctx->ws[kWsTmp].ptr;
    TRY(fan_filter_impl(ctx, g, detector, filter, dp, filt));
    FanParams P = fan_params(ctx, g, flags);
    P.equiangular = detector == CTR_FAN_EQUIANGULAR;
    TRY(check_launch(ctx, launch_fan_fbp_bp(P, filt, di, ctx->stream), "fan FBP backprojection"));
    if (dev) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return finish_out(ctx, dev, di, img, nvox2(g));
}

// ---- parallel beam ---------------------------------------------------------

ctr_status ctr_par_fp(ctr_ctx *ctx, const ctr_par_geom *g, const float *img, float *proj, unsigned flags)
{
    TRY(check_par(g, 0));
    TRY(need_ctx(ctx));
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(par_tables(ctx, g));
    const float *di;
    float *dp;
    TRY(stage_in(ctx, dev, kWsIn, img, npix(g), &di));
    TRY(stage_out(ctx, dev, false, kWsOut, proj, nrayp(g), &dp));
    RayParams R = par_ray_params(ctx, g);
    TRY(fp_transposed(ctx, R, di, ctx->ws[kWsT], ctx->stream));
    TRY(check_launch(ctx, launch_ray_fp(R, di, dp, ctx->stream), "parallel forward projection"));
    return finish_out(ctx, dev, dp, proj, nrayp(g));
}

ctr_status ctr_par_bp_adjoint(ctr_ctx *ctx, const ctr_par_geom *g, const float *proj, float *img, unsigned flags)
{
    TRY(check_par(g, 1));
    TRY(need_ctx(ctx));
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(par_tables(ctx, g));
    const float *dp;
    float *di;
    TRY(stage_in(ctx, dev, kWsIn, proj, nrayp(g), &dp));
    TRY(stage_out(ctx, dev, (flags & CTR_ACCUMULATE) != 0, kWsOut, img, npix(g), &di));
    ParBpParams P = par_bp_params(ctx, g, false, flags);
    TRY(check_launch(ctx, launch_par_adjoint(P, dp, di, nullptr, 0, ctx->stream), "parallel adjoint"));
    return finish_out(ctx, dev, di, img, npix(g));
}

ctr_status ctr_par_filter(ctr_ctx *ctx, const ctr_par_geom *g, int filter, const float *proj, float *filtered,
                          unsigned flags)
{
    TRY(check_par(g, 1));
    TRY(need_ctx(ctx));
    TRY(check_filter_args(ctx, filter));
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(par_tables(ctx, g));
    const float *dp;
    float *df;
    TRY(stage_in(ctx, dev, kWsIn, proj, nrayp(g), &dp));
    TRY(stage_out(ctx, dev, false, kWsTmp, filtered, nrayp(g), &df));
    TRY(par_filter_impl(ctx, g, filter, dp, df));
    return finish_out(ctx, dev, df, filtered, nrayp(g));
}

ctr_status ctr_par_fbp(ctr_ctx *ctx, const ctr_par_geom *g, int filter, const float *proj, float *img, unsigned flags)
{
    TRY(check_par(g, 1));
    TRY(need_ctx(ctx));
    TRY(check_filter_args(ctx, filter));
    if (g->n_theta < 2) return fail(ctx, CTR_EINVAL, "FBP needs at least 2 views (dtheta = |theta[1]-theta[0]|)");
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(par_tables(ctx, g));
    const float *dp;
    float *di;
    TRY(stage_in(ctx, dev, kWsIn, proj, nrayp(g), &dp));
    TRY(stage_out(ctx, dev, (flags & CTR_ACCUMULATE) != 0, kWsOut, img, npix(g), &di));
    TRY(buf_reserve(ctx, ctx->ws[kWsTmp], nrayp(g) * sizeof(float)));
    float *filt = (float *)ctx->ws[kWsTmp].ptr;
    TRY(par_filter_impl(ctx, g, filter, dp, filt));
    ParBpParams P = par_bp_params(ctx, g, true, flags);
    TRY(check_launch(ctx, launch_par_fbp_bp(P, filt, di, ctx->stream), "parallel FBP backprojection"));
    if (dev) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));   // the filtered workspace is reused by the next call
    return finish_out(ctx, dev, di, img, npix(g));
}

// ---- curved detector, short scan ------------------------------------------------

ctr_status ctr_curved_filter(ctr_ctx *ctx, const ctr_curved_geom *g, const float *proj, float *filtered,
                             unsigned flags)
{
    TRY(check_curved(g));
    TRY(need_ctx(ctx));
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(curved_tables(ctx, g));
    const float *dp;
    float *df;
    TRY(stage_in(ctx, dev, kWsIn, proj, nrayc(g), &dp));
    TRY(stage_out(ctx, dev, false, kWsTmp, filtered, nrayc(g), &df));
    TRY(curved_filter_impl(ctx, g, dp, df));
    return finish_out(ctx, dev, df, filtered, nrayc(g));
}

ctr_status ctr_curved_bp(ctr_ctx *ctx, const ctr_curved_geom *g, const float *filtered, float *vol, unsigned flags)
{
    TRY(check_curved(g));
    TRY(need_ctx(ctx));
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(curved_tables(ctx, g));
    const float *df;
    float *dv;
    TRY(stage_in(ctx, dev, kWsIn, filtered, nrayc(g), &df));
    TRY(stage_out(ctx, dev, (flags & CTR_ACCUMULATE) != 0, kWsOut, vol, nvoxc(g), &dv));
    CurvedBpParams P = curved_bp_params(ctx, g, flags);
    TRY(curved_footprints(ctx, P));
    TRY(check_launch(ctx, launch_curved_bp(P, df, dv, ctx->stream), "curved-detector backprojection"));
    return finish_out(ctx, dev, dv, vol, nvoxc(g));
}

ctr_status ctr_curved_fdk(ctr_ctx *ctx, const ctr_curved_geom *g, const float *proj, float *vol, unsigned flags)
{
    TRY(check_curved(g));
    TRY(need_ctx(ctx));
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(curved_tables(ctx, g));
    const float *dp;
    float *dv;
    TRY(stage_in(ctx, dev, kWsIn, proj, nrayc(g), &dp));
    TRY(stage_out(ctx, dev, (flags & CTR_ACCUMULATE) != 0, kWsOut, vol, nvoxc(g), &dv));
    TRY(buf_reserve(ctx, ctx->ws[kWsTmp], nrayc(g) * sizeof(float)));
    float *filt = (float *)ctx->ws[kWsTmp].ptr;
    TRY(curved_filter_impl(ctx, g, dp, filt));
    CurvedBpParams P = curved_bp_params(ctx, g, flags);
    TRY(curved_footprints(ctx, P));
    TRY(check_launch(ctx, launch_curved_bp(P, filt, dv, ctx->stream), "curved-detector backprojection"));
    if (dev) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));   // the filtered workspace is reused by the next call
    return finish_out(ctx, dev, dv, vol, nvoxc(g));
}

// ---- Siddon system matrix -----------------------------------------------------

ctr_status ctr_cone_sysmat(ctr_ctx *ctx, const ctr_cone_geom *g, int64_t *row_ptr, int32_t *col, float *val,
                           int64_t *nnz, unsigned flags)
{
    TRY(check_cone(g, 0));
    TRY(need_ctx(ctx));
    TRY(cone_tables(ctx, g));
    return sysmat_impl(ctx, cone_ray_params(ctx, g), nray3(g), row_ptr, col, val, nnz, flags);
}

ctr_status ctr_fan_sysmat(ctr_ctx *ctx, const ctr_fan_geom *g, int64_t *row_ptr, int32_t *col, float *val,
                          int64_t *nnz, unsigned flags)
{
    TRY(check_fan(g, 0));
    TRY(need_ctx(ctx));
    TRY(fan_tables(ctx, g));
    return sysmat_impl(ctx, fan_ray_params(ctx, g), nray2(g), row_ptr, col, val, nnz, flags);
}

ctr_status ctr_par_sysmat(ctr_ctx *ctx, const ctr_par_geom *g, int64_t *row_ptr, int32_t *col, float *val,
                          int64_t *nnz, unsigned flags)
{
    TRY(check_par(g, 0));
    TRY(need_ctx(ctx));
    TRY(par_tables(ctx, g));
    return sysmat_impl(ctx, par_ray_params(ctx, g), nrayp(g), row_ptr, col, val, nnz, flags);
}

// ---- phantoms and metrics -----------------------------------------------------

ctr_status ctr_phantom_ellipsoids(ctr_ctx *ctx, const double *ell, int n_ell, int nt, int ns, int nz, float *vol,
                                  unsigned flags)
{
    if (nt < 1 || ns < 1 || nz < 1) return fail(ctx, CTR_EINVAL, "volume dimensions must be positive");
    if ((double)nt * ns * nz > (double)std::numeric_limits<int>::max())
        return fail(ctx, CTR_ETOOBIG, "more than 2^31-1 voxels");
    if (n_ell < 0 || n_ell > 65536) return fail(ctx, CTR_EINVAL, "n_ell must be in [0, 65536]");
    if (n_ell > 0 && !ell) return fail(ctx, CTR_EINVAL, "ell is NULL");
    std::vector<PhantomEll> h(n_ell > 0 ? n_ell : 1);
    const double deg = 3.141592653589793 / 180.0;
    for (int k = 0; k < n_ell; ++k) {
        const double *e = ell + 8 * (size_t)k;
        for (int j = 0; j < 8; ++j)
            if (!std::isfinite(e[j])) return fail(ctx, CTR_EINVAL, "ellipsoid %d has a non-finite entry", k);
        if (!(e[1] > 0 && e[2] > 0 && e[3] > 0)) return fail(ctx, CTR_EINVAL, "ellipsoid %d has a non-positive semi-axis", k);
        PhantomEll &p = h[k];
        const double ph = e[7] * deg;
        p.A = e[0];
        p.a = e[1];
        p.b = e[2];
        p.c = e[3];
        p.cp = std::cos(ph);
        p.sp = std::sin(ph);
        p.sx = p.cp * e[4] + p.sp * e[5];    // phantom3d.m:97-104: alpha * [x0; y0; z0]
        p.sy = -p.sp * e[4] + p.cp * e[5];
        p.z0 = e[6];
    }
    TRY(need_ctx(ctx));
    PhantomTable &T = ctx->phantom;
    T.h = std::move(h);
    const bool dev = flags & CTR_DEVICE_PTRS;
    TRY(upload(ctx, T.d, T.h.data(), sizeof(PhantomEll) * T.h.size()));
    const size_t n = (size_t)nt * ns * nz;
    float *dv;
    TRY(stage_out(ctx, dev, false, kWsOut, vol, n, &dv));
    TRY(check_launch(ctx, launch_phantom((const PhantomEll *)T.d.ptr, n_ell, nt, ns, nz, dv, ctx->stream),
                     "phantom rasterizer"));
    if (dev) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));   // the table is rewritten by the next call
    return finish_out(ctx, dev, dv, vol, n);
}

// (sum (x - ref)^2, sum ref^2) of two device or host vectors
static ctr_status diff_sums(ctr_ctx *ctx, const float *x, const float *ref, size_t n, unsigned flags, double out[2])
{
    if (!out) return fail(ctx, CTR_EINVAL, "out is NULL");
    if (n == 0) return fail(ctx, CTR_EINVAL, "n must be positive");
    TRY(need_ctx(ctx));
    const bool dev = flags & CTR_DEVICE_PTRS;
    const float *dx, *dr;
    TRY(stage_in(ctx, dev, kWsIn, x, n, &dx));
    TRY(stage_in(ctx, dev, kWsTmp, ref, n, &dr));
    TRY(buf_reserve(ctx, ctx->ws[kWsPart], sizeof(double) * 2 * (4096 + kReduceSlack)));
    TRY(buf_reserve(ctx, ctx->ws[kWsSums], sizeof(double) * 2));
    int nb = 0;
    double *part = (double *)ctx->ws[kWsPart].ptr, *sums = (double *)ctx->ws[kWsSums].ptr;
    TRY(check_launch(ctx, launch_diff_sums(dx, dr, n, part, &nb, ctx->stream), "difference sums"));
    TRY(check_launch(ctx, launch_reduce_partials(part, nb, sums, ctx->stream), "reduce"));
    HIP_TRY(ctx, hipMemcpyAsync(out, sums, sizeof(double) * 2, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return CTR_OK;
}

ctr_status ctr_rrmse(ctr_ctx *ctx, const float *x, const float *ref, size_t n, double *out, unsigned flags)
{
    double s[2];
    TRY(diff_sums(ctx, x, ref, n, flags, s));
    *out = std::sqrt(s[0] / (double)n) / (std::sqrt(s[1] / (double)n) + std::numeric_limits<double>::epsilon());
    return CTR_OK;
}

ctr_status ctr_local_error(ctr_ctx *ctx, const float *cur, const float *prev, size_t n, double *out, unsigned flags)
{
    double s[2];
    TRY(diff_sums(ctx, cur, prev, n, flags, s));
    *out = std::sqrt(s[0]) / std::sqrt(s[1]);
    return CTR_OK;
}

// ---- iterative FDK ------------------------------------------------------------

ctr_status ctr_fdk_error_slice(ctr_ctx *ctx, const ctr_fdkerr_geom *a, const float *vol, float *err, unsigned flags)
{
    if (!a) return fail(ctx, CTR_EINVAL, "arguments are NULL");
    if (a->nt < 1 || a->ns < 1 || a->nz < 1) return fail(ctx, CTR_EINVAL, "volume dimensions must be positive");
    if ((double)a->nt * a->ns * a->nz > (double)(1 << 30)) return fail(ctx, CTR_ETOOBIG, "more than 2^30 voxels");
    for (int k = 0; k < 3; ++k)
        if (!finite_pos(a->size[k])) return fail(ctx, CTR_EINVAL, "size[%d] must be positive and finite", k);
    if (!finite_pos(a->distance)) return fail(ctx, CTR_EINVAL, "distance must be positive and finite");
    if (a->n_beta < 2 || a->n_u < 2 || a->n_p < 2 || a->n_g < 1)
        return fail(ctx, CTR_EINVAL, "need at least 2 views, 2 rows U, 2 samples P and 1 tap");
    if (a->n_g > 32767) return fail(ctx, CTR_ETOOBIG, "more than 32767 taps (tap indices are 16-bit)");
    if (a->t < 0 || a->t >= a->nt) return fail(ctx, CTR_EINVAL, "slice t = %d is outside [0, %d)", a->t, a->nt);
    if ((double)a->n_beta * a->n_u * a->n_p > (double)(1 << 30) || a->n_beta > 65535 || a->nz > 65535)
        return fail(ctx, CTR_ETOOBIG, "n_beta * n_u * n_p above 2^30, or more than 65535 views or slices");
    const double half = 0.5 * std::sqrt(a->size[0] * a->size[0] + a->size[1] * a->size[1]);
    if (!(a->distance > half * (1.0 + 1e-6)))
        return fail(ctx, CTR_EINVAL, "source orbit (distance %g) intersects the volume (half diagonal %g)", a->distance,
                    half);
    if (!a->beta || !a->u || !a->p || !a->g) return fail(ctx, CTR_EINVAL, "an array argument is NULL");
    for (int i = 0; i < a->n_beta; ++i)
        if (!std::isfinite(a->beta[i])) return fail(ctx, CTR_EINVAL, "beta[%d] is not finite", i);
    for (int i = 0; i < a->n_u; ++i)
        if (!std::isfinite(a->u[i])) return fail(ctx, CTR_EINVAL, "u[%d] is not finite", i);
    for (int i = 0; i < a->n_p; ++i)
        if (!std::isfinite(a->p[i])) return fail(ctx, CTR_EINVAL, "p[%d] is not finite", i);
    for (int i = 0; i < a->n_g; ++i)
        if (!std::isfinite(a->g[i])) return fail(ctx, CTR_EINVAL, "g[%d] is not finite", i);
    const float dbeta = std::fabs(a->beta[1] - a->beta[0]), du = std::fabs(a->u[1] - a->u[0]);
    const double pint = std::fabs(a->p[1] - a->p[0]);
    if (!(pint > 0.0)) return fail(ctx, CTR_EINVAL, "p[1] - p[0] must not be 0");
    TRY(need_ctx(ctx));
    const bool dev = flags & CTR_DEVICE_PTRS;
    const int nb = a->n_beta, nu = a->n_u, np = a->n_p;
    const double D = a->distance;
    // host tables, rounded as the reference's float intermediates are
    std::vector<float> cs(nb), sn(nb), tt((size_t)nu * np);
    for (int b = 0; b < nb; ++b) {
        cs[b] = (float)std::cos((double)a->beta[b]);
        sn[b] = (float)std::sin((double)a->beta[b]);
    }
    for (int iu = 0; iu < nu; ++iu)
        for (int ip = 0; ip < np; ++ip) tt[(size_t)iu * np + ip] = (float)(a->p[ip] * (D - (double)a->u[iu]) / D);
    std::vector<double> cp(np);
    for (int ip = 0; ip < np; ++ip) {
        const double P = a->p[ip];
        cp[ip] = std::sqrt(1 + (P / D) * (P / D)) * D / std::sqrt(D * D + P * P) * (double)du * pint * (double)dbeta;
    }
    FdkErrTables &T = ctx->fdkerr;
    TRY(upload(ctx, T.d_cos, cs.data(), sizeof(float) * nb));
    TRY(upload(ctx, T.d_sin, sn.data(), sizeof(float) * nb));
    TRY(upload(ctx, T.d_u, a->u, sizeof(float) * nu));
    TRY(upload(ctx, T.d_t, tt.data(), sizeof(float) * tt.size()));
    TRY(upload(ctx, T.d_p, a->p, sizeof(double) * np));
    TRY(upload(ctx, T.d_g, a->g, sizeof(double) * a->n_g));
    TRY(upload(ctx, T.d_cp, cp.data(), sizeof(double) * np));
    const size_t nvox = (size_t)a->nt * a->ns * a->nz, nout = (size_t)nb * a->nz * a->ns;
    const float *dv;
    float *de;
    TRY(stage_in(ctx, dev, kWsIn, vol, nvox, &dv));
    TRY(stage_out(ctx, dev, false, kWsOut, err, nout, &de));
    FdkErrParams P{};
    P.nt = a->nt;
    P.ns = a->ns;
    P.nz = a->nz;
    P.nb = nb;
    P.nu = nu;
    P.np = np;
    P.ng = a->n_g;
    P.t = a->t;
    P.r0 = a->size[0] / a->nt;
    P.r1 = a->size[1] / a->ns;
    P.r2 = a->size[2] / a->nz;
    P.ct = a->size[0] / 2;
    P.cs = a->size[1] / 2;
    P.cz = a->size[2] / 2;
    P.D = D;
    P.pint = pint;
    P.cosb = (const float *)T.d_cos.ptr;
    P.sinb = (const float *)T.d_sin.ptr;
    P.u = (const float *)T.d_u.ptr;
    P.tab_t = (const float *)T.d_t.ptr;
    P.p = (const double *)T.d_p.ptr;
    P.g = (const double *)T.d_g.ptr;
    P.cp = (const double *)T.d_cp.ptr;
    TRY(buf_reserve(ctx, ctx->ws[kWsG], launch_fdk_error_slice_scratch(P)));
    TRY(check_launch(ctx, launch_fdk_error_slice(P, dv, (int *)ctx->ws[kWsG].ptr, de, ctx->stream),
                     "iterative-FDK error slice"));
    if (dev) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));   // the tables are rewritten by the next call
    return finish_out(ctx, dev, de, err, nout);
}

}  // extern "C"

// ---- SIRT ------------------------------------------------------------------

namespace {

ctr_status sirt_opts_check(ctr_ctx *ctx, const ctr_sirt_opts *o)
{
    if (!o) return fail(ctx, CTR_EINVAL, "options are NULL");
    if (o->n_iter < 0) return fail(ctx, CTR_EINVAL, "n_iter must be >= 0");
    if (!std::isfinite(o->lambda)) return fail(ctx, CTR_EINVAL, "lambda must be finite");
    if (std::isnan(o->min_val) || std::isnan(o->max_val) || o->min_val > o->max_val)
        return fail(ctx, CTR_EINVAL, "clamp bounds must satisfy min_val <= max_val");
    return CTR_OK;
}

// Shared SIRT driver: fp_resid(x, r, partials, &nblk) and adj_update(r, upd)
// close over the geometry.
// V = 1/(A'1 + 1e-9) into ws[kWsVinv] (reserved by the caller), unless it
// already holds V of the operator `key` (0: always recompute)
template <typename AdjOnes>
ctr_status sirt_norm(ctr_ctx *ctx, uint64_t key, size_t nvox, hipStream_t st, AdjOnes adj_ones)
{
    float *vinv = (float *)ctx->ws[kWsVinv].ptr;
    if (key != 0 && ctx->vinv_key == key) return CTR_OK;
    ctx->vinv_key = 0;
    TRY(adj_ones(vinv));
    TRY(check_launch(ctx, launch_reciprocal_eps(vinv, nvox, 1e-9f, st), "reciprocal"));
    ctx->vinv_key = key;
    return CTR_OK;
}

template <typename FpResid, typename AdjOnes, typename AdjUpdate>
ctr_status sirt_loop(ctr_ctx *ctx, size_t nvox, size_t nray, const float *b_in, float *x_in,
                            const ctr_sirt_opts *o, unsigned flags, int nblk_fp, FpResid fp_resid, AdjOnes adj_ones,
                            AdjUpdate adj_update, uint64_t vkey)
{
    const bool dev = flags & CTR_DEVICE_PTRS;
    const float *b;
    float *x;
    TRY(stage_in(ctx, dev, kWsIn, b_in, nray, &b));
    TRY(stage_out(ctx, dev, o->use_x0 != 0, kWsOut, x_in, nvox, &x));
    TRY(buf_reserve(ctx, ctx->ws[kWsR], nray * sizeof(float)));
    TRY(reserve_vinv(ctx, nvox * sizeof(float)));
    const size_t nsum = (size_t)o->n_iter + 2;
    TRY(buf_reserve(ctx, ctx->ws[kWsPart], sizeof(double) * 2 * ((size_t)std::max(nblk_fp, 4096) + kReduceSlack)));
    TRY(buf_reserve(ctx, ctx->ws[kWsSums], sizeof(double) * 2 * nsum));
    float *r = (float *)ctx->ws[kWsR].ptr, *vinv = (float *)ctx->ws[kWsVinv].ptr;
    double *part = (double *)ctx->ws[kWsPart].ptr, *sums = (double *)ctx->ws[kWsSums].ptr;
    if (!o->use_x0) TRY(check_launch(ctx, launch_fill(x, 0.0f, nvox, ctx->stream), "fill"));
    // V = 1/(A'1 + 1e-9): column sums from the adjoint of ones
    TRY(sirt_norm(ctx, vkey, nvox, ctx->stream, adj_ones));
    // ||b||^2 and sum(b) for the residual normalisations (slot n_iter+1)
    int nb2 = 0;
    TRY(check_launch(ctx, launch_sum_sq_sum(b, nray, part, &nb2, ctx->stream), "norm"));
    TRY(check_launch(ctx, launch_reduce_partials(part, nb2, sums + 2 * (nsum - 1), ctx->stream), "reduce"));
    const bool want_res = o->residual_l2 || o->residual_l1;
    SirtUpdate U{vinv, x, o->lambda, o->min_val, o->max_val};
    for (int k = 0; k <= o->n_iter; ++k) {
        if (k == o->n_iter && !want_res) break;
        int nblk = 0;
        TRY(fp_resid(x, b, r, part, &nblk));
        TRY(check_launch(ctx, launch_reduce_partials(part, nblk, sums + 2 * k, ctx->stream), "reduce"));
        if (k == o->n_iter) break;
        TRY(adj_update(r, &U));
    }
    if (want_res) {
        std::vector<double> hs(2 * nsum);
        HIP_TRY(ctx, hipMemcpyAsync(hs.data(), sums, sizeof(double) * 2 * nsum, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        const double bn = std::sqrt(hs[2 * (nsum - 1)]);
        const double bsum = hs[2 * (nsum - 1) + 1];   // SIRT3D.m:59 Norm = sum(R)
        for (int k = 0; k <= o->n_iter; ++k) {
            if (o->residual_l2) o->residual_l2[k] = std::sqrt(hs[2 * k]) / (bn > 0 ? bn : 1.0);
            if (o->residual_l1) o->residual_l1[k] = hs[2 * k + 1] / ((bsum != 0 ? bsum : 1.0) * (double)nvox);
        }
    }
    return finish_out(ctx, dev, x, x_in, nvox);
}

}  // namespace

extern "C" {

ctr_status ctr_sirt3d(ctr_ctx *ctx, const ctr_cone_geom *g, const float *b, float *x, const ctr_sirt_opts *opts,
                      unsigned flags)
{
    TRY(check_cone(g, 1));
    TRY(need_ctx(ctx));
    TRY(maybe_tune(ctx, g, flags));
    TRY(sirt_opts_check(ctx, opts));
    TRY(cone_tables(ctx, g));
    const RayParams R = cone_ray_params(ctx, g);
    const BpParams P = cone_bp_params(ctx, g, 0);
    TRY(ensure_footprints(ctx, P, ctx->foot_adj, true));
    const BpPlan AL = ctx->foot_adj.plan;
    const int4 *AF = (const int4 *)ctx->foot_adj.buf.ptr;
    const bool use_t = tuning_for(ctx, ctx->cone.key).fp_transpose;
    auto fp = [&](const float *xx, const float *bb, float *rr, double *part, int *nblk) {
        RayParams Rx = R;
        TRY(fp_transposed(ctx, Rx, xx, ctx->ws[kWsT], ctx->stream, use_t));
        return check_launch(ctx, launch_ray_fp_residual(Rx, xx, bb, rr, part, nblk, ctx->stream), "SIRT residual");
    };
    auto ones = [&](float *vinv) {
        return check_launch(ctx, launch_adjoint(P, AL, AF, nullptr, vinv, nullptr, 1, ctx->stream), "SIRT column sums");
    };
    auto upd = [&](const float *rr, SirtUpdate *U) {
        return check_launch(ctx, launch_adjoint(P, AL, AF, rr, nullptr, U, 0, ctx->stream), "SIRT update");
    };
    return sirt_loop(ctx, nvox3(g), nray3(g), b, x, opts, flags, ray_fp_blocks(R), fp, ones, upd,
                     fnv(ctx->foot_adj.key, "cone", 4));
}

ctr_status ctr_sirt2d(ctr_ctx *ctx, const ctr_fan_geom *g, const float *b, float *x, const ctr_sirt_opts *opts,
                      unsigned flags)
{
    TRY(check_fan(g, 0));   // any monotone detector
    TRY(need_ctx(ctx));
    TRY(sirt_opts_check(ctx, opts));
    TRY(fan_tables(ctx, g));
    const RayParams R = fan_ray_params(ctx, g);
    const FanParams P = fan_params(ctx, g, 0);
    auto fp = [&](const float *xx, const float *bb, float *rr, double *part, int *nblk) {
        RayParams Rx = R;
        TRY(fp_transposed(ctx, Rx, xx, ctx->ws[kWsT], ctx->stream));
        return check_launch(ctx, launch_ray_fp_residual(Rx, xx, bb, rr, part, nblk, ctx->stream), "SIRT residual");
    };
    auto ones = [&](float *vinv) {
        return check_launch(ctx, launch_fan_adjoint(P, nullptr, vinv, nullptr, 1, ctx->stream), "SIRT column sums");
    };
    auto upd = [&](const float *rr, SirtUpdate *U) {
        return check_launch(ctx, launch_fan_adjoint(P, rr, nullptr, U, 0, ctx->stream), "SIRT update");
    };
    return sirt_loop(ctx, nvox2(g), nray2(g), b, x, opts, flags, ray_fp_blocks(R), fp, ones, upd,
                     fnv(ctx->fan.key, "fan", 3));
}

}  // extern "C"

// ===========================================================================
// Learned SIRT update: the 3-D CNN and the residual-judged loop
// ===========================================================================
namespace {

// Layer table: real channels, padded channel counts of the kernels and the
// offsets of the layer's weights (followed by its bias) in the flat
// FP32 parameter vector.
struct CnnLayer {
    int ci, co;             // real input / output channels
    int ci_pad;             // forward input channels (16 or 32)
    int co_pad_w;           // output channels of the weight-gradient partials (16 or 32)
    int bwd_cin;            // input channels of the input-gradient convolution (0: not needed)
    size_t woff;            // weights [co][ci][27]; bias [co] follows at woff + co*ci*27
};
constexpr CnnLayer kLayers[4] = {
    {2, 32, 16, 32, 0, 0},
    {32, 32, 32, 32, 32, 1760},
    {32, 32, 32, 32, 32, 29440},
    {32, 1, 32, 16, 16, 57120},
};
constexpr size_t kCnnParams = CTR_CNN_NUM_PARAMS;
static_assert(57120 + 32 * 27 + 1 == CTR_CNN_NUM_PARAMS, "parameter layout");

}  // namespace

struct ctr_cnn {
    ctr_ctx *ctx = nullptr;
    DevBuf params, m, v, grad;
    DevBuf wf[4], wb[4];                    // packed BF16 forward / input-gradient weights
    DevBuf a0, a1, a2, a3, out, d4, da, db; // activations and deltas (one patch or volume)
    DevBuf partial, lsum, xin, gin, tin;
    DevBuf tx, tg, full;                    // tiled pass: a tile's input boxes, the whole output
    int step = 0;
};

namespace {

ctr_status cnn_pack(ctr_cnn *net)
{
    ctr_ctx *ctx = net->ctx;
    for (int l = 0; l < 4; ++l) {
        const CnnLayer &L = kLayers[l];
        const float *w = (const float *)net->params.ptr + L.woff;
        TRY(check_launch(ctx,
                         launch_pack_weights(w, L.ci, L.co, L.ci_pad, (__bf16 *)net->wf[l].ptr, L.bwd_cin,
                                             L.bwd_cin ? (__bf16 *)net->wb[l].ptr : nullptr, ctx->stream),
                         "pack weights"));
    }
    return CTR_OK;
}

const float *layer_bias(const ctr_cnn *net, int l)
{
    const CnnLayer &L = kLayers[l];
    return (const float *)net->params.ptr + L.woff + (size_t)L.co * L.ci * 27;
}

ctr_status cnn_reserve(ctr_cnn *net, size_t nvox, bool train)
{
    ctr_ctx *ctx = net->ctx;
    if (train) TRY(buf_reserve(ctx, net->a0, nvox * 16 * sizeof(__bf16)));
    TRY(buf_reserve(ctx, net->a1, nvox * 32 * sizeof(__bf16)));
    TRY(buf_reserve(ctx, net->a2, nvox * 32 * sizeof(__bf16)));
    TRY(buf_reserve(ctx, net->out, nvox * sizeof(float)));
    if (train) {
        TRY(buf_reserve(ctx, net->a3, nvox * 32 * sizeof(__bf16)));
        TRY(buf_reserve(ctx, net->d4, nvox * 16 * sizeof(__bf16)));
        TRY(buf_reserve(ctx, net->da, nvox * 32 * sizeof(__bf16)));
        TRY(buf_reserve(ctx, net->db, nvox * 32 * sizeof(__bf16)));
    }
    return CTR_OK;
}

// Forward pass over one volume; with keep the four activations stay for the
// backward pass (a0..a3), otherwise a1/a2 are reused.  Result in net->out.
ctr_status cnn_forward_dev(ctr_cnn *net, const CnnShape &S, const float *x, const float *g, float g_scale, bool keep)
{
    ctr_ctx *ctx = net->ctx;
    const size_t n = (size_t)S.nx * S.ny * S.nz;
    __bf16 *a0 = (__bf16 *)net->a0.ptr, *a1 = (__bf16 *)net->a1.ptr, *a2 = (__bf16 *)net->a2.ptr;
    __bf16 *a3 = keep ? (__bf16 *)net->a3.ptr : a1;
    // the first layer reads x and g itself; the [V][16] layout feeds the
    // first layer's weight gradient when the backward pass follows
    if (keep) TRY(check_launch(ctx, launch_pack_input(x, g, g_scale, n, a0, ctx->stream), "CNN input"));
    TRY(check_launch(ctx, launch_conv3d_first(x, g, g_scale, (const __bf16 *)net->wf[0].ptr, layer_bias(net, 0), a1,
                                              S, ctx->stream), "conv 1"));
    TRY(check_launch(ctx, launch_conv3d(32, kEpiBiasRelu, a1, (const __bf16 *)net->wf[1].ptr, layer_bias(net, 1),
                                        nullptr, a2, S, ctx->stream), "conv 2"));
    TRY(check_launch(ctx, launch_conv3d(32, kEpiBiasRelu, a2, (const __bf16 *)net->wf[2].ptr, layer_bias(net, 2),
                                        nullptr, a3, S, ctx->stream), "conv 3"));
    TRY(check_launch(ctx, launch_conv3d(32, kEpiLinearF32, a3, (const __bf16 *)net->wf[3].ptr, layer_bias(net, 3),
                                        nullptr, net->out.ptr, S, ctx->stream), "conv 4"));
    return CTR_OK;
}

// The network over a volume in tiles of T^3 output voxels: each tile's input
// box reaches `halo` voxels past it, cropped at the volume's faces (where the
// whole-volume pass pads with zeros as well).  Inside the volume the zero
// padding at a box face reaches 4 voxels inwards (four 3^3 layers), so with
// halo >= 4 every output voxel comes from the same inputs, in the same
// order, as in the whole-volume pass: the result is identical.  Workspaces
// scale with the tile, not the volume.  Result into out (full volume).
ctr_status cnn_forward_tiled(ctr_cnn *net, const CnnShape &S, const float *x, const float *g, float g_scale, float *out,
                             int T, int halo)
{
    ctr_ctx *ctx = net->ctx;
    hipStream_t st = ctx->stream;
    const int bmax_x = std::min(S.nx, T + 2 * halo), bmax_y = std::min(S.ny, T + 2 * halo),
              bmax_z = std::min(S.nz, T + 2 * halo);
    const size_t nbox = (size_t)bmax_x * bmax_y * bmax_z;
    TRY(buf_reserve(ctx, net->tx, nbox * sizeof(float)));
    TRY(buf_reserve(ctx, net->tg, nbox * sizeof(float)));
    TRY(cnn_reserve(net, nbox, false));
    float *tx = (float *)net->tx.ptr, *tg = (float *)net->tg.ptr;
    for (int oz = 0; oz < S.nz; oz += T)
        for (int oy = 0; oy < S.ny; oy += T)
            for (int ox = 0; ox < S.nx; ox += T) {
                const int wx = std::min(T, S.nx - ox), wy = std::min(T, S.ny - oy), wz = std::min(T, S.nz - oz);
                const int x0 = std::max(0, ox - halo), y0 = std::max(0, oy - halo), z0 = std::max(0, oz - halo);
                const int x1 = std::min(S.nx, ox + wx + halo), y1 = std::min(S.ny, oy + wy + halo),
                          z1 = std::min(S.nz, oz + wz + halo);
                const CnnShape B{x1 - x0, y1 - y0, z1 - z0};
                TRY(check_launch(ctx, launch_box_copy(x, S.nx, S.ny, x0, y0, z0, tx, B.nx, B.ny, 0, 0, 0, B.nx, B.ny,
                                                      B.nz, st), "tile input"));
                TRY(check_launch(ctx, launch_box_copy(g, S.nx, S.ny, x0, y0, z0, tg, B.nx, B.ny, 0, 0, 0, B.nx, B.ny,
                                                      B.nz, st), "tile input"));
                TRY(cnn_forward_dev(net, B, tx, tg, g_scale, false));
                TRY(check_launch(ctx, launch_box_copy((const float *)net->out.ptr, B.nx, B.ny, ox - x0, oy - y0,
                                                      oz - z0, out, S.nx, S.ny, ox, oy, oz, wx, wy, wz, st),
                                 "tile output"));
            }
    return CTR_OK;
}

// Backward pass after cnn_forward_dev(keep): accumulates dloss/dparams into
// net->grad and the patch's sum of squared errors into lsum[slot].
ctr_status cnn_backward_dev(ctr_cnn *net, const CnnShape &S, const float *target, int border, float inv_count,
                            int slot)
{
    ctr_ctx *ctx = net->ctx;
    hipStream_t st = ctx->stream;
    __bf16 *a0 = (__bf16 *)net->a0.ptr, *a1 = (__bf16 *)net->a1.ptr, *a2 = (__bf16 *)net->a2.ptr,
           *a3 = (__bf16 *)net->a3.ptr, *d4 = (__bf16 *)net->d4.ptr, *da = (__bf16 *)net->da.ptr,
           *db = (__bf16 *)net->db.ptr;
    float *grad = (float *)net->grad.ptr, *part = (float *)net->partial.ptr;
    double *lpart = (double *)ctx->ws[kWsPart].ptr, *lsum = (double *)net->lsum.ptr;
    int nb = 0;
    TRY(check_launch(ctx, launch_mse_grad((const float *)net->out.ptr, target, S, border, inv_count, d4, lpart, &nb, st),
                     "loss"));
    TRY(check_launch(ctx, launch_reduce_partials(lpart, nb, lsum + 2 * slot, st), "loss sum"));
    const int nblk = wgrad_blocks(S);
    auto wgrad = [&](int l, const __bf16 *act, const __bf16 *delta) {
        const CnnLayer &L = kLayers[l];
        TRY(check_launch(ctx, launch_conv3d_wgrad(L.ci_pad, L.co_pad_w, act, delta, part, nblk, S, st), "weight grad"));
        return check_launch(ctx, launch_reduce_grad(part, nblk, L.ci_pad, L.co_pad_w, L.ci, L.co, grad + L.woff, 1, st),
                            "grad sum");
    };
    TRY(wgrad(3, a3, d4));
    TRY(check_launch(ctx, launch_conv3d(16, kEpiMask, d4, (const __bf16 *)net->wb[3].ptr, nullptr, a3, da, S, st),
                     "input grad 4"));
    TRY(wgrad(2, a2, da));
    TRY(check_launch(ctx, launch_conv3d(32, kEpiMask, da, (const __bf16 *)net->wb[2].ptr, nullptr, a2, db, S, st),
                     "input grad 3"));
    TRY(wgrad(1, a1, db));
    TRY(check_launch(ctx, launch_conv3d(32, kEpiMask, db, (const __bf16 *)net->wb[1].ptr, nullptr, a1, da, S, st),
                     "input grad 2"));
    TRY(wgrad(0, a0, da));
    return CTR_OK;
}

ctr_status need_net(ctr_cnn *net)
{
    if (!net) return fail(nullptr, CTR_EINVAL, "network is NULL");
    if (!net->ctx) return fail(nullptr, CTR_EINVAL, "network's context was destroyed");
    return CTR_OK;
}

void cnn_release(ctr_cnn *net)
{
    ctr_ctx *ctx = net->ctx;
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (DevBuf *b : {&net->params, &net->m, &net->v, &net->grad, &net->a0, &net->a1, &net->a2, &net->a3, &net->out,
                      &net->d4, &net->da, &net->db, &net->partial, &net->lsum, &net->xin, &net->gin, &net->tin,
                      &net->tx, &net->tg, &net->full})
        buf_free(*b);
    for (int l = 0; l < 4; ++l) {
        buf_free(net->wf[l]);
        buf_free(net->wb[l]);
    }
    ctx->nets.erase(std::remove(ctx->nets.begin(), ctx->nets.end(), net), ctx->nets.end());
    net->ctx = nullptr;
}

ctr_status check_cnn_shape(ctr_ctx *ctx, int nx, int ny, int nz)
{
    if (nx < 1 || ny < 1 || nz < 1) return fail(ctx, CTR_EINVAL, "volume dimensions must be positive");
    if ((double)nx * ny * nz * 32 > (double)std::numeric_limits<int>::max() * 8.0)
        return fail(ctx, CTR_ETOOBIG, "volume too large for the CNN workspaces");
    return CTR_OK;
}

}  // namespace

extern "C" {

ctr_status ctr_cnn_create(ctr_ctx *ctx, ctr_cnn **out)
{
    if (!out) return fail(ctx, CTR_EINVAL, "output pointer is NULL");
    *out = nullptr;
    TRY(need_ctx(ctx));
    ctr_cnn *net = new ctr_cnn;
    net->ctx = ctx;
    ctx->nets.push_back(net);
    ctr_status s = CTR_OK;
    for (DevBuf *b : {&net->params, &net->m, &net->v, &net->grad})
        if (s == CTR_OK) s = buf_reserve(ctx, *b, kCnnParams * sizeof(float));
    for (int l = 0; l < 4 && s == CTR_OK; ++l) {
        s = buf_reserve(ctx, net->wf[l], (size_t)27 * kLayers[l].ci_pad * 32 * sizeof(__bf16));
        if (s == CTR_OK && kLayers[l].bwd_cin)
            s = buf_reserve(ctx, net->wb[l], (size_t)27 * kLayers[l].bwd_cin * 32 * sizeof(__bf16));
    }
    if (s == CTR_OK) s = buf_reserve(ctx, net->lsum, sizeof(double) * 2 * 4096);
    if (s == CTR_OK) {
        for (DevBuf *b : {&net->params, &net->m, &net->v, &net->grad})
            if (s == CTR_OK) s = check_launch(ctx, launch_fill((float *)b->ptr, 0.0f, kCnnParams, ctx->stream), "fill");
    }
    if (s == CTR_OK) s = cnn_pack(net);
    if (s != CTR_OK) {
        ctr_cnn_destroy(net);
        return s;
    }
    *out = net;
    return CTR_OK;
}

void ctr_cnn_destroy(ctr_cnn *net)
{
    if (!net) return;
    cnn_release(net);
    delete net;
}

ctr_status ctr_cnn_set_params(ctr_cnn *net, const float *params, size_t n)
{
    TRY(need_net(net));
    ctr_ctx *ctx = net->ctx;
    if (!params || n != kCnnParams) return fail(ctx, CTR_EINVAL, "expected %zu parameters", kCnnParams);
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(params[i])) return fail(ctx, CTR_EINVAL, "parameter %zu is not finite", i);
    TRY(upload(ctx, net->params, params, n * sizeof(float)));
    for (DevBuf *b : {&net->m, &net->v})
        TRY(check_launch(ctx, launch_fill((float *)b->ptr, 0.0f, kCnnParams, ctx->stream), "fill"));
    net->step = 0;
    TRY(cnn_pack(net));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return CTR_OK;
}

ctr_status ctr_cnn_get_params(ctr_cnn *net, float *params, size_t n)
{
    TRY(need_net(net));
    ctr_ctx *ctx = net->ctx;
    if (!params || n != kCnnParams) return fail(ctx, CTR_EINVAL, "expected %zu parameters", kCnnParams);
    HIP_TRY(ctx, hipMemcpyAsync(params, net->params.ptr, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return CTR_OK;
}

ctr_status ctr_cnn_forward(ctr_cnn *net, int nx, int ny, int nz, const float *x, const float *g, float g_scale,
                           float *out, unsigned flags)
{
    TRY(need_net(net));
    ctr_ctx *ctx = net->ctx;
    TRY(check_cnn_shape(ctx, nx, ny, nz));
    if (!std::isfinite(g_scale)) return fail(ctx, CTR_EINVAL, "g_scale must be finite");
    const bool dev = flags & CTR_DEVICE_PTRS;
    const size_t n = (size_t)nx * ny * nz;
    const float *dx, *dg;
    TRY(stage_in(ctx, dev, kWsIn, x, n, &dx));
    TRY(stage_in(ctx, dev, kWsTmp, g, n, &dg));
    if (!out) return fail(ctx, CTR_EINVAL, "output data pointer is NULL");
    TRY(cnn_reserve(net, n, false));
    const CnnShape S{nx, ny, nz};
    TRY(cnn_forward_dev(net, S, dx, dg, g_scale, false));
    if (dev) {
        HIP_TRY(ctx, hipMemcpyAsync(out, net->out.ptr, n * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
        return CTR_OK;
    }
    return finish_out(ctx, false, (const float *)net->out.ptr, out, n);
}

ctr_status ctr_cnn_forward_tiled(ctr_cnn *net, int nx, int ny, int nz, const float *x, const float *g, float g_scale,
                                 float *out, int tile, int halo, unsigned flags)
{
    TRY(need_net(net));
    ctr_ctx *ctx = net->ctx;
    TRY(check_cnn_shape(ctx, nx, ny, nz));
    if (!std::isfinite(g_scale)) return fail(ctx, CTR_EINVAL, "g_scale must be finite");
    if (tile < 8) return fail(ctx, CTR_EINVAL, "tile must be at least 8 voxels");
    if (halo < 4) return fail(ctx, CTR_EINVAL, "halo must be at least 4 voxels (the network's receptive radius)");
    TRY(check_cnn_shape(ctx, tile + 2 * halo, tile + 2 * halo, tile + 2 * halo));
    const bool dev = flags & CTR_DEVICE_PTRS;
    const size_t n = (size_t)nx * ny * nz;
    const float *dx, *dg;
    float *dout;
    TRY(stage_in(ctx, dev, kWsIn, x, n, &dx));
    TRY(stage_in(ctx, dev, kWsTmp, g, n, &dg));
    TRY(stage_out(ctx, dev, false, kWsOut, out, n, &dout));
    TRY(cnn_forward_tiled(net, CnnShape{nx, ny, nz}, dx, dg, g_scale, dout, tile, halo));
    return finish_out(ctx, dev, dout, out, n);
}

ctr_status ctr_cnn_train_step(ctr_cnn *net, int n_patch, int nx, int ny, int nz, const float *x, const float *g,
                              const float *g_scale, const float *target, int border, float step_lr, double *loss,
                              float *grad, unsigned flags)
{
    TRY(need_net(net));
    ctr_ctx *ctx = net->ctx;
    hipStream_t st = ctx->stream;
    TRY(check_cnn_shape(ctx, nx, ny, nz));
    if (n_patch < 1 || n_patch > 4096) return fail(ctx, CTR_EINVAL, "n_patch must be in 1..4096");
    if (!x || !g || !g_scale || !target) return fail(ctx, CTR_EINVAL, "input data pointer is NULL");
    if (border < 0 || 2 * border >= nx || 2 * border >= ny || 2 * border >= nz)
        return fail(ctx, CTR_EINVAL, "border leaves no interior voxels");
    if (!(step_lr >= 0.0f) || !std::isfinite(step_lr)) return fail(ctx, CTR_EINVAL, "step_lr must be >= 0");
    const bool dev = flags & CTR_DEVICE_PTRS;
    const size_t n = (size_t)nx * ny * nz;
    const CnnShape S{nx, ny, nz};
    TRY(cnn_reserve(net, n, true));
    const int nblk = wgrad_blocks(S);
    TRY(buf_reserve(ctx, net->partial, (size_t)nblk * (32 * 32 * 27 + 32) * sizeof(float)));
    TRY(buf_reserve(ctx, ctx->ws[kWsPart], sizeof(double) * 2 * ((n + 255) / 256 + 1 + kReduceSlack)));
    if (!dev) {
        TRY(buf_reserve(ctx, net->xin, n * sizeof(float)));
        TRY(buf_reserve(ctx, net->gin, n * sizeof(float)));
        TRY(buf_reserve(ctx, net->tin, n * sizeof(float)));
    }
    TRY(check_launch(ctx, launch_fill((float *)net->grad.ptr, 0.0f, kCnnParams, st), "fill"));
    const double interior = (double)(nx - 2 * border) * (ny - 2 * border) * (nz - 2 * border) * n_patch;
    const float inv_count = (float)(1.0 / interior);
    for (int i = 0; i < n_patch; ++i) {
        const float *px = x + (size_t)i * n, *pg = g + (size_t)i * n, *pt = target + (size_t)i * n;
        if (!dev) {
            HIP_TRY(ctx, hipMemcpyAsync(net->xin.ptr, px, n * sizeof(float), hipMemcpyHostToDevice, st));
            HIP_TRY(ctx, hipMemcpyAsync(net->gin.ptr, pg, n * sizeof(float), hipMemcpyHostToDevice, st));
            HIP_TRY(ctx, hipMemcpyAsync(net->tin.ptr, pt, n * sizeof(float), hipMemcpyHostToDevice, st));
            px = (const float *)net->xin.ptr;
            pg = (const float *)net->gin.ptr;
            pt = (const float *)net->tin.ptr;
        }
        if (!std::isfinite(g_scale[i])) return fail(ctx, CTR_EINVAL, "g_scale[%d] is not finite", i);
        TRY(cnn_forward_dev(net, S, px, pg, g_scale[i], true));
        TRY(cnn_backward_dev(net, S, pt, border, inv_count, i));
        if (!dev) HIP_TRY(ctx, hipStreamSynchronize(st));   // the staging buffers are reused by the next patch
    }
    std::vector<double> ls(2 * (size_t)n_patch);
    HIP_TRY(ctx, hipMemcpyAsync(ls.data(), net->lsum.ptr, sizeof(double) * ls.size(), hipMemcpyDeviceToHost, st));
    if (grad) HIP_TRY(ctx, hipMemcpyAsync(grad, net->grad.ptr, kCnnParams * sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    double l = 0.0;
    for (int i = 0; i < n_patch; ++i) l += ls[2 * i];
    if (loss) *loss = l / interior;
    if (step_lr > 0.0f) {
        const float b1 = 0.9f, b2 = 0.999f;
        ++net->step;
        const double lr_t = step_lr * std::sqrt(1.0 - std::pow(b2, net->step)) / (1.0 - std::pow(b1, net->step));
        TRY(check_launch(ctx, launch_adam((float *)net->params.ptr, (float *)net->m.ptr, (float *)net->v.ptr,
                                          (const float *)net->grad.ptr, (int)kCnnParams, (float)lr_t, b1, b2, 1e-8f, st),
                         "adam"));
        TRY(cnn_pack(net));
    }
    return CTR_OK;
}

ctr_status ctr_sirt3d_step(ctr_ctx *ctx, const ctr_cone_geom *g, const float *b_in, const float *x_in, float *s_out,
                           float lambda, unsigned flags)
{
    TRY(check_cone(g, 1));
    TRY(need_ctx(ctx));
    TRY(maybe_tune(ctx, g, flags));
    if (!std::isfinite(lambda)) return fail(ctx, CTR_EINVAL, "lambda must be finite");
    TRY(cone_tables(ctx, g));
    hipStream_t st = ctx->stream;
    const RayParams R = cone_ray_params(ctx, g);
    const BpParams P = cone_bp_params(ctx, g, 0);
    TRY(ensure_footprints(ctx, P, ctx->foot_adj, true));
    const BpPlan AL = ctx->foot_adj.plan;
    const int4 *AF = (const int4 *)ctx->foot_adj.buf.ptr;
    const size_t nvox = nvox3(g), nray = nray3(g);
    const bool dev = flags & CTR_DEVICE_PTRS;
    const float *b, *x;
    float *s;
    TRY(stage_in(ctx, dev, kWsIn, b_in, nray, &b));
    TRY(stage_in(ctx, dev, kWsTmp, x_in, nvox, &x));
    TRY(stage_out(ctx, dev, false, kWsOut, s_out, nvox, &s));
    TRY(reserve_vinv(ctx, nvox * sizeof(float)));
    TRY(buf_reserve(ctx, ctx->ws[kWsG], nvox * sizeof(float)));
    TRY(buf_reserve(ctx, ctx->ws[kWsR], nray * sizeof(float)));
    TRY(buf_reserve(ctx, ctx->ws[kWsPart], sizeof(double) * 2 * ((size_t)std::max(ray_fp_blocks(R), 4096) + kReduceSlack)));
    float *vinv = (float *)ctx->ws[kWsVinv].ptr, *gb = (float *)ctx->ws[kWsG].ptr, *r = (float *)ctx->ws[kWsR].ptr;
    double *part = (double *)ctx->ws[kWsPart].ptr;
    TRY(sirt_norm(ctx, fnv(ctx->foot_adj.key, "cone", 4), nvox, st, [&](float *v) {
        return check_launch(ctx, launch_adjoint(P, AL, AF, nullptr, v, nullptr, 1, st), "SIRT column sums");
    }));
    int nblk = 0;
    RayParams Rx = R;
    TRY(fp_transposed(ctx, Rx, x, ctx->ws[kWsT], st, tuning_for(ctx, ctx->cone.key).fp_transpose));
    TRY(check_launch(ctx, launch_ray_fp_residual(Rx, x, b, r, part, &nblk, st), "SIRT residual"));
    TRY(check_launch(ctx, launch_adjoint(P, AL, AF, r, gb, nullptr, 0, st), "SIRT gradient"));
    TRY(check_launch(ctx, launch_sirt_step(vinv, gb, lambda, nvox, s, st), "SIRT step"));
    return finish_out(ctx, dev, s, s_out, nvox);
}

ctr_status ctr_learned_sirt3d(ctr_ctx *ctx, const ctr_cone_geom *g, const float *b_in, float *x_in, ctr_cnn *net,
                              const ctr_learned_opts *o, unsigned flags)
{
    TRY(check_cone(g, 1));
    TRY(need_ctx(ctx));
    TRY(maybe_tune(ctx, g, flags));
    TRY(need_net(net));
    if (net->ctx != ctx) return fail(ctx, CTR_EINVAL, "network belongs to another context");
    if (!o) return fail(ctx, CTR_EINVAL, "options are NULL");
    if (o->n_steps < 0) return fail(ctx, CTR_EINVAL, "n_steps must be >= 0");
    if (!std::isfinite(o->lambda)) return fail(ctx, CTR_EINVAL, "lambda must be finite");
    if (std::isnan(o->min_val) || std::isnan(o->max_val) || o->min_val > o->max_val)
        return fail(ctx, CTR_EINVAL, "clamp bounds must satisfy min_val <= max_val");
    TRY(cone_tables(ctx, g));
    hipStream_t st = ctx->stream;
    const RayParams R = cone_ray_params(ctx, g);
    const BpParams P = cone_bp_params(ctx, g, 0);
    TRY(ensure_footprints(ctx, P, ctx->foot_adj, true));
    const BpPlan AL = ctx->foot_adj.plan;
    const int4 *AF = (const int4 *)ctx->foot_adj.buf.ptr;
    const size_t nvox = nvox3(g), nray = nray3(g);
    const bool dev = flags & CTR_DEVICE_PTRS;
    const float *b;
    float *x;
    TRY(stage_in(ctx, dev, kWsIn, b_in, nray, &b));
    TRY(stage_out(ctx, dev, o->use_x0 != 0, kWsOut, x_in, nvox, &x));
    TRY(reserve_vinv(ctx, nvox * sizeof(float)));
    for (int s : {kWsG, kWsS, kWsD, kWsX2}) TRY(buf_reserve(ctx, ctx->ws[s], nvox * sizeof(float)));
    TRY(buf_reserve(ctx, ctx->ws[kWsR], nray * sizeof(float)));
    if (o->judge_vs_plain) TRY(buf_reserve(ctx, ctx->ws[kWsR2], nray * sizeof(float)));
    TRY(buf_reserve(ctx, ctx->ws[kWsPart], sizeof(double) * 2 * ((size_t)std::max(ray_fp_blocks(R), 4096) + kReduceSlack)));
    TRY(buf_reserve(ctx, ctx->ws[kWsSums], sizeof(double) * 2 * 4));
    if (o->cnn_tile < 0) return fail(ctx, CTR_EINVAL, "cnn_tile must be >= 0");
    // patch-wise network: tiles of cnn_tile^3 with 8-voxel halos, workspaces of one tile
    const bool tiled = o->cnn_tile > 0 && (g->nt > o->cnn_tile || g->ns > o->cnn_tile || g->nz > o->cnn_tile);
    if (tiled) {
        if (o->cnn_tile < 8) return fail(ctx, CTR_EINVAL, "cnn_tile must be 0 or at least 8");
        TRY(buf_reserve(ctx, net->full, nvox * sizeof(float)));
    } else {
        TRY(cnn_reserve(net, nvox, false));
    }
    float *vinv = (float *)ctx->ws[kWsVinv].ptr, *gb = (float *)ctx->ws[kWsG].ptr, *sb = (float *)ctx->ws[kWsS].ptr;
    float *r = (float *)ctx->ws[kWsR].ptr, *alt = (float *)ctx->ws[kWsX2].ptr;
    float *r_spare = o->judge_vs_plain ? (float *)ctx->ws[kWsR2].ptr : nullptr;
    double *part = (double *)ctx->ws[kWsPart].ptr, *sums = (double *)ctx->ws[kWsSums].ptr;
    if (!o->use_x0) TRY(check_launch(ctx, launch_fill(x, 0.0f, nvox, st), "fill"));
    TRY(sirt_norm(ctx, fnv(ctx->foot_adj.key, "cone", 4), nvox, st, [&](float *v) {
        return check_launch(ctx, launch_adjoint(P, AL, AF, nullptr, v, nullptr, 1, st), "SIRT column sums");
    }));
    // reduction of a vector to (sum of squares, sum) read back to the host
    auto host_sums = [&](double *out2) -> ctr_status {
        HIP_TRY(ctx, hipMemcpyAsync(out2, sums, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipStreamSynchronize(st));
        return CTR_OK;
    };
    double h[2];
    int nb = 0;
    TRY(check_launch(ctx, launch_sum_sq_sum(b, nray, part, &nb, st), "norm"));
    TRY(check_launch(ctx, launch_reduce_partials(part, nb, sums, st), "reduce"));
    TRY(host_sums(h));
    const double bn = std::sqrt(h[0]) > 0 ? std::sqrt(h[0]) : 1.0;
    // residual pass: rr = W (b - A xx); returns ||b - A xx|| / ||b||
    auto residual_into = [&](const float *xx, float *rr, double *res) -> ctr_status {
        int nblk = 0;
        RayParams Rx = R;
        TRY(fp_transposed(ctx, Rx, xx, ctx->ws[kWsT], st, tuning_for(ctx, ctx->cone.key).fp_transpose));
        TRY(check_launch(ctx, launch_ray_fp_residual(Rx, xx, b, rr, part, &nblk, st), "SIRT residual"));
        TRY(check_launch(ctx, launch_reduce_partials(part, nblk, sums, st), "reduce"));
        double hh[2];
        TRY(host_sums(hh));
        *res = std::sqrt(hh[0]) / bn;
        return CTR_OK;
    };
    auto residual = [&](const float *xx, double *res) { return residual_into(xx, r, res); };
    float *cur = x;   // buffer holding x_k
    double res_k = 0.0;
    TRY(residual(cur, &res_k));
    if (o->residual_l2) o->residual_l2[0] = res_k;
    const CnnShape S{g->nt, g->ns, g->nz};
    for (int k = 0; k < o->n_steps; ++k) {
        // plain step s = lambda V A' r
        TRY(check_launch(ctx, launch_adjoint(P, AL, AF, r, gb, nullptr, 0, st), "SIRT gradient"));
        TRY(check_launch(ctx, launch_sirt_step(vinv, gb, o->lambda, nvox, sb, st), "SIRT step"));
        TRY(check_launch(ctx, launch_sum_sq_sum(sb, nvox, part, &nb, st), "step norm"));
        TRY(check_launch(ctx, launch_reduce_partials(part, nb, sums, st), "reduce"));
        TRY(host_sums(h));
        const double rms = std::sqrt(h[0] / (double)nvox);
        float *cand = (cur == x) ? alt : x;
        float scale = 0.0f;
        const float *corr = tiled ? (const float *)net->full.ptr : (const float *)net->out.ptr;
        if (rms > 0.0 && std::isfinite(rms)) {
            if (tiled) TRY(cnn_forward_tiled(net, S, cur, sb, (float)(1.0 / rms), (float *)net->full.ptr, o->cnn_tile, 8));
            else TRY(cnn_forward_dev(net, S, cur, sb, (float)(1.0 / rms), false));
            scale = (float)rms;
        }
        TRY(check_launch(ctx, launch_learned_candidate(cur, sb, corr, scale, o->min_val, o->max_val, nvox, cand, st),
                         "learned step"));
        double res_c = 0.0;
        TRY(residual(cand, &res_c));
        int acc = 1;
        if (o->judge_vs_plain && scale > 0.0f) {
            // the plain step's residual (into the spare volume, then r is recomputed for the winner)
            float *plain = (float *)ctx->ws[kWsD].ptr;
            TRY(check_launch(ctx, launch_learned_candidate(cur, sb, sb, 0.0f, o->min_val, o->max_val, nvox, plain, st),
                             "plain step"));
            double res_p = 0.0;
            TRY(residual_into(plain, r_spare, &res_p));
            if (!(res_c < res_p)) {
                acc = 0;
                HIP_TRY(ctx, hipMemcpyAsync(cand, plain, nvox * sizeof(float), hipMemcpyDeviceToDevice, st));
                std::swap(r, r_spare);          // the plain iterate's weighted residual drives the next step
                res_c = res_p;
            }
        } else if (!(res_c < res_k)) {
            // fall back to the plain step
            acc = 0;
            TRY(check_launch(ctx, launch_learned_candidate(cur, sb, sb, 0.0f, o->min_val, o->max_val, nvox, cand, st),
                             "plain step"));
            TRY(residual(cand, &res_c));
        }
        cur = cand;
        res_k = res_c;
        if (o->residual_l2) o->residual_l2[k + 1] = res_k;
        if (o->accepted) o->accepted[k] = acc;
    }
    if (cur != x) HIP_TRY(ctx, hipMemcpyAsync(x, cur, nvox * sizeof(float), hipMemcpyDeviceToDevice, st));
    return finish_out(ctx, dev, x, x_in, nvox);
}

}  // extern "C"

// ===========================================================================
// SIRT emulated by a restricted Chebyshev polynomial
// ===========================================================================
namespace {

// Chebyshev coefficients c_j (j = 0..K) of q(mu) = p_N(lambda mu) on mu in
// [0, 1], interpolated at the K + 1 Chebyshev nodes, where p_N(s) =
// sum_{i<N} (1 - s)^i = (1 - (1 - s)^N) / s: from x0, N plain SIRT steps
// x <- x + lambda (g - M x) (M = V A'W A, spectrum in [0, 1]) give
// x0 + p_N(lambda M) s0 with s0 the plain step at x0.
std::vector<double> emulation_coeffs(int N, int K, double lambda)
{
    const double pi = 3.141592653589793;
    const int n = K + 1;
    std::vector<double> f(n), c(n);
    for (int k = 0; k < n; ++k) {
        const double mu = 0.5 * (std::cos(pi * (k + 0.5) / n) + 1.0), sv = lambda * mu;
        if (sv < 1e-12) f[k] = N;
        else if (sv < 1.0) f[k] = -std::expm1(N * std::log1p(-sv)) / sv;
        else f[k] = (1.0 - std::pow(1.0 - sv, N)) / sv;
    }
    for (int j = 0; j < n; ++j) {
        double acc = 0.0;
        for (int k = 0; k < n; ++k) acc += f[k] * std::cos(pi * j * (k + 0.5) / n);
        c[j] = 2.0 * acc / n;
    }
    c[0] *= 0.5;
    return c;
}

}  // namespace

extern "C" {

ctr_status ctr_sirt3d_emulated(ctr_ctx *ctx, const ctr_cone_geom *g, const float *b_in, float *x_in, ctr_cnn *net,
                               const ctr_emulated_opts *o, unsigned flags)
{
    TRY(check_cone(g, 1));
    TRY(need_ctx(ctx));
    TRY(maybe_tune(ctx, g, flags));
    if (!o) return fail(ctx, CTR_EINVAL, "options are NULL");
    if (o->n_start < 0 || o->n_polish < 0) return fail(ctx, CTR_EINVAL, "iteration counts must be >= 0");
    if (o->degree < 1 || o->degree > 1024) return fail(ctx, CTR_EINVAL, "degree must be in 1..1024");
    if (o->n_target <= o->n_start) return fail(ctx, CTR_EINVAL, "n_target must exceed n_start");
    if (!(o->lambda > 0.0f && o->lambda < 2.0f)) return fail(ctx, CTR_EINVAL, "lambda must be in (0, 2)");
    if (std::isnan(o->min_val) || std::isnan(o->max_val) || o->min_val > o->max_val)
        return fail(ctx, CTR_EINVAL, "clamp bounds must satisfy min_val <= max_val");
    if (!std::isfinite(o->support_threshold)) return fail(ctx, CTR_EINVAL, "support_threshold must be finite");
    if (o->cnn_tile < 0 || (o->cnn_tile > 0 && o->cnn_tile < 8)) return fail(ctx, CTR_EINVAL, "cnn_tile must be 0 or >= 8");
    if (o->support_degree < 0 || o->support_degree > 1024) return fail(ctx, CTR_EINVAL, "support_degree must be in 0..1024");
    if (net) {
        TRY(need_net(net));
        if (net->ctx != ctx) return fail(ctx, CTR_EINVAL, "network belongs to another context");
    }
    TRY(cone_tables(ctx, g));
    hipStream_t st = ctx->stream;
    const RayParams R = cone_ray_params(ctx, g);
    const BpParams P = cone_bp_params(ctx, g, 0);
    TRY(ensure_footprints(ctx, P, ctx->foot_adj, true));
    const BpPlan AL = ctx->foot_adj.plan;
    const int4 *AF = (const int4 *)ctx->foot_adj.buf.ptr;
    const size_t nvox = nvox3(g), nray = nray3(g);
    const bool dev = flags & CTR_DEVICE_PTRS;
    const float *b;
    float *x;
    TRY(stage_in(ctx, dev, kWsIn, b_in, nray, &b));
    TRY(stage_out(ctx, dev, false, kWsOut, x_in, nvox, &x));
    TRY(reserve_vinv(ctx, nvox * sizeof(float)));
    for (int sl : {kWsG, kWsS, kWsD, kWsX2, kWsMask, kWsXs}) TRY(buf_reserve(ctx, ctx->ws[sl], nvox * sizeof(float)));
    TRY(buf_reserve(ctx, ctx->ws[kWsR], nray * sizeof(float)));
    TRY(buf_reserve(ctx, ctx->ws[kWsR2], nray * sizeof(float)));
    TRY(buf_reserve(ctx, ctx->ws[kWsPart], sizeof(double) * 2 * ((size_t)std::max(ray_fp_blocks(R), 4096) + kReduceSlack)));
    TRY(buf_reserve(ctx, ctx->ws[kWsSums], sizeof(double) * 2 * 4));
    const bool tiled = net && o->cnn_tile > 0 && (g->nt > o->cnn_tile || g->ns > o->cnn_tile || g->nz > o->cnn_tile);
    if (net) {
        if (tiled) TRY(buf_reserve(ctx, net->full, nvox * sizeof(float)));
        else TRY(cnn_reserve(net, nvox, false));
    }
    float *vinv = (float *)ctx->ws[kWsVinv].ptr, *gb = (float *)ctx->ws[kWsG].ptr;
    float *cur = (float *)ctx->ws[kWsS].ptr, *prev = (float *)ctx->ws[kWsD].ptr, *acc = (float *)ctx->ws[kWsX2].ptr;
    float *mask = (float *)ctx->ws[kWsMask].ptr, *xs = (float *)ctx->ws[kWsXs].ptr;
    float *r = (float *)ctx->ws[kWsR].ptr, *zero = (float *)ctx->ws[kWsR2].ptr;
    double *part = (double *)ctx->ws[kWsPart].ptr, *sums = (double *)ctx->ws[kWsSums].ptr;
    TRY(sirt_norm(ctx, fnv(ctx->foot_adj.key, "cone", 4), nvox, st, [&](float *v) {
        return check_launch(ctx, launch_adjoint(P, AL, AF, nullptr, v, nullptr, 1, st), "SIRT column sums");
    }));
    TRY(check_launch(ctx, launch_fill(zero, 0.0f, nray, st), "fill"));
    TRY(check_launch(ctx, launch_fill(x, 0.0f, nvox, st), "fill"));
    const bool fpT = tuning_for(ctx, ctx->cone.key).fp_transpose != 0;
    auto host_sums = [&](double *out2) -> ctr_status {
        HIP_TRY(ctx, hipMemcpyAsync(out2, sums, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipStreamSynchronize(st));
        return CTR_OK;
    };
    double h[2];
    int nb = 0;
    TRY(check_launch(ctx, launch_sum_sq_sum(b, nray, part, &nb, st), "norm"));
    TRY(check_launch(ctx, launch_reduce_partials(part, nb, sums, st), "reduce"));
    TRY(host_sums(h));
    const double bn = std::sqrt(h[0]) > 0 ? std::sqrt(h[0]) : 1.0;
    // r = W (bb - A xx); with res, ||b - A xx|| / ||b|| (a host synchronisation)
    auto residual = [&](const float *xx, const float *bb, double *res) -> ctr_status {
        int nblk = 0;
        RayParams Rx = R;
        TRY(fp_transposed(ctx, Rx, xx, ctx->ws[kWsT], st, fpT));
        TRY(check_launch(ctx, launch_ray_fp_residual(Rx, xx, bb, r, part, &nblk, st), "SIRT residual"));
        if (!res) return CTR_OK;
        TRY(check_launch(ctx, launch_reduce_partials(part, nblk, sums, st), "reduce"));
        double hh[2];
        TRY(host_sums(hh));
        *res = std::sqrt(hh[0]) / bn;
        return CTR_OK;
    };
    // n plain clamped SIRT iterations on x (SIRT3D.m:78-82)
    auto plain = [&](int n) -> ctr_status {
        SirtUpdate U{vinv, x, o->lambda, o->min_val, o->max_val};
        for (int k = 0; k < n; ++k) {
            TRY(residual(x, b, nullptr));
            TRY(check_launch(ctx, launch_adjoint(P, AL, AF, r, nullptr, &U, 0, st), "SIRT update"));
        }
        return CTR_OK;
    };
    TRY(plain(o->n_start));
    double res_start = 0.0;
    TRY(residual(x, b, &res_start));
    // support: the network's prediction from (x_k, s_k / rms(s_k)), or the iterate's own
    if (net) {
        TRY(check_launch(ctx, launch_adjoint(P, AL, AF, r, gb, nullptr, 0, st), "SIRT gradient"));
        TRY(check_launch(ctx, launch_sirt_step(vinv, gb, o->lambda, nvox, cur, st), "SIRT step"));
        TRY(check_launch(ctx, launch_sum_sq_sum(cur, nvox, part, &nb, st), "step norm"));
        TRY(check_launch(ctx, launch_reduce_partials(part, nb, sums, st), "reduce"));
        TRY(host_sums(h));
        const double rms = std::sqrt(h[0] / (double)nvox);
        const float gscale = rms > 0.0 && std::isfinite(rms) ? (float)(1.0 / rms) : 0.0f;
        const CnnShape S{g->nt, g->ns, g->nz};
        const float *prob = (const float *)net->out.ptr;
        if (tiled) {
            TRY(cnn_forward_tiled(net, S, x, cur, gscale, (float *)net->full.ptr, o->cnn_tile, 8));
            prob = (const float *)net->full.ptr;
        } else {
            TRY(cnn_forward_dev(net, S, x, cur, gscale, false));
        }
        TRY(check_launch(ctx, launch_threshold_mask(prob, o->support_threshold, nvox, mask, st), "support"));
    } else {
        TRY(check_launch(ctx, launch_threshold_mask(x, o->min_val, nvox, mask, st), "support"));
    }
    // x_s + p_N(lambda M_S) s_S on the support S, N = n_target - n_start, into out
    auto emulate = [&](int degree, float *out) -> ctr_status {
        const std::vector<double> c = emulation_coeffs(o->n_target - o->n_start, degree, o->lambda);
        float *w = cur, *wp = prev;
        TRY(check_launch(ctx, launch_mul(x, mask, nvox, xs, st), "masked start"));
        TRY(residual(xs, b, nullptr));
        TRY(check_launch(ctx, launch_adjoint(P, AL, AF, r, gb, nullptr, 0, st), "SIRT gradient"));
        TRY(check_launch(ctx, launch_cheb_start(gb, vinv, mask, o->lambda, (float)c[0], nvox, w, acc, st), "emulation"));
        for (int j = 0; j < degree; ++j) {
            TRY(residual(w, zero, nullptr));   // r = -W A w_j
            TRY(check_launch(ctx, launch_adjoint(P, AL, AF, r, gb, nullptr, 0, st), "emulation operator"));
            TRY(check_launch(ctx, launch_cheb_step(gb, vinv, mask, w, wp, acc, (float)c[j + 1], j == 0, nvox, st),
                             "emulation"));
            std::swap(w, wp);
        }
        return check_launch(ctx, launch_add_clamp(xs, acc, o->min_val, o->max_val, nvox, out, st), "emulated iterate");
    };
    // a first, low-degree pass refines the support: where its clamped result is positive
    if (o->support_degree > 0) {
        TRY(emulate(o->support_degree, prev));
        TRY(check_launch(ctx, launch_threshold_mask(prev, o->min_val, nvox, mask, st), "support"));
    }
    float *xe = prev;   // the recurrence's buffers are free once it is done
    TRY(emulate(o->degree, xe));
    double res_emul = 0.0;
    TRY(residual(xe, b, &res_emul));
    // the plain projector judges the emulated iterate: kept only if it lowers the residual
    const int fell_back = res_emul < res_start ? 0 : 1;
    if (!fell_back) HIP_TRY(ctx, hipMemcpyAsync(x, xe, nvox * sizeof(float), hipMemcpyDeviceToDevice, st));
    TRY(plain(o->n_polish));
    double res_end = 0.0;
    if (o->residual_l2 || o->fell_back) TRY(residual(x, b, &res_end));
    if (o->residual_l2) {
        o->residual_l2[0] = res_start;
        o->residual_l2[1] = res_emul;
        o->residual_l2[2] = res_end;
    }
    if (o->fell_back) *o->fell_back = fell_back;
    return finish_out(ctx, dev, x, x_in, nvox);
}

}  // extern "C"

// ===========================================================================
// Batched SIRT3D: independent problems advanced by shared launches
// ===========================================================================
namespace {

struct BatchProblem {
    ConeTables T;
    FootCache F;
    RayParams R{};
    BpParams P{};
    size_t nvox = 0, nray = 0;
    DevBuf b, x, r, vinv, part, vt;
    const float *bp = nullptr;
    float *xp = nullptr;
    int fp_blocks = 0, gx = 0, gy = 0;
};

void free_batch_problem(BatchProblem &q)
{
    for (DevBuf *b : {&q.T.d_view, &q.T.d_p, &q.T.d_xi, &q.T.d_bp, &q.T.d_pf, &q.T.d_xf, &q.T.d_zq, &q.F.buf, &q.b, &q.x,
                      &q.r, &q.vinv, &q.part, &q.vt})
        buf_free(*b);
}

// Owns the batch's device memory, graph and stream; releases them on every exit path.
struct BatchRun {
    ctr_ctx *ctx;
    std::vector<BatchProblem> probs;
    DevBuf fp_jobs, red_jobs, adj_jobs, adj_ones_jobs, sums, norms, iter;
    hipStream_t cs = nullptr;
    hipEvent_t ev = nullptr;
    std::vector<hipStream_t> ws;   // streams of the per-problem launches (joined into cs)
    std::vector<hipEvent_t> wev;   // their join events, then the fork event
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    explicit BatchRun(ctr_ctx *c, int n) : ctx(c), probs(n) {}
    ~BatchRun()
    {
        if (cs) (void)hipStreamSynchronize(cs);
        for (hipStream_t w : ws)
            if (w) (void)hipStreamSynchronize(w);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (ev) (void)hipEventDestroy(ev);
        for (hipEvent_t e : wev)
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t w : ws)
            if (w) (void)hipStreamDestroy(w);
        if (cs) (void)hipStreamDestroy(cs);
        (void)hipStreamSynchronize(ctx->stream);
        for (auto &q : probs) free_batch_problem(q);
        for (DevBuf *b : {&fp_jobs, &red_jobs, &adj_jobs, &adj_ones_jobs, &sums, &norms, &iter}) buf_free(*b);
    }
};

}  // namespace

extern "C" {

ctr_status ctr_sirt3d_batched(ctr_ctx *ctx, const ctr_sirt_problem *pr, int n, const ctr_sirt_opts *o, unsigned flags)
{
    TRY(need_ctx(ctx));
    if (!pr || n < 1 || n > 1024) return fail(ctx, CTR_EINVAL, "need 1..1024 problems");
    TRY(sirt_opts_check(ctx, o));
    for (int p = 0; p < n; ++p) {
        TRY(check_cone(&pr[p].geom, 1));
        if (!pr[p].b || !pr[p].x) return fail(ctx, CTR_EINVAL, "problem %d: data pointer is NULL", p);
    }
    for (int p = 0; p < n; ++p) TRY(maybe_tune(ctx, &pr[p].geom, flags));
    const bool dev = flags & CTR_DEVICE_PTRS;
    BatchRun B(ctx, n);
    hipStream_t st = ctx->stream;

    // per-problem tables, footprints and workspaces
    std::vector<FpJob> fj(n);
    std::vector<ReduceJob> rj(n);
    std::vector<std::vector<AdjointJob>> groups(3), ones_groups(3);   // by LDS pitch 40 / 72 / 136
    std::vector<int> group_blocks(3, 0), group_max(3, 0), group_cm(3, 4), group_fit(3, 1);
    const int pitches[3] = {40, 72, 136};
    int fp_total = 0, fp_max = 0;
    for (int p = 0; p < n; ++p) {
        BatchProblem &q = B.probs[p];
        const ctr_cone_geom *g = &pr[p].geom;
        TRY(cone_tables_into(ctx, q.T, g));
        q.R = cone_ray_params_from(q.T, g);
        q.P = cone_bp_params_from(q.T, g, 0);
        TRY(ensure_footprints_key(ctx, q.P, q.F, true, q.T.key));
        q.nvox = nvox3(g);
        q.nray = nray3(g);
        if (dev) {
            q.bp = pr[p].b;
            q.xp = pr[p].x;
        } else {
            TRY(upload(ctx, q.b, pr[p].b, q.nray * sizeof(float)));
            TRY(buf_reserve(ctx, q.x, q.nvox * sizeof(float)));
            if (o->use_x0)
                HIP_TRY(ctx, hipMemcpyAsync(q.x.ptr, pr[p].x, q.nvox * sizeof(float), hipMemcpyHostToDevice, st));
            q.bp = (const float *)q.b.ptr;
            q.xp = (float *)q.x.ptr;
        }
        TRY(buf_reserve(ctx, q.r, q.nray * sizeof(float)));
        TRY(buf_reserve(ctx, q.vinv, q.nvox * sizeof(float)));
        if (tuning_for(ctx, q.T.key).fp_transpose) {
            TRY(buf_reserve(ctx, q.vt, q.nvox * sizeof(float)));
            q.R.volT = (const float *)q.vt.ptr;   // refreshed from x before every residual pass
        }
        q.fp_blocks = fp_job_blocks(q.R, &q.gx, &q.gy);
        TRY(buf_reserve(ctx, q.part, sizeof(double) * 2 * ((size_t)std::max(q.fp_blocks, 4096) + kReduceSlack)));
        fj[p] = FpJob{q.R, q.xp, q.bp, (float *)q.r.ptr, (double *)q.part.ptr, fp_total, q.gx, q.gy};
        fp_total += q.fp_blocks;
        fp_max = std::max(fp_max, q.fp_blocks);
        rj[p] = ReduceJob{(const double *)q.part.ptr, q.fp_blocks};
        const BpPlan &L = q.F.plan;
        BpParams P = q.P;
        P.tiles_t = L.tiles_t;
        P.tiles_s = L.tiles_s;
        P.tiles_z = L.tiles_z;
        const int nwg = L.tiles_t * L.tiles_s * L.tiles_z;
        int gi = 0;
        while (gi < 2 && pitches[gi] != L.wp) ++gi;
        if (pitches[gi] != L.wp) return fail(ctx, CTR_EINVAL, "problem %d: unsupported footprint pitch %d", p, L.wp);
        const SirtUpdate U{(const float *)q.vinv.ptr, q.xp, o->lambda, o->min_val, o->max_val};
        groups[gi].push_back(AdjointJob{P, (const int4 *)q.F.buf.ptr, (const float *)q.r.ptr, nullptr, U,
                                        group_blocks[gi], nwg});
        ones_groups[gi].push_back(AdjointJob{P, (const int4 *)q.F.buf.ptr, nullptr, (float *)q.vinv.ptr, U,
                                             group_blocks[gi], nwg});
        group_blocks[gi] += nwg;
        group_max[gi] = std::max(group_max[gi], nwg);
        group_cm[gi] = L.cm;
        group_fit[gi] = group_fit[gi] && L.fit_all;
        if (!o->use_x0) TRY(check_launch(ctx, launch_fill(q.xp, 0.0f, q.nvox, st), "fill"));
    }
    // job tables in device memory (adjoint groups packed one after another)
    size_t nadj = 0;
    for (auto &gv : groups) nadj += gv.size();
    std::vector<AdjointJob> adj_all, ones_all;
    std::vector<size_t> gofs(3);
    for (int gi = 0; gi < 3; ++gi) {
        gofs[gi] = adj_all.size();
        adj_all.insert(adj_all.end(), groups[gi].begin(), groups[gi].end());
        ones_all.insert(ones_all.end(), ones_groups[gi].begin(), ones_groups[gi].end());
    }
    TRY(upload(ctx, B.fp_jobs, fj.data(), sizeof(FpJob) * n));
    TRY(upload(ctx, B.red_jobs, rj.data(), sizeof(ReduceJob) * n));
    TRY(upload(ctx, B.adj_jobs, adj_all.data(), sizeof(AdjointJob) * nadj));
    TRY(upload(ctx, B.adj_ones_jobs, ones_all.data(), sizeof(AdjointJob) * nadj));
    const size_t nslots = (size_t)o->n_iter + 1;
    TRY(buf_reserve(ctx, B.sums, sizeof(double) * 2 * nslots * n));
    TRY(buf_reserve(ctx, B.norms, sizeof(double) * 2 * n));
    TRY(buf_reserve(ctx, B.iter, sizeof(int)));
    const AdjointJob *dadj = (const AdjointJob *)B.adj_jobs.ptr, *dones = (const AdjointJob *)B.adj_ones_jobs.ptr;
    const FpJob *dfp = (const FpJob *)B.fp_jobs.ptr;
    const ReduceJob *drj = (const ReduceJob *)B.red_jobs.ptr;
    int *diter = (int *)B.iter.ptr;
    double *dsums = (double *)B.sums.ptr;

    // column sums V = 1/(A'1 + 1e-9) and ||b|| per problem
    for (int gi = 0; gi < 3; ++gi)
        if (!groups[gi].empty())
            TRY(check_launch(ctx, launch_adjoint_batched(pitches[gi], group_cm[gi], group_fit[gi] != 0, dones + gofs[gi], (int)groups[gi].size(),
                                                         group_max[gi], 0, 1, st), "batched column sums"));
    for (int p = 0; p < n; ++p) {
        BatchProblem &q = B.probs[p];
        TRY(check_launch(ctx, launch_reciprocal_eps((float *)q.vinv.ptr, q.nvox, 1e-9f, st), "reciprocal"));
        int nb = 0;
        TRY(check_launch(ctx, launch_sum_sq_sum(q.bp, q.nray, (double *)q.part.ptr, &nb, st), "norm"));
        TRY(check_launch(ctx, launch_reduce_partials((double *)q.part.ptr, nb, (double *)B.norms.ptr + 2 * p, st),
                         "reduce"));
    }
    HIP_TRY(ctx, hipMemsetAsync(diter, 0, sizeof(int), st));   // before capture: not part of the graph

    // one iteration, captured on a private stream that follows the context's stream
    HIP_TRY(ctx, hipStreamCreateWithFlags(&B.cs, hipStreamNonBlocking));
    HIP_TRY(ctx, hipEventCreateWithFlags(&B.ev, hipEventDisableTiming));
    HIP_TRY(ctx, hipEventRecord(B.ev, st));
    HIP_TRY(ctx, hipStreamWaitEvent(B.cs, B.ev, 0));
    // CTR_BATCH_STREAMS = K > 0 (default 4): each problem runs its own
    // single-problem kernels (its tuned adjoint variant) on stream p mod K,
    // forked from and joined into the iteration's stream, so that problems
    // overlap; 0: the job-table kernels, one launch per phase for all problems.
    int nstreams = 4;
    if (const char *e = std::getenv("CTR_BATCH_STREAMS")) nstreams = std::max(0, std::atoi(e));
    nstreams = std::min(nstreams, n);
    for (int j = 0; j < nstreams; ++j) {
        B.ws.push_back(nullptr);
        HIP_TRY(ctx, hipStreamCreateWithFlags(&B.ws.back(), hipStreamNonBlocking));
    }
    for (int j = 0; j <= nstreams && nstreams > 0; ++j) {
        B.wev.push_back(nullptr);
        HIP_TRY(ctx, hipEventCreateWithFlags(&B.wev.back(), hipEventDisableTiming));
    }
    // residual sums: slot k of problem p at 2 (p nslots + k) with streams,
    // at 2 (k n + p) with the job-table reduction
    auto sum_index = [&](int p, size_t k) { return nstreams ? 2 * ((size_t)p * nslots + k) : 2 * (k * n + p); };
    auto iteration = [&](hipStream_t s, bool update) -> ctr_status {
        if (nstreams) {
            HIP_TRY(ctx, hipEventRecord(B.wev[nstreams], s));
            for (int j = 0; j < nstreams; ++j) HIP_TRY(ctx, hipStreamWaitEvent(B.ws[j], B.wev[nstreams], 0));
            for (int p = 0; p < n; ++p) {
                const BatchProblem &q = B.probs[p];
                hipStream_t t = B.ws[p % nstreams];
                if (q.R.volT)
                    TRY(check_launch(ctx, launch_transpose_st(q.xp, (float *)q.vt.ptr, q.R.nt, q.R.ns, q.R.nz, t),
                                     "volume transpose"));
                int nblk = 0;
                TRY(check_launch(ctx, launch_ray_fp_residual(q.R, q.xp, q.bp, (float *)q.r.ptr, (double *)q.part.ptr,
                                                             &nblk, t), "residual"));
                TRY(check_launch(ctx, launch_reduce_jobs(drj + p, 1, diter, dsums + sum_index(p, 0), t), "reduce"));
                if (update) {
                    const SirtUpdate U{(const float *)q.vinv.ptr, q.xp, o->lambda, o->min_val, o->max_val};
                    TRY(check_launch(ctx, launch_adjoint(q.P, q.F.plan, (const int4 *)q.F.buf.ptr,
                                                         (const float *)q.r.ptr, nullptr, &U, 0, t), "update"));
                }
            }
            for (int j = 0; j < nstreams; ++j) {
                HIP_TRY(ctx, hipEventRecord(B.wev[j], B.ws[j]));
                HIP_TRY(ctx, hipStreamWaitEvent(s, B.wev[j], 0));
            }
        } else {
            for (int p = 0; p < n; ++p) {
                const BatchProblem &q = B.probs[p];
                if (q.R.volT)
                    TRY(check_launch(ctx, launch_transpose_st(q.xp, (float *)q.vt.ptr, q.R.nt, q.R.ns, q.R.nz, s),
                                     "volume transpose"));
            }
            TRY(check_launch(ctx, launch_ray_fp_residual_batched(dfp, n, fp_max, s), "batched residual"));
            TRY(check_launch(ctx, launch_reduce_jobs(drj, n, diter, dsums, s), "batched reduce"));
            if (update)
                for (int gi = 0; gi < 3; ++gi)
                    if (!groups[gi].empty())
                        TRY(check_launch(ctx, launch_adjoint_batched(pitches[gi], group_cm[gi], group_fit[gi] != 0,
                                                                     dadj + gofs[gi], (int)groups[gi].size(),
                                                                     group_max[gi], 1, 0, s), "batched update"));
        }
        if (!update) return CTR_OK;
        return check_launch(ctx, launch_iter_inc(diter, s), "iteration counter");
    };
    bool use_graph = true;
    if (const char *e = std::getenv("CTR_BATCH_GRAPH")) use_graph = std::atoi(e) != 0;
    if (o->n_iter > 0 && !use_graph) {
        for (int k = 0; k < o->n_iter; ++k) TRY(iteration(B.cs, true));
    } else if (o->n_iter > 0) {
        HIP_TRY(ctx, hipStreamBeginCapture(B.cs, hipStreamCaptureModeThreadLocal));
        ctr_status cap = iteration(B.cs, true);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(B.cs, &graph);
        B.graph = graph;
        if (cap != CTR_OK) return cap;
        HIP_TRY(ctx, e);
        HIP_TRY(ctx, hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
        for (int k = 0; k < o->n_iter; ++k) HIP_TRY(ctx, hipGraphLaunch(B.exec, B.cs));
    }
    TRY(iteration(B.cs, false));   // residual of the final iterate
    HIP_TRY(ctx, hipStreamSynchronize(B.cs));

    std::vector<double> hs(2 * nslots * n), hn(2 * n);
    HIP_TRY(ctx, hipMemcpy(hs.data(), dsums, sizeof(double) * hs.size(), hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(hn.data(), B.norms.ptr, sizeof(double) * hn.size(), hipMemcpyDeviceToHost));
    for (int p = 0; p < n; ++p) {
        if (pr[p].residual_l2) {
            const double bn = std::sqrt(hn[2 * p]) > 0 ? std::sqrt(hn[2 * p]) : 1.0;
            for (size_t k = 0; k < nslots; ++k) pr[p].residual_l2[k] = std::sqrt(hs[sum_index(p, k)]) / bn;
        }
        if (!dev)
            HIP_TRY(ctx, hipMemcpy(pr[p].x, B.probs[p].xp, B.probs[p].nvox * sizeof(float), hipMemcpyDeviceToHost));
    }
    return CTR_OK;
}

}  // extern "C"

// ===========================================================================
// Autotuner
// ===========================================================================
extern "C" {

ctr_status ctr_cone_bp_plan_stats(ctr_ctx *ctx, const ctr_cone_geom *g, int adjoint, long long counts[6])
{
    TRY(check_cone(g, 1));
    TRY(need_ctx(ctx));
    if (!counts) return fail(ctx, CTR_EINVAL, "counts is NULL");
    TRY(cone_tables(ctx, g));
    const BpParams P = cone_bp_params(ctx, g, 0);
    FootCache &fc = adjoint ? ctx->foot_adj : ctx->foot_bp;
    TRY(ensure_footprints(ctx, P, fc, adjoint != 0));
    std::vector<int4> h(fc.plan.nfoot);
    HIP_TRY(ctx, hipMemcpyAsync(h.data(), fc.buf.ptr, h.size() * sizeof(int4), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < 6; ++k) counts[k] = 0;
    for (const int4 &f : h) {
        ++counts[0];
        if (f.w & kFootEmpty) ++counts[1];
        else if (!(f.w & kFootFits)) ++counts[5];
        else if (f.w & kFootInterior) ++counts[2];
        else if (!adjoint && (f.w & kFootRowsIn)) ++counts[3];
        else ++counts[4];
    }
    return CTR_OK;
}

ctr_status ctr_autotune_cone(ctr_ctx *ctx, const ctr_cone_geom *g, ctr_tune_result *out)
{
    TRY(check_cone(g, 1));
    TRY(need_ctx(ctx));
    TRY(cone_tables(ctx, g));
    const uint64_t key = ctx->cone.key;
    hipStream_t st = ctx->stream;
    const size_t nvox = nvox3(g), nray = nray3(g);
    TRY(buf_reserve(ctx, ctx->ws[kWsIn], nvox * sizeof(float)));
    TRY(buf_reserve(ctx, ctx->ws[kWsTmp], nray * sizeof(float)));
    TRY(buf_reserve(ctx, ctx->ws[kWsOut], nvox * sizeof(float)));
    float *vol = (float *)ctx->ws[kWsIn].ptr, *proj = (float *)ctx->ws[kWsTmp].ptr, *dst = (float *)ctx->ws[kWsOut].ptr;
    // addresses and trip counts do not depend on the values: constant data suffices
    TRY(check_launch(ctx, launch_fill(vol, 1.0f, nvox, st), "fill"));
    TRY(check_launch(ctx, launch_fill(proj, 1.0f, nray, st), "fill"));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(ctx, hipEventCreate(&e0));
    hipError_t ec = hipEventCreate(&e1);
    if (ec != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(ctx, CTR_EHIP, "hipEventCreate: %s", hipGetErrorString(ec));
    }
    struct Events {
        hipEvent_t a, b;
        ~Events()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } ev{e0, e1};
    // one warm run, then the mean of three
    auto timed = [&](auto run, double *ms) -> ctr_status {
        TRY(run());
        HIP_TRY(ctx, hipEventRecord(ev.a, st));
        for (int r = 0; r < 3; ++r) TRY(run());
        HIP_TRY(ctx, hipEventRecord(ev.b, st));
        HIP_TRY(ctx, hipEventSynchronize(ev.b));
        float f = 0.0f;
        HIP_TRY(ctx, hipEventElapsedTime(&f, ev.a, ev.b));
        *ms = f / 3.0;
        return CTR_OK;
    };
    ctr_tune_result res{};
    Tuning best;
    const BpParams P = cone_bp_params(ctx, g, 0);
    const int pitches[3] = {64, 72, 80}, cms[3] = {3, 4, 6};
    double bestv = 1e300;
    for (int i = 0; i < 3; ++i) {
        Tuning t = best;
        t.bp_pitch = pitches[i];
        ctx->tune[key] = t;
        TRY(ensure_footprints(ctx, P, ctx->foot_bp, false));
        TRY(timed([&]() {
            return check_launch(ctx, launch_bp_tiled(kBpFdk, P, ctx->foot_bp.plan, (const int4 *)ctx->foot_bp.buf.ptr,
                                                     proj, dst, st), "FDK backprojection");
        }, &res.bp_ms[i]));
        if (res.bp_ms[i] < bestv) {
            bestv = res.bp_ms[i];
            res.bp_pitch = pitches[i];
        }
    }
    best.bp_pitch = res.bp_pitch;
    bestv = 1e300;
    for (int i = 0; i < 3; ++i) {
        Tuning t = best;
        t.adj_cm = cms[i];
        ctx->tune[key] = t;
        TRY(ensure_footprints(ctx, P, ctx->foot_adj, true));
        TRY(timed([&]() {
            return check_launch(ctx, launch_adjoint(P, ctx->foot_adj.plan, (const int4 *)ctx->foot_adj.buf.ptr, proj,
                                                    dst, nullptr, 0, st), "adjoint");
        }, &res.adj_ms[i]));
        if (res.adj_ms[i] < bestv) {
            bestv = res.adj_ms[i];
            res.adj_cm = cms[i];
        }
    }
    best.adj_cm = res.adj_cm;
    for (int i = 0; i < 2; ++i) {
        TRY(timed([&]() {
            RayParams R = cone_ray_params(ctx, g);
            TRY(fp_transposed(ctx, R, vol, ctx->ws[kWsT], st, i == 1));
            return check_launch(ctx, launch_ray_fp(R, vol, proj, st), "forward projection");
        }, &res.fp_ms[i]));
    }
    res.fp_transpose = res.fp_ms[1] < res.fp_ms[0] ? 1 : 0;
    best.fp_transpose = res.fp_transpose;
    ctx->tune[key] = best;
    if (out) *out = res;
    return CTR_OK;
}

}  // extern "C"
