// k_fan.hip - fan-beam pixel-driven kernels for gfx950: the reference
// backprojection (linear sample x pixel chord), the exact adjoint of the
// Siddon fan projector, and the flat / equiangular FBP backprojection.
// One thread per pixel loops over all views (the reference launches one
// kernel per view, BackprojectionFan2D.cu:243-314); 2-D problems are small,
// so no LDS staging.
#include "ctr_internal.h"

namespace ctr {
namespace {

// Reference fan backprojection (BackprojectionFan2D.cu:7-116).
__global__ __launch_bounds__(256) void k_fan_bp_ref(const float *__restrict__ proj, float *__restrict__ img, FanParams P)
{
    const int i = blockIdx.x * 16 + (threadIdx.x & 15), j = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= P.nt || j >= P.ns) return;
    const float x = ((float)i + 0.5f) * P.rt - P.ct, y = ((float)j + 0.5f) * P.rs - P.cs;
    float acc = 0.0f;
    for (int b = 0; b < P.nb; ++b) {
        float val = 0.0f, chord;
        const float *r = proj + (size_t)b * P.np;
        if (!P.quirks) {
            const float sb = (float)P.view[4 * b + 0], cb = (float)P.view[4 * b + 1];
            const float u = x * cb + y * sb, s = -x * sb + y * cb;
            const float U = (P.D - s) / P.D;
            const float uu = fmaf(u / U, P.u_scale, P.u_off);
            if (uu >= 0.0f && uu < P.umax) {
                const int iu = min((int)floorf(uu), (int)P.umax - 1);
                const float fu = uu - (float)iu;
                val = fmaf(fu, r[iu + 1] - r[iu], r[iu]);
            }
            // the line from the source through the pixel centre: chord = 2 |d| min(h/|d_axis|)
            const float dx = x + P.ct - (float)P.view[4 * b + 2], dy = y + P.cs - (float)P.view[4 * b + 3];
            chord = 2.0f * sqrtf(dx * dx + dy * dy) * fminf(0.5f * P.rt / fabsf(dx), 0.5f * P.rs / fabsf(dy));
        } else {
            // the reference's arithmetic in FP64 (its face test is discontinuous, so
            // rounding must follow it): entry face tested against 0 instead of the
            // pixel's low edge (:95-101), distance taken to the ray's far end point (:109)
            const double Db = P.D, st = P.view[4 * b + 2], ss = P.view[4 * b + 3];
            const double sbd = P.view[4 * b + 0], cbd = P.view[4 * b + 1];
            const double ct = P.ct_d, cs = P.cs_d;
            const double xd = ((double)i + 0.5) * P.rt_d - ct, yd = ((double)j + 0.5) * P.rs_d - cs;
            const double ud = xd * cbd + yd * sbd, sd = -xd * sbd + yd * cbd;
            const double pd = ud / ((Db - sd) / Db);
            const double uu = pd * P.u_scale_d + P.u_off_d;
            if (uu >= 0.0 && uu < (double)P.umax) {
                const int iu = min((int)floor(uu), (int)P.umax - 1);
                const double fu = uu - (double)iu;
                val = (float)((1.0 - fu) * (double)r[iu] + fu * (double)r[iu + 1]);
            }
            const double th = atan(pd / Db), Smax = 2.0 * Db;
            const double te = ct + Smax * sin(th) * cbd - (Db - Smax * cos(th)) * sbd;
            const double se = cs + Smax * sin(th) * sbd + (Db - Smax * cos(th)) * cbd;
            double t2s = (se - ss) / ((te - st) == 0.0 ? 1e-15 : (te - st));
            double s2t = (te - st) / ((se - ss) == 0.0 ? 1e-15 : (se - ss));
            t2s = fmin(fmax(t2s, -1e16), 1e16);
            s2t = fmin(fmax(s2t, -1e16), 1e16);
            const double tlo = (double)i * P.rt_d, thi = (double)(i + 1) * P.rt_d;
            const double slo = (double)j * P.rs_d, shi = (double)(j + 1) * P.rs_d;
            const double tlo_s = ss + (tlo - st) * t2s, slo_t = st + (slo - ss) * s2t;
            double T1, S1;
            if (tlo_s >= 0.0 && tlo_s <= shi) {
                T1 = tlo;
                S1 = tlo_s;
            } else if (slo_t >= 0.0 && slo_t <= thi) {
                T1 = slo_t;
                S1 = slo;
            } else {
                continue;
            }
            chord = (float)(2.0 * sqrt((T1 - te) * (T1 - te) + (S1 - se) * (S1 - se)));
        }
        acc = fmaf(val, chord, acc);
    }
    const size_t o = (size_t)j * P.nt + i;
    img[o] = P.accumulate ? img[o] + acc : acc;
}

// Exact adjoint of the fan Siddon projector.  Per pixel and view the
// candidate rays are the bins whose positions fall inside the projection of
// the pixel's square, found by binary search over the monotone bin
// positions (so equiangular and other non-uniform detectors work).  The ray
// S + lam (a, b) through bin p, a = p cos + D sin, b = p sin - D cos, meets
// the pixel for lam in the intersection of its t and s slabs; its length is
// dlam sqrt(p^2 + D^2).  The slab parameters are measured from the source,
// D away, so they are formed in FP64 with the FP64 bin positions of the
// forward projector: a 2-D image is small enough for that.

// First index c of the monotone array pos[0..n) with pos[c] >= x (ascending)
// or pos[c] <= x (descending); n when there is none.
__device__ __forceinline__ int bin_lower(const float *__restrict__ pos, int n, bool asc, float x)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const float v = pos[mid];
        if (asc ? v < x : v > x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_fan_adjoint(const float *__restrict__ proj, float *__restrict__ img,
                                                     FanParams P, SirtUpdate U, int sirt, int ones)
{
    const int i = blockIdx.x * 16 + (threadIdx.x & 15), j = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= P.nt || j >= P.ns) return;
    const double tlo = (double)i * P.rt_d, thi = tlo + P.rt_d, slo = (double)j * P.rs_d, shi = slo + P.rs_d;
    const double D = (double)P.D;
    const bool asc = P.pf[P.np - 1] >= P.pf[0];
    float acc = 0.0f;
    for (int b = 0; b < P.nb; ++b) {
        const double sb = P.view[4 * b + 0], cb = P.view[4 * b + 1];
        const double st = P.view[4 * b + 2], ss = P.view[4 * b + 3];
        float umin = kInf, umax = -kInf;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float x = (float)(((c & 1) ? thi : tlo) - P.ct_d), y = (float)(((c & 2) ? shi : slo) - P.cs_d);
            const float u = x * (float)cb + y * (float)sb, s = -x * (float)sb + y * (float)cb;
            const float pd = u * P.D / (P.D - s);
            umin = fminf(umin, pd);
            umax = fmaxf(umax, pd);
        }
        // a conservative bin range: a bin outside the pixel's shadow gets length 0 below
        const float eu = 1e-3f + 1e-5f * fmaxf(fabsf(umin), fabsf(umax));
        int c0, c1;
        if (asc) {
            c0 = bin_lower(P.pf, P.np, true, umin - eu);
            c1 = bin_lower(P.pf, P.np, true, umax + eu) - 1;   // last bin <= umax + eu (ties included by the margin)
        } else {
            c0 = bin_lower(P.pf, P.np, false, umax + eu);
            c1 = bin_lower(P.pf, P.np, false, umin - eu) - 1;
        }
        c0 = max(c0, 0);
        c1 = min(c1, P.np - 1);
        const float *r = proj + (size_t)b * P.np;
        for (int c = c0; c <= c1; ++c) {
            const double p = P.pdbl[c];
            const double a = p * cb + D * sb, bq = p * sb - D * cb;
            double lo = -INFINITY, hi = INFINITY;
            if (a != 0.0) {
                const double l1 = (tlo - st) / a, l2 = (thi - st) / a;
                lo = fmax(lo, fmin(l1, l2));
                hi = fmin(hi, fmax(l1, l2));
            } else if (!(st >= tlo && st < thi)) {
                continue;
            }
            if (bq != 0.0) {
                const double l1 = (slo - ss) / bq, l2 = (shi - ss) / bq;
                lo = fmax(lo, fmin(l1, l2));
                hi = fmin(hi, fmax(l1, l2));
            } else if (!(ss >= slo && ss < shi)) {
                continue;
            }
            if (!(hi > lo)) continue;
            const float len = (float)((hi - lo) * sqrt(p * p + D * D));
            acc = fmaf(len, ones ? 1.0f : r[c], acc);
        }
    }
    const size_t o = (size_t)j * P.nt + i;
    if (sirt) {
        const float xn = fmaf(U.lambda * U.vinv[o], acc, U.x[o]);
        U.x[o] = fminf(fmaxf(xn, U.lo), U.hi);
    } else {
        img[o] = P.accumulate ? img[o] + acc : acc;
    }
}

// Fan FBP backprojection.
//   flat (FBPfan.cu:45-96): sum_b lerp(q)(p/U) / U^2 * dbeta, U = (D - s)/D
//   equiangular (FanProjection.m:299-324): sum_b lerp(q)(gamma) / L^2 * dbeta,
//   gamma = atan2(u, D - s), L^2 = u^2 + (D - s)^2
__global__ __launch_bounds__(256) void k_fan_fbp_bp(const float *__restrict__ filt, float *__restrict__ img, FanParams P)
{
    const int i = blockIdx.x * 16 + (threadIdx.x & 15), j = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= P.nt || j >= P.ns) return;
    const float x = ((float)i + 0.5f) * P.rt - P.ct, y = ((float)j + 0.5f) * P.rs - P.cs;
    float acc = 0.0f;
    for (int b = 0; b < P.nb; ++b) {
        const float sb = (float)P.view[4 * b + 0], cb = (float)P.view[4 * b + 1];
        const float u = x * cb + y * sb, s = -x * sb + y * cb;
        const float *r = filt + (size_t)b * P.np;
        if (!P.equiangular) {
            const float U = (P.D - s) / P.D;
            const float uu = fmaf(u / U, P.u_scale, P.u_off);
            if (uu >= 0.0f && uu < P.umax) {
                const int iu = min((int)floorf(uu), (int)P.umax - 1);
                const float fu = uu - (float)iu;
                acc += fmaf(fu, r[iu + 1] - r[iu], r[iu]) / (U * U);
            }
        } else {
            const float w = P.D - s;
            const float g = atan2f(u, fabsf(w));
            const float gg = (g - P.g0) * P.g_scale;
            if (gg >= 0.0f && gg < P.gmax) {
                // linear interpolation in gamma (FanProjection.m:93-98) on the uniform angle grid
                const int ig = min((int)floorf(gg), (int)P.gmax - 1);
                const float fg = gg - (float)ig;
                const float val = fmaf(fg, r[ig + 1] - r[ig], r[ig]);
                acc += val / fmaf(u, u, w * w);
            }
        }
    }
    acc *= P.dbeta;
    const size_t o = (size_t)j * P.nt + i;
    img[o] = P.accumulate ? img[o] + acc : acc;
}

dim3 pix_grid(const FanParams &P) { return dim3((P.nt + 15) / 16, (P.ns + 15) / 16); }

}  // namespace

hipError_t launch_fan_bp_ref(const FanParams &P, const float *proj, float *img, hipStream_t st)
{
    hipLaunchKernelGGL(k_fan_bp_ref, pix_grid(P), dim3(256), 0, st, proj, img, P);
    return hipGetLastError();
}

hipError_t launch_fan_adjoint(const FanParams &P, const float *proj, float *img, const SirtUpdate *upd, int ones_input,
                              hipStream_t st)
{
    SirtUpdate U = {};
    int sirt = 0;
    if (upd) {
        U = *upd;
        sirt = 1;
    }
    hipLaunchKernelGGL(k_fan_adjoint, pix_grid(P), dim3(256), 0, st, proj, img, P, U, sirt, ones_input);
    return hipGetLastError();
}

// Parallel-beam FBP backprojection (FBPparallel.cu:41-89): one thread per
// pixel, all views accumulated in a register and written once (the
// reference launches one kernel per view and adds into the image).
__global__ __launch_bounds__(256) void k_par_fbp_bp(const float *__restrict__ filt, float *__restrict__ img,
                                                    ParBpParams P)
{
    const int i = blockIdx.x * 16 + (threadIdx.x & 15), j = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= P.nt || j >= P.ns) return;
    const float x = ((float)i + 0.5f) * P.rt - P.ct, y = ((float)j + 0.5f) * P.rs - P.cs;
    float acc = 0.0f;
    for (int b = 0; b < P.nb; ++b) {
        const float sb = P.view[2 * b], cb = P.view[2 * b + 1];
        const float t = fmaf(x, cb, y * sb);
        const float u = fmaf(t, P.u_scale, P.u_off);
        if (u >= 0.0f && u < P.umax) {
            const int iu = min((int)u, P.np - 2);
            const float f = u - (float)iu;
            const float *row = filt + (size_t)b * P.np;
            acc = fmaf(fmaf(f, row[iu + 1] - row[iu], row[iu]), P.dtheta, acc);
        }
    }
    const size_t o = (size_t)j * P.nt + i;
    img[o] = P.accumulate ? img[o] + acc : acc;
}

// Exact adjoint of the parallel Siddon projector.  Per pixel and view the
// candidate rays are the bins inside the pixel's shadow on the detector.  A
// ray at signed distance dl (along n = (cos, sin)) from the pixel centre,
// parameterised by mu along (-sin, cos) from the foot of the centre, has
// offsets (dl cos - mu sin, dl sin + mu cos) from the centre; its length in
// the pixel is the mu interval where both stay within the half pitches.  The
// centre's detector coordinate is formed in FP64 so that dl keeps its
// precision far from the rotation centre.
__global__ __launch_bounds__(256) void k_par_adjoint(const float *__restrict__ proj, float *__restrict__ img,
                                                     ParBpParams P, SirtUpdate U, int sirt, int ones)
{
    const int i = blockIdx.x * 16 + (threadIdx.x & 15), j = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= P.nt || j >= P.ns) return;
    const double px = ((double)i + 0.5) * P.rt_d - P.ct_d, py = ((double)j + 0.5) * P.rs_d - P.cs_d;
    const float hx = 0.5f * P.rt, hy = 0.5f * P.rs;
    float acc = 0.0f;
    for (int b = 0; b < P.nb; ++b) {
        const double sbd = P.view_d[4 * b + 0], cbd = P.view_d[4 * b + 1];
        const float sb = (float)sbd, cb = (float)cbd;
        const double uc = fma(px, cbd, py * sbd);
        const float reach = fmaf(fabsf(cb), hx, fabsf(sb) * hy);
        float u0 = fmaf((float)uc - reach, P.u_scale, P.u_off), u1 = fmaf((float)uc + reach, P.u_scale, P.u_off);
        const float ulo = fminf(u0, u1), uhi = fmaxf(u0, u1);
        const float eu = 1e-3f + 4e-6f * fmaxf(fabsf(ulo), fabsf(uhi));
        const int c0 = max((int)ceilf(fminf(fmaxf(ulo - eu, -1.0f), (float)P.np)), 0);
        const int c1 = min((int)floorf(fminf(uhi + eu, (float)P.np)), P.np - 1);
        const float *r = proj + (size_t)b * P.np;
        for (int c = c0; c <= c1; ++c) {
            const float dl = (float)(P.t_d[c] - uc);
            const float ox = dl * cb, oy = dl * sb;
            float lo = -kInf, hi = kInf;
            if (sb != 0.0f) {
                const float inv = 1.0f / sb;
                const float l1 = (ox - hx) * inv, l2 = (ox + hx) * inv;
                lo = fmaxf(lo, fminf(l1, l2));
                hi = fminf(hi, fmaxf(l1, l2));
            } else if (!(ox >= -hx && ox < hx)) {
                continue;
            }
            if (cb != 0.0f) {
                const float inv = 1.0f / cb;
                const float l1 = (-hy - oy) * inv, l2 = (hy - oy) * inv;
                lo = fmaxf(lo, fminf(l1, l2));
                hi = fminf(hi, fmaxf(l1, l2));
            } else if (!(oy >= -hy && oy < hy)) {
                continue;
            }
            acc = fmaf(fmaxf(hi - lo, 0.0f), ones ? 1.0f : r[c], acc);
        }
    }
    const size_t o = (size_t)j * P.nt + i;
    if (sirt) {
        const float xn = fmaf(U.lambda * U.vinv[o], acc, U.x[o]);
        U.x[o] = fminf(fmaxf(xn, U.lo), U.hi);
    } else {
        img[o] = P.accumulate ? img[o] + acc : acc;
    }
}

hipError_t launch_par_fbp_bp(const ParBpParams &P, const float *filt, float *img, hipStream_t st)
{
    hipLaunchKernelGGL(k_par_fbp_bp, dim3((P.nt + 15) / 16, (P.ns + 15) / 16), dim3(256), 0, st, filt, img, P);
    return hipGetLastError();
}

hipError_t launch_par_adjoint(const ParBpParams &P, const float *proj, float *img, const SirtUpdate *U,
                              int ones_input, hipStream_t st)
{
    SirtUpdate u{};
    if (U) u = *U;
    hipLaunchKernelGGL(k_par_adjoint, dim3((P.nt + 15) / 16, (P.ns + 15) / 16), dim3(256), 0, st, proj, img, P, u,
                       U ? 1 : 0, ones_input);
    return hipGetLastError();
}

hipError_t launch_fan_fbp_bp(const FanParams &P, const float *filt, float *img, hipStream_t st)
{
    hipLaunchKernelGGL(k_fan_fbp_bp, pix_grid(P), dim3(256), 0, st, filt, img, P);
    return hipGetLastError();
}

}  // namespace ctr
