// k_backproject.hip - voxel-driven cone-beam backprojectors for gfx950.
//
// Dataflow (FDK and the reference BP).  A workgroup of 256 threads owns a
// tile of 16 x 16 voxel columns x TZ slices and loops over ALL views in one
// launch; each thread keeps the TZ partial sums of its column in registers,
// so the volume is written once (K = n_beta views fused per pass) instead of
// read-modify-written once per view (FDK.cu:508-582, Backprojection3D.cu:445-520).
// Per view the workgroup projects the tile's corners onto the detector,
// stages that footprint of the projection into LDS with coalesced loads,
// and every voxel interpolates its sample from LDS (hand-written bilinear).
// A tile whose footprint is fully inside the detector takes an unmasked
// inner loop; edge tiles take a masked one; a footprint larger than the LDS
// buffer reads the samples straight from global memory.
//
// Workgroup -> tile mapping is XCD-aware: consecutive workgroups go to
// different XCDs round-robin, so tile = (bid % 8) * (nwg / 8) + bid / 8
// gives each XCD a contiguous z-slab of tiles whose footprints share rows
// of the detector in that XCD's L2.
#include "ctr_internal.h"

namespace ctr {
namespace {

constexpr int kTX = 16, kTS = 16;
constexpr int kThreads = kTX * kTS;

__device__ __forceinline__ int xcd_tile(int bid, int nwg)
{
    return (nwg & 7) == 0 ? (bid & 7) * (nwg >> 3) + (bid >> 3) : bid;
}

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

struct Footprint {
    int c0, c1, r0, r1;     // staged sample rectangle (inclusive)
    bool empty, interior;
    float mmin, mmax;
};

// Detector footprint of the voxel centres of a tile for one view.
// u is a linear-fractional function of (i, j) and v is bilinear in
// (z, magnification), so the extremes sit at the tile's corners.
__device__ __forceinline__ Footprint tile_footprint(const BpParams &P, const BpView &vw, int i_lo, int i_hi, int j_lo,
                                                    int j_hi, float zlo, float zhi)
{
    float umin = kInf, umax = -kInf, mmin = kInf, mmax = -kInf;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float fi = (float)((c & 1) ? i_hi : i_lo), fj = (float)((c & 2) ? j_hi : j_lo);
        const float dt = fmaf(vw.at, fi, fmaf(vw.as, fj, vw.a0));
        const float ds = fmaf(vw.bt, fi, fmaf(vw.bs, fj, vw.b0));
        const float m = P.D / (P.D - ds);
        const float u = fmaf(dt * m, P.u_scale, P.u_off);
        umin = fminf(umin, u);
        umax = fmaxf(umax, u);
        mmin = fminf(mmin, m);
        mmax = fmaxf(mmax, m);
    }
    float vmin = kInf, vmax = -kInf;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float z = (c & 1) ? zhi : zlo, m = (c & 2) ? mmax : mmin;
        const float v = fmaf(z * m, P.v_scale, P.v_off);
        vmin = fminf(vmin, v);
        vmax = fmaxf(vmax, v);
    }
    const float eu = 2e-3f + 4e-6f * fmaxf(fabsf(umin), fabsf(umax));
    const float ev = 2e-3f + 4e-6f * fmaxf(fabsf(vmin), fabsf(vmax));
    Footprint f;
    f.mmin = mmin;
    f.mmax = mmax;
    f.interior = (umin - eu >= 0.0f) && (umax + eu < P.umax) && (vmin - ev >= 0.0f) && (vmax + ev < P.vmax);
    // whole tile off the detector?
    f.empty = !(umax + eu >= 0.0f && umin - eu < P.umax && vmax + ev >= 0.0f && vmin - ev < P.vmax);
    int c0 = (int)floorf(fmaxf(umin - eu, -1.0f)), c1 = (int)floorf(fminf(umax + eu, (float)P.np)) + 1;
    int r0 = (int)floorf(fmaxf(vmin - ev, -1.0f)), r1 = (int)floorf(fminf(vmax + ev, (float)P.nx)) + 1;
    c0 = max(c0, 0);
    c1 = min(c1, P.np - 1);
    r0 = max(r0, 0);
    r1 = min(r1, P.nx - 1);
    // at least 2 x 2 samples so that every clamped bilinear stencil is inside
    if (c1 <= c0) {
        if (c0 > 0) c0 = c1 - 1; else c1 = c0 + 1;
    }
    if (r1 <= r0) {
        if (r0 > 0) r0 = r1 - 1; else r1 = r0 + 1;
    }
    f.c0 = c0;
    f.c1 = c1;
    f.r0 = r0;
    f.r1 = r1;
    return f;
}

// Coalesced copy of rows r0..r1, columns c0..c1 of one view into LDS.
__device__ __forceinline__ void stage(float *lds, const float *__restrict__ src, int np, int r0, int H, int c0, int W)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int r = wave; r < H; r += kThreads / 64) {
        const float *row = src + (size_t)(r0 + r) * np + c0;
        float *dst = lds + r * W;
        for (int c = lane; c < W; c += 64) dst[c] = row[c];
    }
}

template <int KIND, int TZ, int CAP>
__global__ __launch_bounds__(kThreads) void k_bp_tiled(const float *__restrict__ proj, float *__restrict__ vol,
                                                       BpParams P)
{
    __shared__ float lds[CAP];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tt = tile % P.tiles_t, rest = tile / P.tiles_t;
    const int ts = rest % P.tiles_s, tz = rest / P.tiles_s;
    const int tx = threadIdx.x & (kTX - 1), ty = threadIdx.x / kTX;
    const int i = tt * kTX + tx, j = ts * kTS + ty, z0 = tz * TZ;
    const bool col_in = i < P.nt && j < P.ns;
    const int i_lo = tt * kTX, i_hi = min(i_lo + kTX - 1, P.nt - 1);
    const int j_lo = ts * kTS, j_hi = min(j_lo + kTS - 1, P.ns - 1);
    const int nzt = min(TZ, P.nz - z0);
    const float zA = fmaf((float)z0, P.rz, P.zc0);              // centre of slice z0 relative to the source plane
    const float zlo = zA, zhi = fmaf((float)(TZ - 1), P.rz, zA);
    const float fi = (float)i, fj = (float)j;

    float acc[TZ];
#pragma unroll
    for (int k = 0; k < TZ; ++k) acc[k] = 0.0f;

    for (int b = 0; b < P.nb; ++b) {
        const BpView vw = P.views[b];
        const Footprint f = tile_footprint(P, vw, i_lo, i_hi, j_lo, j_hi, zlo, zhi);
        if (f.empty) continue;
        const int W = f.c1 - f.c0 + 1, H = f.r1 - f.r0 + 1;
        const bool fits = W * H <= CAP;
        const float *src = proj + (size_t)b * P.nx * P.np;
        if (fits) {
            __syncthreads();
            stage(lds, src, P.np, f.r0, H, f.c0, W);
            __syncthreads();
        }
        if (!col_in) continue;

        // per-column projection of this view
        const float dt = fmaf(vw.at, fi, fmaf(vw.as, fj, vw.a0));
        const float ds = fmaf(vw.bt, fi, fmaf(vw.bs, fj, vw.b0));
        const float m = P.D * fast_rcp(P.D - ds);
        const float u = fmaf(dt * m, P.u_scale, P.u_off);
        const float vA = fmaf(zA * m, P.v_scale, P.v_off), vB = P.rz * m * P.v_scale;
        float w = 1.0f, rho2 = 0.0f, mts = 0.0f;
        if (KIND == kBpFdk) {
            w = m * m * vw.wv;
        } else {
            const float xs = fmaf(fi + 0.5f, P.rt, -vw.st), ys = fmaf(fj + 0.5f, P.rs, -vw.ss);
            rho2 = xs * xs + ys * ys;
            mts = fminf(P.ht / fabsf(xs), P.hs / fabsf(ys));
        }
        const bool uok = u >= 0.0f && u < P.umax;
        const int iu_raw = (int)floorf(u);
        const float fu = u - (float)iu_raw;

        if (fits && f.interior && nzt == TZ) {
            // every sample of the tile lies inside the detector: unmasked loop
            const int iu = min(max(iu_raw, f.c0), f.c1 - 1) - f.c0;
#pragma unroll
            for (int k = 0; k < TZ; ++k) {
                const float v = fmaf((float)k, vB, vA);
                const int ivr = (int)v;
                const float fv = v - (float)ivr;
                const int iv = min(max(ivr, f.r0), f.r1 - 1) - f.r0;
                const float *s = lds + iv * W + iu;
                const float a = s[0], bq = s[1], c = s[W], d = s[W + 1];
                const float top = fmaf(fu, bq - a, a), bot = fmaf(fu, d - c, c);
                const float val = fmaf(fv, bot - top, top);
                if (KIND == kBpFdk) {
                    acc[k] = fmaf(w, val, acc[k]);
                } else {
                    const float zs = fmaf((float)k, P.rz, zA);
                    const float len = 2.0f * sqrtf(fmaf(zs, zs, rho2)) * fminf(mts, P.hz / fabsf(zs));
                    acc[k] = fmaf(len, val, acc[k]);
                }
            }
        } else if (fits) {
            const int iu = min(max(iu_raw, f.c0), f.c1 - 1) - f.c0;
#pragma unroll
            for (int k = 0; k < TZ; ++k) {
                const float v = fmaf((float)k, vB, vA);
                const bool ok = uok && v >= 0.0f && v < P.vmax;
                const int ivr = (int)floorf(v);
                const float fv = v - (float)ivr;
                const int iv = min(max(ivr, f.r0), f.r1 - 1) - f.r0;
                const float *s = lds + iv * W + iu;
                const float a = s[0], bq = s[1], c = s[W], d = s[W + 1];
                const float top = fmaf(fu, bq - a, a), bot = fmaf(fu, d - c, c);
                const float val = ok ? fmaf(fv, bot - top, top) : 0.0f;
                if (KIND == kBpFdk) {
                    acc[k] = fmaf(w, val, acc[k]);
                } else {
                    const float zs = fmaf((float)k, P.rz, zA);
                    const float len = 2.0f * sqrtf(fmaf(zs, zs, rho2)) * fminf(mts, P.hz / fabsf(zs));
                    acc[k] = fmaf(len, val, acc[k]);
                }
            }
        } else {
            // footprint larger than the LDS buffer: gather from global memory
            const int iu = min(max(iu_raw, 0), P.np - 2);
#pragma unroll 4
            for (int k = 0; k < TZ; ++k) {
                const float v = fmaf((float)k, vB, vA);
                const bool ok = uok && v >= 0.0f && v < P.vmax;
                const int ivr = (int)floorf(v);
                const float fv = v - (float)ivr;
                const int iv = min(max(ivr, 0), P.nx - 2);
                const float *s = src + (size_t)iv * P.np + iu;
                const float a = s[0], bq = s[1], c = s[P.np], d = s[P.np + 1];
                const float top = fmaf(fu, bq - a, a), bot = fmaf(fu, d - c, c);
                const float val = ok ? fmaf(fv, bot - top, top) : 0.0f;
                if (KIND == kBpFdk) {
                    acc[k] = fmaf(w, val, acc[k]);
                } else {
                    const float zs = fmaf((float)k, P.rz, zA);
                    const float len = 2.0f * sqrtf(fmaf(zs, zs, rho2)) * fminf(mts, P.hz / fabsf(zs));
                    acc[k] = fmaf(len, val, acc[k]);
                }
            }
        }
    }

    if (!col_in) return;
#pragma unroll
    for (int k = 0; k < TZ; ++k) {
        if (k < nzt) {
            const size_t o = ((size_t)(z0 + k) * P.ns + j) * P.nt + i;
            vol[o] = P.accumulate ? vol[o] + acc[k] : acc[k];
        }
    }
}

// FDK backprojection with the reference's stale-sample behaviour
// (FDK_GPU/FDK.cu:100-123: Display_pBeta keeps the last in-detector value for
// every z above the detector).  One thread per column, z ascending, as the
// reference kernel; used only when CTR_REF_QUIRKS is requested.
__global__ __launch_bounds__(256) void k_fdk_quirk(const float *__restrict__ proj, float *__restrict__ vol, BpParams P)
{
    const int i = blockIdx.x * 16 + (threadIdx.x & 15), j = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= P.nt || j >= P.ns) return;
    const float fi = (float)i, fj = (float)j;
    for (int b = 0; b < P.nb; ++b) {
        const BpView vw = P.views[b];
        const float dt = fmaf(vw.at, fi, fmaf(vw.as, fj, vw.a0));
        const float ds = fmaf(vw.bt, fi, fmaf(vw.bs, fj, vw.b0));
        const float m = P.D / (P.D - ds);
        const float u = fmaf(dt * m, P.u_scale, P.u_off);
        const float w = m * m * vw.wv;
        const bool uok = u >= 0.0f && u < P.umax;
        const int iu = min(max((int)floorf(u), 0), P.np - 2);
        const float fu = u - floorf(u);
        const float *src = proj + (size_t)b * P.nx * P.np;
        float carried = 0.0f;
        for (int k = 0; k < P.nz; ++k) {
            const float zs = fmaf((float)k, P.rz, P.zc0);
            const float v = fmaf(zs * m, P.v_scale, P.v_off);
            if (uok && v >= 0.0f && v < P.vmax) {
                const int iv = min(max((int)floorf(v), 0), P.nx - 2);
                const float fv = v - floorf(v);
                const float *s = src + (size_t)iv * P.np + iu;
                const float top = fmaf(fu, s[1] - s[0], s[0]), bot = fmaf(fu, s[P.np + 1] - s[P.np], s[P.np]);
                carried = w * fmaf(fv, bot - top, top);
            }
            const size_t o = ((size_t)k * P.ns + j) * P.nt + i;
            vol[o] += carried;
        }
    }
}

// ---------------------------------------------------------------------------
// Exact adjoint of the Siddon projector, voxel-driven.  For a voxel and a
// view the candidate rays are the detector samples inside the projection of
// the voxel's box.  Along the ray S + lam*(a, b, xi), a = p cos + D sin,
// b = p sin - D cos, the t and s slabs give a lam-interval that depends on
// the column p only and the z slab one that depends on the row xi only, so a
// column's intervals are computed once per view and reused for every voxel
// of the column.  Segment length = (lam_hi - lam_lo) * |(a, b, xi)|; the
// norm sqrt(p^2 + D^2 + xi^2) is folded into the staged samples.
// ---------------------------------------------------------------------------
constexpr int kCMax = 8;

struct AdjCol {
    float lo[kCMax], hi[kCMax];
    int cfirst, ccount;
};

__device__ __forceinline__ void slab(float pos_lo, float pitch, float src, float dirc, float &lo, float &hi)
{
    if (dirc != 0.0f) {
        const float inv = 1.0f / dirc;
        const float l1 = (pos_lo - src) * inv, l2 = (pos_lo + pitch - src) * inv;
        lo = fmaxf(lo, fminf(l1, l2));
        hi = fminf(hi, fmaxf(l1, l2));
    } else if (!(src >= pos_lo && src < pos_lo + pitch)) {
        lo = kInf;
        hi = -kInf;
    }
}

template <bool FROM_LDS, int TZ>
__device__ __forceinline__ void adjoint_view(const BpParams &P, const BpView &vw, const float *lds, const float *src,
                                             const Footprint &f, int W, int i, int j, int z0, float *acc, int ones)
{
    const float fi = (float)i, fj = (float)j;
    // the column's own footprint (voxel box corners, indices +-0.5)
    float umin = kInf, umax = -kInf, mmin = kInf, mmax = -kInf;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float ci = fi + ((c & 1) ? 0.5f : -0.5f), cj = fj + ((c & 2) ? 0.5f : -0.5f);
        const float dt = fmaf(vw.at, ci, fmaf(vw.as, cj, vw.a0));
        const float ds = fmaf(vw.bt, ci, fmaf(vw.bs, cj, vw.b0));
        const float m = P.D / (P.D - ds);
        const float u = fmaf(dt * m, P.u_scale, P.u_off);
        umin = fminf(umin, u);
        umax = fmaxf(umax, u);
        mmin = fminf(mmin, m);
        mmax = fmaxf(mmax, m);
    }
    const float eu = 1e-3f + 4e-6f * fmaxf(fabsf(umin), fabsf(umax));
    int cu0 = (int)ceilf(fmaxf(umin - eu, -1.0f)), cu1 = (int)floorf(fminf(umax + eu, (float)P.np));
    cu0 = max(cu0, FROM_LDS ? f.c0 : 0);
    cu1 = min(cu1, FROM_LDS ? f.c1 : P.np - 1);
    const float tlo = fi * P.rt, slo = fj * P.rs;
    for (int cb = cu0; cb <= cu1; cb += kCMax) {
        float clo[kCMax], chi[kCMax];
#pragma unroll
        for (int q = 0; q < kCMax; ++q) {
            const int c = min(cb + q, P.np - 1);
            const float p = P.pd[c];
            const float a = fmaf(p, vw.cb, P.D * vw.sb), bq = fmaf(p, vw.sb, -P.D * vw.cb);
            float lo = -kInf, hi = kInf;
            slab(tlo, P.rt, vw.st, a, lo, hi);
            slab(slo, P.rs, vw.ss, bq, lo, hi);
            if (cb + q > cu1) {
                lo = kInf;
                hi = -kInf;
            }
            clo[q] = lo;
            chi[q] = hi;
        }
#pragma unroll
        for (int k = 0; k < TZ; ++k) {
            const float zb = fmaf((float)(z0 + k), P.rz, P.zc0 - 0.5f * P.rz);  // low face of the voxel, relative to Cz
            const float zt = zb + P.rz;
            float vmin = fminf(fminf(zb * mmin, zb * mmax), fminf(zt * mmin, zt * mmax));
            float vmax = fmaxf(fmaxf(zb * mmin, zb * mmax), fmaxf(zt * mmin, zt * mmax));
            float v0 = fmaf(vmin, P.v_scale, P.v_off), v1 = fmaf(vmax, P.v_scale, P.v_off);
            if (v0 > v1) {
                const float tmp = v0;
                v0 = v1;
                v1 = tmp;
            }
            const float ev = 1e-3f + 4e-6f * fmaxf(fabsf(v0), fabsf(v1));
            int rv0 = (int)ceilf(fmaxf(v0 - ev, -1.0f)), rv1 = (int)floorf(fminf(v1 + ev, (float)P.nx));
            rv0 = max(rv0, FROM_LDS ? f.r0 : 0);
            rv1 = min(rv1, FROM_LDS ? f.r1 : P.nx - 1);
            float a_k = 0.0f;
            for (int r = rv0; r <= rv1; ++r) {
                const float xi = P.xd[r];
                float zl = -kInf, zh = kInf;
                if (xi != 0.0f) {
                    const float inv = 1.0f / xi;
                    const float l1 = zb * inv, l2 = zt * inv;
                    zl = fminf(l1, l2);
                    zh = fmaxf(l1, l2);
                } else if (!(zb <= 0.0f && 0.0f < zt)) {
                    zl = kInf;
                    zh = -kInf;
                }
                float rowacc = 0.0f;
                if (FROM_LDS) {
                    const float *s = lds + (r - f.r0) * W + (cb - f.c0);
#pragma unroll
                    for (int q = 0; q < kCMax; ++q) {
                        const float len = fmaxf(fminf(chi[q], zh) - fmaxf(clo[q], zl), 0.0f);
                        if (len > 0.0f) rowacc = fmaf(len, s[q], rowacc);
                    }
                } else {
                    const float xi2 = xi * xi + P.D * P.D;
#pragma unroll
                    for (int q = 0; q < kCMax; ++q) {
                        const float len = fmaxf(fminf(chi[q], zh) - fmaxf(clo[q], zl), 0.0f);
                        if (len > 0.0f) {
                            const int c = min(cb + q, P.np - 1);
                            const float p = P.pd[c];
                            const float nrm = sqrtf(fmaf(p, p, xi2));
                            const float y = ones ? 1.0f : src[(size_t)r * P.np + c];
                            rowacc = fmaf(len, y * nrm, rowacc);
                        }
                    }
                }
                a_k += rowacc;
            }
            acc[k] += a_k;
        }
    }
}

template <int TZ, int CAP>
__global__ __launch_bounds__(kThreads) void k_adjoint(const float *__restrict__ proj, float *__restrict__ vol, BpParams P,
                                                      SirtUpdate U, int sirt, int ones)
{
    // kCMax floats of slack: the unrolled candidate loop may address up to
    // kCMax - 1 entries past the staged rectangle (their segment length is 0)
    __shared__ float lds[CAP + kCMax];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tt = tile % P.tiles_t, rest = tile / P.tiles_t;
    const int ts = rest % P.tiles_s, tz = rest / P.tiles_s;
    const int tx = threadIdx.x & (kTX - 1), ty = threadIdx.x / kTX;
    const int i = tt * kTX + tx, j = ts * kTS + ty, z0 = tz * TZ;
    const bool col_in = i < P.nt && j < P.ns;
    const int i_lo = tt * kTX, i_hi = min(i_lo + kTX - 1, P.nt - 1);
    const int j_lo = ts * kTS, j_hi = min(j_lo + kTS - 1, P.ns - 1);
    const int nzt = min(TZ, P.nz - z0);
    // box of the tile (faces, not centres), z relative to the source plane
    const float zlo = fmaf((float)z0, P.rz, P.zc0 - 0.5f * P.rz);
    const float zhi = fmaf((float)(z0 + TZ), P.rz, P.zc0 - 0.5f * P.rz);

    float acc[TZ];
#pragma unroll
    for (int k = 0; k < TZ; ++k) acc[k] = 0.0f;

    for (int b = 0; b < P.nb; ++b) {
        const BpView vw = P.views[b];
        // footprint of the tile box: centres at indices +-0.5 are the faces
        BpView fb = vw;
        fb.a0 = vw.a0 - 0.5f * (vw.at + vw.as);
        fb.b0 = vw.b0 - 0.5f * (vw.bt + vw.bs);
        Footprint f = tile_footprint(P, fb, i_lo, i_hi + 1, j_lo, j_hi + 1, zlo, zhi);
        // rays sit on integer sample positions: the staged rectangle must hold
        // every sample inside [umin, umax]; tile_footprint already pads by one.
        if (f.empty) continue;
        const int W = f.c1 - f.c0 + 1, H = f.r1 - f.r0 + 1;
        const bool fits = W * H <= CAP;
        const float *src = proj + (size_t)b * P.nx * P.np;
        if (fits) {
            __syncthreads();
            const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
            for (int r = wave; r < H; r += kThreads / 64) {
                const float xi = P.xd[f.r0 + r];
                const float xi2 = fmaf(xi, xi, P.D * P.D);
                const float *row = src + (size_t)(f.r0 + r) * P.np + f.c0;
                for (int c = lane; c < W; c += 64) {
                    const float p = P.pd[f.c0 + c];
                    const float nrm = sqrtf(fmaf(p, p, xi2));
                    lds[r * W + c] = ones ? nrm : row[c] * nrm;
                }
            }
            __syncthreads();
        }
        if (!col_in) continue;
        if (fits) adjoint_view<true, TZ>(P, vw, lds, src, f, W, i, j, z0, acc, ones);
        else adjoint_view<false, TZ>(P, vw, lds, src, f, W, i, j, z0, acc, ones);
    }

    if (!col_in) return;
#pragma unroll
    for (int k = 0; k < TZ; ++k) {
        if (k < nzt) {
            const size_t o = ((size_t)(z0 + k) * P.ns + j) * P.nt + i;
            if (sirt) {
                const float xn = fmaf(U.lambda * U.vinv[o], acc[k], U.x[o]);
                U.x[o] = fminf(fmaxf(xn, U.lo), U.hi);
            } else {
                vol[o] = P.accumulate ? vol[o] + acc[k] : acc[k];
            }
        }
    }
}

constexpr int kBpTZ = 16, kBpCap = 6144;
constexpr int kAdjTZ = 8, kAdjCap = 6144;

}  // namespace

hipError_t launch_bp_tiled(int kind, const BpParams &Pin, const float *proj, float *vol, hipStream_t st)
{
    BpParams P = Pin;
    if (kind == kBpFdk && P.quirks) {
        if (!P.accumulate) {
            hipError_t e = hipMemsetAsync(vol, 0, sizeof(float) * (size_t)P.nt * P.ns * P.nz, st);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k_fdk_quirk, dim3((P.nt + 15) / 16, (P.ns + 15) / 16), dim3(256), 0, st, proj, vol, P);
        return hipGetLastError();
    }
    P.tiles_t = (P.nt + kTX - 1) / kTX;
    P.tiles_s = (P.ns + kTS - 1) / kTS;
    P.tiles_z = (P.nz + kBpTZ - 1) / kBpTZ;
    const int nwg = P.tiles_t * P.tiles_s * P.tiles_z;
    if (kind == kBpFdk)
        hipLaunchKernelGGL((k_bp_tiled<kBpFdk, kBpTZ, kBpCap>), dim3(nwg), dim3(kThreads), 0, st, proj, vol, P);
    else
        hipLaunchKernelGGL((k_bp_tiled<kBpRef, kBpTZ, kBpCap>), dim3(nwg), dim3(kThreads), 0, st, proj, vol, P);
    return hipGetLastError();
}

hipError_t launch_adjoint(const BpParams &Pin, const float *proj, float *vol, const SirtUpdate *upd, int ones_input,
                          hipStream_t st)
{
    BpParams P = Pin;
    P.tiles_t = (P.nt + kTX - 1) / kTX;
    P.tiles_s = (P.ns + kTS - 1) / kTS;
    P.tiles_z = (P.nz + kAdjTZ - 1) / kAdjTZ;
    const int nwg = P.tiles_t * P.tiles_s * P.tiles_z;
    SirtUpdate U = {};
    int sirt = 0;
    if (upd) {
        U = *upd;
        sirt = 1;
    }
    hipLaunchKernelGGL((k_adjoint<kAdjTZ, kAdjCap>), dim3(nwg), dim3(kThreads), 0, st, proj, vol, P, U, sirt,
                       ones_input);
    return hipGetLastError();
}

}  // namespace ctr
