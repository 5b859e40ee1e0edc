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
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "ctr_internal.h"

namespace ctr {
namespace {

constexpr int kTX = 16, kTS = 16;
constexpr int kThreads = kTX * kTS;
static_assert(kTX == 16 && kTS == 16, "the interpolating kernel maps four 8 x 8 waves onto a 16 x 16 tile");

__device__ __forceinline__ int xcd_tile(int bid, int nwg)
{
    return (nwg & 7) == 0 ? (bid & 7) * (nwg >> 3) + (bid >> 3) : bid;
}

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }


// ---------------------------------------------------------------------------
// Per-(tile, view) detector footprints.  The rectangle of detector samples a
// tile's voxel centres interpolate from, with a safety margin, computed once
// per geometry in FP64 and read by the backprojection with one scalar load
// per view.  flags: kFootEmpty (no sample of the tile is on the detector),
// kFootInterior (every sample of the tile is inside the detector, so the
// inner loop needs no masks), kFootFits (the rectangle fits the LDS tile).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bp_footprints(BpParams P, int tz, int wp, int hmax, int faces,
                                                     int4 *__restrict__ out)
{
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    const long ntiles = (long)P.tiles_t * P.tiles_s * P.tiles_z;
    if (idx >= ntiles * P.nb) return;
    const int tile = (int)(idx / P.nb), b = (int)(idx % P.nb);
    const int tt = tile % P.tiles_t, rest = tile / P.tiles_t;
    const int ts = rest % P.tiles_s, tzi = rest / P.tiles_s;
    const int i_lo = tt * kTX, i_hi = min(i_lo + kTX - 1, P.nt - 1);
    const int j_lo = ts * kTS, j_hi = min(j_lo + kTS - 1, P.ns - 1);
    const BpView vw = P.views[b];
    // corners of the voxel centres (interpolating backprojection) or of the
    // voxel boxes (exact adjoint: every ray through a voxel)
    const double h = faces ? 0.5 : 0.0;
    double umin = 1e300, umax = -1e300, mmin = 1e300, mmax = -1e300;
    for (int c = 0; c < 4; ++c) {
        const double fi = (c & 1) ? i_hi + h : i_lo - h, fj = (c & 2) ? j_hi + h : j_lo - h;
        const double dt = (double)vw.at * fi + (double)vw.as * fj + (double)vw.a0;
        const double ds = (double)vw.bt * fi + (double)vw.bs * fj + (double)vw.b0;
        const double m = (double)P.D / ((double)P.D - ds);
        const double u = dt * m * (double)P.u_scale + (double)P.u_off;
        umin = fmin(umin, u);
        umax = fmax(umax, u);
        mmin = fmin(mmin, m);
        mmax = fmax(mmax, m);
    }
    const double zlo = (double)tzi * tz * P.rz + P.zc0 - h * P.rz, zhi = zlo + (double)(tz - 1 + 2 * h) * P.rz;
    double vmin = 1e300, vmax = -1e300;
    for (int c = 0; c < 4; ++c) {
        const double v = ((c & 1) ? zhi : zlo) * ((c & 2) ? mmax : mmin) * (double)P.v_scale + (double)P.v_off;
        vmin = fmin(vmin, v);
        vmax = fmax(vmax, v);
    }
    // float evaluation of the same maps in the backprojection differs by far less than this margin
    const double eu = 2e-3 + 4e-6 * fmax(fabs(umin), fabs(umax));
    const double ev = 2e-3 + 4e-6 * fmax(fabs(vmin), fabs(vmax));
    int flags = 0;
    if (!(umax + eu >= 0.0 && umin - eu < P.umax && vmax + ev >= 0.0 && vmin - ev < P.vmax)) flags |= kFootEmpty;
    if (umin - eu >= 0.0 && umax + eu < P.umax && vmin - ev >= 0.0 && vmax + ev < P.vmax) flags |= kFootInterior;
    if (vmin - ev >= 0.0 && vmax + ev < P.vmax) flags |= kFootRowsIn;   // every sample on the detector rows
    int c0, c1, r0, r1;
    if (faces) {
        // the detector samples (integer positions) inside the projected boxes
        c0 = (int)ceil(fmax(umin - eu, -1.0));
        c1 = (int)floor(fmin(umax + eu, (double)P.np));
        r0 = (int)ceil(fmax(vmin - ev, -1.0));
        r1 = (int)floor(fmin(vmax + ev, (double)P.nx));
        if (c1 < 0 || c0 > P.np - 1 || r1 < 0 || r0 > P.nx - 1 || c1 < c0 || r1 < r0) flags |= kFootEmpty;
    } else {
        // both neighbours of every bilinear sample
        c0 = (int)floor(fmax(umin - eu, -1.0));
        c1 = (int)floor(fmin(umax + eu, (double)P.np)) + 1;
        r0 = (int)floor(fmax(vmin - ev, -1.0));
        r1 = (int)floor(fmin(vmax + ev, (double)P.nx)) + 1;
    }
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
    const int W = c1 - c0 + 1, H = r1 - r0 + 1;
    if (W <= wp && H <= hmax) flags |= kFootFits;
    out[idx] = make_int4(c0, r0, W | (H << 16), flags);
}

// Chord of the source-to-voxel-centre line through the voxel
// (Backprojection3D.cu:243-288): 2 |d| min_a(h_a / |d_a|), with
// mts = min(ht/|dt|, hs/|ds|) per column and (zs^2, hz/|zs|) per slice.
__device__ __forceinline__ float chord_len(float rho2, float mts, float2 zq)
{
    return 2.0f * sqrtf(zq.x + rho2) * fminf(mts, zq.y);
}

// Direct global -> LDS copy of a view's footprint rows: lane l of a wave
// copies column c0 + l (clamped into the row) of row r into buf[r * WP + l]
// through a buffer descriptor of the view, the row selected by the scalar
// offset, so the issue loop costs no vector instructions.  The copies
// complete asynchronously; a vmcnt wait and a barrier publish them.  Offsets
// past the view read as zero (hardware range check).
constexpr int kBufDword3 = 0x00020000;   // raw 32-bit buffer descriptor word 3 on gfx9

template <int WP>
__device__ __forceinline__ void issue_view(const float *__restrict__ proj, const BpParams &P, int b, int4 f, float *buf)
{
    if ((f.w & kFootEmpty) || !(f.w & kFootFits)) return;
    const int H = f.z >> 16;
    const size_t view = (size_t)P.nx * P.np;
    __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void *)(proj + (size_t)b * view), 0,
                                                                  (int)(view * sizeof(float)), kBufDword3);
    const unsigned voff = (unsigned)min(f.x + (int)(threadIdx.x & 63), P.np - 1) * 4u;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int r = wave; r < H; r += kThreads / 64)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, buf + r * WP, 4, voff, (unsigned)((f.y + r) * P.np) * 4u, 0, 0);
}

// Direct global -> LDS copy of n consecutive floats, elements first ..
// first + n - 1 of a buffer (its row selected by the scalar byte offset
// soff), into dst[0 .. n): 64 per instruction, the last chunk under an EXEC
// mask, so nothing past dst[n - 1] is written.  The caller keeps
// first + n - 1 inside the buffer.
__device__ __forceinline__ void dma_span(__amdgpu_buffer_rsrc_t rs, float *dst, unsigned soff, int first, int n,
                                         int lane)
{
    for (int c0 = 0; c0 < n; c0 += 64)
        if (lane < n - c0)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, dst + c0, 4, (unsigned)(first + c0 + lane) * 4u, soff, 0, 0);
}

// Voxel-driven backprojection of all views in one pass.  LDS holds two
// footprint buffers (HMAX rows of pitch WP floats, 64 staged columns per
// row): while the tile interpolates view b from one buffer, the DMA engine
// fills the other with view b+1, so one barrier per view publishes the next
// footprint and retires the previous one.  The view table is read through
// the scalar cache (its own restrict argument), so no vector-memory wait in
// the compute phase drains the prefetch.
template <int KIND, int TZ, int WP, int HMAX>
__global__ __launch_bounds__(kThreads, 5) void k_bp_tiled(const float *__restrict__ proj, float *__restrict__ vol,
                                                       BpParams P, const int4 *__restrict__ foot,
                                                       const BpView *__restrict__ views,
                                                       const float2 *__restrict__ zq)
{
    __shared__ float lds[2][HMAX * WP];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tt = tile % P.tiles_t, rest = tile / P.tiles_t;
    const int ts = rest % P.tiles_s, tz = rest / P.tiles_s;
    // each wave covers 8 x 8 columns of the 16 x 16 tile: its stencils then
    // spread over fewer detector rows per column run than a 16 x 4 strip's,
    // which measured fewer LDS bank conflicts (C2 3.0 -> 2.8 ms)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int tx = (wv & 1) * 8 + (lane & 7), ty = (wv >> 1) * 8 + (lane >> 3);
    const int i = tt * kTX + tx, j = ts * kTS + ty, z0 = tz * TZ;
    const bool col_in = i < P.nt && j < P.ns;
    const int nzt = min(TZ, P.nz - z0);
    const float zA = fmaf((float)z0, P.rz, P.zc0);   // centre of slice z0 relative to the source plane
    const float fi = (float)i, fj = (float)j;
    // largest float below vmax: 0 <= v < vmax  <=>  med3(v, 0, vlast) == v
    const float vlast = __builtin_bit_cast(float, __builtin_bit_cast(int, P.vmax) - 1);
    const int4 *ft = foot + (size_t)tile * P.nb;

    float acc[TZ];
#pragma unroll
    for (int k = 0; k < TZ; ++k) acc[k] = 0.0f;

    int4 fn = ft[0];
    issue_view<WP>(proj, P, 0, fn, lds[0]);
    for (int b = 0; b < P.nb; ++b) {
        const int4 f = fn;
        const float *buf = lds[b & 1];
        // this wave's copies of view b are complete; after the barrier every
        // wave's are, and no wave still reads the other buffer (view b-1)
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (b + 1 < P.nb) {
            fn = ft[b + 1];
            issue_view<WP>(proj, P, b + 1, fn, lds[(b + 1) & 1]);
        }
        if ((f.w & kFootEmpty) || !col_in) continue;
        const bool fits = (f.w & kFootFits) != 0;

        // per-column projection of this view
        const BpView vw = views[b];
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
        const int iu_raw = (int)floorf(u);
        const float fu = u - (float)iu_raw;
        const float wr = w * fu, wl = w - wr;   // FDK weights of the columns iu and iu + 1

        if (fits && (f.w & kFootInterior)) {
            // every sample of the tile is on the detector and inside the staged
            // rectangle: no masks, no clamps; the stencil is two ds_read2_b32
            const int base4 = ((iu_raw - f.x) - f.y * WP) * 4;
            const char *l8 = reinterpret_cast<const char *>(buf);
#pragma unroll
            for (int k = 0; k < TZ; ++k) {
                const float v = fmaf((float)k, vB, vA);
                const int iv = (int)v;
                const float fv = __builtin_amdgcn_fractf(v);
                const float *s = reinterpret_cast<const float *>(l8 + (__mul24(iv, WP * 4) + base4));
                const float a = s[0], bq = s[1], c = s[WP], d = s[WP + 1];
                if (KIND == kBpFdk) {
                    // interpolate the two detector columns in v, weight them by
                    // w(1 - fu) and w fu: 10 VALU per voxel-update
                    acc[k] = fmaf(wl, fmaf(fv, c - a, a), acc[k]);
                    acc[k] = fmaf(wr, fmaf(fv, d - bq, bq), acc[k]);
                } else {
                    const float top = fmaf(fu, bq - a, a), bot = fmaf(fu, d - c, c);
                    acc[k] = fmaf(chord_len(rho2, mts, zq[z0 + k]), fmaf(fv, bot - top, top), acc[k]);
                }
            }
        } else if (KIND == kBpFdk && fits && (f.w & kFootRowsIn) && P.rows_in) {
            // every sample is on the detector rows and only u leaves the
            // detector: the interior loop, with the column clamped into the
            // staged rectangle and the weights of off-detector columns zeroed
            // (bit-identical to the masked path below, 10 VALU instead of 14)
            const int W = f.z & 0xffff;
            const bool uok = u >= 0.0f && u < P.umax;
            const int iu = min(max(iu_raw, f.x), f.x + W - 2);
            const float wlu = uok ? wl : 0.0f, wru = uok ? wr : 0.0f;
            const int base4 = ((iu - f.x) - f.y * WP) * 4;
            const char *l8 = reinterpret_cast<const char *>(buf);
#pragma unroll
            for (int k = 0; k < TZ; ++k) {
                const float v = fmaf((float)k, vB, vA);
                const int iv = (int)v;
                const float fv = __builtin_amdgcn_fractf(v);
                const float *s = reinterpret_cast<const float *>(l8 + (__mul24(iv, WP * 4) + base4));
                const float a = s[0], bq = s[1], c = s[WP], d = s[WP + 1];
                acc[k] = fmaf(wlu, fmaf(fv, c - a, a), acc[k]);
                acc[k] = fmaf(wru, fmaf(fv, d - bq, bq), acc[k]);
            }
        } else if (fits) {
            // part of the tile projects off the detector.  v is clamped into
            // [0, vlast] (the clamped stencil lies in the staged rectangle, whose
            // rows reach the detector edge whenever a sample leaves it) and the
            // weight is zeroed where the clamp moved v or u is off the detector.
            const int W = f.z & 0xffff;
            const bool uok = u >= 0.0f && u < P.umax;
            const int iu = min(max(iu_raw, f.x), f.x + W - 2) - f.x;
            const float *col = buf + iu - f.y * WP;
            const float wlu = uok ? wl : 0.0f, wru = uok ? wr : 0.0f;
#pragma unroll
            for (int k = 0; k < TZ; ++k) {
                const float v = fmaf((float)k, vB, vA);
                const float vc = __builtin_amdgcn_fmed3f(v, 0.0f, vlast);
                const int iv = (int)vc;
                const float fv = __builtin_amdgcn_fractf(vc);
                const float *s = col + __mul24(iv, WP);
                const float a = s[0], bq = s[1], c = s[WP], d = s[WP + 1];
                if (KIND == kBpFdk) {
                    const bool on = v == vc;
                    acc[k] = fmaf(on ? wlu : 0.0f, fmaf(fv, c - a, a), acc[k]);
                    acc[k] = fmaf(on ? wru : 0.0f, fmaf(fv, d - bq, bq), acc[k]);
                } else {
                    const float top = fmaf(fu, bq - a, a), bot = fmaf(fu, d - c, c);
                    const float val = fmaf(fv, bot - top, top);
                    const float wk = chord_len(rho2, mts, zq[z0 + k]);
                    acc[k] = fmaf(uok && v == vc ? wk : 0.0f, val, acc[k]);
                }
            }
        } else {
            // footprint larger than an LDS buffer: gather from global memory
            const float *src = proj + (size_t)b * P.nx * P.np;
            const bool uok = u >= 0.0f && u < P.umax;
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
                if (KIND == kBpFdk) {
                    acc[k] = fmaf(ok ? wl : 0.0f, fmaf(fv, c - a, a), acc[k]);
                    acc[k] = fmaf(ok ? wr : 0.0f, fmaf(fv, d - bq, bq), acc[k]);
                } else {
                    const float top = fmaf(fu, bq - a, a), bot = fmaf(fu, d - c, c);
                    const float val = ok ? fmaf(fv, bot - top, top) : 0.0f;
                    acc[k] = fmaf(chord_len(rho2, mts, zq[z0 + k]), val, acc[k]);
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
constexpr int kCMax = 6;   // LDS slack after each staged row and column table: the widest candidate chunk

// Slab interval of one axis for a ray parameterised S + lam*dir: the voxel
// [lo_rel, lo_rel + pitch] (relative to S) holds lam in [l1, l2], sorted.
// inv = 1/dir from slab_rcp, finite: a ray parallel to the slab (dir 0) has
// inv = 1e30, and its interval, [lo_rel, lo_rel + pitch] * 1e30, spans every
// parameter of the volume (|lam| < 1e4) when S lies inside the slab and
// none of them when it lies outside (an end at exactly 0 when S is on a
// face: a voxel in front of the source has lam > 0 either way).
__device__ __forceinline__ void slab_interval(float lo_rel, float pitch, float inv, float &lo, float &hi)
{
    const float l1 = lo_rel * inv, l2 = fmaf(pitch, inv, l1);
    lo = fminf(l1, l2);
    hi = fmaxf(l1, l2);
}

// 1/x for slab_interval: finite, 1e30 in magnitude at most (x = 0 gives 1e30).
__device__ __forceinline__ float slab_rcp(double x)
{
    const double r = x != 0.0 ? 1.0 / x : 1e30;
    return (float)fmin(fmax(r, -1e30), 1e30);
}

// |[zl, zh] n [lo, hi]| for lo <= hi finite and zl <= zh (either may be
// infinite): the interval [zl, zh] clamped into [lo, hi].  Equal to
// max(min(hi, zh) - max(lo, zl), 0) bit for bit, in three instructions: the
// clamps are v_med3, which the compiler does not guard with NaN quieting of
// operands from other blocks as it does fminf/fmaxf.
__device__ __forceinline__ float overlap(float zl, float zh, float lo, float hi)
{
    return __builtin_amdgcn_fmed3f(zh, lo, hi) - __builtin_amdgcn_fmed3f(zl, lo, hi);
}


// Column footprint of one voxel column (box corners): candidate detector
// columns [cu0, cu1] and the magnification range of the column box.
__device__ __forceinline__ void column_box(const BpParams &P, const BpView &vw, float fi, float fj, int &cu0, int &cu1,
                                           float &mmin, float &mmax)
{
    float umin = kInf, umax = -kInf;
    mmin = kInf;
    mmax = -kInf;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float ci = fi + ((c & 1) ? 0.5f : -0.5f), cj = fj + ((c & 2) ? 0.5f : -0.5f);
        const float dt = fmaf(vw.at, ci, fmaf(vw.as, cj, vw.a0));
        const float ds = fmaf(vw.bt, ci, fmaf(vw.bs, cj, vw.b0));
        const float m = P.D * fast_rcp(P.D - ds);
        const float u = fmaf(dt * m, P.u_scale, P.u_off);
        umin = fminf(umin, u);
        umax = fmaxf(umax, u);
        mmin = fminf(mmin, m);
        mmax = fmaxf(mmax, m);
    }
    const float eu = 1e-3f + 4e-6f * fmaxf(fabsf(umin), fabsf(umax));
    cu0 = (int)ceilf(fmaxf(umin - eu, -1.0f));
    cu1 = (int)floorf(fminf(umax + eu, (float)P.np));
}

// Candidate detector rows of the voxel slice [zb, zt] (relative to the
// source plane) seen through a column with magnification in [mmin, mmax].
__device__ __forceinline__ void slice_rows(const BpParams &P, float zb, float zt, float mmin, float mmax, int &rv0,
                                           int &rv1)
{
    // extreme products of {zb, zt} x {mmin, mmax}.  The magnifications are
    // positive (the geometry check keeps the source orbit outside the
    // volume, so every voxel corner lies in front of the source), zb < zt,
    // and rounding is monotone: the smallest product is zb times mmin or
    // mmax by the sign of zb, the largest zt times mmax or mmin by the sign
    // of zt (uniform over the block), bit for bit the min and max of all four
    const float zlo = zb * (zb >= 0.0f ? mmin : mmax);
    const float zhi = zt * (zt >= 0.0f ? mmax : mmin);
    float v0 = fmaf(zlo, P.v_scale, P.v_off);
    float v1 = fmaf(zhi, P.v_scale, P.v_off);
    if (P.v_scale < 0.0f) {
        const float t = v0;
        v0 = v1;
        v1 = t;
    }
    const float ev = 1e-3f + 4e-6f * fmaxf(fabsf(v0), fabsf(v1));
    rv0 = (int)ceilf(fmaxf(v0 - ev, -1.0f));
    rv1 = (int)floorf(fminf(v1 + ev, (float)P.nx));
}

// Rows of the staged footprint and floats of LDS per adjoint block: the
// footprint rows, the column tables, and the z-slab interval table of
// (row, slice) pairs.
template <int TZ, int WP, int CAPF>
struct AdjLds {
    static constexpr int kHmax = (CAPF - 2 * WP - 2 * kCMax) / (WP + 1);
    static constexpr int kFloats = CAPF + 2 * TZ * kHmax;
};

// Exact adjoint (voxel-driven gather of Siddon segment lengths).  LDS holds,
// for the view's footprint rectangle: y*|d| per sample (|d| = sqrt(p^2 +
// D^2 + xi^2)), 1/a and 1/b per column, and per (row, slice) the z-slab
// parameter interval [zl, zh] of that row's rays (every thread of the block
// covers the same slices), so a candidate (ray, voxel) pair costs 4 VALU and
// one immediate-offset LDS read, and a candidate row one 64-bit read.
template <int TZ, int WP, int CAPF, int CM, bool FITALL>
__device__ __forceinline__ void adjoint_tile(const float *__restrict__ proj, float *__restrict__ vol, const BpParams &P,
                                             const int4 *__restrict__ foot, const BpView *__restrict__ views,
                                             const SirtUpdate &U, int sirt, int ones, int tile, float *lds)
{
    constexpr int HMAX = AdjLds<TZ, WP, CAPF>::kHmax;
    float *yd = lds;                                // [HMAX][WP] (+ kCMax zero slack)
    float *inva = lds + HMAX * WP + kCMax;          // [WP + kCMax]
    float *invb = inva + WP + kCMax;                // [WP]
    float2 *ztab = reinterpret_cast<float2 *>(lds + CAPF);   // [HMAX][TZ]: (zl, zh)
    const int tt = tile % P.tiles_t, rest = tile / P.tiles_t;
    const int ts = rest % P.tiles_s, tz = rest / P.tiles_s;
    const int tx = threadIdx.x & (kTX - 1), ty = threadIdx.x / kTX;
    const int i = tt * kTX + tx, j = ts * kTS + ty, z0 = tz * TZ;
    const bool col_in = i < P.nt && j < P.ns;
    const int nzt = min(TZ, P.nz - z0);
    const float zbase = fmaf((float)z0, P.rz, P.zc0 - 0.5f * P.rz);   // low face of slice z0, relative to the source
    const float fi = (float)i, fj = (float)j;
    const int4 *ft = foot + (size_t)tile * P.nb;
    // zero the whole LDS once: the unrolled candidate reads past a staged row
    // then only ever see finite values (their segment length is 0)
    for (int e = threadIdx.x; e < CAPF; e += kThreads) lds[e] = 0.0f;

    float acc[TZ];
#pragma unroll
    for (int k = 0; k < TZ; ++k) acc[k] = 0.0f;

    for (int b = 0; b < P.nb; ++b) {
        const int4 f = ft[b];
        if (f.w & kFootEmpty) continue;
        // FITALL: the host found every footprint of the plan small enough
        // for LDS and launched the variant without the global-memory path
        if (FITALL && !(f.w & kFootFits)) continue;
        const int W = f.z & 0xffff, H = f.z >> 16;
        const bool fits = FITALL || (f.w & kFootFits) != 0;
        const BpView vw = views[b];
        // the input times the ray norm; A'1 reads the norm table for every view
        const float *src = ones ? P.adj_norm : proj + (size_t)b * P.nx * P.np;
        if (fits) {
            // direct global -> LDS copies of the footprint rows and of the
            // view's geometry tables; every wave's copies are complete after
            // its vmcnt wait, and the barrier publishes them to the block
            __syncthreads();
            const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
            const unsigned vbytes = (unsigned)((size_t)P.nx * P.np * sizeof(float));
            const __amdgpu_buffer_rsrc_t ry = __builtin_amdgcn_make_buffer_rsrc((void *)src, 0, (int)vbytes, kBufDword3);
            for (int r = wave; r < H; r += kThreads / 64)
                dma_span(ry, yd + r * WP, (unsigned)((f.y + r) * P.np) * 4u, f.x, W, lane);
            if (wave == 0) {
                const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
                    (void *)(P.adj_inva + (size_t)b * P.np), 0, P.np * (int)sizeof(float), kBufDword3);
                dma_span(ra, inva, 0u, f.x, W, lane);
            } else if (wave == 1) {
                const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
                    (void *)(P.adj_invb + (size_t)b * P.np), 0, P.np * (int)sizeof(float), kBufDword3);
                dma_span(rb, invb, 0u, f.x, W, lane);
            }
            // z-slab intervals of the footprint rows for the tile's slices,
            // formed as the global path below forms them (bit for bit)
            for (int e = threadIdx.x; e < H * TZ; e += kThreads) {
                const int r = e / TZ, k = e - r * TZ;
                float zl, zh;
                slab_interval(fmaf((float)k, P.rz, zbase), P.rz, P.adj_invxi[f.y + r], zl, zh);
                ztab[e] = make_float2(zl, zh);
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
        if (!col_in) continue;

        int cu0, cu1;
        float mmin, mmax;
        column_box(P, vw, fi, fj, cu0, cu1, mmin, mmax);
        const int clo_lim = fits ? f.x : 0, chi_lim = fits ? f.x + W - 1 : P.np - 1;
        const int rlo_lim = fits ? f.y : 0, rhi_lim = fits ? f.y + H - 1 : P.nx - 1;
        cu0 = max(cu0, clo_lim);
        cu1 = min(cu1, chi_lim);
        // voxel low faces relative to the source, in FP64: for a ray nearly
        // parallel to a slab (small a or b) the slab's parameter interval is
        // this offset divided by that component, so its rounding is
        // magnified; formed in float from a float source position it moved
        // segment ends by up to 3e-3 mm at 256^3 / D = 600 mm
        const float tlo = (float)((double)i * P.rt_d - P.view_d[4 * b + 2]);
        const float slo = (float)((double)j * P.rs_d - P.view_d[4 * b + 3]);
        for (int cb = cu0; cb <= cu1; cb += CM) {
            float clo[CM], chi[CM];
#pragma unroll
            for (int q = 0; q < CM; ++q) {
                float ia, ib;
                if (fits) {
                    ia = inva[cb - f.x + q];
                    ib = invb[cb - f.x + q];
                } else {
                    const size_t c = (size_t)b * P.np + min(cb + q, P.np - 1);
                    ia = P.adj_inva[c];
                    ib = P.adj_invb[c];
                }
                float l0, h0, l1, h1;
                slab_interval(tlo, P.rt, ia, l0, h0);
                slab_interval(slo, P.rs, ib, l1, h1);
                // the ray's interval in the voxel column; invalid or empty
                // candidates become [0, 0], which overlaps nothing
                const float lo_ = fmaxf(l0, l1), hi_ = fminf(h0, h1);
                const bool live = cb + q <= cu1 && lo_ < hi_;
                clo[q] = live ? lo_ : 0.0f;
                chi[q] = live ? hi_ : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < TZ; ++k) {
                const float zb = fmaf((float)k, P.rz, zbase), zt = zb + P.rz;
                int rv0, rv1;
                slice_rows(P, zb, zt, mmin, mmax, rv0, rv1);
                rv0 = max(rv0, rlo_lim);
                rv1 = min(rv1, rhi_lim);
                float a_k = 0.0f;
                for (int r = rv0; r <= rv1; ++r) {
                    float zl, zh;
                    if (fits) {
                        const float2 zz = ztab[(r - f.y) * TZ + k];
                        zl = zz.x;
                        zh = zz.y;
                    } else {
                        slab_interval(zb, P.rz, P.adj_invxi[r], zl, zh);
                    }
                    if (fits) {
                        const float *sy = yd + (r - f.y) * WP + (cb - f.x);
#pragma unroll
                        for (int q = 0; q < CM; ++q) {
                            const float len = overlap(zl, zh, clo[q], chi[q]);
                            a_k = fmaf(len, sy[q], a_k);
                        }
                    } else {
#pragma unroll
                        for (int q = 0; q < CM; ++q) {
                            const float len = overlap(zl, zh, clo[q], chi[q]);
                            if (len > 0.0f) a_k = fmaf(len, src[(size_t)r * P.np + min(cb + q, P.np - 1)], a_k);
                        }
                    }
                }
                acc[k] += a_k;
            }
        }
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

// The per-pair work is short dependent chains (slab clip, LDS read, fma), so
// the adjoint runs latency-bound and occupancy decides its speed.  At 256^3 /
// 360 views / 512^2 with the z-interval table and without the global path
// (79-90 VGPRs): six slices at four waves 21.0 ms, at five 18.3 ms, at six
// (spills) 17.3 ms; eight at four 19.3 ms; five at six 17.3 ms; four at six
// 16.9 ms, at seven (spills) 17.6 ms, at eight 18.9 ms; three at seven or
// eight 19.4 ms; two at eight 22.6 ms.  The variant with the global path
// keeps four waves.
template <int TZ, int WP, int CAPF, int CM, bool FITALL>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(FITALL ? 6 : 4, FITALL ? 6 : 4))) void k_adjoint(const float *__restrict__ proj, float *__restrict__ vol, BpParams P,
                                                      const int4 *__restrict__ foot, const BpView *__restrict__ views,
                                                      SirtUpdate U, int sirt, int ones)
{
    __shared__ __align__(16) float lds[AdjLds<TZ, WP, CAPF>::kFloats];
    adjoint_tile<TZ, WP, CAPF, CM, FITALL>(proj, vol, P, foot, views, U, sirt, ones, xcd_tile(blockIdx.x, gridDim.x),
                                           lds);
}

// Many independent problems in one launch: block b belongs to the job whose
// block range holds it (a scalar scan of the small job table).
template <int TZ, int WP, int CAPF, int CM, bool FITALL>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(FITALL ? 6 : 4, FITALL ? 6 : 4))) void k_adjoint_batched(const AdjointJob *__restrict__ jobs, int njobs, int sirt,
                                                              int ones)
{
    __shared__ __align__(16) float lds[AdjLds<TZ, WP, CAPF>::kFloats];
    // job blockIdx.y, its block blockIdx.x; blocks past the job's tiles leave at once
    const int p = blockIdx.y, local = blockIdx.x;
    if (local >= jobs[p].nblocks) return;
    const BpParams P = jobs[p].P;
    const SirtUpdate U = jobs[p].U;
    adjoint_tile<TZ, WP, CAPF, CM, FITALL>(jobs[p].proj, jobs[p].vol, P, jobs[p].foot, P.views, U, sirt, ones,
                                           xcd_tile(local, jobs[p].nblocks), lds);
}

// Number of (tile, view) footprints of a plan that are not empty and do not
// fit the adjoint's LDS buffers: block g counts its grid-stride share into
// out[g] (the host adds kFootCountBlocks counts).
__global__ __launch_bounds__(256) void k_foot_unfit(const int4 *__restrict__ foot, long n, int *__restrict__ out)
{
    __shared__ int part[4];
    int c = 0;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        const int w = foot[e].w;
        c += !(w & kFootEmpty) && !(w & kFootFits);
    }
    for (int m = 32; m > 0; m >>= 1) c += __shfl_xor(c, m, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Geometry tables of the exact adjoint: per (view, column) the reciprocals
// of the ray direction's t and s components, per row 1/xi, per (row, column)
// the ray norm.  One thread per (view, column) and per (row, column).
__global__ __launch_bounds__(256) void k_adjoint_tables(BpParams P, float *__restrict__ inva,
                                                        float *__restrict__ invb, float *__restrict__ invxi,
                                                        float *__restrict__ norm)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const long nvc = (long)P.nb * P.np, nrc = (long)P.nx * P.np;
    if (e < nvc) {
        // the direction components p cos + D sin and p sin - D cos cancel for
        // rays nearly parallel to a slab: formed in FP64 from the projector's
        // own view table, so that the reciprocal keeps float precision
        const int b = (int)(e / P.np), c = (int)(e % P.np);
        const double sb = P.view_d[4 * b + 0], cb = P.view_d[4 * b + 1], p = P.pdd[c], D = (double)P.D;
        const double a = p * cb + D * sb, bq = p * sb - D * cb;
        inva[e] = slab_rcp(a);
        invb[e] = slab_rcp(bq);
    }
    if (e < nrc) {
        const int r = (int)(e / P.np), c = (int)(e % P.np);
        norm[e] = ray_norm(P.pd[c], P.xd[r], P.D);
        if (c == 0) invxi[r] = slab_rcp((double)P.xd[r]);
    }
}

__global__ __launch_bounds__(256) void k_scale_by_norm(const float *__restrict__ in, const float *__restrict__ norm,
                                                       float *__restrict__ out, long per_view, long n)
{
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256)
        out[e] = in[e] * norm[e % per_view];
}

constexpr int kBpTZ = 16;
// (row pitch, rows) of the two FDK/BP_REF footprint buffers, 2 x 32 KB at most
constexpr int kBpShapes[3][2] = {{64, 64}, {72, 56}, {80, 50}};
// Four z-slices per thread, six waves per SIMD for the all-fit variant (79
// VGPRs): the adjoint is latency-bound, so occupancy pays more than sharing
// the per-(column, view) set-up over more slices.  LDS floats per block by
// footprint pitch: six blocks of 26 KB fit a CU at pitch 136 (42 rows), 16 KB
// at the narrower pitches (53 and 97 rows).
constexpr int kAdjTZ = 4;
constexpr int adj_cap(int wp) { return wp == 136 ? 6144 : 4096; }

}  // namespace

// Detector-column width bound of a tile footprint: the tile's diagonal (plus
// one voxel diagonal for box faces) magnified at the volume edge nearest the
// source, in detector columns, plus the bilinear neighbours.
static int footprint_pitch(const BpParams &P, bool faces)
{
    const double rt = P.rt, rs = P.rs;
    const double diag = std::sqrt((kTX * rt) * (kTX * rt) + (kTS * rs) * (kTS * rs)) +
                        (faces ? std::sqrt(rt * rt + rs * rs) : 0.0);
    const double half = 0.5 * std::sqrt((double)P.nt * rt * P.nt * rt + (double)P.ns * rs * P.ns * rs);
    const double wmax = diag * (P.D / (P.D - half)) * std::fabs((double)P.u_scale) + 4.0;
    return wmax <= 40 ? 40 : (wmax <= 72 ? 72 : 136);
}

BpPlan bp_plan(const BpParams &P, int pitch)
{
    BpPlan L{};
    L.faces = 0;
    L.tz = kBpTZ;
    L.tiles_t = (P.nt + kTX - 1) / kTX;
    L.tiles_s = (P.ns + kTS - 1) / kTS;
    L.tiles_z = (P.nz + L.tz - 1) / L.tz;
    // row pitch of the staged footprint (80 floats puts neighbouring rows 16
    // banks apart); the autotuner's choice, overridden by CTR_BP_PITCH
    int want = pitch;
    if (const char *e = std::getenv("CTR_BP_PITCH")) want = std::atoi(e);
    int shape = 2;
    for (int q = 0; q < 3; ++q)
        if (kBpShapes[q][0] == want) shape = q;
    L.wp = kBpShapes[shape][0];
    L.hmax = kBpShapes[shape][1];
    L.capf = 2 * L.wp * L.hmax;
    L.nfoot = (size_t)L.tiles_t * L.tiles_s * L.tiles_z * P.nb;
    return L;
}

BpPlan adjoint_plan(const BpParams &P, int cm)
{
    BpPlan L{};
    L.faces = 1;
    L.tz = kAdjTZ;
    L.tiles_t = (P.nt + kTX - 1) / kTX;
    L.tiles_s = (P.ns + kTS - 1) / kTS;
    L.tiles_z = (P.nz + L.tz - 1) / L.tz;
    L.wp = footprint_pitch(P, true);
    // candidate detector columns evaluated together per voxel (a voxel's
    // shadow spans 2-4 at unit magnification); the autotuner's choice,
    // overridden by CTR_ADJ_CM (3, 4 or 6)
    int want = cm;
    if (const char *e = std::getenv("CTR_ADJ_CM")) want = std::atoi(e);
    L.cm = (want == 3 || want == 6) ? want : 4;
    L.capf = adj_cap(L.wp);
    L.hmax = (L.capf - 2 * L.wp - 2 * kCMax) / (L.wp + 1);
    L.nfoot = (size_t)L.tiles_t * L.tiles_s * L.tiles_z * P.nb;
    return L;
}

hipError_t launch_bp_footprints(const BpParams &Pin, const BpPlan &L, int4 *foot, hipStream_t st)
{
    BpParams P = Pin;
    P.tiles_t = L.tiles_t;
    P.tiles_s = L.tiles_s;
    P.tiles_z = L.tiles_z;
    const long n = (long)L.nfoot;
    // the interpolating kernels stage 64 columns per row (one DMA lane each)
    // whatever the row pitch; the adjoint stages up to its pitch
    const int wcap = L.faces ? L.wp : 64;
    hipLaunchKernelGGL(k_bp_footprints, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, P, L.tz, wcap, L.hmax,
                       L.faces, foot);
    return hipGetLastError();
}

template <int KIND>
static hipError_t launch_bp_kind(const BpParams &P, const BpPlan &L, const int4 *foot, const float *proj, float *vol,
                                 hipStream_t st)
{
    const int nwg = L.tiles_t * L.tiles_s * L.tiles_z;
    if (L.wp == 64)
        hipLaunchKernelGGL((k_bp_tiled<KIND, kBpTZ, 64, 64>), dim3(nwg), dim3(kThreads), 0, st, proj, vol, P, foot,
                           P.views, P.zq);
    else if (L.wp == 72)
        hipLaunchKernelGGL((k_bp_tiled<KIND, kBpTZ, 72, 56>), dim3(nwg), dim3(kThreads), 0, st, proj, vol, P, foot,
                           P.views, P.zq);
    else
        hipLaunchKernelGGL((k_bp_tiled<KIND, kBpTZ, 80, 50>), dim3(nwg), dim3(kThreads), 0, st, proj, vol, P, foot,
                           P.views, P.zq);
    return hipGetLastError();
}

hipError_t launch_bp_tiled(int kind, const BpParams &Pin, const BpPlan &L, const int4 *foot, const float *proj,
                           float *vol, hipStream_t st)
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
    P.tiles_t = L.tiles_t;
    P.tiles_s = L.tiles_s;
    P.tiles_z = L.tiles_z;
    return kind == kBpFdk ? launch_bp_kind<kBpFdk>(P, L, foot, proj, vol, st)
                          : launch_bp_kind<kBpRef>(P, L, foot, proj, vol, st);
}

template <int CM, bool FITALL>
static hipError_t launch_adjoint_cm(const BpParams &P, const BpPlan &L, const int4 *foot, const float *proj,
                                    float *vol, const SirtUpdate &U, int sirt, int ones_input, hipStream_t st)
{
    const int nwg = L.tiles_t * L.tiles_s * L.tiles_z;
    if (L.wp == 40)
        hipLaunchKernelGGL((k_adjoint<kAdjTZ, 40, adj_cap(40), CM, FITALL>), dim3(nwg), dim3(kThreads), 0, st, proj, vol,
                           P, foot, P.views, U, sirt, ones_input);
    else if (L.wp == 72)
        hipLaunchKernelGGL((k_adjoint<kAdjTZ, 72, adj_cap(72), CM, FITALL>), dim3(nwg), dim3(kThreads), 0, st, proj, vol,
                           P, foot, P.views, U, sirt, ones_input);
    else
        hipLaunchKernelGGL((k_adjoint<kAdjTZ, 136, adj_cap(136), CM, FITALL>), dim3(nwg), dim3(kThreads), 0, st, proj,
                           vol, P, foot, P.views, U, sirt, ones_input);
    return hipGetLastError();
}

template <int CM>
static hipError_t launch_adjoint_fit(const BpParams &P, const BpPlan &L, const int4 *foot, const float *proj,
                                     float *vol, const SirtUpdate &U, int sirt, int ones_input, hipStream_t st)
{
    return L.fit_all ? launch_adjoint_cm<CM, true>(P, L, foot, proj, vol, U, sirt, ones_input, st)
                     : launch_adjoint_cm<CM, false>(P, L, foot, proj, vol, U, sirt, ones_input, st);
}

hipError_t launch_adjoint(const BpParams &Pin, const BpPlan &L, const int4 *foot, const float *proj, float *vol,
                          const SirtUpdate *upd, int ones_input, hipStream_t st)
{
    BpParams P = Pin;
    P.tiles_t = L.tiles_t;
    P.tiles_s = L.tiles_s;
    P.tiles_z = L.tiles_z;
    SirtUpdate U = {};
    int sirt = 0;
    if (upd) {
        U = *upd;
        sirt = 1;
    }
    if (L.cm == 3) return launch_adjoint_fit<3>(P, L, foot, proj, vol, U, sirt, ones_input, st);
    if (L.cm == 4) return launch_adjoint_fit<4>(P, L, foot, proj, vol, U, sirt, ones_input, st);
    return launch_adjoint_fit<6>(P, L, foot, proj, vol, U, sirt, ones_input, st);
}

template <int CM, bool FITALL>
static hipError_t launch_adjoint_batched_cm(int wp, const AdjointJob *jobs, int njobs, int max_blocks, int sirt,
                                            int ones_input, hipStream_t st)
{
    // x: the job's blocks (a multiple of 8, so that blockIdx.x mod 8 stays the XCD), y: the job
    const dim3 grid((unsigned)((max_blocks + 7) / 8 * 8), (unsigned)njobs);
    if (wp == 40)
        hipLaunchKernelGGL((k_adjoint_batched<kAdjTZ, 40, adj_cap(40), CM, FITALL>), grid, dim3(kThreads), 0, st, jobs,
                           njobs, sirt, ones_input);
    else if (wp == 72)
        hipLaunchKernelGGL((k_adjoint_batched<kAdjTZ, 72, adj_cap(72), CM, FITALL>), grid, dim3(kThreads), 0, st, jobs,
                           njobs, sirt, ones_input);
    else if (wp == 136)
        hipLaunchKernelGGL((k_adjoint_batched<kAdjTZ, 136, adj_cap(136), CM, FITALL>), grid, dim3(kThreads), 0, st, jobs,
                           njobs, sirt, ones_input);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <int CM>
static hipError_t launch_adjoint_batched_fit(int wp, bool fit_all, const AdjointJob *jobs, int njobs, int max_blocks,
                                             int sirt, int ones_input, hipStream_t st)
{
    return fit_all ? launch_adjoint_batched_cm<CM, true>(wp, jobs, njobs, max_blocks, sirt, ones_input, st)
                   : launch_adjoint_batched_cm<CM, false>(wp, jobs, njobs, max_blocks, sirt, ones_input, st);
}

hipError_t launch_adjoint_batched(int wp, int cm, bool fit_all, const AdjointJob *jobs, int njobs, int max_blocks,
                                  int sirt, int ones_input, hipStream_t st)
{
    if (cm == 3) return launch_adjoint_batched_fit<3>(wp, fit_all, jobs, njobs, max_blocks, sirt, ones_input, st);
    if (cm == 4) return launch_adjoint_batched_fit<4>(wp, fit_all, jobs, njobs, max_blocks, sirt, ones_input, st);
    return launch_adjoint_batched_fit<6>(wp, fit_all, jobs, njobs, max_blocks, sirt, ones_input, st);
}

hipError_t launch_foot_unfit(const int4 *foot, size_t n, int *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_foot_unfit, dim3(kFootCountBlocks), dim3(256), 0, st, foot, (long)n, out);
    return hipGetLastError();
}

hipError_t launch_adjoint_tables(const BpParams &P, hipStream_t st)
{
    const long n = std::max((long)P.nb * P.np, (long)P.nx * P.np);
    hipLaunchKernelGGL(k_adjoint_tables, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, P,
                       const_cast<float *>(P.adj_inva), const_cast<float *>(P.adj_invb),
                       const_cast<float *>(P.adj_invxi), const_cast<float *>(P.adj_norm));
    return hipGetLastError();
}

hipError_t launch_scale_by_norm(const BpParams &P, const float *in, float *out, hipStream_t st)
{
    const long per_view = (long)P.nx * P.np, n = per_view * P.nb;
    const long blocks = std::min((n + 255) / 256, 65536L);
    hipLaunchKernelGGL(k_scale_by_norm, dim3((unsigned)blocks), dim3(256), 0, st, in, P.adj_norm, out, per_view, n);
    return hipGetLastError();
}

}  // namespace ctr
