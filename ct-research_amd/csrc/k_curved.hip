// k_curved.hip - backprojection of the curved (cylindrical, equiangular)
// cone-beam FDK for gfx950 (My_toolbox/Implement_UnitedImaging_Data/
// ImplementUI.cu:80-153).
//
// The channel angle gamma = atan(u / (D - s)) and the in-plane source
// distance rho depend on (x, y, view) only, and the row coordinate
// xi = Dsd z / rho is linear in z.  A workgroup owns a tile of 16 x 16
// columns x TZ slices and loops over all views (the reference launches one
// kernel per view, ImplementUI.cu:333-339); a thread keeps its column's TZ
// sums in registers.  Per view the tile's detector footprint (a rectangle of
// channels x rows, computed once per geometry in FP64) is copied into LDS by
// the DMA path (buffer_load ... lds) while the previous view is
// interpolated from the other buffer, one barrier per view.  Per (column,
// view) a thread forms gamma, rho and the channel interpolation once; each
// voxel then costs a row coordinate, two LDS pair reads and its weight
// dbeta rho / L2^1.5.  Footprints too large for LDS read global memory.
#include "ctr_internal.h"

namespace ctr {
namespace {

constexpr int kCvTX = 16, kCvTS = 16, kCvThreads = kCvTX * kCvTS;
constexpr int kCvWP = 128;    // staged channels per row: two 64-channel DMA chunks
constexpr int kCvHMAX = 32;   // staged rows per view (2 x 16 KiB of LDS per workgroup)
enum { kCvEmpty = 1, kCvFits = 4 };
constexpr int kBufDword3 = 0x00020000;   // raw 32-bit buffer descriptor word 3 on gfx9

__device__ __forceinline__ int xcd_tile(int bid, int nwg)
{
    return (nwg & 7) == 0 ? (bid & 7) * (nwg >> 3) + (bid >> 3) : bid;
}

// Footprint of (tile, view): the channels between the extreme channel angles
// of the tile's corner columns (the level sets of gamma are rays from the
// source, so the extremes over the convex tile are at corners), and the rows
// between the extreme row coordinates xi = Dsd z / rho over the tile's slice
// range and its rho range (rho_min from the point of the tile nearest the
// source), both widened by a margin and by the bilinear neighbour.
__global__ __launch_bounds__(256) void k_curved_footprints(CurvedBpParams P, int4 *__restrict__ out)
{
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    const long ntiles = (long)P.tiles_t * P.tiles_s * P.tiles_z;
    if (idx >= ntiles * P.nb) return;
    const int tile = (int)(idx / P.nb), b = (int)(idx % P.nb);
    const int tt = tile % P.tiles_t, rest = tile / P.tiles_t;
    const int ts = rest % P.tiles_s, tzi = rest / P.tiles_s;
    const int i_lo = tt * kCvTX, i_hi = min(i_lo + kCvTX - 1, P.nt - 1);
    const int j_lo = ts * kCvTS, j_hi = min(j_lo + kCvTS - 1, P.ns - 1);
    const int k_lo = tzi * P.tz, k_hi = min(k_lo + P.tz - 1, P.nz - 1);
    const double sb = P.view[2 * b], cb = P.view[2 * b + 1], D = P.D;
    const double x_lo = (i_lo + (double)P.off) * P.rt - P.ct, x_hi = (i_hi + (double)P.off) * P.rt - P.ct;
    const double y_lo = (j_lo + (double)P.off) * P.rs - P.cs, y_hi = (j_hi + (double)P.off) * P.rs - P.cs;
    double gmin = 1e300, gmax = -1e300, rho_max = 0.0;
    for (int c = 0; c < 4; ++c) {
        const double x = (c & 1) ? x_hi : x_lo, y = (c & 2) ? y_hi : y_lo;
        const double u = x * cb + y * sb, ds = D - (-x * sb + y * cb);
        const double gi = atan(u / ds) * P.g_scale + P.g_off;
        gmin = fmin(gmin, gi);
        gmax = fmax(gmax, gi);
        rho_max = fmax(rho_max, sqrt(u * u + ds * ds));
    }
    // the source (-D sin b, D cos b) clamped into the tile's box of column centres
    const double xs = -D * sb, ysrc = D * cb;
    const double dx = xs - fmin(fmax(xs, x_lo), x_hi), dy = ysrc - fmin(fmax(ysrc, y_lo), y_hi);
    const double rho_min = sqrt(dx * dx + dy * dy);
    const double z_lo = (k_lo + (double)P.off) * P.rz - P.cz, z_hi = (k_hi + (double)P.off) * P.rz - P.cz;
    double vmin = 1e300, vmax = -1e300;
    for (int c = 0; c < 4; ++c) {
        const double z = (c & 1) ? z_hi : z_lo, rho = (c & 2) ? rho_max : rho_min;
        const double v = (double)P.Dsd * z / rho * P.x_scale + P.x_off;
        vmin = fmin(vmin, v);
        vmax = fmax(vmax, v);
    }
    const double eu = 2e-3 + 4e-6 * fmax(fabs(gmin), fabs(gmax));
    const double ev = 2e-3 + 4e-6 * fmax(fabs(vmin), fabs(vmax));
    int flags = 0;
    if (!(gmax + eu >= 0.0 && gmin - eu < P.gmax)) flags |= kCvEmpty;   // every column skipped
    int c0 = (int)floor(fmax(gmin - eu, -1.0)), c1 = (int)floor(fmin(gmax + eu, (double)P.ng)) + 1;
    int r0 = (int)floor(fmax(vmin - ev, -1.0)), r1 = (int)floor(fmin(vmax + ev, (double)P.nx)) + 1;
    c0 = max(c0, 0);
    c1 = min(c1, P.ng - 1);
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
    if (W <= kCvWP && H <= kCvHMAX) flags |= kCvFits;
    out[idx] = make_int4(c0, r0, W | (H << 16), flags);
}

// DMA copy of view b's footprint rows into buf: lane l of a wave copies
// channel c0 + l (clamped into the row) of row r into buf[r * kCvWP + l], and
// for footprints wider than 64 channels also c0 + 64 + l into
// buf[r * kCvWP + 64 + l].  Offsets past the view read zero (hardware range
// check).
__device__ __forceinline__ void curved_issue(const float *__restrict__ filt, const CurvedBpParams &P, int b, int4 f,
                                             float *buf)
{
    if ((f.w & kCvEmpty) || !(f.w & kCvFits)) return;
    const int H = f.z >> 16;
    const size_t view = (size_t)P.nx * P.ng;
    __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void *)(filt + (size_t)b * view), 0,
                                                                  (int)(view * sizeof(float)), kBufDword3);
    const int lane = (int)(threadIdx.x & 63);
    const unsigned v0 = (unsigned)min(f.x + lane, P.ng - 1) * 4u, v1 = (unsigned)min(f.x + 64 + lane, P.ng - 1) * 4u;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool wide = (f.z & 0xffff) > 64;   // the second chunk only when the footprint needs it
    for (int r = wave; r < H; r += kCvThreads / 64) {
        const unsigned roff = (unsigned)((f.y + r) * P.ng) * 4u;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, buf + r * kCvWP, 4, v0, roff, 0, 0);
        if (wide) __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, buf + r * kCvWP + 64, 4, v1, roff, 0, 0);
    }
}

template <int TZ>
__global__ __launch_bounds__(kCvThreads, 5) void k_curved_tiled(const float *__restrict__ filt, float *__restrict__ vol,
                                                             const float *__restrict__ views, CurvedBpParams P)
{
    __shared__ float lds[2][kCvHMAX * kCvWP];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tt = tile % P.tiles_t, rest = tile / P.tiles_t;
    const int ts = rest % P.tiles_s, tzi = rest / P.tiles_s;
    const int i = tt * kCvTX + (threadIdx.x & (kCvTX - 1)), j = ts * kCvTS + (threadIdx.x / kCvTX);
    const int k0 = tzi * TZ;
    const bool col_in = i < P.nt && j < P.ns;
    const float x = ((float)i + P.off) * P.rt - P.ct, y = ((float)j + P.off) * P.rs - P.cs;
    float zc[TZ], acc[TZ];
#pragma unroll
    for (int c = 0; c < TZ; ++c) {
        // slices past nz (not written) take the last slice's z, inside the footprint
        zc[c] = ((float)min(k0 + c, P.nz - 1) + P.off) * P.rz - P.cz;
        acc[c] = 0.0f;
    }
    const int ng = P.ng;
    const size_t view_elems = (size_t)P.nx * ng;
    const unsigned row_bytes = (unsigned)ng * 4u;
    const float xlast = __builtin_bit_cast(float, __builtin_bit_cast(int, P.xmax) - 1);   // largest float below xmax
    const int4 *ft = P.foot + (size_t)tile * P.nb;

    int4 fn = ft[0];
    curved_issue(filt, P, 0, fn, lds[0]);
    for (int b = 0; b < P.nb; ++b) {
        const int4 f = fn;
        const float *buf = lds[b & 1];
        // this wave's copies of view b are complete; after the barrier every
        // wave's are, and no wave still reads the other buffer (view b-1)
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (b + 1 < P.nb) {
            fn = ft[b + 1];
            curved_issue(filt, P, b + 1, fn, lds[(b + 1) & 1]);
        }
        if ((f.w & kCvEmpty) || !col_in) continue;
        const float sb = views[2 * b], cb = views[2 * b + 1];
        const float u = fmaf(x, cb, y * sb), s = fmaf(-x, sb, y * cb);
        // source (-D sin b, D cos b, 0): in-plane distance^2 = u^2 + (D - s)^2
        const float ds = P.D - s;
        const float rho2 = fmaf(u, u, ds * ds);
        const float gi = fmaf(atanf(u * __builtin_amdgcn_rcpf(ds)), P.g_scale, P.g_off);
        if (!(gi >= 0.0f && gi < P.gmax)) continue;
        const int i0 = min((int)gi, ng - 2);
        const float fg = gi - (float)i0;
        const float inv_rho = __builtin_amdgcn_rsqf(rho2), rho = rho2 * inv_rho;
        const float xs = P.Dsd * inv_rho * P.x_scale;   // row coordinate per unit z
        const float wb = P.dbeta * rho;
        // rows are clamped into the detector and out-of-range voxels get weight
        // 0; the clamped stencil lies in the footprint, whose rows reach the
        // detector edge whenever a sample leaves it
        if (f.w & kCvFits) {
            const float *col = buf + (i0 - f.x) - f.y * kCvWP;
#pragma unroll
            for (int c = 0; c < TZ; ++c) {
                const float xi = fmaf(zc[c], xs, P.x_off);
                const float xcl = __builtin_amdgcn_fmed3f(xi, 0.0f, xlast);
                const bool in = xi == xcl;
                const int j0 = (int)xcl;
                const float fxs = __builtin_amdgcn_fractf(xcl);
                const float *q = col + __mul24(j0, kCvWP);
                const float v0 = q[0], v1 = q[1], v2 = q[kCvWP], v3 = q[kCvWP + 1];
                const float a = fmaf(fg, v1 - v0, v0);
                const float bq = fmaf(fg, v3 - v2, v2);
                // weight dbeta rho / L2^1.5 (ImplementUI.cu:145: / L2 * dbeta / (sqrt(L2) / rho))
                const float L2 = fmaf(zc[c], zc[c], rho2);
                const float r = __builtin_amdgcn_rsqf(L2);
                acc[c] = fmaf(fmaf(fxs, bq - a, a), in ? wb * (r * r * r) : 0.0f, acc[c]);
            }
        } else {
            // footprint larger than an LDS buffer: pairs from global memory
            // (32-bit byte offsets from the view's uniform base)
            const float *Fv = filt + (size_t)b * view_elems;
#pragma unroll
            for (int c = 0; c < TZ; ++c) {
                const float xi = fmaf(zc[c], xs, P.x_off);
                const float xcl = __builtin_amdgcn_fmed3f(xi, 0.0f, xlast);
                const bool in = xi == xcl;
                const int j0 = (int)xcl;
                const float fxs = __builtin_amdgcn_fractf(xcl);
                const unsigned ob = (unsigned)(j0 * ng + i0) * 4u;
                float2 p0, p1;   // 4-byte aligned pairs
                __builtin_memcpy(&p0, reinterpret_cast<const char *>(Fv) + ob, sizeof p0);
                __builtin_memcpy(&p1, reinterpret_cast<const char *>(Fv) + (ob + row_bytes), sizeof p1);
                const float a = fmaf(fg, p0.y - p0.x, p0.x);
                const float bq = fmaf(fg, p1.y - p1.x, p1.x);
                const float L2 = fmaf(zc[c], zc[c], rho2);
                const float r = __builtin_amdgcn_rsqf(L2);
                acc[c] = fmaf(fmaf(fxs, bq - a, a), in ? wb * (r * r * r) : 0.0f, acc[c]);
            }
        }
    }
    if (!col_in) return;
#pragma unroll
    for (int c = 0; c < TZ; ++c) {
        const int k = k0 + c;
        if (k < P.nz) {
            const size_t o = ((size_t)k * P.ns + j) * P.nt + i;
            vol[o] = P.accumulate ? vol[o] + acc[c] : acc[c];
        }
    }
}

// z depth of a tile: the per-view column set-up (about 30 instructions) is
// shared by the tile's slices (about 20 each); slices past nz are wasted work
int choose_tz(int nz)
{
    static const int cand[] = {8, 10, 12, 16};
    int best = 10;
    long best_cost = -1;
    for (int tz : cand) {
        const long cost = (long)((nz + tz - 1) / tz) * (30 + 20 * tz);
        if (best_cost < 0 || cost < best_cost) {
            best = tz;
            best_cost = cost;
        }
    }
    return best;
}

template <int TZ>
hipError_t launch_tz(const CurvedBpParams &P, const float *filt, float *vol, hipStream_t st)
{
    const int nwg = P.tiles_t * P.tiles_s * P.tiles_z;
    hipLaunchKernelGGL((k_curved_tiled<TZ>), dim3(nwg), dim3(kCvThreads), 0, st, filt, vol, P.view, P);
    return hipGetLastError();
}

}  // namespace

size_t curved_plan(CurvedBpParams &P)
{
    P.tz = choose_tz(P.nz);
    P.tiles_t = (P.nt + kCvTX - 1) / kCvTX;
    P.tiles_s = (P.ns + kCvTS - 1) / kCvTS;
    P.tiles_z = (P.nz + P.tz - 1) / P.tz;
    return (size_t)P.tiles_t * P.tiles_s * P.tiles_z * P.nb;
}

hipError_t launch_curved_footprints(const CurvedBpParams &P, int4 *foot, hipStream_t st)
{
    const long n = (long)P.tiles_t * P.tiles_s * P.tiles_z * P.nb;
    hipLaunchKernelGGL(k_curved_footprints, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, P, foot);
    return hipGetLastError();
}

hipError_t launch_curved_bp(const CurvedBpParams &P, const float *filt, float *vol, hipStream_t st)
{
    if (!P.foot) return hipErrorInvalidValue;
    switch (P.tz) {
    case 8: return launch_tz<8>(P, filt, vol, st);
    case 10: return launch_tz<10>(P, filt, vol, st);
    case 12: return launch_tz<12>(P, filt, vol, st);
    case 16: return launch_tz<16>(P, filt, vol, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ctr
