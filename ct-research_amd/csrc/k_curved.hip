// k_curved.hip - backprojection of the curved (cylindrical, equiangular)
// cone-beam FDK for gfx950 (My_toolbox/Implement_UnitedImaging_Data/
// ImplementUI.cu:80-153).
//
// The channel angle gamma = atan(u / (D - s)) and the in-plane source
// distance rho depend on (x, y, view) only, and the row coordinate
// xi = Dsd z / rho is linear in z.  So a thread owns one (t, s) column and a
// chunk of kZc voxels along z: per view it forms gamma, rho and the channel
// interpolation once, then each voxel of the chunk costs a row coordinate, a
// bilinear sample and its weight dbeta rho / L2^1.5.  One launch covers all
// views (the reference launches one kernel per view, ImplementUI.cu:333-339).
#include "ctr_internal.h"

namespace ctr {
namespace {

template <int kZc>
__global__ __launch_bounds__(256) void k_curved_bp(const float *__restrict__ filt, float *__restrict__ vol,
                                                   const float *__restrict__ views, CurvedBpParams P)
{
    const int i = blockIdx.x * 16 + (threadIdx.x & 15), j = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int k0 = blockIdx.z * kZc;
    if (i >= P.nt || j >= P.ns) return;
    const float x = ((float)i + P.off) * P.rt - P.ct, y = ((float)j + P.off) * P.rs - P.cs;
    float zc[kZc], acc[kZc];
#pragma unroll
    for (int c = 0; c < kZc; ++c) {
        zc[c] = ((float)(k0 + c) + P.off) * P.rz - P.cz;
        acc[c] = 0.0f;
    }
    const int ng = P.ng;
    const size_t view_elems = (size_t)P.nx * ng;
    const unsigned row_bytes = (unsigned)ng * 4u;
    const float xlast = __builtin_bit_cast(float, __builtin_bit_cast(int, P.xmax) - 1);   // largest float below xmax
    for (int b = 0; b < P.nb; ++b) {
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
        const float *Fv = filt + (size_t)b * view_elems;
        // branch-free over the chunk, so that all its loads are in flight together:
        // rows are clamped into the detector and out-of-range voxels get weight 0
        float v0[kZc], v1[kZc], v2[kZc], v3[kZc], fxs[kZc];
        bool in[kZc];
#pragma unroll
        for (int c = 0; c < kZc; ++c) {
            // 0 <= xi < xmax exactly when clamping into [0, xlast] leaves xi unchanged;
            // the clamped row index is then at most nx - 2
            const float xi = fmaf(zc[c], xs, P.x_off);
            const float xcl = __builtin_amdgcn_fmed3f(xi, 0.0f, xlast);
            in[c] = xi == xcl;
            const int j0 = (int)xcl;
            fxs[c] = __builtin_amdgcn_fractf(xcl);
            // the view's base pointer is uniform; 32-bit byte offsets within the view
            // (scalar base + vector offset addressing)
            const unsigned ob = (unsigned)(j0 * ng + i0) * 4u;
            float2 p0, p1;   // 4-byte aligned pairs
            __builtin_memcpy(&p0, reinterpret_cast<const char *>(Fv) + ob, sizeof p0);
            __builtin_memcpy(&p1, reinterpret_cast<const char *>(Fv) + (ob + row_bytes), sizeof p1);
            v0[c] = p0.x;
            v1[c] = p0.y;
            v2[c] = p1.x;
            v3[c] = p1.y;
        }
#pragma unroll
        for (int c = 0; c < kZc; ++c) {
            const float a = fmaf(fg, v1[c] - v0[c], v0[c]);
            const float bq = fmaf(fg, v3[c] - v2[c], v2[c]);
            // weight dbeta rho / L2^1.5 (ImplementUI.cu:145: / L2 * dbeta / (sqrt(L2) / rho))
            const float L2 = fmaf(zc[c], zc[c], rho2);
            const float r = __builtin_amdgcn_rsqf(L2);
            acc[c] = fmaf(fmaf(fxs[c], bq - a, a), in[c] ? wb * (r * r * r) : 0.0f, acc[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < kZc; ++c) {
        const int k = k0 + c;
        if (k < P.nz) {
            const size_t o = ((size_t)k * P.ns + j) * P.nt + i;
            vol[o] = P.accumulate ? vol[o] + acc[c] : acc[c];
        }
    }
}

template <int ZC>
hipError_t launch_zc(const CurvedBpParams &P, const float *filt, float *vol, hipStream_t st)
{
    const dim3 grid((P.nt + 15) / 16, (P.ns + 15) / 16, (P.nz + ZC - 1) / ZC);
    hipLaunchKernelGGL((k_curved_bp<ZC>), grid, dim3(256), 0, st, filt, vol, P.view, P);
    return hipGetLastError();
}

// z-chunk per thread: the per-view set-up (about 30 instructions) is shared
// by the chunk's voxels (about 20 each); chunks past nz are wasted work
int choose_zc(int nz)
{
    static const int cand[] = {4, 8, 10, 12, 16};
    int best = 8;
    long best_cost = -1;
    for (int zc : cand) {
        const long cost = (long)((nz + zc - 1) / zc) * (30 + 20 * zc);
        if (best_cost < 0 || cost < best_cost) {
            best = zc;
            best_cost = cost;
        }
    }
    return best;
}

}  // namespace

hipError_t launch_curved_bp(const CurvedBpParams &P, const float *filt, float *vol, hipStream_t st)
{
    switch (choose_zc(P.nz)) {
    case 4: return launch_zc<4>(P, filt, vol, st);
    case 10: return launch_zc<10>(P, filt, vol, st);
    case 12: return launch_zc<12>(P, filt, vol, st);
    case 16: return launch_zc<16>(P, filt, vol, st);
    default: return launch_zc<8>(P, filt, vol, st);
    }
}

}  // namespace ctr
