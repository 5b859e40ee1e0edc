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

constexpr int kZc = 8;

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
    const size_t view_stride = (size_t)P.nx * ng;
    for (int b = 0; b < P.nb; ++b) {
        const float sb = views[2 * b], cb = views[2 * b + 1];
        const float u = fmaf(x, cb, y * sb), s = fmaf(-x, sb, y * cb);
        // source (-D sin b, D cos b, 0): in-plane distance^2 = u^2 + (D - s)^2
        const float ds = P.D - s;
        const float rho2 = fmaf(u, u, ds * ds);
        const float gi = fmaf(atanf(u / ds), P.g_scale, P.g_off);
        if (!(gi >= 0.0f && gi < P.gmax)) continue;
        const int i0 = min((int)gi, ng - 2);
        const float fg = gi - (float)i0;
        const float inv_rho = rsqrtf(rho2), rho = rho2 * inv_rho;
        const float xs = P.Dsd * inv_rho * P.x_scale;   // row coordinate per unit z
        const float wb = P.dbeta * rho;
        const float *F = filt + (size_t)b * view_stride + i0;
#pragma unroll
        for (int c = 0; c < kZc; ++c) {
            const float xi = fmaf(zc[c], xs, P.x_off);
            if (xi >= 0.0f && xi < P.xmax) {
                const int j0 = min((int)xi, P.nx - 2);
                const float fx = xi - (float)j0;
                const float *r0 = F + (size_t)j0 * ng, *r1 = r0 + ng;
                const float a = fmaf(fg, r0[1] - r0[0], r0[0]);
                const float bq = fmaf(fg, r1[1] - r1[0], r1[0]);
                const float L2 = fmaf(zc[c], zc[c], rho2);
                const float w = wb / (L2 * sqrtf(L2));
                acc[c] = fmaf(fmaf(fx, bq - a, a), w, acc[c]);
            }
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

}  // namespace

hipError_t launch_curved_bp(const CurvedBpParams &P, const float *filt, float *vol, hipStream_t st)
{
    const dim3 grid((P.nt + 15) / 16, (P.ns + 15) / 16, (P.nz + kZc - 1) / kZc);
    hipLaunchKernelGGL(k_curved_bp, grid, dim3(256), 0, st, filt, vol, P.view, P);
    return hipGetLastError();
}

}  // namespace ctr
