// k_fdkerr.hip - error slice of the iterative FDK correction for gfx950
// (My_novel_algorithm/IterativeFDK/IterativeFDKerrCompute.cu:12-75).
//
// Output [beta][z][s] for one voxel slice t: for every view, integration row
// U and detector sample P, the difference between the voxel reached at
// (P (D-U)/D, U) in the rotated frame (its z moved along the ray) and the
// voxel of the same (t, s) in slice z, times the ramp tap at the voxel's
// projected offset.  The reference launches one kernel per (view, U) with
// float accumulation; here one launch covers everything and each output
// accumulates in FP64 registers.
//
// Every voxel index is a floor of float/double intermediates, so those keep
// the reference's types and this file is compiled without FMA contraction:
// the indices then depend only on IEEE-rounded operations.  The (view, U, P)
// -> (x, y) voxel map does not depend on the output element and comes from a
// table built by k_fdkerr_xy; the tap indices of each output are formed into
// LDS once per chunk of 32 detector samples.  The volume is read from a copy
// with z fastest: the lanes of a wave share the voxel (x, y) of a sample and
// differ in their moved slice, so their gather falls into a few cache lines
// of one z column instead of one line per slice.
#include "ctr_internal.h"

#pragma clang fp contract(off)

namespace ctr {
namespace {

constexpr int kFeThreads = 64;
// detector samples per LDS tap-index table: 32 samples x 64 threads of
// 16-bit indices is 4 KiB per one-wave block, so the 64 VGPRs, not LDS,
// bound the occupancy (8 waves per SIMD).  A 256-sample table of ints
// (64 KiB) left two waves per CU: 628 ms for the 128^3 / 360-view slice,
// against 149 ms with 64-sample chunks and 112 ms with 32.
constexpr int kFeChunk = 32;
constexpr int kFeUnroll = 4;   // detector samples per step of the inner loop
static_assert(kFeUnroll - 1 <= kFdkErrRowPad, "the unrolled map reads stay inside the scratch buffer");

__global__ __launch_bounds__(256) void k_fdkerr_xy(FdkErrParams P, int *__restrict__ xy)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t n = (size_t)P.nb * P.nu * P.np;
    if (i >= n) return;
    const int ip = (int)(i % P.np), iu = (int)((i / P.np) % P.nu), b = (int)(i / ((size_t)P.np * P.nu));
    const float cb = P.cosb[b], sb = P.sinb[b];
    const float U = P.u[iu], T = P.tab_t[(size_t)iu * P.np + ip];
    const float tcb = T * cb, usb = U * sb, tsb = T * sb, ucb = U * cb;
    const float X = tcb - usb, Y = tsb + ucb;
    const double fx = floor(((double)X + P.ct) / P.r0), fy = floor(((double)Y + P.cs) / P.r1);
    xy[i] = (fx >= 0.0 && fx < (double)P.nt && fy >= 0.0 && fy < (double)P.ns) ? (int)fy * P.nt + (int)fx : -1;
}

// Per (view, U) the samples P whose voxel lies in the image form one
// interval (the points (P (D-U)/D, U) of a line in the rotated frame); its
// first and last index, or (1, 0) when empty.
__global__ __launch_bounds__(256) void k_fdkerr_range(FdkErrParams P, const int *__restrict__ xy, int2 *__restrict__ rng)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= P.nb * P.nu) return;
    const int *row = xy + (size_t)i * P.np;
    int lo = P.np, hi = -1;
    for (int ip = 0; ip < P.np; ++ip)
        if (row[ip] >= 0) {
            lo = min(lo, ip);
            hi = ip;
        }
    rng[i] = lo <= hi ? make_int2(lo, hi) : make_int2(1, 0);
}

// pz[(y nt + x) nz + z] = pic[(z ns + y) nt + x]
__global__ __launch_bounds__(256) void k_fdkerr_zcopy(FdkErrParams P, const float *__restrict__ pic,
                                                      float *__restrict__ pz)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t plane = (size_t)P.nt * P.ns;
    if (i >= plane * P.nz) return;
    const size_t z = i % P.nz, o = i / P.nz;
    pz[i] = pic[z * plane + o];
}

__global__ __launch_bounds__(kFeThreads) void k_fdkerr(FdkErrParams P, const float *__restrict__ pz,
                                                       const int *__restrict__ xy, const int2 *__restrict__ rng,
                                                       float *__restrict__ out)
{
    __shared__ short fil[kFeChunk * kFeThreads];   // tap index of (P, thread) for a chunk of P, -1 outside the taps
    const int s = blockIdx.x * kFeThreads + threadIdx.x;
    const int z = blockIdx.y, b = blockIdx.z;
    const float cb = P.cosb[b], sb = P.sinb[b];
    const float image_t = (float)((P.t + 0.5) * P.r0 - P.ct);
    const float image_s = (float)((s + 0.5) * P.r1 - P.cs);
    const float image_z = (float)((z + 0.5) * P.r2 - P.cz);
    const float ts = image_t * sb, tc = image_t * cb, ss = image_s * sb, sc = image_s * cb;
    const float dect_t = tc + ss, dect_s = sc - ts, dect_z = image_z;
    const double D = P.D;
    const float LR = (float)(D / (D - (double)dect_s));
    const float dtl = dect_t * LR;
    const float *slice = pz + z;   // voxel o of slice z at slice[o * nz]
    const double half_lr2 = 0.5 * ((double)LR * LR);
    double acc = 0.0;
    // each thread reads only its own column of the table: no barrier needed
    for (int p0 = 0; p0 < P.np; p0 += kFeChunk) {
        const int pn = min(kFeChunk, P.np - p0);
        for (int k = 0; k < pn; ++k) {
            const double ff = floor(((double)dtl - P.p[p0 + k]) / P.pint) + P.np;
            fil[k * kFeThreads + threadIdx.x] = (short)((ff >= 0.0 && ff < (double)P.ng) ? (int)ff : -1);
        }
        if (s >= P.ns) continue;
        for (int iu = 0; iu < P.nu; ++iu) {
            // samples of this chunk whose voxel lies in the image
            const int2 r = rng[(size_t)b * P.nu + iu];
            const int k0 = max(r.x - p0, 0), k1 = min(r.y - p0 + 1, pn);
            if (k0 >= k1) continue;
            const float U = P.u[iu];
            const float num = (dect_s - U) * dect_z;
            const float Zd = (float)((double)dect_z + (double)num / (D - (double)dect_s));
            const double fz = floor(((double)Zd + P.cz) / P.r2);
            if (!(fz >= 0.0 && fz < (double)P.nz)) continue;
            const float *moved = pz + (size_t)fz;
            const int *row = xy + ((size_t)b * P.nu + iu) * P.np + p0;
            const double *cp = P.cp + p0;
            // four samples per step, their gathers all in flight before the
            // arithmetic: indices are clamped into range (a sample outside the
            // image or the taps, or past k1, reads element 0) and such terms
            // are not added.  row[k1 .. k1 + kFeUnroll - 2] may lie past the
            // map's last row; the sample intervals and kFdkErrRowPad ints of
            // padding follow it in the same buffer, and those values are never
            // used as indices.
            for (int k = k0; k < k1; k += kFeUnroll) {
                int o[kFeUnroll], f[kFeUnroll];
                bool ok[kFeUnroll];
#pragma unroll
                for (int q = 0; q < kFeUnroll; ++q) {
                    const bool in = k + q < k1;
                    o[q] = row[k + q];
                    f[q] = fil[min(k + q, k1 - 1) * kFeThreads + threadIdx.x];
                    o[q] = in ? o[q] : -1;
                    ok[q] = in && f[q] >= 0 && o[q] >= 0;
                }
                float mv[kFeUnroll], sv[kFeUnroll];
                double gv[kFeUnroll];
#pragma unroll
                for (int q = 0; q < kFeUnroll; ++q) {
                    const size_t oz = (size_t)max(o[q], 0) * P.nz;
                    mv[q] = moved[oz];
                    sv[q] = slice[oz];
                    gv[q] = P.g[max(f[q], 0)];
                }
#pragma unroll
                for (int q = 0; q < kFeUnroll; ++q) {
                    const float diff = mv[q] - sv[q];
                    const double term = (double)diff * half_lr2 * (gv[q] * cp[min(k + q, k1 - 1)]);
                    acc = ok[q] ? acc + term : acc;
                }
            }
        }
    }
    if (s < P.ns) out[((size_t)b * P.nz + z) * P.ns + s] = (float)acc;
}

}  // namespace

hipError_t launch_fdk_error_slice(const FdkErrParams &P, const float *pic, int *xy, float *out, hipStream_t st)
{
    const size_t n = (size_t)P.nb * P.nu * P.np;
    hipLaunchKernelGGL(k_fdkerr_xy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, P, xy);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the (view, U) sample intervals follow the voxel map in the same scratch
    int2 *rng = reinterpret_cast<int2 *>(xy + ((n + 1) & ~(size_t)1));
    const int nbu = P.nb * P.nu;
    hipLaunchKernelGGL(k_fdkerr_range, dim3((nbu + 255) / 256), dim3(256), 0, st, P, xy, rng);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    float *pz = reinterpret_cast<float *>(xy + fdk_error_slice_zcopy_offset(P));
    const size_t nvox = (size_t)P.nt * P.ns * P.nz;
    hipLaunchKernelGGL(k_fdkerr_zcopy, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, st, P, pic, pz);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const dim3 grid((P.ns + kFeThreads - 1) / kFeThreads, P.nz, P.nb);
    hipLaunchKernelGGL(k_fdkerr, grid, dim3(kFeThreads), 0, st, P, pz, xy, rng, out);
    return hipGetLastError();
}

}  // namespace ctr
