// k_util.hip - small element-wise and reduction kernels (fill, reciprocal,
// deterministic two-level reductions of per-block partial sums).
#include "ctr_internal.h"

#include <algorithm>

#include <hipcub/hipcub.hpp>

namespace ctr {
namespace {

__global__ __launch_bounds__(256) void k_fill(float *dst, float value, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = value;
}

__global__ __launch_bounds__(256) void k_reciprocal_eps(float *v, size_t n, float eps)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        v[i] = 1.0f / (v[i] + eps);
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum of the pairs [lo, hi) by one block: per lane in a fixed order, then a
// fixed tree over the lanes and the four waves.  Thread 0 gets the pair.
__device__ __forceinline__ void block_pair_sum(const double *__restrict__ partials, int lo, int hi, double &sa,
                                               double &sb)
{
    __shared__ double red[2][4];
    double a = 0.0, b = 0.0;
    for (int i = lo + (int)threadIdx.x; i < hi; i += 256) {
        a += partials[2 * i];
        b += partials[2 * i + 1];
    }
    a = wave_sum(a);
    b = wave_sum(b);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    sa = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    sb = red[1][0] + red[1][1] + red[1][2] + red[1][3];
}

// Deterministic sum of n pairs.  Up to kReduceDirect pairs one block sums
// them; above, reduce_groups(n) blocks sum consecutive ranges of
// ceil(n / groups) pairs into partials[n ..] (stage 1) and one block sums
// those (stage 2).  The order depends on n only, so every caller that sums
// the same partials gets the same bits.
__host__ __device__ __forceinline__ int reduce_groups(int n)
{
    return n <= kReduceDirect ? 0 : min((int)kReduceSlack, (n + 2047) / 2048);
}

__global__ __launch_bounds__(256) void k_reduce_pairs(const double *partials, int n, double *out2)
{
    double a, b;
    block_pair_sum(partials, 0, n, a, b);
    if (threadIdx.x == 0) {
        out2[0] = a;
        out2[1] = b;
    }
}

// Stage 1 for job blockIdx.y (jobs == nullptr: the single sum of n pairs).
__global__ __launch_bounds__(256) void k_reduce_stage1(double *partials, int n, const ReduceJob *__restrict__ jobs)
{
    double *p = partials;
    int m = n;
    if (jobs) {
        p = const_cast<double *>(jobs[blockIdx.y].partials);
        m = jobs[blockIdx.y].n;
    }
    const int groups = reduce_groups(m);
    if ((int)blockIdx.x >= groups) return;
    const int chunk = (m + groups - 1) / groups;
    const int lo = (int)blockIdx.x * chunk, hi = min(m, lo + chunk);
    double a, b;
    block_pair_sum(p, lo, hi, a, b);
    if (threadIdx.x == 0) {
        p[2 * (m + (int)blockIdx.x)] = a;
        p[2 * (m + (int)blockIdx.x) + 1] = b;
    }
}

// One block per job: sums[(*iter) * njobs + p] = sum of job p's partial pairs
// (of its stage-1 pairs when it has more than kReduceDirect).
__global__ __launch_bounds__(256) void k_reduce_jobs(const ReduceJob *__restrict__ jobs, const int *__restrict__ iter,
                                                     double *__restrict__ sums, int njobs)
{
    const int p = blockIdx.x;
    const int n = jobs[p].n, groups = reduce_groups(n);
    double a, b;
    if (groups) block_pair_sum(jobs[p].partials + 2 * (size_t)n, 0, groups, a, b);
    else block_pair_sum(jobs[p].partials, 0, n, a, b);
    if (threadIdx.x == 0) {
        double *o = sums + 2 * ((size_t)(*iter) * njobs + p);
        o[0] = a;
        o[1] = b;
    }
}

// Advance the iteration counter of a captured SIRT iteration (vector store).
__global__ void k_iter_inc(int *iter)
{
    if (threadIdx.x == 0) iter[0] = iter[0] + 1;
}

// out[z][t][s] = in[z][s][t]: 32 x 32 tiles through LDS (padded row against bank conflicts).
__global__ __launch_bounds__(256) void k_transpose_st(const float *__restrict__ in, float *__restrict__ out, int nt,
                                                      int ns)
{
    __shared__ float tile[32][33];
    const int t0 = blockIdx.x * 32, s0 = blockIdx.y * 32;
    const size_t zo = (size_t)blockIdx.z * nt * ns;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int t = t0 + tx, s = s0 + k;
        if (t < nt && s < ns) tile[k][tx] = in[zo + (size_t)s * nt + t];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int s = s0 + tx, t = t0 + k;
        if (t < nt && s < ns) out[zo + (size_t)t * ns + s] = tile[tx][k];
    }
}

// Per-block (sum x^2, sum x).
__global__ __launch_bounds__(256) void k_sum_sq_sum(const float *x, size_t n, double *partials)
{
    __shared__ double red[2][4];
    double a = 0.0, b = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const double v = x[i];
        a += v * v;
        b += v;
    }
    a = wave_sum(a);
    b = wave_sum(b);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        partials[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

// Per-block (sum (x - ref)^2, sum ref^2) for rRMSE and LocalError.
__global__ __launch_bounds__(256) void k_diff_sums(const float *x, const float *ref, size_t n, double *partials)
{
    __shared__ double red[2][4];
    double a = 0.0, b = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const double r = ref[i], d = (double)x[i] - r;
        a += d * d;
        b += r * r;
    }
    a = wave_sum(a);
    b = wave_sum(b);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        partials[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

// Sum of ellipsoid indicators at the voxel centres (phantom3d.m:71-110), FP64
// in the order of the reference's expression, contraction off so that the
// boundary test is the plain IEEE one.
__global__ __launch_bounds__(256) void k_phantom(const PhantomEll *__restrict__ ell, int n_ell, int nt, int ns, int nz,
                                                 float *__restrict__ vol)
{
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t n = (size_t)nt * ns * nz;
    if (i >= n) return;
    const int it = (int)(i % nt), is = (int)((i / nt) % ns), iz = (int)(i / ((size_t)nt * ns));
    // grid on [-1, 1] with the end points: (k - (n-1)/2) / ((n-1)/2); t is MATLAB's y, s its x
    const double y = nt > 1 ? (it - (nt - 1) / 2.0) / ((nt - 1) / 2.0) : 0.0;
    const double x = ns > 1 ? (is - (ns - 1) / 2.0) / ((ns - 1) / 2.0) : 0.0;
    const double z = nz > 1 ? (iz - (nz - 1) / 2.0) / ((nz - 1) / 2.0) : 0.0;
    double acc = 0.0;
    for (int k = 0; k < n_ell; ++k) {
        const PhantomEll e = ell[k];
        const double p = e.cp * x + e.sp * y - e.sx, q = -e.sp * x + e.cp * y - e.sy, r = z - e.z0;
        const double u = p * p / (e.a * e.a) + q * q / (e.b * e.b);
        const double w = r * r / (e.c * e.c);
        if (u + w <= 1.0) acc += e.A;
    }
    vol[i] = (float)acc;
}

int grid_for(size_t n)
{
    size_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    return (int)g;
}

}  // namespace

// Copy a wx x wy x wz box between two volumes ([z][y][x], x fastest): from
// src (row sx, plane sx*sy) at (sx0, sy0, sz0) to dst (row dx, plane dx*dy)
// at (dx0, dy0, dz0).  Both boxes lie inside their volumes.
__global__ __launch_bounds__(256) void k_box_copy(const float *__restrict__ src, int sx, int sy, int sx0, int sy0,
                                                  int sz0, float *__restrict__ dst, int dx, int dy, int dx0, int dy0,
                                                  int dz0, int wx, int wy, int wz)
{
    const size_t n = (size_t)wx * wy * wz;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) {
        const int x = (int)(e % wx);
        const size_t r = e / wx;
        const int y = (int)(r % wy), z = (int)(r / wy);
        dst[((size_t)(dz0 + z) * dy + (dy0 + y)) * dx + (dx0 + x)] = src[((size_t)(sz0 + z) * sy + (sy0 + y)) * sx + (sx0 + x)];
    }
}

// ---- restricted Chebyshev emulation of SIRT (ctr_sirt3d_emulated) ----------
__global__ __launch_bounds__(256) void k_threshold_mask(const float *__restrict__ v, float thr, size_t n,
                                                        float *__restrict__ mask)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) mask[i] = v[i] > thr ? 1.0f : 0.0f;
}

__global__ __launch_bounds__(256) void k_mul(const float *__restrict__ a, const float *__restrict__ m, size_t n,
                                             float *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = a[i] * m[i];
}

// w0 = mask lambda V g (the masked plain step, g = A' W (b - A x)), acc = c0 w0
__global__ __launch_bounds__(256) void k_cheb_start(const float *__restrict__ g, const float *__restrict__ vinv,
                                                    const float *__restrict__ mask, float lambda, float c0, size_t n,
                                                    float *__restrict__ w0, float *__restrict__ acc)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float w = mask[i] * (lambda * vinv[i] * g[i]);
    w0[i] = w;
    acc[i] = c0 * w;
}

// One term of the recurrence w_{j+1} = 2 (2M - I) w_j - w_{j-1} (w_1 = (2M - I) w_0),
// restricted to the mask: g = A' (-W A w_j), so M w_j = -V g.  w_{j+1}
// overwrites w_{j-1} in prev; acc += c w_{j+1}.
__global__ __launch_bounds__(256) void k_cheb_step(const float *__restrict__ g, const float *__restrict__ vinv,
                                                   const float *__restrict__ mask, const float *__restrict__ cur,
                                                   float *__restrict__ prev, float *__restrict__ acc, float c,
                                                   int first, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float xw = 2.0f * (-vinv[i] * g[i]) - cur[i];   // (2M - I) w_j
    const float wn = mask[i] * (first ? xw : 2.0f * xw - prev[i]);
    prev[i] = wn;
    acc[i] = fmaf(c, wn, acc[i]);
}

__global__ __launch_bounds__(256) void k_add_clamp(const float *__restrict__ a, const float *__restrict__ b, float lo,
                                                   float hi, size_t n, float *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = fminf(fmaxf(a[i] + b[i], lo), hi);
}

hipError_t launch_box_copy(const float *src, int sx, int sy, int sx0, int sy0, int sz0, float *dst, int dx, int dy,
                           int dx0, int dy0, int dz0, int wx, int wy, int wz, hipStream_t st)
{
    const size_t n = (size_t)wx * wy * wz;
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_box_copy, dim3(blocks), dim3(256), 0, st, src, sx, sy, sx0, sy0, sz0, dst, dx, dy, dx0, dy0,
                       dz0, wx, wy, wz);
    return hipGetLastError();
}

static unsigned ew_blocks(size_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_threshold_mask(const float *v, float thr, size_t n, float *mask, hipStream_t st)
{
    hipLaunchKernelGGL(k_threshold_mask, dim3(ew_blocks(n)), dim3(256), 0, st, v, thr, n, mask);
    return hipGetLastError();
}

hipError_t launch_mul(const float *a, const float *m, size_t n, float *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_mul, dim3(ew_blocks(n)), dim3(256), 0, st, a, m, n, out);
    return hipGetLastError();
}

hipError_t launch_cheb_start(const float *g, const float *vinv, const float *mask, float lambda, float c0, size_t n,
                             float *w0, float *acc, hipStream_t st)
{
    hipLaunchKernelGGL(k_cheb_start, dim3(ew_blocks(n)), dim3(256), 0, st, g, vinv, mask, lambda, c0, n, w0, acc);
    return hipGetLastError();
}

hipError_t launch_cheb_step(const float *g, const float *vinv, const float *mask, const float *cur, float *prev,
                            float *acc, float c, int first, size_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_cheb_step, dim3(ew_blocks(n)), dim3(256), 0, st, g, vinv, mask, cur, prev, acc, c, first, n);
    return hipGetLastError();
}

hipError_t launch_add_clamp(const float *a, const float *b, float lo, float hi, size_t n, float *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_add_clamp, dim3(ew_blocks(n)), dim3(256), 0, st, a, b, lo, hi, n, out);
    return hipGetLastError();
}

hipError_t launch_fill(float *dst, float value, size_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, st, dst, value, n);
    return hipGetLastError();
}

hipError_t launch_reciprocal_eps(float *v, size_t n, float eps, hipStream_t st)
{
    hipLaunchKernelGGL(k_reciprocal_eps, dim3(grid_for(n)), dim3(256), 0, st, v, n, eps);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(double *partials, int n, double *out2, hipStream_t st)
{
    const int groups = reduce_groups(n);
    if (groups == 0) {
        hipLaunchKernelGGL(k_reduce_pairs, dim3(1), dim3(256), 0, st, partials, n, out2);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_reduce_stage1, dim3(groups), dim3(256), 0, st, partials, n, (const ReduceJob *)nullptr);
    hipLaunchKernelGGL(k_reduce_pairs, dim3(1), dim3(256), 0, st, partials + 2 * (size_t)n, groups, out2);
    return hipGetLastError();
}

hipError_t launch_reduce_jobs(const ReduceJob *jobs, int njobs, const int *iter, double *sums, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_stage1, dim3(kReduceSlack, njobs), dim3(256), 0, st, (double *)nullptr, 0, jobs);
    hipLaunchKernelGGL(k_reduce_jobs, dim3(njobs), dim3(256), 0, st, jobs, iter, sums, njobs);
    return hipGetLastError();
}

hipError_t launch_transpose_st(const float *in, float *out, int nt, int ns, int nz, hipStream_t st)
{
    hipLaunchKernelGGL(k_transpose_st, dim3((nt + 31) / 32, (ns + 31) / 32, nz), dim3(256), 0, st, in, out, nt, ns);
    return hipGetLastError();
}

hipError_t launch_iter_inc(int *iter, hipStream_t st)
{
    hipLaunchKernelGGL(k_iter_inc, dim3(1), dim3(64), 0, st, iter);
    return hipGetLastError();
}

hipError_t launch_sum_sq_sum(const float *x, size_t n, double *partials, int *nblocks, hipStream_t st)
{
    const int g = grid_for(n);
    *nblocks = g;
    hipLaunchKernelGGL(k_sum_sq_sum, dim3(g), dim3(256), 0, st, x, n, partials);
    return hipGetLastError();
}

hipError_t launch_diff_sums(const float *x, const float *ref, size_t n, double *partials, int *nblocks, hipStream_t st)
{
    const int g = grid_for(n);
    *nblocks = g;
    hipLaunchKernelGGL(k_diff_sums, dim3(g), dim3(256), 0, st, x, ref, n, partials);
    return hipGetLastError();
}

hipError_t launch_phantom(const PhantomEll *ell, int n_ell, int nt, int ns, int nz, float *vol, hipStream_t st)
{
    const size_t n = (size_t)nt * ns * nz;
    hipLaunchKernelGGL(k_phantom, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ell, n_ell, nt, ns, nz, vol);
    return hipGetLastError();
}

hipError_t scan_counts(const long long *counts, long long *out, int n, void *tmp, size_t &tmp_bytes, hipStream_t st)
{
    return hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, counts, out, n, st);
}

}  // namespace ctr
