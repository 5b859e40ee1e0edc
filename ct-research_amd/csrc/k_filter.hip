// k_filter.hip - FDK / fan-beam pre-weight and ramp filter for gfx950.
//
// out[p] = sum_q w[q] * h2[p - q + np - 1],  w[q] = R[q] * weight(q, row)
// with h2[k] = pitch * h[|k - (np-1)|] (the reference's PreWeightFilter,
// FDK_GPU/FDK.cu:21-48, computed race-free: every sample of the row is
// pre-weighted in LDS before any output reads it).
//
// One wavefront filters one detector row.  Lane l owns K consecutive
// outputs; for each input sample q the K taps it needs are a sliding window
// over h2, so each outer step of K samples loads K new taps and K inputs
// (LDS broadcast) for K*K FP64 FMAs.  FP64 FMA issues at the non-packed
// FP32 rate on gfx950, and the reference accumulates in double (:38-45).
#include "ctr_internal.h"

namespace ctr {
namespace {

template <int K>
struct FilterLayout {
    // 2K zeros in front: the window advance after the last input sample reads
    // down to h2 index -(2K-1); 64K + K zeros behind for the padding lanes
    static __host__ __device__ int front() { return 2 * K; }
    static __host__ __device__ int tap_len(int np) { return 2 * K + 2 * np - 1 + 64 * K + K; }
    static __host__ __device__ int row_len(int np) { return ((np + K - 1) / K) * K + K; }
};

template <int K, int RW>
__global__ __launch_bounds__(64 *RW) void k_row_filter(const float *__restrict__ in, float *__restrict__ out,
                                                       FilterParams P)
{
    extern __shared__ double sh[];
    const int np = P.np;
    const int TL = FilterLayout<K>::tap_len(np), RL = FilterLayout<K>::row_len(np);
    double *taps = sh;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double *wrow = sh + TL + wave * RL;
    // two-sided taps with zero padding in front and behind
    for (int t = threadIdx.x; t < TL; t += 64 * RW) {
        const int k = t - FilterLayout<K>::front();
        taps[t] = (k >= 0 && k < 2 * np - 1) ? P.taps[k] : 0.0;
    }
    const int row = blockIdx.x * RW + wave;
    const bool active = row < P.rows;
    if (active) {
        const int ix = row % P.nx;
        const float xi = P.xd[ix];
        const float *src = in + (size_t)row * np;
        for (int q = lane; q < RL; q += 64) {
            double v = 0.0;
            if (q < np) {
                float wgt;
                if (P.mode == 0) {
                    const float pq = P.pd[q];
                    wgt = P.D / sqrtf(P.D * P.D + pq * pq + xi * xi);
                } else {
                    wgt = P.colw[q];
                }
                v = (double)(src[q] * wgt);   // rounded to float like the in-place store of FDK.cu:36
            }
            wrow[q] = v;
        }
    }
    __syncthreads();
    if (!active) return;

    for (int base = 0; base < np; base += 64 * K) {
        const int p0 = base + lane * K;
        double acc[K];
#pragma unroll
        for (int j = 0; j < K; ++j) acc[j] = 0.0;
        // R[m] = h2[p0 - q - (K-1) + np - 1 + m], m = 0 .. 2K-2; output j at sub-step u uses R[K-1+j-u]
        double R[2 * K - 1];
        const int t0 = FilterLayout<K>::front() + p0 + np - 1 - (K - 1);   // taps[] index for q = 0, m = 0
#pragma unroll
        for (int m = 0; m < 2 * K - 1; ++m) R[m] = taps[t0 + m];
        for (int q = 0; q < np; q += K) {
#pragma unroll
            for (int u = 0; u < K; ++u) {
                const double wq = wrow[q + u];
#pragma unroll
                for (int j = 0; j < K; ++j) acc[j] = fma(wq, R[K - 1 + j - u], acc[j]);
            }
            // advance the window by K samples: new R[K..2K-2] = old R[0..K-2]
#pragma unroll
            for (int m = 2 * K - 2; m >= K; --m) R[m] = R[m - K];
            const int tq = t0 - (q + K);
#pragma unroll
            for (int m = 0; m < K; ++m) R[m] = taps[tq + m];
        }
        float *dst = out + (size_t)row * np;
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (p0 + j < np) dst[p0 + j] = (float)acc[j];
    }
}

template <int K, int RW>
hipError_t launch_k(const FilterParams &P, const float *in, float *out, hipStream_t st)
{
    const size_t bytes =
        sizeof(double) * ((size_t)FilterLayout<K>::tap_len(P.np) + (size_t)RW * FilterLayout<K>::row_len(P.np));
    const int nblk = (P.rows + RW - 1) / RW;
    hipLaunchKernelGGL((k_row_filter<K, RW>), dim3(nblk), dim3(64 * RW), bytes, st, in, out, P);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_row_filter(const FilterParams &P, const float *in, float *out, hipStream_t st)
{
    const int need = (P.np + 63) / 64;
    if (need <= 1) return launch_k<1, 4>(P, in, out, st);
    if (need <= 2) return launch_k<2, 4>(P, in, out, st);
    if (need <= 4) return launch_k<4, 4>(P, in, out, st);
    if (need <= 8) return launch_k<8, 4>(P, in, out, st);
    if (P.np <= 1024) return launch_k<16, 2>(P, in, out, st);
    return launch_k<16, 1>(P, in, out, st);
}

}  // namespace ctr
