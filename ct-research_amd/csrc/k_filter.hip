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

#include <algorithm>
#include <cstdlib>
#include <mutex>

namespace ctr {
namespace {

// FDK cosine pre-weight of FDK.cu:36 and the float store of the weighted
// sample: r * D / sqrt(D^2 + p^2 + xi^2) formed in FP64 operation by
// operation (no contraction) from the float detector positions, rounded to
// float once.  The ramp-filtered row is ~300 times smaller in norm than its
// input, so a last-bit difference in the stored samples shows at ~3e-6 in
// the output: the same arithmetic as the FP64 oracle keeps it out.
__device__ __forceinline__ double cone_weight64(double D, float p, float xi)
{
    const double pd = p, xd = xi;
    return D / sqrt(__dadd_rn(__dadd_rn(__dmul_rn(D, D), __dmul_rn(pd, pd)), __dmul_rn(xd, xd)));
}
__device__ __forceinline__ float cone_preweight(float r, double w) { return (float)__dmul_rn((double)r, w); }

__global__ __launch_bounds__(256) void k_cone_weights(double D, const float *__restrict__ pd,
                                                      const float *__restrict__ xd, int np, int nx,
                                                      double *__restrict__ pw)
{
    const long n = (long)np * nx;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256)
        pw[e] = cone_weight64(D, pd[e % np], xd[e / np]);
}

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
        const float xi = P.mode == 0 ? P.xd[ix] : 0.0f;
        const float *src = in + (size_t)row * np;
        for (int q = lane; q < RL; q += 64) {
            double v = 0.0;
            if (q < np) {
                if (P.mode == 0) {
                    const double w = P.pw ? P.pw[(size_t)ix * np + q] : cone_weight64(P.Dd, P.pd[q], xi);
                    v = (double)cone_preweight(src[q], w);
                } else {
                    const float wgt = P.mode == 1 ? P.colw[q] : 1.0f;
                    v = (double)(src[q] * wgt);   // rounded to float like the in-place store of FDK.cu:36
                }
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

// ---------------------------------------------------------------------------
// FFT path.  The linear convolution of a row (np samples) with the two-sided
// taps h2 equals a circular convolution of length L >= 2np - 1 with
// c[m] = c[L-m] = h2[np-1+m] (m < np), zero elsewhere.  c is real and even,
// so its DFT H is real: two real rows packed as z = x0 + i x1 are filtered by
// one complex transform pair, z' = IDFT(DFT(z) H) = y0 + i y1, computed as
// conj(DFT(conj(DFT(z)) H / L)).  FP64 throughout; H / L comes from the host.
//
// Two transform kernels.  k_fft16_filter (below; radix 16, one transform per
// wave for L <= 1024) serves the flat-detector modes.  k_fft_filter serves
// the curved-detector mode, whose rows pair by detector row across views and
// whose Parker weights are staged through LDS: Stockham passes of radix 8,
// then one of radix 2 or 4 when L is not a power of 8, over G = L/8 threads
// that each hold 8 points.  After the last pass thread j holds X[j + G t]
// (t = 0..7), which is exactly the input distribution of the first pass, so
// the spectrum multiply and the second transform follow in registers.  LDS
// holds one padded transform (index i at i + i/8) per group of G threads;
// about 100 registers, four workgroups of 256 threads per CU.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 mul_mi(double2 a) { return make_double2(a.y, -a.x); }   // a * (-i)

// forward DFT (exponent -2 pi i n k / R), natural order in and out
__device__ __forceinline__ void dft2(double2 &a, double2 &b)
{
    const double2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

__device__ __forceinline__ void dft4(double2 &a0, double2 &a1, double2 &a2, double2 &a3)
{
    const double2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_mi(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

constexpr double kR2 = 0.70710678118654752440;   // sqrt(1/2)

__device__ __forceinline__ void dft8(double2 *v)
{
    // even/odd split: X[k] = E[k] + w8^k O[k], X[k+4] = E[k] - w8^k O[k]
    double2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    o1 = cmul(o1, make_double2(kR2, -kR2));
    o2 = mul_mi(o2);
    o3 = cmul(o3, make_double2(-kR2, -kR2));
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
}

__device__ __forceinline__ int fft_pad(int i) { return i + (i >> 3); }

// w = exp(-2 pi i num / den)
__device__ __forceinline__ double2 twiddle(int num, int den)
{
    double s, c;
    sincospi(-2.0 * (double)num / (double)den, &s, &c);
    return make_double2(c, s);
}

constexpr int kFftThreads = 256;

// Pass structure of a length-L transform: radix-8 passes with strides
// Ns = 1, 8 (, 64), then a radix-RL pass (RL = 1: none) with Ns = 8^NP8.
template <int L>
struct FftPlan {
    static constexpr int G = L / 8;                           // threads per transform
    static constexpr int NP8 = L == 256 ? 2 : 3;              // radix-8 passes
    static constexpr int RL = L / (NP8 == 2 ? 64 : 512);      // 4, 1, 2, 4 for L = 256 .. 2048
    static constexpr int NBF = RL > 1 ? 8 / RL : 1;           // last-pass butterflies per thread
};

// Twiddle bases of the passes, in a block-shared LDS table filled once per
// block (kept in registers they held 24 VGPRs across the whole row loop and
// cost the kernel a wave per SIMD): [0, 8) w_64^i (radix-8 pass 2), [8, 72)
// w_512^i (pass 3), [72, 72 + G) w_L^j.  The last pass's butterfly j + G m
// needs w_L^(j + G m) = w_L^j w_8^m (G = L/8).  Powers w^r are formed by
// recurrence in each pass (relative error below 8 ulp).
template <int L>
struct FftTwiddles {
    static constexpr int kOff512 = 8, kOffL = 72;
    static constexpr int kSize = kOffL + (FftPlan<L>::RL > 1 ? FftPlan<L>::G : 0);
};

template <int L>
__device__ __forceinline__ void fft_twiddle_table(double2 *tab)
{
    using T = FftTwiddles<L>;
    for (int i = threadIdx.x; i < T::kSize; i += kFftThreads) {
        if (i < T::kOff512) tab[i] = twiddle(i, 64);
        else if (i < T::kOffL) tab[i] = twiddle(i - T::kOff512, 512);
        else tab[i] = twiddle(i - T::kOffL, L);
    }
}

// w_8^m = exp(-2 pi i m / 8)
__device__ __forceinline__ double2 w8(int m)
{
    switch (m & 7) {
    case 0: return make_double2(1.0, 0.0);
    case 1: return make_double2(kR2, -kR2);
    case 2: return make_double2(0.0, -1.0);
    case 3: return make_double2(-kR2, -kR2);
    case 4: return make_double2(-1.0, 0.0);
    case 5: return make_double2(-kR2, kR2);
    case 6: return make_double2(0.0, 1.0);
    default: return make_double2(kR2, kR2);
    }
}

// v[r] *= w^r for r = 1..n-1
__device__ __forceinline__ void twiddle_powers(double2 *v, int n, double2 w0)
{
    double2 w = w0;
    for (int r = 1; r < n; ++r) {
        v[r] = cmul(v[r], w);
        if (r + 1 < n) w = cmul(w, w0);
    }
}

// Radix-8 pass with stride Ns on v[r] = in[j + G r]; the result goes to the
// group's LDS slot at (j / Ns) 8 Ns + j mod Ns + r Ns, and the next pass's
// input v[r] = out[j + G r] is read back.
template <int L>
__device__ __forceinline__ void fft8_pass(double2 *v, double2 *buf, int j, int Ns)
{
    constexpr int G = FftPlan<L>::G;
    dft8(v);
    const int k = j & (Ns - 1);
    const int d = (j - k) * 8 + k;
    __syncthreads();   // the group's previous readers of buf are done
#pragma unroll
    for (int r = 0; r < 8; ++r) buf[fft_pad(d + r * Ns)] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = buf[fft_pad(j + G * r)];
}

// One forward transform of the 8 points v[t] = x[j + G t] of this thread's
// group; on return v[t] = X[j + G t].  buf is the group's padded LDS slot,
// tab the block's twiddle table.
template <int L>
__device__ __forceinline__ void fft_group(double2 *v, double2 *buf, int j, const double2 *tab)
{
    using F = FftPlan<L>;
    using T = FftTwiddles<L>;
    fft8_pass<L>(v, buf, j, 1);
    if constexpr (F::NP8 == 2) {
        // L = 256: radix 8 (Ns = 8), then radix 4 (Ns = 64)
        twiddle_powers(v, 8, tab[j & 7]);
        fft8_pass<L>(v, buf, j, 8);
    } else {
        twiddle_powers(v, 8, tab[j & 7]);
        fft8_pass<L>(v, buf, j, 8);
        twiddle_powers(v, 8, tab[T::kOff512 + (j & 63)]);
        if constexpr (F::RL == 1) {
            // L = 512: the third radix-8 pass is the last; thread j's outputs are
            // X[j + 64 r] = X[j + G r], already in place
            dft8(v);
        } else {
            fft8_pass<L>(v, buf, j, 64);
        }
    }
    if constexpr (F::RL > 1) {
        // butterflies b = j + G m (m < NBF) read X'[b + r L/RL] = v[m + NBF r] and
        // write X[b + r L/RL] to the same registers
#pragma unroll
        for (int m = 0; m < F::NBF; ++m) {
            double2 u[F::RL];
#pragma unroll
            for (int r = 0; r < F::RL; ++r) u[r] = v[m + F::NBF * r];
            twiddle_powers(u, F::RL, m ? cmul(tab[T::kOffL + j], w8(m)) : tab[T::kOffL + j]);
            if constexpr (F::RL == 2) dft2(u[0], u[1]);
            else dft4(u[0], u[1], u[2], u[3]);
#pragma unroll
            for (int r = 0; r < F::RL; ++r) v[m + F::NBF * r] = u[r];
        }
    }
}

// Curved-detector short scan (mode 3): rows pair by detector row across two
// consecutive views (one spectrum per detector row).
template <int L>
__global__ __launch_bounds__(kFftThreads) void k_fft_filter_curved(const float *__restrict__ in,
                                                                  float *__restrict__ out, FilterParams P)
{
    constexpr int G = FftPlan<L>::G, GPB = kFftThreads / G;
    constexpr int SLOT = L + L / 8;
    __shared__ double2 lds[GPB * SLOT];
    __shared__ double2 tab[FftTwiddles<L>::kSize];
    const int grp = threadIdx.x / G, j = threadIdx.x % G;
    double2 *buf = lds + grp * SLOT;
    const int np = P.np;
    fft_twiddle_table<L>(tab);   // published by the first barrier of the row loop

    const int nviews = P.rows / P.nx;
    const int npairs = P.nx * ((nviews + 1) / 2);
    for (int base = blockIdx.x * GPB; base < npairs; base += gridDim.x * GPB) {
        const int pair = base + grp;   // every group takes part in the barriers, active or not
        const int row0 = (2 * (pair / P.nx)) * P.nx + pair % P.nx;
        const int row1 = row0 + P.nx;
        const bool has0 = pair < npairs, has1 = has0 && row1 < P.rows;
        const double *spec = P.spec + (has0 ? (size_t)(row0 % P.nx) * P.spec_stride : 0);
        // row scale and the scaled view angles of the two rows
        const double prow = has0 ? P.rowp[row0 % P.nx] : 1.0;
        const double bp0 = has0 ? P.betad[row0 / P.nx] * prow : 0.0;
        const double bp1 = has1 ? P.betad[row1 / P.nx] * prow : 0.0;
        const float *src0 = in + (size_t)row0 * np, *src1 = in + (size_t)row1 * np;
        double2 v[8];
        // the Parker-weighted samples go through LDS: a rolled loop keeps the
        // weight arithmetic of one sample live at a time
        __syncthreads();   // the group's previous readers of buf are done
#pragma unroll 1
        for (int t = 0; t < 8; ++t) {
            const int q = j + G * t;
            double a = 0.0, b = 0.0;
            if (q < np) {
                // ImplementUI.cu:43-66: (D/P) cos(g P) Parker(b P, start P, g P, fan P)
                const double gp = P.gamd[q] * prow;
                const float base_w = (float)((double)P.D / prow) * cosf((float)gp);
                const double stp = P.start * prow, fanp = P.fan * prow;
                const float w0 = base_w * parker_weight(bp0, stp, gp, fanp);
                const float w1 = base_w * parker_weight(bp1, stp, gp, fanp);
                // the reference keeps the weighted samples in FP64 (ImplementUI.cu:66)
                if (has0) a = (double)src0[q] * (double)w0;
                if (has1) b = (double)src1[q] * (double)w1;
            }
            buf[fft_pad(q)] = make_double2(a, b);
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 8; ++t) v[t] = buf[fft_pad(j + G * t)];
        fft_group<L>(v, buf, j, tab);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const double h = spec[j + G * t];   // H[k] / L, real
            v[t] = make_double2(v[t].x * h, -v[t].y * h);
        }
        fft_group<L>(v, buf, j, tab);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int q = j + G * t;
            if (q < np) {
                if (has0) out[(size_t)row0 * np + q] = (float)v[t].x;
                if (has1) out[(size_t)row1 * np + q] = (float)(-v[t].y);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Radix-16 transforms of length L = 256 .. 2048 over G = L/16 lanes that each
// hold 16 points: Stockham passes of radix 16 (Ns = 1), 16 (Ns = 16) and,
// for L > 256, one of radix R3 = L/256 (Ns = 256) in registers.  Lane j holds
// v[r] = x[j + G r] before and after a transform.  The two exchanges of a
// transform go through one padded array of 8-byte words (index i at
// i + i/16, no bank conflicts for any of the three access patterns), real
// parts first, then imaginary parts: L = 1024 needs 8.7 KB per wave, so
// sixteen waves fit a CU.  For L <= 1024 a transform lives in one wave (one
// to four transforms per wave), so no wave waits for another; L = 2048 takes
// two waves.  Rows are zero-padded to L >= 2np - 1, so the forward
// transform's first pass sees zeros in v[8..15] (x[j + G r] with
// j + G r >= L/2 >= np), and of the inverse only outputs j + G t with t < 8
// are kept (the compiler drops the rest).  Against the radix-8 kernel above
// at L = 1024: two exchanges per transform instead of three, and no barrier
// between waves.
// ---------------------------------------------------------------------------
namespace f16 {
template <int L>
struct Plan {
    static_assert(L >= 256 && L <= 2048 && (L & (L - 1)) == 0, "length");
    static constexpr int G = L / 16;                     // lanes per transform
    static constexpr int R3 = L / 256;                   // radix of the last pass (1: none)
    static constexpr int NB = 256 / G;                   // last-pass butterflies per lane
    static constexpr int kWords = L + L / 16;            // padded words of one transform
    static constexpr int kThreads = G < 64 ? 64 : G;     // workgroup
    static constexpr int kPerBlock = kThreads / G;       // transforms per workgroup
};

constexpr double kC = 0.92387953251128675613, kS = 0.38268343236508977173;   // cos(pi/8), sin(pi/8)

// a * w16^E with w16 = exp(-2 pi i / 16)
template <int E>
__device__ __forceinline__ double2 mul_w16(double2 a)
{
    constexpr int e = E & 15;
    if constexpr (e == 0) return a;
    else if constexpr (e == 4) return mul_mi(a);
    else if constexpr (e == 8) return make_double2(-a.x, -a.y);
    else if constexpr (e == 12) return make_double2(-a.y, a.x);
    else if constexpr (e == 2) return make_double2((a.x + a.y) * kR2, (a.y - a.x) * kR2);
    else if constexpr (e == 6) return make_double2((a.y - a.x) * kR2, -(a.x + a.y) * kR2);
    else {
        // cos(2 pi e / 16), -sin(2 pi e / 16)
        constexpr double c[16] = {1, kC, kR2, kS, 0, -kS, -kR2, -kC, -1, -kC, -kR2, -kS, 0, kS, kR2, kC};
        constexpr double s[16] = {0, -kS, -kR2, -kC, -1, -kC, -kR2, -kS, 0, kS, kR2, kC, 1, kC, kR2, kS};
        return cmul(a, make_double2(c[e], s[e]));
    }
}

// forward 16-point DFT, natural order in and out, as 4 x 4: DFTs over
// n2 of v[n1 + 4 n2], twiddles w16^(n1 k2), DFTs over n1.  ZERO_HI: the
// inputs v[8..15] are zero.
template <bool ZERO_HI>
__device__ __forceinline__ void dft16(double2 *v)
{
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
        if constexpr (ZERO_HI) {
            const double2 a = v[n1], b = v[n1 + 4], mb = mul_mi(b);
            v[n1] = cadd(a, b);
            v[n1 + 4] = cadd(a, mb);
            v[n1 + 8] = csub(a, b);
            v[n1 + 12] = csub(a, mb);
        } else {
            dft4(v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]);
        }
    }
    v[5] = mul_w16<1>(v[5]);
    v[9] = mul_w16<2>(v[9]);
    v[13] = mul_w16<3>(v[13]);
    v[6] = mul_w16<2>(v[6]);
    v[10] = mul_w16<4>(v[10]);
    v[14] = mul_w16<6>(v[14]);
    v[7] = mul_w16<3>(v[7]);
    v[11] = mul_w16<6>(v[11]);
    v[15] = mul_w16<9>(v[15]);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) dft4(v[4 * k2], v[4 * k2 + 1], v[4 * k2 + 2], v[4 * k2 + 3]);
    // v[4 k2 + k1] holds X[k2 + 4 k1]
    double2 t[16];
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2)
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) t[k2 + 4 * k1] = v[4 * k2 + k1];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = t[r];
}

// Orders the LDS accesses of the lanes of one transform.  Within one wave
// the LDS instructions execute in issue order, so a wave needs only a
// compiler fence (no s_waitcnt, no s_barrier); a transform spanning two
// waves (MULTI) needs the workgroup barrier.
template <bool MULTI>
__device__ __forceinline__ void group_sync()
{
    if constexpr (MULTI) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// words[wbase + WSTEP r] = v[r], then v[r] = words[rbase + RSTEP r], real
// parts first, then imaginary parts.  The last sync keeps the next writes
// to words behind this exchange's reads.
template <int WSTEP, int RSTEP, bool MULTI>
__device__ __forceinline__ void exchange(double2 *v, double *words, int wbase, int rbase)
{
#pragma unroll
    for (int r = 0; r < 16; ++r) words[wbase + WSTEP * r] = v[r].x;
    group_sync<MULTI>();
    double re[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) re[r] = words[rbase + RSTEP * r];
    group_sync<MULTI>();
#pragma unroll
    for (int r = 0; r < 16; ++r) words[wbase + WSTEP * r] = v[r].y;
    group_sync<MULTI>();
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = make_double2(re[r], words[rbase + RSTEP * r]);
    group_sync<MULTI>();
}

// One forward transform of this lane's group; words is the group's LDS
// array, t256 = w_256^(j mod 16), tL = w_L^j.  ZERO_HI: x[j + G r] = 0 for
// r >= 8.
template <int L, bool ZERO_HI>
__device__ __forceinline__ void transform(double2 *v, double *words, int j, double2 t256, double2 tL)
{
    using F = Plan<L>;
    constexpr int RSTEP = F::G + F::G / 16;   // padded stride of x[j + G r]
    const int rbase = j + (j >> 4);
    // the powers of the twiddle bases are formed anew in each transform:
    // hoisted out of the row loop they would hold about 100 registers
    asm volatile("" : "+v"(t256.x), "+v"(t256.y), "+v"(tL.x), "+v"(tL.y));
    dft16<ZERO_HI>(v);
    // pass 1 (Ns = 1): output r of lane j goes to 16 j + r, padded 17 j + r
    exchange<1, RSTEP, (F::G > 64)>(v, words, 17 * j, rbase);
    twiddle_powers(v, 16, t256);
    dft16<false>(v);
    // pass 2 (Ns = 16): to (j / 16) 256 + j mod 16 + 16 r, padded (j / 16) 272 + j mod 16 + 17 r
    exchange<17, RSTEP, (F::G > 64)>(v, words, (j >> 4) * 272 + (j & 15), rbase);
    if constexpr (F::R3 > 1) {
        // pass 3 (Ns = 256, radix R3): butterfly b = j + G m reads v[m + NB r] =
        // x'[b + 256 r] and twiddles by w_L^(b r), where w_L^b = w_L^j w16^m
#pragma unroll
        for (int m = 0; m < F::NB; ++m) {
            double2 base = tL;
            if (m == 1) base = mul_w16<1>(tL);
            if (m == 2) base = mul_w16<2>(tL);
            if (m == 3) base = mul_w16<3>(tL);
            if (m == 4) base = mul_w16<4>(tL);
            if (m == 5) base = mul_w16<5>(tL);
            if (m == 6) base = mul_w16<6>(tL);
            if (m == 7) base = mul_w16<7>(tL);
            double2 u[F::R3];
#pragma unroll
            for (int r = 0; r < F::R3; ++r) u[r] = v[m + F::NB * r];
            twiddle_powers(u, F::R3, base);
            if constexpr (F::R3 == 2) dft2(u[0], u[1]);
            else if constexpr (F::R3 == 4) dft4(u[0], u[1], u[2], u[3]);
            else dft8(u);
#pragma unroll
            for (int r = 0; r < F::R3; ++r) v[m + F::NB * r] = u[r];
        }
    }
}
}  // namespace f16

// MODE as k_fft_filter (0, 1, 2, 4; rows are paired as 2p, 2p + 1); each
// group of G lanes filters one row pair per step.
// Four waves per SIMD (128 registers); the FP64 weights formed per sample
// (MODE 0) need more, so three.
template <int L, int MODE>
__global__ __launch_bounds__(f16::Plan<L>::kThreads)
__attribute__((amdgpu_waves_per_eu(MODE == 0 ? 3 : 4, MODE == 0 ? 3 : 4)))
void k_fft16_filter(const float *__restrict__ in, float *__restrict__ out, FilterParams P)
{
    using F = f16::Plan<L>;
    __shared__ double words[F::kPerBlock * F::kWords];
    // with one group per workgroup the row pair is uniform (scalar index arithmetic)
    const int grp = F::kPerBlock == 1 ? 0 : threadIdx.x / F::G, j = threadIdx.x % F::G, np = P.np;
    double *mine = words + grp * F::kWords;
    const double2 t256 = twiddle(j & 15, 256), tL = twiddle(j, L);
    const int npairs = (P.rows + 1) / 2;
    // every group takes part in the barriers of a two-wave transform; with
    // several groups per wave a group past the last pair computes on zeros
    for (int base = blockIdx.x * F::kPerBlock; base < npairs; base += gridDim.x * F::kPerBlock) {
        const int pair = base + grp;
        const int row0 = 2 * pair, row1 = row0 + 1;
        // one group per workgroup: the loop bound already keeps pair < npairs
        const bool has0 = F::kPerBlock == 1 || pair < npairs, has1 = row1 < P.rows;
        // every load is in bounds (absent rows and columns read a stand-in
        // and are zeroed after), so all of them issue before the first wait
        const int rd0 = has0 ? row0 : 0, rd1 = has1 ? row1 : rd0;
        // the lane index, opaque per row pair: addresses formed from it are
        // recomputed per pair rather than hoisted out of the loop and spilled
        int jr = j;
        asm volatile("" : "+v"(jr));
        const float *src0 = in + (size_t)rd0 * np, *src1 = in + (size_t)rd1 * np;
        const float xi0 = MODE == 0 ? P.xd[rd0 % P.nx] : 0.0f;
        const float xi1 = MODE == 0 ? P.xd[rd1 % P.nx] : 0.0f;
        const double *pw0 = MODE == 4 ? P.pw + (size_t)(rd0 % P.nx) * np : nullptr;
        const double *pw1 = MODE == 4 ? P.pw + (size_t)(rd1 % P.nx) * np : nullptr;
        double2 v[16];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int q = jr + F::G * t, qc = q < np ? q : np - 1;
            double a, b;
            if constexpr (MODE == 0) {
                const float pq = P.pd[qc];
                a = (double)cone_preweight(src0[qc], cone_weight64(P.Dd, pq, xi0));
                b = (double)cone_preweight(src1[qc], cone_weight64(P.Dd, pq, xi1));
            } else if constexpr (MODE == 4) {
                a = (double)cone_preweight(src0[qc], pw0[qc]);
                b = (double)cone_preweight(src1[qc], pw1[qc]);
            } else {
                const float w = MODE == 1 ? P.colw[qc] : 1.0f;
                // pre-weighted samples rounded to float like the in-place store of FDK.cu:36
                a = (double)(src0[qc] * w);
                b = (double)(src1[qc] * w);
            }
            v[t] = make_double2(q < np && has0 ? a : 0.0, q < np && has1 ? b : 0.0);
            v[t + 8] = make_double2(0.0, 0.0);
        }
        f16::transform<L, true>(v, mine, j, t256, tL);
        // H[k] / L (real) at k = j + G t, from a uniform base and a 32-bit offset
        const char *hbase = reinterpret_cast<const char *>(P.spec);
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const double h = *reinterpret_cast<const double *>(hbase + 8u * (unsigned)(jr + F::G * t));
            v[t] = make_double2(v[t].x * h, -v[t].y * h);
        }
        f16::transform<L, false>(v, mine, j, t256, tL);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int q = jr + F::G * t;
            if (q < np) {
                if (has0) out[(size_t)row0 * np + q] = (float)v[t].x;
                if (has1) out[(size_t)row1 * np + q] = (float)(-v[t].y);
            }
        }
    }
}

template <int L>
hipError_t launch_fft(const FilterParams &P, const float *in, float *out, hipStream_t st)
{
    if (P.mode != 3 && !P.spec_stride) {
        using F = f16::Plan<L>;
        const int npairs = (P.rows + 1) / 2, need = (npairs + F::kPerBlock - 1) / F::kPerBlock;
        const int cap = 256 * 32 * 64 / F::kThreads;   // pair stride beyond 32 waves per CU
        const int nblk = need < cap ? need : cap;
        const dim3 grid(nblk), block(F::kThreads);
        switch (P.mode) {
        case 0:
            // mode 4: mode 0 with the FP64 weight table
            if (P.pw) hipLaunchKernelGGL((k_fft16_filter<L, 4>), grid, block, 0, st, in, out, P);
            else hipLaunchKernelGGL((k_fft16_filter<L, 0>), grid, block, 0, st, in, out, P);
            break;
        case 1: hipLaunchKernelGGL((k_fft16_filter<L, 1>), grid, block, 0, st, in, out, P); break;
        default: hipLaunchKernelGGL((k_fft16_filter<L, 2>), grid, block, 0, st, in, out, P); break;
        }
        return hipGetLastError();
    }
    if (P.mode != 3 || !P.spec_stride) return hipErrorInvalidValue;
    constexpr int GPB = kFftThreads / FftPlan<L>::G;
    const int npairs = P.nx * ((P.rows / P.nx + 1) / 2);
    const int need = (npairs + GPB - 1) / GPB;
    const int nblk = need < 256 * 16 ? need : 256 * 16;   // grid-stride beyond 16 blocks per CU
    hipLaunchKernelGGL((k_fft_filter_curved<L>), dim3(nblk), dim3(kFftThreads), 0, st, in, out, P);
    return hipGetLastError();
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

int fft_filter_length(int np)
{
    int L = 256;
    while (L < 2 * np - 1) L *= 2;
    return L <= 2048 ? L : 0;
}

hipError_t launch_cone_weights(double D, const float *pd, const float *xd, int np, int nx, double *pw, hipStream_t st)
{
    const long n = (long)np * nx, blocks = std::min((n + 255) / 256, 8192L);
    hipLaunchKernelGGL(k_cone_weights, dim3((unsigned)blocks), dim3(256), 0, st, D, pd, xd, np, nx, pw);
    return hipGetLastError();
}

hipError_t launch_row_filter(const FilterParams &P, const float *in, float *out, hipStream_t st)
{
    if (P.spec) {
        switch (P.fft_len) {
        case 256: return launch_fft<256>(P, in, out, st);
        case 512: return launch_fft<512>(P, in, out, st);
        case 1024: return launch_fft<1024>(P, in, out, st);
        case 2048: return launch_fft<2048>(P, in, out, st);
        default: return hipErrorInvalidValue;
        }
    }
    const int need = (P.np + 63) / 64;
    if (need <= 1) return launch_k<1, 4>(P, in, out, st);
    if (need <= 2) return launch_k<2, 4>(P, in, out, st);
    if (need <= 4) return launch_k<4, 4>(P, in, out, st);
    if (need <= 8) return launch_k<8, 4>(P, in, out, st);
    if (P.np <= 1024) return launch_k<16, 2>(P, in, out, st);
    return launch_k<16, 1>(P, in, out, st);
}

// ---------------------------------------------------------------------------
// Tables of the allocation-free ramp filter.  `scratch` is a file-scope
// device array: each GPU holds its own copy for the life of the process, and
// every call on that GPU overwrites it.  Host code never passes the array's
// name to a kernel (in host code that is a stand-in, not a device address):
// noalloc_scratch looks up the device address with hipGetSymbolAddress once
// per device, and the callers use that address.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
bool g_scratch_done[kMaxDevices] = {};
void *g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

// Half-scaled flat-detector taps times the pitch, h[k] d (the closed forms of
// the library's tap table: Shepp-Logan rounded through float as the
// reference's float taps, FDK_GPU/FDK.cu:6-19; Ram-Lak zero at even k).
__device__ double flat_tap(int filter, double d, int k)
{
    const double pi = 3.141592653589793;
    double v;
    if (filter == 0) {
        v = -1.0 / ((d * pi) * (d * pi)) / (4.0 * k * k - 1.0);
        v = (double)(float)v;
    } else if (k == 0) {
        v = 1.0 / (8.0 * d * d);
    } else if (k % 2 == 0) {
        v = 0.0;
    } else {
        v = -1.0 / (2.0 * pi * pi * (double)k * k * d * d);
    }
    return v * d;
}

// Block b writes taps[256 b ..] and, with L > 0, spec[256 b ..]: the one-sided
// taps and the cosine table sit in LDS, and each thread sums one spectral
// line H[k] = h[0] + 2 sum_m h[m] cos(2 pi k m / L) in FP64.
__global__ __launch_bounds__(256) void k_filter_tables(int filter, double d, int n, int L, double *__restrict__ taps,
                                                      double *__restrict__ spec)
{
    __shared__ double h[kNoAllocMaxCols / 2];
    __shared__ double cosl[2048];
    const int t = threadIdx.x, e = blockIdx.x * 256 + t;
    if (e < 2 * n - 1) taps[e] = flat_tap(filter, d, abs(e - (n - 1)));
    if (L == 0) return;   // direct convolution (uniform over the grid): no spectrum
    for (int m = t; m < n; m += 256) h[m] = flat_tap(filter, d, m);
    for (int i = t; i < L; i += 256) cosl[i] = cospi(2.0 * i / L);
    __syncthreads();
    if (e >= L) return;
    double acc = 0.0;
    for (int m = n - 1; m >= 1; --m) acc = fma(2.0 * h[m], cosl[(int)(((long long)e * m) & (L - 1))], acc);
    spec[e] = (h[0] + acc) / L;
}
}  // namespace

hipError_t noalloc_scratch(int device, void **ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_done[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_err[device] = e;
        g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
        g_scratch_done[device] = true;
    }
    *ptr = g_scratch_ptr[device];
    return g_scratch_err[device];
}

hipError_t launch_filter_tables(int filter, double d, int n, int L, double *taps, double *spec, hipStream_t st)
{
    if (n < 1 || n > kNoAllocMaxCols || L > 2048 || (L & (L - 1)) != 0) return hipErrorInvalidValue;
    const int blocks = std::max((2 * n - 1 + 255) / 256, L / 256);
    hipLaunchKernelGGL(k_filter_tables, dim3(blocks), dim3(256), 0, st, filter, d, n, L, taps, spec);
    return hipGetLastError();
}

}  // namespace ctr
