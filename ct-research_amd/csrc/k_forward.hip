// k_forward.hip - Siddon ray-driven forward projector for gfx950.
//
// One thread per ray (view, row, column); all views in one launch.  A block
// is 64 columns x 4 rows of one view, so a wavefront holds 64 neighbouring,
// nearly parallel rays that walk neighbouring voxels (shared cache lines).
//
// Per ray the entry/exit of the volume box, the first crossing of each axis
// and the crossing counts are computed in FP64 (near-parallel rays need it).
// The traversal runs in FP32: the next crossing of each axis is formed as
// first + count * step with one fma (no accumulated drift), and the axis
// whose plane comes first advances with selects rather than branches, so a
// wavefront never splits three ways.  The loop runs the FP64 setup's crossing
// count and the last segment ends at the exit parameter.  The volume is read
// through a range-checked buffer descriptor.  (The system-matrix walk keeps
// FP64 parameters: it is an export, not a hot loop.)
//
// Geometry: ProjectionCone_3D.cu:26-42 (source, end point at 2D along the
// ray through (p, xi) of the virtual isocentre detector); segment weight is
// the intersection length in mm, as the reference's :311-356.
#include "ctr_internal.h"

namespace ctr {
namespace {

constexpr int kBX = 64, kBY = 4;
// detector columns per wavefront: a block covers 64 columns x 4 rows, and a
// wavefront 16 columns of all 4 rows (wave w: columns 16 w ..).  At oblique
// views the rays of one row reach the same D plane at heights that spread
// with the column, so a wider row span touches more z slices per gather:
// 16 x 4 measured 18.7 ms at 256^3 / 360 views and 257 ms at 512^3 / 720,
// 32 x 2 18.8 / 259 ms, 64 x 1 19.1 / 266 ms
constexpr int kWaveCols = 16;
static_assert(kWaveCols != 16 || (kBX == 64 && kBY == 4), "16-column wavefronts of 4 rows tile a 64 x 4 block");
static_assert(kWaveCols != 32 || (kBX == 64 && kBY % 2 == 0), "32-column wavefronts of 2 rows need an even block height");

// Column and row of this thread's ray within its block.
__device__ __forceinline__ void ray_of_thread(int &dcol, int &drow)
{
    if (kWaveCols == 32) {
        const int wave = threadIdx.y, lane = threadIdx.x;   // blockDim.x = 64: one wavefront per y
        dcol = (wave & 1) * 32 + (lane & 31);
        drow = (wave >> 1) * 2 + (lane >> 5);
    } else if (kWaveCols == 16) {
        const int wave = threadIdx.y, lane = threadIdx.x;
        dcol = wave * 16 + (lane & 15);
        drow = lane >> 4;
    } else {
        dcol = threadIdx.x;
        drow = threadIdx.y;
    }
}

struct AxisSetup {
    int i0, n;        // start index, number of plane crossings
    double next0;     // first crossing, relative to the entry parameter
    double step;      // parameter step between crossings
    int stride;       // signed index stride
};

__device__ __forceinline__ AxisSetup setup_axis(double s, double d, double lam_in, double lam_out, double r, int nvox,
                                                int stride)
{
    AxisSetup a;
    const double xin = (s + lam_in * d) / r, xout = (s + lam_out * d) / r;
    if (d > 0.0) {
        int i0 = (int)floor(xin), i1 = (int)ceil(xout) - 1;
        i0 = min(max(i0, 0), nvox - 1);
        i1 = min(max(i1, i0), nvox - 1);
        a.i0 = i0;
        a.n = i1 - i0;
        a.next0 = ((i0 + 1) * r - s) / d - lam_in;
        a.step = r / d;
        a.stride = stride;
    } else if (d < 0.0) {
        int i0 = (int)ceil(xin) - 1, i1 = (int)floor(xout);
        i0 = min(max(i0, 0), nvox - 1);
        i1 = max(min(i1, i0), 0);
        a.i0 = i0;
        a.n = i0 - i1;
        a.next0 = (i0 * r - s) / d - lam_in;
        a.step = -r / d;
        a.stride = -stride;
    } else {
        a.i0 = min(max((int)floor(xin), 0), nvox - 1);
        a.n = 0;
        a.next0 = __builtin_huge_val();
        a.step = 0.0;
        a.stride = 0;
    }
    return a;
}

constexpr int kBufDword3 = 0x00020000;   // raw 32-bit buffer descriptor word 3 on gfx9

// Slab clip of the ray s + lam*d against [0, b]; returns false on a miss.
__device__ __forceinline__ bool clip_axis(double s, double d, double b, double &lo, double &hi)
{
    if (d == 0.0) return s >= 0.0 && s < b;
    const double inv = 1.0 / d;
    const double t1 = -s * inv, t2 = (b - s) * inv;
    lo = fmax(lo, fmin(t1, t2));
    hi = fmin(hi, fmax(t1, t2));
    return true;
}

// Walk of one ray through the volume: per-axis crossings after the entry
// point, the exit parameter and the first voxel.
struct RayWalk {
    AxisSetup at, as, az;
    double lam_end;   // chord length through the volume box
    int idx0;         // first voxel (in the layout the strides describe)
};

// Set-up of ray (view ib, column ip, row ix); false on a miss.  With tr the
// t and s strides are those of a [z][t][s] copy of the volume.
__device__ __forceinline__ bool ray_walk(const RayParams &g, int ib, int ip, int ix, bool tr, RayWalk &w)
{
    const double sb = g.view[4 * ib + 0], cb = g.view[4 * ib + 1];
    const double P = g.p[ip], X = g.xi[ix];
    double st, ss, dx, dy, dz;
    if (g.parallel) {
        // foot point of the line and its unit direction (ProjectionParallel_2D.cu:22-28)
        st = 0.5 * g.bt + P * cb;
        ss = 0.5 * g.bs + P * sb;
        dx = -sb;
        dy = cb;
        dz = 0.0;
    } else {
        st = g.view[4 * ib + 2];
        ss = g.view[4 * ib + 3];
        dx = P * cb + g.D * sb;
        dy = P * sb - g.D * cb;
        dz = X;
        const double inv_n = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
        dx *= inv_n;
        dy *= inv_n;
        dz *= inv_n;
    }
    double lo = -__builtin_huge_val(), hi = __builtin_huge_val();
    bool hit = clip_axis(st, dx, g.bt, lo, hi);
    hit = clip_axis(ss, dy, g.bs, lo, hi) && hit;
    hit = clip_axis(g.sz, dz, g.bz, lo, hi) && hit;
    if (!hit || !(hi > lo)) return false;
    const int st_t = tr ? g.ns : 1, st_s = tr ? 1 : g.nt;
    w.at = setup_axis(st, dx, lo, hi, g.rt, g.nt, st_t);
    w.as = setup_axis(ss, dy, lo, hi, g.rs, g.ns, st_s);
    w.az = setup_axis(g.sz, dz, lo, hi, g.rz, g.nz, g.nt * g.ns);
    w.lam_end = hi - lo;
    w.idx0 = w.az.i0 * (g.nt * g.ns) + w.as.i0 * st_s + w.at.i0 * st_t;
    return true;
}

// Minimum / maximum of an int over the wavefront; every lane must be active.
__device__ __forceinline__ int wave_min_int(int v)
{
    for (int m = 32; m > 0; m >>= 1) v = min(v, __shfl_xor(v, m, 64));
    return v;
}
__device__ __forceinline__ int wave_max_int(int v)
{
    for (int m = 32; m > 0; m >>= 1) v = max(v, __shfl_xor(v, m, 64));
    return v;
}

// min(a, b) for finite or infinite operands as one v_med3 (no NaN quieting).
// ninf is -inf held in a register the compiler cannot see through
// (opaque_ninf): with a literal -inf it rewrites the med3 as an IEEE min and
// re-canonicalises every operand defined outside the loop on every slab.
__device__ __forceinline__ float min_nc(float a, float b, float ninf)
{
    return __builtin_amdgcn_fmed3f(a, b, ninf);
}
__device__ __forceinline__ float opaque_ninf()
{
    float v;
    asm volatile("v_mov_b32 %0, 0xff800000" : "=v"(v));
    return v;
}

__device__ __forceinline__ float load_vox(__amdgpu_buffer_rsrc_t rs, unsigned off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
}

// One crossing at a time: the axis whose plane comes first advances (t
// before s before z on ties).  The c-th crossing of an axis is first +
// c * step, one rounding each (an accumulated next += step drifts by the
// square root of the crossing count in ulps).  Parameters are relative to
// the entry point, so they stay below the chord length.
__device__ __forceinline__ float walk_crossings(__amdgpu_buffer_rsrc_t rs, const RayWalk &w)
{
    const float inf = __builtin_huge_valf();
    const float t0 = w.at.n > 0 ? (float)w.at.next0 : inf, dt = (float)w.at.step;
    const float s0 = w.as.n > 0 ? (float)w.as.next0 : inf, ds = (float)w.as.step;
    const float z0 = w.az.n > 0 ? (float)w.az.next0 : inf, dz = (float)w.az.step;
    // byte strides; negative ones wrap, as unsigned offsets do
    const unsigned bt_st = (unsigned)w.at.stride * 4u, bs_st = (unsigned)w.as.stride * 4u,
                   bz_st = (unsigned)w.az.stride * 4u;
    float nt = t0, ns = s0, nz = z0, ct = 0.0f, cs = 0.0f, cz = 0.0f;   // crossings passed per axis
    unsigned off = (unsigned)w.idx0 * 4u;
    float lam = 0.0f, acc = 0.0f;
    // every counted crossing ends a segment; the last segment ends at the exit
    const int crossings = w.at.n + w.as.n + w.az.n;
    for (int k = 0; k < crossings; ++k) {
        const bool bt = nt <= ns && nt <= nz;
        const bool bs = !bt && ns <= nz;
        const float m = bt ? nt : (bs ? ns : nz);
        acc = fmaf(m - lam, load_vox(rs, off), acc);
        lam = m;
        off += bt ? bt_st : (bs ? bs_st : bz_st);
        ct = bt ? ct + 1.0f : ct;
        cs = bs ? cs + 1.0f : cs;
        cz = (!bt && !bs) ? cz + 1.0f : cz;
        nt = fmaf(ct, dt, t0);
        ns = fmaf(cs, ds, s0);
        nz = fmaf(cz, dz, z0);
    }
    return fmaf((float)w.lam_end - lam, load_vox(rs, off), acc);
}

// Line integral of one ray; also returns the chord length of the ray
// through the volume box (the row sum A*1 of the Siddon system matrix).
// Every lane of the wavefront must call it (valid false for lanes without a
// ray): the wavefront decides together how it walks.
//
// A wavefront holds 64 neighbouring rays of one detector row.  When, for all
// of them, the in-plane axis D that the view runs along has the shortest
// crossing step (by 0.1%, so float rounding cannot put two crossings of
// another axis between two D planes), the rays walk D's slabs in lockstep:
// at iteration u every ray is in the same absolute slab, so the wavefront's
// loads fall on a few neighbouring cache lines (one row of the layout whose
// fastest axis is the other in-plane axis).  In a slab a ray crosses each
// other axis at most once, so the slab is at most three segments: before
// the first crossing, between the two, after the second.  The segments and
// voxels are those of the one-crossing-at-a-time walk, which the other
// wavefronts run.  No crossing within rounding of the exit is taken that the
// FP64 set-up did not count, so no weighted load falls outside the volume
// (where the two layouts differ).
template <bool kSkip>
__device__ __forceinline__ float trace_ray(const float *__restrict__ vol, const RayParams &g, int ib, int ip, int ix,
                                           bool valid, float &chord)
{
    // views whose central ray runs mostly along t read the [z][t][s] copy, so
    // that the neighbouring rays of a wavefront (neighbouring p, offset in s)
    // gather from neighbouring addresses
    const bool along_t = fabs(g.view[4 * ib + 0]) > fabs(g.view[4 * ib + 1]);
    const bool tr = g.volT != nullptr && along_t;
    RayWalk w;
    const bool hit = valid && ray_walk(g, ib, ip, ix, tr, w);
    chord = hit ? (float)w.lam_end : 0.0f;

    // num_records is 32 bits: 2^30 voxels (4 GiB) saturate to 2^32 - 1, above every valid offset
    const size_t vbytes = (size_t)g.nt * g.ns * g.nz * sizeof(float);
    const unsigned nbytes = vbytes > 0xFFFFFFFFull ? 0xFFFFFFFFu : (unsigned)vbytes;
    // base and size are the same for the whole block (one view of one
    // problem); stating it keeps the descriptor in scalar registers where the
    // compiler cannot prove it (the batched kernels read them from a job
    // table).  readfirstlane returns int: each half goes through unsigned, so
    // the low half is zero-extended, not sign-extended into the high one.
    const unsigned long long base = (unsigned long long)(tr ? g.volT : vol);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)(base >> 32));
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)base);
    const unsigned long long ubase = ((unsigned long long)hi << 32) | (unsigned long long)lo;
    __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)ubase, 0, (int)__builtin_amdgcn_readfirstlane(nbytes), kBufDword3);

    // lockstep slabs: D is t or s by the view, A the other in-plane axis, B = z
    const AxisSetup aD = along_t ? w.at : w.as;
    const AxisSetup aA = along_t ? w.as : w.at;
    const AxisSetup &aB = w.az;
    bool ok = true;
    if (hit) {
        const double lim = 1.001 * aD.step;
        ok = aD.stride != 0 && (aA.n == 0 || aA.step >= lim) && (aB.n == 0 || aB.step >= lim);
    }
    const unsigned long long bad = __ballot(hit && !ok);
    const unsigned long long up = __ballot(hit && aD.stride > 0), down = __ballot(hit && aD.stride < 0);
    if (bad != 0 || (up != 0 && down != 0)) return hit ? walk_crossings(rs, w) : 0.0f;

    // wavefront-relative slab index u: absolute D index istart + dir * u
    const int dir = up != 0 ? 1 : -1;
    const int first = hit ? aD.i0 : (dir > 0 ? 0x7fffffff : -0x7fffffff);
    const int istart = dir > 0 ? wave_min_int(first) : wave_max_int(first);
    // this ray's first slab, >= 0; a lane without a hit starts 2^30 slabs
    // late, so its k = u - o never enters [0, nD]
    const int o = hit ? dir * (aD.i0 - istart) : (1 << 30);
    const int uend = __builtin_amdgcn_readfirstlane(wave_max_int(hit ? o + aD.n : -1));   // uniform
    const int nD = hit ? aD.n : 0;

    const float inf = __builtin_huge_valf();
    const float D0 = hit ? (float)aD.next0 : inf, dD = hit ? (float)aD.step : 0.0f;
    const float A0 = hit && aA.n > 0 ? (float)aA.next0 : inf, dA = hit ? (float)aA.step : 0.0f;
    const float B0 = hit && aB.n > 0 ? (float)aB.next0 : inf, dB = hit ? (float)aB.step : 0.0f;
    const float lend = hit ? (float)w.lam_end : 0.0f;
    // 8 ulps below the exit: the float crossings of the FP64-counted planes
    // are within about 2 ulps of their exact values
    const float lcnt = __builtin_bit_cast(float, max(__builtin_bit_cast(int, lend) - 8, 0));
    const unsigned sD = (unsigned)aD.stride * 4u, sA = (unsigned)aA.stride * 4u, sB = (unsigned)aB.stride * 4u;
    unsigned off = hit ? (unsigned)w.idx0 * 4u : 0u;
    float pA = A0, pB = B0, cA = 0.0f, cB = 0.0f, kf = 0.0f, s = 0.0f, acc = 0.0f;
    const float ninf = opaque_ninf();
    for (int u = 0; u <= uend; ++u) {
        const int k = u - o;
        if ((unsigned)k > (unsigned)nD) continue;   // k outside [0, nD]
        // slab k ends at the k-th D crossing, the last one at the exit
        const float e = min_nc(fmaf(kf, dD, D0), lend, ninf);
        // crossings count up to ec: one past the last counted crossing of an
        // axis is at or beyond the exit, within rounding of it
        const float ec = min_nc(e, lcnt, ninf);
        const bool hasA = pA < ec, hasB = pB < ec;
        // p1 <= p2: the crossings inside the slab in order, ec where there is none
        const float p1 = min_nc(min_nc(pA, pB, ninf), ec, ninf);
        const float p2 = min_nc(__builtin_amdgcn_fmed3f(pA, pB, ec), ec, ninf);
        const unsigned offm = off + (pA <= pB ? sA : sB);
        const unsigned off3 = off + (hasA ? sA : 0u) + (hasB ? sB : 0u);
        const float w2 = p2 - p1, w3 = e - p2;
        acc = fmaf(p1 - s, load_vox(rs, off), acc);
        if (kSkip) {
            // the second and third voxels are read only by lanes whose
            // segment in them is not empty (a crossing inside the slab); a
            // skipped term is 0 * v, so the sum is unchanged bit for bit for
            // finite volumes (an infinite v would make that term NaN)
            float v2 = 0.0f, v3 = 0.0f;
            if (w2 != 0.0f) v2 = load_vox(rs, offm);
            if (w3 != 0.0f) v3 = load_vox(rs, off3);
            acc = fmaf(w2, v2, acc);
            acc = fmaf(w3, v3, acc);
        } else {
            acc = fmaf(w2, load_vox(rs, offm), acc);
            acc = fmaf(w3, load_vox(rs, off3), acc);
        }
        cA = hasA ? cA + 1.0f : cA;
        cB = hasB ? cB + 1.0f : cB;
        pA = fmaf(cA, dA, A0);
        pB = fmaf(cB, dB, B0);
        off = off3 + sD;
        s = e;
        kf += 1.0f;
    }
    return hit ? acc : 0.0f;
}

// Siddon system-matrix row of one ray: the voxels the walk above visits with
// a segment of positive length, in visiting order (the rows GenSysMatFan.m /
// GenSysMatCircular3D.m assemble; A x is exactly trace_ray).  kFill false
// only counts them; kFill true writes at most `limit` entries (the count).
template <bool kFill>
__device__ __forceinline__ int sysmat_row(const RayParams &g, int ib, int ip, int ix, int *__restrict__ col,
                                          float *__restrict__ val, long long limit)
{
    RayWalk w;
    if (!ray_walk(g, ib, ip, ix, false, w)) return 0;
    int idx = w.idx0, cnt = 0;
    const double inf = __builtin_huge_val();
    double nt = w.at.n > 0 ? w.at.next0 : inf;
    double ns = w.as.n > 0 ? w.as.next0 : inf;
    double nz = w.az.n > 0 ? w.az.next0 : inf;
    double lam = 0.0;
    const int total = w.at.n + w.as.n + w.az.n + 1;
    for (int k = 0; k < total; ++k) {
        const double m = fmin(fmin(nt, ns), fmin(nz, w.lam_end));
        if (m > lam) {
            if (kFill && cnt < limit) {
                col[cnt] = idx;
                val[cnt] = (float)(m - lam);
            }
            ++cnt;
        }
        lam = m;
        const bool bt = nt <= ns && nt <= nz;
        const bool bs = !bt && ns <= nz;
        idx += bt ? w.at.stride : (bs ? w.as.stride : w.az.stride);
        nt = bt ? nt + w.at.step : nt;
        ns = bs ? ns + w.as.step : ns;
        nz = (!bt && !bs) ? nz + w.az.step : nz;
    }
    return cnt;
}

__global__ __launch_bounds__(kBX *kBY) void k_sysmat_count(long long *__restrict__ counts, RayParams g)
{
    const int ip = blockIdx.x * kBX + threadIdx.x;
    const int ix = blockIdx.y * kBY + threadIdx.y;
    const int ib = blockIdx.z;
    if (ip >= g.np || ix >= g.nx) return;
    counts[((size_t)ib * g.nx + ix) * g.np + ip] = sysmat_row<false>(g, ib, ip, ix, nullptr, nullptr, 0);
}

__global__ __launch_bounds__(kBX *kBY) void k_sysmat_fill(const long long *__restrict__ row_ptr, int *__restrict__ col,
                                                         float *__restrict__ val, RayParams g)
{
    const int ip = blockIdx.x * kBX + threadIdx.x;
    const int ix = blockIdx.y * kBY + threadIdx.y;
    const int ib = blockIdx.z;
    if (ip >= g.np || ix >= g.nx) return;
    const size_t ray = ((size_t)ib * g.nx + ix) * g.np + ip;
    const long long o = row_ptr[ray];
    sysmat_row<true>(g, ib, ip, ix, col + o, val + o, row_ptr[ray + 1] - o);
}

__global__ __launch_bounds__(kBX *kBY) void k_ray_fp(const float *__restrict__ vol, float *__restrict__ proj, RayParams g)
{
    int dcol, drow;
    ray_of_thread(dcol, drow);
    const int ip = blockIdx.x * kBX + dcol;
    const int ix = blockIdx.y * kBY + drow;
    const int ib = blockIdx.z;
    const bool valid = ip < g.np && ix < g.nx;
    float chord;
    const float acc = g.skip_empty ? trace_ray<true>(vol, g, ib, ip, ix, valid, chord)
                                  : trace_ray<false>(vol, g, ib, ip, ix, valid, chord);
    if (valid) proj[((size_t)ib * g.nx + ix) * g.np + ip] = acc;
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// SIRT residual pass: r = (b - A x) / (A 1 + 1e-9) and per-block partial sums
// of (b - A x)^2 and |b - A x| (SIRT3D.m:80-81, :87) for block (bx, by) of
// view ib; the block's pair goes to partials[2 blk].  For the cone beam r is
// stored times the ray norm, the form in which the exact adjoint reads it.
__device__ __forceinline__ void fp_residual_block(const float *__restrict__ vol, const float *__restrict__ b,
                                                  float *__restrict__ r, double *__restrict__ partials,
                                                  const RayParams &g, int bx, int by, int ib, size_t blk)
{
    __shared__ double red[2][kBY];
    int dcol, drow;
    ray_of_thread(dcol, drow);
    const int ip = bx * kBX + dcol;
    const int ix = by * kBY + drow;
    double e2 = 0.0, e1 = 0.0;
    const bool valid = ip < g.np && ix < g.nx;
    float chord;
    const float acc = g.skip_empty ? trace_ray<true>(vol, g, ib, ip, ix, valid, chord)
                                  : trace_ray<false>(vol, g, ib, ip, ix, valid, chord);
    if (valid) {
        const size_t ray = ((size_t)ib * g.nx + ix) * g.np + ip;
        const float e = b[ray] - acc;
        const float rr = e / (chord + 1e-9f);
        r[ray] = g.residual_norm ? rr * ray_norm((float)g.p[ip], (float)g.xi[ix], (float)g.D) : rr;
        e2 = (double)e * (double)e;
        e1 = fabs((double)e);
    }
    e2 = wave_sum(e2);
    e1 = wave_sum(e1);
    if (threadIdx.x == 0) {
        red[0][threadIdx.y] = e2;
        red[1][threadIdx.y] = e1;
    }
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        double s2 = 0.0, s1 = 0.0;
        for (int w = 0; w < kBY; ++w) {
            s2 += red[0][w];
            s1 += red[1][w];
        }
        partials[2 * blk + 0] = s2;
        partials[2 * blk + 1] = s1;
    }
}

__global__ __launch_bounds__(kBX *kBY) void k_ray_fp_residual(const float *__restrict__ vol, const float *__restrict__ b,
                                                              float *__restrict__ r, double *__restrict__ partials,
                                                              RayParams g)
{
    const size_t blk = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    fp_residual_block(vol, b, r, partials, g, blockIdx.x, blockIdx.y, blockIdx.z, blk);
}

// Independent problems in one launch: a flat grid over the jobs' block ranges.
__global__ __launch_bounds__(kBX *kBY) void k_ray_fp_residual_batched(const FpJob *__restrict__ jobs, int njobs)
{
    // job blockIdx.y, its block blockIdx.x; blocks past the job's grid leave at once
    const int p = blockIdx.y, local = blockIdx.x;
    const int gx = jobs[p].gx, gy = jobs[p].gy;
    if (local >= gx * gy * jobs[p].g.nb) return;
    const RayParams g = jobs[p].g;
    const int bx = local % gx, by = (local / gx) % gy, ib = local / (gx * gy);
    fp_residual_block(jobs[p].vol, jobs[p].b, jobs[p].r, jobs[p].partials, g, bx, by, ib, (size_t)local);
}

dim3 fp_grid(const RayParams &g) { return dim3((g.np + kBX - 1) / kBX, (g.nx + kBY - 1) / kBY, g.nb); }

}  // namespace

int ray_fp_blocks(const RayParams &g)
{
    const dim3 grid = fp_grid(g);
    return (int)(grid.x * grid.y * grid.z);
}

hipError_t launch_sysmat_count(const RayParams &g, long long *counts, hipStream_t st)
{
    hipLaunchKernelGGL(k_sysmat_count, dim3((g.np + kBX - 1) / kBX, (g.nx + kBY - 1) / kBY, g.nb), dim3(kBX, kBY), 0,
                       st, counts, g);
    return hipGetLastError();
}

hipError_t launch_sysmat_fill(const RayParams &g, const long long *row_ptr, int *col, float *val, hipStream_t st)
{
    hipLaunchKernelGGL(k_sysmat_fill, dim3((g.np + kBX - 1) / kBX, (g.nx + kBY - 1) / kBY, g.nb), dim3(kBX, kBY), 0,
                       st, row_ptr, col, val, g);
    return hipGetLastError();
}

hipError_t launch_ray_fp(const RayParams &g, const float *vol, float *proj, hipStream_t st)
{
    hipLaunchKernelGGL(k_ray_fp, fp_grid(g), dim3(kBX, kBY), 0, st, vol, proj, g);
    return hipGetLastError();
}

hipError_t launch_ray_fp_residual(const RayParams &g, const float *vol, const float *b, float *r, double *partials,
                                  int *nblocks, hipStream_t st)
{
    const dim3 grid = fp_grid(g);
    *nblocks = (int)(grid.x * grid.y * grid.z);
    hipLaunchKernelGGL(k_ray_fp_residual, grid, dim3(kBX, kBY), 0, st, vol, b, r, partials, g);
    return hipGetLastError();
}

int fp_job_blocks(const RayParams &g, int *gx, int *gy)
{
    const dim3 grid = fp_grid(g);
    *gx = (int)grid.x;
    *gy = (int)grid.y;
    return (int)(grid.x * grid.y * grid.z);
}

hipError_t launch_ray_fp_residual_batched(const FpJob *jobs, int njobs, int max_blocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_ray_fp_residual_batched, dim3((unsigned)max_blocks, (unsigned)njobs), dim3(kBX, kBY), 0, st,
                       jobs, njobs);
    return hipGetLastError();
}

}  // namespace ctr
