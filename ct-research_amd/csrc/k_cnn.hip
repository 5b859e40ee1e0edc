// k_cnn.hip - 3-D CNN of the learned SIRT update on gfx950.
//
// Network: conv3d 2->32, ReLU, 32->32, ReLU, 32->32, ReLU, 32->1; 3x3x3
// kernels, zero padding 1.  Activations are BF16, channels-last
// ([z][y][x][c], c fastest), weights BF16, accumulation FP32.
//
// Convolution = implicit GEMM, D[co][voxel] += W[co][ci] * act[ci][voxel].
// 16 input channels (the first layer, and the input gradient of the last):
// k_conv3d on v_mfma_f32_32x32x16_bf16 (M = 32 output channels, N = 32
// voxels along x, K = 16 input channels of one tap).  32 input channels:
// k_conv3d_c32 on v_mfma_f32_16x16x32_bf16 (M = 16 output channels in two
// halves, N = 16 voxels, K = all 32 channels).  A (weights) is staged in
// LDS once per block; B (activations) is read straight from global memory
// as 16-byte channel vectors.  A wave owns R rows of voxels along y and
// loads the R + 2 rows around them once per step, reusing them for the
// 3 dy taps.
//
// The backward pass w.r.t. the input is the same convolution with the taps
// flipped and the channel roles swapped; its epilogue applies the ReLU
// derivative of the layer below.  The weight gradient is a VALU reduction
// over LDS-staged tiles (per-block partial sums, reduced deterministically).
#include "ctr_cnn.h"

namespace ctr {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTaps = 27;
constexpr int kConvZRun = 16;   // consecutive output slices per x-y tile of a conv block
// output rows (y) per wave of k_conv3d: 4 for 16 input channels (two waves
// per SIMD)
constexpr int conv_rows(int cin) { return cin == 32 ? 8 : 4; }

__device__ __forceinline__ float bf2f(__bf16 v) { return (float)v; }

// conv3d, CIN = 16 or 32 input channels (zero padded), 32 output channels
// (zero padded).  Epilogues: see CnnEpi.
template <int CIN, int EPI>
__global__ __launch_bounds__(256) void k_conv3d(const __bf16 *__restrict__ in, const __bf16 *__restrict__ wts,
                                                const float *__restrict__ bias, const __bf16 *__restrict__ mask,
                                                void *__restrict__ out, CnnShape S)
{
    constexpr int NQ = CIN / 16;
    constexpr int NW = kTaps * NQ * 32 * 2;   // bf16x8 vectors of the A layout [tap][q][co][16]
    __shared__ bf16x8 wl[NW];
    for (int i = threadIdx.x; i < NW; i += 256) wl[i] = reinterpret_cast<const bf16x8 *>(wts)[i];
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 31, kh = lane >> 5;
    constexpr int R = conv_rows(CIN);   // output rows per wave
    const int tiles_x = (S.nx + 31) / 32, tiles_y = (S.ny + 4 * R - 1) / (4 * R);
    const int zchunks = (S.nz + kConvZRun - 1) / kConvZRun;
    const int ncols = tiles_x * tiles_y * zchunks;
    // a block walks one x-y tile through kConvZRun consecutive slices, so
    // the three input slices of an output slice were read by this CU for the
    // previous one (its L2 serves the dz reuse)
    for (int tcol = blockIdx.x; tcol < ncols; tcol += gridDim.x)
    for (int z = (tcol / (tiles_x * tiles_y)) * kConvZRun,
             zend = min(S.nz, z + kConvZRun); z < zend; ++z) {
        const int tx = tcol % tiles_x, ty = (tcol / tiles_x) % tiles_y;
        const int x = tx * 32 + col, y0 = ty * 4 * R + wave * R;
        f32x16 acc[R];
#pragma unroll
        for (int g = 0; g < R; ++g)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[g][r] = 0.0f;

        // steps (dz, dx, channel chunk); the ten activation rows of the next
        // step are loaded while the MFMAs of the current one run (3 R per
        // step; with 32 input channels 8 rows per wave at one wave per SIMD
        // measured 1.37 ms per layer at 256^3, 4 rows at two waves 1.54 ms)
        constexpr int NS = 9 * NQ;
        // the rows come through a buffer descriptor over the (at most three)
        // input slices of this output slice; a row outside the volume gets
        // an offset past the descriptor's range, which reads as zero.  With
        // no branch and no select on the loaded values, the compiler waits
        // only for the rows the next MFMAs read.
        const int zlo = max(z - 1, 0), zhi = min(z + 1, S.nz - 1);
        const size_t plane = (size_t)S.ny * S.nx * CIN;   // bf16 elements per slice
        const unsigned long long base = (unsigned long long)(in + zlo * plane);
        const unsigned bhi = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)(base >> 32));
        const unsigned blo = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)base);
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(((unsigned long long)bhi << 32) | blo), 0,
            __builtin_amdgcn_readfirstlane((int)((zhi - zlo + 1) * plane * sizeof(__bf16))), 0x00020000);
        auto load_rows = [&](int st, bf16x8 *rows) {
            const int q = st % NQ, dxi = (st / NQ) % 3, dzi = st / (3 * NQ);
            const int zz = z + dzi - 1, xx = x + dxi - 1;
            const bool zxok = zz >= 0 && zz < S.nz && xx >= 0 && xx < S.nx;
#pragma unroll
            for (int rr = 0; rr < R + 2; ++rr) {
                const int yy = y0 - 1 + rr;
                const bool ok = zxok && yy >= 0 && yy < S.ny;
                const unsigned off =
                    ok ? (unsigned)(((((size_t)(zz - zlo) * S.ny + yy) * S.nx + xx) * CIN + q * 16 + kh * 8) *
                                    sizeof(__bf16))
                       : 0x80000000u;
                rows[rr] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off, 0, 0));
            }
        };
        auto mma = [&](int st, const bf16x8 *rows) {
            const int q = st % NQ, dxi = (st / NQ) % 3, dzi = st / (3 * NQ);
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
                const int t = dzi * 9 + dy * 3 + dxi;
                const bf16x8 a = wl[((t * NQ + q) * 32 + col) * 2 + kh];
#pragma unroll
                for (int g = 0; g < R; ++g)
                    acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, rows[g + dy], acc[g], 0, 0, 0);
            }
        };
        bf16x8 ra[R + 2], rb[R + 2];
        load_rows(0, ra);
#pragma unroll 1
        for (int st = 0; st < NS; st += 2) {
            load_rows(min(st + 1, NS - 1), rb);
            mma(st, ra);
            load_rows(min(st + 2, NS - 1), ra);
            if (st + 1 < NS) mma(st + 1, rb);
        }

        // lane holds D[co][voxel col] for co = 8 s + 4 kh + u (s, u = 0..3)
#pragma unroll
        for (int g = 0; g < R; ++g) {
            const int y = y0 + g;
            if (x >= S.nx || y >= S.ny) continue;
            const size_t vox = ((size_t)z * S.ny + y) * S.nx + x;
            if (EPI == kEpiLinearF32) {
                if (kh == 0) reinterpret_cast<float *>(out)[vox] = acc[g][0] + bias[0];
                continue;
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int c0 = 8 * s + 4 * kh;
                bf16x4 o;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    float v = acc[g][4 * s + u];
                    if (EPI == kEpiBiasRelu) {
                        v = fmaxf(v + bias[c0 + u], 0.0f);
                    } else if (EPI == kEpiMask) {
                        v = bf2f(mask[vox * 32 + c0 + u]) > 0.0f ? v : 0.0f;
                    }
                    o[u] = (__bf16)v;
                }
                *reinterpret_cast<bf16x4 *>(reinterpret_cast<__bf16 *>(out) + vox * 32 + c0) = o;
            }
        }
    }
}

// conv3d with 32 input channels (k_conv3d_c32 below) on
// v_mfma_f32_16x16x32_bf16: M = 16 output channels (two halves), N = 16
// voxels along x, K = all 32 input channels of one tap.  Lane l takes
// channels 8 (l / 16) .. + 8 of voxel x0 + l mod 16, so one load
// instruction of a wave reads 16 whole voxels (1 KB contiguous) for both
// channel halves.  Each input row is loaded once at the centre column
// x0 + l mod 16; the dx = -1 and dx = +1 operands are that row shifted
// across the lanes of each channel group, with the halo voxels (x0 - 1 and
// x0 + 16) of all rows from one more load; each row feeds the 3 dy taps and
// both output halves.  Against three loads per (dz, row), one for each dx:
// 1.13 -> 0.97 ms per layer at 256^3 (vector-memory unit busy 0.79 -> 0.57),
// and against the 32 x 32 x 16 layout above (half voxels, two channel
// steps): 1.37 ms.  The A operands come from the global layout
// [tap][q][32 co][16 ci], staged in LDS as [tap][half][lane] so a wave's
// read is one conflict-free 1 KB run.

// x-neighbour operands from a centre row: a 16-lane DPP row holds the 16
// voxels x0 .. x0 + 15 of one channel group, so a row shift by one lane
// gives voxel x0 + col - 1 (row_shr:1) or x0 + col + 1 (row_shl:1).  The lane
// whose source lies outside the row (col 0, col 15) takes the halo voxel
// instead.  The halo voxels of all R + 2 rows of a slice arrive in one
// register, per channel group: lane 1 + r holds x0 - 1 of row r, lane
// 15 - (R + 2) + r holds x0 + 16 of row r.  A first DPP move brings row r's
// halo voxel to col 0 (row_shl:1 + r) or col 15 (row_shr:R + 2 - r); the row
// shift then leaves that lane as it is (its source is out of the row).
template <int SHIFT, int PICK>
__device__ __forceinline__ bf16x8 shift_cols(bf16x8 halo, bf16x8 centre)
{
    typedef int i32x4 __attribute__((ext_vector_type(4)));
    const i32x4 e = __builtin_bit_cast(i32x4, halo), c = __builtin_bit_cast(i32x4, centre);
    i32x4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        r[i] = __builtin_amdgcn_update_dpp(__builtin_amdgcn_update_dpp(0, e[i], PICK, 0xf, 0xf, true), c[i], SHIFT,
                                           0xf, 0xf, false);
    return __builtin_bit_cast(bf16x8, r);
}
constexpr int kDppRowShl = 0x100, kDppRowShr = 0x110;   // + shift, 1 .. 15
// rows ROW .. NR - 1: left[] = the dx = -1 operands, right[] = the dx = +1 operands
template <int ROW, int NR>
__device__ __forceinline__ void shift_rows(bf16x8 halo, const bf16x8 *centre, bf16x8 *left, bf16x8 *right)
{
    static_assert(2 * NR < 16, "halo lanes of the left and right edges overlap");
    if constexpr (ROW < NR) {
        left[ROW] = shift_cols<kDppRowShr + 1, kDppRowShl + 1 + ROW>(halo, centre[ROW]);
        right[ROW] = shift_cols<kDppRowShl + 1, kDppRowShr + NR - ROW>(halo, centre[ROW]);
        shift_rows<ROW + 1, NR>(halo, centre, left, right);
    }
}

// The 32-channel convolution walks input slices: each input slice zi is
// loaded once (centre and halo rows prefetched while the previous slice
// multiplies) and feeds the three output slices zi + 1, zi, zi - 1 through
// the tap planes dz = -1, 0, +1: three rolling accumulator sets per row, the
// lowest complete (and stored) once slice zi's dz = +1 plane is in.  A wave owns 16 x 4
// outputs; a block runs zrun output slices of its 16 x 16 tile.  Against
// loading the three input slices of every output slice (8 rows per wave):
// vector-memory unit busy 0.52 -> 0.30, 0.96 -> 0.93 ms per layer at 256^3.
// Forming the shifted rows once per slice instead of once per tap plane,
// and the halo of all rows in one load and one register instead of one per
// row: 0.94 -> 0.86 ms on the same GPU.  Running the tap planes in order
// dz = +1, 0, -1 (store between the first and second), so that the rolling
// accumulator sets need no copies from slice to slice: 0.87 -> 0.81 ms;
// the store after the first step of the second plane: 0.83 -> 0.81 ms.
constexpr int kC32zRows = 4;   // output rows per wave
// Output slices per block task (zrun, a launch argument): 64 where that
// still leaves at least kC32MinTasks tasks, else 32 or 16.  At 256^3, 64
// against 32: about 2% less time per layer (fewer halo slices per run).
constexpr int kC32MinTasks = 1024;
template <int EPI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_conv3d_c32(const __bf16 *__restrict__ in, const __bf16 *__restrict__ wts,
                                                    const float *__restrict__ bias, const __bf16 *__restrict__ mask,
                                                    void *__restrict__ out, CnnShape S, int zrun)
{
    constexpr int CIN = 32, R = kC32zRows;
    constexpr int NW = kTaps * 2 * 64;   // bf16x8 vectors, [tap][half][lane]
    __shared__ bf16x8 wl[NW];
    for (int i = threadIdx.x; i < NW; i += 256) {
        const int l = i & 63, h = (i >> 6) & 1, t = i >> 7;
        const int m = l & 15, co = 8 * (m >> 2) + 4 * h + (m & 3);
        wl[i] = reinterpret_cast<const bf16x8 *>(wts)[((t * 2 + (l >> 5)) * 32 + co) * 2 + ((l >> 4) & 1)];
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform
    const int col = lane & 15, kg = lane >> 4;
    constexpr int TY = 4 * R;
    const int tiles_x = (S.nx + 15) / 16, tiles_y = (S.ny + TY - 1) / TY;
    const int zchunks = (S.nz + zrun - 1) / zrun;
    const int ncols = tiles_x * tiles_y * zchunks;
    const size_t plane = (size_t)S.ny * S.nx * CIN;
    for (int tcol = blockIdx.x; tcol < ncols; tcol += gridDim.x) {
        const int tx = tcol % tiles_x, ty = (tcol / tiles_x) % tiles_y;
        const int z0 = (tcol / (tiles_x * tiles_y)) * zrun, z1 = min(S.nz, z0 + zrun);
        const int x0 = tx * 16, x = x0 + col, y0 = ty * TY + wave * R;
        // buffer descriptor over input slice zi (outside the volume: an empty
        // range, every load reads zero)
        auto slice_rsrc = [&](int zi) {
            const bool zok = zi >= 0 && zi < S.nz;
            const unsigned long long base = (unsigned long long)(in + (zok ? zi : 0) * plane);
            const unsigned bhi = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)(base >> 32));
            const unsigned blo = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)base);
            return __builtin_amdgcn_make_buffer_rsrc((void *)(((unsigned long long)bhi << 32) | blo), 0,
                                                     __builtin_amdgcn_readfirstlane(zok ? (int)(plane * sizeof(__bf16)) : 0),
                                                     0x00020000);
        };
        // centre rows y0 - 1 .. y0 + R of input slice zi at voxel x (outside
        // the volume: zero)
        auto load_slice = [&](int zi, bf16x8 *rows) {
            const __amdgpu_buffer_rsrc_t rsrc = slice_rsrc(zi);
            // lane part of the offset, recomputed per slice rather than kept
            // in registers across the loop (the empty asm stops the hoist); a
            // lane outside the volume starts at 2^31, past the descriptor's range
            unsigned xoff = x < S.nx ? (unsigned)((x * CIN + kg * 8) * sizeof(__bf16)) : 0x80000000u;
            asm volatile("" : "+v"(xoff));
#pragma unroll
            for (int rr = 0; rr < R + 2; ++rr) {
                const int yy = y0 - 1 + rr;   // uniform
                const unsigned off = yy >= 0 && yy < S.ny
                                         ? xoff + (unsigned)((size_t)yy * S.nx * CIN * sizeof(__bf16))
                                         : 0x80000000u;
                rows[rr] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off, 0, 0));
            }
        };
        // the 2 (R + 2) halo voxels of slice zi in one register (see shift_cols)
        auto load_halo = [&](int zi) {
            const int rr = col <= R + 2 ? col - 1 : col - (15 - (R + 2));
            const int xx = col <= R + 2 ? x0 - 1 : x0 + 16, yy = y0 - 1 + rr;
            const bool ok = (col >= 1 && col <= R + 2) || (col >= 15 - (R + 2) && col <= 14);
            const unsigned off = ok && xx >= 0 && xx < S.nx && yy >= 0 && yy < S.ny
                                     ? (unsigned)((((size_t)yy * S.nx + xx) * CIN + kg * 8) * sizeof(__bf16))
                                     : 0x80000000u;
            return __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(slice_rsrc(zi), off, 0, 0));
        };
        // starting value of an output slice's accumulators: the bias of this
        // lane's channels 8 kg + 4 h + u (bias and ReLU epilogue), else zero
        f32x4 acc0[2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < 4; ++u) acc0[h][u] = EPI == kEpiBiasRelu ? bias[8 * kg + 4 * h + u] : 0.0f;
        f32x4 acc[3][R][2];   // [0]: output zi - 1, [1]: zi, [2]: zi + 1
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int g = 0; g < R; ++g)
#pragma unroll
                for (int h = 0; h < 2; ++h) acc[q][g][h] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        // taps (dzi, dxi, dy) of every row into the output slice acc[2 - dzi];
        // `rows` holds the operands of column offset dxi - 1.  With `init`
        // (the first step into a fresh output slice) the first product of
        // each accumulator starts from acc0 (the bias, or zero).
        auto mma_dx = [&](int dzi, int dxi, const bf16x8 *rows, bool init) {
            bf16x8 a[3][2];
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int h = 0; h < 2; ++h) a[dy][h] = wl[((dzi * 9 + dy * 3 + dxi) * 2 + h) * 64 + lane];
#pragma unroll
            for (int r = 0; r < R + 2; ++r) {
#pragma unroll
                for (int dy = 0; dy < 3; ++dy) {
                    const int g = r - dy;
                    if (g < 0 || g >= R) continue;
#pragma unroll
                    for (int h = 0; h < 2; ++h)
                        acc[2 - dzi][g][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            a[dy][h], rows[r], init && dy == 0 ? acc0[h] : acc[2 - dzi][g][h],
                            0, 0, 0);
                }
            }
        };
        // The centre and halo rows of slice zi + 1 load while slice zi
        // multiplies.  The shifted rows (dx = -1, +1) are formed once per
        // slice and feed all three tap planes.  The planes run in the order
        // dz = +1 (completing output slice zi - 1, stored right after), 0,
        // -1 (starting output slice zi + 1): the registers of the stored
        // slice then carry on as output slice zi, and those of zi as zi + 1,
        // with no copies between slices.
        bf16x8 cur[R + 2], nxt[R + 2], sl[R + 2], sh[R + 2];
        // output: slices of oplane elements, rows of ostride bytes
        const size_t oplane = (size_t)S.ny * S.nx * 32;
        const unsigned ostride = (unsigned)S.nx * 32 * sizeof(__bf16);
        const int xoff_out = x < S.nx ? (int)((x * 32 + 8 * kg) * sizeof(__bf16)) : (int)0x80000000u;
        load_slice(z0 - 1, cur);
        bf16x8 hal = load_halo(z0 - 1), nhal;
        // step (dzi, dxi) of slice zi: output slice zi + 1 - dzi, skipped
        // (uniform) where that slice lies outside this run (its accumulators
        // are then never stored).  The empty asm statements keep each step's
        // weight reads (and their registers) from being hoisted over the
        // others.
        auto step = [&](int zi, int dzi, int dxi) {
            if (zi + 1 - dzi < z0 || zi + 1 - dzi >= z1) return;
            mma_dx(dzi, dxi, dxi == 1 ? cur : dxi == 0 ? sl : sh, dzi == 0 && dxi == 1);
            asm volatile("" ::: "memory");
        };
#pragma unroll 1
        for (int zi = z0 - 1; zi <= z1; ++zi) {
            const bool more = zi + 1 <= z1;
            if (more) {
                load_slice(zi + 1, nxt);
                nhal = load_halo(zi + 1);
            }
            const bool zin = zi >= 0 && zi < S.nz;   // uniform: a slice outside the volume contributes zero
            if (zin) {
                shift_rows<0, R + 2>(hal, cur, sl, sh);
                step(zi, 2, 1);
                step(zi, 2, 0);
                step(zi, 2, 2);
                // one step of the next plane before the store, whose
                // conversions would otherwise wait on the MFMAs just issued
                step(zi, 1, 1);
            }
            // output slice zi - 1 has all three planes now: one 16-byte store
            // per row through a descriptor over that slice (a lane past the
            // volume's x extent has an offset past its range: no store)
            const int zo = zi - 1;
            if (zo >= z0 && zo < z1) {
                const unsigned long long ob = (unsigned long long)(reinterpret_cast<__bf16 *>(out) + zo * oplane);
                const __amdgpu_buffer_rsrc_t orsrc = __builtin_amdgcn_make_buffer_rsrc(
                    (void *)(((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(ob >> 32)) << 32) |
                             (unsigned)__builtin_amdgcn_readfirstlane((unsigned)ob)),
                    0, (int)(oplane * sizeof(__bf16)), 0x00020000);
#pragma unroll
                for (int g = 0; g < R; ++g) {
                    const int y = y0 + g;   // uniform
                    if (y >= S.ny) continue;
                    typedef float f32x2 __attribute__((ext_vector_type(2)));
                    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
                    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
                    u32x4 o;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        f32x2 v = {acc[0][g][j >> 1][2 * (j & 1)], acc[0][g][j >> 1][2 * (j & 1) + 1]};
                        if (EPI != kEpiBiasRelu) {
                            const size_t vox = ((size_t)zo * S.ny + y) * S.nx + x;
#pragma unroll
                            for (int e = 0; e < 2; ++e)
                                v[e] = x < S.nx && bf2f(mask[vox * 32 + 8 * kg + 2 * j + e]) > 0.0f ? v[e] : 0.0f;
                        }
                        o[j] = __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
                    }
                    if (EPI == kEpiBiasRelu) {
                        // ReLU on the BF16 bit patterns: a negative value has
                        // the sign bit set, so it is a negative int16
                        typedef short i16x8 __attribute__((ext_vector_type(8)));
                        o = __builtin_bit_cast(u32x4, __builtin_elementwise_max(__builtin_bit_cast(i16x8, o),
                                                                                 (i16x8){0, 0, 0, 0, 0, 0, 0, 0}));
                    }
                    __builtin_amdgcn_raw_buffer_store_b128(o, orsrc, xoff_out, (int)((unsigned)y * ostride), 0);
                }
            }
            if (zin) {
                step(zi, 1, 0);
                step(zi, 1, 2);
                step(zi, 0, 1);
                step(zi, 0, 0);
                step(zi, 0, 2);
            } else {
#pragma unroll
                for (int g = 0; g < R; ++g)
#pragma unroll
                    for (int h = 0; h < 2; ++h) acc[2][g][h] = acc0[h];
            }
#pragma unroll
            for (int g = 0; g < R; ++g)
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    acc[0][g][h] = acc[1][g][h];
                    acc[1][g][h] = acc[2][g][h];
                }
            if (more) {
#pragma unroll
                for (int r = 0; r < R + 2; ++r) cur[r] = nxt[r];
                hal = nhal;
            }
        }
    }
}

// First layer (2 -> 32 channels) of the input pair (x_k, g_scale * step):
// BF16 pairs, 4 bytes per voxel, formed while the input is staged.  K of
// v_mfma_f32_16x16x32_bf16 runs over (tap, channel) pairs instead of
// channels: 27 taps x 2 channels in two steps of 32, so a row of 16 voxels
// takes 4 MFMAs where a channel-K layout takes 27 with 14 of every 16
// channels zero.  Lane l (column c = l mod 16, group g = l / 16) holds in
// step s the three taps dx = -1, 0, +1 of tap row (dz, dy) = 4 s + g (three
// consecutive voxels of the padded row, both channels) and, for s = 0 and
// g < 3, the tap (dz, dy, dx) = (+1, +1, g - 1) of the ninth tap row; the
// remaining slots have zero weights.
// A block owns a 16 x 32 tile (8 rows per wave) and walks kC2ZRun output
// slices: each input slice's 18 x 34 window (with the one-voxel halo; zero
// outside the volume)
// is loaded from the two FP32 inputs once, packed to BF16 pairs into a ring
// of four LDS slots, and the operands are read from LDS.  The slice two
// ahead loads while the current output slice is formed.
// Against loading the operands of every output row straight from global
// memory (each input voxel read 27 times through the L1; texture-unit
// bound): 0.45 -> 0.26 ms at 256^3, now bound by its 1 GB of output writes.
// Rows per wave 2 / 4 / 8 / 16: 0.32 / 0.275 / 0.265 / 0.31 ms.
constexpr int kC2Rows = 8;                               // output rows per wave
constexpr int kC2TY = 4 * kC2Rows;                       // tile rows per block
constexpr int kC2ZRun = 32;                              // output slices per block task
constexpr int kC2HX = 18, kC2HY = kC2TY + 2;             // staged window of one input slice
constexpr int kC2Win = kC2HX * kC2HY;                    // 324 voxels (one dword each)
constexpr int kC2Ld = (kC2Win + 255) / 256;              // window dwords per thread
__global__ __launch_bounds__(256) void k_conv3d_c2(const float *__restrict__ xin, const float *__restrict__ gin,
                                                   float g_scale, const __bf16 *__restrict__ wts,
                                                   const float *__restrict__ bias, __bf16 *__restrict__ out, CnnShape S)
{
    constexpr int R = kC2Rows;
    // slice zi in slot (zi + 1) mod 4; window column c is voxel x0 + c - 1
    __shared__ unsigned ring[4][kC2Win];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, kg = lane >> 4;
    // A[s][h]: K index 8 kg + 2 jj + ch; A row col = output channel
    // 8 (col / 4) + 4 h + col mod 4 (a lane's accumulators are then 8
    // consecutive channels)
    bf16x8 A[2][2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int r = 4 * s + kg, dz = r / 3, dy = r % 3;   // tap row of the triplet
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int co = 8 * (col >> 2) + 4 * h + (col & 3);
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                int t = -1;
                if (jj < 3) t = 9 * dz + 3 * dy + jj;
                else if (s == 0 && kg < 3) t = 24 + kg;   // tap row 8: (dz, dy) = (2, 2), dx = kg
#pragma unroll
                for (int ch = 0; ch < 2; ++ch)
                    A[s][h][2 * jj + ch] = t >= 0 ? wts[(t * 32 + co) * 16 + ch] : (__bf16)0.0f;
            }
        }
    }
    float bv[8];   // bias of this lane's 8 output channels
#pragma unroll
    for (int e = 0; e < 8; ++e) bv[e] = bias[8 * kg + e];

    const int tiles_x = (S.nx + 15) / 16, tiles_y = (S.ny + kC2TY - 1) / kC2TY;
    const int zruns = (S.nz + kC2ZRun - 1) / kC2ZRun;
    const int ntasks = tiles_x * tiles_y * zruns;
    const size_t plane = (size_t)S.ny * S.nx;   // voxels per slice
    for (int task = blockIdx.x; task < ntasks; task += gridDim.x) {
        const int x0 = (task % tiles_x) * 16, y0 = ((task / tiles_x) % tiles_y) * kC2TY;
        const int z0 = (task / (tiles_x * tiles_y)) * kC2ZRun, z1 = min(S.nz, z0 + kC2ZRun);
        const int x = x0 + col;
        // window of input slice zi (voxels x0 - 1 .. x0 + 16, rows y0 - 1 ..
        // y0 + 16): voxel i of the window goes to thread i mod 256; outside
        // the volume it reads as zero (offset past the descriptor's range)
        auto slice_rsrc = [&](const float *p, int zi) {
            const bool zok = zi >= 0 && zi < S.nz;
            const unsigned long long base = (unsigned long long)(p + (zok ? zi : 0) * plane);
            const unsigned bhi = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)(base >> 32));
            const unsigned blo = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)base);
            return __builtin_amdgcn_make_buffer_rsrc((void *)(((unsigned long long)bhi << 32) | blo), 0,
                                                     __builtin_amdgcn_readfirstlane(zok ? (int)(plane * sizeof(float)) : 0),
                                                     0x00020000);
        };
        auto fetch = [&](int zi, float (&vx)[kC2Ld], float (&vg)[kC2Ld]) {
            const __amdgpu_buffer_rsrc_t rx = slice_rsrc(xin, zi), rg = slice_rsrc(gin, zi);
#pragma unroll
            for (int k = 0; k < kC2Ld; ++k) {
                const int i = (int)threadIdx.x + 256 * k, hy = i / kC2HX, hx = i - hy * kC2HX;
                const int xx = x0 - 1 + hx, yy = y0 - 1 + hy;
                const bool ok = i < kC2Win && xx >= 0 && xx < S.nx && yy >= 0 && yy < S.ny;
                const unsigned off = ok ? (unsigned)((yy * S.nx + xx) * 4) : 0x80000000u;
                vx[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rx, off, 0, 0));
                vg[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rg, off, 0, 0));
            }
        };
        auto stash = [&](int zi, const float (&vx)[kC2Ld], const float (&vg)[kC2Ld]) {
            typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
#pragma unroll
            for (int k = 0; k < kC2Ld; ++k) {
                const int i = (int)threadIdx.x + 256 * k;
                bf16x2 p;
                p[0] = (__bf16)vx[k];
                p[1] = (__bf16)(vg[k] * g_scale);
                if (i < kC2Win) ring[(zi + 1) & 3][i] = __builtin_bit_cast(unsigned, p);
            }
        };
        float vx[kC2Ld], vg[kC2Ld];
        __syncthreads();   // the previous task's reads of the ring are done
#pragma unroll 1
        for (int zi = z0 - 1; zi <= z0 + 1; ++zi) {
            fetch(zi, vx, vg);
            stash(zi, vx, vg);
        }
        __syncthreads();
#pragma unroll 1
        for (int zo = z0; zo < z1; ++zo) {
            // output slice zo + 1 needs input slice zo + 2
            const bool more = zo + 2 <= z1;
            if (more) fetch(zo + 2, vx, vg);
            f32x4 acc[R][2];
#pragma unroll
            for (int g = 0; g < R; ++g) {
                const int wy = wave * R + g;   // window row of output row y0 + wy, minus one
                bf16x8 B[2];
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
                    const int r = 4 * s + kg, dz = r / 3, dy = r % 3;
                    const unsigned *w = &ring[(zo + dz) & 3][(wy + dy) * kC2HX + col];
                    u32x4 b;
                    b[0] = w[0];
                    b[1] = w[1];
                    b[2] = w[2];
                    b[3] = 0u;
                    // tap (+1, +1, kg - 1) for kg < 3: slice zo + 1, row y + 1, window column col + kg
                    if (s == 0 && kg < 3) b[3] = ring[(zo + 2) & 3][(wy + 2) * kC2HX + col + kg];
                    B[s] = __builtin_bit_cast(bf16x8, b);
                }
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    acc[g][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0][h], B[0], f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
                    acc[g][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1][h], B[1], acc[g][h], 0, 0, 0);
                }
            }
            // lane holds D[co][voxel x] for co = 8 kg + 4 h + u (u = 0..3): one
            // 16-byte store per row
            if (x < S.nx) {
#pragma unroll
                for (int g = 0; g < R; ++g) {
                    const int y = y0 + wave * R + g;
                    if (y >= S.ny) continue;
                    const size_t vox = ((size_t)zo * S.ny + y) * S.nx + x;
                    bf16x8 o;
#pragma unroll
                    for (int h = 0; h < 2; ++h)
#pragma unroll
                        for (int u = 0; u < 4; ++u) o[4 * h + u] = (__bf16)fmaxf(acc[g][h][u] + bv[4 * h + u], 0.0f);
                    *reinterpret_cast<bf16x8 *>(out + vox * 32 + 8 * kg) = o;
                }
            }
            // slot of slice zo + 2 last held slice zo - 2, read by output slice
            // zo - 1 before the previous barrier
            if (more) stash(zo + 2, vx, vg);
            __syncthreads();
        }
    }
}

// Last layer of the network (32 -> 1 channel, linear, FP32 out):
//   out[v] = b + sum_t sum_ci W[t][ci] a[v + off(t)][ci].
// A 32-output convolution would spend 31/32 of its MFMA work on padding.
// Here the taps are the MFMA's M dimension instead: per input slice a block
// forms P[t][u] = sum_ci W[t][ci] a[u][ci] (v_mfma_f32_32x32x16_bf16, A =
// W^T with 27 of 32 rows, B = the activations of 32 voxels, K = 32 channels
// in two steps) for every voxel u of its x-y tile with a one-voxel halo,
// stores P in LDS, and each output gathers its nine taps of that slice.  A
// slice contributes to the outputs one below (taps dz = +1), level (dz = 0)
// and one above (dz = -1): three rolling sums per output, the lowest one
// complete once the slice is done.  The block walks kConvZRun output slices.
// A 16 x 16 tile (38 KB of LDS, four blocks per CU) against 32 x 16 (70 KB,
// two blocks): 0.28 -> 0.27 ms at 256^3, despite the larger halo share.
constexpr int kO1TX = 16, kO1TY = 16, kO1HX = kO1TX + 2, kO1HY = kO1TY + 2;
constexpr int kO1K = kO1TX * kO1TY / 256, kO1Step = 256 / kO1TX;   // outputs per thread, their row step
constexpr int kO1Halo = kO1HX * kO1HY;                  // 324 voxels
constexpr int kO1Tiles = (kO1Halo + 31) / 32;           // 11 MFMA column tiles
constexpr int kO1Stride = kO1Tiles * 32 + 4;            // LDS floats per tap (+4: the two lane halves' taps 4 apart
                                                        // fall in different banks)

__global__ __launch_bounds__(256) void k_conv3d_out1(const __bf16 *__restrict__ in, const __bf16 *__restrict__ wts,
                                                     const float *__restrict__ bias, float *__restrict__ out,
                                                     CnnShape S)
{
    __shared__ float pl[kTaps * kO1Stride];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 31, kh = lane >> 5;
    // A = W^T (rows = taps, 27 of 32; K = input channels), from the packed
    // forward layout [tap][q][co][16] at output channel 0
    bf16x8 a[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = (__bf16)0.0f;
        if (col < kTaps) v = reinterpret_cast<const bf16x8 *>(wts)[((col * 2 + q) * 32) * 2 + kh];
        a[q] = v;
    }
    const float b0 = bias[0];
    const size_t plane = (size_t)S.ny * S.nx * 32;   // bf16 elements per slice
    const int tiles_x = (S.nx + kO1TX - 1) / kO1TX, tiles_y = (S.ny + kO1TY - 1) / kO1TY;
    const int zruns = (S.nz + kConvZRun - 1) / kConvZRun;
    const int ncols = tiles_x * tiles_y * zruns;
    const int ox = threadIdx.x % kO1TX, oy0 = threadIdx.x / kO1TX;   // outputs (ox, oy0 + kO1Step k)
    for (int tcol = blockIdx.x; tcol < ncols; tcol += gridDim.x) {
        const int x0 = (tcol % tiles_x) * kO1TX, y0 = ((tcol / tiles_x) % tiles_y) * kO1TY;
        const int zs = (tcol / (tiles_x * tiles_y)) * kConvZRun, ze = min(S.nz, zs + kConvZRun);
        float acc_lo[kO1K], acc_mid[kO1K], acc_hi[kO1K];
#pragma unroll
        for (int k = 0; k < kO1K; ++k) acc_lo[k] = acc_mid[k] = acc_hi[k] = 0.0f;
        // the halo voxels of a slice for this wave's column tiles T = wave +
        // 4 j, loaded one slice ahead: the loads of slice zi + 1 are in
        // flight while slice zi is multiplied and gathered
        constexpr int NJ = (kO1Tiles + 3) / 4;
        auto load_slice = [&](int zi, bf16x8 (*dst)[2]) {
            const bool zok = zi >= 0 && zi < S.nz;
            const unsigned long long base = (unsigned long long)(in + (zok ? zi : 0) * plane);
            const unsigned bhi = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)(base >> 32));
            const unsigned blo = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)base);
            const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(((unsigned long long)bhi << 32) | blo), 0,
                __builtin_amdgcn_readfirstlane(zok ? (int)(plane * sizeof(__bf16)) : 0), 0x00020000);
#pragma unroll
            for (int jt = 0; jt < NJ; ++jt) {
                const int T = wave + 4 * jt;
                const int i = T * 32 + col, hx = i % kO1HX, hy = i / kO1HX;
                const int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
                const bool ok = T < kO1Tiles && i < kO1Halo && gx >= 0 && gx < S.nx && gy >= 0 && gy < S.ny;
                // a voxel outside the volume reads as zero (offset past the descriptor's range)
                const unsigned off =
                    ok ? (unsigned)((((size_t)gy * S.nx + gx) * 32 + kh * 8) * sizeof(__bf16)) : 0x80000000u;
                dst[jt][0] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off, 0, 0));
                dst[jt][1] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off + 32u, 0, 0));
            }
        };
        bf16x8 cb[NJ][2], nb[NJ][2];
        load_slice(zs - 1, cb);
        for (int zi = zs - 1; zi <= ze; ++zi) {
            if (zi + 1 <= ze) load_slice(zi + 1, nb);
            if (zi >= 0 && zi < S.nz) {   // uniform: slices outside the volume contribute zero
                __syncthreads();   // the previous slice's gathers are done with pl
#pragma unroll
                for (int jt = 0; jt < NJ; ++jt) {
                    const int T = wave + 4 * jt;
                    if (T >= kO1Tiles) continue;
                    const int i = T * 32 + col;
                    f32x16 d;
#pragma unroll
                    for (int r = 0; r < 16; ++r) d[r] = 0.0f;
                    d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], cb[jt][0], d, 0, 0, 0);
                    d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1], cb[jt][1], d, 0, 0, 0);
                    // lane holds P[8 s + 4 kh + u][voxel T*32 + col]
#pragma unroll
                    for (int sg = 0; sg < 4; ++sg)
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int t = 8 * sg + 4 * kh + u;
                            if (t < kTaps) pl[t * kO1Stride + i] = d[4 * sg + u];
                        }
                }
                __syncthreads();
#pragma unroll
                for (int k = 0; k < kO1K; ++k) {
                    const int oy = oy0 + kO1Step * k;
                    float below = 0.0f, level = 0.0f, above = 0.0f;
#pragma unroll
                    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                        for (int dx = 0; dx < 3; ++dx) {
                            const int idx = (oy + dy) * kO1HX + ox + dx;
                            above += pl[(0 * 9 + dy * 3 + dx) * kO1Stride + idx];   // dz = -1: output zi + 1
                            level += pl[(1 * 9 + dy * 3 + dx) * kO1Stride + idx];   // dz = 0
                            below += pl[(2 * 9 + dy * 3 + dx) * kO1Stride + idx];   // dz = +1: output zi - 1
                        }
                    acc_hi[k] += above;
                    acc_mid[k] += level;
                    acc_lo[k] += below;
                }
            }
            const int zo = zi - 1;   // complete now
            if (zo >= zs && zo < ze) {
#pragma unroll
                for (int k = 0; k < kO1K; ++k) {
                    const int x = x0 + ox, y = y0 + oy0 + kO1Step * k;
                    if (x < S.nx && y < S.ny) out[((size_t)zo * S.ny + y) * S.nx + x] = acc_lo[k] + b0;
                }
            }
#pragma unroll
            for (int k = 0; k < kO1K; ++k) {
                acc_lo[k] = acc_mid[k];
                acc_mid[k] = acc_hi[k];
                acc_hi[k] = 0.0f;
            }
            if (zi + 1 <= ze) {
#pragma unroll
                for (int jt = 0; jt < NJ; ++jt) {
                    cb[jt][0] = nb[jt][0];
                    cb[jt][1] = nb[jt][1];
                }
            }
        }
    }
}

// Weight and bias gradient of one layer over a volume:
//   dW[co][ci][t] += sum_v delta[v][co] * act[v + off(t)][ci],  db[co] += sum_v delta[v][co]
// Tiles of 32 x 8 voxels of one slice; the activations with their halo
// (34 x 10 x 3) and the deltas are staged in LDS.  Thread (co, ci block of
// 4) accumulates 4 x 27 weights over every tile it visits; each block
// writes its partial sums, reduced afterwards by k_reduce_grad.
template <int CI, int CO>
__global__ __launch_bounds__(256) void k_conv3d_wgrad(const __bf16 *__restrict__ act, const __bf16 *__restrict__ delta,
                                                      float *__restrict__ partial, CnnShape S)
{
    constexpr int HX = 34, HY = 10, HZ = 3, HV = HX * HY * HZ;
    constexpr int CB = CI / 4;                       // ci blocks
    constexpr int NT = CO * CB;                      // active threads
    static_assert(NT <= 256, "too many weight blocks");
    __shared__ __bf16 sa[HV * CI];
    __shared__ __bf16 sd[256 * CO];
    const int tid = threadIdx.x;
    const int co = tid % CO, cb = tid / CO;
    float acc[4][kTaps], accb = 0.0f;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int t = 0; t < kTaps; ++t) acc[c][t] = 0.0f;

    const int tiles_x = (S.nx + 31) / 32, tiles_y = (S.ny + 7) / 8;
    const int ntiles = tiles_x * tiles_y * S.nz;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int tx = tile % tiles_x, rest = tile / tiles_x;
        const int ty = rest % tiles_y, z = rest / tiles_y;
        const int x0 = tx * 32, y0 = ty * 8;
        __syncthreads();
        // stage activations with halo (zero outside the volume), 8 channels per step
        for (int e = tid; e < HV * (CI / 8); e += 256) {
            const int hv = e / (CI / 8), c8 = e % (CI / 8);
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int xx = x0 - 1 + hx, yy = y0 - 1 + hy, zz = z - 1 + hz;
            bf16x8 v;
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = (__bf16)0.0f;
            if (xx >= 0 && xx < S.nx && yy >= 0 && yy < S.ny && zz >= 0 && zz < S.nz)
                v = *reinterpret_cast<const bf16x8 *>(act + (((size_t)zz * S.ny + yy) * S.nx + xx) * CI + c8 * 8);
            *reinterpret_cast<bf16x8 *>(sa + hv * CI + c8 * 8) = v;
        }
        // stage deltas of the 32 x 8 tile (zero outside)
        for (int e = tid; e < 256 * (CO / 8); e += 256) {
            const int v = e / (CO / 8), c8 = e % (CO / 8);
            const int xx = x0 + (v & 31), yy = y0 + (v >> 5);
            bf16x8 d;
#pragma unroll
            for (int k = 0; k < 8; ++k) d[k] = (__bf16)0.0f;
            if (xx < S.nx && yy < S.ny)
                d = *reinterpret_cast<const bf16x8 *>(delta + (((size_t)z * S.ny + yy) * S.nx + xx) * CO + c8 * 8);
            *reinterpret_cast<bf16x8 *>(sd + v * CO + c8 * 8) = d;
        }
        __syncthreads();
        if (tid < NT) {
            for (int v = 0; v < 256; ++v) {
                const float dv = bf2f(sd[v * CO + co]);
                accb += dv;
                const int vx = v & 31, vy = v >> 5;
#pragma unroll
                for (int t = 0; t < kTaps; ++t) {
                    const int dz = t / 9, dy = (t / 3) % 3, dx = t % 3;   // halo offsets 0..2
                    const int hv = (dz * HY + vy + dy) * HX + vx + dx;
                    const bf16x4 a = *reinterpret_cast<const bf16x4 *>(sa + hv * CI + cb * 4);
#pragma unroll
                    for (int c = 0; c < 4; ++c) acc[c][t] = fmaf(dv, bf2f(a[c]), acc[c][t]);
                }
            }
        }
    }
    // partial layout per block: [CO][CI][27] weights then [CO] bias
    constexpr int NP = CO * CI * kTaps + CO;
    float *P = partial + (size_t)blockIdx.x * NP;
    if (tid < NT) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int t = 0; t < kTaps; ++t) P[(co * CI + cb * 4 + c) * kTaps + t] = acc[c][t];
        if (cb == 0) P[CO * CI * kTaps + co] = accb;
    }
}

// Deterministic reduction of per-block gradient partials into the flat
// parameter layout of one layer ([co][ci][27] weights, then [co] bias):
// partials use padded channel counts ([co_pad][ci_pad][27] + [co_pad]).
__global__ __launch_bounds__(256) void k_reduce_grad(const float *__restrict__ partial, int nblk, int ci_pad,
                                                     int co_pad, int ci_real, int co_real, float *__restrict__ out,
                                                     int accumulate)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int nw = co_real * ci_real * kTaps;
    if (i >= nw + co_real) return;
    int src;
    if (i < nw) {
        const int co = i / (ci_real * kTaps), rem = i % (ci_real * kTaps);
        src = (co * ci_pad + rem / kTaps) * kTaps + rem % kTaps;
    } else {
        src = co_pad * ci_pad * kTaps + (i - nw);
    }
    const size_t np = (size_t)co_pad * ci_pad * kTaps + co_pad;
    float s = 0.0f;
    for (int b = 0; b < nblk; ++b) s += partial[(size_t)b * np + src];
    out[i] = accumulate ? out[i] + s : s;
}

// SIRT step s = lambda * vinv * g.
__global__ __launch_bounds__(256) void k_sirt_step(const float *__restrict__ vinv, const float *__restrict__ g,
                                                   float lambda, size_t n, float *__restrict__ s)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) s[i] = lambda * vinv[i] * g[i];
}

// FP32 parameters (PyTorch conv3d layout [co][ci][kz][ky][kx] + bias) ->
// BF16 MFMA A layouts: forward [tap][q][32 co][16 ci] and, for the input
// gradient, the flipped/transposed convolution [tap'][q'][32 ci][16 co].
__global__ __launch_bounds__(256) void k_pack_weights(const float *__restrict__ w, int ci_real, int co_real, int cin_pad,
                                                      __bf16 *__restrict__ fwd, int cout_pad_t, __bf16 *__restrict__ bwd)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    // forward layout
    const int nq = cin_pad / 16, nf = kTaps * nq * 32 * 16;
    if (i < nf) {
        const int k = i % 16, co = (i / 16) % 32, q = (i / 512) % nq, t = i / (512 * nq);
        const int ci = q * 16 + k;
        float v = 0.0f;
        if (co < co_real && ci < ci_real) v = w[((size_t)co * ci_real + ci) * kTaps + t];
        fwd[i] = (__bf16)v;
    }
    if (!bwd) return;
    // backward-data layout: input channels = forward outputs (padded to cout_pad_t)
    const int nqb = cout_pad_t / 16, nb = kTaps * nqb * 32 * 16;
    if (i < nb) {
        const int k = i % 16, m = (i / 16) % 32, q = (i / 512) % nqb, t = i / (512 * nqb);
        const int co = q * 16 + k, ci = m;   // forward output channel, forward input channel
        float v = 0.0f;
        if (co < co_real && ci < ci_real) v = w[((size_t)co * ci_real + ci) * kTaps + (kTaps - 1 - t)];
        bwd[i] = (__bf16)v;
    }
}

// Adam (decoupled from the gradient scale; bias-corrected step size given).
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                                              const float *__restrict__ g, int n, float lr_t, float b1, float b2,
                                              float eps)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

// Network input: channel 0 = x, channel 1 = g * g_scale, channels 2..15 = 0.
__global__ __launch_bounds__(256) void k_pack_input(const float *__restrict__ x, const float *__restrict__ g,
                                                    float g_scale, size_t n, __bf16 *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    bf16x8 a, b;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        a[k] = (__bf16)0.0f;
        b[k] = (__bf16)0.0f;
    }
    a[0] = (__bf16)x[i];
    a[1] = (__bf16)(g[i] * g_scale);
    reinterpret_cast<bf16x8 *>(out)[2 * i] = a;
    reinterpret_cast<bf16x8 *>(out)[2 * i + 1] = b;
}

// Masked mean-squared error of pred against target over the interior of
// the volume (voxels at least `border` from every face) and its gradient,
// written as the BF16 delta of the last layer (channel 0 of 16).
// partial[block] = (sum of squared errors, 0); inv_count = 1 / (interior voxels of all patches).
__global__ __launch_bounds__(256) void k_mse_grad(const float *__restrict__ pred, const float *__restrict__ target,
                                                  CnnShape S, int border, float inv_count,
                                                  __bf16 *__restrict__ delta, double *__restrict__ partial)
{
    __shared__ double red[256];
    const size_t n = (size_t)S.nx * S.ny * S.nz;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    double e2 = 0.0;
    if (i < n) {
        const int x = (int)(i % S.nx), y = (int)((i / S.nx) % S.ny), z = (int)(i / ((size_t)S.nx * S.ny));
        const bool in = x >= border && x < S.nx - border && y >= border && y < S.ny - border && z >= border &&
                        z < S.nz - border;
        const float e = in ? pred[i] - target[i] : 0.0f;
        e2 = (double)e * e;
        bf16x8 d0, d1;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            d0[k] = (__bf16)0.0f;
            d1[k] = (__bf16)0.0f;
        }
        d0[0] = (__bf16)(2.0f * e * inv_count);
        reinterpret_cast<bf16x8 *>(delta)[2 * i] = d0;
        reinterpret_cast<bf16x8 *>(delta)[2 * i + 1] = d1;
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = red[0];
        partial[2 * blockIdx.x + 1] = 0.0;
    }
}

// Learned-step candidate: x' = clamp(x + step + out_scale * d, lo, hi).
__global__ __launch_bounds__(256) void k_learned_candidate(const float *__restrict__ x, const float *__restrict__ step,
                                                           const float *__restrict__ d, float out_scale, float lo,
                                                           float hi, size_t n, float *__restrict__ xn)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    xn[i] = fminf(fmaxf(x[i] + step[i] + out_scale * d[i], lo), hi);
}

int conv_grid(const CnnShape &S, int cin)
{
    const int ty = 4 * conv_rows(cin);
    const long cols = (long)((S.nx + 31) / 32) * ((S.ny + ty - 1) / ty) * ((S.nz + kConvZRun - 1) / kConvZRun);
    return (int)(cols < 2048 ? cols : 2048);
}

}  // namespace

// The kernels address a slice through a buffer descriptor with a 32-bit
// byte range: one slice of 32 BF16 channels must stay below 2 GiB.
static bool slice_fits(const CnnShape &S) { return (size_t)S.nx * S.ny * 32 * sizeof(__bf16) < (size_t(1) << 31); }

hipError_t launch_conv3d(int cin_pad, int epi, const __bf16 *in, const __bf16 *w, const float *bias, const __bf16 *mask,
                         void *out, const CnnShape &S, hipStream_t st)
{
    if (!slice_fits(S)) return hipErrorInvalidValue;
    const int grid = conv_grid(S, cin_pad);
#define CTR_CONV(CIN, EPI)                                                                                          \
    hipLaunchKernelGGL((k_conv3d<CIN, EPI>), dim3(grid), dim3(256), 0, st, in, w, bias, mask, out, S)
    if (cin_pad == 16) {
        if (epi == kEpiBiasRelu) CTR_CONV(16, kEpiBiasRelu);
        else if (epi == kEpiLinearF32) CTR_CONV(16, kEpiLinearF32);
        else if (epi == kEpiMask) CTR_CONV(16, kEpiMask);
        else return hipErrorInvalidValue;
    } else if (cin_pad == 32) {
        const long tiles = (long)((S.nx + 15) / 16) * ((S.ny + 4 * kC32zRows - 1) / (4 * kC32zRows));
        int zrun = 64;
        while (zrun > 16 && tiles * ((S.nz + zrun - 1) / zrun) < kC32MinTasks) zrun /= 2;
        const long cols = tiles * ((S.nz + zrun - 1) / zrun);
        const dim3 g32((unsigned)std::min(cols, 8192L));
        if (epi == kEpiBiasRelu)
            hipLaunchKernelGGL((k_conv3d_c32<kEpiBiasRelu>), g32, dim3(256), 0, st, in, w, bias, mask, out, S, zrun);
        else if (epi == kEpiLinearF32) {
            // one real output channel: the taps-as-rows kernel
            const long cols = (long)((S.nx + kO1TX - 1) / kO1TX) * ((S.ny + kO1TY - 1) / kO1TY) *
                              ((S.nz + kConvZRun - 1) / kConvZRun);
            hipLaunchKernelGGL(k_conv3d_out1, dim3((unsigned)std::min(cols, 8192L)), dim3(256), 0, st, in, w, bias,
                               (float *)out, S);
        }
        else if (epi == kEpiMask)
            hipLaunchKernelGGL((k_conv3d_c32<kEpiMask>), g32, dim3(256), 0, st, in, w, bias, mask, out, S, zrun);
        else return hipErrorInvalidValue;
    } else {
        return hipErrorInvalidValue;
    }
#undef CTR_CONV
    return hipGetLastError();
}

int wgrad_blocks(const CnnShape &S)
{
    const long tiles = (long)((S.nx + 31) / 32) * ((S.ny + 7) / 8) * S.nz;
    return (int)(tiles < 512 ? tiles : 512);
}

hipError_t launch_conv3d_wgrad(int ci, int co, const __bf16 *act, const __bf16 *delta, float *partial, int nblk,
                               const CnnShape &S, hipStream_t st)
{
    if (ci == 16 && co == 32)
        hipLaunchKernelGGL((k_conv3d_wgrad<16, 32>), dim3(nblk), dim3(256), 0, st, act, delta, partial, S);
    else if (ci == 32 && co == 32)
        hipLaunchKernelGGL((k_conv3d_wgrad<32, 32>), dim3(nblk), dim3(256), 0, st, act, delta, partial, S);
    else if (ci == 32 && co == 16)
        hipLaunchKernelGGL((k_conv3d_wgrad<32, 16>), dim3(nblk), dim3(256), 0, st, act, delta, partial, S);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_reduce_grad(const float *partial, int nblk, int ci_pad, int co_pad, int ci_real, int co_real,
                              float *out, int accumulate, hipStream_t st)
{
    const int n = co_real * ci_real * kTaps + co_real;
    hipLaunchKernelGGL(k_reduce_grad, dim3((n + 255) / 256), dim3(256), 0, st, partial, nblk, ci_pad, co_pad, ci_real,
                       co_real, out, accumulate);
    return hipGetLastError();
}

hipError_t launch_sirt_step(const float *vinv, const float *g, float lambda, size_t n, float *s, hipStream_t st)
{
    hipLaunchKernelGGL(k_sirt_step, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, vinv, g, lambda, n, s);
    return hipGetLastError();
}

hipError_t launch_pack_weights(const float *w, int ci_real, int co_real, int cin_pad, __bf16 *fwd, int cout_pad_t,
                               __bf16 *bwd, hipStream_t st)
{
    const int nf = kTaps * (cin_pad / 16) * 512, nb = bwd ? kTaps * (cout_pad_t / 16) * 512 : 0;
    const int n = nf > nb ? nf : nb;
    hipLaunchKernelGGL(k_pack_weights, dim3((n + 255) / 256), dim3(256), 0, st, w, ci_real, co_real, cin_pad, fwd,
                       cout_pad_t, bwd);
    return hipGetLastError();
}

hipError_t launch_adam(float *p, float *m, float *v, const float *g, int n, float lr_t, float b1, float b2, float eps,
                       hipStream_t st)
{
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, p, m, v, g, n, lr_t, b1, b2, eps);
    return hipGetLastError();
}

hipError_t launch_pack_input(const float *x, const float *g, float g_scale, size_t n, __bf16 *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_pack_input, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, g, g_scale, n, out);
    return hipGetLastError();
}

hipError_t launch_conv3d_first(const float *x, const float *g, float g_scale, const __bf16 *w, const float *bias,
                               __bf16 *out, const CnnShape &S, hipStream_t st)
{
    if (!slice_fits(S)) return hipErrorInvalidValue;
    const long tasks = (long)((S.nx + 15) / 16) * ((S.ny + kC2TY - 1) / kC2TY) * ((S.nz + kC2ZRun - 1) / kC2ZRun);
    const dim3 grid((unsigned)std::min(tasks, 8192L));
    hipLaunchKernelGGL(k_conv3d_c2, grid, dim3(256), 0, st, x, g, g_scale, w, bias, out, S);
    return hipGetLastError();
}

hipError_t launch_mse_grad(const float *pred, const float *target, const CnnShape &S, int border, float inv_count,
                           __bf16 *delta, double *partial, int *nblocks, hipStream_t st)
{
    const size_t n = (size_t)S.nx * S.ny * S.nz;
    const int nb = (int)((n + 255) / 256);
    *nblocks = nb;
    hipLaunchKernelGGL(k_mse_grad, dim3(nb), dim3(256), 0, st, pred, target, S, border, inv_count, delta, partial);
    return hipGetLastError();
}

hipError_t launch_learned_candidate(const float *x, const float *step, const float *d, float out_scale, float lo,
                                    float hi, size_t n, float *xn, hipStream_t st)
{
    hipLaunchKernelGGL(k_learned_candidate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, step, d,
                       out_scale, lo, hi, n, xn);
    return hipGetLastError();
}

}  // namespace ctr
