// ctr_cnn.h - launchers of the 3-D CNN kernels of the learned SIRT update.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace ctr {

struct CnnShape {
    int nx, ny, nz;         // volume dimensions, x fastest (the t, s, z axes of the CT volume)
};

enum CnnEpi {
    kEpiBiasRelu = 0,       // bf16 out [V][32] = relu(acc + bias)
    kEpiLinearF32 = 1,      // float out [V] = acc[channel 0] + bias[0]
    kEpiMask = 2,           // bf16 out [V][32] = acc where mask[V][32] > 0, else 0 (ReLU derivative)
};

// conv3d 3x3x3, zero padding 1, cin_pad in {16, 32} input channels (BF16,
// channels-last), 32 output channels; w in the packed forward layout.
hipError_t launch_conv3d(int cin_pad, int epi, const __bf16 *in, const __bf16 *w, const float *bias,
                         const __bf16 *mask, void *out, const CnnShape &S, hipStream_t st);
// weight/bias gradient partials [nblk][co*ci*27 + co]; (ci, co) in {(16,32), (32,32), (32,16)}
int wgrad_blocks(const CnnShape &S);
hipError_t launch_conv3d_wgrad(int ci, int co, const __bf16 *act, const __bf16 *delta, float *partial, int nblk,
                               const CnnShape &S, hipStream_t st);
// sum of the partials into the flat layer layout [co_real][ci_real][27] + [co_real] (+= when accumulate)
hipError_t launch_reduce_grad(const float *partial, int nblk, int ci_pad, int co_pad, int ci_real, int co_real,
                              float *out, int accumulate, hipStream_t st);
// s = lambda * vinv * g
hipError_t launch_sirt_step(const float *vinv, const float *g, float lambda, size_t n, float *s, hipStream_t st);
// FP32 [co_real][ci_real][27] weights -> packed BF16 forward (cin_pad) and input-gradient (cout_pad_t) layouts
hipError_t launch_pack_weights(const float *w, int ci_real, int co_real, int cin_pad, __bf16 *fwd, int cout_pad_t,
                               __bf16 *bwd, hipStream_t st);
hipError_t launch_adam(float *p, float *m, float *v, const float *g, int n, float lr_t, float b1, float b2, float eps,
                       hipStream_t st);
hipError_t launch_pack_input(const float *x, const float *g, float g_scale, size_t n, __bf16 *out, hipStream_t st);
// first layer (2 -> 32, bias + ReLU) of the input pair (x, g * g_scale)
hipError_t launch_conv3d_first(const float *x, const float *g, float g_scale, const __bf16 *w, const float *bias,
                               __bf16 *out, const CnnShape &S, hipStream_t st);
hipError_t launch_mse_grad(const float *pred, const float *target, const CnnShape &S, int border, float inv_count,
                           __bf16 *delta, double *partial, int *nblocks, hipStream_t st);
hipError_t launch_learned_candidate(const float *x, const float *step, const float *d, float out_scale, float lo,
                                    float hi, size_t n, float *xn, hipStream_t st);

}  // namespace ctr
