// nn.hpp -- launchers of the learned preconditioner's network kernels
// (nn_kernels.hip).  Grids are node counts per direction {x, y, z}.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mfx {

// layer 0 (1 -> 16, FP32 input in32), 1..6 (BF16 MFMA), 7 (16 -> 1, FP32
// output out32); od: output grid, id: input tensor grid
void nn_conv_forward(int layer, const uint16_t *in, const float *in32, const uint16_t *wb,
                     const float *wr, const uint16_t *skip, uint16_t *out, float *out32,
                     const uint32_t *od, const uint32_t *id, hipStream_t s);
void nn_conv_adjoint(int stride, const float *G, const float *W, float *dX, const uint32_t *od,
                     const uint32_t *xd, int cin, int cout, hipStream_t s);
void nn_up_adjoint_mask(const float *fine, const uint32_t *fd, const float *add,
                        const uint16_t *act, float *coarse, const uint32_t *cd, int C,
                        hipStream_t s);
void nn_add_mask(const float *a, const float *b, const uint16_t *act, float *out, uint64_t n,
                 hipStream_t s);
// grad += dL/dW of one layer; part: scratch of n_parts * 27 * cin * cout floats
void nn_conv_wgrad(int stride, bool up, const uint16_t *in, const float *in32, const float *G,
                   float *part, float *grad, const uint32_t *od, const uint32_t *id, int cin,
                   int cout, uint32_t n_parts, hipStream_t s);
void nn_head_backward(const float *dy, const float *W, const uint16_t *act, float *g,
                      const uint32_t *d, hipStream_t s);
uint32_t nn_mse_blocks(uint64_t n);
void nn_mse(const float *y, const float *target, const uint8_t *mask, float scale, float *dy,
            double *part, uint64_t n, hipStream_t s);
void nn_adam(float *w, float *m, float *v, const float *g, uint32_t n, float lr, float b1, float b2,
             float eps, int step, hipStream_t s);
void nn_pack_weights(const float *w, uint16_t *wb, float *wr, int cin, int cout, int kp,
                     hipStream_t s);
void nn_scale_to_f32(const double *r, float *out, double scale, uint64_t n, hipStream_t s);
void nn_scale_to_f64(const float *y, double *z, double scale, uint64_t n, hipStream_t s);
void nn_axpy_f32(float *x, float a, const float *y, uint64_t n, hipStream_t s);

}  // namespace mfx
