// mf_cost.hpp -- shapes and parameter layout of the learned subtree-cost model.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mf {

constexpr int kFeatures = 80;   // 64 state encoding + 16 scalar features
constexpr int kHidden = 128;

// Flat FP32 parameter vector: W1 [128 x 80], b1, W2 [128 x 128], b2,
// W3 [128 x 128], b3, w4 [128], b4 (row-major, out x in).
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHidden * kFeatures;
constexpr int kOffW2 = kOffB1 + kHidden;
constexpr int kOffB2 = kOffW2 + kHidden * kHidden;
constexpr int kOffW3 = kOffB2 + kHidden;
constexpr int kOffB3 = kOffW3 + kHidden * kHidden;
constexpr int kOffW4 = kOffB3 + kHidden;
constexpr int kOffB4 = kOffW4 + kHidden;
constexpr int kParams = kOffB4 + 1;

// BF16 inference image: MFMA A-fragments of the three hidden layers (64 lanes
// x 4 halves per 16x16 tile) and the FP32 biases / head.
constexpr int kFragHalves = (kHidden / 16) * (kFeatures / 16 + 2 * (kHidden / 16)) * 256;
constexpr int kSmallFloats = 4 * kHidden + 4;

static_assert(kFeatures % 16 == 0, "features must fill whole MFMA k blocks");

hipError_t cost_linear_fwd(int rows, int in, const float *x, const float *w, const float *b, float *y,
                           hipStream_t s);
hipError_t cost_linear_bwd_input(int rows, int in, const float *dy, const float *w, const float *x, float *dx,
                                 hipStream_t s);
constexpr int kSplitK = 8;   // weight-gradient products: partial tiles over the batch
hipError_t cost_linear_bwd_weight(int rows, int in, const float *dy, const float *x, float *dw, float *db,
                                  float *part, hipStream_t s);
hipError_t cost_head_fwd(int rows, const float *h, const float *w, float *out, hipStream_t s);
hipError_t cost_mse(int rows, const float *pred, const float *target, float *g, float *partial, float *loss,
                    hipStream_t s);
hipError_t cost_head_bwd(int rows, const float *g, const float *w, const float *h, float *dw, float *dh,
                         hipStream_t s);
hipError_t cost_adam(int n, float *w, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                     int step, hipStream_t s);
hipError_t cost_gather(int rows, int cols, const uint32_t *idx, const float *src, float *dst, hipStream_t s);
hipError_t cost_pack_bf16(const float *params, uint16_t *frag, float *small, hipStream_t s);
uint32_t cost_infer_lds_bytes();
hipError_t cost_infer_bf16(int rows, const float *x, const uint16_t *frag, const float *small, float *out, int cus,
                           hipStream_t s);

}  // namespace mf
