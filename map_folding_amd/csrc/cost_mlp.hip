// cost_mlp.hip -- the learned subtree-cost model on gfx950.
//
// A 4-layer perceptron, kFeatures -> 128 -> 128 -> 128 -> 1 with ReLU, that
// predicts log2(1 + nodes) of a work item's subtree.  It only orders and
// splits work; no count ever depends on it.
//
// Training runs in FP32 on these kernels (no library GEMM):
//   gemm_f32<TA, TB, EPI>  64x64 output tile per 256-thread block, 16-deep
//                          K slices staged through LDS, 4x4 outputs per
//                          thread; epilogues: bias+ReLU (forward), ReLU mask
//                          of the layer input (backward to the input), and
//                          plain store (weight gradients).  One kernel covers
//                          the forward product X W^T (TB), the input gradient
//                          dY W and the weight gradient dY^T X (TA; split
//                          over the batch rows into partial tiles that
//                          sum_partials adds in order).
//   head_fwd / head_bwd    the 128 -> 1 output layer and the MSE gradient.
//   column_sums            bias gradients (deterministic, no atomics).
//   adam_step              Adam with bias correction, one thread per weight.
//
// Inference (mlp_infer_bf16) runs the whole network per 16 items on one wave
// with v_mfma_f32_16x16x16_bf16: BF16 weights and activations, FP32
// accumulation and biases.  The product is taken transposed (neurons x items)
// so that the accumulator layout of one layer -- lane l holds rows
// 4*(l/16)+r of column l%16 -- is exactly the B-operand layout the next
// layer needs; activations never leave the registers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "mf_cost.hpp"

namespace mf {
namespace {

constexpr int kTile = 64;
constexpr int kSlice = 16;

enum Epilogue : int { kStore = 0, kBiasRelu = 1, kMaskRelu = 2 };

// C[M x N] = op(A) * op(B), op(A) is M x K, op(B) is K x N.
//   TA = false: A is row-major M x K (lda);  TA = true: A is row-major K x M.
//   TB = false: B is row-major K x N (ldb);  TB = true: B is row-major N x K.
// kBiasRelu: C = max(0, C + bias[n]).  kMaskRelu: C = C * (mask[m, n] > 0),
// mask row-major M x N with ldc (the forward activation of the layer input).
// Split K: block z covers K rows [z*kchunk, min(K, (z+1)*kchunk)) and writes
// its partial product to C + z*M*ldc (summed afterwards in z order).
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_f32(int M, int N, int K, const float *__restrict__ A, int lda,
                                                const float *__restrict__ B, int ldb, float *__restrict__ C,
                                                int ldc, const float *__restrict__ bias,
                                                const float *__restrict__ mask, int kchunk) {
    __shared__ float as[kSlice][kTile + 4];   // as[k][m]
    __shared__ float bs[kSlice][kTile + 4];   // bs[k][n]
    const int tid = threadIdx.x;
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    const int tm = (tid / 16) * 4, tn = (tid % 16) * 4;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(K, kbeg + kchunk);
    C += static_cast<size_t>(blockIdx.z) * M * ldc;
    float acc[4][4] = {};
    for (int k0 = kbeg; k0 < kend; k0 += kSlice) {
        // 64 x 16 elements of each operand, 4 per thread.
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = tid + e * 256;          // 0..1023
            // A element (m, k) of the slice.
            int mm, kk;
            if (TA) { mm = idx % kTile; kk = idx / kTile; }   // contiguous along m
            else    { kk = idx % kSlice; mm = idx / kSlice; } // contiguous along k
            const int gm = m0 + mm, gk = k0 + kk;
            float va = 0.f;
            if (gm < M && gk < kend) va = TA ? A[static_cast<size_t>(gk) * lda + gm] : A[static_cast<size_t>(gm) * lda + gk];
            as[kk][mm] = va;
            int nn, kb;
            if (TB) { kb = idx % kSlice; nn = idx / kSlice; }
            else    { nn = idx % kTile; kb = idx / kTile; }
            const int gn = n0 + nn, gkb = k0 + kb;
            float vb = 0.f;
            if (gn < N && gkb < kend) vb = TB ? B[static_cast<size_t>(gn) * ldb + gkb] : B[static_cast<size_t>(gkb) * ldb + gn];
            bs[kb][nn] = vb;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kSlice; ++k) {
            float av[4], bv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) av[r] = as[k][tm + r];
#pragma unroll
            for (int c = 0; c < 4; ++c) bv[c] = bs[k][tn + c];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] = __builtin_fmaf(av[r], bv[c], acc[r][c]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int gm = m0 + tm + r;
        if (gm >= M) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int gn = n0 + tn + c;
            if (gn >= N) continue;
            float v = acc[r][c];
            if (EPI == kBiasRelu) v = fmaxf(v + bias[gn], 0.f);
            if (EPI == kMaskRelu) v = mask[static_cast<size_t>(gm) * ldc + gn] > 0.f ? v : 0.f;
            C[static_cast<size_t>(gm) * ldc + gn] = v;
        }
    }
}

// out[b] = h[b, :] . w + w[kHidden] (one wave per row).
__global__ __launch_bounds__(256) void head_fwd(int rows, const float *__restrict__ h, const float *__restrict__ w,
                                                float *__restrict__ out) {
    const int lane = threadIdx.x & 63, row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= rows) return;
    const float *x = h + static_cast<size_t>(row) * kHidden;
    float s = x[lane] * w[lane] + x[lane + 64] * w[lane + 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) out[row] = s + w[kHidden];
}

// MSE on the batch: g[b] = 2 (pred - target) / rows; per-block loss partials.
__global__ __launch_bounds__(256) void mse_grad(int rows, const float *__restrict__ pred,
                                                const float *__restrict__ target, float *__restrict__ g,
                                                float *__restrict__ partial) {
    __shared__ float red[256];
    const int b = blockIdx.x * 256 + threadIdx.x;
    float e = 0.f;
    if (b < rows) {
        const float d = pred[b] - target[b];
        g[b] = 2.f * d / static_cast<float>(rows);
        e = d * d;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] / static_cast<float>(rows);
}

// dH[b, k] = g[b] * w[k] * (h[b, k] > 0)  (gradient into the last hidden layer).
__global__ __launch_bounds__(256) void head_bwd_input(int rows, const float *__restrict__ g,
                                                      const float *__restrict__ w, const float *__restrict__ h,
                                                      float *__restrict__ dh) {
    const size_t idx = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= static_cast<size_t>(rows) * kHidden) return;
    const int k = static_cast<int>(idx % kHidden);
    const int b = static_cast<int>(idx / kHidden);
    dh[idx] = h[idx] > 0.f ? g[b] * w[k] : 0.f;
}

// Column sums over `rows` of a row-major [rows x cols] matrix: 32 columns
// per block, 8 row groups (rows r = g mod 8) summed in row order, then the
// groups in order (deterministic).  Used for bias gradients and, with a row
// weight, for the head weights.
__global__ __launch_bounds__(256) void column_sums(int rows, int cols, const float *__restrict__ x,
                                                   const float *__restrict__ rowscale, float *__restrict__ out) {
    __shared__ float part[8][33];
    const int cl = threadIdx.x & 31, grp = threadIdx.x >> 5;
    const int c = blockIdx.x * 32 + cl;
    float s = 0.f;
    if (c < cols)
        for (int r = grp; r < rows; r += 8) {
            const float v = x[static_cast<size_t>(r) * cols + c];
            s = rowscale ? __builtin_fmaf(v, rowscale[r], s) : s + v;
        }
    part[grp][cl] = s;
    __syncthreads();
    if (grp == 0 && c < cols) {
        float t = 0.f;
#pragma unroll
        for (int g = 0; g < 8; ++g) t += part[g][cl];
        out[c] = t;
    }
}

// out[i] = sum over z of part[z * n + i], in z order.
__global__ __launch_bounds__(256) void sum_partials(int parts, int n, const float *__restrict__ part,
                                                    float *__restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < parts; ++z) s += part[static_cast<size_t>(z) * n + i];
    out[i] = s;
}

__global__ __launch_bounds__(256) void sum_into(int n, const float *__restrict__ x, float *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float s = 0.f;
        for (int i = 0; i < n; ++i) s += x[i];
        *out = s;
    }
}

__global__ __launch_bounds__(256) void adam_step(int n, float *__restrict__ w, const float *__restrict__ g,
                                                 float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                                 float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(256) void gather_rows(int rows, int cols, const uint32_t *__restrict__ idx,
                                                   const float *__restrict__ src, float *__restrict__ dst) {
    const size_t t = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= static_cast<size_t>(rows) * cols) return;
    const int r = static_cast<int>(t / cols), c = static_cast<int>(t % cols);
    dst[t] = src[static_cast<size_t>(idx[r]) * cols + c];
}

// ------------------------------------------------------------ BF16 inference

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Round-to-nearest-even FP32 -> BF16 (finite inputs).
__device__ __forceinline__ uint16_t to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = static_cast<short>(to_bf16(a));
    r[1] = static_cast<short>(to_bf16(b));
    r[2] = static_cast<short>(to_bf16(c));
    r[3] = static_cast<short>(to_bf16(d));
    return r;
}

// Weight fragments, in the order the inference kernel reads them:
// layer L, neuron tile t, k block q: 64 lanes x 4 bf16, lane l holding
// W[16t + l%16][16q + 4(l/16) + r].  Biases and the head stay FP32 (the head
// weights are rounded to BF16 values).
__global__ __launch_bounds__(64) void pack_bf16(const float *__restrict__ params, uint16_t *__restrict__ frag,
                                                float *__restrict__ small) {
    const int l = threadIdx.x;
    int fo = 0;
#pragma unroll 1
    for (int layer = 0; layer < 3; ++layer) {
        const int kin = layer == 0 ? kFeatures : kHidden;
        const float *W = params + (layer == 0 ? kOffW1 : layer == 1 ? kOffW2 : kOffW3);
        const int kq = kin / 16;
        for (int t = 0; t < kHidden / 16; ++t)
            for (int q = 0; q < kq; ++q) {
                const int row = 16 * t + (l % 16);
                for (int r = 0; r < 4; ++r) {
                    const int col = 16 * q + 4 * (l / 16) + r;
                    frag[(static_cast<size_t>(fo) * 64 + l) * 4 + r] = to_bf16(W[row * kin + col]);
                }
                ++fo;
            }
    }
    // small: b1, b2, b3 (fp32), head weights as bf16-rounded fp32, head bias.
    for (int k = l; k < kHidden; k += 64) {
        small[k] = params[kOffB1 + k];
        small[kHidden + k] = params[kOffB2 + k];
        small[2 * kHidden + k] = params[kOffB3 + k];
        small[3 * kHidden + k] = __uint_as_float(static_cast<uint32_t>(to_bf16(params[kOffW4 + k])) << 16);
    }
    if (l == 0) small[4 * kHidden] = params[kOffB4];
}

template <int KQ>
__device__ __forceinline__ void layer_mfma(const uint16_t *__restrict__ frag_lds, const bf16x4 *bin, f32x4 *acc,
                                           int lane) {
#pragma unroll
    for (int t = 0; t < kHidden / 16; ++t) {
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < KQ; ++q) {
            const bf16x4 a = *reinterpret_cast<const bf16x4 *>(frag_lds + ((t * KQ + q) * 64 + lane) * 4);
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bin[q], c, 0, 0, 0);
        }
        acc[t] = c;
    }
}

// One wave = 16 items per pass; a block (4 waves) holds all BF16 weight
// fragments in LDS (kFragHalves halves) plus the FP32 biases and head.
__global__ __launch_bounds__(256) void mlp_infer_bf16(int rows, const float *__restrict__ x,
                                                      const uint16_t *__restrict__ frag,
                                                      const float *__restrict__ small, float *__restrict__ out) {
    extern __shared__ __align__(16) uint8_t smem[];
    uint16_t *wf = reinterpret_cast<uint16_t *>(smem);
    float *sm = reinterpret_cast<float *>(smem + kFragHalves * 2);
    for (int i = threadIdx.x; i < kFragHalves / 8; i += 256)
        reinterpret_cast<uint4 *>(wf)[i] = reinterpret_cast<const uint4 *>(frag)[i];
    for (int i = threadIdx.x; i < kSmallFloats; i += 256) sm[i] = small[i];
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane % 16, quad = lane / 16;
    const uint16_t *f1 = wf;
    const uint16_t *f2 = f1 + (kHidden / 16) * (kFeatures / 16) * 256;
    const uint16_t *f3 = f2 + (kHidden / 16) * (kHidden / 16) * 256;
    const int groups = (rows + 15) / 16;
    for (int grp = blockIdx.x * 4 + wave; grp < groups; grp += gridDim.x * 4) {
        const int item = grp * 16 + col;
        const bool live = item < rows;
        // Layer 1 B operand: features k = 16q + 4*quad + r of this lane's item.
        bf16x4 b1[kFeatures / 16];
#pragma unroll
        for (int q = 0; q < kFeatures / 16; ++q) {
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (live) v = *reinterpret_cast<const float4 *>(x + static_cast<size_t>(item) * kFeatures + 16 * q + 4 * quad);
            b1[q] = pack4(v.x, v.y, v.z, v.w);
        }
        f32x4 acc[kHidden / 16];
        bf16x4 hb[kHidden / 16];
        layer_mfma<kFeatures / 16>(f1, b1, acc, lane);
#pragma unroll
        for (int t = 0; t < kHidden / 16; ++t) {
            const float *bias = sm + 16 * t + 4 * quad;
            hb[t] = pack4(fmaxf(acc[t][0] + bias[0], 0.f), fmaxf(acc[t][1] + bias[1], 0.f),
                          fmaxf(acc[t][2] + bias[2], 0.f), fmaxf(acc[t][3] + bias[3], 0.f));
        }
        layer_mfma<kHidden / 16>(f2, hb, acc, lane);
#pragma unroll
        for (int t = 0; t < kHidden / 16; ++t) {
            const float *bias = sm + kHidden + 16 * t + 4 * quad;
            hb[t] = pack4(fmaxf(acc[t][0] + bias[0], 0.f), fmaxf(acc[t][1] + bias[1], 0.f),
                          fmaxf(acc[t][2] + bias[2], 0.f), fmaxf(acc[t][3] + bias[3], 0.f));
        }
        layer_mfma<kHidden / 16>(f3, hb, acc, lane);
        // Head: BF16-rounded h3 times BF16 head weights, FP32 sums.
        float s = 0.f;
#pragma unroll
        for (int t = 0; t < kHidden / 16; ++t) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 16 * t + 4 * quad + r;
                const float h = fmaxf(acc[t][r] + sm[2 * kHidden + k], 0.f);
                const float hr = __uint_as_float(static_cast<uint32_t>(to_bf16(h)) << 16);
                s = __builtin_fmaf(hr, sm[3 * kHidden + k], s);
            }
        }
        s += __shfl_xor(s, 16, 64);
        s += __shfl_xor(s, 32, 64);
        if (quad == 0 && live) out[item] = s + sm[4 * kHidden];
    }
}

}  // namespace

// ------------------------------------------------------------ host launchers

#define MF_LAUNCH(kernel, grid, block, lds, s, ...) \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), lds, s, __VA_ARGS__)

static inline unsigned cdiv(size_t a, size_t b) { return static_cast<unsigned>((a + b - 1) / b); }

hipError_t cost_linear_fwd(int rows, int in, const float *x, const float *w, const float *b, float *y,
                           hipStream_t s) {
    MF_LAUNCH((gemm_f32<false, true, kBiasRelu>), dim3(cdiv(kHidden, kTile), cdiv(rows, kTile)), 256, 0, s, rows,
              kHidden, in, x, in, w, in, y, kHidden, b, nullptr, in);
    return hipGetLastError();
}

// dX = (dY W) * (X > 0): dY rows x 128, W 128 x in, X (the forward input) rows x in.
hipError_t cost_linear_bwd_input(int rows, int in, const float *dy, const float *w, const float *x, float *dx,
                                 hipStream_t s) {
    MF_LAUNCH((gemm_f32<false, false, kMaskRelu>), dim3(cdiv(in, kTile), cdiv(rows, kTile)), 256, 0, s, rows, in,
              kHidden, dy, kHidden, w, in, dx, in, nullptr, x, kHidden);
    return hipGetLastError();
}

// dW = dY^T X (128 x in), db = column sums of dY.  The product is split over
// the batch rows into up to kSplitK partial tiles (in `part`, kSplitK x 128 x
// in floats) that are then summed in order.
hipError_t cost_linear_bwd_weight(int rows, int in, const float *dy, const float *x, float *dw, float *db,
                                  float *part, hipStream_t s) {
    const int chunk = std::max(kSlice, ((rows + kSplitK - 1) / kSplitK + kSlice - 1) / kSlice * kSlice);
    const int parts = (rows + chunk - 1) / chunk;
    MF_LAUNCH((gemm_f32<true, false, kStore>), dim3(cdiv(in, kTile), cdiv(kHidden, kTile), parts), 256, 0, s,
              kHidden, in, rows, dy, kHidden, x, in, part, in, nullptr, nullptr, chunk);
    MF_LAUNCH(sum_partials, cdiv(static_cast<size_t>(kHidden) * in, 256), 256, 0, s, parts, kHidden * in, part, dw);
    MF_LAUNCH(column_sums, cdiv(kHidden, 32), 256, 0, s, rows, kHidden, dy, nullptr, db);
    return hipGetLastError();
}

hipError_t cost_head_fwd(int rows, const float *h, const float *w, float *out, hipStream_t s) {
    MF_LAUNCH(head_fwd, cdiv(rows, 4), 256, 0, s, rows, h, w, out);
    return hipGetLastError();
}

hipError_t cost_mse(int rows, const float *pred, const float *target, float *g, float *partial, float *loss,
                    hipStream_t s) {
    const unsigned blocks = cdiv(rows, 256);
    MF_LAUNCH(mse_grad, blocks, 256, 0, s, rows, pred, target, g, partial);
    MF_LAUNCH(sum_into, 1, 256, 0, s, static_cast<int>(blocks), partial, loss);
    return hipGetLastError();
}

// Head gradients: dw[k] = sum_b g[b] h[b,k], db = sum_b g[b]; dh = g w * (h > 0).
hipError_t cost_head_bwd(int rows, const float *g, const float *w, const float *h, float *dw, float *dh,
                         hipStream_t s) {
    MF_LAUNCH(column_sums, cdiv(kHidden, 32), 256, 0, s, rows, kHidden, h, g, dw);
    MF_LAUNCH(column_sums, 1, 256, 0, s, rows, 1, g, nullptr, dw + kHidden);
    MF_LAUNCH(head_bwd_input, cdiv(static_cast<size_t>(rows) * kHidden, 256), 256, 0, s, rows, g, w, h, dh);
    return hipGetLastError();
}

hipError_t cost_adam(int n, float *w, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                     int step, hipStream_t s) {
    const float c1 = 1.f - powf(b1, static_cast<float>(step));
    const float c2 = 1.f - powf(b2, static_cast<float>(step));
    MF_LAUNCH(adam_step, cdiv(n, 256), 256, 0, s, n, w, g, m, v, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t cost_gather(int rows, int cols, const uint32_t *idx, const float *src, float *dst, hipStream_t s) {
    MF_LAUNCH(gather_rows, cdiv(static_cast<size_t>(rows) * cols, 256), 256, 0, s, rows, cols, idx, src, dst);
    return hipGetLastError();
}

hipError_t cost_pack_bf16(const float *params, uint16_t *frag, float *small, hipStream_t s) {
    MF_LAUNCH(pack_bf16, 1, 64, 0, s, params, frag, small);
    return hipGetLastError();
}

uint32_t cost_infer_lds_bytes() { return kFragHalves * 2 + kSmallFloats * 4; }

hipError_t cost_infer_bf16(int rows, const float *x, const uint16_t *frag, const float *small, float *out,
                           int cus, hipStream_t s) {
    const unsigned groups = cdiv(rows, 16);
    const unsigned blocks = std::min<unsigned>(cdiv(groups, 4), static_cast<unsigned>(cus) * 2u);
    if (blocks == 0) return hipSuccess;
    const uint32_t lds = cost_infer_lds_bytes();
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&mlp_infer_bf16),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
    MF_LAUNCH(mlp_infer_bf16, blocks, 256, lds, s, rows, x, frag, small, out);
    return hipGetLastError();
}

}  // namespace mf
