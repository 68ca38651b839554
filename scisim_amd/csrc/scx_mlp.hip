// scx_mlp.hip -- the learned warm start: a 16-128-128-3 multilayer perceptron (ReLU)
// in BF16 with FP32 accumulation on the matrix cores, inference over the active set and
// training (forward + backward, weight gradients, Adam) on collected samples.
//
// Every product is a 16x16x16 MFMA tile (v_mfma_f32_16x16x16_bf16) computed in the
// transposed form Out^T = W^T In^T, with samples on the N side: lane l holds sample
// l % 16, and the accumulator of one layer (rows 4 (l / 16) + e, column l % 16) is
// already the B operand of the next layer's K-tile, so activations stay in registers
// between layers. Weights live in LDS in the layouts the A operands read (4 consecutive
// BF16 per lane). Weight gradients are MFMAs with the samples as the K dimension, over
// activations stored unit-major per 16-sample tile.
#include "scx_device.h"
#include "scx_internal.h"

namespace scx {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN kept quiet); BF16 -> FP32 is exact.
__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// BF16 weight copies, built from the FP32 parameters by k_mlp_pack:
//   w1t [128][16]  w1t[u][k] = W1[k][u]        (forward, layer 1 A operand)
//   w2t [128][128] w2t[v][u] = W2[u][v]        (forward, layer 2)
//   w2n [128][128] w2n[u][v] = W2[u][v]        (backward, dH1 = W2 dH2)
//   w3t [16][128]  w3t[o][v] = W3[v][o], o < 3 (forward, layer 3)
//   w3n [128][16]  w3n[v][o] = W3[v][o]        (backward, dH2 = W3 dY)
// biases FP32: b1 [128], b2 [128], b3 [16] (3 used).
constexpr int kW1T = 0, kW2T = kW1T + kMlpHid * kMlpIn, kW2N = kW2T + kMlpHid * kMlpHid,
              kW3T = kW2N + kMlpHid * kMlpHid, kW3N = kW3T + 16 * kMlpHid, kPackedBf16 = kW3N + kMlpHid * 16;
constexpr int kB1 = 0, kB2 = kMlpHid, kB3 = 2 * kMlpHid, kPackedBias = 2 * kMlpHid + 16;

__global__ void __launch_bounds__(kBlock) k_mlp_pack(const float* __restrict__ p, uint16_t* __restrict__ w,
                                                     float* __restrict__ bias) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < kMlpHid * kMlpIn) {  // W1 [k][u], k < 16
    const uint32_t k = i / kMlpHid, u = i % kMlpHid;
    w[kW1T + u * kMlpIn + k] = f2bf(p[kMlpOffW1 + i]);
  }
  if (i < kMlpHid * kMlpHid) {  // W2 [u][v]
    const uint32_t u = i / kMlpHid, v = i % kMlpHid;
    const uint16_t h = f2bf(p[kMlpOffW2 + i]);
    w[kW2T + v * kMlpHid + u] = h;
    w[kW2N + u * kMlpHid + v] = h;
  }
  if (i < kMlpHid * 16) {  // W3 [v][o], padded to 16 outputs
    const uint32_t v = i / 16, o = i % 16;
    const uint16_t h = (o < (uint32_t)kMlpOut) ? f2bf(p[kMlpOffW3 + v * kMlpOut + o]) : (uint16_t)0;
    w[kW3T + o * kMlpHid + v] = h;
    w[kW3N + v * 16 + o] = h;
  }
  if (i < kMlpHid) {
    bias[kB1 + i] = p[kMlpOffB1 + i];
    bias[kB2 + i] = p[kMlpOffB2 + i];
  }
  if (i < 16) bias[kB3 + i] = (i < (uint32_t)kMlpOut) ? p[kMlpOffB3 + i] : 0.f;
}

// A operand: rows r0 + lane % 16, columns c0 + 4 (lane / 16) + 0..3 of a row-major
// BF16 matrix with `ld` columns.
__device__ __forceinline__ bf16x4 a_frag(const uint16_t* m, int ld, int r0, int c0, int lane) {
  return *reinterpret_cast<const bf16x4*>(m + (size_t)(r0 + (lane & 15)) * ld + c0 + 4 * (lane >> 4));
}

// The forward pass of one 16-sample tile. x: the tile's B operand (features
// 4 (lane / 16) + e of sample lane % 16). h1, h2: BF16 activations (B operands of the
// next layer), p1, p2: pre-activations > 0 masks; y: outputs o = 4 (lane / 16) + e of
// sample lane % 16 (o < 3 meaningful).
struct Tile {
  bf16x4 h1[8], h2[8];
  f32x4 y;
};

__device__ __forceinline__ void forward_tile(const uint16_t* w, const float* bias, bf16x4 x, int lane, Tile& T) {
  const int r = 4 * (lane >> 4);
#pragma unroll
  for (int ut = 0; ut < 8; ++ut) {
    f32x4 acc = {0, 0, 0, 0};
    acc = mfma16(a_frag(w + kW1T, kMlpIn, ut * 16, 0, lane), x, acc);
#pragma unroll
    for (int e = 0; e < 4; ++e) T.h1[ut][e] = (short)f2bf(fmaxf(acc[e] + bias[kB1 + ut * 16 + r + e], 0.f));
  }
#pragma unroll
  for (int ut = 0; ut < 8; ++ut) {
    f32x4 acc = {0, 0, 0, 0};
#pragma unroll
    for (int kt = 0; kt < 8; ++kt) acc = mfma16(a_frag(w + kW2T, kMlpHid, ut * 16, kt * 16, lane), T.h1[kt], acc);
#pragma unroll
    for (int e = 0; e < 4; ++e) T.h2[ut][e] = (short)f2bf(fmaxf(acc[e] + bias[kB2 + ut * 16 + r + e], 0.f));
  }
  f32x4 acc = {0, 0, 0, 0};
#pragma unroll
  for (int kt = 0; kt < 8; ++kt) acc = mfma16(a_frag(w + kW3T, kMlpHid, 0, kt * 16, lane), T.h2[kt], acc);
#pragma unroll
  for (int e = 0; e < 4; ++e) T.y[e] = acc[e] + bias[kB3 + r + e];
}

// Features of sample s (row-major [n][16] FP32) as the tile's B operand.
__device__ __forceinline__ bf16x4 x_frag(const float* X, uint32_t n, uint32_t s, int lane) {
  bf16x4 x = {0, 0, 0, 0};
  if (s < n) {
    const float4 v = *reinterpret_cast<const float4*>(X + (size_t)s * kMlpIn + 4 * (lane >> 4));
    x[0] = (short)f2bf(v.x);
    x[1] = (short)f2bf(v.y);
    x[2] = (short)f2bf(v.z);
    x[3] = (short)f2bf(v.w);
  }
  return x;
}

// Weights and biases into LDS (all threads of the block).
__device__ __forceinline__ void stage_weights(const uint16_t* __restrict__ gw, const float* __restrict__ gb,
                                              uint16_t* sw, float* sb, int count) {
  const uint4* src = reinterpret_cast<const uint4*>(gw);
  uint4* dst = reinterpret_cast<uint4*>(sw);
  for (int i = threadIdx.x; i < count / 8; i += blockDim.x) dst[i] = src[i];
  for (int i = threadIdx.x; i < kPackedBias; i += blockDim.x) sb[i] = gb[i];
  __syncthreads();
}

// The perceptron predicts a correction to the transported cached impulse: the prediction
// of a sample is MLP(x) + x[kMlpBase .. kMlpBase + 2] (features 8-10), so an untrained or
// zero network reproduces the transported-cache warm start.
constexpr int kMlpBase = 8;

// Inference: Y [n][3] = MLP(X [n][16]) + X[:, 8:11]. One wave per 16-sample tile, grid-stride.
__global__ void __launch_bounds__(kBlock) k_mlp_infer(uint32_t n, const float* __restrict__ X,
                                                      const uint16_t* __restrict__ gw, const float* __restrict__ gb,
                                                      float* __restrict__ Y) {
  __shared__ __attribute__((aligned(16))) uint16_t sw[kPackedBf16];
  __shared__ float sb[kPackedBias];
  stage_weights(gw, gb, sw, sb, kPackedBf16);
  const int lane = threadIdx.x & 63;
  const uint32_t wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const uint32_t nwaves = gridDim.x * (kBlock / 64);
  const uint32_t ntiles = (n + 15) / 16;
  for (uint32_t t = wave; t < ntiles; t += nwaves) {
    const uint32_t s = t * 16 + (lane & 15);
    Tile T;
    forward_tile(sw, sb, x_frag(X, n, s, lane), lane, T);
    // outputs 0..2 of sample lane are in lanes 0..15, elements 0..2
    if (lane < 16 && s < n)
#pragma unroll
      for (int o = 0; o < kMlpOut; ++o) Y[(size_t)s * kMlpOut + o] = T.y[o] + X[(size_t)s * kMlpIn + kMlpBase + o];
  }
}

// Training, pass 1: forward and backward of each 16-sample tile of a batch (samples
// idx[b0 .. b0 + nb) of the pool, or b0 + i when idx is null). Loss = mean over samples
// and outputs of (y + x[8:11] - t)^2. Stores per tile, BF16, unit-major with the 16 samples
// contiguous: H1T, H2T [tile][128][16], G1T, G2T (gradients w.r.t. the pre-activations)
// [tile][128][16], GY [tile][16][16] (dLoss/dy, outputs padded to 16), XT [tile][16][16]
// (features), and the block's sum of squared errors.
struct TrainBufs {
  uint16_t* h1t;
  uint16_t* h2t;
  uint16_t* g1t;
  uint16_t* g2t;
  uint16_t* gy;
  uint16_t* xt;
  float* sse;  // per block
};

__global__ void __launch_bounds__(kBlock) k_mlp_fwdbwd(uint32_t nb, const uint32_t* __restrict__ idx, uint32_t b0,
                                                       const float* __restrict__ X, const float* __restrict__ Tgt,
                                                       const uint16_t* __restrict__ gw, const float* __restrict__ gb,
                                                       float loss_scale, TrainBufs B) {
  __shared__ __attribute__((aligned(16))) uint16_t sw[kPackedBf16];
  __shared__ float sb[kPackedBias];
  __shared__ float s_sse[kBlock / 64];
  stage_weights(gw, gb, sw, sb, kPackedBf16);
  const int lane = threadIdx.x & 63;
  const int r = 4 * (lane >> 4);
  const uint32_t wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const uint32_t nwaves = gridDim.x * (kBlock / 64);
  const uint32_t ntiles = (nb + 15) / 16;
  float sse = 0.f;
  for (uint32_t t = wave; t < ntiles; t += nwaves) {
    const uint32_t i = t * 16 + (lane & 15);
    const bool valid = i < nb;
    const uint32_t s = valid ? (idx ? idx[b0 + i] : b0 + i) : 0u;
    const bf16x4 x = valid ? x_frag(X, 0xFFFFFFFFu, s, lane) : bf16x4{0, 0, 0, 0};
    Tile T;
    forward_tile(sw, sb, x, lane, T);
    // dLoss/dy for outputs o = r + e < 3
    bf16x4 gy = {0, 0, 0, 0};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int o = r + e;
      if (valid && o < kMlpOut) {
        const float d = (T.y[e] + X[(size_t)s * kMlpIn + kMlpBase + o]) - Tgt[(size_t)s * kMlpOut + o];
        sse += d * d;
        gy[e] = (short)f2bf(loss_scale * 2.f * d);
      }
    }
    // dH2 = W3 dY, masked by h2 > 0
    bf16x4 g2[8];
#pragma unroll
    for (int ut = 0; ut < 8; ++ut) {
      f32x4 acc = {0, 0, 0, 0};
      acc = mfma16(a_frag(sw + kW3N, 16, ut * 16, 0, lane), gy, acc);
#pragma unroll
      for (int e = 0; e < 4; ++e) g2[ut][e] = (short)((T.h2[ut][e] != 0 && (T.h2[ut][e] & 0x8000) == 0) ? f2bf(acc[e]) : 0);
    }
    // dH1 = W2 dH2, masked by h1 > 0
    bf16x4 g1[8];
#pragma unroll
    for (int ut = 0; ut < 8; ++ut) {
      f32x4 acc = {0, 0, 0, 0};
#pragma unroll
      for (int kt = 0; kt < 8; ++kt) acc = mfma16(a_frag(sw + kW2N, kMlpHid, ut * 16, kt * 16, lane), g2[kt], acc);
#pragma unroll
      for (int e = 0; e < 4; ++e) g1[ut][e] = (short)((T.h1[ut][e] != 0 && (T.h1[ut][e] & 0x8000) == 0) ? f2bf(acc[e]) : 0);
    }
    // stores: element (unit ut*16 + r + e, sample lane % 16) of each tile block
    const size_t tb = (size_t)t * kMlpHid * 16;
    const int sl = lane & 15;
#pragma unroll
    for (int ut = 0; ut < 8; ++ut)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const size_t o = tb + (size_t)(ut * 16 + r + e) * 16 + sl;
        B.h1t[o] = (uint16_t)T.h1[ut][e];
        B.h2t[o] = (uint16_t)T.h2[ut][e];
        B.g1t[o] = (uint16_t)g1[ut][e];
        B.g2t[o] = (uint16_t)g2[ut][e];
      }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      B.gy[(size_t)t * 256 + (r + e) * 16 + sl] = (uint16_t)gy[e];
      B.xt[(size_t)t * 256 + (r + e) * 16 + sl] = (uint16_t)x[e];
    }
  }
  sse = wave_reduce_sum_f32(sse);
  if (lane == 0) s_sse[threadIdx.x >> 6] = sse;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < kBlock / 64; ++w) t += s_sse[w];
    B.sse[blockIdx.x] = t;
  }
}

// Training, pass 2: weight gradients, one wave per 16x16 output tile, K = samples:
//   dW1 [16][128]  = X^T G1     (8 tiles: columns of 16 hidden units)
//   dW2 [128][128] = H1^T G2    (64 tiles)
//   dW3 [128][3]   = H2^T GY    (8 tiles: rows of 16 hidden units)
// A operand: rows of the unit-major activations (4 consecutive samples per lane), B
// operand: the gradient tile's rows likewise (element (k = sample, n = unit)).
// Bias gradients (sums over samples of G1, G2, GY) ride along on the first tiles.
__global__ void __launch_bounds__(kBlock) k_mlp_wgrad(uint32_t ntiles, TrainBufs B, float* __restrict__ grad) {
  const int lane = threadIdx.x & 63;
  const int r = 4 * (lane >> 4);
  const uint32_t job = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (job >= 80) return;
  const uint16_t *am, *bm;
  int arow0, brow0, ald, bld;  // row offsets/strides inside one sample tile block
  if (job < 8) {               // dW1 tile (k 0..15, units job*16..)
    am = B.xt;
    ald = 256;
    arow0 = 0;
    bm = B.g1t;
    bld = kMlpHid * 16;
    brow0 = job * 16;
  } else if (job < 72) {  // dW2 tile (mt, nt)
    const int q = job - 8, mt = q / 8, nt = q % 8;
    am = B.h1t;
    ald = kMlpHid * 16;
    arow0 = mt * 16;
    bm = B.g2t;
    bld = kMlpHid * 16;
    brow0 = nt * 16;
  } else {  // dW3 tile (rows of units (job-72)*16.., outputs 0..15)
    am = B.h2t;
    ald = kMlpHid * 16;
    arow0 = (job - 72) * 16;
    bm = B.gy;
    bld = 256;
    brow0 = 0;
  }
  f32x4 acc = {0, 0, 0, 0};
  float bsum = 0.f;  // bias gradient of row brow0 + lane % 16 (first wave of each matrix)
  for (uint32_t t = 0; t < ntiles; ++t) {
    const bf16x4 a = *reinterpret_cast<const bf16x4*>(am + (size_t)t * ald + (size_t)(arow0 + (lane & 15)) * 16 + r);
    const bf16x4 b = *reinterpret_cast<const bf16x4*>(bm + (size_t)t * bld + (size_t)(brow0 + (lane & 15)) * 16 + r);
    acc = mfma16(a, b, acc);
#pragma unroll
    for (int e = 0; e < 4; ++e) bsum += bf2f((uint16_t)b[e]);
  }
  // acc: element (m = arow-local r + e, n = lane % 16)
  const int n = lane & 15;
  if (job < 8) {
#pragma unroll
    for (int e = 0; e < 4; ++e) grad[kMlpOffW1 + (r + e) * kMlpHid + job * 16 + n] = acc[e];
  } else if (job < 72) {
    const int q = job - 8, mt = q / 8, nt = q % 8;
#pragma unroll
    for (int e = 0; e < 4; ++e) grad[kMlpOffW2 + (mt * 16 + r + e) * kMlpHid + nt * 16 + n] = acc[e];
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (n < kMlpOut) grad[kMlpOffW3 + ((job - 72) * 16 + r + e) * kMlpOut + n] = acc[e];
  }
  // bias gradients: lanes with the same lane % 16 hold partial sums over their samples
  bsum += __shfl_xor(bsum, 16, 64);
  bsum += __shfl_xor(bsum, 32, 64);
  if (lane < 16) {
    if (job < 8) grad[kMlpOffB1 + job * 16 + lane] = bsum;
    else if (job >= 8 && job < 16) grad[kMlpOffB2 + (job - 8) * 16 + lane] = bsum;  // (mt 0, nt) tiles
    else if (job == 72 && lane < kMlpOut) grad[kMlpOffB3 + lane] = bsum;
  }
}

// Adam (Kingma & Ba 2015) on the FP32 parameters, bias-corrected.
__global__ void __launch_bounds__(kBlock) k_mlp_adam(float* __restrict__ p, const float* __restrict__ g,
                                                     float* __restrict__ m, float* __restrict__ v, float lr,
                                                     float b1, float b2, float eps, float c1, float c2) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= (uint32_t)kMlpParams) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
hipError_t launch_mlp_pack(scx_ctx* c) {
  const uint32_t nb = (kMlpHid * kMlpHid + kBlock - 1) / kBlock;
  k_mlp_pack<<<nb, kBlock, 0, c->stream>>>(c->d.mlp_params, c->d.mlp_w, c->d.mlp_b);
  return hipGetLastError();
}

hipError_t launch_mlp_infer(scx_ctx* c, uint32_t n, const float* X, float* Y) {
  const uint32_t ntiles = (n + 15) / 16;
  if (ntiles == 0) return hipSuccess;
  const uint32_t blocks = std::min<uint32_t>((ntiles + 3) / 4, 4u * (uint32_t)c->num_cus);
  k_mlp_infer<<<blocks, kBlock, 0, c->stream>>>(n, X, c->d.mlp_w, c->d.mlp_b, Y);
  return hipGetLastError();
}

// One batch: forward/backward over samples idx[b0 .. b0 + nb) (idx null: b0 + i) of
// (X, Tgt), gradients into d.mlp_grad, squared-error partials into d.mlp_sse.
hipError_t launch_mlp_gradients(scx_ctx* c, uint32_t nb, const uint32_t* idx, uint32_t b0, const float* X,
                                const float* Tgt) {
  DeviceBuffers& d = c->d;
  const uint32_t ntiles = (nb + 15) / 16;
  if (ntiles == 0 || ntiles > d.mlp_cap_tiles) return hipErrorInvalidValue;
  const size_t act = (size_t)d.mlp_cap_tiles * kMlpHid * 16;
  TrainBufs B{d.mlp_act, d.mlp_act + act, d.mlp_act + 2 * act, d.mlp_act + 3 * act, d.mlp_act + 4 * act,
              d.mlp_act + 4 * act + (size_t)d.mlp_cap_tiles * 256, d.mlp_sse};
  const uint32_t blocks = std::min<uint32_t>((ntiles + 3) / 4, (uint32_t)kMlpSseBlocks);
  d.mlp_sse_blocks = blocks;  // the loss partials this pass writes
  k_mlp_fwdbwd<<<blocks, kBlock, 0, c->stream>>>(nb, idx, b0, X, Tgt, d.mlp_w, d.mlp_b,
                                                  1.0f / ((float)nb * (float)kMlpOut), B);
  k_mlp_wgrad<<<20, kBlock, 0, c->stream>>>(ntiles, B, d.mlp_grad);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(scx_ctx* c, float lr, uint32_t step) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  const uint32_t nb = (kMlpParams + kBlock - 1) / kBlock;
  k_mlp_adam<<<nb, kBlock, 0, c->stream>>>(c->d.mlp_params, c->d.mlp_grad, c->d.mlp_m, c->d.mlp_v, lr, b1, b2,
                                           1e-8f, c1, c2);
  return hipGetLastError();
}

}  // namespace scx
