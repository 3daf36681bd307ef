// Inference of the learned opening criterion's network (12-64-64-1,
// hard-swish) on FP16 matrix cores: shared by the forward kernel (mlp.hip)
// and the traversal's level kernel (bh.hip), which evaluates it inline.
#pragma once
#include <hip/hip_runtime.h>

#include "flk.h"

namespace flk {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kMlpH = 64;  // hidden width
constexpr int kOffW1 = 0, kOffB1 = 768, kOffW2 = 832, kOffB2 = 4928, kOffW3 = 4992, kOffB3 = 5056;

// Inference: both layers as C = W . X^T (neurons x samples), so lane l ends
// layer 1 holding neurons nb*16 + 4*(l/16) + i of sample l%16 -- exactly the
// B fragment of layer 2 once W2's reduction index is permuted to the same
// order (k-slot 8g + j of step s <-> neuron (2s + j/4)*16 + 4g + j%4).
// The A fragments of both layers and this lane's output weights live in
// VGPRs for the whole kernel, the biases and input normalisation in LDS;
// activations never leave registers.  One tile = 16 samples = 12 MFMAs per
// wave.
//
// FP16 form: inputs, weights and activations are FP16 (v_mfma_f32_16x16x32_f16,
// FP32 accumulation, the biases entering as the accumulator's initial
// value).  A pair of layer outputs is rounded to FP16 (v_cvt_pk_f16_f32, to
// nearest even), its gate min(max(z/6 + 1/2, 0), 1) is one v_pk_fma_f16 with
// the clamp bit, the activation one v_pk_mul_f16, and the output layer one
// v_dot2c_f32_f16 per pair (FP32 sum): 2 instructions per activation of
// layer 1 and 2.5 of layer 2.
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
constexpr float kLog2e = 1.44269504088896340736f;
constexpr float kLn2 = 0.69314718055994530942f;
struct InferRegs {
  f16x8 a1[4], a2[4][2];
  f16x2 w3[4][2];  // output weights of neurons nb*16 + 4g + {0, 1} and {2, 3}
};
struct InferLds {
  float b1[kMlpH], b2[kMlpH];
  float mean[16], istd[16];    // features 12..15 pad: mean 0, istd 0
  float mean2[16], istd2[16];  // the same for log2 features
  float b3;
};

// Hard-swish of two layer outputs in FP16 (one rounding per operation).
__device__ __forceinline__ f16x2 hswish2(float a, float b) {
  const f16x2 z = __builtin_convertvector(f32x2{a, b}, f16x2);
  const f16x2 sixth = {(_Float16)(1.0f / 6.0f), (_Float16)(1.0f / 6.0f)};
  const f16x2 half = {(_Float16)0.5f, (_Float16)0.5f};
  const f16x2 gate = __builtin_elementwise_min(
      __builtin_elementwise_max(__builtin_elementwise_fma(z, sixth, half), f16x2{0, 0}), f16x2{1, 1});
  return z * gate;
}

__device__ void load_infer(const float *__restrict__ P, const float *__restrict__ norm, InferRegs &R,
                           InferLds &L) {
  const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int f = 8 * g + j;
      R.a1[nb][j] = (_Float16)(f < FLK_MLP_NFEAT ? P[kOffW1 + (nb * 16 + c) * FLK_MLP_NFEAT + f] : 0.0f);
    }
#pragma unroll
    for (int st = 0; st < 2; ++st)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int neuron = (2 * st + j / 4) * 16 + 4 * g + (j & 3);
        R.a2[nb][st][j] = (_Float16)P[kOffW2 + (nb * 16 + c) * kMlpH + neuron];
      }
#pragma unroll
    for (int q = 0; q < 2; ++q)
      R.w3[nb][q] = f16x2{(_Float16)P[kOffW3 + nb * 16 + 4 * g + 2 * q],
                          (_Float16)P[kOffW3 + nb * 16 + 4 * g + 2 * q + 1]};
  }
  for (int t = threadIdx.x; t < kMlpH; t += blockDim.x) {
    L.b1[t] = P[kOffB1 + t];
    L.b2[t] = P[kOffB2 + t];
  }
  for (int t = threadIdx.x; t < 16; t += blockDim.x) {
    L.mean[t] = t < FLK_MLP_NFEAT ? norm[t] : 0.f;
    L.istd[t] = t < FLK_MLP_NFEAT ? 1.0f / norm[FLK_MLP_NFEAT + t] : 0.f;
    // features given as log2 values (the level kernel): (log2 x - mean / ln 2) * (istd ln 2)
    L.mean2[t] = t < FLK_MLP_NFEAT ? norm[t] * kLog2e : 0.f;
    L.istd2[t] = t < FLK_MLP_NFEAT ? kLn2 / norm[FLK_MLP_NFEAT + t] : 0.f;
  }
  if (threadIdx.x == 0) L.b3 = P[kOffB3];
  __syncthreads();
}

// The same fragments in LDS, lane-major (one fragment of a wave is 1 KB of
// contiguous LDS), for a kernel that cannot spare 56 VGPRs for them.
struct InferWLds {
  f16x8 a1[4][64];
  f16x8 a2[4][2][64];
  f16x2 w3[4][2][64];
};
__device__ __forceinline__ void infer_w_store(const InferRegs &R, InferWLds &W) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    W.a1[nb][l] = R.a1[nb];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      W.a2[nb][q][l] = R.a2[nb][q];
      W.w3[nb][q][l] = R.w3[nb][q];
    }
  }
}

// The network as a level kernel keeps it in LDS -- weight fragments, biases
// and input normalisation -- packed once per upload of the weights
// (mlp_pack_image: load_infer and infer_w_store, so bit for bit what a block
// would build itself) and copied into each block's LDS with 16-byte loads.
struct InferImage {
  InferWLds W;
  InferLds L;
};
static_assert(sizeof(InferImage) % 16 == 0, "whole 16-byte words");

// Weight sources of a tile: the fragments in VGPRs, or read from LDS where
// each is used (volatile reads in the LDS address space: hoisted out of a
// loop of tiles they would occupy the 56 VGPRs again).
struct RegWeights {
  const InferRegs &R;
  __device__ f16x8 a1(int nb) const { return R.a1[nb]; }
  __device__ f16x8 a2(int nb, int q) const { return R.a2[nb][q]; }
  __device__ f16x2 w3(int nb, int q) const { return R.w3[nb][q]; }
};
typedef const volatile __attribute__((address_space(3))) InferWLds LdsW;
struct LdsWeights {
  LdsW *W;
  int l;
  __device__ f16x8 a1(int nb) const { return W->a1[nb][l]; }
  __device__ f16x8 a2(int nb, int q) const { return W->a2[nb][q][l]; }
  __device__ f16x2 w3(int nb, int q) const { return W->w3[nb][q][l]; }
};

// One tile: bx is this lane's B fragment of the normalised inputs (features
// 8g..8g+7 of sample l%16, FP16).  Returns the prediction of sample l%16 in
// every lane.
template <class Wt>
__device__ __forceinline__ float infer_tile(const Wt &R, const InferLds &L, f16x8 bx) {
  const int l = threadIdx.x & 63, g = l >> 4;
  f16x2 h[4][2];
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    const float4 bb = *reinterpret_cast<const float4 *>(&L.b1[nb * 16 + 4 * g]);
    const f32x4 z = __builtin_amdgcn_mfma_f32_16x16x32_f16(R.a1(nb), bx, f32x4{bb.x, bb.y, bb.z, bb.w}, 0, 0, 0);
    h[nb][0] = hswish2(z[0], z[1]);
    h[nb][1] = hswish2(z[2], z[3]);
  }
  // layer 2's B fragments: steps 0 and 1 hold the outputs of blocks 0, 1 and 2, 3
  const f16x8 bh0 = __builtin_shufflevector(__builtin_shufflevector(h[0][0], h[0][1], 0, 1, 2, 3),
                                            __builtin_shufflevector(h[1][0], h[1][1], 0, 1, 2, 3), 0, 1, 2,
                                            3, 4, 5, 6, 7);
  const f16x8 bh1 = __builtin_shufflevector(__builtin_shufflevector(h[2][0], h[2][1], 0, 1, 2, 3),
                                            __builtin_shufflevector(h[3][0], h[3][1], 0, 1, 2, 3), 0, 1, 2,
                                            3, 4, 5, 6, 7);
  float part = 0.f;
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    const float4 bb = *reinterpret_cast<const float4 *>(&L.b2[nb * 16 + 4 * g]);
    f32x4 z = __builtin_amdgcn_mfma_f32_16x16x32_f16(R.a2(nb, 0), bh0, f32x4{bb.x, bb.y, bb.z, bb.w}, 0, 0, 0);
    z = __builtin_amdgcn_mfma_f32_16x16x32_f16(R.a2(nb, 1), bh1, z, 0, 0, 0);
    part = __builtin_amdgcn_fdot2(hswish2(z[0], z[1]), R.w3(nb, 0), part, false);
    part = __builtin_amdgcn_fdot2(hswish2(z[2], z[3]), R.w3(nb, 1), part, false);
  }
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  return part + L.b3;
}

// NT tiles of one wave through the same weight fragments: each fragment is
// read once for all of them (read from LDS, the fragments are 224 B per lane
// per tile, more LDS time than the tile's 12 MFMAs take) and applied to every
// tile in the order infer_tile applies it, so every prediction is
// bit-identical to infer_tile's.  No tile is skipped: the NT independent
// MFMAs of a step issue back to back instead of each waiting for the last,
// and a tile of zero rows only costs its MFMAs (a sample is one output
// column, so it never touches another's).  bx[t] is tile t's B fragment;
// y[t] the prediction of tile t's sample l%16, in every lane.
template <int NT, class Wt>
__device__ __forceinline__ void infer_tiles(const Wt &R, const InferLds &L, const f16x8 *bx, float *y) {
  const int l = threadIdx.x & 63, g = l >> 4;
  f16x8 bh0[NT], bh1[NT];
  {
    f16x2 h[NT][4][2];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
      const f16x8 a = R.a1(nb);
      const float4 bb = *reinterpret_cast<const float4 *>(&L.b1[nb * 16 + 4 * g]);
      f32x4 z[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t)
        z[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, bx[t], f32x4{bb.x, bb.y, bb.z, bb.w}, 0, 0, 0);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        h[t][nb][0] = hswish2(z[t][0], z[t][1]);
        h[t][nb][1] = hswish2(z[t][2], z[t][3]);
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      bh0[t] = __builtin_shufflevector(__builtin_shufflevector(h[t][0][0], h[t][0][1], 0, 1, 2, 3),
                                       __builtin_shufflevector(h[t][1][0], h[t][1][1], 0, 1, 2, 3), 0, 1, 2, 3, 4,
                                       5, 6, 7);
      bh1[t] = __builtin_shufflevector(__builtin_shufflevector(h[t][2][0], h[t][2][1], 0, 1, 2, 3),
                                       __builtin_shufflevector(h[t][3][0], h[t][3][1], 0, 1, 2, 3), 0, 1, 2, 3, 4,
                                       5, 6, 7);
    }
  }
  float part[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) part[t] = 0.f;
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    const f16x8 a0 = R.a2(nb, 0), a1 = R.a2(nb, 1);
    const f16x2 w0 = R.w3(nb, 0), w1 = R.w3(nb, 1);
    const float4 bb = *reinterpret_cast<const float4 *>(&L.b2[nb * 16 + 4 * g]);
    f32x4 z[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
      z[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, bh0[t], f32x4{bb.x, bb.y, bb.z, bb.w}, 0, 0, 0);
#pragma unroll
    for (int t = 0; t < NT; ++t) z[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, bh1[t], z[t], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      part[t] = __builtin_amdgcn_fdot2(hswish2(z[t][0], z[t][1]), w0, part[t], false);
      part[t] = __builtin_amdgcn_fdot2(hswish2(z[t][2], z[t][3]), w1, part[t], false);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    float p = part[t];
    p += __shfl_xor(p, 16, 64);
    p += __shfl_xor(p, 32, 64);
    y[t] = p + L.b3;
  }
}

}  // namespace
}  // namespace flk
