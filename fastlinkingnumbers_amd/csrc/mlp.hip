// Learned opening criterion: a 12-64-64-1 hard-swish MLP on matrix cores.
//
// The network predicts log10 of the far-field error of a node pair from 12
// dimensionless features (see mlp_features in bh.hip).  Parameters are FP32
// masters (W1[64x12] b1[64] W2[64x64] b2[64] w3[64] b3, 5,057 floats).  The
// activation is hard-swish, h(z) = z min(max(z/6 + 1/2, 0), 1): the
// logistic of SiLU costs two quarter-rate transcendental instructions per
// activation (2^x and a reciprocal), which bounded inference, while
// hard-swish is one packed FP16 FMA with the clamp bit and one packed
// multiply per two activations.
//
// Training: inputs normalised by the training mean/std, padded from 12 to 32
// features; every matrix product of the forward and backward pass runs on
// v_mfma_f32_16x16x32_bf16 tiles of 16 rows with FP32 accumulation, the
// biases, activations and output layer in FP32.  Inference (below) runs on
// FP16 operands, v_mfma_f32_16x16x32_f16.
//
// MFMA 16x16x32 fragments (one wave): lane l supplies A[l%16][8*(l/16)+j] and
// B[8*(l/16)+j][l%16] for j = 0..7 and receives C[4*(l/16)+i][l%16], i=0..3.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "bh_types.h"
#include "flk_internal.h"
#include "mlp_infer.h"

namespace flk {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int H = 64;       // hidden width
constexpr int KP = 32;      // padded input width
constexpr int OFF_W1 = 0, OFF_B1 = 768, OFF_W2 = 832, OFF_B2 = 4928, OFF_W3 = 4992, OFF_B3 = 5056;

// Hard-swish in FP32 (training) and its derivative: 1/2 + z/3 on (-3, 3),
// 0 at and below -3, 1 at and above 3.
__device__ __forceinline__ float hswish(float z) {
  return z * fminf(fmaxf(fmaf(z, 1.0f / 6.0f, 0.5f), 0.f), 1.f);
}
__device__ __forceinline__ float dhswish(float z) {
  return z <= -3.f ? 0.f : (z < 3.f ? fmaf(z, 1.0f / 3.0f, 0.5f) : 1.f);
}

struct WaveLds {          // per-wave staging, bf16 row-major tiles of 16 rows
  __bf16 x[16][KP];
  __bf16 h1[16][H];
  __bf16 dz2[16][H];
  __bf16 dz1[16][H];
  float y[16];
  float dy[16];
};
struct BlockLds {
  __bf16 w1[H][KP];       // W1 padded, row = output neuron
  __bf16 w2[H][H];
  float b1[H], b2[H], w3[H];
  float b3;
  float mean[FLK_MLP_NFEAT], istd[FLK_MLP_NFEAT];
  WaveLds wv[4];
};

__device__ __forceinline__ bf16x8 a_frag(const __bf16 *m, int ld, int k0) {
  const int l = threadIdx.x & 63;
  bf16x8 v;
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = m[(l % 16) * ld + k0 + 8 * (l / 16) + j];
  return v;
}

// Loads the weights into LDS (bf16 copies are made here from FP32 masters).
__device__ void load_weights(BlockLds &L, const float *__restrict__ P, const float *__restrict__ norm) {
  for (int t = threadIdx.x; t < H * KP; t += blockDim.x) {
    const int j = t / KP, i = t % KP;
    L.w1[j][i] = (__bf16)(i < FLK_MLP_NFEAT ? P[OFF_W1 + j * FLK_MLP_NFEAT + i] : 0.0f);
  }
  for (int t = threadIdx.x; t < H * H; t += blockDim.x) L.w2[t / H][t % H] = (__bf16)P[OFF_W2 + t];
  for (int t = threadIdx.x; t < H; t += blockDim.x) {
    L.b1[t] = P[OFF_B1 + t];
    L.b2[t] = P[OFF_B2 + t];
    L.w3[t] = P[OFF_W3 + t];
  }
  if (threadIdx.x == 0) L.b3 = P[OFF_B3];
  for (int t = threadIdx.x; t < FLK_MLP_NFEAT; t += blockDim.x) {
    L.mean[t] = norm[t];
    L.istd[t] = 1.0f / norm[FLK_MLP_NFEAT + t];
  }
  __syncthreads();
}

// Forward pass of 16 rows held in W.x: returns Z1, Z2 fragments (C layout,
// 4 column blocks) and writes W.h1 (bf16) and W.y.
__device__ void forward16(const BlockLds &L, WaveLds &W, f32x4 (&z1)[4], f32x4 (&z2)[4]) {
  const int l = threadIdx.x & 63;
  const bf16x8 ax = a_frag(&W.x[0][0], KP, 0);
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    bf16x8 b;
#pragma unroll
    for (int j = 0; j < 8; ++j) b[j] = L.w1[nb * 16 + l % 16][8 * (l / 16) + j];
    f32x4 acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ax, b, acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      acc[i] += L.b1[nb * 16 + l % 16];
      W.h1[4 * (l / 16) + i][nb * 16 + l % 16] = (__bf16)hswish(acc[i]);
    }
    z1[nb] = acc;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const bf16x8 ah0 = a_frag(&W.h1[0][0], H, 0), ah1 = a_frag(&W.h1[0][0], H, 32);
  float part[4] = {0, 0, 0, 0};
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    bf16x8 b0, b1;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      b0[j] = L.w2[nb * 16 + l % 16][8 * (l / 16) + j];
      b1[j] = L.w2[nb * 16 + l % 16][32 + 8 * (l / 16) + j];
    }
    f32x4 acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah0, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah1, b1, acc, 0, 0, 0);
    const float w3 = L.w3[nb * 16 + l % 16];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      acc[i] += L.b2[nb * 16 + l % 16];
      part[i] += w3 * hswish(acc[i]);
    }
    z2[nb] = acc;
  }
  // output layer: sum the 64 columns (16 lanes x 4 blocks) of each row
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) part[i] += __shfl_xor(part[i], o, 64);
  }
  if (l % 16 == 0)
#pragma unroll
    for (int i = 0; i < 4; ++i) W.y[4 * (l / 16) + i] = part[i] + L.b3;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// Stages 16 rows of raw features (rows >= n are zero) normalised into W.x.
__device__ void stage_x(const BlockLds &L, WaveLds &W, const float *__restrict__ x,
                        const int64_t *__restrict__ idx, int64_t r0, int64_t n) {
  const int l = threadIdx.x & 63;
  for (int t = l; t < 16 * KP; t += 64) {
    const int r = t / KP, i = t % KP;
    float v = 0.f;
    if (i < FLK_MLP_NFEAT && r0 + r < n) {
      const int64_t row = idx ? idx[r0 + r] : r0 + r;
      v = (x[row * FLK_MLP_NFEAT + i] - L.mean[i]) * L.istd[i];
    }
    W.x[r][i] = (__bf16)v;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(const float *__restrict__ P,
                                                          const float *__restrict__ norm,
                                                          const float *__restrict__ x, int64_t n,
                                                          float *__restrict__ y) {
  __shared__ InferLds L;
  InferRegs R;
  load_infer(P, norm, R, L);
  const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
  const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t r0 = wave * 16; r0 < n; r0 += nwaves * 16) {
    const int64_t row = r0 + c;
    f16x8 bx;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int f = 8 * g + j;
      const float v =
          (row < n && f < FLK_MLP_NFEAT) ? (x[row * FLK_MLP_NFEAT + f] - L.mean[f]) * L.istd[f] : 0.f;
      bx[j] = (_Float16)v;
    }
    const float yv = infer_tile(RegWeights{R}, L, bx);
    if (g == 0 && row < n) y[row] = yv;
  }
}

// Forward + backward of the mean squared error over rows idx[0..n) (or 0..n):
// accumulates d(loss)/d(params) into grad (FP32 atomics, one set per wave)
// and the summed squared error into loss.  inv_n = 1 / total rows.
__global__ __launch_bounds__(256) void mlp_grad_kernel(const float *__restrict__ P,
                                                       const float *__restrict__ norm,
                                                       const float *__restrict__ x,
                                                       const float *__restrict__ t,
                                                       const int64_t *__restrict__ idx, int64_t n,
                                                       float inv_n, float *__restrict__ grad,
                                                       float *__restrict__ loss) {
  __shared__ BlockLds L;
  load_weights(L, P, norm);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  WaveLds &W = L.wv[w];
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  // per-lane gradient accumulators in C layout
  f32x4 gW2[4][4], gW1[4];  // gW2[mb][nb]: rows mb*16.. (output j), cols nb*16.. (input k)
  for (int a = 0; a < 4; ++a) {
    gW1[a] = f32x4{0, 0, 0, 0};
    for (int b = 0; b < 4; ++b) gW2[a][b] = f32x4{0, 0, 0, 0};
  }
  float gb1 = 0, gb2[4] = {0, 0, 0, 0}, gw3[4] = {0, 0, 0, 0}, gb3 = 0, lsum = 0;
  float gb1v[4] = {0, 0, 0, 0};
  for (int64_t r0 = ((int64_t)blockIdx.x * 4 + w) * 16; r0 < n; r0 += nwaves * 16) {
    stage_x(L, W, x, idx, r0, n);
    f32x4 z1[4], z2[4];
    forward16(L, W, z1, z2);
    if (l < 16) {
      float dy = 0.f;
      if (r0 + l < n) {
        const int64_t row = idx ? idx[r0 + l] : r0 + l;
        const float e = W.y[l] - t[row];
        lsum += e * e;
        dy = 2.0f * e * inv_n;
      }
      W.dy[l] = dy;
      gb3 += dy;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // dZ2 = dy w3 * h'(z2); dw3 += dy * h2; db2 += dZ2
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
      const int col = nb * 16 + l % 16;
      const float w3 = L.w3[col];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 4 * (l / 16) + i;
        const float dy = W.dy[row];
        const float z = z2[nb][i];
        gw3[nb] += dy * hswish(z);
        const float d = dy * w3 * dhswish(z);
        gb2[nb] += d;
        W.dz2[row][col] = (__bf16)d;
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // dH1 = dZ2 W2 ; dZ1 = dH1 * h'(z1)
    {
      const bf16x8 a0 = a_frag(&W.dz2[0][0], H, 0), a1 = a_frag(&W.dz2[0][0], H, 32);
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        bf16x8 b0, b1;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          b0[j] = L.w2[8 * (l / 16) + j][nb * 16 + l % 16];
          b1[j] = L.w2[32 + 8 * (l / 16) + j][nb * 16 + l % 16];
        }
        f32x4 acc = {0, 0, 0, 0};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float d = acc[i] * dhswish(z1[nb][i]);
          gb1v[nb] += d;
          W.dz1[4 * (l / 16) + i][nb * 16 + l % 16] = (__bf16)d;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // weight gradients over the 16 rows (K padded to 32 with zeros):
    // dW2[j][k] += sum_r dZ2[r][j] H1[r][k];  dW1[j][i] += sum_r dZ1[r][j] X[r][i]
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      bf16x8 a2, a1;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int r = 8 * (l / 16) + j;
        a2[j] = r < 16 ? W.dz2[r][mb * 16 + l % 16] : (__bf16)0.0f;
        a1[j] = r < 16 ? W.dz1[r][mb * 16 + l % 16] : (__bf16)0.0f;
      }
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        bf16x8 b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int r = 8 * (l / 16) + j;
          b[j] = r < 16 ? W.h1[r][nb * 16 + l % 16] : (__bf16)0.0f;
        }
        gW2[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b, gW2[mb][nb], 0, 0, 0);
      }
      bf16x8 bx;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int r = 8 * (l / 16) + j;
        bx[j] = r < 16 ? W.x[r][l % 16] : (__bf16)0.0f;
      }
      gW1[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, bx, gW1[mb], 0, 0, 0);
    }
    __builtin_amdgcn_wave_barrier();
  }
  (void)gb1;
  // flush: C layout row = 4*(l/16)+i within block mb, col = l%16 within nb
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int j = mb * 16 + 4 * (l / 16) + i;
#pragma unroll
      for (int nb = 0; nb < 4; ++nb)
        atomicAdd(grad + OFF_W2 + j * H + nb * 16 + l % 16, gW2[mb][nb][i]);
      if (l % 16 < FLK_MLP_NFEAT) atomicAdd(grad + OFF_W1 + j * FLK_MLP_NFEAT + l % 16, gW1[mb][i]);
    }
  }
  // bias / w3 sums: reduce over the 4 row groups (lane bits 4,5)
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    float a = gb2[nb], b = gw3[nb], c = gb1v[nb];
    a += __shfl_xor(a, 16, 64); a += __shfl_xor(a, 32, 64);
    b += __shfl_xor(b, 16, 64); b += __shfl_xor(b, 32, 64);
    c += __shfl_xor(c, 16, 64); c += __shfl_xor(c, 32, 64);
    if (l < 16) {
      atomicAdd(grad + OFF_B2 + nb * 16 + l, a);
      atomicAdd(grad + OFF_W3 + nb * 16 + l, b);
      atomicAdd(grad + OFF_B1 + nb * 16 + l, c);
    }
  }
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    gb3 += __shfl_xor(gb3, o, 64);
    lsum += __shfl_xor(lsum, o, 64);
  }
  if (l == 0) {
    atomicAdd(grad + OFF_B3, gb3);
    atomicAdd(loss, lsum);
  }
}

__global__ void zero_kernel(float *p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.f;
}

// Adam (FP32 masters): m, v moments, bias-corrected step.
__global__ void adam_kernel(float *__restrict__ P, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, int n, float lr, float b1t, float b2t) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float b1 = 0.9f, b2 = 0.999f;
  m[t] = b1 * m[t] + (1 - b1) * g[t];
  v[t] = b2 * v[t] + (1 - b2) * g[t] * g[t];
  const float mh = m[t] / (1 - b1t), vh = v[t] / (1 - b2t);
  P[t] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

// The level kernel's LDS image of the network, one wave.
__global__ __launch_bounds__(64) void mlp_pack_image(const float *__restrict__ P, const float *__restrict__ norm,
                                                     InferImage *__restrict__ img) {
  InferRegs R;
  load_infer(P, norm, R, img->L);
  infer_w_store(R, img->W);
}

}  // namespace

// ---------------------------------------------------------------- host side
// The weights from the host masters to the device, and the level kernel's
// image of them packed from those.
void mlp_upload(flk_ctx *ctx) {
  MlpState &M = ctx->mlp;
  if (!M.loaded) throw Error(FLK_ERR_STATE, "no MLP weights loaded (flk_mlp_set_weights)");
  M.d_params.reserve(sizeof(float) * FLK_MLP_NPARAM);
  M.d_norm.reserve(sizeof(float) * 2 * FLK_MLP_NFEAT);
  FLK_HIP(hipMemcpyAsync(M.d_params.ptr, M.params.data(), sizeof(float) * FLK_MLP_NPARAM,
                         hipMemcpyHostToDevice, ctx->stream));
  FLK_HIP(hipMemcpyAsync(M.d_norm.ptr, M.norm.data(), sizeof(float) * 2 * FLK_MLP_NFEAT,
                         hipMemcpyHostToDevice, ctx->stream));
  M.d_image.reserve(sizeof(InferImage));
  hipLaunchKernelGGL(mlp_pack_image, dim3(1), dim3(64), 0, ctx->stream, M.d_params.as<float>(),
                     M.d_norm.as<float>(), M.d_image.as<InferImage>());
  FLK_HIP(hipGetLastError());
}

void mlp_forward(flk_ctx *ctx, const float *x, int64_t n, float *y) {
  if (n <= 0) return;
  mlp_upload(ctx);
  hipStream_t s = ctx->stream;
  DevBuf dx, dy;
  dx.reserve(sizeof(float) * FLK_MLP_NFEAT * n);
  dy.reserve(sizeof(float) * n);
  FLK_HIP(hipMemcpyAsync(dx.ptr, x, sizeof(float) * FLK_MLP_NFEAT * n, hipMemcpyHostToDevice, s));
  const int blocks = (int)std::min<int64_t>((n + 63) / 64, 8 * std::max(1, ctx->num_cus));
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks), dim3(256), 0, s, ctx->mlp.d_params.as<float>(),
                     ctx->mlp.d_norm.as<float>(), dx.as<float>(), n, dy.as<float>());
  FLK_HIP(hipGetLastError());
  FLK_HIP(hipMemcpyAsync(y, dy.ptr, sizeof(float) * n, hipMemcpyDeviceToHost, s));
  FLK_HIP(hipStreamSynchronize(s));
}

void mlp_forward_device(flk_ctx *ctx, const float *dx, int64_t n, float *dy) {
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((n + 63) / 64, 8 * std::max(1, ctx->num_cus)));
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                     ctx->mlp.d_params.as<float>(), ctx->mlp.d_norm.as<float>(), dx, n, dy);
  FLK_HIP(hipGetLastError());
}

float mlp_gradient(flk_ctx *ctx, const float *x, const float *y, int64_t n, float *grad) {
  mlp_upload(ctx);
  hipStream_t s = ctx->stream;
  DevBuf dx, dt, dg, dl;
  dx.reserve(sizeof(float) * FLK_MLP_NFEAT * std::max<int64_t>(n, 1));
  dt.reserve(sizeof(float) * std::max<int64_t>(n, 1));
  dg.reserve(sizeof(float) * FLK_MLP_NPARAM);
  dl.reserve(sizeof(float));
  if (n) {
    FLK_HIP(hipMemcpyAsync(dx.ptr, x, sizeof(float) * FLK_MLP_NFEAT * n, hipMemcpyHostToDevice, s));
    FLK_HIP(hipMemcpyAsync(dt.ptr, y, sizeof(float) * n, hipMemcpyHostToDevice, s));
  }
  hipLaunchKernelGGL(zero_kernel, dim3((FLK_MLP_NPARAM + 255) / 256), dim3(256), 0, s, dg.as<float>(),
                     FLK_MLP_NPARAM);
  hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, dl.as<float>(), 1);
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((n + 63) / 64, 8 * std::max(1, ctx->num_cus)));
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(256), 0, s, ctx->mlp.d_params.as<float>(),
                     ctx->mlp.d_norm.as<float>(), dx.as<float>(), dt.as<float>(), nullptr, n,
                     1.0f / (float)std::max<int64_t>(n, 1), dg.as<float>(), dl.as<float>());
  FLK_HIP(hipGetLastError());
  float loss = 0;
  FLK_HIP(hipMemcpyAsync(grad, dg.ptr, sizeof(float) * FLK_MLP_NPARAM, hipMemcpyDeviceToHost, s));
  FLK_HIP(hipMemcpyAsync(&loss, dl.ptr, sizeof(float), hipMemcpyDeviceToHost, s));
  FLK_HIP(hipStreamSynchronize(s));
  return loss / (float)std::max<int64_t>(n, 1);
}

float mlp_train(flk_ctx *ctx, const float *x, const float *y, int64_t n, int epochs, int batch,
                float lr, uint64_t seed) {
  if (n <= 0 || epochs <= 0 || batch <= 0) throw Error(FLK_ERR_ARG, "empty training set");
  mlp_upload(ctx);
  hipStream_t s = ctx->stream;
  DevBuf dx, dt, dg, dl, dm, dv, didx;
  dx.reserve(sizeof(float) * FLK_MLP_NFEAT * n);
  dt.reserve(sizeof(float) * n);
  dg.reserve(sizeof(float) * FLK_MLP_NPARAM);
  dm.reserve(sizeof(float) * FLK_MLP_NPARAM);
  dv.reserve(sizeof(float) * FLK_MLP_NPARAM);
  dl.reserve(sizeof(float));
  didx.reserve(sizeof(int64_t) * n);
  FLK_HIP(hipMemcpyAsync(dx.ptr, x, sizeof(float) * FLK_MLP_NFEAT * n, hipMemcpyHostToDevice, s));
  FLK_HIP(hipMemcpyAsync(dt.ptr, y, sizeof(float) * n, hipMemcpyHostToDevice, s));
  const unsigned zb = (FLK_MLP_NPARAM + 255) / 256;
  hipLaunchKernelGGL(zero_kernel, dim3(zb), dim3(256), 0, s, dm.as<float>(), FLK_MLP_NPARAM);
  hipLaunchKernelGGL(zero_kernel, dim3(zb), dim3(256), 0, s, dv.as<float>(), FLK_MLP_NPARAM);
  std::vector<int64_t> perm(n);
  for (int64_t k = 0; k < n; ++k) perm[k] = k;
  std::mt19937_64 rng(seed);
  float b1t = 1.f, b2t = 1.f, last = 0.f;
  const int blocks_max = 4 * std::max(1, ctx->num_cus);
  for (int e = 0; e < epochs; ++e) {
    for (int64_t k = n - 1; k > 0; --k) std::swap(perm[k], perm[rng() % (uint64_t)(k + 1)]);
    FLK_HIP(hipMemcpyAsync(didx.ptr, perm.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, dl.as<float>(), 1);
    for (int64_t b0 = 0; b0 < n; b0 += batch) {
      const int64_t nb = std::min<int64_t>(batch, n - b0);
      hipLaunchKernelGGL(zero_kernel, dim3(zb), dim3(256), 0, s, dg.as<float>(), FLK_MLP_NPARAM);
      const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((nb + 63) / 64, blocks_max));
      hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(256), 0, s, ctx->mlp.d_params.as<float>(),
                         ctx->mlp.d_norm.as<float>(), dx.as<float>(), dt.as<float>(),
                         didx.as<int64_t>() + b0, nb, 1.0f / (float)nb, dg.as<float>(), dl.as<float>());
      b1t *= 0.9f;
      b2t *= 0.999f;
      hipLaunchKernelGGL(adam_kernel, dim3(zb), dim3(256), 0, s, ctx->mlp.d_params.as<float>(),
                         dg.as<float>(), dm.as<float>(), dv.as<float>(), FLK_MLP_NPARAM, lr, b1t, b2t);
    }
    FLK_HIP(hipGetLastError());
    FLK_HIP(hipMemcpyAsync(&last, dl.ptr, sizeof(float), hipMemcpyDeviceToHost, s));
    FLK_HIP(hipStreamSynchronize(s));
    last /= (float)n;
  }
  FLK_HIP(hipMemcpyAsync(ctx->mlp.params.data(), ctx->mlp.d_params.ptr, sizeof(float) * FLK_MLP_NPARAM,
                         hipMemcpyDeviceToHost, s));
  FLK_HIP(hipStreamSynchronize(s));
  return last;
}

}  // namespace flk

#define FLK_GUARD(ctx, ...)                 \
  try {                                     \
    __VA_ARGS__                             \
  } catch (const flk::Error &e) {           \
    if (ctx) (ctx)->err = e.msg;            \
    return e.code;                          \
  } catch (const std::exception &e) {       \
    if (ctx) (ctx)->err = e.what();         \
    return FLK_ERR_STATE;                   \
  }

extern "C" {

int32_t flk_mlp_set_weights(flk_ctx *ctx, const float *params, const float *norm24) {
  if (!ctx || !params || !norm24) return FLK_ERR_ARG;
  for (int k = 0; k < FLK_MLP_NFEAT; ++k)
    if (!(norm24[FLK_MLP_NFEAT + k] > 0)) {
      ctx->err = "feature std must be positive";
      return FLK_ERR_ARG;
    }
  ctx->mlp.params.assign(params, params + FLK_MLP_NPARAM);
  ctx->mlp.norm.assign(norm24, norm24 + 2 * FLK_MLP_NFEAT);
  ctx->mlp.loaded = true;
  return FLK_OK;
}

int32_t flk_mlp_get_weights(const flk_ctx *ctx, float *params, float *norm24) {
  if (!ctx || !params || !norm24) return FLK_ERR_ARG;
  if (!ctx->mlp.loaded) return FLK_ERR_STATE;
  std::copy(ctx->mlp.params.begin(), ctx->mlp.params.end(), params);
  std::copy(ctx->mlp.norm.begin(), ctx->mlp.norm.end(), norm24);
  return FLK_OK;
}

int32_t flk_mlp_forward(flk_ctx *ctx, const float *x, int64_t n, float *y) {
  if (!ctx || n < 0 || (n > 0 && (!x || !y))) return FLK_ERR_ARG;
  FLK_GUARD(ctx, {
    FLK_HIP(hipSetDevice(ctx->device));
    flk::mlp_forward(ctx, x, n, y);
    return FLK_OK;
  })
}

int32_t flk_mlp_gradient(flk_ctx *ctx, const float *x, const float *y, int64_t n, float *grad,
                         float *loss) {
  if (!ctx || n < 0 || !grad || (n > 0 && (!x || !y))) return FLK_ERR_ARG;
  FLK_GUARD(ctx, {
    FLK_HIP(hipSetDevice(ctx->device));
    const float l = flk::mlp_gradient(ctx, x, y, n, grad);
    if (loss) *loss = l;
    return FLK_OK;
  })
}

int32_t flk_mlp_train(flk_ctx *ctx, const float *x, const float *y, int64_t n, int32_t epochs,
                      int32_t batch, float lr, uint64_t seed, float *final_loss) {
  if (!ctx || n <= 0 || !x || !y) return FLK_ERR_ARG;
  FLK_GUARD(ctx, {
    FLK_HIP(hipSetDevice(ctx->device));
    const float l = flk::mlp_train(ctx, x, y, n, epochs, batch, lr, seed);
    if (final_loss) *final_loss = l;
    return FLK_OK;
  })
}

int32_t flk_mlp_samples(flk_ctx *ctx, int32_t a, int32_t b, int64_t cap, uint64_t seed, float *x,
                        float *y, int64_t *count) {
  if (!ctx || !count || cap < 0 || (cap > 0 && (!x || !y))) return FLK_ERR_ARG;
  FLK_GUARD(ctx, {
    FLK_HIP(hipSetDevice(ctx->device));
    if (a < 0 || b < 0 || a >= ctx->curves.ncurves || b >= ctx->curves.ncurves)
      throw flk::Error(FLK_ERR_ARG, "curve index out of range");
    *count = flk::mlp_samples(ctx, a, b, cap, seed, x, y);
    return FLK_OK;
  })
}

}  // extern "C"
