// Curve boxes, vertex staging and the potential-link search (gfx950).
//
// Potential-link search (reference src/potential_link_search.cpp:88-127): all
// unordered curve pairs whose FP64 bounding boxes overlap as closed intervals.
// The device builds an implicit BVH over the curve boxes: boxes are sorted by
// the 30-bit Morton code of their centres, padded to a power of two, and the
// levels of a complete binary tree (heap layout, root = 1, children 2k and
// 2k+1) are unions of their children.  One thread per curve then walks the
// tree with a private stack (a wavefront of independent queries), appending
// every overlapping pair with a lower first index to a global list, which is
// radix-sorted into the reference's ascending order.  All box tests are FP64
// on the exact input boxes, so the pair set is exactly the reference's.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "flk_internal.h"

namespace flk {

namespace {

struct Box6 {
  double mn[3], mx[3];
};

__device__ __forceinline__ bool box_overlap(const Box6 &a, const Box6 &b) {
  return a.mn[0] <= b.mx[0] && a.mn[1] <= b.mx[1] && a.mn[2] <= b.mx[2] && b.mn[0] <= a.mx[0] &&
         b.mn[1] <= a.mx[1] && b.mn[2] <= a.mx[2];
}

// FP64 min/max of each curve's vertices (the polyline case of reference
// curve.cpp:331-343; an empty curve gets the empty box) in two passes: one
// block per chunk of at most kBoxChunk vertices of one curve, then one thread
// per curve folding its chunks, so a curve of millions of vertices is spread
// over the whole chip.
constexpr int64_t kBoxChunk = 16384;

__global__ __launch_bounds__(256) void curve_box_chunks(const double *__restrict__ xyz,
                                                        const int64_t *__restrict__ chunk_v,
                                                        double *__restrict__ part) {
  const int64_t c = blockIdx.x;
  const int64_t v0 = chunk_v[2 * c], v1 = chunk_v[2 * c + 1];
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int64_t v = v0 + threadIdx.x; v < v1; v += blockDim.x) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double x = xyz[3 * v + d];
      mn[d] = fmin(mn[d], x);
      mx[d] = fmax(mx[d], x);
    }
  }
  __shared__ double smn[3][4], smx[3][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    for (int o = 32; o > 0; o >>= 1) {
      mn[d] = fmin(mn[d], __shfl_xor(mn[d], o, 64));
      mx[d] = fmax(mx[d], __shfl_xor(mx[d], o, 64));
    }
    if (lane == 0) {
      smn[d][w] = mn[d];
      smx[d][w] = mx[d];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int d = 0; d < 3; ++d) {
      double lo = smn[d][0], hi = smx[d][0];
      for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
        lo = fmin(lo, smn[d][k]);
        hi = fmax(hi, smx[d][k]);
      }
      part[6 * c + d] = lo;
      part[6 * c + 3 + d] = hi;
    }
  }
}

__global__ __launch_bounds__(256) void curve_box_merge(const double *__restrict__ part,
                                                       const int64_t *__restrict__ coff, int32_t ncurves,
                                                       double *__restrict__ boxes) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncurves) return;
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int64_t k = coff[c]; k < coff[c + 1]; ++k)
    for (int d = 0; d < 3; ++d) {
      mn[d] = fmin(mn[d], part[6 * k + d]);
      mx[d] = fmax(mx[d], part[6 * k + 3 + d]);
    }
  for (int d = 0; d < 3; ++d) {
    boxes[6 * c + d] = mn[d];
    boxes[6 * c + 3 + d] = mx[d];
  }
}

// Device vertex v (curve c, local index l <= n_c) takes input vertex l, or
// vertex 0 for the closing copy l == n_c; coordinates are (x - origin) * scale.
__global__ __launch_bounds__(256) void stage_vertices(const double *__restrict__ xyz,
                                                      const int64_t *__restrict__ in_off,
                                                      const int64_t *__restrict__ dev_off,
                                                      int32_t ncurves, int64_t ndev, double ox,
                                                      double oy, double oz, double scale,
                                                      double4 *__restrict__ out) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= ndev) return;
  int lo = 0, hi = ncurves;  // last c with dev_off[c] <= v
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (dev_off[mid] <= v) lo = mid;
    else hi = mid;
  }
  const int c = lo;
  const int64_t l = v - dev_off[c];
  const int64_t n = in_off[c + 1] - in_off[c];
  if (n == 0) {
    out[v] = make_double4(0, 0, 0, 0);
    return;
  }
  const int64_t src = in_off[c] + (l < n ? l : 0);
  out[v] = make_double4((xyz[3 * src] - ox) * scale, (xyz[3 * src + 1] - oy) * scale,
                        (xyz[3 * src + 2] - oz) * scale, 0.0);
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ __launch_bounds__(256) void morton_keys(const double *__restrict__ boxes, int32_t n,
                                                   double lx, double ly, double lz, double ix,
                                                   double iy, double iz,
                                                   uint32_t *__restrict__ keys,
                                                   int32_t *__restrict__ ids) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *b = boxes + 6 * (int64_t)i;
  uint32_t k = 0;
  if (b[0] <= b[3]) {  // non-empty
    const double c[3] = {0.5 * (b[0] + b[3]), 0.5 * (b[1] + b[4]), 0.5 * (b[2] + b[5])};
    const double lo[3] = {lx, ly, lz}, inv[3] = {ix, iy, iz};
    uint32_t q[3];
    for (int d = 0; d < 3; ++d) {
      double u = (c[d] - lo[d]) * inv[d];
      u = fmin(fmax(u, 0.0), 1.0);
      q[d] = (uint32_t)fmin(u * 1024.0, 1023.0);
    }
    k = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
  }
  keys[i] = k;
  ids[i] = i;
}

// Heap slot L + i holds the box of the i-th curve in Morton order (empty
// boxes pad to a power of two).
__global__ __launch_bounds__(256) void bvh_leaves(const double *__restrict__ boxes,
                                                  const int32_t *__restrict__ order, int32_t n,
                                                  int32_t L, Box6 *__restrict__ heap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  Box6 b;
  if (i < n) {
    const double *s = boxes + 6 * (int64_t)order[i];
    for (int d = 0; d < 3; ++d) {
      b.mn[d] = s[d];
      b.mx[d] = s[3 + d];
    }
  } else {
    for (int d = 0; d < 3; ++d) {
      b.mn[d] = DBL_MAX;
      b.mx[d] = -DBL_MAX;
    }
  }
  heap[L + i] = b;
}

__global__ __launch_bounds__(256) void bvh_level(Box6 *__restrict__ heap, int32_t first,
                                                 int32_t count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int k = first + i;
  const Box6 a = heap[2 * k], b = heap[2 * k + 1];
  Box6 u;
  for (int d = 0; d < 3; ++d) {
    u.mn[d] = fmin(a.mn[d], b.mn[d]);
    u.mx[d] = fmax(a.mx[d], b.mx[d]);
  }
  heap[k] = u;
}

// One query per curve (Morton position i): stack walk of the heap, append
// (min id, max id) for every overlap whose partner has a larger curve id.
__global__ __launch_bounds__(256) void bvh_query(const Box6 *__restrict__ heap,
                                                 const int32_t *__restrict__ order, int32_t n,
                                                 int32_t L, int64_t cap,
                                                 unsigned long long *__restrict__ count,
                                                 uint64_t *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Box6 q = heap[L + i];
  const int32_t me = order[i];
  int stack[64];
  int sp = 0;
  stack[sp++] = 1;
  while (sp > 0) {
    const int k = stack[--sp];
    if (!box_overlap(heap[k], q)) continue;
    if (k >= L) {
      const int j = k - L;
      if (j < n) {
        const int32_t other = order[j];
        if (other > me) {
          const unsigned long long slot = atomicAdd(count, 1ull);
          if ((int64_t)slot < cap) out[slot] = ((uint64_t)(uint32_t)me << 32) | (uint32_t)other;
        }
      }
    } else if (sp + 2 <= 64) {
      stack[sp++] = 2 * k + 1;
      stack[sp++] = 2 * k;
    }
  }
}

}  // namespace

void curve_boxes_run(flk_ctx *ctx, std::vector<double> &boxes) {
  Curves &cv = ctx->curves;
  boxes.assign(6 * (size_t)cv.ncurves, 0.0);
  if (cv.ncurves == 0) return;
  // chunk table: (first, end) vertex of every chunk, and each curve's chunks
  std::vector<int64_t> meta(cv.ncurves + 1);
  std::vector<int64_t> chunk_v;
  for (int32_t c = 0; c < cv.ncurves; ++c) {
    meta[c] = (int64_t)chunk_v.size() / 2;
    for (int64_t v = cv.in_off[c]; v < cv.in_off[c + 1]; v += kBoxChunk) {
      chunk_v.push_back(v);
      chunk_v.push_back(std::min(v + kBoxChunk, cv.in_off[c + 1]));
    }
  }
  const int64_t nchunks = (int64_t)chunk_v.size() / 2;
  meta[cv.ncurves] = nchunks;
  meta.insert(meta.end(), chunk_v.begin(), chunk_v.end());
  ctx->d_boxes.reserve(sizeof(double) * 6 * (cv.ncurves + std::max<int64_t>(nchunks, 1)));
  ctx->d_misc.reserve(sizeof(int64_t) * meta.size());
  FLK_HIP(hipMemcpyAsync(ctx->d_misc.ptr, meta.data(), sizeof(int64_t) * meta.size(), hipMemcpyHostToDevice,
                         ctx->stream));
  double *d_boxes = ctx->d_boxes.as<double>(), *d_part = d_boxes + 6 * (size_t)cv.ncurves;
  const int64_t *d_coff = ctx->d_misc.as<int64_t>(), *d_chunk = d_coff + cv.ncurves + 1;
  if (nchunks > 0) {
    hipLaunchKernelGGL(curve_box_chunks, dim3((unsigned)nchunks), dim3(256), 0, ctx->stream,
                       cv.raw.as<double>(), d_chunk, d_part);
    FLK_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(curve_box_merge, dim3((unsigned)((cv.ncurves + 255) / 256)), dim3(256), 0, ctx->stream,
                     d_part, d_coff, cv.ncurves, d_boxes);
  FLK_HIP(hipGetLastError());
  FLK_HIP(hipMemcpyAsync(boxes.data(), d_boxes, sizeof(double) * 6 * cv.ncurves, hipMemcpyDeviceToHost,
                         ctx->stream));
  FLK_HIP(hipStreamSynchronize(ctx->stream));
}

void stage_vertices_launch(flk_ctx *ctx) {
  Curves &cv = ctx->curves;
  ctx->d_misc.reserve(sizeof(int64_t) * (cv.ncurves + 1));
  FLK_HIP(hipMemcpyAsync(ctx->d_misc.ptr, cv.in_off.data(), sizeof(int64_t) * (cv.ncurves + 1),
                         hipMemcpyHostToDevice, ctx->stream));
  FLK_HIP(hipMemcpyAsync(cv.d_off.ptr, cv.off.data(), sizeof(int64_t) * (cv.ncurves + 1),
                         hipMemcpyHostToDevice, ctx->stream));
  const int64_t blocks = (cv.nvert_dev + 255) / 256;
  hipLaunchKernelGGL(stage_vertices, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                     cv.raw.as<double>(), ctx->d_misc.as<int64_t>(), cv.d_off.as<int64_t>(),
                     cv.ncurves, cv.nvert_dev, cv.origin[0], cv.origin[1], cv.origin[2],
                     cv.scale, cv.dpts.as<double4>());
  FLK_HIP(hipGetLastError());
}

void pls_run(flk_ctx *ctx, const double *boxes, int32_t n, std::vector<int32_t> &out) {
  out.clear();
  if (n < 2) return;
  hipStream_t s = ctx->stream;
  int32_t L = 1;
  while (L < n) L <<= 1;
  // device scratch layout inside d_pls
  const size_t sz_boxes = sizeof(double) * 6 * (size_t)n;
  const size_t sz_keys = sizeof(uint32_t) * (size_t)n;
  const size_t sz_ids = sizeof(int32_t) * (size_t)n;
  const size_t sz_heap = sizeof(Box6) * 2 * (size_t)L;
  DevBuf d_b, d_k0, d_k1, d_i0, d_i1, d_heap, d_cnt, d_out, d_out2, d_tmp;
  d_b.reserve(sz_boxes);
  d_k0.reserve(sz_keys);
  d_k1.reserve(sz_keys);
  d_i0.reserve(sz_ids);
  d_i1.reserve(sz_ids);
  d_heap.reserve(sz_heap);
  d_cnt.reserve(sizeof(unsigned long long));
  struct Guard {
    DevBuf *b[10];
    ~Guard() {
      for (auto *p : b) p->release();
    }
  } guard{{&d_b, &d_k0, &d_k1, &d_i0, &d_i1, &d_heap, &d_cnt, &d_out, &d_out2, &d_tmp}};
  FLK_HIP(hipMemcpyAsync(d_b.ptr, boxes, sz_boxes, hipMemcpyHostToDevice, s));
  // model box for Morton normalisation
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int32_t i = 0; i < n; ++i)
    if (boxes[6 * i] <= boxes[6 * i + 3])
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], 0.5 * (boxes[6 * i + d] + boxes[6 * i + 3 + d]));
        hi[d] = std::max(hi[d], 0.5 * (boxes[6 * i + d] + boxes[6 * i + 3 + d]));
      }
  double inv[3];
  for (int d = 0; d < 3; ++d) inv[d] = (hi[d] > lo[d]) ? 1.0 / (hi[d] - lo[d]) : 0.0;
  if (!(lo[0] <= hi[0])) lo[0] = lo[1] = lo[2] = 0.0;
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(morton_keys, dim3(nb), dim3(256), 0, s, d_b.as<double>(), n, lo[0], lo[1],
                     lo[2], inv[0], inv[1], inv[2], d_k0.as<uint32_t>(), d_i0.as<int32_t>());
  FLK_HIP(hipGetLastError());
  size_t tmp_bytes = 0;
  FLK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, d_k0.as<uint32_t>(),
                                             d_k1.as<uint32_t>(), d_i0.as<int32_t>(),
                                             d_i1.as<int32_t>(), n, 0, 30, s));
  d_tmp.reserve(std::max<size_t>(tmp_bytes, 16));
  FLK_HIP(hipcub::DeviceRadixSort::SortPairs(d_tmp.ptr, tmp_bytes, d_k0.as<uint32_t>(),
                                             d_k1.as<uint32_t>(), d_i0.as<int32_t>(),
                                             d_i1.as<int32_t>(), n, 0, 30, s));
  const int32_t *order = d_i1.as<int32_t>();
  hipLaunchKernelGGL(bvh_leaves, dim3((unsigned)((L + 255) / 256)), dim3(256), 0, s,
                     d_b.as<double>(), order, n, L, d_heap.as<Box6>());
  FLK_HIP(hipGetLastError());
  for (int32_t first = L / 2; first >= 1; first /= 2) {
    hipLaunchKernelGGL(bvh_level, dim3((unsigned)((first + 255) / 256)), dim3(256), 0, s,
                       d_heap.as<Box6>(), first, first);
    FLK_HIP(hipGetLastError());
  }
  // query: first pass with a guessed capacity, second pass if it overflowed
  int64_t cap = std::max<int64_t>(1024, 8 * (int64_t)n);
  unsigned long long cnt = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    d_out.reserve(sizeof(uint64_t) * cap);
    FLK_HIP(hipMemsetAsync(d_cnt.ptr, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(bvh_query, dim3(nb), dim3(256), 0, s, d_heap.as<Box6>(), order, n, L, cap,
                       d_cnt.as<unsigned long long>(), d_out.as<uint64_t>());
    FLK_HIP(hipGetLastError());
    FLK_HIP(hipMemcpyAsync(&cnt, d_cnt.ptr, sizeof(cnt), hipMemcpyDeviceToHost, s));
    FLK_HIP(hipStreamSynchronize(s));
    if ((int64_t)cnt <= cap) break;
    cap = (int64_t)cnt;
  }
  const int64_t m = (int64_t)cnt;
  if (m == 0) return;
  d_out2.reserve(sizeof(uint64_t) * m);
  tmp_bytes = 0;
  FLK_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, d_out.as<uint64_t>(),
                                            d_out2.as<uint64_t>(), (int)m, 0, 64, s));
  d_tmp.reserve(std::max<size_t>(tmp_bytes, 16));
  FLK_HIP(hipcub::DeviceRadixSort::SortKeys(d_tmp.ptr, tmp_bytes, d_out.as<uint64_t>(),
                                            d_out2.as<uint64_t>(), (int)m, 0, 64, s));
  std::vector<uint64_t> h(m);
  FLK_HIP(hipMemcpyAsync(h.data(), d_out2.ptr, sizeof(uint64_t) * m, hipMemcpyDeviceToHost, s));
  FLK_HIP(hipStreamSynchronize(s));
  out.resize(2 * m);
  for (int64_t k = 0; k < m; ++k) {
    out[2 * k] = (int32_t)(h[k] >> 32);
    out[2 * k + 1] = (int32_t)(h[k] & 0xffffffffu);
  }
}

}  // namespace flk
