// Discretization of spline curves into polylines on the device (reference
// src/discretize.cpp:102-281).
//
// Every segment of every curve starts as one piece (curve, segment, t_lo = 0,
// t_hi = 1) with its exact FP64 box.  Each round builds a Morton-ordered
// implicit BVH over the boxes of the live pieces; one query per piece walks it
// and, for every overlapping piece with a larger index that belongs to a
// different curve of a potentially linked pair, counts the pair and marks both
// pieces.  Marked pieces are bisected at the parameter midpoint (two new
// pieces with their boxes), the others retire.  The loop ends when a round
// finds no overlap, and fails after 55 rounds or when a new piece's box
// diagonal falls below eps * mean |q| / sqrt(3), as the reference does.  The
// retired pieces, radix-sorted by (curve segment, t_lo), give each curve's
// polyline vertices.
//
// Boxes and points use the reference's formulas (cubic bases, cancellation-
// free quadratic for the derivative roots, closed-interval overlap) in FP64
// with floating-point contraction off, so every box and every overlap decision
// is the same as the host restatement's, bit for bit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>

#include "flk_internal.h"

#pragma clang fp contract(off)

namespace flk {

namespace {

// Segment types in the order of VerifyCurves::SegmentType (2 is Bezier).
constexpr int kBSpline = 0, kCatmullRom = 1, kPolyline = 3;

struct DCurve {
  int64_t q0;    // first control point in the packed array
  int32_t n;     // control points
  int32_t type;  // segment type
  int32_t nseg;  // segments (looped: n; open: n - 1 polyline, n - 3 spline)
  int32_t seg0;  // global index of the first segment
};

struct PBox {
  double mn[3], mx[3];
};

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max

// One coordinate of a cubic segment in the power basis and its exact range
// over [t0, t1]: endpoint values and the derivative roots inside
// (reference curve.cpp:197-277; complex roots give NaN, rejected by every
// comparison).
__device__ void coord_range(int type, double a, double b, double c, double d, double t0, double t1, double &lo,
                            double &hi) {
  double k3, k2, k1, k0;
  if (type == kBSpline) {
    k3 = -(1.0 / 6.0) * a + 0.5 * b - 0.5 * c + (1.0 / 6.0) * d;
    k2 = 0.5 * a - b + 0.5 * c;
    k1 = -0.5 * a + 0.5 * c;
    k0 = (1.0 / 6.0) * a + (2.0 / 3.0) * b + (1.0 / 6.0) * c;
  } else if (type == kCatmullRom) {
    k3 = -0.5 * a + 1.5 * b - 1.5 * c + 0.5 * d;
    k2 = 1.0 * a - 2.5 * b + 2.0 * c - 0.5 * d;
    k1 = -0.5 * a + 0.5 * c;
    k0 = 1.0 * b;
  } else {
    k3 = -1.0 * a + 3.0 * b - 3.0 * c + 1.0 * d;
    k2 = 3.0 * a - 6.0 * b + 3.0 * c;
    k1 = -3.0 * a + 3.0 * b;
    k0 = 1.0 * a;
  }
  auto f = [&](double t) { return ((k3 * t + k2) * t + k1) * t + k0; };
  lo = f(t0);
  hi = f(t1);
  if (hi < lo) {
    const double s = lo;
    lo = hi;
    hi = s;
  }
  auto take = [&](double t) {
    if (t >= t0 && t <= t1) {
      const double v = f(t);
      lo = dmin(lo, v);
      hi = dmax(hi, v);
    }
  };
  const double disc = 4 * k2 * k2 - 12 * k1 * k3;
  if (disc == 0) {
    take(-2 * k2 / (6 * k3));
  } else {
    const double r = k2 >= 0 ? -2 * k2 - sqrt(disc) : -2 * k2 + sqrt(disc);
    take(r / (6 * k3));
    take(2 * k1 / r);
  }
}

// Point of segment i at parameter t (reference curve.cpp:99-131).
__device__ void curve_point(const double *__restrict__ q, const DCurve &C, int i, double t, double p[3]) {
  const int n = C.n;
  const double *a = q + 3 * (C.q0 + i), *b = q + 3 * (C.q0 + (i + 1) % n);
  if (C.type == kPolyline) {
    for (int d = 0; d < 3; ++d) p[d] = (1.0 - t) * a[d] + t * b[d];
    return;
  }
  double B[4];
  if (C.type == kBSpline) {
    B[0] = ((-(1.0 / 6) * t + (1.0 / 2)) * t - (1.0 / 2)) * t + (1.0 / 6);
    B[1] = ((1.0 / 2) * t - 1.0) * (t * t) + (2.0 / 3);
    B[2] = ((-(1.0 / 2) * t + (1.0 / 2)) * t + (1.0 / 2)) * t + (1.0 / 6);
    B[3] = (1.0 / 6) * (t * t) * t;
  } else if (C.type == kCatmullRom) {
    B[0] = ((-0.5 * t + 1.0) * t - 0.5) * t;
    B[1] = (1.5 * t - 2.5) * (t * t) + 1.0;
    B[2] = ((-1.5 * t + 2.0) * t + 0.5) * t;
    B[3] = (0.5 * t - 0.5) * (t * t);
  } else {
    B[0] = ((-1.0 * t + 3.0) * t - 3.0) * t + 1.0;
    B[1] = ((3.0 * t - 6.0) * t + 3.0) * t;
    B[2] = (-3.0 * t + 3.0) * (t * t);
    B[3] = 1.0 * t * t * t;
  }
  const double *c = q + 3 * (C.q0 + (i + 2) % n), *e = q + 3 * (C.q0 + (i + 3) % n);
  for (int d = 0; d < 3; ++d) p[d] = B[0] * a[d] + B[1] * b[d] + B[2] * c[d] + B[3] * e[d];
}

// Exact box of segment i over [tlo, thi] (reference curve.cpp:293-324).
__device__ PBox seg_box(const double *__restrict__ q, const DCurve &C, int i, double tlo, double thi) {
  PBox b;
  if (C.type == kPolyline) {
    double pa[3], pb[3];
    curve_point(q, C, i, tlo, pa);
    curve_point(q, C, i, thi, pb);
    for (int d = 0; d < 3; ++d) {
      b.mn[d] = dmin(pa[d], pb[d]);
      b.mx[d] = dmax(pa[d], pb[d]);
    }
  } else {
    const int n = C.n;
    for (int d = 0; d < 3; ++d)
      coord_range(C.type, q[3 * (C.q0 + i) + d], q[3 * (C.q0 + (i + 1) % n) + d], q[3 * (C.q0 + (i + 2) % n) + d],
                  q[3 * (C.q0 + (i + 3) % n) + d], tlo, thi, b.mn[d], b.mx[d]);
  }
  for (int d = 0; d < 3; ++d) {
    b.mn[d] -= 0.0;
    b.mx[d] += 0.0;
  }
  return b;
}

__device__ __forceinline__ double box_diag(const PBox &b) {
  const double x = b.mx[0] - b.mn[0], y = b.mx[1] - b.mn[1], z = b.mx[2] - b.mn[2];
  return sqrt(x * x + y * y + z * z);
}

__device__ __forceinline__ bool overlap(const PBox &a, const PBox &b) {
  return a.mn[0] <= b.mx[0] && a.mn[1] <= b.mx[1] && a.mn[2] <= b.mx[2] && b.mn[0] <= a.mx[0] &&
         b.mn[1] <= a.mx[1] && b.mn[2] <= a.mx[2];
}

__device__ __forceinline__ int curve_of(const DCurve *__restrict__ cs, int nc, int64_t g) {
  int lo = 0, hi = nc;  // last c with seg0 <= g
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (cs[mid].seg0 <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

// One piece per segment: (curve, segment, 0, 1), its box and diagonal.
__global__ __launch_bounds__(256) void disc_init(const double *__restrict__ q, const DCurve *__restrict__ cs,
                                                 int32_t nc, int64_t nseg, int32_t *__restrict__ pc,
                                                 int32_t *__restrict__ ps, double *__restrict__ plo,
                                                 double *__restrict__ phi, PBox *__restrict__ pbox,
                                                 double *__restrict__ diag) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nseg) return;
  const int c = curve_of(cs, nc, g);
  const DCurve C = cs[c];
  const int s = (int)(g - C.seg0);
  const PBox b = seg_box(q, C, s, 0.0, 1.0);
  pc[g] = c;
  ps[g] = s;
  plo[g] = 0.0;
  phi[g] = 1.0;
  pbox[g] = b;
  diag[g] = box_diag(b);
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ __launch_bounds__(256) void disc_morton(const PBox *__restrict__ pbox, int64_t n, double lx, double ly,
                                                   double lz, double ix, double iy, double iz,
                                                   uint32_t *__restrict__ keys, int32_t *__restrict__ ids) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PBox b = pbox[i];
  const double lo[3] = {lx, ly, lz}, inv[3] = {ix, iy, iz};
  uint32_t k[3];
  for (int d = 0; d < 3; ++d) {
    double u = (0.5 * (b.mn[d] + b.mx[d]) - lo[d]) * inv[d];
    u = fmin(fmax(u, 0.0), 1.0);  // NaN -> 0
    k[d] = (uint32_t)fmin(u * 1024.0, 1023.0);
  }
  keys[i] = (spread10(k[0]) << 2) | (spread10(k[1]) << 1) | spread10(k[2]);
  ids[i] = (int32_t)i;
}

// Heap slot L + i: the box of the i-th piece in Morton order (empty boxes pad
// to a power of two); slot k < L: union of slots 2k and 2k + 1.
__global__ __launch_bounds__(256) void disc_leaves(const PBox *__restrict__ pbox, const int32_t *__restrict__ order,
                                                   int64_t n, int64_t L, PBox *__restrict__ heap) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  PBox b;
  if (i < n) {
    b = pbox[order[i]];
  } else {
    for (int d = 0; d < 3; ++d) {
      b.mn[d] = DBL_MAX;
      b.mx[d] = -DBL_MAX;
    }
  }
  heap[L + i] = b;
}

__global__ __launch_bounds__(256) void disc_level(PBox *__restrict__ heap, int64_t first, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int64_t k = first + i;
  const PBox a = heap[2 * k], b = heap[2 * k + 1];
  PBox u;
  for (int d = 0; d < 3; ++d) {
    u.mn[d] = fmin(a.mn[d], b.mn[d]);
    u.mx[d] = fmax(a.mx[d], b.mx[d]);
  }
  heap[k] = u;
}

// Potential link between curves a < b: binary search of the sorted keys.
__device__ __forceinline__ bool linked(const uint64_t *__restrict__ keys, int64_t m, int a, int b) {
  const uint64_t key = ((uint64_t)(uint32_t)min(a, b) << 32) | (uint32_t)max(a, b);
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo < m && keys[lo] == key;
}

// One query per piece (Morton position i): every overlapping piece with a
// larger index on a different, potentially linked curve is counted once and
// both pieces are marked.  The stack holds at most one pending sibling per
// level of a tree of depth <= 40.
__global__ __launch_bounds__(256) void disc_query(const PBox *__restrict__ heap, const int32_t *__restrict__ order,
                                                  int64_t n, int64_t L, const int32_t *__restrict__ pc,
                                                  const uint64_t *__restrict__ links, int64_t nlinks,
                                                  uint8_t *__restrict__ hit, unsigned long long *__restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long found = 0;
  if (i < n) {
    const PBox qb = heap[L + i];
    const int32_t me = order[i];
    const int cme = pc[me];
    int64_t stack[48];
    int sp = 0;
    stack[sp++] = 1;
    while (sp > 0) {
      const int64_t k = stack[--sp];
      if (!overlap(heap[k], qb)) continue;
      if (k >= L) {
        const int64_t j = k - L;
        if (j < n) {
          const int32_t other = order[j];
          const int co = pc[other];
          if (other > me && co != cme && linked(links, nlinks, cme, co)) {
            ++found;
            hit[me] = 1;
            hit[other] = 1;
          }
        }
      } else if (sp + 2 <= 48) {
        stack[sp++] = 2 * k + 1;
        stack[sp++] = 2 * k;
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) found += __shfl_xor(found, off, 64);
  if ((threadIdx.x & 63) == 0 && found) atomicAdd(count, found);
}

__global__ __launch_bounds__(256) void disc_flags(const uint8_t *__restrict__ hit, int64_t n,
                                                  int32_t *__restrict__ f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = hit[i] ? 1 : 0;
}

// Marked pieces split at the midpoint into slots 2 s, 2 s + 1 of the next
// list (s = marked pieces before this one); unmarked pieces retire to slot
// r0 + i - s of the retired list.  too_small is set when a new box's
// diagonal is below the threshold.
__global__ __launch_bounds__(256) void disc_split(const double *__restrict__ q, const DCurve *__restrict__ cs,
                                                  int64_t n, const uint8_t *__restrict__ hit,
                                                  const int32_t *__restrict__ scan, const int32_t *__restrict__ pc,
                                                  const int32_t *__restrict__ ps, const double *__restrict__ plo,
                                                  const double *__restrict__ phi, int32_t *__restrict__ nc_,
                                                  int32_t *__restrict__ ns_, double *__restrict__ nlo,
                                                  double *__restrict__ nhi, PBox *__restrict__ nbox,
                                                  int64_t r0, int32_t *__restrict__ rc, int32_t *__restrict__ rs,
                                                  double *__restrict__ rlo, double *__restrict__ rhi,
                                                  double threshold, int32_t *__restrict__ too_small) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = pc[i], s = ps[i];
  const double lo = plo[i], hi = phi[i];
  const int64_t before = scan[i];
  if (hit[i]) {
    const DCurve C = cs[c];
    const double mid = 0.5 * (lo + hi);
    const PBox b1 = seg_box(q, C, s, lo, mid), b2 = seg_box(q, C, s, mid, hi);
    if (!(box_diag(b1) >= threshold) || !(box_diag(b2) >= threshold)) *too_small = 1;
    const int64_t o = 2 * before;
    nc_[o] = c;
    ns_[o] = s;
    nlo[o] = lo;
    nhi[o] = mid;
    nbox[o] = b1;
    nc_[o + 1] = c;
    ns_[o + 1] = s;
    nlo[o + 1] = mid;
    nhi[o + 1] = hi;
    nbox[o + 1] = b2;
  } else {
    const int64_t o = r0 + (i - before);
    rc[o] = c;
    rs[o] = s;
    rlo[o] = lo;
    rhi[o] = hi;
  }
}

// Sort keys of a retired piece, for two stable passes: t_lo as the integer
// t_lo * 2^R (t_lo is a dyadic fraction with at most R <= 55 bits), then the
// global segment index.  Two passes keep the order (segment, t_lo) for any
// segment count and round count the reference allows.
__global__ __launch_bounds__(256) void disc_keys(const DCurve *__restrict__ cs, const int32_t *__restrict__ rc,
                                                 const int32_t *__restrict__ rs, const double *__restrict__ rlo,
                                                 int64_t n, int R, uint64_t *__restrict__ tkey,
                                                 uint32_t *__restrict__ ids) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  tkey[i] = (uint64_t)ldexp(rlo[i], R);
  ids[i] = (uint32_t)i;
}

// Second pass's keys in the first pass's order: the global segment of piece
// perm[k].
__global__ __launch_bounds__(256) void disc_seg_keys(const DCurve *__restrict__ cs, const int32_t *__restrict__ rc,
                                                     const int32_t *__restrict__ rs, const uint32_t *__restrict__ perm,
                                                     int64_t n, uint32_t *__restrict__ gkey) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = perm[k];
  gkey[k] = (uint32_t)(cs[rc[i]].seg0 + rs[i]);
}

// Output vertex of sorted retired piece k: its start point at out[k + the
// number of open curves before its curve]; an open curve's last piece also
// writes its end point after it.
__global__ __launch_bounds__(256) void disc_points(const double *__restrict__ q, const DCurve *__restrict__ cs,
                                                   const int32_t *__restrict__ order, const int32_t *__restrict__ rc,
                                                   const int32_t *__restrict__ rs, const double *__restrict__ rlo,
                                                   const double *__restrict__ rhi, int64_t n,
                                                   const int32_t *__restrict__ open_before,
                                                   const int32_t *__restrict__ is_open, double *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t p = order[k];
  const int c = rc[p];
  const DCurve C = cs[c];
  double v[3];
  curve_point(q, C, rs[p], rlo[p], v);
  const int64_t o = k + open_before[c];
  for (int d = 0; d < 3; ++d) out[3 * o + d] = v[d];
  if (is_open[c]) {
    const bool last = k + 1 == n || rc[order[k + 1]] != c;
    if (last) {
      curve_point(q, C, rs[p], rhi[p], v);
      for (int d = 0; d < 3; ++d) out[3 * (o + 1) + d] = v[d];
    }
  }
}

unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

}  // namespace

// The output polylines become the context's curve set (device to device);
// res.xyz receives a host copy when host_copy is set.
void discretize_run(flk_ctx *ctx, const double *ctrl, const int64_t *offsets, const int32_t *types,
                    const uint8_t *looped, int32_t nc, const int32_t *pairs, int64_t npairs, DiscResult &res,
                    bool host_copy) {
  hipStream_t s = ctx->stream;
  res = DiscResult{};
  // curves and their segments
  std::vector<DCurve> cs(nc);
  int64_t nseg = 0;
  double coord_sum = 0.0;
  double bmin[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, bmax[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int32_t c = 0; c < nc; ++c) {
    const int64_t n = offsets[c + 1] - offsets[c];
    if (n < 0 || n > INT32_MAX / 2) throw Error(FLK_ERR_ARG, "bad curve offsets");
    if (types[c] < 0 || types[c] > 3) throw Error(FLK_ERR_ARG, "invalid segment type");
    const int32_t ns = looped[c] ? (int32_t)n : (int32_t)std::max<int64_t>(0, types[c] == kPolyline ? n - 1 : n - 3);
    cs[c] = DCurve{offsets[c], (int32_t)n, types[c], ns, (int32_t)nseg};
    for (int32_t k = 0; k < ns; ++k) {
      const double *p = ctrl + 3 * (offsets[c] + k);
      coord_sum += std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    }
    for (int64_t k = 0; k < n; ++k)
      for (int d = 0; d < 3; ++d) {
        bmin[d] = std::min(bmin[d], ctrl[3 * (offsets[c] + k) + d]);
        bmax[d] = std::max(bmax[d], ctrl[3 * (offsets[c] + k) + d]);
      }
    nseg += ns;
    if (nseg > INT32_MAX / 4) throw Error(FLK_ERR_ARG, "too many segments to discretize");
  }
  res.segments = nseg;
  res.out_offsets.assign(nc + 1, 0);
  if (nseg == 0) {
    try {
      curves_layout(ctx, res.out_offsets.data(), nc);
      curves_finish(ctx);
    } catch (...) {
      curves_clear(ctx);
      throw;
    }
    return;
  }
  const double eps = std::numeric_limits<double>::epsilon();
  const double threshold = eps * (coord_sum / (nseg * 1.73205080757));
  // potential links as sorted (min << 32 | max) keys
  std::vector<uint64_t> lk;
  lk.reserve(npairs);
  for (int64_t k = 0; k < npairs; ++k) {
    const int32_t a = pairs[2 * k], b = pairs[2 * k + 1];
    if (a < 0 || b < 0 || a >= nc || b >= nc) throw Error(FLK_ERR_ARG, "pair index out of range");
    if (a != b) lk.push_back(((uint64_t)(uint32_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b));
  }
  std::sort(lk.begin(), lk.end());
  lk.erase(std::unique(lk.begin(), lk.end()), lk.end());

  const int64_t nctrl = offsets[nc];
  DevBuf d_q, d_cs, d_links, d_cnt;
  d_q.reserve(sizeof(double) * 3 * std::max<int64_t>(nctrl, 1));
  d_cs.reserve(sizeof(DCurve) * nc);
  d_links.reserve(sizeof(uint64_t) * std::max<size_t>(lk.size(), 1));
  d_cnt.reserve(sizeof(unsigned long long) * 2);
  {
    DevBuf land;
    land.reserve(sizeof(float) * 3 * std::max<int64_t>(nctrl, 1));
    ingest_points(ctx, ctrl, false, nctrl, d_q.as<double>(), land.as<float>(), land.bytes);
    FLK_HIP(hipStreamSynchronize(s));  // before the landing area is freed
  }
  FLK_HIP(hipMemcpyAsync(d_cs.ptr, cs.data(), sizeof(DCurve) * nc, hipMemcpyHostToDevice, s));
  if (!lk.empty())
    FLK_HIP(hipMemcpyAsync(d_links.ptr, lk.data(), sizeof(uint64_t) * lk.size(), hipMemcpyHostToDevice, s));
  const double *q = d_q.as<double>();
  const DCurve *dcs = d_cs.as<DCurve>();
  unsigned long long *cnt = d_cnt.as<unsigned long long>();  // [0] overlaps, [1] too-small flag

  // live pieces (two generations) and retired pieces
  struct Gen {
    DevBuf c, s, lo, hi, box;
    void reserve(int64_t n) {
      c.reserve(sizeof(int32_t) * n);
      s.reserve(sizeof(int32_t) * n);
      lo.reserve(sizeof(double) * n);
      hi.reserve(sizeof(double) * n);
      box.reserve(sizeof(PBox) * n);
    }
  } gen[2], ret;
  // the retired list accumulates over rounds: grow it keeping its contents
  auto grow_kept = [&](DevBuf &b, size_t used, size_t need) {
    if (need <= b.bytes) return;
    DevBuf nb;
    nb.reserve(std::max(need, 2 * b.bytes));
    if (used) FLK_HIP(hipMemcpyAsync(nb.ptr, b.ptr, used, hipMemcpyDeviceToDevice, s));
    FLK_HIP(hipStreamSynchronize(s));
    b = std::move(nb);
  };
  auto grow_retired = [&](int64_t used, int64_t need) {
    grow_kept(ret.c, sizeof(int32_t) * used, sizeof(int32_t) * need);
    grow_kept(ret.s, sizeof(int32_t) * used, sizeof(int32_t) * need);
    grow_kept(ret.lo, sizeof(double) * used, sizeof(double) * need);
    grow_kept(ret.hi, sizeof(double) * used, sizeof(double) * need);
  };
  gen[0].reserve(nseg);
  {
    DevBuf d_diag;
    d_diag.reserve(sizeof(double) * nseg);
    hipLaunchKernelGGL(disc_init, dim3(blocks_for(nseg)), dim3(256), 0, s, q, dcs, nc, nseg, gen[0].c.as<int32_t>(),
                       gen[0].s.as<int32_t>(), gen[0].lo.as<double>(), gen[0].hi.as<double>(),
                       gen[0].box.as<PBox>(), d_diag.as<double>());
    FLK_HIP(hipGetLastError());
    // the reference's input guard, in its order (first smallest, NaN taken)
    std::vector<double> diag(nseg);
    FLK_HIP(hipMemcpyAsync(diag.data(), d_diag.ptr, sizeof(double) * nseg, hipMemcpyDeviceToHost, s));
    FLK_HIP(hipStreamSynchronize(s));
    double md = std::numeric_limits<double>::infinity();
    int64_t mg = -1;
    for (int64_t g = 0; g < nseg; ++g)
      if (!(diag[g] >= md)) {
        md = diag[g];
        mg = g;
      }
    if (!(md >= threshold)) {
      int c = 0;
      while (c + 1 < nc && cs[c + 1].seg0 <= mg) ++c;
      res.small_curve = c;
      res.small_segment = (int32_t)(mg - cs[c].seg0);
      res.small_size = md;
      throw Error(FLK_ERR_ARG, "Discretize: Smallest input segment is smaller than eps * avg coordinate. Consider "
                               "combining.");
    }
  }
  const double ext[3] = {bmax[0] - bmin[0], bmax[1] - bmin[1], bmax[2] - bmin[2]};
  const double inv[3] = {ext[0] > 0 ? 1.0 / ext[0] : 0.0, ext[1] > 0 ? 1.0 / ext[1] : 0.0,
                         ext[2] > 0 ? 1.0 / ext[2] : 0.0};
  int64_t n = nseg, nret = 0;
  int cur = 0;
  DevBuf d_keys0, d_keys1, d_ids0, d_ids1, d_heap, d_hit, d_flag, d_scan, d_tmp;
  unsigned long long host_cnt[2] = {0, 0};
  for (int round = 0; round < 55; ++round) {
    Gen &G = gen[cur];
    int64_t L = 1;
    while (L < n) L <<= 1;
    d_keys0.reserve(sizeof(uint32_t) * n);
    d_keys1.reserve(sizeof(uint32_t) * n);
    d_ids0.reserve(sizeof(int32_t) * n);
    d_ids1.reserve(sizeof(int32_t) * n);
    d_heap.reserve(sizeof(PBox) * 2 * L);
    d_hit.reserve(n);
    d_flag.reserve(sizeof(int32_t) * (n + 1));
    d_scan.reserve(sizeof(int32_t) * (n + 1));
    hipLaunchKernelGGL(disc_morton, dim3(blocks_for(n)), dim3(256), 0, s, G.box.as<PBox>(), n, bmin[0], bmin[1],
                       bmin[2], inv[0], inv[1], inv[2], d_keys0.as<uint32_t>(), d_ids0.as<int32_t>());
    FLK_HIP(hipGetLastError());
    size_t tb = 0;
    FLK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, d_keys0.as<uint32_t>(), d_keys1.as<uint32_t>(),
                                               d_ids0.as<int32_t>(), d_ids1.as<int32_t>(), (int)n, 0, 30, s));
    d_tmp.reserve(std::max<size_t>(tb, 16));
    FLK_HIP(hipcub::DeviceRadixSort::SortPairs(d_tmp.ptr, tb, d_keys0.as<uint32_t>(), d_keys1.as<uint32_t>(),
                                               d_ids0.as<int32_t>(), d_ids1.as<int32_t>(), (int)n, 0, 30, s));
    const int32_t *order = d_ids1.as<int32_t>();
    PBox *heap = d_heap.as<PBox>();
    hipLaunchKernelGGL(disc_leaves, dim3(blocks_for(L)), dim3(256), 0, s, G.box.as<PBox>(), order, n, L, heap);
    for (int64_t first = L / 2; first >= 1; first /= 2)
      hipLaunchKernelGGL(disc_level, dim3(blocks_for(first)), dim3(256), 0, s, heap, first, first);
    FLK_HIP(hipGetLastError());
    FLK_HIP(hipMemsetAsync(d_hit.ptr, 0, n, s));
    FLK_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * 2, s));
    hipLaunchKernelGGL(disc_query, dim3(blocks_for(n)), dim3(256), 0, s, heap, order, n, L, G.c.as<int32_t>(),
                       d_links.as<uint64_t>(), (int64_t)lk.size(), d_hit.as<uint8_t>(), cnt);
    FLK_HIP(hipGetLastError());
    // positions: scan of the marks (a trailing zero makes scan[n] the count)
    hipLaunchKernelGGL(disc_flags, dim3(blocks_for(n)), dim3(256), 0, s, d_hit.as<uint8_t>(), n,
                       d_flag.as<int32_t>());
    FLK_HIP(hipMemsetAsync(d_flag.as<int32_t>() + n, 0, sizeof(int32_t), s));
    tb = 0;
    FLK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, d_flag.as<int32_t>(), d_scan.as<int32_t>(), (int)(n + 1), s));
    d_tmp.reserve(std::max<size_t>(tb, 16));
    FLK_HIP(hipcub::DeviceScan::ExclusiveSum(d_tmp.ptr, tb, d_flag.as<int32_t>(), d_scan.as<int32_t>(), (int)(n + 1),
                                             s));
    int32_t nhit = 0;
    FLK_HIP(hipMemcpyAsync(&nhit, d_scan.as<int32_t>() + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FLK_HIP(hipMemcpyAsync(host_cnt, cnt, sizeof host_cnt, hipMemcpyDeviceToHost, s));
    FLK_HIP(hipStreamSynchronize(s));
    res.overlaps.push_back((int64_t)host_cnt[0]);
    // bisect the marked pieces, retire the rest
    Gen &N = gen[cur ^ 1];
    N.reserve(std::max<int64_t>(2 * (int64_t)nhit, 1));
    grow_retired(nret, std::max<int64_t>(nret + (n - nhit), 1));
    hipLaunchKernelGGL(disc_split, dim3(blocks_for(n)), dim3(256), 0, s, q, dcs, n, d_hit.as<uint8_t>(),
                       d_scan.as<int32_t>(), G.c.as<int32_t>(), G.s.as<int32_t>(), G.lo.as<double>(),
                       G.hi.as<double>(), N.c.as<int32_t>(), N.s.as<int32_t>(), N.lo.as<double>(), N.hi.as<double>(),
                       N.box.as<PBox>(), nret, ret.c.as<int32_t>(), ret.s.as<int32_t>(), ret.lo.as<double>(),
                       ret.hi.as<double>(), threshold, reinterpret_cast<int32_t *>(cnt + 1));
    FLK_HIP(hipGetLastError());
    FLK_HIP(hipMemcpyAsync(host_cnt + 1, cnt + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    FLK_HIP(hipStreamSynchronize(s));
    nret += n - nhit;
    if (host_cnt[1] != 0)
      throw Error(FLK_ERR_ARG, "Discretize: Smallest segment is smaller than eps * avg coordinate. Perhaps curves "
                               "are too close.");
    if (host_cnt[0] == 0) break;
    n = 2 * (int64_t)nhit;
    cur ^= 1;
  }
  res.rounds = (int32_t)res.overlaps.size();
  if (res.overlaps.back() > 0)
    throw Error(FLK_ERR_ARG, "Reached max recursion (curves are too close) in Discretize().");

  // sort the retired pieces by (global segment, t_lo) and evaluate the vertices
  const int R = res.rounds;
  int gbits = 1;
  while (gbits < 32 && ((int64_t)1 << gbits) < nseg) ++gbits;
  DevBuf d_k0, d_k1, d_i0, d_i1, d_i2, d_ob, d_open, d_out;
  d_k0.reserve(sizeof(uint64_t) * nret);
  d_k1.reserve(sizeof(uint64_t) * nret);
  d_i0.reserve(sizeof(uint32_t) * nret);
  d_i1.reserve(sizeof(uint32_t) * nret);
  d_i2.reserve(sizeof(uint32_t) * nret);
  d_tmp.reserve(std::max<size_t>(radix_temp_bytes(nret), 16));
  hipLaunchKernelGGL(disc_keys, dim3(blocks_for(nret)), dim3(256), 0, s, dcs, ret.c.as<int32_t>(),
                     ret.s.as<int32_t>(), ret.lo.as<double>(), nret, R, d_k0.as<uint64_t>(), d_i0.as<uint32_t>());
  FLK_HIP(hipGetLastError());
  // stable by t_lo (R bits), then stable by segment: d_i1 = the order
  radix_sort_pairs<uint64_t>(d_k0.as<uint64_t>(), d_k1.as<uint64_t>(), d_i0.as<uint32_t>(), d_i1.as<uint32_t>(),
                             d_i2.as<uint32_t>(), nret, std::max(R, 1), d_tmp.ptr, s);
  hipLaunchKernelGGL(disc_seg_keys, dim3(blocks_for(nret)), dim3(256), 0, s, dcs, ret.c.as<int32_t>(),
                     ret.s.as<int32_t>(), d_i2.as<uint32_t>(), nret, d_k0.as<uint32_t>());
  FLK_HIP(hipGetLastError());
  radix_sort_pairs<uint32_t>(d_k0.as<uint32_t>(), d_k1.as<uint32_t>(), d_i2.as<uint32_t>(), d_i0.as<uint32_t>(),
                             d_i1.as<uint32_t>(), nret, gbits, d_tmp.ptr, s);
  // per-curve retired counts from the piece ranges: pieces of curve c are the
  // sorted positions with keys in [seg0(c), seg0(c) + nseg(c)) << R
  std::vector<int32_t> per_curve(nc, 0);
  {
    std::vector<int32_t> rcurve(nret);
    FLK_HIP(hipMemcpyAsync(rcurve.data(), ret.c.ptr, sizeof(int32_t) * nret, hipMemcpyDeviceToHost, s));
    FLK_HIP(hipStreamSynchronize(s));
    for (int32_t c : rcurve) ++per_curve[c];
  }
  std::vector<int32_t> open_before(nc, 0), is_open(nc, 0);
  int32_t opens = 0;
  res.out_offsets.assign(nc + 1, 0);
  for (int32_t c = 0; c < nc; ++c) {
    open_before[c] = opens;
    is_open[c] = (!looped[c] && per_curve[c] > 0) ? 1 : 0;
    opens += is_open[c];
    res.out_offsets[c + 1] = res.out_offsets[c] + per_curve[c] + is_open[c];
  }
  res.retired = nret;
  const int64_t nout = res.out_offsets[nc];
  d_ob.reserve(sizeof(int32_t) * nc);
  d_open.reserve(sizeof(int32_t) * nc);
  d_out.reserve(sizeof(double) * 3 * std::max<int64_t>(nout, 1));
  FLK_HIP(hipMemcpyAsync(d_ob.ptr, open_before.data(), sizeof(int32_t) * nc, hipMemcpyHostToDevice, s));
  FLK_HIP(hipMemcpyAsync(d_open.ptr, is_open.data(), sizeof(int32_t) * nc, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(disc_points, dim3(blocks_for(nret)), dim3(256), 0, s, q, dcs, d_i1.as<int32_t>(),
                     ret.c.as<int32_t>(), ret.s.as<int32_t>(), ret.lo.as<double>(), ret.hi.as<double>(), nret,
                     d_ob.as<int32_t>(), d_open.as<int32_t>(), d_out.as<double>());
  FLK_HIP(hipGetLastError());
  if (host_copy) {
    res.xyz.resize(3 * nout);
    FLK_HIP(hipMemcpyAsync(res.xyz.data(), d_out.ptr, sizeof(double) * 3 * nout, hipMemcpyDeviceToHost, s));
  }
  FLK_HIP(hipStreamSynchronize(s));
  // install the polylines: the output array becomes the curve set's input
  // array (it holds 3 nout doubles, so the layout's reserve keeps it)
  std::swap(ctx->curves.raw, d_out);
  try {
    curves_layout(ctx, res.out_offsets.data(), nc);
    curves_finish(ctx);
  } catch (...) {
    curves_clear(ctx);
    throw;
  }
}

}  // namespace flk
