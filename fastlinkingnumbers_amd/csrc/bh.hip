// Two-tree Barnes–Hut on the device (gfx950).
//
// Reference: src/compute_link.cpp:294-934 (trees, moments, far field, node-pair
// traversal) and src/linking_number_certificate.cpp:102-229 (adaptive beta).
//
// Trees.  All segments of the curves in Barnes–Hut pairs are sorted once by
// (curve, 30-bit Morton code of the segment midpoint in the curve's box), a
// single radix sort.  Each curve's tree is the balanced median split of its
// Morton-ordered segments (left half floor(n/2)), down to leaf buckets of at
// most leaf_size segments; a subtree is a contiguous range of the sorted order
// (the reference's ModifiedKdBVH.h:128-131 property).  Nodes are created level
// by level with scan-assigned child slots (deterministic numbering).
//
// Moments (reference compute_link.cpp:333-389) are built bottom-up in FP64
// about each node's box centre: nodes of at most 8 segments sum their
// segments' closed forms (FP32 terms, FP64 centring), larger nodes shift their
// children's moments in FP64 (Eqs. 21-23), and the records are stored in FP32.  The centre keeps an FP32 hi/lo split so node
// separations are exact to FP32 precision of the separation itself.
//
// Traversal.  Work items (pair, node1, node2) live in level buffers; one
// kernel per level classifies every item: leaf-bucket pairs go to a leaf list
// evaluated exactly by a second kernel, pairs with |r|^2 > beta^2 (r1+r2)^2
// are evaluated with the second-order far field (reference
// compute_link.cpp:410-524, Supplemental Eqs. 29-30) and the Eq. 14 error
// estimate, the rest split the larger node (a leaf is never split) into the
// next level.  Levels are processed in chunks with a LIFO stack, so memory is
// bounded by depth x chunk whatever the frontier width.  Per-pair sums and
// error estimates are FP64; a block of 1024 items reduces its sums and queue
// slots in LDS before one global atomic each.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "bh_types.h"
#include "mlp_infer.h"
#include "flk_internal.h"

namespace flk {

// ---------------------------------------------------------------- records
struct BhRange {    // node range during construction
  int32_t lo, hi;
};
// Nodes of at most this many segments take their moments from their segments
// (no FP64 record round trip; a node reads at most 2 x 8 segments).
constexpr int32_t kDirectSegments = 8;

struct Mom64 {      // FP64 moments during construction
  double box[6];
  double c[3];
  double D[9];
  double Q[18];
};


namespace {

constexpr float kInv2PiF = 0.15915494309189533577f;
constexpr float kInv4PiF = 0.07957747154594766788f;

__device__ __forceinline__ uint64_t spread16(uint64_t v) {
  v &= 0xffff;
  v = (v | (v << 16)) & 0x0000ff0000ffULL;
  v = (v | (v << 8)) & 0x00f00f00f00fULL;
  v = (v | (v << 4)) & 0x0c30c30c30c3ULL;
  v = (v | (v << 2)) & 0x249249249249ULL;
  return v;
}

// Sort keys of the Barnes–Hut segments, one block per tile of the segmented
// sort (a tile lies within one curve, tcurve[b]): the Morton code of the
// segment midpoint in the curve's box, 10 bits per axis; the value is the
// segment's first vertex.  cbox holds per curve the box corner (3 doubles)
// and 1024 / extent per axis (0 for a flat axis).  The segmented sort keeps
// every curve's segments in its own range, and segments sharing a cell keep
// their order along the curve (the sort is stable).  The vertices are read as
// VP: the FP64 coordinates (double4), or, when every stored coordinate is
// exact in FP32, their FP32 copies (float4, half the bytes, the same values).
template <typename VP>
__global__ __launch_bounds__(256) void bh_keys(const VP *__restrict__ pts, const int64_t *__restrict__ seg_off,
                                               const int64_t *__restrict__ vtx_off, const double *__restrict__ cbox,
                                               const RadixTile *__restrict__ tiles,
                                               const int32_t *__restrict__ tcurve, uint32_t *__restrict__ keys,
                                               uint32_t *__restrict__ vals) {
  const RadixTile tl = tiles[blockIdx.x];
  const int c = tcurve[blockIdx.x];
  const int64_t v0 = vtx_off[c] - seg_off[c];
  const double lo[3] = {cbox[6 * c], cbox[6 * c + 1], cbox[6 * c + 2]};
  const double sc[3] = {cbox[6 * c + 3], cbox[6 * c + 4], cbox[6 * c + 5]};
  for (int o = threadIdx.x; o < tl.count; o += 256) {
    const int64_t g = (int64_t)tl.start + o;
    const int64_t v = v0 + g;
    const VP p = pts[v], q = pts[v + 1];
    const double m[3] = {0.5 * ((double)p.x + (double)q.x), 0.5 * ((double)p.y + (double)q.y),
                         0.5 * ((double)p.z + (double)q.z)};
    uint32_t code = 0;
    for (int d = 0; d < 3; ++d) {
      const double u = (m[d] - lo[d]) * sc[d];
      const uint64_t qd = (uint64_t)fmin(fmax(u, 0.0), 1023.0);
      code |= (uint32_t)(spread16(qd) << (2 - d));
    }
    keys[g] = code;
    vals[g] = (uint32_t)v;
  }
}

// Splits every node of one level with more than `leaf` segments (flag[k],
// written by the previous level's split, or from the host for the roots):
// children get slots 2*scan[k], 2*scan[k]+1 of the next level, and their own
// split flags in next_flag.  scan is null on a level where every node splits
// (the exclusive scan of all ones is k).
__global__ __launch_bounds__(256) void bh_split(const BhRange *__restrict__ lev, int32_t n, int32_t base,
                                                int32_t next_base, int32_t leaf, const int32_t *__restrict__ flag,
                                                const int32_t *__restrict__ scan, BhRange *__restrict__ next,
                                                int2 *__restrict__ child, int32_t *__restrict__ next_flag) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const BhRange r = lev[k];
  if (flag[k]) {
    const int s = 2 * (scan ? scan[k] : k);
    const int mid = r.lo + (r.hi - r.lo) / 2;
    next[s] = BhRange{r.lo, mid};
    next[s + 1] = BhRange{mid, r.hi};
    next_flag[s] = mid - r.lo > leaf ? 1 : 0;
    next_flag[s + 1] = r.hi - mid > leaf ? 1 : 0;
    child[base + k] = make_int2(next_base + s, next_base + s + 1);
  } else {
    child[base + k] = make_int2(-1, -1);
  }
}

__host__ __device__ __forceinline__ int qidx(int i, int j) { return sym_idx(i, j); }

// Endpoints of every segment in sorted order, gathered once after the sort
// (six values per segment, copied unchanged from the vertex array), so the
// moment kernels read each node's segments as one contiguous run.  SP is
// double2 (48 bytes per segment), or float2 (24 bytes) when every stored
// coordinate is exact in FP32 (the curve set has no FP32 lo parts): the FP64
// values read back are then the same bits, at half the traffic, and they are
// gathered from the FP32 copies of the vertices (VP float4, 16 bytes a
// vertex instead of 32).
__device__ __forceinline__ double2 mk2(double2, double x, double y) { return make_double2(x, y); }
__device__ __forceinline__ float2 mk2(float2, double x, double y) { return make_float2((float)x, (float)y); }

template <typename SP, typename VP>
__global__ __launch_bounds__(256) void bh_stage_segments(const uint32_t *__restrict__ order,
                                                         const VP *__restrict__ pts, int64_t nseg,
                                                         SP *__restrict__ sp) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const uint32_t v = order[s];
  const VP a = pts[v], b = pts[v + 1];
  sp[3 * s] = mk2(SP{}, a.x, a.y);
  sp[3 * s + 1] = mk2(SP{}, a.z, b.x);
  sp[3 * s + 2] = mk2(SP{}, b.y, b.z);
}

template <typename SP>
__device__ __forceinline__ void seg_ends(const SP *__restrict__ sp, int s, double (&pa)[3], double (&pb)[3]) {
  const SP u = sp[3 * (int64_t)s], v = sp[3 * (int64_t)s + 1], w = sp[3 * (int64_t)s + 2];
  pa[0] = u.x; pa[1] = u.y; pa[2] = v.x;
  pb[0] = v.y; pb[1] = w.x; pb[2] = w.y;
}

// f(pa, pb) for the segments [lo, lo + cnt) in order, their end points read
// kSegChunk segments at a time with every load of a chunk issued before the
// first is used (one round trip per chunk, not per segment).
constexpr int kSegChunk = 4;
template <typename SP, class F>
__device__ __forceinline__ void each_seg(const SP *__restrict__ sp, int lo, int cnt, F &&f) {
  for (int q0 = 0; q0 < cnt; q0 += kSegChunk) {
    const SP *p = sp + 3 * ((int64_t)lo + q0);
    SP v[3 * kSegChunk];
#pragma unroll
    for (int q = 0; q < kSegChunk; ++q)
      if (q0 + q < cnt) {
        v[3 * q] = p[3 * q];
        v[3 * q + 1] = p[3 * q + 1];
        v[3 * q + 2] = p[3 * q + 2];
      }
#pragma unroll
    for (int q = 0; q < kSegChunk; ++q)
      if (q0 + q < cnt) {
        double pa[3] = {(double)v[3 * q].x, (double)v[3 * q].y, (double)v[3 * q + 1].x};
        double pb[3] = {(double)v[3 * q + 1].y, (double)v[3 * q + 2].x, (double)v[3 * q + 2].y};
        f(pa, pb);
      }
  }
}

// Bounding box of the segments [r.lo, r.hi) of a small node.
template <typename SP>
__device__ void leaf_box(BhRange r, const SP *__restrict__ sp, double (&box)[6]) {
  for (int t = 0; t < 3; ++t) {
    box[t] = DBL_MAX;
    box[3 + t] = -DBL_MAX;
  }
  each_seg(sp, r.lo, r.hi - r.lo, [&](const double (&pa)[3], const double (&pb)[3]) {
    for (int d = 0; d < 3; ++d) {
      box[d] = fmin(box[d], fmin(pa[d], pb[d]));
      box[3 + d] = fmax(box[3 + d], fmax(pa[d], pb[d]));
    }
  });
}

// FP64 moments -> FP32 node records (centre hi/lo, radius, children, segment
// range, moment norms for the error estimate; c, D, Q).
__device__ __forceinline__ void node_records(const Mom64 &M, int2 ch, BhRange r, BhGeo &g, BhMom &m) {
  double ctr[3], dsq = 0;
  for (int d = 0; d < 3; ++d) {
    ctr[d] = 0.5 * (M.box[3 + d] + M.box[d]);
    const double e = M.box[3 + d] - M.box[d];
    dsq += e * e;
  }
  const float hx = (float)ctr[0], hy = (float)ctr[1], hz = (float)ctr[2];
  g.cx = hx; g.cy = hy; g.cz = hz;
  g.lx = (float)(ctr[0] - hx); g.ly = (float)(ctr[1] - hy); g.lz = (float)(ctr[2] - hz);
  g.rsq = (float)(0.25 * dsq);
  g.radius = (float)sqrt(0.25 * dsq);
  g.c0 = ch.x;
  g.c1 = ch.y;
  g.seg0 = r.lo;
  g.nseg = r.hi - r.lo;
  double dn = 0, qn = 0;
  for (int t = 0; t < 9; ++t) dn += M.D[t] * M.D[t];
  for (int kk = 0; kk < 3; ++kk)
    for (int t = 0; t < 6; ++t) {
      const double v = M.Q[6 * kk + t];
      qn += ((t == 1 || t == 2 || t == 4) ? 2.0 : 1.0) * v * v;
    }
  g.cn = (float)sqrt(M.c[0] * M.c[0] + M.c[1] * M.c[1] + M.c[2] * M.c[2]);
  g.dn = (float)(sqrt(dn) * 0.57735026918962576451);
  g.qn = (float)((1.0 / 3.0) * sqrt(qn));
  g.pad = 0.f;
  for (int t = 0; t < 3; ++t) m.c[t] = (float)M.c[t];
  for (int t = 0; t < 9; ++t) m.D[t] = (float)M.D[t];
  for (int t = 0; t < 18; ++t) m.Q[t] = (float)M.Q[t];
  m.pad[0] = m.pad[1] = 0.f;
}

// Moments and records of a node of at most `direct` segments, from its
// segments: the box and centre in FP64, each segment's tangent and centred
// midpoint rounded to FP32 from FP64, and the closed forms (reference
// compute_link.cpp:333-356: c += t, D += t (m - ctr)^T,
// Q_k += t_k (t t^T / 12 + (m - ctr)(m - ctr)^T)) accumulated in FP32
// (at most `direct` terms, the precision of the FP32 records they end in).
// Both moment kernels use it, so a node's records depend only on its own
// segments, never on which launch computed them.
template <typename SP>
__device__ __forceinline__ void direct_records(BhRange rr, int2 ch, const SP *__restrict__ sp, BhGeo &gr,
                                               BhMom &m) {
  double box[6];
  leaf_box(rr, sp, box);
  double ctr[3], dsq = 0;
  for (int d = 0; d < 3; ++d) {
    ctr[d] = 0.5 * (box[3 + d] + box[d]);
    const double e = box[3 + d] - box[d];
    dsq += e * e;
  }
  for (int t = 0; t < 3; ++t) m.c[t] = 0.f;
  for (int t = 0; t < 9; ++t) m.D[t] = 0.f;
  for (int t = 0; t < 18; ++t) m.Q[t] = 0.f;
  each_seg(sp, rr.lo, rr.hi - rr.lo, [&](const double (&pa)[3], const double (&pb)[3]) {
    float t[3], w[3];
    for (int d = 0; d < 3; ++d) {
      t[d] = (float)(pb[d] - pa[d]);
      w[d] = (float)(0.5 * (pa[d] + pb[d]) - ctr[d]);
    }
    for (int i = 0; i < 3; ++i) {
      m.c[i] += t[i];
      for (int j = 0; j < 3; ++j) m.D[3 * i + j] = fmaf(t[i], w[j], m.D[3 * i + j]);
    }
    for (int kk = 0; kk < 3; ++kk)
      for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j)
          m.Q[6 * kk + qidx(i, j)] =
              fmaf(t[kk], fmaf((1.0f / 12.0f) * t[i], t[j], w[i] * w[j]), m.Q[6 * kk + qidx(i, j)]);
  });
  m.pad[0] = m.pad[1] = 0.f;
  const float hx = (float)ctr[0], hy = (float)ctr[1], hz = (float)ctr[2];
  gr.cx = hx; gr.cy = hy; gr.cz = hz;
  gr.lx = (float)(ctr[0] - hx); gr.ly = (float)(ctr[1] - hy); gr.lz = (float)(ctr[2] - hz);
  gr.rsq = (float)(0.25 * dsq);
  gr.radius = sqrtf(gr.rsq);
  gr.c0 = ch.x;
  gr.c1 = ch.y;
  gr.seg0 = rr.lo;
  gr.nseg = rr.hi - rr.lo;
  float dn = 0.f, qn = 0.f;
  for (int t = 0; t < 9; ++t) dn = fmaf(m.D[t], m.D[t], dn);
  for (int kk = 0; kk < 3; ++kk)
    for (int t = 0; t < 6; ++t) {
      const float v = m.Q[6 * kk + t];
      qn = fmaf((t == 1 || t == 2 || t == 4) ? 2.f * v : v, v, qn);
    }
  gr.cn = sqrtf(fmaf(m.c[0], m.c[0], fmaf(m.c[1], m.c[1], m.c[2] * m.c[2])));
  gr.dn = sqrtf(dn) * 0.57735026918962576451f;
  gr.qn = (1.0f / 3.0f) * sqrtf(qn);
  gr.pad = 0.f;
  if (gr.nseg == 1) {  // a one-segment leaf: its tangent instead of a BhMom record
    gr.c0 = kLeafSegment;
    gr.c1 = __builtin_bit_cast(int32_t, m.c[0]);
    gr.pad = m.c[1];
    gr.dn = m.c[2];
  }
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// Stores the FP32 records of a wave's nw consecutive nodes (lane l holds
// node l's; geo and mom point at the wave's first node) through the wave's
// LDS buffer sw (at least 256 float4), so each store instruction writes 1 KB
// of contiguous memory.  One-segment leaves write no BhMom record.
__device__ __forceinline__ void wave_store_records(const BhGeo &gr, const BhMom &m, BhGeo *__restrict__ geo,
                                                   BhMom *__restrict__ mom, int nw, float4 *sw) {
  const int lane = threadIdx.x & 63;
  const float4 *g4 = reinterpret_cast<const float4 *>(&gr);
#pragma unroll
  for (int i = 0; i < 4; ++i) sw[4 * lane + i] = g4[i];
  wave_lds_sync();
  float4 *gdst = reinterpret_cast<float4 *>(geo);
  for (int c = lane; c < 4 * nw; c += 64) gdst[c] = sw[c];
  wave_lds_sync();
  const unsigned long long keep = __ballot(lane < nw && !is_leaf_segment(gr));
  const float4 *m4 = reinterpret_cast<const float4 *>(&m);
  float4 *mdst = reinterpret_cast<float4 *>(mom);
#pragma unroll
  for (int h = 0; h < 2; ++h) {  // records of lanes 32h .. 32h + 31
    if (((keep >> (32 * h)) & 0xffffffffull) == 0) continue;  // wave-uniform
    if ((lane >> 5) == h)
#pragma unroll
      for (int i = 0; i < 8; ++i) sw[8 * (lane & 31) + i] = m4[i];
    wave_lds_sync();
    const int lim = 8 * min(32, nw - 32 * h);
    for (int c = lane; c < lim; c += 64)
      if ((keep >> (32 * h + c / 8)) & 1ull) mdst[256 * h + c] = sw[c];
    wave_lds_sync();
  }
}

// The same for FP64 records, kept only by the lanes in `has` (others skip
// theirs): 16 lanes' records (288 float4) per pass through sw.
__device__ __forceinline__ void wave_store_m64(const Mom64 &M, unsigned long long has, Mom64 *__restrict__ dst,
                                               int nw, float4 *sw) {
  constexpr int kChunks = (int)(sizeof(Mom64) / 16);  // 18
  static_assert(sizeof(Mom64) % 16 == 0, "whole 16-byte chunks");
  const int lane = threadIdx.x & 63;
  const float4 *src = reinterpret_cast<const float4 *>(&M);
  float4 *d = reinterpret_cast<float4 *>(dst);
#pragma unroll
  for (int h = 0; h < 4; ++h) {  // lanes 16h .. 16h + 15
    if ((has >> (16 * h)) & 0xffffull) {
      if ((lane >> 4) == h && ((has >> lane) & 1ull))
#pragma unroll
        for (int i = 0; i < kChunks; ++i) sw[kChunks * (lane & 15) + i] = src[i];
      wave_lds_sync();
      for (int c = lane; c < kChunks * 16; c += 64) {
        const int owner = 16 * h + c / kChunks;
        if (owner < nw && ((has >> owner) & 1ull)) d[kChunks * 16 * h + c] = sw[c];
      }
      wave_lds_sync();
    }
  }
}

// The bottom levels, where every node has at most `direct` segments, in one
// launch over all of them; registers stay low enough for full occupancy.
// A wave's 64 consecutive records leave through a 4 KB LDS buffer, so every
// store instruction writes 1 KB of contiguous memory instead of one 16-byte
// piece of each lane's own record.
template <typename SP>
__global__ __launch_bounds__(256) void bh_moments_direct(const BhRange *__restrict__ range, int32_t base,
                                                         int32_t n, const int2 *__restrict__ child,
                                                         const SP *__restrict__ sp, BhGeo *__restrict__ geo,
                                                         BhMom *__restrict__ mom) {
  __shared__ float4 stage[4][256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int k0 = blockIdx.x * blockDim.x + w * 64;  // this wave's first node past base
  if (k0 >= n) return;                                // whole wave: no LDS exchange to join
  const int k = k0 + lane, nw = min(64, n - k0);
  BhGeo gr;
  BhMom m;
  if (k < n) direct_records(range[base + k], child[base + k], sp, gr, m);
  wave_store_records(gr, m, geo + base + k0, mom + base + k0, nw, stage[w]);
}

// FP64 moments of a node above `direct` segments from its two children.
template <typename SP>
__device__ void node_moments(BhRange rr, int2 ch, int32_t direct, const BhRange *__restrict__ range,
                             const SP *__restrict__ sp, const Mom64 *__restrict__ m64,
                             const BhMom *__restrict__ mom, Mom64 &M) {
  for (int t = 0; t < 3; ++t) M.c[t] = 0;
  for (int t = 0; t < 9; ++t) M.D[t] = 0;
  for (int t = 0; t < 18; ++t) M.Q[t] = 0;
  {
    const BhRange rc[2] = {range[ch.x], range[ch.y]};
    const int cid[2] = {ch.x, ch.y};
    double cb[2][6];
    for (int side = 0; side < 2; ++side) {
      if (rc[side].hi - rc[side].lo <= direct) leaf_box(rc[side], sp, cb[side]);
      else
        for (int t = 0; t < 6; ++t) cb[side][t] = m64[cid[side]].box[t];
    }
    for (int d = 0; d < 3; ++d) {
      M.box[d] = fmin(cb[0][d], cb[1][d]);
      M.box[3 + d] = fmax(cb[0][3 + d], cb[1][3 + d]);
    }
    double ctr[3];
    for (int d = 0; d < 3; ++d) ctr[d] = 0.5 * (M.box[3 + d] + M.box[d]);
    // a child's moments about its own box centre (an FP64 record, or the FP32
    // record of a child of at most `direct` segments, written by an earlier
    // launch) shifted to this node's centre
    auto shift_add = [&](const auto &C, const double (&cbox)[6]) {
      double r[3];
      for (int d = 0; d < 3; ++d) r[d] = 0.5 * (cbox[3 + d] + cbox[d]) - ctr[d];
      double cc[3];
      for (int i = 0; i < 3; ++i) cc[i] = C.c[i];
      for (int kk = 0; kk < 3; ++kk) {
        double dk[3];
        for (int j = 0; j < 3; ++j) dk[j] = C.D[3 * kk + j];
        for (int i = 0; i < 3; ++i)
          for (int j = i; j < 3; ++j)
            M.Q[6 * kk + qidx(i, j)] += (double)C.Q[6 * kk + qidx(i, j)] + r[j] * dk[i] + r[i] * dk[j] +
                                        cc[kk] * (r[i] * r[j]);
        for (int j = 0; j < 3; ++j) M.D[3 * kk + j] += dk[j] + cc[kk] * r[j];
        M.c[kk] += cc[kk];
      }
    };
    for (int side = 0; side < 2; ++side) {
      if (rc[side].hi - rc[side].lo <= direct) shift_add(mom[cid[side]], cb[side]);
      else shift_add(m64[cid[side]], cb[side]);
    }
  }
}

// Bottom-up moments of one level (children already done) and the node's FP32
// records.  A node of at most `direct` segments (leaf buckets and the small
// subtrees above them, which are most of the nodes) gets direct_records and
// keeps no FP64 record; a larger node shifts each child's moments to its own
// centre with the shift formulas of the reference (compute_link.cpp:358-389,
// Eqs. 21-23): a larger child's FP64 record, or a small child's FP32 record.  Only nodes above
// `direct` segments write FP64 records, so the bottom levels never round-trip
// them through memory.  Children are read one at a time.
template <typename SP>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3))) void bh_moments(const BhRange *__restrict__ range, int32_t n,
                                                  int32_t base, int32_t direct,
                                                  const int2 *__restrict__ child,
                                                  const SP *__restrict__ sp,
                                                  Mom64 *__restrict__ m64, BhGeo *__restrict__ geo,
                                                  BhMom *__restrict__ mom) {
  __shared__ float4 stage[2][18 * 16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int k0 = blockIdx.x * blockDim.x + w * 64;  // this wave's first node past base
  if (k0 >= n) return;                                // whole wave: no LDS exchange to join
  const int k = k0 + lane, nw = min(64, n - k0);
  BhGeo gr;
  BhMom m;
  Mom64 M;
  bool fp64 = false;
  if (k < n) {
    const int id = base + k;
    const int2 ch = child[id];
    const BhRange rr = range[id];
    if (ch.x < 0 || rr.hi - rr.lo <= direct) {
      direct_records(rr, ch, sp, gr, m);
    } else {
      fp64 = true;
      node_moments(rr, ch, direct, range, sp, m64, mom, M);
      node_records(M, ch, rr, gr, m);
    }
  }
  wave_store_m64(M, __ballot(fp64), m64 + base + k0, nw, stage[w]);
  wave_store_records(gr, m, geo + base + k0, mom + base + k0, nw, stage[w]);
}


// ---------------------------------------------------------------- far field
struct F3 {
  float x, y, z;
};
__device__ __forceinline__ F3 f3(float a, float b, float c) { return F3{a, b, c}; }
__device__ __forceinline__ float dot3(F3 a, F3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
__device__ __forceinline__ F3 cross3(F3 a, F3 b) {
  return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ F3 add3(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 sub3(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 Dv(const float *D, F3 v) {  // D v
  return f3(fmaf(D[0], v.x, fmaf(D[1], v.y, D[2] * v.z)), fmaf(D[3], v.x, fmaf(D[4], v.y, D[5] * v.z)),
            fmaf(D[6], v.x, fmaf(D[7], v.y, D[8] * v.z)));
}
__device__ __forceinline__ F3 DTv(const float *D, F3 v) {  // D^T v
  return f3(fmaf(D[0], v.x, fmaf(D[3], v.y, D[6] * v.z)), fmaf(D[1], v.x, fmaf(D[4], v.y, D[7] * v.z)),
            fmaf(D[2], v.x, fmaf(D[5], v.y, D[8] * v.z)));
}
// Q_k v  (row i: sum_j Q_k,ij v_j) from the symmetric packing
__device__ __forceinline__ F3 Qv(const float *Q, int k, F3 v) {
  const float *q = Q + 6 * k;
  return f3(fmaf(q[0], v.x, fmaf(q[1], v.y, q[2] * v.z)), fmaf(q[1], v.x, fmaf(q[3], v.y, q[4] * v.z)),
            fmaf(q[2], v.x, fmaf(q[4], v.y, q[5] * v.z)));
}
__device__ __forceinline__ float Qel(const float *Q, int k, int i, int j) { return Q[6 * k + qidx(i, j)]; }

// Second-order far field between node A (curve 1) and node B (curve 2) with
// r = ctr_B - ctr_A: returns the linking-number contribution and writes the
// Eq. 14 error estimate (both already multiplied by 1/(4 pi |r|^3)).
__device__ float far_field(const BhMom &A, const BhMom &B, const BhGeo &GA, const BhGeo &GB, F3 r,
                           float dsq, float &err) {
  const F3 ca = f3(A.c[0], A.c[1], A.c[2]), cb = f3(B.c[0], B.c[1], B.c[2]);
  const F3 wa = f3(A.D[5] - A.D[7], A.D[6] - A.D[2], A.D[1] - A.D[3]);
  const F3 wb = f3(B.D[5] - B.D[7], B.D[6] - B.D[2], B.D[1] - B.D[3]);
  const float ir2 = 1.0f / dsq;
  float acc = dot3(r, cross3(cb, ca));  // monopole
  // dipole (Eq. 29)
  acc -= dot3(ca, wb) + dot3(wa, cb);
  acc -= 3.0f * ir2 * (dot3(cross3(Dv(B.D, r), ca), r) + dot3(cross3(Dv(A.D, r), cb), r));
  // quadrupole (Eq. 30)
  const F3 ta = f3(Qel(A.Q, 0, 0, 0) + Qel(A.Q, 0, 1, 1) + Qel(A.Q, 0, 2, 2),
                   Qel(A.Q, 1, 0, 0) + Qel(A.Q, 1, 1, 1) + Qel(A.Q, 1, 2, 2),
                   Qel(A.Q, 2, 0, 0) + Qel(A.Q, 2, 1, 1) + Qel(A.Q, 2, 2, 2));
  const F3 tb = f3(Qel(B.Q, 0, 0, 0) + Qel(B.Q, 0, 1, 1) + Qel(B.Q, 0, 2, 2),
                   Qel(B.Q, 1, 0, 0) + Qel(B.Q, 1, 1, 1) + Qel(B.Q, 1, 2, 2),
                   Qel(B.Q, 2, 0, 0) + Qel(B.Q, 2, 1, 1) + Qel(B.Q, 2, 2, 2));
  const float s2a = -dot3(r, add3(cross3(ta, cb), cross3(ca, tb)));
  auto rd = [](const float *X, int a, const float *Y, int b) {
    return fmaf(X[3 * a], Y[3 * b], fmaf(X[3 * a + 1], Y[3 * b + 1], X[3 * a + 2] * Y[3 * b + 2]));
  };
  const float s2b = 2.0f * (r.x * (rd(A.D, 1, B.D, 2) - rd(A.D, 2, B.D, 1)) +
                            r.y * (rd(A.D, 2, B.D, 0) - rd(A.D, 0, B.D, 2)) +
                            r.z * (rd(A.D, 0, B.D, 1) - rd(A.D, 1, B.D, 0)));
  // u_i = eps_ijk (Q_j r)_k ... as (Q1 r)_2 - (Q2 r)_1 etc.
  const F3 a0 = Qv(A.Q, 0, r), a1 = Qv(A.Q, 1, r), a2 = Qv(A.Q, 2, r);
  const F3 b0 = Qv(B.Q, 0, r), b1 = Qv(B.Q, 1, r), b2 = Qv(B.Q, 2, r);
  const F3 ua = f3(a1.z - a2.y, a2.x - a0.z, a0.y - a1.x);
  const F3 ub = f3(b1.z - b2.y, b2.x - b0.z, b0.y - b1.x);
  float s3 = dot3(cb, ua) - dot3(ca, ub);
  s3 += dot3(r, sub3(DTv(A.D, wb), DTv(B.D, wa)));
  const F3 rra = f3(dot3(r, a0), dot3(r, a1), dot3(r, a2));
  const F3 rrb = f3(dot3(r, b0), dot3(r, b1), dot3(r, b2));
  const float s1 = -dot3(r, cross3(cb, rra)) + dot3(r, cross3(ca, rrb)) +
                   2.0f * dot3(Dv(A.D, r), cross3(r, Dv(B.D, r)));
  acc -= 1.5f * ir2 * ((s2a + s2b + 2.0f * s3) + 5.0f * ir2 * s1);
  const float mult = ir2 * __builtin_amdgcn_rsqf(dsq) * kInv4PiF;
  err = mult * ir2 *
        (GA.qn * (GA.radius * GB.cn + 3.0f * geo_dn(GB)) + GB.qn * (GB.radius * GA.cn + 3.0f * geo_dn(GA)));
  return mult * acc;
}

// Reference leaf term (compute_link.cpp:576-606) from relative midpoint r and
// tangents t1, t2: det == 0 handled with copysign as the reference does.
__device__ __forceinline__ float leaf_term(F3 r, F3 t1, F3 t2) {
  const float det = dot3(r, cross3(t2, t1));
  const F3 a = f3(fmaf(0.5f, t1.x - t2.x, r.x), fmaf(0.5f, t1.y - t2.y, r.y), fmaf(0.5f, t1.z - t2.z, r.z));
  const F3 b = f3(fmaf(-0.5f, t1.x + t2.x, r.x), fmaf(-0.5f, t1.y + t2.y, r.y), fmaf(-0.5f, t1.z + t2.z, r.z));
  const F3 c = f3(fmaf(0.5f, t2.x - t1.x, r.x), fmaf(0.5f, t2.y - t1.y, r.y), fmaf(0.5f, t2.z - t1.z, r.z));
  const F3 d = f3(fmaf(0.5f, t2.x + t1.x, r.x), fmaf(0.5f, t2.y + t1.y, r.y), fmaf(0.5f, t2.z + t1.z, r.z));
  const float la = sqrtf(dot3(a, a)), lb = sqrtf(dot3(b, b)), lc = sqrtf(dot3(c, c)), ld = sqrtf(dot3(d, d));
  const float ac = dot3(a, c);
  const float div1 = la * lb * lc + dot3(a, b) * lc + ac * lb + dot3(b, c) * la;
  const float div2 = la * ld * lc + dot3(a, d) * lc + ac * ld + dot3(d, c) * la;
  const float xp = fmaf(div1, div2, -det * det);
  const float s = det == 0.f ? copysignf(1.f, div1) + copysignf(1.f, div2) : div1 + div2;
  const float yp = det * s;
  const float sg = det == 0.f ? -copysignf(1.f, div1) : copysignf(1.f, det);
  return kInv2PiF * atan2f(yp, xp) + ((s < 0.f || (s == 0.f && det < 0.f)) ? sg : 0.f);
}

// Per-wave FP64 accumulation into per-pair sums: one atomic per wave when
// the wave's active items share a pair, one per item otherwise.
__device__ __forceinline__ void pair_add(int pair, bool active, double v, double e,
                                         double *__restrict__ sum, double *__restrict__ err) {
  const unsigned long long act = __ballot(active);
  if (act == 0) return;
  const int lead = __ffsll((long long)act) - 1;
  const int p0 = __shfl(pair, lead, 64);
  const bool uniform = __ballot(active && pair != p0) == 0;
  if (uniform) {
    double a = active ? v : 0.0, b = active ? e : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off, 64);
      b += __shfl_xor(b, off, 64);
    }
    if ((int)(threadIdx.x & 63) == lead) {
      atomicAdd(sum + p0, a);
      if (b != 0.0) atomicAdd(err + p0, b);
    }
  } else if (active) {
    atomicAdd(sum + pair, v);
    if (e != 0.0) atomicAdd(err + pair, e);
  }
}

// Wave-aggregated append of `cnt` slots; returns this lane's first slot.
__device__ __forceinline__ int64_t wave_append(unsigned long long *counter, int cnt) {
  int incl = cnt;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(incl, off, 64);
    if ((int)(threadIdx.x & 63) >= off) incl += y;
  }
  const int total = __shfl(incl, 63, 64);
  unsigned long long base = 0;
  if ((threadIdx.x & 63) == 63 && total > 0) base = atomicAdd(counter, (unsigned long long)total);
  base = __shfl(base, 63, 64);
  return (int64_t)base + incl - cnt;
}

// Traversal blocks: 256 threads, each classifying IPT items (2, 4 or 8, a
// tunable shape), so one block owns 256 * IPT consecutive items.  Queue slots,
// sums and statistics are reduced in LDS and published with one global atomic
// per block: atomics on a single address serialise across the chip (one per
// wave capped a 4M-item level at ~7 ns per atomic), and one per 1024 items
// removes that ceiling.
constexpr int kTravBlock = 256;

struct TravShared {
  double v[4], e[4];
  int pair[4];
  int scan[4];
  int nsplit, nleaf;
  unsigned long long base[2];
};

// Adds every thread's (pair, v, e) into sum/err.  When all contributions of
// the block share a pair (items of one pair cluster in the queues) this is one
// atomic per quantity per block; mixed waves fall back to one per item.
// Called by every thread of the block.
__device__ void block_pair_add(int pair, bool active, double v, double e, double *__restrict__ sum,
                               double *__restrict__ err, TravShared &sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long act = __ballot(active);
  int p0 = -1;
  bool uniform = true;
  if (act) {
    p0 = __shfl(pair, __ffsll((long long)act) - 1, 64);
    uniform = __ballot(active && pair != p0) == 0;
  }
  if (!uniform && active) {
    atomicAdd(sum + pair, v);
    if (e != 0.0) atomicAdd(err + pair, e);
  }
  double a = (active && uniform) ? v : 0.0, b = (active && uniform) ? e : 0.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  __syncthreads();
  if (lane == 0) {
    sh.pair[w] = (act && uniform) ? p0 : -1;
    sh.v[w] = a;
    sh.e[w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int p = -1;
    bool same = true;
    for (int k = 0; k < kTravBlock / 64; ++k)
      if (sh.pair[k] >= 0) {
        if (p < 0) p = sh.pair[k];
        else if (sh.pair[k] != p) same = false;
      }
    if (p >= 0 && same) {
      double sa = 0.0, sb = 0.0;
      for (int k = 0; k < kTravBlock / 64; ++k)
        if (sh.pair[k] >= 0) {
          sa += sh.v[k];
          sb += sh.e[k];
        }
      atomicAdd(sum + p, sa);
      if (sb != 0.0) atomicAdd(err + p, sb);
    } else if (p >= 0) {
      for (int k = 0; k < kTravBlock / 64; ++k)
        if (sh.pair[k] >= 0) {
          atomicAdd(sum + sh.pair[k], sh.v[k]);
          if (sh.e[k] != 0.0) atomicAdd(err + sh.pair[k], sh.e[k]);
        }
    }
  }
}

// Slot of this lane among the wave's lanes with `pred`, and the wave's
// first slot from an LDS counter (one LDS atomic per wave).
__device__ __forceinline__ int wave_lds_slot(bool pred, int *counter) {
  const unsigned long long m = __ballot(pred);
  if (m == 0) return -1;
  const int lead = __ffsll((long long)m) - 1;
  int base = 0;
  if ((int)(threadIdx.x & 63) == lead) base = atomicAdd(counter, __popcll(m));
  base = __shfl(base, lead, 64);
  const unsigned long long below = m & ((1ull << (threadIdx.x & 63)) - 1ull);
  return pred ? base + __popcll(below) : -1;
}

// Learned criterion inside the level kernel: the network's biases and input
// normalisation, and each wave's rows of features and predictions.
struct MlpShared {
  InferImage img;  // weight fragments in LDS: the level kernel has no VGPRs to spare for them
  __attribute__((aligned(16))) _Float16 xs[kTravBlock / 64][64][16];
  float ys[kTravBlock / 64][64];
  unsigned int evals;  // network evaluations of the block
};
struct NoMlpShared {};

// Predictions of the learned criterion for the wave's 64 items (one per
// lane; log10 of the far-field error, +inf where the decision does not
// depend on it): the features of the items that need the network are
// compacted into the wave's LDS rows and evaluated 16 per tile, all of the
// wave's tiles through each weight fragment read from LDS (infer_tiles), and
// each lane reads its item's prediction back.  Wave-uniform: every lane calls
// it.
__device__ __forceinline__ float wave_predict(bool need, const BhGeo &A, const BhGeo &B, float dsq, MlpShared &ms,
                                              unsigned int &wave_evals) {
  const unsigned long long m = __ballot(need);
  if (m == 0) return __builtin_inff();
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6, g = l >> 4, c = l & 15;
  const int pos = __popcll(m & ((1ull << l) - 1ull));
  if (need) {
    float f[FLK_MLP_NFEAT];
    mlp_features_t<true>(A, B, dsq, f);  // log2 features, normalised with mean2 / istd2
    f16x8 v0, v1;
#pragma unroll
    for (int j = 0; j < 8; ++j) v0[j] = (_Float16)((f[j] - ms.img.L.mean2[j]) * ms.img.L.istd2[j]);
#pragma unroll
    for (int j = 0; j < 8; ++j)
      v1[j] = (_Float16)(j < FLK_MLP_NFEAT - 8 ? (f[8 + j] - ms.img.L.mean2[8 + j]) * ms.img.L.istd2[8 + j] : 0.f);
    *reinterpret_cast<f16x8 *>(&ms.xs[w][pos][0]) = v0;
    *reinterpret_cast<f16x8 *>(&ms.xs[w][pos][8]) = v1;
  }
  wave_lds_sync();
  const int cnt = __popcll(m);
  wave_evals += (unsigned int)cnt;
  // rows past cnt are zero (their predictions are not kept); four tiles when
  // more than two are needed, else two
  f16x8 bx[4];
  float yv[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int row = 16 * t + c;
    if (g < 2 && row < cnt) bx[t] = *reinterpret_cast<const f16x8 *>(&ms.xs[w][row][8 * g]);
    else
#pragma unroll
      for (int j = 0; j < 8; ++j) bx[t][j] = (_Float16)0.f;
  }
  if (cnt > 32) infer_tiles<4>(LdsWeights{(LdsW *)&ms.img.W, l}, ms.img.L, bx, yv);
  else infer_tiles<2>(LdsWeights{(LdsW *)&ms.img.W, l}, ms.img.L, bx, yv);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int row = 16 * t + c;
    if (g == 0 && row < cnt) ms.ys[w][row] = yv[t];
  }
  wave_lds_sync();
  const float p = need ? ms.ys[w][pos] : __builtin_inff();
  wave_lds_sync();  // every lane has its prediction before the rows are reused
  return p;
}

// One traversal level over a chunk of items.  Split items are staged in LDS
// as (pair with the split side in its sign, unsplit node, child 0, child 1)
// from the front of the block's stage, leaf-bucket items from the back; after
// the block has classified its 256 * IPT items the stage is copied out
// coalesced behind one queue atomic each.  MLP: the learned criterion, its
// network evaluated here on the node records the level already holds (the
// weight fragments in LDS, loaded once per block); *evals counts the
// evaluations.
template <int IPT, bool MLP>
__global__ __launch_bounds__(kTravBlock) __attribute__((amdgpu_waves_per_eu(4))) void bh_level(
    const BhItem *__restrict__ items, int64_t n, const unsigned long long *__restrict__ n_dev,
    const BhGeo *__restrict__ geo, const BhMom *__restrict__ mom, const float *__restrict__ beta2,
    float eps, const InferImage *__restrict__ mlp_img, float log_tol, unsigned long long *__restrict__ evals,
    BhItem *__restrict__ next, unsigned long long *__restrict__ next_count, int64_t next_cap,
    BhItem *__restrict__ leaves, unsigned long long *__restrict__ leaf_count, int64_t leaf_cap,
    double *__restrict__ psum, double *__restrict__ perr) {
  constexpr int kTravItems = kTravBlock * IPT;
  // A level enqueued before its count reached the host reads the count from
  // the previous level's queue counter; n is then the bound (the buffer's
  // capacity).  A block takes chunks blockIdx.x, blockIdx.x + gridDim.x, ...:
  // one chunk per block when the grid covers the bound, several on a
  // persistent grid.  The loop condition is uniform across the block.  A
  // count above the bound means the previous level overflowed its buffer
  // (its last item may be unwritten): the level is skipped, and the host,
  // seeing the count, discards the pass.
  if (n_dev) n = level_count(n_dev, n);
  __shared__ int4 stage[kTravItems];
  __shared__ TravShared sh;
  __shared__ std::conditional_t<MLP, MlpShared, NoMlpShared> ms;
  unsigned int wave_evals = 0;
  // Items per chunk: 256 * IPT, or, for a level that one round of the grid
  // covers, the count spread over every block in whole multiples of 256, so
  // a small level costs its threads one or two item iterations (each a chain
  // of dependent loads) instead of IPT.  Uniform across the grid.
  const int64_t gsz = (int64_t)gridDim.x;
  const int64_t csz =
      n <= gsz * kTravItems ? min<int64_t>(kTravItems, ((n + gsz - 1) / gsz + kTravBlock - 1) / kTravBlock * kTravBlock)
                            : kTravItems;
  if constexpr (MLP) {
    if ((int64_t)blockIdx.x * csz >= n) return;  // the whole block, before any barrier
    // the network's LDS image, every load of a thread in flight at once
    constexpr int kWords = (int)(sizeof(InferImage) / 16);
    static_assert(kWords > 3 * kTravBlock && kWords <= 4 * kTravBlock, "four 16-byte words per thread");
    const int4 *src = reinterpret_cast<const int4 *>(mlp_img) + threadIdx.x;
    int4 *dst = reinterpret_cast<int4 *>(&ms.img) + threadIdx.x;
    const bool w3 = 3 * kTravBlock + (int)threadIdx.x < kWords;
    const int4 v0 = src[0], v1 = src[kTravBlock], v2 = src[2 * kTravBlock];
    const int4 v3 = w3 ? src[3 * kTravBlock] : int4{};
    dst[0] = v0;
    dst[kTravBlock] = v1;
    dst[2 * kTravBlock] = v2;
    if (w3) dst[3 * kTravBlock] = v3;
    if (threadIdx.x == 0) ms.evals = 0;  // read after the chunk loop's barriers
    __syncthreads();
  }
  for (int64_t chunk = blockIdx.x; chunk * csz < n; chunk += gridDim.x) {
  if (threadIdx.x == 0) sh.nsplit = sh.nleaf = 0;
  __syncthreads();
  const int64_t b0 = chunk * csz + threadIdx.x;
  int tp = -1;  // this thread's far-field pair and its running sums
  double tv = 0.0, te = 0.0;
#pragma unroll 1
  for (int k = 0; k < IPT && k * kTravBlock < csz; ++k) {
    const int64_t t = b0 + (int64_t)k * kTravBlock;
    const bool live = t < n;
    BhItem it = live ? items[t] : BhItem{0, 0, 0, 0};
    int kind = -1;  // 0 leaf pair, 1 far, 2 split
    int ca = 0, cb = 0;
    bool split_a = false;
    float pv = __builtin_inff();  // learned criterion: predicted log10 far-field error
    // Learned criterion: both node records are loaded once, before the
    // prediction, and kept for the classification (a dead lane reads node 0
    // of its empty item; loading them again after the network's LDS fences
    // cost another round trip: 2x8M learned 11.75 -> 11.15 ms).  Beta loads
    // them inside the live branch (loading them before it measured 4% slower).
    BhGeo A0, B0;
    if constexpr (MLP) {
      A0 = geo[it.a];
      B0 = geo[it.b];
      const float rx = (B0.cx - A0.cx) + (B0.lx - A0.lx), ry = (B0.cy - A0.cy) + (B0.ly - A0.ly),
                  rz = (B0.cz - A0.cz) + (B0.lz - A0.lz);
      const float dsq = fmaf(rx, rx, fmaf(ry, ry, rz * rz)) + eps;
      const float rs = A0.radius + B0.radius;
      pv = wave_predict(live && !(A0.c0 < 0 && B0.c0 < 0) && dsq > rs * rs + eps, A0, B0, dsq, ms, wave_evals);
    }
    if (live) {
      const BhGeo A = MLP ? A0 : geo[it.a], B = MLP ? B0 : geo[it.b];
      const F3 r = f3((B.cx - A.cx) + (B.lx - A.lx), (B.cy - A.cy) + (B.ly - A.ly),
                      (B.cz - A.cz) + (B.lz - A.lz));
      const float dsq = dot3(r, r) + eps;
      const bool la = A.c0 < 0, lb = B.c0 < 0;
      const float rs = A.radius + B.radius;
      // Learned criterion: far field when the spheres are disjoint and the
      // predicted log10 error is below the tolerance; the squared predicted
      // error is accumulated (the pair's estimate is the root of the sum, as
      // for independent signs).  Otherwise the reference's beta test.
      const bool far = MLP ? (dsq > rs * rs + eps && pv < log_tol) : (dsq > rs * rs * beta2[it.pair] + eps);
      if (la && lb) {
        kind = 0;
      } else if (far) {
        kind = 1;
        float e = 0.f;
        const float val = far_field(node_mom(mom, it.a, A), node_mom(mom, it.b, B), A, B, r, dsq, e);
        if (MLP) e = exp10f(2.0f * pv);
        if (tp < 0 || tp == it.pair) {
          tp = it.pair;
          tv += (double)val;
          te += (double)e;
        } else {
          atomicAdd(psum + it.pair, (double)val);
          if (e != 0.f) atomicAdd(perr + it.pair, (double)e);
        }
      } else {
        kind = 2;
        split_a = (!la && !lb) ? !(B.rsq > A.rsq) : lb;
        ca = split_a ? A.c0 : B.c0;
        cb = split_a ? A.c1 : B.c1;
      }
    }
    const int ss = wave_lds_slot(kind == 2, &sh.nsplit);
    if (ss >= 0) stage[ss] = make_int4(split_a ? ~it.pair : it.pair, split_a ? it.b : it.a, ca, cb);
    const int sl = wave_lds_slot(kind == 0, &sh.nleaf);
    if (sl >= 0) stage[kTravItems - 1 - sl] = make_int4(it.pair, it.a, it.b, 0);
  }
  block_pair_add(tp, tp >= 0, tv, te, psum, perr, sh);
  // block_pair_add ends with every stage write visible (it synchronises)
  const int ns = sh.nsplit, nl = sh.nleaf;
  if (threadIdx.x == 0) {
    sh.base[0] = ns ? atomicAdd(next_count, 2ull * (unsigned long long)ns) : 0ull;
    sh.base[1] = nl ? atomicAdd(leaf_count, (unsigned long long)nl) : 0ull;
  }
  __syncthreads();
  const int64_t bn = (int64_t)sh.base[0], bl = (int64_t)sh.base[1];
  for (int i = threadIdx.x; i < ns; i += kTravBlock) {
    const int4 r = stage[i];
    const int64_t s = bn + 2 * (int64_t)i;
    if (s + 1 < next_cap) {
      if (r.x < 0) {  // node 1 split: (child, unsplit node 2)
        next[s] = BhItem{~r.x, r.z, r.y, 0};
        next[s + 1] = BhItem{~r.x, r.w, r.y, 0};
      } else {
        next[s] = BhItem{r.x, r.y, r.z, 0};
        next[s + 1] = BhItem{r.x, r.y, r.w, 0};
      }
    }
  }
  for (int i = threadIdx.x; i < nl; i += kTravBlock) {
    const int64_t s = bl + i;
    if (s < leaf_cap) {
      const int4 r = stage[kTravItems - 1 - i];
      leaves[s] = BhItem{r.x, r.y, r.z, 0};
    }
  }
  __syncthreads();  // the stage and the shared counters are reused by the next chunk
  }
  if constexpr (MLP) {
    if ((threadIdx.x & 63) == 0 && wave_evals) atomicAdd(&ms.evals, wave_evals);
    __syncthreads();
    if (threadIdx.x == 0 && ms.evals) atomicAdd(evals, (unsigned long long)ms.evals);
  }
}

// Exact evaluation of leaf-bucket pairs: every segment pair of the two
// buckets with the reference leaf term, from FP32 hi/lo vertices (two
// one-segment leaves: from their node records, which hold the midpoints and
// the edge vectors).  Blocks of
// 256 * IPT items as in bh_level (chunks strided over the grid; with n_dev
// the count is read on the device, n being the bound); counters[2] gets the
// segment-pair count.
template <int IPT>
__global__ __launch_bounds__(kTravBlock) void bh_leaf_pairs(
    const BhItem *__restrict__ items, int64_t n, const unsigned long long *__restrict__ n_dev,
    const BhGeo *__restrict__ geo, const uint32_t *__restrict__ order, const float4 *__restrict__ hi,
    const float4 *__restrict__ lo, const float4 *__restrict__ ed, double *__restrict__ psum,
    double *__restrict__ perr, unsigned long long *__restrict__ counters) {
  constexpr int kTravItems = kTravBlock * IPT;
  if (n_dev) n = level_count(n_dev, n);
  __shared__ TravShared sh;
  int npairs = 0;
  for (int64_t chunk = blockIdx.x; chunk * kTravItems < n; chunk += gridDim.x) {
  const int64_t b0 = chunk * kTravItems + threadIdx.x;
  int tp = -1;
  double tv = 0.0;
#pragma unroll 1
  for (int k = 0; k < IPT; ++k) {
    const int64_t t = b0 + (int64_t)k * kTravBlock;
    if (t >= n) break;
    const BhItem it = items[t];
    const BhGeo A = geo[it.a], B = geo[it.b];
    float acc = 0.f;
    if (is_leaf_segment(A) && is_leaf_segment(B)) {
      // one segment each: the records already hold the midpoints (the box
      // centres, hi/lo) and the FP32 edge vectors, so no vertex is read
      const F3 r = f3((B.cx - A.cx) + (B.lx - A.lx), (B.cy - A.cy) + (B.ly - A.ly), (B.cz - A.cz) + (B.lz - A.lz));
      acc = leaf_term(r, f3(__builtin_bit_cast(float, A.c1), A.pad, A.dn),
                      f3(__builtin_bit_cast(float, B.c1), B.pad, B.dn));
      ++npairs;
    } else {
      for (int i = A.seg0; i < A.seg0 + A.nseg; ++i) {
        const uint32_t va = order[i];
        const float4 pa = hi[va], pb = hi[va + 1], qa = lo[va], qb = lo[va + 1], ea = ed[va];
        for (int j = B.seg0; j < B.seg0 + B.nseg; ++j) {
          const uint32_t vb = order[j];
          const float4 ra = hi[vb], rb = hi[vb + 1], sa = lo[vb], sb = lo[vb + 1], eb = ed[vb];
          // r = m2 - m1 = ((q0 - p0) + (q1 - p1)) / 2, each difference hi/lo exact
          const F3 r = f3(0.5f * (((ra.x - pa.x) + (sa.x - qa.x)) + ((rb.x - pb.x) + (sb.x - qb.x))),
                          0.5f * (((ra.y - pa.y) + (sa.y - qa.y)) + ((rb.y - pb.y) + (sb.y - qb.y))),
                          0.5f * (((ra.z - pa.z) + (sa.z - qa.z)) + ((rb.z - pb.z) + (sb.z - qb.z))));
          acc += leaf_term(r, f3(ea.x, ea.y, ea.z), f3(eb.x, eb.y, eb.z));
          ++npairs;
        }
      }
    }
    if (tp < 0 || tp == it.pair) {
      tp = it.pair;
      tv += (double)acc;
    } else {
      atomicAdd(psum + it.pair, (double)acc);
    }
  }
  block_pair_add(tp, tp >= 0, tv, 0.0, psum, perr, sh);
  __syncthreads();  // sh is reused by the next chunk
  }
  int tot = npairs;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off, 64);
  if ((threadIdx.x & 63) == 0) sh.scan[threadIdx.x >> 6] = tot;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s = 0;
    for (int k = 0; k < kTravBlock / 64; ++k) s += (unsigned long long)sh.scan[k];
    if (s) atomicAdd(counters + 2, s);
  }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Training-sample traversal level: every node pair with disjoint bounding
// spheres is a candidate sample (kept with probability p when the exact label
// costs at most max_pairs segment pairs); pairs are split further until they
// pass beta_stop, so samples cover 1 < |r|/(r1+r2) <= beta_stop at every depth.
__global__ __launch_bounds__(256) void bh_sample_level(
    const BhItem *__restrict__ items, int64_t n, const BhGeo *__restrict__ geo, float eps,
    float stop2, double p, uint64_t seed, int64_t max_pairs, BhItem *__restrict__ next,
    unsigned long long *__restrict__ next_count, BhItem *__restrict__ samp,
    unsigned long long *__restrict__ samp_count, int64_t samp_cap) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = t < n;
  BhItem it = live ? items[t] : BhItem{0, 0, 0, 0};
  bool split = false, keep = false;
  int ca = 0, cb = 0;
  bool split_a = false;
  if (live) {
    const BhGeo A = geo[it.a], B = geo[it.b];
    const float rx = (B.cx - A.cx) + (B.lx - A.lx), ry = (B.cy - A.cy) + (B.ly - A.ly),
                rz = (B.cz - A.cz) + (B.lz - A.lz);
    const float dsq = rx * rx + ry * ry + rz * rz + eps;
    const bool la = A.c0 < 0, lb = B.c0 < 0;
    const float rs = A.radius + B.radius;
    if (!(la && lb)) {
      if (dsq > rs * rs + eps) {
        const uint64_t h = mix64(seed ^ ((uint64_t)(uint32_t)it.a << 32 | (uint32_t)it.b));
        keep = (int64_t)A.nseg * B.nseg <= max_pairs && (double)(h >> 11) * 0x1.0p-53 < p;
      }
      split = !(dsq > rs * rs * stop2 + eps);
      if (split) {
        split_a = (!la && !lb) ? !(B.rsq > A.rsq) : lb;
        ca = split_a ? A.c0 : B.c0;
        cb = split_a ? A.c1 : B.c1;
      }
    }
  }
  const int64_t s = wave_append(next_count, split ? 2 : 0);
  if (split) {
    if (split_a) {
      next[s] = BhItem{0, ca, it.b, 0};
      next[s + 1] = BhItem{0, cb, it.b, 0};
    } else {
      next[s] = BhItem{0, it.a, ca, 0};
      next[s + 1] = BhItem{0, it.a, cb, 0};
    }
  }
  const int64_t k = wave_append(samp_count, keep ? 1 : 0);
  if (keep && k < samp_cap) samp[k] = it;
}

// FP64 one-arctan segment-pair term (reference compute_link.cpp:151-175) of
// segment p0->p1 against q0->q1, times 4 pi.
__device__ __forceinline__ double pair_term_x4pi_f64(double4 p0, double4 p1, double4 q0, double4 q1) {
  const double ax = q0.x - p0.x, ay = q0.y - p0.y, az = q0.z - p0.z;
  const double bx = q0.x - p1.x, by = q0.y - p1.y, bz = q0.z - p1.z;
  const double cx = q1.x - p1.x, cy = q1.y - p1.y, cz = q1.z - p1.z;
  const double dx = q1.x - p0.x, dy = q1.y - p0.y, dz = q1.z - p0.z;
  const double la = sqrt(ax * ax + ay * ay + az * az), lb = sqrt(bx * bx + by * by + bz * bz);
  const double lc = sqrt(cx * cx + cy * cy + cz * cz), ld = sqrt(dx * dx + dy * dy + dz * dz);
  const double ac = cx * ax + cy * ay + cz * az;
  const double div1 = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + ac * lb + (cx * bx + cy * by + cz * bz) * la;
  const double div2 = la * ld * lc + (ax * dx + ay * dy + az * dz) * lc + ac * ld + (cx * dx + cy * dy + cz * dz) * la;
  const double det = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
  const double xp = div1 * div2 - det * det, yp = det * (div1 + div2);
  double v = 2.0 * atan2(yp, xp);
  if (yp * det < 0) v += (det > 0 ? 1.0 : -1.0) * 12.566370614359172953850573533118;
  return v;
}

// One wave per sample: exact FP64 sum over the two nodes' segment pairs, the
// FP32 far field, features and the label log10 |exact - far field| (>= -14).
__global__ __launch_bounds__(256) void bh_sample_labels(
    const BhItem *__restrict__ samp, int64_t n, const BhGeo *__restrict__ geo,
    const BhMom *__restrict__ mom, const uint32_t *__restrict__ order,
    const double4 *__restrict__ pts, float eps, float *__restrict__ x, float *__restrict__ y) {
  const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int l = threadIdx.x & 63;
  if (w >= n) return;
  const BhItem it = samp[w];
  const BhGeo A = geo[it.a], B = geo[it.b];
  double acc = 0.0;
  const int64_t total = (int64_t)A.nseg * B.nseg;
  for (int64_t k = l; k < total; k += 64) {
    const uint32_t va = order[A.seg0 + (int)(k / B.nseg)], vb = order[B.seg0 + (int)(k % B.nseg)];
    acc += pair_term_x4pi_f64(pts[va], pts[va + 1], pts[vb], pts[vb + 1]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (l == 0) {
    const double exact = acc * 0.07957747154594766788444188168625718;
    const F3 r = f3((B.cx - A.cx) + (B.lx - A.lx), (B.cy - A.cy) + (B.ly - A.ly),
                    (B.cz - A.cz) + (B.lz - A.lz));
    const float dsq = dot3(r, r) + eps;
    float e;
    const float ff = far_field(node_mom(mom, it.a, A), node_mom(mom, it.b, B), A, B, r, dsq, e);
    const double d = fabs(exact - (double)ff);
    y[w] = (float)fmax(log10(d), -14.0);
    float f[FLK_MLP_NFEAT];
    mlp_features(A, B, dsq, f);
    for (int k2 = 0; k2 < FLK_MLP_NFEAT; ++k2) x[w * FLK_MLP_NFEAT + k2] = f[k2];
  }
}

__global__ void bh_roots(const int32_t *__restrict__ ra, const int32_t *__restrict__ rb, int64_t n,
                         BhItem *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = BhItem{(int32_t)k, ra[k], rb[k], 0};
}

// Clears the counters of a device-scheduled pass and sets the root level's
// count (a kernel rather than a memset, so the pass can be captured).
__global__ void bh_pass_init(unsigned long long *__restrict__ cnt, int n, int root_slot,
                             unsigned long long roots) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) cnt[k] = k == root_slot ? roots : 0ull;
}

__global__ void bh_zero_f64(double *__restrict__ p, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0.0;
}

// Exclusive scan of one value per thread over a 256-thread block; *total
// receives the sum.  sw: 4 ints of LDS.  Contains block barriers.
__device__ __forceinline__ int block_scan256(int v, int *total, int *sw) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(inc, d, 64);
    if (lane >= d) inc += u;
  }
  if (lane == 63) sw[w] = inc;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    off += k < w ? sw[k] : 0;
    tot += sw[k];
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// The re-run decision after a batch's first pass, on the device.  Beta
// (reference linking_number_certificate.cpp:165-184): the estimate scaled by
// err_scale gives the target beta = min(beta_init (E / target)^(1/4),
// beta_limit), and a pair whose target exceeds beta_init is re-run with it.
// Learned criterion: the root-sum-square of the accepted predictions is the
// estimate, and a pair above a tenth of the target is re-run (one decade
// stricter, by the caller's second pass).  One block walks the pairs in order
// and compacts the re-runs with a block scan, so the list is deterministic:
// root items (pair = position in the list), beta^2 and the pair's index.  It
// also clears the second pass's counters and sets its root count at
// cnt2[root_slot].
__global__ __launch_bounds__(256) void bh_adapt(int64_t nb, const int32_t *__restrict__ ra,
                                                const int32_t *__restrict__ rb, const double *__restrict__ err1,
                                                int mlp, double err_scale, double target_error, double beta_init,
                                                double beta_limit, BhItem *__restrict__ items2,
                                                float *__restrict__ beta2_2, int32_t *__restrict__ redo,
                                                unsigned long long *__restrict__ cnt2, int nslots2, int root_slot) {
  __shared__ int sw[4];
  for (int k = threadIdx.x; k < nslots2; k += 256) cnt2[k] = 0ull;
  int carry = 0;
  for (int64_t c0 = 0; c0 < nb; c0 += 256) {
    const int64_t t = c0 + threadIdx.x;
    int flag = 0;
    float b2 = 0.f;
    if (t < nb) {
      if (!mlp) {
        const double want = pow(err1[t] * err_scale / target_error, 0.25) * beta_init;
        const double target = beta_limit < want ? beta_limit : want;  // std::min: a NaN estimate stays NaN
        if (target > beta_init) {
          flag = 1;
          b2 = (float)(target * target);
        }
      } else if (sqrt(err1[t]) > 0.1 * target_error) {
        flag = 1;
        b2 = (float)(beta_init * beta_init);
      }
    }
    int tot = 0;
    const int pos = carry + block_scan256(flag, &tot, sw);
    if (flag) {
      items2[pos] = BhItem{pos, ra[t], rb[t], 0};
      beta2_2[pos] = b2;
      redo[pos] = (int32_t)t;
    }
    carry += tot;
  }
  __syncthreads();  // the clear above is done before the count is set
  if (threadIdx.x == 0) cnt2[root_slot] = (unsigned long long)carry;
}

// Per-pair results of a batch: the first pass's sum and estimate (scaled by
// err_scale for beta, the root-sum-square for the learned criterion) ...
__global__ void bh_merge_first(int64_t nb, const double *__restrict__ sum, const double *__restrict__ err, int mlp,
                               double err_scale, double *__restrict__ v, double *__restrict__ e,
                               uint32_t *__restrict__ f) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb) return;
  v[t] = sum[t];
  e[t] = mlp ? sqrt(err[t]) : err[t] * err_scale;
  f[t] = 0u;
}

// ... replaced by the second pass's for the re-run pairs.
__global__ void bh_merge_redo(int64_t nb, const unsigned long long *__restrict__ n_redo,
                              const int32_t *__restrict__ redo, const double *__restrict__ sum2,
                              const double *__restrict__ err2, int mlp, double err_scale, double *__restrict__ v,
                              double *__restrict__ e, uint32_t *__restrict__ f) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb || k >= (int64_t)*n_redo) return;
  const int32_t t = redo[k];
  v[t] = sum2[k];
  e[t] = mlp ? sqrt(err2[k]) : err2[k] * err_scale;
  f[t] = FLK_PAIR_RETRY;
}

}  // namespace

// ---------------------------------------------------------------- host side
namespace {

// Node records of the trees of one call; the buffers are the context's.

struct Forest {
  int32_t nnodes = 0;
  int32_t depth = 0;          // node levels of the deepest tree
  std::vector<int32_t> root;  // per curve (-1 if no tree)
  DevBuf &geo, &mom, &order;
  explicit Forest(BhScratch &b) : geo(b.geo), mom(b.mom), order(b.order) {}
};

// Host plan of a forest build: the curves with trees (rank order), their
// segment and vertex offsets, key boxes, roots and the level structure, all
// known before anything runs on the device.  plan_forest also reserves every
// buffer the build writes; enqueue_forest then only enqueues (no allocation,
// no read-back), so the build can be captured.  The host arrays are read by
// the build's copies when they run: the plan outlives them.
struct ForestPlan {
  int leaf = 1, direct = kDirectSegments;
  std::vector<int32_t> cs;
  std::vector<int64_t> seg_off, vtx_off;
  std::vector<double> cbox;
  std::vector<BhRange> roots;
  std::vector<RadixTile> tiles;  // the segmented sort's tiles (one curve each)
  std::vector<int32_t> tcurve;   // each tile's curve rank
  std::vector<int32_t> root_flag;  // the roots' split flags (more than `leaf` segments)
  int64_t S = 0;
  std::vector<int32_t> level_base, level_n;
  std::vector<int64_t> level_split, level_max;
  int32_t total = 0;
  int lev_top = 0;
  // In a graph: seg_off, vtx_off, cbox, roots, tiles, tcurve and root_flag
  // uploaded once (dsrc, in that order), copied device to device on every
  // replay.
  DevBuf dconst;
  std::vector<const void *> dsrc;
};

// Uploads host arrays into one device buffer, each piece 16-byte aligned, and
// returns the pieces' device addresses (the graph's copies then read these).
std::vector<const void *> upload_pack(DevBuf &dst, const std::vector<std::pair<const void *, size_t>> &parts,
                                      hipStream_t s) {
  size_t total = 0;
  for (const auto &p : parts) total += (p.second + 15) & ~size_t(15);
  dst.reserve(std::max<size_t>(total, 16));
  std::vector<const void *> out;
  size_t off = 0;
  for (const auto &p : parts) {
    unsigned char *d = dst.as<unsigned char>() + off;
    if (p.second) FLK_HIP(hipMemcpyAsync(d, p.first, p.second, hipMemcpyHostToDevice, s));
    out.push_back(d);
    off += (p.second + 15) & ~size_t(15);
  }
  return out;
}

// A plan's host array into device scratch: device to device from its
// uploaded copy (a graph), else from the host.
void put(void *dst, const void *host, const std::vector<const void *> &dsrc, int i, size_t bytes, hipStream_t s) {
  if (dsrc.empty()) FLK_HIP(hipMemcpyAsync(dst, host, bytes, hipMemcpyHostToDevice, s));
  else FLK_HIP(hipMemcpyAsync(dst, dsrc[i], bytes, hipMemcpyDeviceToDevice, s));
}

void plan_forest(flk_ctx *ctx, const std::vector<int32_t> &curves_in, int leaf, ForestPlan &P) {
  Curves &cv = ctx->curves;
  P = ForestPlan{};
  P.leaf = leaf;
  P.direct = std::max(leaf, kDirectSegments);
  for (int32_t c : curves_in)
    if (cv.n[c] >= 2) P.cs.push_back(c);
  const int32_t nc = (int32_t)P.cs.size();
  if (nc == 0) return;
  if (cv.nvert_dev >= (int64_t)UINT32_MAX) throw Error(FLK_ERR_ARG, "too many vertices for Barnes-Hut");
  P.seg_off.assign(nc + 1, 0);
  P.vtx_off.assign(nc, 0);
  P.cbox.assign(6 * (size_t)nc, 0.0);
  for (int32_t k = 0; k < nc; ++k) {
    P.seg_off[k + 1] = P.seg_off[k] + cv.n[P.cs[k]];
    P.vtx_off[k] = cv.off[P.cs[k]];
  }
  // Sort keys: a Morton code of 10 bits per axis relative to the curve's box
  // (a curve's tree never depends on the batch around it), sorted within
  // each curve's range by the segmented sort: 30-bit keys, four passes,
  // whatever the number of curves.
  const double cells = 1024.0;
  for (int32_t k = 0; k < nc; ++k)
    for (int d = 0; d < 3; ++d) {
      const double lo = (cv.box[6 * P.cs[k] + d] - cv.origin[d]) * cv.scale;
      const double hi = (cv.box[6 * P.cs[k] + 3 + d] - cv.origin[d]) * cv.scale;
      P.cbox[6 * k + d] = lo;
      P.cbox[6 * k + 3 + d] = hi > lo ? cells / (hi - lo) : 0.0;
    }
  const int64_t S = P.S = P.seg_off[nc];
  if (S >= INT32_MAX / 2) throw Error(FLK_ERR_ARG, "too many Barnes-Hut segments");
  P.roots.resize(nc);
  for (int32_t k = 0; k < nc; ++k) P.roots[k] = BhRange{(int32_t)P.seg_off[k], (int32_t)P.seg_off[k + 1]};
  P.root_flag.resize(nc);
  for (int32_t k = 0; k < nc; ++k) P.root_flag[k] = P.seg_off[k + 1] - P.seg_off[k] > leaf ? 1 : 0;
  const int64_t ntiles = segmented_tiles(P.seg_off.data(), nc, P.tiles);
  P.tcurve.resize(ntiles);
  for (int32_t k = 0, t = 0; k < nc; ++k)
    for (; t < ntiles && P.tiles[t].start < P.seg_off[k + 1]; ++t) P.tcurve[t] = k;
  // Splits per level from the segment counts alone: a range of m > leaf
  // segments splits into floor(m/2) and m - floor(m/2), so every level of a
  // curve's tree holds at most two distinct range sizes (m and m + 1).  With
  // the counts known, the levels are enqueued without reading anything back.
  {
    std::vector<std::pair<int64_t, int64_t>> sizes;  // (range size, count) of the current level
    std::vector<std::pair<int64_t, int64_t>> next;
    for (int32_t k = 0; k < nc; ++k) sizes.emplace_back(P.seg_off[k + 1] - P.seg_off[k], 1);
    for (;;) {
      int64_t split = 0, mx = 0;
      for (const auto &sc : sizes) mx = std::max(mx, sc.first);
      P.level_max.push_back(mx);
      next.clear();
      for (const auto &sc : sizes) {
        if (sc.first <= leaf) continue;
        split += sc.second;
        next.emplace_back(sc.first / 2, sc.second);
        next.emplace_back(sc.first - sc.first / 2, sc.second);
      }
      P.level_split.push_back(split);
      if (split == 0) break;
      // merge equal sizes (keeps the list at two entries per distinct size pair)
      std::sort(next.begin(), next.end());
      sizes.clear();
      for (const auto &sc : next) {
        if (!sizes.empty() && sizes.back().first == sc.first) sizes.back().second += sc.second;
        else sizes.push_back(sc);
      }
    }
  }
  P.level_base.assign(1, 0);
  P.level_n.assign(1, nc);
  int32_t total = nc;
  for (size_t lev = 0; lev < P.level_split.size(); ++lev) {
    const int32_t nsplit = (int32_t)P.level_split[lev];
    if (nsplit == 0) break;
    P.level_base.push_back(total);
    P.level_n.push_back(2 * nsplit);
    total += 2 * nsplit;
  }
  P.total = total;
  // the bottom levels (every range within `direct`) run as one launch
  P.lev_top = (int)P.level_base.size();
  for (int lev = 0; lev < (int)P.level_base.size() && lev < (int)P.level_max.size(); ++lev)
    if (P.level_max[lev] <= P.direct) {
      P.lev_top = lev;
      break;
    }
  // FP64 records exist only for nodes above `direct` segments: node ids are
  // assigned level by level, so they all fall below the first level whose
  // ranges are all that small
  const int64_t m64_nodes = P.lev_top < (int)P.level_base.size() ? P.level_base[P.lev_top] : total;
  int64_t widest = nc;
  for (int64_t sp : P.level_split) widest = std::max(widest, 2 * sp);
  BhScratch &B = ctx->bh;
  B.seg.reserve(sizeof(int64_t) * (nc + 1));
  B.vtx.reserve(sizeof(int64_t) * nc);
  B.box.reserve(sizeof(double) * 6 * nc);
  B.keys0.reserve(sizeof(uint32_t) * S);
  B.keys1.reserve(sizeof(uint32_t) * S);
  B.rtiles.reserve(sizeof(RadixTile) * ntiles);
  B.rcurve.reserve(sizeof(int32_t) * ntiles);
  B.ids0.reserve(sizeof(uint32_t) * S);
  B.order.reserve(sizeof(uint32_t) * S);
  B.flag.reserve(sizeof(int32_t) * std::max<int64_t>(S + nc + 1, widest));
  B.flag2.reserve(sizeof(int32_t) * std::max<int64_t>(S + nc + 1, widest));
  B.scan.reserve(sizeof(int32_t) * std::max<int64_t>(S + nc + 1, widest));  // also the sort's value scratch
  B.tmp.reserve(std::max(radix_temp_bytes(S, ntiles), scan_temp_bytes(widest)));
  B.sp.reserve(sizeof(double2) * 3 * S);
  B.range.reserve(sizeof(BhRange) * (2 * S + nc));
  B.child.reserve(sizeof(int2) * (2 * S + nc));
  B.m64.reserve(sizeof(Mom64) * std::max<int64_t>(m64_nodes, 1));
  B.geo.reserve(sizeof(BhGeo) * total);
  B.mom.reserve(sizeof(BhMom) * total);
}

// The end-point gather, level-by-level node creation and the moments of a
// planned forest, with the segments' end points staged as SP.
template <typename SP>
void enqueue_moments(flk_ctx *ctx, const ForestPlan &P, SP *sp, hipStream_t s) {
  Curves &cv = ctx->curves;
  BhScratch &B = ctx->bh;
  const int64_t S = P.S;
  const int32_t nc = (int32_t)P.cs.size();
  // FP32 staging means every coordinate is exact in FP32: read the FP32 copies
  if constexpr (std::is_same_v<SP, float2>)
    hipLaunchKernelGGL((bh_stage_segments<SP, float4>), dim3((unsigned)((S + 255) / 256)), dim3(256), 0, s,
                       B.order.as<uint32_t>(), cv.hi.as<float4>(), S, sp);
  else
    hipLaunchKernelGGL((bh_stage_segments<SP, double4>), dim3((unsigned)((S + 255) / 256)), dim3(256), 0, s,
                       B.order.as<uint32_t>(), cv.dpts.as<double4>(), S, sp);
  FLK_HIP(hipGetLastError());
  // level-by-level node creation; split flags alternate between two buffers
  put(B.range.ptr, P.roots.data(), P.dsrc, 3, sizeof(BhRange) * nc, s);
  put(B.flag.ptr, P.root_flag.data(), P.dsrc, 6, sizeof(int32_t) * nc, s);
  int32_t *flags[2] = {B.flag.as<int32_t>(), B.flag2.as<int32_t>()};
  for (size_t lev = 0; lev < P.level_base.size(); ++lev) {
    const int32_t base = P.level_base[lev], n = P.level_n[lev];
    const int32_t next = lev + 1 < P.level_base.size() ? P.level_base[lev + 1] : P.total;
    const unsigned nb = (unsigned)((n + 255) / 256);
    // slots of the split nodes: a scan of the flags, unless the plan shows
    // that every node of the level splits
    const bool all = P.level_split[lev] == n;
    if (!all) exclusive_scan(flags[lev & 1], B.scan.as<int32_t>(), n, B.tmp.as<int32_t>(), s);
    hipLaunchKernelGGL(bh_split, dim3(nb), dim3(256), 0, s, B.range.as<BhRange>() + base, n, base, next, P.leaf,
                       (const int32_t *)flags[lev & 1], all ? nullptr : (const int32_t *)B.scan.as<int32_t>(),
                       B.range.as<BhRange>() + next, B.child.as<int2>(), flags[(lev + 1) & 1]);
    FLK_HIP(hipGetLastError());
  }
  if (P.lev_top < (int)P.level_base.size()) {
    const int32_t b0 = P.level_base[P.lev_top], nd = P.total - b0;
    hipLaunchKernelGGL(bh_moments_direct<SP>, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, s,
                       B.range.as<BhRange>(), b0, nd, B.child.as<int2>(), (const SP *)sp, B.geo.as<BhGeo>(),
                       B.mom.as<BhMom>());
    FLK_HIP(hipGetLastError());
  }
  for (int lev = P.lev_top - 1; lev >= 0; --lev) {
    const int32_t n = P.level_n[lev];
    hipLaunchKernelGGL(bh_moments<SP>, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, B.range.as<BhRange>(),
                       n, P.level_base[lev], P.direct, B.child.as<int2>(), (const SP *)sp, B.m64.as<Mom64>(),
                       B.geo.as<BhGeo>(), B.mom.as<BhMom>());
    FLK_HIP(hipGetLastError());
  }
}

void enqueue_forest(flk_ctx *ctx, const ForestPlan &P, hipStream_t s) {
  const int32_t nc = (int32_t)P.cs.size();
  if (nc == 0) return;
  Curves &cv = ctx->curves;
  BhScratch &B = ctx->bh;
  const int64_t S = P.S;
  put(B.seg.ptr, P.seg_off.data(), P.dsrc, 0, sizeof(int64_t) * (nc + 1), s);
  put(B.vtx.ptr, P.vtx_off.data(), P.dsrc, 1, sizeof(int64_t) * nc, s);
  put(B.box.ptr, P.cbox.data(), P.dsrc, 2, sizeof(double) * 6 * nc, s);
  const int64_t ntiles = (int64_t)P.tiles.size();
  put(B.rtiles.ptr, P.tiles.data(), P.dsrc, 4, sizeof(RadixTile) * ntiles, s);
  put(B.rcurve.ptr, P.tcurve.data(), P.dsrc, 5, sizeof(int32_t) * ntiles, s);
  if (cv.inexact)
    hipLaunchKernelGGL(bh_keys<double4>, dim3((unsigned)ntiles), dim3(256), 0, s, cv.dpts.as<double4>(),
                       B.seg.as<int64_t>(), B.vtx.as<int64_t>(), B.box.as<double>(), B.rtiles.as<RadixTile>(),
                       B.rcurve.as<int32_t>(), B.keys0.as<uint32_t>(), B.ids0.as<uint32_t>());
  else  // every coordinate exact in FP32: the FP32 copies, the same values
    hipLaunchKernelGGL(bh_keys<float4>, dim3((unsigned)ntiles), dim3(256), 0, s, cv.hi.as<float4>(),
                       B.seg.as<int64_t>(), B.vtx.as<int64_t>(), B.box.as<double>(), B.rtiles.as<RadixTile>(),
                       B.rcurve.as<int32_t>(), B.keys0.as<uint32_t>(), B.ids0.as<uint32_t>());
  FLK_HIP(hipGetLastError());
  radix_sort_pairs<uint32_t>(B.keys0.as<uint32_t>(), B.keys1.as<uint32_t>(), B.ids0.as<uint32_t>(),
                             B.scan.as<uint32_t>(), B.order.as<uint32_t>(), S, 30, B.tmp.ptr, s,
                             B.rtiles.as<RadixTile>(), ntiles);
  if (cv.inexact) enqueue_moments(ctx, P, B.sp.as<double2>(), s);
  else enqueue_moments(ctx, P, B.sp.as<float2>(), s);
}

void forest_from_plan(flk_ctx *ctx, const ForestPlan &P, Forest &F) {
  F.root.assign(ctx->curves.ncurves, -1);
  F.nnodes = P.total;
  F.depth = (int32_t)P.level_base.size();
  if (P.cs.empty()) F.depth = 0;
  for (size_t k = 0; k < P.cs.size(); ++k) F.root[P.cs[k]] = (int32_t)k;  // roots are nodes 0..nc-1
}

void build_forest(flk_ctx *ctx, const std::vector<int32_t> &curves_in, int leaf, Forest &F) {
  ForestPlan P;
  plan_forest(ctx, curves_in, leaf, P);
  enqueue_forest(ctx, P, ctx->stream);
  FLK_HIP(hipStreamSynchronize(ctx->stream));
  forest_from_plan(ctx, P, F);
}

// Items per thread of the level and leaf kernels: the tuned value, or by
// criterion (0): 8 for beta, 4 for the learned criterion, whose network
// staging in LDS would hold 8 items per thread to 2 blocks per CU (measured
// on 2 x 8M random walks: beta 21.4 ms at 8 against 21.8 at 4, learned 12.5
// at 4 against 15.1 at 8).
int items_per_thread(const flk_ctx *ctx, bool use_mlp) {
  const int t = ctx->tune.bh_items_per_thread;
  return (t == 2 || t == 4 || t == 8) ? t : (use_mlp ? 4 : 8);
}

// The level kernel for a criterion and a number of items per thread.
using LevelKernel = void (*)(const BhItem *, int64_t, const unsigned long long *, const BhGeo *, const BhMom *,
                             const float *, float, const InferImage *, float, unsigned long long *,
                             BhItem *, unsigned long long *, int64_t, BhItem *, unsigned long long *, int64_t,
                             double *, double *);
LevelKernel level_kernel(int ipt, bool use_mlp) {
  if (use_mlp) return ipt == 2 ? bh_level<2, true> : ipt == 8 ? bh_level<8, true> : bh_level<4, true>;
  return ipt == 2 ? bh_level<2, false> : ipt == 8 ? bh_level<8, false> : bh_level<4, false>;
}

// Traverses the given pairs (indices into the forest roots) with per-pair
// beta; adds into psum/perr (device, per pair).
//
// Items live in level buffers processed in chunks of up to kChunk items; a
// level that fits one chunk is consumed whole and its buffer swapped with the
// next level's (breadth-first, two buffers), a wider one pushes the next
// level on a LIFO stack so memory stays bounded by depth x chunk.  The host
// reads the queue counts (pinned memory) once per round of up to four levels;
// the device counters rotate through a pre-zeroed array instead of a clear
// per launch.
void traverse(flk_ctx *ctx, Forest &F, const std::vector<int32_t> &ra, const std::vector<int32_t> &rb,
              const std::vector<float> &beta2, float eps, double *d_sum, double *d_err,
              unsigned long long stats[4], bool use_mlp, float log_tol, bool timing) {
  hipStream_t s = ctx->stream;
  BhScratch &B = ctx->bh;
  const int64_t np = (int64_t)ra.size();
  if (np == 0) return;
  constexpr int64_t kChunk = 1 << 24;
  constexpr int kSlots = 1024;  // count pairs per zeroing of the counter array
  constexpr int kRound = 4;     // levels per host round trip (at most)
  // items per thread of the level/leaf kernels: 2, 4 or 8 (grid and kernel agree)
  const int ipt = items_per_thread(ctx, use_mlp);
  B.roots_a.reserve(sizeof(int32_t) * np);
  B.roots_b.reserve(sizeof(int32_t) * np);
  B.beta2.reserve(sizeof(float) * np);
  B.cnt.reserve(sizeof(unsigned long long) * (4 + 2 * kSlots));
  if (!B.host_cnt) FLK_HIP(hipHostMalloc((void **)&B.host_cnt, 2 * kRound * sizeof(unsigned long long)));
  FLK_HIP(hipMemcpyAsync(B.roots_a.ptr, ra.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice, s));
  FLK_HIP(hipMemcpyAsync(B.roots_b.ptr, rb.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice, s));
  FLK_HIP(hipMemcpyAsync(B.beta2.ptr, beta2.data(), sizeof(float) * np, hipMemcpyHostToDevice, s));
  // [2] leaf segment pairs, [3] network evaluations, [4 + 2k] queues
  unsigned long long *cnt = B.cnt.as<unsigned long long>();
  FLK_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * (4 + 2 * kSlots), s));

  struct Level {
    DevBuf *buf;
    int64_t n;
  };
  std::vector<Level> stack;
  size_t used = 0;  // level buffers handed out from B.levels
  auto level_buf = [&](int64_t items) -> DevBuf * {
    if (B.levels.size() <= used) B.add_level();
    DevBuf *b = B.levels[used++].get();
    b->reserve(sizeof(BhItem) * std::max<int64_t>(items, 1));
    return b;
  };
  stack.push_back(Level{level_buf(np), np});
  hipLaunchKernelGGL(bh_roots, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s,
                     B.roots_a.as<int32_t>(), B.roots_b.as<int32_t>(), np, stack[0].buf->as<BhItem>());
  FLK_HIP(hipGetLastError());
  DevBuf *spare = level_buf(2 * std::min(np, kChunk));
  if (use_mlp) mlp_upload(ctx);
  const InferImage *mlp_img = use_mlp ? ctx->mlp.d_image.as<InferImage>() : nullptr;
  unsigned long long visits = 0, far = 0;
  int slot_next = 0;
  DevBuf *alt = nullptr;  // second output buffer of a multi-level round
  while (!stack.empty()) {
    Level &top = stack.back();
    if (top.n == 0) {
      spare = top.buf;  // an emptied level's buffer is reused
      stack.pop_back();
      continue;
    }
    const int64_t take = std::min<int64_t>(top.n, kChunk);
    top.n -= take;
    const BhItem *src = top.buf->as<BhItem>() + top.n;
    // Levels enqueued before the next read-back: the first from its exact
    // count and, when it is consumed whole, up to kRound - 1 more, level d
    // sized for its bound take * 2^d (a level at most doubles) and reading
    // its true count from the previous level's counter on the device.  The
    // last level's bound keeps its output within 2 * kChunk items.  With
    // per-launch timing every level is its own round.
    int D = 1;
    if (!timing && top.n == 0)
      while (D < kRound && (take << (D + 1)) <= 2 * kChunk) ++D;
    const int64_t out_cap = take << D;  // items: bound of the last level's output
    spare->reserve(sizeof(BhItem) * out_cap);
    if (D > 1) {
      if (!alt) alt = level_buf(out_cap);
      alt->reserve(sizeof(BhItem) * out_cap);
    }
    const int64_t leaf_cap = (int64_t)((1LL << D) - 1) * take;  // sum of the level bounds
    if (B.leaves.bytes < sizeof(BhItem) * leaf_cap) {
      FLK_HIP(hipStreamSynchronize(s));  // the previous leaf launch may still read it
      B.leaves.reserve(sizeof(BhItem) * leaf_cap);
    }
    if (slot_next + D > kSlots) {
      FLK_HIP(hipMemsetAsync(cnt + 4, 0, sizeof(unsigned long long) * 2 * kSlots, s));
      slot_next = 0;
    }
    unsigned long long *qc = cnt + 4 + 2 * slot_next;  // per level [0] next items; [1] of level 0: leaves
    slot_next += D;
    DevBuf *outb[kRound];
    for (int d = 0; d < D; ++d) {
      outb[d] = (d % 2 == 0) ? spare : alt;
      const BhItem *in = d == 0 ? src : outb[d - 1]->as<BhItem>();
      const int64_t bound = take << d;
      const unsigned long long *n_dev = d == 0 ? nullptr : qc + 2 * (d - 1);
      // with timing, start/stop events ride on the dispatch itself
      const dim3 grid((unsigned)((bound + ipt * kTravBlock - 1) / (ipt * kTravBlock)));
      launch_timed(timing ? ctx->ev1 : nullptr, timing ? ctx->ev2 : nullptr, level_kernel(ipt, use_mlp), grid,
                   dim3(kTravBlock), s, in, bound, n_dev, F.geo.as<BhGeo>(), F.mom.as<BhMom>(),
                   (const float *)B.beta2.as<float>(), eps, mlp_img, log_tol, cnt + 3,
                   outb[d]->as<BhItem>(), qc + 2 * d, 2 * bound, B.leaves.as<BhItem>(), qc + 1, leaf_cap, d_sum,
                   d_err);
      FLK_HIP(hipGetLastError());
    }
    FLK_HIP(hipMemcpyAsync(B.host_cnt, qc, 2 * D * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    FLK_HIP(hipStreamSynchronize(s));
    if (timing) {
      float ms_level = 0.f;
      FLK_HIP(hipEventElapsedTime(&ms_level, ctx->ev1, ctx->ev2));
      if (use_mlp) ctx->stats.ms_mlp += ms_level;  // the network runs inside the level kernels
      ctx->stats.ms_bh_level += ms_level;
    }
    // every item is exactly one of: leaf pair, far field, split (two children)
    int64_t in_n = take;
    for (int d = 0; d < D; ++d) {
      const unsigned long long nn = B.host_cnt[2 * d];
      if ((int64_t)nn > 2 * (take << d)) throw Error(FLK_ERR_STATE, "Barnes-Hut level buffer overflow");
      visits += (unsigned long long)in_n;
      far += (unsigned long long)in_n - nn / 2;
      in_n = (int64_t)nn;
    }
    const unsigned long long nleaf = B.host_cnt[1];
    if ((int64_t)nleaf > leaf_cap) throw Error(FLK_ERR_STATE, "Barnes-Hut leaf buffer overflow");
    far -= nleaf;
    if (nleaf > 0) {
      const Curves &cv = ctx->curves;
      auto leafk = ipt == 2 ? bh_leaf_pairs<2> : ipt == 8 ? bh_leaf_pairs<8> : bh_leaf_pairs<4>;
      hipLaunchKernelGGL(leafk, dim3((unsigned)((nleaf + ipt * kTravBlock - 1) / (ipt * kTravBlock))),
                         dim3(kTravBlock), 0, s, B.leaves.as<BhItem>(), (int64_t)nleaf, nullptr, F.geo.as<BhGeo>(),
                         F.order.as<uint32_t>(), cv.hi.as<float4>(), cv.lo.as<float4>(),
                         cv.ed.as<float4>(), d_sum, d_err, cnt);
      FLK_HIP(hipGetLastError());
    }
    DevBuf *last = outb[D - 1];
    if (top.n == 0) {
      // the level is consumed: the last level's output replaces it, the
      // consumed buffer and the round's other output buffer are free
      DevBuf *consumed = top.buf;
      top.buf = last;
      top.n = in_n;
      if (D > 1) alt = (last == spare) ? alt : spare;
      spare = consumed;
    } else if (in_n > 0) {
      stack.push_back(Level{last, in_n});  // D == 1: last is spare
      spare = level_buf(2 * kChunk);
    }
  }
  unsigned long long st[4];
  FLK_HIP(hipMemcpyAsync(st, cnt, sizeof st, hipMemcpyDeviceToHost, s));
  FLK_HIP(hipStreamSynchronize(s));
  stats[0] += visits;
  stats[1] += far;
  stats[2] += st[2];
  stats[3] += st[3];  // learned-criterion network evaluations
}

// Blocks of `kernel` resident on the whole chip at once (the persistent grid
// of a device-scheduled pass).
int resident_grid(flk_ctx *ctx, const void *kernel) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kTravBlock, 0) != hipSuccess || nb < 1) nb = 1;
  return nb * std::max(1, ctx->num_cus);
}

// Device schedule of a batch's traversal passes.  A pass enqueues every
// level at once: level d reads its item count from the counter that level
// d - 1 incremented, on a persistent grid whose blocks stride over the
// level's chunks, so no count returns to the host until the batch has ended.
// Every split deepens one of the two trees, so a pass has fewer than
// 2 * depth + 1 levels; launched levels past the last nonempty one see a
// count of 0 and exit at once.  Both level buffers and the leaf list hold
// `cap` items.  A level or leaf list beyond that, or items left after the
// last launched level, mark the pass as overflowed: its sums are incomplete,
// and the batch is re-run on the host-driven schedule.
struct PassShape {
  int64_t cap = 0;
  int D = 0, nslots = 0, ipt = 4;
  unsigned grid_lev = 1, grid_leaf = 1;
};

// Counter words of the host-driven schedule (traverse); a batch's re-run
// count lives in the 8 words after them.
constexpr int kHostCounterWords = 4 + 2 * 1024;

// Sizes the device schedule of a batch of nb pairs whose deepest tree has
// `depth` node levels and reserves its buffers; false when it cannot hold
// the batch (the host-driven schedule runs it instead).
bool plan_device_passes(flk_ctx *ctx, int64_t nb, int depth, bool use_mlp, PassShape &P) {
  BhScratch &B = ctx->bh;
  constexpr int kMaxLevels = 1000;
  P.D = 2 * depth + 1;
  if (P.D > kMaxLevels) return false;
  if (B.dev_cap == 0) {
    // two level buffers and the leaf list (16 B per item), within a tenth of
    // the free memory, at most 2^26 items each
    size_t free_b = 0, total_b = 0;
    FLK_HIP(hipMemGetInfo(&free_b, &total_b));
    int64_t cap = int64_t(1) << 26;
    while (cap > (int64_t(1) << 20) && (double)cap * 52.0 > 0.1 * (double)free_b) cap >>= 1;
    B.dev_cap = cap;
  }
  P.cap = ctx->tune.bh_level_capacity > 0 ? (int64_t)ctx->tune.bh_level_capacity : B.dev_cap;
  if (nb > P.cap) return false;
  P.ipt = items_per_thread(ctx, use_mlp);
  P.nslots = 4 + 2 * (P.D + 1);
  while (B.levels.size() < 2) B.add_level();
  B.levels[0]->reserve(sizeof(BhItem) * P.cap);
  B.levels[1]->reserve(sizeof(BhItem) * P.cap);
  B.leaves.reserve(sizeof(BhItem) * P.cap);
  const LevelKernel lev = level_kernel(P.ipt, use_mlp);
  auto leafk = P.ipt == 2 ? bh_leaf_pairs<2> : P.ipt == 8 ? bh_leaf_pairs<8> : bh_leaf_pairs<4>;
  const int64_t items_per_block = (int64_t)P.ipt * kTravBlock;
  const int64_t chunks = (P.cap + items_per_block - 1) / items_per_block;
  P.grid_lev = (unsigned)std::min<int64_t>(resident_grid(ctx, (const void *)lev), chunks);
  P.grid_leaf = (unsigned)std::min<int64_t>(resident_grid(ctx, (const void *)leafk), chunks);
  return true;
}

// Enqueues one pass on the device schedule: counters at cnt (the caller's
// kernel has cleared them and set the root count at cnt[4]), level-0 items at
// items0, per-pair beta^2 and sums indexed by the items' pair field.
void enqueue_pass(flk_ctx *ctx, const PassShape &P, const BhItem *items0, const float *beta2, float eps,
                  bool use_mlp, float log_tol, unsigned long long *cnt, double *d_sum, double *d_err,
                  hipStream_t s) {
  BhScratch &B = ctx->bh;
  const Curves &cv = ctx->curves;
  BhItem *buf[2] = {B.levels[0]->as<BhItem>(), B.levels[1]->as<BhItem>()};
  const InferImage *mlp_img = use_mlp ? ctx->mlp.d_image.as<InferImage>() : nullptr;
  const LevelKernel lev = level_kernel(P.ipt, use_mlp);
  auto leafk = P.ipt == 2 ? bh_leaf_pairs<2> : P.ipt == 8 ? bh_leaf_pairs<8> : bh_leaf_pairs<4>;
  // [2] leaf segment pairs, [3] network evaluations, [4 + 2d] items of level
  // d, [5] leaf items
  for (int d = 0; d < P.D; ++d) {
    const BhItem *in = d == 0 ? items0 : buf[(d - 1) & 1];
    hipLaunchKernelGGL(lev, dim3(P.grid_lev), dim3(kTravBlock), 0, s, in, P.cap,
                       (const unsigned long long *)(cnt + 4 + 2 * d), B.geo.as<BhGeo>(), B.mom.as<BhMom>(), beta2,
                       eps, mlp_img, log_tol, cnt + 3, buf[d & 1], cnt + 4 + 2 * (d + 1), P.cap,
                       B.leaves.as<BhItem>(), cnt + 5, P.cap, d_sum, d_err);
    FLK_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(leafk, dim3(P.grid_leaf), dim3(kTravBlock), 0, s, B.leaves.as<BhItem>(), P.cap,
                     (const unsigned long long *)(cnt + 5), B.geo.as<BhGeo>(), B.order.as<uint32_t>(),
                     cv.hi.as<float4>(), cv.lo.as<float4>(), cv.ed.as<float4>(), d_sum, d_err, cnt);
  FLK_HIP(hipGetLastError());
}

// Adds a device-scheduled pass's visits, far fields, leaf segment pairs and
// network evaluations (from its counters h) to st; false when it overflowed.
// Every item is exactly one of: leaf pair, far field, split (two children).
bool pass_counts(const unsigned long long *h, const PassShape &P, unsigned long long st[4]) {
  const unsigned long long nleaf = h[5];
  if ((int64_t)nleaf > P.cap || h[4 + 2 * P.D] != 0) return false;
  unsigned long long visits = 0, split_children = 0;
  for (int d = 0; d < P.D; ++d) {
    if ((int64_t)h[4 + 2 * d] > P.cap || (int64_t)h[4 + 2 * (d + 1)] > P.cap) return false;
    visits += h[4 + 2 * d];
    split_children += h[4 + 2 * (d + 1)];
  }
  st[0] += visits;
  st[1] += visits - split_children / 2 - nleaf;
  st[2] += h[2];
  st[3] += h[3];
  return true;
}

}  // namespace

int64_t mlp_samples(flk_ctx *ctx, int32_t a, int32_t b, int64_t cap, uint64_t seed, float *x,
                    float *y) {
  Curves &cv = ctx->curves;
  hipStream_t s = ctx->stream;
  if (cap <= 0) return 0;
  if (cv.n[a] < 2 || cv.n[b] < 2) return 0;
  Forest F(ctx->bh);
  std::vector<int32_t> cs{std::min(a, b), std::max(a, b)};
  if (a == b) cs.resize(1);
  build_forest(ctx, cs, 1, F);
  const float eps = (float)(1e-24 * cv.scale * cv.scale);
  const float stop2 = 36.0f;                // samples up to |r| = 6 (r1 + r2)
  const int64_t max_pairs = 1 << 20;        // exact-label cost cap per sample
  const int64_t chunk = 1 << 22;
  DevBuf d_cnt, d_samp, d_a, d_b;
  d_cnt.reserve(sizeof(unsigned long long) * 4);
  d_a.reserve(sizeof(BhItem) * 2 * chunk);
  d_b.reserve(sizeof(BhItem) * 2 * chunk);
  d_samp.reserve(sizeof(BhItem) * cap);
  unsigned long long *cnt = d_cnt.as<unsigned long long>();
  auto run = [&](double p, int64_t samp_cap) -> int64_t {
    // breadth-first over levels; a level wider than the chunk is thinned by
    // keeping a random subset (the samples stay spread over all depths)
    FLK_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * 4, s));
    BhItem root{0, F.root[a], F.root[b], 0};
    FLK_HIP(hipMemcpyAsync(d_a.ptr, &root, sizeof root, hipMemcpyHostToDevice, s));
    int64_t n = 1;
    DevBuf *cur = &d_a, *nxt = &d_b;
    while (n > 0) {
      const int64_t take = std::min(n, chunk);
      FLK_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), s));
      hipLaunchKernelGGL(bh_sample_level, dim3((unsigned)((take + 255) / 256)), dim3(256), 0, s,
                         cur->as<BhItem>(), take, F.geo.as<BhGeo>(), eps, stop2, p, seed, max_pairs,
                         nxt->as<BhItem>(), cnt, d_samp.as<BhItem>(), cnt + 1, samp_cap);
      FLK_HIP(hipGetLastError());
      unsigned long long nn = 0;
      FLK_HIP(hipMemcpyAsync(&nn, cnt, sizeof nn, hipMemcpyDeviceToHost, s));
      FLK_HIP(hipStreamSynchronize(s));
      n = (int64_t)nn;
      std::swap(cur, nxt);
    }
    unsigned long long ns = 0;
    FLK_HIP(hipMemcpyAsync(&ns, cnt + 1, sizeof ns, hipMemcpyDeviceToHost, s));
    FLK_HIP(hipStreamSynchronize(s));
    return (int64_t)ns;
  };
  const int64_t eligible = run(1.0, 0);
  if (eligible == 0) return 0;
  const double p = std::min(1.0, 1.05 * (double)cap / (double)eligible);
  const int64_t got = std::min(run(p, cap), cap);
  DevBuf dx, dy;
  dx.reserve(sizeof(float) * FLK_MLP_NFEAT * got);
  dy.reserve(sizeof(float) * got);
  hipLaunchKernelGGL(bh_sample_labels, dim3((unsigned)((got * 64 + 255) / 256)), dim3(256), 0, s,
                     d_samp.as<BhItem>(), got, F.geo.as<BhGeo>(), F.mom.as<BhMom>(),
                     F.order.as<uint32_t>(), cv.dpts.as<double4>(), eps, dx.as<float>(), dy.as<float>());
  FLK_HIP(hipGetLastError());
  FLK_HIP(hipMemcpyAsync(x, dx.ptr, sizeof(float) * FLK_MLP_NFEAT * got, hipMemcpyDeviceToHost, s));
  FLK_HIP(hipMemcpyAsync(y, dy.ptr, sizeof(float) * got, hipMemcpyDeviceToHost, s));
  FLK_HIP(hipStreamSynchronize(s));
  return got;
}

void tree_export(flk_ctx *ctx, int32_t curve, int32_t leaf, std::vector<flk_tree_node> &nodes,
                 std::vector<int32_t> &seg_order) {
  Curves &cv = ctx->curves;
  hipStream_t s = ctx->stream;
  nodes.clear();
  seg_order.clear();
  if (cv.n[curve] < 2) return;
  Forest F(ctx->bh);
  build_forest(ctx, std::vector<int32_t>{curve}, leaf, F);
  const int32_t nn = F.nnodes, nseg = cv.n[curve];
  std::vector<BhGeo> g(nn);
  std::vector<BhMom> m(nn);
  std::vector<uint32_t> ord(nseg);
  FLK_HIP(hipMemcpyAsync(g.data(), F.geo.ptr, sizeof(BhGeo) * nn, hipMemcpyDeviceToHost, s));
  FLK_HIP(hipMemcpyAsync(m.data(), F.mom.ptr, sizeof(BhMom) * nn, hipMemcpyDeviceToHost, s));
  FLK_HIP(hipMemcpyAsync(ord.data(), F.order.ptr, sizeof(uint32_t) * nseg, hipMemcpyDeviceToHost, s));
  FLK_HIP(hipStreamSynchronize(s));
  // stored = (input - origin) * scale with a power-of-two scale: the
  // conversions back are exact apart from the origin's addition
  const double is = 1.0 / cv.scale, is2 = is * is, is3 = is2 * is;
  nodes.resize(nn);
  for (int32_t k = 0; k < nn; ++k) {
    const BhGeo &a = g[k];
    const BhMom b = is_leaf_segment(a) ? leaf_segment_mom(a) : m[k];
    flk_tree_node &o = nodes[k];
    std::memset(&o, 0, sizeof o);
    const double ctr[3] = {(double)a.cx + (double)a.lx, (double)a.cy + (double)a.ly, (double)a.cz + (double)a.lz};
    for (int d = 0; d < 3; ++d) o.center[d] = cv.origin[d] + ctr[d] * is;
    o.radius = (double)a.radius * is;
    o.rsq = (double)a.rsq * is2;
    for (int d = 0; d < 3; ++d) o.c[d] = (double)b.c[d] * is;
    for (int t = 0; t < 9; ++t) o.D[t] = (double)b.D[t] * is2;
    for (int kk = 0; kk < 3; ++kk)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) o.Q[9 * kk + 3 * i + j] = (double)b.Q[6 * kk + qidx(i, j)] * is3;
    o.cn = (double)a.cn * is;
    o.dn = (double)geo_dn(a) * is2;
    o.qn = (double)a.qn * is3;
    o.child[0] = a.c0 < 0 ? -1 : a.c0;
    o.child[1] = a.c0 < 0 ? -1 : a.c1;
    o.seg0 = a.seg0;
    o.nseg = a.nseg;
  }
  seg_order.resize(nseg);
  for (int32_t t = 0; t < nseg; ++t) seg_order[t] = (int32_t)((int64_t)ord[t] - cv.off[curve]);
}

namespace {

// One batch of Barnes–Hut pairs (indices into the call's pair list): its
// forest plan, root nodes, first-pass beta^2, the device schedule's shape,
// and the results and counters the batch's read-back fills.
struct BatchPlan {
  std::vector<int64_t> idx;
  ForestPlan forest;
  std::vector<int32_t> ra, rb;
  std::vector<float> beta2;
  PassShape shape;
  bool device = false;  // the device schedule holds it
  std::vector<double> v, e;
  std::vector<uint32_t> f;
  std::vector<unsigned long long> counts;  // both passes' counters, nslots each
  // In a graph: ra, rb and beta2 uploaded once (dsrc), and the read-back into
  // pinned memory (res), copied into v, e, f and counts after the launch.
  DevBuf dconst;
  std::vector<const void *> dsrc;
  HostBuf res;
  double *hv = nullptr, *he = nullptr;
  uint32_t *hf = nullptr;
  unsigned long long *hc = nullptr;
  // where the read-back lands: the vectors, or res
  void land_in_vectors() {
    hv = v.data();
    he = e.data();
    hf = f.data();
    hc = counts.data();
  }
  void land_pinned(hipStream_t s) {
    const size_t nb = idx.size(), nc = counts.size();
    res.reserve(16 * nb + 4 * nb + 8 * nc + 64);
    unsigned char *p = res.as<unsigned char>();
    hv = reinterpret_cast<double *>(p);
    he = hv + nb;
    hc = reinterpret_cast<unsigned long long *>(he + nb);
    hf = reinterpret_cast<uint32_t *>(hc + nc);
    dsrc = upload_pack(dconst, {{ra.data(), 4 * nb}, {rb.data(), 4 * nb}, {beta2.data(), 4 * nb}}, s);
    forest.dsrc = upload_pack(forest.dconst,
                              {{forest.seg_off.data(), 8 * forest.seg_off.size()},
                               {forest.vtx_off.data(), 8 * forest.vtx_off.size()},
                               {forest.cbox.data(), 8 * forest.cbox.size()},
                               {forest.roots.data(), sizeof(BhRange) * forest.roots.size()},
                               {forest.tiles.data(), sizeof(RadixTile) * forest.tiles.size()},
                               {forest.tcurve.data(), 4 * forest.tcurve.size()},
                               {forest.root_flag.data(), 4 * forest.root_flag.size()}},
                              s);
  }
  void collect() {  // pinned read-back into the vectors
    if (hv == v.data()) return;
    std::copy(hv, hv + v.size(), v.begin());
    std::copy(he, he + e.size(), e.begin());
    std::copy(hf, hf + f.size(), f.begin());
    std::copy(hc, hc + counts.size(), counts.begin());
  }
};

void plan_batch(flk_ctx *ctx, const int32_t *pairs, std::vector<int64_t> idx, const flk_params &prm,
                BatchPlan &bp) {
  Curves &cv = ctx->curves;
  const bool use_mlp = prm.criterion == FLK_CRIT_MLP;
  bp.idx = std::move(idx);
  std::vector<char> used(cv.ncurves, 0);
  for (int64_t k : bp.idx) used[pairs[2 * k]] = used[pairs[2 * k + 1]] = 1;
  std::vector<int32_t> cs;
  for (int32_t c = 0; c < cv.ncurves; ++c)
    if (used[c]) cs.push_back(c);
  plan_forest(ctx, cs, std::max(1, prm.leaf_size), bp.forest);
  std::vector<int32_t> root(cv.ncurves, -1);
  for (size_t k = 0; k < bp.forest.cs.size(); ++k) root[bp.forest.cs[k]] = (int32_t)k;
  const int64_t nb = (int64_t)bp.idx.size();
  bp.ra.resize(nb);
  bp.rb.resize(nb);
  bp.beta2.assign(nb, (float)(prm.beta_init * prm.beta_init));
  for (int64_t t = 0; t < nb; ++t) {
    bp.ra[t] = root[pairs[2 * bp.idx[t]]];
    bp.rb[t] = root[pairs[2 * bp.idx[t] + 1]];
  }
  bp.device = plan_device_passes(ctx, nb, (int)bp.forest.level_base.size(), use_mlp, bp.shape);
  BhScratch &B = ctx->bh;
  B.roots_a.reserve(sizeof(int32_t) * nb);
  B.roots_b.reserve(sizeof(int32_t) * nb);
  B.beta2.reserve(sizeof(float) * 2 * nb);
  B.items0.reserve(sizeof(BhItem) * nb);
  B.redo.reserve(sizeof(int32_t) * nb);
  B.sum.reserve(sizeof(double) * 2 * nb);
  B.err.reserve(sizeof(double) * 2 * nb);
  B.outv.reserve(sizeof(double) * nb);
  B.oute.reserve(sizeof(double) * nb);
  B.outf.reserve(sizeof(uint32_t) * nb);
  B.cnt.reserve(sizeof(unsigned long long) * std::max(2 * bp.shape.nslots, kHostCounterWords + 8));
  if (use_mlp) mlp_upload(ctx);
  bp.v.assign(nb, 0.0);
  bp.e.assign(nb, 0.0);
  bp.f.assign(nb, 0u);
  bp.counts.assign(2 * (size_t)bp.shape.nslots, 0ull);
  bp.land_in_vectors();
}

// The batch's merge of both passes into per-pair results, and their read-back.
void enqueue_merge(flk_ctx *ctx, BatchPlan &bp, const flk_params &prm, const unsigned long long *n_redo,
                   hipStream_t s) {
  BhScratch &B = ctx->bh;
  const int64_t nb = (int64_t)bp.idx.size();
  const int mlp = prm.criterion == FLK_CRIT_MLP ? 1 : 0;
  const unsigned g = (unsigned)((nb + 255) / 256);
  hipLaunchKernelGGL(bh_merge_first, dim3(g), dim3(256), 0, s, nb, (const double *)B.sum.as<double>(),
                     (const double *)B.err.as<double>(), mlp, prm.err_scale, B.outv.as<double>(),
                     B.oute.as<double>(), B.outf.as<uint32_t>());
  hipLaunchKernelGGL(bh_merge_redo, dim3(g), dim3(256), 0, s, nb, n_redo, (const int32_t *)B.redo.as<int32_t>(),
                     (const double *)(B.sum.as<double>() + nb), (const double *)(B.err.as<double>() + nb), mlp,
                     prm.err_scale, B.outv.as<double>(), B.oute.as<double>(), B.outf.as<uint32_t>());
  FLK_HIP(hipGetLastError());
  FLK_HIP(hipMemcpyAsync(bp.hv, B.outv.ptr, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  FLK_HIP(hipMemcpyAsync(bp.he, B.oute.ptr, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  FLK_HIP(hipMemcpyAsync(bp.hf, B.outf.ptr, sizeof(uint32_t) * nb, hipMemcpyDeviceToHost, s));
}

void enqueue_adapt(flk_ctx *ctx, int64_t nb, const flk_params &prm, unsigned long long *cnt2, int nslots2,
                   hipStream_t s) {
  BhScratch &B = ctx->bh;
  hipLaunchKernelGGL(bh_adapt, dim3(1), dim3(256), 0, s, nb, (const int32_t *)B.roots_a.as<int32_t>(),
                     (const int32_t *)B.roots_b.as<int32_t>(), (const double *)B.err.as<double>(),
                     prm.criterion == FLK_CRIT_MLP ? 1 : 0, prm.err_scale, prm.target_error, prm.beta_init,
                     prm.beta_limit, B.items0.as<BhItem>(), B.beta2.as<float>() + nb, B.redo.as<int32_t>(), cnt2,
                     nslots2, 4);
  FLK_HIP(hipGetLastError());
}

// Traversal of a planned batch on the device schedule, after its forest:
// first pass, the re-run decision (bh_adapt), second pass over the re-run
// list, the merge, and the read-back of results and counters.  Kernels and
// copies only, so it can be captured.
void enqueue_traversal(flk_ctx *ctx, BatchPlan &bp, const flk_params &prm, hipStream_t s) {
  BhScratch &B = ctx->bh;
  const Curves &cv = ctx->curves;
  const bool use_mlp = prm.criterion == FLK_CRIT_MLP;
  const int64_t nb = (int64_t)bp.idx.size();
  const PassShape &P = bp.shape;
  const float eps = (float)(prm.eps_sq * cv.scale * cv.scale);
  const float log_tol = (float)prm.mlp_log_tol;
  double *sum = B.sum.as<double>(), *err = B.err.as<double>();
  hipLaunchKernelGGL(bh_zero_f64, dim3((unsigned)((2 * nb + 255) / 256)), dim3(256), 0, s, sum, 2 * nb);
  hipLaunchKernelGGL(bh_zero_f64, dim3((unsigned)((2 * nb + 255) / 256)), dim3(256), 0, s, err, 2 * nb);
  put(B.roots_a.ptr, bp.ra.data(), bp.dsrc, 0, sizeof(int32_t) * nb, s);
  put(B.roots_b.ptr, bp.rb.data(), bp.dsrc, 1, sizeof(int32_t) * nb, s);
  put(B.beta2.ptr, bp.beta2.data(), bp.dsrc, 2, sizeof(float) * nb, s);
  unsigned long long *c1 = B.cnt.as<unsigned long long>(), *c2 = c1 + P.nslots;
  hipLaunchKernelGGL(bh_pass_init, dim3((unsigned)((P.nslots + 255) / 256)), dim3(256), 0, s, c1, P.nslots, 4,
                     (unsigned long long)nb);
  hipLaunchKernelGGL(bh_roots, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s,
                     (const int32_t *)B.roots_a.as<int32_t>(), (const int32_t *)B.roots_b.as<int32_t>(), nb,
                     B.items0.as<BhItem>());
  FLK_HIP(hipGetLastError());
  enqueue_pass(ctx, P, B.items0.as<BhItem>(), B.beta2.as<float>(), eps, use_mlp, log_tol, c1, sum, err, s);
  enqueue_adapt(ctx, nb, prm, c2, P.nslots, s);
  enqueue_pass(ctx, P, B.items0.as<BhItem>(), B.beta2.as<float>() + nb, eps, use_mlp, log_tol - 1.0f, c2,
               sum + nb, err + nb, s);
  enqueue_merge(ctx, bp, prm, c2 + 4, s);
  FLK_HIP(hipMemcpyAsync(bp.hc, c1, sizeof(unsigned long long) * 2 * P.nslots, hipMemcpyDeviceToHost,
                         s));
}

// Results and statistics of a batch whose read-back has arrived; false when
// a pass overflowed the device schedule (nothing is written then).
bool finish_batch(flk_ctx *ctx, const BatchPlan &bp, std::vector<double> &lk, std::vector<double> &err,
                  std::vector<uint32_t> &flags) {
  unsigned long long st[4] = {0, 0, 0, 0};
  const unsigned long long *h = bp.counts.data();
  if (!pass_counts(h, bp.shape, st) || !pass_counts(h + bp.shape.nslots, bp.shape, st)) return false;
  for (size_t t = 0; t < bp.idx.size(); ++t) {
    lk[bp.idx[t]] = bp.v[t];
    err[bp.idx[t]] = bp.e[t];
    flags[bp.idx[t]] |= bp.f[t];
  }
  ctx->stats.retries += (int64_t)h[bp.shape.nslots + 4];
  ctx->stats.bh_visits += (int64_t)st[0];
  ctx->stats.bh_far += (int64_t)st[1];
  ctx->stats.bh_leaf_pairs += (int64_t)st[2];
  ctx->stats.mlp_evals += (int64_t)st[3];
  return true;
}

// A batch on the host-driven schedule: per-launch timing, or a batch the
// device schedule cannot hold or that overflowed it.  The same re-run
// decision (bh_adapt) and merge; the re-run list comes back to the host.
void bh_batch_host(flk_ctx *ctx, const int32_t *pairs, const std::vector<int64_t> &idx, const flk_params &prm,
                   std::vector<double> &lk, std::vector<double> &err, std::vector<uint32_t> &flags) {
  hipStream_t s = ctx->stream;
  const bool use_mlp = prm.criterion == FLK_CRIT_MLP;
  BatchPlan bp;
  auto t0 = std::chrono::steady_clock::now();
  plan_batch(ctx, pairs, idx, prm, bp);
  enqueue_forest(ctx, bp.forest, s);
  FLK_HIP(hipStreamSynchronize(s));
  ctx->stats.ms_bh_build +=
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  t0 = std::chrono::steady_clock::now();
  Forest F(ctx->bh);
  forest_from_plan(ctx, bp.forest, F);
  BhScratch &B = ctx->bh;
  const int64_t nb = (int64_t)bp.idx.size();
  const float eps = (float)(prm.eps_sq * ctx->curves.scale * ctx->curves.scale);
  double *sum = B.sum.as<double>(), *err_d = B.err.as<double>();
  hipLaunchKernelGGL(bh_zero_f64, dim3((unsigned)((2 * nb + 255) / 256)), dim3(256), 0, s, sum, 2 * nb);
  hipLaunchKernelGGL(bh_zero_f64, dim3((unsigned)((2 * nb + 255) / 256)), dim3(256), 0, s, err_d, 2 * nb);
  unsigned long long st[4] = {0, 0, 0, 0};
  traverse(ctx, F, bp.ra, bp.rb, bp.beta2, eps, sum, err_d, st, use_mlp, (float)prm.mlp_log_tol, prm.timing != 0);
  unsigned long long *c2 = B.cnt.as<unsigned long long>() + kHostCounterWords;
  enqueue_adapt(ctx, nb, prm, c2, 8, s);
  unsigned long long nr = 0;
  FLK_HIP(hipMemcpyAsync(&nr, c2 + 4, sizeof nr, hipMemcpyDeviceToHost, s));
  FLK_HIP(hipStreamSynchronize(s));
  int passes = 1;
  if (nr > 0) {
    std::vector<BhItem> it(nr);
    std::vector<float> b2(nr);
    FLK_HIP(hipMemcpyAsync(it.data(), B.items0.ptr, sizeof(BhItem) * nr, hipMemcpyDeviceToHost, s));
    FLK_HIP(hipMemcpyAsync(b2.data(), B.beta2.as<float>() + nb, sizeof(float) * nr, hipMemcpyDeviceToHost, s));
    FLK_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> ra2(nr), rb2(nr);
    for (size_t k = 0; k < nr; ++k) {
      ra2[k] = it[k].a;
      rb2[k] = it[k].b;
    }
    traverse(ctx, F, ra2, rb2, b2, eps, sum + nb, err_d + nb, st, use_mlp, (float)prm.mlp_log_tol - 1.0f,
             prm.timing != 0);
    passes = 2;
  }
  enqueue_merge(ctx, bp, prm, c2 + 4, s);
  FLK_HIP(hipStreamSynchronize(s));
  for (size_t t = 0; t < bp.idx.size(); ++t) {
    lk[bp.idx[t]] = bp.v[t];
    err[bp.idx[t]] = bp.e[t];
    flags[bp.idx[t]] |= bp.f[t];
  }
  ctx->stats.retries += (int64_t)nr;
  ctx->stats.bh_host_passes += passes;
  ctx->stats.bh_visits += (int64_t)st[0];
  ctx->stats.bh_far += (int64_t)st[1];
  ctx->stats.bh_leaf_pairs += (int64_t)st[2];
  ctx->stats.mlp_evals += (int64_t)st[3];
  ctx->stats.ms_bh_traverse +=
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Trees, traversal and the adaptive-beta / learned-criterion re-run for one
// batch of Barnes–Hut pairs, launched eagerly; results into lk/err/flags.
void bh_batch(flk_ctx *ctx, const int32_t *pairs, const std::vector<int64_t> &idx, const flk_params &prm,
              std::vector<double> &lk, std::vector<double> &err, std::vector<uint32_t> &flags) {
  if (prm.timing) {
    bh_batch_host(ctx, pairs, idx, prm, lk, err, flags);
    return;
  }
  hipStream_t s = ctx->stream;
  BatchPlan bp;
  auto t0 = std::chrono::steady_clock::now();
  plan_batch(ctx, pairs, idx, prm, bp);
  if (!bp.device) {
    bh_batch_host(ctx, pairs, idx, prm, lk, err, flags);
    return;
  }
  enqueue_forest(ctx, bp.forest, s);
  FLK_HIP(hipStreamSynchronize(s));
  ctx->stats.ms_bh_build +=
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  t0 = std::chrono::steady_clock::now();
  enqueue_traversal(ctx, bp, prm, s);
  FLK_HIP(hipStreamSynchronize(s));
  ctx->stats.ms_bh_traverse +=
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (!finish_batch(ctx, bp, lk, err, flags)) bh_batch_host(ctx, pairs, idx, prm, lk, err, flags);
}

// Device bytes the trees and their build records take per segment: two
// 192-byte node records (384), 288-byte FP64 records for the nodes above 8
// segments (about one node in eight: ~72), ranges and children (32), sort keys,
// ids and order (24), the staged segment endpoints (48), scan flags (8) and
// sort scratch, with headroom.
constexpr double kBhBytesPerSegment = 648.0;

// Batches of Barnes–Hut pairs whose curves' trees fit the device: the
// segment budget is bh_batch_ksegs thousand, or what free memory (plus the
// scratch this context already holds) allows when a curve set first runs
// Barnes–Hut -- kept for later calls on the same curves, so repeated
// certificates form the same batches and reuse the same scratch.  A curve
// shared by two batches is built in both.
std::vector<std::vector<int64_t>> form_batches(flk_ctx *ctx, const int32_t *pairs, const std::vector<int64_t> &bh_idx,
                                               const flk_params &prm) {
  Curves &cv = ctx->curves;
  int64_t budget;
  if (prm.bh_batch_ksegs > 0) {
    budget = (int64_t)prm.bh_batch_ksegs * 1000;
  } else {
    if (cv.bh_budget <= 0) {
      size_t free_b = 0, total_b = 0;
      FLK_HIP(hipMemGetInfo(&free_b, &total_b));
      cv.bh_budget = (int64_t)(0.6 * (double)(free_b + ctx->bh.bytes()) / kBhBytesPerSegment);
    }
    budget = cv.bh_budget;
  }
  budget = std::max<int64_t>(budget, 1);
  std::vector<std::vector<int64_t>> out;
  std::vector<char> in_batch(cv.ncurves, 0);
  std::vector<int32_t> members;
  std::vector<int64_t> batch;
  int64_t segs = 0;
  auto flush = [&]() {
    if (batch.empty()) return;
    out.push_back(batch);
    for (int32_t c : members) in_batch[c] = 0;
    members.clear();
    batch.clear();
    segs = 0;
  };
  for (int64_t k : bh_idx) {
    const int32_t a = pairs[2 * k], b = pairs[2 * k + 1];
    const int64_t add = (in_batch[a] ? 0 : cv.n[a]) + (in_batch[b] || a == b ? 0 : cv.n[b]);
    if (!batch.empty() && segs + add > budget) flush();
    for (int32_t c : {a, b})
      if (!in_batch[c]) {
        in_batch[c] = 1;
        members.push_back(c);
        segs += cv.n[c];
      }
    batch.push_back(k);
  }
  flush();
  return out;
}

uint64_t hash_bytes(uint64_t h, const void *p, size_t n) {
  const unsigned char *c = static_cast<const unsigned char *>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 0x100000001b3ULL;
  return h;
}

}  // namespace

// A whole Barnes–Hut certificate as one captured graph: the direct-sum batch
// of the pairs below the threshold (uploads, persistent tile kernel,
// reduction, read-back) and, per Barnes–Hut batch, the forest build (sort,
// levels, moments), both traversal passes with the device re-run decision
// between them, the merge and the read-back.  The device pointers it holds
// belong to the context's scratch: it is replayed only while no DevBuf has
// been allocated or freed since its capture (g_devbuf_epoch), and the host
// arrays its copies read live here.
struct CertGraph {
  uint64_t key = 0, epoch = 0;
  hipGraphExec_t exec = nullptr;
  std::vector<int64_t> ds_idx;
  DsGraph ds;
  std::vector<BatchPlan> batches;
  ~CertGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

namespace {

// Captures (or finds) the certificate graph for these pairs and launches it;
// false when a batch does not fit the device schedule (the caller then runs
// the eager path).
bool bh_run_graph(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, const flk_params &prm,
                  const std::vector<int64_t> &ds_idx, const std::vector<int32_t> &ds_pairs,
                  const std::vector<int64_t> &bh_idx, std::vector<double> &lk, std::vector<double> &err,
                  std::vector<uint32_t> &flags) {
  hipStream_t s = ctx->stream;
  const bool use_mlp = prm.criterion == FLK_CRIT_MLP;
  uint64_t key = hash_bytes(0xcbf29ce484222325ULL, pairs, sizeof(int32_t) * 2 * (size_t)npairs);
  const double pd[7] = {prm.beta_init, prm.beta_limit, prm.eps_sq, prm.target_error, prm.err_scale,
                        prm.mlp_log_tol, 0.0};
  const int64_t pi[8] = {prm.criterion, prm.bh_direct_threshold, prm.leaf_size, prm.fp64, prm.bh_batch_ksegs,
                         (int64_t)ctx->curves_version, ctx->curves.inexact ? 1 : 0, npairs};
  key = hash_bytes(key, pd, sizeof pd);
  key = hash_bytes(key, pi, sizeof pi);
  key = hash_bytes(key, &ctx->tune, sizeof ctx->tune);
  CertGraph *G = ctx->cert_graph.get();
  if (!(G && G->key == key && G->epoch == g_devbuf_epoch.load())) {
    ctx->cert_graph.reset();
    auto NG = std::make_shared<CertGraph>();
    NG->key = key;
    NG->ds_idx = ds_idx;
    for (auto &idx : form_batches(ctx, pairs, bh_idx, prm)) {
      NG->batches.emplace_back();
      BatchPlan &bp = NG->batches.back();
      plan_batch(ctx, pairs, std::move(idx), prm, bp);
      if (!bp.device) return false;
      bp.land_pinned(s);
    }
    if (!ds_idx.empty()) {
      // Beside the Barnes-Hut branch the direct-sum kernel takes a share of
      // its resident blocks, the direct sum's fraction of the two parts'
      // estimated times alone: its segment pairs at 1.2e12 per second (FP64:
      // 1.9e11), and 0.19 ns per Barnes-Hut tree segment for build and
      // traversal.  Both rates are the 10,000-loop lattice's (5.6e10 direct
      // pairs, 3.2e8 tree segments; share 0.43, 3 of the 8 blocks per CU),
      // whose certificate took 109 ms with the parts in sequence, 97 ms as two
      // branches at full residency, 87, 84 and 97 ms with 4, 3 and 2 blocks.
      double share = 1.0;
      if (!NG->batches.empty()) {
        const std::vector<int32_t> &n = ctx->curves.n;
        double work = 0.0, bh_segments = 0.0;
        for (size_t k = 0; k + 1 < ds_pairs.size(); k += 2)
          work += (double)n[ds_pairs[k]] * (double)n[ds_pairs[k + 1]];
        for (const BatchPlan &bp : NG->batches) bh_segments += (double)bp.forest.S;
        const double t_ds = work / (prm.fp64 ? 1.9e11 : 1.2e12), t_bh = bh_segments * 0.19e-9;
        share = t_ds / (t_ds + t_bh);
      }
      ds_capture_prepare(ctx, ds_pairs.data(), (int64_t)ds_idx.size(), prm.fp64 != 0, NG->ds, share);
    }
    FLK_HIP(hipStreamSynchronize(s));
    hipGraph_t graph = nullptr;
    FLK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    t_capturing = true;
    // A hybrid certificate's direct-sum part (VALU-bound) and its Barnes-Hut
    // batches (memory-bound builds and traversals) share nothing they write:
    // the direct sum is a branch of its own, forked from the capture stream
    // onto the context's second stream and joined before the capture ends.
    const bool fork = !ds_idx.empty() && !NG->batches.empty();
    bool forked = false, joined = false;
    try {
      if (fork) {
        FLK_HIP(hipEventRecord(ctx->ev_fork, s));
        FLK_HIP(hipStreamWaitEvent(ctx->aux, ctx->ev_fork, 0));
        forked = true;
        ds_capture_enqueue(ctx, NG->ds, ctx->aux);
      } else if (!ds_idx.empty()) {
        ds_capture_enqueue(ctx, NG->ds, s);
      }
      for (BatchPlan &bp : NG->batches) {
        enqueue_forest(ctx, bp.forest, s);
        enqueue_traversal(ctx, bp, prm, s);
      }
      if (forked) {
        FLK_HIP(hipEventRecord(ctx->ev_join, ctx->aux));
        FLK_HIP(hipStreamWaitEvent(s, ctx->ev_join, 0));
        joined = true;
      }
    } catch (...) {
      t_capturing = false;
      if (forked && !joined) {  // the capture cannot end with a branch left open
        (void)hipEventRecord(ctx->ev_join, ctx->aux);
        (void)hipStreamWaitEvent(s, ctx->ev_join, 0);
      }
      (void)hipStreamEndCapture(s, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    t_capturing = false;
    FLK_HIP(hipStreamEndCapture(s, &graph));
    const hipError_t ie = hipGraphInstantiate(&NG->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      NG->exec = nullptr;
      throw Error(FLK_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    NG->epoch = g_devbuf_epoch.load();
    ctx->cert_graph = NG;
    G = NG.get();
  } else if (use_mlp) {
    mlp_upload(ctx);  // the weights as they are now (same buffers)
  }
  FLK_HIP(hipEventRecord(ctx->ev0, s));
  const auto t0 = std::chrono::steady_clock::now();
  FLK_HIP(hipGraphLaunch(G->exec, s));
  ctx->stats.ms_graph_launch =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  FLK_HIP(hipEventRecord(ctx->ev1, s));
  FLK_HIP(hipStreamSynchronize(s));
  float ms = 0.f;
  FLK_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->stats.ms_kernel = ms;
  ctx->stats.graph_launches += 1;
  if (!G->ds_idx.empty()) {
    for (size_t t = 0; t < G->ds_idx.size(); ++t) lk[G->ds_idx[t]] = G->ds.h_out.as<double>()[t];
    ctx->stats.ds_seg_pairs += G->ds.work;
    ctx->stats.ds_launches += G->ds.tasks.empty() ? 0 : 1;
    ctx->stats.ds_waves = G->ds.waves;
  }
  // keep the graph alive while batches that overflowed re-run eagerly (their
  // allocations retire it for the next call)
  std::shared_ptr<CertGraph> hold = ctx->cert_graph;
  for (BatchPlan &bp : G->batches) {
    bp.collect();
    if (!finish_batch(ctx, bp, lk, err, flags)) bh_batch_host(ctx, pairs, bp.idx, prm, lk, err, flags);
  }
  return true;
}

}  // namespace

void bh_run(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, const flk_params &prm,
            std::vector<double> &lk, std::vector<double> &err, std::vector<uint32_t> &flags) {
  Curves &cv = ctx->curves;
  lk.assign(npairs, 0.0);
  err.assign(npairs, 0.0);
  flags.assign(npairs, 0u);
  if (npairs == 0) return;
  const bool use_mlp = prm.criterion == FLK_CRIT_MLP;
  if (use_mlp && !ctx->mlp.loaded)
    throw Error(FLK_ERR_STATE, "the learned criterion needs weights (flk_mlp_set_weights)");
  // Pairs below the direct threshold (or with a curve too short for a tree)
  // go to the direct sum; the rest to Barnes–Hut.
  std::vector<int64_t> ds_idx, bh_idx;
  std::vector<int32_t> ds_pairs;
  for (int64_t k = 0; k < npairs; ++k) {
    const int32_t a = pairs[2 * k], b = pairs[2 * k + 1];
    const int64_t work = (int64_t)cv.n[a] * cv.n[b];
    if (cv.n[a] < 2 || cv.n[b] < 2 || work < prm.bh_direct_threshold) {
      ds_idx.push_back(k);
      ds_pairs.push_back(a);
      ds_pairs.push_back(b);
      flags[k] |= FLK_PAIR_DIRECT;
    } else {
      bh_idx.push_back(k);
    }
  }
  // Default: the whole certificate (direct-sum part and every Barnes–Hut
  // batch) as one graph launch.
  if (prm.graph && !prm.timing && !bh_idx.empty() &&
      bh_run_graph(ctx, pairs, npairs, prm, ds_idx, ds_pairs, bh_idx, lk, err, flags))
    return;
  if (!ds_idx.empty()) {
    std::vector<double> v;
    float ms = 0.f;
    ds_run(ctx, ds_pairs.data(), (int64_t)ds_idx.size(), prm.fp64 != 0, v, &ms, prm.graph != 0);
    ctx->stats.ms_ds_kernel += ms;
    for (size_t t = 0; t < ds_idx.size(); ++t) lk[ds_idx[t]] = v[t];
  }
  for (const auto &idx : form_batches(ctx, pairs, bh_idx, prm)) bh_batch(ctx, pairs, idx, prm, lk, err, flags);
}

}  // namespace flk
