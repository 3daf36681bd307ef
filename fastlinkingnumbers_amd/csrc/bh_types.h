// Barnes–Hut node records and the learned criterion's node-pair features.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace flk {

struct BhGeo {      // 64 bytes: what the classification reads
  float cx, cy, cz, radius;   // centre (hi part) and half box diagonal
  float lx, ly, lz, rsq;      // centre lo part, radius^2 as 0.25 |diag|^2
  int32_t c0, c1;             // children (c0 < 0: leaf bucket)
  int32_t seg0, nseg;         // range in the sorted segment order
  float cn, dn, qn, pad;      // moment norms |c|, |D|_F/sqrt3, sqrt(sum |Q_k|^2)/3
};
struct BhMom {      // 128 bytes: read by far-field evaluations
  float c[3];
  float D[9];       // D[k*3+j] = int r'_k (r - ctr)_j
  float Q[18];      // per k: xx xy xz yy yz zz of int r'_k (r-ctr)_i (r-ctr)_j
  float pad[2];
};
struct BhItem {
  int32_t pair, a, b, pad;
};

// A leaf of one segment stores no BhMom record (half the nodes of a tree of
// one-segment leaves; their records were a third of the bytes the tree build
// writes).  Its moments about its own centre, the segment's midpoint, are
// c = t, D = 0 and Q_k = t_k t t^T / 12 with t the segment's FP32 tangent,
// so its geometry record carries t in three slots such a leaf does not
// otherwise use -- c1 (the bits of t.x), pad (t.y) and dn (t.z; its true dn
// is 0) -- and c0 = kLeafSegment marks it (c0 < 0 still means "leaf").
constexpr int32_t kLeafSegment = -2;
__host__ __device__ __forceinline__ bool is_leaf_segment(const BhGeo &g) { return g.c0 == kLeafSegment; }
__host__ __device__ __forceinline__ float geo_dn(const BhGeo &g) { return is_leaf_segment(g) ? 0.f : g.dn; }
__host__ __device__ __forceinline__ int sym_idx(int i, int j) {  // symmetric 3x3 -> 6: xx xy xz yy yz zz
  return i <= j ? (i == 0 ? j : (i == 1 ? 2 + j : 5)) : (j == 0 ? i : (j == 1 ? 2 + i : 5));
}
// The moment record of a one-segment leaf, bit for bit what summing its
// single segment gives (the midpoint offset is exactly zero).
__host__ __device__ __forceinline__ BhMom leaf_segment_mom(const BhGeo &g) {
  float t[3];
  t[0] = __builtin_bit_cast(float, g.c1);
  t[1] = g.pad;
  t[2] = g.dn;
  BhMom m;
  for (int i = 0; i < 3; ++i) m.c[i] = t[i];
  for (int k = 0; k < 9; ++k) m.D[k] = 0.f;
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j)
        m.Q[6 * k + sym_idx(i, j)] = fmaf(t[k], fmaf((1.0f / 12.0f) * t[i], t[j], 0.f), 0.f);
  m.pad[0] = m.pad[1] = 0.f;
  return m;
}
__device__ __forceinline__ BhMom node_mom(const BhMom *__restrict__ mom, int32_t id, const BhGeo &g) {
  return is_leaf_segment(g) ? leaf_segment_mom(g) : mom[id];
}

// Items of a level whose count was produced on the device, for a kernel whose
// buffer holds `bound` items: a count above the bound means the producer
// overflowed the buffer (its last slot may be unwritten), and the level is
// skipped (0 items); the host sees the count and discards the pass.
__device__ __forceinline__ int64_t level_count(const unsigned long long *n_dev, int64_t bound) {
  const unsigned long long c = *n_dev;
  return c <= (unsigned long long)bound ? (int64_t)c : 0;
}

// Natural log of a positive normal float (every feature argument is at least
// 1e-12 or a segment count): the hardware log2 times ln 2, without the
// denormal range reduction of the general logf.
__device__ __forceinline__ float ln_normal(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994530942f; }

// 12 dimensionless features of a node pair at centre distance d: both radii,
// their sum, and the monopole/dipole/quadrupole norms scaled by d^-1, d^-2,
// d^-3, the segment counts and the Eq. 14 error estimate of the reference
// (compute_link.cpp:517-522) in linking-number units (all natural logs,
// floored at log 1e-12).  Scale-free, so a network trained on one model size
// applies to another.  kLog2: base-2 logs instead (the inference kernel folds
// the factor ln 2 into its input normalisation).
template <bool kLog2>
__device__ __forceinline__ void mlp_features_t(const BhGeo &A, const BhGeo &B, float dsq, float *f) {
  auto lg = [](float x) { return kLog2 ? __builtin_amdgcn_logf(x) : ln_normal(x); };
  const float id = __builtin_amdgcn_rsqf(dsq), id2 = id * id;
  const float tiny = 1e-12f;
  const float adn = geo_dn(A), bdn = geo_dn(B);
  f[0] = lg(A.radius * id + tiny);
  f[1] = lg(B.radius * id + tiny);
  f[2] = lg((A.radius + B.radius) * id + tiny);
  f[3] = lg(A.cn * id + tiny);
  f[4] = lg(adn * id2 + tiny);
  f[5] = lg(A.qn * id2 * id + tiny);
  f[6] = lg(B.cn * id + tiny);
  f[7] = lg(bdn * id2 + tiny);
  f[8] = lg(B.qn * id2 * id + tiny);
  f[9] = lg((float)A.nseg);
  f[10] = lg((float)B.nseg);
  const float est = 0.0795774715f * id2 * id2 * id *
                    (A.qn * (A.radius * B.cn + 3.0f * bdn) + B.qn * (B.radius * A.cn + 3.0f * adn));
  f[11] = lg(est + tiny);
}
__device__ __forceinline__ void mlp_features(const BhGeo &A, const BhGeo &B, float dsq, float *f) {
  mlp_features_t<false>(A, B, dsq, f);
}

}  // namespace flk
