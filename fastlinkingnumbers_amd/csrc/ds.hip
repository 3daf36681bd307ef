// Direct-summation tile kernels (gfx950).
//
// The Gauss linking integral of two closed polylines is the sum over segment
// pairs of the signed solid angle of the quadrilateral (a, b, c, d) spanned by
// the relative vectors of their endpoints, split into the two triangles
// (a, b, c) and (a, d, c) with the Van Oosterom–Strackee tangent formula and
// recombined into one arctangent (Arai 2013; reference
// src/compute_link.cpp:63-84 and :151-175).
//
// Work layout.  A wave owns 64*R consecutive rows (segments of the longer
// curve, R per lane, kept in VGPRs) and sweeps a range of columns (segments of
// the other curve).  The column vertices are wave-uniform, so they come
// through the scalar cache into SGPRs: no LDS traffic, no bank conflicts.  The
// grid of relative vectors G(i, j) = q_j - p_i is shared between the four
// segment pairs that touch it, so each pair costs one new grid point (one
// sqrt), one row dot, one column dot, one diagonal dot, the triple product
// and one arctangent.  The row and column dots come from squared norms by
// polarization, G.G' = (|G|^2 + |G'|^2 - |G' - G|^2) / 2, where G' - G is a
// segment's edge: -|edge|^2 / 2 is stored with each edge (computed in FP64),
// so such a dot costs two instructions instead of three.  The diagonal dot
// and the triple product use only the new column: with a = G(k,j),
// c = G(k+1,j+1), d = G(k,j+1) = a + e2, a.c = d.c - e2.c, and the triple
// product is the same for every corner X of the quadrilateral,
// det = X.(e2 x e1) = e1.(X x e2); so a step keeps only the norms of the
// previous column in registers, not its vectors, and one cross product
// w = G(k+1,j+1) x e2 serves the pairs k (corner c) and k+1 (corner d).  Per-lane sums are FP32 over short column runs and FP64
// beyond; the integer branch corrections are exact.
//
// Load balance.  All tasks of all pairs form one list with an integer cost
// prefix; each wave of the persistent grid takes the tasks whose starting cost
// falls in its equal share.  Every task writes its own partial, and a second
// kernel reduces the partials of each pair in a fixed order, so results are
// deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <mutex>

#include "flk_internal.h"

namespace flk {

namespace {

constexpr float kPiF = 3.14159265358979323846f;
constexpr float kHalfPiF = 1.57079632679489661923f;

// atan(t) = t * P(t^2) on [0, 1]; coefficients fitted for this engine by a
// relative-error minimax fit (fit_atan_poly.py, Remez exchange) and rounded
// to FP32: max relative error 1.2e-7 in exact arithmetic, 2.1e-7 evaluated in
// FP32 (tests/test_atan_poly.py).
__device__ __forceinline__ float atan_unit(float t) {
  const float s = t * t;
  float p = -0.004693266935646534f;
  p = fmaf(p, s, 0.024252379313111305f);
  p = fmaf(p, s, -0.05948637053370476f);
  p = fmaf(p, s, 0.09914292395114899f);
  p = fmaf(p, s, -0.1401948183774948f);
  p = fmaf(p, s, 0.19969724118709564f);
  p = fmaf(p, s, -0.33331990242004395f);
  p = fmaf(p, s, 0.9999998807907104f);
  return t * p;
}

// Four-quadrant arctangent with IEEE atan2 sign conventions for signed zeros.
__device__ __forceinline__ float atan2_fast(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  const float t = mn * __builtin_amdgcn_rcpf(fmaxf(mx, 1e-37f));
  float r = atan_unit(t);
  r = (ay > ax) ? (kHalfPiF - r) : r;
  r = (x < 0.f) ? (kPiF - r) : r;
  return copysignf(r, y);
}

__device__ __forceinline__ int64_t lower_bound_i64(const int64_t *__restrict__ a, int64_t n,
                                                   int64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ---------------------------------------------------------------- FP32 tiles
template <int R, bool PRECISE>
__global__ __launch_bounds__(256) void ds_tile_f32(
    const float4 *__restrict__ hi, const float4 *__restrict__ lo,
    const float4 *__restrict__ ed, const DsPair *__restrict__ pairs,
    const DsTask *__restrict__ tasks, const int64_t *__restrict__ cost_prefix,
    int64_t ntasks, double *__restrict__ task_theta, int32_t *__restrict__ task_corr) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
  const int64_t total = cost_prefix[ntasks];
  const int64_t tb = lower_bound_i64(cost_prefix, ntasks, total / nwaves * wave +
                                                              (total % nwaves) * wave / nwaves);
  const int64_t te = lower_bound_i64(cost_prefix, ntasks, total / nwaves * (wave + 1) +
                                                              (total % nwaves) * (wave + 1) / nwaves);

  for (int64_t t = tb; t < te; ++t) {
    const DsTask tk = tasks[t];
    const DsPair pr = pairs[tk.pair];
    const int rbase = tk.row0 + lane * R;

    float px[R + 1], py[R + 1], pz[R + 1];
    float lx[R + 1], ly[R + 1], lz[R + 1];
    float ex[R], ey[R], ez[R], he1[R];
#pragma unroll
    for (int k = 0; k <= R; ++k) {
      const int64_t v = pr.off_r + min(rbase + k, pr.n_r);
      const float4 h = hi[v];
      px[k] = h.x; py[k] = h.y; pz[k] = h.z;
      if (PRECISE) {
        const float4 l = lo[v];
        lx[k] = l.x; ly[k] = l.y; lz[k] = l.z;
      } else {
        lx[k] = ly[k] = lz[k] = 0.f;
      }
    }
    // Rows past the curve end become zero-length segments at the closing
    // vertex with a zero edge: det = 0, y' = +-0, x' >= 0, so they add +-0.
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const bool valid = rbase + k < pr.n_r;
      const float4 e = ed[pr.off_r + min(rbase + k, pr.n_r)];
      ex[k] = valid ? e.x : 0.f;
      ey[k] = valid ? e.y : 0.f;
      ez[k] = valid ? e.z : 0.f;
      he1[k] = valid ? e.w : 0.f;  // -|e1|^2 / 2
    }

    // Relative vectors of one grid column: G(k, j) = q_j - p_k, its squared
    // norm and its norm.
    auto column = [&](const float4 &q, const float4 &ql, float (&gx)[R + 1], float (&gy)[R + 1],
                      float (&gz)[R + 1], float (&gn)[R + 1], float (&gl)[R + 1]) {
#pragma unroll
      for (int k = 0; k <= R; ++k) {
        if constexpr (PRECISE) {
          gx[k] = (q.x - px[k]) + (ql.x - lx[k]);
          gy[k] = (q.y - py[k]) + (ql.y - ly[k]);
          gz[k] = (q.z - pz[k]) + (ql.z - lz[k]);
        } else {
          gx[k] = q.x - px[k];
          gy[k] = q.y - py[k];
          gz[k] = q.z - pz[k];
        }
        gn[k] = fmaf(gx[k], gx[k], fmaf(gy[k], gy[k], gz[k] * gz[k]));
        gl[k] = __builtin_amdgcn_sqrtf(gn[k]);
      }
    };
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    // column vertex indices are clamped to the closing vertex (index n_c of
    // the curve) in 32 bits, which stays on the scalar unit (there is no
    // 64-bit scalar compare)
    const float4 *__restrict__ chi = hi + pr.off_c;
    const float4 *__restrict__ clo = lo + pr.off_c;
    const float4 *__restrict__ ced = ed + pr.off_c;
    const int ncl = pr.n_c;
    float an[R + 1], al[R + 1], acol[R];
    float bn[R + 1], bl[R + 1], bcol[R];
    {
      float gx[R + 1], gy[R + 1], gz[R + 1];
      const int64_t v = pr.off_c + tk.col0;
      column(hi[v], PRECISE ? lo[v] : zero4, gx, gy, gz, an, al);
    }
#pragma unroll
    for (int k = 0; k < R; ++k) acol[k] = fmaf(0.5f, an[k] + an[k + 1], he1[k]);

    // Column vertex j+1 and edge j of the next step, loaded one step ahead
    // (scalar loads; the index is clamped to the closing vertex).
    float4 nq = hi[pr.off_c + tk.col0 + 1];
    float4 nql = PRECISE ? lo[pr.off_c + tk.col0 + 1] : zero4;
    float4 ne = ed[pr.off_c + tk.col0];

    // One column step: the new column (j+1), then the R segment pairs (k, j)
    // of the quadrilaterals a=G(k,j) b=G(k+1,j) c=G(k+1,j+1) d=G(k,j+1).
    auto step = [&](int j, const float (&gn)[R + 1], const float (&gl)[R + 1], const float (&gcol)[R],
                    float (&nn)[R + 1], float (&nl)[R + 1], float (&ncol)[R], float &accf, float &sacc,
                    int &corr) {
      const float4 q = nq, ql = nql, e2 = ne;
      {
        const int w = min(j + 2, ncl);
        nq = chi[w];
        if constexpr (PRECISE) nql = clo[w];
        ne = ced[min(j + 1, ncl)];
      }
      float nx[R + 1], ny[R + 1], nz[R + 1];
      column(q, ql, nx, ny, nz, nn, nl);
      float row[R + 1], diag[R];
      // G(k,j).G(k,j+1) and G(k,j+1).G(k+1,j+1) by polarization (e2.w, he1:
      // -|edge|^2 / 2), the diagonal G(k,j).G(k+1,j+1) as G(k,j+1).G(k+1,j+1)
      // - e2.G(k+1,j+1), where e2.G(k+1,j+1) = |G(k+1,j+1)|^2 - G(k+1,j).G(k+1,j+1)
      // since G(k+1,j) = G(k+1,j+1) - e2 (two operations instead of three)
#pragma unroll
      for (int k = 0; k <= R; ++k) row[k] = fmaf(0.5f, gn[k] + nn[k], e2.w);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        ncol[k] = fmaf(0.5f, nn[k] + nn[k + 1], he1[k]);
        diag[k] = ncol[k] + (row[k + 1] - nn[k + 1]);
      }
      // det of the couple (2c, 2c+1) from one cross product at grid point
      // 2c+1 (computed for all couples first: interleaving them with the
      // denominators measured 2% slower)
      float xp[R], yp[R], s12[R], dt[R], d1[R], d2[R];
#pragma unroll
      for (int c = 0; c < R / 2; ++c) {
        const int k1 = 2 * c + 1;
        const float wx = fmaf(ny[k1], e2.z, -nz[k1] * e2.y);
        const float wy = fmaf(nz[k1], e2.x, -nx[k1] * e2.z);
        const float wz = fmaf(nx[k1], e2.y, -ny[k1] * e2.x);
        dt[k1 - 1] = fmaf(ex[k1 - 1], wx, fmaf(ey[k1 - 1], wy, ez[k1 - 1] * wz));
        dt[k1] = fmaf(ex[k1], wx, fmaf(ey[k1], wy, ez[k1] * wz));
      }
      bool small = true;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const float la = gl[k], lb = gl[k + 1], lc = nl[k + 1], ld = nl[k];
        const float s = fmaf(la, lc, diag[k]);
        const float div1 = fmaf(lb, s, fmaf(gcol[k], lc, row[k + 1] * la));
        const float div2 = fmaf(ld, s, fmaf(row[k], lc, ncol[k] * la));
        d1[k] = div1;
        d2[k] = div2;
        xp[k] = fmaf(div1, div2, -dt[k] * dt[k]);
        s12[k] = div1 + div2;
        yp[k] = dt[k] * s12[k];
        // small angle: no branch correction, |y'/x'| <= 1/8 and x' > 8e-18
        // (per-pair tests interleave with the arithmetic; a max/min fold at
        // the end of the step serialises and measured 5-9% slower)
        small = small && (s12[k] > 0.f) && (fmaf(0.125f, xp[k], -fabsf(yp[k])) > 1e-18f);
      }
      // Two small angles add as one: arg(z1) + arg(z2) = arg(z1 z2) with
      // z = x' + i y', so a couple of pairs needs one reciprocal (the
      // transcendental unit is the scarce one).  Both |y'/x'| <= 1/8 keep the
      // product's real part above 63/64 x1 x2 > 0 and its tangent within
      // 16/63, and both x' > 8e-18 keep it above 6e-35, clear of FP32
      // underflow.
      float zr[R / 2], zi[R / 2];
#pragma unroll
      for (int c = 0; c < R / 2; ++c) {
        zr[c] = fmaf(xp[2 * c], xp[2 * c + 1], -yp[2 * c] * yp[2 * c + 1]);
        zi[c] = fmaf(xp[2 * c], yp[2 * c + 1], yp[2 * c] * xp[2 * c + 1]);
      }
      // every lane on the short path: a scalar test of the lane mask
      if (__builtin_expect(__builtin_amdgcn_ballot_w64(!small) == 0, 1)) {
        // atan(t) = t P(t^2) on |t| <= 16/63: minimax coefficients fitted for
        // this engine (fit_atan_poly.py), rounded to FP32: relative error
        // 1.1e-9 in exact arithmetic (mean 3e-10), 8.9e-8 evaluated in FP32
#pragma unroll
        for (int c = 0; c < R / 2; ++c) {
          const float tt = zi[c] * __builtin_amdgcn_rcpf(zr[c]);
          const float s2 = tt * tt;
          float p = fmaf(s2, 0.09799769520759583f, -0.14216716587543488f);
          p = fmaf(s2, p, 0.19998548924922943f);
          p = fmaf(s2, p, -0.33333322405815125f);
          p = fmaf(s2, p, 1.0f);
          accf = fmaf(tt, p, accf);
        }
      } else {
        // the general path keeps its own accumulators, so the frequent short
        // path merges without register copies
#pragma unroll
        for (int k = 0; k < R; ++k) {
          float x = xp[k], y = yp[k];
          // x' and y' are products of two factors of |G|^3 (denominators, and
          // the triple product times their sum): for relative vectors below
          // ~1e-6 of the unit they underflow FP32 although the factors do not;
          // such a pair is redone from its factors scaled by 2^40 (the angle
          // depends on the ratio only).  Only this rare path pays for it.
          if (fabsf(x) < 0x1p-100f && fabsf(y) < 0x1p-100f) {
            const float a = d1[k] * 0x1p40f, b = d2[k] * 0x1p40f, d = dt[k] * 0x1p40f;
            x = fmaf(a, b, -d * d);
            y = d * (a + b);
          }
          sacc += atan2_fast(y, x);
          corr += (s12[k] < 0.f && dt[k] != 0.f) ? (dt[k] > 0.f ? 1 : -1) : 0;
        }
      }
    };

    double acc = 0.0;
    int corr = 0;
    const int jend = tk.col0 + tk.ncols;
    int j = tk.col0;
    while (j < jend) {
      const int jrun = min(jend, j + 64);
      float accf = 0.f, sacc = 0.f;
      for (; j + 1 < jrun; j += 2) {
        step(j, an, al, acol, bn, bl, bcol, accf, sacc, corr);
        step(j + 1, bn, bl, bcol, an, al, acol, accf, sacc, corr);
      }
      if (j < jrun) {
        step(j, an, al, acol, bn, bl, bcol, accf, sacc, corr);
#pragma unroll
        for (int k = 0; k <= R; ++k) {
          an[k] = bn[k]; al[k] = bl[k];
        }
#pragma unroll
        for (int k = 0; k < R; ++k) acol[k] = bcol[k];
        ++j;
      }
      acc += (double)accf + (double)sacc;
    }
    acc = wave_sum(acc);
    corr = wave_sum(corr);
    if (lane == 0) {
      task_theta[t] = acc;
      task_corr[t] = corr;
    }
  }
}

// ---------------------------------------------------------------- FP64 tiles
// The judge: same tiling, FP64 coordinates, library atan2 and sqrt.
template <int R>
__global__ __launch_bounds__(256) void ds_tile_f64(
    const double4 *__restrict__ pts, const DsPair *__restrict__ pairs,
    const DsTask *__restrict__ tasks, const int64_t *__restrict__ cost_prefix, int64_t ntasks,
    double *__restrict__ task_theta, int32_t *__restrict__ task_corr) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
  const int64_t total = cost_prefix[ntasks];
  const int64_t tb = lower_bound_i64(cost_prefix, ntasks, total / nwaves * wave +
                                                              (total % nwaves) * wave / nwaves);
  const int64_t te = lower_bound_i64(cost_prefix, ntasks, total / nwaves * (wave + 1) +
                                                              (total % nwaves) * (wave + 1) / nwaves);
  for (int64_t t = tb; t < te; ++t) {
    const DsTask tk = tasks[t];
    const DsPair pr = pairs[tk.pair];
    const int rbase = tk.row0 + lane * R;
    double px[R + 1], py[R + 1], pz[R + 1];
    bool valid[R];
#pragma unroll
    for (int k = 0; k <= R; ++k) {
      const double4 h = pts[pr.off_r + min(rbase + k, pr.n_r)];
      px[k] = h.x; py[k] = h.y; pz[k] = h.z;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) valid[k] = rbase + k < pr.n_r;
    double acc = 0.0;
    int corr = 0;
    for (int j = tk.col0; j < tk.col0 + tk.ncols; ++j) {
      const double4 q0 = pts[pr.off_c + j];
      const double4 q1 = pts[pr.off_c + j + 1];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const double ax = q0.x - px[k], ay = q0.y - py[k], az = q0.z - pz[k];
        const double bx = q0.x - px[k + 1], by = q0.y - py[k + 1], bz = q0.z - pz[k + 1];
        const double cx = q1.x - px[k + 1], cy = q1.y - py[k + 1], cz = q1.z - pz[k + 1];
        const double dx = q1.x - px[k], dy = q1.y - py[k], dz = q1.z - pz[k];
        const double la = sqrt(ax * ax + ay * ay + az * az);
        const double lb = sqrt(bx * bx + by * by + bz * bz);
        const double lc = sqrt(cx * cx + cy * cy + cz * cz);
        const double ld = sqrt(dx * dx + dy * dy + dz * dz);
        const double ac = cx * ax + cy * ay + cz * az;
        const double div1 = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + ac * lb +
                            (cx * bx + cy * by + cz * bz) * la;
        const double div2 = la * ld * lc + (ax * dx + ay * dy + az * dz) * lc + ac * ld +
                            (cx * dx + cy * dy + cz * dz) * la;
        const double det = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) +
                           az * (bx * cy - by * cx);
        const double xp = div1 * div2 - det * det;
        const double yp = det * (div1 + div2);
        const double th = atan2(yp, xp);
        const int cr = (yp * det < 0.0) ? (det > 0.0 ? 1 : -1) : 0;
        acc += valid[k] ? th : 0.0;
        corr += valid[k] ? cr : 0;
      }
    }
    acc = wave_sum(acc);
    corr = wave_sum(corr);
    if (lane == 0) {
      task_theta[t] = acc;
      task_corr[t] = corr;
    }
  }
}

// One wave per pair: fixed-order reduction of the pair's task partials.
__global__ __launch_bounds__(256) void ds_pair_reduce(const double *__restrict__ task_theta,
                                                      const int32_t *__restrict__ task_corr,
                                                      const int64_t *__restrict__ pair_task0,
                                                      int64_t npairs, double *__restrict__ lk) {
  const int lane = threadIdx.x & 63;
  const int64_t p = (int64_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (p >= npairs) return;
  double th = 0.0;
  int64_t cr = 0;
  for (int64_t t = pair_task0[p] + lane; t < pair_task0[p + 1]; t += 64) {
    th += task_theta[t];
    cr += task_corr[t];
  }
  th = wave_sum(th);
  cr = wave_sum(cr);
  if (lane == 0) lk[p] = th * 0.15915494309189533576888376337251436 + (double)cr;
}

// Converts the uploaded FP64 vertices (already shifted and scaled) into the
// FP32 hi/lo split and FP32 edge vectors; flags any vertex not exactly
// representable in FP32.
__global__ __launch_bounds__(256) void convert_vertices(const double4 *__restrict__ d, int64_t n,
                                                        float4 *__restrict__ hi,
                                                        float4 *__restrict__ lo,
                                                        float4 *__restrict__ ed,
                                                        int32_t *__restrict__ inexact) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (v < n) {
    const double4 p = d[v];
    const float hx = (float)p.x, hy = (float)p.y, hz = (float)p.z;
    const float lx = (float)(p.x - (double)hx), ly = (float)(p.y - (double)hy),
                lz = (float)(p.z - (double)hz);
    hi[v] = make_float4(hx, hy, hz, 0.f);
    lo[v] = make_float4(lx, ly, lz, 0.f);
    bad = (lx != 0.f) || (ly != 0.f) || (lz != 0.f);
    if (v + 1 < n) {
      const double4 q = d[v + 1];
      const double ex = q.x - p.x, ey = q.y - p.y, ez = q.z - p.z;
      ed[v] = make_float4((float)ex, (float)ey, (float)ez, (float)(-0.5 * (ex * ex + ey * ey + ez * ez)));
    } else {
      ed[v] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(inexact, 1);
}

}  // namespace

void convert_vertices_launch(hipStream_t s, const double4 *d, int64_t n, float4 *hi, float4 *lo,
                             float4 *ed, int32_t *inexact) {
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(convert_vertices, dim3((unsigned)blocks), dim3(256), 0, s, d, n, hi, lo, ed,
                     inexact);
  FLK_HIP(hipGetLastError());
}

// Rows per lane of the FP32 tile kernel, and rows per wave.
constexpr int kRowsPerLane64 = 2;

namespace {

// Blocks of 256 threads per CU at which every block of `kernel` is resident
// (the persistent grid gives each wave an equal share of the work, so blocks
// that wait for a slot would run as a second, under-occupied round).
int resident_blocks(const void *kernel) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0) != hipSuccess || nb < 1) nb = 1;
  return nb;
}

// Host plan of one batched direct sum: pair records (rows = longer curve),
// the tasks of every `stride`-th block of 64R rows (stride 1: all of them),
// their cost prefix and each pair's first task.
struct DsPlan {
  int R = 8, blocks = 0;
  bool fp64 = false, precise = false;
  int64_t work = 0;
  std::vector<DsPair> dp;
  std::vector<DsTask> tasks;
  std::vector<int64_t> prefix, task0;
};

// share: the fraction of the resident blocks per CU to launch when the grid
// is sized from residency (a direct sum that runs beside other kernels).
void ds_plan(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, bool fp64, const flk_tune &t, int64_t stride,
             DsPlan &P, double share = 1.0) {
  Curves &cv = ctx->curves;
  P.fp64 = fp64;
  P.precise = cv.inexact;
  P.R = fp64 ? kRowsPerLane64 : (t.ds_rows_per_lane == 2 || t.ds_rows_per_lane == 4 ? t.ds_rows_per_lane : 8);
  const int rows_per_wave = 64 * P.R;
  int bpc = t.ds_blocks_per_cu;
  if (bpc <= 0) {
    static int cached[7] = {0, 0, 0, 0, 0, 0, 0};  // f64; f32 R2, R4, R8 x exact/precise
    const int slot = fp64 ? 0 : 1 + 2 * (P.R == 2 ? 0 : P.R == 4 ? 1 : 2) + (P.precise ? 1 : 0);
    if (!cached[slot]) {
      const void *k = fp64        ? (const void *)ds_tile_f64<kRowsPerLane64>
                      : P.R == 2 ? (P.precise ? (const void *)ds_tile_f32<2, true> : (const void *)ds_tile_f32<2, false>)
                      : P.R == 4 ? (P.precise ? (const void *)ds_tile_f32<4, true> : (const void *)ds_tile_f32<4, false>)
                                 : (P.precise ? (const void *)ds_tile_f32<8, true> : (const void *)ds_tile_f32<8, false>);
      cached[slot] = resident_blocks(k);
    }
    bpc = std::max(1, std::min(cached[slot], (int)std::lround(cached[slot] * share)));
  }
  P.blocks = std::max(1, ctx->num_cus) * bpc;
  const int64_t nwaves = (int64_t)P.blocks * 4;

  P.dp.assign(npairs, DsPair{});
  std::vector<std::pair<int32_t, int32_t>> rowblk;  // (pair, first row) of the blocks run
  int64_t wave_cost = 0, blk = 0;
  P.work = 0;
  for (int64_t k = 0; k < npairs; ++k) {
    int a = pairs[2 * k], b = pairs[2 * k + 1];
    if (cv.n[a] < cv.n[b]) std::swap(a, b);  // rows = longer curve
    P.dp[k] = DsPair{cv.off[a], cv.off[b], cv.n[a], cv.n[b]};
    if (cv.n[a] == 0 || cv.n[b] == 0) continue;
    for (int32_t r0 = 0; r0 < cv.n[a]; r0 += rows_per_wave, ++blk) {
      if (blk % stride) continue;
      rowblk.emplace_back((int32_t)k, r0);
      wave_cost += (int64_t)rows_per_wave * cv.n[b];
      P.work += (int64_t)std::min(rows_per_wave, cv.n[a] - r0) * cv.n[b];
    }
  }
  // Column chunk: aim for ds_tasks_per_wave tasks per wave over the batch,
  // never below 256 columns.
  const int64_t tpw = std::max(1, t.ds_tasks_per_wave);
  const int64_t target_task = std::max<int64_t>(wave_cost / (nwaves * tpw), (int64_t)rows_per_wave * 256);
  const int64_t chunk_cols = std::max<int64_t>(256, target_task / rows_per_wave);
  P.tasks.clear();
  P.prefix.clear();
  P.task0.assign(npairs + 1, 0);
  P.tasks.reserve(rowblk.size() * 2);
  P.prefix.reserve(rowblk.size() * 2 + 1);
  int64_t acc = 0;
  size_t rb = 0;
  for (int64_t k = 0; k < npairs; ++k) {
    P.task0[k] = (int64_t)P.tasks.size();
    for (; rb < rowblk.size() && rowblk[rb].first == k; ++rb) {
      const DsPair &p = P.dp[k];
      for (int64_t c0 = 0; c0 < p.n_c; c0 += chunk_cols) {
        const int32_t nc = (int32_t)std::min<int64_t>(chunk_cols, p.n_c - c0);
        P.tasks.push_back(DsTask{(int32_t)k, rowblk[rb].second, (int32_t)c0, nc});
        P.prefix.push_back(acc);
        acc += (int64_t)nc * rows_per_wave;
      }
    }
  }
  P.task0[npairs] = (int64_t)P.tasks.size();
  P.prefix.push_back(acc);
}

// Device buffers of one batch.
struct DsBufs {
  DsPair *pairs;
  DsTask *tasks;
  int64_t *cost, *task0;
  double *theta, *out;
  int32_t *corr;
};

DsBufs reserve_bufs(DevBuf &pairs, DevBuf &tasks, DevBuf &cost, DevBuf &theta, DevBuf &corr, DevBuf &task0,
                    DevBuf &out, size_t np, size_t nt) {
  pairs.reserve(sizeof(DsPair) * np);
  tasks.reserve(sizeof(DsTask) * nt);
  cost.reserve(sizeof(int64_t) * (nt + 1));
  theta.reserve(sizeof(double) * nt);
  corr.reserve(sizeof(int32_t) * nt);
  task0.reserve(sizeof(int64_t) * (np + 1));
  out.reserve(sizeof(double) * np);
  return DsBufs{pairs.as<DsPair>(), tasks.as<DsTask>(), cost.as<int64_t>(), task0.as<int64_t>(),
                theta.as<double>(), out.as<double>(), corr.as<int32_t>()};
}

// Uploads of the plan arrays, the tile kernel (between ev_a and ev_b when
// given) and, with `out`, the per-pair reduction and read-back: one stream-
// ordered sequence, launched eagerly or captured into a graph.
void ds_upload(const DsPair *dp, const DsTask *tasks, const int64_t *prefix, const int64_t *task0, int64_t npairs,
               int64_t ntasks, const DsBufs &d, hipStream_t s) {
  FLK_HIP(hipMemcpyAsync(d.pairs, dp, sizeof(DsPair) * npairs, hipMemcpyHostToDevice, s));
  FLK_HIP(hipMemcpyAsync(d.tasks, tasks, sizeof(DsTask) * ntasks, hipMemcpyHostToDevice, s));
  FLK_HIP(hipMemcpyAsync(d.cost, prefix, sizeof(int64_t) * (ntasks + 1), hipMemcpyHostToDevice, s));
  FLK_HIP(hipMemcpyAsync(d.task0, task0, sizeof(int64_t) * (npairs + 1), hipMemcpyHostToDevice, s));
}

// With dp == nullptr the plan is already on the device (a graph uploads it
// once, at capture).
void ds_enqueue(flk_ctx *ctx, const DsPlan &P, const DsPair *dp, const DsTask *tasks, const int64_t *prefix,
                const int64_t *task0, int64_t npairs, int64_t ntasks, const DsBufs &d, hipStream_t s,
                hipEvent_t ev_a, hipEvent_t ev_b, double *out) {
  Curves &cv = ctx->curves;
  if (dp) ds_upload(dp, tasks, prefix, task0, npairs, ntasks, d, s);
  if (ev_a) FLK_HIP(hipEventRecord(ev_a, s));
  const float4 *hi = cv.hi.as<float4>(), *lo = cv.lo.as<float4>(), *ed = cv.ed.as<float4>();
  const dim3 grid(P.blocks), block(256);
  if (P.fp64) {
    hipLaunchKernelGGL((ds_tile_f64<kRowsPerLane64>), grid, block, 0, s, cv.dpts.as<double4>(), d.pairs, d.tasks,
                       d.cost, ntasks, d.theta, d.corr);
  } else if (P.R == 2) {
    if (P.precise)
      hipLaunchKernelGGL((ds_tile_f32<2, true>), grid, block, 0, s, hi, lo, ed, d.pairs, d.tasks, d.cost, ntasks,
                         d.theta, d.corr);
    else
      hipLaunchKernelGGL((ds_tile_f32<2, false>), grid, block, 0, s, hi, lo, ed, d.pairs, d.tasks, d.cost, ntasks,
                         d.theta, d.corr);
  } else if (P.R == 4) {
    if (P.precise)
      hipLaunchKernelGGL((ds_tile_f32<4, true>), grid, block, 0, s, hi, lo, ed, d.pairs, d.tasks, d.cost, ntasks,
                         d.theta, d.corr);
    else
      hipLaunchKernelGGL((ds_tile_f32<4, false>), grid, block, 0, s, hi, lo, ed, d.pairs, d.tasks, d.cost, ntasks,
                         d.theta, d.corr);
  } else {
    if (P.precise)
      hipLaunchKernelGGL((ds_tile_f32<8, true>), grid, block, 0, s, hi, lo, ed, d.pairs, d.tasks, d.cost, ntasks,
                         d.theta, d.corr);
    else
      hipLaunchKernelGGL((ds_tile_f32<8, false>), grid, block, 0, s, hi, lo, ed, d.pairs, d.tasks, d.cost, ntasks,
                         d.theta, d.corr);
  }
  FLK_HIP(hipGetLastError());
  if (ev_b) FLK_HIP(hipEventRecord(ev_b, s));
  if (out) {
    const int64_t rblocks = (npairs * 64 + 255) / 256;
    hipLaunchKernelGGL(ds_pair_reduce, dim3((unsigned)rblocks), dim3(256), 0, s, d.theta, d.corr, d.task0,
                       npairs, d.out);
    FLK_HIP(hipGetLastError());
    FLK_HIP(hipMemcpyAsync(out, d.out, sizeof(double) * npairs, hipMemcpyDeviceToHost, s));
  }
}

// Eager run of a plan on the context's shared buffers; returns the tile
// kernel's device time (ms).
float ds_eager(flk_ctx *ctx, const DsPlan &P, std::vector<double> *lk_out) {
  if (P.tasks.empty()) return 0.f;
  const int64_t np = (int64_t)P.dp.size(), nt = (int64_t)P.tasks.size();
  const DsBufs d = reserve_bufs(ctx->d_pairs, ctx->d_tasks, ctx->d_cost, ctx->d_task_theta, ctx->d_task_corr,
                                ctx->d_pair_task0, ctx->d_pair_out, np, nt);
  hipStream_t s = ctx->stream;
  ds_enqueue(ctx, P, P.dp.data(), P.tasks.data(), P.prefix.data(), P.task0.data(), np, nt, d, s, ctx->ev0,
             ctx->ev1, lk_out ? lk_out->data() : nullptr);
  FLK_HIP(hipStreamSynchronize(s));
  float ms = 0.f;
  FLK_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  return ms;
}

uint64_t fnv1a(uint64_t h, const void *p, size_t n) {
  const unsigned char *c = static_cast<const unsigned char *>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 0x100000001b3ULL;
  return h;
}

// Graph run: the batch is captured once per (pairs, shape, precision, curve
// set) and replayed after that.  The graph is timed whole (its uploads and
// the reduction add well under a percent to a batch worth a launch).
float ds_graph_run(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, bool fp64, std::vector<double> &lk_out,
                   int64_t *work) {
  const flk_tune &t = ctx->tune;
  uint64_t key = fnv1a(0xcbf29ce484222325ULL, pairs, sizeof(int32_t) * 2 * (size_t)npairs);
  const int64_t meta[7] = {npairs, t.ds_rows_per_lane, t.ds_blocks_per_cu, t.ds_tasks_per_wave, fp64 ? 1 : 0,
                           (int64_t)ctx->curves_version, ctx->curves.inexact ? 1 : 0};
  key = fnv1a(key, meta, sizeof meta);
  DsGraph &G = ctx->ds_graph[fp64 ? 1 : 0];
  hipStream_t s = ctx->stream;
  if (!(G.exec && G.key == key)) {
    G.reset();
    DsPlan P;
    ds_plan(ctx, pairs, npairs, fp64, t, 1, P);
    if (P.tasks.empty()) {
      lk_out.assign(npairs, 0.0);
      *work = 0;
      return 0.f;
    }
    // the graph's copy nodes read the host arrays kept in G
    G.dp = std::move(P.dp);
    G.tasks = std::move(P.tasks);
    G.prefix = std::move(P.prefix);
    G.task0 = std::move(P.task0);
    G.out.assign(npairs, 0.0);
    G.work = P.work;
    G.waves = P.blocks * 4;
    const int64_t nt = (int64_t)G.tasks.size();
    const DsBufs d =
        reserve_bufs(G.d_pairs, G.d_tasks, G.d_cost, G.d_theta, G.d_corr, G.d_task0, G.d_out, npairs, nt);
    G.h_out.reserve(sizeof(double) * npairs);
    // the plan goes to the device once; the graph reads it there
    ds_upload(G.dp.data(), G.tasks.data(), G.prefix.data(), G.task0.data(), npairs, nt, d, s);
    FLK_HIP(hipStreamSynchronize(s));
    FLK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipGraph_t graph = nullptr;
    try {
      ds_enqueue(ctx, P, nullptr, nullptr, nullptr, nullptr, npairs, nt, d, s, nullptr, nullptr,
                 G.h_out.as<double>());
    } catch (...) {
      (void)hipStreamEndCapture(s, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    FLK_HIP(hipStreamEndCapture(s, &graph));
    const hipError_t ie = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      G.exec = nullptr;
      throw Error(FLK_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    G.key = key;
  }
  FLK_HIP(hipEventRecord(ctx->ev0, s));
  FLK_HIP(hipGraphLaunch(G.exec, s));
  FLK_HIP(hipEventRecord(ctx->ev1, s));
  FLK_HIP(hipStreamSynchronize(s));
  float ms = 0.f;
  FLK_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  lk_out.assign(G.h_out.as<double>(), G.h_out.as<double>() + npairs);
  *work = G.work;
  ctx->stats.ds_waves = G.waves;
  return ms;
}

}  // namespace

void ds_capture_prepare(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, bool fp64, DsGraph &G, double share) {
  DsPlan P;
  ds_plan(ctx, pairs, npairs, fp64, ctx->tune, 1, P, share);
  G.dp = std::move(P.dp);
  G.tasks = std::move(P.tasks);
  G.prefix = std::move(P.prefix);
  G.task0 = std::move(P.task0);
  G.out.assign(npairs, 0.0);
  G.work = P.work;
  G.waves = P.blocks * 4;
  G.npairs = npairs;
  G.R = P.R;
  G.blocks = P.blocks;
  G.fp64 = P.fp64;
  G.precise = P.precise;
  G.h_out.reserve(sizeof(double) * std::max<int64_t>(npairs, 1));
  std::fill(G.h_out.as<double>(), G.h_out.as<double>() + npairs, 0.0);
  if (!G.tasks.empty()) {
    const DsBufs d = reserve_bufs(G.d_pairs, G.d_tasks, G.d_cost, G.d_theta, G.d_corr, G.d_task0, G.d_out, npairs,
                                  G.tasks.size());
    ds_upload(G.dp.data(), G.tasks.data(), G.prefix.data(), G.task0.data(), npairs, (int64_t)G.tasks.size(), d,
              ctx->stream);
  }
}

void ds_capture_enqueue(flk_ctx *ctx, DsGraph &G, hipStream_t s) {
  if (G.tasks.empty()) return;  // G.out stays zero: no segment pairs
  DsPlan P;
  P.R = G.R;
  P.blocks = G.blocks;
  P.fp64 = G.fp64;
  P.precise = G.precise;
  const int64_t nt = (int64_t)G.tasks.size();
  const DsBufs d = DsBufs{G.d_pairs.as<DsPair>(), G.d_tasks.as<DsTask>(), G.d_cost.as<int64_t>(),
                          G.d_task0.as<int64_t>(), G.d_theta.as<double>(), G.d_out.as<double>(),
                          G.d_corr.as<int32_t>()};
  ds_enqueue(ctx, P, nullptr, nullptr, nullptr, nullptr, G.npairs, nt, d, s, nullptr, nullptr, G.h_out.as<double>());
}

void ds_run(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, bool fp64,
            std::vector<double> &lk_out, float *ms_kernel, bool use_graph) {
  lk_out.assign(npairs, 0.0);
  float ms = 0.f;
  int64_t work = 0;
  if (npairs > 0) {
    if (use_graph) {
      ms = ds_graph_run(ctx, pairs, npairs, fp64, lk_out, &work);
    } else {
      DsPlan P;
      ds_plan(ctx, pairs, npairs, fp64, ctx->tune, 1, P);
      work = P.work;
      ms = ds_eager(ctx, P, &lk_out);
      ctx->stats.ds_waves = P.blocks * 4;
    }
    ctx->stats.ds_launches += 1;
  }
  if (ms_kernel) *ms_kernel = ms;
  ctx->stats.ds_seg_pairs += work;
}

// ---------------------------------------------------------------- without allocation
// One batch's device arrays (pair records, per-pair first task and sum, the
// tasks, their cost prefix and partials) carved from a fixed array, for
// callers that cannot allocate.  Every task writes its own partial and the
// reduction writes every pair's sum, so nothing needs clearing between calls.
namespace {
constexpr size_t align256(size_t n) { return (n + 255) & ~size_t(255); }
constexpr size_t kNaPairs = 0;
constexpr size_t kNaTask0 = align256(kNaPairs + sizeof(DsPair) * FLK_NOALLOC_MAX_PAIRS);
constexpr size_t kNaOut = align256(kNaTask0 + sizeof(int64_t) * (FLK_NOALLOC_MAX_PAIRS + 1));
constexpr size_t kNaTasks = align256(kNaOut + sizeof(double) * FLK_NOALLOC_MAX_PAIRS);
constexpr size_t kNaCost = align256(kNaTasks + sizeof(DsTask) * FLK_NOALLOC_MAX_TASKS);
constexpr size_t kNaTheta = align256(kNaCost + sizeof(int64_t) * (FLK_NOALLOC_MAX_TASKS + 1));
constexpr size_t kNaCorr = align256(kNaTheta + sizeof(double) * FLK_NOALLOC_MAX_TASKS);
constexpr size_t kNaBytes = align256(kNaCorr + sizeof(int32_t) * FLK_NOALLOC_MAX_TASKS);
static_assert(kNaBytes <= (size_t)FLK_NOALLOC_SCRATCH_MB << 20, "scratch layout exceeds its stated size");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Each GPU holds its own instance of this array; host code reaches it only
// through the per-device address from hipGetSymbolAddress (the array's name
// is a host-side stand-in in host code).
static __device__ __align__(256) unsigned char scratch[kNaBytes];

void *noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(FLK_ERR_DEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(FLK_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

void ds_run_noalloc(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, bool fp64, std::vector<double> &lk_out,
                    float *ms_kernel) {
  lk_out.assign(npairs, 0.0);
  if (ms_kernel) *ms_kernel = 0.f;
  if (npairs == 0) return;
  if (npairs > FLK_NOALLOC_MAX_PAIRS)
    throw Error(FLK_ERR_ARG, "a batch without allocation holds at most " + std::to_string(FLK_NOALLOC_MAX_PAIRS) +
                                 " pairs (got " + std::to_string(npairs) + ")");
  unsigned char *base = static_cast<unsigned char *>(ctx->noalloc_scratch);
  if (!base) throw Error(FLK_ERR_STATE, "no scratch array on this device");
  DsPlan P;
  ds_plan(ctx, pairs, npairs, fp64, ctx->tune, 1, P);
  const int64_t nt = (int64_t)P.tasks.size();
  if (nt > FLK_NOALLOC_MAX_TASKS)
    throw Error(FLK_ERR_ARG, "a batch without allocation holds at most " + std::to_string(FLK_NOALLOC_MAX_TASKS) +
                                 " tasks (this one needs " + std::to_string(nt) + ")");
  ctx->stats.ds_seg_pairs += P.work;
  ctx->stats.ds_launches += 1;
  ctx->stats.ds_waves = P.blocks * 4;
  if (nt == 0) return;
  const DsBufs d{reinterpret_cast<DsPair *>(base + kNaPairs), reinterpret_cast<DsTask *>(base + kNaTasks),
                 reinterpret_cast<int64_t *>(base + kNaCost),  reinterpret_cast<int64_t *>(base + kNaTask0),
                 reinterpret_cast<double *>(base + kNaTheta),  reinterpret_cast<double *>(base + kNaOut),
                 reinterpret_cast<int32_t *>(base + kNaCorr)};
  hipStream_t s = ctx->stream;
  ds_enqueue(ctx, P, P.dp.data(), P.tasks.data(), P.prefix.data(), P.task0.data(), npairs, nt, d, s, ctx->ev0,
             ctx->ev1, lk_out.data());
  FLK_HIP(hipStreamSynchronize(s));
  float ms = 0.f;
  FLK_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (ms_kernel) *ms_kernel = ms;
}

float ds_time_sample(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, const flk_tune &t, int64_t stride,
                     int64_t *work) {
  DsPlan P;
  ds_plan(ctx, pairs, npairs, false, t, std::max<int64_t>(1, stride), P);
  if (work) *work = P.work;
  return ds_eager(ctx, P, nullptr);
}

int64_t ds_work(flk_ctx *ctx, const int32_t *pairs, int64_t npairs) {
  const Curves &cv = ctx->curves;
  int64_t w = 0;
  for (int64_t k = 0; k < npairs; ++k) w += (int64_t)cv.n[pairs[2 * k]] * cv.n[pairs[2 * k + 1]];
  return w;
}

}  // namespace flk
