// Discretization of spline curves into polylines (reference
// src/discretize.cpp:102-281), restated for this engine.
//
// Each round, the live pieces (curve, segment, t_lo, t_hi) of every
// potentially linked curve pair are tested for box overlap with a
// sweep-and-prune along x (pieces sorted by box min x; a piece meets only the
// pieces of the other curve still active in the sweep).  Overlapping pieces
// are bisected at the parameter midpoint; untouched pieces are retired.  The
// round loop stops when no overlap is left, or fails after 55 rounds or when a
// piece's box diagonal falls below eps * mean |q| / sqrt(3), as the reference
// does.  The retired pieces of each curve, sorted by (segment, t_lo), give the
// polyline's vertices.
#include <algorithm>
#include <atomic>
#include <iostream>
#include <limits>
#include <tuple>

#include "host_internal.h"

namespace VerifyCurves {

namespace {

struct Piece {
  int seg;
  double lo, hi;
  Box3d box;
};

// Sweep-and-prune overlap of two piece lists (indices sorted by box min x);
// marks overlapping pieces on both sides and returns the number of pairs.
int64_t sweep(const std::vector<Piece> &A, const std::vector<int> &ia, const std::vector<Piece> &B,
              const std::vector<int> &ib, std::vector<std::atomic<char>> &hitA,
              std::vector<std::atomic<char>> &hitB) {
  int64_t count = 0;
  std::vector<int> activeA, activeB;
  size_t a = 0, b = 0;
  auto prune = [](std::vector<int> &act, const std::vector<Piece> &P, double x) {
    size_t w = 0;
    for (size_t k = 0; k < act.size(); ++k)
      if (!(P[act[k]].box.max()[0] < x)) act[w++] = act[k];
    act.resize(w);
  };
  while (a < ia.size() || b < ib.size()) {
    const bool takeA =
        b >= ib.size() || (a < ia.size() && A[ia[a]].box.min()[0] <= B[ib[b]].box.min()[0]);
    if (takeA) {
      const Piece &p = A[ia[a]];
      prune(activeB, B, p.box.min()[0]);
      for (int k : activeB)
        if (p.box.intersects(B[k].box)) {
          ++count;
          hitA[ia[a]].store(1, std::memory_order_relaxed);
          hitB[k].store(1, std::memory_order_relaxed);
        }
      activeA.push_back(ia[a]);
      ++a;
    } else {
      const Piece &p = B[ib[b]];
      prune(activeA, A, p.box.min()[0]);
      for (int k : activeA)
        if (p.box.intersects(A[k].box)) {
          ++count;
          hitB[ib[b]].store(1, std::memory_order_relaxed);
          hitA[k].store(1, std::memory_order_relaxed);
        }
      activeB.push_back(ib[b]);
      ++b;
    }
  }
  return count;
}

}  // namespace

std::vector<Curve> Discretize(const std::vector<Curve> &curves,
                              const std::vector<std::vector<int>> &links, bool verbose) {
  const int nc = (int)curves.size();
  const double eps = std::numeric_limits<double>::epsilon();
  std::vector<std::vector<Piece>> live(nc);
  int64_t total = 0;
  double coord_sum = 0.0, min_diag = std::numeric_limits<double>::infinity();
  int min_c = -1, min_s = -1;
  for (int i = 0; i < nc; ++i) {
    const int ns = curves[i].num_segments();
    for (int s = 0; s < ns; ++s) {
      Box3d b = curves[i].get_segment_bounding_box(s);
      live[i].push_back({s, 0.0, 1.0, b});
      coord_sum += curves[i].get_q()[s].norm();
      const double d = b.sizes().norm();
      if (!(d >= min_diag)) {
        min_diag = d;
        min_c = i;
        min_s = s;
      }
    }
    total += std::max(ns, 0);
  }
  const double threshold = eps * (coord_sum / (total * 1.73205080757));
  if (!(min_diag >= threshold)) {
    std::cerr << "Discretize: Smallest input segment, (" << min_c << "," << min_s
              << ") has a size of " << min_diag
              << ", which is smaller than eps * avg coordinate. Consider combining them into "
                 "one vertex."
              << std::endl;
    throw std::runtime_error(
        "Discretize: Smallest input segment is smaller than eps * avg coordinate. Consider "
        "combining.");
  }
  // unordered candidate pairs from the both-direction lists
  std::vector<std::pair<int, int>> cand;
  for (int i = 0; i < nc; ++i)
    for (int j : links[i])
      if (i < j) cand.emplace_back(i, j);

  std::vector<std::vector<std::tuple<int, double, double>>> retired(nc);
  int64_t overlaps = 0;
  for (int round = 0; round < 55; ++round) {
    std::vector<std::vector<int>> order(nc);
    std::vector<std::vector<std::atomic<char>>> hit(nc);
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < nc; ++i) {
      order[i].resize(live[i].size());
      for (size_t k = 0; k < order[i].size(); ++k) order[i][k] = (int)k;
      std::sort(order[i].begin(), order[i].end(), [&](int a, int b) {
        return live[i][a].box.min()[0] < live[i][b].box.min()[0];
      });
      hit[i] = std::vector<std::atomic<char>>(live[i].size());
      for (auto &h : hit[i]) h.store(0, std::memory_order_relaxed);
    }
    overlaps = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : overlaps)
    for (size_t k = 0; k < cand.size(); ++k) {
      const int i = cand[k].first, j = cand[k].second;
      overlaps += sweep(live[i], order[i], live[j], order[j], hit[i], hit[j]);
    }
    std::vector<std::vector<Piece>> next(nc);
    bool too_small = false;
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < nc; ++i) {
      double md = std::numeric_limits<double>::infinity();
      for (size_t k = 0; k < live[i].size(); ++k) {
        const Piece &p = live[i][k];
        if (hit[i][k].load(std::memory_order_relaxed)) {
          const double mid = 0.5 * (p.lo + p.hi);
          Box3d b1 = curves[i].get_segment_bounding_box(p.seg, p.lo, mid);
          Box3d b2 = curves[i].get_segment_bounding_box(p.seg, mid, p.hi);
          md = std::min(md, std::min(b1.sizes().norm(), b2.sizes().norm()));
          next[i].push_back({p.seg, p.lo, mid, b1});
          next[i].push_back({p.seg, mid, p.hi, b2});
        } else {
          retired[i].emplace_back(p.seg, p.lo, p.hi);
        }
      }
      if (!(md >= threshold)) {
#pragma omp critical
        too_small = true;
      }
    }
    if (too_small) {
      std::cout << "Discretize: Smallest segment is smaller than eps * avg coordinate. Perhaps "
                   "curves are too close."
                << std::endl;
      throw std::runtime_error(
          "Discretize: Smallest segment is smaller than eps * avg coordinate. Perhaps curves are "
          "too close.");
    }
    if (verbose) std::cout << "Iteration " << round + 1 << ", " << overlaps << " overlaps." << std::endl;
    if (overlaps == 0) {
      if (verbose) std::cout << "Finished after " << round + 1 << " iteration(s)." << std::endl;
      break;
    }
    live.swap(next);
  }
  if (overlaps > 0) {
    std::cout << "Max recursion (curves are too close)." << std::endl;
    std::cout << "Number of overlaps: " << overlaps << std::endl;
    throw std::runtime_error("Reached max recursion (curves are too close) in Discretize().");
  }
  std::vector<Curve> out(nc);
  int64_t count = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : count)
  for (int i = 0; i < nc; ++i) {
    Curve &o = out[i];
    o.set_looped(true);
    o.set_segment_type(SegmentType::Polyline);
    auto &r = retired[i];
    std::sort(r.begin(), r.end());
    o.get_q().reserve(r.size() + 1);
    for (const auto &p : r) o.get_q().push_back(curves[i].get_point(std::get<0>(p), std::get<1>(p)));
    if (!curves[i].is_looped() && !r.empty()) {
      o.get_q().push_back(curves[i].get_point(std::get<0>(r.back()), std::get<2>(r.back())));
      o.set_looped(false);
    }
    o.ComputeCurveBoundingBox();
    count += (int64_t)r.size();
  }
  if (verbose)
    std::cout << "Original segment count: " << total << ", discretized count: " << count
              << std::endl;
  return out;
}

}  // namespace VerifyCurves
