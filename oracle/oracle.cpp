/*
 * CPU ORACLE — test infrastructure only (see oracle.h).  FP64 restatement of
 * the reference's algorithms; compiled with -ffp-contract=off so every
 * expression rounds as written.  Citations are reference file:line.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

static thread_local std::string g_err;

/* ---------------------------------------------------------------- vectors */
struct V3 {
  double x = 0, y = 0, z = 0;
  V3() {}
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  double &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
static inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 operator*(double s, V3 a) { return V3(s * a.x, s * a.y, s * a.z); }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b) {
  return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double norm2(V3 a) { return dot(a, a); }
static inline double norm(V3 a) { return std::sqrt(norm2(a)); }

/* Axis-aligned box with the closed-interval semantics of Eigen::AlignedBox
 * (a default box is empty: min=+max, max=lowest). */
struct Box {
  V3 mn{std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
        std::numeric_limits<double>::max()};
  V3 mx{std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest(),
        std::numeric_limits<double>::lowest()};
  Box() {}
  Box(V3 a, V3 b) : mn(a), mx(b) {}
  bool intersects(const Box &o) const {
    return mn.x <= o.mx.x && mn.y <= o.mx.y && mn.z <= o.mx.z && o.mn.x <= mx.x &&
           o.mn.y <= mx.y && o.mn.z <= mx.z;
  }
  void extend(const Box &o) {
    mn = V3(std::min(mn.x, o.mn.x), std::min(mn.y, o.mn.y), std::min(mn.z, o.mn.z));
    mx = V3(std::max(mx.x, o.mx.x), std::max(mx.y, o.mx.y), std::max(mx.z, o.mx.z));
  }
  Box merged(const Box &o) const {
    Box b = *this;
    b.extend(o);
    return b;
  }
  V3 center() const { return 0.5 * (mn + mx); }
  V3 sizes() const { return mx - mn; }
};

/* ---------------------------------------------------------------- curves */
struct Curve {
  std::vector<V3> q;
  int type = ORC_BSPLINE;
  bool looped = false;
  int64_t nseg() const { /* curve.h:61-66 */
    if (looped) return (int64_t)q.size();
    return type == ORC_POLYLINE ? (int64_t)q.size() - 1 : (int64_t)q.size() - 3;
  }
};

/* Cubic bases, curve.cpp:20-43 (B-spline, uniform Catmull–Rom, Bézier). */
static void basis(int type, double t, double *B) {
  if (type == ORC_BSPLINE) {
    B[0] = ((-(1.0 / 6) * t + (1.0 / 2)) * t - (1.0 / 2)) * t + (1.0 / 6);
    B[1] = ((1.0 / 2) * t - 1.0) * (t * t) + (2.0 / 3);
    B[2] = ((-(1.0 / 2) * t + (1.0 / 2)) * t + (1.0 / 2)) * t + (1.0 / 6);
    B[3] = (1.0 / 6) * (t * t) * t;
  } else if (type == ORC_CATMULLROM) {
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
}

/* Point on segment seg at parameter t, curve.cpp:156-184. */
static V3 point(const Curve &c, int64_t seg, double t) {
  if (t < 0.0 || t > 1.0 || seg < 0 || seg >= c.nseg())
    throw std::runtime_error("Curve::get_point() has an invalid i or t.");
  const size_t n = c.q.size();
  if (c.type == ORC_POLYLINE) {
    const V3 &a = c.q[seg], &b = c.q[(seg + 1) % n];
    return (1.0 - t) * a + t * b;
  }
  double B[4];
  basis(c.type, t, B);
  V3 r;
  for (int d = 0; d < 3; ++d)
    r[d] = B[0] * c.q[seg][d] + B[1] * c.q[(seg + 1) % n][d] +
           B[2] * c.q[(seg + 2) % n][d] + B[3] * c.q[(seg + 3) % n][d];
  return r;
}

/* Exact 1D extent of a cubic sub-segment from its monomial coefficients and
 * the interior extrema (roots of the derivative), curve.cpp:197-277.  A
 * negative discriminant yields NaN roots, which every comparison rejects. */
static void extent_1d(int type, double a0, double a1, double a2, double a3, double t0,
                      double t1, double &lo, double &hi) {
  double c3, c2, c1, c0;
  if (type == ORC_BSPLINE) {
    c3 = -(1.0 / 6.0) * a0 + 0.5 * a1 - 0.5 * a2 + (1.0 / 6.0) * a3;
    c2 = 0.5 * a0 - a1 + 0.5 * a2;
    c1 = -0.5 * a0 + 0.5 * a2;
    c0 = (1.0 / 6.0) * a0 + (2.0 / 3.0) * a1 + (1.0 / 6.0) * a2;
  } else if (type == ORC_CATMULLROM) {
    c3 = -0.5 * a0 + 1.5 * a1 - 1.5 * a2 + 0.5 * a3;
    c2 = 1.0 * a0 - 2.5 * a1 + 2.0 * a2 - 0.5 * a3;
    c1 = -0.5 * a0 + 0.5 * a2;
    c0 = 1.0 * a1;
  } else {
    c3 = -1.0 * a0 + 3.0 * a1 - 3.0 * a2 + 1.0 * a3;
    c2 = 3.0 * a0 - 6.0 * a1 + 3.0 * a2;
    c1 = -3.0 * a0 + 3.0 * a1;
    c0 = 1.0 * a0;
  }
  auto ev = [&](double t) { return ((c3 * t + c2) * t + c1) * t + c0; };
  lo = ev(t0);
  hi = ev(t1);
  if (hi < lo) std::swap(lo, hi);
  auto consider = [&](double t) {
    if (t >= t0 && t <= t1) {
      double v = ev(t);
      lo = std::min(lo, v);
      hi = std::max(hi, v);
    }
  };
  const double disc = 4 * c2 * c2 - 12 * c1 * c3;
  if (disc == 0) {
    consider(-2 * c2 / (6 * c3));
  } else {
    /* Numerically stable pair of roots of 3 c3 t^2 + 2 c2 t + c1. */
    double rad = (c2 >= 0) ? (-2 * c2 - std::sqrt(disc)) : (-2 * c2 + std::sqrt(disc));
    consider(rad / (6 * c3));
    consider(2 * c1 / rad);
  }
}

/* curve.cpp:293-324 */
static Box segment_box(const Curve &c, int64_t seg, double tlo = 0.0, double thi = 1.0) {
  if (seg < 0 || seg >= c.nseg() || tlo < 0 || thi > 1 || tlo > thi)
    throw std::runtime_error("Curve::get_segment_bounding_box() has an invalid i or t.");
  Box b;
  if (c.type == ORC_POLYLINE) {
    V3 a = point(c, seg, tlo), e = point(c, seg, thi);
    b.mn = V3(std::min(a.x, e.x), std::min(a.y, e.y), std::min(a.z, e.z));
    b.mx = V3(std::max(a.x, e.x), std::max(a.y, e.y), std::max(a.z, e.z));
  } else {
    const size_t n = c.q.size();
    for (int d = 0; d < 3; ++d)
      extent_1d(c.type, c.q[seg][d], c.q[(seg + 1) % n][d], c.q[(seg + 2) % n][d],
                c.q[(seg + 3) % n][d], tlo, thi, b.mn[d], b.mx[d]);
  }
  return b;
}

/* curve.cpp:331-343: union of all segment boxes, empty box for no segments. */
static Box curve_box(const Curve &c) {
  const int64_t ns = c.nseg();
  if (ns <= 0) return Box();
  Box b = segment_box(c, 0);
  for (int64_t i = 1; i < ns; ++i) b.extend(segment_box(c, i));
  return b;
}

/* Start points of a closed curve, curve.cpp:125-134. */
static std::vector<V3> start_points(const Curve &c) {
  if (!c.looped)
    throw std::runtime_error("Curve::GetStartPoints() does not support unlooped curves.");
  std::vector<V3> s(c.q.size());
  for (size_t i = 0; i < c.q.size(); ++i) s[i] = point(c, (int64_t)i, 0.0);
  return s;
}
static std::vector<V3> midpoints(const Curve &c) { /* curve.cpp:99-108 */
  if (!c.looped)
    throw std::runtime_error("Curve::GetMidpoints() does not support unlooped curves.");
  std::vector<V3> s(c.q.size());
  for (size_t i = 0; i < c.q.size(); ++i) s[i] = point(c, (int64_t)i, 0.5);
  return s;
}
static std::vector<V3> tangents(const Curve &c) { /* curve.cpp:110-123 */
  std::vector<V3> s = start_points(c);
  const size_t n = s.size();
  std::vector<V3> t(n);
  for (size_t i = 0; i + 1 < n; ++i) t[i] = s[i + 1] - s[i];
  if (n) t[n - 1] = s[0] - s[n - 1];
  return t;
}

/* ---------------------------------------------------------------- kd-BVH */
/* Median-split kd hierarchy with the layout of ModifiedKdBVH.h:101-237:
 * internal boxes are appended in post-order (root last), children[2k..2k+1]
 * hold a box index (< nobj-1) or nobj-1+object index, and objects are
 * reordered so each subtree covers a contiguous object range. */
struct KdTree {
  std::vector<Box> boxes;        // internal nodes
  std::vector<int> children;     // 2 per internal node
  std::vector<int> obj;          // kd order -> caller's object index
  std::vector<Box> objbox;       // kd order boxes
  int nobj = 0;
  int root() const { return (int)boxes.size() - 1; }
  int nbox() const { return (int)boxes.size(); }

  void build(const std::vector<Box> &in) {
    boxes.clear();
    children.clear();
    nobj = (int)in.size();
    obj.resize(nobj);
    for (int i = 0; i < nobj; ++i) obj[i] = i;
    objbox = in;
    if (nobj < 2) return;
    std::vector<std::pair<V3, int>> centers(nobj);
    for (int i = 0; i < nobj; ++i) centers[i] = {in[i].center(), i};
    boxes.reserve(nobj - 1);
    children.reserve(2 * nobj - 2);
    split(centers, 0, nobj, in, 0);
    for (int i = 0; i < nobj; ++i) {
      obj[i] = centers[i].second;
      objbox[i] = in[obj[i]];
    }
  }

 private:
  void split(std::vector<std::pair<V3, int>> &c, int lo, int hi, const std::vector<Box> &in,
             int axis) {
    auto less = [axis](const std::pair<V3, int> &a, const std::pair<V3, int> &b) {
      return a.first[axis] < b.first[axis];
    };
    const int leafbase = nobj - 1;
    if (hi - lo == 2) {
      boxes.push_back(in[c[lo].second].merged(in[c[lo + 1].second]));
      children.push_back(lo + leafbase);
      children.push_back(lo + 1 + leafbase);
    } else if (hi - lo == 3) {
      const int mid = lo + 2;
      std::nth_element(c.begin() + lo, c.begin() + mid, c.begin() + hi, less);
      split(c, lo, mid, in, (axis + 1) % 3);
      const int left = (int)boxes.size() - 1;
      boxes.push_back(boxes[left].merged(in[c[mid].second]));
      children.push_back(left);
      children.push_back(mid + leafbase);
    } else {
      const int mid = lo + (hi - lo) / 2;
      std::nth_element(c.begin() + lo, c.begin() + mid, c.begin() + hi, less);
      split(c, lo, mid, in, (axis + 1) % 3);
      const int left = (int)boxes.size() - 1;
      split(c, mid, hi, in, (axis + 1) % 3);
      const int right = (int)boxes.size() - 1;
      boxes.push_back(boxes[left].merged(boxes[right]));
      children.push_back(left);
      children.push_back(right);
    }
  }
};

/* All (a, b) object pairs of two kd trees whose boxes intersect (closed
 * intervals), reported in kd order indices; the pair set is what
 * BVAlgorithms.h:94-145 enumerates.  A tree with <2 objects is a flat list. */
template <typename F>
static void tree_pairs(const KdTree &A, const KdTree &B, F &&report) {
  if (A.nobj == 0 || B.nobj == 0) return;
  // node id: >=0 internal box index; <0 encodes leaf object -(k+1).
  auto node_box = [](const KdTree &T, int id) -> const Box & {
    return id >= 0 ? T.boxes[id] : T.objbox[-id - 1];
  };
  auto kids = [](const KdTree &T, int id, int out[2]) {
    for (int s = 0; s < 2; ++s) {
      int ch = T.children[2 * id + s];
      out[s] = ch < T.nobj - 1 ? ch : -((ch - (T.nobj - 1)) + 1);
    }
  };
  const int ra = A.nobj < 2 ? -1 : A.root();
  const int rb = B.nobj < 2 ? -1 : B.root();
  std::vector<std::pair<int, int>> stack{{ra, rb}};
  while (!stack.empty()) {
    const int a = stack.back().first, b = stack.back().second;
    stack.pop_back();
    if (!node_box(A, a).intersects(node_box(B, b))) continue;
    int ka[2], kb[2];
    if (a >= 0 && b >= 0) {
      kids(A, a, ka);
      kids(B, b, kb);
      for (int s = 0; s < 2; ++s)
        for (int t = 0; t < 2; ++t) stack.push_back({ka[s], kb[t]});
    } else if (a >= 0) {
      kids(A, a, ka);
      stack.push_back({ka[0], b});
      stack.push_back({ka[1], b});
    } else if (b >= 0) {
      kids(B, b, kb);
      stack.push_back({a, kb[0]});
      stack.push_back({a, kb[1]});
    } else {
      report(-a - 1, -b - 1);
    }
  }
}

/* ---------------------------------------------------------------- PLS */
/* potential_link_search.cpp:88-127: sorted unique (i<j) pairs of curves whose
 * boxes intersect.  Brute force below 50 curves, tree self-join otherwise; the
 * set is the same either way. */
static std::vector<std::pair<int, int>> pls_unique(const std::vector<Box> &bx, bool use_tree) {
  const int n = (int)bx.size();
  std::vector<std::pair<int, int>> out;
  if (use_tree) {
    KdTree t;
    t.build(bx);
    tree_pairs(t, t, [&](int a, int b) {
      int i = t.obj[a], j = t.obj[b];
      if (i < j) out.emplace_back(i, j);
    });
    std::sort(out.begin(), out.end());
  } else {
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j)
        if (bx[i].intersects(bx[j])) out.emplace_back(i, j);
  }
  return out;
}

/* ---------------------------------------------------------------- Discretize */
/* discretize.cpp:102-281.  A piece is (curve, segment, t_lo, t_hi).  Each round
 * bisects every piece of curve i whose box meets a piece of a potentially
 * linked curve; untouched pieces are retired.  Stops when a round finds no
 * overlap; throws after 55 rounds or when a piece gets smaller than
 * eps * mean |q| / sqrt(3). */
typedef std::tuple<int, int64_t, double, double> Piece;

static std::vector<Curve> discretize(const std::vector<Curve> &curves,
                                     const std::vector<std::vector<int>> &links) {
  const int nc = (int)curves.size();
  const double eps = std::numeric_limits<double>::epsilon();
  std::vector<std::vector<std::pair<Piece, Box>>> live(nc);
  int64_t total = 0;
  double coord_sum = 0.0;
  double min_diag = std::numeric_limits<double>::infinity();
  int min_c = -1;
  int64_t min_s = -1;
  for (int i = 0; i < nc; ++i) {
    const int64_t ns = curves[i].nseg();
    for (int64_t s = 0; s < ns; ++s) {
      Box b = segment_box(curves[i], s);
      live[i].push_back({Piece(i, s, 0.0, 1.0), b});
      coord_sum += norm(curves[i].q[s]);
      double d = norm(b.sizes());
      if (!(d >= min_diag)) {
        min_diag = d;
        min_c = i;
        min_s = s;
      }
    }
    total += std::max<int64_t>(ns, 0);
  }
  coord_sum /= total * 1.73205080757;
  const double threshold = eps * coord_sum;
  if (!(min_diag >= threshold)) {
    std::fprintf(stderr,
                 "Discretize: Smallest input segment, (%d,%lld) has a size of %g, which is "
                 "smaller than eps * avg coordinate.\n",
                 min_c, (long long)min_s, min_diag);
    throw std::runtime_error(
        "Discretize: Smallest input segment is smaller than eps * avg coordinate. "
        "Consider combining.");
  }
  std::vector<std::vector<Piece>> retired(nc);
  int64_t npairs = 0;
  for (int round = 0; round < 55; ++round) {
    std::vector<KdTree> trees(nc);
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < nc; ++i) {
      std::vector<Box> b(live[i].size());
      for (size_t k = 0; k < b.size(); ++k) b[k] = live[i][k].second;
      trees[i].build(b);
    }
    npairs = 0;
    std::vector<std::vector<std::pair<Piece, Box>>> next(nc);
    bool too_small = false;
#pragma omp parallel for schedule(dynamic) reduction(+ : npairs)
    for (int i = 0; i < nc; ++i) {
      std::vector<char> hit(live[i].size(), 0);
      int64_t local = 0;
      for (int j : links[i]) {
        tree_pairs(trees[i], trees[j], [&](int a, int) {
          hit[trees[i].obj[a]] = 1;
          ++local;
        });
      }
      double md = std::numeric_limits<double>::infinity();
      for (size_t k = 0; k < live[i].size(); ++k) {
        const Piece &pc = live[i][k].first;
        if (hit[k]) {
          const int c = std::get<0>(pc);
          const int64_t s = std::get<1>(pc);
          const double lo = std::get<2>(pc), hi = std::get<3>(pc);
          const double mid = 0.5 * (lo + hi);
          Box b1 = segment_box(curves[c], s, lo, mid);
          Box b2 = segment_box(curves[c], s, mid, hi);
          next[i].push_back({Piece(c, s, lo, mid), b1});
          next[i].push_back({Piece(c, s, mid, hi), b2});
          md = std::min(md, std::min(norm(b1.sizes()), norm(b2.sizes())));
        } else {
          retired[i].push_back(pc);
        }
      }
      npairs += local;
      if (!(md >= threshold)) {
#pragma omp critical
        too_small = true;
      }
    }
    if (too_small)
      throw std::runtime_error(
          "Discretize: Smallest segment is smaller than eps * avg coordinate. Perhaps "
          "curves are too close.");
    if (npairs % 2 != 0) std::printf("Error: npairs is odd: %lld.\n", (long long)npairs);
    std::printf("Iteration %d, %lld overlaps.\n", round + 1, (long long)(npairs / 2));
    if (npairs == 0) {
      std::printf("Finished after %d iteration(s).\n", round + 1);
      break;
    }
    live.swap(next);
  }
  if (npairs > 0)
    throw std::runtime_error("Reached max recursion (curves are too close) in Discretize().");
  std::vector<Curve> out(nc);
  int64_t count = 0;
  for (int i = 0; i < nc; ++i) {
    Curve &o = out[i];
    o.type = ORC_POLYLINE;
    o.looped = true;
    std::sort(retired[i].begin(), retired[i].end());
    for (const Piece &pc : retired[i]) o.q.push_back(point(curves[i], std::get<1>(pc), std::get<2>(pc)));
    if (!curves[i].looped && !retired[i].empty()) {
      const Piece &pc = retired[i].back();
      o.q.push_back(point(curves[i], std::get<1>(pc), std::get<3>(pc)));
      o.looped = false;
    }
    count += (int64_t)retired[i].size();
  }
  std::printf("Original segment count: %lld, discretized count: %lld\n", (long long)total,
              (long long)count);
  return out;
}

/* ---------------------------------------------------------------- direct sum */
static const double kInv2Pi = 0.15915494309189533576888376337251436;
static const double kInv4Pi = 0.07957747154594766788444188168625718;
static const double k4Pi = 12.566370614359172953850573533118011536788677597500;

/* The per-segment-pair quantities of compute_link.cpp:63-84 / :151-175 for the
 * quadrilateral a=q0-p0, b=q0-p1, c=q1-p1, d=q1-p0 (Arai 2013): two triangle
 * denominators, the triple product, and the one-arctan combination. */
static inline double pair_term_x4pi(V3 a, V3 b, V3 c, V3 d) {
  const double la = norm(a), lb = norm(b), lc = norm(c), ld = norm(d);
  const double ac = dot(c, a);
  const double div1 = la * lb * lc + dot(a, b) * lc + ac * lb + dot(c, b) * la;
  const double div2 = la * ld * lc + dot(a, d) * lc + ac * ld + dot(c, d) * la;
  const double det = dot(a, cross(b, c));
  const double xp = div1 * div2 - det * det;
  const double yp = det * (div1 + div2);
  double v = 2.0 * std::atan2(yp, xp);
  const double sgn = det > 0 ? 1.0 : (det < 0 ? -1.0 : 0.0);
  if (yp * det < 0) v += k4Pi * sgn;
  return v;  // 4*pi times the linking-number contribution
}

/* compute_link.cpp:126-178, optionally restricted to a row range. */
static double ds_rows(const std::vector<V3> &p, const std::vector<V3> &q, int64_t r0,
                      int64_t r1, bool parallel) {
  const int64_t n1 = (int64_t)p.size(), n2 = (int64_t)q.size();
  double acc = 0.0;
#pragma omp parallel for reduction(+ : acc) if (parallel) schedule(static)
  for (int64_t i = r0; i < r1; ++i) {
    const V3 pi = p[i], pn = p[(i + 1) % n1];
    double row = 0.0;
    for (int64_t j = 0; j < n2; ++j) {
      const V3 qj = q[j], qn = q[(j + 1) % n2];
      row += pair_term_x4pi(qj - pi, qj - pn, qn - pn, qn - pi);
    }
    acc += row;
  }
  return acc * kInv4Pi;
}

/* compute_link.cpp:200-280 (Bertolazzi running product per column j, one
 * atan2 per column; branch crossings counted into integer offsets). */
static double ds_fast(const std::vector<V3> &p, const std::vector<V3> &q) {
  const int64_t n1 = (int64_t)p.size(), n2 = (int64_t)q.size();
  std::vector<double> X(n2, 1.0), Y(n2, 0.0), S(n2, -1.0), off(n2, 0.0);
  for (int64_t i = 0; i < n1; ++i) {
    const V3 pi = p[i], pn = p[(i + 1) % n1];
    for (int64_t j = 0; j < n2; ++j) {
      const V3 qj = q[j], qn = q[(j + 1) % n2];
      const V3 a = qj - pi, b = qj - pn, c = qn - pn, d = qn - pi;
      const double la = norm(a), lb = norm(b), lc = norm(c), ld = norm(d);
      const double ac = dot(c, a);
      const double div1 = la * lb * lc + dot(a, b) * lc + ac * lb + dot(c, b) * la;
      const double div2 = la * ld * lc + dot(a, d) * lc + ac * ld + dot(c, d) * la;
      const double det = dot(a, cross(b, c));
      const double xp = div1 * div2 - det * det;
      const double yp = det * (div1 + div2);
      double ny = X[j] * yp + xp * Y[j];
      double nx = X[j] * xp - Y[j] * yp;
      const double sc = std::max(std::fabs(nx), std::fabs(ny));
      nx /= sc;
      ny /= sc;
      const double sgn = det > 0 ? 1.0 : (det < 0 ? -1.0 : 0.0);
      if (yp * det < 0) off[j] += sgn;
      const double ns = (ny > 0.0 || (ny == 0.0 && nx < 0.0)) ? 1.0 : -1.0;
      if (ns * S[j] < 0 && yp * S[j] >= 0) off[j] += S[j];
      X[j] = nx;
      Y[j] = ny;
      S[j] = ns;
    }
  }
  double lk = 0.0;
  for (int64_t j = 0; j < n2; ++j) lk += (2.0 * kInv4Pi) * std::atan2(Y[j], X[j]);
  for (int64_t j = 0; j < n2; ++j) lk += off[j];
  return lk;
}

/* ---------------------------------------------------------------- Barnes–Hut */
/* compute_link.h:95-114. */
struct Moment {
  double Q[3][3][3] = {};  // Q[k][i][j] = ∫ r'_k (r-r̃)_i (r-r̃)_j
  double D[3][3] = {};     // D[k][j]    = ∫ r'_k (r-r̃)_j
  V3 c, mid;
  double cn = 0, dn = 0, qn = 0;
};

struct SegTree {
  KdTree kd;
  std::vector<Moment> mom;  // internal nodes then leaves (kd order)
  int nbox() const { return kd.nbox(); }
  const Box &box(int node) const {
    return node < kd.nbox() ? kd.boxes[node] : kd.objbox[node - kd.nbox()];
  }
  std::pair<int, int> kids(int node) const {
    return {kd.children[2 * node], kd.children[2 * node + 1]};
  }
};

/* compute_link.cpp:333-389: leaf moments from (midpoint, tangent), then each
 * child shifted to the parent's box centre (Eqs. 21-23). */
static void build_moments(SegTree &T, const std::vector<V3> &mids, const std::vector<V3> &tans,
                          int node) {
  const Box &bb = T.kd.boxes[node];
  const V3 ctr = 0.5 * (bb.mx + bb.mn);
  Moment &M = T.mom[node];
  M.mid = ctr;
  auto ch = T.kids(node);
  const int off = T.nbox();
  for (int s = 0; s < 2; ++s) {
    const int idx = s == 0 ? ch.first : ch.second;
    Moment &C = T.mom[idx];
    if (idx < off) {
      build_moments(T, mids, tans, idx);
    } else {
      const int seg = T.kd.obj[idx - off];
      C.mid = mids[seg];
      const V3 t = tans[seg];
      C.c = t;
      C.cn = norm(t);
      for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) C.Q[k][i][j] = (1.0 / 12.0) * t[k] * (t[i] * t[j]);
      C.qn = (1.0 / 36.0) * C.cn * C.cn * C.cn;
    }
    const V3 cc = C.c;
    M.c = M.c + cc;
    const V3 r = C.mid - ctr;
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 3; ++j) M.D[k][j] += C.D[k][j] + cc[k] * r[j];
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          M.Q[k][i][j] += C.Q[k][i][j] + (r[j] * C.D[k][i]) + (r[i] * C.D[k][j]) +
                          cc[k] * (r[i] * r[j]);
  }
  M.cn = norm(M.c);
  double dsq = 0, qsq = 0;
  for (int k = 0; k < 3; ++k)
    for (int j = 0; j < 3; ++j) dsq += M.D[k][j] * M.D[k][j];
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) qsq += M.Q[k][i][j] * M.Q[k][i][j];
  M.dn = std::sqrt(dsq) * 0.57735026918962576450914878050195745564760175127;
  M.qn = (1.0 / 3.0) * std::sqrt(qsq);
}

/* compute_link.cpp:294-315 + :115-127: tree over segment boxes, then moments. */
static SegTree make_tree(const std::vector<V3> &p) {
  Curve c;
  c.q = p;
  c.type = ORC_POLYLINE;
  c.looped = true;
  const int64_t n = (int64_t)p.size();
  std::vector<Box> bx(n);
  for (int64_t s = 0; s < n; ++s) bx[s] = segment_box(c, s);
  SegTree T;
  T.kd.build(bx);
  T.mom.assign(T.kd.nbox() + n, Moment());
  if (n < 2) throw std::runtime_error("Barnes-Hut needs at least 2 segments per curve.");
  build_moments(T, midpoints(c), tangents(c), T.kd.root());
  return T;
}

/* Dipole/quadrupole far-field terms and the Eq. 14 error estimate,
 * compute_link.cpp:410-524 (Supplemental Eqs. 29-30).  Returns the bracket that
 * is later multiplied by 1/(4 pi |r|^3); e receives the unscaled estimate. */
static double far_terms(const Moment &A, const Moment &B, V3 r, double ir2, double ra,
                        double rb, double &e) {
  // curl-like vectors of the dipole matrices: w_i = eps_ijk D_jk
  auto wvec = [](const double D[3][3]) {
    return V3(D[1][2] - D[2][1], D[2][0] - D[0][2], D[0][1] - D[1][0]);
  };
  auto mulv = [](const double D[3][3], V3 v) {
    return V3(D[0][0] * v.x + D[0][1] * v.y + D[0][2] * v.z,
              D[1][0] * v.x + D[1][1] * v.y + D[1][2] * v.z,
              D[2][0] * v.x + D[2][1] * v.y + D[2][2] * v.z);
  };
  auto mulTv = [](const double D[3][3], V3 v) {
    return V3(D[0][0] * v.x + D[1][0] * v.y + D[2][0] * v.z,
              D[0][1] * v.x + D[1][1] * v.y + D[2][1] * v.z,
              D[0][2] * v.x + D[1][2] * v.y + D[2][2] * v.z);
  };
  const V3 ca = A.c, cb = B.c;
  const V3 wa = wvec(A.D), wb = wvec(B.D);
  double acc = 0;
  // dipole (Eq. 29)
  acc -= dot(ca, wb) + dot(wa, cb);
  acc -= 3.0 * ir2 * (dot(cross(mulv(B.D, r), ca), r) + dot(cross(mulv(A.D, r), cb), r));
  // quadrupole (Eq. 30)
  const V3 ta(A.Q[0][0][0] + A.Q[0][1][1] + A.Q[0][2][2],
              A.Q[1][0][0] + A.Q[1][1][1] + A.Q[1][2][2],
              A.Q[2][0][0] + A.Q[2][1][1] + A.Q[2][2][2]);
  const V3 tb(B.Q[0][0][0] + B.Q[0][1][1] + B.Q[0][2][2],
              B.Q[1][0][0] + B.Q[1][1][1] + B.Q[1][2][2],
              B.Q[2][0][0] + B.Q[2][1][1] + B.Q[2][2][2]);
  const double s2a = -dot(r, cross(ta, cb) + cross(ca, tb));
  auto rowdot = [](const double D1[3][3], int a, const double D2[3][3], int b) {
    return D1[a][0] * D2[b][0] + D1[a][1] * D2[b][1] + D1[a][2] * D2[b][2];
  };
  const double s2b = 2.0 * (r.x * (rowdot(A.D, 1, B.D, 2) - rowdot(A.D, 2, B.D, 1)) +
                            r.y * (rowdot(A.D, 2, B.D, 0) - rowdot(A.D, 0, B.D, 2)) +
                            r.z * (rowdot(A.D, 0, B.D, 1) - rowdot(A.D, 1, B.D, 0)));
  const double s2 = s2a + s2b;
  auto qrow = [](const Moment &M, int k, int i, V3 v) {
    return M.Q[k][i][0] * v.x + M.Q[k][i][1] * v.y + M.Q[k][i][2] * v.z;
  };
  const V3 ua(qrow(A, 1, 2, r) - qrow(A, 2, 1, r), qrow(A, 2, 0, r) - qrow(A, 0, 2, r),
              qrow(A, 0, 1, r) - qrow(A, 1, 0, r));
  const V3 ub(qrow(B, 1, 2, r) - qrow(B, 2, 1, r), qrow(B, 2, 0, r) - qrow(B, 0, 2, r),
              qrow(B, 0, 1, r) - qrow(B, 1, 0, r));
  double s3 = dot(cb, ua);
  s3 += -dot(ca, ub);
  s3 += dot(r, mulTv(A.D, wb) - mulTv(B.D, wa));
  auto qquad = [&](const Moment &M, int k) {
    V3 v(qrow(M, k, 0, r), qrow(M, k, 1, r), qrow(M, k, 2, r));
    return dot(r, v);
  };
  const V3 rra(qquad(A, 0), qquad(A, 1), qquad(A, 2));
  const V3 rrb(qquad(B, 0), qquad(B, 1), qquad(B, 2));
  const double s1 = -dot(r, cross(cb, rra)) + dot(r, cross(ca, rrb)) +
                    2.0 * dot(mulv(A.D, r), cross(r, mulv(B.D, r)));
  acc -= 1.5 * ir2 * ((s2 + 2.0 * s3) + 5.0 * ir2 * s1);
  e = ir2 * (A.qn * (ra * B.cn + 3 * B.dn) + B.qn * (rb * A.cn + 3 * A.dn));
  return acc;
}

/* Leaf term, compute_link.cpp:576-606 (det == 0 handled by copysign). */
static double leaf_term(const Moment &A, const Moment &B) {
  const V3 r = B.mid - A.mid;
  const double det = dot(r, cross(B.c, A.c));
  const V3 a = r + 0.5 * (A.c - B.c);
  const V3 b = r + (-0.5) * (A.c + B.c);
  const V3 c = r + 0.5 * (B.c - A.c);
  const V3 d = r + 0.5 * (B.c + A.c);
  const double la = norm(a), lb = norm(b), lc = norm(c), ld = norm(d);
  const double ac = dot(a, c);
  const double div1 = la * lb * lc + dot(a, b) * lc + ac * lb + dot(b, c) * la;
  const double div2 = la * ld * lc + dot(a, d) * lc + ac * ld + dot(d, c) * la;
  const double xp = div1 * div2 - det * det;
  const double s = det == 0 ? std::copysign(1.0, div1) + std::copysign(1.0, div2) : div1 + div2;
  const double yp = det * s;
  const double sg = det == 0 ? -std::copysign(1.0, div1) : std::copysign(1.0, det);
  return kInv2Pi * std::atan2(yp, xp) + ((s < 0 || (s == 0 && det < 0)) ? sg : 0.0);
}

struct BHCtx {
  double epssq, itolsq;
  int64_t visits = 0;
};

/* Node-pair classification shared by the DFS and BFS passes. Returns 0 leaf,
 * 1 far, 2 split (split1 tells which). */
static int classify(const SegTree &T1, const SegTree &T2, int n1, int n2, const BHCtx &X,
                    double &r1, double &r2, V3 &r, double &dsq, bool &split1) {
  const bool leaf1 = n1 >= T1.nbox(), leaf2 = n2 >= T2.nbox();
  const double r1sq = 0.25 * norm2(T1.box(n1).sizes());
  const double r2sq = 0.25 * norm2(T2.box(n2).sizes());
  r1 = std::sqrt(r1sq);
  r2 = std::sqrt(r2sq);
  const double rsum2 = (r1 + r2) * (r1 + r2);
  r = T2.mom[n2].mid - T1.mom[n1].mid;
  dsq = norm2(r) + X.epssq;
  if (leaf1 && leaf2) return 0;
  if (dsq > rsum2 * X.itolsq + X.epssq) return 1;
  if (!leaf1 && !leaf2) split1 = !(r2sq > r1sq);
  else split1 = leaf2;  // only one of them can be split
  return 2;
}

static double far_value(const SegTree &T1, const SegTree &T2, int n1, int n2, V3 r, double dsq,
                        double r1, double r2, double &err) {
  const Moment &A = T1.mom[n1], &B = T2.mom[n2];
  double v = dot(r, cross(B.c, A.c));
  const double ir2 = 1.0 / dsq;
  const double mult = ir2 / std::sqrt(dsq) * kInv4Pi;
  double e;
  v += far_terms(A, B, r, ir2, r1, r2, e);
  err += e * mult;
  return mult * v;
}

/* compute_link.cpp:547-666, including the argument swap on recursion. */
static double bh_dfs(const SegTree &T1, const SegTree &T2, int n1, int n2, BHCtx &X,
                     double &err) {
  X.visits++;
  double r1, r2, dsq;
  V3 r;
  bool split1 = true;
  const int cls = classify(T1, T2, n1, n2, X, r1, r2, r, dsq, split1);
  if (cls == 0) return leaf_term(T1.mom[n1], T2.mom[n2]);
  if (cls == 1) return far_value(T1, T2, n1, n2, r, dsq, r1, r2, err);
  if (split1) {
    auto k = T1.kids(n1);
    return bh_dfs(T2, T1, n2, k.first, X, err) + bh_dfs(T2, T1, n2, k.second, X, err);
  }
  auto k = T2.kids(n2);
  return bh_dfs(T1, T2, n1, k.first, X, err) + bh_dfs(T1, T2, n1, k.second, X, err);
}

/* compute_link.cpp:695-841: one breadth-first pass over a frontier. */
static double bh_bfs_pass(const SegTree &T1, const SegTree &T2, const std::vector<int> &a,
                          const std::vector<int> &b, std::vector<int> &na, std::vector<int> &nb,
                          BHCtx &X, double &err) {
  const int64_t m = (int64_t)a.size();
  na.assign(2 * m, 0);
  nb.assign(2 * m, 0);
  int64_t next = 0;
  double acc = 0, eacc = 0;
  int64_t vis = 0;
#pragma omp parallel for reduction(+ : acc, eacc, vis) if (m >= 100)
  for (int64_t t = 0; t < m; ++t) {
    vis++;
    double r1, r2, dsq;
    V3 r;
    bool split1 = true;
    const int cls = classify(T1, T2, a[t], b[t], X, r1, r2, r, dsq, split1);
    if (cls == 0) {
      acc += leaf_term(T1.mom[a[t]], T2.mom[b[t]]);
    } else if (cls == 1) {
      acc += far_value(T1, T2, a[t], b[t], r, dsq, r1, r2, eacc);
    } else {
      int64_t slot;
#pragma omp atomic capture
      slot = next++;
      if (split1) {
        auto k = T1.kids(a[t]);
        na[2 * slot] = k.first;
        na[2 * slot + 1] = k.second;
        nb[2 * slot] = nb[2 * slot + 1] = b[t];
      } else {
        auto k = T2.kids(b[t]);
        na[2 * slot] = na[2 * slot + 1] = a[t];
        nb[2 * slot] = k.first;
        nb[2 * slot + 1] = k.second;
      }
    }
  }
  na.resize(2 * next);
  nb.resize(2 * next);
  X.visits += vis;
  err = eacc;
  return acc;
}

/* compute_link.cpp:866-934. */
static double bh_eval(const SegTree &T1, const SegTree &T2, double epssq, double itolsq,
                      double &err_total, bool parallel, int64_t *visits) {
  BHCtx X;
  X.epssq = epssq;
  X.itolsq = itolsq;
  err_total = 0;
  double out;
  if (parallel) {
    std::vector<int> a, b, na{T1.kd.root()}, nb{T2.kd.root()};
    double acc = 0;
    int64_t size = 1;
    while (size > 0 && size < 75000) {
      std::swap(a, na);
      std::swap(b, nb);
      double e = 0;
      acc += bh_bfs_pass(T1, T2, a, b, na, nb, X, e);
      size = (int64_t)na.size();
      err_total += e;
    }
    double e2 = 0;
    int64_t vis = 0;
#pragma omp parallel for reduction(+ : acc, e2, vis) schedule(dynamic)
    for (int64_t t = 0; t < size; ++t) {
      BHCtx Y = X;
      Y.visits = 0;
      acc += bh_dfs(T1, T2, na[t], nb[t], Y, e2);
      vis += Y.visits;
    }
    X.visits += vis;
    err_total += e2;
    out = acc;
  } else {
    double e = 0;
    float f = (float)bh_dfs(T1, T2, T1.kd.root(), T2.kd.root(), X, e);
    err_total += e;
    out = f;
  }
  if (visits) *visits = X.visits;
  return out;
}

/* ---------------------------------------------------------------- certificate */
struct Triplet {
  int i, j, v;
};

static const double kEpsSq = 1e-24;       // linking_number_certificate.cpp:24
static const double kTargetErr = 0.1;     // :25
static const double kErrScale = 0.05;     // :26
static const int kParallelCurves = 6;     // :22

static int to_int(double x) {
  // The reference converts round(LK) to int implicitly; NaN gives INT_MIN on x86.
  if (std::isnan(x)) return std::numeric_limits<int>::min();
  return (int)x;
}

static std::vector<Triplet> certificate(const std::vector<Curve> &in, bool force_ds,
                                        bool disc, double beta_init, double beta_limit,
                                        double &mean_err, double &max_err,
                                        std::vector<double> *lks,
                                        std::vector<std::pair<int, int>> *cands) {
  std::printf("Starting Potential Link Search.\n");
  const int nc = (int)in.size();
  std::vector<Box> bx(nc);
  for (int i = 0; i < nc; ++i) bx[i] = curve_box(in[i]);
  std::vector<Curve> curves;
  std::vector<std::pair<int, int>> uniq;
  if (disc) {
    auto u0 = pls_unique(bx, nc >= 50);
    std::vector<std::vector<int>> both(nc);
    for (auto &pr : u0) {
      both[pr.first].push_back(pr.second);
      both[pr.second].push_back(pr.first);
    }
    for (auto &v : both) std::sort(v.begin(), v.end());
    std::printf("Starting Discretization.\n");
    curves = discretize(in, both);
    std::vector<Box> bx2(nc);
    for (int i = 0; i < nc; ++i) bx2[i] = curve_box(curves[i]);
    uniq = pls_unique(bx2, nc >= 50);
  } else {
    uniq = pls_unique(bx, nc >= 50);
    std::printf("Skipping Discretization.\n");
    curves = in;
  }
  std::vector<char> used(nc, 0);
  for (auto &pr : uniq) used[pr.first] = used[pr.second] = 1;
  std::vector<std::vector<V3>> sp(nc);
  for (int i = 0; i < nc; ++i)
    if (used[i]) sp[i] = start_points(curves[i]);
  const bool small = nc < kParallelCurves;
  std::vector<double> lk(uniq.size(), 0.0);
  if (force_ds) {
    std::printf("Starting Direct Summation.\n");
#pragma omp parallel for schedule(dynamic) if (!small)
    for (int64_t k = 0; k < (int64_t)uniq.size(); ++k) {
      const auto &pr = uniq[k];
      lk[k] = small ? ds_rows(sp[pr.first], sp[pr.second], 0, (int64_t)sp[pr.first].size(), true)
                    : ds_fast(sp[pr.first], sp[pr.second]);
    }
  } else {
    std::printf("Starting Barnes–Hut.\n");
    std::vector<SegTree> trees(nc);
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < nc; ++i)
      if (used[i]) trees[i] = make_tree(sp[i]);
#pragma omp parallel for schedule(dynamic) if (!small)
    for (int64_t k = 0; k < (int64_t)uniq.size(); ++k) {
      const auto &pr = uniq[k];
      double err = 0;
      double v = bh_eval(trees[pr.first], trees[pr.second], kEpsSq, beta_init * beta_init, err,
                         small, nullptr);
      err *= kErrScale;
      const double target = std::min(std::pow(err / kTargetErr, 0.25) * beta_init, beta_limit);
      if (target > beta_init) {
        if (small)
          std::printf("Error estimate of %g exceeded the target of %g; retrying with beta = %g.\n",
                      err, kTargetErr, target);
        err = 0;
        v = bh_eval(trees[pr.first], trees[pr.second], kEpsSq, target * target, err, small,
                    nullptr);
      }
      lk[k] = v;
    }
  }
  std::vector<Triplet> out;
  double sum = 0, mx = 0;
  for (size_t k = 0; k < uniq.size(); ++k) {
    const double r = std::round(lk[k]);
    const double e = std::fabs(lk[k] - r);
    sum += e;
    mx = std::max(mx, e);
    if (!(std::fabs(r) < 0.25)) {
      out.push_back({uniq[k].first, uniq[k].second, to_int(r)});
      if (std::isnan(r))
        std::fprintf(stderr, "Linking number (%d,%d) is NaN.\n", uniq[k].first, uniq[k].second);
    }
  }
  mean_err = uniq.empty() ? 0.0 : sum / (double)uniq.size();
  max_err = mx;
  if (!uniq.empty()) {
    std::printf("Linking-Number Computation, Mean Off-Integer Error: %g\n", mean_err);
    std::printf("Linking-Number Computation, Max  Off-Integer Error: %g\n", max_err);
  }
  std::sort(out.begin(), out.end(), [](const Triplet &a, const Triplet &b) {
    return a.i != b.i ? a.i < b.i : a.j < b.j;
  });
  if (lks) *lks = lk;
  if (cands) *cands = uniq;
  return out;
}

/* ---------------------------------------------------------------- BCC / OBJ */
struct BccHeader {  // model.cpp:268-277, 64 bytes
  char sign[3];
  unsigned char bytes;
  char type[2];
  char dims;
  char up;
  uint64_t ncurves;
  uint64_t npoints;
  char info[40];
};
static_assert(sizeof(BccHeader) == 64, "BCC header must be 64 bytes");

static std::vector<Curve> load_bcc(const std::string &path) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("Unable to load BCC file.");
  BccHeader h;
  auto fail = [&](const char *msg) {
    std::fclose(f);
    std::fprintf(stderr, "%s: %s\n", msg, path.c_str());
    throw std::runtime_error("Unable to load BCC file.");
  };
  if (std::fread(&h, sizeof h, 1, f) != 1) fail("Incomplete header");
  if (h.sign[0] != 'B' || h.sign[1] != 'C' || h.sign[2] != 'C') fail("Invalid header");
  if (h.bytes != 0x44) fail("Invalid bytecount");
  if (h.dims != 3) fail("Invalid dimensions");
  int type;
  if (h.type[0] == 'C' && h.type[1] == '0') type = ORC_CATMULLROM;
  else if (h.type[0] == 'B' && h.type[1] == 'S') type = ORC_BSPLINE;
  else if (h.type[0] == 'B' && h.type[1] == 'Z') type = ORC_BEZIER;
  else if (h.type[0] == 'P' && h.type[1] == 'L') type = ORC_POLYLINE;
  else fail("Invalid valid curve segment type");
  std::vector<Curve> cs(h.ncurves);
  std::vector<float> buf;
  for (auto &c : cs) {
    int32_t cnt;
    if (std::fread(&cnt, 4, 1, f) != 1) fail("Cannot read curve control point count");
    c.looped = cnt < 0;
    c.type = type;
    if (cnt < 0) cnt = -cnt;
    if (cnt) {
      buf.resize(3 * (size_t)cnt);
      if (std::fread(buf.data(), 4, buf.size(), f) != buf.size())
        fail("Invalid number of control points read");
      c.q.resize(cnt);
      for (int32_t k = 0; k < cnt; ++k)
        c.q[k] = V3((double)buf[3 * k], (double)buf[3 * k + 1], (double)buf[3 * k + 2]);
    }
  }
  std::fclose(f);
  return cs;
}

static void save_bcc(const std::vector<Curve> &cs, const std::string &path) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("Unable to export BCC file.");
  BccHeader h;
  std::memset(&h, 0, sizeof h);
  h.sign[0] = 'B'; h.sign[1] = 'C'; h.sign[2] = 'C';
  h.bytes = 0x44;
  h.type[0] = 'C'; h.type[1] = '0';
  if (!cs.empty()) {
    switch (cs[0].type) {
      case ORC_BSPLINE: h.type[0] = 'B'; h.type[1] = 'S'; break;
      case ORC_BEZIER: h.type[0] = 'B'; h.type[1] = 'Z'; break;
      case ORC_POLYLINE: h.type[0] = 'P'; h.type[1] = 'L'; break;
      default: break;
    }
  }
  h.dims = 3;
  h.up = 1;
  h.ncurves = cs.size();
  for (auto &c : cs) h.npoints += c.q.size();
  std::fwrite(&h, sizeof h, 1, f);
  std::vector<float> buf;
  for (auto &c : cs) {
    int32_t cnt = (int32_t)c.q.size();
    if (c.looped) cnt = -cnt;
    std::fwrite(&cnt, 4, 1, f);
    buf.resize(3 * c.q.size());
    for (size_t k = 0; k < c.q.size(); ++k) {
      buf[3 * k] = (float)c.q[k].x;
      buf[3 * k + 1] = (float)c.q[k].y;
      buf[3 * k + 2] = (float)c.q[k].z;
    }
    if (!buf.empty()) std::fwrite(buf.data(), 4, buf.size(), f);
  }
  std::fclose(f);
}

/* model.cpp:124-264: 'v' and 'l' records, Houdini '\\' continuation, Blender
 * line chaining (a new 'l' whose first vertex is not the previous last vertex
 * starts a new curve), closing duplicate removed. */
static bool approx_equal(V3 a, V3 b) {  // Eigen isApprox with 1e-12
  const double p = 1e-12;
  return norm2(a - b) <= p * p * std::min(norm2(a), norm2(b));
}
static std::string trim_ws(const std::string &s) {
  const char *ws = " \t\v\r";
  size_t e = s.find_last_not_of(ws);
  size_t b = s.find_first_not_of(ws);
  if (b == std::string::npos) return "";
  return s.substr(b, e + 1 - b);
}
static std::vector<Curve> load_obj(const std::string &path, int type) {
  std::ifstream f(path);
  if (!f.is_open()) throw std::runtime_error("Import OBJ model: unable to open file.");
  std::vector<V3> verts;
  std::vector<Curve> out;
  Curve cur;
  cur.type = type;
  int first_vid = -1, last_vid = -1;
  auto flush = [&]() {
    if (cur.q.size() > 1) {
      if (last_vid == first_vid || approx_equal(cur.q.back(), cur.q.front())) cur.q.pop_back();
      cur.looped = true;
      out.push_back(cur);
    }
  };
  std::string line;
  while (std::getline(f, line)) {
    line = trim_ws(line);
    if (line.empty() || line[0] == '#') continue;
    while (!line.empty() && line.back() == '\\') {
      std::string more;
      if (!std::getline(f, more)) break;
      line.pop_back();
      line += " " + trim_ws(more);
    }
    if (line.size() < 2 || !(line[1] == ' ' || line[1] == '\t' || line[1] == '\v')) continue;
    std::istringstream ss(line.substr(2));
    if (line[0] == 'v') {
      double x, y, z;
      if (!(ss >> x >> y >> z)) throw std::runtime_error("Obj input: Invalid vertex.");
      verts.push_back(V3(x, y, z));
    } else if (line[0] == 'l') {
      std::string tok;
      bool first = true;
      while (ss >> tok) {
        const int vid = std::stoi(tok.substr(0, tok.find('/')));
        if (first) {
          first = false;
          if (vid != last_vid) {
            if (cur.q.size() > 1) {
              flush();
              cur.q.clear();
              last_vid = -1;
            }
            cur.q.push_back(verts.at(vid - 1));
            last_vid = vid;
            first_vid = vid;
          }
        } else {
          if (vid >= 1 && vid < (int)verts.size() + 1) {
            cur.q.push_back(verts[vid - 1]);
            last_vid = vid;
          } else {
            throw std::runtime_error("Obj Import: illegal vertex ID.");
          }
        }
      }
    }
  }
  flush();
  return out;
}

}  // namespace orc

/* ================================================================ C ABI */
using namespace orc;

struct orc_model {
  std::vector<Curve> curves;
};

#define ORC_TRY(...)                  \
  try {                               \
    __VA_ARGS__                       \
  } catch (const std::exception &e) { \
    orc::g_err = e.what();            \
    return -1;                        \
  }

static Curve make_pl(const double *xyz, int64_t n) {
  Curve c;
  c.type = ORC_POLYLINE;
  c.looped = true;
  c.q.resize(n);
  for (int64_t k = 0; k < n; ++k) c.q[k] = V3(xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2]);
  return c;
}
static std::vector<V3> make_pts(const double *xyz, int64_t n) {
  std::vector<V3> v(n);
  for (int64_t k = 0; k < n; ++k) v[k] = V3(xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2]);
  return v;
}

extern "C" {

const char *orc_last_error(void) { return orc::g_err.c_str(); }

orc_model *orc_model_new(void) { return new orc_model(); }
void orc_model_free(orc_model *m) { delete m; }

int orc_model_add_curve(orc_model *m, const double *xyz, int64_t n, int seg_type, int looped) {
  Curve c;
  c.type = seg_type;
  c.looped = looped != 0;
  c.q = make_pts(xyz, n);
  m->curves.push_back(std::move(c));
  return 0;
}

int orc_model_load_bcc(orc_model *m, const char *path) {
  ORC_TRY({
    std::string p(path);
    size_t dot = p.rfind('.');
    std::string ext = dot == std::string::npos ? "" : p.substr(dot);
    for (auto &ch : ext) ch = (char)std::tolower((unsigned char)ch);
    if (ext != ".bcc") throw std::runtime_error("Input model file extension error.");
    m->curves = load_bcc(p);
    return 0;
  })
}
int orc_model_save_bcc(const orc_model *m, const char *path) {
  ORC_TRY({
    save_bcc(m->curves, path);
    return 0;
  })
}
int orc_model_load_obj(orc_model *m, const char *path, int seg_type) {
  ORC_TRY({
    m->curves = load_obj(path, seg_type);
    return 0;
  })
}
int32_t orc_model_ncurves(const orc_model *m) { return (int32_t)m->curves.size(); }
int64_t orc_model_curve_size(const orc_model *m, int32_t c) {
  return (int64_t)m->curves.at(c).q.size();
}
int orc_model_curve_info(const orc_model *m, int32_t c, int *seg_type, int *looped) {
  *seg_type = m->curves.at(c).type;
  *looped = m->curves.at(c).looped;
  return 0;
}
int orc_model_get_curve(const orc_model *m, int32_t c, double *xyz) {
  const auto &q = m->curves.at(c).q;
  for (size_t k = 0; k < q.size(); ++k) {
    xyz[3 * k] = q[k].x;
    xyz[3 * k + 1] = q[k].y;
    xyz[3 * k + 2] = q[k].z;
  }
  return 0;
}

int orc_point(const double *q, int64_t n, int seg_type, int looped, int64_t seg, double t,
              double *out) {
  ORC_TRY({
    Curve c;
    c.type = seg_type;
    c.looped = looped;
    c.q = make_pts(q, n);
    V3 v = point(c, seg, t);
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
    return 0;
  })
}
int orc_segment_box(const double *q, int64_t n, int seg_type, int looped, int64_t seg,
                    double tlo, double thi, double *out) {
  ORC_TRY({
    Curve c;
    c.type = seg_type;
    c.looped = looped;
    c.q = make_pts(q, n);
    Box b = segment_box(c, seg, tlo, thi);
    for (int d = 0; d < 3; ++d) { out[d] = b.mn[d]; out[3 + d] = b.mx[d]; }
    return 0;
  })
}
int orc_curve_boxes(const orc_model *m, double *out) {
  ORC_TRY({
    for (size_t i = 0; i < m->curves.size(); ++i) {
      Box b = curve_box(m->curves[i]);
      for (int d = 0; d < 3; ++d) { out[6 * i + d] = b.mn[d]; out[6 * i + 3 + d] = b.mx[d]; }
    }
    return 0;
  })
}

int64_t orc_pls(const orc_model *m, int use_tree, int32_t *pairs, int64_t cap) {
  ORC_TRY({
    std::vector<Box> bx(m->curves.size());
    for (size_t i = 0; i < bx.size(); ++i) bx[i] = curve_box(m->curves[i]);
    auto u = pls_unique(bx, use_tree != 0);
    for (int64_t k = 0; k < (int64_t)u.size() && k < cap; ++k) {
      pairs[2 * k] = u[k].first;
      pairs[2 * k + 1] = u[k].second;
    }
    return (int64_t)u.size();
  })
}

int orc_discretize(const orc_model *in, orc_model *out) {
  ORC_TRY({
    const int nc = (int)in->curves.size();
    std::vector<Box> bx(nc);
    for (int i = 0; i < nc; ++i) bx[i] = curve_box(in->curves[i]);
    auto u = pls_unique(bx, nc >= 50);
    std::vector<std::vector<int>> both(nc);
    for (auto &pr : u) {
      both[pr.first].push_back(pr.second);
      both[pr.second].push_back(pr.first);
    }
    for (auto &v : both) std::sort(v.begin(), v.end());
    out->curves = discretize(in->curves, both);
    return 0;
  })
}

double orc_ds_one_arctan(const double *p, int64_t n1, const double *q, int64_t n2, int parallel) {
  auto P = make_pts(p, n1), Q = make_pts(q, n2);
  return ds_rows(P, Q, 0, n1, parallel != 0);
}
double orc_ds_rows(const double *p, int64_t n1, const double *q, int64_t n2, int64_t r0,
                   int64_t r1, int parallel) {
  auto P = make_pts(p, n1), Q = make_pts(q, n2);
  return ds_rows(P, Q, r0, r1, parallel != 0);
}
double orc_ds_fast(const double *p, int64_t n1, const double *q, int64_t n2) {
  auto P = make_pts(p, n1), Q = make_pts(q, n2);
  return ds_fast(P, Q);
}
int orc_ds_pairs(const double *xyz, const int64_t *offsets, int32_t ncurves,
                 const int32_t *pairs, int64_t npairs, double *out) {
  try {
    std::vector<std::vector<V3>> pts(ncurves);
    std::vector<char> used(ncurves, 0);
    for (int64_t k = 0; k < 2 * npairs; ++k) used.at(pairs[k]) = 1;
    for (int32_t c = 0; c < ncurves; ++c)
      if (used[c]) pts[c] = make_pts(xyz + 3 * offsets[c], offsets[c + 1] - offsets[c]);
    const bool few = ncurves < kParallelCurves;
#pragma omp parallel for schedule(dynamic) if (!few)
    for (int64_t k = 0; k < npairs; ++k) {
      const auto &p = pts[pairs[2 * k]], &q = pts[pairs[2 * k + 1]];
      out[k] = few ? ds_rows(p, q, 0, (int64_t)p.size(), true) : ds_fast(p, q);
    }
    return 0;
  } catch (const std::exception &e) {
    orc::g_err = e.what();
    return -1;
  }
}

double orc_segment_pair_term(const double *p0, const double *p1, const double *q0,
                             const double *q1) {
  V3 P0(p0[0], p0[1], p0[2]), P1(p1[0], p1[1], p1[2]);
  V3 Q0(q0[0], q0[1], q0[2]), Q1(q1[0], q1[1], q1[2]);
  return pair_term_x4pi(Q0 - P0, Q0 - P1, Q1 - P1, Q1 - P0) * kInv4Pi;
}
double orc_bh_leaf_term(const double *m1, const double *t1, const double *m2, const double *t2) {
  Moment A, B;
  A.mid = V3(m1[0], m1[1], m1[2]);
  A.c = V3(t1[0], t1[1], t1[2]);
  B.mid = V3(m2[0], m2[1], m2[2]);
  B.c = V3(t2[0], t2[1], t2[2]);
  return leaf_term(A, B);
}

double orc_bh_pair(const double *p, int64_t n1, const double *q, int64_t n2, double beta,
                   int parallel, double *err, int64_t *visits) {
  try {
    SegTree A = make_tree(make_pts(p, n1)), B = make_tree(make_pts(q, n2));
    double e = 0;
    double v = bh_eval(A, B, kEpsSq, beta * beta, e, parallel != 0, visits);
    if (err) *err = e;
    return v;
  } catch (const std::exception &e) {
    orc::g_err = e.what();
    return std::numeric_limits<double>::quiet_NaN();
  }
}

static void pack_moment(const Moment &M, double *o) {
  int k = 0;
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) o[k++] = M.Q[a][i][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o[k++] = M.D[i][j];
  o[k++] = M.c.x; o[k++] = M.c.y; o[k++] = M.c.z;
  o[k++] = M.mid.x; o[k++] = M.mid.y; o[k++] = M.mid.z;
  o[k++] = M.cn; o[k++] = M.dn; o[k++] = M.qn;
}
static Moment unpack_moment(const double *o) {
  Moment M;
  int k = 0;
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) M.Q[a][i][j] = o[k++];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M.D[i][j] = o[k++];
  M.c = V3(o[k], o[k + 1], o[k + 2]); k += 3;
  M.mid = V3(o[k], o[k + 1], o[k + 2]); k += 3;
  M.cn = o[k++]; M.dn = o[k++]; M.qn = o[k++];
  return M;
}

int orc_tree_moments(const double *p, int64_t n, double *mom, double *boxes, int32_t *children,
                     int32_t *leaf_seg) {
  ORC_TRY({
    SegTree T = make_tree(make_pts(p, n));
    for (size_t k = 0; k < T.mom.size(); ++k) pack_moment(T.mom[k], mom + 45 * k);
    for (int k = 0; k < T.nbox(); ++k) {
      const Box &b = T.kd.boxes[k];
      for (int d = 0; d < 3; ++d) { boxes[6 * k + d] = b.mn[d]; boxes[6 * k + 3 + d] = b.mx[d]; }
      children[2 * k] = T.kd.children[2 * k];
      children[2 * k + 1] = T.kd.children[2 * k + 1];
    }
    for (int64_t k = 0; k < n; ++k) leaf_seg[k] = T.kd.obj[k];
    return 0;
  })
}

double orc_far_field(const double *m1, const double *m2, double r1, double r2, double epssq,
                     double *err) {
  Moment A = unpack_moment(m1), B = unpack_moment(m2);
  V3 r = B.mid - A.mid;
  double dsq = norm2(r) + epssq;
  double v = dot(r, cross(B.c, A.c));
  const double ir2 = 1.0 / dsq;
  const double mult = ir2 / std::sqrt(dsq) * kInv4Pi;
  double e;
  v += far_terms(A, B, r, ir2, r1, r2, e);
  if (err) *err = e * mult;
  return mult * v;
}

int64_t orc_certificate(const orc_model *m, int force_ds, int disc, double beta_init,
                        double beta_limit, int32_t *trip, int64_t cap, double *mean_err,
                        double *max_err, double *lk_out, int32_t *cand_out, int64_t cand_cap,
                        int64_t *ncand) {
  ORC_TRY({
    double me = 0, mx = 0;
    std::vector<double> lks;
    std::vector<std::pair<int, int>> cands;
    auto t = certificate(m->curves, force_ds != 0, disc != 0, beta_init, beta_limit, me, mx,
                         &lks, &cands);
    for (int64_t k = 0; k < (int64_t)t.size() && k < cap; ++k) {
      trip[3 * k] = t[k].i;
      trip[3 * k + 1] = t[k].j;
      trip[3 * k + 2] = t[k].v;
    }
    if (mean_err) *mean_err = me;
    if (max_err) *max_err = mx;
    if (ncand) *ncand = (int64_t)cands.size();
    for (int64_t k = 0; k < (int64_t)cands.size() && k < cand_cap; ++k) {
      if (lk_out) lk_out[k] = lks[k];
      if (cand_out) {
        cand_out[2 * k] = cands[k].first;
        cand_out[2 * k + 1] = cands[k].second;
      }
    }
    std::fflush(stdout);
    return (int64_t)t.size();
  })
}

}  // extern "C"
