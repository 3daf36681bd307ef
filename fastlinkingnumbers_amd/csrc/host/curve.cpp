// Curve geometry of the drop-in API (reference src/curve.cpp).
#include "curve.h"

#include <iostream>

namespace VerifyCurves {

namespace {

// Cubic bases B(t) of the three spline types (reference curve.cpp:20-43).
void basis(SegmentType type, double t, double *B) {
  switch (type) {
    case SegmentType::BSpline:
      B[0] = ((-(1.0 / 6) * t + (1.0 / 2)) * t - (1.0 / 2)) * t + (1.0 / 6);
      B[1] = ((1.0 / 2) * t - 1.0) * (t * t) + (2.0 / 3);
      B[2] = ((-(1.0 / 2) * t + (1.0 / 2)) * t + (1.0 / 2)) * t + (1.0 / 6);
      B[3] = (1.0 / 6) * (t * t) * t;
      break;
    case SegmentType::CatmullRom:
      B[0] = ((-0.5 * t + 1.0) * t - 0.5) * t;
      B[1] = (1.5 * t - 2.5) * (t * t) + 1.0;
      B[2] = ((-1.5 * t + 2.0) * t + 0.5) * t;
      B[3] = (0.5 * t - 0.5) * (t * t);
      break;
    case SegmentType::Bezier:
      B[0] = ((-1.0 * t + 3.0) * t - 3.0) * t + 1.0;
      B[1] = ((3.0 * t - 6.0) * t + 3.0) * t;
      B[2] = (-3.0 * t + 3.0) * (t * t);
      B[3] = 1.0 * t * t * t;
      break;
    default:
      throw std::runtime_error("Invalid segment type in Curve::get_point().");
  }
}

// Power-basis coefficients (t^3, t^2, t, 1) of one coordinate of a cubic
// segment, then its exact range over [t0, t1]: the endpoint values and the
// derivative roots inside, found with the cancellation-free quadratic formula
// (reference curve.cpp:197-277; complex roots give NaN and are rejected by
// every comparison).
void coord_range(SegmentType type, double a, double b, double c, double d, double t0, double t1,
                 double &lo, double &hi) {
  double k3, k2, k1, k0;
  if (type == SegmentType::BSpline) {
    k3 = -(1.0 / 6.0) * a + 0.5 * b - 0.5 * c + (1.0 / 6.0) * d;
    k2 = 0.5 * a - b + 0.5 * c;
    k1 = -0.5 * a + 0.5 * c;
    k0 = (1.0 / 6.0) * a + (2.0 / 3.0) * b + (1.0 / 6.0) * c;
  } else if (type == SegmentType::CatmullRom) {
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
  auto take = [&](double t) {
    if (t >= t0 && t <= t1) {
      const double v = f(t);
      lo = std::min(lo, v);
      hi = std::max(hi, v);
    }
  };
  lo = f(t0);
  hi = f(t1);
  if (hi < lo) std::swap(lo, hi);
  const double disc = 4 * k2 * k2 - 12 * k1 * k3;
  if (disc == 0) {
    take(-2 * k2 / (6 * k3));
  } else {
    const double r = k2 >= 0 ? -2 * k2 - std::sqrt(disc) : -2 * k2 + std::sqrt(disc);
    take(r / (6 * k3));
    take(2 * k1 / r);
  }
}

}  // namespace

Curve::Curve() {}

void Curve::Translate(const Point3 &d) {
  for (auto &p : q_) p += d;
}
void Curve::Rotate(const Matrix3d &R) {
  for (auto &p : q_) p = R * p;
}
void Curve::Scale(const Point3 &s) {
  for (auto &p : q_) p = p.cwiseProduct(s);
}

std::vector<Point3> Curve::GetMidpoints() const {
  if (!looped_) throw std::runtime_error("Curve::GetMidpoints() does not support unlooped curves.");
  std::vector<Point3> out(q_.size());
  for (size_t i = 0; i < q_.size(); ++i) out[i] = get_point((int)i, 0.5);
  return out;
}

std::vector<Point3> Curve::GetTangents() const {
  if (!looped_) throw std::runtime_error("Curve::GetTangents() does not support unlooped curves.");
  std::vector<Point3> s = GetStartPoints(), t(s.size());
  for (size_t i = 0; i + 1 < s.size(); ++i) t[i] = s[i + 1] - s[i];
  if (!s.empty()) t.back() = s.front() - s.back();
  return t;
}

std::vector<Point3> Curve::GetStartPoints() const {
  if (!looped_)
    throw std::runtime_error("Curve::GetStartPoints() does not support unlooped curves.");
  if (segment_type_ == SegmentType::Polyline) return q_;
  std::vector<Point3> out(q_.size());
  for (size_t i = 0; i < q_.size(); ++i) out[i] = get_point((int)i, 0.0);
  return out;
}

std::vector<Point3> Curve::GetLoopedStartPoints() const {
  if (!looped_)
    throw std::runtime_error("Curve::GetLoopedStartPoints() does not support unlooped curve.");
  std::vector<Point3> s = GetStartPoints();
  if (!s.empty()) s.push_back(s.front());
  return s;
}

Point3 Curve::get_point(int i, double t) const {
  if (t < 0.0 || t > 1.0 || i < 0 || i >= num_segments()) {
    std::cerr << "Curve::get_point() has an invalid i or t: i: " << i << ", t: " << t
              << ", n: " << q_.size() << "." << std::endl;
    throw std::runtime_error("Curve::get_point() has an invalid i or t.");
  }
  const size_t n = q_.size();
  if (segment_type_ == SegmentType::Polyline) return (1.0 - t) * q_[i] + t * q_[(i + 1) % n];
  double B[4];
  basis(segment_type_, t, B);
  Point3 r;
  for (int d = 0; d < 3; ++d)
    r[d] = B[0] * q_[i][d] + B[1] * q_[(i + 1) % n][d] + B[2] * q_[(i + 2) % n][d] +
           B[3] * q_[(i + 3) % n][d];
  return r;
}

Box3d Curve::get_segment_bounding_box(int i, double t_lo, double t_hi, double pad) const {
  if (i < 0 || i >= num_segments() || t_lo < 0 || t_hi > 1 || t_lo > t_hi) {
    std::cerr << "Curve::get_segment_bounding_box() has an invalid i or t: i: " << i
              << ", tlo: " << t_lo << ", thi: " << t_hi << ", n: " << q_.size() << "." << std::endl;
    throw std::runtime_error("Curve::get_segment_bounding_box() has an invalid i or t.");
  }
  Point3 mn, mx;
  if (segment_type_ == SegmentType::Polyline) {
    const Point3 a = get_point(i, t_lo), b = get_point(i, t_hi);
    mn = a.cwiseMin(b);
    mx = a.cwiseMax(b);
  } else {
    const size_t n = q_.size();
    for (int d = 0; d < 3; ++d)
      coord_range(segment_type_, q_[i][d], q_[(i + 1) % n][d], q_[(i + 2) % n][d],
                  q_[(i + 3) % n][d], t_lo, t_hi, mn[d], mx[d]);
  }
  for (int d = 0; d < 3; ++d) {
    mn[d] -= pad;
    mx[d] += pad;
  }
  return Box3d(mn, mx);
}

void Curve::ComputeCurveBoundingBox() const {
  const int ns = num_segments();
  if (ns <= 0) {
    bounding_box_ = Box3d();
  } else if (segment_type_ == SegmentType::Polyline) {
    // union of segment boxes = vertex min/max over the vertices segments touch
    const int nv = looped_ ? (int)q_.size() : ns + 1;
    Box3d b(q_[0], q_[0]);
    for (int k = 1; k < nv; ++k) {
      b.min() = b.min().cwiseMin(q_[k]);
      b.max() = b.max().cwiseMax(q_[k]);
    }
    bounding_box_ = b;
  } else {
    Box3d b = get_segment_bounding_box(0);
    for (int k = 1; k < ns; ++k) b.extend(get_segment_bounding_box(k));
    bounding_box_ = b;
  }
  bounding_box_initialized_ = true;
}

}  // namespace VerifyCurves
