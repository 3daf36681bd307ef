/*
 * curve.h — drop-in for the reference's include/curve.h (antequ/fastlinkingnumbers).
 *
 * Same class, enum and method names and the same semantics as
 * include/curve.h:35-195.  Eigen is not used: Point3, Box3d and Matrix3d are
 * small project value types exposing the Eigen members the reference API and
 * its callers use (x() y() z() [] data() norm() dot() cross(); min() max()
 * intersects() extend() merged() center() sizes()).
 */
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <vector>

#define PI (3.1415926535897932384626433832795028841971693993751058209749445923078164)

namespace VerifyCurves {

enum class SegmentType { BSpline, CatmullRom, Bezier, Polyline };

/* A 3-vector of doubles (the subset of Eigen::Vector3d the API uses). */
class Point3 {
 public:
  Point3() : v_{0, 0, 0} {}
  Point3(double x, double y, double z) : v_{x, y, z} {}
  double &x() { return v_[0]; }
  double &y() { return v_[1]; }
  double &z() { return v_[2]; }
  double x() const { return v_[0]; }
  double y() const { return v_[1]; }
  double z() const { return v_[2]; }
  double &operator[](int i) { return v_[i]; }
  double operator[](int i) const { return v_[i]; }
  double &operator()(int i) { return v_[i]; }
  double operator()(int i) const { return v_[i]; }
  double *data() { return v_; }
  const double *data() const { return v_; }
  Point3 operator+(const Point3 &o) const { return {v_[0] + o.v_[0], v_[1] + o.v_[1], v_[2] + o.v_[2]}; }
  Point3 operator-(const Point3 &o) const { return {v_[0] - o.v_[0], v_[1] - o.v_[1], v_[2] - o.v_[2]}; }
  Point3 operator-() const { return {-v_[0], -v_[1], -v_[2]}; }
  Point3 operator*(double s) const { return {v_[0] * s, v_[1] * s, v_[2] * s}; }
  Point3 &operator+=(const Point3 &o) { v_[0] += o.v_[0]; v_[1] += o.v_[1]; v_[2] += o.v_[2]; return *this; }
  Point3 &operator-=(const Point3 &o) { v_[0] -= o.v_[0]; v_[1] -= o.v_[1]; v_[2] -= o.v_[2]; return *this; }
  bool operator==(const Point3 &o) const { return v_[0] == o.v_[0] && v_[1] == o.v_[1] && v_[2] == o.v_[2]; }
  double dot(const Point3 &o) const { return v_[0] * o.v_[0] + v_[1] * o.v_[1] + v_[2] * o.v_[2]; }
  Point3 cross(const Point3 &o) const {
    return {v_[1] * o.v_[2] - v_[2] * o.v_[1], v_[2] * o.v_[0] - v_[0] * o.v_[2],
            v_[0] * o.v_[1] - v_[1] * o.v_[0]};
  }
  double squaredNorm() const { return dot(*this); }
  double norm() const { return std::sqrt(squaredNorm()); }
  Point3 cwiseProduct(const Point3 &o) const { return {v_[0] * o.v_[0], v_[1] * o.v_[1], v_[2] * o.v_[2]}; }
  Point3 cwiseMin(const Point3 &o) const {
    return {std::min(v_[0], o.v_[0]), std::min(v_[1], o.v_[1]), std::min(v_[2], o.v_[2])};
  }
  Point3 cwiseMax(const Point3 &o) const {
    return {std::max(v_[0], o.v_[0]), std::max(v_[1], o.v_[1]), std::max(v_[2], o.v_[2])};
  }
  /* Eigen::isApprox with the default double precision (1e-12). */
  bool isApprox(const Point3 &o, double prec = 1e-12) const {
    return (*this - o).squaredNorm() <= prec * prec * std::min(squaredNorm(), o.squaredNorm());
  }

 private:
  double v_[3];
};
inline Point3 operator*(double s, const Point3 &p) { return p * s; }

/* Row-major 3x3 matrix (used by Curve::Rotate). */
class Matrix3d {
 public:
  Matrix3d() : m_{0, 0, 0, 0, 0, 0, 0, 0, 0} {}
  static Matrix3d Identity() {
    Matrix3d r;
    r.m_[0] = r.m_[4] = r.m_[8] = 1.0;
    return r;
  }
  double &operator()(int i, int j) { return m_[3 * i + j]; }
  double operator()(int i, int j) const { return m_[3 * i + j]; }
  Point3 operator*(const Point3 &p) const {
    return {m_[0] * p[0] + m_[1] * p[1] + m_[2] * p[2], m_[3] * p[0] + m_[4] * p[1] + m_[5] * p[2],
            m_[6] * p[0] + m_[7] * p[1] + m_[8] * p[2]};
  }

 private:
  double m_[9];
};

/* Closed axis-aligned box with Eigen::AlignedBox semantics: a default box is
 * empty (min = +max double, max = lowest double); intersects() is closed. */
class Box3d {
 public:
  Box3d()
      : min_(std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
             std::numeric_limits<double>::max()),
        max_(std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest(),
             std::numeric_limits<double>::lowest()) {}
  Box3d(const Point3 &mn, const Point3 &mx) : min_(mn), max_(mx) {}
  Point3 &min() { return min_; }
  Point3 &max() { return max_; }
  const Point3 &min() const { return min_; }
  const Point3 &max() const { return max_; }
  bool isEmpty() const { return !(min_[0] <= max_[0] && min_[1] <= max_[1] && min_[2] <= max_[2]); }
  bool intersects(const Box3d &o) const {
    return min_[0] <= o.max_[0] && min_[1] <= o.max_[1] && min_[2] <= o.max_[2] &&
           o.min_[0] <= max_[0] && o.min_[1] <= max_[1] && o.min_[2] <= max_[2];
  }
  Box3d &extend(const Box3d &o) {
    min_ = min_.cwiseMin(o.min_);
    max_ = max_.cwiseMax(o.max_);
    return *this;
  }
  Box3d merged(const Box3d &o) const {
    Box3d b = *this;
    return b.extend(o);
  }
  Point3 center() const { return (min_ + max_) * 0.5; }
  Point3 sizes() const { return max_ - min_; }

 private:
  Point3 min_, max_;
};

/* include/curve.h:41-195. */
class Curve {
 public:
  Curve();
  int num_dofs() const { return (int)q_.size() * 3; }
  int num_segments() const {
    return looped_ ? (int)q_.size()
                   : ((segment_type_ == SegmentType::Polyline) ? (int)q_.size() - 1
                                                               : (int)q_.size() - 3);
  }
  void Translate(const Point3 &d);
  void Rotate(const Matrix3d &R);
  void Scale(const Point3 &s);
  std::vector<Point3> GetMidpoints() const;
  std::vector<Point3> GetTangents() const;
  std::vector<Point3> GetStartPoints() const;
  std::vector<Point3> GetLoopedStartPoints() const;
  Point3 get_point(int seg_index, double t) const;
  Box3d get_segment_bounding_box(int seg_index, double t_lo = 0.0, double t_hi = 1.0,
                                 double pad = 0.0) const;
  void ComputeCurveBoundingBox() const;
  Box3d get_last_curve_bounding_box() const {
    if (!bounding_box_initialized_)
      throw std::runtime_error(
          "Bounding box uninitialized. Use ComputeCurveBoundingBox() to initialize it.");
    return bounding_box_;
  }
  std::vector<Point3> &get_q() { return q_; }
  const std::vector<Point3> &get_q() const { return q_; }
  bool is_looped() const { return looped_; }
  void set_looped(bool looped) { looped_ = looped; }
  SegmentType get_segment_type() const { return segment_type_; }
  void set_segment_type(SegmentType type) { segment_type_ = type; }

 private:
  std::vector<Point3> q_;
  SegmentType segment_type_ = SegmentType::BSpline;
  bool looped_ = false;
  mutable Box3d bounding_box_;
  mutable bool bounding_box_initialized_ = false;
};

}  // namespace VerifyCurves
