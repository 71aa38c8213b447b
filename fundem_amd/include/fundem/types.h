// Value types and host math of the FunDEM-compatible C++ API.
//
// double3 / int3 are HIP's vector types (24 and 12 bytes, as CUDA's).
// quaternion and symMatrix keep the reference's layouts (myQua.h:10-15,
// myMat.h:15-27). The free functions keep the reference's names and argument
// order for application code; internally every helper evaluates in the
// reference's operation order (component-wise, division by multiplying with
// the reciprocal), so that host-side setup arithmetic rounds exactly as the
// reference's does.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

struct quaternion {
  double q0, q1, q2, q3;
};

inline quaternion make_quaternion(double q0, double q1, double q2, double q3) {
  quaternion q;
  q.q0 = q0;
  q.q1 = q1;
  q.q2 = q2;
  q.q3 = q3;
  return q;
}

struct symMatrix {
  double xx, yy, zz, xy, xz, yz;
  symMatrix() : xx(0.0), yy(0.0), zz(0.0), xy(0.0), xz(0.0), yz(0.0) {}
  symMatrix(double a, double b, double c, double d, double e, double f)
      : xx(a), yy(b), zz(c), xy(d), xz(e), yz(f) {}
};

inline symMatrix make_symMatrix(double xx, double yy, double zz, double xy, double xz, double yz) {
  return symMatrix(xx, yy, zz, xy, xz, yz);
}

constexpr double pi() { return 3.14159265358979323846; }

inline bool isZero(const double x) { return std::fabs(x) <= 1.0e-20; }

namespace fundem {
namespace vec {
// explicit component-wise helpers (no reliance on operator overloads)
inline double3 add(double3 a, double3 b) { return make_double3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline double3 sub(double3 a, double3 b) { return make_double3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline double3 neg(double3 a) { return make_double3(-a.x, -a.y, -a.z); }
inline double3 mul(double3 a, double s) { return make_double3(a.x * s, a.y * s, a.z * s); }
inline double3 divr(double3 a, double s) {
  const double inv = 1.0 / s;
  return make_double3(a.x * inv, a.y * inv, a.z * inv);
}
inline double dot(double3 a, double3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double3 cross(double3 a, double3 b) {
  return make_double3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline double len(double3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
inline double3 unit(double3 v) {
  const double l2 = v.x * v.x + v.y * v.y + v.z * v.z;
  if (isZero(l2)) return make_double3(0.0, 0.0, 0.0);
  const double inv = 1.0 / std::sqrt(l2);
  return make_double3(v.x * inv, v.y * inv, v.z * inv);
}
inline quaternion unit(quaternion q) {
  const double l2 = q.q0 * q.q0 + q.q1 * q.q1 + q.q2 * q.q2 + q.q3 * q.q3;
  if (isZero(l2)) return q;
  const double inv = 1.0 / std::sqrt(l2);
  return make_quaternion(q.q0 * inv, q.q1 * inv, q.q2 * inv, q.q3 * inv);
}
inline double3 rotate(quaternion q, double3 v) {
  const double3 u = make_double3(q.q1, q.q2, q.q3);
  const double3 t = mul(cross(u, v), 2.0);
  return add(add(v, mul(t, q.q0)), cross(u, t));
}
inline double3 rotateT(double3 v, quaternion q) {
  const double3 u = make_double3(-q.q1, -q.q2, -q.q3);
  const double3 t = mul(cross(u, v), 2.0);
  return add(add(v, mul(t, q.q0)), cross(u, t));
}
}  // namespace vec
}  // namespace fundem

inline double dot(double3 a, double3 b) { return fundem::vec::dot(a, b); }
inline double3 cross(double3 a, double3 b) { return fundem::vec::cross(a, b); }
inline double lengthSquared(double3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
inline double length(double3 v) { return fundem::vec::len(v); }
inline double3 normalize(double3 v) { return fundem::vec::unit(v); }
inline quaternion normalize(quaternion q) { return fundem::vec::unit(q); }
inline double3 rotateVectorByQuaternion(quaternion q, double3 v) { return fundem::vec::rotate(q, v); }
inline double3 reverseRotateVectorByQuaternion(double3 v, quaternion q) {
  return fundem::vec::rotateT(v, q);
}
inline int linearIndex3D(const int3 ijk, const int3 dims) {
  return ijk.z * dims.y * dims.x + ijk.y * dims.x + ijk.x;
}
inline double3 operator*(symMatrix m, double3 v) {
  return make_double3(m.xx * v.x + m.xy * v.y + m.xz * v.z, m.xy * v.x + m.yy * v.y + m.yz * v.z,
                      m.xz * v.x + m.yz * v.y + m.zz * v.z);
}
// Inverse of a symmetric 3x3 matrix by cofactors; zero when |det| < 1e-30
// (the reference's threshold, myMat.h:163-194).
inline symMatrix inverse(symMatrix A) {
  const double det = A.xx * (A.yy * A.zz - A.yz * A.yz) + A.xy * (A.yz * A.xz - A.xy * A.zz) +
                     A.xz * (A.xy * A.yz - A.yy * A.xz);
  if (std::fabs(det) < 1e-30) return symMatrix(0, 0, 0, 0, 0, 0);
  const double di = 1.0 / det;
  return symMatrix((A.yy * A.zz - A.yz * A.yz) * di, (A.xx * A.zz - A.xz * A.xz) * di,
                   (A.xx * A.yy - A.xy * A.xy) * di, -(A.xy * A.zz - A.xz * A.yz) * di,
                   (A.xy * A.yz - A.xz * A.yy) * di, -(A.xx * A.yz - A.xz * A.xy) * di);
}
