/*
 * TEST INFRASTRUCTURE ONLY. Serial FP64 restatement of the mesh level set of
 * LevelSetObject::TriangleMeshParticle (kernel/helper/
 * LevelSetObjectGenerator.h) and of its grid sampling in Base::buildGrid, for
 * the tests to check the device path against. Only tests/ use it.
 *
 *   closest point   closestPointOnTriangle      LevelSetObjectGenerator.h:454-495
 *   nearest surface closestPointBVH / traverse  :701-756 (here a scan over all
 *                   triangles keeping the first strict minimum of |p - c|^2:
 *                   the BVH returns the same minimum distance)
 *   inside test     solidAngleFromPointToTriangle, isPointInsideClosedTriangleMesh
 *                                               :758-802
 *   value, gradient evaluateImplicitFunctionValueAndGradient :888-914
 *   grid            Base::buildGrid             :122-224
 * Operation order as written there (built with -ffp-contract=off).
 */
#include <algorithm>
#include <cmath>
#include <limits>

#include "oracle.h"

namespace {

struct P3 {
  double x, y, z;
};
P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
P3 add(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
P3 scale(P3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
P3 cross(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double len(P3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
const double kPi = 3.14159265358979323846;

P3 closest(P3 p, P3 a, P3 b, P3 c) {
  const P3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return a;
  const P3 bp = sub(p, b);
  const double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return b;
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return add(a, scale(ab, d1 / (d1 - d3)));
  const P3 cp = sub(p, c);
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return c;
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return add(a, scale(ac, d2 / (d2 - d6)));
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0)
    return add(b, scale(sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
  const double denom = 1.0 / (va + vb + vc);
  const double v = vb * denom, w = vc * denom;
  return add(add(a, scale(ab, v)), scale(ac, w));
}

double solidAngle(P3 p, P3 a, P3 b, P3 c) {
  const P3 r0 = sub(a, p), r1 = sub(b, p), r2 = sub(c, p);
  const double l0 = len(r0), l1 = len(r1), l2 = len(r2);
  if (l0 < 1e-30 || l1 < 1e-30 || l2 < 1e-30) return 0.0;
  const double num = dot(r0, cross(r1, r2));
  const double den = l0 * l1 * l2 + dot(r0, r1) * l2 + dot(r1, r2) * l0 + dot(r2, r0) * l1;
  return 2.0 * std::atan2(num, den);
}

bool valid(int nv, const int* f, int nf) {
  if (nv <= 0 || nf <= 0) return false;
  for (int t = 0; t < nf; ++t) {
    const int a = f[3 * t], b = f[3 * t + 1], c = f[3 * t + 2];
    if (a < 0 || b < 0 || c < 0 || a >= nv || b >= nv || c >= nv) return false;
    if (a == b || b == c || c == a) return false;
  }
  return true;
}

void evaluate(const double* v, const int* f, int nf, P3 p, int positiveInside, double& value, P3& grad) {
  auto vert = [&](int i) { return P3{v[3 * i], v[3 * i + 1], v[3 * i + 2]}; };
  double best = std::numeric_limits<double>::max();
  P3 cpBest{0.0, 0.0, 0.0};
  double sum = 0.0;
  for (int t = 0; t < nf; ++t) {
    const P3 a = vert(f[3 * t]), b = vert(f[3 * t + 1]), c = vert(f[3 * t + 2]);
    const P3 cp = closest(p, a, b, c);
    const double d2 = dot(sub(p, cp), sub(p, cp));
    if (d2 < best) {
      best = d2;
      cpBest = cp;
    }
  }
  for (int t = 0; t < nf; ++t)
    sum += solidAngle(p, vert(f[3 * t]), vert(f[3 * t + 1]), vert(f[3 * t + 2]));
  const double d = len(sub(p, cpBest));
  const bool inside = std::fabs(sum / (4.0 * kPi)) > 0.5;
  value = inside ? -d : d;
  grad = {0.0, 0.0, 0.0};
  if (d > 1e-30) {
    const double inv = 1.0 / d;
    const P3 g = scale(sub(p, cpBest), inv);
    grad = inside ? scale(g, -1.0) : g;
  }
  if (positiveInside) {
    value = -value;
    grad = scale(grad, -1.0);
  }
}

}  // namespace

extern "C" void oracle_mesh_points(const double* v, int nv, const int* f, int nf, const double* pts,
                                   int np, int positiveInside, double* value, double* grad) {
  const bool ok = valid(nv, f, nf);
  for (int i = 0; i < np; ++i) {
    double val = 1.0;
    P3 g{0.0, 0.0, 0.0};
    if (ok) evaluate(v, f, nf, {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}, positiveInside, val, g);
    value[i] = val;
    grad[3 * i] = g.x;
    grad[3 * i + 1] = g.y;
    grad[3 * i + 2] = g.z;
  }
}

extern "C" long long oracle_mesh_grid(const double* v, int nv, const int* f, int nf, double h,
                                      int padding, int positiveInside, double* origin3, int* size3,
                                      double* sdf) {
  size3[0] = size3[1] = size3[2] = 0;
  if (!valid(nv, f, nf) || h <= 0.0) return 0;
  P3 mn{v[0], v[1], v[2]}, mx = mn;
  for (int i = 1; i < nv; ++i) {
    mn = {std::min(mn.x, v[3 * i]), std::min(mn.y, v[3 * i + 1]), std::min(mn.z, v[3 * i + 2])};
    mx = {std::max(mx.x, v[3 * i]), std::max(mx.y, v[3 * i + 1]), std::max(mx.z, v[3 * i + 2])};
  }
  const double pad = double(std::max(1, padding)) * h;
  const P3 lo{h * std::floor((mn.x - pad) / h), h * std::floor((mn.y - pad) / h),
              h * std::floor((mn.z - pad) / h)};
  const P3 hi{h * std::ceil((mx.x + pad) / h), h * std::ceil((mx.y + pad) / h),
              h * std::ceil((mx.z + pad) / h)};
  const int nx = int(std::floor((hi.x - lo.x) / h + 0.5)) + 1;
  const int ny = int(std::floor((hi.y - lo.y) / h + 0.5)) + 1;
  const int nz = int(std::floor((hi.z - lo.z) / h + 0.5)) + 1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
  origin3[0] = lo.x;
  origin3[1] = lo.y;
  origin3[2] = lo.z;
  size3[0] = nx;
  size3[1] = ny;
  size3[2] = nz;
  const long long N = (long long)nx * ny * nz;
  if (!sdf) return N;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const P3 p{lo.x + double(i) * h, lo.y + double(j) * h, lo.z + double(k) * h};
        double val;
        P3 g;
        evaluate(v, f, nf, p, positiveInside, val, g);
        const double gn = len(g);
        sdf[(long long)k * ny * nx + (long long)j * nx + i] =
            gn > 1e-12 ? val / gn : (val <= 0.0 ? -1.0 : 1.0) * 0.5 * h;
      }
  return N;
}
