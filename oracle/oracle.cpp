/*
 * TEST INFRASTRUCTURE ONLY. Serial FP64 restatement of FunDEM's level-set DEM
 * timestep (kaiqideng/FunDEM). Only tests/, the smoke check and bench.py's CPU
 * baseline use it, and only as the checker. Interface and switches: oracle.h.
 *
 * Arithmetic is written operation by operation in the reference's evaluation
 * order (left-to-right, no fused multiply-add: built with -ffp-contract=off),
 * so that the same FP64 operations on the same inputs give the same bits as
 * the device path. Each function cites the reference code it restates.
 *
 * Canonical force fold (canonical_fold = 1), used by fold_canonical():
 *   pair sums   For each interaction pair (A, target) - target a particle B,
 *               a wall W or a ghost of particle B - in increasing template node
 *               k:  SF += F_k;  STA += (r_c - r_A) x F_k;
 *               STB += (r_c - r_B) x (-F_k)      (r_B: ghost position for ghosts)
 *   particle i  F = T = 0, then add records in this order:
 *               1. particle-particle pairs by partner j ascending:
 *                  j > i: (+SF(i,j), STA(i,j));  j < i: (-SF(j,i), STB(j,i))
 *               2. particle-wall pairs by wall ascending: (SF, STA)
 *               3. ghost sets X, Y, XY in turn; within a set by partner real
 *                  id j ascending: first (i, ghost j) as A: (SF, STA), then
 *                  (j, ghost i) as B if j != i: (-SF, STB)
 *               4. bonds ascending: as master (F_ij, M_ij + (rb_i - r_i) x F_ij),
 *                  as slave (-F_ij, M_ji + (rb_j - r_j) x (-F_ij))
 *               Only pairs with at least one contact contribute.
 */
#include "oracle.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- math ----
// myVec.h:12-235, myQua.h:17-179, myMat.h:15-194 (host branches).
struct V3 { double x, y, z; };
struct Q4 { double q0, q1, q2, q3; };
struct Sym { double xx, yy, zz, xy, xz, yz; };
struct I3 { int x, y, z; };

static const double kPi = 3.14159265358979323846;  // myVec.h:14-17

static inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
// a*s and s*a both evaluate component*s (myVec.h:44-52)
static inline V3 scl(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
// operator/ multiplies by the reciprocal (myVec.h:54-59)
static inline V3 dvs(V3 a, double s) { double inv = 1.0 / s; return {a.x * inv, a.y * inv, a.z * inv}; }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline double len2(V3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
static inline double len(V3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
static inline bool isZero(double x) { return std::fabs(x) <= 1.0e-20; }  // myVec.h:22-25
static inline V3 normalize(V3 v) {                                    // myVec.h:121-137
  double l2 = v.x * v.x + v.y * v.y + v.z * v.z;
  if (isZero(l2)) return {0.0, 0.0, 0.0};
  double inv = 1.0 / std::sqrt(l2);
  return {v.x * inv, v.y * inv, v.z * inv};
}
static inline Q4 qnormalize(Q4 q) {  // myQua.h:51-62
  double l2 = q.q0 * q.q0 + q.q1 * q.q1 + q.q2 * q.q2 + q.q3 * q.q3;
  if (isZero(l2)) return q;
  double inv = 1.0 / std::sqrt(l2);
  return {q.q0 * inv, q.q1 * inv, q.q2 * inv, q.q3 * inv};
}
static inline V3 rotate(Q4 q, V3 v) {  // myQua.h:104-114
  V3 u = {q.q1, q.q2, q.q3};
  V3 t = scl(cross(u, v), 2.0);
  return add(add(v, scl(t, q.q0)), cross(u, t));
}
static inline V3 rotateT(V3 v, Q4 q) {  // myQua.h:120-129
  V3 u = {-q.q1, -q.q2, -q.q3};
  V3 t = scl(cross(u, v), 2.0);
  return add(add(v, scl(t, q.q0)), cross(u, t));
}
static inline Q4 qintegrate(Q4 q, V3 w, double dt) {  // myQua.h:77-93
  V3 v = scl(w, 0.5 * dt);
  Q4 d;
  d.q0 = -q.q1 * v.x - q.q2 * v.y - q.q3 * v.z;
  d.q1 = q.q0 * v.x + q.q3 * v.y - q.q2 * v.z;
  d.q2 = -q.q3 * v.x + q.q0 * v.y + q.q1 * v.z;
  d.q3 = q.q2 * v.x - q.q1 * v.y + q.q0 * v.z;
  return qnormalize({q.q0 + d.q0, q.q1 + d.q1, q.q2 + d.q2, q.q3 + d.q3});
}
static inline V3 symMul(Sym m, V3 v) {  // myMat.h:80-87
  return {m.xx * v.x + m.xy * v.y + m.xz * v.z, m.xy * v.x + m.yy * v.y + m.yz * v.z,
          m.xz * v.x + m.yz * v.y + m.zz * v.z};
}
static Sym rotateTensor(Q4 q, Sym I) {  // myMat.h:107-158
  double q0 = q.q0, q1 = q.q1, q2 = q.q2, q3 = q.q3;
  double q11 = 2 * q1 * q1, q22 = 2 * q2 * q2, q33 = 2 * q3 * q3;
  double a = 1.0 - q22 - q33, b = 2 * (q1 * q2 - q0 * q3), c = 2 * (q1 * q3 + q0 * q2);
  double d = 2 * (q1 * q2 + q0 * q3), e = 1.0 - q11 - q33, f = 2 * (q2 * q3 - q0 * q1);
  double g = 2 * (q1 * q3 - q0 * q2), h = 2 * (q2 * q3 + q0 * q1), i = 1.0 - q11 - q22;
  double r0x = I.xx * a + I.xy * b + I.xz * c, r0y = I.xy * a + I.yy * b + I.yz * c,
         r0z = I.xz * a + I.yz * b + I.zz * c;
  double r1x = I.xx * d + I.xy * e + I.xz * f, r1y = I.xy * d + I.yy * e + I.yz * f,
         r1z = I.xz * d + I.yz * e + I.zz * f;
  double r2x = I.xx * g + I.xy * h + I.xz * i, r2y = I.xy * g + I.yy * h + I.yz * i,
         r2z = I.xz * g + I.yz * h + I.zz * i;
  return {r0x * a + r0y * b + r0z * c, r1x * d + r1y * e + r1z * f, r2x * g + r2y * h + r2z * i,
          r0x * d + r0y * e + r0z * f, r0x * g + r0y * h + r0z * i, r1x * g + r1y * h + r1z * i};
}
static Sym inverse(Sym A) {  // myMat.h:163-194
  const double det = A.xx * (A.yy * A.zz - A.yz * A.yz) + A.xy * (A.yz * A.xz - A.xy * A.zz) +
                     A.xz * (A.xy * A.yz - A.yy * A.xz);
  if (std::fabs(det) < 1e-30) return {0, 0, 0, 0, 0, 0};
  const double di = 1.0 / det;
  return {(A.yy * A.zz - A.yz * A.yz) * di,  (A.xx * A.zz - A.xz * A.xz) * di,
          (A.xx * A.yy - A.xy * A.xy) * di,  -(A.xy * A.zz - A.xz * A.yz) * di,
          (A.xy * A.yz - A.xz * A.yy) * di,  -(A.xx * A.yz - A.xz * A.xy) * di};
}

// Double -> int as the device converts (saturating, NaN -> 0); C++'s (int) is
// undefined out of range, so make the conversion explicit.
static inline int toInt(double v) {
  if (v != v) return 0;
  if (v >= 2147483647.0) return INT_MAX;
  if (v <= -2147483648.0) return INT_MIN;
  return (int)v;
}
static inline int lin3(I3 c, I3 d) { return c.z * d.y * d.x + c.y * d.x + c.x; }  // myVec.h:232-235
static inline I3 gridPos(V3 p, V3 mn, V3 inv) {  // mySpatialGrid.h:24-31
  return {toInt((p.x - mn.x) * inv.x), toInt((p.y - mn.y) * inv.y), toInt((p.z - mn.z) * inv.z)};
}
static inline double r32(double x, bool on) { return on ? (double)(float)x : x; }
static inline V3 r32(V3 v, bool on) { return {r32(v.x, on), r32(v.y, on), r32(v.z, on)}; }
static inline Q4 r32(Q4 q, bool on) { return {r32(q.q0, on), r32(q.q1, on), r32(q.q2, on), r32(q.q3, on)}; }

static inline V3 ld3(const double* p) { return {p[0], p[1], p[2]}; }
static inline void st3(double* p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// ------------------------------------------------------------ grid ----
struct Grid {
  V3 mn{0, 0, 0}, mx{0, 0, 0}, inv{0, 0, 0};
  I3 size{2, 2, 2};
  long long num = 8;
};
static void gridSet(Grid& g, V3 mn, V3 mx, double cs) {  // mySpatialGrid.h:212-240
  if (mx.x <= mn.x) mx.x = mn.x + cs;
  if (mx.y <= mn.y) mx.y = mn.y + cs;
  if (mx.z <= mn.z) mx.z = mn.z + cs;
  g.mn = mn;
  g.mx = mx;
  if (isZero(cs)) { g.size = {2, 2, 2}; g.inv = {0, 0, 0}; g.num = 8; return; }
  V3 L = sub(mx, mn);
  g.size.x = L.x > 2. * cs ? int(L.x / cs) : 2;
  g.size.y = L.y > 2. * cs ? int(L.y / cs) : 2;
  g.size.z = L.z > 2. * cs ? int(L.z / cs) : 2;
  g.inv.x = (L.x > 0.0) ? (double(g.size.x) / L.x) : 0.0;
  g.inv.y = (L.y > 0.0) ? (double(g.size.y) / L.y) : 0.0;
  g.inv.z = (L.z > 0.0) ? (double(g.size.z) / L.z) : 0.0;
  g.num = (long long)g.size.x * g.size.y * g.size.z;
}
static int hashOf(V3 p, const Grid& g) {  // mySpatialGrid.cu:4-26
  if (p.x < g.mn.x) return -1; else if (p.x >= g.mx.x) return -1;
  if (p.y < g.mn.y) return -1; else if (p.y >= g.mx.y) return -1;
  if (p.z < g.mn.z) return -1; else if (p.z >= g.mx.z) return -1;
  return lin3(gridPos(p, g.mn, g.inv), g.size);
}
static int ghostHashOf(V3 p, const Grid& g) {  // mySpatialGrid.cu:214-240
  I3 gp = gridPos(p, g.mn, g.inv);
  bool out = false;
  if (p.x < g.mn.x) { gp.x = 0; out = true; } else if (p.x >= g.mx.x) { gp.x = g.size.x - 1; out = true; }
  if (p.y < g.mn.y) { gp.y = 0; out = true; } else if (p.y >= g.mx.y) { gp.y = g.size.y - 1; out = true; }
  if (p.z < g.mn.z) { gp.z = 0; out = true; } else if (p.z >= g.mx.y) { gp.z = g.size.z - 1; out = true; }
  return out ? lin3(gp, g.size) : -1;
}
struct HashTable {
  std::vector<int> hash, index, start, end;
};
// buildHashStartEnd.cu:18-137: identity index, stable sort by hash (thrust's
// radix sort on int keys is stable; negative keys order first), start/end.
static void sortStartEnd(std::vector<int>& hash, std::vector<int>& index, long long numGrid,
                         std::vector<int>& start, std::vector<int>& end) {
  const size_t n = hash.size();
  index.resize(n);
  std::iota(index.begin(), index.end(), 0);
  std::stable_sort(index.begin(), index.end(), [&](int a, int b) { return hash[a] < hash[b]; });
  std::vector<int> sorted(n);
  for (size_t i = 0; i < n; ++i) sorted[i] = hash[index[i]];
  hash.swap(sorted);
  start.assign((size_t)numGrid, -1);
  end.assign((size_t)numGrid, -1);
  for (size_t i = 0; i < n; ++i) {  // findStartAndEnd (buildHashStartEnd.cu:47-67)
    int h = hash[i];
    if (h < 0 || h >= numGrid) continue;
    if (i == 0 || hash[i - 1] != h) start[h] = (int)i;
    if (i == n - 1 || hash[i + 1] != h) end[h] = (int)(i + 1);
  }
}

// -------------------------------------------------------- templates ----
struct Geoms {
  std::vector<double> radius, invMass, invh;
  std::vector<Sym> I, Iinv;
  std::vector<V3> origin;
  std::vector<I3> size;
  std::vector<long long> sdfStart;
  std::vector<double> sdf;
  std::vector<int> nodeStart, nodeCount;
  std::vector<V3> nodes;
  int num() const { return (int)radius.size(); }
};
// diss: the damping ratio of the contact dashpots, -ln e / sqrtf(ln^2 e + pi^2)
// as contactKernel.cuh:50-51 evaluates it, or given directly (damping-ratio
// materials); a pair uses the one of its material with the smaller e.
struct Materials {
  std::vector<double> kn, ks, mu, e, diss;
};

// ------------------------------------------------ level-set sampling ----
// LSParticleContactDetectionKernel.cuh:4-32
static inline double trilinear(double x, double y, double z, const double* c) {
  const double wx0 = 1.0 - x, wx1 = x, wy0 = 1.0 - y, wy1 = y, wz0 = 1.0 - z, wz1 = z;
  return c[0] * wx0 * wy0 * wz0 + c[1] * wx1 * wy0 * wz0 + c[2] * wx0 * wy1 * wz0 +
         c[3] * wx1 * wy1 * wz0 + c[4] * wx0 * wy0 * wz1 + c[5] * wx1 * wy0 * wz1 +
         c[6] * wx0 * wy1 * wz1 + c[7] * wx1 * wy1 * wz1;
}
// LSParticleContactDetectionKernel.cuh:34-77. The x weight parameter is an
// int there, so the fractional x in [0,1) arrives as 0: wx0 = 1, wx1 = 0.
static inline V3 gradientQ1(double xf, double y, double z, const double* c) {
  const int xi = (int)xf;
  const double wx0 = 1.0 - xi, wx1 = xi, wy0 = 1.0 - y, wy1 = y, wz0 = 1.0 - z, wz1 = z;
  double dx = (c[1] - c[0]) * wy0 * wz0 + (c[3] - c[2]) * wy1 * wz0 + (c[5] - c[4]) * wy0 * wz1 +
              (c[7] - c[6]) * wy1 * wz1;
  double dy = (c[2] - c[0]) * wx0 * wz0 + (c[3] - c[1]) * wx1 * wz0 + (c[6] - c[4]) * wx0 * wz1 +
              (c[7] - c[5]) * wx1 * wz1;
  double dz = (c[4] - c[0]) * wx0 * wy0 + (c[5] - c[1]) * wx1 * wy0 + (c[6] - c[2]) * wx0 * wy1 +
              (c[7] - c[3]) * wx1 * wy1;
  return {dx, dy, dz};
}
// Cell lookup of a local-frame point on one grid (LSParticleContactDetectionKernel.cu:72-92):
// the fractional position in the cell and its eight corner values; false when
// the cell is not inside the grid.
struct CellSample {
  double x, y, z, c[8];
};
static bool cellSample(V3 local, V3 o, double ih, I3 sz, const double* s, CellSample& cs) {
  const double gx = (local.x - o.x) * ih, gy = (local.y - o.y) * ih, gz = (local.z - o.z) * ih;
  const int i0 = toInt(std::floor(gx)), j0 = toInt(std::floor(gy)), k0 = toInt(std::floor(gz));
  if (i0 < 0 || j0 < 0 || k0 < 0) return false;
  if (i0 >= sz.x - 1 || j0 >= sz.y - 1 || k0 >= sz.z - 1) return false;
  const int i1 = i0 + 1, j1 = j0 + 1, k1 = k0 + 1;
  cs.x = gx - (double)i0;
  cs.y = gy - (double)j0;
  cs.z = gz - (double)k0;
  const I3 corner[8] = {{i0, j0, k0}, {i1, j0, k0}, {i0, j1, k0}, {i1, j1, k0},
                        {i0, j0, k1}, {i1, j0, k1}, {i0, j1, k1}, {i1, j1, k1}};
  for (int q = 0; q < 8; ++q) cs.c[q] = s[lin3(corner[q], sz)];
  return true;
}
// Shared body of the four detection kernels from the local-frame position on:
// LSParticleContactDetectionKernel.cu:72-119 (count) and :218-287 (write).
// Returns true on contact (-phi >= 0) and fills overlap and the world normal.
static bool sampleContact(V3 local, Q4 qB, int geoB, const Geoms& G, double& overlap, V3& normal) {
  CellSample cs;
  if (!cellSample(local, G.origin[geoB], G.invh[geoB], G.size[geoB],
                  G.sdf.data() + G.sdfStart[geoB], cs))
    return false;
  const double ov = -trilinear(cs.x, cs.y, cs.z, cs.c);
  if (!(ov >= 0.)) return false;
  overlap = ov;
  normal = normalize(rotate(qB, gradientQ1(cs.x, cs.y, cs.z, cs.c)));
  return true;
}

// ------------------------------------------------------- contact law ----
// contactKernel.cuh:4-34
static V3 integrateSlidingSpring(V3 prev, V3 vt, V3 n, V3 Fn, double mu, double ks, double ct,
                                 double dt) {
  V3 spring = {0., 0., 0.};
  if (mu > 0. && ks > 0.) {
    V3 p1 = sub(prev, scl(n, dot(prev, n)));
    const double a1 = len(p1);
    if (!isZero(a1)) p1 = scl(p1, len(prev) / a1);
    spring = add(p1, scl(vt, dt));
    V3 sf = sub(scl(spring, -ks), scl(vt, ct));
    const double asf = len(sf), afn = len(Fn);
    if (!isZero(asf) && asf > mu * afn) {
      const double ratio = mu * afn / asf;
      sf = scl(sf, ratio);
      spring = dvs(neg(add(sf, scl(vt, ct))), ks);
    }
  }
  return spring;
}
// contactKernel.cuh:36-77 (the FP32 square roots are the reference's sqrtf).
static double dissipationOf(double e) {
  const double logRes = std::log(e);
  return -logRes / (double)sqrtf((float)(logRes * logRes + kPi * kPi));
}
static void linearContact(V3& F, V3& spring, V3 v, V3 n, double delta, double dt, double kn,
                          double ks, double mu, double diss, double effM) {
  if (delta > 0.) {
    const double cn = 2. * diss * (double)sqrtf((float)(effM * kn));
    const V3 vn = scl(n, dot(v, n));
    const V3 Fn = sub(scl(n, kn * delta), scl(vn, cn));
    const double ct = 2. * diss * (double)sqrtf((float)(effM * ks));
    const V3 vt = sub(v, vn);
    spring = integrateSlidingSpring(spring, vt, n, Fn, mu, ks, ct, dt);
    const V3 Fs = sub(scl(spring, -ks), scl(vt, ct));
    F = add(Fn, Fs);
  } else {
    spring = {0., 0., 0.};
  }
}

// ----------------------------------------------------------- bonds ----
struct Bonds {
  std::vector<int> master, slave, active;
  std::vector<double> B1, B2, B3, B4, R, L0, st, coh, mub;
  std::vector<V3> n1i, n2i, n3i, pi_, n1j, n2j, n3j, pj;
  std::vector<double> sigma, tau;
  int num() const { return (int)master.size(); }
};
// contactKernel.cuh:138-199
static int vbond(V3& F, V3& Mij, V3& Mji, double& sig, double& tau, V3 rbi, V3 rbj, V3 n1i,
                 V3 n2i, V3 n3i, V3 n1j, V3 n2j, V3 n3j, double B1, double B2, double B3, double B4,
                 double R, double D0, double st, double coh, double mub) {
  const V3 rbij = sub(rbi, rbj);
  const double D = len(rbij);
  if (isZero(D)) return 1;
  const V3 nij = normalize(rbij);
  const V3 e = neg(nij);
  const V3 dn1 = sub(n1j, n1i);
  F = add(scl(e, B1 * (D - D0)), scl(sub(dn1, scl(e, dot(dn1, e))), B2 / (2 * D)));
  const V3 MTB = sub(scl(cross(n1j, n1i), B3), scl(add(cross(n2j, n2i), cross(n3j, n3i)), 0.5 * B4));
  Mij = add(scl(cross(e, n1i), -0.5 * B2), MTB);
  Mji = sub(scl(cross(e, n1j), 0.5 * B2), MTB);
  const double Fn = dot(F, nij);
  const V3 Fs = sub(F, scl(nij, Fn));
  const double Tt = dot(MTB, nij);
  const V3 Tb = sub(MTB, scl(nij, Tt));
  const double A = R * R * kPi;
  const double I = A * R * R / 4.;
  const double J = 2 * I;
  sig = -Fn / A + len(Tb) / I * R;
  tau = len(Fs) / A + std::fabs(Tt) / J * R;
  if (st > 0. && sig > st) return 0;
  else if (coh > 0. && tau > coh - mub * sig) return 0;
  return 1;
}
// VBondedInteraction.h:10-22
static V3 bondN2(V3 n1) {
  const V3 ref = (std::fabs(n1.x) < 0.9) ? V3{1., 0., 0.} : V3{0., 1., 0.};
  return normalize(cross(ref, n1));
}

// ----------------------------------------------------------- solver ----
struct Contact {
  int A, B, k;
  V3 point, normal, spring;
  double overlap;
  V3 F;
};
enum { SET_PP = 0, SET_PW = 1, SET_GX = 2, SET_GY = 3, SET_GXY = 4, NSETS = 5 };

struct Body {  // SoA state of LSParticle (LSParticle.h:821-834)
  std::vector<V3> x, v, w, F, T;
  std::vector<Q4> q;
  std::vector<int> mat, geo;
  int num() const { return (int)x.size(); }
};
struct GhostSet {
  std::vector<V3> pos, vel;
  std::vector<Q4> q;
  HashTable ht;
};

}  // namespace orc

using namespace orc;

struct OracleSim {
  bool f32 = false, canon = false;
  Materials pmat, wmat;
  Geoms pgeo, wgeo;
  Body P, W;
  Bonds bonds;
  bool perX = false, perY = false;
  bool sector = false;  // PeriodicBoundarySector: ghost sets 0, 1, 2 are R90, R180, R270
  double damping = -1.0;  // the tutorials' global damping (off when < 0)
  int constIndex = -1;    // the tutorials' constant force (off when < 0)
  V3 constF{0, 0, 0}, constT{0, 0, 0};
  double cellSizeOverride = 0.0;
  Grid pgrid, wgrid;
  HashTable pht, wht;
  GhostSet gh[3];
  std::vector<Contact> C[NSETS];
  std::vector<long long> nodeBase;  // first global node id of each particle
  V3 gravity{0, 0, 0};
  double dt = 0.0;
  bool initialized = false;

  static uint64_t key(long long node, int B) { return ((uint64_t)node << 32) | (uint32_t)B; }

  // ------------------------------------------------------- stage 1 ----
  // particleVelocityAngularVelocityIntegrationKernel (particleIntegrationKernel.cu:3-32)
  void kick(double h) {
    for (int i = 0; i < P.num(); ++i) {
      const int g = P.geo[i];
      const double invM = pgeo.invMass[g];
      if (isZero(invM)) continue;
      P.v[i] = r32(add(P.v[i], scl(add(scl(P.F[i], invM), gravity), h)), f32);
      const V3 w = P.w[i];
      const Q4 q = P.q[i];
      const V3 Iw = symMul(rotateTensor(q, pgeo.I[g]), w);
      P.w[i] = r32(add(w, scl(symMul(rotateTensor(q, pgeo.Iinv[g]), sub(P.T[i], cross(w, Iw))), h)), f32);
    }
  }
  // particlePositionOrientationIntegrationKernel (particleIntegrationKernel.cu:34-48)
  void drift(Body& B, double step) {
    for (int i = 0; i < B.num(); ++i) {
      B.x[i] = add(B.x[i], scl(B.v[i], step));
      B.q[i] = r32(qintegrate(B.q[i], B.w[i], step), f32);
    }
  }
  // ------------------------------------------------------- stage 2 ----
  void buildGrid(const Body& B, const Grid& g, HashTable& ht) {  // LSSolver.h:747-778
    ht.hash.resize(B.num());
    for (int i = 0; i < B.num(); ++i) ht.hash[i] = hashOf(B.x[i], g);
    sortStartEnd(ht.hash, ht.index, g.num, ht.start, ht.end);
  }
  // --------------------------------------------------- detection ----
  // One node of A against a target set through the 27-cell stencil
  // (LSParticleContactDetectionKernel.cu:35-123 / :336-426). pp selects the
  // particle-particle filters (index order and bounding sphere).
  template <class Emit>
  void searchNode(V3 p, int A, bool pp, const Grid& g, const HashTable& ht,
                  const std::vector<V3>& pos, const std::vector<Q4>& qs, const std::vector<int>& geo,
                  const Geoms& G, Emit emit) const {
    I3 gp = gridPos(p, g.mn, g.inv);
    I3 s = {-1, -1, -1}, e = {1, 1, 1};
    if (gp.x <= 0) { gp.x = 0; s.x = 0; }
    if (gp.x >= g.size.x - 1) { gp.x = g.size.x - 1; e.x = 0; }
    if (gp.y <= 0) { gp.y = 0; s.y = 0; }
    if (gp.y >= g.size.y - 1) { gp.y = g.size.y - 1; e.y = 0; }
    if (gp.z <= 0) { gp.z = 0; s.z = 0; }
    if (gp.z >= g.size.z - 1) { gp.z = g.size.z - 1; e.z = 0; }
    for (int zz = s.z; zz <= e.z; zz++)
      for (int yy = s.y; yy <= e.y; yy++)
        for (int xx = s.x; xx <= e.x; xx++) {
          const int h = lin3({gp.x + xx, gp.y + yy, gp.z + zz}, g.size);
          const int st = ht.start[h];
          if (st == -1) continue;
          const int en = ht.end[h];
          for (int i = st; i < en; i++) {
            const int B = ht.index[i];
            if (pp && A >= B) continue;
            const V3 rel = sub(p, pos[B]);
            if (pp) {
              const double rB = G.radius[geo[B]];
              if (len2(rel) > rB * rB) continue;
            }
            double ov;
            V3 n;
            if (sampleContact(rotateT(rel, qs[B]), qs[B], geo[B], G, ov, n)) emit(B, ov, n);
          }
        }
  }
  // Count/write kernels of one interaction set, with the history remap of
  // LSParticleContactDetectionKernel.cu:288-297 (lookup by (node, object)).
  void detectSet(int set, const Grid& g, const HashTable& ht, const std::vector<V3>& pos,
                 const std::vector<Q4>& qs, const std::vector<int>& geo, const Geoms& G, bool pp) {
    std::unordered_map<uint64_t, V3> old;
    old.reserve(C[set].size() * 2 + 1);
    for (const Contact& c : C[set]) old[key(nodeBase[c.A] + c.k, c.B)] = c.spring;
    std::vector<Contact> out;
    for (int A = 0; A < P.num(); ++A) {
      const int ga = P.geo[A];
      const int n0 = pgeo.nodeStart[ga], nn = pgeo.nodeCount[ga];
      for (int k = 0; k < nn; ++k) {
        const V3 p = add(rotate(P.q[A], pgeo.nodes[n0 + k]), P.x[A]);
        searchNode(p, A, pp, g, ht, pos, qs, geo, G, [&](int B, double ov, V3 n) {
          Contact c;
          c.A = A; c.B = B; c.k = k;
          c.overlap = ov;
          c.normal = n;
          c.point = add(p, scl(n, 0.5 * ov));
          auto it = old.find(key(nodeBase[A] + k, B));
          c.spring = (it == old.end()) ? V3{0., 0., 0.} : it->second;
          c.F = {0, 0, 0};
          out.push_back(c);
        });
      }
    }
    C[set].swap(out);
  }
  void detectAll(bool withGhosts) {
    detectSet(SET_PP, pgrid, pht, P.x, P.q, P.geo, pgeo, true);      // LSSolver.h:788-842
    detectSet(SET_PW, wgrid, wht, W.x, W.q, W.geo, wgeo, false);     // LSSolver.h:849-910
    if (withGhosts) updateGhosts();
  }
  // ------------------------------------------------------ periodic ----
  // PeriodicBoundaryXY2D::updateGhostSpatialGrid (PeriodicBoundary.h:221-295)
  // and buildGhostInteraction (:297-302).
  void wrapAxis(int axis) {  // mySpatialGrid.cu:28-56
    const V3 off = axis == 0 ? V3{pgrid.mx.x - pgrid.mn.x, 0., 0.} : V3{0., pgrid.mx.y - pgrid.mn.y, 0.};
    for (int i = 0; i < P.num(); ++i) {
      V3 p = P.x[i];
      const double c = axis == 0 ? p.x : p.y;
      const double lo = axis == 0 ? pgrid.mn.x : pgrid.mn.y, hi = axis == 0 ? pgrid.mx.x : pgrid.mx.y;
      if (c < lo) p = add(p, off);
      else if (c >= hi) p = sub(p, off);
      P.x[i] = p;
    }
  }
  void ghostPositions(int set) {  // mySpatialGrid.cu:87-152
    GhostSet& G = gh[set];
    const V3 offX = {pgrid.mx.x - pgrid.mn.x, 0., 0.}, offY = {0., pgrid.mx.y - pgrid.mn.y, 0.};
    G.pos.resize(P.num());
    for (int i = 0; i < P.num(); ++i) {
      const V3 p = P.x[i];
      const I3 c = gridPos(p, pgrid.mn, pgrid.inv);
      V3 gp = p;
      if (set == 0 && c.x == 0) gp = add(gp, offX);
      if (set == 1 && c.y == 0) gp = add(gp, offY);
      if (set == 2 && c.x == 0 && c.y == 0) gp = add(gp, add(offX, offY));
      G.pos[i] = gp;
    }
    G.ht.hash.resize(P.num());
    for (int i = 0; i < P.num(); ++i) G.ht.hash[i] = ghostHashOf(G.pos[i], pgrid);
    sortStartEnd(G.ht.hash, G.ht.index, pgrid.num, G.ht.start, G.ht.end);
    G.q = P.q;
    G.vel = P.v;
  }
  // rotateVector (myVec.h:143-181, host branch) by the angle th about z.
  static V3 rotZ(V3 v, double angle) {
    const volatile double thv = angle;  // sin and cos from the C library at run time
    const double th = thv;
    const V3 om = {0.0, 0.0, th};
    const double theta2 = dot(om, om);
    if (isZero(theta2)) return add(add(v, cross(om, v)), scl(cross(om, cross(om, v)), 0.5));
    const double invTheta = 1.0 / std::sqrt(theta2);
    const double theta = theta2 * invTheta;
    const double s = std::sin(theta), c = std::cos(theta);
    const double A = s * invTheta;
    const double B = (1.0 - c) * (invTheta * invTheta);
    const V3 oxv = cross(om, v);
    const V3 oxoxv = cross(om, oxv);
    return add(add(v, scl(oxv, A)), scl(oxoxv, B));
  }
  // rotateQuaternion (myQua.h:131-175): q rotated by th about z, normalised.
  static Q4 rotQz(Q4 q, double angle) {
    const volatile double thv = angle;
    const double th = thv;
    const double theta2 = 0.0 * 0.0 + 0.0 * 0.0 + th * th;
    if (isZero(theta2)) return q;
    const double theta = std::sqrt(theta2);
    const double invTheta = 1.0 / theta;
    const double ux = 0.0 * invTheta, uy = 0.0 * invTheta, uz = th * invTheta;
    const double half = 0.5 * theta;
    const double c = std::cos(half), s = std::sin(half);
    const double r0 = c, r1 = s * ux, r2 = s * uy, r3 = s * uz;
    Q4 r;
    r.q0 = r0 * q.q0 - r1 * q.q1 - r2 * q.q2 - r3 * q.q3;
    r.q1 = r0 * q.q1 + r1 * q.q0 + r2 * q.q3 - r3 * q.q2;
    r.q2 = r0 * q.q2 - r1 * q.q3 + r2 * q.q0 + r3 * q.q1;
    r.q3 = r0 * q.q3 + r1 * q.q2 - r2 * q.q1 + r3 * q.q0;
    return qnormalize(r);
  }
  // updateVelocityPositionOrientationOutOfSectorKernel (mySpatialGrid.cu:59-85):
  // a particle past the x = min face (with y >= min) turns by -pi/2 about the
  // z axis through min, one past the y = min face (with x >= min) by +pi/2.
  void sectorWrap() {
    const V3 mn = pgrid.mn;
    const double hp = 0.5 * kPi;
    for (int i = 0; i < P.num(); ++i) {
      V3 v = P.v[i], p = P.x[i];
      Q4 q = P.q[i];
      if (p.x < mn.x && p.y >= mn.y) {
        v = rotZ(v, -hp);
        p = add(mn, rotZ(sub(p, mn), -hp));
        q = rotQz(q, -hp);
      }
      if (p.y < mn.y && p.x >= mn.x) {
        v = rotZ(v, hp);
        p = add(mn, rotZ(sub(p, mn), hp));
        q = rotQz(q, hp);
      }
      P.v[i] = r32(v, f32);
      P.x[i] = p;
      P.q[i] = r32(q, f32);
    }
  }
  // calculateSectorGhostVelocityPositionOrientationKernel (mySpatialGrid.cu:154-212):
  // R90 images of the particles in the first y cell row, R270 of the first
  // x column, R180 of their corner column; hashed by the ghost cell rule.
  void sectorGhosts() {
    const V3 mn = pgrid.mn;
    const double ang[3] = {0.5 * kPi, kPi, -0.5 * kPi};
    for (int s = 0; s < 3; ++s) {
      GhostSet& G = gh[s];
      G.pos = P.x;
      G.vel = P.v;
      G.q = P.q;
      for (int i = 0; i < P.num(); ++i) {
        const I3 c = gridPos(P.x[i], pgrid.mn, pgrid.inv);
        const bool on = s == 0 ? c.y == 0 : s == 2 ? c.x == 0 : (c.x == 0 && c.y == 0);
        if (!on) continue;
        G.vel[i] = rotZ(P.v[i], ang[s]);
        G.pos[i] = add(mn, rotZ(sub(P.x[i], mn), ang[s]));
        G.q[i] = rotQz(P.q[i], ang[s]);
      }
      G.ht.hash.resize(P.num());
      for (int i = 0; i < P.num(); ++i) G.ht.hash[i] = ghostHashOf(G.pos[i], pgrid);
      sortStartEnd(G.ht.hash, G.ht.index, pgrid.num, G.ht.start, G.ht.end);
    }
  }
  void updateGhosts() {
    if (sector) {  // PeriodicBoundarySector::updateGhostSpatialGrid, buildGhostInteraction
      sectorWrap();
      sectorGhosts();
      const int sets[3] = {SET_GX, SET_GY, SET_GXY};
      for (int s = 0; s < 3; ++s) detectSet(sets[s], pgrid, gh[s].ht, gh[s].pos, gh[s].q, P.geo, pgeo, false);
      return;
    }
    if (perX) { wrapAxis(0); ghostPositions(0); }
    if (perY) { wrapAxis(1); ghostPositions(1); }
    if (perX && perY) ghostPositions(2);
    const int sets[3] = {SET_GX, SET_GY, SET_GXY};
    const bool on[3] = {perX, perY, perX && perY};
    for (int s = 0; s < 3; ++s)
      if (on[s]) detectSet(sets[s], pgrid, gh[s].ht, gh[s].pos, gh[s].q, P.geo, pgeo, false);
  }
  // -------------------------------------------------------- forces ----
  // Per-contact force of each set: contactKernel.cu:79-119 (PP), :167-200
  // (wall), :243-283 (ghost). Updates the spring; stores the contact force.
  void contactForces(int set) {
    for (Contact& c : C[set]) {
      const int i = c.A, j = c.B;
      const V3 ri = P.x[i], vi = P.v[i], wi = P.w[i];
      const double invMi = pgeo.invMass[P.geo[i]];
      const int mi = P.mat[i];
      V3 rj, vj, wj;
      double kn = 0., ks = 0., effM = 0., diss, mu;
      if (set == SET_PW) {
        rj = W.x[j]; vj = W.v[j]; wj = W.w[j];
        const int mj = W.mat[j];
        if (invMi > 0.) effM = 1. / invMi;
        mu = std::fmin(pmat.mu[mi], wmat.mu[mj]);
        diss = pmat.e[mi] <= wmat.e[mj] ? pmat.diss[mi] : wmat.diss[mj];
        kn = pmat.kn[mi];
        ks = pmat.ks[mi];
      } else {
        if (set == SET_PP) { rj = P.x[j]; vj = P.v[j]; }
        else { const GhostSet& G = gh[set - SET_GX]; rj = G.pos[j]; vj = G.vel[j]; }
        wj = P.w[j];
        const double invMj = pgeo.invMass[P.geo[j]];
        const int mj = P.mat[j];
        const double kni = pmat.kn[mi], knj = pmat.kn[mj], ksi = pmat.ks[mi], ksj = pmat.ks[mj];
        if (kni > 0. && knj > 0.) kn = kni * knj / (kni + knj);
        if (ksi > 0. && ksj > 0.) ks = ksi * ksj / (ksi + ksj);
        if (invMi > 0. || invMj > 0.) effM = 1. / (invMi + invMj);
        diss = pmat.e[mi] <= pmat.e[mj] ? pmat.diss[mi] : pmat.diss[mj];
        mu = std::fmin(pmat.mu[mi], pmat.mu[mj]);
      }
      const V3 vc = sub(add(vi, cross(wi, sub(c.point, ri))), add(vj, cross(wj, sub(c.point, rj))));
      V3 F = {0., 0., 0.};
      V3 s = c.spring;
      linearContact(F, s, vc, c.normal, c.overlap, dt, kn, ks, mu, diss, effM);
      c.spring = r32(s, f32);
      c.F = F;
    }
  }
  V3 contactRi(const Contact& c) const { return P.x[c.A]; }
  V3 contactRj(int set, const Contact& c) const {
    if (set == SET_PP) return P.x[c.B];
    if (set == SET_PW) return W.x[c.B];
    return gh[set - SET_GX].pos[c.B];
  }
  // Bond forces: contactKernel.cu:330-371. Returns per-bond results.
  struct BondOut { V3 F, Ti, Tj; bool on; };
  std::vector<BondOut> bondForces() {
    std::vector<BondOut> out(bonds.num());
    for (int k = 0; k < bonds.num(); ++k) {
      BondOut& o = out[k];
      o.on = false;
      if (bonds.active[k] == 0) { bonds.sigma[k] = 0.; bonds.tau[k] = 0.; continue; }
      const int i = bonds.master[k], j = bonds.slave[k];
      const Q4 qi = P.q[i], qj = P.q[j];
      const V3 n1i = rotate(qi, bonds.n1i[k]), n2i = rotate(qi, bonds.n2i[k]), n3i = rotate(qi, bonds.n3i[k]);
      const V3 n1j = rotate(qj, bonds.n1j[k]), n2j = rotate(qj, bonds.n2j[k]), n3j = rotate(qj, bonds.n3j[k]);
      const V3 ri = P.x[i], rj = P.x[j];
      const V3 rbi = add(rotate(qi, bonds.pi_[k]), ri), rbj = add(rotate(qj, bonds.pj[k]), rj);
      V3 F = {0, 0, 0}, Mij = {0, 0, 0}, Mji = {0, 0, 0};
      double sig = bonds.sigma[k], tau = bonds.tau[k];
      bonds.active[k] = vbond(F, Mij, Mji, sig, tau, rbi, rbj, n1i, n2i, n3i, n1j, n2j, n3j,
                              bonds.B1[k], bonds.B2[k], bonds.B3[k], bonds.B4[k], bonds.R[k],
                              bonds.L0[k], bonds.st[k], bonds.coh[k], bonds.mub[k]);
      bonds.sigma[k] = sig;
      bonds.tau[k] = tau;
      o.on = true;
      o.F = F;
      o.Ti = add(Mij, cross(sub(rbi, ri), F));
      o.Tj = add(Mji, cross(sub(rbj, rj), neg(F)));
    }
    return out;
  }
  // Serial execution of the reference's atomics in contact-list order.
  void foldReference(const std::vector<BondOut>& bo) {
    auto addTo = [&](int idx, V3 f, V3 t) { P.F[idx] = add(P.F[idx], f); P.T[idx] = add(P.T[idx], t); };
    for (int set = 0; set < NSETS; ++set)
      for (const Contact& c : C[set]) {
        const V3 ri = contactRi(c), rj = contactRj(set, c);
        addTo(c.A, c.F, cross(sub(c.point, ri), c.F));
        if (set == SET_PW) continue;
        if (set >= SET_GX && c.A == c.B) continue;
        addTo(c.B, neg(c.F), cross(sub(c.point, rj), neg(c.F)));
      }
    for (int k = 0; k < bonds.num(); ++k) {
      if (!bo[k].on) continue;
      addTo(bonds.master[k], bo[k].F, bo[k].Ti);
      addTo(bonds.slave[k], neg(bo[k].F), bo[k].Tj);
    }
  }
  // Canonical fold (see the header comment of this file).
  struct PairSum { int A, B; V3 SF, STA, STB; };
  std::vector<PairSum> pairSums(int set) const {
    std::vector<int> ord(C[set].size());
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](int a, int b) {
      const Contact &x = C[set][a], &y = C[set][b];
      if (x.A != y.A) return x.A < y.A;
      if (x.B != y.B) return x.B < y.B;
      return x.k < y.k;
    });
    std::vector<PairSum> out;
    for (int idx : ord) {
      const Contact& c = C[set][idx];
      if (out.empty() || out.back().A != c.A || out.back().B != c.B)
        out.push_back({c.A, c.B, {0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}});
      PairSum& s = out.back();
      s.SF = add(s.SF, c.F);
      s.STA = add(s.STA, cross(sub(c.point, contactRi(c)), c.F));
      s.STB = add(s.STB, cross(sub(c.point, contactRj(set, c)), neg(c.F)));
    }
    return out;
  }
  void foldCanonical(const std::vector<BondOut>& bo) {
    struct Rec { int partner, role; V3 f, t; };
    const int N = P.num();
    std::vector<std::vector<Rec>> recs[NSETS];
    for (int set = 0; set < NSETS; ++set) {
      recs[set].assign(N, {});
      for (const PairSum& s : pairSums(set)) {
        recs[set][s.A].push_back({s.B, 0, s.SF, s.STA});
        if (set == SET_PW) continue;
        if (set >= SET_GX && s.A == s.B) continue;
        recs[set][s.B].push_back({s.A, 1, neg(s.SF), s.STB});
      }
      for (auto& r : recs[set])
        std::stable_sort(r.begin(), r.end(), [](const Rec& a, const Rec& b) {
          return a.partner != b.partner ? a.partner < b.partner : a.role < b.role;
        });
    }
    std::vector<std::vector<Rec>> brec(N);
    for (int k = 0; k < bonds.num(); ++k) {
      if (!bo[k].on) continue;
      brec[bonds.master[k]].push_back({k, 0, bo[k].F, bo[k].Ti});
      brec[bonds.slave[k]].push_back({k, 1, neg(bo[k].F), bo[k].Tj});
    }
    for (int i = 0; i < N; ++i) {
      V3 F = {0., 0., 0.}, T = {0., 0., 0.};
      for (int set = 0; set < NSETS; ++set)
        for (const Rec& r : recs[set][i]) { F = add(F, r.f); T = add(T, r.t); }
      for (const Rec& r : brec[i]) { F = add(F, r.f); T = add(T, r.t); }
      P.F[i] = F;
      P.T[i] = T;
    }
  }
  void forcesAndFold() {
    for (int i = 0; i < P.num(); ++i) { P.F[i] = {0., 0., 0.}; P.T[i] = {0., 0., 0.}; }  // LSSolver.h:919-923
    for (int set = 0; set < NSETS; ++set) contactForces(set);
    std::vector<BondOut> bo = bondForces();
    if (canon) foldCanonical(bo); else foldReference(bo);
    for (int i = 0; i < P.num(); ++i) { P.F[i] = r32(P.F[i], f32); P.T[i] = r32(P.T[i], f32); }
  }
  bool periodic() const { return perX || perY || sector; }
  // compute_addBondAndPeriodicBoundary loop body (LSSolver.h:1206-1220) and
  // its simpler variants, which are the same with empty ghost/bond sets.
  // The tutorials' addExternalForceTorque overrides (apps/tutorial/
  // globalDamping.cu:4-49), called between the forces and the second
  // half-kick (LSSolver.h:1176): global damping F -= c |F| v/|v| (myVec.h:120-135
  // normalize, zero for a zero vector), then the constant-force kernel, which
  // adds both of its vectors to the particle's force.
  void externalForces() {
    if (damping >= 0.0)
      for (int i = 0; i < P.num(); ++i) {
        const V3 f = P.F[i], t = P.T[i];
        P.F[i] = r32(sub(f, scl(normalize(P.v[i]), damping * len(f))), f32);
        P.T[i] = r32(sub(t, scl(normalize(P.w[i]), damping * len(t))), f32);
      }
    if (constIndex >= 0 && constIndex < P.num()) {
      P.F[constIndex] = r32(add(P.F[constIndex], constF), f32);
      P.F[constIndex] = r32(add(P.F[constIndex], constT), f32);
    }
  }
  void step() {
    const double h = 0.5 * dt;
    kick(h);
    drift(P, dt);
    drift(W, dt);
    buildGrid(P, pgrid, pht);
    buildGrid(W, wgrid, wht);
    detectAll(periodic());
    forcesAndFold();
    externalForces();
    kick(h);
  }
};

// --------------------------------------------------------------- C API ----
extern "C" {

void oracle_grid_set(double* min3, double* max3, double cs, int* size3, double* inv3) {
  Grid g;
  gridSet(g, ld3(min3), ld3(max3), cs);
  st3(min3, g.mn);
  st3(max3, g.mx);
  size3[0] = g.size.x; size3[1] = g.size.y; size3[2] = g.size.z;
  st3(inv3, g.inv);
}
void oracle_hash(int n, const double* pos, const double* min3, const double* max3, const double* inv3,
                 const int* size3, int* out) {
  Grid g;
  g.mn = ld3(min3); g.mx = ld3(max3); g.inv = ld3(inv3); g.size = {size3[0], size3[1], size3[2]};
  for (int i = 0; i < n; ++i) out[i] = hashOf(ld3(pos + 3 * i), g);
}
void oracle_ghost_hash(int n, const double* pos, const double* min3, const double* max3,
                       const double* inv3, const int* size3, int* out) {
  Grid g;
  g.mn = ld3(min3); g.mx = ld3(max3); g.inv = ld3(inv3); g.size = {size3[0], size3[1], size3[2]};
  for (int i = 0; i < n; ++i) out[i] = ghostHashOf(ld3(pos + 3 * i), g);
}
void oracle_sort_start_end(int n, int* hash, int* index, long long numGrid, int* start, int* end) {
  std::vector<int> h(hash, hash + n), idx, st, en;
  sortStartEnd(h, idx, numGrid, st, en);
  std::memcpy(hash, h.data(), sizeof(int) * n);
  std::memcpy(index, idx.data(), sizeof(int) * n);
  std::memcpy(start, st.data(), sizeof(int) * numGrid);
  std::memcpy(end, en.data(), sizeof(int) * numGrid);
}
void oracle_inclusive_scan(long long n, const int* count, int* prefix) {
  long long acc = 0;
  for (long long i = 0; i < n; ++i) { acc += count[i]; prefix[i] = (int)acc; }
}

OracleSim* osim_create(int store_fp32, int canonical_fold) {
  OracleSim* s = new OracleSim();
  s->f32 = store_fp32 != 0;
  s->canon = canonical_fold != 0;
  return s;
}
void osim_destroy(OracleSim* s) { delete s; }

int osim_add_material(OracleSim* s, double kn, double ks, double mu, double e) {
  s->pmat.kn.push_back(kn > 0. ? kn : 0.0);
  s->pmat.ks.push_back(ks > 0. ? ks : 0.0);
  s->pmat.mu.push_back(mu > 0. ? mu : 0.0);
  s->pmat.e.push_back(e > 0. && e < 1.0 ? e : 1.0);
  s->pmat.diss.push_back(dissipationOf(s->pmat.e.back()));
  return (int)s->pmat.kn.size() - 1;
}
// Additive: a material given by its damping ratio (e orders materials only).
static double restitutionOfDamping(double zeta) {
  if (zeta <= 0.) return 1.0;
  if (zeta >= 1.) return 0.0;
  return std::exp(-kPi * zeta / std::sqrt(1.0 - zeta * zeta));
}
int osim_add_material_damping(OracleSim* s, double kn, double ks, double mu, double zeta) {
  s->pmat.kn.push_back(kn > 0. ? kn : 0.0);
  s->pmat.ks.push_back(ks > 0. ? ks : 0.0);
  s->pmat.mu.push_back(mu > 0. ? mu : 0.0);
  s->pmat.diss.push_back(zeta > 0. ? zeta : 0.0);
  s->pmat.e.push_back(restitutionOfDamping(s->pmat.diss.back()));
  return (int)s->pmat.kn.size() - 1;
}
int osim_add_wall_material_damping(OracleSim* s, double mu, double zeta) {
  s->wmat.kn.push_back(0.0);
  s->wmat.ks.push_back(0.0);
  s->wmat.mu.push_back(mu > 0. ? mu : 0.0);
  s->wmat.diss.push_back(zeta > 0. ? zeta : 0.0);
  s->wmat.e.push_back(restitutionOfDamping(s->wmat.diss.back()));
  return (int)s->wmat.kn.size() - 1;
}
int osim_add_wall_material(OracleSim* s, double mu, double e) {
  s->wmat.kn.push_back(0.0);
  s->wmat.ks.push_back(0.0);
  s->wmat.mu.push_back(mu > 0. ? mu : 0.0);
  s->wmat.e.push_back(e > 0. && e < 1.0 ? e : 1.0);
  s->wmat.diss.push_back(dissipationOf(s->wmat.e.back()));
  return (int)s->wmat.kn.size() - 1;
}
// geometryInfo::pushHost (LSParticle.h:74-117) after LSParticle::addGeometry's checks.
static int pushGeometry(OracleSim* s, Geoms& G, double mass, Sym I, V3 origin, double spacing,
                        I3 size, const double* sdf, const double* nodes, int nnodes) {
  if (size.x < 2 || size.y < 2 || size.z < 2) return -1;
  if (spacing <= 0.0) return -1;
  const bool f = s->f32;
  const double invM = mass > 0.0 ? 1.0 / mass : 0.0;
  const Sym Iu = mass > 0.0 ? I : Sym{0, 0, 0, 0, 0, 0};
  G.invMass.push_back(invM);
  G.I.push_back(Iu);
  G.Iinv.push_back(inverse(Iu));
  G.origin.push_back(origin);
  G.invh.push_back(1.0 / spacing);
  G.size.push_back(size);
  G.sdfStart.push_back((long long)G.sdf.size());
  const long long nsdf = (long long)size.x * size.y * size.z;
  for (long long i = 0; i < nsdf; ++i) G.sdf.push_back(r32(sdf[i], f));
  G.nodeStart.push_back((int)G.nodes.size());
  G.nodeCount.push_back(nnodes);
  double r = 0.;
  for (int k = 0; k < nnodes; ++k) {
    const V3 p = r32(ld3(nodes + 3 * k), f);
    G.nodes.push_back(p);
    r = std::max(r, len(p));
  }
  G.radius.push_back(r);
  return G.num() - 1;
}
int osim_add_geometry(OracleSim* s, double mass, const double* I6, const double* o3, double h,
                      const int* sz, const double* sdf, const double* nodes, int nn) {
  return pushGeometry(s, s->pgeo, mass, {I6[0], I6[1], I6[2], I6[3], I6[4], I6[5]}, ld3(o3), h,
                      {sz[0], sz[1], sz[2]}, sdf, nodes, nn);
}
int osim_add_geometry_density(OracleSim* s, double density, const double* o3, double h,
                              const int* sz, const double* sdf, const double* nodes, int nn,
                              double* off3) {
  // LSSolver.h:470-574
  st3(off3, {0., 0., 0.});
  const I3 size = {sz[0], sz[1], sz[2]};
  if (size.x < 2 || size.y < 2 || size.z < 2 || h <= 0.0) return -1;
  auto H = [](double phi, double eps) -> double {
    if (eps <= 0.0) return (phi > 0.0) ? 0.0 : 1.0;
    if (phi < -eps) return 1.0;
    if (phi > eps) return 0.0;
    const double x = -phi / eps;
    return 0.5 * (1.0 + x + std::sin(kPi * x) / kPi);
  };
  const V3 o = ld3(o3);
  double mass = 0.;
  const double mg = density * h * h * h;
  for (int x = 0; x < size.x; x++)
    for (int y = 0; y < size.y; y++)
      for (int z = 0; z < size.z; z++) mass += H(sdf[lin3({x, y, z}, size)] / h, 1.5);
  mass *= mg;
  V3 c = {0., 0., 0.};
  Sym I = {0, 0, 0, 0, 0, 0};
  if (mass > 0.) {
    for (int x = 0; x < size.x; x++)
      for (int y = 0; y < size.y; y++)
        for (int z = 0; z < size.z; z++) {
          const double Hv = H(sdf[lin3({x, y, z}, size)] / h, 1.5);
          c.x += Hv * (o.x + double(x) * h);
          c.y += Hv * (o.y + double(y) * h);
          c.z += Hv * (o.z + double(z) * h);
        }
    c = scl(c, mg / mass);
    for (int x = 0; x < size.x; x++)
      for (int y = 0; y < size.y; y++)
        for (int z = 0; z < size.z; z++) {
          const double Hv = H(sdf[lin3({x, y, z}, size)] / h, 1.5);
          const V3 r = sub(add(o, scl(V3{double(x), double(y), double(z)}, h)), c);
          I.xx += Hv * (r.y * r.y + r.z * r.z) * mg;
          I.yy += Hv * (r.x * r.x + r.z * r.z) * mg;
          I.zz += Hv * (r.y * r.y + r.x * r.x) * mg;
          I.xy -= Hv * r.x * r.y * mg;
          I.xz -= Hv * r.x * r.z * mg;
          I.yz -= Hv * r.y * r.z * mg;
        }
  }
  std::vector<double> shifted(3 * (size_t)nn);
  for (int k = 0; k < nn; ++k) st3(&shifted[3 * k], sub(ld3(nodes + 3 * k), c));
  const int g = pushGeometry(s, s->pgeo, mass, I, sub(o, c), h, size, sdf, shifted.data(), nn);
  st3(off3, c);
  return g;
}
int osim_add_wall_geometry(OracleSim* s, const double* o3, double h, const int* sz, const double* sdf,
                           const double* nodes, int nn) {
  return pushGeometry(s, s->wgeo, 0.0, {0, 0, 0, 0, 0, 0}, ld3(o3), h, {sz[0], sz[1], sz[2]}, sdf,
                      nodes, nn);
}
static int pushBody(OracleSim* s, Body& B, const Geoms& G, size_t nmat, const double* x3,
                    const double* v3, const double* w3, const double* q4, int mat, int geo) {
  if (mat < 0 || (size_t)mat >= nmat) return -1;
  if (geo < 0 || geo >= G.num()) return -1;
  const bool f = s->f32;
  B.x.push_back(ld3(x3));
  B.v.push_back(r32(ld3(v3), f));
  B.w.push_back(r32(ld3(w3), f));
  B.q.push_back(r32(qnormalize({q4[0], q4[1], q4[2], q4[3]}), f));
  B.F.push_back({0., 0., 0.});
  B.T.push_back({0., 0., 0.});
  B.mat.push_back(mat);
  B.geo.push_back(geo);
  return B.num() - 1;
}
int osim_add_particle(OracleSim* s, const double* x3, const double* v3, const double* w3,
                      const double* q4, int mat, int geo) {
  return pushBody(s, s->P, s->pgeo, s->pmat.kn.size(), x3, v3, w3, q4, mat, geo);
}
int osim_add_wall(OracleSim* s, const double* x3, const double* v3, const double* w3, const double* q4,
                  int mat, int geo) {
  return pushBody(s, s->W, s->wgeo, s->wmat.kn.size(), x3, v3, w3, q4, mat, geo);
}
int osim_add_bond(OracleSim* s, int i, int j, const double* pt3, const double* nrm3, double Rb,
                  double Lb, double Eb, double nu, double st, double coh, double mub) {
  // LSSolver.h:306-337 then VBondedInteraction.h:74-195
  if (i >= s->P.num() || j >= s->P.num()) return -1;
  const V3 gn = ld3(nrm3);
  if (isZero(len2(gn))) return -1;
  if (Rb <= 0. || Lb <= 0. || Eb <= 0. || nu <= -1.0 || nu >= 0.5) return -1;
  const double Gb = Eb / (2. * (1. + nu));
  const double Ab = kPi * Rb * Rb;
  const double Ib = Ab * Rb * Rb / 4.;
  const double Jb = 2. * Ib;
  const double Kn = Eb * Ab / Lb, Kt = Gb * Jb / Lb, Ks = 12. * Eb * Ib / (Lb * Lb * Lb), Kb = Eb * Ib / Lb;
  const double B1 = Kn, B2 = Lb * Lb * Ks, B4 = Kt, B3 = Kb - B2 / 4. - B4 / 2.;
  const V3 n = dvs(gn, len(gn));
  const V3 c = ld3(pt3);
  const V3 gpi = add(c, scl(n, 0.5 * Lb)), gpj = sub(c, scl(n, 0.5 * Lb));
  const Q4 qi = s->P.q[i], qj = s->P.q[j];
  const V3 lpi = rotateT(sub(gpi, s->P.x[i]), qi), lpj = rotateT(sub(gpj, s->P.x[j]), qj);
  const V3 N1i = neg(n), N2i = bondN2(N1i), N3i = normalize(cross(N1i, N2i));
  const V3 N1j = neg(N1i), N2j = N2i, N3j = N3i;
  const bool f = s->f32;
  Bonds& b = s->bonds;
  b.master.push_back(i);
  b.slave.push_back(j);
  b.active.push_back(1);
  b.B1.push_back(r32(B1, f)); b.B2.push_back(r32(B2, f)); b.B3.push_back(r32(B3, f)); b.B4.push_back(r32(B4, f));
  b.R.push_back(r32(Rb, f)); b.L0.push_back(r32(Lb, f));
  b.st.push_back(r32(st, f)); b.coh.push_back(r32(coh, f)); b.mub.push_back(r32(mub, f));
  b.n1i.push_back(r32(normalize(rotateT(N1i, qi)), f));
  b.n2i.push_back(r32(normalize(rotateT(N2i, qi)), f));
  b.n3i.push_back(r32(normalize(rotateT(N3i, qi)), f));
  b.pi_.push_back(r32(lpi, f));
  b.n1j.push_back(r32(normalize(rotateT(N1j, qj)), f));
  b.n2j.push_back(r32(normalize(rotateT(N2j, qj)), f));
  b.n3j.push_back(r32(normalize(rotateT(N3j, qj)), f));
  b.pj.push_back(r32(lpj, f));
  b.sigma.push_back(0.);
  b.tau.push_back(0.);
  return b.num() - 1;
}
void osim_set_periodic(OracleSim* s, int x, int y) { s->perX = x != 0; s->perY = y != 0; }
void osim_set_sector(OracleSim* s, int on) { s->sector = on != 0; }
void osim_set_global_damping(OracleSim* s, double c) { s->damping = c; }
void osim_set_constant_force(OracleSim* s, int index, const double* F3, const double* T3) {
  s->constIndex = index;
  s->constF = ld3(F3);
  s->constT = ld3(T3);
}
void osim_set_cell_size(OracleSim* s, double cs) { s->cellSizeOverride = cs; }

int osim_initialize(OracleSim* s, const double* mn3, const double* mx3, const double* g3, double dt) {
  const V3 mn = ld3(mn3), mx = ld3(mx3), L = sub(mx, mn);
  if (L.x <= 0.0 || L.y <= 0.0 || L.z <= 0.0 || dt <= 0.0) return -1;  // LSSolver.h:646-676
  s->gravity = ld3(g3);
  s->dt = dt;
  double cs = 0.0;  // LSParticle::initialize (LSParticle.h:371-386)
  if (s->P.num() > 0 && s->pgeo.num() > 0)
    cs = *std::max_element(s->pgeo.radius.begin(), s->pgeo.radius.end()) * 2.0;
  if (s->cellSizeOverride > 0.0) cs = s->cellSizeOverride;
  gridSet(s->pgrid, mn, mx, cs);
  double wcs = 0.0;
  if (s->W.num() > 0 && s->wgeo.num() > 0)
    wcs = *std::max_element(s->wgeo.radius.begin(), s->wgeo.radius.end()) * 2.0;
  gridSet(s->wgrid, mn, mx, wcs);
  s->nodeBase.assign(s->P.num() + 1, 0);
  for (int i = 0; i < s->P.num(); ++i)
    s->nodeBase[i + 1] = s->nodeBase[i] + s->pgeo.nodeCount[s->P.geo[i]];
  // pre-loop build (LSSolver.h:1097-1100 and variants)
  s->buildGrid(s->P, s->pgrid, s->pht);
  s->buildGrid(s->W, s->wgrid, s->wht);
  s->detectAll(s->periodic());
  s->initialized = true;
  return 0;
}
int osim_step(OracleSim* s, int nsteps) {
  if (!s->initialized) return -1;
  for (int n = 0; n < nsteps; ++n) s->step();
  return 0;
}
long long osim_bruteforce_pp(OracleSim* s, int* a, int* b, int* k, long long cap) {
  // Every (node of A, B > A) with B inside the domain, without the grid.
  long long cnt = 0;
  const OracleSim& S = *s;
  std::vector<int> h(S.P.num());
  for (int i = 0; i < S.P.num(); ++i) h[i] = hashOf(S.P.x[i], S.pgrid);
  for (int A = 0; A < S.P.num(); ++A) {
    const int ga = S.P.geo[A];
    for (int B = A + 1; B < S.P.num(); ++B) {
      if (h[B] < 0 || h[B] >= S.pgrid.num) continue;
      const double rB = S.pgeo.radius[S.P.geo[B]];
      for (int kk = 0; kk < S.pgeo.nodeCount[ga]; ++kk) {
        const V3 p = add(rotate(S.P.q[A], S.pgeo.nodes[S.pgeo.nodeStart[ga] + kk]), S.P.x[A]);
        const V3 rel = sub(p, S.P.x[B]);
        if (len2(rel) > rB * rB) continue;
        double ov;
        V3 n;
        if (!sampleContact(rotateT(rel, S.P.q[B]), S.P.q[B], S.P.geo[B], S.pgeo, ov, n)) continue;
        if (cnt < cap) { a[cnt] = A; b[cnt] = B; k[cnt] = kk; }
        ++cnt;
      }
    }
  }
  return cnt;
}
int osim_num_particles(const OracleSim* s) { return s->P.num(); }
int osim_num_walls(const OracleSim* s) { return s->W.num(); }
int osim_num_bonds(const OracleSim* s) { return s->bonds.num(); }
void osim_get_state(const OracleSim* s, double* x, double* v, double* w, double* q, double* f, double* t) {
  for (int i = 0; i < s->P.num(); ++i) {
    if (x) st3(x + 3 * i, s->P.x[i]);
    if (v) st3(v + 3 * i, s->P.v[i]);
    if (w) st3(w + 3 * i, s->P.w[i]);
    if (q) { q[4 * i] = s->P.q[i].q0; q[4 * i + 1] = s->P.q[i].q1; q[4 * i + 2] = s->P.q[i].q2; q[4 * i + 3] = s->P.q[i].q3; }
    if (f) st3(f + 3 * i, s->P.F[i]);
    if (t) st3(t + 3 * i, s->P.T[i]);
  }
}
void osim_set_state(OracleSim* s, const double* x, const double* v, const double* w, const double* q,
                    const double* f, const double* t) {
  for (int i = 0; i < s->P.num(); ++i) {
    if (x) s->P.x[i] = ld3(x + 3 * i);
    if (v) s->P.v[i] = ld3(v + 3 * i);
    if (w) s->P.w[i] = ld3(w + 3 * i);
    if (q) s->P.q[i] = {q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
    if (f) s->P.F[i] = ld3(f + 3 * i);
    if (t) s->P.T[i] = ld3(t + 3 * i);
  }
}
void osim_get_walls(const OracleSim* s, double* x, double* q) {
  for (int i = 0; i < s->W.num(); ++i) {
    if (x) st3(x + 3 * i, s->W.x[i]);
    if (q) { q[4 * i] = s->W.q[i].q0; q[4 * i + 1] = s->W.q[i].q1; q[4 * i + 2] = s->W.q[i].q2; q[4 * i + 3] = s->W.q[i].q3; }
  }
}
void osim_get_geometry(const OracleSim* s, int g, double* r, double* invM, double* I6, double* Ii6,
                       double* o3, double* ih) {
  const Geoms& G = s->pgeo;
  if (g < 0 || g >= G.num()) return;
  if (r) *r = G.radius[g];
  if (invM) *invM = G.invMass[g];
  if (I6) { const Sym m = G.I[g]; I6[0] = m.xx; I6[1] = m.yy; I6[2] = m.zz; I6[3] = m.xy; I6[4] = m.xz; I6[5] = m.yz; }
  if (Ii6) { const Sym m = G.Iinv[g]; Ii6[0] = m.xx; Ii6[1] = m.yy; Ii6[2] = m.zz; Ii6[3] = m.xy; Ii6[4] = m.xz; Ii6[5] = m.yz; }
  if (o3) st3(o3, G.origin[g]);
  if (ih) *ih = G.invh[g];
}
void osim_get_grid(const OracleSim* s, int walls, double* mn, double* mx, double* inv, int* size) {
  const Grid& g = walls ? s->wgrid : s->pgrid;
  if (mn) st3(mn, g.mn);
  if (mx) st3(mx, g.mx);
  if (inv) st3(inv, g.inv);
  if (size) { size[0] = g.size.x; size[1] = g.size.y; size[2] = g.size.z; }
}
long long osim_get_contacts(const OracleSim* s, int set, int* a, int* tg, int* k, double* spring,
                            double* point, double* normal, double* overlap, long long cap) {
  if (set < 0 || set >= NSETS) return -1;
  const std::vector<Contact>& C = s->C[set];
  std::vector<int> ord(C.size());
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int x, int y) {
    if (C[x].A != C[y].A) return C[x].A < C[y].A;
    if (C[x].B != C[y].B) return C[x].B < C[y].B;
    return C[x].k < C[y].k;
  });
  const long long n = (long long)C.size();
  for (long long i = 0; i < n && i < cap; ++i) {
    const Contact& c = C[ord[i]];
    if (a) a[i] = c.A;
    if (tg) tg[i] = c.B;
    if (k) k[i] = c.k;
    if (spring) st3(spring + 3 * i, c.spring);
    if (point) st3(point + 3 * i, c.point);
    if (normal) st3(normal + 3 * i, c.normal);
    if (overlap) overlap[i] = c.overlap;
  }
  return n;
}
void osim_get_hash(const OracleSim* s, int walls, int* hash, int* index, int* start, int* end) {
  const HashTable& h = walls ? s->wht : s->pht;
  if (hash) std::memcpy(hash, h.hash.data(), sizeof(int) * h.hash.size());
  if (index) std::memcpy(index, h.index.data(), sizeof(int) * h.index.size());
  if (start) std::memcpy(start, h.start.data(), sizeof(int) * h.start.size());
  if (end) std::memcpy(end, h.end.data(), sizeof(int) * h.end.size());
}
void osim_get_bonds(const OracleSim* s, int* active, double* sig, double* tau) {
  for (int k = 0; k < s->bonds.num(); ++k) {
    if (active) active[k] = s->bonds.active[k];
    if (sig) sig[k] = s->bonds.sigma[k];
    if (tau) tau[k] = s->bonds.tau[k];
  }
}
int osim_set_spring(OracleSim* s, int set, int a, int tg, int k, const double* sp) {
  if (set < 0 || set >= NSETS) return 0;
  for (Contact& c : s->C[set])
    if (c.A == a && c.B == tg && c.k == k) { c.spring = r32(ld3(sp), s->f32); return 1; }
  return 0;
}

int osim_get_ghosts(const OracleSim* s, int set, double* pos, int* hash) {
  if (set < 0 || set >= 3) return 0;
  const GhostSet& G = s->gh[set];
  const int n = (int)G.pos.size();
  for (int i = 0; i < n; ++i) {
    if (pos) st3(pos + 3 * i, G.pos[i]);
    if (hash) hash[i] = ghostHashOf(G.pos[i], s->pgrid);
  }
  return n;
}

}  // extern "C"

// The level-set sampling of the detection kernels on one grid, for the tests:
// trilinear value and Q1 gradient (per cell, before rotation and
// normalisation) at local-frame points; inside[i] = 0 and zeros where the
// point's cell is not inside the grid.
extern "C" void oracle_sample_level_set(const double* sdf, const int* size3, const double* origin3,
                                        double spacing, const double* pts, int np, double* value,
                                        double* grad, int* inside) {
  const I3 sz{size3[0], size3[1], size3[2]};
  const V3 o{origin3[0], origin3[1], origin3[2]};
  const double ih = 1.0 / spacing;
  for (int i = 0; i < np; ++i) {
    CellSample cs;
    const bool in = cellSample({pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}, o, ih, sz, sdf, cs);
    const V3 g = in ? gradientQ1(cs.x, cs.y, cs.z, cs.c) : V3{0., 0., 0.};
    value[i] = in ? trilinear(cs.x, cs.y, cs.z, cs.c) : 0.;
    grad[3 * i] = g.x;
    grad[3 * i + 1] = g.y;
    grad[3 * i + 2] = g.z;
    inside[i] = in ? 1 : 0;
  }
}
