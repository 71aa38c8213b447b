// Device-side FP64 arithmetic for the level-set DEM step.
//
// Every routine evaluates its expression in the reference's operation order
// (kernel/CUDAKernelFunction/myUtility/myVec.h, myQua.h, myMat.h;
// LSParticleContactDetectionKernel.cuh; contactKernel.cuh). The library is
// compiled with -ffp-contract=off, so no multiply-add is fused and each FP64
// operation rounds exactly once, as on the host.
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>

#include "fundem_amd.h"

namespace fdm {

struct V3 {
  double x, y, z;
};
struct Q4 {
  double q0, q1, q2, q3;
};
struct S6 {
  double xx, yy, zz, xy, xz, yz;
};
struct I3 {
  int x, y, z;
};

constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 vneg(V3 a) { return {-a.x, -a.y, -a.z}; }
// vector times scalar, each component times s (the reference's operator* in both orders)
__device__ __forceinline__ V3 vmul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 vdiv(V3 a, double s) {
  const double inv = 1.0 / s;
  return {a.x * inv, a.y * inv, a.z * inv};
}
__device__ __forceinline__ double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 vcross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double vlen2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__device__ __forceinline__ double vlen(V3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ bool nearZero(double x) { return fabs(x) <= 1.0e-20; }
// normalize with a correctly rounded 1/sqrt (not a reciprocal-sqrt estimate)
__device__ __forceinline__ V3 vnormalize(V3 v) {
  const double l2 = v.x * v.x + v.y * v.y + v.z * v.z;
  if (nearZero(l2)) return {0.0, 0.0, 0.0};
  const double inv = 1.0 / sqrt(l2);
  return {v.x * inv, v.y * inv, v.z * inv};
}
__device__ __forceinline__ Q4 qnormalize(Q4 q) {
  const double l2 = q.q0 * q.q0 + q.q1 * q.q1 + q.q2 * q.q2 + q.q3 * q.q3;
  if (nearZero(l2)) return q;
  const double inv = 1.0 / sqrt(l2);
  return {q.q0 * inv, q.q1 * inv, q.q2 * inv, q.q3 * inv};
}
// v' = v + q0 t + u x t with t = 2 u x v (myQua.h:104-114)
__device__ __forceinline__ V3 qrotate(Q4 q, V3 v) {
  const V3 u = {q.q1, q.q2, q.q3};
  const V3 t = vmul(vcross(u, v), 2.0);
  return vadd(vadd(v, vmul(t, q.q0)), vcross(u, t));
}
// rotation by the conjugate (myQua.h:120-129)
__device__ __forceinline__ V3 qrotateT(V3 v, Q4 q) {
  const V3 u = {-q.q1, -q.q2, -q.q3};
  const V3 t = vmul(vcross(u, v), 2.0);
  return vadd(vadd(v, vmul(t, q.q0)), vcross(u, t));
}
// explicit step of dq/dt = 1/2 (0, w) (x) q, then normalise (myQua.h:77-93)
__device__ __forceinline__ Q4 qintegrate(Q4 q, V3 w, double dt) {
  const V3 v = vmul(w, 0.5 * dt);
  Q4 d;
  d.q0 = -q.q1 * v.x - q.q2 * v.y - q.q3 * v.z;
  d.q1 = q.q0 * v.x + q.q3 * v.y - q.q2 * v.z;
  d.q2 = -q.q3 * v.x + q.q0 * v.y + q.q1 * v.z;
  d.q3 = q.q2 * v.x - q.q1 * v.y + q.q0 * v.z;
  return qnormalize({q.q0 + d.q0, q.q1 + d.q1, q.q2 + d.q2, q.q3 + d.q3});
}
__device__ __forceinline__ V3 smul(S6 m, V3 v) {
  return {m.xx * v.x + m.xy * v.y + m.xz * v.z, m.xy * v.x + m.yy * v.y + m.yz * v.z,
          m.xz * v.x + m.yz * v.y + m.zz * v.z};
}
// R I R^T with R from q, in the explicit expansion of myMat.h:107-158
__device__ __forceinline__ S6 srotate(Q4 q, S6 I) {
  const double q0 = q.q0, q1 = q.q1, q2 = q.q2, q3 = q.q3;
  const double q11 = 2 * q1 * q1, q22 = 2 * q2 * q2, q33 = 2 * q3 * q3;
  const double a = 1.0 - q22 - q33, b = 2 * (q1 * q2 - q0 * q3), c = 2 * (q1 * q3 + q0 * q2);
  const double d = 2 * (q1 * q2 + q0 * q3), e = 1.0 - q11 - q33, f = 2 * (q2 * q3 - q0 * q1);
  const double g = 2 * (q1 * q3 - q0 * q2), h = 2 * (q2 * q3 + q0 * q1), i = 1.0 - q11 - q22;
  const double r0x = I.xx * a + I.xy * b + I.xz * c, r0y = I.xy * a + I.yy * b + I.yz * c,
               r0z = I.xz * a + I.yz * b + I.zz * c;
  const double r1x = I.xx * d + I.xy * e + I.xz * f, r1y = I.xy * d + I.yy * e + I.yz * f,
               r1z = I.xz * d + I.yz * e + I.zz * f;
  const double r2x = I.xx * g + I.xy * h + I.xz * i, r2y = I.xy * g + I.yy * h + I.yz * i,
               r2z = I.xz * g + I.yz * h + I.zz * i;
  return {r0x * a + r0y * b + r0z * c, r1x * d + r1y * e + r1z * f, r2x * g + r2y * h + r2z * i,
          r0x * d + r0y * e + r0z * f, r0x * g + r0y * h + r0z * i, r1x * g + r1y * h + r1z * i};
}

// A template whose inverse inertia is exactly zero (the reference's inverse
// returns zero below |det I| = 1e-30, e.g. for millimetre particles): the
// angular half-kick then adds an exact zero to the spin, so it is skipped
// (the spin keeps its bits, up to the sign of a zero component).
__device__ __forceinline__ bool spinFixed(const fdm_template& t) {
  return t.invInertia[0] == 0.0 && t.invInertia[1] == 0.0 && t.invInertia[2] == 0.0 &&
         t.invInertia[3] == 0.0 && t.invInertia[4] == 0.0 && t.invInertia[5] == 0.0;
}

// A load through the constant address space, for data that no lane writes
// while the kernel runs: with a wave-uniform address the compiler issues it
// as a scalar load (into SGPRs), otherwise as an ordinary vector load.
// Structures are read word by word (the words no field in use needs are
// dropped by the compiler).
template <class T>
__device__ __forceinline__ T ldu(const T* p) {
  if constexpr (std::is_arithmetic<T>::value) {
    return *(const __attribute__((address_space(4))) T*)p;
  } else {
    static_assert(sizeof(T) % 4 == 0 && std::is_trivially_copyable<T>::value, "word-sized POD");
    const __attribute__((address_space(4))) unsigned* w = (const __attribute__((address_space(4))) unsigned*)p;
    T r;
    unsigned* d = reinterpret_cast<unsigned*>(&r);
#pragma unroll
    for (unsigned k = 0; k < sizeof(T) / 4; ++k) d[k] = w[k];
    return r;
  }
}

// ---------------------------------------------------------------- state I/O
// State is stored as fdm_real (FP32, or FP64 in the FP64-storage build) and
// computed with in FP64; a store rounds to the storage type.
__device__ __forceinline__ double roundState(double x) { return (double)(fdm_real)x; }
__device__ __forceinline__ V3 ld3(const fdm_real* p, int n, int i) {
  return {(double)p[i], (double)p[n + i], (double)p[2 * n + i]};
}
__device__ __forceinline__ void st3(fdm_real* p, int n, int i, V3 v) {
  p[i] = (fdm_real)v.x;
  p[n + i] = (fdm_real)v.y;
  p[2 * n + i] = (fdm_real)v.z;
}
__device__ __forceinline__ Q4 ld4(const fdm_real* p, int n, int i) {
  return {(double)p[i], (double)p[n + i], (double)p[2 * n + i], (double)p[3 * n + i]};
}
__device__ __forceinline__ void st4(fdm_real* p, int n, int i, Q4 q) {
  p[i] = (fdm_real)q.q0;
  p[n + i] = (fdm_real)q.q1;
  p[2 * n + i] = (fdm_real)q.q2;
  p[3 * n + i] = (fdm_real)q.q3;
}
// positions: planar FP64 (a per-step displacement can be far below one FP32
// ulp of the coordinate)
__device__ __forceinline__ V3 ldp(const double* p, int n, int i) { return {p[i], p[n + i], p[2 * n + i]}; }
__device__ __forceinline__ void stp(double* p, int n, int i, V3 v) {
  p[i] = v.x;
  p[n + i] = v.y;
  p[2 * n + i] = v.z;
}
__device__ __forceinline__ V3 ldd3(const double* p) { return {p[0], p[1], p[2]}; }
// a vector as the state storage holds it
// stored-precision triples and quadruples, widened to FP64 where used
struct R3 {
  fdm_real x, y, z;
};
struct R4 {
  fdm_real q0, q1, q2, q3;
};
__device__ __forceinline__ R3 ld3s(const fdm_real* p, int n, int i) { return {p[i], p[n + i], p[2 * n + i]}; }
__device__ __forceinline__ R4 ld4s(const fdm_real* p, int n, int i) {
  return {p[i], p[n + i], p[2 * n + i], p[3 * n + i]};
}
__device__ __forceinline__ void st3s(fdm_real* p, int n, int i, R3 v) {
  p[i] = v.x;
  p[n + i] = v.y;
  p[2 * n + i] = v.z;
}
__device__ __forceinline__ V3 wide(R3 v) { return {(double)v.x, (double)v.y, (double)v.z}; }
__device__ __forceinline__ Q4 wide(R4 q) { return {(double)q.q0, (double)q.q1, (double)q.q2, (double)q.q3}; }
__device__ __forceinline__ R3 narrow3(V3 v) { return {(fdm_real)v.x, (fdm_real)v.y, (fdm_real)v.z}; }
__device__ __forceinline__ V3 roundState(V3 v) {
  return {roundState(v.x), roundState(v.y), roundState(v.z)};
}
__device__ __forceinline__ Q4 roundState(Q4 q) {
  return {roundState(q.q0), roundState(q.q1), roundState(q.q2), roundState(q.q3)};
}
// one history record: the spring in storage precision and the node id
__device__ __forceinline__ fdm_record makeRecord(V3 spring, int node) {
  fdm_record r;
  r.x = (fdm_real)spring.x;
  r.y = (fdm_real)spring.y;
  r.z = (fdm_real)spring.z;
  r.node = node;
  return r;
}
__device__ __forceinline__ V3 recordSpring(const fdm_record& r) {
  return {(double)r.x, (double)r.y, (double)r.z};
}
__device__ __forceinline__ S6 lds6(const double* p) { return {p[0], p[1], p[2], p[3], p[4], p[5]}; }
__host__ __device__ __forceinline__ V3 fromD3(fdm_double3 d) { return {d.x, d.y, d.z}; }

// ------------------------------------------------------------------ grid
// saturating double -> int (NaN -> 0), the device conversion made explicit
__device__ __forceinline__ int toInt(double v) {
  if (v != v) return 0;
  if (v >= 2147483647.0) return INT_MAX;
  if (v <= -2147483648.0) return INT_MIN;
  return (int)v;
}
__device__ __forceinline__ int lin3(I3 c, I3 d) { return c.z * d.y * d.x + c.y * d.x + c.x; }
__device__ __forceinline__ I3 gridPos(V3 p, const fdm_grid& g) {
  return {toInt((p.x - g.min.x) * g.invCell.x), toInt((p.y - g.min.y) * g.invCell.y),
          toInt((p.z - g.min.z) * g.invCell.z)};
}
__device__ __forceinline__ I3 gsize(const fdm_grid& g) { return {g.size.x, g.size.y, g.size.z}; }
// calculateHashKernel (mySpatialGrid.cu:4-26)
__device__ __forceinline__ int cellHash(V3 p, const fdm_grid& g) {
  if (p.x < g.min.x) return -1;
  else if (p.x >= g.max.x) return -1;
  if (p.y < g.min.y) return -1;
  else if (p.y >= g.max.y) return -1;
  if (p.z < g.min.z) return -1;
  else if (p.z >= g.max.z) return -1;
  return lin3(gridPos(p, g), gsize(g));
}
// calculateGhostHashKernel (mySpatialGrid.cu:214-240); its z test compares with max.y
__device__ __forceinline__ int ghostCellHash(V3 p, const fdm_grid& g) {
  I3 c = gridPos(p, g);
  bool out = false;
  if (p.x < g.min.x) { c.x = 0; out = true; }
  else if (p.x >= g.max.x) { c.x = g.size.x - 1; out = true; }
  if (p.y < g.min.y) { c.y = 0; out = true; }
  else if (p.y >= g.max.y) { c.y = g.size.y - 1; out = true; }
  if (p.z < g.min.z) { c.z = 0; out = true; }
  else if (p.z >= g.max.y) { c.z = g.size.z - 1; out = true; }
  return out ? lin3(c, gsize(g)) : -1;
}
__device__ __forceinline__ V3 offsetX(const fdm_grid& g) { return {g.max.x - g.min.x, 0., 0.}; }
__device__ __forceinline__ V3 offsetY(const fdm_grid& g) { return {0., g.max.y - g.min.y, 0.}; }

// updatePositionOutOfBoundaryXD/YD (mySpatialGrid.cu:28-56)
__device__ __forceinline__ V3 wrapAxis(V3 p, const fdm_grid& g, int axis) {
  if (axis == 0) {
    if (p.x < g.min.x) p = vadd(p, offsetX(g));
    else if (p.x >= g.max.x) p = vsub(p, offsetX(g));
  } else {
    if (p.y < g.min.y) p = vadd(p, offsetY(g));
    else if (p.y >= g.max.y) p = vsub(p, offsetY(g));
  }
  return p;
}

// calculateGhostPosition{X,Y,XY}DKernel (mySpatialGrid.cu:87-152)
__device__ __forceinline__ V3 ghostOf(V3 p, const fdm_grid& g, int mask) {
  const I3 c = gridPos(p, g);
  if (mask == 1) return c.x == 0 ? vadd(p, offsetX(g)) : p;
  if (mask == 2) return c.y == 0 ? vadd(p, offsetY(g)) : p;
  return (c.x == 0 && c.y == 0) ? vadd(p, vadd(offsetX(g), offsetY(g))) : p;
}

// The positions the reference holds at each point of updateGhostSpatialGrid:
// after the X wrap (ghost X is taken here), after the Y wrap (ghost Y and XY).
struct WrapStages {
  V3 afterX, afterY;
};
__device__ __forceinline__ WrapStages wrapStages(V3 x, const fdm_grid& g, fdm_periodic per) {
  WrapStages w;
  w.afterX = per.x ? wrapAxis(x, g, 0) : x;
  w.afterY = per.y ? wrapAxis(w.afterX, g, 1) : w.afterX;
  return w;
}

// ------------------------------------------------------ sector periodicity
// PeriodicBoundarySector (PeriodicBoundary.h:319-426): the quarter plane
// x >= min.x, y >= min.y turns about the z axis through min. Image k: 0 is
// +pi/2 (R90), 1 is pi (R180), 2 is -pi/2 (R270).
__host__ __device__ __forceinline__ bool anyPeriodic(const fdm_periodic& p) {
  return p.x || p.y || p.sector;
}
// whether ghost set s (planar X, Y, XY or sector R90, R180, R270) exists
__host__ __device__ __forceinline__ bool ghostSetOn(const fdm_periodic& p, int s) {
  return p.sector ? true : s == 0 ? p.x != 0 : s == 1 ? p.y != 0 : (p.x != 0 && p.y != 0);
}
// rotateVector (myVec.h:143-181) by image k's angle, with its host-evaluated
// coefficients
__device__ __forceinline__ V3 sectorRotV(V3 v, const fdm_periodic& p, int k) {
  const V3 om = {0.0, 0.0, p.angle[k]};
  const V3 oxv = vcross(om, v);
  const V3 oxoxv = vcross(om, oxv);
  return vadd(vadd(v, vmul(oxv, p.A[k])), vmul(oxoxv, p.B[k]));
}
// rotateQuaternion (myQua.h:131-175): the turn's quaternion times q, normalised
__device__ __forceinline__ Q4 sectorRotQ(Q4 q, const fdm_periodic& p, int k) {
  const double th = p.angle[k];
  const double theta = sqrt(0.0 * 0.0 + 0.0 * 0.0 + th * th);
  const double inv = 1.0 / theta;
  const double ux = 0.0 * inv, uy = 0.0 * inv, uz = th * inv;
  const double r0 = p.halfCos[k], r1 = p.halfSin[k] * ux, r2 = p.halfSin[k] * uy,
               r3 = p.halfSin[k] * uz;
  Q4 r;
  r.q0 = r0 * q.q0 - r1 * q.q1 - r2 * q.q2 - r3 * q.q3;
  r.q1 = r0 * q.q1 + r1 * q.q0 + r2 * q.q3 - r3 * q.q2;
  r.q2 = r0 * q.q2 - r1 * q.q3 + r2 * q.q0 + r3 * q.q1;
  r.q3 = r0 * q.q3 + r1 * q.q2 - r2 * q.q1 + r3 * q.q0;
  return qnormalize(r);
}
__device__ __forceinline__ V3 sectorImagePos(V3 x, const fdm_grid& g, const fdm_periodic& p, int k) {
  const V3 mn = fromD3(g.min);
  return vadd(mn, sectorRotV(vsub(x, mn), p, k));
}
// updateVelocityPositionOrientationOutOfSectorKernel (mySpatialGrid.cu:59-85):
// the position after the sector wrap and which turns it took (bit 0: -pi/2
// through the x = min face, bit 1: +pi/2 through the y = min face, in that
// order, the second test on the turned position)
struct SectorWrap {
  V3 x;
  int turns;
};
__device__ __forceinline__ SectorWrap sectorWrap(V3 x, const fdm_grid& g, const fdm_periodic& p) {
  SectorWrap w{x, 0};
  if (w.x.x < g.min.x && w.x.y >= g.min.y) {
    w.x = sectorImagePos(w.x, g, p, 2);
    w.turns |= 1;
  }
  if (w.x.y < g.min.y && w.x.x >= g.min.x) {
    w.x = sectorImagePos(w.x, g, p, 0);
    w.turns |= 2;
  }
  return w;
}
// The same turns applied to a velocity and an orientation; a turned value is
// stored (FP32), so it is used rounded from the wrap on.
__device__ __forceinline__ V3 sectorTurnV(V3 v, int turns, const fdm_periodic& p) {
  if (!turns) return v;
  if (turns & 1) v = sectorRotV(v, p, 2);
  if (turns & 2) v = sectorRotV(v, p, 0);
  return roundState(v);
}
__device__ __forceinline__ Q4 sectorTurnQ(Q4 q, int turns, const fdm_periodic& p) {
  if (!turns) return q;
  if (turns & 1) q = sectorRotQ(q, p, 2);
  if (turns & 2) q = sectorRotQ(q, p, 0);
  return roundState(q);
}
// calculateSectorGhostVelocityPositionOrientationKernel (mySpatialGrid.cu:154-212):
// image k of a wrapped particle is turned when it lies in the first y row
// (R90), first x column (R270) or both (R180); otherwise the set holds the
// particle's own pose (a ghost only if that lies outside the domain)
__device__ __forceinline__ bool sectorImageOn(V3 xw, const fdm_grid& g, int k) {
  const I3 c = gridPos(xw, g);
  return k == 0 ? c.y == 0 : k == 2 ? c.x == 0 : (c.x == 0 && c.y == 0);
}
// The position after this step's periodic wrap, planar or sector.
__device__ __forceinline__ V3 wrapPosition(V3 x, const fdm_grid& g, const fdm_periodic& p) {
  if (p.sector) return sectorWrap(x, g, p).x;
  return wrapStages(x, g, p).afterY;
}

// Sort key of a body of one instance among several: 1 + global cell, or 0
// when the body is outside its instance's grid (ids rounded up past the last
// cell are kept out too, as the table never holds them).
// Periodic images a particle at x has (its visibility bits: set X 1, Y 2,
// XY 4, or the sector's R90, R180, R270 images 1, 2, 4): an image exists
// when its ghost cell id is valid (mySpatialGrid.cu:87-240). `sets` names the
// sets that exist (any instance periodic on the axes), `per` the particle's
// own instance's periodic axes. Bit 8 tells that the particle wraps.
__device__ __forceinline__ int visibilityBits(V3 x, const fdm_grid& g, const fdm_periodic& sets,
                                              const fdm_periodic& per) {
  int bits = 0;
  if (sets.sector) {
    const SectorWrap sw = sectorWrap(x, g, sets);
    for (int k = 0; k < 3; ++k) {
      const V3 gp = sectorImageOn(sw.x, g, k) ? sectorImagePos(sw.x, g, sets, k) : sw.x;
      bits |= (ghostCellHash(gp, g) >= 0 ? 1 : 0) << k;
    }
    if (sw.turns) bits |= 8;
  } else if (sets.x || sets.y) {
    const WrapStages w = wrapStages(x, g, per);
    if (sets.x && per.x && ghostCellHash(ghostOf(w.afterX, g, 1), g) >= 0) bits |= 1;
    if (sets.y && per.y && ghostCellHash(ghostOf(w.afterY, g, 2), g) >= 0) bits |= 2;
    if (sets.x && sets.y && per.x && per.y && ghostCellHash(ghostOf(w.afterY, g, 3), g) >= 0)
      bits |= 4;
    if (!(w.afterY.x == x.x && w.afterY.y == x.y && w.afterY.z == x.z)) bits |= 8;
  }
  return bits;
}

__device__ __forceinline__ unsigned instanceKey(int h, const fdm_grid& g, long long cellOffset) {
  return (h < 0 || (long long)h >= g.num) ? 0u : (unsigned)(cellOffset + h + 1);
}

// Grid state of a body: 0 outside the grid (hash -1), 1 inside, 2 inside but
// with a cell index rounded up to the grid size on some axis, so that its
// cell id aliases another cell.
__device__ __forceinline__ int bodyState(V3 p, const fdm_grid& g) {
  if (cellHash(p, g) < 0) return 0;
  const I3 c = gridPos(p, g);
  return (c.x >= g.size.x || c.y >= g.size.y || c.z >= g.size.z) ? 2 : 1;
}
// node cell clamped into the grid, as the detection kernels clamp it
// (LSParticleContactDetectionKernel.cu:38-46)
__device__ __forceinline__ I3 clampedCell(V3 p, const fdm_grid& g) {
  I3 c = gridPos(p, g);
  if (c.x <= 0) c.x = 0;
  if (c.x >= g.size.x - 1) c.x = g.size.x - 1;
  if (c.y <= 0) c.y = 0;
  if (c.y >= g.size.y - 1) c.y = g.size.y - 1;
  if (c.z <= 0) c.z = 0;
  if (c.z >= g.size.z - 1) c.z = g.size.z - 1;
  return c;
}
// cell coordinates of a valid hash value
__device__ __forceinline__ I3 decodeHash(int h, const fdm_grid& g) {
  const int nx = g.size.x, ny = g.size.y;
  const int z = h / (nx * ny);
  const int r = h - z * nx * ny;
  const int y = r / nx;
  return {r - y * nx, y, z};
}
// Is a target with hash value h inside the 27-cell stencil of a node whose
// clamped cell is c? The stencil visits exactly the in-range cells within one
// cell of c on every axis, so the target is found iff its cell is.
__device__ __forceinline__ bool inStencil(int h, I3 c, const fdm_grid& g) {
  if (h < 0 || (long long)h >= g.num) return false;
  const I3 t = decodeHash(h, g);
  return abs(t.x - c.x) <= 1 && abs(t.y - c.y) <= 1 && abs(t.z - c.z) <= 1;
}

// --------------------------------------------------------------- level set
// trilinear value (LSParticleContactDetectionKernel.cuh:4-32)
__device__ __forceinline__ double trilinear(double x, double y, double z, const double* c) {
  const double wx0 = 1.0 - x, wx1 = x, wy0 = 1.0 - y, wy1 = y, wz0 = 1.0 - z, wz1 = z;
  return c[0] * wx0 * wy0 * wz0 + c[1] * wx1 * wy0 * wz0 + c[2] * wx0 * wy1 * wz0 +
         c[3] * wx1 * wy1 * wz0 + c[4] * wx0 * wy0 * wz1 + c[5] * wx1 * wy0 * wz1 +
         c[6] * wx0 * wy1 * wz1 + c[7] * wx1 * wy1 * wz1;
}
// gradient weights of LSParticleContactDetectionKernel.cuh:34-77, whose x
// parameter is an int: the fractional x in [0,1) enters as 0.
__device__ __forceinline__ V3 gradientIntX(double xf, double y, double z, const double* c) {
  const int xi = (int)xf;
  const double wx0 = 1.0 - xi, wx1 = xi, wy0 = 1.0 - y, wy1 = y, wz0 = 1.0 - z, wz1 = z;
  const double dx = (c[1] - c[0]) * wy0 * wz0 + (c[3] - c[2]) * wy1 * wz0 +
                    (c[5] - c[4]) * wy0 * wz1 + (c[7] - c[6]) * wy1 * wz1;
  const double dy = (c[2] - c[0]) * wx0 * wz0 + (c[3] - c[1]) * wx1 * wz0 +
                    (c[6] - c[4]) * wx0 * wz1 + (c[7] - c[5]) * wx1 * wz1;
  const double dz = (c[4] - c[0]) * wx0 * wy0 + (c[5] - c[1]) * wx1 * wy0 +
                    (c[6] - c[2]) * wx0 * wy1 + (c[7] - c[3]) * wx1 * wy1;
  return {dx, dy, dz};
}

struct Sample {
  double x, y, z;  // fractional position in the SDF cell
  double c[8];     // corner values phi000, phi100, phi010, phi110, phi001, phi101, phi011, phi111
};
// Grid coordinates and SDF box test of the detection kernels
// (LSParticleContactDetectionKernel.cu:72-105): fractional position in s and
// the offset of the cell's first corner in `base`. False outside the box.
__device__ __forceinline__ bool sdfLocate(V3 local, const fdm_template& t, Sample& s, int& base) {
  const double ih = t.invSpacing;
  const double gx = (local.x - t.origin[0]) * ih;
  const double gy = (local.y - t.origin[1]) * ih;
  const double gz = (local.z - t.origin[2]) * ih;
  const int i0 = toInt(floor(gx)), j0 = toInt(floor(gy)), k0 = toInt(floor(gz));
  if (i0 < 0 || j0 < 0 || k0 < 0) return false;
  if (i0 >= t.size[0] - 1 || j0 >= t.size[1] - 1 || k0 >= t.size[2] - 1) return false;
  s.x = gx - (double)i0;
  s.y = gy - (double)j0;
  s.z = gz - (double)k0;
  base = t.sdfStart + (k0 * t.size[0] * t.size[1] + j0 * t.size[0] + i0);
  return true;
}
// The eight corner values of a located cell.
__device__ __forceinline__ void sdfCorners(const float* b, const fdm_template& t, Sample& s) {
  const int nx = t.size[0], nxy = t.size[0] * t.size[1];
  s.c[0] = b[0];
  s.c[1] = b[1];
  s.c[2] = b[nx];
  s.c[3] = b[nx + 1];
  s.c[4] = b[nxy];
  s.c[5] = b[nxy + 1];
  s.c[6] = b[nxy + nx];
  s.c[7] = b[nxy + nx + 1];
}
__device__ __forceinline__ bool sampleSDF(V3 local, const fdm_template& t, const float* sdf,
                                          Sample& s) {
  int base;
  if (!sdfLocate(local, t, s, base)) return false;
  sdfCorners(sdf + base, t, s);
  return true;
}

// ------------------------------------------------------------ contact law
// contactKernel.cuh:4-34
__device__ __forceinline__ V3 slidingSpring(V3 prev, V3 vt, V3 n, V3 Fn, double mu, double ks,
                                            double ct, double dt) {
  V3 spring = {0., 0., 0.};
  if (mu > 0. && ks > 0.) {
    V3 p1 = vsub(prev, vmul(n, vdot(prev, n)));
    const double a1 = vlen(p1);
    if (!nearZero(a1)) p1 = vmul(p1, vlen(prev) / a1);
    spring = vadd(p1, vmul(vt, dt));
    V3 sf = vsub(vmul(spring, -ks), vmul(vt, ct));
    const double asf = vlen(sf), afn = vlen(Fn);
    if (!nearZero(asf) && asf > mu * afn) {
      const double ratio = mu * afn / asf;
      sf = vmul(sf, ratio);
      spring = vdiv(vneg(vadd(sf, vmul(vt, ct))), ks);
    }
  }
  return spring;
}
// contactKernel.cuh:36-77. `diss` is the material's precomputed dissipation;
// the damping coefficients use FP32 square roots as the reference does.
__device__ __forceinline__ void linearContact(V3& F, V3& spring, V3 v, V3 n, double delta,
                                              double dt, double kn, double ks, double mu,
                                              double diss, double effM) {
  if (delta > 0.) {
    const double cn = 2. * diss * (double)sqrtf((float)(effM * kn));
    const V3 vn = vmul(n, vdot(v, n));
    const V3 Fn = vsub(vmul(n, kn * delta), vmul(vn, cn));
    const double ct = 2. * diss * (double)sqrtf((float)(effM * ks));
    const V3 vt = vsub(v, vn);
    spring = slidingSpring(spring, vt, n, Fn, mu, ks, ct, dt);
    const V3 Fs = vsub(vmul(spring, -ks), vmul(vt, ct));
    F = vadd(Fn, Fs);
  } else {
    spring = {0., 0., 0.};
  }
}

// ------------------------------------------------------- wave64 helpers
__device__ __forceinline__ int laneId() { return __lane_id(); }
__device__ __forceinline__ double readlaneD(double v, int lane) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)bits, lane);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(bits >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ int readlaneI(int v, int lane) {
  return (int)__builtin_amdgcn_readlane((unsigned)v, lane);
}
__device__ __forceinline__ unsigned long long ballot64(bool p) { return __ballot(p); }

}  // namespace fdm
