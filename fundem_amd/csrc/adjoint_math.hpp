// Hand-written reverse-mode derivatives (vector-Jacobian products) of the
// level-set DEM step, for the adjoint timestep.
//
// The primal routines restate the step's arithmetic (contact geometry of
// LSParticleContactDetectionKernel.cu:128-305 with the trilinear value and
// gradient of LSParticleContactDetectionKernel.cuh:4-77; LinearContact and
// integrateSlidingSpring of contactKernel.cuh:4-77; the kick and drift of
// particleIntegrationKernel.cu:3-48 with rotateTensorByQuaternion,
// myMat.h:107-158, and quaternionIntegration, myQua.h:77-93). They are
// templates over the scalar type so that a forward-mode (dual number)
// instantiation can check every *Vjp routine on the host; the device
// instantiates them with double only, as the linearisation point of the
// reverse sweep.
//
// Each *Vjp routine takes the primal inputs and the adjoints of the outputs
// and ADDS the adjoints of the inputs into its output arguments. Branches
// (contact open or closed, stick or slip, zero-length guards) follow the
// primal values, so each routine is the derivative of the branch taken.
//
// The header is self-contained: the including translation unit provides
// __host__ / __device__ (hip_runtime.h, or empty definitions for a host-only
// build) and sqrt / floor for its scalar types.
#pragma once
#include <cmath>

namespace fdm {
namespace adj {

#define FDM_AD __host__ __device__ inline

FDM_AD double val(double x) { return x; }
FDM_AD double fsqrt(double x) { return sqrt(x); }

template <class T>
struct Vec {
  T x, y, z;
};
template <class T>
struct Quat {
  T q0, q1, q2, q3;
};
// symmetric 3x3 as (xx, yy, zz, xy, xz, yz)
template <class T>
struct Sym {
  T xx, yy, zz, xy, xz, yz;
};

using V = Vec<double>;
using Q = Quat<double>;
using S = Sym<double>;

template <class T> FDM_AD Vec<T> add(Vec<T> a, Vec<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> FDM_AD Vec<T> sub(Vec<T> a, Vec<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> FDM_AD Vec<T> neg(Vec<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T, class U> FDM_AD Vec<T> mul(Vec<T> a, U s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> FDM_AD T dot(Vec<T> a, Vec<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> FDM_AD Vec<T> cross(Vec<T> a, Vec<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> FDM_AD T norm(Vec<T> a) { return fsqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
FDM_AD bool nearZero(double x) { return fabs(x) <= 1.0e-20; }
FDM_AD void acc(V& a, V b) { a.x += b.x; a.y += b.y; a.z += b.z; }
FDM_AD void accq(Q& a, Q b) { a.q0 += b.q0; a.q1 += b.q1; a.q2 += b.q2; a.q3 += b.q3; }

// ------------------------------------------------------------ rotations
// v' = v + q0 t + u x t, t = 2 u x v (myQua.h:104-114); the conjugate
// rotation negates u (myQua.h:120-129).
template <class T>
FDM_AD Vec<T> qrot(Quat<T> q, Vec<T> v) {
  const Vec<T> u = {q.q1, q.q2, q.q3};
  const Vec<T> t = mul(cross(u, v), 2.0);
  return add(add(v, mul(t, q.q0)), cross(u, t));
}
template <class T>
FDM_AD Vec<T> qrotT(Vec<T> v, Quat<T> q) {
  return qrot(Quat<T>{q.q0, -q.q1, -q.q2, -q.q3}, v);
}
// adjoint of qrot(q, v) for output adjoint yb
FDM_AD void qrotVjp(Q q, V v, V yb, Q& qb, V& vb) {
  const V u = {q.q1, q.q2, q.q3};
  const V t = mul(cross(u, v), 2.0);
  const V tb = add(mul(yb, q.q0), cross(yb, u));
  qb.q0 += dot(t, yb);
  const V ub = add(cross(t, yb), mul(cross(v, tb), 2.0));
  qb.q1 += ub.x;
  qb.q2 += ub.y;
  qb.q3 += ub.z;
  acc(vb, add(yb, mul(cross(tb, u), 2.0)));
}
FDM_AD void qrotTVjp(V v, Q q, V yb, V& vb, Q& qb) {
  Q c = {0., 0., 0., 0.};
  qrotVjp(Q{q.q0, -q.q1, -q.q2, -q.q3}, v, yb, c, vb);
  qb.q0 += c.q0;
  qb.q1 -= c.q1;
  qb.q2 -= c.q2;
  qb.q3 -= c.q3;
}

// rotation-matrix rows of q as rotateTensorByQuaternion forms them
template <class T>
FDM_AD void rotRows(Quat<T> q, Vec<T> r[3]) {
  const T q0 = q.q0, q1 = q.q1, q2 = q.q2, q3 = q.q3;
  const T q11 = 2. * q1 * q1, q22 = 2. * q2 * q2, q33 = 2. * q3 * q3;
  r[0] = {1.0 - q22 - q33, 2. * (q1 * q2 - q0 * q3), 2. * (q1 * q3 + q0 * q2)};
  r[1] = {2. * (q1 * q2 + q0 * q3), 1.0 - q11 - q33, 2. * (q2 * q3 - q0 * q1)};
  r[2] = {2. * (q1 * q3 - q0 * q2), 2. * (q2 * q3 + q0 * q1), 1.0 - q11 - q22};
}
template <class T, class U>
FDM_AD Vec<T> symMul(Sym<U> m, Vec<T> v) {
  return {m.xx * v.x + m.xy * v.y + m.xz * v.z, m.xy * v.x + m.yy * v.y + m.yz * v.z,
          m.xz * v.x + m.yz * v.y + m.zz * v.z};
}
// R I R^T for a symmetric I (myMat.h:107-158)
template <class T>
FDM_AD Sym<T> srot(Quat<T> q, S I) {
  Vec<T> r[3];
  rotRows(q, r);
  const Vec<T> a0 = symMul(I, r[0]), a1 = symMul(I, r[1]), a2 = symMul(I, r[2]);
  return {dot(a0, r[0]), dot(a1, r[1]), dot(a2, r[2]), dot(a0, r[1]), dot(a0, r[2]), dot(a1, r[2])};
}
// adjoint of srot(q, I) for the six output adjoints Pb
FDM_AD void srotVjp(Q q, S I, S Pb, Q& qb) {
  V r[3];
  rotRows(q, r);
  const V a[3] = {symMul(I, r[0]), symMul(I, r[1]), symMul(I, r[2])};
  // P_jk = r_j . I r_k: adds Pb_jk (I r_k) to row j and Pb_jk (I r_j) to row k
  V rb[3] = {mul(a[0], 2. * Pb.xx), mul(a[1], 2. * Pb.yy), mul(a[2], 2. * Pb.zz)};
  acc(rb[0], mul(a[1], Pb.xy));
  acc(rb[1], mul(a[0], Pb.xy));
  acc(rb[0], mul(a[2], Pb.xz));
  acc(rb[2], mul(a[0], Pb.xz));
  acc(rb[1], mul(a[2], Pb.yz));
  acc(rb[2], mul(a[1], Pb.yz));
  const double q0 = q.q0, q1 = q.q1, q2 = q.q2, q3 = q.q3;
  // row 0: (1 - 2q2^2 - 2q3^2, 2(q1q2 - q0q3), 2(q1q3 + q0q2))
  qb.q2 += -4. * q2 * rb[0].x;
  qb.q3 += -4. * q3 * rb[0].x;
  qb.q1 += 2. * q2 * rb[0].y;
  qb.q2 += 2. * q1 * rb[0].y;
  qb.q0 += -2. * q3 * rb[0].y;
  qb.q3 += -2. * q0 * rb[0].y;
  qb.q1 += 2. * q3 * rb[0].z;
  qb.q3 += 2. * q1 * rb[0].z;
  qb.q0 += 2. * q2 * rb[0].z;
  qb.q2 += 2. * q0 * rb[0].z;
  // row 1: (2(q1q2 + q0q3), 1 - 2q1^2 - 2q3^2, 2(q2q3 - q0q1))
  qb.q1 += 2. * q2 * rb[1].x;
  qb.q2 += 2. * q1 * rb[1].x;
  qb.q0 += 2. * q3 * rb[1].x;
  qb.q3 += 2. * q0 * rb[1].x;
  qb.q1 += -4. * q1 * rb[1].y;
  qb.q3 += -4. * q3 * rb[1].y;
  qb.q2 += 2. * q3 * rb[1].z;
  qb.q3 += 2. * q2 * rb[1].z;
  qb.q0 += -2. * q1 * rb[1].z;
  qb.q1 += -2. * q0 * rb[1].z;
  // row 2: (2(q1q3 - q0q2), 2(q2q3 + q0q1), 1 - 2q1^2 - 2q2^2)
  qb.q1 += 2. * q3 * rb[2].x;
  qb.q3 += 2. * q1 * rb[2].x;
  qb.q0 += -2. * q2 * rb[2].x;
  qb.q2 += -2. * q0 * rb[2].x;
  qb.q2 += 2. * q3 * rb[2].y;
  qb.q3 += 2. * q2 * rb[2].y;
  qb.q0 += 2. * q1 * rb[2].y;
  qb.q1 += 2. * q0 * rb[2].y;
  qb.q1 += -4. * q1 * rb[2].z;
  qb.q2 += -4. * q2 * rb[2].z;
}
// adjoint of symMul(m, v) with respect to m (as six entries) and v
FDM_AD void symMulVjp(S m, V v, V yb, S& mb, V& vb) {
  mb.xx += yb.x * v.x;
  mb.yy += yb.y * v.y;
  mb.zz += yb.z * v.z;
  mb.xy += yb.x * v.y + yb.y * v.x;
  mb.xz += yb.x * v.z + yb.z * v.x;
  mb.yz += yb.y * v.z + yb.z * v.y;
  acc(vb, symMul(m, yb));
}

// ------------------------------------------------------- integration
// Angular half-kick w' = w + h (R I^-1 R^T)(T - w x (R I R^T w))
// (particleIntegrationKernel.cu:3-32).
template <class T>
FDM_AD Vec<T> spin(Vec<T> w, Quat<T> q, Vec<T> tq, S I, S invI, double h) {
  const Vec<T> Iw = symMul(srot(q, I), w);
  return add(w, mul(symMul(srot(q, invI), sub(tq, cross(w, Iw))), h));
}
FDM_AD void spinVjp(V w, Q q, V tq, S I, S invI, double h, V wpb, V& wb, Q& qb, V& tqb) {
  const S Si = srot(q, I), Mi = srot(q, invI);
  const V Iw = symMul(Si, w);
  const V a = sub(tq, cross(w, Iw));
  acc(wb, wpb);
  const V yb = mul(wpb, h);
  S Mb = {0., 0., 0., 0., 0., 0.};
  V ab = {0., 0., 0.};
  symMulVjp(Mi, a, yb, Mb, ab);
  acc(tqb, ab);
  const V cb = neg(ab);  // c = w x Iw
  acc(wb, cross(Iw, cb));
  const V Iwb = cross(cb, w);
  S Sb = {0., 0., 0., 0., 0., 0.};
  symMulVjp(Si, w, Iwb, Sb, wb);
  srotVjp(q, invI, Mb, qb);
  srotVjp(q, I, Sb, qb);
}

// q' = normalise(q + 1/2 dt (0, w) (x) q) (myQua.h:77-93)
template <class T>
FDM_AD Quat<T> qint(Quat<T> q, Vec<T> w, double dt) {
  const Vec<T> v = mul(w, 0.5 * dt);
  const T d0 = -q.q1 * v.x - q.q2 * v.y - q.q3 * v.z;
  const T d1 = q.q0 * v.x + q.q3 * v.y - q.q2 * v.z;
  const T d2 = -q.q3 * v.x + q.q0 * v.y + q.q1 * v.z;
  const T d3 = q.q2 * v.x - q.q1 * v.y + q.q0 * v.z;
  const Quat<T> u = {q.q0 + d0, q.q1 + d1, q.q2 + d2, q.q3 + d3};
  const T l2 = u.q0 * u.q0 + u.q1 * u.q1 + u.q2 * u.q2 + u.q3 * u.q3;
  if (nearZero(val(l2))) return u;
  const T inv = 1.0 / fsqrt(l2);
  return {u.q0 * inv, u.q1 * inv, u.q2 * inv, u.q3 * inv};
}
FDM_AD void qintVjp(Q q, V w, double dt, Q qpb, Q& qb, V& wb) {
  const V v = mul(w, 0.5 * dt);
  const Q u = {q.q0 - q.q1 * v.x - q.q2 * v.y - q.q3 * v.z, q.q1 + q.q0 * v.x + q.q3 * v.y - q.q2 * v.z,
               q.q2 - q.q3 * v.x + q.q0 * v.y + q.q1 * v.z, q.q3 + q.q2 * v.x - q.q1 * v.y + q.q0 * v.z};
  const double l2 = u.q0 * u.q0 + u.q1 * u.q1 + u.q2 * u.q2 + u.q3 * u.q3;
  Q ub = qpb;
  if (!nearZero(l2)) {
    const double l = sqrt(l2);
    const Q n = {u.q0 / l, u.q1 / l, u.q2 / l, u.q3 / l};
    const double p = n.q0 * qpb.q0 + n.q1 * qpb.q1 + n.q2 * qpb.q2 + n.q3 * qpb.q3;
    ub = {(qpb.q0 - n.q0 * p) / l, (qpb.q1 - n.q1 * p) / l, (qpb.q2 - n.q2 * p) / l,
          (qpb.q3 - n.q3 * p) / l};
  }
  accq(qb, ub);
  V vb = {0., 0., 0.};
  // d0 = -q1 vx - q2 vy - q3 vz
  qb.q1 -= v.x * ub.q0; qb.q2 -= v.y * ub.q0; qb.q3 -= v.z * ub.q0;
  vb.x -= q.q1 * ub.q0; vb.y -= q.q2 * ub.q0; vb.z -= q.q3 * ub.q0;
  // d1 = q0 vx + q3 vy - q2 vz
  qb.q0 += v.x * ub.q1; qb.q3 += v.y * ub.q1; qb.q2 -= v.z * ub.q1;
  vb.x += q.q0 * ub.q1; vb.y += q.q3 * ub.q1; vb.z -= q.q2 * ub.q1;
  // d2 = -q3 vx + q0 vy + q1 vz
  qb.q3 -= v.x * ub.q2; qb.q0 += v.y * ub.q2; qb.q1 += v.z * ub.q2;
  vb.x -= q.q3 * ub.q2; vb.y += q.q0 * ub.q2; vb.z += q.q1 * ub.q2;
  // d3 = q2 vx - q1 vy + q0 vz
  qb.q2 += v.x * ub.q3; qb.q1 -= v.y * ub.q3; qb.q0 += v.z * ub.q3;
  vb.x += q.q2 * ub.q3; vb.y -= q.q1 * ub.q3; vb.z += q.q0 * ub.q3;
  acc(wb, mul(vb, 0.5 * dt));
}

// --------------------------------------------------------- contact geometry
// One node-contact: node s of body A (its frame) against the level set of B.
struct GeoConst {
  V s;          // node, A's frame
  V origin;     // B's SDF origin (B's frame)
  double ih;    // 1 / SDF spacing
  double c[8];  // corner values of the node's SDF cell, as in the detection kernels
  int cell[3];  // that cell
};
template <class T>
struct GeoOut {
  Vec<T> pc, n;
  T delta;
};
// trilinear value and its true gradient in cell coordinates
template <class T>
FDM_AD T tri(T x, T y, T z, const double* c) {
  const T wx0 = 1.0 - x, wx1 = x, wy0 = 1.0 - y, wy1 = y, wz0 = 1.0 - z, wz1 = z;
  return c[0] * wx0 * wy0 * wz0 + c[1] * wx1 * wy0 * wz0 + c[2] * wx0 * wy1 * wz0 +
         c[3] * wx1 * wy1 * wz0 + c[4] * wx0 * wy0 * wz1 + c[5] * wx1 * wy0 * wz1 +
         c[6] * wx0 * wy1 * wz1 + c[7] * wx1 * wy1 * wz1;
}
FDM_AD V triGrad(double x, double y, double z, const double* c) {
  const double wx0 = 1.0 - x, wy0 = 1.0 - y, wz0 = 1.0 - z;
  return {(c[1] - c[0]) * wy0 * wz0 + (c[3] - c[2]) * y * wz0 + (c[5] - c[4]) * wy0 * z +
              (c[7] - c[6]) * y * z,
          (c[2] - c[0]) * wx0 * wz0 + (c[3] - c[1]) * x * wz0 + (c[6] - c[4]) * wx0 * z +
              (c[7] - c[5]) * x * z,
          (c[4] - c[0]) * wx0 * wy0 + (c[5] - c[1]) * x * wy0 + (c[6] - c[2]) * wx0 * y +
              (c[7] - c[3]) * x * y};
}
// the detection kernels' gradient, whose x weight is an integer (0 inside a
// cell): it depends on y and z only (LSParticleContactDetectionKernel.cuh:34-77)
template <class T>
FDM_AD Vec<T> gradIntX(T y, T z, const double* c) {
  const T wy0 = 1.0 - y, wz0 = 1.0 - z;
  return {(c[1] - c[0]) * wy0 * wz0 + (c[3] - c[2]) * y * wz0 + (c[5] - c[4]) * wy0 * z +
              (c[7] - c[6]) * y * z,
          (c[2] - c[0]) * wz0 + (c[6] - c[4]) * z, (c[4] - c[0]) * wy0 + (c[6] - c[2]) * y};
}

template <class T>
FDM_AD GeoOut<T> geometry(Vec<T> xA, Quat<T> qA, Vec<T> xB, Quat<T> qB, const GeoConst& k) {
  const Vec<T> sk = {T(k.s.x), T(k.s.y), T(k.s.z)};
  const Vec<T> p = add(qrot(qA, sk), xA);
  const Vec<T> l = qrotT(sub(p, xB), qB);
  const T fx = (l.x - k.origin.x) * k.ih - (double)k.cell[0];
  const T fy = (l.y - k.origin.y) * k.ih - (double)k.cell[1];
  const T fz = (l.z - k.origin.z) * k.ih - (double)k.cell[2];
  const T delta = -tri(fx, fy, fz, k.c);
  const Vec<T> m = qrot(qB, gradIntX(fy, fz, k.c));
  const T lm = norm(m);
  const Vec<T> n = nearZero(val(lm)) ? Vec<T>{T(0.), T(0.), T(0.)} : mul(m, 1.0 / lm);
  return {add(p, mul(n, 0.5 * delta)), n, delta};
}
FDM_AD void geometryVjp(V xA, Q qA, V xB, Q qB, const GeoConst& k, V pcb, V nb, double deltab,
                        V& xAb, Q& qAb, V& xBb, Q& qBb) {
  const V p = add(qrot(qA, k.s), xA);
  const V r = sub(p, xB);
  const V l = qrotT(r, qB);
  const double fx = (l.x - k.origin.x) * k.ih - (double)k.cell[0];
  const double fy = (l.y - k.origin.y) * k.ih - (double)k.cell[1];
  const double fz = (l.z - k.origin.z) * k.ih - (double)k.cell[2];
  const double delta = -tri(fx, fy, fz, k.c);
  const V g = gradIntX(fy, fz, k.c);
  const V m = qrot(qB, g);
  const double lm = norm(m);
  V pb = pcb;
  V n = {0., 0., 0.};
  V mb = {0., 0., 0.};
  if (!nearZero(lm)) {
    n = mul(m, 1.0 / lm);
    const V ntb = add(nb, mul(pcb, 0.5 * delta));
    mb = mul(sub(ntb, mul(n, dot(n, ntb))), 1.0 / lm);
  }
  const double dtb = deltab + 0.5 * dot(n, pcb);
  V gb = {0., 0., 0.};
  qrotVjp(qB, g, mb, qBb, gb);
  const double* c = k.c;
  // gradIntX partials: gx(y, z), gy(z), gz(y)
  double fyb = gb.x * (-(c[1] - c[0]) * (1.0 - fz) + (c[3] - c[2]) * (1.0 - fz) - (c[5] - c[4]) * fz +
                       (c[7] - c[6]) * fz) +
               gb.z * (-(c[4] - c[0]) + (c[6] - c[2]));
  double fzb = gb.x * (-(c[1] - c[0]) * (1.0 - fy) - (c[3] - c[2]) * fy + (c[5] - c[4]) * (1.0 - fy) +
                       (c[7] - c[6]) * fy) +
               gb.y * (-(c[2] - c[0]) + (c[6] - c[4]));
  const V tg = triGrad(fx, fy, fz, c);
  const double fxb = -dtb * tg.x;
  fyb += -dtb * tg.y;
  fzb += -dtb * tg.z;
  const V lb = {fxb * k.ih, fyb * k.ih, fzb * k.ih};
  V rb = {0., 0., 0.};
  qrotTVjp(r, qB, lb, rb, qBb);
  acc(pb, rb);
  acc(xBb, neg(rb));
  acc(xAb, pb);
  V sb = {0., 0., 0.};
  qrotVjp(qA, k.s, pb, qAb, sb);
}

// ------------------------------------------------------------- contact law
// Pair constants; kn, ks, mu, zeta are the pair's combined values.
struct LawConst {
  double kn, ks, mu, zeta, effM, dt;
};
template <class T>
struct LawOut {
  Vec<T> F, spring;
};
// LinearContact with integrateSlidingSpring (contactKernel.cuh:4-77); the
// damping coefficients are 2 zeta sqrt(effM k).
template <class T, class P>
FDM_AD LawOut<T> law(Vec<T> sp, Vec<T> v, Vec<T> n, T delta, P kn, P ks, P mu, P zeta, double effM,
                     double dt) {
  const Vec<T> zero = {T(0.), T(0.), T(0.)};
  if (!(val(delta) > 0.)) return {zero, zero};
  const P cn = 2. * zeta * fsqrt(effM * kn);
  const P ct = 2. * zeta * fsqrt(effM * ks);
  const Vec<T> vn = mul(n, dot(v, n));
  const Vec<T> Fn = sub(mul(n, kn * delta), mul(vn, cn));
  const Vec<T> vt = sub(v, vn);
  Vec<T> s = zero;
  if (val(mu) > 0. && val(ks) > 0.) {
    Vec<T> p1 = sub(sp, mul(n, dot(sp, n)));
    const T a1 = norm(p1);
    if (!nearZero(val(a1))) p1 = mul(p1, norm(sp) / a1);
    s = add(p1, mul(vt, dt));
    Vec<T> sf = sub(mul(s, -ks), mul(vt, ct));
    const T asf = norm(sf), afn = norm(Fn);
    if (!nearZero(val(asf)) && val(asf) > val(mu * afn)) {
      sf = mul(sf, mu * afn / asf);
      s = mul(neg(add(sf, mul(vt, ct))), 1.0 / ks);
    }
  }
  return {add(Fn, sub(mul(s, -ks), mul(vt, ct))), s};
}
struct LawBar {
  double kn, ks, mu, zeta;
};
FDM_AD void lawVjp(V sp, V v, V n, double delta, const LawConst& c, V Fb, V sb, V& spb, V& vb,
                   V& nb, double& deltab, LawBar& pb) {
  if (!(delta > 0.)) return;
  const double kn = c.kn, ks = c.ks, mu = c.mu, zeta = c.zeta, effM = c.effM, dt = c.dt;
  const double sqn = sqrt(effM * kn), sqt = sqrt(effM * ks);
  const double cn = 2. * zeta * sqn, ct = 2. * zeta * sqt;
  const double vdn = dot(v, n);
  const V vn = mul(n, vdn);
  const V Fn = sub(mul(n, kn * delta), mul(vn, cn));
  const V vt = sub(v, vn);
  // primal spring path
  bool springOn = mu > 0. && ks > 0., slip = false, scaled = false;
  V p1 = {0., 0., 0.}, s = {0., 0., 0.}, sf = {0., 0., 0.}, out = {0., 0., 0.};
  double a1 = 0., L = 0., asf = 0., afn = 0., d0 = 0.;
  if (springOn) {
    d0 = dot(sp, n);
    p1 = sub(sp, mul(n, d0));
    a1 = norm(p1);
    L = norm(sp);
    scaled = !nearZero(a1);
    const V p1s = scaled ? mul(p1, L / a1) : p1;
    s = add(p1s, mul(vt, dt));
    sf = sub(mul(s, -ks), mul(vt, ct));
    asf = norm(sf);
    afn = norm(Fn);
    slip = !nearZero(asf) && asf > mu * afn;
    out = s;
    if (slip) out = mul(neg(add(mul(sf, mu * afn / asf), mul(vt, ct))), 1.0 / ks);
  }
  V Fnb = Fb, vtb = {0., 0., 0.};
  double ctb = 0., cnb = 0.;
  // F = Fn + (-ks out - ct vt)
  V outb = sub(sb, mul(Fb, ks));
  pb.ks += -dot(out, Fb);
  ctb += -dot(vt, Fb);
  acc(vtb, mul(Fb, -ct));
  if (springOn) {
    V s_b = {0., 0., 0.};
    if (slip) {
      // out = -(sf2 + ct vt) / ks with sf2 = mu afn sf / asf
      const V sf2 = mul(sf, mu * afn / asf);
      const V X = add(sf2, mul(vt, ct));
      const V sf2b = mul(outb, -1.0 / ks);
      ctb += -dot(vt, outb) / ks;
      acc(vtb, mul(outb, -ct / ks));
      pb.ks += dot(X, outb) / (ks * ks);
      const V sh = mul(sf, 1.0 / asf);
      const double shsf2b = dot(sh, sf2b);
      pb.mu += afn * shsf2b;
      const double afnb = mu * shsf2b;
      const V shb = mul(sf2b, mu * afn);
      const V sfb = mul(sub(shb, mul(sh, dot(sh, shb))), 1.0 / asf);
      if (afn > 0.) acc(Fnb, mul(Fn, afnb / afn));
      // sf = -ks s - ct vt
      s_b = mul(sfb, -ks);
      pb.ks += -dot(s, sfb);
      ctb += -dot(vt, sfb);
      acc(vtb, mul(sfb, -ct));
    } else {
      s_b = outb;
    }
    // s = p1s + vt dt
    acc(vtb, mul(s_b, dt));
    V p1b = s_b;
    if (scaled) {
      const V ph = mul(p1, 1.0 / a1);
      const double phb = dot(ph, s_b);
      p1b = mul(sub(s_b, mul(ph, phb)), L / a1);
      if (L > 0.) acc(spb, mul(sp, phb / L));
    }
    // p1 = sp - n (sp . n)
    acc(spb, p1b);
    acc(nb, mul(p1b, -d0));
    const double d0b = -dot(n, p1b);
    acc(spb, mul(n, d0b));
    acc(nb, mul(sp, d0b));
  }
  // vt = v - vn
  acc(vb, vtb);
  V vnb = neg(vtb);
  // Fn = kn delta n - cn vn
  const double nFnb = dot(n, Fnb);
  pb.kn += delta * nFnb;
  deltab += kn * nFnb;
  acc(nb, mul(Fnb, kn * delta));
  cnb += -dot(vn, Fnb);
  acc(vnb, mul(Fnb, -cn));
  // vn = n (v . n)
  acc(nb, mul(vnb, vdn));
  const double vdnb = dot(n, vnb);
  acc(vb, mul(n, vdnb));
  acc(nb, mul(v, vdnb));
  // cn = 2 zeta sqrt(effM kn), ct = 2 zeta sqrt(effM ks)
  pb.zeta += 2. * sqn * cnb + 2. * sqt * ctb;
  if (sqn > 0.) pb.kn += zeta * effM * cnb / sqn;
  if (sqt > 0.) pb.ks += zeta * effM * ctb / sqt;
}

// --------------------------------------------------------------- one contact
// Force on A, torques on A and B about their positions, and the new spring
// of a node-contact between A and target B (B's force is -F).
template <class T>
struct ContactOut {
  Vec<T> F, TA, TB, spring;
};
template <class T, class P>
FDM_AD ContactOut<T> contact(Vec<T> xA, Quat<T> qA, Vec<T> vA, Vec<T> wA, Vec<T> xB, Quat<T> qB,
                             Vec<T> vB, Vec<T> wB, Vec<T> sp, const GeoConst& g, P kn, P ks, P mu,
                             P zeta, double effM, double dt) {
  const GeoOut<T> geo = geometry(xA, qA, xB, qB, g);
  const Vec<T> rA = sub(geo.pc, xA), rB = sub(geo.pc, xB);
  const Vec<T> vc = sub(add(vA, cross(wA, rA)), add(vB, cross(wB, rB)));
  const LawOut<T> f = law(sp, vc, geo.n, geo.delta, kn, ks, mu, zeta, effM, dt);
  return {f.F, cross(rA, f.F), cross(rB, neg(f.F)), f.spring};
}
struct BodyBar {
  V x;
  Q q;
  V v, w;
};
FDM_AD void contactVjp(V xA, Q qA, V vA, V wA, V xB, Q qB, V vB, V wB, V sp, const GeoConst& g,
                       const LawConst& c, V Fb, V TAb, V TBb, V springb, BodyBar& Ab, BodyBar& Bb,
                       V& spb, LawBar& pb) {
  const GeoOut<double> geo = geometry(xA, qA, xB, qB, g);
  const V rA = sub(geo.pc, xA), rB = sub(geo.pc, xB);
  const V vc = sub(add(vA, cross(wA, rA)), add(vB, cross(wB, rB)));
  const LawOut<double> f = law(sp, vc, geo.n, geo.delta, c.kn, c.ks, c.mu, c.zeta, c.effM, c.dt);
  // torques TA = rA x F, TB = rB x (-F)
  const V Ft = add(add(Fb, cross(TAb, rA)), cross(rB, TBb));
  V rAb = cross(f.F, TAb), rBb = cross(TBb, f.F);
  V vcb = {0., 0., 0.}, nb = {0., 0., 0.};
  double deltab = 0.;
  lawVjp(sp, vc, geo.n, geo.delta, c, Ft, springb, spb, vcb, nb, deltab, pb);
  // vc = vA + wA x rA - vB - wB x rB
  acc(Ab.v, vcb);
  acc(Bb.v, neg(vcb));
  acc(Ab.w, cross(rA, vcb));
  acc(rAb, cross(vcb, wA));
  acc(Bb.w, neg(cross(rB, vcb)));
  acc(rBb, neg(cross(vcb, wB)));
  const V pcb = add(rAb, rBb);
  acc(Ab.x, neg(rAb));
  acc(Bb.x, neg(rBb));
  geometryVjp(xA, qA, xB, qB, g, pcb, nb, deltab, Ab.x, Ab.q, Bb.x, Bb.q);
}

}  // namespace adj
}  // namespace fdm
