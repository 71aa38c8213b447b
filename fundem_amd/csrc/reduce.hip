// Bond forces (VBond) and the deterministic force/torque reduction.
//
// The reference adds every contact and bond contribution with double
// atomicAdds (contactKernel.cu:120-123,203-204,284-290,367-370), so its sums
// depend on scheduling. Here each particle folds its records serially in a
// fixed order: particle pairs by partner (pairs where it is the target, then
// pairs it owns), wall pairs by wall, ghost sets X, Y, XY by partner real id
// (owned before mirrored on a tie), bonds by bond id. Each record is a
// per-pair sum that the narrow phase accumulated in node order. The second
// half-kick follows in the same thread.
#include "fdm_device.hpp"
#include "fdm_launch.hpp"

namespace fdm {

// VBond (contactKernel.cuh:138-199); returns the new bonded flag.
__device__ __forceinline__ int vbond(V3& F, V3& Mij, V3& Mji, double& sig, double& tau, V3 rbi,
                                     V3 rbj, V3 n1i, V3 n2i, V3 n3i, V3 n1j, V3 n2j, V3 n3j,
                                     double B1, double B2, double B3, double B4, double R, double D0,
                                     double st, double coh, double mub) {
  const V3 rbij = vsub(rbi, rbj);
  const double D = vlen(rbij);
  if (nearZero(D)) return 1;
  const V3 nij = vnormalize(rbij);
  const V3 e = vneg(nij);
  const V3 dn1 = vsub(n1j, n1i);
  F = vadd(vmul(e, B1 * (D - D0)), vmul(vsub(dn1, vmul(e, vdot(dn1, e))), B2 / (2 * D)));
  const V3 MTB =
      vsub(vmul(vcross(n1j, n1i), B3), vmul(vadd(vcross(n2j, n2i), vcross(n3j, n3i)), 0.5 * B4));
  Mij = vadd(vmul(vcross(e, n1i), -0.5 * B2), MTB);
  Mji = vsub(vmul(vcross(e, n1j), 0.5 * B2), MTB);
  const double Fn = vdot(F, nij);
  const V3 Fs = vsub(F, vmul(nij, Fn));
  const double Tt = vdot(MTB, nij);
  const V3 Tb = vsub(MTB, vmul(nij, Tt));
  const double A = R * R * kPi;
  const double I = A * R * R / 4.;
  const double J = 2 * I;
  sig = -Fn / A + vlen(Tb) / I * R;
  tau = vlen(Fs) / A + fabs(Tt) / J * R;
  if (st > 0. && sig > st) return 0;
  else if (coh > 0. && tau > coh - mub * sig) return 0;
  return 1;
}

// addBondedForceTorqueKernel (contactKernel.cu:293-371), results per bond.
__global__ void kBonds(fdm_bodies P, fdm_bonds b) {
  if (gateOff(b.gate)) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.n) return;
  if (b.active[k] == 0) {
    b.applied[k] = 0;
    b.maxNormalStress[k] = 0.;
    b.maxShearStress[k] = 0.;
    return;
  }
  const int n = P.n;
  const int i = b.master[k], j = b.slave[k];
  const Q4 qi = ld4(P.q, n, i), qj = ld4(P.q, n, j);
  const fdm_real* L = b.local + 24 * (size_t)k;
  auto lv = [&](int o) { return V3{(double)L[o], (double)L[o + 1], (double)L[o + 2]}; };
  const V3 n1i = qrotate(qi, lv(0)), n2i = qrotate(qi, lv(3)), n3i = qrotate(qi, lv(6));
  const V3 n1j = qrotate(qj, lv(12)), n2j = qrotate(qj, lv(15)), n3j = qrotate(qj, lv(18));
  const V3 ri = ldp(P.x, n, i), rj = ldp(P.x, n, j);
  const V3 rbi = vadd(qrotate(qi, lv(9)), ri), rbj = vadd(qrotate(qj, lv(21)), rj);
  const fdm_real* c = b.coef + 9 * (size_t)k;
  V3 F = {0., 0., 0.}, Mij = {0., 0., 0.}, Mji = {0., 0., 0.};
  double sig = b.maxNormalStress[k], tau = b.maxShearStress[k];
  const int on = vbond(F, Mij, Mji, sig, tau, rbi, rbj, n1i, n2i, n3i, n1j, n2j, n3j, c[0], c[1],
                       c[2], c[3], c[4], c[5], c[6], c[7], c[8]);
  b.active[k] = on;
  if (!on) b.breakStep[k] = b.stepIndex ? *b.stepIndex : 0;
  b.maxNormalStress[k] = sig;
  b.maxShearStress[k] = tau;
  b.applied[k] = 1;
  const V3 Ti = vadd(Mij, vcross(vsub(rbi, ri), F));
  const V3 Tj = vadd(Mji, vcross(vsub(rbj, rj), vneg(F)));
  double* of = b.outF + 3 * (size_t)k;
  double* oi = b.outTi + 3 * (size_t)k;
  double* oj = b.outTj + 3 * (size_t)k;
  of[0] = F.x; of[1] = F.y; of[2] = F.z;
  oi[0] = Ti.x; oi[1] = Ti.y; oi[2] = Ti.z;
  oj[0] = Tj.x; oj[1] = Tj.y; oj[2] = Tj.z;
}

struct ReduceK {
  fdm_bodies P;
  fdm_geometry pgeo;
  fdm_pairs pp, pw, g[3];
  fdm_mirror ppm, gm[3];
  int havePW, haveG[3];
  fdm_bonds bonds;
  int haveBonds;
  V3 gravity;
  double h;
  int kick;
  int* stepCounter;
  const int* gate;
};

__device__ __forceinline__ void addRec(V3& F, V3& T, const double* f, const double* t, bool negF) {
  const V3 fv = {f[0], f[1], f[2]};
  F = vadd(F, negF ? vneg(fv) : fv);
  T = vadd(T, {t[0], t[1], t[2]});
}

// Smallest mirror entry greater than `last` (mirror lists are short and
// unordered; selection keeps the fold order independent of append order).
__device__ __forceinline__ int nextMirror(const int* list, int m, int last) {
  int best = INT_MAX;
  for (int t = 0; t < m; ++t) {
    const int v = list[t];
    if (v > last && v < best) best = v;
  }
  return best;
}

__device__ __forceinline__ int segBegin(const fdm_pairs& s, int i) {
  return min(s.count[i], s.capacity);
}
__device__ __forceinline__ int segEnd(const fdm_pairs& s, int i) {
  return min(s.count[i + 1], s.capacity);
}

__global__ void kReduce(ReduceK a) {
  if (gateOff(a.gate)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = a.P.n;
  if (i == 0 && a.stepCounter) *a.stepCounter += 1;  // read by the bond stage before this launch
  if (i >= n) return;
  V3 F = {0., 0., 0.}, T = {0., 0., 0.};
  // particle pairs: as target (partners below i), then owned (partners above)
  {
    const int m = min(a.ppm.count[i], a.ppm.capacity);
    const int* list = a.ppm.list + (size_t)i * a.ppm.capacity;
    int last = -1;
    for (int r = 0; r < m; ++r) {
      const int p = nextMirror(list, m, last);
      last = p;
      addRec(F, T, a.pp.sumF + 3 * (size_t)p, a.pp.sumTB + 3 * (size_t)p, true);
    }
    if (m > 0) a.ppm.count[i] = 0;
    for (int p = segBegin(a.pp, i); p < segEnd(a.pp, i); ++p)
      if (a.pp.recCnt[p] > 0) addRec(F, T, a.pp.sumF + 3 * (size_t)p, a.pp.sumTA + 3 * (size_t)p, false);
  }
  if (a.havePW)
    for (int p = segBegin(a.pw, i); p < segEnd(a.pw, i); ++p)
      if (a.pw.recCnt[p] > 0) addRec(F, T, a.pw.sumF + 3 * (size_t)p, a.pw.sumTA + 3 * (size_t)p, false);
  for (int s = 0; s < 3; ++s) {
    if (!a.haveG[s]) continue;
    const fdm_pairs& G = a.g[s];
    const fdm_mirror& M = a.gm[s];
    const int m = min(M.count[i], M.capacity);
    const int* list = M.list + (size_t)i * M.capacity;
    int own = segBegin(G, i);
    const int ownEnd = segEnd(G, i);
    int mp = m > 0 ? nextMirror(list, m, -1) : INT_MAX;
    int used = 0;
    while (true) {
      while (own < ownEnd && G.recCnt[own] == 0) ++own;
      const bool haveOwn = own < ownEnd;
      const bool haveMir = used < m;
      if (!haveOwn && !haveMir) break;
      const int partnerOwn = haveOwn ? G.pairT[own] : INT_MAX;
      const int partnerMir = haveMir ? G.pairA[mp] : INT_MAX;
      if (haveOwn && partnerOwn <= partnerMir) {
        addRec(F, T, G.sumF + 3 * (size_t)own, G.sumTA + 3 * (size_t)own, false);
        ++own;
      } else {
        addRec(F, T, G.sumF + 3 * (size_t)mp, G.sumTB + 3 * (size_t)mp, true);
        ++used;
        mp = used < m ? nextMirror(list, m, mp) : INT_MAX;
      }
    }
    if (m > 0) M.count[i] = 0;
  }
  if (a.haveBonds) {
    const fdm_bonds& b = a.bonds;
    for (int e = b.partStart[i]; e < b.partStart[i + 1]; ++e) {
      const int code = b.partList[e];
      const int k = code >> 1;
      if (!b.applied[k]) continue;
      if ((code & 1) == 0) addRec(F, T, b.outF + 3 * (size_t)k, b.outTi + 3 * (size_t)k, false);
      else addRec(F, T, b.outF + 3 * (size_t)k, b.outTj + 3 * (size_t)k, true);
    }
  }
  st3(a.P.F, n, i, F);
  st3(a.P.T, n, i, T);
  if (!a.kick) return;
  // integration2ndHalf: the kick of particleIntegrationKernel.cu:3-32 with the
  // stored (FP32) force and torque
  const fdm_template& t = a.pgeo.info[a.P.geo[i]];
  const double invM = t.invMass;
  if (nearZero(invM)) return;
  const V3 Fs = roundState(F), Ts = roundState(T);
  const V3 v = ld3(a.P.v, n, i);
  st3(a.P.v, n, i, vadd(v, vmul(vadd(vmul(Fs, invM), a.gravity), a.h)));
  if (spinFixed(t)) return;
  const V3 w = ld3(a.P.w, n, i);
  const Q4 q = ld4(a.P.q, n, i);
  const V3 Iw = smul(srotate(q, lds6(t.inertia)), w);
  st3(a.P.w, n, i, vadd(w, vmul(smul(srotate(q, lds6(t.invInertia)), vsub(Ts, vcross(w, Iw))), a.h)));
}

}  // namespace fdm

using namespace fdm;

extern "C" int fdm_launchBondForces(fdm_bodies particles, fdm_bonds bonds, hipStream_t stream) {
  if (bonds.n < 0) return FDM_EINVAL;
  if (bonds.n == 0) return FDM_OK;
  hipLaunchKernelGGL(kBonds, dim3(blocksFor(bonds.n, 256)), dim3(256), 0, stream, particles, bonds);
  return launchStatus();
}

__global__ void kIncrement(int* c) {
  if (threadIdx.x == 0) *c += 1;
}

extern "C" int fdm_incrementCounter(int* counter, hipStream_t stream) {
  if (!counter) return FDM_EINVAL;
  hipLaunchKernelGGL(kIncrement, dim3(1), dim3(64), 0, stream, counter);
  return launchStatus();
}

extern "C" int fdm_launchReduceForceTorque(const fdm_reduce_args* args, hipStream_t stream) {
  if (!args || !args->pp || !args->ppMirror) return FDM_EINVAL;
  const fdm_reduce_args& g = *args;
  ReduceK k{};
  k.P = g.particles;
  k.pgeo = g.pgeo;
  k.pp = *g.pp;
  k.ppm = *g.ppMirror;
  k.havePW = g.pw != nullptr;
  if (g.pw) k.pw = *g.pw;
  for (int s = 0; s < 3; ++s) {
    k.haveG[s] = g.ghost[s] != nullptr && g.ghostMirror[s] != nullptr;
    if (k.haveG[s]) {
      k.g[s] = *g.ghost[s];
      k.gm[s] = *g.ghostMirror[s];
    }
  }
  k.haveBonds = g.bonds != nullptr && g.bonds->n > 0;
  if (k.haveBonds) k.bonds = *g.bonds;
  k.gravity = fromD3(g.gravity);
  k.h = g.h;
  k.kick = g.kick;
  k.stepCounter = g.stepCounter;
  k.gate = g.gate;
  if (g.particles.n <= 0) {
    if (g.stepCounter) hipLaunchKernelGGL(kIncrement, dim3(1), dim3(64), 0, stream, g.stepCounter);
    return launchStatus();
  }
  fdm_tuning tu = tuning();
  if (g.tuning) {
    tu = *g.tuning;
    if (fdm_checkTuning(&tu) != FDM_OK) return FDM_EINVAL;
  }
  const int rb = tu.reduceBlock;
  hipLaunchKernelGGL(kReduce, dim3(blocksFor(g.particles.n, rb)), dim3(rb), 0, stream, k);
  return launchStatus();
}
