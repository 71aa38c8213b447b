// Adjoint timestep: the contact tape of a checkpoint segment (append and
// restore) and one step in reverse. The reverse step is four kernels:
//   kAdjKick     thread per particle: reverse of the second half-kick, which
//                yields the adjoint of every particle's contact force and
//                torque sums;
//   kAdjContacts thread per node-contact: the hand-written VJP of the contact
//                geometry, the contact law and the torque arms (adjoint_math.hpp),
//                writing each contact's share for A and for its target and the
//                adjoint of its previous spring into the previous block;
//   kAdjGather   thread per particle: its shares in a fixed order (as A among
//                particle pairs, as A among wall pairs, as target by ascending
//                contact), then the reverse of the drift and the first half-kick;
//                the wall-material shares of its wall contacts are kept per
//                particle (at most two wall materials) and summed on the host.
// Every sum runs in a fixed order, so the adjoint is deterministic.
#include <algorithm>

#include "fdm_device.hpp"
#include "fdm_launch.hpp"
#include "adjoint_math.hpp"

namespace fdm {
namespace {

constexpr int kStride = 1024;  // blocks of the grid-stride loops over device counts
constexpr int kShare = 17;     // x 3, q 4, v 3, w 3, kn ks mu zeta
constexpr int kBucket = 64;    // particle-pair contacts per target particle and step
constexpr int kParts = 5;      // sets in tape order: particle, wall, ghost X, Y, XY

__device__ __forceinline__ adj::V toV(V3 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ adj::Q toQ(Q4 a) { return {a.q0, a.q1, a.q2, a.q3}; }
__device__ __forceinline__ adj::S toS(const double* p) { return {p[0], p[1], p[2], p[3], p[4], p[5]}; }
__device__ __forceinline__ adj::V ldV(const double* p, int n, int i) { return {p[i], p[n + i], p[2 * n + i]}; }
__device__ __forceinline__ adj::Q ldQ(const double* p, int n, int i) {
  return {p[i], p[n + i], p[2 * n + i], p[3 * n + i]};
}
__device__ __forceinline__ void stV(double* p, int n, int i, adj::V v) {
  p[i] = v.x;
  p[n + i] = v.y;
  p[2 * n + i] = v.z;
}
__device__ __forceinline__ void stQ(double* p, int n, int i, adj::Q q) {
  p[i] = q.q0;
  p[n + i] = q.q1;
  p[2 * n + i] = q.q2;
  p[3 * n + i] = q.q3;
}
// a vector as the state storage holds it
__device__ __forceinline__ adj::V stateRound(adj::V v) {
  return {roundState(v.x), roundState(v.y), roundState(v.z)};
}

// ------------------------------------------------------------ tape append
// The sets of a step in tape order: particle, wall, ghost X, Y, XY. An
// absent set has capacity 0.
struct TapeSets {
  fdm_pairs s[kParts];
  int capOff[kParts + 1];  // prefix of the pair capacities
};

__device__ __forceinline__ int setTotal(const TapeSets& S, int p) {
  return S.capOff[p + 1] > S.capOff[p] ? min(*S.s[p].total, S.s[p].capacity) : 0;
}

// per-pair contact counts of all sets, set after set
__global__ void kTapeCount(TapeSets S, int* cnt) {
  int tot[kParts];
#pragma unroll
  for (int p = 0; p < kParts; ++p) tot[p] = setTotal(S, p);
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < S.capOff[kParts]; g += gridDim.x * blockDim.x) {
    int p = 0;
    while (g >= S.capOff[p + 1]) ++p;
    const int q = g - S.capOff[p];
    cnt[g] = q < tot[p] ? S.s[p].recCnt[q] : 0;
  }
}

__global__ void kTapeBlock(fdm_contact_tape t, int block, TapeSets S, const int* total, const int* off,
                           int maxBlock) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int all = *total;
  int* part = t.blockPart + (size_t)(kParts + 1) * block;
  t.blockStart[block] = block * maxBlock;
  if (all > maxBlock) {
    atomicOr(t.status, FDM_STATUS_RECORD_OVERFLOW);
    for (int p = 0; p <= kParts; ++p) part[p] = 0;
    t.blockCount[block] = 0;
    return;
  }
  for (int p = 0; p < kParts; ++p) part[p] = S.capOff[p] < S.capOff[kParts] ? off[S.capOff[p]] : all;
  part[kParts] = all;
  t.blockCount[block] = all;
}

__device__ __forceinline__ bool keyLess(const fdm_contact_tape& t, int c, int A, int T, int k) {
  if (t.A[c] != A) return t.A[c] < A;
  if (t.T[c] != T) return t.T[c] < T;
  return t.node[c] < k;
}

// the same (A, target, node) in the previous block's part of the same set
__device__ __forceinline__ int prevOf(const fdm_contact_tape& t, int prevBlock, int set, int A, int T,
                                      int k) {
  if (prevBlock < 0) return -1;
  const int ps = t.blockStart[prevBlock];
  const int* part = t.blockPart + (size_t)(kParts + 1) * prevBlock;
  int lo = ps + part[set];
  const int end = ps + part[set + 1];
  int hi = end;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (keyLess(t, mid, A, T, k)) lo = mid + 1;
    else hi = mid;
  }
  return (lo < end && t.A[lo] == A && t.T[lo] == T && t.node[lo] == k) ? lo : -1;
}

// contacts and their history links (thread per pair)
__global__ void kTapeScatter(TapeSets S, const int* off, const int* total, fdm_contact_tape t, int block,
                             int prevBlock) {
  if (t.blockCount[block] != *total) return;  // block did not fit
  const int start = t.blockStart[block];
  int tot[kParts], pre[kParts + 1];
  pre[0] = 0;
#pragma unroll
  for (int p = 0; p < kParts; ++p) {
    tot[p] = setTotal(S, p);
    pre[p + 1] = pre[p] + tot[p];
  }
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < pre[kParts]; g += gridDim.x * blockDim.x) {
    int set = 0;
    while (g >= pre[set + 1]) ++set;
    const fdm_pairs& s = S.s[set];
    const int p = g - pre[set];
    const int n = s.recCnt[p];
    if (n <= 0) continue;
    const int i0 = off[S.capOff[set] + p], A = s.pairA[p], T = s.pairT[p], r0 = s.recOff[p];
    for (int r = 0; r < n; ++r) {
      const fdm_record rec = s.rec[r0 + r];
      const int c = start + i0 + r, k = rec.node;
      t.A[c] = A;
      t.T[c] = T;
      t.node[c] = k;
      t.spring[3 * (size_t)c] = rec.x;
      t.spring[3 * (size_t)c + 1] = rec.y;
      t.spring[3 * (size_t)c + 2] = rec.z;
      t.prev[c] = prevOf(t, prevBlock, set, A, T, k);
    }
  }
}

// ----------------------------------------------------------- tape restore
// pair starts of a block part; zeros up to maxBlock so that the scan total counts pairs
__global__ void kRestoreFlags(fdm_contact_tape t, int block, int set, int* flag, int maxBlock) {
  const int* part = t.blockPart + (size_t)(kParts + 1) * block;
  const int start = t.blockStart[block] + part[set], count = part[set + 1] - part[set];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < maxBlock; i += gridDim.x * blockDim.x) {
    const int c = start + i;
    flag[i] = i < count && (i == 0 || t.A[c] != t.A[c - 1] || t.T[c] != t.T[c - 1]) ? 1 : 0;
  }
}

__global__ void kRestoreScatter(fdm_contact_tape t, int block, int set, const int* flag, const int* idx,
                                const int* numPairs, fdm_pairs s) {
  const int* part = t.blockPart + (size_t)(kParts + 1) * block;
  const int start = t.blockStart[block] + part[set], count = part[set + 1] - part[set];
  if (*numPairs > s.capacity || count > s.recCapacity) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(t.status, FDM_STATUS_PAIR_OVERFLOW);
    return;
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const int c = start + i;
    s.rec[i] = makeRecord({(double)t.spring[3 * c], (double)t.spring[3 * c + 1],
                           (double)t.spring[3 * c + 2]}, t.node[c]);
    if (!flag[i]) continue;
    int j = i + 1;
    while (j < count && t.A[start + j] == t.A[c] && t.T[start + j] == t.T[c]) ++j;
    const int p = idx[i];
    s.pairA[p] = t.A[c];
    s.pairT[p] = t.T[c];
    s.pairG[p] = 0;
    s.recOff[p] = i;
    s.recCnt[p] = j - i;
  }
}

// CSR starts by A over the restored pairs
__global__ void kRestoreCount(fdm_pairs s, int nParticles, const int* numPairs) {
  const int P = min(*numPairs, s.capacity);
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a <= nParticles; a += gridDim.x * blockDim.x) {
    int lo = 0, hi = P;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s.pairA[mid] < a) lo = mid + 1;
      else hi = mid;
    }
    s.count[a] = lo;
    if (a == 0) *s.total = P;
  }
}

// --------------------------------------------------------- reverse step
struct AdjWork {
  double* vh;      // [3n] v after the first half-kick
  double* wh;      // [3n]
  double* Fc;      // [3n] adjoint of the contact force sum
  double* Tc;      // [3n]
  double* shareA;  // [kShare * maxBlock]
  double* shareB;  // [kShare * maxBlock]
  // half-open block-local ranges per particle and set, written by
  // kAdjContacts at run boundaries (empty unless written): contacts with
  // A = i in each set
  int* run;        // [kParts * 2n]
  // contacts whose target is particle i, from every particle-target set (unordered)
  int* bucketCount;  // [n]
  int* bucket;       // [n * kBucket]
  unsigned* status;
};

__device__ __forceinline__ void kickPrimal(const fdm_adjoint_step_args& a, const fdm_bodies& b,
                                           const fdm_template& t, int i, adj::V& v, adj::V& w) {
  const int n = b.n;
  v = toV(ld3(b.v, n, i));
  w = toV(ld3(b.w, n, i));
  if (nearZero(t.invMass)) return;
  const adj::V F = toV(ld3(b.F, n, i)), T = toV(ld3(b.T, n, i));
  const adj::V g = {a.gravity.x, a.gravity.y, a.gravity.z};
  const double h = 0.5 * a.dt;
  v = stateRound(adj::add(v, adj::mul(adj::add(adj::mul(F, t.invMass), g), h)));
  w = stateRound(adj::spin(w, toQ(ld4(b.q, n, i)), T, toS(t.inertia), toS(t.invInertia), h));
}

// reverse of the second half-kick, v' = vh + h (F' invM + g), w' = spin(wh, q', T')
__global__ void __launch_bounds__(256) kAdjKick(fdm_adjoint_step_args a, AdjWork k) {
  const int n = a.before.n;
  const double h = 0.5 * a.dt;
  if (a.prevBlock >= 0) {  // the previous block's spring adjoints are written by kAdjContacts
    const size_t s0 = 3 * (size_t)a.tape.blockStart[a.prevBlock];
    const int m = 3 * a.tape.blockCount[a.prevBlock];
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x)
      a.springBar[s0 + j] = 0.;
  }
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < kParts * 2 * n; j += gridDim.x * blockDim.x) {
    k.run[j] = 0;
    if (j < n) k.bucketCount[j] = 0;
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const fdm_template& t = a.pgeo.info[a.before.geo[i]];
    adj::V vh, wh;
    kickPrimal(a, a.before, t, i, vh, wh);
    stV(k.vh, n, i, vh);
    stV(k.wh, n, i, wh);
    const adj::V vb = ldV(a.vBar, n, i), wb = ldV(a.wBar, n, i);
    adj::V Fc = ldV(a.FBar, n, i), Tc = ldV(a.TBar, n, i);
    if (!nearZero(t.invMass)) {
      adj::acc(Fc, adj::mul(vb, h * t.invMass));
      adj::V whb = {0., 0., 0.}, tqb = {0., 0., 0.};
      adj::Q qb = ldQ(a.qBar, n, i);
      adj::spinVjp(wh, toQ(ld4(a.after.q, n, i)), toV(ld3(a.after.T, n, i)), toS(t.inertia),
                   toS(t.invInertia), h, wb, whb, qb, tqb);
      adj::acc(Tc, tqb);
      stQ(a.qBar, n, i, qb);
      stV(a.wBar, n, i, whb);
    }
    stV(k.Fc, n, i, Fc);
    stV(k.Tc, n, i, Tc);
  }
}

__device__ __forceinline__ void putShare(double* s, const adj::BodyBar& b, double kn, double ks,
                                         double mu, double zeta) {
  s[0] = b.x.x; s[1] = b.x.y; s[2] = b.x.z;
  s[3] = b.q.q0; s[4] = b.q.q1; s[5] = b.q.q2; s[6] = b.q.q3;
  s[7] = b.v.x; s[8] = b.v.y; s[9] = b.v.z;
  s[10] = b.w.x; s[11] = b.w.y; s[12] = b.w.z;
  s[13] = kn; s[14] = ks; s[15] = mu; s[16] = zeta;
}

__global__ void __launch_bounds__(256) kAdjContacts(fdm_adjoint_step_args a, AdjWork k) {
  const fdm_contact_tape& tp = a.tape;
  const int start = tp.blockStart[a.block], count = tp.blockCount[a.block];
  const int* part = tp.blockPart + (size_t)(kParts + 1) * a.block;
  const int n = a.before.n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const int c = start + i;
    int set = 0;
    while (i >= part[set + 1]) ++set;
    const bool wall = set == FDM_SET_WALL;
    const int A = tp.A[c], T = tp.T[c], node = tp.node[c];
    {
      // run boundaries by A within the set, and the target's bucket
      int* run = k.run + (size_t)set * 2 * n;
      if (i == part[set] || tp.A[c - 1] != A) run[2 * A] = i;
      if (i == part[set + 1] - 1 || tp.A[c + 1] != A) run[2 * A + 1] = i + 1;
      if (!wall) {
        const int pos = atomicAdd(k.bucketCount + T, 1);
        if (pos < kBucket) k.bucket[(size_t)T * kBucket + pos] = i;
        else atomicOr(k.status, FDM_STATUS_BUCKET_OVERFLOW);
      }
    }
    const fdm_template& tA = a.pgeo.info[a.after.geo[A]];
    const fdm_material& Mi = a.pmat[a.after.mat[A]];
    const adj::V xA = toV(ldp(a.after.x, n, A));
    const adj::Q qA = toQ(ld4(a.after.q, n, A));
    const adj::V vA = ldV(k.vh, n, A), wA = ldV(k.wh, n, A);
    adj::V xB, vB, wB;
    adj::Q qB;
    const fdm_template* tB;
    const float* sdf;
    adj::LawConst lc;
    lc.dt = a.dt;
    double shKnA = 0., shKnB = 0., shKsA = 0., shKsB = 0.;
    bool muToA = true, zetaToA = true;
    if (wall) {
      const int m = a.walls.n;
      xB = toV(ldp(a.walls.x, m, T));
      qB = toQ(ld4(a.walls.q, m, T));
      vB = toV(ld3(a.walls.v, m, T));
      wB = toV(ld3(a.walls.w, m, T));
      tB = &a.wgeo.info[a.walls.geo[T]];
      sdf = a.wgeo.sdf;
      const fdm_material& Mw = a.wmat[a.walls.mat[T]];
      lc.effM = tA.invMass > 0. ? 1. / tA.invMass : 0.;
      lc.mu = fmin(Mi.mu, Mw.mu);
      lc.zeta = Mi.e <= Mw.e ? Mi.dissipation : Mw.dissipation;
      lc.kn = Mi.kn;
      lc.ks = Mi.ks;
      shKnA = 1.;
      shKsA = 1.;
      muToA = Mi.mu <= Mw.mu;
      zetaToA = Mi.e <= Mw.e;
    } else {
      xB = toV(ldp(a.after.x, n, T));
      if (set >= FDM_SET_GHOST) {
        // the ghost: the real particle shifted by its set's periods
        const fdm_grid& g = a.inst.count > 0 ? a.inst.pgrid[a.inst.particleInstance[T]] : a.pgrid;
        const int mask = set - FDM_SET_GHOST + 1;
        if (mask & 1) xB.x += g.max.x - g.min.x;
        if (mask & 2) xB.y += g.max.y - g.min.y;
      }
      qB = toQ(ld4(a.after.q, n, T));
      vB = ldV(k.vh, n, T);
      wB = ldV(k.wh, n, T);
      tB = &a.pgeo.info[a.after.geo[T]];
      sdf = a.pgeo.sdf;
      const fdm_material& Mj = a.pmat[a.after.mat[T]];
      const double invMj = tB->invMass;
      lc.kn = 0.;
      lc.ks = 0.;
      if (Mi.kn > 0. && Mj.kn > 0.) {
        lc.kn = Mi.kn * Mj.kn / (Mi.kn + Mj.kn);
        const double s2 = (Mi.kn + Mj.kn) * (Mi.kn + Mj.kn);
        shKnA = Mj.kn * Mj.kn / s2;
        shKnB = Mi.kn * Mi.kn / s2;
      }
      if (Mi.ks > 0. && Mj.ks > 0.) {
        lc.ks = Mi.ks * Mj.ks / (Mi.ks + Mj.ks);
        const double s2 = (Mi.ks + Mj.ks) * (Mi.ks + Mj.ks);
        shKsA = Mj.ks * Mj.ks / s2;
        shKsB = Mi.ks * Mi.ks / s2;
      }
      lc.effM = (tA.invMass > 0. || invMj > 0.) ? 1. / (tA.invMass + invMj) : 0.;
      lc.zeta = Mi.e <= Mj.e ? Mi.dissipation : Mj.dissipation;
      lc.mu = fmin(Mi.mu, Mj.mu);
      muToA = Mi.mu <= Mj.mu;
      zetaToA = Mi.e <= Mj.e;
    }
    // the node's SDF cell at the linearisation point
    adj::GeoConst g;
    const fdm_real* nd = a.pgeo.nodes + 3 * ((size_t)tA.nodeStart + node);
    g.s = {(double)nd[0], (double)nd[1], (double)nd[2]};
    g.origin = {tB->origin[0], tB->origin[1], tB->origin[2]};
    g.ih = tB->invSpacing;
    const adj::V p = adj::add(adj::qrot(qA, g.s), xA);
    const adj::V l = adj::qrotT(adj::sub(p, xB), qB);
    const int ci = toInt(floor((l.x - g.origin.x) * g.ih));
    const int cj = toInt(floor((l.y - g.origin.y) * g.ih));
    const int ck = toInt(floor((l.z - g.origin.z) * g.ih));
    adj::BodyBar Ab = {{0., 0., 0.}, {0., 0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    adj::BodyBar Bb = Ab;
    adj::V spb = {0., 0., 0.};
    adj::LawBar pb = {0., 0., 0., 0.};
    const int pc = tp.prev[c];
    if (ci >= 0 && cj >= 0 && ck >= 0 && ci < tB->size[0] - 1 && cj < tB->size[1] - 1 &&
        ck < tB->size[2] - 1) {
      g.cell[0] = ci;
      g.cell[1] = cj;
      g.cell[2] = ck;
      const int nx = tB->size[0], nxy = tB->size[0] * tB->size[1];
      const float* b = sdf + tB->sdfStart + ((size_t)ck * nxy + (size_t)cj * nx + ci);
      g.c[0] = b[0];
      g.c[1] = b[1];
      g.c[2] = b[nx];
      g.c[3] = b[nx + 1];
      g.c[4] = b[nxy];
      g.c[5] = b[nxy + 1];
      g.c[6] = b[nxy + nx];
      g.c[7] = b[nxy + nx + 1];
      const adj::V sp = pc >= 0 ? adj::V{(double)tp.spring[3 * pc], (double)tp.spring[3 * pc + 1],
                                         (double)tp.spring[3 * pc + 2]}
                                : adj::V{0., 0., 0.};
      adj::V Fb = ldV(k.Fc, n, A);
      adj::V TBb = {0., 0., 0.};
      if (!wall) {
        Fb = adj::sub(Fb, ldV(k.Fc, n, T));
        TBb = ldV(k.Tc, n, T);
      }
      const adj::V sb = {a.springBar[3 * (size_t)c], a.springBar[3 * (size_t)c + 1],
                         a.springBar[3 * (size_t)c + 2]};
      adj::contactVjp(xA, qA, vA, wA, xB, qB, vB, wB, sp, g, lc, Fb, ldV(k.Tc, n, A), TBb, sb, Ab,
                      Bb, spb, pb);
    }
    if (pc >= 0) {
      a.springBar[3 * (size_t)pc] = spb.x;
      a.springBar[3 * (size_t)pc + 1] = spb.y;
      a.springBar[3 * (size_t)pc + 2] = spb.z;
    }
    putShare(k.shareA + (size_t)kShare * i, Ab, pb.kn * shKnA, pb.ks * shKsA, muToA ? pb.mu : 0.,
             zetaToA ? pb.zeta : 0.);
    putShare(k.shareB + (size_t)kShare * i, Bb, pb.kn * shKnB, pb.ks * shKsB, muToA ? 0. : pb.mu,
             zetaToA ? 0. : pb.zeta);
  }
}

__device__ __forceinline__ void addShare(double* acc, const double* s) {
#pragma unroll
  for (int j = 0; j < kShare; ++j) acc[j] += s[j];
}

__global__ void __launch_bounds__(256) kAdjGather(fdm_adjoint_step_args a, AdjWork k) {
  const fdm_contact_tape& tp = a.tape;
  const int start = tp.blockStart[a.block];
  const int n = a.before.n;
  const double dt = a.dt, h = 0.5 * a.dt;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s[kShare];
#pragma unroll
    for (int j = 0; j < kShare; ++j) s[j] = 0.;
    for (int set = 0; set < kParts; ++set) {
      const int* run = k.run + (size_t)set * 2 * n + 2 * i;
      for (int c = run[0]; c < run[1]; ++c) addShare(s, k.shareA + (size_t)kShare * c);
    }
    {
      // as target: ascending contact order, selected from the bucket
      const int m = min(k.bucketCount[i], kBucket);
      const int* bk = k.bucket + (size_t)i * kBucket;
      int last = -1;
      for (int r = 0; r < m; ++r) {
        int next = INT_MAX;
        for (int u = 0; u < m; ++u) {
          const int c = bk[u];
          if (c > last && c < next) next = c;
        }
        last = next;
        addShare(s, k.shareB + (size_t)kShare * next);
      }
    }
    // wall-material shares of the wall contacts, per wall material
    const int* wrun = k.run + (size_t)FDM_SET_WALL * 2 * n + 2 * i;
    for (int c = wrun[0]; c < wrun[1]; ++c) {
      const double* sb = k.shareB + (size_t)kShare * c;
      if (sb[15] == 0. && sb[16] == 0.) continue;
      const int wm = a.walls.mat[tp.T[start + c]];
      int slot = -1;
      for (int u = 0; u < 2 && slot < 0; ++u) {
        if (a.wallMat[2 * i + u] == wm) slot = u;
        else if (a.wallMat[2 * i + u] < 0) {
          a.wallMat[2 * i + u] = wm;
          slot = u;
        }
      }
      if (slot < 0) {
        atomicOr(k.status, FDM_STATUS_BUCKET_OVERFLOW);
        continue;
      }
      a.wallShare[4 * i + 2 * slot] += sb[15];
      a.wallShare[4 * i + 2 * slot + 1] += sb[16];
    }
    double* pbar = a.paramBar + 4 * (size_t)i;
    pbar[0] += s[13];
    pbar[1] += s[14];
    pbar[2] += s[15];
    pbar[3] += s[16];
    // adjoint of the state after the drift
    adj::V xb = ldV(a.xBar, n, i);
    adj::Q qb = ldQ(a.qBar, n, i);
    adj::V vhb = ldV(a.vBar, n, i), whb = ldV(a.wBar, n, i);
    adj::acc(xb, {s[0], s[1], s[2]});
    adj::accq(qb, {s[3], s[4], s[5], s[6]});
    adj::acc(vhb, {s[7], s[8], s[9]});
    adj::acc(whb, {s[10], s[11], s[12]});
    // drift: x' = x + dt vh, q' = qint(q, wh, dt)
    adj::acc(vhb, adj::mul(xb, dt));
    const adj::Q q = toQ(ld4(a.before.q, n, i));
    adj::Q q0b = {0., 0., 0., 0.};
    adj::qintVjp(q, ldV(k.wh, n, i), dt, qb, q0b, whb);
    // first half-kick
    const fdm_template& t = a.pgeo.info[a.before.geo[i]];
    adj::V wb = whb, Fb = {0., 0., 0.}, Tb = {0., 0., 0.};
    if (!nearZero(t.invMass)) {
      Fb = adj::mul(vhb, h * t.invMass);
      wb = {0., 0., 0.};
      adj::spinVjp(toV(ld3(a.before.w, n, i)), q, toV(ld3(a.before.T, n, i)), toS(t.inertia),
                   toS(t.invInertia), h, whb, wb, q0b, Tb);
    }
    stV(a.xBar, n, i, xb);
    stQ(a.qBar, n, i, q0b);
    stV(a.vBar, n, i, vhb);
    stV(a.wBar, n, i, wb);
    stV(a.FBar, n, i, Fb);
    stV(a.TBar, n, i, Tb);
  }
}

// state copy of a set of bodies
__global__ void kCopyBodies(fdm_bodies dst, fdm_bodies src) {
  const int n = src.n;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < 4 * n; j += gridDim.x * blockDim.x) {
    dst.q[j] = src.q[j];
    if (j < 3 * n) {
      dst.x[j] = src.x[j];
      dst.v[j] = src.v[j];
      dst.w[j] = src.w[j];
      dst.F[j] = src.F[j];
      dst.T[j] = src.T[j];
    }
  }
}

// workspace carving
size_t up(size_t b) { return (b + 255) & ~size_t(255); }

struct TapeWs {
  int *cnt, *off, *total, *flag, *idx;
  void* sub;
  size_t subBytes;
};
size_t tapeBytes(int maxBlock, long long pairCapacity, TapeWs* w, char* base) {
  const size_t m = std::max(maxBlock, 64);
  const size_t pairs = std::max<size_t>((size_t)pairCapacity, 64);
  size_t o = 0;
  auto take = [&](size_t bytes) {
    char* p = base ? base + o : nullptr;
    o += up(bytes);
    return p;
  };
  int* cnt = (int*)take(sizeof(int) * pairs);
  int* off = (int*)take(sizeof(int) * pairs);
  int* total = (int*)take(sizeof(int) * 16);
  int* flag = (int*)take(sizeof(int) * m);
  int* idx = (int*)take(sizeof(int) * m);
  const size_t sb = scanWorkspace(std::max(pairs, m));
  void* sub = take(sb);
  if (w) *w = {cnt, off, total, flag, idx, sub, sb};
  return o;
}

bool makeSets(const fdm_pairs* const* sets, TapeSets& S, long long& cap) {
  if (!sets || !sets[0]) return false;
  S.capOff[0] = 0;
  cap = 0;
  for (int p = 0; p < kParts; ++p) {
    S.s[p] = sets[p] ? *sets[p] : *sets[0];
    const int c = sets[p] ? sets[p]->capacity : 0;
    if (c < 0) return false;
    cap += c;
    if (cap >= (1ll << 31)) return false;
    S.capOff[p + 1] = (int)cap;
  }
  return true;
}

}  // namespace
}  // namespace fdm

using namespace fdm;

extern "C" size_t fdm_tapeWorkspaceBytes(int maxBlockContacts, long long pairCapacity) {
  return tapeBytes(maxBlockContacts, pairCapacity, nullptr, nullptr);
}

extern "C" int fdm_launchTapeAppend(const fdm_pairs* const* sets, fdm_contact_tape tape, int block,
                                    int prevBlock, int maxBlockContacts, void* ws, size_t wsBytes,
                                    hipStream_t stream) {
  TapeSets S;
  long long cap = 0;
  if (!makeSets(sets, S, cap) || block < 0 || block >= tape.maxBlocks || prevBlock >= block ||
      maxBlockContacts < 1 || (long long)tape.maxBlocks * maxBlockContacts > (long long)tape.capacity)
    return FDM_EINVAL;
  TapeWs w;
  if (wsBytes < tapeBytes(maxBlockContacts, cap, &w, (char*)ws)) return FDM_EINVAL;
  hipLaunchKernelGGL(kTapeCount, dim3(std::max(1u, std::min(blocksFor(cap, 256), 4096u))), dim3(256), 0,
                     stream, S, w.cnt);
  int r = scanInts(w.cnt, w.off, (size_t)std::max(cap, 1ll), false, w.total, w.sub, w.subBytes, stream);
  if (r != FDM_OK) return r;
  hipLaunchKernelGGL(kTapeBlock, dim3(1), dim3(64), 0, stream, tape, block, S, (const int*)w.total,
                     (const int*)w.off, maxBlockContacts);
  hipLaunchKernelGGL(kTapeScatter, dim3(kStride), dim3(256), 0, stream, S, (const int*)w.off,
                     (const int*)w.total, tape, block, prevBlock);
  return launchStatus();
}

extern "C" int fdm_launchTapeRestore(fdm_contact_tape tape, int block, const fdm_pairs* const* sets,
                                     int nParticles, int maxBlockContacts, void* ws, size_t wsBytes,
                                     hipStream_t stream) {
  TapeSets S;
  long long cap = 0;
  if (!makeSets(sets, S, cap) || block < 0 || block >= tape.maxBlocks || nParticles < 0 ||
      maxBlockContacts < 1)
    return FDM_EINVAL;
  TapeWs w;
  if (wsBytes < tapeBytes(maxBlockContacts, cap, &w, (char*)ws)) return FDM_EINVAL;
  for (int set = 0; set < kParts; ++set) {
    if (!sets[set]) continue;
    hipLaunchKernelGGL(kRestoreFlags, dim3(kStride), dim3(256), 0, stream, tape, block, set, w.flag,
                       maxBlockContacts);
    int r = scanInts(w.flag, w.idx, (size_t)maxBlockContacts, false, w.total + 2 + set, w.sub,
                     w.subBytes, stream);
    if (r != FDM_OK) return r;
    hipLaunchKernelGGL(kRestoreScatter, dim3(kStride), dim3(256), 0, stream, tape, block, set,
                       (const int*)w.flag, (const int*)w.idx, (const int*)(w.total + 2 + set), *sets[set]);
    hipLaunchKernelGGL(kRestoreCount, dim3(std::min(blocksFor((long long)nParticles + 1, 256), 4096u)),
                       dim3(256), 0, stream, *sets[set], nParticles, (const int*)(w.total + 2 + set));
  }
  return launchStatus();
}

extern "C" size_t fdm_adjointWorkspaceBytes(int nParticles, int maxBlockContacts) {
  const size_t n = std::max(nParticles, 1), m = std::max(maxBlockContacts, 1);
  return up(4 * 3 * n * sizeof(double)) + 2 * up(kShare * m * sizeof(double)) +
         up(kParts * 2 * n * sizeof(int)) + up(n * sizeof(int)) + up(n * kBucket * sizeof(int));
}

extern "C" int fdm_launchAdjointStep(const fdm_adjoint_step_args* args, void* ws, size_t wsBytes,
                                     hipStream_t stream) {
  if (!args || args->block < 1 || args->block >= args->tape.maxBlocks ||
      args->prevBlock >= args->block || args->before.n != args->after.n || !args->xBar ||
      !args->qBar || !args->vBar || !args->wBar || !args->FBar || !args->TBar ||
      !args->springBar || !args->paramBar || !args->wallShare || !args->wallMat || !args->status)
    return FDM_EINVAL;
  const int n = args->before.n;
  if (wsBytes < fdm_adjointWorkspaceBytes(n, args->maxBlockContacts)) return FDM_EINVAL;
  char* b = (char*)ws;
  AdjWork k;
  const size_t v3 = 3 * (size_t)std::max(n, 1) * sizeof(double);
  k.vh = (double*)b;
  k.wh = (double*)(b + v3);
  k.Fc = (double*)(b + 2 * v3);
  k.Tc = (double*)(b + 3 * v3);
  b += up(4 * v3);
  k.shareA = (double*)b;
  b += up(kShare * (size_t)std::max(args->maxBlockContacts, 1) * sizeof(double));
  k.shareB = (double*)b;
  b += up(kShare * (size_t)std::max(args->maxBlockContacts, 1) * sizeof(double));
  k.run = (int*)b;
  b += up(kParts * 2 * (size_t)std::max(n, 1) * sizeof(int));
  k.bucketCount = (int*)b;
  b += up((size_t)std::max(n, 1) * sizeof(int));
  k.bucket = (int*)b;
  k.status = args->status;
  const fdm_adjoint_step_args a = *args;
  const unsigned pb = std::max(1u, std::min(blocksFor(n, 256), 4096u));
  hipLaunchKernelGGL(kAdjKick, dim3(std::max(pb, 64u)), dim3(256), 0, stream, a, k);
  hipLaunchKernelGGL(kAdjContacts, dim3(kStride), dim3(256), 0, stream, a, k);
  hipLaunchKernelGGL(kAdjGather, dim3(pb), dim3(256), 0, stream, a, k);
  return launchStatus();
}

extern "C" int fdm_launchCopyBodies(fdm_bodies dst, fdm_bodies src, hipStream_t stream) {
  if (src.n != dst.n || src.n < 0) return FDM_EINVAL;
  if (src.n == 0) return FDM_OK;
  hipLaunchKernelGGL(kCopyBodies, dim3(std::min(blocksFor(4ll * src.n, 256), 4096u)), dim3(256), 0,
                     stream, dst, src);
  return launchStatus();
}
