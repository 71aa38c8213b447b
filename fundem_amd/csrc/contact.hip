// Contact detection and contact forces.
//
// The reference (LSParticleContactDetectionKernel.cu:4-607) runs one thread
// per surface node over the 27-cell stencil, twice (count, then write), for
// every interaction set, then one thread per contact for forces. Here:
//
//  broad phase  one thread per particle A lists the targets that any of A's
//               nodes could touch under the reference's rules, sorted by
//               target. Particle targets need |x_A - x_B| <= R_A + R_B + s,
//               ghost and wall targets need |x_A - x_T| <= R_A + (SDF box
//               reach of T) + s. A pair is dropped when the target's
//               interpolated SDF at A's centre exceeds Lip * (R_A + s) (Lip
//               bounds the gradient of the trilinear interpolant), since
//               then no node of A can reach phi <= 0. The skin s lets a list
//               serve later steps while bodies move less than s/2 (the step
//               decides on the device when to rebuild); s = 0 gives lists
//               for the current positions only.
//  narrow phase one wavefront per candidate pair. A's nodes are swept in
//               chunks of 64 (one node per lane). A chunk is skipped when its
//               bounding sphere cannot satisfy the node test (the same three
//               bounds at chunk level: sphere, stencil, SDF box, Lipschitz).
//               Each remaining node evaluates the reference's predicate
//               exactly in FP64: stencil membership of the target's cell,
//               bounding sphere (particle targets, lower index only), SDF box
//               and -phi >= 0. The wave counts its contacts, reserves that
//               many history records, then recomputes each contact's normal
//               and point, takes its previous spring (keyed by (A, target,
//               node)), applies the linear spring-dashpot law and writes the
//               updated spring. Forces and torques are summed per pair in
//               node order, serially, so the result does not depend on the
//               schedule.
//
// All culls are conservative with a relative margin of 1e-5 on reaches (a
// unit quaternion stored in FP32 scales vectors by up to 1 +- 2.4e-7), so
// they never remove a contact that the exact test would accept.
#include <algorithm>

#include "fdm_device.hpp"
#include "fdm_launch.hpp"

namespace fdm {

constexpr double kReach = 1.0 + 1e-5;

struct Target {
  V3 x;       // detection position
  Q4 q;
  bool ident;  // q is exactly (1, 0, 0, 0): the rotations return their input
  fdm_template t;
  const float* sdf;
  int hash;   // hash value in the target grid
};
// Into and out of the target's frame. A rotation by the exact identity
// quaternion returns its input (up to the sign of a zero component), so it
// is skipped; the test is wave-uniform in the sweep.
__device__ __forceinline__ V3 toLocal(V3 v, const Target& tg) { return tg.ident ? v : qrotateT(v, tg.q); }
__device__ __forceinline__ V3 toWorld(V3 v, const Target& tg) { return tg.ident ? v : qrotate(tg.q, v); }


// Every kernel of this file only reads body state, templates, materials,
// candidate lists and the previous step's records: they are read through the
// constant address space (ldu), so that the wave-uniform reads of the sweep
// (one pair per wave) become scalar loads into SGPRs, issued together,
// instead of vector loads that each hold VGPRs and wait in turn.
__device__ __forceinline__ fdm_template tmplOf(const fdm_geometry& g, int geo) {
  return ldu(g.info + geo);
}
__device__ __forceinline__ fdm_material matOf(const fdm_material* m, int k) { return ldu(m + k); }
__device__ __forceinline__ V3 ldpR(const double* p, int n, int i) {
  return {ldu(p + i), ldu(p + n + i), ldu(p + 2 * n + i)};
}
__device__ __forceinline__ V3 ld3R(const fdm_real* p, int n, int i) {
  return {(double)ldu(p + i), (double)ldu(p + n + i), (double)ldu(p + 2 * n + i)};
}
__device__ __forceinline__ Q4 ld4R(const fdm_real* p, int n, int i) {
  return {(double)ldu(p + i), (double)ldu(p + n + i), (double)ldu(p + 2 * n + i),
          (double)ldu(p + 3 * n + i)};
}
__device__ __forceinline__ V3 ldd3R(const double* p) { return {ldu(p), ldu(p + 1), ldu(p + 2)}; }

// True when no point within `reach` of `local` (target frame) can have
// phi <= 0: the point is inside the SDF box and phi - Lip*reach is positive
// beyond a relative margin.
__device__ __forceinline__ bool lipschitzClear(V3 local, double reach, const fdm_template& t,
                                               const float* sdf) {
  Sample s;
  if (!sampleSDF(local, t, sdf, s)) return false;
  const double phi = trilinear(s.x, s.y, s.z, s.c);
  const double lr = t.lipschitz * reach;
  return phi - lr > 1e-6 * (fabs(phi) + lr);
}

// True when phi > 0 at `local` for certain on an affine (plane-wall) SDF:
// the plane's value exceeds the fit's bound on the residual and rounding.
// Such a node cannot touch, so its SDF fetch is skipped.
__device__ __forceinline__ bool planeClear(V3 local, const fdm_template& t) {
  if (!t.affine) return false;
  const double v = t.plane[3] + t.plane[0] * local.x + t.plane[1] * local.y + t.plane[2] * local.z;
  return v > t.planeSlack;
}
// The same for every point within `reach` of `local`: the plane changes by at
// most |a|_1 |d| over an offset d (|a . d| <= sum |a_i| |d_i| <= |a|_1 |d|).
// For an affine SDF this replaces the Lipschitz test without an SDF fetch.
__device__ __forceinline__ bool planeClearBall(V3 local, double reach, const fdm_template& t) {
  const double v = t.plane[3] + t.plane[0] * local.x + t.plane[1] * local.y + t.plane[2] * local.z;
  const double a1 = fabs(t.plane[0]) + fabs(t.plane[1]) + fabs(t.plane[2]);
  return v - a1 * reach > t.planeSlack;
}
// Ball test against the zero level: the plane bound for affine SDFs, the
// Lipschitz bound otherwise.
__device__ __forceinline__ bool levelClear(V3 local, double reach, const fdm_template& t, const float* sdf) {
  return t.affine ? planeClearBall(local, reach, t) : lipschitzClear(local, reach, t, sdf);
}

// True when the ball (local, reach) misses the SDF box of t (target frame).
__device__ __forceinline__ bool outsideBox(V3 local, double reach, const fdm_template& t) {
  const double h = 1.0 / t.invSpacing;
  double d2 = 0.0;
  const double lo[3] = {t.origin[0], t.origin[1], t.origin[2]};
  const double p[3] = {local.x, local.y, local.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double hi = lo[a] + h * (double)(t.size[a] - 1);
    const double m = 1e-6 * (hi - lo[a]);
    double e = 0.0;
    if (p[a] < lo[a] - m) e = lo[a] - m - p[a];
    else if (p[a] > hi + m) e = p[a] - hi - m;
    d2 += e * e;
  }
  const double r = reach * kReach;
  return d2 > r * r;
}

// Clamped cell range of the box [c - r, c + r] (monotone in each coordinate).
__device__ __forceinline__ void cellRange(V3 c, double r, const fdm_grid& g, I3& lo, I3& hi) {
  lo = clampedCell(vsub(c, {r, r, r}), g);
  hi = clampedCell(vadd(c, {r, r, r}), g);
}
__device__ __forceinline__ bool cellNear(I3 t, I3 lo, I3 hi) {
  return t.x >= lo.x - 1 && t.x <= hi.x + 1 && t.y >= lo.y - 1 && t.y <= hi.y + 1 &&
         t.z >= lo.z - 1 && t.z <= hi.z + 1;
}

// ---------------------------------------------------------------- broad phase
struct BroadK {
  int set;
  fdm_bodies P;
  fdm_geometry pgeo;
  fdm_bodies W;
  fdm_geometry wgeo;
  int ghostCapacity;
  const int* ghostCount;
  const int* ghostReal;
  const double* ghostPos;
  const int* ghostBox;
  fdm_grid grid;
  fdm_hash_table table;
  fdm_pairs pairs;
  unsigned* status;
  double skin;
  double maxReach;
  const int* gate;
  fdm_instances inst;
  fdm_periodic periodic;
};

template <int SET>
__device__ __forceinline__ bool acceptCandidate(const BroadK& a, int A, V3 xA, double RA, int T,
                                                int G, V3 xT) {
  const fdm_template tT = SET == FDM_SET_WALL ? tmplOf(a.wgeo, ldu(a.W.geo + T))
                                              : tmplOf(a.pgeo, ldu(a.P.geo + T));
  const Q4 qT = SET == FDM_SET_WALL ? ld4R(a.W.q, a.W.n, T) : ld4R(a.P.q, a.P.n, T);
  const float* sdf = SET == FDM_SET_WALL ? a.wgeo.sdf : a.pgeo.sdf;
  const double reachT = SET == FDM_SET_PARTICLE ? tT.radius : tT.boxReach;
  const V3 d = vsub(xA, xT);
  const double rs = (RA + reachT) * kReach + a.skin;
  if (vlen2(d) > rs * rs) return false;
  const V3 local = qrotateT(d, qT);
  const double reach = RA * kReach + a.skin;
  // no node can enter the SDF box, or none can reach phi <= 0 inside it
  return !outsideBox(local, reach, tT) && !lipschitzClear(local, reach, tT, sdf);
}

// Visit the candidate targets of particle A; `emit(T, G)` for each.
//
// Every accepted pair satisfies a necessary condition of a node contact with
// the reach widened by the skin: a node lies within R_A of x_A, and a contact
// needs it within R_T of x_T (particle targets, sphere test) or inside T's
// SDF box, hence within the box reach of x_T (walls, ghosts). The cells
// visited cover all targets within that distance, so lists built at one step
// stay complete while bodies move less than skin/2:
//  - particles are stored under their cell id; the visited range is the
//    geometric range of A's reach plus the largest target radius, joined with
//    the reference's own stencil (one cell around A's node cells), which also
//    covers a target whose id aliases another cell in the step it is built;
//  - walls are all visited;
//  - ghosts are stored under their clamped geometric cell, which is monotone
//    in the position, so the geometric range alone suffices.
template <int SET, class Emit>
__device__ __forceinline__ void visitCandidates(const BroadK& a, int A, Emit emit) {
  const int n = a.P.n;
  const double RA = tmplOf(a.pgeo, ldu(a.P.geo + A)).radius;
  // with several instances: A's own grid, and its instance's cells and walls
  const bool many = a.inst.count > 0;
  const int in = many ? a.inst.particleInstance[A] : 0;
  // ghost contacts are detected after the periodic wrap
  const V3 xA = SET == FDM_SET_GHOST
                    ? wrapPosition(ldpR(a.P.x, n, A), many ? a.inst.pgrid[in] : a.grid,
                                   many ? a.inst.periodic[in] : a.periodic)
                    : ldpR(a.P.x, n, A);
  if (SET == FDM_SET_WALL) {
    const fdm_grid& wg = many ? a.inst.wgrid[in] : a.grid;
    const int w0 = many ? a.inst.wallStart[in] : 0;
    const int w1 = many ? a.inst.wallStart[in + 1] : a.W.n;
    for (int w = w0; w < w1; ++w) {
      const int h = a.table.hashValue[w];
      if (h < 0 || (long long)h >= wg.num) continue;
      if (acceptCandidate<SET>(a, A, xA, RA, w, w, ldpR(a.W.x, a.W.n, w))) emit(w, w);
    }
    return;
  }
  const fdm_grid& grid = many ? a.inst.pgrid[in] : a.grid;
  const long long cellOffset = many ? a.inst.pcellOffset[in] : 0;
  I3 b0, b1;
  cellRange(xA, (RA + a.maxReach) * kReach + a.skin, grid, b0, b1);
  if (SET == FDM_SET_PARTICLE) {
    I3 lo, hi;
    cellRange(xA, RA * kReach + a.skin, grid, lo, hi);
    b0 = {min(b0.x, max(lo.x - 1, 0)), min(b0.y, max(lo.y - 1, 0)), min(b0.z, max(lo.z - 1, 0))};
    b1 = {max(b1.x, min(hi.x + 1, grid.size.x - 1)), max(b1.y, min(hi.y + 1, grid.size.y - 1)),
          max(b1.z, min(hi.z + 1, grid.size.z - 1))};
  } else {
    if (*a.ghostCount <= 0) return;
    b0 = {max(b0.x, a.ghostBox[0]), max(b0.y, a.ghostBox[1]), max(b0.z, a.ghostBox[2])};
    b1 = {min(b1.x, a.ghostBox[3]), min(b1.y, a.ghostBox[4]), min(b1.z, a.ghostBox[5])};
  }
  const I3 sz = gsize(grid);
  for (int cz = b0.z; cz <= b1.z; ++cz)
    for (int cy = b0.y; cy <= b1.y; ++cy) {
      // The cells of an x-row have consecutive ids and the table is sorted by
      // cell id, so the row's entries form one range: the bounds of its
      // cells are independent loads, and the range is visited in the same
      // order as cell by cell.
      const long long h0 = cellOffset + lin3({b0.x, cy, cz}, sz);
      int s = INT_MAX, e = 0;
#pragma unroll 4
      for (int k = 0; k <= b1.x - b0.x; ++k) {
        const int cs = a.table.cellStart[h0 + k];
        const int ce = a.table.cellEnd[h0 + k];
        if (cs >= 0) {
          s = min(s, cs);
          e = max(e, ce);
        }
      }
      for (int m = s; m < e; ++m) {
        const int idx = a.table.sortIndex[m];
        if (SET == FDM_SET_PARTICLE) {
          if (idx <= A) continue;
          if (acceptCandidate<SET>(a, A, xA, RA, idx, idx, ldpR(a.P.x, n, idx))) emit(idx, idx);
        } else {
          const int j = a.ghostReal[idx];
          if (acceptCandidate<SET>(a, A, xA, RA, j, idx, ldd3R(a.ghostPos + 3 * idx)))
            emit(j, idx);
        }
      }
    }
}

template <int SET>
__global__ void kBroadCount(BroadK a) {
  if (gateOff(a.gate)) return;
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A > a.P.n) return;
  if (A == a.P.n) {
    a.pairs.count[A] = 0;
    return;
  }
  int c = 0;
  visitCandidates<SET>(a, A, [&](int, int) { ++c; });
  a.pairs.count[A] = c;
}

template <int SET>
__global__ void kBroadWrite(BroadK a) {
  if (gateOff(a.gate)) return;
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= a.P.n) return;
  const int begin = a.pairs.count[A];
  const int end = a.pairs.count[A + 1];
  if (begin == end) return;
  if (end > a.pairs.capacity) {
    // mark the slots that fit as empty so the narrow phase skips them
    for (int w = begin; w < a.pairs.capacity; ++w) a.pairs.pairA[w] = -1;
    atomicOr(a.status, FDM_STATUS_PAIR_OVERFLOW);
    return;
  }
  int w = begin;
  visitCandidates<SET>(a, A, [&](int T, int G) {
    if (w < end) {
      a.pairs.pairA[w] = A;
      a.pairs.pairT[w] = T;
      a.pairs.pairG[w] = G;
    }
    ++w;
  });
  // insertion sort of the segment by target id (lists are short)
  for (int i = begin + 1; i < end; ++i) {
    const int t = a.pairs.pairT[i], g = a.pairs.pairG[i];
    int j = i - 1;
    while (j >= begin && a.pairs.pairT[j] > t) {
      a.pairs.pairT[j + 1] = a.pairs.pairT[j];
      a.pairs.pairG[j + 1] = a.pairs.pairG[j];
      --j;
    }
    a.pairs.pairT[j + 1] = t;
    a.pairs.pairG[j + 1] = g;
  }
}

// -------------------------------------------------------------- narrow phase
// Fields shared by every interaction set of a step.
struct NarrowCommon {
  fdm_bodies P;
  fdm_geometry pgeo;
  const fdm_material* pmat;
  fdm_bodies W;
  fdm_geometry wgeo;
  const fdm_material* wmat;
  fdm_grid pgrid;
  fdm_periodic periodic;
  fdm_instances inst;
  double dt;
  int updateSprings;
  unsigned* status;
  double* dbgPoint;
  double* dbgNormal;
  double* dbgOverlap;
};

// One interaction set.
struct NarrowSet {
  int kind;           // FDM_SET_*
  int ghostMask;      // ghost sets: 1 X, 2 Y, 3 XY
  int ghostFromReal;  // ghost sets: derive positions from the real particles
  const double* ghostPos;
  const int* ghostHash;
  const double* ghostQ;    // sector images: orientation and velocity of each ghost
  const double* ghostVel;
  fdm_grid grid;
  const int* targetHash;
  fdm_pairs pairs;
  fdm_pairs prev;
  fdm_mirror mirror;
};

struct NarrowK {
  NarrowCommon c;
  NarrowSet s;
};

constexpr int kMaxSets = 5;
constexpr int kXcds = 8;  // accelerator complex dies of the MI355X, each with its own L2
struct NarrowMulti {
  NarrowCommon c;
  const int* gate;
  int nsets;
  int wallSet;  // set swept by kWallGroups (kCullSweep leaves it out), or -1
  NarrowSet s[kMaxSets];
};

// Position used for force arms: the stored position after this step's
// periodic wrap (the reference computes forces after updateGhostSpatialGrid).
__device__ __forceinline__ V3 wrappedArm(V3 x, const fdm_grid& g, const fdm_periodic& per) {
  return wrapPosition(x, g, per);
}

// The grids and periodic flags that apply to a pair whose A is `A`: its
// instance's, or the single system's.
struct PairFrame {
  const fdm_grid* pgrid;   // particle grid (wrap, ghosts, particle cell ids)
  const fdm_grid* tgrid;   // grid of the target's cell ids (the wall grid for walls)
  fdm_periodic periodic;
};
template <int SET>
__device__ __forceinline__ PairFrame pairFrame(const NarrowCommon& C, const NarrowSet& S, int A) {
  PairFrame f;
  if (C.inst.count > 0) {
    const int in = ldu(C.inst.particleInstance + A);
    f.pgrid = &C.inst.pgrid[in];
    f.tgrid = SET == FDM_SET_WALL ? &C.inst.wgrid[in] : f.pgrid;
    f.periodic = C.inst.periodic[in];
  } else {
    f.pgrid = &C.pgrid;
    f.tgrid = &S.grid;
    f.periodic = C.periodic;
  }
  return f;
}

// A's position as the set's detection sees it: stored (before this step's
// periodic wrap) for particle and wall pairs, wrapped for ghost pairs, which
// the reference detects after updateGhostSpatialGrid (LSSolver.h:1168-1171).
template <int SET>
__device__ __forceinline__ V3 detectPosA(const NarrowCommon& C, const PairFrame& f, int A) {
  const V3 x = ldpR(C.P.x, C.P.n, A);
  return SET == FDM_SET_GHOST ? wrappedArm(x, *f.pgrid, f.periodic) : x;
}
// A's orientation as the set's detection sees it: a sector turn of this
// step's wrap rotates it for ghost pairs.
template <int SET>
__device__ __forceinline__ Q4 detectQA(const NarrowCommon& C, const PairFrame& f, int A) {
  const Q4 q = ld4R(C.P.q, C.P.n, A);
  if (SET != FDM_SET_GHOST || !f.periodic.sector) return q;
  return sectorTurnQ(q, sectorWrap(ldpR(C.P.x, C.P.n, A), *f.pgrid, f.periodic).turns, f.periodic);
}
// A particle's velocity as the contact forces see it: after this step's wrap,
// which turns it in a sector (the force kernels run after
// updateGhostSpatialGrid, LSSolver.h:1170-1175).
__device__ __forceinline__ V3 forceVel(const NarrowCommon& C, const PairFrame& f, int i) {
  const V3 v = ld3R(C.P.v, C.P.n, i);
  if (!f.periodic.sector) return v;
  return sectorTurnV(v, sectorWrap(ldpR(C.P.x, C.P.n, i), *f.pgrid, f.periodic).turns, f.periodic);
}

// The reference's node predicate for one (node, target) candidate: the
// node's cell is in the target's stencil, the node is within RB of a
// particle target, inside the target's SDF box, and phi <= 0 there. The
// conjunction is evaluated with the cheapest and most selective tests first:
// the sphere, the SDF box, phi, and last the cell test (three divisions),
// which then runs for the few nodes inside the target; `tcell` is the
// target's decoded cell.
template <int SET>
__device__ __forceinline__ bool nodeHit(V3 p, const Target& tg, I3 tcell, const fdm_grid& g,
                                        double RB, Sample& s) {
  if (tg.hash < 0 || (long long)tg.hash >= g.num) return false;
  const V3 rel = vsub(p, tg.x);
  if (SET == FDM_SET_PARTICLE) {
    if (vlen2(rel) > RB * RB) return false;
  }
  int base;
  const V3 local = toLocal(rel, tg);
  if (!sdfLocate(local, tg.t, s, base)) return false;
  if (SET == FDM_SET_WALL && planeClear(local, tg.t)) return false;
  sdfCorners(tg.sdf + base, tg.t, s);
  const double ov = -trilinear(s.x, s.y, s.z, s.c);
  if (!(ov >= 0.)) return false;
  const I3 c = clampedCell(p, g);
  return abs(tcell.x - c.x) <= 1 && abs(tcell.y - c.y) <= 1 && abs(tcell.z - c.z) <= 1;
}
// nodeHit in two halves, so that the SDF fetches of several nodes can be in
// flight together: the tests before the fetch (target cell id valid, sphere,
// SDF box) with the cell's location, then the tests after it (phi <= 0, the
// node's cell in the target's stencil).
template <int SET>
__device__ __forceinline__ bool nodeCell(V3 p, const Target& tg, const fdm_grid& g, double RB,
                                         Sample& s, int& base) {
  if (tg.hash < 0 || (long long)tg.hash >= g.num) return false;
  const V3 rel = vsub(p, tg.x);
  if (SET == FDM_SET_PARTICLE) {
    if (vlen2(rel) > RB * RB) return false;
  }
  const V3 local = toLocal(rel, tg);
  if (!sdfLocate(local, tg.t, s, base)) return false;
  return !(SET == FDM_SET_WALL && planeClear(local, tg.t));
}
__device__ __forceinline__ bool nodeHitAt(V3 p, const Sample& s, I3 tcell, const fdm_grid& g) {
  const double ov = -trilinear(s.x, s.y, s.z, s.c);
  if (!(ov >= 0.)) return false;
  const I3 c = clampedCell(p, g);
  return abs(tcell.x - c.x) <= 1 && abs(tcell.y - c.y) <= 1 && abs(tcell.z - c.z) <= 1;
}

__device__ __forceinline__ int findPrevPair(const NarrowSet& a, int p, int A, int T, int& off) {
  // lists kept from the previous step hold the pair at the same index
  if (p < min(ldu(a.prev.total), a.prev.capacity) && ldu(a.prev.pairA + p) == A &&
      ldu(a.prev.pairT + p) == T) {
    off = ldu(a.prev.recOff + p);
    return ldu(a.prev.recCnt + p);
  }
  int lo = min(ldu(a.prev.count + A), a.prev.capacity);
  int hi = min(ldu(a.prev.count + A + 1), a.prev.capacity) - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int t = ldu(a.prev.pairT + mid);
    if (t == T) {
      off = ldu(a.prev.recOff + mid);
      return ldu(a.prev.recCnt + mid);
    }
    if (t < T) lo = mid + 1;
    else hi = mid - 1;
  }
  off = 0;
  return 0;
}

__device__ __forceinline__ V3 findPrevSpring(const NarrowSet& a, int off, int cnt, int k) {
  int lo = off, hi = off + cnt - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const fdm_record r = a.prev.rec[mid];
    const int kk = r.node;
    if (kk == k) return recordSpring(r);
    if (kk < k) lo = mid + 1;
    else hi = mid - 1;
  }
  return {0., 0., 0.};
}

// The target of pair p as this step sees it: position, orientation, template,
// SDF, cell id, and the position its force arm uses.
template <int SET>
__device__ __forceinline__ void resolveTarget(const NarrowCommon& C, const NarrowSet& S, int p, int T,
                                              const PairFrame& f, Target& tg, V3& armB) {
  const int n = C.P.n;
  if (SET == FDM_SET_WALL) {
    tg.x = ldpR(C.W.x, C.W.n, T);
    tg.q = ld4R(C.W.q, C.W.n, T);
    tg.t = tmplOf(C.wgeo, ldu(C.W.geo + T));
    tg.sdf = C.wgeo.sdf;
    tg.hash = ldu(S.targetHash + T);
    armB = tg.x;
  } else if (SET == FDM_SET_PARTICLE) {
    tg.x = ldpR(C.P.x, n, T);
    tg.q = ld4R(C.P.q, n, T);
    tg.t = tmplOf(C.pgeo, ldu(C.P.geo + T));
    tg.sdf = C.pgeo.sdf;
    tg.hash = ldu(S.targetHash + T);
    armB = wrappedArm(tg.x, *f.pgrid, f.periodic);
  } else {
    if (S.ghostFromReal) {
      // no wrap happened this step: the ghost is the real particle shifted
      // by the periods, as the ghost stage computes it
      const WrapStages w = wrapStages(ldpR(C.P.x, n, T), *f.pgrid, f.periodic);
      tg.x = ghostOf(S.ghostMask == 1 ? w.afterX : w.afterY, *f.pgrid, S.ghostMask);
      tg.hash = ghostCellHash(tg.x, *f.pgrid);
    } else {
      const int G = ldu(S.pairs.pairG + p);
      tg.x = ldd3R(S.ghostPos + 3 * G);
      tg.hash = ldu(S.ghostHash + G);
    }
    if (S.ghostQ) {
      const double* gq = S.ghostQ + 4 * (size_t)S.pairs.pairG[p];
      tg.q = {gq[0], gq[1], gq[2], gq[3]};
    } else {
      tg.q = ld4R(C.P.q, n, T);
    }
    tg.t = tmplOf(C.pgeo, ldu(C.P.geo + T));
    tg.sdf = C.pgeo.sdf;
    armB = tg.x;
  }
  tg.ident = tg.q.q0 == 1.0 && tg.q.q1 == 0.0 && tg.q.q2 == 0.0 && tg.q.q3 == 0.0;
}

// Contacts staged per wave in LDS between the node sweep and the contact law.
constexpr int kMaxHits = 64;
struct HitSlots {
  double pc[3][kMaxHits];
  double nrm[3][kMaxHits];
  double ov[kMaxHits];
  int node[kMaxHits];
};

// Nodes of one pair that passed the tests before the SDF fetch (sphere, SDF
// box), per wave in LDS, in node order: position, fractional cell position,
// first corner of the cell, node id; and the hit mask of each chunk.
struct CandSlots {
  double px[kMaxHits], py[kMaxHits], pz[kMaxHits];
  double fx[kMaxHits], fy[kMaxHits], fz[kMaxHits];
  int base[kMaxHits];
  int node[kMaxHits];
  unsigned long long chunkHits[FDM_MAX_CHUNKS];
};

// One wave per candidate pair: chunk culls (one chunk per lane), the node
// sweep over surviving chunks (one node per lane) evaluating the reference's
// predicate, the contact geometry of every hit staged in LDS, then the
// contact law with one hit per lane and the pair sums folded in node order.
// A pair with more than kMaxHits contacts recomputes its hits chunk by chunk
// instead of staging them.
// CULL: first take the particle-level pre-cull test of sweepPair (sphere,
// SDF box, Lipschitz bound at this step's positions) and leave empty record
// ranges when it rejects the pair.
template <int SET, bool CULL = false>
__device__ void processPair(const NarrowCommon& C, const NarrowSet& S, int p, int lane,
                           HitSlots& hs, int pool) {
  const int A = ldu(S.pairs.pairA + p);
  if (A < 0) {  // slot left empty by a capacity overflow
    if (lane == 0) {
      S.pairs.recOff[p] = 0;
      S.pairs.recCnt[p] = 0;
    }
    return;
  }
  const int T = ldu(S.pairs.pairT + p);
  const int n = C.P.n;
  const fdm_template& tA = tmplOf(C.pgeo, ldu(C.P.geo + A));
  const PairFrame fr = pairFrame<SET>(C, S, A);
  const Q4 qA = detectQA<SET>(C, fr, A);
  const V3 xA = detectPosA<SET>(C, fr, A);
  const fdm_grid& tgrid = *fr.tgrid;
  Target tg;
  V3 armB;
  resolveTarget<SET>(C, S, p, T, fr, tg, armB);
  if (CULL) {
    const double RA = tA.radius * kReach;
    const double reachT = SET == FDM_SET_PARTICLE ? tg.t.radius : tg.t.boxReach;
    const V3 d = vsub(xA, tg.x);
    const double rs = RA + reachT * kReach;
    bool keep = vlen2(d) <= rs * rs;
    if (keep) {
      const V3 local = toLocal(d, tg);
      keep = !outsideBox(local, RA, tg.t) && !levelClear(local, RA, tg.t, tg.sdf);
    }
    if (!keep) {
      if (lane == 0) {
        S.pairs.recOff[p] = 0;
        S.pairs.recCnt[p] = 0;
      }
      return;
    }
  }
  const V3 armA = SET == FDM_SET_GHOST ? xA : wrappedArm(xA, *fr.pgrid, fr.periodic);
  const double RB = tg.t.radius;
  const I3 tcell = decodeHash(tg.hash, tgrid);
  const fdm_real* nodes = C.pgeo.nodes + 3 * (size_t)tA.nodeStart;
  const float4* chunks = reinterpret_cast<const float4*>(C.pgeo.chunks) + tA.chunkStart;
  const int nChunks = tA.chunkCount;
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));

  // chunk culls, one chunk per lane: a chunk survives unless its bounding
  // sphere fails the sphere, stencil, SDF-box or Lipschitz bound
  bool live = false;
  if (lane < nChunks) {
    const float4 ch = chunks[lane];
    const V3 cw = vadd(qrotate(qA, {(double)ch.x, (double)ch.y, (double)ch.z}), xA);
    const double rc = (double)ch.w * kReach;
    live = true;
    if (SET == FDM_SET_PARTICLE) {
      const double rr = (RB + rc) * (1.0 + 1e-9);
      if (vlen2(vsub(cw, tg.x)) > rr * rr) live = false;
    }
    if (live) {
      I3 lo, hi;
      cellRange(cw, rc, tgrid, lo, hi);
      if (!cellNear(tcell, lo, hi)) live = false;
    }
    if (live) {
      const V3 lc = toLocal(vsub(cw, tg.x), tg);
      if (outsideBox(lc, rc, tg.t) || levelClear(lc, rc, tg.t, tg.sdf)) live = false;
    }
  }
  const unsigned long long liveChunks = ballot64(live);

  // node sweep: lane c keeps chunk c's hit mask; hits in node order get LDS
  // slots holding their contact point, normal, overlap and node id
  unsigned long long myMask = 0ull;
  int total = 0;
  // the history range of the pair does not depend on the sweep: start it now
  int prevOff = 0;
  const int prevCnt = liveChunks ? findPrevPair(S, p, A, T, prevOff) : 0;
  // node coordinates of the next surviving chunk are loaded one iteration
  // ahead, so their latency overlaps the current chunk's SDF fetches
  fdm_real nx = 0, ny = 0, nz = 0;
  if (liveChunks) {
    const int k0 = __builtin_ctzll(liveChunks) * FDM_CHUNK + lane;
    if (k0 < tA.nodeCount) {
      nx = nodes[3 * k0];
      ny = nodes[3 * k0 + 1];
      nz = nodes[3 * k0 + 2];
    }
  }
  for (unsigned long long lcm = liveChunks; lcm; lcm &= lcm - 1ull) {
    const int c = __builtin_ctzll(lcm);
    const int k = c * FDM_CHUNK + lane;
    const V3 s = {(double)nx, (double)ny, (double)nz};
    const unsigned long long rest = lcm & (lcm - 1ull);
    if (rest) {
      const int k2 = __builtin_ctzll(rest) * FDM_CHUNK + lane;
      if (k2 < tA.nodeCount) {
        nx = nodes[3 * k2];
        ny = nodes[3 * k2 + 1];
        nz = nodes[3 * k2 + 2];
      }
    }
    bool hit = false;
    Sample sm;
    V3 pn;
    if (k < tA.nodeCount) {
      pn = vadd(qrotate(qA, s), xA);
      hit = nodeHit<SET>(pn, tg, tcell, tgrid, RB, sm);
    }
    const unsigned long long m = ballot64(hit);
    if (lane == c) myMask = m;
    if (hit) {
      const int slot = total + __popcll(m & lt);
      if (slot < kMaxHits) {
        const double ov = -trilinear(sm.x, sm.y, sm.z, sm.c);
        const V3 nrm = vnormalize(toWorld(gradientIntX(sm.x, sm.y, sm.z, sm.c), tg));
        const V3 pc = vadd(pn, vmul(nrm, 0.5 * ov));
        hs.pc[0][slot] = pc.x;
        hs.pc[1][slot] = pc.y;
        hs.pc[2][slot] = pc.z;
        hs.nrm[0][slot] = nrm.x;
        hs.nrm[1][slot] = nrm.y;
        hs.nrm[2][slot] = nrm.z;
        hs.ov[slot] = ov;
        hs.node[slot] = k;
      }
    }
    total += __popcll(m);
  }

  int base = 0;
  if (total > 0) {
    const int poolCap = S.pairs.recCapacity / FDM_REC_POOLS;
    if (lane == 0) base = atomicAdd(S.pairs.recTotal + pool * FDM_REC_STRIDE, total);
    base = readlaneI(base, 0);
    if (base + total > poolCap) {
      if (lane == 0) atomicOr(C.status, FDM_STATUS_RECORD_OVERFLOW);
      total = 0;
      base = 0;
    }
    base += pool * poolCap;
  }

  V3 SF = {0., 0., 0.}, STA = {0., 0., 0.}, STB = {0., 0., 0.};
  if (total > 0) {
    // material constants of the pair (contactKernel.cu:79-110,167-193,243-274)
    const int mi = ldu(C.P.mat + A);
    const fdm_material Mi = matOf(C.pmat, mi);
    const double invMi = tA.invMass;
    double kn = 0., ks = 0., effM = 0., mu, diss;
    V3 vB, wB;
    if (SET == FDM_SET_WALL) {
      const fdm_material Mw = matOf(C.wmat, ldu(C.W.mat + T));
      if (invMi > 0.) effM = 1. / invMi;
      mu = fmin(Mi.mu, Mw.mu);
      diss = Mi.e <= Mw.e ? Mi.dissipation : Mw.dissipation;
      kn = Mi.kn;
      ks = Mi.ks;
      vB = ld3R(C.W.v, C.W.n, T);
      wB = ld3R(C.W.w, C.W.n, T);
    } else {
      const fdm_material Mj = matOf(C.pmat, ldu(C.P.mat + T));
      const double invMj = tg.t.invMass;
      if (Mi.kn > 0. && Mj.kn > 0.) kn = Mi.kn * Mj.kn / (Mi.kn + Mj.kn);
      if (Mi.ks > 0. && Mj.ks > 0.) ks = Mi.ks * Mj.ks / (Mi.ks + Mj.ks);
      if (invMi > 0. || invMj > 0.) effM = 1. / (invMi + invMj);
      diss = Mi.e <= Mj.e ? Mi.dissipation : Mj.dissipation;
      mu = fmin(Mi.mu, Mj.mu);
      vB = SET == FDM_SET_GHOST && S.ghostVel
               ? ldd3R(S.ghostVel + 3 * (size_t)S.pairs.pairG[p])
               : forceVel(C, fr, T);
      wB = ld3R(C.P.w, n, T);  // a sector image keeps the real spin (contactKernel.cu:258)
    }
    const V3 vA = forceVel(C, fr, A);
    const V3 wA = ld3R(C.P.w, n, A);
    // contact law of one contact: its spring is updated in place and stored
    // with its node id as record r
    auto law = [&](V3 pc, V3 nrm, double ov, int k, int r, V3& F, V3& TA, V3& TB) {
      V3 spring = findPrevSpring(S, prevOff, prevCnt, k);
      if (C.updateSprings) {
        const V3 vc = vsub(vadd(vA, vcross(wA, vsub(pc, armA))), vadd(vB, vcross(wB, vsub(pc, armB))));
        linearContact(F, spring, vc, nrm, ov, C.dt, kn, ks, mu, diss, effM);
      }
      TA = vcross(vsub(pc, armA), F);
      TB = vcross(vsub(pc, armB), vneg(F));
      S.pairs.rec[r] = makeRecord(spring, k);
      if (C.dbgPoint) {
        C.dbgPoint[3 * r] = pc.x;
        C.dbgPoint[3 * r + 1] = pc.y;
        C.dbgPoint[3 * r + 2] = pc.z;
        C.dbgNormal[3 * r] = nrm.x;
        C.dbgNormal[3 * r + 1] = nrm.y;
        C.dbgNormal[3 * r + 2] = nrm.z;
        C.dbgOverlap[r] = ov;
      }
    };
    if (total <= kMaxHits) {
      // one staged contact per lane (the slots were written by other lanes)
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      V3 F = {0., 0., 0.}, TA = {0., 0., 0.}, TB = {0., 0., 0.};
      if (lane < total) {
        const V3 pc = {hs.pc[0][lane], hs.pc[1][lane], hs.pc[2][lane]};
        const V3 nrm = {hs.nrm[0][lane], hs.nrm[1][lane], hs.nrm[2][lane]};
        law(pc, nrm, hs.ov[lane], hs.node[lane], base + lane, F, TA, TB);
      }
      // pair sums, serially in node order
      for (int b = 0; b < total; ++b) {
        SF = vadd(SF, {readlaneD(F.x, b), readlaneD(F.y, b), readlaneD(F.z, b)});
        STA = vadd(STA, {readlaneD(TA.x, b), readlaneD(TA.y, b), readlaneD(TA.z, b)});
        STB = vadd(STB, {readlaneD(TB.x, b), readlaneD(TB.y, b), readlaneD(TB.z, b)});
      }
    } else {
      // more contacts than slots: recompute each chunk's hits
      const unsigned long long chunksWith = ballot64(myMask != 0ull);
      int running = 0;
      for (unsigned long long cw = chunksWith; cw; cw &= cw - 1ull) {
        const int c = __builtin_ctzll(cw);
        const unsigned lo32 = __builtin_amdgcn_readlane((unsigned)myMask, c);
        const unsigned hi32 = __builtin_amdgcn_readlane((unsigned)(myMask >> 32), c);
        const unsigned long long m = ((unsigned long long)hi32 << 32) | lo32;
        const bool hit = (m >> lane) & 1ull;
        const int k = c * FDM_CHUNK + lane;
        V3 F = {0., 0., 0.}, TA = {0., 0., 0.}, TB = {0., 0., 0.};
        if (hit) {
          const V3 s = {(double)nodes[3 * k], (double)nodes[3 * k + 1], (double)nodes[3 * k + 2]};
          const V3 pn = vadd(qrotate(qA, s), xA);
          Sample sm;
          nodeHit<SET>(pn, tg, tcell, tgrid, RB, sm);
          const double ov = -trilinear(sm.x, sm.y, sm.z, sm.c);
          const V3 nrm = vnormalize(toWorld(gradientIntX(sm.x, sm.y, sm.z, sm.c), tg));
          const V3 pc = vadd(pn, vmul(nrm, 0.5 * ov));
          law(pc, nrm, ov, k, base + running + __popcll(m & lt), F, TA, TB);
        }
        for (unsigned long long mm = m; mm; mm &= mm - 1ull) {
          const int b = __builtin_ctzll(mm);
          SF = vadd(SF, {readlaneD(F.x, b), readlaneD(F.y, b), readlaneD(F.z, b)});
          STA = vadd(STA, {readlaneD(TA.x, b), readlaneD(TA.y, b), readlaneD(TA.z, b)});
          STB = vadd(STB, {readlaneD(TB.x, b), readlaneD(TB.y, b), readlaneD(TB.z, b)});
        }
        running += __popcll(m);
      }
    }
  }
  if (lane == 0) {
    S.pairs.recOff[p] = base;
    S.pairs.recCnt[p] = total;
    double* f = S.pairs.sumF + 3 * (size_t)p;
    double* ta = S.pairs.sumTA + 3 * (size_t)p;
    double* tb = S.pairs.sumTB + 3 * (size_t)p;
    f[0] = SF.x; f[1] = SF.y; f[2] = SF.z;
    ta[0] = STA.x; ta[1] = STA.y; ta[2] = STA.z;
    tb[0] = STB.x; tb[1] = STB.y; tb[2] = STB.z;
    if (SET != FDM_SET_WALL && total > 0 && T != A) {
      const int pos = atomicAdd(&S.mirror.count[T], 1);
      if (pos < S.mirror.capacity) S.mirror.list[(size_t)T * S.mirror.capacity + pos] = p;
      else atomicOr(C.status, FDM_STATUS_BUCKET_OVERFLOW);
    }
  }
}

template <int SET>
__global__ void __launch_bounds__(256) kNarrow(NarrowK a) {
  __shared__ HitSlots slots[4];
  const int lane = laneId();
  const int wpb = blockDim.x >> 6;
  const int total = min(*a.s.pairs.total, a.s.pairs.capacity);
  const int stride = gridDim.x * wpb;
  HitSlots& hs = slots[threadIdx.x >> 6];
  for (int p = blockIdx.x * wpb + (threadIdx.x >> 6); p < total; p += stride)
    processPair<SET>(a.c, a.s, p, lane, hs, p % FDM_REC_POOLS);
}

// ---------------------------------------------------------------------------
// Narrow phase of several sets in three kernels: a pair-level cull (one thread
// per candidate pair), the node sweep (one wave per kept pair) that stages the
// geometry of every contact, and the contact law (one thread per kept pair)
// that walks its contacts in node order.

struct NarrowScratch {
  int* workHit;      // [work item] first staged contact
  // [8 * work item] or null: what the law kernel needs first, written by the
  // sweep so that the law's thread starts with one load instead of a chain:
  // record count, first record, first staged contact, A, target, and the
  // previous step's record range of the pair
  int* lawHead;
  double* hitGeom;   // [8 * hitCapacity]: contact point, normal, overlap, node id
  int hitCapacity;
  int* hitCount;     // [FDM_REC_COUNTERS] pool counters of the staged contacts
};

__device__ __forceinline__ void stageHit(double* g, V3 pc, V3 nrm, double ov, int k) {
  g[0] = pc.x;
  g[1] = pc.y;
  g[2] = pc.z;
  g[3] = nrm.x;
  g[4] = nrm.y;
  g[5] = nrm.z;
  g[6] = ov;
  g[7] = (double)k;
}

// Node sweep of one kept pair (one wave): the culls and exact predicate of
// processPair; the contacts' geometry is staged in node order (through LDS,
// or recomputed chunk by chunk past kMaxHits) for the contact-law kernel,
// and the pair's record range is reserved.
// CULL: first take the particle-level pre-cull test (sphere, SDF box,
// Lipschitz bound at this step's positions) from the same values, and leave
// empty record ranges when it rejects the pair.
// NPL (nodes per lane and iteration): 1 sweeps one surviving chunk at a
// time; 2 two chunks at a time; 3 first collects, chunk by chunk, the nodes
// that pass the tests before the SDF fetch into LDS (in node order) and then
// fetches the SDF cells of up to 64 of them at once, so that a pair whose
// few candidate nodes are spread over several chunks waits for one round of
// fetches instead of one per chunk.
template <int SET, bool CULL = false, int NPL = 1>
__device__ void sweepPair(const NarrowCommon& C, const NarrowSet& S, const NarrowScratch& X, int w,
                          int p, int lane, HitSlots& hs, int pool, CandSlots* cs = nullptr) {
  const int A = ldu(S.pairs.pairA + p);
  const int T = ldu(S.pairs.pairT + p);
  const int n = C.P.n;
  const fdm_template& tA = tmplOf(C.pgeo, ldu(C.P.geo + A));
  const PairFrame fr = pairFrame<SET>(C, S, A);
  const Q4 qA = detectQA<SET>(C, fr, A);
  const V3 xA = detectPosA<SET>(C, fr, A);
  const fdm_grid& tgrid = *fr.tgrid;
  Target tg;
  V3 armB;
  resolveTarget<SET>(C, S, p, T, fr, tg, armB);
  if (CULL) {
    const double RA = tA.radius * kReach;
    const double reachT = SET == FDM_SET_PARTICLE ? tg.t.radius : tg.t.boxReach;
    const V3 d = vsub(xA, tg.x);
    const double rs = RA + reachT * kReach;
    bool keep = vlen2(d) <= rs * rs;
    if (keep) {
      const V3 local = toLocal(d, tg);
      // (no Lipschitz test at the pair level here: it would put an SDF fetch
      // before the chunk culls, whose own Lipschitz tests reject the same
      // pairs with their fetches in flight together)
      keep = !outsideBox(local, RA, tg.t);
    }
    if (!keep) {
      if (lane == 0) {
        S.pairs.recOff[p] = 0;
        S.pairs.recCnt[p] = 0;
        if (X.lawHead) *reinterpret_cast<int4*>(X.lawHead + 8 * (size_t)w) = make_int4(0, 0, 0, 0);
      }
      return;
    }
  }
  const double RB = tg.t.radius;
  const I3 tcell = decodeHash(tg.hash, tgrid);
  const fdm_real* nodes = C.pgeo.nodes + 3 * (size_t)tA.nodeStart;
  const float4* chunks = reinterpret_cast<const float4*>(C.pgeo.chunks) + tA.chunkStart;
  const int nChunks = tA.chunkCount;
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));

  bool live = false;
  if (lane < nChunks) {
    const float4 ch = chunks[lane];
    const V3 cw = vadd(qrotate(qA, {(double)ch.x, (double)ch.y, (double)ch.z}), xA);
    const double rc = (double)ch.w * kReach;
    live = true;
    if (SET == FDM_SET_PARTICLE) {
      const double rr = (RB + rc) * (1.0 + 1e-9);
      if (vlen2(vsub(cw, tg.x)) > rr * rr) live = false;
    }
    if (live) {
      I3 lo, hi;
      cellRange(cw, rc, tgrid, lo, hi);
      if (!cellNear(tcell, lo, hi)) live = false;
    }
    if (live) {
      const V3 lc = toLocal(vsub(cw, tg.x), tg);
      if (outsideBox(lc, rc, tg.t) || levelClear(lc, rc, tg.t, tg.sdf)) live = false;
    }
  }
  const unsigned long long liveChunks = ballot64(live);
  // An affine (plane-wall) target: the plane's value at node s of A is
  // affine in s, plane(toLocal(R_A s + x_A - x_T)) = c0 + u . s with
  // g = R_T a (a: the plane's gradient), u = R_A^T g, c0 = plane_3 +
  // g . (x_A - x_T) (the rotation formula's transpose is the conjugate's,
  // for any quaternion). A node with c0 + u . s above the plane's slack has
  // phi > 0 (the slack covers this evaluation's rounding as well), so its
  // transform and SDF fetch are skipped.
  const bool planeNodes = SET == FDM_SET_WALL && tg.t.affine;
  V3 pu = {0., 0., 0.};
  double pc0 = 0.;
  if (planeNodes && liveChunks) {
    const V3 ga = toWorld({tg.t.plane[0], tg.t.plane[1], tg.t.plane[2]}, tg);
    pu = qrotateT(ga, qA);
    pc0 = tg.t.plane[3] + vdot(ga, vsub(xA, tg.x));
  }
  auto planeSkip = [&](V3 sn) {
    return planeNodes && pc0 + pu.x * sn.x + pu.y * sn.y + pu.z * sn.z > tg.t.planeSlack;
  };
  // the previous step's record range of the pair, for the law's header (its
  // scalar loads complete while the nodes are swept)
  int prevOff = 0;
  const int prevCnt = X.lawHead && liveChunks ? findPrevPair(S, p, A, T, prevOff) : 0;

  unsigned long long myMask = 0ull;
  int total = 0;
  // a hit's contact geometry, into LDS slot `slot` (node order)
  auto stage = [&](int slot, int k, V3 pn, const Sample& sm) {
    const double ov = -trilinear(sm.x, sm.y, sm.z, sm.c);
    const V3 nrm = vnormalize(toWorld(gradientIntX(sm.x, sm.y, sm.z, sm.c), tg));
    const V3 pc = vadd(pn, vmul(nrm, 0.5 * ov));
    hs.pc[0][slot] = pc.x;
    hs.pc[1][slot] = pc.y;
    hs.pc[2][slot] = pc.z;
    hs.nrm[0][slot] = nrm.x;
    hs.nrm[1][slot] = nrm.y;
    hs.nrm[2][slot] = nrm.z;
    hs.ov[slot] = ov;
    hs.node[slot] = k;
  };
  if (NPL == 3) {
    if (lane < nChunks) cs->chunkHits[lane] = 0ull;
    int ncand = 0;
    // SDF fetch and predicate of the collected candidates, their hits staged
    // in node order
    auto flush = [&]() {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      bool hit = false;
      Sample sm;
      V3 pn = {0., 0., 0.};
      int k = 0;
      if (lane < ncand) {
        pn = {cs->px[lane], cs->py[lane], cs->pz[lane]};
        sm.x = cs->fx[lane];
        sm.y = cs->fy[lane];
        sm.z = cs->fz[lane];
        k = cs->node[lane];
        sdfCorners(tg.sdf + cs->base[lane], tg.t, sm);
        hit = nodeHitAt(pn, sm, tcell, tgrid);
      }
      const unsigned long long m = ballot64(hit);
      if (hit) {
        const int slot = total + __popcll(m & lt);
        if (slot < kMaxHits) stage(slot, k, pn, sm);
        atomicOr(&cs->chunkHits[k / FDM_CHUNK], 1ull << (k % FDM_CHUNK));
      }
      total += __popcll(m);
      ncand = 0;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    };
    fdm_real cx = 0, cy = 0, cz = 0;
    if (liveChunks) {
      const int k0 = __builtin_ctzll(liveChunks) * FDM_CHUNK + lane;
      if (k0 < tA.nodeCount) {
        cx = nodes[3 * k0];
        cy = nodes[3 * k0 + 1];
        cz = nodes[3 * k0 + 2];
      }
    }
    for (unsigned long long lcm = liveChunks; lcm; lcm &= lcm - 1ull) {
      const int c = __builtin_ctzll(lcm);
      const int k = c * FDM_CHUNK + lane;
      const V3 sn = {(double)cx, (double)cy, (double)cz};
      const unsigned long long rest = lcm & (lcm - 1ull);
      if (rest) {
        const int k2 = __builtin_ctzll(rest) * FDM_CHUNK + lane;
        if (k2 < tA.nodeCount) {
          cx = nodes[3 * k2];
          cy = nodes[3 * k2 + 1];
          cz = nodes[3 * k2 + 2];
        }
      }
      bool ok = false;
      Sample sm;
      int b = 0;
      V3 pn = {0., 0., 0.};
      if (k < tA.nodeCount && !planeSkip(sn)) {
        pn = vadd(qrotate(qA, sn), xA);
        ok = nodeCell<SET>(pn, tg, tgrid, RB, sm, b);
      }
      const unsigned long long m = ballot64(ok);
      if (ncand + __popcll(m) > kMaxHits) flush();
      if (ok) {
        const int slot = ncand + __popcll(m & lt);
        cs->px[slot] = pn.x;
        cs->py[slot] = pn.y;
        cs->pz[slot] = pn.z;
        cs->fx[slot] = sm.x;
        cs->fy[slot] = sm.y;
        cs->fz[slot] = sm.z;
        cs->base[slot] = b;
        cs->node[slot] = k;
      }
      ncand += __popcll(m);
    }
    if (ncand > 0) flush();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane < nChunks) myMask = cs->chunkHits[lane];
  }
  if (NPL == 2) {
    // two surviving chunks per iteration (128 nodes in flight per wave): each
    // lane tests node `lane` of both, the second chunk's SDF fetches issued
    // before the first chunk's values are used; hits staged in node order
    for (unsigned long long lcm = liveChunks; lcm;) {
      const int c = __builtin_ctzll(lcm);
      lcm &= lcm - 1ull;
      const int c2 = lcm ? __builtin_ctzll(lcm) : -1;
      if (lcm) lcm &= lcm - 1ull;
      const int k = c * FDM_CHUNK + lane, k2 = c2 * FDM_CHUNK + lane;
      V3 s1 = {0., 0., 0.}, s2 = {0., 0., 0.};
      if (k < tA.nodeCount) s1 = {(double)nodes[3 * k], (double)nodes[3 * k + 1], (double)nodes[3 * k + 2]};
      if (c2 >= 0 && k2 < tA.nodeCount)
        s2 = {(double)nodes[3 * k2], (double)nodes[3 * k2 + 1], (double)nodes[3 * k2 + 2]};
      const bool has = k < tA.nodeCount && !planeSkip(s1);
      const bool has2 = c2 >= 0 && k2 < tA.nodeCount && !planeSkip(s2);
      V3 pn = {0., 0., 0.}, pn2 = {0., 0., 0.};
      if (has) pn = vadd(qrotate(qA, s1), xA);
      if (has2) pn2 = vadd(qrotate(qA, s2), xA);
      Sample sm, sm2;
      int b1 = 0, b2 = 0;
      bool in1 = has && nodeCell<SET>(pn, tg, tgrid, RB, sm, b1);
      bool in2 = has2 && nodeCell<SET>(pn2, tg, tgrid, RB, sm2, b2);
      if (in1) sdfCorners(tg.sdf + b1, tg.t, sm);
      if (in2) sdfCorners(tg.sdf + b2, tg.t, sm2);
      const bool hit = in1 && nodeHitAt(pn, sm, tcell, tgrid);
      const bool hit2 = in2 && nodeHitAt(pn2, sm2, tcell, tgrid);
      const unsigned long long m = ballot64(hit), m2 = ballot64(hit2);
      if (lane == c) myMask = m;
      if (c2 >= 0 && lane == c2) myMask = m2;
      if (hit) {
        const int slot = total + __popcll(m & lt);
        if (slot < kMaxHits) stage(slot, k, pn, sm);
      }
      total += __popcll(m);
      if (hit2) {
        const int slot = total + __popcll(m2 & lt);
        if (slot < kMaxHits) stage(slot, k2, pn2, sm2);
      }
      total += __popcll(m2);
    }
  }
  fdm_real nx = 0, ny = 0, nz = 0;
  if (NPL == 1 && liveChunks) {
    const int k0 = __builtin_ctzll(liveChunks) * FDM_CHUNK + lane;
    if (k0 < tA.nodeCount) {
      nx = nodes[3 * k0];
      ny = nodes[3 * k0 + 1];
      nz = nodes[3 * k0 + 2];
    }
  }
  for (unsigned long long lcm = NPL == 1 ? liveChunks : 0ull; lcm; lcm &= lcm - 1ull) {
    const int c = __builtin_ctzll(lcm);
    const int k = c * FDM_CHUNK + lane;
    const V3 s = {(double)nx, (double)ny, (double)nz};
    const unsigned long long rest = lcm & (lcm - 1ull);
    if (rest) {
      const int k2 = __builtin_ctzll(rest) * FDM_CHUNK + lane;
      if (k2 < tA.nodeCount) {
        nx = nodes[3 * k2];
        ny = nodes[3 * k2 + 1];
        nz = nodes[3 * k2 + 2];
      }
    }
    bool hit = false;
    Sample sm;
    V3 pn;
    if (k < tA.nodeCount && !planeSkip(s)) {
      pn = vadd(qrotate(qA, s), xA);
      hit = nodeHit<SET>(pn, tg, tcell, tgrid, RB, sm);
    }
    const unsigned long long m = ballot64(hit);
    if (lane == c) myMask = m;
    if (hit) {
      const int slot = total + __popcll(m & lt);
      if (slot < kMaxHits) {
        const double ov = -trilinear(sm.x, sm.y, sm.z, sm.c);
        const V3 nrm = vnormalize(toWorld(gradientIntX(sm.x, sm.y, sm.z, sm.c), tg));
        const V3 pc = vadd(pn, vmul(nrm, 0.5 * ov));
        hs.pc[0][slot] = pc.x;
        hs.pc[1][slot] = pc.y;
        hs.pc[2][slot] = pc.z;
        hs.nrm[0][slot] = nrm.x;
        hs.nrm[1][slot] = nrm.y;
        hs.nrm[2][slot] = nrm.z;
        hs.ov[slot] = ov;
        hs.node[slot] = k;
      }
    }
    total += __popcll(m);
  }

  // the record and staging reservations and the mirror slot, their atomics
  // issued together by lane 0 (one round trip instead of three)
  int base = 0, gbase = 0, mpos = 0;
  const bool mirrored = SET != FDM_SET_WALL && total > 0 && T != A;
  if (total > 0) {
    int rb = 0, gb = 0;
    if (lane == 0) {
      rb = atomicAdd(S.pairs.recTotal + pool * FDM_REC_STRIDE, total);
      gb = atomicAdd(X.hitCount + pool * FDM_REC_STRIDE, total);
      if (mirrored) mpos = atomicAdd(&S.mirror.count[T], 1);
    }
    rb = readlaneI(rb, 0);
    gb = readlaneI(gb, 0);
    const int rCap = S.pairs.recCapacity / FDM_REC_POOLS, gCap = X.hitCapacity / FDM_REC_POOLS;
    if (rb + total > rCap || gb + total > gCap) {
      if (lane == 0)
        atomicOr(C.status, rb + total > rCap ? FDM_STATUS_RECORD_OVERFLOW : FDM_STATUS_STAGING_OVERFLOW);
      total = 0;
    } else {
      base = pool * rCap + rb;
      gbase = pool * gCap + gb;
    }
  }
  if (total > 0 && total <= kMaxHits) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane < total)
      stageHit(X.hitGeom + 8 * (size_t)(gbase + lane), {hs.pc[0][lane], hs.pc[1][lane], hs.pc[2][lane]},
               {hs.nrm[0][lane], hs.nrm[1][lane], hs.nrm[2][lane]}, hs.ov[lane], hs.node[lane]);
  } else if (total > 0) {
    // more contacts than slots: recompute each chunk's hits
    const unsigned long long chunksWith = ballot64(myMask != 0ull);
    int running = 0;
    for (unsigned long long cw = chunksWith; cw; cw &= cw - 1ull) {
      const int c = __builtin_ctzll(cw);
      const unsigned lo32 = __builtin_amdgcn_readlane((unsigned)myMask, c);
      const unsigned hi32 = __builtin_amdgcn_readlane((unsigned)(myMask >> 32), c);
      const unsigned long long m = ((unsigned long long)hi32 << 32) | lo32;
      if ((m >> lane) & 1ull) {
        const int k = c * FDM_CHUNK + lane;
        const V3 s = {(double)nodes[3 * k], (double)nodes[3 * k + 1], (double)nodes[3 * k + 2]};
        const V3 pn = vadd(qrotate(qA, s), xA);
        Sample sm;
        nodeHit<SET>(pn, tg, tcell, tgrid, RB, sm);
        const double ov = -trilinear(sm.x, sm.y, sm.z, sm.c);
        const V3 nrm = vnormalize(toWorld(gradientIntX(sm.x, sm.y, sm.z, sm.c), tg));
        const V3 pc = vadd(pn, vmul(nrm, 0.5 * ov));
        stageHit(X.hitGeom + 8 * (size_t)(gbase + running + __popcll(m & lt)), pc, nrm, ov, k);
      }
      running += __popcll(m);
    }
  }
  if (lane == 0) {
    S.pairs.recOff[p] = base;
    S.pairs.recCnt[p] = total;
    X.workHit[w] = gbase;
    if (X.lawHead) {
      int4* h = reinterpret_cast<int4*>(X.lawHead + 8 * (size_t)w);
      h[0] = make_int4(total, base, gbase, A);
      if (total > 0) h[1] = make_int4(T, prevOff, prevCnt, 0);
    }
    // (a pair whose reservation overflowed keeps its mirror slot, with no
    // records and zero sums: the slot count and the list agree, and the
    // target's fold adds nothing stale)
    if (mirrored) {
      if (mpos < S.mirror.capacity) S.mirror.list[(size_t)T * S.mirror.capacity + mpos] = p;
      else atomicOr(C.status, FDM_STATUS_BUCKET_OVERFLOW);
      if (total == 0)
        for (int c = 0; c < 3; ++c) {
          S.pairs.sumF[3 * (size_t)p + c] = 0.;
          S.pairs.sumTB[3 * (size_t)p + c] = 0.;
        }
    }
  }
}

// Four wall pairs per wave, sixteen lanes each (pair 4q + j on lanes
// 16j..16j+15): the pre-cull, chunk culls and exact node predicate of
// sweepPair<FDM_SET_WALL, true>, with the per-pair values held per lane. A
// group first collects, in node order, the nodes of its surviving chunks
// that pass the plane test of an affine target (every node of another
// target) into LDS, then evaluates the predicate on them sixteen at a time,
// so that the few nodes near a plane wall are fetched in one round for four
// pairs. Contacts are staged in LDS in node order (past kGroupHits they are
// recomputed straight into the workspace once the range is reserved).
constexpr int kGroupHits = 16;
constexpr int kGroupedWallPairs = 65536;
constexpr int kGroupCands = 64;
struct WallGroupSlots {
  double pc[3][4][kGroupHits];
  double nrm[3][4][kGroupHits];
  double ov[4][kGroupHits];
  int node[4][kGroupHits];
  int cand[4][kGroupCands];
};

__device__ void sweepWallGroup(const NarrowCommon& C, const NarrowSet& S, const NarrowScratch& X,
                               int g0, int nPairs, int q, unsigned gmask, int lane,
                               WallGroupSlots& ws) {
  const int grp = lane >> 4, gl = lane & 15, sh = 16 * grp;
  const int p = 4 * q + grp;
  const int w = g0 + p;
  const unsigned below = gl == 0 ? 0u : (0xFFFFu >> (16 - gl));  // lanes of the group below this one
  auto groupBits = [&](unsigned long long m) __attribute__((always_inline)) {
    return (unsigned)((m >> sh) & 0xFFFFull);
  };
  bool active = p < nPairs && ((gmask >> grp) & 1u);
  const int A = active ? ldu(S.pairs.pairA + p) : 0;
  const int T = active ? ldu(S.pairs.pairT + p) : 0;
  auto emptyRange = [&]() __attribute__((always_inline)) {
    if (gl == 0) {
      S.pairs.recOff[p] = 0;
      S.pairs.recCnt[p] = 0;
      *reinterpret_cast<int4*>(X.lawHead + 8 * (size_t)w) = make_int4(0, 0, 0, 0);
    }
  };
  if (active && A < 0) {  // slot left empty by a capacity overflow
    emptyRange();
    active = false;
  }
  // the caller gives groups whose pairs share the target (and system)
  const unsigned long long act = ballot64(active);
  if (!act) return;
  const int ref = __builtin_ctzll(act);
  const int Tu = readlaneI(T, ref), Au = readlaneI(A, ref);
  const int inst = C.inst.count > 0 ? ldu(C.inst.particleInstance + Au) : 0;
  // lanes of an inactive group read body Au and use nothing
  const int Ai = active ? A : Au;
  const int geoA = ldu(C.P.geo + Ai);
  const double radiusA = ldu(&C.pgeo.info[geoA].radius);
  const PairFrame fr = pairFrame<FDM_SET_WALL>(C, S, Au);
  const Q4 qA = detectQA<FDM_SET_WALL>(C, fr, Ai);
  const V3 xA = detectPosA<FDM_SET_WALL>(C, fr, Ai);
  // the wall grid by value (wave-uniform; a pointer that may point into
  // the kernel arguments would make the compiler copy them to scratch)
  const fdm_grid tgrid = C.inst.count > 0 ? ldu(C.inst.wgrid + inst) : S.grid;
  Target tg;
  V3 armB;
  resolveTarget<FDM_SET_WALL>(C, S, p, Tu, fr, tg, armB);
  if (active) {
    const double RA = radiusA * kReach;
    const double rs = RA + tg.t.boxReach * kReach;
    const V3 d = vsub(xA, tg.x);
    bool keep = vlen2(d) <= rs * rs;
    if (keep) keep = !outsideBox(toLocal(d, tg), RA, tg.t);
    if (!keep) {
      emptyRange();
      active = false;
    }
  }
  const double RB = tg.t.radius;
  const I3 tcell = decodeHash(tg.hash, tgrid);
  const int nChunks = active ? ldu(&C.pgeo.info[geoA].chunkCount) : 0;
  const int nodeCount = active ? ldu(&C.pgeo.info[geoA].nodeCount) : 0;
  const float4* chunks =
      reinterpret_cast<const float4*>(C.pgeo.chunks) + ldu(&C.pgeo.info[geoA].chunkStart);
  const fdm_real* nodes = C.pgeo.nodes + 3 * (size_t)ldu(&C.pgeo.info[geoA].nodeStart);

  // chunk culls, sixteen chunks of the group's pair at a time
  unsigned long long live = 0ull;
  for (int cb = 0; cb < FDM_MAX_CHUNKS && ballot64(cb < nChunks); cb += 16) {
    const int c = cb + gl;
    bool lv = false;
    if (c < nChunks) {
      const float4 ch = chunks[c];
      const V3 cw = vadd(qrotate(qA, {(double)ch.x, (double)ch.y, (double)ch.z}), xA);
      const double rc = (double)ch.w * kReach;
      I3 lo, hi;
      cellRange(cw, rc, tgrid, lo, hi);
      lv = cellNear(tcell, lo, hi);
      if (lv) {
        const V3 lc = toLocal(vsub(cw, tg.x), tg);
        if (outsideBox(lc, rc, tg.t) || levelClear(lc, rc, tg.t, tg.sdf)) lv = false;
      }
    }
    live |= (unsigned long long)groupBits(ballot64(lv)) << cb;
  }
  int prevOff = 0;
  const int prevCnt = active && live ? findPrevPair(S, p, A, T, prevOff) : 0;
  const bool planeNodes = active && tg.t.affine;
  V3 pu = {0., 0., 0.};
  double pc0 = 0.;
  if (planeNodes && live) {
    const V3 ga = toWorld({tg.t.plane[0], tg.t.plane[1], tg.t.plane[2]}, tg);
    pu = qrotateT(ga, qA);
    pc0 = tg.t.plane[3] + vdot(ga, vsub(xA, tg.x));
  }
  auto loadNode = [&](int k) __attribute__((always_inline)) {
    return V3{(double)nodes[3 * k], (double)nodes[3 * k + 1], (double)nodes[3 * k + 2]};
  };
  // a node worth the exact predicate: in range and, on an affine target,
  // not above the plane by more than its slack (as in sweepPair)
  auto candidate = [&](int k, V3 sn) __attribute__((always_inline)) {
    return k < nodeCount &&
           !(planeNodes && pc0 + pu.x * sn.x + pu.y * sn.y + pu.z * sn.z > tg.t.planeSlack);
  };
  auto stageLds = [&](int slot, int k, V3 pn, const Sample& sm) __attribute__((always_inline)) {
    const double ov = -trilinear(sm.x, sm.y, sm.z, sm.c);
    const V3 nrm = vnormalize(toWorld(gradientIntX(sm.x, sm.y, sm.z, sm.c), tg));
    const V3 pc = vadd(pn, vmul(nrm, 0.5 * ov));
    ws.pc[0][grp][slot] = pc.x;
    ws.pc[1][grp][slot] = pc.y;
    ws.pc[2][grp][slot] = pc.z;
    ws.nrm[0][grp][slot] = nrm.x;
    ws.nrm[1][grp][slot] = nrm.y;
    ws.nrm[2][grp][slot] = nrm.z;
    ws.ov[grp][slot] = ov;
    ws.node[grp][slot] = k;
  };

  int ncand = 0, total = 0;
  // the exact predicate of the collected candidates, sixteen per round, hits
  // staged in node order
  auto flush = [&]() __attribute__((always_inline)) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int r = 0; ballot64(r < ncand); r += 16) {
      bool hit = false;
      Sample sm;
      V3 pn = {0., 0., 0.};
      int k = 0;
      if (r + gl < ncand) {
        k = ws.cand[grp][r + gl];
        pn = vadd(qrotate(qA, loadNode(k)), xA);
        hit = nodeHit<FDM_SET_WALL>(pn, tg, tcell, tgrid, RB, sm);
      }
      const unsigned m = groupBits(ballot64(hit));
      if (hit) {
        const int slot = total + __popc(m & below);
        if (slot < kGroupHits) stageLds(slot, k, pn, sm);
      }
      total += __popc(m);
    }
    ncand = 0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  };
  // collection: the i-th surviving chunk of every group per iteration, its
  // 64 nodes four per lane (node 64c + 16u + gl, so u-major is node order)
  for (unsigned long long rem = live; ballot64(rem != 0ull);) {
    const int c = rem ? __builtin_ctzll(rem) : -1;
    if (rem) rem &= rem - 1ull;
    bool pass[4];
    unsigned m[4];
    int add = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = c * FDM_CHUNK + 16 * u + gl;
      pass[u] = c >= 0 && k < nodeCount && candidate(k, loadNode(k));
      m[u] = groupBits(ballot64(pass[u]));
      add += __popc(m[u]);
    }
    if (ballot64(ncand + add > kGroupCands)) flush();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (pass[u]) ws.cand[grp][ncand + __popc(m[u] & below)] = c * FDM_CHUNK + 16 * u + gl;
      ncand += __popc(m[u]);
    }
  }
  if (ballot64(ncand > 0)) flush();

  // record and staging reservations (one lane of each group)
  const int pool = w % FDM_REC_POOLS;
  int rb = 0, gb = 0;
  if (active && gl == 0 && total > 0) {
    rb = atomicAdd(S.pairs.recTotal + pool * FDM_REC_STRIDE, total);
    gb = atomicAdd(X.hitCount + pool * FDM_REC_STRIDE, total);
  }
  rb = __shfl(rb, sh);
  gb = __shfl(gb, sh);
  int base = 0, gbase = 0;
  if (active && total > 0) {
    const int rCap = S.pairs.recCapacity / FDM_REC_POOLS, gCap = X.hitCapacity / FDM_REC_POOLS;
    if (rb + total > rCap || gb + total > gCap) {
      if (gl == 0)
        atomicOr(C.status, rb + total > rCap ? FDM_STATUS_RECORD_OVERFLOW : FDM_STATUS_STAGING_OVERFLOW);
      total = 0;
    } else {
      base = pool * rCap + rb;
      gbase = pool * gCap + gb;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (active && total > 0 && total <= kGroupHits)
    for (int h = gl; h < total; h += 16)
      stageHit(X.hitGeom + 8 * (size_t)(gbase + h), {ws.pc[0][grp][h], ws.pc[1][grp][h], ws.pc[2][grp][h]},
               {ws.nrm[0][grp][h], ws.nrm[1][grp][h], ws.nrm[2][grp][h]}, ws.ov[grp][h], ws.node[grp][h]);
  // more contacts than slots: the group's surviving chunks again, hits
  // written in node order at the reserved range
  const bool again = active && total > kGroupHits;
  if (ballot64(again)) {
    int running = 0;
    for (unsigned long long rem = again ? live : 0ull; ballot64(rem != 0ull);) {
      const int c = rem ? __builtin_ctzll(rem) : -1;
      if (rem) rem &= rem - 1ull;
#pragma unroll 1
      for (int u = 0; u < 4; ++u) {
        const int k = c * FDM_CHUNK + 16 * u + gl;
        bool hit = false;
        Sample sm;
        V3 pn = {0., 0., 0.};
        if (c >= 0 && k < nodeCount) {
          const V3 sn = loadNode(k);
          if (candidate(k, sn)) {
            pn = vadd(qrotate(qA, sn), xA);
            hit = nodeHit<FDM_SET_WALL>(pn, tg, tcell, tgrid, RB, sm);
          }
        }
        const unsigned mh = groupBits(ballot64(hit));
        if (hit) {
          const double ov = -trilinear(sm.x, sm.y, sm.z, sm.c);
          const V3 nrm = vnormalize(toWorld(gradientIntX(sm.x, sm.y, sm.z, sm.c), tg));
          const V3 pc = vadd(pn, vmul(nrm, 0.5 * ov));
          stageHit(X.hitGeom + 8 * (size_t)(gbase + running + __popc(mh & below)), pc, nrm, ov, k);
        }
        running += __popc(mh);
      }
    }
  }
  if (active && gl == 0) {
    S.pairs.recOff[p] = base;
    S.pairs.recCnt[p] = total;
    X.workHit[w] = gbase;
    int4* h = reinterpret_cast<int4*>(X.lawHead + 8 * (size_t)w);
    h[0] = make_int4(total, base, gbase, A);
    if (total > 0) h[1] = make_int4(T, prevOff, prevCnt, 0);
  }
}

// Contact law of one kept pair with contacts (one thread): the previous
// springs by (A, target, node), the linear spring-dashpot law per contact in
// node order, and the pair's force and torque sums in that order.
template <int SET>
__device__ void lawPair(const NarrowCommon& C, const NarrowSet& S, const NarrowScratch& X, int w,
                        int p) {
  int cnt, base, gbase, A, T, prevOff = 0, prevCnt;
  if (X.lawHead) {
    const int4 h0 = ldu(reinterpret_cast<const int4*>(X.lawHead + 8 * (size_t)w));
    cnt = h0.x;
    if (cnt <= 0) return;
    const int4 h1 = ldu(reinterpret_cast<const int4*>(X.lawHead + 8 * (size_t)w + 4));
    base = h0.y;
    gbase = h0.z;
    A = h0.w;
    T = h1.x;
    prevOff = h1.y;
    prevCnt = h1.z;
  } else {
    cnt = S.pairs.recCnt[p];
    if (cnt <= 0) return;
    base = S.pairs.recOff[p];
    gbase = X.workHit[w];
    A = ldu(S.pairs.pairA + p);
    T = ldu(S.pairs.pairT + p);
    prevCnt = findPrevPair(S, p, A, T, prevOff);
  }
  const int n = C.P.n;
  const fdm_template& tA = tmplOf(C.pgeo, ldu(C.P.geo + A));
  const PairFrame fr = pairFrame<SET>(C, S, A);
  const V3 xA = detectPosA<SET>(C, fr, A);
  Target tg;
  V3 armB;
  resolveTarget<SET>(C, S, p, T, fr, tg, armB);
  const V3 armA = SET == FDM_SET_GHOST ? xA : wrappedArm(xA, *fr.pgrid, fr.periodic);
  // material constants of the pair (contactKernel.cu:79-110,167-193,243-274)
  const fdm_material Mi = matOf(C.pmat, ldu(C.P.mat + A));
  const double invMi = tA.invMass;
  double kn = 0., ks = 0., effM = 0., mu, diss;
  V3 vB, wB;
  if (SET == FDM_SET_WALL) {
    const fdm_material Mw = matOf(C.wmat, ldu(C.W.mat + T));
    if (invMi > 0.) effM = 1. / invMi;
    mu = fmin(Mi.mu, Mw.mu);
    diss = Mi.e <= Mw.e ? Mi.dissipation : Mw.dissipation;
    kn = Mi.kn;
    ks = Mi.ks;
    vB = ld3R(C.W.v, C.W.n, T);
    wB = ld3R(C.W.w, C.W.n, T);
  } else {
    const fdm_material Mj = matOf(C.pmat, ldu(C.P.mat + T));
    const double invMj = tg.t.invMass;
    if (Mi.kn > 0. && Mj.kn > 0.) kn = Mi.kn * Mj.kn / (Mi.kn + Mj.kn);
    if (Mi.ks > 0. && Mj.ks > 0.) ks = Mi.ks * Mj.ks / (Mi.ks + Mj.ks);
    if (invMi > 0. || invMj > 0.) effM = 1. / (invMi + invMj);
    diss = Mi.e <= Mj.e ? Mi.dissipation : Mj.dissipation;
    mu = fmin(Mi.mu, Mj.mu);
    vB = SET == FDM_SET_GHOST && S.ghostVel ? ldd3R(S.ghostVel + 3 * (size_t)S.pairs.pairG[p])
                                            : forceVel(C, fr, T);
    wB = ld3R(C.P.w, n, T);  // a sector image keeps the real spin (contactKernel.cu:258)
  }
  const V3 vA = forceVel(C, fr, A);
  const V3 wA = ld3R(C.P.w, n, A);
  V3 SF = {0., 0., 0.}, STA = {0., 0., 0.}, STB = {0., 0., 0.};
  // The pair's previous records are in node order, as its staged contacts
  // are: walk both in step rather than search per contact; and load the next
  // contact's staged geometry one iteration ahead.
  int j = prevOff;
  const int jEnd = prevOff + prevCnt;
  fdm_record pr{};
  if (j < jEnd) pr = ldu(S.prev.rec + j);
  double gn[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) gn[c] = ldu(X.hitGeom + 8 * (size_t)gbase + c);
  for (int h = 0; h < cnt; ++h) {
    double g[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) g[c] = gn[c];
    if (h + 1 < cnt) {
#pragma unroll
      for (int c = 0; c < 8; ++c) gn[c] = ldu(X.hitGeom + 8 * (size_t)(gbase + h + 1) + c);
    }
    const V3 pc = {g[0], g[1], g[2]};
    const V3 nrm = {g[3], g[4], g[5]};
    const double ov = g[6];
    const int k = (int)g[7];
    V3 F = {0., 0., 0.};
    while (j < jEnd && pr.node < k) {
      ++j;
      if (j < jEnd) pr = ldu(S.prev.rec + j);
    }
    V3 spring = j < jEnd && pr.node == k ? recordSpring(pr) : V3{0., 0., 0.};
    if (C.updateSprings) {
      const V3 vc = vsub(vadd(vA, vcross(wA, vsub(pc, armA))), vadd(vB, vcross(wB, vsub(pc, armB))));
      linearContact(F, spring, vc, nrm, ov, C.dt, kn, ks, mu, diss, effM);
    }
    const V3 TA = vcross(vsub(pc, armA), F);
    const V3 TB = vcross(vsub(pc, armB), vneg(F));
    SF = vadd(SF, F);
    STA = vadd(STA, TA);
    STB = vadd(STB, TB);
    const int r = base + h;
    S.pairs.rec[r] = makeRecord(spring, k);
    if (C.dbgPoint) {
      C.dbgPoint[3 * r] = pc.x;
      C.dbgPoint[3 * r + 1] = pc.y;
      C.dbgPoint[3 * r + 2] = pc.z;
      C.dbgNormal[3 * r] = nrm.x;
      C.dbgNormal[3 * r + 1] = nrm.y;
      C.dbgNormal[3 * r + 2] = nrm.z;
      C.dbgOverlap[r] = ov;
    }
  }
  double* f = S.pairs.sumF + 3 * (size_t)p;
  double* ta = S.pairs.sumTA + 3 * (size_t)p;
  double* tb = S.pairs.sumTB + 3 * (size_t)p;
  f[0] = SF.x; f[1] = SF.y; f[2] = SF.z;
  ta[0] = STA.x; ta[1] = STA.y; ta[2] = STA.z;
  tb[0] = STB.x; tb[1] = STB.y; tb[2] = STB.z;
}

// Pair-level cull of every set's candidates at this step's positions (one
// thread per pair): the pair is kept only if a node of A can still reach
// the target (the broad-phase tests without the skin). Culled pairs get
// empty record ranges; kept pairs go to the work list as (p << 3 | set).
// Candidate pairs of all sets in one index space: set k holds the indices
// [start[k], start[k + 1]).
struct SetStarts {
  int v[kMaxSets + 1];
};
__device__ __forceinline__ SetStarts setStarts(const NarrowMulti& a) {
  SetStarts st;
  st.v[0] = 0;
#pragma unroll
  for (int k = 0; k < kMaxSets; ++k)
    st.v[k + 1] = st.v[k] + (k < a.nsets ? min(*a.s[k].pairs.total, a.s[k].pairs.capacity) : 0);
  return st;
}

// The pre-cull and the sweep in one pass: one wave per candidate pair (WAVES
// waves per workgroup). The wave takes the pre-cull's test itself (its
// values are wave-uniform) and sweeps the pair only when it is kept; a culled
// pair gets empty record ranges. No work list: the candidate index is the
// work item.
// Four waves per SIMD (at most 128 VGPRs): the sweep waits on memory for
// most of its life, so the occupancy that hides that wait matters more than
// the few registers the collecting mode then spills.
template <int WAVES, int NPL>
__global__ void __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(4, 4)))
kCullSweep(NarrowMulti a, NarrowScratch x) {
  if (gateOff(a.gate)) return;
  __shared__ HitSlots slots[WAVES];
  __shared__ CandSlots cands[NPL == 3 ? WAVES : 1];
  const int lane = laneId();
  const int wave = WAVES == 1 ? 0 : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const SetStarts st = setStarts(a);
  const int total = st.v[kMaxSets];
  // Work assignment. Workgroups are dealt to the eight XCDs in turn
  // (workgroup b runs on XCD b mod 8), each XCD with an L2 of its own, and
  // neighbouring candidates share bodies, templates and list lines: so each
  // XCD takes runs of 64 consecutive candidates (run q on XCD q mod 8), its
  // waves sharing out its runs. Candidates are taken from the end of the
  // index space: the ghost sets (image pairs, whose spheres hold many nodes)
  // before the wall set, so that the cheap wall pairs fill the last round.
  // (a set swept by kWallGroups is left out of the index space)
  int ws0 = 0, nWall = 0;
#pragma unroll
  for (int q = 0; q < kMaxSets; ++q)
    if (q == a.wallSet) {
      ws0 = st.v[q];
      nWall = st.v[q + 1] - st.v[q];
    }
  const int nOther = total - nWall;
  const bool byXcd = gridDim.x % kXcds == 0;
  const int xcd = byXcd ? blockIdx.x % kXcds : 0;
  const int stride = byXcd ? gridDim.x / kXcds * WAVES : gridDim.x * WAVES;
  for (int i = (byXcd ? blockIdx.x / kXcds : blockIdx.x) * WAVES + wave;; i += stride) {
    const int r = byXcd ? ((i >> 6) * kXcds + xcd) * 64 + (i & 63) : i;
    if (r >= nOther) break;
    const int go = nOther - 1 - r;
    const int g = go < ws0 ? go : go + nWall;
    int k = 0, base = 0;
#pragma unroll
    for (int q = 1; q < kMaxSets; ++q)  // constant indices: the starts stay scalar
      if (g >= st.v[q]) {
        k = q;
        base = st.v[q];
      }
    const NarrowSet& s = a.s[k];
    const int p = g - base;
    if (s.pairs.pairA[p] < 0) {  // an empty slot of an overflowing list
      if (lane == 0) {
        s.pairs.recOff[p] = 0;
        s.pairs.recCnt[p] = 0;
        if (x.lawHead) *reinterpret_cast<int4*>(x.lawHead + 8 * (size_t)g) = make_int4(0, 0, 0, 0);
      }
      continue;
    }
    const int pool = g % FDM_REC_POOLS;
    CandSlots* cs = NPL == 3 ? &cands[wave] : nullptr;
    if (s.kind == FDM_SET_PARTICLE)
      sweepPair<FDM_SET_PARTICLE, true, NPL>(a.c, s, x, g, p, lane, slots[wave], pool, cs);
    else if (s.kind == FDM_SET_WALL)
      sweepPair<FDM_SET_WALL, true, NPL>(a.c, s, x, g, p, lane, slots[wave], pool, cs);
    else
      sweepPair<FDM_SET_GHOST, true, NPL>(a.c, s, x, g, p, lane, slots[wave], pool, cs);
  }
}

// The wall set's pairs four per wave (sweepWallGroup) for the one-node-per-
// lane sweep, launched before kCullSweep, which then leaves the wall set
// out. One wave per workgroup; runs of 16 groups (64 consecutive pairs) per
// XCD as in kCullSweep. A group of four whose pairs do not share the target
// and system is swept one pair at a time (the same code with one group of
// lanes taking part).
__global__ void __launch_bounds__(64) kWallGroups(NarrowMulti a, NarrowScratch x) {
  if (gateOff(a.gate)) return;
  __shared__ WallGroupSlots groups;
  const int lane = laneId();
  const SetStarts st = setStarts(a);
  int k = 0, ws0 = 0, nWall = 0;
#pragma unroll
  for (int q = 0; q < kMaxSets; ++q)  // constant indices: the starts stay scalar
    if (q == a.wallSet) {
      k = q;
      ws0 = st.v[q];
      nWall = st.v[q + 1] - st.v[q];
    }
  const NarrowSet& S = a.s[k];
  const NarrowCommon& C = a.c;
  const int items = (nWall + 3) / 4;
  const bool byXcd = gridDim.x % kXcds == 0;
  const int xcd = byXcd ? blockIdx.x % kXcds : 0;
  const int stride = byXcd ? gridDim.x / kXcds : gridDim.x;
  for (int i = byXcd ? blockIdx.x / kXcds : blockIdx.x;; i += stride) {
    const int r = byXcd ? ((i >> 4) * kXcds + xcd) * 16 + (i & 15) : i;
    if (r >= items) break;
    const int p = 4 * r + (lane >> 4);
    const bool in = p < nWall;
    const int A = in ? ldu(S.pairs.pairA + p) : -1;
    const int T = in ? ldu(S.pairs.pairT + p) : 0;
    const bool live = A >= 0;
    const int inst = live && C.inst.count > 0 ? ldu(C.inst.particleInstance + A) : 0;
    const unsigned long long act = ballot64(live);
    bool uniform = true;
    if (act) {
      const int ref = __builtin_ctzll(act);
      uniform = !ballot64(live && (T != readlaneI(T, ref) || inst != readlaneI(inst, ref)));
    }
    const int passes = uniform ? 1 : 4;
    for (int it = 0; it < passes; ++it)
      sweepWallGroup(C, S, x, ws0, nWall, r, uniform ? 0xFu : (1u << it), lane, groups);
  }
}

// The pre-cull, the sweep and the contact law in one pass: one wave per
// candidate pair (WAVES waves per workgroup) as kCullSweep, whose wave then
// applies the law to its pair's contacts (one contact per lane) and folds
// the pair sums in node order (processPair); no staging, no second kernel.
template <int WAVES>
__global__ void __launch_bounds__(64 * WAVES) kCullSweepLaw(NarrowMulti a) {
  if (gateOff(a.gate)) return;
  __shared__ HitSlots slots[WAVES];
  const int lane = laneId();
  const int wave = WAVES == 1 ? 0 : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const SetStarts st = setStarts(a);
  const int total = st.v[kMaxSets];
  for (int g = blockIdx.x * WAVES + wave; g < total; g += gridDim.x * WAVES) {
    int k = 0, base = 0;
#pragma unroll
    for (int q = 1; q < kMaxSets; ++q)  // constant indices: the starts stay scalar
      if (g >= st.v[q]) {
        k = q;
        base = st.v[q];
      }
    const NarrowSet& s = a.s[k];
    const int p = g - base;
    const int pool = g % FDM_REC_POOLS;
    if (s.kind == FDM_SET_PARTICLE)
      processPair<FDM_SET_PARTICLE, true>(a.c, s, p, lane, slots[wave], pool);
    else if (s.kind == FDM_SET_WALL)
      processPair<FDM_SET_WALL, true>(a.c, s, p, lane, slots[wave], pool);
    else
      processPair<FDM_SET_GHOST, true>(a.c, s, p, lane, slots[wave], pool);
  }
}

// The contact law over the candidate index space of kCullSweep (one thread
// per candidate; culled pairs and pairs without contacts return at once).
__global__ void __launch_bounds__(256) kContactLawAll(NarrowMulti a, NarrowScratch x) {
  if (gateOff(a.gate)) return;
  const SetStarts st = setStarts(a);
  const int total = st.v[kMaxSets];
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
    int k = 0, base = 0;
#pragma unroll
    for (int q = 1; q < kMaxSets; ++q)  // constant indices: the starts stay scalar
      if (g >= st.v[q]) {
        k = q;
        base = st.v[q];
      }
    const NarrowSet& s = a.s[k];
    const int p = g - base;
    if (s.kind == FDM_SET_PARTICLE) lawPair<FDM_SET_PARTICLE>(a.c, s, x, g, p);
    else if (s.kind == FDM_SET_WALL) lawPair<FDM_SET_WALL>(a.c, s, x, g, p);
    else lawPair<FDM_SET_GHOST>(a.c, s, x, g, p);
  }
}

__global__ void kResetCounter(int* c) {
  for (int i = threadIdx.x; i < FDM_REC_COUNTERS; i += blockDim.x) c[i] = 0;
}

}  // namespace fdm

namespace fdm {
namespace {
fdm_tuning g_tuning = {1, 32, 256, 256, 1, 1, 0};
}
const fdm_tuning& tuning() { return g_tuning; }
}  // namespace fdm

// The shapes t asks for, zero fields taken from the process default; false
// when a field is outside the documented shapes.
static bool completeTuning(const fdm_tuning& t, fdm_tuning& n) {
  n = fdm::tuning();
  if (t.sweepWaves) n.sweepWaves = t.sweepWaves;
  if (t.sweepWavesPerCU) n.sweepWavesPerCU = t.sweepWavesPerCU;
  if (t.lawBlock) n.lawBlock = t.lawBlock;
  if (t.reduceBlock) n.reduceBlock = t.reduceBlock;
  if (t.lawMode) n.lawMode = t.lawMode;
  if (t.nodesPerLane) n.nodesPerLane = t.nodesPerLane;
  if (t.wallGroups) n.wallGroups = t.wallGroups;
  if (!(n.sweepWaves == 1 || n.sweepWaves == 2 || n.sweepWaves == 4)) return false;
  if (n.sweepWavesPerCU < 4 || n.sweepWavesPerCU > 128) return false;
  if (!(n.lawBlock == 64 || n.lawBlock == 128 || n.lawBlock == 256)) return false;
  if (n.reduceBlock < 64 || n.reduceBlock > 1024 || n.reduceBlock % 64) return false;
  if (!(n.lawMode == 1 || n.lawMode == 2)) return false;
  if (n.nodesPerLane < 1 || n.nodesPerLane > 3) return false;
  if (n.wallGroups < 0 || n.wallGroups > 2) return false;
  return true;
}


using namespace fdm;

extern "C" int fdm_launchBroadPhase(const fdm_broadphase_args* args, void* workspace,
                                    size_t workspaceBytes, hipStream_t stream) {
  if (!args) return FDM_EINVAL;
  const fdm_broadphase_args& g = *args;
  const int n = g.particles.n;
  if (n < 0) return FDM_EINVAL;
  BroadK k;
  k.set = g.set;
  k.P = g.particles;
  k.pgeo = g.pgeo;
  k.W = g.walls;
  k.wgeo = g.wgeo;
  k.ghostCapacity = g.ghosts ? g.ghosts->capacity : 0;
  k.ghostCount = g.ghosts ? g.ghosts->count : nullptr;
  k.ghostReal = g.ghosts ? g.ghosts->realId : nullptr;
  k.ghostPos = g.ghosts ? g.ghosts->pos : nullptr;
  k.ghostBox = g.ghosts ? g.ghosts->cellBox : nullptr;
  k.grid = g.grid;
  k.table = g.table;
  k.pairs = g.pairs;
  k.status = nullptr;
  k.skin = g.skin;
  k.maxReach = g.maxReach;
  k.gate = g.gate;
  k.inst = g.inst;
  k.periodic = g.periodic;
  if (g.set == FDM_SET_GHOST && !g.ghosts) return FDM_EINVAL;
  if (!(g.skin >= 0.0) || !(g.maxReach >= 0.0)) return FDM_EINVAL;
  const unsigned nb = blocksFor((long long)n + 1, 128);
  switch (g.set) {
    case FDM_SET_PARTICLE:
      hipLaunchKernelGGL(kBroadCount<FDM_SET_PARTICLE>, dim3(nb), dim3(128), 0, stream, k);
      break;
    case FDM_SET_WALL:
      hipLaunchKernelGGL(kBroadCount<FDM_SET_WALL>, dim3(nb), dim3(128), 0, stream, k);
      break;
    case FDM_SET_GHOST:
      hipLaunchKernelGGL(kBroadCount<FDM_SET_GHOST>, dim3(nb), dim3(128), 0, stream, k);
      break;
    default:
      return FDM_EINVAL;
  }
  int r = scanInts(g.pairs.count, g.pairs.count, (size_t)n + 1, false, g.pairs.total, workspace,
                   workspaceBytes, stream, g.gate);
  if (r != FDM_OK) return r;
  return launchStatus();
}

// The write half needs the status word; it is launched by the narrow-phase
// entry's companion below so that the broad-phase signature stays the
// reference-like (count, scan) pair.
extern "C" int fdm_launchBroadPhaseWrite(const fdm_broadphase_args* args, unsigned* status,
                                         hipStream_t stream) {
  if (!args) return FDM_EINVAL;
  const fdm_broadphase_args& g = *args;
  const int n = g.particles.n;
  BroadK k;
  k.set = g.set;
  k.P = g.particles;
  k.pgeo = g.pgeo;
  k.W = g.walls;
  k.wgeo = g.wgeo;
  k.ghostCapacity = g.ghosts ? g.ghosts->capacity : 0;
  k.ghostCount = g.ghosts ? g.ghosts->count : nullptr;
  k.ghostReal = g.ghosts ? g.ghosts->realId : nullptr;
  k.ghostPos = g.ghosts ? g.ghosts->pos : nullptr;
  k.ghostBox = g.ghosts ? g.ghosts->cellBox : nullptr;
  k.grid = g.grid;
  k.table = g.table;
  k.pairs = g.pairs;
  k.status = status;
  k.skin = g.skin;
  k.maxReach = g.maxReach;
  k.gate = g.gate;
  k.inst = g.inst;
  k.periodic = g.periodic;
  if (n <= 0) return FDM_OK;
  const unsigned nb = blocksFor(n, 128);
  switch (g.set) {
    case FDM_SET_PARTICLE:
      hipLaunchKernelGGL(kBroadWrite<FDM_SET_PARTICLE>, dim3(nb), dim3(128), 0, stream, k);
      break;
    case FDM_SET_WALL:
      hipLaunchKernelGGL(kBroadWrite<FDM_SET_WALL>, dim3(nb), dim3(128), 0, stream, k);
      break;
    case FDM_SET_GHOST:
      hipLaunchKernelGGL(kBroadWrite<FDM_SET_GHOST>, dim3(nb), dim3(128), 0, stream, k);
      break;
    default:
      return FDM_EINVAL;
  }
  return launchStatus();
}

namespace {

void fillCommon(NarrowCommon& k, const fdm_narrowphase_args& g) {
  k.P = g.particles;
  k.pgeo = g.pgeo;
  k.pmat = g.pmat;
  k.W = g.walls;
  k.wgeo = g.wgeo;
  k.wmat = g.wmat;
  k.pgrid = g.pgrid;
  k.periodic = g.periodic;
  k.inst = g.inst;
  k.dt = g.dt;
  k.updateSprings = g.updateSprings;
  k.status = g.status;
  k.dbgPoint = g.dbgPoint;
  k.dbgNormal = g.dbgNormal;
  k.dbgOverlap = g.dbgOverlap;
}

void fillSet(NarrowSet& k, const fdm_narrowphase_args& g) {
  k.kind = g.set;
  k.ghostMask = g.ghostMask;
  k.ghostFromReal = g.ghostFromReal;
  k.ghostPos = g.ghosts ? g.ghosts->pos : nullptr;
  k.ghostHash = g.ghosts ? g.ghosts->table.hashValue : nullptr;
  k.ghostQ = g.set == FDM_SET_GHOST && g.periodic.sector && g.ghosts ? g.ghosts->q : nullptr;
  k.ghostVel = g.set == FDM_SET_GHOST && g.periodic.sector && g.ghosts ? g.ghosts->vel : nullptr;
  k.grid = g.grid;
  k.targetHash = g.targetHash;
  k.pairs = g.pairs;
  k.prev = g.prev;
  k.mirror = g.mirror;
}

bool validSet(const fdm_narrowphase_args& g) {
  if (g.set < FDM_SET_PARTICLE || g.set > FDM_SET_GHOST) return false;
  if (g.set == FDM_SET_GHOST) {
    if (g.ghostMask < 1 || g.ghostMask > 3) return false;
    if (!g.ghostFromReal && !g.ghosts) return false;
    // sector images carry their own orientation and velocity
    if (g.periodic.sector && (g.ghostFromReal || !g.ghosts->q || !g.ghosts->vel)) return false;
  }
  if (g.periodic.sector && (g.periodic.x || g.periodic.y || g.inst.count > 0)) return false;
  return true;
}

}  // namespace

extern "C" int fdm_launchNarrowPhase(const fdm_narrowphase_args* args, hipStream_t stream) {
  if (!args || !validSet(*args)) return FDM_EINVAL;
  const fdm_narrowphase_args& g = *args;
  NarrowK k;
  fillCommon(k.c, g);
  fillSet(k.s, g);
  hipLaunchKernelGGL(kResetCounter, dim3(1), dim3(64), 0, stream, g.pairs.recTotal);
  const int blocks = persistentBlocks(4);
  switch (g.set) {
    case FDM_SET_PARTICLE:
      hipLaunchKernelGGL(kNarrow<FDM_SET_PARTICLE>, dim3(blocks), dim3(256), 0, stream, k);
      break;
    case FDM_SET_WALL:
      hipLaunchKernelGGL(kNarrow<FDM_SET_WALL>, dim3(blocks), dim3(256), 0, stream, k);
      break;
    default:
      hipLaunchKernelGGL(kNarrow<FDM_SET_GHOST>, dim3(blocks), dim3(256), 0, stream, k);
      break;
  }
  return launchStatus();
}

extern "C" int fdm_launchNarrowPhaseSets(const fdm_narrowphase_args* sets, int nsets,
                                         const fdm_narrow_workspace* ws, hipStream_t stream) {
  if (!sets || nsets < 1 || nsets > kMaxSets || !ws || !ws->workHit || !ws->hitGeom ||
      !ws->hitCount || ws->hitCapacity < FDM_REC_POOLS)
    return FDM_EINVAL;
  NarrowMulti k;
  fillCommon(k.c, sets[0]);
  k.gate = sets[0].gate;
  k.nsets = nsets;
  long long cap = 0;
  for (int i = 0; i < nsets; ++i) {
    if (!validSet(sets[i])) return FDM_EINVAL;
    fillSet(k.s[i], sets[i]);
    cap += sets[i].pairs.capacity;
  }
  if (cap > ws->workCapacity || cap >= (1ll << 28)) return FDM_EINVAL;
  for (int i = nsets; i < kMaxSets; ++i) k.s[i] = k.s[0];
  NarrowScratch x;
  x.workHit = ws->workHit;
  x.lawHead = ws->workList;
  x.hitGeom = ws->hitGeom;
  x.hitCapacity = ws->hitCapacity;
  x.hitCount = ws->hitCount;
  fdm_tuning tu = tuning();
  if (ws->tuning && !completeTuning(*ws->tuning, tu)) return FDM_EINVAL;
  // The wall set four pairs per wave (one node per lane, law in its own
  // kernel): by default for wall sets sized for at least kGroupedWallPairs
  // pairs, as a small system's step is bound by launch latency, which a
  // second sweep kernel would add to. The grid covers the set's capacity in
  // groups.
  k.wallSet = -1;
  int wallGroups = 0;
  if (tu.lawMode == 1 && tu.nodesPerLane == 1 && x.lawHead)
    for (int i = 0; i < nsets && k.wallSet < 0; ++i)
      if (sets[i].set == FDM_SET_WALL && tu.wallGroups != 2 &&
          (tu.wallGroups == 1 || sets[i].pairs.capacity >= kGroupedWallPairs)) {
        k.wallSet = i;
        wallGroups = (sets[i].pairs.capacity + 3) / 4;
      }
  const int sweepBlocks = std::max(1, persistentBlocks(tu.sweepWaves) * tu.sweepWavesPerCU / 32);
  if (k.wallSet >= 0) {
    const int blocks = std::min(persistentBlocks(1) * 12 / 32, wallGroups);
    hipLaunchKernelGGL(kWallGroups, dim3(std::max(kXcds, blocks / kXcds * kXcds)), dim3(64), 0, stream,
                       k, x);
  }
  if (tu.lawMode == 2) {
    if (tu.sweepWaves == 4)
      hipLaunchKernelGGL(kCullSweepLaw<4>, dim3(sweepBlocks), dim3(256), 0, stream, k);
    else if (tu.sweepWaves == 2)
      hipLaunchKernelGGL(kCullSweepLaw<2>, dim3(sweepBlocks), dim3(128), 0, stream, k);
    else
      hipLaunchKernelGGL(kCullSweepLaw<1>, dim3(sweepBlocks), dim3(64), 0, stream, k);
    return launchStatus();
  }
  if (tu.nodesPerLane == 3) {
    if (tu.sweepWaves == 4)
      hipLaunchKernelGGL((kCullSweep<4, 3>), dim3(sweepBlocks), dim3(256), 0, stream, k, x);
    else if (tu.sweepWaves == 2)
      hipLaunchKernelGGL((kCullSweep<2, 3>), dim3(sweepBlocks), dim3(128), 0, stream, k, x);
    else
      hipLaunchKernelGGL((kCullSweep<1, 3>), dim3(sweepBlocks), dim3(64), 0, stream, k, x);
  } else if (tu.nodesPerLane == 2) {
    if (tu.sweepWaves == 4)
      hipLaunchKernelGGL((kCullSweep<4, 2>), dim3(sweepBlocks), dim3(256), 0, stream, k, x);
    else if (tu.sweepWaves == 2)
      hipLaunchKernelGGL((kCullSweep<2, 2>), dim3(sweepBlocks), dim3(128), 0, stream, k, x);
    else
      hipLaunchKernelGGL((kCullSweep<1, 2>), dim3(sweepBlocks), dim3(64), 0, stream, k, x);
  } else if (tu.sweepWaves == 4) {
    hipLaunchKernelGGL((kCullSweep<4, 1>), dim3(sweepBlocks), dim3(256), 0, stream, k, x);
  } else if (tu.sweepWaves == 2) {
    hipLaunchKernelGGL((kCullSweep<2, 1>), dim3(sweepBlocks), dim3(128), 0, stream, k, x);
  } else {
    hipLaunchKernelGGL((kCullSweep<1, 1>), dim3(sweepBlocks), dim3(64), 0, stream, k, x);
  }
  hipLaunchKernelGGL(kContactLawAll, dim3(persistentBlocks(tu.lawBlock / 64)), dim3(tu.lawBlock), 0,
                     stream, k, x);
  return launchStatus();
}

extern "C" int fdm_setTuning(const fdm_tuning* t) {
  if (!t) return FDM_EINVAL;
  fdm_tuning n;
  if (!completeTuning(*t, n)) return FDM_EINVAL;
  g_tuning = n;
  return FDM_OK;
}

extern "C" int fdm_checkTuning(fdm_tuning* t) {
  if (!t) return FDM_EINVAL;
  fdm_tuning n;
  if (!completeTuning(*t, n)) return FDM_EINVAL;
  *t = n;
  return FDM_OK;
}

extern "C" void fdm_getTuning(fdm_tuning* t) {
  if (t) *t = g_tuning;
}
