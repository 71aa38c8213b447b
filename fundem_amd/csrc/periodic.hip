// Cell ids, planar periodic wrap and ghost images.
//
// Stage launchers restate the reference's kernels one for one
// (mySpatialGrid.cu:4-56,87-152,214-240). The fused stage used by the step
// computes, per particle, the X wrap, the X ghost, the Y wrap, the Y ghost and
// the XY ghost in the reference's order (PeriodicBoundary.h:221-295), keeps
// only ghosts whose ghost hash is valid (the others are never found by any
// node), compacts them in particle order and builds one hash table per set.
#include <algorithm>
#include <climits>
#include <cmath>

#include "fdm_device.hpp"
#include "fdm_launch.hpp"

namespace fdm {

__global__ void kHashOnly(const double* pos, int n, fdm_grid grid, int* hashValue, unsigned* sortKey,
                          int* sortIndex) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int h = cellHash(ldp(pos, n, i), grid);
  hashValue[i] = h;
  sortKey[i] = (unsigned)(h + 1);
  sortIndex[i] = i;
}

__global__ void kGhostHashOnly(const double* gpos, int n, fdm_grid grid, int* hashValue,
                               unsigned* sortKey, int* sortIndex) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int h = ghostCellHash(ldd3(gpos + 3 * i), grid);
  hashValue[i] = h;
  sortKey[i] = (unsigned)(h + 1);
  sortIndex[i] = i;
}

__global__ void kWrap(double* pos, int n, fdm_grid g, int axis) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  stp(pos, n, i, wrapAxis(ldp(pos, n, i), g, axis));
}

__global__ void kGhostPosition(double* gpos, const double* pos, int n, fdm_grid g, int mask) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const V3 p = ghostOf(ldp(pos, n, i), g, mask);
  gpos[3 * i] = p.x;
  gpos[3 * i + 1] = p.y;
  gpos[3 * i + 2] = p.z;
}

__device__ __forceinline__ bool sameV3(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// Visibility flags of every particle's images in the active sets (null
// pointers: sets that do not exist), per particle with its own instance's
// grid and periodic axes.
__global__ void kGhostVisibility(const double* pos, int n, fdm_grid grid, fdm_periodic sets,
                                 fdm_instances I, int* f0, int* f1, int* f2, const int* gate) {
  if (gateOff(gate)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool many = I.count > 0;
  const int in = many ? I.particleInstance[i] : 0;
  const int bits = visibilityBits(ldp(pos, n, i), many ? I.pgrid[in] : grid, sets,
                                  many ? I.periodic[in] : sets);
  if (f0) f0[i] = bits & 1;
  if (f1) f1[i] = (bits >> 1) & 1;
  if (f2) f2[i] = (bits >> 2) & 1;
}

// Pass 1: visibility flag of each particle's ghost in each active set. With
// list reuse (ctl != null) it flags a wrap and any visibility change since
// the last ghost list build.
__global__ void kGhostFlags(const double* pos, int n, fdm_grid g, fdm_periodic per, int* fx, int* fy,
                            int* fxy, const int* atBuild, int* ctl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool changed = false, wrapped = false;
  if (i < n) {
    const V3 x = ldp(pos, n, i);
    const WrapStages w = wrapStages(x, g, per);
    int bits = 0;
    if (per.sector) {
      const SectorWrap sw = sectorWrap(x, g, per);
      int* fl[3] = {fx, fy, fxy};
      for (int k = 0; k < 3; ++k) {
        const V3 gp = sectorImageOn(sw.x, g, k) ? sectorImagePos(sw.x, g, per, k) : sw.x;
        const int f = ghostCellHash(gp, g) >= 0 ? 1 : 0;
        fl[k][i] = f;
        bits |= f << k;
      }
      if (ctl) {
        wrapped = sw.turns != 0;
        changed = atBuild && bits != atBuild[i];
      }
    } else if (per.x) {
      const int f = ghostCellHash(ghostOf(w.afterX, g, 1), g) >= 0 ? 1 : 0;
      fx[i] = f;
      bits |= f;
    }
    if (per.y) {
      const int f = ghostCellHash(ghostOf(w.afterY, g, 2), g) >= 0 ? 1 : 0;
      fy[i] = f;
      bits |= f << 1;
    }
    if (per.x && per.y) {
      const int f = ghostCellHash(ghostOf(w.afterY, g, 3), g) >= 0 ? 1 : 0;
      fxy[i] = f;
      bits |= f << 2;
    }
    if (ctl) {
      wrapped = !sameV3(w.afterY, x);
      changed = atBuild && bits != atBuild[i];
    }
  }
  if (!ctl) return;
  if (ballot64(changed) != 0ull && laneId() == 0) ctl[FDM_CTL_GHOST_CHANGED] = 1;
  if (ballot64(wrapped) != 0ull && laneId() == 0) ctl[FDM_CTL_WRAPPED] = 1;
}

struct GhostOut {
  int* slot;       // exclusive scan of the flags
  const int* gate;
  int* realId;
  double* pos;
  int* hashValue;
  unsigned* sortKey;
  int* sortIndex;
  int capacity;
  unsigned* status;
  double* q;    // sector images: orientation and velocity
  double* vel;
};

__device__ __forceinline__ int emitGhost(const GhostOut& o, int i, V3 gp, const fdm_grid& g,
                                         long long cellOffset) {
  const int s = o.slot[i];
  if (s >= o.capacity) {
    atomicOr(o.status, FDM_STATUS_GHOST_OVERFLOW);
    return -1;
  }
  const int h = ghostCellHash(gp, g);
  o.realId[s] = i;
  o.pos[3 * s] = gp.x;
  o.pos[3 * s + 1] = gp.y;
  o.pos[3 * s + 2] = gp.z;
  o.hashValue[s] = h;
  // lists are built by the clamped geometric cell (monotone in the position);
  // the reference's ghost cell id above decides the stencil rule
  const I3 c = clampedCell(gp, g);
  o.sortKey[s] = (unsigned)(cellOffset + lin3(c, gsize(g)) + 1);
  o.sortIndex[s] = s;
  return s;  // the set's cell box is folded afterwards by kGhostBox
}

constexpr int kBoxBlocks = 128;
constexpr int kBoxThreads = 256;

// Bounding box of the clamped cells of a set's stored ghosts (the cells
// emitGhost keys them by), folded per workgroup and merged with one atomic
// per bound and workgroup. Folding inside the scatter cost one atomic per
// ghost on the same six words, which serialised at L2; min and max are
// exact, so the box does not depend on the order of the merges.
__global__ void __launch_bounds__(kBoxThreads) kGhostBox(const double* gpos, const int* realId,
                                                         const int* count, int capacity,
                                                         fdm_grid g0, fdm_instances I,
                                                         int* cellBox, const int* gate) {
  if (gateOff(gate)) return;
  __shared__ int part[6][kBoxThreads / 64];
  const int c = min(*count, capacity);
  int b[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < c; k += gridDim.x * blockDim.x) {
    const fdm_grid& g = I.count > 0 ? I.pgrid[I.particleInstance[realId[k]]] : g0;
    const I3 cc = clampedCell(ldd3(gpos + 3 * (size_t)k), g);
    b[0] = min(b[0], cc.x);
    b[1] = min(b[1], cc.y);
    b[2] = min(b[2], cc.z);
    b[3] = max(b[3], cc.x);
    b[4] = max(b[4], cc.y);
    b[5] = max(b[5], cc.z);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      b[d] = min(b[d], __shfl_xor(b[d], off, 64));
      b[3 + d] = max(b[3 + d], __shfl_xor(b[3 + d], off, 64));
    }
  }
  const int wave = threadIdx.x >> 6;
  if (laneId() == 0)
    for (int d = 0; d < 6; ++d) part[d][wave] = b[d];
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int w = 1; w < kBoxThreads / 64; ++w)
    for (int d = 0; d < 3; ++d) {
      b[d] = min(b[d], part[d][w]);
      b[3 + d] = max(b[3 + d], part[3 + d][w]);
    }
  if (b[0] > b[3]) return;  // no ghost in this workgroup
  for (int d = 0; d < 3; ++d) {
    atomicMin(&cellBox[d], b[d]);
    atomicMax(&cellBox[3 + d], b[3 + d]);
  }
}

// Pass 2 (after the flags were scanned into slots): scatter the visible
// ghosts of each set. The wrapped positions are stored later, by
// kCommitWrap after the step's detection: the reference detects particle and
// wall contacts before the wrap (LSSolver.h:1168-1171).
// With several instances (I.count > 0) each particle uses its instance's
// grid and periodic flags; `per` then only says which sets exist.
__global__ void kGhostScatter(const double* pos, const fdm_real* vel, const fdm_real* quat, int n,
                              fdm_grid g0, fdm_periodic per, GhostOut ox, GhostOut oy, GhostOut oxy,
                              int* atBuild, fdm_instances I) {
  if (gateOff(ox.gate)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (per.sector) {  // one system: the sector is not combined with instances
    const SectorWrap sw = sectorWrap(ldp(pos, n, i), g0, per);
    const V3 v = sectorTurnV(ld3(vel, n, i), sw.turns, per);
    const Q4 q = sectorTurnQ(ld4(quat, n, i), sw.turns, per);
    const GhostOut* o[3] = {&ox, &oy, &oxy};
    int bits = 0;
    for (int k = 0; k < 3; ++k) {
      // a particle outside its image rows keeps its own pose in every set,
      // and that copy is a ghost too when it lies outside the domain
      const bool turn = sectorImageOn(sw.x, g0, k);
      const V3 gp = turn ? sectorImagePos(sw.x, g0, per, k) : sw.x;
      if (ghostCellHash(gp, g0) < 0) continue;
      bits |= 1 << k;
      const int s = emitGhost(*o[k], i, gp, g0, 0);
      if (s < 0) continue;
      const Q4 gq = turn ? sectorRotQ(q, per, k) : q;
      const V3 gv = turn ? sectorRotV(v, per, k) : v;
      double* oq = o[k]->q + 4 * (size_t)s;
      oq[0] = gq.q0;
      oq[1] = gq.q1;
      oq[2] = gq.q2;
      oq[3] = gq.q3;
      double* ov = o[k]->vel + 3 * (size_t)s;
      ov[0] = gv.x;
      ov[1] = gv.y;
      ov[2] = gv.z;
    }
    if (atBuild) atBuild[i] = bits;
    return;
  }
  const int in = I.count > 0 ? I.particleInstance[i] : 0;
  const fdm_grid& g = I.count > 0 ? I.pgrid[in] : g0;
  const fdm_periodic pi = I.count > 0 ? I.periodic[in] : per;
  const long long off = I.count > 0 ? I.pcellOffset[in] : 0;
  const WrapStages w = wrapStages(ldp(pos, n, i), g, pi);
  int bits = 0;
  if (per.x && pi.x) {
    const V3 gp = ghostOf(w.afterX, g, 1);
    if (ghostCellHash(gp, g) >= 0) {
      emitGhost(ox, i, gp, g, off);
      bits |= 1;
    }
  }
  if (per.y && pi.y) {
    const V3 gp = ghostOf(w.afterY, g, 2);
    if (ghostCellHash(gp, g) >= 0) {
      emitGhost(oy, i, gp, g, off);
      bits |= 2;
    }
  }
  if (per.x && per.y && pi.x && pi.y) {
    const V3 gp = ghostOf(w.afterY, g, 3);
    if (ghostCellHash(gp, g) >= 0) {
      emitGhost(oxy, i, gp, g, off);
      bits |= 4;
    }
  }
  if (atBuild) atBuild[i] = bits;
}

// The periodic wrap of the stored positions (updateGhostSpatialGrid's
// launchUpdatePositionOutOfBoundaryXD / YD, PeriodicBoundary.h:225-231), run
// after the step's detection; gated by *gate (a ghost rebuild, which every
// wrap causes).
__global__ void kCommitWrap(double* pos, fdm_real* vel, fdm_real* quat, int n, fdm_grid g0,
                            fdm_periodic per, fdm_instances I, const int* gate) {
  if (gateOff(gate)) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (per.sector) {  // updateVelocityPositionOrientationOutOfSectorKernel
      const SectorWrap sw = sectorWrap(ldp(pos, n, i), g0, per);
      if (!sw.turns) continue;
      stp(pos, n, i, sw.x);
      st3(vel, n, i, sectorTurnV(ld3(vel, n, i), sw.turns, per));
      st4(quat, n, i, sectorTurnQ(ld4(quat, n, i), sw.turns, per));
      continue;
    }
    const int in = I.count > 0 ? I.particleInstance[i] : 0;
    const fdm_grid& g = I.count > 0 ? I.pgrid[in] : g0;
    const fdm_periodic pi = I.count > 0 ? I.periodic[in] : per;
    const V3 x = ldp(pos, n, i);
    const V3 w = wrapStages(x, g, pi).afterY;
    if (!sameV3(w, x)) stp(pos, n, i, w);
  }
}

// Steps without a ghost rebuild: no particle wrapped and every visibility bit
// is unchanged, so the kept ghosts are exactly this step's; refresh their
// positions and the reference's cell ids.
__global__ void kGhostRefresh(const double* pos, int n, fdm_grid g, fdm_periodic per, int mask,
                              const int* count, int capacity, const int* realId, double* gpos,
                              int* hashValue, const int* gate) {
  if (gate && *gate != 0) return;
  const int c = min(*count, capacity);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < c; k += gridDim.x * blockDim.x) {
    const int i = realId[k];
    const WrapStages w = wrapStages(ldp(pos, n, i), g, per);
    const V3 gp = ghostOf(mask == 1 ? w.afterX : w.afterY, g, mask);
    gpos[3 * k] = gp.x;
    gpos[3 * k + 1] = gp.y;
    gpos[3 * k + 2] = gp.z;
    hashValue[k] = ghostCellHash(gp, g);
  }
}

__global__ void kResetBox(int* box0, int* box1, int* box2, const int* gate) {
  if (gateOff(gate)) return;
  const int t = threadIdx.x;
  if (t < 6) {
    const int v = t < 3 ? INT_MAX : INT_MIN;
    if (box0) box0[t] = v;
    if (box1) box1[t] = v;
    if (box2) box2[t] = v;
  }
}

__global__ void kClampCount(int* count, int capacity, const int* gate) {
  if (gateOff(gate)) return;
  if (threadIdx.x == 0 && *count > capacity) *count = capacity;
}

}  // namespace fdm

using namespace fdm;

// The sector's rotation constants, evaluated as rotateVector's and
// rotateQuaternion's host branches evaluate them (myVec.h:143-181,
// myQua.h:131-175), for angles +pi/2, pi and -pi/2 about z.
extern "C" void fdm_sectorPeriodic(fdm_periodic* p) {
  if (!p) return;
  *p = fdm_periodic{};
  p->sector = 1;
  const double ang[3] = {0.5 * kPi, kPi, -0.5 * kPi};
  for (int k = 0; k < 3; ++k) {
    const volatile double thv = ang[k];  // evaluated by the C library at run time
    const double th = thv;
    const double theta2 = 0.0 * 0.0 + 0.0 * 0.0 + th * th;
    const double invTheta = 1.0 / std::sqrt(theta2);
    const double theta = theta2 * invTheta;
    p->angle[k] = th;
    p->A[k] = std::sin(theta) * invTheta;
    p->B[k] = (1.0 - std::cos(theta)) * (invTheta * invTheta);
    const double half = 0.5 * std::sqrt(theta2);
    p->halfCos[k] = std::cos(half);
    p->halfSin[k] = std::sin(half);
  }
}

extern "C" int fdm_launchUpdateSpatialGridHashStartEnd(fdm_hash_table table, const double* position,
                                                       int n, fdm_grid grid, void* workspace,
                                                       size_t workspaceBytes, hipStream_t stream) {
  if (n < 0 || grid.num <= 0) return FDM_EINVAL;
  if (n > 0)
    hipLaunchKernelGGL(kHashOnly, dim3(blocksFor(n, 256)), dim3(256), 0, stream, position, n, grid,
                       table.hashValue, table.sortKey, table.sortIndex);
  return buildStartEnd(table, n, nullptr, grid, workspace, workspaceBytes, stream);
}

extern "C" int fdm_launchUpdatePositionOutOfBoundaryXD(double* position, int n, fdm_grid grid,
                                                       hipStream_t stream) {
  if (n < 0) return FDM_EINVAL;
  if (n == 0) return FDM_OK;
  hipLaunchKernelGGL(kWrap, dim3(blocksFor(n, 256)), dim3(256), 0, stream, position, n, grid, 0);
  return launchStatus();
}

extern "C" int fdm_launchUpdatePositionOutOfBoundaryYD(double* position, int n, fdm_grid grid,
                                                       hipStream_t stream) {
  if (n < 0) return FDM_EINVAL;
  if (n == 0) return FDM_OK;
  hipLaunchKernelGGL(kWrap, dim3(blocksFor(n, 256)), dim3(256), 0, stream, position, n, grid, 1);
  return launchStatus();
}

extern "C" int fdm_launchCalculateGhostPosition(double* ghostPosition, const double* position, int n,
                                                fdm_grid grid, int axisMask, hipStream_t stream) {
  if (n < 0 || axisMask < 1 || axisMask > 3) return FDM_EINVAL;
  if (n == 0) return FDM_OK;
  hipLaunchKernelGGL(kGhostPosition, dim3(blocksFor(n, 256)), dim3(256), 0, stream, ghostPosition,
                     position, n, grid, axisMask);
  return launchStatus();
}

extern "C" int fdm_launchUpdateGhostSpatialGridHashStartEnd(fdm_hash_table table,
                                                            const double* ghostPosition, int n,
                                                            fdm_grid grid, void* workspace,
                                                            size_t workspaceBytes,
                                                            hipStream_t stream) {
  if (n < 0 || grid.num <= 0) return FDM_EINVAL;
  if (n > 0)
    hipLaunchKernelGGL(kGhostHashOnly, dim3(blocksFor(n, 256)), dim3(256), 0, stream, ghostPosition,
                       n, grid, table.hashValue, table.sortKey, table.sortIndex);
  return buildStartEnd(table, n, nullptr, grid, workspace, workspaceBytes, stream);
}

namespace {

const int* firstAtBuild(const fdm_ghosts* sets, const bool* on) {
  for (int s = 0; s < 3; ++s)
    if (on[s]) return sets[s].atBuild;
  return nullptr;
}

// scans of the flags, scatter of the visible ghosts (storing the wrapped
// positions) and one table per set; every kernel gated by *gate.
int buildGhosts(fdm_bodies particles, fdm_grid grid, fdm_periodic periodic, fdm_ghosts* sets,
                const bool* on, int* atBuild, const int* gate, unsigned* status, void* workspace,
                size_t workspaceBytes, hipStream_t stream, const fdm_instances* inst = nullptr) {
  fdm_instances none{};
  const fdm_instances I = inst ? *inst : none;
  const int n = particles.n;
  hipLaunchKernelGGL(kResetBox, dim3(1), dim3(64), 0, stream, on[0] ? sets[0].cellBox : nullptr,
                     on[1] ? sets[1].cellBox : nullptr, on[2] ? sets[2].cellBox : nullptr, gate);
  hipLaunchKernelGGL(kGhostVisibility, dim3(blocksFor(n, 256)), dim3(256), 0, stream, particles.x,
                     n, grid, periodic, I, on[0] ? sets[0].flag : nullptr,
                     on[1] ? sets[1].flag : nullptr, on[2] ? sets[2].flag : nullptr, gate);
  for (int s = 0; s < 3; ++s) {
    if (!on[s]) continue;
    int r = scanInts(sets[s].flag, sets[s].flag, (size_t)n, false, sets[s].count, workspace,
                     workspaceBytes, stream, gate);
    if (r != FDM_OK) return r;
  }
  GhostOut o[3];
  for (int s = 0; s < 3; ++s) {
    o[s].slot = sets[s].flag;
    o[s].gate = gate;
    o[s].realId = sets[s].realId;
    o[s].pos = sets[s].pos;
    o[s].hashValue = sets[s].table.hashValue;
    o[s].sortKey = sets[s].table.sortKey;
    o[s].sortIndex = sets[s].table.sortIndex;
    o[s].capacity = sets[s].capacity;
    o[s].status = status;
    o[s].q = sets[s].q;
    o[s].vel = sets[s].vel;
    if (periodic.sector && on[s] && (!o[s].q || !o[s].vel)) return FDM_EINVAL;
  }
  hipLaunchKernelGGL(kGhostScatter, dim3(blocksFor(n, 256)), dim3(256), 0, stream, particles.x,
                     particles.v, particles.q, n, grid, periodic, o[0], o[1], o[2], atBuild, I);
  for (int s = 0; s < 3; ++s) {
    if (!on[s]) continue;
    hipLaunchKernelGGL(kGhostBox, dim3(kBoxBlocks), dim3(kBoxThreads), 0, stream, sets[s].pos,
                       sets[s].realId, sets[s].count, sets[s].capacity, grid, I, sets[s].cellBox,
                       gate);
    hipLaunchKernelGGL(kClampCount, dim3(1), dim3(64), 0, stream, sets[s].count, sets[s].capacity,
                       gate);
    int r = buildStartEnd(sets[s].table, sets[s].capacity, sets[s].count, grid, workspace,
                          workspaceBytes, stream, gate);
    if (r != FDM_OK) return r;
  }
  return launchStatus();
}

}  // namespace

extern "C" int fdm_launchPeriodicGhosts(fdm_bodies particles, fdm_grid grid, fdm_periodic periodic,
                                        fdm_ghosts* sets, unsigned* status, void* workspace,
                                        size_t workspaceBytes, hipStream_t stream) {
  const int n = particles.n;
  if (n < 0 || !sets || !status) return FDM_EINVAL;
  if (n == 0 || !anyPeriodic(periodic)) return FDM_OK;
  if (periodic.sector && (periodic.x || periodic.y)) return FDM_EINVAL;
  const bool on[3] = {ghostSetOn(periodic, 0), ghostSetOn(periodic, 1), ghostSetOn(periodic, 2)};
  hipLaunchKernelGGL(kGhostFlags, dim3(blocksFor(n, 256)), dim3(256), 0, stream, particles.x, n,
                     grid, periodic, sets[0].flag, sets[1].flag, sets[2].flag, nullptr, nullptr);
  const int r = buildGhosts(particles, grid, periodic, sets, on, nullptr, nullptr, status, workspace,
                            workspaceBytes, stream);
  if (r != FDM_OK) return r;
  return fdm_launchPeriodicCommitWrap(particles, grid, periodic, nullptr, nullptr, stream);
}

extern "C" int fdm_launchPeriodicPrepare(fdm_bodies particles, fdm_grid grid, fdm_periodic periodic,
                                         fdm_ghosts* sets, const fdm_skin* skin,
                                         hipStream_t stream) {
  const int n = particles.n;
  if (n < 0 || !sets || !skin || !skin->ctl) return FDM_EINVAL;
  if (n == 0 || !anyPeriodic(periodic)) return FDM_OK;
  if (periodic.sector && (periodic.x || periodic.y)) return FDM_EINVAL;
  const bool on[3] = {ghostSetOn(periodic, 0), ghostSetOn(periodic, 1), ghostSetOn(periodic, 2)};
  const int* atBuild = firstAtBuild(sets, on);
  if (!atBuild) return FDM_EINVAL;
  hipLaunchKernelGGL(kGhostFlags, dim3(blocksFor(n, 256)), dim3(256), 0, stream, particles.x, n,
                     grid, periodic, sets[0].flag, sets[1].flag, sets[2].flag, atBuild, skin->ctl);
  return launchStatus();
}

extern "C" int fdm_launchPeriodicBuild(fdm_bodies particles, fdm_grid grid, fdm_periodic periodic,
                                       fdm_ghosts* sets, const int* gate, unsigned* status,
                                       void* workspace, size_t workspaceBytes,
                                       hipStream_t stream) {
  const int n = particles.n;
  if (n < 0 || !sets || !status) return FDM_EINVAL;
  if (n == 0 || !anyPeriodic(periodic)) return FDM_OK;
  if (periodic.sector && (periodic.x || periodic.y)) return FDM_EINVAL;
  const bool on[3] = {ghostSetOn(periodic, 0), ghostSetOn(periodic, 1), ghostSetOn(periodic, 2)};
  return buildGhosts(particles, grid, periodic, sets, on, (int*)firstAtBuild(sets, on), gate, status,
                     workspace, workspaceBytes, stream);
}

extern "C" int fdm_launchPeriodicBuildInstances(fdm_bodies particles, fdm_grid grid,
                                                fdm_periodic periodic, const fdm_instances* inst,
                                                fdm_ghosts* sets, unsigned* status,
                                                void* workspace, size_t workspaceBytes,
                                                hipStream_t stream) {
  const int n = particles.n;
  if (n < 0 || !sets || !status || !inst || inst->count <= 0) return FDM_EINVAL;
  if (n == 0 || !anyPeriodic(periodic)) return FDM_OK;
  if (periodic.sector) return FDM_EINVAL;  // the sector runs as one system
  const bool on[3] = {ghostSetOn(periodic, 0), ghostSetOn(periodic, 1), ghostSetOn(periodic, 2)};
  return buildGhosts(particles, grid, periodic, sets, on, (int*)firstAtBuild(sets, on), nullptr,
                     status, workspace, workspaceBytes, stream, inst);
}

extern "C" int fdm_launchPeriodicCommitWrap(fdm_bodies particles, fdm_grid grid,
                                            fdm_periodic periodic, const fdm_instances* inst,
                                            const int* gate, hipStream_t stream) {
  const int n = particles.n;
  if (n < 0) return FDM_EINVAL;
  if (n == 0 || !anyPeriodic(periodic)) return FDM_OK;
  if (periodic.sector && inst && inst->count > 0) return FDM_EINVAL;
  fdm_instances none{};
  const unsigned nb = std::min<unsigned>(blocksFor(n, 256), 2048);
  hipLaunchKernelGGL(kCommitWrap, dim3(nb), dim3(256), 0, stream, particles.x, particles.v,
                     particles.q, n, grid, periodic, inst ? *inst : none, gate);
  return launchStatus();
}

extern "C" int fdm_launchPeriodicRefresh(fdm_bodies particles, fdm_grid grid, fdm_periodic periodic,
                                         fdm_ghosts* sets, const int* gate, hipStream_t stream) {
  const int n = particles.n;
  if (n < 0 || !sets) return FDM_EINVAL;
  if (n == 0 || !anyPeriodic(periodic)) return FDM_OK;
  // sector images are rebuilt every step (their orientation and velocity
  // are not refreshed here)
  if (periodic.sector) return FDM_EINVAL;
  const bool on[3] = {ghostSetOn(periodic, 0), ghostSetOn(periodic, 1), ghostSetOn(periodic, 2)};
  for (int s = 0; s < 3; ++s) {
    if (!on[s]) continue;
    const unsigned nb = std::min<unsigned>(blocksFor(sets[s].capacity, 256), 256);
    hipLaunchKernelGGL(kGhostRefresh, dim3(nb), dim3(256), 0, stream, particles.x, n, grid,
                       periodic, s == 0 ? 1 : s == 1 ? 2 : 3, sets[s].count, sets[s].capacity,
                       sets[s].realId, sets[s].pos, sets[s].table.hashValue, gate);
  }
  return launchStatus();
}
