// Rigid-body integration: half-kick of v and w, drift of x and q, and the
// fused first half of a step that also computes the cell id and sort key of
// every body. One thread per body; planar FP32 SoA in and out, FP64 inside.
#include <algorithm>

#include "fdm_device.hpp"
#include "fdm_launch.hpp"

namespace fdm {

// particleVelocityAngularVelocityIntegrationKernel (particleIntegrationKernel.cu:3-32)
__device__ __forceinline__ void kickOne(const fdm_bodies& b, const fdm_template* tmpl, int i,
                                        V3 g, double h) {
  const int n = b.n;
  const fdm_template& t = tmpl[b.geo[i]];
  const double invM = t.invMass;
  if (nearZero(invM)) return;
  const V3 F = ld3(b.F, n, i);
  const V3 v = ld3(b.v, n, i);
  st3(b.v, n, i, vadd(v, vmul(vadd(vmul(F, invM), g), h)));
  if (spinFixed(t)) return;
  const V3 w = ld3(b.w, n, i);
  const Q4 q = ld4(b.q, n, i);
  const V3 Iw = smul(srotate(q, lds6(t.inertia)), w);
  const V3 T = ld3(b.T, n, i);
  st3(b.w, n, i, vadd(w, vmul(smul(srotate(q, lds6(t.invInertia)), vsub(T, vcross(w, Iw))), h)));
}

// particlePositionOrientationIntegrationKernel (particleIntegrationKernel.cu:34-48)
__device__ __forceinline__ V3 driftOne(const fdm_bodies& b, int i, double dt) {
  const int n = b.n;
  const V3 x = vadd(ldp(b.x, n, i), vmul(ld3(b.v, n, i), dt));
  stp(b.x, n, i, x);
  st4(b.q, n, i, qintegrate(ld4(b.q, n, i), ld3(b.w, n, i), dt));
  return x;
}

__global__ void kKick(fdm_bodies b, const fdm_template* tmpl, V3 g, double h) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  kickOne(b, tmpl, i, g, h);
}

__global__ void kDrift(fdm_bodies b, double dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  driftOne(b, i, dt);
}

// Half-kick, drift, and the body's cell id (calculateHashKernel,
// mySpatialGrid.cu:4-26) with its sort key (hash + 1, so that -1 sorts first
// as in a signed sort) and the identity index (setHashIndex,
// buildHashStartEnd.cu:18-24). With list reuse (ctl != null) it also takes
// the largest motion rate |v| + |w| rho of the drift (one atomic per block)
// and flags bodies whose grid state differs from the state at the last
// build. Grid-stride, at most kIntegrateBlocks blocks.
constexpr int kIntegrateBlocks = 1024;

__global__ void __launch_bounds__(256) kKickDriftHash(fdm_bodies b, const fdm_template* tmpl, V3 g,
                                                      double dt, int kick, int drift, fdm_grid grid,
                                                      int* hashValue, unsigned* sortKey,
                                                      int* sortIndex, const int* atBuild, int* ctl) {
  __shared__ float waveMax[4];
  float rate = 0.f;
  bool changed = false;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += gridDim.x * blockDim.x) {
    if (kick) kickOne(b, tmpl, i, g, 0.5 * dt);
    V3 x;
    if (drift) x = driftOne(b, i, dt);
    else x = ldp(b.x, b.n, i);
    const int h = cellHash(x, grid);
    hashValue[i] = h;
    sortKey[i] = (unsigned)(h + 1);
    sortIndex[i] = i;
    if (ctl) {
      if (drift) {
        const fdm_template& t = tmpl[b.geo[i]];
        const double rho = fmax(t.radius, t.boxReach);
        const double r = vlen(ld3(b.v, b.n, i)) + vlen(ld3(b.w, b.n, i)) * rho;
        rate = fmaxf(rate, (float)r * (1.0f + 1e-6f));
      }
      if (atBuild) {
        const int st = bodyState(x, grid);
        changed |= st == 2 || st != atBuild[i];
      }
    }
  }
  if (!ctl) return;
  if (ballot64(changed) != 0ull && laneId() == 0) ctl[FDM_CTL_VALID_CHANGED] = 1;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) rate = fmaxf(rate, __shfl_xor(rate, off, 64));
  if (laneId() == 0) waveMax[threadIdx.x >> 6] = rate;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = waveMax[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmaxf(m, waveMax[w]);
    if (m > 0.f) atomicMax(ctl + FDM_CTL_MAX_RATE, __float_as_int(m));
  }
}

// A list build's start: each body's grid state, and the unsorted (key,
// index) lists the build sorts (key hash + 1, 0 outside the grid).
__global__ void kCommitState(const double* x, int n, fdm_grid grid, int* atBuild,
                             unsigned* sortKey, int* sortIndex, const int* gate) {
  if (gateOff(gate)) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const V3 p = ldp(x, n, i);
    atBuild[i] = bodyState(p, grid);
    sortKey[i] = (unsigned)(cellHash(p, grid) + 1);
    sortIndex[i] = i;
  }
}

// The same for bodies of several systems: keys in the global cell space.
__global__ void kCommitStateInstances(const double* x, int n, fdm_instances I, int isWall,
                                      int* atBuild, unsigned* sortKey, int* sortIndex,
                                      const int* gate) {
  if (gateOff(gate)) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int in = isWall ? I.wallInstance[i] : I.particleInstance[i];
    const fdm_grid& g = isWall ? I.wgrid[in] : I.pgrid[in];
    const V3 p = ldp(x, n, i);
    atBuild[i] = bodyState(p, g);
    sortKey[i] = instanceKey(cellHash(p, g), g, isWall ? I.wcellOffset[in] : I.pcellOffset[in]);
    sortIndex[i] = i;
  }
}

// Rebuild decision of one step (one thread). The displacement bound grows by
// dt times the largest motion rate, plus 1e-6 rho per step for the rounding
// of orientations to FP32 (a relative 2^-24 per component moves a point at
// distance rho by well under 1e-6 rho). Ghost lists are also rebuilt after
// a visibility change or a wrap; a wrap (positions jump by a period)
// rebuilds the next step's lists as well. One thread of one workgroup runs
// it, after every other writer of the control words has finished (earlier
// kernels of the stream); `rate` and `flags` (bit 0 grid state changed, 1
// visibility changed, 2 wrapped) add this kernel's own findings. The control
// words are read together (plain loads: the cache is fresh at kernel start),
// and the host's copy of the decision is one word, (number << 2) | rebuild
// << 1 | ghost rebuild, so it needs no ordering fence.
__device__ void decideStep(const fdm_skin& s, float rateIn, int flagsIn) {
  int* c = s.ctl;
  int w[FDM_CTL_WORDS];
#pragma unroll
  for (int k = 0; k < FDM_CTL_WORDS; ++k) w[k] = c[k];
  const double disp = *s.disp;
  const float rate = fmaxf(__int_as_float(w[FDM_CTL_MAX_RATE]), rateIn);
  double d = disp + s.dt * (double)rate * (1.0 + 1e-6) + 1e-6 * s.rhoMax;
  const bool rebuild = w[FDM_CTL_FORCE] != 0 || w[FDM_CTL_PENDING] != 0 ||
                       w[FDM_CTL_VALID_CHANGED] != 0 || (flagsIn & 1) || !(2.0 * d <= s.skin);
  const bool wrapped = w[FDM_CTL_WRAPPED] != 0 || (flagsIn & 4);
  const bool ghost = rebuild || wrapped || w[FDM_CTL_GHOST_CHANGED] != 0 || (flagsIn & 2);
  if (rebuild) {
    d = 0.0;
    c[FDM_CTL_COUNT] = w[FDM_CTL_COUNT] + 1;
  }
  if (ghost) c[FDM_CTL_GHOST_COUNT] = w[FDM_CTL_GHOST_COUNT] + 1;
  *s.disp = d;
  c[FDM_CTL_REBUILD] = rebuild ? 1 : 0;
  c[FDM_CTL_REBUILD_GHOST] = ghost ? 1 : 0;
  c[FDM_CTL_KEEP] = (rebuild || ghost) ? 0 : 1;
  c[FDM_CTL_PENDING] = wrapped ? 1 : 0;
  c[FDM_CTL_MAX_RATE] = 0;
  c[FDM_CTL_FORCE] = 0;
  c[FDM_CTL_VALID_CHANGED] = 0;
  c[FDM_CTL_WRAPPED] = 0;
  c[FDM_CTL_GHOST_CHANGED] = 0;
  const int seq = w[FDM_CTL_SEQ] + 1;
  c[FDM_CTL_SEQ] = seq;
  if (s.hostDecision)
    *(volatile int*)(s.hostDecision + (seq & (FDM_DECISION_RING - 1))) =
        (int)(((unsigned)seq << 2) | (rebuild ? 2u : 0u) | (ghost ? 1u : 0u));
}

__global__ void kSkinDecide(fdm_skin s) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  decideStep(s, 0.f, 0);
}

struct StepBeginK {
  fdm_step_begin_args a;
  fdm_skin s;
  float* blockRate;  // [gridDim.x] largest motion rate of each workgroup
  int* blockFlags;   // [gridDim.x] bit 0 grid state changed, 1 visibility changed, 2 wrapped
};

// Phase A of a step (one thread per particle, then per wall). Each workgroup
// leaves its largest motion rate and its change flags in a slot of its own
// (no atomics on shared words); kStepDecide folds the slots.
// One body of phase A. Everything it reads from the launch arguments is
// copied into registers first (selecting between argument structs by
// reference would make the compiler copy them to per-lane scratch).
// MANY: bodies of several systems (per-body grids and periodic flags); the
// single-system kernel holds its grid in scalar registers.
template <bool WALL, bool MANY>
__device__ __forceinline__ void stepBeginBody(const fdm_step_begin_args& a, int j, float& rate,
                                              int& flags) {
  const fdm_bodies b = WALL ? a.walls : a.particles;
  const fdm_template* tmpl = WALL ? a.wgeo.info : a.pgeo.info;
  const fdm_hash_table t = WALL ? a.wtable : a.ptable;
  const bool many = MANY;
  const int in = !many ? 0 : (WALL ? a.inst.wallInstance[j] : a.inst.particleInstance[j]);
  const fdm_grid grid =
      !many ? (WALL ? a.wgrid : a.pgrid) : (WALL ? a.inst.wgrid[in] : a.inst.pgrid[in]);
  const fdm_periodic perI = many ? a.inst.periodic[in] : a.periodic;
  const bool ghosts = !WALL && anyPeriodic(a.periodic);
  const bool kick = !WALL && a.kick;
  const int n = b.n;
  // Every input is loaded before anything is stored: the arrays may alias as
  // far as the compiler knows, so a load placed after a store could not be
  // issued before it. One round of loads (the torque only where the spin
  // changes), then the half-kick and the drift in registers with the
  // roundings of the stored state (kickOne, driftOne), then the stores.
  // (the stored FP32 values are held as loaded and widened where used)
  const int geo = b.geo[j];
  const V3 x0 = ldp(b.x, n, j);
  R3 v = ld3s(b.v, n, j), w = ld3s(b.w, n, j);
  const R4 q0 = ld4s(b.q, n, j);
  R3 F = {0, 0, 0};
  if (kick) F = ld3s(b.F, n, j);
  const int atBuild = t.validAtBuild[j];
  const int ghostAtBuild = ghosts ? a.ghostAtBuild[j] : 0;
  const fdm_template& tp = tmpl[geo];
  bool vNew = false, wNew = false;
  if (kick && !nearZero(tp.invMass)) {
    // particleVelocityAngularVelocityIntegrationKernel (particleIntegrationKernel.cu:3-32)
    const double h = 0.5 * a.dt;
    v = narrow3(vadd(wide(v), vmul(vadd(vmul(wide(F), tp.invMass), fromD3(a.gravity)), h)));
    vNew = true;
    if (!spinFixed(tp)) {
      const V3 wd = wide(w);
      const Q4 qd = wide(q0);
      const V3 Iw = smul(srotate(qd, lds6(tp.inertia)), wd);
      const V3 T = ld3(b.T, n, j);
      w = narrow3(vadd(wd, vmul(smul(srotate(qd, lds6(tp.invInertia)), vsub(T, vcross(wd, Iw))), h)));
      wNew = true;
    }
  }
  V3 x = x0;
  if (vNew) st3s(b.v, n, j, v);
  if (wNew) st3s(b.w, n, j, w);
  if (a.drift) {
    // particlePositionOrientationIntegrationKernel (particleIntegrationKernel.cu:34-48)
    x = vadd(x0, vmul(wide(v), a.dt));
    stp(b.x, n, j, x);
    st4(b.q, n, j, qintegrate(wide(q0), wide(w), a.dt));
  }
  t.hashValue[j] = cellHash(x, grid);
  if (a.drift) {
    const double rho = fmax(tp.radius, tp.boxReach);
    const double r = vlen(wide(v)) + vlen(wide(w)) * rho;
    rate = (float)r * (1.0f + 1e-6f);
  }
  const int st = bodyState(x, grid);
  if (st == 2 || st != atBuild) flags |= 1;
  if (ghosts) {
    // sets exist where any instance is periodic; an instance contributes
    // only to the sets of its own periodic axes
    const int vb = visibilityBits(x, grid, a.periodic, perI);
    const int bits = vb & 7;
    for (int k = 0; k < 3; ++k)
      if (a.ghostFlag[k]) a.ghostFlag[k][j] = (bits >> k) & 1;
    if (bits != ghostAtBuild) flags |= 2;
    if (vb & 8) flags |= 4;
  }
}

template <bool MANY>
__global__ void __launch_bounds__(256) kStepBegin(StepBeginK k) {
  if (gateOff(k.a.gate)) return;
  const fdm_step_begin_args& a = k.a;
  __shared__ float waveMax[4];
  const int nP = a.particles.n, nW = a.walls.n;
  float rate = 0.f;
  int flags = 0;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nP) stepBeginBody<false, MANY>(a, i, rate, flags);
  else if (i < nP + nW) stepBeginBody<true, MANY>(a, i - nP, rate, flags);
  const int f1 = __syncthreads_or(flags & 1), f2 = __syncthreads_or(flags & 2),
            f4 = __syncthreads_or(flags & 4);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) rate = fmaxf(rate, __shfl_xor(rate, off, 64));
  if (laneId() == 0) waveMax[threadIdx.x >> 6] = rate;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = waveMax[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmaxf(m, waveMax[w]);
    k.blockRate[blockIdx.x] = m;
    k.blockFlags[blockIdx.x] = (f1 ? 1 : 0) | (f2 ? 2 : 0) | (f4 ? 4 : 0);
  }
}

// Clears the given counter arrays, folds the workgroup slots of kStepBegin
// into the control words and takes the list decision (one workgroup).
__global__ void __launch_bounds__(1024) kStepDecide(StepBeginK k, int nblocks) {
  if (gateOff(k.a.gate)) return;
  __shared__ float rmax[16];
  __shared__ int fl[16];
  for (int q = 0; q < 8; ++q)
    if (k.a.clearCounters[q])
      for (int i = threadIdx.x; i < FDM_REC_COUNTERS; i += blockDim.x) k.a.clearCounters[q][i] = 0;
  float r = 0.f;
  int f = 0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
    r = fmaxf(r, k.blockRate[b]);
    f |= k.blockFlags[b];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    r = fmaxf(r, __shfl_xor(r, off, 64));
    f |= __shfl_xor(f, off, 64);
  }
  if (laneId() == 0) {
    rmax[threadIdx.x >> 6] = r;
    fl[threadIdx.x >> 6] = f;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
    r = fmaxf(r, rmax[w]);
    f |= fl[w];
  }
  r = fmaxf(r, rmax[0]);
  f |= fl[0];
  decideStep(k.s, r, f);
}

// Copies the candidate lists of the previous step when no rebuild happens.
__global__ void kCarryPairs(fdm_pairs from, fdm_pairs to, int n, const int* gate) {
  if (gate && *gate != 0) return;
  const int stride = gridDim.x * blockDim.x;
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  for (int i = t0; i <= n; i += stride) to.count[i] = from.count[i];
  const int total = *from.total;
  if (t0 == 0) *to.total = total;
  const int np = min(total, min(from.capacity, to.capacity));
  for (int p = t0; p < np; p += stride) {
    to.pairA[p] = from.pairA[p];
    to.pairT[p] = from.pairT[p];
    to.pairG[p] = from.pairG[p];
  }
}

// addGlobalDampingForceTorque (globalDamping.cu:4-16) on FP32 forces
__device__ __forceinline__ void dampOne(fdm_real* F, fdm_real* T, const fdm_real* v, const fdm_real* w, int n,
                                        int i, double c) {
  const V3 f = ld3(F, n, i), t = ld3(T, n, i);
  st3(F, n, i, vsub(f, vmul(vnormalize(ld3(v, n, i)), c * vlen(f))));
  st3(T, n, i, vsub(t, vmul(vnormalize(ld3(w, n, i)), c * vlen(t))));
}

__global__ void kGlobalDamping(fdm_bodies b, double c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  dampOne(b.F, b.T, b.v, b.w, b.n, i, c);
}

// The same on the caller's launch shape (grid-stride over the bodies).
__global__ void kGlobalDampingArrays(fdm_real* F, fdm_real* T, const fdm_real* v, const fdm_real* w, int n,
                                     double c) {
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dampOne(F, T, v, w, n, i, c);
}

// addGlobalConstantForceTorque (globalDamping.cu:38-49): both vectors go to
// the force
__global__ void kConstantForce(fdm_bodies b, V3 F, V3 T, int index) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  V3 f = ld3(b.F, b.n, index);
  f = roundState(vadd(f, F));
  st3(b.F, b.n, index, vadd(f, T));
}

}  // namespace fdm

using namespace fdm;

extern "C" int fdm_launchAddGlobalDampingForceTorque(fdm_bodies particles, double c,
                                                     hipStream_t stream) {
  if (particles.n < 0) return FDM_EINVAL;
  if (particles.n == 0) return FDM_OK;
  hipLaunchKernelGGL(kGlobalDamping, dim3(blocksFor(particles.n, 256)), dim3(256), 0, stream,
                     particles, c);
  return launchStatus();
}

extern "C" int fdm_launchAddGlobalDampingForceTorqueArrays(fdm_real* force, fdm_real* torque,
                                                           const fdm_real* velocity,
                                                           const fdm_real* angularVelocity,
                                                           double c, size_t num, size_t gridD,
                                                           size_t blockD, hipStream_t stream) {
  if (num == 0) return FDM_OK;
  if (num > (size_t)INT32_MAX || !force || !torque || !velocity || !angularVelocity)
    return FDM_EINVAL;
  if (gridD == 0 || blockD == 0 || blockD > 1024 || gridD > (size_t)INT32_MAX) return FDM_EINVAL;
  hipLaunchKernelGGL(kGlobalDampingArrays, dim3((unsigned)gridD), dim3((unsigned)blockD), 0, stream,
                     force, torque, velocity, angularVelocity, (int)num, c);
  return launchStatus();
}

extern "C" int fdm_launchAddGlobalConstantForceTorque(fdm_bodies particles, fdm_double3 F,
                                                      fdm_double3 T, int index,
                                                      hipStream_t stream) {
  if (index < 0 || index >= particles.n) return FDM_EINVAL;
  hipLaunchKernelGGL(kConstantForce, dim3(1), dim3(64), 0, stream, particles, fromD3(F), fromD3(T),
                     index);
  return launchStatus();
}

extern "C" int fdm_launchParticleVelocityAngularVelocityIntegration(fdm_bodies bodies,
                                                                    fdm_geometry geo,
                                                                    fdm_double3 gravity, double h,
                                                                    hipStream_t stream) {
  if (bodies.n < 0) return FDM_EINVAL;
  if (bodies.n == 0) return FDM_OK;
  hipLaunchKernelGGL(kKick, dim3(blocksFor(bodies.n, 256)), dim3(256), 0, stream, bodies, geo.info,
                     fromD3(gravity), h);
  return launchStatus();
}

extern "C" int fdm_launchParticlePositionOrientationIntegration(fdm_bodies bodies, double dt,
                                                                hipStream_t stream) {
  if (bodies.n < 0) return FDM_EINVAL;
  if (bodies.n == 0) return FDM_OK;
  hipLaunchKernelGGL(kDrift, dim3(blocksFor(bodies.n, 256)), dim3(256), 0, stream, bodies, dt);
  return launchStatus();
}

extern "C" int fdm_launchKickDriftHash(fdm_bodies bodies, fdm_geometry geo, fdm_double3 gravity,
                                       double dt, int kick, int drift, fdm_grid grid,
                                       fdm_hash_table table, const fdm_skin* skin,
                                       hipStream_t stream) {
  if (bodies.n < 0) return FDM_EINVAL;
  if (bodies.n == 0) return FDM_OK;
  const unsigned nb = std::min<unsigned>(blocksFor(bodies.n, 256), kIntegrateBlocks);
  hipLaunchKernelGGL(kKickDriftHash, dim3(nb), dim3(256), 0, stream, bodies, geo.info,
                     fromD3(gravity), dt, kick, drift, grid, table.hashValue, table.sortKey,
                     table.sortIndex, skin ? (const int*)table.validAtBuild : nullptr,
                     skin ? skin->ctl : nullptr);
  return launchStatus();
}

extern "C" size_t fdm_stepBeginWorkspaceBytes(size_t bodies) {
  const size_t nb = std::max<size_t>(1, (bodies + 255) / 256);
  return alignUp(nb * sizeof(float)) + alignUp(nb * sizeof(int));
}

namespace {

int stepBeginSetup(const fdm_step_begin_args* args, const fdm_skin* skin, void* workspace,
                   size_t workspaceBytes, StepBeginK& k, unsigned& nb) {
  if (!args || !skin || !skin->ctl || !skin->disp) return FDM_EINVAL;
  const fdm_step_begin_args& a = *args;
  if (a.particles.n < 0 || a.walls.n < 0) return FDM_EINVAL;
  if (a.particles.n > 0 && !a.ptable.validAtBuild) return FDM_EINVAL;
  if (a.walls.n > 0 && !a.wtable.validAtBuild) return FDM_EINVAL;
  if (anyPeriodic(a.periodic)) {
    if (!a.ghostAtBuild) return FDM_EINVAL;
    if (a.periodic.sector && (a.periodic.x || a.periodic.y || a.inst.count > 0)) return FDM_EINVAL;
  }
  const long long total = (long long)a.particles.n + a.walls.n;
  if (workspaceBytes < fdm_stepBeginWorkspaceBytes((size_t)total)) return FDM_EINVAL;
  nb = std::max<unsigned>(1, blocksFor(total, 256));
  Carve cv(workspace, workspaceBytes);
  k.a = a;
  k.s = *skin;
  k.blockRate = cv.take<float>(nb);
  k.blockFlags = cv.take<int>(nb);
  return cv.ok ? FDM_OK : FDM_EINVAL;
}

}  // namespace

extern "C" int fdm_launchStepBegin(const fdm_step_begin_args* args, const fdm_skin* skin,
                                   void* workspace, size_t workspaceBytes, hipStream_t stream) {
  StepBeginK k;
  unsigned nb = 0;
  const int r = stepBeginSetup(args, skin, workspace, workspaceBytes, k, nb);
  if (r != FDM_OK) return r;
  if (args->inst.count > 0) hipLaunchKernelGGL(kStepBegin<true>, dim3(nb), dim3(256), 0, stream, k);
  else hipLaunchKernelGGL(kStepBegin<false>, dim3(nb), dim3(256), 0, stream, k);
  return launchStatus();
}

extern "C" int fdm_launchStepDecide(const fdm_step_begin_args* args, const fdm_skin* skin,
                                    void* workspace, size_t workspaceBytes, hipStream_t stream) {
  StepBeginK k;
  unsigned nb = 0;
  const int r = stepBeginSetup(args, skin, workspace, workspaceBytes, k, nb);
  if (r != FDM_OK) return r;
  hipLaunchKernelGGL(kStepDecide, dim3(1), dim3(1024), 0, stream, k, (int)nb);
  return launchStatus();
}

extern "C" int fdm_launchSkinDecide(const fdm_skin* skin, hipStream_t stream) {
  if (!skin || !skin->ctl || !skin->disp) return FDM_EINVAL;
  hipLaunchKernelGGL(kSkinDecide, dim3(1), dim3(64), 0, stream, *skin);
  return launchStatus();
}

extern "C" int fdm_launchCommitBodyState(fdm_bodies bodies, fdm_grid grid, fdm_hash_table table,
                                         const fdm_skin* skin, hipStream_t stream) {
  if (bodies.n < 0 || !skin || !table.validAtBuild || !table.sortKey || !table.sortIndex)
    return FDM_EINVAL;
  if (bodies.n == 0) return FDM_OK;
  const unsigned nb = std::min<unsigned>(blocksFor(bodies.n, 256), kIntegrateBlocks);
  hipLaunchKernelGGL(kCommitState, dim3(nb), dim3(256), 0, stream, bodies.x, bodies.n, grid,
                     table.validAtBuild, table.sortKey, table.sortIndex,
                     skin->ctl + FDM_CTL_REBUILD);
  return launchStatus();
}

extern "C" int fdm_launchCommitBodyStateInstances(fdm_bodies bodies, const fdm_instances* inst,
                                                  int isWall, fdm_hash_table table,
                                                  const fdm_skin* skin, hipStream_t stream) {
  if (bodies.n < 0 || !inst || inst->count <= 0 || !skin || !table.validAtBuild ||
      !table.sortKey || !table.sortIndex)
    return FDM_EINVAL;
  if (bodies.n == 0) return FDM_OK;
  const unsigned nb = std::min<unsigned>(blocksFor(bodies.n, 256), kIntegrateBlocks);
  hipLaunchKernelGGL(kCommitStateInstances, dim3(nb), dim3(256), 0, stream, bodies.x, bodies.n,
                     *inst, isWall, table.validAtBuild, table.sortKey, table.sortIndex,
                     skin->ctl + FDM_CTL_REBUILD);
  return launchStatus();
}

extern "C" int fdm_launchCarryPairs(const fdm_pairs* from, const fdm_pairs* to, int n,
                                    const int* gate, hipStream_t stream) {
  if (!from || !to || n < 0) return FDM_EINVAL;
  hipLaunchKernelGGL(kCarryPairs, dim3(kIntegrateBlocks), dim3(256), 0, stream, *from, *to, n, gate);
  return launchStatus();
}
