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

// Records each body's grid state after a list build.
__global__ void kCommitState(const double* x, int n, fdm_grid grid, int* atBuild, const int* gate) {
  if (gateOff(gate)) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    atBuild[i] = bodyState(ldp(x, n, i), grid);
}

// Rebuild decision of one step (one thread). The displacement bound grows by
// dt times the largest motion rate, plus 1e-6 rho per step for the rounding
// of orientations to FP32 (a relative 2^-24 per component moves a point at
// distance rho by well under 1e-6 rho). Ghost lists are also rebuilt after
// a visibility change or a wrap; a wrap (positions jump by a period)
// rebuilds the next step's lists as well.
__global__ void kSkinDecide(fdm_skin s) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int* c = s.ctl;
  const float rate = __int_as_float(c[FDM_CTL_MAX_RATE]);
  double d = *s.disp + s.dt * (double)rate * (1.0 + 1e-6) + 1e-6 * s.rhoMax;
  const bool rebuild = c[FDM_CTL_FORCE] != 0 || c[FDM_CTL_PENDING] != 0 ||
                       c[FDM_CTL_VALID_CHANGED] != 0 || !(2.0 * d <= s.skin);
  const bool wrapped = c[FDM_CTL_WRAPPED] != 0;
  const bool ghost = rebuild || wrapped || c[FDM_CTL_GHOST_CHANGED] != 0;
  if (rebuild) {
    d = 0.0;
    c[FDM_CTL_COUNT] += 1;
  }
  if (ghost) c[FDM_CTL_GHOST_COUNT] += 1;
  *s.disp = d;
  c[FDM_CTL_REBUILD] = rebuild ? 1 : 0;
  c[FDM_CTL_REBUILD_GHOST] = ghost ? 1 : 0;
  c[FDM_CTL_PENDING] = wrapped ? 1 : 0;
  c[FDM_CTL_MAX_RATE] = 0;
  c[FDM_CTL_FORCE] = 0;
  c[FDM_CTL_VALID_CHANGED] = 0;
  c[FDM_CTL_WRAPPED] = 0;
  c[FDM_CTL_GHOST_CHANGED] = 0;
  const int seq = c[FDM_CTL_SEQ] + 1;
  c[FDM_CTL_SEQ] = seq;
  if (s.hostDecision) {
    volatile int* h = s.hostDecision;
    h[1] = rebuild ? 1 : 0;
    h[2] = ghost ? 1 : 0;
    __threadfence_system();
    h[0] = seq;
    __threadfence_system();
  }
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

}  // namespace fdm

using namespace fdm;

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

extern "C" int fdm_launchSkinDecide(const fdm_skin* skin, hipStream_t stream) {
  if (!skin || !skin->ctl || !skin->disp) return FDM_EINVAL;
  hipLaunchKernelGGL(kSkinDecide, dim3(1), dim3(64), 0, stream, *skin);
  return launchStatus();
}

extern "C" int fdm_launchCommitBodyState(fdm_bodies bodies, fdm_grid grid, fdm_hash_table table,
                                         const fdm_skin* skin, hipStream_t stream) {
  if (bodies.n < 0 || !skin || !table.validAtBuild) return FDM_EINVAL;
  if (bodies.n == 0) return FDM_OK;
  const unsigned nb = std::min<unsigned>(blocksFor(bodies.n, 256), kIntegrateBlocks);
  hipLaunchKernelGGL(kCommitState, dim3(nb), dim3(256), 0, stream, bodies.x, bodies.n, grid,
                     table.validAtBuild, skin->ctl + FDM_CTL_REBUILD);
  return launchStatus();
}

extern "C" int fdm_launchCarryPairs(const fdm_pairs* from, const fdm_pairs* to, int n,
                                    const int* gate, hipStream_t stream) {
  if (!from || !to || n < 0) return FDM_EINVAL;
  hipLaunchKernelGGL(kCarryPairs, dim3(kIntegrateBlocks), dim3(256), 0, stream, *from, *to, n, gate);
  return launchStatus();
}
