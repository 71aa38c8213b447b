// Rigid-body integration: half-kick of v and w, drift of x and q, and the
// fused first half of a step that also computes the cell id and sort key of
// every body. One thread per body; planar FP32 SoA in and out, FP64 inside.
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
// buildHashStartEnd.cu:18-24).
__global__ void kKickDriftHash(fdm_bodies b, const fdm_template* tmpl, V3 g, double dt, int kick,
                               int drift, fdm_grid grid, int* hashValue, unsigned* sortKey,
                               int* sortIndex) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  if (kick) kickOne(b, tmpl, i, g, 0.5 * dt);
  V3 x;
  if (drift) x = driftOne(b, i, dt);
  else x = ldp(b.x, b.n, i);
  const int h = cellHash(x, grid);
  hashValue[i] = h;
  sortKey[i] = (unsigned)(h + 1);
  sortIndex[i] = i;
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
                                       fdm_hash_table table, hipStream_t stream) {
  if (bodies.n < 0) return FDM_EINVAL;
  if (bodies.n == 0) return FDM_OK;
  hipLaunchKernelGGL(kKickDriftHash, dim3(blocksFor(bodies.n, 256)), dim3(256), 0, stream, bodies,
                     geo.info, fromD3(gravity), dt, kick, drift, grid, table.hashValue,
                     table.sortKey, table.sortIndex);
  return launchStatus();
}
