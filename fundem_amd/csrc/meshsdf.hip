// Signed distance of a closed triangle mesh sampled on a background grid, or
// at a list of points (LevelSetObject::TriangleMeshParticle,
// LevelSetObjectGenerator.h:382-951, with the grid sampling of
// Base::buildGrid, LevelSetObjectGenerator.h:122-224).
//
// The reference walks a SAH-BVH for the closest point and sums the solid
// angles of all triangles for the inside test, one point at a time on the
// host. Here both reductions run over every triangle: the workgroup stages
// 256 triangles at a time in LDS (SoA) and eight lanes share a point, each
// folding every eighth triangle, before the partials are combined: a grid of
// 70k nodes then runs 560k lanes, enough to keep every SIMD of the chip busy.
// A triangle whose bounding box is farther than the lane's best distance so
// far skips the closest-point test (the BVH's pruning, per lane); the solid
// angle is needed for every triangle. The closest distance is the
// exact minimum over triangles of the FP64 point-triangle distance; ties go
// to the lowest triangle index. Compute bound (FP64 VALU), not HBM: the
// triangles are read once per workgroup from L2.
#include <cfloat>
#include <climits>

#include "fdm_device.hpp"
#include "fdm_launch.hpp"

namespace fdm {

namespace {

constexpr int kMeshTile = 256;

// Closest point of triangle abc to p (Ericson's Voronoi-region walk), in the
// operation order of closestPointOnTriangle (LevelSetObjectGenerator.h:454-495).
__device__ __forceinline__ V3 closestOnTriangle(V3 p, V3 a, V3 b, V3 c) {
  const V3 ab = vsub(b, a), ac = vsub(c, a), ap = vsub(p, a);
  const double d1 = vdot(ab, ap), d2 = vdot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return a;
  const V3 bp = vsub(p, b);
  const double d3 = vdot(ab, bp), d4 = vdot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return b;
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return vadd(a, vmul(ab, d1 / (d1 - d3)));
  const V3 cp = vsub(p, c);
  const double d5 = vdot(ab, cp), d6 = vdot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return c;
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return vadd(a, vmul(ac, d2 / (d2 - d6)));
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0)
    return vadd(b, vmul(vsub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
  const double denom = 1.0 / (va + vb + vc);
  const double v = vb * denom, w = vc * denom;
  return vadd(vadd(a, vmul(ab, v)), vmul(ac, w));
}

// Solid angle of triangle abc seen from p (Van Oosterom and Strackee), as
// solidAngleFromPointToTriangle (LevelSetObjectGenerator.h:758-782).
__device__ __forceinline__ double solidAngle(V3 p, V3 a, V3 b, V3 c) {
  const V3 r0 = vsub(a, p), r1 = vsub(b, p), r2 = vsub(c, p);
  const double l0 = vlen(r0), l1 = vlen(r1), l2 = vlen(r2);
  if (l0 < 1e-30 || l1 < 1e-30 || l2 < 1e-30) return 0.0;
  const double num = vdot(r0, vcross(r1, r2));
  const double den = l0 * l1 * l2 + vdot(r0, r1) * l2 + vdot(r1, r2) * l0 + vdot(r2, r0) * l1;
  return 2.0 * atan2(num, den);
}

struct MeshPoints {
  // grid mode (points == nullptr): node (i, j, k) at origin + (i, j, k) h
  double ox, oy, oz, h;
  int nx, ny, nz;
  // point mode
  const double* points;  // [3*count]
  long long count;
};

__device__ __forceinline__ V3 pointOf(const MeshPoints& m, long long id) {
  if (m.points) return {m.points[3 * id], m.points[3 * id + 1], m.points[3 * id + 2]};
  const long long nxy = (long long)m.nx * m.ny;
  const int k = (int)(id / nxy);
  const long long r = id - (long long)k * nxy;
  const int j = (int)(r / m.nx);
  const int i = (int)(r - (long long)j * m.nx);
  return {m.ox + double(i) * m.h, m.oy + double(j) * m.h, m.oz + double(k) * m.h};
}

// Eight threads per point (kSlices), each folding every eighth triangle of a
// staged tile; the partials are combined in LDS. sdf (grid mode): the
// level-set value of buildGrid, value / |grad| or +-h/2 where the gradient
// vanishes. value/grad (point mode): the implicit value and gradient of
// evaluateImplicitFunctionValueAndGradient (LevelSetObjectGenerator.h:888-914).
// The closest point is the first triangle (lowest index) at the minimum
// squared distance, as a serial scan finds it. The solid angles are summed per
// slice and the slices then in order, so the sum's rounding differs from a
// serial sum; the inside decision |sum| / 4 pi > 1/2 only sees that for
// points within rounding of the surface, where the value is +-0 either way.
constexpr int kSlices = 8;
constexpr int kPointsPerBlock = kMeshTile / kSlices;

__global__ __launch_bounds__(kMeshTile) void kMeshDistance(const double* __restrict__ tri, int nt,
                                                            MeshPoints m, int positiveInside,
                                                            double* sdf, double* value,
                                                            double* grad) {
  __shared__ double sT[9][kMeshTile];
  __shared__ double sBest[kMeshTile], sSum[kMeshTile], sCp[3][kMeshTile];
  __shared__ int sIdx[kMeshTile];
  const int slice = threadIdx.x % kSlices;
  const int local = threadIdx.x / kSlices;
  const long long id = (long long)blockIdx.x * kPointsPerBlock + local;
  const bool active = id < m.count;
  const V3 p = active ? pointOf(m, id) : V3{0.0, 0.0, 0.0};
  double best = DBL_MAX;
  int bestIdx = INT_MAX;
  V3 bestCp = {0.0, 0.0, 0.0};
  double sum = 0.0;
  for (int t0 = 0; t0 < nt; t0 += kMeshTile) {
    __syncthreads();
    const int t = t0 + (int)threadIdx.x;
    if (t < nt) {
#pragma unroll
      for (int c = 0; c < 9; ++c) sT[c][threadIdx.x] = tri[9 * (size_t)t + c];
    }
    __syncthreads();
    if (!active) continue;
    const int cnt = min(kMeshTile, nt - t0);
    for (int u = slice; u < cnt; u += kSlices) {
      const V3 a = {sT[0][u], sT[1][u], sT[2][u]};
      const V3 b = {sT[3][u], sT[4][u], sT[5][u]};
      const V3 c = {sT[6][u], sT[7][u], sT[8][u]};
      // box lower bound on the distance, with a margin so that rounding
      // never skips the triangle that holds the minimum
      const double dx = fmax(fmax(fmin(fmin(a.x, b.x), c.x) - p.x, 0.0), p.x - fmax(fmax(a.x, b.x), c.x));
      const double dy = fmax(fmax(fmin(fmin(a.y, b.y), c.y) - p.y, 0.0), p.y - fmax(fmax(a.y, b.y), c.y));
      const double dz = fmax(fmax(fmin(fmin(a.z, b.z), c.z) - p.z, 0.0), p.z - fmax(fmax(a.z, b.z), c.z));
      const double box = dx * dx + dy * dy + dz * dz;
      if (box <= best * (1.0 + 1e-9)) {
        const V3 cp = closestOnTriangle(p, a, b, c);
        const double d2 = vlen2(vsub(p, cp));
        if (d2 < best) {  // triangles of a slice come in increasing index
          best = d2;
          bestIdx = t0 + u;
          bestCp = cp;
        }
      }
      sum += solidAngle(p, a, b, c);
    }
  }
  sBest[threadIdx.x] = best;
  sIdx[threadIdx.x] = bestIdx;
  sSum[threadIdx.x] = sum;
  sCp[0][threadIdx.x] = bestCp.x;
  sCp[1][threadIdx.x] = bestCp.y;
  sCp[2][threadIdx.x] = bestCp.z;
  __syncthreads();
  if (!active || slice != 0) return;
  for (int k = 1; k < kSlices; ++k) {
    const int o = threadIdx.x + k;
    const double b2 = sBest[o];
    if (b2 < best || (b2 == best && sIdx[o] < bestIdx)) {
      best = b2;
      bestIdx = sIdx[o];
      bestCp = {sCp[0][o], sCp[1][o], sCp[2][o]};
    }
    sum += sSum[o];
  }
  // evaluateImplicitFunctionValueAndGradient
  const V3 dp = vsub(p, bestCp);
  const double d = vlen(dp);
  const bool inside = fabs(sum / (4.0 * kPi)) > 0.5;
  double val = inside ? -d : d;
  V3 g = {0.0, 0.0, 0.0};
  if (d > 1e-30) {
    const V3 u = vdiv(dp, d);
    g = inside ? vneg(u) : u;
  }
  if (positiveInside) {
    val = -val;
    g = vneg(g);
  }
  if (sdf) {
    // Base::buildGrid: signed-distance-like value from value and gradient
    const double gn = vlen(g);
    sdf[id] = gn > 1e-12 ? val / gn : (val <= 0.0 ? -1.0 : 1.0) * 0.5 * m.h;
  } else {
    value[id] = val;
    grad[3 * id] = g.x;
    grad[3 * id + 1] = g.y;
    grad[3 * id + 2] = g.z;
  }
}

}  // namespace
}  // namespace fdm

using namespace fdm;

extern "C" int fdm_launchMeshSignedDistanceGrid(const double* triangles, int numTriangles,
                                                const double* origin, double spacing,
                                                const int* size, int positiveInside, double* sdf,
                                                hipStream_t stream) {
  if (!triangles || numTriangles <= 0 || !origin || !size || !sdf || !(spacing > 0.0))
    return FDM_EINVAL;
  if (size[0] <= 0 || size[1] <= 0 || size[2] <= 0) return FDM_EINVAL;
  MeshPoints m{};
  m.ox = origin[0];
  m.oy = origin[1];
  m.oz = origin[2];
  m.h = spacing;
  m.nx = size[0];
  m.ny = size[1];
  m.nz = size[2];
  m.points = nullptr;
  m.count = (long long)size[0] * size[1] * size[2];
  hipLaunchKernelGGL(kMeshDistance, dim3(blocksFor(m.count, kPointsPerBlock)), dim3(kMeshTile), 0,
                     stream, triangles, numTriangles, m, positiveInside, sdf, nullptr, nullptr);
  return launchStatus();
}

extern "C" int fdm_launchMeshSignedDistancePoints(const double* triangles, int numTriangles,
                                                  const double* points, int numPoints,
                                                  int positiveInside, double* value, double* grad,
                                                  hipStream_t stream) {
  if (!triangles || numTriangles <= 0 || numPoints < 0 || (numPoints > 0 && (!points || !value || !grad)))
    return FDM_EINVAL;
  if (numPoints == 0) return FDM_OK;
  MeshPoints m{};
  m.points = points;
  m.count = numPoints;
  m.h = 0.0;
  hipLaunchKernelGGL(kMeshDistance, dim3(blocksFor(numPoints, kPointsPerBlock)), dim3(kMeshTile), 0,
                     stream, triangles, numTriangles, m, positiveInside, nullptr, value, grad);
  return launchStatus();
}
