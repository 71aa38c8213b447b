// Level-set templates from closed triangle meshes
// (LevelSetObject::TriangleMeshParticle, LevelSetObjectGenerator.h:382-951,
// with Base::buildGrid / buildGridByResolution / outputGridVTU,
// LevelSetObjectGenerator.h:56-379). Same class, method names, arguments and
// results; the grid is sampled on the GPU by fdm_launchMeshSignedDistanceGrid
// (one thread per grid node over all triangles) instead of a host loop over
// a SAH-BVH. Grid placement (padding, snapping to multiples of the spacing,
// node counts) is computed on the host exactly as the reference does.
#pragma once
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/fundem_amd.h"
#include "DeviceBuffer.h"
#include "VTUWriter.h"
#include "types.h"

namespace LevelSetObject {

// Background grid of a template: value < 0 inside, > 0 outside, nodes at
// gridOrigin + (i, j, k) * gridNodeSpacing, x fastest.
class GridInfo {
 public:
  double3 gridOrigin{0.0, 0.0, 0.0};
  int3 gridNodeSize{0, 0, 0};
  double gridNodeSpacing{0.0};
  std::vector<double> gridNodeSignedDistance;
};

class TriangleMeshParticle {
 public:
  TriangleMeshParticle() = default;
  TriangleMeshParticle(const std::vector<double3>& v, const std::vector<int3>& f) { setMesh(v, f); }
  virtual ~TriangleMeshParticle() = default;

  void setMesh(const std::vector<double3>& v, const std::vector<int3>& f) {
    vertexPosition_ = v;
    triangleVertexIndex_ = f;
    updateBoundingBox();
    clearGrid();
  }
  void setLevelSetPositiveInside(bool v) { levelSetPositiveInside_ = v; }
  void clearMesh() {
    vertexPosition_.clear();
    triangleVertexIndex_.clear();
    updateBoundingBox();
    clearGrid();
  }
  void clearGrid() { gridInfo_ = GridInfo{}; }
  bool isGridBuilt() const { return !gridInfo_.gridNodeSignedDistance.empty(); }

  // Non-empty, every index in range, no repeated vertex in a triangle
  // (LevelSetObjectGenerator.h:872-886).
  virtual bool isValid() const {
    if (vertexPosition_.empty() || triangleVertexIndex_.empty()) return false;
    const int n = (int)vertexPosition_.size();
    for (const int3& t : triangleVertexIndex_) {
      if (t.x < 0 || t.y < 0 || t.z < 0 || t.x >= n || t.y >= n || t.z >= n) return false;
      if (t.x == t.y || t.y == t.z || t.z == t.x) return false;
    }
    return true;
  }

  // Signed distance (negative inside) and its gradient (p - closest)/d at one
  // point; value 1 and zero gradient for an invalid mesh
  // (LevelSetObjectGenerator.h:888-914). Runs on the GPU; synchronises.
  virtual void evaluateImplicitFunctionValueAndGradient(double& value, double3& grad,
                                                        const double3& p) const {
    if (!isValid()) {
      value = 1.0;
      grad = make_double3(0.0, 0.0, 0.0);
      return;
    }
    std::vector<double> v(1), g(3);
    const double pt[3] = {p.x, p.y, p.z};
    evaluatePoints(pt, 1, v.data(), g.data());
    value = v[0];
    grad = make_double3(g[0], g[1], g[2]);
  }
  // The same at n points [3n] (values [n], gradients [3n]).
  void evaluatePoints(const double* points, int n, double* values, double* grads) const {
    if (n <= 0) return;
    if (!isValid()) throw std::invalid_argument("TriangleMeshParticle: invalid mesh");
    hipStream_t s = nullptr;
    fundem::DeviceBuffer<double> tri, pts, val, grad;
    uploadTriangles(tri, s);
    pts.upload(points, 3 * (size_t)n, s);
    val.allocate(n, false);
    grad.allocate(3 * (size_t)n, false);
    fundem::fdmCheck(fdm_launchMeshSignedDistancePoints(tri.get(), (int)triangleVertexIndex_.size(),
                                                        pts.get(), n, levelSetPositiveInside_ ? 1 : 0,
                                                        val.get(), grad.get(), s),
                     "mesh signed distance");
    val.download(values, n, s);
    grad.download(grads, 3 * (size_t)n, s);
    fundem::hipCheck(hipStreamSynchronize(s), "mesh signed distance");
  }

  virtual double3 boundingBoxMin() const { return boundingBoxMin_; }
  virtual double3 boundingBoxMax() const { return boundingBoxMax_; }

  // Base::buildGrid (LevelSetObjectGenerator.h:122-224): the mesh's box
  // padded by max(1, padding) spacings, snapped outward to multiples of the
  // spacing, nodes floor(extent / h + 0.5) + 1 per axis; each node's value
  // is value / |gradient| of the signed distance, +-h/2 where the gradient
  // vanishes. Leaves the grid empty for an invalid mesh or spacing <= 0.
  // Synchronises.
  void buildGrid(const double backgroundGridNodeSpacing, const int backgroundGridPaddingLayers = 2) {
    clearGrid();
    if (!isValid()) return;
    if (backgroundGridNodeSpacing <= 0.0) return;
    const double h = backgroundGridNodeSpacing;
    const double pad = double(std::max(1, backgroundGridPaddingLayers)) * h;
    const double3 bmin = boundingBoxMin(), bmax = boundingBoxMax();
    auto down = [&](double c) { return h * std::floor(c / h); };
    auto up = [&](double c) { return h * std::ceil(c / h); };
    const double3 lo = make_double3(down(bmin.x - pad), down(bmin.y - pad), down(bmin.z - pad));
    const double3 hi = make_double3(up(bmax.x + pad), up(bmax.y + pad), up(bmax.z + pad));
    const int nx = int(std::floor((hi.x - lo.x) / h + 0.5)) + 1;
    const int ny = int(std::floor((hi.y - lo.y) / h + 0.5)) + 1;
    const int nz = int(std::floor((hi.z - lo.z) / h + 0.5)) + 1;
    if (nx <= 0 || ny <= 0 || nz <= 0) return;
    GridInfo g;
    g.gridOrigin = lo;
    g.gridNodeSize = make_int3(nx, ny, nz);
    g.gridNodeSpacing = h;
    const size_t N = size_t(nx) * size_t(ny) * size_t(nz);
    hipStream_t s = nullptr;
    fundem::DeviceBuffer<double> tri, sdf;
    uploadTriangles(tri, s);
    sdf.allocate(N, false);
    const double origin[3] = {lo.x, lo.y, lo.z};
    const int size[3] = {nx, ny, nz};
    fundem::fdmCheck(fdm_launchMeshSignedDistanceGrid(tri.get(), (int)triangleVertexIndex_.size(),
                                                      origin, h, size,
                                                      levelSetPositiveInside_ ? 1 : 0, sdf.get(), s),
                     "mesh signed distance grid");
    g.gridNodeSignedDistance = sdf.toHost(s, N);
    gridInfo_ = std::move(g);
  }
  // Base::buildGridByResolution (LevelSetObjectGenerator.h:226-254): the
  // spacing is the largest box extent over the nodes per diameter.
  void buildGridByResolution(const int backgroundGridNodesPerParticleDiameter = 50,
                             const int backgroundGridPaddingLayers = 2) {
    if (!isValid()) return;
    if (backgroundGridNodesPerParticleDiameter <= 0) return;
    const double3 a = boundingBoxMin(), b = boundingBoxMax();
    const double d = std::max(b.x - a.x, std::max(b.y - a.y, b.z - a.z));
    if (d <= 0.0) return;
    buildGrid(d / double(backgroundGridNodesPerParticleDiameter), backgroundGridPaddingLayers);
  }
  // Base::outputGridVTU (LevelSetObjectGenerator.h:256-362): one vertex per
  // grid node with its value ("signedDistance") and indices ("gridIJK"), in
  // <prefix>LSGrid.vtu (inline binary arrays).
  void outputGridVTU(const std::string& fileNamePrefix) const {
    if (!isGridBuilt()) return;
    const GridInfo& g = gridInfo_;
    const int nx = g.gridNodeSize.x, ny = g.gridNodeSize.y, nz = g.gridNodeSize.z;
    const size_t N = size_t(nx) * size_t(ny) * size_t(nz);
    if (g.gridNodeSignedDistance.size() != N) return;
    std::vector<float> pts(3 * N), sd(N);
    std::vector<int32_t> ijk(3 * N), conn(N), offs(N);
    std::vector<uint8_t> types(N, 1);
    size_t id = 0;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i, ++id) {
          pts[3 * id + 0] = float(g.gridOrigin.x + double(i) * g.gridNodeSpacing);
          pts[3 * id + 1] = float(g.gridOrigin.y + double(j) * g.gridNodeSpacing);
          pts[3 * id + 2] = float(g.gridOrigin.z + double(k) * g.gridNodeSpacing);
          ijk[3 * id + 0] = i;
          ijk[3 * id + 1] = j;
          ijk[3 * id + 2] = k;
          sd[id] = float(g.gridNodeSignedDistance[id]);
          conn[id] = int32_t(id);
          offs[id] = int32_t(id + 1);
        }
    fundem::vtu::Piece f(0.0, 0, N, N);
    f.points(pts);
    f.cells(conn, offs, types);
    f.beginPointDataScalars("signedDistance");
    f.pointArray("signedDistance", sd, 1);
    f.pointArray("gridIJK", ijk, 3);
    f.write(fileNamePrefix + "LSGrid.vtu");
  }

  const GridInfo& gridInfo() const { return gridInfo_; }
  GridInfo& gridInfo() { return gridInfo_; }
  const std::vector<double3>& vertexPosition() const { return vertexPosition_; }
  std::vector<double3>& vertexPosition() { return vertexPosition_; }
  const std::vector<int3>& triangleVertexIndex() const { return triangleVertexIndex_; }
  std::vector<int3>& triangleVertexIndex() { return triangleVertexIndex_; }
  const double3& meshBoundingBoxMin() const { return boundingBoxMin_; }
  const double3& meshBoundingBoxMax() const { return boundingBoxMax_; }

 protected:
  void updateBoundingBox() {
    if (vertexPosition_.empty()) {
      boundingBoxMin_ = boundingBoxMax_ = make_double3(0.0, 0.0, 0.0);
      return;
    }
    boundingBoxMin_ = boundingBoxMax_ = vertexPosition_[0];
    for (const double3& v : vertexPosition_) {
      boundingBoxMin_ = make_double3(std::min(boundingBoxMin_.x, v.x), std::min(boundingBoxMin_.y, v.y),
                                     std::min(boundingBoxMin_.z, v.z));
      boundingBoxMax_ = make_double3(std::max(boundingBoxMax_.x, v.x), std::max(boundingBoxMax_.y, v.y),
                                     std::max(boundingBoxMax_.z, v.z));
    }
  }
  // triangles as [a b c] FP64 coordinate triples on the device
  void uploadTriangles(fundem::DeviceBuffer<double>& tri, hipStream_t s) const {
    std::vector<double> t(9 * triangleVertexIndex_.size());
    for (size_t k = 0; k < triangleVertexIndex_.size(); ++k) {
      const int3 f = triangleVertexIndex_[k];
      const double3 v[3] = {vertexPosition_[f.x], vertexPosition_[f.y], vertexPosition_[f.z]};
      for (int c = 0; c < 3; ++c) {
        t[9 * k + 3 * c + 0] = v[c].x;
        t[9 * k + 3 * c + 1] = v[c].y;
        t[9 * k + 3 * c + 2] = v[c].z;
      }
    }
    tri.upload(t, s);
  }

  std::vector<double3> vertexPosition_;
  std::vector<int3> triangleVertexIndex_;
  bool levelSetPositiveInside_{false};
  double3 boundingBoxMin_{0.0, 0.0, 0.0};
  double3 boundingBoxMax_{0.0, 0.0, 0.0};
  GridInfo gridInfo_;
};

}  // namespace LevelSetObject
