// LSParticle: host container of a set of level-set bodies (particles or
// walls): materials, templates, per-body state and the body spatial grid.
// Public methods follow kernel/LSParticle.h:219-834. Device state is planar
// SoA, positions FP64 and the rest FP32 (see include/fundem_amd.h); host
// mirrors stay double3 and are refreshed by finalize().
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <limits>
#include <vector>

#include "../../../include/fundem_amd.h"
#include "DeviceBuffer.h"
#include "types.h"

namespace fundem {

// Grid parameters of spatialGrid::set (mySpatialGrid.h:212-240).
inline fdm_grid makeGrid(double3 mn, double3 mx, double cs) {
  if (mx.x <= mn.x) mx.x = mn.x + cs;
  if (mx.y <= mn.y) mx.y = mn.y + cs;
  if (mx.z <= mn.z) mx.z = mn.z + cs;
  fdm_grid g;
  g.min = {mn.x, mn.y, mn.z};
  g.max = {mx.x, mx.y, mx.z};
  g.invCell = {0.0, 0.0, 0.0};
  g.size = {2, 2, 2};
  g.num = 8;
  if (isZero(cs)) return g;
  const double Lx = mx.x - mn.x, Ly = mx.y - mn.y, Lz = mx.z - mn.z;
  g.size.x = Lx > 2. * cs ? int(Lx / cs) : 2;
  g.size.y = Ly > 2. * cs ? int(Ly / cs) : 2;
  g.size.z = Lz > 2. * cs ? int(Lz / cs) : 2;
  g.invCell.x = Lx > 0.0 ? double(g.size.x) / Lx : 0.0;
  g.invCell.y = Ly > 0.0 ? double(g.size.y) / Ly : 0.0;
  g.invCell.z = Lz > 0.0 ? double(g.size.z) / Lz : 0.0;
  g.num = (long long)g.size.x * g.size.y * g.size.z;
  return g;
}

// Device hash table of one grid.
struct HashTableBuffers {
  DeviceBuffer<int> hashValue, sortIndex, cellStart, cellEnd, prevKey, prevCount, validAtBuild;
  DeviceBuffer<unsigned> sortKey;
  void allocate(size_t n, long long numGrid) {
    const size_t m = std::max<size_t>(n, 1);
    hashValue.allocate(m);
    validAtBuild.allocate(m);
    sortIndex.allocate(m);
    sortKey.allocate(m);
    prevKey.allocate(m);
    cellStart.allocate((size_t)numGrid, false);
    cellEnd.allocate((size_t)numGrid, false);
    hipCheck(hipMemset(cellStart.get(), 0xFF, sizeof(int) * (size_t)numGrid), "hipMemset");
    hipCheck(hipMemset(cellEnd.get(), 0xFF, sizeof(int) * (size_t)numGrid), "hipMemset");
    prevCount.allocate(1);  // 0 entries marked: all cells already -1
  }
  fdm_hash_table view() {
    fdm_hash_table t;
    t.hashValue = hashValue.get();
    t.sortKey = sortKey.get();
    t.sortIndex = sortIndex.get();
    t.cellStart = cellStart.get();
    t.cellEnd = cellEnd.get();
    t.prevKey = prevKey.get();
    t.prevCount = prevCount.get();
    t.validAtBuild = validAtBuild.get();
    return t;
  }
};

struct LSParticle {
 public:
  LSParticle() = default;
  LSParticle(const LSParticle&) = delete;
  LSParticle& operator=(const LSParticle&) = delete;
  LSParticle(LSParticle&&) noexcept = default;
  LSParticle& operator=(LSParticle&&) noexcept = default;

  // LSParticle::addMaterial (LSParticle.h:229-240), with its clamps.
  void addMaterial(double kn, double ks, double mu, double e) {
    fdm_material m{};
    m.kn = kn > 0. ? kn : 0.0;
    m.ks = ks > 0. ? ks : 0.0;
    m.mu = mu > 0. ? mu : 0.0;
    m.e = (e > 0. && e < 1.0) ? e : 1.0;
    // contactKernel.cuh:50-51, evaluated once per material
    const double logRes = std::log(m.e);
    m.dissipation = -logRes / (double)std::sqrt((float)(logRes * logRes + pi() * pi()));
    materials_.push_back(m);
  }
  // A material given by its damping ratio zeta; e is the restitution of a
  // linear dashpot with that ratio, exp(-pi zeta / sqrt(1 - zeta^2)), which
  // only orders materials when a pair picks whose damping applies.
  void addMaterialDampingRatio(double kn, double ks, double mu, double zeta) {
    fdm_material m{};
    m.kn = kn > 0. ? kn : 0.0;
    m.ks = ks > 0. ? ks : 0.0;
    m.mu = mu > 0. ? mu : 0.0;
    m.dissipation = zeta > 0. ? zeta : 0.0;
    m.e = m.dissipation <= 0. ? 1.0
          : m.dissipation >= 1. ? 0.0
                                : std::exp(-pi() * m.dissipation / std::sqrt(1.0 - m.dissipation * m.dissipation));
    materials_.push_back(m);
  }

  // LSParticle::addGeometry (LSParticle.h:242-290). SDF values are stored
  // FP32 on the device, node coordinates at state precision (fdm_real); the
  // template constants stay FP64.
  void addGeometry(double mass, symMatrix inertia, double3 origin, double spacing, int3 size,
                   const std::vector<double>& sdf, const std::vector<double3>& nodes,
                   const std::vector<int3>& connectivity) {
    if (size.x < 2 || size.y < 2 || size.z < 2) {
      std::cerr << "[LSParticle] Invalid level-set host grid size: (" << size.x << ", " << size.y
                << ", " << size.z << ")." << std::endl;
      return;
    }
    if (spacing <= 0.0) {
      std::cerr << "[LSParticle] Invalid level-set host grid spacing: " << spacing << "."
                << std::endl;
      return;
    }
    const size_t expected = size_t(size.x) * size_t(size.y) * size_t(size.z);
    if (expected != sdf.size()) {
      std::cerr << "[LSParticle] Inconsistent level-set host data size. Expected " << expected
                << ", got " << sdf.size() << "." << std::endl;
      return;
    }
    if (nodes.size() > size_t(FDM_CHUNK) * FDM_MAX_CHUNKS) {
      std::cerr << "[LSParticle] Too many boundary nodes in one template: " << nodes.size()
                << " (at most " << FDM_CHUNK * FDM_MAX_CHUNKS << ")." << std::endl;
      return;
    }
    fdm_template t{};
    const symMatrix I = mass > 0.0 ? inertia : symMatrix();
    t.invMass = mass > 0.0 ? 1.0 / mass : 0.0;
    const symMatrix Ii = ::inverse(I);
    const double Iv[6] = {I.xx, I.yy, I.zz, I.xy, I.xz, I.yz};
    const double Iiv[6] = {Ii.xx, Ii.yy, Ii.zz, Ii.xy, Ii.xz, Ii.yz};
    for (int k = 0; k < 6; ++k) {
      t.inertia[k] = Iv[k];
      t.invInertia[k] = Iiv[k];
    }
    t.origin[0] = origin.x;
    t.origin[1] = origin.y;
    t.origin[2] = origin.z;
    t.invSpacing = 1.0 / spacing;
    t.size[0] = size.x;
    t.size[1] = size.y;
    t.size[2] = size.z;
    t.sdfStart = (long long)sdf_.size();
    for (double v : sdf) sdf_.push_back((float)v);
    t.nodeStart = (int)(nodes_.size() / 3);
    t.nodeCount = (int)nodes.size();
    double radius = 0.;
    std::vector<double3> fn(nodes.size());
    for (size_t k = 0; k < nodes.size(); ++k) {
      const fdm_real x = (fdm_real)nodes[k].x, y = (fdm_real)nodes[k].y, z = (fdm_real)nodes[k].z;
      nodes_.push_back(x);
      nodes_.push_back(y);
      nodes_.push_back(z);
      fn[k] = make_double3(x, y, z);
      radius = std::max(radius, length(fn[k]));
    }
    t.radius = radius;
    // culling chunks of 64 consecutive nodes: centre and enclosing radius
    t.chunkStart = (int)(chunks_.size() / 4);
    t.chunkCount = (int)((nodes.size() + FDM_CHUNK - 1) / FDM_CHUNK);
    for (int c = 0; c < t.chunkCount; ++c) {
      const size_t b = size_t(c) * FDM_CHUNK, e = std::min(nodes.size(), b + FDM_CHUNK);
      double cx = 0, cy = 0, cz = 0;
      for (size_t k = b; k < e; ++k) {
        cx += fn[k].x;
        cy += fn[k].y;
        cz += fn[k].z;
      }
      const double inv = 1.0 / double(e - b);
      const float fx = (float)(cx * inv), fy = (float)(cy * inv), fz = (float)(cz * inv);
      double r = 0.0;
      for (size_t k = b; k < e; ++k)
        r = std::max(r, length(make_double3(fn[k].x - fx, fn[k].y - fy, fn[k].z - fz)));
      float fr = (float)r;
      if ((double)fr < r) fr = std::nextafter(fr, std::numeric_limits<float>::infinity());
      chunks_.push_back(fx);
      chunks_.push_back(fy);
      chunks_.push_back(fz);
      chunks_.push_back(fr);
    }
    // Gradient bound of the trilinear interpolant of the stored (FP32) SDF.
    // Inside a cell, d(phi)/dx is a convex combination of the cell's four
    // x-edge differences (likewise y, z), so |grad phi| <= sqrt(ex^2 + ey^2 +
    // ez^2) / h with e* the cell's largest edge difference per axis; the
    // interpolant is continuous, so the largest cell bound is a Lipschitz
    // constant everywhere in the box.
    double Lc = 0.0;
    const float* s = sdf_.data() + t.sdfStart;
    const size_t sx = 1, sy = size_t(size.x), sz = size_t(size.x) * size.y;
    for (int z = 0; z + 1 < size.z; ++z)
      for (int y = 0; y + 1 < size.y; ++y)
        for (int x = 0; x + 1 < size.x; ++x) {
          const size_t i = size_t(z) * sz + size_t(y) * sy + x;
          double e[3] = {0, 0, 0};
          for (int c = 0; c < 4; ++c) {
            const size_t ax = i + ((c & 1) ? sy : 0) + ((c & 2) ? sz : 0);  // x-edges
            const size_t ay = i + ((c & 1) ? sx : 0) + ((c & 2) ? sz : 0);  // y-edges
            const size_t az = i + ((c & 1) ? sx : 0) + ((c & 2) ? sy : 0);  // z-edges
            e[0] = std::max(e[0], std::fabs((double)s[ax + sx] - s[ax]));
            e[1] = std::max(e[1], std::fabs((double)s[ay + sy] - s[ay]));
            e[2] = std::max(e[2], std::fabs((double)s[az + sz] - s[az]));
          }
          Lc = std::max(Lc, std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]));
        }
    t.lipschitz = Lc * t.invSpacing * (1.0 + 1e-9);
    double reach = 0.0;
    for (int c = 0; c < 8; ++c) {
      const double px = origin.x + ((c & 1) ? spacing * (size.x - 1) : 0.0);
      const double py = origin.y + ((c & 2) ? spacing * (size.y - 1) : 0.0);
      const double pz = origin.z + ((c & 4) ? spacing * (size.z - 1) : 0.0);
      reach = std::max(reach, std::sqrt(px * px + py * py + pz * pz));
    }
    t.boxReach = reach * (1.0 + 1e-9);
    // Affine fit of the stored SDF over the grid nodes (least squares; the
    // centred node coordinates are orthogonal on a regular grid). The
    // trilinear interpolant of the residuals is a convex combination of
    // them, so phi >= fit - max |residual| everywhere in the box; a small
    // relative slack covers the rounding of both evaluations. Only a fit
    // within a thousandth of the spacing is kept (plane walls).
    {
      const size_t nn = size_t(size.x) * size.y * size.z;
      const double ci = 0.5 * (size.x - 1), cj = 0.5 * (size.y - 1), ck = 0.5 * (size.z - 1);
      double mean = 0.0, sI = 0.0, sJ = 0.0, sK = 0.0, dI = 0.0, dJ = 0.0, dK = 0.0, vmax = 0.0;
      for (int z = 0; z < size.z; ++z)
        for (int y = 0; y < size.y; ++y)
          for (int x = 0; x < size.x; ++x) {
            const double v = s[size_t(z) * sz + size_t(y) * sy + x];
            mean += v;
            sI += (x - ci) * v;
            sJ += (y - cj) * v;
            sK += (z - ck) * v;
            dI += (x - ci) * (x - ci);
            dJ += (y - cj) * (y - cj);
            dK += (z - ck) * (z - ck);
            vmax = std::max(vmax, std::fabs(v));
          }
      mean /= double(nn);
      const double b = dI > 0 ? sI / dI : 0.0, c = dJ > 0 ? sJ / dJ : 0.0, d = dK > 0 ? sK / dK : 0.0;
      double res = 0.0;
      for (int z = 0; z < size.z; ++z)
        for (int y = 0; y < size.y; ++y)
          for (int x = 0; x < size.x; ++x) {
            const double fit = mean + b * (x - ci) + c * (y - cj) + d * (z - ck);
            res = std::max(res, std::fabs((double)s[size_t(z) * sz + size_t(y) * sy + x] - fit));
          }
      const double h = spacing;
      // phi(l) = mean + b (i - ci) + ... with i = (l.x - origin.x) / h
      t.plane[0] = b / h;
      t.plane[1] = c / h;
      t.plane[2] = d / h;
      t.plane[3] = mean - b * (origin.x / h + ci) - c * (origin.y / h + cj) - d * (origin.z / h + ck);
      const double scale = vmax + (std::fabs(t.plane[3]) + (std::fabs(b) + std::fabs(c) + std::fabs(d)) *
                                                              (reach / h + ci + cj + ck));
      t.planeSlack = res + 1e-9 * scale;
      t.affine = res <= 1e-3 * h ? 1 : 0;
    }
    templates_.push_back(t);
    nodeHost_.push_back(fn);
    connectivity_.push_back(connectivity);
  }

  // LSParticle::addParticle (LSParticle.h:292-342).
  void addParticle(double3 position, double3 velocity, double3 angularVelocity,
                   quaternion orientation, int materialID, int geometryID) {
    if (materialID < 0 || size_t(materialID) >= materials_.size()) {
      std::cerr << "[LSParticle] Invalid material ID: " << materialID << "." << std::endl;
      return;
    }
    if (geometryID < 0 || size_t(geometryID) >= templates_.size()) {
      std::cerr << "[LSParticle] Invalid geometry ID: " << geometryID << "." << std::endl;
      return;
    }
    position_.push_back(position);
    velocity_.push_back(velocity);
    angularVelocity_.push_back(angularVelocity);
    force_.push_back(make_double3(0., 0., 0.));
    torque_.push_back(make_double3(0., 0., 0.));
    orientation_.push_back(normalize(orientation));
    materialID_.push_back(materialID);
    geometryID_.push_back(geometryID);
    system_.push_back(currentSystem_);
  }
  // Bodies added from now on belong to system k (several independent systems
  // in one solver).
  void setCurrentSystem(int k) { currentSystem_ = k; }
  const std::vector<int>& systemOfBody() const { return system_; }

  // Cell size of the grid rule for the bodies of system k (k < 0: all
  // bodies): 2 * the largest radius over every template, as
  // LSParticle::initialize computes it (LSParticle.h:383), or `cellSize` when
  // positive.
  double cellSizeRule(double cellSize) const {
    double cs = 0.0;
    if (num() > 0 && !templates_.empty()) {
      for (const auto& t : templates_) cs = std::max(cs, t.radius);
      cs *= 2.0;
    }
    if (cellSize > 0.0) cs = cellSize;
    return cs;
  }
  // 2 * the largest radius over the templates the bodies of system k use.
  double cellSizeRuleOfSystem(int k, double cellSize) const {
    if (cellSize > 0.0) return cellSize;
    double cs = 0.0;
    for (size_t i = 0; i < num(); ++i)
      if (system_[i] == k) cs = std::max(cs, templates_[geometryID_[i]].radius);
    return 2.0 * cs;
  }

  // Several systems: the table spans every system's cells end to end
  // (globalCells), the grid returned by grid() only carries that count.
  void initializeSystems(long long globalCells, hipStream_t s) {
    grid_ = fdm_grid{};
    grid_.num = std::max<long long>(globalCells, 1);
    grid_.size = {1, 1, 1};
    upload(s);
    table_.allocate(num(), grid_.num);
  }

  void move(size_t index, double3 offset) {
    if (index >= num()) return;
    position_[index] = make_double3(position_[index].x + offset.x, position_[index].y + offset.y,
                                    position_[index].z + offset.z);
  }
  void setVelocity(size_t index, double3 v) {
    if (index < num()) velocity_[index] = v;
  }
  void setAngularVelocity(size_t index, double3 w) {
    if (index < num()) angularVelocity_[index] = w;
  }

  // LSParticle::initialize (LSParticle.h:371-386): upload and grid set-up.
  // cellSize > 0 overrides the 2 * max radius rule.
  void initialize(double3 minDomain, double3 maxDomain, double cellSize, hipStream_t s) {
    double cs = 0.0;
    const size_t n = num();
    if (n > 0 && !templates_.empty()) {
      for (const auto& t : templates_) cs = std::max(cs, t.radius);
      cs *= 2.0;
    }
    if (cellSize > 0.0) cs = cellSize;
    grid_ = makeGrid(minDomain, maxDomain, cs);
    upload(s);
    table_.allocate(n, grid_.num);
  }

  void upload(hipStream_t s) {
    const size_t n = num();
    const size_t m = std::max<size_t>(n, 1);
    std::vector<fdm_real> buf;
    auto put3 = [&](DeviceBuffer<fdm_real>& d, const std::vector<double3>& h) {
      buf.assign(3 * m, fdm_real(0));
      for (size_t i = 0; i < n; ++i) {
        buf[i] = (fdm_real)h[i].x;
        buf[n + i] = (fdm_real)h[i].y;
        buf[2 * n + i] = (fdm_real)h[i].z;
      }
      d.allocate(3 * m, false);
      d.upload(buf, s);
      hipCheck(hipStreamSynchronize(s), "upload");
    };
    {
      std::vector<double> xb(3 * m, 0.0);
      for (size_t i = 0; i < n; ++i) {
        xb[i] = position_[i].x;
        xb[n + i] = position_[i].y;
        xb[2 * n + i] = position_[i].z;
      }
      x_.allocate(3 * m, false);
      x_.upload(xb, s);
      hipCheck(hipStreamSynchronize(s), "upload");
    }
    put3(v_, velocity_);
    put3(w_, angularVelocity_);
    put3(F_, force_);
    put3(T_, torque_);
    buf.assign(4 * m, fdm_real(0));
    for (size_t i = 0; i < n; ++i) {
      buf[i] = (fdm_real)orientation_[i].q0;
      buf[n + i] = (fdm_real)orientation_[i].q1;
      buf[2 * n + i] = (fdm_real)orientation_[i].q2;
      buf[3 * n + i] = (fdm_real)orientation_[i].q3;
    }
    q_.allocate(4 * m, false);
    q_.upload(buf, s);
    mat_.upload(materialID_, s);
    geo_.upload(geometryID_, s);
    dmat_.upload(materials_, s);
    dtmpl_.upload(templates_, s);
    dsdf_.upload(sdf_, s);
    dnodes_.upload(nodes_, s);
    dchunks_.upload(chunks_, s);
    hipCheck(hipStreamSynchronize(s), "upload");
  }

  // LSParticle::finalize: device state back to the host mirrors.
  void finalize(hipStream_t s) {
    const size_t n = num();
    if (n == 0 || !x_.get()) return;
    std::vector<fdm_real> buf(4 * n);
    auto get3 = [&](const DeviceBuffer<fdm_real>& d, std::vector<double3>& h) {
      d.download(buf.data(), 3 * n, s);
      hipCheck(hipStreamSynchronize(s), "download");
      for (size_t i = 0; i < n; ++i) h[i] = make_double3(buf[i], buf[n + i], buf[2 * n + i]);
    };
    {
      std::vector<double> xb(3 * n);
      x_.download(xb.data(), 3 * n, s);
      hipCheck(hipStreamSynchronize(s), "download");
      for (size_t i = 0; i < n; ++i) position_[i] = make_double3(xb[i], xb[n + i], xb[2 * n + i]);
    }
    get3(v_, velocity_);
    get3(w_, angularVelocity_);
    get3(F_, force_);
    get3(T_, torque_);
    q_.download(buf.data(), 4 * n, s);
    hipCheck(hipStreamSynchronize(s), "download");
    for (size_t i = 0; i < n; ++i)
      orientation_[i] = make_quaternion(buf[i], buf[n + i], buf[2 * n + i], buf[3 * n + i]);
  }

  fdm_bodies bodies() {
    fdm_bodies b;
    b.n = (int)num();
    b.x = x_.get();
    b.v = v_.get();
    b.w = w_.get();
    b.q = q_.get();
    b.F = F_.get();
    b.T = T_.get();
    b.mat = mat_.get();
    b.geo = geo_.get();
    return b;
  }
  fdm_geometry geometry() {
    fdm_geometry g;
    g.num = (int)templates_.size();
    g.info = dtmpl_.get();
    g.sdf = dsdf_.get();
    g.nodes = dnodes_.get();
    g.chunks = dchunks_.get();
    return g;
  }
  const fdm_grid& grid() const { return grid_; }
  HashTableBuffers& table() { return table_; }
  const fdm_material* deviceMaterials() const { return dmat_.get(); }

  size_t num() const { return position_.size(); }
  size_t num_device() const { return x_.size() ? num() : 0; }
  // Launch shape of a kernel with one thread per body, for user hooks
  // (LSParticle::gridDim/blockDim, LSParticle.h:765-766), set at
  // initialisation by the rule of LSParticle.h:379-381: blockDim is
  // maxGPUThread, or the body count when that is smaller; gridDim covers
  // every body. Both stay 1 without bodies.
  const size_t& gridDim() const { return gridDim_; }
  const size_t& blockDim() const { return blockDim_; }
  void setLaunchShape(size_t maxGPUThread) {
    const size_t n = num();
    if (n == 0) return;
    if (maxGPUThread > 0) blockDim_ = maxGPUThread;
    if (n < maxGPUThread) blockDim_ = n;
    gridDim_ = (n + blockDim_ - 1) / blockDim_;
  }
  size_t numGeometry() const { return templates_.size(); }
  size_t numMaterial() const { return materials_.size(); }
  const fdm_template& templateInfo(size_t g) const { return templates_[g]; }
  const std::vector<double3>& templateNodes(size_t g) const { return nodeHost_[g]; }
  const std::vector<int3>& templateConnectivity(size_t g) const { return connectivity_[g]; }
  const fdm_material& materialInfo(size_t m) const { return materials_[m]; }
  // concatenated FP32 SDF values and node triples, as uploaded
  const std::vector<float>& sdfHost() const { return sdf_; }
  const std::vector<fdm_real>& nodesHost() const { return nodes_; }
  long long totalNodes() const {
    long long s = 0;
    for (int g : geometryID_) s += templates_[g].nodeCount;
    return s;
  }

  // Device state, planar SoA (n values per component): positions FP64, the
  // rest FP32.
  double* position() { return x_.get(); }
  fdm_real* velocity() { return v_.get(); }
  fdm_real* angularVelocity() { return w_.get(); }
  fdm_real* force() { return F_.get(); }
  fdm_real* torque() { return T_.get(); }
  fdm_real* orientation() { return q_.get(); }
  int* materialID() { return mat_.get(); }
  int* geometryID() { return geo_.get(); }
  int* hashValue() { return table_.hashValue.get(); }
  int* hashIndex() { return table_.sortIndex.get(); }

  const std::vector<double3>& positionHostRef() const { return position_; }
  const std::vector<double3>& velocityHostRef() const { return velocity_; }
  const std::vector<double3>& angularVelocityHostRef() const { return angularVelocity_; }
  const std::vector<double3>& forceHostRef() const { return force_; }
  const std::vector<double3>& torqueHostRef() const { return torque_; }
  const std::vector<quaternion>& orientationHostRef() const { return orientation_; }
  const std::vector<int>& geometryIDHostRef() const { return geometryID_; }
  const std::vector<int>& materialIDHostRef() const { return materialID_; }

  double deviceMemoryGB() const {
    return x_.memoryGB() + v_.memoryGB() + w_.memoryGB() + q_.memoryGB() + F_.memoryGB() +
           T_.memoryGB() + dsdf_.memoryGB() + dnodes_.memoryGB();
  }

 private:
  std::vector<fdm_material> materials_;
  std::vector<fdm_template> templates_;
  std::vector<float> sdf_, chunks_;
  std::vector<fdm_real> nodes_;
  std::vector<std::vector<double3>> nodeHost_;
  std::vector<std::vector<int3>> connectivity_;
  std::vector<double3> position_, velocity_, angularVelocity_, force_, torque_;
  std::vector<quaternion> orientation_;
  std::vector<int> materialID_, geometryID_, system_;
  int currentSystem_ = 0;

  DeviceBuffer<double> x_;
  DeviceBuffer<fdm_real> v_, w_, q_, F_, T_;
  DeviceBuffer<int> mat_, geo_;
  DeviceBuffer<fdm_material> dmat_;
  DeviceBuffer<fdm_template> dtmpl_;
  DeviceBuffer<float> dsdf_, dchunks_;
  DeviceBuffer<fdm_real> dnodes_;
  fdm_grid grid_{};
  HashTableBuffers table_;
  size_t gridDim_ = 1, blockDim_ = 1;
};

}  // namespace fundem
