// Periodic boundaries: planar periodicity in x and/or y
// (PeriodicBoundaryXY2D, kernel/PeriodicBoundary.h:178-317) or the
// quarter-plane sector about the z axis through the domain's min corner
// (PeriodicBoundarySector, :319-426), whose three ghost sets are the R90, R180
// and R270 images with turned orientations and velocities. The reference
// keeps a full ghost copy of every particle per set and hashes all of them;
// only ghosts whose ghost hash is valid can ever be found by a node, so the
// device keeps just those, compacted in particle order, with one hash table
// and one interaction set per active ghost set.
#pragma once
#include <algorithm>

#include "../../../include/fundem_amd.h"
#include "DeviceBuffer.h"
#include "LSParticle.h"
#include "SolidInteraction.h"

namespace fundem {

struct GhostSet {
  int capacity = 0;
  DeviceBuffer<int> count, flag, realId, cellBox;
  DeviceBuffer<double> pos, q, vel;
  HashTableBuffers table;
  SolidInteraction interaction;
  bool posed = false;  // holds image orientations and velocities (sector)
  void allocate(size_t n, int cap, long long numGrid, bool withPose) {
    capacity = std::max(cap, 1);
    count.allocate(1);
    flag.allocate(std::max<size_t>(n, 1));
    realId.allocate(capacity);
    pos.allocate(3 * (size_t)capacity);
    if (withPose) {
      q.allocate(4 * (size_t)capacity);
      vel.allocate(3 * (size_t)capacity);
    }
    posed = withPose;
    cellBox.allocate(6);
    table.allocate(capacity, numGrid);
  }
  fdm_ghosts view(int* atBuild) {
    fdm_ghosts g;
    g.atBuild = atBuild;
    g.capacity = capacity;
    g.count = count.get();
    g.flag = flag.get();
    g.realId = realId.get();
    g.pos = pos.get();
    g.q = posed ? q.get() : nullptr;
    g.vel = posed ? vel.get() : nullptr;
    g.table = table.view();
    g.cellBox = cellBox.get();
    return g;
  }
};

struct PeriodicBoundaryXY2D {
 public:
  void turnXOn() { x_ = true; }
  void turnYOn() { y_ = true; }
  void turnXOff() { x_ = false; }
  void turnYOff() { y_ = false; }
  void turnSectorOn() {
    sector_ = true;
    fdm_sectorPeriodic(&sectorFlags_);
  }
  void turnSectorOff() { sector_ = false; }
  bool isActived() const { return x_ || y_ || sector_; }
  bool isSector() const { return sector_; }
  bool isPlanar() const { return x_ || y_; }
  bool active(int set) const { return sector_ ? true : set == 0 ? x_ : set == 1 ? y_ : (x_ && y_); }
  fdm_periodic flags() const {
    fdm_periodic p{};
    if (sector_) return sectorFlags_;
    p.x = x_ ? 1 : 0;
    p.y = y_ ? 1 : 0;
    return p;
  }

  void initialize(LSParticle& P, int ghostCapacity, int pairCapacity, int recordCapacity,
                  int mirrorCapacity) {
    for (int s = 0; s < 3; ++s) {
      if (!active(s)) continue;
      if (sets_[s].capacity != std::max(ghostCapacity, 1) || allocatedFor_[s] != P.num() ||
          sets_[s].posed != sector_)
        sets_[s].allocate(P.num(), ghostCapacity, P.grid().num, sector_);
      allocatedFor_[s] = P.num();
      sets_[s].interaction.initialize(P.num(), pairCapacity, recordCapacity, mirrorCapacity);
    }
    if (isActived()) atBuild_.allocate(std::max<size_t>(P.num(), 1));
  }
  GhostSet& set(int s) { return sets_[s]; }
  // view of set s; every set shares the per-particle visibility bits
  fdm_ghosts view(int s) { return sets_[s].view(atBuild_.get()); }
  double deviceMemoryGB() const {
    double t = 0;
    for (const auto& s : sets_) t += s.pos.memoryGB() + s.interaction.deviceMemoryGB();
    return t;
  }

 private:
  bool x_ = false, y_ = false, sector_ = false;
  fdm_periodic sectorFlags_{};
  GhostSet sets_[3];
  DeviceBuffer<int> atBuild_;
  size_t allocatedFor_[3] = {(size_t)-1, (size_t)-1, (size_t)-1};
};

}  // namespace fundem
