// LSSolver: the FunDEM-compatible C++ host API (kernel/LSSolver.h). Public
// and protected methods keep the reference's names, arguments and error
// behaviour (messages on std::cerr and a silent return on invalid input);
// the step calls the gfx950 launchers of include/fundem_amd.h.
//
// Additive API (not in the reference): setCellSize, setContactCapacity,
// initialize/advance (step without output), useGraph, contacts, status,
// stage timing.
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <filesystem>
#include <cstdio>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/fundem_amd.h"
#include "DeviceBuffer.h"
#include "LSParticle.h"
#include "PeriodicBoundary.h"
#include "SolidInteraction.h"
#include "VBondedInteraction.h"
#include "VTUWriter.h"
#include "types.h"

namespace fundem {


class LSAdjoint;
class LSAutotuner;

class LSSolver {
  friend class LSAdjoint;
  friend class LSAutotuner;

 public:
  explicit LSSolver(const std::string dir = "Problem", hipStream_t stream = 0,
                    const size_t maxGPUThread = 256, const int device = 0)
      : dir_(dir), stream_(stream), maxGPUThread_(maxGPUThread), device_(device) {
    fdm_getTuning(&tuning_);
  }
  virtual ~LSSolver() {
    destroyGraphs();
    if (hostDecision_) (void)hipHostFree(hostDecision_);
    if (ownStream_) (void)hipStreamDestroy(ownStream_);
  }
  LSSolver(const LSSolver&) = delete;
  LSSolver& operator=(const LSSolver&) = delete;

  // ---------------------------------------------------------------- inputs
  void addParticleMaterial(double normalStiffness, double shearStiffness,
                           double frictionCoefficient, double restitutionCoefficient) {
    particles_.addMaterial(normalStiffness, shearStiffness, frictionCoefficient,
                           restitutionCoefficient);
  }
  void addParticleGeometry(double mass, symMatrix inertiaTensor, double3 gridNodeOrigin,
                           double gridNodeSpacing, int3 gridNodeSize,
                           const std::vector<double>& gridNodeSignedDistance,
                           const std::vector<double3>& boundaryNodePosition,
                           const std::vector<int3>& boundaryNodeConnectivity) {
    particles_.addGeometry(mass, inertiaTensor, gridNodeOrigin, gridNodeSpacing, gridNodeSize,
                           gridNodeSignedDistance, boundaryNodePosition, boundaryNodeConnectivity);
  }
  void addParticle(double3 position, double3 velocity, double3 angularVelocity,
                   quaternion orientation, int materialID, int geometryID) {
    particles_.addParticle(position, velocity, angularVelocity, orientation, materialID, geometryID);
  }
  // LSSolver::addParticleCluster (LSSolver.h:125-165)
  void addParticleCluster(const std::vector<double3>& position, const std::vector<double3>& velocity,
                          const std::vector<double3>& angularVelocity,
                          const std::vector<quaternion>& orientation, int materialID,
                          double density, double3 gridNodeOrigin, double gridNodeSpacing,
                          int3 gridNodeSize, const std::vector<double>& gridNodeSignedDistance,
                          const std::vector<double3>& boundaryNodePosition,
                          const std::vector<int3>& boundaryNodeConnectivity) {
    if (position.size() != velocity.size() || position.size() != angularVelocity.size() ||
        position.size() != orientation.size()) {
      std::cerr << "[LSSolver] Inconsistent particle cluster data size. Position, velocity, "
                   "angular velocity, and orientation vectors must have the same size."
                << std::endl;
      return;
    }
    double3 localOffset;
    addParticleGeometry(localOffset, density, gridNodeOrigin, gridNodeSpacing, gridNodeSize,
                        gridNodeSignedDistance, boundaryNodePosition, boundaryNodeConnectivity);
    for (size_t i = 0; i < position.size(); ++i)
      addParticle(vec::add(position[i], rotateVectorByQuaternion(orientation[i], localOffset)),
                  velocity[i], angularVelocity[i], orientation[i], materialID,
                  int(particles_.numGeometry()) - 1);
  }
  void addWallMaterial(double frictionCoefficient, double restitutionCoefficient) {
    walls_.addMaterial(0.0, 0.0, frictionCoefficient, restitutionCoefficient);
  }
  // Additive: materials given by their damping ratio zeta (the contact
  // dashpots are 2 zeta sqrt(m k)) instead of a restitution coefficient.
  void addParticleMaterialDampingRatio(double normalStiffness, double shearStiffness,
                                       double frictionCoefficient, double dampingRatio) {
    particles_.addMaterialDampingRatio(normalStiffness, shearStiffness, frictionCoefficient,
                                       dampingRatio);
  }
  void addWallMaterialDampingRatio(double frictionCoefficient, double dampingRatio) {
    walls_.addMaterialDampingRatio(0.0, 0.0, frictionCoefficient, dampingRatio);
  }
  void addWallGeometry(double3 gridNodeOrigin, double gridNodeSpacing, int3 gridNodeSize,
                       const std::vector<double>& gridNodeSignedDistance,
                       const std::vector<double3>& boundaryNodePosition,
                       const std::vector<int3>& boundaryNodeConnectivity) {
    walls_.addGeometry(0.0, symMatrix(), gridNodeOrigin, gridNodeSpacing, gridNodeSize,
                       gridNodeSignedDistance, boundaryNodePosition, boundaryNodeConnectivity);
  }
  void addWall(double3 position, double3 velocity, double3 angularVelocity, quaternion orientation,
               int materialID, int geometryID) {
    walls_.addParticle(position, velocity, angularVelocity, orientation, materialID, geometryID);
  }
  void moveLSParticle(size_t index, double3 offSet) { particles_.move(index, offSet); }
  void moveWall(size_t index, double3 offset) { walls_.move(index, offset); }
  void setFixedVelocityToWall(size_t index, double3 velocity) { walls_.setVelocity(index, velocity); }
  void setFixedAngularVelocityToWall(size_t index, double3 angularVelocity) {
    walls_.setAngularVelocity(index, angularVelocity);
  }
  // LSSolver::addBondedInteraction (LSSolver.h:306-337). The bond frames are
  // built from the state the device holds (FP64 positions, FP32 orientations).
  void addBondedInteraction(int masterObjectID, int slaveObjectID, double3 bondPoint,
                            double3 bondNormal, double radius, double initialLength,
                            double YoungsModulus, double poissonRatio, double tensileStrength = 0.,
                            double cohesion = 0., double frictionCoefficient = 0.) {
    if (masterObjectID >= (int)particles_.num() || slaveObjectID >= (int)particles_.num()) return;
    auto r3 = [](double3 v) { return v; };
    auto r4 = [](quaternion q) {
      return make_quaternion((fdm_real)q.q0, (fdm_real)q.q1, (fdm_real)q.q2, (fdm_real)q.q3);
    };
    bonds_.add(masterObjectID, slaveObjectID, r3(particles_.positionHostRef()[masterObjectID]),
               r3(particles_.positionHostRef()[slaveObjectID]),
               r4(particles_.orientationHostRef()[masterObjectID]),
               r4(particles_.orientationHostRef()[slaveObjectID]), bondPoint, bondNormal, radius,
               initialLength, YoungsModulus, poissonRatio, tensileStrength, cohesion,
               frictionCoefficient);
  }
  void addPeriodicBoundaryXD() { periodic_.turnXOn(); }
  void addPeriodicBoundaryYD() { periodic_.turnYOn(); }
  // LSSolver::addPeriodicBoundarySector (LSSolver.h:410-413): the quarter
  // plane x >= min.x, y >= min.y of the domain is one sector of a body of
  // revolution about the z axis through min; a particle leaving through one
  // face re-enters through the other turned by -+pi/2, and R90, R180, R270
  // images stand in for the neighbouring sectors. One system, not combined
  // with planar periodicity; candidate lists are rebuilt every step.
  void addPeriodicBoundarySector() { periodic_.turnSectorOn(); }
  void removePeriodicBoundary() {
    periodic_.turnSectorOff();
    periodic_.turnXOff();
    periodic_.turnYOff();
  }

  // ------------------------------------------------------------- additive
  // Several independent systems in one solver (no counterpart in the
  // reference): bodies, walls and materials added after addSystem(k) belong
  // to system k, which has its own domain, grids and periodic axes; walls
  // must be added system by system. initialize() then ignores its domain
  // arguments. Each system evolves exactly as it would in a solver of its
  // own with the same templates.
  int addSystem(double3 minDomain, double3 maxDomain, bool periodicX, bool periodicY) {
    if (!(maxDomain.x > minDomain.x && maxDomain.y > minDomain.y && maxDomain.z > minDomain.z))
      throw std::invalid_argument("invalid system domain");
    systems_.push_back({minDomain, maxDomain, periodicX, periodicY});
    const int k = (int)systems_.size() - 1;
    particles_.setCurrentSystem(k);
    walls_.setCurrentSystem(k);
    return k;
  }
  size_t numSystems() const { return systems_.size(); }

  // Cell size of the particle grid (the reference derives 2 * max radius,
  // LSParticle.h:383). Must be at least the largest template radius.
  void setCellSize(double cellSize) { cellSize_ = cellSize; }
  void setContactCapacity(int pairsPerParticle, int recordsPerParticle) {
    pairsPerParticle_ = std::max(pairsPerParticle, 1);
    recordsPerParticle_ = std::max(recordsPerParticle, 1);
  }
  void useGraph(bool on) { useGraph_ = on; }
  // Launch shapes of this solver's narrow phase and reduction (results do
  // not depend on them); zero fields take the process default. Starts at
  // the process default of the time of construction.
  void setLaunchTuning(const fdm_tuning& t) {
    fdm_tuning n = t;
    fdmCheck(fdm_checkTuning(&n), "launch tuning");
    tuning_ = n;
    destroyGraphs();
  }
  const fdm_tuning& launchTuning() const { return tuning_; }
  void setStageTiming(bool on) { timing_ = on; }
  // Extra reach of the candidate lists, which are then rebuilt only when a
  // body may have moved by half of it (decided on the device each step).
  // Results do not depend on it. Negative: a tenth of the largest particle
  // radius. Takes effect at the next initialize.
  void setNeighborSkin(double skin) { skinRequest_ = skin; }
  double neighborSkin() const { return skin_; }
  // Candidate-list builds and ghost-list builds since initialize; synchronises.
  std::pair<long, long> listBuilds() {
    if (!ctl_.get()) return {0, 0};
    std::vector<int> c = ctl_.toHost(stream_, FDM_CTL_WORDS);
    return {c[FDM_CTL_COUNT], c[FDM_CTL_GHOST_COUNT]};
  }
  // Accumulated device time per stage (ms) and the number of launches timed.
  const std::map<std::string, double>& stageMilliseconds() {
    resolveTimers();
    return stageMs_;
  }
  const std::map<std::string, long>& stageCalls() {
    resolveTimers();
    return stageCalls_;
  }
  void resetStageTiming() {
    resolveTimers();
    stageMs_.clear();
    stageCalls_.clear();
  }

  // LSSolver::solve (LSSolver.h:440-467).
  void solve(const double3 minDomain, const double3 maxDomain, const double3 gravity,
             const double timeStep, const double maximumTime, const size_t numFrame, const int argc,
             char** argv) {
    if (!checkSolverInput(minDomain, maxDomain, timeStep, maximumTime)) return;
    if (!activateGPUDevice()) return;
    phase_ += 1;
    if (writeVTU_) {
      updateDir(argc, argv);
      removeFrameFiles();
    }
    initialize(minDomain, maxDomain, gravity, timeStep);
    const size_t numStep = size_t(maximumTime / timeStep) + 1;
    size_t frameInterval = numStep;
    if (numFrame > 0) frameInterval = numStep / numFrame;
    if (frameInterval < 1) frameInterval = 1;
    size_t iStep = 0, iFrame = 0;
    double time = 0.;
    output(iFrame, iStep, time);
    while (iStep <= numStep) {
      stepOnce(false, time);
      iStep += 1;
      time += timeStep;
      if (iStep % frameInterval == 0) {
        iFrame++;
        output(iFrame, iStep, time);
      }
    }
    iFrame++;
    output(iFrame, iStep, time);
    std::cout << "[Solver] Computation Completed" << std::endl;
  }

  // Upload and the pre-loop build of the compute* variants
  // (LSSolver.h:1097-1100): grid, contact lists (springs carried over, none
  // updated) and, with periodicity, the wrap and ghost sets.
  void initialize(const double3 minDomain, const double3 maxDomain, const double3 gravity,
                  const double timeStep) {
    if (!checkSolverInput(minDomain, maxDomain, timeStep, 0.0))
      throw std::invalid_argument("invalid domain or time step");
    hipCheck(hipSetDevice(device_), "hipSetDevice");
    if (!stream_) {
      // The null stream cannot be captured into a graph: work on a stream of
      // our own that still synchronises with the null stream, so hooks that
      // launch on stream 0 stay ordered with the step.
      hipCheck(hipStreamCreate(&ownStream_), "hipStreamCreate");
      stream_ = ownStream_;
    }
    gravity_ = gravity;
    dt_ = timeStep;
    domainMin_ = minDomain;
    domainMax_ = maxDomain;
    if (systems_.empty()) {
      particles_.initialize(minDomain, maxDomain, cellSize_, stream_);
      walls_.initialize(minDomain, maxDomain, 0.0, stream_);
      inst_ = fdm_instances{};
    } else {
      initializeSystems();
    }
    particles_.setLaunchShape(maxGPUThread_);
    walls_.setLaunchShape(maxGPUThread_);
    pp_.setLaunchShape(maxGPUThread_);
    pw_.setLaunchShape(maxGPUThread_);
    double maxR = 0.0;
    for (size_t g = 0; g < particles_.numGeometry(); ++g)
      maxR = std::max(maxR, particles_.templateInfo(g).radius);
    if (cellSize_ > 0.0 && cellSize_ < maxR)
      throw std::invalid_argument("cell size smaller than the largest particle radius");
    const size_t N = particles_.num(), W = walls_.num();
    maxRadius_ = maxR;
    maxBoxReach_ = 0.0;
    rhoMax_ = 0.0;
    for (size_t g = 0; g < particles_.numGeometry(); ++g) {
      const fdm_template& t = particles_.templateInfo(g);
      maxBoxReach_ = std::max(maxBoxReach_, t.boxReach);
      rhoMax_ = std::max({rhoMax_, t.radius, t.boxReach});
    }
    for (size_t g = 0; g < walls_.numGeometry(); ++g) {
      const fdm_template& t = walls_.templateInfo(g);
      rhoMax_ = std::max({rhoMax_, t.radius, t.boxReach});
    }
    if (periodic_.isSector() && periodic_.isPlanar())
      throw std::invalid_argument("the sector periodic boundary excludes planar periodicity");
    if (periodic_.isSector() && !systems_.empty())
      throw std::invalid_argument("the sector periodic boundary runs as one system");
    skin_ = skinRequest_ >= 0.0 ? skinRequest_ : 0.1 * maxR;
    if (periodic_.isSector()) skin_ = 0.0;  // images are rebuilt every step
    ctl_.allocate(FDM_CTL_WORDS);
    disp_.allocate(1);
    {
      const int one = 1;  // the first step builds every list
      hipCheck(hipMemcpy(ctl_.get() + FDM_CTL_FORCE, &one, sizeof(int), hipMemcpyHostToDevice),
               "ctl");
    }
    const int pairCap = (int)std::max<size_t>(size_t(pairsPerParticle_) * N, 1024);
    // records come from FDM_REC_POOLS pools: each gets a share plus 1024
    const int recCap = (int)(std::max<size_t>(size_t(recordsPerParticle_) * N, 4096) +
                             size_t(FDM_REC_POOLS) * 1024);
    pp_.initialize(N, pairCap, recCap, mirrorCapacity_);
    // a particle can pair with every wall of its own system
    size_t wallsPerSystem = W;
    if (!systems_.empty()) {
      std::vector<size_t> perSystem(systems_.size(), 0);
      for (int k : walls_.systemOfBody()) perSystem[k] += 1;
      wallsPerSystem = *std::max_element(perSystem.begin(), perSystem.end());
    }
    const long long wallPairCap =
        std::max<long long>((long long)N * (long long)std::max<size_t>(wallsPerSystem, 1), 1024);
    if (wallPairCap >= (1ll << 31)) throw std::invalid_argument("too many wall pairs");
    pw_.initialize(N, (int)wallPairCap, recCap, 1);
    periodic_.initialize(particles_, (int)std::max<size_t>(N, 64),
                         (int)std::max<size_t>(4 * N, 1024), recCap, mirrorCapacity_);
    pp_.resetGenerations();
    pw_.resetGenerations();
    for (int s = 0; s < 3; ++s)
      if (periodic_.active(s)) periodic_.set(s).interaction.resetGenerations();
    {
      // work list of the narrow phase: every candidate pair of every set
      long long cap = (long long)pairCap + wallPairCap;
      for (int s = 0; s < 3; ++s)
        if (periodic_.active(s)) cap += std::max<long long>(4 * (long long)N, 1024);
      if (cap >= (1ll << 28)) throw std::invalid_argument("contact capacities too large");
      workHit_.allocate((size_t)cap, false);
      lawHead_.allocate(8 * (size_t)cap, false);
      // staged contacts of one step, all sets together: room for every
      // active set's records
      int activeSets = 1 + (W ? 1 : 0);
      for (int s = 0; s < 3; ++s)
        if (periodic_.active(s)) activeSets += 1;
      const long long hitCap = (long long)recCap * activeSets;
      if (hitCap >= (1ll << 31)) throw std::invalid_argument("contact record capacity too large");
      hitCount_.allocate(FDM_REC_COUNTERS);
      hitGeom_.allocate(8 * (size_t)hitCap, false);
      hitCapacity_ = (int)hitCap;
    }
    status_.allocate(1);
    stepCounter_.allocate(1);
    bonds_.initialize(N, stepCounter_.get(), stream_);
    size_t ws = fdm_radixSortWorkspaceBytes(std::max<size_t>({N, W, size_t(64)}));
    ws = std::max(ws, fdm_scanWorkspaceBytes(N + 1));
    ws = std::max(ws, fdm_stepBeginWorkspaceBytes(N + W));
    workspace_.allocate(ws / sizeof(double) + 1, false);
    destroyGraphs();
    if (!hostDecision_) {
      hipCheck(hipHostMalloc((void**)&hostDecision_, FDM_DECISION_RING * sizeof(int),
                             hipHostMallocMapped | hipHostMallocCoherent),
               "hipHostMalloc");
      hipCheck(hipHostGetDevicePointer((void**)&hostDecisionDev_, hostDecision_, 0),
               "hipHostGetDevicePointer");
    }
    hipCheck(hipStreamSynchronize(stream_), "initialize");
    for (int k = 0; k < FDM_DECISION_RING; ++k) hostDecision_[k] = 0;
    decisionSeq_ = 0;
    hookMode_ = HookUnknown;
    stepOnce(true, 0.0);
    hipCheck(hipStreamSynchronize(stream_), "initialize");
    checkStatus();
  }

  // Advance nSteps loop iterations without output (the loop body of
  // compute*, LSSolver.h:1101-1110 and variants).
  // Throws if a capacity overflowed during these steps (the device status
  // word is read once, at the end).
  void advance(size_t nSteps, double time = 0.0) {
    if (nSteps == 0) return;
    // the first step after initialize runs alone (it finds out whether
    // addExternalForceTorque acts); hooks and sector images need the
    // host at every step; stage timing keeps one step in flight
    size_t done = 0;
    if (hookMode_ == HookUnknown) {
      stepOnce(false, time);
      time += dt_;
      done = 1;
    }
    if (done < nSteps && hookMode_ == HookNone && !periodic_.isSector() && !timing_ &&
        lookahead_ > 0) {
      advanceAhead(nSteps - done, time);
    } else {
      const bool graphs = useGraph_ && hookMode_ == HookNone && !timing_;
      for (; done < nSteps; ++done) {
        if (graphs) graphStep();
        else stepOnce(false, time);
        time += dt_;
      }
    }
    checkStatus();
  }
  // Steps queued ahead of the list decisions the host has read (advance);
  // 0 waits for every step's decision before queueing the next. At most
  // FDM_DECISION_RING - 1; default 8 (two graphs of four steps).
  void setLookahead(int steps) {
    lookahead_ = std::max(0, std::min(steps, FDM_DECISION_RING - 1));
  }

  // LSSolver's finalize of the containers (copyDeviceToHost): the state,
  // the bonds and the host copies of the particle and wall contacts.
  void download() {
    downloadState();
    if (particles_.num_device() > 0) {
      fillHostContacts(FDM_SET_PARTICLE, pp_);
      fillHostContacts(FDM_SET_WALL, pw_);
    }
  }
  // Bodies and bonds only.
  void downloadState() {
    particles_.finalize(stream_);
    walls_.finalize(stream_);
    bonds_.finalize(stream_);
    hipCheck(hipStreamSynchronize(stream_), "download");
  }

  // Device status word (capacity overflows); synchronises.
  unsigned status() {
    if (!status_.get()) return 0;
    unsigned v = 0;
    hipCheck(hipMemcpyAsync(&v, status_.get(), sizeof(unsigned), hipMemcpyDeviceToHost, stream_),
             "status");
    hipCheck(hipStreamSynchronize(stream_), "status");
    return v;
  }
  void checkStatus() {
    const unsigned s = status();
    if (s != 0)
      throw std::runtime_error("contact capacity exceeded (status " + std::to_string(s) +
                               "); raise it with setContactCapacity");
  }

  // Candidate pairs of each set (particle, wall, ghost X, Y, XY) at the
  // latest step; synchronises.
  std::vector<int> pairTotals() {
    std::vector<int> out(5, 0);
    out[0] = pp_.lastTotal(stream_);
    out[1] = pw_.lastTotal(stream_);
    for (int s = 0; s < 3; ++s)
      if (periodic_.active(s)) out[2 + s] = periodic_.set(s).interaction.lastTotal(stream_);
    return out;
  }

  // Contacts of the latest step, canonical order. set: 0 particle-particle,
  // 1 particle-wall, 2/3/4 ghost X/Y/XY.
  std::vector<ContactRecord> contacts(int set) {
    if (set == 0) return pp_.contacts(stream_);
    if (set == 1) return pw_.contacts(stream_);
    if (set >= 2 && set <= 4 && periodic_.active(set - 2))
      return periodic_.set(set - 2).interaction.contacts(stream_);
    return {};
  }

  // One node-contact of the latest step with its geometry.
  struct ContactGeometry {
    int A, target, node;
    double3 point, normal;
    double overlap;
    double3 spring;
  };
  // Contact points, normals and overlaps of the latest step's contacts of a
  // set (0 particle, 1 wall, 2/3/4 ghost X/Y/XY), recomputed on the host from
  // the downloaded state and the templates with the detection kernels'
  // arithmetic (LSParticleContactDetectionKernel.cu:128-305): the node in the
  // target's frame, the trilinear value and int-x gradient of its SDF cell,
  // the normal and the point p + n delta / 2. The reference keeps these per
  // contact (SolidInteraction contactPoint / contactNormal / overlap).
  // Synchronises.
  std::vector<ContactGeometry> contactGeometry(int set) {
    downloadState();
    return contactGeometryHost(set);
  }
  // The host copies of a set's contacts (SolidInteraction's *HostRef
  // getters) from the state last downloaded.
  void fillHostContacts(int set, SolidInteraction& si) {
    const std::vector<ContactGeometry> cg = contactGeometryHost(set);
    std::vector<long long> nodeBase(particles_.num() + 1, 0);
    const auto& geo = particles_.geometryIDHostRef();
    for (size_t i = 0; i < particles_.num(); ++i)
      nodeBase[i + 1] = nodeBase[i] + particles_.templateInfo(geo[i]).nodeCount;
    std::vector<double3> pt, nr;
    std::vector<int> master, slave;
    pt.reserve(cg.size());
    nr.reserve(cg.size());
    master.reserve(cg.size());
    slave.reserve(cg.size());
    for (const ContactGeometry& c : cg) {
      pt.push_back(c.point);
      nr.push_back(c.normal);
      master.push_back((int)(nodeBase[c.A] + c.node));
      slave.push_back(c.target);
    }
    si.setHostContacts(std::move(pt), std::move(nr), std::move(master), std::move(slave),
                       maxGPUThread_);
  }
  // The same from the state last downloaded.
  std::vector<ContactGeometry> contactGeometryHost(int set) {
    const std::vector<ContactRecord> recs = contacts(set);
    std::vector<ContactGeometry> out;
    out.reserve(recs.size());
    const auto& px = particles_.positionHostRef();
    const auto& pq = particles_.orientationHostRef();
    const auto& geo = particles_.geometryIDHostRef();
    const bool wall = set == FDM_SET_WALL;
    LSParticle& tb = wall ? walls_ : particles_;
    for (const ContactRecord& r : recs) {
      const fdm_template& tA = particles_.templateInfo(geo[r.A]);
      const fdm_real* nd = particles_.nodesHost().data() + 3 * ((size_t)tA.nodeStart + r.node);
      const double3 s = make_double3(nd[0], nd[1], nd[2]);
      const double3 p = vec::add(vec::rotate(pq[r.A], s), px[r.A]);
      double3 xB = tb.positionHostRef()[r.target];
      const quaternion qB = tb.orientationHostRef()[r.target];
      if (set >= FDM_SET_GHOST) {
        const fdm_grid& g = systems_.empty() ? particles_.grid() : sysGrid_[particles_.systemOfBody()[r.target]];
        const int mask = set - FDM_SET_GHOST + 1;
        if (mask & 1) xB.x += g.max.x - g.min.x;
        if (mask & 2) xB.y += g.max.y - g.min.y;
      }
      const fdm_template& tB = tb.templateInfo(tb.geometryIDHostRef()[r.target]);
      const double3 l = vec::rotateT(vec::sub(p, xB), qB);
      const double gx = (l.x - tB.origin[0]) * tB.invSpacing, gy = (l.y - tB.origin[1]) * tB.invSpacing,
                   gz = (l.z - tB.origin[2]) * tB.invSpacing;
      const int i0 = (int)std::floor(gx), j0 = (int)std::floor(gy), k0 = (int)std::floor(gz);
      ContactGeometry c{r.A, r.target, r.node, make_double3(0, 0, 0), make_double3(0, 0, 0), 0.0, r.spring};
      if (i0 >= 0 && j0 >= 0 && k0 >= 0 && i0 < tB.size[0] - 1 && j0 < tB.size[1] - 1 && k0 < tB.size[2] - 1) {
        const double x = gx - (double)i0, y = gy - (double)j0, z = gz - (double)k0;
        const int nx = tB.size[0], nxy = tB.size[0] * tB.size[1];
        const float* b = tb.sdfHost().data() + tB.sdfStart + ((size_t)k0 * nxy + (size_t)j0 * nx + i0);
        const double cc[8] = {b[0], b[1], b[nx], b[nx + 1], b[nxy], b[nxy + 1], b[nxy + nx], b[nxy + nx + 1]};
        const double wx0 = 1.0 - x, wx1 = x, wy0 = 1.0 - y, wy1 = y, wz0 = 1.0 - z, wz1 = z;
        const double phi = cc[0] * wx0 * wy0 * wz0 + cc[1] * wx1 * wy0 * wz0 + cc[2] * wx0 * wy1 * wz0 +
                           cc[3] * wx1 * wy1 * wz0 + cc[4] * wx0 * wy0 * wz1 + cc[5] * wx1 * wy0 * wz1 +
                           cc[6] * wx0 * wy1 * wz1 + cc[7] * wx1 * wy1 * wz1;
        // the int-x gradient weights (LSParticleContactDetectionKernel.cuh:34-77)
        const double ix0 = 1.0, ix1 = 0.0;
        const double3 grad = make_double3(
            (cc[1] - cc[0]) * wy0 * wz0 + (cc[3] - cc[2]) * wy1 * wz0 + (cc[5] - cc[4]) * wy0 * wz1 +
                (cc[7] - cc[6]) * wy1 * wz1,
            (cc[2] - cc[0]) * ix0 * wz0 + (cc[3] - cc[1]) * ix1 * wz0 + (cc[6] - cc[4]) * ix0 * wz1 +
                (cc[7] - cc[5]) * ix1 * wz1,
            (cc[4] - cc[0]) * ix0 * wy0 + (cc[5] - cc[1]) * ix1 * wy0 + (cc[6] - cc[2]) * ix0 * wy1 +
                (cc[7] - cc[3]) * ix1 * wy1);
        c.overlap = -phi;
        c.normal = vec::unit(vec::rotate(qB, grad));
        c.point = vec::add(p, vec::mul(c.normal, 0.5 * c.overlap));
      }
      out.push_back(c);
    }
    return out;
  }

  // LSSolver::createBondsFromLSInteractions (LSSolver.h:347-380): one bond per
  // particle-particle node-contact of the latest step, between its particles,
  // at its contact point with its normal. The bonds join the next initialise.
  void createBondsFromLSInteractions(const double radius, const double initialLength,
                                     const double YoungsModulus, const double poissonRatio,
                                     const double tensileStrength = 0., const double cohesion = 0.,
                                     const double frictionCoefficient = 0.) {
    for (const ContactGeometry& c : contactGeometry(FDM_SET_PARTICLE))
      addBondedInteraction(c.A, c.target, c.point, c.normal, radius, initialLength, YoungsModulus,
                           poissonRatio, tensileStrength, cohesion, frictionCoefficient);
  }

  // The frame files of LSSolver::output (LSSolver.h:717-745) for the current
  // state, under `dir`:
  //   LSParticle/LSObjectMesh_NNNN.vtu         surface nodes and triangles of the particles
  //   LSParticle/Particle_NNNN.vtu             one point per particle
  //   LSParticleInteraction/LSInteraction_NNNN.vtu       particle-particle node contacts
  //   LSParticleInteraction/VBondedInteraction_NNNN.vtu  bonds, when there are any
  //   Wall/LSObjectMesh_NNNN.vtu               wall surfaces
  //   LSParticle-WallInteraction/LSInteraction_NNNN.vtu  particle-wall node contacts
  // Contacts are those of the latest step (none before the first). Ghost
  // contacts of periodic sets are not written, as in the reference.
  // Synchronises.
  void writeVTU(const std::string& dir, size_t iFrame, size_t iStep, double time) {
    downloadState();
    vtu::makeDirs(dir);
    const std::string dp = dir + "/LSParticle", di = dir + "/LSParticleInteraction",
                      dw = dir + "/Wall", dpw = dir + "/LSParticle-WallInteraction";
    writeObjectMesh(particles_, dp, iFrame, iStep, time);
    writeParticlePoints(dp, iFrame, iStep, time);
    writeInteractions(FDM_SET_PARTICLE, di, iFrame, iStep, time);
    writeBonds(di, iFrame, iStep, time);
    writeObjectMesh(walls_, dw, iFrame, iStep, time);
    writeInteractions(FDM_SET_WALL, dpw, iFrame, iStep, time);
  }
  // The directory LSSolver::output writes a frame to: the solver's directory
  // with the phase appended (LSSolver.h:732).
  std::string frameDirectory() const { return dir_ + "_phase" + std::to_string(phase_); }
  // Write (default) or skip the frame files at each output frame of solve().
  void setWriteVTU(bool on) { writeVTU_ = on; }

  LSParticle& getLSParticle() { return particles_; }
  LSParticle& getWall() { return walls_; }
  SolidInteraction& getLSParticleInteraction() { return pp_; }
  SolidInteraction& getLSParticleWallInteraction() { return pw_; }
  VBondedInteraction& getLSParticleBondedInteraction() { return bonds_; }
  hipStream_t stream() const { return stream_; }
  double timeStep() const { return dt_; }

 protected:
  // LSSolver::addParticleGeometry(localOffset, density, ...) (LSSolver.h:470-574):
  // mass, centroid and inertia from the smooth Heaviside of the SDF; nodes
  // and SDF origin shifted by -centroid.
  void addParticleGeometry(double3& localOffset, double density, double3 gridNodeOrigin,
                           double gridNodeSpacing, int3 gridNodeSize,
                           const std::vector<double>& gridNodeSignedDistance,
                           const std::vector<double3>& boundaryNodePosition,
                           const std::vector<int3>& boundaryNodeConnectivity) {
    localOffset = make_double3(0., 0., 0.);
    if (gridNodeSize.x < 2 || gridNodeSize.y < 2 || gridNodeSize.z < 2) {
      std::cerr << "[LSSolver] Invalid grid node size: " << gridNodeSize.x << " x " << gridNodeSize.y
                << " x " << gridNodeSize.z << ". Each dimension must be at least 2." << std::endl;
      return;
    }
    if (gridNodeSpacing <= 0.0) {
      std::cerr << "[LSSolver] Invalid grid node spacing: " << gridNodeSpacing
                << ". Must be greater than 0." << std::endl;
      return;
    }
    const size_t expected =
        size_t(gridNodeSize.x) * size_t(gridNodeSize.y) * size_t(gridNodeSize.z);
    if (expected != gridNodeSignedDistance.size()) {
      std::cerr << "[LSSolver] Inconsistent grid node data size. Expected " << expected
                << " signed distance values, but got " << gridNodeSignedDistance.size() << "."
                << std::endl;
      return;
    }
    auto heaviside = [](double phi, double eps) -> double {
      if (eps <= 0.0) return (phi > 0.0) ? 0.0 : 1.0;
      if (phi < -eps) return 1.0;
      if (phi > eps) return 0.0;
      const double x = -phi / eps;
      return 0.5 * (1.0 + x + std::sin(pi() * x) / pi());
    };
    const double h = gridNodeSpacing;
    const double mg = density * h * h * h;
    const int3 sz = gridNodeSize;
    double mass = 0.;
    for (int x = 0; x < sz.x; x++)
      for (int y = 0; y < sz.y; y++)
        for (int z = 0; z < sz.z; z++)
          mass += heaviside(gridNodeSignedDistance[linearIndex3D(make_int3(x, y, z), sz)] / h, 1.5);
    mass *= mg;
    double3 c = make_double3(0., 0., 0.);
    symMatrix I;
    if (mass > 0.) {
      for (int x = 0; x < sz.x; x++)
        for (int y = 0; y < sz.y; y++)
          for (int z = 0; z < sz.z; z++) {
            const double H = heaviside(gridNodeSignedDistance[linearIndex3D(make_int3(x, y, z), sz)] / h, 1.5);
            c.x += H * (gridNodeOrigin.x + double(x) * h);
            c.y += H * (gridNodeOrigin.y + double(y) * h);
            c.z += H * (gridNodeOrigin.z + double(z) * h);
          }
      c = vec::mul(c, mg / mass);
      for (int x = 0; x < sz.x; x++)
        for (int y = 0; y < sz.y; y++)
          for (int z = 0; z < sz.z; z++) {
            const double H = heaviside(gridNodeSignedDistance[linearIndex3D(make_int3(x, y, z), sz)] / h, 1.5);
            const double3 r = vec::sub(
                vec::add(gridNodeOrigin, vec::mul(make_double3(double(x), double(y), double(z)), h)), c);
            I.xx += H * (r.y * r.y + r.z * r.z) * mg;
            I.yy += H * (r.x * r.x + r.z * r.z) * mg;
            I.zz += H * (r.y * r.y + r.x * r.x) * mg;
            I.xy -= H * r.x * r.y * mg;
            I.xz -= H * r.x * r.z * mg;
            I.yz -= H * r.y * r.z * mg;
          }
    }
    std::vector<double3> shifted;
    shifted.reserve(boundaryNodePosition.size());
    for (const auto& p : boundaryNodePosition) shifted.push_back(vec::sub(p, c));
    particles_.addGeometry(mass, I, vec::sub(gridNodeOrigin, c), gridNodeSpacing, gridNodeSize,
                           gridNodeSignedDistance, shifted, boundaryNodeConnectivity);
    localOffset = c;
  }

  // User hook between the force reduction and the second half-kick
  // (LSSolver.h:588). Device force/torque are FP32 planar SoA.
  virtual void addExternalForceTorque(const double time) {
    (void)time;
    baseHookCalled_ = true;
  }

  // Detect again, at the next step, whether addExternalForceTorque is
  // overridden to act (for a subclass that turns its hook on or off).
  void resetHookDetection() {
    hookMode_ = HookUnknown;
    destroyGraphs();
  }

  // Called at every output frame after the state is downloaded.
  virtual void outputFrame(size_t iFrame, size_t iStep, double time) {
    (void)iFrame;
    (void)iStep;
    (void)time;
  }

 private:
  enum HookMode { HookUnknown, HookNone, HookUser };

  // LSParticle::outputVTU (LSParticle.h:426-600): every surface node at
  // x + R(q) s with its rigid-body velocity v + w x (p - x) and its particle
  // id; one triangle per template triangle (VTK_TRIANGLE), node ids offset by
  // the nodes of the bodies before.
  void writeObjectMesh(const LSParticle& P, const std::string& dir, size_t iFrame, size_t iStep,
                       double time) {
    const size_t n = P.num();
    if (n == 0) return;
    const auto& x = P.positionHostRef();
    const auto& v = P.velocityHostRef();
    const auto& w = P.angularVelocityHostRef();
    const auto& q = P.orientationHostRef();
    const auto& geo = P.geometryIDHostRef();
    size_t N = 0, M = 0;
    for (size_t i = 0; i < n; ++i) {
      N += P.templateNodes(geo[i]).size();
      M += P.templateConnectivity(geo[i]).size();
    }
    std::vector<float> pts(3 * N), vel(3 * N);
    std::vector<int32_t> pid(N), conn(3 * M), offs(M);
    std::vector<uint8_t> types(M, 5);
    size_t a = 0, c = 0;
    for (size_t i = 0; i < n; ++i) {
      for (const int3& t : P.templateConnectivity(geo[i])) {
        conn[3 * c + 0] = int32_t(a + t.x);
        conn[3 * c + 1] = int32_t(a + t.y);
        conn[3 * c + 2] = int32_t(a + t.z);
        offs[c] = int32_t(3 * (c + 1));
        ++c;
      }
      for (const double3& s : P.templateNodes(geo[i])) {
        const double3 pw = vec::add(x[i], vec::rotate(q[i], s));
        const double3 vw = vec::add(v[i], vec::cross(w[i], vec::sub(pw, x[i])));
        put3(pts, a, pw);
        put3(vel, a, vw);
        pid[a] = int32_t(i);
        ++a;
      }
    }
    vtu::makeDirs(dir);
    vtu::Piece f(time, iStep, N, M);
    f.points(pts);
    f.cells(conn, offs, types);
    f.beginPointData("velocity");
    f.pointArray("particleID", pid, 1);
    f.pointArray("velocity", vel, 3);
    f.write(vtu::framePath(dir, "LSObjectMesh", iFrame));
  }
  // LSParticle::outputParticleVTU (LSParticle.h:602-756): one point per
  // particle with velocity, angular velocity and orientation (q0 q1 q2 q3).
  void writeParticlePoints(const std::string& dir, size_t iFrame, size_t iStep, double time) {
    const size_t n = particles_.num();
    if (n == 0) return;
    std::vector<float> pts(3 * n), vel(3 * n), ang(3 * n), ori(4 * n);
    for (size_t i = 0; i < n; ++i) {
      put3(pts, i, particles_.positionHostRef()[i]);
      put3(vel, i, particles_.velocityHostRef()[i]);
      put3(ang, i, particles_.angularVelocityHostRef()[i]);
      const quaternion& q = particles_.orientationHostRef()[i];
      ori[4 * i + 0] = float(q.q0);
      ori[4 * i + 1] = float(q.q1);
      ori[4 * i + 2] = float(q.q2);
      ori[4 * i + 3] = float(q.q3);
    }
    vtu::makeDirs(dir);
    vtu::Piece f(time, iStep, n, 0);
    f.points(pts);
    f.noCells();
    f.beginPointData("velocity");
    f.pointArray("velocity", vel, 3);
    f.pointArray("angularVelocity", ang, 3);
    f.pointArray("orientation", ori, 4);
    f.write(vtu::framePath(dir, "Particle", iFrame));
  }
  // SolidInteraction::outputVTU (SolidInteraction.h:153-316): one vertex per
  // node-contact at its contact point, with the normal, the direction of the
  // sliding spring and the elastic energies 0.5 kn delta^2 and
  // 0.5 ks |spring|^2 (contactKernel.cu:118-119 for particle pairs, with the
  // series stiffnesses; :201-202 for walls, with the particle's own).
  void writeInteractions(int set, const std::string& dir, size_t iFrame, size_t iStep, double time) {
    const std::vector<ContactGeometry> cg =
        particles_.num_device() > 0 ? contactGeometryHost(set) : std::vector<ContactGeometry>{};
    const size_t N = cg.size();
    std::vector<float> pts(3 * N), nrm(3 * N), dir3(3 * N), eN(N), eS(N);
    std::vector<int32_t> conn(N), offs(N);
    std::vector<uint8_t> types(N, 1);
    const auto& matA = particles_.materialIDHostRef();
    for (size_t k = 0; k < N; ++k) {
      const ContactGeometry& c = cg[k];
      const fdm_material& mi = particles_.materialInfo(matA[c.A]);
      double kn = mi.kn, ks = mi.ks;
      if (set != FDM_SET_WALL) {
        const fdm_material& mj = particles_.materialInfo(matA[c.target]);
        kn = (mi.kn > 0. && mj.kn > 0.) ? mi.kn * mj.kn / (mi.kn + mj.kn) : 0.;
        ks = (mi.ks > 0. && mj.ks > 0.) ? mi.ks * mj.ks / (mi.ks + mj.ks) : 0.;
      }
      put3(pts, k, c.point);
      put3(nrm, k, c.normal);
      put3(dir3, k, vec::unit(c.spring));
      eN[k] = float(0.5 * c.overlap * c.overlap * kn);
      eS[k] = float(0.5 * vec::dot(c.spring, c.spring) * ks);
      conn[k] = int32_t(k);
      offs[k] = int32_t(k + 1);
    }
    vtu::makeDirs(dir);
    vtu::Piece f(time, iStep, N, N);
    f.points(pts);
    f.cells(conn, offs, types);
    f.beginPointData("");
    f.pointArray("contactNormal", nrm, 3);
    f.pointArray("slidingDirection", dir3, 3);
    f.pointArray("normalElasticEnergy", eN, 1);
    f.pointArray("slidingElasticEnergy", eS, 1);
    f.write(vtu::framePath(dir, "LSInteraction", iFrame));
  }
  // VBondedInteraction::outputVTU (VBondedInteraction.h:262-448): one vertex
  // per bond at the midpoint of its two bond points, with its ends, state,
  // the elastic energies of VBond (contactKernel.cuh:170-173) evaluated on
  // the downloaded state from the stored FP32 frames, and the last step's
  // stresses (those of the failing step for a bond that just failed, zero
  // after). A failed bond reports zero energies; its point is the midpoint
  // of its bond points in the current state.
  void writeBonds(const std::string& dir, size_t iFrame, size_t iStep, double time) {
    const size_t N = bonds_.numPair();
    if (N == 0) return;
    const auto& x = particles_.positionHostRef();
    const auto& q = particles_.orientationHostRef();
    const auto& act = bonds_.activatedHostRef();
    const bool onDevice = bonds_.numPair_device() > 0;
    const std::vector<double> sig = onDevice ? bonds_.maxNormalStress(stream_) : std::vector<double>(N, 0.);
    const std::vector<double> tau = onDevice ? bonds_.maxShearStress(stream_) : std::vector<double>(N, 0.);
    std::vector<float> pts(3 * N), un(N), us(N), ub(N), ut(N), sn(N), ss(N);
    std::vector<int32_t> conn(N), offs(N), mid(N), sid(N), on(N);
    std::vector<uint8_t> types(N, 1);
    for (size_t k = 0; k < N; ++k) {
      const int i = bonds_.masterObjectIDHostRef()[k], j = bonds_.slaveObjectIDHostRef()[k];
      const fdm_real* L = bonds_.localHost().data() + 24 * k;
      const fdm_real* C = bonds_.coefHost().data() + 9 * k;
      auto lv = [&](int o) { return make_double3(L[o], L[o + 1], L[o + 2]); };
      const double3 n1i = vec::rotate(q[i], lv(0)), n2i = vec::rotate(q[i], lv(3)),
                    n3i = vec::rotate(q[i], lv(6));
      const double3 n1j = vec::rotate(q[j], lv(12)), n2j = vec::rotate(q[j], lv(15)),
                    n3j = vec::rotate(q[j], lv(18));
      const double3 rbi = vec::add(vec::rotate(q[i], lv(9)), x[i]);
      const double3 rbj = vec::add(vec::rotate(q[j], lv(21)), x[j]);
      put3(pts, k, vec::mul(vec::add(rbi, rbj), 0.5));
      const int active = k < act.size() ? act[k] : 1;
      if (active) {
        const double B1 = C[0], B2 = C[1], B3 = C[2], B4 = C[3], D0 = C[5];
        const double3 rij = vec::sub(rbi, rbj);
        const double D = vec::len(rij);
        if (!isZero(D)) {
          const double3 e = vec::neg(vec::unit(rij));
          const double c11 = vec::dot(n1i, n1j);
          un[k] = float(0.5 * B1 * (D - D0) * (D - D0));
          us[k] = float(B2 * (0.5 * vec::dot(vec::sub(n1j, n1i), e) - 0.25 * c11 + 0.75));
          ub[k] = float((0.25 * B2 + B3 + 0.5 * B4) * (c11 + 1.));
          ut[k] = float(-0.5 * B4 * (c11 + vec::dot(n2i, n2j) + vec::dot(n3i, n3j) - 1.));
        }
      }
      sn[k] = float(sig[k]);
      ss[k] = float(tau[k]);
      conn[k] = int32_t(k);
      offs[k] = int32_t(k + 1);
      mid[k] = i;
      sid[k] = j;
      on[k] = active;
    }
    vtu::makeDirs(dir);
    vtu::Piece f(time, iStep, N, N);
    f.points(pts);
    f.cells(conn, offs, types);
    f.beginPointData("");
    f.pointArray("masterObjectID", mid, 1);
    f.pointArray("slaveObjectID", sid, 1);
    f.pointArray("activated", on, 1);
    f.pointArray("normalElasticEnergy", un, 1);
    f.pointArray("shearElasticEnergy", us, 1);
    f.pointArray("bendingElasticEnergy", ub, 1);
    f.pointArray("torsionElasticEnergy", ut, 1);
    f.pointArray("maxNormalStress", sn, 1);
    f.pointArray("maxShearStress", ss, 1);
    f.write(vtu::framePath(dir, "VBondedInteraction", iFrame));
  }
  static void put3(std::vector<float>& a, size_t i, const double3& v) {
    a[3 * i + 0] = float(v.x);
    a[3 * i + 1] = float(v.y);
    a[3 * i + 2] = float(v.z);
  }
  // LSSolver::updateDir (LSSolver.h:611-644): a relative output directory is
  // taken from the nearest enclosing folder named "build" of the executable,
  // else the executable's folder (the working directory without argv).
  void updateDir(int argc, char** argv) {
    namespace fs = std::filesystem;
    const fs::path p(dir_);
    if (p.is_absolute()) return;
    fs::path base = fs::current_path();
    if (argc > 0 && argv && argv[0]) {
      const fs::path exe = fs::weakly_canonical(fs::absolute(fs::path(argv[0])));
      fs::path d = exe.parent_path();
      base = d;
      while (!d.empty() && d.filename() != "build") {
        const fs::path up = d.parent_path();
        if (up == d) break;
        d = up;
      }
      if (!d.empty() && d.filename() == "build") base = d;
    }
    dir_ = (base / p).string();
  }
  // LSSolver::removeFiles (LSSolver.h:678-690): old frames of this phase.
  void removeFrameFiles() {
    namespace fs = std::filesystem;
    const std::string dir = frameDirectory();
    vtu::makeDirs(dir);
    for (const char* sub : {"/LSParticle", "/LSParticleInteraction", "/Wall", "/LSParticle-WallInteraction"}) {
      std::error_code ec;
      if (!fs::is_directory(dir + sub, ec)) continue;
      for (const auto& e : fs::directory_iterator(dir + sub, ec))
        if (e.is_regular_file() && e.path().extension() == ".vtu") fs::remove(e.path(), ec);
    }
  }

  bool activateGPUDevice() {
    if (hipSetDevice(device_) != hipSuccess) {
      std::cout << "hipSetDevice( " << device_ << " ) failed! Do you have a HIP-capable GPU installed?"
                << std::endl;
      std::exit(1);
    }
    return true;
  }
  bool checkSolverInput(double3 mn, double3 mx, double dt, double T) {
    const double3 L = vec::sub(mx, mn);
    if (L.x <= 0.0 || L.y <= 0.0 || L.z <= 0.0) {
      std::cerr << "[Solver] Invalid simulation domain size: (" << L.x << ", " << L.y << ", " << L.z
                << ")." << std::endl;
      return false;
    }
    if (dt <= 0.0) {
      std::cerr << "[Solver] Invalid timeStep: " << dt << "." << std::endl;
      return false;
    }
    if (T < 0.0) {
      std::cerr << "[Solver] Invalid maximumTime: " << T << "." << std::endl;
      return false;
    }
    return true;
  }
  void output(size_t iFrame, size_t iStep, double time) {
    std::cout << "[Solver] Phase " << phase_ << ": ------ Frame " << iFrame << " at Time " << time
              << " ------ " << std::endl;
    download();
    checkStatus();
    outputFrame(iFrame, iStep, time);
    if (writeVTU_) {
      std::cout << "[Solver] Outputting... " << std::endl;
      writeVTU(frameDirectory(), iFrame, iStep, time);
      std::cout << "[Solver] Output Completed" << std::endl;
    }
  }

  // Brackets a stage with two events on the solver stream; no host
  // synchronisation until the times are read (resolveTimers).
  struct Timer {
    LSSolver* s;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    Timer(LSSolver* sv, const char* nm) : s(sv), name(nm) {
      if (!s->timing_) return;
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, s->stream_);
    }
    ~Timer() {
      if (!s->timing_) return;
      (void)hipEventRecord(b, s->stream_);
      s->pending_.push_back({name, a, b});
    }
  };
  struct PendingTimer {
    std::string name;
    hipEvent_t a, b;
  };
  void resolveTimers() {
    for (auto& p : pending_) {
      (void)hipEventSynchronize(p.b);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, p.a, p.b);
      stageMs_[p.name] += ms;
      stageCalls_[p.name] += 1;
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    pending_.clear();
  }

  struct SystemDesc {
    double3 min, max;
    bool px, py;
  };

  // Grids, cell offsets and instance tables of the systems; the bodies'
  // tables span every system's cells end to end.
  void initializeSystems() {
    const int K = (int)systems_.size();
    const double pcs = particles_.cellSizeRule(cellSize_);
    std::vector<fdm_grid> pg(K), wg(K);
    std::vector<long long> po(K), wo(K);
    std::vector<fdm_periodic> per(K);
    long long pc = 0, wc = 0;
    bool anyX = false, anyY = false;
    for (int k = 0; k < K; ++k) {
      const SystemDesc& d = systems_[k];
      pg[k] = makeGrid(d.min, d.max, pcs);
      wg[k] = makeGrid(d.min, d.max, walls_.cellSizeRuleOfSystem(k, 0.0));
      po[k] = pc;
      wo[k] = wc;
      pc += pg[k].num;
      wc += wg[k].num;
      per[k].x = d.px ? 1 : 0;
      per[k].y = d.py ? 1 : 0;
      anyX |= d.px;
      anyY |= d.py;
    }
    if (pc >= (1ll << 31) || wc >= (1ll << 31)) throw std::invalid_argument("too many cells");
    const std::vector<int>& ws = walls_.systemOfBody();
    std::vector<int> wallStart(K + 1, 0);
    for (size_t i = 0; i < ws.size(); ++i) {
      if (i > 0 && ws[i] < ws[i - 1]) throw std::invalid_argument("walls must be added system by system");
      wallStart[ws[i] + 1] += 1;
    }
    for (int k = 0; k < K; ++k) wallStart[k + 1] += wallStart[k];
    particles_.initializeSystems(pc, stream_);
    walls_.initializeSystems(wc, stream_);
    removePeriodicBoundary();
    if (anyX) addPeriodicBoundaryXD();
    if (anyY) addPeriodicBoundaryYD();
    instParticle_.upload(particles_.systemOfBody(), stream_);
    if (particles_.systemOfBody().empty()) instParticle_.allocate(1);
    instWall_.upload(ws, stream_);
    if (ws.empty()) instWall_.allocate(1);
    instWallStart_.upload(wallStart, stream_);
    sysGrid_ = pg;
    instPGrid_.upload(pg, stream_);
    instWGrid_.upload(wg, stream_);
    instPOff_.upload(po, stream_);
    instWOff_.upload(wo, stream_);
    instPeriodic_.upload(per, stream_);
    hipCheck(hipStreamSynchronize(stream_), "systems");
    inst_.count = K;
    inst_.particleInstance = instParticle_.get();
    inst_.wallInstance = instWall_.get();
    inst_.wallStart = instWallStart_.get();
    inst_.pgrid = instPGrid_.get();
    inst_.wgrid = instWGrid_.get();
    inst_.pcellOffset = instPOff_.get();
    inst_.wcellOffset = instWOff_.get();
    inst_.periodic = instPeriodic_.get();
  }

  fdm_skin skinView() {
    fdm_skin k;
    k.skin = skin_;
    k.rhoMax = rhoMax_;
    k.dt = dt_;
    k.ctl = ctl_.get();
    k.disp = disp_.get();
    k.hostDecision = hostDecisionDev_;
    return k;
  }

  // Broad phase of one set into its (new) current list generation.
  void broadPhase(int set, SolidInteraction& si, const fdm_ghosts* ghosts, fdm_grid grid,
                  fdm_hash_table table) {
    fdm_broadphase_args b{};
    b.skin = skin_;
    b.maxReach = set == FDM_SET_PARTICLE ? maxRadius_ : set == FDM_SET_GHOST ? maxBoxReach_ : 0.0;
    b.periodic = periodic_.flags();
    b.gate = nullptr;
    b.inst = inst_;
    b.set = set;
    b.particles = particles_.bodies();
    b.pgeo = particles_.geometry();
    b.walls = walls_.bodies();
    b.wgeo = walls_.geometry();
    b.ghosts = ghosts;
    b.grid = grid;
    b.table = table;
    b.pairs = si.current();
    Timer t(this, set == 0 ? "broad_pp" : set == 1 ? "broad_pw" : "broad_ghost");
    fdmCheck(fdm_launchBroadPhase(&b, workspace_.get(), workspace_.size() * sizeof(double), stream_),
             "broad phase");
    fdmCheck(fdm_launchBroadPhaseWrite(&b, status_.get(), stream_), "broad phase write");
  }

  fdm_narrowphase_args narrowArgs(int set, SolidInteraction& si, const fdm_ghosts* ghosts,
                                  fdm_grid grid, const int* targetHash, int ghostMask,
                                  bool ghostFromReal, bool prime) {
    fdm_narrowphase_args n{};
    n.set = set;
    n.particles = particles_.bodies();
    n.pgeo = particles_.geometry();
    n.pmat = particles_.deviceMaterials();
    n.walls = walls_.bodies();
    n.wgeo = walls_.geometry();
    n.wmat = walls_.deviceMaterials();
    n.ghosts = ghosts;
    n.grid = grid;
    n.targetHash = targetHash;
    n.pgrid = particles_.grid();
    n.periodic = periodic_.flags();
    n.pairs = si.current();
    n.prev = si.previous();
    n.mirror = si.mirror();
    n.ghostMask = ghostMask;
    n.ghostFromReal = ghostFromReal ? 1 : 0;
    n.inst = inst_;
    n.dt = dt_;
    n.updateSprings = prime ? 0 : 1;
    n.status = status_.get();
    return n;
  }

  // A step runs in two phases. Phase A (one launch: kick, drift, cell ids,
  // periodic visibility flags) ends with the device's decision whether the
  // candidate lists are rebuilt, written to host-visible memory. The host
  // waits for that word and launches phase B with only the work the step
  // needs: list builds when decided, then the narrow phase of every set in
  // one launch, bonds and the reduction with the second kick.
  void phaseA(bool prime, const int* gate = nullptr) {
    fdm_step_begin_args a{};
    a.gate = gate;
    a.particles = particles_.bodies();
    a.pgeo = particles_.geometry();
    a.pgrid = particles_.grid();
    a.ptable = particles_.table().view();
    a.walls = walls_.bodies();
    a.walls.n = (int)walls_.num();
    a.wgeo = walls_.geometry();
    a.wgrid = walls_.grid();
    a.wtable = walls_.table().view();
    a.gravity = {gravity_.x, gravity_.y, gravity_.z};
    a.dt = dt_;
    a.kick = prime ? 0 : 1;
    a.drift = prime ? 0 : 1;
    a.periodic = periodic_.flags();
    a.inst = inst_;
    if (periodic_.isActived()) {
      for (int s = 0; s < 3; ++s)
        a.ghostFlag[s] = nullptr;  // the ghost list build computes them
      a.ghostAtBuild = periodic_.view(periodic_.active(0) ? 0 : 1).atBuild;
    }
    int q = 0;
    a.clearCounters[q++] = pp_.nextRecordCounter();
    if (walls_.num()) a.clearCounters[q++] = pw_.nextRecordCounter();
    for (int s = 0; s < 3; ++s)
      if (periodic_.active(s)) a.clearCounters[q++] = periodic_.set(s).interaction.nextRecordCounter();
    a.clearCounters[q++] = hitCount_.get();
    const fdm_skin sk = skinView();
    const size_t wsBytes = workspace_.size() * sizeof(double);
    {
      Timer t(this, "step_begin");
      fdmCheck(fdm_launchStepBegin(&a, &sk, workspace_.get(), wsBytes, stream_), "step begin");
    }
    Timer t(this, "decide");
    fdmCheck(fdm_launchStepDecide(&a, &sk, workspace_.get(), wsBytes, stream_), "step decision");
  }

  struct Decision {
    bool rebuild, ghost;
  };
  // Spin on the host-visible decision of the phase A just launched.
  Decision waitDecision() { return waitDecisionFor(++decisionSeq_); }
  // Spin on the ring slot of decision number `number`.
  Decision waitDecisionFor(int number) {
    const unsigned seq = (unsigned)number & 0x3fffffffu;
    volatile int* h = hostDecision_ + (number & (FDM_DECISION_RING - 1));
    unsigned v = 0;
    for (long spin = 1;; ++spin) {
      v = (unsigned)h[0];
      if ((v >> 2) == seq) break;
      if ((spin & 1023) == 0) {
        const hipError_t e = hipStreamQuery(stream_);
        if (e == hipSuccess) {
          v = (unsigned)h[0];
          if ((v >> 2) == seq) break;
          throw std::runtime_error("list decision missing after the stream drained");
        }
        if (e != hipErrorNotReady) hipCheck(e, "step");
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return {(v & 2u) != 0, (v & 1u) != 0};
  }

  // New result generations for every set, new list generations where rebuilt.
  void beginSets(Decision d) {
    pp_.beginStep(d.rebuild);
    pw_.beginStep(d.rebuild);
    for (int s = 0; s < 3; ++s)
      if (periodic_.active(s)) periodic_.set(s).interaction.beginStep(d.ghost);
  }

  // List builds of a step (the part of phase B that runs only when the
  // decision asks for it).
  void phaseBRebuild(Decision d) {
    fdm_bodies P = particles_.bodies();
    const size_t N = particles_.num(), W = walls_.num();
    void* ws = workspace_.get();
    const size_t wsBytes = workspace_.size() * sizeof(double);
    const fdm_skin sk = skinView();
    if (d.rebuild) {
      {
        Timer t(this, "sort_cells");
        if (inst_.count > 0)
          fdmCheck(fdm_launchCommitBodyStateInstances(P, &inst_, 0, particles_.table().view(), &sk,
                                                      stream_),
                   "body state");
        else
          fdmCheck(fdm_launchCommitBodyState(P, particles_.grid(), particles_.table().view(), &sk,
                                             stream_),
                   "body state");
        fdmCheck(fdm_buildHashStartEnd(particles_.table().view(), (int)N, particles_.grid(), ws,
                                       wsBytes, stream_),
                 "hash table");
        if (W) {
          if (inst_.count > 0)
            fdmCheck(fdm_launchCommitBodyStateInstances(walls_.bodies(), &inst_, 1,
                                                        walls_.table().view(), &sk, stream_),
                     "wall state");
          else
            fdmCheck(fdm_launchCommitBodyState(walls_.bodies(), walls_.grid(),
                                               walls_.table().view(), &sk, stream_),
                     "wall state");
          fdmCheck(fdm_buildHashStartEnd(walls_.table().view(), (int)W, walls_.grid(), ws, wsBytes,
                                         stream_),
                   "wall hash table");
        }
      }
      broadPhase(FDM_SET_PARTICLE, pp_, nullptr, particles_.grid(), particles_.table().view());
      if (W) broadPhase(FDM_SET_WALL, pw_, nullptr, walls_.grid(), walls_.table().view());
    }
    if (periodic_.isActived() && d.ghost) {
      fdm_ghosts gv[3];
      for (int s = 0; s < 3; ++s) gv[s] = periodic_.active(s) ? periodic_.view(s) : fdm_ghosts{};
      {
        Timer t(this, "periodic_build");
        if (inst_.count > 0)
          fdmCheck(fdm_launchPeriodicBuildInstances(P, particles_.grid(), periodic_.flags(), &inst_,
                                                    gv, status_.get(), ws, wsBytes, stream_),
                   "periodic ghosts");
        else
          fdmCheck(fdm_launchPeriodicBuild(P, particles_.grid(), periodic_.flags(), gv, nullptr,
                                           status_.get(), ws, wsBytes, stream_),
                   "periodic ghosts");
      }
      for (int s = 0; s < 3; ++s)
        if (periodic_.active(s))
          broadPhase(FDM_SET_GHOST, periodic_.set(s).interaction, &gv[s], particles_.grid(),
                     periodic_.set(s).table.view());
    }
  }

  // The part of phase B every step runs: the narrow phase of all sets, bonds
  // and the reduction with the second kick. With `gate` each kernel runs
  // only if *gate != 0 (queued before the decision is known).
  void phaseBCommon(bool prime, double time, Decision d, const int* gate) {
    fdm_bodies P = particles_.bodies();
    fdm_geometry pgeo = particles_.geometry();
    const fdm_double3 g = {gravity_.x, gravity_.y, gravity_.z};
    const size_t N = particles_.num(), W = walls_.num();
    fdm_narrowphase_args na[5];
    int ns = 0;
    na[ns++] = narrowArgs(FDM_SET_PARTICLE, pp_, nullptr, particles_.grid(),
                          particles_.table().hashValue.get(), 0, false, prime);
    if (W)
      na[ns++] = narrowArgs(FDM_SET_WALL, pw_, nullptr, walls_.grid(),
                            walls_.table().hashValue.get(), 0, false, prime);
    fdm_ghosts gv[3];
    const fdm_pairs* gpairs[3] = {nullptr, nullptr, nullptr};
    const fdm_mirror* gmir[3] = {nullptr, nullptr, nullptr};
    fdm_pairs gcur[3];
    fdm_mirror gm[3];
    for (int s = 0; s < 3; ++s) {
      if (!periodic_.active(s)) continue;
      gv[s] = periodic_.view(s);
      GhostSet& G = periodic_.set(s);
      na[ns++] = narrowArgs(FDM_SET_GHOST, G.interaction, &gv[s], particles_.grid(),
                            G.table.hashValue.get(), s + 1, !d.ghost, prime);
      gcur[s] = G.interaction.current();
      gm[s] = G.interaction.mirror();
      gpairs[s] = &gcur[s];
      gmir[s] = &gm[s];
    }
    na[0].gate = gate;
    {
      Timer t(this, "narrow");
      fdm_narrow_workspace w;
      w.workList = lawHead_.get();
      w.workCapacity = (int)workHit_.size();
      w.workCount = nullptr;
      w.workHit = workHit_.get();
      w.hitGeom = hitGeom_.get();
      w.hitCapacity = hitCapacity_;
      w.hitCount = hitCount_.get();
      w.tuning = &tuning_;
      fdmCheck(fdm_launchNarrowPhaseSets(na, ns, &w, stream_), "narrow phase");
    }
    if (periodic_.isActived() && !gate) {
      // the wrap of this step's positions, after its detection (a step that
      // keeps every list, the gated one, has no wrap)
      fdmCheck(fdm_launchPeriodicCommitWrap(P, particles_.grid(), periodic_.flags(),
                                            inst_.count > 0 ? &inst_ : nullptr,
                                            ctl_.get() + FDM_CTL_REBUILD_GHOST, stream_),
               "periodic wrap");
    }
    if (prime) {
      // no forces in the pre-loop build: drop the mirror entries
      fdm_mirror m = pp_.mirror();
      fdmCheck(fdm_fillInt(m.count, std::max<size_t>(N, 1), 0, stream_), "clear");
      for (int s = 0; s < 3; ++s)
        if (gmir[s]) fdmCheck(fdm_fillInt(gmir[s]->count, std::max<size_t>(N, 1), 0, stream_), "clear");
      return;
    }
    fdm_bonds bv = bonds_.view();
    bv.gate = gate;
    if (bv.n > 0) {
      Timer t(this, "bonds");
      fdmCheck(fdm_launchBondForces(P, bv, stream_), "bonds");
    }
    fdm_pairs ppc = pp_.current(), pwc = pw_.current();
    fdm_mirror ppm = pp_.mirror();
    fdm_reduce_args r{};
    r.particles = P;
    r.pgeo = pgeo;
    r.pp = &ppc;
    r.ppMirror = &ppm;
    r.pw = W ? &pwc : nullptr;
    for (int s = 0; s < 3; ++s) {
      r.ghost[s] = gpairs[s];
      r.ghostMirror[s] = gmir[s];
    }
    r.bonds = bv.n > 0 ? &bv : nullptr;
    r.gravity = g;
    r.h = 0.5 * dt_;
    r.stepCounter = stepCounter_.get();
    r.gate = gate;
    r.tuning = &tuning_;
    {
      Timer t(this, "reduce_kick");
      if (hookMode_ == HookNone) {
        r.kick = 1;
        fdmCheck(fdm_launchReduceForceTorque(&r, stream_), "reduce");
      } else {
        r.kick = 0;
        fdmCheck(fdm_launchReduceForceTorque(&r, stream_), "reduce");
        baseHookCalled_ = false;
        addExternalForceTorque(time);
        if (hookMode_ == HookUnknown) hookMode_ = baseHookCalled_ ? HookNone : HookUser;
        fdmCheck(fdm_launchParticleVelocityAngularVelocityIntegration(P, pgeo, g, 0.5 * dt_, stream_),
                 "kick");
      }
    }
  }

  void phaseB(bool prime, double time, Decision d) {
    beginSets(d);
    if (d.rebuild || d.ghost) phaseBRebuild(d);
    phaseBCommon(prime, time, d, nullptr);
  }

  struct GenState {
    SolidInteraction::Generations g[5];
  };
  GenState saveGenerations() {
    GenState s;
    s.g[0] = pp_.generations();
    s.g[1] = pw_.generations();
    for (int k = 0; k < 3; ++k) s.g[2 + k] = periodic_.set(k).interaction.generations();
    return s;
  }
  void restoreGenerations(const GenState& s) {
    pp_.setGenerations(s.g[0]);
    pw_.setGenerations(s.g[1]);
    for (int k = 0; k < 3; ++k) periodic_.set(k).interaction.setGenerations(s.g[2 + k]);
  }

  void stepOnce(bool prime, double time) {
    phaseA(prime);
    // sector images are rebuilt every step: no step keeps its lists
    if (prime || hookMode_ != HookNone || periodic_.isSector()) {
      phaseB(prime, time, waitDecision());
      return;
    }
    // Queue the step that keeps every list before the decision is known
    // (its kernels test the keep word), so the device does not wait for the
    // host; when the decision asks for a rebuild, queue the full step after.
    const GenState before = saveGenerations();
    const Decision keep = {false, false};
    beginSets(keep);
    phaseBCommon(false, time, keep, ctl_.get() + FDM_CTL_KEEP);
    const Decision d = waitDecision();
    if (d.rebuild || d.ghost) {
      restoreGenerations(before);
      phaseB(false, time, d);
    }
  }

  int ghostListIndex() {
    for (int s = 0; s < 3; ++s)
      if (periodic_.active(s)) return periodic_.set(s).interaction.listIndex();
    return 0;
  }

  // ---------------------------------------------------------- lookahead
  // advance() queues each step before the decisions of the steps ahead of
  // it are known: phase A runs only when the control word FDM_CTL_KEEP is
  // set, i.e. when the previous step kept its lists (its own decision sets
  // the word) or the host completed it; the keep variant of phase B runs
  // under the same word, now set by this step's decision. A step whose
  // decision asks for a rebuild clears the word, so the device skips its
  // keep variant and every step queued after it. When the host reads that
  // decision it rolls its buffer generations back to the step, queues the
  // rebuild variant, sets the word and queues the skipped steps again. So
  // the device never waits for the host, and the host waits only when it is
  // `lookahead_` steps ahead.
  struct Queued {
    int number;          // decision number of the step
    GenState before;     // generations before the step's phase B
    double time;
  };
  // Queues `count` steps (1, or kGraphSteps from a captured graph of that
  // many steps when graphs are on).
  void queueSteps(int count, double time, std::vector<Queued>& q) {
    const int* keep = ctl_.get() + FDM_CTL_KEEP;
    const Decision kd = {false, false};
    GenState before[kGraphSteps];
    if (useGraph_) {
      const int stepNext = pp_.nextStepIndex();
      hipGraphExec_t& ge = (count == 1 ? graphsQueued1_ : graphsQueuedK_)[stepNext][pp_.listIndex()]
                                                                        [ghostListIndex()];
      auto body = [&] {
        for (int j = 0; j < count; ++j) {
          phaseA(false, keep);
          before[j] = saveGenerations();
          beginSets(kd);
          phaseBCommon(false, 0.0, kd, keep);
        }
      };
      if (!ge) {
        ge = capture(body);
      } else {
        for (int j = 0; j < count; ++j) {
          before[j] = saveGenerations();
          beginSets(kd);
        }
      }
      hipCheck(hipGraphLaunch(ge, stream_), "graph launch");
    } else {
      for (int j = 0; j < count; ++j) {
        phaseA(false, keep);
        before[j] = saveGenerations();
        beginSets(kd);
        phaseBCommon(false, time + dt_ * double(j), kd, keep);
      }
    }
    for (int j = 0; j < count; ++j) {
      Queued e;
      e.time = time + dt_ * double(j);
      e.before = before[j];
      e.number = ++decisionSeq_;
      q.push_back(e);
    }
  }
  // Completes the oldest queued step. Returns false when it needed a
  // rebuild: then the steps queued after it were skipped on the device and
  // are dropped from the queue.
  bool resolveOldest(std::vector<Queued>& q) {
    const Queued e = q.front();
    q.erase(q.begin());
    const Decision d = waitDecisionFor(e.number);
    if (!(d.rebuild || d.ghost)) return true;
    restoreGenerations(e.before);
    if (useGraph_) {
      const int stepNext = pp_.nextStepIndex();
      const int mainNext = pp_.listIndex() ^ (d.rebuild ? 1 : 0);
      const int ghostNext = ghostListIndex() ^ (d.ghost ? 1 : 0);
      hipGraphExec_t& gb =
          graphsB_[stepNext][mainNext][ghostNext][d.rebuild ? 1 : 0][d.ghost ? 1 : 0];
      if (!gb) gb = capture([&] { phaseB(false, 0.0, d); });
      else beginSets(d);
      hipCheck(hipGraphLaunch(gb, stream_), "graph launch");
    } else {
      phaseB(false, e.time, d);
    }
    fdmCheck(fdm_fillInt(ctl_.get() + FDM_CTL_KEEP, 1, 1, stream_), "keep word");
    decisionSeq_ = e.number;
    q.clear();
    return false;
  }
  void advanceAhead(size_t nSteps, double time) {
    // the step before is complete: let the first queued phase A run
    fdmCheck(fdm_fillInt(ctl_.get() + FDM_CTL_KEEP, 1, 1, stream_), "keep word");
    std::vector<Queued> q;
    size_t next = 0;       // next step to queue
    size_t completed = 0;  // steps whose decision has been read
    while (completed < nSteps) {
      // with graphs, kGraphSteps steps per launch where they fit
      if (useGraph_ && next + kGraphSteps <= nSteps && (int)q.size() + kGraphSteps <= lookahead_) {
        queueSteps(kGraphSteps, time + dt_ * double(next), q);
        next += kGraphSteps;
        continue;
      }
      if (next < nSteps && (int)q.size() < lookahead_) {
        queueSteps(1, time + dt_ * double(next), q);
        ++next;
        continue;
      }
      const size_t first = completed;
      if (!resolveOldest(q)) next = first + 1;  // queue the skipped steps again
      completed = first + 1;
    }
  }

  // One step from captured graphs: phase A (keyed by the result generation
  // it clears) and the phase B variant for this step's decision and buffer
  // generations, each captured at first use.
  void graphStep() {
    const int stepNext = pp_.nextStepIndex();
    hipGraphExec_t& ga = graphsA_[stepNext];
    if (!ga) ga = capture([&] { phaseA(false); });
    hipCheck(hipGraphLaunch(ga, stream_), "graph launch");
    const GenState before = saveGenerations();
    const Decision keep = {false, false};
    if (!periodic_.isSector()) {  // sector images are rebuilt every step
      hipGraphExec_t& gs = graphsKeep_[stepNext][pp_.listIndex()][ghostListIndex()];
      if (!gs) {
        gs = capture([&] {
          beginSets(keep);
          phaseBCommon(false, 0.0, keep, ctl_.get() + FDM_CTL_KEEP);
        });
      } else {
        beginSets(keep);
      }
      hipCheck(hipGraphLaunch(gs, stream_), "graph launch");
    }
    const Decision d = waitDecision();
    if (!(d.rebuild || d.ghost)) return;
    restoreGenerations(before);
    const int mainNext = pp_.listIndex() ^ (d.rebuild ? 1 : 0);
    const int ghostNext = ghostListIndex() ^ (d.ghost ? 1 : 0);
    hipGraphExec_t& gb = graphsB_[stepNext][mainNext][ghostNext][d.rebuild ? 1 : 0][d.ghost ? 1 : 0];
    if (!gb) gb = capture([&] { phaseB(false, 0.0, d); });  // advances the generations
    else beginSets(d);
    hipCheck(hipGraphLaunch(gb, stream_), "graph launch");
  }
  template <class F>
  hipGraphExec_t capture(F&& body) {
    hipGraph_t graph;
    hipGraphExec_t exec = nullptr;
    hipCheck(hipStreamBeginCapture(stream_, hipStreamCaptureModeGlobal), "capture");
    body();
    hipCheck(hipStreamEndCapture(stream_, &graph), "capture end");
    hipCheck(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate");
    (void)hipGraphDestroy(graph);
    return exec;
  }
  void destroyGraphs() {
    for (auto& g : graphsA_) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
    for (auto* set : {&graphsQueued1_[0][0][0], &graphsQueuedK_[0][0][0]})
      for (int k = 0; k < 8; ++k) {
        if (set[k]) (void)hipGraphExecDestroy(set[k]);
        set[k] = nullptr;
      }
    hipGraphExec_t* b = &graphsB_[0][0][0][0][0];
    for (size_t k = 0; k < sizeof(graphsB_) / sizeof(graphsB_[0][0][0][0][0]); ++k) {
      if (b[k]) (void)hipGraphExecDestroy(b[k]);
      b[k] = nullptr;
    }
    hipGraphExec_t* c = &graphsKeep_[0][0][0];
    for (size_t k = 0; k < sizeof(graphsKeep_) / sizeof(graphsKeep_[0][0][0]); ++k) {
      if (c[k]) (void)hipGraphExecDestroy(c[k]);
      c[k] = nullptr;
    }
  }

  std::string dir_;
  size_t phase_ = 0;
  bool writeVTU_ = true;
  hipStream_t stream_;
  hipStream_t ownStream_ = nullptr;
  size_t maxGPUThread_;
  int device_;
  LSParticle particles_, walls_;
  SolidInteraction pp_, pw_;
  VBondedInteraction bonds_;
  PeriodicBoundaryXY2D periodic_;
  double3 gravity_ = make_double3(0., 0., 0.);
  double dt_ = 0.0;
  double cellSize_ = 0.0;
  int pairsPerParticle_ = 16, recordsPerParticle_ = 48, mirrorCapacity_ = 48;
  DeviceBuffer<double> workspace_;
  DeviceBuffer<unsigned> status_;
  DeviceBuffer<int> stepCounter_;
  DeviceBuffer<int> ctl_;
  std::vector<SystemDesc> systems_;
  std::vector<fdm_grid> sysGrid_;  // particle grid of each system (host copy)
  fdm_instances inst_{};
  DeviceBuffer<int> instParticle_, instWall_, instWallStart_;
  DeviceBuffer<fdm_grid> instPGrid_, instWGrid_;
  DeviceBuffer<long long> instPOff_, instWOff_;
  DeviceBuffer<fdm_periodic> instPeriodic_;
  DeviceBuffer<int> workHit_, hitCount_, lawHead_;
  DeviceBuffer<double> hitGeom_;
  int hitCapacity_ = 0;
  fdm_tuning tuning_{};
  DeviceBuffer<double> disp_;
  double skinRequest_ = -1.0, skin_ = 0.0;
  double3 domainMin_{}, domainMax_{};
  double maxRadius_ = 0.0, maxBoxReach_ = 0.0, rhoMax_ = 0.0;
  bool useGraph_ = false, timing_ = false;
  bool baseHookCalled_ = false;
  HookMode hookMode_ = HookUnknown;
  hipGraphExec_t graphsA_[2] = {};
  // queued steps as graphs: one step, or kGraphSteps steps (an even number,
  // so a graph starts and ends at the same generation parity), keyed by the
  // result generation, list generation and ghost list generation
  static constexpr int kGraphSteps = 4;
  hipGraphExec_t graphsQueued1_[2][2][2] = {};
  hipGraphExec_t graphsQueuedK_[2][2][2] = {};
  int lookahead_ = 2 * kGraphSteps;
  hipGraphExec_t graphsB_[2][2][2][2][2] = {};
  hipGraphExec_t graphsKeep_[2][2][2] = {};
  int* hostDecision_ = nullptr;     // pinned, host view
  int* hostDecisionDev_ = nullptr;  // the same memory, device view
  int decisionSeq_ = 0;
  std::map<std::string, double> stageMs_;
  std::map<std::string, long> stageCalls_;
  std::vector<PendingTimer> pending_;
};

}  // namespace fundem

using fundem::LSSolver;
