// SolidInteraction: one interaction set (particle-particle, particle-wall or
// one ghost set). The reference stores per-node counts and prefix sums plus
// contact arrays, copied to "previous" arrays every step
// (kernel/SolidInteraction.h:92-151). Here the set holds
//  - two generations of candidate lists (CSR counts and pairs), of which one
//    is current; a step that rebuilds the lists writes the other one and
//    makes it current, while the kept one still resolves the history;
//  - two generations of per-step results (history records, per-pair record
//    ranges and force/torque sums); each step writes one and reads the other.
#pragma once
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../../include/fundem_amd.h"
#include "DeviceBuffer.h"
#include "types.h"

namespace fundem {

struct ListBuffers {
  int capacity = 0;
  DeviceBuffer<int> count, total, pairA, pairT, pairG;
  void allocate(size_t nBodies, int cap) {
    capacity = std::max(cap, 1);
    count.allocate(nBodies + 1);
    total.allocate(1);
    pairA.allocate(capacity);
    pairT.allocate(capacity);
    pairG.allocate(capacity);
  }
  double memoryGB() const { return count.memoryGB() + pairA.memoryGB() * 3; }
};

struct StepBuffers {
  int recCapacity = 0;
  DeviceBuffer<int> recOff, recCnt, recTotal;
  DeviceBuffer<double> sumF, sumTA, sumTB;
  DeviceBuffer<fdm_record> rec;
  void allocate(int pairCap, int recCap) {
    const int cap = std::max(pairCap, 1);
    recCapacity = std::max(recCap, 1);
    recOff.allocate(cap);
    recCnt.allocate(cap);
    recTotal.allocate(FDM_REC_COUNTERS);
    sumF.allocate(3 * (size_t)cap);
    sumTA.allocate(3 * (size_t)cap);
    sumTB.allocate(3 * (size_t)cap);
    rec.allocate(recCapacity);
  }
  double memoryGB() const { return recOff.memoryGB() * 2 + sumF.memoryGB() * 3 + rec.memoryGB(); }
};

inline fdm_pairs pairsView(ListBuffers& l, StepBuffers& s) {
  fdm_pairs p;
  p.capacity = l.capacity;
  p.count = l.count.get();
  p.total = l.total.get();
  p.pairA = l.pairA.get();
  p.pairT = l.pairT.get();
  p.pairG = l.pairG.get();
  p.recOff = s.recOff.get();
  p.recCnt = s.recCnt.get();
  p.sumF = s.sumF.get();
  p.sumTA = s.sumTA.get();
  p.sumTB = s.sumTB.get();
  p.recCapacity = s.recCapacity;
  p.recTotal = s.recTotal.get();
  p.rec = s.rec.get();
  return p;
}

// One contact of a set, in canonical order (A, target, node).
struct ContactRecord {
  int A, target, node;
  double3 spring;
};

struct SolidInteraction {
 public:
  SolidInteraction() = default;
  SolidInteraction(const SolidInteraction&) = delete;
  SolidInteraction& operator=(const SolidInteraction&) = delete;

  // Buffers for nBodies masters; capacities in pairs and records. History
  // survives a re-initialisation with the same sizes (as the reference's
  // SolidInteraction::initialize keeps its arrays, SolidInteraction.h:23-49).
  void initialize(size_t nBodies, int pairCapacity, int recordCapacity, int mirrorCapacity) {
    if (nBodies == nBodies_ && pairCapacity == list_[0].capacity &&
        recordCapacity == step_[0].recCapacity && mirrorCapacity == mirrorCap_)
      return;
    nBodies_ = nBodies;
    for (auto& l : list_) l.allocate(nBodies, pairCapacity);
    for (auto& s : step_) s.allocate(pairCapacity, recordCapacity);
    mirrorCap_ = std::max(mirrorCapacity, 1);
    mirrorCount_.allocate(std::max<size_t>(nBodies, 1));
    mirrorList_.allocate(std::max<size_t>(nBodies, 1) * mirrorCap_);
    list_cur_ = list_prev_ = step_cur_ = 0;
  }

  // Generation roles back to their initial state (every set of a solver
  // then flips in step).
  void resetGenerations() { list_cur_ = list_prev_ = step_cur_ = 0; }
  struct Generations {
    int listCur, listPrev, stepCur;
  };
  Generations generations() const { return {list_cur_, list_prev_, step_cur_}; }
  void setGenerations(Generations g) {
    list_cur_ = g.listCur;
    list_prev_ = g.listPrev;
    step_cur_ = g.stepCur;
  }

  // Start of a step: a new results generation, and a new list generation
  // when the lists are rebuilt.
  void beginStep(bool rebuild) {
    list_prev_ = list_cur_;
    if (rebuild) list_cur_ ^= 1;
    step_cur_ ^= 1;
  }
  // Generations the step after beginStep(rebuild) would use, without
  // changing anything (keys of captured graphs).
  int nextStepIndex() const { return step_cur_ ^ 1; }
  int listIndex() const { return list_cur_; }

  fdm_pairs current() { return pairsView(list_[list_cur_], step_[step_cur_]); }
  fdm_pairs previous() { return pairsView(list_[list_prev_], step_[step_cur_ ^ 1]); }
  // record counter of the results generation the next step writes
  int* nextRecordCounter() { return step_[step_cur_ ^ 1].recTotal.get(); }
  fdm_mirror mirror() {
    fdm_mirror m;
    m.capacity = mirrorCap_;
    m.count = mirrorCount_.get();
    m.list = mirrorList_.get();
    return m;
  }

  // Candidate pairs of the latest step; synchronises.
  int lastTotal(hipStream_t s) {
    ListBuffers& l = list_[list_cur_];
    if (!l.total.get()) return 0;
    return std::min(l.total.toHost(s, 1)[0], l.capacity);
  }

  // Contacts of the most recent step, sorted by (A, target, node).
  // Synchronises the stream.
  std::vector<ContactRecord> contacts(hipStream_t s) {
    ListBuffers& l = list_[list_cur_];
    StepBuffers& g = step_[step_cur_];
    std::vector<ContactRecord> out;
    if (!l.total.get()) return out;
    std::vector<int> total = l.total.toHost(s, 1);
    const int np = std::min(total[0], l.capacity);
    if (np <= 0) return out;
    std::vector<int> A = l.pairA.toHost(s, np), T = l.pairT.toHost(s, np);
    std::vector<int> off = g.recOff.toHost(s, np), cnt = g.recCnt.toHost(s, np);
    int nr = 0;  // records are spread over pools: read up to the last one used
    for (int p = 0; p < np; ++p)
      if (cnt[p] > 0) nr = std::max(nr, off[p] + cnt[p]);
    nr = std::min(nr, g.recCapacity);
    std::vector<fdm_record> rec = g.rec.toHost(s, std::max(nr, 0));
    for (int p = 0; p < np; ++p)
      for (int r = off[p]; r < off[p] + cnt[p] && r < nr; ++r) {
        const fdm_record& v = rec[r];
        out.push_back({A[p], T[p], v.node, make_double3(v.x, v.y, v.z)});
      }
    return out;
  }
  double deviceMemoryGB() const {
    return list_[0].memoryGB() + list_[1].memoryGB() + step_[0].memoryGB() + step_[1].memoryGB();
  }

  // ------------------------------------------------------------ getters
  // The reference's getter set (SolidInteraction.h:323-372) on this
  // layout. Device view of the latest step: the candidate pairs of the
  // current list (master body, target body) and, per pair, the first
  // history record and the number of its node contacts; a record is a
  // fdm_record (sliding spring xyz, node id), a pair's records are
  // contiguous in node order. The previous step's view resolves history.
  int* masterID() { return list_[list_cur_].pairA.get(); }
  int* slaveID() { return list_[list_cur_].pairT.get(); }
  int* contactStart() { return step_[step_cur_].recOff.get(); }
  int* contactCount() { return step_[step_cur_].recCnt.get(); }
  fdm_record* slidingSpring() { return step_[step_cur_].rec.get(); }
  int* previousMasterID() { return list_[list_prev_].pairA.get(); }
  int* previousSlaveID() { return list_[list_prev_].pairT.get(); }
  fdm_record* previousSlidingSpring() { return step_[step_cur_ ^ 1].rec.get(); }
  // Candidate pairs of the current list (counts stay on the device between
  // steps; this reads the count back and synchronises).
  size_t numCandidatePair_device(hipStream_t s) { return (size_t)std::max(lastTotal(s), 0); }

  // Node contacts of the step last downloaded (LSSolver::download, or an
  // output frame), the length of the host copies below; launch shape over
  // them and over the master bodies by the rules of SolidInteraction.h:
  // 45-48 and 108-118.
  size_t numPair_device() const { return masterIDHost_.size(); }
  const size_t& pairGridDim() const { return pairGridDim_; }
  const size_t& pairBlockDim() const { return pairBlockDim_; }
  size_t numMaster() const { return nBodies_ == (size_t)-1 ? 0 : nBodies_; }
  size_t numMaster_device() const { return numMaster(); }
  const size_t& masterGridDim() const { return masterGridDim_; }
  const size_t& masterBlockDim() const { return masterBlockDim_; }

  // Host copies of the node contacts of the step last downloaded, in
  // canonical order (master body, target, node): contact point and normal,
  // the master's boundary node as a global index (nodes numbered particle by
  // particle, as boundaryNodeInfo numbers them) and the target body.
  const std::vector<double3> contactPointHostRef() const { return pointHost_; }
  const std::vector<double3> contactNormalHostRef() const { return normalHost_; }
  const std::vector<int> masterIDHostRef() const { return masterIDHost_; }
  const std::vector<int> slaveIDHostRef() const { return slaveIDHost_; }

  void setHostContacts(std::vector<double3> point, std::vector<double3> normal,
                       std::vector<int> masterNode, std::vector<int> slave, size_t maxGPUThread) {
    pointHost_ = std::move(point);
    normalHost_ = std::move(normal);
    masterIDHost_ = std::move(masterNode);
    slaveIDHost_ = std::move(slave);
    const size_t np = masterIDHost_.size();
    if (np > 0) {
      pairBlockDim_ = np;
      if (maxGPUThread > 0 && maxGPUThread < np) pairBlockDim_ = maxGPUThread;
      pairGridDim_ = (np + pairBlockDim_ - 1) / pairBlockDim_;
    } else {
      pairGridDim_ = 1;
      pairBlockDim_ = 1;
    }
  }
  void setLaunchShape(size_t maxGPUThread) {
    const size_t n = numMaster();
    if (n == 0) return;
    if (maxGPUThread > 0) masterBlockDim_ = maxGPUThread;
    if (n < maxGPUThread) masterBlockDim_ = n;
    masterGridDim_ = (n + masterBlockDim_ - 1) / masterBlockDim_;
  }

 private:
  size_t nBodies_ = (size_t)-1;
  ListBuffers list_[2];
  StepBuffers step_[2];
  int list_cur_ = 0, list_prev_ = 0, step_cur_ = 0;
  int mirrorCap_ = 0;
  DeviceBuffer<int> mirrorCount_, mirrorList_;
  std::vector<double3> pointHost_, normalHost_;
  std::vector<int> masterIDHost_, slaveIDHost_;
  size_t pairGridDim_ = 1, pairBlockDim_ = 1, masterGridDim_ = 1, masterBlockDim_ = 1;
};

}  // namespace fundem
