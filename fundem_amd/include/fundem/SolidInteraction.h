// SolidInteraction: one interaction set (particle-particle, particle-wall or
// one ghost set). The reference stores per-node counts and prefix sums plus
// contact arrays, copied to "previous" arrays every step
// (kernel/SolidInteraction.h:92-151). Here the set holds two generations of
// candidate pairs with their tangential-history records; a step writes one
// generation while reading the other, and the roles swap afterwards.
#pragma once
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../../include/fundem_amd.h"
#include "DeviceBuffer.h"
#include "types.h"

namespace fundem {

struct PairBuffers {
  int capacity = 0, recCapacity = 0;
  DeviceBuffer<int> count, total, pairA, pairT, pairG, recOff, recCnt, recTotal;
  DeviceBuffer<double> sumF, sumTA, sumTB;
  DeviceBuffer<float4> rec;
  void allocate(size_t nBodies, int cap, int recCap) {
    capacity = std::max(cap, 1);
    recCapacity = std::max(recCap, 1);
    count.allocate(nBodies + 1);
    total.allocate(1);
    recTotal.allocate(1);
    pairA.allocate(capacity);
    pairT.allocate(capacity);
    pairG.allocate(capacity);
    recOff.allocate(capacity);
    recCnt.allocate(capacity);
    sumF.allocate(3 * (size_t)capacity);
    sumTA.allocate(3 * (size_t)capacity);
    sumTB.allocate(3 * (size_t)capacity);
    rec.allocate(recCapacity);
  }
  fdm_pairs view() {
    fdm_pairs p;
    p.capacity = capacity;
    p.count = count.get();
    p.total = total.get();
    p.pairA = pairA.get();
    p.pairT = pairT.get();
    p.pairG = pairG.get();
    p.recOff = recOff.get();
    p.recCnt = recCnt.get();
    p.sumF = sumF.get();
    p.sumTA = sumTA.get();
    p.sumTB = sumTB.get();
    p.recCapacity = recCapacity;
    p.recTotal = recTotal.get();
    p.rec = rec.get();
    return p;
  }
  double memoryGB() const {
    return count.memoryGB() + pairA.memoryGB() * 5 + sumF.memoryGB() * 3 + rec.memoryGB();
  }
};

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
    if (nBodies == nBodies_ && pairCapacity == gen_[0].capacity &&
        recordCapacity == gen_[0].recCapacity && mirrorCapacity == mirrorCap_)
      return;
    nBodies_ = nBodies;
    for (auto& g : gen_) g.allocate(nBodies, pairCapacity, recordCapacity);
    mirrorCap_ = std::max(mirrorCapacity, 1);
    mirrorCount_.allocate(std::max<size_t>(nBodies, 1));
    mirrorList_.allocate(std::max<size_t>(nBodies, 1) * mirrorCap_);
    cur_ = 0;
  }
  fdm_pairs current() { return gen_[cur_].view(); }
  fdm_pairs previous() { return gen_[cur_ ^ 1].view(); }
  fdm_mirror mirror() {
    fdm_mirror m;
    m.capacity = mirrorCap_;
    m.count = mirrorCount_.get();
    m.list = mirrorList_.get();
    return m;
  }
  void swap() { cur_ ^= 1; }
  int currentIndex() const { return cur_; }
  void setCurrentIndex(int c) { cur_ = c & 1; }

  // Contacts of the most recent step (the generation written last), sorted
  // by (A, target, node). Synchronises the stream.
  std::vector<ContactRecord> contacts(hipStream_t s) {
    PairBuffers& g = gen_[cur_ ^ 1];
    std::vector<int> total = g.total.toHost(s, 1);
    const int np = std::min(total[0], g.capacity);
    std::vector<ContactRecord> out;
    if (np <= 0) return out;
    std::vector<int> A = g.pairA.toHost(s, np), T = g.pairT.toHost(s, np);
    std::vector<int> off = g.recOff.toHost(s, np), cnt = g.recCnt.toHost(s, np);
    std::vector<int> rt = g.recTotal.toHost(s, 1);
    const int nr = std::min(rt[0], g.recCapacity);
    std::vector<float4> rec = g.rec.toHost(s, std::max(nr, 0));
    for (int p = 0; p < np; ++p)
      for (int r = off[p]; r < off[p] + cnt[p] && r < nr; ++r) {
        float4 v = rec[r];
        int k;
        std::memcpy(&k, &v.w, sizeof(int));
        out.push_back({A[p], T[p], k, make_double3(v.x, v.y, v.z)});
      }
    return out;
  }
  // Candidate pairs of the generation written last; synchronises.
  int lastTotal(hipStream_t s) {
    PairBuffers& g = gen_[cur_ ^ 1];
    if (!g.total.get()) return 0;
    return std::min(g.total.toHost(s, 1)[0], g.capacity);
  }
  double deviceMemoryGB() const { return gen_[0].memoryGB() + gen_[1].memoryGB(); }

 private:
  size_t nBodies_ = (size_t)-1;
  PairBuffers gen_[2];
  int cur_ = 0;
  int mirrorCap_ = 0;
  DeviceBuffer<int> mirrorCount_, mirrorList_;
};

}  // namespace fundem
