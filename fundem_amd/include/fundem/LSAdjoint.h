// LSAdjoint: reverse-mode gradients of a scalar outcome of a solver's run
// (no counterpart in the reference). Function-style use:
//
//   LSAdjoint adj(solver);                 // solver initialised
//   adj.forward(steps, checkpointEvery);   // advances the solver, keeps checkpoints
//   AdjointGradients g = adj.backward(dJ/dx, dJ/dq, dJ/dv, dJ/dw of the final state);
//
// backward() restores each checkpoint in turn from the last, re-runs its
// segment to record the states and contacts of every step (a contact tape),
// and runs the hand-written reverse step (fdm_launchAdjointStep) over it. The
// results are the adjoints of the initial positions, orientations,
// velocities and angular velocities and of each particle material's kn, ks,
// mu and damping ratio, and of each wall material's mu and damping ratio.
// The derivative is that of the branch each contact took (open or closed,
// stick or slip); FP32 storage of velocities and springs is differentiated
// as the identity.
//
// Supported: particle-particle, particle-wall (walls with prescribed motion)
// and periodic ghost contacts, several independent systems in one solver.
// Not supported: bonds, external-force hooks, the sector periodic boundary.
#pragma once
#include <algorithm>
#include <array>
#include <stdexcept>
#include <vector>

#include "LSSolver.h"

namespace fundem {

struct AdjointGradients {
  std::vector<double3> position, velocity, angularVelocity;  // initial state
  std::vector<quaternion> orientation;
  std::vector<double> kn, ks, mu, dampingRatio;              // per particle material
  std::vector<double> wallMu, wallDampingRatio;              // per wall material
};

class LSAdjoint {
 public:
  explicit LSAdjoint(LSSolver& solver) : s_(solver) {}
  LSAdjoint(const LSAdjoint&) = delete;
  LSAdjoint& operator=(const LSAdjoint&) = delete;

  // Node-contacts per particle per step that one tape block can hold.
  void setContactsPerParticle(int c) { contactsPerParticle_ = std::max(c, 1); }

  // Advance the solver `steps` steps from its current state, keeping a
  // checkpoint every `checkpointEvery` steps.
  void forward(size_t steps, size_t checkpointEvery) {
    if (steps < 1) throw std::invalid_argument("forward needs at least one step");
    if (s_.particles_.num_device() == 0) throw std::logic_error("initialise the solver first");
    if (s_.bonds_.view().n > 0) throw std::invalid_argument("bonds are not differentiated");
    if (s_.periodic_.isSector())
      throw std::invalid_argument("the sector periodic boundary is not differentiated");
    steps_ = steps;
    every_ = std::max<size_t>(1, std::min(checkpointEvery, steps));
    setup();
    checkpoints_.clear();
    checkpoints_.resize((steps_ + every_ - 1) / every_);
    save(checkpoints_[0]);
    for (size_t n = 1; n <= steps_; ++n) {
      s_.stepOnce(false, 0.0);
      if (n % every_ == 0 && n < steps_) save(checkpoints_[n / every_]);
    }
    save(final_);
    if (s_.hookMode_ == LSSolver::HookUser)
      throw std::invalid_argument("external-force hooks are not differentiated");
    s_.checkStatus();
  }

  AdjointGradients backward(const std::vector<double3>& xBar, const std::vector<quaternion>& qBar,
                            const std::vector<double3>& vBar, const std::vector<double3>& wBar) {
    if (checkpoints_.empty()) throw std::logic_error("run forward first");
    const size_t N = N_, W = W_;
    if (xBar.size() != N || qBar.size() != N || vBar.size() != N || wBar.size() != N)
      throw std::invalid_argument("one adjoint per particle");
    hipStream_t st = s_.stream_;
    {
      std::vector<double> h(4 * N, 0.0);
      auto up3 = [&](DeviceBuffer<double>& d, const std::vector<double3>& v) {
        for (size_t i = 0; i < N; ++i) {
          h[i] = v[i].x;
          h[N + i] = v[i].y;
          h[2 * N + i] = v[i].z;
        }
        d.upload(h.data(), 3 * N, st);
        hipCheck(hipStreamSynchronize(st), "upload");
      };
      up3(xBar_, xBar);
      up3(vBar_, vBar);
      up3(wBar_, wBar);
      for (size_t i = 0; i < N; ++i) {
        h[i] = qBar[i].q0;
        h[N + i] = qBar[i].q1;
        h[2 * N + i] = qBar[i].q2;
        h[3 * N + i] = qBar[i].q3;
      }
      qBar_.upload(h.data(), 4 * N, st);
      hipCheck(hipStreamSynchronize(st), "upload");
    }
    zero(FBar_, st);
    zero(TBar_, st);
    zero(paramBar_, st);
    zero(wallShare_, st);
    hipCheck(hipMemsetAsync(wallMat_.get(), 0xff, wallMat_.size() * sizeof(int), st), "clear");
    std::vector<double> boundary;  // spring adjoints of the first block of the later segment
    for (size_t seg = checkpoints_.size(); seg-- > 0;) {
      const size_t n0 = seg * every_, n1 = std::min(n0 + every_, steps_);
      const int len = (int)(n1 - n0);
      restore(checkpoints_[seg]);
      // record the segment
      copyState(slots_[0], true);
      appendBlock(0, -1);
      for (int j = 1; j <= len; ++j) {
        s_.stepOnce(false, 0.0);
        copyState(slots_[j], true);
        appendBlock(j, j - 1);
      }
      s_.checkStatus();
      checkTape();
      zero(springBar_, st);
      if (!boundary.empty()) {
        const std::vector<int> st0 = tape_.blockStart.toHost(st, len + 1);
        const std::vector<int> cnt = tape_.blockCount.toHost(st, len + 1);
        if ((size_t)cnt[len] * 3 != boundary.size())
          throw std::runtime_error("recomputed segment differs from the forward run");
        hipCheck(hipMemcpyAsync(springBar_.get() + 3 * (size_t)st0[len], boundary.data(),
                                boundary.size() * sizeof(double), hipMemcpyHostToDevice, st),
                 "spring adjoint");
      }
      for (int j = len; j >= 1; --j) reverseStep(j);
      {
        const std::vector<int> st0 = tape_.blockStart.toHost(st, 1);
        const std::vector<int> cnt = tape_.blockCount.toHost(st, 1);
        boundary.assign(3 * (size_t)cnt[0], 0.0);
        if (!boundary.empty())
          hipCheck(hipMemcpyAsync(boundary.data(), springBar_.get() + 3 * (size_t)st0[0],
                                  boundary.size() * sizeof(double), hipMemcpyDeviceToHost, st),
                   "spring adjoint");
        hipCheck(hipStreamSynchronize(st), "spring adjoint");
      }
    }
    restore(final_);
    if (s_.status() & FDM_STATUS_BUCKET_OVERFLOW)
      throw std::runtime_error("more than 64 contacts on one target in a step, or contacts with "
                               "walls of more than two materials on one particle");
    // results
    AdjointGradients g;
    std::vector<double> x = xBar_.toHost(st, 3 * N), v = vBar_.toHost(st, 3 * N),
                        w = wBar_.toHost(st, 3 * N), q = qBar_.toHost(st, 4 * N),
                        p = paramBar_.toHost(st, 4 * N);
    g.position.resize(N);
    g.velocity.resize(N);
    g.angularVelocity.resize(N);
    g.orientation.resize(N);
    for (size_t i = 0; i < N; ++i) {
      g.position[i] = make_double3(x[i], x[N + i], x[2 * N + i]);
      g.velocity[i] = make_double3(v[i], v[N + i], v[2 * N + i]);
      g.angularVelocity[i] = make_double3(w[i], w[N + i], w[2 * N + i]);
      g.orientation[i] = make_quaternion(q[i], q[N + i], q[2 * N + i], q[3 * N + i]);
    }
    const size_t M = s_.particles_.numMaterial();
    g.kn.assign(M, 0.0);
    g.ks.assign(M, 0.0);
    g.mu.assign(M, 0.0);
    g.dampingRatio.assign(M, 0.0);
    const std::vector<int>& mat = s_.particles_.materialIDHostRef();
    for (size_t i = 0; i < N; ++i) {
      const int m = mat[i];
      g.kn[m] += p[4 * i];
      g.ks[m] += p[4 * i + 1];
      g.mu[m] += p[4 * i + 2];
      g.dampingRatio[m] += p[4 * i + 3];
    }
    const size_t WM = s_.walls_.numMaterial();
    g.wallMu.assign(WM, 0.0);
    g.wallDampingRatio.assign(WM, 0.0);
    if (W > 0) {
      const std::vector<double> ws = wallShare_.toHost(st, 4 * N);
      const std::vector<int> wm = wallMat_.toHost(st, 2 * N);
      for (size_t i = 0; i < N; ++i)
        for (int u = 0; u < 2; ++u) {
          const int m = wm[2 * i + u];
          if (m < 0) continue;
          g.wallMu[m] += ws[4 * i + 2 * u];
          g.wallDampingRatio[m] += ws[4 * i + 2 * u + 1];
        }
    }
    return g;
  }

  size_t steps() const { return steps_; }
  size_t checkpoints() const { return checkpoints_.size(); }

  // A point the solver can return to: mark() saves its state and contacts,
  // rollback() restores them (with a forced list rebuild), after which the
  // solver continues exactly as it would have from the mark.
  void mark() {
    if (slots_.empty()) {
      every_ = 1;
      setup();
    }
    save(mark_);
  }
  void rollback() { restore(mark_); }

 private:
  struct StateBuf {
    DeviceBuffer<double> x, wx;
    DeviceBuffer<fdm_real> q, v, w, F, T, wq, wv, ww, wF, wT;
  };
  struct TapeBuf {
    int capacity = 0, maxBlocks = 0;
    DeviceBuffer<int> blockStart, blockPart, blockCount, A, T, node, prev;
    DeviceBuffer<fdm_real> spring;
    void allocate(int cap, int blocks) {
      capacity = std::max(cap, 1);
      maxBlocks = std::max(blocks, 1);
      blockStart.allocate(maxBlocks);
      blockPart.allocate(6 * (size_t)maxBlocks);
      blockCount.allocate(maxBlocks);
      A.allocate(capacity, false);
      T.allocate(capacity, false);
      node.allocate(capacity, false);
      prev.allocate(capacity, false);
      spring.allocate(3 * (size_t)capacity, false);
    }
    fdm_contact_tape view(unsigned* status) {
      fdm_contact_tape t;
      t.capacity = capacity;
      t.maxBlocks = maxBlocks;
      t.blockStart = blockStart.get();
      t.blockPart = blockPart.get();
      t.blockCount = blockCount.get();
      t.A = A.get();
      t.T = T.get();
      t.node = node.get();
      t.spring = spring.get();
      t.prev = prev.get();
      t.status = status;
      return t;
    }
  };
  struct Checkpoint {
    StateBuf state;
    int part[6] = {0, 0, 0, 0, 0, 0};
    int count = 0;
    DeviceBuffer<int> A, T, node;
    DeviceBuffer<fdm_real> spring;
  };

  static void zero(DeviceBuffer<double>& d, hipStream_t s) {
    if (d.size()) hipCheck(hipMemsetAsync(d.get(), 0, d.size() * sizeof(double), s), "clear");
  }
  static void zeroInt(DeviceBuffer<int>& d, hipStream_t s) {
    if (d.size()) hipCheck(hipMemsetAsync(d.get(), 0, d.size() * sizeof(int), s), "clear");
  }
  template <class T>
  static void copy(T* dst, const T* src, size_t n, hipStream_t s) {
    if (n) hipCheck(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, s), "copy");
  }

  void setup() {
    N_ = s_.particles_.num();
    W_ = s_.walls_.num();
    long long recs = 0;
    pairCap_ = 0;
    for (const fdm_pairs& p : currentSets(setBuf_))
      if (p.recTotal) {
        recs += p.recCapacity;
        pairCap_ += p.capacity;
      }
    maxBlock_ = (int)std::min<long long>((long long)contactsPerParticle_ * (long long)std::max<size_t>(N_, 1),
                                         recs);
    maxBlock_ = std::max(maxBlock_, 64);
    const int len = (int)every_;
    slots_.resize(len + 1);
    for (auto& b : slots_) allocState(b);
    const long long cap = (long long)maxBlock_ * (len + 1);
    if (cap >= (1ll << 31)) throw std::invalid_argument("contact tape too large");
    tape_.allocate((int)cap, len + 1);
    staging_.allocate(maxBlock_, 1);
    tapeWs_.allocate(fdm_tapeWorkspaceBytes(maxBlock_, pairCap_), false);
    adjWs_.allocate(fdm_adjointWorkspaceBytes((int)N_, maxBlock_), false);
    xBar_.allocate(3 * N_);
    qBar_.allocate(4 * N_);
    vBar_.allocate(3 * N_);
    wBar_.allocate(3 * N_);
    FBar_.allocate(3 * N_);
    TBar_.allocate(3 * N_);
    paramBar_.allocate(4 * N_);
    wallShare_.allocate(4 * N_);
    wallMat_.allocate(2 * N_);
    springBar_.allocate(3 * (size_t)tape_.capacity);
  }
  void allocState(StateBuf& b) {
    b.x.allocate(3 * N_, false);
    b.q.allocate(4 * N_, false);
    b.v.allocate(3 * N_, false);
    b.w.allocate(3 * N_, false);
    b.F.allocate(3 * N_, false);
    b.T.allocate(3 * N_, false);
    b.wx.allocate(std::max<size_t>(3 * W_, 1), false);
    b.wq.allocate(std::max<size_t>(4 * W_, 1), false);
    b.wv.allocate(std::max<size_t>(3 * W_, 1), false);
    b.ww.allocate(std::max<size_t>(3 * W_, 1), false);
    b.wF.allocate(std::max<size_t>(3 * W_, 1), false);
    b.wT.allocate(std::max<size_t>(3 * W_, 1), false);
  }
  // solver state to a slot (out) or back (!out)
  void copyState(StateBuf& b, bool out) {
    hipStream_t st = s_.stream_;
    const fdm_bodies P = s_.particles_.bodies(), S = slotView(b);
    fdmCheck(out ? fdm_launchCopyBodies(S, P, st) : fdm_launchCopyBodies(P, S, st), "state copy");
    if (W_ > 0) {
      fdm_bodies Wb = s_.walls_.bodies(), Ws = Wb;
      Ws.x = b.wx.get();
      Ws.q = b.wq.get();
      Ws.v = b.wv.get();
      Ws.w = b.ww.get();
      Ws.F = b.wF.get();
      Ws.T = b.wT.get();
      fdmCheck(out ? fdm_launchCopyBodies(Ws, Wb, st) : fdm_launchCopyBodies(Wb, Ws, st), "state copy");
    }
  }
  fdm_bodies slotView(StateBuf& b) {
    fdm_bodies v = s_.particles_.bodies();
    v.x = b.x.get();
    v.q = b.q.get();
    v.v = b.v.get();
    v.w = b.w.get();
    v.F = b.F.get();
    v.T = b.T.get();
    return v;
  }
  // the current generation of every set, in tape order (absent: recTotal NULL)
  std::array<fdm_pairs, 5>& currentSets(std::array<fdm_pairs, 5>& out) {
    for (auto& p : out) p = fdm_pairs{};
    out[0] = s_.pp_.current();
    if (W_ || s_.walls_.num()) out[1] = s_.pw_.current();
    for (int k = 0; k < 3; ++k)
      if (s_.periodic_.active(k)) out[2 + k] = s_.periodic_.set(k).interaction.current();
    return out;
  }
  std::array<const fdm_pairs*, 5> setPointers(std::array<fdm_pairs, 5>& sets) {
    std::array<const fdm_pairs*, 5> ptr;
    currentSets(sets);
    for (int k = 0; k < 5; ++k) ptr[k] = sets[k].recTotal ? &sets[k] : nullptr;
    return ptr;
  }
  void appendTo(TapeBuf& t, int block, int prevBlock) {
    const auto ptr = setPointers(setBuf_);
    fdmCheck(fdm_launchTapeAppend(ptr.data(), t.view(s_.status_.get()), block, prevBlock, maxBlock_,
                                  tapeWs_.get(), tapeWs_.size(), s_.stream_),
             "contact tape");
  }
  void appendBlock(int block, int prevBlock) { appendTo(tape_, block, prevBlock); }
  void checkTape() {
    const unsigned st = s_.status();
    if (st & FDM_STATUS_RECORD_OVERFLOW)
      throw std::runtime_error("contact tape full; raise setContactsPerParticle");
  }
  // state and history of the solver's current step
  void save(Checkpoint& c) {
    hipStream_t st = s_.stream_;
    allocState(c.state);
    copyState(c.state, true);
    appendTo(staging_, 0, -1);
    checkTape();
    const std::vector<int> part = staging_.blockPart.toHost(st, 6);
    for (int k = 0; k < 6; ++k) c.part[k] = part[k];
    c.count = staging_.blockCount.toHost(st, 1)[0];
    const size_t m = std::max(c.count, 1);
    c.A.allocate(m, false);
    c.T.allocate(m, false);
    c.node.allocate(m, false);
    c.spring.allocate(3 * m, false);
    copy(c.A.get(), staging_.A.get(), c.count, st);
    copy(c.T.get(), staging_.T.get(), c.count, st);
    copy(c.node.get(), staging_.node.get(), c.count, st);
    copy(c.spring.get(), staging_.spring.get(), 3 * (size_t)c.count, st);
    hipCheck(hipStreamSynchronize(st), "checkpoint");
  }
  // the solver continues from a checkpoint: its state, its contacts as the
  // history of the next step, and a forced list rebuild
  void restore(Checkpoint& c) {
    hipStream_t st = s_.stream_;
    copyState(c.state, false);
    copy(staging_.A.get(), c.A.get(), c.count, st);
    copy(staging_.T.get(), c.T.get(), c.count, st);
    copy(staging_.node.get(), c.node.get(), c.count, st);
    copy(staging_.spring.get(), c.spring.get(), 3 * (size_t)c.count, st);
    const int zero = 0;
    hipCheck(hipMemcpyAsync(staging_.blockStart.get(), &zero, sizeof(int), hipMemcpyHostToDevice, st), "restore");
    hipCheck(hipMemcpyAsync(staging_.blockPart.get(), c.part, 6 * sizeof(int), hipMemcpyHostToDevice, st),
             "restore");
    hipCheck(hipMemcpyAsync(staging_.blockCount.get(), &c.count, sizeof(int), hipMemcpyHostToDevice, st),
             "restore");
    const auto ptr = setPointers(setBuf_);
    fdmCheck(fdm_launchTapeRestore(staging_.view(s_.status_.get()), 0, ptr.data(), (int)N_, maxBlock_,
                                   tapeWs_.get(), tapeWs_.size(), st),
             "history restore");
    const int one = 1;
    hipCheck(hipMemcpyAsync(s_.ctl_.get() + FDM_CTL_FORCE, &one, sizeof(int), hipMemcpyHostToDevice, st),
             "force rebuild");
    hipCheck(hipStreamSynchronize(st), "restore");
  }
  void reverseStep(int j) {
    fdm_adjoint_step_args a{};
    a.before = slotView(slots_[j - 1]);
    a.after = slotView(slots_[j]);
    a.pgeo = s_.particles_.geometry();
    a.pmat = s_.particles_.deviceMaterials();
    a.walls = s_.walls_.bodies();
    a.walls.x = slots_[j].wx.get();
    a.walls.q = slots_[j].wq.get();
    a.wgeo = s_.walls_.geometry();
    a.wmat = s_.walls_.deviceMaterials();
    a.gravity = {s_.gravity_.x, s_.gravity_.y, s_.gravity_.z};
    a.dt = s_.dt_;
    a.pgrid = s_.particles_.grid();
    a.inst = s_.inst_;
    a.tape = tape_.view(s_.status_.get());
    a.block = j;
    a.prevBlock = j - 1;
    a.maxBlockContacts = maxBlock_;
    a.xBar = xBar_.get();
    a.qBar = qBar_.get();
    a.vBar = vBar_.get();
    a.wBar = wBar_.get();
    a.FBar = FBar_.get();
    a.TBar = TBar_.get();
    a.springBar = springBar_.get();
    a.paramBar = paramBar_.get();
    a.wallMat = wallMat_.get();
    a.wallShare = wallShare_.get();
    a.status = s_.status_.get();
    fdmCheck(fdm_launchAdjointStep(&a, adjWs_.get(), adjWs_.size(), s_.stream_), "adjoint step");
  }

  LSSolver& s_;
  int contactsPerParticle_ = 16;
  size_t steps_ = 0, every_ = 1, N_ = 0, W_ = 0;
  int maxBlock_ = 0;
  long long pairCap_ = 0;
  std::array<fdm_pairs, 5> setBuf_{};
  std::vector<Checkpoint> checkpoints_;
  Checkpoint final_, mark_;
  std::vector<StateBuf> slots_;
  TapeBuf tape_, staging_;
  DeviceBuffer<char> tapeWs_, adjWs_;
  DeviceBuffer<double> xBar_, qBar_, vBar_, wBar_, FBar_, TBar_, paramBar_, wallShare_, springBar_;
  DeviceBuffer<int> wallMat_;
};

}  // namespace fundem
