// LSAutotuner: picks the settings that do not change results for the
// fastest step on the solver's own system (no counterpart in the reference):
//   - the candidate-list skin (rebuild distance), as a fraction of the
//     largest particle radius;
//   - the particle grid's cell size, as a multiple of the configured one
//     (only for one non-periodic system without bonds: there every candidate
//     is at least the largest radius, so the 27-cell stencil finds every
//     particle within reach and the contact set is the same);
//   - the launch shapes of fdm_tuning: node-sweep waves (kept pairs in
//     flight) per workgroup and sweep grid size, contact-law and reduction
//     workgroup sizes.
// Each candidate is timed from the same saved point (LSAdjoint::mark /
// rollback) one knob at a time, keeping the best value of each knob before
// the next; the solver is left at the saved point with the best settings.
// The cost of a candidate is the time of a list-building step plus the
// time of kept-list steps, amortised over the rebuild interval that the
// measured motion gives for its skin.
#pragma once
#include <chrono>
#include <cstdio>
#include <limits>
#include <string>
#include <vector>

#include "LSAdjoint.h"

namespace fundem {

struct TuneChoice {
  double skin = 0.0;      // m
  double cellSize = 0.0;  // m (0: the solver's own rule)
  fdm_tuning launch{};
  double msPerStep = 0.0;
  std::vector<std::string> trials;  // one line per timed candidate
};

class LSAutotuner {
 public:
  explicit LSAutotuner(LSSolver& s) : s_(s), snap_(s) {}

  TuneChoice run(size_t trialSteps = 20, bool tuneCellSize = true) {
    if (s_.particles_.num_device() == 0) throw std::logic_error("initialise the solver first");
    trialSteps = std::max<size_t>(trialSteps, 4);
    snap_.mark();
    TuneChoice best;
    fdm_getTuning(&best.launch);
    best.skin = s_.skin_;
    best.cellSize = s_.cellSize_;
    double bestCost = cost(trialSteps, best, "start");
    const double R = s_.maxRadius_;
    for (double f : {0.0, 0.05, 0.1, 0.2, 0.35}) {
      TuneChoice c = best;
      c.skin = f * R;
      tryCandidate(c, trialSteps, best, bestCost, "skin");
    }
    const bool cellOk = tuneCellSize && s_.systems_.empty() && !s_.periodic_.isActived() &&
                        s_.bonds_.view().n == 0;
    if (cellOk) {
      const fdm_grid g = s_.particles_.grid();
      const double base = s_.cellSize_ > 0.0 ? s_.cellSize_ : 1.0 / g.invCell.x;
      for (double f : {1.25, 1.5, 2.0}) {
        TuneChoice c = best;
        c.cellSize = f * base;
        tryCandidate(c, trialSteps, best, bestCost, "cell");
      }
    }
    for (int w : {1, 2, 4}) {
      TuneChoice c = best;
      c.launch.sweepWaves = w;
      tryCandidate(c, trialSteps, best, bestCost, "sweepWaves");
    }
    for (int w : {16, 32, 64}) {
      TuneChoice c = best;
      c.launch.sweepWavesPerCU = w;
      tryCandidate(c, trialSteps, best, bestCost, "sweepWavesPerCU");
    }
    for (int b : {64, 128, 256}) {
      TuneChoice c = best;
      c.launch.lawBlock = b;
      tryCandidate(c, trialSteps, best, bestCost, "lawBlock");
    }
    for (int b : {128, 256, 512}) {
      TuneChoice c = best;
      c.launch.reduceBlock = b;
      tryCandidate(c, trialSteps, best, bestCost, "reduceBlock");
    }
    apply(best);
    snap_.rollback();
    best.msPerStep = bestCost;
    best.trials = log_;
    return best;
  }

 private:
  void apply(const TuneChoice& c) {
    fdmCheck(fdm_setTuning(&c.launch), "tuning");
    s_.destroyGraphs();
    s_.skinRequest_ = c.skin;
    s_.skin_ = c.skin;
    if (c.cellSize != s_.cellSize_) {
      s_.setCellSize(c.cellSize);
      s_.initialize(s_.domainMin_, s_.domainMax_, s_.gravity_, s_.dt_);
      s_.skin_ = c.skin;  // initialize derives the skin from the request, which is c.skin
    }
  }
  void tryCandidate(const TuneChoice& c, size_t steps, TuneChoice& best, double& bestCost,
                    const char* knob) {
    const double t = cost(steps, c, knob);
    if (t < bestCost) {
      bestCost = t;
      best = c;
    }
  }
  // ms per step of a candidate, amortising its list builds
  double cost(size_t steps, const TuneChoice& c, const char* knob) {
    apply(c);
    snap_.rollback();
    hipStream_t st = s_.stream_;
    using clk = std::chrono::steady_clock;
    // warm-up (and graph capture): a forced build and kept-list steps
    s_.advance(3);
    snap_.rollback();
    hipCheck(hipStreamSynchronize(st), "tune");
    const auto t0 = clk::now();
    s_.advance(1);  // forced list build
    hipCheck(hipStreamSynchronize(st), "tune");
    const auto t1 = clk::now();
    const auto b0 = s_.listBuilds().first;
    s_.advance(steps - 1);
    hipCheck(hipStreamSynchronize(st), "tune");
    const auto t2 = clk::now();
    const auto b1 = s_.listBuilds().first;
    double disp = 0.0;
    if (s_.disp_.get())
      hipCheck(hipMemcpy(&disp, s_.disp_.get(), sizeof(double), hipMemcpyDeviceToHost), "tune");
    const unsigned status = s_.status();
    const double build = std::chrono::duration<double, std::milli>(t1 - t0).count();
    const double rest = std::chrono::duration<double, std::milli>(t2 - t1).count() / double(steps - 1);
    double total;
    char line[256];
    if (status != 0) {
      total = std::numeric_limits<double>::infinity();
      hipCheck(hipMemsetAsync(s_.status_.get(), 0, sizeof(unsigned), st), "tune");
    } else if (c.skin <= 0.0) {
      total = (build + rest * double(steps - 1)) / double(steps);
    } else {
      // steps until the displacement bound reaches half the skin, at the
      // motion measured since the last build of the trial
      const long kept = std::max<long>(1, (long)(steps - 1) - (long)(b1 - b0));
      const double perStep = disp / double(kept);
      const double interval = perStep > 0.0 ? std::max(1.0, 0.5 * c.skin / perStep) : 1e6;
      total = (build + rest * (interval - 1.0)) / interval;
      if (b1 > b0) total = (build + rest * double(steps - 1)) / double(steps);
    }
    std::snprintf(line, sizeof line,
                  "%-16s skin %.3e cell %.3e sweep %d/%d law %d reduce %d: build %.3f ms, step %.3f ms"
                  " -> %.4f ms/step",
                  knob, c.skin, c.cellSize, c.launch.sweepWaves, c.launch.sweepWavesPerCU,
                  c.launch.lawBlock, c.launch.reduceBlock, build, rest, total);
    log_.push_back(line);
    return total;
  }

  LSSolver& s_;
  LSAdjoint snap_;
  std::vector<std::string> log_;
};

}  // namespace fundem
