// LSAutotuner: picks the settings that do not change results for the
// fastest step on the solver's own system (no counterpart in the reference):
//   - the candidate-list skin (rebuild distance), as a fraction of the
//     largest particle radius (not with the sector periodic boundary, whose
//     images are rebuilt every step);
//   - the particle grid's cell size, as a multiple (1.25, 1.5, 2) of the
//     configured one or of the rule's 2 x the largest radius. A node reaches
//     a particle target within the target's radius, so any cell of at least
//     the largest radius puts every target within reach in the node's
//     27-cell stencil. With planar periodic faces an image exists for the
//     bodies of the first cell column (row) only, and two bodies touch
//     across a face only when their centres are within the sum of their
//     radii, at most twice the largest radius: when the configured cell is
//     at least that, every candidate (coarser) cell keeps every body that can
//     touch across a face in the first column, and a clamped ghost cell stays
//     in the stencil of every node within reach. A configured cell below
//     twice the largest radius (the pour's 2.5 mm) leaves some such bodies
//     without an image, as the reference does, and a coarser cell would add
//     their contacts, so periodic systems configured so keep their cell. The
//     contact set is therefore the same; with canonical sums the results are
//     the same bits. Not with the sector boundary (its images are rebuilt
//     each step by their own cell rule), nor for batched systems (their
//     packings are configured with cells below twice the largest radius: a
//     coarser candidate changed their contacts and overflowed a mirror list);
//   - the solver's launch shapes (fdm_tuning): node-sweep waves (kept pairs
//     in flight) per workgroup and sweep grid size, nodes per lane (64 or
//     128 nodes in flight per sweep wave), where the contact law runs (a
//     kernel of its own, or the sweep's wave), contact-law and reduction
//     workgroup sizes.
// Each candidate is timed from the same saved point (LSAdjoint::mark /
// rollback) one knob at a time, keeping the best value of each knob before
// the next; the solver is left at the saved point with the best settings.
// The cost of a candidate is its mean step time from the saved point (one
// forced list build, then the solver's own decisions), the better of two
// timings; the skin is timed over a longer window that holds real list
// rebuilds. A candidate replaces the current choice only when it is 2 %
// faster.
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
    // candidates are timed on eager launches: a cell-size candidate
    // re-initialises the solver, and every candidate re-captures its graphs
    struct GraphsBack {
      LSSolver& s;
      bool on;
      ~GraphsBack() { s.useGraph_ = on; }
    } graphsBack{s_, s_.useGraph_};
    s_.useGraph_ = false;
    snap_.mark();
    TuneChoice best;
    best.launch = s_.launchTuning();
    best.skin = s_.skin_;
    best.cellSize = s_.cellSize_;
    // the skin trades list builds for longer lists: time it over a window
    // that holds real rebuilds
    const size_t window = std::max<size_t>(5 * trialSteps, 60);
    const double R = s_.maxRadius_;
    std::vector<TuneChoice> group;
    double bestCost = 0.0;
    if (!s_.periodic_.isSector()) {  // sector images are rebuilt every step: no skin
      for (double f : {0.05, 0.1, 0.2, 0.35}) {
        TuneChoice c = best;
        c.skin = f * R;
        group.push_back(c);
      }
      bestCost = pick(group, window, best, "skin");
    }
    const double base = s_.particles_.cellSizeRule(s_.cellSize_);
    const bool periodicOk = !s_.periodic_.isActived() || base >= 2.0 * s_.maxRadius_;
    const bool cellOk = tuneCellSize && !s_.periodic_.isSector() && s_.systems_.empty() && periodicOk;
    if (cellOk) {
      group.clear();
      for (double f : {1.25, 1.5, 2.0}) {
        TuneChoice c = best;
        c.cellSize = f * base;
        group.push_back(c);
      }
      bestCost = pick(group, window, best, "cell");
    }
    group.clear();
    for (int w : {1, 2, 4})
      for (int pc : {16, 32, 64}) {
        TuneChoice c = best;
        c.launch.sweepWaves = w;
        c.launch.sweepWavesPerCU = pc;
        group.push_back(c);
      }
    bestCost = pick(group, trialSteps, best, "sweep");
    group.clear();
    for (int m : {1, 2}) {
      TuneChoice c = best;
      c.launch.lawMode = m;
      group.push_back(c);
    }
    bestCost = pick(group, trialSteps, best, "lawMode");
    group.clear();
    if (best.launch.lawMode == 1) {
      for (int npl : {1, 2, 3}) {
        TuneChoice c = best;
        c.launch.nodesPerLane = npl;
        group.push_back(c);
      }
      bestCost = pick(group, trialSteps, best, "nodesPerLane");
      group.clear();
    }
    for (int b : {64, 128, 256}) {
      TuneChoice c = best;
      c.launch.lawBlock = b;
      group.push_back(c);
    }
    bestCost = pick(group, trialSteps, best, "lawBlock");
    group.clear();
    for (int b : {128, 256, 512}) {
      TuneChoice c = best;
      c.launch.reduceBlock = b;
      group.push_back(c);
    }
    bestCost = pick(group, trialSteps, best, "reduceBlock");
    apply(best);
    snap_.rollback();
    best.msPerStep = bestCost;
    best.trials = log_;
    return best;
  }

 private:
  void apply(const TuneChoice& c) {
    s_.setLaunchTuning(c.launch);
    s_.skinRequest_ = c.skin;
    s_.skin_ = c.skin;
    if (c.cellSize != s_.cellSize_) {
      s_.setCellSize(c.cellSize);
      s_.initialize(s_.domainMin_, s_.domainMax_, s_.gravity_, s_.dt_);
      s_.skin_ = c.skin;  // initialize derives the skin from the request, which is c.skin
    }
  }
  // the current choice against a group of candidates, each timed twice
  // (the smaller time counts); a candidate must be 2 % faster to be taken
  double pick(const std::vector<TuneChoice>& group, size_t steps, TuneChoice& best, const char* knob) {
    double bestCost = std::min(cost(steps, best, knob), cost(steps, best, knob));
    TuneChoice choice = best;
    double choiceCost = bestCost;
    for (const TuneChoice& c : group) {
      const double t = std::min(cost(steps, c, knob), cost(steps, c, knob));
      if (t < choiceCost) {
        choiceCost = t;
        choice = c;
      }
    }
    if (choiceCost < 0.98 * bestCost) {
      best = choice;
      bestCost = choiceCost;
    }
    return bestCost;
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
    char line[320];
    if (status != 0) {
      total = std::numeric_limits<double>::infinity();
      hipCheck(hipMemsetAsync(s_.status_.get(), 0, sizeof(unsigned), st), "tune");
    } else {
      total = (build + rest * double(steps - 1)) / double(steps);
    }
    std::snprintf(line, sizeof line,
                  "%-12s skin %.3e cell %.3e sweep %d/%d x%d law %d/%d reduce %d: first %.3f ms, "
                  "then %.3f ms/step, %ld builds, displacement bound %.2e -> %.4f ms/step",
                  knob, c.skin, c.cellSize, c.launch.sweepWaves, c.launch.sweepWavesPerCU,
                  c.launch.nodesPerLane, c.launch.lawMode, c.launch.lawBlock, c.launch.reduceBlock,
                  build, rest, (long)(b1 - b0), disp, total);
    log_.push_back(line);
    return total;
  }

  LSSolver& s_;
  LSAdjoint snap_;
  std::vector<std::string> log_;
};

}  // namespace fundem
