// VBondedInteraction: beam-like bonds between particles (kernel/
// VBondedInteraction.h:61-566). Bond constants and local frames are computed
// on the host in FP64 exactly as the reference's add() does, then stored FP32
// on the device. Per-bond results feed the deterministic reduction.
#pragma once
#include <algorithm>
#include <iostream>
#include <vector>

#include "../../../include/fundem_amd.h"
#include "DeviceBuffer.h"
#include "types.h"

namespace fundem {

struct VBondedInteraction {
 public:
  VBondedInteraction() = default;
  VBondedInteraction(const VBondedInteraction&) = delete;
  VBondedInteraction& operator=(const VBondedInteraction&) = delete;

  // VBondedInteraction::add (VBondedInteraction.h:74-195). Positions and
  // orientations are those the device holds (FP32), so that the local frames
  // match the device state.
  void add(int master, int slave, double3 xi, double3 xj, quaternion qi, quaternion qj,
           double3 point, double3 normal, double radius, double length0, double youngs,
           double poisson, double tensile, double cohesion, double friction) {
    if (isZero(lengthSquared(normal))) {
      std::cerr << "[VBondedInteraction] Invalid bond normal: (" << normal.x << ", " << normal.y
                << ", " << normal.z << ")." << std::endl;
      return;
    }
    if (radius <= 0.) {
      std::cerr << "[VBondedInteraction] Invalid bond radius: " << radius << "." << std::endl;
      return;
    }
    if (length0 <= 0.) {
      std::cerr << "[VBondedInteraction] Invalid bond length: " << length0 << "." << std::endl;
      return;
    }
    if (youngs <= 0.) {
      std::cerr << "[VBondedInteraction] Invalid Young's modulus: " << youngs << "." << std::endl;
      return;
    }
    if (poisson <= -1.0 || poisson >= 0.5) {
      std::cerr << "[VBondedInteraction] Invalid Poisson ratio: " << poisson << "." << std::endl;
      return;
    }
    const double Rb = radius, Lb = length0, Eb = youngs;
    const double Gb = Eb / (2. * (1. + poisson));
    const double Ab = pi() * Rb * Rb;
    const double Ib = Ab * Rb * Rb / 4.;
    const double Jb = 2. * Ib;
    const double Kn = Eb * Ab / Lb, Kt = Gb * Jb / Lb, Ks = 12. * Eb * Ib / (Lb * Lb * Lb),
                 Kb = Eb * Ib / Lb;
    const double B1 = Kn, B2 = Lb * Lb * Ks, B4 = Kt, B3 = Kb - B2 / 4. - B4 / 2.;
    const double3 n = vec::divr(normal, vec::len(normal));
    const double3 gpi = vec::add(point, vec::mul(n, 0.5 * Lb)), gpj = vec::sub(point, vec::mul(n, 0.5 * Lb));
    const double3 lpi = vec::rotateT(vec::sub(gpi, xi), qi), lpj = vec::rotateT(vec::sub(gpj, xj), qj);
    const double3 N1i = vec::neg(n);
    const double3 ref = (std::fabs(N1i.x) < 0.9) ? make_double3(1., 0., 0.) : make_double3(0., 1., 0.);
    const double3 N2i = vec::unit(vec::cross(ref, N1i));
    const double3 N3i = vec::unit(vec::cross(N1i, N2i));
    const double3 N1j = vec::neg(N1i), N2j = N2i, N3j = N3i;
    const double coef[9] = {B1, B2, B3, B4, Rb, Lb, tensile, cohesion, friction};
    for (double c : coef) coef_.push_back((fdm_real)c);
    const double3 loc[8] = {vec::unit(vec::rotateT(N1i, qi)), vec::unit(vec::rotateT(N2i, qi)), vec::unit(vec::rotateT(N3i, qi)),
                            lpi,
                            vec::unit(vec::rotateT(N1j, qj)), vec::unit(vec::rotateT(N2j, qj)), vec::unit(vec::rotateT(N3j, qj)),
                            lpj};
    for (const double3& v : loc) {
      local_.push_back((fdm_real)v.x);
      local_.push_back((fdm_real)v.y);
      local_.push_back((fdm_real)v.z);
    }
    master_.push_back(master);
    slave_.push_back(slave);
    center_.push_back(point);
  }

  // Upload and build the per-particle bond lists (ascending bond id).
  void initialize(size_t numParticles, const int* stepCounter, hipStream_t s) {
    const size_t n = master_.size();
    if (n == 0) return;
    std::vector<int> cnt(numParticles + 1, 0);
    for (size_t k = 0; k < n; ++k) {
      cnt[master_[k] + 1]++;
      cnt[slave_[k] + 1]++;
    }
    for (size_t i = 0; i < numParticles; ++i) cnt[i + 1] += cnt[i];
    std::vector<int> list(cnt[numParticles]), fill(cnt.begin(), cnt.end() - 1);
    for (size_t k = 0; k < n; ++k) {
      list[fill[master_[k]]++] = int(k) * 2;
      list[fill[slave_[k]]++] = int(k) * 2 + 1;
    }
    dmaster_.upload(master_, s);
    dslave_.upload(slave_, s);
    if (active_.empty()) {
      active_.assign(n, 1);
      breakStep_.assign(n, -1);
    }
    dactive_.upload(active_, s);
    dbreak_.upload(breakStep_, s);
    dcoef_.upload(coef_, s);
    dlocal_.upload(local_, s);
    outF_.allocate(3 * n);
    outTi_.allocate(3 * n);
    outTj_.allocate(3 * n);
    applied_.allocate(n);
    sigma_.allocate(n);
    tau_.allocate(n);
    partStart_.upload(cnt, s);
    partList_.upload(list, s);
    stepCounter_ = stepCounter;
    hipCheck(hipStreamSynchronize(s), "bond upload");
  }

  void finalize(hipStream_t s) {
    if (master_.empty() || !dactive_.get()) return;
    active_ = dactive_.toHost(s, master_.size());
    breakStep_ = dbreak_.toHost(s, master_.size());
  }

  fdm_bonds view() {
    fdm_bonds b;
    b.n = (int)master_.size();
    b.master = dmaster_.get();
    b.slave = dslave_.get();
    b.active = dactive_.get();
    b.breakStep = dbreak_.get();
    b.coef = dcoef_.get();
    b.local = dlocal_.get();
    b.outF = outF_.get();
    b.outTi = outTi_.get();
    b.outTj = outTj_.get();
    b.applied = applied_.get();
    b.maxNormalStress = sigma_.get();
    b.maxShearStress = tau_.get();
    b.partStart = partStart_.get();
    b.partList = partList_.get();
    b.stepIndex = stepCounter_;
    return b;
  }

  size_t numPair() const { return master_.size(); }
  size_t numPair_device() const { return dactive_.get() ? master_.size() : 0; }
  const std::vector<int>& activatedHostRef() const { return active_; }
  const std::vector<int>& breakStepHostRef() const { return breakStep_; }
  const std::vector<int>& masterObjectIDHostRef() const { return master_; }
  const std::vector<int>& slaveObjectIDHostRef() const { return slave_; }
  // per bond: B1 B2 B3 B4 R L0 tensile cohesion friction; local frames n1 n2 n3 p of
  // master then slave (FP32, as uploaded)
  const std::vector<fdm_real>& coefHost() const { return coef_; }
  const std::vector<fdm_real>& localHost() const { return local_; }
  std::vector<double> maxNormalStress(hipStream_t s) { return sigma_.toHost(s, master_.size()); }
  std::vector<double> maxShearStress(hipStream_t s) { return tau_.toHost(s, master_.size()); }
  double deviceMemoryGB() const {
    return dcoef_.memoryGB() + dlocal_.memoryGB() + outF_.memoryGB() * 3 + sigma_.memoryGB() * 2;
  }

 private:
  std::vector<int> master_, slave_, active_, breakStep_;
  std::vector<fdm_real> coef_, local_;
  std::vector<double3> center_;
  DeviceBuffer<int> dmaster_, dslave_, dactive_, dbreak_, applied_, partStart_, partList_;
  DeviceBuffer<fdm_real> dcoef_, dlocal_;
  DeviceBuffer<double> outF_, outTi_, outTj_, sigma_, tau_;
  const int* stepCounter_ = nullptr;
};

}  // namespace fundem
