// The tutorials' external-force hook launcher (apps/tutorial/globalDamping.cuh:6-15)
// with its own name and argument list, so that an LSSolver subclass written for
// the reference, e.g.
//
//   void addExternalForceTorque(const double time) override {
//     launchAddGlobalDampingForceTorque(getLSParticle().force(), getLSParticle().torque(),
//         getLSParticle().velocity(), getLSParticle().angularVelocity(), dampingCoefficient,
//         getLSParticle().num_device(), getLSParticle().gridDim(), getLSParticle().blockDim(), 0);
//   }
//
// compiles unchanged against this header. The arrays are this build's
// planar SoA device state (fdm_real: FP32, or FP64 in the FP64-storage build). The tutorial's constant-force launcher has no
// body count in its arguments, which planar SoA needs to address a body's
// y and z components; fdm_launchAddGlobalConstantForceTorque takes the bodies.
#pragma once
#include <stdexcept>
#include <string>

#include "../../../include/fundem_amd.h"

inline void launchAddGlobalDampingForceTorque(fdm_real* force, fdm_real* torque,
                                              const fdm_real* velocity,
                                              const fdm_real* angularVelocity,
                                              const double dampingCoefficient, const size_t num,
                                              const size_t gridD, const size_t blockD,
                                              hipStream_t stream) {
  const int r = fdm_launchAddGlobalDampingForceTorqueArrays(
      force, torque, velocity, angularVelocity, dampingCoefficient, num, gridD, blockD, stream);
  if (r != FDM_OK)
    throw std::runtime_error("launchAddGlobalDampingForceTorque failed (" + std::to_string(r) + ")");
}
