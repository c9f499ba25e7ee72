// factory_hip.h -- the Factory that qsim's drivers are templated on.
//
// qsim's runners never name a backend; they take a Factory with
// Simulator/StateSpace types and CreateSimulator()/CreateStateSpace()
// (lib/run_qsim.h:33-40, lib/qtrajectory.h:34-44, lib/hybrid.h:302-311;
// example: apps/qsim_base_cuda.cu:123-138).  FactoryHIP plugs the gfx950
// backend into them.

#ifndef QSIM_AMD_FACTORY_HIP_H_
#define QSIM_AMD_FACTORY_HIP_H_

#include "qsim_amd/simulator_hip.h"

namespace qsim {

template <typename FP = float>
struct FactoryHIP {
  using Simulator = SimulatorHIP<FP>;
  using StateSpace = typename Simulator::StateSpace;
  using fp_type = FP;

  FactoryHIP() {}
  explicit FactoryHIP(const typename StateSpace::Parameter&) {}

  StateSpace CreateStateSpace() const { return StateSpace(); }
  Simulator CreateSimulator() const { return Simulator(); }
};

}  // namespace qsim

#endif  // QSIM_AMD_FACTORY_HIP_H_
