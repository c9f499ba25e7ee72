// simulator_hip.h -- qsim Simulator over the gfx950 backend.
//
// Drop-in counterpart of qsim's SimulatorCUDA (lib/simulator_cuda.h:34-260):
// ApplyGate, ApplyControlledGate, ExpectationValue and SIMDRegisterSize with
// the reference's argument meaning (qs ascending, row/column bit j <-> qs[j],
// cvals bit i <-> cqs[i]) and error conventions (gates on more qubits than
// supported are a silent no-op, lib/simulator_cuda.h:96-98; device errors end
// the process).  Host-only C++ over the C ABI of qsim_hip.h.

#ifndef QSIM_AMD_SIMULATOR_HIP_H_
#define QSIM_AMD_SIMULATOR_HIP_H_

#include <complex>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "qsim_amd/statespace_hip.h"
#include "qsim_hip.h"

namespace qsim {

template <typename FP = float>
class SimulatorHIP final {
 public:
  using StateSpace = StateSpaceHIP<FP>;
  using State = typename StateSpace::State;
  using fp_type = typename StateSpace::fp_type;

  SimulatorHIP() {}

  // psi <- U psi on the ascending qubits qs (0 <= qs.size() <= 6).
  void ApplyGate(const std::vector<unsigned>& qs, const fp_type* matrix,
                 State& state) const {
    Apply(qs, nullptr, 0, 0, matrix, hip_detail::Prec<fp_type>(), state);
  }

  // The same with a matrix in another precision (host-built FP64 fused
  // matrices for complex64 states).  No reference counterpart.
  template <typename M, typename = std::enable_if_t<
                            !std::is_same<M, fp_type>::value &&
                            (std::is_same<M, float>::value ||
                             std::is_same<M, double>::value)>>
  void ApplyGate(const std::vector<unsigned>& qs, const M* matrix,
                 State& state) const {
    Apply(qs, nullptr, 0, 0, matrix, hip_detail::Prec<M>(), state);
  }

  // Controlled U: targets qs, ascending controls cqs with values cvals.
  void ApplyControlledGate(const std::vector<unsigned>& qs,
                           const std::vector<unsigned>& cqs, uint64_t cvals,
                           const fp_type* matrix, State& state) const {
    Apply(qs, cqs.data(), cqs.size(), cvals, matrix,
          hip_detail::Prec<fp_type>(), state);
  }

  // <psi|U|psi> with FP64 accumulation; 0 for unsupported sizes, as
  // lib/simulator_cuda.h:216-260 returns.
  std::complex<double> ExpectationValue(const std::vector<unsigned>& qs,
                                        const fp_type* matrix,
                                        const State& state) const {
    double re = 0, im = 0;
    int rc = qsim_hip_expectation_value(
        hip_detail::Prec<fp_type>(), state.get(), state.num_qubits(),
        qs.data(), qs.size(), matrix, hip_detail::Prec<fp_type>(), &re, &im);
    if (rc == QSIM_HIP_ERR_UNSUPPORTED) return 0;
    hip_detail::Check(rc, "ExpectationValue");
    return {re, im};
  }

  // The test fixtures size their problems by it
  // (tests/simulator_testfixture.h:1153): one wave64 of amplitudes.
  static unsigned SIMDRegisterSize() { return 64; }

 private:
  void Apply(const std::vector<unsigned>& qs, const unsigned* cqs, unsigned nc,
             uint64_t cvals, const void* matrix, int mprec,
             State& state) const {
    int rc = qsim_hip_apply_controlled_gate(
        hip_detail::Prec<fp_type>(), state.get(), state.num_qubits(),
        qs.data(), qs.size(), cqs, nc, cvals, matrix, mprec);
    if (rc == QSIM_HIP_ERR_UNSUPPORTED) return;
    hip_detail::Check(rc, "ApplyGate");
  }
};

}  // namespace qsim

#endif  // QSIM_AMD_SIMULATOR_HIP_H_
