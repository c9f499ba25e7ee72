// statespace_hip.h -- qsim StateSpace over the gfx950 backend.
//
// Drop-in counterpart of qsim's StateSpaceCUDA (lib/statespace_cuda.h:42-468):
// the same members with the same meaning, implemented by the C ABI of
// qsim_hip.h.  The device layout is interleaved complex (normal order), so
// InternalToNormalOrder / NormalToInternalOrder do nothing.  Needs qsim's
// lib/ on the include path for the CRTP base (lib/statespace.h) and the
// sampling draws (lib/util.h).

#ifndef QSIM_AMD_STATESPACE_HIP_H_
#define QSIM_AMD_STATESPACE_HIP_H_

#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

#include "qsim_amd/vectorspace_hip.h"
#include "qsim_hip.h"
#include "statespace.h"
#include "util.h"

namespace qsim {

template <typename FP = float>
class StateSpaceHIP
    : public StateSpace<StateSpaceHIP<FP>, VectorSpaceHIP, FP> {
 private:
  using Base = StateSpace<StateSpaceHIP<FP>, qsim::VectorSpaceHIP, FP>;
  static constexpr int kPrec = hip_detail::Prec<FP>();

 public:
  using State = typename Base::State;
  using fp_type = typename Base::fp_type;
  using MeasurementResult = typename Base::MeasurementResult;

  // Accepted for source compatibility with StateSpaceCUDA::Parameter
  // (lib/statespace_cuda.h:59-70); launch shapes are chosen per kernel.
  struct Parameter {
    unsigned num_threads = 512;
    unsigned num_dblocks = 16;
  };

  StateSpaceHIP() {}
  explicit StateSpaceHIP(const Parameter&) {}

  static uint64_t MinSize(unsigned num_qubits) {
    return qsim_hip_min_size(num_qubits);
  }

  void InternalToNormalOrder(State& state) const {
    hip_detail::Check(qsim_hip_internal_to_normal_order(
                          kPrec, state.get(), state.num_qubits()),
                      "InternalToNormalOrder");
  }

  void NormalToInternalOrder(State& state) const {
    hip_detail::Check(qsim_hip_normal_to_internal_order(
                          kPrec, state.get(), state.num_qubits()),
                      "NormalToInternalOrder");
  }

  void SetAllZeros(State& state) const {
    hip_detail::Check(
        qsim_hip_set_all_zeros(kPrec, state.get(), state.num_qubits()),
        "SetAllZeros");
  }

  void SetStateUniform(State& state) const {
    hip_detail::Check(
        qsim_hip_set_state_uniform(kPrec, state.get(), state.num_qubits()),
        "SetStateUniform");
  }

  void SetStateZero(State& state) const {
    hip_detail::Check(
        qsim_hip_set_state_zero(kPrec, state.get(), state.num_qubits()),
        "SetStateZero");
  }

  // Thread-safe single-amplitude access (one 8/16-byte copy).
  static std::complex<fp_type> GetAmpl(const State& state, uint64_t i) {
    double re, im;
    hip_detail::Check(qsim_hip_get_ampl(kPrec, state.get(), state.num_qubits(),
                                        i, &re, &im),
                      "GetAmpl");
    return std::complex<fp_type>(fp_type(re), fp_type(im));
  }

  // Many amplitudes in one gather pass (no reference counterpart: it
  // replaces loops of GetAmpl, e.g. apps/qsim_amplitudes.cc:138-153).
  static std::vector<std::complex<fp_type>> GetAmpls(
      const State& state, const std::vector<uint64_t>& indices) {
    std::vector<std::complex<fp_type>> out(indices.size());
    hip_detail::Check(
        qsim_hip_get_ampls(kPrec, state.get(), state.num_qubits(),
                           indices.data(), indices.size(), out.data()),
        "GetAmpls");
    return out;
  }

  static void SetAmpl(State& state, uint64_t i,
                      const std::complex<fp_type>& ampl) {
    SetAmpl(state, i, std::real(ampl), std::imag(ampl));
  }

  static void SetAmpl(State& state, uint64_t i, fp_type re, fp_type im) {
    hip_detail::Check(qsim_hip_set_ampl(kPrec, state.get(), state.num_qubits(),
                                        i, re, im),
                      "SetAmpl");
  }

  void BulkSetAmpl(State& state, uint64_t mask, uint64_t bits,
                   const std::complex<fp_type>& val,
                   bool exclude = false) const {
    BulkSetAmpl(state, mask, bits, std::real(val), std::imag(val), exclude);
  }

  void BulkSetAmpl(State& state, uint64_t mask, uint64_t bits, fp_type re,
                   fp_type im, bool exclude = false) const {
    hip_detail::Check(
        qsim_hip_bulk_set_ampl(kPrec, state.get(), state.num_qubits(), mask,
                               bits, re, im, exclude ? 1 : 0),
        "BulkSetAmpl");
  }

  bool Add(const State& src, State& dest) const {
    if (src.num_qubits() != dest.num_qubits()) return false;
    hip_detail::Check(
        qsim_hip_add(kPrec, src.get(), dest.get(), src.num_qubits()), "Add");
    return true;
  }

  void Multiply(fp_type a, State& state) const {
    hip_detail::Check(
        qsim_hip_multiply(kPrec, state.get(), state.num_qubits(), a),
        "Multiply");
  }

  std::complex<double> InnerProduct(const State& state1,
                                    const State& state2) const {
    if (state1.num_qubits() != state2.num_qubits()) return std::nan("");
    double re, im;
    hip_detail::Check(qsim_hip_inner_product(kPrec, state1.get(), state2.get(),
                                             state1.num_qubits(), &re, &im),
                      "InnerProduct");
    return {re, im};
  }

  double RealInnerProduct(const State& state1, const State& state2) const {
    if (state1.num_qubits() != state2.num_qubits()) return std::nan("");
    double r;
    hip_detail::Check(
        qsim_hip_real_inner_product(kPrec, state1.get(), state2.get(),
                                    state1.num_qubits(), &r),
        "RealInnerProduct");
    return r;
  }

  double Norm(const State& state) const {
    double r;
    hip_detail::Check(
        qsim_hip_norm(kPrec, state.get(), state.num_qubits(), &r), "Norm");
    return r;
  }

  // Draws as lib/statespace_cuda.h:243-312: FP64 device norm, host draws of
  // GenerateRandomValues (lib/util.h:67-85), device inverse-CDF search.
  template <typename DistrRealType = double>
  std::vector<uint64_t> Sample(const State& state, uint64_t num_samples,
                               unsigned seed) const {
    std::vector<uint64_t> bitstrings;
    if (num_samples == 0) return bitstrings;
    double norm = Norm(state);
    auto rs = GenerateRandomValues<DistrRealType>(num_samples, seed, norm);
    std::vector<double> draws(rs.begin(), rs.begin() + num_samples);
    bitstrings.resize(num_samples, 0);
    hip_detail::Check(qsim_hip_sample(kPrec, state.get(), state.num_qubits(),
                                      draws.data(), num_samples,
                                      bitstrings.data()),
                      "Sample");
    return bitstrings;
  }

  void Collapse(const MeasurementResult& mr, State& state) const {
    hip_detail::Check(qsim_hip_collapse(kPrec, state.get(), state.num_qubits(),
                                        mr.mask, mr.bits),
                      "Collapse");
  }

  std::vector<double> PartialNorms(const State& state) const {
    std::vector<double> norms(qsim_hip_num_partial_norms(state.num_qubits()));
    hip_detail::Check(qsim_hip_partial_norms(kPrec, state.get(),
                                             state.num_qubits(), norms.data()),
                      "PartialNorms");
    return norms;
  }

  uint64_t FindMeasuredBits(unsigned m, double r, uint64_t mask,
                            const State& state) const {
    uint64_t bits;
    hip_detail::Check(
        qsim_hip_find_measured_bits(kPrec, state.get(), state.num_qubits(), m,
                                    r, mask, &bits),
        "FindMeasuredBits");
    return bits;
  }
};

}  // namespace qsim

#endif  // QSIM_AMD_STATESPACE_HIP_H_
