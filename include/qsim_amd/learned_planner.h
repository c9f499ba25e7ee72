// learned_planner.h -- features of a pass for the pass-time cost model, and
// the candidate partitions the learned planner scores.
//
// qsim's MultiQubitGateFuser (lib/fuser_mqubit.h:185-526) fixes the
// partition by max_fused_size alone.  The planner instead builds several
// candidate partitions of a circuit -- qsim's fuser at max_fused_size 2..6
// with each gate arithmetic, and cache-blocked passes (block_planner.h) of
// several tile shapes -- describes every pass of each by the feature vector
// below, asks the cost model (a small MLP evaluated on the GPU,
// qsim_hip_costmodel_forward) for each pass's log time, and keeps the
// candidate with the least predicted total.  Without a model it falls back
// to qsim's fuser at max_fused_size 4.
#ifndef QSIM_AMD_LEARNED_PLANNER_H_
#define QSIM_AMD_LEARNED_PLANNER_H_

#include <algorithm>
#include <cmath>
#include <vector>

#include "qsim_hip.h"

namespace qsim {
namespace hip {

constexpr unsigned kCostFeatures = 19;
constexpr unsigned kCostHidden = 64;

// Kernel family a pass runs on (gate_kernels.hip / mfma_kernels.hip /
// f16_kernels.hip dispatch).
enum KernelFamily { kVector = 0, kMfmaF32 = 1, kMfmaBf16 = 2, kMfmaF16 = 3 };

inline int GateFamily(unsigned k, unsigned num_controls, int method,
                      unsigned n) {
  if (method == QSIM_HIP_METHOD_VECTOR || num_controls > 0 || k < 4 ||
      k > 6 || n < k + 5) {
    return kVector;
  }
  if (method == QSIM_HIP_METHOD_MFMA_BF16X3 && k >= 5) return kMfmaBf16;
  // The FP16 kernels take 5- and 6-qubit gates on 12 qubits or more (the
  // FP32 MFMA kernels run the rest).
  if (method == QSIM_HIP_METHOD_MFMA_FP16X2 && k >= 5 && n >= 12) {
    return kMfmaF16;
  }
  return kMfmaF32;
}

// Target: log(pass ms) - (n - 24) ln 2, the log time per 2^24 amplitudes.
inline float CostTargetOffset(unsigned n) {
  return float((double(n) - 24.0) * std::log(2.0));
}

// Feature vector of a gate pass:
//   0 n, 1 blocked flag, 2 k, 3..8 sorted targets / n (0 past k),
//   9 targets below qubit 5 (the LDS-tile boundary of the gate kernels),
//   10 targets below qubit 13, 11 controlled flag, 12 controls,
//   13 diagonal flag, 14..17 family one-hot (vector, FP32 MFMA, BF16 MFMA,
//   FP16 MFMA), 18 lowest target / n.
inline void GateFeatures(unsigned n, const std::vector<unsigned>& qs,
                         unsigned num_controls, bool diagonal, int family,
                         float* f) {
  std::fill(f, f + kCostFeatures, 0.0f);
  const float fn = float(n);
  f[0] = fn;
  f[2] = float(qs.size());
  unsigned below5 = 0, below13 = 0;
  for (size_t i = 0; i < qs.size(); ++i) {
    if (i < 6) f[3 + i] = float(qs[i]) / fn;
    below5 += qs[i] < 5;
    below13 += qs[i] < 13;
  }
  f[9] = float(below5);
  f[10] = float(below13);
  f[11] = num_controls > 0 ? 1.0f : 0.0f;
  f[12] = float(num_controls);
  f[13] = diagonal ? 1.0f : 0.0f;
  f[14 + std::min(family, 3)] = 1.0f;
  f[18] = qs.empty() ? 0.0f : float(qs[0]) / fn;
}

// Feature vector of a cache-blocked pass: 0 n, 1 blocked flag, 2 tile
// qubits, 3..8 the lowest six tile qubits above the low run / n,
// 9 low-run length, 10 fused gates in the pass, 18 lowest high qubit / n.
inline void BlockFeatures(unsigned n, const std::vector<unsigned>& tile,
                          unsigned num_ops, float* f) {
  std::fill(f, f + kCostFeatures, 0.0f);
  const float fn = float(n);
  f[0] = fn;
  f[1] = 1.0f;
  f[2] = float(tile.size());
  unsigned lc = 0;
  while (lc < tile.size() && tile[lc] == lc) ++lc;
  for (size_t i = lc; i < tile.size() && i < lc + 6; ++i) {
    f[3 + (i - lc)] = float(tile[i]) / fn;
  }
  f[9] = float(lc);
  f[10] = float(num_ops);
  f[18] = lc < tile.size() ? float(tile[lc]) / fn : 1.0f;
}

// A candidate partition: qsim's fuser at max_fused_size `a` with gate method
// `method`, or blocked passes of `a` tile qubits over the `b` lowest.
struct PlanCandidate {
  int blocked;
  unsigned a, b;
  int method;
};

inline std::vector<PlanCandidate> DefaultCandidates() {
  std::vector<PlanCandidate> c;
  for (unsigned k = 2; k <= 6; ++k) {
    c.push_back({0, k, 0, QSIM_HIP_METHOD_AUTO});
    if (k >= 4) c.push_back({0, k, 0, QSIM_HIP_METHOD_VECTOR});
    if (k >= 5) c.push_back({0, k, 0, QSIM_HIP_METHOD_MFMA_BF16X3});
    if (k >= 5) c.push_back({0, k, 0, QSIM_HIP_METHOD_MFMA_FP16X2});
  }
  const unsigned shapes[][2] = {{14, 4}, {14, 5}, {14, 6}, {13, 4},
                                {13, 5}, {13, 6}, {12, 5}, {11, 4}};
  for (const auto& s : shapes) c.push_back({1, s[0], s[1], QSIM_HIP_METHOD_AUTO});
  return c;
}

// The fallback partition: qsim's fuser at max_fused_size 4.
inline PlanCandidate FallbackCandidate() {
  return {0, 4, 0, QSIM_HIP_METHOD_AUTO};
}

}  // namespace hip
}  // namespace qsim

#endif  // QSIM_AMD_LEARNED_PLANNER_H_
