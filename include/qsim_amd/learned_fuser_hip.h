// learned_fuser_hip.h -- the learned planner as a qsim Fuser.
//
// qsim's drivers take the gate partition from a Fuser type: QSimRunner
// calls Fuser::FuseGates(param, num_qubits, ops[, times]) and applies the
// returned fused gates one by one (lib/run_qsim.h:109, :246, :264-280;
// lib/gate_appl.h:35-48); its Parameter derives from Fuser::Parameter
// (lib/run_qsim.h:45).  LearnedFuserHIP has the Fuser shape of qsim's
// MultiQubitGateFuser (lib/fuser_mqubit.h:152-290) and chooses the
// partition per circuit with the cost model:
//
//  - every candidate of DefaultCandidates() that runs with the library's
//    default gate arithmetic (qsim's fuser at max_fused_size 2..6, blocked
//    passes of eight tile shapes) is planned, each pass described by its
//    features (learned_planner.h), the network predicts each pass's time
//    (qsim_hip_costmodel_forward, on the GPU) and the candidate with the
//    least predicted total is kept;
//  - a fuser pick is qsim's MultiQubitGateFuser output itself;
//  - a blocked pick is returned as its fused (<= 4-qubit) gates, pass by
//    pass, with their component gates and FP64-built matrices, and the
//    backend's gate queue is set to that tile shape
//    (qsim_hip_set_gate_queue): SimulatorHIP::ApplyGate's calls then come
//    back together as the planned blocked passes (greedy grouping of
//    consecutive gates gives no more passes than the plan).
//
// Without a model (none given and no library default,
// qsim_hip_costmodel_default), for circuits below 24 qubits (under the
// model's training range, where planning costs more than the circuit), for
// calls with time boundaries or without matrices, and for lists of
// operation pointers (qtrajectory re-fuses each trajectory's deferred
// gates that way, lib/qtrajectory.h:399-408, too often to plan each), it
// is qsim's MultiQubitGateFuser at param.max_fused_size with the queue
// off.  The candidates are planned on host threads in parallel.
#ifndef QSIM_AMD_LEARNED_FUSER_HIP_H_
#define QSIM_AMD_LEARNED_FUSER_HIP_H_

#include <cmath>
#include <future>
#include <limits>
#include <type_traits>
#include <variant>
#include <vector>

#include "fuser_mqubit.h"
#include "gate.h"
#include "operation.h"
#include "qsim_amd/block_planner.h"
#include "qsim_amd/learned_planner.h"
#include "qsim_hip.h"

namespace qsim {
namespace hip {

// Blocked-planner settings of every blocked candidate: 64 greedy restarts
// with scores scaled in 1 +- 0.25 (on the 32-qubit grid circuit 21 passes
// instead of 26 for 14 tile qubits over a 4-qubit low run), at most as many
// fused gates per pass as the pass kernel keeps resident.
inline BlockPlannerParameter CandidateBlockParameter(unsigned tile_bits,
                                                     unsigned low_bits) {
  BlockPlannerParameter bp;
  bp.tile_bits = tile_bits;
  bp.low_bits = low_bits;
  bp.restarts = 64;
  bp.jitter = 0.5;
  bp.max_ops = qsim_hip_block_resident_ops(tile_bits);
  return bp;
}

template <typename FP>
bool IsDiagonalMatrix(const std::vector<FP>& m, unsigned k) {
  const unsigned D = 1u << k;
  if (m.size() < 2 * size_t(D) * D) return false;
  for (unsigned r = 0; r < D; ++r)
    for (unsigned c = 0; c < D; ++c)
      if (r != c && (m[2 * (r * D + c)] != 0 || m[2 * (r * D + c) + 1] != 0))
        return false;
  return true;
}

// Cost-model features of one entry of a fused partition (a fused gate, a
// measurement or an unfused operation) run with gate method `method`.
template <typename FP, typename FusedOp>
void FusedOpFeatures(unsigned n, const FusedOp& op, int method, float* f) {
  if (const auto* fg = std::get_if<FusedGate<FP>>(&op)) {
    const unsigned k = fg->qubits.size();
    GateFeatures(n, fg->qubits, 0, IsDiagonalMatrix(fg->matrix, k),
                 GateFamily(k, 0, method, n), f);
    return;
  }
  const Operation<FP>* o = nullptr;
  if (const auto* pp = std::get_if<const Operation<FP>*>(&op)) o = *pp;
  if (o != nullptr) {
    if (const auto* cg = OpGetAlternative<ControlledGate<FP>>(*o)) {
      GateFeatures(n, cg->qubits, cg->controlled_by.size(), false, kVector,
                   f);
      return;
    }
  }
  // Measurements and other operations: one reduction pass over the qubits.
  static const std::vector<unsigned> none;
  const std::vector<unsigned>* qs = nullptr;
  if (const auto* m = std::get_if<Measurement>(&op)) qs = &m->qubits;
  else if (o != nullptr) qs = &OpQubits(*o);
  GateFeatures(n, qs ? *qs : none, 0, false, kVector, f);
}

// A scored candidate partition.
struct CandidateResult {
  bool ok = false;
  unsigned passes = 0;
  double predicted_ms = std::numeric_limits<double>::quiet_NaN();
  BlockPlan plan;  // blocked candidates: the plan that was scored
};

// Feature rows of every pass of candidate c (and, blocked, its plan).
template <typename IO, typename FP>
bool CandidateRows(const PlanCandidate& c, unsigned n,
                   const std::vector<Operation<FP>>& ops,
                   std::vector<float>& rows, CandidateResult& r) {
  rows.clear();
  auto add = [&](auto&& fill) {
    const size_t o = rows.size();
    rows.resize(o + kCostFeatures);
    fill(&rows[o]);
  };
  typename MultiQubitGateFuser<IO>::Parameter p;
  p.verbosity = 0;
  if (!c.blocked) {
    if (c.a < 1 || c.a > 6) return false;
    p.max_fused_size = c.a;
    const auto fused = MultiQubitGateFuser<IO>::FuseGates(p, n, ops);
    if (fused.empty() && !ops.empty()) return false;
    for (const auto& e : fused) {
      add([&](float* f) { FusedOpFeatures<FP>(n, e, c.method, f); });
    }
    r.passes = fused.size();
    return true;
  }
  if (!PlanBlocks<IO>(CandidateBlockParameter(c.a, c.b), n, ops, r.plan)) {
    return false;
  }
  p.max_fused_size = 2;
  const std::vector<unsigned> no_splits;
  for (const auto& s : r.plan.steps) {
    if (s.blocked) {
      const BlockPass& b = r.plan.blocks[s.index];
      add([&](float* f) { BlockFeatures(n, b.tile, b.ks.size(), f); });
      ++r.passes;
      continue;
    }
    // An operation outside the blocked passes, as qsim's fuser hands it
    // back on its own.
    const auto f1 = MultiQubitGateFuser<IO>::template FuseGates<Operation<FP>>(
        p, n, ops.cbegin() + s.index, ops.cbegin() + s.index + 1, no_splits,
        true);
    for (const auto& e : f1) {
      add([&](float* f) { FusedOpFeatures<FP>(n, e, c.method, f); });
      ++r.passes;
    }
  }
  return true;
}

// Scores every candidate with the model (one forward call over all passes)
// and returns the index of the least predicted time, or -1 if none could
// be planned or the model call failed (*rc then holds its status).
template <typename IO, typename FP>
int ScoreCandidates(void* model, unsigned n,
                    const std::vector<Operation<FP>>& ops,
                    const std::vector<PlanCandidate>& cands,
                    std::vector<CandidateResult>& out, int* rc = nullptr) {
  if (rc) *rc = QSIM_HIP_OK;
  out.assign(cands.size(), CandidateResult{});
  std::vector<std::vector<float>> rows(cands.size());
  std::vector<std::future<bool>> jobs;
  for (size_t j = 0; j < cands.size(); ++j) {
    // Deferred (run on this thread) where no thread can be started.
    jobs.push_back(std::async(std::launch::async | std::launch::deferred,
                              [&, j] {
      return CandidateRows<IO>(cands[j], n, ops, rows[j], out[j]);
    }));
  }
  std::vector<float> feats;
  std::vector<size_t> owner;
  for (size_t j = 0; j < cands.size(); ++j) {
    out[j].ok = jobs[j].get();
    if (!out[j].ok) continue;
    feats.insert(feats.end(), rows[j].begin(), rows[j].end());
    owner.insert(owner.end(), rows[j].size() / kCostFeatures, j);
  }
  if (model == nullptr || owner.empty()) return -1;
  std::vector<float> y(owner.size());
  const int r = qsim_hip_costmodel_forward(model, feats.data(), owner.size(),
                                           y.data());
  if (r != QSIM_HIP_OK) {
    if (rc) *rc = r;
    return -1;
  }
  for (auto& c : out)
    if (c.ok) c.predicted_ms = 0.0;
  for (size_t i = 0; i < owner.size(); ++i) {
    out[owner[i]].predicted_ms +=
        std::exp(double(y[i]) + CostTargetOffset(n));
  }
  int best = -1;
  for (size_t j = 0; j < cands.size(); ++j) {
    if (out[j].ok && (best < 0 || out[j].predicted_ms < out[best].predicted_ms))
      best = int(j);
  }
  return best;
}

}  // namespace hip

template <typename IO>
class LearnedFuserHIP {
 public:
  struct Parameter {
    // The fallback: qsim's fuser at this size (no model, small circuits).
    unsigned max_fused_size = 4;
    unsigned verbosity = 0;
    // Cost-model network (qsim_hip_costmodel_*); nullptr: the library's
    // default model (qsim_hip_costmodel_default), if any.
    void* cost_model = nullptr;
    // > 0: no model; blocked passes of tile_bits tile qubits over the
    // low_bits lowest qubits.
    unsigned tile_bits = 0;
    unsigned low_bits = 4;
  };

  // Circuits below this many qubits take the fallback.
  static constexpr unsigned kMinLearnedQubits = 24;

  template <typename Operation>
  static auto FuseGates(const Parameter& param, unsigned max_qubit1,
                        const std::vector<Operation>& ops,
                        bool fuse_matrix = true) {
    return FuseGates<Operation>(param, max_qubit1, ops.cbegin(), ops.cend(),
                                {}, fuse_matrix);
  }

  template <typename Operation>
  static auto FuseGates(const Parameter& param, unsigned max_qubit1,
                        const std::vector<Operation>& ops,
                        const std::vector<unsigned>& times_to_split_at,
                        bool fuse_matrix = true) {
    return FuseGates<Operation>(param, max_qubit1, ops.cbegin(), ops.cend(),
                                times_to_split_at, fuse_matrix);
  }

  template <typename Operation>
  static auto FuseGates(const Parameter& param, unsigned max_qubit1,
                        typename std::vector<Operation>::const_iterator obeg,
                        typename std::vector<Operation>::const_iterator oend,
                        bool fuse_matrix = true) {
    return FuseGates<Operation>(param, max_qubit1, obeg, oend, {},
                                fuse_matrix);
  }

  template <typename OperationP>
  static auto FuseGates(const Parameter& param, unsigned max_qubit1,
                        typename std::vector<OperationP>::const_iterator obeg,
                        typename std::vector<OperationP>::const_iterator oend,
                        const std::vector<unsigned>& times_to_split_at,
                        bool fuse_matrix = true) {
    using Op = std::remove_const_t<
        std::remove_pointer_t<std::remove_const_t<OperationP>>>;
    using fp_type = OpFpType<Op>;
    using Fused = qsim::FusedGate<fp_type>;
    typename MultiQubitGateFuser<IO>::Parameter mp;
    mp.max_fused_size = param.max_fused_size;
    mp.verbosity = param.verbosity;
    auto fallback = [&]() {
      qsim_hip_set_gate_queue(0, 0);
      return MultiQubitGateFuser<IO>::template FuseGates<OperationP>(
          mp, max_qubit1, obeg, oend, times_to_split_at, fuse_matrix);
    };
    using Result = decltype(fallback());
    if constexpr (!std::is_same_v<Op, qsim::Operation<fp_type>> ||
                  std::is_pointer_v<OperationP>) {
      return fallback();
    } else {
      if (!times_to_split_at.empty() || !fuse_matrix || obeg >= oend) {
        return fallback();
      }
      auto deref = [](const auto& e) -> const Op& {
        if constexpr (std::is_pointer_v<std::decay_t<decltype(e)>>) {
          return *e;
        } else {
          return e;
        }
      };
      std::vector<Op> ops;
      ops.reserve(oend - obeg);
      for (auto it = obeg; it != oend; ++it) ops.push_back(deref(*it));

      hip::PlanCandidate pick{1, param.tile_bits, param.low_bits,
                              QSIM_HIP_METHOD_AUTO};
      hip::BlockPlan plan;
      bool planned = false;
      if (param.tile_bits == 0) {
        void* model = param.cost_model != nullptr
                          ? param.cost_model
                          : qsim_hip_costmodel_default();
        if (model == nullptr || max_qubit1 < kMinLearnedQubits) {
          return fallback();
        }
        std::vector<hip::PlanCandidate> cands;
        for (const auto& c : hip::DefaultCandidates())
          if (c.method == QSIM_HIP_METHOD_AUTO) cands.push_back(c);
        std::vector<hip::CandidateResult> res;
        const int best =
            hip::ScoreCandidates<IO>(model, max_qubit1, ops, cands, res);
        if (best < 0) return fallback();
        pick = cands[best];
        if (param.verbosity > 0) {
          IO::messagef("learned fuser: %s %u/%u, %u passes, %.3f ms "
                       "predicted\n",
                       pick.blocked ? "blocked" : "fuser", pick.a, pick.b,
                       res[best].passes, res[best].predicted_ms);
        }
        if (!pick.blocked) {
          mp.max_fused_size = pick.a;
          qsim_hip_set_gate_queue(0, 0);
          return MultiQubitGateFuser<IO>::template FuseGates<OperationP>(
              mp, max_qubit1, obeg, oend, times_to_split_at, fuse_matrix);
        }
        plan = std::move(res[best].plan);
        planned = true;
      }
      if (!planned &&
          !hip::PlanBlocks<IO>(
              hip::CandidateBlockParameter(pick.a, pick.b), max_qubit1, ops,
              plan)) {
        return fallback();
      }
      if (qsim_hip_set_gate_queue(pick.a, pick.b) != QSIM_HIP_OK) {
        return fallback();
      }

      // The plan's fused gates pass by pass; operations outside blocked
      // passes as qsim's fuser returns them on their own.
      Result out;
      auto gate_at = [&](size_t i) {
        return OpGetAlternative<Gate<fp_type>>(deref(*(obeg + i)));
      };
      typename MultiQubitGateFuser<IO>::Parameter one;
      one.max_fused_size = 2;
      one.verbosity = 0;
      const std::vector<unsigned> no_splits;
      for (const auto& s : plan.steps) {
        if (!s.blocked) {
          auto f1 = MultiQubitGateFuser<IO>::template FuseGates<OperationP>(
              one, max_qubit1, obeg + s.index, obeg + s.index + 1, no_splits,
              true);
          for (auto& e : f1) out.push_back(std::move(e));
          continue;
        }
        const hip::BlockPass& b = plan.blocks[s.index];
        size_t off = 0;
        for (size_t g = 0; g < b.ks.size(); ++g) {
          const unsigned k = b.ks[g];
          const size_t len = size_t{2} << (2 * k);
          Fused fg{};
          fg.qubits.assign(b.qs.begin() + 4 * g, b.qs.begin() + 4 * g + k);
          fg.matrix.assign(b.m.begin() + off, b.m.begin() + off + len);
          off += len;
          for (size_t i : b.comps[g]) {
            const Gate<fp_type>* gp = gate_at(i);
            if (gp == nullptr) continue;
            if (fg.gates.empty()) {
              fg.parent = gp;
              fg.kind = gp->kind;
            }
            fg.time = std::max(fg.time, gp->time);
            fg.gates.push_back(gp);
          }
          out.push_back(std::move(fg));
        }
      }
      return out;
    }
  }
};

}  // namespace qsim

#endif  // QSIM_AMD_LEARNED_FUSER_HIP_H_
