// driver.cc -- whole-circuit driver: qsim's own host code over the gfx950
// backend (C ABI in include/qsim_hip_driver.h).
//
// Parsing is qsim's CircuitQsimParser, the partition is qsim's
// MultiQubitGateFuser or the blocked-pass planner (qsim_amd/block_planner.h),
// and the apply loop is the one of QSimRunner::Run (lib/run_qsim.h:264-280)
// with gate_appl.h's dispatch, on SimulatorHIP / StateSpaceHIP.  Compiled
// against qsim's lib/ headers.

#include <cstdint>
#include <cstring>
#include <memory>
#include <new>
#include <random>
#include <sstream>
#include <string>
#include <variant>
#include <vector>

#include "circuit_qsim_parser.h"
#include "fuser_mqubit.h"
#include "gate_appl.h"
#include "io.h"
#include "qsim_amd/block_planner.h"
#include "qsim_amd/learned_fuser_hip.h"
#include "qsim_amd/learned_planner.h"
#include "qsim_amd/simulator_hip.h"
#include "qsim_hip.h"
#include "qsim_hip_driver.h"

namespace {

using namespace qsim;

// Cost model used by qsim_driver_fuse(..., QSIM_DRIVER_FUSER_LEARNED, ...).
void* g_cost_model = nullptr;

template <typename FP>
using FusedOps =
    std::vector<std::variant<FusedGate<FP>, Measurement, const Operation<FP>*>>;

// A pass of the current partition: entry `index` of f32 / f64 (a fused
// gate, a measurement or an unfused operation), or blocked pass `index`.
struct Step {
  bool blocked;
  size_t index;
};

struct DriverCircuit {
  Circuit<Operation<float>> c32;
  Circuit<Operation<double>> c64;
  int matrix_prec = QSIM_HIP_F32;
  FusedOps<float> f32;
  FusedOps<double> f64;
  std::vector<hip::BlockPass> blocks;
  std::vector<Step> steps;
  // Gate method the partition runs with (-1: the library's current one).
  int method = -1;
  // Arithmetic of the blocked passes (QSIM_DRIVER_BLOCK_*).
  int block_arith = QSIM_DRIVER_BLOCK_F16;
  // Set while qsim_driver_capture records and captures the passes.
  bool capturing = false;
  // Prepared blocked passes (qsim_hip_block_prepare), one per block, made at
  // the first run.
  std::vector<void*> prepared;

  void ReleasePrepared() {
    for (void* h : prepared) qsim_hip_block_free(h);
    prepared.clear();
  }
  ~DriverCircuit() { ReleasePrepared(); }
};

void PlainSteps(DriverCircuit& d) {
  d.ReleasePrepared();
  d.blocks.clear();
  d.steps.clear();
  for (size_t i = 0; i < d.f32.size(); ++i) d.steps.push_back({false, i});
}

template <typename FP>
bool Parse(const std::string& text, unsigned maxtime,
           Circuit<Operation<FP>>& c) {
  std::stringstream ss(text);
  return CircuitQsimParser<IO>::FromStream(maxtime, "circuit", ss, c);
}

template <typename FP>
int Fuse(const Circuit<Operation<FP>>& c, unsigned max_fused, int fuser,
         FusedOps<FP>& out) {
  if (fuser == QSIM_DRIVER_FUSER_MQUBIT) {
    if (max_fused < 1 || max_fused > 6) return QSIM_HIP_ERR_ARG;
    MultiQubitGateFuser<IO>::Parameter p;
    p.max_fused_size = max_fused;
    p.verbosity = 0;
    out = MultiQubitGateFuser<IO>::FuseGates(p, c.num_qubits, c.ops);
  } else if (fuser == QSIM_DRIVER_FUSER_NONE) {
    // Each operation as qsim's fuser hands back an unfused one: a pointer
    // into the circuit, applied by ApplyGate (lib/gate_appl.h:30-46).
    out.clear();
    out.reserve(c.ops.size());
    for (const auto& op : c.ops) out.push_back(&op);
  } else {
    return QSIM_HIP_ERR_ARG;
  }
  if (out.empty() && !c.ops.empty()) return QSIM_HIP_ERR_ARG;
  return QSIM_HIP_OK;
}

template <typename FP, typename FusedOp>
void PassInfo(const FusedOp& op, unsigned* k, unsigned* qubits,
              unsigned* num_gates) {
  const std::vector<unsigned>* qs = nullptr;
  unsigned ng = 1;
  if (const auto* fg = std::get_if<FusedGate<FP>>(&op)) {
    qs = &fg->qubits;
    ng = fg->gates.size();
  } else if (const auto* m = std::get_if<Measurement>(&op)) {
    qs = &m->qubits;
  } else {
    qs = &OpQubits(*std::get<const Operation<FP>*>(op));
  }
  *k = qs->size();
  if (qubits) {
    for (unsigned i = 0; i < qs->size() && i < 64; ++i) qubits[i] = (*qs)[i];
  }
  if (num_gates) *num_gates = ng;
}

// Blocked pass `index`: prepared (fragments on the device) at its first run,
// one kernel launch per run after that.
int RunBlock(DriverCircuit& d, size_t index, float* p, unsigned n) {
  // Split-BF16: the circuit's first blocked pass runs on a structured state
  // (a basis state and its first gates), where the two-part amplitude split
  // would move the norm coherently, so it uses the accurate product.
  const int mprec =
      d.block_arith == QSIM_DRIVER_BLOCK_BF16
          ? QSIM_HIP_F64 | int(QSIM_HIP_BLOCK_BF16) |
                (index == 0 ? int(QSIM_HIP_BLOCK_ACCURATE) : 0)
          : QSIM_HIP_F64;
  if (d.capturing) {
    // A captured graph owns its data: the fragments go through the
    // library's staging (the graph plan's buffers), not the circuit's
    // prepared handles, which a later fuse may release.
    const hip::BlockPass& b = d.blocks[index];
    return qsim_hip_apply_gate_block(QSIM_HIP_F32, p, n, b.tile.data(),
                                     b.tile.size(), b.ks.size(), b.ks.data(),
                                     b.qs.data(), b.m.data(), mprec);
  }
  if (d.prepared.size() != d.blocks.size()) {
    d.ReleasePrepared();
    d.prepared.assign(d.blocks.size(), nullptr);
  }
  if (d.prepared[index] == nullptr) {
    const hip::BlockPass& b = d.blocks[index];
    int rc = qsim_hip_block_prepare(n, b.tile.data(), b.tile.size(),
                                    b.ks.size(), b.ks.data(), b.qs.data(),
                                    b.m.data(), mprec, &d.prepared[index]);
    if (rc != QSIM_HIP_OK) return rc;
  }
  return qsim_hip_block_run(d.prepared[index], QSIM_HIP_F32, p, n);
}

// Complex64 state, FP32 matrices: exactly QSimRunner's loop.
int RunF32(DriverCircuit& d, float* p, unsigned first, unsigned count,
           unsigned seed) {
  StateSpaceHIP<float> ss;
  SimulatorHIP<float> sim;
  auto state = ss.Create(p, d.c32.num_qubits);
  std::mt19937 rgen(seed);
  for (unsigned i = first; i < first + count; ++i) {
    const Step& st = d.steps[i];
    if (st.blocked) {
      int rc = RunBlock(d, st.index, p, d.c32.num_qubits);
      if (rc != QSIM_HIP_OK) return rc;
    } else if (!ApplyGate(ss, sim, d.f32[st.index], rgen, state)) {
      return QSIM_HIP_ERR_ARG;
    }
  }
  return QSIM_HIP_OK;
}

// Complex64 state, FP64 fused matrices.
int RunF32M64(DriverCircuit& d, float* p, unsigned first, unsigned count,
              unsigned seed) {
  StateSpaceHIP<float> ss;
  SimulatorHIP<float> sim;
  auto state = ss.Create(p, d.c64.num_qubits);
  std::mt19937 rgen(seed);
  const unsigned n = d.c64.num_qubits;
  for (unsigned i = first; i < first + count; ++i) {
    const Step& st = d.steps[i];
    if (st.blocked) {
      int rc = RunBlock(d, st.index, p, n);
      if (rc != QSIM_HIP_OK) return rc;
      continue;
    }
    const auto& op = d.f64[st.index];
    if (const auto* fg = std::get_if<FusedGate<double>>(&op)) {
      sim.ApplyGate(fg->qubits, fg->matrix.data(), state);
    } else if (const auto* m = std::get_if<Measurement>(&op)) {
      auto r = ss.Measure(m->qubits, rgen, state);
      if (!r.valid) return QSIM_HIP_ERR_ARG;
    } else {
      // Unfusible operations: measurements (qsim's fuser hands them back as
      // pointers to the circuit's operations) and controlled gates.
      const Operation<double>* o = std::get<const Operation<double>*>(op);
      if (const auto* m = OpGetAlternative<Measurement>(*o)) {
        auto r = ss.Measure(m->qubits, rgen, state);
        if (!r.valid) return QSIM_HIP_ERR_ARG;
      } else if (const auto* g = OpGetAlternative<Gate<double>>(*o)) {
        sim.ApplyGate(g->qubits, g->matrix.data(), state);
      } else if (const auto* cg = OpGetAlternative<ControlledGate<double>>(*o)) {
        int rc = qsim_hip_apply_controlled_gate(
            QSIM_HIP_F32, p, n, cg->qubits.data(), cg->qubits.size(),
            cg->controlled_by.data(), cg->controlled_by.size(), cg->cmask,
            cg->matrix.data(), QSIM_HIP_F64);
        if (rc != QSIM_HIP_OK && rc != QSIM_HIP_ERR_UNSUPPORTED) return rc;
      }
    }
  }
  return QSIM_HIP_OK;
}

// Complex128 state.  Blocked passes run gate by gate (FP64 arithmetic).
int RunF64(DriverCircuit& d, double* p, unsigned first, unsigned count,
           unsigned seed) {
  StateSpaceHIP<double> ss;
  SimulatorHIP<double> sim;
  auto state = ss.Create(p, d.c64.num_qubits);
  std::mt19937 rgen(seed);
  const unsigned n = d.c64.num_qubits;
  for (unsigned i = first; i < first + count; ++i) {
    const Step& st = d.steps[i];
    if (st.blocked) {
      const hip::BlockPass& b = d.blocks[st.index];
      size_t off = 0;
      for (size_t g = 0; g < b.ks.size(); ++g) {
        int rc = qsim_hip_apply_gate(QSIM_HIP_F64, p, n, &b.qs[4 * g],
                                     b.ks[g], &b.m[off], QSIM_HIP_F64);
        if (rc != QSIM_HIP_OK) return rc;
        off += size_t{2} << (2 * b.ks[g]);
      }
    } else if (!ApplyGate(ss, sim, d.f64[st.index], rgen, state)) {
      return QSIM_HIP_ERR_ARG;
    }
  }
  return QSIM_HIP_OK;
}

// *trial: in, the planner restart to use (-1: search); out, the one used.
template <typename FP>
int PlanBlocked(const Circuit<Operation<FP>>& c, unsigned tile_bits,
                unsigned low_bits, std::vector<hip::BlockPass>& blocks,
                std::vector<Step>& steps, FusedOps<FP>& single, int* trial) {
  hip::BlockPlannerParameter bp =
      hip::CandidateBlockParameter(tile_bits, low_bits);
  bp.trial = *trial;
  hip::BlockPlan plan;
  if (!hip::PlanBlocks<IO>(bp, c.num_qubits, c.ops, plan)) {
    return QSIM_HIP_ERR_UNSUPPORTED;
  }
  *trial = int(plan.trial);
  blocks = std::move(plan.blocks);
  steps.clear();
  single.clear();
  // Operations outside blocked passes go through qsim's fuser one by one,
  // which turns them into the entries its apply loop expects.
  MultiQubitGateFuser<IO>::Parameter p;
  p.max_fused_size = 2;
  p.verbosity = 0;
  for (const auto& s : plan.steps) {
    if (s.blocked) {
      steps.push_back({true, s.index});
      continue;
    }
    const std::vector<unsigned> no_splits;
    auto f = MultiQubitGateFuser<IO>::template FuseGates<Operation<FP>>(
        p, c.num_qubits, c.ops.cbegin() + s.index,
        c.ops.cbegin() + s.index + 1, no_splits, true);
    for (auto& e : f) {
      steps.push_back({false, single.size()});
      single.push_back(std::move(e));
    }
  }
  return QSIM_HIP_OK;
}

// Cost-model features of step i when run with gate method `method`.
void PassFeatures(const DriverCircuit& d, unsigned i, int method, float* f) {
  const unsigned n = d.c32.num_qubits;
  const Step& st = d.steps[i];
  if (st.blocked) {
    const hip::BlockPass& b = d.blocks[st.index];
    hip::BlockFeatures(n, b.tile, b.ks.size(), f);
    return;
  }
  hip::FusedOpFeatures<float>(n, d.f32[st.index], method, f);
}

int FuseBlocked(DriverCircuit& d, unsigned tile_bits, unsigned low_bits,
                int trial);

// trial >= 0: a blocked candidate's restart, as its scoring chose it.
int ApplyCandidate(DriverCircuit& d, const hip::PlanCandidate& c,
                   int trial = -1) {
  int rc = c.blocked ? FuseBlocked(d, c.a, c.b, trial)
                     : qsim_driver_fuse(&d, c.a, QSIM_DRIVER_FUSER_MQUBIT,
                                        QSIM_HIP_F64);
  d.method = rc == QSIM_HIP_OK ? c.method : -1;
  return rc;
}

int FuseBlocked(DriverCircuit& d, unsigned tile_bits, unsigned low_bits,
                int trial) {
  if (tile_bits < 9 || tile_bits > 14 || low_bits < 3 ||
      low_bits > tile_bits) {
    return QSIM_HIP_ERR_ARG;
  }
  d.method = -1;
  d.matrix_prec = QSIM_HIP_F64;
  d.ReleasePrepared();
  std::vector<hip::BlockPass> b32;
  std::vector<Step> s32;
  // The FP32 copy of the circuit has the same structure: it takes the
  // restart the FP64 search chose.
  int rc = PlanBlocked(d.c64, tile_bits, low_bits, d.blocks, d.steps, d.f64,
                       &trial);
  if (rc == QSIM_HIP_OK) {
    rc = PlanBlocked(d.c32, tile_bits, low_bits, b32, s32, d.f32, &trial);
  }
  if (rc == QSIM_HIP_OK && s32.size() != d.steps.size()) {
    rc = QSIM_HIP_ERR_RUNTIME;
  }
  if (rc != QSIM_HIP_OK) {
    d.f32.clear();
    d.f64.clear();
    PlainSteps(d);
  }
  return rc;
}

// Runs an entry point's body: a failed host allocation becomes an error
// return instead of an exception crossing the C boundary.
template <typename F>
int GuardHost(F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return QSIM_HIP_ERR_RUNTIME;
  }
}

}  // namespace

extern "C" {

void* qsim_driver_load(const char* text, unsigned maxtime) {
  try {
    auto d = std::make_unique<DriverCircuit>();
    std::string t(text);
    if (!Parse<float>(t, maxtime, d->c32)) return nullptr;
    if (!Parse<double>(t, maxtime, d->c64)) return nullptr;
    return d.release();
  } catch (const std::bad_alloc&) {
    return nullptr;
  }
}

void qsim_driver_free(void* c) { delete static_cast<DriverCircuit*>(c); }

unsigned qsim_driver_num_qubits(void* c) {
  return static_cast<DriverCircuit*>(c)->c32.num_qubits;
}

unsigned qsim_driver_num_gates(void* c) {
  return static_cast<DriverCircuit*>(c)->c32.ops.size();
}

void qsim_driver_set_cost_model(void* model) { g_cost_model = model; }

int qsim_driver_fuse(void* c, unsigned max_fused_size, int fuser,
                     int matrix_prec) {
  return GuardHost([&]() -> int {
    auto* d = static_cast<DriverCircuit*>(c);
    d->method = -1;
    if (fuser == QSIM_DRIVER_FUSER_LEARNED) {
      return qsim_driver_plan_learned(c, g_cost_model, 0, nullptr, nullptr,
                                      nullptr, nullptr, nullptr, nullptr,
                                      nullptr, nullptr);
    }
    d->matrix_prec = matrix_prec;
    int rc = Fuse(d->c32, max_fused_size, fuser, d->f32);
    if (rc == QSIM_HIP_OK) rc = Fuse(d->c64, max_fused_size, fuser, d->f64);
    if (rc != QSIM_HIP_OK) {
      d->f32.clear();
      d->f64.clear();
    }
    PlainSteps(*d);
    return rc;
  });
}

int qsim_driver_fuse_blocked(void* c, unsigned tile_bits, unsigned low_bits) {
  return GuardHost([&]() -> int {
    return FuseBlocked(*static_cast<DriverCircuit*>(c), tile_bits, low_bits, -1);
  });
}

unsigned qsim_driver_num_passes(void* c) {
  return static_cast<DriverCircuit*>(c)->steps.size();
}

int qsim_driver_pass_kind(void* c, unsigned i) {
  auto* d = static_cast<DriverCircuit*>(c);
  if (i >= d->steps.size()) return QSIM_HIP_ERR_ARG;
  const Step& st = d->steps[i];
  if (st.blocked) return QSIM_DRIVER_PASS_BLOCKED;
  const auto& op = d->f32[st.index];
  if (std::get_if<FusedGate<float>>(&op)) return QSIM_DRIVER_PASS_GATE;
  if (std::get_if<Measurement>(&op)) return QSIM_DRIVER_PASS_MEASUREMENT;
  // Unfusible operations (measurements included) come back as pointers to
  // the circuit's own operations.
  const Operation<float>* o = std::get<const Operation<float>*>(op);
  if (OpGetAlternative<Measurement>(*o)) return QSIM_DRIVER_PASS_MEASUREMENT;
  return QSIM_DRIVER_PASS_OPERATION;
}

int qsim_driver_pass_info(void* c, unsigned i, unsigned* k, unsigned* qubits,
                          unsigned* num_gates) {
  auto* d = static_cast<DriverCircuit*>(c);
  if (i >= d->steps.size()) return QSIM_HIP_ERR_ARG;
  const Step& st = d->steps[i];
  if (st.blocked) {
    const hip::BlockPass& b = d->blocks[st.index];
    *k = b.tile.size();
    if (qubits) {
      for (size_t j = 0; j < b.tile.size(); ++j) qubits[j] = b.tile[j];
    }
    if (num_gates) *num_gates = b.num_gates;
    return QSIM_HIP_OK;
  }
  PassInfo<float>(d->f32[st.index], k, qubits, num_gates);
  return QSIM_HIP_OK;
}

int qsim_driver_block_info(void* c, unsigned i, unsigned* num_ops,
                           unsigned* op_k, unsigned* op_qubits,
                           double* op_matrices) {
  auto* d = static_cast<DriverCircuit*>(c);
  if (i >= d->steps.size() || !d->steps[i].blocked) return QSIM_HIP_ERR_ARG;
  const hip::BlockPass& b = d->blocks[d->steps[i].index];
  *num_ops = b.ks.size();
  for (size_t g = 0; g < b.ks.size(); ++g) {
    if (op_k) op_k[g] = b.ks[g];
    if (op_qubits) {
      for (int j = 0; j < 4; ++j) op_qubits[4 * g + j] = b.qs[4 * g + j];
    }
  }
  if (op_matrices) std::memcpy(op_matrices, b.m.data(), b.m.size() * 8);
  return QSIM_HIP_OK;
}

int qsim_driver_run_passes(void* c, void* state, int prec, unsigned first,
                           unsigned count, unsigned seed) {
  return GuardHost([&]() -> int {
    auto* d = static_cast<DriverCircuit*>(c);
    if (state == nullptr || first > d->steps.size() ||
        count > d->steps.size() - first || d->f32.size() != d->f64.size()) {
      return QSIM_HIP_ERR_ARG;
    }
    if (prec == QSIM_HIP_F64) {
      return RunF64(*d, static_cast<double*>(state), first, count, seed);
    }
    const int old_method = qsim_hip_get_gate_method();
    if (d->method >= 0 && d->method != old_method) {
      int rc = qsim_hip_set_gate_method(d->method);
      if (rc != QSIM_HIP_OK) return rc;
    }
    int rc = d->matrix_prec == QSIM_HIP_F64
                 ? RunF32M64(*d, static_cast<float*>(state), first, count, seed)
                 : RunF32(*d, static_cast<float*>(state), first, count, seed);
    if (d->method >= 0 && d->method != old_method) {
      qsim_hip_set_gate_method(old_method);
    }
    return rc;
  });
}

int qsim_driver_set_method(void* c, int method) {
  auto* d = static_cast<DriverCircuit*>(c);
  if (method < -1 || method > QSIM_HIP_METHOD_MFMA_FP16X2) {
    return QSIM_HIP_ERR_ARG;
  }
  d->method = method;
  return QSIM_HIP_OK;
}

int qsim_driver_set_block_arith(void* c, int arith) {
  auto* d = static_cast<DriverCircuit*>(c);
  if (arith != QSIM_DRIVER_BLOCK_F16 && arith != QSIM_DRIVER_BLOCK_BF16) {
    return QSIM_HIP_ERR_ARG;
  }
  if (arith != d->block_arith) d->ReleasePrepared();
  d->block_arith = arith;
  return QSIM_HIP_OK;
}

int qsim_driver_pass_features(void* c, unsigned i, int method, float* out) {
  return GuardHost([&]() -> int {
    auto* d = static_cast<DriverCircuit*>(c);
    if (i >= d->steps.size() || out == nullptr) return QSIM_HIP_ERR_ARG;
    PassFeatures(*d, i, method, out);
    return QSIM_HIP_OK;
  });
}

int qsim_driver_plan_learned(void* c, void* model, unsigned max_candidates,
                             int* blocked, unsigned* a, unsigned* b,
                             int* methods, double* predicted_ms,
                             unsigned* num_passes, unsigned* num_candidates,
                             int* chosen) {
  return GuardHost([&]() -> int {
    auto* d = static_cast<DriverCircuit*>(c);
    const unsigned n = d->c64.num_qubits;
    const std::vector<hip::PlanCandidate> cands = hip::DefaultCandidates();
    std::vector<hip::CandidateResult> res(cands.size());
    int best = -1;
    if (model != nullptr) {
      int rc = QSIM_HIP_OK;
      best = hip::ScoreCandidates<IO>(model, n, d->c64.ops, cands, res, &rc);
      if (rc != QSIM_HIP_OK) return rc;
    }
    const hip::PlanCandidate pick =
        best >= 0 ? cands[best] : hip::FallbackCandidate();
    const int trial =
        best >= 0 && pick.blocked ? int(res[best].plan.trial) : -1;
    int rc = ApplyCandidate(*d, pick, trial);
    if (rc != QSIM_HIP_OK) return rc;
    if (best < 0) {
      // No model: the fallback, reported by its index in the list.
      for (size_t j = 0; j < cands.size(); ++j) {
        if (!cands[j].blocked && cands[j].a == pick.a &&
            cands[j].method == pick.method) {
          best = int(j);
        }
      }
    }
    const unsigned m = std::min<unsigned>(max_candidates, cands.size());
    for (unsigned j = 0; j < m; ++j) {
      if (blocked) blocked[j] = cands[j].blocked;
      if (a) a[j] = cands[j].a;
      if (b) b[j] = cands[j].b;
      if (methods) methods[j] = cands[j].method;
      if (predicted_ms) predicted_ms[j] = res[j].predicted_ms;
      if (num_passes) num_passes[j] = res[j].passes;
    }
    if (num_candidates) *num_candidates = cands.size();
    if (chosen) *chosen = best;
    return QSIM_HIP_OK;
  });
}

int qsim_driver_run(void* c, void* state, int prec, unsigned seed) {
  return qsim_driver_run_passes(c, state, prec, 0, qsim_driver_num_passes(c),
                                seed);
}

int qsim_driver_capture(void* c, void* state, int prec, void** plan) {
  return GuardHost([&]() -> int {
    if (c == nullptr || state == nullptr || plan == nullptr) {
      return QSIM_HIP_ERR_ARG;
    }
    *plan = nullptr;
    const unsigned np = qsim_driver_num_passes(c);
    for (unsigned i = 0; i < np; ++i) {
      if (qsim_driver_pass_kind(c, i) == QSIM_DRIVER_PASS_MEASUREMENT) {
        return QSIM_HIP_ERR_UNSUPPORTED;
      }
    }
    auto* d = static_cast<DriverCircuit*>(c);
    void* p = nullptr;
    int rc = qsim_hip_graph_record_begin(&p);
    if (rc != QSIM_HIP_OK) return rc;
    d->capturing = true;
    rc = qsim_driver_run(c, state, prec, 0);
    int rc2 = qsim_hip_graph_record_end(p);
    if (rc == QSIM_HIP_OK) rc = rc2;
    if (rc == QSIM_HIP_OK) rc = qsim_hip_graph_capture_begin(p);
    if (rc == QSIM_HIP_OK) {
      rc = qsim_driver_run(c, state, prec, 0);
      rc2 = qsim_hip_graph_capture_end(p);
      if (rc == QSIM_HIP_OK) rc = rc2;
    }
    d->capturing = false;
    if (rc != QSIM_HIP_OK) {
      qsim_hip_graph_free(p);
      return rc;
    }
    *plan = p;
    return QSIM_HIP_OK;
  });
}

}  // extern "C"
