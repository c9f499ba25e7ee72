// driver.cc -- whole-circuit driver: qsim's own host code over the gfx950
// backend (C ABI in include/qsim_hip_driver.h).
//
// Parsing is qsim's CircuitQsimParser, the partition is qsim's
// MultiQubitGateFuser (or the cost-model fuser), and the apply loop is the one
// of QSimRunner::Run (lib/run_qsim.h:264-280) with gate_appl.h's dispatch, on
// SimulatorHIP / StateSpaceHIP.  Compiled against qsim's lib/ headers.

#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <variant>
#include <vector>

#include "circuit_qsim_parser.h"
#include "fuser_mqubit.h"
#include "gate_appl.h"
#include "io.h"
#include "qsim_amd/simulator_hip.h"
#include "qsim_hip.h"
#include "qsim_hip_driver.h"

namespace {

using namespace qsim;

template <typename FP>
using FusedOps =
    std::vector<std::variant<FusedGate<FP>, Measurement, const Operation<FP>*>>;

struct DriverCircuit {
  Circuit<Operation<float>> c32;
  Circuit<Operation<double>> c64;
  int matrix_prec = QSIM_HIP_F32;
  FusedOps<float> f32;
  FusedOps<double> f64;
};

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
    MultiQubitGateFuser<IO>::Parameter p;
    p.max_fused_size = max_fused;
    p.verbosity = 0;
    out = MultiQubitGateFuser<IO>::FuseGates(p, c.num_qubits, c.ops);
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

// Complex64 state, FP32 matrices: exactly QSimRunner's loop.
int RunF32(DriverCircuit& d, float* p, unsigned first, unsigned count,
           unsigned seed) {
  StateSpaceHIP<float> ss;
  SimulatorHIP<float> sim;
  auto state = ss.Create(p, d.c32.num_qubits);
  std::mt19937 rgen(seed);
  for (unsigned i = first; i < first + count; ++i) {
    if (!ApplyGate(ss, sim, d.f32[i], rgen, state)) return QSIM_HIP_ERR_ARG;
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
    const auto& op = d.f64[i];
    if (const auto* fg = std::get_if<FusedGate<double>>(&op)) {
      sim.ApplyGate(fg->qubits, fg->matrix.data(), state);
    } else if (const auto* m = std::get_if<Measurement>(&op)) {
      auto r = ss.Measure(m->qubits, rgen, state);
      if (!r.valid) return QSIM_HIP_ERR_ARG;
    } else {
      const Operation<double>* o = std::get<const Operation<double>*>(op);
      if (const auto* g = OpGetAlternative<Gate<double>>(*o)) {
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

// Complex128 state.
int RunF64(DriverCircuit& d, double* p, unsigned first, unsigned count,
           unsigned seed) {
  StateSpaceHIP<double> ss;
  SimulatorHIP<double> sim;
  auto state = ss.Create(p, d.c64.num_qubits);
  std::mt19937 rgen(seed);
  for (unsigned i = first; i < first + count; ++i) {
    if (!ApplyGate(ss, sim, d.f64[i], rgen, state)) return QSIM_HIP_ERR_ARG;
  }
  return QSIM_HIP_OK;
}

}  // namespace

extern "C" {

void* qsim_driver_load(const char* text, unsigned maxtime) {
  auto d = std::make_unique<DriverCircuit>();
  std::string t(text);
  if (!Parse<float>(t, maxtime, d->c32)) return nullptr;
  if (!Parse<double>(t, maxtime, d->c64)) return nullptr;
  return d.release();
}

void qsim_driver_free(void* c) { delete static_cast<DriverCircuit*>(c); }

unsigned qsim_driver_num_qubits(void* c) {
  return static_cast<DriverCircuit*>(c)->c32.num_qubits;
}

unsigned qsim_driver_num_gates(void* c) {
  return static_cast<DriverCircuit*>(c)->c32.ops.size();
}

int qsim_driver_fuse(void* c, unsigned max_fused_size, int fuser,
                     int matrix_prec) {
  auto* d = static_cast<DriverCircuit*>(c);
  d->matrix_prec = matrix_prec;
  int rc = Fuse(d->c32, max_fused_size, fuser, d->f32);
  if (rc != QSIM_HIP_OK) return rc;
  return Fuse(d->c64, max_fused_size, fuser, d->f64);
}

unsigned qsim_driver_num_passes(void* c) {
  return static_cast<DriverCircuit*>(c)->f32.size();
}

int qsim_driver_pass_info(void* c, unsigned i, unsigned* k, unsigned* qubits,
                          unsigned* num_gates) {
  auto* d = static_cast<DriverCircuit*>(c);
  if (i >= d->f32.size()) return QSIM_HIP_ERR_ARG;
  PassInfo<float>(d->f32[i], k, qubits, num_gates);
  return QSIM_HIP_OK;
}

int qsim_driver_run_passes(void* c, void* state, int prec, unsigned first,
                           unsigned count, unsigned seed) {
  auto* d = static_cast<DriverCircuit*>(c);
  if (state == nullptr || first > d->f32.size() ||
      count > d->f32.size() - first || d->f32.size() != d->f64.size()) {
    return QSIM_HIP_ERR_ARG;
  }
  if (prec == QSIM_HIP_F64) {
    return RunF64(*d, static_cast<double*>(state), first, count, seed);
  }
  if (d->matrix_prec == QSIM_HIP_F64) {
    return RunF32M64(*d, static_cast<float*>(state), first, count, seed);
  }
  return RunF32(*d, static_cast<float*>(state), first, count, seed);
}

int qsim_driver_run(void* c, void* state, int prec, unsigned seed) {
  return qsim_driver_run_passes(c, state, prec, 0, qsim_driver_num_passes(c),
                                seed);
}

}  // extern "C"
