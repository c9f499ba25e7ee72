// qsim_hip_module.cc -- qsim's Python extension `qsim_hip` on the gfx950
// backend.
//
// qsim's HIP extension (pybind_interface/hip/pybind_main_hip.cpp:18-51) is a
// Factory over SimulatorCUDA<float> compiled into the backend-independent
// bindings of pybind_interface/pybind_main.{h,cpp}.  This file supplies the
// same Factory shape over this backend's SimulatorHIP / StateSpaceHIP and
// compiles the same bindings (read from the qsim sources at build time), so
// qsimcirq's `qsim_hip` import (`qsimcirq/__init__.py:34-42`) and every call
// it makes (`qsim_simulate`, `qsim_simulate_fullstate`, samplers, expectation
// values, qtrajectory and qsimh entry points) run here unchanged.

//
// The runner's fuser is LearnedFuserHIP (include/qsim_amd/learned_fuser_hip.h):
// for circuits of 24 qubits and more it picks the partition per circuit
// with the cost model committed beside the package
// (qsim_amd/data/cost_model.json, loaded as the library's default model
// when the module is imported), including cache-blocked passes; otherwise,
// and without the model file, it is qsim's MultiQubitGateFuser at the
// caller's max_fused_size, as in qsim's own extension.

#include <dlfcn.h>

#include <string>

#include "pybind_main.h"

#include "fuser_mqubit.h"
#include "gates_cirq.h"
#include "io.h"
#include "qsim_amd/factory_hip.h"
#include "qsim_amd/learned_fuser_hip.h"
#include "qtrajectory.h"
#include "run_qsim.h"

namespace {

// qsim_amd/lib/qsim_hip<ext> -> qsim_amd/data/cost_model.json.
void LoadDefaultCostModel() {
  Dl_info info{};
  if (dladdr(reinterpret_cast<void*>(&LoadDefaultCostModel), &info) == 0 ||
      info.dli_fname == nullptr) {
    return;
  }
  std::string path(info.dli_fname);
  const size_t slash = path.rfind('/');
  path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) +
         "/../data/cost_model.json";
  void* model = nullptr;
  if (qsim_hip_costmodel_default() == nullptr &&
      qsim_hip_costmodel_load(path.c_str(), &model) == QSIM_HIP_OK) {
    qsim_hip_costmodel_set_default(model);
  }
}

}  // namespace

PYBIND11_MODULE(qsim_hip, m) {
  GPU_MODULE_BINDINGS
  LoadDefaultCostModel();
}

namespace qsim {

struct Factory {
  // The CUDA extension reads "gsst" / "gdb" (CUDA launch shape); the gfx950
  // kernels size their launches themselves, so the options are accepted and
  // not used.
  explicit Factory(const py::dict&) {}

  using Simulator = SimulatorHIP<float>;
  using StateSpace = Simulator::StateSpace;

  using Gate = Cirq::GateCirq<float>;
  using Operation = qsim::Operation<float>;
  using Runner = QSimRunner<IO, LearnedFuserHIP<IO>, Factory>;
  using RunnerParameter = Runner::Parameter;
  using NoisyRunner = qsim::QuantumTrajectorySimulator<IO, Runner>;
  using NoisyRunnerParameter = NoisyRunner::Parameter;

  StateSpace CreateStateSpace() const { return StateSpace(); }
  Simulator CreateSimulator() const { return Simulator(); }
};

// Denormal control is a host-CPU setting of the CPU simulators.
inline void SetFlushToZeroAndDenormalsAreZeros() {}
inline void ClearFlushToZeroAndDenormalsAreZeros() {}

namespace detail {
// The bindings name the backend's state deleter `detail::free` (the CUDA
// backend's is lib/vectorspace_cuda.h:34-36); only the branch for states
// that do not need a host copy calls it, which a device state never takes.
inline void free(void* ptr) { hip_detail::free_device(ptr); }
}  // namespace detail

}  // namespace qsim

#include "pybind_main.cpp"
