// qsim_base_hip.cc -- command-line circuit runner on the gfx950 backend.
//
// The shape of qsim's apps/qsim_base_cuda.cu (:107-172): parse a qsim circuit
// file, run QSimRunner with MultiQubitGateFuser on FactoryHIP, print the first
// amplitudes.  Extra options write the normal-order state and samples.
//
//   qsim_base_hip -c FILE [-d maxtime] [-s seed] [-f max_fused_size]
//                 [-v verbosity] [-o STATE.bin] [-n NUM_SAMPLES -x SEED
//                 -b BITS.bin]

#include <unistd.h>

#include <algorithm>
#include <complex>
#include <cstdio>
#include <limits>
#include <string>
#include <vector>

#include "circuit_qsim_parser.h"
#include "fuser_mqubit.h"
#include "io_file.h"
#include "operation.h"
#include "qsim_amd/factory_hip.h"
#include "run_qsim.h"

int main(int argc, char* argv[]) {
  using namespace qsim;
  std::string file, state_out, bits_out;
  unsigned maxtime = std::numeric_limits<unsigned>::max();
  unsigned seed = 1, max_fused = 2, verbosity = 0, sample_seed = 0;
  uint64_t num_samples = 0;
  int k;
  while ((k = getopt(argc, argv, "c:d:s:f:v:o:n:x:b:")) != -1) {
    switch (k) {
      case 'c': file = optarg; break;
      case 'd': maxtime = std::atoi(optarg); break;
      case 's': seed = std::atoi(optarg); break;
      case 'f': max_fused = std::atoi(optarg); break;
      case 'v': verbosity = std::atoi(optarg); break;
      case 'o': state_out = optarg; break;
      case 'n': num_samples = std::strtoull(optarg, nullptr, 10); break;
      case 'x': sample_seed = std::atoi(optarg); break;
      case 'b': bits_out = optarg; break;
      default:
        IO::errorf("usage: qsim_base_hip -c circuit [-d maxtime] [-s seed] "
                   "[-f max_fused_size] [-v verbosity]\n");
        return 1;
    }
  }
  if (file.empty()) {
    IO::errorf("circuit file is not provided.\n");
    return 1;
  }

  Circuit<Operation<float>> circuit;
  if (!CircuitQsimParser<IOFile>::FromFile(maxtime, file, circuit)) return 1;

  using Factory = FactoryHIP<float>;
  using Runner = QSimRunner<IO, MultiQubitGateFuser<IO>, Factory>;
  Factory factory;
  auto state_space = factory.CreateStateSpace();
  auto state = state_space.Create(circuit.num_qubits);
  if (state_space.IsNull(state)) {
    IO::errorf("not enough memory: is the number of qubits too large?\n");
    return 1;
  }
  state_space.SetStateZero(state);

  Runner::Parameter param;
  param.max_fused_size = max_fused;
  param.seed = seed;
  param.verbosity = verbosity;
  if (!Runner::Run(param, factory, circuit, state)) return 1;

  uint64_t size = std::min(uint64_t{8}, uint64_t{1} << circuit.num_qubits);
  for (uint64_t i = 0; i < size; ++i) {
    auto a = state_space.GetAmpl(state, i);
    IO::messagef("%llu:%16.8g%16.8g%16.8g\n", (unsigned long long)i,
                 std::real(a), std::imag(a), std::norm(a));
  }
  if (!state_out.empty()) {
    std::vector<float> v(state_space.MinSize(circuit.num_qubits));
    state_space.Copy(state, v.data());
    FILE* f = std::fopen(state_out.c_str(), "wb");
    if (!f) return 1;
    std::fwrite(v.data(), sizeof(float), size_t{2} << circuit.num_qubits, f);
    std::fclose(f);
  }
  if (num_samples > 0) {
    auto bits = state_space.Sample(state, num_samples, sample_seed);
    if (!bits_out.empty()) {
      FILE* f = std::fopen(bits_out.c_str(), "wb");
      if (!f) return 1;
      std::fwrite(bits.data(), sizeof(uint64_t), bits.size(), f);
      std::fclose(f);
    }
  }
  return 0;
}
