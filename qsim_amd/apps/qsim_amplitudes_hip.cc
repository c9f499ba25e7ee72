// qsim_amplitudes_hip.cc -- qsim's qsim_amplitudes app on the gfx950 backend.
//
// Same command line and output as apps/qsim_amplitudes.cc (options :27-98,
// checks :100-136, output :138-153, main :155-214): run the circuit with
// QSimRunner (lib/run_qsim.h:79-174) on FactoryHIP and, at each requested
// time, write the amplitudes of the bitstrings listed in the matching input
// file (qsim's text format, lib/bitstring.h) to the matching output file.
// The amplitudes of a file come from one batched gather
// (StateSpaceHIP::GetAmpls) instead of a GetAmpl per bitstring.
//
//   qsim_amplitudes_hip -c circuit -d t1,t2,.. -i in1,in2,.. -o out1,out2,..
//                       [-s seed] [-t threads] [-f max_fused_size]
//                       [-v verbosity] [-z]
// (-t and -z select host threads and denormal flushing for the CPU
// simulator; they are accepted and have no effect here.)

#include <unistd.h>

#include <complex>
#include <cstdlib>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "bitstring.h"
#include "circuit_qsim_parser.h"
#include "fuser_mqubit.h"
#include "io_file.h"
#include "operation.h"
#include "qsim_amd/factory_hip.h"
#include "run_qsim.h"
#include "util.h"

namespace {

constexpr char kUsage[] =
    "usage:\n  ./qsim_amplitudes_hip -c circuit_file -d times_to_save_results "
    "-i input_files -o output_files -s seed -t num_threads "
    "-f max_fused_size -v verbosity -z\n";

struct Args {
  std::string circuit;
  std::vector<unsigned> times{std::numeric_limits<unsigned>::max()};
  std::vector<std::string> inputs, outputs;
  unsigned seed = 1, max_fused = 2, verbosity = 0;
};

bool Parse(int argc, char* argv[], Args& a) {
  auto to_uint = [](const std::string& w) -> unsigned {
    return unsigned(std::atoi(w.c_str()));
  };
  int k;
  while ((k = getopt(argc, argv, "c:d:i:o:s:t:f:v:z")) != -1) {
    switch (k) {
      case 'c': a.circuit = optarg; break;
      case 'd': qsim::SplitString(optarg, ',', to_uint, a.times); break;
      case 'i': qsim::SplitString(optarg, ',', a.inputs); break;
      case 'o': qsim::SplitString(optarg, ',', a.outputs); break;
      case 's': a.seed = to_uint(optarg); break;
      case 'f': a.max_fused = to_uint(optarg); break;
      case 'v': a.verbosity = to_uint(optarg); break;
      case 't': case 'z': break;
      default: qsim::IO::errorf(kUsage); return false;
    }
  }
  const char* err = nullptr;
  if (a.circuit.empty()) {
    err = "circuit file is not provided.\n";
  } else if (a.inputs.empty()) {
    err = "input files are not provided.\n";
  } else if (a.outputs.empty()) {
    err = "output files are not provided.\n";
  } else if (a.times.size() != a.inputs.size() ||
             a.times.size() != a.outputs.size()) {
    err = "the number of times is not the same as the number of input or "
          "output files.\n";
  }
  for (size_t i = 1; err == nullptr && i < a.times.size(); ++i) {
    if (a.times[i] == a.times[i - 1]) err = "duplicate times to save results.\n";
    if (a.times[i] < a.times[i - 1]) err = "times to save results are not sorted.\n";
  }
  if (err != nullptr) {
    qsim::IO::errorf(err);
    if (a.circuit.empty() || a.inputs.empty() || a.outputs.empty())
      qsim::IO::errorf(kUsage);
    return false;
  }
  return true;
}

// The reference's text layout: 9 significant digits, each value right-aligned
// in 17 columns, one bitstring per line.
template <typename StateSpace, typename State>
bool Write(const std::string& file, const State& state,
           const std::vector<qsim::Bitstring>& bits) {
  const auto amps = StateSpace::GetAmpls(state, bits);
  constexpr int kDigits = 2 * sizeof(float) + 1;
  std::ostringstream os;
  os << std::setprecision(kDigits);
  for (const auto& a : amps) {
    os << std::setw(kDigits + 8) << std::real(a) << std::setw(kDigits + 8)
       << std::imag(a) << "\n";
  }
  return qsim::IOFile::WriteToFile(file, os.str());
}

}  // namespace

int main(int argc, char* argv[]) {
  using namespace qsim;
  Args a;
  if (!Parse(argc, argv, a)) return 1;

  Circuit<Operation<float>> circuit;
  if (!CircuitQsimParser<IOFile>::FromFile(a.times.back(), a.circuit, circuit)) {
    return 1;
  }

  using Factory = FactoryHIP<float>;
  using StateSpace = Factory::StateSpace;
  using State = StateSpace::State;
  using Runner = QSimRunner<IO, MultiQubitGateFuser<IO>, Factory>;

  auto at_time = [&](unsigned k, const StateSpace&, const State& state) {
    std::vector<Bitstring> bits;
    BitstringsFromFile<IOFile>(circuit.num_qubits, a.inputs[k], bits);
    if (!bits.empty()) Write<StateSpace>(a.outputs[k], state, bits);
  };

  Runner::Parameter param;
  param.max_fused_size = a.max_fused;
  param.seed = a.seed;
  param.verbosity = a.verbosity;
  Runner::Run(param, Factory(), a.times, circuit, at_time);

  IO::messagef("all done.\n");
  return 0;
}
