// ref_driver.cc -- TEST INFRASTRUCTURE ONLY.
//
// Drives qsim's own CPU simulators, compiled from the reference headers where
// they lie (oracle/Makefile adds the reference's lib/ to the include path), so
// that the oracle restatement and the HIP backend can be checked against the
// reference itself.  Output goes to oracle/_ref/ only.
//
//   qsim_ref -c FILE [-m avx|basic32|basic64] [-f MAX_FUSED] [-t THREADS]
//            [-s SEED] [-o STATE.bin] [-n NUM_SAMPLES] [-x SAMPLE_SEED]
//            [-b BITSTRINGS.bin] [-d MAXTIME] [-P PASSES]
//   qsim_ref -R NUM -x SEED -M MAX -b DRAWS.bin     (GenerateRandomValues)
//   qsim_ref -c FILE -T R0:R1 -g ROWS:COLS [-p P] [-G GAMMA] [-f MAX_FUSED]
//            [-t THREADS] [-N]  (noisy trajectories, QuantumTrajectorySimulator;
//                                -N: FILE is noisy-circuit text, see NoisyText)
//
// The simulation is QSimRunner::Run with MultiQubitGateFuser
// (lib/run_qsim.h:225-289, lib/fuser_mqubit.h:185-526) on the state of the
// chosen backend; the final state is written in normal order
// (InternalToNormalOrder) as interleaved complex of the backend's fp type.
// One JSON line on stdout reports timings and sizes.
//
// -P PASSES runs only the first PASSES fused operations (a bounded CPU sample
// for the benchmark's baseline leg) and reports their time.

#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "circuit_qsim_parser.h"
#include "fuser_mqubit.h"
#include "gate_appl.h"
#include "io_file.h"
#include "parfor.h"
#include "run_qsim.h"
#include "simmux.h"
#include "simulator_basic.h"
#include "util.h"
#include "channels_cirq.h"
#include "qtrajectory.h"

namespace {

// The simulator lib/simmux.h selected for this build's instruction set.
#if defined(__AVX512F__)
constexpr const char* kSimmuxSimulator = "SimulatorAVX512";
#elif defined(__AVX2__)
constexpr const char* kSimmuxSimulator = "SimulatorAVX";
#elif defined(__SSE4_1__)
constexpr const char* kSimmuxSimulator = "SimulatorSSE";
#else
constexpr const char* kSimmuxSimulator = "SimulatorBasic";
#endif

struct Options {
  std::string circuit_file;
  std::string mode = "avx";
  unsigned max_fused = 2;
  unsigned threads = 1;
  unsigned seed = 1;
  unsigned maxtime = ~0u;
  std::string state_out;
  uint64_t num_samples = 0;
  unsigned sample_seed = 0;
  std::string bits_out;
  uint64_t draws = 0;
  double draws_max = 1.0;
  long passes = -1;
  // Trajectory mode (-T R0:R1 -g ROWS:COLS -p P_DEPOL -G GAMMA).
  bool trajectories = false;
  uint64_t rep0 = 0, rep1 = 0;
  unsigned rows = 0, cols = 0;
  double p_depol = 1e-3, gamma = 5e-3;
  bool noisy_text = false;
};

template <typename Sim, typename FP>
struct Factory {
  using Simulator = Sim;
  using StateSpace = typename Sim::StateSpace;
  explicit Factory(unsigned nt) : nt(nt) {}
  StateSpace CreateStateSpace() const { return StateSpace(nt); }
  Simulator CreateSimulator() const { return Simulator(nt); }
  unsigned nt;
};

bool WriteFile(const std::string& path, const void* p, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  size_t w = std::fwrite(p, 1, bytes, f);
  std::fclose(f);
  return w == bytes;
}

template <typename Sim, typename FP>
int Run(const Options& opt) {
  using namespace qsim;
  using Fac = Factory<Sim, FP>;
  using StateSpace = typename Fac::StateSpace;
  using Fuser = MultiQubitGateFuser<IO>;
  using Runner = QSimRunner<IO, Fuser, Fac>;

  Circuit<Operation<FP>> circuit;
  if (!CircuitQsimParser<IOFile>::FromFile(opt.maxtime, opt.circuit_file,
                                           circuit)) {
    std::fprintf(stderr, "cannot parse %s\n", opt.circuit_file.c_str());
    return 1;
  }

  Fac factory(opt.threads);
  StateSpace ss = factory.CreateStateSpace();
  auto state = ss.Create(circuit.num_qubits);
  if (ss.IsNull(state)) {
    std::fprintf(stderr, "not enough memory\n");
    return 1;
  }
  ss.SetStateZero(state);

  typename Runner::Parameter param;
  param.max_fused_size = opt.max_fused;
  param.seed = opt.seed;
  param.verbosity = 0;

  double t0 = GetTime();
  double t_fuse = 0, t_sim = 0;
  size_t num_fused = 0;
  bool ok = true;

  if (opt.passes < 0) {
    // The whole circuit through QSimRunner::Run (its fusion included); the
    // fused-pass count from a separate fusion of the same circuit.
    double tf = GetTime();
    num_fused = Fuser::FuseGates(param, circuit.num_qubits, circuit.ops).size();
    t_fuse = GetTime() - tf;
    Sim sim = factory.CreateSimulator();
    std::vector<typename StateSpace::MeasurementResult> mres;
    t0 = GetTime();
    ok = Runner::Run(param, circuit, ss, sim, state, mres);
    t_sim = GetTime() - t0;
  } else {
    // Bounded sample: fuse the whole circuit, apply the first P fused ops.
    Sim sim = factory.CreateSimulator();
    double tf = GetTime();
    auto fused = Fuser::FuseGates(param, circuit.num_qubits, circuit.ops);
    t_fuse = GetTime() - tf;
    num_fused = fused.size();
    std::mt19937 rgen(opt.seed);
    size_t p = std::min<size_t>(fused.size(), size_t(opt.passes));
    t0 = GetTime();
    for (size_t i = 0; i < p; ++i) {
      ApplyGate(ss, sim, fused[i], rgen, state);
    }
    t_sim = GetTime() - t0;
  }

  if (!ok) {
    std::fprintf(stderr, "simulation failed\n");
    return 1;
  }

  ss.InternalToNormalOrder(state);
  uint64_t amps = uint64_t{1} << circuit.num_qubits;
  if (!opt.state_out.empty()) {
    std::vector<FP> v(StateSpace::MinSize(circuit.num_qubits));
    ss.Copy(state, v.data());
    if (!WriteFile(opt.state_out, v.data(), 2 * amps * sizeof(FP))) return 1;
  }

  double t_samp = 0;
  if (opt.num_samples > 0) {
    // Sample on the normal-order copy so the walk follows amplitude order.
    using Basic = SimulatorBasic<ParallelFor, FP>;
    using BSS = typename Basic::StateSpace;
    BSS bss(1);
    auto bstate = bss.Create(circuit.num_qubits);
    ss.Copy(state, bstate.get());
    double ts = GetTime();
    auto bits = bss.Sample(bstate, opt.num_samples, opt.sample_seed);
    t_samp = GetTime() - ts;
    if (!opt.bits_out.empty() &&
        !WriteFile(opt.bits_out, bits.data(), bits.size() * sizeof(uint64_t)))
      return 1;
  }

  double norm = ss.Norm(state);
  std::printf(
      "{\"mode\": \"%s\", \"simulator\": \"%s\", "
      "\"num_qubits\": %u, \"num_gates\": %zu, "
      "\"max_fused\": %u, \"threads\": %u, \"fused_ops\": %zu, "
      "\"passes\": %ld, \"fuse_s\": %.9g, \"sim_s\": %.9g, "
      "\"sample_s\": %.9g, \"norm\": %.17g}\n",
      opt.mode.c_str(),
      opt.mode == "avx" ? kSimmuxSimulator
                        : (opt.mode == "basic64" ? "SimulatorBasic<double>"
                                                 : "SimulatorBasic<float>"),
      circuit.num_qubits, circuit.ops.size(),
      opt.max_fused, opt.threads, num_fused, opt.passes, t_fuse, t_sim, t_samp,
      norm);
  return 0;
}

// Noisy grid circuit, by the rule the trajectory engine states
// (qsim_amd/apps/trajectories.cc): gate times t -> 3t; after every gate a
// depolarising channel on each of its qubits (time 3t + 1); after the last
// gate of cycle c (gate times 2c, 2c + 1) an amplitude-damping channel on
// every qubit (time 3(2c + 1) + 2).
qsim::Circuit<qsim::Operation<float>> NoisyGrid(
    const qsim::Circuit<qsim::Operation<float>>& c, double p, double gamma) {
  using namespace qsim;
  Circuit<Operation<float>> nc;
  nc.num_qubits = c.num_qubits;
  for (size_t i = 0; i < c.ops.size(); ++i) {
    Gate<float> g = *OpGetAlternative<Gate<float>>(c.ops[i]);
    const unsigned t = g.time;
    g.time = 3 * t;
    const std::vector<unsigned> qs = g.qubits;
    nc.ops.push_back(std::move(g));
    for (unsigned q : qs) {
      nc.ops.push_back(Cirq::DepolarizingChannel<float>::Create(3 * t + 1, q, p));
    }
    if (i + 1 == c.ops.size() || OpTime(c.ops[i + 1]) / 2 != t / 2) {
      for (unsigned q = 0; q < c.num_qubits; ++q) {
        nc.ops.push_back(Cirq::AmplitudeDampingChannel<float>::Create(
            3 * (2 * (t / 2) + 1) + 2, q, gamma));
      }
    }
  }
  return nc;
}

// Noisy-circuit text: qsim circuit lines, and lines "<time> channel <name>
// <qubit> <params>" (name: depolarize, asymmetric_depolarize,
// amplitude_damp, generalized_amplitude_damp, phase_damp, phase_flip,
// bit_flip, reset; built by qsim's factories, lib/channels_cirq.h) or
// "<time> channel kraus <nq> <qubits> <nk> {<unitary> <prob> <matrix>}".
// Operations in file order, gates and channels with their given times.
bool NoisyText(const std::string& path,
               qsim::Circuit<qsim::Operation<float>>& nc) {
  using namespace qsim;
  std::ifstream f(path);
  if (!f) return false;
  std::string line;
  unsigned n = 0;
  while (std::getline(f, line)) {
    std::istringstream in(line);
    if (in >> n) break;
  }
  if (n == 0) return false;
  nc.num_qubits = n;
  while (std::getline(f, line)) {
    std::istringstream in(line);
    unsigned time;
    std::string word;
    if (!(in >> time >> word)) continue;
    if (word != "channel") {
      Circuit<Operation<float>> one;
      std::stringstream gs;
      gs << n << "\n" << line << "\n";
      if (!CircuitQsimParser<IO>::FromStream(~0u, "line", gs, one)) {
        return false;
      }
      for (auto& op : one.ops) nc.ops.push_back(std::move(op));
      continue;
    }
    std::string name;
    in >> name;
    if (name == "kraus") {
      unsigned nq, nk;
      in >> nq;
      std::vector<unsigned> qs(nq);
      for (auto& q : qs) in >> q;
      in >> nk;
      Channel<float> ch{{kChannel, time, qs}, {}};
      for (unsigned i = 0; i < nk; ++i) {
        int unitary;
        double prob;
        in >> unitary >> prob;
        Matrix<float> m(size_t{2} << (2 * nq));
        for (auto& v : m) {
          double x;
          in >> x;
          v = float(x);
        }
        KrausOperator<float> kop;
        kop.unitary = unitary != 0;
        kop.prob = prob;
        kop.ops.push_back(
            CreateGate<Gate<float>, Cirq::MatrixGate<float>>(time, qs,
                                                              std::move(m)));
        kop.CalculateKdKMatrix();
        ch.kops.push_back(std::move(kop));
      }
      if (!in) return false;
      nc.ops.push_back(std::move(ch));
      continue;
    }
    unsigned q;
    double a = 0, b = 0, c = 0;
    in >> q;
    if (name == "depolarize" && (in >> a)) {
      nc.ops.push_back(Cirq::DepolarizingChannel<float>::Create(time, q, a));
    } else if (name == "asymmetric_depolarize" && (in >> a >> b >> c)) {
      nc.ops.push_back(
          Cirq::AsymmetricDepolarizingChannel<float>::Create(time, q, a, b, c));
    } else if (name == "amplitude_damp" && (in >> a)) {
      nc.ops.push_back(Cirq::AmplitudeDampingChannel<float>::Create(time, q, a));
    } else if (name == "generalized_amplitude_damp" && (in >> a >> b)) {
      nc.ops.push_back(
          Cirq::GeneralizedAmplitudeDampingChannel<float>::Create(time, q, a, b));
    } else if (name == "phase_damp" && (in >> a)) {
      nc.ops.push_back(Cirq::PhaseDampingChannel<float>::Create(time, q, a));
    } else if (name == "phase_flip" && (in >> a)) {
      nc.ops.push_back(Cirq::PhaseFlipChannel<float>::Create(time, q, a));
    } else if (name == "bit_flip" && (in >> a)) {
      nc.ops.push_back(Cirq::BitFlipChannel<float>::Create(time, q, a));
    } else if (name == "reset") {
      nc.ops.push_back(Cirq::ResetChannel<float>::Create(time, q));
    } else {
      return false;
    }
  }
  return true;
}

// qsim's QuantumTrajectorySimulator (lib/qtrajectory.h) on SimulatorAVX for
// repetitions [r0, r1): one JSON line per trajectory with the Kraus index of
// every channel (collect_kop_stat, gates' zeros dropped) and <Z_a Z_b> of the
// grid-neighbour pairs (horizontal then vertical per cell, row by row).
int RunTrajectories(const Options& opt) {
  using namespace qsim;
  using Sim = qsim::Simulator<ParallelFor>;
  using Fac = Factory<Sim, float>;
  using Runner = QSimRunner<IO, MultiQubitGateFuser<IO>, Fac>;
  using QT = QuantumTrajectorySimulator<IO, Runner>;
  Circuit<Operation<float>> circuit, nc;
  if (opt.noisy_text) {
    if (!NoisyText(opt.circuit_file, nc)) return 1;
    circuit.num_qubits = nc.num_qubits;
  } else {
    if (!CircuitQsimParser<IOFile>::FromFile(opt.maxtime, opt.circuit_file,
                                             circuit)) {
      return 1;
    }
    nc = NoisyGrid(circuit, opt.p_depol, opt.gamma);
  }
  const unsigned rows = opt.rows, cols = opt.cols;
  if (rows * cols != circuit.num_qubits) return 1;
  std::vector<char> is_channel;
  for (const auto& op : nc.ops) {
    is_channel.push_back(OpGetAlternative<Channel<float>>(op) != nullptr);
  }
  std::vector<std::pair<unsigned, unsigned>> pairs;
  for (unsigned r = 0; r < rows; ++r) {
    for (unsigned c = 0; c < cols; ++c) {
      if (c + 1 < cols) pairs.push_back({r * cols + c, r * cols + c + 1});
      if (r + 1 < rows) pairs.push_back({r * cols + c, (r + 1) * cols + c});
    }
  }
  const std::vector<float> zz = {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, -1, 0, 0, 0, 0, 0,
                                 0, 0, 0, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0};
  Fac factory(opt.threads);
  auto ss = factory.CreateStateSpace();
  auto sim = factory.CreateSimulator();
  typename QT::Parameter param;
  param.max_fused_size = opt.max_fused;
  param.verbosity = 0;
  param.collect_kop_stat = true;
  double t_sim = 0;
  auto measure = [&](uint64_t rep, const typename Fac::StateSpace::State& st,
                     const typename QT::Stat& stat) {
    std::printf("{\"rep\": %llu, \"kraus\": [", (unsigned long long)rep);
    bool first = true;
    for (size_t i = 0; i < stat.samples.size() && i < is_channel.size(); ++i) {
      if (!is_channel[i]) continue;
      std::printf("%s%llu", first ? "" : ", ",
                  (unsigned long long)stat.samples[i]);
      first = false;
    }
    std::printf("], \"zz\": [");
    for (size_t j = 0; j < pairs.size(); ++j) {
      auto v = sim.ExpectationValue({pairs[j].first, pairs[j].second},
                                    zz.data(), st);
      std::printf("%s%.17g", j ? ", " : "", std::real(v));
    }
    std::printf("]}\n");
  };
  double t0 = GetTime();
  bool ok = QT::RunBatch(param, nc, opt.rep0, opt.rep1, ss, sim, measure);
  t_sim = GetTime() - t0;
  std::printf("{\"trajectories\": %llu, \"num_qubits\": %u, \"sim_s\": %.9g, "
              "\"threads\": %u}\n",
              (unsigned long long)(opt.rep1 - opt.rep0), circuit.num_qubits,
              t_sim, opt.threads);
  return ok ? 0 : 1;
}

}  // namespace

int main(int argc, char* argv[]) {
  Options opt;
  int k;
  while ((k = getopt(argc, argv, "c:m:f:t:s:o:n:x:b:d:R:M:P:T:g:p:G:N")) != -1) {
    switch (k) {
      case 'N': opt.noisy_text = true; break;
      case 'T':
        opt.trajectories = true;
        if (std::sscanf(optarg, "%llu:%llu",
                        reinterpret_cast<unsigned long long*>(&opt.rep0),
                        reinterpret_cast<unsigned long long*>(&opt.rep1)) != 2)
          return 1;
        break;
      case 'g':
        if (std::sscanf(optarg, "%u:%u", &opt.rows, &opt.cols) != 2) return 1;
        break;
      case 'p': opt.p_depol = std::strtod(optarg, nullptr); break;
      case 'G': opt.gamma = std::strtod(optarg, nullptr); break;
      case 'c': opt.circuit_file = optarg; break;
      case 'm': opt.mode = optarg; break;
      case 'f': opt.max_fused = std::atoi(optarg); break;
      case 't': opt.threads = std::atoi(optarg); break;
      case 's': opt.seed = std::atoi(optarg); break;
      case 'o': opt.state_out = optarg; break;
      case 'n': opt.num_samples = std::strtoull(optarg, nullptr, 10); break;
      case 'x': opt.sample_seed = std::atoi(optarg); break;
      case 'b': opt.bits_out = optarg; break;
      case 'd': opt.maxtime = std::atoi(optarg); break;
      case 'R': opt.draws = std::strtoull(optarg, nullptr, 10); break;
      case 'M': opt.draws_max = std::strtod(optarg, nullptr); break;
      case 'P': opt.passes = std::atol(optarg); break;
      default:
        std::fprintf(stderr, "bad option\n");
        return 1;
    }
  }

  if (opt.draws > 0) {
    auto rs = qsim::GenerateRandomValues<double>(opt.draws, opt.sample_seed,
                                                 opt.draws_max);
    return WriteFile(opt.bits_out, rs.data(), rs.size() * sizeof(double)) ? 0
                                                                          : 1;
  }

  if (opt.circuit_file.empty()) {
    std::fprintf(stderr, "no circuit file\n");
    return 1;
  }

  using For = qsim::ParallelFor;
  if (opt.trajectories) return RunTrajectories(opt);
  if (opt.mode == "avx") return Run<qsim::Simulator<For>, float>(opt);
  if (opt.mode == "basic32") return Run<qsim::SimulatorBasic<For, float>, float>(opt);
  if (opt.mode == "basic64") return Run<qsim::SimulatorBasic<For, double>, double>(opt);
  std::fprintf(stderr, "unknown mode %s\n", opt.mode.c_str());
  return 1;
}
