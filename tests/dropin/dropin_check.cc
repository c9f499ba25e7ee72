// dropin_check.cc -- qsim's own drivers on the gfx950 backend.
//
// Every scenario instantiates a qsim driver twice, once with qsim's CPU
// SimulatorBasic Factory and once with FactoryHIP, and compares the results
// (and, where qsim's test suite holds known answers, checks those):
//
//   runner      QSimRunner::Run with MultiQubitGateFuser and mid-circuit
//               measurements (lib/run_qsim.h:225-289)
//   callback    QSimRunner::Run with measurement times and a callback that
//               takes expectation values (lib/run_qsim.h:61-150)
//   expect      ExpectationValue<IO, Fuser> over Pauli-string sums, in place
//               and on a temporary ket (lib/expect.h:48-153); the known
//               answers of TestExpectationValue2
//   qtrajectory QuantumTrajectorySimulator::RunBatch on a noisy grid circuit
//               (depolarising + amplitude damping; lib/qtrajectory.h:112-387)
//   hybrid      HybridSimulator::Run amplitudes (lib/hybrid.h:304-453) with the
//               known answers of TestHybrid2
//   sample      StateSpace::Sample on an identical state (lib/statespace.h)
//   learned     QSimRunner with LearnedFuserHIP: blocked passes of 14 tile
//               qubits (gates reach SimulatorHIP::ApplyGate and the gate
//               queue groups them) on the 20-qubit grid against
//               SimulatorBasic<double>, with and without a mid-circuit
//               measurement
//   learned-model  (with a cost-model file as argv[1]) the model's own pick
//               on a 24-qubit grid against the unfused run on FactoryHIP
//
// Prints one JSON object per scenario and exits 0 only if all pass.

#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "channels_cirq.h"
#include "circuit.h"
#include "circuit_noisy.h"
#include "circuit_qsim_parser.h"
#include "expect.h"
#include "fuser_basic.h"
#include "fuser_mqubit.h"
#include "gates_cirq.h"
#include "gates_qsim.h"
#include "hybrid.h"
#include "io.h"
#include "parfor.h"
#include "qsim_amd/factory_hip.h"
#include "qsim_amd/learned_fuser_hip.h"
#include "qtrajectory.h"
#include "run_qsim.h"
#include "seqfor.h"
#include "simulator_basic.h"

namespace {

using namespace qsim;

template <typename FP>
struct FactoryCPU {
  using Simulator = SimulatorBasic<ParallelFor, FP>;
  using StateSpace = typename Simulator::StateSpace;
  using fp_type = FP;
  StateSpace CreateStateSpace() const { return StateSpace(4); }
  Simulator CreateSimulator() const { return Simulator(4); }
};

int g_failures = 0;

void Report(const char* name, bool ok, const std::string& detail) {
  std::printf("{\"scenario\": \"%s\", \"pass\": %s, %s}\n", name,
              ok ? "true" : "false", detail.c_str());
  std::fflush(stdout);
  if (!ok) ++g_failures;
}

template <typename Factory>
std::vector<std::complex<double>> Amplitudes(
    const Factory& f, const typename Factory::StateSpace::State& s) {
  auto ss = f.CreateStateSpace();
  std::vector<std::complex<double>> a;
  uint64_t size = uint64_t{1} << s.num_qubits();
  std::vector<typename Factory::fp_type> v(ss.MinSize(s.num_qubits()));
  ss.Copy(s, v.data());
  // Normal order for both backends (SimulatorBasic and the HIP layout are
  // interleaved complex).
  for (uint64_t i = 0; i < size; ++i) a.emplace_back(v[2 * i], v[2 * i + 1]);
  return a;
}

double RelL2(const std::vector<std::complex<double>>& a,
             const std::vector<std::complex<double>>& b) {
  double num = 0, den = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    num += std::norm(a[i] - b[i]);
    den += std::norm(b[i]);
  }
  return std::sqrt(num / den);
}

std::string GridText(unsigned rows, unsigned cols, unsigned cycles,
                     unsigned seed, bool measure_mid) {
  // The grid construction of qsim_amd/circuits.py, with optional
  // measurements after the middle cycle.
  std::mt19937 rng(seed);
  const char* one[3] = {"x_1_2", "y_1_2", "hz_1_2"};
  unsigned n = rows * cols;
  std::vector<int> last(n, -1);
  std::ostringstream os;
  os << n << "\n";
  unsigned t = 0;
  for (unsigned c = 0; c < cycles; ++c) {
    for (unsigned q = 0; q < n; ++q) {
      int g;
      do { g = rng() % 3; } while (g == last[q]);
      last[q] = g;
      os << t << " " << one[g] << " " << q << "\n";
    }
    ++t;
    unsigned p = c % 4;
    if (p < 2) {
      for (unsigned r = 0; r < rows; ++r)
        for (unsigned cc = p; cc + 1 < cols; cc += 2)
          os << t << " fs " << r * cols + cc << " " << r * cols + cc + 1
             << " 1.5707963267948966 0.5235987755982988\n";
    } else {
      for (unsigned r = p - 2; r + 1 < rows; r += 2)
        for (unsigned cc = 0; cc < cols; ++cc)
          os << t << " fs " << r * cols + cc << " " << (r + 1) * cols + cc
             << " 1.5707963267948966 0.5235987755982988\n";
    }
    ++t;
    if (measure_mid && c == cycles / 2) {
      os << t << " m 0 3 5\n";
      ++t;
    }
  }
  return os.str();
}

template <typename FP>
Circuit<Operation<FP>> Parse(const std::string& text) {
  Circuit<Operation<FP>> c;
  std::stringstream ss(text);
  if (!CircuitQsimParser<IO>::FromStream(~0u, "dropin", ss, c)) {
    std::fprintf(stderr, "parse failed\n");
    std::exit(2);
  }
  return c;
}

// ---- runner: fused run with mid-circuit measurements --------------------
template <typename Factory>
std::pair<std::vector<std::complex<double>>, std::vector<uint64_t>> RunRunner(
    const Factory& f, const Circuit<Operation<float>>& c, unsigned k) {
  using Runner = QSimRunner<IO, MultiQubitGateFuser<IO>, Factory>;
  typename Runner::Parameter p;
  p.max_fused_size = k;
  p.seed = 7;
  p.verbosity = 0;
  auto ss = f.CreateStateSpace();
  auto s = ss.Create(c.num_qubits);
  ss.SetStateZero(s);
  std::vector<typename Factory::StateSpace::MeasurementResult> mres;
  bool ok = Runner::Run(p, f, c, s, mres);
  std::vector<uint64_t> bits;
  for (auto& m : mres) bits.push_back(ok ? m.bits : ~uint64_t{0});
  return {Amplitudes(f, s), bits};
}

void Scenario_Runner() {
  auto c = Parse<float>(GridText(3, 4, 10, 3, true));
  FactoryCPU<float> cpu;
  FactoryHIP<float> hip;
  bool ok = true;
  std::ostringstream d;
  d << "\"cases\": [";
  for (unsigned k : {2u, 4u, 6u}) {
    auto a = RunRunner(cpu, c, k);
    auto b = RunRunner(hip, c, k);
    double e = RelL2(b.first, a.first);
    bool same_bits = a.second == b.second && !a.second.empty();
    ok = ok && e < 1e-5 && same_bits;
    d << (k == 2 ? "" : ", ") << "{\"max_fused\": " << k << ", \"rel_l2\": "
      << e << ", \"measured_bits_equal\": " << (same_bits ? "true" : "false")
      << "}";
  }
  d << "]";
  Report("runner", ok, d.str());
}

// ---- callback: measurement times + expectation values -------------------
template <typename Factory>
std::vector<double> RunCallback(const Factory& f,
                                const Circuit<Operation<float>>& c) {
  using Runner = QSimRunner<IO, MultiQubitGateFuser<IO>, Factory>;
  typename Runner::Parameter p;
  p.max_fused_size = 3;
  p.seed = 1;
  p.verbosity = 0;
  std::vector<double> out;
  auto measure = [&](unsigned k, const typename Factory::StateSpace& ss,
                     typename Factory::StateSpace::State& s) {
    auto sim = f.CreateSimulator();
    // <Z_0 Z_1> through the simulator's ExpectationValue.
    std::vector<float> zz = {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, -1, 0, 0, 0, 0, 0,
                             0, 0, 0, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0};
    out.push_back(std::real(sim.ExpectationValue({0, 1}, zz.data(), s)));
    out.push_back(ss.Norm(s));
  };
  Runner::Run(p, f, {3, 9, 15}, c, measure);
  return out;
}

void Scenario_Callback() {
  auto c = Parse<float>(GridText(3, 4, 8, 5, false));
  auto a = RunCallback(FactoryCPU<float>(), c);
  auto b = RunCallback(FactoryHIP<float>(), c);
  bool ok = a.size() == b.size() && !a.empty();
  double worst = 0;
  for (size_t i = 0; ok && i < a.size(); ++i)
    worst = std::max(worst, std::abs(a[i] - b[i]));
  ok = ok && worst < 1e-5;
  std::ostringstream d;
  d << "\"values\": " << a.size() << ", \"max_abs_diff\": " << worst;
  Report("callback", ok, d.str());
}

// ---- expect: TestExpectationValue2 ---------------------------------------
template <typename Factory>
std::vector<std::complex<double>> RunExpect(const Factory& f, bool temp_ket) {
  using fp_type = typename Factory::fp_type;
  using Fuser = MultiQubitGateFuser<IO>;
  unsigned n = 16, depth = 16;
  auto ss = f.CreateStateSpace();
  auto sim = f.CreateSimulator();
  auto s = ss.Create(n);
  ss.SetStateZero(s);
  std::vector<GateQSim<fp_type>> circuit;
  for (unsigned k = 0; k < n; ++k) circuit.push_back(GateHd<fp_type>::Create(0, k));
  unsigned t = 1;
  for (unsigned i = 0; i < depth / 2; ++i) {
    for (unsigned k = 0; k < n; ++k) circuit.push_back(GateRX<fp_type>::Create(t, k, 0.1 * k));
    ++t;
    for (unsigned k = 0; k < n / 2; ++k) circuit.push_back(GateIS<fp_type>::Create(t, 2 * k, 2 * k + 1));
    ++t;
    for (unsigned k = 0; k < n; ++k) circuit.push_back(GateRY<fp_type>::Create(t, k, 0.1 * k));
    ++t;
    for (unsigned k = 0; k < n / 2; ++k) circuit.push_back(GateIS<fp_type>::Create(t, 2 * k + 1, (2 * k + 2) % n));
    ++t;
  }
  for (const auto& g : circuit) ApplyGate(sim, g, s);
  std::vector<std::complex<double>> res;
  for (unsigned k = 1; k <= 6; ++k) {
    std::vector<OpString<fp_type>> strings;
    for (unsigned i = 0; i <= n - k; ++i) {
      strings.push_back({{0.1 + 0.2 * i, 0}, {}});
      for (unsigned j = 0; j < k; ++j) {
        switch (j % 3) {
          case 0: strings.back().ops.push_back(GateX<fp_type>::Create(0, i + j)); break;
          case 1: strings.back().ops.push_back(GateY<fp_type>::Create(0, i + j)); break;
          default: strings.back().ops.push_back(GateZ<fp_type>::Create(0, i + j)); break;
        }
      }
    }
    if (temp_ket) {
      auto tmp = ss.Create(n);
      typename Fuser::Parameter fp;
      fp.max_fused_size = 4;
      res.push_back(ExpectationValue<IO, Fuser>(fp, strings, ss, sim, s, tmp));
    } else {
      res.push_back(ExpectationValue<IO, Fuser>(strings, sim, s));
    }
  }
  return res;
}

void Scenario_Expect() {
  // expected_real of tests/simulator_testfixture.h:1590-1597.
  const double expected[6] = {0.014314421865856278, 0.021889885055134076,
                              -0.006954622792545706, 0.013091871136566622,
                              0.004322795104235413, -0.008040613483171907};
  bool ok = true;
  double worst = 0, worst_im = 0, worst_cpu = 0;
  for (bool temp : {false, true}) {
    auto b = RunExpect(FactoryHIP<float>(), temp);
    auto a = RunExpect(FactoryCPU<float>(), temp);
    for (int k = 0; k < 6; ++k) {
      worst = std::max(worst, std::abs(std::real(b[k]) - expected[k]));
      worst_im = std::max(worst_im, std::abs(std::imag(b[k])));
      worst_cpu = std::max(worst_cpu, std::abs(b[k] - a[k]));
    }
  }
  ok = worst < 1e-6 && worst_im < 1e-8 && worst_cpu < 1e-6;
  std::ostringstream d;
  d << "\"max_abs_err_vs_known\": " << worst << ", \"max_abs_imag\": "
    << worst_im << ", \"max_abs_diff_vs_cpu\": " << worst_cpu;
  Report("expect", ok, d.str());
}

// ---- qtrajectory ---------------------------------------------------------
struct NoiseFactory {
  double p;
  Channel<float> Create(unsigned time, unsigned q) const {
    return Cirq::DepolarizingChannel<float>::Create(time, q, p);
  }
};

template <typename Factory>
void RunTrajectories(const Factory& f, const Circuit<Operation<float>>& nc,
                     unsigned reps, std::vector<std::vector<uint64_t>>& kops,
                     std::vector<double>& zz) {
  using Runner = QSimRunner<IO, MultiQubitGateFuser<IO>, Factory>;
  using QT = QuantumTrajectorySimulator<IO, Runner>;
  typename QT::Parameter p;
  p.max_fused_size = 2;
  p.verbosity = 0;
  p.collect_kop_stat = true;
  auto ss = f.CreateStateSpace();
  auto sim = f.CreateSimulator();
  std::vector<float> zzm = {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, -1, 0, 0, 0, 0, 0,
                            0, 0, 0, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0};
  auto measure = [&](uint64_t r, const typename Factory::StateSpace::State& s,
                     const typename QT::Stat& stat) {
    kops.push_back(stat.samples);
    zz.push_back(std::real(sim.ExpectationValue({0, 1}, zzm.data(), s)));
  };
  QT::RunBatch(p, nc, 0, reps, ss, sim, measure);
}

void Scenario_Trajectories() {
  auto c = Parse<float>(GridText(2, 5, 6, 11, false));
  auto nc = MakeNoisy<float>(c.num_qubits, c.ops.begin(), c.ops.end(),
                             NoiseFactory{0.02});
  // Amplitude damping on every qubit at the end (non-unitary Kraus branch
  // choice through ExpectationValue).
  unsigned tend = 0;
  for (const auto& op : nc.ops) tend = std::max(tend, OpTime(op));
  for (unsigned q = 0; q < nc.num_qubits; ++q)
    nc.ops.push_back(Cirq::AmplitudeDampingChannel<float>::Create(tend + 1, q, 0.05));
  const unsigned reps = 300;
  std::vector<std::vector<uint64_t>> ka, kb;
  std::vector<double> za, zb;
  RunTrajectories(FactoryCPU<float>(), nc, reps, ka, za);
  RunTrajectories(FactoryHIP<float>(), nc, reps, kb, zb);
  unsigned same = 0;
  double worst = 0, mean_a = 0, mean_b = 0;
  for (unsigned r = 0; r < reps; ++r) {
    if (ka[r] == kb[r]) {
      ++same;
      worst = std::max(worst, std::abs(za[r] - zb[r]));
    }
    mean_a += za[r] / reps;
    mean_b += zb[r] / reps;
  }
  bool ok = ka.size() == reps && kb.size() == reps && same >= reps - 3 &&
            worst < 1e-4;
  std::ostringstream d;
  d << "\"trajectories\": " << reps << ", \"same_kraus_sequence\": " << same
    << ", \"max_abs_diff_zz_same_sequence\": " << worst
    << ", \"mean_zz_cpu\": " << mean_a << ", \"mean_zz_hip\": " << mean_b;
  Report("qtrajectory", ok, d.str());
}

// ---- hybrid: TestHybrid2 -------------------------------------------------
template <typename Factory>
std::vector<std::complex<float>> RunHybrid(const Factory& f,
                                           unsigned root_gatexs) {
  const char* text =
      "2\n0 h 0\n0 h 1\n1 cz 0 1\n2 t 0\n2 t 1\n3 cz 0 1\n4 x_1_2 0\n"
      "4 y_1_2 1\n5 cz 0 1\n6 t 0\n6 t 1\n7 cz 0 1\n8 rx 0 0.7\n8 ry 1 0.3\n"
      "9 cz 0 1\n10 t 0\n10 t 1\n11 x_1_2 0\n11 y_1_2 1\n12 cz 0 1\n"
      "13 x_1_2 0\n13 y_1_2 1\n14 sw 0 1\n";
  auto circuit = Parse<float>(text);
  using Fuser = BasicGateFuser<IO>;
  using HS = HybridSimulator<IO, SequentialFor>;
  std::vector<unsigned> parts = {0, 1};
  typename HS::template HybridData<float> hd;
  HS::SplitLattice(parts, circuit.ops, hd);
  typename HS::template Parameter<Fuser::Parameter> param;
  param.prefix = 1;
  param.num_prefix_gatexs = 0;
  param.num_root_gatexs = root_gatexs;
  param.num_threads = 1;
  param.verbosity = 0;
  auto f0 = Fuser::FuseGates(param, hd.num_qubits0, hd.ops0);
  auto f1 = Fuser::FuseGates(param, hd.num_qubits1, hd.ops1);
  std::vector<uint64_t> bits = {0, 1, 2, 3};
  std::vector<std::complex<float>> res(4, 0);
  HS(1).Run(param, f, hd, parts, f0, f1, bits, res);
  return res;
}

void Scenario_Hybrid() {
  // tests/hybrid_testfixture.h:113-121.
  const std::complex<double> expected[4] = {{-0.16006945, -0.04964612},
                                            {0.22667059, -0.03155057},
                                            {-0.03155057, 0.56567556},
                                            {0.28935891, 0.71751291}};
  double worst = 0;
  for (unsigned root : {0u, 1u}) {
    auto r = RunHybrid(FactoryHIP<float>(), root);
    for (int i = 0; i < 4; ++i)
      worst = std::max(worst, std::abs(std::complex<double>(r[i]) - expected[i]));
  }
  std::ostringstream d;
  d << "\"max_abs_err_vs_known\": " << worst;
  Report("hybrid", worst < 1e-6, d.str());
}

// ---- sample: identical state, identical draws ----------------------------
void Scenario_Sample() {
  unsigned n = 14;
  FactoryCPU<float> cpu;
  FactoryHIP<float> hip;
  auto css = cpu.CreateStateSpace();
  auto hss = hip.CreateStateSpace();
  auto cs = css.Create(n);
  auto hs = hss.Create(n);
  std::mt19937 rng(3);
  std::normal_distribution<float> nd;
  std::vector<float> v(2 << n);
  double norm = 0;
  for (auto& x : v) {
    x = nd(rng);
    norm += double(x) * x;
  }
  for (auto& x : v) x = float(x / std::sqrt(norm));
  css.Copy(v.data(), cs);
  hss.Copy(v.data(), hs);
  auto a = css.Sample(cs, 5000, 17);
  auto b = hss.Sample(hs, 5000, 17);
  // Measure through the base class (VirtualMeasure + Collapse).
  std::mt19937 r1(5), r2(5);
  auto ma = css.Measure({1, 4, 9}, r1, cs);
  auto mb = hss.Measure({1, 4, 9}, r2, hs);
  auto aa = Amplitudes(cpu, cs);
  auto ab = Amplitudes(hip, hs);
  bool ok = a == b && ma.bits == mb.bits && RelL2(ab, aa) < 1e-6;
  std::ostringstream d;
  unsigned diff = 0;
  for (size_t i = 0; i < a.size(); ++i) diff += a[i] != b[i];
  d << "\"samples\": " << a.size() << ", \"differing\": " << diff
    << ", \"measured_bits_equal\": " << (ma.bits == mb.bits ? "true" : "false")
    << ", \"collapsed_rel_l2\": " << RelL2(ab, aa);
  Report("sample", ok, d.str());
}

}  // namespace

// ---- learned: QSimRunner with LearnedFuserHIP -----------------------------
template <typename Factory, typename Fuser>
std::pair<std::vector<std::complex<double>>, std::vector<uint64_t>> RunWith(
    const Factory& f, const Circuit<Operation<typename Factory::fp_type>>& c,
    const typename QSimRunner<IO, Fuser, Factory>::Parameter& p) {
  using Runner = QSimRunner<IO, Fuser, Factory>;
  auto ss = f.CreateStateSpace();
  auto s = ss.Create(c.num_qubits);
  ss.SetStateZero(s);
  std::vector<typename Factory::StateSpace::MeasurementResult> mres;
  bool ok = Runner::Run(p, f, c, s, mres);
  std::vector<uint64_t> bits;
  for (auto& m : mres) bits.push_back(ok ? m.bits : ~uint64_t{0});
  if (!ok) bits.push_back(~uint64_t{0});
  return {Amplitudes(f, s), bits};
}

void Scenario_Learned() {
  bool ok = true;
  std::ostringstream d;
  d << "\"cases\": [";
  for (bool mid : {false, true}) {
    const std::string text = GridText(4, 5, 14, 0, mid);
    auto c32 = Parse<float>(text);
    auto c64 = Parse<double>(text);
    using RunnerCpu = QSimRunner<IO, MultiQubitGateFuser<IO>, FactoryCPU<double>>;
    typename RunnerCpu::Parameter pc;
    pc.max_fused_size = 2;
    pc.seed = 11;
    pc.verbosity = 0;
    auto ref = RunWith<FactoryCPU<double>, MultiQubitGateFuser<IO>>(
        FactoryCPU<double>(), c64, pc);
    using Learned = LearnedFuserHIP<IO>;
    typename QSimRunner<IO, Learned, FactoryHIP<float>>::Parameter pl;
    pl.tile_bits = 14;
    pl.low_bits = 4;
    pl.seed = 11;
    pl.verbosity = 0;
    uint64_t p0 = 0, g0 = 0, p1 = 0, g1 = 0;
    qsim_hip_get_gate_queue(nullptr, nullptr, &p0, &g0);
    auto got = RunWith<FactoryHIP<float>, Learned>(FactoryHIP<float>(), c32,
                                                   pl);
    qsim_hip_get_gate_queue(nullptr, nullptr, &p1, &g1);
    qsim_hip_set_gate_queue(0, 0);
    const double e = RelL2(got.first, ref.first);
    const bool same_bits = got.second == ref.second;
    // The fused gates came back as blocked passes (several gates each).
    const bool blocked = p1 > p0 && g1 - g0 >= 2 * (p1 - p0);
    ok = ok && e < 1e-5 && same_bits && blocked;
    d << (mid ? ", " : "") << "{\"mid_measurement\": " << (mid ? "true" : "false")
      << ", \"rel_l2_vs_fp64\": " << e << ", \"blocked_passes\": " << (p1 - p0)
      << ", \"gates_in_blocked_passes\": " << (g1 - g0)
      << ", \"measured_bits_equal\": " << (same_bits ? "true" : "false") << "}";
  }
  d << "]";
  Report("learned", ok, d.str());
}

void Scenario_LearnedModel(const char* model_path) {
  void* model = nullptr;
  if (qsim_hip_costmodel_load(model_path, &model) != QSIM_HIP_OK) {
    Report("learned-model", false,
           std::string("\"error\": \"") + qsim_hip_last_error() + "\"");
    return;
  }
  qsim_hip_costmodel_set_default(model);
  const std::string text = GridText(4, 6, 12, 5, false);
  auto c = Parse<float>(text);
  // The unfused run: every gate of the circuit on its own.
  FactoryHIP<float> f;
  auto ss = f.CreateStateSpace();
  auto sim = f.CreateSimulator();
  auto s = ss.Create(c.num_qubits);
  ss.SetStateZero(s);
  for (const auto& op : c.ops) ApplyGate(sim, op, s);
  auto ref = Amplitudes(f, s);
  using Learned = LearnedFuserHIP<IO>;
  typename QSimRunner<IO, Learned, FactoryHIP<float>>::Parameter pl;
  pl.seed = 1;
  pl.verbosity = 0;  // the library's default model
  unsigned tb = 0, lb = 0;
  uint64_t p0 = 0, g0 = 0, p1 = 0, g1 = 0;
  qsim_hip_get_gate_queue(nullptr, nullptr, &p0, &g0);
  auto got = RunWith<FactoryHIP<float>, Learned>(f, c, pl);
  qsim_hip_get_gate_queue(&tb, &lb, &p1, &g1);
  qsim_hip_set_gate_queue(0, 0);
  qsim_hip_costmodel_free(model);
  const double e = RelL2(got.first, ref);
  std::ostringstream d;
  d << "\"rel_l2_vs_unfused\": " << e << ", \"tile_bits\": " << tb
    << ", \"low_bits\": " << lb << ", \"blocked_passes\": " << (p1 - p0)
    << ", \"gates_in_blocked_passes\": " << (g1 - g0);
  Report("learned-model", e < 1e-5, d.str());
}

int main(int argc, char** argv) {
  Scenario_Runner();
  Scenario_Callback();
  Scenario_Expect();
  Scenario_Trajectories();
  Scenario_Hybrid();
  Scenario_Sample();
  Scenario_Learned();
  if (argc > 1) Scenario_LearnedModel(argv[1]);
  return g_failures == 0 ? 0 : 1;
}
