// trajectories.cc -- host side of the batched trajectory engine
// (include/qsim_hip_driver.h, qsim_traj_*).
//
// For every trajectory this replays qsim's RunIteration
// (lib/qtrajectory.h:240-387) on the host as far as the host can: the same
// mt19937(rep) stream with one uniform draw per operation, the same walk over
// each channel's Kraus operators against their probability bounds, the same
// deferral of the chosen operators.  Runs of deferred gates are fused with
// qsim's MultiQubitGateFuser (FP64 products) into the trajectory's operation
// stream -- or, for the tiled kernel, partitioned into tile passes by the
// blocked-pass planner (qsim_amd/block_planner.h): a pass header followed by
// its <= 4-qubit gates.  Where the bounds do not decide (an amplitude-damping draw past
// 1 - gamma) the stream gets a decision operation that the device takes from
// the state itself.  All streams then run in one launch
// (qsim_hip_traj_run).
//
// A noisy circuit is a sequence of gates and qsim channels (Channel,
// lib/channel.h:30-125), either
//  - built from a qsim circuit by one fixed rule (qsim_traj_create): after
//    every gate, a depolarising channel (DepolarizingChannel,
//    lib/channels_cirq.h:127-200) on each of its qubits; after the last gate
//    of every cycle (gate times 2c and 2c + 1), an amplitude-damping channel
//    (AmplitudeDampingChannel, lib/channels_cirq.h:274-300) on every qubit;
//  - or read from noisy-circuit text (qsim_traj_create_noisy): qsim's
//    circuit lines, and channel lines built by qsim's own channel factories
//    (lib/channels_cirq.h) or given as explicit Kraus operators.
// Any channel's operators go through the same sampling; those the bounds
// do not decide are decided on the device (QSIM_HIP_TRAJ_CHANNEL), for
// channels on at most 2 qubits with at most 16 operators.

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <set>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <new>
#include <random>
#include <sstream>
#include <string>
#include <unordered_map>
#include <variant>
#include <vector>

#include "channel.h"
#include "channels_cirq.h"
#include "circuit_qsim_parser.h"
#include "fuser_mqubit.h"
#include "io.h"
#include "qsim_amd/block_planner.h"
#include "qsim_hip.h"
#include "qsim_hip_driver.h"

namespace {

using namespace qsim;

struct NoisyOp {
  int kind;  // 0 gate (index into the circuit), 1 channel (index into chans)
  size_t index;
  unsigned time;  // gate: the time of its deferred copy
};

// Most Kraus operators and qubits of a channel that the device decides
// (traj_kernels.hip).
constexpr unsigned kMaxDeviceKraus = 16;
constexpr unsigned kMaxDeviceChannelQubits = 2;

// A channel in the form the device decides it in: its qubits and, per Kraus
// operator, K on the operator's own qubits (the product of its gates, FP64,
// rounded to FP32; empty for the identity) and K^dagger K over the
// channel's qubits (FP32; empty for a unitary operator).
struct DeviceChannel {
  bool ok = false;  // within the device's limits
  std::vector<unsigned> qubits;
  std::vector<std::vector<unsigned>> kq;
  std::vector<std::vector<float>> k, kdk;
};

struct NoisyCircuit {
  unsigned num_qubits = 0;
  Circuit<Operation<double>> circuit;
  std::vector<Channel<double>> chans;
  std::vector<DeviceChannel> dchans;  // per channel
  std::vector<NoisyOp> ops;
  std::vector<uint32_t> pairs;  // grid-neighbour pairs (a, b)
  unsigned num_channels = 0;
};

// K^dagger K of `kop` (over kop.qubits) as a matrix over `qubits` (which
// contain them): the tensor product with the identity on the others.
std::vector<double> ExpandKdK(const KrausOperator<double>& kop,
                              const std::vector<unsigned>& qubits) {
  const unsigned nq = qubits.size(), D = 1u << nq;
  unsigned pos[kMaxDeviceChannelQubits] = {0, 0};
  for (unsigned j = 0; j < kop.qubits.size(); ++j) {
    for (unsigned b = 0; b < nq; ++b)
      if (qubits[b] == kop.qubits[j]) pos[j] = b;
  }
  const unsigned kd = 1u << kop.qubits.size();
  unsigned mask = 0;
  for (unsigned j = 0; j < kop.qubits.size(); ++j) mask |= 1u << pos[j];
  auto sub = [&](unsigned x) {  // x's bits at the operator's qubits
    unsigned v = 0;
    for (unsigned j = 0; j < kop.qubits.size(); ++j) v |= ((x >> pos[j]) & 1u) << j;
    return v;
  };
  std::vector<double> m(2 * size_t(D) * D, 0.0);
  for (unsigned r = 0; r < D; ++r) {
    for (unsigned c = 0; c < D; ++c) {
      if ((r & ~mask) != (c & ~mask)) continue;
      const unsigned rr = sub(r), cc = sub(c);
      m[2 * (r * D + c)] = kop.kd_k[2 * (rr * kd + cc)];
      m[2 * (r * D + c) + 1] = kop.kd_k[2 * (rr * kd + cc) + 1];
    }
  }
  return m;
}

DeviceChannel MakeDeviceChannel(const Channel<double>& ch) {
  DeviceChannel dc;
  std::set<unsigned> qs;
  for (const auto& kop : ch.kops) {
    for (const auto& g : kop.ops) qs.insert(g.qubits.begin(), g.qubits.end());
    if (!kop.unitary) qs.insert(kop.qubits.begin(), kop.qubits.end());
  }
  dc.qubits.assign(qs.begin(), qs.end());
  if (dc.qubits.empty() || dc.qubits.size() > kMaxDeviceChannelQubits ||
      ch.kops.empty() || ch.kops.size() > kMaxDeviceKraus) {
    return dc;
  }
  for (const auto& kop : ch.kops) {
    std::set<unsigned> kset;
    for (const auto& g : kop.ops) kset.insert(g.qubits.begin(), g.qubits.end());
    std::vector<unsigned> kq(kset.begin(), kset.end());
    std::vector<float> k;
    if (!kop.ops.empty()) {
      // The operator's gates as one fused gate (lib/fuser.h:149-179).
      FusedGate<double> fg{};
      fg.kind = kop.ops[0].kind;
      fg.time = kop.ops[0].time;
      fg.qubits = kq;
      fg.parent = &kop.ops[0];
      for (const auto& g : kop.ops) fg.gates.push_back(&g);
      CalculateFusedMatrix(fg);
      k.assign(fg.matrix.begin(), fg.matrix.end());
    }
    std::vector<float> kdk;
    if (!kop.unitary) {
      if (kop.kd_k.empty() || kop.qubits.empty()) return dc;
      const auto m = ExpandKdK(kop, dc.qubits);
      kdk.assign(m.begin(), m.end());
    }
    dc.kq.push_back(std::move(kq));
    dc.k.push_back(std::move(k));
    dc.kdk.push_back(std::move(kdk));
  }
  dc.ok = true;
  return dc;
}

void GridPairs(unsigned rows, unsigned cols, NoisyCircuit& nc) {
  for (unsigned r = 0; r < rows; ++r) {
    for (unsigned c = 0; c < cols; ++c) {
      if (c + 1 < cols) {
        nc.pairs.push_back(r * cols + c);
        nc.pairs.push_back(r * cols + c + 1);
      }
      if (r + 1 < rows) {
        nc.pairs.push_back(r * cols + c);
        nc.pairs.push_back((r + 1) * cols + c);
      }
    }
  }
}

void AddChannel(Channel<double>&& ch, NoisyCircuit& nc) {
  nc.ops.push_back({1, nc.chans.size(), 0});
  nc.dchans.push_back(MakeDeviceChannel(ch));
  nc.chans.push_back(std::move(ch));
}

// One trajectory's plan.
struct TrajPlan {
  unsigned circuit = 0;
  uint64_t rep = 0;
  std::vector<qsim_hip_traj_op> ops;
  std::vector<std::vector<float>> mats;  // per gate op, FP32 matrix
  std::vector<uint64_t> mhash;           // per op, MatrixHash of mats[i]
  std::vector<int32_t> kraus;            // per channel: index, or -2 - slot
  unsigned num_decisions = 0;
  bool normalize_last = false;           // non-unitary at the end
};

struct TrajBatchHost {
  std::vector<NoisyCircuit> circuits;
  double p = 0, gamma = 0;
  std::vector<TrajPlan> plans;      // caller order
  std::vector<unsigned> order;      // device order -> caller index
  std::vector<uint32_t> pair_off;   // caller order
  std::vector<uint32_t> dec_off;    // caller order
  void* handle = nullptr;
  uint32_t total_pairs = 0, total_decisions = 0;
  std::vector<double> zz;
  std::vector<int32_t> dec;
  bool have_results = false;
};

bool Build(const std::string& text, unsigned rows, unsigned cols, double p,
           double gamma, NoisyCircuit& nc) {
  std::stringstream ss(text);
  if (!CircuitQsimParser<IO>::FromStream(~0u, "circuit", ss, nc.circuit)) {
    return false;
  }
  const unsigned n = nc.circuit.num_qubits;
  if (rows * cols != n) return false;
  nc.num_qubits = n;
  const auto& ops = nc.circuit.ops;
  for (size_t i = 0; i < ops.size(); ++i) {
    const Gate<double>* g = OpGetAlternative<Gate<double>>(ops[i]);
    if (g == nullptr || g->qubits.empty() || g->qubits.size() > 4) {
      return false;
    }
    // Deferred gates carry time 3t, so that the gates of a trajectory stay
    // time-ordered per qubit with the depolarising Paulis (3t + 1) and the
    // damping operators (3(2c + 1) + 2).
    const unsigned t = g->time;
    nc.ops.push_back({0, i, 3 * t});
    for (unsigned q : g->qubits) {
      AddChannel(Cirq::DepolarizingChannel<double>::Create(3 * t + 1, q, p),
                 nc);
    }
    const bool cycle_end =
        i + 1 == ops.size() || OpTime(ops[i + 1]) / 2 != t / 2;
    if (cycle_end) {
      const unsigned c = t / 2;
      for (unsigned q = 0; q < n; ++q) {
        AddChannel(Cirq::AmplitudeDampingChannel<double>::Create(
                       3 * (2 * c + 1) + 2, q, gamma),
                   nc);
      }
    }
  }
  nc.num_channels = nc.chans.size();
  GridPairs(rows, cols, nc);
  return true;
}

// A channel line of noisy-circuit text, its first two fields (time and the
// word "channel") already read:
//   <name> <qubit> <params>   name: depolarize p | asymmetric_depolarize
//       px py pz | amplitude_damp gamma | generalized_amplitude_damp p gamma
//       | phase_damp gamma | phase_flip p | bit_flip p | reset  (qsim's
//       factories, lib/channels_cirq.h)
//   kraus <nq> <qubits, ascending> <number of operators>, then per
//       operator: <unitary 0|1> <probability bound> <2 4^nq numbers: the
//       matrix, row-major complex, bit j of its index <-> the j-th qubit>
bool ParseChannel(unsigned time, unsigned n, std::istringstream& in,
                  Channel<double>& ch) {
  std::string name;
  unsigned q = 0;
  if (!(in >> name)) return false;
  if (name == "kraus") {
    unsigned nq = 0, nk = 0;
    if (!(in >> nq) || nq < 1 || nq > 4) return false;
    std::vector<unsigned> qs(nq);
    for (unsigned j = 0; j < nq; ++j) {
      if (!(in >> qs[j]) || qs[j] >= n || (j > 0 && qs[j] <= qs[j - 1])) {
        return false;
      }
    }
    if (!(in >> nk) || nk < 1 || nk > 64) return false;
    ch = Channel<double>{{kChannel, time, qs}, {}};
    const size_t len = size_t{2} << (2 * nq);
    for (unsigned i = 0; i < nk; ++i) {
      int unitary = 0;
      double prob = 0;
      Matrix<double> m(len);
      if (!(in >> unitary >> prob)) return false;
      for (auto& v : m) {
        if (!(in >> v)) return false;
      }
      KrausOperator<double> kop;
      kop.unitary = unitary != 0;
      kop.prob = prob;
      kop.ops.push_back(CreateGate<Gate<double>, Cirq::MatrixGate<double>>(
          time, qs, std::move(m)));
      kop.CalculateKdKMatrix();
      ch.kops.push_back(std::move(kop));
    }
    return true;
  }
  if (!(in >> q) || q >= n) return false;
  double a = 0, b = 0, c = 0;
  if (name == "depolarize" && (in >> a)) {
    ch = Cirq::DepolarizingChannel<double>::Create(time, q, a);
  } else if (name == "asymmetric_depolarize" && (in >> a >> b >> c)) {
    ch = Cirq::AsymmetricDepolarizingChannel<double>::Create(time, q, a, b, c);
  } else if (name == "amplitude_damp" && (in >> a)) {
    ch = Cirq::AmplitudeDampingChannel<double>::Create(time, q, a);
  } else if (name == "generalized_amplitude_damp" && (in >> a >> b)) {
    ch = Cirq::GeneralizedAmplitudeDampingChannel<double>::Create(time, q, a,
                                                                 b);
  } else if (name == "phase_damp" && (in >> a)) {
    ch = Cirq::PhaseDampingChannel<double>::Create(time, q, a);
  } else if (name == "phase_flip" && (in >> a)) {
    ch = Cirq::PhaseFlipChannel<double>::Create(time, q, a);
  } else if (name == "bit_flip" && (in >> a)) {
    ch = Cirq::BitFlipChannel<double>::Create(time, q, a);
  } else if (name == "reset") {
    ch = Cirq::ResetChannel<double>::Create(time, q);
  } else {
    return false;
  }
  return true;
}

// Noisy-circuit text: qsim's circuit format (lib/circuit_qsim_parser.h)
// with channel lines ("<time> channel ...", ParseChannel) among the gate
// lines, in the order the operations apply.  Gates keep their times.
bool BuildNoisy(const std::string& text, unsigned rows, unsigned cols,
                NoisyCircuit& nc) {
  std::istringstream lines(text);
  std::string line;
  unsigned n = 0;
  while (std::getline(lines, line)) {
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
    std::istringstream in(line);
    if (!(in >> n) || n == 0 || n > 34) return false;
    break;
  }
  if (n == 0 || rows * cols != n) return false;
  nc.num_qubits = n;
  nc.circuit.num_qubits = n;
  while (std::getline(lines, line)) {
    const size_t first = line.find_first_not_of(" \t\r");
    if (first == std::string::npos || line[first] == '#') continue;
    std::istringstream in(line);
    unsigned time = 0;
    std::string word;
    if (!(in >> time >> word)) return false;
    if (word == "channel") {
      Channel<double> ch;
      if (!ParseChannel(time, n, in, ch)) return false;
      std::string extra;
      if (in >> extra) return false;
      AddChannel(std::move(ch), nc);
      continue;
    }
    // A gate line, through qsim's parser.
    Circuit<Operation<double>> one;
    std::stringstream gs;
    gs << n << "\n" << line << "\n";
    if (!CircuitQsimParser<IO>::FromStream(~0u, "circuit", gs, one) ||
        one.ops.size() != 1) {
      return false;
    }
    const Gate<double>* g = OpGetAlternative<Gate<double>>(one.ops[0]);
    if (g == nullptr || g->qubits.empty() || g->qubits.size() > 4) {
      return false;
    }
    nc.ops.push_back({0, nc.circuit.ops.size(), g->time});
    nc.circuit.ops.push_back(std::move(one.ops[0]));
  }
  nc.num_channels = nc.chans.size();
  GridPairs(rows, cols, nc);
  return true;
}

// A gate of the noisy circuit with its time remapped so that the deferred
// gates of a trajectory are time-ordered per qubit (gate 3t, depolarising
// Paulis 3t + 1, damping 3(2c + 1) + 2).
Gate<double> Timed(const Gate<double>& g, unsigned time) {
  Gate<double> h = g;
  h.time = time;
  return h;
}

// How a trajectory's runs of deferred gates become operations: fused gates
// of at most max_fused qubits, or (tile_bits >= 9) tile passes of
// tile_bits qubits, low_bits of them the lowest.
struct FusionSpec {
  unsigned max_fused = 4;
  unsigned tile_bits = 0;
  unsigned low_bits = 5;
};

// Fused matrices by fused gate: its qubits and its component gates' qubits
// and matrices.  The trajectories of one circuit share most of their fused
// gates (they differ where a Kraus operator other than the likeliest was
// drawn), so qsim's product (CalculateFusedMatrix, lib/fuser.h:150-179) runs
// once per distinct gate and the others are copies of its result.  One
// cache per planning thread.
class FusedMatrixCache {
 public:
  void Fill(FusedGate<double>& fg) {
    key_.clear();
    Put(fg.qubits.data(), fg.qubits.size() * sizeof(unsigned));
    for (const auto& pgate : fg.gates) {
      const auto* pg = OpGetAlternative<Gate<double>>(pgate);
      const auto& pq = OpQubits(pgate);
      const auto& pm =
          pg ? pg->matrix : OpGetAlternative<DecomposedGate<double>>(pgate)->matrix;
      const size_t nq = pq.size(), nm = pm.size();
      Put(&nq, sizeof(nq));
      Put(pq.data(), nq * sizeof(unsigned));
      Put(&nm, sizeof(nm));
      Put(pm.data(), nm * sizeof(double));
    }
    auto it = memo_.find(key_);
    if (it != memo_.end()) {
      fg.matrix = it->second;
      return;
    }
    CalculateFusedMatrix(fg);
    if (memo_.size() >= kMaxEntries) memo_.clear();
    memo_.emplace(key_, fg.matrix);
  }

 private:
  static constexpr size_t kMaxEntries = 1 << 16;
  void Put(const void* p, size_t bytes) {
    key_.append(static_cast<const char*>(p), bytes);
  }
  std::string key_;
  std::unordered_map<std::string, Matrix<double>> memo_;
};

// Tile passes for a run of gates; false if the state is too small for them.
bool FlushTiled(const std::vector<Operation<double>>& deferred, unsigned n,
                const FusionSpec& spec, FusedMatrixCache& cache,
                TrajPlan& plan) {
  hip::BlockPlannerParameter bp;
  bp.tile_bits = spec.tile_bits;
  bp.low_bits = spec.low_bits;
  // No more fused gates per pass than the tile kernel keeps resident.
  bp.max_ops = qsim_hip_block_resident_ops(std::min(spec.tile_bits, n));
  hip::BlockPlan blocks;
  const std::function<void(FusedGate<double>&)> fill =
      [&cache](FusedGate<double>& fg) { cache.Fill(fg); };
  if (!hip::PlanBlocks<IO>(bp, n, deferred, blocks, &fill)) return false;
  for (const auto& step : blocks.steps) {
    if (!step.blocked) return false;  // only gates are deferred
  }
  for (const auto& pass : blocks.blocks) {
    qsim_hip_traj_op h{};
    h.kind = QSIM_HIP_TRAJ_TILE_PASS;
    h.k = pass.ks.size();
    uint64_t mask = 0;
    for (unsigned q : pass.tile) mask |= uint64_t{1} << q;
    h.q[0] = uint32_t(mask);
    h.q[1] = uint32_t(mask >> 32);
    plan.ops.push_back(h);
    plan.mats.emplace_back();  // no matrix of its own
    size_t mo = 0;
    for (size_t j = 0; j < pass.ks.size(); ++j) {
      qsim_hip_traj_op o{};
      o.kind = QSIM_HIP_TRAJ_GATE;
      o.k = pass.ks[j];
      for (unsigned b = 0; b < o.k; ++b) o.q[b] = pass.qs[4 * j + b];
      const size_t nf = size_t{2} << (2 * o.k);
      plan.ops.push_back(o);
      plan.mats.emplace_back(pass.m.begin() + mo, pass.m.begin() + mo + nf);
      mo += nf;
    }
  }
  return true;
}

void Flush(std::vector<Operation<double>>& deferred, unsigned n,
           const FusionSpec& spec, FusedMatrixCache& cache, TrajPlan& plan) {
  if (deferred.empty()) return;
  if (spec.tile_bits >= 9) {
    const size_t nops = plan.ops.size();
    if (FlushTiled(deferred, n, spec, cache, plan)) {
      deferred.clear();
      return;
    }
    plan.ops.resize(nops);
    plan.mats.resize(nops);
  }
  const unsigned max_fused = spec.max_fused;
  MultiQubitGateFuser<IO>::Parameter fp;
  fp.max_fused_size = max_fused;
  fp.verbosity = 0;
  auto fused = MultiQubitGateFuser<IO>::FuseGates(fp, n, deferred, false);
  for (auto& f : fused) {
    auto* fg = std::get_if<FusedGate<double>>(&f);
    if (fg == nullptr) continue;
    cache.Fill(*fg);
    qsim_hip_traj_op o{};
    o.kind = QSIM_HIP_TRAJ_GATE;
    o.k = fg->qubits.size();
    for (unsigned b = 0; b < o.k && b < 4; ++b) o.q[b] = fg->qubits[b];
    plan.ops.push_back(o);
    plan.mats.emplace_back(fg->matrix.begin(), fg->matrix.end());
  }
  deferred.clear();
}

// 64-bit hash of an FP32 matrix's bits (for sharing identical matrices
// across the batch; equal hashes are confirmed by comparing the bits).
uint64_t MatrixHash(const std::vector<float>& m) {
  uint64_t h = 0x9e3779b97f4a7c15ull ^ m.size();
  for (float f : m) {
    uint32_t u;
    std::memcpy(&u, &f, sizeof(u));
    h = (h ^ u) * 0xff51afd7ed558ccdull;
    h ^= h >> 33;
  }
  return h;
}

// RunIteration's sampling (lib/qtrajectory.h:270-370) for one trajectory;
// false if a channel the bounds do not decide is beyond the device's
// limits.
bool PlanTrajectory(const NoisyCircuit& nc, uint64_t rep,
                    const FusionSpec& spec, FusedMatrixCache& cache,
                    TrajPlan& plan) {
  const unsigned n = nc.num_qubits;
  std::mt19937 rgen(rep);
  std::uniform_real_distribution<double> distr(0.0, 1.0);
  std::vector<Operation<double>> deferred;
  plan.ops.clear();
  plan.mats.clear();
  plan.kraus.clear();
  plan.num_decisions = 0;
  bool unitary = true;
  for (const NoisyOp& op : nc.ops) {
    const double r = distr(rgen);
    if (op.kind == 0) {
      const auto& g = *OpGetAlternative<Gate<double>>(nc.circuit.ops[op.index]);
      deferred.push_back(Timed(g, op.time));
      continue;
    }
    const Channel<double>& ch = nc.chans[op.index];
    double cp = 0;
    int picked = -1;
    for (size_t i = 0; i < ch.kops.size(); ++i) {
      cp += ch.kops[i].prob;
      if (r < cp) {
        picked = int(i);
        break;
      }
    }
    if (picked >= 0) {
      for (const auto& g : ch.kops[picked].ops) deferred.push_back(g);
      plan.kraus.push_back(picked);
      unitary = unitary && ch.kops[picked].unitary;
      continue;
    }
    // The bounds do not decide: the state does, on the device (the walk
    // over the operators' <psi|K^dagger K|psi> of lib/qtrajectory.h:330-370,
    // on the normalised state where the trajectory is not unitary).
    const DeviceChannel& dc = nc.dchans[op.index];
    if (!dc.ok) return false;
    Flush(deferred, n, spec, cache, plan);
    qsim_hip_traj_op h{};
    h.kind = QSIM_HIP_TRAJ_CHANNEL;
    h.k = ch.kops.size();
    h.moff2 = dc.qubits.size();
    for (unsigned b = 0; b < dc.qubits.size(); ++b) h.q[b] = dc.qubits[b];
    h.r = r;
    h.cp = cp;
    h.slot = plan.num_decisions++;
    h.pad = unitary ? 0 : 1;
    plan.ops.push_back(h);
    plan.mats.emplace_back();  // no matrix of its own
    for (size_t i = 0; i < ch.kops.size(); ++i) {
      qsim_hip_traj_op ko{};
      ko.kind = QSIM_HIP_TRAJ_KRAUS;
      ko.k = dc.k[i].empty() ? 0 : dc.kq[i].size();
      for (unsigned b = 0; b < ko.k; ++b) ko.q[b] = dc.kq[i][b];
      ko.r = ch.kops[i].prob;
      ko.pad = ch.kops[i].unitary ? 1 : 0;
      // K then K^dagger K; their offsets are set when the batch is
      // assembled (moff2 - moff = the size of K).
      std::vector<float> m(dc.k[i]);
      m.insert(m.end(), dc.kdk[i].begin(), dc.kdk[i].end());
      ko.moff2 = dc.k[i].size();
      plan.ops.push_back(ko);
      plan.mats.push_back(std::move(m));
    }
    plan.kraus.push_back(-2 - int32_t(h.slot));
    // A chosen operator makes the state non-unitary; where none is chosen
    // (possible only if the last operator is unitary) the state was just
    // normalised.  Either way the next decision normalising first is
    // harmless, so the flag is cleared.
    unitary = false;
  }
  Flush(deferred, n, spec, cache, plan);
  plan.normalize_last = !unitary;
  plan.mhash.resize(plan.mats.size());
  for (size_t i = 0; i < plan.mats.size(); ++i) {
    plan.mhash[i] = MatrixHash(plan.mats[i]);
  }
  return true;
}

}  // namespace

extern "C" {

void* qsim_traj_create(const char* const* circuits, unsigned num_circuits,
                       const unsigned* rows, const unsigned* cols,
                       double p_depol, double gamma_damp) {
  if (circuits == nullptr || rows == nullptr || cols == nullptr ||
      num_circuits == 0 || p_depol < 0 || p_depol > 1 || gamma_damp < 0 ||
      gamma_damp > 1) {
    return nullptr;
  }
  try {
    auto b = std::make_unique<TrajBatchHost>();
    b->p = p_depol;
    b->gamma = gamma_damp;
    b->circuits.resize(num_circuits);
    for (unsigned i = 0; i < num_circuits; ++i) {
      if (circuits[i] == nullptr ||
          !Build(circuits[i], rows[i], cols[i], p_depol, gamma_damp,
                 b->circuits[i])) {
        return nullptr;
      }
    }
    return b.release();
  } catch (const std::exception&) {  // qsim's parser and allocations
    return nullptr;
  }
}

void* qsim_traj_create_noisy(const char* const* texts, unsigned num_circuits,
                             const unsigned* rows, const unsigned* cols) {
  if (texts == nullptr || rows == nullptr || cols == nullptr ||
      num_circuits == 0) {
    return nullptr;
  }
  try {
    auto b = std::make_unique<TrajBatchHost>();
    b->circuits.resize(num_circuits);
    for (unsigned i = 0; i < num_circuits; ++i) {
      if (texts[i] == nullptr ||
          !BuildNoisy(texts[i], rows[i], cols[i], b->circuits[i])) {
        return nullptr;
      }
    }
    return b.release();
  } catch (const std::exception&) {  // qsim's parser and allocations
    return nullptr;
  }
}

unsigned qsim_traj_num_channels(void* batch, unsigned circuit) {
  auto* b = static_cast<TrajBatchHost*>(batch);
  return circuit < b->circuits.size() ? b->circuits[circuit].num_channels : 0;
}

}  // extern "C"

namespace {

int PlanBatchImpl(TrajBatchHost* b, const unsigned* circuit_of,
                  const uint64_t* reps, unsigned num_traj,
                  const FusionSpec& spec) {
  if (b == nullptr || circuit_of == nullptr || reps == nullptr ||
      num_traj == 0) {
    return QSIM_HIP_ERR_ARG;
  }
  for (unsigned t = 0; t < num_traj; ++t) {
    if (circuit_of[t] >= b->circuits.size()) return QSIM_HIP_ERR_ARG;
  }
  if (b->handle) {
    qsim_hip_traj_free(b->handle);
    b->handle = nullptr;
  }
  b->have_results = false;
  b->plans.assign(num_traj, TrajPlan{});
  std::vector<FusedMatrixCache> caches(omp_get_max_threads());
  // An exception may not leave the parallel region: a failed host
  // allocation is caught per trajectory and returned as an error.
  std::atomic<bool> out_of_memory{false}, beyond_device{false};
#pragma omp parallel for schedule(dynamic, 4)
  for (int t = 0; t < int(num_traj); ++t) {
    b->plans[t].circuit = circuit_of[t];
    b->plans[t].rep = reps[t];
    try {
      if (!PlanTrajectory(b->circuits[circuit_of[t]], reps[t], spec,
                          caches[omp_get_thread_num()], b->plans[t])) {
        beyond_device = true;
      }
    } catch (const std::bad_alloc&) {
      out_of_memory = true;
    }
  }
  if (out_of_memory) {
    b->plans.clear();
    return QSIM_HIP_ERR_RUNTIME;
  }
  if (beyond_device) {
    // A channel on more than 2 qubits or with more than 16 operators whose
    // draw its bounds did not decide.
    b->plans.clear();
    return QSIM_HIP_ERR_UNSUPPORTED;
  }
  // Device order: largest states first.
  b->order.resize(num_traj);
  for (unsigned t = 0; t < num_traj; ++t) b->order[t] = t;
  std::stable_sort(b->order.begin(), b->order.end(), [&](unsigned x, unsigned y) {
    return b->circuits[b->plans[x].circuit].num_qubits >
           b->circuits[b->plans[y].circuit].num_qubits;
  });
  // Concatenate streams; identical matrices share one copy (found by their
  // hashes, computed in the parallel planning, and confirmed bit for bit).
  std::vector<float> pool;
  std::unordered_map<uint64_t, std::vector<uint32_t>> seen;  // hash -> offsets
  std::vector<qsim_hip_traj_op> ops;
  std::vector<qsim_hip_traj_desc> descs;
  std::vector<uint32_t> pairs;
  b->pair_off.assign(num_traj, 0);
  b->dec_off.assign(num_traj, 0);
  uint32_t npairs = 0, ndec = 0;
  for (unsigned t = 0; t < num_traj; ++t) {
    b->pair_off[t] = npairs;
    b->dec_off[t] = ndec;
    const auto& nc = b->circuits[b->plans[t].circuit];
    pairs.insert(pairs.end(), nc.pairs.begin(), nc.pairs.end());
    npairs += nc.pairs.size() / 2;
    ndec += b->plans[t].num_decisions;
  }
  for (unsigned t : b->order) {
    TrajPlan& pl = b->plans[t];
    const auto& nc = b->circuits[pl.circuit];
    qsim_hip_traj_desc d{};
    d.num_qubits = nc.num_qubits;
    d.num_pairs = nc.pairs.size() / 2;
    d.pair_offset = b->pair_off[t];
    d.pad = pl.normalize_last ? 1 : 0;
    d.op_begin = ops.size();
    for (size_t i = 0; i < pl.ops.size(); ++i) {
      qsim_hip_traj_op o = pl.ops[i];
      if (o.kind == QSIM_HIP_TRAJ_GATE) {
        const auto& m = pl.mats[i];
        const size_t bytes = m.size() * sizeof(float);
        auto& offs = seen[pl.mhash[i]];
        uint32_t off = ~0u;
        for (uint32_t c : offs) {
          if (c + m.size() <= pool.size() &&
              std::memcmp(pool.data() + c, m.data(), bytes) == 0) {
            off = c;
            break;
          }
        }
        if (off == ~0u) {
          off = uint32_t(pool.size());
          offs.push_back(off);
          pool.insert(pool.end(), m.begin(), m.end());
        }
        o.moff = off;
      } else if (o.kind == QSIM_HIP_TRAJ_TILE_PASS) {
        o.moff = 0;
      } else if (o.kind == QSIM_HIP_TRAJ_CHANNEL) {
        o.slot += b->dec_off[t];
      } else if (o.kind == QSIM_HIP_TRAJ_KRAUS) {
        // K at moff, K^dagger K right after it (the plan's moff2 holds the
        // size of K).
        const auto& m = pl.mats[i];
        const uint32_t ksize = o.moff2;
        o.moff = uint32_t(pool.size());
        o.moff2 = o.moff + ksize;
        pool.insert(pool.end(), m.begin(), m.end());
      }
      ops.push_back(o);
    }
    d.op_end = ops.size();
    descs.push_back(d);
    pl.ops.clear();
    pl.mats.clear();
    pl.mhash.clear();
  }
  b->total_pairs = npairs;
  b->total_decisions = ndec;
  return qsim_hip_traj_prepare(descs.data(), descs.size(), ops.data(),
                               ops.size(), pool.data(), pool.size(),
                               pairs.data(), npairs, ndec, &b->handle);
}

// A failed host allocation while planning becomes an error return (an
// exception may not cross the C ABI); the batch keeps no partial plan.
int PlanBatch(TrajBatchHost* b, const unsigned* circuit_of,
              const uint64_t* reps, unsigned num_traj,
              const FusionSpec& spec) {
  try {
    return PlanBatchImpl(b, circuit_of, reps, num_traj, spec);
  } catch (const std::bad_alloc&) {
    if (b != nullptr) b->plans.clear();
    return QSIM_HIP_ERR_RUNTIME;
  }
}

}  // namespace

extern "C" {

int qsim_traj_plan(void* batch, const unsigned* circuit_of,
                   const uint64_t* reps, unsigned num_traj,
                   unsigned max_fused_size) {
  if (max_fused_size < 1 || max_fused_size > 4) return QSIM_HIP_ERR_ARG;
  FusionSpec spec;
  spec.max_fused = max_fused_size;
  return PlanBatch(static_cast<TrajBatchHost*>(batch), circuit_of, reps,
                   num_traj, spec);
}

int qsim_traj_plan_tiled(void* batch, const unsigned* circuit_of,
                         const uint64_t* reps, unsigned num_traj,
                         unsigned tile_bits, unsigned low_bits) {
  if (tile_bits < 9 || tile_bits > 14 || low_bits < 3 || low_bits > 9 ||
      low_bits > tile_bits) {
    return QSIM_HIP_ERR_ARG;
  }
  FusionSpec spec;
  spec.tile_bits = tile_bits;
  spec.low_bits = low_bits;
  return PlanBatch(static_cast<TrajBatchHost*>(batch), circuit_of, reps,
                   num_traj, spec);
}

int qsim_traj_run(void* batch, unsigned threads, int use_graph) {
  auto* b = static_cast<TrajBatchHost*>(batch);
  if (b == nullptr || b->handle == nullptr) return QSIM_HIP_ERR_ARG;
  b->have_results = false;
  return qsim_hip_traj_run(b->handle, threads, use_graph);
}

static int Fetch(TrajBatchHost* b) {
  if (b->have_results) return QSIM_HIP_OK;
  try {
    b->zz.assign(b->total_pairs, 0.0);
    b->dec.assign(b->total_decisions, -1);
  } catch (const std::bad_alloc&) {
    return QSIM_HIP_ERR_RUNTIME;
  }
  int rc = qsim_hip_traj_results(b->handle, b->zz.data(), b->total_pairs,
                                 b->dec.data(), b->total_decisions);
  if (rc == QSIM_HIP_OK) b->have_results = true;
  return rc;
}

int qsim_traj_zz(void* batch, unsigned t, double* zz, unsigned max,
                 unsigned* num) {
  auto* b = static_cast<TrajBatchHost*>(batch);
  if (b == nullptr || b->handle == nullptr || t >= b->plans.size()) {
    return QSIM_HIP_ERR_ARG;
  }
  int rc = Fetch(b);
  if (rc != QSIM_HIP_OK) return rc;
  const unsigned np = b->circuits[b->plans[t].circuit].pairs.size() / 2;
  for (unsigned i = 0; i < np && i < max; ++i) zz[i] = b->zz[b->pair_off[t] + i];
  if (num) *num = np;
  return QSIM_HIP_OK;
}

int qsim_traj_kraus(void* batch, unsigned t, int32_t* kidx, unsigned max,
                    unsigned* num) {
  auto* b = static_cast<TrajBatchHost*>(batch);
  if (b == nullptr || b->handle == nullptr || t >= b->plans.size()) {
    return QSIM_HIP_ERR_ARG;
  }
  int rc = Fetch(b);
  if (rc != QSIM_HIP_OK) return rc;
  const auto& k = b->plans[t].kraus;
  for (size_t i = 0; i < k.size() && i < max; ++i) {
    kidx[i] = k[i] >= 0 ? k[i] : b->dec[b->dec_off[t] + uint32_t(-2 - k[i])];
  }
  if (num) *num = k.size();
  return QSIM_HIP_OK;
}

int qsim_traj_state(void* batch, unsigned t, float* out) {
  auto* b = static_cast<TrajBatchHost*>(batch);
  if (b == nullptr || b->handle == nullptr || t >= b->plans.size()) {
    return QSIM_HIP_ERR_ARG;
  }
  for (unsigned d = 0; d < b->order.size(); ++d) {
    if (b->order[d] == t) return qsim_hip_traj_state(b->handle, d, out);
  }
  return QSIM_HIP_ERR_ARG;
}

void qsim_traj_free(void* batch) {
  auto* b = static_cast<TrajBatchHost*>(batch);
  if (b == nullptr) return;
  if (b->handle) qsim_hip_traj_free(b->handle);
  delete b;
}

}  // extern "C"
