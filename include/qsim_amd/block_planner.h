// block_planner.h -- partitions a qsim circuit into cache-blocked passes.
//
// A plain fused pass (MultiQubitGateFuser, lib/fuser_mqubit.h:185-526) costs
// one stream of the state through HBM per fused gate of <= 6 qubits.  A
// blocked pass (qsim_hip_apply_gate_block) costs one stream per tile qubit
// set of up to 13 qubits: the Lc lowest qubits plus chosen higher ones.  The
// planner walks the circuit's dependency frontier and, pass by pass, grows
// the tile set greedily -- adding the single qubit, or the missing qubits of
// a frontier gate, that unlocks the most gates per added qubit -- until the
// set is full; every gate whose qubits lie in the set and whose
// predecessors are applied joins the pass.  The gates of a pass are then
// fused into <= 4-qubit gates by qsim's own MultiQubitGateFuser, with
// matrices multiplied in the operations' FP type (FP64 for the blocked path).
//
// Operations that are not uncontrolled gates on <= 4 qubits (measurements,
// controlled gates, channels, larger gates) end the current segment and are
// handed back as single operations, in order.
#ifndef QSIM_AMD_BLOCK_PLANNER_H_
#define QSIM_AMD_BLOCK_PLANNER_H_

#include <algorithm>
#include <cstdint>
#include <functional>
#include <set>
#include <variant>
#include <vector>

#include "fuser_mqubit.h"
#include "gate.h"
#include "operation.h"

namespace qsim {
namespace hip {

struct BlockPass {
  std::vector<unsigned> tile;  // tile qubits, ascending
  std::vector<unsigned> ks;    // per fused gate: qubit count (<= 4)
  std::vector<unsigned> qs;    // per fused gate: 4 slots, ascending qubits
  std::vector<double> m;       // fused matrices, concatenated (2 * 4^k each)
  // per fused gate: indices (into the planned ops) of its circuit gates, in
  // the order they act
  std::vector<std::vector<size_t>> comps;
  unsigned num_gates = 0;      // circuit gates in the pass
};

// One step of a blocked plan: a blocked pass, or operation `op` of the
// circuit applied on its own.
struct BlockStep {
  bool blocked;
  size_t index;  // into BlockPlan::blocks, or into the circuit's ops
};

struct BlockPlan {
  std::vector<BlockPass> blocks;
  std::vector<BlockStep> steps;
  unsigned trial = 0;  // the restart this plan comes from (0: greedy)
};

struct BlockPlannerParameter {
  unsigned tile_bits = 13;  // tile qubits per pass (9..14)
  unsigned low_bits = 5;    // lowest qubits always in the tile (>= 3)
  unsigned max_ops = 32;    // fused gates per pass
  // Greedy restarts: trial 0 is the plain greedy plan; trials 1.. scale each
  // candidate's score by a pseudo-random factor in 1 +- jitter / 2 (seeded
  // by `seed` and the trial), and the plan with the least
  // passes + op_cost * fused gates is kept (on MI355X a blocked pass costs
  // about 7.75 ms of streaming plus 2.25 ms per fused gate at 32 qubits).
  unsigned restarts = 1;
  double jitter = 0.3;
  double op_cost = 0.29;
  uint64_t seed = 0;
  // >= 0: plan exactly this restart (as an earlier search chose it, e.g.
  // for a second copy of the same circuit) instead of searching.
  int trial = -1;
};

namespace detail {

// Pseudo-random score factors for a restart (splitmix64).
struct Jitter {
  uint64_t state;
  double amount;
  double Next() {
    uint64_t z = (state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return 1.0 + amount * (double(z >> 11) * 0x1.0p-53 - 0.5);
  }
};

template <typename FP>
const Gate<FP>* BlockableGate(const Operation<FP>& op) {
  const Gate<FP>* g = OpGetAlternative<Gate<FP>>(op);
  if (g == nullptr || g->qubits.empty() || g->qubits.size() > 4) return nullptr;
  return g;
}

// Gate dependency frontier over a segment of blockable gates.
class Frontier {
 public:
  Frontier(unsigned n, const std::vector<std::vector<unsigned>>& gq)
      : gq_(gq), per_q_(n), ptr_(n, 0) {
    for (size_t i = 0; i < gq.size(); ++i)
      for (unsigned q : gq[i]) per_q_[q].push_back(i);
  }

  bool Done() const {
    for (size_t q = 0; q < per_q_.size(); ++q)
      if (ptr_[q] < per_q_[q].size()) return false;
    return true;
  }

  // Gates that become applicable with qubit set `in` (a bit mask over
  // qubits), in closure order; advances `ptr` accordingly.
  size_t Closure(const std::vector<char>& in, std::vector<size_t>& ptr,
                 std::vector<size_t>* taken) const {
    size_t cnt = 0;
    bool changed = true;
    while (changed) {
      changed = false;
      for (size_t q = 0; q < per_q_.size(); ++q) {
        if (!in[q]) continue;
        while (ptr[q] < per_q_[q].size()) {
          const size_t i = per_q_[q][ptr[q]];
          bool ok = true;
          for (unsigned x : gq_[i]) {
            ok = ok && in[x] && ptr[x] < per_q_[x].size() &&
                 per_q_[x][ptr[x]] == i;
          }
          if (!ok) break;
          for (unsigned x : gq_[i]) ++ptr[x];
          if (taken) taken->push_back(i);
          ++cnt;
          changed = true;
        }
      }
    }
    return cnt;
  }

  // The closure continued from `ptr` (already closed over `in` minus
  // `seeds`) after the qubits `seeds` joined `in`: only frontier gates of
  // qubits whose frontier moved (or that joined) can become applicable, so a
  // worklist of those qubits reaches the same fixpoint as Closure.
  size_t ClosureFrom(const std::vector<char>& in, std::vector<size_t>& ptr,
                     const std::vector<unsigned>& seeds) const {
    std::vector<unsigned> work(seeds);
    size_t cnt = 0;
    while (!work.empty()) {
      const unsigned q = work.back();
      work.pop_back();
      if (!in[q]) continue;
      while (ptr[q] < per_q_[q].size()) {
        const size_t i = per_q_[q][ptr[q]];
        bool ok = true;
        for (unsigned x : gq_[i]) {
          ok = ok && in[x] && ptr[x] < per_q_[x].size() &&
               per_q_[x][ptr[x]] == i;
        }
        if (!ok) break;
        for (unsigned x : gq_[i]) {
          ++ptr[x];
          if (x != q) work.push_back(x);
        }
        ++cnt;
      }
    }
    return cnt;
  }

  // Chooses a tile set (low qubits + greedy additions) and takes the gates it
  // unlocks.  Returns them in circuit order (ascending index).
  std::vector<size_t> NextPass(unsigned tile_bits, unsigned low_bits,
                               std::vector<unsigned>& tile,
                               Jitter* jitter = nullptr) {
    const unsigned n = per_q_.size();
    std::vector<char> in(n, 0);
    unsigned size = 0;
    for (unsigned q = 0; q < low_bits && q < n; ++q) {
      in[q] = 1;
      ++size;
    }
    // p0: the frontier after every gate the current set closes over.
    std::vector<size_t> p0 = ptr_;
    Closure(in, p0, nullptr);
    while (size < tile_bits) {
      // Candidates: single qubits, and the missing qubits of frontier gates.
      std::vector<std::vector<unsigned>> cands;
      for (unsigned q = 0; q < n; ++q)
        if (!in[q]) cands.push_back({q});
      std::set<std::vector<unsigned>> seen;
      for (unsigned q = 0; q < n; ++q) {
        if (p0[q] >= per_q_[q].size()) continue;
        std::vector<unsigned> add;
        for (unsigned x : gq_[per_q_[q][p0[q]]])
          if (!in[x]) add.push_back(x);
        if (add.size() > 1 && seen.insert(add).second) cands.push_back(add);
      }
      double best = -1;
      const std::vector<unsigned>* pick = nullptr;
      for (const auto& c : cands) {
        if (size + c.size() > tile_bits) continue;
        for (unsigned x : c) in[x] = 1;
        // The closure is a unique fixpoint and grows with the qubit set, so
        // the candidate's extra gates are the closure continued from p0.
        std::vector<size_t> p1 = p0;
        double gain = double(ClosureFrom(in, p1, c)) / double(c.size());
        if (jitter != nullptr) gain *= jitter->Next();
        for (unsigned x : c) in[x] = 0;
        if (gain > best) {
          best = gain;
          pick = &c;
        }
      }
      if (pick == nullptr) break;
      for (unsigned x : *pick) in[x] = 1;
      size += pick->size();
      ClosureFrom(in, p0, *pick);
    }
    std::vector<size_t> taken;
    Closure(in, ptr_, &taken);
    std::sort(taken.begin(), taken.end());
    tile.clear();
    for (unsigned q = 0; q < n; ++q)
      if (in[q]) tile.push_back(q);
    return taken;
  }

 private:
  const std::vector<std::vector<unsigned>>& gq_;
  std::vector<std::vector<size_t>> per_q_;
  std::vector<size_t> ptr_;
};

// A fused gate of a blocked pass: ascending qubits, FP64 row-major
// interleaved matrix (bit j of an index <-> qubits[j]).
struct TileGate {
  std::vector<unsigned> qubits;
  std::vector<double> matrix;
  std::vector<size_t> comps;  // circuit gates in it (indices), in order
};

// m over `qs` (ascending) embedded in the space of `u` (ascending superset),
// identity on the other qubits of u.
inline std::vector<double> Embed(const std::vector<double>& m,
                                 const std::vector<unsigned>& qs,
                                 const std::vector<unsigned>& u) {
  const unsigned D = 1u << u.size(), d = 1u << qs.size();
  unsigned pos[4] = {0, 0, 0, 0}, mask = 0;
  for (size_t i = 0; i < qs.size(); ++i) {
    for (size_t j = 0; j < u.size(); ++j)
      if (u[j] == qs[i]) pos[i] = j;
    mask |= 1u << pos[i];
  }
  std::vector<double> out(2 * size_t(D) * D, 0.0);
  for (unsigned r = 0; r < D; ++r) {
    for (unsigned c = 0; c < D; ++c) {
      if ((r & ~mask) != (c & ~mask)) continue;
      unsigned mr = 0, mc = 0;
      for (size_t i = 0; i < qs.size(); ++i) {
        mr |= ((r >> pos[i]) & 1u) << i;
        mc |= ((c >> pos[i]) & 1u) << i;
      }
      out[2 * (size_t(r) * D + c)] = m[2 * (size_t(mr) * d + mc)];
      out[2 * (size_t(r) * D + c) + 1] = m[2 * (size_t(mr) * d + mc) + 1];
    }
  }
  return out;
}

// Every fused gate of a blocked pass costs one MFMA sweep of the tile
// whatever its size (<= 4 qubits), so gates are merged while the union of
// their qubits stays within 4: gate b joins an earlier gate a when no gate
// between them touches b's qubits (b commutes back to a), as the product
// U_b U_a on the union.  dry: qubit sets only (no matrices), for counting.
inline void MergeTileGates(std::vector<TileGate>& g, bool dry = false) {
  auto overlaps = [](const std::vector<unsigned>& x,
                     const std::vector<unsigned>& y) {
    for (unsigned a : x)
      for (unsigned b : y)
        if (a == b) return true;
    return false;
  };
  bool changed = true;
  while (changed) {
    changed = false;
    for (size_t a = 0; a < g.size() && !changed; ++a) {
      for (size_t b = a + 1; b < g.size(); ++b) {
        std::vector<unsigned> u = g[a].qubits;
        for (unsigned q : g[b].qubits)
          if (std::find(u.begin(), u.end(), q) == u.end()) u.push_back(q);
        if (u.size() > 4) continue;
        bool blocked = false;
        for (size_t k = a + 1; k < b && !blocked; ++k)
          blocked = overlaps(g[k].qubits, g[b].qubits);
        if (blocked) continue;
        std::sort(u.begin(), u.end());
        g[a].comps.insert(g[a].comps.end(), g[b].comps.begin(),
                          g[b].comps.end());
        if (dry) {
          g[a].qubits = u;
          g.erase(g.begin() + b);
          changed = true;
          break;
        }
        const auto ma = Embed(g[a].matrix, g[a].qubits, u);
        const auto mb = Embed(g[b].matrix, g[b].qubits, u);
        const unsigned D = 1u << u.size();
        std::vector<double> prod(2 * size_t(D) * D, 0.0);
        for (unsigned r = 0; r < D; ++r)
          for (unsigned k = 0; k < D; ++k) {
            const double br = mb[2 * (size_t(r) * D + k)];
            const double bi = mb[2 * (size_t(r) * D + k) + 1];
            if (br == 0 && bi == 0) continue;
            for (unsigned c = 0; c < D; ++c) {
              const double ar = ma[2 * (size_t(k) * D + c)];
              const double ai = ma[2 * (size_t(k) * D + c) + 1];
              prod[2 * (size_t(r) * D + c)] += br * ar - bi * ai;
              prod[2 * (size_t(r) * D + c) + 1] += br * ai + bi * ar;
            }
          }
        g[a].qubits = u;
        g[a].matrix = std::move(prod);
        g.erase(g.begin() + b);
        changed = true;
        break;
      }
    }
  }
}

}  // namespace detail

namespace detail {

// One greedy plan (with `jitter`, a restart's perturbed one).  dry: the
// plan's structure only (tiles and fused-gate qubits, no matrices), for
// comparing restarts.
template <typename IO, typename FP>
bool PlanBlocksOnce(const BlockPlannerParameter& param, unsigned num_qubits,
                    const std::vector<Operation<FP>>& ops, BlockPlan& plan,
                    const std::function<void(FusedGate<FP>&)>* fill,
                    Jitter* jitter, bool dry) {
  plan = BlockPlan{};
  const unsigned tb = std::min(param.tile_bits, num_qubits);
  if (tb < 9 || tb > 14) return false;
  const unsigned lc = std::min(std::max(param.low_bits, 3u), tb);

  size_t i = 0;
  while (i < ops.size()) {
    // A segment of blockable gates [i, j).
    size_t j = i;
    while (j < ops.size() && BlockableGate(ops[j]) != nullptr) ++j;
    if (j > i) {
      std::vector<std::vector<unsigned>> gq;
      for (size_t t = i; t < j; ++t)
        gq.push_back(BlockableGate(ops[t])->qubits);
      Frontier fr(num_qubits, gq);
      while (!fr.Done()) {
        std::vector<unsigned> tile;
        std::vector<size_t> taken = fr.NextPass(tb, lc, tile, jitter);
        if (taken.empty()) return false;  // cannot happen for k <= 4 gates
        std::vector<const Operation<FP>*> ptrs;
        for (size_t t : taken) ptrs.push_back(&ops[i + t]);
        typename MultiQubitGateFuser<IO>::Parameter fp;
        fp.max_fused_size = 4;
        fp.verbosity = 0;
        const std::vector<unsigned> no_splits;
        auto fused =
            MultiQubitGateFuser<IO>::template FuseGates<const Operation<FP>*>(
                fp, num_qubits, ptrs.cbegin(), ptrs.cend(), no_splits,
                !dry && fill == nullptr);
        // The circuit index of each component gate.
        std::vector<std::pair<const Gate<FP>*, size_t>> where;
        for (size_t t : taken) where.push_back({BlockableGate(ops[i + t]), i + t});
        auto index_of = [&](const typename FusedGate<FP>::PGate& pg) {
          const Gate<FP>* const* gp = std::get_if<const Gate<FP>*>(&pg);
          for (const auto& w : where)
            if (gp != nullptr && w.first == *gp) return w.second;
          return size_t(-1);
        };
        std::vector<TileGate> tg;
        for (auto& f : fused) {
          auto* fg = std::get_if<FusedGate<FP>>(&f);
          if (fg == nullptr) return false;  // only gates were passed in
          if (!dry && fill != nullptr) (*fill)(*fg);
          TileGate t{fg->qubits,
                     std::vector<double>(fg->matrix.begin(), fg->matrix.end()),
                     {}};
          if (!dry) {
            for (const auto& pg : fg->gates) t.comps.push_back(index_of(pg));
          }
          tg.push_back(std::move(t));
        }
        MergeTileGates(tg, dry);
        BlockPass bp;
        bp.tile = tile;
        bp.num_gates = taken.size();
        auto flush = [&]() {
          plan.steps.push_back({true, plan.blocks.size()});
          plan.blocks.push_back(bp);
          bp.ks.clear();
          bp.qs.clear();
          bp.m.clear();
          bp.comps.clear();
          bp.num_gates = 0;
        };
        for (auto& t : tg) {
          if (bp.ks.size() == param.max_ops) flush();
          bp.ks.push_back(t.qubits.size());
          for (unsigned b = 0; b < 4; ++b)
            bp.qs.push_back(b < t.qubits.size() ? t.qubits[b] : 0);
          bp.m.insert(bp.m.end(), t.matrix.begin(), t.matrix.end());
          bp.comps.push_back(std::move(t.comps));
        }
        if (!bp.ks.empty()) flush();
      }
    }
    if (j < ops.size()) plan.steps.push_back({false, j});
    i = j + 1;
  }
  return true;
}

}  // namespace detail

// Plans blocked passes for ops of a circuit on `num_qubits` qubits.  Returns
// false if the circuit is too small for blocked passes (num_qubits < 9).
// `fill`, if given, supplies each fused gate's matrix (qsim's fuser then
// skips its own products): a caller planning many similar gate lists can
// reuse matrices it has already formed.  With param.restarts > 1 the
// greedy plan is compared with perturbed ones by their structure (passes,
// fused gates) and the best is planned in full.
template <typename IO, typename FP>
bool PlanBlocks(const BlockPlannerParameter& param, unsigned num_qubits,
                const std::vector<Operation<FP>>& ops, BlockPlan& plan,
                const std::function<void(FusedGate<FP>&)>* fill = nullptr) {
  auto jitter_of = [&](unsigned trial) {
    return detail::Jitter{param.seed * 0x100000001b3ull + trial, param.jitter};
  };
  if (param.trial > 0) {
    detail::Jitter jt = jitter_of(unsigned(param.trial));
    const bool ok = detail::PlanBlocksOnce<IO>(param, num_qubits, ops, plan,
                                               fill, &jt, false);
    plan.trial = unsigned(param.trial);
    return ok;
  }
  if (param.restarts <= 1 || param.trial == 0) {
    return detail::PlanBlocksOnce<IO>(param, num_qubits, ops, plan, fill,
                                      nullptr, false);
  }
  double best = 0;
  unsigned best_trial = 0;
  for (unsigned trial = 0; trial < param.restarts; ++trial) {
    detail::Jitter jt = jitter_of(trial);
    BlockPlan p;
    if (!detail::PlanBlocksOnce<IO>(param, num_qubits, ops, p, fill,
                                    trial == 0 ? nullptr : &jt, true)) {
      return false;
    }
    size_t nops = 0;
    for (const auto& b : p.blocks) nops += b.ks.size();
    const double cost = double(p.blocks.size()) + param.op_cost * double(nops);
    if (trial == 0 || cost < best) {
      best = cost;
      best_trial = trial;
    }
  }
  detail::Jitter jt = jitter_of(best_trial);
  const bool ok = detail::PlanBlocksOnce<IO>(
      param, num_qubits, ops, plan, fill, best_trial == 0 ? nullptr : &jt,
      false);
  plan.trial = best_trial;
  return ok;
}

}  // namespace hip
}  // namespace qsim

#endif  // QSIM_AMD_BLOCK_PLANNER_H_
