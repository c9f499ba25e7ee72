// traj_kernels.hip -- batched noisy trajectories in one launch.
//
// qsim's trajectory simulator (lib/qtrajectory.h:112-387) runs trajectories
// one after another, each gate batch as its own set of kernel launches with a
// host synchronisation at every channel whose Kraus operator depends on the
// state.  Here a whole batch -- hundreds of trajectories of uneven qubit
// counts -- is ONE launch: workgroup t owns trajectory t, its state in a
// device pool and its operation stream, and walks the stream with nothing
// but workgroup barriers between operations:
//
//   gate      a fused gate (k <= 4, FP32 matrix in LDS) over all 2^(n-k)
//             groups of the trajectory's state;
//   tile pass a run of gates applied tile by tile through LDS with the
//             FP16 MFMA product of the circuit kernel's blocked passes (the
//             tiled kernel; the per-gate kernels run the gates one by one);
//   channel   a channel whose draw fell past the probability bounds: FP64
//             sweeps for <psi|K^dagger K|psi> operator by operator (the
//             first also giving the norm), the choice exactly as
//             RunIteration makes it (lib/qtrajectory.h:330-370), then the
//             chosen operator applied (scaled to the normalised state where
//             the trajectory is not unitary).
//
// The Kraus operators that the bounds decide (depolarising Paulis, damping
// K0 when the draw is below its bound, any unitary operator whose
// cumulative bound covers the draw) are sampled on the host from the same
// mt19937(rep) draws and folded into the fused gates.  At the end each
// trajectory writes <Z_a Z_b> of its qubit pairs, FP64-accumulated and
// divided by the norm (the normalisation of lib/qtrajectory.h:377-383).
//
// Workgroups are dispatched in trajectory order, so the host puts the
// largest states first (longest-processing-time-first balancing across CUs).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <system_error>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "block_tile.h"
#include "internal.h"
#include "reduce.h"

namespace qhip {

namespace {

using namespace tileop;

// Pairs per <Z_a Z_b> sweep (FP64 accumulators per thread: 32 would not fit
// beside the tile kernel's registers) and amplitudes per thread per step of
// a whole-state sweep (64 KiB of loads in flight per workgroup).
template <int NT>
constexpr unsigned kPairChunk = 16;
template <int NT>
constexpr unsigned kSweep = 8192 / NT;

template <int NT, int K>
__device__ void ApplyGateK(float2* __restrict__ s, unsigned n,
                           const qsim_hip_traj_op& o, const float* lm) {
  constexpr unsigned D = 1u << K;
  const uint64_t ngroups = uint64_t{1} << (n - K);
  unsigned q[K];
#pragma unroll
  for (int b = 0; b < K; ++b) q[b] = o.q[b];
  for (uint64_t g = threadIdx.x; g < ngroups; g += NT) {
    uint64_t base = g;
#pragma unroll
    for (int b = 0; b < K; ++b) {
      base = ((base >> q[b]) << (q[b] + 1)) | (base & ((uint64_t{1} << q[b]) - 1));
    }
    float2 v[D];
#pragma unroll
    for (unsigned j = 0; j < D; ++j) {
      uint64_t off = 0;
#pragma unroll
      for (int b = 0; b < K; ++b) off |= uint64_t((j >> b) & 1u) << q[b];
      v[j] = s[base + off];
    }
    // Rows stay a loop for k >= 3: unrolled, the compiler would hoist the
    // whole LDS matrix (2 D^2 floats) into registers across the group loop.
#pragma unroll(K <= 2 ? D : 1)
    for (unsigned r = 0; r < D; ++r) {
      float re = 0.0f, im = 0.0f;
#pragma unroll
      for (unsigned j = 0; j < D; ++j) {
        const float mr = lm[2 * (r * D + j)], mi = lm[2 * (r * D + j) + 1];
        re = fmaf(mr, v[j].x, re);
        re = fmaf(-mi, v[j].y, re);
        im = fmaf(mr, v[j].y, im);
        im = fmaf(mi, v[j].x, im);
      }
      uint64_t off = 0;
#pragma unroll
      for (int b = 0; b < K; ++b) off |= uint64_t((r >> b) & 1u) << q[b];
      s[base + off] = float2{re, im};
    }
  }
}

// LDS scratch of the reductions: NT / 64 wave partials, then two broadcast
// slots.
template <int NT>
constexpr unsigned kRedWords = NT / 64 + 2;

template <int NT>
__device__ void ZeroState(float2* __restrict__ s, uint64_t N) {
  for (uint64_t i = threadIdx.x; i < N; i += NT) {
    s[i] = float2{i == 0 ? 1.0f : 0.0f, 0.0f};
  }
}

// A fused gate: its FP32 matrix staged in LDS (lm), then one sweep.
template <int NT>
__device__ void TrajGate(float2* __restrict__ s, unsigned n,
                         const qsim_hip_traj_op& o,
                         const float* __restrict__ mats, float* lm) {
  const unsigned D = 1u << o.k;
  for (unsigned i = threadIdx.x; i < 2 * D * D; i += NT) lm[i] = mats[o.moff + i];
  __syncthreads();
  switch (o.k) {
    case 1: ApplyGateK<NT, 1>(s, n, o, lm); break;
    case 2: ApplyGateK<NT, 2>(s, n, o, lm); break;
    case 3: ApplyGateK<NT, 3>(s, n, o, lm); break;
    default: ApplyGateK<NT, 4>(s, n, o, lm); break;
  }
  __syncthreads();
}

// Most Kraus operators of a channel decided on the device, and its most
// qubits.
constexpr unsigned kMaxKraus = 16;
constexpr unsigned kMaxChannelQubits = 2;

// Scales the state by a (a sweep).
template <int NT>
__device__ void ScaleState(float2* __restrict__ s, uint64_t N, float a) {
  for (uint64_t i = threadIdx.x; i < N; i += NT) {
    const float2 v = s[i];
    s[i] = float2{v.x * a, v.y * a};
  }
}

// A channel that the draw's bounds did not decide (header ops[oi], its
// Kraus operators at ops[oi + 1 ..]), decided as RunIteration decides it
// (lib/qtrajectory.h:330-370) and in the same order: the non-unitary
// operators one at a time, each with a sweep for <psi|K^dagger K|psi>
// (each group's value in FP32, summed in FP64) -- the first sweep also
// giving the norm where the trajectory is not unitary, so that the
// probabilities are those of the normalised state (NormalizeState,
// lib/qtrajectory.h:449-456) -- until the draw falls below the running
// sum; the chosen operator is then applied, scaled by the normalisation.
// lm: LDS of 2 x 4^NQ floats (at least one 4 x 4 complex matrix); vals:
// LDS of 2 doubles.
template <int NT, int NQ>
__device__ void TrajChannelQ(float2* __restrict__ s, unsigned n,
                             const qsim_hip_traj_op* __restrict__ ops,
                             uint64_t oi, const float* __restrict__ mats,
                             int32_t* __restrict__ dec_out, double* red,
                             float* lm, double* vals) {
  constexpr unsigned D = 1u << NQ;
  const qsim_hip_traj_op h = ops[oi];
  const unsigned M = h.k;
  const uint64_t N = uint64_t{1} << n;
  const uint64_t ngroups = N >> NQ;
  const unsigned q0 = h.q[0], q1 = NQ > 1 ? h.q[1] : 0;
  const bool normalise = h.pad & 1u;
  bool have_norm = !normalise;
  double norm = 1.0;
  double cp = h.cp, max_prob = 0.0;
  int max_index = 0, pick = -1;
  for (unsigned i = 0; i < M; ++i) {
    const qsim_hip_traj_op c = ops[oi + 1 + i];
    if (c.pad & 1u) continue;  // unitary: decided by its bound alone
    __syncthreads();           // earlier reads of lm are done
    for (unsigned e = threadIdx.x; e < 2 * D * D; e += NT)
      lm[e] = mats[c.moff2 + e];
    __syncthreads();
    double acc = 0.0, nacc = 0.0;
    for (uint64_t g = threadIdx.x; g < ngroups; g += NT) {
      uint64_t base =
          ((g >> q0) << (q0 + 1)) | (g & ((uint64_t{1} << q0) - 1));
      if (NQ > 1) {
        base = ((base >> q1) << (q1 + 1)) | (base & ((uint64_t{1} << q1) - 1));
      }
      float2 v[D];
#pragma unroll
      for (unsigned j = 0; j < D; ++j) {
        const uint64_t off = ((j & 1u) ? uint64_t{1} << q0 : 0) |
                             ((j & 2u) ? uint64_t{1} << q1 : 0);
        v[j] = s[base + off];
      }
      float w = 0.0f;  // Re v^dagger (K^dagger K) v
#pragma unroll
      for (unsigned r = 0; r < D; ++r) {
        float re = 0.0f, im = 0.0f;
#pragma unroll
        for (unsigned k = 0; k < D; ++k) {
          const float mr = lm[2 * (r * D + k)], mi = lm[2 * (r * D + k) + 1];
          re += mr * v[k].x - mi * v[k].y;
          im += mr * v[k].y + mi * v[k].x;
        }
        w += v[r].x * re + v[r].y * im;
      }
      acc += double(w);
      if (!have_norm) {
#pragma unroll
        for (unsigned j = 0; j < D; ++j)
          nacc += double(v[j].x) * v[j].x + double(v[j].y) * v[j].y;
      }
    }
    acc = BlockSum<NT>(acc, red);
    if (threadIdx.x == 0) vals[0] = acc;
    if (!have_norm) {
      nacc = BlockSum<NT>(nacc, red);
      if (threadIdx.x == 0) vals[1] = nacc;
    }
    __syncthreads();
    if (!have_norm) {
      norm = vals[1];
      have_norm = true;
    }
    // The walk, the same in every thread.
    const double prob = vals[0] / norm;
    if (prob > max_prob) {
      max_prob = prob;
      max_index = int(i);
    }
    cp += prob - c.r;
    if (h.r < cp || i == M - 1) {
      pick = h.r < cp ? int(i) : max_index;
      break;
    }
  }
  if (!have_norm) {
    // Every operator unitary: nothing chosen, the state only normalised.
    double nacc = 0.0;
    for (uint64_t i = threadIdx.x; i < N; i += NT) {
      const float2 v = s[i];
      nacc += double(v.x) * v.x + double(v.y) * v.y;
    }
    nacc = BlockSum<NT>(nacc, red);
    if (threadIdx.x == 0) vals[1] = nacc;
    __syncthreads();
    norm = vals[1];
  }
  if (threadIdx.x == 0) dec_out[h.slot] = pick;
  __syncthreads();  // every thread is done with lm and vals
  const float a = normalise ? float(1.0 / std::sqrt(norm)) : 1.0f;
  const qsim_hip_traj_op c = ops[oi + 1 + (pick >= 0 ? pick : 0)];
  if (pick < 0 || c.k == 0) {
    // Nothing applied (or the identity): the state is only normalised.
    if (normalise) ScaleState<NT>(s, N, a);
    __syncthreads();
    return;
  }
  const unsigned DK = 1u << c.k;
  for (unsigned e = threadIdx.x; e < 2 * DK * DK; e += NT)
    lm[e] = mats[c.moff + e] * a;
  __syncthreads();
  if (c.k == 1) ApplyGateK<NT, 1>(s, n, c, lm);
  else ApplyGateK<NT, 2>(s, n, c, lm);
  __syncthreads();
}

template <int NT>
__device__ void TrajChannel(float2* __restrict__ s, unsigned n,
                            const qsim_hip_traj_op* __restrict__ ops,
                            uint64_t oi, const float* __restrict__ mats,
                            int32_t* __restrict__ dec_out, double* red,
                            float* lm, double* vals) {
  static_assert(kMaxChannelQubits == 2, "channel sizes");
  if (ops[oi].moff2 == 1)
    TrajChannelQ<NT, 1>(s, n, ops, oi, mats, dec_out, red, lm, vals);
  else
    TrajChannelQ<NT, 2>(s, n, ops, oi, mats, dec_out, red, lm, vals);
}

// Signs (-1)^(bit a + bit b) of index i for the pairs (pa[j], pb[j]),
// j < kPC, as bit j of a mask.  The map is linear over XOR, so the sign mask
// of i = x | y (disjoint bits) is the XOR of those of x and y.
template <unsigned kPC>
__device__ __forceinline__ uint32_t PairSigns(uint64_t i,
                                              const unsigned (&pa)[kPC],
                                              const unsigned (&pb)[kPC]) {
  uint32_t m = 0;
#pragma unroll
  for (unsigned j = 0; j < kPC; ++j) {
    m |= uint32_t(((i >> pa[j]) ^ (i >> pb[j])) & 1u) << j;
  }
  return m;
}

// <Z_a Z_b> = sum_i |s_i|^2 (-1)^(bit a + bit b) / sum_i |s_i|^2 for the
// trajectory's pairs, FP64-accumulated.  A non-unitary trajectory ends
// normalised (lib/qtrajectory.h:377-383): its norm goes to norm_out and the
// scaling is applied where the state is read (qsim_hip_traj_state), which
// saves a read-write sweep of every state.
template <int NT>
__device__ void TrajFinish(const float2* __restrict__ s,
                           const qsim_hip_traj_desc& d,
                           const uint32_t* __restrict__ pairs,
                           double* __restrict__ zz_out,
                           double* __restrict__ norm_out, double* red) {
  constexpr unsigned kS = kSweep<NT>;
  constexpr unsigned kPC = kPairChunk<NT>;
  const uint64_t N = uint64_t{1} << d.num_qubits;
  double total = 0.0;  // the norm, in thread 0
  bool have_norm = false;
  for (unsigned p0 = 0; p0 < d.num_pairs; p0 += kPC) {
    const unsigned np = std::min(kPC, d.num_pairs - p0);
    unsigned pa[kPC], pb[kPC];
    double acc[kPC];
#pragma unroll
    for (unsigned j = 0; j < kPC; ++j) {
      pa[j] = j < np ? pairs[2 * (d.pair_offset + p0 + j)] : 0;
      pb[j] = j < np ? pairs[2 * (d.pair_offset + p0 + j) + 1] : 0;
      acc[j] = 0.0;
    }
    // Element u of a step is i0 + u NT = i0 | u NT (i0's bits from
    // log2 NT to log2 (kS NT) are zero): per-element sign masks are
    // uniform, per-step ones one per thread.
    uint32_t su[kS];
#pragma unroll
    for (unsigned u = 0; u < kS; ++u) su[u] = PairSigns(uint64_t(u) * NT, pa, pb);
    double nrm = 0.0;
    for (uint64_t i0 = threadIdx.x; i0 < N; i0 += kS * NT) {
      const uint32_t s0 = PairSigns(i0, pa, pb);
      float2 a[kS];
#pragma unroll
      for (unsigned u = 0; u < kS; ++u) {
        const uint64_t i = i0 + u * NT;
        a[u] = i < N ? s[i] : float2{0.0f, 0.0f};
      }
#pragma unroll
      for (unsigned u = 0; u < kS; ++u) {
        const double w = double(a[u].x) * a[u].x + double(a[u].y) * a[u].y;
        const uint32_t m = s0 ^ su[u];
        nrm += w;
#pragma unroll
        for (unsigned j = 0; j < kPC; ++j) acc[j] += (m >> j) & 1u ? -w : w;
      }
    }
    nrm = BlockSum<NT>(nrm, red);
    total = nrm;
    have_norm = true;
    // Unrolled, so that acc stays in registers.
#pragma unroll
    for (unsigned j = 0; j < kPC; ++j) {
      const double v = BlockSum<NT>(acc[j], red);
      if (threadIdx.x == 0 && j < np) zz_out[d.pair_offset + p0 + j] = v / nrm;
    }
  }
  if (!(d.pad & 1u)) return;
  if (!have_norm) {
    double nrm = 0.0;
    for (uint64_t i0 = threadIdx.x; i0 < N; i0 += kS * NT) {
      float2 a[kS];
#pragma unroll
      for (unsigned u = 0; u < kS; ++u) {
        const uint64_t i = i0 + u * NT;
        a[u] = i < N ? s[i] : float2{0.0f, 0.0f};
      }
#pragma unroll
      for (unsigned u = 0; u < kS; ++u) {
        nrm += double(a[u].x) * a[u].x + double(a[u].y) * a[u].y;
      }
    }
    total = BlockSum<NT>(nrm, red);
  }
  if (threadIdx.x == 0) norm_out[blockIdx.x] = total;
}

// Workgroup t runs trajectory t, one sweep of the state per operation.  A
// tile-pass header is skipped: the gates of its pass follow it in the
// stream and run here one sweep each.
template <int NT>
__global__ void __launch_bounds__(NT)
TrajectoryKernel(const qsim_hip_traj_desc* __restrict__ descs,
                 const qsim_hip_traj_op* __restrict__ ops,
                 const float* __restrict__ mats,
                 const uint32_t* __restrict__ pairs, float2* __restrict__ pool,
                 double* __restrict__ zz_out, double* __restrict__ norm_out,
                 int32_t* __restrict__ dec_out) {
  __shared__ float lm[2 * 16 * 16];
  __shared__ double red[kRedWords<NT>];
  __shared__ double vals[2];
  const qsim_hip_traj_desc d = descs[blockIdx.x];
  const unsigned n = d.num_qubits;
  float2* s = pool + d.state_offset;
  ZeroState<NT>(s, uint64_t{1} << n);
  __syncthreads();
  for (uint64_t oi = d.op_begin; oi < d.op_end; ++oi) {
    const qsim_hip_traj_op o = ops[oi];
    if (o.kind == QSIM_HIP_TRAJ_GATE) {
      TrajGate<NT>(s, n, o, mats, lm);
    } else if (o.kind == QSIM_HIP_TRAJ_CHANNEL) {
      TrajChannel<NT>(s, n, ops, oi, mats, dec_out, red, lm, vals);
      oi += o.k;  // its Kraus operators
    }
  }
  TrajFinish<NT>(s, d, pairs, zz_out, norm_out, red);
}

// A tile pass of a trajectory on the device: geometry, program, and the
// index of its first op record (kF16RecWords words each) in the record pool.
struct TrajTilePass {
  BlockGeom g;
  BlockProgram p;
  uint64_t rec;
};

// The tiled kernel's shapes (block_tile.h): TB = 13, tiles of up to 13
// qubits, 256 threads per trajectory, two workgroups per CU; TB = 14, tiles
// of up to 14 qubits, 512 threads, one workgroup per CU (twice the bytes in
// flight per trajectory, fewer passes, half as many trajectories at once).
// One tile pass of the workgroup's trajectory, with the FP16 arithmetic of
// the circuit kernel's blocked passes (block_kernels.hip): the pass's op
// records are copied into LDS once, then for every tile of the state the
// tile is staged in LDS scaled by a power of two (TileScale), the ops run
// on it (FP16 MFMA, two parts per operand), and it is stored back scaled
// the other way, with the next tile's loads in flight meanwhile (in two
// halves around the first ops).  A pass's ops are unitary or Kraus
// operators (spectral norm <= 1), so the tile never grows.  The tile sits at
// LDS offset 0 (no static LDS in the kernel, checked on the host).  FULL:
// tiles of TB qubits.
template <int TB, bool FULL>
__device__ __forceinline__ void TrajTileRun(float2* __restrict__ s,
                                            const TrajTilePass* __restrict__ tp,
                                            const uint32_t* __restrict__ recs,
                                            char* lds_raw) {
  constexpr int NT = BlockShape<TB>::NT;
  constexpr int NW = BlockShape<TB>::kWaves;
  constexpr int PER = 32;
  static_assert((1 << TB) / NT == PER, "shape");
  float2* lds = reinterpret_cast<float2*>(lds_raw);
  float* slots = reinterpret_cast<float*>(lds_raw + F16Lds<TB>::kMaxOff);
  uint32_t* rec = reinterpret_cast<uint32_t*>(lds_raw + F16Lds<TB>::kRecOff);
  const BlockGeom& g = tp->g;
  const BlockProgram& p = tp->p;
  const unsigned nops = p.nops;
  const int l = threadIdx.x & 63;
  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned nbatch = (1u << g.tb) >> 9;
  const unsigned nper = (1u << g.tb) / NT;
  __syncthreads();  // the record area was scratch until now
  {
    const unsigned nq = nops * unsigned(kF16RecWords / 4);
    const uint4* src =
        reinterpret_cast<const uint4*>(recs + tp->rec * kF16RecWords);
    uint4* dst = reinterpret_cast<uint4*>(rec);
    for (unsigned q = threadIdx.x; q < nq; q += NT) dst[q] = src[q];
  }
  TileIo<PER, NT> io;
  io.Init(g);
  float2 buf[PER];
  io.template Fetch<FULL, 15u>(s, 0, g, buf);  // tile 0 starts at amplitude 0
  int e = TileScale<NW>(LocalMax<PER, FULL>(buf, nper), slots, p.headroom);
  io.template PutScaled<FULL>(g, buf, lds, e);
  __syncthreads();
  for (uint64_t tile = 0; tile < g.ntiles; ++tile) {
    const uint64_t base = InsertZeros(tile << g.Lc, g.outer, g.nouter);
    // The next tile, or this one again after the last (an in-bounds read
    // that is never used).
    const uint64_t next = tile + 1 < g.ntiles ? tile + 1 : tile;
    const uint64_t nb = InsertZeros(next << g.Lc, g.outer, g.nouter);
    io.template Fetch<FULL, 3u>(s, nb, g, buf);
    for (unsigned op = 0; op < nops; ++op) {
      const uint32_t* r = rec + op * unsigned(kF16RecWords);
      v8h f[4];
      LoadFragF16(r, l, f);
      const uint32_t* tab = r + kF16TabOff + 16 * (l >> 5);
      const uint32_t sj = r[kF16TabOff + kTableWords + (l & 31)];
      uint32_t mo[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) mo[q] = tab[q] ^ sj;
      ApplyTileOpF16(reinterpret_cast<const uint8_t*>(r + kF16BposOff), f,
                     mo, nbatch, wave, NW);
      __syncthreads();
      if (op == 1) io.template Fetch<FULL, 12u>(s, nb, g, buf);
    }
    if (nops == 1) io.template Fetch<FULL, 12u>(s, nb, g, buf);
    io.template StoreScaled<16, FULL>(s, base, g, lds, e);
    e = TileScale<NW>(LocalMax<PER, FULL>(buf, nper), slots, p.headroom);
    io.template PutScaled<FULL>(g, buf, lds, e);
    __syncthreads();
  }
}

// Workgroup t runs trajectory t with its gate runs as tile passes: one
// stream of the state per pass instead of one per fused gate.  Gates outside
// a pass (states under 9 qubits) and channel decisions run as sweeps over
// the state with the workgroup's threads; their scratch (the gate matrix,
// the reduction partials) reuses the record area of the LDS.
template <int TB>
__global__ void __launch_bounds__(BlockShape<TB>::NT)
__attribute__((amdgpu_waves_per_eu(2)))
TrajTileKernel(const qsim_hip_traj_desc* __restrict__ descs,
               const qsim_hip_traj_op* __restrict__ ops,
               const float* __restrict__ mats,
               const TrajTilePass* __restrict__ passes,
               const uint32_t* __restrict__ recs,
               const uint32_t* __restrict__ pairs, float2* __restrict__ pool,
               double* __restrict__ zz_out, double* __restrict__ norm_out,
               int32_t* __restrict__ dec_out) {
  constexpr int NT = BlockShape<TB>::NT;
  extern __shared__ char lds_raw[];
  // The record area holds the channel and gate scratch between tile passes.
  char* aux = lds_raw + F16Lds<TB>::kRecOff;
  constexpr size_t kLmBytes = 2 * 16 * 16 * sizeof(float);
  static_assert(kLmBytes + (kRedWords<NT> + 2) * sizeof(double) <=
                    F16Lds<TB>::Bytes(F16Lds<TB>::kMaxOps) - F16Lds<TB>::kRecOff,
                "scratch fits the record area");
  float* lm = reinterpret_cast<float*>(aux);
  double* red = reinterpret_cast<double*>(aux + kLmBytes);
  double* vals = red + kRedWords<NT>;
  const qsim_hip_traj_desc d = descs[blockIdx.x];
  const unsigned n = d.num_qubits;
  float2* s = pool + d.state_offset;
  ZeroState<NT>(s, uint64_t{1} << n);
  __syncthreads();
  for (uint64_t oi = d.op_begin; oi < d.op_end; ++oi) {
    const qsim_hip_traj_op o = ops[oi];
    if (o.kind == QSIM_HIP_TRAJ_GATE) {
      TrajGate<NT>(s, n, o, mats, lm);
    } else if (o.kind == QSIM_HIP_TRAJ_TILE_PASS) {
      if (passes[o.moff].g.tb == TB) {
        TrajTileRun<TB, true>(s, passes + o.moff, recs, lds_raw);
      } else {
        TrajTileRun<TB, false>(s, passes + o.moff, recs, lds_raw);
      }
      oi += o.k;  // the pass's gates
    } else if (o.kind == QSIM_HIP_TRAJ_CHANNEL) {
      TrajChannel<NT>(s, n, ops, oi, mats, dec_out, red, lm, vals);
      oi += o.k;  // its Kraus operators
    }
  }
  TrajFinish<NT>(s, d, pairs, zz_out, norm_out, red);
}

// Dynamic LDS of the tiled kernel: the tile, the scale slots and the
// records of the longest pass its shape keeps resident.
template <int TB>
constexpr size_t TrajLds() {
  return F16Lds<TB>::Bytes(F16Lds<TB>::kMaxOps);
}

struct TrajBatch {
  unsigned num_traj = 0;
  uint32_t num_pairs = 0, num_decisions = 0;
  std::vector<unsigned> max_qubits;
  qsim_hip_traj_desc* descs = nullptr;
  qsim_hip_traj_op* ops = nullptr;
  float* mats = nullptr;
  uint32_t* pairs = nullptr;
  float2* pool = nullptr;
  double* zz = nullptr;
  double* norms = nullptr;  // per trajectory (device order)
  std::vector<uint8_t> normalize;
  int32_t* dec = nullptr;
  TrajTilePass* passes = nullptr;
  uint32_t* recs = nullptr;
  unsigned max_tile_bits = 0;  // over the batch's tile passes
  std::vector<uint64_t> state_offset;
  std::vector<unsigned> nq;
  hipStream_t gstream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  unsigned graph_threads = 0;
};

template <int NT>
void LaunchTraj(TrajBatch& b, hipStream_t st) {
  hipLaunchKernelGGL(TrajectoryKernel<NT>, dim3(b.num_traj), dim3(NT), 0, st,
                     b.descs, b.ops, b.mats, b.pairs, b.pool, b.zz, b.norms,
                     b.dec);
}

// The tiled kernel needs more than 64 KiB of dynamic LDS (allowed once per
// process, outside any stream capture) and addresses its tile from LDS
// offset 0, which holds only with no static LDS ahead of the dynamic block.
int TileKernelReady() {
  static std::once_flag once;
  static hipError_t attr = hipSuccess;
  static size_t static_lds = 0;
  std::call_once(once, [] {
    const struct {
      const void* fn;
      size_t lds;
    } fns[2] = {
        {reinterpret_cast<const void*>(&TrajTileKernel<13>), TrajLds<13>()},
        {reinterpret_cast<const void*>(&TrajTileKernel<14>), TrajLds<14>()}};
    for (const auto& f : fns) {
      hipFuncAttributes fa{};
      if (attr == hipSuccess) attr = hipFuncGetAttributes(&fa, f.fn);
      static_lds = std::max(static_lds, fa.sharedSizeBytes);
      if (attr == hipSuccess) {
        attr = hipFuncSetAttribute(
            f.fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(f.lds));
      }
    }
  });
  QHIP_CHECK(attr);
  if (static_lds != 0) {
    SetError("tiled trajectory kernel has %zu bytes of static LDS", static_lds);
    return QSIM_HIP_ERR_RUNTIME;
  }
  return QSIM_HIP_OK;
}

int Launch(TrajBatch& b, unsigned threads, hipStream_t st) {
  switch (threads) {
    case 256: LaunchTraj<256>(b, st); break;
    case 512: LaunchTraj<512>(b, st); break;
    case 1024: LaunchTraj<1024>(b, st); break;
    case QSIM_HIP_TRAJ_TILED:
      QHIP_TRY(TileKernelReady());
      // The 14-qubit shape when a pass needs it.
      if (b.max_tile_bits == 14) {
        hipLaunchKernelGGL(TrajTileKernel<14>, dim3(b.num_traj),
                           dim3(BlockShape<14>::NT), TrajLds<14>(), st,
                           b.descs, b.ops, b.mats, b.passes, b.recs, b.pairs,
                           b.pool, b.zz, b.norms, b.dec);
      } else {
        hipLaunchKernelGGL(TrajTileKernel<13>, dim3(b.num_traj),
                           dim3(BlockShape<13>::NT), TrajLds<13>(), st,
                           b.descs, b.ops, b.mats, b.passes, b.recs, b.pairs,
                           b.pool, b.zz, b.norms, b.dec);
      }
      QHIP_LAUNCHED("TrajTileKernel");
      return QSIM_HIP_OK;
    default:
      SetError("trajectory batches run with 256, 512 or 1024 threads or "
               "tiled (%u), not %u", unsigned(QSIM_HIP_TRAJ_TILED), threads);
      return QSIM_HIP_ERR_ARG;
  }
  QHIP_LAUNCHED("TrajectoryKernel");
  return QSIM_HIP_OK;
}

// A distinct tile pass to build: the header of its first occurrence (its
// gates o[1..k] follow) and the state size.
struct TilePassJob {
  unsigned n;
  const qsim_hip_traj_op* o;
};

// Index of the tile pass whose header is o[0] among the batch's distinct
// passes, adding it if new.  Passes with the same state size, tile, gate
// qubits and matrix offsets -- most passes of the trajectories of one
// circuit, when the caller shares identical matrices as the batch planner
// does -- share one device copy.
uint32_t TilePassIndex(unsigned n, const qsim_hip_traj_op* o,
                       std::vector<TilePassJob>& jobs,
                       std::unordered_map<std::string, uint32_t>& seen) {
  std::string key;
  auto put = [&](const void* p, size_t bytes) {
    key.append(static_cast<const char*>(p), bytes);
  };
  put(&n, sizeof(n));
  put(o[0].q, 2 * sizeof(uint32_t));
  for (unsigned j = 1; j <= o[0].k; ++j) {
    put(&o[j].k, sizeof(uint32_t));
    put(o[j].q, o[j].k * sizeof(uint32_t));
    put(&o[j].moff, sizeof(uint32_t));
  }
  auto it = seen.emplace(std::move(key), uint32_t(jobs.size())).first;
  if (it->second == jobs.size()) jobs.push_back({n, o});
  return it->second;
}

// The device form of a tile pass (header and gates validated, including the
// conditions PrepareGateBlock checks): geometry, program and FP16 op
// records as the circuit kernel's blocked passes build them, from the
// gates' FP32 matrices (exact in FP64); the records go to
// rec[0 .. k kF16RecWords).
int BuildTilePass(const TilePassJob& job, const float* mats, TrajTilePass& tp,
                  uint32_t* rec) {
  const qsim_hip_traj_op* o = job.o;
  const unsigned n = job.n, nops = o[0].k;
  const uint64_t mask = uint64_t(o[0].q[0]) | (uint64_t(o[0].q[1]) << 32);
  unsigned tile[kBlockMaxBits], nt = 0;
  for (unsigned q = 0; q < n; ++q) {
    if ((mask >> q) & 1u) tile[nt++] = q;
  }
  std::vector<unsigned> ks(nops), qs(4 * size_t(nops), 0);
  std::vector<double> m64;
  std::vector<size_t> moff(nops);
  for (unsigned j = 0; j < nops; ++j) {
    const qsim_hip_traj_op& gop = o[j + 1];
    ks[j] = gop.k;
    for (unsigned b = 0; b < gop.k; ++b) qs[4 * j + b] = gop.q[b];
    moff[j] = m64.size();
    const size_t nf = size_t{2} << (2 * gop.k);
    m64.insert(m64.end(), mats + gop.moff, mats + gop.moff + nf);
  }
  std::vector<uint16_t> fb;
  const int rc = PrepareGateBlock(n, tile, nt, nops, ks.data(), qs.data(),
                                  m64.data(), moff.data(), tp.g, tp.p, fb,
                                  kBlockArithF16);
  if (rc != QSIM_HIP_OK) return rc;
  if (tp.p.arith != kBlockArithF16) {
    // Matrices that could grow the state by 2^10: not Kraus operators.
    SetError("tile pass gates grow the state by more than 2^10 (run the "
             "per-gate kernels)");
    return QSIM_HIP_ERR_UNSUPPORTED;
  }
  std::memcpy(rec, fb.data(), fb.size() * sizeof(uint16_t));
  return QSIM_HIP_OK;
}

// Builds the distinct passes, spread over a few host threads that take jobs
// from a shared counter (each pass is independent; the lane-position memo is
// shared under its lock).  A thread that cannot be started leaves its share
// to the others, and a failed host allocation in a worker becomes an error
// return instead of ending the process.
int BuildTilePasses(const std::vector<TilePassJob>& jobs, const float* mats,
                    std::vector<TrajTilePass>& passes,
                    std::vector<uint32_t>& recs) {
  constexpr size_t kRec = kF16RecWords;  // words per record
  passes.assign(jobs.size(), TrajTilePass{});
  size_t total = 0;
  for (size_t j = 0; j < jobs.size(); ++j) {
    passes[j].rec = total;
    total += jobs[j].o[0].k;
  }
  recs.assign(total * kRec, 0);
  constexpr int kHostAlloc = 1;  // worker-side marker, never returned
  std::vector<int> rc(jobs.size(), QSIM_HIP_OK);
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const unsigned nw = unsigned(std::min<size_t>(
      std::min(hw, 16u), std::max<size_t>(1, jobs.size() / 64)));
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (size_t j; (j = next.fetch_add(1)) < jobs.size();) {
      try {
        rc[j] = BuildTilePass(jobs[j], mats, passes[j],
                              recs.data() + passes[j].rec * kRec);
      } catch (const std::bad_alloc&) {
        rc[j] = kHostAlloc;
      }
    }
  };
  std::vector<std::thread> pool;
  pool.reserve(nw);
  for (unsigned w = 1; w < nw; ++w) {
    try {
      pool.emplace_back(work);
    } catch (const std::system_error&) {
      break;
    }
  }
  work();
  for (auto& th : pool) th.join();
  for (size_t j = 0; j < jobs.size(); ++j) {
    if (rc[j] == kHostAlloc) {
      SetError("traj_prepare: out of host memory building tile pass %zu", j);
      return QSIM_HIP_ERR_RUNTIME;
    }
    if (rc[j] != QSIM_HIP_OK) {
      // Again on this thread, for its error message.
      return BuildTilePass(jobs[j], mats, passes[j],
                           recs.data() + passes[j].rec * kRec);
    }
  }
  return QSIM_HIP_OK;
}

__global__ void FillInt(int32_t* __restrict__ p, unsigned n, int32_t v) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
int Upload(T** dst, const T* src, size_t count) {
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  QHIP_CHECK(hipMalloc(dst, bytes));
  if (count) {
    QHIP_CHECK(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  }
  return QSIM_HIP_OK;
}

void Release(TrajBatch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->gstream) (void)hipStreamDestroy(b->gstream);
  (void)hipFree(b->descs);
  (void)hipFree(b->ops);
  (void)hipFree(b->mats);
  (void)hipFree(b->pairs);
  (void)hipFree(b->pool);
  (void)hipFree(b->zz);
  (void)hipFree(b->norms);
  (void)hipFree(b->dec);
  (void)hipFree(b->passes);
  (void)hipFree(b->recs);
  delete b;
}

}  // namespace

}  // namespace qhip

using namespace qhip;

extern "C" {

static int TrajPrepare(const qsim_hip_traj_desc* trajs, unsigned num_traj,
                       const qsim_hip_traj_op* ops, uint64_t num_ops,
                       const float* matrices, uint64_t num_floats,
                       const uint32_t* pairs, uint32_t num_pairs,
                       uint32_t num_decisions, void** handle) {
  if (handle == nullptr || trajs == nullptr || num_traj == 0 ||
      (num_ops && ops == nullptr) || (num_floats && matrices == nullptr) ||
      (num_pairs && pairs == nullptr)) {
    SetError("trajectory batch: null or empty operand");
    return QSIM_HIP_ERR_ARG;
  }
  *handle = nullptr;
  QHIP_TRY(EnsureInit());
  // Validate every stream against the kernel's assumptions before anything
  // reaches the device.
  uint64_t pool = 0;
  std::vector<qsim_hip_traj_desc> d(trajs, trajs + num_traj);
  for (unsigned t = 0; t < num_traj; ++t) {
    const unsigned n = d[t].num_qubits;
    if (n < 4 || n > 34 || d[t].op_begin > d[t].op_end ||
        d[t].op_end > num_ops ||
        uint64_t(d[t].pair_offset) + d[t].num_pairs > num_pairs) {
      SetError("trajectory %u: bad size, operation range or pair range", t);
      return QSIM_HIP_ERR_ARG;
    }
    d[t].state_offset = pool;
    pool += uint64_t{1} << n;
    for (uint64_t i = d[t].op_begin; i < d[t].op_end; ++i) {
      const qsim_hip_traj_op& o = ops[i];
      bool ok;
      if (o.kind == QSIM_HIP_TRAJ_GATE) {
        ok = o.k >= 1 && o.k <= 4 &&
             uint64_t(o.moff) + (size_t{2} << (2 * o.k)) <= num_floats;
        for (unsigned b = 0; ok && b < o.k; ++b)
          ok = o.q[b] < n && (b == 0 || o.q[b] > o.q[b - 1]);
      } else if (o.kind == QSIM_HIP_TRAJ_CHANNEL) {
        // The header of a channel decision and its Kraus operators: each K
        // on qubits of the channel, K^dagger K over the channel's qubits.
        const unsigned nq = o.moff2;
        ok = o.k >= 1 && o.k <= kMaxKraus && i + o.k < d[t].op_end &&
             nq >= 1 && nq <= kMaxChannelQubits && o.slot < num_decisions;
        uint64_t cmask = 0;
        for (unsigned b = 0; ok && b < nq; ++b) {
          ok = o.q[b] < n && (b == 0 || o.q[b] > o.q[b - 1]);
          if (ok) cmask |= uint64_t{1} << o.q[b];
        }
        for (unsigned j = 1; ok && j <= o.k; ++j) {
          const qsim_hip_traj_op& c = ops[i + j];
          ok = c.kind == QSIM_HIP_TRAJ_KRAUS && c.k <= nq &&
               (c.k == 0 ||
                uint64_t(c.moff) + (size_t{2} << (2 * c.k)) <= num_floats) &&
               ((c.pad & 1u) ||
                uint64_t(c.moff2) + (size_t{2} << (2 * nq)) <= num_floats);
          for (unsigned b = 0; ok && b < c.k; ++b) {
            ok = c.q[b] < 64 && ((cmask >> c.q[b]) & 1u) &&
                 (b == 0 || c.q[b] > c.q[b - 1]);
          }
        }
        if (ok) i += o.k;  // the Kraus operators are part of the header
      } else if (o.kind == QSIM_HIP_TRAJ_TILE_PASS) {
        // The header of a pass over the gates that follow it (each checked
        // as a gate in its turn; their tile membership when the pass is
        // built).
        // The tile holds qubits 0, 1 and 2 (the contiguous low run) and
        // every qubit of the pass's gates.
        const uint64_t mask = uint64_t(o.q[0]) | (uint64_t(o.q[1]) << 32);
        const unsigned nt = unsigned(__builtin_popcountll(mask));
        // At most as many gates as the tile kernel keeps resident.
        ok = o.k >= 1 && o.k <= BlockResidentOps(nt) && i + o.k < d[t].op_end &&
             (mask >> n) == 0 && nt >= kBlockMinBits && nt <= kBlockMaxBits &&
             (mask & 7u) == 7u;
        for (unsigned j = 1; ok && j <= o.k; ++j) {
          const qsim_hip_traj_op& g = ops[i + j];
          ok = g.kind == QSIM_HIP_TRAJ_GATE && g.k >= 1 && g.k <= 4;
          for (unsigned b = 0; ok && b < g.k; ++b)
            ok = g.q[b] < 64 && ((mask >> g.q[b]) & 1u);
        }
      } else {
        ok = false;
      }
      if (!ok) {
        SetError("trajectory %u: operation %llu is malformed", t,
                 (unsigned long long)(i - d[t].op_begin));
        return QSIM_HIP_ERR_ARG;
      }
    }
    for (uint32_t p = 0; p < d[t].num_pairs; ++p) {
      const uint32_t* pp = pairs + 2 * (size_t(d[t].pair_offset) + p);
      if (pp[0] >= n || pp[1] >= n || pp[0] == pp[1]) {
        SetError("trajectory %u: bad qubit pair", t);
        return QSIM_HIP_ERR_ARG;
      }
    }
  }
  // Tile passes to their device form; the header's moff becomes the pass
  // index.
  std::vector<qsim_hip_traj_op> ov(ops, ops + num_ops);
  std::vector<TrajTilePass> passes;
  std::vector<uint32_t> recs;
  {
    std::vector<TilePassJob> jobs;
    std::unordered_map<std::string, uint32_t> seen;
    for (unsigned t = 0; t < num_traj; ++t) {
      for (uint64_t i = d[t].op_begin; i < d[t].op_end; ++i) {
        if (ov[i].kind != QSIM_HIP_TRAJ_TILE_PASS) continue;
        ov[i].moff = TilePassIndex(d[t].num_qubits, ops + i, jobs, seen);
      }
    }
    QHIP_TRY(BuildTilePasses(jobs, matrices, passes, recs));
    if (!passes.empty()) QHIP_TRY(TileKernelReady());
  }
  // Released (device buffers included) on every early return or exception.
  std::unique_ptr<TrajBatch, void (*)(TrajBatch*)> b(new TrajBatch, Release);
  b->num_traj = num_traj;
  for (const TrajTilePass& tp : passes) {
    b->max_tile_bits = std::max(b->max_tile_bits, tp.g.tb);
  }
  b->num_pairs = num_pairs;
  b->num_decisions = num_decisions;
  for (unsigned t = 0; t < num_traj; ++t) {
    b->state_offset.push_back(d[t].state_offset);
    b->nq.push_back(d[t].num_qubits);
    b->normalize.push_back(d[t].pad & 1u);
  }
  int rc = Upload(&b->descs, d.data(), d.size());
  if (rc == QSIM_HIP_OK) rc = Upload(&b->ops, ov.data(), num_ops);
  if (rc == QSIM_HIP_OK) rc = Upload(&b->passes, passes.data(), passes.size());
  if (rc == QSIM_HIP_OK) rc = Upload(&b->recs, recs.data(), recs.size());
  if (rc == QSIM_HIP_OK) rc = Upload(&b->mats, matrices, num_floats);
  if (rc == QSIM_HIP_OK) rc = Upload(&b->pairs, pairs, size_t(2) * num_pairs);
  if (rc == QSIM_HIP_OK) {
    hipError_t e = hipMalloc(&b->pool, pool * sizeof(float2));
    if (e == hipSuccess) e = hipMalloc(&b->zz, std::max<size_t>(num_pairs, 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&b->norms, size_t(num_traj) * 8);
    if (e == hipSuccess)
      e = hipMalloc(&b->dec, std::max<size_t>(num_decisions, 1) * 4);
    if (e != hipSuccess) rc = HipFail(e, "trajectory batch buffers");
  }
  if (rc != QSIM_HIP_OK) return rc;
  *handle = b.release();
  return QSIM_HIP_OK;
}

// A failed host allocation (the validated copies of the streams, the tile
// passes, the batch record) becomes an error return: an exception may not
// cross the C ABI.
int qsim_hip_traj_prepare(const qsim_hip_traj_desc* trajs, unsigned num_traj,
                          const qsim_hip_traj_op* ops, uint64_t num_ops,
                          const float* matrices, uint64_t num_floats,
                          const uint32_t* pairs, uint32_t num_pairs,
                          uint32_t num_decisions, void** handle) {
  try {
    return TrajPrepare(trajs, num_traj, ops, num_ops, matrices, num_floats,
                       pairs, num_pairs, num_decisions, handle);
  } catch (const std::bad_alloc&) {
    if (handle) *handle = nullptr;
    SetError("traj_prepare: out of host memory");
    return QSIM_HIP_ERR_RUNTIME;
  }
}

int qsim_hip_traj_run(void* handle, unsigned threads, int use_graph) {
  QHIP_TRY(FlushGateQueue());
  auto* b = static_cast<TrajBatch*>(handle);
  if (b == nullptr) return QSIM_HIP_ERR_ARG;
  // Undecided slots read -1 (the decision kernel writes the reached ones).
  hipLaunchKernelGGL(FillInt, dim3((b->num_decisions + 255) / 256 + 1),
                     dim3(256), 0, Stream(), b->dec, b->num_decisions, -1);
  QHIP_LAUNCHED("FillInt");
  if (!use_graph) {
    PassTimerBegin();
    QHIP_TRY(Launch(*b, threads, Stream()));
    PassTimerEnd();
    QHIP_CHECK(hipStreamSynchronize(Stream()));
    return QSIM_HIP_OK;
  }
  QHIP_CHECK(hipStreamSynchronize(Stream()));
  if (b->exec == nullptr || b->graph_threads != threads) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    if (b->gstream == nullptr) QHIP_CHECK(hipStreamCreate(&b->gstream));
    QHIP_CHECK(hipStreamBeginCapture(b->gstream, hipStreamCaptureModeThreadLocal));
    int rc = Launch(*b, threads, b->gstream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->gstream, &g);
    if (rc != QSIM_HIP_OK) return rc;
    if (e != hipSuccess) return HipFail(e, "hipStreamEndCapture");
    b->graph = g;
    QHIP_CHECK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    b->graph_threads = threads;
  }
  QHIP_CHECK(hipGraphLaunch(b->exec, b->gstream));
  QHIP_CHECK(hipStreamSynchronize(b->gstream));
  return QSIM_HIP_OK;
}

int qsim_hip_traj_results(void* handle, double* zz, uint32_t num_pairs,
                          int32_t* decisions, uint32_t num_decisions) {
  auto* b = static_cast<TrajBatch*>(handle);
  if (b == nullptr || num_pairs > b->num_pairs ||
      num_decisions > b->num_decisions) {
    SetError("trajectory batch: result sizes exceed the batch");
    return QSIM_HIP_ERR_ARG;
  }
  if (zz && num_pairs)
    QHIP_CHECK(hipMemcpy(zz, b->zz, num_pairs * 8ull, hipMemcpyDeviceToHost));
  if (decisions && num_decisions)
    QHIP_CHECK(hipMemcpy(decisions, b->dec, num_decisions * 4ull,
                         hipMemcpyDeviceToHost));
  return QSIM_HIP_OK;
}

int qsim_hip_traj_state(void* handle, unsigned t, float* out) {
  auto* b = static_cast<TrajBatch*>(handle);
  if (b == nullptr || t >= b->num_traj || out == nullptr) {
    SetError("trajectory batch: bad trajectory index");
    return QSIM_HIP_ERR_ARG;
  }
  const size_t count = size_t{2} << b->nq[t];
  QHIP_CHECK(hipMemcpy(out, b->pool + b->state_offset[t],
                       count * sizeof(float), hipMemcpyDeviceToHost));
  if (b->normalize[t]) {
    // The end-of-trajectory normalisation (the norm from the final sweep).
    double nrm = 0.0;
    QHIP_CHECK(hipMemcpy(&nrm, b->norms + t, sizeof(double),
                         hipMemcpyDeviceToHost));
    const float a = float(1.0 / std::sqrt(nrm));
    for (size_t i = 0; i < count; ++i) out[i] *= a;
  }
  return QSIM_HIP_OK;
}

int qsim_hip_traj_free(void* handle) {
  auto* b = static_cast<TrajBatch*>(handle);
  if (b == nullptr) return QSIM_HIP_OK;
  QHIP_CHECK(hipStreamSynchronize(Stream()));
  Release(b);
  return QSIM_HIP_OK;
}

}  // extern "C"
