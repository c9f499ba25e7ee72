// gate_kernels.hip -- k-qubit gate passes and <psi|U|psi> on gfx950.
//
// Two kernel families apply psi <- U psi on every 2^k-amplitude group
// (the operation of SimulatorCUDA::ApplyGate / ApplyControlledGate,
// lib/simulator_cuda.h:70-207; arithmetic as lib/simulator_basic.h:167-275):
//
//  * GateRegKernel  -- all targets and controls at qubit >= 5.  One group per
//    work item, held in registers; the 64 lanes of a wave read 64 consecutive
//    float2/double2 for every group member, so each member load is one
//    contiguous 512 B (complex64) stream.  No LDS, no barriers.
//  * GateTileKernel -- some target or control below qubit 5.  A workgroup
//    stages a tile of 2^tb amplitudes (the Lc lowest qubits, contiguous in
//    memory, times the high targets) in LDS with coalesced loads, applies the
//    groups from LDS and streams the tile back.
//
// The gate matrix (row-major, row/column bit j <-> qs[j]) travels by value in
// the kernel arguments when it fits in 2 KiB (scalar loads, uniform across the
// wave) and otherwise through a device buffer.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.h"
#include "reduce.h"
#include "tile.h"

namespace qhip {

namespace {

template <typename FP, int K>
struct MatVal {
  FP v[2 << (2 * K)];
  __device__ __forceinline__ FP operator()(int i) const { return v[i]; }
};

template <typename FP>
struct MatPtr {
  const FP* __restrict__ p;
  __device__ __forceinline__ FP operator()(int i) const { return p[i]; }
};

// Largest tile per element type: 64 KiB of LDS (2^13 complex64 or 2^12
// complex128).  Passes with k <= 5 use half of it unless the larger tile
// splits into fewer contiguous runs; k = 6 always uses it.
template <typename FP, int K>
constexpr unsigned TileBitsMax() {
  return sizeof(FP) == 4 ? 13u : 12u;
}

template <int K>
__device__ __forceinline__ uint64_t MemberOffset(unsigned j,
                                                 const uint64_t* tbit) {
  uint64_t o = 0;
#pragma unroll
  for (int b = 0; b < K; ++b) o += ((j >> b) & 1) ? tbit[b] : 0;
  return o;
}

// out = U v for one group; writes row r through `store(r, value)`.
template <typename FP, int K, typename M, typename Store>
__device__ __forceinline__ void MatVec(const typename Cplx<FP>::T* v,
                                       const M& m, Store store) {
  using C = typename Cplx<FP>::T;
  constexpr int D = 1 << K;
  if constexpr (K <= 4) {
#pragma unroll
    for (int r = 0; r < D; ++r) {
      FP re = 0, im = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const FP mr = m(2 * (r * D + c));
        const FP mi = m(2 * (r * D + c) + 1);
        re = fma(mr, v[c].x, re);
        re = fma(-mi, v[c].y, re);
        im = fma(mi, v[c].x, im);
        im = fma(mr, v[c].y, im);
      }
      store(r, C{re, im});
    }
  } else {
#pragma unroll 1
    for (int r = 0; r < D; ++r) {
      FP re = 0, im = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const FP mr = m(2 * (r * D + c));
        const FP mi = m(2 * (r * D + c) + 1);
        re = fma(mr, v[c].x, re);
        re = fma(-mi, v[c].y, re);
        im = fma(mi, v[c].x, im);
        im = fma(mr, v[c].y, im);
      }
      store(r, C{re, im});
    }
  }
}

template <typename FP, int K, typename M>
__global__ void __launch_bounds__(256)
GateRegKernel(typename Cplx<FP>::T* __restrict__ s, GateGeom g, M m) {
  using C = typename Cplx<FP>::T;
  constexpr int D = 1 << K;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
       t < g.ngroups; t += stride) {
    const uint64_t base = InsertZeros(t, g.ins, g.nins) | g.cbits;
    C v[D];
#pragma unroll
    for (int j = 0; j < D; ++j) v[j] = s[base + MemberOffset<K>(j, g.tbit)];
    MatVec<FP, K>(v, m, [&](int r, C val) {
      s[base + MemberOffset<K>(r, g.tbit)] = val;
    });
  }
}

template <int K>
__device__ __forceinline__ unsigned TileMemberOffset(unsigned j,
                                                     const unsigned* tpos) {
  unsigned o = 0;
#pragma unroll
  for (int b = 0; b < K; ++b) o |= ((j >> b) & 1u) << tpos[b];
  return o;
}

template <typename FP, int K, typename M>
__global__ void __launch_bounds__(256)
GateTileKernel(typename Cplx<FP>::T* __restrict__ s, TileGeom g, M m) {
  using C = typename Cplx<FP>::T;
  constexpr int D = 1 << K;
  extern __shared__ char lds_raw[];
  C* lds = reinterpret_cast<C*>(lds_raw);
  const unsigned T = 1u << g.tb;
  const unsigned ngroups = T >> K;
  constexpr int PER = (size_t{1} << TileBitsMax<FP, K>()) / 256;
  for (uint64_t tile = blockIdx.x; tile < g.ntiles; tile += gridDim.x) {
    const uint64_t base =
        InsertZeros(tile << g.Lc, g.outer, g.nouter) | g.cbits_high;
    TileLoad<PER>(s, base, g, lds);
    __syncthreads();
    for (unsigned gr = threadIdx.x; gr < ngroups; gr += 256) {
      unsigned tb0 = gr;
#pragma unroll
      for (int b = 0; b < K; ++b) {
        const unsigned p = g.tpos[b];
        tb0 = ((tb0 >> p) << (p + 1)) | (tb0 & ((1u << p) - 1));
      }
      if ((tb0 & g.low_cmask) != g.low_cvals) continue;
      C v[D];
#pragma unroll
      for (int j = 0; j < D; ++j)
        v[j] = lds[TileSwz(tb0 + TileMemberOffset<K>(j, g.tpos))];
      MatVec<FP, K>(v, m, [&](int r, C val) {
        lds[TileSwz(tb0 + TileMemberOffset<K>(r, g.tpos))] = val;
      });
    }
    __syncthreads();
    TileStore<PER>(s, base, g, lds);
    __syncthreads();
  }
}

// <psi|U|psi> partial sums; any target positions (reads through the cache
// hierarchy, like the gate kernels' register path).
template <typename FP, int K, typename M>
__global__ void __launch_bounds__(256)
ExpectationKernel(const typename Cplx<FP>::T* __restrict__ s, GateGeom g, M m,
                  Dbl2* __restrict__ partials) {
  using C = typename Cplx<FP>::T;
  constexpr int D = 1 << K;
  Dbl2 acc;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
       t < g.ngroups; t += stride) {
    const uint64_t base = InsertZeros(t, g.ins, g.nins);
    C v[D];
#pragma unroll
    for (int j = 0; j < D; ++j) v[j] = s[base + MemberOffset<K>(j, g.tbit)];
    MatVec<FP, K>(v, m, [&](int r, C u) {
      // conj(v_r) * (U v)_r: products in FP, sums in FP64
      // (lib/simulator_basic.h:316-317).
      acc.x += double(v[r].x * u.x + v[r].y * u.y);
      acc.y += double(v[r].x * u.y - v[r].y * u.x);
    });
  }
  acc = BlockSum<256>(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// ---- host side -------------------------------------------------------------

uint64_t Deposit(uint64_t v, uint64_t mask) {
  uint64_t r = 0;
  for (unsigned i = 0; i < 64; ++i) {
    if ((mask >> i) & 1) {
      r |= (v & 1) << i;
      v >>= 1;
    }
  }
  return r;
}

int CheckQubits(unsigned n, const unsigned* qs, unsigned k, const unsigned* cqs,
                unsigned nc) {
  for (unsigned i = 0; i < k; ++i) {
    if (qs[i] >= n || (i > 0 && qs[i] <= qs[i - 1])) {
      SetError("target qubits must be ascending, distinct and < %u", n);
      return QSIM_HIP_ERR_ARG;
    }
  }
  for (unsigned i = 0; i < nc; ++i) {
    if (cqs[i] >= n || (i > 0 && cqs[i] <= cqs[i - 1])) {
      SetError("control qubits must be ascending, distinct and < %u", n);
      return QSIM_HIP_ERR_ARG;
    }
    for (unsigned j = 0; j < k; ++j) {
      if (cqs[i] == qs[j]) {
        SetError("qubit %u is both target and control", cqs[i]);
        return QSIM_HIP_ERR_ARG;
      }
    }
  }
  if (k + nc > kMaxInserted) {
    SetError("too many gate qubits (%u)", k + nc);
    return QSIM_HIP_ERR_UNSUPPORTED;
  }
  return QSIM_HIP_OK;
}

template <typename FP, int K>
constexpr bool kByValue = sizeof(FP) * (2u << (2 * K)) <= 2048;



template <typename FP, int K, typename M>
int LaunchReg(typename Cplx<FP>::T* s, const GateGeom& g, const M& m) {
  uint64_t blocks = (g.ngroups + 255) / 256;
  blocks = std::min<uint64_t>(std::max<uint64_t>(blocks, 1), uint64_t{1} << 20);
  PassTimerBegin();
  hipLaunchKernelGGL((GateRegKernel<FP, K, M>), dim3(unsigned(blocks)),
                     dim3(256), 0, Stream(), s, g, m);
  PassTimerEnd();
  QHIP_LAUNCHED("GateRegKernel");
  return QSIM_HIP_OK;
}

template <typename FP, int K, typename M>
int LaunchTile(typename Cplx<FP>::T* s, const TileGeom& g, const M& m) {
  using C = typename Cplx<FP>::T;
  size_t lds = (size_t{1} << g.tb) * sizeof(C);
  if (lds > 65536) {
    SetError("tile of %zu bytes exceeds the 64 KiB LDS budget", lds);
    return QSIM_HIP_ERR_RUNTIME;
  }
  uint64_t blocks = std::min<uint64_t>(g.ntiles, uint64_t{1} << 20);
  PassTimerBegin();
  hipLaunchKernelGGL((GateTileKernel<FP, K, M>), dim3(unsigned(blocks)),
                     dim3(256), lds, Stream(), s, g, m);
  PassTimerEnd();
  QHIP_LAUNCHED("GateTileKernel");
  return QSIM_HIP_OK;
}


template <typename FP, int K>
int DispatchK(typename Cplx<FP>::T* s, unsigned n, const unsigned* qs,
              const unsigned* cqs, unsigned nc, uint64_t cvals, const FP* m) {
  // Lowest qubit the pass touches decides the kernel family.
  unsigned lowest = 64;
  if (K > 0) lowest = std::min(lowest, qs[0]);
  if (nc > 0) lowest = std::min(lowest, cqs[0]);

  MatVal<FP, K> mv;
  MatPtr<FP> mp{nullptr};
  constexpr bool byval = kByValue<FP, K>;
  if constexpr (byval) {
    std::memcpy(mv.v, m, sizeof(mv.v));
  } else {
    size_t bytes = sizeof(FP) * (size_t{2} << (2 * K));
    void* d = MatrixBuffer(bytes);
    if (d == nullptr) return QSIM_HIP_ERR_OOM;
    QHIP_TRY(StageMatrix(d, m, bytes));
    mp.p = static_cast<const FP*>(d);
  }

  if (lowest >= 5 || n < 5) {
    GateGeom g{};
    std::vector<unsigned> ins(qs, qs + K);
    ins.insert(ins.end(), cqs, cqs + nc);
    std::sort(ins.begin(), ins.end());
    g.nins = ins.size();
    for (unsigned i = 0; i < g.nins; ++i) g.ins[i] = ins[i];
    uint64_t cmask = 0;
    for (unsigned i = 0; i < nc; ++i) cmask |= uint64_t{1} << cqs[i];
    g.cbits = Deposit(cvals, cmask);
    for (int b = 0; b < K; ++b) g.tbit[b] = uint64_t{1} << qs[b];
    g.ngroups = uint64_t{1} << (n - K - nc);
    if constexpr (byval) return LaunchReg<FP, K>(s, g, mv);
    else return LaunchReg<FP, K>(s, g, mp);
  }

  // Tile geometry: find tb <= n, h with h = #targets >= tb - h, and enough
  // outer bits left for the high controls.  The tile is 2^h contiguous runs;
  // the larger tile (64 KiB) is taken when it leaves fewer runs, or for k = 6.
  TileGeom g{};
  auto fit = [&](unsigned TB, unsigned* hout, unsigned* lcout,
                 unsigned* tbout) {
    for (;;) {
      unsigned h = 0, Lc = TB;
      for (unsigned hh = 0; hh <= unsigned(K); ++hh) {
        if (hh > TB) break;
        unsigned lc = TB - hh;
        unsigned cnt = 0;
        for (int b = 0; b < K; ++b) cnt += qs[b] >= lc;
        if (cnt == hh) {
          h = hh;
          Lc = lc;
          break;
        }
      }
      unsigned hc = 0;
      for (unsigned i = 0; i < nc; ++i) hc += cqs[i] >= Lc;
      if (n >= TB + hc) {
        *hout = h;
        *lcout = Lc;
        *tbout = TB;
        return;
      }
      --TB;
    }
  };
  const unsigned small_tb = TileBitsMax<FP, K>() - (K >= 6 ? 0u : 1u);
  fit(std::min(small_tb, n), &g.h, &g.Lc, &g.tb);
  if (K < 6 && n > small_tb) {
    unsigned h2, lc2, tb2;
    fit(std::min(TileBitsMax<FP, K>(), n), &h2, &lc2, &tb2);
    if (h2 < g.h) {
      g.h = h2;
      g.Lc = lc2;
      g.tb = tb2;
    }
  }
  std::vector<unsigned> outer;
  unsigned hi = 0;
  for (int b = 0; b < K; ++b) {
    if (qs[b] >= g.Lc) {
      g.hbit[hi] = uint64_t{1} << qs[b];
      g.tpos[b] = g.Lc + hi;
      ++hi;
      outer.push_back(qs[b]);
    } else {
      g.tpos[b] = qs[b];
    }
  }
  uint64_t cmask_high = 0;
  unsigned low_cmask = 0;
  uint64_t cvals_high = 0, cvals_low = 0;
  {
    unsigned lo_i = 0, hi_i = 0;
    for (unsigned i = 0; i < nc; ++i) {
      uint64_t bit = (cvals >> i) & 1;
      if (cqs[i] >= g.Lc) {
        cmask_high |= uint64_t{1} << cqs[i];
        cvals_high |= bit << cqs[i];
        outer.push_back(cqs[i]);
        ++hi_i;
      } else {
        low_cmask |= 1u << cqs[i];
        cvals_low |= bit << cqs[i];
        ++lo_i;
      }
    }
    (void)lo_i;
    (void)hi_i;
  }
  std::sort(outer.begin(), outer.end());
  g.nouter = outer.size();
  for (unsigned i = 0; i < g.nouter; ++i) g.outer[i] = outer[i];
  g.cbits_high = cvals_high;
  g.low_cmask = low_cmask;
  g.low_cvals = unsigned(cvals_low);
  unsigned hc = 0;
  for (unsigned i = 0; i < nc; ++i) hc += cqs[i] >= g.Lc;
  // Every tile-local position must lie inside the tile and every amplitude
  // index inside the state; refuse to launch otherwise.
  bool ok = hi == g.h && g.tb == g.Lc + g.h && g.tb <= n && n >= g.tb + hc;
  for (int b = 0; b < K; ++b) ok = ok && g.tpos[b] < g.tb;
  for (int b = 1; b < K; ++b) ok = ok && g.tpos[b] > g.tpos[b - 1];
  if (!ok) {
    SetError("internal: inconsistent tile geometry (n=%u k=%d)", n, K);
    return QSIM_HIP_ERR_RUNTIME;
  }
  g.ntiles = uint64_t{1} << (n - g.tb - hc);
  if constexpr (byval) return LaunchTile<FP, K>(s, g, mv);
  else return LaunchTile<FP, K>(s, g, mp);
}

template <typename FP, int K>
int DispatchExpectation(const typename Cplx<FP>::T* s, unsigned n,
                        const unsigned* qs, const FP* m, double* re,
                        double* im) {
  GateGeom g{};
  g.nins = K;
  for (int b = 0; b < K; ++b) {
    g.ins[b] = qs[b];
    g.tbit[b] = uint64_t{1} << qs[b];
  }
  g.ngroups = uint64_t{1} << (n - K);
  g.cbits = 0;

  std::lock_guard<std::mutex> lock(ScratchLock());
  const unsigned max_blocks = 1024;
  uint64_t blocks = std::min<uint64_t>((g.ngroups + 255) / 256, max_blocks);
  blocks = std::max<uint64_t>(blocks, 1);
  Dbl2* partials = static_cast<Dbl2*>(DeviceScratch(sizeof(Dbl2) * (blocks + 1)));
  if (partials == nullptr) return QSIM_HIP_ERR_OOM;

  MatVal<FP, K> mv;
  MatPtr<FP> mp{nullptr};
  constexpr bool byval = kByValue<FP, K>;
  if constexpr (byval) {
    std::memcpy(mv.v, m, sizeof(mv.v));
    hipLaunchKernelGGL((ExpectationKernel<FP, K, MatVal<FP, K>>),
                       dim3(unsigned(blocks)), dim3(256), 0, Stream(), s, g, mv,
                       partials + 1);
  } else {
    size_t bytes = sizeof(FP) * (size_t{2} << (2 * K));
    void* d = MatrixBuffer(bytes);
    if (d == nullptr) return QSIM_HIP_ERR_OOM;
    QHIP_TRY(StageMatrix(d, m, bytes));
    mp.p = static_cast<const FP*>(d);
    hipLaunchKernelGGL((ExpectationKernel<FP, K, MatPtr<FP>>),
                       dim3(unsigned(blocks)), dim3(256), 0, Stream(), s, g, mp,
                       partials + 1);
  }
  QHIP_LAUNCHED("ExpectationKernel");
  hipLaunchKernelGGL((FinalSumKernel<Dbl2>), dim3(1), dim3(256), 0, Stream(),
                     partials + 1, unsigned(blocks), partials);
  QHIP_LAUNCHED("FinalSumKernel");
  Dbl2 r;
  QHIP_CHECK(hipMemcpyAsync(&r, partials, sizeof(Dbl2), hipMemcpyDeviceToHost,
                            Stream()));
  QHIP_CHECK(hipStreamSynchronize(Stream()));
  *re = r.x;
  *im = r.y;
  return QSIM_HIP_OK;
}

}  // namespace

template <typename FP>
int LaunchGate(typename Cplx<FP>::T* s, unsigned n, const unsigned* qs,
               unsigned k, const unsigned* cqs, unsigned nc, uint64_t cvals,
               const FP* m) {
  if (k > 6) {
    SetError("gates on %u > 6 qubits are not supported", k);
    return QSIM_HIP_ERR_UNSUPPORTED;
  }
  if (k + nc > n) {
    SetError("gate on %u qubits does not fit a %u-qubit state", k + nc, n);
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_TRY(CheckQubits(n, qs, k, cqs, nc));
  switch (k) {
    case 0: return DispatchK<FP, 0>(s, n, qs, cqs, nc, cvals, m);
    case 1: return DispatchK<FP, 1>(s, n, qs, cqs, nc, cvals, m);
    case 2: return DispatchK<FP, 2>(s, n, qs, cqs, nc, cvals, m);
    case 3: return DispatchK<FP, 3>(s, n, qs, cqs, nc, cvals, m);
    case 4: return DispatchK<FP, 4>(s, n, qs, cqs, nc, cvals, m);
    case 5: return DispatchK<FP, 5>(s, n, qs, cqs, nc, cvals, m);
    default: return DispatchK<FP, 6>(s, n, qs, cqs, nc, cvals, m);
  }
}

template <typename FP>
int LaunchExpectation(const typename Cplx<FP>::T* s, unsigned n,
                      const unsigned* qs, unsigned k, const FP* m, double* re,
                      double* im) {
  if (k == 0 || k > 6) {
    SetError("expectation values need 1..6 qubits (got %u)", k);
    return QSIM_HIP_ERR_UNSUPPORTED;
  }
  if (k > n) {
    SetError("operator on %u qubits does not fit a %u-qubit state", k, n);
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_TRY(CheckQubits(n, qs, k, nullptr, 0));
  switch (k) {
    case 1: return DispatchExpectation<FP, 1>(s, n, qs, m, re, im);
    case 2: return DispatchExpectation<FP, 2>(s, n, qs, m, re, im);
    case 3: return DispatchExpectation<FP, 3>(s, n, qs, m, re, im);
    case 4: return DispatchExpectation<FP, 4>(s, n, qs, m, re, im);
    case 5: return DispatchExpectation<FP, 5>(s, n, qs, m, re, im);
    default: return DispatchExpectation<FP, 6>(s, n, qs, m, re, im);
  }
}

template int LaunchGate<float>(float2*, unsigned, const unsigned*, unsigned,
                               const unsigned*, unsigned, uint64_t,
                               const float*);
template int LaunchGate<double>(double2*, unsigned, const unsigned*, unsigned,
                                const unsigned*, unsigned, uint64_t,
                                const double*);
template int LaunchExpectation<float>(const float2*, unsigned, const unsigned*,
                                      unsigned, const float*, double*,
                                      double*);
template int LaunchExpectation<double>(const double2*, unsigned,
                                       const unsigned*, unsigned,
                                       const double*, double*, double*);

}  // namespace qhip
