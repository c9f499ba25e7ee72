// tile.h -- staging of LDS tiles for the low-target gate kernels.
//
// A tile holds 2^tb amplitudes of the state: the Lc lowest qubits (one
// contiguous run of memory) times h high qubits (TileGeom, BlockGeom).  Each
// of the NT threads moves PER = max_tile / NT amplitudes; all of a thread's
// global loads of a chunk are issued before any LDS store, so a chunk costs
// one memory round trip.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.h"

namespace qhip {

// Slot of amplitude a in LDS: XOR of the low 4 bits with bits 4..11 folded,
// so that the 32-64 groups a wave reads at once spread over the banks even
// when the targets occupy the low bits.
__device__ __forceinline__ unsigned TileSwz(unsigned a) {
  return a ^ (((a >> 4) ^ (a >> 8)) & 15u);
}

template <typename G>
__device__ __forceinline__ uint64_t TileAddr(uint64_t base, unsigned a,
                                             const G& g) {
  uint64_t gi = base + (a & ((1u << g.Lc) - 1));
  const unsigned hi = a >> g.Lc;
  for (unsigned b = 0; b < g.h; ++b) gi += ((hi >> b) & 1u) ? g.hbit[b] : 0;
  return gi;
}

// Loads are issued in chunks of at most MAXCH per thread (one round trip per
// chunk) to bound the registers a chunk keeps live.
template <int PER, int NT = 256, int MAXCH = 16, typename C, typename G>
__device__ __forceinline__ void TileLoad(const C* __restrict__ s, uint64_t base,
                                         const G& g, C* lds) {
  constexpr int CH = PER < MAXCH ? PER : MAXCH;
  const unsigned T = 1u << g.tb;
#pragma unroll 1
  for (int i0 = 0; i0 < PER; i0 += CH) {
    if (threadIdx.x + NT * i0 >= T) break;
    C v[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const unsigned a = threadIdx.x + NT * (i0 + i);
      if (a < T) v[i] = s[TileAddr(base, a, g)];
    }
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const unsigned a = threadIdx.x + NT * (i0 + i);
      if (a < T) lds[TileSwz(a)] = v[i];
    }
  }
}

template <int PER, int NT = 256, int MAXCH = 16, typename C, typename G>
__device__ __forceinline__ void TileStore(C* __restrict__ s, uint64_t base,
                                          const G& g, const C* lds) {
  constexpr int CH = PER < MAXCH ? PER : MAXCH;
  const unsigned T = 1u << g.tb;
#pragma unroll 1
  for (int i0 = 0; i0 < PER; i0 += CH) {
    if (threadIdx.x + NT * i0 >= T) break;
    C v[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const unsigned a = threadIdx.x + NT * (i0 + i);
      if (a < T) v[i] = lds[TileSwz(a)];
    }
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const unsigned a = threadIdx.x + NT * (i0 + i);
      if (a < T) s[TileAddr(base, a, g)] = v[i];
    }
  }
}

}  // namespace qhip
