// f16_kernels.hip -- k = 5, 6 gate passes on FP16 MFMA with scaling, and
// the exact-FP32 and split-BF16 k = 5, 6 passes on the same work split
// (further below).
//
// The pass computes, per 2^k-amplitude group, out = U in as the real
// product C[2D x 32] = A[2D x 2D] B[2D x 32] of mfma_kernels.hip (D = 2^k,
// 32 groups side by side), on v_mfma_f32_32x32x16_f16:
//
//  - the matrix is scaled by a power of two so that its largest real or
//    imaginary part lies in [1, 2), and split A = Ah + Al into two fp16
//    parts (22 significant bits; entries below 2^-3 of the largest keep an
//    absolute error under 2^-25 of it);
//  - each group is scaled by its own power of two so that its largest real
//    or imaginary part lies in [2^5, 2^6) and split B = Bh + Bl the same
//    way (fp16's 5-bit exponent then covers the group, and no value or
//    product overflows);
//  - Al Bh + Ah Bl + Ah Bh is accumulated in FP32 (the dropped Al Bl is
//    2^-22 relative), three MFMAs per K block of 16 reals, and the result
//    is scaled back.  Scaling by powers of two is exact.
//
// Work split: a 256-thread workgroup (4 waves) takes 4 / RB batches of 32
// groups at a time, RB = D / 16 being the 32-row blocks of A; wave w
// computes row block w % RB of batch w / RB.  So a wave needs only its row
// block's fragments (2 parts x D / 8 K blocks x 16 bytes per lane: 64
// registers at k = 6), which it loads once and keeps in registers, and
// every batch's inputs are read once from HBM into LDS and shared by the
// RB waves that use them.
//
// Two data paths:
//  - every target at qubit >= 5 (GateF16RegKernel): a workgroup's 4 / RB
//    batches are 128 / RB consecutive groups whose members are runs of
//    256 bytes or more; they are copied to LDS (member-major), the next set
//    is fetched into registers while this one is computed, and results go
//    straight to HBM (the whole set is in LDS before any of it is written);
//  - a target below qubit 5 (GateF16SetKernel): the same sets of batches,
//    which then span 2^11 amplitudes over the targets and the lowest other
//    qubits, are copied to LDS, computed in place there and copied back
//    (coalesced both ways), with the next set's loads in flight meanwhile.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

namespace qhip {

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef _Float16 v2h __attribute__((ext_vector_type(2)));

template <int K>
struct F16Dims {
  static constexpr int D = 1 << K;    // members per group
  static constexpr int RB = D / 16;   // 32-row blocks of the 2D real rows
  static constexpr int KB = D / 8;    // K blocks of 16 reals (8 members)
  static constexpr int BPI = 4 / RB;  // batches per workgroup step
  static constexpr int G = 32 * BPI;  // groups per workgroup step
  static constexpr int MPL = D / 2;   // members a lane reads per batch
};

// One wave's row block of A: f[kb][part] holds, for lane l, elements
// A[32 rb + l%32][16 kb + 8 (l/32) + e], e = 0..7 (part 0 high, 1 low).
template <int K>
struct F16Frag {
  v8h f[F16Dims<K>::KB][2];
  __device__ __forceinline__ void Load(const uint4* __restrict__ frag,
                                       int rb) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int kb = 0; kb < F16Dims<K>::KB; ++kb)
#pragma unroll
      for (int p = 0; p < 2; ++p)
        f[kb][p] = __builtin_bit_cast(
            v8h, frag[((rb * F16Dims<K>::KB + kb) * 2 + p) * 64 + l]);
  }
};

// A lane reads its batch's inputs through load(t), t < MPL: member
// 8 (t / 4) + 4 (l / 32) + t % 4 of its group l % 32 (lane half h supplies
// members 8 kb + 4 h + q of K block kb).  The inputs are read twice, for
// the group's scale and then K block by K block into the products, so that
// only one K block's values are live at a time.

// The group's scale exponent: its largest |re| or |im| (this lane's and its
// partner lane's l ^ 32 members) to [2^5, 2^6); 0 for an all-zero or
// non-finite group.
template <int K, typename Load>
__device__ __forceinline__ int GroupScale(Load load) {
  float m = 0.0f;
#pragma unroll
  for (int t = 0; t < F16Dims<K>::MPL; ++t) {
    const float2 v = load(t);
    m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
  }
  m = fmaxf(m, __shfl_xor(m, 32));
  // The products read the inputs again (keeping them from here on would
  // hold every K block's values in registers at once).
  asm volatile("" ::: "memory");
  if (!(m > 0.0f) || !(m <= 3.4e38f)) return 0;
  int e;
  (void)frexpf(m, &e);  // m in [2^(e-1), 2^e)
  return 6 - e;
}

// This wave's 32 output rows (32 rb .. 32 rb + 31) of the batch: K block
// kb's inputs scaled by 2^e and split into fp16 (high, low) B operands,
// then Al Bh + Ah Bl + Ah Bh into the FP32 accumulators.
template <int K, typename Load>
__device__ __forceinline__ v16f RowBlock(const F16Frag<K>& a, Load load,
                                         int e) {
  v16f c = v16f{};
#pragma unroll
  for (int kb = 0; kb < F16Dims<K>::KB; ++kb) {
    v8h bh, bl;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float2 v = load(4 * kb + q);
      const v2f x = v2f{ldexpf(v.x, e), ldexpf(v.y, e)};
      // High parts rounded to nearest (one packed conversion); the
      // residuals x - hi are exact in FP32 and v_fma_mix{lo,hi}_f16 round
      // them to fp16 straight from the fp16 high half.
      const v2h hi = __builtin_convertvector(x, v2h);
      uint32_t lw;
      asm("v_fma_mixlo_f16 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]"
          : "=v"(lw)
          : "v"(hi), "v"(x[0]));
      asm("v_fma_mixhi_f16 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
          : "+v"(lw)
          : "v"(hi), "v"(x[1]));
      const v2h lo = __builtin_bit_cast(v2h, lw);
      bh[2 * q] = hi[0];
      bh[2 * q + 1] = hi[1];
      bl[2 * q] = lo[0];
      bl[2 * q + 1] = lo[1];
    }
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.f[kb][1], bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.f[kb][0], bl, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.f[kb][0], bh, c, 0, 0, 0);
  }
  return c;
}

// Lane l's results: the (re, im) pairs of members 16 rb + 4 s + 2 (l/32)
// + u (s < 4, u < 2) of its group, handed to store(member, value) scaled
// by 2^-us.
template <typename Store>
__device__ __forceinline__ void StoreRows(const v16f& c, int rb, int us,
                                          Store store) {
  const int h = (threadIdx.x >> 5) & 1;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int m = 16 * rb + 4 * s + 2 * h + u;
      store(m, float2{ldexpf(c[4 * s + 2 * u], -us),
                      ldexpf(c[4 * s + 2 * u + 1], -us)});
    }
}

__device__ __forceinline__ uint64_t MemberOffset(unsigned m,
                                                 const uint64_t* tbit, int K) {
  uint64_t o = 0;
  for (int b = 0; b < K; ++b) o += ((m >> b) & 1) ? tbit[b] : 0;
  return o;
}

// Every target at qubit >= 5.  LDS: two buffers of G x D amplitudes
// (member-major: slot m G + g), 16 KiB each.  sm: the matrix's scale
// exponent.
template <int K>
__global__ void __launch_bounds__(256)
GateF16RegKernel(float2* __restrict__ s, GateGeom g,
                 const uint4* __restrict__ frag, int sm) {
  using Dm = F16Dims<K>;
  constexpr int G = Dm::G, D = Dm::D;
  constexpr int PERT = G * D / 256;  // amplitudes each thread copies: 8
  __shared__ float2 buf[2][G * D];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int rb = wave % Dm::RB, bsub = wave / Dm::RB;
  const int j = lane & 31, h = lane >> 5;
  F16Frag<K> a;
  a.Load(frag, rb);
  const uint64_t nsets = g.ngroups / G;
  // Copy map: thread t moves group t % G, members t / G + (256 / G) r.
  const unsigned cg = threadIdx.x % G, cm0 = threadIdx.x / G;
  constexpr unsigned kMStep = 256 / G;
  auto fetch = [&](uint64_t set, float2 (&x)[PERT]) {
    const uint64_t base = InsertZeros(set * G + cg, g.ins, g.nins);
#pragma unroll
    for (int r = 0; r < PERT; ++r)
      x[r] = s[base + MemberOffset(cm0 + kMStep * r, g.tbit, K)];
  };
  auto put = [&](const float2 (&x)[PERT], float2* b) {
#pragma unroll
    for (int r = 0; r < PERT; ++r) b[(cm0 + kMStep * r) * G + cg] = x[r];
  };
  float2 pre[PERT];
  uint64_t set = blockIdx.x;
  if (set < nsets) {
    fetch(set, pre);
    put(pre, buf[0]);
  }
  __syncthreads();
  int cur = 0;
  for (; set < nsets; set += gridDim.x) {
    const uint64_t nxt = set + gridDim.x;
    if (nxt < nsets) fetch(nxt, pre);
    // This wave's batch: groups bsub * 32 + j of the set, read once.
    const float2* b = buf[cur] + bsub * 32 + j;
    float2 v[Dm::MPL];
#pragma unroll
    for (int t = 0; t < Dm::MPL; ++t)
      v[t] = b[(8 * (t >> 2) + 4 * h + (t & 3)) * G];
    auto load = [&](int t) { return v[t]; };
    const int e = GroupScale<K>(load);
    const v16f c = RowBlock<K>(a, load, e);
    const uint64_t base = InsertZeros(set * G + bsub * 32 + j, g.ins, g.nins);
    StoreRows(c, rb, e + sm, [&](int m, float2 x) {
      s[base + MemberOffset(m, g.tbit, K)] = x;
    });
    if (nxt < nsets) put(pre, buf[cur ^ 1]);
    __syncthreads();
    cur ^= 1;
  }
}

// A target below qubit 5.  Any 32 consecutive groups of a k-qubit gate
// span the 2^(k+5) amplitudes over the targets and the 5 lowest other
// qubits; a workgroup's 4 / RB batches span the 2^11 amplitudes over the
// targets and the 11 - k lowest other qubits (the set's qubits, which
// always include qubits 0..4, so a set is runs of 32 or more contiguous
// amplitudes).  A set is copied to LDS member-major (slot m G + g, the
// group index XOR-ed with m's low bits so that both the batch reads and
// the copies spread over the banks), its batches are computed, the results
// go back to their slots and the set is copied out; the next set's loads
// are in flight in registers meanwhile.
struct SetGeom {
  uint64_t nsets;
  unsigned pos[11];  // the set's qubits, ascending
  unsigned tmask;    // bit b: pos[b] is a target (member bits in order)
};

template <int K>
__global__ void __launch_bounds__(256)
GateF16SetKernel(float2* __restrict__ s, SetGeom g,
                 const uint4* __restrict__ frag, int sm) {
  using Dm = F16Dims<K>;
  constexpr int G = Dm::G;
  constexpr int PERT = 2048 / 256;  // amplitudes each thread copies
  __shared__ float2 buf[2][2048];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int rb = wave % Dm::RB, bsub = wave / Dm::RB;
  const int j = lane & 31, h = lane >> 5;
  F16Frag<K> a;
  a.Load(frag, rb);
  auto swz = [](unsigned m, unsigned grp) {
    return m * unsigned(G) + (grp ^ (m & 31u));
  };
  // Copy map: element r of this thread is set-local amplitude
  // x = t + 256 r; its offset in the set and its LDS slot.
  uint64_t goff[PERT];
  unsigned slot[PERT];
#pragma unroll
  for (int r = 0; r < PERT; ++r) {
    const unsigned x = threadIdx.x + 256u * r;
    uint64_t o = 0;
    unsigned m = 0, grp = 0, mi = 0, gi = 0;
    for (int b = 0; b < 11; ++b) {
      const unsigned bit = (x >> b) & 1u;
      o |= uint64_t(bit) << g.pos[b];
      if ((g.tmask >> b) & 1u) {
        m |= bit << mi++;
      } else {
        grp |= bit << gi++;
      }
    }
    goff[r] = o;
    slot[r] = swz(m, grp);
  }
  float2 pre[PERT];
  auto fetch = [&](uint64_t set) {
    const float2* p = s + InsertZeros(set, g.pos, 11);
#pragma unroll
    for (int r = 0; r < PERT; ++r) pre[r] = p[goff[r]];
  };
  auto put = [&](float2* b) {
#pragma unroll
    for (int r = 0; r < PERT; ++r) b[slot[r]] = pre[r];
  };
  uint64_t set = blockIdx.x;
  if (set < g.nsets) {
    fetch(set);
    put(buf[0]);
  }
  __syncthreads();
  int cur = 0;
  const unsigned grp = unsigned(bsub) * 32u + unsigned(j);
  for (; set < g.nsets; set += gridDim.x) {
    const uint64_t nxt = set + gridDim.x;
    if (nxt < g.nsets) fetch(nxt);
    float2* b = buf[cur];
    float2 v[Dm::MPL];
#pragma unroll
    for (int t = 0; t < Dm::MPL; ++t)
      v[t] = b[swz(8u * (t >> 2) + 4u * h + (t & 3), grp)];
    auto load = [&](int t) { return v[t]; };
    const int e = GroupScale<K>(load);
    const v16f c = RowBlock<K>(a, load, e);
    // Every wave has read the set before any result overwrites it.
    __syncthreads();
    StoreRows(c, rb, e + sm, [&](int m, float2 x) { b[swz(m, grp)] = x; });
    __syncthreads();
    float2* p = s + InsertZeros(set, g.pos, 11);
#pragma unroll
    for (int r = 0; r < PERT; ++r) p[goff[r]] = b[slot[r]];
    if (nxt < g.nsets) put(buf[cur ^ 1]);
    __syncthreads();
    cur ^= 1;
  }
}

// ---- Exact FP32 on the same work split ------------------------------------
//
// v_mfma_f32_32x32x2f32 takes one member per K step: lane l supplies
// A[32 rb + l%32][2 t + l/32] (the FP32 fragments of mfma_kernels.hip,
// frag[(rb D + t) 64 + l]) and component l/32 of member t of group l%32,
// and its C layout is the FP16 product's, so StoreRows applies.  A wave
// keeps its row block's D fragment values in registers (64 at k = 6) and
// reads component h of each member from the set in LDS: the matrix never
// comes from the cache hierarchy again, which the per-batch fragment
// reloads of GateMfmaRegKernel / GateMfmaTileKernel at k = 5, 6 do.

template <int K>
struct F32Frag {
  float f[F16Dims<K>::D];
  __device__ __forceinline__ void Load(const float* __restrict__ frag,
                                       int rb) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int t = 0; t < F16Dims<K>::D; ++t)
      f[t] = frag[(rb * F16Dims<K>::D + t) * 64 + l];
  }
};

// This wave's 32 output rows of the batch whose member t, component h,
// is at b[stride t] (floats).
template <int K>
__device__ __forceinline__ v16f RowBlockF32(const F32Frag<K>& a,
                                            const float* b, int stride) {
  v16f c = v16f{};
#pragma unroll
  for (int t = 0; t < F16Dims<K>::D; ++t)
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a.f[t], b[stride * t], c, 0, 0,
                                             0);
  return c;
}

template <typename Store>
__device__ __forceinline__ void StoreRowsF32(const v16f& c, int rb,
                                             Store store) {
  const int h = (threadIdx.x >> 5) & 1;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int u = 0; u < 2; ++u)
      store(16 * rb + 4 * s + 2 * h + u,
            float2{c[4 * s + 2 * u], c[4 * s + 2 * u + 1]});
}

// Every target at qubit >= 5 (sets and copies as GateF16RegKernel).
template <int K>
__global__ void __launch_bounds__(256)
GateF32RowRegKernel(float2* __restrict__ s, GateGeom g,
                    const float* __restrict__ frag) {
  using Dm = F16Dims<K>;
  constexpr int G = Dm::G, D = Dm::D;
  constexpr int PERT = G * D / 256;
  __shared__ float2 buf[2][G * D];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int rb = wave % Dm::RB, bsub = wave / Dm::RB;
  const int j = lane & 31, h = lane >> 5;
  F32Frag<K> a;
  a.Load(frag, rb);
  const uint64_t nsets = g.ngroups / G;
  const unsigned cg = threadIdx.x % G, cm0 = threadIdx.x / G;
  constexpr unsigned kMStep = 256 / G;
  auto fetch = [&](uint64_t set, float2 (&x)[PERT]) {
    const uint64_t base = InsertZeros(set * G + cg, g.ins, g.nins);
#pragma unroll
    for (int r = 0; r < PERT; ++r)
      x[r] = s[base + MemberOffset(cm0 + kMStep * r, g.tbit, K)];
  };
  auto put = [&](const float2 (&x)[PERT], float2* b) {
#pragma unroll
    for (int r = 0; r < PERT; ++r) b[(cm0 + kMStep * r) * G + cg] = x[r];
  };
  float2 pre[PERT];
  uint64_t set = blockIdx.x;
  if (set < nsets) {
    fetch(set, pre);
    put(pre, buf[0]);
  }
  __syncthreads();
  int cur = 0;
  for (; set < nsets; set += gridDim.x) {
    const uint64_t nxt = set + gridDim.x;
    if (nxt < nsets) fetch(nxt, pre);
    const float* b =
        reinterpret_cast<const float*>(buf[cur] + bsub * 32 + j) + h;
    const v16f c = RowBlockF32<K>(a, b, 2 * G);
    const uint64_t base = InsertZeros(set * G + bsub * 32 + j, g.ins, g.nins);
    StoreRowsF32(c, rb, [&](int m, float2 x) {
      s[base + MemberOffset(m, g.tbit, K)] = x;
    });
    if (nxt < nsets) put(pre, buf[cur ^ 1]);
    __syncthreads();
    cur ^= 1;
  }
}

// A target below qubit 5 (sets, slots and copies as GateF16SetKernel).
template <int K>
__global__ void __launch_bounds__(256)
GateF32RowSetKernel(float2* __restrict__ s, SetGeom g,
                    const float* __restrict__ frag) {
  using Dm = F16Dims<K>;
  constexpr int G = Dm::G, D = Dm::D;
  constexpr int PERT = 2048 / 256;
  __shared__ float2 buf[2][2048];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int rb = wave % Dm::RB, bsub = wave / Dm::RB;
  const int j = lane & 31, h = lane >> 5;
  F32Frag<K> a;
  a.Load(frag, rb);
  auto swz = [](unsigned m, unsigned grp) {
    return m * unsigned(G) + (grp ^ (m & 31u));
  };
  uint64_t goff[PERT];
  unsigned slot[PERT];
#pragma unroll
  for (int r = 0; r < PERT; ++r) {
    const unsigned x = threadIdx.x + 256u * r;
    uint64_t o = 0;
    unsigned m = 0, grp = 0, mi = 0, gi = 0;
    for (int b = 0; b < 11; ++b) {
      const unsigned bit = (x >> b) & 1u;
      o |= uint64_t(bit) << g.pos[b];
      if ((g.tmask >> b) & 1u) {
        m |= bit << mi++;
      } else {
        grp |= bit << gi++;
      }
    }
    goff[r] = o;
    slot[r] = swz(m, grp);
  }
  float2 pre[PERT];
  auto fetch = [&](uint64_t set) {
    const float2* p = s + InsertZeros(set, g.pos, 11);
#pragma unroll
    for (int r = 0; r < PERT; ++r) pre[r] = p[goff[r]];
  };
  auto put = [&](float2* b) {
#pragma unroll
    for (int r = 0; r < PERT; ++r) b[slot[r]] = pre[r];
  };
  uint64_t set = blockIdx.x;
  if (set < g.nsets) {
    fetch(set);
    put(buf[0]);
  }
  __syncthreads();
  int cur = 0;
  const unsigned grp = unsigned(bsub) * 32u + unsigned(j);
  for (; set < g.nsets; set += gridDim.x) {
    const uint64_t nxt = set + gridDim.x;
    if (nxt < g.nsets) fetch(nxt);
    float2* b = buf[cur];
    const float* bf = reinterpret_cast<const float*>(b);
    v16f c = v16f{};
#pragma unroll
    for (int t = 0; t < D; ++t)
      c = __builtin_amdgcn_mfma_f32_32x32x2f32(
          a.f[t], bf[2 * swz(unsigned(t), grp) + h], c, 0, 0, 0);
    // Every wave has read the set before any result overwrites it.
    __syncthreads();
    StoreRowsF32(c, rb, [&](int m, float2 x) { b[swz(m, grp)] = x; });
    __syncthreads();
    float2* p = s + InsertZeros(set, g.pos, 11);
#pragma unroll
    for (int r = 0; r < PERT; ++r) p[goff[r]] = b[slot[r]];
    if (nxt < g.nsets) put(buf[cur ^ 1]);
    __syncthreads();
    cur ^= 1;
  }
}

// ---- Split-BF16 on the same work split --------------------------------------
//
// v_mfma_f32_32x32x16_bf16 has the FP16 product's operand maps (lane half h
// supplies members 8 kb + 4 h + q of K block kb).  A wave keeps its row
// block's three bf16 parts (hi, mid, lo per K block: 96 registers at
// k = 6) and splits each K block of amplitudes in two (hi, lo) as it reads
// them from the set in LDS; Al Bh + Am Bh + Ah Bl + Ah Bh per K block, as
// the split-BF16 kernels of mfma_kernels.hip accumulate them.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

template <int K>
struct Bf16Frag {
  v8bf f[F16Dims<K>::KB][3];
  __device__ __forceinline__ void Load(const uint4* __restrict__ frag,
                                       int rb) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int kb = 0; kb < F16Dims<K>::KB; ++kb)
#pragma unroll
      for (int p = 0; p < 3; ++p)
        f[kb][p] = __builtin_bit_cast(
            v8bf, frag[((rb * F16Dims<K>::KB + kb) * 3 + p) * 64 + l]);
  }
};

// This wave's 32 output rows; read(t) gives the lane's member
// 8 (t / 4) + 4 h + t % 4 of its group.
template <int K, typename Read>
__device__ __forceinline__ v16f RowBlockBf16(const Bf16Frag<K>& a,
                                             Read read) {
  v16f c = v16f{};
#pragma unroll
  for (int kb = 0; kb < F16Dims<K>::KB; ++kb) {
    v8bf bh, bl;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float2 v = read(4 * kb + q);
      const __bf16 hr = (__bf16)v.x, hi = (__bf16)v.y;
      bh[2 * q] = hr;
      bh[2 * q + 1] = hi;
      bl[2 * q] = (__bf16)(v.x - (float)hr);
      bl[2 * q + 1] = (__bf16)(v.y - (float)hi);
    }
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.f[kb][2], bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.f[kb][1], bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.f[kb][0], bl, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.f[kb][0], bh, c, 0, 0, 0);
  }
  return c;
}

// Every target at qubit >= 5 (sets and copies as GateF16RegKernel).
template <int K>
__global__ void __launch_bounds__(256)
GateBf16RowRegKernel(float2* __restrict__ s, GateGeom g,
                     const uint4* __restrict__ frag) {
  using Dm = F16Dims<K>;
  constexpr int G = Dm::G, D = Dm::D;
  constexpr int PERT = G * D / 256;
  __shared__ float2 buf[2][G * D];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int rb = wave % Dm::RB, bsub = wave / Dm::RB;
  const int j = lane & 31, h = lane >> 5;
  Bf16Frag<K> a;
  a.Load(frag, rb);
  const uint64_t nsets = g.ngroups / G;
  const unsigned cg = threadIdx.x % G, cm0 = threadIdx.x / G;
  constexpr unsigned kMStep = 256 / G;
  auto fetch = [&](uint64_t set, float2 (&x)[PERT]) {
    const uint64_t base = InsertZeros(set * G + cg, g.ins, g.nins);
#pragma unroll
    for (int r = 0; r < PERT; ++r)
      x[r] = s[base + MemberOffset(cm0 + kMStep * r, g.tbit, K)];
  };
  auto put = [&](const float2 (&x)[PERT], float2* b) {
#pragma unroll
    for (int r = 0; r < PERT; ++r) b[(cm0 + kMStep * r) * G + cg] = x[r];
  };
  float2 pre[PERT];
  uint64_t set = blockIdx.x;
  if (set < nsets) {
    fetch(set, pre);
    put(pre, buf[0]);
  }
  __syncthreads();
  int cur = 0;
  for (; set < nsets; set += gridDim.x) {
    const uint64_t nxt = set + gridDim.x;
    if (nxt < nsets) fetch(nxt, pre);
    const float2* b = buf[cur] + bsub * 32 + j;
    const v16f c = RowBlockBf16<K>(
        a, [&](int t) { return b[(8 * (t >> 2) + 4 * h + (t & 3)) * G]; });
    const uint64_t base = InsertZeros(set * G + bsub * 32 + j, g.ins, g.nins);
    StoreRowsF32(c, rb, [&](int m, float2 x) {
      s[base + MemberOffset(m, g.tbit, K)] = x;
    });
    if (nxt < nsets) put(pre, buf[cur ^ 1]);
    __syncthreads();
    cur ^= 1;
  }
}

// A target below qubit 5 (sets, slots and copies as GateF16SetKernel).
template <int K>
__global__ void __launch_bounds__(256)
GateBf16RowSetKernel(float2* __restrict__ s, SetGeom g,
                     const uint4* __restrict__ frag) {
  using Dm = F16Dims<K>;
  constexpr int G = Dm::G;
  constexpr int PERT = 2048 / 256;
  __shared__ float2 buf[2][2048];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int rb = wave % Dm::RB, bsub = wave / Dm::RB;
  const int j = lane & 31, h = lane >> 5;
  Bf16Frag<K> a;
  a.Load(frag, rb);
  auto swz = [](unsigned m, unsigned grp) {
    return m * unsigned(G) + (grp ^ (m & 31u));
  };
  uint64_t goff[PERT];
  unsigned slot[PERT];
#pragma unroll
  for (int r = 0; r < PERT; ++r) {
    const unsigned x = threadIdx.x + 256u * r;
    uint64_t o = 0;
    unsigned m = 0, grp = 0, mi = 0, gi = 0;
    for (int b = 0; b < 11; ++b) {
      const unsigned bit = (x >> b) & 1u;
      o |= uint64_t(bit) << g.pos[b];
      if ((g.tmask >> b) & 1u) {
        m |= bit << mi++;
      } else {
        grp |= bit << gi++;
      }
    }
    goff[r] = o;
    slot[r] = swz(m, grp);
  }
  float2 pre[PERT];
  auto fetch = [&](uint64_t set) {
    const float2* p = s + InsertZeros(set, g.pos, 11);
#pragma unroll
    for (int r = 0; r < PERT; ++r) pre[r] = p[goff[r]];
  };
  auto put = [&](float2* b) {
#pragma unroll
    for (int r = 0; r < PERT; ++r) b[slot[r]] = pre[r];
  };
  uint64_t set = blockIdx.x;
  if (set < g.nsets) {
    fetch(set);
    put(buf[0]);
  }
  __syncthreads();
  int cur = 0;
  const unsigned grp = unsigned(bsub) * 32u + unsigned(j);
  for (; set < g.nsets; set += gridDim.x) {
    const uint64_t nxt = set + gridDim.x;
    if (nxt < g.nsets) fetch(nxt);
    float2* b = buf[cur];
    const v16f c = RowBlockBf16<K>(a, [&](int t) {
      return b[swz(8u * unsigned(t >> 2) + 4u * unsigned(h) + unsigned(t & 3),
                   grp)];
    });
    // Every wave has read the set before any result overwrites it.
    __syncthreads();
    StoreRowsF32(c, rb, [&](int m, float2 x) { b[swz(m, grp)] = x; });
    __syncthreads();
    float2* p = s + InsertZeros(set, g.pos, 11);
#pragma unroll
    for (int r = 0; r < PERT; ++r) p[goff[r]] = b[slot[r]];
    if (nxt < g.nsets) put(buf[cur ^ 1]);
    __syncthreads();
    cur ^= 1;
  }
}

// Host: the fp16 (high, low) fragments of A = the real embedding of
// 2^-sm... scaled matrix, and the scale exponent sm (the kernels multiply
// their results by 2^-(e + sm)).  frag: RB x KB x 2 x 64 lanes x 8 halves.
template <int K>
int BuildFragF16(const double* m, std::vector<uint16_t>& frag) {
  using Dm = F16Dims<K>;
  constexpr int D = Dm::D;
  double mx = 0.0;
  for (int i = 0; i < 2 * D * D; ++i) mx = std::max(mx, std::fabs(m[i]));
  int sm = 0;
  if (mx > 0.0 && std::isfinite(mx)) {
    int ex;
    (void)std::frexp(mx, &ex);  // mx in [2^(ex-1), 2^ex)
    sm = 1 - ex;                // scaled largest part in [1, 2)
  }
  auto A = [&](int i, int k) -> double {
    const int r = i >> 1, ci = i & 1, mm = k >> 1, ck = k & 1;
    const double re = m[2 * (r * D + mm)], im = m[2 * (r * D + mm) + 1];
    if (ci == 0) return ck == 0 ? re : -im;
    return ck == 0 ? im : re;
  };
  frag.assign(size_t(Dm::RB) * Dm::KB * 2 * 64 * 8, 0);
  for (int rb = 0; rb < Dm::RB; ++rb)
    for (int kb = 0; kb < Dm::KB; ++kb)
      for (int l = 0; l < 64; ++l)
        for (int e = 0; e < 8; ++e) {
          const double x =
              std::ldexp(A(32 * rb + (l & 31), 16 * kb + 8 * (l >> 5) + e), sm);
          const _Float16 hi = (_Float16)x;
          const _Float16 lo = (_Float16)(x - (double)hi);
          const size_t o = (((size_t(rb) * Dm::KB + kb) * 2) * 64 + l) * 8 + e;
          std::memcpy(&frag[o], &hi, 2);
          std::memcpy(&frag[o + 64 * 8], &lo, 2);
        }
  return sm;
}

// The set's qubits: the targets and the 11 - k lowest other qubits; false
// when the state has fewer than 11 qubits.
bool MakeSetGeom(unsigned n, const unsigned* qs, int K, SetGeom& g) {
  if (n < 11) return false;
  g = SetGeom{};
  g.nsets = uint64_t{1} << (n - 11);
  unsigned np = 0, others = 0, ti = 0;
  for (unsigned q = 0; q < n && np < 11; ++q) {
    const bool target = ti < unsigned(K) && qs[ti] == q;
    if (target) {
      g.tmask |= 1u << np;
      g.pos[np++] = q;
      ++ti;
    } else if (others < 11u - unsigned(K)) {
      g.pos[np++] = q;
      ++others;
    }
  }
  return np == 11 && ti == unsigned(K);
}

template <int K>
int DispatchF32Rows(float2* s, unsigned n, const unsigned* qs,
                    const float* frag) {
  using Dm = F16Dims<K>;
  if (qs[0] >= 5) {
    GateGeom g{};
    g.nins = K;
    for (int b = 0; b < K; ++b) {
      g.ins[b] = qs[b];
      g.tbit[b] = uint64_t{1} << qs[b];
    }
    g.ngroups = uint64_t{1} << (n - K);
    g.cbits = 0;
    if (g.ngroups < uint64_t(Dm::G)) return QSIM_HIP_ERR_UNSUPPORTED;
    const uint64_t nsets = g.ngroups / Dm::G;
    const uint64_t blocks =
        std::max<uint64_t>(1, std::min<uint64_t>(nsets, uint64_t(NumCUs()) * 4));
    PassTimerBegin();
    hipLaunchKernelGGL(GateF32RowRegKernel<K>, dim3(unsigned(blocks)),
                       dim3(256), 0, Stream(), s, g, frag);
    PassTimerEnd();
    QHIP_LAUNCHED("GateF32RowRegKernel");
    return QSIM_HIP_OK;
  }
  SetGeom g{};
  if (!MakeSetGeom(n, qs, K, g)) return QSIM_HIP_ERR_UNSUPPORTED;
  const uint64_t blocks = std::max<uint64_t>(
      1, std::min<uint64_t>(g.nsets, uint64_t(NumCUs()) * 4));
  PassTimerBegin();
  hipLaunchKernelGGL(GateF32RowSetKernel<K>, dim3(unsigned(blocks)),
                     dim3(256), 0, Stream(), s, g, frag);
  PassTimerEnd();
  QHIP_LAUNCHED("GateF32RowSetKernel");
  return QSIM_HIP_OK;
}

template <int K>
int DispatchBf16Rows(float2* s, unsigned n, const unsigned* qs,
                     const uint4* frag) {
  using Dm = F16Dims<K>;
  if (qs[0] >= 5) {
    GateGeom g{};
    g.nins = K;
    for (int b = 0; b < K; ++b) {
      g.ins[b] = qs[b];
      g.tbit[b] = uint64_t{1} << qs[b];
    }
    g.ngroups = uint64_t{1} << (n - K);
    g.cbits = 0;
    if (g.ngroups < uint64_t(Dm::G)) return QSIM_HIP_ERR_UNSUPPORTED;
    const uint64_t nsets = g.ngroups / Dm::G;
    const uint64_t blocks =
        std::max<uint64_t>(1, std::min<uint64_t>(nsets, uint64_t(NumCUs()) * 4));
    PassTimerBegin();
    hipLaunchKernelGGL(GateBf16RowRegKernel<K>, dim3(unsigned(blocks)),
                       dim3(256), 0, Stream(), s, g, frag);
    PassTimerEnd();
    QHIP_LAUNCHED("GateBf16RowRegKernel");
    return QSIM_HIP_OK;
  }
  SetGeom g{};
  if (!MakeSetGeom(n, qs, K, g)) return QSIM_HIP_ERR_UNSUPPORTED;
  const uint64_t blocks = std::max<uint64_t>(
      1, std::min<uint64_t>(g.nsets, uint64_t(NumCUs()) * 4));
  PassTimerBegin();
  hipLaunchKernelGGL(GateBf16RowSetKernel<K>, dim3(unsigned(blocks)),
                     dim3(256), 0, Stream(), s, g, frag);
  PassTimerEnd();
  QHIP_LAUNCHED("GateBf16RowSetKernel");
  return QSIM_HIP_OK;
}

template <int K>
int DispatchF16(float2* s, unsigned n, const unsigned* qs, const double* m64) {
  using Dm = F16Dims<K>;
  std::vector<uint16_t> frag;
  const int sm = BuildFragF16<K>(m64, frag);
  const size_t fbytes = frag.size() * sizeof(uint16_t);
  void* dfrag = MatrixBuffer(fbytes);
  if (dfrag == nullptr) return QSIM_HIP_ERR_OOM;
  QHIP_TRY(StageMatrix(dfrag, frag.data(), fbytes));
  const uint4* f4 = static_cast<const uint4*>(dfrag);

  if (qs[0] >= 5) {
    GateGeom g{};
    g.nins = K;
    for (int b = 0; b < K; ++b) {
      g.ins[b] = qs[b];
      g.tbit[b] = uint64_t{1} << qs[b];
    }
    g.ngroups = uint64_t{1} << (n - K);
    g.cbits = 0;
    if (g.ngroups < uint64_t(Dm::G)) return QSIM_HIP_ERR_UNSUPPORTED;
    const uint64_t nsets = g.ngroups / Dm::G;
    const uint64_t blocks =
        std::max<uint64_t>(1, std::min<uint64_t>(nsets, uint64_t(NumCUs()) * 4));
    PassTimerBegin();
    hipLaunchKernelGGL(GateF16RegKernel<K>, dim3(unsigned(blocks)), dim3(256),
                       0, Stream(), s, g, f4, sm);
    PassTimerEnd();
    QHIP_LAUNCHED("GateF16RegKernel");
    return QSIM_HIP_OK;
  }

  SetGeom g{};
  if (!MakeSetGeom(n, qs, K, g)) return QSIM_HIP_ERR_UNSUPPORTED;
  const uint64_t blocks = std::max<uint64_t>(
      1, std::min<uint64_t>(g.nsets, uint64_t(NumCUs()) * 4));
  PassTimerBegin();
  hipLaunchKernelGGL(GateF16SetKernel<K>, dim3(unsigned(blocks)), dim3(256), 0,
                     Stream(), s, g, f4, sm);
  PassTimerEnd();
  QHIP_LAUNCHED("GateF16SetKernel");
  return QSIM_HIP_OK;
}

}  // namespace

int LaunchGateF16(float2* s, unsigned n, const unsigned* qs, unsigned k,
                  const double* m64) {
  if (k < 5 || k > 6) return QSIM_HIP_ERR_UNSUPPORTED;
  for (unsigned i = 0; i < k; ++i) {
    if (qs[i] >= n || (i > 0 && qs[i] <= qs[i - 1])) {
      SetError("target qubits must be ascending, distinct and < %u", n);
      return QSIM_HIP_ERR_ARG;
    }
  }
  return k == 5 ? DispatchF16<5>(s, n, qs, m64) : DispatchF16<6>(s, n, qs, m64);
}

int LaunchRowsF32(float2* s, unsigned n, const unsigned* qs, unsigned k,
                  const float* frag) {
  if (k < 5 || k > 6) return QSIM_HIP_ERR_UNSUPPORTED;
  return k == 5 ? DispatchF32Rows<5>(s, n, qs, frag)
                : DispatchF32Rows<6>(s, n, qs, frag);
}

int LaunchRowsBf16(float2* s, unsigned n, const unsigned* qs, unsigned k,
                   const uint16_t* frag) {
  if (k < 5 || k > 6) return QSIM_HIP_ERR_UNSUPPORTED;
  const uint4* f4 = reinterpret_cast<const uint4*>(frag);
  return k == 5 ? DispatchBf16Rows<5>(s, n, qs, f4)
                : DispatchBf16Rows<6>(s, n, qs, f4);
}

}  // namespace qhip
