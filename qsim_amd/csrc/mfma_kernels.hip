// mfma_kernels.hip -- k = 4..6 gate passes on the gfx950 matrix cores.
//
// A k-qubit pass is, per 2^k-amplitude group, out = U in.  Embedding complex
// numbers as real pairs (re, im) turns it into a real 2D x 2D matrix
// (D = 2^k) times a 2D-vector, and 32 groups side by side make the MFMA
// product C[2D x 32] = A[2D x 2D] B[2D x 32]:
//
//   A[2r  ][2m] = Re U[r][m]   A[2r  ][2m+1] = -Im U[r][m]
//   A[2r+1][2m] = Im U[r][m]   A[2r+1][2m+1] =  Re U[r][m]
//   B[2m + c][j] = (c ? Im : Re) of member m of group j.
//
// FP32 mode uses v_mfma_f32_32x32x2f32 (exact FP32 products, FP32 sums, twice
// the non-packed vector FP32 rate).  Operand maps (verified on gfx950):
//   a = A[l%32][2t + l/32], b = B[2t + l/32][l%32] for K-step t (member t),
//   C register r of lane l = C[8(r/4) + 4(l/32) + r%4][l%32].
// So lane j < 32 feeds the real parts and lane j + 32 the imaginary parts of
// group j, and lane l receives (re, im) pairs of members
// 16 rb + 4 s + 2 (l/32) + {0, 1} (rb = row block, s = 0..3): each result is
// one float2, and 32 lanes store 32 consecutive amplitudes.
//
// BF16 mode uses v_mfma_f32_32x32x16_bf16 with the matrix split in three
// bf16 parts A = Ah + Am + Al (relative error 2^-27) and the amplitudes in
// two, B = Bh + Bl (2^-18), and A B ~ Al Bh + Am Bh + Ah Bl + Ah Bh
// accumulated in FP32.  The matrix error is the same for every group of a
// pass, so it adds up coherently (a two-part matrix drifts the norm by up
// to 2e-6 per pass on structured states); the amplitude error differs from
// group to group and averages out.  Four MFMAs per K block.
// Its operand maps: a[e] = A[l%32][8(l/32) + e], b[e] = B[8(l/32) + e][l%32]
// for e = 0..7 over a K block of 16, i.e. 4 members per lane half.
//
// Two data paths, as for the vector kernels: targets all at qubit >= 5 read
// the operands straight from HBM (32 lanes = 32 consecutive amplitudes); a
// target below qubit 5 stages a tile in LDS (XOR-swizzled to spread banks).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.h"
#include "tile.h"

namespace qhip {

// Round-to-nearest-even FP32 -> BF16 bits.
inline uint16_t Bf16Bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7fff + ((u >> 16) & 1);
  return uint16_t(u >> 16);
}
inline float Bf16ToFloat(uint16_t b) {
  uint32_t u = uint32_t(b) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// BF16 fragments (hi, mid, lo) per K block kb of 16 reals:
// frag[((rb * KB + kb) * 3 + part) * 64 * 8 + l * 8 + e] = part of
// A[32 rb + l%32][16 kb + 8 (l/32) + e].
void BuildFragBf16(int K, const double* m, uint16_t* frag) {
  const int D = 1 << K, RB = D / 16, KB = 2 * D / 16;
  auto A = [&](int i, int k) -> double {
    int r = i >> 1, ci = i & 1, mm = k >> 1, ck = k & 1;
    double re = m[2 * (r * D + mm)], im = m[2 * (r * D + mm) + 1];
    if (ci == 0) return ck == 0 ? re : -im;
    return ck == 0 ? im : re;
  };
  for (int rb = 0; rb < RB; ++rb)
    for (int kb = 0; kb < KB; ++kb)
      for (int l = 0; l < 64; ++l)
        for (int e = 0; e < 8; ++e) {
          double x = A(32 * rb + (l & 31), 16 * kb + 8 * (l >> 5) + e);
          const uint16_t hi = Bf16Bits(float(x));
          const double r1 = x - double(Bf16ToFloat(hi));
          const uint16_t mid = Bf16Bits(float(r1));
          const uint16_t lo = Bf16Bits(float(r1 - double(Bf16ToFloat(mid))));
          frag[((rb * KB + kb) * 3 + 0) * 512 + l * 8 + e] = hi;
          frag[((rb * KB + kb) * 3 + 1) * 512 + l * 8 + e] = mid;
          frag[((rb * KB + kb) * 3 + 2) * 512 + l * 8 + e] = lo;
        }
}

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int kWaves = 4;  // waves per 256-thread workgroup

template <int K>
struct Dims {
  static constexpr int D = 1 << K;     // members per group
  static constexpr int RB = D / 16;    // 32-row blocks of the 2D rows
};

// FP32 fragments: frag[(rb * D + t) * 64 + l] = A[32 rb + l%32][2t + l/32].
void BuildFragF32(int K, const double* m, float* frag) {
  const int D = 1 << K, RB = D / 16;
  auto A = [&](int i, int k) -> double {
    int r = i >> 1, ci = i & 1, mm = k >> 1, ck = k & 1;
    double re = m[2 * (r * D + mm)], im = m[2 * (r * D + mm) + 1];
    if (ci == 0) return ck == 0 ? re : -im;
    return ck == 0 ? im : re;
  };
  for (int rb = 0; rb < RB; ++rb)
    for (int t = 0; t < D; ++t)
      for (int l = 0; l < 64; ++l)
        frag[(rb * D + t) * 64 + l] = float(A(32 * rb + (l & 31), 2 * t + (l >> 5)));
}

__device__ __forceinline__ uint64_t MemberOff(unsigned m, const uint64_t* tbit,
                                              int K) {
  uint64_t o = 0;
  for (int b = 0; b < K; ++b) o += ((m >> b) & 1) ? tbit[b] : 0;
  return o;
}

// Hides a pointer's value from the optimiser so loads through it are not
// hoisted out of the enclosing loop (k = 6 fragments would not fit).
template <typename T>
__device__ __forceinline__ const T* Launder(const T* p) {
  asm volatile("" : "+s"(p));
  return p;
}

// x = hi + lo + O(2^-18 |x|), both parts bf16 (round to nearest even).
__device__ __forceinline__ __bf16 Bf16Hi(float x) { return (__bf16)x; }
__device__ __forceinline__ __bf16 Bf16Lo(float x) {
  return (__bf16)(x - (float)(__bf16)x);
}

// ---------------------------------------------------------------------------
// FP32 MFMA.  B operands for one batch of 32 groups are gathered by `load(t)`
// (member t; real part on lanes < 32, imaginary on lanes >= 32); results are
// handed to `store(m, float2)` for the members this lane owns.

template <int K, typename AFrag, typename Load, typename Store>
__device__ __forceinline__ void BatchF32(const AFrag& afrag, Load load,
                                         Store store) {
  constexpr int D = Dims<K>::D, RB = Dims<K>::RB;
  const int h = (threadIdx.x >> 5) & 1;
  v16f c[RB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) c[rb] = v16f{};
#pragma unroll
  for (int t = 0; t < D; ++t) {
    const float b = load(t);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
      c[rb] = __builtin_amdgcn_mfma_f32_32x32x2f32(afrag(rb, t), b, c[rb],
                                                   0, 0, 0);
    // k = 6 streams operands and fragments: keep the scheduler from hoisting
    // all 64 steps' loads at once (they would not fit the register file).
    if constexpr (K >= 6) {
      if ((t & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
  }
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int m = 16 * rb + 4 * s + 2 * h + u;
        store(m, float2{c[rb][4 * s + 2 * u], c[rb][4 * s + 2 * u + 1]});
      }
}

// A fragments held in registers (k = 4, 5).
template <int K>
struct RegFragF32 {
  float v[Dims<K>::RB][Dims<K>::D];
  __device__ void Load(const float* __restrict__ frag) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int rb = 0; rb < Dims<K>::RB; ++rb)
#pragma unroll
      for (int t = 0; t < Dims<K>::D; ++t)
        v[rb][t] = frag[(rb * Dims<K>::D + t) * 64 + l];
  }
  __device__ __forceinline__ float operator()(int rb, int t) const {
    return v[rb][t];
  }
};

// A fragments read through the cache hierarchy (k = 6).
template <int K>
struct MemFragF32 {
  const float* p;
  __device__ void Load(const float* frag) { p = Launder(frag); }
  __device__ __forceinline__ float operator()(int rb, int t) const {
    return p[(rb * Dims<K>::D + t) * 64 + (threadIdx.x & 63)];
  }
};

template <int K>
using FragF32 = typename std::conditional<(K <= 5), RegFragF32<K>,
                                          MemFragF32<K>>::type;

template <int K>
__global__ void __launch_bounds__(256)
GateMfmaRegKernel(float* __restrict__ s, GateGeom g,
                  const float* __restrict__ frag) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 31, h = lane >> 5;
  FragF32<K> af;
  af.Load(frag);
  const uint64_t nbatch = g.ngroups >> 5;
  const uint64_t wstride = uint64_t(gridDim.x) * kWaves;
  for (uint64_t b = uint64_t(blockIdx.x) * kWaves + wave; b < nbatch;
       b += wstride) {
    const uint64_t base = InsertZeros(b * 32 + j, g.ins, g.nins) | g.cbits;
    if constexpr (K >= 6) af.Load(frag);
    BatchF32<K>(
        af,
        [&](int t) { return s[2 * (base + MemberOff(t, g.tbit, K)) + h]; },
        [&](int m, float2 v) {
          *reinterpret_cast<float2*>(s + 2 * (base + MemberOff(m, g.tbit, K))) = v;
        });
  }
}

template <int K>
__global__ void __launch_bounds__(256)
GateMfmaTileKernel(float2* __restrict__ s, TileGeom g,
                   const float* __restrict__ frag) {
  extern __shared__ char lds_raw[];
  float2* lds = reinterpret_cast<float2*>(lds_raw);
  const float* ldsf = reinterpret_cast<const float*>(lds_raw);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 31, h = lane >> 5;
  FragF32<K> af;
  af.Load(frag);
  const unsigned T = 1u << g.tb;
  const unsigned nbatch = (T >> K) >> 5;
  constexpr int PER = 8192 / 256;  // tiles of up to 2^13 amplitudes
  for (uint64_t tile = blockIdx.x; tile < g.ntiles; tile += gridDim.x) {
    const uint64_t base =
        InsertZeros(tile << g.Lc, g.outer, g.nouter) | g.cbits_high;
    TileLoad<PER>(s, base, g, lds);
    __syncthreads();
    for (unsigned bt = wave; bt < nbatch; bt += kWaves) {
      unsigned p0 = bt * 32 + j;
#pragma unroll
      for (int b = 0; b < K; ++b) {
        const unsigned p = g.tpos[b];
        p0 = ((p0 >> p) << (p + 1)) | (p0 & ((1u << p) - 1));
      }
      const bool active = (p0 & g.low_cmask) == g.low_cvals;
      auto toff = [&](int m) {
        unsigned o = 0;
#pragma unroll
        for (int b = 0; b < K; ++b) o |= ((unsigned(m) >> b) & 1u) << g.tpos[b];
        return o;
      };
      if constexpr (K >= 6) af.Load(frag);
      BatchF32<K>(
          af, [&](int t) { return ldsf[2 * TileSwz(p0 + toff(t)) + h]; },
          [&](int m, float2 v) {
            if (active) lds[TileSwz(p0 + toff(m))] = v;
          });
    }
    __syncthreads();
    TileStore<PER>(s, base, g, lds);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Split-BF16 MFMA: per K block of 16 reals (8 members), lane half h supplies
// members 8 kb + 4 h + 0..3 as (re, im) pairs.

template <int K, typename Load, typename Store>
__device__ __forceinline__ void BatchBf16(const uint16_t* __restrict__ frag,
                                          Load load, Store store) {
  constexpr int D = Dims<K>::D, RB = Dims<K>::RB, KB = 2 * D / 16;
  const int l = threadIdx.x & 63;
  const int h = l >> 5;
  v16f c[RB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) c[rb] = v16f{};
  const v8bf* f = reinterpret_cast<const v8bf*>(frag);
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    // One K block of operands, split once and reused by every row block.
    v8bf bh, bl;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float2 v = load(8 * kb + 4 * h + q);
      bh[2 * q] = Bf16Hi(v.x);
      bl[2 * q] = Bf16Lo(v.x);
      bh[2 * q + 1] = Bf16Hi(v.y);
      bl[2 * q + 1] = Bf16Lo(v.y);
    }
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      const v8bf ah = f[((rb * KB + kb) * 3 + 0) * 64 + l];
      const v8bf am = f[((rb * KB + kb) * 3 + 1) * 64 + l];
      const v8bf al = f[((rb * KB + kb) * 3 + 2) * 64 + l];
      c[rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, c[rb], 0, 0, 0);
      c[rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bh, c[rb], 0, 0, 0);
      c[rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, c[rb], 0, 0, 0);
      c[rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, c[rb], 0, 0, 0);
    }
    if constexpr (K >= 6) __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int m = 16 * rb + 4 * s + 2 * h + u;
        store(m, float2{c[rb][4 * s + 2 * u], c[rb][4 * s + 2 * u + 1]});
      }
}

template <int K>
__global__ void __launch_bounds__(256)
GateBf16RegKernel(float2* __restrict__ s, GateGeom g,
                  const uint16_t* __restrict__ frag) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 31;
  const uint64_t nbatch = g.ngroups >> 5;
  const uint64_t wstride = uint64_t(gridDim.x) * kWaves;
  for (uint64_t b = uint64_t(blockIdx.x) * kWaves + wave; b < nbatch;
       b += wstride) {
    const uint64_t base = InsertZeros(b * 32 + j, g.ins, g.nins) | g.cbits;
    BatchBf16<K>(
        Launder(frag), [&](int t) { return s[base + MemberOff(t, g.tbit, K)]; },
        [&](int m, float2 v) { s[base + MemberOff(m, g.tbit, K)] = v; });
  }
}

template <int K>
__global__ void __launch_bounds__(256)
GateBf16TileKernel(float2* __restrict__ s, TileGeom g,
                   const uint16_t* __restrict__ frag) {
  extern __shared__ char lds_raw[];
  float2* lds = reinterpret_cast<float2*>(lds_raw);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 31;
  const unsigned T = 1u << g.tb;
  const unsigned nbatch = (T >> K) >> 5;
  constexpr int PER = 8192 / 256;  // tiles of up to 2^13 amplitudes
  for (uint64_t tile = blockIdx.x; tile < g.ntiles; tile += gridDim.x) {
    const uint64_t base =
        InsertZeros(tile << g.Lc, g.outer, g.nouter) | g.cbits_high;
    TileLoad<PER>(s, base, g, lds);
    __syncthreads();
    for (unsigned bt = wave; bt < nbatch; bt += kWaves) {
      unsigned p0 = bt * 32 + j;
#pragma unroll
      for (int b = 0; b < K; ++b) {
        const unsigned p = g.tpos[b];
        p0 = ((p0 >> p) << (p + 1)) | (p0 & ((1u << p) - 1));
      }
      const bool active = (p0 & g.low_cmask) == g.low_cvals;
      auto toff = [&](int m) {
        unsigned o = 0;
#pragma unroll
        for (int b = 0; b < K; ++b) o |= ((unsigned(m) >> b) & 1u) << g.tpos[b];
        return o;
      };
      BatchBf16<K>(
          Launder(frag), [&](int t) { return lds[TileSwz(p0 + toff(t))]; },
          [&](int m, float2 v) {
            if (active) lds[TileSwz(p0 + toff(m))] = v;
          });
    }
    __syncthreads();
    TileStore<PER>(s, base, g, lds);
    __syncthreads();
  }
}

// ---- host side -------------------------------------------------------------

uint64_t DepositBits(uint64_t v, uint64_t mask) {
  uint64_t r = 0;
  for (unsigned i = 0; i < 64; ++i) {
    if ((mask >> i) & 1) {
      r |= (v & 1) << i;
      v >>= 1;
    }
  }
  return r;
}

template <int K>
int DispatchMfma(float2* s, unsigned n, const unsigned* qs, const double* m64,
                 bool bf16) {
  constexpr int D = 1 << K;
  // Fragments go through the matrix staging ring (64 KiB slots).
  size_t fbytes;
  std::vector<float> f32;
  std::vector<uint16_t> fb;
  if (!bf16) {
    f32.resize(size_t(Dims<K>::RB) * D * 64);
    BuildFragF32(K, m64, f32.data());
    fbytes = f32.size() * sizeof(float);
  } else {
    fb.resize(size_t(Dims<K>::RB) * (2 * D / 16) * 3 * 512);
    BuildFragBf16(K, m64, fb.data());
    fbytes = fb.size() * sizeof(uint16_t);
  }
  void* dfrag = MatrixBuffer(fbytes);
  if (dfrag == nullptr) return QSIM_HIP_ERR_OOM;
  QHIP_TRY(StageMatrix(dfrag, bf16 ? (const void*)fb.data()
                                   : (const void*)f32.data(),
                       fbytes));
  // k = 5, 6 (exact FP32 or split-BF16): each wave keeps its row block of
  // the matrix in registers (no per-batch fragment reloads through the
  // cache).
  if (K >= 5) {
    const int rc =
        bf16 ? LaunchRowsBf16(s, n, qs, unsigned(K),
                              static_cast<const uint16_t*>(dfrag))
             : LaunchRowsF32(s, n, qs, unsigned(K),
                             static_cast<const float*>(dfrag));
    if (rc != QSIM_HIP_ERR_UNSUPPORTED) return rc;
  }

  if (qs[0] >= 5) {
    GateGeom g{};
    g.nins = K;
    for (int b = 0; b < K; ++b) {
      g.ins[b] = qs[b];
      g.tbit[b] = uint64_t{1} << qs[b];
    }
    g.ngroups = uint64_t{1} << (n - K);
    g.cbits = 0;
    // Every wave batch is 32 consecutive groups (no target below qubit 5).
    uint64_t nbatch = g.ngroups >> 5;
    // About one batch per wave: equal short waves leave no partial last
    // round of resident waves.
    uint64_t blocks = std::min<uint64_t>((nbatch + kWaves - 1) / kWaves,
                                         uint64_t{1} << 20);
    blocks = std::max<uint64_t>(blocks, 1);
    PassTimerBegin();
    if (!bf16)
      hipLaunchKernelGGL(GateMfmaRegKernel<K>, dim3(unsigned(blocks)),
                         dim3(256), 0, Stream(), reinterpret_cast<float*>(s), g,
                         static_cast<const float*>(dfrag));
    else
      hipLaunchKernelGGL(GateBf16RegKernel<K>, dim3(unsigned(blocks)),
                         dim3(256), 0, Stream(), s, g,
                         static_cast<const uint16_t*>(dfrag));
    PassTimerEnd();
    QHIP_LAUNCHED("GateMfmaRegKernel");
    return QSIM_HIP_OK;
  }

  // Tile path: 2^12 amplitudes (32 KiB) or 2^13 (64 KiB).  The tile is 2^h
  // contiguous runs of 2^Lc amplitudes (h = targets at or above Lc); 13 bits
  // are used for k = 6, and for k <= 5 whenever they leave fewer runs.
  TileGeom g{};
  auto split = [&](unsigned tb, unsigned* hout, unsigned* lcout) {
    for (unsigned hh = 0; hh <= unsigned(K); ++hh) {
      unsigned lc = tb - hh, cnt = 0;
      for (int b = 0; b < K; ++b) cnt += qs[b] >= lc;
      if (cnt == hh) {
        *hout = hh;
        *lcout = lc;
        return;
      }
    }
    *hout = 0;
    *lcout = tb;
  };
  unsigned TB = std::min(12u, n), h = 0, Lc = TB;
  split(TB, &h, &Lc);
  if (n >= 13) {
    unsigned h13, lc13;
    split(13, &h13, &lc13);
    if (K >= 6 || h13 < h) {
      TB = 13;
      h = h13;
      Lc = lc13;
    }
  }
  g.Lc = Lc;
  g.h = h;
  g.tb = TB;
  unsigned hi = 0;
  for (int b = 0; b < K; ++b) {
    if (qs[b] >= Lc) {
      g.hbit[hi] = uint64_t{1} << qs[b];
      g.outer[hi] = qs[b];
      g.tpos[b] = Lc + hi;
      ++hi;
    } else {
      g.tpos[b] = qs[b];
    }
  }
  g.nouter = hi;
  g.cbits_high = 0;
  g.low_cmask = 0;
  g.low_cvals = 0;
  bool ok = hi == h && g.tb == g.Lc + g.h && g.tb <= n &&
            ((1u << g.tb) >> K) >= 32 && ((1u << g.tb) >> K) % 32 == 0;
  for (int b = 0; b < K; ++b) ok = ok && g.tpos[b] < g.tb;
  for (int b = 1; b < K; ++b) ok = ok && g.tpos[b] > g.tpos[b - 1];
  if (!ok) return QSIM_HIP_ERR_UNSUPPORTED;  // the vector kernels take it
  g.ntiles = uint64_t{1} << (n - g.tb);
  size_t lds = (size_t{1} << g.tb) * sizeof(float2);
  uint64_t blocks = std::min<uint64_t>(g.ntiles, uint64_t{1} << 20);
  PassTimerBegin();
  if (!bf16)
    hipLaunchKernelGGL(GateMfmaTileKernel<K>, dim3(unsigned(blocks)), dim3(256),
                       lds, Stream(), s, g, static_cast<const float*>(dfrag));
  else
    hipLaunchKernelGGL(GateBf16TileKernel<K>, dim3(unsigned(blocks)), dim3(256),
                       lds, Stream(), s, g,
                       static_cast<const uint16_t*>(dfrag));
  PassTimerEnd();
  QHIP_LAUNCHED("GateMfmaTileKernel");
  return QSIM_HIP_OK;
}

}  // namespace

int LaunchGateMfma(float2* s, unsigned n, const unsigned* qs, unsigned k,
                   const double* m64, int method) {
  // Enough groups for whole 32-group batches, and a 4..6-qubit gate.
  if (k < 4 || k > 6 || n < k + 5) return QSIM_HIP_ERR_UNSUPPORTED;
  for (unsigned i = 0; i < k; ++i) {
    if (qs[i] >= n || (i > 0 && qs[i] <= qs[i - 1])) {
      SetError("target qubits must be ascending, distinct and < %u", n);
      return QSIM_HIP_ERR_ARG;
    }
  }
  if (method == QSIM_HIP_METHOD_MFMA_FP16X2 && k >= 5) {
    const int rc = LaunchGateF16(s, n, qs, k, m64);
    if (rc != QSIM_HIP_ERR_UNSUPPORTED) return rc;
  }
  const bool bf16 = method == QSIM_HIP_METHOD_MFMA_BF16X3 && k >= 5;
  switch (k) {
    case 4: return DispatchMfma<4>(s, n, qs, m64, false);
    case 5: return DispatchMfma<5>(s, n, qs, m64, bf16);
    default: return DispatchMfma<6>(s, n, qs, m64, bf16);
  }
}

}  // namespace qhip
