// f16_kernels.hip -- k = 5, 6 gate passes on FP16 MFMA with scaling.
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
//  - a target below qubit 5 (GateF16TileKernel): a tile of 2^12 or 2^13
//    amplitudes (the Lc lowest qubits times the h targets above them, as in
//    the FP32 tile kernels) is staged in LDS, its batches are computed in
//    place there, and the tile is written back.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"
#include "tile.h"

namespace qhip {

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

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

// Largest |re| or |im| of this lane's and its partner lane's (l ^ 32)
// members: the group's; then the group's scale exponent (largest value to
// [2^5, 2^6); 0 for an all-zero or non-finite group).
template <int N>
__device__ __forceinline__ int GroupScale(const float2 (&v)[N]) {
  float m = 0.0f;
#pragma unroll
  for (int i = 0; i < N; ++i) m = fmaxf(m, fmaxf(fabsf(v[i].x), fabsf(v[i].y)));
  m = fmaxf(m, __shfl_xor(m, 32));
  if (!(m > 0.0f) || !(m <= 3.4e38f)) return 0;
  int e;
  (void)frexpf(m, &e);  // m in [2^(e-1), 2^e)
  return 6 - e;
}

// The scaled inputs of K block kb as fp16 (high, low) B operands: lane half
// h supplies members 8 kb + 4 h + q, q = 0..3, as (re, im) pairs.
template <int N>
__device__ __forceinline__ void SplitB(const float2 (&v)[N], int kb, int e,
                                       v8h& bh, v8h& bl) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float re = ldexpf(v[4 * kb + q].x, e);
    const float im = ldexpf(v[4 * kb + q].y, e);
    const _Float16 hr = (_Float16)re, hi = (_Float16)im;
    bh[2 * q] = hr;
    bh[2 * q + 1] = hi;
    bl[2 * q] = (_Float16)(re - (float)hr);
    bl[2 * q + 1] = (_Float16)(im - (float)hi);
  }
}

// This wave's 32 output rows of one batch: rows 32 rb + r; lane l receives
// the (re, im) pairs of members 16 rb + 4 s + 2 (l/32) + u (s < 4, u < 2)
// of its group l % 32, handed to store(member, value) scaled by 2^-us.
template <int K, typename Store>
__device__ __forceinline__ void RowBlock(const F16Frag<K>& a,
                                         const float2 (&v)[F16Dims<K>::MPL],
                                         int rb, int e, int us, Store store) {
  const int h = (threadIdx.x >> 5) & 1;
  v16f c = v16f{};
#pragma unroll
  for (int kb = 0; kb < F16Dims<K>::KB; ++kb) {
    v8h bh, bl;
    SplitB(v, kb, e, bh, bl);
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.f[kb][1], bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.f[kb][0], bl, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.f[kb][0], bh, c, 0, 0, 0);
  }
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
    // This wave's batch: groups bsub * 32 + j of the set.
    const float2* b = buf[cur];
    float2 v[Dm::MPL];
#pragma unroll
    for (int kb = 0; kb < Dm::KB; ++kb)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        v[4 * kb + q] = b[(8 * kb + 4 * h + q) * G + bsub * 32 + j];
    const int e = GroupScale(v);
    const uint64_t base = InsertZeros(set * G + bsub * 32 + j, g.ins, g.nins);
    RowBlock<K>(a, v, rb, e, e + sm, [&](int m, float2 x) {
      s[base + MemberOffset(m, g.tbit, K)] = x;
    });
    if (nxt < nsets) put(pre, buf[cur ^ 1]);
    __syncthreads();
    cur ^= 1;
  }
}

// A target below qubit 5: LDS tiles of 2^tb amplitudes (tb = 12 or 13).
template <int K>
__global__ void __launch_bounds__(256)
GateF16TileKernel(float2* __restrict__ s, TileGeom g,
                  const uint4* __restrict__ frag, int sm) {
  using Dm = F16Dims<K>;
  extern __shared__ char lds_raw[];
  float2* lds = reinterpret_cast<float2*>(lds_raw);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int rb = wave % Dm::RB, bsub = wave / Dm::RB;
  const int j = lane & 31, h = lane >> 5;
  F16Frag<K> a;
  a.Load(frag, rb);
  const unsigned nbatch = ((1u << g.tb) >> K) >> 5;
  constexpr int PER = 8192 / 256;  // tiles of up to 2^13 amplitudes
  auto toff = [&](int m) {
    unsigned o = 0;
#pragma unroll
    for (int b = 0; b < K; ++b) o |= ((unsigned(m) >> b) & 1u) << g.tpos[b];
    return o;
  };
  for (uint64_t tile = blockIdx.x; tile < g.ntiles; tile += gridDim.x) {
    const uint64_t base =
        InsertZeros(tile << g.Lc, g.outer, g.nouter) | g.cbits_high;
    TileLoad<PER>(s, base, g, lds);
    __syncthreads();
    for (unsigned bt0 = 0; bt0 < nbatch; bt0 += Dm::BPI) {
      unsigned p0 = (bt0 + bsub) * 32 + j;
#pragma unroll
      for (int b = 0; b < K; ++b) {
        const unsigned p = g.tpos[b];
        p0 = ((p0 >> p) << (p + 1)) | (p0 & ((1u << p) - 1));
      }
      float2 v[Dm::MPL];
#pragma unroll
      for (int kb = 0; kb < Dm::KB; ++kb)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          v[4 * kb + q] = lds[TileSwz(p0 + toff(8 * kb + 4 * h + q))];
      // Every wave has read this step's batches before any is overwritten.
      __syncthreads();
      const int e = GroupScale(v);
      RowBlock<K>(a, v, rb, e, e + sm, [&](int m, float2 x) {
        lds[TileSwz(p0 + toff(m))] = x;
      });
    }
    __syncthreads();
    TileStore<PER>(s, base, g, lds);
    __syncthreads();
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

  // Tile of 2^12 amplitudes, or 2^13 where that leaves fewer memory runs
  // (and always for k = 6): 2^h runs of 2^Lc, h = targets at or above Lc.
  auto split = [&](unsigned tb, unsigned* hout, unsigned* lcout) {
    for (unsigned hh = 0; hh <= unsigned(K); ++hh) {
      const unsigned lc = tb - hh;
      unsigned cnt = 0;
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
  if (n < 12) return QSIM_HIP_ERR_UNSUPPORTED;
  unsigned TB = 12, h = 0, Lc = 12;
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
  TileGeom g{};
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
  const unsigned groups = (1u << g.tb) >> K;
  bool ok = hi == h && g.tb == g.Lc + g.h && g.tb <= n &&
            groups % unsigned(Dm::G) == 0;
  for (int b = 0; b < K; ++b) ok = ok && g.tpos[b] < g.tb;
  for (int b = 1; b < K; ++b) ok = ok && g.tpos[b] > g.tpos[b - 1];
  if (!ok) return QSIM_HIP_ERR_UNSUPPORTED;
  g.ntiles = uint64_t{1} << (n - g.tb);
  const size_t lds = (size_t{1} << g.tb) * sizeof(float2);
  const uint64_t blocks = std::min<uint64_t>(g.ntiles, uint64_t{1} << 20);
  PassTimerBegin();
  hipLaunchKernelGGL(GateF16TileKernel<K>, dim3(unsigned(blocks)), dim3(256),
                     lds, Stream(), s, g, f4, sm);
  PassTimerEnd();
  QHIP_LAUNCHED("GateF16TileKernel");
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

}  // namespace qhip
