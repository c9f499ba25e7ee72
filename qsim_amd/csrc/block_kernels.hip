// block_kernels.hip -- cache-blocked gate passes on gfx950.
//
// A plain gate pass streams the whole state through HBM once per fused gate
// (16 B per amplitude).  A blocked pass streams it once for a whole run of
// gates: the state is cut into tiles of 2^tb amplitudes (tb <= 13, 64 KiB of
// complex64) spanned by the Lc lowest qubits (contiguous runs in memory) and
// h = tb - Lc chosen higher qubits; each workgroup stages a tile in LDS,
// applies every gate of the program to it there, and writes it back.  Any
// gate sequence whose qubits all lie in the tile set is exact under this
// reordering: the tiles are independent subspaces of the gates' action.
//
// Each program op is a 4-qubit gate on tile-local positions (smaller gates
// are padded with identity factors on the host) applied as the split-BF16
// MFMA product of the gate kernels (mfma_kernels.hip): the 16 x 16 complex
// matrix becomes a real 32 x 32 A operand, 32 groups side by side the B
// operand of v_mfma_f32_32x32x16_bf16; the matrix is split in three bf16
// parts, the amplitudes in two, and Al Bh + Am Bh + Ah Bl + Ah Bh is
// accumulated in FP32.  Operand maps (verified on
// gfx950): a[e] = A[l%32][8(l/32) + e], b[e] = B[8(l/32) + e][l%32],
// C register r of lane l = C[8(r/4) + 4(l/32) + r%4][l%32].
//
// The grid is a tile-stride loop; 512 threads (8 waves) share a tile, and
// the 64 KiB tile lets two workgroups share a CU's 160 KiB of LDS, so one
// workgroup's HBM phase overlaps the other's MFMA phase.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.h"
#include "tile.h"

namespace qhip {

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;
constexpr int kPipeThreads = 512;
constexpr int kPipeWaves = kPipeThreads / 64;
constexpr unsigned kPipeMaxOps = 16;  // 16 x 6 KiB of fragments beside the tile
constexpr size_t kTileBytes = (size_t{1} << kBlockMaxBits) * sizeof(float2);
// BF16 fragment elements of one 4-qubit op: 2 K blocks x (hi, mid, lo) x 512.
constexpr size_t kOpFrag = 2 * 3 * 512;

__device__ __forceinline__ __bf16 Hi(float x) { return (__bf16)x; }
__device__ __forceinline__ __bf16 Lo(float x) {
  return (__bf16)(x - (float)(__bf16)x);
}

// One 4-qubit op over the LDS tile: wave `wave` of `nwaves` takes batches
// of 32 groups; `f` holds the op's fragments (hi, mid, lo per K block).
__device__ __forceinline__ void ApplyTileOp(float2* lds, const uint8_t* tp,
                                            const v8bf (&f)[6], unsigned nbatch,
                                            int wave, int nwaves) {
  const int l = threadIdx.x & 63;
  const int j = l & 31, h = l >> 5;
  const unsigned t0 = tp[0], t1 = tp[1], t2 = tp[2], t3 = tp[3];
  auto toff = [&](unsigned m) {
    return ((m & 1u) << t0) | (((m >> 1) & 1u) << t1) |
           (((m >> 2) & 1u) << t2) | (((m >> 3) & 1u) << t3);
  };
  for (unsigned bt = wave; bt < nbatch; bt += nwaves) {
    unsigned p0 = bt * 32 + j;
    p0 = ((p0 >> t0) << (t0 + 1)) | (p0 & ((1u << t0) - 1));
    p0 = ((p0 >> t1) << (t1 + 1)) | (p0 & ((1u << t1) - 1));
    p0 = ((p0 >> t2) << (t2 + 1)) | (p0 & ((1u << t2) - 1));
    p0 = ((p0 >> t3) << (t3 + 1)) | (p0 & ((1u << t3) - 1));
    // Lane half h supplies members 8 kb + 4 h + 0..3 of group j.
    float2 v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const unsigned m = 8 * (q >> 2) + 4 * h + (q & 3);
      v[q] = lds[TileSwz(p0 + toff(m))];
    }
    v8bf bh0, bl0, bh1, bl1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      bh0[2 * q] = Hi(v[q].x);
      bl0[2 * q] = Lo(v[q].x);
      bh0[2 * q + 1] = Hi(v[q].y);
      bl0[2 * q + 1] = Lo(v[q].y);
      bh1[2 * q] = Hi(v[4 + q].x);
      bl1[2 * q] = Lo(v[4 + q].x);
      bh1[2 * q + 1] = Hi(v[4 + q].y);
      bl1[2 * q + 1] = Lo(v[4 + q].y);
    }
    // Two accumulators (one per K block) halve the dependent MFMA chain.
    v16f c = v16f{}, d = v16f{};
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[2], bh0, c, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[5], bh1, d, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[1], bh0, c, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[4], bh1, d, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[0], bl0, c, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[3], bl1, d, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[0], bh0, c, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[3], bh1, d, 0, 0, 0);
    c += d;
    // Lane l holds (re, im) of members 4 s + 2 h + {0, 1}.
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const unsigned m = 4 * q + 2 * h + u;
        lds[TileSwz(p0 + toff(m))] =
            float2{c[4 * q + 2 * u], c[4 * q + 2 * u + 1]};
      }
  }
}

// Programs of more than kPipeMaxOps gates: fragments read from global memory
// per op, two 512-thread workgroups per CU.
__global__ void __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(4)))
GateBlockKernel(float2* __restrict__ s, BlockGeom g, BlockProgram p,
                const uint16_t* __restrict__ frag) {
  extern __shared__ char lds_raw[];
  float2* lds = reinterpret_cast<float2*>(lds_raw);
  const int l = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned nbatch = (1u << g.tb) >> 9;  // 32 groups of 16 per batch
  constexpr int PER = (1 << kBlockMaxBits) / kThreads;
  for (uint64_t tile = blockIdx.x; tile < g.ntiles; tile += gridDim.x) {
    const uint64_t base = InsertZeros(tile << g.Lc, g.outer, g.nouter);
    TileLoad<PER, kThreads, 8>(s, base, g, lds);
    __syncthreads();
    for (unsigned op = 0; op < p.nops; ++op) {
      const v8bf* fp = reinterpret_cast<const v8bf*>(frag + op * kOpFrag);
      v8bf f[6];
#pragma unroll
      for (int r = 0; r < 6; ++r) f[r] = fp[r * 64 + l];
      ApplyTileOp(lds, p.tpos[op], f, nbatch, wave, kWaves);
      __syncthreads();
    }
    TileStore<PER, kThreads, 8>(s, base, g, lds);
    __syncthreads();
  }
}

// Programs of up to kPipeMaxOps gates: one 1024-thread workgroup per CU
// keeps the program's fragments in LDS beside the tile and software-
// pipelines the tiles: the next tile's loads are issued into registers
// before the current tile's gates run, so HBM latency hides behind the
// MFMA work.  (No global loads happen between the prefetch and its use, so
// the in-order vector-memory counter never makes a gate wait for it.)
__global__ void __launch_bounds__(kPipeThreads)
__attribute__((amdgpu_waves_per_eu(2)))
GateBlockPipeKernel(float2* __restrict__ s, BlockGeom g, BlockProgram p,
                    const uint16_t* __restrict__ frag) {
  extern __shared__ char lds_raw[];
  float2* lds = reinterpret_cast<float2*>(lds_raw);
  v8bf* lfrag = reinterpret_cast<v8bf*>(lds_raw + kTileBytes);
  const int l = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned T = 1u << g.tb;
  const unsigned nbatch = T >> 9;
  constexpr int PER = (1 << kBlockMaxBits) / kPipeThreads;

  const unsigned nfv = p.nops * unsigned(kOpFrag / 8);
  const v8bf* gf = reinterpret_cast<const v8bf*>(frag);
  for (unsigned i = threadIdx.x; i < nfv; i += kPipeThreads) lfrag[i] = gf[i];

  float2 r[PER];
  auto prefetch = [&](uint64_t t) {
    const uint64_t b = InsertZeros(t << g.Lc, g.outer, g.nouter);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const unsigned a = threadIdx.x + kPipeThreads * i;
      if (a < T) r[i] = s[TileAddr(b, a, g)];
    }
  };
  uint64_t tile = blockIdx.x;
  if (tile < g.ntiles) prefetch(tile);
  while (tile < g.ntiles) {
    const uint64_t base = InsertZeros(tile << g.Lc, g.outer, g.nouter);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const unsigned a = threadIdx.x + kPipeThreads * i;
      if (a < T) lds[TileSwz(a)] = r[i];
    }
    __syncthreads();
    const uint64_t next = tile + gridDim.x;
    if (next < g.ntiles) prefetch(next);
    for (unsigned op = 0; op < p.nops; ++op) {
      const v8bf* fp = lfrag + op * unsigned(kOpFrag / 8);
      v8bf f[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) f[k] = fp[k * 64 + l];
      ApplyTileOp(lds, p.tpos[op], f, nbatch, wave, kPipeWaves);
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const unsigned a = threadIdx.x + kPipeThreads * i;
      if (a < T) s[TileAddr(base, a, g)] = lds[TileSwz(a)];
    }
    __syncthreads();
    tile = next;
  }
}

// U (2^k x 2^k over positions `pos`, ascending) embedded into the 16 x 16
// matrix over `pad` (4 ascending positions containing `pos`), identity on
// the extra positions.
void PadTo4(const double* u, unsigned k, const unsigned* pos,
            const unsigned* pad, double* out) {
  unsigned which[4];  // bit of pad[b] in U's index, or ~0u for an extra
  for (int b = 0; b < 4; ++b) {
    which[b] = ~0u;
    for (unsigned i = 0; i < k; ++i)
      if (pos[i] == pad[b]) which[b] = i;
  }
  const unsigned D = 1u << k;
  for (unsigned r = 0; r < 16; ++r) {
    for (unsigned c = 0; c < 16; ++c) {
      unsigned ur = 0, uc = 0;
      bool same = true;
      for (int b = 0; b < 4; ++b) {
        const unsigned rb = (r >> b) & 1, cb = (c >> b) & 1;
        if (which[b] == ~0u) {
          same = same && rb == cb;
        } else {
          ur |= rb << which[b];
          uc |= cb << which[b];
        }
      }
      const size_t o = 2 * (size_t(r) * 16 + c);
      out[o] = same ? u[2 * (size_t(ur) * D + uc)] : 0.0;
      out[o + 1] = same ? u[2 * (size_t(ur) * D + uc) + 1] : 0.0;
    }
  }
}

}  // namespace

int PrepareGateBlock(unsigned n, const unsigned* tile, unsigned nt,
                     unsigned nops, const unsigned* ks, const unsigned* qs,
                     const double* m64, const size_t* moff, BlockGeom& g,
                     BlockProgram& p, std::vector<uint16_t>& fb) {
  if (nt < kBlockMinBits || nt > kBlockMaxBits || nt > n) {
    SetError("a blocked pass needs %u..%u tile qubits (at most n = %u), got %u",
             kBlockMinBits, kBlockMaxBits, n, nt);
    return QSIM_HIP_ERR_ARG;
  }
  if (nops > kBlockMaxOps) {
    SetError("a blocked pass holds at most %u gates, got %u", kBlockMaxOps,
             nops);
    return QSIM_HIP_ERR_ARG;
  }
  for (unsigned i = 0; i < nt; ++i) {
    if (tile[i] >= n || (i > 0 && tile[i] <= tile[i - 1])) {
      SetError("tile qubits must be ascending, distinct and < %u", n);
      return QSIM_HIP_ERR_ARG;
    }
  }
  g = BlockGeom{};
  g.Lc = 0;
  while (g.Lc < nt && tile[g.Lc] == g.Lc) ++g.Lc;
  g.tb = nt;
  g.h = nt - g.Lc;
  if (g.Lc < 3 || g.h > kBlockMaxHigh) {
    SetError("tile qubits must start with the run 0..%u of at least 3 qubits "
             "and add at most %u higher ones", 2u, kBlockMaxHigh);
    return QSIM_HIP_ERR_ARG;
  }
  g.nouter = g.h;
  for (unsigned b = 0; b < g.h; ++b) {
    g.outer[b] = tile[g.Lc + b];
    g.hbit[b] = uint64_t{1} << tile[g.Lc + b];
  }
  g.ntiles = uint64_t{1} << (n - nt);

  // Tile-local position of a global qubit (tile qubits are ascending, so
  // the map keeps order).
  auto local = [&](unsigned q) -> int {
    for (unsigned i = 0; i < nt; ++i)
      if (tile[i] == q) return int(i);
    return -1;
  };

  p = BlockProgram{};
  p.nops = nops;
  fb.assign(std::max<size_t>(nops, 1) * kOpFrag, 0);
  std::vector<double> u16(2 * 16 * 16);
  for (unsigned i = 0; i < nops; ++i) {
    const unsigned k = ks[i];
    if (k < 1 || k > 4) {
      SetError("blocked-pass gate %u acts on %u qubits (1..4 supported)", i, k);
      return QSIM_HIP_ERR_ARG;
    }
    unsigned pos[4], pad[4];
    for (unsigned b = 0; b < k; ++b) {
      const int lp = local(qs[4 * i + b]);
      if (lp < 0 || (b > 0 && unsigned(lp) <= pos[b - 1])) {
        SetError("blocked-pass gate %u: qubits must be ascending and inside "
                 "the tile", i);
        return QSIM_HIP_ERR_ARG;
      }
      pos[b] = unsigned(lp);
    }
    // Pad with the lowest free tile positions.
    unsigned np = 0, next = 0;
    unsigned merged[4];
    for (unsigned b = 0; b < k; ++b) merged[np++] = pos[b];
    while (np < 4) {
      bool used = false;
      for (unsigned b = 0; b < np; ++b) used = used || merged[b] == next;
      if (!used) merged[np++] = next;
      ++next;
    }
    std::sort(merged, merged + 4);
    for (int b = 0; b < 4; ++b) {
      pad[b] = merged[b];
      p.tpos[i][b] = uint8_t(merged[b]);
    }
    PadTo4(m64 + moff[i], k, pos, pad, u16.data());
    BuildFragBf16(4, u16.data(), fb.data() + i * kOpFrag);
  }

  return QSIM_HIP_OK;
}

int RunGateBlock(float2* s, const BlockGeom& g, const BlockProgram& p,
                 const uint16_t* dfrag) {
  if (p.nops <= kPipeMaxOps) {
    static bool attr = false;
    const size_t lds = kTileBytes + kPipeMaxOps * kOpFrag * sizeof(uint16_t);
    if (!attr) {
      QHIP_CHECK(hipFuncSetAttribute(
          reinterpret_cast<const void*>(&GateBlockPipeKernel),
          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
      attr = true;
    }
    // One workgroup per CU, each streaming its share of the tiles.
    const uint64_t blocks = std::min<uint64_t>(g.ntiles, uint64_t(NumCUs()));
    const size_t used = kTileBytes + p.nops * kOpFrag * sizeof(uint16_t);
    PassTimerBegin();
    hipLaunchKernelGGL(GateBlockPipeKernel, dim3(unsigned(blocks)),
                       dim3(kPipeThreads), used, Stream(), s, g, p, dfrag);
    PassTimerEnd();
    QHIP_LAUNCHED("GateBlockPipeKernel");
    return QSIM_HIP_OK;
  }
  const size_t lds = (size_t{1} << g.tb) * sizeof(float2);
  // Two workgroups per CU fit in LDS; a few tiles each.
  const uint64_t blocks =
      std::min<uint64_t>(g.ntiles, uint64_t(NumCUs()) * 2 * 8);
  PassTimerBegin();
  hipLaunchKernelGGL(GateBlockKernel, dim3(unsigned(blocks)), dim3(kThreads),
                     lds, Stream(), s, g, p, dfrag);
  PassTimerEnd();
  QHIP_LAUNCHED("GateBlockKernel");
  return QSIM_HIP_OK;
}

int LaunchGateBlock(float2* s, unsigned n, const unsigned* tile, unsigned nt,
                    unsigned nops, const unsigned* ks, const unsigned* qs,
                    const double* m64, const size_t* moff) {
  BlockGeom g{};
  BlockProgram p{};
  std::vector<uint16_t> fb;
  QHIP_TRY(PrepareGateBlock(n, tile, nt, nops, ks, qs, m64, moff, g, p, fb));
  const size_t fbytes = fb.size() * sizeof(uint16_t);
  void* dfrag = MatrixBuffer(fbytes);
  if (dfrag == nullptr) return QSIM_HIP_ERR_OOM;
  QHIP_TRY(StageMatrix(dfrag, fb.data(), fbytes));
  return RunGateBlock(s, g, p, static_cast<const uint16_t*>(dfrag));
}

// A blocked pass prepared once: geometry, program and device fragments.
struct PreparedBlock {
  unsigned n;
  BlockGeom g;
  BlockProgram p;
  uint16_t* dfrag;
};

int PrepareBlockHandle(unsigned n, const unsigned* tile, unsigned nt,
                       unsigned nops, const unsigned* ks, const unsigned* qs,
                       const double* m64, const size_t* moff, void** out) {
  auto* b = new PreparedBlock{};
  std::vector<uint16_t> fb;
  int rc = PrepareGateBlock(n, tile, nt, nops, ks, qs, m64, moff, b->g, b->p,
                            fb);
  if (rc != QSIM_HIP_OK) {
    delete b;
    return rc;
  }
  b->n = n;
  const size_t bytes = fb.size() * sizeof(uint16_t);
  hipError_t e = hipMalloc(&b->dfrag, bytes);
  if (e == hipSuccess) {
    e = hipMemcpy(b->dfrag, fb.data(), bytes, hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    (void)hipFree(b->dfrag);
    delete b;
    return HipFail(e, "prepared blocked pass");
  }
  *out = b;
  return QSIM_HIP_OK;
}

int RunBlockHandle(void* h, float2* s, unsigned n) {
  auto* b = static_cast<PreparedBlock*>(h);
  if (b->n != n) {
    SetError("prepared blocked pass is for %u qubits, state has %u", b->n, n);
    return QSIM_HIP_ERR_ARG;
  }
  return RunGateBlock(s, b->g, b->p, b->dfrag);
}

int FreeBlockHandle(void* h) {
  auto* b = static_cast<PreparedBlock*>(h);
  if (b == nullptr) return QSIM_HIP_OK;
  // Launches that read the fragments may still be in flight.
  QHIP_CHECK(hipStreamSynchronize(Stream()));
  QHIP_CHECK(hipFree(b->dfrag));
  delete b;
  return QSIM_HIP_OK;
}

}  // namespace qhip
