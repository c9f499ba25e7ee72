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
// the 64 KiB tile (plus a 12 KiB fragment ring) lets two workgroups share a
// CU's 160 KiB of LDS, so one workgroup's HBM phase overlaps the other's
// MFMA phase.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "internal.h"
#include "tile.h"

namespace qhip {

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;
constexpr size_t kTileBytes = (size_t{1} << kBlockMaxBits) * sizeof(float2);
// BF16 fragment elements of one 4-qubit op: 2 K blocks x (hi, mid, lo) x 512.
constexpr size_t kOpFrag = 2 * 3 * 512;
constexpr int kOpWords = int(kOpFrag / 2);       // 32-bit words per op (6 KiB)
constexpr int kRingWords = kOpWords / kThreads;  // per thread per op
static_assert(kOpWords % kThreads == 0, "fragment ring split");
constexpr size_t kBlockLds = kTileBytes + 2 * kOpWords * sizeof(uint32_t);

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
  // TileSwz is linear over GF(2) and a member's offset has no bit in common
  // with its group base, so slot(base + off) = TileSwz(base) ^ TileSwz(off):
  // per member one XOR with a per-op constant (the lane half's part of the
  // member index folded in once).
  const unsigned sh = TileSwz(toff(4u * unsigned(h)));
  unsigned rd[8], wr[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    rd[q] = sh ^ TileSwz(toff(8u * unsigned(q >> 2) + unsigned(q & 3)));
  }
  const unsigned shw = TileSwz(toff(2u * unsigned(h)));
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int u = 0; u < 2; ++u) wr[2 * q + u] = shw ^ TileSwz(toff(4u * q + u));
  for (unsigned bt = wave; bt < nbatch; bt += nwaves) {
    unsigned p0 = bt * 32 + j;
    p0 = ((p0 >> t0) << (t0 + 1)) | (p0 & ((1u << t0) - 1));
    p0 = ((p0 >> t1) << (t1 + 1)) | (p0 & ((1u << t1) - 1));
    p0 = ((p0 >> t2) << (t2 + 1)) | (p0 & ((1u << t2) - 1));
    p0 = ((p0 >> t3) << (t3 + 1)) | (p0 & ((1u << t3) - 1));
    const unsigned sp = TileSwz(p0);
    // Lane half h supplies members 8 kb + 4 h + 0..3 of group j, split into
    // bf16 (hi, lo) pairs two values at a time.
    v8bf bh[2], bl[2];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float2 v = lds[sp ^ rd[q]];
      const v2f x = {v.x, v.y};
      const v2bf hi = __builtin_convertvector(x, v2bf);
      const v2f r = x - __builtin_convertvector(hi, v2f);
      const v2bf lo = __builtin_convertvector(r, v2bf);
      bh[q >> 2][2 * (q & 3)] = hi[0];
      bh[q >> 2][2 * (q & 3) + 1] = hi[1];
      bl[q >> 2][2 * (q & 3)] = lo[0];
      bl[q >> 2][2 * (q & 3) + 1] = lo[1];
    }
    // One accumulator: the vector ALU, not MFMA latency, bounds this loop
    // (other waves cover the dependent chain), so no extra adds.
    v16f c = v16f{};
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[2], bh[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[1], bh[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[0], bl[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[0], bh[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[5], bh[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[4], bh[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[3], bl[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[3], bh[1], c, 0, 0, 0);
    // Lane l holds (re, im) of members 4 s + 2 h + {0, 1}.
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        lds[sp ^ wr[2 * q + u]] = float2{c[4 * q + 2 * u], c[4 * q + 2 * u + 1]};
      }
  }
}

// Offset (in amplitudes) of tile-local high index x: x's bits deposited at
// the tile's high qubits.
__device__ __forceinline__ uint64_t DepositHigh(unsigned x, const BlockGeom& g) {
  uint64_t o = 0;
  for (unsigned b = 0; b < g.h; ++b) o += ((x >> b) & 1u) ? g.hbit[b] : 0;
  return o;
}

// Tile staging with addresses split into a per-thread part and a per-element
// part.  Element i of thread t is tile-local amplitude a = t + NT i.  The
// host keeps the low run within the thread bits (Lc <= log2 NT = 9), so a's
// low bits come from t alone and its high index is
// (t >> Lc) | (i << (9 - Lc)) with disjoint bits: its offset is
// D(t >> Lc) + D(i << (9 - Lc)), one add per element instead of a deposit
// loop.  The element parts are wave-uniform and kept in units of 2^Lc
// amplitudes (32 bits: no tile qubit is 32 above the low run).
template <int PER, int NT>
struct TileIo {
  static_assert(NT == 512, "thread bits");
  static constexpr unsigned kLogNT = 9;
  uint64_t tofs;       // this thread's part
  uint32_t eofs[PER];  // the element parts >> Lc

  __device__ __forceinline__ void Init(const BlockGeom& g) {
    const unsigned t = threadIdx.x;
    tofs = (t & ((1u << g.Lc) - 1)) + DepositHigh(t >> g.Lc, g);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      eofs[i] = uint32_t(DepositHigh(unsigned(i) << (kLogNT - g.Lc), g) >> g.Lc);
    }
  }

  template <int CH>
  __device__ __forceinline__ void Load(const float2* __restrict__ s,
                                       uint64_t base, const BlockGeom& g,
                                       float2* lds) const {
    const unsigned T = 1u << g.tb;
    const float2* p = s + base + tofs;
#pragma unroll
    for (int i0 = 0; i0 < PER; i0 += CH) {
      float2 v[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const unsigned a = threadIdx.x + NT * (i0 + i);
        if (a < T) v[i] = p[uint64_t(eofs[i0 + i]) << g.Lc];
      }
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const unsigned a = threadIdx.x + NT * (i0 + i);
        if (a < T) lds[TileSwz(a)] = v[i];
      }
    }
  }

  template <int CH>
  __device__ __forceinline__ void Store(float2* __restrict__ s, uint64_t base,
                                        const BlockGeom& g,
                                        const float2* lds) const {
    const unsigned T = 1u << g.tb;
    float2* p = s + base + tofs;
#pragma unroll
    for (int i0 = 0; i0 < PER; i0 += CH) {
      float2 v[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const unsigned a = threadIdx.x + NT * (i0 + i);
        if (a < T) v[i] = lds[TileSwz(a)];
      }
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const unsigned a = threadIdx.x + NT * (i0 + i);
        if (a < T) p[uint64_t(eofs[i0 + i]) << g.Lc] = v[i];
      }
    }
  }
};

// Two 512-thread workgroups per CU, each with a 64 KiB tile and a two-slot
// ring of op fragments in LDS: while op i runs from slot i % 2, each thread
// holds its 12 bytes of op i + 1's fragments in flight from global memory
// (L2) and writes them to the other slot before the op's barrier, so the
// fragment latency hides behind the op's MFMA work.  The program repeats for
// every tile, so the ring wraps from the last op to the first.
__global__ void __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(4)))
GateBlockKernel(float2* __restrict__ s, BlockGeom g, BlockProgram p,
                const uint16_t* __restrict__ frag) {
  extern __shared__ char lds_raw[];
  float2* lds = reinterpret_cast<float2*>(lds_raw);
  uint32_t* ring = reinterpret_cast<uint32_t*>(lds_raw + kTileBytes);
  const int l = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned nbatch = (1u << g.tb) >> 9;  // 32 groups of 16 per batch
  constexpr int PER = (1 << kBlockMaxBits) / kThreads;
  const uint32_t* gfrag = reinterpret_cast<const uint32_t*>(frag);
  // Op 0's fragments into slot 0.
#pragma unroll
  for (int w = 0; w < kRingWords; ++w) {
    ring[threadIdx.x + kThreads * w] = gfrag[threadIdx.x + kThreads * w];
  }
  TileIo<PER, kThreads> io;
  io.Init(g);
  unsigned k = 0;  // ops run so far: the current op's slot is k % 2
  for (uint64_t tile = blockIdx.x; tile < g.ntiles; tile += gridDim.x) {
    const uint64_t base = InsertZeros(tile << g.Lc, g.outer, g.nouter);
    io.Load<16>(s, base, g, lds);
    __syncthreads();
    for (unsigned op = 0; op < p.nops; ++op, ++k) {
      const unsigned nxt = op + 1 == p.nops ? 0 : op + 1;
      const uint32_t* gn = gfrag + size_t(nxt) * kOpWords;
      uint32_t pre[kRingWords];
#pragma unroll
      for (int w = 0; w < kRingWords; ++w) pre[w] = gn[threadIdx.x + kThreads * w];
      const v8bf* fp =
          reinterpret_cast<const v8bf*>(ring + (k & 1u) * kOpWords);
      v8bf f[6];
#pragma unroll
      for (int r = 0; r < 6; ++r) f[r] = fp[r * 64 + l];
      ApplyTileOp(lds, p.tpos[op], f, nbatch, wave, kWaves);
      uint32_t* slot = ring + ((k + 1) & 1u) * kOpWords;
#pragma unroll
      for (int w = 0; w < kRingWords; ++w) slot[threadIdx.x + kThreads * w] = pre[w];
      __syncthreads();
    }
    io.Store<16>(s, base, g, lds);
    __syncthreads();
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
  // The kernel's staging keeps the contiguous run within its 9 thread bits;
  // further low qubits are handled as tile qubits above the run (the same
  // tile-local positions and addresses).
  if (g.Lc > 9) {
    g.Lc = 9;
    g.h = nt - 9;
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
  // More than 64 KiB of dynamic LDS must be allowed once per process.
  static std::once_flag once;
  static hipError_t attr = hipSuccess;
  std::call_once(once, [] {
    attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&GateBlockKernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(kBlockLds));
  });
  QHIP_CHECK(attr);
  // Two workgroups per CU fit in LDS (2 x 76 KiB); a few tiles each.
  const uint64_t blocks =
      std::min<uint64_t>(g.ntiles, uint64_t(NumCUs()) * 2 * 8);
  PassTimerBegin();
  hipLaunchKernelGGL(GateBlockKernel, dim3(unsigned(blocks)), dim3(kThreads),
                     kBlockLds, Stream(), s, g, p, dfrag);
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
