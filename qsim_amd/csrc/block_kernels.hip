// block_kernels.hip -- cache-blocked gate passes on gfx950.
//
// A plain gate pass streams the whole state through HBM once per fused gate
// (16 B per amplitude).  A blocked pass streams it once for a whole run of
// gates: the state is cut into tiles of 2^tb amplitudes (tb <= 14, 128 KiB of
// complex64) spanned by the Lc lowest qubits (contiguous runs in memory) and
// h = tb - Lc chosen higher qubits; each workgroup stages a tile in LDS,
// applies every gate of the program to it there, and writes it back.  Any
// gate sequence whose qubits all lie in the tile set is exact under this
// reordering: the tiles are independent subspaces of the gates' action.
//
// Each program op is a 4-qubit gate on tile-local positions, applied as an
// MFMA product (block_tile.h): FP16 with a power-of-two scale per tile
// (GateBlockF16Kernel, the default; every op record of the pass resident in
// LDS), or split-BF16 (GateBlockKernel, for programs that could grow the
// state by more than 2^10 and on request; op records through a two-slot
// ring).
//
// The grid is a tile-stride loop.  Tiles of up to 13 qubits: 256 threads
// share a tile, and the 64 KiB tile plus its op records lets two workgroups
// share a CU's 160 KiB of LDS, so one workgroup's HBM phase overlaps the
// other's MFMA phase.  14-qubit tiles: 512 threads, one workgroup per CU;
// fewer passes for a circuit, with the overlap coming from the next tile's
// prefetch alone.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "block_tile.h"
#include "internal.h"
#include "tile.h"

namespace qhip {

namespace {

using namespace tileop;

// First tile of workgroup b in a tile-stride loop of stride G.  Tiles that
// differ only in the lowest non-tile qubits hold neighbouring memory runs
// (with a short contiguous run, parts of the same cache lines and DRAM
// pages); the dispatcher places workgroup b on XCD b mod 8, so each group
// of kTileGroup such tiles goes to workgroups b, b + 8, b + 16, ... (one
// XCD, one L2) instead of b, b + 1, ...  A bijection on [0, G) when
// 8 kTileGroup divides G.
constexpr unsigned kTileGroup = 4;
__device__ __forceinline__ uint64_t PairedTile(unsigned b, unsigned G) {
  if (G % (8u * kTileGroup) != 0) return b;
  const unsigned q = b / (8u * kTileGroup), r = b % (8u * kTileGroup);
  return kTileGroup * (8u * q + (r & 7u)) + (r >> 3);
}

// Each workgroup has its tile and a two-slot ring of op records in LDS:
// while op i runs from slot i % 2, each thread holds its share (32 or 16
// bytes) of op i + 1's record in flight from global memory (L2) and writes
// it to the other slot before the op's barrier, so the record latency hides
// behind the op's MFMA work.  The program repeats for every tile, so the
// ring wraps from the last op to the first.  Both shapes (BlockShape) run
// two waves per SIMD, so each wave has 256 registers: enough to keep two
// batches of an op in flight and the next tile's amplitudes (64 registers)
// in flight from HBM while this tile's ops run.  Vector memory counters
// complete in order, so the prefetch goes out in quarters, each right after
// an op's record fetch: the record an op waits for is never behind a
// quarter issued during that op, and each quarter has at least one whole op
// to arrive (quarters 0..2 with ops 0..2, the rest with the last op).  The
// op sequence is straight-line per program length so that the compiler's
// wait counts stay exact.
// TB: the kernel's tile shape; FULL: the tile holds 2^TB amplitudes, every
// thread's 32 elements (else fewer, 2^tb with tb < TB); ACC: the accurate
// op product (ApplyTileOp).
template <int TB, bool FULL, bool ACC>
__global__ void __launch_bounds__(BlockShape<TB>::NT)
__attribute__((amdgpu_waves_per_eu(2)))
GateBlockKernel(float2* __restrict__ s, BlockGeom g, BlockProgram p,
                const uint16_t* __restrict__ frag) {
  constexpr int NT = BlockShape<TB>::NT;
  constexpr int kRingQuads = kOpWords / 4 / NT;  // uint4 per thread per op
  static_assert(kOpWords % (4 * NT) == 0, "record split");
  extern __shared__ char lds_raw[];
  float2* lds = reinterpret_cast<float2*>(lds_raw);
  uint32_t* ring =
      reinterpret_cast<uint32_t*>(lds_raw + BlockShape<TB>::kTileBytes);
  const int l = threadIdx.x & 63;
  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned nbatch = (1u << g.tb) >> 9;  // 32 groups of 16 per batch
  constexpr int PER = (1 << TB) / NT;
  static_assert(PER == 32, "prefetch quarters");
  const uint4* gfrag = reinterpret_cast<const uint4*>(frag);
  uint4* ring4 = reinterpret_cast<uint4*>(ring);
  // Op 0's record into slot 0.
#pragma unroll
  for (int w = 0; w < kRingQuads; ++w) {
    ring4[threadIdx.x + NT * w] = gfrag[threadIdx.x + NT * w];
  }
  TileIo<PER, NT> io;
  io.Init(g);
  unsigned k = 0;  // ops run so far: the current op's slot is k % 2
  // Op `op` on the tile in LDS; `extra` runs once op nxt's record loads are
  // issued.
  auto run_op = [&](unsigned op, unsigned nxt, auto&& extra) {
      const uint4* gn = gfrag + size_t(nxt) * (kOpWords / 4);
      const uint4 pre0 = gn[threadIdx.x];
      const uint4 pre1 = kRingQuads > 1 ? gn[threadIdx.x + NT] : uint4{};
      extra();
      const uint32_t* cur = ring + (k & 1u) * kOpWords;
      const v8bf* fp = reinterpret_cast<const v8bf*>(cur);
      v8bf f[6];
#pragma unroll
      for (int r = 0; r < 6; ++r) f[r] = fp[r * 64 + l];
      const uint32_t* tab = cur + kFragWords + 16 * (l >> 5);
      const uint32_t sj = cur[kFragWords + kTableWords + (l & 31)];
      uint32_t rd[8], wr[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        rd[q] = tab[q] ^ sj;
        wr[q] = tab[8 + q] ^ sj;
      }
      ApplyTileOp<ACC>(reinterpret_cast<const uint8_t*>(cur + kBposWord), f,
                       rd, wr, nbatch, wave,
                       BlockShape<TB>::kWaves);
      uint4* slot = ring4 + ((k + 1) & 1u) * (kOpWords / 4);
      slot[threadIdx.x] = pre0;
      if (kRingQuads > 1) slot[threadIdx.x + NT] = pre1;
      __syncthreads();
      ++k;
  };
  float2 buf[PER];
  uint64_t tile = PairedTile(blockIdx.x, gridDim.x);
  if (tile < g.ntiles) {
    io.template Fetch<FULL, 15u>(
        s, InsertZeros(tile << g.Lc, g.outer, g.nouter), g, buf);
    io.template Put<FULL>(g, buf, lds);
  }
  __syncthreads();
  // Each iteration leaves the next tile in LDS, so the waits on the
  // prefetch are counted within one iteration.
  for (; tile < g.ntiles; tile += gridDim.x) {
    const uint64_t base = InsertZeros(tile << g.Lc, g.outer, g.nouter);
    // The next tile of this workgroup, or this one again past the end (an
    // in-bounds read that is never used), so the prefetch is unconditional.
    const uint64_t next = tile + gridDim.x < g.ntiles ? tile + gridDim.x : tile;
    const uint64_t nbase = InsertZeros(next << g.Lc, g.outer, g.nouter);
    auto fetch = [&](auto mask) {
      return [&, mask] {
        io.template Fetch<FULL, decltype(mask)::value>(s, nbase, g, buf);
      };
    };
    using Q0 = std::integral_constant<unsigned, 1u>;
    using Q1 = std::integral_constant<unsigned, 2u>;
    using Q2 = std::integral_constant<unsigned, 4u>;
    // One op, and partial tiles (fewer elements per thread, guarded):
    // everything during the last op.
    switch (FULL ? p.nops : 1u) {
      case 1:
        for (unsigned op = 0; op + 1 < p.nops; ++op) run_op(op, op + 1, [] {});
        run_op(p.nops - 1, 0, fetch(std::integral_constant<unsigned, 15u>{}));
        break;
      case 2:
        run_op(0, 1, fetch(Q0{}));
        run_op(1, 0, fetch(std::integral_constant<unsigned, 14u>{}));
        break;
      case 3:
        run_op(0, 1, fetch(Q0{}));
        run_op(1, 2, fetch(Q1{}));
        run_op(2, 0, fetch(std::integral_constant<unsigned, 12u>{}));
        break;
      default:
        run_op(0, 1, fetch(Q0{}));
        run_op(1, 2, fetch(Q1{}));
        run_op(2, 3, fetch(Q2{}));
        for (unsigned op = 3; op + 1 < p.nops; ++op) run_op(op, op + 1, [] {});
        run_op(p.nops - 1, 0, fetch(std::integral_constant<unsigned, 8u>{}));
        break;
    }
    io.template Store<16, FULL>(s, base, g, lds);
    __syncthreads();
    io.template Put<FULL>(g, buf, lds);
    __syncthreads();
  }
}

// The FP16-with-scaling pass (block_tile.h).  Every op record of the pass
// is copied to LDS once per workgroup, so the ops issue no global loads:
// the only vector memory operations are the tile's loads and stores, and
// the wait for the next tile's loads (issued in full at the start of the
// current tile's ops) never stands behind a record fetch issued after the
// previous tile's stores.  The stores of a tile drain while the next
// tile's ops run.  Measured on MI355X, an op costs about 2 ms per 32-qubit
// pass whatever the HBM traffic: rewriting the tile in LDS (8-byte LDS
// writes run at half the read rate) bounds it, not the MFMA or VALU work.  The tile is scaled by 2^e when staged and by 2^-e when
// stored (TileScale, exact); one more barrier per tile for the maximum.
template <int TB, bool FULL>
__global__ void __launch_bounds__(BlockShape<TB>::NT)
__attribute__((amdgpu_waves_per_eu(2)))
GateBlockF16Kernel(float2* __restrict__ s, BlockGeom g, BlockProgram p,
                   const uint32_t* __restrict__ recs) {
  constexpr int NT = BlockShape<TB>::NT;
  constexpr int NW = BlockShape<TB>::kWaves;
  constexpr int PER = (1 << TB) / NT;
  static_assert(PER == 32, "tile elements per thread");
  extern __shared__ char lds_raw[];
  float2* lds = reinterpret_cast<float2*>(lds_raw);
  float* slots = reinterpret_cast<float*>(lds_raw + F16Lds<TB>::kMaxOff);
  uint32_t* rec = reinterpret_cast<uint32_t*>(lds_raw + F16Lds<TB>::kRecOff);
  const int l = threadIdx.x & 63;
  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned nbatch = (1u << g.tb) >> 9;  // 32 groups of 16 per batch
  const unsigned nper = (1u << g.tb) / NT;
  {
    const unsigned nq = p.nops * unsigned(kF16RecWords / 4);
    const uint4* src = reinterpret_cast<const uint4*>(recs);
    uint4* dst = reinterpret_cast<uint4*>(rec);
    for (unsigned q = threadIdx.x; q < nq; q += NT) dst[q] = src[q];
  }
  TileIo<PER, NT> io;
  io.Init(g);
  float2 buf[PER];
  uint64_t tile = PairedTile(blockIdx.x, gridDim.x);
  int e = 0;
  if (tile < g.ntiles) {
    io.template Fetch<FULL, 15u>(
        s, InsertZeros(tile << g.Lc, g.outer, g.nouter), g, buf);
    e = TileScale<NW>(LocalMax<PER, FULL>(buf, nper), slots, p.headroom);
    io.template PutScaled<FULL>(g, buf, lds, e);
  }
  __syncthreads();
  for (; tile < g.ntiles; tile += gridDim.x) {
    const uint64_t base = InsertZeros(tile << g.Lc, g.outer, g.nouter);
    // The next tile of this workgroup, or this one again past the end (an
    // in-bounds read that is never used), so the prefetch is unconditional.
    // Its loads go out in two halves, the second after the second op (or
    // after a single op): a whole tile's loads at once exceed what a CU keeps
    // in flight, and the waves would stall issuing them instead of running
    // the ops.
    const uint64_t next = tile + gridDim.x < g.ntiles ? tile + gridDim.x : tile;
    const uint64_t nb = InsertZeros(next << g.Lc, g.outer, g.nouter);
    io.template Fetch<FULL, 3u>(s, nb, g, buf);
    for (unsigned op = 0; op < p.nops; ++op) {
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
    if (p.nops == 1) io.template Fetch<FULL, 12u>(s, nb, g, buf);
    io.template StoreScaled<16, FULL>(s, base, g, lds, e);
    e = TileScale<NW>(LocalMax<PER, FULL>(buf, nper), slots, p.headroom);
    io.template PutScaled<FULL>(g, buf, lds, e);
    __syncthreads();
  }
}

// TileSwz on the host (the slot of tile-local amplitude a).
unsigned HostTileSwz(unsigned a) { return a ^ (((a >> 4) ^ (a >> 8)) & 15u); }

// x's low k bits placed at positions pos[0..k).
unsigned Deposit(unsigned x, const unsigned* pos, unsigned k) {
  unsigned o = 0;
  for (unsigned b = 0; b < k; ++b) o |= ((x >> b) & 1u) << pos[b];
  return o;
}

// Splits the tb - 4 non-target positions of an op (targets `pad`) into 5
// lane positions and tb - 9 batch positions (ascending; unused entries 0).
// The 32 groups of a batch are read in quarter-waves of 16 lanes and half
// waves of 32; the choice minimises first the largest number of lanes of a
// quarter-wave whose group slots share a bank (slot mod 16: 8-byte slots
// over 32 four-byte banks), then the same over the half-wave (slot mod 32),
// so that the op's LDS reads and writes are conflict-free where the
// targets allow it.
void ChooseLanePositions(const unsigned* pad, unsigned tb, unsigned* lane,
                         unsigned* batch) {
  unsigned np = 0, nonk[kBlockMaxBits];
  for (unsigned q = 0; q < tb; ++q) {
    if (q != pad[0] && q != pad[1] && q != pad[2] && q != pad[3]) nonk[np++] = q;
  }
  auto worst = [](const unsigned* sl, unsigned n, unsigned mod) {
    unsigned cnt[32] = {}, w = 0;
    for (unsigned j = 0; j < n; ++j) w = std::max(w, ++cnt[sl[j] % mod]);
    return w;
  };
  unsigned best = ~0u;
  unsigned cand[5];
  // Ordered choice: a 4-subset for lane bits 0..3 (kept ascending), then
  // one more position for lane bit 4.
  for (unsigned a = 0; a < np && best != 0; ++a)
    for (unsigned b = a + 1; b < np && best != 0; ++b)
      for (unsigned c = b + 1; c < np && best != 0; ++c)
        for (unsigned d = c + 1; d < np && best != 0; ++d)
          for (unsigned e = 0; e < np && best != 0; ++e) {
            if (e == a || e == b || e == c || e == d) continue;
            cand[0] = nonk[a];
            cand[1] = nonk[b];
            cand[2] = nonk[c];
            cand[3] = nonk[d];
            cand[4] = nonk[e];
            unsigned sl[32];
            for (unsigned j = 0; j < 32; ++j) sl[j] = HostTileSwz(Deposit(j, cand, 5));
            const unsigned score =
                ((std::max(worst(sl, 16, 16), worst(sl + 16, 16, 16)) - 1) << 8) |
                (worst(sl, 32, 32) - 1);
            if (score < best) {
              best = score;
              std::copy(cand, cand + 5, lane);
            }
          }
  unsigned nb = 0;
  for (unsigned q = 0; q < np; ++q) {
    if (std::find(lane, lane + 5, nonk[q]) == lane + 5) batch[nb++] = nonk[q];
  }
  for (; nb < kBlockMaxBits - 9; ++nb) batch[nb] = 0;
}

// ChooseLanePositions depends on the padded targets and tb alone (all below
// 16), and a batch of trajectories repeats the same few thousand
// combinations across its passes: the choice is memoised per process.
void LanePositions(const unsigned* pad, unsigned tb, unsigned* lane,
                   unsigned* batch) {
  constexpr unsigned kNb = kBlockMaxBits - 9;
  static std::mutex mu;
  static std::unordered_map<uint32_t, std::array<uint8_t, 5 + kNb>> memo;
  const uint32_t key =
      pad[0] | (pad[1] << 4) | (pad[2] << 8) | (pad[3] << 12) | (tb << 16);
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = memo.find(key);
    if (it != memo.end()) {
      for (unsigned b = 0; b < 5; ++b) lane[b] = it->second[b];
      for (unsigned b = 0; b < kNb; ++b) batch[b] = it->second[5 + b];
      return;
    }
  }
  ChooseLanePositions(pad, tb, lane, batch);
  std::array<uint8_t, 5 + kNb> v{};
  for (unsigned b = 0; b < 5; ++b) v[b] = uint8_t(lane[b]);
  for (unsigned b = 0; b < kNb; ++b) v[5 + b] = uint8_t(batch[b]);
  std::lock_guard<std::mutex> lock(mu);
  memo.emplace(key, v);
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

// Round-to-nearest-even double -> IEEE fp16 bits (overflow to infinity).
uint16_t HalfBits(double x) {
  const uint16_t sign = std::signbit(x) ? 0x8000u : 0u;
  const double a = std::fabs(x);
  if (a == 0.0) return sign;
  if (!(a < 65520.0)) return uint16_t(sign | 0x7c00u);
  int e;
  (void)std::frexp(a, &e);  // a in [2^(e-1), 2^e)
  if (e - 1 < -14) {        // subnormal: multiples of 2^-24
    return uint16_t(sign | unsigned(std::nearbyint(std::ldexp(a, 24))));
  }
  double m = std::nearbyint(std::ldexp(a, 11 - e));  // [2^10, 2^11]
  int E = e - 1;
  if (m == 2048.0) {
    m = 1024.0;
    ++E;
  }
  if (E > 15) return uint16_t(sign | 0x7c00u);
  return uint16_t(sign | unsigned((E + 15) << 10) | (unsigned(m) - 1024u));
}

double HalfValue(uint16_t b) {
  const int E = (b >> 10) & 31, m = b & 1023;
  const double v = E == 0 ? std::ldexp(double(m), -24)
                          : std::ldexp(double(1024 + m), E - 25);
  return (b & 0x8000u) ? -v : v;
}

// Upper bound of the spectral norm of a 16 x 16 complex matrix: the square
// root of the largest absolute row sum of U^dagger U (which bounds its
// largest eigenvalue); 1 up to rounding for a unitary matrix.
double NormBound16(const double* u) {
  double best = 0.0;
  for (unsigned i = 0; i < 16; ++i) {
    double row = 0.0;
    for (unsigned j = 0; j < 16; ++j) {
      double re = 0.0, im = 0.0;
      for (unsigned r = 0; r < 16; ++r) {
        const double ar = u[2 * (16 * r + i)], ai = -u[2 * (16 * r + i) + 1];
        const double br = u[2 * (16 * r + j)], bi = u[2 * (16 * r + j) + 1];
        re += ar * br - ai * bi;
        im += ar * bi + ai * br;
      }
      row += std::hypot(re, im);
    }
    best = std::max(best, row);
  }
  return std::sqrt(best);
}

// The member a lane half h reads as the q-th B value (record column order)
// is the one its q-th C pair holds.
constexpr bool F16ReadsAreWrites() {
  for (unsigned h = 0; h < 2; ++h)
    for (unsigned q = 0; q < 8; ++q)
      if (F16Col(4 * h + 8 * (q >> 2) + (q & 3)) != 2 * h + 4 * (q >> 1) + (q & 1))
        return false;
  return true;
}
static_assert(F16ReadsAreWrites(), "FP16 member order");

// FP16 op record of the padded matrix u (FP64) and the tables: record
// column p of row r holds u[r][F16Col(p)].
void BuildRecF16(const double* u, const uint32_t* tab, uint32_t* rec) {
  for (unsigned e = 0; e < 256; ++e) {
    const unsigned src = (e & ~15u) | F16Col(e & 15u);
    const uint16_t hr = HalfBits(u[2 * src]), hi = HalfBits(u[2 * src + 1]);
    const uint16_t lr = HalfBits(u[2 * src] - HalfValue(hr));
    const uint16_t li = HalfBits(u[2 * src + 1] - HalfValue(hi));
    rec[e] = uint32_t(hr) | (uint32_t(hi) << 16);
    rec[kF16MatWords + e] = uint32_t(lr) | (uint32_t(li) << 16);
  }
  std::memcpy(rec + kF16TabOff, tab,
              sizeof(uint32_t) * (kTableWords + kLaneWords));
}

}  // namespace

int PrepareGateBlock(unsigned n, const unsigned* tile, unsigned nt,
                     unsigned nops, const unsigned* ks, const unsigned* qs,
                     const double* m64, const size_t* moff, BlockGeom& g,
                     BlockProgram& p, std::vector<uint16_t>& fb, int arith) {
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
  // The kernel's staging keeps the contiguous run within its thread bits
  // (8 for the 256-thread shape, 9 for the 512-thread one); further low
  // qubits are handled as tile qubits above the run (the same tile-local
  // positions and addresses).
  const unsigned max_lc = nt == 14 ? 9 : 8;
  if (g.Lc > max_lc) {
    g.Lc = max_lc;
    g.h = nt - max_lc;
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
  p.arith = unsigned(arith);
  const size_t kOpElems =  // uint16 per op record
      2 * size_t(arith == kBlockArithF16 ? kF16RecWords : kOpWords);
  fb.assign(std::max<size_t>(nops, 1) * kOpElems, 0);
  std::vector<double> u16(2 * 16 * 16);
  double growth = 1.0, max_growth = 1.0;  // products of op norm bounds
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
    uint16_t* rec = fb.data() + i * kOpElems;
    if (arith == kBlockArithF16) {
      growth *= NormBound16(u16.data());
      max_growth = std::max(max_growth, growth);
    } else {
      BuildFragBf16(4, u16.data(), rec);
    }
    // Lane tables: member m of a group sits at tile offset toff(m) (bit b of
    // m at position pad[b]); lane half h reads members 4h + 8(q/4) + q%4 and
    // writes members 2h + 4(q/2) + q%2 (the operand maps of the MFMA
    // product).  FP16 records store the matrix columns permuted (F16Col),
    // which makes the members read those written.
    auto toff = [&](unsigned m) {
      unsigned o = 0;
      for (int b = 0; b < 4; ++b) o |= ((m >> b) & 1u) << pad[b];
      return o;
    };
    uint32_t tab[kTableWords + kLaneWords];
    for (unsigned h = 0; h < 2; ++h) {
      for (unsigned q = 0; q < 8; ++q) {
        const unsigned rm = arith == kBlockArithF16
                                ? F16Col(4 * h + 8 * (q >> 2) + (q & 3))
                                : 4 * h + 8 * (q >> 2) + (q & 3);
        tab[16 * h + q] = uint32_t(sizeof(float2)) * HostTileSwz(toff(rm));
        tab[16 * h + 8 + q] = uint32_t(sizeof(float2)) *
                              HostTileSwz(toff(2 * h + 4 * (q >> 1) + (q & 1)));
      }
    }
    // Group table: lane j's group has j's bits at the lane positions.
    unsigned lane[5], batch[kBlockMaxBits - 9];
    LanePositions(pad, nt, lane, batch);
    for (unsigned j = 0; j < 32; ++j) {
      tab[kTableWords + j] =
          uint32_t(sizeof(float2)) * HostTileSwz(Deposit(j, lane, 5));
    }
    for (unsigned b = 0; b < kBlockMaxBits - 9; ++b) p.bpos[i][b] = uint8_t(batch[b]);
    uint32_t bw[2] = {0, 0};
    std::memcpy(bw, p.bpos[i], kBlockMaxBits - 9);
    if (arith == kBlockArithF16) {
      uint32_t r32[kF16RecWords];
      BuildRecF16(u16.data(), tab, r32);
      r32[kF16BposOff] = bw[0];
      r32[kF16BposOff + 1] = bw[1];
      r32[kF16BposOff + 2] = 0;
      r32[kF16BposOff + 3] = 0;
      std::memcpy(rec, r32, sizeof(r32));
    } else {
      std::memcpy(rec + 2 * size_t(kFragWords), tab, sizeof(tab));
      std::memcpy(rec + 2 * size_t(kBposWord), bw, sizeof(bw));
    }
  }
  if (arith == kBlockArithF16) {
    // The tile's largest value is staged 2^headroom lower when the ops can
    // grow the tile's norm (non-unitary matrices): the 2^13.5 bound of the
    // scaled values leaves a factor 4 below fp16's 65504 (TileScale).
    // Beyond 2^10 of growth fp16 would lose the small values: such a
    // program runs on the split-BF16 product (FP32 exponent range).
    const int lg = int(std::ceil(std::log2(max_growth * (1.0 + 1e-9))));
    if (lg > 10) {
      return PrepareGateBlock(n, tile, nt, nops, ks, qs, m64, moff, g, p, fb,
                              kBlockArithBf16);
    }
    p.headroom = unsigned(std::max(0, lg - 2));
  }

  return QSIM_HIP_OK;
}

namespace {

// The kernel shape for a pass's tile size.
template <bool ACC>
void LaunchShapes(float2* s, const BlockGeom& g, const BlockProgram& p,
                  const uint16_t* dfrag) {
  if (g.tb == 14) {
    // One workgroup per CU fits in LDS (144 KiB); a long tile loop each.
    const uint64_t blocks = std::min<uint64_t>(g.ntiles, uint64_t(NumCUs()) * 4);
    hipLaunchKernelGGL((GateBlockKernel<14, true, ACC>), dim3(unsigned(blocks)),
                       dim3(BlockShape<14>::NT), BlockLds<14>(), Stream(), s,
                       g, p, dfrag);
  } else {
    // Two workgroups per CU fit in LDS (2 x 80 KiB); a few tiles each.
    const uint64_t blocks =
        std::min<uint64_t>(g.ntiles, uint64_t(NumCUs()) * 2 * 8);
    if (g.tb == 13) {
      hipLaunchKernelGGL((GateBlockKernel<13, true, ACC>), dim3(unsigned(blocks)),
                         dim3(BlockShape<13>::NT), BlockLds<13>(), Stream(), s,
                         g, p, dfrag);
    } else {
      hipLaunchKernelGGL((GateBlockKernel<13, false, ACC>),
                         dim3(unsigned(blocks)), dim3(BlockShape<13>::NT),
                         BlockLds<13>(), Stream(), s, g, p, dfrag);
    }
  }
}

// The FP16 pass: programs longer than the shape's LDS holds run as
// consecutive launches of at most F16Lds<TB>::kMaxOps ops each.
void LaunchF16(float2* s, const BlockGeom& g, const BlockProgram& p,
               const uint32_t* drec) {
  const unsigned maxops =
      g.tb == 14 ? F16Lds<14>::kMaxOps : F16Lds<13>::kMaxOps;
  for (unsigned c0 = 0; c0 < p.nops; c0 += maxops) {
    BlockProgram q = p;
    q.nops = std::min(maxops, p.nops - c0);
    for (unsigned i = 0; i < q.nops; ++i) {
      std::memcpy(q.tpos[i], p.tpos[c0 + i], sizeof(q.tpos[i]));
      std::memcpy(q.bpos[i], p.bpos[c0 + i], sizeof(q.bpos[i]));
    }
    const uint32_t* r = drec + size_t(c0) * kF16RecWords;
    if (g.tb == 14) {
      const uint64_t blocks =
          std::min<uint64_t>(g.ntiles, uint64_t(NumCUs()) * 4);
      hipLaunchKernelGGL((GateBlockF16Kernel<14, true>), dim3(unsigned(blocks)),
                         dim3(BlockShape<14>::NT), F16Lds<14>::Bytes(q.nops),
                         Stream(), s, g, q, r);
    } else {
      const uint64_t blocks =
          std::min<uint64_t>(g.ntiles, uint64_t(NumCUs()) * 2 * 8);
      if (g.tb == 13) {
        hipLaunchKernelGGL((GateBlockF16Kernel<13, true>),
                           dim3(unsigned(blocks)), dim3(BlockShape<13>::NT),
                           F16Lds<13>::Bytes(q.nops), Stream(), s, g, q, r);
      } else {
        hipLaunchKernelGGL((GateBlockF16Kernel<13, false>),
                           dim3(unsigned(blocks)), dim3(BlockShape<13>::NT),
                           F16Lds<13>::Bytes(q.nops), Stream(), s, g, q, r);
      }
    }
  }
}

}  // namespace

unsigned BlockResidentOps(unsigned tile_bits) {
  return tile_bits == 14 ? F16Lds<14>::kMaxOps : F16Lds<13>::kMaxOps;
}

int RunGateBlock(float2* s, const BlockGeom& g, const BlockProgram& p,
                 const uint16_t* dfrag) {
  // More than 64 KiB of dynamic LDS must be allowed once per process, and
  // the kernels address LDS from offset 0, which holds only with no static
  // LDS ahead of the dynamic block.
  static std::once_flag once;
  static hipError_t attr = hipSuccess;
  static size_t static_lds = 0;
  std::call_once(once, [] {
    const struct {
      const void* fn;
      size_t lds;
    } fns[] = {
        {reinterpret_cast<const void*>(&GateBlockKernel<13, true, false>), BlockLds<13>()},
        {reinterpret_cast<const void*>(&GateBlockKernel<13, false, false>), BlockLds<13>()},
        {reinterpret_cast<const void*>(&GateBlockKernel<14, true, false>), BlockLds<14>()},
        {reinterpret_cast<const void*>(&GateBlockKernel<13, true, true>), BlockLds<13>()},
        {reinterpret_cast<const void*>(&GateBlockKernel<13, false, true>), BlockLds<13>()},
        {reinterpret_cast<const void*>(&GateBlockKernel<14, true, true>), BlockLds<14>()},
        {reinterpret_cast<const void*>(&GateBlockF16Kernel<13, true>),
         F16Lds<13>::Bytes(F16Lds<13>::kMaxOps)},
        {reinterpret_cast<const void*>(&GateBlockF16Kernel<13, false>),
         F16Lds<13>::Bytes(F16Lds<13>::kMaxOps)},
        {reinterpret_cast<const void*>(&GateBlockF16Kernel<14, true>),
         F16Lds<14>::Bytes(F16Lds<14>::kMaxOps)}};
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
    SetError("blocked-pass kernel has %zu bytes of static LDS", static_lds);
    return QSIM_HIP_ERR_RUNTIME;
  }
  PassTimerBegin();
  if (p.arith == kBlockArithF16) {
    LaunchF16(s, g, p, reinterpret_cast<const uint32_t*>(dfrag));
  } else if (p.accurate) {
    LaunchShapes<true>(s, g, p, dfrag);
  } else {
    LaunchShapes<false>(s, g, p, dfrag);
  }
  PassTimerEnd();
  QHIP_LAUNCHED("GateBlockKernel");
  return QSIM_HIP_OK;
}

int LaunchGateBlock(float2* s, unsigned n, const unsigned* tile, unsigned nt,
                    unsigned nops, const unsigned* ks, const unsigned* qs,
                    const double* m64, const size_t* moff, bool accurate,
                    int arith) {
  BlockGeom g{};
  BlockProgram p{};
  std::vector<uint16_t> fb;
  QHIP_TRY(
      PrepareGateBlock(n, tile, nt, nops, ks, qs, m64, moff, g, p, fb, arith));
  p.accurate = accurate ? 1u : 0u;
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
                       const double* m64, const size_t* moff, void** out,
                       bool accurate, int arith) {
  auto* b = new PreparedBlock{};
  std::vector<uint16_t> fb;
  int rc = PrepareGateBlock(n, tile, nt, nops, ks, qs, m64, moff, b->g, b->p,
                            fb, arith);
  b->p.accurate = accurate ? 1u : 0u;
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
