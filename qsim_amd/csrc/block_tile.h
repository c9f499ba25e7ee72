// block_tile.h -- the tile machinery of blocked passes, shared by the
// circuit kernels (block_kernels.hip) and the batched trajectory kernel
// (traj_kernels.hip): kernel shapes, the op-record layouts, the MFMA
// product of one 4-qubit op over an LDS tile (FP16 with scaling, the
// default, and split-BF16), and tile staging between HBM and LDS.
//
// Each op is a 4-qubit gate on tile-local positions (smaller gates are
// padded with identity factors on the host).  The split-BF16 form is the
// product of the gate kernels (mfma_kernels.hip): the 16 x 16 complex
// matrix becomes a real 32 x 32 A operand, 32 groups side by side the B
// operand of v_mfma_f32_32x32x16_bf16; the matrix is split in three bf16
// parts, the amplitudes in two, and Al Bh + Am Bh + Ah Bl + Ah Bh is
// accumulated in FP32.  Operand maps (verified on gfx950):
// a[e] = A[l%32][8(l/32) + e], b[e] = B[8(l/32) + e][l%32],
// C register r of lane l = C[8(r/4) + 4(l/32) + r%4][l%32].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.h"
#include "tile.h"

namespace qhip {
namespace tileop {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));

// Kernel shapes by tile qubits TB: up to 13 (64 KiB tiles), two 256-thread
// workgroups per CU; 14 (128 KiB), one 512-thread workgroup per CU.  Both
// give every thread 32 tile elements and every wave 256 registers.
template <int TB>
struct BlockShape {
  static_assert(TB == 13 || TB == 14, "tile shape");
  static constexpr int NT = TB == 13 ? 256 : 512;
  static constexpr int kWaves = NT / 64;
  static constexpr size_t kTileBytes = (size_t{1} << TB) * sizeof(float2);
};
// BF16 fragment elements of one 4-qubit op: 2 K blocks x (hi, mid, lo) x 512.
constexpr size_t kOpFrag = 2 * 3 * 512;
// One op record in 32-bit words: the fragments (1536 words); the member
// tables -- LDS byte offsets, relative to a group's slot, of the members a
// lane reads (8) and writes (8), for lane half 0 then 1 (32 words); the
// group table -- LDS byte offset of lane j's group within a batch, j < 32
// (32 words); the op's five batch-bit positions (one byte each, 2 words);
// padded to 8 KiB.  The positions travel in the record so that an op reads
// them from LDS: read from the kernel arguments (bytes at a run-time index)
// they would be vector memory loads, and waiting for them would wait for
// every tile load issued before them.
constexpr int kFragWords = int(kOpFrag / 2);
constexpr int kTableWords = 32;
constexpr int kLaneWords = 32;
constexpr int kBposWord = kFragWords + kTableWords + kLaneWords;
constexpr int kOpWords = 2048;
static_assert(kBposWord + 2 <= kOpWords, "op record");
template <int TB>
constexpr size_t BlockLds() {
  return BlockShape<TB>::kTileBytes + 2 * kOpWords * sizeof(uint32_t);
}

// The tile at LDS byte offset `b`: the kernel has no static LDS, so its
// dynamic block starts at offset 0 (checked on the host before launch), and
// an address is the offset itself, with no base to add.
typedef __attribute__((address_space(3))) v2f LdsV2f;
__device__ __forceinline__ LdsV2f* LdsAt(unsigned b) {
  return reinterpret_cast<LdsV2f*>(size_t{b});
}

// One 4-qubit op over the LDS tile.  A group (the 16 amplitudes the op
// mixes) is fixed by its bits at the tile's non-target positions; the host
// splits those positions into 5 lane positions (chosen so that the 32
// groups of a batch fall into distinct LDS banks) and tb - 9 batch
// positions `bp`.  Wave `wave` (wave-uniform) of `nwaves` takes batches
// bt = wave, wave + nwaves, ...; lanes j and j + 32 serve the batch's j-th
// group.  `f` holds the op's fragments (hi, mid, lo per K block); rd / wr
// are this lane's member offsets with its group's offset already folded in
// (LDS bytes).  TileSwz is linear over GF(2) and the lane part, batch part
// and member offsets occupy disjoint bits, so a member's slot is the XOR of
// the three parts' swizzled slots: the batch part is scalar work and an
// address is one XOR (the tile sits at LDS offset 0).
__device__ __forceinline__ unsigned BatchSlot(unsigned bt,
                                              const unsigned (&b)[5]) {
  unsigned pb = 0;
#pragma unroll
  for (int i = 0; i < 5; ++i) pb |= ((bt >> i) & 1u) << b[i];
  return TileSwz(pb) * unsigned(sizeof(float2));
}

// Lane half h's 8 members of its group as bf16 (hi, lo) pairs.
__device__ __forceinline__ void SplitBf16(const v2f (&v)[8], v8bf (&bh)[2],
                                          v8bf (&bl)[2]) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const v2bf hi = __builtin_convertvector(v[q], v2bf);
    const v2f r = v[q] - __builtin_convertvector(hi, v2f);
    const v2bf lo = __builtin_convertvector(r, v2bf);
    bh[q >> 2][2 * (q & 3)] = hi[0];
    bh[q >> 2][2 * (q & 3) + 1] = hi[1];
    bl[q >> 2][2 * (q & 3)] = lo[0];
    bl[q >> 2][2 * (q & 3) + 1] = lo[1];
  }
}

// The op on one batch: Al Bh + Am Bh + Ah Bl + Ah Bh per K block, one
// accumulator chain.
__device__ __forceinline__ v16f MfmaChain(const v8bf (&f)[6],
                                          const v8bf (&bh)[2],
                                          const v8bf (&bl)[2]) {
  v16f c = v16f{};
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[2], bh[0], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[1], bh[0], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[0], bl[0], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[0], bh[0], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[5], bh[1], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[4], bh[1], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[3], bl[1], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f[3], bh[1], c, 0, 0, 0);
  return c;
}

// Lane half h's 8 members of its group in three bf16 parts (hi, mid, lo):
// the amplitudes to about 2^-25 instead of 2^-18.
__device__ __forceinline__ void SplitBf16x3(const v2f (&v)[8], v8bf (&bh)[2],
                                            v8bf (&bm)[2], v8bf (&bl)[2]) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const v2bf hi = __builtin_convertvector(v[q], v2bf);
    const v2f r1 = v[q] - __builtin_convertvector(hi, v2f);
    const v2bf mid = __builtin_convertvector(r1, v2bf);
    const v2f r2 = r1 - __builtin_convertvector(mid, v2f);
    const v2bf lo = __builtin_convertvector(r2, v2bf);
    bh[q >> 2][2 * (q & 3)] = hi[0];
    bh[q >> 2][2 * (q & 3) + 1] = hi[1];
    bm[q >> 2][2 * (q & 3)] = mid[0];
    bm[q >> 2][2 * (q & 3) + 1] = mid[1];
    bl[q >> 2][2 * (q & 3)] = lo[0];
    bl[q >> 2][2 * (q & 3) + 1] = lo[1];
  }
}

// The accurate product: every term down to 2^-16 relative (Al Bh + Am Bm +
// Ah Bl + Am Bh + Ah Bm + Ah Bh per K block, smallest first), 12 MFMAs.
__device__ __forceinline__ v16f MfmaChain6(const v8bf (&f)[6],
                                           const v8bf (&bh)[2],
                                           const v8bf (&bm)[2],
                                           const v8bf (&bl)[2]) {
  v16f c = v16f{};
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    const v8bf& ah = f[3 * kb];
    const v8bf& am = f[3 * kb + 1];
    const v8bf& al = f[3 * kb + 2];
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh[kb], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bm[kb], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl[kb], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bh[kb], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bm[kb], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh[kb], c, 0, 0, 0);
  }
  return c;
}

// Lane l holds (re, im) of members 4 s + 2 h + {0, 1}.
__device__ __forceinline__ void WriteBatch(unsigned sp, const uint32_t (&wr)[8],
                                           const v16f& c) {
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      *LdsAt(sp ^ wr[2 * q + u]) = v2f{c[4 * q + 2 * u], c[4 * q + 2 * u + 1]};
    }
}

// ACCURATE: amplitudes in three parts and the 12-MFMA product (one batch at
// a time), for passes on structured states such as the first passes from a
// basis state, where the two-part split's rounding is the same for many
// amplitudes and shifts the norm coherently.
template <bool ACCURATE = false>
__device__ __forceinline__ void ApplyTileOp(const uint8_t* bp,
                                            const v8bf (&f)[6],
                                            const uint32_t (&rd)[8],
                                            const uint32_t (&wr)[8],
                                            unsigned nbatch, unsigned wave,
                                            unsigned nwaves) {
  unsigned b[5];
#pragma unroll
  for (int i = 0; i < 5; ++i)
    b[i] = unsigned(__builtin_amdgcn_readfirstlane(int(bp[i])));
  if (ACCURATE) {
    for (unsigned bt = wave; bt < nbatch; bt += nwaves) {
      const unsigned sp0 = BatchSlot(bt, b);
      v2f v0[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v0[q] = *LdsAt(sp0 ^ rd[q]);
      v8bf bh0[2], bm0[2], bl0[2];
      SplitBf16x3(v0, bh0, bm0, bl0);
      WriteBatch(sp0, wr, MfmaChain6(f, bh0, bm0, bl0));
    }
    return;
  }
  // Two batches at a time (their groups are disjoint): batch 1's reads and
  // split overlap batch 0's MFMA chain.
  unsigned bt = wave;
  for (; bt + nwaves < nbatch; bt += 2 * nwaves) {
    const unsigned sp0 = BatchSlot(bt, b);
    const unsigned sp1 = BatchSlot(bt + nwaves, b);
    v2f v0[8], v1[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v0[q] = *LdsAt(sp0 ^ rd[q]);
    v8bf bh0[2], bl0[2], bh1[2], bl1[2];
    SplitBf16(v0, bh0, bl0);
#pragma unroll
    for (int q = 0; q < 8; ++q) v1[q] = *LdsAt(sp1 ^ rd[q]);
    const v16f c0 = MfmaChain(f, bh0, bl0);
    SplitBf16(v1, bh1, bl1);
    const v16f c1 = MfmaChain(f, bh1, bl1);
    WriteBatch(sp0, wr, c0);
    WriteBatch(sp1, wr, c1);
  }
  if (bt < nbatch) {
    const unsigned sp0 = BatchSlot(bt, b);
    v2f v0[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v0[q] = *LdsAt(sp0 ^ rd[q]);
    v8bf bh0[2], bl0[2];
    SplitBf16(v0, bh0, bl0);
    WriteBatch(sp0, wr, MfmaChain(f, bh0, bl0));
  }
}

// ---- FP16-with-scaling ops ---------------------------------------------
//
// The same product on v_mfma_f32_32x32x16_f16: the matrix in two fp16 parts
// (Ah + Al, 22 bits), the amplitudes in two (Bh + Bl, 22 bits) and
// Al Bh + Ah Bl + Ah Bh accumulated in FP32 (6 MFMAs per 32 groups instead
// of 8; the dropped Al Bl is 2^-22 relative).  fp16 has a 5-bit exponent,
// so the tile is scaled by a power of two when it is staged (TileScale):
// its largest real or imaginary part goes to [2^5, 2^6) / 2^headroom.  The
// ops of a pass keep the tile's norm (unitary, or contractive for headroom
// 0), so no value passes sqrt(2 * 2^14) * 2^6 < 2^14 < 65504; values below
// 2^-8 of the tile's largest keep an absolute error under 2^-30 of it.
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef _Float16 v2h __attribute__((ext_vector_type(2)));

// FP16 op record, 32-bit words: the 16 x 16 complex matrix as packed
// (re, im) fp16 pairs, row r at words 16 r .. 16 r + 15 with record column
// p holding matrix column F16Col(p), hi part then lo part; the member
// tables, the group table and the batch-bit positions as in the bf16
// record; padded to a multiple of 4 words.  With that column order the 8
// members a lane reads for the B operand are the 8 members its C registers
// hold, in the same order (both: 8 kb + 4 (c / 2) + 2 h + c % 2 for
// q = 4 kb + c), so one LDS address per member serves the read and the
// write of a batch (the read and write halves of the member table are the
// same).
constexpr int kF16MatWords = 256;
constexpr int kF16TabOff = 2 * kF16MatWords;
constexpr int kF16BposOff = kF16TabOff + kTableWords + kLaneWords;
constexpr int kF16RecWords = kF16BposOff + 4;  // 580
static_assert(kF16RecWords % 4 == 0, "records are copied as uint4");

// Matrix column stored at record column p: p with bits 1 and 2 swapped.
__host__ __device__ constexpr unsigned F16Col(unsigned p) {
  return (p & 9u) | ((p & 2u) << 1) | ((p & 4u) >> 1);
}

// LDS of an FP16 blocked-pass workgroup: the tile, 16 floats for the
// per-wave maxima, then the pass's op records (all of them, loaded once per
// workgroup).  14-qubit tiles: one workgroup per CU (160 KiB); 13: two.
template <int TB>
struct F16Lds {
  static constexpr size_t kLimit = TB == 14 ? 163840 : 81920;
  static constexpr size_t kMaxOff = BlockShape<TB>::kTileBytes;
  static constexpr size_t kRecOff = kMaxOff + 64;
  static constexpr unsigned kMaxOps =
      unsigned((kLimit - kRecOff) / (kF16RecWords * sizeof(uint32_t)));
  static constexpr size_t Bytes(unsigned nops) {
    return kRecOff + size_t(nops) * kF16RecWords * sizeof(uint32_t);
  }
};

// Lane l's A fragments of the op whose record is at r (LDS): for K block
// kb, f[2 kb] = hi and f[2 kb + 1] = lo of A[l%32][16 kb + 8 (l/32) + e].
// Row l%32 = 2 i + p reads record columns 8 kb + 4 (l/32) + c, c < 4, of
// row i, and forms (Re, -Im) for p = 0 or (Im, Re) for p = 1 from the
// stored (Re, Im) pair.
__device__ __forceinline__ void LoadFragF16(const uint32_t* r, int l,
                                            v8h (&f)[4]) {
  const unsigned i = unsigned(l & 31) >> 1, hh = unsigned(l) >> 5;
  const unsigned rot = (l & 1) ? 16u : 0u;
  const uint32_t sx = (l & 1) ? 0u : 0x80000000u;
#pragma unroll
  for (int kb = 0; kb < 2; ++kb)
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      const uint4 w = *reinterpret_cast<const uint4*>(
          r + part * kF16MatWords + 16 * i + 8 * kb + 4 * hh);
      uint4 x;
      x.x = __builtin_amdgcn_alignbit(w.x, w.x, rot) ^ sx;
      x.y = __builtin_amdgcn_alignbit(w.y, w.y, rot) ^ sx;
      x.z = __builtin_amdgcn_alignbit(w.z, w.z, rot) ^ sx;
      x.w = __builtin_amdgcn_alignbit(w.w, w.w, rot) ^ sx;
      f[2 * kb + part] = __builtin_bit_cast(v8h, x);
    }
}

// Lane half h's 8 members as fp16 (hi, lo) pairs: hi rounds to nearest;
// the residual v - hi is exact in FP32 and v_fma_mix{lo,hi}_f16 round it
// to fp16 straight from the fp16 half, one instruction per value.
__device__ __forceinline__ void SplitF16(const v2f (&v)[8], v8h (&bh)[2],
                                         v8h (&bl)[2]) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const v2h hi = __builtin_convertvector(v[q], v2h);
    uint32_t lw;
    asm("v_fma_mixlo_f16 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]"
        : "=v"(lw)
        : "v"(hi), "v"(v[q][0]));
    asm("v_fma_mixhi_f16 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
        : "+v"(lw)
        : "v"(hi), "v"(v[q][1]));
    const v2h lo = __builtin_bit_cast(v2h, lw);
    bh[q >> 2][2 * (q & 3)] = hi[0];
    bh[q >> 2][2 * (q & 3) + 1] = hi[1];
    bl[q >> 2][2 * (q & 3)] = lo[0];
    bl[q >> 2][2 * (q & 3) + 1] = lo[1];
  }
}

// Al Bh + Ah Bl + Ah Bh per K block, smallest first, one accumulator chain.
__device__ __forceinline__ v16f MfmaChainF16(const v8h (&f)[4],
                                             const v8h (&bh)[2],
                                             const v8h (&bl)[2]) {
  v16f c = v16f{};
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(f[2 * kb + 1], bh[kb], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(f[2 * kb], bl[kb], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(f[2 * kb], bh[kb], c, 0, 0, 0);
  }
  return c;
}

// Results of a batch back to the slots its inputs came from: C registers
// (4 q + 2 u, 4 q + 2 u + 1) are member index 2 q + u of the lane's list.
__device__ __forceinline__ void WriteBack(const uint32_t (&a)[8],
                                          const v16f& c) {
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      *LdsAt(a[2 * q + u]) = v2f{c[4 * q + 2 * u], c[4 * q + 2 * u + 1]};
    }
}

// One FP16 op over the (scaled) LDS tile; batches as in ApplyTileOp, two
// at a time.  mo: the lane's member offsets (record column order, so that
// reads and writes share them) XOR its group's slot.
__device__ __forceinline__ void ApplyTileOpF16(const uint8_t* bp,
                                               const v8h (&f)[4],
                                               const uint32_t (&mo)[8],
                                               unsigned nbatch, unsigned wave,
                                               unsigned nwaves) {
  unsigned b[5];
#pragma unroll
  for (int i = 0; i < 5; ++i)
    b[i] = unsigned(__builtin_amdgcn_readfirstlane(int(bp[i])));
  unsigned bt = wave;
  for (; bt + nwaves < nbatch; bt += 2 * nwaves) {
    const unsigned sp0 = BatchSlot(bt, b);
    const unsigned sp1 = BatchSlot(bt + nwaves, b);
    uint32_t a0[8], a1[8];
    v2f v0[8], v1[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      a0[q] = sp0 ^ mo[q];
      v0[q] = *LdsAt(a0[q]);
    }
    v8h bh0[2], bl0[2], bh1[2], bl1[2];
    SplitF16(v0, bh0, bl0);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      a1[q] = sp1 ^ mo[q];
      v1[q] = *LdsAt(a1[q]);
    }
    const v16f c0 = MfmaChainF16(f, bh0, bl0);
    SplitF16(v1, bh1, bl1);
    const v16f c1 = MfmaChainF16(f, bh1, bl1);
    WriteBack(a0, c0);
    WriteBack(a1, c1);
  }
  if (bt < nbatch) {
    const unsigned sp0 = BatchSlot(bt, b);
    uint32_t a0[8];
    v2f v0[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      a0[q] = sp0 ^ mo[q];
      v0[q] = *LdsAt(a0[q]);
    }
    v8h bh0[2], bl0[2];
    SplitF16(v0, bh0, bl0);
    WriteBack(a0, MfmaChainF16(f, bh0, bl0));
  }
}

// Largest |re| or |im| over the first `nper` of a thread's PER elements
// (FULL: all of them).
template <int PER, bool FULL>
__device__ __forceinline__ float LocalMax(const float2 (&v)[PER],
                                          unsigned nper) {
  float m = 0.0f;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    if (FULL || unsigned(i) < nper) {
      m = fmaxf(m, fmaxf(fabsf(v[i].x), fabsf(v[i].y)));
    }
  }
  return m;
}

// Workgroup maximum of `m` through `slots` (one float per wave), then the
// tile's scale exponent: the largest value to [2^5, 2^6) / 2^headroom, or
// 0 for an all-zero (or non-finite) tile.  Ends with a barrier before the
// slots are read; the slots are written again only after a later barrier.
template <int NWAVES>
__device__ __forceinline__ int TileScale(float m, float* slots,
                                         unsigned headroom) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_xor(m, d));
  if ((threadIdx.x & 63) == 0) slots[threadIdx.x >> 6] = m;
  __syncthreads();
  float t = 0.0f;
#pragma unroll
  for (int w = 0; w < NWAVES; ++w) t = fmaxf(t, slots[w]);
  if (!(t > 0.0f) || !(t <= 3.0e38f)) return 0;
  int e;
  (void)frexpf(t, &e);  // t in [2^(e-1), 2^e)
  return 6 - int(headroom) - e;  // |e| <= 149: ldexpf takes any such shift
}

// Offset (in amplitudes) of tile-local high index x: x's bits deposited at
// the tile's high qubits.
__device__ __forceinline__ uint64_t DepositHigh(unsigned x, const BlockGeom& g) {
  uint64_t o = 0;
  for (unsigned b = 0; b < g.h; ++b) o += ((x >> b) & 1u) ? g.hbit[b] : 0;
  return o;
}

// Tile staging with addresses split into a per-thread part and a per-element
// part.  Element i of thread t is tile-local amplitude a = t + NT i (valid
// for i < 2^tb / NT, a wave-uniform test).  The host keeps the low run
// within the thread bits (Lc <= log2 NT), so a's low bits come from t alone
// and its high index is (t >> Lc) | (i << (log2 NT - Lc)) with disjoint
// bits: its offset is D(t >> Lc) + D(i << (log2 NT - Lc)), one add per
// element instead of a deposit loop.  The element parts are wave-uniform
// and kept in units of 2^Lc amplitudes (32 bits: no tile qubit is 32 above
// the low run).
template <int PER, int NT>
struct TileIo {
  static_assert(NT == 256 || NT == 512, "thread bits");
  static constexpr unsigned kLogNT = NT == 256 ? 8 : 9;
  uint64_t tofs;       // this thread's part
  uint32_t eofs[PER];  // the element parts >> Lc

  // LDS slot of element i: TileSwz is linear and changes only the low four
  // bits, so TileSwz(t + NT i) = (TileSwz(t) ^ (TileSwz(NT i) & 15)) + NT i
  // with TileSwz(t) < NT: the NT i part is an instruction offset and the
  // rest takes one XOR with a constant.
  __device__ __forceinline__ unsigned Slot(int i) const {
    return (TileSwz(threadIdx.x) ^ (TileSwz(NT * unsigned(i)) & 15u)) +
           NT * unsigned(i);
  }

  __device__ __forceinline__ void Init(const BlockGeom& g) {
    const unsigned t = threadIdx.x;
    tofs = (t & ((1u << g.Lc) - 1)) + DepositHigh(t >> g.Lc, g);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      eofs[i] = uint32_t(DepositHigh(unsigned(i) << (kLogNT - g.Lc), g) >> g.Lc);
    }
  }

  // Issues the loads of the tile at `base` into v (this thread's elements),
  // those of the quarters of v whose bits are set in MASK (quarter c holds
  // elements c PER/4 .. (c+1) PER/4 - 1).
  template <bool FULL, unsigned MASK>
  __device__ __forceinline__ void Fetch(const float2* __restrict__ s,
                                        uint64_t base, const BlockGeom& g,
                                        float2 (&v)[PER]) const {
    const unsigned nper = (1u << g.tb) >> kLogNT;  // elements per thread
    const float2* p = s + base + tofs;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      if (((MASK >> (4 * i / PER)) & 1u) && (FULL || unsigned(i) < nper)) {
        v[i] = p[uint64_t(eofs[i]) << g.Lc];
      }
    }
  }

  // Writes fetched elements to their LDS slots.
  template <bool FULL>
  __device__ __forceinline__ void Put(const BlockGeom& g,
                                      const float2 (&v)[PER],
                                      float2* lds) const {
    const unsigned nper = (1u << g.tb) >> kLogNT;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      if (FULL || unsigned(i) < nper) lds[Slot(i)] = v[i];
    }
  }

  // Put with every value multiplied by 2^e (exact for normal results).
  template <bool FULL>
  __device__ __forceinline__ void PutScaled(const BlockGeom& g,
                                            const float2 (&v)[PER],
                                            float2* lds, int e) const {
    const unsigned nper = (1u << g.tb) >> kLogNT;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      if (FULL || unsigned(i) < nper) {
        lds[Slot(i)] = float2{ldexpf(v[i].x, e), ldexpf(v[i].y, e)};
      }
    }
  }

  // Store with every value multiplied by 2^-e.
  template <int CH, bool FULL>
  __device__ __forceinline__ void StoreScaled(float2* __restrict__ s,
                                              uint64_t base,
                                              const BlockGeom& g,
                                              const float2* lds,
                                              int e) const {
    const unsigned nper = (1u << g.tb) >> kLogNT;
    float2* p = s + base + tofs;
#pragma unroll
    for (int i0 = 0; i0 < PER; i0 += CH) {
      float2 v[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        if (FULL || unsigned(i0 + i) < nper) {
          const float2 x = lds[Slot(i0 + i)];
          v[i] = float2{ldexpf(x.x, -e), ldexpf(x.y, -e)};
        }
      }
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        if (FULL || unsigned(i0 + i) < nper) p[uint64_t(eofs[i0 + i]) << g.Lc] = v[i];
      }
    }
  }

  template <int CH, bool FULL>
  __device__ __forceinline__ void Store(float2* __restrict__ s, uint64_t base,
                                        const BlockGeom& g,
                                        const float2* lds) const {
    const unsigned nper = (1u << g.tb) >> kLogNT;
    float2* p = s + base + tofs;
#pragma unroll
    for (int i0 = 0; i0 < PER; i0 += CH) {
      float2 v[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        if (FULL || unsigned(i0 + i) < nper) v[i] = lds[Slot(i0 + i)];
      }
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        if (FULL || unsigned(i0 + i) < nper) p[uint64_t(eofs[i0 + i]) << g.Lc] = v[i];
      }
    }
  }
};

}  // namespace tileop
}  // namespace qhip
