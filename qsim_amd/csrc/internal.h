// internal.h -- shared definitions of the gfx950 state-vector backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <new>
#include <vector>

#include "qsim_hip.h"

namespace qhip {

// ---- error reporting -------------------------------------------------------

void SetError(const char* fmt, ...) __attribute__((format(printf, 1, 2)));

// Runs an entry point's body: a failed host allocation becomes an error
// return instead of an exception crossing the C boundary.
template <typename F>
int GuardHost(F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    SetError("host memory allocation failed");
    return QSIM_HIP_ERR_RUNTIME;
  }
}
int HipFail(hipError_t e, const char* what);

#define QHIP_CHECK(expr)                                   \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return ::qhip::HipFail(e_, #expr); \
  } while (0)

#define QHIP_LAUNCHED(name)                                   \
  do {                                                        \
    hipError_t e_ = hipGetLastError();                        \
    if (e_ != hipSuccess) return ::qhip::HipFail(e_, name);   \
  } while (0)

#define QHIP_TRY(expr)            \
  do {                            \
    int rc_ = (expr);             \
    if (rc_ != QSIM_HIP_OK) return rc_; \
  } while (0)

// ---- context ---------------------------------------------------------------

int EnsureInit();
hipStream_t Stream();

// Grow-only device scratch of the current device (not re-entrant: callers
// hold ScratchLock()).  Returns nullptr and sets the error on failure.
void* DeviceScratch(size_t bytes);
std::mutex& ScratchLock();

// Device buffer for gate matrices too large to travel as kernel arguments.
void* MatrixBuffer(size_t bytes);

// Copies `bytes` of a host matrix into the device slot `dev` (from
// MatrixBuffer) through pinned staging, ordered on the library stream.
int StageMatrix(void* dev, const void* host, size_t bytes);

int NumCUs();

// Runs the gates the gate queue holds (qsim_hip_set_gate_queue).
int FlushGateQueue();
// The gate queue's tile qubits (0: off) and the number of gates it holds.
int QueuedGates(unsigned* tile_bits, uint64_t* gates);

// Op records an FP16 blocked pass of tile_bits tile qubits keeps resident in
// LDS (longer programs stream the state once per such chunk).
unsigned BlockResidentOps(unsigned tile_bits);

// Per-pass timer hooks around gate kernel launches (no-ops unless started).
void PassTimerBegin();
void PassTimerEnd();

// ---- complex element types -------------------------------------------------

template <typename FP> struct Cplx;
template <> struct Cplx<float> { using T = float2; };
template <> struct Cplx<double> { using T = double2; };

struct Dbl2 {
  double x, y;
  __host__ __device__ Dbl2() : x(0), y(0) {}
  __host__ __device__ Dbl2(double a, double b) : x(a), y(b) {}
  __host__ __device__ Dbl2& operator+=(const Dbl2& o) {
    x += o.x;
    y += o.y;
    return *this;
  }
};

constexpr unsigned kMaxInserted = 40;  // targets + controls per gate

// Geometry of a gate pass in the register ("high target") kernels: thread
// item t -> group base = t with zero bits inserted at ins[0..nins) (ascending)
// OR cbits; group member j at base + sum_b bit(j,b) * tbit[b].
struct GateGeom {
  uint64_t ngroups;
  uint64_t cbits;
  uint64_t tbit[6];
  unsigned nins;
  unsigned ins[kMaxInserted];
};

// Geometry of the LDS tile kernels.  A tile holds 2^tb amplitudes: the Lc
// lowest qubits (contiguous in memory) and the h targets at or above Lc.
// Tile index -> base = (tile << Lc) with zero bits inserted at the outer
// positions (high targets and high controls) OR cbits_high.
struct TileGeom {
  uint64_t ntiles;
  uint64_t cbits_high;
  uint64_t hbit[6];
  unsigned nouter;
  unsigned outer[kMaxInserted];
  unsigned Lc;
  unsigned h;
  unsigned tb;
  unsigned tpos[6];     // tile-local target positions, ascending
  unsigned low_cmask;   // controls below Lc (tile-local == global bits)
  unsigned low_cvals;
};

// Cache-blocked passes: a tile of 2^tb amplitudes over the Lc lowest qubits
// and h = tb - Lc chosen higher qubits is staged in LDS once, and a program
// of up to kBlockMaxOps 4-qubit gates on tile-local positions runs on it
// before it is written back.  Tile index -> base = (tile << Lc) with zero
// bits inserted at the high qubits.
constexpr unsigned kBlockMaxHigh = 11;
constexpr unsigned kBlockMaxOps = 32;
constexpr unsigned kBlockMinBits = 9;   // one 32-group batch of 16 members
constexpr unsigned kBlockMaxBits = 14;  // 128 KiB of complex64 per tile

struct BlockGeom {
  uint64_t ntiles;
  uint64_t hbit[kBlockMaxHigh];
  unsigned nouter;
  unsigned outer[kBlockMaxHigh];
  unsigned Lc;
  unsigned h;
  unsigned tb;
};

// Arithmetic of a blocked pass's ops: split-BF16 (matrix in 3 parts,
// amplitudes in 2, or 3 when `accurate`) or FP16 with per-tile scaling
// (matrix and amplitudes in 2 parts each).
constexpr unsigned kBlockArithBf16 = 0;
constexpr unsigned kBlockArithF16 = 1;

struct BlockProgram {
  unsigned nops;
  unsigned accurate;  // bf16: amplitudes in three parts (ApplyTileOp)
  unsigned arith;     // kBlockArith*
  unsigned headroom;  // f16: the tile is staged 2^headroom lower
  uint8_t tpos[kBlockMaxOps][4];  // tile-local targets of each op, ascending
  // Tile-local positions that carry the batch index of each op's groups
  // (tb - 9 of them; unused entries are 0 and only ever see zero bits).
  uint8_t bpos[kBlockMaxOps][kBlockMaxBits - 9];
};

__host__ __device__ inline uint64_t InsertZeros(uint64_t i, const unsigned* pos,
                                                unsigned k) {
  for (unsigned j = 0; j < k; ++j) {
    uint64_t low = i & ((uint64_t{1} << pos[j]) - 1);
    i = ((i >> pos[j]) << (pos[j] + 1)) | low;
  }
  return i;
}

// Host-side launch helpers implemented in gate_kernels.hip / state_kernels.hip.
template <typename FP>
int LaunchGate(typename Cplx<FP>::T* s, unsigned n, const unsigned* qs,
               unsigned k, const unsigned* cqs, unsigned nc, uint64_t cvals,
               const FP* m);

template <typename FP>
int LaunchExpectation(const typename Cplx<FP>::T* s, unsigned n,
                      const unsigned* qs, unsigned k, const FP* m, double* re,
                      double* im);

template <typename FP> using CT = typename Cplx<FP>::T;

// State-space passes (state_kernels.hip).  Fill: 0 zeros, 1 |0>, 2 uniform.
template <typename FP> int Fill(CT<FP>* s, unsigned n, int what);
template <typename FP>
int BulkSet(CT<FP>* s, unsigned n, uint64_t mask, uint64_t bits, double re,
            double im, bool exclude);
template <typename FP> int Add(const CT<FP>* src, CT<FP>* dst, unsigned n);
template <typename FP> int Scale(CT<FP>* s, unsigned n, double a);
template <typename FP>
int Dot(const CT<FP>* a, const CT<FP>* b, unsigned n, bool real, double* re,
        double* im);
// Dot on the fixed device scratch instead of the grow-only one: no device
// allocation.  InitStaticScratch takes the scratch's device address (once,
// at library initialisation, with the device current).
template <typename FP>
int DotStatic(const CT<FP>* a, const CT<FP>* b, unsigned n, bool real,
              double* re, double* im);
int InitStaticScratch();
unsigned PartialNormCount(unsigned n);
template <typename FP> int PartialNorms(const CT<FP>* s, unsigned n, double* out);
template <typename FP>
int FindMeasuredBits(const CT<FP>* s, unsigned n, unsigned m, double r,
                     uint64_t mask, uint64_t* out);
template <typename FP>
int Collapse(CT<FP>* s, unsigned n, uint64_t mask, uint64_t bits);
template <typename FP>
int Sample(const CT<FP>* s, unsigned n, const double* draws, uint64_t num,
           uint64_t* out);

template <typename FP>
int Gather(const CT<FP>* s, const uint64_t* idx, uint64_t count, CT<FP>* out);

int Convert(void* dst, int dprec, const void* src, int sprec, unsigned n);

// MFMA path for uncontrolled k>=5 gates on complex64 states.  `m64` is the
// matrix in FP64 (rounded products are formed from it); returns
// QSIM_HIP_ERR_UNSUPPORTED when the method cannot handle the gate.
int LaunchGateMfma(float2* s, unsigned n, const unsigned* qs, unsigned k,
                   const double* m64, int method);

// FP16 MFMA with per-group scaling for uncontrolled k = 5, 6 gates on
// complex64 states (f16_kernels.hip); QSIM_HIP_ERR_UNSUPPORTED when the
// state is too small for its batches.
int LaunchGateF16(float2* s, unsigned n, const unsigned* qs, unsigned k,
                  const double* m64);

// Exact-FP32 MFMA for uncontrolled k = 5, 6 gates with each wave's row
// block of the matrix held in registers (f16_kernels.hip); `frag`: the
// device FP32 fragments of mfma_kernels.hip.  QSIM_HIP_ERR_UNSUPPORTED when
// the state is too small for its sets.
int LaunchRowsF32(float2* s, unsigned n, const unsigned* qs, unsigned k,
                  const float* frag);
// The same for the split-BF16 product; `frag`: BuildFragBf16's layout.
int LaunchRowsBf16(float2* s, unsigned n, const unsigned* qs, unsigned k,
                   const uint16_t* frag);

// Split-BF16 A-operand fragments of a 2^K x 2^K complex matrix (FP64,
// row-major interleaved): 2^K/16 row blocks x 2^(K+1)/16 K blocks x
// (hi, mid, lo) x 512.
void BuildFragBf16(int K, const double* m, uint16_t* frag);

// One cache-blocked pass (block_kernels.hip).  `tile` lists the tile qubits
// ascending; op i acts on ks[i] <= 4 qubits qs[4i ..] (ascending, inside the
// tile) with the FP64 matrix at m64 + moff[i].
int LaunchGateBlock(float2* s, unsigned n, const unsigned* tile, unsigned nt,
                    unsigned nops, const unsigned* ks, const unsigned* qs,
                    const double* m64, const size_t* moff,
                    bool accurate = false, int arith = kBlockArithF16);

// The host half of a blocked pass: validates the program and builds its
// geometry, per-op tables and op records (kOpWords 32-bit words each for
// split-BF16, kF16RecWords for FP16, in `fb` as 16-bit halves).  An FP16
// program whose matrices can grow the state by more than 2^10 is built as
// split-BF16 instead (p.arith says which).
int PrepareGateBlock(unsigned n, const unsigned* tile, unsigned nt,
                     unsigned nops, const unsigned* ks, const unsigned* qs,
                     const double* m64, const size_t* moff, BlockGeom& g,
                     BlockProgram& p, std::vector<uint16_t>& fb,
                     int arith = kBlockArithBf16);

// The same pass prepared once (device fragments kept) and run many times.
int PrepareBlockHandle(unsigned n, const unsigned* tile, unsigned nt,
                       unsigned nops, const unsigned* ks, const unsigned* qs,
                       const double* m64, const size_t* moff, void** out,
                       bool accurate = false, int arith = kBlockArithF16);
int RunBlockHandle(void* h, float2* s, unsigned n);
int FreeBlockHandle(void* h);

}  // namespace qhip
