// state_kernels.hip -- state-space passes of the gfx950 backend: fills,
// elementwise updates, FP64 reductions, measurement and parallel sampling.
//
// Counterparts: StateSpaceCUDA (lib/statespace_cuda.h) and its kernels
// (lib/statespace_cuda_kernels.h), with the arithmetic of StateSpaceBasic
// (lib/statespace_basic.h).  The layout is interleaved complex, so the
// reference's layout conversions are no-ops here.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "internal.h"
#include "reduce.h"

namespace qhip {

namespace {

constexpr unsigned kBS = 256;

inline unsigned GridFor(uint64_t items, uint64_t cap = uint64_t{1} << 16) {
  uint64_t b = (items + kBS - 1) / kBS;
  return unsigned(std::max<uint64_t>(1, std::min(b, cap)));
}

template <typename C>
__global__ void __launch_bounds__(kBS)
FillKernel(C* __restrict__ s, uint64_t size, C v, uint64_t one_index, C one) {
  const uint64_t stride = uint64_t(gridDim.x) * kBS;
  for (uint64_t i = uint64_t(blockIdx.x) * kBS + threadIdx.x; i < size;
       i += stride)
    s[i] = i == one_index ? one : v;
}

template <typename C, typename FP>
__global__ void __launch_bounds__(kBS)
BulkSetKernel(C* __restrict__ s, uint64_t size, uint64_t mask, uint64_t bits,
              FP re, FP im, bool exclude) {
  const uint64_t stride = uint64_t(gridDim.x) * kBS;
  for (uint64_t i = uint64_t(blockIdx.x) * kBS + threadIdx.x; i < size;
       i += stride) {
    bool in = ((i & mask) == bits) != exclude;
    if (in) s[i] = C{re, im};
  }
}

template <typename C>
__global__ void __launch_bounds__(kBS)
AddKernel(const C* __restrict__ src, C* __restrict__ dst, uint64_t size) {
  const uint64_t stride = uint64_t(gridDim.x) * kBS;
  for (uint64_t i = uint64_t(blockIdx.x) * kBS + threadIdx.x; i < size;
       i += stride) {
    C a = src[i], b = dst[i];
    b.x += a.x;
    b.y += a.y;
    dst[i] = b;
  }
}

template <typename C, typename FP>
__global__ void __launch_bounds__(kBS)
ScaleKernel(C* __restrict__ s, uint64_t size, FP a) {
  const uint64_t stride = uint64_t(gridDim.x) * kBS;
  for (uint64_t i = uint64_t(blockIdx.x) * kBS + threadIdx.x; i < size;
       i += stride) {
    C v = s[i];
    v.x *= a;
    v.y *= a;
    s[i] = v;
  }
}

// Zero amplitudes outside (i & mask) == bits, scale the rest
// (lib/statespace_cuda_kernels.h:273-285 semantics).
template <typename C, typename FP>
__global__ void __launch_bounds__(kBS)
CollapseKernel(C* __restrict__ s, uint64_t size, uint64_t mask, uint64_t bits,
               FP renorm) {
  const uint64_t stride = uint64_t(gridDim.x) * kBS;
  for (uint64_t i = uint64_t(blockIdx.x) * kBS + threadIdx.x; i < size;
       i += stride) {
    C v = s[i];
    bool keep = (i & mask) == bits;
    v.x = keep ? v.x * renorm : FP(0);
    v.y = keep ? v.y * renorm : FP(0);
    s[i] = v;
  }
}

// Partial sums of |s_i|^2 over contiguous ranges of `part` amplitudes, one
// workgroup per range; sums in FP64.  Products in FP (Norm, PartialNorms,
// Collapse: lib/statespace_basic.h:237-273) or in FP64 (DBL: Sample,
// lib/statespace_basic.h:203-222).  MASKED keeps (i & mask) == bits only.
template <typename C, bool MASKED, bool DBL = false>
__global__ void __launch_bounds__(kBS)
NormPartsKernel(const C* __restrict__ s, uint64_t part, uint64_t mask,
                uint64_t bits, double* __restrict__ out) {
  const uint64_t i0 = uint64_t(blockIdx.x) * part;
  double acc = 0;
  for (uint64_t i = i0 + threadIdx.x; i < i0 + part; i += kBS) {
    C v = s[i];
    if (MASKED && (i & mask) != bits) continue;
    if (DBL)
      acc += double(v.x) * double(v.x) + double(v.y) * double(v.y);
    else
      acc += double(v.x * v.x + v.y * v.y);
  }
  acc = BlockSum<kBS>(acc);
  if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

// sum conj(a_i) b_i (REAL: real part only) with FP64 accumulation
// (lib/statespace_cuda_kernels.h:181-190 via lib/util_cuda.h:101-122).
template <typename C, bool REAL>
__global__ void __launch_bounds__(kBS)
DotKernel(const C* __restrict__ a, const C* __restrict__ b, uint64_t size,
          Dbl2* __restrict__ out) {
  Dbl2 acc;
  const uint64_t stride = uint64_t(gridDim.x) * kBS;
  for (uint64_t i = uint64_t(blockIdx.x) * kBS + threadIdx.x; i < size;
       i += stride) {
    C u = a[i], v = b[i];
    acc.x += double(u.x * v.x + u.y * v.y);
    if (!REAL) acc.y += double(u.x * v.y - u.y * v.x);
  }
  acc = BlockSum<kBS>(acc);
  if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

// Finds, inside amplitudes [i0, i1), the first k with r < running sum of
// |s|^2 (FP64, in index order); writes k or ~0 when the range ends first.
// One workgroup: 256-amplitude chunks, workgroup prefix sums with a carry.
template <typename C>
__global__ void __launch_bounds__(kBS)
FindFirstKernel(const C* __restrict__ s, uint64_t i0, uint64_t i1, double r,
                uint64_t* __restrict__ out) {
  __shared__ double wave_tot[kBS / 64];
  __shared__ uint64_t found;
  __shared__ double carry_s;
  if (threadIdx.x == 0) {
    found = ~uint64_t{0};
    carry_s = 0;
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (uint64_t c0 = i0; c0 < i1; c0 += kBS) {
    uint64_t i = c0 + threadIdx.x;
    double p = 0;
    if (i < i1) {
      C v = s[i];
      p = double(v.x * v.x + v.y * v.y);
    }
    double incl = WaveInclusiveScan(p);
    if (l == 63) wave_tot[w] = incl;
    __syncthreads();
    double before = carry_s;
    for (int j = 0; j < w; ++j) before += wave_tot[j];
    double csum = before + incl;
    // The first index of the chunk that satisfies r < csum.
    bool hit = i < i1 && r < csum;
    unsigned long long ballot = __ballot(hit);
    if (l == 0 && ballot != 0) {
      uint64_t cand = c0 + uint64_t(w) * 64 + __builtin_ctzll(ballot);
      atomicMin(reinterpret_cast<unsigned long long*>(&found),
                (unsigned long long)cand);
    }
    __syncthreads();
    if (found != ~uint64_t{0}) break;
    if (threadIdx.x == 0) {
      double t = carry_s;
      for (int j = 0; j < kBS / 64; ++j) t += wave_tot[j];
      carry_s = t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = found;
}

// Exclusive FP64 prefix sums of `count` values with one workgroup:
// prefix[0] = 0, prefix[i+1] = prefix[i] + v[i].
__global__ void __launch_bounds__(1024)
ExclusiveScanKernel(const double* __restrict__ v, uint64_t count,
                    double* __restrict__ prefix) {
  __shared__ double wave_tot[16];
  __shared__ double carry_s;
  if (threadIdx.x == 0) {
    carry_s = 0;
    prefix[0] = 0;
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (uint64_t c0 = 0; c0 < count; c0 += 1024) {
    uint64_t i = c0 + threadIdx.x;
    double p = i < count ? v[i] : 0.0;
    double incl = WaveInclusiveScan(p);
    if (l == 63) wave_tot[w] = incl;
    __syncthreads();
    double before = carry_s;
    for (int j = 0; j < w; ++j) before += wave_tot[j];
    if (i < count) prefix[i + 1] = before + incl;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = carry_s;
      for (int j = 0; j < 16; ++j) t += wave_tot[j];
      carry_s = t;
    }
    __syncthreads();
  }
}

// One wave per draw: binary search of the sample-block prefix sums, then an
// in-order walk of the block with wave prefix sums (FP64).  A draw that the
// rounding of the block sums places past its block continues in the next
// block; a draw beyond the last amplitude maps to size - 1
// (lib/statespace_basic.h:227-229).
template <typename C>
__global__ void __launch_bounds__(64)
SampleKernel(const C* __restrict__ s, uint64_t size, uint64_t sblock,
             const double* __restrict__ prefix, uint64_t nblocks,
             const double* __restrict__ draws, uint64_t* __restrict__ out) {
  const uint64_t d = blockIdx.x;
  const double r = draws[d];
  const int l = threadIdx.x;
  // Largest b in [0, nblocks) with prefix[b] <= r.
  uint64_t lo = 0, hi = nblocks;  // invariant: prefix[lo] <= r or lo == 0
  while (hi - lo > 1) {
    uint64_t mid = (lo + hi) / 2;
    if (prefix[mid] <= r) lo = mid;
    else hi = mid;
  }
  double carry = prefix[lo];
  uint64_t result = size - 1;
  bool done = false;
  // The match lies in block lo, or in lo + 1 when rounding of the block sums
  // moved it; the walk never goes further (a NaN state ends here too).
  const uint64_t iend = std::min(size, (lo + 2) * sblock);
  for (uint64_t i0 = lo * sblock; i0 < iend && !done; i0 += 64) {
    uint64_t i = i0 + l;
    double p = 0;
    if (i < iend) {
      C v = s[i];
      p = double(v.x) * double(v.x) + double(v.y) * double(v.y);
    }
    double incl = WaveInclusiveScan(p);
    bool hit = i < iend && r < carry + incl;
    unsigned long long ballot = __ballot(hit);
    if (ballot != 0) {
      result = i0 + __builtin_ctzll(ballot);
      done = true;
    } else {
      carry += __shfl(incl, 63, 64);
    }
  }
  if (l == 0) out[d] = result;
}

template <typename FP>
using C_t = typename Cplx<FP>::T;

// Blocks of a reduction pass: at most kReduceBlocks partial sums.
constexpr unsigned kReduceBlocks = 1024;

// Partial sums (part + 1) and the final sum (part[0]) of a dot product
// into `part`, which holds at least GridFor(size, kReduceBlocks) + 1 Dbl2.
template <typename FP>
int ReduceInto(const C_t<FP>* a, const C_t<FP>* b, uint64_t size, bool real,
               Dbl2* part, double* re, double* im) {
  const unsigned blocks = GridFor(size, kReduceBlocks);
  if (real)
    hipLaunchKernelGGL((DotKernel<C_t<FP>, true>), dim3(blocks), dim3(kBS), 0,
                       Stream(), a, b, size, part + 1);
  else
    hipLaunchKernelGGL((DotKernel<C_t<FP>, false>), dim3(blocks), dim3(kBS), 0,
                       Stream(), a, b, size, part + 1);
  QHIP_LAUNCHED("DotKernel");
  hipLaunchKernelGGL((FinalSumKernel<Dbl2>), dim3(1), dim3(256), 0, Stream(),
                     part + 1, blocks, part);
  QHIP_LAUNCHED("FinalSumKernel");
  Dbl2 r;
  QHIP_CHECK(hipMemcpyAsync(&r, part, sizeof(Dbl2), hipMemcpyDeviceToHost,
                            Stream()));
  QHIP_CHECK(hipStreamSynchronize(Stream()));
  *re = r.x;
  if (im) *im = r.y;
  return QSIM_HIP_OK;
}

template <typename FP>
int Reduce(const C_t<FP>* a, const C_t<FP>* b, uint64_t size, bool real,
           double* re, double* im) {
  std::lock_guard<std::mutex> lock(ScratchLock());
  const unsigned blocks = GridFor(size, kReduceBlocks);
  Dbl2* part = static_cast<Dbl2*>(DeviceScratch(sizeof(Dbl2) * (blocks + 1)));
  if (part == nullptr) return QSIM_HIP_ERR_OOM;
  return ReduceInto<FP>(a, b, size, real, part, re, im);
}

// ---- fixed scratch of the no-allocation reductions -------------------------
//
// The partial sums of DotStatic live in this array (one per device, held
// until the process ends), as (re, im) pairs.  Its name is a host-side stand-in: host code
// uses only the device address hipGetSymbolAddress gives for it
// (g_static_scratch, taken once when the library is initialised).
constexpr unsigned kStaticScratchDbl2 = 2048;
static_assert(kStaticScratchDbl2 >= kReduceBlocks + 1, "reduction scratch");
static __device__ __align__(256) double scratch[2 * kStaticScratchDbl2];
Dbl2* g_static_scratch = nullptr;
std::mutex g_static_scratch_mu;

// Number of equal index ranges for partial norms.
unsigned NumParts(unsigned n) {
  uint64_t size = uint64_t{1} << n;
  uint64_t p = size / 1024;
  return unsigned(std::max<uint64_t>(1, std::min<uint64_t>(p, 1024)));
}

template <typename FP>
int MaskedNorm(const C_t<FP>* s, unsigned n, uint64_t mask, uint64_t bits,
               double* out) {
  std::lock_guard<std::mutex> lock(ScratchLock());
  unsigned parts = NumParts(n);
  uint64_t part = (uint64_t{1} << n) / parts;
  double* d = static_cast<double*>(DeviceScratch(sizeof(double) * (parts + 1)));
  if (d == nullptr) return QSIM_HIP_ERR_OOM;
  hipLaunchKernelGGL((NormPartsKernel<C_t<FP>, true>), dim3(parts), dim3(kBS),
                     0, Stream(), s, part, mask, bits, d + 1);
  QHIP_LAUNCHED("NormPartsKernel");
  hipLaunchKernelGGL((FinalSumKernel<double>), dim3(1), dim3(256), 0, Stream(),
                     d + 1, parts, d);
  QHIP_LAUNCHED("FinalSumKernel");
  QHIP_CHECK(hipMemcpyAsync(out, d, sizeof(double), hipMemcpyDeviceToHost,
                            Stream()));
  QHIP_CHECK(hipStreamSynchronize(Stream()));
  return QSIM_HIP_OK;
}

}  // namespace

// ---- entry points used by api.hip -----------------------------------------

template <typename FP>
int Fill(C_t<FP>* s, unsigned n, int what) {
  // what: 0 zeros, 1 |0>, 2 uniform
  uint64_t size = uint64_t{1} << n;
  C_t<FP> v{0, 0}, one{0, 0};
  uint64_t one_index = ~uint64_t{0};
  if (what == 1) {
    one = C_t<FP>{1, 0};
    one_index = 0;
  } else if (what == 2) {
    // lib/statespace_cuda.h:115-127: 1/sqrt(2^n) rounded to the element type.
    v = C_t<FP>{FP(1.0 / std::sqrt(double(size))), 0};
  }
  hipLaunchKernelGGL((FillKernel<C_t<FP>>), dim3(GridFor(size)), dim3(kBS), 0,
                     Stream(), s, size, v, one_index, one);
  QHIP_LAUNCHED("FillKernel");
  return QSIM_HIP_OK;
}

template <typename FP>
int BulkSet(C_t<FP>* s, unsigned n, uint64_t mask, uint64_t bits, double re,
            double im, bool exclude) {
  uint64_t size = uint64_t{1} << n;
  hipLaunchKernelGGL((BulkSetKernel<C_t<FP>, FP>), dim3(GridFor(size)),
                     dim3(kBS), 0, Stream(), s, size, mask, bits, FP(re),
                     FP(im), exclude);
  QHIP_LAUNCHED("BulkSetKernel");
  return QSIM_HIP_OK;
}

template <typename FP>
int Add(const C_t<FP>* src, C_t<FP>* dst, unsigned n) {
  uint64_t size = uint64_t{1} << n;
  hipLaunchKernelGGL((AddKernel<C_t<FP>>), dim3(GridFor(size)), dim3(kBS), 0,
                     Stream(), src, dst, size);
  QHIP_LAUNCHED("AddKernel");
  return QSIM_HIP_OK;
}

template <typename FP>
int Scale(C_t<FP>* s, unsigned n, double a) {
  uint64_t size = uint64_t{1} << n;
  hipLaunchKernelGGL((ScaleKernel<C_t<FP>, FP>), dim3(GridFor(size)),
                     dim3(kBS), 0, Stream(), s, size, FP(a));
  QHIP_LAUNCHED("ScaleKernel");
  return QSIM_HIP_OK;
}

template <typename FP>
int Dot(const C_t<FP>* a, const C_t<FP>* b, unsigned n, bool real, double* re,
        double* im) {
  return Reduce<FP>(a, b, uint64_t{1} << n, real, re, im);
}

template <typename FP>
int DotStatic(const C_t<FP>* a, const C_t<FP>* b, unsigned n, bool real,
              double* re, double* im) {
  const uint64_t size = uint64_t{1} << n;
  const unsigned blocks = GridFor(size, kReduceBlocks);
  if (g_static_scratch == nullptr) {
    SetError("no-allocation reduction: the fixed scratch was not set up "
             "(library not initialised)");
    return QSIM_HIP_ERR_RUNTIME;
  }
  if (size_t(blocks) + 1 > kStaticScratchDbl2) {
    SetError("no-allocation reduction: %u partial sums exceed the fixed "
             "scratch of %u", blocks + 1, kStaticScratchDbl2 - 1);
    return QSIM_HIP_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_static_scratch_mu);
  return ReduceInto<FP>(a, b, size, real, g_static_scratch, re, im);
}

int InitStaticScratch() {
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return HipFail(e, "hipGetSymbolAddress(scratch)");
  g_static_scratch = static_cast<Dbl2*>(p);
  return QSIM_HIP_OK;
}

unsigned PartialNormCount(unsigned n) { return NumParts(n); }

template <typename FP>
int PartialNorms(const C_t<FP>* s, unsigned n, double* out) {
  std::lock_guard<std::mutex> lock(ScratchLock());
  unsigned parts = NumParts(n);
  uint64_t part = (uint64_t{1} << n) / parts;
  double* d = static_cast<double*>(DeviceScratch(sizeof(double) * parts));
  if (d == nullptr) return QSIM_HIP_ERR_OOM;
  hipLaunchKernelGGL((NormPartsKernel<C_t<FP>, false>), dim3(parts), dim3(kBS),
                     0, Stream(), s, part, 0, 0, d);
  QHIP_LAUNCHED("NormPartsKernel");
  QHIP_CHECK(hipMemcpyAsync(out, d, sizeof(double) * parts,
                            hipMemcpyDeviceToHost, Stream()));
  QHIP_CHECK(hipStreamSynchronize(Stream()));
  return QSIM_HIP_OK;
}

template <typename FP>
int FindMeasuredBits(const C_t<FP>* s, unsigned n, unsigned m, double r,
                     uint64_t mask, uint64_t* out) {
  unsigned parts = NumParts(n);
  if (m >= parts) {
    SetError("partial-norm index %u out of range (%u)", m, parts);
    return QSIM_HIP_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(ScratchLock());
  uint64_t part = (uint64_t{1} << n) / parts;
  uint64_t* d = static_cast<uint64_t*>(DeviceScratch(sizeof(uint64_t)));
  if (d == nullptr) return QSIM_HIP_ERR_OOM;
  hipLaunchKernelGGL((FindFirstKernel<C_t<FP>>), dim3(1), dim3(kBS), 0,
                     Stream(), s, uint64_t(m) * part, uint64_t(m + 1) * part, r,
                     d);
  QHIP_LAUNCHED("FindFirstKernel");
  uint64_t k;
  QHIP_CHECK(hipMemcpyAsync(&k, d, sizeof(uint64_t), hipMemcpyDeviceToHost,
                            Stream()));
  QHIP_CHECK(hipStreamSynchronize(Stream()));
  // Past the end of the range (rounding): the last index of the range, as
  // lib/statespace_basic.h:291-294 returns.
  if (k == ~uint64_t{0}) k = uint64_t(m + 1) * part - 1;
  *out = k & mask;
  return QSIM_HIP_OK;
}

template <typename FP>
int Collapse(C_t<FP>* s, unsigned n, uint64_t mask, uint64_t bits) {
  double norm = 0;
  QHIP_TRY(MaskedNorm<FP>(s, n, mask, bits, &norm));
  FP renorm = FP(1.0 / std::sqrt(norm));
  uint64_t size = uint64_t{1} << n;
  hipLaunchKernelGGL((CollapseKernel<C_t<FP>, FP>), dim3(GridFor(size)),
                     dim3(kBS), 0, Stream(), s, size, mask, bits, renorm);
  QHIP_LAUNCHED("CollapseKernel");
  return QSIM_HIP_OK;
}

template <typename FP>
int Sample(const C_t<FP>* s, unsigned n, const double* draws, uint64_t num,
           uint64_t* out) {
  if (num == 0) return QSIM_HIP_OK;
  const uint64_t size = uint64_t{1} << n;
  const uint64_t sblock = std::min<uint64_t>(size, 4096);
  const uint64_t nblocks = size / sblock;
  std::lock_guard<std::mutex> lock(ScratchLock());
  size_t bytes = sizeof(double) * (nblocks + nblocks + 1) +
                 sizeof(double) * num + sizeof(uint64_t) * num;
  char* base = static_cast<char*>(DeviceScratch(bytes));
  if (base == nullptr) return QSIM_HIP_ERR_OOM;
  double* bsum = reinterpret_cast<double*>(base);
  double* prefix = bsum + nblocks;
  double* d_draws = prefix + nblocks + 1;
  uint64_t* d_out = reinterpret_cast<uint64_t*>(d_draws + num);

  hipLaunchKernelGGL((NormPartsKernel<C_t<FP>, false, true>),
                     dim3(unsigned(nblocks)), dim3(kBS), 0, Stream(), s,
                     sblock, 0, 0, bsum);
  QHIP_LAUNCHED("NormPartsKernel");
  hipLaunchKernelGGL(ExclusiveScanKernel, dim3(1), dim3(1024), 0, Stream(),
                     bsum, nblocks, prefix);
  QHIP_LAUNCHED("ExclusiveScanKernel");
  QHIP_CHECK(hipMemcpyAsync(d_draws, draws, sizeof(double) * num,
                            hipMemcpyHostToDevice, Stream()));
  hipLaunchKernelGGL((SampleKernel<C_t<FP>>), dim3(unsigned(num)), dim3(64), 0,
                     Stream(), s, size, sblock, prefix, nblocks, d_draws,
                     d_out);
  QHIP_LAUNCHED("SampleKernel");
  QHIP_CHECK(hipMemcpyAsync(out, d_out, sizeof(uint64_t) * num,
                            hipMemcpyDeviceToHost, Stream()));
  QHIP_CHECK(hipStreamSynchronize(Stream()));
  return QSIM_HIP_OK;
}

// out[j] = s[idx[j]]: the amplitudes of a batch of bitstrings in one pass.
template <typename C>
__global__ void __launch_bounds__(kBS)
GatherKernel(const C* __restrict__ s, const uint64_t* __restrict__ idx,
             uint64_t count, C* __restrict__ out) {
  const uint64_t stride = uint64_t(gridDim.x) * kBS;
  for (uint64_t j = uint64_t(blockIdx.x) * kBS + threadIdx.x; j < count;
       j += stride) {
    out[j] = s[idx[j]];
  }
}

// Amplitudes of `count` indices (host arrays; indices checked by the caller)
// in chunks that fit the scratch buffer: one upload, one gather launch and
// one download per chunk.
template <typename FP>
int Gather(const C_t<FP>* s, const uint64_t* idx, uint64_t count,
           C_t<FP>* out) {
  constexpr uint64_t kChunk = uint64_t{1} << 20;
  std::lock_guard<std::mutex> lock(ScratchLock());
  const uint64_t cap = std::min(count, kChunk);
  char* base = static_cast<char*>(
      DeviceScratch(cap * (sizeof(uint64_t) + sizeof(C_t<FP>))));
  if (base == nullptr) return QSIM_HIP_ERR_OOM;
  uint64_t* d_idx = reinterpret_cast<uint64_t*>(base);
  C_t<FP>* d_out = reinterpret_cast<C_t<FP>*>(d_idx + cap);
  for (uint64_t j0 = 0; j0 < count; j0 += cap) {
    const uint64_t m = std::min(cap, count - j0);
    QHIP_CHECK(hipMemcpyAsync(d_idx, idx + j0, sizeof(uint64_t) * m,
                              hipMemcpyHostToDevice, Stream()));
    hipLaunchKernelGGL((GatherKernel<C_t<FP>>), dim3(GridFor(m)), dim3(kBS), 0,
                       Stream(), s, d_idx, m, d_out);
    QHIP_LAUNCHED("GatherKernel");
    QHIP_CHECK(hipMemcpyAsync(out + j0, d_out, sizeof(C_t<FP>) * m,
                              hipMemcpyDeviceToHost, Stream()));
    QHIP_CHECK(hipStreamSynchronize(Stream()));
  }
  return QSIM_HIP_OK;
}

#define QHIP_INSTANTIATE(FP)                                                   \
  template int Fill<FP>(C_t<FP>*, unsigned, int);                              \
  template int BulkSet<FP>(C_t<FP>*, unsigned, uint64_t, uint64_t, double,     \
                           double, bool);                                      \
  template int Add<FP>(const C_t<FP>*, C_t<FP>*, unsigned);                    \
  template int Scale<FP>(C_t<FP>*, unsigned, double);                          \
  template int Dot<FP>(const C_t<FP>*, const C_t<FP>*, unsigned, bool,         \
                       double*, double*);                                      \
  template int DotStatic<FP>(const C_t<FP>*, const C_t<FP>*, unsigned, bool,   \
                             double*, double*);                                \
  template int PartialNorms<FP>(const C_t<FP>*, unsigned, double*);            \
  template int FindMeasuredBits<FP>(const C_t<FP>*, unsigned, unsigned,        \
                                    double, uint64_t, uint64_t*);              \
  template int Collapse<FP>(C_t<FP>*, unsigned, uint64_t, uint64_t);           \
  template int Sample<FP>(const C_t<FP>*, unsigned, const double*, uint64_t,   \
                          uint64_t*);                                          \
  template int Gather<FP>(const C_t<FP>*, const uint64_t*, uint64_t, C_t<FP>*);

template <typename D, typename S>
__global__ void __launch_bounds__(kBS)
ConvertKernel(D* __restrict__ dst, const S* __restrict__ src, uint64_t size) {
  const uint64_t stride = uint64_t(gridDim.x) * kBS;
  for (uint64_t i = uint64_t(blockIdx.x) * kBS + threadIdx.x; i < size;
       i += stride) {
    S v = src[i];
    dst[i] = D{decltype(dst->x)(v.x), decltype(dst->x)(v.y)};
  }
}

int Convert(void* dst, int dprec, const void* src, int sprec, unsigned n) {
  const uint64_t size = uint64_t{1} << n;
  const dim3 grid(GridFor(size));
  if (dprec == QSIM_HIP_F64 && sprec == QSIM_HIP_F32) {
    hipLaunchKernelGGL((ConvertKernel<double2, float2>), grid, dim3(kBS), 0,
                       Stream(), (double2*)dst, (const float2*)src, size);
  } else if (dprec == QSIM_HIP_F32 && sprec == QSIM_HIP_F64) {
    hipLaunchKernelGGL((ConvertKernel<float2, double2>), grid, dim3(kBS), 0,
                       Stream(), (float2*)dst, (const double2*)src, size);
  } else if (dprec == QSIM_HIP_F32) {
    hipLaunchKernelGGL((ConvertKernel<float2, float2>), grid, dim3(kBS), 0,
                       Stream(), (float2*)dst, (const float2*)src, size);
  } else {
    hipLaunchKernelGGL((ConvertKernel<double2, double2>), grid, dim3(kBS), 0,
                       Stream(), (double2*)dst, (const double2*)src, size);
  }
  QHIP_LAUNCHED("ConvertKernel");
  return QSIM_HIP_OK;
}

QHIP_INSTANTIATE(float)
QHIP_INSTANTIATE(double)

}  // namespace qhip
