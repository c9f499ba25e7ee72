// api.hip -- the C ABI of include/qsim_hip.h: context, memory, dispatch.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "internal.h"

namespace qhip {

namespace {

thread_local char g_error[512] = "";

struct Device {
  bool ready = false;
  int id = 0;
  int cus = 0;
  void* scratch = nullptr;
  size_t scratch_size = 0;
  // Matrix ring: device slots and pinned host staging, one event per slot.
  static constexpr int kSlots = 64;
  // A k = 6 FP64 matrix or BF16 fragments, or the fragments of a full
  // cache-blocked program (32 op records of 8 KiB).
  static constexpr size_t kSlotBytes = 256 * 1024;
  char* mat_dev = nullptr;
  char* mat_host = nullptr;
  hipEvent_t mat_ev[kSlots] = {};
  bool mat_ev_used[kSlots] = {};
  int mat_next = 0;
  hipEvent_t ev[64] = {};
};

std::mutex g_init_mu;
std::mutex g_scratch_mu;
std::mutex g_mat_mu;
Device g_dev;
hipStream_t g_stream = nullptr;
std::atomic<int> g_method{QSIM_HIP_METHOD_AUTO};

struct PassTimer {
  bool on = false;
  unsigned capacity = 0;
  unsigned count = 0;
  std::vector<hipEvent_t> ev;
};
PassTimer g_timer;

// A recorded and captured sequence of passes (qsim_hip_graph_*).  While
// recording, every matrix the passes stage gets a device buffer of its own
// (the plan cache) instead of a slot of the shared ring; while capturing,
// the same sequence takes those buffers back in order and stages nothing,
// so the graph holds only kernel launches on device-resident data.
struct GraphPlan {
  std::vector<void*> bufs;
  std::vector<size_t> sizes;
  size_t next = 0;
  hipStream_t stream = nullptr;
  hipStream_t saved = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool mismatch = false;
};
enum PlanMode { kPlanOff, kPlanRecord, kPlanCapture };
PlanMode g_plan_mode = kPlanOff;
GraphPlan* g_plan = nullptr;

}  // namespace

void SetError(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

int HipFail(hipError_t e, const char* what) {
  SetError("%s: %s", what, hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? QSIM_HIP_ERR_OOM : QSIM_HIP_ERR_RUNTIME;
}

static int InitDevice(int device) {
  std::lock_guard<std::mutex> lock(g_init_mu);
  if (g_dev.ready) {
    if (device != g_dev.id) {
      SetError("library already initialised on device %d", g_dev.id);
      return QSIM_HIP_ERR_ARG;
    }
    return QSIM_HIP_OK;
  }
  int count = 0;
  QHIP_CHECK(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) {
    SetError("device %d not present (%d devices)", device, count);
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_CHECK(hipSetDevice(device));
  hipDeviceProp_t prop;
  QHIP_CHECK(hipGetDeviceProperties(&prop, device));
  g_dev.id = device;
  g_dev.cus = prop.multiProcessorCount;
  QHIP_CHECK(hipMalloc(&g_dev.mat_dev, Device::kSlots * Device::kSlotBytes));
  QHIP_CHECK(hipHostMalloc(&g_dev.mat_host, Device::kSlots * Device::kSlotBytes,
                           hipHostMallocDefault));
  for (int i = 0; i < Device::kSlots; ++i) {
    QHIP_CHECK(hipEventCreateWithFlags(&g_dev.mat_ev[i],
                                       hipEventDisableTiming));
  }
  for (int i = 0; i < 64; ++i) QHIP_CHECK(hipEventCreate(&g_dev.ev[i]));
  QHIP_TRY(InitStaticScratch());
  g_dev.ready = true;
  return QSIM_HIP_OK;
}

int EnsureInit() {
  if (g_dev.ready) return QSIM_HIP_OK;
  return InitDevice(0);
}

hipStream_t Stream() { return g_stream; }

int NumCUs() { return g_dev.cus; }

std::mutex& ScratchLock() { return g_scratch_mu; }

void* DeviceScratch(size_t bytes) {
  if (bytes <= g_dev.scratch_size) return g_dev.scratch;
  if (g_dev.scratch != nullptr) {
    // Earlier work may still read the old buffer.
    if (hipStreamSynchronize(g_stream) != hipSuccess ||
        hipFree(g_dev.scratch) != hipSuccess) {
      SetError("cannot release scratch buffer");
      return nullptr;
    }
    g_dev.scratch = nullptr;
    g_dev.scratch_size = 0;
  }
  size_t sz = std::max<size_t>(bytes, 1 << 20);
  hipError_t e = hipMalloc(&g_dev.scratch, sz);
  if (e != hipSuccess) {
    HipFail(e, "hipMalloc(scratch)");
    g_dev.scratch = nullptr;
    return nullptr;
  }
  g_dev.scratch_size = sz;
  return g_dev.scratch;
}

void* MatrixBuffer(size_t bytes) {
  if (bytes > Device::kSlotBytes) {
    SetError("matrix of %zu bytes exceeds the staging slot", bytes);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_mat_mu);
  if (g_plan_mode == kPlanRecord) {
    void* d = nullptr;
    hipError_t e = hipMalloc(&d, bytes);
    if (e != hipSuccess) {
      HipFail(e, "hipMalloc(plan matrix)");
      return nullptr;
    }
    g_plan->bufs.push_back(d);
    g_plan->sizes.push_back(bytes);
    return d;
  }
  if (g_plan_mode == kPlanCapture) {
    GraphPlan& pl = *g_plan;
    if (pl.next >= pl.bufs.size() || pl.sizes[pl.next] != bytes) {
      pl.mismatch = true;
      SetError("captured passes differ from the recorded ones");
      return nullptr;
    }
    return pl.bufs[pl.next++];
  }
  int slot = g_dev.mat_next;
  g_dev.mat_next = (slot + 1) % Device::kSlots;
  return g_dev.mat_dev + size_t(slot) * Device::kSlotBytes;
}

void PassTimerBegin() {
  if (g_plan_mode == kPlanCapture) return;
  if (g_timer.on && g_timer.count < g_timer.capacity)
    (void)hipEventRecord(g_timer.ev[2 * g_timer.count], g_stream);
}

void PassTimerEnd() {
  if (g_plan_mode == kPlanCapture) return;
  if (g_timer.on && g_timer.count < g_timer.capacity) {
    (void)hipEventRecord(g_timer.ev[2 * g_timer.count + 1], g_stream);
    ++g_timer.count;
  }
}

int StageMatrix(void* dev, const void* host, size_t bytes) {
  std::lock_guard<std::mutex> lock(g_mat_mu);
  // Plan buffers are fresh: a blocking copy while recording (the host
  // matrix may be a temporary), nothing while capturing (the recorded
  // contents are the captured pass's).
  if (g_plan_mode == kPlanRecord) {
    QHIP_CHECK(hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice));
    return QSIM_HIP_OK;
  }
  if (g_plan_mode == kPlanCapture) return QSIM_HIP_OK;
  size_t slot = (static_cast<char*>(dev) - g_dev.mat_dev) / Device::kSlotBytes;
  char* h = g_dev.mat_host + slot * Device::kSlotBytes;
  // The pinned slot may still be the source of an earlier, unfinished copy.
  if (g_dev.mat_ev_used[slot])
    QHIP_CHECK(hipEventSynchronize(g_dev.mat_ev[slot]));
  std::memcpy(h, host, bytes);
  QHIP_CHECK(hipMemcpyAsync(dev, h, bytes, hipMemcpyHostToDevice, g_stream));
  QHIP_CHECK(hipEventRecord(g_dev.mat_ev[slot], g_stream));
  g_dev.mat_ev_used[slot] = true;
  return QSIM_HIP_OK;
}

}  // namespace qhip

using namespace qhip;

namespace {

int CheckStateArgs(int prec, const void* s, unsigned n) {
  if (prec != QSIM_HIP_F32 && prec != QSIM_HIP_F64) {
    SetError("unknown precision %d", prec);
    return QSIM_HIP_ERR_ARG;
  }
  if (s == nullptr) {
    SetError("null state");
    return QSIM_HIP_ERR_ARG;
  }
  if (n > 40) {
    SetError("%u qubits exceed the supported 40", n);
    return QSIM_HIP_ERR_ARG;
  }
  return EnsureInit();
}

// Every entry point that touches a state (or orders work on the stream)
// first runs the gates the queue holds.
int CheckState(int prec, const void* s, unsigned n) {
  QHIP_TRY(CheckStateArgs(prec, s, n));
  return FlushGateQueue();
}

uint64_t ElemBytes(int prec) { return prec == QSIM_HIP_F32 ? 4 : 8; }

// Offsets and FP64 copies of a blocked pass's concatenated gate matrices.
int BlockMatrices(unsigned num_gates, const unsigned* ks, const void* matrices,
                  int matrix_prec, std::vector<size_t>& moff,
                  std::vector<double>& m64) {
  if (num_gates > kBlockMaxOps) {
    SetError("a blocked pass holds at most %u gates, got %u", kBlockMaxOps,
             num_gates);
    return QSIM_HIP_ERR_ARG;
  }
  if (matrix_prec != QSIM_HIP_F32 && matrix_prec != QSIM_HIP_F64) {
    SetError("blocked-pass matrices must be QSIM_HIP_F32 or QSIM_HIP_F64");
    return QSIM_HIP_ERR_ARG;
  }
  moff.resize(num_gates);
  size_t total = 0;
  for (unsigned i = 0; i < num_gates; ++i) {
    if (ks[i] < 1 || ks[i] > 4) {
      SetError("blocked-pass gate %u acts on %u qubits (1..4 supported)", i,
               ks[i]);
      return QSIM_HIP_ERR_ARG;
    }
    moff[i] = total;
    total += size_t{2} << (2 * ks[i]);
  }
  m64.resize(total);
  if (matrix_prec == QSIM_HIP_F64) {
    std::memcpy(m64.data(), matrices, total * sizeof(double));
  } else {
    const float* mf = static_cast<const float*>(matrices);
    for (size_t i = 0; i < total; ++i) m64[i] = mf[i];
  }
  return QSIM_HIP_OK;
}

}  // namespace

extern "C" {

const char* qsim_hip_last_error(void) { return g_error; }

int qsim_hip_init(int device) { return InitDevice(device); }

int qsim_hip_device_info(char* name, size_t name_len, uint64_t* total_mem,
                         int* num_cus) {
  QHIP_TRY(EnsureInit());
  hipDeviceProp_t prop;
  QHIP_CHECK(hipGetDeviceProperties(&prop, g_dev.id));
  if (name && name_len) {
    std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (total_mem) *total_mem = prop.totalGlobalMem;
  if (num_cus) *num_cus = prop.multiProcessorCount;
  return QSIM_HIP_OK;
}

int qsim_hip_set_stream(void* stream) {
  QHIP_TRY(EnsureInit());
  QHIP_TRY(FlushGateQueue());
  g_stream = static_cast<hipStream_t>(stream);
  return QSIM_HIP_OK;
}

int qsim_hip_synchronize(void) {
  QHIP_TRY(EnsureInit());
  QHIP_TRY(FlushGateQueue());
  QHIP_CHECK(hipStreamSynchronize(g_stream));
  return QSIM_HIP_OK;
}

uint64_t qsim_hip_min_size(unsigned num_qubits) {
  return std::max<uint64_t>(64, uint64_t{2} << num_qubits);
}

int qsim_hip_alloc(int prec, unsigned num_qubits, void** out) {
  if (out == nullptr) return QSIM_HIP_ERR_ARG;
  *out = nullptr;
  if (num_qubits > 40) {
    SetError("%u qubits exceed the supported 40", num_qubits);
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_TRY(EnsureInit());
  uint64_t bytes = ElemBytes(prec) * qsim_hip_min_size(num_qubits);
  hipError_t e = hipMalloc(out, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    SetError("hipMalloc(%llu bytes): %s", (unsigned long long)bytes,
             hipGetErrorString(e));
    return QSIM_HIP_ERR_OOM;
  }
  return QSIM_HIP_OK;
}

int qsim_hip_free(void* ptr) {
  if (ptr == nullptr) return QSIM_HIP_OK;
  QHIP_TRY(EnsureInit());
  QHIP_TRY(FlushGateQueue());
  QHIP_CHECK(hipStreamSynchronize(g_stream));
  QHIP_CHECK(hipFree(ptr));
  return QSIM_HIP_OK;
}

static int Copy(void* dst, const void* src, uint64_t bytes, hipMemcpyKind k) {
  QHIP_TRY(EnsureInit());
  QHIP_TRY(FlushGateQueue());
  if (bytes == 0) return QSIM_HIP_OK;
  if (dst == nullptr || src == nullptr) {
    SetError("null copy operand");
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_CHECK(hipMemcpyAsync(dst, src, bytes, k, g_stream));
  QHIP_CHECK(hipStreamSynchronize(g_stream));
  return QSIM_HIP_OK;
}

int qsim_hip_copy_d2d(void* dst, const void* src, uint64_t bytes) {
  return Copy(dst, src, bytes, hipMemcpyDeviceToDevice);
}
int qsim_hip_copy_d2h(void* dst, const void* src, uint64_t bytes) {
  return Copy(dst, src, bytes, hipMemcpyDeviceToHost);
}
int qsim_hip_copy_h2d(void* dst, const void* src, uint64_t bytes) {
  return Copy(dst, src, bytes, hipMemcpyHostToDevice);
}

int qsim_hip_copy_convert(void* dst, int dst_prec, const void* src,
                          int src_prec, unsigned n) {
  QHIP_TRY(CheckState(dst_prec, dst, n));
  QHIP_TRY(CheckState(src_prec, src, n));
  return Convert(dst, dst_prec, src, src_prec, n);
}

#define QHIP_DISPATCH(prec, F32CALL, F64CALL) \
  ((prec) == QSIM_HIP_F32 ? (F32CALL) : (F64CALL))

int qsim_hip_set_all_zeros(int prec, void* s, unsigned n) {
  QHIP_TRY(CheckState(prec, s, n));
  return QHIP_DISPATCH(prec, Fill<float>((float2*)s, n, 0),
                       Fill<double>((double2*)s, n, 0));
}

int qsim_hip_set_state_zero(int prec, void* s, unsigned n) {
  QHIP_TRY(CheckState(prec, s, n));
  return QHIP_DISPATCH(prec, Fill<float>((float2*)s, n, 1),
                       Fill<double>((double2*)s, n, 1));
}

int qsim_hip_set_state_uniform(int prec, void* s, unsigned n) {
  QHIP_TRY(CheckState(prec, s, n));
  return QHIP_DISPATCH(prec, Fill<float>((float2*)s, n, 2),
                       Fill<double>((double2*)s, n, 2));
}

int qsim_hip_internal_to_normal_order(int prec, void* s, unsigned n) {
  return CheckState(prec, s, n);
}

int qsim_hip_normal_to_internal_order(int prec, void* s, unsigned n) {
  return CheckState(prec, s, n);
}

int qsim_hip_get_ampl(int prec, const void* s, unsigned n, uint64_t i,
                      double* re, double* im) {
  QHIP_TRY(CheckState(prec, s, n));
  if (i >> n) {
    SetError("amplitude index out of range");
    return QSIM_HIP_ERR_ARG;
  }
  // A blocking copy on the legacy stream: safe from concurrent host threads.
  if (prec == QSIM_HIP_F32) {
    float2 v;
    QHIP_CHECK(hipStreamSynchronize(g_stream));
    QHIP_CHECK(hipMemcpy(&v, (const float2*)s + i, sizeof(v),
                         hipMemcpyDeviceToHost));
    *re = v.x;
    *im = v.y;
  } else {
    double2 v;
    QHIP_CHECK(hipStreamSynchronize(g_stream));
    QHIP_CHECK(hipMemcpy(&v, (const double2*)s + i, sizeof(v),
                         hipMemcpyDeviceToHost));
    *re = v.x;
    *im = v.y;
  }
  return QSIM_HIP_OK;
}

int qsim_hip_get_ampls(int prec, const void* s, unsigned n,
                       const uint64_t* indices, uint64_t count, void* out) {
  QHIP_TRY(CheckState(prec, s, n));
  if (count == 0) return QSIM_HIP_OK;
  if (indices == nullptr || out == nullptr) {
    SetError("null amplitude index or output array");
    return QSIM_HIP_ERR_ARG;
  }
  for (uint64_t j = 0; j < count; ++j) {
    if (indices[j] >> n) {
      SetError("amplitude index %llu out of range for %u qubits",
               (unsigned long long)indices[j], n);
      return QSIM_HIP_ERR_ARG;
    }
  }
  return QHIP_DISPATCH(
      prec, Gather<float>((const float2*)s, indices, count, (float2*)out),
      Gather<double>((const double2*)s, indices, count, (double2*)out));
}

int qsim_hip_set_ampl(int prec, void* s, unsigned n, uint64_t i, double re,
                      double im) {
  QHIP_TRY(CheckState(prec, s, n));
  if (i >> n) {
    SetError("amplitude index out of range");
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_CHECK(hipStreamSynchronize(g_stream));
  if (prec == QSIM_HIP_F32) {
    float2 v{float(re), float(im)};
    QHIP_CHECK(hipMemcpy((float2*)s + i, &v, sizeof(v), hipMemcpyHostToDevice));
  } else {
    double2 v{re, im};
    QHIP_CHECK(
        hipMemcpy((double2*)s + i, &v, sizeof(v), hipMemcpyHostToDevice));
  }
  return QSIM_HIP_OK;
}

int qsim_hip_bulk_set_ampl(int prec, void* s, unsigned n, uint64_t mask,
                           uint64_t bits, double re, double im, int exclude) {
  QHIP_TRY(CheckState(prec, s, n));
  return QHIP_DISPATCH(
      prec, BulkSet<float>((float2*)s, n, mask, bits, re, im, exclude != 0),
      BulkSet<double>((double2*)s, n, mask, bits, re, im, exclude != 0));
}

int qsim_hip_add(int prec, const void* src, void* dst, unsigned n) {
  QHIP_TRY(CheckState(prec, src, n));
  QHIP_TRY(CheckState(prec, dst, n));
  return QHIP_DISPATCH(prec, Add<float>((const float2*)src, (float2*)dst, n),
                       Add<double>((const double2*)src, (double2*)dst, n));
}

int qsim_hip_multiply(int prec, void* s, unsigned n, double a) {
  QHIP_TRY(CheckState(prec, s, n));
  return QHIP_DISPATCH(prec, Scale<float>((float2*)s, n, a),
                       Scale<double>((double2*)s, n, a));
}

int qsim_hip_inner_product(int prec, const void* a, const void* b, unsigned n,
                           double* re, double* im) {
  QHIP_TRY(CheckState(prec, a, n));
  QHIP_TRY(CheckState(prec, b, n));
  return QHIP_DISPATCH(
      prec, Dot<float>((const float2*)a, (const float2*)b, n, false, re, im),
      Dot<double>((const double2*)a, (const double2*)b, n, false, re, im));
}

int qsim_hip_real_inner_product(int prec, const void* a, const void* b,
                                unsigned n, double* out) {
  QHIP_TRY(CheckState(prec, a, n));
  QHIP_TRY(CheckState(prec, b, n));
  return QHIP_DISPATCH(
      prec, Dot<float>((const float2*)a, (const float2*)b, n, true, out, nullptr),
      Dot<double>((const double2*)a, (const double2*)b, n, true, out, nullptr));
}

int qsim_hip_norm(int prec, const void* s, unsigned n, double* out) {
  return qsim_hip_real_inner_product(prec, s, s, n, out);
}

int qsim_hip_inner_product_noalloc(int prec, const void* a, const void* b,
                                   unsigned n, double* re, double* im) {
  QHIP_TRY(CheckStateArgs(prec, a, n));
  QHIP_TRY(CheckStateArgs(prec, b, n));
  if (re == nullptr) {
    SetError("null result pointer");
    return QSIM_HIP_ERR_ARG;
  }
  // Running queued gates could stage a blocked pass; this entry point
  // leaves that to the caller (qsim_hip_flush) instead.
  unsigned tb = 0;
  uint64_t queued = 0;
  QHIP_TRY(QueuedGates(&tb, &queued));
  if (queued != 0) {
    SetError("no-allocation inner product: the gate queue holds %llu gates; "
             "call qsim_hip_flush first", (unsigned long long)queued);
    return QSIM_HIP_ERR_ARG;
  }
  return QHIP_DISPATCH(
      prec,
      DotStatic<float>((const float2*)a, (const float2*)b, n, im == nullptr,
                       re, im),
      DotStatic<double>((const double2*)a, (const double2*)b, n,
                        im == nullptr, re, im));
}

int qsim_hip_sample(int prec, const void* s, unsigned n, const double* draws,
                    uint64_t num, uint64_t* out) {
  QHIP_TRY(CheckState(prec, s, n));
  if (num > 0 && (draws == nullptr || out == nullptr)) {
    SetError("null sample buffers");
    return QSIM_HIP_ERR_ARG;
  }
  if (num > (uint64_t{1} << 31)) {
    SetError("too many samples");
    return QSIM_HIP_ERR_ARG;
  }
  return QHIP_DISPATCH(
      prec, Sample<float>((const float2*)s, n, draws, num, out),
      Sample<double>((const double2*)s, n, draws, num, out));
}

int qsim_hip_sample_seeded(int prec, const void* s, unsigned n, uint64_t num,
                           unsigned seed, uint64_t* out) {
  return GuardHost([&]() -> int {
    if (num == 0) return CheckState(prec, s, n);
    double norm = 0;
    QHIP_TRY(qsim_hip_norm(prec, s, n, &norm));
    // The draws of lib/util.h:67-85 (mt19937, uniform_real_distribution, sort).
    std::vector<double> rs;
    rs.reserve(num + 1);
    std::mt19937 rgen(seed);
    std::uniform_real_distribution<double> distr(0.0, norm);
    for (uint64_t i = 0; i < num; ++i) rs.push_back(distr(rgen));
    std::sort(rs.begin(), rs.end());
    rs.push_back(norm);
    return qsim_hip_sample(prec, s, n, rs.data(), num, out);
  });
}

unsigned qsim_hip_num_partial_norms(unsigned n) { return PartialNormCount(n); }

int qsim_hip_partial_norms(int prec, const void* s, unsigned n, double* out) {
  QHIP_TRY(CheckState(prec, s, n));
  return QHIP_DISPATCH(prec, PartialNorms<float>((const float2*)s, n, out),
                       PartialNorms<double>((const double2*)s, n, out));
}

int qsim_hip_find_measured_bits(int prec, const void* s, unsigned n,
                                unsigned m, double r, uint64_t mask,
                                uint64_t* out) {
  QHIP_TRY(CheckState(prec, s, n));
  return QHIP_DISPATCH(
      prec, FindMeasuredBits<float>((const float2*)s, n, m, r, mask, out),
      FindMeasuredBits<double>((const double2*)s, n, m, r, mask, out));
}

int qsim_hip_collapse(int prec, void* s, unsigned n, uint64_t mask,
                      uint64_t bits) {
  return GuardHost([&]() -> int {
    QHIP_TRY(CheckState(prec, s, n));
    return QHIP_DISPATCH(prec, Collapse<float>((float2*)s, n, mask, bits),
                         Collapse<double>((double2*)s, n, mask, bits));
  });
}

// One gate pass now (no queue).
static int ApplyGateNow(int prec, void* s, unsigned n, const unsigned* qs,
                        unsigned k, const unsigned* cqs, unsigned nc,
                        uint64_t cvals, const void* matrix, int matrix_prec) {
  if ((k > 0 && qs == nullptr) || (nc > 0 && cqs == nullptr) ||
      matrix == nullptr) {
    SetError("null gate operand");
    return QSIM_HIP_ERR_ARG;
  }
  if (k > 6) {
    SetError("gates on %u > 6 qubits are not supported", k);
    return QSIM_HIP_ERR_UNSUPPORTED;
  }
  const size_t melems = size_t{2} << (2 * k);
  int method = g_method.load();
  if (prec == QSIM_HIP_F32 && nc == 0 && k >= 4 &&
      method != QSIM_HIP_METHOD_VECTOR) {
    std::vector<double> m64(melems);
    if (matrix_prec == QSIM_HIP_F64) {
      std::memcpy(m64.data(), matrix, melems * sizeof(double));
    } else {
      const float* mf = static_cast<const float*>(matrix);
      for (size_t i = 0; i < melems; ++i) m64[i] = mf[i];
    }
    int rc = LaunchGateMfma((float2*)s, n, qs, k, m64.data(), method);
    if (rc != QSIM_HIP_ERR_UNSUPPORTED) return rc;
  }
  if (prec == QSIM_HIP_F32) {
    std::vector<float> mf;
    const float* m = static_cast<const float*>(matrix);
    if (matrix_prec == QSIM_HIP_F64) {
      mf.resize(melems);
      const double* md = static_cast<const double*>(matrix);
      for (size_t i = 0; i < melems; ++i) mf[i] = float(md[i]);
      m = mf.data();
    }
    return LaunchGate<float>((float2*)s, n, qs, k, cqs, nc, cvals, m);
  } else {
    std::vector<double> md;
    const double* m = static_cast<const double*>(matrix);
    if (matrix_prec == QSIM_HIP_F32) {
      md.resize(melems);
      const float* mf = static_cast<const float*>(matrix);
      for (size_t i = 0; i < melems; ++i) md[i] = mf[i];
      m = md.data();
    }
    return LaunchGate<double>((double2*)s, n, qs, k, cqs, nc, cvals, m);
  }
}

int qsim_hip_apply_controlled_gate(int prec, void* s, unsigned n,
                                   const unsigned* qs, unsigned k,
                                   const unsigned* cqs, unsigned nc,
                                   uint64_t cvals, const void* matrix,
                                   int matrix_prec) {
  return GuardHost([&]() -> int {
    // Without controls this is a plain gate, which the gate queue may take.
    if (nc == 0) {
      return qsim_hip_apply_gate(prec, s, n, qs, k, matrix, matrix_prec);
    }
    QHIP_TRY(CheckState(prec, s, n));
    return ApplyGateNow(prec, s, n, qs, k, cqs, nc, cvals, matrix, matrix_prec);
  });
}

}  // extern "C"

namespace qhip {

// ---- gate queue ------------------------------------------------------------
//
// While on, uncontrolled complex64 gates on 1..4 qubits are collected into
// cache-blocked passes instead of running one pass each: consecutive gates
// on one state whose qubits, with the low_bits lowest qubits, fit in
// tile_bits qubits (and whose number fits the pass kernel's resident op
// records) run as one blocked pass when the next gate does not fit, or at
// the first other call of the library.  Grouping consecutive gates
// greedily this way gives the fewest passes any grouping into consecutive
// runs could (a run that fits stays fitting when shortened), so a fuser
// that emits its gates pass by pass gets its passes back.
struct GateQueue {
  unsigned tb = 0, lc = 0;  // tb == 0: off
  void* state = nullptr;
  unsigned n = 0;
  uint64_t mask = 0;        // qubits of the queued gates and the low run
  std::vector<unsigned> ks, qs;
  std::vector<double> m;
  std::vector<size_t> moff;
  uint64_t passes = 0, gates = 0;  // blocked passes run and gates in them
};
static std::mutex g_queue_mu;
static GateQueue g_queue;

static int FlushLocked() {
  GateQueue& q = g_queue;
  const unsigned ng = unsigned(q.ks.size());
  if (ng == 0) return QSIM_HIP_OK;
  int rc = QSIM_HIP_OK;
  if (ng == 1 || q.n < kBlockMinBits) {
    for (unsigned i = 0; i < ng && rc == QSIM_HIP_OK; ++i) {
      rc = ::ApplyGateNow(QSIM_HIP_F32, q.state, q.n, &q.qs[4 * i], q.ks[i],
                        nullptr, 0, 0, &q.m[q.moff[i]], QSIM_HIP_F64);
    }
  } else {
    // The tile: the queued gates' qubits and the low run, filled up to
    // tile_bits with the lowest other qubits (extending the contiguous low
    // run where it can).
    uint64_t mask = q.mask;
    const unsigned want = std::min(q.tb, q.n);
    for (unsigned b = 0; b < q.n && unsigned(__builtin_popcountll(mask)) < want;
         ++b) {
      mask |= uint64_t{1} << b;
    }
    std::vector<unsigned> tile;
    for (unsigned b = 0; b < q.n; ++b)
      if ((mask >> b) & 1) tile.push_back(b);
    rc = LaunchGateBlock(static_cast<float2*>(q.state), q.n, tile.data(),
                         unsigned(tile.size()), ng, q.ks.data(), q.qs.data(),
                         q.m.data(), q.moff.data(), false, kBlockArithF16);
    q.passes += 1;
    q.gates += ng;
  }
  q.ks.clear();
  q.qs.clear();
  q.m.clear();
  q.moff.clear();
  q.mask = 0;
  q.state = nullptr;
  return rc;
}

int FlushGateQueue() {
  std::lock_guard<std::mutex> lock(g_queue_mu);
  return FlushLocked();
}

int QueuedGates(unsigned* tile_bits, uint64_t* gates) {
  std::lock_guard<std::mutex> lock(g_queue_mu);
  *tile_bits = g_queue.tb;
  *gates = g_queue.ks.size();
  return QSIM_HIP_OK;
}

// Queues the gate if the queue is on and takes it; *queued says whether it
// did.
static int Enqueue(int prec, void* s, unsigned n, const unsigned* qs,
                   unsigned k, const void* matrix, int matrix_prec,
                   bool* queued) {
  *queued = false;
  std::lock_guard<std::mutex> lock(g_queue_mu);
  GateQueue& q = g_queue;
  if (q.tb == 0 || prec != QSIM_HIP_F32 || k < 1 || k > 4 ||
      (matrix_prec != QSIM_HIP_F32 && matrix_prec != QSIM_HIP_F64) ||
      qs == nullptr || matrix == nullptr || n < kBlockMinBits) {
    return FlushLocked();
  }
  uint64_t gm = 0;
  for (unsigned i = 0; i < k; ++i) {
    if (qs[i] >= n || (i > 0 && qs[i] <= qs[i - 1])) return FlushLocked();
    gm |= uint64_t{1} << qs[i];
  }
  const uint64_t low = (uint64_t{1} << std::min(q.lc, n)) - 1;
  const uint64_t mask = q.mask | gm | low;
  const bool fits = q.state == s && q.n == n &&
                    unsigned(__builtin_popcountll(mask)) <= std::min(q.tb, n) &&
                    q.ks.size() < BlockResidentOps(q.tb);
  if (!fits) {
    QHIP_TRY(FlushLocked());
    q.state = s;
    q.n = n;
    q.mask = gm | low;
  } else {
    q.mask = mask;
  }
  const size_t melems = size_t{2} << (2 * k);
  q.moff.push_back(q.m.size());
  if (matrix_prec == QSIM_HIP_F64) {
    const double* md = static_cast<const double*>(matrix);
    q.m.insert(q.m.end(), md, md + melems);
  } else {
    const float* mf = static_cast<const float*>(matrix);
    q.m.insert(q.m.end(), mf, mf + melems);
  }
  q.ks.push_back(k);
  for (unsigned i = 0; i < 4; ++i) q.qs.push_back(i < k ? qs[i] : 0);
  *queued = true;
  return QSIM_HIP_OK;
}

}  // namespace qhip

extern "C" {

int qsim_hip_set_gate_queue(unsigned tile_bits, unsigned low_bits) {
  if (tile_bits != 0 &&
      (tile_bits < kBlockMinBits || tile_bits > kBlockMaxBits ||
       low_bits < 3 || low_bits > tile_bits)) {
    SetError("gate queue: tile_bits %u..%u (or 0 for off) and low_bits 3.."
             "tile_bits, got %u, %u", kBlockMinBits, kBlockMaxBits, tile_bits,
             low_bits);
    return QSIM_HIP_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_queue_mu);
  QHIP_TRY(FlushLocked());
  g_queue.tb = tile_bits;
  g_queue.lc = tile_bits == 0 ? 0 : low_bits;
  return QSIM_HIP_OK;
}

int qsim_hip_get_gate_queue(unsigned* tile_bits, unsigned* low_bits,
                            uint64_t* passes, uint64_t* gates) {
  std::lock_guard<std::mutex> lock(g_queue_mu);
  if (tile_bits) *tile_bits = g_queue.tb;
  if (low_bits) *low_bits = g_queue.lc;
  if (passes) *passes = g_queue.passes;
  if (gates) *gates = g_queue.gates;
  return QSIM_HIP_OK;
}

int qsim_hip_flush(void) { return FlushGateQueue(); }

unsigned qsim_hip_block_resident_ops(unsigned tile_bits) {
  return BlockResidentOps(tile_bits);
}

int qsim_hip_apply_gate(int prec, void* s, unsigned n, const unsigned* qs,
                        unsigned k, const void* matrix, int matrix_prec) {
  return GuardHost([&]() -> int {
    QHIP_TRY(CheckStateArgs(prec, s, n));
    bool queued = false;
    QHIP_TRY(Enqueue(prec, s, n, qs, k, matrix, matrix_prec, &queued));
    if (queued) return QSIM_HIP_OK;
    return ApplyGateNow(prec, s, n, qs, k, nullptr, 0, 0, matrix, matrix_prec);
  });
}

int qsim_hip_apply_gate_block(int prec, void* s, unsigned n,
                              const unsigned* tile_qubits,
                              unsigned num_tile_qubits, unsigned num_gates,
                              const unsigned* gate_num_qubits,
                              const unsigned* gate_qubits,
                              const void* matrices, int matrix_prec) {
  return GuardHost([&]() -> int {
    QHIP_TRY(CheckState(prec, s, n));
    if (tile_qubits == nullptr ||
        (num_gates > 0 && (gate_num_qubits == nullptr ||
                           gate_qubits == nullptr || matrices == nullptr))) {
      SetError("null blocked-pass operand");
      return QSIM_HIP_ERR_ARG;
    }
    if (prec != QSIM_HIP_F32) {
      SetError("blocked passes run on complex64 states only");
      return QSIM_HIP_ERR_UNSUPPORTED;
    }
    if (num_gates == 0) return QSIM_HIP_OK;
    const bool accurate = (matrix_prec & QSIM_HIP_BLOCK_ACCURATE) != 0;
    const int arith = (matrix_prec & QSIM_HIP_BLOCK_BF16) != 0
                          ? int(kBlockArithBf16)
                          : int(kBlockArithF16);
    std::vector<size_t> moff;
    std::vector<double> m64;
    QHIP_TRY(BlockMatrices(
        num_gates, gate_num_qubits, matrices,
        matrix_prec & ~(QSIM_HIP_BLOCK_ACCURATE | QSIM_HIP_BLOCK_BF16), moff,
        m64));
    return LaunchGateBlock(static_cast<float2*>(s), n, tile_qubits,
                           num_tile_qubits, num_gates, gate_num_qubits,
                           gate_qubits, m64.data(), moff.data(), accurate,
                           arith);
  });
}

int qsim_hip_block_prepare(unsigned n, const unsigned* tile_qubits,
                           unsigned num_tile_qubits, unsigned num_gates,
                           const unsigned* gate_num_qubits,
                           const unsigned* gate_qubits, const void* matrices,
                           int matrix_prec, void** handle) {
  return GuardHost([&]() -> int {
    if (handle == nullptr || tile_qubits == nullptr || num_gates == 0 ||
        gate_num_qubits == nullptr || gate_qubits == nullptr ||
        matrices == nullptr) {
      SetError("null or empty blocked-pass operand");
      return QSIM_HIP_ERR_ARG;
    }
    *handle = nullptr;
    QHIP_TRY(EnsureInit());
    const bool accurate = (matrix_prec & QSIM_HIP_BLOCK_ACCURATE) != 0;
    const int arith = (matrix_prec & QSIM_HIP_BLOCK_BF16) != 0
                          ? int(kBlockArithBf16)
                          : int(kBlockArithF16);
    std::vector<size_t> moff;
    std::vector<double> m64;
    QHIP_TRY(BlockMatrices(
        num_gates, gate_num_qubits, matrices,
        matrix_prec & ~(QSIM_HIP_BLOCK_ACCURATE | QSIM_HIP_BLOCK_BF16), moff,
        m64));
    return PrepareBlockHandle(n, tile_qubits, num_tile_qubits, num_gates,
                              gate_num_qubits, gate_qubits, m64.data(),
                              moff.data(), handle, accurate, arith);
  });
}

int qsim_hip_block_run(void* handle, int prec, void* s, unsigned n) {
  QHIP_TRY(CheckState(prec, s, n));
  if (handle == nullptr) {
    SetError("null blocked-pass handle");
    return QSIM_HIP_ERR_ARG;
  }
  if (prec != QSIM_HIP_F32) {
    SetError("blocked passes run on complex64 states only");
    return QSIM_HIP_ERR_UNSUPPORTED;
  }
  return RunBlockHandle(handle, static_cast<float2*>(s), n);
}

int qsim_hip_block_free(void* handle) {
  if (handle == nullptr) return QSIM_HIP_OK;
  return FreeBlockHandle(handle);
}

int qsim_hip_expectation_value(int prec, const void* s, unsigned n,
                               const unsigned* qs, unsigned k,
                               const void* matrix, int matrix_prec, double* re,
                               double* im) {
  return GuardHost([&]() -> int {
    QHIP_TRY(CheckState(prec, s, n));
    if (qs == nullptr || matrix == nullptr || re == nullptr || im == nullptr) {
      SetError("null operand");
      return QSIM_HIP_ERR_ARG;
    }
    if (k > 6) {
      SetError("operators on %u > 6 qubits are not supported", k);
      return QSIM_HIP_ERR_UNSUPPORTED;
    }
    const size_t melems = size_t{2} << (2 * k);
    if (prec == QSIM_HIP_F32) {
      std::vector<float> mf;
      const float* m = static_cast<const float*>(matrix);
      if (matrix_prec == QSIM_HIP_F64) {
        mf.resize(melems);
        for (size_t i = 0; i < melems; ++i)
          mf[i] = float(static_cast<const double*>(matrix)[i]);
        m = mf.data();
      }
      return LaunchExpectation<float>((const float2*)s, n, qs, k, m, re, im);
    } else {
      std::vector<double> md;
      const double* m = static_cast<const double*>(matrix);
      if (matrix_prec == QSIM_HIP_F32) {
        md.resize(melems);
        for (size_t i = 0; i < melems; ++i)
          md[i] = static_cast<const float*>(matrix)[i];
        m = md.data();
      }
      return LaunchExpectation<double>((const double2*)s, n, qs, k, m, re, im);
    }
  });
}

int qsim_hip_set_gate_method(int method) {
  if (method < QSIM_HIP_METHOD_AUTO || method > QSIM_HIP_METHOD_MFMA_FP16X2) {
    SetError("unknown gate method %d", method);
    return QSIM_HIP_ERR_ARG;
  }
  g_method.store(method);
  return QSIM_HIP_OK;
}

int qsim_hip_get_gate_method(void) { return g_method.load(); }

int qsim_hip_event_record(int slot) {
  QHIP_TRY(EnsureInit());
  QHIP_TRY(FlushGateQueue());
  if (slot < 0 || slot >= 64) {
    SetError("event slot %d out of range", slot);
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_CHECK(hipEventRecord(g_dev.ev[slot], g_stream));
  return QSIM_HIP_OK;
}

int qsim_hip_event_elapsed_ms(int a, int b, float* ms) {
  QHIP_TRY(EnsureInit());
  if (a < 0 || a >= 64 || b < 0 || b >= 64 || ms == nullptr) {
    SetError("bad event slots");
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_CHECK(hipEventSynchronize(g_dev.ev[b]));
  QHIP_CHECK(hipEventElapsedTime(ms, g_dev.ev[a], g_dev.ev[b]));
  return QSIM_HIP_OK;
}

int qsim_hip_graph_record_begin(void** plan) {
  return GuardHost([&]() -> int {
    QHIP_TRY(EnsureInit());
    QHIP_TRY(FlushGateQueue());
    if (plan == nullptr) {
      SetError("null plan handle");
      return QSIM_HIP_ERR_ARG;
    }
    if (g_plan_mode != kPlanOff) {
      SetError("a pass sequence is already being recorded or captured");
      return QSIM_HIP_ERR_ARG;
    }
    g_plan = new GraphPlan{};
    g_plan_mode = kPlanRecord;
    *plan = g_plan;
    return QSIM_HIP_OK;
  });
}

int qsim_hip_graph_record_end(void* plan) {
  if (plan == nullptr || plan != g_plan || g_plan_mode != kPlanRecord) {
    SetError("no recording in progress for this plan");
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_TRY(FlushGateQueue());
  g_plan_mode = kPlanOff;
  g_plan = nullptr;
  QHIP_CHECK(hipStreamSynchronize(g_stream));
  return QSIM_HIP_OK;
}

int qsim_hip_graph_capture_begin(void* plan) {
  auto* pl = static_cast<GraphPlan*>(plan);
  if (pl == nullptr || g_plan_mode != kPlanOff || pl->exec != nullptr) {
    SetError("plan is not ready for capture");
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_TRY(FlushGateQueue());
  QHIP_CHECK(hipStreamSynchronize(g_stream));
  if (pl->stream == nullptr) QHIP_CHECK(hipStreamCreate(&pl->stream));
  pl->next = 0;
  pl->mismatch = false;
  pl->saved = g_stream;
  QHIP_CHECK(hipStreamBeginCapture(pl->stream, hipStreamCaptureModeRelaxed));
  g_stream = pl->stream;
  g_plan = pl;
  g_plan_mode = kPlanCapture;
  return QSIM_HIP_OK;
}

int qsim_hip_graph_capture_end(void* plan) {
  auto* pl = static_cast<GraphPlan*>(plan);
  if (pl == nullptr || pl != g_plan || g_plan_mode != kPlanCapture) {
    SetError("no capture in progress for this plan");
    return QSIM_HIP_ERR_ARG;
  }
  // Gates still queued belong to the captured sequence.
  const int frc = FlushGateQueue();
  g_plan_mode = kPlanOff;
  g_plan = nullptr;
  g_stream = pl->saved;
  hipGraph_t graph = nullptr;
  QHIP_CHECK(hipStreamEndCapture(pl->stream, &graph));
  pl->graph = graph;
  if (frc != QSIM_HIP_OK) return frc;
  if (pl->mismatch || pl->next != pl->bufs.size()) {
    SetError("captured passes differ from the recorded ones");
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_CHECK(hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0));
  return QSIM_HIP_OK;
}

int qsim_hip_graph_launch(void* plan) {
  auto* pl = static_cast<GraphPlan*>(plan);
  if (pl == nullptr || pl->exec == nullptr) {
    SetError("plan has no captured graph");
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_TRY(FlushGateQueue());
  QHIP_CHECK(hipGraphLaunch(pl->exec, g_stream));
  return QSIM_HIP_OK;
}

int qsim_hip_graph_free(void* plan) {
  auto* pl = static_cast<GraphPlan*>(plan);
  if (pl == nullptr) return QSIM_HIP_OK;
  if (pl == g_plan) {
    SetError("plan is being recorded or captured");
    return QSIM_HIP_ERR_ARG;
  }
  // Launches of the graph may still be in flight.
  QHIP_CHECK(hipStreamSynchronize(g_stream));
  if (pl->exec) QHIP_CHECK(hipGraphExecDestroy(pl->exec));
  if (pl->graph) QHIP_CHECK(hipGraphDestroy(pl->graph));
  if (pl->stream) QHIP_CHECK(hipStreamDestroy(pl->stream));
  for (void* b : pl->bufs) QHIP_CHECK(hipFree(b));
  delete pl;
  return QSIM_HIP_OK;
}

int qsim_hip_pass_timer_start(unsigned capacity) {
  return GuardHost([&]() -> int {
    QHIP_TRY(EnsureInit());
    QHIP_TRY(FlushGateQueue());
    if (capacity == 0 || capacity > (1u << 20)) {
      SetError("pass timer capacity %u out of range", capacity);
      return QSIM_HIP_ERR_ARG;
    }
    while (g_timer.ev.size() < 2 * size_t(capacity)) {
      hipEvent_t e;
      QHIP_CHECK(hipEventCreate(&e));
      g_timer.ev.push_back(e);
    }
    g_timer.capacity = capacity;
    g_timer.count = 0;
    g_timer.on = true;
    return QSIM_HIP_OK;
  });
}

int qsim_hip_pass_timer_stop(float* ms, unsigned max_count, unsigned* count) {
  QHIP_TRY(EnsureInit());
  QHIP_TRY(FlushGateQueue());
  g_timer.on = false;
  QHIP_CHECK(hipStreamSynchronize(g_stream));
  unsigned n = std::min(g_timer.count, max_count);
  for (unsigned i = 0; i < n; ++i) {
    QHIP_CHECK(hipEventElapsedTime(&ms[i], g_timer.ev[2 * i],
                                   g_timer.ev[2 * i + 1]));
  }
  if (count) *count = g_timer.count;
  return QSIM_HIP_OK;
}

}  // extern "C"
