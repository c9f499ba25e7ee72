// traj_kernels.hip -- batched noisy trajectories in one launch.
//
// qsim's trajectory simulator (lib/qtrajectory.h:112-387) runs trajectories
// one after another, each gate batch as its own set of kernel launches with a
// host synchronisation at every channel whose Kraus operator depends on the
// state.  Here a whole batch -- hundreds of trajectories of uneven qubit
// counts -- is ONE launch: workgroup t owns trajectory t, its state in a
// device pool and its operation stream, and walks the stream with nothing
// but workgroup barriers between operations:
//
//   gate      a fused gate (k <= 4, FP32 matrix in LDS) over all 2^(n-k)
//             groups of the trajectory's state;
//   decision  an amplitude-damping channel whose draw fell past the
//             probability bound: FP64 norm and |1>-population of the qubit,
//             the choice between K0 and K1 exactly as RunIteration makes it
//             (lib/qtrajectory.h:322-356), then the chosen operator applied.
//
// The Kraus operators that the bounds decide (depolarising Paulis, damping
// K0 when the draw is below 1 - gamma) are sampled on the host from the same
// mt19937(rep) draws and folded into the fused gates.  At the end each
// trajectory writes <Z_a Z_b> of its qubit pairs, FP64-accumulated and
// divided by the norm (the normalisation of lib/qtrajectory.h:377-383).
//
// Workgroups are dispatched in trajectory order, so the host puts the
// largest states first (longest-processing-time-first balancing across CUs).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"
#include "reduce.h"

namespace qhip {

namespace {

constexpr unsigned kPairChunk = 16;

template <int NT, int K>
__device__ void ApplyGateK(float2* __restrict__ s, unsigned n,
                           const qsim_hip_traj_op& o, const float* lm) {
  constexpr unsigned D = 1u << K;
  const uint64_t ngroups = uint64_t{1} << (n - K);
  unsigned q[K];
#pragma unroll
  for (int b = 0; b < K; ++b) q[b] = o.q[b];
  for (uint64_t g = threadIdx.x; g < ngroups; g += NT) {
    uint64_t base = g;
#pragma unroll
    for (int b = 0; b < K; ++b) {
      base = ((base >> q[b]) << (q[b] + 1)) | (base & ((uint64_t{1} << q[b]) - 1));
    }
    float2 v[D];
#pragma unroll
    for (unsigned j = 0; j < D; ++j) {
      uint64_t off = 0;
#pragma unroll
      for (int b = 0; b < K; ++b) off |= uint64_t((j >> b) & 1u) << q[b];
      v[j] = s[base + off];
    }
    // Rows stay a loop for k >= 3: unrolled, the compiler would hoist the
    // whole LDS matrix (2 D^2 floats) into registers across the group loop.
#pragma unroll(K <= 2 ? D : 1)
    for (unsigned r = 0; r < D; ++r) {
      float re = 0.0f, im = 0.0f;
#pragma unroll
      for (unsigned j = 0; j < D; ++j) {
        const float mr = lm[2 * (r * D + j)], mi = lm[2 * (r * D + j) + 1];
        re = fmaf(mr, v[j].x, re);
        re = fmaf(-mi, v[j].y, re);
        im = fmaf(mr, v[j].y, im);
        im = fmaf(mi, v[j].x, im);
      }
      uint64_t off = 0;
#pragma unroll
      for (int b = 0; b < K; ++b) off |= uint64_t((r >> b) & 1u) << q[b];
      s[base + off] = float2{re, im};
    }
  }
}

template <int NT>
__global__ void __launch_bounds__(NT)
TrajectoryKernel(const qsim_hip_traj_desc* __restrict__ descs,
                 const qsim_hip_traj_op* __restrict__ ops,
                 const float* __restrict__ mats,
                 const uint32_t* __restrict__ pairs, float2* __restrict__ pool,
                 double* __restrict__ zz_out, int32_t* __restrict__ dec_out) {
  __shared__ float lm[2 * 16 * 16];
  __shared__ double shared_d[2];
  const qsim_hip_traj_desc d = descs[blockIdx.x];
  const unsigned n = d.num_qubits;
  const uint64_t N = uint64_t{1} << n;
  float2* s = pool + d.state_offset;

  for (uint64_t i = threadIdx.x; i < N; i += NT) {
    s[i] = float2{i == 0 ? 1.0f : 0.0f, 0.0f};
  }
  __syncthreads();

  for (uint64_t oi = d.op_begin; oi < d.op_end; ++oi) {
    const qsim_hip_traj_op o = ops[oi];
    if (o.kind == QSIM_HIP_TRAJ_GATE) {
      const unsigned D = 1u << o.k;
      for (unsigned i = threadIdx.x; i < 2 * D * D; i += NT) lm[i] = mats[o.moff + i];
      __syncthreads();
      switch (o.k) {
        case 1: ApplyGateK<NT, 1>(s, n, o, lm); break;
        case 2: ApplyGateK<NT, 2>(s, n, o, lm); break;
        case 3: ApplyGateK<NT, 3>(s, n, o, lm); break;
        default: ApplyGateK<NT, 4>(s, n, o, lm); break;
      }
      __syncthreads();
      continue;
    }
    // Amplitude-damping decision on qubit q: FP64 populations of |0> and |1>.
    const unsigned q = o.q[0];
    double p0 = 0.0, p1 = 0.0;
    for (uint64_t i = threadIdx.x; i < N; i += NT) {
      const float2 a = s[i];
      const double w = double(a.x) * a.x + double(a.y) * a.y;
      if ((i >> q) & 1) p1 += w; else p0 += w;
    }
    p0 = BlockSum<NT>(p0);
    p1 = BlockSum<NT>(p1);
    if (threadIdx.x == 0) {
      shared_d[0] = p0;
      shared_d[1] = p1;
    }
    __syncthreads();
    p0 = shared_d[0];
    p1 = shared_d[1];
    __syncthreads();
    // RunIteration normalises a non-unitary state first, then samples K0 iff
    // r < <psi|K0^dagger K0|psi> (the bound 1 - gamma cancels out of cp).
    const double norm = p0 + p1;
    const double prob0 = (p0 + double(o.kdk1) * p1) / norm;
    const int pick = double(o.r) < prob0 ? 0 : 1;
    const float scale = o.pad & 1u ? float(1.0 / std::sqrt(norm)) : 1.0f;
    const float* km = mats + (pick == 0 ? o.moff : o.moff2);
    const float m00r = km[0] * scale, m00i = km[1] * scale;
    const float m01r = km[2] * scale, m01i = km[3] * scale;
    const float m10r = km[4] * scale, m10i = km[5] * scale;
    const float m11r = km[6] * scale, m11i = km[7] * scale;
    const uint64_t half = N >> 1, qb = uint64_t{1} << q;
    for (uint64_t g = threadIdx.x; g < half; g += NT) {
      const uint64_t i0 = ((g >> q) << (q + 1)) | (g & (qb - 1));
      const float2 a = s[i0], b = s[i0 + qb];
      s[i0] = float2{m00r * a.x - m00i * a.y + m01r * b.x - m01i * b.y,
                     m00r * a.y + m00i * a.x + m01r * b.y + m01i * b.x};
      s[i0 + qb] = float2{m10r * a.x - m10i * a.y + m11r * b.x - m11i * b.y,
                          m10r * a.y + m10i * a.x + m11r * b.y + m11i * b.x};
    }
    if (threadIdx.x == 0) dec_out[o.slot] = pick;
    __syncthreads();
  }

  // <Z_a Z_b> = sum_i |s_i|^2 (-1)^(bit a + bit b) / sum_i |s_i|^2.
  for (unsigned p0 = 0; p0 < d.num_pairs; p0 += kPairChunk) {
    const unsigned np = std::min(kPairChunk, d.num_pairs - p0);
    unsigned pa[kPairChunk], pb[kPairChunk];
    double acc[kPairChunk];
#pragma unroll
    for (unsigned j = 0; j < kPairChunk; ++j) {
      pa[j] = j < np ? pairs[2 * (d.pair_offset + p0 + j)] : 0;
      pb[j] = j < np ? pairs[2 * (d.pair_offset + p0 + j) + 1] : 0;
      acc[j] = 0.0;
    }
    double nrm = 0.0;
    for (uint64_t i = threadIdx.x; i < N; i += NT) {
      const float2 a = s[i];
      const double w = double(a.x) * a.x + double(a.y) * a.y;
      nrm += w;
#pragma unroll
      for (unsigned j = 0; j < kPairChunk; ++j) {
        const unsigned par = unsigned((i >> pa[j]) ^ (i >> pb[j])) & 1u;
        acc[j] += par ? -w : w;
      }
    }
    nrm = BlockSum<NT>(nrm);
    for (unsigned j = 0; j < kPairChunk; ++j) {
      const double v = BlockSum<NT>(acc[j]);
      if (threadIdx.x == 0 && j < np) zz_out[d.pair_offset + p0 + j] = v / nrm;
    }
  }

  // A non-unitary trajectory ends normalised (lib/qtrajectory.h:377-383).
  if (d.pad & 1u) {
    double nrm = 0.0;
    for (uint64_t i = threadIdx.x; i < N; i += NT) {
      const float2 a = s[i];
      nrm += double(a.x) * a.x + double(a.y) * a.y;
    }
    nrm = BlockSum<NT>(nrm);
    if (threadIdx.x == 0) shared_d[0] = nrm;
    __syncthreads();
    const float a = float(1.0 / std::sqrt(shared_d[0]));
    for (uint64_t i = threadIdx.x; i < N; i += NT) {
      const float2 v = s[i];
      s[i] = float2{v.x * a, v.y * a};
    }
  }
}

struct TrajBatch {
  unsigned num_traj = 0;
  uint32_t num_pairs = 0, num_decisions = 0;
  std::vector<unsigned> max_qubits;
  qsim_hip_traj_desc* descs = nullptr;
  qsim_hip_traj_op* ops = nullptr;
  float* mats = nullptr;
  uint32_t* pairs = nullptr;
  float2* pool = nullptr;
  double* zz = nullptr;
  int32_t* dec = nullptr;
  std::vector<uint64_t> state_offset;
  std::vector<unsigned> nq;
  hipStream_t gstream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  unsigned graph_threads = 0;
};

template <int NT>
void LaunchTraj(TrajBatch& b, hipStream_t st) {
  hipLaunchKernelGGL(TrajectoryKernel<NT>, dim3(b.num_traj), dim3(NT), 0, st,
                     b.descs, b.ops, b.mats, b.pairs, b.pool, b.zz, b.dec);
}

int Launch(TrajBatch& b, unsigned threads, hipStream_t st) {
  switch (threads) {
    case 256: LaunchTraj<256>(b, st); break;
    case 512: LaunchTraj<512>(b, st); break;
    case 1024: LaunchTraj<1024>(b, st); break;
    default:
      SetError("trajectory batches run with 256, 512 or 1024 threads, not %u",
               threads);
      return QSIM_HIP_ERR_ARG;
  }
  QHIP_LAUNCHED("TrajectoryKernel");
  return QSIM_HIP_OK;
}

__global__ void FillInt(int32_t* __restrict__ p, unsigned n, int32_t v) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
int Upload(T** dst, const T* src, size_t count) {
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  QHIP_CHECK(hipMalloc(dst, bytes));
  if (count) {
    QHIP_CHECK(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  }
  return QSIM_HIP_OK;
}

void Release(TrajBatch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->gstream) (void)hipStreamDestroy(b->gstream);
  (void)hipFree(b->descs);
  (void)hipFree(b->ops);
  (void)hipFree(b->mats);
  (void)hipFree(b->pairs);
  (void)hipFree(b->pool);
  (void)hipFree(b->zz);
  (void)hipFree(b->dec);
  delete b;
}

}  // namespace

}  // namespace qhip

using namespace qhip;

extern "C" {

int qsim_hip_traj_prepare(const qsim_hip_traj_desc* trajs, unsigned num_traj,
                          const qsim_hip_traj_op* ops, uint64_t num_ops,
                          const float* matrices, uint64_t num_floats,
                          const uint32_t* pairs, uint32_t num_pairs,
                          uint32_t num_decisions, void** handle) {
  if (handle == nullptr || trajs == nullptr || num_traj == 0 ||
      (num_ops && ops == nullptr) || (num_floats && matrices == nullptr) ||
      (num_pairs && pairs == nullptr)) {
    SetError("trajectory batch: null or empty operand");
    return QSIM_HIP_ERR_ARG;
  }
  *handle = nullptr;
  QHIP_TRY(EnsureInit());
  // Validate every stream against the kernel's assumptions before anything
  // reaches the device.
  uint64_t pool = 0;
  std::vector<qsim_hip_traj_desc> d(trajs, trajs + num_traj);
  for (unsigned t = 0; t < num_traj; ++t) {
    const unsigned n = d[t].num_qubits;
    if (n < 4 || n > 34 || d[t].op_begin > d[t].op_end ||
        d[t].op_end > num_ops ||
        uint64_t(d[t].pair_offset) + d[t].num_pairs > num_pairs) {
      SetError("trajectory %u: bad size, operation range or pair range", t);
      return QSIM_HIP_ERR_ARG;
    }
    d[t].state_offset = pool;
    pool += uint64_t{1} << n;
    for (uint64_t i = d[t].op_begin; i < d[t].op_end; ++i) {
      const qsim_hip_traj_op& o = ops[i];
      bool ok;
      if (o.kind == QSIM_HIP_TRAJ_GATE) {
        ok = o.k >= 1 && o.k <= 4 &&
             uint64_t(o.moff) + (size_t{2} << (2 * o.k)) <= num_floats;
        for (unsigned b = 0; ok && b < o.k; ++b)
          ok = o.q[b] < n && (b == 0 || o.q[b] > o.q[b - 1]);
      } else if (o.kind == QSIM_HIP_TRAJ_DAMP_DECIDE) {
        ok = o.q[0] < n && uint64_t(o.moff) + 8 <= num_floats &&
             uint64_t(o.moff2) + 8 <= num_floats && o.slot < num_decisions;
      } else {
        ok = false;
      }
      if (!ok) {
        SetError("trajectory %u: operation %llu is malformed", t,
                 (unsigned long long)(i - d[t].op_begin));
        return QSIM_HIP_ERR_ARG;
      }
    }
    for (uint32_t p = 0; p < d[t].num_pairs; ++p) {
      const uint32_t* pp = pairs + 2 * (size_t(d[t].pair_offset) + p);
      if (pp[0] >= n || pp[1] >= n || pp[0] == pp[1]) {
        SetError("trajectory %u: bad qubit pair", t);
        return QSIM_HIP_ERR_ARG;
      }
    }
  }
  auto* b = new TrajBatch;
  b->num_traj = num_traj;
  b->num_pairs = num_pairs;
  b->num_decisions = num_decisions;
  for (unsigned t = 0; t < num_traj; ++t) {
    b->state_offset.push_back(d[t].state_offset);
    b->nq.push_back(d[t].num_qubits);
  }
  int rc = Upload(&b->descs, d.data(), d.size());
  if (rc == QSIM_HIP_OK) rc = Upload(&b->ops, ops, num_ops);
  if (rc == QSIM_HIP_OK) rc = Upload(&b->mats, matrices, num_floats);
  if (rc == QSIM_HIP_OK) rc = Upload(&b->pairs, pairs, size_t(2) * num_pairs);
  if (rc == QSIM_HIP_OK) {
    hipError_t e = hipMalloc(&b->pool, pool * sizeof(float2));
    if (e == hipSuccess) e = hipMalloc(&b->zz, std::max<size_t>(num_pairs, 1) * 8);
    if (e == hipSuccess)
      e = hipMalloc(&b->dec, std::max<size_t>(num_decisions, 1) * 4);
    if (e != hipSuccess) rc = HipFail(e, "trajectory batch buffers");
  }
  if (rc != QSIM_HIP_OK) {
    Release(b);
    return rc;
  }
  *handle = b;
  return QSIM_HIP_OK;
}

int qsim_hip_traj_run(void* handle, unsigned threads, int use_graph) {
  auto* b = static_cast<TrajBatch*>(handle);
  if (b == nullptr) return QSIM_HIP_ERR_ARG;
  // Undecided slots read -1 (the decision kernel writes the reached ones).
  hipLaunchKernelGGL(FillInt, dim3((b->num_decisions + 255) / 256 + 1),
                     dim3(256), 0, Stream(), b->dec, b->num_decisions, -1);
  QHIP_LAUNCHED("FillInt");
  if (!use_graph) {
    PassTimerBegin();
    QHIP_TRY(Launch(*b, threads, Stream()));
    PassTimerEnd();
    QHIP_CHECK(hipStreamSynchronize(Stream()));
    return QSIM_HIP_OK;
  }
  QHIP_CHECK(hipStreamSynchronize(Stream()));
  if (b->exec == nullptr || b->graph_threads != threads) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    if (b->gstream == nullptr) QHIP_CHECK(hipStreamCreate(&b->gstream));
    QHIP_CHECK(hipStreamBeginCapture(b->gstream, hipStreamCaptureModeThreadLocal));
    int rc = Launch(*b, threads, b->gstream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->gstream, &g);
    if (rc != QSIM_HIP_OK) return rc;
    if (e != hipSuccess) return HipFail(e, "hipStreamEndCapture");
    b->graph = g;
    QHIP_CHECK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    b->graph_threads = threads;
  }
  QHIP_CHECK(hipGraphLaunch(b->exec, b->gstream));
  QHIP_CHECK(hipStreamSynchronize(b->gstream));
  return QSIM_HIP_OK;
}

int qsim_hip_traj_results(void* handle, double* zz, uint32_t num_pairs,
                          int32_t* decisions, uint32_t num_decisions) {
  auto* b = static_cast<TrajBatch*>(handle);
  if (b == nullptr || num_pairs > b->num_pairs ||
      num_decisions > b->num_decisions) {
    SetError("trajectory batch: result sizes exceed the batch");
    return QSIM_HIP_ERR_ARG;
  }
  if (zz && num_pairs)
    QHIP_CHECK(hipMemcpy(zz, b->zz, num_pairs * 8ull, hipMemcpyDeviceToHost));
  if (decisions && num_decisions)
    QHIP_CHECK(hipMemcpy(decisions, b->dec, num_decisions * 4ull,
                         hipMemcpyDeviceToHost));
  return QSIM_HIP_OK;
}

int qsim_hip_traj_state(void* handle, unsigned t, float* out) {
  auto* b = static_cast<TrajBatch*>(handle);
  if (b == nullptr || t >= b->num_traj || out == nullptr) {
    SetError("trajectory batch: bad trajectory index");
    return QSIM_HIP_ERR_ARG;
  }
  QHIP_CHECK(hipMemcpy(out, b->pool + b->state_offset[t],
                       (size_t{1} << b->nq[t]) * sizeof(float2),
                       hipMemcpyDeviceToHost));
  return QSIM_HIP_OK;
}

int qsim_hip_traj_free(void* handle) {
  auto* b = static_cast<TrajBatch*>(handle);
  if (b == nullptr) return QSIM_HIP_OK;
  QHIP_CHECK(hipStreamSynchronize(Stream()));
  Release(b);
  return QSIM_HIP_OK;
}

}  // extern "C"
