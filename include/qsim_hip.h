/* qsim_hip.h -- C ABI of the MI355X (gfx950) state-vector backend for qsim.
 *
 * Plain pointers and sizes only.  State vectors are device pointers to
 * interleaved complex numbers ("normal order": amplitude i is (s[2i], s[2i+1]))
 * of either complex64 (QSIM_HIP_F32) or complex128 (QSIM_HIP_F64); a state of
 * n qubits holds qsim_hip_min_size(n) real elements.  All functions return
 * QSIM_HIP_OK (0) or a negative code; qsim_hip_last_error() describes the last
 * failure of the calling thread.  Work is issued on the library stream
 * (the legacy default stream unless qsim_hip_set_stream() chose another);
 * gate applications are asynchronous, reductions and host copies synchronous.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * the qsim sources).  The C++ adapter in include/qsim_amd/ maps these calls
 * onto qsim's Simulator / StateSpace / VectorSpace template API.
 */
#ifndef QSIM_HIP_H_
#define QSIM_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  QSIM_HIP_OK = 0,
  QSIM_HIP_ERR_ARG = -1,         /* invalid argument (sizes, qubits) */
  QSIM_HIP_ERR_RUNTIME = -2,     /* HIP runtime / kernel launch failure */
  QSIM_HIP_ERR_OOM = -3,         /* device allocation failed */
  QSIM_HIP_ERR_UNSUPPORTED = -4  /* gate size beyond what the path supports */
};

/* Element type of a state vector. */
enum { QSIM_HIP_F32 = 0, QSIM_HIP_F64 = 1 };

/* Arithmetic used for an uncontrolled gate pass (qsim_hip_set_gate_method). */
enum {
  /* complex64, uncontrolled k >= 4: FP32 MFMA (exact FP32 products and
   * sums); everything else: the FP32 / FP64 vector kernels */
  QSIM_HIP_METHOD_AUTO = 0,
  QSIM_HIP_METHOD_VECTOR = 1,       /* always the vector kernels */
  QSIM_HIP_METHOD_MFMA_BF16X3 = 2,  /* as AUTO, but k >= 5 on split-BF16
                                       MFMA (matrix in 3 bf16 parts, state in
                                       2, 4 products) with FP32 accumulation */
  QSIM_HIP_METHOD_MFMA_FP16X2 = 3   /* as AUTO, but k >= 5 on FP16 MFMA
                                       with a power-of-two scale per matrix
                                       and per group (matrix and state in 2
                                       fp16 parts each, 3 products) and FP32
                                       accumulation */
};

/* ---- context ----------------------------------------------------------- */

/* Last error message of the calling thread ("" if none). */
const char* qsim_hip_last_error(void);

/* Selects the device and creates the library's per-device resources.
 * Called implicitly by every other entry point with device 0. */
int qsim_hip_init(int device);

/* Device name, total memory (bytes) and compute-unit count. */
int qsim_hip_device_info(char* name, size_t name_len, uint64_t* total_mem,
                         int* num_cus);

/* Issues work on `stream` (a hipStream_t; NULL = legacy default stream). */
int qsim_hip_set_stream(void* stream);

/* Waits for all work on the library stream.
 * Replaces VectorSpaceCUDA::DeviceSync (lib/vectorspace_cuda.h:162-164). */
int qsim_hip_synchronize(void);

/* ---- VectorSpace (lib/vectorspace_cuda.h:42-167) ----------------------- */

/* Real elements of an n-qubit state: 2 * 2^n, at least 64 so that tiny
 * states keep a 256-byte footprint.  Replaces StateSpaceCUDA::MinSize
 * (lib/statespace_cuda.h:81-83). */
uint64_t qsim_hip_min_size(unsigned num_qubits);

/* Allocates an n-qubit state (uninitialised).  On OOM returns
 * QSIM_HIP_ERR_OOM and *out = NULL.  Replaces VectorSpaceCUDA::Create
 * (lib/vectorspace_cuda.h:87-96). */
int qsim_hip_alloc(int prec, unsigned num_qubits, void** out);

/* Frees a pointer from qsim_hip_alloc (NULL is a no-op).  Replaces
 * VectorSpaceCUDA::Free / detail::free (lib/vectorspace_cuda.h:31-33). */
int qsim_hip_free(void* ptr);

/* Byte copies device<-device, host<-device and device<-host.  Replace the
 * four VectorSpaceCUDA::Copy overloads (lib/vectorspace_cuda.h:116-160). */
int qsim_hip_copy_d2d(void* dst, const void* src, uint64_t bytes);
int qsim_hip_copy_d2h(void* dst, const void* src, uint64_t bytes);
int qsim_hip_copy_h2d(void* dst, const void* src, uint64_t bytes);

/* Copies an n-qubit state between element types (complex64 <-> complex128,
 * or the same type), rounding to nearest.  No reference counterpart: it lets
 * a complex64 run be compared with a complex128 run on the device. */
int qsim_hip_copy_convert(void* dst, int dst_prec, const void* src,
                          int src_prec, unsigned num_qubits);

/* ---- StateSpace (lib/statespace_cuda.h) -------------------------------- */

/* All amplitudes 0 (lib/statespace_cuda.h:109-112). */
int qsim_hip_set_all_zeros(int prec, void* state, unsigned num_qubits);

/* |0...0> (lib/statespace_cuda.h:130-135). */
int qsim_hip_set_state_zero(int prec, void* state, unsigned num_qubits);

/* Uniform superposition 2^(-n/2) (lib/statespace_cuda.h:115-127). */
int qsim_hip_set_state_uniform(int prec, void* state, unsigned num_qubits);

/* Layout conversion.  The device layout is already normal order, so both
 * are no-ops kept for API parity (lib/statespace_cuda.h:85-107). */
int qsim_hip_internal_to_normal_order(int prec, void* state,
                                      unsigned num_qubits);
int qsim_hip_normal_to_internal_order(int prec, void* state,
                                      unsigned num_qubits);

/* One amplitude; thread-safe (lib/statespace_cuda.h:138-164). */
int qsim_hip_get_ampl(int prec, const void* state, unsigned num_qubits,
                      uint64_t i, double* re, double* im);
int qsim_hip_set_ampl(int prec, void* state, unsigned num_qubits, uint64_t i,
                      double re, double im);

/* out[j] = state[indices[j]] for j < count: one gather pass and one copy
 * instead of a GetAmpl per bitstring.  `out` is host memory of `count`
 * complex values in the state's precision (float or double pairs); every
 * index must be < 2^num_qubits (QSIM_HIP_ERR_ARG otherwise, nothing read).
 * Replaces the per-bitstring GetAmpl loops of apps/qsim_amplitudes.cc:138-153
 * (WriteAmplitudes) and lib/hybrid.h:428-438 (the results collection);
 * no single reference function corresponds. */
int qsim_hip_get_ampls(int prec, const void* state, unsigned num_qubits,
                       const uint64_t* indices, uint64_t count, void* out);

/* state[i] = re + i im where (i & mask) == bits (or != bits if exclude)
 * (lib/statespace_cuda.h:168-187). */
int qsim_hip_bulk_set_ampl(int prec, void* state, unsigned num_qubits,
                           uint64_t mask, uint64_t bits, double re, double im,
                           int exclude);

/* dest += src (lib/statespace_cuda.h:190-205). */
int qsim_hip_add(int prec, const void* src, void* dest, unsigned num_qubits);

/* state *= a; `a` is rounded to the state's element type first, as the
 * reference's fp_type argument is (lib/statespace_cuda.h:208-217). */
int qsim_hip_multiply(int prec, void* state, unsigned num_qubits, double a);

/* sum_i conj(a_i) b_i with FP64 accumulation (lib/statespace_cuda.h:219-229). */
int qsim_hip_inner_product(int prec, const void* a, const void* b,
                           unsigned num_qubits, double* re, double* im);

/* Re sum_i conj(a_i) b_i (lib/statespace_cuda.h:231-237). */
int qsim_hip_real_inner_product(int prec, const void* a, const void* b,
                                unsigned num_qubits, double* out);

/* sum_i |s_i|^2 in FP64 (lib/statespace_cuda.h:239-241). */
int qsim_hip_norm(int prec, const void* state, unsigned num_qubits,
                  double* out);

/* qsim_hip_inner_product (im != NULL) or qsim_hip_real_inner_product
 * (im == NULL) without any device allocation: the partial sums go to a
 * fixed device array of the library (2047 partial sums, 32 KiB, held until
 * the process ends; every reduction of any state size fits it) instead of
 * the grow-only scratch.  No counterpart in the reference: for callers that
 * cannot allocate device memory (graph capture, serving loops).  Needs an
 * initialised library (qsim_hip_init or any earlier call); refuses with
 * QSIM_HIP_ERR_ARG while the gate queue holds gates (qsim_hip_flush first,
 * since running them could stage a pass).  Calls that overlap on one GPU
 * (other processes, or graphs replayed beside it) race on the array. */
int qsim_hip_inner_product_noalloc(int prec, const void* a, const void* b,
                                   unsigned num_qubits, double* re,
                                   double* im);

/* Inverse-CDF sampling with caller-supplied draws: `draws` holds `num`
 * ascending values in [0, norm) (lib/util.h:67-85 order); out[j] = first index
 * k with draws[j] < sum_{i<=k} |s_i|^2 (FP64 prefix sums), or 2^n - 1 if none
 * (lib/statespace_cuda.h:243-312, lib/statespace_basic.h:195-233). */
int qsim_hip_sample(int prec, const void* state, unsigned num_qubits,
                    const double* draws, uint64_t num, uint64_t* out);

/* As qsim_hip_sample with the draws of GenerateRandomValues(num, seed, norm)
 * (lib/util.h:67-85): mt19937(seed), uniform_real_distribution<double>(0,
 * norm), sorted.  Replaces StateSpaceCUDA::Sample (lib/statespace_cuda.h:
 * 243-312). */
int qsim_hip_sample_seeded(int prec, const void* state, unsigned num_qubits,
                           uint64_t num, unsigned seed, uint64_t* out);

/* Number of partial norms qsim_hip_partial_norms returns for n qubits. */
unsigned qsim_hip_num_partial_norms(unsigned num_qubits);

/* Norms of qsim_hip_num_partial_norms(n) equal, contiguous index ranges
 * (lib/statespace_cuda.h:331-355). */
int qsim_hip_partial_norms(int prec, const void* state, unsigned num_qubits,
                           double* out);

/* First index k in partial range m with r < running sum, masked
 * (lib/statespace_cuda.h:357-373, lib/statespace_basic.h:275-295). */
int qsim_hip_find_measured_bits(int prec, const void* state,
                                unsigned num_qubits, unsigned m, double r,
                                uint64_t mask, uint64_t* out);

/* Zeroes amplitudes with (i & mask) != bits and rescales the rest by
 * 1/sqrt(their norm) (lib/statespace_cuda.h:316-329). */
int qsim_hip_collapse(int prec, void* state, unsigned num_qubits,
                      uint64_t mask, uint64_t bits);

/* ---- Simulator (lib/simulator_cuda.h) ---------------------------------- */

/* psi <- U psi.  qs: k ascending qubit indices (k <= 6); matrix: row-major
 * interleaved complex 2^k x 2^k, row/column bit j <-> qs[j], element type
 * matrix_prec.  k = 0 multiplies by the 1x1 matrix.  k > 6 returns
 * QSIM_HIP_ERR_UNSUPPORTED.  Replaces SimulatorCUDA::ApplyGate
 * (lib/simulator_cuda.h:70-125). */
int qsim_hip_apply_gate(int prec, void* state, unsigned num_qubits,
                        const unsigned* qs, unsigned k, const void* matrix,
                        int matrix_prec);

/* Controlled U on k <= 4 targets (any k when nc == 0), nc ascending control
 * qubits, cvals bit i <-> cqs[i].  Replaces SimulatorCUDA::
 * ApplyControlledGate (lib/simulator_cuda.h:135-207). */
int qsim_hip_apply_controlled_gate(int prec, void* state, unsigned num_qubits,
                                   const unsigned* qs, unsigned k,
                                   const unsigned* cqs, unsigned nc,
                                   uint64_t cvals, const void* matrix,
                                   int matrix_prec);

/* A cache-blocked pass: applies num_gates (<= 32) gates in order with one
 * stream of the state through HBM (programs longer than the kernel's LDS
 * holds, 14 ops for 14 tile qubits and 7 otherwise, stream once per such
 * chunk).  The state is cut into tiles over the tile qubits (ascending,
 * 9..14 of them, starting with the run 0, 1, 2, ... of at least 3 qubits, at
 * most 11 beyond it); gate i acts on gate_num_qubits[i] (1..4) ascending
 * qubits gate_qubits[4 i ...], all tile qubits, with the next 2 * 4^k matrix
 * elements of `matrices` (row-major interleaved, bit j <-> qubit j;
 * matrix_prec QSIM_HIP_F32 or QSIM_HIP_F64, optionally OR-ed with the flags
 * below).  Arithmetic, MFMA products with FP32 accumulation:
 *  - default: FP16 with scaling -- each tile is scaled by a power of two
 *    when staged, matrix and amplitudes are split in 2 fp16 parts each, 3
 *    products (6 MFMAs per 32 groups); about 2^-22 relative per gate.
 *    Programs whose matrices could grow the state by more than 2^10 run on
 *    split-BF16 instead;
 *  - QSIM_HIP_BLOCK_BF16: split-BF16, the matrix in 3 bf16 parts and the
 *    state in 2 (8 MFMAs, about 2^-17 per gate), or 3 with
 *    QSIM_HIP_BLOCK_ACCURATE (12 MFMAs; for passes on structured states --
 *    the first passes from a basis state -- where the 2-part split's
 *    rounding repeats across amplitudes and moves the norm coherently).
 * Complex64 states only (else QSIM_HIP_ERR_UNSUPPORTED).  Equivalent to one
 * SimulatorCUDA::ApplyGate (lib/simulator_cuda.h:70-125) per gate, in the
 * apply loop of QSimRunner::Run (lib/run_qsim.h:264-280). */
enum { QSIM_HIP_BLOCK_ACCURATE = 0x100, QSIM_HIP_BLOCK_BF16 = 0x200 };
int qsim_hip_apply_gate_block(int prec, void* state, unsigned num_qubits,
                              const unsigned* tile_qubits,
                              unsigned num_tile_qubits, unsigned num_gates,
                              const unsigned* gate_num_qubits,
                              const unsigned* gate_qubits,
                              const void* matrices, int matrix_prec);

/* The same pass prepared once for n-qubit states: validation, tile geometry
 * and the gates' MFMA operand fragments are built and uploaded at prepare
 * time, so a run is one kernel launch.  A handle is not thread-safe;
 * qsim_hip_block_free waits for launches that may still read it. */
int qsim_hip_block_prepare(unsigned num_qubits, const unsigned* tile_qubits,
                           unsigned num_tile_qubits, unsigned num_gates,
                           const unsigned* gate_num_qubits,
                           const unsigned* gate_qubits, const void* matrices,
                           int matrix_prec, void** handle);
int qsim_hip_block_run(void* handle, int prec, void* state,
                       unsigned num_qubits);
int qsim_hip_block_free(void* handle);

/* Gate queue.  While on (tile_bits 9..14; 0 turns it off), complex64
 * qsim_hip_apply_gate calls on 1..4 qubits are not run one pass each but
 * collected into cache-blocked passes (qsim_hip_apply_gate_block, FP16
 * with per-tile scaling): consecutive gates on one state whose qubits,
 * together with qubits 0 .. low_bits - 1, fit in tile_bits qubits form one
 * pass (at most 14 gates for 14 tile qubits, 7 otherwise), run when the
 * next gate does not fit or goes to another state, or at the first other
 * call of the library -- every other entry point runs the queued gates
 * first, so results read through the library are those of the gates run
 * one by one, up to the blocked pass's rounding (about 2^-22 relative per
 * gate).  A state must not be read or freed behind the library's back
 * while gates for it are queued (qsim_hip_flush runs them).  Grouping
 * consecutive gates greedily gives the fewest passes of any grouping into
 * consecutive runs, so gates emitted pass by pass by a blocked-pass
 * planner (LearnedFuserHIP) come back as those passes.  No reference
 * counterpart: it lets qsim's unchanged apply loop (lib/run_qsim.h:264-280,
 * lib/gate_appl.h:35-48) reach blocked passes through
 * SimulatorCUDA::ApplyGate's interface (lib/simulator_cuda.h:70-125). */
int qsim_hip_set_gate_queue(unsigned tile_bits, unsigned low_bits);
/* The queue's setting and the blocked passes (and gates in them) it has run
 * in this process. */
int qsim_hip_get_gate_queue(unsigned* tile_bits, unsigned* low_bits,
                            uint64_t* passes, uint64_t* gates);
/* Runs the queued gates now. */
int qsim_hip_flush(void);

/* Gates an FP16 blocked pass over tile_bits tile qubits keeps resident (a
 * longer program streams the state once per such chunk): 14 for 14 tile
 * qubits, 7 below. */
unsigned qsim_hip_block_resident_ops(unsigned tile_bits);

/* <psi|U|psi> for k <= 6 with FP64 accumulation.  Replaces SimulatorCUDA::
 * ExpectationValue (lib/simulator_cuda.h:216-260). */
int qsim_hip_expectation_value(int prec, const void* state,
                               unsigned num_qubits, const unsigned* qs,
                               unsigned k, const void* matrix,
                               int matrix_prec, double* re, double* im);

/* Arithmetic of subsequent uncontrolled k >= 5 passes on complex64 states
 * (QSIM_HIP_METHOD_*).  No reference counterpart: qsim's GPU path has only
 * FP32 vector arithmetic (lib/simulator_cuda_kernels.h:30-202). */
int qsim_hip_set_gate_method(int method);
int qsim_hip_get_gate_method(void);

/* ---- captured pass sequences ---------------------------------------------
 * A sequence of gate passes (any mix of qsim_hip_apply_gate,
 * apply_controlled_gate, apply_gate_block, block_run and state fills; no
 * call that returns data to the host) recorded once and replayed as one HIP
 * graph, with every gate matrix resident on the device.  Use:
 *   record_begin(&plan); <run the passes>; record_end(plan);
 *   capture_begin(plan); <run the same passes again>; capture_end(plan);
 *   launch(plan) ... launch(plan); free(plan);
 * The recorded run executes normally; while recording, each staged matrix
 * gets a device buffer of its own (kept by the plan) instead of a slot of
 * the shared staging ring.  The captured run must issue the same passes in
 * the same order (the same matrices): it reuses those buffers, stages
 * nothing and only records kernel launches (on a stream of the plan's own;
 * the state pointers are part of the graph).  capture_end fails with
 * QSIM_HIP_ERR_ARG if the sequences differ.  launch enqueues the graph on
 * the library stream (asynchronous, like a pass).  One recording or capture
 * at a time per process.  No single reference counterpart: it replaces
 * QSimRunner's per-gate host work (matrix upload and launch per fused gate,
 * lib/simulator_cuda.h:338-343, lib/run_qsim.h:128-160) for repeated runs
 * of a noiseless circuit. */
int qsim_hip_graph_record_begin(void** plan);
int qsim_hip_graph_record_end(void* plan);
int qsim_hip_graph_capture_begin(void* plan);
int qsim_hip_graph_capture_end(void* plan);
int qsim_hip_graph_launch(void* plan);
int qsim_hip_graph_free(void* plan);

/* ---- timing ------------------------------------------------------------ */

/* Records event `slot` (0..63) on the library stream. */
int qsim_hip_event_record(int slot);

/* Milliseconds between two recorded events (waits for the second). */
int qsim_hip_event_elapsed_ms(int slot_start, int slot_end, float* ms);

/* Per-pass timer: while started, every gate / controlled-gate pass records a
 * HIP event pair around its kernel on the library stream (up to `capacity`
 * passes).  qsim_hip_pass_timer_stop waits for the work, writes the kernel
 * durations in ms (in launch order) and stops recording.  No reference
 * counterpart: qsim times passes on the host with a device sync per gate
 * (lib/run_qsim.h:129-142). */
int qsim_hip_pass_timer_start(unsigned capacity);
int qsim_hip_pass_timer_stop(float* ms, unsigned max_count, unsigned* count);

/* ---- batched noisy trajectories -------------------------------------------
 * Many quantum trajectories (of possibly different qubit counts) run in ONE
 * kernel launch, one workgroup per trajectory, each with its own state in a
 * device pool and its own operation stream: fused gates (FP32 matrices, the
 * Kraus operators already sampled on the host folded in) and channel
 * decisions that need the state (taken on the device, in FP64).  At the end
 * every trajectory writes <Z_a Z_b> for its qubit pairs (FP64).  Replaces
 * the serial loop of QuantumTrajectorySimulator::RunBatch / RunIteration
 * (lib/qtrajectory.h:112-387), with the same sampling rule: a channel's
 * Kraus operator is chosen from the draw against probability bounds, and
 * against the probabilities <psi|K^dagger K|psi> of the (normalised) state
 * when the bounds do not decide (lib/qtrajectory.h:322-370). */
enum {
  QSIM_HIP_TRAJ_GATE = 0,
  /* A channel the bounds did not decide, followed by its k Kraus operators
   * (QSIM_HIP_TRAJ_KRAUS, 1 <= k <= 16, in the channel's order): q[0..moff2)
   * are the channel's qubits (moff2 = 1 or 2, ascending), r the draw, cp the
   * sum of the operators' probability bounds, slot the decision record
   * index, pad bit 0 = normalise the state first (the trajectory has
   * applied a non-unitary operator).  Walks the operators as RunIteration
   * does (lib/qtrajectory.h:330-370): unitary ones are skipped; for the
   * others cp += p_i - bound_i with p_i = <psi|K_i^dagger K_i|psi> (FP64
   * sums), the first with r < cp is applied -- or, at the last operator,
   * the likeliest one -- scaled to the normalised state. */
  QSIM_HIP_TRAJ_CHANNEL = 1,
  /* The header of a tile pass over the k gate operations that follow it
   * (1 <= k <= 32): q[0] | q[1] << 32 is the mask of its tile qubits (9..14
   * of them, starting with qubits 0, 1, 2), and every gate lies inside the
   * tile.  The tiled kernel applies the k gates with one stream of the
   * state (tiles staged in LDS, split-BF16 MFMA, as
   * qsim_hip_apply_gate_block does); the per-gate kernels ignore the header
   * and apply the gates one by one.  moff is not read. */
  QSIM_HIP_TRAJ_TILE_PASS = 2,
  /* A Kraus operator of the channel before it: K on q[0..k) (k = 0: the
   * identity) at matrix offset moff, K^dagger K over the channel's qubits at
   * moff2 (not read for a unitary operator), r its probability bound, pad
   * bit 0 = unitary. */
  QSIM_HIP_TRAJ_KRAUS = 3
};
/* qsim_hip_traj_run's `threads` value for the tiled kernel: 256 threads
 * per trajectory, or 512 when a tile pass has 14 qubits. */
enum { QSIM_HIP_TRAJ_TILED = 1 };

typedef struct {
  uint32_t kind;   /* QSIM_HIP_TRAJ_* */
  uint32_t k;      /* gate, Kraus: qubit count; channel, tile pass: number of
                      operations that follow it */
  uint32_t q[4];   /* gate, Kraus: ascending qubits; channel: its qubits;
                      tile pass: tile-qubit mask */
  uint32_t moff;   /* gate, Kraus: matrix offset (floats) */
  uint32_t moff2;  /* Kraus: K^dagger K offset; channel: qubit count */
  double r;        /* channel: the draw; Kraus: probability bound */
  double cp;       /* channel: sum of the probability bounds */
  uint32_t slot;   /* channel: index in the decision record */
  uint32_t pad;    /* channel: bit 0 normalise first; Kraus: bit 0 unitary */
} qsim_hip_traj_op;

typedef struct {
  uint32_t num_qubits;
  uint32_t num_pairs;     /* <Z_a Z_b> pairs at pairs[2 * pair_offset] */
  uint64_t state_offset;  /* complex amplitudes from the pool start */
  uint64_t op_begin;      /* operations [op_begin, op_end) */
  uint64_t op_end;
  uint32_t pair_offset;
  uint32_t pad;           /* bit 0: normalise the final state */
} qsim_hip_traj_desc;

/* Copies a batch to the device and allocates its state pool (states of all
 * trajectories at once).  Trajectories run in the order given: put the
 * largest first.  num_decisions: size of the decision record. */
int qsim_hip_traj_prepare(const qsim_hip_traj_desc* trajs, unsigned num_traj,
                          const qsim_hip_traj_op* ops, uint64_t num_ops,
                          const float* matrices, uint64_t num_floats,
                          const uint32_t* pairs, uint32_t num_pairs,
                          uint32_t num_decisions, void** handle);
/* One launch of the whole batch with `threads` (256, 512 or 1024) per
 * trajectory, or with the tiled kernel (QSIM_HIP_TRAJ_TILED); with
 * use_graph the launch is captured once in a hipGraph and replayed after.
 * Waits for completion. */
int qsim_hip_traj_run(void* handle, unsigned threads, int use_graph);
/* <Z_a Z_b> of every pair (FP64, pair order of the batch) and the decision
 * record (the chosen Kraus operator's index in its channel; -1: none chosen
 * or not reached). */
int qsim_hip_traj_results(void* handle, double* zz, uint32_t num_pairs,
                          int32_t* decisions, uint32_t num_decisions);
/* Final state of trajectory t (complex64, 2^n pairs) to host memory,
 * normalised when the descriptor asks for it (the scaling by the norm of
 * the last run is applied to the copy). */
int qsim_hip_traj_state(void* handle, unsigned t, float* out);
int qsim_hip_traj_free(void* handle);

/* ---- pass-time cost model ------------------------------------------------
 * A multilayer perceptron num_features -> hidden -> hidden -> hidden -> 1
 * (ReLU hidden layers) that predicts a pass's log time from features of the
 * pass; trained and evaluated on the device by the library's own kernels
 * (FP32, Adam, mean squared error).  Parameters are one flat array, layer by
 * layer: W (in x out, row-major), then b.  No reference counterpart: qsim's
 * MultiQubitGateFuser (lib/fuser_mqubit.h:185-526) picks its partition from
 * max_fused_size alone; the planner scores candidate partitions with this
 * model.  Not thread-safe per model. */
int qsim_hip_costmodel_create(unsigned num_features, unsigned hidden,
                              void** model);
int qsim_hip_costmodel_free(void* model);
uint64_t qsim_hip_costmodel_num_params(void* model);
int qsim_hip_costmodel_set_params(void* model, const float* params,
                                  uint64_t num_params);
int qsim_hip_costmodel_get_params(void* model, float* params,
                                  uint64_t num_params);
/* A model saved as JSON {"num_features": F, "hidden": H, "params": [...]}
 * (qsim_amd/data/cost_model.json), created on the device. */
int qsim_hip_costmodel_load(const char* path, void** model);
/* The process's default model, which the learned fuser uses when its
 * parameter names none (NULL: none; freeing the model clears it). */
int qsim_hip_costmodel_set_default(void* model);
void* qsim_hip_costmodel_default(void);
/* y[i] = network(x[i * num_features ...]) for i < batch (host arrays). */
int qsim_hip_costmodel_forward(void* model, const float* x, unsigned batch,
                               float* y);
/* Mean squared error of the batch against targets t and its gradient with
 * respect to the parameters (host arrays; grad may be NULL). */
int qsim_hip_costmodel_gradient(void* model, const float* x, const float* t,
                                unsigned batch, double* loss, float* grad);
/* Adam training over num_samples samples for `epochs` epochs of seeded
 * shuffled minibatches; final_loss = mean squared error over the set. */
int qsim_hip_costmodel_train(void* model, const float* x, const float* t,
                             unsigned num_samples, unsigned epochs,
                             unsigned batch, float lr, unsigned seed,
                             double* final_loss);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif  /* QSIM_HIP_H_ */
