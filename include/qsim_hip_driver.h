/* qsim_hip_driver.h -- C ABI of the circuit driver built on qsim's own host
 * code (circuit parser, MultiQubitGateFuser, gate dispatch) and the gfx950
 * backend.  It is what bench.py and the tests call to run whole circuits.
 *
 * A circuit handle owns the parsed circuit (in FP32 and FP64 gate matrices)
 * and the current fused partition.  States are device pointers from
 * qsim_hip_alloc.
 */
#ifndef QSIM_HIP_DRIVER_H_
#define QSIM_HIP_DRIVER_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Fuser selection for qsim_driver_fuse. */
enum {
  QSIM_DRIVER_FUSER_MQUBIT = 0,  /* qsim's MultiQubitGateFuser */
  QSIM_DRIVER_FUSER_LEARNED = 1, /* cost-model planner (qsim_driver_plan_learned) */
  QSIM_DRIVER_FUSER_NONE = 2     /* no fusion: every operation of the circuit
                                    is its own pass, as a QSimRunner loop over
                                    the unfused operations would apply them
                                    (max_fused_size is ignored) */
};

/* Parses qsim's circuit text format up to `maxtime` (CircuitQsimParser::
 * FromStream, lib/circuit_qsim_parser.h:52-147).  NULL on a parse error. */
void* qsim_driver_load(const char* text, unsigned maxtime);
void qsim_driver_free(void* circuit);
unsigned qsim_driver_num_qubits(void* circuit);
unsigned qsim_driver_num_gates(void* circuit);

/* Fuses the circuit into passes of at most max_fused_size qubits
 * (MultiQubitGateFuser::FuseGates, lib/fuser_mqubit.h:185-526).  With
 * matrix_prec == QSIM_HIP_F64 the fused matrices are the products of the
 * FP64 gate matrices (lib/fuser.h:149-179 evaluated in double). */
int qsim_driver_fuse(void* circuit, unsigned max_fused_size, int fuser,
                     int matrix_prec);

/* Partitions the circuit into cache-blocked passes of `tile_bits` (9..14)
 * tile qubits that always include the `low_bits` (3..tile_bits) lowest
 * qubits (qsim_amd/block_planner.h); the gates of a pass are fused to
 * <= 4 qubits by MultiQubitGateFuser with FP64 matrices.  Operations that
 * are not uncontrolled gates on <= 4 qubits stay single passes.  No
 * reference counterpart: it replaces the partition of FuseGates
 * (lib/fuser_mqubit.h:185-526) and the per-gate loop of QSimRunner::Run
 * (lib/run_qsim.h:264-280).  QSIM_HIP_ERR_UNSUPPORTED below 9 qubits. */
int qsim_driver_fuse_blocked(void* circuit, unsigned tile_bits,
                             unsigned low_bits);

unsigned qsim_driver_num_passes(void* circuit);

/* Kinds of passes (qsim_driver_pass_kind). */
enum {
  QSIM_DRIVER_PASS_GATE = 0,         /* a fused gate */
  QSIM_DRIVER_PASS_MEASUREMENT = 1,  /* a (fused) measurement */
  QSIM_DRIVER_PASS_OPERATION = 2,    /* an unfused operation (controlled gate) */
  QSIM_DRIVER_PASS_BLOCKED = 3       /* a cache-blocked pass */
};
int qsim_driver_pass_kind(void* circuit, unsigned i);

/* Pass i: its qubit count, qubits (ascending, up to 64) and gate count.  For
 * a blocked pass: the tile qubits and the circuit gates it applies. */
int qsim_driver_pass_info(void* circuit, unsigned i, unsigned* k,
                          unsigned* qubits, unsigned* num_gates);

/* Blocked pass i: its fused gates (<= 32): qubit counts, 4 qubit slots
 * each and, if op_matrices is not NULL, their FP64 matrices concatenated
 * (2 * 4^k doubles each; at most 32 * 512). */
int qsim_driver_block_info(void* circuit, unsigned i, unsigned* num_ops,
                           unsigned* op_k, unsigned* op_qubits,
                           double* op_matrices);

/* Gate method the current partition runs with (QSIM_HIP_METHOD_*, or -1 for
 * the library's current one); set for the run only and restored after. */
int qsim_driver_set_method(void* circuit, int method);

/* Arithmetic of the circuit's blocked passes: FP16 with per-tile scaling
 * (the default) or split-BF16 (three-part amplitudes in the circuit's first
 * blocked pass); see qsim_hip_apply_gate_block. */
enum { QSIM_DRIVER_BLOCK_F16 = 0, QSIM_DRIVER_BLOCK_BF16 = 1 };
int qsim_driver_set_block_arith(void* circuit, int arith);

/* Cost-model features (include/qsim_amd/learned_planner.h, 19 floats) of
 * pass i of the current partition when run with gate method `method`. */
int qsim_driver_pass_features(void* circuit, unsigned i, int method,
                              float* out);

/* The learned planner: builds the candidate partitions of
 * qsim::hip::DefaultCandidates(), predicts each pass's time with the cost
 * model (qsim_hip_costmodel_*, log ms per 2^24 amplitudes), and leaves the
 * circuit partitioned by the candidate with the least predicted total.
 * model == NULL selects the fallback, qsim's fuser at max_fused_size 4.
 * Per candidate (up to max_candidates entries, any pointer may be NULL):
 * blocked flag, a (max_fused_size or tile qubits), b (low tile qubits),
 * gate method, predicted ms (NaN without a model), pass count.  *chosen is
 * the index of the candidate in use.  No reference counterpart: it replaces
 * the fixed max_fused_size of MultiQubitGateFuser::FuseGates
 * (lib/fuser_mqubit.h:185-526) as called by QSimRunner::Run
 * (lib/run_qsim.h:61-289). */
int qsim_driver_plan_learned(void* circuit, void* model,
                             unsigned max_candidates, int* blocked,
                             unsigned* a, unsigned* b, int* methods,
                             double* predicted_ms, unsigned* num_passes,
                             unsigned* num_candidates, int* chosen);

/* Model used by qsim_driver_fuse(..., QSIM_DRIVER_FUSER_LEARNED, ...)
 * (NULL: the fallback partition). */
void qsim_driver_set_cost_model(void* model);

/* ---- batched noisy trajectories (qsim_amd/apps/trajectories.cc) ---------
 * Noisy circuits on rows x cols grids of qubits: gates and qsim channels
 * (Channel, lib/channel.h:30-125).  Each trajectory replays RunIteration's
 * sampling (lib/qtrajectory.h:240-387) from mt19937(rep) on the host where
 * the probability bounds decide it; the rest is decided on the device
 * (QSIM_HIP_TRAJ_CHANNEL; channels on at most 2 qubits with at most 16
 * Kraus operators).  All trajectories then run in one launch
 * (qsim_hip_traj_run).  Replaces QuantumTrajectorySimulator::RunBatch
 * (lib/qtrajectory.h:112-175) over reps; results are <Z_a Z_b> on grid
 * neighbours (horizontal pairs then vertical ones, row by row) and the
 * Kraus index of every channel (collect_kop_stat order without the gates'
 * zeros; -1 where a device-decided channel chose none).
 *
 * qsim_traj_create: qsim circuits made noisy by one fixed rule -- a
 * depolarising channel (p_depol) on each qubit of every gate after it, an
 * amplitude-damping channel (gamma_damp) on every qubit after each cycle
 * (gate times 2c, 2c + 1).  NULL on a parse error. */
void* qsim_traj_create(const char* const* circuits, unsigned num_circuits,
                       const unsigned* rows, const unsigned* cols,
                       double p_depol, double gamma_damp);
/* Noisy circuits given as text: qsim's circuit format
 * (lib/circuit_qsim_parser.h) with channel lines among the gate lines, in
 * the order the operations apply:
 *   <time> channel <name> <qubit> <params>, name one of depolarize p,
 *       asymmetric_depolarize px py pz, amplitude_damp gamma,
 *       generalized_amplitude_damp p gamma, phase_damp gamma, phase_flip p,
 *       bit_flip p, reset (built by qsim's factories, lib/channels_cirq.h);
 *   <time> channel kraus <nq> <qubits, ascending> <number of operators>,
 *       then per operator <unitary 0|1> <probability bound> and its matrix
 *       (2 4^nq numbers, row-major complex; bit j of the index <-> the j-th
 *       qubit) -- K^dagger K as KrausOperator::CalculateKdKMatrix forms it.
 * Gates and channels keep their times; gates act on at most 4 qubits.  A
 * channel beyond the device's limits plans only where its bounds decide
 * (QSIM_HIP_ERR_UNSUPPORTED from qsim_traj_plan otherwise).  NULL on a
 * parse error. */
void* qsim_traj_create_noisy(const char* const* texts, unsigned num_circuits,
                             const unsigned* rows, const unsigned* cols);
unsigned qsim_traj_num_channels(void* batch, unsigned circuit);
/* Trajectory t runs circuit circuit_of[t] with repetition id reps[t]; runs
 * of deferred gates are fused to at most max_fused_size (1..4) qubits. */
int qsim_traj_plan(void* batch, const unsigned* circuit_of,
                   const uint64_t* reps, unsigned num_traj,
                   unsigned max_fused_size);
/* The same sampling with every run of deferred gates partitioned into tile
 * passes (QSIM_HIP_TRAJ_TILE_PASS) of tile_bits (9..14) qubits, low_bits
 * (3..9) of them the lowest, by the blocked-pass planner; states under 9
 * qubits keep fused gates of at most 4 qubits.  Run such a batch with
 * threads = QSIM_HIP_TRAJ_TILED (the per-gate kernels also accept it). */
int qsim_traj_plan_tiled(void* batch, const unsigned* circuit_of,
                         const uint64_t* reps, unsigned num_traj,
                         unsigned tile_bits, unsigned low_bits);
/* threads: 256, 512, 1024 (per-gate kernels) or QSIM_HIP_TRAJ_TILED. */
int qsim_traj_run(void* batch, unsigned threads, int use_graph);
int qsim_traj_zz(void* batch, unsigned t, double* zz, unsigned max,
                 unsigned* num);
int qsim_traj_kraus(void* batch, unsigned t, int32_t* kidx, unsigned max,
                    unsigned* num);
int qsim_traj_state(void* batch, unsigned t, float* out);
void qsim_traj_free(void* batch);

/* Applies passes [first, first + count) to a state of the circuit's size
 * (the apply loop of QSimRunner::Run, lib/run_qsim.h:264-280); measurement
 * passes draw from mt19937(seed).  Asynchronous for gate passes. */
int qsim_driver_run_passes(void* circuit, void* state, int prec,
                           unsigned first, unsigned count, unsigned seed);

/* All passes. */
int qsim_driver_run(void* circuit, void* state, int prec, unsigned seed);

/* The whole circuit as one replayable HIP graph on `state` (complex64 or
 * complex128): runs every pass once (recording, the state advances as in
 * qsim_driver_run), then captures the same passes with their matrices and
 * blocked-pass fragments resident on the device (qsim_hip_graph_*).  Each
 * qsim_hip_graph_launch(*plan) then applies the circuit to `state` again in
 * one launch.  Circuits with measurements are refused
 * (QSIM_HIP_ERR_UNSUPPORTED: their passes depend on host draws).  The plan
 * owns every matrix and blocked-pass fragment the graph reads, so it stays
 * valid after the circuit is re-fused or freed; free it with
 * qsim_hip_graph_free. */
int qsim_driver_capture(void* circuit, void* state, int prec, void** plan);

#ifdef __cplusplus
}
#endif

#endif /* QSIM_HIP_DRIVER_H_ */
