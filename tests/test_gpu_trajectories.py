"""Batched noisy trajectories (qsim_traj_*, one launch for the whole batch)
against qsim's own QuantumTrajectorySimulator on the CPU (oracle/_ref,
SimulatorAVX) with the same repetition ids.

Bars: the Kraus index of every channel identical to the CPU run (the host
replays the same mt19937 draws; the state-dependent channel decisions are
taken on the device); <Z_a Z_b> on grid neighbours within 2e-5 of the CPU
values (FP32 states on both sides, FP64 sums); graph replay bit-identical to
the plain launch.  The tiled kernel (tile passes: runs of gates applied to
LDS-staged tiles in split-BF16 MFMA arithmetic, one stream of the state per
pass) is held to the same bars.  Noisy circuits come from the fixed
depolarise + damp rule and from noisy-circuit text with every channel kind
qsim builds (single-qubit channels, reset, 2-qubit Kraus sets); the full
1024-trajectory batch is checked on a sample of its trajectories.
"""

import ctypes
import os

import numpy as np
import pytest

import oracle
import qsim_amd
from helpers import random_matrix, to_interleaved
from qsim_amd import circuits, trajectories

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(oracle.REF_BIN),
                                 reason="reference driver not built")]


def reference(tmp_path, n, cycles, reps, p, gamma, seed=0):
    rows, cols = circuits.grid_shape(n)
    f = tmp_path / f"g{n}.txt"
    f.write_text(circuits.grid_circuit(n, cycles, seed))
    out = {}
    lo, hi = min(reps), max(reps) + 1
    runs, _ = oracle.run_reference_trajectories(str(f), rows, cols, lo, hi,
                                                p, gamma, max_fused=4,
                                                threads=4)
    for r in runs:
        out[r["rep"]] = r
    return out


def check_batch(tmp_path, sizes, cycles, reps_per_size, p, gamma,
                threads=512, kmax=4):
    b = trajectories.TrajectoryBatch(sizes, cycles, p_depol=p,
                                     gamma_damp=gamma)
    circuit_of, reps = [], []
    for ci, n in enumerate(sizes):
        for r in range(reps_per_size):
            circuit_of.append(ci)
            reps.append(1000 * ci + r)
    b.plan(circuit_of, reps, kmax)
    b.run(threads)
    refs = {ci: reference(tmp_path, n, cycles,
                          [1000 * ci + r for r in range(reps_per_size)], p,
                          gamma) for ci, n in enumerate(sizes)}
    worst, decisions = 0.0, 0
    for t, (ci, rep) in enumerate(zip(circuit_of, reps)):
        ref = refs[ci][rep]
        k = b.kraus(t)
        assert len(k) == b.num_channels(ci) == len(ref["kraus"])
        assert list(k) == ref["kraus"], (sizes[ci], rep)
        zz = b.zz(t)
        assert len(zz) == len(trajectories.grid_pairs(sizes[ci]))
        worst = max(worst, float(np.max(np.abs(zz - np.asarray(ref["zz"])))))
        decisions += int(np.sum(k[k >= 0] > 0))
    print(f"sizes {sizes}: {len(reps)} trajectories, {decisions} non-trivial "
          f"Kraus picks, max |dZZ| {worst:.2e}", flush=True)
    assert worst < 2e-5
    return b


def test_heavy_noise_matches_qtrajectory(tmp_path):
    # Strong noise: many Pauli errors and many state-dependent damping
    # decisions per trajectory.
    check_batch(tmp_path, [12, 14, 16], 8, 12, p=0.05, gamma=0.1)


@pytest.mark.parametrize("threads,kmax", [(256, 2), (1024, 3)])
def test_launch_shapes_agree(tmp_path, threads, kmax):
    check_batch(tmp_path, [12, 16], 6, 6, p=0.02, gamma=0.05,
                threads=threads, kmax=kmax)


def test_mixed_18_20_22_batch(tmp_path):
    # The configuration's sizes and noise, a few trajectories of each.
    check_batch(tmp_path, [18, 20, 22], 12, 4, p=1e-3, gamma=5e-3)


def test_graph_replay_is_identical():
    b = trajectories.TrajectoryBatch([12, 14], 6, p_depol=0.05, gamma_damp=0.1)
    b.plan([0, 1, 0, 1], [5, 6, 7, 8], 4)
    b.run(512)
    ref = [b.zz(t) for t in range(4)]
    kref = [b.kraus(t) for t in range(4)]
    for _ in range(3):
        b.run(512, use_graph=True)
        for t in range(4):
            assert np.array_equal(b.zz(t), ref[t])
            assert np.array_equal(b.kraus(t), kref[t])


def test_final_state_is_normalised():
    b = trajectories.TrajectoryBatch([12], 6, p_depol=0.05, gamma_damp=0.1)
    b.plan([0, 0], [1, 2], 4)
    b.run(256)
    for t in range(2):
        s = b.state(t).astype(np.float64)
        assert abs(np.sum(s * s) - 1.0) < 1e-5


TILED = trajectories.TILED


# 14-qubit tiles select the 512-thread shape; there the 12- and 13-qubit
# states run partial tiles.
@pytest.mark.parametrize("plan", [("tile", 9, 5), ("tile", 13, 5),
                                  ("tile", 11, 3), ("tile", 14, 4),
                                  ("tile", 14, 5), ("tile", 14, 9)])
def test_tiled_kernel_matches_qtrajectory(tmp_path, plan):
    check_batch(tmp_path, [12, 14, 16], 8, 8, p=0.05, gamma=0.1,
                threads=TILED, kmax=plan)


@pytest.mark.parametrize("tile_bits,low_bits", [(13, 5), (14, 5), (14, 4)])
def test_tiled_mixed_18_20_22_batch(tmp_path, tile_bits, low_bits):
    check_batch(tmp_path, [18, 20, 22], 12, 4, p=1e-3, gamma=5e-3,
                threads=TILED, kmax=("tile", tile_bits, low_bits))


def test_per_gate_kernels_run_tiled_plans():
    # The per-gate kernels skip the pass headers and apply the gates one by
    # one; a 6-qubit trajectory keeps fused gates in a tiled plan.  Same
    # Kraus choices and <ZZ> as the tiled kernel; final states within 1e-5
    # relative L2 (split-BF16 with three-part amplitudes against FP32
    # arithmetic over some 30 gates).
    b = trajectories.TrajectoryBatch([6, 12, 16], 6, p_depol=0.05,
                                     gamma_damp=0.1)
    co, reps = [0, 1, 2, 0, 1, 2], [11, 12, 13, 14, 15, 16]
    b.plan(co, reps, ("tile", 12, 5))
    b.run(TILED)
    tiled = [(b.zz(t), b.kraus(t), b.state(t)) for t in range(6)]
    b.run(512)
    for t in range(6):
        zz, k, st = tiled[t]
        assert np.array_equal(b.kraus(t), k)
        assert np.max(np.abs(b.zz(t) - zz)) < 2e-5
        ref = b.state(t)
        assert np.linalg.norm(st - ref) < 1e-5 * np.linalg.norm(ref)


@pytest.mark.parametrize("tile_bits", [10, 14])
def test_tiled_graph_replay_is_identical(tile_bits):
    # 10-qubit tiles run the 256-thread shape, 14-qubit tiles the 512-thread
    # one (the 12-qubit states as partial tiles).
    b = trajectories.TrajectoryBatch([12, 14], 6, p_depol=0.05, gamma_damp=0.1)
    b.plan([0, 1, 0, 1], [5, 6, 7, 8], ("tile", tile_bits, 4))
    b.run(TILED)
    ref = [b.zz(t) for t in range(4)]
    kref = [b.kraus(t) for t in range(4)]
    for _ in range(3):
        b.run(TILED, use_graph=True)
        for t in range(4):
            assert np.array_equal(b.zz(t), ref[t])
            assert np.array_equal(b.kraus(t), kref[t])


def test_replanning_reuses_the_batch():
    # A batch planned and run one way, then re-planned the other way (its
    # device buffers replaced), gives the same Kraus choices and <ZZ>.
    b = trajectories.TrajectoryBatch([12, 16], 6, p_depol=0.05, gamma_damp=0.1)
    co, reps = [0, 1, 1, 0], [21, 22, 23, 24]
    out = []
    for threads, plan in ((TILED, ("tile", 14, 5)), (512, 4), (TILED, ("tile", 12, 5))):
        b.plan(co, reps, plan)
        b.run(threads, use_graph=True)
        out.append([(b.kraus(t), b.zz(t)) for t in range(4)])
    for other in out[1:]:
        for (k0, z0), (k1, z1) in zip(out[0], other):
            assert np.array_equal(k0, k1)
            assert np.max(np.abs(z0 - z1)) < 2e-5
    b.free()


def _fmt_matrix(m):
    return " ".join(f"{x:.17g}" for v in np.asarray(m).reshape(-1)
                    for x in (v.real, v.imag))


def _damp(g):
    a0 = np.array([[1, 0], [0, np.sqrt(1 - g)]], complex)
    a1 = np.array([[0, np.sqrt(g)], [0, 0]], complex)
    return a0, a1


def _kron2(a, b):
    # Bit 0 of the index <-> the first (lower) qubit: K[r0 + 2 r1, c0 + 2 c1]
    # = a[r0, c0] b[r1, c1].
    return np.kron(b, a)


SINGLE = ["amplitude_damp 0.08", "phase_damp 0.1",
          "generalized_amplitude_damp 0.3 0.1", "bit_flip 0.05",
          "asymmetric_depolarize 0.02 0.03 0.04", "depolarize 0.05",
          "phase_flip 0.05", "amplitude_damp 0.2"]


def noisy_text(n, cycles, seed=0):
    """A noisy grid circuit as text (qsim_traj_create_noisy's format), the
    noise placed as qsim's MakeNoisy places it (lib/circuit_noisy.h:33-63):
    every gate is followed by a channel on each of its qubits, cycling
    through qsim's single-qubit channels (a reset on every 29th); after
    every cycle two 2-qubit Kraus channels on neighbouring qubits --
    correlated amplitude damping (four non-unitary operators) and a
    mixed-unitary II / ZZ / XX channel.  Times keep every qubit's operations
    in strict time order, as qsim's fuser requires of deferred gates
    (lib/fuser_mqubit.h:1220-1224): gate 4t, its channels 4t + 1, the
    cycle's two Kraus channels 4t + 2 and 4t + 3."""
    rows, cols = circuits.grid_shape(n)
    _, ops = circuits.grid_circuit_ops(rows, cols, cycles, seed)
    a0, a1 = _damp(0.1)
    damp2 = [(0, (1 - 0.1) ** 2, _kron2(a0, a0)), (0, 0.0, _kron2(a0, a1)),
             (0, 0.0, _kron2(a1, a0)), (0, 0.0, _kron2(a1, a1))]
    z, x, i2 = np.diag([1, -1]).astype(complex), np.array([[0, 1], [1, 0]],
                                                          complex), np.eye(2)
    mixed = [(1, 0.9, _kron2(i2, i2)), (1, 0.05, _kron2(z, z)),
             (1, 0.05, _kron2(x, x))]
    lines = [str(n)]
    count = 0
    last_t = None
    for t, name, qs, prm in ops:
        if last_t is not None and t // 2 != last_t // 2:
            c = last_t // 2
            a, b = (2 * c) % (n - 1), (2 * c) % (n - 1) + 1
            for dt, kops in ((2, damp2), (3, mixed)):
                body = " ".join(f"{u} {p!r} {_fmt_matrix(k)}"
                                for u, p, k in kops)
                lines.append(f"{4 * last_t + dt} channel kraus 2 {a} {b} "
                             f"{len(kops)} {body}")
        last_t = t
        lines.append(" ".join([str(4 * t), name] + [str(q) for q in qs] +
                              [repr(v) for v in prm]))
        for q in qs:
            count += 1
            kind = "reset" if count % 29 == 0 else SINGLE[count % len(SINGLE)]
            nm, *params = kind.split()
            lines.append(" ".join([str(4 * t + 1), "channel", nm, str(q)] +
                                  params))
    return "\n".join(lines) + "\n"


def check_noisy(tmp_path, sizes, cycles, reps_per_size, threads, kmax):
    texts = [noisy_text(n, cycles, seed=3) for n in sizes]
    shapes = [circuits.grid_shape(n) for n in sizes]
    b = trajectories.TrajectoryBatch.from_noisy(texts, shapes)
    circuit_of, reps = [], []
    for ci in range(len(sizes)):
        for r in range(reps_per_size):
            circuit_of.append(ci)
            reps.append(500 * ci + r)
    b.plan(circuit_of, reps, kmax)
    b.run(threads)
    worst, nontrivial = 0.0, 0
    for ci, (n, (rows, cols)) in enumerate(zip(sizes, shapes)):
        f = tmp_path / f"noisy{n}.txt"
        f.write_text(texts[ci])
        lo = 500 * ci
        runs, _ = oracle.run_reference_trajectories(
            str(f), rows, cols, lo, lo + reps_per_size, max_fused=4,
            threads=4, noisy_text=True)
        ref = {r["rep"]: r for r in runs}
        for t in range(len(reps)):
            if circuit_of[t] != ci:
                continue
            k = b.kraus(t)
            assert list(k) == ref[reps[t]]["kraus"], (n, reps[t])
            nontrivial += int(np.sum(k > 0))
            zz = b.zz(t)
            worst = max(worst, float(np.max(np.abs(
                zz - np.asarray(ref[reps[t]]["zz"])))))
    print(f"noisy {sizes}: {len(reps)} trajectories, {nontrivial} "
          f"non-trivial Kraus picks, max |dZZ| {worst:.2e}", flush=True)
    assert worst < 2e-5
    b.free()


@pytest.mark.parametrize("threads,kmax", [(512, 4), (TILED, ("tile", 10, 4))])
def test_any_channel_matches_qtrajectory(tmp_path, threads, kmax):
    # qsim's single-qubit channels, reset and 2-qubit Kraus sets (non-unitary
    # and mixed-unitary): identical Kraus choices (many of them decided on
    # the device from <psi|K^dagger K|psi>) and <Z_a Z_b> within 2e-5.
    check_noisy(tmp_path, [10, 12], 6, 12, threads, kmax)


def test_channels_beyond_the_device_are_refused():
    # A 3-qubit Kraus channel whose bounds do not decide every draw.
    n = 9
    k = np.eye(8, dtype=complex) * np.sqrt(0.5)
    body = " ".join(f"0 0 {_fmt_matrix(k)}" for _ in range(2))
    text = (f"{n}\n0 h 0\n1 channel kraus 3 0 1 2 2 {body}\n")
    b = trajectories.TrajectoryBatch.from_noisy([text], [(3, 3)])
    with pytest.raises(qsim_amd.QsimHipError) as e:
        b.plan([0], [7], 4)
    assert e.value.code == qsim_amd.ERR_UNSUPPORTED
    b.free()


def test_full_1024_batch_samples_match_qtrajectory(tmp_path):
    # The benchmark's batch as it runs there (1024 trajectories of 18, 20
    # and 22 qubits, 12 cycles, tiled 14/4 passes, one graph-replayed
    # launch); a sample of its trajectories against qsim's CPU runs of the
    # same repetitions.
    sizes = [18, 20, 22]
    b = trajectories.TrajectoryBatch(sizes, 12, p_depol=1e-3, gamma_damp=5e-3)
    ntraj = 1024
    circuit_of = [i % 3 for i in range(ntraj)]
    reps = list(range(ntraj))
    b.plan(circuit_of, reps, ("tile", 14, 4))
    b.run(TILED, use_graph=True)
    b.run(TILED, use_graph=True)
    sample = sorted(set([0, 1, 2, 3, 4, 5, 340, 341, 342, 511, 512, 513,
                         700, 701, 702, 1021, 1022, 1023] +
                        list(range(100, 1000, 97))))
    worst, checked = 0.0, 0
    for ci, n in enumerate(sizes):
        mine = [t for t in sample if circuit_of[t] == ci]
        rows, cols = circuits.grid_shape(n)
        f = tmp_path / f"g{n}.txt"
        f.write_text(circuits.grid_circuit(n, 12, 0))
        ref = {}
        for t in mine:
            runs, _ = oracle.run_reference_trajectories(
                str(f), rows, cols, t, t + 1, 1e-3, 5e-3, max_fused=4,
                threads=8)
            ref[t] = runs[0]
        for t in mine:
            assert list(b.kraus(t)) == ref[t]["kraus"], (n, t)
            worst = max(worst, float(np.max(np.abs(
                b.zz(t) - np.asarray(ref[t]["zz"])))))
            checked += 1
    print(f"1024-trajectory batch: {checked} sampled trajectories against "
          f"qsim's QuantumTrajectorySimulator, max |dZZ| {worst:.2e}",
          flush=True)
    assert worst < 2e-5
    b.free()


# The C-ABI structs (include/qsim_hip.h).
OP = np.dtype([("kind", "<u4"), ("k", "<u4"), ("q", "<u4", (4,)),
               ("moff", "<u4"), ("moff2", "<u4"), ("r", "<f8"),
               ("cp", "<f8"), ("slot", "<u4"), ("pad", "<u4")])
assert OP.itemsize == 56
DESC = np.dtype([("num_qubits", "<u4"), ("num_pairs", "<u4"),
                 ("state_offset", "<u8"), ("op_begin", "<u8"),
                 ("op_end", "<u8"), ("pair_offset", "<u4"), ("pad", "<u4")])


def raw_batch(n, ops, mats):
    d = np.zeros(1, DESC)
    d[0]["num_qubits"] = n
    d[0]["op_end"] = len(ops)
    h = ctypes.c_void_p()
    pairs = np.zeros(2, np.uint32)
    rc = qsim_amd.lib().qsim_hip_traj_prepare(
        d.ctypes.data, 1, ops.ctypes.data, len(ops), mats.ctypes.data,
        len(mats), pairs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 0,
        0, ctypes.byref(h))
    return rc, h


def tile_stream(mask, gates):
    """A tile-pass header (QSIM_HIP_TRAJ_TILE_PASS) and its gates."""
    ops = np.zeros(1 + len(gates), OP)
    ops["kind"][0] = 2
    ops["k"][0] = len(gates)
    ops["q"][0, :2] = [mask & 0xFFFFFFFF, mask >> 32]
    mats = []
    rng = np.random.default_rng(len(gates))
    for i, qs in enumerate(gates):
        ops["kind"][1 + i] = 0
        ops["k"][1 + i] = len(qs)
        ops["q"][1 + i, :len(qs)] = qs
        ops["moff"][1 + i] = sum(len(m) for m in mats)
        mats.append(to_interleaved(random_matrix(len(qs), rng), np.float32))
    return ops, np.concatenate(mats).astype(np.float32)


def test_tile_pass_header_runs_raw():
    # A hand-built stream: the tiled kernel and the per-gate kernel give
    # the same state.
    n = 12
    ops, mats = tile_stream(0b1111111111, [[3], [4, 9], [0, 1, 2, 5]])
    rc, h = raw_batch(n, ops, mats)
    assert rc == 0
    L = qsim_amd.lib()
    out = []
    for threads in (TILED, 256):
        assert L.qsim_hip_traj_run(h, threads, 0) == 0
        st = np.zeros(2 << n, np.float32)
        assert L.qsim_hip_traj_state(h, 0, st.ctypes.data) == 0
        out.append(st)
    L.qsim_hip_traj_free(h)
    assert np.linalg.norm(out[0] - out[1]) < 1e-5 * np.linalg.norm(out[1])


@pytest.mark.parametrize("bad", ["k0", "past_end", "mask_above_n", "small",
                                 "big", "no_low_run", "outside", "decision"])
def test_malformed_tile_passes_are_rejected(bad):
    n = 16
    mask = (1 << 10) - 1
    gates = [[3], [4, 9]]
    ops, mats = tile_stream(mask, gates)
    if bad == "k0":
        ops["k"][0] = 0
    elif bad == "past_end":
        ops["k"][0] = 3
    elif bad == "mask_above_n":
        ops["q"][0, 0] = mask | (1 << 16)
    elif bad == "small":
        ops["q"][0, 0] = 0b1000011111  # 6 tile qubits
    elif bad == "big":
        ops["q"][0, 0] = (1 << 15) - 1  # 15 tile qubits
    elif bad == "no_low_run":
        ops["q"][0, 0] = mask & ~0b100 | (1 << 12)
    elif bad == "outside":
        ops["q"][2, 1] = 11
    elif bad == "decision":
        ops["kind"][2] = 1
    rc, h = raw_batch(n, ops, mats)
    assert rc == qsim_amd.ERR_ARG, bad
    assert not h.value


@pytest.mark.parametrize("tile_bits", [13, 14])
def test_tile_pass_holds_the_resident_op_count(tile_bits):
    # A pass holds at most as many gates as the tiled kernel keeps resident
    # in LDS (qsim_hip_block_resident_ops): that many run and agree with the
    # per-gate kernel, one more is refused.
    n = 16
    L = qsim_amd.lib()
    limit = L.qsim_hip_block_resident_ops(tile_bits)
    assert limit >= 7
    mask = (1 << tile_bits) - 1
    rng = np.random.default_rng(tile_bits)
    gates = [sorted(rng.choice(tile_bits, 2, replace=False).tolist())
             for _ in range(limit + 1)]
    ops, mats = tile_stream(mask, gates[:limit])
    rc, h = raw_batch(n, ops, mats)
    assert rc == 0
    out = []
    for threads in (TILED, 256):
        assert L.qsim_hip_traj_run(h, threads, 0) == 0
        st = np.zeros(2 << n, np.float32)
        assert L.qsim_hip_traj_state(h, 0, st.ctypes.data) == 0
        out.append(st)
    L.qsim_hip_traj_free(h)
    assert np.linalg.norm(out[0] - out[1]) < 1e-5 * np.linalg.norm(out[1])
    ops, mats = tile_stream(mask, gates)
    rc, h = raw_batch(n, ops, mats)
    assert rc == qsim_amd.ERR_ARG
    assert not h.value


def test_tile_pass_that_could_grow_the_state_is_refused():
    # Tile passes hold unitary gates and Kraus operators (norm <= 1); gates
    # that could grow the state by 2^10 or more do not fit the FP16 tile
    # arithmetic and are refused at prepare time.
    n = 16
    ops, mats = tile_stream((1 << 10) - 1, [[3], [4, 9]])
    rc, h = raw_batch(n, ops, mats * np.float32(4096))
    assert rc == qsim_amd.ERR_UNSUPPORTED
    assert not h.value


def channel_stream(n, nq=1, kraus=2):
    """A channel decision (QSIM_HIP_TRAJ_CHANNEL) on qubits 0.. nq - 1 with
    `kraus` amplitude-damping-like operators (QSIM_HIP_TRAJ_KRAUS)."""
    ops = np.zeros(1 + kraus, OP)
    ops["kind"][0] = 1
    ops["k"][0] = kraus
    ops["moff2"][0] = nq
    ops["q"][0, :nq] = np.arange(nq)
    ops["r"][0] = 0.999
    ops["cp"][0] = 0.9
    dim = 1 << nq
    mats = []
    for i in range(kraus):
        ops["kind"][1 + i] = 3
        ops["k"][1 + i] = 1
        ops["q"][1 + i, 0] = 0
        ops["moff"][1 + i] = sum(len(m) for m in mats)
        mats.append(to_interleaved(np.eye(2) * 0.5, np.float32))
        ops["moff2"][1 + i] = sum(len(m) for m in mats)
        mats.append(to_interleaved(np.eye(dim) * 0.25, np.float32))
        ops["r"][1 + i] = 0.45
    return ops, np.concatenate(mats).astype(np.float32)


def raw_channel_batch(n, ops, mats, num_decisions):
    d = np.zeros(1, DESC)
    d[0]["num_qubits"] = n
    d[0]["op_end"] = len(ops)
    h = ctypes.c_void_p()
    pairs = np.zeros(2, np.uint32)
    rc = qsim_amd.lib().qsim_hip_traj_prepare(
        d.ctypes.data, 1, ops.ctypes.data, len(ops), mats.ctypes.data,
        len(mats), pairs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 0,
        num_decisions, ctypes.byref(h))
    return rc, h


@pytest.mark.parametrize("bad", ["ok", "k0", "past_end", "nq3", "slot",
                                 "kraus_alone", "outside", "matrix_end",
                                 "kraus_too_big"])
def test_malformed_channel_streams_are_rejected(bad):
    n = 8
    ops, mats = channel_stream(n)
    ndec = 1
    if bad == "k0":
        ops["k"][0] = 0
    elif bad == "past_end":
        ops["k"][0] = 3
    elif bad == "nq3":
        ops["moff2"][0] = 3
        ops["q"][0, :3] = [0, 1, 2]
    elif bad == "slot":
        ndec = 0
    elif bad == "kraus_alone":
        ops = ops[1:].copy()
    elif bad == "outside":
        ops["q"][2, 0] = 5
    elif bad == "matrix_end":
        ops["moff2"][2] = len(mats) - 2
    elif bad == "kraus_too_big":
        ops["k"][1] = 2
        ops["q"][1, :2] = [0, 1]
    rc, h = raw_channel_batch(n, ops, mats, ndec)
    if bad == "ok":
        assert rc == qsim_amd.OK
        qsim_amd.lib().qsim_hip_traj_free(h)
    else:
        assert rc == qsim_amd.ERR_ARG, bad
        assert not h.value
