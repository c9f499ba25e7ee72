"""Batched noisy trajectories (qsim_traj_*, one launch for the whole batch)
against qsim's own QuantumTrajectorySimulator on the CPU (oracle/_ref,
SimulatorAVX) with the same repetition ids.

Bars: the Kraus index of every channel identical to the CPU run (the host
replays the same mt19937 draws; the state-dependent damping decisions are
taken on the device); <Z_a Z_b> on grid neighbours within 2e-5 of the CPU
values (FP32 states on both sides, FP64 sums); graph replay bit-identical to
the plain launch.
"""

import os

import numpy as np
import pytest

import oracle
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
