"""Host side of the batched trajectory engine (no GPU): the noisy circuit it
builds has the channels qsim's trajectory simulator sees for the same rule
(one Kraus record per channel in oracle/_ref's collect_kop_stat output)."""

import os

import numpy as np
import pytest

import oracle
from qsim_amd import circuits, driver, trajectories

pytestmark = pytest.mark.skipif(not os.path.exists(driver.DRIVER_PATH),
                                reason="driver library not built")


@pytest.mark.parametrize("n,cycles", [(12, 4), (16, 7), (20, 3)])
def test_channel_count_matches_reference(tmp_path, n, cycles):
    b = trajectories.TrajectoryBatch([n], cycles)
    _, ops = circuits.grid_circuit_ops(*circuits.grid_shape(n), cycles, 0)
    gate_qubits = sum(len(q) for _, _, q, _ in ops)
    assert b.num_channels(0) == gate_qubits + n * cycles
    if os.path.exists(oracle.REF_BIN):
        f = tmp_path / "c.txt"
        f.write_text(circuits.grid_circuit(n, cycles, 0))
        runs, _ = oracle.run_reference_trajectories(
            str(f), *circuits.grid_shape(n), 0, 1, threads=2)
        assert len(runs[0]["kraus"]) == b.num_channels(0)
        assert len(runs[0]["zz"]) == len(trajectories.grid_pairs(n))


def test_bad_batches_are_rejected():
    with pytest.raises(ValueError):
        trajectories.TrajectoryBatch([12], 4, p_depol=1.5)


def _noisy_text():
    import test_gpu_trajectories as tg
    return tg.noisy_text


@pytest.mark.parametrize("n,cycles", [(10, 4), (12, 6)])
def test_noisy_text_channels_match_reference(tmp_path, n, cycles):
    # Noisy-circuit text with every channel kind: the batch parses it into
    # the channels qsim's trajectory simulator builds from the same text
    # (one Kraus record per channel), and every reference Kraus index is a
    # valid operator of its channel.
    text = _noisy_text()(n, cycles, seed=3)
    shape = circuits.grid_shape(n)
    b = trajectories.TrajectoryBatch.from_noisy([text], [shape])
    nch = sum(1 for line in text.splitlines() if " channel " in line)
    assert b.num_channels(0) == nch
    if os.path.exists(oracle.REF_BIN):
        f = tmp_path / "noisy.txt"
        f.write_text(text)
        runs, _ = oracle.run_reference_trajectories(
            str(f), *shape, 0, 3, max_fused=4, threads=2, noisy_text=True)
        for r in runs:
            assert len(r["kraus"]) == nch
            assert min(r["kraus"]) >= 0
            assert len(r["zz"]) == len(trajectories.grid_pairs(n))
    b.free()


@pytest.mark.parametrize("bad", [
    "5 channel depolarize 0\n",             # missing parameter
    "5 channel unknown_channel 0 0.1\n",    # unknown name
    "5 channel bit_flip 9 0.1\n",           # qubit out of range
    "5 channel kraus 2 1 0 1 1 0.5 " + " ".join(["0"] * 32) + "\n",  # order
    "5 channel kraus 1 0 1 1 0.5 1 0 0\n",  # short matrix
    "5 channel bit_flip 0 0.1 extra\n",     # trailing field
    "5 hh 0\n",                             # unknown gate
])
def test_malformed_noisy_text_is_refused(bad):
    with pytest.raises(ValueError):
        trajectories.TrajectoryBatch.from_noisy(["9\n0 h 0\n" + bad], [(3, 3)])
