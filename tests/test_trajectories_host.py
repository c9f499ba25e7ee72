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
