"""The circuit driver (qsim's parser + MultiQubitGateFuser over the backend):
host-side checks that need no GPU."""

import os

import numpy as np
import pytest

import oracle
from qsim_amd import circuits, driver

pytestmark = pytest.mark.skipif(not os.path.exists(driver.DRIVER_PATH),
                                reason="driver library not built")


def test_parse_and_fuse_grid_circuit():
    text = circuits.grid_circuit(20, 14, seed=0)
    c = driver.Circuit(text)
    assert c.num_qubits == 20
    assert c.num_gates == oracle.Circuit(text).num_ops
    for k in (2, 3, 4, 5, 6):
        n = c.fuse(k)
        assert n > 0
        gates = 0
        for qs, ng in c.passes():
            assert len(qs) <= k and qs == sorted(qs)
            gates += ng
        assert gates == c.num_gates   # every gate lands in exactly one pass


@pytest.mark.skipif(not os.path.exists(oracle.REF_BIN), reason="no reference")
def test_partition_matches_reference_fuser(tmp_path):
    # Same fuser, same circuit: the driver's pass count equals qsim's own.
    text = circuits.grid_circuit(24, 10, seed=3)
    cf = tmp_path / "c.txt"
    cf.write_text(text)
    c = driver.Circuit(text)
    for k in (2, 4, 6):
        n = c.fuse(k)
        info = oracle.run_reference(str(cf), mode="avx", max_fused=k, passes=0)
        assert info["fused_ops"] == n


def test_bad_circuit_rejected():
    with pytest.raises(ValueError):
        driver.Circuit("3\n0 h 7\n")
