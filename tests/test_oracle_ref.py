"""Pins the oracle against qsim's own compiled CPU simulators.

Fixtures in tests/golden/reference_runs.json were written by
tests/golden/make_fixtures.py with oracle/_ref/qsim_ref (SimulatorAVX and
SimulatorBasic<double> driven by QSimRunner + MultiQubitGateFuser).
"""

import os

import numpy as np
import pytest

import oracle
from helpers import from_interleaved, reference_runs, rel_l2
from qsim_amd import circuits


def test_random_all_gate_circuits_fp64():
    # Every gate name of the qsim text format, with and without controls;
    # the reference ran them fused (max_fused_size 4) in FP64.
    for case in reference_runs()["random10"]:
        c = oracle.Circuit(case["circuit"], np.float64)
        s, _ = c.run()
        ref = np.array(case["state"])
        ref = ref[:, 0] + 1j * ref[:, 1]
        assert rel_l2(from_interleaved(s), ref) < 1e-12, case["seed"]


def test_grid20_state_fp32_and_fp64():
    g = reference_runs()["grid20"]
    txt = circuits.grid_circuit(20, 14, seed=0)
    for dtype, key, tol in ((np.float32, "avx", 2e-5), (np.float64, "basic64", 1e-12)):
        s, _ = oracle.Circuit(txt, dtype).run()
        c = from_interleaved(s)
        amps = g[key]["amplitudes"]
        idx = np.array([int(i) for i in amps])
        ref = np.array([amps[str(i)][0] + 1j * amps[str(i)][1] for i in idx])
        scale = np.sqrt(len(c) / 1.0) * np.abs(ref).max()
        assert np.max(np.abs(c[idx] - ref)) < tol * max(1.0, np.abs(ref).max()), key
        assert abs(np.sum(np.abs(c) ** 2) - g[key]["norm"]) < 1e-5
        assert abs(c.real.sum() - g[key]["sum_re"]) < 1e-3 * max(1, abs(g[key]["sum_re"]))


@pytest.mark.skipif(not os.path.exists(oracle.REF_BIN),
                    reason="reference driver not built (needs the qsim sources)")
def test_live_reference_random_circuit(tmp_path):
    txt = circuits.random_circuit(12, 200, seed=42, controlled_fraction=0.15)
    cf = tmp_path / "c.txt"
    cf.write_text(txt)
    sf = tmp_path / "s.bin"
    oracle.run_reference(str(cf), mode="basic64", max_fused=3, state_out=str(sf))
    ref = np.fromfile(sf, dtype=np.float64)
    s, _ = oracle.Circuit(txt, np.float64).run()
    assert rel_l2(s, ref) < 1e-12
    oracle.run_reference(str(cf), mode="avx", max_fused=5, state_out=str(sf))
    ref32 = np.fromfile(sf, dtype=np.float32)
    s32, _ = oracle.Circuit(txt, np.float32).run()
    assert rel_l2(s32, ref32) < 1e-5
