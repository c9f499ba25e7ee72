"""qsim's own drivers (QSimRunner, ExpectationValue, QuantumTrajectorySimulator,
HybridSimulator, StateSpace::Sample/Measure) running on FactoryHIP, compared
in one binary against qsim's CPU SimulatorBasic Factory and qsim's known
answers (tests/dropin/dropin_check.cc)."""

import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from helpers import rel_l2
from qsim_amd import circuits

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "qsim_amd", "bin")


@pytest.fixture(scope="module")
def dropin_results():
    exe = os.path.join(ROOT, "tests", "dropin", "bin", "qsim_dropin_check")
    if not os.path.exists(exe):
        pytest.skip("drop-in binary not built (needs the qsim headers)")
    model = os.path.join(ROOT, "qsim_amd", "data", "cost_model.json")
    r = subprocess.run([exe, model], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout, r.stderr, flush=True)
    res = {}
    for line in r.stdout.splitlines():
        if line.startswith("{"):
            d = json.loads(line)
            res[d["scenario"]] = d
    return r.returncode, res


# learned: QSimRunner<IO, LearnedFuserHIP<IO>, FactoryHIP<float>> reaches
# blocked passes through SimulatorHIP::ApplyGate (the library's gate queue)
# and matches SimulatorBasic<double> within 1e-5 relative L2 on the 20-qubit
# grid, measurements included; learned-model: the committed cost model's
# own pick at 24 qubits within 1e-5 of the unfused run.
@pytest.mark.parametrize("scenario", ["runner", "callback", "expect",
                                      "qtrajectory", "hybrid", "sample",
                                      "learned", "learned-model"])
def test_qsim_driver_on_hip_factory(dropin_results, scenario):
    rc, res = dropin_results
    assert scenario in res, f"{scenario} did not report"
    assert res[scenario]["pass"], res[scenario]


def test_qsim_base_app(tmp_path):
    exe = os.path.join(BIN, "qsim_base_hip")
    if not os.path.exists(exe):
        pytest.skip("app not built")
    text = circuits.grid_circuit(20, 14, seed=0)
    cf = tmp_path / "c.txt"
    cf.write_text(text)
    sf = tmp_path / "s.bin"
    r = subprocess.run([exe, "-c", str(cf), "-f", "4", "-o", str(sf)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(sf, dtype=np.float32)
    ref, _ = oracle.Circuit(text, np.float32).run()
    assert rel_l2(got, ref) < 1e-5
    assert len(r.stdout.strip().splitlines()) == 8   # first 8 amplitudes
