"""Writes the golden fixtures under tests/golden/ (run in the build container).

Sources, all read as text or run as the compiled reference driver:

* known-answer tests of qsim's C++ suite (tests/simulator_testfixture.h):
  the gate sequences and expected amplitudes of TestApplyGate1/2/3/5, the
  gates and ``expected_results`` of TestCircuitWithControlledGates, and the
  ``expected_real`` values of TestExpectationValue2 -- extracted as data;
* the compiled reference simulators (oracle/_ref/qsim_ref): final states of
  seeded grid circuits and random all-gate circuits (as amplitude subsets and
  checksums), sampled bitstrings, and mt19937 draws.

Usage: python tests/golden/make_fixtures.py [--ref /root/reference]
"""

from __future__ import annotations

import argparse
import json
import os
import re
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from qsim_amd import circuits  # noqa: E402

GATE_NAMES = {
    "GateHd": ("h", 1), "GateT": ("t", 1), "GateX": ("x", 1),
    "GateY": ("y", 1), "GateZ": ("z", 1), "GateX2": ("x_1_2", 1),
    "GateY2": ("y_1_2", 1), "GateRX": ("rx", 1), "GateRY": ("ry", 1),
    "GateRZ": ("rz", 1), "GateRXY": ("rxy", 1), "GateHZ2": ("hz_1_2", 1),
    "GateS": ("s", 1), "GateId1": ("id1", 1), "GateId2": ("id2", 2),
    "GateCZ": ("cz", 2), "GateCNot": ("cnot", 2), "GateSwap": ("sw", 2),
    "GateIS": ("is", 2), "GateFS": ("fs", 2), "GateCP": ("cp", 2),
    "GateGPh": ("gph", 0),
}


def _nums(s):
    return [float(x) for x in re.findall(r"-?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?", s)]


def _gate_from_create(cls, args):
    name, nq = GATE_NAMES[cls]
    vals = _nums(args)
    if name == "gph":
        # GateGPh::Create(time, phi) or (time, cos, sin)
        # (lib/gates_qsim.h:68-84).
        prm = vals[1:]
        if len(prm) == 1:
            prm = [float(np.cos(prm[0])), float(np.sin(prm[0]))]
        return {"name": "gph", "qubits": [], "params": prm}
    qubits = [int(v) for v in vals[1:1 + nq]]
    return {"name": name, "qubits": qubits, "params": vals[1 + nq:]}


def _body(text, func):
    i = text.index(f"void {func}(")
    j = text.find("\ntemplate <typename Factory>", i)
    return text[i:] if j < 0 else text[i:j]


def extract_apply_gate_tests(text):
    out = {}
    for func in ("TestApplyGate1", "TestApplyGate2", "TestApplyGate3",
                 "TestApplyGate5"):
        body = _body(text, func)
        nq = int(re.search(r"num_qubits = (\d+);", body).group(1))
        gates = {}
        for m in re.finditer(r"auto (gate\d+) = (Gate\w+)<fp_type>::Create\(([^;]*)\);",
                             body):
            gates[m.group(1)] = _gate_from_create(m.group(2), m.group(3))
        fused = {}
        for m in re.finditer(r"FusedGate<fp_type> (fgate\d+)\{[^{]*\{[^}]*\},\s*&gate\d+,"
                             r"\s*\{([^}]*)\}", body, re.S):
            fused[m.group(1)] = [g.strip().lstrip("&") for g in m.group(2).split(",")]
        steps = []
        tokens = re.finditer(
            r"ApplyGate\(simulator, (\w+), state\);|"
            r"GetAmpl\(state, (\d+)\);\s*EXPECT_NEAR\(std::real\(ampl\d+\), "
            r"([-\d.eE]+), [^)]*\);\s*EXPECT_NEAR\(std::imag\(ampl\d+\), ([-\d.eE]+)",
            body)
        for t in tokens:
            if t.group(1):
                names = fused.get(t.group(1), [t.group(1)])
                steps.append({"apply": [gates[n] for n in names]})
            else:
                steps.append({"check": [int(t.group(2)), float(t.group(3)),
                                        float(t.group(4))]})
        out[func] = {"num_qubits": nq, "steps": steps}
    return out


def extract_controlled_circuit(text):
    body = _body(text, "TestCircuitWithControlledGates")
    gates = []
    for m in re.finditer(r"gates\.push_back\((Gate\w+)<fp_type>::Create\(([^)]*)\)"
                         r"(?:\.ControlledBy\(\{([^}]*)\}(?:,\s*\{([^}]*)\})?\))?\);",
                         body):
        g = _gate_from_create(m.group(1), m.group(2))
        if m.group(3):
            g["controls"] = [int(x) for x in _nums(m.group(3))]
            if m.group(4):
                g["cvals"] = [int(x) for x in _nums(m.group(4))]
            else:
                g["cvals"] = [1] * len(g["controls"])
        gates.append(g)
    exp = re.search(r"expected_results = \{(.*?)\};", body, re.S).group(1)
    pairs = re.findall(r"\{([-\d.eE]+),\s*([-\d.eE]+)\}", exp)
    nq = int(re.search(r"num_qubits = (\d+);", body).group(1))
    return {"num_qubits": nq, "gates": gates,
            "expected": [[float(a), float(b)] for a, b in pairs]}


def extract_expectation2(text):
    body = _body(text, "TestExpectationValue2")
    exp = re.search(r"expected_real\[6\] = \{(.*?)\};", body, re.S).group(1)
    return {"num_qubits": 16, "depth": 16, "expected_real": _nums(exp)}


def ref_state(circuit_text, mode, max_fused, tmp, threads=8, samples=0,
              sample_seed=0):
    cf = os.path.join(tmp, "c.txt")
    sf = os.path.join(tmp, "s.bin")
    bf = os.path.join(tmp, "b.bin")
    with open(cf, "w") as f:
        f.write(circuit_text)
    info = oracle.run_reference(cf, mode=mode, max_fused=max_fused,
                                threads=threads, state_out=sf,
                                num_samples=samples, sample_seed=sample_seed,
                                bits_out=bf if samples else None)
    dt = np.float64 if mode == "basic64" else np.float32
    s = np.fromfile(sf, dtype=dt)
    bits = np.fromfile(bf, dtype=np.uint64) if samples else None
    return s, bits, info


def state_summary(s, picks):
    c = s.astype(np.float64).view(np.complex128)
    return {
        "amplitudes": {str(i): [float(c[i].real), float(c[i].imag)] for i in picks},
        "sum_abs2_weighted": float(np.sum(np.abs(c) ** 2 * np.arange(len(c)) / len(c))),
        "sum_re": float(c.real.sum()), "sum_im": float(c.imag.sum()),
        "norm": float(np.sum(np.abs(c) ** 2)),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    text = open(os.path.join(args.ref, "tests", "simulator_testfixture.h")).read()

    fx = {"source": "tests/simulator_testfixture.h of the qsim sources"}
    fx["apply_gate"] = extract_apply_gate_tests(text)
    fx["controlled_circuit"] = extract_controlled_circuit(text)
    fx["expectation2"] = extract_expectation2(text)
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(fx, f, indent=1)

    rng = np.random.default_rng(1234)
    ref = {"source": "oracle/_ref/qsim_ref (qsim SimulatorAVX / SimulatorBasic)"}
    with tempfile.TemporaryDirectory() as tmp:
        # Config-1 workload: 20q grid, 14 cycles, seed 0, fusion 2, samples.
        txt = circuits.grid_circuit(20, 14, seed=0)
        s32, bits, info = ref_state(txt, "avx", 2, tmp, samples=1000,
                                    sample_seed=0)
        s64, _, _ = ref_state(txt, "basic64", 2, tmp)
        picks = sorted(set(int(x) for x in rng.integers(0, 1 << 20, 64)) | {0, (1 << 20) - 1})
        ref["grid20"] = {
            "num_qubits": 20, "cycles": 14, "seed": 0, "max_fused": 2,
            "avx": state_summary(s32, picks), "basic64": state_summary(s64, picks),
            "samples_seed0": [int(b) for b in bits],
            "samples_norm": info["norm"],
        }
        # Random circuits over every gate name (and controls), 10 qubits.
        rnd = []
        for seed in range(4):
            txt = circuits.random_circuit(10, 120, seed=seed,
                                          controlled_fraction=0.2 if seed % 2 else 0.0)
            s64, _, _ = ref_state(txt, "basic64", 4, tmp)
            c = s64.view(np.complex128)
            rnd.append({"seed": seed, "circuit": txt,
                        "state": [[float(a.real), float(a.imag)] for a in c]})
        ref["random10"] = rnd
        # mt19937 draws of GenerateRandomValues (lib/util.h:67-85).
        dr = oracle.reference_draws(50, 7, 0.75, os.path.join(tmp, "d.bin"))
        ref["draws_seed7_max0.75"] = [float(x) for x in dr]
    with open(os.path.join(HERE, "reference_runs.json"), "w") as f:
        json.dump(ref, f)
    print("wrote", os.path.join(HERE, "known_answers.json"), "and reference_runs.json")


if __name__ == "__main__":
    main()
