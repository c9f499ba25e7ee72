"""Circuits for qsim's Python extensions, built through an extension's own
circuit API (pybind_interface/pybind_main.cpp:230-310: Circuit, add_gate,
add_matrix_gate, add_channel, control_last_gate, add_measurement, OpString),
as qsimcirq does, plus their dense FP64 meaning for small n.

Matrix conventions (lib/gates_cirq.h): a matrix gate's first qubit is the
most significant bit of its matrix index (MatrixGate reverses the list before
qsim's ascending sort); a state index has bit q <-> qubit q.
"""

from __future__ import annotations

import importlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
HIP_DIR = os.path.join(ROOT, "qsim_amd", "lib")

# The names GPU_MODULE_BINDINGS registers (pybind_interface/pybind_main.h:368-482).
GPU_BINDINGS = {
    "qsim_simulate", "qtrajectory_simulate", "qsim_simulate_fullstate",
    "qtrajectory_simulate_fullstate", "qsim_sample", "qsim_sample_final",
    "qtrajectory_sample", "qtrajectory_sample_final",
    "qsim_simulate_expectation_values",
    "qsim_simulate_moment_expectation_values",
    "qtrajectory_simulate_expectation_values",
    "qtrajectory_simulate_moment_expectation_values", "qsimh_simulate",
}


def load(name: str, directory: str):
    """Imports extension `name` from `directory` (None if not built)."""
    if not os.path.isdir(directory) or not any(
            f.startswith(name + ".") and f.endswith(".so")
            for f in os.listdir(directory)):
        return None
    if directory not in sys.path:
        sys.path.insert(0, directory)
    return importlib.import_module(name)


def random_unitary(k: int, rng) -> np.ndarray:
    a = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
    q, r = np.linalg.qr(a)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def flat(u: np.ndarray):
    return [float(x) for x in np.stack([u.real, u.imag], -1).reshape(-1)]


def _named(name, params):
    """Dense matrix (first qubit = most significant bit) of Cirq gates."""
    c, s = np.cos, np.sin
    if name == "kH":
        return np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    if name == "kX":
        return np.array([[0, 1], [1, 0]])
    if name == "krx":
        p = params["phi"]
        return np.array([[c(p / 2), -1j * s(p / 2)], [-1j * s(p / 2), c(p / 2)]])
    if name == "kT":
        return np.diag([1, np.exp(1j * np.pi / 4)])
    if name == "kCZ":
        return np.diag([1, 1, 1, -1])
    if name == "kCX":
        return np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]])
    if name == "kFSimGate":
        t, p = params["theta"], params["phi"]
        return np.array([[1, 0, 0, 0], [0, c(t), -1j * s(t), 0],
                         [0, -1j * s(t), c(t), 0], [0, 0, 0, np.exp(-1j * p)]])
    raise KeyError(name)


def build(m, n: int, depth: int, seed: int, measure_at_end: bool = False,
          controlled: bool = False, cut: int = None):
    """A random layered circuit on module m.  Returns (circuit, ops) with ops
    [(qubits, dense matrix, controls)] for the dense check.  With `cut`, two-
    qubit gates that cross the cut are CZ only (hybrid Schmidt decomposition)."""
    rng = np.random.default_rng(seed)
    circ = m.Circuit()
    circ.num_qubits = n
    ops = []
    t = 0
    for layer in range(depth):
        for q in range(n):
            kind = ["kH", "kX", "krx", "kT", "mat"][rng.integers(5)]
            if kind == "mat":
                u = random_unitary(1, rng)
                m.add_matrix_gate(t, [q], flat(u), circ)
            else:
                params = {"phi": float(rng.uniform(0, 6))} if kind == "krx" else {}
                m.add_gate(getattr(m.GateKind, kind), t, [q], params, circ)
                u = _named(kind, params)
            ops.append(([q], u, []))
        t += 1
        start = layer % 2
        for q0 in range(start, n - 1, 2):
            q1 = q0 + 1
            crosses = cut is not None and q0 < cut <= q1
            kind = "kCZ" if crosses else ["kCZ", "kCX", "kFSimGate", "mat"][rng.integers(4)]
            if kind == "mat":
                u = random_unitary(2, rng)
                m.add_matrix_gate(t, [q0, q1], flat(u), circ)
            else:
                params = ({"theta": float(rng.uniform(0, 3)),
                           "phi": float(rng.uniform(0, 3))}
                          if kind == "kFSimGate" else {})
                m.add_gate(getattr(m.GateKind, kind), t, [q0, q1], params, circ)
                u = _named(kind, params)
            ops.append(([q0, q1], u, []))
        t += 1
        if controlled and n >= 4 and layer % 3 == 2:
            m.add_gate(m.GateKind.kX, t, [n - 1], {}, circ)
            m.control_last_gate([0, 1], [1, 0], circ)
            ops.append(([n - 1], _named("kX", {}), [(0, 1), (1, 0)]))
            t += 1
    if measure_at_end:
        m.add_measurement(t, list(range(n)), circ)
    return circ, ops


def dense_run(n: int, ops) -> np.ndarray:
    """FP64 state of the ops from |0...0> (controls as (qubit, value))."""
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    for qs, u, ctrl in ops:
        t = psi.reshape([2] * n)          # axis a <-> bit n - 1 - a
        axes = [n - 1 - q for q in qs]    # qs[0] is the matrix's top bit
        k = len(qs)
        moved = np.moveaxis(t, axes, range(k))
        shape = moved.shape
        new = (u @ moved.reshape(1 << k, -1)).reshape(shape)
        new = np.moveaxis(new, range(k), axes).reshape(-1)
        if ctrl:
            idx = np.arange(1 << n)
            mask = np.ones(1 << n, dtype=bool)
            for q, v in ctrl:
                mask &= ((idx >> q) & 1) == v
            psi = np.where(mask, new, psi)
        else:
            psi = new
    return psi


def options(circ, **kw):
    o = {"c": circ, "f": 2, "v": 0, "s": 1, "z": 0, "t": 4}
    o.update(kw)
    return o
