"""Shared test helpers: golden fixtures, gate specs, dense references."""

from __future__ import annotations

import functools
import json
import os

import numpy as np

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@functools.lru_cache(None)
def known_answers():
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)


@functools.lru_cache(None)
def reference_runs():
    with open(os.path.join(GOLDEN, "reference_runs.json")) as f:
        return json.load(f)


def gate_op(spec, dtype):
    """(sorted qubits, matrix, sorted controls, cvals mask) of a fixture gate.

    Controls follow MakeControlledGate (lib/gate.h:94-148): (qubit, value)
    pairs sorted by qubit, value bit i <-> i-th sorted control.
    """
    if spec["name"] == "gph":
        qs = np.zeros(0, dtype=np.uint32)
        m = np.array(spec["params"], dtype=dtype)
    else:
        qs, m = oracle.gate_matrix(spec["name"], spec["qubits"],
                                   spec["params"], dtype=dtype)
    ctrl = spec.get("controls", [])
    vals = spec.get("cvals", [1] * len(ctrl))
    pairs = sorted(zip(ctrl, vals))
    cqs = np.array([p[0] for p in pairs], dtype=np.uint32)
    cmask = 0
    for i, (_, v) in enumerate(pairs):
        cmask |= (v & 1) << i
    return qs, m, cqs, cmask


def expectation2_circuit():
    """Gate list of TestExpectationValue2 (tests/simulator_testfixture.h:1474-1532):
    H on all 16 qubits, then depth/2 rounds of RX(0.1k), iSWAP pairs, RY(0.1k),
    shifted iSWAP pairs."""
    n, depth = 16, 16
    ops = [{"name": "h", "qubits": [k], "params": []} for k in range(n)]
    for _ in range(depth // 2):
        ops += [{"name": "rx", "qubits": [k], "params": [0.1 * k]} for k in range(n)]
        ops += [{"name": "is", "qubits": [2 * k, 2 * k + 1], "params": []}
                for k in range(n // 2)]
        ops += [{"name": "ry", "qubits": [k], "params": [0.1 * k]} for k in range(n)]
        ops += [{"name": "is", "qubits": [2 * k + 1, (2 * k + 2) % n], "params": []}
                for k in range(n // 2)]
    return n, ops


PAULI = {
    0: np.array([[0, 1], [1, 0]], dtype=np.complex128),
    1: np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    2: np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


def pauli_string_matrix(k):
    """Matrix of X,Y,Z,X,... on k consecutive qubits; bit j <-> j-th qubit."""
    m = np.array([[1.0]], dtype=np.complex128)
    for j in range(k):
        m = np.kron(PAULI[j % 3], m)
    return m


def to_interleaved(m, dtype):
    m = np.asarray(m, dtype=np.complex128)
    out = np.empty(m.size * 2, dtype=dtype)
    out[0::2] = m.real.reshape(-1)
    out[1::2] = m.imag.reshape(-1)
    return out


def from_interleaved(a):
    a = np.asarray(a, dtype=np.float64)
    return a[0::2] + 1j * a[1::2]


def dense_gate(n, qs, m):
    """Full 2^n x 2^n matrix of a gate: matrix bit j <-> qubit qs[j]."""
    k = len(qs)
    dim = 1 << k
    mc = from_interleaved(m).reshape(dim, dim)
    size = 1 << n
    full = np.zeros((size, size), dtype=np.complex128)
    for col in range(size):
        cj = 0
        for b, q in enumerate(qs):
            cj |= ((col >> q) & 1) << b
        base = col
        for q in qs:
            base &= ~(1 << q)
        for rj in range(dim):
            row = base
            for b, q in enumerate(qs):
                row |= ((rj >> b) & 1) << q
            full[row, col] += mc[rj, cj]
    return full


def random_matrix(k, rng, unitary=True):
    dim = 1 << k
    a = rng.normal(size=(dim, dim)) + 1j * rng.normal(size=(dim, dim))
    if unitary:
        q, r = np.linalg.qr(a)
        return q * (np.diag(r) / np.abs(np.diag(r)))
    return a / np.sqrt(dim)


def random_state(n, rng, dtype=np.float32):
    c = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    c /= np.linalg.norm(c)
    return to_interleaved(c, dtype)


def rel_l2(a, b):
    """||a - b|| / ||b|| for real or complex arrays (complex kept complex)."""
    a = np.asarray(a)
    b = np.asarray(b)
    if np.iscomplexobj(a) or np.iscomplexobj(b):
        a = a.astype(np.complex128)
        b = b.astype(np.complex128)
    else:
        a = a.astype(np.float64)
        b = b.astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
