"""Seeded random grid circuits in the qsim text format.

The format is qsim's circuit file format (docs/input_format.md in the qsim
sources; parsed by lib/circuit_qsim_parser.h:52-447): a qubit count line, then
``time gate qubits params`` lines.  Both the HIP backend's drivers and the CPU
checkers read the same text, so a circuit is defined once.

Construction (one cycle = two time steps):

* time ``2c``: one single-qubit gate per qubit, uniform over
  {x_1_2, y_1_2, hz_1_2} (sqrt(X), sqrt(Y), sqrt(W); lib/gates_qsim.h:190-220,
  :312-326), never the gate that qubit received in the previous cycle;
* time ``2c+1``: ``fs(theta=pi/2, phi=pi/6)`` (lib/gates_qsim.h:509-533) on
  every coupler of pattern ``c mod 4``, the patterns being
  A = horizontal couplers starting at an even column,
  B = horizontal couplers starting at an odd column,
  C = vertical couplers starting at an even row,
  D = vertical couplers starting at an odd row.

Qubit (r, c) of an R x C grid has index ``r*C + c``.  The RNG is Python's
``random.Random(seed)``, which is stable across platforms and versions.
"""

from __future__ import annotations

import math
import random
from typing import List, Tuple

ONE_QUBIT_GATES = ("x_1_2", "y_1_2", "hz_1_2")
FSIM_THETA = math.pi / 2
FSIM_PHI = math.pi / 6

# Grid shapes used for each qubit count (rows, cols).
GRID_SHAPES = {
    12: (3, 4), 16: (4, 4), 18: (3, 6), 20: (4, 5), 22: (2, 11), 24: (4, 6),
    25: (5, 5), 26: (2, 13), 27: (3, 9), 28: (4, 7), 29: (1, 29), 30: (5, 6),
    31: (1, 31), 32: (4, 8), 33: (3, 11), 34: (2, 17),
}


def grid_shape(num_qubits: int) -> Tuple[int, int]:
    if num_qubits in GRID_SHAPES:
        return GRID_SHAPES[num_qubits]
    # Most square factorisation.
    best = (1, num_qubits)
    for r in range(1, int(math.isqrt(num_qubits)) + 1):
        if num_qubits % r == 0:
            best = (r, num_qubits // r)
    return best


def coupler_pattern(rows: int, cols: int, pattern: int) -> List[Tuple[int, int]]:
    """Couplers of pattern 0..3 (A, B, C, D above) as (q0, q1) pairs."""
    pairs = []
    if pattern in (0, 1):
        for r in range(rows):
            for c in range(pattern, cols - 1, 2):
                pairs.append((r * cols + c, r * cols + c + 1))
    else:
        start = pattern - 2
        for r in range(start, rows - 1, 2):
            for c in range(cols):
                pairs.append((r * cols + c, (r + 1) * cols + c))
    return pairs


def grid_circuit_ops(rows: int, cols: int, cycles: int, seed: int = 0):
    """Returns (num_qubits, ops) with ops = [(time, name, qubits, params)]."""
    rng = random.Random(seed)
    n = rows * cols
    last = [None] * n
    ops = []
    for cyc in range(cycles):
        t = 2 * cyc
        for q in range(n):
            choices = [g for g in ONE_QUBIT_GATES if g != last[q]]
            g = choices[rng.randrange(len(choices))]
            last[q] = g
            ops.append((t, g, (q,), ()))
        for q0, q1 in coupler_pattern(rows, cols, cyc % 4):
            ops.append((t + 1, "fs", (q0, q1), (FSIM_THETA, FSIM_PHI)))
    return n, ops


def format_circuit(num_qubits: int, ops) -> str:
    lines = [str(num_qubits)]
    for t, name, qs, params in ops:
        parts = [str(t), name] + [str(q) for q in qs]
        parts += [repr(float(p)) for p in params]
        lines.append(" ".join(parts))
    return "\n".join(lines) + "\n"


def grid_circuit(num_qubits: int, cycles: int, seed: int = 0) -> str:
    """The random grid circuit on ``num_qubits`` qubits as qsim text."""
    rows, cols = grid_shape(num_qubits)
    n, ops = grid_circuit_ops(rows, cols, cycles, seed)
    return format_circuit(n, ops)


def random_circuit(num_qubits: int, num_gates: int, seed: int = 0,
                   gate_names=("h", "t", "x_1_2", "y_1_2", "hz_1_2", "rx", "ry",
                               "rz", "rxy", "s", "x", "y", "z", "cz", "cnot",
                               "sw", "is", "fs", "cp", "id1", "id2"),
                   controlled_fraction: float = 0.0,
                   max_controls: int = 2) -> str:
    """A random circuit over many gate types (for parity tests).

    Every gate gets its own time step, so any qubit order is legal.
    """
    rng = random.Random(seed)
    arity = {"h": (1, 0), "t": (1, 0), "x": (1, 0), "y": (1, 0), "z": (1, 0),
             "x_1_2": (1, 0), "y_1_2": (1, 0), "hz_1_2": (1, 0), "s": (1, 0),
             "rx": (1, 1), "ry": (1, 1), "rz": (1, 1), "rxy": (1, 2),
             "id1": (1, 0), "id2": (2, 0), "cz": (2, 0), "cnot": (2, 0),
             "sw": (2, 0), "is": (2, 0), "fs": (2, 2), "cp": (2, 1)}
    lines = [str(num_qubits)]
    for t in range(num_gates):
        name = gate_names[rng.randrange(len(gate_names))]
        nq, npar = arity[name]
        qs = rng.sample(range(num_qubits), nq)
        params = [round(rng.uniform(-3.0, 3.0), 6) for _ in range(npar)]
        ctrl = []
        if controlled_fraction > 0 and rng.random() < controlled_fraction:
            free = [q for q in range(num_qubits) if q not in qs]
            nc = rng.randint(1, min(max_controls, len(free)))
            ctrl = rng.sample(free, nc)
        parts = [str(t)]
        if ctrl:
            parts += ["c"] + [str(c) for c in ctrl]
        parts += [name] + [str(q) for q in qs] + [repr(p) for p in params]
        lines.append(" ".join(parts))
    return "\n".join(lines) + "\n"
