"""Batched noisy trajectories on the GPU (include/qsim_hip_driver.h,
qsim_traj_*): many trajectories of uneven qubit counts in one launch.

A noisy circuit is either a qsim grid circuit with a depolarising channel
after every gate (on each of its qubits) and an amplitude-damping channel
on every qubit after each cycle (TrajectoryBatch), or any noisy-circuit text
of qsim gates and channels (TrajectoryBatch.from_noisy; the format is
qsim_traj_create_noisy's).  Trajectory sampling follows qsim's RunIteration
(lib/qtrajectory.h:240-387) with mt19937(rep) draws, so the Kraus choices
equal the CPU trajectory simulator's for the same repetition ids.
"""

from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _check, circuits
from .driver import driver_lib


# qsim_hip_traj_run's kernel choice for tile passes (QSIM_HIP_TRAJ_TILED).
TILED = 1

# Autotuner candidates: (threads, plan); plan is a fusion size or a tile
# spec ("tile", tile_bits, low_bits).
DEFAULT_CANDIDATES = ((TILED, ("tile", 14, 5)), (TILED, ("tile", 14, 4)),
                      (TILED, ("tile", 14, 3)), (TILED, ("tile", 13, 4)),
                      (TILED, ("tile", 12, 5)), (256, 4), (512, 4),
                      (1024, 4), (1024, 3))


class TrajectoryBatch:
    """A set of noisy grid circuits and a planned batch of trajectories."""

    def __init__(self, num_qubits_list, cycles, seed=0, p_depol=1e-3,
                 gamma_damp=5e-3):
        self.sizes = list(num_qubits_list)
        texts, rows, cols = [], [], []
        for n in self.sizes:
            r, c = circuits.grid_shape(n)
            texts.append(circuits.grid_circuit(n, cycles, seed).encode())
            rows.append(r)
            cols.append(c)
        arr = (ctypes.c_char_p * len(texts))(*texts)
        r = np.asarray(rows, np.uint32)
        c = np.asarray(cols, np.uint32)
        self.h = driver_lib().qsim_traj_create(
            ctypes.cast(arr, ctypes.c_void_p),
            len(texts), r.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
            c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), p_depol,
            gamma_damp)
        if not self.h:
            raise ValueError("cannot build the noisy circuits")
        self.texts = [t.decode() for t in texts]
        self.cycles, self.seed = cycles, seed
        self.p_depol, self.gamma_damp = p_depol, gamma_damp
        self.circuit_of = None
        self.reps = None

    @classmethod
    def from_noisy(cls, texts, shapes):
        """A batch over noisy-circuit texts (qsim_traj_create_noisy), each
        on a rows x cols grid (``shapes``: (rows, cols) per text) whose
        neighbour pairs get <Z_a Z_b>."""
        self = cls.__new__(cls)
        enc = [t.encode() for t in texts]
        arr = (ctypes.c_char_p * len(enc))(*enc)
        r = np.asarray([s[0] for s in shapes], np.uint32)
        c = np.asarray([s[1] for s in shapes], np.uint32)
        self.h = driver_lib().qsim_traj_create_noisy(
            ctypes.cast(arr, ctypes.c_void_p), len(enc),
            r.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
            c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        if not self.h:
            raise ValueError("cannot parse the noisy circuits")
        self.sizes = [int(a) * int(b) for a, b in shapes]
        self.texts = list(texts)
        self.cycles = self.seed = None
        self.p_depol = self.gamma_damp = None
        self.circuit_of = None
        self.reps = None
        return self

    def num_channels(self, circuit: int) -> int:
        return driver_lib().qsim_traj_num_channels(self.h, circuit)

    def plan(self, circuit_of, reps, max_fused_size=4) -> float:
        """Host planning (sampling + fusion); returns its wall seconds.

        ``max_fused_size`` is 1..4 (fused gates, one sweep of the state
        each), or a tile spec ``("tile", tile_bits, low_bits)``: runs of
        gates become tile passes (one sweep per pass) for the tiled kernel
        (``run(TILED)``)."""
        co = np.ascontiguousarray(circuit_of, np.uint32)
        rp = np.ascontiguousarray(reps, np.uint64)
        cp = co.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        t0 = time.time()
        if isinstance(max_fused_size, tuple):
            kind, tile_bits, low_bits = max_fused_size
            if kind != "tile":
                raise ValueError(f"unknown plan {max_fused_size!r}")
            _check(driver_lib().qsim_traj_plan_tiled(
                self.h, cp, rp.ctypes.data, len(co), tile_bits, low_bits))
        else:
            _check(driver_lib().qsim_traj_plan(
                self.h, cp, rp.ctypes.data, len(co), max_fused_size))
        self.circuit_of, self.reps = co, rp
        return time.time() - t0

    def run(self, threads=512, use_graph=False):
        """One launch of the planned batch: ``threads`` 256, 512 or 1024
        (per-gate kernels) or ``TILED`` (tile passes, 256 threads)."""
        _check(driver_lib().qsim_traj_run(self.h, threads, int(use_graph)))

    def zz(self, t: int) -> np.ndarray:
        out = np.zeros(256, np.float64)
        num = ctypes.c_uint32(0)
        _check(driver_lib().qsim_traj_zz(self.h, t, out.ctypes.data, 256,
                                         ctypes.byref(num)))
        return out[:num.value].copy()

    def kraus(self, t: int) -> np.ndarray:
        cap = 1 << 16
        out = np.zeros(cap, np.int32)
        num = ctypes.c_uint32(0)
        _check(driver_lib().qsim_traj_kraus(self.h, t, out.ctypes.data, cap,
                                            ctypes.byref(num)))
        return out[:num.value].copy()

    def state(self, t: int) -> np.ndarray:
        n = self.sizes[int(self.circuit_of[t])]
        out = np.zeros(2 << n, np.float32)
        _check(driver_lib().qsim_traj_state(self.h, t, out.ctypes.data))
        return out

    def autotune(self, candidates=DEFAULT_CANDIDATES, sample=None, log=None):
        """Times each (threads, plan) -- a per-gate kernel shape with a
        fusion size, or the tiled kernel with a tile shape -- on a sample
        of the planned trajectories (the whole batch by default: one
        workgroup runs one trajectory, so a small sample measures latency,
        not throughput) and returns the fastest; the batch is left planned
        that way."""
        co, rp = self.circuit_of, self.reps
        if sample is not None and sample < len(co):
            idx = np.linspace(0, len(co) - 1, sample).astype(int)
            sco, srp = co[idx], rp[idx]
        else:
            sco, srp = co, rp
        best = None
        results = []
        for threads, plan in candidates:
            self.plan(sco, srp, plan)
            self.run(threads)
            t0 = time.time()
            self.run(threads)
            dt = time.time() - t0
            kernel = "tiled" if threads == TILED else f"{threads} threads"
            results.append({"kernel": kernel, "plan": describe_plan(plan),
                            "s": dt})
            if log:
                log(f"autotune {kernel}, {describe_plan(plan)}: "
                    f"{dt * 1e3:.1f} ms")
            if best is None or dt < best[2]:
                best = (threads, plan, dt)
        self.plan(co, rp, best[1])
        return best[0], best[1], results

    def free(self):
        if getattr(self, "h", None):
            driver_lib().qsim_traj_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def describe_plan(plan) -> str:
    if isinstance(plan, tuple):
        return f"tile passes {plan[1]}/{plan[2]}"
    return f"fused k={plan}"


def grid_pairs(num_qubits):
    """<Z_a Z_b> pair order of the batch: per grid cell, the horizontal then
    the vertical neighbour."""
    rows, cols = circuits.grid_shape(num_qubits)
    pairs = []
    for r in range(rows):
        for c in range(cols):
            if c + 1 < cols:
                pairs.append((r * cols + c, r * cols + c + 1))
            if r + 1 < rows:
                pairs.append((r * cols + c, (r + 1) * cols + c))
    return pairs
