"""Whole-circuit runs through libqsim_hip_driver.so (include/qsim_hip_driver.h).

The driver library is qsim's own host code -- CircuitQsimParser,
MultiQubitGateFuser and the QSimRunner apply loop -- compiled over the gfx950
backend's C++ adapter (include/qsim_amd/).  This module is its ctypes face.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import F32, F64, State, _check, lib

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER_PATH = os.path.join(HERE, "lib", "libqsim_hip_driver.so")

FUSER_MQUBIT = 0
FUSER_LEARNED = 1
FUSER_NONE = 2

NUM_FEATURES = 19  # include/qsim_amd/learned_planner.h

PASS_GATE = 0
PASS_MEASUREMENT = 1
PASS_OPERATION = 2
PASS_BLOCKED = 3

_dll = None


def driver_lib():
    global _dll
    if _dll is None:
        lib()  # loads libqsim_hip.so first (the driver links to it)
        if not os.path.exists(DRIVER_PATH):
            raise FileNotFoundError(
                f"{DRIVER_PATH} is missing: it is built by "
                "`python -m qsim_amd.build` where qsim's headers are present")
        d = ctypes.CDLL(DRIVER_PATH)
        u, i, vp = ctypes.c_uint, ctypes.c_int, ctypes.c_void_p
        up = ctypes.POINTER(ctypes.c_uint32)
        for name, res, args in (
                ("qsim_driver_load", vp, [ctypes.c_char_p, u]),
                ("qsim_driver_free", None, [vp]),
                ("qsim_driver_num_qubits", u, [vp]),
                ("qsim_driver_num_gates", u, [vp]),
                ("qsim_driver_fuse", i, [vp, u, i, i]),
                ("qsim_driver_fuse_blocked", i, [vp, u, u]),
                ("qsim_driver_num_passes", u, [vp]),
                ("qsim_driver_pass_kind", i, [vp, u]),
                ("qsim_driver_pass_info", i, [vp, u, up, up, up]),
                ("qsim_driver_block_info", i, [vp, u, up, up, up,
                                               ctypes.POINTER(ctypes.c_double)]),
                ("qsim_driver_set_method", i, [vp, i]),
                ("qsim_driver_set_block_arith", i, [vp, i]),
                ("qsim_driver_pass_features", i, [vp, u, i, vp]),
                ("qsim_driver_plan_learned", i, [vp, vp, u, vp, vp, vp, vp,
                                                 vp, vp, up, vp]),
                ("qsim_driver_set_cost_model", None, [vp]),
                ("qsim_traj_create", vp, [vp, u, up, up, ctypes.c_double,
                                          ctypes.c_double]),
                ("qsim_traj_create_noisy", vp, [vp, u, up, up]),
                ("qsim_traj_num_channels", u, [vp, u]),
                ("qsim_traj_plan", i, [vp, up, vp, u, u]),
                ("qsim_traj_plan_tiled", i, [vp, up, vp, u, u, u]),
                ("qsim_traj_run", i, [vp, u, i]),
                ("qsim_traj_zz", i, [vp, u, vp, u, up]),
                ("qsim_traj_kraus", i, [vp, u, vp, u, up]),
                ("qsim_traj_state", i, [vp, u, vp]),
                ("qsim_traj_free", None, [vp]),
                ("qsim_driver_run_passes", i, [vp, vp, i, u, u, u]),
                ("qsim_driver_run", i, [vp, vp, i, u]),
                ("qsim_driver_capture", i, [vp, vp, i,
                                            ctypes.POINTER(ctypes.c_void_p)])):
            f = getattr(d, name)
            f.restype = res
            f.argtypes = args
        _dll = d
    return _dll


class Circuit:
    """A qsim text circuit parsed, fused and run by qsim's host code."""

    def __init__(self, text: str, maxtime: int = 0xFFFFFFFF):
        self.h = driver_lib().qsim_driver_load(text.encode(), maxtime)
        if not self.h:
            raise ValueError("qsim's parser rejected the circuit")
        self.num_qubits = driver_lib().qsim_driver_num_qubits(self.h)
        self.num_gates = driver_lib().qsim_driver_num_gates(self.h)

    def fuse(self, max_fused_size: int, fuser: int = FUSER_MQUBIT,
             f64_matrices: bool = False):
        _check(driver_lib().qsim_driver_fuse(self.h, max_fused_size, fuser,
                                             F64 if f64_matrices else F32))
        return self.num_passes

    def fuse_blocked(self, tile_bits: int = 13, low_bits: int = 5):
        """Cache-blocked passes (FP64 fused matrices; MFMA products in the
        arithmetic set_block_arith chose, scaled FP16 by default)."""
        _check(driver_lib().qsim_driver_fuse_blocked(self.h, tile_bits,
                                                     low_bits))
        return self.num_passes

    def set_block_arith(self, arith: str):
        """Blocked-pass arithmetic: "f16" (scaled two-part FP16, the
        default) or "bf16" (split-BF16)."""
        if arith not in ("f16", "bf16"):
            raise ValueError(f"blocked-pass arithmetic {arith!r}")
        _check(driver_lib().qsim_driver_set_block_arith(
            self.h, 0 if arith == "f16" else 1))

    def set_method(self, method: int):
        """Gate method the partition runs with (-1: the library's)."""
        _check(driver_lib().qsim_driver_set_method(self.h, method))

    def pass_features(self, method: int) -> np.ndarray:
        """Cost-model features of every pass, (num_passes, 19) float32."""
        out = np.zeros((self.num_passes, NUM_FEATURES), np.float32)
        for i in range(self.num_passes):
            _check(driver_lib().qsim_driver_pass_features(
                self.h, i, method, out[i].ctypes.data))
        return out

    def plan_learned(self, model=None):
        """Runs the learned planner (model: a qsim_amd.costmodel.CostModel
        or None for the fallback); the circuit keeps the chosen partition.
        Returns (chosen index, [candidate dicts])."""
        m = 32
        blocked = np.zeros(m, np.int32)
        a = np.zeros(m, np.uint32)
        b = np.zeros(m, np.uint32)
        meth = np.zeros(m, np.int32)
        pred = np.zeros(m, np.float64)
        npass = np.zeros(m, np.uint32)
        count = ctypes.c_uint32(0)
        chosen = ctypes.c_int(0)
        _check(driver_lib().qsim_driver_plan_learned(
            self.h, model.h if model is not None else None, m,
            blocked.ctypes.data, a.ctypes.data, b.ctypes.data,
            meth.ctypes.data, pred.ctypes.data, npass.ctypes.data,
            ctypes.byref(count), ctypes.byref(chosen)))
        cands = [{"blocked": bool(blocked[j]), "a": int(a[j]), "b": int(b[j]),
                  "method": int(meth[j]), "predicted_ms": float(pred[j]),
                  "passes": int(npass[j])} for j in range(count.value)]
        return chosen.value, cands

    def apply_candidate(self, cand):
        """Partitions the circuit as planner candidate `cand` describes."""
        if cand["blocked"]:
            self.fuse_blocked(cand["a"], cand["b"])
        else:
            self.fuse(cand["a"], FUSER_MQUBIT, f64_matrices=True)
        self.set_method(cand["method"])
        return self.num_passes

    def pass_kind(self, i: int) -> int:
        r = driver_lib().qsim_driver_pass_kind(self.h, i)
        if r < 0:
            _check(r)
        return r

    def block_info(self, i: int):
        """[(qubits, FP64 interleaved matrix), ...]: the fused gates of
        blocked pass i."""
        n = ctypes.c_uint32(0)
        ks = (ctypes.c_uint32 * 64)()
        qs = (ctypes.c_uint32 * 256)()
        ms = np.zeros(32 * 512, np.float64)
        _check(driver_lib().qsim_driver_block_info(
            self.h, i, ctypes.byref(n), ks, qs,
            ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        out, off = [], 0
        for g in range(n.value):
            size = 2 << (2 * ks[g])
            out.append((list(qs[4 * g:4 * g + ks[g]]), ms[off:off + size].copy()))
            off += size
        return out

    @property
    def num_passes(self) -> int:
        return driver_lib().qsim_driver_num_passes(self.h)

    def pass_info(self, i: int):
        k = ctypes.c_uint32(0)
        ng = ctypes.c_uint32(0)
        qs = (ctypes.c_uint32 * 64)()
        _check(driver_lib().qsim_driver_pass_info(
            self.h, i, ctypes.byref(k), qs, ctypes.byref(ng)))
        return list(qs[:k.value]), ng.value

    def passes(self):
        return [self.pass_info(i) for i in range(self.num_passes)]

    def run(self, state: State, seed: int = 1, first: int = 0,
            count: int = None):
        if count is None:
            count = self.num_passes - first
        _check(driver_lib().qsim_driver_run_passes(
            self.h, state.ptr, state.prec, first, count, seed))

    def capture(self, state: State) -> "Graph":
        """Runs the circuit once on `state` and captures it as one HIP graph
        with every matrix resident on the device (qsim_driver_capture)."""
        p = ctypes.c_void_p()
        _check(driver_lib().qsim_driver_capture(self.h, state.ptr, state.prec,
                                                ctypes.byref(p)))
        return Graph(p.value)

    def __del__(self):
        try:
            if self.h:
                driver_lib().qsim_driver_free(self.h)
        except Exception:
            pass


class Graph:
    """A captured circuit (qsim_hip_graph_*): launch() applies the circuit
    again to the state it was captured on, in one graph launch.  The graph
    owns its matrices; it does not depend on the Circuit object."""

    def __init__(self, handle: int):
        self.h = handle

    def launch(self):
        _check(lib().qsim_hip_graph_launch(self.h))

    def free(self):
        if self.h:
            _check(lib().qsim_hip_graph_free(self.h))
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
