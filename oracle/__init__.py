"""TEST INFRASTRUCTURE ONLY: ctypes bindings to the CPU restatement of qsim.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker.  The product (qsim_amd) never imports
it.  See qsim_oracle.cc for what each function restates and how it is pinned.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "qsim_ref")
REF_BIN_AVX512 = os.path.join(HERE, "_ref", "qsim_ref_avx512")


def host_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    fl = line.split()
                    return all(x in fl for x in ("avx512f", "avx512dq",
                                                 "avx512bw", "avx512vl"))
    except OSError:
        pass
    return False

_lib = None

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        for suf, fp in (("f32", _f32p), ("f64", _f64p)):
            f = getattr(_lib, f"orc_apply_gate_{suf}")
            f.argtypes = [ctypes.c_uint, fp, _u32p, ctypes.c_uint, fp]
            f = getattr(_lib, f"orc_apply_controlled_gate_{suf}")
            f.argtypes = [ctypes.c_uint, fp, _u32p, ctypes.c_uint, _u32p,
                          ctypes.c_uint, ctypes.c_uint64, fp]
            f = getattr(_lib, f"orc_expectation_value_{suf}")
            f.argtypes = [ctypes.c_uint, fp, _u32p, ctypes.c_uint, fp, _f64p]
            for name in ("norm", "serial_norm"):
                f = getattr(_lib, f"orc_{name}_{suf}")
                f.argtypes = [ctypes.c_uint, fp]
                f.restype = ctypes.c_double
            f = getattr(_lib, f"orc_inner_product_{suf}")
            f.argtypes = [ctypes.c_uint, fp, fp, _f64p]
            f = getattr(_lib, f"orc_real_inner_product_{suf}")
            f.argtypes = [ctypes.c_uint, fp, fp]
            f.restype = ctypes.c_double
            f = getattr(_lib, f"orc_sample_{suf}")
            f.argtypes = [ctypes.c_uint, fp, ctypes.c_uint64, ctypes.c_uint, _u64p]
            f = getattr(_lib, f"orc_sample_with_draws_{suf}")
            f.argtypes = [ctypes.c_uint, fp, _f64p, ctypes.c_uint64, _u64p]
            f = getattr(_lib, f"orc_collapse_{suf}")
            f.argtypes = [ctypes.c_uint, fp, ctypes.c_uint64, ctypes.c_uint64]
            f = getattr(_lib, f"orc_circuit_parse_{suf}")
            f.argtypes = [ctypes.c_char_p, ctypes.c_uint]
            f.restype = ctypes.c_void_p
            for name in ("num_qubits", "num_ops"):
                f = getattr(_lib, f"orc_circuit_{name}_{suf}")
                f.argtypes = [ctypes.c_void_p]
                f.restype = ctypes.c_uint
            f = getattr(_lib, f"orc_circuit_free_{suf}")
            f.argtypes = [ctypes.c_void_p]
            f = getattr(_lib, f"orc_circuit_run_{suf}")
            f.argtypes = [ctypes.c_void_p, fp, ctypes.c_uint, _u64p,
                          ctypes.c_uint, _u32p]
            f.restype = ctypes.c_int
            f = getattr(_lib, f"orc_gate_matrix_{suf}")
            f.argtypes = [ctypes.c_char_p, fp, ctypes.c_uint, _u32p,
                          ctypes.c_uint, _u32p, fp]
            f.restype = ctypes.c_int
        _lib.orc_generate_random_values.argtypes = [
            ctypes.c_uint64, ctypes.c_uint, ctypes.c_double, _f64p]
        _lib.orc_num_threads.restype = ctypes.c_int
        _lib.orc_set_num_threads.argtypes = [ctypes.c_int]
        _lib.orc_mlp_num_params.argtypes = [ctypes.c_uint, ctypes.c_uint]
        _lib.orc_mlp_num_params.restype = ctypes.c_ulong
        _lib.orc_mlp_forward.argtypes = [ctypes.c_uint, ctypes.c_uint, _f32p,
                                         _f32p, ctypes.c_uint, _f64p]
        _lib.orc_mlp_gradient.argtypes = [ctypes.c_uint, ctypes.c_uint, _f32p,
                                          _f32p, _f32p, ctypes.c_uint, _f64p]
        _lib.orc_mlp_gradient.restype = ctypes.c_double
    return _lib


def mlp_num_params(num_features, hidden):
    """Parameter count of the cost-model network (mlp_oracle.cc)."""
    return int(lib().orc_mlp_num_params(num_features, hidden))


def mlp_forward(num_features, hidden, params, x):
    """FP64 forward pass of the cost-model network on rows of x."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, num_features)
    y = np.zeros(len(xx), np.float64)
    lib().orc_mlp_forward(num_features, hidden, _ptr(p), _ptr(xx), len(xx),
                          _ptr(y))
    return y


def mlp_gradient(num_features, hidden, params, x, t):
    """(mean squared error, its FP64 gradient w.r.t. the parameters)."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, num_features)
    tt = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(len(p), np.float64)
    loss = lib().orc_mlp_gradient(num_features, hidden, _ptr(p), _ptr(xx),
                                  _ptr(tt), len(xx), _ptr(g))
    return float(loss), g


def _suf(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def _ptr(a, fp=None):
    if a.dtype == np.float32:
        return a.ctypes.data_as(_f32p)
    if a.dtype == np.float64:
        return a.ctypes.data_as(_f64p)
    if a.dtype == np.uint32:
        return a.ctypes.data_as(_u32p)
    if a.dtype == np.uint64:
        return a.ctypes.data_as(_u64p)
    raise TypeError(a.dtype)


def state_zero(n, dtype=np.float32):
    s = np.zeros(2 << n, dtype=dtype)
    s[0] = 1
    return s


def state_uniform(n, dtype=np.float32):
    s = np.zeros(2 << n, dtype=dtype)
    s[0::2] = dtype(1.0) / np.sqrt(dtype(1 << n)) if dtype == np.float32 \
        else 1.0 / np.sqrt(float(1 << n))
    return s


def apply_gate(n, state, qubits, matrix):
    qs = np.ascontiguousarray(qubits, dtype=np.uint32)
    m = np.ascontiguousarray(matrix, dtype=state.dtype)
    getattr(lib(), f"orc_apply_gate_{_suf(state.dtype)}")(
        n, _ptr(state), _ptr(qs), len(qs), _ptr(m))


def apply_controlled_gate(n, state, qubits, cqubits, cvals, matrix):
    qs = np.ascontiguousarray(qubits, dtype=np.uint32)
    cqs = np.ascontiguousarray(cqubits, dtype=np.uint32)
    m = np.ascontiguousarray(matrix, dtype=state.dtype)
    getattr(lib(), f"orc_apply_controlled_gate_{_suf(state.dtype)}")(
        n, _ptr(state), _ptr(qs), len(qs), _ptr(cqs), len(cqs), cvals, _ptr(m))


def expectation_value(n, state, qubits, matrix):
    qs = np.ascontiguousarray(qubits, dtype=np.uint32)
    m = np.ascontiguousarray(matrix, dtype=state.dtype)
    out = np.zeros(2, dtype=np.float64)
    getattr(lib(), f"orc_expectation_value_{_suf(state.dtype)}")(
        n, _ptr(state), _ptr(qs), len(qs), _ptr(m), _ptr(out))
    return complex(out[0], out[1])


def norm(n, state):
    return getattr(lib(), f"orc_norm_{_suf(state.dtype)}")(n, _ptr(state))


def serial_norm(n, state):
    return getattr(lib(), f"orc_serial_norm_{_suf(state.dtype)}")(n, _ptr(state))


def inner_product(n, a, b):
    out = np.zeros(2, dtype=np.float64)
    getattr(lib(), f"orc_inner_product_{_suf(a.dtype)}")(n, _ptr(a), _ptr(b),
                                                          _ptr(out))
    return complex(out[0], out[1])


def real_inner_product(n, a, b):
    return getattr(lib(), f"orc_real_inner_product_{_suf(a.dtype)}")(
        n, _ptr(a), _ptr(b))


def sample(n, state, num, seed):
    out = np.zeros(num, dtype=np.uint64)
    getattr(lib(), f"orc_sample_{_suf(state.dtype)}")(n, _ptr(state), num, seed,
                                                      _ptr(out))
    return out


def sample_with_draws(n, state, draws):
    draws = np.ascontiguousarray(draws, dtype=np.float64)
    num = len(draws) - 1
    out = np.zeros(max(num, 0), dtype=np.uint64)
    getattr(lib(), f"orc_sample_with_draws_{_suf(state.dtype)}")(
        n, _ptr(state), _ptr(draws), num, _ptr(out))
    return out


def collapse(n, state, mask, bits):
    getattr(lib(), f"orc_collapse_{_suf(state.dtype)}")(n, _ptr(state), mask,
                                                        bits)


def generate_random_values(num, seed, max_value):
    out = np.zeros(num + 1, dtype=np.float64)
    lib().orc_generate_random_values(num, seed, max_value, _ptr(out))
    return out


def gate_matrix(name, qubits, params=(), dtype=np.float64):
    """(sorted qubits, row-major interleaved matrix) of a qsim gate."""
    suf = _suf(dtype)
    p = np.ascontiguousarray(params, dtype=dtype)
    qs = np.ascontiguousarray(qubits, dtype=np.uint32)
    qs_out = np.zeros(max(len(qs), 1), dtype=np.uint32)
    m = np.zeros(2 << (2 * len(qs)), dtype=dtype)
    rc = getattr(lib(), f"orc_gate_matrix_{suf}")(
        name.encode(), _ptr(p) if len(p) else ctypes.cast(None, _f64p if suf == "f64" else _f32p),
        len(p), _ptr(qs), len(qs), _ptr(qs_out), _ptr(m))
    if rc != 0:
        raise ValueError(f"unknown gate {name} / arity")
    return qs_out[:len(qs)].copy(), m


class Circuit:
    """A parsed qsim text circuit (restated parser)."""

    def __init__(self, text, dtype=np.float32, maxtime=0xFFFFFFFF):
        self.suf = _suf(dtype)
        self.dtype = np.dtype(dtype)
        self.h = getattr(lib(), f"orc_circuit_parse_{self.suf}")(
            text.encode(), maxtime)
        if not self.h:
            raise ValueError("cannot parse circuit")
        self.num_qubits = getattr(lib(), f"orc_circuit_num_qubits_{self.suf}")(
            self.h)
        self.num_ops = getattr(lib(), f"orc_circuit_num_ops_{self.suf}")(self.h)

    def run(self, state=None, seed=1):
        if state is None:
            state = state_zero(self.num_qubits, self.dtype)
        mres = np.zeros(1024, dtype=np.uint64)
        cnt = ctypes.c_uint32(0)
        rc = getattr(lib(), f"orc_circuit_run_{self.suf}")(
            self.h, _ptr(state), seed, _ptr(mres), 1024, ctypes.byref(cnt))
        if rc != 0:
            raise RuntimeError("measurement failed")
        return state, mres[:cnt.value].copy()

    def __del__(self):
        try:
            if self.h:
                getattr(lib(), f"orc_circuit_free_{self.suf}")(self.h)
        except Exception:
            pass


def run_reference(circuit_file, mode="avx", max_fused=2, threads=1, seed=1,
                  state_out=None, num_samples=0, sample_seed=0, bits_out=None,
                  passes=None, timeout=3600, widest_simd=False):
    """Runs the compiled reference driver (oracle/_ref/qsim_ref; with
    widest_simd, its AVX-512 build where the host has AVX-512, so that
    lib/simmux.h selects SimulatorAVX512 as a native build would)."""
    import json
    exe = REF_BIN
    if widest_simd and os.path.exists(REF_BIN_AVX512) and host_has_avx512():
        exe = REF_BIN_AVX512
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    cmd = [exe, "-c", circuit_file, "-m", mode, "-f", str(max_fused),
           "-t", str(threads), "-s", str(seed)]
    if state_out:
        cmd += ["-o", state_out]
    if num_samples:
        cmd += ["-n", str(num_samples), "-x", str(sample_seed)]
        if bits_out:
            cmd += ["-b", bits_out]
    if passes is not None:
        cmd += ["-P", str(passes)]
    out = subprocess.run(cmd, check=True, capture_output=True, text=True,
                         timeout=timeout)
    return json.loads(out.stdout.strip().splitlines()[-1])


def reference_draws(num, seed, max_value, path):
    subprocess.run([REF_BIN, "-R", str(num), "-x", str(seed), "-M",
                    repr(float(max_value)), "-b", path], check=True)
    return np.fromfile(path, dtype=np.float64)


def run_reference_trajectories(circuit_file, rows, cols, rep0, rep1,
                               p_depol=1e-3, gamma=5e-3, max_fused=4,
                               threads=1, timeout=3600, widest_simd=False,
                               noisy_text=False):
    """qsim's QuantumTrajectorySimulator on the noisy grid circuit
    (oracle/_ref/qsim_ref -T; with noisy_text, the file is noisy-circuit
    text of gates and channels, -N): per repetition the Kraus index of
    every channel and <Z_a Z_b> of the grid-neighbour pairs; plus a
    summary."""
    import json
    exe = REF_BIN
    if widest_simd and os.path.exists(REF_BIN_AVX512) and host_has_avx512():
        exe = REF_BIN_AVX512
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    cmd = [exe, "-c", circuit_file, "-T", f"{rep0}:{rep1}", "-g",
           f"{rows}:{cols}", "-p", repr(float(p_depol)), "-G",
           repr(float(gamma)), "-f", str(max_fused), "-t", str(threads)]
    if noisy_text:
        cmd.append("-N")
    out = subprocess.run(cmd, check=True, capture_output=True, text=True,
                         timeout=timeout)
    lines = [json.loads(s) for s in out.stdout.strip().splitlines()]
    return lines[:-1], lines[-1]
