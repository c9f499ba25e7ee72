"""qsim_amd -- MI355X (gfx950) state-vector backend for qsim.

The compute path is the C ABI of ``include/qsim_hip.h`` implemented by
hand-written HIP kernels in ``qsim_amd/csrc`` and built into
``qsim_amd/lib/libqsim_hip.so``.  This module loads that library and mirrors
qsim's StateSpace / Simulator interface for Python callers (tests, bench):
the method names, argument meaning and error behaviour follow
``lib/statespace_cuda.h`` and ``lib/simulator_cuda.h`` of qsim.

There is no CPU fallback: if the library is missing or cannot be loaded,
``lib()`` raises.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libqsim_hip.so")

F32 = 0
F64 = 1
BLOCK_ACCURATE = 0x100  # QSIM_HIP_BLOCK_ACCURATE, OR-ed into a matrix precision
BLOCK_BF16 = 0x200  # QSIM_HIP_BLOCK_BF16: split-BF16 instead of scaled FP16
# Status codes of include/qsim_hip.h.
OK = 0
ERR_ARG = -1
ERR_RUNTIME = -2
ERR_OOM = -3
ERR_UNSUPPORTED = -4
METHOD_AUTO = 0
METHOD_VECTOR = 1
METHOD_MFMA_BF16X3 = 2
METHOD_MFMA_FP16X2 = 3

_lib = None

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_u = ctypes.c_uint
_u64 = ctypes.c_uint64
_d = ctypes.c_double

# (name, restype, argtypes) of every exported entry point of qsim_hip.h.
SIGNATURES = [
    ("qsim_hip_last_error", ctypes.c_char_p, []),
    ("qsim_hip_init", _i, [_i]),
    ("qsim_hip_device_info", _i, [ctypes.c_char_p, ctypes.c_size_t, _u64p,
                                  ctypes.POINTER(ctypes.c_int)]),
    ("qsim_hip_set_stream", _i, [_vp]),
    ("qsim_hip_synchronize", _i, []),
    ("qsim_hip_min_size", _u64, [_u]),
    ("qsim_hip_alloc", _i, [_i, _u, ctypes.POINTER(_vp)]),
    ("qsim_hip_free", _i, [_vp]),
    ("qsim_hip_copy_d2d", _i, [_vp, _vp, _u64]),
    ("qsim_hip_copy_d2h", _i, [_vp, _vp, _u64]),
    ("qsim_hip_copy_h2d", _i, [_vp, _vp, _u64]),
    ("qsim_hip_copy_convert", _i, [_vp, _i, _vp, _i, _u]),
    ("qsim_hip_set_all_zeros", _i, [_i, _vp, _u]),
    ("qsim_hip_set_state_zero", _i, [_i, _vp, _u]),
    ("qsim_hip_set_state_uniform", _i, [_i, _vp, _u]),
    ("qsim_hip_internal_to_normal_order", _i, [_i, _vp, _u]),
    ("qsim_hip_normal_to_internal_order", _i, [_i, _vp, _u]),
    ("qsim_hip_get_ampl", _i, [_i, _vp, _u, _u64, _f64p, _f64p]),
    ("qsim_hip_set_ampl", _i, [_i, _vp, _u, _u64, _d, _d]),
    ("qsim_hip_get_ampls", _i, [_i, _vp, _u, _vp, _u64, _vp]),
    ("qsim_hip_graph_record_begin", _i, [ctypes.POINTER(ctypes.c_void_p)]),
    ("qsim_hip_graph_record_end", _i, [_vp]),
    ("qsim_hip_graph_capture_begin", _i, [_vp]),
    ("qsim_hip_graph_capture_end", _i, [_vp]),
    ("qsim_hip_graph_launch", _i, [_vp]),
    ("qsim_hip_graph_free", _i, [_vp]),
    ("qsim_hip_bulk_set_ampl", _i, [_i, _vp, _u, _u64, _u64, _d, _d, _i]),
    ("qsim_hip_add", _i, [_i, _vp, _vp, _u]),
    ("qsim_hip_multiply", _i, [_i, _vp, _u, _d]),
    ("qsim_hip_inner_product", _i, [_i, _vp, _vp, _u, _f64p, _f64p]),
    ("qsim_hip_real_inner_product", _i, [_i, _vp, _vp, _u, _f64p]),
    ("qsim_hip_norm", _i, [_i, _vp, _u, _f64p]),
    ("qsim_hip_inner_product_noalloc", _i, [_i, _vp, _vp, _u, _f64p, _f64p]),
    ("qsim_hip_sample", _i, [_i, _vp, _u, _f64p, _u64, _u64p]),
    ("qsim_hip_set_gate_queue", _i, [_u, _u]),
    ("qsim_hip_get_gate_queue", _i, [_vp, _vp, _vp, _vp]),
    ("qsim_hip_flush", _i, []),
    ("qsim_hip_block_resident_ops", _u, [_u]),
    ("qsim_hip_costmodel_load", _i, [ctypes.c_char_p, _vp]),
    ("qsim_hip_costmodel_set_default", _i, [_vp]),
    ("qsim_hip_costmodel_default", _vp, []),
    ("qsim_hip_sample_seeded", _i, [_i, _vp, _u, _u64, _u, _u64p]),
    ("qsim_hip_num_partial_norms", _u, [_u]),
    ("qsim_hip_partial_norms", _i, [_i, _vp, _u, _f64p]),
    ("qsim_hip_find_measured_bits", _i, [_i, _vp, _u, _u, _d, _u64, _u64p]),
    ("qsim_hip_collapse", _i, [_i, _vp, _u, _u64, _u64]),
    ("qsim_hip_apply_gate", _i, [_i, _vp, _u, _u32p, _u, _vp, _i]),
    ("qsim_hip_apply_controlled_gate", _i, [_i, _vp, _u, _u32p, _u, _u32p, _u,
                                            _u64, _vp, _i]),
    ("qsim_hip_apply_gate_block", _i, [_i, _vp, _u, _u32p, _u, _u, _u32p,
                                       _u32p, _vp, _i]),
    ("qsim_hip_expectation_value", _i, [_i, _vp, _u, _u32p, _u, _vp, _i,
                                        _f64p, _f64p]),
    ("qsim_hip_set_gate_method", _i, [_i]),
    ("qsim_hip_get_gate_method", _i, []),
    ("qsim_hip_event_record", _i, [_i]),
    ("qsim_hip_event_elapsed_ms", _i, [_i, _i, ctypes.POINTER(ctypes.c_float)]),
    ("qsim_hip_pass_timer_start", _i, [_u]),
    ("qsim_hip_pass_timer_stop", _i, [ctypes.POINTER(ctypes.c_float), _u, _u32p]),
    ("qsim_hip_block_prepare", _i, [_u, _u32p, _u, _u, _u32p, _u32p, _vp, _i,
                                    ctypes.POINTER(_vp)]),
    ("qsim_hip_block_run", _i, [_vp, _i, _vp, _u]),
    ("qsim_hip_block_free", _i, [_vp]),
    ("qsim_hip_traj_prepare", _i, [_vp, _u, _vp, _u64, _vp, _u64, _u32p, _u,
                                   _u, ctypes.POINTER(_vp)]),
    ("qsim_hip_traj_run", _i, [_vp, _u, _i]),
    ("qsim_hip_traj_results", _i, [_vp, _f64p, _u, _vp, _u]),
    ("qsim_hip_traj_state", _i, [_vp, _u, _vp]),
    ("qsim_hip_traj_free", _i, [_vp]),
    ("qsim_hip_costmodel_create", _i, [_u, _u, ctypes.POINTER(_vp)]),
    ("qsim_hip_costmodel_free", _i, [_vp]),
    ("qsim_hip_costmodel_num_params", _u64, [_vp]),
    ("qsim_hip_costmodel_set_params", _i, [_vp, _vp, _u64]),
    ("qsim_hip_costmodel_get_params", _i, [_vp, _vp, _u64]),
    ("qsim_hip_costmodel_forward", _i, [_vp, _vp, _u, _vp]),
    ("qsim_hip_costmodel_gradient", _i, [_vp, _vp, _vp, _u, _f64p, _vp]),
    ("qsim_hip_costmodel_train", _i, [_vp, _vp, _vp, _u, _u, _u,
                                      ctypes.c_float, _u, _f64p]),
]


class QsimHipError(RuntimeError):
    """A failed call into the HIP backend (the C ABI's error code + text)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"qsim_hip error {code}: {msg}")
        self.code = code


def load_library(path: str = LIB_PATH):
    """Loads libqsim_hip.so and declares every entry point (no GPU needed)."""
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `python -m qsim_amd.build`")
    dll = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        f = getattr(dll, name)
        f.restype = res
        f.argtypes = args
    return dll


def lib():
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def _check(rc: int):
    if rc != 0:
        msg = lib().qsim_hip_last_error().decode(errors="replace")
        raise QsimHipError(rc, msg)


def device_info():
    name = ctypes.create_string_buffer(256)
    mem = ctypes.c_uint64(0)
    cus = ctypes.c_int(0)
    _check(lib().qsim_hip_device_info(name, 256, ctypes.byref(mem),
                                      ctypes.byref(cus)))
    return name.value.decode(), mem.value, cus.value


def synchronize():
    _check(lib().qsim_hip_synchronize())


def set_gate_method(method: int):
    _check(lib().qsim_hip_set_gate_method(method))


def get_gate_method() -> int:
    return lib().qsim_hip_get_gate_method()


def set_gate_queue(tile_bits: int, low_bits: int = 4):
    """Collects apply_gate calls into blocked passes (0: off);
    qsim_hip_set_gate_queue."""
    _check(lib().qsim_hip_set_gate_queue(tile_bits, low_bits))


def gate_queue_stats():
    """(tile_bits, low_bits, blocked passes run, gates in them)."""
    tb, lc = ctypes.c_uint(0), ctypes.c_uint(0)
    passes, gates = ctypes.c_uint64(0), ctypes.c_uint64(0)
    _check(lib().qsim_hip_get_gate_queue(ctypes.byref(tb), ctypes.byref(lc),
                                         ctypes.byref(passes),
                                         ctypes.byref(gates)))
    return tb.value, lc.value, passes.value, gates.value


def flush():
    _check(lib().qsim_hip_flush())


class PassTimer:
    """Kernel durations of every gate pass between start() and stop()."""

    @staticmethod
    def start(capacity: int = 65536):
        _check(lib().qsim_hip_pass_timer_start(capacity))

    @staticmethod
    def stop(max_count: int = 65536) -> np.ndarray:
        out = np.zeros(max_count, dtype=np.float32)
        cnt = ctypes.c_uint32(0)
        _check(lib().qsim_hip_pass_timer_stop(
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), max_count,
            ctypes.byref(cnt)))
        return out[:min(cnt.value, max_count)].copy()


class Timer:
    """HIP events on the library stream (slots 0..63)."""

    @staticmethod
    def record(slot: int):
        _check(lib().qsim_hip_event_record(slot))

    @staticmethod
    def elapsed_ms(a: int, b: int) -> float:
        ms = ctypes.c_float(0)
        _check(lib().qsim_hip_event_elapsed_ms(a, b, ctypes.byref(ms)))
        return ms.value


def _prec_of(dtype) -> int:
    dt = np.dtype(dtype)
    if dt in (np.float32, np.complex64):
        return F32
    if dt in (np.float64, np.complex128):
        return F64
    raise TypeError(f"unsupported state type {dt}")


def _u32(a: Sequence[int]):
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.uint32).reshape(-1))
    return arr, arr.ctypes.data_as(_u32p)


class State:
    """A device-resident state vector (VectorSpaceCUDA::Vector counterpart,
    lib/vectorspace_cuda.h:42-82).  Owns its device memory."""

    def __init__(self, num_qubits: int, dtype=np.float32, ptr: int = None,
                 owner: bool = True):
        self.num_qubits = int(num_qubits)
        self.prec = _prec_of(dtype)
        self.fp = np.float32 if self.prec == F32 else np.float64
        self.owner = owner
        if ptr is None:
            p = ctypes.c_void_p(0)
            _check(lib().qsim_hip_alloc(self.prec, self.num_qubits,
                                        ctypes.byref(p)))
            self.ptr = p.value
        else:
            self.ptr = ptr

    @property
    def min_size(self) -> int:
        return int(lib().qsim_hip_min_size(self.num_qubits))

    @property
    def nbytes(self) -> int:
        return 2 * (1 << self.num_qubits) * np.dtype(self.fp).itemsize

    def free(self):
        if self.ptr and self.owner:
            _check(lib().qsim_hip_free(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def to_numpy(self) -> np.ndarray:
        """Interleaved (re, im) copy of the 2^n amplitudes."""
        out = np.empty(2 << self.num_qubits, dtype=self.fp)
        _check(lib().qsim_hip_copy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def from_numpy(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.fp).reshape(-1)
        if a.size != 2 << self.num_qubits:
            raise ValueError("size mismatch")
        _check(lib().qsim_hip_copy_h2d(self.ptr, a.ctypes.data, a.nbytes))


class StateSpace:
    """Python mirror of qsim's StateSpace API (lib/statespace_cuda.h,
    lib/statespace.h) over the C ABI."""

    def __init__(self, dtype=np.float32):
        self.prec = _prec_of(dtype)
        self.dtype = np.float32 if self.prec == F32 else np.float64

    # VectorSpace part
    def create(self, num_qubits: int) -> Optional[State]:
        try:
            return State(num_qubits, self.dtype)
        except QsimHipError as e:
            if e.code == -3:   # OOM -> Null(), lib/vectorspace_cuda.h:87-96
                return None
            raise

    @staticmethod
    def min_size(num_qubits: int) -> int:
        return int(lib().qsim_hip_min_size(num_qubits))

    def copy(self, src: State, dest: State) -> bool:
        if src.num_qubits != dest.num_qubits:
            return False
        if src.prec != dest.prec:
            _check(lib().qsim_hip_copy_convert(dest.ptr, dest.prec, src.ptr,
                                               src.prec, src.num_qubits))
            return True
        _check(lib().qsim_hip_copy_d2d(dest.ptr, src.ptr, src.nbytes))
        return True

    # StateSpace part
    def set_all_zeros(self, s: State):
        _check(lib().qsim_hip_set_all_zeros(s.prec, s.ptr, s.num_qubits))

    def set_state_zero(self, s: State):
        _check(lib().qsim_hip_set_state_zero(s.prec, s.ptr, s.num_qubits))

    def set_state_uniform(self, s: State):
        _check(lib().qsim_hip_set_state_uniform(s.prec, s.ptr, s.num_qubits))

    def get_ampl(self, s: State, i: int) -> complex:
        re, im = ctypes.c_double(0), ctypes.c_double(0)
        _check(lib().qsim_hip_get_ampl(s.prec, s.ptr, s.num_qubits, i,
                                       ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def get_ampls(self, s: State, indices) -> np.ndarray:
        """Amplitudes of many basis states in one gather (complex64 or
        complex128, as the state)."""
        idx = np.ascontiguousarray(indices, dtype=np.uint64).reshape(-1)
        out = np.empty(idx.size, dtype=np.complex64 if s.prec == F32
                       else np.complex128)
        _check(lib().qsim_hip_get_ampls(s.prec, s.ptr, s.num_qubits,
                                        idx.ctypes.data, idx.size,
                                        out.ctypes.data))
        return out

    def set_ampl(self, s: State, i: int, val: complex):
        _check(lib().qsim_hip_set_ampl(s.prec, s.ptr, s.num_qubits, i,
                                       val.real, val.imag))

    def bulk_set_ampl(self, s: State, mask: int, bits: int, val: complex,
                      exclude: bool = False):
        _check(lib().qsim_hip_bulk_set_ampl(s.prec, s.ptr, s.num_qubits, mask,
                                            bits, val.real, val.imag,
                                            int(exclude)))

    def add(self, src: State, dest: State) -> bool:
        if src.num_qubits != dest.num_qubits:
            return False
        _check(lib().qsim_hip_add(src.prec, src.ptr, dest.ptr, src.num_qubits))
        return True

    def multiply(self, a: float, s: State):
        _check(lib().qsim_hip_multiply(s.prec, s.ptr, s.num_qubits, a))

    def inner_product(self, a: State, b: State) -> complex:
        if a.num_qubits != b.num_qubits:
            return complex(float("nan"), 0.0)
        re, im = ctypes.c_double(0), ctypes.c_double(0)
        _check(lib().qsim_hip_inner_product(a.prec, a.ptr, b.ptr, a.num_qubits,
                                            ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def real_inner_product(self, a: State, b: State) -> float:
        if a.num_qubits != b.num_qubits:
            return float("nan")
        r = ctypes.c_double(0)
        _check(lib().qsim_hip_real_inner_product(a.prec, a.ptr, b.ptr,
                                                 a.num_qubits, ctypes.byref(r)))
        return r.value

    def norm(self, s: State) -> float:
        r = ctypes.c_double(0)
        _check(lib().qsim_hip_norm(s.prec, s.ptr, s.num_qubits, ctypes.byref(r)))
        return r.value

    def sample(self, s: State, num_samples: int, seed: int) -> np.ndarray:
        out = np.zeros(num_samples, dtype=np.uint64)
        _check(lib().qsim_hip_sample_seeded(s.prec, s.ptr, s.num_qubits,
                                            num_samples, seed,
                                            out.ctypes.data_as(_u64p)))
        return out

    def sample_with_draws(self, s: State, draws: np.ndarray) -> np.ndarray:
        draws = np.ascontiguousarray(draws, dtype=np.float64)
        num = len(draws)
        out = np.zeros(num, dtype=np.uint64)
        _check(lib().qsim_hip_sample(s.prec, s.ptr, s.num_qubits,
                                     draws.ctypes.data_as(_f64p), num,
                                     out.ctypes.data_as(_u64p)))
        return out

    def partial_norms(self, s: State) -> np.ndarray:
        cnt = lib().qsim_hip_num_partial_norms(s.num_qubits)
        out = np.zeros(cnt, dtype=np.float64)
        _check(lib().qsim_hip_partial_norms(s.prec, s.ptr, s.num_qubits,
                                            out.ctypes.data_as(_f64p)))
        return out

    def find_measured_bits(self, s: State, m: int, r: float, mask: int) -> int:
        out = ctypes.c_uint64(0)
        _check(lib().qsim_hip_find_measured_bits(s.prec, s.ptr, s.num_qubits,
                                                 m, r, mask, ctypes.byref(out)))
        return out.value

    def collapse(self, s: State, mask: int, bits: int):
        _check(lib().qsim_hip_collapse(s.prec, s.ptr, s.num_qubits, mask, bits))

    def virtual_measure(self, qubits, rgen_draw, s: State):
        """StateSpace::VirtualMeasure (lib/statespace.h:97-140); `rgen_draw`
        maps the total norm to one uniform draw in [0, norm)."""
        mask = 0
        for q in qubits:
            if q >= s.num_qubits:
                return None
            mask |= 1 << q
        pn = np.cumsum(self.partial_norms(s))
        r = rgen_draw(pn[-1])
        m = 0
        while r > pn[m]:
            m += 1
        if m > 0:
            r -= pn[m - 1]
        bits = self.find_measured_bits(s, m, r, mask)
        return mask, bits

    def measure(self, qubits, rgen_draw, s: State):
        res = self.virtual_measure(qubits, rgen_draw, s)
        if res is not None:
            self.collapse(s, res[0], res[1])
        return res


def _block_args(tile_qubits, gates):
    tarr, tp = _u32(sorted(tile_qubits))
    ks, qs, ms = [], [], []
    for q, m in gates:
        q = list(q)
        ks.append(len(q))
        qs.extend(q + [0] * (4 - len(q)))
        ms.append(np.asarray(m, dtype=np.float64).reshape(-1))
    karr, kp = _u32(ks)
    qarr, qp = _u32(qs)
    mat = (np.ascontiguousarray(np.concatenate(ms)) if ms
           else np.zeros(2, np.float64))
    return (tarr, tp, karr, kp, qarr, qp, mat)


def _block_flags(accurate, arith):
    if arith not in ("f16", "bf16"):
        raise ValueError(f"blocked-pass arithmetic {arith!r}: 'f16' or 'bf16'")
    return (F64 | (BLOCK_ACCURATE if accurate else 0)
            | (BLOCK_BF16 if arith == "bf16" else 0))


class PreparedBlock:
    """A cache-blocked pass whose fragments live on the device
    (qsim_hip_block_prepare / _run / _free)."""

    def __init__(self, num_qubits, tile_qubits, gates, accurate=False,
                 arith="f16"):
        a = _block_args(tile_qubits, gates)
        tarr, tp, karr, kp, qarr, qp, mat = a
        h = ctypes.c_void_p()
        mprec = _block_flags(accurate, arith)
        _check(lib().qsim_hip_block_prepare(num_qubits, tp, len(tarr),
                                            len(karr), kp, qp,
                                            mat.ctypes.data, mprec,
                                            ctypes.byref(h)))
        self.h = h
        self.num_qubits = num_qubits

    def run(self, s: "State"):
        _check(lib().qsim_hip_block_run(self.h, s.prec, s.ptr, s.num_qubits))

    def free(self):
        if getattr(self, "h", None):
            _check(lib().qsim_hip_block_free(self.h))
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Simulator:
    """Python mirror of SimulatorCUDA (lib/simulator_cuda.h:34-260)."""

    def apply_gate(self, qs, matrix, s: State):
        qarr, qp = _u32(qs)
        m = np.ascontiguousarray(matrix).reshape(-1)
        mprec = _prec_of(m.dtype)
        _check(lib().qsim_hip_apply_gate(s.prec, s.ptr, s.num_qubits, qp,
                                         len(qarr), m.ctypes.data, mprec))

    def apply_controlled_gate(self, qs, cqs, cvals, matrix, s: State):
        qarr, qp = _u32(qs)
        carr, cp = _u32(cqs)
        m = np.ascontiguousarray(matrix).reshape(-1)
        mprec = _prec_of(m.dtype)
        _check(lib().qsim_hip_apply_controlled_gate(
            s.prec, s.ptr, s.num_qubits, qp, len(qarr), cp, len(carr), cvals,
            m.ctypes.data, mprec))

    def apply_gate_block(self, tile_qubits, gates, s: State, accurate=False,
                         arith="f16"):
        """Applies `gates` [(qs, matrix), ...] (k <= 4, qubits inside
        `tile_qubits`) in one cache-blocked pass; FP64 or FP32 matrices.
        arith: "f16" (scaled two-part FP16 products, the default) or "bf16"
        (QSIM_HIP_BLOCK_BF16: split-BF16); accurate: with "bf16",
        amplitudes split in three bf16 parts (12 MFMAs per 32 groups instead
        of 8; QSIM_HIP_BLOCK_ACCURATE)."""
        tarr, tp = _u32(sorted(tile_qubits))
        ks, qs, ms = [], [], []
        for q, m in gates:
            q = list(q)
            ks.append(len(q))
            qs.extend(q + [0] * (4 - len(q)))
            ms.append(np.asarray(m, dtype=np.float64).reshape(-1))
        karr, kp = _u32(ks)
        qarr, qp = _u32(qs)
        mat = (np.ascontiguousarray(np.concatenate(ms)) if ms
               else np.zeros(2, np.float64))
        _check(lib().qsim_hip_apply_gate_block(
            s.prec, s.ptr, s.num_qubits, tp, len(tarr), len(ks), kp, qp,
            mat.ctypes.data, _block_flags(accurate, arith)))

    def prepare_gate_block(self, num_qubits, tile_qubits, gates,
                           accurate=False, arith="f16"):
        """A blocked pass prepared once (qsim_hip_block_prepare)."""
        return PreparedBlock(num_qubits, tile_qubits, gates, accurate, arith)

    def expectation_value(self, qs, matrix, s: State) -> complex:
        qarr, qp = _u32(qs)
        m = np.ascontiguousarray(matrix).reshape(-1)
        mprec = _prec_of(m.dtype)
        re, im = ctypes.c_double(0), ctypes.c_double(0)
        _check(lib().qsim_hip_expectation_value(
            s.prec, s.ptr, s.num_qubits, qp, len(qarr), m.ctypes.data, mprec,
            ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    @staticmethod
    def simd_register_size() -> int:
        # Tests size their problems with it (tests/simulator_testfixture.h:1153).
        return 64
