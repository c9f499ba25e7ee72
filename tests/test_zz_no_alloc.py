"""The no-allocation inner product (qsim_hip_inner_product_noalloc) through
the C ABI: on the fixed device scratch it gives exactly what
qsim_hip_inner_product / qsim_hip_real_inner_product give (same kernels, same
reduction order), on consecutive calls of shrinking size, and it refuses to
run while the gate queue holds gates."""

import ctypes

import numpy as np
import pytest

import qsim_amd

pytestmark = pytest.mark.gpu


def _random_state(ss, n, dtype, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(2 << n).astype(dtype)
    st = ss.create(n)
    st.from_numpy(v)
    return st


def _noalloc(prec, a, b, n, real):
    re, im = ctypes.c_double(0), ctypes.c_double(0)
    rc = qsim_amd.lib().qsim_hip_inner_product_noalloc(
        prec, a, b, n, ctypes.byref(re), None if real else ctypes.byref(im))
    return rc, complex(re.value, 0.0 if real else im.value)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_matches_inner_product(dtype):
    ss = qsim_amd.StateSpace(dtype)
    # Two calls in a row, the second on a smaller state (fewer partial sums
    # than the first left in the fixed scratch).
    for n, seed in ((22, 1), (13, 2), (6, 3)):
        a = _random_state(ss, n, dtype, seed)
        b = _random_state(ss, n, dtype, seed + 100)
        want = ss.inner_product(a, b)
        rc, got = _noalloc(a.prec, a.ptr, b.ptr, n, real=False)
        assert rc == qsim_amd.OK, qsim_amd.lib().qsim_hip_last_error()
        assert got == want, (n, got, want)
        want_re = ss.real_inner_product(a, b)
        rc, got_re = _noalloc(a.prec, a.ptr, b.ptr, n, real=True)
        assert rc == qsim_amd.OK, qsim_amd.lib().qsim_hip_last_error()
        assert got_re.real == want_re
        # Against a host FP64 sum of the same values.  The device rounds
        # each product to the state's type before the FP64 sum (as the
        # reference does), so the bound is a few roundings of sum |a||b|.
        x = a.to_numpy().astype(np.float64)
        y = b.to_numpy().astype(np.float64)
        xc, yc = x[0::2] + 1j * x[1::2], y[0::2] + 1j * y[1::2]
        host = np.vdot(xc, yc)
        eps = np.finfo(dtype).eps
        bound = 4 * eps * float(np.sum(np.abs(xc) * np.abs(yc)))
        assert abs(got - host) <= bound, (n, abs(got - host), bound)
        a.free()
        b.free()


def test_noalloc_refuses_while_gates_are_queued():
    ss = qsim_amd.StateSpace(np.float32)
    n = 12
    a = _random_state(ss, n, np.float32, 5)
    h = np.array([1, 0, 1, 0, 1, 0, -1, 0], np.float32) / np.sqrt(2.0)
    qsim_amd.set_gate_queue(10, 4)
    try:
        sim = qsim_amd.Simulator()
        sim.apply_gate([3], h, a)
        rc, _ = _noalloc(a.prec, a.ptr, a.ptr, n, real=True)
        assert rc == qsim_amd.ERR_ARG
        assert b"gate queue" in qsim_amd.lib().qsim_hip_last_error()
        qsim_amd.flush()
        rc, got = _noalloc(a.prec, a.ptr, a.ptr, n, real=True)
        assert rc == qsim_amd.OK
        assert got.real == ss.norm(a)
    finally:
        qsim_amd.set_gate_queue(0, 0)
        a.free()
