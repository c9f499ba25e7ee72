"""qsim's Python extensions on the host: the reference's own CPU extension
(pybind_interface/avx2, built from the reference sources into oracle/_ref)
pinned against dense FP64 products, and this backend's `qsim_hip` extension
checked to export exactly the GPU bindings (no device calls here)."""

import numpy as np
import pytest

import pybind_circuits as pc


@pytest.fixture(scope="module")
def avx2():
    m = pc.load("qsim_avx2", pc.REF_DIR)
    if m is None:
        pytest.skip("reference extension not built (needs qsim sources and pybind11)")
    return m


@pytest.mark.parametrize("controlled", [False, True])
def test_reference_extension_is_the_dense_product(avx2, controlled):
    n = 8
    circ, ops = pc.build(avx2, n, 6, seed=11, controlled=controlled)
    got = np.asarray(avx2.qsim_simulate_fullstate(pc.options(circ), 0))
    got = got.view(np.complex64)
    want = pc.dense_run(n, ops)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < 3e-6, err


def test_dense_check_catches_a_wrong_convention(avx2):
    # The dense restatement must fail when a matrix gate's qubit order is
    # flipped: the pin above is not vacuous.
    n = 5
    circ, ops = pc.build(avx2, n, 3, seed=12)
    got = np.asarray(avx2.qsim_simulate_fullstate(pc.options(circ), 0))
    got = got.view(np.complex64)
    flipped = [(qs[::-1], u, c) if len(qs) == 2 else (qs, u, c)
               for qs, u, c in ops]
    wrong = pc.dense_run(n, flipped)
    assert np.linalg.norm(got - wrong) / np.linalg.norm(wrong) > 1e-2


def test_hip_extension_exports_the_gpu_bindings():
    m = pc.load("qsim_hip", pc.HIP_DIR)
    if m is None:
        pytest.skip("qsim_hip extension not built (needs qsim sources and pybind11)")
    names = {x for x in dir(m) if not x.startswith("_")}
    assert names == pc.GPU_BINDINGS
