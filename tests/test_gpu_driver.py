"""Whole circuits through qsim's host code (parser, MultiQubitGateFuser,
apply loop) on the gfx950 backend, against the CPU oracle."""

import os

import numpy as np
import pytest

import oracle
import qsim_amd
from helpers import from_interleaved, reference_runs, rel_l2
from qsim_amd import circuits, driver

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grid20():
    text = circuits.grid_circuit(20, 14, seed=0)
    s32, _ = oracle.Circuit(text, np.float32).run()
    s64, _ = oracle.Circuit(text, np.float64).run()
    return text, s32, s64


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6])
def test_grid20_fused_vs_oracle(grid20, k):
    text, s32, s64 = grid20
    c = driver.Circuit(text)
    c.fuse(k)
    ss = qsim_amd.StateSpace(np.float32)
    st = ss.create(20)
    ss.set_state_zero(st)
    c.run(st)
    got = st.to_numpy()
    assert rel_l2(got, s32) < 1e-5
    # qsim's FP32 gate matrices are unitary only to ~3e-8 (sqrt(W) holds
    # float(1/sqrt 2)), so the FP32-matrix state loses norm: its direction
    # is measured by the normalised fidelity.
    g, f = from_interleaved(got), from_interleaved(s64)
    fid_raw = abs(np.vdot(f, g)) ** 2
    fid_norm = fid_raw / (np.vdot(g, g).real * np.vdot(f, f).real)
    print(f"k={k}: {c.num_passes} passes, rel-L2 vs oracle "
          f"{rel_l2(got, s32):.2e}, 1-F vs FP64 {1 - fid_raw:.2e} "
          f"(normalised {1 - fid_norm:.2e})", flush=True)
    assert fid_norm > 1 - 1e-9
    # FP64 fused matrices (built on the host in double) on the complex64
    # state meet the unnormalised bar F >= 1 - 1e-6.
    c.fuse(k, f64_matrices=True)
    ss.set_state_zero(st)
    c.run(st)
    got64m = st.to_numpy()
    assert rel_l2(got64m, s64) < 1e-5
    fid = abs(np.vdot(f, from_interleaved(got64m))) ** 2
    print(f"k={k} FP64 matrices: 1-F vs FP64 {1 - fid:.2e}", flush=True)
    assert fid > 1 - 1e-6


def test_grid20_complex128(grid20):
    text, _, s64 = grid20
    c = driver.Circuit(text)
    c.fuse(4)
    ss = qsim_amd.StateSpace(np.float64)
    st = ss.create(20)
    ss.set_state_zero(st)
    c.run(st)
    assert rel_l2(st.to_numpy(), s64) < 1e-12


def test_grid20_matches_reference_fixture(grid20):
    text, s32, _ = grid20
    g = reference_runs()["grid20"]
    c = driver.Circuit(text)
    c.fuse(2)
    ss = qsim_amd.StateSpace(np.float32)
    st = ss.create(20)
    ss.set_state_zero(st)
    c.run(st)
    got = from_interleaved(st.to_numpy())
    for i, (re, im) in g["avx"]["amplitudes"].items():
        assert abs(got[int(i)] - complex(re, im)) < 1e-6
    assert abs(ss.norm(st) - g["avx"]["norm"]) < 1e-6


def test_random_all_gate_circuits(grid20):
    # Every gate name of the format, controls included, fused up to k=4.
    for case in reference_runs()["random10"]:
        c = driver.Circuit(case["circuit"])
        c.fuse(4)
        ref = np.array(case["state"])
        ref = ref[:, 0] + 1j * ref[:, 1]
        for dtype, tol in ((np.float32, 1e-5), (np.float64, 1e-12)):
            ss = qsim_amd.StateSpace(dtype)
            st = ss.create(c.num_qubits)
            ss.set_state_zero(st)
            c.run(st)
            assert rel_l2(from_interleaved(st.to_numpy()), ref) < tol


def test_measurement_circuit_matches_oracle():
    # Mid-circuit measurements: same mt19937 draws, same collapsed state.
    text = ("6\n0 h 0\n0 h 1\n0 h 2\n0 h 3\n0 h 4\n0 h 5\n1 cz 0 1\n"
            "1 cz 2 3\n2 m 0 2 4\n3 fs 1 5 0.4 0.2\n4 rx 3 0.3\n5 m 1 3 5\n")
    for seed in (1, 2, 3, 7):
        s_or, mres = oracle.Circuit(text, np.float64).run(seed=seed)
        c = driver.Circuit(text)
        c.fuse(2)
        ss = qsim_amd.StateSpace(np.float64)
        st = ss.create(6)
        ss.set_state_zero(st)
        c.run(st, seed=seed)
        assert rel_l2(st.to_numpy(), s_or) < 1e-12, seed


@pytest.mark.parametrize("tile_bits,low_bits",
                         [(14, 4), (14, 5), (13, 5), (12, 6), (10, 3)])
def test_grid20_blocked_vs_oracle(grid20, tile_bits, low_bits):
    # Cache-blocked passes (scaled FP16 MFMA, FP64 fused matrices): the
    # state meets F >= 1 - 1e-6 against the FP64 oracle and stays within
    # 1e-5 relative L2 of it and of the unfused FP32 run.
    text, s32, s64 = grid20
    c = driver.Circuit(text)
    npass = c.fuse_blocked(tile_bits, low_bits)
    ss = qsim_amd.StateSpace(np.float32)
    st = ss.create(20)
    ss.set_state_zero(st)
    c.run(st)
    got = st.to_numpy()
    fid = abs(np.vdot(from_interleaved(s64), from_interleaved(got))) ** 2
    err = rel_l2(got, s64)
    print(f"blocked {tile_bits}/{low_bits}: {npass} passes, rel-L2 vs FP64 "
          f"{err:.2e}, 1-F {1 - fid:.2e}", flush=True)
    assert err < 1e-5
    assert fid > 1 - 1e-6
    # Against the unfused complex64 run (every gate its own pass).
    cu = driver.Circuit(text)
    cu.fuse(1, driver.FUSER_NONE)
    su = ss.create(20)
    ss.set_state_zero(su)
    cu.run(su)
    err_u = rel_l2(got, su.to_numpy())
    su.free()
    print(f"blocked {tile_bits}/{low_bits}: rel-L2 vs unfused {err_u:.2e}",
          flush=True)
    assert err_u < 1e-5
    # The complex128 replay of the same partition is exact.
    ss64 = qsim_amd.StateSpace(np.float64)
    st64 = ss64.create(20)
    ss64.set_state_zero(st64)
    c.run(st64)
    assert rel_l2(st64.to_numpy(), s64) < 1e-12


def test_blocked_random_circuit_with_controls_and_measurement():
    # Controlled gates and measurements stay single passes between blocked
    # ones; the result equals the oracle's (same measurement draws).
    text = circuits.random_circuit(
        16, 400, seed=11, controlled_fraction=0.05,
        gate_names=("h", "t", "rx", "ry", "rz", "cz", "cnot", "fs", "is"))
    lines = text.strip().split("\n")
    lines.insert(200, "199 m 3 7 11")
    text = "\n".join(lines) + "\n"
    want, mres = oracle.Circuit(text, np.float64).run(seed=3)
    c = driver.Circuit(text)
    c.fuse_blocked(12, 5)
    kinds = [c.pass_kind(i) for i in range(c.num_passes)]
    assert driver.PASS_MEASUREMENT in kinds and driver.PASS_OPERATION in kinds
    ss = qsim_amd.StateSpace(np.float32)
    st = ss.create(16)
    ss.set_state_zero(st)
    c.run(st, seed=3)
    got = st.to_numpy()
    fid = abs(np.vdot(from_interleaved(want), from_interleaved(got))) ** 2
    assert fid > 1 - 1e-6, fid
