"""The circuit driver (qsim's parser + MultiQubitGateFuser over the backend):
host-side checks that need no GPU."""

import os

import numpy as np
import pytest

import oracle
import qsim_amd
from qsim_amd import circuits, driver

pytestmark = pytest.mark.skipif(not os.path.exists(driver.DRIVER_PATH),
                                reason="driver library not built")


def test_parse_and_fuse_grid_circuit():
    text = circuits.grid_circuit(20, 14, seed=0)
    c = driver.Circuit(text)
    assert c.num_qubits == 20
    assert c.num_gates == oracle.Circuit(text).num_ops
    for k in (2, 3, 4, 5, 6):
        n = c.fuse(k)
        assert n > 0
        gates = 0
        for qs, ng in c.passes():
            assert len(qs) <= k and qs == sorted(qs)
            gates += ng
        assert gates == c.num_gates   # every gate lands in exactly one pass


@pytest.mark.skipif(not os.path.exists(oracle.REF_BIN), reason="no reference")
def test_partition_matches_reference_fuser(tmp_path):
    # Same fuser, same circuit: the driver's pass count equals qsim's own.
    text = circuits.grid_circuit(24, 10, seed=3)
    cf = tmp_path / "c.txt"
    cf.write_text(text)
    c = driver.Circuit(text)
    for k in (2, 4, 6):
        n = c.fuse(k)
        info = oracle.run_reference(str(cf), mode="avx", max_fused=k, passes=0)
        assert info["fused_ops"] == n


def test_bad_circuit_rejected():
    with pytest.raises(ValueError):
        driver.Circuit("3\n0 h 7\n")


def _run_blocked_fp64(c, n):
    """The blocked partition replayed gate by gate on the CPU in FP64."""
    s = oracle.state_zero(n, np.float64)
    for i in range(c.num_passes):
        assert c.pass_kind(i) == driver.PASS_BLOCKED
        for qs, m in c.block_info(i):
            oracle.apply_gate(n, s, qs, m)
    return s


@pytest.mark.parametrize("tile_bits,low_bits", [(13, 5), (11, 3), (9, 4),
                                                (12, 7), (14, 5)])
def test_blocked_plan_structure(tile_bits, low_bits):
    text = circuits.grid_circuit(20, 14, seed=0)
    c = driver.Circuit(text)
    npass = c.fuse_blocked(tile_bits, low_bits)
    assert npass > 0
    gates = 0
    for i in range(npass):
        assert c.pass_kind(i) == driver.PASS_BLOCKED
        tile, ng = c.pass_info(i)
        gates += ng
        assert tile == sorted(tile) and len(tile) <= tile_bits
        assert tile[:low_bits] == list(range(low_bits))
        ops = c.block_info(i)
        assert 1 <= len(ops) <= 32
        for qs, m in ops:
            assert 1 <= len(qs) <= 4 and set(qs) <= set(tile)
            assert qs == sorted(qs)
            d = 1 << len(qs)
            u = m[0::2].reshape(d, d) + 1j * m[1::2].reshape(d, d)
            assert np.allclose(u.conj().T @ u, np.eye(d), atol=1e-12)
    assert gates == c.num_gates


@pytest.mark.parametrize("n,seed,kind", [(12, 1, "grid"), (14, 2, "grid"),
                                         (12, 3, "random"), (13, 4, "random")])
def test_blocked_plan_reproduces_circuit(n, seed, kind):
    # The blocked partition (gate order inside and across passes, and the
    # FP64 fused matrices) is the same unitary as the circuit: replayed on
    # the oracle it matches the oracle's own gate-by-gate run to 1e-12.
    if kind == "grid":
        text = circuits.grid_circuit(n, 12, seed=seed)
    else:
        names = ("h", "t", "x_1_2", "y_1_2", "hz_1_2", "rx", "ry", "rz",
                 "rxy", "s", "cz", "cnot", "sw", "is", "fs", "cp")
        text = circuits.random_circuit(n, 300, seed=seed, gate_names=names)
    c = driver.Circuit(text)
    c.fuse_blocked(9 if n < 13 else 11, 4)
    got = _run_blocked_fp64(c, n)
    want, _ = oracle.Circuit(text, dtype=np.float64).run()
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < 1e-12, err


def test_blocked_plan_keeps_controlled_gates_in_place():
    text = circuits.random_circuit(12, 200, seed=5, controlled_fraction=0.1,
                                   gate_names=("h", "rx", "cz", "fs", "t"))
    c = driver.Circuit(text)
    npass = c.fuse_blocked(10, 4)
    kinds = [c.pass_kind(i) for i in range(npass)]
    assert driver.PASS_OPERATION in kinds and driver.PASS_BLOCKED in kinds
    gates = sum(c.pass_info(i)[1] for i in range(npass))
    assert gates == c.num_gates


def test_blocked_plan_cuts_hbm_passes_at_32_qubits():
    # 32-qubit depth-20 grid circuit: qsim's fuser needs 82 passes at k = 4;
    # 13-qubit blocked passes need far fewer.
    text = circuits.grid_circuit(32, 20, seed=0)
    c = driver.Circuit(text)
    k4 = c.fuse(4)
    blocked = c.fuse_blocked(13, 5)
    print("32q passes: mqubit k=4", k4, "blocked 13/5", blocked)
    assert blocked * 2 < k4


def test_blocked_plan_is_deterministic_and_restarts_pay():
    # The restart search (seeded jitter over the greedy frontier) gives the
    # same partition on every call; at 32 qubits, 14-qubit tiles with a
    # 4-qubit low run, it needs at most 21 passes.
    text = circuits.grid_circuit(32, 20, seed=0)
    plans = []
    for _ in range(2):
        c = driver.Circuit(text)
        npass = c.fuse_blocked(14, 4)
        plans.append([(c.pass_info(i), [(qs, m.tobytes())
                                        for qs, m in c.block_info(i)])
                      for i in range(npass)])
    assert plans[0] == plans[1]
    print("32q blocked 14/4 passes:", len(plans[0]))
    assert len(plans[0]) <= 21


def test_blocked_plan_rejects_small_circuits():
    c = driver.Circuit(circuits.grid_circuit(8, 4, seed=0))
    with pytest.raises(qsim_amd.QsimHipError) as e:
        c.fuse_blocked(13, 5)
    assert e.value.code == -4 and c.num_passes == 0
