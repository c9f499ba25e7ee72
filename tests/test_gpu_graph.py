"""Whole circuits captured as one HIP graph (qsim_driver_capture,
qsim_hip_graph_*): every replay must reproduce the eager run bit for bit,
with the program's own eager work and synchronisations between replays, and
the replayed state must match the CPU oracle's FP64 run of the same circuit
(1e-5 relative L2 for complex64, 1e-12 for complex128)."""

import numpy as np
import pytest

import oracle
import qsim_amd
from qsim_amd import circuits, driver


def _oracle_rel_l2(text, got, dtype):
    ref, _ = oracle.Circuit(text, np.float64).run()
    got = got.astype(np.float64)
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))

pytestmark = pytest.mark.gpu


def _partitioned(text, how):
    c = driver.Circuit(text)
    if how == "k4":
        c.fuse(4)
    elif how == "k6-f64":
        c.fuse(6, f64_matrices=True)
        c.set_method(qsim_amd.METHOD_MFMA_FP16X2)
    elif how == "blocked":
        c.fuse_blocked(13, 5)
    elif how == "k2":
        c.fuse(2)
    return c


def _replay_matches(c, n, dtype, text, replays=3):
    ss = qsim_amd.StateSpace(dtype)
    ref = ss.create(n)
    ss.set_state_zero(ref)
    c.run(ref)
    want = ref.to_numpy()
    st = ss.create(n)
    ss.set_state_zero(st)
    g = c.capture(st)
    # The recording run is an ordinary run.
    assert np.array_equal(st.to_numpy(), want)
    for r in range(replays):
        ss.set_state_zero(st)
        g.launch()
        qsim_amd.synchronize()
        got = st.to_numpy()
        assert np.array_equal(got, want), f"replay {r}"
        # Eager work between replays (a reduction with its own sync).
        assert abs(ss.norm(st) - 1.0) < 1e-4
    err = _oracle_rel_l2(text, got, dtype)
    assert err < (1e-5 if dtype == np.float32 else 1e-12), err
    g.free()
    st.free()
    ref.free()


@pytest.mark.parametrize("how", ["k4", "k6-f64", "blocked", "k2"])
def test_graph_replays_the_circuit(how):
    n = 20
    print(f"graph replay n={n} partition={how}", flush=True)
    text = circuits.grid_circuit(n, 14, seed=2)
    c = _partitioned(text, how)
    _replay_matches(c, n, np.float32, text)


def test_graph_complex128_and_controlled_gates():
    n = 16
    text = circuits.random_circuit(n, 200, seed=5, controlled_fraction=0.2)
    print(f"graph replay n={n} controlled, complex64 and complex128", flush=True)
    c = _partitioned(text, "k4")
    _replay_matches(c, n, np.float32, text)
    _replay_matches(c, n, np.float64, text)


def test_graph_refuses_measurements():
    n = 12
    text = circuits.grid_circuit(n, 4, seed=0).rstrip("\n")
    text += "\n9 m 0 1 2\n"
    c = _partitioned(text, "k4")
    ss = qsim_amd.StateSpace(np.float32)
    st = ss.create(n)
    ss.set_state_zero(st)
    with pytest.raises(qsim_amd.QsimHipError) as e:
        c.capture(st)
    assert e.value.code == qsim_amd.ERR_UNSUPPORTED
    # The library is usable afterwards (no recording left open).
    text2 = circuits.grid_circuit(n, 4, seed=0)
    c2 = _partitioned(text2, "k4")
    _replay_matches(c2, n, np.float32, text2, replays=1)


def test_graph_keeps_working_after_other_passes():
    # Eager passes after the capture reuse the staging ring; the graph's
    # matrices live in its own buffers and are unaffected.
    n = 18
    c = _partitioned(circuits.grid_circuit(n, 10, seed=3), "k6-f64")
    ss = qsim_amd.StateSpace(np.float32)
    ref = ss.create(n)
    ss.set_state_zero(ref)
    c.run(ref)
    want = ref.to_numpy()
    st = ss.create(n)
    ss.set_state_zero(st)
    g = c.capture(st)
    other = _partitioned(circuits.grid_circuit(n, 10, seed=4), "k6-f64")
    for _ in range(3):
        other.run(ref)            # 100+ stagings through the ring
    ss.set_state_zero(st)
    g.launch()
    qsim_amd.synchronize()
    assert np.array_equal(st.to_numpy(), want)
    g.free()


def test_graph_outlives_its_circuit():
    # The plan owns the blocked-pass fragments: re-fusing and freeing the
    # circuit leaves the graph valid.
    n = 18
    text = circuits.grid_circuit(n, 10, seed=6)
    c = _partitioned(text, "blocked")
    ss = qsim_amd.StateSpace(np.float32)
    ref = ss.create(n)
    ss.set_state_zero(ref)
    c.run(ref)
    want = ref.to_numpy()
    st = ss.create(n)
    ss.set_state_zero(st)
    g = c.capture(st)
    c.fuse(4)          # releases the circuit's prepared blocked passes
    del c
    ss.set_state_zero(st)
    g.launch()
    qsim_amd.synchronize()
    assert np.array_equal(st.to_numpy(), want)
    g.free()
