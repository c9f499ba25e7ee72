"""qsim's Python extension `qsim_hip` on the gfx950 backend
(qsim_amd/pybind/qsim_hip_module.cc) against the reference's own CPU
extension (pybind_interface/avx2, oracle/_ref) on the same circuit objects.

The circuits are built with the reference extension's circuit API and handed
to both extensions, as qsimcirq hands its circuits to whichever simulator
module it selected.  Every entry point family of the GPU bindings is run:
full state (from a basis state and from a given vector), amplitudes,
samplers, expectation values (whole circuit and per moment), trajectories
and the hybrid (qsimh) simulator."""

import numpy as np
import pytest

import pybind_circuits as pc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    ref = pc.load("qsim_avx2", pc.REF_DIR)
    hip = pc.load("qsim_hip", pc.HIP_DIR)
    if ref is None or hip is None:
        pytest.skip("extensions not built (need qsim sources and pybind11)")
    return ref, hip


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("n,controlled", [(10, False), (12, True), (20, False)])
def test_fullstate(mods, n, controlled):
    ref, hip = mods
    print(f"pybind fullstate n={n} controlled={controlled}", flush=True)
    circ, _ = pc.build(ref, n, 8, seed=n, controlled=controlled)
    o = pc.options(circ)
    want = np.asarray(ref.qsim_simulate_fullstate(o, 0))
    got = np.asarray(hip.qsim_simulate_fullstate(o, 0))
    assert got.dtype == np.float32 and got.shape == want.shape
    assert rel(got, want) < 1e-5
    # From a basis state other than |0> and from a given vector.
    want = np.asarray(ref.qsim_simulate_fullstate(o, 5))
    got = np.asarray(hip.qsim_simulate_fullstate(o, 5))
    assert rel(got, want) < 1e-5
    rng = np.random.default_rng(1)
    v = rng.normal(size=2 << n).astype(np.float32)
    v /= np.linalg.norm(v)
    want = np.asarray(ref.qsim_simulate_fullstate(o, v))
    got = np.asarray(hip.qsim_simulate_fullstate(o, v))
    assert rel(got, want) < 1e-5


def test_amplitudes(mods):
    ref, hip = mods
    n = 14
    circ, _ = pc.build(ref, n, 8, seed=3)
    rng = np.random.default_rng(2)
    idx = rng.integers(0, 1 << n, size=64)
    bits = "\n".join("".join(str((int(i) >> p) & 1) for p in range(n))
                     for i in idx)
    o = pc.options(circ, i=bits)
    want = np.asarray(ref.qsim_simulate(o))
    got = np.asarray(hip.qsim_simulate(o))
    assert got.shape == (64,)
    assert rel(got, want) < 1e-5


def test_samplers(mods):
    ref, hip = mods
    n = 12
    circ, _ = pc.build(ref, n, 6, seed=4)
    o = pc.options(circ, s=7)
    want = np.asarray(ref.qsim_sample_final(o, 2000))
    got = np.asarray(hip.qsim_sample_final(o, 2000))
    # Same draws on FP32 states that differ by rounding: a draw within
    # ~1e-7 of a cumulative boundary may land on the neighbour.
    assert got.shape == want.shape
    assert np.mean(got == want) > 0.99
    mcirc, _ = pc.build(ref, n, 6, seed=4, measure_at_end=True)
    o = pc.options(mcirc, s=7)
    want = np.asarray(ref.qsim_sample(o))
    got = np.asarray(hip.qsim_sample(o))
    assert got.shape == want.shape == (n,)
    assert np.array_equal(got, want)


def _opsums(m, n):
    zz = m.OpString()
    zz.weight = 0.5 + 0.25j
    m.add_gate_to_opstring(m.GateKind.kZ, [0], zz)
    m.add_gate_to_opstring(m.GateKind.kZ, [1], zz)
    x = m.OpString()
    x.weight = 1.0
    m.add_gate_to_opstring(m.GateKind.kX, [n - 1], x)
    return [([zz, x], 3), ([x], 1)]


def test_expectation_values(mods):
    ref, hip = mods
    n = 12
    circ, ops = pc.build(ref, n, 6, seed=5)
    o = pc.options(circ)
    want = np.asarray(ref.qsim_simulate_expectation_values(o, _opsums(ref, n), 0))
    got = np.asarray(hip.qsim_simulate_expectation_values(o, _opsums(ref, n), 0))
    assert np.abs(got - want).max() < 2e-6
    # Per moment: after op 20 and after the whole circuit.
    steps = [(20, _opsums(ref, n)), (len(ops), _opsums(ref, n))]
    want = np.asarray(ref.qsim_simulate_moment_expectation_values(o, steps, 0))
    got = np.asarray(hip.qsim_simulate_moment_expectation_values(o, steps, 0))
    assert got.shape == want.shape == (2, 2)
    assert np.abs(got - want).max() < 2e-6


def _noisy(m, n, depth, seed, p=0.02, gamma=0.01):
    """Layers of 1q gates, depolarising, 2q gates, amplitude damping."""
    rng = np.random.default_rng(seed)
    c = m.Circuit()
    c.num_qubits = n
    eye = np.eye(2)
    paulis = [np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]),
              np.diag([1, -1])]
    k0 = np.diag([1, np.sqrt(1 - gamma)])
    k1 = np.array([[0, np.sqrt(gamma)], [0, 0]])
    t = 0
    for layer in range(depth):
        for q in range(n):
            m.add_matrix_gate(t, [q], pc.flat(pc.random_unitary(1, rng)), c)
        for q in range(n):
            m.add_channel(t + 1, [q], [(1 - p, pc.flat(eye), True)] +
                          [(p / 3, pc.flat(s), True) for s in paulis], c)
        for q0 in range(layer % 2, n - 1, 2):
            m.add_gate(m.GateKind.kCZ, t + 2, [q0, q0 + 1], {}, c)
        for q in range(n):
            m.add_channel(t + 3, [q], [(1 - gamma, pc.flat(k0), False),
                                       (0.0, pc.flat(k1), False)], c)
        t += 4
    return c


def test_trajectories(mods):
    ref, hip = mods
    n = 10
    circ = _noisy(ref, n, 5, seed=6)
    o = pc.options(circ, r=1, s=3)
    print("pybind qtrajectory", flush=True)
    want = np.asarray(ref.qtrajectory_simulate_fullstate(o, 0))
    got = np.asarray(hip.qtrajectory_simulate_fullstate(o, 0))
    assert rel(got, want) < 1e-5
    o = pc.options(circ, r=40, s=3)
    want = np.asarray(ref.qtrajectory_simulate_expectation_values(
        o, _opsums(ref, n), 0))
    got = np.asarray(hip.qtrajectory_simulate_expectation_values(
        o, _opsums(ref, n), 0))
    assert np.abs(got - want).max() < 1e-5
    o = pc.options(circ, r=1, s=9)
    want = np.asarray(ref.qtrajectory_sample_final(o, 500))
    got = np.asarray(hip.qtrajectory_sample_final(o, 500))
    assert np.mean(got == want) > 0.99


def test_hybrid(mods):
    ref, hip = mods
    n, cut = 12, 6
    circ, _ = pc.build(ref, n, 4, seed=7, cut=cut)
    rng = np.random.default_rng(8)
    idx = rng.integers(0, 1 << n, size=32)
    bits = "\n".join("".join(str((int(i) >> p) & 1) for p in range(n))
                     for i in idx)
    o = pc.options(circ, i=bits, k=list(range(cut, n)), w=0, p=0, r=0, t=2)
    print("pybind qsimh", flush=True)
    want = np.asarray(ref.qsimh_simulate(o))
    got = np.asarray(hip.qsimh_simulate(o))
    assert got.shape == want.shape == (32,)
    assert rel(got, want) < 1e-5
    # And the hybrid amplitudes are the full state's.
    full = np.asarray(hip.qsim_simulate_fullstate(pc.options(circ), 0))
    full = full.view(np.complex64)
    assert rel(got, full[idx]) < 1e-5
