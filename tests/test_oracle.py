"""Pins the CPU oracle (oracle/qsim_oracle.cc) before it is trusted.

Each check is chosen so that a plausible mistake in the restatement (a wrong
sign, a transposed matrix, a swapped qubit order, a missing control value)
fails it: qsim's own known-answer tests, analytic closed forms, and dense
2^n x 2^n products built independently with numpy.
"""

import numpy as np
import pytest

import oracle
from helpers import (dense_gate, expectation2_circuit, from_interleaved,
                     gate_op, known_answers, pauli_string_matrix,
                     random_matrix, random_state, rel_l2, to_interleaved)

DTYPES = [np.float32, np.float64]


def run_steps(n, steps, dtype):
    s = oracle.state_zero(n, dtype)
    checks = []
    for st in steps:
        if "apply" in st:
            for spec in st["apply"]:
                qs, m, cqs, cmask = gate_op(spec, dtype)
                oracle.apply_controlled_gate(n, s, qs, cqs, cmask, m)
        else:
            i, re, im = st["check"]
            checks.append((s[2 * i], s[2 * i + 1], re, im))
    return s, checks


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("test", ["TestApplyGate1", "TestApplyGate2",
                                  "TestApplyGate3", "TestApplyGate5"])
def test_known_answer_apply_gate(test, dtype):
    # tests/simulator_testfixture.h:36-358: amplitudes to 1e-6.
    t = known_answers()["apply_gate"][test]
    s, checks = run_steps(t["num_qubits"], t["steps"], dtype)
    assert checks
    for got_re, got_im, re, im in checks:
        assert abs(got_re - re) < 1e-6 and abs(got_im - im) < 1e-6
    assert abs(oracle.norm(t["num_qubits"], s) - 1) < 1e-6


@pytest.mark.parametrize("dtype", DTYPES)
def test_known_answer_controlled_circuit(dtype):
    # tests/simulator_testfixture.h:360-766 (expected_results :661-731).
    t = known_answers()["controlled_circuit"]
    n = t["num_qubits"]
    s = oracle.state_zero(n, dtype)
    for spec in t["gates"]:
        qs, m, cqs, cmask = gate_op(spec, dtype)
        oracle.apply_controlled_gate(n, s, qs, cqs, cmask, m)
    exp = np.array(t["expected"])
    got = np.stack([s[0:2 * len(exp):2], s[1:2 * len(exp):2]], axis=1)
    assert np.max(np.abs(got - exp)) < 1e-6


@pytest.mark.parametrize("dtype", DTYPES)
def test_multi_qubit_gates_closed_form(dtype):
    # tests/simulator_testfixture.h:1146-1213: uniform state, matrix entries
    # 1, 2, ... ; the result is known analytically for every qubit offset.
    for n in range(2, 12):
        size = 1 << n
        inorm = np.sqrt(1.0 / size)
        for q in range(1, min(6, n - 1) + 1):
            size1 = 1 << q
            size2 = size1 * size1
            m = np.arange(1, 2 * size2 + 1, dtype=dtype)
            for k in range(0, n - q + 1):
                qs = list(range(k, k + q))
                s = oracle.state_uniform(n, dtype)
                oracle.apply_gate(n, s, qs, m)
                j = (np.arange(size) >> k) & (size1 - 1)
                er = size2 * (1 + 2 * j) * inorm
                ei = er + size1 * inorm
                np.testing.assert_allclose(s[0::2] / er, 1.0, atol=1e-6)
                np.testing.assert_allclose(s[1::2] / ei, 1.0, atol=1e-6)


@pytest.mark.parametrize("dtype", DTYPES)
def test_apply_gate_dense(dtype):
    rng = np.random.default_rng(7)
    tol = 2e-6 if dtype == np.float32 else 1e-12
    for n in range(1, 9):
        for k in range(0, min(n, 6) + 1):
            qs = sorted(rng.choice(n, size=k, replace=False).tolist())
            u = random_matrix(k, rng, unitary=bool(rng.integers(2)))
            m = to_interleaved(u, dtype)
            s = random_state(n, rng, dtype)
            ref = dense_gate(n, qs, to_interleaved(u, np.float64)) @ from_interleaved(s)
            oracle.apply_gate(n, s, qs, m)
            assert rel_l2(from_interleaved(s), ref) < tol


@pytest.mark.parametrize("dtype", DTYPES)
def test_controlled_gate_dense(dtype):
    rng = np.random.default_rng(8)
    tol = 2e-6 if dtype == np.float32 else 1e-12
    for n in range(2, 9):
        for _ in range(6):
            k = int(rng.integers(1, min(4, n - 1) + 1))
            nc = int(rng.integers(1, n - k + 1))
            perm = rng.permutation(n)
            qs = sorted(perm[:k].tolist())
            cqs = sorted(perm[k:k + nc].tolist())
            cvals = int(rng.integers(0, 1 << nc))
            u = random_matrix(k, rng)
            # Dense controlled operator: U on the targets where controls match.
            full = np.eye(1 << n, dtype=np.complex128)
            blk = dense_gate(n, qs, to_interleaved(u, np.float64))
            sel = np.ones(1 << n, dtype=bool)
            for i, c in enumerate(cqs):
                sel &= ((np.arange(1 << n) >> c) & 1) == ((cvals >> i) & 1)
            full[np.ix_(sel, sel)] = blk[np.ix_(sel, sel)]
            s = random_state(n, rng, dtype)
            ref = full @ from_interleaved(s)
            oracle.apply_controlled_gate(n, s, qs, cqs, cvals,
                                         to_interleaved(u, dtype))
            assert rel_l2(from_interleaved(s), ref) < tol


@pytest.mark.parametrize("dtype", DTYPES)
def test_expectation_value_dense(dtype):
    rng = np.random.default_rng(9)
    for n in range(1, 9):
        for k in range(1, min(n, 6) + 1):
            qs = sorted(rng.choice(n, size=k, replace=False).tolist())
            u = random_matrix(k, rng, unitary=False)
            s = random_state(n, rng, dtype)
            psi = from_interleaved(s)
            ref = np.vdot(psi, dense_gate(n, qs, to_interleaved(u, np.float64)) @ psi)
            got = oracle.expectation_value(n, s, qs, to_interleaved(u, dtype))
            assert abs(got - ref) < (5e-6 if dtype == np.float32 else 1e-12)


def test_expectation_value2_known_answer():
    # tests/simulator_testfixture.h:1474-1625, expected_real to 1e-6, imag 1e-8.
    exp = known_answers()["expectation2"]["expected_real"]
    n, ops = expectation2_circuit()
    for dtype in DTYPES:
        s = oracle.state_zero(n, dtype)
        for spec in ops:
            qs, m, _, _ = gate_op(spec, dtype)
            oracle.apply_gate(n, s, qs, m)
        for k in range(1, 7):
            total = 0j
            pm = to_interleaved(pauli_string_matrix(k), dtype)
            for i in range(n - k + 1):
                ev = oracle.expectation_value(n, s, list(range(i, i + k)), pm)
                total += (0.1 + 0.2 * i) * ev
            assert abs(total.real - exp[k - 1]) < 1e-6
            assert abs(total.imag) < 1e-8


@pytest.mark.parametrize("dtype", DTYPES)
def test_norm_and_inner_product(dtype):
    # tests/statespace_testfixture.h:466-512 closed form: states H^n|0> and
    # (H^(n-1) x T H)|0> have <a|b> = 0.5 + 0.5/sqrt(2) (+ i 0.5/sqrt(2)).
    n = 8
    a = oracle.state_zero(n, dtype)
    b = oracle.state_zero(n, dtype)
    for q in range(n):
        qs, m = oracle.gate_matrix("h", [q], dtype=dtype)
        oracle.apply_gate(n, a, qs, m)
        oracle.apply_gate(n, b, qs, m)
    qs, m = oracle.gate_matrix("t", [0], dtype=dtype)
    oracle.apply_gate(n, b, qs, m)
    assert abs(oracle.norm(n, a) - 1) < 1e-6
    ip = oracle.inner_product(n, a, b)
    assert abs(ip.real - (0.5 + 0.5 / np.sqrt(2))) < 1e-6
    assert abs(ip.imag - 0.5 / np.sqrt(2)) < 1e-6
    assert abs(oracle.real_inner_product(n, a, b) - ip.real) < 1e-12


@pytest.mark.parametrize("dtype", DTYPES)
def test_sample_matches_searchsorted(dtype):
    rng = np.random.default_rng(3)
    for n in (1, 3, 6, 10, 14):
        s = random_state(n, rng, dtype)
        p = s[0::2].astype(np.float64) ** 2 + s[1::2].astype(np.float64) ** 2
        csum = np.cumsum(p)   # sequential FP64 running sum
        norm = oracle.serial_norm(n, s)
        draws = oracle.generate_random_values(2000, 11, norm)
        got = oracle.sample_with_draws(n, s, draws)
        ref = np.searchsorted(csum, draws[:-1], side="right")
        ref = np.minimum(ref, (1 << n) - 1)
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(oracle.sample(n, s, 2000, 11), got)


def test_sampling_small_distribution():
    # tests/statespace_testfixture.h:514-558: a 6-qubit state with given
    # probabilities, 2M samples, frequencies within 4e-4.
    n = 6
    probs = np.arange(1, 65, dtype=np.float64)
    probs /= probs.sum()
    s = to_interleaved(np.sqrt(probs), np.float32)
    bits = oracle.sample(n, s, 2_000_000, 1)
    freq = np.bincount(bits.astype(np.int64), minlength=64) / 2e6
    assert np.max(np.abs(freq - probs)) < 4e-4


def test_draws_match_reference_fixture():
    from helpers import reference_runs
    ref = np.array(reference_runs()["draws_seed7_max0.75"])
    got = oracle.generate_random_values(50, 7, 0.75)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("dtype", DTYPES)
def test_collapse_and_measure(dtype):
    rng = np.random.default_rng(5)
    n = 7
    s = random_state(n, rng, dtype)
    c = from_interleaved(s)
    mask, bits = 0b1010010, 0b0010000
    keep = (np.arange(1 << n) & mask) == bits
    ref = np.where(keep, c, 0)
    ref /= np.linalg.norm(ref)
    oracle.collapse(n, s, mask, bits)
    assert rel_l2(from_interleaved(s), ref) < 2e-6


def test_parser_matches_gate_calls():
    text = "3\n0 h 0\n0 h 1\n1 cnot 2 0\n2 fs 1 2 0.3 -0.2\n3 c 0 2 rx 1 0.7\n"
    c = oracle.Circuit(text, np.float64)
    s, _ = c.run()
    ref = oracle.state_zero(3, np.float64)
    for name, qs, prm in (("h", [0], []), ("h", [1], []), ("cnot", [2, 0], []),
                          ("fs", [1, 2], [0.3, -0.2])):
        q, m = oracle.gate_matrix(name, qs, prm, np.float64)
        oracle.apply_gate(3, ref, q, m)
    q, m = oracle.gate_matrix("rx", [1], [0.7], np.float64)
    oracle.apply_controlled_gate(3, ref, q, [0, 2], 3, m)
    np.testing.assert_allclose(s, ref, atol=1e-15)


# ---- cost-model network (mlp_oracle.cc) --------------------------------------

def _mlp_layers(F, H):
    return [(F, H), (H, H), (H, H), (H, 1)]


def _mlp_pack(layers):
    return np.concatenate([np.concatenate([w.reshape(-1), b]) for w, b in
                           layers]).astype(np.float32)


def test_mlp_forward_closed_form():
    # Identity-like hidden layers pass the positive part of the input through;
    # the output sums it: y = sum(relu(x + b1)) + b_out.
    F, H = 3, 4
    w1 = np.zeros((F, H))
    w1[np.arange(F), np.arange(F)] = 1.0
    b1 = np.array([0.0, 0.0, -1.0, 0.25])
    eye = np.eye(H)
    p = _mlp_pack([(w1, b1), (eye, np.zeros(H)), (eye, np.zeros(H)),
                   (np.ones((H, 1)), np.array([0.5]))])
    assert len(p) == oracle.mlp_num_params(F, H)
    x = np.array([[1.0, -2.0, 3.0], [-1.0, 0.5, 0.5], [0.0, 0.0, 0.0]])
    want = [1 + 0 + 2 + 0.25 + 0.5, 0 + 0.5 + 0 + 0.25 + 0.5, 0.25 + 0.5]
    assert np.allclose(oracle.mlp_forward(F, H, p, x), want, atol=1e-12)


def test_mlp_gradient_matches_finite_differences():
    # Central differences of the FP64 loss, with parameters on a 2^-20 grid
    # and a 2^-14 step so that p +- h is exact in float32; every component of
    # the analytic gradient must agree.
    rng = np.random.default_rng(17)
    F, H, B = 5, 6, 7
    n = oracle.mlp_num_params(F, H)
    p = np.round(rng.normal(scale=0.6, size=n) * 2**20) / 2**20
    p = p.astype(np.float32)
    x = rng.normal(size=(B, F)).astype(np.float32)
    t = rng.normal(size=B).astype(np.float32)
    loss, g = oracle.mlp_gradient(F, H, p, x, t)
    h = 2.0 ** -14
    fd = np.zeros(n)
    for i in range(n):
        pp, pm = p.copy(), p.copy()
        pp[i] += np.float32(h)
        pm[i] -= np.float32(h)
        fd[i] = (oracle.mlp_gradient(F, H, pp, x, t)[0] -
                 oracle.mlp_gradient(F, H, pm, x, t)[0]) / (2 * h)
    err = np.abs(fd - g)
    assert np.all(err <= 1e-6 * np.abs(g) + 1e-7), (err.max(), np.argmax(err))
    assert abs(loss - np.mean((oracle.mlp_forward(F, H, p, x) - t) ** 2)) < 1e-12


def test_mlp_gradient_check_catches_a_transposed_weight():
    # A planted mistake (one layer's weight read transposed) is caught by the
    # finite-difference check above: the losses differ.
    rng = np.random.default_rng(3)
    F, H = 4, 4
    layers = [(rng.normal(size=(i, o)), rng.normal(size=o))
              for i, o in _mlp_layers(F, H)]
    p = _mlp_pack(layers)
    bad = _mlp_pack([layers[0], (layers[1][0].T, layers[1][1])] + layers[2:])
    x = rng.normal(size=(5, F)).astype(np.float32)
    assert not np.allclose(oracle.mlp_forward(F, H, p, x),
                           oracle.mlp_forward(F, H, bad, x))
