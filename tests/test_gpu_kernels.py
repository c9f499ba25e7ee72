"""Parity of the HIP kernels (through the C ABI) with the CPU oracle.

Bar: complex64 results within 1e-6 relative L2 (per pass) / 1e-5 (whole
circuits) of the oracle's FP32 result, complex128 within 1e-12; sampled
bitstrings and measured bits identical for identical draws.
"""

import numpy as np
import pytest

import oracle
import qsim_amd
from helpers import (dense_gate, expectation2_circuit, from_interleaved,
                     gate_op, known_answers, pauli_string_matrix,
                     random_matrix, random_state, reference_runs, rel_l2,
                     to_interleaved)
from qsim_amd import circuits

pytestmark = pytest.mark.gpu

DTYPES = [np.float32, np.float64]
TOL = {np.float32: 1e-6, np.float64: 1e-12}


@pytest.fixture(scope="module")
def ss32():
    return qsim_amd.StateSpace(np.float32)


@pytest.fixture(scope="module")
def sim():
    return qsim_amd.Simulator()


def upload(a, n, dtype):
    st = qsim_amd.State(n, dtype)
    st.from_numpy(a)
    return st


def test_device_is_gfx950():
    name, mem, cus = qsim_amd.device_info()
    print("device:", name, mem, cus, flush=True)
    assert "gfx950" in name


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("test", ["TestApplyGate1", "TestApplyGate2",
                                  "TestApplyGate3", "TestApplyGate5"])
def test_known_answer_apply_gate(test, dtype, sim):
    t = known_answers()["apply_gate"][test]
    n = t["num_qubits"]
    ss = qsim_amd.StateSpace(dtype)
    st = ss.create(n)
    ss.set_state_zero(st)
    checks = 0
    for step in t["steps"]:
        if "apply" in step:
            for spec in step["apply"]:
                qs, m, cqs, cmask = gate_op(spec, dtype)
                sim.apply_controlled_gate(qs, cqs, cmask, m, st)
        else:
            i, re, im = step["check"]
            a = ss.get_ampl(st, i)
            assert abs(a.real - re) < 1e-6 and abs(a.imag - im) < 1e-6
            checks += 1
    assert checks > 0
    assert abs(ss.norm(st) - 1) < 1e-6


@pytest.mark.parametrize("dtype", DTYPES)
def test_known_answer_controlled_circuit(dtype, sim):
    t = known_answers()["controlled_circuit"]
    n = t["num_qubits"]
    ss = qsim_amd.StateSpace(dtype)
    st = ss.create(n)
    ss.set_state_zero(st)
    for spec in t["gates"]:
        qs, m, cqs, cmask = gate_op(spec, dtype)
        sim.apply_controlled_gate(qs, cqs, cmask, m, st)
    s = st.to_numpy()
    exp = np.array(t["expected"])
    got = np.stack([s[0:2 * len(exp):2], s[1:2 * len(exp):2]], axis=1)
    assert np.max(np.abs(got - exp)) < 1e-6
    # Same gates twice give bit-identical states (the fixture's determinism
    # check, tests/simulator_testfixture.h:739-765).
    st2 = ss.create(n)
    ss.set_state_zero(st2)
    for spec in t["gates"]:
        qs, m, cqs, cmask = gate_op(spec, dtype)
        sim.apply_controlled_gate(qs, cqs, cmask, m, st2)
    np.testing.assert_array_equal(st2.to_numpy(), s)


@pytest.mark.parametrize("dtype", DTYPES)
def test_multi_qubit_gates_closed_form(dtype, sim):
    # tests/simulator_testfixture.h:1146-1213 up to 10 + log2(64) qubits.
    ss = qsim_amd.StateSpace(dtype)
    for n in range(2, 17):
        st = ss.create(n)
        size = 1 << n
        inorm = np.sqrt(1.0 / size)
        for q in range(1, min(6, n - 1) + 1):
            size1 = 1 << q
            size2 = size1 * size1
            m = np.arange(1, 2 * size2 + 1, dtype=dtype)
            for k in range(0, n - q + 1):
                ss.set_state_uniform(st)
                sim.apply_gate(list(range(k, k + q)), m, st)
                s = st.to_numpy()
                j = (np.arange(size) >> k) & (size1 - 1)
                er = size2 * (1 + 2 * j) * inorm
                ei = er + size1 * inorm
                assert np.max(np.abs(s[0::2] / er - 1)) < 1e-6, (n, q, k)
                assert np.max(np.abs(s[1::2] / ei - 1)) < 1e-6, (n, q, k)


def _target_sets(n, k, rng):
    """Low, high, mixed and random target sets (exercises both kernels)."""
    sets = set()
    if k == 0:
        return [[]]
    sets.add(tuple(range(k)))
    sets.add(tuple(range(n - k, n)))
    for _ in range(4):
        sets.add(tuple(sorted(rng.choice(n, size=k, replace=False).tolist())))
    return [list(s) for s in sets]


@pytest.mark.parametrize("dtype", DTYPES)
def test_apply_gate_vs_oracle(dtype, sim):
    rng = np.random.default_rng(11)
    for n in list(range(1, 13)) + [14, 17, 20]:
        for k in range(0, min(n, 6) + 1):
            for qs in _target_sets(n, k, rng):
                u = random_matrix(k, rng, unitary=bool(rng.integers(2)))
                m = to_interleaved(u, dtype)
                s = random_state(n, rng, dtype)
                st = upload(s, n, dtype)
                sim.apply_gate(qs, m, st)
                oracle.apply_gate(n, s, qs, m)
                err = rel_l2(st.to_numpy(), s)
                assert err < TOL[dtype], (n, k, qs, err)


@pytest.mark.parametrize("dtype", DTYPES)
def test_apply_gate_f64_matrix_on_f32_state(dtype, sim):
    # Matrices may arrive in FP64 (host-built fused matrices).
    rng = np.random.default_rng(12)
    n = 12
    for k in range(1, 7):
        qs = sorted(rng.choice(n, size=k, replace=False).tolist())
        u = random_matrix(k, rng)
        s = random_state(n, rng, np.float32)
        st = upload(s, n, np.float32)
        sim.apply_gate(qs, to_interleaved(u, np.float64), st)
        oracle.apply_gate(n, s, qs, to_interleaved(u, np.float32))
        assert rel_l2(st.to_numpy(), s) < 1e-6


@pytest.mark.parametrize("dtype", DTYPES)
def test_controlled_gate_vs_oracle(dtype, sim):
    rng = np.random.default_rng(13)
    for n in (2, 3, 5, 7, 9, 12, 16, 19):
        for _ in range(12):
            k = int(rng.integers(0, min(4, n - 1) + 1))
            nc = int(rng.integers(1, min(n - k, 5) + 1))
            perm = rng.permutation(n)
            qs = sorted(perm[:k].tolist())
            cqs = sorted(perm[k:k + nc].tolist())
            cvals = int(rng.integers(0, 1 << nc))
            u = random_matrix(k, rng)
            m = to_interleaved(u, dtype)
            s = random_state(n, rng, dtype)
            st = upload(s, n, dtype)
            sim.apply_controlled_gate(qs, cqs, cvals, m, st)
            oracle.apply_controlled_gate(n, s, qs, cqs, cvals, m)
            err = rel_l2(st.to_numpy(), s)
            assert err < TOL[dtype], (n, qs, cqs, cvals, err)


@pytest.mark.parametrize("dtype", DTYPES)
def test_expectation_value_vs_oracle(dtype, sim):
    rng = np.random.default_rng(14)
    for n in (1, 2, 4, 6, 9, 13, 18):
        for k in range(1, min(n, 6) + 1):
            for qs in _target_sets(n, k, rng)[:3]:
                u = random_matrix(k, rng, unitary=False)
                m = to_interleaved(u, dtype)
                s = random_state(n, rng, dtype)
                st = upload(s, n, dtype)
                got = sim.expectation_value(qs, m, st)
                ref = oracle.expectation_value(n, s, qs, m)
                tol = 2e-6 if dtype == np.float32 else 1e-12
                assert abs(got - ref) < tol, (n, k, qs, got, ref)


def test_expectation_value2_known_answer(sim):
    exp = known_answers()["expectation2"]["expected_real"]
    n, ops = expectation2_circuit()
    for dtype in DTYPES:
        ss = qsim_amd.StateSpace(dtype)
        st = ss.create(n)
        ss.set_state_zero(st)
        for spec in ops:
            qs, m, _, _ = gate_op(spec, dtype)
            sim.apply_gate(qs, m, st)
        for k in range(1, 7):
            pm = to_interleaved(pauli_string_matrix(k), dtype)
            total = 0j
            for i in range(n - k + 1):
                total += (0.1 + 0.2 * i) * sim.expectation_value(
                    list(range(i, i + k)), pm, st)
            assert abs(total.real - exp[k - 1]) < 1e-6
            assert abs(total.imag) < 1e-8


@pytest.mark.parametrize("dtype", DTYPES)
def test_reductions_vs_oracle(dtype):
    rng = np.random.default_rng(15)
    ss = qsim_amd.StateSpace(dtype)
    for n in (0, 1, 3, 8, 11, 16, 21):
        a = random_state(n, rng, dtype) * dtype(1.3)
        b = random_state(n, rng, dtype)
        sa, sb = upload(a, n, dtype), upload(b, n, dtype)
        tol = 1e-12
        assert abs(ss.norm(sa) - oracle.norm(n, a)) < tol
        assert abs(ss.inner_product(sa, sb) - oracle.inner_product(n, a, b)) < tol
        assert abs(ss.real_inner_product(sa, sb)
                   - oracle.real_inner_product(n, a, b)) < tol
    # Size mismatch -> NaN (lib/statespace_cuda.h:219-223).
    assert np.isnan(ss.real_inner_product(upload(a, n, dtype), ss.create(n - 1)))


@pytest.mark.parametrize("dtype", DTYPES)
def test_sample_identical_to_oracle(dtype):
    rng = np.random.default_rng(16)
    ss = qsim_amd.StateSpace(dtype)
    for n in (1, 2, 5, 9, 12, 13, 16, 20):
        s = random_state(n, rng, dtype)
        st = upload(s, n, dtype)
        norm = oracle.serial_norm(n, s)
        for num, seed in ((1, 3), (1000, 0), (20000, 9)):
            draws = oracle.generate_random_values(num, seed, norm)
            got = ss.sample_with_draws(st, draws[:-1])
            ref = oracle.sample_with_draws(n, s, draws)
            np.testing.assert_array_equal(got, ref)
        got = ss.sample(st, 1000, 5)
        np.testing.assert_array_equal(got, oracle.sample(n, s, 1000, 5))


def test_sample_draws_past_the_end(ss32):
    # Draws at or beyond the total map to 2^n - 1 (lib/statespace_basic.h:227-229).
    n = 10
    rng = np.random.default_rng(17)
    s = random_state(n, rng, np.float32)
    st = upload(s, n, np.float32)
    total = oracle.serial_norm(n, s)
    draws = np.array([0.0, total * 0.5, total, total * 2])
    got = ss32.sample_with_draws(st, draws)
    ref = oracle.sample_with_draws(n, s, np.append(draws, total * 2))
    np.testing.assert_array_equal(got, ref)
    assert got[-1] == (1 << n) - 1


@pytest.mark.parametrize("dtype", DTYPES)
def test_measurement_vs_oracle(dtype):
    rng = np.random.default_rng(18)
    ss = qsim_amd.StateSpace(dtype)
    for n in (3, 10, 14, 18):
        s = random_state(n, rng, dtype)
        st = upload(s, n, dtype)
        pn = ss.partial_norms(st)
        assert abs(pn.sum() - oracle.norm(n, s)) < 1e-12
        c = from_interleaved(s)
        p = np.abs(c.astype(np.complex64 if dtype == np.float32 else np.complex128)) ** 2
        csum = np.cumsum(p.astype(np.float64))
        for trial in range(20):
            qubits = sorted(rng.choice(n, size=int(rng.integers(1, n + 1)),
                                       replace=False).tolist())
            mask = sum(1 << q for q in qubits)
            r_total = rng.uniform(0, csum[-1])
            res = ss.virtual_measure(qubits, lambda norm: r_total * norm / csum[-1], st)
            k_ref = int(np.searchsorted(csum, r_total * pn.sum() / csum[-1], side="right"))
            k_ref = min(k_ref, (1 << n) - 1)
            # Allow the boundary index only when the draw sits on a rounding tie.
            assert res[1] == (k_ref & mask) or abs(csum[k_ref] - r_total) < 1e-12
        mask, bits = 0b101, 0b100
        ss.collapse(st, mask, bits)
        oracle.collapse(n, s, mask, bits)
        assert rel_l2(st.to_numpy(), s) < TOL[dtype] * 4


@pytest.mark.parametrize("dtype", DTYPES)
def test_state_fills_and_elementwise(dtype):
    ss = qsim_amd.StateSpace(dtype)
    rng = np.random.default_rng(19)
    n = 13
    st = ss.create(n)
    ss.set_state_zero(st)
    s = st.to_numpy()
    assert s[0] == 1 and np.count_nonzero(s) == 1
    ss.set_state_uniform(st)
    s = st.to_numpy()
    assert np.all(s[0::2] == dtype(1.0 / np.sqrt(1 << n))) and not s[1::2].any()
    ss.set_all_zeros(st)
    assert not st.to_numpy().any()
    a = random_state(n, rng, dtype)
    b = random_state(n, rng, dtype)
    sa, sb = upload(a, n, dtype), upload(b, n, dtype)
    assert ss.add(sa, sb)
    np.testing.assert_array_equal(sb.to_numpy(), (a + b).astype(dtype))
    ss.multiply(0.37, sa)
    np.testing.assert_array_equal(sa.to_numpy(), a * dtype(0.37))
    for exclude in (False, True):
        sc = upload(a, n, dtype)
        ss.bulk_set_ampl(sc, 0b110, 0b010, complex(0.25, -0.5), exclude)
        ref = from_interleaved(a).copy()
        sel = ((np.arange(1 << n) & 0b110) == 0b010) != exclude
        ref[sel] = complex(0.25, -0.5)
        np.testing.assert_array_equal(from_interleaved(sc.to_numpy()), ref)
    ss.set_ampl(sa, 77, complex(1.5, -2.5))
    assert ss.get_ampl(sa, 77) == complex(1.5, -2.5)
    sd = ss.create(n)
    assert ss.copy(sa, sd)
    np.testing.assert_array_equal(sd.to_numpy(), sa.to_numpy())
    assert not ss.copy(sa, ss.create(n + 1))
    assert not ss.add(sa, ss.create(n + 1))


def test_grid20_circuit_vs_oracle_and_reference(ss32, sim):
    # The 20-qubit / 14-cycle grid circuit, gate by gate on the GPU.
    txt = circuits.grid_circuit(20, 14, seed=0)
    c = oracle.Circuit(txt, np.float32)
    s_or, _ = c.run()
    st = ss32.create(20)
    ss32.set_state_zero(st)
    n, ops = circuits.grid_circuit_ops(4, 5, 14, 0)
    for t, name, qs, prm in ops:
        q, m = oracle.gate_matrix(name, list(qs), list(prm), np.float32)
        sim.apply_gate(q, m, st)
    s = st.to_numpy()
    assert rel_l2(s, s_or) < 1e-5
    g = reference_runs()["grid20"]
    amps = g["avx"]["amplitudes"]
    cplx = from_interleaved(s)
    for i, (re, im) in amps.items():
        assert abs(cplx[int(i)] - complex(re, im)) < 1e-6
    # Identical draws on the identical state: identical bitstrings.
    bits = ss32.sample(st, 1000, 0)
    np.testing.assert_array_equal(bits, oracle.sample(20, s, 1000, 0))
    # Against the reference's own, independently evolved FP32 state the
    # cumulative sums differ by ~1e-9, which moves a few of 1000 draws across
    # an amplitude boundary (each such flip lands on a neighbouring index).
    ref_bits = np.array(g["samples_seed0"], dtype=np.uint64)
    flips = np.flatnonzero(bits != ref_bits)
    print(f"grid20: {len(flips)} of 1000 samples differ from the reference's "
          f"independently evolved state", flush=True)
    assert len(flips) <= 10


def test_large_state_round_trip(sim):
    # 27 qubits: gates followed by their inverses restore the state; the norm
    # is preserved by every pass (size-independent properties).
    n = 27
    rng = np.random.default_rng(20)
    ss = qsim_amd.StateSpace(np.float32)
    st = ss.create(n)
    ss.set_state_uniform(st)
    ref = ss.create(n)
    ss.set_state_uniform(ref)
    gates = []
    for k in (1, 2, 3, 4, 5, 6, 2, 3):
        qs = sorted(rng.choice(n, size=k, replace=False).tolist())
        gates.append((qs, random_matrix(k, rng)))
    for qs, u in gates:
        sim.apply_gate(qs, to_interleaved(u, np.float32), st)
        assert abs(ss.norm(st) - 1) < 1e-5
    for qs, u in reversed(gates):
        sim.apply_gate(qs, to_interleaved(u.conj().T, np.float32), st)
    ip = ss.inner_product(ref, st)
    assert abs(ip - 1) < 1e-5


def test_error_paths(ss32, sim):
    st = ss32.create(5)
    with pytest.raises(qsim_amd.QsimHipError) as e:
        sim.apply_gate(list(range(7)), np.zeros(2 << 14, np.float32), ss32.create(8))
    assert e.value.code == -4
    with pytest.raises(qsim_amd.QsimHipError) as e:
        sim.apply_gate([3, 1], np.zeros(32, np.float32), st)
    assert e.value.code == -1
    with pytest.raises(qsim_amd.QsimHipError):
        sim.apply_gate([5], np.zeros(8, np.float32), st)
    with pytest.raises(qsim_amd.QsimHipError):
        sim.apply_controlled_gate([1], [1], 1, np.zeros(8, np.float32), st)
    # Allocation beyond the device -> Null (None), not an exception.
    assert qsim_amd.StateSpace(np.float64).create(40) is None


@pytest.mark.parametrize("method", [qsim_amd.METHOD_VECTOR, qsim_amd.METHOD_AUTO])
def test_fp32_methods_vs_oracle(method, sim):
    # FP32 vector kernels and FP32 MFMA (k >= 4) meet the same 1e-6 bar, for
    # low, high, mixed and random targets.
    rng = np.random.default_rng(21)
    qsim_amd.set_gate_method(method)
    try:
        for n in (9, 10, 11, 12, 13, 16, 19):
            for k in (4, 5, 6):
                if k + 5 > n:
                    continue
                for qs in _target_sets(n, k, rng):
                    u = random_matrix(k, rng)
                    m = to_interleaved(u, np.float32)
                    s = random_state(n, rng, np.float32)
                    st = upload(s, n, np.float32)
                    sim.apply_gate(qs, m, st)
                    oracle.apply_gate(n, s, qs, m)
                    err = rel_l2(st.to_numpy(), s)
                    assert err < 1e-6, (method, n, k, qs, err)
    finally:
        qsim_amd.set_gate_method(qsim_amd.METHOD_AUTO)


def test_bf16x3_mfma_accuracy(sim):
    # Split-BF16 x3 products carry ~2^-16 relative error per product: per pass
    # the state stays within 5e-5 relative L2 of the FP64 result.
    rng = np.random.default_rng(22)
    qsim_amd.set_gate_method(qsim_amd.METHOD_MFMA_BF16X3)
    try:
        for n in (11, 12, 16, 20):
            for k in (5, 6):
                for qs in _target_sets(n, k, rng):
                    u = random_matrix(k, rng)
                    s = random_state(n, rng, np.float64)
                    st = upload(s.astype(np.float32), n, np.float32)
                    sim.apply_gate(qs, to_interleaved(u, np.float64), st)
                    oracle.apply_gate(n, s, qs, to_interleaved(u, np.float64))
                    err = rel_l2(st.to_numpy(), s)
                    assert err < 5e-5, (n, k, qs, err)
    finally:
        qsim_amd.set_gate_method(qsim_amd.METHOD_AUTO)


def test_fp16x2_mfma_accuracy(sim):
    # FP16 with a power-of-two scale per matrix and per group, matrix and
    # amplitudes in two fp16 parts each (22 bits), FP32 accumulation: per
    # pass the state stays within 1e-6 relative L2 of the FP64 result, for
    # low, high, mixed and random targets (LDS-tile and register paths).
    rng = np.random.default_rng(23)
    qsim_amd.set_gate_method(qsim_amd.METHOD_MFMA_FP16X2)
    try:
        for n in (11, 12, 13, 16, 20):
            for k in (5, 6):
                for qs in _target_sets(n, k, rng):
                    u = random_matrix(k, rng)
                    s = random_state(n, rng, np.float64)
                    st = upload(s.astype(np.float32), n, np.float32)
                    sim.apply_gate(qs, to_interleaved(u, np.float64), st)
                    oracle.apply_gate(n, s, qs, to_interleaved(u, np.float64))
                    err = rel_l2(st.to_numpy(), s)
                    assert err < 1e-6, (n, k, qs, err)
    finally:
        qsim_amd.set_gate_method(qsim_amd.METHOD_AUTO)


@pytest.mark.parametrize("method,bar", [
    (qsim_amd.METHOD_AUTO, 1e-6),           # exact FP32, row blocks
    (qsim_amd.METHOD_MFMA_BF16X3, 5e-5),    # split-BF16, row blocks
    (qsim_amd.METHOD_MFMA_FP16X2, 1e-6)])   # FP16 with scaling
def test_k56_row_block_kernels_over_many_sets(sim, method, bar):
    # At 22 qubits every k = 5, 6 kernel has more sets of groups than
    # workgroups (2048 against at most 4 per CU), so each workgroup runs
    # its double-buffered loop: the next set's loads in flight while the
    # current one is computed, for low (LDS set) and high (register)
    # targets.
    rng = np.random.default_rng(24)
    n = 22
    qsim_amd.set_gate_method(method)
    try:
        for k in (5, 6):
            for qs in ([0, 1, 2, 3, 4, 5][:k], [3, 7, 11, 15, 19, 21][:k],
                       [5, 9, 12, 14, 17, 20][:k]):
                u = random_matrix(k, rng)
                s = random_state(n, rng, np.float64)
                st = upload(s.astype(np.float32), n, np.float32)
                sim.apply_gate(qs, to_interleaved(u, np.float64), st)
                oracle.apply_gate(n, s, qs, to_interleaved(u, np.float64))
                err = rel_l2(st.to_numpy(), s)
                print(f"method {method} k={k} qs={qs} relL2 {err:.2e}",
                      flush=True)
                assert err < bar, (method, k, qs, err)
    finally:
        qsim_amd.set_gate_method(qsim_amd.METHOD_AUTO)


@pytest.mark.parametrize("scale", [1e-30, 1e-3, 1.0, 1e3, 1e30])
def test_fp16x2_scaling_over_magnitudes(sim, scale):
    # The per-group scale keeps fp16's 5-bit exponent out of the way: groups
    # of very different size in one state (one group all zero, one tiny, the
    # rest at `scale`), and non-unitary matrices scaled by 1e-3 and 1e3.
    rng = np.random.default_rng(24)
    n = 14
    qsim_amd.set_gate_method(qsim_amd.METHOD_MFMA_FP16X2)
    try:
        for k, qs in ((5, [0, 3, 6, 9, 12]), (6, [5, 6, 7, 8, 9, 13]),
                      (6, [0, 1, 2, 3, 4, 5])):
            for mscale in (1e-3, 1.0, 1e3):
                u = random_matrix(k, rng) * mscale
                s = random_state(n, rng, np.float64) * scale
                s[: 2 << k] = 0.0           # a zero group (low targets)
                s[2 << k: 4 << k] *= 1e-20  # a tiny one
                s = s.astype(np.float32).astype(np.float64)
                st = upload(s.astype(np.float32), n, np.float32)
                sim.apply_gate(qs, to_interleaved(u, np.float64), st)
                oracle.apply_gate(n, s, qs, to_interleaved(u, np.float64))
                err = rel_l2(st.to_numpy(), s)
                assert err < 1e-6, (k, qs, mscale, err)
    finally:
        qsim_amd.set_gate_method(qsim_amd.METHOD_AUTO)
