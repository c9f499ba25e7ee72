"""Cache-blocked passes (qsim_hip_apply_gate_block) against the CPU oracle.

A blocked pass applies a run of up to 32 gates (k <= 4, all inside one tile
qubit set) with one stream of the state through HBM, in split-BF16 MFMA
arithmetic (matrix in three bf16 parts, amplitudes in two).  The oracle
applies the same gates one by one in FP64.  Bar: the complex64 result within
1e-4 relative L2 of the FP64 result after up to 32 gates (the amplitude
split leaves about 2^-18 relative per gate), i.e. fidelity above 1 - 1e-8.
"""

import numpy as np
import pytest

import oracle
import qsim_amd
from helpers import random_matrix, random_state, rel_l2, to_interleaved

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sim():
    return qsim_amd.Simulator()


def upload(a, n, dtype=np.float32):
    st = qsim_amd.State(n, dtype)
    st.from_numpy(a)
    return st


def random_tile(n, lc, tb, rng):
    high = sorted(rng.choice(np.arange(lc, n), size=tb - lc, replace=False))
    return list(range(lc)) + [int(q) for q in high]


def random_program(tile, nops, rng, kmax=4):
    gates = []
    for _ in range(nops):
        k = int(rng.integers(1, kmax + 1))
        qs = sorted(int(q) for q in rng.choice(tile, size=k, replace=False))
        gates.append((qs, to_interleaved(random_matrix(k, rng), np.float64)))
    return gates


def run_case(sim, n, tile, gates, rng):
    s = random_state(n, rng, np.float64)
    st = upload(s.astype(np.float32), n)
    sim.apply_gate_block(tile, gates, st)
    for qs, m in gates:
        oracle.apply_gate(n, s, qs, m)
    got = st.to_numpy()
    err = rel_l2(got, s)
    a = got[0::2].astype(np.float64) + 1j * got[1::2]
    b = s[0::2] + 1j * s[1::2]
    fid = abs(np.vdot(b, a)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)
    return err, fid


@pytest.mark.parametrize("n,lc,tb", [(9, 9, 9), (12, 5, 9), (13, 13, 13),
                                     (14, 3, 13), (16, 5, 13), (16, 7, 12),
                                     (18, 6, 11), (20, 5, 13), (22, 8, 13),
                                     (14, 14, 14), (17, 5, 14), (20, 9, 14),
                                     (22, 3, 14)])
def test_block_random_programs(sim, n, lc, tb):
    rng = np.random.default_rng(1000 * n + 10 * lc + tb)
    for trial in range(3):
        tile = random_tile(n, lc, tb, rng)
        nops = int(rng.integers(1, 33))
        gates = random_program(tile, nops, rng)
        err, fid = run_case(sim, n, tile, gates, rng)
        print(f"n={n} tile={tile} ops={nops} relL2={err:.2e} 1-F={1 - fid:.2e}",
              flush=True)
        assert err < 1e-4, (n, tile, nops, err)
        assert fid > 1 - 1e-8


def test_block_max_program_and_single_gates(sim):
    rng = np.random.default_rng(5)
    n = 17
    tile = list(range(5)) + [6, 9, 11, 13, 14, 15, 16, 5][:8]
    tile = sorted(tile)
    gates = random_program(tile, 32, rng)
    err, fid = run_case(sim, n, tile, gates, rng)
    assert err < 1e-4 and fid > 1 - 1e-8
    # One-qubit gates on every tile qubit, including the highest and lowest.
    gates = [([q], to_interleaved(random_matrix(1, rng), np.float64))
             for q in tile]
    err, fid = run_case(sim, n, tile, gates, rng)
    assert err < 5e-5 and fid > 1 - 1e-8


def test_block_matches_gate_passes_on_grid_layers(sim):
    # The same gates through plain gate passes (FP32 vector kernels) and
    # through one blocked pass agree to the BF16x3 bar.
    rng = np.random.default_rng(9)
    n = 20
    tile = list(range(5)) + [7, 8, 12, 13, 17, 18, 19, 16]
    tile = sorted(tile)
    gates = random_program(tile, 24, rng, kmax=2)
    s = random_state(n, rng, np.float32)
    a = upload(s, n)
    b = upload(s, n)
    sim.apply_gate_block(tile, gates, a)
    for qs, m in gates:
        sim.apply_gate(qs, m, b)
    assert rel_l2(a.to_numpy(), b.to_numpy()) < 1e-4


@pytest.mark.parametrize("bad", ["outside", "unsorted", "k5", "too_many",
                                 "short_run", "small_tile", "big_tile"])
def test_block_rejects_bad_programs(sim, bad):
    rng = np.random.default_rng(3)
    n = 16
    tile = list(range(5)) + [8, 9, 10, 12, 14, 15]
    gates = random_program(tile, 4, rng)
    if bad == "outside":
        gates[1] = ([6], gates[1][1][:8])
    elif bad == "unsorted":
        gates[0] = ([9, 2], to_interleaved(random_matrix(2, rng), np.float64))
    elif bad == "k5":
        gates[0] = ([0, 1, 2, 3, 4], to_interleaved(random_matrix(5, rng),
                                                    np.float64))
    elif bad == "too_many":
        gates = random_program(tile, 33, rng)
    elif bad == "short_run":
        tile = [0, 1, 3, 4, 5, 6, 7, 8, 9, 10]
        gates = [([3], to_interleaved(random_matrix(1, rng), np.float64))]
    elif bad == "small_tile":
        tile = list(range(8))
        gates = [([3], to_interleaved(random_matrix(1, rng), np.float64))]
    elif bad == "big_tile":
        tile = list(range(15))
        gates = [([3], to_interleaved(random_matrix(1, rng), np.float64))]
    st = upload(random_state(n, rng), n)
    before = st.to_numpy()
    with pytest.raises(qsim_amd.QsimHipError) as e:
        sim.apply_gate_block(tile, gates, st)
    assert e.value.code == -1
    assert np.array_equal(st.to_numpy(), before)


def test_block_rejects_complex128(sim):
    rng = np.random.default_rng(4)
    st = upload(random_state(12, rng, np.float64), 12, np.float64)
    gates = [([0], to_interleaved(random_matrix(1, rng), np.float64))]
    with pytest.raises(qsim_amd.QsimHipError) as e:
        sim.apply_gate_block(list(range(10)), gates, st)
    assert e.value.code == -4


def test_prepared_block_runs_repeatedly(sim):
    # A prepared pass (fragments uploaded once) gives the one-shot pass's
    # result bit for bit, run after run, and refuses a state of another size.
    rng = np.random.default_rng(12)
    n = 18
    tile = list(range(5)) + [6, 9, 10, 13, 15, 16, 17, 12]
    gates = random_program(sorted(tile), 20, rng)
    s = random_state(n, rng)
    a, b = upload(s, n), upload(s, n)
    pb = sim.prepare_gate_block(n, tile, gates)
    for _ in range(3):
        pb.run(a)
        sim.apply_gate_block(tile, gates, b)
        assert np.array_equal(a.to_numpy(), b.to_numpy())
    other = upload(random_state(n - 1, rng), n - 1)
    with pytest.raises(qsim_amd.QsimHipError) as e:
        pb.run(other)
    assert e.value.code == -1
    pb.free()


def test_accurate_product_on_a_structured_state(sim):
    # The first pass from |0>: every amplitude of the 16-qubit state after
    # one layer of identical single-qubit gates has the same magnitude, so
    # the two-part amplitude split rounds them all alike and the norm moves
    # coherently; the three-part split (QSIM_HIP_BLOCK_ACCURATE) keeps it
    # within FP32 rounding.  Both stay within the blocked-pass bar.
    rng = np.random.default_rng(21)
    n = 16
    tile = list(range(5)) + [7, 9, 10, 12, 13, 14, 15, 11, 6]
    tile = sorted(tile)
    h = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    gates = [([q], to_interleaved(h.astype(complex), np.float64))
             for q in tile]
    gates += random_program(tile, 20, rng, kmax=2)
    s0 = np.zeros(2 << n)
    s0[0] = 1.0
    ref = s0.copy()
    for qs, m in gates[:32]:
        oracle.apply_gate(n, ref, qs, m)
    drift = {}
    for accurate in (False, True):
        st = upload(s0.astype(np.float32), n)
        sim.apply_gate_block(tile, gates[:32], st, accurate=accurate)
        got = st.to_numpy().astype(np.float64)
        drift[accurate] = abs(np.sum(got * got) - 1.0)
        assert rel_l2(got, ref) < 1e-4
    print(f"norm drift: two-part {drift[False]:.2e}, three-part "
          f"{drift[True]:.2e}", flush=True)
    assert drift[True] < 1e-6


def test_accurate_prepared_pass_matches_oracle(sim):
    rng = np.random.default_rng(22)
    n = 18
    tile = sorted(list(range(5)) + [6, 9, 10, 13, 15, 16, 17, 12])
    gates = random_program(tile, 24, rng)
    s = random_state(n, rng, np.float64)
    st = upload(s.astype(np.float32), n)
    pb = sim.prepare_gate_block(n, tile, gates, accurate=True)
    pb.run(st)
    pb.free()
    for qs, m in gates:
        oracle.apply_gate(n, s, qs, m)
    # Three-part amplitudes: FP32-level agreement, well inside the
    # two-part bar of 1e-4.
    assert rel_l2(st.to_numpy(), s) < 2e-6
