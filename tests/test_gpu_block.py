"""Cache-blocked passes (qsim_hip_apply_gate_block) against the CPU oracle.

A blocked pass applies a run of up to 32 gates (k <= 4, all inside one tile
qubit set) with one stream of the state through HBM, as MFMA products with
FP32 accumulation.  The oracle applies the same gates one by one in FP64.

Two arithmetics:
- "f16" (the default): each tile is scaled by a power of two when staged;
  matrix and amplitudes in two fp16 parts each (22 bits), three products.
  About 2^-22 relative per gate: bar 1e-5 relative L2 after up to 32
  gates (the bar every complex64 circuit result meets against the FP32
  unfused run), fidelity above 1 - 1e-10.
- "bf16" (QSIM_HIP_BLOCK_BF16): matrix in three bf16 parts, amplitudes in
  two (2^-17 relative per gate): bar 1e-4, fidelity above 1 - 1e-8.
"""

import numpy as np
import pytest

import oracle
import qsim_amd
from helpers import random_matrix, random_state, rel_l2, to_interleaved

pytestmark = pytest.mark.gpu

# arithmetic -> (relative L2 bar, 1 - fidelity bar) after up to 32 gates
BARS = {"f16": (1e-5, 1e-10), "bf16": (1e-4, 1e-8)}


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


def run_case(sim, n, tile, gates, rng, arith="f16", s=None):
    if s is None:
        s = random_state(n, rng, np.float64)
    st = upload(s.astype(np.float32), n)
    sim.apply_gate_block(tile, gates, st, arith=arith)
    for qs, m in gates:
        oracle.apply_gate(n, s, qs, m)
    got = st.to_numpy()
    err = rel_l2(got, s)
    a = got[0::2].astype(np.float64) + 1j * got[1::2]
    b = s[0::2] + 1j * s[1::2]
    fid = abs(np.vdot(b, a)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)
    return err, fid


@pytest.mark.parametrize("arith", ["f16", "bf16"])
@pytest.mark.parametrize("n,lc,tb", [(9, 9, 9), (12, 5, 9), (13, 13, 13),
                                     (14, 3, 13), (16, 5, 13), (16, 7, 12),
                                     (18, 6, 11), (20, 5, 13), (22, 8, 13),
                                     (14, 14, 14), (17, 5, 14), (20, 9, 14),
                                     (22, 3, 14), (22, 4, 14), (21, 4, 13)])
def test_block_random_programs(sim, n, lc, tb, arith):
    rng = np.random.default_rng(1000 * n + 10 * lc + tb)
    bar, fbar = BARS[arith]
    for trial in range(3):
        tile = random_tile(n, lc, tb, rng)
        nops = int(rng.integers(1, 33))
        gates = random_program(tile, nops, rng)
        err, fid = run_case(sim, n, tile, gates, rng, arith)
        print(f"{arith} n={n} tile={tile} ops={nops} relL2={err:.2e} "
              f"1-F={1 - fid:.2e}", flush=True)
        assert err < bar, (n, tile, nops, err)
        assert fid > 1 - fbar


@pytest.mark.parametrize("arith", ["f16", "bf16"])
def test_block_max_program_and_single_gates(sim, arith):
    rng = np.random.default_rng(5)
    bar, fbar = BARS[arith]
    n = 17
    tile = list(range(5)) + [6, 9, 11, 13, 14, 15, 16, 5][:8]
    tile = sorted(tile)
    gates = random_program(tile, 32, rng)
    err, fid = run_case(sim, n, tile, gates, rng, arith)
    assert err < bar and fid > 1 - fbar
    # One-qubit gates on every tile qubit, including the highest and lowest.
    gates = [([q], to_interleaved(random_matrix(1, rng), np.float64))
             for q in tile]
    err, fid = run_case(sim, n, tile, gates, rng, arith)
    assert err < bar / 2 and fid > 1 - fbar


@pytest.mark.parametrize("arith", ["f16", "bf16"])
def test_block_matches_gate_passes_on_grid_layers(sim, arith):
    # The same gates through plain gate passes (FP32 vector kernels) and
    # through one blocked pass agree to the arithmetic's bar.
    rng = np.random.default_rng(9)
    n = 20
    tile = list(range(5)) + [7, 8, 12, 13, 17, 18, 19, 16]
    tile = sorted(tile)
    gates = random_program(tile, 24, rng, kmax=2)
    s = random_state(n, rng, np.float32)
    a = upload(s, n)
    b = upload(s, n)
    sim.apply_gate_block(tile, gates, a, arith=arith)
    for qs, m in gates:
        sim.apply_gate(qs, m, b)
    assert rel_l2(a.to_numpy(), b.to_numpy()) < BARS[arith][0]


@pytest.mark.parametrize("scale", [1e-30, 1e-12, 1.0, 3e4])
def test_f16_scaling_over_magnitudes(sim, scale):
    # The per-tile power-of-two scale makes the FP16 products independent
    # of the state's magnitude: unnormalised states far outside fp16's
    # range give the same relative error as a normalised one.
    rng = np.random.default_rng(31)
    n = 18
    tile = random_tile(n, 5, 14, rng)
    gates = random_program(tile, 16, rng)
    s = random_state(n, rng, np.float64) * scale
    err, fid = run_case(sim, n, tile, gates, rng, "f16", s)
    print(f"scale {scale:g}: relL2 {err:.2e}", flush=True)
    assert err < 1e-5 and fid > 1 - 1e-10


def test_f16_tiles_of_very_different_norms(sim):
    # Tiles whose norms differ by 1e20 each keep their own relative
    # precision: the small tiles' error is not bounded by the large ones'.
    rng = np.random.default_rng(32)
    n = 16
    tile = list(range(5)) + [5, 6, 7, 8, 9, 10, 11, 12, 13]
    gates = random_program(tile, 12, rng)
    s = random_state(n, rng, np.float64)
    z = s[0::2] + 1j * s[1::2]
    z[: 1 << 14] *= 1e-20   # the first of the four tiles
    s = to_interleaved(z, np.float64)
    small = z[: 1 << 14].copy()
    st = upload(s.astype(np.float32), n)
    sim.apply_gate_block(tile, gates, st)
    for qs, m in gates:
        oracle.apply_gate(n, s, qs, m)
    got = st.to_numpy()
    assert rel_l2(got, s) < 1e-5
    g = got[0::2][: 1 << 14].astype(np.float64) + 1j * got[1::2][: 1 << 14]
    ref = (s[0::2] + 1j * s[1::2])[: 1 << 14]
    assert np.linalg.norm(g - ref) / np.linalg.norm(ref) < 1e-5
    assert np.linalg.norm(small) > 0


def test_f16_growing_and_contracting_matrices(sim):
    # Non-unitary programs: the host bounds the growth of the state from
    # the matrices' norms and stages tiles lower (headroom) so that no value
    # leaves fp16's range; contractions need none.
    rng = np.random.default_rng(33)
    n = 16
    tile = list(range(5)) + [6, 8, 9, 10, 12, 13, 14, 15]
    gates = []
    for i in range(10):
        k = 1 + i % 3
        qs = sorted(int(q) for q in rng.choice(tile, size=k, replace=False))
        u = random_matrix(k, rng) * (3.0 if i % 2 == 0 else 0.5)
        gates.append((qs, to_interleaved(u, np.float64)))
    s = random_state(n, rng, np.float64)
    err, fid = run_case(sim, n, tile, gates, rng, "f16", s)
    assert err < 1e-5, err
    # Growth beyond 2^10 (ten gates scaled by 3): the pass runs on the
    # split-BF16 product instead, at that arithmetic's bar.
    grow = [(qs, m * 3.0) for qs, m in random_program(tile, 10, rng)]
    err, fid = run_case(sim, n, tile, grow, rng, "f16")
    assert err < BARS["bf16"][0], err
    # A concentrating program: a state spread evenly over the tile (with
    # re = im) is refocused by Hadamards on every tile qubit into one
    # amplitude that carries the whole tile's norm, the largest value the
    # scaled tile can hold.
    h = to_interleaved(np.array([[1, 1], [1, -1]]) / np.sqrt(2), np.float64)
    z = np.zeros(1 << n, complex)
    z[0] = (1 + 1j) / np.sqrt(2)
    s = to_interleaved(z, np.float64)
    for q in tile:
        oracle.apply_gate(n, s, [q], h)
    gates = [([q], h) for q in tile]
    err, fid = run_case(sim, n, tile, gates, rng, "f16", s)
    assert err < 1e-5, err


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
    # the two-part bf16 amplitude split rounds them all alike and the norm
    # moves coherently; the three-part split (QSIM_HIP_BLOCK_ACCURATE)
    # keeps it within FP32 rounding, and so does the FP16 default (22-bit
    # parts).  All stay within their bars.
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
    for mode in ("bf16", "bf16-accurate", "f16"):
        st = upload(s0.astype(np.float32), n)
        sim.apply_gate_block(tile, gates[:32], st,
                             accurate=mode == "bf16-accurate",
                             arith=mode[:4].rstrip("-"))
        got = st.to_numpy().astype(np.float64)
        drift[mode] = abs(np.sum(got * got) - 1.0)
        assert rel_l2(got, ref) < (1e-5 if mode != "bf16" else 1e-4)
    print(f"norm drift: bf16 two-part {drift['bf16']:.2e}, three-part "
          f"{drift['bf16-accurate']:.2e}, f16 {drift['f16']:.2e}", flush=True)
    assert drift["bf16-accurate"] < 1e-6
    assert drift["f16"] < 1e-6


def test_accurate_prepared_pass_matches_oracle(sim):
    rng = np.random.default_rng(22)
    n = 18
    tile = sorted(list(range(5)) + [6, 9, 10, 13, 15, 16, 17, 12])
    gates = random_program(tile, 24, rng)
    s = random_state(n, rng, np.float64)
    st = upload(s.astype(np.float32), n)
    pb = sim.prepare_gate_block(n, tile, gates, accurate=True, arith="bf16")
    pb.run(st)
    pb.free()
    for qs, m in gates:
        oracle.apply_gate(n, s, qs, m)
    # Three-part amplitudes: FP32-level agreement, well inside the
    # two-part bar of 1e-4.
    assert rel_l2(st.to_numpy(), s) < 2e-6
