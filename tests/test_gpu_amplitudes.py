"""Batched amplitude gather (qsim_hip_get_ampls) and the qsim_amplitudes app
on the gfx950 backend, against the state itself, the oracle and qsim's own
apps/qsim_amplitudes.cc built from the reference sources (oracle/_ref)."""

import os
import subprocess

import numpy as np
import pytest

import oracle
import qsim_amd
from qsim_amd import circuits

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "qsim_amd", "bin")
REF_APP = os.path.join(ROOT, "oracle", "_ref", "qsim_amplitudes")


def random_state(ss, n, dtype, seed):
    rng = np.random.default_rng(seed)
    st = ss.create(n)
    ct = np.complex64 if dtype == np.float32 else np.complex128
    v = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(ct)
    st.from_numpy(v.view(dtype))
    return st, v


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gather_is_the_state(dtype):
    n = 14
    print(f"gather n={n} dtype={np.dtype(dtype).name}", flush=True)
    ss = qsim_amd.StateSpace(dtype)
    st, v = random_state(ss, n, dtype, 3)
    rng = np.random.default_rng(4)
    idx = rng.integers(0, 1 << n, size=5000, dtype=np.uint64)
    idx[:3] = [0, (1 << n) - 1, 0]           # ends and a repeat
    got = ss.get_ampls(st, idx)
    assert got.dtype == v.dtype
    assert np.array_equal(got, v[idx.astype(np.int64)])   # a copy: bit-exact
    for i in idx[:16]:
        assert ss.get_ampl(st, int(i)) == complex(v[int(i)])


def test_gather_in_chunks():
    # More indices than one scratch chunk (2^20), bit-exact.
    n = 22
    print(f"gather chunks n={n}", flush=True)
    ss = qsim_amd.StateSpace(np.float32)
    st, v = random_state(ss, n, np.float32, 5)
    rng = np.random.default_rng(6)
    idx = rng.integers(0, 1 << n, size=(1 << 21) + 17, dtype=np.uint64)
    got = ss.get_ampls(st, idx)
    assert np.array_equal(got, v[idx.astype(np.int64)])


def test_gather_arguments():
    ss = qsim_amd.StateSpace(np.float32)
    st, v = random_state(ss, 10, np.float32, 7)
    assert ss.get_ampls(st, []).size == 0
    with pytest.raises(qsim_amd.QsimHipError) as e:
        ss.get_ampls(st, [1, 1 << 10])
    assert e.value.code == qsim_amd.ERR_ARG


def _bitstring_lines(idx, n):
    # qsim's text format: character p is bit p (lib/bitstring.h).
    return ["".join("1" if (int(i) >> p) & 1 else "0" for p in range(n))
            for i in idx]


def _read_amplitudes(path):
    lines = open(path).read().splitlines()
    vals = np.array([[float(x) for x in ln.split()] for ln in lines])
    return lines, vals[:, 0] + 1j * vals[:, 1]


def test_amplitudes_app_matches_reference_app(tmp_path):
    exe = os.path.join(BIN, "qsim_amplitudes_hip")
    if not os.path.exists(exe):
        pytest.skip("app not built (needs the qsim headers)")
    if not os.path.exists(REF_APP):
        pytest.skip("reference qsim_amplitudes not built")
    n, cycles = 20, 14
    text = circuits.grid_circuit(n, cycles, seed=0)
    cf = tmp_path / "c.txt"
    cf.write_text(text)
    rng = np.random.default_rng(8)
    times = [9, 2 * cycles - 1]
    ins, outs, refs, idxs = [], [], [], []
    for k, t in enumerate(times):
        idx = rng.integers(0, 1 << n, size=300, dtype=np.uint64)
        f = tmp_path / f"in{k}.txt"
        f.write_text("\n".join(_bitstring_lines(idx, n)) + "\n")
        ins.append(str(f))
        outs.append(str(tmp_path / f"out{k}.txt"))
        refs.append(str(tmp_path / f"ref{k}.txt"))
        idxs.append(idx)
    d = ",".join(str(t) for t in times)
    print(f"qsim_amplitudes n={n} times={d}", flush=True)
    r = subprocess.run([exe, "-c", str(cf), "-d", d, "-i", ",".join(ins),
                        "-o", ",".join(outs), "-f", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "all done" in r.stdout
    rr = subprocess.run([REF_APP, "-c", str(cf), "-d", d, "-i", ",".join(ins),
                         "-o", ",".join(refs), "-f", "2", "-t", "8"],
                        capture_output=True, text=True, timeout=300)
    assert rr.returncode == 0, rr.stderr
    for k in range(len(times)):
        got_lines, got = _read_amplitudes(outs[k])
        ref_lines, ref = _read_amplitudes(refs[k])
        assert len(got_lines) == len(ref_lines) == len(idxs[k])
        # Same text layout: two right-aligned 17-column fields per line.
        assert all(len(a) == len(b) == 34 for a, b in zip(got_lines, ref_lines))
        # Both are FP32 runs of the same fused circuit.
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err < 1e-5, err
    # The final time against the oracle's FP64 state.
    full, _ = oracle.Circuit(text, np.float64).run()
    want = full.view(np.complex128)[idxs[-1].astype(np.int64)]
    _, got = _read_amplitudes(outs[-1])
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5


def test_amplitudes_app_rejects_bad_options(tmp_path):
    exe = os.path.join(BIN, "qsim_amplitudes_hip")
    if not os.path.exists(exe):
        pytest.skip("app not built (needs the qsim headers)")
    cf = tmp_path / "c.txt"
    cf.write_text(circuits.grid_circuit(12, 2, seed=0))
    # Times must be sorted and match the file lists (as the reference app).
    r = subprocess.run([exe, "-c", str(cf), "-d", "3,1", "-i", "a,b",
                        "-o", "c,d"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "not sorted" in r.stderr
    r = subprocess.run([exe, "-c", str(cf), "-d", "1,3", "-i", "a",
                        "-o", "c,d"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
