"""The learned planner's cost model on the GPU: forward pass and gradients
against the FP64 CPU oracle (oracle/mlp_oracle.cc), training, and what the
trained network gains over the fallback partition.

Bars: forward within 1e-5 relative of the FP64 oracle, gradients within
1e-4 relative (FP32 kernels against FP64); training reduces the loss and is
deterministic for a seed; the planner's pick runs at least as fast as the
fallback (qsim's fuser at max_fused_size 4) and its state keeps
F >= 1 - 1e-6 against the FP64 run.
"""

import os
import time

import numpy as np
import pytest

import oracle
import qsim_amd
from qsim_amd import circuits, costmodel, driver

pytestmark = pytest.mark.gpu

F, H = costmodel.NUM_FEATURES, costmodel.HIDDEN


def random_model(rng, scale=0.3):
    m = costmodel.CostModel(F, H)
    m.params = rng.normal(scale=scale, size=m.num_params).astype(np.float32)
    return m


def test_forward_matches_oracle():
    rng = np.random.default_rng(1)
    m = random_model(rng)
    x = rng.normal(size=(300, F)).astype(np.float32)
    got = m.forward(x)
    want = oracle.mlp_forward(F, H, m.params, x)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"forward max rel err {err:.2e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("batch", [1, 7, 256, 1000])
def test_gradient_matches_oracle(batch):
    rng = np.random.default_rng(batch)
    m = random_model(rng)
    x = rng.normal(size=(batch, F)).astype(np.float32)
    t = rng.normal(size=batch).astype(np.float32)
    loss, g = m.gradient(x, t)
    loss_o, g_o = oracle.mlp_gradient(F, H, m.params, x, t)
    rel = np.linalg.norm(g - g_o) / np.linalg.norm(g_o)
    print(f"batch {batch}: loss {loss:.6e} vs {loss_o:.6e}, grad rel err "
          f"{rel:.2e}", flush=True)
    assert abs(loss - loss_o) <= 1e-5 * abs(loss_o)
    assert rel < 1e-4
    assert np.abs(g - g_o).max() <= 1e-4 * np.abs(g_o).max()


def _synthetic(rng, n):
    x = rng.normal(size=(n, F)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] ** 2 - x[:, 2] * x[:, 3]).astype(
        np.float32)
    return x, t


def test_training_reduces_loss_and_is_deterministic():
    rng = np.random.default_rng(2)
    x, t = _synthetic(rng, 4096)
    a = costmodel.CostModel(F, H)
    b = costmodel.CostModel(F, H)
    init = a.params
    b.params = init
    l0, _ = a.gradient(x, t)
    la = a.train(x, t, epochs=60, batch=256, lr=2e-3, seed=9)
    lb = b.train(x, t, epochs=60, batch=256, lr=2e-3, seed=9)
    print(f"training loss {l0:.4f} -> {la:.4f}", flush=True)
    assert la < 0.1 * l0
    assert la == lb and np.array_equal(a.params, b.params)


def test_folded_normalisation_matches_standardised_model():
    rng = np.random.default_rng(4)
    x = (rng.normal(size=(500, F)) * 3 + 7).astype(np.float32)
    mu, sd = x.mean(0).astype(np.float64), x.std(0).astype(np.float64)
    m = random_model(rng)
    y_std = m.forward(((x - mu) / sd).astype(np.float32)) * 2.5 - 1.25
    folded = costmodel.CostModel(F, H)
    folded.params = costmodel.fold_normalisation(m.params, mu, sd, -1.25, 2.5)
    assert np.allclose(folded.forward(x), y_std, rtol=1e-4, atol=1e-4)


def _run_ms(circ, ss, st, reps=2):
    ss.set_state_zero(st)
    circ.run(st)
    qsim_amd.synchronize()
    best = None
    for _ in range(reps):
        ss.set_state_zero(st)
        qsim_amd.synchronize()
        t0 = time.time()
        circ.run(st)
        qsim_amd.synchronize()
        dt = time.time() - t0
        best = dt if best is None else min(best, dt)
    return best * 1e3


def test_small_model_trained_on_fresh_timings_plans_well():
    # End to end at small sizes: record timings, train on the GPU, plan a
    # held-out seed.  The pick is never slower than the fallback by more than
    # timing noise, and its state is correct.
    x, y, runs = costmodel.collect(sizes=(20, 22), seeds=(0, 1), cycles=(12,))
    print(f"collected {len(y)} pass timings from {len(runs)} runs", flush=True)
    m = costmodel.train(x, y, epochs=40, batch=256, lr=2e-3, seed=0)
    print(f"rmse(log) {m.meta['rmse_log']:.3f}", flush=True)
    n = 22
    text = circuits.grid_circuit(n, 12, seed=77)
    c = driver.Circuit(text)
    chosen, cands = c.plan_learned(m)
    ss = qsim_amd.StateSpace(np.float32)
    st = ss.create(n)
    t_pick = _run_ms(c, ss, st)
    ss.set_state_zero(st)
    c.run(st)
    got = st.to_numpy()
    c.plan_learned(None)  # fallback
    t_fb = _run_ms(c, ss, st)
    st.free()
    print(f"pick {cands[chosen]}: {t_pick:.2f} ms, fallback {t_fb:.2f} ms",
          flush=True)
    assert t_pick <= 1.25 * t_fb
    # The pick's state against the unfused run.
    cu = driver.Circuit(text)
    cu.fuse(1, driver.FUSER_NONE)
    su = ss.create(n)
    ss.set_state_zero(su)
    cu.run(su)
    unf = su.to_numpy()
    su.free()
    assert np.linalg.norm(got - unf) / np.linalg.norm(unf) < 1e-5
    s64, _ = oracle.Circuit(text, np.float64).run()
    a = got[0::2] + 1j * got[1::2].astype(np.float64)
    b = s64[0::2] + 1j * s64[1::2]
    fid = abs(np.vdot(b, a)) ** 2 / np.vdot(a, a).real
    assert fid > 1 - 1e-6


@pytest.mark.skipif(not os.path.exists(costmodel.MODEL_PATH),
                    reason="no trained cost model committed")
@pytest.mark.parametrize("n,seed", [(28, 101), (30, 102)])
def test_committed_model_gains_over_fallback(n, seed):
    # The committed model (trained on grid circuits at 24-32 qubits, seeds
    # 0-2) plans held-out seeds; measured wall time decides.
    m = costmodel.CostModel.load()
    text = circuits.grid_circuit(n, 20, seed=seed)
    c = driver.Circuit(text)
    chosen, cands = c.plan_learned(m)
    ss = qsim_amd.StateSpace(np.float32)
    st = ss.create(n)
    su = ss.create(n)
    try:
        t_pick = _run_ms(c, ss, st)
        pred = cands[chosen]["predicted_ms"]
        ss.set_state_zero(st)
        c.run(st)
        # The pick's state against the unfused run (every gate its own FP32
        # pass), on the device: ||pick - unfused|| / ||unfused||.
        cu = driver.Circuit(text)
        cu.fuse(1, driver.FUSER_NONE)
        ss.set_state_zero(su)
        cu.run(su)
        nu = ss.norm(su)
        ss.multiply(-1.0, su)
        ss.add(su, st)
        err_u = float(np.sqrt(ss.norm(st) / nu))
        c.plan_learned(None)
        t_fb = _run_ms(c, ss, st)
        # The best fixed choice among the blocked shapes the planner picks
        # from most often, each planned for this circuit.
        fixed = {}
        for tb, lb in ((14, 4), (14, 5), (13, 5)):
            c.fuse_blocked(tb, lb)
            fixed[f"blocked {tb}/{lb}"] = _run_ms(c, ss, st)
    finally:
        st.free()
        su.free()
    best_fixed = min(fixed, key=fixed.get)
    print(f"{n}q seed {seed}: pick {cands[chosen]} measured {t_pick:.1f} ms "
          f"(predicted {pred:.1f}), fallback {t_fb:.1f} ms, gain "
          f"{t_fb / t_pick:.2f}x; best fixed {best_fixed} "
          f"{fixed[best_fixed]:.1f} ms ({fixed[best_fixed] / t_pick:.3f}x); "
          f"rel-L2 vs unfused {err_u:.2e}", flush=True)
    assert t_pick < t_fb
    assert abs(pred - t_pick) / t_pick < 0.5
    # Never worse than the best fixed blocked shape beyond timing noise.
    assert t_pick <= 1.05 * fixed[best_fixed]
    assert err_u < 1e-5
