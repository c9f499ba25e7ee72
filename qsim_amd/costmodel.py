"""The learned planner's pass-time cost model.

A small MLP (features -> 64 -> 64 -> 64 -> 1, include/qsim_hip.h
qsim_hip_costmodel_*) predicts the log time of one pass -- a fused gate or a
cache-blocked pass -- from the features the circuit driver computes for it
(include/qsim_amd/learned_planner.h).  Training data are the backend's own
per-pass kernel timings, recorded with HIP events while seeded random grid
circuits run under every candidate partition; training runs on the GPU with
the library's forward/backward/Adam kernels.  The planner
(qsim_driver_plan_learned) sums the predicted times of each candidate
partition and keeps the fastest; qsim's fuser at max_fused_size 4 is the
fallback.

The stored model consumes raw features and returns the raw target
(log ms - (n - 24) ln 2): the input standardisation and the target scaling
used in training are folded into the first and last layers.
"""

from __future__ import annotations

import ctypes
import json
import math
import os
import time

import numpy as np

from . import F32, PassTimer, StateSpace, _check, lib, synchronize

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "data", "cost_model.json")

NUM_FEATURES = 19
HIDDEN = 64


def target_offset(n: int) -> float:
    return (n - 24) * math.log(2.0)


def layer_shapes(num_features=NUM_FEATURES, hidden=HIDDEN):
    """[(in, out), ...] of the four layers; parameters are W then b each."""
    return [(num_features, hidden), (hidden, hidden), (hidden, hidden),
            (hidden, 1)]


def split_params(p, num_features=NUM_FEATURES, hidden=HIDDEN):
    out, off = [], 0
    for i, o in layer_shapes(num_features, hidden):
        w = p[off:off + i * o].reshape(i, o)
        off += i * o
        b = p[off:off + o]
        off += o
        out.append((w, b))
    return out


def join_params(layers):
    return np.concatenate([np.concatenate([w.reshape(-1), b]) for w, b in
                           layers]).astype(np.float32)


def fold_normalisation(p, mu_x, sd_x, mu_y, sd_y,
                       num_features=NUM_FEATURES, hidden=HIDDEN):
    """Parameters trained on ((x - mu_x) / sd_x, (y - mu_y) / sd_y) turned
    into parameters on raw x and y."""
    layers = [(w.astype(np.float64).copy(), b.astype(np.float64).copy())
              for w, b in split_params(np.asarray(p), num_features, hidden)]
    w0, b0 = layers[0]
    layers[0] = (w0 / sd_x[:, None], b0 - (mu_x / sd_x) @ w0)
    wl, bl = layers[-1]
    layers[-1] = (wl * sd_y, bl * sd_y + mu_y)
    return join_params(layers)


class CostModel:
    """A cost-model network on the GPU (the library's own kernels)."""

    def __init__(self, num_features=NUM_FEATURES, hidden=HIDDEN):
        self.num_features = num_features
        self.hidden = hidden
        h = ctypes.c_void_p()
        _check(lib().qsim_hip_costmodel_create(num_features, hidden,
                                               ctypes.byref(h)))
        self.h = h
        self.num_params = int(lib().qsim_hip_costmodel_num_params(h))
        self.meta = {}

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(self.num_params, np.float32)
        _check(lib().qsim_hip_costmodel_get_params(self.h, p.ctypes.data,
                                                   self.num_params))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().qsim_hip_costmodel_set_params(self.h, p.ctypes.data,
                                                   len(p)))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(
            -1, self.num_features)
        y = np.zeros(len(x), np.float32)
        _check(lib().qsim_hip_costmodel_forward(self.h, x.ctypes.data, len(x),
                                                y.ctypes.data))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(
            -1, self.num_features)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.num_params, np.float32)
        loss = ctypes.c_double(0)
        _check(lib().qsim_hip_costmodel_gradient(
            self.h, x.ctypes.data, t.ctypes.data, len(x), ctypes.byref(loss),
            g.ctypes.data))
        return loss.value, g

    def train(self, x, t, epochs=100, batch=256, lr=1e-3, seed=0) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(
            -1, self.num_features)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = ctypes.c_double(0)
        _check(lib().qsim_hip_costmodel_train(
            self.h, x.ctypes.data, t.ctypes.data, len(x), epochs, batch,
            ctypes.c_float(lr), seed, ctypes.byref(loss)))
        return loss.value

    def save(self, path=MODEL_PATH):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump({"num_features": self.num_features,
                       "hidden": self.hidden,
                       "params": [float(v) for v in self.params],
                       "meta": self.meta}, f)

    @classmethod
    def load(cls, path=MODEL_PATH) -> "CostModel":
        with open(path) as f:
            d = json.load(f)
        m = cls(d["num_features"], d["hidden"])
        m.params = np.asarray(d["params"], np.float32)
        m.meta = d.get("meta", {})
        return m

    def predict_ms(self, x, n) -> np.ndarray:
        return np.exp(self.forward(x).astype(np.float64) + target_offset(n))

    def free(self):
        if getattr(self, "h", None):
            lib().qsim_hip_costmodel_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def candidates():
    """The planner's candidate partitions (qsim::hip::DefaultCandidates)."""
    from . import circuits, driver
    c = driver.Circuit(circuits.grid_circuit(12, 2, 0))
    _, cands = c.plan_learned(None)
    return cands


def collect(sizes, seeds, cycles=(20,), cands=None, log=None):
    """Per-pass (features, target) pairs: every candidate partition of seeded
    grid circuits, timed pass by pass with HIP events (one warm-up run, one
    timed run).  Returns X (N, 19), y (N,), and a record per run."""
    from . import circuits, driver
    cands = cands or candidates()
    xs, ys, runs = [], [], []
    ss = StateSpace(np.float32)
    for n in sizes:
        st = ss.create(n)
        if st is None:
            raise MemoryError(f"cannot allocate a {n}-qubit state")
        try:
            for cyc in cycles:
                for seed in seeds:
                    circ = driver.Circuit(circuits.grid_circuit(n, cyc, seed))
                    for cand in cands:
                        try:
                            npass = circ.apply_candidate(cand)
                        except Exception:
                            continue  # e.g. blocked passes below 9 qubits
                        feats = circ.pass_features(cand["method"])
                        ss.set_state_zero(st)
                        circ.run(st)
                        synchronize()
                        ss.set_state_zero(st)
                        synchronize()
                        PassTimer.start(npass + 16)
                        t0 = time.time()
                        circ.run(st)
                        synchronize()
                        wall = time.time() - t0
                        ms = PassTimer.stop(npass + 16)
                        if len(ms) != npass:
                            continue
                        xs.append(feats)
                        ys.append(np.log(np.maximum(ms, 1e-4)) -
                                  target_offset(n))
                        runs.append({"n": n, "cycles": cyc, "seed": seed,
                                     "cand": cand, "passes": npass,
                                     "kernel_ms": float(np.sum(ms)),
                                     "wall_s": wall})
                        if log:
                            log(f"collect n={n} cyc={cyc} seed={seed} "
                                f"{cand} passes={npass} "
                                f"sum={np.sum(ms):.2f} ms")
        finally:
            st.free()
    return (np.concatenate(xs).astype(np.float32),
            np.concatenate(ys).astype(np.float32), runs)


def train(x, y, epochs=200, batch=256, lr=1e-3, seed=0):
    """Trains a model on standardised data and returns it with the
    normalisation folded in (raw features -> raw target)."""
    mu_x = x.mean(axis=0).astype(np.float64)
    sd_x = x.std(axis=0).astype(np.float64)
    sd_x[sd_x < 1e-6] = 1.0
    mu_y = float(y.mean())
    sd_y = float(y.std()) or 1.0
    xn = ((x - mu_x) / sd_x).astype(np.float32)
    yn = ((y - mu_y) / sd_y).astype(np.float32)
    m = CostModel(x.shape[1], HIDDEN)
    loss = m.train(xn, yn, epochs=epochs, batch=batch, lr=lr, seed=seed)
    m.params = fold_normalisation(m.params, mu_x, sd_x, mu_y, sd_y,
                                  x.shape[1], HIDDEN)
    m.meta = {"train_loss_standardised": loss, "samples": int(len(x)),
              "epochs": epochs, "batch": batch, "lr": lr, "seed": seed,
              "rmse_log": float(np.sqrt(loss) * sd_y)}
    return m
