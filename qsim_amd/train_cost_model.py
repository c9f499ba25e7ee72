"""Records pass timings on the GPU, trains the planner's cost model on the
GPU, and writes qsim_amd/data/cost_model.json.

    python -m qsim_amd.train_cost_model [--sizes 24,26,28,30,32]
        [--seeds 0,1,2] [--cycles 14,20] [--epochs 200]
        [--holdout-sizes 28,30] [--holdout-seeds 100] [--out PATH]
        [--save-data PATH.npz]

Every candidate partition of the planner (qsim's fuser at max_fused_size
2..6 with each gate method, and blocked passes of several tile shapes) runs
on seeded random grid circuits; each pass's kernel time (HIP events) and its
features form one sample.  Held-out seeds then measure the model: the error
of its per-pass predictions, of its per-partition totals against the
measured kernel time, and the time of the partition it picks against the
fallback (max_fused_size 4).
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def _ints(s):
    return [int(v) for v in s.split(",") if v]


def log(*a):
    print(*a, flush=True)


def evaluate(model, x, y, runs):
    """Per-pass and per-partition prediction errors on a held-out set."""
    pred = model.forward(x)
    rmse = float(np.sqrt(np.mean((pred - y) ** 2)))
    out = {"passes": int(len(y)), "rmse_log": rmse,
           "median_abs_rel_pass": float(np.median(np.abs(np.exp(pred - y) - 1)))}
    # Per-run totals (runs are stored in collection order, passes contiguous).
    off, rel = 0, []
    for r in runs:
        k = r["passes"]
        from .costmodel import target_offset
        p_ms = float(np.sum(np.exp(pred[off:off + k] + target_offset(r["n"]))))
        rel.append(abs(p_ms - r["kernel_ms"]) / r["kernel_ms"])
        off += k
    out["median_abs_rel_partition"] = float(np.median(rel))
    out["max_abs_rel_partition"] = float(np.max(rel))
    return out


def planner_gain(model, sizes, seeds, cycles):
    """Measured time of the planner's pick against the fallback."""
    import qsim_amd
    from . import circuits, driver
    res = []
    ss = qsim_amd.StateSpace(np.float32)
    for n in sizes:
        st = ss.create(n)
        try:
            for seed in seeds:
                c = driver.Circuit(circuits.grid_circuit(n, cycles, seed))
                times = {}
                for name in ("learned", "fallback"):
                    chosen, cands = c.plan_learned(
                        model if name == "learned" else None)
                    best = None
                    for rep in range(3):
                        ss.set_state_zero(st)
                        qsim_amd.synchronize()
                        t0 = time.time()
                        c.run(st)
                        qsim_amd.synchronize()
                        dt = time.time() - t0
                        if rep > 0:
                            best = dt if best is None else min(best, dt)
                    times[name] = (best, cands[chosen])
                r = {"n": n, "seed": seed, "cycles": cycles,
                     "learned_s": times["learned"][0],
                     "learned_pick": times["learned"][1],
                     "fallback_s": times["fallback"][0],
                     "gain": times["fallback"][0] / times["learned"][0]}
                log(f"planner n={n} seed={seed}: {r['learned_pick']} "
                    f"{r['learned_s']:.3f}s vs fallback {r['fallback_s']:.3f}s "
                    f"gain {r['gain']:.2f}x")
                res.append(r)
        finally:
            st.free()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default="24,25,26,27,28,29,30,32")
    ap.add_argument("--seeds", default="0,1,2")
    ap.add_argument("--cycles", default="14,20")
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--holdout-sizes", default="28,30")
    ap.add_argument("--holdout-seeds", default="100")
    ap.add_argument("--gain-sizes", default="28,30,32,33")
    ap.add_argument("--gain-seeds", default="101")
    ap.add_argument("--out", default=None)
    ap.add_argument("--save-data", default=None)
    args = ap.parse_args(argv)

    from . import costmodel
    t0 = time.time()
    x, y, runs = costmodel.collect(_ints(args.sizes), _ints(args.seeds),
                                   _ints(args.cycles), log=log)
    log(f"collected {len(y)} pass timings in {len(runs)} runs, "
        f"{time.time() - t0:.0f}s")
    if args.save_data:
        np.savez_compressed(args.save_data, x=x, y=y)
    t1 = time.time()
    m = costmodel.train(x, y, epochs=args.epochs, batch=args.batch,
                        lr=args.lr, seed=0)
    log(f"trained on the GPU in {time.time() - t1:.1f}s: {m.meta}")
    hx, hy, hruns = costmodel.collect(_ints(args.holdout_sizes),
                                      _ints(args.holdout_seeds), (20,))
    m.meta["holdout"] = evaluate(m, hx, hy, hruns)
    m.meta["holdout"]["sizes"] = _ints(args.holdout_sizes)
    m.meta["holdout"]["seeds"] = _ints(args.holdout_seeds)
    log(f"held-out: {m.meta['holdout']}")
    m.meta["train_sizes"] = _ints(args.sizes)
    m.meta["train_seeds"] = _ints(args.seeds)
    m.meta["train_cycles"] = _ints(args.cycles)
    m.meta["planner_gain"] = planner_gain(m, _ints(args.gain_sizes),
                                          _ints(args.gain_seeds), 20)
    m.save(args.out or costmodel.MODEL_PATH)
    log(json.dumps({"samples": len(y), "holdout": m.meta["holdout"],
                    "planner_gain": m.meta["planner_gain"]}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
