"""Benchmark: a 32-qubit random grid circuit on one MI355X.

One step = one full circuit (|0..0> preparation plus every fused pass) on a
complex64 state resident in HBM.  The circuit is the 4x8 grid, 20 cycles,
seed 0 construction of qsim_amd/circuits.py, fused by qsim's
MultiQubitGateFuser (max_fused_size 4) and applied by qsim's own apply loop
over the gfx950 kernels.

Prints one JSON line: seconds per circuit (value), gate passes/s, the
roofline of the dominant gate-pass kernel from HIP events recorded around
every pass inside the timed region, and qsim's AVX/OpenMP CPU path timed on a
bounded sample of the same circuit on this host.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--qubits 32]
                    [--cycles 20] [--max-fused 4] [--cpu-passes P | --no-cpu]
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = ("32q circuit time + HBM GB/s, learned/MFMA vs plain FP32 fusion "
          "at F>=1-1e-6")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_family(qubits):
    """Which gate kernel a pass runs: the register path when every target is
    at qubit >= 5, the LDS tile path otherwise (qsim_amd/csrc/gate_kernels.hip)."""
    return "GateRegKernel" if (not qubits or min(qubits) >= 5) else "GateTileKernel"


def run_gpu(args, text):
    import qsim_amd
    from qsim_amd import driver

    name, mem, cus = qsim_amd.device_info()
    log(f"device: {name}, {mem / 2**30:.0f} GiB, {cus} CUs")
    circ = driver.Circuit(text)
    t0 = time.time()
    npass = circ.fuse(args.max_fused, driver.FUSER_MQUBIT)
    fuse_s = time.time() - t0
    passes = circ.passes()
    ks = [len(q) for q, _ in passes]
    log(f"circuit: {circ.num_qubits} qubits, {circ.num_gates} gates -> "
        f"{npass} fused passes (k histogram {np.bincount(ks).tolist()}), "
        f"fuse {fuse_s:.3f}s")

    ss = qsim_amd.StateSpace(np.float32)
    st = ss.create(circ.num_qubits)
    if st is None:
        raise RuntimeError("cannot allocate the state")
    for w in range(args.warmup):
        ss.set_state_zero(st)
        circ.run(st)
        qsim_amd.synchronize()
        log(f"warmup {w + 1}/{args.warmup} done")

    qsim_amd.synchronize()
    qsim_amd.PassTimer.start(npass * args.steps + 16)
    qsim_amd.Timer.record(0)
    t_wall = time.time()
    for _ in range(args.steps):
        ss.set_state_zero(st)
        circ.run(st)
    qsim_amd.Timer.record(1)
    qsim_amd.synchronize()
    wall = time.time() - t_wall
    total_ms = qsim_amd.Timer.elapsed_ms(0, 1)
    pass_ms = qsim_amd.PassTimer.stop(npass * args.steps + 16)
    log(f"timed: {args.steps} circuits in {total_ms:.1f} ms (wall {wall:.3f}s), "
        f"{len(pass_ms)} passes timed")
    norm = ss.norm(st)
    log(f"final norm {norm:.9f}")

    n = circ.num_qubits
    bytes_per_pass = 16.0 * (1 << n)   # read + write every complex64 once
    fam = [kernel_family(q) for q, _ in passes] * args.steps
    per_fam = {}
    for f, ms in zip(fam, pass_ms):
        per_fam.setdefault(f, []).append(float(ms))
    fam_stats = {}
    for f, v in per_fam.items():
        avg = float(np.mean(v))
        fam_stats[f] = {"passes": len(v), "avg_ms": avg, "total_ms": float(np.sum(v)),
                        "achieved_gbps": bytes_per_pass / (avg * 1e-3) / 1e9}
    dominant = max(fam_stats, key=lambda f: fam_stats[f]["total_ms"])
    d = fam_stats[dominant]
    all_avg = float(np.mean(pass_ms))
    per_k = {}
    for (q, _), ms in zip(passes * args.steps, pass_ms):
        per_k.setdefault(len(q), []).append(float(ms))
    k_stats = {str(k): {"passes": len(v), "avg_ms": float(np.mean(v)),
                        "gbps": bytes_per_pass / (float(np.mean(v)) * 1e-3) / 1e9}
               for k, v in sorted(per_k.items())}
    return {
        "num_qubits": n, "num_gates": circ.num_gates, "passes": npass,
        "circuit_ms": total_ms / args.steps, "wall_s": wall, "norm": norm,
        "fuse_s": fuse_s, "pass_ms_sum": float(np.sum(pass_ms)) / args.steps,
        "dominant": dominant, "families": fam_stats, "per_k": k_stats,
        "all_pass_avg_ms": all_avg,
        "roofline": {
            "bound": "hbm", "kernel": dominant,
            "achieved": d["achieved_gbps"], "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": d["achieved_gbps"] / HBM_PEAK_GBPS,
            "traffic": None,
            "bytes_per_launch": bytes_per_pass,
            "avg_launch_ms": d["avg_ms"],
        },
    }


def run_cpu_baseline(args, circuit_file, gpu_passes):
    import oracle
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    p = args.cpu_passes
    log(f"cpu baseline: qsim SimulatorAVX, {threads} threads, first {p} passes")
    t0 = time.time()
    info = oracle.run_reference(circuit_file, mode="avx", max_fused=args.max_fused,
                                threads=threads, passes=p, timeout=1800)
    wall = time.time() - t0
    per_pass = info["sim_s"] / max(1, min(p, info["fused_ops"]))
    per_circuit = per_pass * info["fused_ops"]
    log(f"cpu baseline: {info['sim_s']:.2f}s for {p} passes "
        f"({per_pass:.3f}s/pass), {info['fused_ops']} passes/circuit, "
        f"process wall {wall:.1f}s")
    return {
        "value": per_circuit, "unit": "s/circuit", "cores": threads,
        "kind": "reference",
        "sample": (f"first {p} of {info['fused_ops']} fused passes "
                   f"(MultiQubitGateFuser k={args.max_fused}) of the same "
                   f"{info['num_qubits']}q circuit on qsim's SimulatorAVX "
                   f"(oracle/_ref), {threads} OpenMP threads; seconds per "
                   f"pass x passes per circuit"),
        "sample_s": info["sim_s"], "passes_per_circuit_cpu": info["fused_ops"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=32)
    ap.add_argument("--cycles", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--max-fused", type=int, default=4)
    ap.add_argument("--cpu-passes", type=int, default=6)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not "
              "supported (the state-vector shards by its top qubits across GPUs "
              "only in design; see the sharding notes).", file=sys.stderr)
        sys.exit(2)

    from qsim_amd import circuits
    text = circuits.grid_circuit(args.qubits, args.cycles, args.seed)
    rows, cols = circuits.grid_shape(args.qubits)

    gpu = run_gpu(args, text)

    cpu = None
    if not args.no_cpu:
        with tempfile.TemporaryDirectory() as tmp:
            cf = os.path.join(tmp, "circuit.txt")
            with open(cf, "w") as f:
                f.write(text)
            try:
                cpu = run_cpu_baseline(args, cf, gpu["passes"])
            except Exception as e:  # reported, never silently replaced
                log(f"cpu baseline failed: {e!r}")
                cpu = {"value": None, "error": repr(e)}

    circuit_s = gpu["circuit_ms"] / 1e3
    out = {
        "metric": METRIC,
        "value": circuit_s,
        "unit": "s/circuit",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": gpu["circuit_ms"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.qubits}q random grid circuit {rows}x{cols}, "
                         f"{args.cycles} cycles, seed {args.seed}, complex64, "
                         f"MultiQubitGateFuser max_fused_size={args.max_fused}, "
                         "plain FP32 vector kernels"),
            "num_qubits": args.qubits, "cycles": args.cycles,
            "max_fused_size": args.max_fused, "method": "plain-fp32",
            "num_gates": gpu["num_gates"], "fused_passes": gpu["passes"],
        },
        "gate_passes_per_s": gpu["passes"] / circuit_s,
        "hbm_gbps_avg_pass": 16.0 * (1 << gpu["num_qubits"]) /
                             (gpu["all_pass_avg_ms"] * 1e-3) / 1e9,
        "pass_ms_sum": gpu["pass_ms_sum"],
        "final_norm": gpu["norm"],
        "kernels": gpu["families"],
        "per_k": gpu["per_k"],
        "roofline": gpu["roofline"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
