"""Benchmark: simulation steps per second of the ADMM contact step on one MI355X,
run to ADMM residual <= 1e-4 (capped at admm_max_iter), with the roofline of
the dominant kernel and the FP64 CPU oracle timed on the host beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cloth256|tetbox|cloth32]

Prints one JSON line.  A "step" is one simulation time step of the whole scene
(detection, local/contact/global ADMM iterations) on synthetic geometry with
positions resident in HBM.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def spmv_bytes(n: int, nnz_off: int) -> int:
    """Algorithmic bytes of one launch of the fused PCG direction + SpMV kernel
    (k_cg_dir_spmv: p' = z + beta p, q = A p', 3 interleaved FP32 columns):
    off-diagonal values + column ids (8 B each), row pointers (4(n+1)), diagonal
    (4n), z and p read once (24n), p' and q written (24n); the per-block
    partial sums are negligible."""
    return 8 * nnz_off + 4 * (n + 1) + 4 * n + 24 * n + 24 * n


def cpu_baseline(workload: str, scene, cfg, seconds: float = 20.0) -> dict:
    """FP64 oracle (prefactored sparse LLT + OpenMP local step) on the host cores,
    bounded to about `seconds` of stepping; returns steps/s."""
    from oracle.pyoracle import Oracle

    cores = os.cpu_count() or 1
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    o = Oracle(scene.mesh(), cfg)
    o.step()  # warm-up frame
    t0 = time.perf_counter()
    steps = 0
    while True:
        o.step()
        steps += 1
        if time.perf_counter() - t0 > seconds or steps >= 200:
            break
    dt = time.perf_counter() - t0
    return {"value": steps / dt, "unit": "steps/s", "cores": int(os.environ.get("OMP_NUM_THREADS", cores)),
            "kind": "port", "sample": f"{workload}: {steps} FP64 oracle steps after 1 warm-up step ({dt:.1f} s)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cloth256", choices=["cloth256", "tetbox", "cloth32"])
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(batched scenes and a single scene both run on one device here)", file=sys.stderr)
        return 2

    import parallel_admm_amd as pa
    from parallel_admm_amd.configs import SCENE_KIND, config_for

    scene = pa.Scene.generate(SCENE_KIND[args.workload])
    cfg = config_for(args.workload)
    cfg.residual_tol = args.tol
    cfg.residual_interval = 1
    backend = pa.ADMMBackendHIP.from_scene(scene, cfg)
    backend.set_sync_to_host(False)
    for _ in range(args.warmup):
        backend.step()
    print(f"[bench] {args.workload}: warmup {args.warmup} steps done", file=sys.stderr, flush=True)
    backend.set_profiling(1)
    backend.kernel_time(reset=True)
    stats = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        stats.append(backend.step())
    t1 = time.perf_counter()
    spmv_ms, spmv_launches = backend.kernel_time()
    backend.set_profiling(0)
    elapsed = t1 - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    iters = [s["iterations"] for s in stats]
    conv = sum(s["converged"] for s in stats)
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s; ADMM iters mean {np.mean(iters):.1f}, "
          f"converged {conv}/{args.steps}, contacts last {stats[-1]['n_contacts']}", file=sys.stderr, flush=True)

    # one extra (untimed) step with phase events, for the breakdown on stderr
    backend.set_profiling(2)
    st = backend.step()
    ph = backend.phase_times()
    backend.set_profiling(0)
    print("[bench] phase ms of one step (detection, local, contact, rhs, global, other): "
          + ", ".join(f"{v:.2f}" for v in ph) + f"; iterations {st['iterations']}, PCG {st['global_iters']}",
          file=sys.stderr, flush=True)

    # roofline of the PCG SpMV (the kernel moving the most bytes per launch in the global solve)
    sizes = backend.sizes()
    n = sizes["n_dyn"]
    nnz_off = sizes["nnz_a_off"]
    bytes_launch = spmv_bytes(n, nnz_off)
    achieved = bytes_launch / (spmv_ms * 1e-3) / 1e9 if spmv_ms > 0 else 0.0

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.workload, scene, cfg, args.cpu_seconds)

    out = {
        "metric": "sim steps/s at ADMM residual <=1e-4, plain vs learned warm-start; % HBM peak",
        "value": args.steps / elapsed,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "vertices": scene.info["n_all"],
            "dynamic_vertices": n,
            "constraint_rows": scene.info["m_rows"],
            "residual_tol": args.tol,
            "admm_max_iter": cfg.admm_max_iter,
            "warm_start": "plain (cold-start duals per detection)",
            "admm_iters_mean": float(np.mean(iters)),
            "converged_steps": conv,
            "contacts_last": stats[-1]["n_contacts"],
            "pcg_iters_per_step": float(np.mean([s["global_iters"] for s in stats])),
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_cg_dir_spmv",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_launch": bytes_launch,
            "avg_launch_ms": spmv_ms,
            "launches": spmv_launches,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
