#!/usr/bin/env python3
"""Benchmark of the MI355X per-step evolution loop (sort -> M(r) -> force ->
adaptive leapfrog / Levi-Civita), one GPU.

Workload (default): isotropic Plummer sphere, G = M = a = 1, N = 1.2e7
equal-mass particles (radii by inverse CDF, truncated at r = 100; speeds by
Aarseth-Henon-Wielen rejection; isotropic directions), 5 % of particles picked
at random with ell scaled by 1e-3; method-5 constants (tol 1e-7, Richardson
output, N_taumin 10, alpha 0.05), dt = 1e-3.  A step is one full pass of the
loop over all particles; inputs are resident in HBM before timing starts.

Prints one JSON line (see the contract in DESIGN.md).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0

# Algorithmic HBM bytes per particle for each kernel family of one step
# (float64 state, 64-bit keys, uint32 permutation / ids, 8-bit digits).
BYTES_PER_PARTICLE = {
    # read 8 B key
    "sort_histogram": 8.0,
    # average over 8 passes: first reads key 8 + writes key/val 12; six middle
    # passes read 12 + write 12; the last reads 12 + writes the 4 B permutation
    "sort_onesweep": (20.0 + 6 * 24.0 + 16.0) / 8.0,
    # read perm 4 + r, v, ell 24 + id 4; write r, v, ell 24 + id 4 + key 8 + level 1
    "integrate": 32.0 + 37.0,
    "radius_keys": 16.0,
}


def cpu_info() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(r, v, ell, ids, params, budget_s: float) -> dict:
    """The oracle (reference flags, OpenMP on all host cores, the reference's
    default chunked sort) timed on whole steps of the same workload."""
    import oracle

    threads = oracle.lib(True).orc_num_threads()
    po = oracle.from_step_params(params)
    r, v, ell, ids = r.copy(), v.copy(), ell.copy(), ids.copy()
    n = r.size
    steps = 0
    t0 = time.perf_counter()
    while True:
        oracle.step(r, v, ell, ids, po, chunked=True, fast=True)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 50:
            break
    return {
        "value": n * steps / el,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} whole steps of the same N={n} workload from the same initial state, "
                  f"C restatement built with the reference's -Ofast flags, chunked quadsort-order "
                  f"sort (--sort 1), OpenMP {threads} threads on {cpu_info()}",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=12_000_000)
    ap.add_argument("--seed", type=int, default=20250401)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the evolution loop here is built and timed on a "
              "single MI355X; run with --gpus 1", file=sys.stderr)
        return 2

    import nsphere_amd as ns

    n = args.n
    t_ic = time.perf_counter()
    r, v, ell, ids = ns.generate_ic(ns.IC_PLUMMER, n, args.seed, 100.0, 0.05, 1e-3)
    print(f"[bench] ICs N={n} in {time.perf_counter() - t_ic:.1f}s", file=sys.stderr, flush=True)
    p = ns.unit_params()

    ctx = ns.GpuContext(n)
    ctx.upload(r, v, ell, ids)
    ctx.sort_particles()  # the reference's exact initial sort, nsphere.c:5339-5342
    ctx.sync()
    K0, W0 = ctx.energy(p)
    # CPU baseline starts from the same sorted initial state.
    r0, v0, l0, id0 = ctx.download()

    ctx.run(p, args.warmup)
    ctx.sync()
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)

    ctx.set_option(ns.NS_OPT_PROFILE, 1)
    ctx.reset_kernel_stats()
    ctx.sync()
    t0 = time.perf_counter()
    ctx.run(p, args.steps)
    ctx.sync()
    el = time.perf_counter() - t0
    stats = ctx.kernel_stats()
    ctx.set_option(ns.NS_OPT_PROFILE, 0)
    K1, W1 = ctx.energy(p)
    lv = ctx.levels()
    print(f"[bench] timed {args.steps} steps in {el:.3f}s", file=sys.stderr, flush=True)

    value = n * args.steps / el
    kernels = {}
    for name, (launches, ms) in stats.items():
        avg_ms = ms / max(launches, 1)
        d = {"launches": launches, "total_ms": round(ms, 4), "avg_us": round(avg_ms * 1e3, 3)}
        if name in BYTES_PER_PARTICLE:
            gbs = BYTES_PER_PARTICLE[name] * n / (avg_ms * 1e-3) / 1e9
            d["algorithmic_GBps"] = round(gbs, 1)
        kernels[name] = d
    dom = max(stats.items(), key=lambda kv: kv[1][1])[0]
    dom_launches, dom_ms = stats[dom]
    dom_avg_s = dom_ms / dom_launches * 1e-3
    bpp = BYTES_PER_PARTICLE.get(dom)
    roofline = None
    if bpp is not None:
        achieved = bpp * n / dom_avg_s / 1e9
        roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                    "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                    "bytes_per_particle": bpp}

    e0, e1 = K0 + W0, K1 + W1
    lev = lv & ns.NS_LVL_MASK
    out = {
        "metric": "particle-steps/s at N=1.2e7, chosen form vs plain FP64 GPU at |dE/E|<=1e-6",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plummer_n12e6_lowl5pct_dt1e-3_plain_fp64", "n": n, "dt": p.dt,
                   "form": "plain-fp64", "lowl_frac": 0.05, "lowl_scale": 1e-3, "seed": args.seed},
        "roofline": roofline,
        "kernels": kernels,
        "dE_over_E": (e1 - e0) / abs(e0),
        "levels": {"hist": np.bincount(lev, minlength=8)[:16].tolist(),
                   "levi_civita": int(((lv & ns.NS_LVL_LC) != 0).sum())},
    }
    ctx.close()
    if not args.no_cpu:
        print("[bench] timing CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(r0, v0, l0, id0, p, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
