#!/usr/bin/env python3
"""Benchmark of the MI355X per-step evolution loop (sort -> M(r) -> force ->
adaptive leapfrog / Levi-Civita), one GPU.

Workload: isotropic Plummer sphere, G = M = a = 1, N = 1.2e7 equal-mass
particles (radii by inverse CDF truncated at r = 100; speeds by
Aarseth-Henon-Wielen rejection; isotropic directions), 5 % of particles picked
at random with ell scaled by 1e-3; method-5 constants (tol 1e-7, 4F-3C
output, N_taumin 10, alpha 0.05), dt = 1e-3.  A step is one full pass of the
loop over all particles; the state is resident in HBM before timing starts.

Both forms run from the same sorted initial state: the plain float64 form (one
lane per particle) and the chosen form; `value` is the chosen form's whole-job
particle-steps/s, the plain form's figure and the ratio are reported beside it.
Prints one JSON line.
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
PEAK_FP64_TFLOPS = 78.6       # MI355X FP64 (vector and matrix) peak
FLOPS_PER_FORCE_EVAL = 16     # FP64 flops per force evaluation + kick/drift, a divide counted as one
# HBM bytes per particle-step of the whole step.  Canonical (8-bit digits
# over keys as wide as the state, so 8 passes of 64-bit keys for float64):
# float64 304, float32 124.  This build's algorithm: between steps the
# float64 sort buckets 24-bit quantile keys (one scatter of key and position
# into per-bucket regions, 12 B, then one pass ordering each bucket, 20 B:
# 32 B instead of 200 B), and the integrate writes a 4-byte key instead of 8,
# so a float64 step moves 132 B.
CANONICAL_STEP_BYTES = {"f64": 304, "f32": 124}
STEP_BYTES = {"f64": 132, "f32": 124}


def bytes_per_particle(state_bytes: int) -> dict:
    """Algorithmic HBM bytes per particle of each kernel family of one step,
    for float64 (8) or float32 (4) state, uint32 permutation and ids, 8-bit
    digits.  The float32 sort passes 4 times over its 32-bit radius keys.
    The float64 sort between steps passes 3 times over the 24-bit quantile
    keys the integrate wrote (4 bytes each); its last pass also writes its
    keys for the fix-up that orders runs of equal keys by radius (read key 4
    and permutation 4, the 8-byte radius of the half of the positions that
    share their key, write the permutation 4: 16)."""
    s = float(state_bytes)
    wide = state_bytes == 8
    passes = 3 if wide else 4
    key = 4.0                        # the sort key and the key the integrate writes
    first = 4 + (4 + 4)              # read key; write key + payload
    middle = 2 * (4 + 4)             # read and write key + payload
    last = (4 + 4) + (8 if wide else 4)  # read key + payload; write (key +) permutation
    # read perm 4, r v ell 3s, id 4; write r v ell 3s, id 4, key, level 1
    integrate = (4 + 3 * s + 4) + (3 * s + 4 + key + 1)
    out = {
        "sort_histogram": 4,                                     # read the key (onesweep)
        "sort_upsweep": 4,                                       # read the key (reduce-then-scan)
        "sort_scatter": (first + (passes - 2) * middle + last) / passes,
        "integrate": integrate,                                  # plain form
        "integrate_bulk": integrate,                             # binned forms: every particle's gather + write
        "radius_keys": s + s,                                    # read r; write the key
        # read perm 4 and the mass by it; write the mass in rank order and M_enc (float64)
        "mass_scan": 4 + 2 * s + 8,
    }
    if wide:
        out["sort_fixup"] = 16
        # bucket sort of the quantile keys: read the key, write (key, position);
        # then per bucket read (key, position) and the radius at the position,
        # write the permutation
        out["qsort_scatter"] = 4 + 8
        out["qsort_bucket"] = 8 + 8 + 4
    return out


# HBM bytes per launch of the benchmark workload (N = 1.2e7), from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE runs of the same kernels, scaled by the counters' reading for kernels of known traffic;
# the profile each figure comes from is named beside it.  These are not measured in the bench run
# itself (a counter run is a separate rocprofv3 invocation), and a kernel whose code changed since
# its profile has no entry.
TRAFFIC_PROFILES = {
    # sort_scatter: mean of the 4 passes' read + write
    ("f32", 12_000_000): ("profiles/hbm_traffic_plummer_n1.2e7_binned_fast_f32_overlap.md",
                          {"sort_scatter": 171.4e6, "sort_upsweep": 50.04e6, "integrate_bulk": 496.6e6,
                           "integrate_lc": 7.51e6, "integrate_deep": 0.41e6, "integrate_late": 0.20e6}),
    ("f64", 12_000_000): ("profiles/hbm_traffic_plummer_n1.2e7_binned_fast_f64_overlap.md",
                          {"qsort_scatter": 145.1e6, "qsort_bucket": 278.9e6,
                           "integrate_bulk": 785.2e6, "integrate_lc": 9.99e6, "integrate_deep": 0.41e6,
                           "integrate_late": 0.46e6}),
}

WORK_ROW = {"integrate": "bulk_or_plain", "integrate_bulk": "bulk_or_plain", "integrate_deep": "deep",
            "integrate_late": "late"}
# Kernel families timed with HIP events inside the timed region (two events
# per launch): the integrate kernels and the sort's scatter, which the
# rooflines use.  The other families are timed over a short run after it.
SORT_KERNELS = ["sort_upsweep", "sort_scan", "sort_scatter", "sort_fixup", "qsort_scatter", "qsort_bucket"]
ROOFLINE_KERNELS = ["integrate", "integrate_bulk", "integrate_lc", "integrate_deep", "integrate_late", *SORT_KERNELS]
AFTER_STEPS = 20
# The late kernel's chains run at best at the rate of one wave alone on an idle
# device running its fast Levi-Civita micro-step (ns_gpu_chain_floor, measured
# at the start of each run).
CHAIN_FLOOR_STEPS = 1 << 22



# IC recipes (kind, default N, truncation radius, low-ell fraction; ell x 1e-3, dt = 1e-3): the
# benchmark's Plummer workload, and the Hernquist and Plummer+Hernquist workloads also run by the tests
WORKLOADS = {
    "plummer_n1.2e7_lowl5pct_dt1e-3": (0, 12_000_000, 100.0, 0.05),
    "hernquist_n5e7_lowl5pct_dt1e-3": (1, 50_000_000, 200.0, 0.05),
    "mix_n1.2e7_lowl10pct_dt1e-3": (2, 12_000_000, 100.0, 0.10),
    # 256 independent halos, N_h log-uniform in [1e3, 1e6], alternating Plummer / Hernquist
    "ensemble_256halos_lowl2pct_dt1e-3": (-1, 0, 0.0, 0.02),
}


def ensemble_ic(ns, seed: int, nhalos: int = 256, lowl: float = 0.02):
    """Halos of N_h = 10^U(3, 6) particles, Plummer (even h, r < 100) or Hernquist
    (odd h, r < 200), each in its own radius order, ids unique over the ensemble."""
    sizes = np.round(10.0 ** np.random.default_rng(seed).uniform(3.0, 6.0, nhalos)).astype(np.int64)
    parts, base = [], 0
    for h, nh in enumerate(sizes):
        kind = h % 2
        r, v, ell, ids = ns.generate_ic(kind, int(nh), seed + 1 + h, 100.0 if kind == 0 else 200.0, lowl, 1e-3)
        q = np.argsort(r, kind="stable")
        parts.append((r[q], v[q], ell[q], (ids[q] + base).astype(np.uint32)))
        base += int(nh)
    return [np.concatenate([pt[k] for pt in parts]) for k in range(4)], sizes

FORMS = {"plain": 0, "binned": 1, "binned-fast": 2}
PRECISIONS = {"f64": 0, "f32": 1}


def cpu_info() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(state, params, budget_s: float, work_cap: int, halos=None) -> dict:
    """The oracle with the reference's -Ofast flags, OpenMP on the host cores
    and the reference's default chunked sort, timed on whole steps of the same
    workload from the same initial state, under the same work cap.  An
    ensemble is stepped halo by halo, each as its own system (the reference
    runs one halo per process), for whole passes over all halos."""
    import oracle

    threads = oracle.lib(True).orc_num_threads()
    po = oracle.from_step_params(params, work_cap)
    r, v, ell, ids = (a.copy() for a in state)
    n = r.size
    if halos is None:
        bounds = [(0, n)]
    else:
        off = np.concatenate([[0], np.cumsum(halos)])
        bounds = list(zip(off[:-1], off[1:]))
        ids = np.concatenate([ids[a:b] - a for a, b in bounds]).astype(np.uint32)  # per-halo ids
    steps = 0
    t0 = time.perf_counter()
    while True:
        for a, b in bounds:
            hr, hv, hl, hi = r[a:b], v[a:b], ell[a:b], ids[a:b]
            oracle.step(hr, hv, hl, hi, po, chunked=True, fast=True)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 200:
            break
    return {
        "value": n * steps / el,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} whole steps of the N={n} workload from the benchmark's initial state "
                  f"({el:.1f} s); C restatement with the reference's -Ofast flags, chunked sort "
                  f"(--sort 1), work cap {work_cap}; {threads} OpenMP threads on {cpu_info()}",
    }


def cpu_p1(ns, seed: int, n: int = 100_000, steps: int = 1000, wall_limit_s: float = 120.0) -> dict:
    """The reference's CPU case run in full: Plummer, N = 1e5 equal masses
    (no low-ell tail), 1000 steps at dt = 1e-3, on the oracle with the
    reference's -Ofast flags, its default chunked sort and OpenMP on the host
    cores, with no work cap, as the reference runs it.  The initial exact sort
    (nsphere.c:5339-5342) is outside the timing; a wall limit stops the run
    early and says so."""
    import oracle

    r, v, ell, ids = ns.generate_ic(0, n, seed, 100.0, 0.0, 1e-3)
    q = oracle.sort_perm(r)
    r, v, ell, ids = (a[q].copy() for a in (r, v, ell, ids))
    po = oracle.params(dt=1e-3, budget=0)
    threads = oracle.lib(True).orc_num_threads()
    done = 0
    t0 = time.perf_counter()
    while done < steps and time.perf_counter() - t0 < wall_limit_s:
        oracle.step(r, v, ell, ids, po, chunked=True, fast=True)
        done += 1
    el = time.perf_counter() - t0
    return {"value": n * done / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "steps": done, "seconds": round(el, 3),
            "sample": f"N={n} Plummer (no low-ell tail), {done} of {steps} steps at dt=1e-3 run in full "
                      f"({el:.1f} s); C restatement with the reference's -Ofast flags, chunked sort "
                      f"(--sort 1), no work cap; {threads} OpenMP threads on {cpu_info()}"}


def graph_steps_for(warmup: int) -> int:
    """Overlapped steps per graph launch of a run that is not captured whole
    (odd, or past 128 steps): the most (up to 16, even) that the warmup run
    replays once after its first eager step, so that the capture happens
    there and not in the timed region."""
    return max(2, min(16, (warmup - 1) // 2 * 2))


def measure(ns, form: str, precision: str, state, p, steps: int, warmup: int, work_cap: int,
            mass=None, halos=None, graph=False, autotune=0, mlp=None, overlap=True, sort_keys=1) -> dict:
    """Time `steps` steps after `warmup`.  HIP-event timing of the roofline
    kernel families (ROOFLINE_KERNELS) runs over the timed region, unless
    graph: then the warmup captures the steps, the timed region replays them
    (graph nodes carry no per-kernel events) and the roofline families are
    timed over profiled eager steps right after it.  Every other family is
    timed over AFTER_STEPS profiled steps after the timed region."""
    n = state[0].size
    ctx = ns.GpuContext(n)
    # one host sync (and progress line) per chunk: each ns_gpu_run call ends
    # by joining its streams, so chunks are long enough to keep that rare
    chunk = 100
    bpp = bytes_per_particle(4 if precision == "f32" else 8)
    try:
        ctx.set_option(ns.NS_OPT_PRECISION, PRECISIONS[precision])
        ctx.set_option(ns.NS_OPT_INTEGRATOR, FORMS[form])
        ctx.set_option(ns.NS_OPT_BUDGET, work_cap)
        ctx.set_option(ns.NS_OPT_OVERLAP, int(overlap))
        ctx.set_option(ns.NS_OPT_SORT_KEYS, sort_keys)
        if graph:
            # every option before the warmup (setting one drops the captured
            # graph): the warmup run captures the graph, the timed region
            # only replays it
            ctx.set_option(ns.NS_OPT_GRAPH, 1)
            ctx.set_option(ns.NS_OPT_GRAPH_STEPS, graph_steps_for(warmup))
            ctx.set_option(ns.NS_OPT_PROFILE, 0)
        if halos is not None:
            ctx.set_halos(halos)
        ctx.upload(*state)
        if mass is not None:
            ctx.set_masses(mass)
        if mlp is not None:
            ctx.set_mlp(mlp)
        K0, W0 = ctx.energy(p)
        ctx.run(p, warmup)
        ctx.sync()
        tune = ctx.autotune(p, autotune) if autotune else None
        ctx.cap_hits(reset=True)
        if graph:  # an even run is one graph: capture it before the timed region
            ctx.graph_prepare(p, min(chunk, steps))
        else:
            ctx.set_option(ns.NS_OPT_PROFILE, 1)
            ctx.profile_kernels(ROOFLINE_KERNELS)
        ctx.reset_kernel_stats()
        ctx.sync()
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            k = min(chunk, steps - done)
            ctx.run(p, k)
            ctx.sync()
            done += k
            if done % 100 == 0 or done == steps:
                print(f"[bench] {form}: {done}/{steps} steps, {time.perf_counter() - t0:.2f}s",
                      file=sys.stderr, flush=True)
        el = time.perf_counter() - t0
        hits = ctx.cap_hits()
        K1, W1 = ctx.energy(p)
        lv = ctx.levels()
        if graph:
            # graph nodes carry no per-kernel events: the roofline kernels'
            # durations come from as many profiled eager steps (up to 20)
            # right after the timed region
            ctx.set_option(ns.NS_OPT_PROFILE, 1)
            ctx.profile_kernels(ROOFLINE_KERNELS)
            ctx.reset_kernel_stats()
            ctx.run(p, min(steps, 20))
            ctx.sync()
        stats = ctx.kernel_stats()
        evals = ctx.force_evals()
        chain = ctx.chain_steps()
        # every kernel family, over a short run after the timed region
        ctx.set_option(ns.NS_OPT_PROFILE, 1)
        ctx.profile_kernels(None)
        ctx.reset_kernel_stats()
        ctx.run(p, AFTER_STEPS)
        ctx.sync()
        after = ctx.kernel_stats()
        after_evals = ctx.force_evals()
        ctx.set_option(ns.NS_OPT_PROFILE, 0)
    finally:
        ctx.close()

    def table(st, ev_rows):
        out = {}
        for name, (launches, ms) in st.items():
            avg_ms = ms / max(launches, 1)
            d = {"launches": launches, "total_ms": round(ms, 4), "avg_us": round(avg_ms * 1e3, 3)}
            if name in bpp:
                d["algorithmic_GBps"] = round(bpp[name] * n / (avg_ms * 1e-3) / 1e9, 1)
            if name in WORK_ROW and ms > 0:
                ev = ev_rows[WORK_ROW[name]]
                d["force_evals_per_launch"] = ev / max(launches, 1)
                d["fp64_TFLOPs"] = round(FLOPS_PER_FORCE_EVAL * ev / (ms * 1e-3) / 1e12, 4)
            out[name] = d
        return out

    kernels = table(stats, evals)
    kernels_after = table(after, after_evals)
    lev = lv & ns.NS_LVL_MASK
    return {
        "precision": precision,
        "bpp": bpp,
        "seconds": el,
        "value": n * steps / el,
        "ms_per_step": el / steps * 1e3,
        "stats": stats,
        "evals": evals,
        "chain": chain,
        "kernels": kernels,
        "kernels_after": kernels_after,
        "E0": K0 + W0,
        "E1": K1 + W1,
        "cap_hits": hits,
        "graph": graph,
        "timing": (f"HIP events over {min(steps, 20)} profiled eager steps right after the graph-replayed "
                   "timed region (graph nodes carry no per-kernel events)") if graph else
                  "HIP events over the timed region",
        "autotune": tune,
        "levels": {"hist": np.bincount(lev, minlength=8)[:24].tolist(),
                   "levi_civita": int(((lv & ns.NS_LVL_LC) != 0).sum())},
    }


def train_level_predictor(ns, kind: int, n_log: int, rmax: float, lowl: float, seed: int, p, work_cap: int,
                          log_steps: int = 10, epochs: int = 4):
    """Levels logged by the plain adaptive integrator (FP64) over `log_steps`
    steps of an n_log-particle run of the recipe, the network trained on them
    by the library's own kernels; a second, smaller run gives held-out pairs."""
    t0 = time.perf_counter()

    def logged(n, s, steps):
        r, v, ell, ids = ns.generate_ic(max(kind, 0) if kind >= 0 else 2, n, s, rmax or 100.0, lowl, 1e-3)
        with ns.GpuContext(n) as ctx:
            ctx.set_option(ns.NS_OPT_INTEGRATOR, ns.FORM_PLAIN)
            ctx.set_option(ns.NS_OPT_BUDGET, work_cap)
            ctx.upload(r, v, ell, ids)
            return ns.level_samples(ctx, p, steps)

    X, y = logged(n_log, seed + 101, log_steps)
    Xt, yt = logged(max(n_log // 4, 1000), seed + 202, 2)
    t_log = time.perf_counter() - t0
    mlp = ns.Mlp(seed=seed)
    t1 = time.perf_counter()
    loss = mlp.train(X, y, epochs=epochs, batch=16384, lr=3e-3)
    t_train = time.perf_counter() - t1
    pred = mlp.predict(Xt)
    conf = np.zeros((8, 8), np.int64)
    np.add.at(conf, (yt.astype(np.int64), pred.astype(np.int64)), 1)
    deep = yt > 0
    print(f"[bench] level predictor: {X.shape[0]} pairs logged in {t_log:.1f}s, trained in {t_train:.1f}s",
          file=sys.stderr, flush=True)
    return mlp, {"log_particles": n_log, "log_steps": log_steps, "train_pairs": int(X.shape[0]),
                 "epochs": epochs, "train_seconds": round(t_train, 3), "final_loss": loss,
                 "held_out_pairs": int(Xt.shape[0]),
                 "held_out_level_gt0_recall": float(np.mean(pred[deep] > 0)) if deep.any() else None,
                 "held_out_level0_false_positive": float(np.mean(pred[~deep] > 0)),
                 "held_out_confusion": conf.tolist()}


def measure_snapshots(ns, precision: str, state, p, every: int, steps: int, work_cap: int, path: str) -> dict:
    """The snapshot path beside the step: `steps` steps of the chosen form
    with an all_particle_data snapshot (ns_apd_writer) every `every` steps,
    against the same steps without; the records kernel timed with HIP events."""
    n = state[0].size
    out = {"every": every, "steps": steps, "path": path, "bytes_per_snapshot": 16 * n}
    for with_snap in (False, True):
        ctx = ns.GpuContext(n)
        try:
            ctx.set_option(ns.NS_OPT_PRECISION, PRECISIONS[precision])
            ctx.set_option(ns.NS_OPT_INTEGRATOR, FORMS["binned-fast"])
            ctx.set_option(ns.NS_OPT_BUDGET, work_cap)
            ctx.upload(*state)
            ctx.run(p, 2)
            ctx.sync()
            w = ns.ApdWriter(ctx, path, block_size=1) if with_snap else None
            ctx.set_option(ns.NS_OPT_PROFILE, 1)
            ctx.reset_kernel_stats()
            t0 = time.perf_counter()
            for _ in range(steps // every):
                ctx.run(p, every)
                if w is not None:
                    w.write()
            kept = w.close() if w is not None else 0
            ctx.sync()
            el = time.perf_counter() - t0
            st = ctx.kernel_stats()
        finally:
            ctx.close()
        key = "with" if with_snap else "without"
        out[f"ms_per_step_{key}"] = el / (steps // every * every) * 1e3
        if with_snap:
            launches, ms = st.get("snapshot_records", (0, 0.0))
            avg = ms / max(launches, 1) * 1e-3
            out["snapshots"] = kept
            out["records_kernel_us"] = round(avg * 1e6, 2)
            # algorithmic bytes: read r, v, ell, id at the state's width, write a 16 B record
            bpp = (3 * (4 if precision == "f32" else 8) + 4) + 16
            out["records_kernel_GBps"] = round(bpp * n / avg / 1e9, 1) if avg > 0 else None
            out["records_bytes_per_particle"] = bpp
    out["cost_per_snapshot_ms"] = (out["ms_per_step_with"] - out["ms_per_step_without"]) * every
    return out


def measure_potential(ns, precision: str, state, reps: int = 20) -> dict:
    """Dynamic potential (ns_gpu_potential) of the benchmark state: the three
    scan kernels timed together with HIP events; algorithmic bytes = two reads
    of r and one write of psi (float64)."""
    n = state[0].size
    p = ns.default_params()
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_PRECISION, PRECISIONS[precision])
        ctx.upload(*state)
        ctx.potential(p)
        ctx.set_option(ns.NS_OPT_PROFILE, 1)
        ctx.reset_kernel_stats()
        for _ in range(reps):
            ctx.potential(p)
        launches, ms = ctx.kernel_stats().get("potential", (0, 0.0))
    avg = ms / max(launches, 1) * 1e-3
    bpp = 2 * (4 if precision == "f32" else 8) + 8
    return {"calls": launches, "avg_us": round(avg * 1e6, 2), "bytes_per_particle": bpp,
            "GBps": round(bpp * n / avg / 1e9, 1) if avg > 0 else None,
            "note": "includes the copy of psi to the host only outside the timed kernels"}


def measure_density(ns, precision: str, state, reps: int = 3) -> dict:
    """The snapshot density estimate (ns_gpu_density) of the benchmark state,
    timed end to end with HIP events (device kernels plus the two host
    tridiagonal solves between them)."""
    n = state[0].size
    p = ns.default_params()
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_PRECISION, PRECISIONS[precision])
        ctx.upload(*state)
        _, info = ctx.density(p)
        ctx.set_option(ns.NS_OPT_PROFILE, 1)
        ctx.reset_kernel_stats()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.density(p)
        wall = (time.perf_counter() - t0) / reps
        launches, ms = ctx.kernel_stats().get("density", (0, 0.0))
    return {"calls": launches, "avg_ms": round(ms / max(launches, 1), 3), "wall_ms_per_call": round(wall * 1e3, 3),
            "info": info}


def measure_rank_files(ns, precision: str, state, path: str = "/dev/null") -> dict:
    """The two Rank_Mass_Rad_VRad files of the benchmark state (ns_gpu_rank_files),
    wall time per call with the rows written to `path`."""
    n = state[0].size
    p = ns.default_params()
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_PRECISION, PRECISIONS[precision])
        ctx.upload(*state)
        ctx.rank_files(p, path, path)
        t0 = time.perf_counter()
        ctx.rank_files(p, path, path)
        wall = time.perf_counter() - t0
    return {"wall_ms": round(wall * 1e3, 2), "rows": 2 * n, "bytes": 60 * n, "path": path}


def traffic_for(run: dict, n: int, kernel: str):
    """(bytes per launch, profile) for the kernel, or (None, None)."""
    src, table = TRAFFIC_PROFILES.get((run["precision"], n), (None, {}))
    b = table.get(kernel)
    return (b, src) if b is not None else (None, None)


def roofline_for(run: dict, n: int, kernel: str) -> dict | None:
    """HBM roofline of a streaming kernel, or, for the deep and Levi-Civita
    wave kernels (a few thousand particles, latency-bound chains of dependent
    FP64 operations), its FP64 rate against the FP64 peak."""
    stats = run["stats"]
    if kernel not in stats:
        return None
    launches, ms = stats[kernel]
    avg_s = ms / launches * 1e-3
    if kernel in run["bpp"]:
        bpp = run["bpp"][kernel]
        achieved = bpp * n / avg_s / 1e9
        traffic, src = traffic_for(run, n, kernel)
        return {"kernel": kernel, "bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                "traffic_source": src, "bytes_per_particle": bpp, "avg_us": round(avg_s * 1e6, 2),
                "timing": run["timing"]}
    if kernel in WORK_ROW:
        ev = run["evals"][WORK_ROW[kernel]] / launches
        flops = FLOPS_PER_FORCE_EVAL * ev
        achieved = flops / avg_s / 1e12
        traffic, src = traffic_for(run, n, kernel)
        out = {"kernel": kernel, "bound": "fp64", "achieved": round(achieved, 4), "peak": PEAK_FP64_TFLOPS,
               "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 6),
               "traffic": traffic, "traffic_source": src,
               "force_evals_per_launch": ev, "flops_per_force_eval": FLOPS_PER_FORCE_EVAL,
               "avg_us": round(avg_s * 1e6, 2), "timing": run["timing"]}
        if kernel == "integrate_late" and run.get("chain"):
            # the launch lasts as long as its longest Levi-Civita chain: time
            # per micro-step of that chain against one wave running the same
            # micro-step alone on an idle GPU
            ns_step = ms * 1e6 / run["chain"]
            floor = run.get("chain_floor_ns")
            out["latency"] = {"chain_micro_steps_per_launch": round(run["chain"] / launches, 1),
                              "ns_per_micro_step": round(ns_step, 2), "floor_ns": floor,
                              "floor_source": "ns_gpu_chain_floor, one wave alone, measured in this run",
                              "frac": round(floor / ns_step, 3) if floor else None}
            if floor and run.get("ms_per_step"):
                # the timed region's step period against that chain at the
                # floor: the late kernel bounds the step when this is near 1
                chain_ms = run["chain"] / launches * floor * 1e-6
                out["latency"]["timed_step_vs_chain_floor"] = {
                    "chain_at_floor_ms": round(chain_ms, 4), "ms_per_step": round(run["ms_per_step"], 4),
                    "frac": round(chain_ms / run["ms_per_step"], 3),
                    "note": "chain length from the profiled steps, step period from the timed region"}
        return out
    return None


SIDE_STREAM = {"integrate_late", "overlap_plan", "overlap_order"}


def dominant_kernel(run: dict) -> str:
    """The kernel family with the most time on the stream that bounds the
    step: in an overlapped run the late kernel's stream when its kernels add
    up to more than the context stream's, else the context stream."""
    stats = run["stats"]
    side = {k: v[1] for k, v in stats.items() if k in SIDE_STREAM}
    main = {k: v[1] for k, v in stats.items() if k not in SIDE_STREAM}
    pool = side if side and sum(side.values()) > sum(main.values()) else main
    return max(pool.items(), key=lambda kv: kv[1])[0]


def sort_roofline(run: dict, n: int) -> dict | None:
    """The whole radius sort of one step -- tile counts, their scan, the
    scatter passes and (float64 keys) the run fix-up -- as one HBM-bound unit:
    the algorithmic bytes of all its launches over their summed durations
    (the scan moves a few MB of tile counts; counted as zero)."""
    stats = run["stats"]
    if "sort_scatter" not in stats and "qsort_scatter" not in stats:
        return None
    steps = (stats.get("qsort_scatter", (0,))[0] + stats.get("sort_fixup", (0,))[0]) or stats["sort_scatter"][0] / 4
    total_ms, total_bytes, parts = 0.0, 0.0, {}
    for k in SORT_KERNELS:
        if k not in stats:
            continue
        launches, ms = stats[k]
        b = run["bpp"].get(k, 0.0) * n * launches
        total_ms += ms
        total_bytes += b
        parts[k] = {"launches": launches, "avg_us": round(ms / launches * 1e3, 2),
                    "bytes_per_particle": run["bpp"].get(k, 0.0)}
    achieved = total_bytes / (total_ms * 1e-3) / 1e9
    return {"kernel": "sort (all passes)", "bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
            "us_per_sort": round(total_ms / steps * 1e3, 1),
            "bytes_per_particle_per_sort": round(total_bytes / n / steps, 1), "kernels": parts}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000,
                    help="timed steps (the metric is quoted over 1000 steps at dt = 1e-3)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="plummer_n1.2e7_lowl5pct_dt1e-3")
    ap.add_argument("--n", type=int, default=None, help="particles (default: the workload's)")
    ap.add_argument("--masses", choices=["none", "equal", "random"], default="none",
                    help="per-particle masses, so that M(r) comes from the enclosed-mass scan instead of "
                         "the rank (equal: M/N each; random: uniform in [0.5, 1.5] M/N, normalised)")
    ap.add_argument("--seed", type=int, default=20250401)
    ap.add_argument("--dt", type=float, default=1e-3,
                    help="base time step (the workloads are quoted at 1e-3)")
    ap.add_argument("--form", choices=sorted(FORMS), default="binned-fast")
    ap.add_argument("--precision", choices=sorted(PRECISIONS), default="f64",
                    help="state precision of the chosen form (f64 by default: over 1000 steps the float32 "
                         "state drifts 4e-6 from the plain form's energy, past the 1e-6 bound; f32: float32 "
                         "state and 32-bit keys, float64 arithmetic); the plain form is always f64")
    ap.add_argument("--plain-steps", type=int, default=None,
                    help="steps for the plain-form comparison (default: --steps)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-overlap", action="store_true",
                    help="every kernel on one stream (NS_OPT_OVERLAP=0): the same results, for comparison")
    ap.add_argument("--sort-keys", type=int, choices=[0, 1, 2], default=1,
                    help="NS_OPT_SORT_KEYS of the chosen form: 1 (default) buckets the 24-bit quantile keys between "
                         "steps, 2 sorts them with three radix passes and a fix-up, 0 sorts the 64-bit radius keys")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every step's kernels eagerly (by default the timed region replays HIP "
                         "graphs of captured steps, NS_OPT_GRAPH, captured during the warmup)")
    ap.add_argument("--mlp", type=int, default=0, metavar="N_LOG",
                    help="learned binning: log levels of the plain adaptive integrator on a separate "
                         "N_LOG-particle run of the same recipe (10 steps), train the level predictor on "
                         "them with the library's kernels, and time the chosen form with and without it")
    ap.add_argument("--snapshot-every", type=int, default=0, metavar="K",
                    help="also time the snapshot path: an all_particle_data snapshot every K steps "
                         "(ns_apd_writer, written to --snapshot-path) against the same steps without")
    ap.add_argument("--snapshot-path", default="/dev/null")
    ap.add_argument("--rank-files", action="store_true",
                    help="also time the Rank_Mass_Rad_VRad files (ns_gpu_rank_files, rows to /dev/null)")
    ap.add_argument("--density", action="store_true",
                    help="also time the snapshot density estimate (ns_gpu_density) on the benchmark state")
    ap.add_argument("--potential", action="store_true",
                    help="also time the dynamic potential (ns_gpu_potential) on the benchmark state")
    ap.add_argument("--ic-on-gpu", action="store_true",
                    help="draw the initial conditions on the device (ns_gpu_generate_ic) instead of the host")
    ap.add_argument("--autotune", type=int, default=0, metavar="STEPS",
                    help="autotune the chosen form's launch configuration first (STEPS timed steps per trial)")
    ap.add_argument("--budget", type=int, default=1 << 18,
                    help="force evaluations allowed per particle-step, the same for every form and "
                         "the CPU baseline (the reference has no cap: its first-order Levi-Civita tau steps "
                         "can need 2^28 subdivisions on plunging orbits)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the evolution loop is built and timed on a single "
              "MI355X; run it with --gpus 1", file=sys.stderr)
        return 2

    import nsphere_amd as ns

    chain_floor = ns.chain_floor(0, CHAIN_FLOOR_STEPS)  # on the idle device, before any other work
    kind, n_default, rmax, lowl = WORKLOADS[args.workload]
    t_ic = time.perf_counter()
    p = ns.unit_params(dt=args.dt)
    halos = None
    t_drawn = None
    if kind < 0:
        state, halos = ensemble_ic(ns, args.seed, lowl=lowl)
        n = state[0].size
    else:
        n = args.n or n_default
        # the reference's exact initial sort (nsphere.c:5339-5342), done once
        with ns.GpuContext(n) as ctx:
            if args.ic_on_gpu:
                ctx.generate_ic(kind, args.seed, rmax, lowl, 1e-3)
            else:
                ctx.upload(*ns.generate_ic(kind, n, args.seed, rmax, lowl, 1e-3))
            ctx.sync()
            t_drawn = time.perf_counter() - t_ic
            ctx.sort_particles()
            state = ctx.download()
    print(f"[bench] N={n} initial state in {time.perf_counter() - t_ic:.1f}s", file=sys.stderr, flush=True)

    mass = None
    if args.masses != "none":
        by_id = np.full(n, 1.0 / n) if args.masses == "equal" else \
            np.random.default_rng(args.seed + 1).uniform(0.5, 1.5, n)
        by_id /= by_id.sum()
        mass = by_id[state[3]]  # in the order of the sorted initial state
    mlp, learned = None, None
    if args.mlp:
        mlp, learned = train_level_predictor(ns, kind, args.mlp, rmax, lowl, args.seed, p, args.budget)
    chosen = measure(ns, args.form, args.precision, state, p, args.steps, args.warmup, args.budget, mass, halos,
                     not args.no_graph, args.autotune, mlp, overlap=not args.no_overlap, sort_keys=args.sort_keys)
    if mlp is not None:  # the same form without the network
        without = measure(ns, args.form, args.precision, state, p, args.steps, args.warmup, args.budget, mass,
                          halos, not args.no_graph, args.autotune, overlap=not args.no_overlap,
                          sort_keys=args.sort_keys)
        learned.update({"value_without": without["value"], "ms_per_step_without": without["ms_per_step"],
                        "gain": chosen["value"] / without["value"],
                        "kernels_without": without["kernels"],
                        "identical_energy": chosen["E1"] == without["E1"]})
    plain = None
    if not args.no_plain and (args.form, args.precision) != ("plain", "f64"):
        plain = measure(ns, "plain", "f64", state, p, args.plain_steps or args.steps, args.warmup, args.budget,
                        mass, halos)

    dom = dominant_kernel(chosen)
    chosen["chain_floor_ns"] = chain_floor
    out = {
        "metric": "particle-steps/s at N=1.2e7, chosen form vs plain FP64 GPU at |dE/E|<=1e-6",
        "value": chosen["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": chosen["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "n": n, "dt": p.dt, "form": args.form,
                   "state": args.precision, "lowl_frac": lowl, "lowl_scale": 1e-3, "seed": args.seed,
                   "masses": args.masses, "halos": None if halos is None else int(len(halos)),
                   "graph": not args.no_graph, "overlap": not args.no_overlap and args.form != "plain",
                   "sort_keys": args.sort_keys},
        "autotune": chosen["autotune"],
        "roofline": roofline_for(chosen, n, dom),
        "roofline_sort": sort_roofline(chosen, n),
        "roofline_late": roofline_for(chosen, n, "integrate_late"),
        "roofline_scatter": roofline_for(chosen, n, "qsort_scatter") or roofline_for(chosen, n, "sort_scatter"),
        "roofline_bucket": roofline_for(chosen, n, "qsort_bucket"),
        "roofline_step": {"bound": "hbm", "bytes_per_particle_step": STEP_BYTES[args.precision],
                          "canonical_bytes_per_particle_step": CANONICAL_STEP_BYTES[args.precision],
                          "achieved": round(chosen["value"] * STEP_BYTES[args.precision] / 1e9, 1),
                          "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": round(chosen["value"] * STEP_BYTES[args.precision] / 1e9 / PEAK_HBM_GBS, 4)},
        "dominant_kernel": dom,
        "kernels": chosen["kernels"],
        "kernels_after": {"steps": AFTER_STEPS, "note": "every kernel family, timed over the steps after the "
                                                        "timed region", **chosen["kernels_after"]},
        "dE_over_E": (chosen["E1"] - chosen["E0"]) / abs(chosen["E0"]),
        "levels": chosen["levels"],
        "work_cap": {"force_evals_per_particle_step": args.budget, "capped_particle_steps": chosen["cap_hits"]},
        "initial_conditions": {"drawn_on": "gpu" if args.ic_on_gpu else "host",
                               "seconds": round(t_drawn, 4) if t_drawn is not None else None},
    }
    if learned is not None:
        out["config"]["learned_binning"] = True
        out["learned_binning"] = learned
    if args.potential:
        out["potential"] = measure_potential(ns, args.precision, state)
    if args.density:
        out["density"] = measure_density(ns, args.precision, state)
    if args.rank_files:
        out["rank_files"] = measure_rank_files(ns, args.precision, state)
    if args.snapshot_every:
        print("[bench] timing the snapshot path", file=sys.stderr, flush=True)
        out["snapshot"] = measure_snapshots(ns, args.precision, state, p, args.snapshot_every,
                                            max(args.snapshot_every, min(args.steps, 50)), args.budget,
                                            args.snapshot_path)
    if plain is not None:
        out["plain_fp64"] = {"value": plain["value"], "ms_per_step": plain["ms_per_step"],
                             "kernels": plain["kernels"],
                             "dE_over_E": (plain["E1"] - plain["E0"]) / abs(plain["E0"]),
                             "capped_particle_steps": plain["cap_hits"]}
        out["speedup_vs_plain"] = chosen["value"] / plain["value"]
        if (args.plain_steps or args.steps) == args.steps:
            out["energy_vs_plain"] = abs(chosen["E1"] - plain["E1"]) / abs(plain["E1"])
    if not args.no_cpu:
        print("[bench] timing CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(state, p, args.cpu_seconds, args.budget, halos)
        out["cpu_baseline"]["p1_full_run"] = cpu_p1(ns, args.seed)
        if mass is not None:
            out["cpu_baseline"]["sample"] += "; equal masses from the rank (the oracle step has no mass scan)"
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
