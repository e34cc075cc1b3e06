#!/usr/bin/env python3
"""Benchmark: (omega, l) minor-ODE solves per second on one MI355X.

A *solve* is computeToroidalMinors + computeSpheroidalMinors for one (omega, l)
at the workload's node set.  One *step* is one pass over the whole workload:
819 frequencies (T = 4096 s, f <= 0.2 Hz) x l = 0..2500 = 2,048,319 solves on the
seeded 200-knot global model with 16 source nodes 0-700 km deep and the receiver
node at the surface.  Inputs are resident in HBM and results stay there.

Prints one JSON line (see the field list at the bottom of main()).
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

METRIC = "minor-ODE (ω,l) solves/s, learned vs plain step control at equal rel-err 1e-6"
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector peak


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(name):
    from gemini_unified_amd import ExternalRadialNodes, FrequencyGrid, global_model
    nem = global_model(1, 200)
    if name == "c2":
        nodes = ExternalRadialNodes.from_blocks(nem, [15], [700000.0 / 15])
        grid = FrequencyGrid.global_fixed_lmax(4096.0, 0.2, 2500)
        desc = ("C2: seeded 200-knot global model, T=4096 s, f<=0.2 Hz (819 frequencies), l=0..2500, "
                "16 source nodes 0-700 km, surface receiver, eps=1e-6")
    else:
        nodes = ExternalRadialNodes.from_blocks(nem, [1], [30000.0])
        grid = FrequencyGrid.global_fixed_lmax(2048.0, 0.05, 600)
        desc = ("C1: seeded 200-knot global model, T=2048 s, f<=0.05 Hz (102 frequencies), l=0..600, "
                "source 30 km, surface receiver, eps=1e-6")
    jre = nodes.node_index(nem.rearth)
    return nem, nodes, grid, jre, desc


def cpu_baseline(nem, nodes, grid, jre, eps, seconds, threads, seed=4):
    """Oracle (scalar C++ restatement, OpenMP over l) on a seeded sample of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from gemini_unified_amd import MC_PI
    orc = oracle.Oracle(nem, nodes)
    rng = np.random.default_rng(seed)
    order = rng.permutation(grid.freqs.size)
    per_f = 32
    done, t_used, nf = 0, 0.0, 0
    for q in order:
        f = grid.freqs[q]
        orc.set_frequency(f, 3)
        # a block of consecutive degrees at a seeded random start keeps every thread busy
        l0 = int(rng.integers(0, max(1, int(grid.nwn[q]) - per_f)))
        t0 = time.perf_counter()
        st = orc.minors_degrees(2.0 * MC_PI * f, -grid.sigma, l0, per_f, jre, jre, eps, threads)
        t_used += time.perf_counter() - t0
        done += per_f
        nf += 1
        if t_used > seconds:
            break
    assert np.all(st == 0)
    return done / t_used, done, nf


def spectrum_error(eng, nem, nodes, grid, jre, eps, nsample=24, seed=5):
    """Per-(omega,l) relative L2 error of the FK columns against the oracle at eps=1e-10."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from gemini_unified_amd import MC_PI
    orc = oracle.Oracle(nem, nodes)
    rng = np.random.default_rng(seed)
    qs = np.sort(rng.choice(grid.freqs.size, size=nsample, replace=False))
    ls = np.array([int(rng.integers(1, int(grid.nwn[q]))) for q in qs])
    r = eng.gfk(grid.freqs[qs], ls, 1, grid.sigma, eps, jre, 1)
    errs = []
    for k, (q, l) in enumerate(zip(qs, ls)):
        f = grid.freqs[q]
        orc.set_frequency(f, 3)
        rc, ref, status = orc.gfk_degrees(2.0 * MC_PI * f, -grid.sigma, int(l), 1, jre, 1, 1e-10)
        assert rc == 0
        g = r["gf"][:, :, k]
        o = ref[:, :, 0]
        errs.append(np.linalg.norm(g - o) / np.linalg.norm(o))
    errs = np.array(errs)
    return float(errs.max()), float(np.quantile(errs, 0.99)), int(nsample)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["c1", "c2"], default="c2")
    ap.add_argument("--variant", choices=["exact", "fast"], default="fast")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build measures one GPU; the (omega,l) grid shards over frequencies "
                                   "with no exchange during integration, multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from gemini_unified_amd import GreenFKEngine
    nem, nodes, grid, jre, desc = workload(args.workload)
    eps = 1e-6
    ninst = grid.nkf
    log(f"workload {args.workload}: {ninst} (omega,l) solves per step, {nodes.nnod} nodes, receiver node {jre}")
    eng = GreenFKEngine(nem, nodes, device=0, variant=args.variant)
    freqs, nwn = grid.freqs, grid.nwn

    def step():
        return eng.run_minors_resident(freqs, 0, nwn, grid.sigma, eps, jre, jre)

    for w in range(args.warmup):
        t0 = time.perf_counter()
        step()
        log(f"warmup {w + 1}/{args.warmup}: {time.perf_counter() - t0:.2f} s")
    t_steps = []
    stats = None
    for k in range(args.steps):
        t0 = time.perf_counter()
        stats = step()
        t_steps.append(time.perf_counter() - t0)
        log(f"step {k + 1}/{args.steps}: {t_steps[-1]:.3f} s")
    ms_per_step = 1e3 * sum(t_steps) / len(t_steps)
    value = ninst / (ms_per_step / 1e3)
    tm = eng.timing()          # per-stage kernel ms of the last step (HIP events on the engine stream)
    sph_flops = sum(int(eng.system_stats(s)[7]) for s in (0, 2, 3))
    tor_flops = int(eng.system_stats(1)[7])
    sph_ms, tor_ms = tm[3], tm[2]
    achieved = sph_flops / (sph_ms * 1e-3) / 1e12
    roofline = {"bound": "fp64-valu", "kernel": "k_spheroidal", "achieved": round(achieved, 4),
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
                "traffic": None, "kernel_ms": round(sph_ms, 3), "algorithmic_flops": sph_flops,
                "toroidal": {"kernel_ms": round(tor_ms, 3),
                             "achieved": round(tor_flops / (tor_ms * 1e-3) / 1e12, 4)}}
    log(f"plain: {value:.1f} solves/s; k_spheroidal {sph_ms:.1f} ms, {achieved:.3f} TFLOP/s algorithmic")

    err = None
    if not args.no_error:
        t0 = time.perf_counter()
        emax, ep99, ns = spectrum_error(eng, nem, nodes, grid, jre, eps)
        err = {"max": emax, "p99": ep99, "sample": ns, "truth": "CPU oracle at eps=1e-10"}
        log(f"spectrum rel-L2 vs truth: max {emax:.2e} p99 {ep99:.2e} ({time.perf_counter() - t0:.1f} s)")

    cpu = None
    if not args.no_cpu:
        threads = max(1, min(16, os.cpu_count() or 1))
        log(f"cpu baseline: oracle on {threads} threads for ~{args.cpu_seconds:.0f} s")
        rate, nsolves, nfreq = cpu_baseline(nem, nodes, grid, jre, eps, args.cpu_seconds, threads)
        cpu = {"value": round(rate, 3), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": f"{nsolves} solves: {nfreq} seeded random frequencies x 32 consecutive degrees at a "
                         f"seeded random start, same workload and eps"}
        log(f"cpu baseline: {rate:.2f} solves/s")

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": desc, "solves_per_step": ninst, "nodes": nodes.nnod, "variant": args.variant,
                   "step_control": "plain"},
        "plain": {"value": round(value, 3), "ms_per_step": round(ms_per_step, 3)},
        "learned": None,
        "spectrum_rel_err": err,
        "stage_ms": {"tables": round(tm[0], 3), "setup": round(tm[1], 3), "toroidal": round(tm[2], 3),
                     "spheroidal": round(tm[3], 3)},
        "work": {k: int(v) for k, v in zip(["propagations", "accepted_steps", "rejected_steps", "mmp_sequences",
                                            "matvecs", "sysmat_evals", "neville_cells", "flops"], stats[:8])},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
