#!/usr/bin/env python3
"""Body-steps/s of the MI355X Barnes-Hut engine on one GPU.

Workload: 300,149 bodies from the synthetic heliocentric generator (central
mass, 8 planets, main-belt asteroids; seed 0), theta = 1.05, dt = 1 h, the
drift-first KDK step of the reference's `--fc tree` path.  A step is one pass
of the whole path over the resident bodies: half kick, drift, box, Morton
keys, radix sort + gather, octree build, walk, half kick.  Inputs are in HBM
before the timed region starts.

Prints one JSON line.  `value` is body-steps/s of the form given by --form
(default: the fastest form built), with the plain FP64 walk reported beside
it, the maximum relative force error against an FP64 direct sum on 4,096
seeded sample bodies plus the Sun and planets, the roofline of the walk kernel
(algorithmic flops per launch / measured walk time) and a CPU baseline: the
reference's own code (oracle/_ref, Release-like flags, OpenMP) on a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
HOUR = 1.0 / 24.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
METRIC = "body-steps/s @300k θ=1.05: mixed-prec+learned vs plain FP64; max rel force err"


def load_bh():
    spec = importlib.util.spec_from_file_location("bh_amd", ROOT / "barnes-hut_amd" / "__init__.py",
                                                  submodule_search_locations=[])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bh_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_form(bh, form, pos, vel, mass, theta, steps, warmup):
    eng = bh.Engine(form=form, theta=theta)
    eng.set_bodies(pos, vel, mass)
    eng.step(warmup, HOUR)
    t0 = time.perf_counter()
    eng.step(steps, HOUR)          # synchronous on return
    t1 = time.perf_counter()
    phases, nph = eng.step_phase_ms()
    return eng, (t1 - t0), phases


def force_error(bh, eng, n, samples, seed=0):
    st = eng.get_bodies()
    ids = st["ids"].astype(np.int64)
    where = np.empty(n, np.int64)
    where[ids] = np.arange(n)
    rng = np.random.default_rng(seed)
    sample_ids = np.concatenate([np.arange(9), rng.choice(np.arange(9, n), samples, replace=False)])
    si = where[sample_ids]
    ds = eng.direct_sum(si)
    err = np.linalg.norm(st["acc"][si] - ds, axis=1) / np.linalg.norm(ds, axis=1)
    return dict(max_all=float(err.max()), max_asteroids=float(err[9:].max()),
                max_central_and_planets=float(err[:9].max()), median=float(np.median(err)),
                p99=float(np.quantile(err, 0.99)), samples=int(samples), includes_ids_0_8=True)


def walk_work(bh, eng, theta):
    """Algorithmic flops of one plain walk launch: 9 per tested node (3 sub,
    5 for r^2 + soft2, 1 for theta^2 r^2) + 12 per accepted node (sqrt, div,
    2 mul for 1/r^3, 2 mul for G m / r^3, 3 mul + 3 add)."""
    eng.set_counting(True)
    eng.compute_accelerations(form=bh.FORM_PLAIN, theta=theta)
    acc, opened, tested = eng.counters()
    eng.set_counting(False)
    flops = 9.0 * tested.sum() + 12.0 * acc.sum()
    return flops, float(acc.mean()), float(tested.mean()), float(opened.mean())


def cpu_baseline(pos, vel, mass, theta, steps):
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    cores = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    n = len(mass)
    kind = None
    for k in ("fast", "strict"):
        if oracle.ref_available(k):
            kind = k
            break
    t0 = time.perf_counter()
    if kind:
        oracle.ref_simulate(pos, vel, mass, steps, HOUR, theta=theta, kind=kind)
        what = f"reference sources (src/morton_keys.cpp, linear_octree.cpp, traversal.cpp) built {kind}"
        k = "reference"
    else:
        oracle.simulate(pos, vel, mass, steps, HOUR, theta=theta)
        what = "C restatement oracle/bh_oracle.c"
        k = "port"
    dt = time.perf_counter() - t0
    return {"value": n * steps / dt, "unit": "body-steps/s", "cores": cores, "kind": k,
            "sample": f"{steps} KDK steps of the {n}-body workload (theta {theta}, dt 1 h) from t=0, "
                      f"{what}, OpenMP threads {cores}, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=300149)
    ap.add_argument("--theta", type=float, default=1.05)
    ap.add_argument("--form", choices=["plain", "mixed", "learned"], default="plain")
    ap.add_argument("--samples", type=int, default=4096)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU (one GPU holds the whole problem); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    bh = load_bh()
    forms = {"plain": bh.FORM_PLAIN, "mixed": bh.FORM_MIXED, "learned": bh.FORM_LEARNED}
    log(f"bench: generating {args.n} bodies")
    pos, vel, mass = bh.generate(bh.GEN_BELT, 0, args.n)

    log(f"bench: plain FP64, warmup {args.warmup}, steps {args.steps}")
    eng_p, wall_p, ph_p = run_form(bh, bh.FORM_PLAIN, pos, vel, mass, args.theta, args.steps, args.warmup)
    plain_rate = args.n * args.steps / wall_p
    err_p = force_error(bh, eng_p, args.n, args.samples)
    flops, mean_acc, mean_tested, mean_open_full = walk_work(bh, eng_p, args.theta)
    log(f"bench: plain {plain_rate:.4g} body-steps/s, phases ms {np.round(ph_p, 4).tolist()}")

    form = args.form
    if form == "plain":
        wall, ph, err, eng = wall_p, ph_p, err_p, eng_p
    else:
        log(f"bench: {form}, warmup {args.warmup}, steps {args.steps}")
        eng, wall, ph = run_form(bh, forms[form], pos, vel, mass, args.theta, args.steps, args.warmup)
        err = force_error(bh, eng, args.n, args.samples)
    rate = args.n * args.steps / wall

    walk_ms = float(ph_p[3])
    achieved = flops / (walk_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                "kernel": "k_walk_plain (plain FP64 walk + fused half kick)",
                "note": "compute roofline at the FP64 vector peak, which equals the dense FP64 MFMA figure; "
                        "the walk runs on VALU (a gather, not a contraction).  Work = 9 flops per tested node "
                        "+ 12 per accepted node, counted by the walk itself",
                "flops_per_launch": flops, "walk_ms": walk_ms,
                "per_body": {"accepted": mean_acc, "tested": mean_tested, "opened_full_chain": mean_open_full}}

    cpu = None
    if not args.no_cpu:
        log(f"bench: CPU baseline, {args.cpu_steps} steps")
        cpu = cpu_baseline(pos, vel, mass, args.theta, args.cpu_steps)

    line = {
        "metric": METRIC, "value": rate, "unit": "body-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if form == "plain" else "f32+f64",
        "data": "synthetic (project generator, seed 0)",
        "config": {"workload": f"heliocentric belt system, {args.n} bodies, theta {args.theta}, dt 1 h, KDK",
                   "n_bodies": args.n, "theta": args.theta, "dt_days": HOUR, "form": form},
        "max_rel_force_err": err["max_all"],
        "force_error": err,
        "plain_fp64": {"value": plain_rate, "ms_per_step": wall_p / args.steps * 1e3,
                       "phase_ms": {"kick_drift_box_keys": ph_p[0], "sort_gather": ph_p[1], "tree": ph_p[2],
                                    "walk_kick": ph_p[3], "total_device": ph_p[4]},
                       "force_error": err_p},
        "speedup_vs_plain": rate / plain_rate,
        "phase_ms": {"kick_drift_box_keys": ph[0], "sort_gather": ph[1], "tree": ph[2], "walk_kick": ph[3],
                     "total_device": ph[4]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
