#!/usr/bin/env python3
"""Body-steps/s of the MI355X Barnes-Hut engine on one GPU.

Default workload (--config c1): 300,149 bodies from the synthetic
heliocentric generator (central mass, 8 planets, main-belt asteroids; seed 0),
theta = 1.05, dt = 1 h, the drift-first KDK step of the reference's
`--fc tree` path.  A step is one pass of the whole path over the resident
bodies: half kick, drift, box, Morton keys, radix sort + gather, octree
build, walk, half kick.  Inputs are in HBM before the timed region starts;
the timed steps are the library's bh_step, which replays the step as a HIP
graph.

Prints one JSON line.  `value` is body-steps/s of the form given by --form
(default: LEARNED, the mixed-precision walk with the learned schedule, whose
cost model trains on the first 50 steps before the timed region), with the
plain FP64 walk (bit-identical to the reference), the mixed walk in key order
and the mixed walk with the schedule the learned one replaces (heaviest first
by the previous step's work, on the same states: gain_vs_prev_cost) reported
beside it; the maximum relative force error of each form against an FP64
direct sum on 4,096 seeded sample bodies plus the Sun and planets; the
roofline of the dominant walk kernel (algorithmic flops per launch, W = 21 I
+ 9 O_c from the plain walk's counters on the same state, / its mean
duration, HIP events on the library stream over eager steps that follow the
timed region); and a CPU baseline: the reference's own code (oracle/_ref,
Release-like flags, OpenMP) on a bounded sample of the same workload, timed
after one warm-up step.

Other configurations (--config): c0 (19,054 bodies, plain), c1-theta0.5,
c2 (4,194,304 bodies with a Kuiper belt, theta 0.7), c3 (1,048,576 bodies
with Trojan clouds, learned schedule), c4 (64 systems of log-uniform size in
[1e4, 2e5], one batched step, autotuned group shape, error per system on 512
samples each).
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


PHASE_STEPS = 5
MOMENT_CUT = None      # --moment-cut: the engine's tree-moment path (None: its default)


def run_form(bh, form, pos, vel, mass, theta, steps, warmup, train_steps=50, systems=None, tune=False,
             sched=None, shape=None, pre_steps=0):
    """Fresh engine; the LEARNED form first trains its cost model on the
    first `train_steps` steps (untimed), then every form runs `warmup`
    untimed and `steps` timed steps (graph replay).  PHASE_STEPS eager steps
    after the timed region give the per-phase device times.  sched / shape
    (a tuned group shape) / pre_steps (untimed steps before the warmup) let a
    MIXED run time another schedule on the states a LEARNED run times."""
    eng = bh.Engine(form=form, theta=theta)
    if MOMENT_CUT is not None:
        eng.set_moment_cutoff(MOMENT_CUT)
    eng.set_bodies(pos, vel, mass)
    if systems is not None:
        eng.set_systems(systems)
    if sched is not None:
        eng.set_schedule(sched)
    if shape is not None:
        eng.set_group_params(shape["gmax"], shape["efac"], shape["occ"], shape["walk_blocks"])
    tuned = eng.autotune(reps=3) if tune else None   # a true median of three timings per candidate
    if form == bh.FORM_LEARNED:
        eng.model_configure(train_steps=train_steps)
        eng.step(train_steps, HOUR)
    if pre_steps:
        eng.step(pre_steps, HOUR)
    eng.step(warmup, HOUR)
    r0 = eng.graph_replays()
    t0 = time.perf_counter()
    eng.step(steps, HOUR)          # synchronous on return
    t1 = time.perf_counter()
    replays = eng.graph_replays() - r0
    eng.set_graph(False)
    e0 = time.perf_counter()
    eng.step(PHASE_STEPS, HOUR)
    e1 = time.perf_counter()
    eng.set_graph(True)
    phases, nph = eng.step_phase_ms()
    return eng, (t1 - t0), phases, dict(graph_steps=replays, eager_ms_per_step=(e1 - e0) / PHASE_STEPS * 1e3,
                                        autotune=tuned)


def force_error(bh, eng, n, samples, seed=0):
    st = eng.get_bodies()
    ids = st["ids"].astype(np.int64)
    where = np.empty(n, np.int64)
    where[ids] = np.arange(n)
    rng = np.random.default_rng(seed)
    sample_ids = np.concatenate([np.arange(9), rng.choice(np.arange(9, n), min(samples, n - 9), replace=False)])
    si = where[sample_ids]
    ds = eng.direct_sum(si)
    err = np.linalg.norm(st["acc"][si] - ds, axis=1) / np.linalg.norm(ds, axis=1)
    return dict(max_all=float(err.max()), max_asteroids=float(err[9:].max()),
                max_central_and_planets=float(err[:9].max()), median=float(np.median(err)),
                p99=float(np.quantile(err, 0.99)), samples=int(len(sample_ids) - 9), includes_ids_0_8=True)


def ensemble_error(bh, eng, sizes, samples, seed=0):
    """Per system: ids 0-8 (central mass, planets) plus `samples` seeded
    asteroids, against the device FP64 direct sum over that system."""
    st = eng.get_bodies()
    off = eng.systems()
    rng = np.random.default_rng(seed)
    per = []
    for k, ns in enumerate(sizes):
        lo = int(off[k])
        ids = st["ids"][lo:lo + ns].astype(np.int64) - lo
        where = np.empty(ns, np.int64)
        where[ids] = np.arange(ns)
        pick = np.concatenate([np.arange(9), 9 + rng.choice(ns - 9, min(samples, ns - 9), replace=False)])
        si = lo + where[pick]
        ds = eng.direct_sum(si)
        err = np.linalg.norm(st["acc"][si] - ds, axis=1) / np.linalg.norm(ds, axis=1)
        per.append(dict(n=int(ns), max_all=float(err.max()), max_asteroids=float(err[9:].max()),
                        median=float(np.median(err))))
    return dict(max_all=max(p["max_all"] for p in per), max_asteroids=max(p["max_asteroids"] for p in per),
                median_of_medians=float(np.median([p["median"] for p in per])), samples_per_system=samples,
                includes_ids_0_8=True, per_system=per)


def algorithmic_work(bh, eng, theta):
    """Algorithmic flops of one walk over the engine's current state, the
    same for every form: W = 21 I + 9 O_c per body, I = nodes the reference's
    criterion accepts (src/traversal.cpp:63-98), O_c = the other compressed
    nodes it tests (opened branching nodes and the body's own leaf).  A
    tested node costs 9 (3 sub, 5 for r^2 + soft2, 1 for theta^2 r^2), an
    accepted one 12 more (sqrt, div, 2 mul for 1/r^3, 2 mul for G m / r^3,
    3 mul + 3 add).  I and O_c come from the plain walk's counters
    (bh_get_counters) in one counted PLAIN evaluation of the same state.  The
    compressed node count gives the walk's compulsory bytes (each 48-byte
    node record read once, and per body 24 B position + 8 B mass + 24 B
    velocity read, 24 B velocity + 24 B acceleration written)."""
    eng.set_counting(True)
    eng.compute_accelerations(form=bh.FORM_PLAIN, theta=theta)
    acc, opened, tested = eng.counters()
    eng.set_counting(False)
    n_full, n_comp = eng.tree_size()
    I = float(acc.sum())
    Oc = float(tested.sum()) - I
    n = len(acc)
    return dict(flops_per_launch=21.0 * I + 9.0 * Oc, accepted_per_body=I / n, opened_compressed_per_body=Oc / n,
                opened_full_chain_per_body=float(opened.mean()), tested_per_body=float(tested.mean()),
                compressed_nodes=int(n_comp), compulsory_bytes=48.0 * n_comp + 104.0 * n)


def algorithmic_work_any(bh, eng, theta, systems):
    """algorithmic_work of the engine's state; an ensemble (which the plain
    form does not run) is counted system by system on engines of its own."""
    if systems is None:
        return algorithmic_work(bh, eng, theta)
    st = eng.get_bodies()
    off = eng.systems()
    tot = None
    for k in range(len(systems)):
        sl = slice(int(off[k]), int(off[k + 1]))
        with bh.Engine(form=bh.FORM_PLAIN, theta=theta) as e1:
            e1.set_bodies(st["pos"][sl], st["vel"][sl], st["mass"][sl])
            w = algorithmic_work(bh, e1, theta)
        ns = sl.stop - sl.start
        if tot is None:
            tot = {k_: 0.0 for k_ in w}
        for k_, v in w.items():
            tot[k_] += v * ns if k_.endswith("_per_body") else v
    n = len(st["mass"])
    for k_ in tot:
        if k_.endswith("_per_body"):
            tot[k_] /= n
    tot["compressed_nodes"] = int(tot["compressed_nodes"])
    return tot


def mixed_walk_work(eng):
    """Flops the mixed walk executes in one launch, from its last groups:
    every body of a group evaluates the group's whole list, 20 FP32 flops per
    entry (3 sub, 3 FMA for r^2, rsqrt, 2 mul for 1/r^3, 1 mul by G m,
    3 FMA), and every group tests each visited node once, 12 FP64 flops
    (3 sub, 3 abs-sub, 3 max, 5 for d^2, 1 mul; counted as 12).  A group's
    list holds what the least-accepting body of the group needs, so this
    exceeds the algorithmic work of the reference's per-body criterion."""
    first, inter, visit = eng.groups()
    cnt = np.diff(first.astype(np.int64))
    flops = 20.0 * float(np.sum(cnt * inter.astype(np.float64))) + 12.0 * float(visit.sum())
    return flops, dict(groups=int(len(inter)), bodies_per_group=float(cnt.mean()),
                       list_per_group=float(inter.mean()), visited_per_group=float(visit.mean()),
                       list_entries_per_body=float(np.sum(cnt * inter.astype(np.float64)) / cnt.sum()),
                       lane_utilisation=float(cnt.sum() / (64.0 * len(inter))))


def cpu_model():
    """The host CPU's model name (lscpu), or the platform's processor string."""
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                return ln.split(":", 1)[1].strip()
    except Exception:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(pos, vel, mass, theta, steps):
    """The reference's force path on the host cores: one untimed warm-up KDK
    step from t = 0 (the first, unsorted build), then `steps` timed steps from
    that state (src/main.cpp:185,323,331-334 time the step loop the same way,
    wall clock around the steps).  It runs the reference's own sources
    (src/morton_keys.cpp, linear_octree.cpp, traversal.cpp) as built in
    oracle/_ref with Release-like flags; the built library travels with the
    tree, the sources do not.  Without it the oracle's C restatement runs
    (kind "port")."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    cores = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    n = len(mass)
    kind = None
    for k in ("fast", "strict"):
        if oracle.ref_available(k):
            kind = k
            break
    if kind:
        run = lambda p, v, m, k_, t0=0.0, acc=None: oracle.ref_simulate(p, v, m, k_, HOUR, theta=theta, kind=kind,
                                                                         t0=t0, acc=acc)
        what = f"reference sources (src/morton_keys.cpp, linear_octree.cpp, traversal.cpp) built {kind}"
        k = "reference"
    else:
        run = lambda p, v, m, k_, t0=0.0, acc=None: oracle.simulate(p, v, m, k_, HOUR, theta=theta, t0=t0, acc=acc)
        what = "C restatement oracle/bh_oracle.c"
        k = "port"
    w0 = time.perf_counter()
    st = run(pos, vel, mass, 1)
    w1 = time.perf_counter()
    t0 = time.perf_counter()
    run(st["pos"], st["vel"], st["mass"], steps, t0=HOUR, acc=st["acc"])
    dt = time.perf_counter() - t0
    model = cpu_model()
    return {"value": n * steps / dt, "unit": "body-steps/s", "cores": cores, "kind": k, "cpu_model": model,
            "sample": f"{steps} timed KDK steps of the {n}-body workload (theta {theta}, dt 1 h) after one untimed "
                      f"warm-up step from t=0 ({w1 - w0:.1f} s), {what}, OpenMP threads {cores} on {model}, "
                      f"{dt:.1f} s"}


CONFIGS = {
    "c1": dict(kind="belt", n=300149, theta=1.05, steps=100, form="learned", forms="plain,mixed,learned",
               workload="heliocentric belt system, 300149 bodies (central mass, 8 planets, asteroids), theta 1.05, "
                        "dt 1 h, drift-first KDK"),
    "c1-theta0.5": dict(kind="belt", n=300149, theta=0.5, steps=100, form="learned", forms="plain,mixed,learned",
                        workload="heliocentric belt system, 300149 bodies, theta 0.5, dt 1 h, drift-first KDK"),
    "c0": dict(kind="belt", n=19054, theta=1.05, steps=200, form="plain", forms="plain",
               workload="heliocentric belt system, 19054 bodies, plain FP64 walk, theta 1.05, dt 1 h, 200 KDK steps"),
    "c2": dict(kind="kuiper", n=4194304, theta=0.7, steps=48, form="learned", forms="mixed,learned",
               workload="belt + 30% Kuiper belt, 4194304 bodies, theta 0.7, dt 1 h, precision form at scale"),
    "c3": dict(kind="trojan", n=1048576, theta=1.05, steps=200, form="learned", forms="plain,mixed,learned",
               workload="belt 60% + Jupiter Trojans 2 x 15% + Hildas 10%, 1048576 bodies, theta 1.05, learned "
                        "schedule"),
    "c4": dict(kind="belt", ensemble=64, nmin=10000, nmax=200000, theta=1.05, steps=100, form="mixed",
               forms="mixed,learned",
               workload="64 independent belt systems (seeds 0..63), N log-uniform in [1e4, 2e5], one batched "
                        "graph-replayed step, autotuned group shape, theta 1.05, dt 1 h"),
}


def traffic_record(kernel, config):
    """Calibrated HBM bytes per launch of `kernel` on `config`, from a
    committed rocprofv3 summary (profiles/<kernel>.<config>.traffic.json), or
    None."""
    f = ROOT / "profiles" / f"{kernel}.{config}.traffic.json"
    if not f.exists():
        return None
    rec = json.loads(f.read_text())
    return rec


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--config", choices=sorted(CONFIGS), default="c1")
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: the config's)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=None, help="override the body count")
    ap.add_argument("--theta", type=float, default=None)
    ap.add_argument("--form", choices=["plain", "mixed", "learned"], default=None)
    ap.add_argument("--samples", type=int, default=4096)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-prev-cost", action="store_true",
                    help="skip the MIXED run with the previous-step schedule beside LEARNED")
    ap.add_argument("--cpu-only", action="store_true",
                    help="only time the CPU baseline of the config (--cpu-steps steps; an ensemble: that many "
                         "steps of each system in turn) and print its record")
    ap.add_argument("--forms", default=None,
                    help="forms to run besides --form (comma list); the plain form gives speedup_vs_plain")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB ten times (calibrates rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE against known bytes) and exit")
    ap.add_argument("--moment-cut", type=int, default=None,
                    help="deep cutoff of the tree moments: -1 chosen at graph capture (default), 0..21 fixed")
    args = ap.parse_args()
    global MOMENT_CUT
    MOMENT_CUT = args.moment_cut
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU (one GPU holds the whole problem); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch   # plumbing only: a plain copy kernel of known size
        n = args.copy_probe * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 10}))
        return

    cfg = CONFIGS[args.config]
    theta = cfg["theta"] if args.theta is None else args.theta
    steps = cfg["steps"] if args.steps is None else args.steps
    warmup = args.warmup
    form = cfg["form"] if args.form is None else args.form
    forms_arg = cfg["forms"] if args.forms is None else args.forms

    bh = load_bh()
    kinds = {"belt": bh.GEN_BELT, "kuiper": bh.GEN_KUIPER, "trojan": bh.GEN_TROJAN}
    systems = None
    if "ensemble" in cfg:
        sizes = [int(x) for x in bh.ensemble_sizes(0, cfg["ensemble"], cfg["nmin"], cfg["nmax"])]
        log(f"bench: generating {len(sizes)} systems, {sum(sizes)} bodies")
        parts = [bh.generate(kinds[cfg["kind"]], s, ns) for s, ns in enumerate(sizes)]
        pos = np.concatenate([q[0] for q in parts])
        vel = np.concatenate([q[1] for q in parts])
        mass = np.concatenate([q[2] for q in parts])
        systems = sizes
        n = int(sum(sizes))
    else:
        n = cfg["n"] if args.n is None else args.n
        log(f"bench: generating {n} bodies")
        pos, vel, mass = bh.generate(kinds[cfg["kind"]], 0, n)

    if args.cpu_only:
        # the reference's path has no ensembles: each system in turn, rates combined
        log(f"bench: CPU baseline only, {args.cpu_steps} steps")
        if systems is None:
            cpu = cpu_baseline(pos, vel, mass, theta, args.cpu_steps)
        else:
            secs, lo = 0.0, 0
            for ns in systems:
                r = cpu_baseline(pos[lo:lo + ns], vel[lo:lo + ns], mass[lo:lo + ns], theta, args.cpu_steps)
                secs += ns * args.cpu_steps / r["value"]
                lo += ns
            cpu = {"value": n * args.cpu_steps / secs, "unit": "body-steps/s", "cores": r["cores"],
                   "kind": r["kind"], "sample": f"{args.cpu_steps} KDK steps of each of the {len(systems)} "
                                                f"systems ({n} bodies) in turn, {secs:.1f} s"}
        print(json.dumps({"config": {"workload": cfg["workload"], "name": args.config, "n_bodies": n,
                                     "theta": theta}, "cpu_baseline": cpu}))
        return

    forms = {"plain": bh.FORM_PLAIN, "mixed": bh.FORM_MIXED, "learned": bh.FORM_LEARNED}
    results = {}
    wanted = set(f for f in forms_arg.split(",") if f) | {form}
    if systems is not None:
        wanted.discard("plain")
    work = None   # algorithmic walk work of the headline form's final state
    for name in ("plain", "mixed", "learned", "prev_cost"):
        if name == "prev_cost":
            # the schedule LEARNED replaces: heaviest first by the previous
            # step's recorded work, on the same states (MIXED and LEARNED
            # forces are bit-identical) with LEARNED's tuned group shape
            if "learned" not in results or args.no_prev_cost:
                continue
            log(f"bench: mixed with the previous-step schedule, warmup {warmup}, steps {steps}")
            shape = results["learned"]["autotune"]
            eng, wall, ph, extra = run_form(bh, bh.FORM_MIXED, pos, vel, mass, theta, steps, warmup,
                                            systems=systems, sched=bh.SCHED_PREV_COST, shape=shape, pre_steps=50)
        else:
            if name not in wanted:
                continue
            log(f"bench: {name}, warmup {warmup}, steps {steps}")
            eng, wall, ph, extra = run_form(bh, forms[name], pos, vel, mass, theta, steps, warmup, systems=systems,
                                            tune=name != "plain")
        r = dict(value=n * steps / wall, ms_per_step=wall / steps * 1e3,
                 phase_ms={"kick_drift_box_keys": ph[0], "sort_gather": ph[1], "tree_groups_schedule": ph[2],
                           "walk_kick": ph[3], "total_device": ph[4]},
                 phase_note=f"device phases from {PHASE_STEPS} eager steps after the timed region", **extra)
        if name == "prev_cost":
            r["schedule"] = "PREV_COST (heaviest first by the previous step's walk work), LEARNED's group shape"
        if systems is not None:
            r["force_error"] = ensemble_error(bh, eng, systems, 512)
        else:
            r["force_error"] = force_error(bh, eng, n, args.samples)
        if name != "plain":
            flops, gs = mixed_walk_work(eng)
            r["executed_work"] = dict(flops_per_launch=flops, **gs)
        w = algorithmic_work_any(bh, eng, theta, systems)
        r["algorithmic_work"] = w
        r["walk_tflops_algorithmic"] = w["flops_per_launch"] / (ph[3] * 1e-3) / 1e12
        if name == form:
            work = w
        if name == "learned":
            ready, samples, updates, loss = eng.model_status()
            r["model"] = dict(ready=ready, samples=samples, updates=updates, train_loss=loss,
                              trained_steps=50, note="trained on the first 50 steps before the timed region")
        results[name] = r
        log(f"bench: {name} {r['value']:.4g} body-steps/s, {r['ms_per_step']:.3f} ms/step "
            f"(eager {extra['eager_ms_per_step']:.3f}), phases {np.round(ph, 3).tolist()}, "
            f"max err {r['force_error']['max_all']:.3g}")
        eng.close()

    head = results[form]
    plain = results.get("plain")
    walk_ms = head["phase_ms"]["walk_kick"]
    achieved = work["flops_per_launch"] / (walk_ms * 1e-3) / 1e12
    if form == "plain":
        kname, kern, peak, bound = "k_walk_plain", "k_walk_plain (FP64 walk + fused half kick)", \
            FP64_PEAK_TFLOPS, "valu_fp64"
        peak_note = "FP64 vector peak"
    else:
        kname, kern, peak, bound = "k_walk_mixed", \
            "k_walk_mixed (FP32 group walk + FP64 heavy term + fused half kick)", FP32_PEAK_TFLOPS, "valu_fp32"
        peak_note = ("FP32 vector peak as published (it counts packed FP32 instructions, which the library's kernels "
                     "are compiled without: the non-packed ceiling is 78.6)")
    tr = traffic_record(kname, args.config)
    roofline = {"bound": bound, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": tr["bytes_per_launch"] if tr else None,
                "kernel": kern,
                "note": f"compute roofline at the {peak_note}; the walk runs on VALU (a gather with a data-dependent "
                        "opening decision, no dense contraction, so no MFMA).  achieved = algorithmic flops per "
                        "launch (W = 21 I + 9 O_c per body from the plain walk's counters on the same state, "
                        "identical for every form) / the walk phase's mean time over eager steps right after the "
                        "timed region (HIP events on the library stream; the phase is the walk kernel plus the "
                        "small heavy-body kernels)",
                "algorithmic_flops_per_launch": work["flops_per_launch"], "walk_ms": walk_ms,
                "work": work}
    if form != "plain":
        ex = head["executed_work"]["flops_per_launch"]
        roofline["executed_flops_per_launch"] = ex
        roofline["executed_over_algorithmic"] = ex / work["flops_per_launch"]
        roofline["executed_tflops"] = ex / (walk_ms * 1e-3) / 1e12
    if tr:
        roofline["traffic_source"] = tr.get("source")
        roofline["traffic_gbs"] = tr["bytes_per_launch"] / (walk_ms * 1e-3) / 1e9
        roofline["compulsory_bytes"] = work["compulsory_bytes"]
        roofline["traffic_ratio"] = tr["bytes_per_launch"] / work["compulsory_bytes"]

    cpu = None
    if not args.no_cpu and systems is None:
        log(f"bench: CPU baseline, {args.cpu_steps} steps")
        cpu = cpu_baseline(pos, vel, mass, theta, args.cpu_steps)

    line = {
        "metric": METRIC, "value": head["value"], "unit": "body-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if form == "plain" else "f32+f64",
        "data": "synthetic (project generator" + (", seeds 0..63)" if systems else ", seed 0)"),
        "config": {"workload": cfg["workload"], "name": args.config, "n_bodies": n, "theta": theta,
                   "dt_days": HOUR, "form": form, **({"systems": len(systems)} if systems else {})},
        "max_rel_force_err": head["force_error"]["max_all"],
        "speedup_vs_plain": head["value"] / plain["value"] if plain else None,
        "gain_vs_prev_cost": (head["value"] / results["prev_cost"]["value"]
                              if form == "learned" and "prev_cost" in results else None),
        "forms": results,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
