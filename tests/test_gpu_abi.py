"""GPU tests of the C ABI's calling patterns: the per-call drop-in for the
reference's force path inside a host-side step loop (src/main.cpp:191-281),
both through ctypes and as the compiled integration sketch
(barnes-hut_amd/bh_dropin), bit for bit against the oracle's trajectory; the
device guard word staying raised across the steps of one call; and the
ensemble split refused once the bodies have been sorted."""
import pathlib
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HOUR = 1.0 / 24.0
DROPIN = pathlib.Path(__file__).resolve().parent.parent / "barnes-hut_amd" / "bh_dropin"


def say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def test_per_call_loop_matches_oracle_trajectory(bh, orc):
    """The host loop of src/main.cpp:191-281 with FP64 half kicks and drift
    on the host and one bh_set_bodies -> bh_compute_accelerations ->
    bh_get_bodies round per step (the host arrays take the engine's Morton
    order every step, as sortBodiesByMortonKey permutes them in the
    reference): 6 steps of 19,054 bodies equal oracle.simulate bit for bit."""
    n, steps, theta = 19054, 6, 1.05
    say("per-call loop", n, "bodies", steps, "steps")
    pos0, vel0, m0 = bh.generate(bh.GEN_BELT, 2, n)
    pos, vel, m = pos0.copy(), vel0.copy(), m0.copy()
    ids = np.arange(n, dtype=np.uint64)
    acc = np.zeros((n, 3))
    dt, t = HOUR, 0.0
    with bh.Engine(theta=theta) as eng:
        for _ in range(steps):
            if t > 0:
                vel = vel + (0.5 * dt) * acc      # main.cpp:193-200
            pos = pos + dt * vel                  # main.cpp:202-208
            eng.set_bodies(pos, vel, m, ids)
            eng.compute_accelerations(form=bh.FORM_PLAIN, theta=theta)
            st = eng.get_bodies()
            pos, vel, m, ids, acc = st["pos"], st["vel"], st["mass"], st["ids"], st["acc"]
            vel = vel + (0.5 * dt) * acc          # main.cpp:276-281
            t += dt
    o = orc.simulate(pos0, vel0, m0, steps, dt, theta=theta)
    assert np.array_equal(ids, o["ids"])
    for k, v in (("pos", pos), ("vel", vel), ("acc", acc), ("mass", m)):
        assert np.array_equal(bits(v), bits(o[k])), k


@pytest.mark.skipif(not DROPIN.exists(), reason="bh_dropin not built (make -C barnes-hut_amd)")
def test_compiled_integration_sketch_matches_oracle(bh, orc, tmp_path):
    """barnes-hut_amd/bh_dropin (the loop INTEGRATION.md prints) from a CSV:
    hourly steps up to 6 h of 19,054 bodies, final state bit-identical to the
    oracle run from the same CSV bits."""
    n = 19054
    pos, vel, m = bh.generate(bh.GEN_BELT, 3, n)
    csv = tmp_path / "belt.csv"
    bh.write_csv(csv, pos, vel, m)
    out = tmp_path / "final.csv"
    say("bh_dropin", n, "bodies, 6 steps")
    r = subprocess.run([str(DROPIN), str(csv), "1h", "6h", "1.05", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    got = np.loadtxt(out, delimiter=",", skiprows=1, dtype=np.float64)
    ids = np.loadtxt(out, delimiter=",", skiprows=1, usecols=0, dtype=np.uint64)
    rid, rm, rpos, rvel = bh.read_csv(csv)
    # `for (t = 0; t < 6h; t += 1h)` with t accumulated in days runs 7 steps
    # (the reference's loop, src/main.cpp:191)
    steps = bh.count_steps(bh.parse_time("1h"), bh.parse_time("6h"))
    assert steps >= 6
    o = orc.simulate(rpos, rvel, rm, steps, bh.parse_time("1h"), theta=1.05, ids=rid)
    assert np.array_equal(ids, o["ids"])
    assert np.array_equal(bits(got[:, 1:4]), bits(o["pos"]))
    assert np.array_equal(bits(got[:, 4:7]), bits(o["vel"]))
    assert np.array_equal(bits(got[:, 7:10]), bits(o["acc"]))


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_guard_bits_survive_the_steps_of_one_call(bh, form):
    """A guard bit raised before (or in) the first step of a multi-step
    bh_step is still reported when the call returns, although later steps
    raise nothing; the check clears it, so the next call succeeds and
    matches a run that never raised it."""
    f = {"plain": bh.FORM_PLAIN, "mixed": bh.FORM_MIXED}[form]
    pos, vel, m = bh.generate(bh.GEN_BELT, 4, 5000)
    say("sticky guard bits", form)
    with bh.Engine(form=f, theta=1.05) as eng:
        eng.set_bodies(pos, vel, m)
        eng.step(1, HOUR)
        eng.debug_raise_error(2)
        with pytest.raises(bh.BHError) as ei:
            eng.step(4, HOUR)   # eager step, then graph replays
        assert ei.value.code == bh.ERR_STATE
        eng.step(2, HOUR)       # cleared by the report
        a = eng.get_bodies()
    with bh.Engine(form=f, theta=1.05) as eng:
        eng.set_bodies(pos, vel, m)
        eng.step(7, HOUR)
        b = eng.get_bodies()
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(bits(a[k]), bits(b[k])), k


def test_set_systems_after_an_evaluation_is_refused(bh):
    """The first force evaluation sorts one system's bodies over the whole
    index range, so an ensemble split afterwards would mix systems."""
    pos, vel, m = bh.generate(bh.GEN_BELT, 5, 3000)
    with bh.Engine(form=bh.FORM_MIXED, theta=1.05) as eng:
        eng.set_bodies(pos, vel, m)
        eng.set_systems([1000, 2000])         # right after the upload: fine
        eng.set_bodies(pos, vel, m)
        eng.compute_accelerations(form=bh.FORM_MIXED, theta=1.05)
        with pytest.raises(bh.BHError) as ei:
            eng.set_systems([1000, 2000])
        assert ei.value.code == bh.ERR_STATE
        eng.set_bodies(pos, vel, m)           # a new upload allows it again
        eng.set_systems([1000, 2000])
