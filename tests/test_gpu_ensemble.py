"""GPU tests of the batched ensemble form and of graph replay.

An ensemble steps many independent systems in one launch sequence.  Each
system's keys, tree, walk groups and heavy set depend on its own bodies only,
so its accelerations must be bit-identical to a run of that system alone; the
device direct sum (restricted to the sample's system) bounds the error per
system.  Graph replay of steady steps must reproduce eager steps bit for bit,
replay after replay."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HOUR = 1.0 / 24.0


def say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def make_systems(bh, sizes, seed0=11):
    kinds = [bh.GEN_BELT, bh.GEN_KUIPER, bh.GEN_TROJAN]
    out = []
    for k, n in enumerate(sizes):
        if n >= 9:
            out.append(bh.generate(kinds[k % 3], seed0 + k, n))
        else:   # tiny systems: the first n bodies of a generated one
            p, v, m = bh.generate(bh.GEN_BELT, seed0 + k, 9)
            out.append((p[:n].copy(), v[:n].copy(), m[:n].copy()))
    return out


def concat(systems):
    pos = np.concatenate([s[0] for s in systems])
    vel = np.concatenate([s[1] for s in systems])
    m = np.concatenate([s[2] for s in systems])
    return pos, vel, m


SIZES = [3000, 12000, 9, 25000, 1, 700]


def single_forces(bh, systems, form, theta):
    res = []
    for p, v, m in systems:
        with bh.Engine(form=form, theta=theta) as e:
            e.set_bodies(p, v, m)
            a = e.compute_accelerations(form=form, theta=theta)
            res.append((a, e.get_bodies()["ids"]))
    return res


@pytest.mark.parametrize("theta", [1.05, 0.5])
def test_ensemble_forces_bit_identical_to_single_runs(bh, theta):
    say("ensemble vs single", SIZES, theta)
    systems = make_systems(bh, SIZES)
    ref = single_forces(bh, systems, bh.FORM_MIXED, theta)
    pos, vel, m = concat(systems)
    with bh.Engine(form=bh.FORM_MIXED, theta=theta) as e:
        e.set_bodies(pos, vel, m)
        e.set_systems(SIZES)
        off = e.systems()
        assert list(np.diff(off)) == SIZES
        acc = e.compute_accelerations(form=bh.FORM_MIXED, theta=theta)
        ids = e.get_bodies()["ids"]
    for k, (a_ref, id_ref) in enumerate(ref):
        sl = slice(off[k], off[k + 1])
        assert np.array_equal(ids[sl] - off[k], id_ref), k
        assert np.array_equal(bits(acc[sl]), bits(a_ref)), k


def test_ensemble_steps_match_single_runs(bh):
    sizes = [5000, 800, 20000]
    say("ensemble steps", sizes)
    systems = make_systems(bh, sizes, seed0=30)
    ref = []
    for p, v, mm in systems:
        with bh.Engine(form=bh.FORM_MIXED, theta=1.05) as e:
            e.set_bodies(p, v, mm)
            e.step(5, HOUR)
            ref.append(e.get_bodies())
    pos, vel, m = concat(systems)
    with bh.Engine(form=bh.FORM_MIXED, theta=1.05) as e:
        e.set_bodies(pos, vel, m)
        e.set_systems(sizes)
        e.step(5, HOUR)
        assert e.graph_replays() > 0
        st = e.get_bodies()
        off = e.systems()
    for k, r in enumerate(ref):
        sl = slice(off[k], off[k + 1])
        for key in ("pos", "vel", "acc"):
            assert np.array_equal(bits(st[key][sl]), bits(r[key])), (k, key)


@pytest.mark.parametrize("form,n", [("mixed", 40000), ("learned", 40000), ("mixed", 1 << 20)])
def test_graph_replay_matches_eager(bh, form, n):
    """From 2^20 bodies the tree phase forks onto a second stream, so the
    captured graph has two branches; replays must still match eager steps."""
    f = bh.FORM_MIXED if form == "mixed" else bh.FORM_LEARNED
    say("graph vs eager", form, n)
    pos, vel, m = bh.generate(bh.GEN_TROJAN, 5, n)
    engines = []
    for graph in (False, True):
        e = bh.Engine(form=f, theta=1.05)
        e.model_configure(train_steps=2, iters_per_step=4, batch=512, lr=3e-3)
        e.set_graph(graph)
        e.set_bodies(pos, vel, m)
        engines.append(e)
    try:
        for k in range(6):
            states = []
            for e in engines:
                e.step(1, HOUR)
                states.append(e.get_bodies())
            for key in ("pos", "vel", "acc"):
                assert np.array_equal(bits(states[0][key]), bits(states[1][key])), (k, key)
        assert engines[0].graph_replays() == 0
        assert engines[1].graph_replays() >= 3
    finally:
        for e in engines:
            e.close()


def test_ensemble_error_vs_direct_sum(bh):
    sizes = [int(x) for x in bh.ensemble_sizes(3, 12, 10000, 60000)]
    say("ensemble error", sizes)
    systems = make_systems(bh, sizes, seed0=50)
    pos, vel, m = concat(systems)
    rng = np.random.default_rng(4)
    with bh.Engine(form=bh.FORM_MIXED, theta=1.05) as e:
        e.set_bodies(pos, vel, m)
        e.set_systems(sizes)
        acc = e.compute_accelerations(form=bh.FORM_MIXED, theta=1.05)
        off = e.systems()
        worst = 0.0
        for k in range(len(sizes)):
            idx = np.concatenate([off[k] + np.arange(9),
                                  off[k] + 9 + rng.choice(sizes[k] - 9, 64, replace=False)]).astype(np.int64)
            ds = e.direct_sum(idx)
            err = np.linalg.norm(acc[idx] - ds, axis=1) / np.linalg.norm(ds, axis=1)
            worst = max(worst, err.max())
        say("  worst", worst)
        assert worst < 1e-5


def test_ensemble_rejects_plain_and_bad_layouts(bh):
    systems = make_systems(bh, [400, 300])
    pos, vel, m = concat(systems)
    with bh.Engine(form=bh.FORM_PLAIN) as e:
        e.set_bodies(pos, vel, m)
        with pytest.raises(bh.BHError):
            e.set_systems([400, 299])
        with pytest.raises(bh.BHError):
            e.set_systems([700, 0])
        e.set_systems([400, 300])
        with pytest.raises(bh.BHError):
            e.compute_accelerations(form=bh.FORM_PLAIN)
        e.compute_accelerations(form=bh.FORM_MIXED)
        e.set_bodies(pos, vel, m)   # back to one system: the plain form runs again
        assert len(e.systems()) == 2
        e.compute_accelerations(form=bh.FORM_PLAIN)


def test_autotune_picks_a_candidate_and_keeps_accuracy(bh):
    sizes = [20000, 5000, 40000]
    systems = make_systems(bh, sizes, seed0=70)
    pos, vel, m = concat(systems)
    with bh.Engine(form=bh.FORM_MIXED, theta=1.05) as e:
        e.set_bodies(pos, vel, m)
        e.set_systems(sizes)
        r = e.autotune(reps=2)
        say("autotune", r)
        assert r["n_tried"] == 24
        assert r["gmax"] in (32, 48, 64) and r["efac"] in (1.0, 2.0, 4.0, 8.0)
        assert 0 < r["best_ms"] <= r["default_ms"]
        assert e.group_params()["gmax"] == r["gmax"]
        acc = e.compute_accelerations(form=bh.FORM_MIXED, theta=1.05)
        off = e.systems()
        idx = np.concatenate([off[k] + np.arange(9, 73) for k in range(3)]).astype(np.int64)
        ds = e.direct_sum(idx)
        err = np.linalg.norm(acc[idx] - ds, axis=1) / np.linalg.norm(ds, axis=1)
        assert err.max() < 1e-5
