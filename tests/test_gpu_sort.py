"""GPU tests of the near-sorted key sort.

The step's key sort must give the stable order by Morton key (and by system
first in an ensemble), whichever path it takes.  The near-sorted path radix
sorts only the keys that moved out of order since the previous step and
merges them with the rest; the full path radix sorts every key.  Both must
produce bit-identical bodies, step after step, and the near path must be the
one taken once the bodies are in sorted order."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HOUR = 1.0 / 24.0


def say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def run(bh, mode, form, pos, vel, m, steps, theta=1.05, systems=None, graph=True):
    with bh.Engine(form=form, theta=theta) as e:
        e.set_sort_mode(mode)
        e.set_graph(graph)
        e.set_bodies(pos, vel, m)
        if systems is not None:
            e.set_systems(systems)
        e.compute_accelerations(form=form, theta=theta)
        first = e.sort_stats()
        e.step(steps, HOUR)
        last = e.sort_stats()
        return e.get_bodies(), first, last


def same(a, b):
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(bits(a[k]), bits(b[k])), k
    assert np.array_equal(a["ids"], b["ids"])


@pytest.mark.parametrize("form_name", ["plain", "mixed"])
def test_near_sorted_sort_matches_full_sort(bh, form_name):
    form = bh.FORM_PLAIN if form_name == "plain" else bh.FORM_MIXED
    n = 19054 if form_name == "plain" else 300149
    say("near vs full sort", form_name, n)
    pos, vel, m = bh.generate(bh.GEN_BELT, 4, n)
    near, first, last = run(bh, bh.SORT_NEAR, form, pos, vel, m, 6)
    full, _, _ = run(bh, bh.SORT_FULL, form, pos, vel, m, 6)
    same(near, full)
    # generated bodies come in no particular order: the first sort sorts every key
    assert first == (False, n)
    took_near, keys = last
    say("last step: near", took_near, "keys radix sorted", keys, "of", n)
    assert took_near and 0 < keys < n // 10


@pytest.mark.parametrize("n", [1 << 20, 3 << 19])
def test_near_sorted_sort_many_moved_keys(bh, n):
    """A million Kuiper-belt bodies: tens of thousands of moved keys, many
    radix tiles chained by look-back.  2^20 keys take the 512-key mark tiles,
    1.5M the 2048-key tiles with the separate tile-count scan."""
    say("near vs full sort, many moved keys", n)
    pos, vel, m = bh.generate(bh.GEN_KUIPER, 6, n)
    near, _, last = run(bh, bh.SORT_NEAR, bh.FORM_MIXED, pos, vel, m, 3, theta=0.7)
    full, _, _ = run(bh, bh.SORT_FULL, bh.FORM_MIXED, pos, vel, m, 3, theta=0.7)
    same(near, full)
    say("last step", last)
    assert last[0] and 12288 < last[1] < n // 2


def test_near_sorted_sort_ensemble(bh):
    sizes = [3000, 40000, 9, 25000, 1, 700]
    kinds = [bh.GEN_BELT, bh.GEN_KUIPER, bh.GEN_TROJAN]
    parts = []
    for k, s in enumerate(sizes):
        if s >= 9:
            parts.append(bh.generate(kinds[k % 3], 20 + k, s))
        else:
            p, v, mm = bh.generate(bh.GEN_BELT, 20 + k, 9)
            parts.append((p[:s].copy(), v[:s].copy(), mm[:s].copy()))
    pos = np.concatenate([p[0] for p in parts])
    vel = np.concatenate([p[1] for p in parts])
    m = np.concatenate([p[2] for p in parts])
    say("near vs full sort, ensemble", sizes)
    near, _, last = run(bh, bh.SORT_NEAR, bh.FORM_MIXED, pos, vel, m, 5, systems=sizes)
    full, _, _ = run(bh, bh.SORT_FULL, bh.FORM_MIXED, pos, vel, m, 5, systems=sizes)
    same(near, full)
    assert last[0]


def test_near_sorted_sort_falls_back_on_disorder(bh):
    """Bodies handed over in sorted order except for a reversed block and a
    few far jumps: whichever path the sort takes, the order is the stable one
    (checked against the full sort), eager and graph-replayed."""
    n = 60000
    pos, vel, m = bh.generate(bh.GEN_TROJAN, 9, n)
    with bh.Engine(form=bh.FORM_MIXED) as e:
        e.set_bodies(pos, vel, m)
        e.compute_accelerations(form=bh.FORM_MIXED)
        st = e.get_bodies()   # sorted order
    ps, vs, ms = st["pos"].copy(), st["vel"].copy(), st["mass"].copy()
    ps[20000:20500] = ps[20000:20500][::-1].copy()
    vs[20000:20500] = vs[20000:20500][::-1].copy()
    ms[20000:20500] = ms[20000:20500][::-1].copy()
    for a, b in ((100, 50000), (30000, 7), (59000, 123)):
        for arr in (ps, vs):
            arr[[a, b]] = arr[[b, a]]
        ms[[a, b]] = ms[[b, a]]
    for graph in (False, True):
        say("disorder case, graph", graph)
        near, first, _ = run(bh, bh.SORT_NEAR, bh.FORM_MIXED, ps, vs, ms, 4, graph=graph)
        full, _, _ = run(bh, bh.SORT_FULL, bh.FORM_MIXED, ps, vs, ms, 4, graph=graph)
        same(near, full)
        say("first sort near", first)
