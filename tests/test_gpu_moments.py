"""GPU tests of the tree moments' deep launch.

The moments are summed bottom-up, one launch per depth, except that the
branching nodes at depth >= a cutoff are summed by one workgroup, depth by
depth, in a single launch.  Every cutoff must give the same bits: 21 (one
launch per depth), the automatic choice made at graph capture, cutoffs that
leave the deep list in LDS, and cutoffs whose list is larger than LDS holds,
in the plain form (bit-exact with the reference) and the mixed form."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HOUR = 1.0 / 24.0


def say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("form_name", ["plain", "mixed"])
def test_moment_cutoffs_give_the_same_bits(bh, form_name):
    form = bh.FORM_PLAIN if form_name == "plain" else bh.FORM_MIXED
    n = 60000 if form_name == "plain" else 300149
    # Trojan clouds: dense clusters make deep trees
    pos, vel, m = bh.generate(bh.GEN_TROJAN, 3, n)
    ref = None
    for cut in (21, -1, 12, 6, 1, 0):
        say("moment cutoff", form_name, n, cut)
        with bh.Engine(form=form, theta=1.05) as e:
            e.set_moment_cutoff(cut)
            e.set_bodies(pos, vel, m)
            e.step(4, HOUR)   # the first step eager, the rest graph-replayed
            st = e.get_bodies()
            assert e.graph_replays() >= 2
        if ref is None:
            ref = st
            continue
        for k in ("pos", "vel", "acc"):
            assert np.array_equal(bits(st[k]), bits(ref[k])), (cut, k)
        assert np.array_equal(st["ids"], ref["ids"])


def test_moment_cutoff_rejects_out_of_range(bh):
    with bh.Engine(form=bh.FORM_MIXED) as e:
        for bad in (-2, 22):
            with pytest.raises(bh.BHError):
                e.set_moment_cutoff(bad)
