"""CPU tests that pin the oracle (oracle/bh_oracle.c): known answers, closed
forms, brute force on tiny inputs, the golden fixtures written from the
reference's own code (tests/golden/), and the live reference build where it
exists (oracle/_ref, development container only)."""
import hashlib
import pathlib

import numpy as np
import pytest

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
G = 1.48812e-34
HOUR = 1.0 / 24.0


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def tree_digest(prefix, depth, mass, com):
    h = hashlib.sha256()
    for a in (np.ascontiguousarray(prefix, "<u8"), np.ascontiguousarray(depth, "u1"),
              np.ascontiguousarray(mass, "<f8"), np.ascontiguousarray(com, "<f8")):
        h.update(a.tobytes())
    return h.hexdigest()


# ------------------------------------------------------------- Morton keys

def test_morton_known_answers(orc):
    L = orc.lib()
    assert L.or_morton63(1, 0, 0) == 1
    assert L.or_morton63(0, 1, 0) == 2
    assert L.or_morton63(0, 0, 1) == 4
    assert L.or_morton63(2, 0, 0) == 8
    full = (1 << 21) - 1
    assert L.or_morton63(full, full, full) == (1 << 63) - 1
    assert L.or_prefix(0x7FFF_FFFF_FFFF_FFFF, 0) == 0
    assert L.or_prefix(0x7FFF_FFFF_FFFF_FFFF, 21) == 0x7FFF_FFFF_FFFF_FFFF
    assert L.or_prefix(0x7FFF_FFFF_FFFF_FFFF, 1) == 0x7 << 60


def test_spread_compact_round_trip(orc):
    L = orc.lib()
    rng = np.random.default_rng(0)
    for v in list(rng.integers(0, 1 << 21, 200)) + [0, 1, (1 << 21) - 1]:
        s = L.or_spread21(int(v))
        assert s & ~0x1249249249249249 == 0
        assert L.or_compact21(s) == v


def test_keys_box_corners_and_degenerate_axis(orc):
    pos = np.array([[0.0, 5.0, 7.0], [2.0, 5.0, 9.0], [1.0, 5.0, 8.0]])
    bb = orc.bbox(pos)
    assert list(bb) == [0.0, 5.0, 7.0, 2.0, 5.0, 9.0]
    k = orc.morton_keys(pos, bb)
    # min corner -> 0; max corner clamps to nextafter(1,0) -> q = 2^21-1 on x and z,
    # y is degenerate (span 1.0, t = 0); centre -> q = 2^20 on x and z
    L = orc.lib()
    assert k[0] == 0
    assert k[1] == L.or_morton63((1 << 21) - 1, 0, (1 << 21) - 1)
    assert k[2] == L.or_morton63(1 << 20, 0, 1 << 20)


def test_keys_match_independent_numpy(orc):
    rng = np.random.default_rng(1)
    pos = rng.normal(size=(5000, 3)) * np.array([3.0, 3.0, 0.2])
    bb = orc.bbox(pos)
    k = orc.morton_keys(pos, bb)
    lo, hi = bb[:3], bb[3:]
    span = np.where(hi == lo, 1.0, hi - lo)
    t = (pos - lo) / span
    t = np.where(t <= 0, 0.0, np.where(t >= 1.0, np.nextafter(1.0, 0.0), t))
    q = (t * 2097152.0).astype(np.uint64)
    ref = np.zeros(len(pos), np.uint64)
    for b in range(21):
        for a in range(3):
            ref |= ((q[:, a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + a)
    assert np.array_equal(k, ref)


# ----------------------------------------------------------------- sort

def test_sort_is_stable_argsort(orc):
    rng = np.random.default_rng(2)
    n = 3000
    pos = rng.normal(size=(n, 3))
    mass = rng.uniform(1, 2, n)
    keys = rng.integers(0, 50, n).astype(np.uint64)   # many ties
    k = keys.copy()
    p = pos.copy()
    m = mass.copy()
    ids = np.arange(n, dtype=np.uint64)
    perm = np.zeros(n, np.int64)
    import ctypes as C
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    orc.lib().or_sort(n, P(k), P(p), P(m), None, P(ids), P(perm))
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(perm, ref)
    assert np.array_equal(k, keys[ref]) and np.array_equal(p, pos[ref]) and np.array_equal(ids, ref.astype(np.uint64))


# ----------------------------------------------------------------- tree

def test_topology_closed_form(orc, bh):
    pos, vel, m = bh.generate(bh.GEN_BELT, 5, 2500)
    r = orc.force_path(pos, m, tree=True)
    prefix, depth, mass, com = r["tree"]
    keys = r["keys"]
    want_p, want_d = [], []
    for d in range(22):
        pre = keys & (np.uint64(0) if d == 0 else np.uint64(((1 << 64) - 1) << (63 - 3 * d) & ((1 << 64) - 1)))
        u = np.unique(pre)
        want_p.append(u)
        want_d.append(np.full(len(u), d, np.uint8))
    assert np.array_equal(prefix, np.concatenate(want_p))
    assert np.array_equal(depth, np.concatenate(want_d))
    # root holds the total mass; every depth holds it too
    for d in range(22):
        assert np.isclose(mass[depth == d].sum(), m.sum(), rtol=1e-14)


def test_com_brute_force_small(orc):
    rng = np.random.default_rng(3)
    pos = rng.uniform(-1, 1, (7, 3))
    mass = rng.uniform(1e20, 1e22, 7)
    r = orc.force_path(pos, mass, tree=True)
    prefix, depth, tm, com = r["tree"]
    keys, sp, sm = r["keys"], r["pos"], r["mass"]
    for k in range(len(prefix)):
        d = int(depth[k])
        mask = np.uint64(0) if d == 0 else np.uint64((((1 << 64) - 1) << (63 - 3 * d)) & ((1 << 64) - 1))
        sel = (keys & mask) == prefix[k]
        mm = sm[sel].sum()
        c = (sm[sel, None] * sp[sel]).sum(0) / mm
        assert np.isclose(tm[k], mm, rtol=1e-15)
        assert np.allclose(com[k], c, rtol=1e-14, atol=0)
    # a single-body cell keeps that body's mass exactly
    leaves = depth == 21
    assert np.array_equal(np.sort(tm[leaves]), np.sort(sm))


# ----------------------------------------------------------------- walk

def test_walk_theta0_is_direct_sum_over_other_leaves(orc, bh):
    pos, vel, m = bh.generate(bh.GEN_BELT, 4, 800)
    r = orc.force_path(pos, m, theta=0.0)
    ds = orc.direct(r["pos"], r["mass"], np.arange(800))
    err = np.linalg.norm(r["acc"] - ds, axis=1) / np.linalg.norm(ds, axis=1)
    assert err.max() < 1e-12


def test_walk_huge_theta_is_root_monopole(orc, bh):
    pos, vel, m = bh.generate(bh.GEN_BELT, 4, 500)
    r = orc.force_path(pos, m, theta=1e6, tree=True)
    prefix, depth, tm, com = r["tree"]
    root = np.where(depth == 0)[0][0]
    d = com[root] - r["pos"]
    r2 = (d * d).sum(1)
    want = (G * tm[root]) * (1.0 / np.sqrt(r2))[:, None] ** 3 * d
    assert np.allclose(r["acc"], want, rtol=1e-13, atol=0)


def test_walk_two_body_exact(orc):
    pos = np.array([[0.1, 0.2, -0.3], [1.7, -0.4, 0.25]])
    mass = np.array([2.0e30, 3.0e24])
    # theta = 0.5 opens the root for both bodies (at 1.05 the light body
    # would accept the root monopole, which includes its own mass)
    r = orc.force_path(pos, mass, theta=0.5)
    sp, sm = r["pos"], r["mass"]
    for i in range(2):
        j = 1 - i
        # the other body's leaf: raw = 0 + x m, com = raw / m, then the
        # one-rank exchange round trip (0*0 + m com) / (0 + m)
        com = (0.0 + sp[j] * sm[j]) / sm[j]
        com = (0.0 * 0.0 + sm[j] * com) / (0.0 + sm[j])
        dx = com - sp[i]
        r2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2] + 0.0
        inv = 1.0 / np.sqrt(r2)
        s = G * sm[j] * (inv * inv * inv)
        assert np.array_equal(r["acc"][i], s * dx)
        # and it differs from using the raw position by at most a few ulps
        assert np.allclose(r["acc"][i], s * (sp[j] - sp[i]), rtol=1e-14)


def test_walk_two_body_softened(orc):
    """soft2 is added to r^2 once (src/traversal.cpp:75): with eps^2 = 0.25
    the two-body force is G m d / (|d|^2 + eps^2)^(3/2), and the direct sum
    agrees to rounding."""
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    mass = np.array([2.0e30, 3.0e24])
    eps2 = 0.25
    r = orc.force_path(pos, mass, theta=0.5, soft2=eps2)
    sp, sm = r["pos"], r["mass"]
    for i in range(2):
        j = 1 - i
        d = sp[j] - sp[i]
        want = G * sm[j] * d / (d @ d + eps2) ** 1.5
        assert np.allclose(r["acc"][i], want, rtol=1e-14, atol=0)
        assert not np.allclose(r["acc"][i], G * sm[j] * d / (d @ d) ** 1.5, rtol=1e-3, atol=0)
    ds = orc.direct(sp, sm, np.arange(2), soft2=eps2)
    assert np.allclose(ds, r["acc"], rtol=1e-14, atol=0)


def test_direct_three_body_line(orc):
    pos = np.array([[0.0, 0, 0], [1.0, 0, 0], [3.0, 0, 0]])
    mass = np.array([1e30, 2e30, 4e30])
    a = orc.direct(pos, mass, np.arange(3))
    want0 = G * (2e30 / 1.0 + 4e30 / 9.0)
    want1 = G * (-1e30 / 1.0 + 4e30 / 4.0)
    want2 = -G * (1e30 / 9.0 + 2e30 / 4.0)
    assert np.allclose(a[:, 0], [want0, want1, want2], rtol=1e-15)
    assert np.all(a[:, 1:] == 0)


# ------------------------------------------------------------ step loop

def test_count_steps_trap(orc):
    L = orc.lib()
    assert L.or_count_steps(HOUR, 100 * HOUR) == 101
    assert L.or_count_steps(HOUR, 200 * HOUR) == 200
    assert L.or_count_steps(HOUR, 168 * HOUR) == 168
    assert L.or_count_steps(HOUR, 48 * HOUR) == 48


def test_one_step_closed_form(orc):
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    vel = np.array([[0.0, 0.0, 0.0], [0.0, 0.0172, 0.0]])
    mass = np.array([1.989e30, 5.97e24])
    dt = HOUR
    st = orc.simulate(pos, vel, mass, 1, dt, theta=0.5)
    # step 0: no first kick (t == 0), drift, force, half kick
    x = pos + dt * vel
    order = np.argsort(orc.morton_keys(x, orc.bbox(x)), kind="stable")
    x, v, mm = x[order], vel[order], mass[order]
    a = np.zeros((2, 3))
    for i in range(2):
        j = 1 - i
        com = (0.0 + x[j] * mm[j]) / mm[j]
        com = (0.0 * 0.0 + mm[j] * com) / (0.0 + mm[j])
        d = com - x[i]
        r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + 0.0
        inv = 1.0 / np.sqrt(r2)
        a[i] = (G * mm[j] * (inv * inv * inv)) * d
    v = v + 0.5 * dt * a
    assert np.array_equal(st["pos"], x) and np.array_equal(st["vel"], v) and np.array_equal(st["acc"], a)


# ------------------------------------------- golden fixtures (reference)

FORCE_FIXTURES = ["belt_600", "belt_4000", "kuiper_3000", "trojan_3000", "scenario1_3000", "dups_12"]


@pytest.mark.parametrize("name", FORCE_FIXTURES)
def test_oracle_matches_reference_fixture(orc, name):
    f = np.load(GOLD / f"{name}.npz")
    for k, th in enumerate(f["thetas"]):
        r = orc.force_path(f["pos"], f["mass"], theta=float(th), tree=(k == 0))
        if k == 0:
            assert np.array_equal(r["keys"], f["keys"])
            assert np.array_equal(r["perm"], f["perm"])
            assert np.array_equal(r["bb"], f["bb"])
            prefix, depth, tm, com = r["tree"]
            assert len(prefix) == int(f["tree_nodes"])
            assert np.array_equal(np.bincount(depth, minlength=22), f["tree_depth_counts"])
            assert tree_digest(*r["tree"]) == str(f["tree_sha256"])
            if "tree_prefix" in f:
                assert np.array_equal(prefix, f["tree_prefix"]) and np.array_equal(bits(com), bits(f["tree_com"]))
        assert np.array_equal(bits(r["acc"]), bits(f[f"acc_{k}"])), f"theta={th}"


def test_oracle_matches_reference_trajectory(orc):
    f = np.load(GOLD / "traj_belt_1500.npz")
    st = orc.simulate(f["pos0"], f["vel0"], f["mass0"], int(f["nsteps"]), float(f["dt"]), theta=float(f["theta"]))
    for k in ("pos", "vel", "acc", "mass"):
        assert np.array_equal(bits(st[k]), bits(f[k])), k
    assert np.array_equal(st["ids"], f["ids"])


# -------------------------------------- live reference (dev container)

@pytest.mark.parametrize("seed,n,theta", [(11, 3000, 1.05), (12, 2000, 0.5), (13, 1000, 0.0)])
def test_oracle_matches_live_reference(orc, bh, seed, n, theta):
    if not orc.ref_available("strict"):
        pytest.skip("oracle/_ref not built here (needs the reference checkout)")
    pos, vel, m = bh.generate(bh.GEN_TROJAN, seed, n)
    a = orc.force_path(pos, m, theta=theta, tree=True)
    b = orc.ref_force_path(pos, m, theta=theta, tree=True)
    assert np.array_equal(a["keys"], b["keys"]) and np.array_equal(a["perm"], b["perm"])
    assert tree_digest(*a["tree"]) == tree_digest(*b["tree"])
    assert np.array_equal(bits(a["acc"]), bits(b["acc"]))


def test_oracle_matches_live_reference_softened(orc, bh):
    if not orc.ref_available("strict"):
        pytest.skip("oracle/_ref not built here (needs the reference checkout)")
    pos, vel, m = bh.generate(bh.GEN_BELT, 14, 2000)
    a = orc.force_path(pos, m, theta=1.05, soft2=1e-6)
    b = orc.ref_force_path(pos, m, theta=1.05, soft2=1e-6)
    assert np.array_equal(bits(a["acc"]), bits(b["acc"]))
    c = orc.force_path(pos, m, theta=1.05)
    assert not np.array_equal(bits(a["acc"]), bits(c["acc"]))


def test_oracle_counts_are_plausible(orc, bh):
    pos, vel, m = bh.generate(bh.GEN_BELT, 0, 3000)
    r = orc.force_path(pos, m, theta=1.05, counts=True)
    # an outlying body may accept the root monopole at once (no opened node)
    assert r["n_accept"].min() >= 1 and r["n_open"].mean() > 10
    r0 = orc.force_path(pos, m, theta=0.0, counts=True)
    assert np.all(r0["n_accept"] == 2999)   # every other leaf, no duplicates here
