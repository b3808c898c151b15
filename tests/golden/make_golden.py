"""Writes the golden fixtures of tests/golden/ from the reference's own force
path (oracle/_ref/libref_strict.so: src/morton_keys.cpp, src/linear_octree.cpp,
src/traversal.cpp of YonatanGM/Barnes-Hut compiled with strict flags, plus
the glue of oracle/ref_harness.cpp).

Run in the development container after `make -C oracle all ref` and
`make -C barnes-hut_amd` (the generator lives in libbh.so):

    python tests/golden/make_golden.py

Each fixture holds inputs and the reference's outputs only: Morton keys,
sort permutation, box, accelerations, and the node set of the tree (prefix,
depth, mass, com) either in full (small cases) or as a SHA-256 digest plus
per-depth node counts.  Inputs named `scenario1_*` are the first rows of the
reference's data file data/state_vectors_csvs/scenario1_19054.csv (mass,
position, velocity columns), which hold close moon-planet pairs.
"""
from __future__ import annotations

import hashlib
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))

import oracle  # noqa: E402
from conftest import load_bh  # noqa: E402

SCENARIO1 = pathlib.Path("/root/reference/data/state_vectors_csvs/scenario1_19054.csv")
HOUR = 1.0 / 24.0


def tree_digest(prefix, depth, mass, com):
    h = hashlib.sha256()
    for a in (np.ascontiguousarray(prefix, "<u8"), np.ascontiguousarray(depth, "u1"),
              np.ascontiguousarray(mass, "<f8"), np.ascontiguousarray(com, "<f8")):
        h.update(a.tobytes())
    return h.hexdigest()


def depth_counts(depth):
    return np.bincount(np.asarray(depth, np.int64), minlength=22)[:22]


def force_fixture(name, pos, vel, mass, thetas, full_tree):
    out = dict(pos=pos, vel=vel, mass=mass, thetas=np.array(thetas, np.float64))
    tree = None
    for k, th in enumerate(thetas):
        r = oracle.ref_force_path(pos, mass, theta=th, tree=True)
        if k == 0:
            out["keys"], out["perm"], out["bb"] = r["keys"], r["perm"], r["bb"]
            tree = r["tree"]
        else:
            assert np.array_equal(out["keys"], r["keys"]) and np.array_equal(out["perm"], r["perm"])
        out[f"acc_{k}"] = r["acc"]
    out["tree_sha256"] = np.array(tree_digest(*tree))
    out["tree_depth_counts"] = depth_counts(tree[1])
    out["tree_nodes"] = np.array(len(tree[0]))
    if full_tree:
        out["tree_prefix"], out["tree_depth"], out["tree_mass"], out["tree_com"] = tree
    np.savez_compressed(HERE / f"{name}.npz", **out)
    print(f"{name}: n={len(mass)} nodes={len(tree[0])} thetas={thetas}")


def trajectory_fixture(name, pos, vel, mass, nsteps, dt, theta):
    st = oracle.ref_simulate(pos, vel, mass, nsteps, dt, theta=theta)
    np.savez_compressed(HERE / f"{name}.npz", pos0=pos, vel0=vel, mass0=mass, nsteps=np.array(nsteps),
                        dt=np.array(dt), theta=np.array(theta), pos=st["pos"], vel=st["vel"], mass=st["mass"],
                        ids=st["ids"], acc=st["acc"])
    print(f"{name}: n={len(mass)} steps={nsteps}")


def main():
    if not oracle.ref_available("strict"):
        raise SystemExit("oracle/_ref/libref_strict.so is missing: make -C oracle ref")
    bh = load_bh()
    # synthetic heliocentric systems (generator of this repository)
    pos, vel, m = bh.generate(bh.GEN_BELT, 0, 600)
    force_fixture("belt_600", pos, vel, m, [0.0, 0.5, 1.05, 1.0e6], full_tree=True)
    pos, vel, m = bh.generate(bh.GEN_BELT, 0, 4000)
    force_fixture("belt_4000", pos, vel, m, [0.5, 1.05], full_tree=False)
    pos, vel, m = bh.generate(bh.GEN_KUIPER, 2, 3000)
    force_fixture("kuiper_3000", pos, vel, m, [0.7], full_tree=False)
    pos, vel, m = bh.generate(bh.GEN_TROJAN, 3, 3000)
    force_fixture("trojan_3000", pos, vel, m, [1.05], full_tree=False)
    # the reference's own data file: first 3000 bodies (Sun, planets, moons, ...)
    rows = np.genfromtxt(SCENARIO1, delimiter=",", skip_header=1, usecols=range(3, 10), max_rows=3000)
    force_fixture("scenario1_3000", np.ascontiguousarray(rows[:, 1:4]), np.ascontiguousarray(rows[:, 4:7]),
                  np.ascontiguousarray(rows[:, 0]), [0.5, 1.05], full_tree=False)
    # duplicate positions (shared leaves, self-leaf skip); n <= 16 keeps
    # std::sort on insertion sort, whose tie order is the stable order
    rng = np.random.default_rng(7)
    p = rng.uniform(-1, 1, (12, 3))
    p[5] = p[2]
    p[9] = p[2]
    p[11] = p[7]
    v = np.zeros((12, 3))
    mm = rng.uniform(1e20, 1e22, 12)
    force_fixture("dups_12", p, v, mm, [0.0, 1.05], full_tree=True)
    # trajectory: 24 KDK steps of dt = 1 h
    pos, vel, m = bh.generate(bh.GEN_BELT, 1, 1500)
    trajectory_fixture("traj_belt_1500", pos, vel, m, 24, HOUR, 1.05)


if __name__ == "__main__":
    main()
