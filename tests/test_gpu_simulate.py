"""GPU tests of the command-line driver (the reference's simulate interface):
a reference run (-r) reproduces the oracle's trajectory bit for bit, a second
run measures SUMMED_DIST_ERROR 0 against it, the snapshots are valid VTK XML
at the reference's snapshot times, and the mixed form lands close."""
import pathlib
import subprocess
import sys
import xml.etree.ElementTree as ET

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
BIN = pathlib.Path(__file__).resolve().parent.parent / "barnes-hut_amd" / "bh_simulate"


def say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def run(cwd, *args):
    r = subprocess.run([str(BIN), *args], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


def field(out, name):
    line = [ln for ln in out.splitlines() if ln.startswith(name)][-1]
    return float(line.split()[-1])


def test_simulate_reference_run_matches_oracle(bh, orc, tmp_path):
    n = 3000
    pos, vel, m = bh.generate(bh.GEN_BELT, 21, n)
    csv = tmp_path / "belt.csv"
    bh.write_csv(csv, pos, vel, m)
    say("simulate -r", n)
    out = run(tmp_path, "-f", str(csv), "--dt", "1h", "--tend", "10h", "--vs", "3h", "-o", "snap", "-r", "--fc", "tree")
    assert "Loaded 3000 bodies" in out and "Reference positions saved" in out
    ref = np.loadtxt(tmp_path / "reference" / "final_ref.csv", delimiter=",", skiprows=1)
    ids = ref[:, 0].astype(np.int64)
    # the oracle from the same CSV bits
    rid, rm, rpos, rvel = bh.read_csv(csv)
    steps = bh.count_steps(1 / 24, 10 / 24)
    st = orc.simulate(rpos, rvel, rm, steps, 1 / 24)
    want = np.empty((n, 3))
    want[st["ids"].astype(np.int64)] = st["pos"]
    assert np.array_equal(ref[:, 1:].view(np.uint64), want[ids].view(np.uint64))
    # snapshots where src/main.cpp:284 takes them: after step k when t_k >= next
    # or k == 0, next += 3h, with t and next accumulated by repeated addition
    t, nxt, k, want_t = 0.0, 0.0, 0, []
    while t < 10 / 24:
        if t >= nxt or k == 0:
            want_t.append(t)
            nxt += 3 / 24
        k, t = k + 1, t + 1 / 24
    frames = sorted(p.name for p in (tmp_path / "snap" / "0").glob("sim.*.vtp"))
    assert frames == sorted(f"sim.{k}.vtp" for k in range(len(want_t)))
    pvd = list((tmp_path / "snap").glob("*.pvd"))
    assert len(pvd) == 1
    sets = ET.parse(pvd[0]).getroot().find("Collection").findall("DataSet")
    assert [s.get("file") for s in sets] == [f"0/sim.{k}.vtp" for k in range(len(want_t))]
    assert [float(s.get("timestep")) for s in sets] == want_t
    piece = ET.parse(tmp_path / "snap" / "0" / "sim.1.vtp").getroot().find("PolyData").find("Piece")
    assert piece.get("NumberOfPoints") == str(n)
    names = [d.get("Name") for d in piece.find("PointData").findall("DataArray")]
    assert names == ["rank", "body_id", "velocity", "mass", "acceleration"]
    # a second, normal run compares against the reference file
    out2 = run(tmp_path, "-f", str(csv), "--dt", "1h", "--tend", "10h", "--vs", "100d", "-o", "snap2")
    assert field(out2, "SUMMED_DIST_ERROR") == 0.0
    out3 = run(tmp_path, "-f", str(csv), "--dt", "1h", "--tend", "10h", "--vs", "100d", "-o", "snap3",
               "--form", "mixed")
    err = field(out3, "SUMMED_DIST_ERROR")   # printed with 4 decimals, as the reference does
    say("mixed summed distance error (AU)", err)
    assert 0.0 <= err < 1e-6 * n
