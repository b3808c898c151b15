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


def read_vtp(path):
    """Arrays of one snapshot by name: floats parsed from their 17-digit text
    (exact FP64), integers as int64 / uint64, vectors as (n, 3)."""
    piece = ET.parse(path).getroot().find("PolyData").find("Piece")
    out = {}
    arrays = list(piece.find("Points").findall("DataArray")) + list(piece.find("PointData").findall("DataArray")) + \
        list(piece.find("Verts").findall("DataArray"))
    for da in arrays:
        words = da.text.split()
        typ = da.get("type")
        if typ == "Float64":
            v = np.array([float(w) for w in words])
        elif typ == "UInt64":
            v = np.array([int(w) for w in words], np.uint64)
        else:
            v = np.array([int(w) for w in words], np.int64)
        comps = int(da.get("NumberOfComponents", "1"))
        out[da.get("Name")] = v.reshape(-1, 3) if comps == 3 else v
    return out


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
    # frame contents (src/io.cpp:84-232): every frame's positions, ids,
    # velocities, masses and |a| equal the oracle's state after the step that
    # wrote it, bit for bit (17 significant digits round-trip FP64)
    snap_steps, t, nxt, k = [], 0.0, 0.0, 0
    while t < 10 / 24:
        if t >= nxt or k == 0:
            snap_steps.append(k)
            nxt += 3 / 24
        k, t = k + 1, t + 1 / 24
    for f, k in enumerate(snap_steps):
        say("frame", f, "after step", k)
        fr = read_vtp(tmp_path / "snap" / "0" / f"sim.{f}.vtp")
        o = orc.simulate(rpos, rvel, rm, k + 1, 1 / 24, ids=rid)
        assert np.array_equal(fr["body_id"], o["ids"])
        assert np.array_equal(fr["rank"], np.zeros(n, np.int64))
        assert np.array_equal(fr["position"].view(np.uint64), o["pos"].view(np.uint64))
        assert np.array_equal(fr["velocity"].view(np.uint64), o["vel"].view(np.uint64))
        assert np.array_equal(fr["mass"].view(np.uint64), o["mass"].view(np.uint64))
        a = o["acc"]
        amag = np.sqrt(a[:, 0] * a[:, 0] + a[:, 1] * a[:, 1] + a[:, 2] * a[:, 2])
        assert np.array_equal(fr["acceleration"].view(np.uint64), amag.view(np.uint64))
        assert np.array_equal(fr["offsets"], np.arange(1, n + 1)) and np.array_equal(fr["connectivity"], np.arange(n))
    # a second, normal run compares against the reference file
    out2 = run(tmp_path, "-f", str(csv), "--dt", "1h", "--tend", "10h", "--vs", "100d", "-o", "snap2")
    assert field(out2, "SUMMED_DIST_ERROR") == 0.0
    out3 = run(tmp_path, "-f", str(csv), "--dt", "1h", "--tend", "10h", "--vs", "100d", "-o", "snap3",
               "--form", "mixed")
    err = field(out3, "SUMMED_DIST_ERROR")   # printed with 4 decimals, as the reference does
    say("mixed summed distance error (AU)", err)
    assert 0.0 <= err < 1e-6 * n
