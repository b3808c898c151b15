"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

The generator and CSV code of the library (barnes-hut_amd/host/bh_gen.cpp)
and the oracle's C restatement (oracle/bh_oracle.c) are compiled with
-fsanitize=address,undefined and run by small drivers under tests/sanitize/.
A sanitizer finding aborts the driver.  The sanitized oracle must also give
the same bits as the oracle library itself, so the checks run on the code
the parity tests rely on."""
import os
import pathlib
import shutil
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="needs gcc and g++")


def _build(tmp_path, name, cmd):
    exe = tmp_path / name
    r = subprocess.run(cmd + ["-o", str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return exe


def test_generator_and_csv_under_sanitizers(tmp_path):
    exe = _build(tmp_path, "gen_driver",
                 ["g++", "-std=c++17", "-ffp-contract=off", *SAN, str(ROOT / "tests/sanitize/gen_driver.cpp"),
                  str(ROOT / "barnes-hut_amd/host/bh_gen.cpp")])
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=300, env=ENV)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"


def test_oracle_step_loop_under_sanitizers(tmp_path):
    import sys
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    exe = _build(tmp_path, "oracle_driver",
                 ["gcc", "-std=c11", "-ffp-contract=off", "-fno-fast-math", *SAN,
                  str(ROOT / "tests/sanitize/oracle_driver.c"), str(ROOT / "oracle/bh_oracle.c"), "-lm"])
    rng = np.random.default_rng(11)
    n, steps, dt, theta = 700, 3, 1.0 / 24.0, 1.05
    pos = rng.normal(size=(n, 3)) * [3.0, 3.0, 0.3]
    vel = rng.normal(size=(n, 3)) * 1e-3
    mass = rng.uniform(1e-12, 1e-9, n)
    mass[0] = 1.0
    pos[5] = pos[6]   # a co-located pair: one key, one leaf of two bodies
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        np.array([n, steps], np.int64).tofile(f)
        np.array([dt, theta, orc.G], np.float64).tofile(f)
        pos.astype(np.float64).tofile(f)
        vel.astype(np.float64).tofile(f)
        mass.astype(np.float64).tofile(f)
    out = tmp_path / "out.bin"
    r = subprocess.run([str(exe), str(inp), str(out), "ok"], capture_output=True, text=True, timeout=300, env=ENV)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, np.float64, count=9 * n)
    ids = np.fromfile(out, np.uint64, offset=8 * 9 * n)
    want = orc.simulate(pos, vel, mass, steps, dt, theta=theta)
    assert np.array_equal(ids, want["ids"])
    for k, off in (("pos", 0), ("vel", 3 * n), ("acc", 6 * n)):
        got = raw[off:off + 3 * n].reshape(n, 3)
        assert np.array_equal(got.view(np.uint64), np.ascontiguousarray(want[k]).view(np.uint64)), k
