"""The command-line driver's interface without a GPU: help text, refused
force policies and bad options exit before any device call."""
import pathlib
import subprocess
import sys

import pytest

BIN = pathlib.Path(__file__).resolve().parent.parent / "barnes-hut_amd" / "bh_simulate"
pytestmark = pytest.mark.skipif(not BIN.exists(), reason="bh_simulate not built (make -C barnes-hut_amd)")


def run(*args):
    return subprocess.run([str(BIN), *args], capture_output=True, text=True, timeout=60)


def test_help_lists_the_reference_options():
    r = run("--help")
    assert r.returncode == 0
    for opt in ("--file", "--dt", "--tend", "--vs", "--outdir", "--bodies", "--fc", "--theta", "--reference"):
        assert opt in r.stdout


@pytest.mark.parametrize("fc", ["let", "let_direct"])
def test_let_policies_are_refused(fc):
    r = run("--fc", fc, "-f", "x.csv")
    assert r.returncode == 2 and "only the full tree" in r.stderr


def test_bad_options():
    assert run("--bogus").returncode == 2
    assert run("--dt").returncode == 2
    assert run("--form", "fast", "-f", "x.csv").returncode == 2
    assert run("--dt", "3w", "-f", "x.csv").returncode == 1
    assert run().returncode == 1   # no input file


def test_bench_refuses_several_gpus():
    """bench.py runs on one GPU: --gpus above 1 exits with status 2 and says so,
    before touching any device."""
    root = pathlib.Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=root)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr


def test_integration_sketch_is_the_compiled_dropin():
    """The per-call loop printed in INTEGRATION.md is the part of
    barnes-hut_amd/host/bh_dropin.cpp between its sketch markers, verbatim (up
    to the first indent), so the documented integration is the one that
    compiles and that the GPU tests run."""
    root = pathlib.Path(__file__).resolve().parents[1]
    src = (root / "barnes-hut_amd" / "host" / "bh_dropin.cpp").read_text()
    body = src.split("// [sketch begin]\n", 1)[1].split("    // [sketch end]", 1)[0]
    want = [ln[4:] if ln.startswith("    ") else ln for ln in body.rstrip().splitlines()]
    doc = (root / "INTEGRATION.md").read_text()
    sect = doc.split("## 1. Drop-in", 1)[1].split("## 2.", 1)[0]
    block = sect.split("```cpp\n", 1)[1].split("```", 1)[0]
    got = block.rstrip().splitlines()
    assert got[:2] == ['#include "bh.h"', ""]
    assert got[2:] == want
    # every call the loop makes per step, in the order that keeps the host
    # arrays in the engine's Morton order
    loop = "\n".join(want)
    assert loop.index("bh_set_bodies") < loop.index("bh_compute_accelerations") < loop.index("bh_get_bodies")
