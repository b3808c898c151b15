"""The command-line driver's interface without a GPU: help text, refused
force policies and bad options exit before any device call."""
import pathlib
import subprocess

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
