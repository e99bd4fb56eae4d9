"""Argument checks of the command-line drivers build/gpp.ex and build/ff.ex.  They run
before any array is allocated or any GPU call is made, so these tests need no GPU:
positional sizes must be whole decimal integers in range (the reference's atoi would
read "12abc" as 12 and overflow on large values), a seed a non-negative integer, and
a run whose host arrays exceed the host's physical memory is refused before it
allocates."""
import subprocess
from pathlib import Path

import pytest

import bgw

BUILD = Path(bgw.__file__).resolve().parents[1] / "build"


def _run(exe, *args):
    p = subprocess.run([str(BUILD / exe), *map(str, args)], capture_output=True, text=True, timeout=60)
    return p.returncode, p.stdout + p.stderr


@pytest.mark.parametrize("args", [
    ("12abc", 2, 8, 1),                # trailing characters
    ("", 2, 8, 1),                     # empty
    (99999999999, 2, 8, 1),            # beyond int
    (8, 2, 8, "1e3"),                  # not a decimal integer
    (8, 2, 8, 1, "--random", -3),      # negative seed
    (8, 2, 8, 1, "--random", "7x"),
])
def test_gpp_driver_refuses_malformed_numbers(args):
    rc, out = _run("gpp.ex", *args)
    assert rc == 1 and "correct form" in out, out


@pytest.mark.parametrize("args", [
    (8, 2, 10, 5, 15, "x"),
    (99999999999, 2, 10, 5, 15, 10),
    (8, 2, 10, 5, 15, 10, "--random", "zz"),
])
def test_ff_driver_refuses_malformed_numbers(args):
    rc, out = _run("ff.ex", *args)
    assert rc == 1 and "ff.ex number_bands nvband" in out, out  # the usage text


def test_drivers_refuse_runs_larger_than_host_memory():
    rc, out = _run("gpp.ex", 2147483647, 2, 2147483647, 1)  # ~2.7e11 GB of arrays
    assert rc == 1 and "more than the host's" in out, out
    rc, out = _run("ff.ex", 2000000000, 2, 2000000000, 1, 15, 10)
    assert rc == 1 and "more than the host's" in out, out
