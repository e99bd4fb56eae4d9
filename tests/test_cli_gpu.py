"""The reference-compatible command-line drivers build/gpp.ex and build/ff.ex
(the reference's argv and print format, GPP/gpp.cpp:133-305 and FF/ff.cpp:453-764).
gpp.ex calls noflagOCC_solver through the reference-signature C++ shim
(host/bgw_reference_api.hpp) and ff.ex calls the three FF kernels the same way, so
these tests also cover that shim.  The drivers check the reference constants against
exact closed forms; with --random SEED they generate the same inputs as bgw.gpp_inputs /
bgw.ff_inputs, which the tests compare with the oracle.  --dump FILE and --input FILE
write and read input sets (bgw_inputs_write / bgw_inputs_open)."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import bgw
import oracle as O

pytestmark = pytest.mark.gpu
BUILD = Path(bgw.__file__).resolve().parents[1] / "build"
NUM = r"(-?[0-9.]+(?:e[-+]?[0-9]+)?)"


def _run(exe, *args):
    p = subprocess.run([str(BUILD / exe), *map(str, args)], capture_output=True, text=True, timeout=240)
    return p.returncode, p.stdout


def _pair(out, label):
    m = re.search(re.escape(label) + r"\s*=?\s*\(" + NUM + "," + NUM + r"\)", out)
    assert m, f"{label} not printed:\n{out}"
    return complex(float(m.group(1)), float(m.group(2)))


@pytest.mark.parametrize("precision", ["fp64", "fp32", "moments"])
def test_gpp_driver_reference_constants(precision):
    rc, out = _run("gpp.ex", 64, 2, 4096, 8, "--precision", precision)
    assert rc == 0
    assert "SUCCESS" in out and "FAILURE" not in out, out
    # achtemp[0] = nb*ngpown*ncouls * 0.125i / (3 - 0.5 - 0.5i) for the constant inputs
    cf = 64 * 512 * 4096 * 0.125j / (3 - 0.5 - 0.5j)
    assert abs(_pair(out, "Final achtemp\n") - cf) <= 1e-6 * abs(cf)


@pytest.mark.parametrize("form", [[], ["--precision", "moments"]])
def test_gpp_driver_random_inputs_vs_oracle(form):
    nb, nvband, ncouls, nodes = 24, 2, 800, 100  # ngpown = 8
    rc, out = _run("gpp.ex", nb, nvband, ncouls, nodes, "--random", 4321, "--all", *form)
    assert rc == 0
    d = bgw.gpp_inputs(nb, nvband, ncouls, ncouls // nodes, 4321)
    o = O.gpp_forward(d)
    # printed with six decimals
    tol = lambda ref: 1e-6 + 1e-9 * abs(ref)  # noqa: E731
    assert abs(_pair(out, "Final achtemp\n") - o["achtemp"][0]) <= tol(o["achtemp"][0])
    assert abs(_pair(out, "sum_n1 acht_n1_loc[n1][0]") - o["achtemp"][0]) <= tol(o["achtemp"][0])
    assert abs(_pair(out, "achstemp") - o["achstemp"][0]) <= tol(o["achstemp"][0])
    assert abs(_pair(out, "asxtemp[0]") - o["asxtemp"][0]) <= tol(o["asxtemp"][0])


def test_gpp_driver_usage_errors():
    assert _run("gpp.ex", 1, 2, 3)[0] == 1  # three positional arguments
    assert _run("gpp.ex", 8, 9, 100, 4)[0] == 1  # nvband > number_bands
    assert _run("gpp.ex", 8, 2, 10, 20)[0] == 1  # ngpown = 0
    assert _run("gpp.ex", 8, 2, 64, 4, "--precision", "fp16")[0] == 1  # unknown precision form
    assert _run("ff.ex", 10, 2, 20, 5, 5, 2, "--precision", "fp16")[0] == 1


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ff_driver_reference_constants(precision):
    rc, out = _run("ff.ex", 60, 12, 600, 30, 15, 4, "--precision", precision)
    assert rc == 0
    assert out.count("SUCCESS") == 3 and "FAILURE" not in out, out


def test_ff_driver_random_inputs_vs_oracle():
    shape = (40, 3, 300, 20, 15, 2)
    rc, out = _run("ff.ex", *shape, "--random", 777)
    assert rc == 0
    f = bgw.ff_inputs(*shape, 777)
    for label, ref in (("achsDtemp", O.achsDtemp(f)), ("asxDtemp", O.asxDtemp(f)[0]),
                       ("achDtemp_cor", O.achDtemp_cor(f)[0])):
        assert abs(_pair(out, label + " ") - ref) <= 1e-6 + 1e-9 * abs(ref), label


def test_ff_driver_rejects_ngpown_above_ncouls():
    rc, out = _run("ff.ex", 10, 2, 20, 30, 5, 2)
    assert rc == 1 and "ngpown cannot be greater than ncouls" in out
    for bad in ((0, 0, 0, 0, 0, 0), (10, 2, 20, 5, 1, 2), (10, 12, 20, 5, 5, 2)):  # sizes, nFreq < 2, nvband
        assert _run("ff.ex", *bad)[0] == 1, bad


def test_gpp_driver_dump_then_input_reproduces_random_run(tmp_path):
    f = tmp_path / "gpp.bgwin"
    rc, out1 = _run("gpp.ex", 24, 2, 800, 100, "--random", 99, "--all", "--dump", f)
    assert rc == 0 and "inputs written" in out1
    rc, out2 = _run("gpp.ex", "--input", f, "--all")
    assert rc == 0, out2
    for label in ("Final achtemp\n", "sum_n1 acht_n1_loc[n1][0]", "achstemp", "asxtemp[0]"):
        assert _pair(out1, label) == _pair(out2, label), label
    d = bgw.load_inputs(f)
    o = O.gpp_forward(d)
    assert abs(_pair(out2, "Final achtemp\n") - o["achtemp"][0]) <= 1e-6 + 1e-9 * abs(o["achtemp"][0])


def test_gpp_driver_reads_the_constant_inputs_back(tmp_path):
    f = tmp_path / "c.bgwin"
    rc, out1 = _run("gpp.ex", 16, 2, 512, 8, "--dump", f)
    assert rc == 0 and "SUCCESS" in out1
    rc, out2 = _run("gpp.ex", "--input", f)
    assert rc == 0 and _pair(out1, "Final achtemp\n") == _pair(out2, "Final achtemp\n")


def test_ff_driver_dump_then_input_reproduces_random_run(tmp_path):
    f = tmp_path / "ff.bgwin"
    shape = (40, 3, 300, 20, 15, 2)
    rc, out1 = _run("ff.ex", *shape, "--random", 555, "--dump", f)
    assert rc == 0 and "inputs written" in out1
    rc, out2 = _run("ff.ex", "--input", f)
    assert rc == 0, out2
    d = bgw.load_inputs(f)
    for label, ref in (("achsDtemp", O.achsDtemp(d)), ("asxDtemp", O.asxDtemp(d)[0]),
                       ("achDtemp_cor", O.achDtemp_cor(d)[0])):
        assert _pair(out1, label + " ") == _pair(out2, label + " "), label
        assert abs(_pair(out2, label + " ") - ref) <= 1e-6 + 1e-9 * abs(ref), label


def test_drivers_refuse_bad_input_sets(tmp_path):
    rc, out = _run("gpp.ex", "--input", tmp_path / "missing.bgwin")
    assert rc == 1 and "cannot read input set" in out
    f = tmp_path / "ff.bgwin"
    bgw.save_inputs(f, bgw.ff_inputs(6, 2, 30, 4, 5, 3, 77), "ff")
    rc, out = _run("gpp.ex", "--input", f)
    assert rc == 1 and "not a GPP input set" in out
    rc, out = _run("gpp.ex", 8, 2, 64, 4, "--input", f)
    assert rc == 1
