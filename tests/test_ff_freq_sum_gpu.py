"""GPU parity of the FF real-axis frequency sum (bgw_achDtemp_Kernel), the term the
reference prepares (its weights pref and pref_zb, FF/ff.cpp:468,564-565,619-626, and the
orphan achDtemp timers, :503) without applying it.  Its definition is this library's
(include/bgw.h); it is pinned by exact-rational fixtures, the closed form at the reference
driver's constants and the oracle.  Tolerance 1e-12 relative (FP64 MFMA), 1e-6 in the
FP32-operand form."""
import numpy as np
import pytest

import bgw
import oracle as O
from conftest import rel_err
from helpers import ff_reference_achDtemp, load_exact_ff

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.mark.parametrize("case", range(2))
def test_exact_fixture(ctx, case):
    f, exp = load_exact_ff()[case]
    got = ctx.achDtemp_Kernel(f)
    assert np.all(np.abs(got - exp["achDtemp"]) <= 1e-13 * abs(exp["achDtemp"]))


@pytest.mark.parametrize("nb,nvband,ncouls,ngpown,nFreq", [(20, 4, 64, 16, 15), (150, 40, 300, 70, 6),
                                                          (9, 9, 10, 10, 2)])
def test_closed_form_reference_constants(ctx, nb, nvband, ncouls, ngpown, nFreq):
    f = bgw.ff_inputs(nb, nvband, ncouls, ngpown, nFreq, 3, 0, kind="reference")
    ref = ff_reference_achDtemp(nb, nvband, ncouls, ngpown, nFreq)
    got = ctx.achDtemp_Kernel(f)
    assert np.all(np.abs(got - ref) <= TOL * abs(ref))


# row tiles of 64 bands: one partial tile, several tiles with a ragged last one, every
# band occupied; odd G' and G counts
@pytest.mark.parametrize("shape", [(40, 3, 96, 24, 15, 2), (130, 5, 1000, 65, 15, 3), (77, 77, 530, 33, 4, 1)])
def test_random_vs_oracle(ctx, shape):
    f = bgw.ff_inputs(*shape, 1900 + shape[0])
    got = ctx.achDtemp_Kernel(f)
    ref = O.achDtemp(f)
    assert rel_err(got, ref) < TOL
    assert np.all(got == got[0])  # wx has no iw term (FF/ff.cpp:245,337)


def test_device_form_equals_host_form(ctx):
    import torch

    f = bgw.ff_inputs(70, 6, 400, 30, 8, 2, 77)
    keys = ("inv", "indinv", "ekq", "aqsm", "aqsn", "vcoul", "I_epsR", "I_epsA", "dFreqGrid", "pref")
    T = {k: torch.from_numpy(np.ascontiguousarray(f[k])).to("cuda:0") for k in keys}
    out = torch.zeros(2, dtype=torch.complex128, device="cuda:0")
    ctx.ff_device("achDtemp", f, {k: v.data_ptr() for k, v in T.items()}, out.data_ptr())
    ctx.synchronize()
    # the host form subtracts R - A while staging, the device form in the contraction:
    # the same IEEE subtraction, the same sums
    assert np.array_equal(out.cpu().numpy(), ctx.achDtemp_Kernel(f))


def test_linear_in_the_weights(ctx):
    """Doubling pref doubles every weight exactly, so the result doubles bit for bit."""
    f = bgw.ff_inputs(50, 4, 200, 20, 7, 1, 5)
    a = ctx.achDtemp_Kernel(f)
    b = ctx.achDtemp_Kernel(dict(f, pref=2.0 * f["pref"]))
    assert np.array_equal(b, 2.0 * a)


def test_zero_denominator_is_left_out(ctx):
    """A band whose wx falls exactly on a grid frequency drops that one term."""
    f = bgw.ff_inputs(6, 2, 48, 8, 5, 1, 9)
    f["ekq"] = f["ekq"].copy()
    f["ekq"][3] = 0.5 - f["dFreqGrid"][2]  # unoccupied band 3: wx == dFreqGrid[2]
    got = ctx.achDtemp_Kernel(f)
    assert np.all(np.isfinite(got))
    assert rel_err(got, O.achDtemp(f)) < TOL


def test_validation(ctx):
    f = bgw.ff_inputs(10, 2, 40, 8, 4, 1, 3)
    with pytest.raises(bgw.BgwError) as e:
        ctx.achDtemp_Kernel(dict(f, nvband=11))
    assert e.value.code == bgw.BGW_EINVAL
    with pytest.raises(ValueError, match="pref"):
        ctx.achDtemp_Kernel(dict(f, pref=f["pref"][:2]))


@pytest.mark.slow
def test_config4_slices_reduced_bands(ctx):
    """The config-4 slices at full size (15 x 819 x 16384) with the first 8 bands."""
    f = bgw.ff_config4_inputs()
    nc = f["ncouls"]
    sub = dict(f, nb=8, aqsm=f["aqsm"][: 8 * nc].copy(), aqsn=f["aqsn"][: 8 * nc].copy(), ekq=f["ekq"][:8].copy())
    assert rel_err(ctx.achDtemp_Kernel(sub), O.achDtemp(sub)) < TOL


@pytest.mark.slow
def test_config4_full_vs_oracle(ctx):
    """Config 4 at full size (512 bands, 15 slices of 819 x 16384): both precision forms of
    the frequency sum against the oracle on the same inputs (the oracle's compensated FP64
    sum takes a few tens of seconds on the host cores)."""
    f = bgw.ff_config4_inputs()
    ref = O.achDtemp(f)
    assert rel_err(ctx.achDtemp_Kernel(f), ref) < TOL
    ctx.set_precision(bgw.BGW_FP32_BULK)
    try:
        got32 = ctx.achDtemp_Kernel(f)
    finally:
        ctx.set_precision(bgw.BGW_FP64)
    assert rel_err(got32, ref) < 1e-6
    assert rel_err(got32, ref) > 1e-13  # the FP32-operand path ran


def test_fp32_operand_form(ctx):
    f = bgw.ff_inputs(130, 5, 1000, 65, 15, 3, 2030)
    ref = O.achDtemp(f)
    ctx.set_precision(bgw.BGW_FP32_BULK)
    try:
        got = ctx.achDtemp_Kernel(f)
    finally:
        ctx.set_precision(bgw.BGW_FP64)
    assert rel_err(got, ref) < 1e-6
    assert rel_err(got, ref) > 1e-13  # the FP32-operand MFMA ran
