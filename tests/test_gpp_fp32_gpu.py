"""FP32-bulk precision form of the GPP forward (bgw_set_precision(BGW_FP32_BULK)):
bulk complex products in FP32, FP64 pole denominators near poles and FP64
reductions.  Target: <= 1e-6 relative error per output vs the FP64 oracle."""
import numpy as np
import pytest

import bgw
import oracle as O
from conftest import rel_err
from helpers import load_exact_gpp

pytestmark = pytest.mark.gpu
TOL32 = 1e-6


@pytest.fixture
def ctx32(ctx):
    ctx.set_precision(bgw.BGW_FP32_BULK)
    yield ctx
    ctx.set_precision(bgw.BGW_FP64)


def test_config0_fp32_bulk_vs_oracle(ctx32):
    d = bgw.gpp_config_inputs(0)
    r = ctx32.gpp_solver(d)
    o = O.gpp_forward(d)
    assert rel_err(r["achtemp"], o["achtemp"]) < TOL32
    assert rel_err(r["acht_n1_loc"], o["acht_n1_loc"]) < TOL32
    assert rel_err(r["achstemp"], o["achstemp"]) < TOL32  # FP32-operand MFMA, FP64 stage folds
    assert rel_err(r["asxtemp"], o["asxtemp"]) < TOL32
    # and it is not silently the FP64 path: the error is visibly above FP64 rounding
    assert rel_err(r["achtemp"], o["achtemp"]) > 1e-13


@pytest.mark.parametrize("nb", [5, 130, 300, 600])
def test_band_counts_fp32_bulk(ctx32, nb):
    d = bgw.gpp_inputs(nb, 2, 1500, 21, 40 + nb)
    r = ctx32.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    o = O.gpp_forward(d, want=("achtemp", "acht_n1_loc"))
    assert rel_err(r["acht_n1_loc"], o["acht_n1_loc"]) < TOL32


def test_achstemp_cut_fixture_fp32(ctx32):
    """The exact case whose I_eps sit on both sides of achstemp's |I_eps| > to1 cut: the
    FP32-operand contraction applies the cut like the FP64 one (a dropped or inverted
    cut changes achstemp by order one)."""
    d, exp = load_exact_gpp()[3]
    r = ctx32.gpp_solver(d, occ=1.0)
    o = O.gpp_forward(d, occ=1.0)
    for k in ("achtemp", "acht_n1_loc", "achstemp", "asxtemp"):
        assert rel_err(r[k], exp[k]) < TOL32, k
        assert rel_err(r[k], o[k]) < TOL32, k


def test_near_pole_tiles_use_fp64(ctx32):
    # wtilde close to a band frequency: the FP32 guard sends those tiles to FP64
    d = bgw.gpp_inputs(16, 2, 2000, 6, 77)
    wx = d["wx"].reshape(16, 3)
    wt = d["wtilde"].reshape(6, 2000)
    wt[:, ::97] = wx[5, 1] + 1e-3 - 1e-4j
    d["wtilde"] = wt.reshape(-1)
    r = ctx32.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    o = O.gpp_forward(d, want=("achtemp", "acht_n1_loc"))
    assert rel_err(r["acht_n1_loc"], o["acht_n1_loc"]) < TOL32


@pytest.mark.slow
def test_config2_reduced_bands_fp32_bulk(ctx32):
    d = bgw.gpp_config_inputs(2, nb=12)
    r = ctx32.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    o = O.gpp_forward(d, want=("achtemp", "acht_n1_loc"))
    assert rel_err(r["achtemp"], o["achtemp"]) < TOL32
    assert rel_err(r["acht_n1_loc"], o["acht_n1_loc"]) < TOL32


@pytest.mark.slow
def test_config2_full_fp32_bulk_vs_fp64(ctx):
    """Config 2 at full size (512 bands, ngpown 1638, ncouls 32768, every output): the
    FP32-bulk form against the FP64 path on the same inputs (which is itself checked
    against the reference's own kernel at this size), per output at most 1e-6 relative
    and visibly above FP64 rounding."""
    d = bgw.gpp_config_inputs(2)
    r64 = ctx.gpp_solver(d)
    ctx.set_precision(bgw.BGW_FP32_BULK)
    try:
        r32 = ctx.gpp_solver(d)
    finally:
        ctx.set_precision(bgw.BGW_FP64)
    for k in ("achtemp", "acht_n1_loc", "achstemp", "asxtemp"):
        assert rel_err(r32[k], r64[k]) < TOL32, k
    assert rel_err(r32["achtemp"], r64["achtemp"]) > 1e-13


def test_ff_fp32_operand_vs_oracle(ctx32):
    f = bgw.ff_inputs(130, 5, 1000, 65, 15, 3, 1030)
    achs = ctx32.achsDtemp_Kernel(f)
    asx = ctx32.asxDtemp_Kernel(f)
    cor = ctx32.achDtemp_cor_Kernel(f)
    assert abs(achs - O.achsDtemp(f)) <= TOL32 * abs(O.achsDtemp(f))
    assert rel_err(asx, O.asxDtemp(f)) < TOL32
    assert rel_err(cor, O.achDtemp_cor(f)) < TOL32
    # and the FP32 MFMA path ran: the error is visibly above FP64 rounding
    assert abs(achs - O.achsDtemp(f)) > 1e-13 * abs(O.achsDtemp(f))
    assert rel_err(cor, O.achDtemp_cor(f)) > 1e-13


def test_ff_fp32_operand_long_valence_lists(ctx32):
    """nvband above 64: asxDtemp's slice lists can exceed the square row tile, so the
    FP32 form takes the 80-row tile for it too (as for achDtemp_cor)."""
    f = bgw.ff_inputs(100, 70, 600, 40, 15, 2, 4711)
    asx = ctx32.asxDtemp_Kernel(f)
    cor = ctx32.achDtemp_cor_Kernel(f)
    assert rel_err(asx, O.asxDtemp(f)) < TOL32
    assert rel_err(cor, O.achDtemp_cor(f)) < TOL32
    assert rel_err(asx, O.asxDtemp(f)) > 1e-13


def test_band_kernels_streaming_form_in_fp32_mode(ctx32):
    """A band kernel is a one-row plan that streams its two slices once: bound by that
    reading, it takes the FP64 streaming form in both precision forms, so its FP64
    accuracy holds with the FP32-bulk form selected."""
    f = bgw.ff_inputs(40, 3, 500, 70, 15, 2, 1031)
    for kernel in (1, 2):
        got = ctx32.schDttt_corKernel(kernel, f, 5, 4, 0.6, 0.4)
        ref = O.schDttt_cor(f, kernel, 5, 4, 0.6, 0.4)
        assert abs(got - ref) <= 1e-12 * abs(ref)


@pytest.mark.slow
def test_ff_config4_shape_fp32_operand(ctx32):
    f = bgw.ff_config4_inputs(ncouls=2048)
    achs = ctx32.achsDtemp_Kernel(f)
    cor = ctx32.achDtemp_cor_Kernel(f)
    assert abs(achs - O.achsDtemp(f)) <= TOL32 * abs(O.achsDtemp(f))
    assert rel_err(cor, O.achDtemp_cor(f)) < TOL32
