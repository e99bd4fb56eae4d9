"""Moment form of the GPP forward and adjoint (bgw_set_precision(BGW_FP64_MOMENTS)).  Each
work item's G-sum sum_g I_eps(p,g)/(wx - wtilde(p,g)) is evaluated from 16 moments of its
far poles, those with |phi(w)| h/D <= 0.095 about the chunk's centre; the far chunks of a
row are re-expanded about one centre and evaluated together.  A chunk's near poles, up to
128, are listed and summed directly, and a chunk with more takes the direct FP64 sum as a
whole.  The truncation is below 2^-54 of the sum of the terms' magnitudes, so the FP64
tolerance applies: 1e-12 relative error per complex output against the oracle."""
import numpy as np
import pytest

import bgw
import oracle as O
from conftest import rel_err
from helpers import gpp_reference_closed_form, load_exact_gpp

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture
def ctxm(ctx):
    ctx.set_precision(bgw.BGW_FP64_MOMENTS)
    yield ctx
    ctx.set_precision(bgw.BGW_FP64)


def _direct(ctx, d, **kw):
    ctx.set_precision(bgw.BGW_FP64)
    try:
        return ctx.gpp_solver(d, **kw)
    finally:
        ctx.set_precision(bgw.BGW_FP64_MOMENTS)


def _near_poles(d, every):
    """Moves one pole in every `every` G-vectors of each row onto a band frequency's
    neighbourhood (|wx - wtilde| ~ 1e-3): a chunk holding one of them is not separated
    from the band frequencies and takes the direct sum."""
    wx = d["wx"].reshape(d["nb"], 3)
    wt = d["wtilde"].reshape(d["ngpown"], d["ncouls"]).copy()
    wt[:, ::every] = wx[d["nb"] // 2, 1] + 1e-3 - 1e-4j
    d["wtilde"] = wt.reshape(-1)
    return d


def test_config0_vs_oracle(ctxm):
    d = bgw.gpp_config_inputs(0)
    r = ctxm.gpp_solver(d, occ=1.0)
    o = O.gpp_forward(d, occ=1.0)
    for k in ("achtemp", "acht_n1_loc", "achstemp", "asxtemp"):
        assert rel_err(r[k], o[k]) < TOL, k
    # the moments were used: the sums are rounded differently from the direct form's
    f = _direct(ctxm, d, occ=1.0)
    assert not np.array_equal(r["acht_n1_loc"], f["acht_n1_loc"])
    assert rel_err(r["acht_n1_loc"], f["acht_n1_loc"]) < 1e-13


@pytest.mark.parametrize("nb", [2, 63, 65, 129, 257, 513])
def test_band_counts_vs_oracle(ctxm, nb):
    # every (threads, bands-per-lane) variant, partial waves and band tiles; 4500 G-vectors
    # per row give two full chunks and a short one
    d = bgw.gpp_inputs(nb, 2, 4500, 19, 3000 + nb)
    r = ctxm.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    o = O.gpp_forward(d, want=("achtemp", "acht_n1_loc"))
    assert rel_err(r["achtemp"], o["achtemp"]) < TOL
    assert rel_err(r["acht_n1_loc"], o["acht_n1_loc"]) < TOL


@pytest.mark.parametrize("ncouls,ngpown", [(1, 1), (2, 3), (255, 3), (2049, 5)])
def test_short_chunks_vs_oracle(ctxm, ncouls, ngpown):
    # chunks of 1, 2, 255 and 1 G-vectors: samples and wave sums over partly empty waves
    d = bgw.gpp_inputs(40, 2, ncouls, ngpown, 70 + ncouls)
    r = ctxm.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    o = O.gpp_forward(d, want=("achtemp", "acht_n1_loc"))
    assert rel_err(r["acht_n1_loc"], o["acht_n1_loc"]) < TOL


def test_reference_constants_closed_form(ctxm):
    # every pole at 0.5+0.5i: R = 0 and the expansion is exact in its first term
    nb, ngpown, ncouls = 512, 7, 3000
    d = bgw.gpp_inputs(nb, 2, ncouls, ngpown, 0, kind="reference")
    r = ctxm.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    assert rel_err(r["achtemp"], gpp_reference_closed_form(nb, ngpown, ncouls)) < TOL


@pytest.mark.parametrize("case", range(4))
def test_exact_fixtures(ctxm, case):
    d, exp = load_exact_gpp()[case]
    r = ctxm.gpp_solver(d, occ=1.0)
    for k in ("achtemp", "acht_n1_loc", "achstemp", "asxtemp"):
        assert rel_err(r[k], exp[k]) < 1e-13, k


def test_every_chunk_near_falls_back_bitwise(ctxm):
    """Every 7th pole near a band frequency: more near poles in every chunk than its list
    holds (128), so every item takes the direct sum, in the same launch shape and order as
    the FP64 form, and the results are equal bit for bit."""
    d = _near_poles(bgw.gpp_inputs(300, 2, 5000, 12, 91), 7)
    r = ctxm.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    f = _direct(ctxm, d, outputs=("achtemp", "acht_n1_loc"))
    assert np.array_equal(r["achtemp"], f["achtemp"])
    assert np.array_equal(r["acht_n1_loc"], f["acht_n1_loc"])
    o = O.gpp_forward(d, want=("achtemp", "acht_n1_loc"))
    assert rel_err(r["acht_n1_loc"], o["acht_n1_loc"]) < TOL


@pytest.mark.parametrize("every", [997, 61])
def test_sparse_near_poles_listed_vs_oracle(ctxm, every):
    """A few poles near a band frequency in every chunk (2 or 34 of 2048): they are listed
    and summed directly, the rest of the chunk from its moments; the sample's centre
    ignores them."""
    d = _near_poles(bgw.gpp_inputs(96, 2, 5000, 12, 90 + every), every)
    r = ctxm.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    o = O.gpp_forward(d, want=("achtemp", "acht_n1_loc"))
    assert rel_err(r["achtemp"], o["achtemp"]) < TOL
    assert rel_err(r["acht_n1_loc"], o["acht_n1_loc"]) < TOL
    f = _direct(ctxm, d, outputs=("achtemp", "acht_n1_loc"))
    assert not np.array_equal(r["acht_n1_loc"], f["acht_n1_loc"])  # the moments were used


def test_mixed_near_and_far_chunks_vs_oracle(ctxm):
    # near poles in the first chunk of each row only (G-vectors 0 and 1000 of 6000)
    d = bgw.gpp_inputs(64, 2, 6000, 9, 92)
    wx = d["wx"].reshape(64, 3)
    wt = d["wtilde"].reshape(9, 6000)
    wt[:, 0] = wx[7, 2] - 2e-3 - 1e-5j
    wt[:, 1000] = wx[50, 0] + 5e-4 - 1e-3j
    d["wtilde"] = wt.reshape(-1)
    r = ctxm.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    o = O.gpp_forward(d, want=("achtemp", "acht_n1_loc"))
    assert rel_err(r["acht_n1_loc"], o["acht_n1_loc"]) < TOL
    f = _direct(ctxm, d, outputs=("achtemp", "acht_n1_loc"))
    assert not np.array_equal(r["acht_n1_loc"], f["acht_n1_loc"])  # the far chunks used the moments


def test_separate_clusters_evaluate_their_own_moments(ctxm):
    """The second chunk of each row holds its poles near Re w = -29, far from the band
    frequencies but outside the disc about the other chunks' centre: depending on which
    chunk evaluates the row (chunk p mod 3 on), it or the others are evaluated from their
    own moments while the rest are re-expanded about the row's centre."""
    d = bgw.gpp_inputs(48, 2, 5000, 6, 96)
    wt = d["wtilde"].reshape(6, 5000)
    wt[:, 2048:4096] -= 30.0
    d["wtilde"] = wt.reshape(-1)
    r = ctxm.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    o = O.gpp_forward(d, want=("achtemp", "acht_n1_loc"))
    assert rel_err(r["acht_n1_loc"], o["acht_n1_loc"]) < TOL
    f = _direct(ctxm, d, outputs=("achtemp", "acht_n1_loc"))
    assert not np.array_equal(r["acht_n1_loc"], f["acht_n1_loc"])


def test_wide_pole_cloud_falls_back(ctxm):
    # poles spread over [0.5, 7]: the chunks' radius exceeds 0.26 x their distance to
    # wx >= 8, so the direct sum runs everywhere
    d = bgw.gpp_inputs(32, 2, 3000, 5, 93)
    wt = d["wtilde"]
    d["wtilde"] = 0.5 + (wt.real - 0.5) * 2.6 + 1j * wt.imag
    r = ctxm.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    f = _direct(ctxm, d, outputs=("achtemp", "acht_n1_loc"))
    assert np.array_equal(r["acht_n1_loc"], f["acht_n1_loc"])


def test_exact_path_tiles(ctxm):
    # poles at a band frequency with a tiny imaginary part: those chunks take the direct
    # sum's reference-arithmetic tiles
    d = bgw.gpp_inputs(8, 2, 600, 4, 8080)
    wx = d["wx"].reshape(8, 3)
    wt = d["wtilde"].reshape(4, 600)
    wt[1, 10] = wx[3, 1] + 1j * 2.0**-170
    wt[2, 599] = wx[0, 0] - 1j * 2.0**-160
    d["wtilde"] = wt.reshape(-1)
    r = ctxm.gpp_solver(d)
    o = O.gpp_forward(d)
    assert np.all(np.isfinite(r["achtemp"]))
    assert rel_err(r["acht_n1_loc"], o["acht_n1_loc"]) < TOL


def test_non_finite_pole_propagates_like_direct(ctxm):
    d = bgw.gpp_inputs(16, 2, 3000, 4, 94)
    wt = d["wtilde"].reshape(4, 3000)
    wt[2, 2500] = np.nan
    d["wtilde"] = wt.reshape(-1)
    r = ctxm.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    f = _direct(ctxm, d, outputs=("achtemp", "acht_n1_loc"))
    assert np.array_equal(np.isnan(r["acht_n1_loc"]), np.isnan(f["acht_n1_loc"]))
    assert np.all(np.isnan(r["achtemp"]))


def test_bitwise_reproducible_and_device_equals_host(ctxm):
    import torch

    d = bgw.gpp_inputs(100, 2, 5000, 64, 95)
    a = ctxm.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    b = ctxm.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    assert np.array_equal(a["achtemp"], b["achtemp"]) and np.array_equal(a["acht_n1_loc"], b["acht_n1_loc"])
    dev = torch.device("cuda:0")
    T = {k: torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in
         ("inv", "indinv", "wx", "wtilde", "aqsm", "aqsn", "I_eps", "vcoul")}
    ach = torch.zeros(3, dtype=torch.complex128, device=dev)
    loc = torch.zeros(100 * 3, dtype=torch.complex128, device=dev)
    torch.cuda.synchronize()
    P = [T[k].data_ptr() for k in ("inv", "indinv", "wx", "wtilde", "aqsm", "aqsn", "I_eps", "vcoul")]
    ctxm.gpp_solver_device(100, 2, 64, 5000, *P, 1.0, ach.data_ptr(), loc.data_ptr(), None, None)
    ctxm.synchronize()
    # the host call stages the rows in ranges and continues the totals across launches
    assert np.array_equal(ach.cpu().numpy(), a["achtemp"])
    assert np.array_equal(loc.cpu().numpy().reshape(100, 3), a["acht_n1_loc"])


@pytest.mark.slow
def test_config1_full_vs_direct_and_oracle(ctxm):
    """The benchmark workload at full size (512 bands, ngpown 1638, ncouls 32768): the
    moment form against the direct FP64 form, and 12 bands against the oracle."""
    d = bgw.gpp_config_inputs(1)
    r = ctxm.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    f = _direct(ctxm, d, outputs=("achtemp", "acht_n1_loc"))
    assert rel_err(r["achtemp"], f["achtemp"]) < TOL
    assert rel_err(r["acht_n1_loc"], f["acht_n1_loc"]) < TOL
    assert rel_err(r["acht_n1_loc"].sum(0), r["achtemp"]) < 1e-14
    sub = dict(d, nb=12, aqsm=d["aqsm"][: 12 * 32768].copy(), aqsn=d["aqsn"][: 12 * 32768].copy(),
               wx=d["wx"][:36].copy())
    o = O.gpp_forward(sub, want=("achtemp", "acht_n1_loc"))
    assert rel_err(r["acht_n1_loc"][:12], o["acht_n1_loc"]) < TOL


# ------------------------------------------------------------------ adjoint in the moment form
def _bwd_direct(ctx, d, occ):
    ctx.set_precision(bgw.BGW_FP64)
    try:
        return ctx.gpp_backward(d, occ=occ)
    finally:
        ctx.set_precision(bgw.BGW_FP64_MOMENTS)


def _grad_close(g, dw, dI, ds, tol=1e-12):
    assert np.max(np.abs(g["dL_dwtilde"] - dw)) <= tol * np.max(np.abs(dw))
    assert np.max(np.abs(g["dL_dI_eps"] - dI)) <= tol * np.max(np.abs(dI))
    assert rel_err(g["dsigma_domega"], ds) < tol


@pytest.mark.parametrize("nb,ncouls,ngpown", [(70, 300, 20), (1100, 600, 4), (33, 1500, 7)])
def test_adjoint_vs_oracle(ctxm, nb, ncouls, ngpown):
    """The band sums of every pole from the row's 16 target moments (1100 bands: more than
    the direct loop's LDS chunk; 1500 G-vectors: three blocks per row, the last one short)."""
    d = bgw.gpp_inputs(nb, 3, ncouls, ngpown, 555 + nb)
    g = ctxm.gpp_backward(d, occ=0.8)
    e = np.arange(ngpown * ncouls, dtype=np.int64)
    dw, dI = O.gpp_grad_entries(d, e, occ=0.8)
    _grad_close(g, dw, dI, O.gpp_dsigma_domega(d, occ=0.8))
    f = _bwd_direct(ctxm, d, 0.8)
    assert not np.array_equal(g["dL_dI_eps"], f["dL_dI_eps"])  # the moments were used


def test_adjoint_near_poles_fall_back_bitwise(ctxm):
    # a pole near a band frequency in every wave of 64 poles: the band loop everywhere
    d = _near_poles(bgw.gpp_inputs(64, 2, 2000, 6, 97), 61)
    g = ctxm.gpp_backward(d, occ=1.0)
    f = _bwd_direct(ctxm, d, 1.0)
    for k in g:
        assert np.array_equal(g[k], f[k]), k


def test_adjoint_sparse_near_poles_vs_oracle(ctxm):
    # one pole in 499 near a band frequency: its wave takes the band loop, the others the
    # expansion
    d = _near_poles(bgw.gpp_inputs(64, 2, 2000, 6, 96), 499)
    g = ctxm.gpp_backward(d, occ=1.0)
    e = np.arange(6 * 2000, dtype=np.int64)
    dw, dI = O.gpp_grad_entries(d, e, occ=1.0)
    _grad_close(g, dw, dI, O.gpp_dsigma_domega(d, occ=1.0))
    f = _bwd_direct(ctxm, d, 1.0)
    assert not np.array_equal(g["dL_dI_eps"], f["dL_dI_eps"])


def test_adjoint_mixed_blocks_vs_oracle(ctxm):
    # a near pole in one wave of the first 512-G block of each row
    d = bgw.gpp_inputs(48, 2, 1800, 5, 98)
    wx = d["wx"].reshape(48, 3)
    wt = d["wtilde"].reshape(5, 1800)
    wt[:, 300] = wx[9, 1] - 1e-3 - 2e-4j
    d["wtilde"] = wt.reshape(-1)
    g = ctxm.gpp_backward(d, occ=1.0)
    e = np.arange(5 * 1800, dtype=np.int64)
    dw, dI = O.gpp_grad_entries(d, e, occ=1.0)
    _grad_close(g, dw, dI, O.gpp_dsigma_domega(d, occ=1.0))


def test_adjoint_bitwise_reproducible(ctxm):
    d = bgw.gpp_inputs(40, 2, 700, 9, 99)
    a = ctxm.gpp_backward(d)
    b = ctxm.gpp_backward(d)
    for k in a:
        assert np.array_equal(a[k], b[k])


@pytest.mark.slow
def test_adjoint_config3_vs_direct(ctxm):
    """The adjoint workload at full size (256 bands, ngpown 819, ncouls 16384): the moment
    form against the direct FP64 backward (itself checked against the oracle's analytic
    gradient and finite differences) on the same inputs."""
    d = bgw.gpp_config_inputs(3)
    g = ctxm.gpp_backward(d, occ=1.0)
    f = _bwd_direct(ctxm, d, 1.0)
    _grad_close(g, f["dL_dwtilde"], f["dL_dI_eps"], f["dsigma_domega"])
