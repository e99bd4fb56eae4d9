"""GPU adjoint (bgw_gpp_backward) against the oracle's analytic gradient and
central finite differences of the GPU forward, for
L = Re sum_iw (asxtemp[iw] + sum_n1 acht_n1_loc[n1][iw])."""
import numpy as np
import pytest

import bgw
import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _entries(d, n, seed):
    rng = np.random.default_rng(seed)
    npg = d["ngpown"] * d["ncouls"]
    igp = d["indinv"][d["inv"]]
    e = list(rng.choice(npg, n - n // 4, replace=False))
    ps = rng.choice(d["ngpown"], n // 4, replace=False)
    e += [int(p) * d["ncouls"] + int(igp[p]) for p in ps]  # the G' columns carry the extra T(p) term
    return np.asarray(e, np.int64)


def test_small_full_gradient_vs_oracle(ctx):
    d = bgw.gpp_inputs(70, 3, 300, 20, 555)
    g = ctx.gpp_backward(d, occ=0.8)
    e = np.arange(d["ngpown"] * d["ncouls"], dtype=np.int64)
    dw, dI = O.gpp_grad_entries(d, e, occ=0.8)
    assert np.max(np.abs(g["dL_dwtilde"] - dw)) <= 1e-12 * np.max(np.abs(dw))
    assert np.max(np.abs(g["dL_dI_eps"] - dI)) <= 1e-12 * np.max(np.abs(dI))
    assert rel_err(g["dsigma_domega"], O.gpp_dsigma_domega(d, occ=0.8)) < 1e-12


def test_near_pole_branches_vs_oracle(ctx):
    from helpers import load_exact_gpp

    d, _ = load_exact_gpp()[2]  # near-pole, zero-branch and cut-off terms in asxtemp
    g = ctx.gpp_backward(d, occ=1.0)
    e = np.arange(d["ngpown"] * d["ncouls"], dtype=np.int64)
    dw, dI = O.gpp_grad_entries(d, e, occ=1.0)
    assert np.max(np.abs(g["dL_dwtilde"] - dw)) <= 1e-12 * np.max(np.abs(dw))
    assert np.max(np.abs(g["dL_dI_eps"] - dI)) <= 1e-12 * np.max(np.abs(dI))
    assert rel_err(g["dsigma_domega"], O.gpp_dsigma_domega(d, occ=1.0)) < 1e-12


def test_ssx_mixed_branches_within_a_band_vs_oracle(ctx):
    # The screened-exchange terms of a band's three frequencies are formed together when all
    # three are surely regular and >= 0; these inputs break that for some G-vectors: 40 poles
    # next to valence band 0's middle frequency (|wdiff|^2 = 5e-4 < limittwo there: the
    # near-pole branch, the other two frequencies regular), and valence band 1 with a negative
    # first frequency (the cut-off test).  Forward asxtemp and the gradients against the oracle.
    d = bgw.gpp_inputs(24, 2, 256, 8, 31)
    ngpown, ncouls = d["ngpown"], d["ncouls"]
    wx = d["wx"].reshape(-1, 3).copy()
    wx[1, 0] = -0.3
    d["wx"] = wx.reshape(-1).copy()
    wt = d["wtilde"].copy()
    idx = np.random.default_rng(5).choice(ngpown * ncouls, 40, replace=False)
    wt[idx] = wx[0, 1] + 0.02 - 0.01j
    d["wtilde"] = wt
    r = ctx.gpp_solver(d, occ=1.0)
    o = O.gpp_forward(d, occ=1.0)
    assert rel_err(r["asxtemp"], o["asxtemp"]) < 1e-12
    assert rel_err(r["achtemp"], o["achtemp"]) < 1e-12
    g = ctx.gpp_backward(d, occ=1.0)
    e = np.arange(ngpown * ncouls, dtype=np.int64)
    dw, dI = O.gpp_grad_entries(d, e, occ=1.0)
    assert np.max(np.abs(g["dL_dwtilde"] - dw)) <= 1e-12 * np.max(np.abs(dw))
    assert np.max(np.abs(g["dL_dI_eps"] - dI)) <= 1e-12 * np.max(np.abs(dI))
    assert rel_err(g["dsigma_domega"], O.gpp_dsigma_domega(d, occ=1.0)) < 1e-12


def test_more_bands_than_one_lds_chunk_vs_oracle(ctx):
    # 1100 bands: the band loop stages c(n1,p) and wx in LDS chunks of 1024
    d = bgw.gpp_inputs(1100, 3, 300, 4, 1100)
    g = ctx.gpp_backward(d, occ=1.0)
    e = np.arange(d["ngpown"] * d["ncouls"], dtype=np.int64)
    dw, dI = O.gpp_grad_entries(d, e, occ=1.0)
    assert np.max(np.abs(g["dL_dwtilde"] - dw)) <= 1e-12 * np.max(np.abs(dw))
    assert np.max(np.abs(g["dL_dI_eps"] - dI)) <= 1e-12 * np.max(np.abs(dI))
    assert rel_err(g["dsigma_domega"], O.gpp_dsigma_domega(d, occ=1.0)) < 1e-12


def test_bitwise_reproducible(ctx):
    d = bgw.gpp_inputs(40, 2, 700, 9, 1)
    a = ctx.gpp_backward(d)
    b = ctx.gpp_backward(d)
    for k in a:
        assert np.array_equal(a[k], b[k])


def _sub_problem(d, p, g):
    """The terms of L that depend on wtilde(p,g) / I_eps(p,g): row p, and, unless
    g is the G' column igp(p) of row p (whose wtilde enters sch_store1 of every
    term of the row), only the columns {g, igp}.  Small sub-sums keep the
    finite-difference rounding noise far below the derivative."""
    nc = d["ncouls"]
    igp = int(d["indinv"][d["inv"][p]])
    if g == igp:
        cols = np.arange(nc)
        pos = igp
    else:
        cols = np.array([g, igp])
        pos = 1
    k = len(cols)
    A = lambda x: np.ascontiguousarray(x.reshape(-1, nc)[:, cols]).reshape(-1)  # noqa: E731
    return dict(d, ncouls=k, ngpown=1, inv=np.array([pos], np.int32),
                indinv=np.concatenate([np.arange(k, dtype=np.int32), np.array([k - 1], np.int32)]),
                aqsm=A(d["aqsm"]), aqsn=A(d["aqsn"]), vcoul=np.ascontiguousarray(d["vcoul"][cols]),
                wtilde=np.ascontiguousarray(d["wtilde"][p * nc + cols]),
                I_eps=np.ascontiguousarray(d["I_eps"][p * nc + cols])), (0 if g != igp else g)


def _L(ctx, d):
    r = ctx.gpp_solver(d, occ=1.0, outputs=("achtemp", "asxtemp"))
    return float(np.sum(r["achtemp"].real) + np.sum(r["asxtemp"].real))


@pytest.mark.slow
def test_config3_entries_vs_oracle_and_finite_differences(ctx):
    # config 3: number_bands 256, nvband 2, ncouls 16384, ngpown 819
    d = bgw.gpp_config_inputs(3)
    g = ctx.gpp_backward(d, occ=1.0)
    e = _entries(d, 64, 3)
    dw, dI = O.gpp_grad_entries(d, e, occ=1.0)
    assert rel_err(g["dL_dwtilde"][e], dw) < 1e-11
    assert rel_err(g["dL_dI_eps"][e], dI) < 1e-11
    # central finite differences (h = 1e-6) of the GPU forward on 64 seeded entries
    h = 1e-6
    nc = d["ncouls"]
    for ent in e:
        p, gg = divmod(int(ent), nc)
        sub, j = _sub_problem(d, p, gg)
        Lmag = abs(_L(ctx, sub))
        for name, ana in (("wtilde", g["dL_dwtilde"][ent]), ("I_eps", g["dL_dI_eps"][ent])):
            fd = []
            for unit in (1.0, 1j):
                rp = dict(sub, **{name: sub[name].copy()})
                rm = dict(sub, **{name: sub[name].copy()})
                rp[name][j] += h * unit
                rm[name][j] -= h * unit
                fd.append((_L(ctx, rp) - _L(ctx, rm)) / (2 * h))
            fdc = fd[0] + 1j * fd[1]
            # truncation O(h^2) plus rounding of L (~1e-16 |L| per evaluation) / h
            tol = 1e-6 * abs(ana) + 4e-15 * Lmag / h
            assert abs(fdc - ana) <= tol, (name, p, gg, fdc, ana, tol)


@pytest.mark.slow
def test_config3_dsigma_domega_vs_oracle_and_fd(ctx):
    d = bgw.gpp_config_inputs(3)
    g = ctx.gpp_backward(d, occ=1.0)
    sub = dict(d, nb=8, aqsm=d["aqsm"][: 8 * 16384].copy(), aqsn=d["aqsn"][: 8 * 16384].copy(), wx=d["wx"][:24].copy())
    gs = ctx.gpp_backward(sub, occ=1.0)
    assert rel_err(gs["dsigma_domega"], O.gpp_dsigma_domega(sub)) < 1e-12
    h = 1e-5
    rp = ctx.gpp_solver(dict(d, wx=d["wx"] + h), outputs=("achtemp", "asxtemp"))
    rm = ctx.gpp_solver(dict(d, wx=d["wx"] - h), outputs=("achtemp", "asxtemp"))
    fd = ((rp["achtemp"] + rp["asxtemp"]) - (rm["achtemp"] + rm["asxtemp"])) / (2 * h)
    assert rel_err(fd, g["dsigma_domega"]) < 1e-6
    # quasiparticle renormalisation factor Z = 1/(1 - Re dSigma/domega) is finite
    Z = 1.0 / (1.0 - g["dsigma_domega"].real)
    assert np.all(np.isfinite(Z))
