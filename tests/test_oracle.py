"""Pins the CPU oracle (oracle/bgw_oracle.c) before it is trusted as the checker.

Each oracle function is checked against something it was not written from:
  * exact rational evaluations of the formulas (tests/golden/exact_*.json);
  * closed forms for the reference drivers' constant inputs, including the
    reference's own printed goldens (GPP/gpp.cpp:119-125, FF/ff.cpp:72-93);
  * the reference's own kernels compiled from its sources (oracle/_ref), on
    seeded random inputs;
  * invariants that a dropped term, wrong sign, wrong index or swapped operand
    breaks (band-sum, linearity, separable case, finite differences).
"""
import numpy as np
import pytest

import bgw
import oracle as O
from conftest import rel_err
from helpers import (ff_band_terms, ff_reference_achDtemp, ff_reference_closed_form, ff_with_kernel2_band,
                     gpp_reference_achstemp, gpp_reference_closed_form, load_exact_ff, load_exact_gpp)

# ------------------------------------------------------------------------- GPP


@pytest.mark.parametrize("case", range(4))
def test_gpp_oracle_exact_fixture(case):
    d, exp = load_exact_gpp()[case]
    r = O.gpp_forward(d)
    assert rel_err(r["achtemp"], exp["achtemp"]) < 1e-14
    assert rel_err(r["acht_n1_loc"], exp["acht_n1_loc"]) < 1e-14
    assert rel_err(r["achstemp"], exp["achstemp"]) < 1e-14
    assert rel_err(r["asxtemp"], exp["asxtemp"]) < 1e-14  # case 2 covers near-pole and cut-off


def _achstemp_numpy(d, keep):
    """-1/2 sum_{n1,p} vcoul(igp) conj(aqsm(n1,igp)) sum_g aqsn(n1,g) I(p,g) keep(I(p,g))."""
    nb, ng, nc = d["nb"], d["ngpown"], d["ncouls"]
    igp = d["indinv"][d["inv"]]
    I = d["I_eps"].reshape(ng, nc)
    Im = np.where(keep(I), I, 0.0)
    am, an = d["aqsm"].reshape(nb, nc), d["aqsn"].reshape(nb, nc)
    m = an @ Im.T  # [nb, ng]
    return -0.5 * np.sum(d["vcoul"][igp][None, :] * np.conj(am[:, igp]) * m)


def test_achstemp_cut_fixture_discriminates():
    """Case 3 of the exact GPP fixtures has every I_eps at zero or near the static-COH cut
    |I_eps| > to1.  Its achstemp pins the cut: the right mask reproduces it, and a removed
    mask, an inverted one, or one that compares |I|^2 with to1 instead of to1^2 miss it
    by far more than the FP32 form's tolerance."""
    d, exp = load_exact_gpp()[3]
    to1 = 1e-6
    n2 = lambda I: I.real ** 2 + I.imag ** 2  # noqa: E731
    right = _achstemp_numpy(d, lambda I: n2(I) > to1 * to1)
    assert rel_err([right], exp["achstemp"]) < 1e-14
    kept = np.sum(n2(d["I_eps"]) > to1 * to1)
    assert 0 < kept < np.count_nonzero(d["I_eps"])  # both sides of the cut are present
    for wrong in (lambda I: np.ones(I.shape, bool), lambda I: ~(n2(I) > to1 * to1), lambda I: n2(I) > to1):
        assert rel_err([_achstemp_numpy(d, wrong)], exp["achstemp"]) > 1e-2


@pytest.mark.parametrize("nb,ngpown,ncouls", [(1, 1, 1), (3, 5, 17), (7, 33, 130), (12, 64, 1000)])
def test_gpp_oracle_closed_form_reference_constants(nb, ngpown, ncouls):
    d = bgw.gpp_inputs(nb, 2, ncouls, ngpown, 0, kind="reference")
    r = O.gpp_forward(d)
    assert rel_err(r["achtemp"], gpp_reference_closed_form(nb, ngpown, ncouls)) < 1e-15
    assert rel_err(r["achstemp"], [gpp_reference_achstemp(nb, ngpown, ncouls)]) < 1e-15


def test_gpp_reference_default_golden_vs_closed_form():
    # Default run 512 2 32768 20 (GPP/gpp.cpp:145-149): ngpown = 1638.
    cf = gpp_reference_closed_form(512, 1638, 32768)
    assert cf[0] == complex(-264241152.0, 1321205760.0)
    # the reference's own OpenMP and serial goldens are ~1e-8 and ~3e-7 from it
    omp = complex(-264241149.849658, 1321205773.349384)
    ser = complex(-264241220.914570, 1321205332.084101)
    assert 1e-9 < abs(omp - cf[0]) / abs(cf[0]) < 1e-7
    assert 1e-8 < abs(ser - cf[0]) / abs(cf[0]) < 1e-6


def test_gpp_oracle_vs_reference_build_per_band():
    if not O.have_ref():
        pytest.skip("reference checker not built (needs the reference tree)")
    d = bgw.gpp_inputs(9, 2, 700, 300, 424242)
    r = O.gpp_forward(d)
    ref = O.ref_gpp_per_band(d)
    # the reference sums naively (error ~1e-14 here); the oracle is compensated
    assert rel_err(ref, r["acht_n1_loc"]) < 1e-11
    assert rel_err(ref.sum(0), r["achtemp"]) < 1e-11


def test_gpp_oracle_vs_reference_build_noflag():
    if not O.have_ref():
        pytest.skip("reference checker not built (needs the reference tree)")
    d = bgw.gpp_inputs(6, 2, 512, 256, 99)
    wx3 = d["wx"][:3].copy()
    d["wx"] = np.tile(wx3, d["nb"])
    r = O.gpp_forward(d)
    assert rel_err(O.ref_noflagOCC_solver(d, wx3), r["achtemp"]) < 1e-11


def test_gpp_oracle_vs_committed_reference_vectors():
    """Vectors the reference build wrote (tests/golden/make_ref_vectors.py)."""
    import json

    from helpers import GOLDEN

    path = GOLDEN / "ref_vectors.json"
    vec = json.loads(path.read_text())
    for c in vec["gpp"]:
        d = bgw.gpp_inputs(c["nb"], c["nvband"], c["ncouls"], c["ngpown"], c["seed"])
        r = O.gpp_forward(d)
        ref = np.asarray(c["acht_n1_loc"])
        assert rel_err(ref[..., 0] + 1j * ref[..., 1], r["acht_n1_loc"]) < 1e-11


def test_gpp_oracle_invariants():
    d = bgw.gpp_inputs(5, 2, 333, 70, 777)
    r = O.gpp_forward(d)
    # band-sum
    assert rel_err(r["acht_n1_loc"].sum(0), r["achtemp"]) < 1e-15
    # linearity in I_eps (achtemp, achstemp, asxtemp are all linear in I)
    d2 = dict(d, I_eps=2.0 * d["I_eps"])
    r2 = O.gpp_forward(d2)
    for k in ("achtemp", "achstemp", "asxtemp"):
        assert rel_err(r2[k], 2.0 * r[k]) < 1e-15
    # separable case: band-independent wx -> achtemp = sum_p (sum_n1 s) T(p,iw), numpy
    wx3 = d["wx"][:3].copy()
    d3 = dict(d, wx=np.tile(wx3, d["nb"]))
    r3 = O.gpp_forward(d3)
    nb, ngpown, ncouls = d["nb"], d["ngpown"], d["ncouls"]
    am = d["aqsm"].reshape(nb, ncouls)
    an = d["aqsn"].reshape(nb, ncouls)
    wt = d["wtilde"].reshape(ngpown, ncouls)
    I = d["I_eps"].reshape(ngpown, ncouls)
    igp = d["indinv"][d["inv"]]
    s = np.conj(am[:, igp]) * an[:, igp] * 0.5 * d["vcoul"][igp] * wt[np.arange(ngpown), igp]  # [nb, ngpown]
    T = np.stack([(I / (w - wt)).sum(1) for w in wx3], 1)  # [ngpown, 3]
    assert rel_err(r3["achtemp"], s.sum(0) @ T) < 1e-12


def test_gpp_asx_oracle_matches_numpy_regular_branch():
    # with the benchmark distributions every term is in the regular branch:
    # ssx = wt^2 I / (wx^2 - wt^2)
    d = bgw.gpp_inputs(4, 3, 200, 50, 31337)
    r = O.gpp_forward(d, occ=0.75)
    nb, ngpown, ncouls = d["nb"], d["ngpown"], d["ncouls"]
    am = d["aqsm"].reshape(nb, ncouls)
    an = d["aqsn"].reshape(nb, ncouls)
    wt = d["wtilde"].reshape(ngpown, ncouls)
    I = d["I_eps"].reshape(ngpown, ncouls)
    igp = d["indinv"][d["inv"]]
    vc = d["vcoul"][igp]
    wx = d["wx"].reshape(nb, 3)
    exp = np.zeros(3, complex)
    for n1 in range(d["nvband"]):
        m = 0.75 * vc * np.conj(am[n1, igp])
        for iw in range(3):
            ssx = wt**2 * I / (wx[n1, iw] ** 2 - wt**2)
            exp[iw] += np.sum(m * (ssx * an[n1][None, :]).sum(1))
    assert rel_err(r["asxtemp"], exp) < 1e-12


def _L(d, occ=1.0):
    r = O.gpp_forward(d, occ=occ, want=("achtemp", "asxtemp"))
    return float(np.sum(r["achtemp"].real) + np.sum(r["asxtemp"].real))


def test_gpp_oracle_gradient_vs_finite_differences():
    d = bgw.gpp_inputs(4, 2, 24, 6, 2024)
    rng = np.random.default_rng(5)
    igp = d["indinv"][d["inv"]]
    entries = list(rng.integers(0, d["ngpown"] * d["ncouls"], 10)) + [p * d["ncouls"] + igp[p] for p in range(3)]
    entries = np.asarray(entries, np.int64)
    dw, dI = O.gpp_grad_entries(d, entries)
    h = 1e-6
    for k, e in enumerate(entries):
        for name, grad in (("wtilde", dw), ("I_eps", dI)):
            fd = []
            for unit in (1.0, 1j):
                dp = {**d, name: d[name].copy()}
                dm = {**d, name: d[name].copy()}
                dp[name][e] += h * unit
                dm[name][e] -= h * unit
                fd.append((_L(dp) - _L(dm)) / (2 * h))
            fdc = fd[0] + 1j * fd[1]
            assert abs(fdc - grad[k]) <= 1e-6 * max(1.0, abs(grad[k])), (name, e, fdc, grad[k])


def test_gpp_oracle_gradient_near_pole_branches_vs_finite_differences():
    # exact fixture 2: terms in the near-pole branch, the zero branch and the cut-off
    d, _ = load_exact_gpp()[2]
    n = d["ngpown"] * d["ncouls"]
    entries = np.arange(n, dtype=np.int64)
    dw, dI = O.gpp_grad_entries(d, entries)
    h = 1e-7
    for k in entries:
        for name, grad in (("wtilde", dw), ("I_eps", dI)):
            fd = []
            for unit in (1.0, 1j):
                dp = {**d, name: d[name].copy()}
                dm = {**d, name: d[name].copy()}
                dp[name][k] += h * unit
                dm[name][k] -= h * unit
                fd.append((_L(dp) - _L(dm)) / (2 * h))
            fdc = fd[0] + 1j * fd[1]
            assert abs(fdc - grad[k]) <= 1e-6 * max(1.0, abs(grad[k])), (name, k, fdc, grad[k])
    ds = O.gpp_dsigma_domega(d)
    hw = 1e-7
    rp = O.gpp_forward(dict(d, wx=d["wx"] + hw), want=("achtemp", "asxtemp"))
    rm = O.gpp_forward(dict(d, wx=d["wx"] - hw), want=("achtemp", "asxtemp"))
    fd = ((rp["achtemp"] + rp["asxtemp"]) - (rm["achtemp"] + rm["asxtemp"])) / (2 * hw)
    assert np.max(np.abs(fd - ds)) <= 1e-6 * np.max(np.abs(ds))


def test_gpp_oracle_dsigma_domega_vs_finite_differences():
    d = bgw.gpp_inputs(4, 2, 40, 8, 4048)
    ds = O.gpp_dsigma_domega(d)
    h = 1e-5
    dp = dict(d, wx=d["wx"] + h)
    dm = dict(d, wx=d["wx"] - h)
    rp = O.gpp_forward(dp, want=("achtemp", "asxtemp"))
    rm = O.gpp_forward(dm, want=("achtemp", "asxtemp"))
    fd = ((rp["achtemp"] + rp["asxtemp"]) - (rm["achtemp"] + rm["asxtemp"])) / (2 * h)
    assert rel_err(fd, ds) < 1e-7


# -------------------------------------------------------------------------- FF


@pytest.mark.parametrize("case", range(2))
def test_ff_oracle_exact_fixture(case):
    f, exp = load_exact_ff()[case]
    assert abs(O.achsDtemp(f) - exp["achsDtemp"]) <= 1e-14 * abs(exp["achsDtemp"])
    asx = O.asxDtemp(f)
    assert np.all(np.abs(asx - exp["asxDtemp"]) <= 1e-14 * abs(exp["asxDtemp"]))
    cor = O.achDtemp_cor(f)
    assert np.all(np.abs(cor - exp["achDtemp_cor"]) <= 1e-14 * abs(exp["achDtemp_cor"]))
    achD = O.achDtemp(f)
    assert np.all(np.abs(achD - exp["achDtemp"]) <= 1e-14 * abs(exp["achDtemp"]))


@pytest.mark.parametrize("nb,nvband,ncouls,ngpown,nFreq", [(20, 4, 64, 16, 15), (40, 25, 33, 7, 6), (9, 9, 10, 10, 3)])
def test_ff_oracle_closed_form_reference_constants(nb, nvband, ncouls, ngpown, nFreq):
    f = bgw.ff_inputs(nb, nvband, ncouls, ngpown, nFreq, 3, 0, kind="reference")
    achs, asx, cor = ff_reference_closed_form(nb, nvband, ncouls, ngpown, nFreq)
    assert abs(O.achsDtemp(f) - achs) <= 1e-15 * abs(achs)
    assert np.all(np.abs(O.asxDtemp(f) - asx) <= 1e-15 * abs(asx))
    assert np.all(np.abs(O.achDtemp_cor(f) - cor) <= 1e-15 * abs(cor))
    achD = ff_reference_achDtemp(nb, nvband, ncouls, ngpown, nFreq)
    assert np.all(np.abs(O.achDtemp(f) - achD) <= 1e-13 * abs(achD))


def test_pref_weights():
    """FF/ff.cpp:619-626 on a uniform grid of spacing 1/2 and on an uneven one: interior
    weights are the spacing over 3.14, the last point repeats the spacing before it, the
    two ends are halved; the binding's helper and the oracle's agree."""
    g = np.array([0.0, 0.5, 1.0, 1.5, 2.0])
    assert np.array_equal(O.pref_weights(g), np.array([0.25, 0.5, 0.5, 0.5, 0.25]) / 3.14)
    u = np.array([0.0, 0.25, 1.0, 3.0])
    assert np.allclose(O.pref_weights(u), np.array([0.125, 0.75, 2.0, 1.0]) / 3.14, rtol=1e-15, atol=0)
    assert np.array_equal(bgw.ff_pref(u), O.pref_weights(u))
    with pytest.raises(ValueError):
        bgw.ff_pref([1.0])


def test_achDtemp_occupation_sign():
    """The pole of an occupied band sits at -dFreqGrid, of an unoccupied one at
    +dFreqGrid: with one frequency slice carrying R - A, moving band 0 between
    occupied and unoccupied changes its term by exactly the ratio of its denominators."""
    f = bgw.ff_inputs(3, 0, 16, 4, 4, 1, 31)
    f["I_epsA"] = f["I_epsA"].copy()
    npg = 4 * 16
    for k in (0, 1, 3):  # only slice 2 has R != A
        f["I_epsA"][k * npg:(k + 1) * npg] = f["I_epsR"][k * npg:(k + 1) * npg]
    sub = dict(f, nb=1, aqsm=f["aqsm"][:16].copy(), aqsn=f["aqsn"][:16].copy(), ekq=f["ekq"][:1].copy())
    unocc = O.achDtemp(dict(sub, nvband=0))[0]
    occ = O.achDtemp(dict(sub, nvband=1))[0]
    wx = f["freqevalmin"] - f["ekq"][0] + f["freqevalstep"]
    w2 = f["dFreqGrid"][2]
    assert abs(occ * (wx + w2) - unocc * (wx - w2)) <= 1e-14 * abs(unocc * (wx - w2))


def test_ff_reference_default_goldens_vs_closed_form():
    # default run 15023 1998 22401 66 15 10 (FF/ff.cpp:471-477)
    achs, asx, cor = ff_reference_closed_form(15023, 1998, 22401, 66, 15)
    assert achs == complex(-2776374339.75, -2776374339.75)
    # FF/ff.cpp:72-73 prints +2776374339.75 for the imaginary part: that golden has
    # the wrong sign and passes only because the check is one-sided (:75)
    assert complex(-2776374339.75, 2776374339.75) != achs
    assert asx == complex(738493767.0, -730362204.0)  # FF/ff.cpp:82-83
    assert cor == complex(0.0, -734427985.5)  # FF/ff.cpp:92-93


def test_ff_oracle_vs_reference_build():
    if not O.have_ref():
        pytest.skip("reference checker not built (needs the reference tree)")
    for seed, shape in ((5, (40, 3, 96, 24, 15, 2)), (6, (33, 5, 70, 9, 7, 4))):
        f = bgw.ff_inputs(*shape, seed)
        assert abs(O.ref_achsDtemp(f) - O.achsDtemp(f)) <= 1e-12 * abs(O.achsDtemp(f))
        assert rel_err(O.ref_asxDtemp(f), O.asxDtemp(f)) < 1e-12
        assert rel_err(O.ref_achDtemp_cor(f), O.achDtemp_cor(f)) < 1e-12


# table of compute_fact cases: every branch and quirk of FF/ff.cpp:103-152
GRID = np.array([0.0, 0.5, 1.0, 1.5, 2.0])
FACT_CASES = [
    (1.2, 1, False),   # loop 1 inside interval 2
    (0.2, 1, False),   # loop 1 first interval -> remapped to nFreq-2 (extrapolated)
    (1.0, 1, False),   # loop 1 exactly on a grid point: not found -> remap
    (3.0, 1, False),   # loop 1 beyond grid -> remap
    (-0.7, 1, False),  # loop 1 negative: bracket -wx, weights (g1+wx)/d, (-g0-wx)/d
    (-0.1, 1, False),  # loop 1 negative first interval -> remap
    (0.0, 1, False),   # loop 1 wx == 0 takes the negative branch
    (0.7, 2, False),   # loop 2 wx>0: -1/2 weights
    (0.3, 2, False),   # loop 2 wx>0 first interval: NO remap (ifreq == -1 never holds)
    (1.5, 2, True),    # loop 2 wx>0 on grid point: stays 0
    (5.0, 2, True),    # loop 2 beyond grid: interval 0 extrapolation
    (-0.7, 2, True),   # loop 2 occupied, wx<=0
    (-0.2, 2, True),   # loop 2 occupied first interval -> remap
    (-0.7, 2, False),  # loop 2 unoccupied wx<=0: nothing written
]


@pytest.mark.parametrize("wx,loop,occ", FACT_CASES)
def test_compute_fact_oracle_vs_exact(wx, loop, occ):
    from fractions import Fraction

    from golden.make_exact_fixtures import compute_fact_exact

    f1, f2, i = O.compute_fact(wx, GRID, loop, occ)
    e1, e2, ei = compute_fact_exact(Fraction(wx), [Fraction(g) for g in GRID], loop, occ)
    assert i == ei
    assert abs(f1 - float(e1)) <= 2e-16 * max(1, abs(float(e1)))
    assert abs(f2 - float(e2)) <= 2e-16 * max(1, abs(float(e2)))
    # in/out semantics: nothing written for loop 2 unoccupied wx<=0
    if loop == 2 and not occ and wx <= 0:
        assert O.compute_fact(wx, GRID, loop, occ, ifreq=7, fact1=3.0, fact2=4.0) == (3.0, 4.0, 7)


def test_schDttt_cor_oracle_composes_achDtemp_cor():
    """achDtemp_cor is the sum of schDttt_corKernel1/2 over its bands: the per-band
    oracle must reproduce the (reference-pinned) achDtemp_cor oracle."""
    for seed, shape in ((11, (40, 3, 96, 24, 15, 2)), (12, (33, 5, 70, 9, 7, 4))):
        f = ff_with_kernel2_band(bgw.ff_inputs(*shape, seed))
        terms = ff_band_terms(f)
        assert {t[0] for t in terms} == {1, 2}  # both kernels take part
        total = sum(O.schDttt_cor(f, *t) for t in terms)
        ref = O.achDtemp_cor(f)[0]
        assert abs(total - ref) <= 1e-13 * abs(ref)


@pytest.mark.parametrize("kernel,scale", [(1, 1.0), (2, -0.5)])
def test_schDttt_cor_oracle_closed_form_reference_constants(kernel, scale):
    # FF/ff.cpp:584-633 constants: aqsn conj(aqsm) = 1/2, vcoul = 1, I_epsR - I_epsA = i in
    # every slice, so the sum is ngpown ncouls (1/2) i (fact1 + fact2) times the kernel's scale
    nb, ngpown, ncouls, nFreq = 9, 7, 50, 6
    f = bgw.ff_inputs(nb, 2, ncouls, ngpown, nFreq, 2, 0, kind="reference")
    fact1, fact2 = 0.375, -1.25
    want = ngpown * ncouls * 0.5j * (fact1 + fact2) * scale
    got = O.schDttt_cor(f, kernel, nb - 1, nFreq - 2, fact1, fact2)
    assert abs(got - want) <= 1e-15 * abs(want)
