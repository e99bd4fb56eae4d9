"""GPU parity of the FF kernels (bgw_achsDtemp_Kernel, bgw_asxDtemp_Kernel,
bgw_achDtemp_cor_Kernel) against the oracle, exact fixtures and closed forms.
Tolerance 1e-12 relative (FP64 MFMA path)."""
import json

import numpy as np
import pytest

import bgw
import oracle as O
from conftest import rel_err
from helpers import GOLDEN, ff_band_terms, ff_reference_closed_form, ff_with_kernel2_band, load_exact_ff

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _all(ctx, f):
    return ctx.achsDtemp_Kernel(f), ctx.asxDtemp_Kernel(f), ctx.achDtemp_cor_Kernel(f)


@pytest.mark.parametrize("case", range(2))
def test_exact_fixture(ctx, case):
    f, exp = load_exact_ff()[case]
    achs, asx, cor = _all(ctx, f)
    assert abs(achs - exp["achsDtemp"]) <= 1e-13 * abs(exp["achsDtemp"])
    assert np.all(np.abs(asx - exp["asxDtemp"]) <= 1e-13 * abs(exp["asxDtemp"]))
    assert np.all(np.abs(cor - exp["achDtemp_cor"]) <= 1e-13 * abs(exp["achDtemp_cor"]))


@pytest.mark.parametrize("nb,nvband,ncouls,ngpown,nFreq", [(20, 4, 64, 16, 15), (150, 40, 300, 70, 6)])
def test_closed_form_reference_constants(ctx, nb, nvband, ncouls, ngpown, nFreq):
    f = bgw.ff_inputs(nb, nvband, ncouls, ngpown, nFreq, 4, 0, kind="reference")
    achs, asx, cor = ff_reference_closed_form(nb, nvband, ncouls, ngpown, nFreq)
    g = _all(ctx, f)
    assert abs(g[0] - achs) <= TOL * abs(achs)
    assert np.all(np.abs(g[1] - asx) <= TOL * abs(asx))
    assert np.all(np.abs(g[2] - cor) <= TOL * abs(cor))


# nvband up to 4 takes asxDtemp's streaming form, above it the MFMA tiles; (60, 4, ...)
# is the largest streaming case, with an odd G' count and a ragged last G chunk
@pytest.mark.parametrize("shape", [(40, 3, 96, 24, 15, 2), (60, 4, 700, 51, 9, 2), (130, 5, 1000, 65, 15, 3),
                                   (77, 77, 530, 33, 4, 1)])
def test_random_vs_oracle(ctx, shape):
    f = bgw.ff_inputs(*shape, 900 + shape[0])
    achs, asx, cor = _all(ctx, f)
    assert abs(achs - O.achsDtemp(f)) <= TOL * abs(O.achsDtemp(f))
    assert rel_err(asx, O.asxDtemp(f)) < TOL
    assert rel_err(cor, O.achDtemp_cor(f)) < TOL
    # every evaluation frequency carries the same value (wx has no iw term, FF/ff.cpp:245,337)
    assert np.all(asx == asx[0]) and np.all(cor == cor[0])


def test_committed_reference_vectors(ctx):
    vec = json.loads((GOLDEN / "ref_vectors.json").read_text())
    for c in vec["ff"]:
        f = bgw.ff_inputs(*c["shape"], c["seed"])
        achs, asx, cor = _all(ctx, f)
        ra = complex(*c["achsDtemp"])
        assert abs(achs - ra) <= 1e-11 * abs(ra)
        rx = np.asarray(c["asxDtemp"])
        assert rel_err(asx, rx[:, 0] + 1j * rx[:, 1]) < 1e-11
        rc = np.asarray(c["achDtemp_cor"])
        assert rel_err(cor, rc[:, 0] + 1j * rc[:, 1]) < 1e-11


def test_large_frequency_grid_vs_oracle(ctx):
    # the reference's "compute bound" regime has many frequencies (FF/ff.cpp:56-57):
    # 2 * 5000 jobs in the plan and the tile prefix
    f = bgw.ff_inputs(8, 2, 100, 10, 5000, 7, 5000)
    achs, asx, cor = _all(ctx, f)
    assert abs(achs - O.achsDtemp(f)) <= TOL * abs(O.achsDtemp(f))
    assert rel_err(asx, O.asxDtemp(f)) < TOL
    assert rel_err(cor, O.achDtemp_cor(f)) < TOL


@pytest.mark.parametrize("shape", [(10, 0, 64, 16, 5, 2), (12, 3, 100, 20, 2, 3), (1, 1, 50, 7, 6, 2),
                                   (1, 0, 50, 7, 6, 2)])
def test_edge_shapes_vs_oracle(ctx, shape):
    # no valence bands (asxDtemp's band loop is empty: exactly zero), the two-point
    # frequency grid, and a single band (occupied or not)
    f = bgw.ff_inputs(*shape, 4242)
    achs, asx, cor = _all(ctx, f)
    assert abs(achs - O.achsDtemp(f)) <= TOL * abs(O.achsDtemp(f))
    oa, oc = O.asxDtemp(f), O.achDtemp_cor(f)
    if shape[1] == 0:
        assert np.all(asx == 0)
    else:
        assert rel_err(asx, oa) < TOL
    if np.all(oc == 0):
        assert np.all(cor == 0)
    else:
        assert rel_err(cor, oc) < TOL


def test_validation(ctx):
    f = bgw.ff_inputs(10, 2, 20, 30, 5, 2, 3)
    with pytest.raises(bgw.BgwError) as e:  # ngpown > ncouls (FF/ff.cpp:492-499)
        ctx.achDtemp_cor_Kernel(f)
    assert e.value.code == bgw.BGW_EINVAL


@pytest.mark.slow
def test_config4_shape_vs_oracle(ctx):
    # config 4 (nb 512, ngpown 819, nFreq 15, nfreqeval 10) with ncouls cut to 2048
    f = bgw.ff_config4_inputs(ncouls=2048)
    achs, asx, cor = _all(ctx, f)
    assert abs(achs - O.achsDtemp(f)) <= TOL * abs(O.achsDtemp(f))
    assert rel_err(asx, O.asxDtemp(f)) < TOL
    assert rel_err(cor, O.achDtemp_cor(f)) < TOL


@pytest.mark.skipif(not O.REF_OMP.exists(), reason="reference build (oracle/_ref) absent")
def test_full_config4_vs_reference_build(ctx):
    """The full config-4 inputs (nb 512, ncouls 16384, ngpown 819, nFreq 15, nfreqeval 10)
    through the host entry points, against the reference's own FF kernels
    (FF/ff.cpp:171-451, built from its sources into oracle/_ref) on the same inputs.
    The reference sums naively in FP64, about 1e-13 relative from the exact sum here;
    its asxDtemp / achDtemp_cor repeat one value for every frequency, so it runs them
    with nfreqeval 1."""
    f = bgw.ff_config4_inputs()
    f1 = dict(f, nfreqeval=1)
    achs, asx, cor = _all(ctx, f)
    r_achs = O.ref_achsDtemp(f, kind="omp")
    r_asx = O.ref_asxDtemp(f1, kind="omp")[0]
    r_cor = O.ref_achDtemp_cor(f1, kind="omp")[0]
    assert abs(achs - r_achs) <= 1e-11 * abs(r_achs)
    assert np.all(np.abs(asx - r_asx) <= 1e-11 * abs(r_asx))
    assert np.all(np.abs(cor - r_cor) <= 1e-11 * abs(r_cor))


@pytest.mark.parametrize("shape", [(40, 3, 96, 24, 15, 2), (33, 5, 70, 9, 7, 4), (6, 2, 1000, 65, 4, 1)])
def test_band_kernels_vs_oracle(ctx, shape):
    """bgw_schDttt_corKernel1/2 (FF/ff.cpp:381-451) on host arrays against the oracle, on
    the first and the last slice pair and band (G' counts 24, 9 and 65: partial tiles)."""
    f = bgw.ff_inputs(*shape, 2027)
    nb, nFreq = f["nb"], f["nFreq"]
    for kernel in (1, 2):
        for n1, ifreq, f1, f2 in ((0, 0, 0.3, 0.7), (nb - 1, nFreq - 2, -1.25, 0.5)):
            got = ctx.schDttt_corKernel(kernel, f, n1, ifreq, f1, f2)
            ref = O.schDttt_cor(f, kernel, n1, ifreq, f1, f2)
            assert abs(got - ref) <= TOL * abs(ref), (kernel, n1, ifreq)


def test_band_kernels_compose_achDtemp_cor(ctx):
    """Summed over achDtemp_cor's bands (compute_fact's intervals and weights), the
    per-band entry points give achDtemp_cor."""
    f = ff_with_kernel2_band(bgw.ff_inputs(40, 3, 96, 24, 15, 2, 11))
    terms = ff_band_terms(f)
    assert {t[0] for t in terms} == {1, 2}
    total = sum(ctx.schDttt_corKernel(k, f, n1, ifreq, f1, f2) for k, n1, ifreq, f1, f2 in terms)
    cor = ctx.achDtemp_cor_Kernel(f)[0]
    assert abs(total - cor) <= TOL * abs(cor)


def test_band_kernels_device_form_and_validation(ctx):
    import torch

    f = bgw.ff_inputs(20, 2, 300, 30, 6, 1, 5)
    keys = ("inv", "indinv", "aqsm", "aqsn", "vcoul", "I_epsR", "I_epsA")
    T = {k: torch.from_numpy(np.ascontiguousarray(f[k])).to("cuda:0") for k in keys}
    out = torch.zeros(1, dtype=torch.complex128, device="cuda:0")
    torch.cuda.synchronize()
    ptrs = {k: v.data_ptr() for k, v in T.items()}
    for kernel in (1, 2):
        h = ctx.schDttt_corKernel(kernel, f, 7, 3, 0.25, -0.75)
        ctx.schDttt_corKernel_device(kernel, f, ptrs, 7, 3, 0.25, -0.75, out.data_ptr())
        ctx.synchronize()
        assert out.cpu().numpy()[0] == h  # same kernels on the same data: bit for bit
    for bad in ({"ifreq": -1}, {"n1": -1}):
        args = dict(n1=0, ifreq=0) | bad
        with pytest.raises(bgw.BgwError) as ei:
            ctx.schDttt_corKernel(1, f, args["n1"], args["ifreq"], 0.5, 0.5)
        assert ei.value.code == bgw.BGW_EINVAL
    # ngpown > ncouls (FF/ff.cpp:492-499): the C ABI refuses it before it reads an array
    g = dict(f, ngpown=f["ncouls"] + 1)
    with pytest.raises(bgw.BgwError) as ei:
        ctx.schDttt_corKernel_device(2, g, ptrs, 0, 0, 0.5, 0.5, out.data_ptr())
    assert ei.value.code == bgw.BGW_EINVAL
    # and the host form's binding refuses the arrays, shorter than such a call would read
    with pytest.raises(ValueError, match="elements, the call reads"):
        ctx.schDttt_corKernel(2, g, 0, 0, 0.5, 0.5)
    bad_inv = dict(f, inv=f["inv"].copy())
    bad_inv["inv"][3] = f["ncouls"]
    with pytest.raises(bgw.BgwError) as ei:
        ctx.schDttt_corKernel(1, bad_inv, 0, 0, 0.5, 0.5)
    assert ei.value.code == bgw.BGW_EINVAL

