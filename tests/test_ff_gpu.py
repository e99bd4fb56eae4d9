"""GPU parity of the FF kernels (bgw_achsDtemp_Kernel, bgw_asxDtemp_Kernel,
bgw_achDtemp_cor_Kernel) against the oracle, exact fixtures and closed forms.
Tolerance 1e-12 relative (FP64 MFMA path)."""
import json

import numpy as np
import pytest

import bgw
import oracle as O
from conftest import rel_err
from helpers import GOLDEN, ff_reference_closed_form, load_exact_ff

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


@pytest.mark.parametrize("shape", [(40, 3, 96, 24, 15, 2), (130, 5, 1000, 65, 15, 3), (77, 77, 530, 33, 4, 1)])
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
    # the reference's "compute bound" regime has many frequencies (FF/ff.cpp:56-57);
    # 2*nFreq slices exceed the plan kernel's LDS counters and use its global path
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
