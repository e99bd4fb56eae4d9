"""GPU parity of the GPP forward (bgw_noflagOCC_solver / bgw_gpp_solver) against
the compensated CPU oracle, closed forms and size-independent properties.

Tolerance: 1e-12 relative error per complex output (FP64 path)."""
import numpy as np
import pytest

import bgw
import oracle as O
from conftest import rel_err
from helpers import gpp_reference_achstemp, gpp_reference_closed_form, load_exact_gpp

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _check_all(r, o, tol=TOL):
    assert rel_err(r["achtemp"], o["achtemp"]) < tol
    assert rel_err(r["acht_n1_loc"], o["acht_n1_loc"]) < tol
    assert rel_err(r["achstemp"], o["achstemp"]) < tol
    assert rel_err(r["asxtemp"], o["asxtemp"]) < tol


@pytest.mark.parametrize("case", range(4))  # case 3: I_eps on both sides of achstemp's |I_eps| > to1 cut
def test_exact_fixture(ctx, case):
    d, exp = load_exact_gpp()[case]
    r = ctx.gpp_solver(d, occ=1.0)
    assert rel_err(r["achtemp"], exp["achtemp"]) < 1e-13
    assert rel_err(r["acht_n1_loc"], exp["acht_n1_loc"]) < 1e-13
    assert rel_err(r["achstemp"], exp["achstemp"]) < 1e-13
    assert rel_err(r["asxtemp"], exp["asxtemp"]) < 1e-13


@pytest.mark.parametrize("nb,ngpown,ncouls", [(1, 1, 1), (40, 50, 1000), (512, 7, 3000)])
def test_closed_form_reference_constants(ctx, nb, ngpown, ncouls):
    d = bgw.gpp_inputs(nb, min(2, nb), ncouls, ngpown, 0, kind="reference")
    r = ctx.gpp_solver(d)
    assert rel_err(r["achtemp"], gpp_reference_closed_form(nb, ngpown, ncouls)) < TOL
    assert rel_err(r["acht_n1_loc"].sum(0), r["achtemp"]) < 1e-14
    assert rel_err(r["achstemp"], [gpp_reference_achstemp(nb, ngpown, ncouls)]) < TOL
    ach = ctx.noflagOCC_solver(nb, ngpown, ncouls, d["inv"], d["indinv"], np.array([3.0, 4.0, 5.0]), d["wtilde"],
                               d["aqsm"], d["aqsn"], d["I_eps"], d["vcoul"])
    assert rel_err(ach, gpp_reference_closed_form(nb, ngpown, ncouls)) < TOL


def test_config0_vs_oracle(ctx):
    d = bgw.gpp_config_inputs(0)
    r = ctx.gpp_solver(d, occ=1.0)
    o = O.gpp_forward(d, occ=1.0)
    _check_all(r, o)


@pytest.mark.parametrize("nb", [2, 63, 65, 129, 200, 257, 300, 513])
def test_band_counts_vs_oracle(ctx, nb):
    # exercises every (threads, bands-per-lane) variant, partial waves and band tiles
    d = bgw.gpp_inputs(nb, min(3, nb), 777, 37, 1000 + nb)
    r = ctx.gpp_solver(d, occ=0.5)
    o = O.gpp_forward(d, occ=0.5)
    _check_all(r, o)


@pytest.mark.parametrize("ncouls,ngpown", [(1, 1), (255, 3), (2049, 5), (4097, 2)])
def test_g_tails_vs_oracle(ctx, ncouls, ngpown):
    d = bgw.gpp_inputs(70, 2, ncouls, ngpown, 77 + ncouls)
    _check_all(ctx.gpp_solver(d), O.gpp_forward(d))


def test_noflag_reference_signature_vs_reference_semantics(ctx):
    d = bgw.gpp_inputs(24, 2, 1500, 90, 5150)
    wx3 = d["wx"][:3].copy()
    ach = ctx.noflagOCC_solver(d["nb"], d["ngpown"], d["ncouls"], d["inv"], d["indinv"], wx3, d["wtilde"], d["aqsm"],
                               d["aqsn"], d["I_eps"], d["vcoul"])
    o = O.gpp_forward(dict(d, wx=np.tile(wx3, d["nb"])))
    assert rel_err(ach, o["achtemp"]) < TOL


def test_more_gprime_rows_than_one_grid_dimension(ctx):
    # ngpown above 65535 (a grid's row limit): the S-table prep launches in chunks
    d = bgw.gpp_inputs(3, 2, 4, 66000, 66)
    _check_all(ctx.gpp_solver(d), O.gpp_forward(d))


def test_reference_index_init_reaches_ncouls(ctx):
    # inv_igp_index[p] = (p+1)*ncouls/ngpown reaches ncouls for the last p
    # (GPP/gpp.cpp:263-269): indinv[ncouls] = ncouls-1 must be honoured
    d = bgw.gpp_inputs(16, 2, 600, 60, 4)
    d["inv"] = ((np.arange(60) + 1) * 600 // 60).astype(np.int32)
    assert d["inv"][-1] == 600
    _check_all(ctx.gpp_solver(d), O.gpp_forward(d))


def test_exact_path_tiles(ctx):
    # a wtilde exactly at a band frequency with a tiny imaginary part puts the
    # pole denominator below 2^-300: those tiles take the reference-arithmetic path
    d = bgw.gpp_inputs(8, 2, 600, 4, 8080)
    wx = d["wx"].reshape(8, 3)
    wt = d["wtilde"].reshape(4, 600)
    wt[1, 10] = wx[3, 1] + 1j * 2.0**-170
    wt[2, 599] = wx[0, 0] - 1j * 2.0**-160
    d["wtilde"] = wt.reshape(-1)
    r = ctx.gpp_solver(d)
    o = O.gpp_forward(d)
    assert np.all(np.isfinite(r["achtemp"]))
    _check_all(r, o)


def test_bitwise_reproducible(ctx):
    d = bgw.gpp_inputs(100, 2, 5000, 64, 11)
    a = ctx.gpp_solver(d)
    b = ctx.gpp_solver(d)
    for k in a:
        assert np.array_equal(a[k], b[k])


def test_host_validation(ctx):
    d = bgw.gpp_inputs(4, 2, 64, 8, 1)
    bad = dict(d, inv=d["inv"].copy())
    bad["inv"][3] = 65
    with pytest.raises(bgw.BgwError) as e:
        ctx.gpp_solver(bad)
    assert e.value.code == bgw.BGW_EINVAL
    bad = dict(d, indinv=d["indinv"].copy())
    bad["indinv"][5] = -1
    with pytest.raises(bgw.BgwError):
        ctx.gpp_solver(bad)
    with pytest.raises(bgw.BgwError):
        ctx.gpp_solver(dict(d, nvband=5))
    # the context stays usable after errors
    assert rel_err(ctx.gpp_solver(d)["achtemp"], O.gpp_forward(d)["achtemp"]) < TOL


def _dev(t):
    return t.data_ptr()


def test_device_api_matches_host_and_reports_bad_index(ctx):
    import torch

    d = bgw.gpp_inputs(64, 2, 900, 40, 321)
    dev = torch.device("cuda:0")
    T = {k: torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in
         ("inv", "indinv", "wx", "wtilde", "aqsm", "aqsn", "I_eps", "vcoul")}
    out = {k: torch.zeros(n, dtype=torch.complex128, device=dev) for k, n in
           (("achtemp", 3), ("acht_n1_loc", 64 * 3), ("achstemp", 1), ("asxtemp", 3))}
    torch.cuda.synchronize()
    ctx.gpp_solver_device(64, 2, 40, 900, _dev(T["inv"]), _dev(T["indinv"]), _dev(T["wx"]), _dev(T["wtilde"]),
                          _dev(T["aqsm"]), _dev(T["aqsn"]), _dev(T["I_eps"]), _dev(T["vcoul"]), 1.0,
                          _dev(out["achtemp"]), _dev(out["acht_n1_loc"]), _dev(out["achstemp"]), _dev(out["asxtemp"]))
    ctx.synchronize()
    h = ctx.gpp_solver(d)
    for k in h:
        assert np.array_equal(out[k].cpu().numpy().reshape(h[k].shape), h[k])
    T["inv"][0] = 10**6  # out of range: clamped on the device, reported at synchronize
    torch.cuda.synchronize()
    ctx.gpp_solver_device(64, 2, 40, 900, _dev(T["inv"]), _dev(T["indinv"]), _dev(T["wx"]), _dev(T["wtilde"]),
                          _dev(T["aqsm"]), _dev(T["aqsn"]), _dev(T["I_eps"]), _dev(T["vcoul"]), 1.0,
                          _dev(out["achtemp"]), None, None, None)
    with pytest.raises(bgw.BgwError) as e:
        ctx.synchronize()
    assert e.value.code == bgw.BGW_EINVAL
    ctx.synchronize()  # flag cleared


@pytest.mark.slow
def test_config1_reduced_bands_vs_oracle(ctx):
    # the benchmark shape (ngpown 1638, ncouls 32768) with 12 of its 512 bands
    d = bgw.gpp_config_inputs(1, nb=12)
    _check_all(ctx.gpp_solver(d), O.gpp_forward(d))


@pytest.mark.slow
def test_config1_full_closed_form(ctx):
    # the reference's default run (512 2 32768 20) with its constant inputs:
    # exact achtemp[0] = (-264241152, 1321205760)
    d = bgw.gpp_inputs(512, 2, 32768, 1638, 0, kind="reference")
    r = ctx.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    cf = gpp_reference_closed_form(512, 1638, 32768)
    assert cf[0] == complex(-264241152.0, 1321205760.0)
    assert rel_err(r["achtemp"], cf) < TOL
    assert rel_err(r["acht_n1_loc"], np.tile(cf / 512, (512, 1))) < TOL


@pytest.mark.slow
def test_config1_full_properties(ctx):
    d = bgw.gpp_config_inputs(1)
    a = ctx.gpp_solver(d)
    b = ctx.gpp_solver(d)
    for k in a:
        assert np.array_equal(a[k], b[k])  # run-to-run bitwise
    assert rel_err(a["acht_n1_loc"].sum(0), a["achtemp"]) < 1e-14
    # per-band values of 12 bands against the oracle on the same inputs
    sub = dict(d, nb=12, aqsm=d["aqsm"][: 12 * 32768].copy(), aqsn=d["aqsn"][: 12 * 32768].copy(),
               wx=d["wx"][:36].copy())
    o = O.gpp_forward(sub, want=("achtemp", "acht_n1_loc"))
    assert rel_err(a["acht_n1_loc"][:12], o["acht_n1_loc"]) < TOL


@pytest.mark.slow
def test_config1_full_all_outputs_vs_oracle(ctx):
    """The benchmark workload at full size with all four outputs (the static COH term's
    masked MFMA contraction and the SX term included) against the oracle on the same
    inputs; the oracle's compensated sums take a few tens of seconds on the host cores."""
    d = bgw.gpp_config_inputs(1)
    _check_all(ctx.gpp_solver(d), O.gpp_forward(d))


@pytest.mark.slow
@pytest.mark.skipif(not O.REF_OMP.exists(), reason="reference build (oracle/_ref) absent")
def test_config1_full_vs_reference_build(ctx):
    """The benchmark workload at full size (512 bands, ngpown 1638, ncouls 32768, wx per
    band): every band's acht_n1_loc against the reference's own noflagOCC_solver
    (GPP/gpp.cpp:307-402, built from its sources into oracle/_ref, OpenMP) called once
    per band on the same inputs.  The reference sums 5.4e7 terms per band naively, about
    1e-13 relative from the exact sum."""
    d = bgw.gpp_config_inputs(1)
    a = ctx.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    r = O.ref_gpp_per_band(d, kind="omp")
    assert np.max(np.abs(a["acht_n1_loc"] - r) / np.abs(r)) < 1e-11
    assert rel_err(a["achtemp"], r.sum(0)) < 1e-11


def test_binding_takes_int64_indices_and_strided_arrays(ctx):
    """The Python binding converts int64 index arrays and non-contiguous views before the
    call (and keeps the copies alive through it): results equal those of canonical inputs."""
    d = bgw.gpp_inputs(16, 2, 600, 40, 3)
    ref = ctx.gpp_solver(d)
    e = dict(d, inv=d["inv"].astype(np.int64), indinv=d["indinv"].astype(np.int64))
    for k in ("wtilde", "aqsm", "aqsn", "I_eps"):
        big = np.zeros(2 * d[k].size, np.complex128)
        big[::2] = d[k]
        e[k] = big[::2]
        assert not e[k].flags["C_CONTIGUOUS"]
    got = ctx.gpp_solver(e)
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k
    f = bgw.ff_inputs(20, 2, 96, 24, 6, 2, 3)
    g = dict(f, inv=f["inv"].astype(np.int64), I_epsA=np.asfortranarray(f["I_epsA"].reshape(6, 24, 96)))
    assert np.array_equal(ctx.achDtemp_cor_Kernel(g), ctx.achDtemp_cor_Kernel(f))
