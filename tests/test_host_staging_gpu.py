"""Host-pointer entry points with arrays large enough to go through the context's
pinned staging ring (several 32 MiB chunks, the last one partial): their results
must equal, bit for bit, those of the device-pointer entry points on the same
data, for uploads (GPP forward, FF achDtemp_cor) and downloads (GPP adjoint
gradients of ngpown x ncouls entries).  The GPP host path also overlaps the upload
of wtilde / I_eps with main-kernel launches over row ranges (each continuing the
per-workgroup totals of the last), and without achstemp / asxtemp uploads only the
columns of aqsm / aqsn that sch_store1 reads."""
import numpy as np
import pytest

import bgw

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def test_staged_gpp_forward_matches_device_api(ctx):
    import torch

    # I_eps and wtilde 1000 x 24000 complex = 384 MB each (12 chunks); aqs 96 x 24000 (2 chunks)
    nb, nvband, ncouls, ngpown = 96, 2, 24000, 1000
    d = bgw.gpp_inputs(nb, nvband, ncouls, ngpown, 2024)
    h = ctx.gpp_solver(d)
    T = {k: _dev(torch, d[k]) for k in ("inv", "indinv", "wx", "wtilde", "aqsm", "aqsn", "I_eps", "vcoul")}
    out = {k: torch.zeros(n, dtype=torch.complex128, device="cuda:0") for k, n in
           (("achtemp", 3), ("acht_n1_loc", nb * 3), ("achstemp", 1), ("asxtemp", 3))}
    torch.cuda.synchronize()
    P = {k: v.data_ptr() for k, v in T.items()}
    ctx.gpp_solver_device(nb, nvband, ngpown, ncouls, P["inv"], P["indinv"], P["wx"], P["wtilde"], P["aqsm"],
                          P["aqsn"], P["I_eps"], P["vcoul"], 1.0, out["achtemp"].data_ptr(),
                          out["acht_n1_loc"].data_ptr(), out["achstemp"].data_ptr(), out["asxtemp"].data_ptr())
    ctx.synchronize()
    for k in h:
        assert np.array_equal(out[k].cpu().numpy().reshape(h[k].shape), h[k]), k


def test_gathered_columns_path_matches_device_api(ctx):
    """Without achstemp/asxtemp the host path uploads only the columns igp(p) of
    aqsm/aqsn (gathered on the host); noflagOCC_solver takes that path too."""
    import torch

    nb, nvband, ncouls, ngpown = 80, 2, 20000, 900
    d = bgw.gpp_inputs(nb, nvband, ncouls, ngpown, 515)
    h = ctx.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
    T = {k: _dev(torch, d[k]) for k in ("inv", "indinv", "wx", "wtilde", "aqsm", "aqsn", "I_eps", "vcoul")}
    ach = torch.zeros(3, dtype=torch.complex128, device="cuda:0")
    loc = torch.zeros(nb * 3, dtype=torch.complex128, device="cuda:0")
    torch.cuda.synchronize()
    P = {k: v.data_ptr() for k, v in T.items()}
    ctx.gpp_solver_device(nb, nvband, ngpown, ncouls, P["inv"], P["indinv"], P["wx"], P["wtilde"], P["aqsm"],
                          P["aqsn"], P["I_eps"], P["vcoul"], 1.0, ach.data_ptr(), loc.data_ptr())
    ctx.synchronize()
    assert np.array_equal(ach.cpu().numpy(), h["achtemp"])
    assert np.array_equal(loc.cpu().numpy().reshape(nb, 3), h["acht_n1_loc"])
    # the reference signature (wx[3], achtemp only)
    d3 = dict(d, wx=np.tile(d["wx"][:3], nb))
    T3 = _dev(torch, d3["wx"])
    torch.cuda.synchronize()
    ctx.gpp_solver_device(nb, nvband, ngpown, ncouls, P["inv"], P["indinv"], T3.data_ptr(), P["wtilde"],
                          P["aqsm"], P["aqsn"], P["I_eps"], P["vcoul"], 1.0, ach.data_ptr())
    ctx.synchronize()
    r = ctx.noflagOCC_solver(nb, ngpown, ncouls, d["inv"], d["indinv"], d["wx"][:3], d["wtilde"], d["aqsm"],
                             d["aqsn"], d["I_eps"], d["vcoul"])
    assert np.array_equal(ach.cpu().numpy(), r)


def test_staged_gpp_backward_downloads_match_device_api(ctx):
    import torch

    # dL/dwtilde and dL/dI_eps: 400 x 24000 complex = 154 MB each (5 chunks, the last partial)
    nb, nvband, ncouls, ngpown = 32, 2, 24000, 400
    d = bgw.gpp_inputs(nb, nvband, ncouls, ngpown, 77)
    h = ctx.gpp_backward(d)
    T = {k: _dev(torch, d[k]) for k in ("inv", "indinv", "wx", "wtilde", "aqsm", "aqsn", "I_eps", "vcoul")}
    npg = ngpown * ncouls
    ds = torch.zeros(3, dtype=torch.complex128, device="cuda:0")
    dw = torch.zeros(npg, dtype=torch.complex128, device="cuda:0")
    dI = torch.zeros(npg, dtype=torch.complex128, device="cuda:0")
    torch.cuda.synchronize()
    P = {k: v.data_ptr() for k, v in T.items()}
    ctx.gpp_backward_device(nb, nvband, ngpown, ncouls, P["inv"], P["indinv"], P["wx"], P["wtilde"], P["aqsm"],
                            P["aqsn"], P["I_eps"], P["vcoul"], 1.0, ds.data_ptr(), dw.data_ptr(), dI.data_ptr())
    ctx.synchronize()
    assert np.array_equal(ds.cpu().numpy(), h["dsigma_domega"])
    assert np.array_equal(dw.cpu().numpy(), h["dL_dwtilde"])
    assert np.array_equal(dI.cpu().numpy(), h["dL_dI_eps"])
    assert np.count_nonzero(h["dL_dI_eps"]) == npg


def test_staged_ff_upload_matches_device_api(ctx):
    import torch

    # I_epsR / I_epsA: 6 x 200 x 20000 complex = 384 MB each
    f = bgw.ff_inputs(64, 2, 20000, 200, 6, 3, 9)
    h = ctx.achDtemp_cor_Kernel(f)
    keys = ("inv", "indinv", "ekq", "aqsm", "aqsn", "vcoul", "I_epsR", "I_epsA", "dFreqGrid")
    T = {k: _dev(torch, f[k]) for k in keys}
    out = torch.zeros(f["nfreqeval"], dtype=torch.complex128, device="cuda:0")
    torch.cuda.synchronize()
    ctx.ff_device("achDtemp_cor", f, {k: v.data_ptr() for k, v in T.items()}, out.data_ptr())
    ctx.synchronize()
    assert np.array_equal(out.cpu().numpy(), h)


def test_ff_host_transfers_only_the_slices_its_plan_reads(ctx):
    """The FF host path uploads only the frequency slices its bands interpolate
    between (and, for asxDtemp, only the valence rows of aqsm / aqsn).  A first call
    fills every slice of the staging buffers, and the non-valence rows of aqsm /
    aqsn, with NaN; a second call of the same shape with finite data must then read
    nothing but what it uploaded."""
    import torch

    nb, nvband, ncouls, ngpown, nFreq, nfe = 20, 2, 4000, 50, 6, 2
    f = bgw.ff_inputs(nb, nvband, ncouls, ngpown, nFreq, nfe, 31)
    # call 1: unoccupied bands with wx inside every interval of the grid (loop 2, wx > 0)
    grid = f["dFreqGrid"]
    mids = 0.5 * (grid[:-1] + grid[1:])
    nan = dict(f)
    nan["ekq"] = f["ekq"].copy()
    nan["ekq"][nvband:nvband + len(mids)] = f["freqevalmin"] + f["freqevalstep"] - mids
    nan["I_epsR"] = np.full_like(f["I_epsR"], np.nan)
    nan["I_epsA"] = np.full_like(f["I_epsA"], np.nan)
    # and the rows of aqsm / aqsn beyond the valence bands, which asxDtemp does not transfer
    for k in ("aqsm", "aqsn"):
        nan[k] = f[k].copy()
        nan[k][nvband * ncouls:] = np.nan
    ctx.achDtemp_cor_Kernel(nan)
    # call 2: finite data, two valence bands touching at most four slices
    h = ctx.asxDtemp_Kernel(f)
    assert np.all(np.isfinite(h))
    keys = ("inv", "indinv", "ekq", "aqsm", "aqsn", "vcoul", "I_epsR", "I_epsA", "dFreqGrid")
    T = {k: _dev(torch, f[k]) for k in keys}
    out = torch.zeros(nfe, dtype=torch.complex128, device="cuda:0")
    torch.cuda.synchronize()
    ctx.ff_device("asxDtemp", f, {k: v.data_ptr() for k, v in T.items()}, out.data_ptr())
    ctx.synchronize()
    assert np.array_equal(out.cpu().numpy(), h)
    # and achDtemp_cor on finite data after the NaN call
    h2 = ctx.achDtemp_cor_Kernel(f)
    assert np.all(np.isfinite(h2))


def test_staged_fp32_form_and_caller_stream(ctx):
    """The FP32-bulk main kernel through the ranged, carried launches of the host
    path equals its single device launch bit for bit, also when the context runs on
    a caller's stream and the host call follows an asynchronous device call."""
    import torch

    nb, nvband, ncouls, ngpown = 96, 2, 24000, 700
    d = bgw.gpp_inputs(nb, nvband, ncouls, ngpown, 99)
    T = {k: _dev(torch, d[k]) for k in ("inv", "indinv", "wx", "wtilde", "aqsm", "aqsn", "I_eps", "vcoul")}
    P = {k: v.data_ptr() for k, v in T.items()}
    s = torch.cuda.Stream()
    ctx.set_stream(s.cuda_stream)
    ctx.set_precision(bgw.BGW_FP32_BULK)
    try:
        ach = torch.zeros(3, dtype=torch.complex128, device="cuda:0")
        loc = torch.zeros(nb * 3, dtype=torch.complex128, device="cuda:0")
        torch.cuda.synchronize()
        ctx.gpp_solver_device(nb, nvband, ngpown, ncouls, P["inv"], P["indinv"], P["wx"], P["wtilde"], P["aqsm"],
                              P["aqsn"], P["I_eps"], P["vcoul"], 1.0, ach.data_ptr(), loc.data_ptr())
        # no synchronisation here: the host call is ordered after it on the same stream
        h = ctx.gpp_solver(d, outputs=("achtemp", "acht_n1_loc"))
        ctx.synchronize()
        assert np.array_equal(ach.cpu().numpy(), h["achtemp"])
        assert np.array_equal(loc.cpu().numpy().reshape(nb, 3), h["acht_n1_loc"])
    finally:
        ctx.set_precision(bgw.BGW_FP64)
        ctx.set_stream(None)


def test_two_contexts_in_two_threads(ctx):
    """Distinct contexts may be used from distinct threads (include/bgw.h): two threads,
    each with its own context, run the host GPP and FF entry points at the same time
    (ctypes releases the GIL during the calls) through their own pinned rings and copy
    streams, and get what one context gets alone, bit for bit."""
    import threading

    d = bgw.gpp_inputs(64, 2, 4000, 200, 7)  # wtilde / I_eps 12.8 MB each: staged
    f = bgw.ff_inputs(40, 3, 96, 24, 15, 2, 8)
    ref_g = ctx.gpp_solver(d)
    ref_f = ctx.achDtemp_cor_Kernel(f)
    out, errs = {}, []

    def work(i):
        try:
            with bgw.Context(0) as c:
                for _ in range(3):
                    g = c.gpp_solver(d)
                    ff = c.achDtemp_cor_Kernel(f)
                out[i] = (g, ff)
        except Exception as e:  # reported below, in the main thread
            errs.append(repr(e))

    ts = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=100)
    assert not errs, errs
    assert sorted(out) == [0, 1]
    for i in range(2):
        g, ff = out[i]
        for k in ref_g:
            assert np.array_equal(g[k], ref_g[k]), (i, k)
        assert np.array_equal(ff, ref_f), i
