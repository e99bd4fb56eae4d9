"""FF kernels on a long frequency grid with few touched slices: the shape of the reference's
compute-bound example (FF/ff.cpp:56-57, `ffkernel.x 8 2 1000 60 20000 30000`: nvband 2,
ngpown 60, nFreq 20000), here with ncouls 64 so the slices fit the test.  asxDtemp's plan
then has 40000 jobs of which at most 2*nvband hold rows: the streaming contraction launches
workgroups for the live jobs only, the sparse MFMA plans bound their row tiles by the live
jobs, and the host path stages only the touched slices (packed).  Tolerance 1e-12 relative
against the oracle; host and device forms equal bit for bit."""
import numpy as np
import pytest

import bgw
import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12
NFREQ = 20000
DEV_KEYS = ("inv", "indinv", "ekq", "aqsm", "aqsn", "vcoul", "I_epsR", "I_epsA", "dFreqGrid", "pref")


def _device_call(ctx, op, f):
    """The _device entry point on full arrays (every slice resident), as a caller holding
    them in HBM makes it; the plan reads slice j % nFreq of them."""
    import torch

    T = {k: torch.from_numpy(np.ascontiguousarray(f[k])).to("cuda:0") for k in DEV_KEYS}
    out = torch.zeros(f["nfreqeval"], dtype=torch.complex128, device="cuda:0")
    torch.cuda.synchronize()
    ctx.ff_device(op, f, {k: v.data_ptr() for k, v in T.items()}, out.data_ptr())
    ctx.synchronize()
    r = out.cpu().numpy()
    del T, out
    torch.cuda.empty_cache()
    return r


def _timed(ctx, name, call):
    ctx.reset_timing()
    ctx.set_timing(True)
    try:
        r = call()
    finally:
        ctx.set_timing(False)
    return r, ctx.kernel_time(name)


@pytest.fixture(scope="module")
def grid_nv2():
    return bgw.ff_inputs(8, 2, 64, 60, NFREQ, 3, 20000)


def test_asxDtemp_streaming_long_grid(ctx, grid_nv2):
    f = grid_nv2
    got, (ms, n) = _timed(ctx, "contract_ff_asx", lambda: ctx.asxDtemp_Kernel(f))
    assert rel_err(got, O.asxDtemp(f)) < TOL
    # 40000 jobs, at most 4 with rows: 4 x 30 workgroups, microseconds of work (launching a
    # workgroup per job would be 1.2M workgroups)
    assert n == 1 and ms < 2.0, ms
    assert np.array_equal(_device_call(ctx, "asxDtemp", f), got)


def test_achDtemp_cor_long_grid(ctx, grid_nv2):
    f = grid_nv2
    got, (ms, n) = _timed(ctx, "contract_ff_cor", lambda: ctx.achDtemp_cor_Kernel(f))
    assert rel_err(got, O.achDtemp_cor(f)) < TOL
    assert n == 1 and ms < 5.0, ms  # 20000 jobs, 8 bands: the row tiles of at most 16 live jobs
    assert np.array_equal(_device_call(ctx, "achDtemp_cor", f), got)


def test_asxDtemp_mfma_long_grid(ctx):
    """nvband 6 takes the sparse MFMA plan (more rows per job than the streaming form
    handles): its row-tile bound is the live jobs plus the row spill, not the job count."""
    f = bgw.ff_inputs(8, 6, 64, 60, NFREQ, 2, 20001)
    got, (ms, n) = _timed(ctx, "contract_ff_asx", lambda: ctx.asxDtemp_Kernel(f))
    assert rel_err(got, O.asxDtemp(f)) < TOL
    assert n == 1 and ms < 5.0, ms
    assert np.array_equal(_device_call(ctx, "asxDtemp", f), got)
