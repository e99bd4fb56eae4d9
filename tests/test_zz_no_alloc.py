"""bgw_gpp_solver_device_noalloc: the allocation-free form of bgw_gpp_solver_device.
Its results are compared with that entry point's on the same device arrays, in every
arithmetic form, over two calls in a row (the second one smaller), and the device's free
memory is checked to be unchanged across the allocation-free calls."""
import numpy as np
import pytest

import bgw

pytestmark = pytest.mark.gpu


def _device_inputs(torch, d):
    dev = torch.device("cuda:0")
    return {k: torch.from_numpy(np.ascontiguousarray(d[k])).to(dev) for k in
            ("inv", "indinv", "wx", "wtilde", "aqsm", "aqsn", "I_eps", "vcoul")}


def _outputs(torch, d):
    return (torch.zeros(3, dtype=torch.complex128, device="cuda:0"),
            torch.zeros(d["nb"] * 3, dtype=torch.complex128, device="cuda:0"))


def _run(torch, ctx, d, T, out, noalloc):
    nb, ng, nc = d["nb"], d["ngpown"], d["ncouls"]
    ach, loc = out
    P = [T[k].data_ptr() for k in ("inv", "indinv", "wx", "wtilde", "aqsm", "aqsn", "I_eps", "vcoul")]
    if noalloc:
        ctx.gpp_solver_device_noalloc(nb, ng, nc, *P, ach.data_ptr(), loc.data_ptr())
    else:
        ctx.gpp_solver_device(nb, d["nvband"], ng, nc, *P, 1.0, ach.data_ptr(), loc.data_ptr(), None, None)
    ctx.synchronize()
    return ach.cpu().numpy(), loc.cpu().numpy()


@pytest.mark.parametrize("precision", [bgw.BGW_FP64, bgw.BGW_FP32_BULK, bgw.BGW_FP64_MOMENTS])
def test_noalloc_matches_device_entry_point(precision):
    torch = pytest.importorskip("torch")
    cases = [bgw.gpp_inputs(300, 2, 4000, 160, 501), bgw.gpp_inputs(40, 2, 1100, 37, 502)]  # second call smaller
    with bgw.Context(0) as ctx:
        ctx.set_precision(precision)
        Ts = [_device_inputs(torch, d) for d in cases]
        outs = [_outputs(torch, d) for d in cases]
        refs = [_run(torch, ctx, d, T, o, noalloc=False) for d, T, o in zip(cases, Ts, outs)]
        for o in outs:
            o[0].fill_(np.nan)
            o[1].fill_(np.nan)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        got = [_run(torch, ctx, d, T, o, noalloc=True) for d, T, o in zip(cases, Ts, outs)]
        free1 = torch.cuda.mem_get_info()[0]
        for (a_ref, l_ref), (a, l) in zip(refs, got):
            # same kernels, same launch shape, same summation order: equal bit for bit
            assert np.array_equal(a, a_ref)
            assert np.array_equal(l, l_ref)
            assert np.all(np.isfinite(a)) and np.max(np.abs(a)) > 0
        assert free1 == free0


def test_noalloc_rejects_what_does_not_fit():
    torch = pytest.importorskip("torch")
    with bgw.Context(0) as ctx:
        # 4096 bands x 4096 G' needs a 256 MiB S table alone: refused before any launch
        with pytest.raises(bgw.BgwError, match="scratch") as e:
            ctx.gpp_solver_device_noalloc(4096, 4096, 8192, *([1] * 8), 1, None)
        assert e.value.code == bgw.BGW_EINVAL
        with pytest.raises(bgw.BgwError) as e:
            ctx.gpp_solver_device_noalloc(8, 4, 16, *([None] * 8), None, None)
        assert e.value.code == bgw.BGW_EINVAL
        ctx.synchronize()


@pytest.mark.slow
@pytest.mark.parametrize("precision", [bgw.BGW_FP64, bgw.BGW_FP32_BULK, bgw.BGW_FP64_MOMENTS])
def test_noalloc_benchmark_workload(precision):
    """The benchmark workload (512 bands, ngpown 1638, ncouls 32768) fits the 64 MiB scratch
    array in every arithmetic form (37-51 MiB) and matches the device entry point bit for bit."""
    torch = pytest.importorskip("torch")
    d = bgw.gpp_config_inputs(1)
    with bgw.Context(0) as ctx:
        ctx.set_precision(precision)
        T = _device_inputs(torch, d)
        out = _outputs(torch, d)
        ref = _run(torch, ctx, d, T, out, noalloc=False)
        out[0].fill_(np.nan)
        out[1].fill_(np.nan)
        torch.cuda.synchronize()
        got = _run(torch, ctx, d, T, out, noalloc=True)
        assert np.array_equal(got[0], ref[0])
        assert np.array_equal(got[1], ref[1])
