"""C-ABI library checks that need no GPU: it loads, exports every symbol that
include/bgw.h declares, and its host-only logic (seeded generator, compute_fact)
behaves; without a GPU every context creation fails loudly."""
import ctypes as C
import os

import numpy as np
import pytest

import bgw
import oracle as O
from test_oracle import FACT_CASES


def test_library_exports_every_header_symbol():
    L = bgw.lib()
    syms = bgw.header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_no_gpu_means_enodev():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert bgw.lib().bgw_create(C.byref(h), 0) == bgw.BGW_ENODEV
    with pytest.raises(bgw.BgwError):
        bgw.Context(0)


def test_null_context_is_einval():
    L = bgw.lib()
    assert L.bgw_set_precision(None, 0) == bgw.BGW_EINVAL
    assert L.bgw_synchronize(None) == bgw.BGW_EINVAL
    assert L.bgw_noflagOCC_solver(None, 1, 1, 1, *([None] * 9)) == bgw.BGW_EINVAL
    assert L.bgw_gpp_solver(None, 1, 0, 1, 1, *([None] * 8), 1.0, None, None, None, None) == bgw.BGW_EINVAL


def _unopened_context():
    """A Context whose handle is null: every method that reaches the C ABI would get
    BGW_EINVAL, so a ValueError from it can only come from the binding's own checks."""
    ctx = bgw.Context.__new__(bgw.Context)
    ctx._h = C.c_void_p()
    return ctx


@pytest.mark.parametrize("key,cut", [("I_eps", 1), ("wtilde", 17), ("aqsm", 1), ("aqsn", 5), ("vcoul", 1),
                                     ("inv", 1), ("indinv", 1), ("wx", 3)])
def test_short_gpp_arrays_are_rejected_before_the_call(key, cut):
    d = bgw.gpp_inputs(4, 2, 64, 8, 11)
    d[key] = d[key][:-cut]
    ctx = _unopened_context()
    with pytest.raises(ValueError, match="elements, the call reads"):
        ctx.gpp_solver(d)
    with pytest.raises(ValueError, match="elements, the call reads"):
        ctx.gpp_backward(d)


def test_short_ff_arrays_are_rejected_before_the_call():
    f = bgw.ff_inputs(6, 2, 64, 8, 4, 2, 21)
    ctx = _unopened_context()
    for key, op in (("I_epsR", "achsDtemp_Kernel"), ("I_epsA", "asxDtemp_Kernel"), ("aqsm", "achDtemp_cor_Kernel"),
                    ("ekq", "achDtemp_cor_Kernel"), ("dFreqGrid", "asxDtemp_Kernel")):
        g = dict(f)
        if key in ("I_epsR", "I_epsA"):
            g[key] = g[key][: 8 * 64]  # one slice only
        elif key == "aqsm":
            g[key] = g[key][: 64 * 5]
        else:
            g[key] = g[key][:1]
        with pytest.raises(ValueError, match="elements, the call reads"):
            getattr(ctx, op)(g)
    # schDttt_corKernel reads slices ifreq, ifreq + 1 and band row n1
    with pytest.raises(ValueError, match="I_epsR_array"):
        ctx.schDttt_corKernel(1, f, 2, 3, 0.5, 0.5)
    with pytest.raises(ValueError, match="aqsmtemp"):
        ctx.schDttt_corKernel(1, f, 6, 0, 0.5, 0.5)
    # a short wx_array in the reference-signature solver
    d = bgw.gpp_inputs(4, 2, 64, 8, 11)
    with pytest.raises(ValueError, match="wx_array"):
        ctx.noflagOCC_solver(4, 8, 64, d["inv"], d["indinv"], d["wx"][:2], d["wtilde"], d["aqsm"], d["aqsn"],
                             d["I_eps"], d["vcoul"])


def test_synth_is_deterministic_and_in_range():
    a = bgw.synth_complex(17, (0.5, 3.0), (-0.1, 0.0), 10000)
    b = bgw.synth_complex(17, (0.5, 3.0), (-0.1, 0.0), 10000)
    assert np.array_equal(a, b)
    assert a.real.min() >= 0.5 and a.real.max() < 3.0 and a.imag.min() >= -0.1 and a.imag.max() < 0.0
    # prefix property of the counter-based stream
    assert np.array_equal(bgw.synth_uniform(3, 0, 1, 50), bgw.synth_uniform(3, 0, 1, 100)[:50])
    p = bgw.synth_permutation(9, 1000)
    assert sorted(p.tolist()) == list(range(1000))
    idx = bgw.synth_index(4, 37, 5000)
    assert idx.min() >= 0 and idx.max() < 37
    # uniformity sanity
    u = bgw.synth_uniform(5, 0, 1, 200000)
    assert abs(u.mean() - 0.5) < 5e-3


def test_gpp_inputs_match_benchmark_distributions():
    d = bgw.gpp_inputs(8, 2, 512, 256, 1)
    assert d["indinv"][-1] == 511 and sorted(d["indinv"][:-1].tolist()) == list(range(512))
    assert np.all(d["inv"] < 512)
    wx = d["wx"].reshape(8, 3)
    assert np.allclose(wx[:, 1] - wx[:, 0], 1.0) and np.allclose(wx[:, 1], 10 - d["ek"])
    assert d["I_eps"].real.min() >= 0 and d["I_eps"].real.max() < 0.5


@pytest.mark.parametrize("wx,loop,occ", FACT_CASES)
def test_product_compute_fact_matches_oracle(wx, loop, occ):
    grid = np.array([0.0, 0.5, 1.0, 1.5, 2.0])
    for init in ((0, 0.0, 0.0), (2, 3.0, 4.0), (7, 3.0, 4.0)):
        o = O.compute_fact(wx, grid, loop, occ, *init)
        if o is None:  # unbracketed wx with a stale ifreq beyond the grid: refused, nothing read
            with pytest.raises(bgw.BgwError) as e:
                bgw.compute_fact(wx, grid, loop, occ, *init)
            assert e.value.code == bgw.BGW_EINVAL
        else:
            assert bgw.compute_fact(wx, grid, loop, occ, *init) == o


def test_compute_fact_never_reads_past_the_grid():
    """wx on a grid point leaves the incoming ifreq in place; the reference would then
    index the grid with it.  An out-of-grid ifreq is refused (BGW_EINVAL) every time,
    an in-grid one is used, like the reference."""
    grid = np.array([0.0, 0.5, 1.0, 1.5, 2.0])
    for _ in range(50):
        with pytest.raises(bgw.BgwError):
            bgw.compute_fact(1.0, grid, 2, False, 7, 3.0, 4.0)  # loop 2, wx > 0 on a grid point
        assert O.compute_fact(1.0, grid, 2, False, 7, 3.0, 4.0) is None
    f1, f2, i = bgw.compute_fact(1.0, grid, 2, False, 1, 3.0, 4.0)
    assert i == 1 and (f1, f2) == (-0.5 * (1.0 - 1.0) / 0.5, -0.5 * (1.0 - 0.5) / 0.5)


def test_product_compute_fact_sweep_matches_oracle():
    rng = np.random.default_rng(0)
    grid = np.sort(rng.uniform(0, 3, 9))
    grid[0] = 0.0
    for wx in np.concatenate([rng.uniform(-4, 4, 300), grid, -grid]):
        for loop in (1, 2):
            for occ in (False, True):
                assert bgw.compute_fact(wx, grid, loop, occ) == O.compute_fact(wx, grid, loop, occ)


def test_reference_shim_links_after_the_reference_headers():
    """host/bgw_reference_api.hpp included after the reference's common/Defines.h, which
    declares noflagOCC_solver and schDttt_corKernel1/2 as plain functions (:36-62): every
    kernel call of the reference API compiles and links against libbgw, in both of the
    reference's complex modes (oracle/Makefile, target shim-check; nothing is run)."""
    import subprocess
    from pathlib import Path

    odir = Path(O.__file__).resolve().parent
    ref = subprocess.run(["make", "-s", "-C", str(odir), "print-ref"], capture_output=True, text=True).stdout.strip()
    if not (Path(ref) / "common" / "Defines.h").exists():
        pytest.skip("the reference tree is not present")
    r = subprocess.run(["make", "-s", "-B", "-C", str(odir), "shim-check"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_binding_converts_and_keeps_call_arguments():
    """bgw._Hold, which every host-array method of bgw.Context passes its arrays through:
    it rejects dtypes the C ABI cannot take as they are (an interleaved float64 buffer as a
    complex array, a complex array as a real one, indices beyond int32) and keeps the
    contiguous copies it makes alive, so the pointer it returns stays valid for the call."""
    import gc

    h = bgw._Hold()
    with pytest.raises(TypeError):
        h.c128(np.zeros(4))
    with pytest.raises(TypeError):
        h.f64(np.zeros(4, np.complex128))
    with pytest.raises(TypeError):
        h.i32(np.zeros(4))
    with pytest.raises(ValueError):
        h.i32(np.array([1, 2**40]))
    z = (np.arange(8) + 1j * np.arange(8))[::2]  # strided: a copy is made
    p = h.c128(z)
    ip = h.i32(np.arange(5, dtype=np.int64))
    del z
    gc.collect()
    got = np.ctypeslib.as_array((C.c_double * 8).from_address(p))
    assert np.array_equal(got, [0, 0, 2, 2, 4, 4, 6, 6])
    assert list((C.c_int * 5).from_address(ip)) == [0, 1, 2, 3, 4]
    assert h.c128(None) is None and h.i32(None) is None
