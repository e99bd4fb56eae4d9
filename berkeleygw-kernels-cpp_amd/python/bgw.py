"""ctypes binding of libbgw (the C ABI declared in include/bgw.h).

Host-pointer calls take numpy arrays in the reference layout (row-major,
interleaved complex128); ``*_device`` calls take raw device addresses (ints), for
example ``torch.Tensor.data_ptr()``.  There is no CPU fallback: without the
built library, or without a GPU, every compute call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parents[1]
LIB_PATH = PKG / "build" / "libbgw.so"
HEADER = PKG.parent / "include" / "bgw.h"

BGW_OK, BGW_EINVAL, BGW_ENOMEM, BGW_EHIP, BGW_ENODEV, BGW_EIO = 0, -22, -12, -5, -19, -74
BGW_FP64, BGW_FP32_BULK, BGW_FP64_MOMENTS = 0, 1, 2
BGW_INPUTS_GPP, BGW_INPUTS_FF, BGW_INPUTS_OTHER = 1, 2, 3
BGW_DT_F64, BGW_DT_C128, BGW_DT_I32, BGW_DT_I64 = 1, 2, 3, 4

_ERRNAMES = {BGW_EINVAL: "BGW_EINVAL", BGW_ENOMEM: "BGW_ENOMEM", BGW_EHIP: "BGW_EHIP", BGW_ENODEV: "BGW_ENODEV",
             BGW_EIO: "BGW_EIO"}

_lib = None


class BgwError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


def _share_torch_hip_runtime() -> None:
    """PyTorch-ROCm ships its own libamdhip64 with the same soname as the system
    one.  Loading torch first makes libbgw bind to that already-loaded runtime, so
    the process has one HIP runtime and torch tensors/streams are valid in libbgw
    calls.  Without torch, libbgw uses the system runtime."""
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def lib() -> C.CDLL:
    """Loads build/libbgw.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} not built; run __graft_entry__.build()")
        _share_torch_hip_runtime()
        _lib = C.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


_P = C.c_void_p
_SZ = C.c_size_t
_D = C.c_double
_I = C.c_int


class _ArrayDesc(C.Structure):
    _fields_ = [("name", C.c_char_p), ("dtype", C.c_int), ("ndim", C.c_int), ("dims", C.c_size_t * 4),
                ("data", C.c_void_p)]


def _declare(L: C.CDLL) -> None:
    L.bgw_create.argtypes = [C.POINTER(_P), _I]
    L.bgw_destroy.argtypes = [_P]
    L.bgw_destroy.restype = None
    L.bgw_last_error.argtypes = [_P]
    L.bgw_last_error.restype = C.c_char_p
    L.bgw_set_stream.argtypes = [_P, _P]
    L.bgw_get_stream.argtypes = [_P]
    L.bgw_get_stream.restype = _P
    L.bgw_set_precision.argtypes = [_P, _I]
    L.bgw_synchronize.argtypes = [_P]
    L.bgw_set_timing.argtypes = [_P, _I]
    L.bgw_kernel_time.argtypes = [_P, C.c_char_p, C.POINTER(_D), C.POINTER(C.c_long)]
    L.bgw_reset_timing.argtypes = [_P]
    L.bgw_device_cus.argtypes = [_P]
    gpp12 = [_P, _SZ, _SZ, _SZ, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    L.bgw_noflagOCC_solver.argtypes = gpp12
    ext = [_P, _SZ, _SZ, _SZ, _SZ, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P, _P, _P, _P]
    L.bgw_gpp_solver.argtypes = ext
    L.bgw_gpp_solver_device.argtypes = ext
    L.bgw_gpp_solver_device_noalloc.argtypes = [_P, _SZ, _SZ, _SZ, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    L.bgw_gpp_backward.argtypes = ext[:-1]
    L.bgw_gpp_backward_device.argtypes = ext[:-1]
    L.bgw_compute_fact.argtypes = [_D, _I, _P, C.POINTER(_D), C.POINTER(_D), C.POINTER(_I), _I, _I]
    achs = [_P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P]
    L.bgw_achsDtemp_Kernel.argtypes = achs
    L.bgw_achsDtemp_Kernel_device.argtypes = achs
    asx = [_P, _I, _I, _I, _I, _I, _D, _D, _D, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    L.bgw_asxDtemp_Kernel.argtypes = asx
    L.bgw_asxDtemp_Kernel_device.argtypes = asx
    cor = [_P, _I, _I, _I, _I, _I, _I, _D, _D, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    L.bgw_achDtemp_cor_Kernel.argtypes = cor
    L.bgw_achDtemp_cor_Kernel_device.argtypes = cor
    achD = [_P, _I, _I, _I, _I, _I, _I, _D, _D, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    L.bgw_achDtemp_Kernel.argtypes = achD
    L.bgw_achDtemp_Kernel_device.argtypes = achD
    band = [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _D, _D]
    for k in ("1", "2"):
        getattr(L, f"bgw_schDttt_corKernel{k}").argtypes = band
        getattr(L, f"bgw_schDttt_corKernel{k}_device").argtypes = band
    L.bgw_synth_uniform.argtypes = [C.c_ulonglong, _D, _D, _P, _SZ]
    L.bgw_synth_uniform.restype = None
    L.bgw_synth_complex.argtypes = [C.c_ulonglong, _D, _D, _D, _D, _P, _SZ]
    L.bgw_synth_complex.restype = None
    L.bgw_synth_index.argtypes = [C.c_ulonglong, _I, _P, _SZ]
    L.bgw_synth_index.restype = None
    L.bgw_synth_permutation.argtypes = [C.c_ulonglong, _P, _SZ]
    L.bgw_synth_permutation.restype = None
    L.bgw_inputs_write.argtypes = [C.c_char_p, _I, C.POINTER(_ArrayDesc), _I]
    L.bgw_inputs_open.argtypes = [C.c_char_p, _I, C.POINTER(_P)]
    L.bgw_inputs_kind.argtypes = [_P]
    L.bgw_inputs_count.argtypes = [_P]
    L.bgw_inputs_entry.argtypes = [_P, _I, C.POINTER(_ArrayDesc)]
    L.bgw_inputs_get.argtypes = [_P, C.c_char_p, C.POINTER(_ArrayDesc)]
    L.bgw_inputs_close.argtypes = [_P]
    L.bgw_inputs_close.restype = None
    L.bgw_io_last_error.argtypes = []
    L.bgw_io_last_error.restype = C.c_char_p


def header_symbols() -> list[str]:
    """Names of every function declared in include/bgw.h."""
    import re

    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bgw_[a-zA-Z0-9_]+)\s*\(", text)))


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


class _Hold:
    """The array arguments of one foreign call, in the layout and dtype the C ABI reads
    (C-contiguous int32, float64 or complex128).  A converted copy must outlive the call,
    so the holder keeps every array it returns a pointer to; it lives as long as the
    calling method.

    The C ABI mirrors the reference's pointer-only signatures and cannot see array
    lengths, so each argument is given the number of elements the call will read (`n`)
    and a shorter array raises ValueError here, before any foreign call."""

    def __init__(self):
        self._keep = []

    def _ptr(self, a, n, name):
        if n is not None and a.size < n:
            raise ValueError(f"{name}: {a.size} elements, the call reads {n}")
        self._keep.append(a)
        return a.ctypes.data

    def i32(self, a, n=None, name="index array"):
        if a is None:
            return None
        a = np.asarray(a)
        if not np.issubdtype(a.dtype, np.integer):
            raise TypeError(f"{name}: expected an integer index array, got {a.dtype}")
        b = np.ascontiguousarray(a, np.int32)
        if b.dtype != a.dtype and not np.array_equal(b, a):
            raise ValueError(f"{name}: values outside the int32 range")
        return self._ptr(b, n, name)

    def f64(self, a, n=None, name="real array"):
        if a is None:
            return None
        a = np.asarray(a)
        if np.iscomplexobj(a):
            raise TypeError(f"{name}: expected a real array, got {a.dtype}")
        return self._ptr(np.ascontiguousarray(a, np.float64), n, name)

    def c128(self, a, n=None, name="complex array"):
        if a is None:
            return None
        a = np.asarray(a)
        if not np.iscomplexobj(a):
            raise TypeError(f"{name}: expected a complex array, got {a.dtype}")
        return self._ptr(np.ascontiguousarray(a, np.complex128), n, name)


# ----------------------------------------------------------------- synthetic inputs
def synth_uniform(seed: int, lo: float, hi: float, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    lib().bgw_synth_uniform(seed, lo, hi, _ptr(out), n)
    return out


def synth_complex(seed: int, re: tuple, im: tuple, n: int) -> np.ndarray:
    out = np.empty(n, np.complex128)
    lib().bgw_synth_complex(seed, re[0], re[1], im[0], im[1], _ptr(out), n)
    return out


def synth_index(seed: int, n_max: int, n: int) -> np.ndarray:
    out = np.empty(n, np.int32)
    lib().bgw_synth_index(seed, n_max, _ptr(out), n)
    return out


def synth_permutation(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.int32)
    lib().bgw_synth_permutation(seed, _ptr(out), n)
    return out


# The benchmark configurations (number_bands, nvband, ncouls, nodes_per_group).
GPP_CONFIGS = {
    0: dict(nb=32, nvband=2, ncouls=2048, nodes_per_group=1),
    1: dict(nb=512, nvband=2, ncouls=32768, nodes_per_group=20),
    2: dict(nb=512, nvband=2, ncouls=32768, nodes_per_group=20),
    3: dict(nb=256, nvband=2, ncouls=16384, nodes_per_group=20),
}
FF_CONFIG4 = dict(nb=512, nvband=2, ncouls=16384, nodes_per_group=20, nFreq=15, nfreqeval=10)


def seed_base(config: int) -> int:
    return 0xB6600000 + 16 * config


def gpp_inputs(nb: int, nvband: int, ncouls: int, ngpown: int, seed: int, kind: str = "random") -> dict:
    """GPP inputs in the reference layout.

    kind="random": the benchmark distributions (aqsm/aqsn re,im U[-1,1]; I_eps re
    U[0,.5], im U[-.05,.05]; wtilde re U[.5,3], im U[-.1,0]; vcoul U[.1,1];
    inv_igp_index U{0..ncouls-1}; indinv a permutation with indinv[ncouls] =
    ncouls-1; ek U[-1,1]; wx(n1,iw) = max(10 - ek[n1] + (iw-1), 1e-6)).
    kind="reference": the reference driver's constants (GPP/gpp.cpp:246-276):
    every complex input 0.5+0.5i, vcoul 1, inv_igp_index[p] = (p+1)*ncouls/ngpown,
    indinv[g] = g, wx = {3,4,5} for every band.
    """
    if kind == "reference":
        c = np.full(nb * ncouls, 0.5 + 0.5j)
        e = np.full(ngpown * ncouls, 0.5 + 0.5j)
        inv = ((np.arange(ngpown, dtype=np.int64) + 1) * ncouls // ngpown).astype(np.int32)
        indinv = np.concatenate([np.arange(ncouls, dtype=np.int32), np.array([ncouls - 1], np.int32)])
        wx = np.tile(np.array([3.0, 4.0, 5.0]), nb)
        return dict(nb=nb, nvband=nvband, ncouls=ncouls, ngpown=ngpown, aqsm=c, aqsn=c.copy(), I_eps=e,
                    wtilde=e.copy(), vcoul=np.ones(ncouls), inv=inv, indinv=indinv, wx=wx, ek=None)
    aqsm = synth_complex(seed + 1, (-1, 1), (-1, 1), nb * ncouls)
    aqsn = synth_complex(seed + 2, (-1, 1), (-1, 1), nb * ncouls)
    I_eps = synth_complex(seed + 3, (0, 0.5), (-0.05, 0.05), ngpown * ncouls)
    wtilde = synth_complex(seed + 4, (0.5, 3.0), (-0.1, 0.0), ngpown * ncouls)
    vcoul = synth_uniform(seed + 5, 0.1, 1.0, ncouls)
    inv = synth_index(seed + 6, ncouls, ngpown)
    indinv = np.concatenate([synth_permutation(seed + 7, ncouls), np.array([ncouls - 1], np.int32)])
    ek = synth_uniform(seed + 8, -1.0, 1.0, nb)
    e_lk, dw, to1 = 10.0, 1.0, 1e-6
    wx = e_lk - ek[:, None] + dw * (np.arange(3)[None, :] - 1.0)
    wx = np.where(wx < to1, to1, wx).reshape(-1).copy()
    return dict(nb=nb, nvband=nvband, ncouls=ncouls, ngpown=ngpown, aqsm=aqsm, aqsn=aqsn, I_eps=I_eps,
                wtilde=wtilde, vcoul=vcoul, inv=inv, indinv=indinv, wx=wx, ek=ek)


def gpp_config_inputs(config: int, nb: int | None = None) -> dict:
    c = GPP_CONFIGS[config]
    ngpown = c["ncouls"] // c["nodes_per_group"]
    return gpp_inputs(nb or c["nb"], c["nvband"], c["ncouls"], ngpown, seed_base(config))


def ff_pref(dFreqGrid) -> np.ndarray:
    """The reference driver's trapezoid weights of the frequency grid (FF/ff.cpp:619-626):
    pref[f] = (grid[f+1] - grid[f]) / 3.14, the last repeating the one before, both ends
    halved.  The input of bgw_achDtemp_Kernel."""
    g = np.asarray(dFreqGrid, np.float64)
    if g.size < 2:
        raise ValueError("the frequency grid needs at least two points")
    p = np.empty(g.size)
    p[:-1] = (g[1:] - g[:-1]) / 3.14
    p[-1] = p[-2]
    p[0] *= 0.5
    p[-1] *= 0.5
    return p


def ff_inputs(nb: int, nvband: int, ncouls: int, ngpown: int, nFreq: int, nfreqeval: int, seed: int,
              kind: str = "random") -> dict:
    """FF inputs.  kind="reference" reproduces FF/ff.cpp:584-633 (inv/indinv identity,
    ekq = -10 + n1, aqs 0.5+0.5i, vcoul 1, I_epsR 0.5+0.5i, I_epsA 0.5-0.5i,
    dFreqGrid = 2f, freqevalmin 0, freqevalstep 0.5, occ 1).  kind="random": aqs as
    the GPP inputs, I_eps{R,A} re U[0,.5], im U[-.1,.1], ekq U[-1,1], dFreqGrid
    uniform on [0,2], same freqevalmin/step/occ.  Both carry pref = ff_pref(dFreqGrid)."""
    npg = nFreq * ngpown * ncouls
    if kind == "reference":
        return dict(nb=nb, nvband=nvband, ncouls=ncouls, ngpown=ngpown, nFreq=nFreq, nfreqeval=nfreqeval,
                    inv=np.arange(ngpown, dtype=np.int32), indinv=np.arange(ncouls, dtype=np.int32),
                    ekq=-10.0 + np.arange(nb, dtype=np.float64), aqsm=np.full(nb * ncouls, 0.5 + 0.5j),
                    aqsn=np.full(nb * ncouls, 0.5 + 0.5j), vcoul=np.ones(ncouls),
                    I_epsR=np.full(npg, 0.5 + 0.5j), I_epsA=np.full(npg, 0.5 - 0.5j),
                    dFreqGrid=2.0 * np.arange(nFreq, dtype=np.float64), freqevalmin=0.0, freqevalstep=0.5,
                    occ=1.0, pref=ff_pref(2.0 * np.arange(nFreq, dtype=np.float64)))
    return dict(nb=nb, nvband=nvband, ncouls=ncouls, ngpown=ngpown, nFreq=nFreq, nfreqeval=nfreqeval,
                inv=synth_index(seed + 6, ncouls, ngpown), indinv=synth_permutation(seed + 7, ncouls),
                ekq=synth_uniform(seed + 8, -1.0, 1.0, nb),
                aqsm=synth_complex(seed + 1, (-1, 1), (-1, 1), nb * ncouls),
                aqsn=synth_complex(seed + 2, (-1, 1), (-1, 1), nb * ncouls),
                vcoul=synth_uniform(seed + 5, 0.1, 1.0, ncouls),
                I_epsR=synth_complex(seed + 9, (0, 0.5), (-0.1, 0.1), npg),
                I_epsA=synth_complex(seed + 10, (0, 0.5), (-0.1, 0.1), npg),
                dFreqGrid=2.0 * np.arange(nFreq, dtype=np.float64) / (nFreq - 1), freqevalmin=0.0,
                freqevalstep=0.5, occ=1.0, pref=ff_pref(2.0 * np.arange(nFreq, dtype=np.float64) / (nFreq - 1)))


def ff_config4_inputs(nb: int | None = None, ngpown: int | None = None, ncouls: int | None = None) -> dict:
    c = FF_CONFIG4
    nc = ncouls or c["ncouls"]
    return ff_inputs(nb or c["nb"], c["nvband"], nc, ngpown or c["ncouls"] // c["nodes_per_group"], c["nFreq"],
                     c["nfreqeval"], seed_base(4))


# ----------------------------------------------------------------------- input sets on disk
_DT = {BGW_DT_F64: np.float64, BGW_DT_C128: np.complex128, BGW_DT_I32: np.int32, BGW_DT_I64: np.int64}
_DT_OF = {np.dtype(v): k for k, v in _DT.items()}

# file entry name -> key of the dicts made by gpp_inputs / ff_inputs
_GPP_FILE = {"aqsmtemp": "aqsm", "aqsntemp": "aqsn", "I_eps_array": "I_eps", "wtilde_array": "wtilde",
             "vcoul": "vcoul", "inv_igp_index": "inv", "indinv": "indinv", "wx_array": "wx"}
_FF_FILE = {"aqsmtemp": "aqsm", "aqsntemp": "aqsn", "I_epsR_array": "I_epsR", "I_epsA_array": "I_epsA",
            "vcoul": "vcoul", "inv_igp_index": "inv", "indinv": "indinv", "ekq": "ekq", "dFreqGrid": "dFreqGrid"}


def save_inputs(path, d: dict, kind: str) -> None:
    """Writes a GPP or FF input dict (the form gpp_inputs / ff_inputs return) as an
    input set (bgw_inputs_write).  Complex arrays get their 2-D / 3-D shapes."""
    nb, nc, ng = int(d["nb"]), int(d["ncouls"]), int(d["ngpown"])
    arrays = {"number_bands": np.int64(nb), "nvband": np.int64(d["nvband"]), "ncouls": np.int64(nc),
              "ngpown": np.int64(ng), "occ": np.float64(d.get("occ", 1.0))}
    if kind == "gpp":
        code, names = BGW_INPUTS_GPP, _GPP_FILE
        shapes = {"aqsm": (nb, nc), "aqsn": (nb, nc), "I_eps": (ng, nc), "wtilde": (ng, nc)}
        wx = np.asarray(d["wx"], np.float64)
        shapes["wx"] = (3,) if wx.size == 3 else (nb, 3)
    elif kind == "ff":
        code, names = BGW_INPUTS_FF, _FF_FILE
        nf = int(d["nFreq"])
        arrays.update(nFreq=np.int64(nf), nfreqeval=np.int64(d["nfreqeval"]),
                      freqevalmin=np.float64(d["freqevalmin"]), freqevalstep=np.float64(d["freqevalstep"]))
        shapes = {"aqsm": (nb, nc), "aqsn": (nb, nc), "I_epsR": (nf, ng, nc), "I_epsA": (nf, ng, nc)}
    else:
        raise ValueError("kind must be 'gpp' or 'ff'")
    for fname, key in names.items():
        a = np.asarray(d[key])
        arrays[fname] = a.reshape(shapes[key]) if key in shapes else a
    write_input_set(path, code, arrays)


def write_input_set(path, kind: int, arrays: dict) -> None:
    """Low-level writer: `arrays` maps entry names to numpy arrays or numpy scalars
    (float64, complex128, int32 or int64; a numpy scalar becomes a rank-0 entry)."""
    keep, descs = [], []
    for name, a in arrays.items():
        rank0 = np.ndim(a) == 0
        a = np.ascontiguousarray(a)
        if a.dtype not in _DT_OF:
            raise TypeError(f"{name}: unsupported dtype {a.dtype}")
        if a.ndim > 4:
            raise ValueError(f"{name}: rank above 4")
        keep.append(a)
        dims = (C.c_size_t * 4)(*(list(a.shape) + [1] * (4 - a.ndim)))
        descs.append(_ArrayDesc(name.encode(), _DT_OF[a.dtype], 0 if rank0 else a.ndim, dims, a.ctypes.data))
    arr = (_ArrayDesc * len(descs))(*descs)
    rc = lib().bgw_inputs_write(str(path).encode(), kind, arr, len(descs))
    if rc != BGW_OK:
        raise BgwError(rc, lib().bgw_io_last_error().decode(errors="replace"))


class _Mapping:
    """Owns a bgw_inputs handle; unmapped when the last view of it is gone."""

    def __init__(self, h):
        self.h = h

    def __del__(self):
        if self.h and _lib is not None:
            _lib.bgw_inputs_close(self.h)
            self.h = None


class InputSet:
    """A memory-mapped input set (bgw_inputs_open).  `arrays` maps entry names to
    read-only numpy views of the mapping.  Each view keeps the mapping alive, so
    views stay valid after close(); the file is unmapped once the set and every
    view of it are gone."""

    def __init__(self, path, verify: bool = True):
        h = _P()
        rc = lib().bgw_inputs_open(str(path).encode(), int(verify), C.byref(h))
        if rc != BGW_OK:
            raise BgwError(rc, lib().bgw_io_last_error().decode(errors="replace"))
        self._m = _Mapping(h)
        self.kind = lib().bgw_inputs_kind(h)
        self.arrays = {}
        for i in range(lib().bgw_inputs_count(h)):
            e = _ArrayDesc()
            lib().bgw_inputs_entry(h, i, C.byref(e))
            shape = tuple(e.dims[:e.ndim])
            dt = np.dtype(_DT[e.dtype])
            n = int(np.prod(shape)) if shape else 1
            buf = (C.c_char * (n * dt.itemsize)).from_address(e.data)
            buf._mapping = self._m  # the numpy view's base chain holds the mapping
            v = np.frombuffer(buf, dtype=dt, count=n)
            v = v.reshape(shape) if shape else v[0]
            if isinstance(v, np.ndarray):
                v.flags.writeable = False
            self.arrays[e.name.decode()] = v

    def close(self) -> None:
        """Drops this set's references; the mapping goes with the last view."""
        self.arrays = {}
        self._m = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def to_dict(self) -> dict:
        """Copies the set into the dict form gpp_inputs / ff_inputs return."""
        a = self.arrays
        d = dict(nb=int(a["number_bands"]), nvband=int(a["nvband"]), ncouls=int(a["ncouls"]),
                 ngpown=int(a["ngpown"]), occ=float(a["occ"]))
        if self.kind == BGW_INPUTS_GPP:
            names = _GPP_FILE
        elif self.kind == BGW_INPUTS_FF:
            names = _FF_FILE
            d.update(nFreq=int(a["nFreq"]), nfreqeval=int(a["nfreqeval"]), freqevalmin=float(a["freqevalmin"]),
                     freqevalstep=float(a["freqevalstep"]))
        else:
            raise ValueError("to_dict needs a GPP or FF input set")
        for fname, key in names.items():
            d[key] = np.array(a[fname]).reshape(-1)
        return d


def load_inputs(path, verify: bool = True) -> dict:
    """Reads a GPP or FF input set into the dict form gpp_inputs / ff_inputs return."""
    with InputSet(path, verify) as s:
        return s.to_dict()


def _gpp_arrays(h: _Hold, d: dict):
    """The eight array arguments of the extended GPP entry points, length-checked."""
    nb, ng, nc = d["nb"], d["ngpown"], d["ncouls"]
    return (h.i32(d["inv"], ng, "inv_igp_index"), h.i32(d["indinv"], nc + 1, "indinv"),
            h.f64(d["wx"], 3 * nb, "wx_array"), h.c128(d["wtilde"], ng * nc, "wtilde_array"),
            h.c128(d["aqsm"], nb * nc, "aqsmtemp"), h.c128(d["aqsn"], nb * nc, "aqsntemp"),
            h.c128(d["I_eps"], ng * nc, "I_eps_array"), h.f64(d["vcoul"], nc, "vcoul"))


def _ff_arrays(h: _Hold, f: dict, band_rows: int, slices: int):
    """Length-checked FF arrays: band arrays of `band_rows` rows, `slices` frequency slices."""
    ng, nc = f["ngpown"], f["ncouls"]
    return dict(inv=h.i32(f["inv"], ng, "inv_igp_index"), indinv=h.i32(f["indinv"], nc, "indinv"),
                aqsm=h.c128(f["aqsm"], band_rows * nc, "aqsmtemp"),
                aqsn=h.c128(f["aqsn"], band_rows * nc, "aqsntemp"),
                vcoul=h.f64(f["vcoul"], nc, "vcoul"), R=h.c128(f["I_epsR"], slices * ng * nc, "I_epsR_array"),
                A=h.c128(f.get("I_epsA"), slices * ng * nc, "I_epsA_array"))


# ----------------------------------------------------------------------- context
class Context:
    """One GPU + one stream (bgw_ctx)."""

    def __init__(self, device: int = 0):
        self._h = _P()
        rc = lib().bgw_create(C.byref(self._h), device)
        if rc != BGW_OK:
            raise BgwError(rc, "bgw_create failed (no usable GPU?)")

    def close(self) -> None:
        if self._h:
            lib().bgw_destroy(self._h)
            self._h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def _check(self, rc: int) -> None:
        if rc != BGW_OK:
            raise BgwError(rc, lib().bgw_last_error(self._h).decode(errors="replace"))

    def set_stream(self, stream_ptr: int | None) -> None:
        self._check(lib().bgw_set_stream(self._h, stream_ptr))

    def set_precision(self, precision: int) -> None:
        self._check(lib().bgw_set_precision(self._h, precision))

    def synchronize(self) -> None:
        self._check(lib().bgw_synchronize(self._h))

    def set_timing(self, on: bool) -> None:
        self._check(lib().bgw_set_timing(self._h, int(on)))

    def kernel_time(self, name: str) -> tuple[float, int]:
        ms, n = _D(), C.c_long()
        self._check(lib().bgw_kernel_time(self._h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def reset_timing(self) -> None:
        self._check(lib().bgw_reset_timing(self._h))

    @property
    def cus(self) -> int:
        return lib().bgw_device_cus(self._h)

    # ------------------------------------------------------------ GPP (host)
    def noflagOCC_solver(self, nb, ngpown, ncouls, inv, indinv, wx3, wtilde, aqsm, aqsn, I_eps, vcoul):
        ach = np.zeros(3, np.complex128)
        h = _Hold()
        npg, nbg = ngpown * ncouls, nb * ncouls
        self._check(lib().bgw_noflagOCC_solver(
            self._h, nb, ngpown, ncouls, h.i32(inv, ngpown, "inv_igp_index"), h.i32(indinv, ncouls + 1, "indinv"),
            h.f64(wx3, 3, "wx_array"), h.c128(wtilde, npg, "wtilde_array"), h.c128(aqsm, nbg, "aqsmtemp"),
            h.c128(aqsn, nbg, "aqsntemp"), h.c128(I_eps, npg, "I_eps_array"), h.f64(vcoul, ncouls, "vcoul"),
            _ptr(ach)))
        return ach

    def gpp_solver(self, d: dict, occ: float = 1.0, outputs=("achtemp", "acht_n1_loc", "achstemp", "asxtemp")):
        nb = d["nb"]
        res = {"achtemp": np.zeros(3, np.complex128)}
        if "acht_n1_loc" in outputs:
            res["acht_n1_loc"] = np.zeros(nb * 3, np.complex128)
        if "achstemp" in outputs:
            res["achstemp"] = np.zeros(1, np.complex128)
        if "asxtemp" in outputs:
            res["asxtemp"] = np.zeros(3, np.complex128)
        h = _Hold()
        self._check(lib().bgw_gpp_solver(
            self._h, nb, d["nvband"], d["ngpown"], d["ncouls"], *_gpp_arrays(h, d), occ, _ptr(res["achtemp"]),
            _ptr(res.get("acht_n1_loc")), _ptr(res.get("achstemp")), _ptr(res.get("asxtemp"))))
        if "acht_n1_loc" in res:
            res["acht_n1_loc"] = res["acht_n1_loc"].reshape(nb, 3)
        return res

    def gpp_backward(self, d: dict, occ: float = 1.0):
        npg = d["ngpown"] * d["ncouls"]
        ds = np.zeros(3, np.complex128)
        dw = np.zeros(npg, np.complex128)
        dI = np.zeros(npg, np.complex128)
        h = _Hold()
        self._check(lib().bgw_gpp_backward(
            self._h, d["nb"], d["nvband"], d["ngpown"], d["ncouls"], *_gpp_arrays(h, d), occ, _ptr(ds), _ptr(dw),
            _ptr(dI)))
        return {"dsigma_domega": ds, "dL_dwtilde": dw, "dL_dI_eps": dI}

    # ------------------------------------------------------------- GPP (device)
    def gpp_solver_device(self, nb, nvband, ngpown, ncouls, inv, indinv, wx, wtilde, aqsm, aqsn, I_eps, vcoul,
                          occ, achtemp, acht_n1_loc=None, achstemp=None, asxtemp=None):
        self._check(lib().bgw_gpp_solver_device(self._h, nb, nvband, ngpown, ncouls, inv, indinv, wx, wtilde, aqsm,
                                                aqsn, I_eps, vcoul, occ, achtemp, acht_n1_loc, achstemp, asxtemp))

    def gpp_solver_device_noalloc(self, nb, ngpown, ncouls, inv, indinv, wx, wtilde, aqsm, aqsn, I_eps, vcoul, achtemp,
                                  acht_n1_loc=None):
        """bgw_gpp_solver_device_noalloc: achtemp (+ acht_n1_loc) on device arrays with no
        device allocation (workspace from the library's per-GPU scratch array)."""
        self._check(lib().bgw_gpp_solver_device_noalloc(self._h, nb, ngpown, ncouls, inv, indinv, wx, wtilde, aqsm,
                                                        aqsn, I_eps, vcoul, achtemp, acht_n1_loc))

    def gpp_backward_device(self, nb, nvband, ngpown, ncouls, inv, indinv, wx, wtilde, aqsm, aqsn, I_eps, vcoul,
                            occ, dsigma=None, dwt=None, dI=None):
        self._check(lib().bgw_gpp_backward_device(self._h, nb, nvband, ngpown, ncouls, inv, indinv, wx, wtilde,
                                                  aqsm, aqsn, I_eps, vcoul, occ, dsigma, dwt, dI))

    # -------------------------------------------------------------------- FF
    def achsDtemp_Kernel(self, f: dict):
        out = np.zeros(1, np.complex128)
        h = _Hold()
        a = _ff_arrays(h, dict(f, I_epsA=None), f["nb"], 2)  # slice 1 of I_epsR is read
        self._check(lib().bgw_achsDtemp_Kernel(
            self._h, f["nb"], f["ngpown"], f["ncouls"], a["inv"], a["indinv"], a["aqsn"], a["aqsm"], a["R"],
            a["vcoul"], _ptr(out)))
        return out[0]

    def asxDtemp_Kernel(self, f: dict):
        out = np.zeros(f["nfreqeval"], np.complex128)
        h = _Hold()
        nv, nF = f["nvband"], f["nFreq"]
        a = _ff_arrays(h, f, nv, nF)  # asxDtemp reads the valence rows n1 < nvband
        self._check(lib().bgw_asxDtemp_Kernel(
            self._h, nv, f["nfreqeval"], f["ncouls"], f["ngpown"], nF, f["freqevalmin"], f["freqevalstep"], f["occ"],
            h.f64(f["ekq"], nv, "ekq"), h.f64(f["dFreqGrid"], nF, "dFreqGrid"), a["inv"], a["indinv"], a["aqsm"],
            a["aqsn"], a["vcoul"], a["R"], a["A"], _ptr(out)))
        return out

    def achDtemp_cor_Kernel(self, f: dict):
        out = np.zeros(f["nfreqeval"], np.complex128)
        h = _Hold()
        nb, nF = f["nb"], f["nFreq"]
        a = _ff_arrays(h, f, nb, nF)
        self._check(lib().bgw_achDtemp_cor_Kernel(
            self._h, nb, f["nvband"], f["nfreqeval"], f["ncouls"], f["ngpown"], nF, f["freqevalmin"],
            f["freqevalstep"], h.f64(f["ekq"], nb, "ekq"), h.f64(f["dFreqGrid"], nF, "dFreqGrid"), a["inv"],
            a["indinv"], a["aqsm"], a["aqsn"], a["vcoul"], a["R"], a["A"], _ptr(out)))
        return out

    def achDtemp_Kernel(self, f: dict):
        """The real-axis frequency sum (bgw_achDtemp_Kernel); f["pref"] defaults to
        ff_pref(f["dFreqGrid"])."""
        out = np.zeros(f["nfreqeval"], np.complex128)
        h = _Hold()
        nb, nF = f["nb"], f["nFreq"]
        a = _ff_arrays(h, f, nb, nF)
        pref = f["pref"] if f.get("pref") is not None else ff_pref(f["dFreqGrid"])
        self._check(lib().bgw_achDtemp_Kernel(
            self._h, nb, f["nvband"], f["nfreqeval"], f["ncouls"], f["ngpown"], nF, f["freqevalmin"],
            f["freqevalstep"], h.f64(f["ekq"], nb, "ekq"), h.f64(f["dFreqGrid"], nF, "dFreqGrid"),
            h.f64(pref, nF, "pref"), a["inv"], a["indinv"], a["aqsm"], a["aqsn"], a["vcoul"], a["R"], a["A"],
            _ptr(out)))
        return out

    def schDttt_corKernel(self, kernel: int, f: dict, n1: int, ifreq: int, fact1: float, fact2: float) -> complex:
        """schDttt_corKernel1 (kernel=1) or schDttt_corKernel2 (kernel=2) on host arrays:
        band n1's term of achDtemp_cor on slices ifreq, ifreq + 1."""
        out = np.zeros(1, np.complex128)
        fn = getattr(lib(), f"bgw_schDttt_corKernel{int(kernel)}")
        h = _Hold()
        a = _ff_arrays(h, f, n1 + 1, ifreq + 2)  # the call reads row n1 and slices ifreq, ifreq + 1
        self._check(fn(self._h, _ptr(out), a["inv"], a["indinv"], a["R"], a["A"], a["aqsm"], a["aqsn"], a["vcoul"],
                       f["ncouls"], ifreq, f["ngpown"], n1, fact1, fact2))
        return out[0]

    def schDttt_corKernel_device(self, kernel: int, f: dict, ptrs: dict, n1: int, ifreq: int, fact1: float,
                                 fact2: float, out_ptr: int):
        fn = getattr(lib(), f"bgw_schDttt_corKernel{int(kernel)}_device")
        self._check(fn(self._h, out_ptr, ptrs["inv"], ptrs["indinv"], ptrs["I_epsR"], ptrs["I_epsA"], ptrs["aqsm"],
                       ptrs["aqsn"], ptrs["vcoul"], f["ncouls"], ifreq, f["ngpown"], n1, fact1, fact2))

    def ff_device(self, op: str, f: dict, ptrs: dict, out_ptr: int):
        L = lib()
        if op == "achsDtemp":
            rc = L.bgw_achsDtemp_Kernel_device(self._h, f["nb"], f["ngpown"], f["ncouls"], ptrs["inv"],
                                               ptrs["indinv"], ptrs["aqsn"], ptrs["aqsm"], ptrs["I_epsR"],
                                               ptrs["vcoul"], out_ptr)
        elif op == "asxDtemp":
            rc = L.bgw_asxDtemp_Kernel_device(self._h, f["nvband"], f["nfreqeval"], f["ncouls"], f["ngpown"],
                                              f["nFreq"], f["freqevalmin"], f["freqevalstep"], f["occ"],
                                              ptrs["ekq"], ptrs["dFreqGrid"], ptrs["inv"], ptrs["indinv"],
                                              ptrs["aqsm"], ptrs["aqsn"], ptrs["vcoul"], ptrs["I_epsR"],
                                              ptrs["I_epsA"], out_ptr)
        elif op == "achDtemp":
            rc = L.bgw_achDtemp_Kernel_device(self._h, f["nb"], f["nvband"], f["nfreqeval"], f["ncouls"],
                                              f["ngpown"], f["nFreq"], f["freqevalmin"], f["freqevalstep"],
                                              ptrs["ekq"], ptrs["dFreqGrid"], ptrs["pref"], ptrs["inv"],
                                              ptrs["indinv"], ptrs["aqsm"], ptrs["aqsn"], ptrs["vcoul"],
                                              ptrs["I_epsR"], ptrs["I_epsA"], out_ptr)
        else:
            rc = L.bgw_achDtemp_cor_Kernel_device(self._h, f["nb"], f["nvband"], f["nfreqeval"], f["ncouls"],
                                                  f["ngpown"], f["nFreq"], f["freqevalmin"], f["freqevalstep"],
                                                  ptrs["ekq"], ptrs["dFreqGrid"], ptrs["inv"], ptrs["indinv"],
                                                  ptrs["aqsm"], ptrs["aqsn"], ptrs["vcoul"], ptrs["I_epsR"],
                                                  ptrs["I_epsA"], out_ptr)
        self._check(rc)


def compute_fact(wx: float, dFreqGrid: np.ndarray, loop: int, flag_occ: bool, ifreq: int = 0,
                 fact1: float = 0.0, fact2: float = 0.0):
    g = np.ascontiguousarray(dFreqGrid, np.float64)
    f1, f2, i = _D(fact1), _D(fact2), _I(ifreq)
    rc = lib().bgw_compute_fact(wx, len(g), _ptr(g), C.byref(f1), C.byref(f2), C.byref(i), loop, int(flag_occ))
    if rc != BGW_OK:
        raise BgwError(rc, "bgw_compute_fact")
    return f1.value, f2.value, i.value


def gpp_flops(nb: int, ngpown: int, ncouls: int) -> float:
    """Algorithmic FP64 work of achtemp: 15 FLOP per (n1, p, g, iw) term."""
    return 15.0 * nb * ngpown * ncouls * 3


