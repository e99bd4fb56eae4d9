"""Python access to the CPU oracle (build/liboracle.so) and to the reference's own
kernels built as a checker (_ref/libbgwref_*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the measured or
shipped path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "build" / "liboracle.so"
REF_SERIAL = HERE / "_ref" / "libbgwref_serial.so"
REF_OMP = HERE / "_ref" / "libbgwref_omp.so"

_P, _SZ, _D, _I = C.c_void_p, C.c_size_t, C.c_double, C.c_int
_o = None
_refs: dict = {}


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def olib() -> C.CDLL:
    global _o
    if _o is None:
        if not ORACLE_LIB.exists():
            raise FileNotFoundError(f"{ORACLE_LIB} not built (make -C oracle)")
        L = C.CDLL(str(ORACLE_LIB))
        gpp = [_SZ, _SZ, _SZ, _SZ, _P, _P, _P, _P, _P, _P, _P, _P, _D]
        L.oracle_gpp_forward.argtypes = gpp + [_P, _P, _P, _P, _I]
        L.oracle_gpp_forward.restype = None
        L.oracle_gpp_grad_entries.argtypes = gpp + [_P, _SZ, _P, _P, _I]
        L.oracle_gpp_grad_entries.restype = None
        L.oracle_gpp_dsigma_domega.argtypes = gpp + [_P, _I]
        L.oracle_gpp_dsigma_domega.restype = None
        L.oracle_compute_fact.argtypes = [_D, _I, _P, C.POINTER(_D), C.POINTER(_D), C.POINTER(_I), _I, _I]
        L.oracle_compute_fact.restype = _I
        L.oracle_achsDtemp.argtypes = [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _I]
        L.oracle_achsDtemp.restype = None
        ff = [_I, _I, _I, _I, _I, _D, _D]
        L.oracle_asxDtemp.argtypes = ff + [_D, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I]
        L.oracle_asxDtemp.restype = None
        L.oracle_achDtemp_cor.argtypes = [_I] + ff + [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I]
        L.oracle_achDtemp_cor.restype = None
        L.oracle_achDtemp.argtypes = [_I] + ff + [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I]
        L.oracle_achDtemp.restype = None
        L.oracle_schDttt_cor.argtypes = [_I, _I, _I, _I, _I, _D, _D] + [_P] * 8 + [_I]
        L.oracle_schDttt_cor.restype = None
        _o = L
    return _o


def have_ref() -> bool:
    return REF_SERIAL.exists()


def rlib(kind: str = "serial") -> C.CDLL:
    if kind not in _refs:
        path = REF_SERIAL if kind == "serial" else REF_OMP
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        L = C.CDLL(str(path))
        L.ref_noflagOCC_solver.argtypes = [_SZ, _SZ, _SZ] + [_P] * 9
        L.ref_noflagOCC_solver.restype = None
        L.ref_gpp_per_band.argtypes = [_SZ, _SZ, _SZ] + [_P] * 9
        L.ref_gpp_per_band.restype = None
        L.ref_achsDtemp.argtypes = [_I, _I, _I, _I] + [_P] * 7
        L.ref_achsDtemp.restype = None
        ff = [_I, _I, _I, _I, _I, _I, _D, _D, _D] + [_P] * 10
        L.ref_asxDtemp.argtypes = ff
        L.ref_asxDtemp.restype = None
        L.ref_achDtemp_cor.argtypes = [_I, _I, _I, _I, _I, _I, _D, _D] + [_P] * 10
        L.ref_achDtemp_cor.restype = None
        L.ref_last_seconds.argtypes = []
        L.ref_last_seconds.restype = _D
        _refs[kind] = L
    return _refs[kind]


def nthreads_default() -> int:
    return int(os.environ.get("BGW_ORACLE_THREADS", "0"))


# ------------------------------------------------------------------------- GPP
def gpp_forward(d: dict, occ: float = 1.0, nthreads: int | None = None, want=("achtemp", "acht_n1_loc",
                                                                              "achstemp", "asxtemp")) -> dict:
    nb = d["nb"]
    res = {"achtemp": np.zeros(3, np.complex128)}
    if "acht_n1_loc" in want:
        res["acht_n1_loc"] = np.zeros(nb * 3, np.complex128)
    if "achstemp" in want:
        res["achstemp"] = np.zeros(1, np.complex128)
    if "asxtemp" in want:
        res["asxtemp"] = np.zeros(3, np.complex128)
    olib().oracle_gpp_forward(nb, d["nvband"], d["ngpown"], d["ncouls"], _p(d["inv"]), _p(d["indinv"]), _p(d["wx"]),
                              _p(d["wtilde"]), _p(d["aqsm"]), _p(d["aqsn"]), _p(d["I_eps"]), _p(d["vcoul"]), occ,
                              _p(res["achtemp"]), _p(res.get("acht_n1_loc")), _p(res.get("achstemp")),
                              _p(res.get("asxtemp")), nthreads_default() if nthreads is None else nthreads)
    if "acht_n1_loc" in res:
        res["acht_n1_loc"] = res["acht_n1_loc"].reshape(nb, 3)
    return res


def gpp_grad_entries(d: dict, entries: np.ndarray, occ: float = 1.0, nthreads: int | None = None):
    e = np.ascontiguousarray(entries, np.int64)
    dw = np.zeros(len(e), np.complex128)
    dI = np.zeros(len(e), np.complex128)
    olib().oracle_gpp_grad_entries(d["nb"], d["nvband"], d["ngpown"], d["ncouls"], _p(d["inv"]), _p(d["indinv"]),
                                   _p(d["wx"]), _p(d["wtilde"]), _p(d["aqsm"]), _p(d["aqsn"]), _p(d["I_eps"]),
                                   _p(d["vcoul"]), occ, _p(e), len(e), _p(dw), _p(dI),
                                   nthreads_default() if nthreads is None else nthreads)
    return dw, dI


def gpp_dsigma_domega(d: dict, occ: float = 1.0, nthreads: int | None = None):
    out = np.zeros(3, np.complex128)
    olib().oracle_gpp_dsigma_domega(d["nb"], d["nvband"], d["ngpown"], d["ncouls"], _p(d["inv"]), _p(d["indinv"]),
                                    _p(d["wx"]), _p(d["wtilde"]), _p(d["aqsm"]), _p(d["aqsn"]), _p(d["I_eps"]),
                                    _p(d["vcoul"]), occ, _p(out), nthreads_default() if nthreads is None else nthreads)
    return out


# -------------------------------------------------------------------------- FF
def compute_fact(wx, dFreqGrid, loop, flag_occ, ifreq=0, fact1=0.0, fact2=0.0):
    g = np.ascontiguousarray(dFreqGrid, np.float64)
    f1, f2, i = _D(fact1), _D(fact2), _I(ifreq)
    if olib().oracle_compute_fact(wx, len(g), _p(g), C.byref(f1), C.byref(f2), C.byref(i), loop, int(flag_occ)):
        return None  # stale ifreq outside the grid, no bracket: the reference would read out of bounds
    return f1.value, f2.value, i.value


def achsDtemp(f: dict, nthreads: int | None = None) -> complex:
    out = np.zeros(1, np.complex128)
    olib().oracle_achsDtemp(f["nb"], f["ngpown"], f["ncouls"], _p(f["inv"]), _p(f["indinv"]), _p(f["aqsn"]),
                            _p(f["aqsm"]), _p(f["I_epsR"]), _p(f["vcoul"]), _p(out),
                            nthreads_default() if nthreads is None else nthreads)
    return out[0]


def asxDtemp(f: dict, nthreads: int | None = None) -> np.ndarray:
    out = np.zeros(f["nfreqeval"], np.complex128)
    olib().oracle_asxDtemp(f["nvband"], f["nfreqeval"], f["ncouls"], f["ngpown"], f["nFreq"], f["freqevalmin"],
                           f["freqevalstep"], f["occ"], _p(f["ekq"]), _p(f["dFreqGrid"]), _p(f["inv"]),
                           _p(f["indinv"]), _p(f["aqsm"]), _p(f["aqsn"]), _p(f["vcoul"]), _p(f["I_epsR"]),
                           _p(f["I_epsA"]), _p(out), nthreads_default() if nthreads is None else nthreads)
    return out


def achDtemp_cor(f: dict, nthreads: int | None = None) -> np.ndarray:
    out = np.zeros(f["nfreqeval"], np.complex128)
    olib().oracle_achDtemp_cor(f["nb"], f["nvband"], f["nfreqeval"], f["ncouls"], f["ngpown"], f["nFreq"],
                               f["freqevalmin"], f["freqevalstep"], _p(f["ekq"]), _p(f["dFreqGrid"]), _p(f["inv"]),
                               _p(f["indinv"]), _p(f["aqsm"]), _p(f["aqsn"]), _p(f["vcoul"]), _p(f["I_epsR"]),
                               _p(f["I_epsA"]), _p(out), nthreads_default() if nthreads is None else nthreads)
    return out


def pref_weights(dFreqGrid) -> np.ndarray:
    """The reference driver's trapezoid weights, FF/ff.cpp:619-626: (grid[f+1]-grid[f])/3.14,
    the last repeating the one before, both ends halved."""
    g = np.asarray(dFreqGrid, np.float64)
    n = g.size
    p = np.empty(n)
    for f in range(n):
        p[f] = (g[f + 1] - g[f]) / 3.14 if f < n - 1 else p[f - 1]
    p[0] *= 0.5
    p[n - 1] *= 0.5
    return p


def achDtemp(f: dict, nthreads: int | None = None) -> np.ndarray:
    """The real-axis frequency sum with the weights pref (oracle_achDtemp)."""
    out = np.zeros(f["nfreqeval"], np.complex128)
    pref = np.ascontiguousarray(f["pref"] if "pref" in f else pref_weights(f["dFreqGrid"]), np.float64)
    olib().oracle_achDtemp(f["nb"], f["nvband"], f["nfreqeval"], f["ncouls"], f["ngpown"], f["nFreq"],
                           f["freqevalmin"], f["freqevalstep"], _p(f["ekq"]), _p(f["dFreqGrid"]), _p(pref),
                           _p(f["inv"]), _p(f["indinv"]), _p(f["aqsm"]), _p(f["aqsn"]), _p(f["vcoul"]),
                           _p(f["I_epsR"]), _p(f["I_epsA"]), _p(out),
                           nthreads_default() if nthreads is None else nthreads)
    return out


def schDttt_cor(f: dict, kernel: int, n1: int, ifreq: int, fact1: float, fact2: float,
                nthreads: int | None = None) -> complex:
    """schDttt_corKernel1 (kernel=1) / schDttt_corKernel2 (kernel=2), FF/ff.cpp:381-451."""
    out = np.zeros(1, np.complex128)
    olib().oracle_schDttt_cor(int(kernel), f["ncouls"], f["ngpown"], ifreq, n1, fact1, fact2, _p(f["inv"]),
                              _p(f["indinv"]), _p(f["I_epsR"]), _p(f["I_epsA"]), _p(f["aqsm"]), _p(f["aqsn"]),
                              _p(f["vcoul"]), _p(out), nthreads_default() if nthreads is None else nthreads)
    return out[0]


# -------------------------------------------------------- the reference itself
def ref_noflagOCC_solver(d: dict, wx3: np.ndarray, kind: str = "serial") -> np.ndarray:
    out = np.zeros(3, np.complex128)
    w = np.ascontiguousarray(wx3, np.float64)
    rlib(kind).ref_noflagOCC_solver(d["nb"], d["ngpown"], d["ncouls"], _p(d["inv"]), _p(d["indinv"]), _p(w),
                                    _p(d["wtilde"]), _p(d["aqsm"]), _p(d["aqsn"]), _p(d["I_eps"]), _p(d["vcoul"]),
                                    _p(out))
    return out


def ref_gpp_per_band(d: dict, kind: str = "serial") -> np.ndarray:
    out = np.zeros(d["nb"] * 3, np.complex128)
    rlib(kind).ref_gpp_per_band(d["nb"], d["ngpown"], d["ncouls"], _p(d["inv"]), _p(d["indinv"]), _p(d["wx"]),
                                _p(d["wtilde"]), _p(d["aqsm"]), _p(d["aqsn"]), _p(d["I_eps"]), _p(d["vcoul"]),
                                _p(out))
    return out.reshape(d["nb"], 3)


def ref_achsDtemp(f: dict, kind: str = "serial") -> complex:
    out = np.zeros(1, np.complex128)
    rlib(kind).ref_achsDtemp(f["nb"], f["ngpown"], f["ncouls"], f["nFreq"], _p(f["inv"]), _p(f["indinv"]), _p(f["aqsn"]),
                         _p(f["aqsm"]), _p(f["I_epsR"]), _p(f["vcoul"]), _p(out))
    return out[0]


def ref_asxDtemp(f: dict, kind: str = "serial") -> np.ndarray:
    out = np.zeros(f["nfreqeval"], np.complex128)
    rlib(kind).ref_asxDtemp(f["nb"], f["nvband"], f["nfreqeval"], f["ncouls"], f["ngpown"], f["nFreq"], f["freqevalmin"],
                        f["freqevalstep"], f["occ"], _p(f["ekq"]), _p(f["dFreqGrid"]), _p(f["inv"]), _p(f["indinv"]),
                        _p(f["aqsm"]), _p(f["aqsn"]), _p(f["vcoul"]), _p(f["I_epsR"]), _p(f["I_epsA"]), _p(out))
    return out


def ref_achDtemp_cor(f: dict, kind: str = "serial") -> np.ndarray:
    out = np.zeros(f["nfreqeval"], np.complex128)
    rlib(kind).ref_achDtemp_cor(f["nb"], f["nvband"], f["nfreqeval"], f["ncouls"], f["ngpown"], f["nFreq"],
                            f["freqevalmin"], f["freqevalstep"], _p(f["ekq"]), _p(f["dFreqGrid"]), _p(f["inv"]),
                            _p(f["indinv"]), _p(f["aqsm"]), _p(f["aqsn"]), _p(f["vcoul"]), _p(f["I_epsR"]),
                            _p(f["I_epsA"]), _p(out))
    return out


def ref_last_seconds(kind: str = "serial") -> float:
    """Wall time of the last reference kernel call of that build, timed inside the
    adapter around the kernel alone (not the copies into the reference's containers)."""
    return float(rlib(kind).ref_last_seconds())
