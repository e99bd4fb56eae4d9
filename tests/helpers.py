"""Shared test helpers: fixture loading and the reference-constant closed forms."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def _c(v):
    a = np.asarray(v, np.float64)
    return np.ascontiguousarray(a[..., 0] + 1j * a[..., 1])


def load_exact_gpp():
    out = []
    for c in json.loads((GOLDEN / "exact_gpp.json").read_text()):
        d = dict(nb=c["nb"], nvband=c["nvband"], ngpown=c["ngpown"], ncouls=c["ncouls"], aqsm=_c(c["aqsm"]),
                 aqsn=_c(c["aqsn"]), I_eps=_c(c["I_eps"]), wtilde=_c(c["wtilde"]),
                 vcoul=np.asarray(c["vcoul"], np.float64), inv=np.asarray(c["inv"], np.int32),
                 indinv=np.asarray(c["indinv"], np.int32), wx=np.asarray(c["wx"], np.float64))
        e = c["expect"]
        exp = dict(achtemp=_c(e["achtemp"]), acht_n1_loc=_c(e["acht_n1_loc"]), achstemp=_c([e["achstemp"]]),
                   asxtemp=_c(e["asxtemp"]))
        out.append((d, exp))
    return out


def load_exact_ff():
    out = []
    for c in json.loads((GOLDEN / "exact_ff.json").read_text()):
        f = dict(nb=c["nb"], nvband=c["nvband"], ngpown=c["ngpown"], ncouls=c["ncouls"], nFreq=c["nFreq"],
                 nfreqeval=c["nfreqeval"], aqsm=_c(c["aqsm"]), aqsn=_c(c["aqsn"]), I_epsR=_c(c["I_epsR"]),
                 I_epsA=_c(c["I_epsA"]), vcoul=np.asarray(c["vcoul"], np.float64), inv=np.asarray(c["inv"], np.int32),
                 indinv=np.asarray(c["indinv"], np.int32), ekq=np.asarray(c["ekq"], np.float64),
                 dFreqGrid=np.asarray(c["dFreqGrid"], np.float64), freqevalmin=c["freqevalmin"],
                 freqevalstep=c["freqevalstep"], occ=c["occ"], pref=np.asarray(c["pref"], np.float64))
        e = c["expect"]
        out.append((f, {k: complex(*v) for k, v in e.items()}))
    return out


# --- closed forms for the reference drivers' constant inputs --------------------
def gpp_reference_closed_form(nb, ngpown, ncouls, wx=(3.0, 4.0, 5.0)):
    """Every input 0.5+0.5i, vcoul=1 (GPP/gpp.cpp:246-276): each term is
    s*I/(wx-wt) with s = conj(a)a*0.5*1*wt = 0.25(1+i), s*I = 0.25(1+i)(0.5+0.5i)
    = 0.125i, so achtemp[iw] = nb*ngpown*ncouls * 0.125i / (wx[iw] - 0.5 - 0.5i)."""
    wx = np.asarray(wx, np.float64)
    return nb * ngpown * ncouls * 0.125j / (wx - (0.5 + 0.5j))


def gpp_reference_achstemp(nb, ngpown, ncouls):
    """-1/2 sum vcoul conj(aqsm) sum_g aqsn I = -1/2 nb ngpown ncouls (0.5-0.5i)(0.5+0.5i)^2."""
    return -0.5 * nb * ngpown * ncouls * (0.5 - 0.5j) * (0.5 + 0.5j) ** 2


def ff_reference_closed_form(nb, nvband, ncouls, ngpown, nFreq):
    """FF/ff.cpp:584-633 constants.  wx(n1) = 10.5 - n1 (freqevalmin 0, step 0.5,
    ekq = -10 + n1); interpolation weights sum to 1 in loop 1 and loop-2-occupied
    and to -1/2 in loop 2 with wx>0, whatever interval compute_fact picks.
      achsDtemp = -nb ngpown ncouls (0.125 + 0.125i)
      asxDtemp  = ngpown ncouls 0.5 [#(n1<nvband, wx>0) R + #(n1<nvband, wx<=0) A],  R=.5+.5i, A=.5-.5i
      achDtemp_cor = ngpown ncouls [ #(wx>0, unocc) * 0.5 * (-1/2) i + #(wx<=0, occ) * 0.5 * (-1/2) i ]."""
    n = np.arange(nb)
    wx = 10.5 - n
    achs = -nb * ngpown * ncouls * (0.125 + 0.125j)
    occ = n < nvband
    asx = ngpown * ncouls * 0.5 * (np.sum(occ & (wx > 0)) * (0.5 + 0.5j) + np.sum(occ & ~(wx > 0)) * (0.5 - 0.5j))
    k1 = np.sum((wx > 0) & ~occ)
    k2 = np.sum(~(wx > 0) & occ)
    cor = ngpown * ncouls * (k1 * 0.5 * (-0.5) * 1j + k2 * 0.5 * (-0.5) * 1j)
    return achs, asx, cor


def ff_reference_achDtemp(nb, nvband, ncouls, ngpown, nFreq):
    """The real-axis frequency sum at the FF/ff.cpp:584-633 constants: aqsn conj(aqsm) = 0.5,
    vcoul = 1 and R - A = i make every (n1, p, g, f) term 0.5 (-i/2 pref(f)/den) i
    = pref(f) / (4 den), den = wx(n1) - s(n1) 2f with wx = 10.5 - n1 and s = -1 for
    n1 < nvband; pref from FF/ff.cpp:619-626 on the grid 2f (spacing 2, so 2/3.14 with
    halved ends).  Real-valued."""
    pref = np.full(nFreq, 2.0 / 3.14)
    pref[0] *= 0.5
    pref[-1] *= 0.5
    tot = 0.0
    for n1 in range(nb):
        wx = 10.5 - n1
        s = -1.0 if n1 < nvband else 1.0
        tot += np.sum(pref / (wx - s * 2.0 * np.arange(nFreq)))
    return ngpown * ncouls * tot / 4.0


def ff_with_kernel2_band(f):
    """The same FF inputs with occupied band 0 at wx = -0.4 (ekq 0.9), so that
    achDtemp_cor has a schDttt_corKernel2 term (FF/ff.cpp:347-350)."""
    f = dict(f)
    f["ekq"] = f["ekq"].copy()
    f["ekq"][0] = 0.9
    return f


def ff_band_terms(f):
    """(kernel, n1, ifreq, fact1, fact2) of every band achDtemp_cor_Kernel sums
    (FF/ff.cpp:337-352): Kernel1 for wx > 0 unoccupied, Kernel2 for wx <= 0 occupied,
    with compute_fact's loop-2 interval and weights from the oracle."""
    import oracle as O

    out = []
    for n1 in range(f["nb"]):
        wx = f["freqevalmin"] - f["ekq"][n1] + f["freqevalstep"]
        occ = n1 < f["nvband"]
        f1, f2, ifreq = O.compute_fact(wx, f["dFreqGrid"], 2, occ)
        if wx > 0 and not occ:
            out.append((1, n1, ifreq, f1, f2))
        elif not wx > 0 and occ:
            out.append((2, n1, ifreq, f1, f2))
    return out

