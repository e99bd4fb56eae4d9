"""Writes tests/golden/exact_*.json: tiny GPP/FF cases evaluated in exact rational
arithmetic (fractions.Fraction) straight from the formulas, independent of the
reference and of the oracle code.

Inputs are dyadic rationals (k/64), so they are exact doubles; expected outputs
are the exact sums rounded once to double.  Formulas (reference lines):
  achtemp[iw]  = sum_{n1,p,g} s(n1,p) I(p,g) conj(wd)/|wd|^2, wd = wx(n1,iw) - wt(p,g),
                 s = conj(aqsm(n1,igp)) aqsn(n1,igp) 0.5 vcoul(igp) wt(p,igp)   GPP/gpp.cpp:362-390
  acht_n1_loc  = per-band split of achtemp
  achstemp     = -1/2 sum_{n1,p} vcoul(igp) conj(aqsm(n1,igp)) sum_g aqsn(n1,g) I(p,g) [|I(p,g)| > to1]
  achsDtemp    = sum_{n1,p} 0.5 vcoul(igp) (-sum_g aqsn(n1,g) conj(aqsm(n1,igp)) R(1,p,g))  FF/ff.cpp:191-208
  achDtemp_cor = FF/ff.cpp:332-451 with fact1/fact2 from compute_fact (exact rationals)
  achDtemp     = sum_{n1,p,g} aqsn(n1,g) conj(aqsm(n1,igp)) vcoul(igp) sum_f -i/2 pref(f) (R-A)(f,p,g)
                 / (wx(n1) - s(n1) dFreqGrid(f)), s = -1 for n1 < nvband, +1 otherwise; pref the
                 reference driver's weights FF/ff.cpp:619-626, evaluated in double as it does

Run:  python tests/golden/make_exact_fixtures.py
"""
from __future__ import annotations

import json
import random
from fractions import Fraction as F
from pathlib import Path

OUT = Path(__file__).resolve().parent


class Cx:
    __slots__ = ("re", "im")

    def __init__(self, re, im=F(0)):
        self.re, self.im = F(re), F(im)

    def __add__(self, o):
        o = o if isinstance(o, Cx) else Cx(o)
        return Cx(self.re + o.re, self.im + o.im)

    __radd__ = __add__

    def __sub__(self, o):
        o = o if isinstance(o, Cx) else Cx(o)
        return Cx(self.re - o.re, self.im - o.im)

    def __rsub__(self, o):
        return Cx(o) - self

    def __mul__(self, o):
        o = o if isinstance(o, Cx) else Cx(o)
        return Cx(self.re * o.re - self.im * o.im, self.re * o.im + self.im * o.re)

    __rmul__ = __mul__

    def conj(self):
        return Cx(self.re, -self.im)

    def norm2(self):
        return self.re * self.re + self.im * self.im

    def out(self):
        return [float(self.re), float(self.im)]


def dy(rng, lo, hi):
    """dyadic k/64 in [lo, hi]"""
    return F(rng.randint(int(lo * 64), int(hi * 64)), 64)


TO1 = F(1, 10**6)
LIMITTWO = F(1, 4)
LIMITONE = 1 / (4 * TO1)
SEXCUT = F(4)


def ssx_exact(wx, wt, I):
    """BerkeleyGW occupied-band ssx in exact arithmetic (three branches + cut-off);
    |z| comparisons are done on squares."""
    wt2 = wt * wt
    Omega2 = wt2 * I
    wdiff = Cx(wx) - wt
    wdiffr = wdiff.norm2()
    delw = wt * wdiff.conj() * (F(1) / wdiffr)
    delwr = delw.norm2()
    if wdiffr > LIMITTWO and delwr < LIMITONE:
        cden = Cx(wx * wx) - wt2
        ssx = Omega2 * cden.conj() * (F(1) / cden.norm2())
    elif delwr > TO1:
        cden = F(4) * wt2 * (delw + F(1, 2))
        ssx = (Cx(0) - Omega2) * cden.conj() * (F(1) / cden.norm2()) * delw
    else:
        ssx = Cx(0)
    if ssx.norm2() > SEXCUT * SEXCUT * I.norm2() and wx < 0:
        ssx = Cx(0)
    return ssx


# I_eps values of the cut case: zero, below to1 and above it, each at least 5% from
# the boundary |I| = to1 (so the double comparison cannot differ from the exact one)
TINY_I = [Cx(0), Cx(F(1, 2**21)), Cx(0, F(1, 2**21)), Cx(F(1, 2**21), F(1, 2**21)), Cx(F(15, 2**24)),
          Cx(F(9, 2**23)), Cx(0, F(-9, 2**23)), Cx(F(1, 2**19), F(1, 2**20)), Cx(F(3, 2**20), F(-1, 2**20)),
          Cx(F(-7, 2**23), F(1, 2**22)), Cx(F(-5, 2**22), F(3, 2**22))]


def cut_keeps(I):
    """The static-COH cut |I_eps| > to1, exactly (on squares)."""
    return I.norm2() > TO1 * TO1


def gpp_case(seed, nb, nvband, ngpown, ncouls, near_pole=False, tiny_I=False):
    rng = random.Random(seed)
    aqsm = [Cx(dy(rng, -1, 1), dy(rng, -1, 1)) for _ in range(nb * ncouls)]
    aqsn = [Cx(dy(rng, -1, 1), dy(rng, -1, 1)) for _ in range(nb * ncouls)]
    I = [Cx(dy(rng, 0, 0.5), dy(rng, -0.05, 0.05)) for _ in range(ngpown * ncouls)]
    # |I| > to1 for every entry (re >= 1/64), so the static-COH cut never applies ...
    I = [Cx(max(x.re, F(1, 64)), x.im) for x in I]
    if tiny_I:
        # ... except here: every entry is zero or within a factor 4 of to1, on both sides
        I = [TINY_I[k % len(TINY_I)] for k in range(ngpown * ncouls)]
        for x in I:
            r = x.norm2() / (TO1 * TO1)
            assert x.norm2() == 0 or not F(9, 10) < r < F(11, 10)
    wt = [Cx(dy(rng, 0.5, 3), dy(rng, -0.1, 0)) for _ in range(ngpown * ncouls)]
    vcoul = [dy(rng, 0.1, 1) for _ in range(ncouls)]
    inv = [rng.randrange(ncouls + 1) for _ in range(ngpown)]  # may reach ncouls
    perm = list(range(ncouls))
    rng.shuffle(perm)
    indinv = perm + [ncouls - 1]
    wx = [dy(rng, 8, 12) for _ in range(nb * 3)]
    if near_pole:
        # band 0 at negative frequencies (the cut-off applies there) and a few
        # wtilde within 1/2 of a band frequency (the near-pole branch)
        wx[0], wx[1], wx[2] = F(-1, 2), F(-3, 16), F(-5, 4)
        wt[0] = Cx(wx[3] + F(1, 8), F(-1, 16))
        wt[ncouls + 1] = Cx(wx[4] - F(3, 16), F(-1, 32))
        wt[2] = Cx(F(-5, 16), F(-1, 8))
        wt[3] = Cx(F(1, 8192), F(0))  # |delw|^2 <= to1 next to wx = -3/16: the zero branch
    ach = [[Cx(0) for _ in range(3)] for _ in range(nb)]
    achs = Cx(0)
    for p in range(ngpown):
        igp = indinv[inv[p]]
        for n1 in range(nb):
            s = aqsm[n1 * ncouls + igp].conj() * aqsn[n1 * ncouls + igp] * F(1, 2) * vcoul[igp] * wt[p * ncouls + igp]
            for iw in range(3):
                acc = Cx(0)
                for g in range(ncouls):
                    wd = wx[n1 * 3 + iw] - wt[p * ncouls + g]
                    delw = wd.conj() * (F(1) / wd.norm2())
                    acc = acc + delw * I[p * ncouls + g]
                ach[n1][iw] = ach[n1][iw] + s * acc
            m = Cx(0)
            for g in range(ncouls):
                if cut_keeps(I[p * ncouls + g]):
                    m = m + aqsn[n1 * ncouls + g] * I[p * ncouls + g]
            achs = achs + F(-1, 2) * vcoul[igp] * aqsm[n1 * ncouls + igp].conj() * m
    achtemp = [sum((ach[n1][iw] for n1 in range(nb)), Cx(0)) for iw in range(3)]
    asx = [Cx(0) for _ in range(3)]
    for n1 in range(nvband):
        for p in range(ngpown):
            igp = indinv[inv[p]]
            m = aqsm[n1 * ncouls + igp].conj() * vcoul[igp]
            for iw in range(3):
                t = Cx(0)
                for g in range(ncouls):
                    t = t + aqsn[n1 * ncouls + g] * ssx_exact(wx[n1 * 3 + iw], wt[p * ncouls + g], I[p * ncouls + g])
                asx[iw] = asx[iw] + m * t
    return {
        "nb": nb, "nvband": nvband, "ngpown": ngpown, "ncouls": ncouls,
        "aqsm": [c.out() for c in aqsm], "aqsn": [c.out() for c in aqsn], "I_eps": [c.out() for c in I],
        "wtilde": [c.out() for c in wt], "vcoul": [float(v) for v in vcoul], "inv": inv, "indinv": indinv,
        "wx": [float(v) for v in wx],
        "expect": {"achtemp": [c.out() for c in achtemp], "acht_n1_loc": [[c.out() for c in r] for r in ach],
                   "achstemp": achs.out(), "asxtemp": [c.out() for c in asx]},
    }


def compute_fact_exact(wx, grid, loop, occ):
    """FF/ff.cpp:103-152 in exact arithmetic (same branch structure)."""
    f1 = f2 = F(0)
    ifreq = 0
    n = len(grid)
    if loop == 1:
        w = wx if wx > 0 else -wx
        for k in range(n - 1):
            if grid[k] < w < grid[k + 1]:
                ifreq = k
        if ifreq == 0:
            ifreq = n - 2
        d = grid[ifreq + 1] - grid[ifreq]
        if wx > 0:
            f1, f2 = (grid[ifreq + 1] - wx) / d, (wx - grid[ifreq]) / d
        else:
            f1, f2 = (grid[ifreq + 1] + wx) / d, (-grid[ifreq] - wx) / d
    elif wx > 0:
        for k in range(n - 1):
            if grid[k] < wx < grid[k + 1]:
                ifreq = k
        d = grid[ifreq + 1] - grid[ifreq]
        f1, f2 = F(-1, 2) * (grid[ifreq + 1] - wx) / d, F(-1, 2) * (wx - grid[ifreq]) / d
    elif occ:
        w = -wx
        for k in range(n - 1):
            if grid[k] < w < grid[k + 1]:
                ifreq = k
        if ifreq == 0:
            ifreq = n - 2
        d = grid[ifreq + 1] - grid[ifreq]
        f1, f2 = (grid[ifreq + 1] - w) / d, (w - grid[ifreq]) / d
    return f1, f2, ifreq


def ff_case(seed, nb, nvband, ngpown, ncouls, nFreq, nfreqeval):
    rng = random.Random(seed)
    aqsm = [Cx(dy(rng, -1, 1), dy(rng, -1, 1)) for _ in range(nb * ncouls)]
    aqsn = [Cx(dy(rng, -1, 1), dy(rng, -1, 1)) for _ in range(nb * ncouls)]
    R = [Cx(dy(rng, 0, 0.5), dy(rng, -0.1, 0.1)) for _ in range(nFreq * ngpown * ncouls)]
    A = [Cx(dy(rng, 0, 0.5), dy(rng, -0.1, 0.1)) for _ in range(nFreq * ngpown * ncouls)]
    vcoul = [dy(rng, 0.1, 1) for _ in range(ncouls)]
    inv = [rng.randrange(ncouls) for _ in range(ngpown)]
    indinv = list(range(ncouls))
    rng.shuffle(indinv)
    ekq = [dy(rng, -1, 1) for _ in range(nb)]
    grid = [F(2 * f, nFreq - 1) for f in range(nFreq)]
    fmin, fstep, occ = F(0), F(1, 2), F(1)

    def sl(X, f, p, g):
        return X[(f * ngpown + p) * ncouls + g]

    achs = Cx(0)
    for n1 in range(nb):
        for p in range(ngpown):
            igp = indinv[inv[p]]
            t = Cx(0)
            for g in range(ncouls):
                t = t - aqsn[n1 * ncouls + g] * aqsm[n1 * ncouls + igp].conj() * sl(R, 1, p, g)
            achs = achs + t * vcoul[igp] * F(1, 2)
    asx = Cx(0)
    for n1 in range(nvband):
        wx = fmin - ekq[n1] + fstep
        f1, f2, fi = compute_fact_exact(wx, grid, 1, False)
        X = R if wx > 0 else A
        for p in range(ngpown):
            igp = indinv[inv[p]]
            for g in range(ncouls):
                asx = asx + aqsn[n1 * ncouls + g] * aqsm[n1 * ncouls + igp].conj() * (
                    sl(X, fi, p, g) * f1 + sl(X, fi + 1, p, g) * f2) * vcoul[igp] * occ
    cor = Cx(0)
    for n1 in range(nb):
        occd = n1 < nvband
        wx = fmin - ekq[n1] + fstep
        f1, f2, fi = compute_fact_exact(wx, grid, 2, occd)
        if wx > 0 and not occd:
            sc = F(1)
        elif not wx > 0 and occd:
            sc = F(-1, 2)
        else:
            continue
        for p in range(ngpown):
            igp = indinv[inv[p]]
            for g in range(ncouls):
                e = (sl(R, fi, p, g) - sl(A, fi, p, g)) * f1 + (sl(R, fi + 1, p, g) - sl(A, fi + 1, p, g)) * f2
                cor = cor + aqsn[n1 * ncouls + g] * aqsm[n1 * ncouls + igp].conj() * e * vcoul[igp] * sc
    # the real-axis frequency sum: grid and weights as the double values the kernels see
    grid_d = [F(float(x)) for x in grid]
    pref_d = []
    for f in range(nFreq):
        pref_d.append((float(grid_d[f + 1]) - float(grid_d[f])) / 3.14 if f < nFreq - 1 else pref_d[f - 1])
    pref_d[0] *= 0.5
    pref_d[nFreq - 1] *= 0.5
    achD = Cx(0)
    for n1 in range(nb):
        wx = fmin - ekq[n1] + fstep
        for f in range(nFreq):
            den = wx + grid_d[f] if n1 < nvband else wx - grid_d[f]
            if den == 0:
                continue
            w = Cx(0, F(-1, 2) * F(pref_d[f]) / den)
            for p in range(ngpown):
                igp = indinv[inv[p]]
                m = Cx(0)
                for g in range(ncouls):
                    m = m + aqsn[n1 * ncouls + g] * (sl(R, f, p, g) - sl(A, f, p, g))
                achD = achD + m * aqsm[n1 * ncouls + igp].conj() * vcoul[igp] * w
    return {
        "nb": nb, "nvband": nvband, "ngpown": ngpown, "ncouls": ncouls, "nFreq": nFreq, "nfreqeval": nfreqeval,
        "aqsm": [c.out() for c in aqsm], "aqsn": [c.out() for c in aqsn], "I_epsR": [c.out() for c in R],
        "I_epsA": [c.out() for c in A], "vcoul": [float(v) for v in vcoul], "inv": inv, "indinv": indinv,
        "ekq": [float(v) for v in ekq], "dFreqGrid": [float(v) for v in grid], "freqevalmin": 0.0,
        "freqevalstep": 0.5, "occ": 1.0, "pref": pref_d,
        "expect": {"achsDtemp": achs.out(), "asxDtemp": asx.out(), "achDtemp_cor": cor.out(), "achDtemp": achD.out()},
    }


def main():
    cases = [gpp_case(11, 3, 1, 4, 8), gpp_case(12, 5, 2, 3, 7), gpp_case(13, 4, 2, 3, 6, near_pole=True),
             gpp_case(14, 4, 2, 3, 11, tiny_I=True)]
    (OUT / "exact_gpp.json").write_text(json.dumps(cases))
    ff = [ff_case(21, 6, 2, 3, 8, 5, 2), ff_case(22, 9, 3, 2, 6, 6, 3)]
    (OUT / "exact_ff.json").write_text(json.dumps(ff))
    print("wrote", OUT / "exact_gpp.json", OUT / "exact_ff.json")


if __name__ == "__main__":
    main()
