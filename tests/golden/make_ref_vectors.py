"""Writes tests/golden/ref_vectors.json: outputs of the REFERENCE's own kernels
(compiled from its sources into oracle/_ref by oracle/Makefile) on seeded
synthetic inputs.  Inputs are not stored: they are regenerated from the seed by
the library's generator (bgw.gpp_inputs / bgw.ff_inputs).

Run in the container that has the reference tree:
    make -C oracle ref && python tests/golden/make_ref_vectors.py
"""
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "berkeleygw-kernels-cpp_amd" / "python"))
sys.path.insert(0, str(ROOT / "oracle"))

import numpy as np  # noqa: E402

import bgw  # noqa: E402
import oracle as O  # noqa: E402


def pair(a):
    a = np.asarray(a)
    return np.stack([a.real, a.imag], -1).tolist()


def main():
    gpp = []
    for seed, (nb, nvband, ncouls, ngpown) in ((1001, (7, 2, 300, 120)), (1002, (3, 1, 257, 64))):
        d = bgw.gpp_inputs(nb, nvband, ncouls, ngpown, seed)
        gpp.append(dict(seed=seed, nb=nb, nvband=nvband, ncouls=ncouls, ngpown=ngpown,
                        acht_n1_loc=pair(O.ref_gpp_per_band(d))))
    ff = []
    for seed, shape in ((2001, (24, 3, 80, 20, 15, 2)), (2002, (17, 4, 50, 11, 6, 3))):
        f = bgw.ff_inputs(*shape, seed)
        ff.append(dict(seed=seed, shape=list(shape), achsDtemp=pair([O.ref_achsDtemp(f)])[0],
                       asxDtemp=pair(O.ref_asxDtemp(f)), achDtemp_cor=pair(O.ref_achDtemp_cor(f))))
    out = Path(__file__).resolve().parent / "ref_vectors.json"
    out.write_text(json.dumps({"gpp": gpp, "ff": ff}, indent=1))
    print("wrote", out)


if __name__ == "__main__":
    main()
