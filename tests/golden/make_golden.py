"""Write the golden fixtures under tests/golden/.

* ``rys_prim_eri.json``: primitive Cartesian ERIs from openLOWDIN's own
  Rys-quadrature routine ``coul_rep_`` (src/ints/RysQInts.cpp:123-276), built
  from the reference sources by ``oracle/Makefile`` into
  ``oracle/_ref/librysq.so``.  One random primitive quartet per class
  (la lb|lc ld), l <= 2 (81 classes), with labels chosen so every component
  is written (increasing, gapped labels, indexb != indexs; the routine skips
  entries otherwise, RysQInts.cpp:185-214).  Classes with more than 216
  components keep a fixed subsample of 64 components.
* ``boys_mpmath.json``: F_m(T) = gamma(m+1/2, T) / (2 T^(m+1/2)) at 50
  digits (mpmath), m = 0..16, on T values spanning every branch.

Run in this container (needs the reference tree and mpmath):

    make -C oracle ref && python tests/golden/make_golden.py
"""
import ctypes as C
import itertools
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def ncart(l):
    return (l + 1) * (l + 2) // 2


def rys_fixture():
    lib = C.CDLL(os.path.join(REPO, "oracle", "_ref", "librysq.so"))
    rng = np.random.default_rng(20261202)
    cases = []
    d = lambda x: C.byref(C.c_double(float(x)))
    i = lambda x: C.byref(C.c_int(int(x)))
    for ls in itertools.product(range(3), repeat=4):
        cen = rng.uniform(-1.5, 1.5, size=(4, 3))
        ex = np.exp(rng.uniform(np.log(0.15), np.log(30.0), size=4))
        n = int(np.prod([ncart(l) for l in ls]))
        out = np.zeros(n)
        la, lb, lc, ld = ls
        lib.coul_rep_(d(cen[0, 0]), d(cen[0, 1]), d(cen[0, 2]), i(la), d(ex[0]), i(0),
                      d(cen[1, 0]), d(cen[1, 1]), d(cen[1, 2]), i(lb), d(ex[1]), i(100),
                      d(cen[2, 0]), d(cen[2, 1]), d(cen[2, 2]), i(lc), d(ex[2]), i(200),
                      d(cen[3, 0]), d(cen[3, 1]), d(cen[3, 2]), i(ld), d(ex[3]), i(300),
                      i(1), i(2), out.ctypes.data_as(C.POINTER(C.c_double)))
        idx = np.arange(n) if n <= 216 else np.sort(rng.choice(n, 64, replace=False))
        cases.append({"l": list(ls), "centers": cen.tolist(), "exps": ex.tolist(),
                      "index": idx.tolist(), "value": [float("%.17g" % v) for v in out[idx]]})
    return {"source": "openLOWDIN src/ints/RysQInts.cpp coul_rep_ (Rys quadrature), "
                      "unnormalised primitive Cartesian Gaussians x^l exp(-a r^2); "
                      "component index ((ia*nb+ib)*nc+ic)*nd+id, Cartesian order xx,xy,xz,yy,yz,zz",
            "cases": cases}


def boys_fixture():
    import mpmath as mp
    mp.mp.dps = 50
    Ts = [0.0, 1e-14, 1e-9, 1e-5, 1e-3, 0.02, 0.1, 0.37, 0.5, 1.0, 1.7, 2.5, 3.3, 4.9, 6.0,
          7.77, 9.5, 11.0, 13.3, 15.0, 17.9, 20.0, 23.4, 26.1, 29.0, 31.3, 33.9, 34.99, 35.0,
          35.01, 38.5, 40.0, 45.0, 52.5, 60.0, 75.0, 100.0, 150.0, 300.0, 1000.0]
    vals = []
    for T in Ts:
        row = []
        for m in range(17):
            if T == 0.0:
                v = mp.mpf(1) / (2 * m + 1)
            else:
                v = mp.gammainc(m + mp.mpf(1) / 2, 0, T) / (2 * mp.mpf(T) ** (m + mp.mpf(1) / 2))
            row.append(float("%.17g" % float(v)))
        vals.append(row)
    return {"source": "mpmath 50-digit lower incomplete gamma", "T": Ts, "F": vals}


if __name__ == "__main__":
    with open(os.path.join(HERE, "rys_prim_eri.json"), "w") as fh:
        json.dump(rys_fixture(), fh, separators=(",", ":"))
    with open(os.path.join(HERE, "boys_mpmath.json"), "w") as fh:
        json.dump(boys_fixture(), fh, separators=(",", ":"))
    print("fixtures written")
