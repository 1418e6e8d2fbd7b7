"""Reference values of the Faddeeva function w(z) = exp(-z^2) erfc(-iz),
computed with 400 digits (mpmath), for the device w(z) test: 2400 points over
the regions of the first quadrant (origin disk, the trapezoidal-rule bands
above and next to the real axis, continued-fraction and far zones), points
within 1e-300 .. 1e-3 of either axis, and the lower half-plane where
2 exp(-z^2) - w(-z) does not cancel badly.  Written to wofz_mpmath.json as
strings (z and w, 17 significant digits).

    python tests/golden/make_wofz_mpmath.py
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 400  # the tiny components near the axes need ~330 digits
HERE = os.path.dirname(os.path.abspath(__file__))


def points():
    g = np.random.default_rng(20261203)
    out = []
    U = lambda a, b, n: g.uniform(a, b, n)  # noqa: E731
    LU = lambda a, b, n: np.exp(g.uniform(np.log(a), np.log(b), n))  # noqa: E731
    n = 200
    out += list(U(0, 1, n) * np.exp(1j * U(0, np.pi / 2, n)))           # |z| < 1
    out += list(U(0, 15, n) + 1j * U(0.5, 8, n))                        # nodes about 0
    out += list(U(0.7, 15, n) + 1j * U(0, 0.5, n))                      # nodes about x
    out += list(U(1, 15, n) + 1j * LU(1e-300, 0.5, n))                  # next to the real axis
    out += list(LU(1e-300, 1e-3, n) + 1j * U(0.2, 30, n))               # next to the imaginary axis
    out += list(U(0, 30, n) + 1j * U(8, 100, n))                        # continued fraction, y >= 8
    out += list(U(15, 300, n) + 1j * U(0, 8, n))                        # continued fraction, x >= 15
    out += list(U(15, 27, n) + 1j * LU(1e-300, 1e-3, n))                # ... with exp(-x^2) in Re w
    out += list(LU(1e2, 1e12, n) * np.exp(1j * U(0, np.pi / 2, n)))     # far out
    out += list(U(-6, 6, n) + 1j * U(-3, 0, n))                         # lower half-plane
    out += list(U(-20, 20, n) - 1j * LU(1e-12, 0.5, n))                 # just below the real axis
    out += list(U(-1.2, 1.2, n) - 1j * U(1, 6, n))                      # below, exp(-z^2) dominant
    return np.array(out)


def main():
    z = points()
    rows = []
    for zz in z:
        w = mp.exp(-mp.mpc(zz) ** 2) * mp.erfc(-1j * mp.mpc(zz))
        rows.append([repr(float(zz.real)), repr(float(zz.imag)), mp.nstr(w.real, 17), mp.nstr(w.imag, 17)])
    with open(os.path.join(HERE, "wofz_mpmath.json"), "w") as f:
        json.dump({"source": "mpmath %s, 400 digits" % mp.__version__, "points": rows}, f)


if __name__ == "__main__":
    main()
