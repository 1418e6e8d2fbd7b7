"""The device's w(z) (behind every Bassetti-Erskine kick) against the
reference's known answers, 400-digit values and the oracle's restatement of
the reference's package, and the field the beam-ion kicks apply against the
oracle's Bassetti-Erskine field.

faddeeva_w is the Maclaurin series / trapezoidal rule with pole correction /
continued fraction of cts_faddeeva.hpp; the kicks take Weideman's rational
approximation near the beam (|w| relative error 3e-14).  Tolerances: each
component 1e-13 relative (the reference's own KAT criterion), 2e-13 below the
real axis where 2 exp(-z^2) - w(-z) cancels; |w| 1e-13 against the oracle.
"""
import json
import math
import os

import numpy as np
import pytest

from cetasim_amd import capi, configs
from cetasim_amd.tracker import Tracker

import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def dev():
    ring, secs, _ = configs.lattice()
    tr = Tracker(ring, secs)
    tr.set_bunches([64], [1.0], [1.0])
    yield tr
    tr.close()


def device_w(tr, z):
    z = np.asarray(z, dtype=complex)
    re, im = np.ascontiguousarray(z.real), np.ascontiguousarray(z.imag)
    wr, wi = np.empty(len(z)), np.empty(len(z))
    capi.check(tr.lib.cts_debug_faddeeva(tr.h, len(z), capi.dptr(re), capi.dptr(im), capi.dptr(wr), capi.dptr(wi)))
    return wr + 1j * wi


def oracle_w(z):
    L = O.lib()
    out = np.empty(len(z), dtype=complex)
    import ctypes
    a, b = ctypes.c_double(), ctypes.c_double()
    for i, zz in enumerate(z):
        L.orc_faddeeva_w(zz.real, zz.imag, a, b)
        out[i] = complex(a.value, b.value)
    return out


def _num(s):
    table = {"Inf": math.inf, "-Inf": -math.inf, "NaN": math.nan, "-NaN": math.nan}
    return table[s.strip()] if s.strip() in table else float(s)


def _component_ok(g, e, tol):
    """Componentwise relative error within tol; NaN and infinities exact."""
    for a, b in ((e.real, g.real), (e.imag, g.imag)):
        if math.isnan(a):
            if not math.isnan(b):
                return False
        elif math.isinf(a) or a == 0.0:
            if b != a:
                return False
        elif abs(b - a) > tol * abs(a):
            return False
    return True


def test_known_answers(dev):
    """The reference's 57 known answers (src/Faddeeva.cpp:1922-2108), each
    component to 1e-13 relative (the reference's own criterion), the
    infinities and NaNs included."""
    kat = json.load(open(os.path.join(HERE, "golden", "faddeeva_kat.json")))
    z = np.array([complex(_num(a), _num(b)) for a, b in kat["z"]])
    w_exp = np.array([complex(_num(a), _num(b)) for a, b in kat["w"]])
    got = device_w(dev, z)
    for i in range(len(z)):
        assert _component_ok(got[i], w_exp[i], 1e-13), (i, z[i], got[i], w_exp[i])


def test_against_mpmath(dev):
    """2400 points over every region of the method and both half-planes
    against 400-digit values (tests/golden/wofz_mpmath.json): each component
    to 1e-13 relative, 2e-13 below the real axis where 2 exp(-z^2) - w(-z)
    cancels (the bound of the formula, not of the method)."""
    pts = json.load(open(os.path.join(HERE, "golden", "wofz_mpmath.json")))["points"]
    z = np.array([complex(float(a), float(b)) for a, b, _, _ in pts])
    w_exp = np.array([complex(float(c), float(d)) for _, _, c, d in pts])
    got = device_w(dev, z)
    worst = 0.0
    for i in range(len(z)):
        tol = 2e-13 if z[i].imag < 0 else 1e-13
        assert _component_ok(got[i], w_exp[i], tol), (z[i], got[i], w_exp[i])
        for a, b in ((w_exp[i].real, got[i].real), (w_exp[i].imag, got[i].imag)):
            if a != 0:
                worst = max(worst, abs(b - a) / abs(a))
    print(f"w(z) vs 400-digit values: worst component relative error {worst:.3g}")


def test_against_oracle_grid(dev):
    """A dense grid of the first quadrant (where the kicks evaluate w) plus
    the lower half-plane and the axes, against the oracle's restatement of
    the reference's package: |w| relative error 1e-13."""
    rng = np.random.default_rng(5)
    z = np.concatenate([
        rng.uniform(0, 12, 20000) + 1j * rng.uniform(0, 9, 20000),
        rng.uniform(0, 10, 5000) + 1j * 10 ** rng.uniform(-12, 0, 5000),
        10 ** rng.uniform(-6, 1, 5000) + 1j * rng.uniform(0, 7, 5000),
        rng.uniform(-8, 8, 3000) + 1j * rng.uniform(-5, 0, 3000),
        rng.uniform(0, 30, 500) + 0j, 1j * rng.uniform(-5, 30, 500),
    ])
    got = device_w(dev, z)
    exp = oracle_w(z)
    rel = np.abs(got - exp) / np.abs(exp)
    print(f"max |w| relative error against the oracle: {np.max(rel):.3g}")
    assert np.max(rel) <= 1e-13, z[np.argmax(rel)]


def test_kick_field_against_oracle(dev):
    """The kicks' field (per-beam constants, Weideman's approximation for
    both w(z) on Re z < 100, Im z < 1000) against the oracle's
    MPBunch::SSIonBunchInteraction branches (src/MPBunch.cpp:1128-1147):
    flat, inverted, near-round and round beams; points from the centre to far
    outside (the package's continued-fraction and large-x regions), on the
    axes and at the zero-field cut.  Tolerance: 1e-12 |f| + 2e-13 / max(sx, sy)
    (w's |w| relative error of 3e-14 times the Bassetti-Erskine prefactor)."""
    import ctypes
    rng = np.random.default_rng(17)
    beams = [(1e-4, 1e-5), (2e-5, 4e-6), (1e-5, 1e-4), (1e-5, 9e-6), (2e-4, 4e-6), (1e-5, 1.00005e-5),
             (3e-5, 3e-5)]
    pts = []
    for sx, sy in beams:
        u = np.concatenate([rng.uniform(-12, 12, 3000), np.sign(rng.uniform(-1, 1, 1000)) * 10 ** rng.uniform(-7, 0, 1000),
                            rng.uniform(-150, 150, 500), np.zeros(100), rng.uniform(-5, 5, 100)])
        v = np.concatenate([rng.uniform(-12, 12, 3000), np.sign(rng.uniform(-1, 1, 1000)) * 10 ** rng.uniform(-7, 0, 1000),
                            rng.uniform(-150, 150, 500), rng.uniform(-5, 5, 100), np.zeros(100)])
        v[:20] = 0.0
        u[20:25] = 3e-6  # |x/sx| + |y/sy| below the zero-field cut
        v[20:25] = 1e-6
        pts.append(np.stack([u * sx, v * sy, np.full(len(u), sx), np.full(len(u), sy)]))
    P = np.concatenate(pts, axis=1)
    n = P.shape[1]
    inp = np.ascontiguousarray(P.reshape(-1))
    out = np.empty(2 * n)
    capi.check(dev.lib.cts_debug_kick_field(dev.h, n, capi.dptr(inp), capi.dptr(out)))
    L = O.lib()
    a, b = ctypes.c_double(), ctypes.c_double()
    exp = np.empty((2, n))
    for i in range(n):
        L.orc_ion_field(P[0, i], P[1, i], P[2, i], P[3, i], a, b)
        exp[0, i], exp[1, i] = a.value, b.value
    got = out.reshape(2, n)
    scale = 1.0 / np.maximum(P[2], P[3])
    err = np.abs(got - exp)
    tol = 1e-12 * np.abs(exp) + 2e-13 * scale
    worst = np.max(err / (np.abs(exp) + scale))
    print(f"kick field: max |df| / (|f| + 1/max(sx, sy)) = {worst:.3g}")
    bad = np.argwhere(err > tol)
    assert len(bad) == 0, (bad[:5], P[:, bad[0][1]], got[:, bad[0][1]], exp[:, bad[0][1]])
