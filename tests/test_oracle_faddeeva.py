"""Pins the oracle's w(z) and Bassetti-Erskine field (CPU only)."""
import ctypes
import json
import math
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def w(z, L=None):
    L = L or O.lib()
    a, b = ctypes.c_double(), ctypes.c_double()
    (L.orc_faddeeva_w if hasattr(L, "orc_faddeeva_w") else L.ref_faddeeva_w)(z.real, z.imag, a, b)
    return complex(a.value, b.value)


def relerr(a, b):
    """src/Faddeeva.cpp:1904-1916: a = expected, b = computed."""
    if math.isnan(a) or math.isnan(b) or math.isinf(a) or math.isinf(b):
        if (math.isnan(a) != math.isnan(b)) or (math.isinf(a) != math.isinf(b)) or \
                (math.isinf(a) and math.isinf(b) and a * b < 0):
            return math.inf
        return 0.0
    if a == 0:
        return 0.0 if b == 0 else math.inf
    return abs((b - a) / a)


def parse(s):
    s = s.strip()
    table = {"Inf": math.inf, "-Inf": -math.inf, "NaN": math.nan, "-NaN": math.nan}
    return table[s] if s in table else float(s)


def test_kat_57_points():
    """The reference's own known answers (src/Faddeeva.cpp:1922-2108): max
    relative error <= 1e-13 on each component."""
    kat = json.load(open(os.path.join(HERE, "golden", "faddeeva_kat.json")))
    worst = 0.0
    for (zr, zi), (wr, wi) in zip(kat["z"], kat["w"]):
        z = complex(parse(zr), parse(zi))
        got = w(z)
        e = max(relerr(parse(wr), got.real), relerr(parse(wi), got.imag))
        assert e <= 1e-13, (zr, zi, got, wr, wi, e)
        worst = max(worst, e)
    assert worst <= kat["tolerance"]


def test_against_mpmath_values():
    """2400 points of tests/golden/wofz_mpmath.json (400-digit values): the
    oracle's w(z) to 1e-13 in |w| and 5e-13 per component (the reference
    package's own accuracy next to the imaginary axis is about 2e-13)."""
    pts = json.load(open(os.path.join(HERE, "golden", "wofz_mpmath.json")))["points"]
    worst_abs = worst_comp = 0.0
    for a, b, c, d in pts:
        z = complex(float(a), float(b))
        e = complex(float(c), float(d))
        got = w(z)
        worst_abs = max(worst_abs, abs(got - e) / abs(e))
        worst_comp = max(worst_comp, relerr(e.real, got.real), relerr(e.imag, got.imag))
    assert worst_abs <= 1e-13, worst_abs
    assert worst_comp <= 5e-13, worst_comp


def test_against_reference_build():
    """oracle/_ref is the reference's Faddeeva.cpp compiled as shipped."""
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    rng = np.random.default_rng(7)
    worst = 0.0
    for scale in (1e-5, 1e-2, 0.3, 1, 3, 6.5, 9, 15, 40, 300, 1e5):
        for _ in range(400):
            z = complex(rng.normal() * scale, rng.normal() * scale)
            b = w(z, R)
            if not (math.isfinite(b.real) and math.isfinite(b.imag)) or abs(b) == 0:
                continue
            a = w(z)
            worst = max(worst, abs(a - b) / abs(b))
    assert worst < 5e-14, worst
    for x in np.concatenate([rng.uniform(-26, 30, 500), rng.uniform(-1e-3, 1e-3, 200)]):
        assert abs(O.lib().orc_erfcx(x) - R.ref_erfcx(x)) <= 2e-13 * abs(R.ref_erfcx(x))
        rw = R.ref_w_im(x)
        assert abs(O.lib().orc_w_im(x) - rw) <= 5e-14 * abs(rw)
    assert O.lib().orc_erfcx(-29.5) == R.ref_erfcx(-29.5) == math.inf


def test_identities():
    """w(-z) = 2 exp(-z^2) - w(z); w(iy) = erfcx(y); w(z) ~ i/(sqrt(pi) z)."""
    rng = np.random.default_rng(3)
    for _ in range(500):
        z = complex(rng.normal() * 3, rng.normal() * 3)
        lhs = w(-z)
        e2 = 2 * np.exp(-z * z)
        rhs = e2 - w(z)
        # the right side cancels: its rounding scales with |2 exp(-z^2)|
        assert abs(lhs - rhs) <= 1e-13 * max(abs(lhs), abs(e2), 1.0)
    for y in (0.0, 0.1, 1.0, 5.0, 30.0, -0.5):
        assert abs(w(complex(0.0, y)).real - O.lib().orc_erfcx(y)) <= 1e-15 * O.lib().orc_erfcx(y)
    for z in (complex(1e4, 3e3), complex(-2e3, 5e3), complex(3e8, 1.0)):
        asym = 1j / (math.sqrt(math.pi) * z)
        assert abs(w(z) - asym) <= 1e-6 * abs(asym)
    # w(0) = 1, Re w(x) = exp(-x^2) on the real axis
    assert w(complex(1e-300, 0.0)).real == pytest.approx(1.0, rel=1e-15)
    assert w(complex(1.7, 0.0)).real == pytest.approx(math.exp(-1.7 ** 2), rel=1e-15)


def be(x, y, sx, sy, fn="orc_bassetti_erskine"):
    a, b = ctypes.c_double(), ctypes.c_double()
    getattr(O.lib(), fn)(x, y, sx, sy, a, b)
    return a.value, b.value


def field_quadrature(x, y, sx, sy):
    """Field of a unit 2-D Gaussian charge in the reference's units
    (F -> -(x, y)/r^2 far away): -(x,y) * int_0^inf exp(-x^2/(2sx^2+q)
    - y^2/(2sy^2+q)) / (sqrt(2sx^2+q) sqrt(2sy^2+q)) * (1/(2sx^2+q), 1/(2sy^2+q)) dq."""
    from scipy.integrate import quad

    s = 2 * sx * sx  # q = s*u makes the integrand dimensionless

    def g(u, which):
        ax, ay = 1.0 + u, (sy * sy) / (sx * sx) + u
        e = math.exp(-x * x / (s * ax) - y * y / (s * ay)) / math.sqrt(ax * ay)
        return e / (ax if which == 0 else ay)

    def integral(which):
        edges = [0.0, 1e-3, 1e-2, 0.1, 1.0, 10.0, 1e2, 1e3, 1e4, 1e5, 1e6]
        tot = sum(quad(g, a, b, args=(which,), epsabs=0, epsrel=1e-13, limit=200)[0]
                  for a, b in zip(edges[:-1], edges[1:]))
        return tot + quad(g, edges[-1], math.inf, args=(which,), epsabs=0, epsrel=1e-13, limit=200)[0]

    return -x * integral(0) / s, -y * integral(1) / s


def test_bassetti_erskine_brute_force():
    """Quadrature of the Gaussian-charge field integral fixes magnitude,
    sign and the x<->y roles (a swapped operand or a dropped term fails)."""
    for (x, y, sx, sy) in [(1e-5, 2e-6, 2e-5, 4e-6), (-3e-5, 1e-6, 2e-5, 4e-6), (4e-5, -9e-6, 2e-5, 4e-6),
                           (1e-6, 1e-6, 1e-5, 8e-6), (2e-4, 1e-4, 2e-5, 4e-6)]:
        fx, fy = be(x, y, sx, sy)
        qx, qy = field_quadrature(x, y, sx, sy)
        assert fx == pytest.approx(qx, rel=1e-8, abs=1e-10 * abs(qy))
        assert fy == pytest.approx(qy, rel=1e-8, abs=1e-10 * abs(qx))


def test_bassetti_erskine_symmetry_and_limits():
    sx, sy = 2e-5, 4e-6
    fx, fy = be(1e-5, 3e-6, sx, sy)
    assert be(-1e-5, 3e-6, sx, sy) == pytest.approx((-fx, fy), rel=1e-15)
    assert be(1e-5, -3e-6, sx, sy) == pytest.approx((fx, -fy), rel=1e-15)
    # far field -> -(x, y)/r^2
    x, y = 3e-3, 2e-3
    fx, fy = be(x, y, sx, sy)
    r2 = x * x + y * y
    assert fx == pytest.approx(-x / r2, rel=1e-4)
    assert fy == pytest.approx(-y / r2, rel=1e-4)
    # nearly round beam -> GaussianField
    sxr, syr = 1e-5, 1e-5 * (1 - 2e-5)
    gx, gy = be(7e-6, 5e-6, sxr, sxr, "orc_gaussian_field")
    bx, by = be(7e-6, 5e-6, sxr, syr * 0.999)
    assert bx == pytest.approx(gx, rel=2e-3)
    assert by == pytest.approx(gy, rel=2e-3)


def test_ion_field_branches():
    sx, sy = 2e-5, 4e-6
    # x/y swap branch equals the transposed call
    fx1, fy1 = be(3e-6, 1e-5, sy, sx, "orc_ion_field")
    fy2, fx2 = be(1e-5, 3e-6, sx, sy, "orc_ion_field")
    assert (fx1, fy1) == (fx2, fy2)
    # at the centre the field is zero
    assert be(0.0, 0.0, sx, sy, "orc_ion_field") == (0.0, 0.0)
