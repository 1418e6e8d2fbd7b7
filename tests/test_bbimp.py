"""The oracle's broadband impedance (longitudinal plane) pinned by what the
mathematics fixes: a constant resistive impedance R gives V = -R I bin by
bin; a reactive Z_k = i a k (an inductance on the file's frequency grid)
gives the spectral derivative of a cosine profile, V_j = a m sin(2 pi m j/n)
for I_j = cos(2 pi m j/n); the smoothing leaves a constant profile as it is,
keeps the sum, and spreads a single spike into the normalised Gaussian
window exp(-alpha^2 x^2 / 2), x = (j - 25)/25, alpha = 10; a quadrupolar
table Z_Q = -i R gives V_Q = +R I and px += V_Q x / (E beta^2 beta_x)."""
import numpy as np

import oracle as O


def potential(prof, zre, zim):
    n = len(prof)
    w = np.zeros(n)
    O.lib().orc_bbi_potential(O.p(np.ascontiguousarray(prof, float)), (n + 1) // 2,
                              O.p(np.ascontiguousarray(zre, float)), O.p(np.ascontiguousarray(zim, float)), O.p(w))
    return w


def test_resistive_and_reactive_impedance():
    nimp = 40
    n = 2 * nimp - 1
    rng = np.random.default_rng(1)
    prof = rng.random(n)
    w = potential(prof, np.full(nimp, 250.0), np.zeros(nimp))
    np.testing.assert_allclose(w, -250.0 * prof, rtol=1e-12, atol=1e-12 * 250)
    # Z_k = i a k on a cosine of m periods: V_j = a m sin(2 pi m j / n)
    a, m = 3.0, 5
    j = np.arange(n)
    prof = np.cos(2 * np.pi * m * j / n)
    w = potential(prof, np.zeros(nimp), a * np.arange(nimp))
    np.testing.assert_allclose(w, a * m * np.sin(2 * np.pi * m * j / n), rtol=0, atol=1e-12 * a * m)
    # the imaginary part of Z_0 drops out (c2r takes the real part at k = 0)
    zim = np.zeros(nimp); zim[0] = 7.0
    np.testing.assert_array_equal(potential(prof, np.zeros(nimp), zim), 0 * prof)


def test_smoothing():
    L = O.lib()
    p = np.full(80, 2.5)
    L.orc_bbi_smooth(O.p(p), 80)
    np.testing.assert_allclose(p, 2.5, rtol=1e-14)
    p = np.zeros(121)
    p[60] = 1.0
    L.orc_bbi_smooth(O.p(p), 121)
    x = (np.arange(51) - 25) / 25
    k = np.exp(-0.5 * 100 * x ** 2)
    k /= k.sum()
    np.testing.assert_allclose(p[35:86], k, rtol=1e-13, atol=1e-16)
    assert np.all(p[:35] == 0) and np.all(p[86:] == 0)
    np.testing.assert_allclose(p.sum(), 1.0, rtol=1e-14)
    # end values are padded: a step at the left edge keeps its level
    p = np.concatenate([np.full(30, 1.0), np.zeros(60)])
    L.orc_bbi_smooth(O.p(p), 90)
    assert abs(p[0] / p[1] - 1) < 1e-12


def test_quadrupolar_plane_sign_and_kick():
    """One bunch through orc_bbi_kick with the longitudinal plane off and
    Z_Qx = -i R: i Z_Qx = R, so V_Qx = +R I (the reference's + sign,
    src/MPBunch.cpp:2218-2233), and px changes by V_Qx(bin) x/(E beta^2 beta_x)."""
    import ctypes
    rng = np.random.default_rng(2)
    n0 = 3000
    a = {c: np.zeros(n0) for c in O.COORDS}
    a["z"] = rng.normal(0, 4e-3, n0)
    a["x"] = rng.normal(0, 1e-4, n0)
    b = O.Bunch(a)
    nimp, fmax, R = 261, 1.0e11, 500.0
    zero = np.zeros(nimp)
    zre = np.concatenate([zero, zero, zero])
    zim = np.concatenate([zero, np.full(nimp, -R), zero])
    planes = np.array([0, 1, 0], np.int32)
    n = 2 * nimp - 1
    wake, prof = np.zeros(3 * n), np.zeros(n)
    rbeta, t0, h, E, bx = 0.99999, 2e-6, 1000, 6e9, 5.0
    O.lib().orc_bbi_kick(ctypes.byref(b.s), nimp, fmax, planes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                         O.p(zre), O.p(zim), bx, 1.0, 1e-15, rbeta, t0, h, E, O.p(wake), O.p(prof))
    wake = wake.reshape(3, n)
    assert np.all(wake[0] == 0) and np.all(wake[2] == 0)
    np.testing.assert_allclose(wake[1], R * prof, rtol=0, atol=1e-12 * R * prof.max())
    dz = 299792458.0 * rbeta / (2 * fmax)
    k = ((a["z"] + dz * (nimp - 1) + dz / 2) / dz).astype(int)
    np.testing.assert_allclose(b["px"], wake[1][k] / E / rbeta ** 2 * a["x"] / bx, rtol=1e-13, atol=0)
    assert np.all(b["dp"] == 0)
