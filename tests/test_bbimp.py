"""The oracle's broadband impedance (longitudinal plane) pinned by what the
mathematics fixes: a constant resistive impedance R gives V = -R I bin by
bin; a reactive Z_k = i a k (an inductance on the file's frequency grid)
gives the spectral derivative of a cosine profile, V_j = a m sin(2 pi m j/n)
for I_j = cos(2 pi m j/n); the smoothing leaves a constant profile as it is,
keeps the sum, and spreads a single spike into the normalised Gaussian
window exp(-alpha^2 x^2 / 2), x = (j - 25)/25, alpha = 10."""
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
