"""Coupled-bunch-mode growth rates (host code of the library) against the
oracle's line fit, closed forms and a numpy DFT."""
import ctypes

import numpy as np

from cetasim_amd import capi

import oracle as O


def lib():
    return capi.load()


def test_fit_matches_oracle_and_closed_form():
    rng = np.random.default_rng(0)
    x = np.arange(50, dtype=float) * 10
    y = 0.37 + 2.5e-3 * x + rng.normal(0, 1e-3, 50)
    w = np.ones(50)
    c0, c1 = ctypes.c_double(), ctypes.c_double()
    assert lib().cts_fit_wlinear(capi.dptr(x), capi.dptr(w), capi.dptr(y), 50, ctypes.byref(c0), ctypes.byref(c1)) == 0
    o0, o1 = ctypes.c_double(), ctypes.c_double()
    O.lib().orc_fit_wlinear(O.p(x), O.p(w), O.p(y), 50, ctypes.byref(o0), ctypes.byref(o1))
    assert (c0.value, c1.value) == (o0.value, o1.value)
    # ordinary least squares closed form
    A = np.stack([np.ones(50), x], 1)
    ref = np.linalg.lstsq(A, y, rcond=None)[0]
    assert abs(c1.value - ref[1]) <= 1e-12 * abs(ref[1]) and abs(c0.value - ref[0]) <= 1e-12


def moments_array(mean):
    """cts_moments records with the given centroids mean[n][b][6]."""
    ns, nb, _ = mean.shape
    m = (capi.Moments * (ns * nb))()
    for n in range(ns):
        for b in range(nb):
            m[n * nb + b].mean[:] = list(mean[n, b])
    return m


def growth(mean, harm, h, tunes, twiss, interval, fs, fe, t0):
    ns, nb, _ = mean.shape
    m = moments_array(mean)
    rates = np.empty(3 * nb)
    amps = np.empty(ns * 3 * nb)
    hh = np.ascontiguousarray(harm, dtype=np.int32)
    rc = lib().cts_cbm_growth_from_moments(m, ns, nb, capi.i32ptr(hh), h, capi.dptr(np.array(tunes, float)),
                                           capi.dptr(np.array(twiss, float)), interval, fs, fe, t0,
                                           capi.dptr(rates), capi.dptr(amps))
    assert rc == 0
    return rates.reshape(3, nb), amps.reshape(ns, 3, nb)


def test_pure_mode_growth_closed_form():
    """Centroids that carry one coupled-bunch mode growing as e^{lam turn}
    give that mode the rate lam / T0."""
    nb, h, interval, ns = 16, 1000, 10, 60
    harm = np.arange(nb) * 2
    nux, nuy, nuz = 0.18, 0.27, 0.005
    bx, by, bz, ax, ay = 10.0, 8.0, 3.0, 0.0, 0.0
    lam, t0, k = 3e-4, 2e-6, 5
    mean = np.zeros((ns, nb, 6))
    for n in range(ns):
        turn = n * interval
        for b in range(nb):
            for plane, (beta, nu) in enumerate(((bx, nux), (by, nuy))):
                phase = -2 * np.pi * nu * harm[b] / h
                zb = 1e-6 * np.exp(lam * turn) * np.exp(2j * np.pi * k * b / nb)
                w = zb * np.exp(-1j * phase)   # = x/sqrt(beta) - i sqrt(beta) px
                mean[n, b, 2 * plane] = np.sqrt(beta) * w.real
                mean[n, b, 2 * plane + 1] = -w.imag / np.sqrt(beta)
            mean[n, b, 4] = 1e-4 * np.cos(0.3 * b)
            mean[n, b, 5] = 1e-5
    rates, amps = growth(mean, harm, h, (nux, nuy, nuz), (ax, bx, ay, by, bz), interval, 50, 550, t0)
    assert abs(rates[0, k] - lam / t0) <= 1e-8 * lam / t0
    assert abs(rates[1, k] - lam / t0) <= 1e-8 * lam / t0
    # the other transverse modes carry nothing but rounding
    others = [i for i in range(nb) if i != k]
    assert np.max(amps[:, 0, others]) <= 1e-12 * np.max(amps[:, 0, k])


def test_matches_numpy_dft():
    rng = np.random.default_rng(4)
    nb, ns, interval, h = 12, 30, 5, 1000
    harm = np.sort(rng.choice(500, nb, replace=False)) * 2
    mean = rng.normal(0, 1e-6, (ns, nb, 6)) * np.exp(np.linspace(0, 1, ns))[:, None, None]
    tw = (0.3, 9.0, -0.2, 7.0, 2.5)
    tunes = (0.18, 0.27, 0.005)
    rates, amps = growth(mean, harm, h, tunes, tw, interval, 10, 140, 2e-6)
    ax, bx, ay, by, bz = tw
    ph = lambda nu: np.exp(1j * (-2 * np.pi * nu * harm / h))
    zx = (mean[..., 0] / np.sqrt(bx) - 1j * (np.sqrt(bx) * mean[..., 1] + ax / np.sqrt(bx) * mean[..., 0])) * ph(tunes[0])
    zy = (mean[..., 2] / np.sqrt(by) - 1j * (np.sqrt(by) * mean[..., 3] + ay / np.sqrt(by) * mean[..., 2])) * ph(tunes[1])
    zz = (mean[..., 4] / np.sqrt(bz) + 1j * (np.sqrt(bz) * mean[..., 5])) * ph(tunes[2])
    A = np.abs(np.stack([np.fft.fft(zx, axis=1), np.fft.fft(zy, axis=1), np.fft.fft(zz, axis=1)], 1))
    assert np.allclose(amps, A, rtol=1e-12, atol=0)
    i0, i1 = 10 // interval, 140 // interval
    t = np.arange(i1 - i0) * interval
    for pl in range(3):
        for k in range(nb):
            slope = np.polyfit(t, np.log(A[i0:i1, pl, k]), 1)[0]
            assert abs(rates[pl, k] - slope / 2e-6) <= 1e-9 * abs(slope / 2e-6) + 1e-6
