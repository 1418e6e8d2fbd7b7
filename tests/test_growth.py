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


# ---------------------------------------------------------------- full analysis
KEYS = ("ideal_amp", "ideal_arg", "iq_amp", "iq_arg", "ana_re", "ana_im", "hil_amp", "hil_arg")
RATES = ("rate_ideal", "rate_iq", "rate_hil", "rate_bunch")


def analysis(mean, harm, h, tunes, twiss, interval, fs, fe, t0):
    ns, nb, _ = mean.shape
    m = moments_array(mean)
    out = {k: np.zeros(ns * 3 * nb) for k in KEYS}
    out.update({k: np.zeros(3 * nb) for k in RATES})
    r = capi.CbmResult(*[capi.dptr(out[k]) for k in KEYS + RATES])
    hh = np.ascontiguousarray(harm, dtype=np.int32)
    rc = lib().cts_cbm_analysis(m, ns, nb, capi.i32ptr(hh), h, capi.dptr(np.array(tunes, float)),
                                capi.dptr(np.array(twiss, float)), interval, fs, fe, t0, ctypes.byref(r))
    assert rc == 0
    return {k: v.reshape(ns, 3, nb) for k, v in out.items() if k in KEYS} | \
        {k: v.reshape(3, nb) for k, v in out.items() if k in RATES}


def oracle_analysis(mean, harm, h, tunes, twiss, interval, fs, fe, t0):
    ns, nb, _ = mean.shape
    out = {k: np.zeros(ns * 3 * nb) for k in KEYS}
    out.update({k: np.zeros(3 * nb) for k in RATES})
    hh = (ctypes.c_int * nb)(*[int(x) for x in harm])
    mm = np.ascontiguousarray(mean, dtype=float)
    rc = O.lib().orc_cbm_analysis(O.p(mm), ns, nb, hh, h, O.p(np.array(tunes, float)), O.p(np.array(twiss, float)),
                                  interval, fs, fe, t0, *[O.p(out[k]) for k in KEYS + RATES])
    assert rc == 0
    return {k: v.reshape(ns, 3, nb) for k, v in out.items() if k in KEYS} | \
        {k: v.reshape(3, nb) for k, v in out.items() if k in RATES}


import pytest  # noqa: E402


@pytest.mark.parametrize("nb,ns", [(12, 40), (7, 41), (30, 64)])
def test_cbm_analysis_matches_oracle(nb, ns):
    """Mixed-radix FFTs (prime, composite, power of two) against the
    oracle's direct DFTs, every output of the analysis."""
    rng = np.random.default_rng(nb * 100 + ns)
    h = 600
    harm = np.sort(rng.choice(h, nb, replace=False))
    tunes, twiss, interval, t0 = (0.18, 0.27, 0.005), (0.3, 8.0, -0.2, 12.0, 2.5), 5, 2e-6
    turns = np.arange(ns)[:, None] * interval
    mean = rng.normal(0, 1e-5, (ns, nb, 6)) * np.exp(2e-3 * turns)[..., None]
    got = analysis(mean, harm, h, tunes, twiss, interval, 20, 20 + 25 * interval, t0)
    exp = oracle_analysis(mean, harm, h, tunes, twiss, interval, 20, 20 + 25 * interval, t0)
    for k in ("ideal_amp", "iq_amp", "hil_amp"):
        np.testing.assert_allclose(got[k], exp[k], rtol=1e-11, atol=1e-11 * np.max(np.abs(exp[k])))
    for k in ("ana_re", "ana_im"):
        np.testing.assert_allclose(got[k], exp[k], rtol=0, atol=1e-11 * np.max(np.abs(mean)))
    for k in ("ideal_arg", "iq_arg", "hil_arg"):
        d = np.angle(np.exp(1j * (got[k] - exp[k])))
        assert np.max(np.abs(d)) < 1e-9, k
    for k in RATES:
        np.testing.assert_allclose(got[k], exp[k], rtol=1e-8, atol=1e-6)


def test_analytic_signal_closed_form():
    """A centroid history c + A cos(2 pi k n / N + phi) with k whole cycles
    has the analytic signal c + A e^{i(2 pi k n / N + phi)}, for N even and
    odd."""
    for ns in (64, 45):
        nb = 3
        n = np.arange(ns)
        mean = np.zeros((ns, nb, 6))
        cases = [(1e-4, 2e-5, 5, 0.3), (-3e-5, 1e-5, 1, -1.1), (0.0, 4e-6, ns // 2 - 1, 2.0)]
        for b, (c, A, k, phi) in enumerate(cases):
            mean[:, b, 0] = c + A * np.cos(2 * np.pi * k * n / ns + phi)
        got = analysis(mean, [0, 10, 20], 100, (0.18, 0.27, 0.005), (0, 1, 0, 1, 1), 1, 0, ns, 1e-6)
        for b, (c, A, k, phi) in enumerate(cases):
            exp = c + A * np.exp(1j * (2 * np.pi * k * n / ns + phi))
            np.testing.assert_allclose(got["ana_re"][:, 0, b], exp.real, atol=1e-18)
            np.testing.assert_allclose(got["ana_im"][:, 0, b], exp.imag, atol=1e-18)


def test_iq_closed_form():
    """Centroids A cos(2 pi f_j t_i) of bunches equally spaced around the
    ring: the IQ component at f_j is A/2 (1 + G), G the geometric sum
    (1 - r^nb) / (nb (1 - r)), r = e^{-4 pi i (j + nu) / nb}."""
    nb, h, nu, t0 = 16, 320, 0.27, 2e-6
    harm = np.arange(nb) * (h // nb)
    f0, trf = 1 / t0, t0 / h
    A = 3e-5
    for j in (0, 3, 9):
        fj = (j + nu) * f0
        mean = np.zeros((2, nb, 6))
        mean[:, :, 2] = A * np.cos(2 * np.pi * fj * harm * trf)
        got = analysis(mean, harm, h, (0.18, nu, 0.005), (0, 1, 0, 1, 1), 1, 0, 2, t0)
        r = np.exp(-4j * np.pi * (j + nu) / nb)
        G = (1 - r ** nb) / (nb * (1 - r))
        assert got["iq_amp"][0, 1, j] == pytest.approx(abs(A / 2 * (1 + G)), rel=1e-12)
        assert got["iq_arg"][0, 1, j] == pytest.approx(np.angle(A / 2 * (1 + G)), abs=1e-12)


def test_hilbert_modes_follow_ideal_modes():
    """With unit beta and no alpha, centroids rotating in phase space
    (x = cos theta, px = -sin theta) over whole cycles give Hilbert modes
    equal to the ideal modes, and growth rates of an exponentially growing
    rotation agree."""
    nb, ns, h, interval = 8, 64, 80, 1
    harm = np.arange(nb) * (h // nb)
    nu = 5 / ns  # whole cycles over the samples
    n = np.arange(ns)[:, None]
    phi = 2 * np.pi * 0.37 * np.arange(nb)[None, :]
    theta = 2 * np.pi * nu * n + phi
    mean = np.zeros((ns, nb, 6))
    mean[:, :, 0] = 1e-5 * np.cos(theta)
    mean[:, :, 1] = -1e-5 * np.sin(theta)
    got = analysis(mean, harm, h, (nu, 0.27, 0.005), (0, 1, 0, 1, 1), interval, 0, ns, 1e-6)
    np.testing.assert_allclose(got["hil_amp"][:, 0], got["ideal_amp"][:, 0], rtol=1e-9,
                               atol=1e-9 * got["ideal_amp"][:, 0].max())
