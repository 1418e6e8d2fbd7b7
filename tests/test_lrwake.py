"""The oracle's long-range wake (WakeFunction::InitialLRWake's table and
MPBeam::LRWakeBeamIntaction's sum) pinned by what the mathematics fixes:
Yokoya factors of a round pipe (1) and of a flat chamber (1, pi^2/24,
pi^2/12, -pi^2/24, pi^2/24), the resistive wall's tau^-1/2 and tau^-3/2
laws and its scaling with length, sector count and conductivity, the
resonator's values at tau = 0 and its decay, and closed forms of the
per-turn sum (a single bunch; sources at rest; linearity)."""
import ctypes

import numpy as np
import pytest

import oracle as O

PI = np.pi


def yokoya(rx, ry):
    out = (ctypes.c_double * 5)()
    O.lib().orc_lrw_yokoya(rx, ry, out)
    return np.array(out[:])


def table(nturns, harmonic, t0, sectors=(), res=(), bx=5.0, by=7.0):
    L = O.lib()
    sec = np.ascontiguousarray(np.asarray(sectors, float).reshape(-1, 7))
    rs = np.ascontiguousarray(np.asarray(res, float).reshape(-1, 9))
    cap = 2 * nturns * harmonic + 4
    t, w, rw, bb = np.zeros(cap), np.zeros(3 * cap), np.zeros(3 * cap), np.zeros(3 * cap)
    n = L.orc_lrw_table(nturns, harmonic, t0, len(sec), O.p(sec), len(rs), O.p(rs), bx, by, cap, O.p(t), O.p(w),
                        O.p(rw), O.p(bb))
    assert n > 0
    return t[:n], w[:3 * n].reshape(3, n), rw[:3 * n].reshape(3, n), bb[:3 * n].reshape(3, n)


def test_yokoya_round_and_flat():
    np.testing.assert_array_equal(yokoya(0.01, 0.01), np.ones(5))
    flat = np.array([1.0, PI ** 2 / 24, PI ** 2 / 12, -PI ** 2 / 24, PI ** 2 / 24])
    np.testing.assert_allclose(yokoya(1.0, 0.05), flat, rtol=4e-3)
    # an upright ellipse exchanges the planes
    a, b = yokoya(0.02, 0.01), yokoya(0.01, 0.02)
    np.testing.assert_allclose(b[[0, 2, 1, 4, 3]], a, rtol=1e-15)
    # close to round: the dipole factors approach 1, the quadrupole ones 0
    c = yokoya(1.0, 0.999)
    assert np.all(np.abs(c[:3] - 1) < 5e-3) and np.all(np.abs(c[3:]) < 2e-3)


def test_resistive_wall_laws():
    t0, h = 1e-6, 20
    sec = [0.01, 0.01, 100.0, 4.0, 6.0, 1, 3.5e7]
    t, w, rw, bb = table(3, h, t0, sectors=[sec])
    np.testing.assert_array_equal(bb, 0.0)
    assert np.all(rw[:, 0] == 0.0)  # tau = 0 returns zero
    tau = t[1:]
    # W_x, W_y ~ -tau^{-1/2} (defocusing), W_z ~ tau^{-3/2}
    for k, pw in ((0, 0.5), (1, 0.5), (2, 1.5)):
        c = rw[k, 1:] * tau ** pw
        np.testing.assert_allclose(c, c[0], rtol=1e-13)
    assert rw[0, 1] < 0 and rw[1, 1] < 0 and rw[2, 1] > 0
    # ratio of the planes: sector beta over the interaction-point beta
    np.testing.assert_allclose(rw[0, 1:] / rw[1, 1:], (4.0 / 5.0) / (6.0 / 7.0), rtol=1e-14)
    # two identical sectors = one sector of twice the count; conductivity^-1/2
    _, _, rw2, _ = table(3, h, t0, sectors=[sec, sec])
    sec2 = list(sec); sec2[5] = 2
    _, _, rw3, _ = table(3, h, t0, sectors=[sec2])
    np.testing.assert_allclose(rw2, 2 * rw, rtol=1e-14)
    np.testing.assert_allclose(rw3, 2 * rw, rtol=1e-14)
    sec4 = list(sec); sec4[6] = 4 * sec[6]
    _, _, rw4, _ = table(3, h, t0, sectors=[sec4])
    np.testing.assert_allclose(rw4, 0.5 * rw, rtol=1e-14)
    # the grid: dt = T_rf / 2 and np = 2 nturns h + 2
    assert len(t) == 2 * 3 * h + 2
    np.testing.assert_allclose(np.diff(t), t0 / h / 2, rtol=1e-12)


def test_resonator_values():
    t0, h = 1e-6, 20
    rs, q, f = 1e4, 2.0, 1.3e9
    res = [rs, q, f, 2e5, 3.0, 1.1e9, 3e5, 4.0, 1.2e9]
    t, w, rw, bb = table(2, h, t0, res=[res], bx=5.0, by=7.0)
    np.testing.assert_array_equal(rw, 0.0)
    om = 2 * PI * f
    alpha = om / (2 * q)
    # W_z(0) = alpha Rs (half the jump of 2 alpha Rs); W_x(0) = W_y(0) = 0
    np.testing.assert_allclose(bb[2, 0], alpha * rs, rtol=1e-14)
    assert bb[0, 0] == 0.0 and bb[1, 0] == 0.0
    # the envelope decays as exp(-alpha |tau|): |W| <= 2 alpha Rs sqrt(1 + alpha^2/omega_b^2) e^{-alpha tau}
    omb = np.sqrt(om ** 2 - alpha ** 2)
    env = 2 * alpha * rs * np.sqrt(1 + (alpha / omb) ** 2) * np.exp(-alpha * t[1:])
    assert np.all(np.abs(bb[2, 1:]) <= env * (1 + 1e-12))
    # transverse: c Rs omega / (Q omega_b) e^{-alpha tau} sin(-omega_b tau) / beta
    ox = 2 * PI * 1.1e9
    ax = ox / 6.0
    obx = np.sqrt(ox ** 2 - ax ** 2)
    exp_x = 299792458.0 * 2e5 * ox / 3.0 / obx * np.exp(-ax * t) * np.sin(-obx * t) / 5.0
    np.testing.assert_allclose(bb[0], exp_x, rtol=1e-9, atol=1e-12 * np.max(np.abs(exp_x)))


def lrw_kick(nturns, harm, harmonic, t0, rbeta, energy, ne, pos, w):
    L = O.lib()
    nb = len(harm)
    h = np.ascontiguousarray(harm, np.int32)
    kick = np.zeros(3 * nb)
    rc = L.orc_lrw_kick(nturns, nb, h.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), harmonic, t0, rbeta, energy,
                        O.p(np.ascontiguousarray(ne, float)), O.p(np.ascontiguousarray(pos, float).reshape(-1)),
                        O.p(np.ascontiguousarray(w.reshape(-1))), w.shape[1], O.p(kick))
    return rc, kick.reshape(nb, 3)


def test_kick_closed_forms():
    t0, h, T = 1e-6, 20, 4
    sec = [0.01, 0.01, 100.0, 4.0, 6.0, 1, 3.5e7]
    res = [1e4, 2.0, 1.3e9, 2e5, 3.0, 1.1e9, 3e5, 4.0, 1.2e9]
    t, w, _, _ = table(T, h, t0, sectors=[sec], res=[res])
    rbeta, energy = 0.99999, 3e9
    scale = 1.602017733e-19 / energy / rbeta ** 2
    # one bunch, at rest on the axis: the earlier turns' longitudinal wake
    # at tau = -n h T_rf (table points 2 n h)
    pos = np.zeros((T, 1, 3))
    rc, k = lrw_kick(T, [0], h, t0, rbeta, energy, [1e10], pos, w)
    assert rc == 0
    exp_z = -sum(w[2, 2 * n * h] for n in range(1, T)) * 1e10 * scale
    np.testing.assert_allclose(k[0, 2], exp_z, rtol=1e-12)
    assert k[0, 0] == 0.0 and k[0, 1] == 0.0
    # three bunches: the transverse kick is linear in the source centroids,
    # and the present turn acts only from the bunches ahead
    harm = [0, 3, 7]
    ne = [1e10, 2e10, 1.5e10]
    rng = np.random.default_rng(3)
    pos = np.zeros((T, 3, 3))
    pos[:, :, :2] = rng.normal(0, 1e-4, (T, 3, 2))
    rc, k1 = lrw_kick(T, harm, h, t0, rbeta, energy, ne, pos, w)
    pos2 = pos.copy(); pos2[:, :, :2] *= 2
    rc2, k2 = lrw_kick(T, harm, h, t0, rbeta, energy, ne, pos2, w)
    assert rc == 0 and rc2 == 0
    np.testing.assert_allclose(k2[:, :2], 2 * k1[:, :2], rtol=1e-13)
    np.testing.assert_allclose(k2[:, 2], k1[:, 2], rtol=1e-15)
    # bunch 2 receives from bunch 0 of the present turn: tau = -7 T_rf, table point 14
    pos3 = np.zeros((T, 3, 3)); pos3[T - 1, 0, 0] = 1e-4
    _, k3 = lrw_kick(T, harm, h, t0, rbeta, energy, ne, pos3, w)
    np.testing.assert_allclose(k3[2, 0], -w[0, 14] * ne[0] * 1e-4 * scale, rtol=1e-12)
    np.testing.assert_allclose(k3[1, 0], -w[0, 6] * ne[0] * 1e-4 * scale, rtol=1e-12)
    assert k3[0, 0] == 0.0  # bunch 0 sees bunch 0 of the present turn from nobody
    # a source ahead in time is an error (the reference exits)
    pos4 = np.zeros((T, 3, 3)); pos4[T - 1, 0, 2] = -1.0  # same bucket, 1 m apart: tau > 0 for a follower
    rc4, _ = lrw_kick(T, [0, 0, 0], h, t0, rbeta, energy, ne, pos4, w)
    assert rc4 == -1
