"""The oracle's beam-loaded (rfMode 1) cavities pinned by closed forms: the
one-cavity phase that restores U0; without beam the kick is the ideal
cavity's at the bin centres; the generator settles at the fixed point
V = (omega_r R_s / 2 Q_0) I_g / (sigma - i delta_omega) of its dynamics; a
passive cavity driven by one bunch per turn settles at
V_b = V_b0 D / (1 - D), D = exp(-T0/t_F) exp(2 pi i f_r T0)."""
import ctypes

import numpy as np

import oracle as O

PI = np.pi
E = 1.602017733e-19


def tables(cav, nb, ne, gap, nturns, harmonic=100, f0=5e5, u0=1.2e6, rbeta=0.99999):
    L = O.lib()
    nres = len(cav["mode"])
    ip = np.ascontiguousarray(np.concatenate([cav[k] for k in ("mode", "type", "harm", "cold")]), np.int32)
    dp = np.ascontiguousarray(np.concatenate([np.asarray(cav[k], float)
                                              for k in ("rs", "q0", "coupling", "detune", "vol", "phase")]))
    nbl = int(np.sum(np.asarray(cav["mode"]) == 1))
    dev = np.zeros(nturns * nb * nbl * 3)
    info = np.zeros(nturns * nb * nres * 8)
    PIc = ctypes.POINTER(ctypes.c_int)
    g = np.ascontiguousarray(gap, np.int32)
    r = L.orc_bl_tables(nres, ip.ctypes.data_as(PIc), O.p(dp), harmonic, f0, 1 / f0, u0, rbeta, nb,
                        O.p(np.ascontiguousarray(ne, float)), g.ctypes.data_as(PIc), nturns, O.p(dev), O.p(info))
    assert r == nbl
    return dev.reshape(nturns, nb, nbl, 3), info.reshape(nturns, nb, nres, 8)


def cavity(type_=1, rs=1e5, q0=2e4, beta=2.0, detune=-3e4, vol=3e6, harm=1):
    return dict(mode=[1], type=[type_], harm=[harm], cold=[0], rs=[rs], q0=[q0], coupling=[beta],
                detune=[detune], vol=[vol], phase=[0.0])


def test_no_beam_is_the_ideal_cavity():
    h, f0, u0, vol = 100, 5e5, 1.2e6, 3e6
    dev, info = tables(cavity(vol=vol), 1, [0.0], [h], 3, harmonic=h, f0=f0, u0=u0)
    # the one-cavity phase pi - asin(U0/V) - pi/2: Re V = U0 at z = 0
    v = dev[0, 0, 0, 0] + 1j * dev[0, 0, 0, 1]
    np.testing.assert_allclose(v.real, u0, rtol=1e-14)
    np.testing.assert_allclose(abs(v), vol, rtol=1e-14)
    assert dev[0, 0, 0, 2] == 0.0
    # the kick at the bin centres equals the ideal cavity's there
    rng = np.random.default_rng(4)
    z = rng.normal(0, 3e-3, 2000)
    a = {c: np.zeros(2000) for c in O.COORDS}
    a["z"] = z
    b1, b2 = O.Bunch(a), O.Bunch(a)
    zmin, zmax, nbins = z.min() - 1e-6, z.max() + 1e-6, 50
    O.lib().orc_bl_kick(ctypes.byref(b1.s), v.real, v.imag, 0.0, zmin, zmax, nbins, 1, f0 * h, 0.99999, 6e9)
    dz = (zmax - zmin) / nbins
    k = np.floor((zmax - z) / dz)
    b2["z"][:] = zmax - k * dz - dz / 2
    O.lib().orc_rf_kick(ctypes.byref(b2.s), v.real, v.imag, 1, f0 * h, 0.99999, 6e9)
    np.testing.assert_allclose(b1["dp"], b2["dp"], rtol=1e-15, atol=0)


def test_generator_fixed_point():
    h, f0 = 100, 5e5
    cav = cavity(rs=1e5, q0=2e4, beta=2.0, detune=-3e4)
    dev, info = tables(cav, 1, [0.0], [h], 400, harmonic=h, f0=f0)
    gen = info[:, 0, 0, 4] + 1j * info[:, 0, 0, 5]
    # recompute I_g as GetInitialResonatorGenIg does from the turn-0 generator voltage
    fr = h * f0 + cav["detune"][0]
    ql = cav["q0"][0] / (1 + cav["coupling"][0])
    psi = np.arctan(ql * (fr / (h * f0) - h * f0 / fr))
    g0 = gen[0]
    ig = abs(g0) / abs(np.cos(psi)) * np.exp(1j * (np.angle(g0) - psi)) * (1 + cav["coupling"][0]) / cav["rs"][0]
    sigma = 2 * PI * fr / (2 * ql)
    dom = 2 * PI * cav["detune"][0]
    vstar = (2 * PI * fr * cav["rs"][0] / (2 * cav["q0"][0])) * ig / (sigma - 1j * dom)
    np.testing.assert_allclose(gen[-1], vstar, rtol=1e-9)
    assert abs(gen[1] - vstar) > 1e-6 * abs(vstar)  # it does move


def test_passive_cavity_steady_state():
    h, f0 = 100, 5e5
    cav = cavity(type_=0, rs=1e5, q0=2e4, beta=0.0, detune=4e4)
    ne = [6.0e9]
    # 4000 turns = 63 filling times: the start-up transient has decayed
    dev, info = tables(cav, 1, ne, [h], 4000, harmonic=h, f0=f0)
    fr = h * f0 + cav["detune"][0]
    ql = cav["q0"][0]
    tf = 2 * ql / (2 * PI * fr)
    T = h / (h * f0)
    D = np.exp(-T / tf) * np.exp(2j * PI * fr * T)
    vb0 = -2 * PI * fr * cav["rs"][0] / cav["q0"][0] * ne[0] * E
    vb = vb0 * D / (1 - D)
    # the voltage the bunch sees (no generator), and its self loss V_b0 / 2
    cav_v = dev[-1, 0, 0, 0] + 1j * dev[-1, 0, 0, 1]
    np.testing.assert_allclose(cav_v, vb, rtol=1e-9)
    np.testing.assert_allclose(dev[-1, 0, 0, 2], vb0, rtol=1e-15)
    np.testing.assert_allclose(info[-1, 0, 0, 6], vb0 / 2, rtol=1e-15)


def test_passive_cavity_uneven_train():
    # three bunches of different charge with gaps 10, 30, 60 buckets: the
    # voltages seen at arrival are the periodic solution of
    # V_{b+1} = (V_b + V_b0,b) D(gap_b), D(g) = exp(-g T_rf/t_F + 2 pi i f_r g T_rf)
    h, f0 = 100, 5e5
    cav = cavity(type_=0, rs=1e5, q0=2e4, beta=0.0, detune=4e4)
    ne = [6.0e9, 2.0e9, 9.0e9]
    gap = [10, 30, 60]
    dev, info = tables(cav, 3, ne, gap, 4000, harmonic=h, f0=f0)
    fr = h * f0 + cav["detune"][0]
    tf = 2 * cav["q0"][0] / (2 * PI * fr)
    trf = 1 / f0 / h
    D = [np.exp(-g * trf / tf) * np.exp(2j * PI * fr * g * trf) for g in gap]
    vb0 = [-2 * PI * fr * cav["rs"][0] / cav["q0"][0] * n * E for n in ne]
    # V_0 = (((V_0 + a0) D0 + a1) D1 + a2) D2  ->  V_0 (1 - D0 D1 D2) = ((a0 D0 + a1) D1 + a2) D2
    v0 = ((vb0[0] * D[0] + vb0[1]) * D[1] + vb0[2]) * D[2] / (1 - D[0] * D[1] * D[2])
    v1 = (v0 + vb0[0]) * D[0]
    v2 = (v1 + vb0[1]) * D[1]
    got = dev[-1, :, 0, 0] + 1j * dev[-1, :, 0, 1]
    np.testing.assert_allclose(got, [v0, v1, v2], rtol=1e-9)
    np.testing.assert_allclose(dev[-1, :, 0, 2], vb0, rtol=1e-15)
    # info: cavity = induced - V_b0 for a passive cavity, generator zero
    ind = info[-1, :, 0, 2] + 1j * info[-1, :, 0, 3]
    np.testing.assert_allclose(ind - np.array(vb0), got, rtol=1e-12)
    assert np.all(info[:, :, 0, 4:6] == 0.0)
