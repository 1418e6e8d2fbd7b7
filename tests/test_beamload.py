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


# ---------------------------------------------------------------------------
# The walk over the train (orc_bl_state / orc_bl_walk): the variants whose
# cavity state depends on the particles.

def walk_setup(cav, ne, gap, harmonic=100, f0=5e5, u0=1.2e6):
    L = O.lib()
    nres = len(cav["mode"])
    ip = np.ascontiguousarray(np.concatenate([cav[k] for k in ("mode", "type", "harm", "cold", "excite", "dirfb")]),
                              np.int32)
    dp = np.ascontiguousarray(np.concatenate([np.asarray(cav[k], float) for k in
                                              ("rs", "q0", "coupling", "detune", "vol", "phase", "fb_gain",
                                               "fb_phase", "fb_delay")]))
    st = np.zeros(nres * L.orc_bl_state_size())
    PIc = ctypes.POINTER(ctypes.c_int)
    g = np.ascontiguousarray(gap, np.int32)
    L.orc_bl_state(nres, ip.ctypes.data_as(PIc), O.p(dp), harmonic, f0, u0, len(ne),
                   O.p(np.ascontiguousarray(ne, float)), g.ctypes.data_as(PIc), O.p(st))
    return st


def walk(st, nres, bunches, gap, ne, mq, nbins, harmonic=100, f0=5e5, rbeta=0.99999, energy=6e9):
    L = O.lib()
    nb = len(bunches)
    arr = (O.OrcBunch * nb)(*[b.s for b in bunches])
    info, fb = np.zeros(nb * nres * 8), np.zeros(nb * nres)
    g = np.ascontiguousarray(gap, np.int32)
    L.orc_bl_walk(O.p(st), nres, arr, nb, g.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                  O.p(np.ascontiguousarray(ne, float)), O.p(np.ascontiguousarray(mq, float)), nbins, harmonic,
                  f0, 1 / f0, rbeta, energy, O.p(info), O.p(fb))
    return info.reshape(nb, nres, 8), fb.reshape(nb, nres)


def wcav(excite=0, dirfb=0, gain=1.0, fbp=0.0, fbd=0.0, **kw):
    c = cavity(**kw)
    c.update(excite=[excite], dirfb=[dirfb], fb_gain=[gain], fb_phase=[fbp], fb_delay=[fbd])
    return c


def train(nb, n, seed, dz=0.0):
    rng = np.random.default_rng(seed)
    out = []
    for b in range(nb):
        a = {c: np.zeros(n) for c in O.COORDS}
        a["z"] = rng.normal(dz * b, 3e-3, n)
        a["dp"] = rng.normal(0, 1e-3, n)
        out.append(a)
    return out


def test_walk_stable_equals_tables():
    """With the stable variant and no feedback the walk kicks every particle
    exactly as the tabulated path does, turn after turn."""
    cav = wcav()
    nb, n, nbins, h = 3, 1500, 40, 100
    gap = [30, 30, 40]
    ne = [4e10, 5e10, 3e10]
    parts = train(nb, n, 5)
    dev, _ = tables(cav, nb, ne, gap, 3, harmonic=h)
    st = walk_setup(cav, ne, gap, harmonic=h)
    bt = [O.Bunch(a) for a in parts]
    bw = [O.Bunch(a) for a in parts]
    for t in range(3):
        for b in range(nb):
            zmin, zmax = ctypes.c_double(), ctypes.c_double()
            O.lib().orc_zminmax(ctypes.byref(bt[b].s), ctypes.byref(zmin), ctypes.byref(zmax))
            vr, vi, vb0 = dev[t, b, 0]
            O.lib().orc_bl_kick(ctypes.byref(bt[b].s), vr, vi, vb0, zmin.value, zmax.value, nbins, 1, 5e5 * h,
                                0.99999, 6e9)
        walk(st, 1, bw, gap, ne, [ne[b] / n for b in range(nb)], nbins, harmonic=h)
        for b in range(nb):
            np.testing.assert_array_equal(bw[b]["dp"], bt[b]["dp"])


def test_walk_bin_by_bin_passive_cavity():
    """Bin-by-bin variant, a cold passive cavity (no generator, nothing
    induced before the first bunch): bin k of the first bunch sees the sum of
    the earlier bins' voltages, V_k = sum_{m<k} V0_m D^{k-m}, with
    V0_m = -2 pi f_r R_s / Q_0 q n_m e and D = exp(-dt/t_F + 2 pi i f_r dt),
    dt the bin length in time; each particle gets Re V_k + Re V0_k / 2."""
    rs, q0, beta, det = 2e5, 1e4, 1.0, -5e4
    cav = wcav(excite=1, type_=0, rs=rs, q0=q0, beta=beta, detune=det)
    cav["cold"] = [1]
    n, nbins, h, f0, rbeta, energy = 4000, 30, 100, 5e5, 0.99999, 6e9
    ne = [6e10]
    mq = ne[0] / n
    a = train(1, n, 8)[0]
    b = O.Bunch(a)
    st = walk_setup(cav, ne, [h], harmonic=h, f0=f0)
    walk(st, 1, [b], [h], ne, [mq], nbins, harmonic=h, f0=f0, rbeta=rbeta, energy=energy)
    z = a["z"]
    zmin, zmax = z.min() - 1e-6, z.max() + 1e-6
    dz = (zmax - zmin) / nbins
    k = np.floor((zmax - z) / dz).astype(int)
    cnt = np.bincount(k, minlength=nbins)
    fre = h * f0 + det
    tf = 2 * (q0 / (1 + beta)) / (2 * PI * fre)
    dt = dz / 299792458.0 / rbeta
    D = np.exp(-dt / tf) * np.exp(2j * PI * fre * dt)
    v0 = -2 * PI * fre * rs / q0 * mq * cnt * E
    vk = np.array([sum(v0[m] * D ** (j - m) for m in range(j)) for j in range(nbins)])
    kick = (vk.real[k] + v0[k] / 2) / energy / rbeta ** 2
    # the kicks to 1e-12 of their scale, beside the rounding of delta + kick
    tol = 1e-12 * np.max(np.abs(kick)) + 4 * np.spacing(np.max(np.abs(a["dp"])))
    np.testing.assert_allclose(b["dp"] - a["dp"], kick, rtol=0, atol=tol)
    assert np.max(np.abs(vk)) > 0


def test_walk_time_to_next_bunch():
    """Bin-by-bin variant: the voltage the second bunch's head meets decays
    and turns over the time from the first bunch's tail to its head,
    gap tRF + (zmin_1 - zmax_2)/(c beta): moving the second bunch by dz
    multiplies it by exp(-T/t_F) exp(2 pi i f_r T) with T = -dz/(c beta)
    (the head bin's own charge enters only the next bins)."""
    rs, q0, beta, det = 2e5, 1e4, 1.0, -5e4
    cav = wcav(excite=1, type_=0, rs=rs, q0=q0, beta=beta, detune=det)
    cav["cold"] = [1]
    n, nbins, h, f0, rbeta, energy = 3000, 20, 100, 5e5, 0.99999, 6e9
    ne, gap = [5e10, 5e10], [10, 90]
    mq = [ne[0] / n, ne[1] / n]
    heads = []
    shift = 2e-3
    for s in (0.0, shift):
        a0, a1 = train(2, n, 3)
        a1 = dict(a1)
        a1["z"] = a1["z"] + s
        bs = [O.Bunch(a0), O.Bunch(a1)]
        st = walk_setup(cav, ne, gap, harmonic=h, f0=f0)
        info, _ = walk(st, 1, bs, gap, ne, mq, nbins, harmonic=h, f0=f0, rbeta=rbeta, energy=energy)
        # the head particle of bunch 1 (its first bin sees the carried voltage only)
        z1 = a1["z"]
        i_head = np.argmax(z1)
        kick = (bs[1]["dp"][i_head] - a1["dp"][i_head]) * energy * rbeta ** 2
        zmin1, zmax1 = z1.min() - 1e-6, z1.max() + 1e-6
        n0 = np.sum(np.floor((zmax1 - z1) / ((zmax1 - zmin1) / nbins)) == 0)
        v00 = -2 * PI * (h * f0 + det) * rs / q0 * mq[1] * n0 * E
        heads.append(kick - v00 / 2)  # Re of the carried voltage
        # and the complex voltage itself from the phase of two variants is not
        # observable from one kick; keep the carried voltage's real part
    fre = h * f0 + det
    tf = 2 * (q0 / (1 + beta)) / (2 * PI * fre)
    T = -shift / 299792458.0 / rbeta
    # the carried voltage V at the head for shift 0; its real part is heads[0];
    # for the shifted bunch Re(V exp(-T/tf) exp(2 pi i fre T))
    # the imaginary part is fixed by the first bunch alone: rebuild V from bunch 0
    a0 = train(2, n, 3)[0]
    z0 = a0["z"]
    zmin0, zmax0 = z0.min() - 1e-6, z0.max() + 1e-6
    dz0 = (zmax0 - zmin0) / nbins
    cnt0 = np.bincount(np.floor((zmax0 - z0) / dz0).astype(int), minlength=nbins)
    D0 = np.exp(-dz0 / 299792458.0 / rbeta / tf) * np.exp(2j * PI * fre * dz0 / 299792458.0 / rbeta)
    v0 = -2 * PI * fre * rs / q0 * mq[0] * cnt0 * E
    vb = 0j
    for k in range(nbins):
        vb = (vb + v0[k]) * D0
    a1 = train(2, n, 3)[1]
    T0 = gap[0] / (h * f0) + (zmin0 - (a1["z"].max() + 1e-6)) / 299792458.0 / rbeta
    V = vb * np.exp(-T0 / tf) * np.exp(2j * PI * fre * T0)
    np.testing.assert_allclose(heads[0], V.real, rtol=1e-10)
    V2 = V * np.exp(-T / tf) * np.exp(2j * PI * fre * T)
    np.testing.assert_allclose(heads[1], V2.real, rtol=1e-10)


def test_walk_direct_feedback():
    """The direct-feedback kick: without beam the cavity sits at the
    requested voltage and the feedback kicks nothing; with beam it is linear
    in the gain, turns sign with a phase shift of pi or a delay of half an
    RF period of the cavity, and reaches every particle, lost ones too."""
    h, f0 = 100, 5e5
    n, nbins = 2000, 30
    gap, ne0 = [h], [0.0]
    a = train(1, n, 2)[0]
    st = walk_setup(wcav(dirfb=1), ne0, gap, harmonic=h, f0=f0)
    b = O.Bunch(a)
    a_z = dict(a)
    _, fb = walk(st, 1, [b], gap, ne0, [0.0], nbins, harmonic=h, f0=f0)
    # no beam: generator = required voltage; the sample at the bunch's bucket
    # is taken at t = -zmin/c, where only the beam-induced part would move
    assert abs(fb[0, 0]) <= 1e-9 * 3e6
    ne = [5e10]
    res = {}
    fre = h * f0 - 3e4
    for key, kw in (("g1", {}), ("g2", dict(gain=2.0)), ("pi", dict(fbp=PI)), ("half", dict(fbd=0.5 / fre))):
        lost = np.zeros(n, np.int32)
        lost[:10] = 1
        b = O.Bunch(a, lost)
        st = walk_setup(wcav(dirfb=1, **kw), ne, gap, harmonic=h, f0=f0)
        _, fb = walk(st, 1, [b], gap, ne, [ne[0] / n], nbins, harmonic=h, f0=f0)
        res[key] = (fb[0, 0], b)
    f1 = res["g1"][0]
    assert abs(f1) > 0
    np.testing.assert_allclose(res["g2"][0], 2 * f1, rtol=1e-13)
    np.testing.assert_allclose(res["pi"][0], -f1, rtol=1e-12)
    np.testing.assert_allclose(res["half"][0], -f1, rtol=1e-9)
    # lost particles get the feedback kick and nothing else
    b1 = res["g1"][1]
    np.testing.assert_array_equal(b1["dp"][:10], a["dp"][:10] + f1 / 6e9 / 0.99999 ** 2)
