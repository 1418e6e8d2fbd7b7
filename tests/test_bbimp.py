"""The oracle's broadband impedance pinned by what the
mathematics fixes: a constant resistive impedance R gives V = -R I bin by
bin; a reactive Z_k = i a k (an inductance on the file's frequency grid)
gives the spectral derivative of a cosine profile, V_j = a m sin(2 pi m j/n)
for I_j = cos(2 pi m j/n); the smoothing leaves a constant profile as it is,
keeps the sum, and spreads a single spike into the normalised Gaussian
window exp(-alpha^2 x^2 / 2), x = (j - 25)/25, alpha = 10; a quadrupolar
table Z_Q = -i R gives V_Q = +R I and px += V_Q x / (E beta^2 beta_x); a
dipolar table Z_D = -i R gives V_D = +R I_x, I_x the smoothed x-weighted
profile formed here from the particles, and px += V_D / (E beta^2 beta_x);
at a common offset x0 of every particle the dipolar kick equals the
quadrupolar one for the same table."""
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


CLIGHT = 299792458.0
ECHARGE = 1.602017733e-19  # include/Global.h (the oracle's constant)


def kick(b, nimp, fmax, planes, ztab, bx=5.0, by=3.0, q=1e-15, rbeta=0.99999, t0=2e-6, h=1000, E=6e9):
    """orc_bbi_kick on one bunch; ztab {plane: (re, im)} in flag order
    (0 Z, 1 Dx, 2 Dy, 3 Qx, 4 Qy).  Returns (wake [5][n], prof [3][n])."""
    import ctypes
    n = 2 * nimp - 1
    zre, zim = np.zeros((5, nimp)), np.zeros((5, nimp))
    for k, (re, im) in ztab.items():
        zre[k], zim[k] = re, im
    pl = np.zeros(5, np.int32)
    pl[list(planes)] = 1
    wake, prof = np.zeros(5 * n), np.zeros(3 * n)
    O.lib().orc_bbi_kick(ctypes.byref(b.s), nimp, fmax, pl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                         O.p(zre.ravel()), O.p(zim.ravel()), bx, by, q, rbeta, t0, h, E, O.p(wake), O.p(prof))
    return wake.reshape(5, n), prof.reshape(3, n)


def bunch_xyz(seed, n0=3000, x0=0.0, y0=0.0, sx=1e-4, sy=5e-5):
    rng = np.random.default_rng(seed)
    a = {c: np.zeros(n0) for c in O.COORDS}
    a["z"] = rng.normal(0, 4e-3, n0)
    a["x"] = x0 + (rng.normal(0, sx, n0) if sx else np.zeros(n0))
    a["y"] = y0 + (rng.normal(0, sy, n0) if sy else np.zeros(n0))
    return a, O.Bunch(a)


def grid(nimp, fmax, rbeta=0.99999, t0=2e-6, h=1000):
    n = 2 * nimp - 1
    dz = CLIGHT * rbeta / (2 * fmax)
    zmin = -dz * (nimp - 1)
    rf = t0 / h * CLIGHT
    return n, dz, zmin, int((-rf / 2 - zmin + dz / 2) / dz), int((rf / 2 - zmin + dz / 2) / dz)


def test_quadrupolar_plane_sign_and_kick():
    """One bunch through orc_bbi_kick with only the Qx plane on and
    Z_Qx = -i R: i Z_Qx = R, so V_Qx = +R I (the reference's + sign,
    src/MPBunch.cpp:2218-2233), and px changes by V_Qx(bin) x/(E beta^2 beta_x)."""
    a, b = bunch_xyz(2, sy=0)
    nimp, fmax, R = 261, 1.0e11, 500.0
    wake, prof = kick(b, nimp, fmax, (3,), {3: (np.zeros(nimp), np.full(nimp, -R))})
    assert all(np.all(wake[q] == 0) for q in (0, 1, 2, 4))
    np.testing.assert_allclose(wake[3], R * prof[0], rtol=0, atol=1e-12 * R * prof[0].max())
    n, dz, zmin, _, _ = grid(nimp, fmax)
    k = ((a["z"] - zmin + dz / 2) / dz).astype(int)
    np.testing.assert_allclose(b["px"], wake[3][k] / 6e9 / 0.99999 ** 2 * a["x"] / 5.0, rtol=1e-13, atol=0)
    assert np.all(b["dp"] == 0) and np.all(b["py"] == 0)


def test_dipolar_plane_sign_profile_and_kick():
    """Only the Dx plane on, Z_Dx = -i R: V_Dx = +R I_x with I_x the
    smoothed x-weighted profile (charge times x summed per bin,
    src/MPBunch.cpp:2252-2281), and px changes by V_Dx(bin)/(E beta^2
    beta_x), independently of the particle's own x.  The raw weighted
    profile is formed here from the particles, the smoothing is the one
    test_smoothing pins."""
    a, b = bunch_xyz(4, x0=3e-5)
    nimp, fmax, R = 261, 1.0e11, 700.0
    q, E, rbeta, bx = 2e-15, 6e9, 0.99999, 5.0
    wake, prof = kick(b, nimp, fmax, (1,), {1: (np.zeros(nimp), np.full(nimp, -R))}, bx=bx, q=q)
    n, dz, zmin, i0, i1 = grid(nimp, fmax)
    k = ((a["z"] - zmin + dz / 2) / dz).astype(int)
    raw = np.zeros(n)
    np.add.at(raw, k, q * ECHARGE / dz * CLIGHT * a["x"])
    bucket = np.ascontiguousarray(raw[i0:i1])
    O.lib().orc_bbi_smooth(O.p(bucket), i1 - i0)
    exp = raw.copy()
    exp[i0:i1] = bucket
    np.testing.assert_allclose(prof[1], exp, rtol=0, atol=1e-12 * np.abs(exp).max())
    assert all(np.all(wake[p] == 0) for p in (0, 2, 3, 4))
    np.testing.assert_allclose(wake[1], R * prof[1], rtol=0, atol=1e-12 * R * np.abs(prof[1]).max())
    np.testing.assert_allclose(b["px"], wake[1][k] / E / rbeta ** 2 / bx, rtol=1e-13, atol=0)
    assert np.all(b["py"] == 0) and np.all(b["dp"] == 0)


def test_dipolar_equals_quadrupolar_at_a_common_offset():
    """Every particle at the same (x0, y0): the x-weighted profile is x0
    times the charge profile, so with the same impedance table the dipolar
    kick V_Dx/(E beta^2 beta_x) equals the quadrupolar kick
    V_Qx x0/(E beta^2 beta_x) (and likewise in y)."""
    nimp, fmax = 261, 1.0e11
    f = np.arange(nimp) * fmax / (nimp - 1)
    zt = np.zeros(nimp, complex)
    zt[1:] = (1.5e10 / f[1:]) * 1e3 / (1 + 1j * (f[1:] / 1.5e10 - 1.5e10 / f[1:]))
    tab = (zt.real.copy(), zt.imag.copy())
    x0, y0 = 2e-4, -1.5e-4
    a, bq = bunch_xyz(6, x0=x0, y0=y0, sx=0, sy=0)
    _, bd = bunch_xyz(6, x0=x0, y0=y0, sx=0, sy=0)
    wq, _ = kick(bq, nimp, fmax, (3, 4), {3: tab, 4: tab})
    wd, _ = kick(bd, nimp, fmax, (1, 2), {1: tab, 2: tab})
    np.testing.assert_allclose(wd[1], x0 * wq[3], rtol=0, atol=1e-12 * np.abs(x0 * wq[3]).max())
    np.testing.assert_allclose(wd[2], y0 * wq[4], rtol=0, atol=1e-12 * np.abs(y0 * wq[4]).max())
    for c in ("px", "py"):
        scale = np.abs(bq[c]).max()
        assert scale > 0
        np.testing.assert_allclose(bd[c], bq[c], rtol=0, atol=1e-12 * scale)


# ---- the time-domain method (MPBunch::BBImpBunchInteractionTD) ----

T0, H = 2e-6, 1000  # bucket L = T0/h c = 0.599584916 m


def td_kick(b, w, dz=1e-3, bx=5.0, by=3.0, q=1e-15, rbeta=0.99999, E=6e9):
    """orc_bbi_td_kick on one bunch; w {plane: table} in flag order, all
    tables of one length.  Returns (V [5][n], prof [3][n])."""
    import ctypes
    nw = len(next(iter(w.values())))
    tabs = np.zeros((5, nw))
    pl = np.zeros(5, np.int32)
    for k, t in w.items():
        tabs[k] = t
        pl[k] = 1
    pos = np.arange(nw) * dz
    L = O.lib()
    n = L.orc_bbi_td_bins(T0, H, dz)
    V, prof = np.zeros(5 * n), np.zeros(3 * n)
    L.orc_bbi_td_kick(ctypes.byref(b.s), nw, O.p(pos), O.p(tabs.ravel()), pl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                      bx, by, q, T0, H, rbeta, E, O.p(V), O.p(prof))
    return V.reshape(5, n), prof.reshape(3, n)


def filter_kernel():
    x = (np.arange(51) - 25) / 25
    k = np.exp(-0.5 * 100 * x ** 2)
    return k / k.sum()


def test_td_grid_and_single_particle():
    """(int)(L/dz) bins from -L/2; one particle at z = 12.8 mm lands in bin
    (int)((0.0128 + 0.299792458)/1e-3) = (int)312.59 = 312 (a centred grid
    would give 313), its charge spread by the
    filter kernel; a wake of one nonzero row w_0 gives V = -w_0 rho and
    delta += V(312)/(E beta^2), nothing else moves."""
    assert O.lib().orc_bbi_td_bins(T0, H, 1e-3) == 599
    assert O.lib().orc_bbi_td_bins(T0, H, 2.5e-4) == 2398
    a = {c: np.zeros(1) for c in O.COORDS}
    a["z"][0], a["x"][0] = 0.0128, 1e-3
    b = O.Bunch(a)
    q, w0 = 3e-15, 2.0e12
    w = np.zeros(10)
    w[0] = w0
    V, prof = td_kick(b, {0: w}, q=q)
    assert V.shape == (5, 599)
    rho = np.zeros(599)
    rho[312 - 25:312 + 26] = q * ECHARGE * filter_kernel()
    np.testing.assert_allclose(prof[0], rho, rtol=1e-13, atol=1e-40)
    np.testing.assert_allclose(V[0], -w0 * rho, rtol=1e-13, atol=0)
    assert np.all(V[1:] == 0)
    np.testing.assert_allclose(b["dp"], V[0][312] / 6e9 / 0.99999 ** 2, rtol=1e-15)
    assert np.all(b["px"] == 0) and np.all(b["py"] == 0)


def test_td_sum_runs_over_the_bins_ahead():
    """A wake whose only nonzero row is at lag d gives V_i = -w_d rho_{i+d}
    (the sum over j >= i, src/MPBunch.cpp:2100-2104): a shift towards
    lower bins; rows past the table's end add nothing."""
    a, b = bunch_xyz(11)
    d, wd = 7, 1.5e12
    w = np.zeros(20)
    w[d] = wd
    V, prof = td_kick(b, {0: w})
    rho = prof[0]
    exp = np.zeros_like(rho)
    exp[:-d] = -wd * rho[d:]
    np.testing.assert_allclose(V[0], exp, rtol=1e-14, atol=0)
    # a constant wake of 5 rows: V_i = -w (rho_i + ... + rho_{i+4})
    a, b = bunch_xyz(12)
    V, prof = td_kick(b, {0: np.full(5, 1.0e12)})
    rho = prof[0]
    exp = -1.0e12 * np.array([rho[i:i + 5].sum() for i in range(len(rho))])
    np.testing.assert_allclose(V[0], exp, rtol=1e-13, atol=1e-13 * np.abs(exp).max())


def test_td_constant_wake_sees_the_whole_charge():
    """A constant wake longer than the grid: V_0 = -w Q, Q the bunch's
    charge (the smoothing keeps the sum), V rises to 0 behind the tail."""
    a, b = bunch_xyz(13, n0=5000)
    q, wc = 2e-15, 3.0e12
    V, prof = td_kick(b, {0: np.full(700, wc)}, q=q)
    Q = 5000 * q * ECHARGE
    np.testing.assert_allclose(V[0][0], -wc * Q, rtol=1e-13)
    assert np.all(np.diff(V[0]) >= -1e-12 * wc * Q)
    assert V[0][-1] == 0


def lag_sums(w, f):
    """-sum_{j >= i, j - i < len(w)} w[j - i] f_j for every i, by brute force."""
    n = len(f)
    return np.array([-sum(w[j - i] * f[j] for j in range(i, min(n, i + len(w)))) for i in range(n)])


def test_td_dipolar_and_quadrupolar_planes():
    """<x>_j is the mean x of bin j's particles (brute force here), zero in
    empty bins; V_Dx = -sum w rho_j <x>_j over the bins ahead, so at one
    common offset x0 it equals x0 times the quadrupolar sum over the
    occupied bins only (the smoothed rho of an empty bin meets <x> = 0);
    px += V_Dx(bin)/(E beta^2 beta_x) + V_Qx(bin)/(E beta^2 beta_x) x."""
    a, b = bunch_xyz(14, x0=2e-4, y0=-1e-4)
    w = 1e12 * np.exp(-np.arange(60) / 15.0) * np.sin(np.arange(60) / 4.0 + 0.3)
    V, prof = td_kick(b, {1: w, 2: 0.5 * w})
    k = ((a["z"] + 0.299792458) / 1e-3).astype(int)
    cnt = np.bincount(k, minlength=599)
    mx = np.bincount(k, weights=a["x"], minlength=599)
    my = np.bincount(k, weights=a["y"], minlength=599)
    occ = cnt > 0
    np.testing.assert_allclose(prof[1][occ], mx[occ] / cnt[occ], rtol=1e-12)
    np.testing.assert_allclose(prof[2][occ], my[occ] / cnt[occ], rtol=1e-12)
    assert np.all(prof[1][~occ] == 0) and np.all(prof[2][~occ] == 0)
    exp = lag_sums(w, prof[0] * prof[1])
    np.testing.assert_allclose(V[1], exp, rtol=0, atol=1e-13 * np.abs(exp).max())
    assert np.all(b["dp"] == 0)
    x0, y0 = 3e-4, -2e-4
    a, bq = bunch_xyz(15, x0=x0, y0=y0, sx=0, sy=0)
    k = ((a["z"] + 0.299792458) / 1e-3).astype(int)
    occ = np.bincount(k, minlength=599) > 0
    V, pd = td_kick(bq, {1: w, 2: w, 3: 2 * w, 4: 3 * w}, bx=5.0, by=3.0)
    np.testing.assert_allclose(pd[1][occ], x0, rtol=1e-14)
    np.testing.assert_allclose(pd[2][occ], y0, rtol=1e-14)
    for plane, c0 in ((1, x0), (2, y0)):
        exp = c0 * lag_sums(w, pd[0] * occ)
        np.testing.assert_allclose(V[plane], exp, rtol=0, atol=1e-13 * np.abs(exp).max())
    E, rb = 6e9, 0.99999
    np.testing.assert_allclose(bq["px"], V[1][k] / E / rb ** 2 / 5.0 + V[3][k] / E / rb ** 2 / 5.0 * x0, rtol=1e-14)
    np.testing.assert_allclose(bq["py"], V[2][k] / E / rb ** 2 / 3.0 + V[4][k] / E / rb ** 2 / 3.0 * y0, rtol=1e-14)


def test_td_lost_and_off_grid_particles():
    """Lost particles and particles beyond the bucket's end (bin >= n) are
    neither binned nor kicked."""
    a, _ = bunch_xyz(16, n0=400)
    a["z"][:3] = [0.3005, 0.31, -0.35]
    lost = np.zeros(400, np.int32)
    lost[3:10] = 1
    b = O.Bunch(a, lost=lost)
    w = np.zeros(8)
    w[0] = 1e12
    V, prof = td_kick(b, {0: w}, q=1e-15)
    np.testing.assert_allclose(prof[0].sum(), 390 * 1e-15 * ECHARGE, rtol=1e-13)
    assert np.all(b["dp"][:10] == 0) and np.all(b["dp"][10:] != 0)


def wake_rows(z, cols):
    return np.ascontiguousarray(np.column_stack([z] + list(cols)), dtype=float)


def td_wake(rows):
    n = len(rows)
    out = np.zeros(6 * n)
    cen = O.lib().orc_bbi_td_wake(n, O.p(rows.ravel()), O.p(out))
    return cen, out.reshape(6, n)


def test_td_wake_file_preparation():
    """GetQuasiWakePoten's reading (src/MPBeam.cpp:748-812): columns times
    -1, positions from 0, the longitudinal wake folded about the z = 0 row;
    the transverse ones only negated.  Hand values: z = -2..3, wz = 1..6
    gives cen 2 and wz' = (-3, -4 - 2, -5, -6, 0, 0)."""
    z = np.arange(-2.0, 4.0)
    cols = [np.arange(1.0, 7.0), 10 + np.arange(6.0), 20 + np.arange(6.0), 30 + np.arange(6.0), 40 + np.arange(6.0)]
    cen, out = td_wake(wake_rows(z, cols))
    assert cen == 2
    np.testing.assert_array_equal(out[0], [0, 1, 2, 3, 4, 5])
    np.testing.assert_array_equal(out[1], [-3, -6, -5, -6, 0, 0])
    for c in range(2, 6):
        np.testing.assert_array_equal(out[c], -cols[c - 1])
    # a causal wake (zero ahead of z = 0) is only shifted to start at z = 0;
    # an even one is doubled behind z = 0
    z = np.arange(-4.0, 9.0) * 1e-4
    causal = np.where(z >= 0, np.cos(z * 1e4), 0.0)
    _, out = td_wake(wake_rows(z, [causal] + [np.zeros(13)] * 4))
    np.testing.assert_array_equal(out[1][:9], -causal[4:])
    assert np.all(out[1][9:] == 0)
    even = np.exp(-(z * 1e4) ** 2)
    _, out = td_wake(wake_rows(z, [even] + [np.zeros(13)] * 4))
    np.testing.assert_allclose(out[1][1:4], -2 * even[5:8], rtol=1e-15)
    assert out[1][0] == -even[4]
    # no row at z = 0: cen 0, the table as read (negated)
    z = np.arange(6.0) + 0.5
    cen, out = td_wake(wake_rows(z, cols))
    assert cen == 0
    np.testing.assert_array_equal(out[1], -cols[0])
