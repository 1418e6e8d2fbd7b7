"""Pins the oracle's tracking functions with closed forms, invariants and
brute force (CPU only).  Each check is chosen so that a plausible slip (a
dropped term, a wrong sign or index, a transposed operand) fails it."""
import ctypes
import math

import numpy as np
import pytest

import oracle as O

L = O.lib()
C = 299792458.0
E_CHARGE = 1.602017733e-19  # include/Global.h:28


def gaussian(n, seed, sx=20e-6, sy=4e-6, sz=3e-3, sd=1e-3, beta=10.0):
    rng = np.random.default_rng(seed)
    return dict(x=rng.standard_normal(n) * sx, px=rng.standard_normal(n) * sx / beta,
                y=rng.standard_normal(n) * sy, py=rng.standard_normal(n) * sy / beta,
                z=rng.standard_normal(n) * sz, dp=rng.standard_normal(n) * sd)


def twiss(**kw):
    t = dict(beta_x=10.0, alpha_x=0.0, eta_x=0.0, etap_x=0.0, beta_y=10.0, alpha_y=0.0, eta_y=0.0,
             etap_y=0.0, beta_z=0.5)
    t.update(kw)
    return O.OrcTwiss(**t)


def make_map(t1, t2, nux, nuy, chrom=(0, 0), adtx=(0,) * 5, adty=(0,) * 5):
    b1h1, h1b1, h2b2 = O.section_matrices(t1, t2)
    m = O.OrcMap(O.p(b1h1), O.p(h2b2), t1.eta_x, t1.etap_x, t1.eta_y, t1.etap_y, t1.alpha_x, t1.beta_x,
                 t1.alpha_y, t1.beta_y, 2 * math.pi * nux, 2 * math.pi * nuy, nux, nuy, chrom[0], chrom[1])
    m.adtx[:] = adtx
    m.adty[:] = adty
    m._keep = (b1h1, h1b1, h2b2)
    return m


def test_dgemm_matches_numpy_and_skips_zero_terms():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((6, 6)); A[A < -0.5] = 0.0
    B = rng.standard_normal((6, 4))
    Cm = np.zeros(24)
    L.orc_dgemm(6, 4, 6, O.p(np.ascontiguousarray(A).ravel()), O.p(np.ascontiguousarray(B).ravel()), O.p(Cm))
    assert np.allclose(Cm.reshape(6, 4), A @ B, rtol=1e-15, atol=1e-15)
    # a zero entry times an infinite operand is skipped, as gslcblas does
    A2 = np.eye(2).ravel(); B2 = np.array([1.0, np.inf])
    C2 = np.zeros(2)
    L.orc_dgemm(2, 1, 2, O.p(A2), O.p(B2), O.p(C2))
    assert C2[0] == 1.0 and C2[1] == np.inf


def test_matrices_are_inverse_pairs():
    t1 = twiss(alpha_x=0.7, eta_x=0.05, etap_x=-0.01, alpha_y=-0.4, eta_y=0.003, etap_y=0.002, beta_x=8, beta_y=12)
    b1h1, h1b1, h2b2 = [m.reshape(6, 6) for m in O.section_matrices(t1, t1)]
    assert np.allclose(b1h1 @ h1b1, np.eye(6), atol=1e-15)
    assert np.array_equal(h1b1, h2b2)


def test_pure_rotation_returns_after_period():
    """eta = 0, no chromaticity/ADTS, nu = 1/4 and 1/5: the particle comes
    back after 4 (x) and 5 (y) turns, and the normalized amplitude is kept."""
    t = twiss(alpha_x=0.3, alpha_y=-0.2)
    m = make_map(t, t, 0.25, 0.2)
    parts = gaussian(200, 3)
    b = O.Bunch(parts)
    J0x = (parts["x"] ** 2 + (0.3 * parts["x"] + 10 * parts["px"]) ** 2) / 10
    for _ in range(20):
        L.orc_section_map(ctypes.byref(b.s), ctypes.byref(m))
    for c in ("x", "px", "y", "py"):
        assert np.allclose(b[c], parts[c], rtol=0, atol=1e-13 * np.max(np.abs(parts[c])))
    Jx = (b["x"] ** 2 + (0.3 * b["x"] + 10 * b["px"]) ** 2) / 10
    assert np.allclose(Jx, J0x, rtol=1e-12)
    # z and dp pass through B^-1 B: unchanged up to rounding
    assert np.allclose(b["z"], parts["z"], rtol=1e-14) and np.allclose(b["dp"], parts["dp"], rtol=1e-14)


def test_one_turn_tune_and_symplecticity():
    """One-turn matrix from finite differences: det 1, symplectic, traces
    2 cos(2 pi nu)."""
    t = twiss(alpha_x=0.5, beta_x=7.0, alpha_y=-0.2, beta_y=13.0, eta_x=0.1, etap_x=0.02)
    nux, nuy = 0.18, 0.27
    m = make_map(t, t, nux, nuy)
    eps = 1e-7
    M = np.zeros((6, 6))
    for k in range(6):
        base = {c: np.zeros(2) for c in ("x", "px", "y", "py", "z", "dp")}
        key = ("x", "px", "y", "py", "z", "dp")[k]
        base[key] = np.array([eps, -eps])
        b = O.Bunch(base)
        L.orc_section_map(ctypes.byref(b.s), ctypes.byref(m))
        for r, c in enumerate(("x", "px", "y", "py", "z", "dp")):
            M[r, k] = (b[c][0] - b[c][1]) / (2 * eps)
    J = np.zeros((6, 6))
    for k in (0, 2, 4):
        J[k, k + 1], J[k + 1, k] = 1, -1
    assert np.allclose(M.T @ J @ M, J, atol=1e-9)
    assert np.linalg.det(M) == pytest.approx(1.0, abs=1e-9)
    assert np.trace(M[0:2, 0:2]) == pytest.approx(2 * math.cos(2 * math.pi * nux), abs=1e-7)
    assert np.trace(M[2:4, 2:4]) == pytest.approx(2 * math.cos(2 * math.pi * nuy), abs=1e-9)


def test_chromatic_tune_shift():
    """A particle with momentum offset d advances by 2*pi*(nu + xi*d)."""
    t = twiss()
    nux, xi, d = 0.18, 2.0, 1e-3
    m = make_map(t, t, nux, 0.27, chrom=(xi, 0))
    b = O.Bunch(dict(x=np.array([1e-5]), px=np.array([0.0]), y=np.array([0.0]), py=np.array([0.0]),
                     z=np.array([0.0]), dp=np.array([d])))
    L.orc_section_map(ctypes.byref(b.s), ctypes.byref(m))
    phi = math.atan2(-b["px"][0] * 10, b["x"][0])  # x' = -sin(phi) x / beta
    assert (phi % (2 * math.pi)) == pytest.approx((2 * math.pi * (nux + xi * d)) % (2 * math.pi), abs=1e-12)


def test_rf_drift_u0_closed_forms():
    n = 50
    parts = gaussian(n, 5)
    b = O.Bunch(parts)
    vr, vi, h, frf, beta, E = 1.2e6, 3.1e7, 1, 5e8, 0.9999999, 6e9
    L.orc_rf_kick(ctypes.byref(b.s), vr, vi, h, frf, beta, E)
    th = -parts["z"] / C / beta * 2 * math.pi * h * frf
    exp = parts["dp"] + (vr * np.cos(th) - vi * np.sin(th)) / E / beta ** 2
    assert np.allclose(b["dp"], exp, rtol=1e-14, atol=0)
    # at z = 0 the kick is Vr/(E beta^2): the single-cavity phase makes it U0
    b0 = O.Bunch(dict(x=[0.0], px=[0.0], y=[0.0], py=[0.0], z=[0.0], dp=[0.0]))
    L.orc_rf_kick(ctypes.byref(b0.s), vr, vi, h, frf, beta, E)
    assert b0["dp"][0] == pytest.approx(vr / E / beta ** 2, rel=1e-15)
    a = np.array([3e-5, 1e-4, 2e-3])
    b2 = O.Bunch(parts)
    L.orc_drift(ctypes.byref(b2.s), 600.0, O.p(a))
    d = parts["dp"]
    assert np.allclose(b2["z"], parts["z"] - 600.0 * (a[0] * d + a[1] * d ** 2 + a[2] * d ** 3), rtol=1e-14)
    b3 = O.Bunch(parts)
    L.orc_energy_loss(ctypes.byref(b3.s), 1.2e6, E, beta)
    assert np.allclose(b3["dp"], parts["dp"] - 1.2e6 / E / beta ** 2, rtol=0, atol=1e-18)


def test_synrad_damping_is_geometric():
    """Without excitation each normalized coordinate decays by lambda (x, y)
    and lambda^2 (dp); z is untouched in normalized form."""
    t = twiss(alpha_x=0.3, eta_x=0.02, etap_x=0.001)
    b1h1, h1b1, _ = O.section_matrices(t, t)
    tau = np.array([100.0, 200.0, 50.0]); en = np.array([4e-11, 1.6e-12, 5e-7])
    parts = gaussian(30, 7)
    b = O.Bunch(parts)
    L.orc_synrad(ctypes.byref(b.s), O.p(b1h1), O.p(h1b1), O.p(tau), O.p(en), None)
    B = b1h1.reshape(6, 6)
    X0 = np.stack([parts[c] for c in ("x", "px", "y", "py", "z", "dp")])
    X1 = np.stack([b[c] for c in ("x", "px", "y", "py", "z", "dp")])
    N0, N1 = B @ X0, B @ X1
    lam = np.exp(-1.0 / tau)
    assert np.allclose(N1[0], lam[0] * N0[0], rtol=1e-12, atol=1e-20)
    assert np.allclose(N1[2], lam[1] * N0[2], rtol=1e-12, atol=1e-20)
    assert np.allclose(N1[4], N0[4], rtol=1e-12, atol=1e-20)
    assert np.allclose(N1[5], lam[2] ** 2 * N0[5], rtol=1e-12, atol=1e-20)
    # with excitation: the added normalized variance is (1 - lambda^2) eps
    n = 200_000
    b = O.Bunch(dict(x=np.zeros(n), px=np.zeros(n), y=np.zeros(n), py=np.zeros(n), z=np.zeros(n), dp=np.zeros(n)))
    noise = np.random.default_rng(1).standard_normal((5, n))
    L.orc_synrad(ctypes.byref(b.s), O.p(b1h1), O.p(h1b1), O.p(tau), O.p(en), O.p(noise.ravel().copy()))
    Nx = (B @ np.stack([b[c] for c in ("x", "px", "y", "py", "z", "dp")]))[0]
    assert np.var(Nx) == pytest.approx((1 - lam[0] ** 2) * en[0], rel=0.01)


def test_mark_lost_brute_force():
    n = 1000
    rng = np.random.default_rng(9)
    parts = dict(x=rng.uniform(-3e-3, 3e-3, n), px=np.zeros(n), y=rng.uniform(-2e-3, 2e-3, n), py=np.zeros(n),
                 z=rng.uniform(-0.4, 0.4, n), dp=np.zeros(n))
    lost0 = np.zeros(n, np.int32); lost0[:10] = 2
    b = O.Bunch(parts, lost0)
    t0, h, ax, ay = 2e-6, 1000, 2.5e-3, 1.5e-3
    cnt = ctypes.c_int64()
    tr = L.orc_mark_lost(ctypes.byref(b.s), t0, h, ax, ay, ctypes.byref(cnt))
    zl = np.abs(parts["z"]) > t0 * C / h / 2
    tl = (parts["x"] / ax) ** 2 + (parts["y"] / ay) ** 2 > 1
    exp = lost0.copy(); exp[zl] = 2; exp[tl] = 1
    assert np.array_equal(b.lost, exp)
    assert cnt.value == zl.sum() + tl.sum()  # double counting kept
    assert tr == (n - cnt.value) / n


def test_rms_two_pass_and_eigen():
    parts = gaussian(5000, 11)
    parts["x"] += 3e-6
    lost = np.zeros(5000, np.int32); lost[::17] = 1
    b = O.Bunch(parts, lost)
    m = O.OrcMoments()
    L.orc_rms(ctypes.byref(b.s), 10.0, 10.0, 0.0, 0.0, 0.0, 0.0, m)
    a = lost == 0
    x, px = parts["x"][a], parts["px"][a]
    assert m.n_alive == a.sum()
    assert m.mean[0] == pytest.approx(x.mean(), rel=1e-13)
    xx, pp, xp = np.var(x), np.var(px), np.mean((x - x.mean()) * (px - px.mean()))
    assert m.emit[0] == pytest.approx(math.sqrt(xx * pp - xp * xp), rel=1e-11)
    assert m.rms_x == pytest.approx(math.sqrt(m.emit[0] * 10.0), rel=1e-14)
    assert m.sigma_z == pytest.approx(np.std(parts["z"][a]), rel=1e-12)
    # eigen emittances of uncoupled beams equal the projected (uncentred,
    # all particles: the int-valued accumulate leaves the means at 0)
    X = np.stack([parts["x"], parts["px"], parts["y"], parts["py"]])
    S = X @ X.T / X.shape[1]
    ex = math.sqrt(S[0, 0] * S[1, 1] - S[0, 1] ** 2)
    assert m.eigen_x == pytest.approx(ex, rel=2e-2)
    assert m.emit4d ** 2 == pytest.approx(math.sqrt(np.linalg.det(S)), rel=1e-9)


def test_det_lu():
    rng = np.random.default_rng(2)
    for _ in range(20):
        A = rng.standard_normal((4, 4))
        assert L.orc_det(4, O.p(A.ravel().copy())) == pytest.approx(np.linalg.det(A), rel=1e-12)
    assert L.orc_det(4, O.p(np.zeros(16))) == 0.0
    P = np.eye(4)[[1, 0, 2, 3]]
    assert L.orc_det(4, O.p(P.ravel().copy())) == -1.0


def test_wake_binning_brute_force_and_lag_table():
    n = 3000
    parts = gaussian(n, 13)
    lost = np.zeros(n, np.int32); lost[::11] = 2
    b = O.Bunch(parts, lost)
    nb = 50
    rs, q, om = np.array([1e4]), np.array([1.0]), np.array([2 * math.pi * 1e10])
    poten = np.zeros(3 * nb); counts = np.zeros(nb, np.int32); edges = np.zeros(2)
    L.orc_sr_wake(ctypes.byref(b.s), nb, 0, 3e-3, 0, 0, 6e4, 6e9, 1, O.p(rs), O.p(q), O.p(om), O.p(poten),
                  counts.ctypes.data_as(O.PI32), O.p(edges), None)
    a = lost == 0
    zmin, zmax = parts["z"][a].min() - 1e-6, parts["z"][a].max() + 1e-6
    assert edges[0] == zmin - 2 * 3e-3
    idx = ((parts["z"] - edges[0]) / edges[1]).astype(int)
    ref = np.bincount(idx[(idx >= 0) & (idx < nb)], minlength=nb)
    assert np.array_equal(counts, ref)
    assert counts.sum() == n  # every particle, lost ones included, is binned
    # kick of survivors only
    kick = poten[2 * nb:]
    assert np.array_equal(b["dp"][~a], parts["dp"][~a])
    assert np.array_equal(b["dp"][a], parts["dp"][a] + kick[idx[a]])
    # the lag table equals per-pair evaluation exactly (tau = (i-j)*dt)
    tab = np.zeros(2 * nb - 1)
    L.orc_wake_table(nb, edges[1] / C, 1, O.p(rs), O.p(q), O.p(om), O.p(tab))
    w = np.zeros(3)
    for (i, j) in ((3, 7), (40, 2), (10, 10)):
        L.orc_bbr_wake((i - j) * (edges[1] / C), 1, O.p(rs), O.p(q), O.p(om), None, None, None, None, None, None,
                       O.p(w))
        assert w[2] == tab[i - j + nb - 1]


def test_bbr_wake_against_quadrature():
    """The Riemann sum approximates Chao Eq. 3.14 for a 1 mm Gaussian bunch
    where the integrand lies inside the truncated range (leading side)."""
    from scipy.integrate import quad
    rs, Q, fr = 1e4, 1.0, 1e10
    w = np.zeros(3)
    om = 2 * math.pi * fr
    sig = 1e-3
    for u in (0.5, 1.0, 2.0):
        L.orc_bbr_wake(u * sig / C, 1, O.p(np.array([rs])), O.p(np.array([Q])), O.p(np.array([om])), None, None,
                       None, None, None, None, O.p(w))
        v = om * sig / C
        sq = math.sqrt(4 * Q * Q - 1)
        f = lambda z: math.exp(-0.5 * (u + 2 * Q / v * z) ** 2 - z) * (math.cos(z * sq) - math.sin(z * sq) / sq)
        zmax = 3 * v / (2 * Q)
        integ = quad(f, 0, zmax, limit=200, epsabs=0, epsrel=1e-13)[0]
        # a left Riemann sum of 1000 steps = integral + dz/2 (f(0) - f(zmax)) + O(dz^2)
        dz = zmax / 1000
        left = integ + 0.5 * dz * (f(0.0) - f(zmax))
        exact = math.sqrt(2 / math.pi) * rs * C / sig * left
        assert w[2] == pytest.approx(exact, rel=1e-5)


def test_sr_wake_kick_is_the_convolution():
    """The bin kick is -e q / E times the discrete convolution of the bin
    counts with the lag table (src/MPBunch.cpp:1983-1998), computed here by
    numpy, plus the previous call's potential carried over (the reference
    never zeroes srWakePoten, src/MPBunch.cpp:1894-1897); the table's sign
    makes the bunch lose energy."""
    n, nb = 4000, 60
    parts = gaussian(n, 23)
    rs, q, om = np.array([1e4]), np.array([1.0]), np.array([2 * math.pi * 1e10])
    qmp, energy = 6.24e4, 6.0e9
    grid = (-0.015, 0.015)
    dz = (grid[1] - grid[0]) / nb
    tab = np.zeros(2 * nb - 1)
    L.orc_wake_table(nb, dz / C, 1, O.p(rs), O.p(q), O.p(om), O.p(tab))
    for carry in (0.0, 1.0):
        b = O.Bunch(parts)
        poten = np.zeros(3 * nb)
        poten[2 * nb:] = carry * np.linspace(-1e-3, 2e-3, nb)
        prev = poten[2 * nb:].copy()
        counts = np.zeros(nb, np.int32); edges = np.zeros(2)
        L.orc_sr_wake(ctypes.byref(b.s), nb, 1, 0.0, grid[0], grid[1], qmp, energy, 1, O.p(rs), O.p(q), O.p(om),
                      O.p(poten), counts.ctypes.data_as(O.PI32), O.p(edges), O.p(tab))
        conv = np.convolve(counts.astype(float), tab)[nb - 1:2 * nb - 1]
        expect = (prev + conv) * (-E_CHARGE * qmp / energy)
        got = poten[2 * nb:]
        assert np.allclose(got, expect, rtol=1e-12, atol=1e-12 * np.max(np.abs(expect)))
        if carry == 0.0:
            # energy loss: the count-weighted mean kick is negative
            assert np.dot(counts, got) < 0


def test_wake_pairs_equal_lag_table():
    """The reference's per-pair evaluation of the wake sum (the CPU
    baseline's cost model) gives the lag table's sums bit for bit."""
    nb, dt = 40, 2.5e-12
    rs, q, om = np.array([1e4, 3e3]), np.array([1.0, 2.0]), np.array([2 * math.pi * 1e10, 2 * math.pi * 4e9])
    counts = np.random.default_rng(3).integers(0, 500, nb).astype(np.int32)
    tab = np.zeros(2 * nb - 1)
    L.orc_wake_table(nb, dt, 2, O.p(rs), O.p(q), O.p(om), O.p(tab))
    exp = np.zeros(nb)
    for i in range(nb):
        for j in range(nb):
            exp[i] += tab[i - j + nb - 1] * counts[j]
    got = np.zeros(nb)
    L.orc_wake_pairs(nb, dt, 2, O.p(rs), O.p(q), O.p(om), counts.ctypes.data_as(O.PI32), O.p(got))
    assert np.array_equal(got, exp)


def test_bbr_wake_trailing_side_quadrature():
    """u < 0 (a receiver bin behind its source): the Riemann sum against
    the integral over the same truncated range [0, 3 v / (2 Q)]; the
    Gaussian's peak moves to z* = -u v / (2 Q), up to the cut-off at u = -3
    (the range the reference sums, not Chao's untruncated integral)."""
    from scipy.integrate import quad
    rs, Q, fr = 1e4, 1.0, 1e10
    w = np.zeros(3)
    om = 2 * math.pi * fr
    sig = 1e-3
    v = om * sig / C
    sq = math.sqrt(4 * Q * Q - 1)
    zmax = 3 * v / (2 * Q)
    vals = []
    for u in (-3.0, -1.5, -0.5):
        L.orc_bbr_wake(u * sig / C, 1, O.p(np.array([rs])), O.p(np.array([Q])), O.p(np.array([om])), None, None,
                       None, None, None, None, O.p(w))
        f = lambda z: math.exp(-0.5 * (u + 2 * Q / v * z) ** 2 - z) * (math.cos(z * sq) - math.sin(z * sq) / sq)
        integ = quad(f, 0, zmax, limit=200, epsabs=0, epsrel=1e-13)[0]
        dz = zmax / 1000
        left = integ + 0.5 * dz * (f(0.0) - f(zmax))
        vals.append((w[2], math.sqrt(2 / math.pi) * rs * C / sig * left))
    scale = max(abs(e) for _, e in vals)
    for got, exact in vals:
        assert abs(got - exact) <= 1e-5 * scale
    # the truncation: at u = -3 the untruncated integral is far larger
    u = -3.0
    f = lambda z: math.exp(-0.5 * (u + 2 * Q / v * z) ** 2 - z) * (math.cos(z * sq) - math.sin(z * sq) / sq)
    full = quad(f, 0, 40 * zmax, limit=400)[0]
    assert abs(full) > 1.2 * abs(vals[0][1] / (math.sqrt(2 / math.pi) * rs * C / sig))


def test_philox_known_answers():
    """Random123 philox4x32-10 known-answer vectors."""
    cases = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
             ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
             ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
              (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, exp in cases:
        c = (ctypes.c_uint32 * 4)(*ctr); k = (ctypes.c_uint32 * 2)(*key); o = (ctypes.c_uint32 * 4)()
        L.orc_philox4x32_10(c, k, o)
        assert tuple(o) == exp


def test_qe_normals_are_standard_normal():
    z = O.qe_fill(2, 0, 0, 400_000)
    for r in z:
        assert abs(r.mean()) < 0.006
        assert r.var() == pytest.approx(1.0, abs=0.008)
        assert abs(np.mean(r ** 3)) < 0.02
        assert np.mean(r ** 4) == pytest.approx(3.0, abs=0.05)
    # independent components and turns
    assert abs(np.corrcoef(z[0], z[1])[0, 1]) < 0.01
    z2 = O.qe_fill(2, 1, 0, 400_000)
    assert abs(np.corrcoef(z[0], z2[0])[0, 1]) < 0.01
    # float Box-Muller against libm for the same uniforms
    f0, f1 = ctypes.c_float(), ctypes.c_float()
    for a, b in ((0x12345678, 0x9abcdef0), (0xffffff00, 0x0), (0x100, 0x80000000)):
        L.orc_box_muller(a, b, f0, f1)
        u1 = ((a >> 8) + 1) * 2.0 ** -24
        u2 = (b >> 8) * 2.0 ** -24
        r = math.sqrt(-2 * math.log(u1))
        assert f0.value == pytest.approx(r * math.cos(2 * math.pi * u2), abs=3e-6)
        assert f1.value == pytest.approx(r * math.sin(2 * math.pi * u2), abs=3e-6)


def test_qe_transform_accuracy():
    """The table-driven Box-Muller pair of the excitation normals against
    libm in double: every table interval, both ends of the radius range
    (including the fourth-order branch next to u1 = 1) and random points;
    the tables against their defining values."""
    L.orc_box_muller_u21.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float),
                                     ctypes.POINTER(ctypes.c_float)]
    L.orc_normal_tables.argtypes = [ctypes.POINTER(ctypes.c_float)]
    tab = (ctypes.c_float * 1024)()
    L.orc_normal_tables(tab)
    t = np.array(tab[:], dtype=np.float64)
    j = np.arange(256)
    np.testing.assert_allclose(t[0:512:2], np.sin(2 * np.pi * j / 256), atol=6e-8)
    np.testing.assert_allclose(t[1:512:2], np.cos(2 * np.pi * j / 256), atol=6e-8)
    np.testing.assert_allclose(t[512::2], np.log1p(j / 256), rtol=6e-8, atol=1e-30)
    np.testing.assert_allclose(t[513::2], 256 / (256 + j), rtol=6e-8)
    rng = np.random.default_rng(1)
    a = np.concatenate([np.arange(0, 2 ** 21, 2053), np.arange(2 ** 21 - 9000, 2 ** 21), np.arange(300),
                        rng.integers(0, 2 ** 21, 3000)])
    b = np.concatenate([rng.integers(0, 2 ** 21, len(a) - 512), np.arange(256) << 13, (np.arange(256) << 13) + 8191])
    f0, f1 = ctypes.c_float(), ctypes.c_float()
    worst = 0.0
    for ai, bi in zip(a.tolist(), b.tolist()):
        L.orc_box_muller_u21(ai, bi, f0, f1)
        r = math.sqrt(-2 * math.log((ai + 1) * 2.0 ** -21))
        th = 2 * math.pi * bi * 2.0 ** -21
        worst = max(worst, abs(f0.value - r * math.cos(th)), abs(f1.value - r * math.sin(th)))
    # float results: a few units of the last place of values up to 5.4
    assert worst < 1.5e-6, worst


def test_fit_wlinear_exact_line():
    x = np.arange(10, dtype=float)
    y = 3.5 - 0.25 * x
    w = np.ones(10)
    c0, c1 = ctypes.c_double(), ctypes.c_double()
    L.orc_fit_wlinear(O.p(x), O.p(w), O.p(y), 10, c0, c1)
    assert c1.value == pytest.approx(-0.25, rel=1e-14)
    assert c0.value == pytest.approx(3.5, rel=1e-14)
    w2 = np.ones(10); w2[3] = 0.0; y2 = y.copy(); y2[3] = 100.0  # zero weight skipped
    L.orc_fit_wlinear(O.p(x), O.p(w2), O.p(y2), 10, c0, c1)
    assert c1.value == pytest.approx(-0.25, rel=1e-13)


def test_ion_generation_statistics():
    n = 20000
    gx, gy = np.zeros(n), np.zeros(n)
    L.orc_ion_generate(5, 3, 7, 0, n, 1e-6, -2e-6, 2e-5, 4e-6, O.p(gx), O.p(gy))
    u, v = (gx - 1e-6) / 2e-5, (gy + 2e-6) / 4e-6
    assert np.all(u * u + v * v <= 4.0)
    # truncated 2-D Gaussian at r = 2: <r^2> = 2 - 4 e^-2 / (1 - e^-2)
    e2 = math.exp(-2)
    assert np.mean(u * u + v * v) == pytest.approx(2 - 4 * e2 / (1 - e2), rel=0.02)
    assert abs(np.mean(u)) < 0.02 and abs(np.mean(v)) < 0.02
