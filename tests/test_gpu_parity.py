"""HIP path vs the CPU oracle on identical seeded macroparticles (MI355X).

Tolerances: coordinates 1e-12 relative to each coordinate's largest
magnitude (the only differences are libm last bits of sin/cos/exp/pow),
loss codes and bin counts bit-exact, moments 1e-9 relative to the rms size.
"""
import ctypes
import math

import numpy as np
import pytest

from cetasim_amd import capi, configs
from cetasim_amd.tracker import Tracker

import oracle as O

pytestmark = pytest.mark.gpu

COORDS = ("x", "px", "y", "py", "z", "dp")
BBR = configs.BASE["bbr"]


def lattice(**over):
    p = dict(configs.BASE)
    p.update(over)
    return configs.lattice(p), p


def tracker(ring, secs, parts, lost=None, p=configs.BASE):
    n = len(parts["x"])
    tr = Tracker(ring, secs)
    q, ne = configs.bunch_scalars([n], ring, p)
    tr.set_bunches([n], q, ne)
    tr.upload(parts, lost)
    return tr


def assert_coords(got, ob, tol=1e-12, mask=None):
    for c in COORDS:
        a, b = got[c], ob[c]
        if mask is not None:
            a, b = a[mask], b[mask]
        scale = max(np.max(np.abs(b)), 1e-300)
        err = np.max(np.abs(a - b)) / scale
        assert err <= tol, (c, err)


def random_losses(n, seed, frac=0.05):
    rng = np.random.default_rng(seed)
    lost = np.zeros(n, np.int32)
    idx = rng.choice(n, int(frac * n), replace=False)
    lost[idx] = rng.integers(1, 3, len(idx))
    return lost


def dispersive_section(ring, der):
    """A section with alpha, eta and eta' all non-zero (every matrix entry)."""
    lib = capi.load()
    t1 = capi.Twiss(beta_x=8.0, alpha_x=0.7, eta_x=0.05, etap_x=-0.01, beta_y=12.0, alpha_y=-0.4,
                    eta_y=0.003, etap_y=0.002, beta_z=der.beta_z)
    t2 = capi.Twiss(beta_x=11.0, alpha_x=-0.3, eta_x=0.02, etap_x=0.004, beta_y=6.0, alpha_y=0.2,
                    eta_y=-0.001, etap_y=0.001, beta_z=der.beta_z)
    sec = capi.Section()
    capi.check(lib.cts_build_section(t1, t2, 2 * math.pi * 0.18, 2 * math.pi * 0.27, sec))
    return sec


def test_qe_normals_bit_exact():
    (ring, secs, _), _ = lattice()
    tr = Tracker(ring, secs)
    for seed, turn, base in ((2, 0, 0), (2, 917, 123456789), (0xDEADBEEFCAFE, 5, 2 ** 33)):
        got = tr.qe_normals(seed, turn, base, 50_000)
        exp = O.qe_fill(seed, turn, base, 50_000)
        assert np.array_equal(got, exp)
    tr.close()


@pytest.mark.parametrize("dispersive", [False, True])
def test_section_map(dispersive):
    (ring, secs, der), p = lattice()
    if dispersive:
        secs = [dispersive_section(ring, der)]
    parts = configs.gaussian_bunch(20_000, 11)
    lost = random_losses(20_000, 12)
    tr = tracker(ring, secs, parts, lost)
    tr.seam("section_map", 0)
    got = tr.download()
    ob = O.Bunch(parts, lost)
    O.lib().orc_section_map(ctypes.byref(ob.s), ctypes.byref(O.make_map(secs[0], ring)))
    assert_coords(got, ob)
    assert np.array_equal(got["lost"], lost)
    # lost particles are frozen
    assert np.array_equal(got["x"][lost != 0], parts["x"][lost != 0])
    tr.close()


@pytest.mark.parametrize("big", [False, True])
def test_eigen_emit(big):
    """GetEigenEmit on the device against the oracle: a coupled bunch (x-y
    correlated), a dispersive section 0, 5% lost particles (included, as the
    reference includes them); big: one lost particle at |x| > 1 m, where the
    reference's integer accumulate gives a non-zero "mean"."""
    (ring, secs, der), p = lattice()
    secs = [dispersive_section(ring, der)]
    n = 30_000
    parts = configs.gaussian_bunch(n, 51)
    parts["y"] = parts["y"] + 0.3 * parts["x"]
    parts["py"] = parts["py"] - 0.2 * parts["px"] + 0.1 * parts["x"] / 10.0
    lost = random_losses(n, 52)
    if big:
        i = int(np.argmax(lost != 0))
        parts["x"][i] = 1.7
    tr = tracker(ring, secs, parts, lost)
    got = tr.eigen_emit()[0]
    ob = O.Bunch(parts, lost)
    r = O.rms(ob, secs[0])
    exp = np.array([r.eigen_x, r.eigen_y, r.emit4d, r.xy_alpha])
    assert np.all(np.isfinite(exp[:3])) and exp[1] > 0
    np.testing.assert_allclose(got[[0, 2, 3]], exp[[0, 2, 3]], rtol=1e-10)
    # eigenEmitY^2 = (-tr - sqrt(tr^2 - 16 det)) / 4 cancels by (X/Y)^2: its
    # error is bounded against eigenEmitX^2, not against itself
    assert abs(got[1] ** 2 - exp[1] ** 2) <= 1e-12 * exp[0] ** 2
    tr.close()


def test_rf_drift_u0_skew():
    (ring, secs, _), p = lattice(skew_k=0.03)
    parts = configs.gaussian_bunch(20_000, 21)
    lost = random_losses(20_000, 22)
    tr = tracker(ring, secs, parts, lost)
    tr.seam("skew_quad")
    tr.seam("rf_kick")
    tr.seam("drift")
    tr.seam("energy_loss")
    got = tr.download()
    ob = O.Bunch(parts, lost)
    L = O.lib()
    L.orc_skew_quad(ctypes.byref(ob.s), ring.skew_k)
    L.orc_rf_kick(ctypes.byref(ob.s), ring.rf_vr[0], ring.rf_vi[0], ring.rf_harm[0], ring.f0 * ring.harmonic,
                  ring.rbeta, ring.energy)
    L.orc_drift(ctypes.byref(ob.s), ring.circ, O.p(np.array(ring.alphac[:])))
    L.orc_energy_loss(ctypes.byref(ob.s), ring.u0, ring.energy, ring.rbeta)
    assert_coords(got, ob)
    tr.close()


@pytest.mark.parametrize("noise_mode", [0, 1, 2])
def test_synrad(noise_mode):
    (ring, secs, _), p = lattice()
    n = 20_000
    parts = configs.gaussian_bunch(n, 31)
    lost = random_losses(n, 32)
    tr = tracker(ring, secs, parts, lost)
    tr.set_synrad(True, noise_mode, 2)
    noise = None
    if noise_mode == 1:
        noise = np.random.default_rng(33).standard_normal((5, n))
        tr.push_noise(noise)
    elif noise_mode == 2:
        noise = O.qe_fill(2, 7, 0, n)
    tr.seam("synrad", 7)
    got = tr.download()
    ob = O.Bunch(parts, lost)
    O.lib().orc_synrad(ctypes.byref(ob.s), O.p(np.array(secs[0].b1h1[:])), O.p(np.array(secs[0].h1b1[:])),
                       O.p(np.array(ring.synrad_tau[:])), O.p(np.array(ring.emit_nat[:])),
                       None if noise is None else O.p(np.ascontiguousarray(noise).ravel()))
    assert_coords(got, ob)
    tr.close()


def test_mark_lost_and_rms():
    (ring, secs, _), p = lattice(ap_x=5e-5, ap_y=1.2e-5)
    n = 30_000
    parts = configs.gaussian_bunch(n, 41)
    parts["z"][:50] = 0.31  # beyond c*T0/h/2 = 0.2998 m
    lost = random_losses(n, 42, 0.02)
    tr = tracker(ring, secs, parts, lost)
    t, rc = tr.mark_lost()
    got = tr.download()
    ob = O.Bunch(parts, lost)
    cnt = ctypes.c_int64()
    t_ref = O.lib().orc_mark_lost(ctypes.byref(ob.s), ring.t0, ring.harmonic, ring.ap_x, ring.ap_y, ctypes.byref(cnt))
    assert np.array_equal(got["lost"], ob.lost)
    assert t[0] == t_ref
    assert (rc == capi.CTS_LOW_TRANSMISSION) == (t_ref < 0.5)
    m = tr.rms()[0]
    om = O.rms(ob, secs[0])
    assert m["n_alive"] == om.n_alive
    sig = np.sqrt(np.array([om.c2[0], om.c2[2], om.c2[3], om.c2[5], om.c2[6], om.c2[8]]))
    assert np.all(np.abs(m["mean"] - np.array(om.mean[:])) <= 1e-12 * sig)
    assert np.allclose(m["c2"], np.array(om.c2[:]), rtol=1e-11, atol=0)
    tr.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_sr_wake(mode):
    (ring, secs, _), p = lattice()
    n = 50_000
    parts = configs.gaussian_bunch(n, 51)
    lost = random_losses(n, 52)
    tr = tracker(ring, secs, parts, lost)
    nbins = 100
    sz = p["sig"][2]
    tr.set_wake(nbins, mode, -5 * sz, 5 * sz, BBR["rs"], BBR["q"], BBR["fr"])
    m0 = tr.rms()[0]
    sig_stale = math.sqrt(m0["c2"][6])
    tr.sr_wake([sig_stale] if mode == 0 else None)
    counts, edges, kick = tr.wake_state()
    got = tr.download()
    ob = O.Bunch(parts, lost)
    q, _ = configs.bunch_scalars([n], ring, p)
    rs, qq, om = np.array(BBR["rs"]), np.array(BBR["q"]), np.array(BBR["fr"]) * 2 * np.pi
    poten = np.zeros(3 * nbins)
    oc = np.zeros(nbins, np.int32)
    oe = np.zeros(2)
    O.lib().orc_sr_wake(ctypes.byref(ob.s), nbins, mode, sig_stale, -5 * sz, 5 * sz, q[0], ring.energy, 1,
                        O.p(rs), O.p(qq), O.p(om), O.p(poten), oc.ctypes.data_as(O.PI32), O.p(oe), None)
    assert np.array_equal(edges[0], oe)
    assert np.array_equal(counts[0], oc)
    kz = poten[2 * nbins:]
    assert np.max(np.abs(kick[0] - kz)) <= 1e-12 * np.max(np.abs(kz))
    assert_coords(got, ob)
    tr.close()


def run_both(ring, secs, p, parts, turns, nbins, mode, synrad, noise_seed, check_every=1):
    n = len(parts["x"])
    tr = tracker(ring, secs, parts, p=p)
    tr.set_history(turns)
    sz = p["sig"][2]
    if nbins:
        tr.set_wake(nbins, mode, -5 * sz, 5 * sz, BBR["rs"], BBR["q"], BBR["fr"])
    if synrad:
        tr.set_synrad(True, 2, noise_seed)
    q, _ = configs.bunch_scalars([n], ring, p)
    ot = O.OracleTurn(ring, secs[0], nbins=nbins, mode=mode, grid=(-5 * sz, 5 * sz), bbr=BBR,
                      macro_charge=q[0], synrad=synrad, noise_seed=noise_seed)
    ob = O.Bunch(parts)
    worst = 0.0
    for t in range(turns):
        tr.turn(t)
        ot.turn(ob, t)
        if check_every and (t % check_every == 0 or t == turns - 1):
            got = tr.download()
            assert np.array_equal(got["lost"], ob.lost)
            for c in COORDS:
                scale = np.max(np.abs(ob[c]))
                worst = max(worst, np.max(np.abs(got[c] - ob[c])) / scale)
    return tr, ot, ob, worst


@pytest.mark.parametrize("mode,synrad", [(1, True), (0, False)])
def test_fused_turn_matches_oracle(mode, synrad):
    (ring, secs, _), p = lattice()
    parts = configs.gaussian_bunch(20_000, 61)
    tr, ot, ob, worst = run_both(ring, secs, p, parts, 6, 100, mode, synrad, 2)
    # errors grow by the map's sensitivity over a few turns; 1 turn is 1e-12
    assert worst <= 1e-11, worst
    counts, edges, kick = tr.wake_state()
    assert np.array_equal(counts[0], ot.counts)
    tr.close()


def test_one_turn_1e12():
    (ring, secs, _), p = lattice()
    parts = configs.gaussian_bunch(20_000, 71)
    tr, ot, ob, worst = run_both(ring, secs, p, parts, 1, 100, 1, True, 2)
    assert worst <= 1e-12, worst
    tr.close()


def test_turn_by_turn_moments():
    """Centroid and rms turn by turn (reference-edge binning, as the
    reference bins) to 1e-9 relative to the rms size."""
    (ring, secs, _), p = lattice()
    n, turns = 10_000, 200
    parts = configs.gaussian_bunch(n, 1)
    tr = tracker(ring, secs, parts, p=p)
    tr.set_history(turns)
    sz = p["sig"][2]
    tr.set_wake(100, 0, 0.0, 0.0, BBR["rs"], BBR["q"], BBR["fr"])
    q, _ = configs.bunch_scalars([n], ring, p)
    ot = O.OracleTurn(ring, secs[0], nbins=100, mode=0, bbr=BBR, macro_charge=q[0])
    ob = O.Bunch(parts)
    ref = []
    for t in range(turns):
        tr.turn(t)
        ot.turn(ob, t)
        ref.append(O.rms(ob, secs[0]))
    hist = tr.history(0, turns)
    for t in range(turns):
        g, r = hist[t], ref[t]
        sig = np.sqrt(np.array([r.c2[0], r.c2[2], r.c2[3], r.c2[5], r.c2[6], r.c2[8]]))
        assert g["n_alive"] == r.n_alive
        assert np.all(np.abs(g["mean"] - np.array(r.mean[:])) <= 1e-9 * sig), t
        rg = np.sqrt(np.array(g["c2"])[[0, 2, 3, 5, 6, 8]])
        assert np.all(np.abs(rg - sig) <= 1e-9 * sig), t
    tr.close()


def test_full_size_properties():
    """1e7 macroparticles (the scale-up workload): histogram conservation,
    finite moments and determinism of two identical runs."""
    (ring, secs, _), p = lattice()
    n = 10_000_000
    parts = configs.gaussian_bunch(n, 2)
    outs = []
    for rep in range(2):
        tr = tracker(ring, secs, parts, p=p)
        sz = p["sig"][2]
        tr.set_wake(400, 1, -5 * sz, 5 * sz, BBR["rs"], BBR["q"], BBR["fr"])
        tr.set_synrad(True, 2, 2)
        for t in range(3):
            tr.turn(t)
        counts, edges, kick = tr.wake_state()
        m = tr.moments()[0]
        got = tr.download()
        inside = np.sum((got["z"] >= -5 * sz) & (got["z"] < 5 * sz))
        # counts of the last binning: z before the last kick is the same z
        assert counts[0].sum() <= n
        assert abs(int(counts[0].sum()) - int(inside)) <= 2
        assert np.all(np.isfinite(m["c2"]))
        outs.append((counts.copy(), m["mean"].copy(), got["x"][:1000].copy()))
        tr.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][2], outs[1][2])


@pytest.mark.parametrize("wake_mode", [1, 0])
def test_layouts_agree(wake_mode):
    """Every layout the autotuner chooses between runs the same turn:
    contiguous chunks (one workgroup, several) and round-robin tiles
    (tiles of 256 particles, tiles holding several bunches, more tiles than
    workgroups, more workgroups than tiles) on an uneven 6-bunch train with
    losses, device excitation and the short-range wake.  Loss codes and bin
    counts bit-identical; on the fixed grid (wake mode 1) the coordinates
    bit-identical too; with the reference's edges (mode 0) the edges carry
    the rms bunch length, a sum whose order follows the layout, so the
    coordinates agree to 1e-12; moments to 1e-12 (partial sums regrouped)."""
    (ring, secs, _), p = lattice()
    counts = np.array([700, 5000, 300, 2600, 9000, 40])
    nb = len(counts)
    gap = configs.train_gaps(nb, 3, ring.harmonic)
    q, ne = configs.bunch_scalars(counts, ring, p)
    parts = configs.train_bunches(counts, 17, p)
    lost = random_losses(len(parts["x"]), 5)
    sz = p["sig"][2]
    layouts = [None, (1, 0), (5, 0), (2, 256), (3, 1024), (40, 2048), (1000, 4096)]
    outs = []
    for lay in layouts:
        print(f"layouts agree: wake mode {wake_mode}, layout {lay}", flush=True)
        tr = Tracker(ring, secs)
        tr.set_bunches(counts, q, ne, gap)
        tr.upload(parts, lost)
        tr.set_wake(100, wake_mode, -5 * sz, 5 * sz, BBR["rs"], BBR["q"], BBR["fr"])
        tr.set_synrad(True, 2, 3)
        if lay:
            tr.set_layout(*lay)
        for t in range(3):
            tr.turn(t)
        d = tr.download()
        hist = tr.wake_state()[0]
        outs.append((d, hist, tr.moments()))
        tr.close()
    d0, h0, m0 = outs[0]
    for lay, (d, h, m) in zip(layouts[1:], outs[1:]):
        assert np.array_equal(d["lost"], d0["lost"]), lay
        assert np.array_equal(h, h0), lay
        for c in COORDS:
            if wake_mode == 1:
                assert np.array_equal(d[c], d0[c]), (lay, c)
            else:
                scale = np.max(np.abs(d0[c]))
                assert np.max(np.abs(d[c] - d0[c])) <= 1e-12 * scale, (lay, c)
        for b in range(nb):
            assert m[b]["n_alive"] == m0[b]["n_alive"]
            np.testing.assert_allclose(m[b]["mean"], m0[b]["mean"], rtol=1e-12, atol=1e-12 * np.sqrt(
                np.abs(m0[b]["c2"]).max()))
            np.testing.assert_allclose(m[b]["c2"], m0[b]["c2"], rtol=1e-12, atol=1e-13 * np.abs(m0[b]["c2"]).max())


def test_autotuned_layout_keeps_parity():
    """The autotuner only changes the workgroup layout: turns after it still
    match the oracle (and its timing runs leave the particles untouched)."""
    (ring, secs, _), p = lattice()
    parts = configs.gaussian_bunch(30_000, 31, p)
    tr = tracker(ring, secs, parts, p=p)
    sz = p["sig"][2]
    q, _ = configs.bunch_scalars([30_000], ring, p)
    tr.set_wake(100, 1, -5 * sz, 5 * sz, BBR["rs"], BBR["q"], BBR["fr"])
    tr.set_synrad(True, 2, 3)
    nwg = tr.autotune()
    assert nwg >= 1
    before = tr.download()
    for c in COORDS:
        assert np.array_equal(before[c], parts[c])
    ot = O.OracleTurn(ring, secs[0], nbins=100, mode=1, grid=(-5 * sz, 5 * sz), bbr=BBR, macro_charge=q[0],
                      synrad=True, noise_seed=3)
    ob = O.Bunch(parts)
    for t in range(3):
        tr.turn(t)
        ot.turn(ob, t)
    assert_coords(tr.download(), ob, 1e-11)
    tr.close()
