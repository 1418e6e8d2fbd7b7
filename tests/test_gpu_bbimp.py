"""The broadband impedance on the HIP path against the oracle: trains of
uneven bunches with broadband-resonator impedance tables (longitudinal
alone, with the short-range wake, with the quadrupolar planes, and with all
five planes on bunches at transverse offsets), over 5 turns.  The device forms the
induced voltage as a circular convolution with the impedance's real kernel,
the oracle as the reference does (DFT, product, inverse DFT): voltages agree
to 1e-10 of their largest value, profiles to 1e-12, coordinates to 1e-11 of
each coordinate's scale as the other train tests.  The time-domain method
(the reference's default) likewise, with a wake table of damped
oscillations, and on a fine grid (10 um bins) where a bunch spans several
LDS windows of the binning.  Graph replays equal the turns run one by
one."""
import numpy as np
import pytest

from cetasim_amd import configs

import oracle as O
from test_gpu_graph import make, same
from test_gpu_train import BBR, compare_coords, setup_train

pytestmark = pytest.mark.gpu

SZ = configs.BASE["sig"][2]


def impedance(nimp=261, f_max=1.0e11, rs=2.0e3, q=1.0, fr=2.0e10):
    """Broadband resonator Z(f) = Rs / (1 + i Q (f/fr - fr/f)) on the file's
    grid f_k = k f_max / (nimp - 1)."""
    f = np.arange(nimp) * f_max / (nimp - 1)
    z = np.zeros(nimp, complex)
    z[1:] = rs / (1 + 1j * q * (f[1:] / fr - fr / f[1:]))
    return z.real.copy(), z.imag.copy(), f_max


def quad_impedance(nimp=261, f_max=1.0e11, rs=1.0e3, q=1.0, fr=1.5e10):
    """Transverse broadband resonator Z_Q(f) = (fr/f) Rs / (1 + i Q (f/fr - fr/f))."""
    f = np.arange(nimp) * f_max / (nimp - 1)
    z = np.zeros(nimp, complex)
    z[1:] = (fr / f[1:]) * rs / (1 + 1j * q * (f[1:] / fr - fr / f[1:]))
    return z.real.copy(), z.imag.copy()


@pytest.mark.parametrize("sr_wake,quad,dip", [(False, False, False), (True, False, False), (False, True, False),
                                               (False, True, True)])
def test_bbimp_turns_match_oracle(sr_wake, quad, dip):
    counts = [3000, 7001, 1500, 4096]
    ring, secs, counts, gap, q, ne, _, tr, bunches, off = setup_train(counts, 13)
    if dip:
        # transverse offsets per bunch, so the x- and y-weighted profiles
        # carry the bunches' positions rather than their noise
        parts = {c: np.concatenate([b[c] for b in bunches]) for c in O.COORDS}
        for b in range(len(counts)):
            parts["x"][off[b]:off[b + 1]] += 4e-5 * (b + 1)
            parts["y"][off[b]:off[b + 1]] -= 1.5e-5 * (b + 1)
        tr.upload(parts)
        bunches = [O.Bunch({c: parts[c][off[b]:off[b + 1]] for c in O.COORDS}) for b in range(len(counts))]
    zre, zim, fmax = impedance()
    qx, qy = quad_impedance(), quad_impedance(rs=1.5e3, fr=1.2e10)
    dx, dy = quad_impedance(rs=2.5e3, fr=1.8e10), quad_impedance(rs=0.8e3, fr=1.0e10)
    beta = (secs[0].beta_x, secs[0].beta_y)
    # planes in the reference's flag order: Z, Dx, Dy, Qx, Qy
    flags = (1, int(dip), int(dip), int(quad), int(quad))
    planes = [k for k in range(5) if flags[k]]
    tr.set_bbimp(zre, zim, fmax, flags=flags, zqx=qx if quad else None, zqy=qy if quad else None, beta=beta,
                 zdx=dx if dip else None, zdy=dy if dip else None)
    tr.set_synrad(True, 2, 8)
    kw = {}
    if sr_wake:
        tr.set_wake(100, 1, -5 * SZ, 5 * SZ, BBR["rs"], BBR["q"], BBR["fr"])
        kw = dict(nbins=100, mode=1, grid=(-5 * SZ, 5 * SZ), bbr=BBR, macro_charge=list(q))
    bbi = dict(nimp=len(zre), f_max=fmax, zre=zre, zim=zim, q=list(q), beta=beta)
    if quad:
        bbi.update(zqx=qx, zqy=qy)
    if dip:
        bbi.update(zdx=dx, zdy=dy)
    ot = O.OracleTrain(ring, secs[0], counts, ne, gap, None, 1, bbi=bbi, synrad=True, noise_seed=8, **kw)
    print(f"broadband impedance train, short-range wake {sr_wake}, planes {flags}", flush=True)
    for t in range(5):
        tr.turn(t)
        ot.turn(bunches, t)
        w, pr = tr.bbimp_wake()
        for b in range(len(counts)):
            ow, op = ot.turns[b].bbi_wake, ot.turns[b].bbi_prof
            for r in range(3 if dip else 1):
                np.testing.assert_allclose(pr[r][b], op[r], rtol=0, atol=1e-12 * np.max(np.abs(op[r])))
            for k in range(5):
                scale = np.max(np.abs(ow[k]))
                if k in planes:
                    assert scale > 0
                else:
                    assert scale == 0 and np.all(w[k][b] == 0)
                np.testing.assert_allclose(w[k][b], ow[k], rtol=0, atol=1e-10 * scale)
    got = tr.download()
    assert not np.any(got["lost"])
    compare_coords(got, bunches, off, 1e-11)
    m = tr.moments()
    for b, ob in enumerate(bunches):
        r = O.rms(ob, secs[0])
        sd = abs(m[b]["c2"][8]) ** 0.5
        assert abs(m[b]["mean"][5] - r.mean[5]) <= 1e-9 * sd, b
    # the kick is not negligible against the tolerance: |V|/E well above 1e-11 of sigma_delta
    assert np.max(np.abs(w)) / ring.energy > 1e-7
    tr.close()


def test_bbimp_settings_are_checked():
    ring, secs, counts, gap, q, ne, _, tr, bunches, off = setup_train([2000], 3)
    zre, zim, fmax = impedance()
    # a plane on without its table; the time domain without its wake table;
    # a method other than 0 and 1; a grid above 16383 bins
    zbig = np.zeros(8193)
    for args, kw in (((zre, zim, fmax), dict(flags=(1, 1, 0, 0, 0))), ((zre, zim, fmax), dict(time_domain=1)),
                     ((zre, zim, fmax), dict(time_domain=2)), ((zre, zim, fmax), dict(flags=(1, 0, 0, 1, 0))),
                     ((zbig, zbig, fmax), dict(flags=(0, 1, 0, 0, 0), zdx=(zbig, zbig)))):
        with pytest.raises(RuntimeError):
            tr.set_bbimp(*args, **kw)
    # time domain: one row; a step that cuts the bucket (0.6 m) into more
    # than 65535 bins; a zero step
    pos, wakes = td_table(nw=50)
    for p_, w_ in ((pos[:1], [w[:1] for w in wakes]), (pos * 0.9e-3, wakes), (np.zeros(50), wakes)):
        with pytest.raises(RuntimeError):
            tr.set_bbimp_td(p_, w_, beta=(5.0, 3.0))
    with pytest.raises(ValueError):
        tr.set_bbimp_td(pos, [wakes[0][:10]] + [None] * 4)
    # a grid shorter than one RF bucket (z_max < c T_rf / 2)
    with pytest.raises(RuntimeError):
        tr.set_bbimp(zre[:50], zim[:50], fmax)
    tr.set_bbimp(zre, zim, fmax, flags=(0, 0, 0, 0, 0))  # off
    tr.turn(0)
    tr.close()


@pytest.mark.parametrize("wake_mode,td", [(1, False), (None, False), (None, True)])
def test_graph_replays_with_bbimp(wake_mode, td):
    e, g = make([3000, 5001, 2500, 1800], wake_mode, False)
    zre, zim, fmax = impedance()
    pos, wakes = td_table()
    for tr in (e, g):
        if td:
            tr.set_bbimp_td(pos, wakes, beta=(5.0, 3.0))
        else:
            tr.set_bbimp(zre, zim, fmax)
    print(f"graph case with broadband impedance, wake={wake_mode}, time domain {td}", flush=True)
    for t in range(9):
        e.turn(t)
    g.run_turns(0, 9)
    same(e, g, 0, 8, "turns 0-8")
    for t in range(9, 14):
        e.turn(t)
    g.run_turns(9, 5)
    same(e, g, 9, 13, "turns 9-13")
    e.close()
    g.close()


def td_table(nw=400, dz=5.0e-4, a=-2.0e12, b=-3.0e13, ell=0.02, lam=4.0e-3):
    """A quasi-green wake table as GetQuasiWakePoten leaves it (positions
    from 0, step dz): wz a damped cosine, the transverse wakes damped sines
    (dipolar x, y at 1 and 0.6 of b, quadrupolar at 0.3 and 0.2)."""
    z = np.arange(nw) * dz
    wz = a * np.exp(-z / ell) * np.cos(z / lam)
    wt = b * np.exp(-z / ell) * np.sin(z / lam)
    return z, [wz, wt, 0.6 * wt, 0.3 * wt, 0.2 * wt]


def run_td(counts, seed, flags, pos, wakes, turns, sr_wake=False, x_offsets=False, stretch=None):
    """Train through the device and the oracle with the time-domain
    broadband impedance; compares V and the profiles every turn and the
    coordinates at the end."""
    ring, secs, counts, gap, q, ne, _, tr, bunches, off = setup_train(counts, seed)
    if x_offsets or stretch:
        parts = {c: np.concatenate([b[c] for b in bunches]) for c in O.COORDS}
        for b in range(len(counts)):
            if x_offsets:
                parts["x"][off[b]:off[b + 1]] += 4e-5 * (b + 1)
                parts["y"][off[b]:off[b + 1]] -= 1.5e-5 * (b + 1)
            if stretch and b == 0:
                parts["z"][off[b]:off[b + 1]] = np.random.default_rng(seed).uniform(-stretch, stretch, counts[b])
        tr.upload(parts)
        bunches = [O.Bunch({c: parts[c][off[b]:off[b + 1]] for c in O.COORDS}) for b in range(len(counts))]
    beta = (secs[0].beta_x, secs[0].beta_y)
    w = [wakes[k] if flags[k] else None for k in range(5)]
    tr.set_bbimp_td(pos, w, beta=beta)
    tr.set_synrad(True, 2, 8)
    kw = {}
    if sr_wake:
        tr.set_wake(100, 1, -5 * SZ, 5 * SZ, BBR["rs"], BBR["q"], BBR["fr"])
        kw = dict(nbins=100, mode=1, grid=(-5 * SZ, 5 * SZ), bbr=BBR, macro_charge=list(q))
    bbi = dict(td=1, pos=pos, w=w, q=list(q), beta=beta)
    ot = O.OracleTrain(ring, secs[0], counts, ne, gap, None, 1, bbi=bbi, synrad=True, noise_seed=8, **kw)
    planes = [k for k in range(5) if flags[k]]
    dip = flags[1] or flags[2]
    for t in range(turns):
        tr.turn(t)
        ot.turn(bunches, t)
        V, pr = tr.bbimp_wake()
        for b in range(len(counts)):
            oV, op = ot.turns[b].bbi_wake, ot.turns[b].bbi_prof
            assert V.shape[2] == oV.shape[1]
            for r in range(3 if dip else 1):
                if r and not flags[r]:
                    continue
                np.testing.assert_allclose(pr[r][b], op[r], rtol=0, atol=1e-12 * np.max(np.abs(op[r])))
            for k in range(5):
                scale = np.max(np.abs(oV[k]))
                if k in planes:
                    assert scale > 0
                else:
                    assert scale == 0 and np.all(V[k][b] == 0)
                np.testing.assert_allclose(V[k][b], oV[k], rtol=0, atol=1e-10 * scale)
    got = tr.download()
    assert not np.any(got["lost"])
    compare_coords(got, bunches, off, 1e-11)
    tr.close()
    return V, ring


@pytest.mark.parametrize("sr_wake,flags", [(False, (1, 0, 0, 0, 0)), (True, (1, 0, 0, 0, 0)),
                                           (False, (1, 0, 0, 1, 1)), (False, (1, 1, 1, 1, 1))])
def test_bbimp_td_turns_match_oracle(sr_wake, flags):
    pos, wakes = td_table()
    print(f"time-domain broadband impedance train, short-range wake {sr_wake}, planes {flags}", flush=True)
    V, ring = run_td([3000, 7001, 1500, 4096], 13, flags, pos, wakes, 5, sr_wake=sr_wake,
                     x_offsets=bool(flags[1] or flags[2]))
    assert V.shape[2] == 1199  # (int)(0.599584916 / 5e-4)
    # the kick is not negligible against the tolerance
    assert np.max(np.abs(V)) / ring.energy > 1e-7


@pytest.mark.parametrize("flags,stretch", [((1, 0, 0, 0, 0), 0.06), ((1, 1, 1, 0, 0), None)])
def test_bbimp_td_fine_grid_spans_several_windows(flags, stretch):
    """10 um bins: 59958 over the bucket.  A bunch stretched over +-60 mm
    spans 12000 bins (two windows of the 8192-bin count histogram); with
    the dipolar planes a bunch of sigma_z 3 mm spans ~2400 bins (two
    windows of 1535)."""
    pos, wakes = td_table(nw=2000, dz=1e-5, ell=4e-3, lam=1e-3)
    print(f"time-domain fine grid, planes {flags}, stretch {stretch}", flush=True)
    V, _ = run_td([6000, 3000, 2000], 17, flags, pos, wakes, 2, x_offsets=bool(flags[1]), stretch=stretch)
    assert V.shape[2] == 59958


@pytest.mark.parametrize("td", [False, True])
def test_bbimp_with_lost_and_off_grid_particles(td):
    """All five planes with particles lost at the first turn's marking (60
    beyond the aperture: binned and kicked in that turn, in no later one)
    and 25 off the grid (time domain: past the bucket's last bin, which ends
    8.5e-5 m short of L/2; frequency domain: beyond z_max, lost at the same
    marking): neither binned nor kicked, on the device as in the oracle."""
    counts = [3000, 4000, 2500]
    ring, secs, counts, gap, q, ne, _, tr, bunches, off = setup_train(counts, 23)
    parts = {c: np.concatenate([b[c] for b in bunches]) for c in O.COORDS}
    for b in range(len(counts)):
        parts["x"][off[b]:off[b + 1]] += 4e-5 * (b + 1)
    parts["x"][off[0]:off[0] + 60] = 0.05
    if td:
        parts["z"][off[1]:off[1] + 25] = 0.29975
        parts["dp"][off[1]:off[1] + 25] = 0.0
    else:
        parts["z"][off[1]:off[1] + 25] = -0.4
    tr.upload(parts)
    bunches = [O.Bunch({c: parts[c][off[b]:off[b + 1]] for c in O.COORDS}) for b in range(len(counts))]
    beta = (secs[0].beta_x, secs[0].beta_y)
    if td:
        pos, wakes = td_table()
        tr.set_bbimp_td(pos, wakes, beta=beta)
        bbi = dict(td=1, pos=pos, w=wakes, q=list(q), beta=beta)
    else:
        zre, zim, fmax = impedance()
        qx, qy = quad_impedance(), quad_impedance(rs=1.5e3, fr=1.2e10)
        dx, dy = quad_impedance(rs=2.5e3, fr=1.8e10), quad_impedance(rs=0.8e3, fr=1.0e10)
        tr.set_bbimp(zre, zim, fmax, flags=(1, 1, 1, 1, 1), zqx=qx, zqy=qy, zdx=dx, zdy=dy, beta=beta)
        bbi = dict(nimp=len(zre), f_max=fmax, zre=zre, zim=zim, q=list(q), beta=beta, zqx=qx, zqy=qy, zdx=dx, zdy=dy)
    tr.set_synrad(True, 2, 8)
    ot = O.OracleTrain(ring, secs[0], counts, ne, gap, None, 1, bbi=bbi, synrad=True, noise_seed=8)
    print(f"broadband impedance with lost and off-grid particles, time domain {td}", flush=True)
    for t in range(3):
        tr.turn(t)
        ot.turn(bunches, t)
        w, pr = tr.bbimp_wake()
        for b in range(len(counts)):
            ow, op = ot.turns[b].bbi_wake, ot.turns[b].bbi_prof
            for r in range(3):
                np.testing.assert_allclose(pr[r][b], op[r], rtol=0, atol=1e-12 * np.max(np.abs(op[r])))
            for k in range(5):
                scale = np.max(np.abs(ow[k]))
                assert scale > 0
                np.testing.assert_allclose(w[k][b], ow[k], rtol=0, atol=1e-10 * scale)
    got = tr.download()
    lost = got["lost"]
    assert np.all(lost[off[0]:off[0] + 60] != 0)
    if td:
        assert not np.any(lost[off[1]:off[1] + 25])
    else:
        assert np.all(lost[off[1]:off[1] + 25] != 0)
    compare_coords(got, bunches, off, 1e-11)
    tr.close()
