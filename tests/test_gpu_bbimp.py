"""The broadband impedance on the HIP path against the oracle: trains of
uneven bunches with broadband-resonator impedance tables (longitudinal
alone, with the short-range wake, with the quadrupolar planes, and with all
five planes on bunches at transverse offsets), over 5 turns.  The device forms the
induced voltage as a circular convolution with the impedance's real kernel,
the oracle as the reference does (DFT, product, inverse DFT): voltages agree
to 1e-10 of their largest value, profiles to 1e-12, coordinates to 1e-11 of
each coordinate's scale as the other train tests.  Graph replays equal the
turns run one by one."""
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
    # a plane on without its table; the time-domain method; the dipolar
    # planes on a grid above 1535 bins
    zbig = np.zeros(800)
    for args, kw in (((zre, zim, fmax), dict(flags=(1, 1, 0, 0, 0))), ((zre, zim, fmax), dict(time_domain=1)),
                     ((zre, zim, fmax), dict(flags=(1, 0, 0, 1, 0))),
                     ((zbig, zbig, fmax), dict(flags=(0, 1, 0, 0, 0), zdx=(zbig, zbig)))):
        with pytest.raises(RuntimeError):
            tr.set_bbimp(*args, **kw)
    # a grid shorter than one RF bucket (z_max < c T_rf / 2)
    with pytest.raises(RuntimeError):
        tr.set_bbimp(zre[:50], zim[:50], fmax)
    tr.set_bbimp(zre, zim, fmax, flags=(0, 0, 0, 0, 0))  # off
    tr.turn(0)
    tr.close()


@pytest.mark.parametrize("wake_mode", [1, None])
def test_graph_replays_with_bbimp(wake_mode):
    e, g = make([3000, 5001, 2500, 1800], wake_mode, False)
    zre, zim, fmax = impedance()
    for tr in (e, g):
        tr.set_bbimp(zre, zim, fmax)
    print(f"graph case with broadband impedance, wake={wake_mode}", flush=True)
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
