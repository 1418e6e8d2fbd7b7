"""Beam-loaded (rfMode 1) cavities on the HIP path against the oracle: the
per-turn cavity, beam-induced, generator and self-loss voltages of every
bunch (the host tables; both sides run the same recurrences, so 1e-12 of
the voltage scale), and bunch trains tracked with one active beam-loaded
cavity in place of the ideal one, with and without a short-range wake:
coordinates 1e-11 of each coordinate's scale, as the other train tests.
The graph replay of such turns equals the turns run one by one."""
import numpy as np
import pytest

from cetasim_amd import configs
from cetasim_amd.tracker import Tracker

import oracle as O
from test_gpu_graph import same
from test_gpu_train import BBR, COORDS, compare_coords

pytestmark = pytest.mark.gpu

SZ = configs.BASE["sig"][2]


def cavity(ring, nturns, nbins=60, rs=5e6, q0=2e4, beta=3.0, detune=-2e4):
    return dict(nres=1, mode=[1], type=[1], harm=[1], cold=[0], rs=[rs], q0=[q0], coupling=[beta],
                detune=[detune], vol=[configs.rf_voltage(configs.BASE)], phase=[0.0], nbins=nbins,
                nturns=nturns)


def setup(counts, seed, keep_ideal=False):
    ring, secs, _ = configs.lattice()
    if not keep_ideal:
        ring.n_rf = 0  # the beam-loaded cavity replaces the ideal one
    counts = np.asarray(counts, dtype=np.int64)
    nb = len(counts)
    gap = configs.train_gaps(nb, 3, ring.harmonic)
    q, ne = configs.bunch_scalars(counts, ring)
    parts = configs.train_bunches(counts, seed)
    tr = Tracker(ring, secs)
    tr.set_bunches(counts, q, ne, gap)
    tr.upload(parts)
    off = np.concatenate([[0], np.cumsum(counts)])
    bunches = [O.Bunch({c: parts[c][off[b]:off[b + 1]] for c in COORDS}) for b in range(nb)]
    return ring, secs, counts, gap, q, ne, tr, bunches, off


def test_cavity_tables_match_oracle():
    counts = [800 + 97 * b for b in range(9)]
    ring, secs, counts, gap, q, ne, tr, bunches, off = setup(counts, 2)
    cfg = cavity(ring, 12)
    tr.set_beam_loading(cfg)
    ot = O.OracleTrain(ring, secs[0], counts, ne, gap, None, 1, bl=cfg)
    scale = cfg["vol"][0]
    for t in (0, 5, 11):
        got = tr.cavity_info(t)
        np.testing.assert_allclose(got, ot.bl_info[t], rtol=0, atol=1e-12 * scale)
    # the train loads the cavity: induced voltage well above the tolerance
    assert np.max(np.abs(ot.bl_info[11, :, 0, 2:4])) > 1e3
    with pytest.raises(RuntimeError):
        tr.cavity_info(12)
    tr.close()


@pytest.mark.parametrize("sr_wake", [False, True])
def test_beam_loaded_turns_match_oracle(sr_wake):
    counts = [1500 + 173 * (b % 4) for b in range(10)]
    ring, secs, counts, gap, q, ne, tr, bunches, off = setup(counts, 17)
    cfg = cavity(ring, 8)
    tr.set_beam_loading(cfg)
    tr.set_synrad(True, 2, 6)
    kw = {}
    if sr_wake:
        tr.set_wake(100, 1, -5 * SZ, 5 * SZ, BBR["rs"], BBR["q"], BBR["fr"])
        kw = dict(nbins=100, mode=1, grid=(-5 * SZ, 5 * SZ), bbr=BBR, macro_charge=list(q))
    ot = O.OracleTrain(ring, secs[0], counts, ne, gap, None, 1, bl=cfg, synrad=True, noise_seed=6, **kw)
    print(f"beam-loaded train, short-range wake {sr_wake}", flush=True)
    for t in range(6):
        tr.turn(t)
        ot.turn(bunches, t)
    got = tr.download()
    assert not np.any(got["lost"])
    compare_coords(got, bunches, off, 1e-11)
    m = tr.moments()
    for b, ob in enumerate(bunches):
        r = O.rms(ob, secs[0])
        assert m[b]["n_alive"] == r.n_alive
        for i in (0, 2, 4, 5):
            sd = abs(m[b]["c2"][3 * (i // 2)]) ** 0.5
            assert abs(m[b]["mean"][i] - r.mean[i]) <= 1e-9 * max(sd, 1e-30), (b, i)
    tr.close()


def test_turn_range_is_checked():
    ring, secs, counts, gap, q, ne, tr, bunches, off = setup([2000, 2000], 3)
    tr.set_beam_loading(cavity(ring, 3))
    for t in range(3):
        tr.turn(t)
    with pytest.raises(RuntimeError):
        tr.turn(3)
    # the per-bin kicks of a bunch must fit in LDS
    with pytest.raises(RuntimeError):
        tr.set_beam_loading(cavity(ring, 3, nbins=4096))
    tr.close()


@pytest.mark.parametrize("wake_mode", [1, None])
def test_graph_replays_with_beam_loading(wake_mode):
    """Two passes per turn (the bin range needs the post-map extent); the
    table row follows the device turn counter."""
    pair = []
    for _ in range(2):
        ring, secs, counts, gap, q, ne, tr, bunches, off = setup([3000, 5001, 2500, 1800], 5)
        tr.set_beam_loading(cavity(ring, 16))
        if wake_mode is not None:
            tr.set_wake(100, wake_mode, -5 * SZ, 5 * SZ, BBR["rs"], BBR["q"], BBR["fr"])
        tr.set_synrad(True, 2, 21)
        tr.set_history(64)
        pair.append(tr)
    e, g = pair
    print(f"graph case with beam loading, wake={wake_mode}", flush=True)
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


def walked(ring, q, nturns, excite, dirfb, nbins=50):
    cfg = cavity(ring, nturns, nbins=nbins)
    cfg.update(excite=[excite], dirfb=[dirfb], fb_gain=[0.3], fb_phase=[0.2], fb_delay=[1e-10],
               macro_charge=list(q))
    return cfg


@pytest.mark.parametrize("excite,dirfb,sr_wake", [(1, 0, False), (0, 1, False), (1, 1, True)])
def test_walked_cavity_matches_oracle(excite, dirfb, sr_wake):
    """The bin-by-bin variant and the direct feedback: the cavity state is
    walked over the train on the device every turn (MPBeam::
    BeamMomtumUpdateDueToRFTest with MPBunch::BunchMomentumUpdateDueToRFMode
    and Bunch::GetLongiKickDueToCavFB) against the oracle's walk: every
    bunch's cavity, induced, generator and self-loss voltages to 1e-10 of
    the voltage scale each turn, coordinates to 1e-11 of each coordinate's
    scale after 6 turns, moments as the other train tests.  Lost particles
    (a tenth of one bunch) get the feedback kick and nothing else."""
    counts = [1200 + 131 * b for b in range(7)]
    ring, secs, counts, gap, q, ne, tr, bunches, off = setup(counts, 9)
    lost = np.zeros(int(np.sum(counts)), np.int32)
    lost[off[2]:off[2] + counts[2] // 10] = 1
    parts = {c: np.concatenate([b[c] for b in bunches]) for c in COORDS}
    tr.upload(parts, lost)
    bunches = [O.Bunch({c: parts[c][off[b]:off[b + 1]] for c in COORDS}, lost[off[b]:off[b + 1]])
               for b in range(len(counts))]
    cfg = walked(ring, q, 8, excite, dirfb)
    tr.set_beam_loading(cfg)
    tr.set_synrad(True, 2, 6)
    kw = {}
    if sr_wake:
        tr.set_wake(100, 1, -5 * SZ, 5 * SZ, BBR["rs"], BBR["q"], BBR["fr"])
        kw = dict(nbins=100, mode=1, grid=(-5 * SZ, 5 * SZ), bbr=BBR, macro_charge=list(q))
    ot = O.OracleTrain(ring, secs[0], counts, ne, gap, None, 1, bl=cfg, synrad=True, noise_seed=6, **kw)
    scale = cfg["vol"][0]
    print(f"walked cavity: excite {excite}, feedback {dirfb}, short-range wake {sr_wake}", flush=True)
    for t in range(6):
        tr.turn(t)
        ot.turn(bunches, t)
        info = tr.cavity_info(t)
        np.testing.assert_allclose(info, ot.bl_last_info, rtol=0, atol=1e-10 * scale)
    if dirfb:
        assert np.max(np.abs(ot.bl_last_fb)) > 0
    got = tr.download()
    compare_coords(got, bunches, off, 1e-11)
    m = tr.moments()
    for b, ob in enumerate(bunches):
        r = O.rms(ob, secs[0])
        assert m[b]["n_alive"] == r.n_alive
        for i in (0, 2, 4, 5):
            sd = abs(m[b]["c2"][3 * (i // 2)]) ** 0.5
            assert abs(m[b]["mean"][i] - r.mean[i]) <= 1e-9 * max(sd, 1e-30), (b, i)
    tr.close()


def test_walked_cavity_settings_are_checked():
    ring, secs, counts, gap, q, ne, tr, bunches, off = setup([2000, 2000], 3)
    # the walked cavities keep two kicks per bin in LDS
    with pytest.raises(RuntimeError):
        tr.set_beam_loading(walked(ring, q, 3, 1, 0, nbins=2100))
    tr.set_beam_loading(walked(ring, q, 3, 1, 1))
    for t in range(5):  # no turn limit; the record keeps turns 0-2
        tr.turn(t)
    tr.cavity_info(2)
    with pytest.raises(RuntimeError):
        tr.cavity_info(3)
    tr.close()


@pytest.mark.parametrize("wake_mode", [1, None])
def test_graph_replays_with_walked_cavity(wake_mode):
    """Three passes and the walk per turn; the walk reads its turn from the
    device counter for the record."""
    pair = []
    for _ in range(2):
        ring, secs, counts, gap, q, ne, tr, bunches, off = setup([3000, 5001, 2500, 1800], 5)
        tr.set_beam_loading(walked(ring, q, 16, 1, 1))
        if wake_mode is not None:
            tr.set_wake(100, wake_mode, -5 * SZ, 5 * SZ, BBR["rs"], BBR["q"], BBR["fr"])
        tr.set_synrad(True, 2, 21)
        tr.set_history(64)
        pair.append(tr)
    e, g = pair
    print(f"graph case with a walked cavity, wake={wake_mode}", flush=True)
    for t in range(9):
        e.turn(t)
    g.run_turns(0, 9)
    same(e, g, 0, 8, "turns 0-8")
    for t in range(9, 14):
        e.turn(t)
    g.run_turns(9, 5)
    same(e, g, 9, 13, "turns 9-13")
    for t in (3, 8, 13):
        assert np.array_equal(e.cavity_info(t), g.cavity_info(t)), t
    e.close()
    g.close()
