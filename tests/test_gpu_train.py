"""Bunch trains on the HIP path vs the CPU oracle (MI355X): uneven
macroparticle counts in one launch with per-bunch bins, and the beam-ion
phase (ion generation, cloud moments, Bassetti-Erskine kicks both ways,
ion drift and loss, bunch by bunch in train order).

Tolerances: bin counts and loss codes bit-exact; coordinates 1e-11 relative
to each coordinate's largest magnitude; ion cloud 1e-9 relative (its moments
are summed in a different order than the oracle's sequential loops, and the
drift step C/h*gap/c is folded into one factor).
"""
import numpy as np
import pytest

from cetasim_amd import capi, configs
from cetasim_amd.tracker import Tracker

import oracle as O

pytestmark = pytest.mark.gpu

COORDS = O.COORDS
BBR = configs.BASE["bbr"]


def setup_train(counts, seed, ions=None, spacing=2, gap=None, p=configs.BASE, ne_scale=1.0):
    ring, secs, _ = configs.lattice(p)
    counts = np.asarray(counts, dtype=np.int64)
    nb = len(counts)
    gap = configs.train_gaps(nb, spacing, ring.harmonic) if gap is None else np.asarray(gap, dtype=np.int32)
    q, ne = configs.bunch_scalars(counts, ring, p)
    ne = ne * ne_scale
    parts = configs.train_bunches(counts, seed, p)
    tr = Tracker(ring, secs)
    tr.set_bunches(counts, q, ne, gap)
    tr.upload(parts)
    cfg = None
    if ions is not None:
        cfg = configs.ion_config(ring, ions, p)
        tr.set_ions(cfg)
    off = np.concatenate([[0], np.cumsum(counts)])
    bunches = [O.Bunch({c: parts[c][off[b]:off[b + 1]] for c in COORDS}) for b in range(nb)]
    return ring, secs, counts, gap, q, ne, cfg, tr, bunches, off


def compare_coords(got, bunches, off, tol):
    for b, ob in enumerate(bunches):
        for c in COORDS:
            a = got[c][off[b]:off[b + 1]]
            e = ob[c]
            scale = max(np.max(np.abs(e)), 1e-300)
            err = np.max(np.abs(a - e)) / scale
            assert err <= tol, (b, c, err)
        assert np.array_equal(got["lost"][off[b]:off[b + 1]], ob.lost), b


ION_TEST = dict(mass_number=[28.0], cross_section=[2.0e-22], pressure=[1.0e-5], temperature=300.0,
                n_new=300, max_ions=50_000, loss_boundary=10.0, seed=11)


@pytest.mark.parametrize("mode", [1, 0])
def test_train_uneven_counts(mode):
    """Five bunches of uneven size, per-bunch SR-wake bins (fixed grid or
    the reference's edges), device quantum excitation: 4 turns."""
    counts = [3000, 7001, 1500, 4096, 2222]
    ring, secs, counts, gap, q, ne, _, tr, bunches, off = setup_train(counts, 5)
    sz = configs.BASE["sig"][2]
    nbins = 120
    tr.set_wake(nbins, mode, -5 * sz, 5 * sz, BBR["rs"], BBR["q"], BBR["fr"])
    tr.set_synrad(True, 2, 9)
    ot = O.OracleTrain(ring, secs[0], counts, ne, gap, None, 0, nbins=nbins, mode=mode, grid=(-5 * sz, 5 * sz),
                       bbr=BBR, macro_charge=list(q), synrad=True, noise_seed=9)
    for t in range(4):
        tr.turn(t)
        ot.turn(bunches, t)
        cnt, _, _ = tr.wake_state()
        for b in range(len(counts)):
            assert np.array_equal(cnt[b], ot.turns[b].counts[:nbins]), (t, b)
    got = tr.download()
    compare_coords(got, bunches, off, 1e-11)
    m = tr.moments()
    for b, ob in enumerate(bunches):
        om = O.rms(ob, secs[0])
        assert m[b]["n_alive"] == om.n_alive
        for i in (0, 2, 4):
            assert abs(m[b]["mean"][i] - om.mean[i]) <= 1e-9 * abs(m[b]["c2"][3 * (i // 2)]) ** 0.5
    tr.close()


def test_ion_phase_matches_oracle():
    """One ion phase (4 bunch passages) from an empty cloud: the cloud, the
    per-passage cloud moments and the electrons' kicks."""
    counts = [2000, 3500, 1200, 2800]
    # short gaps throughout, so that part of the cloud survives the phase
    ring, secs, counts, gap, q, ne, cfg, tr, bunches, off = setup_train(counts, 7, ions=ION_TEST, gap=[2, 2, 2, 3])
    ot = O.OracleTrain(ring, secs[0], counts, ne, gap, cfg, ION_TEST["seed"])
    before = {c: np.concatenate([b[c] for b in bunches]) for c in COORDS}
    tr.ion_phase(0)
    exp = ot.ion_phase(bunches, 0)
    got = tr.ion_passages()
    for b in range(len(counts)):
        # ion number, mean x/y, rms x/y seen by passage b
        np.testing.assert_allclose(got[b, 0, 5], exp[b, 0, 0], rtol=1e-12)
        np.testing.assert_allclose(got[b, 0, 1:5], exp[b, 0, 1:5], rtol=1e-9, atol=1e-18)
        # macro-ion count: survivors of the earlier passages plus n_new
        assert got[b, 0, 0] <= (b + 1) * ION_TEST["n_new"]
    assert got[0, 0, 0] == ION_TEST["n_new"]
    cloud = tr.ion_cloud(0)
    ec = ot.cloud.species(0)
    assert len(cloud["x"]) == len(ec["x"]) > 0
    for k in ("x", "y", "vx", "vy"):
        np.testing.assert_allclose(cloud[k], ec[k], rtol=1e-9, atol=1e-9 * np.max(np.abs(ec[k])))
    after = tr.download()
    compare_coords(after, bunches, off, 1e-12)
    # the kick is far above the tolerance (the comparison is not vacuous)
    dpx = np.max(np.abs(after["px"] - before["px"]))
    assert dpx > 1e-6 * np.max(np.abs(before["px"])), dpx
    tr.close()


def test_ion_train_turns():
    """Three full turns with the ion phase, wake and excitation on."""
    counts = [2500, 1800, 3300]
    ring, secs, counts, gap, q, ne, cfg, tr, bunches, off = setup_train(counts, 8, ions=ION_TEST)
    sz = configs.BASE["sig"][2]
    nbins = 100
    tr.set_wake(nbins, 1, -5 * sz, 5 * sz, BBR["rs"], BBR["q"], BBR["fr"])
    tr.set_synrad(True, 2, 4)
    ot = O.OracleTrain(ring, secs[0], counts, ne, gap, cfg, ION_TEST["seed"], nbins=nbins, mode=1,
                       grid=(-5 * sz, 5 * sz), bbr=BBR, macro_charge=list(q), synrad=True, noise_seed=4)
    for t in range(3):
        tr.turn(t)
        ot.turn(bunches, t)
        got = tr.ion_passages()
        np.testing.assert_allclose(got[:, 0, 5], ot.last_cloud[:, 0, 0], rtol=1e-12)
        np.testing.assert_allclose(got[:, 0, 3:5], ot.last_cloud[:, 0, 3:5], rtol=1e-8)
    got = tr.download()
    compare_coords(got, bunches, off, 1e-11)
    cloud = tr.ion_cloud(0)
    assert len(cloud["x"]) == ot.cloud.c.n[0]
    tr.close()


@pytest.mark.parametrize("nb", [69, 64])
def test_ion_segments_match_oracle(nb):
    """A train longer than one ion segment (32 passages, the library's
    kIonSegment): the cloud is compacted between segments, in slot order, so
    passages, cloud and kicks still follow the oracle's bunch-by-bunch walk;
    two turns (the second starts from the first's cloud).  69 bunches: three
    segments (an odd number of buffer swaps per turn), 64: two (even).

    The bunches carry a tenth of the nominal intensity: at full intensity
    the ions' motion over a long train is chaotic (on the oracle, a 1e-15
    relative change of one bunch size moves the mean y seen by passage 36
    by 1e-3), so rounding differences would swamp the comparison; at a
    tenth the same change stays below 4e-14 over a 37-passage train."""
    counts = [900 + 37 * (b % 7) for b in range(nb)]
    ring, secs, counts, gap, q, ne, cfg, tr, bunches, off = setup_train(counts, 12, ions=ION_TEST, ne_scale=0.1)
    ot = O.OracleTrain(ring, secs[0], counts, ne, gap, cfg, ION_TEST["seed"])
    print(f"ion segments nb={nb}", flush=True)
    for t in range(2):
        tr.turn(t)
        ot.turn(bunches, t)
        got = tr.ion_passages()
        np.testing.assert_allclose(got[:, 0, 5], ot.last_cloud[:, 0, 0], rtol=1e-12)
        np.testing.assert_allclose(got[:, 0, 1:5], ot.last_cloud[:, 0, 1:5], rtol=1e-9, atol=1e-18)
    # ions survive from segment to segment (the comparison covers carried
    # ions); the train's closing gap empties most of the cloud
    assert got[40, 0, 0] > 2 * ION_TEST["n_new"]
    cloud = tr.ion_cloud(0)
    ec = ot.cloud.species(0)
    assert len(cloud["x"]) == len(ec["x"])
    for k in ("x", "y", "vx", "vy"):
        np.testing.assert_allclose(cloud[k], ec[k], rtol=1e-9, atol=1e-9 * np.max(np.abs(ec[k])))
    compare_coords(tr.download(), bunches, off, 1e-11)
    tr.close()


def test_ion_overflow_status():
    """A cloud that would pass max_ions ends the turn with CTS_ION_OVERFLOW,
    as the reference exits (src/LatticeInterActionPoint.cpp:1041-1049)."""
    counts = [1000, 1000, 1000]
    ions = dict(ION_TEST, max_ions=500, n_new=300)
    ring, secs, counts, gap, q, ne, cfg, tr, bunches, off = setup_train(counts, 9, ions=ions)
    tr.ion_phase(0)
    with pytest.raises(capi.CtsError) as e:
        tr.sync()
    assert e.value.code == capi.CTS_ION_OVERFLOW
    tr.close()


def test_growth_rates_match_oracle():
    """Coupled-bunch growth rates (MPBeam::GetCBMGR's ideal modes) fitted to
    the device's turn-by-turn centroids against the same fit on the oracle's
    centroids: within 1%."""
    from cetasim_amd import capi as C
    counts = [1800, 2600, 1500, 2200, 3000, 1700, 2400, 2000]
    nb = len(counts)
    p = configs.BASE
    ring, secs, der = configs.lattice(p)
    ions = dict(ION_TEST, pressure=[2.0e-6], n_new=200)
    ring, secs, counts, gap, q, ne, cfg, tr, bunches, off = setup_train(counts, 21, ions=ions)
    turns = 120
    tr.set_history(turns)
    ot = O.OracleTrain(ring, secs[0], counts, ne, gap, cfg, ions["seed"])
    omean = np.zeros((turns, nb, 6))
    for t in range(turns):
        tr.turn(t)
        ot.turn(bunches, t)
        for b in range(nb):
            omean[t, b] = O.rms(bunches[b], secs[0]).mean[:]
    rates = np.empty(3 * nb)
    capi.check(tr.lib.cts_growth_rates(tr.h, 0, 1, turns, 20, 120, der.workqz, der.beta_z, capi.dptr(rates)))
    rates = rates.reshape(3, nb)
    # the same fit on the oracle's centroids
    m = (C.Moments * (turns * nb))()
    for t in range(turns):
        for b in range(nb):
            m[t * nb + b].mean[:] = list(omean[t, b])
    harm = np.concatenate([[0], np.cumsum(gap[:-1])]).astype(np.int32)
    eref = np.empty(3 * nb)
    assert tr.lib.cts_cbm_growth_from_moments(
        m, turns, nb, capi.i32ptr(harm), ring.harmonic, capi.dptr(np.array([ring.nux, ring.nuy, der.workqz])),
        capi.dptr(np.array([secs[0].alpha_x, secs[0].beta_x, secs[0].alpha_y, secs[0].beta_y, der.beta_z])),
        1, 20, 120, ring.t0, capi.dptr(eref), None) == 0
    eref = eref.reshape(3, nb)
    for pl in range(2):
        scale = np.max(np.abs(eref[pl]))
        assert scale > 0
        tol = 0.01 * np.maximum(np.abs(eref[pl]), 0.1 * scale)
        assert np.all(np.abs(rates[pl] - eref[pl]) <= tol), (pl, rates[pl], eref[pl])
    tr.close()
