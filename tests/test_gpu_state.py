"""Handle state across re-layouts and restarts (MI355X).

* Host-supplied excitation normals (noise mode 1) stay usable after
  cts_set_bunches changes the particle count: the buffer follows the new
  layout and the synrad seam still matches the oracle.
* The long-range wake, the broadband impedance (both methods) and the beam-loaded cavities
  stay switched on across cts_set_bunches (set up again for the new bunch
  layout), giving the same turn as a handle configured from scratch.
* cts_upload starts a new run: the long-range centroid history and any
  pending long-range kick are cleared, so the first turn after a re-upload
  equals the first turn of a fresh handle, bit for bit.
"""
import ctypes

import numpy as np
import pytest

from cetasim_amd import configs
from cetasim_amd.tracker import Tracker

import oracle as O
from test_gpu_lrwake import lrw_cfg
from test_gpu_train import setup_train

pytestmark = pytest.mark.gpu

COORDS = O.COORDS


def test_host_noise_after_new_layout():
    ring, secs, _ = configs.lattice()
    tr = Tracker(ring, secs)
    q, ne = configs.bunch_scalars([5000], ring)
    tr.set_bunches([5000], q, ne)
    tr.set_synrad(True, 1, 0)
    n = 12_345
    q, ne = configs.bunch_scalars([n], ring)
    tr.set_bunches([n], q, ne)
    parts = configs.gaussian_bunch(n, 41)
    tr.upload(parts)
    noise = np.random.default_rng(42).standard_normal((5, n))
    tr.push_noise(noise)
    tr.seam("synrad", 3)
    got = tr.download()
    ob = O.Bunch(parts)
    O.lib().orc_synrad(ctypes.byref(ob.s), O.p(np.array(secs[0].b1h1[:])), O.p(np.array(secs[0].h1b1[:])),
                       O.p(np.array(ring.synrad_tau[:])), O.p(np.array(ring.emit_nat[:])),
                       O.p(np.ascontiguousarray(noise).ravel()))
    for c in COORDS:
        scale = np.max(np.abs(ob[c]))
        assert np.max(np.abs(got[c] - ob[c])) <= 1e-12 * scale, c
    tr.close()


def _one_turn(tr):
    tr.turn(0)
    got = tr.download()
    return got, tr.lrwake_kick()


def test_lrwake_survives_set_bunches_and_upload_restarts():
    counts = [900, 1300, 700]
    ring, secs, counts, gap, q, ne, _, tr, bunches, off = setup_train(counts, 17)
    parts = {c: np.concatenate([b[c] for b in bunches]).copy() for c in COORDS}
    c = lrw_cfg(secs, nturns=3)
    # configured before a re-layout to the same counts: must still be on
    tr.set_lrwake(c["nturns"], c["sectors"], c["res"], c["beta_x"], c["beta_y"])
    tr.set_bunches(counts, q, ne, gap)
    tr.upload(parts)
    # a run of a few turns fills the centroid history and leaves a kick pending
    for t in range(4):
        tr.turn(t)
    assert np.any(tr.lrwake_kick() != 0)
    # a restart: the first turn must equal a fresh handle's first turn
    tr.upload(parts)
    got, kick = _one_turn(tr)

    ring2, secs2, _, _, _, _, _, fresh, _, _ = setup_train(counts, 17)
    fresh.set_lrwake(c["nturns"], c["sectors"], c["res"], c["beta_x"], c["beta_y"])
    ref, kref = _one_turn(fresh)
    assert np.any(kref != 0)
    assert np.array_equal(kick, kref)
    for cc in COORDS:
        assert np.array_equal(got[cc], ref[cc]), cc
    tr.close()
    fresh.close()


def test_bbimp_survives_set_bunches():
    from test_gpu_bbimp import impedance
    counts = [2000, 3100]
    zre, zim, fmax = impedance()
    ring, secs, counts, gap, q, ne, _, tr, bunches, off = setup_train(counts, 19)
    parts = {c: np.concatenate([b[c] for b in bunches]).copy() for c in COORDS}
    tr.set_bbimp(zre, zim, fmax)
    tr.set_bunches(counts, q, ne, gap)
    tr.upload(parts)
    tr.turn(0)
    got = tr.download()
    wake, prof = tr.bbimp_wake()

    _, _, _, _, _, _, _, fresh, _, _ = setup_train(counts, 19)
    fresh.set_bbimp(zre, zim, fmax)
    fresh.turn(0)
    ref = fresh.download()
    wref, pref = fresh.bbimp_wake()
    assert np.max(np.abs(wref)) > 0
    assert np.array_equal(wake, wref) and np.array_equal(prof, pref)
    for c in COORDS:
        assert np.array_equal(got[c], ref[c]), c
    tr.close()
    fresh.close()


def test_time_domain_bbimp_survives_set_bunches():
    """The same for the time-domain impedance (its wake table is kept)."""
    from test_gpu_bbimp import td_table
    counts = [2000, 3100]
    pos, wakes = td_table()
    ring, secs, counts, gap, q, ne, _, tr, bunches, off = setup_train(counts, 19)
    parts = {c: np.concatenate([b[c] for b in bunches]).copy() for c in COORDS}
    for b in range(len(counts)):
        parts["x"][off[b]:off[b + 1]] += 3e-5 * (b + 1)
    tr.set_bbimp_td(pos, wakes, beta=(5.0, 3.0))
    tr.set_bunches(counts, q, ne, gap)
    tr.upload(parts)
    tr.turn(0)
    got = tr.download()
    wake, prof = tr.bbimp_wake()

    _, _, _, _, _, _, _, fresh, _, _ = setup_train(counts, 19)
    fresh.upload(parts)
    fresh.set_bbimp_td(pos, wakes, beta=(5.0, 3.0))
    fresh.turn(0)
    ref = fresh.download()
    wref, pref = fresh.bbimp_wake()
    assert wref.shape[2] == 1199 and all(np.max(np.abs(wref[k])) > 0 for k in range(5))
    assert np.array_equal(wake, wref) and np.array_equal(prof, pref)
    for c in COORDS:
        assert np.array_equal(got[c], ref[c]), c
    tr.close()
    fresh.close()
