"""The long-range wake on the HIP path against the oracle: the wake table
(resistive wall of round and elliptic sectors plus broadband resonators),
and bunch trains tracked with it, with and without a short-range wake, over
several turns (the centroid history fills up).  Tolerances as the other
train tests: coordinates 1e-11 of each coordinate's scale, kicks 1e-10 of
the largest kick; the table 1e-13 relative (both sides evaluate the same
closed forms on the host)."""
import numpy as np
import pytest

from cetasim_amd import configs

import oracle as O
from test_gpu_train import BBR, compare_coords, setup_train

pytestmark = pytest.mark.gpu

SECTORS = [[0.01, 0.01, 2304.0, 2.7, 4.25, 1, 3.5e5],      # round (Yokoya factors 1)
           [0.03, 0.012, 300.0, 5.0, 8.0, 2, 1.4e6]]      # elliptic
RES = [[2e4, 1.5, 1.1e9, 3e6, 2.0, 0.9e9, 5e6, 2.5, 1.05e9]]


def lrw_cfg(secs, nturns=4):
    return dict(nturns=nturns, sectors=SECTORS, res=RES, beta_x=secs[0].beta_x, beta_y=secs[0].beta_y)


def test_table_matches_oracle():
    counts = [500, 700]
    ring, secs, counts, gap, q, ne, cfg, tr, bunches, off = setup_train(counts, 3)
    c = lrw_cfg(secs, nturns=3)
    tr.set_lrwake(c["nturns"], c["sectors"], c["res"], c["beta_x"], c["beta_y"])
    t, w, rw, bbr = tr.lrwake_table()
    ot = O.OracleTrain(ring, secs[0], counts, ne, gap, None, 1, lrw=c)
    assert len(t) == ot.lrw["np"] == 2 * 3 * ring.harmonic + 2
    np.testing.assert_allclose(t, ot.lrw["time"], rtol=1e-15)
    ow = ot.lrw["w"].reshape(3, -1)
    for k in range(3):
        np.testing.assert_allclose(w[k], ow[k], rtol=1e-13, atol=1e-13 * np.max(np.abs(ow[k])))
    np.testing.assert_allclose(w, rw + bbr, rtol=1e-15)
    tr.close()


@pytest.mark.parametrize("sr_wake", [False, True])
def test_lrwake_turns_match_oracle(sr_wake):
    counts = [1500 + 211 * (b % 5) for b in range(12)]
    ring, secs, counts, gap, q, ne, cfg, tr, bunches, off = setup_train(counts, 31)
    c = lrw_cfg(secs)
    tr.set_lrwake(c["nturns"], c["sectors"], c["res"], c["beta_x"], c["beta_y"])
    tr.set_synrad(True, 2, 4)
    kw = {}
    if sr_wake:
        sz = configs.BASE["sig"][2]
        tr.set_wake(100, 1, -5 * sz, 5 * sz, BBR["rs"], BBR["q"], BBR["fr"])
        kw = dict(nbins=100, mode=1, grid=(-5 * sz, 5 * sz), bbr=BBR, macro_charge=list(q))
    ot = O.OracleTrain(ring, secs[0], counts, ne, gap, None, 1, lrw=c, synrad=True, noise_seed=4, **kw)
    print(f"long-range wake train, short-range wake {sr_wake}", flush=True)
    px0 = np.concatenate([b["px"] for b in bunches]).copy()
    for t in range(6):
        tr.turn(t)
        ot.turn(bunches, t)
        k = tr.lrwake_kick()
        ok = ot.last_lrw_kick
        for pl in range(3):
            scale = np.max(np.abs(ok[:, pl]))
            assert scale > 0
            np.testing.assert_allclose(k[:, pl], ok[:, pl], rtol=0, atol=1e-10 * scale)
    got = tr.download()
    assert not np.any(got["lost"])
    compare_coords(got, bunches, off, 1e-11)
    # the turn's moment record includes the kick (px, py means)
    m = tr.moments()
    for b, ob in enumerate(bunches):
        r = O.rms(ob, secs[0])
        np.testing.assert_allclose(m[b]["mean"][1], r.mean[1], rtol=1e-9, atol=1e-12 * abs(r.mean[1]) + 1e-18)
    # the wake is not negligible against the tolerance
    assert np.max(np.abs(ok[:, 0])) > 1e-6 * np.max(np.abs(px0))
    tr.close()
