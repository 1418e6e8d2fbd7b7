"""What the trained field surrogate gains over the exact FP64 Faddeeva path
it replaces (src/Bunch.cpp:119-211), on the same train (MI355X).

The surrogate is trained as the benchmark trains it (4e6 samples of the
FP64 path, seed 4) for 100000 Adam steps, then the same 8-bunch train
(1500-3000 macroparticles per bunch, CO ions at 2e-6 Pa, so that the
ion-driven coupled-bunch motion grows over the run) is tracked for 120
turns three times: exact field both ways (field model 0), the surrogate for
the electrons' kick from the cloud (2), and the surrogate both ways (3).

Reported, as one JSON line: the network's field error over 2^20
validation points of its domain (max and rms of |G_sur - G| / |G|), the
coupled-bunch growth rates of the x and y planes fitted over turns 20-120
(cts_growth_rates: GetCBMGR / FitALinear), their deviation from the exact
path's, and each path's time per turn in graph replays.

Growth-rate deviation of a plane: max over the modes of
|r_sur - r_exact| / max |r_exact| (a mode whose rate is near zero has no
meaningful relative error; the largest rate is the instability's).

Asserted: the field error the training reaches (max <= 2e-2, rms <= 1e-3),
the electron-side growth-rate deviation <= 1% in both planes, finite rates
on both sides.  The ion side (mode 3) is reported, not bounded: the ions'
motion in the beam's field is chaotic, and field errors of ~1e-3 change the
y-plane rates by several percent.  The time ratio is reported as measured;
the network costs more per point than the exact rational w(z), so it is
below 1.
"""
import json
import time

import numpy as np
import pytest

from cetasim_amd import capi, configs
from cetasim_amd.surrogate import Surrogate
from cetasim_amd.tracker import Tracker

pytestmark = pytest.mark.gpu

TURNS = 120
STEPS = 100_000


def trained():
    s = Surrogate()
    s.generate(4_000_000, 1 << 20, 4)
    s.init_params(4)
    k = STEPS // 10
    for steps, lmax, lmin, sd in ((2 * k, 3e-3, 1e-3, 4), (4 * k, 1e-3, 1e-4, 5), (4 * k, 1e-4, 1e-6, 6)):
        s.train(steps, nwg=512, chunks=2, lr_max=lmax, lr_min=lmin, seed=sd)
    return s


def track(sur, mode):
    p = configs.BASE
    ring, secs, der = configs.lattice(p)
    counts = np.array([1800, 2600, 1500, 2200, 3000, 1700, 2400, 2000])
    nb = len(counts)
    gap = configs.train_gaps(nb, 2, ring.harmonic)
    q, ne = configs.bunch_scalars(counts, ring, p)
    parts = configs.train_bunches(counts, 21, p)
    ions = dict(mass_number=[28.0], cross_section=[2.0e-22], pressure=[2.0e-6], temperature=300.0, n_new=200,
                max_ions=50_000, loss_boundary=10.0, seed=11)
    tr = Tracker(ring, secs)
    tr.set_bunches(counts, q, ne, gap)
    tr.upload(parts)
    tr.set_ions(configs.ion_config(ring, ions, p))
    if mode:
        tr.set_field_model(sur, mode)
    tr.set_history(TURNS)
    tr.sync()
    t0 = time.perf_counter()
    tr.run_turns(0, TURNS)
    tr.sync()
    secs_per_turn = (time.perf_counter() - t0) / TURNS
    rates = np.empty(3 * nb)
    capi.check(tr.lib.cts_growth_rates(tr.h, 0, 1, TURNS, 20, TURNS, der.workqz, der.beta_z, capi.dptr(rates)))
    tr.close()
    return rates.reshape(3, nb), secs_per_turn


def deviation(r, r0):
    return [float(np.max(np.abs(r[k] - r0[k])) / np.max(np.abs(r0[k]))) for k in (0, 1)]


def test_surrogate_against_exact_on_train():
    print("surrogate gain: training 100000 steps on 4e6 samples", flush=True)
    s = trained()
    mx, rms = s.validate()
    print(f"surrogate gain: field error max {mx:.3g} rms {rms:.3g}", flush=True)
    r0, t0 = track(None, 0)
    r2, t2 = track(s, 2)
    r3, t3 = track(s, 3)
    s.close()
    dev2, dev3 = deviation(r2, r0), deviation(r3, r0)
    print(json.dumps({"field_err_max": mx, "field_err_rms": rms, "rates_exact_xy": r0[:2].tolist(),
                      "growth_dev_electron_side_xy": dev2, "growth_dev_both_sides_xy": dev3,
                      "ms_per_turn": {"exact": 1e3 * t0, "electron_side": 1e3 * t2, "both_sides": 1e3 * t3},
                      "rate_ratio_electron_side": t0 / t2, "rate_ratio_both_sides": t0 / t3}), flush=True)
    assert mx <= 2e-2 and rms <= 1e-3, (mx, rms)
    assert np.all(np.isfinite(r2)) and np.all(np.isfinite(r3))
    # the train's motion grows: the fitted rates are not noise
    assert np.max(np.abs(r0[0])) > 1e3 and np.max(np.abs(r0[1])) > 1e3
    assert max(dev2) <= 0.01, dev2
    assert t0 > 0 and t2 > 0 and t3 > 0
