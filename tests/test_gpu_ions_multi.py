"""Beam-ion interaction at every interaction point of a three-section ring
(MI355X), against the oracle.

Each section k has its own twiss functions at its start, its own gas
pressure, temperature, length, aperture and ion stream, and its own cloud.
A turn runs, as MPBeam::Run does (src/MPBeam.cpp:1059-1070): for k = 0, 1,
2 the ion phase at point k on the moments after map k - 1 (bunch sizes
sqrt(emittance x beta_k)), then section map k; then RF, drift, energy loss
and loss marking.  Two turns: every point's passage moments (1e-9), its
cloud (1e-9) and the coordinates (1e-11 of each coordinate's scale) against
the oracle walking the same sequence.  The bunches carry a tenth of the
nominal intensity (the ions' chaos at full intensity, see
test_ion_segments_match_oracle).
"""
import ctypes
import math

import numpy as np
import pytest

from cetasim_amd import capi, configs
from cetasim_amd.tracker import Tracker

import oracle as O
from test_gpu_train import ION_TEST, compare_coords

pytestmark = pytest.mark.gpu

COORDS = O.COORDS


def three_sections(der):
    lib = capi.load()
    tw = [capi.Twiss(beta_x=10.0, alpha_x=0.0, eta_x=0.0, etap_x=0.0, beta_y=10.0, alpha_y=0.0, eta_y=0.0,
                     etap_y=0.0, beta_z=der.beta_z),
          capi.Twiss(beta_x=15.0, alpha_x=0.5, eta_x=0.0, etap_x=0.0, beta_y=6.0, alpha_y=-0.3, eta_y=0.0,
                     etap_y=0.0, beta_z=der.beta_z),
          capi.Twiss(beta_x=7.0, alpha_x=-0.2, eta_x=0.0, etap_x=0.0, beta_y=14.0, alpha_y=0.4, eta_y=0.0,
                     etap_y=0.0, beta_z=der.beta_z)]
    mux = [0.05, 0.07, 0.06]  # sums to the tunes 0.18 / 0.27
    muy = [0.09, 0.10, 0.08]
    secs = []
    for k in range(3):
        sec = capi.Section()
        capi.check(lib.cts_build_section(tw[k], tw[(k + 1) % 3], 2 * math.pi * mux[k], 2 * math.pi * muy[k], sec))
        secs.append(sec)
    return secs


def ion_cfgs(ring):
    cfgs = []
    for k, (pressure, temp, length, ap) in enumerate(((1.0e-5, 300.0, 200.0, (0.02, 0.01)),
                                                      (2.0e-5, 350.0, 150.0, (0.015, 0.012)),
                                                      (0.6e-5, 280.0, 249.585, (0.02, 0.01)))):
        ions = dict(ION_TEST, pressure=[pressure], temperature=temp, length=length, seed=11 + 7 * k)
        c = configs.ion_config(ring, ions)
        c.ap_x, c.ap_y = ap
        cfgs.append(c)
    return cfgs


def test_ions_at_three_interaction_points():
    p = configs.BASE
    ring, _, der = configs.lattice(p)
    secs = three_sections(der)
    counts = np.array([1500, 2200, 1800, 2500], np.int64)
    nb = len(counts)
    gap = configs.train_gaps(nb, 2, ring.harmonic)
    q, ne = configs.bunch_scalars(counts, ring, p)
    ne = ne * 0.1
    parts = configs.train_bunches(counts, 41, p)
    tr = Tracker(ring, secs)
    tr.set_bunches(counts, q, ne, gap)
    tr.upload(parts)
    cfgs = ion_cfgs(ring)
    for k, c in enumerate(cfgs):
        tr.set_ions(c, ip=k)
    off = np.concatenate([[0], np.cumsum(counts)])
    bunches = [O.Bunch({c: parts[c][off[b]:off[b + 1]] for c in COORDS}) for b in range(nb)]
    # one oracle train per interaction point: its cloud, its section's twiss
    ots = [O.OracleTrain(ring, secs[k], counts, ne, gap, cfgs[k], cfgs[k].seed) for k in range(3)]
    maps = [O.make_map(secs[k], ring) for k in range(3)]
    L = O.lib()
    for t in range(2):
        print(f"three interaction points: turn {t}", flush=True)
        for k in range(3):
            tr.ion_phase(t, ip=k)
            tr.seam("section_map", k)
            ots[k].ion_phase(bunches, t)
            for b in bunches:
                L.orc_section_map(ctypes.byref(b.s), ctypes.byref(maps[k]))
        tr.seam("rf_kick")
        tr.seam("drift")
        tr.seam("energy_loss")
        tr.mark_lost()
        for b in bunches:
            for r in range(ring.n_rf):
                L.orc_rf_kick(ctypes.byref(b.s), ring.rf_vr[r], ring.rf_vi[r], ring.rf_harm[r], ring.f0 * ring.harmonic,
                              ring.rbeta, ring.energy)
            L.orc_drift(ctypes.byref(b.s), ring.circ, O.p(np.array(ring.alphac[:])))
            L.orc_energy_loss(ctypes.byref(b.s), ring.u0, ring.energy, ring.rbeta)
            cnt = O.I64()
            L.orc_mark_lost(ctypes.byref(b.s), ring.t0, ring.harmonic, ring.ap_x, ring.ap_y, ctypes.byref(cnt))
        for k in range(3):
            got = tr.ion_passages(ip=k)
            exp = ots[k].last_cloud
            np.testing.assert_allclose(got[:, 0, 5], exp[:, 0, 0], rtol=1e-12)
            np.testing.assert_allclose(got[:, 0, 1:5], exp[:, 0, 1:5], rtol=1e-9, atol=1e-18)
    # the points see different clouds (the settings differ)
    n_ion = [tr.ion_passages(ip=k)[-1, 0, 5] for k in range(3)]
    assert len(set(n_ion)) == 3
    for k in range(3):
        cloud = tr.ion_cloud(0, ip=k)
        ec = ots[k].cloud.species(0)
        assert len(cloud["x"]) == len(ec["x"]) > 0
        for c in ("x", "y", "vx", "vy"):
            np.testing.assert_allclose(cloud[c], ec[c], rtol=1e-9, atol=1e-9 * np.max(np.abs(ec[c])))
    compare_coords(tr.download(), bunches, off, 1e-11)
    tr.close()
