"""Pins the oracle's ion passage (orc_ion_passage, oracle/ions.c) to the
physics it restates, independently of its own formulas:
src/LatticeInterActionPoint.cpp:974-1210 and src/MPBunch.cpp:1094-1210.

* Production: one passage through an empty cloud creates
  sigma P L N_e / (k_B T) ions (gas density P / (k_B T) times the ionisation
  cross-section times the bunch population, over the interaction length).
* Ion kick: an ion at rest far from the bunch (r >> sigma) receives the
  impulse of a line charge of N_e electrons, dv = -2 N_e r_e c (m_e / A m_u)
  (x, y) / r^2 (the 2-D field of a line charge integrated over the passage,
  with r_e = e^2 / (4 pi eps0 m_e c^2)); a round bunch gives the exact
  Gaussian-charge field (1 - exp(-r^2 / 2 sigma^2)) at any r.
* Drift: the ion then moves by dv times the time to the next bunch,
  C / h * gap / c.
* Loss: ions beyond min(L x largest bunch size, aperture) are removed.
* Electron kick: an electron far from a cloud of N ions gets
  dp = -2 N r_e / gamma (x, y) / r^2; lost electrons keep their stale force
  (BunchTransferDueToIon applies it to every particle, src/Bunch.cpp:213-220).

The constants are the reference's (include/Global.h:26-36); r_e is also
checked against e, eps0, m_e and c from the same header.
"""
import ctypes
import math
from types import SimpleNamespace

import numpy as np
import pytest

import oracle as O

L = O.lib()

# include/Global.h:26-36
EPS0 = 8.854187817e-12
C = 299792458.0
E_CHARGE = 1.602017733e-19
E_MASS = 9.10938215e-31
E_MASS_EV = 0.51099906e6
U_MASS_EV = 931.49432e6
R_E = 2.81794092e-15
K_B = 1.3806505e-23

CIRC, HARM = 600.0, 1000


def cloud_cfg(**over):
    c = dict(nspec=1, max_ions=100_000, n_new=50, mass_number=[28.0, 0, 0, 0], cross_section=[2.0e-22, 0, 0, 0],
             pressure=[1.0e-7, 0, 0, 0], temperature=300.0, length=CIRC, loss_boundary=1e6, ap_x=10.0, ap_y=10.0)
    c.update(over)
    return SimpleNamespace(**c)


def bunch(xs, ys, lost=None):
    n = len(xs)
    parts = {k: np.zeros(n) for k in O.COORDS}
    parts["x"][:] = xs
    parts["y"][:] = ys
    return O.Bunch(parts, lost)


def passage(cloud, b, view, smax=(1.0, 1.0), seed=7, turn=0, pas=0):
    out = np.zeros(5 * cloud.nspec)
    rc = L.orc_ion_passage(ctypes.byref(cloud.c), ctypes.byref(b.s), ctypes.byref(view), smax[0], smax[1], seed,
                           turn, pas, O.p(out))
    assert rc == 0
    return out


def put_ion(cloud, xs, ys, sp=0):
    n0 = cloud.c.n[sp]
    for i, (x, y) in enumerate(zip(xs, ys)):
        cloud.arr[sp]["x"][n0 + i] = x
        cloud.arr[sp]["y"][n0 + i] = y
        for k in ("vx", "vy", "fx", "fy"):
            cloud.arr[sp][k][n0 + i] = 0.0
    cloud.c.n[sp] = n0 + len(xs)


def test_classical_radius_from_header_constants():
    re = E_CHARGE ** 2 / (4 * math.pi * EPS0 * E_MASS * C ** 2)
    # the header's e (1.602017733e-19) is not CODATA's, hence 2e-4
    assert re == pytest.approx(R_E, rel=5e-4)


def test_production_count():
    cfg = cloud_cfg()
    cloud = O.IonCloud(cfg, CIRC, HARM, 1)
    ne = 6.24e9
    view = O.OrcBunchView(0.0, 0.0, 2e-5, 4e-6, ne, 1, 1.0e4)
    out = passage(cloud, bunch([0.0], [0.0]), view)
    n_ions = cfg.cross_section[0] * cfg.pressure[0] / (K_B * cfg.temperature) * ne * cfg.length
    assert out[0] == pytest.approx(n_ions, rel=1e-14)
    assert cloud.c.macro_charge[0] == pytest.approx(n_ions / cfg.n_new, rel=1e-14)
    assert cloud.c.n[0] == cfg.n_new


@pytest.mark.parametrize("round_beam", [False, True])
def test_ion_kick_drift_far_field(round_beam):
    """Far ions against the line-charge impulse; new ions are generated at
    the bunch centre too (they stay near it)."""
    cfg = cloud_cfg(n_new=4, mass_number=[18.0, 0, 0, 0])
    cloud = O.IonCloud(cfg, CIRC, HARM, 1)
    far = [(3e-3, 1e-3), (-2e-3, 2.5e-3), (4e-3, -4e-3)]
    put_ion(cloud, [p[0] for p in far], [p[1] for p in far])
    ne, gap = 1.0e10, 3
    sx, sy = (2e-5, 2e-5) if round_beam else (2e-5, 4e-6)
    view = O.OrcBunchView(0.0, 0.0, sx, sy, ne, gap, 1.0e4)
    passage(cloud, bunch([0.0], [0.0]), view)
    coef = 2 * ne * R_E * C * E_MASS_EV / (cfg.mass_number[0] * U_MASS_EV)
    t_gap = CIRC / HARM * gap / C
    sp = cloud.species(0)
    for i, (x, y) in enumerate(far):
        r2 = x * x + y * y
        shape = 1.0 - math.exp(-r2 / (2 * sx * sx)) if round_beam else 1.0
        dvx, dvy = -coef * x / r2 * shape, -coef * y / r2 * shape
        # far field: Bassetti-Erskine -> line charge to O(sigma^2 / r^2)
        tol = 1e-13 if round_beam else 1e-4
        assert sp["vx"][i] == pytest.approx(dvx, rel=tol)
        assert sp["vy"][i] == pytest.approx(dvy, rel=tol)
        # the drift over the time to the next bunch, with the ion's velocity
        assert sp["x"][i] == pytest.approx(x + sp["vx"][i] * t_gap, rel=1e-14)
        assert sp["y"][i] == pytest.approx(y + sp["vy"][i] * t_gap, rel=1e-14)


def test_round_beam_field_at_sigma():
    """Round bunch, ion at r = sigma: the Gaussian charge field exactly."""
    cfg = cloud_cfg(n_new=1)
    cloud = O.IonCloud(cfg, CIRC, HARM, 1)
    s = 3e-5
    x, y = 0.6 * s, 0.8 * s
    put_ion(cloud, [x + 1e-4], [y - 2e-4])  # bunch centre at (1e-4, -2e-4)
    ne = 3.0e10
    view = O.OrcBunchView(1e-4, -2e-4, s, s, ne, 1, 1.0e4)
    passage(cloud, bunch([0.0], [0.0]), view)
    coef = 2 * ne * R_E * C * E_MASS_EV / (cfg.mass_number[0] * U_MASS_EV)
    r2 = x * x + y * y
    g = (1 - math.exp(-r2 / (2 * s * s))) / r2
    sp = cloud.species(0)
    assert sp["vx"][0] == pytest.approx(-coef * x * g, rel=1e-13)
    assert sp["vy"][0] == pytest.approx(-coef * y * g, rel=1e-13)


def test_ion_loss_boundary():
    cfg = cloud_cfg(n_new=2, loss_boundary=10.0, ap_x=5e-4, ap_y=1.0)
    cloud = O.IonCloud(cfg, CIRC, HARM, 1)
    # boundary x: min(10 * 1e-4, 5e-4) = 5e-4 (aperture); y: min(10 * 2e-5, 1) = 2e-4
    xs = [4.0e-4, 6.0e-4, -6.0e-4, 1.0e-4, 1.0e-4, 1.0e-4]
    ys = [0.0, 0.0, 0.0, 1.5e-4, 2.5e-4, -2.5e-4]
    put_ion(cloud, xs, ys)
    # a tiny bunch population: the kicks do not move the ions across the edges
    view = O.OrcBunchView(0.0, 0.0, 1e-6, 1e-6, 1.0, 1, 1.0e4)
    passage(cloud, bunch([0.0], [0.0]), view, smax=(1e-4, 2e-5))
    sp = cloud.species(0)
    kept = {(round(x, 12), round(y, 12)) for x, y in zip(sp["x"], sp["y"])}
    assert (round(4e-4, 12), 0.0) in kept
    assert (round(1e-4, 12), round(1.5e-4, 12)) in kept
    for x, y in zip(xs, ys):
        if abs(x) > 5e-4 or abs(y) > 2e-4:
            assert (round(x, 12), round(y, 12)) not in kept
    # the two new ions (near the centre) and the two inside ones survive
    assert cloud.c.n[0] == 4


def test_electron_kick_from_cloud():
    cfg = cloud_cfg(n_new=2000)
    cloud = O.IonCloud(cfg, CIRC, HARM, 1)
    ne, gamma = 6.0e9, 11742.0
    view = O.OrcBunchView(2e-6, -1e-6, 2e-5, 4e-6, ne, 1, gamma)
    xs, ys = [8e-3, -5e-3, 1e-2], [3e-3, 6e-3, -7e-3]
    lost = np.array([0, 0, 1], np.int32)
    b = bunch(xs, ys, lost)
    b.fx[2], b.fy[2] = 1e-9, -2e-9  # stale force of a lost particle
    out = passage(cloud, b, view)
    n_ion, mx, my, rx, ry = out[:5]
    assert n_ion == pytest.approx(cfg.cross_section[0] * cfg.pressure[0] / (K_B * cfg.temperature) * ne * cfg.length,
                                  rel=1e-14)
    # the cloud moments are those of the generated ions (before their drift)
    gx, gy = np.zeros(cfg.n_new), np.zeros(cfg.n_new)
    L.orc_ion_generate(7, 0, 0, 0, cfg.n_new, view.xmean, view.ymean, view.rms_x, view.rms_y, O.p(gx), O.p(gy))
    assert mx == pytest.approx(gx.mean(), rel=1e-12)
    assert rx == pytest.approx(gx.std(), rel=1e-12)
    assert ry == pytest.approx(gy.std(), rel=1e-12)
    for i in range(2):
        dx, dy = xs[i] - mx, ys[i] - my
        r2 = dx * dx + dy * dy
        k = 2 * n_ion * R_E / gamma
        assert b["px"][i] == pytest.approx(-k * dx / r2, rel=1e-5)
        assert b["py"][i] == pytest.approx(-k * dy / r2, rel=1e-5)
    assert b["px"][2] == 1e-9 and b["py"][2] == -2e-9
