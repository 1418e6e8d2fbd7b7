"""Synthetic workloads of the multi-particle tracking turn.

The ring is a single-section storage ring (one interaction point) with
h = 1000 at about 500 MHz (T0 = 2 us), 6 GeV, tunes 0.18/0.27, Qs = 0.005,
chromaticity 2/2, ADTS coefficients 1e3 1/m, alpha1 = 3e-5, alpha2 = 1e-4,
and a broadband resonator (Rs = 10 kOhm, fr = 10 GHz, Q = 1).  The RF
voltage is solved so that the synchrotron tune is 0.005 with the formula of
src/ReadInputSettings.cpp:1023-1027.  Bunches are 6-D Gaussians with
sigma_x = 20 um, sigma_y = 4 um, sigma_z = 3 mm, sigma_dp = 1e-3.
"""
import math

import numpy as np

from . import capi

C_LIGHT = 299792458.0
E_CHARGE = 1.602017733e-19
ME_EV = 0.51099906e6

BASE = dict(
    energy=6.0e9, harmonic=1000, t0=2.0e-6, u0=1.2e6, sdelta0=1.0e-3,
    alphac=(3.0e-5, 1.0e-4, 0.0), nux=0.18, nuy=0.27, qs=0.005,
    chrom=(2.0, 2.0), adtx=(1e3, 1e3, 0.0, 0.0, 0.0), adty=(1e3, 1e3, 0.0, 0.0, 0.0),
    skew_k=0.0, synrad_tau=(1.0e4, 1.0e4, 5.0e3), natural_emit=4.16e-11, coupling=0.04,
    ap_x=0.02, ap_y=0.01, beta_x=10.0, beta_y=10.0,
    sig=(20e-6, 4e-6, 3e-3, 1e-3), bunch_charge=1e-9,
    bbr=dict(rs=[1.0e4], q=[1.0], fr=[1.0e10]),
)

WORKLOADS = {
    # single bunch, 1e5 macroparticles, no noise, 1000 turns, 100 bins
    "c1": dict(n=[100_000], seed=1, nbins=100, turns=1000, synrad=False, noise_seed=None),
    # single bunch, 1e7 macroparticles, quantum excitation, 10k turns, 400 bins
    "c2": dict(n=[10_000_000], seed=2, nbins=400, turns=10_000, synrad=True, noise_seed=2),
}


def rf_voltage(p):
    """V such that sqrt(-eta*h*V*cos(phi_s)/(2*pi*beta*E)) = Qs, phi_s = pi - asin(U0/V)."""
    gamma = p["energy"] / ME_EV
    beta = math.sqrt(1.0 - gamma ** -2)
    eta = p["alphac"][0] - 1.0 / gamma ** 2
    k = p["qs"] ** 2 * 2 * math.pi * beta * p["energy"] / (eta * p["harmonic"])
    return math.sqrt(p["u0"] ** 2 + k ** 2)


def ring_input(p=BASE):
    gamma = p["energy"] / ME_EV
    beta = math.sqrt(1.0 - gamma ** -2)
    ri = capi.RingInput()
    ri.circ = C_LIGHT * beta * p["t0"]
    ri.energy = p["energy"]
    ri.u0 = p["u0"]
    ri.sdelta0 = p["sdelta0"]
    ri.alphac[:] = p["alphac"]
    ri.nux, ri.nuy = p["nux"], p["nuy"]
    ri.chrom[:] = p["chrom"]
    ri.adtx[:] = p["adtx"]
    ri.adty[:] = p["adty"]
    ri.skew_k = p["skew_k"]
    ri.synrad_tau[:] = p["synrad_tau"]
    ri.natural_emit = p["natural_emit"]
    ri.coupling = p["coupling"]
    ri.harmonic = p["harmonic"]
    ri.rf_volt = rf_voltage(p)
    ri.ap_x, ri.ap_y = p["ap_x"], p["ap_y"]
    return ri


def lattice(p=BASE):
    """(Ring, [Section], RingDerived) for the one-section ring."""
    lib = capi.load()
    ri = ring_input(p)
    ring = capi.Ring()
    der = capi.RingDerived()
    capi.check(lib.cts_derive_ring(ri, ring, der))
    tw = capi.Twiss(beta_x=p["beta_x"], alpha_x=0.0, eta_x=0.0, etap_x=0.0,
                    beta_y=p["beta_y"], alpha_y=0.0, eta_y=0.0, etap_y=0.0, beta_z=der.beta_z)
    sec = capi.Section()
    # one section: mu = 2*pi*nu - psi_last + psi_0 with psi = 0
    # (src/LatticeInterActionPoint.cpp:274-278)
    capi.check(lib.cts_build_section(tw, tw, 2 * math.pi * p["nux"], 2 * math.pi * p["nuy"], sec))
    return ring, [sec], der


def gaussian_bunch(n, seed, p=BASE):
    """6-D Gaussian macroparticles, zero means, px/py spreads sigma/beta."""
    rng = np.random.default_rng(seed)
    sx, sy, sz, sd = p["sig"]
    x = rng.standard_normal(n) * sx
    px = rng.standard_normal(n) * (sx / p["beta_x"])
    y = rng.standard_normal(n) * sy
    py = rng.standard_normal(n) * (sy / p["beta_y"])
    z = rng.standard_normal(n) * sz
    dp = rng.standard_normal(n) * sd
    return dict(x=x, px=px, y=y, py=py, z=z, dp=dp)


def bunch_scalars(counts, ring, p=BASE):
    """macroEleCharge and electronNumPerBunch of each bunch (src/Bunch.cpp:69-70)."""
    n_e = np.full(len(counts), p["bunch_charge"] / E_CHARGE)
    q = n_e / np.asarray(counts, dtype=np.float64)
    return q, n_e


# Batched train with residual-gas ions: 80 bunches 4 ns apart (2 buckets at
# h = 1000, T0 = 2 us), macroparticle counts uniform in [1e4, 2e5] (seed 3),
# 1 nC per bunch, one interaction point with CO (mass 28, cross section
# 2.0 Mb) at 1e-9 Pa and 300 K over the whole circumference, ion loss at
# 10 rms bunch sizes, at most 5e5 macro-ions.
TRAIN = dict(nb=80, spacing=2, count_range=(10_000, 200_000), seed=3, turns=2000,
             ions=dict(mass_number=[28.0], cross_section=[2.0e-22], pressure=[1.0e-9], temperature=300.0,
                       n_new=6000, max_ions=500_000, loss_boundary=10.0, seed=3))


def train_counts(nb=TRAIN["nb"], lo=TRAIN["count_range"][0], hi=TRAIN["count_range"][1], seed=TRAIN["seed"]):
    """Macroparticle count of every bunch, uniform integers in [lo, hi]."""
    return np.random.default_rng(seed).integers(lo, hi + 1, nb).astype(np.int64)


def train_gaps(nb, spacing, harmonic):
    """bunchGap of every bunch: buckets to the next bunch, the last one to the
    first bunch of the next turn (src/MPBeam.cpp:130-136)."""
    g = np.full(nb, spacing, dtype=np.int32)
    g[-1] = harmonic - spacing * (nb - 1)
    return g


def train_bunches(counts, seed, p=BASE):
    """Concatenated 6-D Gaussians, one independent stream per bunch."""
    parts = [gaussian_bunch(int(n), seed * 1000 + b, p) for b, n in enumerate(counts)]
    return {c: np.concatenate([q[c] for q in parts]) for c in ("x", "px", "y", "py", "z", "dp")}


def ion_config(ring, ions=TRAIN["ions"], p=BASE):
    """capi.IonConfig of one interaction point spanning the circumference."""
    cfg = capi.IonConfig()
    cfg.nspec = len(ions["mass_number"])
    cfg.n_new = ions["n_new"]
    cfg.max_ions = ions["max_ions"]
    for i in range(cfg.nspec):
        cfg.mass_number[i] = ions["mass_number"][i]
        cfg.cross_section[i] = ions["cross_section"][i]
        cfg.pressure[i] = ions["pressure"][i]
    cfg.temperature = ions["temperature"]
    cfg.length = ions.get("length", ring.circ)
    cfg.loss_boundary = ions["loss_boundary"]
    cfg.ap_x, cfg.ap_y = p["ap_x"], p["ap_y"]
    cfg.seed = ions["seed"]
    return cfg


# Large batched train: 480 bunches 4 ns apart with a 40-bucket gap
# (480 * 2 + 40 = 1000 = h), macroparticle counts uniform in [5e4, 2e5]
# (seed 5), BBR wake per bunch on 100 fixed bins over +-5 sigma_z, the ion
# cloud of TRAIN with 10 000 new macro-ions per passage (at most 5e6).
LARGE = dict(nb=480, spacing=2, count_range=(50_000, 200_000), seed=5, turns=1000, nbins=100,
             ions=dict(TRAIN["ions"], n_new=10_000, max_ions=5_000_000, seed=5))
