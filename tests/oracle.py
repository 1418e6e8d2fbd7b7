"""ctypes bindings of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg load it.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
# the same sources built at -O0 (the reference Makefile's level), for the CPU baseline
ORACLE_O0 = os.path.join(ROOT, "oracle", "liboracle_O0.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libfaddeeva_ref.so")

D = ctypes.c_double
PD = ctypes.POINTER(D)
PI32 = ctypes.POINTER(ctypes.c_int32)
I64 = ctypes.c_int64
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64


class OrcBunch(ctypes.Structure):
    _fields_ = [("n", I64), ("x", PD), ("px", PD), ("y", PD), ("py", PD), ("z", PD), ("dp", PD),
                ("lost", PI32), ("fx_ion", PD), ("fy_ion", PD)]


class OrcMap(ctypes.Structure):
    _fields_ = [("b1h1", PD), ("h2b2", PD), ("eta_x", D), ("etap_x", D), ("eta_y", D), ("etap_y", D),
                ("alpha_x", D), ("beta_x", D), ("alpha_y", D), ("beta_y", D),
                ("phase_adv_x", D), ("phase_adv_y", D), ("nux", D), ("nuy", D),
                ("chrom_x", D), ("chrom_y", D), ("adtx", D * 5), ("adty", D * 5)]


class OrcTwiss(ctypes.Structure):
    _fields_ = [("beta_x", D), ("alpha_x", D), ("eta_x", D), ("etap_x", D),
                ("beta_y", D), ("alpha_y", D), ("eta_y", D), ("etap_y", D), ("beta_z", D)]


class OrcMoments(ctypes.Structure):
    _fields_ = [("n_alive", I64), ("mean", D * 6), ("emit", D * 3), ("emit_eff", D * 3),
                ("rms_x", D), ("rms_y", D), ("rms_eff_x", D), ("rms_eff_y", D),
                ("sigma_z", D), ("sigma_dp", D), ("zmin", D), ("zmax", D),
                ("eigen_x", D), ("eigen_y", D), ("emit4d", D), ("xy_alpha", D), ("c2", D * 9)]


class OrcIonCloud(ctypes.Structure):
    _fields_ = [("nspec", ctypes.c_int), ("cap", I64), ("n", I64 * 4),
                ("x", PD * 4), ("y", PD * 4), ("vx", PD * 4), ("vy", PD * 4), ("fx", PD * 4), ("fy", PD * 4),
                ("macro_charge", D * 4), ("mean_x", D * 4), ("mean_y", D * 4), ("rms_x", D * 4),
                ("rms_y", D * 4), ("mass_number", D * 4), ("cross_section", D * 4), ("pressure", D * 4),
                ("temperature", D), ("length", D), ("n_new", ctypes.c_int), ("loss_boundary", D),
                ("ap_x", D), ("ap_y", D), ("circ", D), ("harm", ctypes.c_int)]


class OrcBunchView(ctypes.Structure):
    _fields_ = [("xmean", D), ("ymean", D), ("rms_x", D), ("rms_y", D), ("n_e", D),
                ("gap", ctypes.c_int), ("gamma", D)]


_libs = {}
_build = "O2"


def set_build(name):
    """Which oracle build lib() returns from now on: "O2" (default) or "O0"."""
    global _build
    assert name in ("O2", "O0")
    _build = name


def lib():
    if _build not in _libs:
        path = ORACLE if _build == "O2" else ORACLE_O0
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        _libs[_build] = _bind(ctypes.CDLL(path))
    return _libs[_build]


def _bind(L):
    L.orc_faddeeva_w.argtypes = [D, D, PD, PD]
    I = ctypes.c_int
    L.orc_lrw_yokoya.argtypes = [D, D, PD]
    PIc = ctypes.POINTER(ctypes.c_int)
    L.orc_bl_tables.restype = I
    L.orc_bl_tables.argtypes = [I, PIc, PD, I, D, D, D, D, I, PD, PIc, I, PD, PD]
    L.orc_bl_kick.argtypes = [ctypes.POINTER(OrcBunch), D, D, D, D, D, I, I, D, D, D]
    L.orc_bl_state_size.restype = I
    L.orc_bl_state.argtypes = [I, PIc, PD, I, D, D, I, PD, PIc, PD]
    L.orc_bl_walk.argtypes = [PD, I, ctypes.POINTER(OrcBunch), I, PIc, PD, PD, I, I, D, D, D, D, PD, PD]
    L.orc_bbi_smooth.argtypes = [PD, I]
    L.orc_bbi_potential.argtypes = [PD, I, PD, PD, PD]
    L.orc_bbi_kick.argtypes = [ctypes.POINTER(OrcBunch), I, D, PIc, PD, PD, D, D, D, D, D, I, D, PD, PD]
    L.orc_bbi_td_wake.argtypes = [I, PD, PD]
    L.orc_bbi_td_wake.restype = I
    L.orc_bbi_td_bins.argtypes = [D, I, D]
    L.orc_bbi_td_bins.restype = I
    L.orc_bbi_td_kick.argtypes = [ctypes.POINTER(OrcBunch), I, PD, PD, PIc, D, D, D, D, I, D, D, PD, PD]
    L.orc_lrw_table.restype = I
    L.orc_lrw_table.argtypes = [I, I, D, I, PD, I, PD, D, D, I, PD, PD, PD, PD]
    L.orc_lrw_kick.restype = I
    L.orc_lrw_kick.argtypes = [I, I, ctypes.POINTER(I), I, D, D, D, PD, PD, PD, I, PD]
    L.orc_erfcx.restype = D; L.orc_erfcx.argtypes = [D]
    L.orc_w_im.restype = D; L.orc_w_im.argtypes = [D]
    L.orc_bassetti_erskine.argtypes = [D, D, D, D, PD, PD]
    L.orc_gaussian_field.argtypes = [D, D, D, D, PD, PD]
    L.orc_ion_field.argtypes = [D, D, D, D, PD, PD]
    L.orc_section_matrices.argtypes = [ctypes.POINTER(OrcTwiss), ctypes.POINTER(OrcTwiss), PD, PD, PD]
    L.orc_dgemm.argtypes = [ctypes.c_int] * 3 + [PD, PD, PD]
    L.orc_section_map.argtypes = [ctypes.POINTER(OrcBunch), ctypes.POINTER(OrcMap)]
    L.orc_skew_quad.argtypes = [ctypes.POINTER(OrcBunch), D]
    L.orc_rf_kick.argtypes = [ctypes.POINTER(OrcBunch), D, D, ctypes.c_int, D, D, D]
    L.orc_drift.argtypes = [ctypes.POINTER(OrcBunch), D, PD]
    L.orc_energy_loss.argtypes = [ctypes.POINTER(OrcBunch), D, D, D]
    L.orc_synrad.argtypes = [ctypes.POINTER(OrcBunch), PD, PD, PD, PD, PD]
    L.orc_mark_lost.restype = D
    L.orc_mark_lost.argtypes = [ctypes.POINTER(OrcBunch), D, ctypes.c_int, D, D, ctypes.POINTER(I64)]
    L.orc_zminmax.argtypes = [ctypes.POINTER(OrcBunch), PD, PD]
    L.orc_rms.argtypes = [ctypes.POINTER(OrcBunch), D, D, D, D, D, D, ctypes.POINTER(OrcMoments)]
    L.orc_det.restype = D; L.orc_det.argtypes = [ctypes.c_int, PD]
    L.orc_bbr_wake.argtypes = [D, ctypes.c_int] + [PD] * 9 + [PD]
    L.orc_wake_table.argtypes = [ctypes.c_int, D, ctypes.c_int, PD, PD, PD, PD]
    L.orc_sr_wake.restype = I64
    L.orc_sr_wake.argtypes = [ctypes.POINTER(OrcBunch), ctypes.c_int, ctypes.c_int, D, D, D, D, D,
                              ctypes.c_int, PD, PD, PD, PD, PI32, PD, PD]
    L.orc_philox4x32_10.argtypes = [ctypes.POINTER(U32), ctypes.POINTER(U32), ctypes.POINTER(U32)]
    L.orc_qe_normals.argtypes = [U64, U32, U64, PD]
    L.orc_box_muller.argtypes = [U32, U32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    L.orc_qe_fill.argtypes = [U64, U32, U64, I64, PD]
    L.orc_ion_generate.argtypes = [U64, U32, U32, ctypes.c_int, ctypes.c_int, D, D, D, D, PD, PD]
    L.orc_ion_rms.argtypes = [ctypes.POINTER(OrcIonCloud), ctypes.c_int]
    L.orc_ion_passage.restype = ctypes.c_int
    L.orc_ion_passage.argtypes = [ctypes.POINTER(OrcIonCloud), ctypes.POINTER(OrcBunch),
                                  ctypes.POINTER(OrcBunchView), D, D, U64, U32, U32, PD]
    L.orc_fit_wlinear.argtypes = [PD, PD, PD, ctypes.c_int, PD, PD]
    L.orc_cbm_analysis.argtypes = ([PD, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, PD,
                                    PD, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double] + [PD] * 12)
    L.orc_wake_pairs.argtypes = [ctypes.c_int, D, ctypes.c_int, PD, PD, PD, PI32, PD]
    return L


def ref_lib():
    """Reference Faddeeva.cpp build (oracle/_ref), or None when absent."""
    if not os.path.exists(REF):
        return None
    L = ctypes.CDLL(REF)
    L.ref_faddeeva_w.argtypes = [D, D, PD, PD]
    L.ref_erfcx.restype = D; L.ref_erfcx.argtypes = [D]
    L.ref_w_im.restype = D; L.ref_w_im.argtypes = [D]
    return L


def p(a):
    return a.ctypes.data_as(PD)


COORDS = ("x", "px", "y", "py", "z", "dp")


class Bunch:
    """numpy-owned copy of one bunch with an OrcBunch view."""

    def __init__(self, parts, lost=None, with_force=True):
        self.a = {c: np.array(parts[c], dtype=np.float64, copy=True) for c in COORDS}
        n = len(self.a["x"])
        self.lost = np.zeros(n, np.int32) if lost is None else np.array(lost, dtype=np.int32, copy=True)
        self.fx = np.zeros(n) if with_force else None
        self.fy = np.zeros(n) if with_force else None
        self.s = OrcBunch(n, *[p(self.a[c]) for c in COORDS], self.lost.ctypes.data_as(PI32),
                          p(self.fx) if with_force else None, p(self.fy) if with_force else None)

    def __getitem__(self, k):
        return self.lost if k == "lost" else self.a[k]


def section_matrices(t1, t2):
    b1h1, h1b1, h2b2 = np.zeros(36), np.zeros(36), np.zeros(36)
    lib().orc_section_matrices(t1, t2, p(b1h1), p(h1b1), p(h2b2))
    return b1h1, h1b1, h2b2


def make_map(sec, ring):
    """OrcMap for a capi.Section + capi.Ring (keeps the matrix arrays alive)."""
    b1 = np.array(sec.b1h1[:]); h2 = np.array(sec.h2b2[:])
    m = OrcMap(p(b1), p(h2), sec.eta_x, sec.etap_x, sec.eta_y, sec.etap_y, sec.alpha_x, sec.beta_x,
               sec.alpha_y, sec.beta_y, sec.phase_adv_x, sec.phase_adv_y, ring.nux, ring.nuy,
               ring.chrom[0], ring.chrom[1])
    m.adtx[:] = ring.adtx[:]
    m.adty[:] = ring.adty[:]
    m._keep = (b1, h2)
    return m


def rms(b, sec):
    m = OrcMoments()
    lib().orc_rms(ctypes.byref(b.s), sec.beta_x, sec.beta_y, sec.eta_x, sec.etap_x, sec.eta_y, sec.etap_y, m)
    return m


def qe_fill(seed, turn, base, n):
    out = np.empty(5 * n)
    lib().orc_qe_fill(seed, turn, base, n, p(out))
    return out.reshape(5, n)


class OracleTurn:
    """The reference turn order (src/MPBeam.cpp:1059-1102) on one bunch,
    with the bin edges of the reference (mode 0) or a fixed grid (mode 1)."""

    def __init__(self, ring, sec, nbins=0, mode=1, grid=(0, 0), bbr=None, macro_charge=1.0,
                 synrad=False, noise_seed=None, base=0, bl=None, bbi=None):
        self.L = lib()
        self.ring, self.sec = ring, sec
        self.map = make_map(sec, ring)
        self.nbins, self.mode, self.grid = nbins, mode, grid
        self.q = macro_charge
        self.synrad, self.noise_seed, self.base = synrad, noise_seed, base
        self.poten = np.zeros(3 * max(nbins, 1))
        self.counts = np.zeros(max(nbins, 1), np.int32)
        self.edges = np.zeros(2)
        if nbins:
            self.rs = np.array(bbr["rs"], float); self.qq = np.array(bbr["q"], float)
            self.om = np.array(bbr["fr"], float) * 2 * np.pi
            self.table = None
            if mode == 1:
                dz = (grid[1] - grid[0]) / nbins
                self.table = np.zeros(2 * nbins - 1)
                self.L.orc_wake_table(nbins, dz / 299792458.0, len(self.rs), p(self.rs), p(self.qq), p(self.om),
                                      p(self.table))
        # beam-loaded cavities: callable(turn) -> [(vr, vi, vb0, harm)], nbins
        self.bl = bl
        # broadband impedance: dict(nimp, f_max, zre, zim, q = this bunch's macro charge;
        # optional zqx, zqy = (re, im) and beta = (beta_x, beta_y) for the quadrupolar planes)
        self.bbi = bbi
        self.bbi_wake = self.bbi_prof = None
        self.b1h1 = np.array(sec.b1h1[:]); self.h1b1 = np.array(sec.h1b1[:])
        self.tau = np.array(ring.synrad_tau[:]); self.en = np.array(ring.emit_nat[:])

    def turn(self, b, turn, noise=None):
        self.first_half(b, turn, noise)
        return self.second_half(b)

    def first_half(self, b, turn, noise=None):
        """Map, skew, RF, drift, U0, synrad; returns the post-map centroid
        (x, y, z) of the survivors (MPBeamRMSCal, src/MPBeam.cpp:1068)."""
        cen = self.pre_rf(b)
        self.post_rf(b, turn, noise)
        return cen

    def pre_rf(self, b):
        """Map and skew; returns the post-map centroid."""
        L, R = self.L, self.ring
        L.orc_section_map(ctypes.byref(b.s), ctypes.byref(self.map))
        m = rms(b, self.sec)
        sig_stale = m.sigma_z
        self._sig_stale = sig_stale
        cen = (m.mean[0], m.mean[2], m.mean[4])
        if R.skew_k != 0:
            L.orc_skew_quad(ctypes.byref(b.s), R.skew_k)
        if self.bl is not None:
            # GetZMinMax at the RF stage (src/MPBeam.cpp:2637, 2716-2752)
            zmin, zmax = ctypes.c_double(), ctypes.c_double()
            L.orc_zminmax(ctypes.byref(b.s), ctypes.byref(zmin), ctypes.byref(zmax))
            self._bl_z = (zmin, zmax)
        return cen

    def post_rf(self, b, turn, noise=None):
        """The ideal cavities, the tabulated beam-loaded ones, drift, U0,
        synrad."""
        L, R = self.L, self.ring
        if self.bl is not None:
            zmin, zmax = self._bl_z
        for r in range(R.n_rf):
            L.orc_rf_kick(ctypes.byref(b.s), R.rf_vr[r], R.rf_vi[r], R.rf_harm[r], R.f0 * R.harmonic, R.rbeta,
                          R.energy)
        if self.bl is not None:
            cavs, nb_bins = self.bl(turn)
            for vr, vi, vb0, harm in cavs:
                L.orc_bl_kick(ctypes.byref(b.s), vr, vi, vb0, zmin.value, zmax.value, nb_bins, harm,
                              R.f0 * R.harmonic, R.rbeta, R.energy)
        L.orc_drift(ctypes.byref(b.s), R.circ, p(np.array(R.alphac[:])))
        L.orc_energy_loss(ctypes.byref(b.s), R.u0, R.energy, R.rbeta)
        if self.synrad:
            n = len(b.lost)
            if noise is None and self.noise_seed is not None:
                noise = qe_fill(self.noise_seed, turn, self.base, n)
            L.orc_synrad(ctypes.byref(b.s), p(self.b1h1), p(self.h1b1), p(self.tau), p(self.en),
                         None if noise is None else p(np.ascontiguousarray(noise).ravel()))

    def second_half(self, b):
        """Short-range wake and loss marking."""
        L, R = self.L, self.ring
        sig_stale = self._sig_stale
        if self.bbi is not None and self.bbi.get("td"):
            # the time-domain method (src/MPBeam.cpp:1337-1340): bbi["w"] =
            # the five wakes in flag order (None: plane off), bbi["pos"]
            c = self.bbi
            pos = np.ascontiguousarray(c["pos"], np.float64)
            nw = len(pos)
            n = L.orc_bbi_td_bins(R.t0, R.harmonic, abs(pos[1] - pos[0]))
            w = np.ascontiguousarray(np.concatenate([np.zeros(nw) if t is None else np.asarray(t, np.float64)
                                                     for t in c["w"]]))
            planes = np.array([int(t is not None) for t in c["w"]], np.int32)
            bx, by = c.get("beta", (1.0, 1.0))
            self.bbi_wake, self.bbi_prof = np.zeros(5 * n), np.zeros(3 * n)
            L.orc_bbi_td_kick(ctypes.byref(b.s), nw, p(pos), p(w), planes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                              bx, by, c["q"], R.t0, R.harmonic, R.rbeta, R.energy, p(self.bbi_wake),
                              p(self.bbi_prof))
            self.bbi_wake = self.bbi_wake.reshape(5, n)
            self.bbi_prof = self.bbi_prof.reshape(3, n)
        elif self.bbi is not None:
            # BBImpBeamInteraction, after the radiation stage (src/MPBeam.cpp:1084)
            c = self.bbi
            n, ni = 2 * c["nimp"] - 1, c["nimp"]
            self.bbi_wake, self.bbi_prof = np.zeros(5 * n), np.zeros(3 * n)
            zero = np.zeros(ni)
            # the reference's flag order: Z, Dx, Dy, Qx, Qy
            tabs = [(c["zre"], c["zim"])] + [c.get(k) for k in ("zdx", "zdy", "zqx", "zqy")]
            zre = np.ascontiguousarray(np.concatenate([t[0] if t else zero for t in tabs]))
            zim = np.ascontiguousarray(np.concatenate([t[1] if t else zero for t in tabs]))
            planes = np.array([int(c.get("long", 1))] + [int(t is not None) for t in tabs[1:]], np.int32)
            bx, by = c.get("beta", (1.0, 1.0))
            L.orc_bbi_kick(ctypes.byref(b.s), ni, c["f_max"], planes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                           p(zre), p(zim), bx, by,
                           c["q"], R.rbeta, R.t0, R.harmonic, R.energy, p(self.bbi_wake), p(self.bbi_prof))
            self.bbi_wake = self.bbi_wake.reshape(5, n)
            self.bbi_prof = self.bbi_prof.reshape(3, n)
        if self.nbins:
            tab = self.table
            L.orc_sr_wake(ctypes.byref(b.s), self.nbins, self.mode, sig_stale, self.grid[0], self.grid[1],
                          self.q, R.energy, len(self.rs), p(self.rs), p(self.qq), p(self.om), p(self.poten),
                          self.counts.ctypes.data_as(PI32), p(self.edges), None if tab is None else p(tab))
        cnt = I64()
        t = L.orc_mark_lost(ctypes.byref(b.s), R.t0, R.harmonic, R.ap_x, R.ap_y, ctypes.byref(cnt))
        return t, cnt.value


KBOLTZ = 1.3806505E-23


class IonCloud:
    """numpy-owned orc_ion_cloud (one interaction point)."""

    def __init__(self, cfg, circ, harmonic, nb):
        self.cap = int(cfg.max_ions)
        self.nspec = cfg.nspec
        size = self.cap + int(cfg.n_new) + 1
        self.arr = [{k: np.zeros(size) for k in ("x", "y", "vx", "vy", "fx", "fy")} for _ in range(cfg.nspec)]
        c = OrcIonCloud()
        c.nspec = cfg.nspec
        c.cap = self.cap
        for p_ in range(cfg.nspec):
            for k in ("x", "y", "vx", "vy", "fx", "fy"):
                getattr(c, k)[p_] = p(self.arr[p_][k])
            c.mass_number[p_] = cfg.mass_number[p_]
            c.cross_section[p_] = cfg.cross_section[p_]
            c.pressure[p_] = cfg.pressure[p_]
        c.temperature = cfg.temperature
        c.length = cfg.length
        c.n_new = cfg.n_new
        c.loss_boundary = cfg.loss_boundary
        c.ap_x, c.ap_y = cfg.ap_x, cfg.ap_y
        c.circ = circ
        c.harm = harmonic
        self.c = c

    def species(self, p_=0):
        n = self.c.n[p_]
        return {k: self.arr[p_][k][:n].copy() for k in ("x", "y", "vx", "vy")}


class OracleTrain:
    """The reference turn of a bunch train with the beam-ion phase at
    interaction point 0 (src/MPBeam.cpp:1059-1102, 2585-2611): moments of
    every bunch, passages in train order, then each bunch's own turn."""

    def __init__(self, ring, sec, counts, n_e, gap, ion_cfg, seed, lrw=None, bl=None, bbi=None, **turn_kw):
        self.L = lib()
        # beam loading: dict(nres, mode, type, harm, cold, rs, q0, coupling, detune, vol, phase, nbins, nturns)
        self.bl_dev = self.bl_info = None
        self.bl_walk = None
        if bl is not None and any(bl["mode"][i] == 1 and (bl.get("excite", [0] * bl["nres"])[i] or
                                                        bl.get("dirfb", [0] * bl["nres"])[i])
                                  for i in range(bl["nres"])):
            # the cavity state depends on the particles: walked over the
            # train every turn (orc_bl_walk)
            nres = bl["nres"]
            z0 = [0] * nres
            ip = np.ascontiguousarray(np.concatenate([bl[k] for k in ("mode", "type", "harm", "cold")] +
                                                     [bl.get("excite", z0), bl.get("dirfb", z0)]), np.int32)
            dp = np.ascontiguousarray(np.concatenate(
                [np.asarray(bl[k], float) for k in ("rs", "q0", "coupling", "detune", "vol", "phase")] +
                [np.asarray(bl.get(k, [0.0] * nres), float) for k in ("fb_gain", "fb_phase", "fb_delay")]))
            st = np.zeros(nres * self.L.orc_bl_state_size())
            PIc = ctypes.POINTER(ctypes.c_int)
            gi = np.ascontiguousarray(gap, np.int32)
            self.L.orc_bl_state(nres, ip.ctypes.data_as(PIc), p(dp), ring.harmonic, ring.f0, ring.u0, len(counts),
                                p(np.ascontiguousarray(n_e, float)), gi.ctypes.data_as(PIc), p(st))
            self.bl_walk = dict(state=st, nres=nres, nbins=bl["nbins"],
                                mq=np.ascontiguousarray(bl["macro_charge"], float))
            bl = None
        if bl is not None:
            nb, nres = len(counts), bl["nres"]
            ip = np.ascontiguousarray(np.concatenate([bl[k] for k in ("mode", "type", "harm", "cold")]), np.int32)
            dp = np.ascontiguousarray(np.concatenate([np.asarray(bl[k], float) for k in
                                                      ("rs", "q0", "coupling", "detune", "vol", "phase")]))
            nbl = int(np.sum(np.asarray(bl["mode"]) == 1))
            dev = np.zeros(bl["nturns"] * nb * nbl * 3)
            info = np.zeros(bl["nturns"] * nb * nres * 8)
            gi = np.ascontiguousarray(gap, np.int32)
            PIc = ctypes.POINTER(ctypes.c_int)
            r = self.L.orc_bl_tables(nres, ip.ctypes.data_as(PIc), p(dp), ring.harmonic, ring.f0, ring.t0, ring.u0,
                                     ring.rbeta, nb, p(np.ascontiguousarray(n_e, float)), gi.ctypes.data_as(PIc),
                                     bl["nturns"], p(dev), p(info))
            assert r == nbl
            self.bl_dev = dev.reshape(bl["nturns"], nb, nbl, 3)
            self.bl_info = info.reshape(bl["nturns"], nb, nres, 8)
            bl_harm = [bl["harm"][i] for i in range(nres) if bl["mode"][i] == 1]
        # long-range wake: dict(nturns, sectors [n][7], res [n][9], beta_x, beta_y)
        self.lrw = None
        if lrw is not None:
            cap = 2 * lrw["nturns"] * ring.harmonic + 4
            t, w = np.zeros(cap), np.zeros(3 * cap)
            sec_a = np.ascontiguousarray(np.asarray(lrw.get("sectors", np.zeros((0, 7))), float).reshape(-1, 7))
            res_a = np.ascontiguousarray(np.asarray(lrw.get("res", np.zeros((0, 9))), float).reshape(-1, 9))
            npt = self.L.orc_lrw_table(lrw["nturns"], ring.harmonic, ring.t0, len(sec_a), p(sec_a), len(res_a),
                                       p(res_a), lrw["beta_x"], lrw["beta_y"], cap, p(t), p(w), None, None)
            assert npt > 0
            w = np.ascontiguousarray(np.concatenate([w[k * npt:(k + 1) * npt] for k in range(3)]))
            self.lrw = dict(lrw, np=npt, w=w, time=t[:npt].copy(),
                            pos=np.zeros((lrw["nturns"], len(counts), 3)))
            self.lrw_harm = np.concatenate([[0], np.cumsum(np.asarray(gap)[:-1])]).astype(np.int32)
            self.last_lrw_kick = None
        self.ring, self.sec = ring, sec
        self.counts = np.asarray(counts, dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)])
        self.n_e = np.asarray(n_e, dtype=np.float64)
        self.gap = np.asarray(gap, dtype=np.int32)
        self.cloud = IonCloud(ion_cfg, ring.circ, ring.harmonic, len(counts)) if ion_cfg is not None else None
        self.seed = seed
        def bl_of(b):
            if self.bl_dev is None:
                return None
            return lambda t: ([(*self.bl_dev[t, b, q], bl_harm[q]) for q in range(len(bl_harm))], bl["nbins"])
        self.turns = [OracleTurn(ring, sec, base=int(self.offsets[b]), bl=bl_of(b),
                                 bbi=None if bbi is None else dict(bbi, q=bbi["q"][b]), **{**turn_kw, **(
            {"macro_charge": turn_kw["macro_charge"][b]} if "macro_charge" in turn_kw else {})})
            for b in range(len(counts))]
        self.last_cloud = None

    def ion_phase(self, bunches, turn):
        ms = [rms(b, self.sec) for b in bunches]
        smax_x = max(m.rms_x for m in ms)
        smax_y = max(m.rms_y for m in ms)
        out = np.zeros((len(bunches), self.cloud.nspec, 5))
        for j, b in enumerate(bunches):
            m = ms[j]
            v = OrcBunchView(m.mean[0], m.mean[2], m.rms_x, m.rms_y, self.n_e[j], int(self.gap[j]),
                             self.ring.rgamma)
            o = np.zeros(5 * self.cloud.nspec)
            rc = self.L.orc_ion_passage(ctypes.byref(self.cloud.c), ctypes.byref(b.s), ctypes.byref(v),
                                        smax_x, smax_y, self.seed, turn, j, p(o))
            if rc != 0:
                raise RuntimeError("ion cloud above its maximum")
            out[j] = o.reshape(self.cloud.nspec, 5)
        self.last_cloud = out
        return out

    def turn(self, bunches, turn):
        if self.cloud is not None:
            self.ion_phase(bunches, turn)
        if self.bl_walk is not None:
            # every bunch up to the skew quad, the cavities' walk over the
            # train (MPBeam::BeamMomtumUpdateDueToRFTest), then the rest
            for t, b in zip(self.turns, bunches):
                t.pre_rf(b)
            W, nb, R = self.bl_walk, len(bunches), self.ring
            arr = (OrcBunch * nb)(*[b.s for b in bunches])
            info, fb = np.zeros(nb * W["nres"] * 8), np.zeros(nb * W["nres"])
            self.L.orc_bl_walk(p(W["state"]), W["nres"], arr, nb,
                               self.gap.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), p(self.n_e), p(W["mq"]),
                               W["nbins"], R.harmonic, R.f0, R.t0, R.rbeta, R.energy, p(info), p(fb))
            self.bl_last_info = info.reshape(nb, W["nres"], 8)
            self.bl_last_fb = fb.reshape(nb, W["nres"])
            for t, b in zip(self.turns, bunches):
                t.post_rf(b, turn)
            return [t.second_half(b) for t, b in zip(self.turns, bunches)]
        if self.lrw is None:
            return [t.turn(b, turn) for t, b in zip(self.turns, bunches)]
        # every bunch up to synrad, then MPBeam::LRWakeBeamIntaction
        # (src/MPBeam.cpp:1085, 3232-3351), then the short-range wake and
        # the loss marking
        cen = np.array([t.first_half(b, turn) for t, b in zip(self.turns, bunches)])
        W = self.lrw
        W["pos"] = np.concatenate([W["pos"][1:], cen[None]], axis=0)
        nb = len(bunches)
        kick = np.zeros(3 * nb)
        pos = np.ascontiguousarray(W["pos"].reshape(-1))
        harm = self.lrw_harm.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        rc = self.L.orc_lrw_kick(W["nturns"], nb, harm, self.ring.harmonic, self.ring.t0, self.ring.rbeta,
                                 self.ring.energy, p(np.ascontiguousarray(self.n_e)), p(pos), p(W["w"]), W["np"],
                                 p(kick))
        if rc != 0:
            raise RuntimeError(f"long-range wake: tau > 0 ({rc})")
        kick = kick.reshape(nb, 3)
        self.last_lrw_kick = kick
        for j, b in enumerate(bunches):
            # Bunch::BunchTransferDueToWake (src/Bunch.cpp:1091-1101): every particle
            b["px"][:] += kick[j, 0]
            b["py"][:] += kick[j, 1]
            b["dp"][:] += kick[j, 2]
        return [t.second_half(b) for t, b in zip(self.turns, bunches)]
