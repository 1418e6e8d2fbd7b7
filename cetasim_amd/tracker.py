"""Thin Python handle over the C ABI (plumbing for tests and the benchmark)."""
import ctypes

import numpy as np

from . import capi

COORDS = ("x", "px", "y", "py", "z", "dp")


class Tracker:
    """One device handle: ring + sections + bunches."""

    def __init__(self, ring, sections):
        self.lib = capi.load()
        self.h = ctypes.c_void_p()
        secs = (capi.Section * len(sections))(*sections)
        capi.check(self.lib.cts_create(ring, secs, len(sections), ctypes.byref(self.h)))
        self.ring = ring
        self.nb = 0
        self.ntot = 0

    def close(self):
        if self.h:
            self.lib.cts_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_bunches(self, counts, macro_charge, n_e, gap=None):
        counts = np.ascontiguousarray(counts, dtype=np.int64)
        q = np.ascontiguousarray(macro_charge, dtype=np.float64)
        ne = np.ascontiguousarray(n_e, dtype=np.float64)
        g = None if gap is None else np.ascontiguousarray(gap, dtype=np.int32)
        capi.check(self.lib.cts_set_bunches(self.h, len(counts), capi.i64ptr(counts), capi.dptr(q),
                                            capi.dptr(ne), None if g is None else capi.i32ptr(g)))
        self.nb = len(counts)
        self.counts = counts
        self.ntot = int(counts.sum())

    def upload(self, parts, lost=None):
        arrs = [np.ascontiguousarray(parts[c], dtype=np.float64) for c in COORDS]
        lp = None
        if lost is not None:
            lost = np.ascontiguousarray(lost, dtype=np.int32)
            lp = capi.i32ptr(lost)
        capi.check(self.lib.cts_upload(self.h, *[capi.dptr(a) for a in arrs], lp))

    def download(self):
        out = {c: np.empty(self.ntot) for c in COORDS}
        lost = np.empty(self.ntot, dtype=np.int32)
        capi.check(self.lib.cts_download(self.h, *[capi.dptr(out[c]) for c in COORDS], capi.i32ptr(lost)))
        out["lost"] = lost
        return out

    def set_wake(self, nbins, mode, grid_lo, grid_hi, rs, q, fr):
        rs = np.ascontiguousarray(rs, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        fr = np.ascontiguousarray(fr, dtype=np.float64)
        capi.check(self.lib.cts_set_wake(self.h, nbins, mode, grid_lo, grid_hi, len(rs), capi.dptr(rs),
                                         capi.dptr(q), capi.dptr(fr)))
        self.nbins = nbins

    def set_synrad(self, on, noise_mode=0, seed=0):
        capi.check(self.lib.cts_set_synrad(self.h, int(on), noise_mode, seed))

    def push_noise(self, noise):
        noise = np.ascontiguousarray(noise, dtype=np.float64)
        assert noise.size == 5 * self.ntot
        capi.check(self.lib.cts_push_noise(self.h, capi.dptr(noise)))

    def set_lrwake(self, nturns, sectors=None, res=None, beta_x=1.0, beta_y=1.0):
        """Long-range wake: nturns = lrwNTurnsWakeTrunction + 1 (0: off);
        sectors [n][7] (radiusX, radiusY, length, betaX, betaY, number,
        conductivity), res [n][9] (Rs, Q, f, x Rs, Q, f, y Rs, Q, f)."""
        self._lrw_keep = []
        cfg = capi.LrwakeConfig()
        cfg.nturns = nturns
        for name, arr, width in (("sectors", sectors, 7), ("res", res, 9)):
            if arr is not None and len(arr):
                a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).reshape(-1, width))
                self._lrw_keep.append(a)
                setattr(cfg, name, capi.dptr(a))
                setattr(cfg, "n_sectors" if name == "sectors" else "n_res", a.shape[0])
        cfg.beta_x, cfg.beta_y = beta_x, beta_y
        capi.check(self.lib.cts_set_lrwake(self.h, ctypes.byref(cfg)))

    def lrwake_table(self):
        n = ctypes.c_int32(0)
        rc = self.lib.cts_get_lrwake_table(self.h, ctypes.byref(n), None, None, None, None)
        if rc != capi.CTS_ERR_ARG:
            capi.check(rc)
        t = np.empty(n.value)
        w, rw, bbr = (np.empty(3 * n.value) for _ in range(3))
        capi.check(self.lib.cts_get_lrwake_table(self.h, ctypes.byref(n), capi.dptr(t), capi.dptr(w), capi.dptr(rw),
                                                 capi.dptr(bbr)))
        return t, w.reshape(3, -1), rw.reshape(3, -1), bbr.reshape(3, -1)

    def lrwake_kick(self):
        out = np.empty(3 * self.nb)
        capi.check(self.lib.cts_get_lrwake_kick(self.h, capi.dptr(out)))
        return out.reshape(self.nb, 3)

    def set_beam_loading(self, bl):
        """bl: dict(nres, mode, type, harm, cold, rs, q0, coupling, detune, vol,
        phase, nbins, nturns), per-cavity lists, and optionally excite,
        dirfb, fb_gain, fb_phase, fb_delay (the bin-by-bin variant and the
        direct feedback); None switches it off."""
        cfg = capi.BeamLoadingConfig()
        if bl is not None:
            cfg.nres = bl["nres"]
            for i in range(bl["nres"]):
                cfg.rf_mode[i] = bl["mode"][i]
                cfg.res_type[i] = bl["type"][i]
                cfg.res_harm[i] = bl["harm"][i]
                cfg.cold[i] = bl["cold"][i]
                cfg.rs[i], cfg.q0[i], cfg.coupling[i] = bl["rs"][i], bl["q0"][i], bl["coupling"][i]
                cfg.detune_hz[i], cfg.vol[i], cfg.phase[i] = bl["detune"][i], bl["vol"][i], bl["phase"][i]
                cfg.excite_instability[i] = bl.get("excite", [0] * bl["nres"])[i]
                cfg.dir_fb[i] = bl.get("dirfb", [0] * bl["nres"])[i]
                cfg.dir_fb_gain[i] = bl.get("fb_gain", [0.0] * bl["nres"])[i]
                cfg.dir_fb_phase[i] = bl.get("fb_phase", [0.0] * bl["nres"])[i]
                cfg.dir_fb_delay[i] = bl.get("fb_delay", [0.0] * bl["nres"])[i]
            cfg.nbins, cfg.nturns = bl["nbins"], bl["nturns"]
        capi.check(self.lib.cts_set_beam_loading(self.h, ctypes.byref(cfg)))
        self.bl_nres = cfg.nres

    def set_bbimp(self, zre, zim, f_max, flags=(1, 0, 0, 0, 0), time_domain=0, zqx=None, zqy=None,
                  beta=(1.0, 1.0), zdx=None, zdy=None):
        """Broadband impedance table Z(f_k) (the file's rows, last frequency
        f_max); quadrupolar and dipolar tables zqx, zqy, zdx, zdy = (re, im);
        flags in the reference's order (Z, Dx, Dy, Qx, Qy); zre None
        switches it off."""
        cfg = capi.BbimpConfig()
        if zre is not None:
            a = lambda v: np.ascontiguousarray(v, np.float64)  # noqa: E731
            self._bbi_z = [a(zre), a(zim)]
            for t in (zqx, zqy, zdx, zdy):
                self._bbi_z += [a(t[0]), a(t[1])] if t else [None, None]
            cfg.nimp = len(self._bbi_z[0])
            cfg.f_max = f_max
            cfg.zz_re, cfg.zz_im = capi.dptr(self._bbi_z[0]), capi.dptr(self._bbi_z[1])
            names = (("zqx_re", "zqx_im"), ("zqy_re", "zqy_im"), ("zdx_re", "zdx_im"), ("zdy_re", "zdy_im"))
            for k, (nre, nim) in enumerate(names):
                if self._bbi_z[2 + 2 * k] is not None:
                    setattr(cfg, nre, capi.dptr(self._bbi_z[2 + 2 * k]))
                    setattr(cfg, nim, capi.dptr(self._bbi_z[3 + 2 * k]))
            cfg.beta_x, cfg.beta_y = beta
            cfg.flags[:] = list(flags)
            cfg.time_domain = time_domain
        capi.check(self.lib.cts_set_bbimp(self.h, ctypes.byref(cfg)))

    def set_bbimp_td(self, pos, wakes, beta=(1.0, 1.0)):
        """Broadband impedance in the time domain (the reference's default,
        BBImpBunchInteractionTD): pos the wake table's positions binPosZ,
        wakes the five tables in the reference's flag order (wz, wDx, wDy,
        wQx, wQy; None switches a plane off) as read_quasi_wake returns
        them."""
        cfg = capi.BbimpConfig()
        a = lambda v: np.ascontiguousarray(v, np.float64)  # noqa: E731
        self._bbi_w = [a(pos)] + [None if w is None else a(w) for w in wakes]
        cfg.time_domain = 1
        cfg.nwake = len(self._bbi_w[0])
        cfg.wake_pos = capi.dptr(self._bbi_w[0])
        for k, name in enumerate(("wake_z", "wake_dx", "wake_dy", "wake_qx", "wake_qy")):
            cfg.flags[k] = int(self._bbi_w[1 + k] is not None)
            if self._bbi_w[1 + k] is not None:
                if len(self._bbi_w[1 + k]) != cfg.nwake:
                    raise ValueError("set_bbimp_td: every wake table needs len(pos) rows")
                setattr(cfg, name, capi.dptr(self._bbi_w[1 + k]))
        cfg.beta_x, cfg.beta_y = beta
        capi.check(self.lib.cts_set_bbimp(self.h, ctypes.byref(cfg)))

    def bbimp_wake(self):
        """(wake [5][nb][2 nimp - 1] in the flags' order: longitudinal,
        dipolar x, y, quadrupolar x, y; profiles [3][nb][2 nimp - 1]: charge,
        x-weighted, y-weighted) of every bunch from the last turn."""
        n = capi.I32()
        capi.check(self.lib.cts_get_bbimp_wake(self.h, ctypes.byref(n), None, None))
        w, pr = np.empty(5 * self.nb * n.value), np.empty(3 * self.nb * n.value)
        capi.check(self.lib.cts_get_bbimp_wake(self.h, ctypes.byref(n), capi.dptr(w), capi.dptr(pr)))
        return w.reshape(5, self.nb, -1), pr.reshape(3, self.nb, -1)

    def cavity_info(self, turn):
        out = np.empty(self.nb * self.bl_nres * 8)
        capi.check(self.lib.cts_get_cavity_info(self.h, turn, capi.dptr(out)))
        return out.reshape(self.nb, self.bl_nres, 8)

    def set_ions(self, cfg, ip=0):
        """cfg: capi.IonConfig, or None to switch ions off, at interaction point ip."""
        if cfg is None:
            cfg = capi.IonConfig()
        capi.check(self.lib.cts_set_ions_ip(self.h, ip, ctypes.byref(cfg)))
        self.nspec = cfg.nspec

    def set_field_model(self, sur, mode):
        """sur: surrogate.Surrogate or None; mode bit 0 ion side, bit 1 electron side."""
        capi.check(self.lib.cts_set_field_model(self.h, sur.h if sur is not None else None, mode))
        self._sur = sur

    def ion_phase(self, turn, ip=0):
        return capi.check(self.lib.cts_ion_phase_ip(self.h, turn, ip))

    def ion_cloud(self, species=0, ip=0):
        n = ctypes.c_int64()
        capi.check(self.lib.cts_get_ion_cloud_ip(self.h, ip, species, 0, ctypes.byref(n), None, None, None, None))
        out = {c: np.empty(n.value) for c in ("x", "y", "vx", "vy")}
        capi.check(self.lib.cts_get_ion_cloud_ip(self.h, ip, species, n.value, ctypes.byref(n),
                                                 *[capi.dptr(out[c]) for c in ("x", "y", "vx", "vy")]))
        return out

    def ion_passages(self, ip=0):
        """[nb][MAX_SPECIES][6]: count, mean x, mean y, rms x, rms y, ion number."""
        out = np.empty(self.nb * capi.MAX_SPECIES * 6)
        capi.check(self.lib.cts_get_ion_passages_ip(self.h, ip, capi.dptr(out)))
        return out.reshape(self.nb, capi.MAX_SPECIES, 6)

    def turn(self, n):
        return capi.check(self.lib.cts_turn(self.h, n))

    def run_turns(self, first, nturns, graph=True):
        return capi.check(self.lib.cts_run_turns(self.h, first, nturns, int(graph)))

    def autotune(self):
        n = ctypes.c_int32()
        capi.check(self.lib.cts_autotune(self.h, ctypes.byref(n)))
        return n.value

    def set_layout(self, nwg, tile=0):
        capi.check(self.lib.cts_set_layout(self.h, nwg, tile))

    def layout(self):
        """(workgroups, tile) of the streaming pass (tile 0: contiguous chunks)."""
        n, t = ctypes.c_int32(), ctypes.c_int32()
        capi.check(self.lib.cts_get_layout(self.h, ctypes.byref(n), ctypes.byref(t)))
        return n.value, t.value

    def flush(self):
        capi.check(self.lib.cts_flush(self.h))

    def sync(self):
        capi.check(self.lib.cts_sync(self.h))

    def seam(self, name, *args):
        return capi.check(getattr(self.lib, "cts_" + name)(self.h, *args))

    def mark_lost(self):
        t = np.empty(self.nb)
        rc = self.lib.cts_mark_lost(self.h, capi.dptr(t))
        if rc not in (capi.CTS_OK, capi.CTS_LOW_TRANSMISSION):
            capi.check(rc)
        return t, rc

    def sr_wake(self, sigma_z_stale=None):
        if sigma_z_stale is None:
            capi.check(self.lib.cts_sr_wake(self.h, None))
        else:
            s = np.ascontiguousarray(sigma_z_stale, dtype=np.float64)
            capi.check(self.lib.cts_sr_wake(self.h, capi.dptr(s)))

    def rms(self):
        m = (capi.Moments * self.nb)()
        capi.check(self.lib.cts_rms(self.h, m))
        return [moments_dict(x) for x in m]

    def moments(self):
        m = (capi.Moments * self.nb)()
        capi.check(self.lib.cts_get_moments(self.h, m))
        return [moments_dict(x) for x in m]

    def eigen_emit(self):
        """[nb][4]: eigenEmitX, eigenEmitY, emitXY4D, xyCouplingAlpha (GetEigenEmit)."""
        out = np.empty(4 * self.nb)
        capi.check(self.lib.cts_eigen_emit(self.h, capi.dptr(out)))
        return out.reshape(self.nb, 4)

    def set_history(self, cap):
        capi.check(self.lib.cts_set_history(self.h, cap))

    def history(self, first, nturns):
        m = (capi.Moments * (self.nb * nturns))()
        capi.check(self.lib.cts_get_history(self.h, first, nturns, m))
        return [moments_dict(x) for x in m]

    def wake_state(self):
        counts = np.empty(self.nb * self.nbins, dtype=np.int32)
        edges = np.empty(self.nb * 2)
        kick = np.empty(self.nb * self.nbins)
        capi.check(self.lib.cts_get_wake_state(self.h, capi.i32ptr(counts), capi.dptr(edges), capi.dptr(kick)))
        return counts.reshape(self.nb, self.nbins), edges.reshape(self.nb, 2), kick.reshape(self.nb, self.nbins)

    def set_timing(self, on):
        capi.check(self.lib.cts_set_timing(self.h, int(on)))

    def kernel_time(self, name):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        capi.check(self.lib.cts_kernel_time(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def qe_normals(self, seed, turn, base, n):
        out = np.empty(5 * n)
        capi.check(self.lib.cts_debug_qe_normals(self.h, seed, turn, base, n, capi.dptr(out)))
        return out.reshape(5, n)


def read_quasi_wake(path, flags=(1, 1, 1, 1, 1)):
    """The wake file of the time-domain broadband impedance as
    MPBeam::GetQuasiWakePoten reads it (src/MPBeam.cpp:748-812): five
    header lines, then rows z, wz, wDx, wDy, wQx, wQy; every wake times -1,
    positions moved to start at 0, and the longitudinal wake folded about
    the row whose z is exactly 0.  Returns (pos, [wz, wDx, wDy, wQx, wQy])
    with None for the planes whose flag is 0."""
    rows = []
    with open(path) as f:
        for ln, line in enumerate(f):
            if ln < 5 or not line.split():
                continue
            v = [float(x) for x in line.split()]
            if len(v) < 6:
                raise ValueError(f"{path}: row {ln + 1} has {len(v)} columns, 6 needed")
            rows.append(v[:6])
    if len(rows) < 2:
        raise ValueError(f"{path}: fewer than two wake rows")
    r = np.array(rows)
    pos = r[:, 0] - r[0, 0]
    zero = np.flatnonzero(r[:, 0] == 0)
    cen = int(zero[-1]) if len(zero) else 0
    t = -1 * r[:, 1]
    n = len(t)
    if 2 * cen > n:
        # the fold would read past the table's end
        raise ValueError(f"{path}: the z = 0 row must lie in the first half of the table")
    wz = np.zeros(n)
    wz[0] = t[cen]
    for i in range(1, cen):
        wz[i] = t[cen + i] + t[cen - i]
    wz[cen:max(cen, n - cen)] = t[2 * cen:max(2 * cen, n)]
    wakes = [wz] + [-1 * r[:, c] for c in range(2, 6)]
    return pos, [w if flags[k] else None for k, w in enumerate(wakes)]


def moments_dict(m):
    return dict(n_alive=m.n_alive, mean=np.array(m.mean[:]), c2=np.array(m.c2[:]), zmin=m.zmin,
                zmax=m.zmax, loss_count=m.loss_count, transmission=m.transmission)
