"""Throughput of the multi-particle tracking turn on one MI355X.

Workload (default, --workload c5): a 480-bunch train (5.9e7 FP64
macroparticles, counts uniform in [5e4, 2e5], seed 5) with a broadband-
resonator wake on 100 fixed bins per bunch, quantum excitation from the
device Philox stream, and residual-gas ions at one interaction point (10 000
new macro-ions per passage, at most 5e6) kicked both ways by the
Bassetti-Erskine field; the turns run as HIP-graph replays.  It is timed
with the exact FP64 field (plain) and with the learned field surrogate
(trained on the GPU first); `value` is the rate of the faster of the two
whose field is within 1e-4 of the FP64 field (`value_path` names it; the
plain path is that field).  --workload c3 is
the 80-bunch train, c2 the single 1e7-macroparticle bunch without ions.  A
step is one full turn.

Prints one JSON line: macroparticle-turns/s, the streaming pass's HBM
roofline fraction (algorithmic 96 B per macroparticle-turn over the
kernel's HIP-event duration), per-kernel times, the surrogate's field error
and the CPU oracle's rate extrapolated from bounded samples.
"""
import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "macroparticle-turns/s on 1 MI355X: surrogate vs plain kick at \u22641e-4 field err"
BYTES_PER_MPT = 96.0  # 6 FP64 read + 6 FP64 written per macroparticle-turn
FP64_PEAK_TFLOPS = 78.6
# FP64 operations per Bassetti-Erskine evaluation as the counters count them
# (SQ_INSTS_VALU_FLOPS_FP64 x 64 over the evaluations of a launch): the
# electron-side kick, and an ion's passage (the field plus drift and moments)
FP64_SOURCE = "profiles/ion_kernels_train_480_bunches_fp64_counters.md"
FP64_FLOP_PER_KICK = 450.0
FP64_FLOP_PER_ION_PASSAGE = 610.0


def measured_traffic_per_mpt(name="k_pass_single_bunch_1e7_hbm_traffic.txt"):
    """HBM bytes per macroparticle-turn of k_pass from a committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as fh:
            for line in fh:
                if "traffic per launch" in line and "B per macroparticle-turn" in line:
                    return float(line.split("(")[1].split()[0]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--nbins", type=int, default=400)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--workload", choices=("c2", "c3", "c5"), default="c5",
                    help="c5 (default): 480-bunch train (6e7 macroparticles) with wake and ions, plain vs "
                         "surrogate kicks; c3: 80-bunch train with the beam-ion phase; "
                         "c2: single bunch 1e7 macroparticles with wake and excitation (plain path)")
    ap.add_argument("--eager", action="store_true",
                    help="time turns launched one by one instead of HIP-graph replays")
    ap.add_argument("--plain-only", action="store_true", help="c3: skip the surrogate run")
    ap.add_argument("--lrwake", type=int, default=0, metavar="NTURNS",
                    help="c3/c5: add the long-range wake (lrwNTurnsWakeTrunction = NTURNS - 1; a 2304 m round "
                         "resistive-wall pipe and the workload's resonator)")
    ap.add_argument("--beam-loading", action="store_true",
                    help="c3/c5: the cavity beam-loaded (rfMode 1: active, Rs 5 MOhm, Q0 2e4, coupling 3, "
                         "detuned -20 kHz, 60 bins per bunch) in place of the ideal one")
    ap.add_argument("--bbimp", action="store_true",
                    help="c3/c5: add the broadband impedance (longitudinal; a broadband resonator Rs 2 kOhm, Q 1, "
                         "20 GHz tabulated at 261 frequencies up to 100 GHz: 521 bins of 1.5 mm per bunch)")
    ap.add_argument("--bbimp-td", action="store_true",
                    help="c3/c5: add the broadband impedance in the time domain (longitudinal; the broadband "
                         "resonator's wake function W(z), Rs 2 kOhm, Q 1, 20 GHz, tabulated every 0.5 mm over "
                         "0.2 m: the 0.6 m bucket on 1199 bins)")
    ap.add_argument("--sur-steps", type=int, default=20000, help="c3: Adam steps of the surrogate training")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (reads MIB, writes MIB) three times, "
                         "to scale FETCH_SIZE/WRITE_SIZE counters")
    a = ap.parse_args()
    if a.bbimp and a.bbimp_td:
        ap.error("--bbimp and --bbimp-td are the two methods of one impedance: choose one")
    return a


def progress(msg):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def timed_turns(tr, a, names):
    """Warm-up turns (the turn graphs are captured there), then the timed
    turns, then as many turns again launched one by one with HIP events
    around every kernel on the handle's stream: the kernel durations.
    Returns (timed seconds, event-stretch seconds, {name: (avg ms, launches)})."""
    graph = not a.eager
    tr.run_turns(0, a.warmup, graph)
    tr.sync()
    t0 = time.perf_counter()
    tr.run_turns(a.warmup, a.steps, graph)
    tr.sync()
    dt = time.perf_counter() - t0
    tr.set_timing(True)
    t0 = time.perf_counter()
    for t in range(a.warmup + a.steps, a.warmup + 2 * a.steps):
        tr.turn(t)
    tr.sync()
    dt_ev = time.perf_counter() - t0
    kt = {k: tr.kernel_time(k) for k in names}
    tr.set_timing(False)
    return dt, dt_ev, kt


def copy_calibration(mib):
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "launches": 3}))


def host_cpu():
    """CPU model and logical CPU counts (all, and this process's affinity)."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def wake_pairs_seconds(O, nbins, dt, bbr):
    """One bunch's wake sum evaluated as the reference does, one pseudo-Green
    call per bin pair (src/MPBunch.cpp:1983-1993): seconds per call,
    timed at 100 bins and scaled by the pair count."""
    import numpy as np
    nb0 = min(nbins, 100)
    rs, q = np.array(bbr["rs"], float), np.array(bbr["q"], float)
    om = np.array(bbr["fr"], float) * 2 * np.pi
    counts = np.full(nb0, 1000, np.int32)
    pz = np.zeros(nb0)
    t0 = time.perf_counter()
    O.lib().orc_wake_pairs(nb0, dt, len(rs), O.p(rs), O.p(q), O.p(om), counts.ctypes.data_as(O.PI32), O.p(pz))
    return (time.perf_counter() - t0) * (nbins / nb0) ** 2


def cpu_baseline(ring, secs, p, budget_s, n_gpu, nbins):
    """The oracle (C restatement, one core) on a sample of the same workload:
    turns of 20 000 macroparticles with the same ring, bins, wake and
    excitation, and the wake sum evaluated per bin pair as the reference
    does; built at -O0 (the reference Makefile's level; `value`) and -O2."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    from cetasim_amd import configs

    n = 20_000
    sz = p["sig"][2]
    out = {}
    for build in ("O0", "O2"):
        O.set_build(build)
        parts = configs.gaussian_bunch(n, 2, p)
        q, _ = configs.bunch_scalars([n], ring, p)
        ot = O.OracleTurn(ring, secs[0], nbins=nbins, mode=1, grid=(-5 * sz, 5 * sz), bbr=p["bbr"],
                          macro_charge=q[0], synrad=True, noise_seed=2)
        ob = O.Bunch(parts)
        ot.turn(ob, 0)
        turns = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < budget_s / 2 or turns < 2:
            ot.turn(ob, turns + 1)
            turns += 1
        t_mp = (time.perf_counter() - t0) / (n * turns)
        t_pairs = wake_pairs_seconds(O, nbins, 10 * sz / nbins / configs.C_LIGHT, p["bbr"])
        out[build] = (n_gpu / (n_gpu * t_mp + t_pairs), t_mp, t_pairs, turns)
    O.set_build("O2")
    v0, t_mp0, t_pairs0, turns0 = out["O0"]
    return {"value": v0, "unit": "macroparticle-turns/s", "cores": 1, "kind": "port",
            "sample": f"restatement of the reference's path (oracle/, not the reference build), -O0 as the "
                      f"reference Makefile: {t_mp0 * 1e9:.0f} ns per macroparticle-turn ({n} macroparticles x "
                      f"{turns0} turns, same ring, wake, excitation) x {n_gpu} macroparticles + the wake sum per "
                      f"bin pair as the reference evaluates it ({nbins}^2 pairs, {t_pairs0:.2f} s per turn)",
            "value_O2": out["O2"][0], "value_O2_lag_table_wake": 1.0 / out["O2"][1], **host_cpu()}


def cpu_baseline_train(ring, secs, p, budget_s, T, ion_evals, ntot, nb):
    """The oracle (C restatement, one core) on bounded samples of the train,
    extrapolated to the workload: the ion phase per macro-ion passage (a
    40-bunch train of 300-macroparticle bunches, 1000 new macro-ions per
    passage), the electron-side kick per macroparticle (one 200 000-
    macroparticle bunch through a 10-ion cloud), the rest of the turn per
    macroparticle (turns of a 20 000-macroparticle bunch with the workload's
    wake and excitation) and, with a wake, each bunch's wake sum evaluated
    per bin pair as the reference does; ion_evals = macro-ion passages per
    turn of the GPU run.  Built at -O0 (the reference Makefile's level;
    `value`) and -O2."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle as O
    from cetasim_amd import configs

    def phase(counts, ions):
        counts = np.asarray(counts)
        nbp = len(counts)
        gap = configs.train_gaps(nbp, T["spacing"], ring.harmonic)
        cfg = configs.ion_config(ring, ions, p)
        q, ne = configs.bunch_scalars(counts, ring, p)
        parts = configs.train_bunches(counts, 3, p)
        off = np.concatenate([[0], np.cumsum(counts)])
        bunches = [O.Bunch({c: parts[c][off[b]:off[b + 1]] for c in O.COORDS}) for b in range(nbp)]
        ot = O.OracleTrain(ring, secs[0], counts, ne, gap, cfg, ions["seed"])
        t0 = time.perf_counter()
        out = ot.ion_phase(bunches, 0)
        return time.perf_counter() - t0, float(out[:, :, 0].sum())

    n = 20_000
    sz = p["sig"][2]
    res = {}
    for build in ("O0", "O2"):
        O.set_build(build)
        t_ph, evals = phase(np.full(40, 300), dict(T["ions"], n_new=1000))
        t_eval = t_ph / evals
        ne_big = 200_000
        t_k, _ = phase([ne_big], dict(T["ions"], n_new=10))
        t_e = t_k / ne_big
        kw = dict(nbins=T["nbins"], mode=1, grid=(-5 * sz, 5 * sz), bbr=p["bbr"],
                  macro_charge=configs.bunch_scalars([n], ring, p)[0][0]) if T.get("nbins") else {}
        ot = O.OracleTurn(ring, secs[0], synrad=True, noise_seed=2, **kw)
        ob = O.Bunch(configs.gaussian_bunch(n, 2, p))
        turns = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < budget_s / 2 or turns < 2:
            ot.turn(ob, turns)
            turns += 1
        t_rest = (time.perf_counter() - t0) / (n * turns)
        t_pairs = wake_pairs_seconds(O, T["nbins"], 10 * sz / T["nbins"] / configs.C_LIGHT, p["bbr"]) \
            if T.get("nbins") else 0.0
        t_turn = ion_evals * t_eval + ntot * (t_e + t_rest) + nb * t_pairs
        res[build] = dict(value=ntot / t_turn, t_eval=t_eval, evals=evals, t_ph=t_ph, t_e=t_e, t_rest=t_rest,
                          turns=turns, t_pairs=t_pairs, lag=ntot / (t_turn - nb * t_pairs))
    O.set_build("O2")
    r = res["O0"]
    return {"value": r["value"], "unit": "macroparticle-turns/s", "cores": 1, "kind": "port",
            "sample": f"restatement of the reference's path (oracle/, not the reference build), -O0 as the "
                      f"reference Makefile: ion phase {r['t_eval'] * 1e9:.0f} ns per macro-ion passage (40-bunch "
                      f"sample, {r['evals']:.0f} passages, {r['t_ph']:.2f} s) x {ion_evals:.3g} passages per turn "
                      f"of the GPU run; electron-side kick {r['t_e'] * 1e9:.0f} ns and rest of the turn "
                      f"{r['t_rest'] * 1e9:.0f} ns per macroparticle ({r['turns']} turns of {n}) x {ntot} "
                      f"macroparticles; wake sum per bin pair as the reference evaluates it "
                      f"{r['t_pairs']:.2f} s per bunch-turn x {nb} bunches",
            "value_O2": res["O2"]["value"], "value_O2_lag_table_wake": res["O2"]["lag"], **host_cpu()}


def train_surrogate(steps):
    """The field surrogate trained on the GPU (untimed set-up): 4e6 samples of
    the FP64 field (seed 4), Adam with a cosine schedule in three stages."""
    from cetasim_amd.surrogate import Surrogate

    s = Surrogate()
    t0 = time.perf_counter()
    s.generate(4_000_000, 1_000_000, 4)
    s.init_params(4)
    k = max(steps // 10, 1)
    s.train(2 * k, nwg=512, chunks=2, lr_max=3e-3, lr_min=1e-3, seed=4)
    s.train(4 * k, nwg=512, chunks=2, lr_max=1e-3, lr_min=1e-4, seed=5)
    loss = s.train(4 * k, nwg=512, chunks=2, lr_max=1e-4, lr_min=1e-6, seed=6)
    mx, rms = s.validate()
    return s, {"samples": 4_000_000, "validation_points": 1_000_000, "adam_steps": 10 * k,
               "batch": 512 * 2 * 64, "train_s": time.perf_counter() - t0, "final_loss": loss,
               "max_rel_field_error": mx, "rms_rel_field_error": rms}


_PARTS = {}


def workload_train(a):
    from cetasim_amd import configs
    return configs.LARGE if a.workload == "c5" else configs.TRAIN


def fp64_roofline(per_turn, n, ion_passages):
    """FP64 rate of the electron-side kick (one evaluation per
    macroparticle and ion species) and of the ion segments (one per ion
    passage) over their HIP-event times per turn."""
    out = {"bound": "fp64", "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "flop_source": FP64_SOURCE}
    for name, key, flop, count in (("k_ion_kick", "ion_kick", FP64_FLOP_PER_KICK, n),
                                   ("k_ion_turn", "ion_turn", FP64_FLOP_PER_ION_PASSAGE, ion_passages)):
        ms = per_turn.get(key)
        if ms:
            achieved = flop * count / (ms * 1e-3) / 1e12
            out[name] = {"achieved": achieved, "frac": achieved / FP64_PEAK_TFLOPS, "kernel_ms_per_turn": ms,
                         "flop_per_evaluation": flop, "evaluations_per_turn": count}
    return out


def train_turns(a, sur=None, mode=0):
    """Time a.steps turns of the train; returns (s, kernel times, cloud, moments, n, nb, event-stretch s)."""
    from cetasim_amd import configs
    from cetasim_amd.tracker import Tracker

    p = configs.BASE
    T = workload_train(a)
    ring, secs, _ = configs.lattice(p)
    if a.beam_loading:
        ring.n_rf = 0  # the beam-loaded cavity replaces the ideal one
    counts = configs.train_counts(T["nb"], T["count_range"][0], T["count_range"][1], T["seed"])
    nb = len(counts)
    gap = configs.train_gaps(nb, T["spacing"], ring.harmonic)
    q, ne = configs.bunch_scalars(counts, ring, p)
    key = (a.workload, T["seed"])
    if key not in _PARTS:  # the same synthetic bunches for every run of one invocation
        _PARTS.clear()
        _PARTS[key] = configs.train_bunches(counts, T["seed"], p)
    tr = Tracker(ring, secs)
    tr.set_bunches(counts, q, ne, gap)
    tr.upload(_PARTS[key])
    tr.set_ions(configs.ion_config(ring, T["ions"], p))
    tr.set_synrad(True, 2, 2)
    if T.get("nbins"):
        sz = p["sig"][2]
        bbr = p["bbr"]
        tr.set_wake(T["nbins"], 1, -5 * sz, 5 * sz, bbr["rs"], bbr["q"], bbr["fr"])
    if a.lrwake:
        bbr = p["bbr"]
        res = [[bbr["rs"][r], bbr["q"][r], bbr["fr"][r], 0.0, 1.0, bbr["fr"][r], 0.0, 1.0, bbr["fr"][r]]
               for r in range(len(bbr["rs"]))]
        tr.set_lrwake(a.lrwake, [[0.01, 0.01, 2304.0, 2.7, 4.25, 1, 3.5e7]], res, secs[0].beta_x, secs[0].beta_y)
    if a.bbimp:
        import numpy as np
        nimp, fmax = 261, 1.0e11
        f = np.arange(nimp) * fmax / (nimp - 1)
        z = np.zeros(nimp, complex)
        z[1:] = 2.0e3 / (1 + 1j * (f[1:] / 2.0e10 - 2.0e10 / f[1:]))
        tr.set_bbimp(z.real.copy(), z.imag.copy(), fmax)
    if a.bbimp_td:
        import numpy as np
        # W(tau) = 2 alpha Rs e^(-alpha tau) (cos wb tau - alpha/wb sin wb tau), W(0) = alpha Rs
        nw, dzw, rs, wr = 400, 5.0e-4, 2.0e3, 2 * math.pi * 2.0e10
        al = wr / 2.0
        wb = math.sqrt(wr * wr - al * al)
        tau = np.arange(nw) * dzw / configs.C_LIGHT
        w = 2 * al * rs * np.exp(-al * tau) * (np.cos(wb * tau) - al / wb * np.sin(wb * tau))
        w[0] = al * rs
        tr.set_bbimp_td(np.arange(nw) * dzw, [w, None, None, None, None])
    if a.beam_loading:
        tr.set_beam_loading(dict(nres=1, mode=[1], type=[1], harm=[1], cold=[0], rs=[5e6], q0=[2e4],
                                 coupling=[3.0], detune=[-2e4], vol=[configs.rf_voltage(p)], phase=[0.0],
                                 nbins=60, nturns=a.warmup + 2 * a.steps))
    if sur is not None:
        tr.set_field_model(sur, mode)
    tr.autotune()
    layout = tr.layout()
    progress(f"train of {nb} bunches set up (field model {mode}), layout {layout}; timing")
    dt, dt_ev, ktimes = timed_turns(tr, a, ("pass", "pass_a", "post", "finalize", "ion_prep", "ion_turn", "ion_compact",
                                            "ion_reduce", "ion_finish", "ion_kick", "lrwake",
                                            "bbi_bin", "bbi_wake", "bbi_kick"))
    ktimes = {k: v for k, v in ktimes.items() if v[1] > 0}
    cloud = tr.ion_passages()
    m = tr.moments()
    tr.close()
    return dt, ktimes, cloud, m, int(counts.sum()), nb, dt_ev, layout


def run_train(a):
    """80-bunch train with the beam-ion phase at one interaction point: the
    plain FP64 Faddeeva path, then (unless --plain-only) the surrogate path."""
    from cetasim_amd import configs

    p = configs.BASE
    T = workload_train(a)
    ring, secs, _ = configs.lattice(p)
    dt, ktimes, cloud, m_plain, n, nb, dt_ev, layout = train_turns(a)
    plain = n * a.steps / dt
    sur_info = None
    # value: the faster kick path whose field is within 1e-4 of the FP64
    # field (the plain path is that field)
    value, value_path = plain, "plain: FP64 Faddeeva kicks both ways"
    per_turn = {k: v[0] * v[1] / a.steps for k, v in ktimes.items()}
    if not a.plain_only:
        progress(f"plain path timed; training the surrogate ({a.sur_steps} Adam steps)")
        sur, sur_info = train_surrogate(a.sur_steps)
        # electron-side kicks from the surrogate (mode 2), and both sides (mode 3)
        dts, kts, cloud_s, m_sur, _, _, _, _ = train_turns(a, sur, 2)
        dt3, kt3, _, _, _, _, _, _ = train_turns(a, sur, 3)
        sur.close()
        rate_sur = n * a.steps / dts
        sur_info.update({"mode": "electron-side kicks (ions kicked by the exact field)",
                         "value": rate_sur, "speedup_vs_plain": rate_sur / plain,
                         "ms_per_step": dts / a.steps * 1e3,
                         "kernels_ms_per_turn": {k: v[0] * v[1] / a.steps for k, v in kts.items()},
                         "value_both_sides": n * a.steps / dt3,
                         "kernels_ms_per_turn_both_sides": {k: v[0] * v[1] / a.steps for k, v in kt3.items()},
                         "centroid_x_last_bunch": [m_plain[-1]["mean"][0], m_sur[-1]["mean"][0]]})
        if sur_info["max_rel_field_error"] <= 1e-4 and rate_sur > plain:
            value, value_path, dt = rate_sur, "surrogate electron-side kicks", dts
        elif sur_info["max_rel_field_error"] > 1e-4:
            value_path += (f" (the surrogate's max field error {sur_info['max_rel_field_error']:.2g} "
                           "exceeds 1e-4)")
        else:
            value_path += " (faster than the surrogate)"
    pass_ms = ktimes["pass"][0]
    achieved = BYTES_PER_MPT * n / (pass_ms * 1e-3) / 1e9
    tpm, tsrc = measured_traffic_per_mpt(f"k_pass_train_{nb}_bunches_hbm_traffic.txt")
    out = {
        "metric": METRIC,
        "value": value,
        "value_path": value_path,
        "unit": "macroparticle-turns/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": dt / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{nb} bunches 4 ns apart, counts uniform in "
                               f"[{T['count_range'][0]:.0e}, {T['count_range'][1]:.0e}] (seed {T['seed']}, "
                               f"total {n}), 1 nC, CO ions at 1e-9 Pa / 300 K over the circumference, "
                               f"{T['ions']['n_new']} macro-ions per passage, <= {T['ions']['max_ions']}, "
                               "Bassetti-Erskine both ways (FP64 Faddeeva; with the surrogate inside its "
                               "domain when value_path names it), excitation on (device Philox)"
                               + (f", BBR wake on {T['nbins']} fixed bins per bunch" if T.get("nbins") else "")
                               + (f", long-range wake over {a.lrwake} turns (2304 m round resistive-wall pipe "
                                  "and the resonator)" if a.lrwake else "")
                               + (", broadband impedance (longitudinal, 521 bins of 1.5 mm per bunch)"
                                  if a.bbimp else "")
                               + (", broadband impedance in the time domain (longitudinal, a 400-row wake "
                                  "table, 1199 bins of 0.5 mm over the bucket)" if a.bbimp_td else "")
                               + (", beam-loaded cavity (rfMode 1, stable model, 60 bins per bunch; the turn "
                                  "adds a read-only pass for the bunch extent)" if a.beam_loading else ""),
                   "macroparticles": n, "bunches": nb, "parallelism": "single GPU",
                   "launch": "eager" if a.eager else "HIP graph replays of two turns"},
        "ms_per_step_kernel_timing_run_plain": dt_ev / a.steps * 1e3,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": tpm * n if tpm else None,
                     "traffic_source": tsrc, "kernel": "k_pass (fused turn)",
                     "kernel_ms": pass_ms, "launches": ktimes["pass"][1],
                     "layout": {"workgroups": layout[0], "tile": layout[1]},
                     # the pass's algorithmic bytes over the whole timed turn (ions included)
                     "turn_frac": BYTES_PER_MPT * n / (dt / a.steps) / 1e9 / HBM_PEAK_GBS,
                     "other_kernels_ms_per_turn": per_turn},
        # the FP64 kernels of the turn: the ion segments and the electron-side
        # kicks, at the FP64 operation counts of the counters' profile
        "roofline_fp64": fp64_roofline(per_turn, n * len(T["ions"]["mass_number"]), float(cloud[:, :, 0].sum())),
        "ion_macroparticles_last_passage": float(cloud[-1, 0, 0]),
        # Bassetti-Erskine evaluations of the ion side per turn (ions met by
        # each bunch, summed over the train)
        "ion_macroparticle_passages_per_turn": float(cloud[:, :, 0].sum()),
        "surrogate": sur_info,
    }
    if not a.no_cpu_baseline:
        progress("GPU runs done; timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline_train(ring, secs, p, a.cpu_seconds, T, float(cloud[:, :, 0].sum()), n,
                                                 nb)
    print(json.dumps(out))


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported "
                                   "(multi-GPU sharding is a design note only)"}))
        sys.exit(2)
    if a.workload in ("c3", "c5"):
        return run_train(a)
    import numpy as np

    from cetasim_amd import configs
    from cetasim_amd.tracker import Tracker

    p = configs.BASE
    ring, secs, _ = configs.lattice(p)
    n = a.n
    parts = configs.gaussian_bunch(n, 2, p)
    tr = Tracker(ring, secs)
    q, ne = configs.bunch_scalars([n], ring, p)
    tr.set_bunches([n], q, ne)
    tr.upload(parts)
    del parts
    sz = p["sig"][2]
    bbr = p["bbr"]
    tr.set_wake(a.nbins, 1, -5 * sz, 5 * sz, bbr["rs"], bbr["q"], bbr["fr"])
    tr.set_synrad(True, 2, 2)
    nwg = tr.autotune()
    layout = tr.layout()

    dt, dt_ev, kt = timed_turns(tr, a, ("pass", "post"))
    pass_ms, n_pass = kt["pass"]
    post_ms, _ = kt["post"]
    m = tr.moments()[0]
    tr.close()

    value = n * a.steps / dt
    achieved = BYTES_PER_MPT * n / (pass_ms * 1e-3) / 1e9
    tpm, tsrc = measured_traffic_per_mpt()
    out = {
        "metric": "macroparticle-turns/s on 1 MI355X (plain FP64 path, single-bunch scale-up)",
        "value": value,
        "unit": "macroparticle-turns/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"single bunch, {n:.0e} macroparticles FP64 SoA, {a.nbins} fixed bins "
                               "+-5 sigma_z, BBR wake Rs=10kOhm fr=10GHz Q=1, quantum excitation "
                               "(device Philox, seed 2), one-section ring Qx/Qy/Qs=0.18/0.27/0.005",
                   "macroparticles": n, "bins": a.nbins, "parallelism": "single GPU",
                   "workgroups": nwg,
                   "launch": "eager" if a.eager else "HIP graph replays of two turns"},
        "ms_per_step_kernel_timing_run": dt_ev / a.steps * 1e3,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None if tpm is None else tpm * n,
                     "traffic_source": tsrc,
                     "kernel": "k_pass (fused turn)", "kernel_ms": pass_ms, "launches": n_pass,
                     "layout": {"workgroups": layout[0], "tile": layout[1]},
                     "turn_frac": BYTES_PER_MPT * n / (dt / a.steps) / 1e9 / HBM_PEAK_GBS,
                     "other_kernels_ms": {"post (moments + wake)": post_ms}},
        "transmission": m["transmission"],
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ring, secs, p, a.cpu_seconds, n, a.nbins)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
