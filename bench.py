"""Throughput of the multi-particle tracking turn on one MI355X.

Workload (default): single bunch, 1e7 FP64 macroparticles, the one-section
ring of cetasim_amd.configs, quantum excitation from the on-device Philox
stream (seed 2), broadband-resonator short-range wake on 400 fixed bins over
+-5 sigma_z.  A step is one full turn (section map, RF, drift, U0, radiation
damping + excitation, binning, wake kick, loss marking, moments).

Prints one JSON line: macroparticle-turns/s, the dominant kernel's HBM
roofline fraction (algorithmic 96 B per macroparticle-turn over the kernel's
HIP-event duration) and the CPU oracle's rate on a bounded sample.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_MPT = 96.0  # 6 FP64 read + 6 FP64 written per macroparticle-turn


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--nbins", type=int, default=400)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (reads MIB, writes MIB) three times, "
                         "to scale FETCH_SIZE/WRITE_SIZE counters")
    return ap.parse_args()


def copy_calibration(mib):
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "launches": 3}))


def cpu_baseline(ring, secs, p, budget_s):
    """The oracle (-O2 C restatement, one core) on a sample of the same
    workload: 20 000 macroparticles, same ring, bins, wake and excitation."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np  # noqa: F401
    import oracle as O
    from cetasim_amd import configs

    n = 20_000
    parts = configs.gaussian_bunch(n, 2, p)
    q, _ = configs.bunch_scalars([n], ring, p)
    sz = p["sig"][2]
    ot = O.OracleTurn(ring, secs[0], nbins=400, mode=1, grid=(-5 * sz, 5 * sz), bbr=p["bbr"],
                      macro_charge=q[0], synrad=True, noise_seed=2)
    ob = O.Bunch(parts)
    ot.turn(ob, 0)
    turns = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        ot.turn(ob, turns + 1)
        turns += 1
    dt = time.perf_counter() - t0
    return {"value": n * turns / dt, "unit": "macroparticle-turns/s", "cores": 1, "kind": "port",
            "sample": f"{n} macroparticles x {turns} turns of the same ring/wake/excitation "
                      f"(oracle/ -O2 C restatement, {platform.processor() or platform.machine()})"}


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported "
                                   "(multi-GPU sharding is a design note only)"}))
        sys.exit(2)
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

    for t in range(a.warmup):
        tr.turn(t)
    tr.sync()
    tr.set_timing(True)
    t0 = time.perf_counter()
    for t in range(a.warmup, a.warmup + a.steps):
        tr.turn(t)
    tr.sync()
    dt = time.perf_counter() - t0
    pass_ms, n_pass = tr.kernel_time("pass")
    post_ms, _ = tr.kernel_time("post")
    tr.set_timing(False)
    m = tr.moments()[0]
    tr.close()

    value = n * a.steps / dt
    achieved = BYTES_PER_MPT * n / (pass_ms * 1e-3) / 1e9
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
                   "macroparticles": n, "bins": a.nbins, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_pass (fused turn)", "kernel_ms": pass_ms, "launches": n_pass,
                     "other_kernels_ms": {"post (moments + wake)": post_ms}},
        "transmission": m["transmission"],
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ring, secs, p, a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
