"""Graph replays of the turn (cts_run_turns with use_graph = 1) against the
same turns run one by one (cts_turn): the kernels are the same, so the
particles, loss codes and turn-by-turn moments must agree bit for bit.
Every case replays the graphs several times, with eager work between the
replays (downloads, which apply the pending wake kick, and single turns),
and compares after each stretch.
"""
import numpy as np
import pytest

from cetasim_amd import configs
from cetasim_amd.tracker import Tracker

pytestmark = pytest.mark.gpu

BBR = configs.BASE["bbr"]
SZ = configs.BASE["sig"][2]
ION = dict(mass_number=[28.0], cross_section=[2.0e-22], pressure=[1.0e-5], temperature=300.0,
           n_new=300, max_ions=50_000, loss_boundary=10.0, seed=11)


def make(counts, wake_mode, ions, seed=5):
    ring, secs, _ = configs.lattice()
    counts = np.asarray(counts, dtype=np.int64)
    gap = configs.train_gaps(len(counts), 2, ring.harmonic)
    q, ne = configs.bunch_scalars(counts, ring)
    parts = configs.train_bunches(counts, seed)
    out = []
    for _ in range(2):
        tr = Tracker(ring, secs)
        tr.set_bunches(counts, q, ne, gap)
        tr.upload(parts)
        if wake_mode is not None:
            tr.set_wake(100, wake_mode, -5 * SZ, 5 * SZ, BBR["rs"], BBR["q"], BBR["fr"])
        tr.set_synrad(True, 2, 21)
        if ions:
            tr.set_ions(configs.ion_config(ring, ION))
        tr.set_history(64)
        out.append(tr)
    return out


def same(e, g, first, last, what):
    de, dg = e.download(), g.download()
    for c in ("x", "px", "y", "py", "z", "dp", "lost"):
        assert np.array_equal(de[c], dg[c]), (what, c)
    he, hg = e.history(first, last - first + 1), g.history(first, last - first + 1)
    for i, (a, b) in enumerate(zip(he, hg)):
        for k in ("n_alive", "zmin", "zmax", "loss_count", "transmission"):
            assert a[k] == b[k], (what, i, k)
        assert np.array_equal(a["mean"], b["mean"]), (what, i)
        assert np.array_equal(a["c2"], b["c2"]), (what, i)


@pytest.mark.parametrize("counts,wake_mode,ions", [
    ([20000], 1, False),             # fixed-grid wake: pass + post
    ([3000, 5001, 2500], 0, False),  # reference edges: pass + finalize + bin + table + wake
    ([3000, 5001, 2500], 1, True),   # ion train: ion phase + pass + post
    ([4000, 2000], None, True),      # no wake: the turn closes its own moments
    ([1500] * 40, 1, True),          # two ion segments: the cloud buffer returns each turn
])
def test_graph_replays_match_eager(counts, wake_mode, ions):
    e, g = make(counts, wake_mode, ions)
    print(f"graph case counts={counts} wake={wake_mode} ions={ions}", flush=True)
    # stretch 1: turn 0 runs eagerly (nothing pending), then 1-2 and 3-4 replay
    for t in range(5):
        e.turn(t)
    g.run_turns(0, 5)
    same(e, g, 0, 4, "turns 0-4")
    # the downloads applied the pending kicks on both; stretch 2 starts
    # eagerly again and ends with an odd turn
    for t in range(5, 12):
        e.turn(t)
    g.run_turns(5, 7)
    same(e, g, 5, 11, "turns 5-11")
    # stretch 3: one eager turn on both, then replays straight after it
    e.turn(12)
    g.turn(12)
    for t in range(13, 21):
        e.turn(t)
    g.run_turns(13, 8)
    same(e, g, 12, 20, "turns 12-20")
    if ions:
        np.testing.assert_array_equal(e.ion_passages(), g.ion_passages())
        ce, cg = e.ion_cloud(), g.ion_cloud()
        for k in ce:
            assert np.array_equal(ce[k], cg[k]), k
    e.close()
    g.close()


@pytest.mark.parametrize("wake_mode", [1, None])
def test_graph_replays_with_lrwake(wake_mode):
    """The long-range wake's kick is deferred like the short-range one and
    its centroid history is indexed by the device turn counter."""
    e, g = make([3000, 5001, 2500, 1800], wake_mode, False)
    sec = [[0.01, 0.01, 2304.0, 2.7, 4.25, 1, 3.5e5]]
    res = [[2e4, 1.5, 1.1e9, 3e6, 2.0, 0.9e9, 5e6, 2.5, 1.05e9]]
    for tr in (e, g):
        tr.set_lrwake(4, sec, res, 5.0, 7.0)
    print(f"graph case with long-range wake, wake={wake_mode}", flush=True)
    for t in range(9):
        e.turn(t)
    g.run_turns(0, 9)
    same(e, g, 0, 8, "turns 0-8")
    np.testing.assert_array_equal(e.lrwake_kick(), g.lrwake_kick())
    for t in range(9, 14):
        e.turn(t)
    g.run_turns(9, 5)
    same(e, g, 9, 13, "turns 9-13")
    e.close()
    g.close()


def test_graph_rebuilt_after_settings_change():
    """A replay after cts_set_wake uses the new settings (the graphs are
    dropped and captured again)."""
    e, g = make([6000], 1, False)
    for t in range(4):
        e.turn(t)
    g.run_turns(0, 4)
    for tr in (e, g):
        tr.set_wake(64, 1, -4 * SZ, 4 * SZ, BBR["rs"], BBR["q"], BBR["fr"])
    for t in range(4, 9):
        e.turn(t)
    g.run_turns(4, 5)
    same(e, g, 4, 8, "after set_wake")
    e.close()
    g.close()
