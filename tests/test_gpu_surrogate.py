"""The learned field surrogate on the GPU against its CPU restatement
(tests/mlp_ref.py) and the oracle's FP64 field.

Tolerances: training/validation samples' field values within 1e-13
relative plus 2e-14 absolute of the oracle (normalized fields are O(1);
the device evaluates w(z) near the beam by Weideman's rational
approximation, |w| relative error below 4e-14, the oracle by the
reference's Algorithm 916 series); network gradient and
outputs within the bf16x3 matrix-core precision (~2^-16 relative per
product): gradients 2e-3 relative to the largest component, outputs 2e-4
absolute on O(1) outputs.
"""
import ctypes

import numpy as np
import pytest

from cetasim_amd.surrogate import Surrogate

import mlp_ref as M
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sur():
    s = Surrogate()
    s.generate(1 << 16, 1 << 14, 4)
    s.init_params(7)
    yield s
    s.close()


def test_samples_match_oracle(sur):
    val = sur.samples(0, 2000, train=False)
    L = O.lib()
    fx, fy = ctypes.c_double(), ctypes.c_double()
    for au, av, r, gx, gy in val[:2000:7]:
        L.orc_ion_field(au, av * r, 1.0, r, fx, fy)
        assert abs(gx - fx.value) <= 1e-13 * abs(fx.value) + 2e-14
        assert abs(gy - fy.value) <= 1e-13 * abs(fy.value) + 2e-14
    assert np.all((val[:, 0] > 0) & (val[:, 0] <= 6) & (val[:, 1] > 0) & (val[:, 1] <= 6))
    assert np.all((val[:, 2] >= 0.02 * (1 - 1e-12)) & (val[:, 2] <= 1))
    # training records encode the same kind of points: features in [-1, 1],
    # and t = log(G / (u b)) is consistent with the decoding
    tr = sur.samples(0, 4000)
    assert np.all(np.abs(tr[:, :3]) <= 1.0 + 1e-6)
    au = (tr[:, 0].astype(np.float64) + 1) * 3
    av = (tr[:, 1].astype(np.float64) + 1) * 3
    r = np.exp(tr[:, 2].astype(np.float64) * (-M.LN_RMIN / 2) + M.LN_RMIN / 2)
    gx, gy = M.decode(au, av, r, np.zeros((len(tr), 2)), (0.0, 1.0, 0.0, 1.0))
    gx, gy = gx * np.exp(tr[:, 3]), gy * np.exp(tr[:, 4])
    for i in range(0, 4000, 97):
        L.orc_ion_field(au[i], av[i] * r[i], 1.0, r[i], fx, fy)
        # float32 features and targets: ~1e-6 relative
        assert abs(gx[i] - fx.value) <= 2e-5 * abs(fx.value)
        assert abs(gy[i] - fy.value) <= 2e-5 * abs(fy.value)


def test_gradient_matches_cpu(sur):
    p, norm = sur.params()
    recs = sur.samples(0, 3000)
    loss, g = sur.gradient(recs)
    eloss, eg = M.loss_grad(p.astype(np.float64), recs.astype(np.float64), norm.astype(np.float64))
    assert abs(loss - eloss) <= 1e-4 * abs(eloss)
    scale = np.max(np.abs(eg))
    err = np.abs(g - eg)
    assert np.max(err) <= 2e-3 * scale, (np.argmax(err), np.max(err) / scale)
    # every block of the gradient is exercised (no block identically zero)
    for a, b in ((0, 192), (192, 256), (256, 4352), (4352, 4416), (4416, 4544), (4544, 4546)):
        assert np.max(np.abs(g[a:b])) > 0


def test_eval_matches_cpu(sur):
    p, _ = sur.params()
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (5000, 3)).astype(np.float32)
    out = sur.eval(X)
    exp, _ = M.forward(p.astype(np.float64), X.astype(np.float64))
    assert np.max(np.abs(out - exp)) <= 2e-4 * max(1.0, np.max(np.abs(exp)))


def test_training_lowers_field_error(sur):
    mx0, rms0 = sur.validate()
    sur.train(400, nwg=128, chunks=2, lr_max=3e-3, lr_min=3e-4)
    mx1, rms1 = sur.validate()
    assert rms1 < 0.5 * rms0, (rms0, rms1)
    assert mx1 < mx0


def test_kicks_with_surrogate(sur):
    """Electron-side kicks from the trained surrogate against the exact FP64
    kicks (same cloud): points outside the surrogate's domain get the exact
    kick bit for bit; inside, the kick error follows the field error."""
    from cetasim_amd import configs
    from cetasim_amd.tracker import Tracker

    sur.train(1500, nwg=128, chunks=2, lr_max=3e-3, lr_min=1e-4)
    mx, rms = sur.validate()
    ions = dict(mass_number=[28.0], cross_section=[2.0e-22], pressure=[1.0e-5], temperature=300.0,
                n_new=400, max_ions=50_000, loss_boundary=10.0, seed=11)
    p = configs.BASE
    ring, secs, _ = configs.lattice(p)
    counts = np.array([3000, 2000, 2500])
    gap = np.array([2, 2, 3], np.int32)
    q, ne = configs.bunch_scalars(counts, ring, p)
    parts = configs.train_bunches(counts, 12, p)
    out = {}
    for mode in (0, 2):
        tr = Tracker(ring, secs)
        tr.set_bunches(counts, q, ne, gap)
        tr.upload(parts)
        tr.set_ions(configs.ion_config(ring, ions, p))
        tr.set_field_model(sur if mode else None, mode)
        tr.ion_phase(0)
        out[mode] = (tr.download(), tr.ion_passages())
        tr.close()
    (d0, c0), (d2, c2) = out[0], out[2]
    # the ion side is exact in both runs: identical clouds
    assert np.array_equal(c0, c2)
    kick0 = np.stack([d0["px"] - parts["px"], d0["py"] - parts["py"]])
    kick2 = np.stack([d2["px"] - parts["px"], d2["py"] - parts["py"]])
    # domain of each electron w.r.t. the cloud seen by its bunch
    off = np.concatenate([[0], np.cumsum(counts)])
    inside = np.zeros(len(parts["x"]), bool)
    for b in range(3):
        _, mx_, my_, rx, ry, _ = c0[b, 0]
        sl = slice(off[b], off[b + 1])
        big, small = max(rx, ry), min(rx, ry)
        u = (parts["x"][sl] - mx_) / rx
        v = (parts["y"][sl] - my_) / ry
        inside[sl] = (np.abs(u) <= 6) & (np.abs(v) <= 6) & (np.abs(u) + np.abs(v) >= 1e-3) & (small / big >= 0.02)
    assert inside.sum() > 0.5 * len(inside)
    assert np.array_equal(kick0[:, ~inside], kick2[:, ~inside])
    rel = np.hypot(*(kick2 - kick0)[:, inside]) / np.hypot(*kick0[:, inside])
    # the validation set samples the domain uniformly; the electrons sit
    # near the cloud centre, where the network's error is largest
    assert np.max(rel) <= max(5 * mx, 0.05), (np.max(rel), mx)
    assert np.sqrt(np.mean(rel ** 2)) <= 5 * rms, (np.sqrt(np.mean(rel ** 2)), rms)
