"""The CPU restatement of the surrogate network: gradient against central
finite differences, output decoding against the field's closed limits."""
import numpy as np

import mlp_ref as M
import oracle as O


def rand_params(seed):
    rng = np.random.default_rng(seed)
    p = np.concatenate([rng.uniform(-0.6, 0.6, 192), rng.uniform(-0.2, 0.2, 64), rng.uniform(-0.125, 0.125, 4096),
                        rng.uniform(-0.2, 0.2, 64), rng.uniform(-0.125, 0.125, 128), rng.uniform(-0.1, 0.1, 2)])
    assert len(p) == M.NPARAM
    return p


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = rand_params(2)
    recs = np.concatenate([rng.uniform(-1, 1, (40, 3)), rng.normal(0, 1, (40, 2))], 1)
    norm = (0.1, 1.3, -0.2, 0.7)
    _, g = M.loss_grad(p, recs, norm)
    idx = np.concatenate([rng.choice(192, 8, False), 192 + rng.choice(64, 4, False),
                          256 + rng.choice(4096, 20, False), 4352 + rng.choice(64, 4, False),
                          4416 + rng.choice(128, 6, False), [4544, 4545]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_grad(pp, recs, norm)[0] - M.loss_grad(pm, recs, norm)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(fd)), (i, fd, g[i])


def test_gelu_derivative():
    z = np.linspace(-6, 6, 1001)
    _, dg = M.gelu(z)
    h = 1e-6
    fd = (M.gelu(z + h)[0] - M.gelu(z - h)[0]) / (2 * h)
    assert np.max(np.abs(fd - dg)) < 1e-8


def test_envelope_limits():
    """The horizontal envelope -> -1/(1+r) at the centre (the linear field
    of the beam), -> the line-charge field far out, and is exact for a
    round beam."""
    r = np.array([0.02, 0.3, 1.0])
    bx, by = M.envelope(np.zeros(3) + 1e-9, np.zeros(3) + 1e-9, r)
    assert np.allclose(bx, -1 / (1 + r)) and np.all(by == -1)
    L = O.lib()
    fx, fy = O.D(), O.D()
    for au, av in ((0.3, 0.2), (2.0, 1.5), (5.0, 4.0)):
        L.orc_gaussian_field(au, av, 1.0, 1.0, fx, fy)
        bx, _ = M.envelope(np.array([au]), np.array([av]), np.array([1.0]))
        assert abs(au * bx[0] - fx.value) <= 1e-14 * abs(fx.value)
    for r_, au in ((0.05, 80.0), (0.5, 60.0)):
        L.orc_ion_field(au, 0.0, 1.0, r_, fx, fy)
        bx, _ = M.envelope(np.array([au]), np.array([0.0]), np.array([r_]))
        assert abs(au * bx[0] - fx.value) <= 1e-3 * abs(fx.value)
