"""cts_sur_eval_noalloc (no device allocation; static per-GPU staging)
against cts_sur_eval on the same points: bit-identical outputs on two calls
in a row, the second with fewer points (the buffer keeps the first call's
contents beyond them), plus the size limit's error."""

import numpy as np
import pytest

from cetasim_amd import capi
from cetasim_amd.surrogate import Surrogate

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 18


def features(n, seed):
    g = np.random.default_rng(seed)
    f = np.stack([g.uniform(0, 6, n), g.uniform(0, 6, n), np.exp(g.uniform(np.log(0.02), 0, n))], axis=1)
    return np.ascontiguousarray(f, dtype=np.float32)


def eval_noalloc(s, feats):
    out = np.empty((len(feats), 2), np.float32)
    capi.check_sur(s.lib.cts_sur_eval_noalloc(s.h, capi.fptr(feats), len(feats), capi.fptr(out)))
    return out


def test_noalloc_matches_eval_twice():
    s = Surrogate()
    s.init_params(7)
    for n, seed in ((MAX_POINTS, 1), (1000, 2)):
        f = features(n, seed)
        print(f"cts_sur_eval_noalloc: {n} points", flush=True)
        got = eval_noalloc(s, f)
        exp = s.eval(f)
        assert np.all(np.isfinite(exp))
        assert np.array_equal(got, exp), n
    s.close()


def test_noalloc_refuses_above_limit():
    s = Surrogate()
    s.init_params(7)
    f = features(4, 3)
    out = np.empty((4, 2), np.float32)
    # the size check comes before any copy: the arrays only need n rows for n <= limit
    rc = s.lib.cts_sur_eval_noalloc(s.h, capi.fptr(f), MAX_POINTS + 1, capi.fptr(out))
    assert rc == capi.CTS_ERR_ARG
    assert b"262144" in s.lib.cts_sur_last_error()
    s.close()
