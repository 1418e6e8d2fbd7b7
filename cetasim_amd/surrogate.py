"""Python handle of the learned field surrogate (C ABI cts_sur_*)."""
import ctypes

import numpy as np

from . import capi

NPARAM = 4546


class Surrogate:
    """Device-side MLP, its training set and its Adam state."""

    def __init__(self):
        self.lib = capi.load()
        self.h = ctypes.c_void_p()
        capi.check_sur(self.lib.cts_sur_create(ctypes.byref(self.h)))

    def close(self):
        if self.h:
            self.lib.cts_sur_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_params(self, seed):
        capi.check_sur(self.lib.cts_sur_init_params(self.h, seed))

    def params(self):
        p = np.empty(NPARAM, np.float32)
        norm = np.empty(4, np.float32)
        capi.check_sur(self.lib.cts_sur_get_params(self.h, capi.fptr(p), capi.fptr(norm)))
        return p, norm

    def set_params(self, p, norm):
        p = np.ascontiguousarray(p, dtype=np.float32)
        norm = np.ascontiguousarray(norm, dtype=np.float32)
        assert p.size == NPARAM and norm.size == 4
        capi.check_sur(self.lib.cts_sur_set_params(self.h, capi.fptr(p), capi.fptr(norm)))

    def generate(self, n_train, n_val, seed):
        capi.check_sur(self.lib.cts_sur_generate(self.h, n_train, n_val, seed))

    def samples(self, first, n, train=True):
        if train:
            out = np.empty((n, 5), np.float32)
            capi.check_sur(self.lib.cts_sur_get_samples(self.h, first, n, capi.fptr(out), None))
        else:
            out = np.empty((n, 5), np.float64)
            capi.check_sur(self.lib.cts_sur_get_samples(self.h, first, n, None, capi.dptr(out)))
        return out

    def train(self, steps, nwg=256, chunks=4, lr_max=3e-3, lr_min=1e-5, seed=4):
        loss = ctypes.c_float()
        capi.check_sur(self.lib.cts_sur_train(self.h, steps, nwg, chunks, lr_max, lr_min, seed, ctypes.byref(loss)))
        return loss.value

    def gradient(self, recs):
        recs = np.ascontiguousarray(recs, dtype=np.float32)
        g = np.empty(NPARAM, np.float32)
        loss = ctypes.c_float()
        capi.check_sur(self.lib.cts_sur_gradient(self.h, capi.fptr(recs), len(recs), capi.fptr(g), ctypes.byref(loss)))
        return loss.value, g

    def eval(self, feats):
        feats = np.ascontiguousarray(feats, dtype=np.float32)
        out = np.empty((len(feats), 2), np.float32)
        capi.check_sur(self.lib.cts_sur_eval(self.h, capi.fptr(feats), len(feats), capi.fptr(out)))
        return out

    def validate(self, per_point=False, n_val=None):
        mx, rms = ctypes.c_double(), ctypes.c_double()
        err = np.empty(n_val) if per_point else None
        capi.check_sur(self.lib.cts_sur_validate(self.h, ctypes.byref(mx), ctypes.byref(rms),
                                                 capi.dptr(err) if per_point else None))
        return (mx.value, rms.value, err) if per_point else (mx.value, rms.value)
