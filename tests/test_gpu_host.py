"""The tracking program (cetasim_amd/cetasim_run) on the GPU: its turn-by-turn
output and growth rates against the same run driven through the C ABI from
Python with the inputs the program's --dump wrote (same library, same
kernels: bit-identical).  The C-ABI turn itself is checked against the
oracle in test_gpu_parity.py and test_gpu_train.py.

The fixture (tests/golden/host) is 6 bunches x 8000 macroparticles in two
trains, 40 turns with ions, short-range wake and radiation; the dumped
ring, sections, bunches and settings are checked against the input in
tests/test_host.py.
"""
import ctypes
import os

import numpy as np
import pytest

from cetasim_amd import capi
from cetasim_amd.tracker import Tracker

import host_dump as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def program_run(tmp_path_factory):
    d = tmp_path_factory.mktemp("run")
    H.copy_fixture(str(d))
    (d / "dump").mkdir()
    r = H.run(["--dump", "dump", "input.dat"], str(d))
    assert r.returncode == 0, r.stderr
    print("cetasim_run input.dat (6 bunches x 8000, 40 turns)", flush=True)
    r = H.run(["input.dat"], str(d), timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "End of Tracking 40 Turns" in r.stdout
    return str(d), H.load_dump(str(d / "dump"))


def test_program_matches_capi_run(program_run):
    d, D = program_run
    s = D["settings"]
    tr = Tracker(D["ring"], D["sections"])
    tr.set_bunches(D["counts"], D["q"], D["ne"], D["gap"])
    tr.upload(D["parts"])
    tr.set_wake(s["nbins"], 0, 0.0, 0.0, s["rs"], s["q"], s["fr"])
    tr.set_synrad(True, 2, s["seed"])
    tr.set_ions(H.ion_config(s))
    tr.set_history(s["nTurns"])
    for n in range(s["nTurns"]):
        tr.turn(n)
    nb, interval = s["nb"], s["interval"]
    turns = list(range(0, s["nTurns"], interval))
    tbt = H.read_table(os.path.join(d, s["tbt"] + ".dat"))
    assert tbt.shape == (len(turns) * nb, 15)
    lib = capi.load()
    samples = []
    for i, t in enumerate(turns):
        m = (capi.Moments * nb)()
        capi.check(lib.cts_get_history(tr.h, t, 1, m))
        samples.append(m)
        for b in range(nb):
            row = tbt[i * nb + b]
            assert row[0] == t and row[1] == b
            mm = m[b]
            np.testing.assert_array_equal(row[2:8], np.array(mm.mean[:]))
            der = capi.Derived()
            capi.check(lib.cts_derive_moments(tr.h, ctypes.byref(mm), 0, ctypes.byref(der)))
            np.testing.assert_array_equal(row[8:14], [der.sigma_z, der.sigma_dp, der.rms_x, der.rms_y,
                                                      der.emit[0], der.emit[1]])
            assert row[14] == mm.transmission
    # growth rates of the program against the same fit of the same samples
    allm = (capi.Moments * (len(turns) * nb))()
    for i, m in enumerate(samples):
        for b in range(nb):
            allm[i * nb + b] = m[b]
    harm = np.ascontiguousarray(D["harm"], np.int32)
    tunes = np.array([D["ring"].nux, D["ring"].nuy, s["workQz"]])
    sec = D["sections"][0]
    twiss = np.array([sec.alpha_x, sec.beta_x, sec.alpha_y, sec.beta_y, s["betaZ"]])
    rates = np.empty(3 * nb)
    capi.check(lib.cts_cbm_growth_from_moments(allm, len(turns), nb, capi.i32ptr(harm), D["ring"].harmonic,
                                               capi.dptr(tunes), capi.dptr(twiss), interval, s["fit_start"],
                                               s["fit_end"], D["ring"].t0, capi.dptr(rates), None))
    gr = H.read_table(os.path.join(d, s["cbmgr"] + ".dat"))
    assert gr.shape == (nb, 4)
    np.testing.assert_array_equal(gr[:, 0], np.arange(nb))
    np.testing.assert_array_equal(gr[:, 1:].T.ravel(), rates)
    # the ions were on: the cloud is not empty after the run
    assert tr.ion_passages()[:, 0, 0].sum() > 0
    tr.close()
