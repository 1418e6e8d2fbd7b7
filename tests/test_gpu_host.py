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


PIPE = ("radiusX radiusY length betaX betaY num matSigma\n"
        "0.01 0.01 2304.0 2.71263 4.25053 1 3.5e+5\n"
        "0.03 0.012 300 5 8 2 1.4e6\n")
LRW_EDITS = {"runLongRangeWake": "1", "LRWPipeGeoInput": "pipe.dat", "LRWBBRInput": "bbr.dat",
             "LRWNTurnsWakeTrunction": "3", "LRWOutput": "lrw.sdds"}
BL_EDITS = {"rfMode": "1", "rfResType": "1", "rfResShuntImpRs": "5.0e6", "rfResQualityQ0": "2.0e4",
            "rfResCouplingBeta": "3.0", "rfResDetuneFre": "-2.0e4", "rfResCold": "0", "rfBunchBinNum": "60"}


def run_program(d, edits=None):
    H.copy_fixture(str(d), edits=edits)
    if edits:
        (d / "pipe.dat").write_text(PIPE)
        (d / "imp.dat").write_text(H.impedance_file()[0])
        (d / "wake.dat").write_text(H.wake_file()[0])
    (d / "dump").mkdir()
    r = H.run(["--dump", "dump", "input.dat"], str(d))
    assert r.returncode == 0, r.stderr
    print("cetasim_run input.dat (6 bunches x 8000, 40 turns)", flush=True)
    r = H.run(["input.dat"], str(d), timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "End of Tracking 40 Turns" in r.stdout
    return str(d), H.load_dump(str(d / "dump"))


@pytest.fixture(scope="module")
def program_run(tmp_path_factory):
    return run_program(tmp_path_factory.mktemp("run"))


@pytest.fixture(scope="module")
def program_run_lrw(tmp_path_factory):
    return run_program(tmp_path_factory.mktemp("run_lrw"), LRW_EDITS)


BBI_EDITS = {"runBBIFlag": "1", "BBIimpedInput": "imp.dat", "BBIimpedSimLandTFlag": "1 1 1 1 1",
             "BBIimpedSimTimeOrFreFlag": "0"}
BBI_TD_EDITS = {"runBBIFlag": "1", "BBIWakeInput": "wake.dat", "BBIimpedSimLandTFlag": "1 1 1 1 1"}


@pytest.fixture(scope="module")
def program_run_bbi(tmp_path_factory):
    return run_program(tmp_path_factory.mktemp("run_bbi"), BBI_EDITS)


@pytest.fixture(scope="module")
def program_run_bbi_td(tmp_path_factory):
    return run_program(tmp_path_factory.mktemp("run_bbi_td"), BBI_TD_EDITS)


@pytest.fixture(scope="module")
def program_run_bl(tmp_path_factory):
    return run_program(tmp_path_factory.mktemp("run_bl"), BL_EDITS)


BL_WALK_EDITS = dict(BL_EDITS, rfResExciteIntability="1", rfResDirFB="1", rfResDirFBGain="0.3",
                     rfResDirFBPhaseShift="0.2", rfResDirFBDelay="1e-10")


@pytest.fixture(scope="module")
def program_run_bl_walk(tmp_path_factory):
    return run_program(tmp_path_factory.mktemp("run_bl_walk"), BL_WALK_EDITS)


def read_sdds_pages(path):
    """Data pages of an ASCII SDDS file as lists of token rows (header up to
    the &data line skipped; '! page number' lines start a page; parameter
    values written between &data and the first page line are skipped)."""
    pages, cur, data = [], None, False
    for line in open(path):
        if not data:
            data = line.startswith("&data")
            continue
        if line.startswith("! page number"):
            cur = []
            pages.append(cur)
            continue
        if line.strip() and cur is not None:
            cur.append(line.split())
    return pages


def test_program_matches_capi_run(program_run):
    check_against_capi(*program_run)


def test_program_with_lrwake_matches_capi_run(program_run_lrw):
    """The same with the long-range wake on (resistive wall of a round and
    an elliptic sector, the fixture's resonators); the program's wake-table
    file against the library's table."""
    d, D = program_run_lrw
    assert D["settings"]["lrw"] == 1 and D["settings"]["lrw_nturns"] == 4
    check_against_capi(d, D)


def test_program_with_beam_loading_matches_capi_run(program_run_bl):
    """The same with the cavity beam-loaded (rfMode 1, active, detuned):
    the cavity, generator and beam-induced columns of the turn-by-turn file
    against the library's cavity tables."""
    d, D = program_run_bl
    assert D["settings"]["bl"] == 1 and D["ring"].n_rf == 0
    check_against_capi(d, D)


def test_program_with_walked_cavity_matches_capi_run(program_run_bl_walk):
    """The same with the cavity's bin-by-bin variant and its direct
    feedback on (the state walked on the device every turn)."""
    d, D = program_run_bl_walk
    assert D["settings"]["bl_int"][0][4:] == [1, 1]
    check_against_capi(d, D)


def test_program_with_bbimp_matches_capi_run(program_run_bbi):
    """The same with the broadband impedance on (all five planes:
    longitudinal, dipolar and quadrupolar, resonator tables in the
    reference's file layout)."""
    d, D = program_run_bbi
    assert D["settings"]["bbi"] == 1
    check_against_capi(d, D)


def test_program_with_time_domain_bbimp_matches_capi_run(program_run_bbi_td):
    """The same with the time-domain broadband impedance (the reference's
    default method; the wake file read as GetQuasiWakePoten does, all five
    planes)."""
    d, D = program_run_bbi_td
    assert D["settings"]["bbi"] == 1 and D["settings"]["bbi_td"] == 1
    check_against_capi(d, D)


def check_against_capi(d, D):
    s = D["settings"]
    nb, interval, nturns = s["nb"], s["interval"], s["nTurns"]
    tr = Tracker(D["ring"], D["sections"])
    tr.set_bunches(D["counts"], D["q"], D["ne"], D["gap"])
    tr.upload(D["parts"])
    tr.set_wake(s["nbins"], 0, 0.0, 0.0, s["rs"], s["q"], s["fr"])
    tr.set_synrad(True, 2, s["seed"])
    tr.set_ions(H.ion_config(s))
    if s.get("lrw"):
        tr.set_lrwake(s["lrw_nturns"], np.array(s["lrw_sectors"]).reshape(-1, 7),
                      np.array(s["lrw_res"]).reshape(-1, 9), *s["lrw_beta"])
        t, w, rw, bbr = tr.lrwake_table()
        pages = read_sdds_pages(os.path.join(d, "lrw.sdds"))
        rows = pages[0]
        assert int(rows[0][0]) == len(t)
        g = lambda v: float("%g" % v)  # noqa: E731
        for i in (0, 1, 2, len(t) // 2, len(t) - 1):
            row = [float(x) for x in rows[1 + i]]
            assert row[0] == g(-t[i])
            assert row[1:] == [g(v) for v in np.concatenate([rw[:, i], bbr[:, i], w[:, i]])]
    if s.get("bbi") and s["bbi_td"] == 1:
        wk = [s[k] if s["bbi_flags"][i] else None
              for i, k in enumerate(("bbi_wz", "bbi_wdx", "bbi_wdy", "bbi_wqx", "bbi_wqy"))]
        tr.set_bbimp_td(s["bbi_wpos"], wk, beta=tuple(s["bbi_beta"]))
    elif s.get("bbi"):
        tr.set_bbimp(s["bbi_re"], s["bbi_im"], s["bbi_fmax"], flags=s["bbi_flags"], time_domain=s["bbi_td"],
                     zqx=(s["bbi_qx_re"], s["bbi_qx_im"]), zqy=(s["bbi_qy_re"], s["bbi_qy_im"]),
                     zdx=(s["bbi_dx_re"], s["bbi_dx_im"]), zdy=(s["bbi_dy_re"], s["bbi_dy_im"]),
                     beta=tuple(s["bbi_beta"]))
    if s.get("bl"):
        nres = s["bl_nres"]
        bi, br = np.array(s["bl_int"]), np.array(s["bl_real"])
        tr.set_beam_loading(dict(nres=nres, mode=bi[:, 0], type=bi[:, 1], harm=bi[:, 2], cold=bi[:, 3],
                                 excite=bi[:, 4], dirfb=bi[:, 5], rs=br[:, 0], q0=br[:, 1], coupling=br[:, 2],
                                 detune=br[:, 3], vol=br[:, 4], phase=br[:, 5], fb_gain=br[:, 6],
                                 fb_phase=br[:, 7], fb_delay=br[:, 8], nbins=s["bl_nbins"],
                                 nturns=s["bl_nturns"]))
    # the program's turn loop with result.sdds on: each turn, then the
    # end-of-turn moments (the pending kicks applied), eigen-emittances and
    # ion charge
    turns = list(range(0, nturns, interval))
    lib = capi.load()
    ionq, samples, rows = [], [], []
    for t in range(nturns):
        tr.turn(t)
        m = (capi.Moments * nb)()
        capi.check(lib.cts_rms(tr.h, m))
        eig = tr.eigen_emit()
        ionc = ctypes.c_double()
        capi.check(lib.cts_get_ion_charge(tr.h, ctypes.byref(ionc)))
        rows.append((m, eig, ionc.value))
        if t % interval == 0:
            ionq.append(tr.ion_passages()[:, :, 5].sum(axis=1))
            samples.append(m)
    check_result_sdds(d, D, tr, rows)
    # TBTBunchAverData.sdds: one page per print turn, one row per bunch
    pages = read_sdds_pages(os.path.join(d, s["tbt"] + ".sdds"))
    assert len(pages) == len(turns)
    for i, t in enumerate(turns):
        assert int(pages[i][0][0]) == nb
        rows = pages[i][1:]
        assert len(rows) == nb
        for b in range(nb):
            row = rows[b]
            mm = samples[i][b]
            der = capi.Derived()
            capi.check(lib.cts_derive_moments(tr.h, ctypes.byref(mm), 0, ctypes.byref(der)))
            assert (int(row[0]), int(row[1]), int(row[2])) == (t, int(D["harm"][b]), b)
            exp = [ionq[i][b], mm.mean[0], mm.mean[2], mm.mean[4], mm.mean[1], mm.mean[3], mm.mean[5],
                   der.emit[0], der.emit[1], der.emit[2], der.rms_x, der.rms_y, der.sigma_z, der.sigma_dp]
            np.testing.assert_allclose([float(x) for x in row[3:17]], exp, rtol=2e-15, atol=0)
            assert len(row) == 17 + 7  # one cavity: 7 phasor columns
            if s.get("bl"):
                f = tr.cavity_info(t)[b, 0]
                cav, ind, gen = complex(f[0], f[1]), complex(f[2], f[3]), complex(f[4], f[5])
                exp = [abs(cav), np.angle(cav), cav.real, abs(gen), np.angle(gen), abs(ind), np.angle(ind)]
                np.testing.assert_allclose([float(x) for x in row[17:24]], exp, rtol=2e-15, atol=0)
    # runCBMGR.sdds: parameters and per-sample columns of every mode, printed
    # with six significant digits
    allm = (capi.Moments * (len(turns) * nb))()
    for i, m in enumerate(samples):
        for b in range(nb):
            allm[i * nb + b] = m[b]
    harm = np.ascontiguousarray(D["harm"], np.int32)
    tunes = np.array([D["ring"].nux, D["ring"].nuy, s["workQz"]])
    sec = D["sections"][0]
    twiss = np.array([sec.alpha_x, sec.beta_x, sec.alpha_y, sec.beta_y, s["betaZ"]])
    ns = len(turns)
    keys = ("ideal_amp", "ideal_arg", "iq_amp", "iq_arg", "ana_re", "ana_im", "hil_amp", "hil_arg")
    rk = ("rate_ideal", "rate_iq", "rate_hil", "rate_bunch")
    out = {k: np.zeros(ns * 3 * nb) for k in keys}
    out.update({k: np.zeros(3 * nb) for k in rk})
    res = capi.CbmResult(*[capi.dptr(out[k]) for k in keys + rk])
    capi.check(lib.cts_cbm_analysis(allm, ns, nb, capi.i32ptr(harm), D["ring"].harmonic, capi.dptr(tunes),
                                    capi.dptr(twiss), interval, s["fit_start"], s["fit_end"], D["ring"].t0,
                                    ctypes.byref(res)))
    A = {k: out[k].reshape(ns, 3, nb) for k in keys}
    R = {k: out[k].reshape(3, nb) for k in rk}
    g = lambda v: float("%g" % v)  # noqa: E731  (iostream default: 6 significant digits)
    cb = read_sdds_pages(os.path.join(d, s["cbmgr"] + ".sdds"))
    assert len(cb) == nb
    for i in range(nb):
        page = cb[i]
        dim = s["fit_end"] // interval - s["fit_start"] // interval
        assert int(page[0][0]) == dim and int(page[1][0]) == i
        par = [float(r[0]) for r in page[2:11]]
        exp = [g(R[k][pl, i]) for k in ("rate_ideal", "rate_iq", "rate_hil") for pl in range(3)]
        assert par == exp
        assert int(page[11][0]) == int(D["harm"][i])
        assert [float(r[0]) for r in page[12:15]] == [g(R["rate_bunch"][pl, i]) for pl in range(3)]
        assert int(page[15][0]) == ns
        for n in range(ns):
            row = [float(x) for x in page[16 + n]]
            assert row[0] == n * interval
            assert row[1:4] == [g(A["ideal_amp"][n, pl, i]) for pl in range(3)]
            assert row[7:10] == [g(A["iq_amp"][n, pl, i]) for pl in range(3)]
            assert row[19:22] == [g(samples[n][i].mean[2 * pl]) for pl in range(3)]
    # the ions were on: the cloud is not empty after the run
    assert tr.ion_passages()[:, 0, 0].sum() > 0
    tr.close()


def check_result_sdds(d, D, tr, rows):
    """result.sdds: the header's parameters and one row per turn (iostream
    default formatting, six significant digits) from the same end-of-turn
    moments, eigen-emittances and ion charge."""
    s = D["settings"]
    nb, nturns = s["nb"], s["nTurns"]
    nprint = s["tbt_print"]
    g = lambda v: float("%g" % v)  # noqa: E731
    lines = open(os.path.join(d, "result.sdds")).read().splitlines()
    cols = [ln for ln in lines if ln.startswith("&column")]
    assert len(cols) == 31 + 16 * nprint
    i0 = lines.index("&data mode=ascii, &end")
    params = [float(x) for x in lines[i0 + 1:i0 + 19]]
    assert params[0] == g(s["ring_current"])
    assert all(v == 0.0 for v in params[9:18])
    page = read_sdds_pages(os.path.join(d, "result.sdds"))[0]
    assert int(page[0][0]) == nturns
    lib = capi.load()
    for t in range(nturns):
        row = [float(x) for x in page[1 + t]]
        m, eig, ionc = rows[t]
        means = np.array([[m[b].mean[k] for k in (0, 2, 4, 1, 3, 5)] for b in range(nb)])
        exp = [t, g(ionc), g(means[:, 0].max()), g(means[:, 1].max())]
        exp += [g(v) for v in means.sum(axis=0) / nb] + [g(v) for v in np.sqrt((means ** 2).sum(axis=0) / nb)]
        assert row[:16] == exp, t
        assert row[19:31] == [0.0] * 12
        for j in range(nprint):
            der = capi.Derived()
            capi.check(lib.cts_derive_moments(tr.h, ctypes.byref(m[j]), 0, ctypes.byref(der)))
            r = row[31 + 16 * j:31 + 16 * (j + 1)]
            assert r[:6] == [g(m[j].mean[k]) for k in range(6)], (t, j)
            assert r[6:12] == [g(v) for v in (der.sigma_z, der.sigma_dp, der.rms_x, der.rms_y, der.emit[0],
                                              der.emit[1])], (t, j)
            assert r[12:14] == [g(eig[j, 0]), g(eig[j, 1])], (t, j)
            assert r[14] == g(m[j].transmission) and r[15] == g(eig[j, 3])
    # the ions were on: the bunches met ion charge; the column itself holds
    # the ions left after the last bunch (here the train gap drifts them
    # past the loss boundary), never a negative charge
    assert tr.ion_passages()[:, :, 5].sum() > 0
    assert all(float(page[1 + t][1]) >= 0 for t in range(nturns))
