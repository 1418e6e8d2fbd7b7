"""The host layer (cetasim_amd/host): the input.dat reader, Train,
LatticeInterActionPoint, WakeFunction and the initial distribution of
MPBunch, through the tracking program's host-only modes (--check, --dump).

Expected values come from the fixture's numbers by the reference's rules
(src/ReadInputSettings.cpp, src/Train.cpp, src/LatticeInterActionPoint.cpp,
src/WakeFunction.cpp, src/MPBeam.cpp:64-161, src/MPBunch.cpp:98-338) and,
for the distribution, from the moments of a truncated Gaussian.
"""
import ctypes
import math
import os

import numpy as np
import pytest

from cetasim_amd import capi

import host_dump as H

E_CHARGE = 1.602017733e-19  # include/Global.h:28
# per-axis variance factor of a 2-D standard normal cut at radius 3:
# E[R^2 | R^2 <= 9] / 2 with R^2 ~ Exp(mean 2)
CUT3 = (2.0 - 9.0 * math.exp(-4.5) / (1.0 - math.exp(-4.5))) / 2.0


@pytest.fixture(scope="module")
def fixture_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("host")
    H.copy_fixture(str(d))
    return str(d)


@pytest.fixture(scope="module")
def chk(fixture_dir):
    return H.check(fixture_dir)


def ring_input():
    ri = capi.RingInput()
    ri.circ, ri.energy, ri.u0, ri.sdelta0 = 2304.0, 6.0e9, 4.17e6, 8.97e-4
    ri.alphac[0] = 3.33e-5
    ri.nux, ri.nuy = 135.18, 86.27
    ri.chrom[0], ri.chrom[1] = 1.5, 1.0
    ri.synrad_tau[0], ri.synrad_tau[1], ri.synrad_tau[2] = 2310, 2880, 1642
    ri.natural_emit, ri.coupling = 19.4e-12, 0.1
    ri.harmonic, ri.rf_volt = 3840, 8.0e6
    ri.ap_x, ri.ap_y = 0.15, 0.05
    return ri


def test_ring_quantities(chk):
    lib = capi.load()
    ring, der = capi.Ring(), capi.RingDerived()
    capi.check(lib.cts_derive_ring(ctypes.byref(ring_input()), ctypes.byref(ring), ctypes.byref(der)))
    assert chk["f0"] == ring.f0 and chk["t0"] == ring.t0 and chk["rGamma"] == ring.rgamma
    assert chk["workQz"] == der.workqz and chk["naturalBunchLength"] == der.natural_bunch_length
    assert chk["harmonics"] == 3840
    # the input workQz (0.0048) is replaced by the RF value
    assert chk["workQz"] != 0.0048
    # bunchEmittance is split by kappa
    assert chk["emittanceX"] == pytest.approx(22e-12 / 1.1, rel=1e-15)
    assert chk["emittanceY"] == pytest.approx(22e-12 / 1.1 * 0.1, rel=1e-15)
    assert chk["nTurns"] == 40 and chk["seed"] == 7
    assert (chk["beamIonFlag"], chk["sRWakeFlag"], chk["synRadDampingFlag"]) == (1, 1, 1)


def test_filling_pattern(chk):
    # Train::Initial sums bunchNumberPerTrain[j] * (bunchGaps[i] + 1): train 1
    # starts at 2 * (1 + 1) + 1000 = 1004, not 2 * (3 + 1) + 1000
    assert chk["trainStart"] == [0, 1004]
    assert chk["bunchHarmNum"] == [0, 4, 1004, 1006, 1008, 1010]
    assert chk["bunchGap"] == [4, 1000, 2, 2, 2, 3840 - 1010]
    assert chk["normCurrent"] == [2, 1, 1, 3, 1, 1]
    cur = np.array([2, 1, 1, 3, 1, 1]) * 0.02 / 9
    np.testing.assert_allclose(chk["current"], cur, rtol=1e-15)
    ne = cur / chk["f0"] / E_CHARGE
    np.testing.assert_allclose(chk["electronNumPerBunch"], ne, rtol=1e-15)
    np.testing.assert_allclose(chk["macroEleCharge"], ne / 8000, rtol=1e-15)


def test_lattice_and_wake(chk):
    np.testing.assert_allclose(chk["phaseAdvX12"], [2 * math.pi * 135.18], rtol=1e-15)
    np.testing.assert_allclose(chk["phaseAdvY12"], [2 * math.pi * 86.27], rtol=1e-15)
    assert chk["interactionLength"] == [2304.0]
    assert chk["vacuumPressure"] == pytest.approx([1.0e-9 * 133.3224], rel=1e-15)
    assert chk["pipeAperatureX"] == [0.15]     # min(twiss 0.2, input 0.15)
    assert chk["pipeAperatureY"] == [0.05]     # min(twiss 0.05, input 0.10)
    assert chk["corssSectionEI"] == pytest.approx([2.0e-22], rel=1e-15)
    np.testing.assert_allclose(chk["lRs"], [1e4, 3e3])
    np.testing.assert_allclose(chk["lQ"], [1.0, 2.0])
    np.testing.assert_allclose(chk["lOmega"], [2 * math.pi * 1e10, 2 * math.pi * 5e9], rtol=1e-15)
    # the section map is built from the twiss columns in file order
    lib = capi.load()
    bz = chk["twissBetaZ"]
    t = capi.Twiss(2.71, 0.05, 0.001, 0.0002, 4.25, -0.03, 0.0, 0.0, bz)
    sec = capi.Section()
    capi.check(lib.cts_build_section(ctypes.byref(t), ctypes.byref(t), 2 * math.pi * 135.18,
                                     2 * math.pi * 86.27, ctypes.byref(sec)))
    assert list(sec.b1h1) == chk["b1h1"]
    assert bz == chk["naturalBunchLength"] / 8.97e-4


def test_initial_distribution(chk):
    m = np.array(chk["moments"]).reshape(6, 6, 2)  # bunch, coordinate, (mean, rms)
    n = 8000
    ex, ey = 22e-12 / 1.1, 22e-12 / 1.1 * 0.1
    bx, ax, dx, dpx = 2.71, 0.05, 0.001, 0.0002
    by, ay = 4.25, -0.03
    beta = math.sqrt(1 - (0.51099906e6 / 6e9) ** 2)
    sd = math.sqrt(CUT3) * 8.97e-4 / beta ** 2
    rms = np.array([
        math.sqrt(CUT3 * ex * bx + dx ** 2 * sd ** 2),
        math.sqrt((ax ** 2 * CUT3 * ex + ex) / bx + dpx ** 2 * sd ** 2),
        math.sqrt(CUT3 * ey * by),
        math.sqrt((ay ** 2 * CUT3 * ey + ey) / by),
        math.sqrt(CUT3) * 2.5e-3,
        sd,
    ])
    for b in range(6):
        np.testing.assert_allclose(m[b, :, 1], rms, rtol=0.035)
        assert np.all(np.abs(m[b, :, 0]) <= 4.5 * rms / math.sqrt(n))
    # bunches are independent draws
    assert len({m[b, 0, 0] for b in range(6)}) == 6


def test_two_sections(tmp_path):
    H.copy_fixture(str(tmp_path), {"RingTwiss": "twiss2.dat", "ringSectNum": "2"})
    c = H.check(str(tmp_path))
    np.testing.assert_allclose(c["phaseAdvX12"], [400.0, 2 * math.pi * 135.18 - 400.0], rtol=1e-15)
    np.testing.assert_allclose(c["phaseAdvY12"], [250.0, 2 * math.pi * 86.27 - 250.0], rtol=1e-15)
    assert c["interactionLength"] == [1000.0, 1304.0]
    assert c["vacuumPressure"] == pytest.approx([1.0e-9 * 133.3224, 2.0e-9 * 133.3224], rel=1e-15)
    assert c["pipeAperatureX"] == [0.15, 0.01]
    assert c["pipeAperatureY"] == [0.05, 0.1]
    assert len(c["b1h1"]) == 72


@pytest.mark.parametrize("edit,needle", [
    ({"ringADTX": "0 0    // two coefficients"}, "ringadtx"),
    ({"fpTrainGaps": "1000 2800"}, "filling pattern"),
    ({"ionCalIonGasPercent": "0.5"}, "gas"),
    ({"ringSectNum": "2"}, "does not match"),
    ({"BunchDistributionType": "4"}, "distribution type"),
    ({"fpSetBunchChargeIndex": "0 9"}, "not a bunch index"),
])
def test_input_errors(tmp_path, edit, needle):
    H.copy_fixture(str(tmp_path), edit)
    r = H.run(["--check", "input.dat"], str(tmp_path))
    assert r.returncode != 0
    assert needle in r.stderr.lower(), r.stderr


def test_last_occurrence_wins_and_case(tmp_path):
    # keys are case-insensitive and blanks inside a key are ignored; a later
    # line overrides an earlier one (the reference example files rely on it)
    H.copy_fixture(str(tmp_path), {"ring Current": "0.04"})
    c = H.check(str(tmp_path))
    np.testing.assert_allclose(sum(c["current"]), 0.04, rtol=1e-14)


def test_dump(tmp_path):
    H.copy_fixture(str(tmp_path))
    d = tmp_path / "dump"
    d.mkdir()
    r = H.run(["--dump", str(d), "input.dat"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    D = H.load_dump(str(d))
    s = D["settings"]
    assert (s["nb"], s["nsec"], s["ntot"]) == (6, 1, 48000)
    assert D["counts"].tolist() == [8000] * 6
    assert D["harm"].tolist() == [0, 4, 1004, 1006, 1008, 1010]
    assert D["gap"].tolist() == [4, 1000, 2, 2, 2, 2830]
    ring = D["ring"]
    assert ring.harmonic == 3840 and ring.n_rf == 1 and ring.nux == 135.18 and ring.chrom[0] == 1.5
    assert (ring.ap_x, ring.ap_y) == (0.15, 0.05)
    assert s["wake"] == 1 and s["nbins"] == 100
    np.testing.assert_allclose(s["fr"], [1e10, 5e9], rtol=1e-15)
    assert s["rs"] == [1e4, 3e3] and s["q"] == [1.0, 2.0]
    assert s["synrad"] == 1 and s["seed"] == 7 and s["ions"] == 1
    assert (s["ion_nspec"], s["ion_n_new"], s["ion_max"]) == (1, 300, 200000)
    assert s["ion_mass"][0] == 28.0 and s["ion_cross"][0] == pytest.approx(2e-22, rel=1e-15)
    assert s["ion_pressure"][0] == pytest.approx(1.333224e-7, rel=1e-15)
    assert (s["ion_temperature"], s["ion_length"], s["ion_loss"]) == (300.0, 2304.0, 10.0)
    assert s["ion_ap"] == [0.15, 0.05] and s["ion_seed"] == 7
    assert s["betaZ"] == 2.5e-3 / 8.97e-4
    # the dumped particles are the ones --check summarised
    c = H.check(str(tmp_path))
    m = np.array(c["moments"]).reshape(6, 6, 2)
    X = np.stack([D["parts"][k] for k in ("x", "px", "y", "py", "z", "dp")])
    for b in range(6):
        seg = X[:, b * 8000:(b + 1) * 8000]
        np.testing.assert_allclose(seg.mean(axis=1), m[b, :, 0], rtol=1e-12, atol=1e-20)


def test_lrwake_settings(tmp_path):
    """runLongRangeWake and the LRW* keys (src/ReadInputSettings.cpp:688-705,
    812-815): the stored truncation is the input plus one; the pipe file
    has one header line and 7 columns (src/WakeFunction.cpp:292-320), the
    resonator file the 12 header lines and 9 columns of the short-range
    one, frequencies in Hz; beta at the first twiss point."""
    H.copy_fixture(str(tmp_path), {"runLongRangeWake": "1", "LRWPipeGeoInput": "pipe.dat", "LRWBBRInput": "bbr.dat",
                                   "LRWNTurnsWakeTrunction": "3"})
    (tmp_path / "pipe.dat").write_text("radiusX radiusY length betaX betaY num matSigma\n"
                                       "0.01 0.02 2304.0 2.5 4.5 1 3.5e+7\n\n0.03 0.012 300 5 8 2 1.4e6\n")
    d = tmp_path / "dump"
    d.mkdir()
    r = H.run(["--dump", str(d), "input.dat"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    s = H.load_dump(str(d))["settings"]
    assert s["lrw"] == 1 and s["lrw_nturns"] == 4
    assert s["lrw_sectors"] == [0.01, 0.02, 2304.0, 2.5, 4.5, 1, 3.5e7, 0.03, 0.012, 300, 5, 8, 2, 1.4e6]
    np.testing.assert_allclose(s["lrw_res"], [1e4, 1, 1e10, 0, 1, 1e10, 0, 1, 1e10,
                                              3e3, 2, 5e9, 0, 1, 5e9, 0, 1, 5e9], rtol=1e-15)
    assert s["lrw_beta"] == [2.71, 4.25]  # betax, betay of the first row of twiss.dat
    # off by default; more than one interaction point is refused
    H.copy_fixture(str(tmp_path), {"runLongRangeWake": "1", "LRWPipeGeoInput": "pipe.dat", "RingTwiss": "twiss2.dat",
                                   "ringSectNum": "2"})
    (tmp_path / "pipe.dat").write_text("h\n0.01 0.01 1 1 1 1 1e7\n")
    r = H.run(["--dump", str(d), "input.dat"], str(tmp_path))
    assert r.returncode != 0 and "long-range wake" in r.stderr


def test_rf_modes(tmp_path):
    # rfMode 1 (beam loading): the cavity keys reach cts_set_beam_loading
    # (src/ReadInputSettings.cpp:194-368) and the ring keeps no ideal cavity
    BL = {"rfMode": "1", "rfResType": "1", "rfResShuntImpRs": "5.0e6", "rfResQualityQ0": "2.0e4",
          "rfResCouplingBeta": "3.0", "rfResDetuneFre": "-2.0e4", "rfResCold": "0", "rfBunchBinNum": "60"}
    H.copy_fixture(str(tmp_path), BL)
    d = tmp_path / "d0"
    d.mkdir()
    r = H.run(["--dump", str(d), "input.dat"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    dump = H.load_dump(str(d))
    s = dump["settings"]
    assert s["bl"] == 1 and s["bl_nres"] == 1 and s["bl_nbins"] == 60 and s["bl_nturns"] == s["nTurns"]
    assert s["bl_int"] == [[1, 1, 1, 0, 0, 0]]
    assert s["bl_real"] == [[5.0e6, 2.0e4, 3.0, -2.0e4, 8.0e6, 1.0, 0.0, 0.0, 0.0]]
    assert dump["ring"].n_rf == 0
    # the bin-by-bin variant and the direct feedback (src/ReadInputSettings.cpp:215-256)
    H.copy_fixture(str(tmp_path), dict(BL, rfResExciteIntability="1", rfResDirFB="1", rfResDirFBGain="0.5",
                                       rfResDirFBPhaseShift="0.1", rfResDirFBDelay="1e-9"))
    r = H.run(["--dump", str(d), "input.dat"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    s = H.load_dump(str(d))["settings"]
    assert s["bl_int"] == [[1, 1, 1, 0, 1, 1]]
    assert s["bl_real"] == [[5.0e6, 2.0e4, 3.0, -2.0e4, 8.0e6, 1.0, 0.5, 0.1, 1e-9]]
    # an unknown mode is refused; a mixed pair keeps the ideal one in the ring
    H.copy_fixture(str(tmp_path), {"rfMode": "2"})
    r = H.run(["--dump", str(d), "input.dat"], str(tmp_path))
    assert r.returncode != 0 and "rfmode" in r.stderr.lower()
    H.copy_fixture(str(tmp_path), {"rfResNum": "2", "rfMode": "1 0", "rfResHarm": "1 3",
                                   "rfResVol": "8.0E+6 2.0E+6", "rfResPhase": "1.0 -1.6", "rfResType": "1 0",
                                   "rfResShuntImpRs": "5e6 1e6", "rfResQualityQ0": "2e4 1e4",
                                   "rfResCouplingBeta": "3 0", "rfResDetuneFre": "-2e4 1e5", "rfResCold": "0 0"})
    r = H.run(["--dump", str(d), "input.dat"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    dump = H.load_dump(str(d))
    assert dump["settings"]["bl_int"] == [[1, 1, 1, 0, 0, 0], [0, 0, 3, 0, 0, 0]]
    ring = dump["ring"]
    assert ring.n_rf == 1 and ring.rf_harm[0] == 3
    np.testing.assert_allclose([ring.rf_vr[0], ring.rf_vi[0]], [2e6 * math.cos(-1.6), 2e6 * math.sin(-1.6)],
                               rtol=1e-15)
    # several ideal cavities keep their input phases
    # (CavityResonator::Initial, src/CavityResonator.cpp:60-139)
    H.copy_fixture(str(tmp_path), {"rfResNum": "2", "rfMode": "0 0", "rfResHarm": "1 3",
                                   "rfResVol": "8.0E+6 2.0E+6", "rfResPhase": "1.0 -1.6"})
    d = tmp_path / "d2"
    d.mkdir()
    r = H.run(["--dump", str(d), "input.dat"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    ring = H.load_dump(str(d))["ring"]
    assert ring.n_rf == 2 and list(ring.rf_harm)[:2] == [1, 3]
    np.testing.assert_allclose(list(ring.rf_vr)[:2], [8e6 * math.cos(1.0), 2e6 * math.cos(-1.6)], rtol=1e-15)
    np.testing.assert_allclose(list(ring.rf_vi)[:2], [8e6 * math.sin(1.0), 2e6 * math.sin(-1.6)], rtol=1e-15)


def test_exe_exists():
    assert os.access(H.exe(), os.X_OK)


def test_bbimp_settings(tmp_path):
    """runBBIFlag and the &BoardBandImpedance keys (src/ReadInputSettings.cpp:
    726-752, 808-811); the impedance file's Z_|| = column 2 - i column 7 and
    its last frequency (src/BoardBandImp.cpp:42-70)."""
    text, zre, zim, fmax = H.impedance_file(nimp=12, f_max=3.0e10)
    (tmp_path / "imp.dat").write_text(text)
    H.copy_fixture(str(tmp_path), {"runBBIFlag": "1", "BBIimpedInput": "imp.dat",
                                   "BBIimpedSimLandTFlag": "1 0 0 0 0", "BBIimpedSimTimeOrFreFlag": "0"})
    d = tmp_path / "dump"
    d.mkdir()
    r = H.run(["--dump", str(d), "input.dat"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    s = H.load_dump(str(d))["settings"]
    assert s["bbi"] == 1 and s["bbi_flags"] == [1, 0, 0, 0, 0] and s["bbi_td"] == 0
    assert s["bbi_fmax"] == 3.0e10
    np.testing.assert_array_equal(s["bbi_re"], zre)
    np.testing.assert_array_equal(s["bbi_im"], zim)
    # Z_Qx = -column 10 - i column 5 (src/BoardBandImp.cpp:63-64); beta at the first twiss row
    rows = [[float(v) for v in ln.split()] for ln in text.splitlines()[5:] if ln.strip()]
    np.testing.assert_array_equal(s["bbi_qx_re"], [-r[9] for r in rows])
    np.testing.assert_array_equal(s["bbi_qx_im"], [-r[4] for r in rows])
    np.testing.assert_array_equal(s["bbi_qy_re"], [-r[10] for r in rows])
    np.testing.assert_array_equal(s["bbi_qy_im"], [-r[5] for r in rows])
    # zDxImp = (-col 8, -col 3), zDyImp = (-col 9, -col 4) (src/BoardBandImp.cpp:65-66)
    np.testing.assert_array_equal(s["bbi_dx_re"], [-r[7] for r in rows])
    np.testing.assert_array_equal(s["bbi_dx_im"], [-r[2] for r in rows])
    np.testing.assert_array_equal(s["bbi_dy_re"], [-r[8] for r in rows])
    np.testing.assert_array_equal(s["bbi_dy_im"], [-r[3] for r in rows])
    assert s["bbi_beta"] == [2.71, 4.25]
    # a missing file is an error
    H.copy_fixture(str(tmp_path), {"runBBIFlag": "1", "BBIimpedInput": "none.dat", "BBIimpedSimTimeOrFreFlag": "0"})
    r = H.run(["--dump", str(d), "input.dat"], str(tmp_path))
    assert r.returncode != 0 and "impedance file" in r.stderr


def test_bbimp_time_domain_settings(tmp_path):
    """BBIimpedSimTimeOrFreFlag defaults to 1 (include/ReadInputSettings.h:
    213): the time domain, with the wake file BBIWakeInput read as
    GetQuasiWakePoten does (src/MPBeam.cpp:748-812); the program's table
    equals read_quasi_wake's and the oracle's restatement, bit for bit."""
    import oracle as O
    from cetasim_amd.tracker import read_quasi_wake
    text, rows = H.wake_file()
    (tmp_path / "wake.dat").write_text(text)
    d = tmp_path / "dump"
    d.mkdir()
    H.copy_fixture(str(tmp_path), {"runBBIFlag": "1", "BBIWakeInput": "wake.dat",
                                   "BBIimpedSimLandTFlag": "1 1 0 0 1"})
    r = H.run(["--dump", str(d), "input.dat"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    s = H.load_dump(str(d))["settings"]
    assert s["bbi"] == 1 and s["bbi_td"] == 1 and s["bbi_flags"] == [1, 1, 0, 0, 1]
    assert s["bbi_re"] == []  # the impedance file is not read
    pos, wakes = read_quasi_wake(str(tmp_path / "wake.dat"))
    rr = np.ascontiguousarray(np.array([[float(repr(v)) for v in r] for r in rows]))
    out = np.zeros(6 * len(rows))
    assert O.lib().orc_bbi_td_wake(len(rows), O.p(rr.ravel()), O.p(out)) == 40
    out = out.reshape(6, -1)
    np.testing.assert_array_equal(pos, out[0])
    np.testing.assert_array_equal(s["bbi_wpos"], out[0])
    for k, name in enumerate(("bbi_wz", "bbi_wdx", "bbi_wdy", "bbi_wqx", "bbi_wqy")):
        np.testing.assert_array_equal(wakes[k], out[1 + k])
        np.testing.assert_array_equal(s[name], out[1 + k])
    # pos starts at 0 with the file's step; wz behind z = 0 folds in the
    # Gaussian ahead of it
    assert pos[0] == 0 and abs(pos[1] - 5e-4) < 1e-18
    assert wakes[0][0] == -rows[40][1] and wakes[0][1] == -(rows[41][1] + rows[39][1])
    # a missing wake file is an error
    H.copy_fixture(str(tmp_path), {"runBBIFlag": "1", "BBIWakeInput": "none.dat"})
    r = H.run(["--dump", str(d), "input.dat"], str(tmp_path))
    assert r.returncode != 0 and "wake file" in r.stderr
