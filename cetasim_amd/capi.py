"""ctypes view of the C ABI in include/cetasim_hip.h and include/cetasim_host.h.

The shared library cetasim_amd/libcetasim_hip.so is built in-tree by
``make -C cetasim_amd/csrc`` (``__graft_entry__.build()``).  Loading fails
loudly when it is missing: there is no fallback path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcetasim_hip.so")

D = ctypes.c_double
I32 = ctypes.c_int32
I64 = ctypes.c_int64
U64 = ctypes.c_uint64
PD = ctypes.POINTER(ctypes.c_double)
PI32 = ctypes.POINTER(ctypes.c_int32)
PI64 = ctypes.POINTER(ctypes.c_int64)
PF = ctypes.POINTER(ctypes.c_float)

CTS_OK = 0
CTS_ERR_ARG = -1
CTS_ERR_HIP = -2
CTS_ERR_STATE = -3
CTS_LOW_TRANSMISSION = 1
CTS_ION_OVERFLOW = 2
CTS_LRWAKE_TAU = 3
MAX_RF = 4
MAX_RES = 8
MAX_SPECIES = 4


class Ring(ctypes.Structure):
    _fields_ = [("circ", D), ("energy", D), ("rbeta", D), ("rgamma", D), ("t0", D), ("f0", D),
                ("harmonic", I32), ("n_rf", I32), ("u0", D), ("alphac", D * 3),
                ("nux", D), ("nuy", D), ("chrom", D * 2), ("adtx", D * 5), ("adty", D * 5),
                ("skew_k", D), ("synrad_tau", D * 3), ("emit_nat", D * 3),
                ("rf_vr", D * MAX_RF), ("rf_vi", D * MAX_RF), ("rf_harm", I32 * MAX_RF),
                ("ap_x", D), ("ap_y", D)]


class Section(ctypes.Structure):
    _fields_ = [("b1h1", D * 36), ("h1b1", D * 36), ("h2b2", D * 36),
                ("eta_x", D), ("etap_x", D), ("eta_y", D), ("etap_y", D),
                ("alpha_x", D), ("beta_x", D), ("alpha_y", D), ("beta_y", D),
                ("phase_adv_x", D), ("phase_adv_y", D)]


class Moments(ctypes.Structure):
    _fields_ = [("n_alive", D), ("mean", D * 6), ("c2", D * 9), ("zmin", D), ("zmax", D),
                ("loss_count", D), ("transmission", D)]


class Derived(ctypes.Structure):
    _fields_ = [("emit", D * 3), ("emit_eff", D * 3), ("rms_x", D), ("rms_y", D),
                ("rms_eff_x", D), ("rms_eff_y", D), ("sigma_z", D), ("sigma_dp", D)]


class Twiss(ctypes.Structure):
    _fields_ = [("beta_x", D), ("alpha_x", D), ("eta_x", D), ("etap_x", D),
                ("beta_y", D), ("alpha_y", D), ("eta_y", D), ("etap_y", D), ("beta_z", D)]


class RingInput(ctypes.Structure):
    _fields_ = [("circ", D), ("energy", D), ("u0", D), ("sdelta0", D), ("alphac", D * 3),
                ("nux", D), ("nuy", D), ("chrom", D * 2), ("adtx", D * 5), ("adty", D * 5),
                ("skew_k", D), ("synrad_tau", D * 3), ("natural_emit", D), ("coupling", D),
                ("harmonic", I32), ("rf_volt", D), ("ap_x", D), ("ap_y", D)]


class IonConfig(ctypes.Structure):
    _fields_ = [("nspec", I32), ("n_new", I32), ("max_ions", I64), ("mass_number", D * MAX_SPECIES),
                ("cross_section", D * MAX_SPECIES), ("pressure", D * MAX_SPECIES), ("temperature", D),
                ("length", D), ("loss_boundary", D), ("ap_x", D), ("ap_y", D), ("seed", U64)]


class LrwakeConfig(ctypes.Structure):
    _fields_ = [("nturns", I32), ("n_sectors", I32), ("sectors", PD), ("n_res", I32), ("res", PD),
                ("beta_x", D), ("beta_y", D)]


class BbimpConfig(ctypes.Structure):
    _fields_ = [("nimp", I32), ("f_max", D), ("zz_re", PD), ("zz_im", PD), ("flags", I32 * 5),
                ("time_domain", I32), ("zqx_re", PD), ("zqx_im", PD), ("zqy_re", PD), ("zqy_im", PD),
                ("beta_x", D), ("beta_y", D), ("zdx_re", PD), ("zdx_im", PD), ("zdy_re", PD), ("zdy_im", PD),
                ("nwake", I32), ("wake_pos", PD), ("wake_z", PD), ("wake_dx", PD), ("wake_dy", PD),
                ("wake_qx", PD), ("wake_qy", PD)]


class BeamLoadingConfig(ctypes.Structure):
    _fields_ = [("nres", I32), ("rf_mode", I32 * MAX_RF), ("res_type", I32 * MAX_RF), ("res_harm", I32 * MAX_RF),
                ("rs", D * MAX_RF), ("q0", D * MAX_RF), ("coupling", D * MAX_RF), ("detune_hz", D * MAX_RF),
                ("vol", D * MAX_RF), ("phase", D * MAX_RF), ("cold", I32 * MAX_RF),
                ("excite_instability", I32 * MAX_RF), ("dir_fb", I32 * MAX_RF), ("nbins", I32), ("nturns", I32),
                ("dir_fb_gain", D * MAX_RF), ("dir_fb_phase", D * MAX_RF), ("dir_fb_delay", D * MAX_RF)]


class CbmResult(ctypes.Structure):
    _fields_ = [(n, PD) for n in ("ideal_amp", "ideal_arg", "iq_amp", "iq_arg", "ana_re", "ana_im", "hil_amp",
                                  "hil_arg", "rate_ideal", "rate_iq", "rate_hil", "rate_bunch")]


class RingDerived(ctypes.Structure):
    _fields_ = [("eta", D), ("workqz", D), ("natural_bunch_length", D), ("beta_z", D)]


# every entry point the headers declare: name -> (restype, argtypes)
SIGNATURES = {
    "cts_create": (I32, [ctypes.POINTER(Ring), ctypes.POINTER(Section), I32, ctypes.POINTER(ctypes.c_void_p)]),
    "cts_destroy": (I32, [ctypes.c_void_p]),
    "cts_last_error": (ctypes.c_char_p, []),
    "cts_device_info": (I32, [ctypes.c_char_p, I32]),
    "cts_set_bunches": (I32, [ctypes.c_void_p, I32, PI64, PD, PD, PI32]),
    "cts_upload": (I32, [ctypes.c_void_p, PD, PD, PD, PD, PD, PD, PI32]),
    "cts_download": (I32, [ctypes.c_void_p, PD, PD, PD, PD, PD, PD, PI32]),
    "cts_set_wake": (I32, [ctypes.c_void_p, I32, I32, D, D, I32, PD, PD, PD]),
    "cts_set_synrad": (I32, [ctypes.c_void_p, I32, I32, U64]),
    "cts_push_noise": (I32, [ctypes.c_void_p, PD]),
    "cts_set_ions": (I32, [ctypes.c_void_p, ctypes.POINTER(IonConfig)]),
    "cts_set_ions_ip": (I32, [ctypes.c_void_p, I32, ctypes.POINTER(IonConfig)]),
    "cts_set_lrwake": (I32, [ctypes.c_void_p, ctypes.POINTER(LrwakeConfig)]),
    "cts_get_lrwake_table": (I32, [ctypes.c_void_p, PI32, PD, PD, PD, PD]),
    "cts_get_lrwake_kick": (I32, [ctypes.c_void_p, PD]),
    "cts_set_beam_loading": (I32, [ctypes.c_void_p, ctypes.POINTER(BeamLoadingConfig)]),
    "cts_get_cavity_info": (I32, [ctypes.c_void_p, I32, PD]),
    "cts_set_bbimp": (I32, [ctypes.c_void_p, ctypes.POINTER(BbimpConfig)]),
    "cts_get_bbimp_wake": (I32, [ctypes.c_void_p, ctypes.POINTER(I32), PD, PD]),
    "cts_ion_phase": (I32, [ctypes.c_void_p, I32]),
    "cts_ion_phase_ip": (I32, [ctypes.c_void_p, I32, I32]),
    "cts_get_ion_cloud": (I32, [ctypes.c_void_p, I32, I64, PI64, PD, PD, PD, PD]),
    "cts_get_ion_cloud_ip": (I32, [ctypes.c_void_p, I32, I32, I64, PI64, PD, PD, PD, PD]),
    "cts_get_ion_passages": (I32, [ctypes.c_void_p, PD]),
    "cts_get_ion_passages_ip": (I32, [ctypes.c_void_p, I32, PD]),
    "cts_get_ion_charge": (I32, [ctypes.c_void_p, PD]),
    "cts_turn": (I32, [ctypes.c_void_p, I32]),
    "cts_flush": (I32, [ctypes.c_void_p]),
    "cts_run_turns": (I32, [ctypes.c_void_p, I32, I32, I32]),
    "cts_autotune": (I32, [ctypes.c_void_p, PI32]),
    "cts_set_layout": (I32, [ctypes.c_void_p, I32, I32]),
    "cts_get_layout": (I32, [ctypes.c_void_p, PI32, PI32]),
    "cts_section_map": (I32, [ctypes.c_void_p, I32]),
    "cts_skew_quad": (I32, [ctypes.c_void_p]),
    "cts_rf_kick": (I32, [ctypes.c_void_p]),
    "cts_drift": (I32, [ctypes.c_void_p]),
    "cts_energy_loss": (I32, [ctypes.c_void_p]),
    "cts_synrad": (I32, [ctypes.c_void_p, I32]),
    "cts_sr_wake": (I32, [ctypes.c_void_p, PD]),
    "cts_mark_lost": (I32, [ctypes.c_void_p, PD]),
    "cts_rms": (I32, [ctypes.c_void_p, ctypes.POINTER(Moments)]),
    "cts_get_moments": (I32, [ctypes.c_void_p, ctypes.POINTER(Moments)]),
    "cts_set_history": (I32, [ctypes.c_void_p, I32]),
    "cts_get_history": (I32, [ctypes.c_void_p, I32, I32, ctypes.POINTER(Moments)]),
    "cts_get_wake_state": (I32, [ctypes.c_void_p, PI32, PD, PD]),
    "cts_derive_moments": (I32, [ctypes.c_void_p, ctypes.POINTER(Moments), I32, ctypes.POINTER(Derived)]),
    "cts_eigen_emit": (I32, [ctypes.c_void_p, PD]),
    "cts_sync": (I32, [ctypes.c_void_p]),
    "cts_set_timing": (I32, [ctypes.c_void_p, I32]),
    "cts_kernel_time": (I32, [ctypes.c_void_p, ctypes.c_char_p, PD, PI64]),
    "cts_debug_qe_normals": (I32, [ctypes.c_void_p, U64, I32, U64, I64, PD]),
    "cts_debug_faddeeva": (I32, [ctypes.c_void_p, I64, PD, PD, PD, PD]),
    "cts_debug_kick_field": (I32, [ctypes.c_void_p, I64, PD, PD]),
    "cts_sur_create": (I32, [ctypes.POINTER(ctypes.c_void_p)]),
    "cts_sur_destroy": (I32, [ctypes.c_void_p]),
    "cts_sur_last_error": (ctypes.c_char_p, []),
    "cts_sur_init_params": (I32, [ctypes.c_void_p, U64]),
    "cts_sur_set_params": (I32, [ctypes.c_void_p, PF, PF]),
    "cts_sur_get_params": (I32, [ctypes.c_void_p, PF, PF]),
    "cts_sur_generate": (I32, [ctypes.c_void_p, I64, I64, U64]),
    "cts_sur_get_samples": (I32, [ctypes.c_void_p, I64, I64, PF, PD]),
    "cts_sur_train": (I32, [ctypes.c_void_p, I32, I32, I32, ctypes.c_float, ctypes.c_float, U64, PF]),
    "cts_sur_gradient": (I32, [ctypes.c_void_p, PF, I64, PF, PF]),
    "cts_sur_eval": (I32, [ctypes.c_void_p, PF, I64, PF]),
    "cts_sur_eval_noalloc": (I32, [ctypes.c_void_p, PF, I64, PF]),
    "cts_sur_validate": (I32, [ctypes.c_void_p, PD, PD, PD]),
    "cts_set_field_model": (I32, [ctypes.c_void_p, ctypes.c_void_p, I32]),
    "cts_cbm_growth_from_moments": (I32, [ctypes.POINTER(Moments), I32, I32, PI32, I32, PD, PD, I32, I32, I32, D,
                                          PD, PD]),
    "cts_fit_wlinear": (I32, [PD, PD, PD, I32, PD, PD]),
    "cts_cbm_analysis": (I32, [ctypes.POINTER(Moments), I32, I32, PI32, I32, PD, PD, I32, I32, I32, D,
                               ctypes.POINTER(CbmResult)]),
    "cts_growth_rates": (I32, [ctypes.c_void_p, I32, I32, I32, I32, I32, D, D, PD]),
    "cts_build_section": (I32, [ctypes.POINTER(Twiss), ctypes.POINTER(Twiss), D, D, ctypes.POINTER(Section)]),
    "cts_derive_ring": (I32, [ctypes.POINTER(RingInput), ctypes.POINTER(Ring), ctypes.POINTER(RingDerived)]),
}

_lib = None


def load():
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C cetasim_amd/csrc` "
                           "(or __graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class CtsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cetasim_hip error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != CTS_OK:
        raise CtsError(rc, load().cts_last_error().decode())
    return rc


def check_sur(rc):
    if rc != CTS_OK:
        raise CtsError(rc, load().cts_sur_last_error().decode())
    return rc


def fptr(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(PF)


def dptr(a):
    return a.ctypes.data_as(PD)


def i32ptr(a):
    return a.ctypes.data_as(PI32)


def i64ptr(a):
    return a.ctypes.data_as(PI64)


def arr(ctype_array):
    return np.ctypeslib.as_array(ctype_array)
