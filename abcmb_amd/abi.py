"""ctypes mirror of ``include/abcmb.h`` and loaders for the native libraries.

The GPU library (``libabcmb_hip.so``) is mandatory for every hot-path call:
``hip_lib()`` raises when it is missing instead of falling back to any CPU
code.  The host background producer (``libabcmb_host.so``) is plain C++.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(_HERE, "libabcmb_hip.so")
HOST_LIB_PATH = os.path.join(_HERE, "libabcmb_host.so")

TAU_TAB_N = 10000
TAU_TAB_LO = -33.0
N_PT_FIELDS = 21
PT_FIELDS = (
    "delta_m", "theta_b_prime", "metric_eta", "metric_h_prime", "metric_eta_prime",
    "metric_alpha", "metric_alpha_prime",
    "cdm_delta", "b_delta", "b_theta",
    "g_delta", "g_theta", "g_sigma", "g_G0", "g_G2",
    "ur_delta", "ur_theta", "ur_sigma",
    "ncdm_delta", "ncdm_theta", "ncdm_sigma",
)
PT_INDEX = {n: i for i, n in enumerate(PT_FIELDS)}

_D = C.c_double
_I = C.c_int32
_P = C.c_void_p


class AbcmbCosmo(C.Structure):
    _fields_ = [(n, _D) for n in (
        "h", "omega_b", "omega_cdm", "omega_Lambda",
        "A_s", "n_s", "k_pivot", "TCMB0", "YHe",
        "N_ur", "T_ur", "N_ncdm", "T_ncdm", "m_ncdm",
        "R_nu", "R_b", "om", "omega_r", "omega_m",
        "tau0",
        "xe_x0", "xe_dx", "xe_lna_last", "xe_last",
        "tm_x0", "tm_dx", "tm_lna_last", "tm_last",
        "lna_rec", "lna_transfer_start", "z_reion")] + [
        ("fluid_par", _D * 8),
        ("has_ncdm", _I), ("n_q", _I), ("xe_n", _I), ("tm_n", _I),
        ("n_save", _I), ("nk", _I), ("species_off", _I), ("pad0", _I),
        ("tau_tab", _P), ("xe_tab", _P), ("tm_tab", _P),
        ("lna_save", _P), ("k", _P), ("pt", _P), ("pt_fluid", _P)]


class AbcmbPertItem(C.Structure):
    _fields_ = [("k", _D), ("lna_start", _D), ("rtol", _D), ("atol", _D),
                ("cosmo", _I), ("ik", _I), ("pad0", _I), ("pad1", _I)]


class AbcmbOdeSettings(C.Structure):
    _fields_ = [(n, _D) for n in ("pcoeff", "icoeff", "dcoeff", "dt0", "safety",
                                  "factormin", "factormax", "newton_kappa", "jac_offset")] + [
        (n, _I) for n in ("max_steps", "max_newton", "l_g", "l_pol", "l_ur", "l_ncdm", "n_q", "n", "n_fluid",
                          "pad0")]


N_TANGENT_DIRS = 5
FLUID_PARAMS = 8
COSMO_DOUBLES = 31 + FLUID_PARAMS
NO_LAMBDA, NO_CDM = 1, 2    # AbcmbCosmo.species_off bits
REC_STAGES = 7              # recombination stage ends (ABCMB_REC_STAGES)


class AbcmbReion(C.Structure):
    _fields_ = [(n, _D) for n in ("tau_reion", "Delta_z_reion", "z_reion_He", "Delta_z_reion_He", "exp_reion")]


class AbcmbTangent(C.Structure):
    _fields_ = [("par", _P), ("tau_tab", _P), ("xe_tab", _P), ("tm_tab", _P), ("pt", _P), ("pt_fluid", _P)]


class AbcmbCotangent(C.Structure):
    _fields_ = [("pt", _P), ("par", _P), ("tau_tab", _P), ("xe_tab", _P), ("tm_tab", _P)]


assert C.sizeof(AbcmbCosmo) == COSMO_DOUBLES * 8 + 8 * 4 + 7 * 8
assert AbcmbCosmo.fluid_par.offset == 31 * 8
assert C.sizeof(AbcmbTangent) == 48
assert C.sizeof(AbcmbPertItem) == 48

_host = None
_hip = None
_libs = {}        # path -> loaded library (builds for user-fluid sets live beside the default)
_last_hip = None  # the HIP library most recently handed out (check() reads its error)


class HipLibraryMissing(RuntimeError):
    pass


def host_lib(path: str = None):
    """The host producer; ``path``: a build for a user-fluid set (default: the standard one)."""
    global _host
    if path is not None and path != HOST_LIB_PATH:
        if path not in _libs:
            if not os.path.exists(path):
                raise RuntimeError(f"{path} is not built")
            _libs[path] = _host_setup(C.CDLL(path))
        return _libs[path]
    if _host is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise RuntimeError(f"{HOST_LIB_PATH} is not built; run __graft_entry__.build()")
        _host = _host_setup(C.CDLL(HOST_LIB_PATH))
    return _host


def _host_setup(lib):
    lib.abcmb_host_background.argtypes = [C.POINTER(AbcmbCosmo)] + [_D] * 7 + [_P]
    lib.abcmb_host_recomb_points.argtypes = []
    lib.abcmb_host_tm_points.argtypes = []
    lib.abcmb_host_los_background.argtypes = [C.POINTER(AbcmbCosmo), _P, C.c_int, _P]
    lib.abcmb_host_lna_start.argtypes = [C.POINTER(AbcmbCosmo), _P, C.c_int, _D, _D, _P]
    lib.abcmb_host_eval.argtypes = [C.POINTER(AbcmbCosmo), _P, C.c_int, _P]
    lib.abcmb_host_drag.argtypes = [C.POINTER(AbcmbCosmo), C.POINTER(_D), C.POINTER(_D)]
    lib.abcmb_host_recombination.argtypes = [C.POINTER(AbcmbCosmo), _D, _P]
    lib.abcmb_host_background_tangent.argtypes = ([C.POINTER(AbcmbCosmo), _D, _P] + [_D] * 6 + [_P, _P, C.c_int]
                                                  + [_P] * 4)
    lib.abcmb_host_los_background_tangent.argtypes = [C.POINTER(AbcmbCosmo)] + [_P] * 5 + [C.c_int, _P, _P]
    lib.abcmb_host_eval_tangent.argtypes = [C.POINTER(AbcmbCosmo)] + [_P] * 5 + [C.c_int, _P, _P]
    lib.abcmb_host_fluid_signature.restype = C.c_char_p
    lib.abcmb_host_fluid_signature.argtypes = []
    return lib


HIP_SYMBOLS = (
    "abcmb_last_error", "abcmb_device_init", "abcmb_pert_workspace_size", "abcmb_pert_fwd", "abcmb_pert_tangent",
    "abcmb_sources_workspace_size", "abcmb_sources", "abcmb_los_workspace_size", "abcmb_los_cl", "abcmb_apply_operator",
    "abcmb_los_tangent_workspace_size", "abcmb_los_tangent",
    "abcmb_lensing_clpp", "abcmb_lens_workspace_size", "abcmb_lens", "abcmb_pk_lin", "abcmb_pert_fwd_noalloc",
    "abcmb_pk", "abcmb_lens_dmat_size", "abcmb_lens_dmat", "abcmb_lens_mfma_workspace_size", "abcmb_lens_mfma",
    "abcmb_tune_set", "abcmb_tune_get", "abcmb_reionize", "abcmb_los_background",
    "abcmb_pert_adjoint_workspace_size", "abcmb_pert_adjoint", "abcmb_los_adjoint_workspace_size",
    "abcmb_los_adjoint", "abcmb_lens_adjoint_workspace_size", "abcmb_lens_adjoint", "abcmb_lensing_clpp_adjoint",
    "abcmb_fluid_signature",
)
TUNE_PERT_PERSISTENT, TUNE_LENS_GEMM_COLS = 0, 1
PK_M, PK_CB = 0, 1


def hip_lib(path: str = None):
    """Load libabcmb_hip.so (or, with ``path``, a build of it for a user-fluid
    set); raise (never fall back) when it is absent."""
    global _hip, _last_hip
    if path is not None and path != HIP_LIB_PATH:
        if path not in _libs:
            if not os.path.exists(path):
                raise HipLibraryMissing(f"{path} is missing: build the model's fluid library first")
            _libs[path] = _hip_setup(C.CDLL(path))
        _last_hip = _libs[path]
        return _last_hip
    if _hip is None:
        if not os.path.exists(HIP_LIB_PATH):
            raise HipLibraryMissing(
                f"{HIP_LIB_PATH} is missing: the perturbation/LOS path only runs through the "
                "HIP library (build it with __graft_entry__.build())")
        _hip = _hip_setup(C.CDLL(HIP_LIB_PATH))
    _last_hip = _hip
    return _hip


def _hip_setup(lib):
    lib.abcmb_last_error.restype = C.c_char_p
    lib.abcmb_pert_workspace_size.restype = C.c_size_t
    lib.abcmb_pert_workspace_size.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.abcmb_pert_fwd.argtypes = [_P, C.c_int, _P, C.c_int, C.POINTER(AbcmbOdeSettings), C.c_int,
                                   _P, _P, _P, _P, _P, C.c_size_t, _P]
    lib.abcmb_pert_fwd_noalloc.argtypes = [_P, C.c_int, _P, C.c_int, C.POINTER(AbcmbOdeSettings), C.c_int,
                                           _P, _P, _P, _P, _P]
    lib.abcmb_pert_tangent.argtypes = [_P, _P, C.c_int, _P, C.c_int, C.POINTER(AbcmbOdeSettings), C.c_int,
                                       _P, _P, _P, _P, C.c_size_t, _P]
    lib.abcmb_pert_adjoint_workspace_size.restype = C.c_size_t
    lib.abcmb_pert_adjoint_workspace_size.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(AbcmbOdeSettings)]
    lib.abcmb_pert_adjoint.argtypes = [_P, _P, C.c_int, _P, C.c_int, C.POINTER(AbcmbOdeSettings), C.c_int,
                                       C.c_int, _P, _P, _P, _P, C.c_size_t, _P]
    lib.abcmb_los_adjoint_workspace_size.restype = C.c_size_t
    lib.abcmb_los_adjoint_workspace_size.argtypes = [C.c_int] * 5
    lib.abcmb_los_adjoint.argtypes = [_P, C.c_int, C.c_int, C.c_int, _P, _P, C.c_int, _P, _P, _P, _P, C.c_int,
                                      C.c_int, _P, C.c_int, _P, _P, _P, _P, _P, C.c_size_t, _P]
    lib.abcmb_lens_adjoint_workspace_size.restype = C.c_size_t
    lib.abcmb_lens_adjoint_workspace_size.argtypes = [C.c_int] * 3
    lib.abcmb_lens_adjoint.argtypes = [C.c_int, _P, _P, C.c_int, _P, _P, C.c_int, _P, _P, _P, _P, C.c_size_t, _P]
    lib.abcmb_lensing_clpp_adjoint.argtypes = [_P, C.c_int, _P, C.c_int, _P, C.c_int, _P, _P, _P]
    lib.abcmb_sources_workspace_size.restype = C.c_size_t
    lib.abcmb_sources_workspace_size.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.abcmb_sources.argtypes = [_P, C.c_int, C.c_int, C.c_int, _P, _P, C.c_int, _P, _P, _P, C.c_size_t, _P]
    lib.abcmb_los_workspace_size.restype = C.c_size_t
    lib.abcmb_los_workspace_size.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.abcmb_los_cl.argtypes = [_P, C.c_int, _P, C.c_int, _P, _P, C.c_int, _P, _P, _P, C.c_int, C.c_int,
                                 _P, C.c_int, C.c_int, _P, _P, _P, C.c_size_t, _P]
    lib.abcmb_los_tangent_workspace_size.restype = C.c_size_t
    lib.abcmb_los_tangent_workspace_size.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    lib.abcmb_los_tangent.argtypes = [_P, C.c_int, _P, _P, _P, C.c_int, C.c_int, _P, _P, C.c_int, _P, _P, _P,
                                      C.c_int, C.c_int, _P, C.c_int, C.c_int, _P, _P, C.c_size_t, _P]
    lib.abcmb_apply_operator.argtypes = [_P, C.c_int, C.c_int, _P, C.c_int, _P, _P]
    lib.abcmb_lensing_clpp.argtypes = [_P, C.c_int, _P, C.c_int, _P, C.c_int, _P, _P]
    lib.abcmb_lens_workspace_size.restype = C.c_size_t
    lib.abcmb_lens_workspace_size.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.abcmb_lens.argtypes = [C.c_int, _P, _P, C.c_int, _P, _P, C.c_int, _P, _P, C.c_size_t, _P]
    lib.abcmb_pk_lin.argtypes = [_P, C.c_int, _P, C.c_int, _P, _P]
    lib.abcmb_pk.argtypes = [_P, C.c_int, _P, C.c_int, _D, C.c_int, _P, _P]
    lib.abcmb_lens_dmat_size.restype = C.c_size_t
    lib.abcmb_lens_dmat_size.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.abcmb_lens_dmat.argtypes = [_P, C.c_int, C.c_int, C.c_int, _P, C.c_size_t, _P]
    lib.abcmb_tune_set.argtypes = [C.c_int, C.c_int]
    lib.abcmb_reionize.argtypes = [_P, _P, C.c_int, _P, _P, C.c_int, _P, _P, _P, _P, _P, _P]
    lib.abcmb_los_background.argtypes = [_P, _P, C.c_int, _P, C.c_int, _P, _P, _P]
    lib.abcmb_tune_get.argtypes = [C.c_int]
    lib.abcmb_lens_mfma_workspace_size.restype = C.c_size_t
    lib.abcmb_lens_mfma_workspace_size.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    lib.abcmb_lens_mfma.argtypes = [C.c_int, _P, _P, C.c_int, _P, _P, C.c_int, _P, C.c_int, _P, _P, C.c_size_t,
                                    _P]
    lib.abcmb_fluid_signature.restype = C.c_char_p
    lib.abcmb_fluid_signature.argtypes = []
    return lib


def check(rc: int, what: str):
    if rc != 0:
        msg = ""
        if _last_hip is not None:
            m = _last_hip.abcmb_last_error()
            msg = m.decode() if m else ""
        raise RuntimeError(f"{what} failed with status {rc}: {msg}")
