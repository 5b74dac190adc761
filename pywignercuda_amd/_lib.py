"""ctypes binding of libwignerhip.so (C ABI declared in include/wignerhip.h).

The library is loaded from this package directory only; there is no fallback
implementation.  A missing or stale library raises WignerHipLibraryError.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_void_p
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libwignerhip.so"

WH_C128, WH_C64, WH_C64_WIDE = 0, 1, 2
WH_FORWARD, WH_INVERSE = -1, 1
WH_SCHEME_LIE, WH_SCHEME_STRANG = 0, 1
WH_MAX_DEG = 8
WH_POT_TABLE = -1
WH_POT_EXPR = -2
WH_EXPR_MAX, WH_EXPR_STACK = 64, 16
WH_NOBS = 16
WH_NOALLOC_MAX_X = 2048
OBS_COLUMNS = ["X", "X2", "P", "P2", "XP", "dV", "PdV", "XdV", "H", "neg", "norm",
               "rawnorm_re", "rawnorm_im", "probsq"]


# ---------------------------------------------------------------- errors
# Codes 1-9 and class names mirror cufft_wrapper.py:15-73 (cufftExceptions).
class WignerHipError(Exception):
    """Error reported by libwignerhip (cufftError counterpart)."""


class WignerHipInvalidPlan(WignerHipError):
    """An invalid plan handle was passed."""


class WignerHipAllocFailed(WignerHipError):
    """Device memory allocation failed."""


class WignerHipInvalidType(WignerHipError):
    """Unsupported transform type / precision."""


class WignerHipInvalidValue(WignerHipError):
    """Invalid argument value (e.g. a bad pointer)."""


class WignerHipInternalError(WignerHipError):
    """Internal error."""


class WignerHipExecFailed(WignerHipError):
    """Execution on the GPU failed."""


class WignerHipSetupFailed(WignerHipError):
    """Library set-up failed."""


class WignerHipInvalidSize(WignerHipError):
    """Unsupported transform size."""


class WignerHipUnalignedData(WignerHipError):
    """Input or output is not aligned for the transform."""


class WignerHipInvalidDescriptor(WignerHipError):
    """Invalid propagator descriptor."""


class WignerHipLaunchFailed(WignerHipError):
    """A kernel launch was rejected."""


class WignerHipNoDevice(WignerHipError):
    """No HIP device is visible."""


class WignerHipLibraryError(WignerHipError):
    """libwignerhip.so is missing or cannot be loaded."""


EXCEPTIONS = {
    1: WignerHipInvalidPlan, 2: WignerHipAllocFailed, 3: WignerHipInvalidType,
    4: WignerHipInvalidValue, 5: WignerHipInternalError, 6: WignerHipExecFailed,
    7: WignerHipSetupFailed, 8: WignerHipInvalidSize, 9: WignerHipUnalignedData,
    10: WignerHipInvalidDescriptor, 11: WignerHipLaunchFailed, 12: WignerHipNoDevice,
}


# ---------------------------------------------------------------- descriptor
_DP = POINTER(c_double)


class Wigner2DDesc(ctypes.Structure):
    """wh_wigner2d_desc of include/wignerhip.h (field order and types must match)."""

    _fields_ = [
        ("struct_size", c_int), ("precision", c_int), ("x_dim", c_int), ("p_dim", c_int),
        ("scheme", c_int),
        ("dt", c_double), ("dx", c_double), ("dp", c_double), ("dtheta", c_double),
        ("dlambda", c_double), ("hbar", c_double), ("mass", c_double),
        ("kinetic_poly", c_int), ("kinetic_deg", c_int), ("kin_coef", c_double * (WH_MAX_DEG + 1)),
        ("pot_deg", c_int), ("pot_rows", c_int), ("pot_coef", _DP),
        ("theta_decay", _DP), ("lambda_decay", _DP), ("x_mask", _DP),
        ("damping", c_int), ("gamma_dt", c_double), ("sigma_p", _DP), ("theta_host", _DP),
        ("renormalize", c_int),
        ("gp_coeff", c_double), ("gp_rho_factor", c_double), ("lambda_host", _DP),
        ("x_host", _DP), ("p_host", _DP), ("dv_host", _DP), ("v0_host", _DP),
        ("dxdp", c_double), ("inv_2m", c_double), ("dx_host", c_double),
        ("pot_dv", _DP), ("odm_g", _DP), ("odm_x", _DP), ("odm_th", _DP), ("odm_sign", c_double),
        ("odm_channels", c_int), ("odm_weight1", c_double), ("odm_weight2", c_double),
        ("pot_expr", _DP), ("pot_expr_len", c_int), ("pot_expr_kvn", c_int), ("pot_t_step", c_double),
    ]


class EnsembleDesc(ctypes.Structure):
    """wh_ensemble_desc of include/wignerhip.h."""

    _fields_ = [
        ("struct_size", c_int), ("precision", c_int), ("scheme", c_int),
        ("dt", c_double), ("hbar", c_double), ("mass", c_double),
        ("kinetic_poly", c_int), ("kinetic_deg", c_int), ("kin_coef", c_double * (WH_MAX_DEG + 1)),
        ("damping", c_int), ("gamma_dt", c_double), ("renormalize", c_int), ("gp_coeff", c_double),
        ("odm_channels", c_int), ("odm_weight1", c_double), ("odm_weight2", c_double),
    ]


class MemberDesc(ctypes.Structure):
    """wh_member_desc of include/wignerhip.h."""

    _fields_ = [
        ("struct_size", c_int), ("x_dim", c_int), ("p_dim", c_int),
        ("dx", c_double), ("dp", c_double), ("dtheta", c_double), ("dlambda", c_double),
        ("pot_deg", c_int), ("pot_rows", c_int), ("pot_coef", _DP),
        ("theta_decay", _DP), ("lambda_decay", _DP), ("x_mask", _DP), ("sigma_p", _DP),
        ("theta_host", _DP), ("lambda_host", _DP),
        ("x_host", _DP), ("p_host", _DP), ("dv_host", _DP), ("v0_host", _DP),
        ("dxdp", c_double), ("inv_2m", c_double), ("dx_host", c_double), ("gp_rho_factor", c_double),
        ("pot_dv", _DP), ("odm_g", _DP), ("odm_x", _DP), ("odm_th", _DP), ("odm_sign", c_double),
        ("pot_expr", _DP), ("pot_expr_len", c_int), ("pot_expr_kvn", c_int), ("pot_t_step", c_double),
    ]


class DiracDesc(ctypes.Structure):
    """wh_dirac_desc of include/wignerhip.h."""

    _fields_ = [
        ("struct_size", c_int), ("precision", c_int), ("x_dim", c_int), ("p_dim", c_int),
        ("dt", c_double), ("dx", c_double), ("dp", c_double), ("dtheta", c_double), ("dlambda", c_double),
        ("mass", c_double), ("mass_half", c_double), ("c", c_double),
        ("pot_deg", c_int), ("pot_rows", c_int), ("pot_coef", _DP), ("pot_dv", _DP),
        ("theta_decay", _DP), ("x_mask", _DP), ("x_host", _DP), ("p_host", _DP), ("theta_host", _DP),
        ("dv0", _DP), ("vsum", _DP),
        ("gamma_dt", c_double), ("dxdp", c_double), ("c_host", c_double), ("mc2_host", c_double),
        ("transmission_x", c_double), ("antiparticle", c_int), ("energy", c_int),
        ("mass_dm", _DP), ("mass_sm", _DP), ("gp_coeff", c_double), ("lambda_host", _DP),
        ("a_deg", c_int * 3), ("a_coef", (c_double * (WH_MAX_DEG + 1)) * 3),
        ("odm", c_int), ("odm_gamma", c_double), ("lambda_bar", c_double),
        ("kvn", c_int), ("step_filter", c_int), ("a_coef_rows", _DP), ("a_rows", c_int),
        ("pot_expr", _DP), ("pot_expr_len", c_int), ("dpot_expr", _DP), ("dpot_expr_len", c_int),
        ("pot_t_step", c_double), ("mass_expr", _DP), ("mass_expr_len", c_int), ("mass_t_step", c_double),
        ("a_expr", _DP * 3), ("a_expr_len", c_int * 3), ("a_t_step", c_double),
    ]


class ObserveDesc(ctypes.Structure):
    """wh_observe_desc of include/wignerhip.h (pointers are device addresses)."""

    _fields_ = [
        ("struct_size", c_int), ("precision", c_int), ("x_dim", c_int), ("p_dim", c_int),
        ("x_host", c_void_p), ("p_host", c_void_p), ("dv_host", c_void_p), ("v0_host", c_void_p),
        ("dxdp", c_double), ("inv_2m", c_double),
    ]


class BlochDesc(ctypes.Structure):
    """wh_bloch_desc of include/wignerhip.h."""

    _fields_ = [
        ("struct_size", c_int), ("x_dim", c_int), ("p_dim", c_int),
        ("dx", c_double), ("dp", c_double), ("dtheta", c_double), ("dlambda", c_double), ("hbar", c_double),
        ("pot_deg", c_int), ("pot_rows", c_int), ("pot_coef", _DP), ("pot_sum", _DP),
        ("kinetic_deg", c_int), ("kin_coef", c_double * (WH_MAX_DEG + 1)),
        ("gp_coeff", c_double),
        ("theta_host", _DP), ("lambda_host", _DP), ("x_host", _DP), ("p_host", _DP),
        ("k_host", _DP), ("v0_host", _DP),
        ("dxdp", c_double), ("dx_host", c_double), ("dp_host", c_double), ("hbar_host", c_double),
        ("gp_coeff_host", c_double),
    ]


WH_BNHIST = 12
BLOCH_HIST = {"dt": 0, "X": 1, "X2": 2, "P": 3, "P2": 4, "H": 5, "E": 6, "NLE": 7, "purity": 8, "purity_im": 9,
              "kept": 10, "E_im": 11}

WH_DNOBS = 32
DIRAC_OBS = {"X": 0, "P": 2, "XP": 4, "XX": 6, "PP": 8, "Alpha1": 10, "D1V0": 12, "XD1V0": 14, "PD1V0": 16,
             "PAlpha1": 18, "XAlpha1": 20, "energy": 22}
DIRAC_OBS_REAL = {"negativity": 24, "transmission": 25, "antiparticle": 26, "norm": 27, "rawnorm": 28}

_lib = None


def _declare(lib):
    vp, ip = c_void_p, c_int
    sig = {
        "wh_last_error": ([], ctypes.c_char_p),
        "wh_status_string": ([ip], ctypes.c_char_p),
        "wh_version": ([], ip),
        "wh_fft_plan_many": ([POINTER(vp), ip, ip, ip, ip, ip], ip),
        "wh_cufft_plan_many": ([POINTER(vp), ip, vp, vp, ip, ip, vp, ip, ip, ip, ip], ip),
        "wh_fft_exec": ([vp, vp, vp, ip, vp], ip),
        "wh_fft_exec_normalized": ([vp, vp, vp, ip, vp], ip),
        "wh_fft_destroy": ([vp], ip),
        "wh_wigner2d_create": ([POINTER(Wigner2DDesc), POINTER(vp)], ip),
        "wh_wigner2d_destroy": ([vp], ip),
        "wh_wigner2d_run": ([vp, vp, ip, ip, vp, ip, ip, vp], ip),
        "wh_wigner2d_kinetic": ([vp, vp, vp], ip),
        "wh_wigner2d_potential": ([vp, vp, ip, ip, vp], ip),
        "wh_wigner2d_observe": ([vp, vp, vp, vp], ip),
        "wh_wigner2d_observe_noalloc": ([POINTER(ObserveDesc), vp, vp, vp], ip),
        "wh_wigner2d_ensemble": ([vp], vp),
        "wh_wigner2d_adjoint_step": ([vp, vp, vp, ip, vp, vp], ip),
        "wh_ensemble_create": ([POINTER(EnsembleDesc), POINTER(MemberDesc), ip, POINTER(vp)], ip),
        "wh_ensemble_destroy": ([vp], ip),
        "wh_ensemble_run": ([vp, POINTER(vp), ip, ip, vp, ip, ip, vp], ip),
        "wh_ensemble_set_variants": ([vp, ip, ip, ip], ip),
        "wh_ensemble_set_potential": ([vp, ip, vp, ip, ip, vp], ip),
        "wh_variant_info": ([ip, ip, POINTER(ip), POINTER(ip)], ip),
        "wh_bloch_create": ([POINTER(BlochDesc), POINTER(vp)], ip),
        "wh_bloch_destroy": ([vp], ip),
        "wh_bloch_run": ([vp, vp, vp, vp, ip, c_double, vp, vp], ip),
        "wh_dirac_create": ([POINTER(DiracDesc), POINTER(vp)], ip),
        "wh_dirac_destroy": ([vp], ip),
        "wh_dirac_run": ([vp, vp, ip, ip, vp, ip, vp], ip),
        "wh_dirac_kinetic": ([vp, vp, vp], ip),
        "wh_dirac_potential": ([vp, vp, ip, vp], ip),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


def lib():
    """The loaded library (loaded once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise WignerHipLibraryError(
                f"{LIB_PATH} not found: build it with `python -m pywignercuda_amd._build`")
        try:
            _lib = _declare(ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL))
        except OSError as e:  # pragma: no cover - depends on the environment
            raise WignerHipLibraryError(f"cannot load {LIB_PATH}: {e}") from e
    return _lib


def check(status: int, what: str = "") -> None:
    """Raises the exception mapped to a nonzero status (cufftCheckStatus, cufft_wrapper.py:67-73)."""
    if status != 0:
        msg = lib().wh_last_error().decode(errors="replace")
        exc = EXCEPTIONS.get(status, WignerHipError)
        raise exc(f"{what}: {msg}" if what else msg)


def last_error() -> str:
    """Message of the calling thread's last failed library call."""
    return lib().wh_last_error().decode(errors="replace")


def dptr(a) -> ctypes.c_void_p:
    """Device pointer of a torch tensor (no copy)."""
    return ctypes.c_void_p(int(a.data_ptr()))


def current_stream_handle():
    import torch

    return ctypes.c_void_p(int(torch.cuda.current_stream().cuda_stream))
