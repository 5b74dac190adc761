"""Drop-in for the reference's cufft_wrapper.py, backed by libwignerhip's own FFT.

Same names and semantics as cufft_wrapper.py: status -> exception map
(:15-73), cufftPlanMany with its own argument list (:118-145, through
wh_cufft_plan_many), Plan_Z2Z (1-D and 2-D) / Plan_Z2Z_2D_Axis0 /
Plan_Z2Z_2D_Axis1 and their C2C siblings (:147-250), cufftExecZ2Z /
cufftExecC2C / cufftDestroy (:81-115), fft_Z2Z / ifft_Z2Z (:396-403).  In
place, inverse unnormalised.

Arrays are torch CUDA tensors (complex128 for Z2Z, complex64 for C2C) or raw
integer device addresses, as the reference passed ``int(gpuarray.gpudata)``.
Plans own their device twiddle table and free it on ``cufftDestroy`` or when
garbage collected.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

CUFFT_Z2Z = 0x69
CUFFT_C2C = 0x29
CUFFT_FORWARD = -1
CUFFT_INVERSE = 1

# exception classes under the reference's names
cufftError = _lib.WignerHipError
cufftInvalidPlan = _lib.WignerHipInvalidPlan
cufftAllocFailed = _lib.WignerHipAllocFailed
cufftInvalidType = _lib.WignerHipInvalidType
cufftInvalidValue = _lib.WignerHipInvalidValue
cufftInternalError = _lib.WignerHipInternalError
cufftExecFailed = _lib.WignerHipExecFailed
cufftSetupFailed = _lib.WignerHipSetupFailed
cufftInvalidSize = _lib.WignerHipInvalidSize
cufftUnalignedData = _lib.WignerHipUnalignedData
cufftExceptions = dict(_lib.EXCEPTIONS)


def cufftCheckStatus(status):
    """Raise the exception mapped to a nonzero status (cufft_wrapper.py:67-73)."""
    _lib.check(status)


class _Plan:
    """Owns one plan handle of libwignerhip (wh_cufft_plan_many's argument list, which is
    cufftPlanMany's: cufft_wrapper.py:118-145)."""

    def __init__(self, rank, n, inembed, istride, idist, fft_type, batch):
        self.fft_type = fft_type
        self.batch = batch
        nn = np.ascontiguousarray(np.asarray(n, dtype=np.int32).reshape(-1))
        emb = None if inembed is None else np.ascontiguousarray(np.asarray(inembed, dtype=np.int32).reshape(-1))
        addr = lambda a: None if a is None else ctypes.c_void_p(a.ctypes.data)  # noqa: E731
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().wh_cufft_plan_many(ctypes.byref(h), int(rank), addr(nn), addr(emb), int(istride),
                                                 int(idist), addr(emb), int(istride), int(idist), int(fft_type),
                                                 int(batch)), "cufftPlanMany")
        self.handle = h

    def destroy(self):
        if self.handle is not None and self.handle.value:
            _lib.check(_lib.lib().wh_fft_destroy(self.handle), "cufftDestroy")
        self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class _RawPlan(_Plan):
    """A plan created from raw cufftPlanMany arguments: n, inembed and onembed as addresses
    of int arrays (what the reference passes, e.g. ``n.ctypes.data``) or as sequences."""

    def __init__(self, rank, n, inembed, istride, idist, onembed, ostride, odist, fft_type, batch):
        self.fft_type = fft_type
        self.batch = batch
        held = [_int_array(a) for a in (n, inembed, onembed)]  # kept alive over the call
        ptr = [None if a is None else ctypes.c_void_p(a if isinstance(a, int) else a.ctypes.data) for a in held]
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().wh_cufft_plan_many(ctypes.byref(h), int(rank), ptr[0], ptr[1], int(istride), int(idist),
                                                 ptr[2], int(ostride), int(odist), int(fft_type), int(batch)),
                   "cufftPlanMany")
        self.handle = h


def _int_array(a):
    """None, an address (int), or an int32 copy of a sequence."""
    if a is None or isinstance(a, int):
        return a
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))


def cufftPlanMany(rank, n, inembed, istride, idist, onembed, ostride, odist, fft_type, batch):
    """Batched plan with cufftPlanMany's arguments (cufft_wrapper.py:131-145): n, inembed and
    onembed as addresses of int arrays (what the reference passes) or sequences."""
    return _RawPlan(rank, n, inembed, istride, idist, onembed, ostride, odist, fft_type, batch)


class Plan_Z2Z(_Plan):
    """complex128 plan (cufft_wrapper.py:147-180): shape (n,) or (n0, n1), batch contiguous
    transforms, cuFFT's basic layout."""

    def __init__(self, shape, batch=1, fft_type=CUFFT_Z2Z):
        if len(shape) == 0:
            raise ValueError("invalid transform size")
        self.shape = shape
        super().__init__(len(shape), shape, None, 1, 0, fft_type, batch)


class Plan_Z2Z_2D_Axis0(_Plan):
    """Transform along axis 0 of a (P, X) array (cufft_wrapper.py:183-215)."""

    def __init__(self, shape, fft_type=CUFFT_Z2Z):
        if len(shape) != 2:
            raise ValueError("invalid transform dimension")
        self.shape = shape
        super().__init__(1, [shape[0]], [shape[0], shape[1]], shape[1], 1, fft_type, shape[1])


class Plan_Z2Z_2D_Axis1(_Plan):
    """Transform along axis 1 of a (P, X) array (cufft_wrapper.py:218-250)."""

    def __init__(self, shape, fft_type=CUFFT_Z2Z):
        if len(shape) != 2:
            raise ValueError("invalid transform dimension")
        self.shape = shape
        super().__init__(1, [shape[1]], list(shape), 1, shape[1], fft_type, shape[0])


class Plan_C2C(Plan_Z2Z):
    def __init__(self, shape, batch=1):
        super().__init__(shape, batch, CUFFT_C2C)


class Plan_C2C_2D_Axis0(Plan_Z2Z_2D_Axis0):
    def __init__(self, shape):
        super().__init__(shape, CUFFT_C2C)


class Plan_C2C_2D_Axis1(Plan_Z2Z_2D_Axis1):
    def __init__(self, shape):
        super().__init__(shape, CUFFT_C2C)


def _addr(a):
    if isinstance(a, int):
        return ctypes.c_void_p(a)
    return _lib.dptr(a)


def _exec(plan, idata, odata, direction, normalized=False):
    h = plan.handle if isinstance(plan, _Plan) else plan
    f = _lib.lib().wh_fft_exec_normalized if normalized else _lib.lib().wh_fft_exec
    _lib.check(f(h, _addr(idata), _addr(odata), int(direction), _lib.current_stream_handle()), "cufftExec")


def cufftExecZ2Z(plan, idata, odata, direction):
    """cufft_wrapper.py:87-91."""
    _exec(plan, idata, odata, direction)


def cufftExecC2C(plan, idata, odata, direction):
    """cufft_wrapper.py:99-103."""
    _exec(plan, idata, odata, direction)


def cufftDestroy(plan):
    """cufft_wrapper.py:109-115 (accepts a plan object or its handle)."""
    if isinstance(plan, _Plan):
        plan.destroy()
    else:
        _lib.check(_lib.lib().wh_fft_destroy(plan), "cufftDestroy")


def fft_Z2Z(x_input_gpu, y_output_gpu, plan):
    """Forward transform, in place (cufft_wrapper.py:396-399)."""
    _exec(plan, x_input_gpu, y_output_gpu, CUFFT_FORWARD)


def ifft_Z2Z(x_input_gpu, y_output_gpu, plan):
    """Inverse transform, unnormalised, in place (cufft_wrapper.py:400-403)."""
    _exec(plan, x_input_gpu, y_output_gpu, CUFFT_INVERSE)


def ifft_Z2Z_normalized(x_input_gpu, y_output_gpu, plan):
    """Inverse transform with the callers' 1/n folded in (GPU_Wigner2D_FFT.py:790-801)."""
    _exec(plan, x_input_gpu, y_output_gpu, CUFFT_INVERSE, normalized=True)


fft_C2C = fft_Z2Z
ifft_C2C = ifft_Z2Z
