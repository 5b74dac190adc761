"""ctypes binding of libwignerhip's FFT for the reference's own cufft_wrapper.py.

Pasted in place of cufft_wrapper.py:76-145 (the ``libcufft.so`` load and the bindings of
cufftExecZ2Z, cufftExecC2C, cufftDestroy and cufftPlanMany), these lines run the
reference's Plan_Z2Z / Plan_Z2Z_2D_Axis0 / Plan_Z2Z_2D_Axis1 classes and fft_Z2Z /
ifft_Z2Z (:147-250, :396-411) unchanged on libwignerhip.so:

* the plan call receives n, inembed and onembed as the addresses the classes pass
  (``n.ctypes.data``; the axis classes pass int64 arrays, whose first element reads
  correctly as an int on a little-endian host; inembed's contents are not read);
* the status codes are cuFFT's, so the module's own cufftCheckStatus (:67-73) raises
  its own exception classes;
* plan handles are pointers (c_void_p) where cuFFT's are unsigned ints;
* transforms run in place on the null stream, as cuFFT's do without cufftSetStream.

It uses the cufftCheckStatus of the module it is pasted into.  The library path comes
from WIGNERHIP_LIB (default: the loader's search path).
"""
import ctypes
import os

_libcufft = ctypes.CDLL(os.environ.get("WIGNERHIP_LIB", "libwignerhip.so"))

# Execution
_libcufft.wh_fft_exec.restype = ctypes.c_int
_libcufft.wh_fft_exec.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                  ctypes.c_void_p]


def cufftExecZ2Z(plan, idata, odata, direction):
    """Double precision complex-to-complex transform, in place."""
    cufftCheckStatus(_libcufft.wh_fft_exec(plan, idata, odata, direction, None))  # noqa: F821


def cufftExecC2C(plan, idata, odata, direction):
    """Single precision complex-to-complex transform, in place."""
    cufftCheckStatus(_libcufft.wh_fft_exec(plan, idata, odata, direction, None))  # noqa: F821


# Destroy plan
_libcufft.wh_fft_destroy.restype = ctypes.c_int
_libcufft.wh_fft_destroy.argtypes = [ctypes.c_void_p]


def cufftDestroy(plan):
    """Destroy an FFT plan."""
    cufftCheckStatus(_libcufft.wh_fft_destroy(plan))  # noqa: F821


# Plan for multiple dimensions (cufftPlanMany's argument list)
_libcufft.wh_cufft_plan_many.restype = ctypes.c_int
_libcufft.wh_cufft_plan_many.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int]


def cufftPlanMany(rank, n, inembed, istride, idist, onembed, ostride, odist, fft_type, batch):
    """Create a batched FFT plan."""
    plan = ctypes.c_void_p()
    status = _libcufft.wh_cufft_plan_many(ctypes.byref(plan), rank, n, inembed, istride, idist, onembed,
                                          ostride, odist, fft_type, batch)
    cufftCheckStatus(status)  # noqa: F821
    return plan
