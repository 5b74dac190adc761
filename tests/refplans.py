"""The argument values the reference's plan classes pass to cufftPlanMany
(cufft_wrapper.py:147-250), rebuilt here the way those classes build them, and the
reference-side ctypes binding loaded into a namespace with a status check."""
import ctypes
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
CUFFT_Z2Z, CUFFT_C2C = 0x69, 0x29


def reference_plan_calls(P, X, fft_type=CUFFT_Z2Z):
    """[(name, args, keep)]: the cufftPlanMany arguments of Plan_Z2Z((P, X), batch=1) and
    Plan_Z2Z((X,), batch=1) (:160-167, as SetFFT_Plans creates them, GPU_Wigner2D_FFT.py:859-865),
    Plan_Z2Z_2D_Axis0((P, X)) (:190-201) and Plan_Z2Z_2D_Axis1((P, X)) (:225-236); keep holds
    the arrays whose addresses are passed."""
    calls = []
    n = np.asarray((P, X), np.int32)
    calls.append(("Plan_Z2Z((P,X))", (2, n.ctypes.data, None, 1, 0, None, 1, 0, fft_type, 1), [n]))
    n = np.asarray((X,), np.int32)
    calls.append(("Plan_Z2Z((X,))", (1, n.ctypes.data, None, 1, 0, None, 1, 0, fft_type, 1), [n]))
    # the axis classes build n and the embeddings with numpy's default integer type
    n = np.array([P])
    stride, idist = X, 1
    inembed, onembed = np.array([P, stride]), np.array([P, stride])
    calls.append(("Plan_Z2Z_2D_Axis0((P,X))", (1, n.ctypes.data, inembed.ctypes.data, stride, idist,
                                               onembed.ctypes.data, stride, idist, fft_type, X),
                  [n, inembed, onembed]))
    n = np.array([X])
    stride, idist = 1, X
    inembed, onembed = np.array((P, X)), np.array((P, X))
    calls.append(("Plan_Z2Z_2D_Axis1((P,X))", (1, n.ctypes.data, inembed.ctypes.data, stride, idist,
                                               onembed.ctypes.data, stride, idist, fft_type, P),
                  [n, inembed, onembed]))
    return calls


class StatusError(Exception):
    def __init__(self, status):
        super().__init__(f"status {status}")
        self.status = status


def load_binding(lib_path):
    """Executes pywignercuda_amd/cufft_ctypes_binding.py in a namespace that provides the
    cufftCheckStatus of the module it is pasted into (here: raise StatusError)."""
    import os

    def cufftCheckStatus(status):
        if status != 0:
            raise StatusError(status)

    src = (ROOT / "pywignercuda_amd" / "cufft_ctypes_binding.py").read_text()
    ns = {"cufftCheckStatus": cufftCheckStatus, "__name__": "cufft_wrapper"}
    old = os.environ.get("WIGNERHIP_LIB")
    os.environ["WIGNERHIP_LIB"] = str(lib_path)
    try:
        exec(compile(src, "cufft_ctypes_binding.py", "exec"), ns)
    finally:
        if old is None:
            del os.environ["WIGNERHIP_LIB"]
        else:
            os.environ["WIGNERHIP_LIB"] = old
    return ns
