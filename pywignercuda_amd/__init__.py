"""pywignercuda_amd: the Wigner phase-space split-operator propagator on MI355X (gfx950).

Drop-in for the propagator classes of cabrer7/PyWignerCUDA:

    from pywignercuda_amd.GPU_Wigner2D_FFT import *           # GPU_Wigner2D_FFT
    from pywignercuda_amd.GPU_Wigner2D_GPitaevskii import *   # GPU_Wigner2D_GPitaevskii
    from pywignercuda_amd.GPU_WignerDirac2D_4x4 import *      # GPU_WignerDirac2D_4x4
    from pywignercuda_amd.GPU_WignerDirac2D_MassPotential_4x4 import *   # ..._MassPotential_4x4
    from pywignercuda_amd import cufft_wrapper                 # plan / exec / destroy

The compute path is libwignerhip.so (hand-written HIP for gfx950, C ABI in
include/wignerhip.h), loaded from this directory; build it with
``python -m pywignercuda_amd._build``.  There is no CPU fallback.
"""
from ._lib import (  # noqa: F401
    WignerHipError,
    WignerHipLibraryError,
    WH_C128,
    WH_C64,
    WH_NOBS,
    OBS_COLUMNS,
    lib,
)

__version__ = "0.3.0"


def __getattr__(name):
    if name == "Wigner2D":
        from .wigner2d import Wigner2D

        return Wigner2D
    if name == "WignerEnsemble":
        from .wigner2d import WignerEnsemble

        return WignerEnsemble
    if name == "WignerDirac4x4":
        from .dirac import WignerDirac4x4

        return WignerDirac4x4
    raise AttributeError(name)
