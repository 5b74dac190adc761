"""Drop-in for GPU_Wigner2D_GPitaevskii.py: the Strang propagator with the
Gross-Pitaevskii nonlinearity (GPU_Wigner2D_GPitaevskii.py:1767-2055).

Same usage as GPU_Wigner2D_FFT with ``self.GPitaevskiiCoeff`` instead of
``grossPitaevskiiCoefficient``; observables are recorded after each step and
``ProbabilitySquare_average`` holds dX sum rho^2 of the density after the kinetic
step (GP.py:1933-1937).  The imaginary-time (Bloch) solver of that file is not
built (GPU_Wigner2D_GPitaevskii_Bloch raises).
"""
from __future__ import annotations

import numpy as np

from . import _io
from .GPU_Wigner2D_FFT import Propagator_Base


class GPU_Wigner2D_GPitaevskii(Propagator_Base):
    GPitaevskiiCoeff = 0.0

    def __init__(self, X_gridDIM, P_gridDIM, X_amplitude, P_amplitude, hBar, mass, D_Theta, D_Lambda,
                 gammaDamping, potentialString, dPotentialString, kinematicString, normalization="Wigner"):
        self.normalization = normalization
        self.D_Theta = D_Theta
        self.D_Lambda = D_Lambda
        self.gammaDamping = gammaDamping
        self.potentialString = potentialString
        self.dPotentialString = dPotentialString
        self.kinematicString = kinematicString
        self.SetPhaseSpaceBox2D(X_gridDIM, P_gridDIM, X_amplitude, P_amplitude)
        self.hBar = hBar
        self.mass = mass
        self.plan = self._make_plan("strang", self.GPitaevskiiCoeff, kinematicString)
        self.Hamiltonian = self.P ** 2 / (2.0 * self.mass) + self.Potential(0, self.X)
        self.Hamiltonian = self.fft_shift2D(self.Hamiltonian)

    def NonLinearEnergy(self, ProbabilityX):
        """GP.py:1757-1758."""
        return self.dX * 0.5 * self.GPitaevskiiCoeff * np.sum(ProbabilityX ** 2)

    def Run(self):
        """GP.py:1809-2055; returns 0."""
        self.file = _io.RunFile(self.fileName or "Wigner2D_GPitaevskii_run")
        self.WriteHDF5_variables()
        self.file.create_dataset("Hamiltonian", data=self.Hamiltonian.real)
        rec, saved = self._run_loop(self.plan, post_step=True)
        extra = {}
        if self.GPitaevskiiCoeff != 0.0:
            self.ProbabilitySquare_average = rec[:, 13]
            extra["/Ehrenfest/ProbabilitySquare_average"] = self.ProbabilitySquare_average
        else:
            self.ProbabilitySquare_average = np.array([])
        self._store_results(rec, saved, extra)
        return 0


class GPU_Wigner2D_GPitaevskii_Bloch(Propagator_Base):
    """Imaginary-time ground-state solver (GP.py:2064-2543) -- not built."""

    def __init__(self, *args, **kwargs):
        raise NotImplementedError("the Bloch (imaginary-time) Gross-Pitaevskii solver is not implemented")
