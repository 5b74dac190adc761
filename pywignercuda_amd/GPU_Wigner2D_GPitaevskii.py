"""Drop-in for GPU_Wigner2D_GPitaevskii.py: the Strang propagator with the
Gross-Pitaevskii nonlinearity (GPU_Wigner2D_GPitaevskii.py:1767-2055).

Same usage as GPU_Wigner2D_FFT with ``self.GPitaevskiiCoeff`` instead of
``grossPitaevskiiCoefficient``; observables are recorded after each step and
``ProbabilitySquare_average`` holds dX sum rho^2 of the density after the kinetic
step (GP.py:1933-1937).  GPU_Wigner2D_GPitaevskii_Bloch is the imaginary-time
ground-state solver of the same file (GP.py:2064-2543).
"""
from __future__ import annotations

import numpy as np

from . import _io
from .GPU_Wigner2D_FFT import Propagator_Base
from .physics import heaviside_host, host_eval


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
    """Wigner Bloch (imaginary-time) propagator, GPU_Wigner2D_GPitaevskii.py:2064-2543.

    Same usage as the reference: set ``GPitaevskiiCoeff``, call ``SetTimeTrack`` and
    ``__init__(X_gridDIM, P_gridDIM, X_amplitude, P_amplitude, hBar, mass, potentialString,
    kinematicString)``, set ``W_init``, then ``Run()`` (the ground state ``W_0``) and
    optionally ``Run_ExitedState1()`` (``W_1``).  Each step is kept or refused on the device
    (energy must decrease and |purity| stay below 1; a refused step halves dt), so a run is
    one launch sequence without host round trips."""

    GPitaevskiiCoeff = 0.0

    def __init__(self, X_gridDIM, P_gridDIM, X_amplitude, P_amplitude, hBar, mass, potentialString,
                 kinematicString):
        self.D_Theta = 0.0
        self.D_Lambda = 0.0
        self.fp_Damping_String = "0."
        self.potentialString = potentialString
        self.dPotentialString = "0."
        self.kinematicString = kinematicString
        self.SetPhaseSpaceBox2D(X_gridDIM, P_gridDIM, X_amplitude, P_amplitude)
        self.hBar = hBar
        self.mass = mass
        from .bloch import GPBloch

        self.plan = GPBloch(X_gridDIM, P_gridDIM, X_amplitude, P_amplitude, hBar, mass, potentialString,
                            kinematicString, gp_coeff=self.GPitaevskiiCoeff, dt=self.dt,
                            time_rows=self.timeSteps + 2, truncate_constants=self.truncateConstants)
        self.Hamiltonian = self.fft_shift2D(self.plan.k_host[:, None] + self.plan.v0_host[None, :])

    def Potential(self, t, x):
        """Host-side Potential (GP.py:687-697; its Heaviside is 1/2 at 0)."""
        return host_eval(self.potentialString, heaviside=heaviside_host, t=t, x=x)

    def KineticEnergy(self, p):
        return host_eval(self.kinematicString, p=p)

    def NonLinearEnergy(self, ProbabilityX):
        """GP.py:1757-1758."""
        return self.dX * 0.5 * self.GPitaevskiiCoeff * np.sum(ProbabilityX ** 2)

    def _bloch(self, timeSteps, W_ground):
        import torch

        n_iter = (self.timeSteps if timeSteps is None else timeSteps) + 1
        W = self.plan.to_device(self.W_init)
        hist, _ = self.plan.run(W, n_iter, self.dt, W_ground=W_ground)
        h = self.plan.history(hist)
        torch.cuda.synchronize()
        kept = h["kept"] == 1.0
        self.timeRange = np.arange(n_iter) * self.dt
        self.X_average, self.X2_average = h["X"], h["X2"]
        self.P_average, self.P2_average = h["P"], h["P2"]
        self.Hamiltonian_average = h["H"]
        self.TotalEnergyHistory = h["E"][kept]
        if self.GPitaevskiiCoeff != 0.0:
            self.NonLinearEnergyHistory = h["NLE"][kept]
        self.purity = h["purity"][kept]
        self.dtHistory = h["dt"]
        self.stepKept = h["kept"]
        return W

    def Run(self):
        """GP.py:2105-2340: the ground state W_0 (returns 0)."""
        self.file = _io.RunFile(self.fileName or "Wigner2D_GPitaevskii_Bloch_run")
        self.WriteHDF5_variables()
        self.file.create_dataset("Hamiltonian", data=self.Hamiltonian.real)
        W = self._bloch(None, None)
        f = self.file
        f["timeRange"] = self.timeRange
        f["timeRangeIndexSaved"] = np.array([], dtype=np.int64)
        f["/Ehrenfest/X_Ehrenfest"] = self.X_average
        f["/Ehrenfest/X2_Ehrenfest"] = self.X2_average
        f["/Ehrenfest/P_Ehrenfest"] = self.P_average
        f["/Ehrenfest/P2_Ehrenfest"] = self.P2_average
        f["/Ehrenfest/Hamiltonian_average"] = self.Hamiltonian_average
        f["/Potential"] = self.Potential(0, self.X_range)
        f["/PotentialString"] = self.potentialString
        f["W_init"] = np.asarray(self.W_init).real
        self.W_0_GPU = W
        self.W_0 = W.cpu().numpy().real
        f["W_0"] = self.W_0
        self._saved = dict(f._data)
        f.close()
        return 0

    def Run_ExitedState1(self, timeSteps=None):
        """GP.py:2343-2535: the first excited state W_1, the ground state W_0 of Run removed
        after every step (returns 0).  The p-marginal refusal of the reference's source is not
        applied: its kernel writes out of bounds, and the recorded run
        (Wigner2D_GPitaevskii_X2_Bloch.ipynb cell 20) matches the loop without it."""
        if not hasattr(self, "W_0_GPU"):
            raise RuntimeError("Run_ExitedState1 needs the ground state of Run()")
        W = self._bloch(timeSteps, self.W_0_GPU)
        self.W_1 = W.cpu().numpy().real
        self.file = _io.RunFile(self.fileName or "Wigner2D_GPitaevskii_Bloch_run")
        for k, v in getattr(self, "_saved", {}).items():
            self.file[k] = v
        self.file["W_1"] = self.W_1
        self.file.close()
        return 0
