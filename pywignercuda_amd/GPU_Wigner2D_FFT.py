"""Drop-in for GPU_Wigner2D_FFT.py: Propagator_Base, GPU_Wigner2D_FFT, GPU_Wigner2D_ODM2 and
the classical KvN propagator GPU_FokkerPlank2D_FFT.

Usage is the reference's (instances/Wigner2D/*.ipynb): subclass, set
``dampingFunction``, ``grossPitaevskiiCoefficient``, ``fp_Damping_String``, call
``SetTimeTrack(...)`` and ``GPU_Wigner2D_FFT.__init__(self, X_gridDIM, P_gridDIM,
X_amplitude, P_amplitude, hBar, mass, D_Theta, D_Lambda, gammaDamping,
potentialString, dPotentialString, kinematicString)``, set ``W_init`` (a P x X
array in FFT order) and call ``Run()``.  ``Run`` returns 0 and sets the same
result attributes (timeRange, X_average, ..., negativeArea, W_end) and writes the
same datasets (HDF5 when h5py is importable, else ``fileName + '.npz'``).

The time loop runs on the GPU through libwignerhip (fused sweeps, observables
reduced on the device, one host copy per saved frame).  Options beyond the
reference, set before ``__init__``:

    precision          'fp64' (default, complex128) or 'fp32' (complex64 bulk)
    truncateConstants  True (default): kernel constants spliced as '%f', exactly
                       as the reference's CUDA sources receive them
                       (GPU_Wigner2D_FFT.py:833-842); False: full precision
    absorbingMask      True (default): the potential kernel's edge mask
                       (GPU_Wigner2D_FFT.py:121-123); False: no mask
    useGraph           True (default): replay the step loop from a hipGraph
"""
from __future__ import annotations

import numpy as np
import scipy.fftpack as fftpack

from . import _io
from .physics import host_eval
from .wigner2d import Wigner2D


class Propagator_Base:
    """Host-side part of Propagator_Base (GPU_Wigner2D_FFT.py:529-1705) kept for drop-in use."""

    precision = "fp64"
    truncateConstants = True
    absorbingMask = True
    useGraph = True
    grossPitaevskiiCoefficient = 0.0
    dampingFunction = "CaldeiraLeggett"
    fp_Damping_String = "p"

    # ---------------------------------------------------------------- set-up
    def SetTimeTrack(self, dt, timeSteps=128, skipFrames=1, fileName="", compression=None):
        """GPU_Wigner2D_FFT.py:530-546 (no kernels are compiled here)."""
        self.runTime = float(dt) * timeSteps
        self.timeSteps = timeSteps
        self.dt = dt
        self.skipFrames = skipFrames
        self.fileName = fileName
        self.compression = compression
        self.__x_p_representation__ = "x_p"

    def SetPhaseSpaceBox2D(self, X_gridDIM, P_gridDIM, X_amplitude, P_amplitude):
        """GPU_Wigner2D_FFT.py:549-567."""
        self.X_gridDIM = X_gridDIM
        self.P_gridDIM = P_gridDIM
        self.X_amplitude = X_amplitude
        self.P_amplitude = P_amplitude
        self.dX = 2.0 * X_amplitude / float(X_gridDIM)
        self.dP = 2.0 * P_amplitude / float(P_gridDIM)
        self.dTheta = 2.0 * np.pi / (2.0 * P_amplitude)
        self.Theta_amplitude = self.dTheta * P_gridDIM / 2.0
        self.dLambda = 2.0 * np.pi / (2.0 * X_amplitude)
        self.Lambda_amplitude = self.dLambda * X_gridDIM / 2.0
        self.SetPhaseSpaceGrid2D()

    def SetPhaseSpaceGrid2D(self):
        """GPU_Wigner2D_FFT.py:569-596; the device grids are computed in the kernels."""
        self.X_range = np.linspace(-self.X_amplitude, self.X_amplitude - self.dX, self.X_gridDIM)
        self.Lambda_range = np.linspace(-self.Lambda_amplitude, self.Lambda_amplitude - self.dLambda, self.X_gridDIM)
        self.Theta_range = np.linspace(-self.Theta_amplitude, self.Theta_amplitude - self.dTheta, self.P_gridDIM)
        self.P_range = np.linspace(-self.P_amplitude, self.P_amplitude - self.dP, self.P_gridDIM)
        self.X = fftpack.fftshift(self.X_range)[np.newaxis, :]
        self.Theta = fftpack.fftshift(self.Theta_range)[:, np.newaxis]
        self.Lambda = fftpack.fftshift(self.Lambda_range)[np.newaxis, :]
        self.P = fftpack.fftshift(self.P_range)[:, np.newaxis]

    # ---------------------------------------------------------------- physics on the host
    def Potential(self, t, x):
        return host_eval(self.potentialString, t=t, x=x)

    def dPotential(self, t, x):
        return host_eval(self.dPotentialString, t=t, x=x)

    def fp_Damping(self, p):
        return host_eval(self.fp_Damping_String, p=p)

    # ---------------------------------------------------------------- classical reference trajectories
    def symplecticExpK(self, mass, dt, c, Z):
        return Z + np.array([dt * c * Z[1] / mass, 0])

    def symplecticExpV(self, mass, dt, d, dVdx, Z):
        return Z + np.array([0, -dt * d * dVdx(0.0, Z[0])])

    def symplecticExpKExpV2(self, mass, dt, dVdx, Z0):
        """Second-order classical symplectic step (GPU_Wigner2D_FFT.py:609-622)."""
        Z = self.symplecticExpK(mass, dt, 0.5, self.symplecticExpV(mass, dt, 1.0, dVdx, Z0))
        return self.symplecticExpK(mass, dt, 0.5, self.symplecticExpV(mass, dt, 0.0, dVdx, Z))

    def SymplecticPropagator(self, dt, n_iterations, Z0, gammaDamping):
        """Classical trajectory with linear friction (GPU_Wigner2D_FFT.py:624-640)."""
        Z = np.asarray(Z0, dtype=float)
        traj = []
        for i in range(n_iterations):
            traj.append(np.append(Z, dt * (i + 1)))
            Z = self.symplecticExpKExpV2(self.mass, dt, self.dPotential, Z)
            Z[1] = np.exp(-2.0 * gammaDamping * dt) * Z[1]
        return np.array(traj)

    def SymplecticPropagator_SmoothedP(self, dt, n_iterations, Z0, gammaDamping):
        Z = np.asarray(Z0, dtype=float)
        traj = []
        for i in range(n_iterations):
            traj.append(np.append(Z, dt * (i + 1)))
            Z = self.symplecticExpKExpV2(self.mass, dt, self.dPotential, Z)
            Z[1] = Z[1] - 2.0 * gammaDamping * dt * np.sign(Z[1]) * self.fp_Damping(Z[1])
        return np.array(traj)

    # ---------------------------------------------------------------- initial states and helpers
    def fft_shift2D(self, X):
        return fftpack.fftshift(fftpack.fftshift(X, axes=1), axes=0)

    def Wigner_HarmonicOscillator(self, n, omega, x, p):
        """GPU_Wigner2D_FFT.py:715-728."""
        from scipy.special import laguerre

        self.representation = "x_p"
        r2 = self.mass * omega ** 2 * (self.X - x) ** 2 + (self.P - p) ** 2 / self.mass
        W = (-1) ** n * laguerre(n)(2 * r2) * np.exp(-r2)
        return W / (np.sum(W) * self.dX * self.dP)

    def Heaviside(self, x):
        return 0.5 * (np.sign(x) + 1.0)

    def Psi_Half_HarmonicOscillator(self, n, omega, x0, p0, X):
        from scipy.special import legendre

        k = np.sqrt(self.mass * omega / self.hBar)
        return (np.exp(1j * p0) * np.exp(-0.5 * k ** 2 * (X - x0) ** 2) * legendre(n)(k * (X - x0))
                * self.Heaviside(-(X - x0)))

    def Wigner_Half_HarmonicOscillator(self, n, omega, x0, p0):
        """GPU_Wigner2D_FFT.py:743-771: W from psi(x - hbar theta/2) psi*(x + hbar theta/2)."""
        Xm = self.X_range[np.newaxis, :] - 0.5 * self.hBar * self.Theta_range[:, np.newaxis]
        Xp = self.X_range[np.newaxis, :] + 0.5 * self.hBar * self.Theta_range[:, np.newaxis]
        W = self.Psi_Half_HarmonicOscillator(n, omega, x0, p0, Xm) * \
            self.Psi_Half_HarmonicOscillator(n, omega, x0, p0, Xp).conj()
        W = self.Fourier_Theta_To_P_CPU(fftpack.fftshift(W))
        return W / (np.sum(W) * self.dX * self.dP)

    def Fourier_XTheta_to_XP(self, M):
        return fftpack.ifft(M, axis=0)

    def Fourier_Theta_To_P_CPU(self, W):
        return fftpack.ifft(W, axis=0)

    def Fourier_P_To_Theta_CPU(self, W):
        return fftpack.fft(W, axis=0)

    def Fourier_X_To_Lambda_CPU(self, W):
        return fftpack.fft(W, axis=1)

    def Fourier_Lambda_To_X_CPU(self, W):
        return fftpack.ifft(W, axis=1)

    # ---------------------------------------------------------------- device transforms (torch tensors)
    def _axis_plans(self, dtype):
        from . import cufft_wrapper as cw

        key = "_plans_" + str(dtype)
        if not hasattr(self, key):
            t = cw.CUFFT_Z2Z if "128" in str(dtype) else cw.CUFFT_C2C
            shape = (self.P_gridDIM, self.X_gridDIM)
            setattr(self, key, (cw.Plan_Z2Z_2D_Axis0(shape, t), cw.Plan_Z2Z_2D_Axis1(shape, t)))
        return getattr(self, key)

    def Fourier_X_To_Lambda_GPU(self, W_GPU):
        from . import cufft_wrapper as cw

        cw.fft_Z2Z(W_GPU, W_GPU, self._axis_plans(W_GPU.dtype)[1])

    def Fourier_Lambda_To_X_GPU(self, W_GPU):
        from . import cufft_wrapper as cw

        cw.ifft_Z2Z_normalized(W_GPU, W_GPU, self._axis_plans(W_GPU.dtype)[1])

    def Fourier_P_To_Theta_GPU(self, W_GPU):
        from . import cufft_wrapper as cw

        cw.fft_Z2Z(W_GPU, W_GPU, self._axis_plans(W_GPU.dtype)[0])

    def Fourier_Theta_To_P_GPU(self, W_GPU):
        from . import cufft_wrapper as cw

        cw.ifft_Z2Z_normalized(W_GPU, W_GPU, self._axis_plans(W_GPU.dtype)[0])

    def Fourier_XTheta_To_LambdaP_GPU(self, W_GPU):
        self.Fourier_X_To_Lambda_GPU(W_GPU)
        self.Fourier_Theta_To_P_GPU(W_GPU)

    def Fourier_LambdaP_To_XTheta_GPU(self, W_GPU):
        self.Fourier_Lambda_To_X_GPU(W_GPU)
        self.Fourier_P_To_Theta_GPU(W_GPU)

    # ---------------------------------------------------------------- output files
    def WriteHDF5_variables(self):
        """GPU_Wigner2D_FFT.py:844-857."""
        f = self.file
        f["dx"] = self.dX
        f["dtheta"] = self.dTheta
        f["dp"] = self.dP
        f["dlambda"] = self.dLambda
        f["x_gridDIM"] = self.X_gridDIM
        f["p_gridDIM"] = self.P_gridDIM
        f["x_min"] = -self.X_amplitude
        f["x_max"] = self.X_amplitude - self.dX
        f["p_min"] = -self.P_amplitude
        f["p_max"] = self.P_amplitude - self.dP
        f["lambda_min"] = self.Lambda_range.min()
        f["lambda_max"] = self.Lambda_range.max()
        f["dt"] = self.dt
        f["timeSteps"] = self.timeSteps
        f["skipFrames"] = self.skipFrames

    def save_Frame(self, t, W):
        """GPU_Wigner2D_FFT.py:660-663 (W a device tensor or an array)."""
        arr = W.cpu().numpy() if hasattr(W, "cpu") else np.asarray(W)
        self.file.create_dataset(str(t), data=arr)

    def WignerFunctionFromFile(self, n, fileName=None):
        return _io.read(fileName or self.fileName)[str(n)]

    def LoadEhrenfestFromFile(self):
        """GPU_Wigner2D_FFT.py:919-942."""
        f = _io.read(self.fileName)
        self.X_average = f["Ehrenfest/X_Ehrenfest"]
        self.X2_average = f["Ehrenfest/X2_Ehrenfest"]
        self.P_average = f["Ehrenfest/P_Ehrenfest"]
        self.P2_average = f["Ehrenfest/P2_Ehrenfest"]
        self.XP_average = f["Ehrenfest/XP_Ehrenfest"]
        self.dPotentialdX_average = f["Ehrenfest/dPotentialdX_Ehrenfest"]
        self.PdPotentialdX_average = f["Ehrenfest/PdPotentialdX_average"]
        self.XdPotentialdX_average = f["Ehrenfest/XdPotentialdX_average"]
        self.Hamiltonian_average = f["Ehrenfest/Hamiltonian_average"]
        self.W_init = f["W_init"] + 0j
        self.W_end = f["W_end"] + 0j

    def _ehrenfest_series(self, name, fileName):
        """The Ehrenfest_*_FromFile readers (GPU_Wigner2D_FFT.py:944-1048): (t, record) with
        t = k dt, k = 0..timeSteps.  They read '/Ehrenfest/<name>_average'; for X, X2, P, P2
        and XP a file written by Run holds '<name>_Ehrenfest' instead (:1939-1943), which is
        read when the first is absent."""
        f = _io.read(fileName or self.fileName)
        key = f"Ehrenfest/{name}_average"
        if key not in f:
            key = f"Ehrenfest/{name}_Ehrenfest"
        return np.arange(int(f["timeSteps"]) + 1) * float(f["dt"]), f[key]

    def Ehrenfest_X_FromFile(self, fileName=None):
        return self._ehrenfest_series("X", fileName)

    def Ehrenfest_P_FromFile(self, fileName=None):
        return self._ehrenfest_series("P", fileName)

    def Ehrenfest_X2_FromFile(self, fileName=None):
        return self._ehrenfest_series("X2", fileName)

    def Ehrenfest_P2_FromFile(self, fileName=None):
        return self._ehrenfest_series("P2", fileName)

    def Ehrenfest_XP_FromFile(self, fileName=None):
        return self._ehrenfest_series("XP", fileName)

    def Ehrenfest_Hamiltonian_FromFile(self, fileName=None):
        return self._ehrenfest_series("Hamiltonian", fileName)

    def WignerMarginal_Probability_x(self, W):
        """Position marginal: sum over the p axis (axis 0) times dP (GPU_Wigner2D_FFT.py:1050-1051)."""
        return np.sum(W, axis=0) * self.dP

    def WignerMarginal_Probability_p(self, W):
        """Momentum marginal: sum over the x axis (axis 1) times dX (GPU_Wigner2D_FFT.py:1053-1054)."""
        return np.sum(W, axis=1) * self.dX

    def OperatorX_XBasis(self):
        """diag(X_range) as complex128 (GPU_Wigner2D_FFT.py:1187-1193)."""
        return np.diag(self.X_range).astype(np.complex128)

    def OperatorP_XBasis(self):
        """The momentum operator on the X grid (GPU_Wigner2D_FFT.py:1159-1185): the matrix of
        -i d/dx for the band-limited interpolant through the grid points, at half-integer
        indices j = n - (N-1)/2.  For k = j_n - j_m (an integer, nonzero off the diagonal):
        P[n, m] = i pi / (dX N^2) * (sin(pi k) cos(pi k/N) - N cos(pi k) sin(pi k/N)) / sin^2(pi k/N),
        zero on the diagonal.  It equals sum_q p_q |q><q| over the plane waves of the
        half-integer momenta p_q = 2 pi q / (N dX), q = -(N-1)/2 .. (N-1)/2."""
        N = self.X_gridDIM
        j = (2.0 * np.arange(N) - N + 1.0) / 2.0
        k = j[:, None] - j[None, :]
        off = k != 0
        s = np.sin(np.pi * np.where(off, k, 1.0) / N)
        val = (np.sin(np.pi * k) * np.cos(np.pi * k / N) - N * np.cos(np.pi * k) * s) / s ** 2
        return np.pi * 1j / (self.dX * N ** 2) * np.where(off, val, 0.0)

    # ---------------------------------------------------------------- device plan
    def _make_plan(self, scheme, gp_coeff, kinetic, odm_weights=None, kvn=False, renormalize=None):
        if kvn:
            # the Fokker-Planck Run applies only the Caldeira-Leggett dissipator (FFT.py:2404-2406)
            damping = "CaldeiraLeggett" if self.gammaDamping != 0.0 and \
                self.dampingFunction == "CaldeiraLeggett" else None
        else:
            if self.gammaDamping != 0.0 and self.dampingFunction not in ("CaldeiraLeggett", "ODM"):
                raise NotImplementedError(f"dampingFunction {self.dampingFunction!r} is not implemented")
            damping = self.dampingFunction if self.gammaDamping != 0.0 else None
        return Wigner2D(
            self.X_gridDIM, self.P_gridDIM, self.X_amplitude, self.P_amplitude, self.hBar, self.mass, self.dt,
            potential=self.potentialString, dpotential=self.dPotentialString, kinetic=kinetic, scheme=scheme,
            D_theta=self.D_Theta, D_lambda=self.D_Lambda, gamma=self.gammaDamping, damping=damping,
            L_material=getattr(self, "L_material", None), kvn=kvn, renormalize=renormalize,
            fp_damping=self.fp_Damping_String, gp_coeff=gp_coeff, mask=self.absorbingMask,
            truncate_constants=self.truncateConstants, precision=self.precision,
            time_rows=self.timeSteps + 3,
            odm_weights=odm_weights if self.dampingFunction == "ODM" else None)

    def _run_loop(self, plan, post_step):
        """Runs timeSteps+1 iterations in chunks between saved frames; returns the record."""
        import torch

        W = plan.to_device(self.W_init)
        total = self.timeSteps + 1
        skip = max(int(self.skipFrames), 1)
        rows, saved = [], []
        it = 0
        while it < total:
            if post_step:  # frames after iteration f (GP.py:1997-1999): chunk ends after f
                end = min(total, (it // skip) * skip + 1 if it % skip == 0 else (it // skip + 1) * skip + 1)
            else:  # frames before iteration f (FFT.py:1880-1882): chunk starts at f
                if it % skip == 0:
                    saved.append(it)
                    self.save_Frame(it, W if it > 0 else self._normalised_init())
                end = min(total, (it // skip + 1) * skip)
            n = end - it
            obs = plan.run(W, n, first_step=it, normalize_initial=(it == 0), use_graph=self.useGraph)
            rows.append(obs[:n].cpu().numpy())
            it = end
            if post_step and (it - 1) % skip == 0:
                saved.append(it - 1)
                self.save_Frame(it - 1, W)
        torch.cuda.synchronize()
        self._W_device = W
        return np.concatenate(rows, axis=0), saved

    def _normalised_init(self):
        W = np.asarray(self.W_init, dtype=np.complex128)
        return W / (W.sum() * self.dX * self.dP)

    def _store_results(self, rec, saved, extra=None):
        n = rec.shape[0]
        self.timeRange = np.arange(n) * self.dt
        self.X_average, self.X2_average = rec[:, 0], rec[:, 1]
        self.P_average, self.P2_average = rec[:, 2], rec[:, 3]
        self.XP_average = rec[:, 4]
        self.dPotentialdX_average = rec[:, 5]
        self.PdPotentialdX_average = rec[:, 6]
        self.XdPotentialdX_average = rec[:, 7]
        self.Hamiltonian_average = rec[:, 8]
        self.negativeArea = rec[:, 9]
        f = self.file
        f["timeRange"] = self.timeRange
        f["timeRangeIndexSaved"] = np.array(saved)
        f["/Ehrenfest/X_Ehrenfest"] = self.X_average
        f["/Ehrenfest/X2_Ehrenfest"] = self.X2_average
        f["/Ehrenfest/P_Ehrenfest"] = self.P_average
        f["/Ehrenfest/P2_Ehrenfest"] = self.P2_average
        f["/Ehrenfest/XP_Ehrenfest"] = self.XP_average
        f["/Ehrenfest/dPotentialdX_Ehrenfest"] = self.dPotentialdX_average
        f["/Ehrenfest/PdPotentialdX_average"] = self.PdPotentialdX_average
        f["/Ehrenfest/XdPotentialdX_average"] = self.XdPotentialdX_average
        f["/Ehrenfest/Hamiltonian_average"] = self.Hamiltonian_average
        for k, v in (extra or {}).items():
            f[k] = v
        f["W_init"] = np.asarray(self.W_init).real
        self.W_end = self._W_device.cpu().numpy().astype(np.complex128).real
        f["W_end"] = self.W_end
        f["negativeArea"] = self.negativeArea
        f.close()


class GPU_Wigner2D_FFT(Propagator_Base):
    """Wigner propagator in 2D phase space with diffusion and amplitude damping
    (GPU_Wigner2D_FFT.py:1713-1964): Lie step K(dt); V(dt) [+ GP]; CL friction; renormalisation."""

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
        self.plan = self._make_plan("lie", self.grossPitaevskiiCoefficient, None)
        self.Hamiltonian = self.P ** 2 / (2.0 * self.mass) + self.Potential(0, self.X)
        self.Hamiltonian = self.fft_shift2D(self.Hamiltonian)

    def Run(self):
        """GPU_Wigner2D_FFT.py:1754-1964; returns 0."""
        self.file = _io.RunFile(self.fileName or "Wigner2D_run")
        self.WriteHDF5_variables()
        self.file.create_dataset("Hamiltonian", data=self.Hamiltonian.real)
        rec, saved = self._run_loop(self.plan, post_step=False)
        self._store_results(rec, saved)
        return 0


class GPU_Wigner2D_ODM2(Propagator_Base):
    """Wigner propagator with two ODM Lindblad channels (GPU_Wigner2D_FFT.py:1973-2230): the Lie
    step of GPU_Wigner2D_FFT, then W += L(W) with (weightLindbland1, sign 1) and with
    (weightLindbland2, sign -1) on the result, then the renormalisation.  Requires
    ``dampingFunction = 'ODM'`` (or 'CaldeiraLeggett', which takes the order-3 composite),
    ``L_material`` and ``epsilon``/``fp_Damping_String`` as in the reference."""

    def __init__(self, X_gridDIM, P_gridDIM, X_amplitude, P_amplitude, hBar, mass, D_Theta, D_Lambda,
                 gammaDamping, weightLindbland1, weightLindbland2, potentialString, dPotentialString,
                 kinematicString, normalization="Wigner"):
        self.normalization = normalization
        self.D_Theta = D_Theta
        self.D_Lambda = D_Lambda
        self.gammaDamping = gammaDamping
        self.weightLindbland1 = weightLindbland1
        self.weightLindbland2 = weightLindbland2
        self.potentialString = potentialString
        self.dPotentialString = dPotentialString
        self.kinematicString = kinematicString
        self.SetPhaseSpaceBox2D(X_gridDIM, P_gridDIM, X_amplitude, P_amplitude)
        self.hBar = hBar
        self.mass = mass
        self.plan = self._make_plan("lie", self.grossPitaevskiiCoefficient, None,
                                    odm_weights=(weightLindbland1, weightLindbland2))
        self.Hamiltonian = self.P ** 2 / (2.0 * self.mass) + self.Potential(0, self.X)
        self.Hamiltonian = self.fft_shift2D(self.Hamiltonian)

    def Run(self):
        """GPU_Wigner2D_FFT.py:2016-2230; returns 0."""
        self.file = _io.RunFile(self.fileName or "Wigner2D_ODM2_run")
        self.WriteHDF5_variables()
        self.file.create_dataset("Hamiltonian", data=self.Hamiltonian.real)
        rec, saved = self._run_loop(self.plan, post_step=False)
        self._store_results(rec, saved)
        return 0


class GPU_FokkerPlank2D_FFT(Propagator_Base):
    """Classical Koopman-von Neumann Fokker-Planck propagator in 2D phase space
    (GPU_Wigner2D_FFT.py:2242-2459): the Lie step of GPU_Wigner2D_FFT with the potential factor
    exp(i dt dV(x) theta) exp(-D_Theta theta^2 dt) and the edge mask of exp_PotentialKvN
    (:27-75) in place of the Bopp difference of V, the Caldeira-Leggett dissipator when
    gammaDamping != 0 and dampingFunction == 'CaldeiraLeggett', and the renormalisation when
    gammaDamping != 0 and normalization == 'Wigner' (:2403-2410).  A dPotentialString that
    depends on t is evaluated per cell and step on the device.  The fields the reference reads off ``self`` (dampingFunction,
    fp_Damping_String, SetTimeTrack) are set as for GPU_Wigner2D_FFT; no kinematicString is
    needed."""

    def __init__(self, X_gridDIM, P_gridDIM, X_amplitude, P_amplitude, hBar, mass, D_Theta, D_Lambda,
                 gammaDamping, potentialString, dPotentialString, normalization="Wigner"):
        self.normalization = normalization
        self.D_Theta = D_Theta
        self.D_Lambda = D_Lambda
        self.gammaDamping = gammaDamping
        self.potentialString = potentialString
        self.dPotentialString = dPotentialString
        self.SetPhaseSpaceBox2D(X_gridDIM, P_gridDIM, X_amplitude, P_amplitude)
        self.hBar = hBar
        self.mass = mass
        self.plan = self._make_plan("lie", 0.0, None, kvn=True,
                                    renormalize=gammaDamping != 0.0 and normalization == "Wigner")
        self.Hamiltonian = self.P ** 2 / (2.0 * self.mass) + self.Potential(0, self.X)
        self.Hamiltonian = self.fft_shift2D(self.Hamiltonian)

    def Run(self):
        """GPU_Wigner2D_FFT.py:2282-2459; returns 0."""
        self.file = _io.RunFile(self.fileName or "FokkerPlanck2D_run")
        self.WriteHDF5_variables()
        self.file.create_dataset("Hamiltonian", data=self.Hamiltonian.real)
        rec, saved = self._run_loop(self.plan, post_step=False)
        self._store_results(rec, saved)
        return 0
