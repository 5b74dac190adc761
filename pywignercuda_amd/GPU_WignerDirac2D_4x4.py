"""Drop-in GPU_WignerDirac2D_4x4 (GPU_WignerDirac2D_4x4.py:1880-3720) on libwignerhip.

Same constructor, the same attributes read before __init__ (Potential_0..3_String,
D_Theta, D_Lambda, grossPitaevskiiCoefficient, gammaDamping, dampingModel) and
before Run (W_init[4, 4, P, X], fileName), the same host helpers the notebooks
call (GaussianSpinor_*, Fourier_4X4_Theta_To_P, Wigner_4x4_Norm, ...) and the same
result attributes after Run (X_Average, ..., antiParticle_population, Dirac_energy,
negativity, transmission, TakabayashiAngle_init/_end, W_end).  The state lives on
the device as 16 planes; the step loop runs in libwignerhip (two fused sweeps per
step, observables reduced on the device, copied once per saved frame).

Extra attributes read if present: precision ('fp64' default or 'fp32'),
truncateConstants (default True, the reference's '%f' splicing), useGraph
(default True).  Physics strings (Potential_0..3, Mass_String) may depend on t and
need not be polynomials in x; strings calling C functions outside the expression
programs' ops raise NotImplementedError.  GPU_WignerDiracKvN2D_4x4 (GPU_WignerDiracKvN2D_4x4.py) is the KvN
variant of the same class.
"""
from __future__ import annotations

import numpy as np

from ._io import RunFile, read
from .dirac import WignerDirac4x4
from .physics import host_eval


class GPU_WignerDirac2D_4x4:
    """Propagator in the X-Theta representation of the 16 components of the Wigner-Dirac
    function (GPU_WignerDirac2D_4x4.py:1880-1885)."""

    def __init__(self, X_gridDIM, P_gridDIM, X_amplitude, P_amplitude, mass, c, dt, timeSteps, skipFrames=1,
                 frameSaveMode="Density", antiParticleNorm=True, antiParticleStepFiltering=False,
                 computeEnergy="False"):
        self.mass, self.c, self.dt = mass, c, dt
        self.timeSteps, self.skipFrames = timeSteps, skipFrames
        self.frameSaveMode = frameSaveMode
        self.antiParticleNorm = antiParticleNorm
        self.antiParticleStepFiltering = antiParticleStepFiltering
        self.computeEnergy = computeEnergy
        self.X_gridDIM, self.P_gridDIM = X_gridDIM, P_gridDIM
        self.X_amplitude, self.P_amplitude = X_amplitude, P_amplitude
        # grids (:1909-1927)
        self.dX = 2.0 * X_amplitude / float(X_gridDIM)
        self.dP = 2.0 * P_amplitude / float(P_gridDIM)
        self.dTheta = 2.0 * np.pi / (2.0 * P_amplitude)
        self.Theta_amplitude = self.dTheta * P_gridDIM / 2.0
        self.dLambda = 2.0 * np.pi / (2.0 * X_amplitude)
        self.Lambda_amplitude = self.dLambda * X_gridDIM / 2.0
        self.X_range = np.linspace(-X_amplitude, X_amplitude - self.dX, X_gridDIM)
        self.Lambda_range = np.linspace(-self.Lambda_amplitude, self.Lambda_amplitude - self.dLambda, X_gridDIM)
        self.Theta_range = np.linspace(-self.Theta_amplitude, self.Theta_amplitude - self.dTheta, P_gridDIM)
        self.P_range = np.linspace(-P_amplitude, P_amplitude - self.dP, P_gridDIM)
        self.X = np.fft.fftshift(self.X_range)[np.newaxis, :]
        self.Theta = np.fft.fftshift(self.Theta_range)[:, np.newaxis]
        self.Lambda = np.fft.fftshift(self.Lambda_range)[np.newaxis, :]
        self.P = np.fft.fftshift(self.P_range)[:, np.newaxis]
        for k in range(4):
            if not hasattr(self, f"Potential_{k}_String"):
                setattr(self, f"Potential_{k}_String", " 0.")
        self.D_Theta = getattr(self, "D_Theta", 0.0)
        self.D_Lambda = getattr(self, "D_Lambda", 0.0)
        self.grossPitaevskiiCoefficient = getattr(self, "grossPitaevskiiCoefficient", 0.0)
        self.gammaDamping = getattr(self, "gammaDamping", 0.0)
        self.dampingModel = getattr(self, "dampingModel", "CaldeiraLegget")
        self.precision = getattr(self, "precision", "fp64")
        self.truncateConstants = getattr(self, "truncateConstants", True)
        self.useGraph = getattr(self, "useGraph", True)
        self._plan = self._make_plan()

    def _make_plan(self, **variant):
        return WignerDirac4x4(
            self.X_gridDIM, self.P_gridDIM, self.X_amplitude, self.P_amplitude, self.mass, self.c, self.dt,
            potentials=tuple(getattr(self, f"Potential_{k}_String") for k in range(4)),
            D_theta=self.D_Theta, D_lambda=self.D_Lambda, gamma=self.gammaDamping,
            damping=self.dampingModel if self.gammaDamping > 0.0 else None,
            gp_coeff=self.grossPitaevskiiCoefficient, absorb=True, truncate_constants=self.truncateConstants,
            antiparticle=bool(self.antiParticleNorm), energy=(self.computeEnergy is True),
            precision=self.precision, lambda_bar=getattr(self, "lambdaBar", None),
            step_filter=bool(self.antiParticleStepFiltering),
            time_rows=(self.timeSteps + 3) if getattr(self, "timeSteps", None) else None, **variant)

    # ------------------------------------------------------------------ potentials (host, :2170-2228)
    def PotentialFunction_0(self, t, x):
        return host_eval(self.Potential_0_String, t=t, x=x)

    def PotentialFunction_1(self, t, x):
        return host_eval(self.Potential_1_String, t=t, x=x)

    def PotentialFunction_2(self, t, x):
        return host_eval(self.Potential_2_String, t=t, x=x)

    def PotentialFunction_3(self, t, x):
        return host_eval(self.Potential_3_String, t=t, x=x)

    # ------------------------------------------------------------------ spinors (:2231-2310)
    def _rho(self, x, px, s, X):
        return np.exp(1j * X * px) * np.exp(-0.5 * ((X - x) / s) ** 2) + 0j

    def GaussianSpinor_ParticleUp(self, x, px, s, X):
        rho = self._rho(x, px, s, X)
        p0 = np.sqrt(px * px + self.mass * self.mass * self.c * self.c)
        return np.array([rho * (p0 + self.mass * self.c), X * 0j, X * 0j, rho * px])

    def GaussianSpinor_ParticleDown(self, x, px, s, X):
        rho = self._rho(x, px, s, X)
        p0 = np.sqrt(px * px + self.mass * self.mass * self.c * self.c)
        return np.array([X * 0j, 1j * rho * (p0 + self.mass * self.c), 1j * rho * px, X * 0j])

    def GaussianSpinor_AntiParticleDown(self, x, px, s, X):
        rho = self._rho(x, px, s, X)
        p0 = np.sqrt(px * px + self.mass * self.mass * self.c * self.c)
        return -1j * np.array([X * 0j, rho * px, rho * (-p0 - self.mass * self.c), X * 0j])

    def GaussianSpinor_AntiParticleUp(self, x, px, s, X):
        rho = self._rho(x, px, s, X)
        p0 = np.sqrt(px * px + self.mass * self.mass * self.c * self.c)
        return -1j * np.array([rho * px, X * 0j, X * 0j, rho * (-p0 - self.mass * self.c)])

    def SpinorNorm(self, Psi):
        return sum(np.sum(np.abs(Psi[k]) ** 2) for k in range(4)) * self.dX * self.dP

    def _majorana(self, Psi, sign):
        """(Psi + sign * C Psi) / 2 with the charge conjugate C Psi = (-Psi4*, Psi3*, Psi2*, -Psi1*)
        of the standard representation (MajoranaSpinorPlus / Minus, D44.py:2337-2355)."""
        Psi = np.asarray(Psi) + 0j
        conj = np.array([-Psi[3].conj(), Psi[2].conj(), Psi[1].conj(), -Psi[0].conj()])
        return 0.5 * (Psi + sign * conj)

    def MajoranaSpinorPlus(self, Psi):
        return self._majorana(Psi, 1.0)

    def MajoranaSpinorMinus(self, Psi):
        return self._majorana(Psi, -1.0)

    # ------------------------------------------------------------------ host transforms (:2535-2556)
    def Fourier_X_To_Lambda(self, W):
        return np.fft.fft(W, axis=1)

    def Fourier_Lambda_To_X(self, W):
        return np.fft.ifft(W, axis=1)

    def Fourier_Theta_To_P(self, W):
        return np.fft.ifft(W, axis=0)

    def Fourier_P_To_Theta(self, W):
        return np.fft.fft(W, axis=0)

    def Fourier_4X4_Theta_To_P(self, W):
        for i in range(4):
            for j in range(4):
                W[i, j][:, :] = self.Fourier_Theta_To_P(W[i, j])

    def Fourier_4X4_P_To_Theta(self, W):
        for i in range(4):
            for j in range(4):
                W[i, j][:, :] = self.Fourier_P_To_Theta(W[i, j])

    def fftshift(self, M):
        return np.fft.fftshift(M)

    # ------------------------------------------------------------------ host diagnostics (:2311-2420, 2677-2738)
    def Wigner_4x4_Norm(self, W):
        return float(np.real(np.sum(W[0, 0]) + np.sum(W[1, 1]) + np.sum(W[2, 2]) + np.sum(W[3, 3]))) * self.dX * self.dP

    def Wigner_4X4__SpinTrace(self, W):
        return W[0, 0] + W[1, 1] + W[2, 2] + W[3, 3]

    def TakabayashiAngle_CPU(self, W):
        s = 2.0 * np.imag(W[0, 2]) + 2.0 * np.imag(W[1, 3])
        c = np.real(W[0, 0]) + np.real(W[1, 1]) - np.real(W[2, 2]) - np.real(W[3, 3])
        return np.arctan2(s, np.abs(c))

    def _takabayashi(self, W):
        """TakabayashiAngle_source (:1197-1227): atan2 without the abs of the CPU helper."""
        s = 2.0 * np.imag(W[0, 2]) + 2.0 * np.imag(W[1, 3])
        c = np.real(W[0, 0]) + np.real(W[1, 1]) - np.real(W[2, 2]) - np.real(W[3, 3])
        return np.arctan2(s, c)

    def WignerDirac_MatrixProduct(self, outC, inA, inB):
        for i in range(4):
            for j in range(4):
                outC[i, j][:, :] = sum(inA[i, k] * inB[k, j] for k in range(4))

    def DiracElectronProjector(self, p, sign):
        sq = 2 * np.sqrt((self.mass * self.c ** 2) ** 2 + (self.c * p) ** 2)
        aa = 0.5 + sign * self.mass * self.c * self.c / sq
        bb = self.c * sign * p / sq
        dd = 0.5 - sign * self.mass * self.c * self.c / sq
        z = np.zeros([self.P_gridDIM, self.X_gridDIM])
        return np.array([[aa, z, z, bb], [z, aa, bb, z], [z, bb, dd, z], [bb, z, z, dd]])

    def FilterElectrons(self, W, sign):
        L = self.DiracElectronProjector(self.P + 0.5 * self.Lambda, sign)
        R = self.DiracElectronProjector(self.P - 0.5 * self.Lambda, sign)
        for i in range(4):
            for j in range(4):
                W[i, j][:, :] = self.Fourier_X_To_Lambda(W[i, j])
        W_ = np.empty_like(W)
        self.WignerDirac_MatrixProduct(W_, L, W)
        self.WignerDirac_MatrixProduct(W, W_, R)
        for i in range(4):
            for j in range(4):
                W[i, j][:, :] = self.Fourier_Lambda_To_X(W[i, j])

    def WignerDiracPurity(self, W_4x4):
        s = sum(np.sum(W_4x4[i, j] * W_4x4[j, i]) for i in range(4) for j in range(4))
        return 2 * np.pi * s.real * self.dX * self.dP

    # ------------------------------------------------------------------ frames
    def save_Density(self, f, t_index, W):
        f.create_dataset(str(t_index), data=np.real(W[0] + W[5] + W[10] + W[15]).astype(np.float64))

    def save_WignerFunction(self, f, t, W):
        for i in range(4):
            for j in range(4):
                Wij = self.fftshift(W[4 * i + j])
                f.create_dataset(f"W_real_{i + 1}{j + 1}/{t}", data=np.real(Wij))
                f.create_dataset(f"W_imag_{i + 1}{j + 1}/{t}", data=np.imag(Wij))

    def Load_Density(self, fileName, n):
        """Frame n of a run saved with frameSaveMode 'Density' (D44.py:2915-2922)."""
        return read(fileName)[str(n)]

    def Load_WignerFunction(self, fileName, n):
        """Frame n of a run saved with frameSaveMode 'Wigner_4X4' as a [4][4][P][X] array
        (D44.py:2925-2951)."""
        f = read(fileName)
        return np.array([[f[f"W_real_{i}{j}/{n}"] + 1j * f[f"W_imag_{i}{j}/{n}"] for j in range(1, 5)]
                         for i in range(1, 5)])

    def _save_frame(self, f, t_index, W_dev):
        W = W_dev.cpu().numpy().astype(np.complex128)
        if self.frameSaveMode == "Wigner_4X4":
            self.save_WignerFunction(f, t_index, W)
        if self.frameSaveMode == "Density":
            self.save_Density(f, t_index, W)

    # ------------------------------------------------------------------ Run (:3248-3720)
    def Run(self):
        import torch

        f = RunFile(self.fileName)
        for k, v in (("x_gridDIM", self.X_gridDIM), ("x_amplitude", self.X_amplitude),
                     ("p_amplitude", self.P_amplitude), ("dx", self.dX), ("dtheta", self.dTheta), ("dp", self.dP),
                     ("p_gridDIM", self.P_gridDIM), ("mass", self.mass), ("dt", self.dt), ("c", self.c),
                     ("x_range", self.X_range)):
            f[k] = v
        for k in range(4):
            f[f"potential_{k}_String"] = getattr(self, f"Potential_{k}_String")
        plan = self._plan
        W = plan.to_device(self.W_init)
        self.TakabayashiAngle_init = self._takabayashi(np.asarray(self.W_init))
        saved = [0]
        self._save_frame(f, 0, W)
        skip = max(int(self.skipFrames), 1)
        rows = []
        t = 1
        while t <= self.timeSteps:
            end = min(self.timeSteps, ((t + skip - 1) // skip) * skip)  # next saved frame
            n = end - t + 1
            obs = plan.run(W, n, first_step=t, use_graph=self.useGraph)
            rows.append(obs[:n].cpu().numpy())
            t = end + 1
            if end % skip == 0:
                saved.append(end)
                self._save_frame(f, end, W)
        torch.cuda.synchronize()
        rec = WignerDirac4x4.observables(np.concatenate(rows, axis=0)) if rows else None
        self.timeRange = np.arange(self.timeSteps + 1) * self.dt
        W_end = W.cpu().numpy().astype(np.complex128).reshape(4, 4, self.P_gridDIM, self.X_gridDIM)
        self.W_end = W_end
        self.TakabayashiAngle_end = self._takabayashi(W_end)
        self.timeSavedIndexRange = np.array(saved)
        names = [("X_Average", "X"), ("P_Average", "P"), ("XP_Average", "XP"), ("XX_Average", "XX"),
                 ("PP_Average", "PP"), ("Alpha_1_Average", "Alpha1"), ("P1_Alpha_1_Average", "PAlpha1"),
                 ("X1_Alpha_1_Average", "XAlpha1"), ("D_1_Potential_0_Average", "D1V0"),
                 ("X1_D_1_Potential_0_Average", "XD1V0"), ("P1_D_1_Potential_0_Average", "PD1V0")]
        for attr, key in names:
            setattr(self, attr, np.real(rec[key]) if rec is not None else np.zeros(0))
            f[attr] = getattr(self, attr)
        self.negativity = rec["negativity"] if rec is not None else np.zeros(0)
        self.transmission = rec["transmission"] if rec is not None else np.zeros(0)
        self.antiParticle_population = rec["antiparticle"] if (rec is not None and self.antiParticleNorm) \
            else np.zeros(0)
        self.Dirac_energy = np.real(rec["energy"]) if (rec is not None and self.computeEnergy is True) \
            else np.zeros(0)
        f["timeRange"] = self.timeRange
        f["antiParticle_population"] = self.antiParticle_population
        f["Dirac_energy"] = self.Dirac_energy
        f["timeSavedIndexRange"] = self.timeSavedIndexRange
        f["Negativity"] = self.negativity
        f["transmission"] = self.transmission
        f["TakabayashiAngle_end"] = self.TakabayashiAngle_end
        f["TakabayashiAngle_init"] = self.TakabayashiAngle_init
        f.close()
        self.W_end_device = W
        return 0


class GPU_WignerDiracKvN2D_4x4(GPU_WignerDirac2D_4x4):
    """GPU_WignerDiracKvN2D_4x4 (GPU_WignerDiracKvN2D_4x4.py:1911-3780): the same API; kinetic
    factors at q = +/- lambda/2 with mass_half = 1e-5 and no diagonal mass term (:183-189,
    :283-289), the potential step e^{-dt D_Theta theta^2 + i dt V0'(x) theta} on every plane
    (KvN_XTheta_Kernel, :356-420), the antiparticle filter with projectors at p after every
    step (:775, :3599-3640); its loop has no damping and no Gross-Pitaevskii term."""

    def __init__(self, *args, **kwargs):
        kwargs.setdefault("antiParticleStepFiltering", True)
        super().__init__(*args, **kwargs)

    def _make_plan(self, **variant):
        return WignerDirac4x4(
            self.X_gridDIM, self.P_gridDIM, self.X_amplitude, self.P_amplitude, self.mass, self.c, self.dt,
            potentials=(self.Potential_0_String, "0.", "0.", "0."), D_theta=self.D_Theta, D_lambda=self.D_Lambda,
            absorb=True, truncate_constants=self.truncateConstants, antiparticle=bool(self.antiParticleNorm),
            energy=(self.computeEnergy is True), precision=self.precision, kvn=True, **variant)
