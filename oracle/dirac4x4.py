"""Oracle: NumPy complex128 restatement of the Wigner-Dirac 4x4 split step.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Never shipped, never the
thing measured.

Follows GPU_WignerDirac2D_4x4.py in behaviour (not in code):

* grids, '%f' constants ........... :1886-1942 (mass, mass_half = mass/2, c, dt, dx, dlambda, dp,
                                     dtheta, D_Theta spliced with '%f')
* kernel coordinate maps .......... :174-181 (i = (row + P/2) % P, p = dp (i - P/2), lambda = dlambda (j - X/2))
* axis transforms ................. :2557-2569 (forward e^{-i}; inverse unnormalised, then *= 1/N)
* kinetic factors ................. :159-351  W <- U(p + lambda/2) W U(p - lambda/2)^dagger, with
                                     U = cos(dt c s) - i sin(dt c s) (mass_half c beta + q alpha_1) / s,
                                     s = sqrt((mass_half c)^2 + q^2)
* potential factors ............... :357-670  W <- e^{V-} P-(x - theta/2) W P+(x + theta/2) e^{V+}
* Caldeira-Leggett order 2 ........ :3011-3058 (dampingModel 'CaldeiraLegget'; sigma(p) = p, host grids)
* ODM damping ..................... :952-1195, applied at :3608-3615 (dampingModel 'ODM', lambdaBar :1947)
* antiparticle step filter ........ :2500-2540, 3636-3678 (projectors at p -/+ lambda/2, sign +1)
* KvN variant ..................... GPU_WignerDiracKvN2D_4x4.py:183-189, 283-289 (kinetic), 356-420
                                    (potential), 775 (filter at lambda = 0), 1961 (mass_half 1e-5)
* edge absorber ................... :111-155 (w = 24 cells, array index), applied at :3676-3683
* trace normalisation ............. :2740-2775
* Gross-Pitaevskii terms .......... :1791-1840, 2695-2715 (rho = Re sum_p trace, no dP)
* observables (pre-step) .......... :3429-3492 and the reduction kernels :43-105, :1229-1790
* antiparticle population ......... :2369-2466 (projector filter, sign -1, in the (p, lambda) picture)
* driver order .................... :3248-3690 (timeSteps iterations, t = dt * t_index, t_index >= 1)

Layout: W[i, j, p_index, x_index] (i, j = spinor indices 0..3), FFT ordered, as
the reference's W_init[4, 4, P, X].

Pins (tests/test_oracle_dirac.py): unitarity of every per-cell factor, trace-norm
conservation, Hermiticity W_ij = conj W_ji, the free Dirac kinetic step against
the exact (p, lambda) propagator of a single spinor, and the pure-state
cross-check against an independent 1-D Dirac spinor split step.
"""
from __future__ import annotations

import numpy as np

from .wigner2d import fft_order_axis, host_axis, percent_f

# plane index k = 4 i + j
DIAG = (0, 5, 10, 15)
ALPHA1 = (3, 6, 9, 12)  # W14, W23, W32, W41


class Dirac4x4Oracle:
    """Restatement of GPU_WignerDirac2D_4x4 (scalar and vector potentials, GP, CL order 2).

    potentials: four callables V_k(t, x) (Potential_0..3_String); dpotential0: dV_0/dx.
    """

    def __init__(self, X, P, X_amp, P_amp, mass, c, dt, potentials, dpotential0=None, D_theta=0.0,
                 gamma=0.0, damping=None, gp_coeff=0.0, absorb=True, truncate_constants=True,
                 antiparticle=True, compute_energy=False, mass_potential=None, kinetic_mass_half=None,
                 lambda_bar=None, kvn=False, step_filter=False):
        """kvn: GPU_WignerDiracKvN2D_4x4 (mass_half 1e-5 unless given, kinetic factors at
        q = +/- lambda/2 without the diagonal mass term, potential e^{-dt D theta^2 + i dt V0' theta},
        step filter on with projectors at p).  step_filter: antiParticleStepFiltering."""
        """mass_potential(t, x), kinetic_mass_half: the MassPotential variant
        (GPU_WignerDirac2D_MassPotential_4x4.py:387-391, 422-427, 1668-1669): Mass(t, x) in the
        potential factors, a fixed mass_half (1e-4 there) in the kinetic factors, the bare mass
        in the projectors."""
        self.X, self.P = int(X), int(P)
        self.X_amp, self.P_amp = float(X_amp), float(P_amp)
        self.mass, self.c, self.dt = float(mass), float(c), float(dt)
        self.kvn = bool(kvn)
        self.step_filter = bool(step_filter) or self.kvn
        if self.kvn and kinetic_mass_half is None:
            kinetic_mass_half = 0.00001  # GPU_WignerDiracKvN2D_4x4.py:1961
        self.dX = 2.0 * self.X_amp / self.X
        self.dP = 2.0 * self.P_amp / self.P
        self.dTheta = 2.0 * np.pi / (2.0 * self.P_amp)
        self.dLambda = 2.0 * np.pi / (2.0 * self.X_amp)
        # host grids (:1916-1927)
        self.x_host = host_axis(self.X_amp, self.X, self.dX)[None, :]
        self.p_host = host_axis(self.P_amp, self.P, self.dP)[:, None]
        self.theta_host = host_axis(self.dTheta * self.P / 2.0, self.P, self.dTheta)[:, None]
        self.lambda_host = host_axis(self.dLambda * self.X / 2.0, self.X, self.dLambda)[None, :]
        tr = percent_f if truncate_constants else float
        self.k_mass, self.k_c, self.k_dt = tr(mass), tr(c), tr(dt)
        self.k_mh = tr(mass / 2.0 if kinetic_mass_half is None else kinetic_mass_half)
        self.mass_potential = mass_potential
        self.k_dx, self.k_dp = tr(self.dX), tr(self.dP)
        self.k_dtheta, self.k_dlambda = tr(self.dTheta), tr(self.dLambda)
        self.k_D = tr(D_theta)
        self.kx = fft_order_axis(self.X, self.k_dx)[None, :]
        self.kp = fft_order_axis(self.P, self.k_dp)[:, None]
        self.ktheta = fft_order_axis(self.P, self.k_dtheta)[:, None]
        self.klambda = fft_order_axis(self.X, self.k_dlambda)[None, :]
        self.V = list(potentials)
        self.dV0 = dpotential0 if dpotential0 is not None else (lambda t, x: 0.0 * x)
        self.gamma = float(gamma)
        self.damping = damping
        self.k_lambda_bar = tr(lambda_bar) if lambda_bar is not None else None  # ODM (:1947)
        self.gp = float(gp_coeff)
        self.absorb_on = absorb
        self.antiparticle = antiparticle
        self.compute_energy = compute_energy
        self._kin = None
        self._pot = {}

    # ------------------------------------------------------------------ transforms
    @staticmethod
    def x_to_lambda(W):
        return np.fft.fft(W, axis=-1)

    @staticmethod
    def lambda_to_x(W):
        return np.fft.ifft(W, axis=-1)

    @staticmethod
    def p_to_theta(W):
        return np.fft.fft(W, axis=-2)

    @staticmethod
    def theta_to_p(W):
        return np.fft.ifft(W, axis=-2)

    # ------------------------------------------------------------------ per-cell factors
    def kinetic_factor(self, q, sign):
        """U(q) of DiracPropagator_P_plus_Lambda (sign +1) / P_minus_Lambda (sign -1, dt -> -dt),
        :183-199 and :282-298.  Returns [4, 4, ...]."""
        mh, c, dt = self.k_mh, self.k_c, self.k_dt
        s = np.sqrt((mh * c) ** 2 + q ** 2)
        cs, sn = np.cos(dt * c * s), np.sin(sign * c * dt * s)
        z = np.zeros(np.broadcast(q, q).shape, dtype=np.complex128)
        U = np.zeros((4, 4) + z.shape, dtype=np.complex128)
        mass_term = 0.0 if self.kvn else 1.0  # the KvN diagonal is cos alone (KvN.py:188-189)
        U11 = cs - 1j * mass_term * mh * c * sn / s
        U44 = cs + 1j * mass_term * mh * c * sn / s
        U14 = -1j * q * sn / s
        U[0, 0], U[3, 3], U[0, 3], U[3, 0] = U11, U44, U14, U14
        U[1, 1], U[1, 2], U[2, 1], U[2, 2] = U11, U14, U14, U44
        return U

    def kinetic_step(self, W):
        """In the (p, lambda) picture: W <- U(p + lambda/2) W (U(p - lambda/2) with dt -> -dt), :3537-3551."""
        if self._kin is None:
            q0 = 0.0 * self.kp if self.kvn else self.kp  # KvN: q = +/- lambda/2 (KvN.py:184, 284)
            self._kin = (self.kinetic_factor(q0 + 0.5 * self.klambda, +1),
                         self.kinetic_factor(q0 - 0.5 * self.klambda, -1))
        L, R = self._kin
        W = np.einsum("ik...,kj...->ij...", L, W)
        return np.einsum("ik...,kj...->ij...", W, R)

    def potential_factor(self, t, sign, B=None):
        """P- (sign -1, x - theta/2) and P+ (sign +1, x + theta/2) with their scalar e^{...}, :406-441, :566-596."""
        x = self.kx + sign * 0.5 * self.ktheta
        c, dt = self.k_c, self.k_dt
        mh = self.k_mh if self.mass_potential is None else np.asarray(self.mass_potential(t, x), dtype=np.float64)
        A1, A2, A3 = (np.broadcast_to(np.asarray(self.V[k](t, x), dtype=np.float64), x.shape) for k in (1, 2, 3))
        F = np.sqrt((mh * c * c * dt) ** 2 + (A1 * A1 + A2 * A2 + A3 * A3) * dt * dt)
        sF = np.sin(F) / F
        s = float(sign)  # the + factor flips the sign of every off-diagonal and mass term
        U = np.zeros((4, 4) + x.shape, dtype=np.complex128)
        U11 = np.cos(F) + s * 1j * mh * c * c * dt * sF
        U33 = np.cos(F) - s * 1j * mh * c * c * dt * sF
        U13 = -s * 1j * dt * A3 * sF
        U14 = -s * dt * (1j * A1 + A2) * sF
        U23 = -s * dt * (1j * A1 - A2) * sF
        U[0, 0], U[1, 1], U[2, 2], U[3, 3] = U11, U11, U33, U33
        U[0, 2], U[0, 3], U[1, 2], U[1, 3] = U13, U14, U23, -U13
        U[2, 0], U[2, 1], U[3, 0], U[3, 1] = U13, U14, U23, -U13
        V0 = np.asarray(self.V[0](t, x), dtype=np.float64)
        gp = 0.0 if B is None else self.gp * np.real(B)
        e = np.exp(-dt * self.k_D * self.ktheta ** 2 / 2.0 + s * 1j * dt * (V0 + gp))
        return U, e

    def potential_step(self, W, t, Bm=None, Bp=None):
        """In the (theta, x) picture: W <- e- P- W ; W <- e+ W P+, :3576-3591.  KvN: every plane
        times e^{-dt D theta^2 + i dt V0'(x) theta} (KvN_XTheta_Kernel, KvN.py:356-420)."""
        if self.kvn:
            dv = np.asarray(self.dV0(t, self.kx), dtype=np.float64)
            th = self.ktheta
            return W * np.exp(-self.k_dt * self.k_D * th * th + 1j * self.k_dt * dv * th)
        Um, em = self.potential_factor(t, -1, Bm)
        Up, ep = self.potential_factor(t, +1, Bp)
        W = em * np.einsum("ik...,kj...->ij...", Um, W)
        return ep * np.einsum("ik...,kj...->ij...", W, Up)

    # ------------------------------------------------------------------ damping, GP, absorber
    def cl_order2(self, W):
        """CaldeiraDissipatorOrder2 per plane (:3019-3058): D(Y) = F_p^-1[theta F_p[p Y]] on host grids;
        LW2 = W + a D W; W += 2 a D LW2, a = i dt gamma."""
        a = 1j * self.dt * self.gamma

        def D(Y):
            return self.theta_to_p(self.p_to_theta(Y * self.p_host) * self.theta_host)

        LW2 = W + a * D(W)
        return W + 2.0 * a * D(LW2)

    def absorber(self):
        """1 - exp(-((j - X/2)/24)^2) over the array index j (CUDAsource_AbsorbBoundary_x, :111-155)."""
        j = np.arange(self.X)
        return (1.0 - np.exp(-(((j - self.X // 2) / 24.0) ** 2)))[None, :]

    def odm(self, W):
        """DampingODM_Kernel (:952-1195) in the (theta, x) picture.  Every one of its 16 outputs
        has the form (with i' = 5 - i, j' = 5 - j, kernel coordinates, t = dt, F^2 = lambdaBar):
          W'_ij = 1/2 [ A cx (W_ij - W_i'j') + B ct (W_ij + W_i'j')
                        + i A sx (W_ij' - W_i'j) + i B st (W_ij' + W_i'j) ],
          A = e^{-8 m c t g x^2 / F^2}, cx + i sx = e^{4 i m c t g x},
          B = e^{-2 m c t g theta^2 / F^2}, ct + i st = e^{2 i m c t g theta}."""
        k = self.k_mass * self.k_c * self.k_dt * self.gamma
        lb = self.k_lambda_bar
        x, th = self.kx, self.ktheta
        A, cx, sx = np.exp(-8.0 * k * x * x / lb), np.cos(4.0 * k * x), np.sin(4.0 * k * x)
        B, ct, st = np.exp(-2.0 * k * th * th / lb), np.cos(2.0 * k * th), np.sin(2.0 * k * th)
        out = np.empty_like(W)
        for i in range(4):
            for j in range(4):
                ib, jb = 3 - i, 3 - j
                out[i, j] = 0.5 * (A * cx * (W[i, j] - W[ib, jb]) + B * ct * (W[i, j] + W[ib, jb])
                                   + 1j * A * sx * (W[i, jb] - W[ib, j]) + 1j * B * st * (W[i, jb] + W[ib, j]))
        return out

    def filter_particles(self, W):
        """FilterAntiParticles (:2500-2540, applied at :3636-3678): in the (p, lambda) picture
        W <- Lambda_+(p + lambda/2) W Lambda_+(p - lambda/2) (KvN: lambda = 0 in the projectors,
        KvN.py:775), back to x."""
        h = 0.0 if self.kvn else 0.5 * self.klambda
        L = self.projector(self.kp + h, +1)
        R = self.projector(self.kp - h, +1)
        F = np.einsum("ik...,kj...->ij...", L, self.x_to_lambda(W))
        return self.lambda_to_x(np.einsum("ik...,kj...->ij...", F, R))

    def gp_terms(self, W):
        """B-(theta,x) = rho(x - theta/2), B+ = rho(x + theta/2) via the shift theorem (:2695-2715)."""
        rho = np.real(sum(W[k // 4, k % 4] for k in DIAG).sum(axis=0))
        rt = np.fft.fft(rho.astype(np.complex128))[None, :]
        ph = np.exp(0.5j * self.lambda_host * self.theta_host)
        return np.fft.ifft(rt / ph, axis=1), np.fft.ifft(rt * ph, axis=1)

    # ------------------------------------------------------------------ diagnostics
    def trace(self, W):
        return W[0, 0] + W[1, 1] + W[2, 2] + W[3, 3]

    def norm(self, W):
        return float(np.real(self.trace(W).sum())) * self.dX * self.dP

    def observables(self, W, t):
        """Pre-step record of the Run loop (:3429-3492).  Complex sums as the reference keeps them."""
        d = self.dX * self.dP
        x, p = self.kx, self.kp
        tr = self.trace(W)
        a1 = W[0, 3] + W[1, 2] + W[2, 1] + W[3, 0]
        dv = np.asarray(self.dV0(t, x), dtype=np.float64)
        out = {
            "X": d * np.sum(tr * x), "P": d * np.sum(tr * p), "XP": d * np.sum(tr * x * p),
            "XX": d * np.sum(tr * x * x), "PP": d * np.sum(tr * p * p),
            "Alpha1": d * np.sum(a1), "D1V0": d * np.sum(tr * dv), "XD1V0": d * np.sum(tr * x * dv),
            "PD1V0": d * np.sum(tr * p * dv), "PAlpha1": d * np.sum(a1 * p), "XAlpha1": d * np.sum(a1 * x),
            "negativity": d * float(np.sum(np.minimum(np.real(tr), 0.0))),
            "transmission": d * float(np.real(np.sum(np.where(np.broadcast_to(x, tr.shape) > 10.0, tr, 0.0)))),
        }
        if self.compute_energy:
            c, m = self.c, self.mass
            e = c * d * np.sum(a1 * self.p_host)
            e += c * c * m * d * np.sum(W[0, 0] + W[1, 1] - W[2, 2] - W[3, 3])
            for k in range(4):
                e += d * np.sum(tr * np.asarray(self.V[k](t, x), dtype=np.float64))
            out["energy"] = e
        return out

    def projector(self, q, sign):
        """DiracElectronProjector at kernel coordinates (Filter_Kernel, :674-760)."""
        m, c = self.k_mass, self.k_c
        sq = 2.0 * np.sqrt(m * m * c ** 4 + c * c * q * q)
        aa = 0.5 + sign * m * c * c / sq
        bb = c * sign * q / sq
        dd = 0.5 - sign * m * c * c / sq
        Z = np.zeros((4, 4) + np.broadcast(q, q).shape)
        Z[0, 0] = Z[1, 1] = aa
        Z[2, 2] = Z[3, 3] = dd
        Z[0, 3] = Z[2, 1] = Z[3, 0] = Z[1, 2] = bb
        return Z

    def antiparticle_population(self, W_lambda, sign=-1):
        """AntiParticlePopulation (sign -1) / ParticlePopulation (sign +1), :2436-2498: filter with
        the projectors in the (p, lambda) picture, back to x on the diagonal planes, trace norm."""
        L = self.projector(self.kp + 0.5 * self.klambda, sign)
        R = self.projector(self.kp - 0.5 * self.klambda, sign)
        F = np.einsum("ik...,kj...->ij...", L, W_lambda)
        F = np.einsum("ik...,kj...->ij...", F, R)
        tr = sum(self.lambda_to_x(F[i, i]) for i in range(4))
        return float(np.real(tr.sum())) * self.dX * self.dP

    # ------------------------------------------------------------------ driver
    def step(self, W, t_index, record=None):
        """One iteration of the Run loop (:3409-3690) at t = dt * t_index."""
        t = self.dt * t_index
        Bm = Bp = None
        if self.gp != 0.0:
            Bm, Bp = self.gp_terms(W)
        if record is not None:
            record.append(self.observables(W, t))
        W = self.x_to_lambda(W)
        if record is not None and self.antiparticle:
            record[-1]["antiparticle"] = self.antiparticle_population(W)
        W = self.kinetic_step(W)
        W = self.lambda_to_x(W)
        W = self.p_to_theta(W)
        W = self.potential_step(W, t, Bm, Bp)
        if self.gamma > 0.0 and self.damping == "ODM":
            W = self.odm(W)
        W = self.theta_to_p(W)
        if self.gamma > 0.0 and self.damping == "CaldeiraLegget":
            W = self.cl_order2(W)
        if self.step_filter:
            W = self.filter_particles(W)
        if self.absorb_on:
            W = W * self.absorber()
        return W / self.norm(W)

    def run(self, W0, time_steps):
        """timeSteps iterations (t_index = 1..timeSteps).  Returns (W, list of per-step records)."""
        W = np.array(W0, dtype=np.complex128)
        rec = []
        for n in range(1, int(time_steps) + 1):
            W = self.step(W, n, rec)
        return W, rec


def gaussian_spinor_particle_up(x0, px, s, X, mass, c):
    """GaussianSpinor_ParticleUp (:2231-2247)."""
    rho = np.exp(1j * X * px) * np.exp(-0.5 * ((X - x0) / s) ** 2) + 0j
    p0 = np.sqrt(px * px + mass * mass * c * c)
    return np.array([rho * (p0 + mass * c), X * 0j, X * 0j, rho * px])


def wigner_from_spinor(spinor, orc: Dirac4x4Oracle):
    """W_init of the notebooks (instances/WignerDirac2D_4x4/Free.ipynb cell 2): the spinor
    evaluated at x -/+ theta/2 on the host grids, W_ij = psiL_i conj(psiR_j) in (x, theta),
    then theta -> p (Fourier_4X4_Theta_To_P, :2545-2548; fftpack.ifft along axis 0) and
    division by the trace norm (:2677-2683)."""
    psiL = spinor(orc.x_host - 0.5 * orc.theta_host)
    psiR = spinor(orc.x_host + 0.5 * orc.theta_host)
    W = np.empty((4, 4, orc.P, orc.X), dtype=np.complex128)
    for i in range(4):
        for j in range(4):
            W[i, j] = np.fft.ifft(psiL[i] * np.conj(psiR[j]), axis=0)
    return W / orc.norm(W)
