"""Oracle: NumPy complex128 restatement of the imaginary-time (Bloch) Gross-Pitaevskii
ground-state solver, GPU_Wigner2D_GPitaevskii_Bloch (GPU_Wigner2D_GPitaevskii.py:2064-2543).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows in behaviour (not in code):

* Bloch potential factor ........ Kernel_Bloch of expPotential, GP.py:56-59, 88-110:
                                  W(theta, x) *= exp(-dt/4 [V(t, x - hbar theta/2) + V(t, x + hbar theta/2)])
                                  (no mask, no D_Theta; dt is the current, possibly halved, step)
* Bloch kinetic factor .......... Kernel_Bloch of expPLambdaKinetic, GP.py:127-130, 160-175:
                                  W(p, lambda) *= exp(-dt/2 [K(p + hbar lambda/2) + K(p - hbar lambda/2)])
* Bloch Gross-Pitaevskii factor . GP.py:498-517: W *= exp(-dt/4 a (B- + B+)), B-/+ = rho(x -/+ theta/2),
                                  rho = dP sum_p W (ProbabilityX, :1743-1748; shift theorem :1717-1735)
* step ...................... Run, GP.py:2183-2245: [GP terms] F_p, V/4, [GP], F_p^-1, F_x, K/2, F_x^-1,
                              [GP terms], F_p, V/4, [GP], F_p^-1, W /= dXdP sum W
* energies and purity ........... GP.py:2170-2176, 2252-2268: Hamiltonian = K(P) + V(0, X) on the host
                                  grids (host Heaviside), E = dXdP sum W H + dX a/2 sum rho^2,
                                  purity = 2 pi hbar dXdP sum W^2 (complex, no conjugate)
* acceptance ................ GP.py:2276-2292: from the second iteration on, keep the step if
                              E_step < E (complex order: real part first) and |purity| < 1,
                              else halve dt and restore the last kept state (initially W_init as
                              given, GP.py:2146)
* excited state ............. Run_ExitedState1, GP.py:2343-2535: after each step W -= c W_0 with
                              c = 2 pi hbar dXdP sum W W_0, before the normalisation.  Its source
                              also refuses a step when the p-marginal ProbabilityP is negative
                              (:2478-2480), but that marginal's kernel (gpu_sum_axis1, :422-441)
                              stores row p's sum at index p*X of a P-element array (out of bounds
                              for p > 0), and the recorded run (Wigner2D_GPitaevskii_X2_Bloch.ipynb
                              cell 20: four refusals, at dt = 0.01 .. 0.00125 after step 700, and
                              the min/max of W_1) is reproduced exactly by the loop without that
                              test (with it, refusals start at step 153, where the p = 0 marginal
                              turns negative).  The oracle follows the recorded run.

Kernel constants (dx, dp, dtheta, dlambda, hbar, a_GP) are '%f'-spliced in the reference
(GP.py:854-863, 2087-2090); dt is a full-precision kernel argument.  Pinned in
tests/test_oracle_bloch.py: the exact fixed point of the split Bloch map for the harmonic
oscillator (a = 0: a Gaussian whose width solves the composed Gaussian maps), monotone
energies, and (slow) the recorded minima/maxima and refusals of the X2 notebook runs.
"""
from __future__ import annotations

import numpy as np

from .wigner2d import Grid, fft, fft_order_axis, ifft, percent_f


def _lex_less(a: complex, b: complex) -> bool:
    """Order of numpy complex scalars (real part, then imaginary part)."""
    a, b = complex(a), complex(b)
    return a.real < b.real or (a.real == b.real and a.imag < b.imag)


class GPBlochOracle:
    """potential(t, x), kinetic(p): vectorised callables evaluated where the reference's kernels
    evaluate the spliced strings; potential_host(t, x), kinetic_host(p): the host-side
    Potential / KineticEnergy of the Hamiltonian grid (they differ only in Heaviside(0))."""

    def __init__(self, X, P, X_amp, P_amp, hbar, mass, dt, potential, kinetic, gp_coeff=0.0,
                 potential_host=None, kinetic_host=None, truncate_constants=True):
        self.g = Grid(X, P, X_amp, P_amp)
        g = self.g
        tr = percent_f if truncate_constants else float
        self.hbar, self.mass, self.dt = float(hbar), float(mass), float(dt)
        self.a = float(gp_coeff)
        self.k_dx, self.k_dp = tr(g.dX), tr(g.dP)
        self.k_dtheta, self.k_dlambda = tr(g.dTheta), tr(g.dLambda)
        self.k_hbar, self.k_a = tr(hbar), tr(gp_coeff)
        self.potential, self.kinetic = potential, kinetic
        self.potential_host = potential_host or potential
        self.kinetic_host = kinetic_host or kinetic
        self.kx = fft_order_axis(g.X, self.k_dx)[None, :]
        self.kp = fft_order_axis(g.P, self.k_dp)[:, None]
        self.ktheta = fft_order_axis(g.P, self.k_dtheta)[:, None]
        self.klambda = fft_order_axis(g.X, self.k_dlambda)[None, :]
        self.H = self.kinetic_host(g.p_host) + self.potential_host(0.0, g.x_host)  # GP.py:2094

    # ---------------------------------------------------------------- factors
    def vsum(self, t):
        h, th, x = self.k_hbar, self.ktheta, self.kx
        return self.potential(t, x - h * th / 2.0) + self.potential(t, x + h * th / 2.0)

    def ksum(self):
        h, lam, p = self.k_hbar, self.klambda, self.kp
        return self.kinetic(p + h * lam / 2.0) + self.kinetic(p - h * lam / 2.0)

    def density(self, W):
        """ProbabilityX: dP sum_p W (complex, GP.py:1743-1748)."""
        return W.sum(axis=0) * self.g.dP

    def shifted_sum(self, rho):
        """B- + B+ with B-/+ = rho(x -/+ theta/2) (MakeGrossPitaevskiiTerms, GP.py:1717-1735;
        e^{i lambda theta / 2} on the host grids without hbar, FFT.py:594-595)."""
        g = self.g
        rt = fft(rho)[None, :]
        ph = np.exp(0.5j * g.lambda_host * g.theta_host)
        return ifft(rt / ph, axis=1) + ifft(rt * ph, axis=1)

    def potential_half(self, W, dt, t, B):
        Y = fft(W, axis=0) * np.exp(-0.25 * dt * self.vsum(t))
        if B is not None:
            Y = Y * np.exp(-0.25 * dt * self.k_a * B)
        return ifft(Y, axis=0)

    def step(self, W, dt, t):
        """One Bloch step before the normalisation (GP.py:2183-2238)."""
        gp = self.a != 0.0
        B = self.shifted_sum(self.density(W)) if gp else None
        W = self.potential_half(W, dt, t, B)
        W = ifft(fft(W, axis=1) * np.exp(-0.5 * dt * self.ksum()), axis=1)
        B = self.shifted_sum(self.density(W)) if gp else None
        return self.potential_half(W, dt, t, B)

    # ---------------------------------------------------------------- diagnostics
    def energies(self, W):
        """(total energy, non-linear energy), complex as the reference forms them."""
        g = self.g
        dxdp = g.dX * g.dP
        E = dxdp * np.sum(W * self.H)
        rho = self.density(W)
        nle = g.dX * 0.5 * self.a * np.sum(rho ** 2) if self.a != 0.0 else 0.0
        return E + nle, nle

    def purity(self, W):
        return 2.0 * np.pi * self.hbar * np.sum(W * W) * self.g.dX * self.g.dP

    def observables(self, W):
        g = self.g
        dxdp = g.dX * g.dP
        x, p = g.x_host, g.p_host
        return {"X": dxdp * np.sum(W * x), "X2": dxdp * np.sum(W * x * x), "P": dxdp * np.sum(W * p),
                "P2": dxdp * np.sum(W * p * p), "H": dxdp * np.sum(W * self.H)}

    # ---------------------------------------------------------------- drivers
    def run(self, W_init, time_steps, W_ground=None):
        """Run (W_ground None, GP.py:2105-2340) or Run_ExitedState1 (GP.py:2343-2535):
        time_steps + 1 iterations.  Returns (W_end, record dict)."""
        g = self.g
        dxdp = g.dX * g.dP
        W0 = np.array(W_init, dtype=np.complex128)
        W = W0 / (W0.sum() * dxdp)
        W_step = W0.copy()  # the last kept state starts as W_init as given (GP.py:2146, 2361)
        E_acc, _ = self.energies(W)
        dt = self.dt
        rec = {k: [] for k in ("time", "X", "X2", "P", "P2", "H", "E", "NLE", "purity", "accepted", "dt")}
        # the kept steps' series (TotalEnergyHistory, NonLinearEnergyHistory, purity of the reference)
        hist = {"TotalEnergyHistory": [], "NonLinearEnergyHistory": [], "purityHistory": []}
        for ti in range(time_steps + 1):
            t = ti * self.dt
            W = self.step(W, dt, t)
            if W_ground is not None:
                c = 2.0 * np.pi * self.hbar * dxdp * np.sum(W * W_ground)
                W = W - c * W_ground
            W = W / (W.sum() * dxdp)
            ob = self.observables(W)
            E, nle = self.energies(W)
            pur = self.purity(W)
            rec["time"].append(t)
            for k in ("X", "X2", "P", "P2", "H"):
                rec[k].append(ob[k])
            rec["E"].append(E)
            rec["NLE"].append(nle)
            rec["purity"].append(pur)
            rec["dt"].append(dt)
            acc = -1
            if ti > 0:
                if _lex_less(E, E_acc) and abs(pur) < 1.0:
                    W_step = W.copy()
                    hist["purityHistory"].append(pur)
                    hist["TotalEnergyHistory"].append(E)
                    hist["NonLinearEnergyHistory"].append(nle)
                    E_acc = E
                    acc = 1
                else:
                    dt = dt / 2.0
                    W = W_step.copy()
                    acc = 0
            rec["accepted"].append(acc)
        out = {k: np.array(v) for k, v in rec.items()}
        out.update({k: np.array(v) for k, v in hist.items()})
        return W, out
