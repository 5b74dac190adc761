"""Wigner2D: one P x X Wigner function propagated by libwignerhip on one GPU.

This is the layer under the drop-in classes (GPU_Wigner2D_FFT,
GPU_Wigner2D_GPitaevskii).  It turns the reference's physical set-up into the
C descriptor: kernel constants (optionally '%f'-truncated as the reference's
source splicing does, GPU_Wigner2D_FFT.py:833-842), polynomial coefficient
tables for V(t,x) and K(p), and the 1-D host tables (decays, absorbing mask,
friction sigma(p), host grids and dV/dx for the observables).

State tensors are torch complex tensors on the GPU, [P][X] row-major,
FFT ordered, exactly like the reference's W_GPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import WH_C128, WH_C64, WH_MAX_DEG, WH_NOBS, WH_SCHEME_LIE, WH_SCHEME_STRANG
from .physics import coefficient_table, depends_on_time, host_eval, polynomial_coefficients


def percent_f(v: float) -> float:
    """A constant after '%f' splicing into kernel source (GPU_Wigner2D_FFT.py:833-842)."""
    return float("%f" % float(v))


def fft_order_axis(n: int, step: float) -> np.ndarray:
    """Kernel coordinate of FFT-ordered index k (GPU_Wigner2D_FFT.py:110-114)."""
    k = np.arange(n)
    return step * (((k + n // 2) % n) - 0.5 * n)


def host_axis(amplitude: float, n: int, step: float) -> np.ndarray:
    """fftshift(linspace(-A, A - d, n)) (GPU_Wigner2D_FFT.py:571-581)."""
    return np.fft.fftshift(np.linspace(-amplitude, amplitude - step, n))


class Wigner2D:
    """Fused split-operator propagator for one Wigner function.

    Parameters follow GPU_Wigner2D_FFT.__init__ (GPU_Wigner2D_FFT.py:1718-1720)
    and the attributes its Run() reads:

    potential, dpotential, kinetic, fp_damping : reference-style C expression strings
        (V(t,x), dV/dx(t,x), K(p), f(p)); V and K must be polynomials of degree
        <= 8 in x / p (coefficients may depend on t).
    scheme      : 'lie' (K; V -- GPU_Wigner2D_FFT) or 'strang' (V/2; K; V/2 --
                  GPU_Wigner2D_GPitaevskii).
    kinetic     : None in the Lie scheme = the reference's hard-wired dt*lambda*p/m.
    damping     : None or 'CaldeiraLeggett' (order-3 composite).
    gp_coeff    : Gross-Pitaevskii coefficient a_GP (0 = off).
    mask        : absorbing edge mask of the potential kernel (on in the reference).
    truncate_constants : emulate the '%f' splicing of kernel constants.
    precision   : 'fp64' (complex128 state) or 'fp32' (complex64 state, FP64 phases,
                  twiddles computed in FP64, FP64 observable accumulation).
    time_rows   : rows of the coefficient table for a time-dependent V (steps + 2).
    pot_table   : optional explicit [rows, 9] coefficient table overriding `potential`
                  on the device (the string is still used for host observables).
    """

    def __init__(self, X, P, X_amp, P_amp, hbar, mass, dt, *, potential, dpotential=None,
                 kinetic=None, scheme="lie", D_theta=0.0, D_lambda=0.0, gamma=0.0,
                 damping=None, fp_damping=None, gp_coeff=0.0, gp_rho_factor=None, mask=True,
                 truncate_constants=False, precision="fp64", time_rows=None, pot_table=None,
                 device=None):
        import torch

        self.X, self.P = int(X), int(P)
        self.X_amp, self.P_amp = float(X_amp), float(P_amp)
        self.hbar, self.mass, self.dt = float(hbar), float(mass), float(dt)
        self.scheme = scheme
        self.precision = precision
        self.dtype = torch.complex128 if precision == "fp64" else torch.complex64
        self.device = torch.device(device if device is not None else "cuda:0")
        self.potential_str, self.dpotential_str = potential, dpotential
        self.gamma, self.damping, self.gp_coeff = float(gamma), damping, float(gp_coeff)
        self.renormalize = self.gamma != 0.0
        tr = percent_f if truncate_constants else float
        # host grids (SetPhaseSpaceBox2D / SetPhaseSpaceGrid2D)
        self.dX = 2.0 * self.X_amp / self.X
        self.dP = 2.0 * self.P_amp / self.P
        self.dTheta = 2.0 * np.pi / (2.0 * self.P_amp)
        self.dLambda = 2.0 * np.pi / (2.0 * self.X_amp)
        x_host = host_axis(self.X_amp, self.X, self.dX)
        p_host = host_axis(self.P_amp, self.P, self.dP)
        theta_host = host_axis(self.dTheta * self.P / 2.0, self.P, self.dTheta)
        lambda_host = host_axis(self.dLambda * self.X / 2.0, self.X, self.dLambda)
        self.x_host, self.p_host = x_host, p_host
        # kernel constants
        k_dt, k_dx, k_dp = tr(dt), tr(self.dX), tr(self.dP)
        k_dtheta, k_dlambda = tr(self.dTheta), tr(self.dLambda)
        k_mass, k_hbar = tr(mass), tr(hbar)
        k_Dtheta, k_Dlambda, k_agp = tr(D_theta), tr(D_lambda), tr(gp_coeff)
        kx = fft_order_axis(self.X, k_dx)
        kp = fft_order_axis(self.P, k_dp)
        ktheta = fft_order_axis(self.P, k_dtheta)
        klambda = fft_order_axis(self.X, k_dlambda)
        frac = 1.0 if scheme == "lie" else 0.5
        theta_decay = np.exp(-frac * k_Dtheta * ktheta * ktheta * k_dt)
        lambda_decay = np.exp(-k_Dlambda * klambda * klambda)
        xm = k_dx * (self.X - 1.0) / 2.0

        def edge(w):
            w2 = (w * k_dx) ** 2
            return (1.0 - np.exp(-((kx - xm) ** 2) / w2)) * (1.0 - np.exp(-((kx + xm) ** 2) / w2))

        x_mask = None
        if mask:
            x_mask = edge(10.0)
            if self.gp_coeff != 0.0 and scheme == "lie":
                x_mask = x_mask * edge(20.0)  # GPU_Wigner2D_FFT.py:514-516
        # potential coefficients (time rows)
        self.time_dependent = depends_on_time(potential, "x") if pot_table is None else pot_table.shape[0] > 1
        if pot_table is not None:
            deg, _ = polynomial_coefficients(potential, "x")
            tab = np.ascontiguousarray(pot_table, dtype=np.float64)
            deg = max(deg, int(np.max(np.nonzero(np.any(tab != 0, axis=0))[0], initial=0)))
        elif self.time_dependent:
            rows = int(time_rows) if time_rows else 4096
            deg, tab = coefficient_table(potential, "x", np.arange(rows) * self.dt)
        else:
            deg, tab = coefficient_table(potential, "x", [0.0])
        self.pot_deg, self.pot_table = deg, np.ascontiguousarray(tab)
        # kinetic
        kin_poly, kdeg, kcoef = 0, 0, np.zeros(WH_MAX_DEG + 1)
        if kinetic is not None and scheme != "lie":
            kdeg, ktab = coefficient_table(kinetic, "p", [0.0])
            kin_poly, kcoef = 1, ktab[0]
        elif scheme != "lie":
            kin_poly, kdeg = 1, 2
            kcoef[2] = 0.5 / k_mass
        # damping tables
        sigma = None
        if damping == "CaldeiraLeggett" and self.gamma != 0.0:
            f = np.broadcast_to(np.asarray(host_eval(fp_damping, p=kp), dtype=np.float64), kp.shape)
            sigma = np.where(kp >= 0.0, f, -f)
        # observables
        dv = np.broadcast_to(np.asarray(host_eval(dpotential, t=0.0, x=x_host) if dpotential else 0.0 * x_host,
                                        dtype=np.float64), x_host.shape)
        v0 = np.broadcast_to(np.asarray(host_eval(potential, t=0.0, x=x_host), dtype=np.float64), x_host.shape)
        self._keep = []  # host arrays referenced by the descriptor until create() copied them

        def ptr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(a)
            return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

        d = _lib.Wigner2DDesc()
        d.struct_size = ctypes.sizeof(_lib.Wigner2DDesc)
        d.precision = WH_C128 if precision == "fp64" else WH_C64
        d.x_dim, d.p_dim = self.X, self.P
        d.scheme = WH_SCHEME_LIE if scheme == "lie" else WH_SCHEME_STRANG
        d.dt, d.dx, d.dp, d.dtheta, d.dlambda, d.hbar, d.mass = k_dt, k_dx, k_dp, k_dtheta, k_dlambda, k_hbar, k_mass
        d.kinetic_poly, d.kinetic_deg = kin_poly, kdeg
        for n in range(WH_MAX_DEG + 1):
            d.kin_coef[n] = float(kcoef[n])
        d.pot_deg, d.pot_rows = self.pot_deg, self.pot_table.shape[0]
        d.pot_coef = ptr(self.pot_table)
        d.theta_decay = ptr(theta_decay) if k_Dtheta != 0.0 else None
        d.lambda_decay = ptr(lambda_decay) if k_Dlambda != 0.0 else None
        d.x_mask = ptr(x_mask)
        d.damping = 1 if sigma is not None else 0
        d.gamma_dt = self.dt * self.gamma
        d.sigma_p = ptr(sigma)
        d.theta_host = ptr(theta_host)
        d.renormalize = 1 if self.renormalize else 0
        d.gp_coeff = k_agp
        d.gp_rho_factor = float(gp_rho_factor) if gp_rho_factor is not None else (1.0 if scheme == "lie" else self.dP)
        d.lambda_host = ptr(lambda_host)
        d.x_host, d.p_host = ptr(x_host), ptr(p_host)
        d.dv_host, d.v0_host = ptr(dv), ptr(v0)
        d.dxdp = self.dX * self.dP
        d.inv_2m = 1.0 / (2.0 * self.mass)
        d.dx_host = self.dX
        self.desc = d
        self._plan = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().wh_wigner2d_create(ctypes.byref(d), ctypes.byref(self._plan)),
                       "wh_wigner2d_create")
        self._keep = []

    # ------------------------------------------------------------------ lifetime
    def close(self):
        if getattr(self, "_plan", None) and self._plan.value:
            _lib.lib().wh_wigner2d_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ state helpers
    def to_device(self, W) -> "torch.Tensor":
        import torch

        t = torch.as_tensor(np.ascontiguousarray(W)).to(self.dtype)
        if t.shape != (self.P, self.X):
            raise ValueError(f"W must have shape (P, X) = {(self.P, self.X)}, got {tuple(t.shape)}")
        return t.to(self.device).contiguous()

    def _check_state(self, W):
        if W.dtype != self.dtype or not W.is_cuda or not W.is_contiguous() or W.shape != (self.P, self.X):
            raise ValueError(f"state must be a contiguous {self.dtype} CUDA tensor of shape {(self.P, self.X)}")

    # ------------------------------------------------------------------ stepping
    def run(self, W, n_steps: int, first_step: int = 0, normalize_initial: bool = True,
            use_graph: bool = True, obs=None):
        """Advances W in place by n_steps iterations; returns the [n_steps+1, 16] record (device)."""
        import torch

        self._check_state(W)
        if self.time_dependent and first_step + n_steps + 1 >= self.pot_table.shape[0]:
            raise ValueError(f"time-dependent potential table has {self.pot_table.shape[0]} rows; "
                             f"need {first_step + n_steps + 2} (pass time_rows)")
        if obs is None:
            obs = torch.zeros((n_steps + 1, WH_NOBS), dtype=torch.float64, device=W.device)
        _lib.check(_lib.lib().wh_wigner2d_run(self._plan, _lib.dptr(W), int(first_step), int(n_steps),
                                              _lib.dptr(obs), int(bool(normalize_initial)),
                                              int(bool(use_graph)), _lib.current_stream_handle()),
                   "wh_wigner2d_run")
        return obs

    def kinetic(self, W):
        self._check_state(W)
        _lib.check(_lib.lib().wh_wigner2d_kinetic(self._plan, _lib.dptr(W), _lib.current_stream_handle()),
                   "wh_wigner2d_kinetic")

    def potential(self, W, step: int = 0, fused: bool = False):
        self._check_state(W)
        _lib.check(_lib.lib().wh_wigner2d_potential(self._plan, _lib.dptr(W), int(step), int(bool(fused)),
                                                    _lib.current_stream_handle()),
                   "wh_wigner2d_potential")

    def observe(self, W):
        import torch

        self._check_state(W)
        obs = torch.zeros((1, WH_NOBS), dtype=torch.float64, device=W.device)
        _lib.check(_lib.lib().wh_wigner2d_observe(self._plan, _lib.dptr(W), _lib.dptr(obs),
                                                  _lib.current_stream_handle()),
                   "wh_wigner2d_observe")
        return obs[0]
