"""Fused Wigner-Dirac 4x4 propagator (libwignerhip wh_dirac_*).

    plan = WignerDirac4x4(512, 512, 18.0, 16.0, mass=1.0, c=1.0, dt=0.002,
                          potentials=("-2/sqrt(x*x+1)", "0.", "0.", "0."))
    W = plan.to_device(W_init)            # [4, 4, P, X] -> device [16, P, X]
    obs = plan.run(W, 1000)               # [1001, 32] record rows (device)

Follows GPU_WignerDirac2D_4x4 (GPU_WignerDirac2D_4x4.py:1880-3720): the kinetic
factor U(p + lambda/2) W U(p - lambda/2)^dagger with half the mass, the potential
factors with the other half, D_Theta decay, the edge absorber (24 cells) and the
trace normalisation after every step; per-step observables as in Run.  With
Potential_1..3 = 0 the potential factors are diagonal in the spinor indices and
the step is two fused sweeps; vector potentials (time-independent polynomials in
x, e.g. the constant magnetic field of the ConstantB notebooks, A2 = -x) couple
all 16 planes and take one more sweep.  The Gross-Pitaevskii term (density
Re sum_p trace W, shifted by -/+ theta/2) and both damping models (order-2
Caldeira-Leggett, ODM) are supported.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import DIRAC_OBS, DIRAC_OBS_REAL, WH_DNOBS, WH_MAX_DEG, WH_POT_EXPR, WH_POT_TABLE
from .physics import NonPolynomialPhysics, coefficient_table, compile_expression, depends_on_time, host_eval
from .wigner2d import PRECISIONS, fft_order_axis, host_axis, percent_f


def _sympy_derivative(expr: str) -> str:
    """d/dx of a physics string, as the reference builds D_1_Potential_0_String
    (sympy.diff + ccode, GPU_WignerDirac2D_4x4.py:1964-1966)."""
    import sympy as sp

    from .physics import _sympy_expr

    e, v, _ = _sympy_expr(expr, "x")
    return str(sp.diff(e, v))


def _is_zero(expr: str) -> bool:
    import sympy as sp

    from .physics import _sympy_expr

    e, _, _ = _sympy_expr(expr, "x")
    e = sp.simplify(e)
    return bool(e.is_zero) or (e.is_number and float(e) == 0.0)


class WignerDirac4x4:
    """One 16-plane Wigner-Dirac state on a P x X grid.

    potentials : four reference-style C expression strings Potential_0..3(t, x).
    D_theta    : D_Theta of the potential kernels (exp(-dt D theta^2 / 2) on each side).
    D_lambda   : accepted and unused, as in the reference (spliced, read by no kernel).
    gamma, damping : gammaDamping and dampingModel; 'CaldeiraLegget' = order-2 dissipator,
                 'ODM' = DampingODM_Kernel between the potential factors and Theta_To_P.
    lambda_bar : lambdaBar of the ODM model.
    kvn        : the KvN variant (GPU_WignerDiracKvN2D_4x4.py): kinetic factors at q = +/- lambda/2
                 without the diagonal mass term (mass_half 1e-5), potential step
                 e^{-dt D theta^2 + i dt V0'(x) theta}, the step filter always on.
    step_filter: antiParticleStepFiltering (W <- Lambda_+ W Lambda_+ after each potential half,
                 then the absorber).
    absorb     : the 24-cell edge absorber applied every step (always on in the reference).
    truncate_constants : emulate the '%f' splicing of the kernel constants.
    antiparticle, energy : record AntiParticlePopulation / DiracEnergy.
    precision  : 'fp64' (complex128 planes), 'fp32' (complex64 planes, FP64 arithmetic in the
                 sweeps) or 'fp32fast' (complex64 planes, FP32 transforms, FP64 factors and sums).
    time_rows  : rows of the coefficient tables for a time-dependent polynomial V0 or
                 time-dependent polynomial vector potentials (>= steps + 2).
    mass_potential, kinetic_mass_half : the MassPotential variant -- Mass(t, x) string of the
                 potential factors and the fixed mass_half of the kinetic factors.
    """

    def __init__(self, X, P, X_amp, P_amp, mass, c, dt, *, potentials=("0.", "0.", "0.", "0."),
                 D_theta=0.0, D_lambda=0.0, gamma=0.0, damping=None, gp_coeff=0.0, absorb=True,
                 truncate_constants=True, antiparticle=True, energy=False, precision="fp64",
                 time_rows=None, device=None, transmission_x=10.0, mass_potential=None,
                 kinetic_mass_half=None, lambda_bar=None, kvn=False, step_filter=False):
        import torch

        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        a_deg, a_coef = [-1, -1, -1], np.zeros((3, WH_MAX_DEG + 1))
        # time-dependent vector potentials: one coefficient row per time index t = dt * index,
        # as V0's rows (the reference evaluates Potential_k(t, x) per cell, :417-442, 575-601)
        a_time = any(not _is_zero(potentials[k]) and depends_on_time(potentials[k], "x") for k in (1, 2, 3))
        a_times = np.arange(int(time_rows) if time_rows else 4096) * float(dt) if a_time else [0.0]
        a_rows = np.zeros((len(a_times), 3, WH_MAX_DEG + 1))
        a_nonpoly = False
        for k in (1, 2, 3):
            if _is_zero(potentials[k]):
                continue
            try:
                deg, tab = coefficient_table(potentials[k], "x", a_times)
            except NonPolynomialPhysics:
                a_nonpoly = True
                continue
            a_deg[k - 1], a_coef[k - 1] = deg, tab[0]
            a_rows[:, k - 1, :] = tab
        # a vector potential that is not a polynomial in x: all of them as programs the device
        # evaluates per cell and step (the reference splices the strings, :417-442, 575-601)
        self._a_expr = None
        if a_nonpoly:
            try:
                self._a_expr = [None if _is_zero(potentials[k]) else compile_expression(potentials[k], "x")
                                for k in (1, 2, 3)]
            except NonPolynomialPhysics as exc:
                raise NotImplementedError(f"vector potentials {potentials[1:]!r}: {exc}") from exc
            a_deg = [-1 if p is None else 0 for p in self._a_expr]
            a_coef[:] = 0.0
            a_time = False
        if kvn:
            # GPU_WignerDiracKvN2D_4x4: V0 only, no damping or GP terms in its loop (:3409-3690)
            if max(a_deg) >= 0 or mass_potential is not None:
                raise NotImplementedError("the KvN variant takes Potential_0 only")
            gamma, damping, gp_coeff, step_filter = 0.0, None, 0.0, True
            if kinetic_mass_half is None:
                kinetic_mass_half = 0.00001  # mass_half of GPU_WignerDiracKvN2D_4x4.py:1961
        if gamma != 0.0 and damping not in (None, "CaldeiraLegget", "ODM"):
            raise NotImplementedError(f"dampingModel {damping!r} is not supported ('CaldeiraLegget' or 'ODM')")
        if gamma > 0.0 and damping == "ODM":
            if lambda_bar is None:
                raise ValueError("ODM damping needs lambdaBar")
        self.precision = precision
        self.dtype = torch.complex128 if precision == "fp64" else torch.complex64
        self.device = torch.device(device if device is not None else "cuda:0")
        self.X, self.P = int(X), int(P)
        self.X_amp, self.P_amp = float(X_amp), float(P_amp)
        self.mass, self.c, self.dt = float(mass), float(c), float(dt)
        self.dX = 2.0 * self.X_amp / self.X
        self.dP = 2.0 * self.P_amp / self.P
        self.dTheta = 2.0 * np.pi / (2.0 * self.P_amp)
        self.dLambda = 2.0 * np.pi / (2.0 * self.X_amp)
        self.x_host = host_axis(self.X_amp, self.X, self.dX)
        self.p_host = host_axis(self.P_amp, self.P, self.dP)
        self.theta_host = host_axis(self.dTheta * self.P / 2.0, self.P, self.dTheta)
        tr = percent_f if truncate_constants else float
        k_dt, k_dx, k_dp = tr(dt), tr(self.dX), tr(self.dP)
        k_dtheta, k_dlambda = tr(self.dTheta), tr(self.dLambda)
        kx = fft_order_axis(self.X, k_dx)
        ktheta = fft_order_axis(self.P, k_dtheta)
        keep = []

        def ptr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

        d = _lib.DiracDesc()
        d.struct_size = ctypes.sizeof(_lib.DiracDesc)
        d.precision = PRECISIONS[precision]
        d.x_dim, d.p_dim = self.X, self.P
        d.dt, d.dx, d.dp, d.dtheta, d.dlambda = k_dt, k_dx, k_dp, k_dtheta, k_dlambda
        k_c = tr(c)
        d.mass, d.c = tr(mass), k_c
        d.mass_half = tr(mass / 2.0 if kinetic_mass_half is None else kinetic_mass_half)
        if mass_potential is not None and depends_on_time(mass_potential, "x"):
            # Mass(t, x) as a program: the device refills the mass tables before each potential sweep
            self._mass_expr = compile_expression(mass_potential, "x")
            d.mass_expr, d.mass_expr_len, d.mass_t_step = ptr(self._mass_expr), self._mass_expr.shape[0], self.dt
        elif mass_potential is not None:
            # Mass(t, x) of the potential factors (GPU_WignerDirac2D_MassPotential_4x4.py:422-427, 585-591)
            u = 0.5 * ktheta[:, None]
            shape = (self.P, self.X)
            Fm = np.broadcast_to(np.asarray(host_eval(mass_potential, t=0.0, x=kx[None, :] - u), dtype=np.float64),
                                 shape) * k_c * k_c * k_dt
            Fp = np.broadcast_to(np.asarray(host_eval(mass_potential, t=0.0, x=kx[None, :] + u), dtype=np.float64),
                                 shape) * k_c * k_c * k_dt
            d.mass_dm, d.mass_sm = ptr(Fm - Fp), ptr(Fm + Fp)
        V0 = potentials[0]
        self.time_dependent = depends_on_time(V0, "x")
        self.a_table = a_rows if a_time else None
        if a_time:
            d.a_coef_rows, d.a_rows = ptr(a_rows), a_rows.shape[0]
        if self._a_expr is not None:
            for k, prog in enumerate(self._a_expr):
                if prog is not None:
                    d.a_expr[k], d.a_expr_len[k] = ptr(prog), prog.shape[0]
            d.a_t_step = self.dt
        try:
            times = np.arange(int(time_rows) if time_rows else 4096) * self.dt if self.time_dependent else [0.0]
            deg, tab = coefficient_table(V0, "x", times)
            self.pot_table = np.ascontiguousarray(tab)
            d.pot_deg, d.pot_rows = deg, tab.shape[0]
            d.pot_coef = ptr(self.pot_table)
        except NonPolynomialPhysics:
            self.pot_table = None
            if self.time_dependent:
                # V0(t, x) and dV0/dx(t, x) as programs the device evaluates per step (the
                # reference splices both strings into its kernels, :1964-1966)
                self._pot_expr = compile_expression(V0, "x")
                self._dpot_expr = compile_expression(_sympy_derivative(V0), "x")
                d.pot_deg, d.pot_rows = WH_POT_EXPR, 1
                d.pot_expr, d.pot_expr_len = ptr(self._pot_expr), self._pot_expr.shape[0]
                d.dpot_expr, d.dpot_expr_len = ptr(self._dpot_expr), self._dpot_expr.shape[0]
                d.pot_t_step = self.dt
            else:
                u = 0.5 * ktheta[:, None]  # x -/+ theta/2 (no hbar: GPU_WignerDirac2D_4x4.py:408, 568)
                vm = host_eval(V0, t=0.0, x=kx[None, :] - u)
                vp = host_eval(V0, t=0.0, x=kx[None, :] + u)
                d.pot_deg, d.pot_rows = WH_POT_TABLE, 1
                d.pot_dv = ptr(np.broadcast_to(np.asarray(vm - vp, dtype=np.float64), (self.P, self.X)))
                dv0 = host_eval(_sympy_derivative(V0) + " + 0.*x", t=0.0, x=kx)
                d.dv0 = ptr(np.broadcast_to(np.asarray(dv0, dtype=np.float64), kx.shape))
                # V0 + A1 + A2 + A3 of the energy record; time-dependent A are added per step on the device
                vs = sum(np.broadcast_to(np.asarray(host_eval(v, t=0.0, x=kx), dtype=np.float64), kx.shape)
                         for v in (potentials[:1] if (a_time or self._a_expr is not None) else potentials))
                d.vsum = ptr(vs)
        k_D = tr(D_theta)
        d.theta_decay = ptr(np.exp(-k_dt * k_D * ktheta * ktheta)) if k_D != 0.0 else None
        if absorb:
            j = np.arange(self.X)
            d.x_mask = ptr(1.0 - np.exp(-(((j - self.X // 2) / 24.0) ** 2)))
        d.x_host, d.p_host, d.theta_host = ptr(self.x_host), ptr(self.p_host), ptr(self.theta_host)
        d.gamma_dt = self.dt * float(gamma) if (gamma > 0.0 and damping == "CaldeiraLegget") else 0.0
        d.dxdp = self.dX * self.dP
        d.c_host, d.mc2_host = self.c, self.mass * self.c * self.c
        d.transmission_x = float(transmission_x)
        d.antiparticle, d.energy = int(bool(antiparticle)), int(bool(energy))
        for k in range(3):
            d.a_deg[k] = a_deg[k]
            for n in range(WH_MAX_DEG + 1):
                d.a_coef[k][n] = float(a_coef[k][n])
        self.vector_potential = max(a_deg) >= 0
        d.gp_coeff = float(gp_coeff)  # passed as a kernel argument, not spliced (:3401, 3583)
        d.kvn, d.step_filter = int(bool(kvn)), int(bool(step_filter))
        self.kvn, self.step_filter = bool(kvn), bool(step_filter)
        if gamma > 0.0 and damping == "ODM":
            # DampingODM_Kernel (:952-1195): mass, c, dt and lambdaBar spliced ('%f', :1947),
            # gammaDamping a full-precision kernel argument (:3609-3615)
            d.odm, d.odm_gamma, d.lambda_bar = 1, float(gamma), tr(lambda_bar)
        d.lambda_host = ptr(host_axis(self.dLambda * self.X / 2.0, self.X, self.dLambda))
        self._plan = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().wh_dirac_create(ctypes.byref(d), ctypes.byref(self._plan)), "wh_dirac_create")
        del keep

    def close(self):
        if getattr(self, "_plan", None) is not None and self._plan.value:
            _lib.lib().wh_dirac_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def to_device(self, W):
        """[4, 4, P, X] (host or device) -> contiguous device [16, P, X] in the state precision."""
        import torch

        t = torch.as_tensor(np.ascontiguousarray(W) if isinstance(W, np.ndarray) else W)
        if tuple(t.shape) != (4, 4, self.P, self.X):
            raise ValueError(f"W must have shape (4, 4, {self.P}, {self.X}), got {tuple(t.shape)}")
        return t.to(self.device).to(self.dtype).reshape(16, self.P, self.X).contiguous()

    def _check_state(self, W):
        if W.dtype != self.dtype or not W.is_cuda or not W.is_contiguous() or tuple(W.shape) != (16, self.P, self.X):
            raise ValueError(f"state must be a contiguous {self.dtype} CUDA tensor of shape {(16, self.P, self.X)}")

    def run(self, W, n_steps: int, first_step: int = 1, use_graph: bool = True, obs=None):
        """Advances W (in place) by n_steps iterations with time indices first_step.. (t = dt * index);
        returns the [n_steps + 1, 32] record (device; row k = state before iteration first_step + k)."""
        import torch

        self._check_state(W)
        rows = [t.shape[0] for t in ((self.pot_table if self.time_dependent else None), self.a_table)
                if t is not None]
        if rows and first_step + n_steps + 1 >= min(rows):
            raise ValueError(f"time-dependent potential table has {min(rows)} rows; "
                             f"need {first_step + n_steps + 2} (pass time_rows)")
        if obs is None:
            obs = torch.zeros((n_steps + 1, WH_DNOBS), dtype=torch.float64, device=W.device)
        _lib.check(_lib.lib().wh_dirac_run(self._plan, _lib.dptr(W), int(first_step), int(n_steps), _lib.dptr(obs),
                                           int(bool(use_graph)), _lib.current_stream_handle()), "wh_dirac_run")
        return obs

    def kinetic(self, W):
        """W <- U(p + lambda/2) W U(p - lambda/2)^dagger between the x transforms (one sweep)."""
        self._check_state(W)
        _lib.check(_lib.lib().wh_dirac_kinetic(self._plan, _lib.dptr(W), _lib.current_stream_handle()),
                   "wh_dirac_kinetic")

    def potential(self, W, step: int = 1):
        """The potential factors at time index `step` (with D_Theta decay, absorber and damping)."""
        self._check_state(W)
        _lib.check(_lib.lib().wh_dirac_potential(self._plan, _lib.dptr(W), int(step), _lib.current_stream_handle()),
                   "wh_dirac_potential")

    @staticmethod
    def observables(obs):
        """Record rows (host array [n, 32]) -> dict of complex / real series."""
        o = np.asarray(obs)
        out = {k: o[:, i] + 1j * o[:, i + 1] for k, i in DIRAC_OBS.items()}
        out.update({k: o[:, i] for k, i in DIRAC_OBS_REAL.items()})
        return out
