"""Wigner2D and WignerEnsemble: Wigner functions propagated by libwignerhip on one GPU.

This is the layer under the drop-in classes (GPU_Wigner2D_FFT,
GPU_Wigner2D_GPitaevskii).  It turns the reference's physical set-up into the
C descriptors: kernel constants (optionally '%f'-truncated as the reference's
source splicing does, GPU_Wigner2D_FFT.py:833-842), polynomial coefficient
tables for V(t,x) and K(p), and the 1-D host tables (decays, absorbing mask,
friction sigma(p), host grids and dV/dx for the observables).

State tensors are torch complex tensors on the GPU, [P][X] row-major,
FFT ordered, exactly like the reference's W_GPU.
"""
from __future__ import annotations

import ctypes
import json
import warnings
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import (WH_C64, WH_C64_WIDE, WH_C128, WH_MAX_DEG, WH_NOBS, WH_POT_EXPR, WH_POT_TABLE, WH_SCHEME_LIE,
                   WH_SCHEME_STRANG)

# state precision names: complex128; complex64 state with FP64 arithmetic inside the sweeps
# (one FP32 rounding of the state per sweep); complex64 state with FP32 arithmetic
PRECISIONS = {"fp64": WH_C128, "fp32": WH_C64_WIDE, "fp32fast": WH_C64}
from .physics import (NonPolynomialPhysics, coefficient_table, compile_expression, depends_on_time, heaviside_host,
                      host_eval)

TUNING_FILE = Path(__file__).resolve().parent / "tuning.json"


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


def _load_tuning():
    try:
        return json.loads(TUNING_FILE.read_text())
    except (OSError, ValueError):
        return {}


def tuned_variants(precision: str, sizes_x, sizes_p, cl: bool, batch: bool = False):
    """Variant indices (x, p, p+CL) chosen by pywignercuda_amd/tune.py for the largest grid.

    batch: the plan holds enough cells to fill the GPU (the tuner's "batch" regime)."""
    tab = _load_tuning()
    t = tab.get(precision + "_batch", {}) if batch else {}
    if not t:
        t = tab.get(precision, {})
    nx, npp = str(max(sizes_x)), str(max(sizes_p))
    xv = t.get("x", {}).get(nx, -1)
    pv = t.get("p", {}).get(npp, -1)
    pcl = t.get("pcl", {}).get(npp, -1)
    return int(xv), int(pv), int(pcl)


class _Tables:
    """Host-side set-up of one member (kept alive until the library copied it)."""

    def __init__(self, X, P, X_amp, P_amp, hbar, mass, dt, *, potential, dpotential, scheme,
                 D_theta, D_lambda, gamma, damping, fp_damping, gp_coeff, gp_rho_factor, mask,
                 truncate_constants, time_rows, pot_table, L_material=None, odm_sign=1.0, kvn=False):
        self.X, self.P = int(X), int(P)
        self.X_amp, self.P_amp = float(X_amp), float(P_amp)
        tr = percent_f if truncate_constants else float
        # host grids (SetPhaseSpaceBox2D / SetPhaseSpaceGrid2D, GPU_Wigner2D_FFT.py:549-596)
        self.dX = 2.0 * self.X_amp / self.X
        self.dP = 2.0 * self.P_amp / self.P
        self.dTheta = 2.0 * np.pi / (2.0 * self.P_amp)
        self.dLambda = 2.0 * np.pi / (2.0 * self.X_amp)
        self.x_host = host_axis(self.X_amp, self.X, self.dX)
        self.p_host = host_axis(self.P_amp, self.P, self.dP)
        self.theta_host = host_axis(self.dTheta * self.P / 2.0, self.P, self.dTheta)
        self.lambda_host = host_axis(self.dLambda * self.X / 2.0, self.X, self.dLambda)
        # kernel constants
        self.k_dt, self.k_dx, self.k_dp = tr(dt), tr(self.dX), tr(self.dP)
        self.k_dtheta, self.k_dlambda = tr(self.dTheta), tr(self.dLambda)
        self.k_mass, self.k_hbar = tr(mass), tr(hbar)
        k_Dtheta, k_Dlambda = tr(D_theta), tr(D_lambda)
        self.k_agp = tr(gp_coeff)
        kx = fft_order_axis(self.X, self.k_dx)
        kp = fft_order_axis(self.P, self.k_dp)
        ktheta = fft_order_axis(self.P, self.k_dtheta)
        klambda = fft_order_axis(self.X, self.k_dlambda)
        frac = 1.0 if scheme == "lie" else 0.5
        self.theta_decay = np.exp(-frac * k_Dtheta * ktheta * ktheta * self.k_dt) if k_Dtheta != 0.0 else None
        self.lambda_decay = np.exp(-k_Dlambda * klambda * klambda) if k_Dlambda != 0.0 else None
        xm = self.k_dx * (self.X - 1.0) / 2.0

        def edge(w):
            w2 = (w * self.k_dx) ** 2
            return (1.0 - np.exp(-((kx - xm) ** 2) / w2)) * (1.0 - np.exp(-((kx + xm) ** 2) / w2))

        self.pot_dv = None
        self.pot_expr = None
        self.pot_expr_kvn = 0
        self.pot_t_step = float(dt)
        self.x_mask = None
        if mask:
            self.x_mask = edge(10.0)  # expPotential, GPU_Wigner2D_FFT.py:121-123
            if gp_coeff != 0.0 and scheme == "lie":
                self.x_mask = self.x_mask * edge(20.0)  # expPotential_GrossPitaevskii, :514-516
        # potential coefficient rows
        if kvn:
            # classical KvN (Fokker-Planck) phase -dt dV(x) theta (exp_PotentialKvN, FFT.py:37-64), as the
            # table read by the p sweep: phase = dt/hbar * pot_dv, pot_dv = -hbar theta dV(x)
            if scheme != "lie" or pot_table is not None or not dpotential:
                raise ValueError("the KvN phase belongs to the Lie scheme and needs dpotential")
            self.time_dependent = depends_on_time(dpotential, "x")
            if self.time_dependent:
                # dPotential(t, x) as a program: the device tabulates -hbar theta dV/dx(t, x) per step
                self.pot_expr, self.pot_expr_kvn = compile_expression(dpotential, "x"), 1
                deg = WH_POT_EXPR
            else:
                dvk = np.broadcast_to(np.asarray(host_eval(dpotential, t=0.0, x=kx), dtype=np.float64), kx.shape)
                self.pot_dv = np.ascontiguousarray(-self.k_hbar * ktheta[:, None] * dvk[None, :])
                deg = WH_POT_TABLE
            tab = np.zeros((1, WH_MAX_DEG + 1))
        elif pot_table is not None:
            # explicit device coefficients: the degree is the highest non-zero column
            tab = np.ascontiguousarray(pot_table, dtype=np.float64)
            nz = np.nonzero(np.any(tab != 0, axis=0))[0]
            deg = int(nz.max()) if nz.size else 0
            self.time_dependent = tab.shape[0] > 1
        else:
            self.time_dependent = depends_on_time(potential, "x")
            times = np.arange(int(time_rows) if time_rows else 4096) * dt if self.time_dependent else [0.0]
            try:
                deg, tab = coefficient_table(potential, "x", times)
            except NonPolynomialPhysics:
                if self.time_dependent:
                    # arbitrary V(t, x): a program evaluated per cell on the device at t = step * dt, as
                    # expPotential evaluates the spliced string (FFT.py:88-94); NonPolynomialPhysics for
                    # syntax the device evaluator does not have
                    self.pot_expr = compile_expression(potential, "x")
                    deg = WH_POT_EXPR
                else:
                    # arbitrary time-independent V: the Bopp difference V(x - hbar theta/2) - V(x + hbar theta/2)
                    # at kernel coordinates, as expPotential evaluates it per cell (FFT.py:88-94)
                    u = 0.5 * self.k_hbar * ktheta[:, None]
                    vm = host_eval(potential, t=0.0, x=kx[None, :] - u)
                    vp = host_eval(potential, t=0.0, x=kx[None, :] + u)
                    self.pot_dv = np.ascontiguousarray(
                        np.broadcast_to(np.asarray(vm - vp, dtype=np.float64), (self.P, self.X)))
                    deg = WH_POT_TABLE
                tab = np.zeros((1, WH_MAX_DEG + 1))
        self.pot_deg, self.pot_table = deg, np.ascontiguousarray(tab)
        # a time-dependent V read from coefficient rows limits the run to the rows present
        self.time_table = self.time_dependent and self.pot_expr is None
        # Caldeira-Leggett friction sigma(p) = sign(p) f(p) at kernel p (theta_fp, FFT.py:424-429)
        self.sigma = None
        if damping == "CaldeiraLeggett" and gamma != 0.0:
            f = np.broadcast_to(np.asarray(host_eval(fp_damping, p=kp), dtype=np.float64), kp.shape)
            self.sigma = np.where(kp >= 0.0, f, -f)
        # ODM tables (SetA_ODM, GPU_Wigner2D_FFT.py:1196-1244) on host grids, full precision
        self.odm_g = self.odm_x = self.odm_th = None
        self.odm_sign = float(odm_sign)
        if damping == "ODM" and gamma != 0.0:
            if L_material is None or fp_damping is None:
                raise ValueError("ODM damping needs L_material and fp_damping")
            h, L = float(hbar), float(L_material)

            def gfun(q):
                f = np.broadcast_to(np.asarray(host_eval(fp_damping, p=np.abs(q) + h / (2.0 * L)),
                                               dtype=np.float64), q.shape)
                return np.sqrt(L * f)

            lam = self.lambda_host[None, :]
            qp = self.p_host[:, None] + 0.5 * h * lam
            qm = self.p_host[:, None] - 0.5 * h * lam
            gp_, gm_ = gfun(qp), gfun(qm)
            self.odm_g = np.ascontiguousarray(np.stack([gp_, gm_, gp_ * np.sign(qp), gm_ * np.sign(qm)]))
            self.odm_x = np.ascontiguousarray(np.stack([np.cos(2.0 * self.x_host), np.sin(2.0 * self.x_host)]))
            ht = h * self.theta_host
            self.odm_th = np.ascontiguousarray(np.stack([np.cos(ht), np.sin(ht)]))
        # observables (FFT.py:1747, 1785-1792): host grids, dV/dx and V at t = 0
        if dpotential:
            dv = host_eval(dpotential, t=0.0, x=self.x_host)
        else:
            dv = 0.0 * self.x_host
        self.dv = np.ascontiguousarray(np.broadcast_to(np.asarray(dv, dtype=np.float64), self.x_host.shape))
        # the host-side Potential of the reference (its Heaviside is 1/2 at 0: GP.py:675-697)
        v0 = host_eval(potential, heaviside=heaviside_host, t=0.0, x=self.x_host)
        self.v0 = np.ascontiguousarray(np.broadcast_to(np.asarray(v0, dtype=np.float64), self.x_host.shape))
        self.dxdp = self.dX * self.dP
        self.inv_2m = 1.0 / (2.0 * float(mass))
        self.rho_factor = float(gp_rho_factor) if gp_rho_factor is not None else (1.0 if scheme == "lie" else self.dP)
        self._keep = []

    def ptr(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._keep.append(a)
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def fill(self, d):
        """Fills the per-member fields of a Wigner2DDesc or MemberDesc."""
        d.x_dim, d.p_dim = self.X, self.P
        d.dx, d.dp, d.dtheta, d.dlambda = self.k_dx, self.k_dp, self.k_dtheta, self.k_dlambda
        d.pot_deg, d.pot_rows = self.pot_deg, self.pot_table.shape[0]
        d.pot_coef = self.ptr(self.pot_table)
        d.pot_dv = self.ptr(self.pot_dv)
        d.pot_expr = self.ptr(self.pot_expr)
        d.pot_expr_len = 0 if self.pot_expr is None else int(self.pot_expr.shape[0])
        d.pot_expr_kvn = self.pot_expr_kvn
        d.pot_t_step = self.pot_t_step
        d.odm_g, d.odm_x, d.odm_th = self.ptr(self.odm_g), self.ptr(self.odm_x), self.ptr(self.odm_th)
        d.odm_sign = self.odm_sign
        d.theta_decay = self.ptr(self.theta_decay)
        d.lambda_decay = self.ptr(self.lambda_decay)
        d.x_mask = self.ptr(self.x_mask)
        d.sigma_p = self.ptr(self.sigma)
        d.theta_host = self.ptr(self.theta_host)
        d.lambda_host = self.ptr(self.lambda_host)
        d.x_host, d.p_host = self.ptr(self.x_host), self.ptr(self.p_host)
        d.dv_host, d.v0_host = self.ptr(self.dv), self.ptr(self.v0)
        d.dxdp, d.inv_2m, d.dx_host = self.dxdp, self.inv_2m, self.dX
        d.gp_rho_factor = self.rho_factor


def _common(d, *, precision, scheme, t: _Tables, dt, gamma, damping, kinetic, gp_coeff, odm_weights=None,
            renormalize=None):
    d.precision = PRECISIONS[precision]
    d.scheme = WH_SCHEME_LIE if scheme == "lie" else WH_SCHEME_STRANG
    d.dt, d.hbar, d.mass = t.k_dt, t.k_hbar, t.k_mass
    kin_poly, kdeg, kcoef = 0, 0, np.zeros(WH_MAX_DEG + 1)
    if scheme != "lie":
        kin_poly = 1
        if kinetic is not None:
            kdeg, ktab = coefficient_table(kinetic, "p", [0.0])
            kcoef = ktab[0]
        else:
            kdeg, kcoef[2] = 2, 0.5 / t.k_mass
    d.kinetic_poly, d.kinetic_deg = kin_poly, kdeg
    for n in range(WH_MAX_DEG + 1):
        d.kin_coef[n] = float(kcoef[n])
    cl = damping == "CaldeiraLeggett" and gamma != 0.0
    d.damping = 1 if cl else (2 if (damping == "ODM" and gamma != 0.0) else 0)
    d.gamma_dt = float(dt) * float(gamma)  # host constants (FFT.py:1150-1156)
    d.renormalize = int(gamma != 0.0 if renormalize is None else bool(renormalize))
    d.gp_coeff = t.k_agp
    # ODM channels: one with weight 1 (GPU_Wigner2D_FFT.Run) or the two of GPU_Wigner2D_ODM2
    w = (1.0,) if odm_weights is None else tuple(float(v) for v in odm_weights)
    d.odm_channels = len(w)
    d.odm_weight1 = w[0]
    d.odm_weight2 = w[1] if len(w) > 1 else 0.0


class _PlanBase:
    def _set_variants(self, ens_handle, sizes_x, sizes_p, variants):
        cells = sum(int(x) * int(p) for x, p in zip(sizes_x, sizes_p))
        batch = cells * (16 if self.precision == "fp64" else 8) >= (256 << 20)
        if variants is not None:
            xv, pv, pcl = variants
        else:
            xv, pv, pcl = tuned_variants(self.precision, sizes_x, sizes_p, self.cl, batch)
            if self.gp:
                xv = -1  # the x variant of a Gross-Pitaevskii plan is fixed at creation
        if (xv, pv, pcl) != (-1, -1, -1):
            rc = _lib.lib().wh_ensemble_set_variants(ens_handle, xv, pv, pcl)
            if rc != 0:
                if variants is not None:
                    _lib.check(rc, "wh_ensemble_set_variants")
                # a tuned choice that does not fit this plan: the defaults stay, say so
                warnings.warn(f"tuned kernel variants {(xv, pv, pcl)} not applied: "
                              f"{_lib.last_error()}", RuntimeWarning, stacklevel=3)

    def to_device(self, W, shape=None):
        import torch

        t = torch.as_tensor(np.ascontiguousarray(W)).to(self.dtype)
        if shape is not None and tuple(t.shape) != tuple(shape):
            raise ValueError(f"W must have shape {shape}, got {tuple(t.shape)}")
        return t.to(self.device).contiguous()


class Wigner2D(_PlanBase):
    """Fused split-operator propagator for one Wigner function.

    Parameters follow GPU_Wigner2D_FFT.__init__ (GPU_Wigner2D_FFT.py:1718-1720)
    and the attributes its Run() reads:

    potential, dpotential, kinetic, fp_damping : reference-style C expression strings
        (V(t,x), dV/dx(t,x), K(p), f(p)); V and K must be polynomials of degree
        <= 8 in x / p (coefficients may depend on t).
    scheme      : 'lie' (K; V -- GPU_Wigner2D_FFT) or 'strang' (V/2; K; V/2 --
                  GPU_Wigner2D_GPitaevskii).
    kinetic     : ignored by the Lie scheme (the reference hard-wires dt*lambda*p/m).
    damping     : None, 'CaldeiraLeggett' (order-3 composite) or 'ODM' (Lindblad dissipator of
                  order 1, after K; V in the Lie scheme and after V/2; K; V/2 in the Strang
                  scheme, GP.py:1958-1968; needs L_material and fp_damping; odm_sign -1 pumps).
    odm_weights : ODM channel weights: None = one channel of weight 1 (GPU_Wigner2D_FFT);
                  (w1, w2) = GPU_Wigner2D_ODM2 (W += L(W) with (w1, +1), then (w2, -1)).
    gp_coeff    : Gross-Pitaevskii coefficient a_GP (0 = off).
    mask        : absorbing edge mask of the potential kernel (on in the reference).
    truncate_constants : emulate the '%f' splicing of kernel constants.
    precision   : 'fp64' (complex128 state), 'fp32' (complex64 state, FP64 arithmetic in the
                  sweeps: half the HBM bytes, one FP32 rounding per sweep) or 'fp32fast'
                  (complex64 state and FP32 arithmetic with FP64 phases, compensated twiddles
                  and FP64 observable sums).
    time_rows   : rows of the coefficient table for a time-dependent V (>= steps + 2).
    pot_table   : explicit [rows, 9] coefficient table for the device (the string
                  is still used for the host observables).
    kvn         : classical Koopman-von Neumann (Fokker-Planck) potential step
                  exp(i dt dV(t, x) theta) in place of the Bopp difference of V
                  (GPU_FokkerPlank2D_FFT, FFT.py:27-75, 2242-2459); Lie scheme.
    renormalize : divide by dXdP sum(W) after each step; None = when gamma != 0.
    variants    : (x, p, p+CL) kernel variant indices; None = tuned choice.
    """

    def __init__(self, X, P, X_amp, P_amp, hbar, mass, dt, *, potential, dpotential=None,
                 kinetic=None, scheme="lie", D_theta=0.0, D_lambda=0.0, gamma=0.0,
                 damping=None, fp_damping=None, gp_coeff=0.0, gp_rho_factor=None, mask=True,
                 truncate_constants=False, precision="fp64", time_rows=None, pot_table=None,
                 device=None, variants=None, L_material=None, odm_sign=1.0, odm_weights=None,
                 kvn=False, renormalize=None):
        import torch

        if odm_weights is not None and not 1 <= len(odm_weights) <= 2:
            raise ValueError("odm_weights: one or two Lindblad channel weights")
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.precision = precision
        self.dtype = torch.complex128 if precision == "fp64" else torch.complex64
        self.device = torch.device(device if device is not None else "cuda:0")
        self.scheme = scheme
        self.cl = damping == "CaldeiraLeggett" and gamma != 0.0
        self.gp = gp_coeff != 0.0
        t = _Tables(X, P, X_amp, P_amp, hbar, mass, dt, potential=potential, dpotential=dpotential,
                    scheme=scheme, D_theta=D_theta, D_lambda=D_lambda, gamma=gamma, damping=damping,
                    fp_damping=fp_damping, gp_coeff=gp_coeff, gp_rho_factor=gp_rho_factor, mask=mask,
                    truncate_constants=truncate_constants, time_rows=time_rows, pot_table=pot_table,
                    L_material=L_material, odm_sign=odm_sign, kvn=kvn)
        self.tables = t
        self.X, self.P = t.X, t.P
        self.dX, self.dP = t.dX, t.dP
        self.x_host, self.p_host = t.x_host, t.p_host
        self.dt = float(dt)
        self.time_dependent = t.time_dependent
        self.time_table = t.time_table
        self.pot_table = t.pot_table
        self.renormalize = gamma != 0.0 if renormalize is None else bool(renormalize)
        d = _lib.Wigner2DDesc()
        d.struct_size = ctypes.sizeof(_lib.Wigner2DDesc)
        _common(d, precision=precision, scheme=scheme, t=t, dt=dt, gamma=gamma, damping=damping,
                kinetic=kinetic, gp_coeff=gp_coeff, odm_weights=odm_weights, renormalize=self.renormalize)
        t.fill(d)
        self._plan = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().wh_wigner2d_create(ctypes.byref(d), ctypes.byref(self._plan)),
                       "wh_wigner2d_create")
        t._keep = []
        self._ens = ctypes.c_void_p(_lib.lib().wh_wigner2d_ensemble(self._plan))
        self._set_variants(self._ens, [self.X], [self.P], variants)

    # ------------------------------------------------------------------ lifetime
    def close(self):
        if getattr(self, "_plan", None) is not None and self._plan.value:
            _lib.lib().wh_wigner2d_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def to_device(self, W):
        return super().to_device(W, (self.P, self.X))

    def _check_state(self, W):
        if W.dtype != self.dtype or not W.is_cuda or not W.is_contiguous() or tuple(W.shape) != (self.P, self.X):
            raise ValueError(f"state must be a contiguous {self.dtype} CUDA tensor of shape {(self.P, self.X)}")

    def set_variants(self, xvar=-1, pvar=-1, pclvar=-1):
        _lib.check(_lib.lib().wh_ensemble_set_variants(self._ens, int(xvar), int(pvar), int(pclvar)),
                   "wh_ensemble_set_variants")

    # ------------------------------------------------------------------ stepping
    def run(self, W, n_steps: int, first_step: int = 0, normalize_initial: bool = True,
            use_graph: bool = True, obs=None):
        """Advances W in place by n_steps iterations; returns the [n_steps+1, 16] record (device)."""
        import torch

        self._check_state(W)
        if self.time_table and first_step + n_steps + 1 >= self.pot_table.shape[0]:
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

    def set_potential(self, coef_table):
        """Replaces the device coefficient rows of V(t,x) ([rows, 9] float64, row = time index)."""
        import torch

        tab = torch.as_tensor(coef_table, dtype=torch.float64).to(self.device).contiguous()
        if tab.dim() != 2 or tab.shape[1] != WH_MAX_DEG + 1:
            raise ValueError(f"coefficient table must be [rows, {WH_MAX_DEG + 1}]")
        nz = torch.nonzero(torch.any(tab != 0, dim=0)).flatten()
        deg = int(nz.max()) if nz.numel() else 0
        _lib.check(_lib.lib().wh_ensemble_set_potential(self._ens, 0, _lib.dptr(tab), int(tab.shape[0]), deg,
                                                        _lib.current_stream_handle()),
                   "wh_ensemble_set_potential")
        self._pot_device = tab  # the copy is asynchronous: keep the source alive
        self.pot_table = tab.cpu().numpy()
        self.time_dependent = self.time_table = tab.shape[0] > 1

    def adjoint_step(self, Lam, W_next, step: int, grad_row):
        """Lambda <- U_step^dagger Lambda; grad_row (device, 9 doubles) <- dL/da_k(step)."""
        self._check_state(Lam)
        self._check_state(W_next)
        _lib.check(_lib.lib().wh_wigner2d_adjoint_step(self._plan, _lib.dptr(Lam), _lib.dptr(W_next), int(step),
                                                       _lib.dptr(grad_row), _lib.current_stream_handle()),
                   "wh_wigner2d_adjoint_step")

    def observe(self, W):
        import torch

        self._check_state(W)
        obs = torch.zeros((1, WH_NOBS), dtype=torch.float64, device=W.device)
        _lib.check(_lib.lib().wh_wigner2d_observe(self._plan, _lib.dptr(W), _lib.dptr(obs),
                                                  _lib.current_stream_handle()),
                   "wh_wigner2d_observe")
        return obs[0]


class WignerEnsemble(_PlanBase):
    """Many Wigner functions (uneven grid sizes and potentials) propagated together.

    members : list of dicts with the per-member keys X, P, X_amp, P_amp, potential,
              and optionally dpotential, pot_table, D_theta, D_lambda, fp_damping, mask,
              L_material and odm_sign (damping 'ODM').
    The remaining keyword arguments (odm_weights included) are shared by all members, as
    for Wigner2D.
    All sweeps of all members are one launch each; the step loop is graph-captured.
    """

    def __init__(self, members, hbar, mass, dt, *, scheme="lie", gamma=0.0, damping=None,
                 kinetic=None, gp_coeff=0.0, truncate_constants=False, precision="fp64",
                 time_rows=None, device=None, variants=None, odm_weights=None):
        import torch

        if odm_weights is not None and not 1 <= len(odm_weights) <= 2:
            raise ValueError("odm_weights: one or two Lindblad channel weights")
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.precision = precision
        self.dtype = torch.complex128 if precision == "fp64" else torch.complex64
        self.device = torch.device(device if device is not None else "cuda:0")
        self.scheme = scheme
        self.cl = damping == "CaldeiraLeggett" and gamma != 0.0
        self.gp = gp_coeff != 0.0
        self.dt = float(dt)
        self.tables = []
        for m in members:
            self.tables.append(_Tables(
                m["X"], m["P"], m["X_amp"], m["P_amp"], hbar, mass, dt, potential=m["potential"],
                dpotential=m.get("dpotential"), scheme=scheme, D_theta=m.get("D_theta", 0.0),
                D_lambda=m.get("D_lambda", 0.0), gamma=gamma, damping=damping,
                fp_damping=m.get("fp_damping"), gp_coeff=gp_coeff, gp_rho_factor=m.get("gp_rho_factor"),
                mask=m.get("mask", False), truncate_constants=truncate_constants, time_rows=time_rows,
                pot_table=m.get("pot_table"), L_material=m.get("L_material"),
                odm_sign=m.get("odm_sign", 1.0)))
        self.n = len(self.tables)
        self.shapes = [(t.P, t.X) for t in self.tables]
        self.time_dependent = any(t.time_dependent for t in self.tables)
        self.time_table = any(t.time_table for t in self.tables)
        self.time_rows = min((t.pot_table.shape[0] for t in self.tables if t.time_table), default=1)
        d = _lib.EnsembleDesc()
        d.struct_size = ctypes.sizeof(_lib.EnsembleDesc)
        _common(d, precision=precision, scheme=scheme, t=self.tables[0], dt=dt, gamma=gamma,
                damping=damping, kinetic=kinetic, gp_coeff=gp_coeff,
                odm_weights=odm_weights if damping == "ODM" else None)
        arr = (_lib.MemberDesc * self.n)()
        for i, t in enumerate(self.tables):
            arr[i].struct_size = ctypes.sizeof(_lib.MemberDesc)
            t.fill(arr[i])
        self._ens = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().wh_ensemble_create(ctypes.byref(d), arr, self.n, ctypes.byref(self._ens)),
                       "wh_ensemble_create")
        for t in self.tables:
            t._keep = []
        self._set_variants(self._ens, [s[1] for s in self.shapes], [s[0] for s in self.shapes], variants)

    def close(self):
        if getattr(self, "_ens", None) is not None and self._ens.value:
            _lib.lib().wh_ensemble_destroy(self._ens)
            self._ens = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_variants(self, xvar=-1, pvar=-1, pclvar=-1):
        _lib.check(_lib.lib().wh_ensemble_set_variants(self._ens, int(xvar), int(pvar), int(pclvar)),
                   "wh_ensemble_set_variants")

    def states_to_device(self, Ws):
        return [super(WignerEnsemble, self).to_device(W, s) for W, s in zip(Ws, self.shapes)]

    def run(self, states, n_steps: int, first_step: int = 0, normalize_initial: bool = True,
            use_graph: bool = True, obs=None):
        """Advances every member in place; returns the [n, n_steps+1, 16] record (device)."""
        import torch

        if len(states) != self.n:
            raise ValueError(f"expected {self.n} states")
        for W, s in zip(states, self.shapes):
            if W.dtype != self.dtype or not W.is_cuda or not W.is_contiguous() or tuple(W.shape) != s:
                raise ValueError(f"states must be contiguous {self.dtype} CUDA tensors of shapes {self.shapes}")
        if self.time_table and first_step + n_steps + 1 >= self.time_rows:
            raise ValueError("time-dependent potential table too short (pass time_rows)")
        if obs is None:
            obs = torch.zeros((self.n, n_steps + 1, WH_NOBS), dtype=torch.float64, device=self.device)
        ptrs = (ctypes.c_void_p * self.n)(*[int(W.data_ptr()) for W in states])
        _lib.check(_lib.lib().wh_ensemble_run(self._ens, ptrs, int(first_step), int(n_steps), _lib.dptr(obs),
                                              int(bool(normalize_initial)), int(bool(use_graph)),
                                              _lib.current_stream_handle()),
                   "wh_ensemble_run")
        return obs
