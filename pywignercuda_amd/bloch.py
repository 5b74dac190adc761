"""GPBloch: the imaginary-time (Bloch) Gross-Pitaevskii solver on one GPU through libwignerhip.

The layer under the drop-in GPU_Wigner2D_GPitaevskii_Bloch
(GPU_Wigner2D_GPitaevskii.py:2064-2543): kernel constants ('%f'-truncated as the reference's
source splicing does, GP.py:854-863), V(t, x) and K(p) as polynomial coefficient rows (or,
for a time-independent V that is not a polynomial, its Bopp sum V(x - hbar theta/2) +
V(x + hbar theta/2) tabulated at kernel coordinates), and the Hamiltonian grid K(P) + V(0, X)
on the host grids (GP.py:2094).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import BLOCH_HIST, WH_BNHIST, WH_MAX_DEG, WH_POT_TABLE
from .physics import NonPolynomialPhysics, coefficient_table, depends_on_time, heaviside_host, host_eval
from .wigner2d import fft_order_axis, host_axis, percent_f


class GPBloch:
    """One P x X grid.  potential / kinetic: the reference's C expression strings."""

    def __init__(self, X, P, X_amp, P_amp, hbar, mass, potential, kinetic, gp_coeff=0.0, dt=None,
                 time_rows=4096, truncate_constants=True, device=None):
        import torch

        self.X, self.P = int(X), int(P)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        tr = percent_f if truncate_constants else float
        self.dX = 2.0 * X_amp / self.X
        self.dP = 2.0 * P_amp / self.P
        self.dTheta = 2.0 * np.pi / (2.0 * P_amp)
        self.dLambda = 2.0 * np.pi / (2.0 * X_amp)
        self.x_host = host_axis(X_amp, self.X, self.dX)
        self.p_host = host_axis(P_amp, self.P, self.dP)
        self.theta_host = host_axis(self.dTheta * self.P / 2.0, self.P, self.dTheta)
        self.lambda_host = host_axis(self.dLambda * self.X / 2.0, self.X, self.dLambda)
        k_dx, k_dp = tr(self.dX), tr(self.dP)
        k_dtheta, k_dlambda = tr(self.dTheta), tr(self.dLambda)
        k_hbar, k_a = tr(hbar), tr(gp_coeff)
        self.hbar, self.mass, self.gp_coeff = float(hbar), float(mass), float(gp_coeff)
        self._keep = []

        def arr(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(a)
            return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

        d = _lib.BlochDesc()
        d.struct_size = ctypes.sizeof(d)
        d.x_dim, d.p_dim = self.X, self.P
        d.dx, d.dp, d.dtheta, d.dlambda, d.hbar = k_dx, k_dp, k_dtheta, k_dlambda, k_hbar
        # V(t, x): polynomial rows, one per iteration index when V depends on t (the reference
        # evaluates it at t = tIndex dt, GP.py:2180-2181), or the tabulated Bopp sum of a
        # time-independent V that is not a polynomial
        timedep = depends_on_time(potential, "x")
        if timedep and dt is None:
            raise ValueError("a time-dependent V needs dt (the time of iteration k is k dt)")
        try:
            times = np.arange(int(time_rows)) * float(dt) if timedep else [0.0]
            deg, tab = coefficient_table(potential, "x", times)
            d.pot_deg, d.pot_rows, d.pot_coef = deg, len(times), arr(tab)
        except NonPolynomialPhysics:
            if timedep:
                raise NonPolynomialPhysics(
                    f"{potential!r}: a time-dependent V must be a polynomial in x (degree <= {WH_MAX_DEG})")
            kx = fft_order_axis(self.X, k_dx)[None, :]
            u = 0.5 * k_hbar * fft_order_axis(self.P, k_dtheta)[:, None]
            vs = host_eval(potential, t=0.0, x=kx - u) + host_eval(potential, t=0.0, x=kx + u)
            d.pot_deg, d.pot_rows = WH_POT_TABLE, 1
            d.pot_sum = arr(np.broadcast_to(np.asarray(vs, dtype=np.float64), (self.P, self.X)))
        kdeg, ktab = coefficient_table(kinetic, "p", [0.0])
        d.kinetic_deg = kdeg
        for n in range(WH_MAX_DEG + 1):
            d.kin_coef[n] = ktab[0, n]
        d.gp_coeff = k_a
        d.theta_host, d.lambda_host = arr(self.theta_host), arr(self.lambda_host)
        d.x_host, d.p_host = arr(self.x_host), arr(self.p_host)
        # the Hamiltonian grid of the reference: KineticEnergy(P) + Potential(0, X) on the host
        # (its host-side Heaviside is 1/2 at 0, GP.py:675-697)
        kh = np.broadcast_to(np.asarray(host_eval(kinetic, p=self.p_host), dtype=np.float64), self.p_host.shape)
        vh = np.broadcast_to(np.asarray(host_eval(potential, heaviside=heaviside_host, t=0.0, x=self.x_host),
                                        dtype=np.float64), self.x_host.shape)
        self.k_host, self.v0_host = kh, vh
        d.k_host, d.v0_host = arr(kh), arr(vh)
        d.dxdp, d.dx_host, d.dp_host, d.hbar_host, d.gp_coeff_host = (
            self.dX * self.dP, self.dX, self.dP, self.hbar, self.gp_coeff)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().wh_bloch_create(ctypes.byref(d), ctypes.byref(h)), "wh_bloch_create")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().wh_bloch_destroy(h)
            except Exception:
                pass

    def to_device(self, W):
        import torch

        return torch.tensor(np.ascontiguousarray(W, dtype=np.complex128), device=self.device)

    def run(self, W, n_iter, dt, W_step=None, W_ground=None):
        """n_iter Bloch iterations from dt (Run / Run_ExitedState1, GP.py:2105-2535) on the device
        tensor W (in place).  Returns (history [n_iter, WH_BNHIST] device tensor, W_step)."""
        import torch

        if W_step is None:
            W_step = torch.empty_like(W)
        hist = torch.zeros((max(int(n_iter), 1), WH_BNHIST), dtype=torch.float64, device=W.device)
        g = _lib.dptr(W_ground) if W_ground is not None else None
        _lib.check(_lib.lib().wh_bloch_run(self._h, _lib.dptr(W), _lib.dptr(W_step), g, int(n_iter), float(dt),
                                           _lib.dptr(hist), _lib.current_stream_handle()), "wh_bloch_run")
        return hist[: int(n_iter)], W_step

    @staticmethod
    def history(hist) -> dict:
        h = hist.cpu().numpy() if hasattr(hist, "cpu") else np.asarray(hist)
        return {k: h[:, i] for k, i in BLOCH_HIST.items()}
