"""Differentiable Wigner propagation: torch.autograd over the hand-written adjoint step.

    Wfinal_real = propagate(prop, coef_table, W0, n_steps, checkpoint_every=50)

`prop` is a Lie-scheme Wigner2D without damping; `coef_table` a float64 tensor
[rows, 9] of the coefficients a_k(t_n) of V(t_n, x) (row n used at time index n,
the last row for later steps); the result is Re W after n_steps (float64, [P, X]).
Gradients flow to coef_table (dL/da_k(n)) and to W0.

Forward: the fused run, keeping the state every `checkpoint_every` steps.
Backward: per segment, last to first, the forward states are recomputed from the
segment's checkpoint (two single-sweep launches per step) and the adjoint state is carried
back through the reverse-time step with conjugated phases (wh_wigner2d_adjoint_step), which also
accumulates dL/da_k(n).  Memory: n_steps/K checkpoints + K states.

The reference has no gradients; there is no reference counterpart to this module.
"""
from __future__ import annotations

import numpy as np

from ._lib import WH_MAX_DEG


def _torch():
    import torch

    return torch


class _WignerPropagate:
    """Namespace for the autograd.Function (built lazily so torch is imported on use)."""

    fn = None


def _make_function():
    torch = _torch()

    class WignerPropagate(torch.autograd.Function):
        @staticmethod
        def forward(ctx, coef_table, W0, prop, n_steps, K):
            prop.set_potential(coef_table.detach())
            W = W0.detach().to(prop.dtype).contiguous().clone()
            ckpts = [W.clone()]
            done = 0
            while done < n_steps:
                n = min(K, n_steps - done)
                prop.run(W, n, first_step=done, normalize_initial=False, use_graph=True)
                done += n
                if done < n_steps:
                    ckpts.append(W.clone())
            ctx.prop, ctx.n_steps, ctx.K = prop, n_steps, K
            ctx.rows = coef_table.shape[0]
            ctx.ckpts = ckpts
            ctx.coef_device = coef_table.device
            return W.real.to(torch.float64)

        @staticmethod
        def backward(ctx, grad_real):
            prop, n_steps, K = ctx.prop, ctx.n_steps, ctx.K
            Lam = grad_real.to(prop.device).to(prop.dtype).contiguous().clone()
            grads = torch.zeros((n_steps, WH_MAX_DEG + 1), dtype=torch.float64, device=prop.device)
            nseg = len(ctx.ckpts)
            for s in reversed(range(nseg)):
                n0 = s * K
                n = min(K, n_steps - n0)
                W = ctx.ckpts[s].clone()
                traj = []
                for k in range(n):
                    # one Lie step as two single sweeps (bit-identical to run(W, 1) without
                    # renormalisation, without its observable launches)
                    prop.kinetic(W)
                    prop.potential(W, n0 + k, fused=False)
                    traj.append(W.clone())
                for k in reversed(range(n)):
                    prop.adjoint_step(Lam, traj[k], n0 + k, grads[n0 + k])
                del traj
            # coefficient rows: row r receives the steps that read it (row = min(step, rows - 1))
            rows = ctx.rows
            dcoef = torch.zeros((rows, WH_MAX_DEG + 1), dtype=torch.float64, device=prop.device)
            m = min(rows - 1, n_steps)
            dcoef[:m] = grads[:m]
            if n_steps >= rows:
                dcoef[rows - 1] = grads[rows - 1:].sum(dim=0)
            elif m < rows:
                dcoef[m:n_steps] = grads[m:n_steps]
            return dcoef.to(ctx.coef_device), Lam, None, None, None

    return WignerPropagate


def propagate(prop, coef_table, W0, n_steps: int, checkpoint_every: int = 50):
    """Re W after n_steps Lie steps of `prop`, differentiable in coef_table and W0."""
    if prop.scheme != "lie" or prop.cl or getattr(prop, "renormalize", False):
        raise ValueError("propagate() needs a Lie-scheme Wigner2D without damping")
    if _WignerPropagate.fn is None:
        _WignerPropagate.fn = _make_function()
    return _WignerPropagate.fn.apply(coef_table, W0, prop, int(n_steps), int(checkpoint_every))


def fourier_control_table(params, n_steps: int, dt: float, T: float, base):
    """Coefficient rows of V(t,x) = base(x) - E(t) x with the control field
    E(t) = sum_j c_j cos(2 pi j t / T) + s_j sin(2 pi j t / T), j = 1..M,
    params = [c_1..c_M, s_1..s_M] (torch, differentiable); base = 9 constant coefficients."""
    torch = _torch()
    M = params.shape[0] // 2
    t = torch.arange(n_steps + 2, dtype=torch.float64, device=params.device) * dt
    j = torch.arange(1, M + 1, dtype=torch.float64, device=params.device)
    arg = 2.0 * np.pi * t[:, None] * j[None, :] / T
    E = torch.cos(arg) @ params[:M] + torch.sin(arg) @ params[M:]
    base = torch.as_tensor(base, dtype=torch.float64, device=params.device)
    tab = base.expand(n_steps + 2, WH_MAX_DEG + 1).clone()
    tab[:, 1] = tab[:, 1] - E
    return tab


def reference_propagate_torch(coef_table, W0, X, P, X_amp, P_amp, hbar, mass, dt, n_steps):
    """FP64 torch implementation of the same Lie steps (torch.fft, autograd): the
    independent gradient check of the hand-written adjoint.  Kinetic multiplier
    exp(-i dt lambda p / m) (GPU_Wigner2D_FFT.py:131-158), potential multiplier
    exp(-i dt (V(x - hbar theta/2) - V(x + hbar theta/2)) / hbar) (:80-126, no mask)."""
    torch = _torch()
    dev = coef_table.device
    k = lambda n, step: torch.tensor(((np.arange(n) + n // 2) % n - n / 2) * step, dtype=torch.float64, device=dev)  # noqa: E731
    dX, dP = 2 * X_amp / X, 2 * P_amp / P
    x, p = k(X, dX)[None, :], k(P, dP)[:, None]
    lam, th = k(X, np.pi / X_amp)[None, :], k(P, np.pi / P_amp)[:, None]
    Kmul = torch.exp(-1j * dt * lam * p / mass)
    xm, xp = x - hbar * th / 2, x + hbar * th / 2
    pw_m = torch.stack([xm ** q for q in range(WH_MAX_DEG + 1)])
    pw_p = torch.stack([xp ** q for q in range(WH_MAX_DEG + 1)])
    g = (pw_m - pw_p).to(torch.complex128)  # [9, P, X]
    W = W0.to(torch.complex128)
    rows = coef_table.shape[0]
    for n in range(n_steps):
        a = coef_table[min(n, rows - 1)].to(torch.complex128)
        phase = torch.einsum("k,kpx->px", a, g) * (dt / hbar)
        W = torch.fft.ifft(torch.fft.fft(W, dim=1) * Kmul, dim=1)
        W = torch.fft.ifft(torch.fft.fft(W, dim=0) * torch.exp(-1j * phase), dim=0)
    return W.real
