"""GPU: the hand-written adjoint step (reverse-time split operator with conjugated
phases, dL/dV coefficients accumulated in the p sweep) behind torch.autograd.

Checked against (1) autograd of an independent FP64 torch implementation of the
same Lie steps (CPU, torch.fft) and (2) central finite differences (h = 1e-6,
FP64) at the differentiable-control configuration: 512^2, V = 0.25x^4 - x^2 - E(t)x,
E(t) a 64-term Fourier series over T = 20 with N(0, 0.1) coefficients, dt = 0.01,
2000 steps, loss 1 - 2 pi hbar int W_T W_target (target Gaussian at x = +2),
checkpoints every 50 steps.  Bar: relative agreement 1e-6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(cuda, X, n_steps, n_params, seed, X_amp=10.0):
    import torch

    from pywignercuda_amd.adjoint import fourier_control_table
    from pywignercuda_amd.wigner2d import Wigner2D

    dt, T = 0.01, 20.0
    base = np.zeros(9)
    base[4], base[2] = 0.25, -1.0
    prop = Wigner2D(X, X, X_amp, X_amp, 1.0, 1.0, dt, potential="0.25*pow(x,4) - pow(x,2)",
                    dpotential="pow(x,3) - 2*x", mask=False, device=cuda,
                    pot_table=np.tile(base, (n_steps + 2, 1)))
    rng = np.random.default_rng(seed)
    params = torch.tensor(rng.normal(0.0, 0.1, n_params), dtype=torch.float64, device=cuda, requires_grad=True)
    x, p = prop.x_host[None, :], prop.p_host[:, None]
    W0 = np.exp(-((x + 1.0) ** 2) - p ** 2) / np.pi
    W0 = W0 / (W0.sum() * prop.dX * prop.dP)
    Wt = np.exp(-((x - 2.0) ** 2) - p ** 2) / np.pi
    Wt = torch.tensor(Wt, dtype=torch.float64, device=cuda)
    c = 2 * np.pi * 1.0 * prop.dX * prop.dP

    def table(par):
        return fourier_control_table(par, n_steps, dt, T, base)

    def loss_of(Wreal):
        return 1.0 - c * (Wreal * Wt).sum()

    return prop, params, prop.to_device(W0), table, loss_of


def test_adjoint_matches_torch_autograd_reference(cuda):
    import torch

    from pywignercuda_amd.adjoint import propagate, reference_propagate_torch

    X, n_steps = 128, 120
    prop, params, W0, table, loss_of = _setup(cuda, X, n_steps, 8, 5, X_amp=8.0)
    tab = table(params.detach()).clone().requires_grad_(True)
    L = loss_of(propagate(prop, tab, W0, n_steps, checkpoint_every=25))
    (g_hip,) = torch.autograd.grad(L, tab)
    tab_cpu = tab.detach().cpu().clone().requires_grad_(True)
    Wr = reference_propagate_torch(tab_cpu, W0.cpu(), X, X, 8.0, 8.0, 1.0, 1.0, 0.01, n_steps)
    Wt = torch.tensor(np.exp(-((prop.x_host[None, :] - 2.0) ** 2) - prop.p_host[:, None] ** 2) / np.pi)
    Lr = 1.0 - 2 * np.pi * prop.dX * prop.dP * (Wr * Wt).sum()
    assert abs(float(L) - float(Lr)) < 1e-11
    (g_ref,) = torch.autograd.grad(Lr, tab_cpu)
    g_hip = g_hip.cpu()
    scale = g_ref.abs().max()
    err = (g_hip - g_ref).abs().max() / scale
    print(f"adjoint vs torch autograd: max rel err {float(err):.3e} (scale {float(scale):.3e})")
    assert err < 1e-8


def test_adjoint_matches_finite_differences_control_config(cuda):
    """512^2, 2000 steps, 64 Fourier coefficients (all checked), checkpoints every 50 steps."""
    import torch

    from pywignercuda_amd.adjoint import propagate

    n_steps = 2000
    prop, params, W0, table, loss_of = _setup(cuda, 512, n_steps, 64, 20261203)
    L = loss_of(propagate(prop, table(params), W0, n_steps, checkpoint_every=50))
    (g,) = torch.autograd.grad(L, params)
    h = 1e-6
    with torch.no_grad():
        for j in range(64):  # every control parameter
            e = torch.zeros_like(params)
            e[j] = h
            Lp = loss_of(propagate(prop, table(params + e), W0, n_steps))
            Lm = loss_of(propagate(prop, table(params - e), W0, n_steps))
            fd = float((Lp - Lm) / (2 * h))
            print(f"param {j}: adjoint {float(g[j]):+.10e}  central FD {fd:+.10e}")
            assert abs(float(g[j]) - fd) < 1e-6 * max(abs(fd), 1e-3), (j, float(g[j]), fd)
