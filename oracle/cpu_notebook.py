"""Oracle: the reference's own CPU propagator, CPU_GPU_Wigner2D.ipynb cells 6-11.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The notebook propagates a real W on a natural-order grid (index 0 = -amplitude)
with real FFTs (pyFFTW r2c/c2r, FFTW_ESTIMATE, single thread) and the Strang
sequence  p->theta, B*=expV, theta->p, x->lambda, Z*=expK, lambda->x,
p->theta, B*=expV, theta->p  (cell 10, raw lines :246-270), hbar = 1:

    expV = exp(-dt*0.5j*(V(X - Theta/2) - V(X + Theta/2)))      cell 7, :139-160
    expK = exp(-dt*1j *(K(P + Lambda/2) - K(P - Lambda/2)))
    Lambda = fftfreq(X, dX/2pi)[:X//2+1],  Theta = fftfreq(P, dP/2pi)[:P//2+1]   cell 6

pyFFTW is absent; numpy.fft.rfft/irfft (pocketfft) compute the same transforms.
The notebook calls its FFTW objects (``Theta_To_P_FFTW()``), and pyFFTW's
``FFTW.__call__`` normalises the backward transform by 1/N, as numpy.fft.irfft
does; the recorded norm 1.0000000000000249 confirms that normalisation.

Known answers recorded by the notebook (cells 24-26) for: 512x512, X_amp=P_amp=12,
dt=0.05, V=0.5x^2+0.01x^4, K=p^2/2, W0 ~ exp(-((p-3)^2 + x^2)) normalised,
500 steps:  min(W) = -0.16857913679981978 (:557),  sum(W) dX dP = 1.0000000000000249 (:577).
"""
from __future__ import annotations

import numpy as np

K1_MIN_W = -0.16857913679981978
K1_NORM = 1.0000000000000249


def set_grid(X_amp, P_amp, X, P):
    """Cell 6 (SetGrid)."""
    dX = 2.0 * X_amp / X
    dP = 2.0 * P_amp / P
    Xg = np.linspace(-X_amp, X_amp - dX, X)[None, :]
    Lam = np.fft.fftfreq(X, dX / (2 * np.pi))[: (1 + X // 2)][None, :]
    Pg = np.linspace(-P_amp, P_amp - dP, P)[:, None]
    Th = np.fft.fftfreq(P, dP / (2 * np.pi))[: (1 + P // 2)][:, None]
    return Xg, Pg, Th, Lam, dX, dP


def set_exponentials(Xg, Pg, Th, Lam, dt, V, K):
    """Cell 7 (SetExponentials)."""
    expV = np.exp(-dt * 0.5j * (V(Xg - 0.5 * Th) - V(Xg + 0.5 * Th)))
    expK = np.exp(-dt * 1j * (K(Pg + 0.5 * Lam) - K(Pg - 0.5 * Lam)))
    return expK, expV


def propagation_step(W, expK, expV, P, X):
    """Cell 10 (Propagation_Step): V/2, K, V/2 on a real W (natural order)."""
    B = np.fft.rfft(W, axis=0) * expV
    W = np.fft.irfft(B, n=P, axis=0)
    Z = np.fft.rfft(W, axis=1) * expK
    W = np.fft.irfft(Z, n=X, axis=1)
    B = np.fft.rfft(W, axis=0) * expV
    return np.fft.irfft(B, n=P, axis=0)


def k1_setup(X=512, P=512):
    """Cells 15-22: the configuration behind the recorded known answers."""
    X_amp = P_amp = 12.0
    dt = 0.05
    V = lambda x: 0.5 * np.power(x, 2) + 0.01 * np.power(x, 4)
    K = lambda p: np.power(p, 2) / 2.0
    Xg, Pg, Th, Lam, dX, dP = set_grid(X_amp, P_amp, X, P)
    expK, expV = set_exponentials(Xg, Pg, Th, Lam, dt, V, K)
    W = np.exp(-((Pg - 3) ** 2 / 1.0 + 1.0 * 1.0 ** 2 * Xg ** 2))  # cell 21
    W = W / (np.sum(W) * dX * dP)
    return dict(W=W, expK=expK, expV=expV, dX=dX, dP=dP, X=X, P=P, dt=dt, V=V, K=K,
                X_amp=X_amp, P_amp=P_amp)


def run(W, expK, expV, P, X, time_steps):
    """Cell 11 (Run): time_steps Strang steps (no extra +1 iteration here)."""
    for _ in range(time_steps):
        W = propagation_step(W, expK, expV, P, X)
    return W
