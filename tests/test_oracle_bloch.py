"""CPU: pins of the imaginary-time (Bloch) Gross-Pitaevskii oracle (oracle/gp_bloch.py,
GPU_Wigner2D_GPitaevskii.py:2064-2543).

* a = 0, V = x^2/2, K = p^2/2: one Bloch step maps a pure Gaussian state psi = exp(-c x^2)
  to exp(-c' x^2) by the composed Gaussian maps c -> c + dt/8 (potential quarter step),
  c -> c / (1 + dt c) (kinetic half step in the p picture), c -> c + dt/8; the iteration's
  fixed point is the exact ground state of the split map, W = (1/pi) exp(-2c x^2 - p^2/(2c)),
  energy (1/(4c) + c)/2.  A factor slip (dt/4 vs dt/2, a lost hbar/2 in the shift) moves c.
* energies never increase along kept steps; a refused step halves dt.
* slow: the X2 notebook runs (Wigner2D_GPitaevskii_X2_Bloch.ipynb cells 11, 20, 21): W_0
  min/max, W_1 min/max and the four refusals at dt = 0.01 .. 0.00125 after step 700.
"""
import numpy as np
import pytest

from oracle.gp_bloch import GPBlochOracle


def _split_fixed_point(dt):
    c = 0.5
    for _ in range(100000):
        c1 = c + dt / 8.0
        c2 = c1 / (1.0 + dt * c1)  # 1/(4c') = 1/(4 c1) + dt/4
        c_new = c2 + dt / 8.0
        if abs(c_new - c) < 1e-16:
            break
        c = c_new
    return c_new


@pytest.mark.parametrize("dt", [0.05, 0.2])
def test_harmonic_ground_state_is_the_split_fixed_point(dt):
    o = GPBlochOracle(128, 128, 8.0, 8.0, 1.0, 1.0, dt, potential=lambda t, x: 0.5 * x * x,
                      kinetic=lambda p: 0.5 * p * p, gp_coeff=0.0, truncate_constants=False)
    g = o.g
    # the step map itself (a run's acceptance test refuses steps of a pure state once its
    # purity rounds to 1, and the energy passes its minimum at c = 1/2 before c* is reached)
    W = np.exp(-g.x_host ** 2 / 4.0 - g.p_host ** 2 / 4.0) + 0j
    for k in range(int(40.0 / dt)):
        W = o.step(W, dt, k * dt)
        W = W / (W.sum() * g.dX * g.dP)
    E, _ = o.energies(W)
    rec = {"E": [E], "purity": [o.purity(W)]}
    c = _split_fixed_point(dt)
    Wx = np.exp(-2.0 * c * g.x_host ** 2 - g.p_host ** 2 / (2.0 * c)) / np.pi
    err = np.linalg.norm(W - Wx) / np.linalg.norm(Wx)
    assert err < 1e-11, err
    assert abs(rec["E"][-1].real - 0.5 * (1.0 / (4.0 * c) + c)) < 1e-11
    assert abs(rec["purity"][-1] - 1.0) < 1e-11  # a pure state
    # the split fixed point is not the ground state of H (c = 1/2): it differs by ~dt^2/64,
    # far above the tolerance of the comparison
    assert abs(c - 0.5) > 1e3 * 1e-11 / dt


def test_energy_decreases_and_refusals_halve_dt():
    o = GPBlochOracle(128, 128, 8.0, 8.0, 1.0, 1.0, 0.05, potential=lambda t, x: 0.5 * x * x + 0.05 * x ** 4,
                      kinetic=lambda p: 0.5 * p * p, gp_coeff=1.0)
    g = o.g
    W0 = np.exp(-(g.x_host - 1.0) ** 2 / 3.0 - g.p_host ** 2 / 3.0)
    W, rec = o.run(W0, 200)
    acc = rec["accepted"]
    kept = np.real(rec["E"])[acc == 1]
    assert np.all(np.diff(kept) < 0)
    dts = rec["dt"]
    for k in range(1, len(acc)):
        if acc[k - 1] == 0:
            assert dts[k] == dts[k - 1] / 2.0
        else:
            assert dts[k] == dts[k - 1]
    assert len(rec["TotalEnergyHistory"]) == int((acc == 1).sum())
    # the Gross-Pitaevskii term widens the ground state: <x^2> above the linear one
    lin = GPBlochOracle(128, 128, 8.0, 8.0, 1.0, 1.0, 0.05, potential=lambda t, x: 0.5 * x * x + 0.05 * x ** 4,
                        kinetic=lambda p: 0.5 * p * p, gp_coeff=0.0)
    _, rec0 = lin.run(W0, 200)
    assert rec["X2"][-1].real > rec0["X2"][-1].real + 0.01


@pytest.mark.slow
def test_x2_notebook_ground_and_first_excited_state():
    """Wigner2D_GPitaevskii_X2_Bloch.ipynb: 512^2, amplitudes 16, hbar = m = 1, dt = 0.01,
    800 steps, a = 1, V = 0.5 x^2, W_init = exp(-(x^2 + p^2)/20) normalised."""
    o = GPBlochOracle(512, 512, 16.0, 16.0, 1.0, 1.0, 0.01, potential=lambda t, x: 0.5 * x * x,
                      kinetic=lambda p: 0.5 * p * p / 1.0, gp_coeff=1.0)
    g = o.g
    W0 = np.exp(-g.x_host ** 2 / 20.0 - g.p_host ** 2 / 20.0)
    W0 /= W0.sum() * g.dX * g.dP
    Wg, rec = o.run(W0, 800)
    W = np.fft.fftshift(Wg.real)
    assert abs(W.min() - (-0.000683046801767)) < 1e-12
    assert abs(W.max() - 0.317703452498) < 1e-11
    assert len(rec["NonLinearEnergyHistory"]) == 800
    W1, rec1 = o.run(W0, 800, W_ground=Wg)
    W = np.fft.fftshift(W1.real)
    assert abs(W.min() - (-0.303893689631)) < 1e-11
    assert abs(W.max() - 0.18269849465) < 1e-11
    refused = np.nonzero(rec1["accepted"] == 0)[0]
    assert list(rec1["dt"][refused]) == [0.01, 0.005, 0.0025, 0.00125]
    assert refused.min() > 700
