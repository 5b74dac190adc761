"""Pins of the Wigner-Dirac 4x4 oracle (oracle/dirac4x4.py) by closed forms and invariants.

The reference has no test suite and no recorded numbers for the Dirac classes, so the
oracle is pinned by what the mathematics fixes:
* each per-cell factor is unitary and of the form cos F - i (sin F / F) G with G
  Hermitian and G^2 = F^2 (a dropped term or a wrong sign breaks one of these);
* the kinetic step maps the pure-state form Psi(p + lambda/2) Psi(p - lambda/2)^dagger to
  (U Psi)(p + lambda/2) (U Psi)(p - lambda/2)^dagger exactly, and the potential step maps
  Phi(x - theta/2) Phi(x + theta/2)^dagger likewise (a transposed operand, a missing
  dagger or a swapped side fails this);
* without absorber the trace norm is conserved and W_ij = conj W_ji is preserved;
* particle + antiparticle populations add up to the trace norm;
* the initial Gaussian spinor state has closed-form moments <x> = x0, <x^2> = x0^2 + s^2/2,
  <p> = px, <p^2> = px^2 + 1/(2 s^2).
"""
import numpy as np
import pytest

from oracle.dirac4x4 import Dirac4x4Oracle, gaussian_spinor_particle_up, wigner_from_spinor


def zero(t, x):
    return 0.0 * x


def coulomb(t, x):
    return -2.0 / np.sqrt(x * x + 1.0)


def make(X=64, P=64, potentials=None, **kw):
    pots = potentials or [coulomb, zero, zero, zero]
    return Dirac4x4Oracle(X, P, 18.0, 16.0, 1.0, 1.0, 0.01, pots,
                          dpotential0=lambda t, x: 2.0 * x / (x * x + 1.0) ** 1.5, **kw)


def initial(orc, x0=-2.0, px=3.0, s=1.0):
    return wigner_from_spinor(lambda X: gaussian_spinor_particle_up(x0, px, s, X, orc.mass, orc.c), orc)


def unitary_err(U):
    Uh = np.conj(np.swapaxes(U, 0, 1))
    I = np.einsum("ik...,kj...->ij...", U, Uh)
    eye = np.eye(4).reshape(4, 4, *([1] * (U.ndim - 2)))
    return np.abs(I - eye).max()


def test_kinetic_factor_unitary_and_adjoint_pair():
    orc = make()
    q = np.linspace(-20, 20, 101)
    Up, Um = orc.kinetic_factor(q, +1), orc.kinetic_factor(q, -1)
    assert unitary_err(Up) < 1e-14
    # the - factor is the + factor with dt -> -dt, i.e. its adjoint
    np.testing.assert_allclose(Um, np.conj(np.swapaxes(Up, 0, 1)), atol=1e-15)
    # cos F - i (sin F/F) G with G = mass_half c beta + q alpha_1 (G^2 = F^2)
    mh, c, dt = orc.k_mh, orc.k_c, orc.k_dt
    beta = np.diag([1.0, 1.0, -1.0, -1.0])
    alpha1 = np.zeros((4, 4))
    alpha1[0, 3] = alpha1[3, 0] = alpha1[1, 2] = alpha1[2, 1] = 1.0
    for k in (0, 37, 100):
        G = c * (mh * c * beta + q[k] * alpha1)
        F = dt * c * np.sqrt((mh * c) ** 2 + q[k] ** 2)
        U = np.cos(F) * np.eye(4) - 1j * np.sin(F) / F * dt * G
        np.testing.assert_allclose(Up[:, :, k], U, atol=1e-15)


@pytest.mark.parametrize("A", [(0.0, 0.0, 0.0), (0.3, -0.7, 0.2)])
def test_potential_factor_unitary_hermitian_generator(A):
    pots = [coulomb] + [(lambda a: (lambda t, x: a + 0.1 * x))(a) for a in A]
    orc = make(potentials=pots)
    Um, em = orc.potential_factor(0.0, -1)
    Up, ep = orc.potential_factor(0.0, +1)
    assert unitary_err(Um) < 1e-14 and unitary_err(Up) < 1e-14
    np.testing.assert_allclose(np.abs(em), 1.0, atol=1e-15)
    # U = cos F - i (sin F / F) G with G Hermitian, G^2 = F^2 I
    j, i = 5, 3
    U = Um[:, :, i, j]
    F = np.arccos(np.real(U[0, 0]) if abs(np.real(U[0, 0])) <= 1 else 1.0)
    G = (U - np.cos(F) * np.eye(4)) * (1j * F / np.sin(F))
    np.testing.assert_allclose(G, np.conj(G.T), atol=1e-13)
    np.testing.assert_allclose(G @ G, F * F * np.eye(4), atol=1e-13)


def test_kinetic_step_exact_on_pure_state_form():
    orc = make()
    q = lambda p, l: p + 0.5 * l  # noqa: E731
    p, lam = orc.kp, orc.klambda

    def Psi(qq):
        return np.array([np.exp(-qq * qq / 8.0), 0.3j * qq * np.exp(-qq * qq / 6.0),
                         np.cos(qq) * np.exp(-qq * qq / 10.0), (0.2 + 0.1j) * np.exp(-(qq - 1) ** 2 / 5.0)])

    a, b = Psi(q(p, lam)), Psi(p - 0.5 * lam)
    W = np.einsum("i...,j...->ij...", a, np.conj(b))
    out = orc.kinetic_step(W)
    Ua = np.einsum("ik...,k...->i...", orc.kinetic_factor(p + 0.5 * lam, +1), a)
    Ub = np.einsum("ik...,k...->i...", orc.kinetic_factor(p - 0.5 * lam, +1), b)
    np.testing.assert_allclose(out, np.einsum("i...,j...->ij...", Ua, np.conj(Ub)), atol=1e-14)


@pytest.mark.parametrize("mass_potential", [None, lambda t, x: 0.5 + 0.4 * np.tanh(x / 2.0)])
def test_potential_step_exact_on_pure_state_form(mass_potential):
    pots = [coulomb, lambda t, x: 0.2 * np.sin(x), lambda t, x: -0.1 * x, lambda t, x: 0.05 + 0 * x]
    orc = make(potentials=pots, mass_potential=mass_potential)
    x, th = orc.kx, orc.ktheta

    def Phi(y):
        return np.array([np.exp(-y * y / 8.0), 0.3j * y * np.exp(-y * y / 6.0),
                         np.cos(y) * np.exp(-y * y / 10.0), (0.2 + 0.1j) * np.exp(-(y - 1) ** 2 / 5.0)])

    a, b = Phi(x - 0.5 * th), Phi(x + 0.5 * th)
    W = np.einsum("i...,j...->ij...", a, np.conj(b))
    out = orc.potential_step(W, 0.0)
    Um, em = orc.potential_factor(0.0, -1)  # at x - theta/2
    # the + factor at x + theta/2 is the adjoint of the - factor there (D_theta = 0)
    orc2 = make(potentials=pots, mass_potential=mass_potential)
    orc2.kx = orc.kx + orc.ktheta  # evaluate the - factor at x + theta/2
    Ub, eb = orc2.potential_factor(0.0, -1)
    A = em * np.einsum("ik...,k...->i...", Um, a)
    B = eb * np.einsum("ik...,k...->i...", Ub, b)
    np.testing.assert_allclose(out, np.einsum("i...,j...->ij...", A, np.conj(B)), atol=1e-13)


def test_trace_norm_and_hermiticity_conserved_without_absorber():
    # 128^2: the state has no weight at the Nyquist theta / lambda, where the discrete
    # step cannot keep W_ij = conj W_ji (theta_N and -theta_N share one FFT row)
    # (a smooth Klein step; the Coulomb cusp of V0 feeds the Nyquist rows within two steps)
    step = lambda t, x: 0.5 * (1.0 + np.tanh((x - 5.0) / 0.5))  # noqa: E731
    orc = make(X=128, P=128, absorb=False, potentials=[step, zero, zero, zero])
    W = initial(orc)
    for n in range(1, 6):
        t = orc.dt * n
        W = orc.theta_to_p(orc.potential_step(orc.p_to_theta(orc.lambda_to_x(orc.kinetic_step(orc.x_to_lambda(W)))), t))
        assert abs(orc.norm(W) - 1.0) < 1e-13
        assert np.abs(W - np.conj(np.swapaxes(W, 0, 1))).max() < 1e-13


def test_populations_sum_to_norm_and_initial_moments():
    orc = make(X=128, P=128)
    x0, px, s = -2.0, 3.0, 1.0
    W = initial(orc, x0, px, s)
    Wl = orc.x_to_lambda(W)
    assert abs(orc.antiparticle_population(Wl, -1) + orc.antiparticle_population(Wl, +1) - 1.0) < 1e-12
    o = orc.observables(W, 0.0)
    np.testing.assert_allclose(np.real(o["X"]), x0, atol=1e-10)
    np.testing.assert_allclose(np.real(o["XX"]), x0 * x0 + s * s / 2.0, atol=1e-9)
    np.testing.assert_allclose(np.real(o["P"]), px, atol=1e-10)
    np.testing.assert_allclose(np.real(o["PP"]), px * px + 1.0 / (2 * s * s), atol=1e-8)
    # a positive-energy spinor at px: the antiparticle population is tiny
    assert orc.antiparticle_population(Wl, -1) < 1e-2


def test_run_normalises_and_records():
    orc = make(compute_energy=True)
    W, rec = orc.run(initial(orc), 3)
    assert len(rec) == 3 and abs(orc.norm(W) - 1.0) < 1e-14
    assert {"X", "Alpha1", "negativity", "antiparticle", "energy"} <= set(rec[0])
