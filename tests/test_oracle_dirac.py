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


# ---------------------------------------------------------------------------- closed-form pins of
# the record columns, the order-2 friction and the Gross-Pitaevskii shift (each fails on a wrong
# plane, a flipped sign or a wrong coefficient)
ALPHA1 = np.zeros((4, 4))
ALPHA1[0, 3] = ALPHA1[3, 0] = ALPHA1[1, 2] = ALPHA1[2, 1] = 1.0
BETA = np.diag([1.0, 1.0, -1.0, -1.0])


def test_ehrenfest_velocity_is_c_alpha1():
    """d<x>/dt = c <alpha_1> (Alpha_1_Average record, D44.py:3074-3100, 3435): central differences
    of the X record against c Re Alpha1 over a run in the soft Coulomb well (no absorber).  Another
    plane choice (alpha_2, alpha_3) or a sign slip misses by O(1)."""
    dt = 0.0005
    orc = make(X=128, P=128, absorb=False)
    orc.dt = orc.k_dt = dt
    W0 = initial(orc, x0=-2.0, px=3.0)
    _, rec = orc.run(W0, 40)
    X = np.array([np.real(r["X"]) for r in rec])
    a1 = np.array([r["Alpha1"] for r in rec])
    v = (X[2:] - X[:-2]) / (2.0 * dt)
    np.testing.assert_allclose(v, orc.c * np.real(a1[1:-1]), rtol=2e-3)
    assert np.abs(np.imag(a1)).max() < 1e-12  # alpha_1 is Hermitian: a real expectation value
    assert 0.8 < np.real(a1[0]) < 1.0  # px = 3, m = c = 1: velocity ~ 3/sqrt(10) ~ 0.95

    # the other record columns through their own Ehrenfest identities (Weyl symbols):
    #   d<p>/dt = -<V0'>,  d<x^2>/dt = 2c<x alpha_1>,  d<xp>/dt = c<p alpha_1> - <x V0'>,
    #   d<p^2>/dt = -2<p V0'>
    def col(k):
        return np.real(np.array([r[k] for r in rec]))

    def ddt(a):
        return (a[2:] - a[:-2]) / (2.0 * dt)

    def close(lhs, rhs):
        scale = np.abs(rhs).max()
        assert np.abs(lhs - rhs).max() < 3e-3 * scale, (np.abs(lhs - rhs).max(), scale)

    close(ddt(col("P")), -col("D1V0")[1:-1])
    close(ddt(col("XX")), 2.0 * orc.c * col("XAlpha1")[1:-1])
    close(ddt(col("XP")), orc.c * col("PAlpha1")[1:-1] - col("XD1V0")[1:-1])
    close(ddt(col("PP")), -2.0 * col("PD1V0")[1:-1])


def _spinor_energy(orc, x0, px, s, V0):
    """<psi|c alpha_1 p + beta m c^2 + V0(x)|psi> of the Gaussian spinor, from its own 1-D
    momentum-space and position-space representations (independent of the Wigner planes)."""
    n = 4096
    xs = np.linspace(-40.0, 40.0, n, endpoint=False)
    dx = xs[1] - xs[0]
    psi = gaussian_spinor_particle_up(x0, px, s, xs, orc.mass, orc.c)
    norm = float(np.sum(np.abs(psi) ** 2)) * dx
    k = 2.0 * np.pi * np.fft.fftfreq(n, dx)
    pk = np.fft.fft(psi, axis=1)
    nk = float(np.sum(np.abs(pk) ** 2))
    kin = orc.c * np.real(np.einsum("ik,ij,jk->", np.conj(pk), ALPHA1, pk * k)) / nk
    rest = orc.mass * orc.c ** 2 * np.real(np.einsum("ik,ij,jk->", np.conj(pk), BETA, pk)) / nk
    pot = float(np.sum(np.abs(psi) ** 2 * V0(0.0, xs))) * dx / norm
    return kin + rest + pot


def test_energy_record_equals_spinor_hamiltonian_and_is_conserved():
    """DiracEnergy (D44.py:2777-2805) of the initial Wigner function equals <psi|H|psi> of the
    spinor it was formed from (c alpha_1 p + beta m c^2 + V0), and with a time-independent V0, no
    absorber and no damping the record is conserved to O(dt^2) while the particle accelerates
    into the well (a flipped beta or a missing V term is not conserved)."""
    orc = make(X=128, P=128, absorb=False, compute_energy=True)
    W0 = initial(orc, x0=-2.0, px=3.0, s=1.0)
    E0 = np.real(orc.observables(W0, 0.0)["energy"])
    np.testing.assert_allclose(E0, _spinor_energy(orc, -2.0, 3.0, 1.0, coulomb), rtol=1e-9)
    orc.dt = orc.k_dt = 0.002
    _, rec = orc.run(W0, 300)
    E = np.array([np.real(r["energy"]) for r in rec])
    drift = np.abs(E - E[0]).max()
    assert drift < 2e-4, drift
    # the kinetic part did change: the test would see a wrong split between the terms
    kin = np.array([np.real(r["PAlpha1"]) for r in rec])
    assert abs(kin[-1] - kin[0]) > 50 * drift


def test_caldeira_leggett_order2_friction_factor():
    """CaldeiraDissipatorOrder2 (D44.py:3019-3058) with sigma(p) = p: D(W) = -i d/dp (p W), so
    W + 2aDW + 2a^2 D^2 W (a = i gamma dt) multiplies <p> by exactly 1 - 2 gamma dt + 2 (gamma dt)^2
    per step for a free spinor (kinetic and mass steps keep <p>; the trace norm is unchanged).
    A wrong coefficient of the second-order term shifts the factor by (gamma dt)^2 = 2.5e-5."""
    g, dt = 0.5, 0.01
    orc = make(X=128, P=128, potentials=[zero, zero, zero, zero], gamma=g, damping="CaldeiraLegget",
               absorb=False)
    orc.dt = orc.k_dt = dt
    W0 = initial(orc, x0=-2.0, px=3.0)
    _, rec = orc.run(W0, 5)
    P = np.array([np.real(r["P"]) for r in rec])
    np.testing.assert_allclose(P[1:] / P[:-1], 1.0 - 2.0 * g * dt + 2.0 * (g * dt) ** 2, rtol=1e-10)
    # and the direct operator: D(p-Gaussian) against -i d/dp (p W) in closed form
    p = orc.p_host
    Y = np.exp(-((p - 1.0) ** 2)) * np.ones((1, orc.X))
    D = orc.theta_to_p(orc.p_to_theta(Y * p) * orc.theta_host)
    dY = -1j * (np.exp(-((p - 1.0) ** 2)) * (1.0 - 2.0 * p * (p - 1.0))) * np.ones((1, orc.X))
    np.testing.assert_allclose(D, dY, atol=1e-10)


def test_gp_terms_shift_theorem_on_trace_density():
    """B-(theta,x) = rho(x - theta/2), B+ = rho(x + theta/2) with rho = Re sum_p tr W (no dP)
    (MakeGrossPitaevskiiTerms, D44.py:2695-2715): a state whose trace density is an analytic
    Gaussian, with unrelated content in the off-diagonal planes (which must not contribute)."""
    orc = make(X=256, P=64, gp_coeff=1.0)
    x, th = orc.x_host, orc.theta_host
    W = np.zeros((4, 4, orc.P, orc.X), dtype=np.complex128)
    rho = lambda y: np.exp(-((y - 0.7) ** 2) / 1.5)  # noqa: E731
    # split the density over the four diagonal planes and put it in the p = 0 row only
    for k, w in zip(range(4), (0.4, 0.3, 0.2, 0.1)):
        W[k, k, 0] = w * rho(x[0])
    W[0, 3] = 5.0 + 2.0j  # off-diagonal planes: ignored by the trace
    W[1, 2, 3] = -7.0
    Bm, Bp = orc.gp_terms(W)
    inside = np.abs(th) < 8.0
    sel = np.broadcast_to(inside, Bm.shape)
    np.testing.assert_allclose(Bm[sel], np.broadcast_to(rho(x - 0.5 * th), Bm.shape)[sel], atol=1e-10)
    np.testing.assert_allclose(Bp[sel], np.broadcast_to(rho(x + 0.5 * th), Bp.shape)[sel], atol=1e-10)


def test_absorber_projector_and_population_pins():
    """Absorber 1 - exp(-((j - X/2)/24)^2) over the array index (D44.py:111-155): zero at the
    array middle (x = +-X_amp in FFT order), -> 1 far from it; the electron projectors
    (Filter_Kernel, :674-760) are complementary orthogonal projectors for every q."""
    orc = make(X=128, P=64)
    a = orc.absorber()[0]
    assert a[64] == 0.0 and a[0] > 0.99 and abs(a[64 + 24] - (1.0 - np.exp(-1.0))) < 1e-15
    q = np.linspace(-30.0, 30.0, 61)
    Pp, Pm = orc.projector(q, +1), orc.projector(q, -1)
    eye = np.eye(4)[:, :, None]
    np.testing.assert_allclose(Pp + Pm, np.broadcast_to(eye, Pp.shape), atol=1e-15)
    np.testing.assert_allclose(np.einsum("ik...,kj...->ij...", Pp, Pp), Pp, atol=1e-14)
    np.testing.assert_allclose(np.einsum("ik...,kj...->ij...", Pp, Pm), 0.0, atol=1e-14)
    # the positive-energy projector commutes with the free Hamiltonian c alpha_1 q + beta m c^2
    H = orc.c * ALPHA1[:, :, None] * q + BETA[:, :, None] * orc.mass * orc.c ** 2
    np.testing.assert_allclose(np.einsum("ik...,kj...->ij...", H, Pp), np.einsum("ik...,kj...->ij...", Pp, H),
                               atol=1e-12)


# ---------------------------------------------------------------------------- ODM damping
def _theta_state(orc):
    W = initial(orc, x0=-1.0, px=2.0, s=1.2)
    return orc.p_to_theta(W)  # the (theta, x) picture the damping acts in


def test_odm_identity_at_zero_damping_and_first_order_generator():
    """In matrix form the damping map is W' = 1/2[A cx (W - a W a) + B ct (W + a W a)
    + i A sx (W a - a W) + i B st (W a + a W)] with a = alpha_1 (the spinor pairing i <-> 5-i).
    gamma = 0 gives the identity; to first order in k = m c dt gamma it is
    W + k{ -4 x^2/lb (W - a W a) - theta^2/lb (W + a W a) - 2 i x [a, W] + i theta {a, W} }
    (a wrong partner plane or a sign slip in any of the 16 outputs fails this)."""
    orc0 = make(X=64, P=64, gamma=0.0, damping="ODM", lambda_bar=4.0)
    W = _theta_state(orc0)
    np.testing.assert_allclose(orc0.odm(W), W, atol=1e-15)
    g = 1e-6
    orc = make(X=64, P=64, gamma=g, damping="ODM", lambda_bar=4.0)
    k = orc.k_mass * orc.k_c * orc.k_dt * g
    a = ALPHA1.astype(np.complex128)
    aWa = np.einsum("ik,kj...,jl->il...", a, W, a)
    Wa = np.einsum("ik...,kj->ij...", W, a)
    aW = np.einsum("ik,kj...->ij...", a, W)
    x, th = orc.kx, orc.ktheta
    lin = W + k * (-4.0 * x * x / 4.0 * (W - aWa) - th * th / 4.0 * (W + aWa)
                   - 2j * x * (aW - Wa) + 1j * th * (aW + Wa))
    err = np.abs(orc.odm(W) - lin).max() / np.abs(W).max()
    assert err < 1e-9, err  # second order: ~ (k x^2)^2 ~ 1e-14 .. 1e-10 here


def test_odm_conserves_position_density_and_hermiticity():
    """The map leaves the theta = 0 row of the trace (the position density sum_p tr W) exactly
    unchanged, and keeps W_ij(theta, x) = conj W_ji(-theta, x) (the (theta, x) form of
    W_ij(x, p) = conj W_ji(x, p))."""
    orc = make(X=64, P=64, gamma=0.05, damping="ODM", lambda_bar=4.0)
    W = _theta_state(orc)
    Wp = orc.odm(W)
    tr, trp = orc.trace(W), orc.trace(Wp)
    np.testing.assert_allclose(trp[0], tr[0], rtol=1e-13, atol=1e-15)
    assert np.abs(trp - tr).max() > 1e-6  # it does act away from theta = 0
    flip = np.roll(Wp[:, :, ::-1, :], 1, axis=2)  # theta -> -theta in FFT order
    dev = np.abs(Wp - np.conj(np.swapaxes(flip, 0, 1)))
    dev[:, :, orc.P // 2, :] = 0.0  # the Nyquist row (-theta_max) has no +theta_max partner
    assert dev.max() < 1e-13 * np.abs(Wp).max()


# ---------------------------------------------------------------------------- step filter and KvN
def test_step_filter_is_an_idempotent_projection():
    """W <- Lambda_+(p + lambda/2) W Lambda_+(p - lambda/2): applying it twice equals once (the
    projectors are idempotent); it keeps a positive-energy spinor state almost whole, and its
    trace norm and the antiparticle population add up to the norm."""
    orc = make(X=128, P=128, step_filter=True)
    W = initial(orc)
    F1 = orc.filter_particles(W)
    F2 = orc.filter_particles(F1)
    assert np.abs(F2 - F1).max() < 1e-13 * np.abs(F1).max()
    assert abs(orc.norm(F1) - 1.0) < 5e-3  # the positive-energy Gaussian spinor (~0.2% antiparticle)
    # the trace norm splits: tr(L+ W L+) + tr(L- W L-) = tr W at lambda = 0 (L+ + L- = 1, L^2 = L)
    anti = orc.antiparticle_population(orc.x_to_lambda(W), -1)
    assert abs(orc.norm(F1) + anti - orc.norm(W)) < 1e-13


def test_kvn_potential_step_is_the_classical_force():
    """KvN potential step e^{-dt D theta^2 + i dt V0' theta}: for V0 = F x it shifts the
    momentum marginal by -F dt exactly (dp/dt = -V0'); the trace norm is unchanged.  Exact
    constants: the '%f'-truncated dtheta would scale the shift by k_dtheta / dtheta."""
    F = 1.5
    orc = make(X=128, P=128, potentials=[lambda t, x: F * x, zero, zero, zero], kvn=True,
               truncate_constants=False)
    orc.dV0 = lambda t, x: F + 0.0 * x
    W = initial(orc, x0=-1.0, px=2.0)
    p0 = np.real(orc.observables(W, 0.0)["P"])
    W1 = orc.theta_to_p(orc.potential_step(orc.p_to_theta(W), 0.0))
    p1 = np.real(orc.observables(W1, 0.0)["P"])
    assert abs((p1 - p0) - (-F * orc.k_dt)) < 1e-10
    assert abs(orc.norm(W1) - orc.norm(W)) < 1e-13


def test_kvn_kinetic_factor_unitary_without_mass_term():
    orc = make(X=64, P=64, kvn=True)
    assert abs(orc.k_mh - 0.00001) < 1e-18
    q = np.linspace(-5, 5, 41)
    for sgn in (+1, -1):
        U = orc.kinetic_factor(q, sgn)
        assert unitary_err(U) < 1e-13
        np.testing.assert_allclose(U[0, 0].imag, 0.0, atol=0)  # cos(c dt s) alone on the diagonal
