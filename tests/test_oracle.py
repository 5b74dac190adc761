"""CPU: the oracle pinned to the reference's recorded outputs and to closed forms.

Pins (each chosen so that a plausible slip fails it):
  * K1, the only numbers the reference records for its CPU propagator
    (CPU_GPU_Wigner2D.ipynb:557, :577) -> oracle.cpu_notebook;
  * the complex FFT-ordered oracle (the form the GPU classes use) against that
    notebook propagator;
  * exact split-step maps: for quadratic V every Lie / Strang step maps a Gaussian
    Wigner function to the Gaussian with the linear (leapfrog) map applied, and the
    free particle shears -- a sign or factor error in either multiplier fails these;
  * observables of Gaussians (closed-form moments);
  * Caldeira-Leggett: the friction sign and factor via d<p>/dt = -2 gamma <p>, and the
    reference's nested accumulation expanded as a polynomial in D;
  * Gross-Pitaevskii: the shifted densities equal rho(x -/+ theta/2) (shift theorem);
  * KvN (Fokker-Planck) potential step: the classical momentum kick
    W(x, p) -> W(x, p + dt V'(x)) of a Gaussian in closed form, and for quadratic V the
    classical and quantum propagators coincide;
  * '%f' splicing of constants.
"""
import numpy as np
import pytest

from oracle import cpu_notebook as nb
from oracle import wigner2d as ow


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ----------------------------------------------------------------------------- notebook K1
def test_notebook_known_answer_k1():
    s = nb.k1_setup()
    W = nb.run(s["W"], s["expK"], s["expV"], s["P"], s["X"], 500)
    assert abs(W.min() - nb.K1_MIN_W) < 1e-11 * abs(nb.K1_MIN_W)
    norm = W.sum() * s["dX"] * s["dP"]
    assert abs(norm - nb.K1_NORM) < 1e-12


def test_complex_oracle_matches_notebook_propagator():
    """Same physics as K1 on 256^2 for 100 steps: complex FFT-ordered Strang vs real-FFT notebook."""
    s = nb.k1_setup(256, 256)
    Wn = nb.run(s["W"], s["expK"], s["expV"], s["P"], s["X"], 100)
    o = ow.Wigner2DOracle(256, 256, 12.0, 12.0, 1.0, 1.0, 0.05, scheme="strang", mask=False,
                          potential=lambda t, x: 0.5 * x ** 2 + 0.01 * x ** 4,
                          kinetic=lambda p: p ** 2 / 2.0)
    W, obs, _ = o.run(np.fft.ifftshift(s["W"]), 99)
    assert rel_l2(np.fft.fftshift(W.real), Wn) < 1e-12
    assert np.abs(W.imag).max() < 1e-13


# ----------------------------------------------------------------------------- closed forms
def gaussian_on(grid, mu, cov):
    """Normalised Gaussian Wigner function with mean mu=(x,p) and covariance cov on the host grid."""
    x, p = grid.x_host, grid.p_host
    inv = np.linalg.inv(cov)
    dx, dp = x - mu[0], p - mu[1]
    q = inv[0, 0] * dx * dx + 2 * inv[0, 1] * dx * dp + inv[1, 1] * dp * dp
    return np.exp(-0.5 * q) / (2 * np.pi * np.sqrt(np.linalg.det(cov)))


def leapfrog_matrix(scheme, dt, omega2, m=1.0):
    K = np.array([[1.0, dt / m], [0.0, 1.0]])  # x -> x + p dt/m
    V = lambda h: np.array([[1.0, 0.0], [-h * omega2, 1.0]])  # noqa: E731  p -> p - h V'(x)
    if scheme == "lie":
        return V(dt) @ K  # K first, then V (GPU_Wigner2D_FFT.py:1884-1899)
    return V(dt / 2) @ K @ V(dt / 2)


@pytest.mark.parametrize("scheme", ["lie", "strang"])
def test_harmonic_split_step_is_exact_linear_map(scheme):
    X = P = 256
    dt, steps, omega2 = 0.05, 60, 1.0
    o = ow.Wigner2DOracle(X, P, 10.0, 10.0, 1.0, 1.0, dt, scheme=scheme, mask=False,
                          potential=lambda t, x: 0.5 * omega2 * x * x, dpotential=lambda t, x: omega2 * x,
                          kinetic=lambda p: 0.5 * p * p)
    mu0, cov0 = np.array([1.5, -0.5]), np.array([[0.5, 0.0], [0.0, 0.5]])
    W0 = gaussian_on(o.g, mu0, cov0)
    W, obs, _ = o.run(W0, steps - 1)  # `steps` iterations
    M = np.linalg.matrix_power(leapfrog_matrix(scheme, dt, omega2), steps)
    W_exact = gaussian_on(o.g, M @ mu0, M @ cov0 @ M.T)
    assert rel_l2(W.real, W_exact) < 1e-10
    # first moments follow the map exactly; Lie records pre-step states
    M1 = leapfrog_matrix(scheme, dt, omega2)
    k = steps - 1 if scheme == "lie" else steps
    mu = np.linalg.matrix_power(M1, k) @ mu0
    assert abs(obs["X"][-1] - mu[0]) < 1e-12 and abs(obs["P"][-1] - mu[1]) < 1e-12


def test_free_particle_shear():
    X = P = 256
    dt, steps = 0.02, 50
    o = ow.Wigner2DOracle(X, P, 12.0, 12.0, 1.0, 2.0, dt, scheme="lie", mask=False,
                          potential=lambda t, x: 0.0 * x)
    mu0, cov0 = np.array([-2.0, 1.0]), np.array([[0.6, 0.1], [0.1, 0.4]])
    W, _, _ = o.run(gaussian_on(o.g, mu0, cov0), steps - 1)
    S = np.array([[1.0, steps * dt / 2.0], [0.0, 1.0]])  # mass 2
    assert rel_l2(W.real, gaussian_on(o.g, S @ mu0, S @ cov0 @ S.T)) < 1e-10


def test_observables_of_a_gaussian():
    o = ow.Wigner2DOracle(256, 256, 10.0, 10.0, 1.0, 1.0, 0.01, mask=False,
                          potential=lambda t, x: 0.5 * x * x + 0.1 * x ** 4,
                          dpotential=lambda t, x: x + 0.4 * x ** 3)
    mu, cov = np.array([0.7, -1.1]), np.array([[0.4, 0.05], [0.05, 0.6]])
    ob = o.observables(o.normalise(gaussian_on(o.g, mu, cov)))
    sx, sp, sxp = cov[0, 0], cov[1, 1], cov[0, 1]
    x4 = mu[0] ** 4 + 6 * mu[0] ** 2 * sx + 3 * sx ** 2
    x3 = mu[0] ** 3 + 3 * mu[0] * sx
    exact = {"X": mu[0], "P": mu[1], "X2": sx + mu[0] ** 2, "P2": sp + mu[1] ** 2,
             "XP": sxp + mu[0] * mu[1], "dV": mu[0] + 0.4 * x3,
             "H": 0.5 * (sp + mu[1] ** 2) + 0.5 * (sx + mu[0] ** 2) + 0.1 * x4, "norm": 1.0}
    for k, v in exact.items():
        assert abs(ob[k] - v) < 1e-10, (k, ob[k], v)
    assert ob["neg"] == 0.0


# ----------------------------------------------------------------------------- Caldeira-Leggett
def test_caldeira_leggett_friction_decays_momentum():
    """Free particle, sigma(p) = p: the friction gives d<p>/dt = -2 gamma <p> (FFT.py:1145-1157)."""
    gamma, dt, steps = 0.2, 0.01, 100
    o = ow.Wigner2DOracle(256, 256, 16.0, 16.0, 1.0, 1.0, dt, gamma=gamma, damping="CaldeiraLeggett",
                          fp_damping=lambda p: np.abs(p), mask=False, potential=lambda t, x: 0.0 * x)
    W0 = gaussian_on(o.g, np.array([0.0, 2.0]), np.array([[0.5, 0.0], [0.0, 0.5]]))
    _, obs, _ = o.run(W0, steps)
    t = np.arange(steps + 1) * dt
    assert np.allclose(obs["P"], 2.0 * np.exp(-2 * gamma * t), rtol=2e-3)
    assert obs["P"][-1] < obs["P"][0]


def test_caldeira_leggett_composite_expansion():
    o = ow.Wigner2DOracle(128, 128, 10.0, 10.0, 1.0, 1.0, 0.05, gamma=0.3, damping="CaldeiraLeggett",
                          fp_damping=lambda p: p * p / np.sqrt(p * p + 0.5), mask=False,
                          potential=lambda t, x: 0.5 * x * x)
    rng = np.random.default_rng(3)
    W = o.normalise(gaussian_on(o.g, np.array([1.0, 0.5]), np.eye(2) * 0.7)) * (1 + 0.01 * rng.standard_normal((128, 128)))
    a = 1j * o.dt * o.gamma
    D1 = o.cl_operator(W)
    D2 = o.cl_operator(D1)
    D3 = o.cl_operator(D2)
    expected = W + 2 * a * D1 + (10.0 / 3.0) * a ** 2 * D2 + (4.0 / 3.0) * a ** 3 * D3
    assert rel_l2(o.caldeira_order3(W), expected) < 1e-13


# ----------------------------------------------------------------------------- Gross-Pitaevskii
def test_gp_shifted_densities_shift_theorem():
    o = ow.Wigner2DOracle(256, 128, 12.0, 12.0, 1.0, 1.0, 0.01, gp_coeff=1.0, mask=False,
                          potential=lambda t, x: 0.0 * x)
    x = o.g.x_host[0]
    rho = np.exp(-((x - 0.5) ** 2) / 0.8)
    Bm, Bp = o.gp_shifted(rho)
    th = o.g.theta_host[:, 0]
    for i in range(0, 128, 7):
        if abs(th[i]) > 8:  # keep the shifted Gaussian inside the periodic box
            continue
        for B, sgn in ((Bm, -1), (Bp, +1)):
            exact = np.exp(-((x + sgn * th[i] / 2 - 0.5) ** 2) / 0.8)
            assert np.abs(B[i].real - exact).max() < 1e-12


# ----------------------------------------------------------------------------- conventions
def test_percent_f_splicing():
    assert ow.percent_f(np.pi / 12) == 0.261799
    assert ow.percent_f(0.005) == 0.005


def test_fft_ordered_axes_match_reference_kernel_map():
    # i = (k + P/2) % P; theta = dtheta*(i - P/2)  (GPU_Wigner2D_FFT.py:110-114)
    P, d = 8, 0.5
    expect = [d * (((k + P // 2) % P) - P / 2) for k in range(P)]
    assert np.allclose(ow.fft_order_axis(P, d), expect)
    assert np.allclose(ow.host_axis(2.0, P, d), np.fft.fftshift(np.linspace(-2.0, 2.0 - d, P)))


def test_odm_six_terms_split_into_x_and_theta_factors():
    """_Lindbladian_ODM (FFT.py:1516-1634) as restated equals its split into x-only and
    theta-only factors (product-to-sum identities of cos/sin(x +- hbar theta/2)), the form the
    fused ODM sweeps compute; and the host tables of the product path equal the oracle's."""
    from pywignercuda_amd.wigner2d import _Tables

    orc = ow.Wigner2DOracle(128, 64, 12.0, 10.0, 0.7, 1.0, 0.005, gamma=0.0175, potential=lambda t, x: 0.5 * x * x,
                            fp_damping=lambda p: p * p / np.sqrt(p * p + 0.5), damping="ODM", L_material=3.0)
    W = ow.harmonic_oscillator_wigner(orc.g, 0, 1.0, 0.5, 2.0) * (1 + 0.1j)
    ref = orc.odm_operator(W)
    T = orc.odm_tables()
    g, h = orc.g, orc.hbar
    A = T["gp"] * T["gm"] - T["gsp"] * T["gsm"]
    B = T["gp"] * T["gsm"] - T["gsp"] * T["gm"]
    C = -0.5 * (T["gp"] ** 2 + T["gm"] ** 2)
    E = T["gp"] * T["gm"] + T["gsp"] * T["gsm"]
    F = T["gp"] * T["gsm"] + T["gsp"] * T["gm"]
    Wh = np.fft.fft(W, axis=1)
    ix = lambda Y: np.fft.ifft(Y, axis=1)  # noqa: E731
    Lx = 0.5 * np.cos(2 * g.x_host) * ix(A * Wh) + 0.5j * np.sin(2 * g.x_host) * ix(B * Wh) + ix(C * Wh)
    Q = 0.5 * (np.cos(h * g.theta_host) * np.fft.fft(E * Wh, axis=0)
               + 1j * np.sin(h * g.theta_host) * np.fft.fft(F * Wh, axis=0))
    got = (Lx + ix(np.fft.ifft(Q, axis=0))) * (orc.dt * 2 * orc.gamma)
    assert np.abs(got - ref).max() < 1e-14 * np.abs(ref).max() * 100
    t = _Tables(128, 64, 12.0, 10.0, 0.7, 1.0, 0.005, potential="0.5*pow(x,2)", dpotential="x", scheme="lie",
                D_theta=0, D_lambda=0, gamma=0.0175, damping="ODM", fp_damping=" p*p/sqrt( p*p + 0.5  ) ",
                gp_coeff=0, gp_rho_factor=None, mask=False, truncate_constants=True, time_rows=None,
                pot_table=None, L_material=3.0)
    for k, name in enumerate(["gp", "gm", "gsp", "gsm"]):
        np.testing.assert_array_equal(t.odm_g[k], np.broadcast_to(T[name], t.odm_g[k].shape))


# ----------------------------------------------------------------------------- KvN Fokker-Planck
def test_kvn_potential_step_is_the_classical_momentum_kick():
    """exp_PotentialKvN (FFT.py:27-75) multiplies the theta picture by exp(i dt V'(x) theta):
    the state becomes W(x, p + dt V'(x)), the kick p -> p - dt V'(x) of every classical
    trajectory.  For a Gaussian the kicked state is known in closed form; a wrong sign, a
    factor 2 or hbar (the Bopp difference's expansion) fails it."""
    X = P = 128
    dt = 0.01
    dV = lambda t, x: 0.4 * x ** 3 - 2.0 * x  # noqa: E731  (V = 0.1 x^4 - x^2)
    orc = ow.Wigner2DOracle(X, P, 8.0, 8.0, 0.7, 1.0, dt, potential=lambda t, x: 0.1 * x ** 4 - x ** 2,
                            dpotential=dV, mask=False, kvn=True)
    g = orc.g
    x0, p0, s = 0.5, 0.3, 1.0
    W0 = ow.gaussian_wigner(g, x0=x0, p0=p0, s=s).astype(np.complex128)
    W1 = orc.potential_step(W0, 0.0)
    x, p = g.x_host, g.p_host
    want = (1.0 / np.pi) * np.exp(-((x - x0) ** 2) / s ** 2 - s ** 2 * (p + dt * dV(0, x) - p0) ** 2)
    assert np.abs(W1 - want).max() < 1e-12
    # the quantum (Bopp) step differs from the kick for a quartic V: the pin separates the two
    Wq = ow.Wigner2DOracle(X, P, 8.0, 8.0, 0.7, 1.0, dt, potential=lambda t, x: 0.1 * x ** 4 - x ** 2,
                           dpotential=dV, mask=False).potential_step(W0, 0.0)
    assert np.abs(Wq - want).max() > 1e-6


def test_kvn_equals_the_wigner_propagator_for_quadratic_v():
    """For quadratic V the Bopp difference is exactly -hbar theta V'(x): the Fokker-Planck run
    (damped, diffusive, renormalised, edge mask on) equals the Wigner run step for step."""
    X = P = 128
    V = lambda t, x: 0.5 * x * x  # noqa: E731
    dV = lambda t, x: x  # noqa: E731
    fp = lambda p: p * p / np.sqrt(p * p + 0.25)  # noqa: E731
    kw = dict(D_theta=0.0143, gamma=0.07, potential=V, dpotential=dV, fp_damping=fp, damping="CaldeiraLeggett",
              truncate_constants=True)
    a = ow.Wigner2DOracle(X, P, 10.0, 10.0, 1.0, 1.0, 0.005, kvn=True, **kw)
    b = ow.Wigner2DOracle(X, P, 10.0, 10.0, 1.0, 1.0, 0.005, **kw)
    W0 = ow.gaussian_wigner(a.g, x0=0.0, p0=3.0)
    Wa, oa, _ = a.run(W0, 30)
    Wb, ob, _ = b.run(W0, 30)
    assert rel_l2(Wa, Wb) < 1e-12
    np.testing.assert_allclose(oa["P"], ob["P"], rtol=1e-11, atol=1e-13)
