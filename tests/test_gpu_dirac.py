"""GPU parity of the fused Wigner-Dirac 4x4 step (libwignerhip wh_dirac_*, through its C ABI)
against the oracle (oracle/dirac4x4.py).

Bar: FP64 relative L2 <= 1e-10 in the 16 planes; observables within 1e-10 relative
(1e-12 absolute near 0).  FP32 planes (FP64 factors and sums): relative L2 <= 1e-5.
"""
import numpy as np
import pytest

from oracle.dirac4x4 import Dirac4x4Oracle, gaussian_spinor_particle_up, wigner_from_spinor

pytestmark = pytest.mark.gpu

KEYS = ["X", "P", "XP", "XX", "PP", "Alpha1", "D1V0", "XD1V0", "PD1V0", "PAlpha1", "XAlpha1"]


def zero(t, x):
    return 0.0 * x


CASES = {
    # Coulomb-like V0 of the hydrogen-like notebooks: tabulated (non-polynomial)
    "coulomb": ("-2/sqrt(x*x+1)", lambda t, x: -2.0 / np.sqrt(x * x + 1.0),
                lambda t, x: 2.0 * x / (x * x + 1.0) ** 1.5),
    # Klein step (instances/WignerDirac2D_4x4/Klein.ipynb): tabulated
    "klein": ("1.5*0.5*( 1. + tanh((x-5.)/0.25)   )", lambda t, x: 0.75 * (1.0 + np.tanh((x - 5.0) / 0.25)),
              lambda t, x: 0.75 / (0.25 * np.cosh((x - 5.0) / 0.25) ** 2)),
    # polynomial V0 (Horner path), linear ramp of Free_Caldeira-type runs
    "ramp": ("1.5*x + 0.01*pow(x,2)", lambda t, x: 1.5 * x + 0.01 * x * x, lambda t, x: 1.5 + 0.02 * x),
}


def make_pair(cuda, case, X=128, P=128, dt=0.01, precision="fp64", D_theta=0.0, gamma=0.0, absorb=True,
              energy=True):
    from pywignercuda_amd.dirac import WignerDirac4x4

    vs, vf, dvf = CASES[case]
    damping = "CaldeiraLegget" if gamma else None
    hip = WignerDirac4x4(X, P, 18.0, 16.0, 1.0, 1.0, dt, potentials=(vs, "0.", " 0.", "0.*x"), D_theta=D_theta,
                         gamma=gamma, damping=damping, absorb=absorb, energy=energy, precision=precision,
                         device=cuda)
    orc = Dirac4x4Oracle(X, P, 18.0, 16.0, 1.0, 1.0, dt, [vf, zero, zero, zero], dpotential0=dvf, D_theta=D_theta,
                         gamma=gamma, damping=damping, absorb=absorb, compute_energy=energy)
    W0 = wigner_from_spinor(lambda Xg: gaussian_spinor_particle_up(-2.0, 3.0, 1.0, Xg, 1.0, 1.0), orc)
    return hip, orc, W0


def run_both(hip, orc, W0, steps, use_graph=False):
    import torch

    W_ref, rec = orc.run(W0, steps)
    W = hip.to_device(W0)
    obs = hip.run(W, steps, first_step=1, use_graph=use_graph)
    torch.cuda.synchronize()
    return W.cpu().numpy().reshape(4, 4, hip.P, hip.X), obs.cpu().numpy(), W_ref, rec


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def check_obs(obs, rec, energy=True, rtol=1e-10, atol=1e-12):
    from pywignercuda_amd.dirac import WignerDirac4x4

    o = WignerDirac4x4.observables(obs)
    n = len(rec)
    for k in KEYS + (["energy"] if energy else []):
        ref = np.array([r[k] for r in rec])
        np.testing.assert_allclose(o[k][:n], ref, rtol=rtol, atol=atol, err_msg=k)
    for k in ["negativity", "transmission", "antiparticle"]:
        ref = np.array([r[k] for r in rec])
        np.testing.assert_allclose(o[k][:n], ref, rtol=rtol, atol=atol, err_msg=k)


@pytest.mark.parametrize("case", ["coulomb", "klein", "ramp"])
def test_dirac_fp64_matches_oracle(cuda, case):
    hip, orc, W0 = make_pair(cuda, case)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 25)
    err = rel_l2(W, W_ref)
    print(f"{case}: rel L2 {err:.3e}")
    assert err < 1e-10
    check_obs(obs, rec)
    # the final row is the normalised final state
    assert abs(obs[-1, 27] - 1.0) < 1e-12


def test_dirac_rectangular_damped(cuda):
    """256 x 128 grid, D_Theta decay and the Caldeira-Leggett order-2 dissipator."""
    hip, orc, W0 = make_pair(cuda, "klein", X=256, P=128, D_theta=0.02, gamma=0.05)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 15)
    assert rel_l2(W, W_ref) < 1e-10
    check_obs(obs, rec)


def test_dirac_graph_replay_matches_eager(cuda):
    import torch

    hip, orc, W0 = make_pair(cuda, "coulomb")
    Wa = hip.to_device(W0)
    oa = hip.run(Wa, 37, use_graph=False)
    for _ in range(3):
        Wb = hip.to_device(W0)
        ob = hip.run(Wb, 37, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(Wa, Wb)
        assert torch.equal(oa, ob)


@pytest.mark.parametrize("precision,bound", [("fp32", 2e-6), ("fp32fast", 2e-5)])
def test_dirac_fp32_error(cuda, precision, bound):
    hip, orc, W0 = make_pair(cuda, "coulomb", precision=precision, X=256, P=256)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 100)
    err = rel_l2(W, W_ref)
    print(f"{precision} dirac: rel L2 {err:.3e}")
    assert err < bound


def test_dropin_dirac_class_free_notebook_setup(cuda, tmp_path):
    """instances/WignerDirac2D_4x4/Free.ipynb cell 2 shape (smaller grid and fewer steps):
    subclass sets the strings, builds W_init from GaussianSpinor_ParticleUp at x -/+ theta/2,
    Run() fills the reference's result attributes; checked against the oracle."""
    from pywignercuda_amd.GPU_WignerDirac2D_4x4 import GPU_WignerDirac2D_4x4

    class Klein(GPU_WignerDirac2D_4x4):
        def __init__(self):
            self.gammaDamping = 0.0
            self.D_Theta = 0.0
            self.D_Lambda = 0.0
            self.grossPitaevskiiCoefficient = 0.0
            self.Potential_0_String = "{0}*0.5*( 1. + tanh((x-5.)/{1})   )".format(1.5, 0.25)
            self.Potential_1_String = " 0."
            self.Potential_2_String = " 0."
            self.Potential_3_String = " 0."
            GPU_WignerDirac2D_4x4.__init__(self, 128, 128, 18, 16, 1.0, 1.0, 0.01, 20, 10,
                                           frameSaveMode="Density", antiParticleNorm=True, computeEnergy=True)

    inst = Klein()
    inst.fileName = str(tmp_path / "Free.hdf5")
    W = np.empty([4, 4, 128, 128], dtype=np.complex128)
    psiL = inst.GaussianSpinor_ParticleUp(-2, 6.0, 1, inst.X - 0.5 * inst.Theta)
    psiR = inst.GaussianSpinor_ParticleUp(-2, 6.0, 1, inst.X + 0.5 * inst.Theta)
    for i in range(4):
        for j in range(4):
            W[i, j][:, :] = psiL[i] * psiR[j].conj()
    inst.Fourier_4X4_Theta_To_P(W)
    W *= 1.0 / inst.Wigner_4x4_Norm(W)
    inst.W_init = W
    assert inst.Run() == 0
    orc = Dirac4x4Oracle(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01,
                         [lambda t, x: 0.75 * (1.0 + np.tanh((x - 5.0) / 0.25)), zero, zero, zero],
                         dpotential0=lambda t, x: 0.75 / (0.25 * np.cosh((x - 5.0) / 0.25) ** 2),
                         compute_energy=True)
    W_ref, rec = orc.run(W, 20)
    assert rel_l2(inst.W_end, W_ref) < 1e-10
    np.testing.assert_allclose(inst.X_Average, [np.real(r["X"]) for r in rec], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(inst.Dirac_energy, [np.real(r["energy"]) for r in rec], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(inst.antiParticle_population, [r["antiparticle"] for r in rec], rtol=1e-9,
                               atol=1e-13)
    assert list(inst.timeSavedIndexRange) == [0, 10, 20]
    assert inst.TakabayashiAngle_end.shape == (128, 128)


def test_dirac_mass_potential_variant(cuda):
    """GPU_WignerDirac2D_MassPotential_4x4: Mass(x) in the potential factors (tabulated phases),
    mass_half = 1e-4 in the kinetic factors, bare mass in the projectors."""
    from pywignercuda_amd.dirac import WignerDirac4x4

    mass_s = "0.5 + 0.4*tanh(x/2.)"
    mass_f = lambda t, x: 0.5 + 0.4 * np.tanh(x / 2.0)  # noqa: E731
    vs, vf, dvf = CASES["klein"]
    hip = WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=(vs, "0.", "0.", "0."), energy=True,
                         mass_potential=mass_s, kinetic_mass_half=0.0001, device=cuda)
    orc = Dirac4x4Oracle(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, [vf, zero, zero, zero], dpotential0=dvf,
                         compute_energy=True, mass_potential=mass_f, kinetic_mass_half=0.0001)
    W0 = wigner_from_spinor(lambda Xg: gaussian_spinor_particle_up(-2.0, 3.0, 1.0, Xg, 1.0, 1.0), orc)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 20)
    assert rel_l2(W, W_ref) < 1e-10
    check_obs(obs, rec)


def test_dirac_mass_potential_with_vector_potential(cuda):
    """Mass(x) together with A: the two potential sweeps take the mass term of each cell from the
    mass tables, F(x -/+ theta/2) (D44M.py:387-427 with Potential_1..3 of D44.py:417-442)."""
    from pywignercuda_amd.dirac import WignerDirac4x4

    mass_s = "0.5 + 0.4*tanh(x/2.)"
    mass_f = lambda t, x: 0.5 + 0.4 * np.tanh(x / 2.0)  # noqa: E731
    vs, vf, dvf = CASES["ramp"]
    pots = (vs, "0.05*x", " -x ", "0.1 + 0.02*pow(x,2)")
    hip = WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=pots, energy=True,
                         mass_potential=mass_s, kinetic_mass_half=0.0001, device=cuda)
    orc = Dirac4x4Oracle(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01,
                         [vf, lambda t, x: 0.05 * x, lambda t, x: -x, lambda t, x: 0.1 + 0.02 * x * x],
                         dpotential0=dvf, compute_energy=True, mass_potential=mass_f, kinetic_mass_half=0.0001)
    W0 = wigner_from_spinor(lambda Xg: gaussian_spinor_particle_up(-2.0, 3.0, 1.0, Xg, 1.0, 1.0), orc)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 20)
    err = rel_l2(W, W_ref)
    print(f"mass + A: rel L2 {err:.3e}")
    assert err < 1e-10
    check_obs(obs, rec)


@pytest.mark.parametrize("with_a", [False, True])
def test_dirac_time_dependent_mass(cuda, with_a):
    """Mass(t, x) that depends on t (MassPotential variant, D44M.py:387-427): the device refills
    the mass tables F(x -/+ theta/2) before each potential sweep from the program, with and
    without a vector potential, against the oracle."""
    from pywignercuda_amd.dirac import WignerDirac4x4

    mass_s = "0.5 + 0.4*tanh(x/2.)*cos(3*t)"
    mass_f = lambda t, x: 0.5 + 0.4 * np.tanh(x / 2.0) * np.cos(3 * t)  # noqa: E731
    vs, vf, dvf = CASES["ramp"]
    a_s = ("0.05*x", " -x ", "0.1 + 0.02*pow(x,2)") if with_a else ("0.", "0.", "0.")
    a_f = [lambda t, x: 0.05 * x, lambda t, x: -x, lambda t, x: 0.1 + 0.02 * x * x] if with_a else [zero] * 3
    hip = WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=(vs,) + a_s, energy=True,
                         mass_potential=mass_s, kinetic_mass_half=0.0001, device=cuda)
    orc = Dirac4x4Oracle(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, [vf] + a_f, dpotential0=dvf, compute_energy=True,
                         mass_potential=mass_f, kinetic_mass_half=0.0001)
    W0 = wigner_from_spinor(lambda Xg: gaussian_spinor_particle_up(-2.0, 3.0, 1.0, Xg, 1.0, 1.0), orc)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 20)
    err = rel_l2(W, W_ref)
    print(f"time-dependent mass (A: {with_a}): rel L2 {err:.3e}")
    assert err < 1e-10
    check_obs(obs, rec)


def test_dirac_gross_pitaevskii(cuda):
    """ConstantB_GP0.02-style nonlinearity (A = 0 here): phase dt a Re(B- - B+) from the trace density."""
    from pywignercuda_amd.dirac import WignerDirac4x4

    vs, vf, dvf = CASES["klein"]
    hip = WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=(vs, "0.", "0.", "0."), gp_coeff=0.5,
                         device=cuda)
    orc = Dirac4x4Oracle(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, [vf, zero, zero, zero], dpotential0=dvf,
                         gp_coeff=0.5)
    W0 = wigner_from_spinor(lambda Xg: gaussian_spinor_particle_up(-2.0, 3.0, 1.0, Xg, 1.0, 1.0), orc)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 20)
    assert rel_l2(W, W_ref) < 1e-10
    check_obs(obs, rec, energy=False)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32fast"])
def test_dirac_vector_potential_constant_field(cuda, precision):
    """instances/WignerDirac2D_4x4/ConstantB.ipynb form: Potential_2 = -x (a constant field) plus
    a small A1, A3 so that every coupling of the 4x4 factors is exercised; two p sweeps."""
    from pywignercuda_amd.dirac import WignerDirac4x4

    vs, vf, dvf = CASES["klein"]
    pots = (vs, "0.05*x", " -x ", "0.1 + 0.02*pow(x,2)")
    hip = WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=pots, energy=True, D_theta=0.01,
                         precision=precision, device=cuda)
    assert hip.vector_potential
    orc = Dirac4x4Oracle(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01,
                         [vf, lambda t, x: 0.05 * x, lambda t, x: -x, lambda t, x: 0.1 + 0.02 * x * x],
                         dpotential0=dvf, compute_energy=True, D_theta=0.01)
    W0 = wigner_from_spinor(lambda Xg: gaussian_spinor_particle_up(-2.0, 3.0, 1.0, Xg, 1.0, 1.0), orc)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 20)
    err = rel_l2(W, W_ref)
    print(f"vector potential {precision}: rel L2 {err:.3e}")
    if precision == "fp64":
        assert err < 1e-10
        check_obs(obs, rec)
    else:
        assert err < 1e-5


@pytest.mark.parametrize("case,precision", [("klein", "fp64"), ("ramp", "fp64"), ("ramp", "fp32")])
def test_dirac_time_dependent_vector_potential(cuda, case, precision):
    """Vector potentials whose polynomial coefficients depend on t (a driving field): the two
    potential sweeps read the coefficient row of the step's time index, the energy record adds
    A1..A3 at the state's time (D44.py:417-442, 575-601, 2801-2804); V0 tabulated (klein) or
    polynomial (ramp)."""
    from pywignercuda_amd.dirac import WignerDirac4x4

    vs, vf, dvf = CASES[case]
    pots = (vs, "0.05*x*cos(3*t)", " -x + 0.2*sin(2*t) ", "0.1*t + 0.02*pow(x,2)")
    hip = WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=pots, energy=True, D_theta=0.01,
                         precision=precision, time_rows=64, device=cuda)
    assert hip.vector_potential and hip.a_table is not None
    orc = Dirac4x4Oracle(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01,
                         [vf, lambda t, x: 0.05 * x * np.cos(3 * t), lambda t, x: -x + 0.2 * np.sin(2 * t),
                          lambda t, x: 0.1 * t + 0.02 * x * x],
                         dpotential0=dvf, compute_energy=True, D_theta=0.01)
    W0 = wigner_from_spinor(lambda Xg: gaussian_spinor_particle_up(-2.0, 3.0, 1.0, Xg, 1.0, 1.0), orc)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 20)
    err = rel_l2(W, W_ref)
    print(f"time-dependent A {case} {precision}: rel L2 {err:.3e}")
    if precision == "fp64":
        assert err < 1e-10
        check_obs(obs, rec)
    else:
        assert err < 1e-5
    with pytest.raises(ValueError):  # beyond the coefficient rows
        hip.run(hip.to_device(W0), 80)


@pytest.mark.parametrize("with_a", [False, True])
def test_dirac_time_dependent_expression_v0(cuda, with_a):
    """A time-dependent non-polynomial V0(t, x) (a modulated Klein step): the device tabulates
    its Bopp difference before each potential sweep and dV0/dx, V0 before each record row,
    from the programs of V0 and of its sympy derivative, against the oracle (FP64)."""
    from pywignercuda_amd.dirac import WignerDirac4x4

    vs = "1.5*0.5*( 1. + tanh((x-5.)/0.25) )*(1. + 0.3*sin(2*t))"
    vf = lambda t, x: 0.75 * (1.0 + np.tanh((x - 5.0) / 0.25)) * (1.0 + 0.3 * np.sin(2 * t))  # noqa: E731
    dvf = lambda t, x: 0.75 / (0.25 * np.cosh((x - 5.0) / 0.25) ** 2) * (1.0 + 0.3 * np.sin(2 * t))  # noqa: E731
    a_s = ("0.05*x", " -x ", "0.1 + 0.02*pow(x,2)") if with_a else ("0.", "0.", "0.")
    a_f = [lambda t, x: 0.05 * x, lambda t, x: -x, lambda t, x: 0.1 + 0.02 * x * x] if with_a else [zero] * 3
    hip = WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=(vs,) + a_s, energy=True, D_theta=0.01,
                         device=cuda)
    orc = Dirac4x4Oracle(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, [vf] + a_f, dpotential0=dvf, compute_energy=True,
                         D_theta=0.01)
    W0 = wigner_from_spinor(lambda Xg: gaussian_spinor_particle_up(-2.0, 3.0, 1.0, Xg, 1.0, 1.0), orc)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 20)
    err = rel_l2(W, W_ref)
    print(f"time-dependent V0 program (A: {with_a}): rel L2 {err:.3e}")
    assert err < 1e-10
    check_obs(obs, rec)


@pytest.mark.parametrize("case,mass", [("klein", False), ("ramp", True)])
def test_dirac_non_polynomial_vector_potential(cuda, case, mass):
    """Vector potentials that are not polynomials in x (a localised, driven A1 and a saturating
    A2): their programs are tabulated at x -/+ theta/2 before each potential sweep and summed
    at x for the energy record (D44.py:417-442, 575-601, 2801-2804), against the oracle; with
    a position-dependent mass in the second case."""
    from pywignercuda_amd.dirac import WignerDirac4x4

    vs, vf, dvf = CASES[case]
    pots = (vs, "0.3*exp(-x*x/8.)*cos(2*t)", " -2*tanh(x/4.) ", "0.")
    extra_h = dict(mass_potential="1. + 0.1*x*x", kinetic_mass_half=0.0001) if mass else {}
    extra_o = dict(mass_potential=lambda t, x: 1.0 + 0.1 * x * x, kinetic_mass_half=0.0001) if mass else {}
    hip = WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=pots, energy=True, D_theta=0.01,
                         device=cuda, **extra_h)
    assert hip.vector_potential
    orc = Dirac4x4Oracle(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01,
                         [vf, lambda t, x: 0.3 * np.exp(-x * x / 8.0) * np.cos(2 * t),
                          lambda t, x: -2.0 * np.tanh(x / 4.0), zero],
                         dpotential0=dvf, compute_energy=True, D_theta=0.01, **extra_o)
    W0 = wigner_from_spinor(lambda Xg: gaussian_spinor_particle_up(-2.0, 3.0, 1.0, Xg, 1.0, 1.0), orc)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 20)
    err = rel_l2(W, W_ref)
    print(f"non-polynomial A ({case}, mass {mass}): rel L2 {err:.3e}")
    assert err < 1e-10
    check_obs(obs, rec)


def test_dirac_single_sweeps_match_oracle_factors(cuda):
    """wh_dirac_kinetic / wh_dirac_potential (one sweep each, no normalisation) against the
    oracle's kinetic and potential steps between the transforms."""
    import torch

    hip, orc, W0 = make_pair(cuda, "coulomb", D_theta=0.02)
    W = hip.to_device(W0)
    hip.kinetic(W)
    torch.cuda.synchronize()
    ref = orc.lambda_to_x(orc.kinetic_step(orc.x_to_lambda(W0)))
    assert rel_l2(W.cpu().numpy().reshape(4, 4, hip.P, hip.X), ref) < 1e-12
    hip.potential(W, step=3)
    torch.cuda.synchronize()
    ref = orc.theta_to_p(orc.potential_step(orc.p_to_theta(ref), orc.dt * 3)) * orc.absorber()
    assert rel_l2(W.cpu().numpy().reshape(4, 4, hip.P, hip.X), ref) < 1e-12


# ---------------------------------------------------------------------------- ODM damping
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("mass_potential", [False, True])
def test_dirac_odm_matches_oracle(cuda, precision, mass_potential):
    """dampingModel 'ODM' (DampingODM_Kernel, GPU_WignerDirac2D_4x4.py:952-1195): the quad-mixing
    p sweep against the oracle, Klein step, D_Theta and absorber on, gamma = 0.05, lambdaBar = 4."""
    from pywignercuda_amd.dirac import WignerDirac4x4

    vs, vf, dvf = CASES["klein"]
    extra_h = dict(mass_potential="1. + 0.1*x*x", kinetic_mass_half=0.0001) if mass_potential else {}
    extra_o = dict(mass_potential=lambda t, x: 1.0 + 0.1 * x * x, kinetic_mass_half=0.0001) if mass_potential else {}
    hip = WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=(vs, "0.", "0.", "0."), D_theta=0.001,
                         gamma=0.05, damping="ODM", lambda_bar=4.0, energy=True, precision=precision, device=cuda,
                         **extra_h)
    orc = Dirac4x4Oracle(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, [vf, zero, zero, zero], dpotential0=dvf,
                         D_theta=0.001, gamma=0.05, damping="ODM", lambda_bar=4.0, compute_energy=True, **extra_o)
    W0 = wigner_from_spinor(lambda Xg: gaussian_spinor_particle_up(-2.0, 3.0, 1.0, Xg, 1.0, 1.0), orc)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 25)
    err = rel_l2(W, W_ref)
    print(f"Dirac ODM {precision} mass_potential={mass_potential}: rel L2 {err:.3e}")
    if precision == "fp64":
        assert err < 1e-10
        check_obs(obs, rec)
    else:
        assert err < 1e-5


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dirac_odm_with_vector_potential(cuda, precision):
    """ODM damping together with a vector potential (and a position-dependent mass): the two
    factor sweeps leave the planes in the theta picture and the ODM sweep mixes the quads and
    transforms back (D44.py:3576-3615), against the oracle."""
    from pywignercuda_amd.dirac import WignerDirac4x4

    vs, vf, dvf = CASES["klein"]
    pots = (vs, "0.05*x", " -x ", "0.1 + 0.02*pow(x,2)")
    hip = WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=pots, D_theta=0.001, gamma=0.05,
                         damping="ODM", lambda_bar=4.0, energy=True, mass_potential="1. + 0.1*x*x",
                         kinetic_mass_half=0.0001, precision=precision, device=cuda)
    orc = Dirac4x4Oracle(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01,
                         [vf, lambda t, x: 0.05 * x, lambda t, x: -x, lambda t, x: 0.1 + 0.02 * x * x],
                         dpotential0=dvf, D_theta=0.001, gamma=0.05, damping="ODM", lambda_bar=4.0,
                         compute_energy=True, mass_potential=lambda t, x: 1.0 + 0.1 * x * x,
                         kinetic_mass_half=0.0001)
    W0 = wigner_from_spinor(lambda Xg: gaussian_spinor_particle_up(-2.0, 3.0, 1.0, Xg, 1.0, 1.0), orc)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 25)
    err = rel_l2(W, W_ref)
    print(f"Dirac ODM + A {precision}: rel L2 {err:.3e}")
    if precision == "fp64":
        assert err < 1e-10
        check_obs(obs, rec)
    else:
        assert err < 1e-5


def _odm_notebook(tmp_path, V0, steps, pX, D_Theta, epsilon):
    from pywignercuda_amd.GPU_WignerDirac2D_4x4 import GPU_WignerDirac2D_4x4

    class Klein(GPU_WignerDirac2D_4x4):
        def __init__(self):
            self.gammaDamping = 0.01
            self.fp_Damping_String = " p*p/sqrt( p*p + {epsilon}  ) ".format(epsilon=epsilon)
            self.dampingModel = "ODM"
            self.lambdaBar = 4.0
            self.D_Theta = D_Theta
            self.D_Lambda = 0.0
            self.grossPitaevskiiCoefficient = 0.0
            self.Potential_0_String = "{0}*0.5*( 1. + tanh((x-5.)/{1})   )".format(V0, 0.25)
            self.Potential_1_String = " 0."
            self.Potential_2_String = " 0."
            self.Potential_3_String = " 0."
            GPU_WignerDirac2D_4x4.__init__(self, 512, 512, 18, 16, 1.0, 1.0, 0.01, steps, 100,
                                           frameSaveMode="Density", antiParticleNorm=True, computeEnergy=True)

    inst = Klein()
    inst.fileName = str(tmp_path / "Klein.hdf5")
    W = np.empty([4, 4, 512, 512], dtype=np.complex128)
    psiL = inst.GaussianSpinor_ParticleUp(-2, pX, 1, inst.X - 0.5 * inst.Theta)
    psiR = inst.GaussianSpinor_ParticleUp(-2, pX, 1, inst.X + 0.5 * inst.Theta)
    for i in range(4):
        for j in range(4):
            W[i, j][:, :] = psiL[i] * psiR[j].conj()
    inst.Fourier_4X4_Theta_To_P(W)
    W *= 1.0 / inst.Wigner_4x4_Norm(W)
    inst.W_init = W
    assert inst.Run() == 0
    W0 = np.fft.fftshift(inst.Wigner_4X4__SpinTrace(10 * inst.W_end).real)
    return float(W0.min()), float(W0.max())


@pytest.mark.parametrize("nb", ["Free_ODM0.01", "Klein_ODM01"])
def test_dirac_odm_notebooks(cuda, tmp_path, nb):
    """instances/WignerDirac2D_4x4/Free_ODM0.01.ipynb (V0 = 0, px = 6, 1000 steps) and
    Klein_ODM01.ipynb (V0 = 9, px = 7, D_Theta = 0.001, 1300 steps), 512^2, gamma = 0.01,
    lambdaBar = 4: min / max of the spin trace of 10 W_end (cell 8 of each notebook)."""
    if nb == "Free_ODM0.01":
        lo, hi = _odm_notebook(tmp_path, 0.0, 1000, 6.0, 0.0, 0.01)
        rec = (-0.000151901096206, 2.79279961979)
    else:
        lo, hi = _odm_notebook(tmp_path, 9.0, 1300, 7.0, 0.001, 0.5)
        rec = (-0.0852455133993, 0.945834335445)
    print(f"{nb}: min {lo!r} (recorded {rec[0]}), max {hi!r} (recorded {rec[1]})")
    assert abs(hi - rec[1]) < 1e-9 * abs(rec[1])
    assert abs(lo - rec[0]) < 1e-6 * abs(rec[0])


# ---------------------------------------------------------------------------- step filter, KvN
@pytest.mark.parametrize("case", ["klein", "ramp"])
def test_dirac_step_filter_matches_oracle(cuda, case):
    """antiParticleStepFiltering (GPU_WignerDirac2D_4x4.py:3636-3678): the filter sweep (projectors at
    p -/+ lambda/2), the absorber after it and the column sums of the filtered state."""
    from pywignercuda_amd.dirac import WignerDirac4x4

    vs, vf, dvf = CASES[case]
    hip = WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=(vs, "0.", "0.", "0."), D_theta=0.01,
                         energy=True, step_filter=True, device=cuda)
    orc = Dirac4x4Oracle(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, [vf, zero, zero, zero], dpotential0=dvf,
                         D_theta=0.01, compute_energy=True, step_filter=True)
    W0 = wigner_from_spinor(lambda Xg: gaussian_spinor_particle_up(-2.0, 3.0, 1.0, Xg, 1.0, 1.0), orc)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 20)
    assert rel_l2(W, W_ref) < 1e-10
    check_obs(obs, rec)


@pytest.mark.parametrize("case", ["ramp", "klein"])
def test_kvn_matches_oracle(cuda, case):
    """GPU_WignerDiracKvN2D_4x4: kinetic factors at +/- lambda/2 without the mass diagonal, the
    potential step e^{-dt D theta^2 + i dt V0' theta}, the filter at lambda = 0 every step."""
    from pywignercuda_amd.dirac import WignerDirac4x4

    vs, vf, dvf = CASES[case]
    hip = WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=(vs, "0.", "0.", "0."), D_theta=0.01,
                         energy=True, kvn=True, device=cuda)
    orc = Dirac4x4Oracle(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, [vf, zero, zero, zero], dpotential0=dvf,
                         D_theta=0.01, compute_energy=True, kvn=True)
    W0 = orc.filter_particles(
        wigner_from_spinor(lambda Xg: gaussian_spinor_particle_up(-2.0, 3.0, 1.0, Xg, 1.0, 1.0), orc))
    W0 = W0 / orc.norm(W0)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 20)
    err = rel_l2(W, W_ref)
    print(f"KvN {case}: rel L2 {err:.3e}")
    assert err < 1e-10
    check_obs(obs, rec)


def test_kvn_time_dependent_expression_v0(cuda):
    """GPU_WignerDiracKvN2D_4x4 with a time-dependent non-polynomial V0 (a modulated Klein step):
    the device evaluates the program of dV0/dx at each potential sweep's time index for the
    KvN phase e^{i dt V0'(t, x) theta} (KvN_XTheta_Kernel, D44K.py:356-420), against the oracle."""
    from pywignercuda_amd.dirac import WignerDirac4x4

    vs = "1.5*0.5*( 1. + tanh((x-5.)/0.25) )*(1. + 0.3*sin(2*t))"
    vf = lambda t, x: 0.75 * (1.0 + np.tanh((x - 5.0) / 0.25)) * (1.0 + 0.3 * np.sin(2 * t))  # noqa: E731
    dvf = lambda t, x: 0.75 / (0.25 * np.cosh((x - 5.0) / 0.25) ** 2) * (1.0 + 0.3 * np.sin(2 * t))  # noqa: E731
    hip = WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=(vs, "0.", "0.", "0."), D_theta=0.01,
                         energy=True, kvn=True, device=cuda)
    orc = Dirac4x4Oracle(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, [vf, zero, zero, zero], dpotential0=dvf,
                         D_theta=0.01, compute_energy=True, kvn=True)
    W0 = orc.filter_particles(
        wigner_from_spinor(lambda Xg: gaussian_spinor_particle_up(-2.0, 3.0, 1.0, Xg, 1.0, 1.0), orc))
    W0 = W0 / orc.norm(W0)
    W, obs, W_ref, rec = run_both(hip, orc, W0, 20)
    err = rel_l2(W, W_ref)
    print(f"KvN time-dependent V0 program: rel L2 {err:.3e}")
    assert err < 1e-10
    check_obs(obs, rec)


def test_kvn_x1_notebook(cuda, tmp_path):
    """instances/WignerDiracKvN2D_4x4/X1.ipynb: 512^2, amplitudes 14, m = c = 1, dt = 0.01, 800 steps,
    V0 = 1.5 x, spinor at x0 = -1, px = 5, s = 1 filtered to positive energy (FilterElectrons) and
    normalised; cells 9-10: min / max of the spin trace of 10 W_init and 10 W_end."""
    from pywignercuda_amd.GPU_WignerDirac2D_4x4 import GPU_WignerDiracKvN2D_4x4

    class Klein(GPU_WignerDiracKvN2D_4x4):
        def __init__(self):
            self.gammaDamping = 0.0
            self.D_Theta = 0.0
            self.D_Lambda = 0.0
            self.Potential_0_String = "1.5*x"
            self.Potential_1_String = " 0.*x"
            self.Potential_2_String = " 0.*x"
            self.Potential_3_String = " 0.*x"
            GPU_WignerDiracKvN2D_4x4.__init__(self, 512, 512, 14, 14, 1.0, 1.0, 0.01, 800, 10,
                                              frameSaveMode="Density", antiParticleNorm=True, computeEnergy=True)

    inst = Klein()
    inst.fileName = str(tmp_path / "X1.hdf5")
    W = np.empty([4, 4, 512, 512], dtype=np.complex128)
    psiL = inst.GaussianSpinor_ParticleUp(-1, 5, 1.0, inst.X - 0.5 * inst.Theta)
    psiR = inst.GaussianSpinor_ParticleUp(-1, 5, 1.0, inst.X + 0.5 * inst.Theta)
    for i in range(4):
        for j in range(4):
            W[i, j][:, :] = psiL[i] * psiR[j].conj()
    inst.Fourier_4X4_Theta_To_P(W)
    inst.FilterElectrons(W, 1)
    W *= 1.0 / inst.Wigner_4x4_Norm(W)
    inst.W_init = W
    Wi = np.fft.fftshift(inst.Wigner_4X4__SpinTrace(10 * W).real)
    assert abs(Wi.max() - 3.17920519628) < 1e-9 * 3.17920519628
    assert inst.Run() == 0
    We = np.fft.fftshift(inst.Wigner_4X4__SpinTrace(10 * inst.W_end).real)
    print(f"KvN X1: W_end min {We.min()!r} max {We.max()!r}")
    assert abs(We.max() - 3.15212339853) < 1e-9 * 3.15212339853
    assert abs(We.min() - (-7.04369343969e-09)) < 1e-3 * 7.04369343969e-09
