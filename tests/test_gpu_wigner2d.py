"""GPU parity of the fused Wigner2D step (libwignerhip, through its C ABI) against the oracle.

Bar (north star): the FP64 form matches the complex128 oracle within relative L2
1e-10 in W; observables within 1e-10 relative (absolute 1e-12 for values near 0).
FP32-bulk form: relative L2 <= 1e-5 (errors stated per observable in the test).
"""
import numpy as np
import pytest

from oracle import wigner2d as ow

pytestmark = pytest.mark.gpu

OBS_KEYS = ["X", "X2", "P", "P2", "XP", "dV", "PdV", "XdV", "H", "neg", "norm"]
OBS_COL = {k: i for i, k in enumerate(OBS_KEYS)}


def V_quartic(t, x):
    return 0.5 * np.power(x, 2) + 0.1 * np.power(x, 4)


def dV_quartic(t, x):
    return x + 0.4 * np.power(x, 3)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def make_pair(cuda, X=256, P=256, amp=10.0, dt=0.01, scheme="lie", mask=True, truncate=False,
              D_theta=0.0, D_lambda=0.0, gamma=0.0, damping=None, fp_str=None, fp_fn=None, gp=0.0,
              precision="fp64", pot=("0.5*pow(x,2)+0.1*pow(x,4)", "x+0.4*pow(x,3)", V_quartic, dV_quartic),
              L_material=None, hbar=1.0, odm_weights=None):
    from pywignercuda_amd.wigner2d import Wigner2D

    vs, dvs, vf, dvf = pot
    hip = Wigner2D(X, P, amp, amp, hbar, 1.0, dt, potential=vs, dpotential=dvs,
                   kinetic="0.5*p*p" if scheme == "strang" else None, scheme=scheme,
                   D_theta=D_theta, D_lambda=D_lambda, gamma=gamma, damping=damping,
                   fp_damping=fp_str, gp_coeff=gp, mask=mask, truncate_constants=truncate,
                   precision=precision, device=cuda, L_material=L_material, odm_weights=odm_weights)
    orc = ow.Wigner2DOracle(X, P, amp, amp, hbar, 1.0, dt, D_theta=D_theta, D_lambda=D_lambda,
                            gamma=gamma, potential=vf, dpotential=dvf,
                            kinetic=(lambda p: 0.5 * p * p), fp_damping=fp_fn, damping=damping,
                            gp_coeff=gp, scheme=scheme, mask=mask, truncate_constants=truncate,
                            L_material=L_material, odm_weights=odm_weights)
    return hip, orc


def run_both(hip, orc, W0, steps, lie=True):
    import torch

    W_ref, obs_ref, extras = orc.run(W0, steps)
    W = hip.to_device(W0)
    obs = hip.run(W, steps + 1, use_graph=False)
    torch.cuda.synchronize()
    return W.cpu().numpy(), obs.cpu().numpy()[: steps + 1], W_ref, obs_ref, extras


def check_obs(obs, obs_ref, rtol=1e-10, atol=1e-12):
    for k in OBS_KEYS:
        np.testing.assert_allclose(obs[:, OBS_COL[k]], obs_ref[k], rtol=rtol, atol=atol, err_msg=k)


@pytest.mark.parametrize("truncate,mask", [(False, False), (True, True)])
def test_lie_quartic_fp64(cuda, truncate, mask):
    hip, orc = make_pair(cuda, truncate=truncate, mask=mask)
    W0 = ow.gaussian_wigner(orc.g, x0=2.0, p0=0.0, s=1.0)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 40)
    assert rel_l2(W, W_ref) < 1e-10
    check_obs(obs, obs_ref)


def test_lie_rectangular_grid(cuda):
    hip, orc = make_pair(cuda, X=512, P=128, mask=True)
    W0 = ow.gaussian_wigner(orc.g, x0=-1.0, p0=1.0, s=1.2)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 20)
    assert rel_l2(W, W_ref) < 1e-10
    check_obs(obs, obs_ref)


def test_strang_quartic_fp64(cuda):
    hip, orc = make_pair(cuda, scheme="strang", mask=True)
    W0 = ow.gaussian_wigner(orc.g, x0=2.0, p0=0.0, s=1.0)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 40)
    assert rel_l2(W, W_ref) < 1e-10
    check_obs(obs, obs_ref)


def test_lie_caldeira_leggett_fp64(cuda):
    fp_str = " p*p/sqrt( p*p + 0.5  ) "
    fp_fn = lambda p: p * p / np.sqrt(p * p + 0.5)  # noqa: E731
    hip, orc = make_pair(cuda, amp=12.0, dt=0.005, D_theta=0.0143, gamma=0.07,
                         damping="CaldeiraLeggett", fp_str=fp_str, fp_fn=fp_fn, truncate=True)
    W0 = ow.harmonic_oscillator_wigner(orc.g, 0, 1.0, 0.0, 3.0)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 30)
    assert rel_l2(W, W_ref) < 1e-10
    check_obs(obs, obs_ref)


def test_caldeira_leggett_every_variant_fp64(cuda):
    """Every Caldeira-Leggett p-sweep variant (radix, lanes, table or register twiddles)
    against the oracle, 256^2, 15 Lie steps."""
    import torch

    from pywignercuda_amd import _lib

    fp_fn = lambda p: np.abs(p)  # noqa: E731
    hip, orc = make_pair(cuda, amp=12.0, dt=0.005, D_theta=0.1, gamma=0.1,
                         damping="CaldeiraLeggett", fp_str="fabs(p)", fp_fn=fp_fn)
    W0 = ow.cat_state_wigner(orc.g, a=2.0)
    W_ref, obs_ref, _ = orc.run(W0, 15)
    n = _lib.lib().wh_variant_info(2, -1, None, None)
    tried = 0
    for v in range(n):
        try:
            hip.set_variants(-1, -1, v)
        except _lib.WignerHipInvalidValue:
            continue
        W = hip.to_device(W0)
        obs = hip.run(W, 16, use_graph=False).cpu().numpy()[:16]
        torch.cuda.synchronize()
        assert rel_l2(W.cpu().numpy(), W_ref) < 1e-10, v
        check_obs(obs, obs_ref)
        tried += 1
    assert tried == n


def test_strang_caldeira_leggett_fp64(cuda):
    fp_fn = lambda p: np.abs(p)  # noqa: E731
    hip, orc = make_pair(cuda, scheme="strang", amp=12.0, dt=0.005, D_theta=0.1, gamma=0.1,
                         damping="CaldeiraLeggett", fp_str="fabs(p)", fp_fn=fp_fn)
    W0 = ow.cat_state_wigner(orc.g, a=2.0)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 20)
    assert rel_l2(W, W_ref) < 1e-10
    check_obs(obs, obs_ref)


def test_lie_gross_pitaevskii_fp64(cuda):
    hip, orc = make_pair(cuda, gp=0.1, truncate=True, amp=12.0, dt=0.005,
                         pot=("0.5*pow(x,2)", "x", lambda t, x: 0.5 * x * x, lambda t, x: x))
    W0 = ow.harmonic_oscillator_wigner(orc.g, 0, 1.0, 1.0, 0.0)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 20)
    assert rel_l2(W, W_ref) < 1e-10
    check_obs(obs, obs_ref)


def test_strang_gross_pitaevskii_fp64(cuda):
    hip, orc = make_pair(cuda, scheme="strang", gp=1.0, amp=16.0, dt=0.01,
                         pot=("0.5*pow(x,2)", "x", lambda t, x: 0.5 * x * x, lambda t, x: x))
    W0 = ow.harmonic_oscillator_wigner(orc.g, 0, 1.0, 1.0, 0.0)
    W, obs, W_ref, obs_ref, extras = run_both(hip, orc, W0, 20)
    assert rel_l2(W, W_ref) < 1e-10
    check_obs(obs, obs_ref)


# complex64 state: 'fp32' keeps FP64 arithmetic inside the sweeps (one FP32 rounding of the
# state per sweep: ~6e-7 after 300 steps here); 'fp32fast' also does the transforms in FP32,
# whose rounding error a unitary step accumulates (~3e-5 after 300 steps -- the same as
# numpy's single-precision FFT gives for this workload, i.e. the FP32 arithmetic floor).
# Observables: the moments weight the state error by up to x^2, p^2 ~ 100 at the grid
# edges, so 'fp32fast' records <x^2> etc. to ~6e-4 relative (1e-3 bound); 'fp32' to 1e-5.
@pytest.mark.parametrize("precision,bound,obs_tol", [("fp32", 1e-5, 1e-5), ("fp32fast", 1e-4, 1e-3)])
@pytest.mark.parametrize("scheme", ["lie", "strang"])
def test_fp32_bulk_error(cuda, scheme, precision, bound, obs_tol):
    hip, orc = make_pair(cuda, scheme=scheme, precision=precision, mask=False)
    W0 = ow.gaussian_wigner(orc.g, x0=2.0, p0=0.0, s=1.0)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 300)
    err = rel_l2(W, W_ref)
    print(f"{precision} {scheme}: rel L2 after 300 steps {err:.3e}")
    assert err < bound, err
    for k in ["X", "P", "X2", "P2", "H", "norm"]:
        np.testing.assert_allclose(obs[:, OBS_COL[k]], obs_ref[k], rtol=obs_tol, atol=obs_tol, err_msg=k)


def test_graph_replay_matches_eager(cuda):
    import torch

    hip, orc = make_pair(cuda, mask=True)
    W0 = ow.gaussian_wigner(orc.g, x0=2.0, p0=0.0, s=1.0)
    Wa = hip.to_device(W0)
    Wb = hip.to_device(W0)
    oa = hip.run(Wa, 37, use_graph=False)
    for _ in range(3):  # replays with eager work and syncs in between
        Wb.copy_(hip.to_device(W0))
        ob = hip.run(Wb, 37, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(Wa, Wb)
        assert torch.equal(oa, ob)


def test_norm_conservation_1000_steps(cuda):
    """No mask, no damping: the step is unitary, the norm is conserved to 1e-12 per 1000 steps."""
    hip, orc = make_pair(cuda, mask=False)
    W0 = ow.gaussian_wigner(orc.g, x0=2.0, p0=0.0, s=1.0)
    W = hip.to_device(W0)
    obs = hip.run(W, 1000).cpu().numpy()
    drift = np.abs(obs[:, OBS_COL["norm"]] - obs[0, OBS_COL["norm"]]).max()
    assert drift < 1e-12, drift


@pytest.mark.parametrize("scheme", ["lie", "strang"])
def test_non_polynomial_potential_table(cuda, scheme):
    """V given as an arbitrary expression (the Klein-step form of the Dirac notebooks):
    the p sweep reads the tabulated Bopp difference instead of Horner coefficients."""
    vs = "1.5*0.5*( 1. + tanh((x-5.)/0.3)   ) - 2/sqrt(x*x+1)"
    dvs = "1.5*0.5/(0.3*pow(cosh((x-5.)/0.3),2)) + 2*x/pow(x*x+1,1.5)"

    def vf(t, x):
        return 1.5 * 0.5 * (1.0 + np.tanh((x - 5.0) / 0.3)) - 2.0 / np.sqrt(x * x + 1.0)

    def dvf(t, x):
        return 1.5 * 0.5 / (0.3 * np.cosh((x - 5.0) / 0.3) ** 2) + 2.0 * x / (x * x + 1.0) ** 1.5

    hip, orc = make_pair(cuda, scheme=scheme, amp=12.0, dt=0.005, D_theta=0.05, truncate=True,
                         pot=(vs, dvs, vf, dvf))
    assert hip.tables.pot_deg == -1
    W0 = ow.gaussian_wigner(orc.g, x0=-2.0, p0=2.0, s=1.0)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 40)
    assert rel_l2(W, W_ref) < 1e-10
    check_obs(obs, obs_ref)


DRIVEN_EXPR = "0.5*x*x + 0.3*sin(2*t)*exp(-0.5*x*x) + 0.05*Heaviside(x - 1.5)*cos(t)"


def driven_V(t, x):
    return 0.5 * x * x + 0.3 * np.sin(2 * t) * np.exp(-0.5 * x * x) + 0.05 * np.where(x - 1.5 < 0, 0.0, 1.0) * np.cos(t)


@pytest.mark.parametrize("scheme,damping", [("lie", None), ("strang", None), ("lie", "CaldeiraLeggett"),
                                            ("strang", "CaldeiraLeggett")])
def test_time_dependent_expression_potential(cuda, scheme, damping):
    """A time-dependent non-polynomial V(t, x) (a driven Gaussian bump on a harmonic well with
    a switched step): the p sweep evaluates its program per cell at t = step * dt, as the
    reference evaluates the spliced string (FFT.py:88-94), against the time-dependent oracle."""
    from pywignercuda_amd.wigner2d import Wigner2D

    X = P = 256
    dt, gamma = 0.01, (0.05 if damping else 0.0)
    fp_str = " p*p/sqrt( p*p + 0.5  ) " if damping else None
    fp_fn = (lambda p: p * p / np.sqrt(p * p + 0.5)) if damping else None
    hip = Wigner2D(X, P, 10.0, 10.0, 1.0, 1.0, dt, potential=DRIVEN_EXPR, dpotential="x",
                   kinetic="0.5*p*p" if scheme == "strang" else None, scheme=scheme, D_theta=0.01,
                   gamma=gamma, damping=damping, fp_damping=fp_str, mask=True, device=cuda)
    assert hip.tables.pot_deg == -2 and not hip.tables.time_table
    orc = ow.Wigner2DOracle(X, P, 10.0, 10.0, 1.0, 1.0, dt, D_theta=0.01, gamma=gamma, potential=driven_V,
                            dpotential=lambda t, x: x, kinetic=(lambda p: 0.5 * p * p), fp_damping=fp_fn,
                            damping=damping, scheme=scheme, mask=True, time_dependent=True)
    W0 = ow.gaussian_wigner(orc.g, x0=1.0, p0=1.0, s=1.0)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 40)
    err = rel_l2(W, W_ref)
    print(f"expression V {scheme} {damping}: rel L2 {err:.3e}")
    assert err < 1e-10
    check_obs(obs, obs_ref)
    # the graph path and a run past any coefficient-table length (no time rows needed)
    W2 = hip.to_device(W0)
    hip.run(W2, 41, use_graph=True)
    assert rel_l2(W2.cpu().numpy(), W_ref) < 1e-10
    hip.run(W2, 5000, first_step=41, use_graph=True)
    assert np.isfinite(W2.cpu().numpy()).all()


@pytest.mark.parametrize("precision,bound", [("fp32", 2e-5), ("fp32fast", 2e-5)])
def test_time_dependent_expression_potential_fp32(cuda, precision, bound):
    """The complex64-state forms run the same program (arithmetic in FP64 / FP32)."""
    from pywignercuda_amd.wigner2d import Wigner2D

    X = P = 256
    dt = 0.01
    hip = Wigner2D(X, P, 10.0, 10.0, 1.0, 1.0, dt, potential=DRIVEN_EXPR, dpotential="x", mask=True,
                   precision=precision, device=cuda)
    orc = ow.Wigner2DOracle(X, P, 10.0, 10.0, 1.0, 1.0, dt, potential=driven_V, dpotential=lambda t, x: x,
                            mask=True, time_dependent=True)
    W0 = ow.gaussian_wigner(orc.g, x0=1.0, p0=1.0, s=1.0)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 40)
    err = rel_l2(W.astype(np.complex128), W_ref)
    print(f"expression V {precision}: rel L2 {err:.3e}")
    assert err < bound


def test_expression_potential_single_sweep(cuda):
    """wh_wigner2d_potential at time index 7 of the driven expression V against the oracle's
    potential_step at t = 7 dt (FFT.py:80-126)."""
    import torch

    from pywignercuda_amd.wigner2d import Wigner2D

    X, P, dt = 256, 128, 0.01
    hip = Wigner2D(X, P, 10.0, 10.0, 1.0, 1.0, dt, potential=DRIVEN_EXPR, dpotential="x", mask=True, device=cuda)
    orc = ow.Wigner2DOracle(X, P, 10.0, 10.0, 1.0, 1.0, dt, potential=driven_V, mask=True, time_dependent=True)
    W0 = ow.gaussian_wigner(orc.g, x0=1.0, p0=0.5, s=1.0).astype(np.complex128)
    W = hip.to_device(W0)
    hip.potential(W, step=7)
    torch.cuda.synchronize()
    assert rel_l2(W.cpu().numpy(), orc.potential_step(W0, 7 * dt)) < 1e-12


@pytest.mark.parametrize("precision,bound", [("fp64", 1e-10), ("fp32", 2e-6)])
def test_kvn_fokker_planck(cuda, precision, bound):
    """The classical KvN (Fokker-Planck) step of GPU_FokkerPlank2D_FFT (FFT.py:27-75, 2242-2459)
    on a quartic double well, where it differs from the Wigner step: Caldeira-Leggett friction,
    diffusion, renormalisation, '%f' constants, against the KvN oracle."""
    from pywignercuda_amd.wigner2d import Wigner2D

    X = P = 256
    dt = 0.005
    fp_str = " p*p/sqrt( p*p + 0.25  ) "
    hip = Wigner2D(X, P, 10.0, 10.0, 1.0, 1.0, dt, potential="0.1*pow(x,4) - x*x",
                   dpotential="0.4*pow(x,3) - 2*x", D_theta=0.0143, gamma=0.07, damping="CaldeiraLeggett",
                   fp_damping=fp_str, mask=True, truncate_constants=True, precision=precision, kvn=True,
                   device=cuda)
    orc = ow.Wigner2DOracle(X, P, 10.0, 10.0, 1.0, 1.0, dt, D_theta=0.0143, gamma=0.07,
                            potential=lambda t, x: 0.1 * x ** 4 - x * x,
                            dpotential=lambda t, x: 0.4 * x ** 3 - 2 * x,
                            fp_damping=lambda p: p * p / np.sqrt(p * p + 0.25), damping="CaldeiraLeggett",
                            mask=True, truncate_constants=True, kvn=True)
    W0 = ow.gaussian_wigner(orc.g, x0=1.5, p0=1.0, s=0.8)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 40)
    err = rel_l2(W.astype(np.complex128), W_ref)
    print(f"KvN {precision}: rel L2 {err:.3e}")
    assert err < bound
    if precision == "fp64":
        check_obs(obs, obs_ref)


def test_kvn_fokker_planck_driven(cuda):
    """KvN (Fokker-Planck) with a time-dependent force dPotential(t, x) (a driven double well):
    the device tabulates -hbar theta dV/dx(t, x) per step, against the KvN oracle."""
    from pywignercuda_amd.wigner2d import Wigner2D

    X = P = 256
    dt = 0.005
    hip = Wigner2D(X, P, 10.0, 10.0, 1.0, 1.0, dt, potential="0.1*pow(x,4) - x*x",
                   dpotential="0.4*pow(x,3) - 2*x + 0.5*sin(3*t)", D_theta=0.0143, gamma=0.07,
                   damping="CaldeiraLeggett", fp_damping=" p*p/sqrt( p*p + 0.25  ) ", mask=True,
                   truncate_constants=True, kvn=True, device=cuda)
    assert hip.tables.pot_expr_kvn == 1
    orc = ow.Wigner2DOracle(X, P, 10.0, 10.0, 1.0, 1.0, dt, D_theta=0.0143, gamma=0.07,
                            potential=lambda t, x: 0.1 * x ** 4 - x * x,
                            dpotential=lambda t, x: 0.4 * x ** 3 - 2 * x + 0.5 * np.sin(3 * t),
                            fp_damping=lambda p: p * p / np.sqrt(p * p + 0.25), damping="CaldeiraLeggett",
                            mask=True, truncate_constants=True, kvn=True, time_dependent=True)
    W0 = ow.gaussian_wigner(orc.g, x0=1.5, p0=1.0, s=0.8)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 40)
    err = rel_l2(W, W_ref)
    print(f"KvN driven: rel L2 {err:.3e}")
    assert err < 1e-10
    # the dV record is the reference's dPotential(0, x) average (FFT.py:2310-2311)
    check_obs(obs, obs_ref)


@pytest.mark.parametrize("gp,hbar", [(0.0, 1.0), (0.5, 0.8)])
def test_lie_odm_fp64(cuda, gp, hbar):
    """ODM Lindblad dissipator (order 1) after the V step, with and without the GP term."""
    fp_str = " p*p/sqrt( p*p + 0.5  ) "
    fp_fn = lambda p: p * p / np.sqrt(p * p + 0.5)  # noqa: E731
    hip, orc = make_pair(cuda, amp=12.0, dt=0.005, gamma=0.07 / 4.0, damping="ODM", fp_str=fp_str, fp_fn=fp_fn,
                         truncate=True, gp=gp, L_material=3.0, hbar=hbar,
                         pot=("0.5*pow(x,2)", "x", lambda t, x: 0.5 * x * x, lambda t, x: x))
    W0 = ow.harmonic_oscillator_wigner(orc.g, 0, 1.0, 0.0, 3.0)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 30)
    assert rel_l2(W, W_ref) < 1e-10
    check_obs(obs, obs_ref)


@pytest.mark.parametrize("gp,hbar", [(0.0, 1.0), (0.5, 0.8)])
def test_strang_odm_fp64(cuda, gp, hbar):
    """GPU_Wigner2D_GPitaevskii with dampingFunction 'ODM' (GP.py:1942-1968): V/2; K; V/2;
    Lindbladian_ODM_Order1; renormalisation -- the half steps of consecutive iterations are not
    merged around the dissipator.  With and without the GP term, against the oracle."""
    fp_str = " p*p/sqrt( p*p + 0.5  ) "
    fp_fn = lambda p: p * p / np.sqrt(p * p + 0.5)  # noqa: E731
    hip, orc = make_pair(cuda, scheme="strang", amp=12.0, dt=0.005, gamma=0.07 / 4.0, damping="ODM", fp_str=fp_str,
                         fp_fn=fp_fn, truncate=True, gp=gp, L_material=3.0, hbar=hbar,
                         pot=("0.5*pow(x,2)", "x", lambda t, x: 0.5 * x * x, lambda t, x: x))
    W0 = ow.harmonic_oscillator_wigner(orc.g, 0, 1.0, 0.0, 3.0)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 30)
    err = rel_l2(W, W_ref)
    print(f"Strang ODM gp={gp}: rel L2 {err:.3e}")
    assert err < 1e-10
    check_obs(obs, obs_ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_lie_odm2_two_channels(cuda, precision):
    """GPU_Wigner2D_ODM2 (GPU_Wigner2D_FFT.py:1973-2230): W += L(W) with (weight 1, sign 1), then
    with (weight 0.9, sign -1) on the result, as the ODM2 notebook sets them."""
    fp_str = " p*p/sqrt( p*p + 0.5  ) "
    fp_fn = lambda p: p * p / np.sqrt(p * p + 0.5)  # noqa: E731
    hip, orc = make_pair(cuda, amp=12.0, dt=0.005, gamma=0.07 / 4.0, damping="ODM", fp_str=fp_str, fp_fn=fp_fn,
                         truncate=True, L_material=3.0, odm_weights=(1.0, 0.9), precision=precision,
                         pot=("0.5*pow(x,2)", "x", lambda t, x: 0.5 * x * x, lambda t, x: x))
    W0 = ow.harmonic_oscillator_wigner(orc.g, 0, 1.0, 0.0, 3.0)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 30)
    err = rel_l2(W.astype(np.complex128), W_ref)
    print(f"ODM2 {precision}: rel L2 {err:.3e}")
    if precision == "fp64":
        assert err < 1e-10
        check_obs(obs, obs_ref)
    else:
        assert err < 1e-5


def test_lie_odm_fp32(cuda):
    fp_fn = lambda p: p * p / np.sqrt(p * p + 0.5)  # noqa: E731
    hip, orc = make_pair(cuda, amp=12.0, dt=0.005, gamma=0.07 / 4.0, damping="ODM",
                         fp_str=" p*p/sqrt( p*p + 0.5  ) ", fp_fn=fp_fn, L_material=3.0, precision="fp32",
                         pot=("0.5*pow(x,2)", "x", lambda t, x: 0.5 * x * x, lambda t, x: x))
    W0 = ow.harmonic_oscillator_wigner(orc.g, 0, 1.0, 0.0, 3.0)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 100)
    err = rel_l2(W, W_ref)
    print(f"odm fp32 rel L2 {err:.3e}")
    assert err < 1e-5


@pytest.mark.parametrize("precision", ["fp32", "fp32fast"])
def test_strang_odm_complex64(cuda, precision):
    """The Strang scheme with ODM damping and the GP term (GP.py:1942-1968) in both complex64
    state forms, 100 steps against the FP64 oracle (bars 1e-5 for 'fp32', as the Lie ODM FP32
    test, and 1e-4 for 'fp32fast', whose transforms run in FP32)."""
    fp_fn = lambda p: p * p / np.sqrt(p * p + 0.5)  # noqa: E731
    hip, orc = make_pair(cuda, scheme="strang", amp=12.0, dt=0.005, gamma=0.07 / 4.0, damping="ODM",
                         fp_str=" p*p/sqrt( p*p + 0.5  ) ", fp_fn=fp_fn, L_material=3.0, precision=precision,
                         gp=0.5, pot=("0.5*pow(x,2)", "x", lambda t, x: 0.5 * x * x, lambda t, x: x))
    W0 = ow.harmonic_oscillator_wigner(orc.g, 0, 1.0, 0.0, 3.0)
    W, obs, W_ref, obs_ref, _ = run_both(hip, orc, W0, 100)
    err = rel_l2(W.astype(np.complex128), W_ref)
    print(f"Strang ODM {precision}: rel L2 {err:.3e}")
    assert err < (1e-5 if precision == "fp32" else 1e-4)


def test_strang_odm_graph_replay_matches_direct_launches(cuda):
    """The Strang + ODM iteration (with GP) captured into the step graph and replayed (a run of
    80 steps: four 16-iteration chunks) gives the state and record of the same run launched
    directly, bit for bit."""
    import torch

    fp_fn = lambda p: p * p / np.sqrt(p * p + 0.5)  # noqa: E731
    hip, orc = make_pair(cuda, scheme="strang", amp=12.0, dt=0.005, gamma=0.07 / 4.0, damping="ODM",
                         fp_str=" p*p/sqrt( p*p + 0.5  ) ", fp_fn=fp_fn, L_material=3.0, gp=0.5,
                         pot=("0.5*pow(x,2)", "x", lambda t, x: 0.5 * x * x, lambda t, x: x))
    W0 = ow.harmonic_oscillator_wigner(orc.g, 0, 1.0, 0.0, 3.0)
    Wa, Wb = hip.to_device(W0), hip.to_device(W0)
    oa = hip.run(Wa, 80, use_graph=True).cpu().numpy()
    ob = hip.run(Wb, 80, use_graph=False).cpu().numpy()
    torch.cuda.synchronize()
    assert np.array_equal(Wa.cpu().numpy(), Wb.cpu().numpy())
    assert np.array_equal(oa, ob)


def test_single_sweeps_match_oracle_factors(cuda):
    """wh_wigner2d_kinetic / wh_wigner2d_potential (caller-composed loops): one K factor and
    one V factor at time index 7 of a driven potential V = 0.5 x^2 + 0.3 sin(t) x^3, against
    the oracle's kinetic_step / potential_step (FFT.py:131-158, 80-126)."""
    import torch

    from pywignercuda_amd.wigner2d import Wigner2D

    X = P = 256
    dt = 0.01
    vf = lambda t, x: 0.5 * x * x + 0.3 * np.sin(t) * x ** 3  # noqa: E731
    hip = Wigner2D(X, P, 10.0, 10.0, 1.0, 1.0, dt, potential="0.5*pow(x,2) + 0.3*sin(t)*pow(x,3)",
                   dpotential="x + 0.9*sin(t)*pow(x,2)", mask=True, device=cuda, time_rows=32)
    orc = ow.Wigner2DOracle(X, P, 10.0, 10.0, 1.0, 1.0, dt, potential=vf, mask=True, time_dependent=True)
    W0 = ow.gaussian_wigner(orc.g, x0=1.0, p0=0.5, s=1.0).astype(np.complex128)
    W = hip.to_device(W0)
    hip.kinetic(W)
    torch.cuda.synchronize()
    K_ref = orc.kinetic_step(W0)
    assert rel_l2(W.cpu().numpy(), K_ref) < 1e-12
    hip.potential(W, step=7)
    torch.cuda.synchronize()
    V_ref = orc.potential_step(K_ref, 7 * dt)
    err = rel_l2(W.cpu().numpy(), V_ref)
    print(f"single sweeps: rel L2 {err:.3e}")
    assert err < 1e-12
