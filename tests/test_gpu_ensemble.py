"""GPU: ensembles of uneven Wigner functions in one launch per sweep, and every kernel variant.

Each member is checked against the oracle run on its own (FP64 bar: rel-L2 1e-10).
"""
import numpy as np
import pytest

from oracle import wigner2d as ow

pytestmark = pytest.mark.gpu


def _member(X, a, b, x0, s):
    return dict(X=X, P=X, X_amp=10.0, P_amp=10.0, potential=f"{a!r}*pow(x,4) - {b!r}*pow(x,2)",
                dpotential=f"4*{a!r}*pow(x,3) - 2*{b!r}*x", x0=x0, s=s, a=a, b=b)


def _oracle(m, dt, steps, scheme="lie"):
    a, b = m["a"], m["b"]
    orc = ow.Wigner2DOracle(m["X"], m["P"], m["X_amp"], m["P_amp"], 1.0, 1.0, dt, scheme=scheme,
                            potential=lambda t, x: a * x ** 4 - b * x ** 2,
                            dpotential=lambda t, x: 4 * a * x ** 3 - 2 * b * x,
                            kinetic=lambda p: 0.5 * p * p, mask=False)
    W0 = ow.gaussian_wigner(orc.g, x0=m["x0"], s=m["s"])
    W, obs, _ = orc.run(W0, steps)
    return W0, W, obs


@pytest.mark.parametrize("scheme", ["lie", "strang"])
def test_uneven_ensemble_matches_oracle(cuda, scheme):
    import torch

    from pywignercuda_amd.wigner2d import WignerEnsemble

    members = [_member(128, 0.1, 1.0, 1.0, 1.0), _member(256, 0.05, 0.5, -2.0, 0.8),
               _member(512, 0.2, 2.0, 0.5, 1.2), _member(128, 0.15, 1.5, -1.0, 0.9)]
    dt, steps = 0.01, 25
    ens = WignerEnsemble(members, 1.0, 1.0, dt, scheme=scheme, kinetic="0.5*p*p", device=cuda)
    refs = [_oracle(m, dt, steps, scheme) for m in members]
    states = ens.states_to_device([r[0] for r in refs])
    obs = ens.run(states, steps + 1, use_graph=True).cpu().numpy()
    torch.cuda.synchronize()
    for i, (W0, W_ref, obs_ref) in enumerate(refs):
        W = states[i].cpu().numpy()
        err = np.linalg.norm(W - W_ref) / np.linalg.norm(W_ref)
        assert err < 1e-10, (i, err)
        np.testing.assert_allclose(obs[i, : steps + 1, 0], obs_ref["X"], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(obs[i, : steps + 1, 8], obs_ref["H"], rtol=1e-10, atol=1e-12)


def test_ensemble_with_an_expression_potential_member(cuda):
    """An ensemble whose second member has a time-dependent non-polynomial V (expression
    program) beside polynomial members: the plan runs the expression p kernels for all of
    them, every member against its own oracle."""
    import torch

    from pywignercuda_amd.wigner2d import WignerEnsemble

    members = [_member(128, 0.1, 1.0, 1.0, 1.0), _member(256, 0.05, 0.5, -2.0, 0.8)]
    expr = "0.5*x*x + 0.2*cos(3*t)*tanh(x)"
    members[1].update(potential=expr, dpotential="x")
    dt, steps = 0.01, 25
    ens = WignerEnsemble(members, 1.0, 1.0, dt, scheme="strang", kinetic="0.5*p*p", device=cuda)
    assert ens.tables[1].pot_deg == -2 and ens.tables[0].pot_deg >= 0
    refs = [_oracle(members[0], dt, steps, "strang")]
    m = members[1]
    orc = ow.Wigner2DOracle(m["X"], m["P"], m["X_amp"], m["P_amp"], 1.0, 1.0, dt, scheme="strang",
                            potential=lambda t, x: 0.5 * x * x + 0.2 * np.cos(3 * t) * np.tanh(x),
                            dpotential=lambda t, x: x, kinetic=lambda p: 0.5 * p * p, mask=False,
                            time_dependent=True)
    W0 = ow.gaussian_wigner(orc.g, x0=m["x0"], s=m["s"])
    W_ref, obs_ref, _ = orc.run(W0, steps)
    refs.append((W0, W_ref, obs_ref))
    states = ens.states_to_device([r[0] for r in refs])
    obs = ens.run(states, steps + 1, use_graph=True).cpu().numpy()
    torch.cuda.synchronize()
    for i, (W0, W_ref, obs_ref) in enumerate(refs):
        W = states[i].cpu().numpy()
        err = np.linalg.norm(W - W_ref) / np.linalg.norm(W_ref)
        assert err < 1e-10, (i, err)
        np.testing.assert_allclose(obs[i, : steps + 1, 8], obs_ref["H"], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("scheme", ["lie", "strang"])
def test_uneven_ensemble_with_odm_damping(cuda, scheme):
    """ODM damping (Lindbladian_ODM_Order1, GPU_Wigner2D_FFT.py:1516-1634; after the second half
    step in the Strang scheme, GP.py:1958-1968) over uneven members with their own L_material,
    one launch per ODM sweep for all of them; every member against its own oracle."""
    import torch

    from pywignercuda_amd.wigner2d import WignerEnsemble

    fp_str = " p*p/sqrt( p*p + 0.5  ) "
    members = [_member(128, 0.1, 1.0, 1.0, 1.0), _member(256, 0.05, 0.5, -2.0, 0.8),
               _member(128, 0.15, 1.5, -1.0, 0.9)]
    for m, L in zip(members, (3.0, 2.0, 4.0)):
        m.update(fp_damping=fp_str, L_material=L)
    dt, steps, gamma = 0.01, 40, 0.05  # 41 iterations: the step graph's 16-iteration chunks replay
    ens = WignerEnsemble(members, 1.0, 1.0, dt, scheme=scheme, kinetic="0.5*p*p", gamma=gamma,
                         damping="ODM", device=cuda)
    refs = []
    for m in members:
        a, b = m["a"], m["b"]
        orc = ow.Wigner2DOracle(m["X"], m["P"], m["X_amp"], m["P_amp"], 1.0, 1.0, dt, scheme=scheme,
                                potential=lambda t, x, a=a, b=b: a * x ** 4 - b * x ** 2,
                                dpotential=lambda t, x, a=a, b=b: 4 * a * x ** 3 - 2 * b * x,
                                kinetic=lambda p: 0.5 * p * p, mask=False, gamma=gamma, damping="ODM",
                                fp_damping=lambda p: p * p / np.sqrt(p * p + 0.5), L_material=m["L_material"])
        W0 = ow.gaussian_wigner(orc.g, x0=m["x0"], s=m["s"])
        W_ref, obs_ref, _ = orc.run(W0, steps)
        refs.append((W0, W_ref, obs_ref))
    states = ens.states_to_device([r[0] for r in refs])
    obs = ens.run(states, steps + 1, use_graph=True).cpu().numpy()
    torch.cuda.synchronize()
    for i, (W0, W_ref, obs_ref) in enumerate(refs):
        W = states[i].cpu().numpy()
        err = np.linalg.norm(W - W_ref) / np.linalg.norm(W_ref)
        print(f"ODM {scheme} member {i}: rel L2 {err:.3e}")
        assert err < 1e-10, (i, err)
        np.testing.assert_allclose(obs[i, : steps + 1, 0], obs_ref["X"], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(obs[i, : steps + 1, 8], obs_ref["H"], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp32fast"])
def test_every_variant_matches_oracle(cuda, prec):
    import ctypes

    import torch

    from pywignercuda_amd import _lib
    from pywignercuda_amd.wigner2d import Wigner2D

    m = _member(512, 0.1, 1.0, 1.5, 1.0)
    W0, W_ref, obs_ref = _oracle(m, 0.01, 10)
    L = _lib.lib()
    nx, npv = L.wh_variant_info(0, -1, None, None), L.wh_variant_info(1, -1, None, None)
    tol = {"fp64": 1e-10, "fp32": 5e-7, "fp32fast": 2e-6}[prec]
    hip = Wigner2D(512, 512, 10.0, 10.0, 1.0, 1.0, 0.01, potential=m["potential"], dpotential=m["dpotential"],
                   mask=False, precision=prec, device=cuda)
    for xv in range(nx):
        for pv in range(npv):
            try:
                hip.set_variants(xv, pv, -1)
            except _lib.WignerHipInvalidValue:
                continue
            W = hip.to_device(W0)
            hip.run(W, 11, use_graph=False)
            torch.cuda.synchronize()
            err = np.linalg.norm(W.cpu().numpy() - W_ref) / np.linalg.norm(W_ref)
            assert err < tol, (xv, pv, err)


def test_fp32_error_1000_steps(cuda):
    """complex64 state (FP64 arithmetic in the sweeps) vs the FP64 form after 1000 unitary
    steps at 512^2: the 1e-5 bound of the FP32 form."""
    import torch

    from pywignercuda_amd.wigner2d import Wigner2D

    kw = dict(potential="0.5*pow(x,2)+0.1*pow(x,4)", dpotential="x+0.4*pow(x,3)", mask=False, device=cuda)
    p64 = Wigner2D(512, 512, 10.0, 10.0, 1.0, 1.0, 0.01, precision="fp64", **kw)
    p32 = Wigner2D(512, 512, 10.0, 10.0, 1.0, 1.0, 0.01, precision="fp32", **kw)
    x, p = p64.x_host[None, :], p64.p_host[:, None]
    W0 = np.exp(-((x - 2.0) ** 2) - p ** 2) / np.pi
    A, B = p64.to_device(W0), p32.to_device(W0)
    oa = p64.run(A, 1000).cpu().numpy()
    ob = p32.run(B, 1000).cpu().numpy()
    torch.cuda.synchronize()
    a, b = A.cpu().numpy(), B.cpu().numpy().astype(np.complex128)
    err = np.linalg.norm(a - b) / np.linalg.norm(a)
    print(f"fp32 rel-L2 after 1000 steps: {err:.3e}")
    assert err < 1e-5, err
    for col, name in [(0, "X"), (2, "P"), (8, "H"), (10, "norm")]:
        d = np.abs(oa[:, col] - ob[:, col]).max()
        print(f"  max |<{name}>_fp32 - <{name}>_fp64| = {d:.3e}")
        assert d < 1e-5, (name, d)


def test_fp32_throughput_variants_error_1000_steps(cuda):
    """The batch throughput sweeps (x variant 4, p variant 3: exchanges between transform
    passes held in complex64) against the FP64 ensemble after 1000 unitary steps: the 1e-5
    bound of the FP32 form, for the uneven sizes that select them by default."""
    import torch

    from pywignercuda_amd import _lib
    from pywignercuda_amd.wigner2d import WignerEnsemble

    L = _lib.lib()
    xtp, ptp = L.wh_variant_info(0, -1, None, None) - 1, L.wh_variant_info(1, -1, None, None) - 1
    members = [_member(256, 0.1, 1.0, 1.0, 1.0), _member(512, 0.05, 0.5, -2.0, 0.8),
               _member(128, 0.2, 2.0, 0.5, 1.2)]
    e64 = WignerEnsemble(members, 1.0, 1.0, 0.01, kinetic="0.5*p*p", precision="fp64", device=cuda)
    e32 = WignerEnsemble(members, 1.0, 1.0, 0.01, kinetic="0.5*p*p", precision="fp32", device=cuda,
                         variants=(xtp, ptp, 0))
    W0 = []
    for t, m in zip(e64.tables, members):
        x, p = t.x_host[None, :], t.p_host[:, None]
        W0.append(np.exp(-((x - m["x0"]) ** 2) / m["s"] ** 2 - m["s"] ** 2 * p ** 2) / np.pi)
    A, B = e64.states_to_device(W0), e32.states_to_device(W0)
    e64.run(A, 1000, use_graph=True)
    e32.run(B, 1000, use_graph=True)
    torch.cuda.synchronize()
    for i in range(len(members)):
        a, b = A[i].cpu().numpy(), B[i].cpu().numpy().astype(np.complex128)
        err = np.linalg.norm(a - b) / np.linalg.norm(a)
        print(f"member {i} ({members[i]['X']}^2): fp32 throughput-variant rel-L2 after 1000 steps {err:.3e}")
        assert err < 1e-5, (i, err)
