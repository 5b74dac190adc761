"""GPU: the axis transforms that replace cuFFT, against the DFT definition (numpy.fft).

Pins the FFT kernels to the mathematics (the reference's cuFFT results are not
available: parity at that boundary is pinned to the DFT, not to cuFFT).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [64, 128, 256, 512, 1024, 2048]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("n", SIZES)
def test_axis_transforms(cuda, prec, n):
    import torch

    from pywignercuda_amd import cufft_wrapper as cw

    rng = np.random.default_rng(n)
    other = 128
    a = rng.standard_normal((other, n)) + 1j * rng.standard_normal((other, n))
    dt = torch.complex128 if prec == "fp64" else torch.complex64
    tol = 1e-13 if prec == "fp64" else 2e-6
    ft = cw.CUFFT_Z2Z if prec == "fp64" else cw.CUFFT_C2C
    # rows (axis 1)
    plan = cw.Plan_Z2Z_2D_Axis1((other, n), fft_type=ft)
    t = torch.tensor(a, dtype=dt, device=cuda)
    cw.fft_Z2Z(t, t, plan)
    ref = np.fft.fft(a, axis=1)
    assert np.abs(t.cpu().numpy() - ref).max() / np.abs(ref).max() < tol
    cw.ifft_Z2Z(t, t, plan)
    assert np.abs(t.cpu().numpy() / n - a).max() / np.abs(a).max() < tol
    cw.cufftDestroy(plan)
    # columns (axis 0) of an [n][other] array
    b = a.T.copy()
    plan = cw.Plan_Z2Z_2D_Axis0((n, other), fft_type=ft)
    t = torch.tensor(b, dtype=dt, device=cuda)
    cw.fft_Z2Z(t, t, plan)
    ref = np.fft.fft(b, axis=0)
    assert np.abs(t.cpu().numpy() - ref).max() / np.abs(ref).max() < tol
    cw.ifft_Z2Z_normalized(t, t, plan)
    assert np.abs(t.cpu().numpy() - b).max() / np.abs(b).max() < tol
    torch.cuda.synchronize()


def test_errors_map_to_reference_exceptions(cuda):
    import torch

    from pywignercuda_amd import cufft_wrapper as cw

    with pytest.raises(cw.cufftInvalidSize):
        cw.Plan_Z2Z((100,))
    plan = cw.Plan_Z2Z((256,), batch=4)
    t = torch.zeros((4, 256), dtype=torch.complex128, device=cuda)
    u = torch.zeros_like(t)
    with pytest.raises(cw.cufftInvalidValue):
        cw.cufftExecZ2Z(plan, t, u, cw.CUFFT_FORWARD)  # out of place is not supported


def test_reference_binding_runs_the_reference_plans(cuda):
    """The reference-side ctypes binding, called as the reference's plan classes and
    fft_Z2Z / ifft_Z2Z call cuFFT (cufft_wrapper.py:160-236, 396-403; raw device addresses,
    null stream), at 512^2: the 2-D plan, the 1-D plan of SetFFT_Plans and both axis plans
    against numpy.fft; inverse unnormalised."""
    import torch

    from pywignercuda_amd import _lib
    from refplans import load_binding, reference_plan_calls

    ns = load_binding(_lib.LIB_PATH)
    P = X = 512
    rng = np.random.default_rng(5)
    a = rng.standard_normal((P, X)) + 1j * rng.standard_normal((P, X))
    fwd = {"Plan_Z2Z((P,X))": lambda z: np.fft.fft2(z), "Plan_Z2Z((X,))": lambda z: np.fft.fft(z[0]),
           "Plan_Z2Z_2D_Axis0((P,X))": lambda z: np.fft.fft(z, axis=0),
           "Plan_Z2Z_2D_Axis1((P,X))": lambda z: np.fft.fft(z, axis=1)}
    norm = {"Plan_Z2Z((P,X))": P * X, "Plan_Z2Z((X,))": X, "Plan_Z2Z_2D_Axis0((P,X))": P,
            "Plan_Z2Z_2D_Axis1((P,X))": X}
    for name, args, keep in reference_plan_calls(P, X):
        plan = ns["cufftPlanMany"](*args)
        data = a[:1] if name == "Plan_Z2Z((X,))" else a
        t = torch.tensor(data, dtype=torch.complex128, device=cuda)
        torch.cuda.synchronize()
        addr = int(t.data_ptr())
        ns["cufftExecZ2Z"](plan, addr, addr, -1)
        torch.cuda.synchronize()
        ref = fwd[name](data).reshape(t.shape)
        err = np.abs(t.cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err < 1e-13, (name, err)
        ns["cufftExecZ2Z"](plan, addr, addr, 1)
        torch.cuda.synchronize()
        err = np.abs(t.cpu().numpy() / norm[name] - data).max() / np.abs(data).max()
        assert err < 1e-13, (name, err)
        ns["cufftDestroy"](plan)
