"""GPU: wh_wigner2d_observe_noalloc, the observables of a state without a plan and without
any device allocation (its column sums live in a fixed device array of the library),
against wh_wigner2d_observe of a plan on the same state.  Both run the same column-sum and
reduce kernels, so the rows agree bit for bit.  Two calls in a row share the device array:
the second on a smaller grid must not see the first call's columns."""
import ctypes

import numpy as np
import pytest

from pywignercuda_amd import _lib

pytestmark = pytest.mark.gpu


def _plan_and_state(cuda, X, P, precision, seed):
    """A Wigner2D plan and a state it has propagated (complex, with negative parts)."""
    import torch

    from oracle import wigner2d as ow
    from pywignercuda_amd.wigner2d import Wigner2D

    hip = Wigner2D(X, P, 8.0, 8.0, 1.0, 1.0, 0.01, potential="0.5*pow(x,2)+0.1*pow(x,4)",
                   dpotential="x+0.4*pow(x,3)", precision=precision, device=cuda)
    g = ow.Grid(X, P, 8.0, 8.0)
    W0 = ow.gaussian_wigner(g, x0=1.5, p0=-0.5, s=1.0)
    rng = np.random.default_rng(seed)
    W0 = W0 + 1e-3 * (rng.standard_normal(W0.shape) + 1j * rng.standard_normal(W0.shape))
    W = hip.to_device(W0)
    hip.run(W, 6, use_graph=False)
    torch.cuda.synchronize()
    return hip, W


def _noalloc(hip, W, prec_code):
    import torch

    dev = W.device
    t = hip.tables  # the host grids the plan was built from
    grids = [torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
             for a in (t.x_host, t.p_host, t.dv, t.v0)]
    d = _lib.ObserveDesc()
    d.struct_size = ctypes.sizeof(_lib.ObserveDesc)
    d.precision = prec_code
    d.x_dim, d.p_dim = int(W.shape[1]), int(W.shape[0])
    d.x_host, d.p_host, d.dv_host, d.v0_host = (int(g.data_ptr()) for g in grids)
    d.dxdp, d.inv_2m = t.dxdp, t.inv_2m
    obs = torch.zeros(_lib.WH_NOBS, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    rc = _lib.lib().wh_wigner2d_observe_noalloc(ctypes.byref(d), _lib.dptr(W), _lib.dptr(obs),
                                                _lib.current_stream_handle())
    _lib.check(rc, "wh_wigner2d_observe_noalloc")
    torch.cuda.synchronize()
    return obs.cpu().numpy(), grids


@pytest.mark.parametrize("precision,code", [("fp64", _lib.WH_C128), ("fp32", _lib.WH_C64)])
def test_noalloc_matches_plan_observe_twice(cuda, precision, code):
    import torch

    rows = []
    for X, P, seed in ((512, 512, 1), (256, 128, 2)):  # the second call on a smaller grid
        print(f"observe_noalloc {precision} P={P} X={X}", flush=True)
        hip, W = _plan_and_state(cuda, X, P, precision, seed)
        ref = hip.observe(W).cpu().numpy()
        got, _ = _noalloc(hip, W, code)
        np.testing.assert_array_equal(got, ref)
        # the row is the state's: norm and <x> against a host sum of the same state
        Wh = W.cpu().numpy().astype(np.complex128)
        dxdp = hip.tables.dxdp
        norm = dxdp * Wh.real.sum()
        assert abs(got[_lib.OBS_COLUMNS.index("norm")] - norm) < 1e-9 * abs(norm)
        xm = dxdp * (Wh.real * hip.x_host[None, :]).sum()
        assert abs(got[0] - xm) < 1e-9 * max(abs(xm), 1.0)
        rows.append(got)
    assert not np.array_equal(rows[0], rows[1])


def test_noalloc_rejects_sizes_beyond_the_maximum(cuda):
    import torch

    hip, W = _plan_and_state(cuda, 256, 256, "fp64", 3)
    d = _lib.ObserveDesc()
    d.struct_size = ctypes.sizeof(_lib.ObserveDesc)
    d.precision = _lib.WH_C128
    d.x_dim, d.p_dim = _lib.WH_NOALLOC_MAX_X * 2, 4
    dummy = torch.zeros(8, dtype=torch.float64, device=cuda)
    d.x_host = d.p_host = d.dv_host = d.v0_host = int(dummy.data_ptr())
    obs = torch.zeros(_lib.WH_NOBS, dtype=torch.float64, device=cuda)
    rc = _lib.lib().wh_wigner2d_observe_noalloc(ctypes.byref(d), _lib.dptr(W), _lib.dptr(obs), None)
    assert rc == 8  # WH_INVALID_SIZE, nothing launched
    assert b"x_dim" in _lib.lib().wh_last_error()
    d.x_dim = 256
    d.struct_size = 8
    assert _lib.lib().wh_wigner2d_observe_noalloc(ctypes.byref(d), _lib.dptr(W), _lib.dptr(obs), None) == 10
    torch.cuda.synchronize()
    assert float(obs.abs().sum()) == 0.0
