"""CPU: the C-ABI library loads, exports every declared entry point, validates its
arguments before touching a device, and maps statuses to the reference's exception
classes (cufft_wrapper.py:15-73).  No compute call is made here."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared_functions():
    text = (ROOT / "include" / "wignerhip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from pywignercuda_amd import _lib

    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = _declared_functions()
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings_and_version():
    from pywignercuda_amd import _lib

    L = _lib.lib()
    assert L.wh_version() > 0
    for code in range(0, 13):
        assert L.wh_status_string(code)
    assert b"size" in L.wh_status_string(8)


def test_fft_plan_argument_validation_maps_to_reference_exceptions():
    from pywignercuda_amd import _lib
    from pywignercuda_amd import cufft_wrapper as cw

    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.wh_fft_plan_many(ctypes.byref(h), 100, 1, 100, 0, 4) == 8
    with pytest.raises(cw.cufftInvalidSize):
        _lib.check(8)
    assert L.wh_fft_plan_many(ctypes.byref(h), 256, 1, 256, 7, 4) == 3  # invalid type
    assert L.wh_fft_plan_many(ctypes.byref(h), 256, 3, 5, 0, 4) == 4  # unsupported layout
    assert L.wh_fft_plan_many(ctypes.byref(h), 256, 1, 256, 0, 0) == 4  # batch < 1
    assert L.wh_fft_exec(None, None, None, -1, None) == 1
    assert L.wh_fft_destroy(None) == 1
    assert issubclass(cw.cufftInvalidPlan, cw.cufftError)
    assert cw.cufftExceptions[9] is cw.cufftUnalignedData


def _valid_wigner_desc():
    from pywignercuda_amd import _lib

    X = P = 256
    keep = []

    def arr(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    d = _lib.Wigner2DDesc()
    d.struct_size = ctypes.sizeof(d)
    d.precision, d.x_dim, d.p_dim, d.scheme = 0, X, P, 0
    d.dt, d.dx, d.dp, d.dtheta, d.dlambda, d.hbar, d.mass = 0.01, 0.078, 0.078, 0.31, 0.31, 1.0, 1.0
    coef = np.zeros(9)
    coef[2] = 0.5
    d.pot_deg, d.pot_rows, d.pot_coef = 2, 1, arr(coef)
    d.x_host, d.p_host = arr(np.zeros(X)), arr(np.zeros(P))
    d.dxdp, d.inv_2m, d.dx_host = 0.006, 0.5, 0.078
    return d, keep


def test_descriptor_layout_matches_the_library():
    """A descriptor the library accepts up to the device query: struct sizes agree."""
    from pywignercuda_amd import _lib

    L = _lib.lib()
    d, keep = _valid_wigner_desc()
    h = ctypes.c_void_p()
    rc = L.wh_wigner2d_create(ctypes.byref(d), ctypes.byref(h))
    assert rc in (0, 12), L.wh_last_error()  # 12 = no device (CPU container)
    if rc == 0:
        L.wh_wigner2d_destroy(h)
    d.struct_size -= 8
    assert L.wh_wigner2d_create(ctypes.byref(d), ctypes.byref(h)) == 10
    assert b"struct_size" in L.wh_last_error()


@pytest.mark.parametrize("field,value,code", [
    ("x_dim", 100, 8), ("p_dim", 4096, 8), ("precision", 5, 3), ("scheme", 9, 10), ("pot_deg", 9, 10),
    ("pot_rows", 0, 10), ("hbar", 0.0, 10), ("damping", 1, 10), ("gp_coeff", 0.5, 10)])
def test_descriptor_validation_errors(field, value, code):
    from pywignercuda_amd import _lib

    L = _lib.lib()
    d, keep = _valid_wigner_desc()
    setattr(d, field, value)
    h = ctypes.c_void_p()
    assert L.wh_wigner2d_create(ctypes.byref(d), ctypes.byref(h)) == code, L.wh_last_error()


def test_expression_program_validation_before_device():
    """The library checks an expression program (length, op codes, stack depth) before it
    touches the device: the evaluator on the device trusts all three."""
    from pywignercuda_amd import _lib
    from pywignercuda_amd.physics import compile_expression

    L = _lib.lib()
    h = ctypes.c_void_p()

    def create(code):
        d, keep = _valid_wigner_desc()
        d.pot_deg = _lib.WH_POT_EXPR
        if code is not None:
            code = np.ascontiguousarray(code, dtype=np.float64)
            keep.append(code)
            d.pot_expr = code.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            d.pot_expr_len = code.shape[0]
        d.pot_t_step = 0.01
        rc = L.wh_wigner2d_create(ctypes.byref(d), ctypes.byref(h))
        if rc == 0:
            L.wh_wigner2d_destroy(h)
        return rc, L.wh_last_error()

    rc, msg = create(compile_expression("0.5*x*x*cos(t)"))
    assert rc in (0, 12), msg  # 12 = no device (CPU container)
    for code, what in [(None, b"ops"), (np.zeros((0, 2)), b"ops"), (np.zeros((65, 2)), b"ops"),
                       ([[1, 0], [4, 0]], b"depth"),             # x + <nothing>
                       ([[1, 0], [1, 0]], b"leaves 2"),          # two values left
                       ([[3, 0]], b"depth"),                     # negate an empty stack
                       ([[1, 0], [21, 0]], b"op"),               # no op 21
                       ([[1.5, 0]], b"op"), ([[np.nan, 0]], b"op"),
                       ([[1, 0]] * 17 + [[4, 0]] * 16, b"depth")]:  # 17 values on a 16-value stack
        rc, msg = create(code)
        assert rc == 10 and b"expression" in msg and what in msg, (code, rc, msg)


def test_ensemble_validation_before_device():
    from pywignercuda_amd import _lib

    L = _lib.lib()
    e = _lib.EnsembleDesc()
    e.struct_size = ctypes.sizeof(e)
    e.dt, e.hbar, e.mass = 0.01, 1.0, 1.0
    m = (_lib.MemberDesc * 2)()
    h = ctypes.c_void_p()
    for i in range(2):
        m[i].struct_size = ctypes.sizeof(_lib.MemberDesc)
        m[i].x_dim = m[i].p_dim = 128
    # no potential coefficients -> invalid descriptor for member 0
    assert L.wh_ensemble_create(ctypes.byref(e), m, 2, ctypes.byref(h)) == 10
    assert b"member 0" in L.wh_last_error()
    assert L.wh_ensemble_create(ctypes.byref(e), m, 0, ctypes.byref(h)) == 4
    assert L.wh_ensemble_run(None, None, 0, 1, None, 1, 0, None) == 1


def test_variant_tables():
    from pywignercuda_amd import _lib

    L = _lib.lib()
    for table in range(3):
        n = L.wh_variant_info(table, -1, None, None)
        assert n >= 2
        r, b = ctypes.c_int(), ctypes.c_int()
        for i in range(n):
            assert L.wh_variant_info(table, i, ctypes.byref(r), ctypes.byref(b)) == 0
            assert r.value in (4, 8, 16) and b.value in (256, 512, 1024)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from pywignercuda_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "libwignerhip.so")
    with pytest.raises(_lib.WignerHipLibraryError):
        _lib.lib()


# ----------------------------------------------------------------------------- physics strings
def test_polynomial_parsing_of_reference_strings():
    from pywignercuda_amd.physics import NonPolynomialPhysics, coefficient_table, polynomial_coefficients

    deg, tab = coefficient_table("0.5*pow(x,2) + 0.1*pow(x,4)", "x", [0.0])
    assert deg == 4 and np.allclose(tab[0, :5], [0, 0, 0.5, 0, 0.1])
    deg, tab = coefficient_table("0.25*pow(x,4) - pow(x,2) - (0.3*cos(2*t))*x", "x", [0.0, 1.0])
    assert np.allclose(tab[:, 1], [-0.3, -0.3 * np.cos(2.0)])
    deg, tab = coefficient_table("0.5*p*p/1.0", "p", [0.0])
    assert deg == 2 and tab[0, 2] == 0.5
    with pytest.raises(NonPolynomialPhysics):
        polynomial_coefficients("-2/sqrt(x*x+1)", "x")
    with pytest.raises(NonPolynomialPhysics):
        polynomial_coefficients("pow(x,10)", "x")


def test_host_eval_uses_numpy_semantics():
    from pywignercuda_amd.physics import host_eval

    x = np.linspace(-1, 1, 5)
    assert np.allclose(host_eval(" p*p/sqrt( p*p + 0.5  ) ", p=x), x * x / np.sqrt(x * x + 0.5))
    assert np.allclose(host_eval("pow(x,3)", x=x), x ** 3)


def test_non_polynomial_potential_tabulated_like_oracle():
    """A non-polynomial V is tabulated as the Bopp difference at kernel coordinates; times
    dt/hbar it must give the oracle's potential-multiplier phase (FFT.py:88-94)."""
    from oracle import wigner2d as ow
    from pywignercuda_amd.wigner2d import _Tables

    vs = "1.5*0.5*( 1. + tanh((x-5.)/0.3)   ) - 2/sqrt(x*x+1)"
    vf = lambda t, x: 1.5 * 0.5 * (1.0 + np.tanh((x - 5.0) / 0.3)) - 2.0 / np.sqrt(x * x + 1.0)  # noqa: E731
    X, P, hbar, dt = 128, 64, 0.7, 0.01
    t = _Tables(X, P, 12.0, 9.0, hbar, 1.0, dt, potential=vs, dpotential=None, scheme="lie", D_theta=0,
                D_lambda=0, gamma=0, damping=None, fp_damping=None, gp_coeff=0, gp_rho_factor=None, mask=False,
                truncate_constants=True, time_rows=None, pot_table=None)
    assert t.pot_deg == -1 and t.pot_dv.shape == (P, X)
    orc = ow.Wigner2DOracle(X, P, 12.0, 9.0, hbar, 1.0, dt, potential=vf, mask=False, truncate_constants=True)
    m = orc.potential_multiplier(0.0)
    phase = t.k_dt / t.k_hbar * t.pot_dv
    np.testing.assert_allclose(np.cos(phase) - 1j * np.sin(phase), m, rtol=0, atol=1e-14)
    # a time-dependent non-polynomial V becomes an expression program (no time rows)
    t2 = _Tables(X, P, 12.0, 9.0, hbar, 1.0, dt, potential="tanh(x)*cos(t)", dpotential=None, scheme="lie",
                 D_theta=0, D_lambda=0, gamma=0, damping=None, fp_damping=None, gp_coeff=0, gp_rho_factor=None,
                 mask=False, truncate_constants=True, time_rows=None, pot_table=None)
    assert t2.pot_deg == -2 and t2.pot_table.shape[0] == 1 and t2.time_dependent and not t2.time_table
    assert t2.pot_t_step == dt and t2.pot_expr.shape == (5, 2)
    # syntax the device evaluator does not have is refused with a clear error
    from pywignercuda_amd.physics import NonPolynomialPhysics
    with pytest.raises(NonPolynomialPhysics):
        _Tables(X, P, 12.0, 9.0, hbar, 1.0, dt, potential="erf(x)*cos(t)", dpotential=None, scheme="lie",
                D_theta=0, D_lambda=0, gamma=0, damping=None, fp_damping=None, gp_coeff=0, gp_rho_factor=None,
                mask=False, truncate_constants=True, time_rows=None, pot_table=None)


def test_kvn_phase_tabulated_like_oracle():
    """The Fokker-Planck (KvN) potential step is the p sweep's table path with the phase
    -dt dV(x) theta of exp_PotentialKvN (FFT.py:37-64): table times dt/hbar against the oracle's
    KvN multiplier, '%f' constants and D_Theta decay included."""
    from oracle import wigner2d as ow
    from pywignercuda_amd.wigner2d import _Tables

    X, P, hbar, dt = 128, 64, 0.8, 0.005
    t = _Tables(X, P, 12.0, 9.0, hbar, 1.0, dt, potential="0.1*pow(x,4) - x*x", dpotential="0.4*pow(x,3) - 2*x",
                scheme="lie", D_theta=0.0143, D_lambda=0, gamma=0, damping=None, fp_damping=None, gp_coeff=0,
                gp_rho_factor=None, mask=True, truncate_constants=True, time_rows=None, pot_table=None, kvn=True)
    assert t.pot_deg == -1 and not t.time_dependent
    orc = ow.Wigner2DOracle(X, P, 12.0, 9.0, hbar, 1.0, dt, D_theta=0.0143,
                            potential=lambda t_, x: 0.1 * x ** 4 - x * x,
                            dpotential=lambda t_, x: 0.4 * x ** 3 - 2 * x, mask=True, truncate_constants=True,
                            kvn=True)
    phase = t.k_dt / t.k_hbar * t.pot_dv
    m = np.cos(phase) - 1j * np.sin(phase)
    m = m * t.theta_decay[:, None] * t.x_mask[None, :]
    np.testing.assert_allclose(m, orc.potential_multiplier(0.0), rtol=0, atol=1e-14)
    # a time-dependent dPotential: its program, tabulated on the device per step in the KvN form
    t3 = _Tables(X, P, 12.0, 9.0, hbar, 1.0, dt, potential="x*x", dpotential="2*x*cos(t)", scheme="lie", D_theta=0,
                 D_lambda=0, gamma=0, damping=None, fp_damping=None, gp_coeff=0, gp_rho_factor=None, mask=True,
                 truncate_constants=True, time_rows=None, pot_table=None, kvn=True)
    assert t3.pot_deg == -2 and t3.pot_expr_kvn == 1 and t3.time_dependent and not t3.time_table


@pytest.mark.parametrize("expr", [
    "0.5*x*x + 0.2*sin(3*t)*exp(-x*x)",
    "1.5*0.5*( 1. + tanh((x-5.)/0.3)   ) - 2/sqrt(x*x+1)*cos(2*M_PI*t)",
    "abs(x)*log(1+t) + atan(x) + sinh(x/4) + cosh(x/5) + tan(x/9) - pow(x, 3)*t/50",
    "Heaviside(x - 1.0)*(1 - exp(-t)) - -x/(2.0+t)",
])
def test_expression_programs_evaluate_like_the_string(expr):
    """The stack program of a physics string, run op by op, gives what the string's numpy
    evaluation gives (the reference splices the string into C and evaluates it per cell)."""
    from pywignercuda_amd.physics import EXPR_OPS, compile_expression, eval_program, host_eval

    code = compile_expression(expr)
    assert code.ndim == 2 and code.shape[1] == 2 and code.shape[0] <= 64
    assert set(code[:, 0].astype(int)) <= set(EXPR_OPS.values())
    x = np.linspace(-7.0, 7.0, 257)
    for t in (0.0, 0.013, 1.7):
        want = np.broadcast_to(host_eval(expr, t=t, x=x), x.shape)
        np.testing.assert_allclose(np.broadcast_to(eval_program(code, x, t), x.shape), want, rtol=1e-13, atol=1e-13)


def test_expression_programs_fold_constants_and_bound_the_stack():
    from pywignercuda_amd.physics import EXPR_OPS, NonPolynomialPhysics, compile_expression

    code = compile_expression("2*M_PI*sqrt(4.0)*x*cos(t)")
    # 2*pi*sqrt(4) folds to one constant: const, x, mul, t, cos, mul
    assert code.shape[0] == 6 and code[0, 0] == EXPR_OPS["const"] and code[0, 1] == pytest.approx(4 * np.pi)
    deep = "x" + "".join(f"+(x*{i}" for i in range(1, 20)) + ")" * 19
    with pytest.raises(NonPolynomialPhysics):
        compile_expression(deep + "*t")  # nesting beyond the device stack
    with pytest.raises(NonPolynomialPhysics):
        compile_expression("+".join(["x*t"] * 40))  # beyond the program length


def test_dirac_descriptor_validation_before_device():
    """wh_dirac_create validates its descriptor before the device query (struct sizes agree)."""
    from pywignercuda_amd import _lib

    L = _lib.lib()
    X = P = 128
    keep = []

    def arr(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    d = _lib.DiracDesc()
    d.struct_size = ctypes.sizeof(d)
    d.x_dim, d.p_dim = X, P
    d.dt, d.dx, d.dp, d.dtheta, d.dlambda, d.mass, d.mass_half, d.c = 0.01, 0.28, 0.25, 0.19, 0.17, 1.0, 0.5, 1.0
    d.pot_deg, d.pot_rows, d.pot_coef = 0, 1, arr(np.zeros(9))
    d.x_host, d.p_host, d.theta_host = arr(np.zeros(X)), arr(np.zeros(P)), arr(np.zeros(P))
    h = ctypes.c_void_p()
    assert L.wh_dirac_create(ctypes.byref(d), ctypes.byref(h)) in (0, 12), L.wh_last_error()
    if h.value:
        L.wh_dirac_destroy(h)
    for field, value, code in [("x_dim", 96, 8), ("precision", 4, 3), ("pot_deg", -1, 10), ("c", 0.0, 10),
                               ("struct_size", 8, 10), ("a_rows", -1, 10), ("a_rows", 3, 10),
                               ("pot_deg", -2, 10),  # WH_POT_EXPR without programs
                               ("mass_expr_len", 3, 10)]:  # a mass program without its ops
        old = getattr(d, field)
        setattr(d, field, value)
        assert L.wh_dirac_create(ctypes.byref(d), ctypes.byref(h)) == code, (field, L.wh_last_error())
        setattr(d, field, old)
    assert L.wh_dirac_run(None, None, 0, 1, None, 0, None) == 1


def test_dirac_plan_rejects_unsupported_physics():
    from pywignercuda_amd.dirac import WignerDirac4x4

    with pytest.raises(NotImplementedError):
        WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=("0.", "0.", " fmod(x, 2.) ", "0."))
    with pytest.raises(NotImplementedError):  # a vector potential outside the expression programs' ops
        WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=("0.", "erf(x)*cos(t)", "0.", "0."))
    with pytest.raises(NotImplementedError):
        WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, gamma=0.1, damping="Lindblad")
    with pytest.raises(NotImplementedError):  # KvN takes Potential_0 only
        WignerDirac4x4(128, 128, 18.0, 16.0, 1.0, 1.0, 0.01, potentials=("x", "0.", " -x ", "0."), kvn=True)


def test_reference_binding_accepts_the_reference_plan_arguments():
    """The reference-side ctypes binding with the exact cufftPlanMany arguments of the
    reference's plan classes at 512^2: every layout passes validation (on a host without a
    GPU the call then fails only at the device allocation), and layouts outside the set
    are refused with cuFFT's codes."""
    from pywignercuda_amd import _lib
    from refplans import StatusError, load_binding, reference_plan_calls

    ns = load_binding(_lib.LIB_PATH)
    device_codes = {0, 2, 5, 6, 7, 12}  # success or a failure of the device allocation
    for name, args, keep in reference_plan_calls(512, 512):
        try:
            plan = ns["cufftPlanMany"](*args)
            ns["cufftDestroy"](plan)
        except StatusError as e:
            assert e.status in device_codes, (name, e.status, _lib.last_error())
    for name, args, keep in reference_plan_calls(512, 512, fft_type=0x29):
        try:
            ns["cufftDestroy"](ns["cufftPlanMany"](*args))
        except StatusError as e:
            assert e.status in device_codes, (name, e.status, _lib.last_error())
    n = np.asarray((512, 512, 512), np.int32)
    bad = [
        ((3, n.ctypes.data, None, 1, 0, None, 1, 0, 0x69, 1), 4),  # rank 3
        ((1, n.ctypes.data, None, 1, 0, None, 1, 0, 0x2A, 1), 3),  # C2R
        ((1, n.ctypes.data, n.ctypes.data, 2, 7, n.ctypes.data, 2, 7, 0x69, 4), 4),  # strides
        ((1, n.ctypes.data, n.ctypes.data, 1, 512, n.ctypes.data, 512, 1, 0x69, 4), 4),  # out layout
        ((2, n.ctypes.data, n.ctypes.data, 1, 512, n.ctypes.data, 1, 512, 0x69, 1), 4),  # 2-D advanced
    ]
    m = np.asarray((100,), np.int32)
    bad.append(((1, m.ctypes.data, None, 1, 0, None, 1, 0, 0x69, 1), 8))
    for args, code in bad:
        with pytest.raises(StatusError) as ei:
            ns["cufftPlanMany"](*args)
        assert ei.value.status == code, (args, ei.value.status)


def test_observe_noalloc_validation_before_device():
    """wh_wigner2d_observe_noalloc checks the descriptor and the size limit before any
    device call: struct size, precision, x_dim beyond WH_NOALLOC_MAX_X, missing grids."""
    from pywignercuda_amd import _lib

    L = _lib.lib()
    d = _lib.ObserveDesc()
    d.struct_size = ctypes.sizeof(d)
    d.precision = _lib.WH_C128
    d.x_dim, d.p_dim = 256, 256
    d.x_host = d.p_host = d.dv_host = d.v0_host = 0x1000  # never dereferenced on these paths
    d.dxdp, d.inv_2m = 0.01, 0.5
    W, obs = ctypes.c_void_p(0x2000), ctypes.c_void_p(0x3000)
    assert L.wh_wigner2d_observe_noalloc(None, W, obs, None) == 4
    bad = _lib.ObserveDesc.from_buffer_copy(d)
    bad.struct_size -= 8
    assert L.wh_wigner2d_observe_noalloc(ctypes.byref(bad), W, obs, None) == 10
    bad = _lib.ObserveDesc.from_buffer_copy(d)
    bad.precision = 7
    assert L.wh_wigner2d_observe_noalloc(ctypes.byref(bad), W, obs, None) == 3
    for x_dim in (0, _lib.WH_NOALLOC_MAX_X + 1, 4096):
        bad = _lib.ObserveDesc.from_buffer_copy(d)
        bad.x_dim = x_dim
        assert L.wh_wigner2d_observe_noalloc(ctypes.byref(bad), W, obs, None) == 8
        assert b"x_dim" in L.wh_last_error()
    bad = _lib.ObserveDesc.from_buffer_copy(d)
    bad.v0_host = None
    assert L.wh_wigner2d_observe_noalloc(ctypes.byref(bad), W, obs, None) == 10
    assert L.wh_wigner2d_observe_noalloc(ctypes.byref(d), None, obs, None) == 4


def test_dropin_host_helpers(tmp_path):
    """Off-path host helpers of Propagator_Base (GPU_Wigner2D_FFT.py:944-1054, 1159-1193), no
    device: OperatorP_XBasis equals the spectral construction sum_q p_q |q><q| over the
    half-integer momenta p_q = 2 pi q / (N dX), q = -(N-1)/2 .. (N-1)/2 (so it is Hermitian and
    differentiates those plane waves exactly); the marginals integrate to the norm and are the
    x / p densities of a product state; the Ehrenfest readers return k*dt and the run file's
    record whether it is stored as '<name>_average' or as Run's '<name>_Ehrenfest'."""
    from pywignercuda_amd import _io
    from pywignercuda_amd.GPU_Wigner2D_FFT import GPU_Wigner2D_FFT

    o = GPU_Wigner2D_FFT.__new__(GPU_Wigner2D_FFT)
    o.SetPhaseSpaceBox2D(32, 16, 5.0, 4.0)
    N, dX = o.X_gridDIM, o.dX
    q = (2.0 * np.arange(N) - N + 1.0) / 2.0
    x, p = q * dX, 2.0 * np.pi * q / (N * dX)
    P_spec = (np.exp(1j * np.outer(x, p)) * p) @ np.exp(-1j * np.outer(p, x)) / N
    P = o.OperatorP_XBasis()
    np.testing.assert_allclose(P, P_spec, atol=1e-13)
    np.testing.assert_allclose(P, P.conj().T, atol=1e-13)
    wave = np.exp(1j * p[N // 2 + 3] * o.X_range)
    np.testing.assert_allclose(P @ wave, p[N // 2 + 3] * wave, atol=1e-12)
    np.testing.assert_allclose(np.diag(o.OperatorX_XBasis()), o.X_range)

    fx = np.exp(-(o.X_range - 1.0) ** 2)
    gp = np.exp(-0.5 * o.P_range ** 2)
    W = gp[:, None] * fx[None, :]
    W /= W.sum() * o.dX * o.dP
    px, pp = o.WignerMarginal_Probability_x(W), o.WignerMarginal_Probability_p(W)
    assert px.shape == (32,) and pp.shape == (16,)
    np.testing.assert_allclose(px.sum() * o.dX, 1.0)
    np.testing.assert_allclose(pp.sum() * o.dP, 1.0)
    np.testing.assert_allclose(px / px.max(), fx / fx.max())
    np.testing.assert_allclose(pp / pp.max(), gp / gp.max())

    f = _io.RunFile(str(tmp_path / "rec"))
    f["dt"], f["timeSteps"] = 0.25, 4
    f["/Ehrenfest/X_Ehrenfest"] = np.arange(5.0)
    f["/Ehrenfest/Hamiltonian_average"] = np.arange(5.0) + 10
    f.close()
    o.fileName = str(tmp_path / "rec")
    t, X = o.Ehrenfest_X_FromFile()
    np.testing.assert_allclose(t, [0.0, 0.25, 0.5, 0.75, 1.0])
    np.testing.assert_allclose(X, np.arange(5.0))
    np.testing.assert_allclose(o.Ehrenfest_Hamiltonian_FromFile(str(tmp_path / "rec"))[1], np.arange(5.0) + 10)


def test_dirac_dropin_host_helpers(tmp_path):
    """GPU_WignerDirac2D_4x4 host helpers, no device.  MajoranaSpinorPlus/Minus (D44.py:2337-2355)
    split Psi into the even and odd parts of the antiunitary charge conjugation C, C^2 = 1; C maps
    the reference's own particle-up spinor at px onto i times its antiparticle-up spinor at -px.
    Load_Density / Load_WignerFunction (D44.py:2915-2951) read back what Run's frame writers store."""
    from pywignercuda_amd import _io
    from pywignercuda_amd.GPU_WignerDirac2D_4x4 import GPU_WignerDirac2D_4x4

    o = GPU_WignerDirac2D_4x4.__new__(GPU_WignerDirac2D_4x4)
    o.mass, o.c = 1.0, 137.036
    X = np.linspace(-5.0, 5.0, 64, endpoint=False)
    rng = np.random.default_rng(7)
    Psi = rng.normal(size=(4, 64)) + 1j * rng.normal(size=(4, 64))
    plus, minus = o.MajoranaSpinorPlus(Psi), o.MajoranaSpinorMinus(Psi)
    np.testing.assert_allclose(plus + minus, Psi)
    C = lambda s: 2.0 * o.MajoranaSpinorPlus(s) - s  # noqa: E731  (C s = 2 Plus(s) - s)
    np.testing.assert_allclose(C(C(Psi)), Psi, atol=1e-14)
    np.testing.assert_allclose(C(plus), plus, atol=1e-14)
    np.testing.assert_allclose(C(minus), -minus, atol=1e-14)
    up = o.GaussianSpinor_ParticleUp(0.5, 3.0, 1.2, X)
    np.testing.assert_allclose(C(up), 1j * o.GaussianSpinor_AntiParticleUp(0.5, -3.0, 1.2, X), atol=1e-10)

    W = rng.normal(size=(16, 8, 4)) + 1j * rng.normal(size=(16, 8, 4))
    f = _io.RunFile(str(tmp_path / "d"))
    o.save_Density(f, 3, W)
    o.save_WignerFunction(f, 5, W)
    f.close()
    np.testing.assert_allclose(o.Load_Density(str(tmp_path / "d"), 3), np.real(W[0] + W[5] + W[10] + W[15]))
    W44 = o.Load_WignerFunction(str(tmp_path / "d"), 5)
    assert W44.shape == (4, 4, 8, 4)
    np.testing.assert_allclose(W44[2, 1], np.fft.fftshift(W[9]))
