"""Physics strings of the reference API -> host evaluators and device coefficient tables.

The reference takes V(t,x), dV/dx, K(p) and the friction f(p) as C expression
strings.  It evaluates them on the host with ``eval`` in the numpy namespace
(Propagator_Base.Potential / dPotential / fp_Damping, GPU_Wigner2D_FFT.py:674-692,
1646-1654) and splices them into CUDA kernels (:869-906).

Here the host side is kept (``host_eval``); on the device a string that is a
polynomial in its variable (coefficients may depend on ``t``) becomes a
coefficient table evaluated in-kernel by Horner's rule.
"""
from __future__ import annotations

import functools

import numpy as np

from ._lib import WH_MAX_DEG


class NonPolynomialPhysics(ValueError):
    """The expression is not a polynomial of degree <= WH_MAX_DEG in its variable."""


def heaviside_kernel(x):
    """The device function the reference splices into its kernels: 0 for x < 0, else 1
    (GPU_Wigner2D_GPitaevskii.py:34-40)."""
    return np.where(np.asarray(x) < 0.0, 0.0, 1.0)


def heaviside_host(x):
    """The host-side step of Propagator_Base.heaviside: 1/2 at x = 0
    (GPU_Wigner2D_GPitaevskii.py:675-685)."""
    x = np.asarray(x, dtype=np.float64)
    return np.where(x > 0.0, 1.0, np.where(x == 0.0, 0.5, 0.0))


def host_eval(expr: str, heaviside=heaviside_kernel, **variables):
    """eval of a reference physics string with numpy (GPU_Wigner2D_FFT.py:674-692,
    GPU_Wigner2D_GPitaevskii.py:687-697).  ``Heaviside`` follows the kernels' definition by
    default (values that the reference computes on the device); pass
    ``heaviside=heaviside_host`` for the values its host-side Potential computes."""
    ns = dict(variables)
    ns.update(pow=np.power, atan=np.arctan, M_E=np.e, M_PI=np.pi, sqrt=np.sqrt, exp=np.exp,
              fabs=np.abs, abs=np.abs, Heaviside=heaviside)
    return eval(expr, np.__dict__, ns)  # noqa: S307 - the reference's own semantics


def _sympy_expr(expr: str, var: str):
    import sympy as sp

    v, t = sp.symbols(f"{var} t", real=True)
    loc = {
        var: v, "t": t, "pow": sp.Pow, "sqrt": sp.sqrt, "exp": sp.exp, "sin": sp.sin,
        "cos": sp.cos, "tan": sp.tan, "tanh": sp.tanh, "atan": sp.atan, "log": sp.log,
        "fabs": sp.Abs, "M_E": sp.E, "M_PI": sp.pi, "Heaviside": sp.Heaviside,
    }
    return sp.sympify(expr, locals=loc), v, t


@functools.lru_cache(maxsize=256)
def polynomial_coefficients(expr: str, var: str = "x"):
    """Coefficients c_n(t) of expr = sum_n c_n(t) var^n (cached per string: ensembles
    parse the same expression for every member).

    Returns (degree, (callable(t) -> float, ...)).  Raises NonPolynomialPhysics.
    """
    import sympy as sp

    e, v, t = _sympy_expr(expr, var)
    try:
        poly = sp.Poly(sp.expand(e), v)
    except sp.PolynomialError as exc:
        raise NonPolynomialPhysics(f"{expr!r} is not a polynomial in {var}") from exc
    deg = poly.degree()
    if deg > WH_MAX_DEG:
        raise NonPolynomialPhysics(f"{expr!r}: degree {deg} > {WH_MAX_DEG}")
    if deg < 0:
        deg = 0
    coeffs = []
    for n in range(deg + 1):
        c = poly.coeff_monomial(v ** n)
        if c.free_symbols - {t}:
            raise NonPolynomialPhysics(f"{expr!r}: coefficient {c} depends on {c.free_symbols}")
        coeffs.append(sp.lambdify(t, c, "numpy"))
    return deg, tuple(coeffs)


def coefficient_table(expr: str, var: str, times) -> tuple[int, np.ndarray]:
    """[len(times), WH_MAX_DEG+1] table of the polynomial coefficients at each time."""
    deg, fns = polynomial_coefficients(expr, var)
    times = np.atleast_1d(np.asarray(times, dtype=np.float64))
    tab = np.zeros((len(times), WH_MAX_DEG + 1), dtype=np.float64)
    for n, f in enumerate(fns):
        tab[:, n] = np.broadcast_to(np.asarray(f(times), dtype=np.float64), times.shape)
    return deg, tab


@functools.lru_cache(maxsize=256)
def depends_on_time(expr: str, var: str = "x") -> bool:
    import sympy as sp

    e, v, t = _sympy_expr(expr, var)
    return t in e.free_symbols
