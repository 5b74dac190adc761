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

from ._lib import WH_EXPR_MAX, WH_EXPR_STACK, WH_MAX_DEG


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


# ---------------------------------------------------------------- expression programs
# Op codes of the device evaluator (WH_EXPR_* of include/wignerhip.h).
EXPR_OPS = {"const": 0, "x": 1, "t": 2, "neg": 3, "add": 4, "sub": 5, "mul": 6, "div": 7, "pow": 8,
            "exp": 9, "sin": 10, "cos": 11, "tan": 12, "sqrt": 13, "fabs": 14, "tanh": 15, "sinh": 16,
            "cosh": 17, "log": 18, "atan": 19, "Heaviside": 20}
_EXPR_FUNCS = {"exp", "sin", "cos", "tan", "sqrt", "fabs", "tanh", "sinh", "cosh", "log", "atan", "Heaviside"}
_EXPR_NP = {"exp": np.exp, "sin": np.sin, "cos": np.cos, "tan": np.tan, "sqrt": np.sqrt, "fabs": np.abs,
            "tanh": np.tanh, "sinh": np.sinh, "cosh": np.cosh, "log": np.log, "atan": np.arctan,
            "Heaviside": heaviside_kernel}


def compile_expression(expr: str, var: str = "x") -> np.ndarray:
    """The device program of a physics string in `var` and t: [n, 2] float64 (op, value) pairs
    for the stack evaluator of the p sweep (pot_deg = WH_POT_EXPR), constant subexpressions
    folded.  Accepts what the reference's kernels accept of C: numbers, + - * /, unary minus,
    pow, the functions of EXPR_OPS (abs as fabs), M_PI and M_E.  Raises NonPolynomialPhysics
    for anything else or for a program beyond WH_EXPR_MAX ops or WH_EXPR_STACK stack values."""
    import ast

    consts = {"M_PI": np.pi, "M_E": np.e}
    code: list[tuple[int, float]] = []

    def fold(node):
        """A Python float when the subtree is constant, else None."""
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return float(node.value)
        if isinstance(node, ast.Name) and node.id in consts:
            return float(consts[node.id])
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            v = fold(node.operand)
            return None if v is None else (-v if isinstance(node.op, ast.USub) else v)
        if isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Add, ast.Sub, ast.Mult, ast.Div, ast.Pow)):
            a, b = fold(node.left), fold(node.right)
            if a is None or b is None:
                return None
            return {ast.Add: a + b, ast.Sub: a - b, ast.Mult: a * b, ast.Div: a / b if b != 0 else None,
                    ast.Pow: float(np.power(a, b))}[type(node.op)]
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Name):
            args = [fold(a) for a in node.args]
            if any(a is None for a in args):
                return None
            name = "fabs" if node.func.id == "abs" else node.func.id
            if name == "pow" and len(args) == 2:
                return float(np.power(args[0], args[1]))
            if name in _EXPR_FUNCS and len(args) == 1:
                return float(_EXPR_NP[name](args[0]))
        return None

    def emit_pow(base, expo):
        """pow(base, k) of a leaf base (x, t) and an integer 1 <= k <= 4 as k - 1 products
        (within an ulp of pow, and without its cost on the device); else the pow op."""
        k = fold(expo)
        leaf = isinstance(base, ast.Name) and base.id in (var, "t")
        if leaf and k is not None and float(k).is_integer() and 1 <= k <= 4:
            for _ in range(int(k)):
                emit(base)
            for _ in range(int(k) - 1):
                code.append((EXPR_OPS["mul"], 0.0))
            return
        emit(base)
        emit(expo)
        code.append((EXPR_OPS["pow"], 0.0))

    def emit(node):
        v = fold(node)
        if v is not None:
            code.append((EXPR_OPS["const"], v))
            return
        if isinstance(node, ast.Name):
            if node.id == var:
                code.append((EXPR_OPS["x"], 0.0))
            elif node.id == "t":
                code.append((EXPR_OPS["t"], 0.0))
            else:
                raise NonPolynomialPhysics(f"{expr!r}: unknown name {node.id!r}")
        elif isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            emit(node.operand)
            if isinstance(node.op, ast.USub):
                code.append((EXPR_OPS["neg"], 0.0))
        elif isinstance(node, ast.BinOp) and isinstance(node.op, ast.Pow):
            emit_pow(node.left, node.right)
        elif isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Add, ast.Sub, ast.Mult, ast.Div)):
            emit(node.left)
            emit(node.right)
            code.append((EXPR_OPS[{ast.Add: "add", ast.Sub: "sub", ast.Mult: "mul", ast.Div: "div"}[type(node.op)]],
                         0.0))
        elif isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and not node.keywords:
            name = "fabs" if node.func.id == "abs" else node.func.id
            if name == "pow" and len(node.args) == 2:
                emit_pow(node.args[0], node.args[1])
            elif name in _EXPR_FUNCS and len(node.args) == 1:
                emit(node.args[0])
                code.append((EXPR_OPS[name], 0.0))
            else:
                raise NonPolynomialPhysics(f"{expr!r}: unsupported call {node.func.id}")
        else:
            raise NonPolynomialPhysics(f"{expr!r}: unsupported syntax {ast.dump(node)[:60]}")

    try:
        tree = ast.parse(expr.strip(), mode="eval")
    except SyntaxError as exc:
        raise NonPolynomialPhysics(f"{expr!r}: {exc}") from exc
    emit(tree.body)
    depth = peak = 0
    for op, _ in code:
        depth += 1 if op <= EXPR_OPS["t"] else -1 if EXPR_OPS["add"] <= op <= EXPR_OPS["pow"] else 0
        peak = max(peak, depth)
    if len(code) > WH_EXPR_MAX or peak > WH_EXPR_STACK:
        raise NonPolynomialPhysics(f"{expr!r}: {len(code)} ops, stack depth {peak} "
                                   f"(at most {WH_EXPR_MAX} and {WH_EXPR_STACK})")
    return np.ascontiguousarray(np.array(code, dtype=np.float64).reshape(-1, 2))


def eval_program(code, x, t):
    """numpy evaluation of a program of compile_expression, op by op as the device runs it."""
    st = []
    inv = {v: k for k, v in EXPR_OPS.items()}
    for op, v in np.asarray(code):
        name = inv[int(op)]
        if name == "const":
            st.append(np.float64(v))
        elif name == "x":
            st.append(np.asarray(x, dtype=np.float64))
        elif name == "t":
            st.append(np.float64(t))
        elif name == "neg":
            st[-1] = -st[-1]
        elif name in ("add", "sub", "mul", "div", "pow"):
            b = st.pop()
            a = st.pop()
            st.append({"add": np.add, "sub": np.subtract, "mul": np.multiply, "div": np.divide,
                       "pow": np.power}[name](a, b))
        else:
            st[-1] = _EXPR_NP[name](st[-1])
    return st[0]
