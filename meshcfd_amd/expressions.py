"""Safe mathematical expressions of the cell centroid (x, y) for
initial-condition blocks (reference `src/meshcfd/core/expressions.py`, and
`initial_conditions.py` for their evaluation).

Grammar and checks follow the reference: Python syntax with "^" read as
"**"; numbers, x, y, e, pi; + - * / % ** and unary +/-; comparisons, and /
or / not (domains only); the functions acos, acosh, asin, asinh, atan,
atan2, atanh, ceil, cos, cosh, degrees, exp, fabs, floor, hypot, log,
log10, radians, sin, sinh, sqrt, tan, tanh, abs, max, min, pow; exponents
of magnitude above 1024 are rejected; at most 512 characters and 128
syntax nodes. A domain must be a boolean predicate, a field must be
numeric.

Evaluation is over arrays: one pass of the syntax tree evaluates an
expression at every cell at once. Arithmetic runs in NumPy (the same IEEE
operations as Python floats), the functions call `math` element by element,
so each cell gets the value the reference's scalar evaluation gives.
"""

from __future__ import annotations

import ast
import math

import numpy as np


class ExpressionError(ValueError):
    """Raised when a mathematical expression is invalid or cannot be evaluated."""


MAX_LENGTH = 512
MAX_NODES = 128
FUNCTIONS = ("acos", "acosh", "asin", "asinh", "atan", "atan2", "atanh", "ceil", "cos", "cosh", "degrees",
             "exp", "fabs", "floor", "hypot", "log", "log10", "radians", "sin", "sinh", "sqrt", "tan", "tanh",
             "abs", "max", "min", "pow")
CONSTANTS = {"e": math.e, "pi": math.pi}
_BIN = (ast.Add, ast.Sub, ast.Mult, ast.Div, ast.Pow, ast.Mod)
_CMP = (ast.Eq, ast.NotEq, ast.Lt, ast.LtE, ast.Gt, ast.GtE)


def _check(node: ast.AST) -> None:
    """Whitelist of syntax (the reference's `_validate_node`)."""
    if isinstance(node, ast.Constant):
        if not isinstance(node.value, (int, float, bool)):
            raise ExpressionError("Unsupported constant in expression.")
        return
    if isinstance(node, ast.Name):
        if node.id not in ("x", "y") and node.id not in CONSTANTS:
            raise ExpressionError(f"Unknown expression name '{node.id}'.")
        return
    if isinstance(node, ast.BinOp):
        if not isinstance(node.op, _BIN):
            raise ExpressionError("Unsupported arithmetic operator.")
        _check(node.left)
        _check(node.right)
        return
    if isinstance(node, ast.UnaryOp):
        if not isinstance(node.op, (ast.UAdd, ast.USub, ast.Not)):
            raise ExpressionError("Unsupported unary operator.")
        _check(node.operand)
        return
    if isinstance(node, ast.BoolOp):
        for v in node.values:
            _check(v)
        return
    if isinstance(node, ast.Compare):
        if not all(isinstance(op, _CMP) for op in node.ops):
            raise ExpressionError("Unsupported comparison operator.")
        _check(node.left)
        for c in node.comparators:
            _check(c)
        return
    if isinstance(node, ast.Call):
        if not isinstance(node.func, ast.Name) or node.func.id not in FUNCTIONS or node.keywords:
            raise ExpressionError("Unsupported function call.")
        if not node.args:
            raise ExpressionError("Function calls need arguments.")
        for a in node.args:
            if isinstance(a, ast.Starred):
                raise ExpressionError("Unsupported function call.")
            _check(a)
        return
    raise ExpressionError("Unsupported expression syntax.")


def _is_bool(v) -> bool:
    return isinstance(v, (bool, np.bool_)) or (isinstance(v, np.ndarray) and v.dtype == bool)


def _elementwise(fn, *args):
    """math function applied per element (scalars stay scalars)."""
    if not any(isinstance(a, np.ndarray) for a in args):
        return fn(*args)
    shape = np.broadcast_shapes(*(np.shape(a) for a in args))
    cols = [np.broadcast_to(np.asarray(a), shape).ravel() for a in args]
    out = [fn(*vals) for vals in zip(*(c.tolist() for c in cols))]
    return np.asarray(out, dtype=np.float64).reshape(shape)


def _power(a, b):
    if np.any(np.abs(b) > 1024):
        raise ValueError("Exponent magnitude must not exceed 1024.")
    if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
        return _elementwise(lambda p, q: p ** q, a, b)
    return a ** b


def _call(name: str, args):
    if name == "pow":
        return _power(*args)
    if name == "abs":
        return np.abs(args[0]) if isinstance(args[0], np.ndarray) else abs(args[0])
    if name in ("max", "min"):
        if len(args) == 1:
            raise TypeError(f"{name} expected more than one number")
        if not any(isinstance(a, np.ndarray) for a in args):
            return (max if name == "max" else min)(*args)
        red = np.maximum if name == "max" else np.minimum
        out = np.asarray(args[0], dtype=np.float64)
        for a in args[1:]:
            # Python's max keeps the first of equal values; fmax-style NaN
            # handling never arises (non-finite values are rejected)
            out = red(out, a)
        return out
    return _elementwise(getattr(math, name), *args)


class MathExpression:
    def __init__(self, source: str, tree: ast.Expression, boolean: bool):
        self.source = source
        self.tree = tree
        self.boolean = boolean

    def evaluate(self, x, y):
        """Value at one point (floats) or at every point (arrays)."""
        with np.errstate(all="ignore"):
            try:
                result = self._eval(self.tree.body, {"x": x, "y": y})
            except (ArithmeticError, TypeError, ValueError) as error:
                raise ExpressionError(f"Could not evaluate expression '{self.source}': {error}") from error
        if self.boolean:
            if not _is_bool(result):
                raise ExpressionError(f"Domain expression '{self.source}' must evaluate to true or false.")
            return np.broadcast_to(result, np.shape(x)).copy() if isinstance(x, np.ndarray) else bool(result)
        if _is_bool(result):
            raise ExpressionError(f"Field expression '{self.source}' must evaluate to a number.")
        value = np.broadcast_to(np.asarray(result, dtype=np.float64), np.shape(x)).copy()
        if not np.isfinite(value).all():
            raise ExpressionError(f"Field expression '{self.source}' must evaluate to a finite number.")
        return value if isinstance(x, np.ndarray) else float(value)

    def _eval(self, node, env):
        if isinstance(node, ast.Constant):
            return node.value
        if isinstance(node, ast.Name):
            return env[node.id] if node.id in env else CONSTANTS[node.id]
        if isinstance(node, ast.BinOp):
            a, b = self._eval(node.left, env), self._eval(node.right, env)
            op = node.op
            if isinstance(op, ast.Add):
                return a + b
            if isinstance(op, ast.Sub):
                return a - b
            if isinstance(op, ast.Mult):
                return a * b
            # over arrays a zero divisor gives a non-finite value (a field
            # then fails its finiteness check, a comparison is false): the
            # reference's scalar evaluation may skip such cells behind a
            # false 'and' operand, so an array cannot raise for them
            if isinstance(op, ast.Div):
                return np.true_divide(a, b) if isinstance(a, np.ndarray) or isinstance(b, np.ndarray) else a / b
            if isinstance(op, ast.Mod):
                return np.mod(a, b) if isinstance(a, np.ndarray) or isinstance(b, np.ndarray) else a % b
            return _power(a, b)
        if isinstance(node, ast.UnaryOp):
            v = self._eval(node.operand, env)
            if isinstance(node.op, ast.Not):
                if not _is_bool(v):
                    raise ExpressionError("The 'not' operator needs a boolean value.")
                return ~v if isinstance(v, np.ndarray) else not v
            return -v if isinstance(node.op, ast.USub) else +v
        if isinstance(node, ast.BoolOp):
            values = [self._eval(v, env) for v in node.values]
            name = "and" if isinstance(node.op, ast.And) else "or"
            if not all(_is_bool(v) for v in values):
                raise ExpressionError(f"The '{name}' operator needs boolean values.")
            out = values[0]
            for v in values[1:]:
                out = (out & v) if name == "and" else (out | v)
            return out
        if isinstance(node, ast.Compare):
            left = self._eval(node.left, env)
            out = True
            for op, comp in zip(node.ops, node.comparators):
                right = self._eval(comp, env)
                r = {ast.Eq: np.equal, ast.NotEq: np.not_equal, ast.Lt: np.less, ast.LtE: np.less_equal,
                     ast.Gt: np.greater, ast.GtE: np.greater_equal}[type(op)](left, right)
                out = out & r
                left = right
            return out if isinstance(out, np.ndarray) else bool(out)
        if isinstance(node, ast.Call):
            return _call(node.func.id, [self._eval(a, env) for a in node.args])
        raise ExpressionError("Unsupported expression syntax.")


def compile_expression(source: str, *, boolean: bool = False) -> MathExpression:
    if not isinstance(source, str) or not source.strip():
        raise ExpressionError("Expression must not be empty.")
    if len(source) > MAX_LENGTH:
        raise ExpressionError("Expression is too long.")
    try:
        tree = ast.parse(source.replace("^", "**"), mode="eval")
    except SyntaxError as error:
        raise ExpressionError(f"Invalid expression '{source}': {error.msg}.") from error
    nodes = list(ast.walk(tree))
    if len(nodes) > MAX_NODES:
        raise ExpressionError("Expression is too complex.")
    if not boolean and any(
        (isinstance(n, ast.Constant) and isinstance(n.value, bool))
        or (isinstance(n, ast.UnaryOp) and isinstance(n.op, ast.Not)) for n in nodes
    ):
        raise ExpressionError("Field expression must be numeric.")
    _check(tree.body)
    body = tree.body
    predicate = (isinstance(body, (ast.Compare, ast.BoolOp))
                 or (isinstance(body, ast.UnaryOp) and isinstance(body.op, ast.Not))
                 or (isinstance(body, ast.Constant) and isinstance(body.value, bool)))
    if boolean and not predicate:
        raise ExpressionError("Domain expression must be a boolean predicate.")
    if not boolean and predicate:
        raise ExpressionError("Field expression must be numeric.")
    return MathExpression(source, tree, boolean)


def evaluate_initial_conditions(centroids: np.ndarray, settings) -> np.ndarray:
    """(C, 6) initial states (rho, u, v, p, k, omega) at the cell centroids:
    the base state, replaced in each cell by the last block (in list order)
    whose domain holds there (reference `initial_conditions.py:15-53`)."""
    x = np.ascontiguousarray(centroids[:, 0], dtype=np.float64)
    y = np.ascontiguousarray(centroids[:, 1], dtype=np.float64)
    base = settings.base_state
    _validate_state(base)
    n = len(x)
    out = np.tile([base.density, base.velocity_x, base.velocity_y, base.pressure, base.turbulence_k,
                   base.turbulence_omega], (n, 1)).astype(np.float64)
    todo = np.ones(n, dtype=bool)
    for block in reversed(settings.blocks):
        domain = compile_expression(block.domain, boolean=True)
        fields = [compile_expression(getattr(block, k)) for k in ("density", "velocity_x", "velocity_y", "pressure")]
        try:
            hit = domain.evaluate(x, y) & todo
        except ExpressionError as error:
            raise ExpressionError(f"Initial-condition block '{block.name}': {error}") from error
        if not hit.any():
            continue
        idx = np.flatnonzero(hit)
        try:
            vals = [f.evaluate(x[idx], y[idx]) for f in fields]
        except ExpressionError as error:
            raise ExpressionError(f"Initial-condition block '{block.name}', cell {int(idx[0])}: {error}") from error
        rho, u, v, p = vals
        bad = ~((rho > 0.0) & (p > 0.0))
        if bad.any():
            cell = int(idx[np.flatnonzero(bad)[0]])
            raise ExpressionError(f"Initial-condition block '{block.name}', cell {cell}: "
                                  "Initial density and pressure must be positive.")
        out[idx, 0], out[idx, 1], out[idx, 2], out[idx, 3] = rho, u, v, p
        todo &= ~hit
    return out


def _validate_state(state) -> None:
    values = (state.density, state.velocity_x, state.velocity_y, state.pressure, state.turbulence_k,
              state.turbulence_omega)
    if not all(math.isfinite(v) for v in values):
        raise ValueError("Initial-state values must be finite.")
    if state.density <= 0.0 or state.pressure <= 0.0:
        raise ValueError("Initial density and pressure must be positive.")
    if state.turbulence_k < 0.0 or state.turbulence_omega <= 0.0:
        raise ValueError("Initial turbulence k must be non-negative and omega positive.")
