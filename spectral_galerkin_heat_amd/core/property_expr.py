"""Whitelisted temperature expressions for material properties.

Same accepted language as the reference's ``PropertyExpr``
(``src/spectral_galerkin_heat/core/property_expr.py:69-147,226-300``): a bare number,
or an expression of ``T`` using ``+ - * / **``, unary ``+``/``-``, the constants ``pi``
and ``e`` and a fixed list of math functions. Anything else raises ``ValueError`` at
construction.

The reference generates Numba/CUDA source text per expression. Here an expression is
compiled once into a small stack program (``to_program``) that the HIP property kernels
interpret per cell, so one compiled library serves any config without runtime codegen.
"""

from __future__ import annotations

import ast
import math

import numpy as np

__all__ = ["PropertyExpr", "FUNC_IDS", "OPCODES"]

# function name -> numpy ufunc name (host evaluation)
_NUMPY_FUNC_NAME = {
    "sqrt": "sqrt", "exp": "exp", "log": "log", "log2": "log2", "log10": "log10",
    "sin": "sin", "cos": "cos", "tan": "tan",
    "asin": "arcsin", "acos": "arccos", "atan": "arctan", "atan2": "arctan2",
    "sinh": "sinh", "cosh": "cosh", "tanh": "tanh",
    "asinh": "arcsinh", "acosh": "arccosh", "atanh": "arctanh",
    "fabs": "abs", "ceil": "ceil", "floor": "floor", "trunc": "trunc",
    "log1p": "log1p", "expm1": "expm1",
}
_ALLOWED_CONSTANTS = {"pi": math.pi, "e": math.e}
_ALLOWED_BINOPS = (ast.Add, ast.Sub, ast.Mult, ast.Div, ast.Pow)
_ALLOWED_UNARYOPS = (ast.UAdd, ast.USub)

# Device stack-program encoding (must match sgh_props.h).
OPCODES = {
    "CONST": 0, "T": 1, "ADD": 2, "SUB": 3, "MUL": 4, "DIV": 5, "POW": 6,
    "NEG": 7, "FN": 8, "ATAN2": 9, "SQR": 10, "CUBE": 11, "RECIP": 12,
}
FUNC_IDS = {name: i for i, name in enumerate(
    ["sqrt", "exp", "log", "log2", "log10", "sin", "cos", "tan", "asin", "acos",
     "atan", "sinh", "cosh", "tanh", "asinh", "acosh", "atanh", "fabs", "ceil",
     "floor", "trunc", "log1p", "expm1"])}
MAX_OPS = 64
MAX_CONSTS = 24
MAX_STACK = 16


def _validate(node):
    if isinstance(node, ast.Constant):
        if isinstance(node.value, bool) or not isinstance(node.value, (int, float)):
            raise ValueError(f"Only numeric literals are allowed, got {node.value!r}.")
        return
    if isinstance(node, ast.Name):
        if node.id != "T" and node.id not in _ALLOWED_CONSTANTS:
            raise ValueError(
                f"Unknown name {node.id!r} in property expression "
                f"(only 'T' and {sorted(_ALLOWED_CONSTANTS)} are allowed).")
        return
    if isinstance(node, ast.BinOp):
        if not isinstance(node.op, _ALLOWED_BINOPS):
            raise ValueError(f"Operator {type(node.op).__name__} is not allowed.")
        _validate(node.left)
        _validate(node.right)
        return
    if isinstance(node, ast.UnaryOp):
        if not isinstance(node.op, _ALLOWED_UNARYOPS):
            raise ValueError(f"Unary operator {type(node.op).__name__} is not allowed.")
        _validate(node.operand)
        return
    if isinstance(node, ast.Call):
        if (not isinstance(node.func, ast.Name) or node.func.id not in _NUMPY_FUNC_NAME
                or node.keywords or not node.args):
            name = getattr(node.func, "id", "<expr>")
            raise ValueError(f"Function {name!r} is not allowed in a property expression.")
        want = 2 if node.func.id == "atan2" else 1
        if len(node.args) != want:
            raise ValueError(f"{node.func.id} takes {want} argument(s).")
        for arg in node.args:
            _validate(arg)
        return
    raise ValueError(f"{type(node).__name__} is not allowed in a property expression.")


def _eval(node, T):
    if isinstance(node, ast.Constant):
        return node.value
    if isinstance(node, ast.Name):
        return T if node.id == "T" else _ALLOWED_CONSTANTS[node.id]
    if isinstance(node, ast.BinOp):
        left, right = _eval(node.left, T), _eval(node.right, T)
        op = type(node.op)
        if op is ast.Add:
            return left + right
        if op is ast.Sub:
            return left - right
        if op is ast.Mult:
            return left * right
        if op is ast.Div:
            return left / right
        return left ** right
    if isinstance(node, ast.UnaryOp):
        val = _eval(node.operand, T)
        return -val if isinstance(node.op, ast.USub) else val
    fn = getattr(np, _NUMPY_FUNC_NAME[node.func.id])
    return fn(*(_eval(a, T) for a in node.args))


class _Compiler:
    def __init__(self):
        self.ops: list[int] = []
        self.consts: list[float] = []
        self.depth = 0
        self.max_depth = 0

    def _push(self):
        self.depth += 1
        self.max_depth = max(self.max_depth, self.depth)

    def const(self, v: float):
        v = float(v)
        if v in self.consts:
            idx = self.consts.index(v)
        else:
            idx = len(self.consts)
            self.consts.append(v)
        self.ops += [OPCODES["CONST"], idx]
        self._push()

    def emit(self, node):
        if isinstance(node, ast.Constant):
            self.const(node.value)
        elif isinstance(node, ast.Name):
            if node.id == "T":
                self.ops.append(OPCODES["T"])
                self._push()
            else:
                self.const(_ALLOWED_CONSTANTS[node.id])
        elif isinstance(node, ast.BinOp):
            if isinstance(node.op, ast.Pow) and isinstance(node.right, ast.Constant) \
                    and float(node.right.value) in (2.0, 0.5, -1.0, 1.0):
                # numpy's scalar-power fast paths: square, sqrt, reciprocal, identity.
                p = float(node.right.value)
                self.emit(node.left)
                if p == 2.0:
                    self.ops.append(OPCODES["SQR"])
                elif p == 0.5:
                    self.ops += [OPCODES["FN"], FUNC_IDS["sqrt"]]
                elif p == -1.0:
                    self.ops.append(OPCODES["RECIP"])
                return
            self.emit(node.left)
            self.emit(node.right)
            op = type(node.op)
            code = {ast.Add: "ADD", ast.Sub: "SUB", ast.Mult: "MUL",
                    ast.Div: "DIV", ast.Pow: "POW"}[op]
            self.ops.append(OPCODES[code])
            self.depth -= 1
        elif isinstance(node, ast.UnaryOp):
            self.emit(node.operand)
            if isinstance(node.op, ast.USub):
                self.ops.append(OPCODES["NEG"])
        else:  # ast.Call
            for a in node.args:
                self.emit(a)
            if node.func.id == "atan2":
                self.ops.append(OPCODES["ATAN2"])
                self.depth -= 1
            else:
                self.ops += [OPCODES["FN"], FUNC_IDS[node.func.id]]


class PropertyExpr:
    """A validated math expression of ``T`` (or a bare constant)."""

    __slots__ = ("_const", "_tree", "source")

    def __init__(self, expr):
        self.source = str(expr).strip()
        try:
            self._const = float(self.source)
            self._tree = None
            return
        except ValueError:
            pass
        try:
            tree = ast.parse(self.source, mode="eval").body
        except SyntaxError as exc:
            raise ValueError(f"Invalid property expression {self.source!r}: {exc}") from None
        _validate(tree)
        self._const = None
        self._tree = tree

    @property
    def is_constant(self) -> bool:
        return self._const is not None

    @property
    def constant_value(self) -> float:
        if self._const is None:
            raise ValueError(f"{self.source!r} is not a constant expression.")
        return self._const

    def __call__(self, T):
        """Evaluate on the host (Python scalar or NumPy array)."""
        if self._const is not None:
            return T * 0.0 + self._const
        return _eval(self._tree, T)

    def to_program(self) -> tuple[list[int], list[float]]:
        """Compile to the device stack program ``(ops, consts)``."""
        c = _Compiler()
        if self._const is not None:
            c.const(self._const)
        else:
            c.emit(self._tree)
        if len(c.ops) > MAX_OPS or len(c.consts) > MAX_CONSTS or c.max_depth > MAX_STACK:
            raise ValueError(
                f"Property expression {self.source!r} is too large for the device "
                f"evaluator (ops {len(c.ops)}/{MAX_OPS}, constants "
                f"{len(c.consts)}/{MAX_CONSTS}, stack {c.max_depth}/{MAX_STACK}).")
        return c.ops, c.consts

    def __repr__(self):
        return f"PropertyExpr({self.source!r})"
