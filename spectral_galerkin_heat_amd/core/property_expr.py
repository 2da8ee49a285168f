"""Temperature expressions for material properties, compiled once to a postfix program.

Accepted language (the interface of the reference's ``PropertyExpr``,
``src/spectral_galerkin_heat/core/property_expr.py:69-121``): a bare number, or an
expression of ``T`` built from numeric literals, the names ``T``, ``pi`` and ``e``, the
binary operators ``+ - * / **``, unary ``+``/``-``, and calls of the math functions listed
in ``FUNC_IDS`` (``atan2`` with two arguments, the others with one, no keywords).
Anything else is refused with ``ValueError`` when the expression is built.

One representation serves the host and the device. The syntax tree is lowered, in one
pass that also does the checking, into a postfix program: a list of integer words
(an opcode, followed by an operand word for ``CONST`` and ``FN``) over a pool of
constants. The HIP property kernels interpret that program per cell
(``csrc/sgh_props.h``) and ``PropertyExpr.__call__`` interprets the same program on
NumPy values, so both sides evaluate the same sequence of operations. Powers by the
literals 2, 0.5, -1 and 1 become square, square root, reciprocal and nothing, which is
what NumPy's ``power`` does for those exponents.
"""

from __future__ import annotations

import ast
import math
import operator

import numpy as np

__all__ = ["PropertyExpr", "FUNC_IDS", "OPCODES"]

# Opcodes of the device program (the values are fixed by csrc/sgh_props.h).
OPCODES = {
    "CONST": 0, "T": 1, "ADD": 2, "SUB": 3, "MUL": 4, "DIV": 5, "POW": 6,
    "NEG": 7, "FN": 8, "ATAN2": 9, "SQR": 10, "CUBE": 11, "RECIP": 12,
}

# Whitelisted one-argument functions: device function id -> (name, NumPy ufunc).
_FUNCTIONS = (
    ("sqrt", np.sqrt), ("exp", np.exp), ("log", np.log), ("log2", np.log2),
    ("log10", np.log10), ("sin", np.sin), ("cos", np.cos), ("tan", np.tan),
    ("asin", np.arcsin), ("acos", np.arccos), ("atan", np.arctan), ("sinh", np.sinh),
    ("cosh", np.cosh), ("tanh", np.tanh), ("asinh", np.arcsinh), ("acosh", np.arccosh),
    ("atanh", np.arctanh), ("fabs", np.abs), ("ceil", np.ceil), ("floor", np.floor),
    ("trunc", np.trunc), ("log1p", np.log1p), ("expm1", np.expm1),
)
FUNC_IDS = {name: i for i, (name, _) in enumerate(_FUNCTIONS)}
_NAMED_CONSTANTS = {"pi": math.pi, "e": math.e}

# Device evaluator capacity (csrc/sgh_props.h: kProgMaxOps, kProgMaxConsts, kProgMaxStack).
MAX_OPS = 64
MAX_CONSTS = 24
MAX_STACK = 16

_BINARY_CODES = {ast.Add: OPCODES["ADD"], ast.Sub: OPCODES["SUB"], ast.Mult: OPCODES["MUL"],
                 ast.Div: OPCODES["DIV"], ast.Pow: OPCODES["POW"]}
# powers NumPy evaluates by a cheaper exact operation: exponent -> replacement words
_POWER_SHORTCUTS = {2.0: [OPCODES["SQR"]], 0.5: [OPCODES["FN"], FUNC_IDS["sqrt"]],
                    -1.0: [OPCODES["RECIP"]], 1.0: []}


class _Lowering:
    """Checks a syntax tree against the whitelist while emitting its postfix program."""

    def __init__(self, source: str):
        self.source = source
        self.words: list[int] = []
        self.pool: list[float] = []

    def refuse(self, what: str):
        raise ValueError(f"property expression {self.source!r}: {what} is not allowed")

    def constant(self, value) -> None:
        value = float(value)
        if value not in self.pool:
            self.pool.append(value)
        self.words += [OPCODES["CONST"], self.pool.index(value)]

    def lower(self, node) -> None:
        handler = getattr(self, "_" + type(node).__name__, None)
        if handler is None:
            self.refuse(f"a {type(node).__name__} node")
        handler(node)

    def _Constant(self, node):
        if type(node.value) not in (int, float):
            self.refuse(f"the literal {node.value!r} (numbers only)")
        self.constant(node.value)

    def _Name(self, node):
        if node.id == "T":
            self.words.append(OPCODES["T"])
        elif node.id in _NAMED_CONSTANTS:
            self.constant(_NAMED_CONSTANTS[node.id])
        else:
            self.refuse(f"the name {node.id!r} (T, pi and e only)")

    def _UnaryOp(self, node):
        if not isinstance(node.op, (ast.UAdd, ast.USub)):
            self.refuse(f"the unary operator {type(node.op).__name__}")
        self.lower(node.operand)
        if isinstance(node.op, ast.USub):
            self.words.append(OPCODES["NEG"])

    def _BinOp(self, node):
        code = _BINARY_CODES.get(type(node.op))
        if code is None:
            self.refuse(f"the operator {type(node.op).__name__}")
        right = node.right
        if (code == OPCODES["POW"] and isinstance(right, ast.Constant)
                and type(right.value) in (int, float) and float(right.value) in _POWER_SHORTCUTS):
            self.lower(node.left)
            self.words += _POWER_SHORTCUTS[float(right.value)]
            return
        self.lower(node.left)
        self.lower(right)
        self.words.append(code)

    def _Call(self, node):
        name = node.func.id if isinstance(node.func, ast.Name) else None
        arity = 2 if name == "atan2" else 1
        if (name is None or (name not in FUNC_IDS and name != "atan2") or node.keywords
                or len(node.args) != arity):
            self.refuse(f"the call of {name or 'an expression'!s} with "
                        f"{len(node.args)} argument(s)")
        for arg in node.args:
            self.lower(arg)
        if name == "atan2":
            self.words.append(OPCODES["ATAN2"])
        else:
            self.words += [OPCODES["FN"], FUNC_IDS[name]]


def _stack_depth(words) -> int:
    """Deepest evaluation stack the postfix program needs."""
    depth = deepest = 0
    i = 0
    while i < len(words):
        op = words[i]
        if op in (OPCODES["CONST"], OPCODES["FN"]):
            i += 1
        if op in (OPCODES["CONST"], OPCODES["T"]):
            depth += 1
        elif op in (OPCODES["ADD"], OPCODES["SUB"], OPCODES["MUL"], OPCODES["DIV"],
                    OPCODES["POW"], OPCODES["ATAN2"]):
            depth -= 1
        deepest = max(deepest, depth)
        i += 1
    return deepest


_HOST_BINARY = {OPCODES["ADD"]: operator.add, OPCODES["SUB"]: operator.sub,
                OPCODES["MUL"]: operator.mul, OPCODES["DIV"]: operator.truediv,
                OPCODES["POW"]: operator.pow, OPCODES["ATAN2"]: np.arctan2}
_HOST_UNARY = {OPCODES["NEG"]: operator.neg, OPCODES["SQR"]: lambda v: v * v,
               OPCODES["CUBE"]: lambda v: v * v * v, OPCODES["RECIP"]: lambda v: 1.0 / v}


def _interpret(words, pool, T):
    """Run a postfix program on the host: T is a Python scalar or a NumPy array."""
    stack = []
    i = 0
    while i < len(words):
        op = words[i]
        if op == OPCODES["CONST"]:
            i += 1
            stack.append(pool[words[i]])
        elif op == OPCODES["T"]:
            stack.append(T)
        elif op == OPCODES["FN"]:
            i += 1
            stack[-1] = _FUNCTIONS[words[i]][1](stack[-1])
        elif op in _HOST_UNARY:
            stack[-1] = _HOST_UNARY[op](stack[-1])
        else:
            rhs = stack.pop()
            stack[-1] = _HOST_BINARY[op](stack[-1], rhs)
        i += 1
    return stack[0]


class PropertyExpr:
    """A property branch: a bare number, or a whitelisted expression of ``T``."""

    __slots__ = ("source", "_value", "_words", "_pool")

    def __init__(self, expr):
        self.source = str(expr).strip()
        try:
            self._value = float(self.source)
        except ValueError:
            self._value = None
        if self._value is not None:
            self._words, self._pool = [OPCODES["CONST"], 0], [self._value]
            return
        try:
            tree = ast.parse(self.source, mode="eval")
        except SyntaxError as exc:
            raise ValueError(f"property expression {self.source!r} does not parse: "
                             f"{exc.msg}") from None
        low = _Lowering(self.source)
        low.lower(tree.body)
        self._words, self._pool = low.words, low.pool

    @property
    def is_constant(self) -> bool:
        """True for a bare number (the constant-property fast path)."""
        return self._value is not None

    @property
    def constant_value(self) -> float:
        if self._value is None:
            raise ValueError(f"property expression {self.source!r} depends on T")
        return self._value

    def __call__(self, T):
        """Value at ``T`` (a Python scalar or a NumPy array); a constant keeps T's shape."""
        if self._value is not None:
            return T * 0.0 + self._value
        return _interpret(self._words, self._pool, T)

    def to_program(self) -> tuple[list[int], list[float]]:
        """The device program ``(words, constants)``; refused when it exceeds the device
        evaluator's capacity."""
        depth = _stack_depth(self._words)
        if len(self._words) > MAX_OPS or len(self._pool) > MAX_CONSTS or depth > MAX_STACK:
            raise ValueError(
                f"property expression {self.source!r} needs {len(self._words)} program "
                f"words, {len(self._pool)} constants and a stack of {depth}; the device "
                f"evaluator holds {MAX_OPS}, {MAX_CONSTS} and {MAX_STACK}")
        return list(self._words), list(self._pool)

    def __repr__(self):
        return f"PropertyExpr({self.source!r})"
