"""Initial-condition expressions of ALPACA input files, evaluated over arrays.

The reference compiles each ``<initialConditions><materialN>`` text with the
exprtk library through ``UserExpression`` (user_expression.cpp:33-79): the
cell-centre coordinates x, y, z are input variables, the prime-state names
(density, velocityX, velocityY, velocityZ, pressure) are output variables
created at zero, unknown symbols resolve to new zero variables, and exprtk's
constants (pi, epsilon, inf) are registered.  This module implements the
subset of exprtk's language that input files use, evaluated for every cell at
once with numpy:

  statements   ``name := expr;``  (also ``name : = expr`` as exprtk's
               tokenizer joins ``:`` ``=``), ``var name := expr;``,
               ``if (cond) { ... } else if (cond) { ... } else { ... }``
               (a single statement may stand without braces), ``{ ... }``
  expressions  ``?:``; ``or | || xor nor``; ``and & && nand``;
               ``< <= > >= == = != <>``; ``+ -``; ``* / %``; unary ``- + not``;
               ``^`` (right-associative, binds tighter than unary minus);
               calls of abs, acos, acosh, asin, asinh, atan, atan2, cbrt, ceil,
               clamp, cos, cosh, erf, exp, expm1, floor, frac, hypot, log,
               log10, log1p, log2, max, min, pow, round, sgn, sin, sinh, sqrt,
               tan, tanh, trunc, not, inrange; numbers; variables.

Branches are evaluated for all cells and merged by their condition mask, so
an ``if`` costs both branches; a comparison or logical operator yields 1.0 or
0.0 as in exprtk.  ``rand`` (a random draw per evaluation) is rejected: a
seeded state is drawn by the caller instead.
"""
from __future__ import annotations

import math
import re

import numpy as np

__all__ = ["ExpressionError", "Program", "compile_expression", "evaluate"]


class ExpressionError(ValueError):
    pass


_TOKEN = re.compile(r"""
    (?P<ws>\s+|//[^\n]*|\#[^\n]*|/\*.*?\*/)
  | (?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?)
  | (?P<name>[A-Za-z_][A-Za-z0-9_.]*)
  | (?P<op>:=|==|!=|<>|<=|>=|&&|\|\||[-+*/%^<>=!?:;,(){}\[\]&|])
""", re.VERBOSE | re.DOTALL)


def _tokenize(text):
    toks = []
    pos = 0
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise ExpressionError(f"unexpected character {text[pos]!r} at {pos}")
        pos = m.end()
        kind = m.lastgroup
        if kind == "ws":
            continue
        val = m.group(kind)
        # exprtk joins ':' '=' written apart into the assignment operator.
        if kind == "op" and val == "=" and toks and toks[-1] == ("op", ":"):
            toks[-1] = ("op", ":=")
            continue
        toks.append((kind, val))
    toks.append(("end", ""))
    return toks


_FUNCS1 = {
    "abs": np.abs, "acos": np.arccos, "acosh": np.arccosh, "asin": np.arcsin,
    "asinh": np.arcsinh, "atan": np.arctan, "cbrt": np.cbrt, "ceil": np.ceil,
    "cos": np.cos, "cosh": np.cosh, "exp": np.exp, "expm1": np.expm1, "floor": np.floor,
    "log": np.log, "log10": np.log10, "log1p": np.log1p, "log2": np.log2, "sin": np.sin,
    "sinh": np.sinh, "sqrt": np.sqrt, "tan": np.tan, "tanh": np.tanh, "trunc": np.trunc,
    "sgn": np.sign, "frac": lambda a: a - np.trunc(a),
    "round": lambda a: np.where(a < 0, -np.floor(-a + 0.5), np.floor(a + 0.5)),
    "erf": np.vectorize(math.erf, otypes=[float]),
    "not": lambda a: (a == 0).astype(float),
}
_FUNCS2 = {
    "atan2": np.arctan2, "hypot": np.hypot, "pow": np.power,
    "min": np.minimum, "max": np.maximum,
}
_CONSTANTS = {"pi": math.pi, "epsilon": np.finfo(float).eps, "inf": math.inf}


class _Parser:
    def __init__(self, text):
        self.toks = _tokenize(text)
        self.i = 0

    def peek(self, k=0):
        return self.toks[self.i + k]

    def take(self):
        t = self.toks[self.i]
        self.i += 1
        return t

    def accept(self, val):
        if self.peek()[1] == val and self.peek()[0] in ("op", "name"):
            self.i += 1
            return True
        return False

    def expect(self, val):
        if not self.accept(val):
            raise ExpressionError(f"expected {val!r}, found {self.peek()[1]!r}")

    # ---- statements
    def program(self):
        body = self.statements(until="end")
        return ("block", body)

    def statements(self, until):
        body = []
        while True:
            t = self.peek()
            if (until == "end" and t[0] == "end") or (until == "}" and t == ("op", "}")):
                return body
            if t[0] == "end":
                raise ExpressionError("unexpected end of expression")
            if self.accept(";"):
                continue
            body.append(self.statement())

    def statement(self):
        t = self.peek()
        if t == ("name", "if"):
            return self.if_statement()
        if t == ("op", "{"):
            self.take()
            body = self.statements(until="}")
            self.expect("}")
            return ("block", body)
        if t == ("name", "var"):
            self.take()
        if self.peek()[0] == "name" and self.peek(1) == ("op", ":="):
            name = self.take()[1]
            self.take()
            return ("assign", name, self.expr())
        return ("expr", self.expr())

    def if_statement(self):
        self.expect("if")
        self.expect("(")
        cond = self.expr()
        self.expect(")")
        then = self.statement()
        other = None
        while self.accept(";"):
            pass
        if self.accept("else"):
            other = self.statement()
        return ("if", cond, then, other)

    # ---- expressions (lowest precedence first)
    def expr(self):
        cond = self.logical_or()
        if self.accept("?"):
            a = self.expr()
            self.expect(":")
            b = self.expr()
            return ("ternary", cond, a, b)
        return cond

    def logical_or(self):
        node = self.logical_and()
        while self.peek()[1] in ("or", "|", "||", "xor", "nor"):
            op = self.take()[1]
            node = ({"|": "or", "||": "or"}.get(op, op), node, self.logical_and())
        return node

    def logical_and(self):
        node = self.comparison()
        while self.peek()[1] in ("and", "&", "&&", "nand"):
            op = self.take()[1]
            node = ({"&": "and", "&&": "and"}.get(op, op), node, self.comparison())
        return node

    def comparison(self):
        node = self.additive()
        while self.peek()[0] == "op" and self.peek()[1] in ("<", "<=", ">", ">=", "==", "=",
                                                              "!=", "<>"):
            op = self.take()[1]
            node = ({"=": "==", "<>": "!="}.get(op, op), node, self.additive())
        return node

    def additive(self):
        node = self.multiplicative()
        while self.peek()[0] == "op" and self.peek()[1] in ("+", "-"):
            op = self.take()[1]
            node = (op, node, self.multiplicative())
        return node

    def multiplicative(self):
        node = self.unary()
        while self.peek()[0] == "op" and self.peek()[1] in ("*", "/", "%"):
            op = self.take()[1]
            node = (op, node, self.unary())
        return node

    def unary(self):
        if self.peek()[0] == "op" and self.peek()[1] in ("-", "+"):
            op = self.take()[1]
            operand = self.unary()
            return ("neg", operand) if op == "-" else operand
        if self.peek() == ("name", "not") and self.peek(1) != ("op", "("):
            self.take()
            return ("not", self.unary())
        return self.power()

    def power(self):
        base = self.primary()
        if self.accept("^"):
            return ("^", base, self.unary())
        return base

    def primary(self):
        kind, val = self.take()
        if kind == "num":
            return ("num", float(val))
        if kind == "op" and val == "(":
            node = self.expr()
            self.expect(")")
            return node
        if kind == "name":
            if self.peek() == ("op", "("):
                self.take()
                args = []
                if not self.accept(")"):
                    args.append(self.expr())
                    while self.accept(","):
                        args.append(self.expr())
                    self.expect(")")
                return ("call", val, args)
            return ("var", val)
        raise ExpressionError(f"unexpected token {val!r}")


class Program:
    """A compiled expression text; run() evaluates it over arrays."""

    def __init__(self, text):
        self.text = text
        self.ast = _Parser(text).program()

    def run(self, inputs, outputs, shape=None):
        """inputs: name -> array (or scalar); outputs: names created at zero.
        Returns a dict with every assigned or created variable, broadcast to
        the common shape of the inputs."""
        if shape is None:
            shape = np.broadcast_shapes(*[np.shape(v) for v in inputs.values()]) if inputs else ()
        env = {k: np.broadcast_to(np.asarray(v, dtype=np.float64), shape).copy()
               for k, v in inputs.items()}
        for name in outputs:
            env.setdefault(name, np.zeros(shape))
        self._exec(self.ast, env, np.ones(shape, dtype=bool), shape)
        return env

    # ---- evaluation
    def _exec(self, node, env, mask, shape):
        kind = node[0]
        if kind == "block":
            for st in node[1]:
                self._exec(st, env, mask, shape)
        elif kind == "assign":
            val = np.broadcast_to(self._eval(node[2], env, shape), shape)
            old = env.get(node[1])
            if old is None:
                old = np.zeros(shape)
            env[node[1]] = np.where(mask, val, old)
        elif kind == "if":
            cond = np.broadcast_to(self._eval(node[1], env, shape) != 0, shape)
            self._exec(node[2], env, mask & cond, shape)
            if node[3] is not None:
                self._exec(node[3], env, mask & ~cond, shape)
        elif kind == "expr":
            self._eval(node[1], env, shape)
        else:
            raise ExpressionError(f"bad statement {kind}")

    def _eval(self, node, env, shape):
        kind = node[0]
        if kind == "num":
            return node[1]
        if kind == "var":
            name = node[1]
            if name in env:
                return env[name]
            if name in _CONSTANTS:
                return _CONSTANTS[name]
            env[name] = np.zeros(shape)  # unknown symbols resolve to zero variables
            return env[name]
        if kind == "neg":
            return -self._eval(node[1], env, shape)
        if kind == "not":
            return (np.asarray(self._eval(node[1], env, shape)) == 0).astype(float)
        if kind == "ternary":
            c = np.asarray(self._eval(node[1], env, shape)) != 0
            return np.where(c, self._eval(node[2], env, shape), self._eval(node[3], env, shape))
        if kind == "call":
            name, args = node[1], [self._eval(a, env, shape) for a in node[2]]
            if name == "rand":
                raise ExpressionError("rand() is not reproducible; seed the state instead")
            if name in _FUNCS1 and len(args) == 1:
                return _FUNCS1[name](np.asarray(args[0], dtype=float))
            if name in _FUNCS2 and len(args) == 2:
                return _FUNCS2[name](args[0], args[1])
            if name in ("min", "max") and len(args) > 2:
                f = _FUNCS2[name]
                out = args[0]
                for a in args[1:]:
                    out = f(out, a)
                return out
            if name == "clamp" and len(args) == 3:
                return np.minimum(np.maximum(args[1], args[0]), args[2])
            if name == "inrange" and len(args) == 3:
                return ((args[0] <= args[1]) & (args[1] <= args[2])).astype(float)
            raise ExpressionError(f"unknown function {name}/{len(args)}")
        a = self._eval(node[1], env, shape)
        b = self._eval(node[2], env, shape)
        if kind == "+":
            return a + b
        if kind == "-":
            return a - b
        if kind == "*":
            return a * b
        if kind == "/":
            return np.asarray(a, dtype=float) / b
        if kind == "%":
            return np.fmod(a, b)
        if kind == "^":
            return np.power(np.asarray(a, dtype=float), b)
        truth = {"<": np.less, "<=": np.less_equal, ">": np.greater, ">=": np.greater_equal,
                 "==": np.equal, "!=": np.not_equal}
        if kind in truth:
            return truth[kind](a, b).astype(float)
        a0 = np.asarray(a) != 0
        b0 = np.asarray(b) != 0
        logic = {"and": a0 & b0, "or": a0 | b0, "xor": a0 ^ b0, "nand": ~(a0 & b0),
                 "nor": ~(a0 | b0)}
        if kind in logic:
            return logic[kind].astype(float)
        raise ExpressionError(f"bad operator {kind}")


def compile_expression(text):
    return Program(text)


def evaluate(text, inputs, outputs):
    return Program(text).run(inputs, outputs)
