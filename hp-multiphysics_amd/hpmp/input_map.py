"""input_map: ``key: value`` configuration with block-prefixed overrides.

Mirrors the reference's input_map (input_map/input_map.h:18-113,
input_map/input_map.cpp): a file of ``key: value`` lines, ``#`` comments, and
lookups that try the block-prefixed key (``b0_mu``) before the global one (``mu``),
as tri_hp/ins/allocate.cpp:20-21 and tri_hp/allocate.cpp:275-276 do.  Values that
are lists (``alpha: .25 .1667``) are read as whitespace-separated entries.

A number that does not parse as a plain number is evaluated as a formula, as the
reference does with muParser (input_map.cpp:80-137 for one value, :164-227 for a
list, one formula per whitespace-separated entry): ``+ - * / ^``, unary signs,
parentheses, comparisons, ``&&``/``||``, ``a ? b : c``, muParser's built-in
functions (plus ``erf``/``erfc``, input_map.cpp:95-104) and the constants ``_pi``,
``_e``; every other name is looked up in the map itself and evaluated the same way,
so ``cfls: C*2.5 C*1.5 C*1.0`` with ``C: 1.0`` gives 2.5, 1.5, 1.0.  Formulas are
parsed into a whitelisted syntax tree and evaluated node by node; nothing in a file
is executed as Python.
"""
from __future__ import annotations

import ast
import math
import re


class FormulaError(ValueError):
    pass


def _sign(x):
    return (x > 0) - (x < 0)


_FUNCS = {
    "sin": math.sin, "cos": math.cos, "tan": math.tan, "asin": math.asin, "acos": math.acos,
    "atan": math.atan, "sinh": math.sinh, "cosh": math.cosh, "tanh": math.tanh, "asinh": math.asinh,
    "acosh": math.acosh, "atanh": math.atanh, "log2": math.log2, "log10": math.log10, "log": math.log,
    "ln": math.log, "exp": math.exp, "sqrt": math.sqrt, "sign": _sign, "rint": lambda x: float(round(x)),
    "abs": abs, "erf": math.erf, "erfc": math.erfc,
    "min": lambda *a: min(a), "max": lambda *a: max(a), "sum": lambda *a: sum(a),
    "avg": lambda *a: sum(a) / len(a),
}
_CONSTS = {"_pi": math.pi, "_e": math.e}
_NUM = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?$")


def _to_python(expr: str) -> str:
    """muParser operators to Python syntax: ^ -> **, && / || -> and / or, c ? a : b ->
    (a if c else b) (right-associative, as muParser)."""
    s = expr.replace("^", "**").replace("&&", " and ").replace("||", " or ")
    if "?" not in s:
        return s
    depth, q = 0, -1
    for k, ch in enumerate(s):
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        elif ch == "?" and depth == 0:
            q = k
            break
    if q < 0:
        # a ternary inside parentheses: convert each parenthesised group
        out, k = [], 0
        while k < len(s):
            if s[k] == "(":
                d, j = 1, k + 1
                while j < len(s) and d:
                    d += {"(": 1, ")": -1}.get(s[j], 0)
                    j += 1
                out.append("(" + _to_python(s[k + 1:j - 1]) + ")")
                k = j
            else:
                out.append(s[k])
                k += 1
        return "".join(out)
    depth, c = 0, -1
    for k in range(q + 1, len(s)):
        ch = s[k]
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        elif ch == "?" and depth == 0:
            depth += 100  # a nested ternary in the true branch owns the next ':'
        elif ch == ":":
            if depth == 0:
                c = k
                break
            if depth >= 100:
                depth -= 100
    if c < 0:
        raise FormulaError(f"'?' without ':' in {expr!r}")
    return f"(({_to_python(s[q + 1:c])}) if ({_to_python(s[:q])}) else ({_to_python(s[c + 1:])}))"


def evaluate(expr: str, lookup=lambda name: None) -> float:
    """Value of a muParser-style formula; `lookup(name)` returns the value of a variable
    (or None when it is undefined)."""
    try:
        tree = ast.parse(_to_python(expr.strip()), mode="eval")
    except SyntaxError as e:
        raise FormulaError(f"cannot parse {expr!r}: {e.msg}") from None

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)) and not isinstance(n.value, bool):
            return float(n.value)
        if isinstance(n, ast.BinOp):
            a, b = ev(n.left), ev(n.right)
            ops = {ast.Add: lambda: a + b, ast.Sub: lambda: a - b, ast.Mult: lambda: a * b,
                   ast.Div: lambda: a / b, ast.Pow: lambda: a ** b}
            if type(n.op) not in ops:
                raise FormulaError(f"operator not allowed in {expr!r}")
            return float(ops[type(n.op)]())
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            v = ev(n.operand)
            return -v if isinstance(n.op, ast.USub) else v
        if isinstance(n, ast.Compare):
            left = ev(n.left)
            for op, rc in zip(n.ops, n.comparators):
                right = ev(rc)
                ok = {ast.Lt: left < right, ast.LtE: left <= right, ast.Gt: left > right,
                      ast.GtE: left >= right, ast.Eq: left == right, ast.NotEq: left != right}.get(type(op))
                if ok is None:
                    raise FormulaError(f"comparison not allowed in {expr!r}")
                if not ok:
                    return 0.0
                left = right
            return 1.0
        if isinstance(n, ast.BoolOp):
            vals = [ev(v) for v in n.values]
            return float(all(vals)) if isinstance(n.op, ast.And) else float(any(vals))
        if isinstance(n, ast.IfExp):
            return ev(n.body) if ev(n.test) != 0.0 else ev(n.orelse)
        if isinstance(n, ast.Name):
            if n.id in _CONSTS:
                return _CONSTS[n.id]
            v = lookup(n.id)
            if v is None:
                raise FormulaError(f"undefined variable {n.id!r} in {expr!r}")
            return float(v)
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _FUNCS and not n.keywords:
            return float(_FUNCS[n.func.id](*[ev(a) for a in n.args]))
        raise FormulaError(f"element {type(n).__name__} not allowed in {expr!r}")

    return ev(tree)


class InputMap(dict):
    @classmethod
    def from_text(cls, text: str) -> "InputMap":
        m = cls()
        for raw in text.splitlines():
            line = raw.split("#", 1)[0].strip()
            if not line or ":" not in line:
                continue
            k, v = line.split(":", 1)
            m[" ".join(k.split())] = " ".join(v.split())
        return m

    @classmethod
    def from_file(cls, path: str) -> "InputMap":
        with open(path) as f:
            return cls.from_text(f.read())

    def lookup(self, key: str, prefix: str = "b0"):
        if prefix and f"{prefix}_{key}" in self:
            return self[f"{prefix}_{key}"]
        return self.get(key)

    def value(self, key: str, _seen: frozenset = frozenset()) -> float | None:
        """The double value of `key` (input_map::get(keyword, double&)): a plain number,
        else its formula evaluated with the map's own entries as variables; None when the
        key is absent.  A formula that refers back to itself raises FormulaError."""
        if key not in self:
            return None
        return self._eval(self[key], _seen | {key})

    def _eval(self, text: str, seen: frozenset) -> float:
        if _NUM.match(text):
            return float(text)

        def var(name):
            if name in seen:
                raise FormulaError(f"formula for {name!r} refers to itself")
            return self.value(name, seen)

        return evaluate(text, var)

    def getf(self, key: str, default: float, prefix: str = "b0") -> float:
        k = f"{prefix}_{key}" if prefix and f"{prefix}_{key}" in self else key
        v = self.value(k)
        return float(v) if v is not None else float(default)

    def geti(self, key: str, default: int, prefix: str = "b0") -> int:
        # an int is read directly, else as a double and truncated (input_map.cpp:139-162)
        return int(self.getf(key, default, prefix))

    def getlist(self, key: str, default, n: int | None = None, prefix: str = "b0") -> list:
        v = self.lookup(key, prefix)
        vals = [self._eval(x, frozenset()) for x in v.split()] if v is not None else list(default)
        return vals[:n] if n is not None else vals


# ins edge-boundary types whose residual is the generic flux of bdry_ins::generic
# (tri_hp/ins/bdry_ins.h:31-77): rho (u.n) u + p_bc n and rho u.n, with the essential
# variables of each type (tri_hp/ins/getnewbdry.cpp:95-113): inflow holds u and v
# (bdry_ins.h:79-92), outflow is bdry_ins::generic itself, symmetry holds velocity
# component `dir` (bdry_ins.h:282-297; its tadvance sets that component to zero on the
# boundary, bdry.cpp:215-242)
INS_FLUX_TYPES = ("inflow", "outflow", "symmetry")


def ins_boundary_essential(inmap: InputMap, ids, prefix: str = "b0") -> dict:
    """Essential-variable masks {boundary id: mask} of the edge boundaries `ids` from the
    deck's ``<prefix>_s<id>_hp_type`` entries (tri_hp/allocate.cpp:182-186).  A type that
    needs more than the generic flux (characteristic, euler, plain, applied_stress, ...)
    and a periodic mesh boundary (``<prefix>_s<id>_type: prdc``) raise ValueError naming
    the boundary, rather than run with a different boundary condition."""
    out = {}
    for i in ids:
        key = f"{prefix}_s{i}"
        if inmap.get(f"{key}_type", "").split()[:1] == ["prdc"]:
            raise ValueError(f"{key}: periodic boundaries are not supported")
        typ = inmap.get(f"{key}_hp_type", "plain").split()[0]
        if typ not in INS_FLUX_TYPES:
            raise ValueError(f"{key}_hp_type: '{typ}' is not supported (supported: {', '.join(INS_FLUX_TYPES)})")
        if typ == "inflow":
            out[i] = 0b011
        elif typ == "outflow":
            out[i] = 0
        else:
            d = inmap.geti(f"{key}_dir", 0, prefix="")
            if d not in (0, 1):
                raise ValueError(f"{key}_dir must be 0 or 1")
            out[i] = 1 << d
    return out
