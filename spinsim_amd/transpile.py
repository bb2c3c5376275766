"""Translate a numba-style Python field function into C for hipRTC.

The reference JIT-compiles the user's ``get_field(time_sample,
sweep_parameters, field_sample)`` with numba as a device function
(spinsim/__init__.py:620; usage in example.py:7-10, :33-38).  Here the same
Python function is translated, statement by statement, into a C function

    __device__ void get_field(double t, const double* sweep, double* field[, SS_CONTEXT])

that hipRTC compiles into the integrator, so no Python runs per sample.

Supported: assignments to locals and to ``field_sample[k]``, ``+= -= *= /=``,
``if/elif/else``, ``for i in range(...)``, ``return``, ``pass``; arithmetic,
comparisons, ``and/or/not``, conditional expressions; ``math``/``cmath``-free
real functions of ``math`` and ``numpy`` (sin, cos, exp, sqrt, atan2, ...),
``abs/min/max/float/int``; numeric globals and closure constants (inlined as
literals, frozen at compile time like numba freezes them); interpolation
samplers from ``generate_interpolation_sampler`` (read at run time from
device tables; a multi-axis sampler's components as ``v = sampler(t)`` then
``v[a]``, or ``sampler(t)[a]``); and other module-level Python functions that return a value
and obey the same rules (emitted as ``__device__`` helpers).  Anything else
raises ``TranspileError`` naming the line.
"""
from __future__ import annotations

import ast
import builtins
import inspect
import math
import textwrap
from typing import Callable

import numpy as np


class TranspileError(ValueError):
    """The Python field function uses a construct with no C translation."""


_MATH_FUNCS = {
    "sin": "sin", "cos": "cos", "tan": "tan", "asin": "asin", "acos": "acos", "atan": "atan",
    "atan2": "atan2", "sinh": "sinh", "cosh": "cosh", "tanh": "tanh", "asinh": "asinh",
    "acosh": "acosh", "atanh": "atanh", "exp": "exp", "expm1": "expm1", "log": "log",
    "log2": "log2", "log10": "log10", "log1p": "log1p", "sqrt": "sqrt", "fabs": "fabs",
    "floor": "floor", "ceil": "ceil", "trunc": "trunc", "erf": "erf", "erfc": "erfc",
    "hypot": "hypot", "pow": "pow", "fmod": "fmod", "copysign": "copysign", "gamma": "tgamma",
    "lgamma": "lgamma", "cbrt": "cbrt", "exp2": "exp2",
    # numpy spellings
    "arcsin": "asin", "arccos": "acos", "arctan": "atan", "arctan2": "atan2",
    "arcsinh": "asinh", "arccosh": "acosh", "arctanh": "atanh", "absolute": "fabs",
    "abs": "fabs", "power": "pow", "square": None, "sign": None, "radians": None,
    "degrees": None,
}
_MATH_CONSTS = {"pi": math.pi, "tau": math.tau, "e": math.e, "inf": math.inf}


def _lit(v) -> str:
    if isinstance(v, bool):
        return "1.0" if v else "0.0"
    v = float(v)
    if math.isinf(v):
        return "(1.0/0.0)" if v > 0 else "(-1.0/0.0)"
    if math.isnan(v):
        return "(0.0/0.0)"
    r = repr(v)
    if "e" not in r and "." not in r and "inf" not in r:
        r += ".0"
    return f"({r})" if v < 0 else r


class _Fn:
    """One translated function (the field function or a helper)."""

    def __init__(self, func: Callable, ctx: "_Context", role: str):
        self.func = func
        self.ctx = ctx
        self.role = role  # "field" or "helper"
        try:
            src = textwrap.dedent(inspect.getsource(func))
        except (OSError, TypeError) as exc:
            raise TranspileError(f"cannot read the source of {func!r}: {exc}") from exc
        tree = ast.parse(src)
        fdef = next((n for n in tree.body if isinstance(n, (ast.FunctionDef, ast.Lambda))), None)
        if fdef is None:
            raise TranspileError(f"{func!r} is not a plain function definition")
        self.node = fdef
        self.first_line = func.__code__.co_firstlineno
        self.params = [a.arg for a in fdef.args.args]
        self.globals = dict(func.__globals__)
        if func.__closure__:
            for name, cell in zip(func.__code__.co_freevars, func.__closure__):
                try:
                    self.globals[name] = cell.cell_contents
                except ValueError:
                    pass
        self.locals: dict[str, str] = {}  # name -> C type
        self.int_vars: set[str] = set()
        self.arrays: dict[str, int] = {}  # name -> components (multi-axis sampler results)

    def err(self, node, msg):
        line = self.first_line + getattr(node, "lineno", 1) - 1
        return TranspileError(f"{self.func.__name__} (line {line}): {msg}")

    # ---- declarations -----------------------------------------------------
    def _sampler_of(self, node):
        """The sampler a Call node calls, if any."""
        from .interpolation import InterpolationSampler

        if not isinstance(node, ast.Call):
            return None
        f = node.func
        obj = None
        if isinstance(f, ast.Name):
            obj = self.globals.get(f.id)
        elif isinstance(f, ast.Attribute) and self._module_of(f.value) is None:
            obj = self._resolve(f)
        return obj if isinstance(obj, InterpolationSampler) else None

    def _collect(self):
        for node in ast.walk(self.node):
            if isinstance(node, ast.Assign) and len(node.targets) == 1 \
                    and isinstance(node.targets[0], ast.Name):
                sampler = self._sampler_of(node.value)
                if sampler is not None and sampler.axes > 1:
                    name = node.targets[0].id
                    if self.arrays.get(name, sampler.axes) != sampler.axes:
                        raise self.err(node, f"{name!r} holds samples of different axis counts")
                    self.arrays[name] = sampler.axes
        for node in ast.walk(self.node):
            if isinstance(node, ast.For):
                if isinstance(node.target, ast.Name):
                    self.int_vars.add(node.target.id)
            targets = []
            if isinstance(node, ast.Assign):
                targets = node.targets
            elif isinstance(node, (ast.AugAssign, ast.AnnAssign)):
                targets = [node.target]
            for t in targets:
                for n in ast.walk(t):
                    if isinstance(n, ast.Name) and isinstance(n.ctx, ast.Store):
                        if n.id in self.params:
                            raise self.err(n, f"assignment to parameter {n.id!r}")
                        self.locals.setdefault(n.id, "double")
        for v in self.int_vars:
            self.locals[v] = "long long"
        for n in self.arrays:
            self.locals.pop(n, None)

    def emit(self) -> str:
        self._collect()
        body = self._block(self.node.body, 1)
        decls = "".join(f"    {t} {n} = 0;\n" for n, t in sorted(self.locals.items()))
        decls += "".join(f"    double {n}[{a}] = {{0}};\n" for n, a in sorted(self.arrays.items()))
        if self.role == "field":
            if len(self.params) != 3:
                raise self.err(self.node, "get_field must take (time_sample, sweep_parameters, "
                                          "field_sample)")
            p = self.params
            ctx = ", SS_CONTEXT" if self.ctx.uses_context else ""
            head = (f"__device__ void get_field(double {p[0]}, const double* {p[1]}, "
                    f"double* {p[2]}{ctx})")
            return head + "\n{\n" + decls + body + "}\n"
        args = ", ".join(f"double {a}" for a in self.params)
        ctx = (", SS_CONTEXT" if args else "SS_CONTEXT") if self.ctx.uses_context else ""
        return (f"__device__ static double ss_helper_{self.func.__name__}({args}{ctx})\n{{\n"
                + decls + body + "    return 0.0;\n}\n")

    # ---- statements ---------------------------------------------------------
    def _block(self, stmts, depth) -> str:
        out = []
        for st in stmts:
            out.append(self._stmt(st, depth))
        return "".join(out)

    def _stmt(self, st, depth) -> str:
        pad = "    " * depth
        if isinstance(st, ast.Expr):
            if isinstance(st.value, ast.Constant) and isinstance(st.value.value, str):
                return ""  # docstring
            return pad + self._expr(st.value) + ";\n"
        if isinstance(st, ast.Pass):
            return ""
        if isinstance(st, ast.Return):
            if st.value is None or (isinstance(st.value, ast.Constant) and st.value.value is None):
                return pad + ("return;\n" if self.role == "field" else "return 0.0;\n")
            if self.role == "field":
                raise self.err(st, "get_field returns nothing; write into field_sample")
            return pad + f"return {self._expr(st.value)};\n"
        if isinstance(st, ast.Assign):
            if len(st.targets) == 1 and isinstance(st.targets[0], ast.Name) \
                    and st.targets[0].id in self.arrays:
                name = st.targets[0].id
                sampler = self._sampler_of(st.value)
                if sampler is None or sampler.axes != self.arrays[name] or len(st.value.args) != 1:
                    raise self.err(st, f"{name!r} holds multi-axis samples; assign it only "
                                       "from one call of that sampler")
                slot = self.ctx.sampler_slot(sampler)
                t = self._expr(st.value.args[0])
                return pad + f"SS_SAMPLE_AXES({slot}, {t}, {name}, {sampler.axes});\n"
            value = self._expr(st.value)
            if len(st.targets) == 1 and isinstance(st.targets[0], ast.Tuple):
                raise self.err(st, "tuple assignment is not supported")
            return "".join(pad + f"{self._target(t)} = {value};\n" for t in st.targets)
        if isinstance(st, ast.AnnAssign):
            if st.value is None:
                return ""
            return pad + f"{self._target(st.target)} = {self._expr(st.value)};\n"
        if isinstance(st, ast.AugAssign):
            tgt = self._target(st.target)
            op = type(st.op)
            val = self._expr(st.value)
            if op in (ast.Add, ast.Sub, ast.Mult, ast.Div):
                sym = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/"}[op]
                return pad + f"{tgt} = {tgt} {sym} ({val});\n"
            if op is ast.Pow:
                return pad + f"{tgt} = pow({tgt}, {val});\n"
            raise self.err(st, f"augmented operator {op.__name__} is not supported")
        if isinstance(st, ast.If):
            s = pad + f"if ({self._cond(st.test)}) {{\n" + self._block(st.body, depth + 1) + pad + "}"
            if st.orelse:
                s += " else {\n" + self._block(st.orelse, depth + 1) + pad + "}"
            return s + "\n"
        if isinstance(st, ast.For):
            if not (isinstance(st.iter, ast.Call) and isinstance(st.iter.func, ast.Name)
                    and st.iter.func.id == "range" and not st.orelse):
                raise self.err(st, "only `for i in range(...)` loops are supported")
            if not isinstance(st.target, ast.Name):
                raise self.err(st, "loop target must be a name")
            a = [self._int_expr(x) for x in st.iter.args]
            if len(a) == 1:
                lo, hi, step = "0", a[0], "1"
            elif len(a) == 2:
                lo, hi, step = a[0], a[1], "1"
            elif len(a) == 3:
                lo, hi, step = a
            else:
                raise self.err(st, "range takes 1 to 3 arguments")
            v = st.target.id
            cond = f"({step}) > 0 ? {v} < ({hi}) : {v} > ({hi})"
            return (pad + f"for ({v} = {lo}; {cond}; {v} += {step}) {{\n"
                    + self._block(st.body, depth + 1) + pad + "}\n")
        if isinstance(st, ast.While):
            raise self.err(st, "while loops are not supported (use for ... in range)")
        raise self.err(st, f"statement {type(st).__name__} is not supported")

    def _target(self, t) -> str:
        if isinstance(t, ast.Name):
            if t.id in self.arrays:
                raise self.err(t, f"{t.id!r} holds multi-axis samples and cannot be updated")
            return t.id
        if isinstance(t, ast.Subscript) and isinstance(t.value, ast.Name):
            if self.role == "field" and t.value.id == self.params[2]:
                return f"{self.params[2]}[{self._int_expr(self._index(t))}]"
        raise self.err(t, "can only assign to local names and field_sample[k]")

    def _index(self, sub):
        s = sub.slice
        if isinstance(s, ast.Index):  # python < 3.9
            s = s.value
        if isinstance(s, (ast.Slice, ast.Tuple)):
            raise self.err(sub, "slices and multi-dimensional indices are not supported")
        return s

    # ---- expressions ------------------------------------------------------
    def _cond(self, node) -> str:
        return self._expr(node)

    def _int_expr(self, node) -> str:
        if isinstance(node, ast.Constant) and isinstance(node.value, int) and not isinstance(
                node.value, bool):
            return str(node.value)
        if isinstance(node, ast.Name) and node.id in self.int_vars:
            return node.id
        return f"(long long)({self._expr(node)})"

    def _expr(self, node) -> str:
        if isinstance(node, ast.Constant):
            if isinstance(node.value, (int, float, bool)):
                return _lit(node.value)
            raise self.err(node, f"constant {node.value!r} is not numeric")
        if isinstance(node, ast.Name):
            return self._name(node)
        if isinstance(node, ast.Attribute):
            return self._attribute(node)
        if isinstance(node, ast.BinOp):
            a, b = self._expr(node.left), self._expr(node.right)
            op = type(node.op)
            if op is ast.Add:
                return f"({a} + {b})"
            if op is ast.Sub:
                return f"({a} - {b})"
            if op is ast.Mult:
                return f"({a} * {b})"
            if op is ast.Div:
                return f"({a} / {b})"
            if op is ast.Pow:
                if isinstance(node.right, ast.Constant) and node.right.value == 2:
                    return f"(({a}) * ({a}))"
                return f"pow({a}, {b})"
            if op is ast.Mod:  # Python's floor modulo
                return f"({a} - floor(({a}) / ({b})) * ({b}))"
            if op is ast.FloorDiv:
                return f"floor(({a}) / ({b}))"
            raise self.err(node, f"operator {op.__name__} is not supported")
        if isinstance(node, ast.UnaryOp):
            v = self._expr(node.operand)
            if isinstance(node.op, ast.USub):
                return f"(-{v})"
            if isinstance(node.op, ast.UAdd):
                return v
            if isinstance(node.op, ast.Not):
                return f"(!({v}))"
            raise self.err(node, "unsupported unary operator")
        if isinstance(node, ast.BoolOp):
            sym = " && " if isinstance(node.op, ast.And) else " || "
            return "(" + sym.join(f"({self._expr(v)})" for v in node.values) + ")"
        if isinstance(node, ast.Compare):
            parts = []
            left = node.left
            for op, right in zip(node.ops, node.comparators):
                sym = {ast.Lt: "<", ast.LtE: "<=", ast.Gt: ">", ast.GtE: ">=", ast.Eq: "==",
                       ast.NotEq: "!="}.get(type(op))
                if sym is None:
                    raise self.err(node, f"comparison {type(op).__name__} is not supported")
                parts.append(f"({self._expr(left)} {sym} {self._expr(right)})")
                left = right
            return "(" + " && ".join(parts) + ")"
        if isinstance(node, ast.IfExp):
            return (f"(({self._expr(node.test)}) ? ({self._expr(node.body)}) : "
                    f"({self._expr(node.orelse)}))")
        if isinstance(node, ast.Subscript):
            if isinstance(node.value, ast.Name) and self.role == "field" and \
                    node.value.id in self.params[1:]:
                return f"{node.value.id}[{self._int_expr(self._index(node))}]"
            if isinstance(node.value, ast.Name) and node.value.id in self.arrays:
                return self._axis_read(node, node.value.id, self.arrays[node.value.id])
            sampler = self._sampler_of(node.value)
            if sampler is not None and sampler.axes > 1:
                if len(node.value.args) != 1:
                    raise self.err(node, "a sampler takes one argument (the time)")
                slot = self.ctx.sampler_slot(sampler)
                axis = self._index(node)
                if isinstance(axis, ast.Constant):
                    self._axis_constant(node, axis, sampler.axes)
                return (f"SS_SAMPLE_AXIS({slot}, (int)({self._int_expr(axis)}), "
                        f"{self._expr(node.value.args[0])})")
            raise self.err(node, "only sweep_parameters[k], field_sample[k] and multi-axis "
                                 "samples can be indexed")
        if isinstance(node, ast.Call):
            return self._call(node)
        raise self.err(node, f"expression {type(node).__name__} is not supported")

    def _axis_constant(self, node, axis, axes):
        if not (isinstance(axis.value, int) and not isinstance(axis.value, bool)
                and 0 <= axis.value < axes):
            raise self.err(node, f"axis index {axis.value!r} out of range for {axes} axes")

    def _axis_read(self, node, name, axes) -> str:
        axis = self._index(node)
        if isinstance(axis, ast.Constant):
            self._axis_constant(node, axis, axes)
            return f"{name}[{axis.value}]"
        i = self._int_expr(axis)
        return f"(({i}) >= 0 && ({i}) < {axes} ? {name}[{i}] : 0.0)"

    def _name(self, node) -> str:
        n = node.id
        if n in self.arrays:
            raise self.err(node, f"{n!r} holds multi-axis samples; read them as {n}[a]")
        if n in self.params:
            if self.role == "field" and n != self.params[0]:
                raise self.err(node, f"{n!r} can only be used with an index")
            return n
        if n in self.locals:
            return n
        if n in ("True", "False"):
            return _lit(n == "True")
        if n in self.globals:
            v = self.globals[n]
            if isinstance(v, (int, float, np.floating, np.integer)) and not isinstance(v, bool):
                return _lit(v)
            if isinstance(v, bool):
                return _lit(v)
            raise self.err(node, f"global {n!r} of type {type(v).__name__} cannot be inlined")
        raise self.err(node, f"unknown name {n!r}")

    def _module_of(self, node):
        """The Python module object an Attribute's base refers to, if any."""
        if isinstance(node, ast.Name):
            v = self.globals.get(node.id)
            if inspect.ismodule(v):
                return v
        return None

    def _attribute(self, node) -> str:
        mod = self._module_of(node.value)
        if mod is not None and mod.__name__ in ("math", "numpy", "cmath"):
            if node.attr in _MATH_CONSTS:
                return _lit(_MATH_CONSTS[node.attr])
        if mod is not None:
            v = getattr(mod, node.attr, None)
            if isinstance(v, (int, float)) and not isinstance(v, bool):
                return _lit(v)
        raise self.err(node, f"attribute {ast.unparse(node)!r} is not supported")

    def _call(self, node) -> str:
        if node.keywords:
            raise self.err(node, "keyword arguments are not supported")
        args = [self._expr(a) for a in node.args]
        f = node.func
        # math.f / np.f
        if isinstance(f, ast.Attribute):
            mod = self._module_of(f.value)
            if mod is not None and mod.__name__ in ("math", "numpy"):
                return self._math(node, f.attr, args)
            target = self._resolve(f)
            if target is not None:
                return self._call_object(node, target, args)
            raise self.err(node, f"call to {ast.unparse(f)!r} is not supported")
        if isinstance(f, ast.Name):
            if f.id in ("abs",):
                return f"fabs({args[0]})"
            if f.id == "min":
                return self._fold("fmin", args)
            if f.id == "max":
                return self._fold("fmax", args)
            if f.id == "float":
                return f"((double)({args[0]}))"
            if f.id == "int":
                return f"((double)(long long)({args[0]}))"
            if f.id == "round":
                if len(args) != 1:
                    raise self.err(node, "round() with digits is not supported")
                return f"rint({args[0]})"  # round-half-even like Python
            if f.id == "pow":
                return f"pow({args[0]}, {args[1]})"
            if f.id in self.globals:
                return self._call_object(node, self.globals[f.id], args)
            if hasattr(builtins, f.id):
                raise self.err(node, f"builtin {f.id}() is not supported")
            raise self.err(node, f"unknown function {f.id!r}")
        raise self.err(node, "unsupported call")

    def _resolve(self, attr):
        base = attr.value
        if isinstance(base, ast.Name) and base.id in self.globals:
            obj = self.globals[base.id]
            return getattr(obj, attr.attr, None)
        return None

    @staticmethod
    def _fold(fn, args):
        out = args[0]
        for a in args[1:]:
            out = f"{fn}({out}, {a})"
        return out

    def _math(self, node, name, args) -> str:
        if name not in _MATH_FUNCS:
            raise self.err(node, f"function {name!r} is not supported")
        c = _MATH_FUNCS[name]
        if name == "square":
            return f"(({args[0]}) * ({args[0]}))"
        if name == "sign":
            return f"(({args[0]}) > 0 ? 1.0 : (({args[0]}) < 0 ? -1.0 : 0.0))"
        if name == "radians":
            return f"(({args[0]}) * {_lit(math.pi / 180)})"
        if name == "degrees":
            return f"(({args[0]}) * {_lit(180 / math.pi)})"
        return f"{c}({', '.join(args)})"

    def _call_object(self, node, obj, args) -> str:
        from .interpolation import InterpolationSampler

        if isinstance(obj, InterpolationSampler):
            if len(args) != 1:
                raise self.err(node, "a sampler takes one argument (the time)")
            if obj.axes > 1:
                raise self.err(node, f"a multi-axis sampler gives {obj.axes} values: assign them "
                                     "to a name (v = sampler(t), then v[a]) or index the call")
            slot = self.ctx.sampler_slot(obj)
            return f"SS_SAMPLE({slot}, {args[0]})"
        if inspect.isfunction(obj):
            name = self.ctx.helper(obj)
            call_args = list(args)
            if self.ctx.uses_context:
                call_args.append("ss_ctx_ref")
            return f"ss_helper_{name}({', '.join(call_args)})"
        if isinstance(obj, np.ufunc) or inspect.isbuiltin(obj):
            return self._math(node, obj.__name__, args)
        raise self.err(node, f"cannot call {obj!r} from a field function")


class _Context:
    def __init__(self):
        self.samplers: list = []
        self.helpers: dict = {}
        self.helper_src: list[str] = []
        self.uses_context = False
        self._pending: list = []

    def sampler_slot(self, sampler) -> int:
        for k, s in enumerate(self.samplers):
            if s is sampler:
                return k
        self.samplers.append(sampler)
        self.uses_context = True
        return len(self.samplers) - 1

    def helper(self, func) -> str:
        if func not in self.helpers:
            self.helpers[func] = func.__name__
            self._pending.append(func)
        return func.__name__


def _run(get_field: Callable, uses_context: bool):
    ctx = _Context()
    ctx.uses_context = uses_context
    field = _Fn(get_field, ctx, "field").emit()
    helpers = []
    done = set()
    while ctx._pending:
        fn = ctx._pending.pop(0)
        if fn in done:
            continue
        done.add(fn)
        helpers.append(_Fn(fn, ctx, "helper").emit())
    protos = "".join(h.split("\n", 1)[0] + ";\n" for h in helpers)
    return protos + "".join(helpers) + field, ctx


def transpile(get_field: Callable):
    """Translate ``get_field`` to C.  Returns (source, samplers, uses_context).

    A first pass discovers whether any sampler is read; if so every emitted
    function takes the sampler context, and a second pass emits that text.
    """
    source, ctx = _run(get_field, False)
    if ctx.uses_context:
        source, ctx = _run(get_field, True)
    return source, list(ctx.samplers), ctx.uses_context
