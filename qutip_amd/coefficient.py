"""Time-dependent coefficients of QuTiP's list form
``QobjEvo([H0, [H1, c1], [H2, c2]])`` (qutip/core/coefficient.py:63-229,
qutip/core/cy/coefficient.pyx).

A coefficient is called on the host as the reference calls it
(``Coefficient.__call__(t)``, coefficient.pyx:118-137) and, when it can be,
also compiled to a small program that the GPU kernels evaluate themselves at
every Runge-Kutta stage time (``device_program``): a stack machine over
complex128 values whose instruction set is declared in include/qutip_amd.h
(``qa_rk_set_coefficients``).  That keeps the whole step loop of a
time-dependent solve on the device.

* ``StrCoefficient``: a Python expression in ``t`` and the ``args`` with
  the names of the reference's string environment (coefficient.py:367-395:
  sin, cos, exp, sqrt, abs, real, conj, norm, pi, np.*, ...).  The host
  call evaluates it as StrFunctionCoefficient does (coefficient.pyx:
  292-410: ``eval`` in that environment); the device program is compiled
  from its syntax tree with the arguments folded in as constants.
* ``InterCoefficient``: an array sampled at ``tlist``, interpolated with
  order 0, 1 or a spline of order k (coefficient.pyx:412-560: the same
  piecewise polynomials, built with scipy's make_interp_spline).
* ``FunctionCoefficient``: a Python function (coefficient.pyx:176-290);
  host only, so a generator holding one runs the plain integrator form.
* ``ConstantCoefficient``, sums, products and conjugates of the above.
"""
import ast
import inspect
import math

import numpy as np
import scipy.special

__all__ = ["Coefficient", "StrCoefficient", "InterCoefficient", "FunctionCoefficient",
           "ConstantCoefficient", "coefficient", "DeviceProgram"]

# ---- instruction set (include/qutip_amd.h, qa_coef.h) -----------------------
CO_CONST, CO_T, CO_ADD, CO_SUB, CO_MUL, CO_DIV, CO_POW, CO_NEG, CO_FN, CO_INTERP = range(10)
CO_LT, CO_LE, CO_GT, CO_GE, CO_EQ, CO_NE, CO_SELECT = range(10, 17)

FUNCTIONS = ["sin", "cos", "tan", "exp", "log", "log10", "sqrt", "abs", "real", "imag", "conj",
             "norm", "arg", "sinh", "cosh", "tanh", "asin", "acos", "atan", "asinh", "acosh",
             "atanh", "erf"]
FN_CODE = {name: i for i, name in enumerate(FUNCTIONS)}
# the numpy / scipy spellings the string environment also offers
ALIASES = {"arcsin": "asin", "arccos": "acos", "arctan": "atan", "arcsinh": "asinh",
           "arccosh": "acosh", "arctanh": "atanh", "absolute": "abs", "angle": "arg",
           "zerf": "erf", "conjugate": "conj"}

_BINOPS = {ast.Add: CO_ADD, ast.Sub: CO_SUB, ast.Mult: CO_MUL, ast.Div: CO_DIV, ast.Pow: CO_POW}
_CMPOPS = {ast.Lt: CO_LT, ast.LtE: CO_LE, ast.Gt: CO_GT, ast.GtE: CO_GE, ast.Eq: CO_EQ,
           ast.NotEq: CO_NE}


def _proj(x):
    if np.isfinite(x):
        return x
    return np.inf + 0j * np.imag(x)


# The names a string coefficient may use (coefficient.py:367-395).
STR_ENV = {
    "sin": np.sin, "cos": np.cos, "tan": np.tan, "asin": np.arcsin, "acos": np.arccos,
    "atan": np.arctan, "pi": np.pi, "sinh": np.sinh, "cosh": np.cosh, "tanh": np.tanh,
    "asinh": np.arcsinh, "acosh": np.arccosh, "atanh": np.arctanh, "exp": np.exp,
    "log": np.log, "log10": np.log10, "erf": scipy.special.erf, "zerf": scipy.special.erf,
    "sqrt": np.sqrt, "real": np.real, "imag": np.imag, "conj": np.conj, "abs": np.abs,
    "norm": lambda x: np.abs(x) ** 2, "arg": np.angle, "proj": _proj, "np": np,
    "spe": scipy.special,
}


class DeviceProgram:
    """Code (int32 instructions) and constants (complex128) of one or more
    coefficient programs; ``start`` delimits the programs."""

    def __init__(self, code, consts, start=None):
        self.code = [int(c) for c in code]
        self.consts = [complex(c) for c in consts]
        self.start = list(start) if start is not None else [0, len(self.code)]

    @staticmethod
    def _relocate(code, off):
        out = []
        for ins in code:
            op, arg = ins & 0xFF, ins >> 8
            if op in (CO_CONST, CO_INTERP):
                arg += off
            out.append(op | (arg << 8))
        return out

    @staticmethod
    def combine(progs, op):
        """One program: the programs' values combined by ``op`` (a binary
        instruction applied left to right)."""
        code, consts = [], []
        for i, p in enumerate(progs):
            code += DeviceProgram._relocate(p.code, len(consts))
            consts += p.consts
            if i:
                code.append(op)
        return DeviceProgram(code, consts)

    @staticmethod
    def concat(progs):
        """A set of programs, one per element (for qa_rk_set_coefficients)."""
        code, consts, start = [], [], [0]
        for p in progs:
            code += DeviceProgram._relocate(p.code, len(consts))
            consts += p.consts
            start.append(len(code))
        return DeviceProgram(code, consts, start)

    def arrays(self):
        """(start int32, code int32, consts complex128) numpy arrays."""
        return (np.asarray(self.start, np.int32), np.asarray(self.code, np.int32),
                np.asarray(self.consts if self.consts else [0j], np.complex128))

    def evaluate(self, t, k=0):
        """Host evaluation of program k, instruction for instruction as the
        device runs it (complex128)."""
        st = []
        cs = self.consts
        for ins in self.code[self.start[k]:self.start[k + 1]]:
            op, arg = ins & 0xFF, ins >> 8
            if op == CO_CONST:
                st.append(cs[arg])
            elif op == CO_T:
                st.append(complex(t))
            elif op == CO_INTERP:
                st.append(_interp_table(cs, arg, t))
            elif op == CO_NEG:
                st[-1] = -st[-1]
            elif op == CO_FN:
                st[-1] = _host_fn(arg, st[-1])
            elif op == CO_SELECT:
                b, a, c = st.pop(), st.pop(), st.pop()
                st.append(a if c.real != 0 else b)
            else:
                b, a = st.pop(), st.pop()
                st.append(_host_binop(op, a, b))
        return st[-1]


def _host_binop(op, a, b):
    if op == CO_ADD:
        return a + b
    if op == CO_SUB:
        return a - b
    if op == CO_MUL:
        return a * b
    if op == CO_DIV:
        return a / b
    if op == CO_POW:
        return complex(a ** b) if not (b.imag == 0 and a.imag == 0 and a.real > 0) \
            else complex(a.real ** b.real)
    cmp = {CO_LT: a.real < b.real, CO_LE: a.real <= b.real, CO_GT: a.real > b.real,
           CO_GE: a.real >= b.real, CO_EQ: a == b, CO_NE: a != b}[op]
    return complex(1.0 if cmp else 0.0)


def _host_fn(code, z):
    name = FUNCTIONS[code]
    if name == "abs":
        return complex(abs(z))
    if name == "norm":
        return complex(abs(z) ** 2)
    if name == "real":
        return complex(z.real)
    if name == "imag":
        return complex(z.imag)
    if name == "conj":
        return z.conjugate()
    if name == "arg":
        return complex(math.atan2(z.imag, z.real))
    if name == "erf":
        return complex(scipy.special.erf(z.real))
    return complex(STR_ENV[name](z))


def _interp_table(cs, o, t):
    npts, nrow = int(cs[o].real), int(cs[o].imag)
    dt, t0 = cs[o + 1].real, cs[o + 1].imag
    tl = [c.real for c in cs[o + 2:o + 2 + npts]]
    poly = cs[o + 2 + npts:o + 2 + npts + nrow * npts]
    if t <= tl[0]:
        return poly[(nrow - 1) * npts]
    if t >= tl[-1]:
        return poly[(nrow - 1) * npts + npts - 1]
    if dt:
        idx = int((t - t0) / dt)
    else:
        lo, hi = 0, npts
        while lo + 1 != hi:
            mid = (lo + hi) // 2
            if t < tl[mid]:
                hi = mid
            else:
                lo = mid
        idx = lo
    idx = min(max(idx, 0), npts - 1)
    if nrow == 1:
        return poly[idx]
    f = t - tl[idx]
    out = 0j
    for i in range(nrow):
        out = out * f + poly[i * npts + idx]
    return out


# ---- coefficients -------------------------------------------------------------
class Coefficient:
    """A function of time (coefficient.pyx:88-175)."""

    args = {}

    def __call__(self, t, _args=None, **kwargs):
        if _args or kwargs:
            return self.replace_arguments(_args, **kwargs)(t)
        return complex(self._call(float(t)))

    def _call(self, t):
        raise NotImplementedError

    def device_program(self):
        """The device program, or None when the coefficient can only run on
        the host."""
        return None

    def replace_arguments(self, _args=None, **kwargs):
        return self

    def copy(self):
        return self

    def conj(self):
        return ConjCoefficient(self)

    def __add__(self, other):
        if not isinstance(other, Coefficient):
            other = ConstantCoefficient(other)
        return SumCoefficient(self, other)

    __radd__ = __add__

    def __mul__(self, other):
        if not isinstance(other, Coefficient):
            other = ConstantCoefficient(other)
        return MulCoefficient(self, other)

    __rmul__ = __mul__


class ConstantCoefficient(Coefficient):
    """coefficient.pyx:916-960."""

    def __init__(self, value):
        self.value = complex(value)

    def _call(self, t):
        return self.value

    def device_program(self):
        return DeviceProgram([CO_CONST], [self.value])


class SumCoefficient(Coefficient):
    """coefficient.pyx:688-746."""

    def __init__(self, a, b):
        self.a, self.b = a, b

    def _call(self, t):
        return self.a._call(t) + self.b._call(t)

    def replace_arguments(self, _args=None, **kwargs):
        return SumCoefficient(self.a.replace_arguments(_args, **kwargs),
                              self.b.replace_arguments(_args, **kwargs))

    def device_program(self):
        pa, pb = self.a.device_program(), self.b.device_program()
        if pa is None or pb is None:
            return None
        return DeviceProgram.combine([pa, pb], CO_ADD)


class MulCoefficient(SumCoefficient):
    """coefficient.pyx:748-806."""

    def _call(self, t):
        return self.a._call(t) * self.b._call(t)

    def replace_arguments(self, _args=None, **kwargs):
        return MulCoefficient(self.a.replace_arguments(_args, **kwargs),
                              self.b.replace_arguments(_args, **kwargs))

    def device_program(self):
        pa, pb = self.a.device_program(), self.b.device_program()
        if pa is None or pb is None:
            return None
        return DeviceProgram.combine([pa, pb], CO_MUL)


class ConjCoefficient(Coefficient):
    """coefficient.pyx:808-861."""

    def __init__(self, a):
        self.a = a

    def _call(self, t):
        return np.conj(self.a._call(t))

    def replace_arguments(self, _args=None, **kwargs):
        return ConjCoefficient(self.a.replace_arguments(_args, **kwargs))

    def device_program(self):
        p = self.a.device_program()
        if p is None:
            return None
        return DeviceProgram(p.code + [CO_FN | (FN_CODE["conj"] << 8)], p.consts)


class StrCoefficient(Coefficient):
    """A string coefficient (coefficient.py:399-453 compiles it with
    Cython; StrFunctionCoefficient, coefficient.pyx:292-410, evaluates it
    with ``eval`` -- the host call here).  The device program is compiled
    from the expression's syntax tree."""

    def __init__(self, base, args=None):
        self.base = base
        self.args = dict(args or {})
        code = ("def coeff(t, args):\n"
                + "".join(f"    {k} = args['{k}']\n" for k in self.args)
                + f"    return {base}")
        lc = {}
        try:
            exec(code, dict(STR_ENV), lc)
            lc["coeff"](0.0, self.args)
        except Exception as err:
            raise Exception("Invalid string coefficient") from err
        self._func = lc["coeff"]
        self._program = None
        self._compiled = False

    def _call(self, t):
        return self._func(t, self.args)

    def replace_arguments(self, _args=None, **kwargs):
        if _args:
            kwargs.update(_args)
        if kwargs:
            return StrCoefficient(self.base, {**self.args, **kwargs})
        return self

    def copy(self):
        return StrCoefficient(self.base, self.args.copy())

    def device_program(self):
        if not self._compiled:
            self._compiled = True
            try:
                self._program = _Compiler(self.args).compile(self.base)
            except _NotCompilable:
                self._program = None
        return self._program


class _NotCompilable(Exception):
    pass


class _Compiler:
    """Syntax tree of a string coefficient -> stack program."""

    def __init__(self, args):
        self.args = args
        self.code = []
        self.consts = []

    def const(self, v):
        v = complex(v)
        for i, c in enumerate(self.consts):
            if c == v and math.copysign(1, c.real) == math.copysign(1, v.real) \
                    and math.copysign(1, c.imag) == math.copysign(1, v.imag):
                self.code.append(CO_CONST | (i << 8))
                return
        self.consts.append(v)
        self.code.append(CO_CONST | ((len(self.consts) - 1) << 8))

    def compile(self, text):
        tree = ast.parse(text.strip(), mode="eval")
        self.visit(tree.body)
        return DeviceProgram(self.code, self.consts)

    def fn_name(self, node):
        if isinstance(node, ast.Name):
            name = node.id
        elif (isinstance(node, ast.Attribute) and isinstance(node.value, ast.Name)
              and node.value.id in ("np", "spe")):
            name = node.attr
        else:
            raise _NotCompilable("callee")
        name = ALIASES.get(name, name)
        if name not in FN_CODE:
            raise _NotCompilable(name)
        return name

    def visit(self, node):
        if isinstance(node, ast.Constant):
            if isinstance(node.value, bool) or not isinstance(node.value, (int, float, complex)):
                raise _NotCompilable("constant")
            self.const(node.value)
        elif isinstance(node, ast.Name):
            if node.id == "t":
                self.code.append(CO_T)
            elif node.id in self.args:
                v = self.args[node.id]
                if not np.isscalar(v) or isinstance(v, (str, bytes)):
                    raise _NotCompilable(node.id)
                self.const(complex(v))
            elif node.id == "pi":
                self.const(np.pi)
            else:
                raise _NotCompilable(node.id)
        elif isinstance(node, ast.Attribute):
            if isinstance(node.value, ast.Name) and node.value.id == "np" and node.attr in (
                    "pi", "e"):
                self.const(getattr(np, node.attr))
            else:
                raise _NotCompilable("attribute")
        elif isinstance(node, ast.BinOp) and type(node.op) in _BINOPS:
            self.visit(node.left)
            self.visit(node.right)
            self.code.append(_BINOPS[type(node.op)])
        elif isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            self.visit(node.operand)
            if isinstance(node.op, ast.USub):
                self.code.append(CO_NEG)
        elif isinstance(node, ast.Compare) and len(node.ops) == 1 and type(node.ops[0]) in _CMPOPS:
            self.visit(node.left)
            self.visit(node.comparators[0])
            self.code.append(_CMPOPS[type(node.ops[0])])
        elif isinstance(node, ast.IfExp):
            self.visit(node.test)
            self.visit(node.body)
            self.visit(node.orelse)
            self.code.append(CO_SELECT)
        elif isinstance(node, ast.Call) and len(node.args) == 1 and not node.keywords:
            name = self.fn_name(node.func)
            self.visit(node.args[0])
            self.code.append(CO_FN | (FN_CODE[name] << 8))
        else:
            raise _NotCompilable(type(node).__name__)


class InterCoefficient(Coefficient):
    """An interpolated array (coefficient.pyx:412-560): order 0 takes the
    value on the left, order 1 linear segments, order >= 2 the spline of
    scipy.interpolate.make_interp_spline turned into piecewise polynomials
    (Horner form, constant term last)."""

    def __init__(self, coeff_arr, tlist, order=3, boundary_conditions=None):
        tlist = np.array(tlist, dtype=np.float64)
        coeff_arr = np.array(coeff_arr, dtype=np.complex128)
        if coeff_arr.ndim != 1:
            raise ValueError("The array to interpolate must be a 1D array")
        if coeff_arr.shape != tlist.shape:
            raise ValueError("tlist must be the same len as the array to interpolate")
        if order < 0:
            raise ValueError("order must be a positive integer")
        order = min(order, len(tlist) - 1)
        if order == 0:
            poly = coeff_arr.reshape((1, -1))
        elif order == 1:
            poly = np.vstack([np.diff(coeff_arr, append=-1) / np.diff(tlist, append=-1),
                              coeff_arr])
        else:
            from scipy.interpolate import make_interp_spline
            spline = make_interp_spline(tlist, coeff_arr, k=order, bc_type=boundary_conditions)
            tlist = np.sort(np.unique(np.concatenate([spline.t, tlist])))
            a = np.arange(spline.k + 1)
            a[0] = 1
            fact = np.cumprod(a)
            poly = np.concatenate([spline(tlist, i) / fact[i]
                                   for i in range(spline.k, -1, -1)]).reshape((spline.k + 1, -1))
        self.tlist = tlist
        self.poly = np.asarray(poly, dtype=np.complex128)
        self.order = self.poly.shape[0] - 1
        diff = np.diff(tlist)
        self.dt = float(diff[0]) if len(diff) >= 1 and np.allclose(diff[0], diff) else 0.0

    def _table(self):
        npts = len(self.tlist)
        return ([complex(npts, self.order + 1), complex(self.dt, self.tlist[0])]
                + [complex(t) for t in self.tlist] + list(self.poly.ravel()))

    def _call(self, t):
        return _interp_table(self._table_cache(), 0, t)

    def _table_cache(self):
        if getattr(self, "_tab", None) is None:
            self._tab = self._table()
        return self._tab

    def device_program(self):
        return DeviceProgram([CO_INTERP], self._table_cache())


class FunctionCoefficient(Coefficient):
    """A Python function of t (coefficient.pyx:176-290): ``f(t, **args)``
    (pythonic style) or ``f(t, args)`` (dict style).  Evaluated on the
    host only."""

    def __init__(self, func, args=None, style=None):
        self.func = func
        self.args = dict(args or {})
        if style is None:
            try:
                params = list(inspect.signature(func).parameters.values())
            except (TypeError, ValueError):
                params = []
            style = "dict" if (len(params) == 2 and params[1].name == "args"
                               and params[1].kind == params[1].POSITIONAL_OR_KEYWORD) \
                else "pythonic"
        self.style = style
        if style == "pythonic":
            try:
                params = inspect.signature(func).parameters
            except (TypeError, ValueError):
                params = {}
            if any(p.kind == p.VAR_KEYWORD for p in params.values()):
                self._kw = None
            else:
                self._kw = set(params) - {next(iter(params), None)}

    def _call(self, t):
        if self.style == "dict":
            return self.func(t, self.args)
        if self._kw is None:
            return self.func(t, **self.args)
        return self.func(t, **{k: v for k, v in self.args.items() if k in self._kw})

    def replace_arguments(self, _args=None, **kwargs):
        if _args:
            kwargs.update(_args)
        if kwargs:
            return FunctionCoefficient(self.func, {**self.args, **kwargs}, self.style)
        return self


def coefficient(base, *, tlist=None, args=None, order=3, boundary_conditions=None,
                function_style=None, **kwargs):
    """Build a Coefficient from a string, an array (with ``tlist``), a
    function, a number or a Coefficient (coefficient.py:63-229)."""
    args = dict(args or {})
    if isinstance(base, Coefficient):
        return base.replace_arguments(args) if args else base
    if isinstance(base, str):
        return StrCoefficient(base, args)
    if isinstance(base, np.ndarray) or isinstance(base, (list, tuple)):
        if tlist is None:
            raise TypeError("an array coefficient needs tlist")
        return InterCoefficient(base, tlist, order, boundary_conditions)
    if callable(base):
        return FunctionCoefficient(base, args, function_style)
    if np.isscalar(base):
        return ConstantCoefficient(base)
    raise TypeError(f"coefficient cannot be made from {type(base)}")
