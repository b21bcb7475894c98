"""Trace an out-of-place right-hand side f(u, p, t) into HIP C.

The reference's solvers call the user's Julia function f(u, p, t) -> SVector
inside the solve loop (gpuatsit5.jl:241,259-270).  For the GPU the function is
evaluated once on tracer values: every arithmetic operation on them records a
node, in the order Python evaluates it, and the recorded program is emitted as
straight-line C (one `const T vK = ...;` per node).  The library compiles the
integrators with floating-point contraction off, so the device evaluates
exactly the recorded operations in the recorded order: a traced RHS gives the
same bits as the same function evaluated in binary64 on the CPU.

Supported inside f: + - * / unary -, ** with a small integer exponent
(expanded into multiplications, as Julia's literal powers are), abs, and the
math functions sqrt, exp, log, sin, cos, tan, tanh, sinh, cosh, atan, plus
numeric constants and captured numpy arrays (emitted as literals).  Control
flow that depends on u, p or t cannot be traced and raises.
"""
import math
import numbers

import numpy as np

_FUNCS = ("sqrt", "exp", "log", "sin", "cos", "tan", "tanh", "sinh", "cosh", "atan")


class TraceError(TypeError):
    pass


class _Tape:
    def __init__(self):
        self.lines = []
        self.count = 0

    def emit(self, expr):
        name = "v%d" % self.count
        self.count += 1
        self.lines.append("  const T %s = %s;" % (name, expr))
        return Sym(self, name)


def _lit(x):
    x = float(x)
    if math.isnan(x) or math.isinf(x):
        raise TraceError("non-finite constant in the right-hand side")
    if x == int(x) and abs(x) < 2 ** 53:
        return "T(%d.0)" % int(x)
    return "T(%s)" % x.hex()


def _operand(tape, x):
    if isinstance(x, Sym):
        return x.name
    if isinstance(x, (numbers.Real, np.floating, np.integer)):
        return _lit(x)
    raise TraceError("cannot use %r inside a traced right-hand side" % type(x).__name__)


class Sym:
    """A traced scalar."""

    __array_priority__ = 1000

    def __init__(self, tape, name):
        self.tape = tape
        self.name = name

    def _bin(self, other, op, rev=False):
        a, b = _operand(self.tape, self), _operand(self.tape, other)
        if rev:
            a, b = b, a
        return self.tape.emit("%s %s %s" % (a, op, b))

    def __add__(self, o):
        return self._bin(o, "+")

    def __radd__(self, o):
        return self._bin(o, "+", True)

    def __sub__(self, o):
        return self._bin(o, "-")

    def __rsub__(self, o):
        return self._bin(o, "-", True)

    def __mul__(self, o):
        return self._bin(o, "*")

    def __rmul__(self, o):
        return self._bin(o, "*", True)

    def __truediv__(self, o):
        return self._bin(o, "/")

    def __rtruediv__(self, o):
        return self._bin(o, "/", True)

    def __neg__(self):
        return self.tape.emit("-%s" % self.name)

    def __pos__(self):
        return self

    def __abs__(self):
        return self.tape.emit("fabs(%s)" % self.name)

    def __pow__(self, k):
        if isinstance(k, (int, np.integer)) and 1 <= int(k) <= 8:
            # repeated multiplication, left to right (x*x*x for k = 3)
            r = self
            for _ in range(int(k) - 1):
                r = r * self
            return r
        raise TraceError("only small positive integer powers can be traced")

    def __bool__(self):
        raise TraceError("control flow on traced values is not supported")

    def _cmp(self, other):
        raise TraceError("comparisons of traced values (control flow) are not supported")

    __lt__ = __le__ = __gt__ = __ge__ = _cmp

    def __float__(self):
        raise TraceError("a traced value cannot be converted to a number")

    # math functions: tracer.sin(x) or numpy's np.sin(x) (which calls x.sin())
    def _fn(self, name):
        return self.tape.emit("%s(%s)" % (name, self.name))

    def sqrt(self):
        return self._fn("sqrt")

    def exp(self):
        return self._fn("exp")

    def log(self):
        return self._fn("log")

    def sin(self):
        return self._fn("sin")

    def cos(self):
        return self._fn("cos")

    def tan(self):
        return self._fn("tan")

    def tanh(self):
        return self._fn("tanh")

    def sinh(self):
        return self._fn("sinh")

    def cosh(self):
        return self._fn("cosh")

    def arctan(self):
        return self._fn("atan")


def _make_fn(name):
    def f(x):
        if isinstance(x, Sym):
            return x._fn(name)
        return getattr(math, name)(x)
    f.__name__ = name
    return f


# Functions a traced RHS may call: sdeh.sin(u[0]), ...
for _n in _FUNCS:
    globals()[_n] = _make_fn(_n)


class _Vec:
    def __init__(self, tape, base, n, kind):
        self.tape, self.base, self.n, self.kind = tape, base, n, kind
        self.cache = {}

    def __len__(self):
        return self.n

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[j] for j in range(*i.indices(self.n))]
        i = int(i)
        if i < 0:
            i += self.n
        if not 0 <= i < self.n:
            raise IndexError("%s[%d] out of range (size %d)" % (self.base, i, self.n))
        if i not in self.cache:
            self.cache[i] = Sym(self.tape, "%s[%d]" % (self.base, i))
        return self.cache[i]

    def __iter__(self):
        return (self[i] for i in range(self.n))


def trace(f, n_state, n_param, fn_name="sdeh_user_f", scalar_u=False, scalar_p=False):
    """Trace f(u, p, t) and return (hip_c_source, fn_name).  With scalar_u /
    scalar_p (a problem whose u0 / p is a scalar, as the reference allows), f
    receives u / p as a single value instead of a vector.

    The source defines
        template <class T> __device__ inline void fn_name(T* du, const T* u,
            const T* p, const T* ps, double t)
    as sdeh_compile_rhs() expects."""
    tape = _Tape()
    u = _Vec(tape, "u", n_state, "u")
    p = _Vec(tape, "p", n_param, "p")
    if scalar_u:
        if n_state != 1:
            raise TraceError("a scalar state has one component")
        u = u[0]
    if scalar_p:
        if n_param != 1:
            raise TraceError("a scalar parameter has one component")
        p = p[0]
    t = Sym(tape, "(T)t")
    out = f(u, p, t)
    out = list(out) if not isinstance(out, Sym) else [out]
    if len(out) != n_state:
        raise TraceError("f returned %d components, expected %d" % (len(out), n_state))
    body = list(tape.lines)
    for i, o in enumerate(out):
        body.append("  du[%d] = %s;" % (i, _operand(tape, o)))
    src = ["// Traced right-hand side (straight-line, operation order as recorded).",
           "template <class T>",
           "__device__ inline void %s(T* du, const T* u, const T* p, const T* ps, double t) {"
           % fn_name]
    src += body
    src.append("}")
    return "\n".join(src) + "\n", fn_name
