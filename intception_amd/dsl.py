"""Input-script DSL: the names an Intception input script uses, for this build.

Mirrors the reference's public input-script vocabulary
(src/intception/__init__.py:28-34): ``dsl_cartesian_gaussian``,
``dsl_integer_index``, ``dsl_scalar``, ``dsl_position``, ``dsl_value``,
``dsl_binop``/``dsl_unop``/``dsl_function``, ``exp``/``sqrt``/``norm``/``nint``,
``boys_f``, ``dsl_integral`` with ``set_base``/``add_vrr``/``add_hrr``/``int``.
An input script written for the reference (e.g. the Coulomb ERI script of
``intception_amd.eri_script``) runs unchanged against this module.

The objects form an expression tree that can be *evaluated numerically* in a
context (exponents, centres, target index components, Cartesian direction).
The HIP backend (``intception_amd.generator``) uses that to check an
integral's base expression and recurrences against the Obara-Saika /
Head-Gordon-Pople family its kernels implement, instead of emitting code from
the tree.

Semantics follow the reference: a recurrence expression is written for the
target integral and one Cartesian direction i; positions (PA, WP, ...) stand
for their i-th component; a Cartesian index used as a coefficient stands for
the i-th component of the target's angular momentum (``ga - 1`` -> a_i - 1);
``integral.int(ga - 1, gb, m + 1)`` is the integral with a lowered by 1_i and m
raised by 1 (src/intception/dsl_extensions.py:822-950,
src/intception/dsl.py:926-1046).
"""
import math
import operator


class dsl_base:
    """Expression node."""

    def _wrap(self, o):
        return o if isinstance(o, dsl_base) else dsl_value(None, o)

    def __add__(self, o): return dsl_binop("+", self, self._wrap(o))
    def __radd__(self, o): return dsl_binop("+", self._wrap(o), self)
    def __sub__(self, o): return dsl_binop("-", self, self._wrap(o))
    def __rsub__(self, o): return dsl_binop("-", self._wrap(o), self)
    def __mul__(self, o): return dsl_binop("*", self, self._wrap(o))
    def __rmul__(self, o): return dsl_binop("*", self._wrap(o), self)
    def __truediv__(self, o): return dsl_binop("/", self, self._wrap(o))
    def __rtruediv__(self, o): return dsl_binop("/", self._wrap(o), self)
    def __pow__(self, o): return dsl_binop("**", self, self._wrap(o))
    def __neg__(self): return dsl_unop("-", self)

    def ev(self, ctx):  # numeric value in a context
        raise NotImplementedError(type(self).__name__)

    def lin(self, ctx):
        """Linear form over integral references: {ref_key: coefficient}."""
        return {None: self.ev(ctx)}

    def name(self):
        return getattr(self, "_name", None)


_OPS = {"+": operator.add, "-": operator.sub, "*": operator.mul, "/": operator.truediv, "**": operator.pow}


def _lin_add(a, b, sign=1.0):
    out = dict(a)
    for k, v in b.items():
        out[k] = out.get(k, 0.0) + sign * v
    return out


class dsl_binop(dsl_base):
    def __init__(self, op, left, right):
        self._op, self._l, self._r = op, left, right

    def op(self): return self._op
    def left(self): return self._l
    def right(self): return self._r

    def ev(self, ctx):
        return _OPS[self._op](self._l.ev(ctx), self._r.ev(ctx))

    def lin(self, ctx):
        if self._op in "+-":
            return _lin_add(self._l.lin(ctx), self._r.lin(ctx), 1.0 if self._op == "+" else -1.0)
        if self._op == "*":
            a, b = self._l.lin(ctx), self._r.lin(ctx)
            if set(a) == {None}:
                return {k: a[None] * v for k, v in b.items()}
            if set(b) == {None}:
                return {k: b[None] * v for k, v in a.items()}
            raise ValueError("product of two integrals is not a recurrence term")
        if self._op == "/":
            a, b = self._l.lin(ctx), self._r.lin(ctx)
            assert set(b) == {None}, "division by an integral"
            return {k: v / b[None] for k, v in a.items()}
        return {None: self.ev(ctx)}

    def __str__(self):
        return "(%s %s %s)" % (self._l, self._op, self._r)


class dsl_unop(dsl_base):
    def __init__(self, op, arg):
        self._op, self._a = op, arg

    def ev(self, ctx):
        v = self._a.ev(ctx)
        return {"-": lambda x: -x, "exp": math.exp, "sqrt": math.sqrt, "int": int,
                "round": round}[self._op](v)

    def lin(self, ctx):
        if self._op == "-":
            return {k: -v for k, v in self._a.lin(ctx).items()}
        return {None: self.ev(ctx)}

    def __str__(self):
        return "%s(%s)" % (self._op, self._a)


class dsl_value(dsl_base):
    def __init__(self, name=None, value=0.0, vartype="double"):
        self._name, self._value, self._vartype = name, value, vartype

    def value(self):
        return self._value

    def ev(self, ctx):
        v = self._value
        return float(v) if not isinstance(v, str) else float(eval(v, {"__builtins__": {}}, {}))

    def __str__(self):
        return self._name or repr(self._value)


class dsl_zero(dsl_value):
    def __init__(self):
        super().__init__("0", 0.0)


class dsl_scalar(dsl_base):
    def __init__(self, name, expr=None, vartype="double"):
        self._name, self._expr, self._vartype = name, expr, vartype

    def expr(self):
        return self._expr

    def ev(self, ctx):
        if self._expr is None:
            return ctx["vars"][self._name]
        return self._expr.ev(ctx)

    def __str__(self):
        return self._name


class dsl_position(dsl_base):
    """3-vector; evaluates to its component in the context's direction."""

    def __init__(self, name=None, expr=None):
        self._name, self._expr = name, expr

    def expr(self):
        return self._expr

    def ev(self, ctx):
        if self._expr is None:
            return ctx["vars"][self._name][ctx["dir"]]
        return self._expr.ev(ctx)

    def __getitem__(self, k):
        return _component(self, k)

    def __str__(self):
        return self._name or str(self._expr)


class _component(dsl_base):
    def __init__(self, pos, k):
        self._pos, self._k = pos, k

    def ev(self, ctx):
        c = dict(ctx, dir=self._k)
        return self._pos.ev(c)

    def __str__(self):
        return "%s[%d]" % (self._pos, self._k)


class dsl_function(dsl_base):
    def __init__(self, name, args, vartype="double"):
        self._name, self._args = name, list(args)

    def ev(self, ctx):
        f = {"exp": math.exp, "sqrt": math.sqrt}[self._name]
        return f(*[self._wrap(a).ev(ctx) for a in self._args])

    def __str__(self):
        return "%s(%s)" % (self._name, ", ".join(str(a) for a in self._args))


def exp(x):
    return dsl_function("exp", [x])


def sqrt(x):
    return dsl_function("sqrt", [x])


def norm(v):
    return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])


def nint(x):
    return dsl_unop("int", dsl_unop("round", x))


class _Shifted(dsl_base):
    """index + k: a shifted index inside int(...), and (as a coefficient) the
    target's component plus k."""

    def __init__(self, index, k):
        self._index, self._k = index, k

    def ev(self, ctx):
        return self._index.ev(ctx) + self._k

    def __add__(self, o):
        if isinstance(o, int):
            return _Shifted(self._index, self._k + o)
        return dsl_base.__add__(self, o)

    def __sub__(self, o):
        if isinstance(o, int):
            return _Shifted(self._index, self._k - o)
        return dsl_base.__sub__(self, o)

    def __str__(self):
        return "%s%+d" % (self._index.name(), self._k)


class _Index(dsl_base):
    def __add__(self, o):
        if isinstance(o, int):
            return _Shifted(self, o)
        return dsl_base.__add__(self, o)

    def __sub__(self, o):
        if isinstance(o, int):
            return _Shifted(self, -o)
        return dsl_base.__sub__(self, o)

    def __str__(self):
        return self._name


class dsl_cartesian_gaussian(_Index):
    """Cartesian Gaussian index with exponent and centre attachments
    (src/intception/dsl.py:926-1046)."""

    def __init__(self, name, constant=True):
        self._name = name
        self._exp = dsl_scalar("x" + name)
        self._cen = dsl_position("c" + name)

    def exp(self):
        return self._exp

    def cen(self):
        return self._cen

    def ev(self, ctx):
        return ctx["index"][self._name][ctx["dir"]]


class dsl_integer_index(_Index):
    """Auxiliary integer index, e.g. m (src/intception/dsl.py:860-923)."""

    def __init__(self, name):
        self._name = name

    def ev(self, ctx):
        return ctx["aux"][self._name]


class boys_f(dsl_base):
    """Boys function F_m(x) (src/intception/boys.py:209-233)."""

    def __init__(self, m, x, varname="boys_f"):
        self._m, self._x = m, x

    def m(self):
        return self._m

    def x(self):
        return self._x

    def ev(self, ctx):
        return ctx["boys"](int(self._m.ev(ctx)), self._x.ev(ctx))

    def __str__(self):
        return "F_%s(%s)" % (self._m, self._x)


class boys_function:
    """Placeholder for the reference's table-backed Boys implementation
    object; this build's Boys function lives in the device library."""


class _IntRef(dsl_base):
    def __init__(self, integral, shifts):
        self._integral, self._shifts = integral, shifts

    def lin(self, ctx):
        return {self._shifts: 1.0}

    def ev(self, ctx):
        raise ValueError("integral reference has no scalar value")

    def __str__(self):
        return "%s(%s)" % (self._integral.name(), ",".join("%s%+d" % kv for kv in self._shifts))


class dsl_rr:
    def __init__(self, name, rrtype, index, expr, move_from=None):
        self._name, self._type, self._index, self._expr, self._from = name, rrtype, index, expr, move_from

    def name(self): return self._name
    def rrtype(self): return self._type
    def changing_index(self): return self._index
    def expr(self): return self._expr
    def move_from(self): return self._from


class dsl_integral(dsl_base):
    """An integral class: Cartesian indices, auxiliary indices, extra
    arguments, a base expression and recurrences
    (src/intception/dsl_extensions.py:369-962)."""

    def __init__(self, *args, name=None, **kw):
        self._name = name
        self._index_list = [a for a in args if isinstance(a, dsl_cartesian_gaussian)]
        self._aux_list = [a for a in args if isinstance(a, dsl_integer_index)]
        self._arg_list = [a for a in args if not isinstance(a, (dsl_cartesian_gaussian, dsl_integer_index))]
        self._base = None
        self._rr_list = []

    def name(self):
        return self._name

    def index_list(self):
        return list(self._index_list)

    def aux_list(self):
        return list(self._aux_list)

    def arg_list(self):
        return list(self._arg_list)

    def base(self):
        return self._base

    def rr_list(self):
        return list(self._rr_list)

    def full_name(self):
        parts = ["intception_integral_", self._name + "_"]
        for l in (self._index_list, self._aux_list, self._arg_list):
            for o in l:
                parts.append(o.name())
            if l:
                parts.append("_")
        if parts[-1] == "_":
            parts.pop()
        return "".join(parts)

    def set_base(self, expr):
        self._base = expr

    def add_vrr(self, name, index, expr):
        assert index in self._index_list, "index being incremented must be in integral definition"
        assert not any(r.rrtype() == "hrr" for r in self._rr_list), "A VRR may not follow a HRR."
        self._rr_list.append(dsl_rr(name, "vrr", index, expr))

    def add_hrr(self, name, move_from, move_to):
        """HGP HRR (a, b) = (a + 1_i, b - 1_i) + (A - B)_i (a, b - 1_i),
        moving angular momentum from move_from to move_to."""
        assert any(r.changing_index() is move_from for r in self._rr_list), \
            "move_from must be incremented by a preceding RR operation"
        R = dsl_position(expr=move_from.cen() - move_to.cen())
        expr = self.int(move_from + 1, move_to - 1) + R * self.int(move_from, move_to - 1)
        self._rr_list.append(dsl_rr(name, "hrr", move_to, expr, move_from))

    def int(self, *args):
        shifts = {}
        for a in args:
            if isinstance(a, _Shifted):
                shifts[a._index.name()] = a._k
            elif isinstance(a, _Index):
                shifts[a.name()] = 0
            else:
                raise ValueError("int() takes indices of the integral")
        names = [i.name() for i in self._index_list + self._aux_list]
        for k in shifts:
            assert k in names, "index %s not in integral %s" % (k, self._name)
        key = tuple((n, shifts.get(n, 0)) for n in names)
        return _IntRef(self, key)
