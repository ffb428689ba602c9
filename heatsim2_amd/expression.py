"""Symbolic expressions for boundary face laws.

Boundary plugins build their face flux from placeholder arguments with
ordinary arithmetic, ``[a, b]`` indexing of tensor conductivities and
``group()`` around the temperature differences that the time integrator
treats implicitly -- the protocol of heatsim2's expression engine
(expression.py:15-188 for the operators, :108-131 for indexing,
:703-721 for group / eliminate_groups / no_groups, :724-750 for subst).

This module keeps that protocol but not the upstream's RPN rewriting
machinery: an expression is a small immutable tree, and
:func:`linear_form` evaluates it numerically once the conductivity
placeholders are bound, returning the coefficient of every temperature
variable together with the group it sits in.  heatsim2_amd.crank_nicolson
uses that to turn each face law into one conductance per
(face class, axis, material pair).
"""
import numbers

import numpy as np

__all__ = ["linear_expression", "group", "eliminate_groups", "no_groups", "subst",
           "linear_form", "NonLinearError"]


class NonLinearError(ValueError):
    """The expression is not linear in the temperature variables."""


_ADD, _MUL, _DIV, _NEG, _GROUP, _INDEX, _CONST, _VAR = range(8)


def _wrap(v):
    if isinstance(v, linear_expression):
        return v
    return linear_expression(v)


class linear_expression(object):
    """Immutable expression tree.  ``linear_expression(3.0)`` is a constant,
    ``linear_expression('T555')`` a variable, ``linear_expression()`` the
    constant 0."""

    __slots__ = ("node",)

    def __init__(self, *args):
        if len(args) == 0:
            self.node = (_CONST, 0.0)
        elif len(args) == 1:
            v = args[0]
            if isinstance(v, linear_expression):
                self.node = v.node
            elif isinstance(v, (numbers.Number, np.ndarray)):
                self.node = (_CONST, v)
            elif isinstance(v, str):
                self.node = (_VAR, v)
            else:
                raise ValueError("Unknown argument type")
        else:
            raise ValueError("Too many arguments (%d)" % len(args))

    @classmethod
    def _make(cls, node):
        e = cls.__new__(cls)
        e.node = node
        return e

    # arithmetic ---------------------------------------------------------
    def __add__(self, o):
        return self._make((_ADD, self, _wrap(o)))

    def __radd__(self, o):
        return self._make((_ADD, _wrap(o), self))

    def __sub__(self, o):
        return self._make((_ADD, self, self._make((_NEG, _wrap(o)))))

    def __rsub__(self, o):
        return self._make((_ADD, _wrap(o), self._make((_NEG, self))))

    def __neg__(self):
        return self._make((_NEG, self))

    def __mul__(self, o):
        return self._make((_MUL, self, _wrap(o)))

    def __rmul__(self, o):
        return self._make((_MUL, _wrap(o), self))

    def __truediv__(self, o):
        return self._make((_DIV, self, _wrap(o)))

    def __rtruediv__(self, o):
        return self._make((_DIV, _wrap(o), self))

    __div__ = __truediv__
    __rdiv__ = __rtruediv__

    def __getitem__(self, index):
        if not (isinstance(index, tuple) and len(index) == 2):
            raise ValueError("index with two subscripts, e.g. k[0,0]")
        return self._make((_INDEX, self, (int(index[0]), int(index[1]))))

    def __repr__(self):
        return "linear_expression(%s)" % _fmt(self)

    __str__ = __repr__

    def exprstr(self):
        return _fmt(self)


def _fmt(e):
    n = e.node
    op = n[0]
    if op == _CONST:
        return "%g" % n[1] if not isinstance(n[1], np.ndarray) else str(n[1].tolist())
    if op == _VAR:
        return n[1]
    if op == _ADD:
        return "(%s + %s)" % (_fmt(n[1]), _fmt(n[2]))
    if op == _MUL:
        return "(%s * %s)" % (_fmt(n[1]), _fmt(n[2]))
    if op == _DIV:
        return "(%s / %s)" % (_fmt(n[1]), _fmt(n[2]))
    if op == _NEG:
        return "-%s" % _fmt(n[1])
    if op == _GROUP:
        return "GROUP(%s)" % _fmt(n[1])
    if op == _INDEX:
        return "%s[%d,%d]" % (_fmt(n[1]), n[2][0], n[2][1])
    raise ValueError(op)


def group(expr):
    """Mark a sub-expression (normally a temperature difference) as one unit
    that the integrator may treat implicitly (expression.py:703-710)."""
    return linear_expression._make((_GROUP, _wrap(expr)))


def eliminate_groups(expr):
    """The same expression without group markers (expression.py:712-716)."""
    e = _wrap(expr)
    n = e.node
    if n[0] == _GROUP:
        return eliminate_groups(n[1])
    if n[0] in (_ADD, _MUL, _DIV):
        return linear_expression._make((n[0], eliminate_groups(n[1]), eliminate_groups(n[2])))
    if n[0] in (_NEG,):
        return linear_expression._make((n[0], eliminate_groups(n[1])))
    if n[0] == _INDEX:
        return linear_expression._make((n[0], eliminate_groups(n[1]), n[2]))
    return e


def no_groups(expr):
    """True if the expression holds no group marker (expression.py:718-721)."""
    n = _wrap(expr).node
    if n[0] == _GROUP:
        return False
    if n[0] in (_ADD, _MUL, _DIV):
        return no_groups(n[1]) and no_groups(n[2])
    if n[0] in (_NEG, _INDEX):
        return no_groups(n[1])
    return True


def subst(expr, name, value):
    """Replace variable ``name`` by a variable name (str) or a value
    (expression.py:724-750)."""
    e = _wrap(expr)
    n = e.node
    if n[0] == _VAR:
        if n[1] != name:
            return e
        return linear_expression(value)
    if n[0] in (_ADD, _MUL, _DIV):
        return linear_expression._make((n[0], subst(n[1], name, value), subst(n[2], name, value)))
    if n[0] in (_NEG, _GROUP):
        return linear_expression._make((n[0], subst(n[1], name, value)))
    if n[0] == _INDEX:
        return linear_expression._make((n[0], subst(n[1], name, value), n[2]))
    return e


# ---------------------------------------------------------- evaluation
# A value during evaluation is (const, terms) with terms a dict
# {(varname, group_id): coefficient}; group_id None means "not grouped".

def _scale(val, s):
    c, terms = val
    return (c * s, {k: v * s for k, v in terms.items()})


def _is_const(val):
    return not val[1]


def _eval(e, env, gid):
    n = e.node
    op = n[0]
    if op == _CONST:
        return (n[1], {})
    if op == _VAR:
        if n[1] in env:
            return (env[n[1]], {})
        return (0.0, {(n[1], gid[0]): 1.0})
    if op == _NEG:
        return _scale(_eval(n[1], env, gid), -1.0)
    if op == _ADD:
        a = _eval(n[1], env, gid)
        b = _eval(n[2], env, gid)
        terms = dict(a[1])
        for k, v in b[1].items():
            terms[k] = terms.get(k, 0.0) + v
        return (a[0] + b[0], terms)
    if op == _MUL:
        a = _eval(n[1], env, gid)
        b = _eval(n[2], env, gid)
        if _is_const(a):
            return _scale(b, a[0])
        if _is_const(b):
            return _scale(a, b[0])
        raise NonLinearError("product of two temperature-dependent factors")
    if op == _DIV:
        a = _eval(n[1], env, gid)
        b = _eval(n[2], env, gid)
        if not _is_const(b):
            raise NonLinearError("division by a temperature-dependent factor")
        return _scale(a, 1.0 / b[0])
    if op == _GROUP:
        outer = gid[0]
        gid[1] += 1
        gid[0] = gid[1]
        try:
            v = _eval(n[1], env, gid)
        finally:
            gid[0] = outer
        return v
    if op == _INDEX:
        a = _eval(n[1], env, gid)
        if not _is_const(a):
            raise NonLinearError("indexing a temperature-dependent expression")
        base = a[0]
        # a scalar conductivity cannot be indexed: the upstream raises a
        # TypeError here too (expression.py:737)
        return (base[n[2][0], n[2][1]], {})
    raise ValueError("unknown node %r" % (op,))


def linear_form(expr, env):
    """Evaluate ``expr`` with the placeholders in ``env`` bound to numbers.

    Returns ``(constant, terms)`` where terms maps (variable, group_id) to a
    float coefficient; group_id is None for a term outside every group,
    otherwise an integer shared by the terms of one group.  Zero
    coefficients are dropped.  Raises NonLinearError for non-linear use of
    temperatures and ValueError when a coefficient is not a real scalar
    (e.g. a tensor conductivity given to an isotropic face law).
    """
    if isinstance(expr, numbers.Number):
        return float(expr), {}
    c, terms = _eval(_wrap(expr), env, [None, 0])
    out = {}
    for k, v in terms.items():
        if isinstance(v, np.ndarray):
            if v.size != 1:
                raise ValueError("coefficient of %s is an array; tensor conductivities need "
                                 "boundary_conducting_anisotropic" % k[0])
            v = v.item()
        v = float(v)
        if v != 0.0:
            out[k] = v
    if isinstance(c, np.ndarray):
        if c.size != 1:
            raise ValueError("constant term is an array")
        c = c.item()
    return float(c), out
