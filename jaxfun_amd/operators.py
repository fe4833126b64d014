"""Cartesian differential operators on SymPy expressions (operators.py:586-671).

``Grad`` returns the tuple of first derivatives, ``Div`` sums the derivatives
of a vector's components, ``Dot`` contracts two vectors; all over the base
scalars of the expression's function space (or x, y, z).
"""

from __future__ import annotations

import sympy as sp

from .coordinates import x, y, z


def _symbols_of(expr) -> tuple:
    for f in sp.preorder_traversal(expr):
        fs = getattr(getattr(f, "func", None), "functionspace", None)
        if fs is not None:
            return fs.system.base_scalars()
    free = expr.free_symbols if hasattr(expr, "free_symbols") else set()
    syms = tuple(s for s in (x, y, z) if s in free)
    return syms if syms else (x,)


class Vector(tuple):
    """A Cartesian vector of SymPy expressions."""

    symbols: tuple = ()


def Grad(f) -> Vector:
    syms = _symbols_of(f)
    v = Vector(sp.diff(f, s) for s in syms)
    v.symbols = syms
    return v


def Div(v) -> sp.Expr:
    if isinstance(v, Vector):
        syms = v.symbols or _symbols_of(sum(v))
        return sp.Add(*[sp.diff(c, s) for c, s in zip(v, syms)])
    raise TypeError("Div expects a vector (e.g. Grad(u))")


def Dot(a, b) -> sp.Expr:
    if isinstance(a, Vector) and isinstance(b, Vector):
        return sp.Add(*[ai * bi for ai, bi in zip(a, b)])
    raise TypeError("Dot expects two vectors")
