"""Test and trial functions as SymPy applied functions (reference:
galerkin/arguments.py:342-460).  Each carries its function space and an
argument tag (0 = test, 1 = trial) as class attributes."""

from __future__ import annotations

import sympy as sp


class ArgumentTag:
    TEST = 0
    TRIAL = 1


t = sp.Symbol("t", real=True)


def _space_symbols(V):
    return tuple(V.system.base_scalars())


def TestFunction(V, name: str | None = None):
    f = sp.Function(name or "v", functionspace=V, argument=ArgumentTag.TEST)
    return f(*_space_symbols(V))


def TrialFunction(V, name: str | None = None, transient: bool = False):
    f = sp.Function(name or "u", functionspace=V, argument=ArgumentTag.TRIAL, transient=transient)
    args = _space_symbols(V) + ((t,) if transient else ())
    return f(*args)


def get_arg(e) -> int | None:
    """ArgumentTag of a (possibly differentiated) test/trial function, else None."""
    if isinstance(e, sp.Derivative):
        e = e.expr
    return getattr(getattr(e, "func", None), "argument", None)


def functionspace_of(e):
    if isinstance(e, sp.Derivative):
        e = e.expr
    return getattr(getattr(e, "func", None), "functionspace", None)
