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


def _lifted(V) -> bool:
    from .composite import DirectSum
    from .tensorproductspace import DirectSumTPS

    return isinstance(V, (DirectSum, DirectSumTPS))


def TestFunction(V, name: str | None = None):
    """Test function; on a lifted space it is the homogeneous part's
    (galerkin/arguments.py:342-390)."""
    if _lifted(V):
        V = V.get_homogeneous()
    f = sp.Function(name or "v", functionspace=V, argument=ArgumentTag.TEST)
    return f(*_space_symbols(V))


def TrialFunction(V, name: str | None = None, transient: bool = False):
    """Trial function.  On a DirectSum it is the sum of a homogeneous and a
    lifting part; on a DirectSumTPS the sum over every tensor product of
    homogeneous and lifting factors (``TrialFunction.doit``,
    galerkin/arguments.py:446-456).  Forms split the lifting parts into the
    right-hand side."""
    args = _space_symbols(V) + ((t,) if transient else ())
    name = name or "u"
    if _lifted(V):
        parts = list(V.tpspaces.values()) if hasattr(V, "tpspaces") else list(V.basespaces)
        terms = []
        for k, P in enumerate(parts):
            nm = name if k == 0 else f"{name}_b{k}"
            f = sp.Function(nm, functionspace=P, argument=ArgumentTag.TRIAL, transient=transient)
            terms.append(f(*args))
        return sp.Add(*terms)
    f = sp.Function(name, functionspace=V, argument=ArgumentTag.TRIAL, transient=transient)
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
