"""Test and trial functions as SymPy applied functions (reference:
galerkin/arguments.py:342-460).  Each carries its function space and an
argument tag (0 = test, 1 = trial, 2 = function with known coefficients) as
class attributes."""

from __future__ import annotations

import sympy as sp


class ArgumentTag:
    TEST = 0
    TRIAL = 1
    JAXFUNC = 2


class _Data:
    """Identity-compared holder for a device array inside SymPy metadata."""

    def __init__(self, value):
        self.value = value


class _DataDummy(sp.Dummy):
    """The extra argument of a JAXFunction: a Dummy that carries the
    function's (array, space, name), so that SymPy can rebuild the function
    from its arguments (``func(*args)``) without losing its coefficients,
    and the data lives exactly as long as expressions that use it."""


class JAXFunction(sp.Function):
    """A function of V with explicit expansion coefficients
    (galerkin/arguments.py:709-900): ``u(x) = sum_i array_i phi_i(x)``.

    ``array`` is a coefficient array (or a SymPy expression, projected onto
    V).  In a form, a JAXFunction that appears linearly (to the first power,
    possibly differentiated, in a term without a trial function) contributes
    ``A @ array`` with the Galerkin matrix A of that term; anywhere else
    (products with the trial function, powers, ...) it is evaluated on the
    quadrature mesh as a variable coefficient.
    """

    argument = ArgumentTag.JAXFUNC

    def __new__(cls, array, V=None, name=None, *rest):
        if isinstance(V, sp.Basic) or V is None:
            # SymPy rebuilding the object from its arguments
            args = (array, V, name) + rest if V is not None else (array,)
            args = tuple(a for a in args if a is not None)
            data = args[-1].jaxf_data
            obj = sp.Function.__new__(cls, *args)
            obj.array, obj.functionspace, obj.name = data
            return obj
        from .inner import project
        from .orthogonal import as_device

        if isinstance(array, sp.Basic):
            array = project(array, V)
        d = _DataDummy()
        name = name if name is not None else "JAXFunction"
        arr = as_device(array)
        d.jaxf_data = (arr, V, name)
        obj = sp.Function.__new__(cls, *(tuple(V.system.base_scalars()) + (d,)))
        obj.array, obj.functionspace, obj.name = arr, V, name
        return obj

    def get_array(self):
        return self.array

    def backward(self, N=None):
        return self.functionspace.backward(self.array, N)

    def __call__(self, *x):
        """Value at a point of a 1D space (galerkin/arguments.py:850-880)."""
        import numpy as np

        return self.functionspace.evaluate(np.array(x, dtype=float).reshape(-1), self.array)[0]

    def doit(self, **hints):
        return self

    def _sympystr(self, printer):
        return self.name

    def __matmul__(self, a):
        return self.array @ a

    def __rmatmul__(self, a):
        return a @ self.array


def is_jaxfunction(e) -> bool:
    if isinstance(e, sp.Derivative):
        e = e.expr
    return isinstance(e, JAXFunction)


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
    if isinstance(e, JAXFunction):
        return e.functionspace
    return getattr(getattr(e, "func", None), "functionspace", None)
