"""FunctionSpace factory (reference: galerkin/functionspace.py:63-173)."""

from __future__ import annotations

from .Chebyshev import CGComposite, Chebyshev
from .composite import BoundaryConditions, Composite
from .Legendre import Legendre, LGComposite


def FunctionSpace(N: int, space: type, bcs=None, domain=None, system=None, name: str | None = None,
                  fun_str: str | None = None, **kw):
    """Return ``space(N)`` or its homogeneous-Dirichlet composite.

    Same signature as the reference; inhomogeneous boundary values (the
    reference's DirectSum lifting) raise NotImplementedError.
    """
    if name is not None:
        kw["name"] = name
    if fun_str is not None:
        kw["fun_str"] = fun_str
    if bcs is None:
        return space(N, domain=domain, system=system, **kw)
    bcs = BoundaryConditions(bcs, domain=domain)
    if not bcs.is_homogeneous():
        raise NotImplementedError("inhomogeneous boundary conditions (DirectSum lifting) are not built")
    comp = LGComposite if issubclass(space, Legendre) else CGComposite if issubclass(space, Chebyshev) \
        else Composite
    return comp(N, space, bcs=bcs, domain=domain, system=system, **kw)
