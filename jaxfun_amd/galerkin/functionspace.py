"""FunctionSpace factory (reference: galerkin/functionspace.py:63-173)."""

from __future__ import annotations

from .Chebyshev import CGComposite, Chebyshev
from .composite import BCGeneric, BoundaryConditions, Composite, DirectSum
from .Jacobi import Jacobi
from .Legendre import Legendre, LGComposite


def FunctionSpace(N: int, space: type, bcs=None, domain=None, system=None, name: str | None = None,
                  fun_str: str | None = None, **kw):
    """Return ``space(N)``, its homogeneous-Dirichlet composite, or for
    inhomogeneous boundary values ``DirectSum(composite, BCGeneric)``
    (galerkin/functionspace.py:130-173; the lifting space takes the name and
    function string with a ``_b`` suffix and the composite's name gains ``0``).
    Jacobi-family parameters (alpha, beta, lambda_) go to the composite and
    to the lifting space alike, so the lift is expanded in the composite's
    own orthogonal basis.
    """
    if name is not None:
        kw["name"] = name
    if fun_str is not None:
        kw["fun_str"] = fun_str
    if bcs is None:
        return space(N, domain=domain, system=system, **kw)
    bcs = BoundaryConditions(bcs, domain=domain)
    if not (issubclass(space, (Legendre, Chebyshev, Jacobi))):
        raise TypeError(f"boundary conditions need a polynomial space, got {space.__name__}")
    comp = LGComposite if issubclass(space, Legendre) else CGComposite if issubclass(space, Chebyshev) \
        else Composite
    C = comp(N, space, bcs=bcs.get_homogeneous(), domain=domain, system=system, **kw)
    if bcs.is_homogeneous():
        return C
    kwb = {k: kw[k] for k in ("alpha", "beta", "lambda_") if k in kw}
    if "name" in kw:
        kwb["name"] = kw["name"] + "_b"
    if "fun_str" in kw:
        kwb["fun_str"] = kw["fun_str"] + "_b"
    B = BCGeneric(bcs.num_bcs(), space, bcs=bcs, domain=domain, system=system, num_quad_points=N, **kwb)
    C.name += "0"
    return DirectSum(C, B)
