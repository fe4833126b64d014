"""Galerkin spaces, forms and assembly (reference: src/jaxfun/galerkin)."""

from .arguments import TestFunction, TrialFunction
from .Chebyshev import CGComposite, Chebyshev
from .composite import BoundaryConditions, Composite
from .functionspace import FunctionSpace
from .inner import InnerKind, TensorMatrix, inner, project
from .Legendre import Legendre, LGComposite
from .orthogonal import MeshKind, OrthogonalSpace
from .tensorproductspace import TensorProduct, TensorProductSpace

__all__ = [
    "BoundaryConditions", "CGComposite", "Chebyshev", "Composite", "FunctionSpace", "InnerKind",
    "LGComposite", "Legendre", "MeshKind", "OrthogonalSpace", "TensorMatrix", "TensorProduct",
    "TensorProductSpace", "TestFunction", "TrialFunction", "inner", "project",
]
