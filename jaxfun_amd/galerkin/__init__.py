"""Galerkin spaces, forms and assembly (reference: src/jaxfun/galerkin)."""

from .arguments import TestFunction, TrialFunction
from .Chebyshev import CGComposite, Chebyshev
from .composite import BCGeneric, BoundaryConditions, Composite, DirectSum, get_bc_basis
from .Fourier import Fourier
from .functionspace import FunctionSpace
from .inner import InnerKind, TensorMatrix, inner, project
from .Legendre import Legendre, LGComposite
from .orthogonal import MeshKind, OrthogonalSpace
from .tensorproductspace import DirectSumTPS, TensorProduct, TensorProductSpace

__all__ = [
    "BCGeneric", "BoundaryConditions", "CGComposite", "DirectSum", "DirectSumTPS", "Fourier", "get_bc_basis", "Chebyshev", "Composite", "FunctionSpace", "InnerKind",
    "LGComposite", "Legendre", "MeshKind", "OrthogonalSpace", "TensorMatrix", "TensorProduct",
    "TensorProductSpace", "TestFunction", "TrialFunction", "inner", "project",
]
