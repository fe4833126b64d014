"""Galerkin spaces, forms and assembly (reference: src/jaxfun/galerkin)."""

from .arguments import TestFunction, TrialFunction
from .Chebyshev import CGComposite, Chebyshev
from .ChebyshevU import ChebyshevU
from .composite import BCGeneric, BoundaryConditions, Composite, DirectSum, get_bc_basis
from .Fourier import Fourier
from .functionspace import FunctionSpace
from .inner import InnerKind, TensorMatrix, inner, project, project1D
from .Jacobi import Jacobi
from .Legendre import Legendre, LGComposite
from .orthogonal import MeshKind, OrthogonalSpace
from .tensorproductspace import DirectSumTPS, TensorProduct, TensorProductSpace
from .Ultraspherical import Ultraspherical

__all__ = [
    "BCGeneric", "BoundaryConditions", "CGComposite", "ChebyshevU", "DirectSum", "DirectSumTPS", "Fourier", "get_bc_basis", "Chebyshev", "Composite", "FunctionSpace", "InnerKind",
    "Jacobi", "LGComposite", "Legendre", "MeshKind", "OrthogonalSpace", "TensorMatrix", "TensorProduct",
    "TensorProductSpace", "TestFunction", "TrialFunction", "Ultraspherical", "inner", "project", "project1D",
]
