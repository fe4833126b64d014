"""Linear algebra on the device (reference: la/)."""

from .matrix import BaseMatrix, DiaMatrix, Matrix, SolverNotApplicable, diags
from .tpmatrix import (
    TPLUFactors,
    TPMatrices,
    TPMatricesDenseLUFactors,
    TPMatricesLUFactors,
    TPMatricesWavenumberSolver,
    TPMatrix,
    tpmats_dense_lu_factor,
    tpmats_lu_factor,
    tpmats_to_kron,
    tpmats_wavenumber_factor,
)

__all__ = [
    "BaseMatrix",
    "DiaMatrix",
    "Matrix",
    "SolverNotApplicable",
    "diags",
    "TPLUFactors",
    "TPMatrices",
    "TPMatricesDenseLUFactors",
    "TPMatricesLUFactors",
    "TPMatricesWavenumberSolver",
    "TPMatrix",
    "tpmats_dense_lu_factor",
    "tpmats_lu_factor",
    "tpmats_to_kron",
    "tpmats_wavenumber_factor",
]
