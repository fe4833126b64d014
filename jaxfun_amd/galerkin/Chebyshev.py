"""Chebyshev (first kind) basis (reference: galerkin/Chebyshev.py)."""

from __future__ import annotations

import numpy as np

from .composite import Composite
from .orthogonal import OrthogonalSpace


def _dense_derivative_matrix(Nt: int, Nu: int, j: int) -> np.ndarray:
    """(T_m, T_n^(j))_w for j in (1, 2) (galerkin/Chebyshev.py:22-41)."""
    r = np.arange(Nt)[:, None]
    c = np.arange(Nu)[None, :]
    off = c - r
    mask = (off >= j) & ((off - j) % 2 == 0)
    val = (np.pi * c + 0 * r).astype(float) if j == 1 else c * (c**2 - r**2) * np.pi / 2
    return np.where(mask, val, 0.0)


class Chebyshev(OrthogonalSpace):
    """Chebyshev polynomials T_n with Gauss-Chebyshev quadrature (galerkin/Chebyshev.py:44-84)."""

    family = "chebyshev"

    def __init__(self, N: int, domain=None, system=None, name: str = "Chebyshev", fun_str: str = "T",
                 **kw):
        super().__init__(N, domain=domain, system=system, name=name, fun_str=fun_str, **kw)

    def norm_squared(self) -> np.ndarray:
        """pi, pi/2, pi/2, ... (galerkin/Chebyshev.py:331-338)."""
        h = np.full(self.N, np.pi / 2)
        h[0] = np.pi
        return h

    def _matrices(self, i, trial, q=0):
        """Closed-form orthogonal matrices (galerkin/Chebyshev.py:399-447)."""
        from ..la import DiaMatrix, Matrix

        u, j = trial
        if not isinstance(u, Chebyshev) or q != 0:
            return None
        if (i, j) not in ((0, 0), (0, 1), (1, 0), (0, 2), (2, 0)):
            return None
        if i == 0 and j == 0:
            m = min(self.N, u.N)
            A = np.zeros((self.N, u.N))
            A[np.arange(m), np.arange(m)] = self.norm_squared()[:m]
            return DiaMatrix.from_dense(A)
        if i in (1, 2) and j == 0:
            m = u._matrices(j, (self, i), q)
            return None if m is None else m.T
        if j >= u.N:
            return None
        return Matrix(_dense_derivative_matrix(self.N, u.N, j))


class CGComposite(Composite):
    """Chebyshev Dirichlet composite basis (galerkin/Chebyshev.py:453-527)."""
