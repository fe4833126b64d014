"""Legendre basis (reference: galerkin/Legendre.py)."""

from __future__ import annotations

import numpy as np

from .composite import Composite
from .orthogonal import OrthogonalSpace


def _dense_derivative_matrix(Nt: int, Nu: int, j: int) -> np.ndarray:
    """(P_m, P_n^(j)) for j in (1, 2) (galerkin/Legendre.py:20-33)."""
    r = np.arange(Nt)[:, None]
    c = np.arange(Nu)[None, :]
    off = c - r
    mask = (off >= j) & ((off - j) % 2 == 0)
    val = np.full((Nt, Nu), 2.0) if j == 1 else (c * (c + 1) - r * (r + 1)).astype(float)
    return np.where(mask, val, 0.0)


class Legendre(OrthogonalSpace):
    """Legendre polynomials P_n on [-1, 1] (galerkin/Legendre.py:42-67)."""

    family = "legendre"

    def __init__(self, N: int, domain=None, system=None, name: str = "Legendre", fun_str: str = "P",
                 **kw):
        super().__init__(N, domain=domain, system=system, name=name, fun_str=fun_str, **kw)

    def norm_squared(self) -> np.ndarray:
        """h_n = 2 / (2n + 1) (galerkin/Jacobi.py:249-251, alpha = beta = 0)."""
        return 2.0 / (2 * np.arange(self.N) + 1)

    def _matrices(self, i, trial, q=0):
        """Closed-form orthogonal matrices (galerkin/Legendre.py:220-265)."""
        from ..la import DiaMatrix, Matrix

        u, j = trial
        if not isinstance(u, Legendre) or q != 0:
            return None
        if i == 0 and j == 0:
            m = min(self.N, u.N)
            A = np.zeros((self.N, u.N))
            A[np.arange(m), np.arange(m)] = self.norm_squared()[:m]
            return DiaMatrix.from_dense(A)
        if i == 0 and j in (1, 2):
            if j >= u.N:
                return None
            return Matrix(_dense_derivative_matrix(self.N, u.N, j))
        if j == 0 and i in (1, 2):
            m = u._matrices(j, (self, i), q)
            return None if m is None else m.T
        return None


class LGComposite(Composite):
    """Legendre Dirichlet composite basis (galerkin/Legendre.py:271-495)."""

    def _matrices(self, i, trial, q=0):
        from ..la import DiaMatrix

        u, j = trial
        if not isinstance(u, Composite) or u.bcs != self.bcs or self.num_dofs != u.num_dofs:
            return None
        if self.bcs != {"left": {"D": 0}, "right": {"D": 0}} or q != 0:
            return None
        M = self.num_dofs
        k = np.arange(M, dtype=float)
        s_test = self.scaling_values()
        s_trial = u.scaling_values()
        if i == 0 and j == 2:
            # (phi_k, phi_l'') = -(4k + 6) delta_kl (galerkin/Legendre.py:439-443)
            return DiaMatrix.from_dense(np.diag(-(4 * k + 6) / (s_test * s_trial)))
        if i == 2 and j == 0:
            return self._matrices(0, (u, 2), q).T
        if i == 1 and j == 1:
            return self._matrices(0, (u, 2), q) * -1.0
        if i == 0 and j == 1:
            A = np.zeros((M, M))
            A[np.arange(1, M), np.arange(M - 1)] = -2 / (s_test[1:] * s_trial[:-1])
            A[np.arange(M - 1), np.arange(1, M)] = 2 / (s_test[:-1] * s_trial[1:])
            return DiaMatrix.from_dense(A)
        if i == 1 and j == 0:
            return self._matrices(0, (u, 1), q).T
        return None
