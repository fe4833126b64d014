"""Chebyshev polynomials of the second kind (reference: galerkin/ChebyshevU.py).

U_n = g_n P_n^(1/2,1/2) with g_n = (n + 1) / P_n^(1/2,1/2)(1), so
U_n(1) = n + 1 and (U_m, U_n)_w = pi/2 delta_mn under w = sqrt(1 - x^2).
The reference transforms with a DST-I; here the transforms are the same
Vandermonde contractions as every other polynomial axis.
"""

from __future__ import annotations

import mpmath as mp
import numpy as np
import sympy as sp
import torch

from .. import kernels as K
from .Jacobi import Jacobi


class ChebyshevU(Jacobi):
    """U_n, n < N (galerkin/ChebyshevU.py:15-54)."""

    family = "chebyshevu"

    def __init__(self, N: int, domain=None, system=None, name: str = "ChebyshevU", fun_str: str = "U", **kw):
        kw.pop("alpha", None)
        kw.pop("beta", None)
        super().__init__(N, domain=domain, system=system, name=name, fun_str=fun_str,
                         alpha=sp.S.Half, beta=sp.S.Half, **kw)

    def gn(self, n: int) -> mp.mpf:
        """(n + 1) / P_n^(1/2,1/2)(1) (galerkin/ChebyshevU.py:211-221)."""
        return (n + 1) / mp.binomial(n + mp.mpf(1) / 2, n)

    def norm_squared(self) -> np.ndarray:
        """pi / 2 for every n (galerkin/ChebyshevU.py:159-161)."""
        return np.full(self.N, np.pi / 2)

    def quad_points_and_weights(self, N: int | None = None):
        """Gauss-Chebyshev nodes of the second kind, ascending,
        x_k = cos((k + 1) pi / (N + 1) + pi), w_k = pi / (N + 1) (1 - x_k^2)
        (galerkin/ChebyshevU.py:84-105)."""
        N = self.num_quad_points if N is None else int(N)
        key = ("quad", N)
        if key not in self._ops:
            x = np.cos((np.arange(N) + 1) * np.pi / (N + 1) + np.pi)
            w = np.full(N, np.pi / (N + 1)) * (1 - x ** 2)
            self._ops[key] = (torch.tensor(x, dtype=torch.float64, device=K.device()),
                              torch.tensor(w, dtype=torch.float64, device=K.device()))
        return self._ops[key]

    def _same_basis(self, u) -> bool:
        return isinstance(u, ChebyshevU)


__all__ = ["ChebyshevU"]
