"""Ultraspherical (Gegenbauer) polynomials (reference: galerkin/Ultraspherical.py).

C_n^(lambda) = g_n P_n^(a,a), a = lambda - 1/2, scaled so that
C_n(+-1) = (+-1)^n: g_n = 1 / P_n^(a,a)(1).  lambda = 0 (a = -1/2) is the
Chebyshev basis T_n; the recurrence table's cancelled first entries and the
Jacobi norms cover it without the reference's special cases
(galerkin/Ultraspherical.py:65-101).
"""

from __future__ import annotations

import mpmath as mp
import sympy as sp

from .Jacobi import Jacobi


class Ultraspherical(Jacobi):
    """C_n^(lambda), n < N (galerkin/Ultraspherical.py:12-60)."""

    family = "ultraspherical"

    def __init__(self, N: int, domain=None, system=None, name: str = "Ultraspherical", fun_str: str = "C",
                 lambda_=1, **kw):
        kw.pop("alpha", None)
        kw.pop("beta", None)
        lam = sp.sympify(lambda_)
        super().__init__(N, domain=domain, system=system, name=name, fun_str=fun_str,
                         alpha=lam - sp.S.Half, beta=lam - sp.S.Half, **kw)

    @property
    def lambda_(self):
        return self.alpha + sp.S.Half

    def gn(self, n: int) -> mp.mpf:
        """1 / P_n^(a,a)(1) (galerkin/Ultraspherical.py:74-83)."""
        return 1 / mp.binomial(n + self._params()[0], n)


__all__ = ["Ultraspherical"]
