"""Jacobi polynomial spaces (reference: galerkin/Jacobi.py).

Basis Q_n = g_n P_n^(alpha,beta), orthogonal under the weight
(1 - x)^alpha (1 + x)^beta on [-1, 1]; g_n = 1 here, other scalings in the
subclasses ``ChebyshevU`` and ``Ultraspherical``.  The device sees a basis
only through its three-term recurrence: the host tabulates the recurrence
coefficients once per space in 40-digit arithmetic and
``jfx_vandermonde_rec_f64`` evaluates the (derivative) Vandermonde matrix from
them; everything downstream (transforms, ``inner``, lifting, solves) is the
same device path as for Legendre and Chebyshev.
"""

from __future__ import annotations

import mpmath as mp
import numpy as np
import sympy as sp
import torch
from scipy.special import roots_jacobi

from .. import kernels as K
from .orthogonal import OrthogonalSpace

_DPS = 40


def _mpf(v) -> mp.mpf:
    return mp.mpf(str(sp.sympify(v).evalf(_DPS)))


def jacobi_end_derivative(a, b, g, j: int, side: str, k: int = 0) -> mp.mpf:
    """d^k/dx^k of g P_j^(a,b) at x = +1 ("R") or x = -1 ("L"), from
    d^k P_j = psi P_{j-k}^(a+k,b+k), psi = (j+a+b+1)_k / 2^k, and
    P_n^(a,b)(1) = C(n+a, n), P_n^(a,b)(-1) = (-1)^n C(n+b, n)
    (galerkin/Jacobi.py:257-304)."""
    if j < k:
        return mp.mpf(0)
    psi = mp.rf(j + a + b + 1, k) / mp.mpf(2) ** k
    if side == "R":
        return g * psi * mp.binomial(j + a, j - k)
    return g * (-1) ** (j + k) * psi * mp.binomial(j + b, j - k)


class Jacobi(OrthogonalSpace):
    """P_n^(alpha,beta), n < N (galerkin/Jacobi.py:25-63)."""

    family = "jacobi"

    def __init__(self, N: int, domain=None, system=None, name: str = "Jacobi", fun_str: str = "J",
                 alpha=0, beta=0, **kw):
        super().__init__(N, domain=domain, system=system, name=name, fun_str=fun_str, **kw)
        self.alpha = sp.sympify(alpha)
        self.beta = sp.sympify(beta)
        if not (float(self.alpha) > -1 and float(self.beta) > -1):
            raise ValueError(f"Jacobi parameters must exceed -1 (got alpha={alpha}, beta={beta})")

    # ------------------------------------------------------------ exact scalars
    def gn(self, n: int) -> mp.mpf:
        """Scaling g_n of Q_n = g_n P_n^(alpha,beta) (galerkin/Jacobi.py:344-357)."""
        return mp.mpf(1)

    def _params(self) -> tuple[mp.mpf, mp.mpf]:
        return _mpf(self.alpha), _mpf(self.beta)

    def _same_basis(self, u) -> bool:
        return type(u) is type(self) and u.alpha == self.alpha and u.beta == self.beta

    def recurrence_table(self, N: int | None = None) -> np.ndarray:
        """(3, N) [r0 | r1 | r2] with Q_n = r0[n] ((x - r1[n]) Q_{n-1} - r2[n] Q_{n-2}).

        From x Q_m = A(m+1, m) Q_{m+1} + A(m, m) Q_m + A(m-1, m) Q_{m-1}, the
        normalised recursion matrix of galerkin/Jacobi.py:359-374, 421-447:
        r0[n] = 1 / A(n, n-1), r1[n] = A(n-1, n-1), r2[n] = A(n-2, n-1).
        The m = 0 entries are taken in the cancelled forms 2 / (a + b + 2) and
        (b - a) / (a + b + 2), which hold also where a + b = -1 or 0.
        """
        N = self.N if N is None else int(N)
        out = np.zeros((3, N))
        with mp.workdps(_DPS):
            a, b = self._params()
            s = a + b
            g = [self.gn(m) for m in range(N + 1)]
            for n in range(1, N):
                m = n - 1
                if m == 0:
                    A1, AA = 2 / (s + 2), (b - a) / (s + 2)
                else:
                    A1 = 2 * (m + 1) * (m + s + 1) / ((2 * m + s + 2) * (2 * m + s + 1))
                    AA = -(a * a - b * b) / ((2 * m + s + 2) * (2 * m + s))
                out[0, n] = float(g[m + 1] / (g[m] * A1))
                out[1, n] = float(AA)
                if n >= 2:
                    k = n - 2
                    A2 = 2 * (k + 1 + a) * (k + 1 + b) / ((2 * k + s + 3) * (2 * k + s + 2))
                    out[2, n] = float(g[k + 1] / g[k] * A2)
        return out

    def _rec_dev(self) -> torch.Tensor:
        if "rec" not in self._ops:
            self._ops["rec"] = torch.tensor(self.recurrence_table(), dtype=torch.float64,
                                            device=K.device()).contiguous()
        return self._ops["rec"]

    def _vander(self, X: torch.Tensor, k: int = 0, scale: float = 1.0) -> torch.Tensor:
        return K.vandermonde_rec(self._rec_dev(), X.contiguous(), self.N, k, scale)

    def norm_squared(self) -> np.ndarray:
        """h_n = g_n^2 (Q_n, Q_n)_w (galerkin/Jacobi.py:249-251, 306-342)."""
        with mp.workdps(_DPS):
            a, b = self._params()
            s = a + b
            h = []
            for n in range(self.N):
                if n == 0:
                    h0 = 2 ** (s + 1) * mp.gamma(a + 1) * mp.gamma(b + 1) / mp.gamma(s + 2)
                else:
                    h0 = (2 ** (s + 1) / (2 * n + s + 1) * mp.gamma(n + a + 1) * mp.gamma(n + b + 1)
                          / (mp.gamma(n + 1) * mp.gamma(n + s + 1)))
                h.append(float(self.gn(n) ** 2 * h0))
        return np.array(h)

    def bnd_value(self, j: int, side: str, k: int = 0, exact: bool = False):
        """d^k Q_j at x = -1 ("L") or x = +1 ("R") on the reference domain
        (galerkin/Jacobi.py:257-288); an mpmath number when ``exact``."""
        with mp.workdps(_DPS):
            a, b = self._params()
            v = jacobi_end_derivative(a, b, self.gn(j), j, side, k)
            return +v if exact else float(v)

    # ------------------------------------------------------------ quadrature
    def quad_points_and_weights(self, N: int | None = None):
        """Gauss-Jacobi nodes (ascending) and weights (galerkin/Jacobi.py:112-124):
        set-up on the host by scipy's ``roots_jacobi``, the routine the
        reference calls, then kept on the device."""
        N = self.num_quad_points if N is None else int(N)
        key = ("quad", N)
        if key not in self._ops:
            x, w = roots_jacobi(N, float(self.alpha), float(self.beta))
            self._ops[key] = (torch.tensor(x, dtype=torch.float64, device=K.device()),
                              torch.tensor(w, dtype=torch.float64, device=K.device()))
        return self._ops[key]

    # ------------------------------------------------------------ matrices
    def _matrices(self, i, trial, q=0):
        """Weighted mass diag(h) for (i, j) = (0, 0), else None (quadrature)
        (galerkin/Jacobi.py:549-571)."""
        from ..la import DiaMatrix

        u, j = trial
        if q != 0 or i != 0 or j != 0 or not self._same_basis(u):
            return None
        h = self.norm_squared()
        n = min(self.N, u.N)
        A = np.zeros((self.N, u.N))
        A[np.arange(n), np.arange(n)] = h[:n]
        return DiaMatrix.from_dense(A)

    def __repr__(self) -> str:
        return f"{type(self).__name__}(N={self.N}, alpha={self.alpha}, beta={self.beta}, domain={tuple(self.domain)})"


__all__ = ["Jacobi"]
