"""FP64 numpy restatement of jaxfun's spectral-Galerkin hot path (the oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg, never by the product package.

Each function cites the reference file:line (paths relative to the
reference's ``src/jaxfun``) whose algorithm it restates.  How each part is
pinned by checks that do not retype the formula (all in ``tests/test_oracle.py``):

* Vandermonde + derivatives: ``numpy.polynomial.{legendre,chebyshev}``
  ``legvander/chebvander`` + ``legder/chebder`` (what the reference's own
  ``tests/galerkin/test_approximations.py:48-87`` compares against).
* General boundary conditions: the reference's closed-form stencils
  (galerkin/composite.py:786-800, the Neumann example of its docstring),
  basis functions meeting every condition under ``numpy.polynomial``
  differentiation, lifting rows meeting theirs, exact recovery of
  polynomial biharmonic and mixed Neumann solutions.
* Jacobi family: ``scipy.special.eval_jacobi`` / ``eval_chebyu`` /
  ``eval_gegenbauer``; derivatives by exact power-basis differentiation;
  norms by Gauss-Jacobi orthogonality; nodes as zeros of P_N with weights
  summing to the weight's integral; the Dirichlet stencil vanishing at both
  ends and its closed forms (Legendre for (0, 0), -(k+1)/(k+3) for U).
* Gauss nodes: ``numpy.polynomial.legendre.leggauss``; exactness of the
  rule on monomials up to degree 2N-1; Chebyshev closed form.
* Galerkin matrices: the literal shenfun golden diagonals of the
  reference's ``tests/galerkin/test_inner.py:104-229``; sympy exact integrals;
  closed forms (Legendre-Dirichlet K = -(4k+6) I).
* Transforms: forward(backward(c)) == c; exactness on
  polynomials; Chebyshev backward vs ``chebval``.
* Banded solve: dense ``numpy.linalg.solve``.
* Per-axis contraction: ``numpy.einsum``.
* FD / Kronecker solves: ``scipy.linalg.solve`` on the assembled
  Kronecker matrix; manufactured solution (1-x^2)(1-y^2).
* IMEX ARS443: exact modal decay exp(-nu*lambda*t) of the linear heat
  equation (third-order convergence), and the tableau's row sums.
* Burgers nonlinear term (``Burgers2D.nonlinear_sp``): exact
  Chebyshev-weighted integrals of -phi_k phi_l u (u_x + u_y) for a
  non-symmetric polynomial u (sympy moments), where the 3/2-padded rule is
  exact; wrong sign, axis, dropped term and unpadded variants fail the check.
* Variable-coefficient operator / PCG / MLP: dense assembled
  operator vs the matrix-free one (and the constant-coefficient stiffness at
  k = 1); load ``VarCoeffPoisson.rhs`` and field ``kfield`` vs
  ``numpy.polynomial.legendre`` evaluation and quadrature (transposed data
  fails); PCG vs direct solve; MLP gradients vs central finite differences.
* Fourier transforms: ``numpy.fft`` is the transform itself (the reference
  calls ``jnp.fft``); pinned further by exactness on trigonometric
  polynomials, the reference's truncation/padding test
  (``tests/galerkin/test_fourier.py``) and the manufactured solutions of
  ``tests/la/test_tpmatrices_solvers.py:50-160`` (per-wavenumber solve vs
  the Kronecker solve and vs cos(2x)(1-y^2)).
* Inhomogeneous Dirichlet lifting: the lifting functions' end values by
  ``numpy.polynomial`` evaluation; the boundary values the reference's own
  tests pin (``tests/galerkin/test_composite.py:66-80``,
  ``test_additional_coverage.py:101-114``); exact recovery of polynomial
  manufactured solutions (1D and 2D, one or two lifted axes, domains other
  than [-1, 1]); the accuracy bound of the reference's
  ``examples/helmholtz2D.py`` for exp(-(x^2+y^2)) on (0, 1)^2.
"""

from __future__ import annotations

import math

import numpy as np
import scipy.linalg as sla

# --------------------------------------------------------------------------
# quadrature
# --------------------------------------------------------------------------


def leggauss(N: int) -> tuple[np.ndarray, np.ndarray]:
    """Gauss-Legendre nodes (ascending) and weights.

    Restates ``Legendre.quad_points_and_weights`` (galerkin/Legendre.py:126-137)
    -> ``utils/fastgl.py:561-567`` (nodes ascending, ``GLPair(N, N-i)``).  The
    reference uses Bogaert's iteration-free formulas; here Newton's method on
    the three-term recurrence is iterated to convergence in FP64.
    """
    if N == 1:
        return np.zeros(1), np.full(1, 2.0)
    k = np.arange(N)
    x = -np.cos(np.pi * (k + 0.75) / (N + 0.5))
    for _ in range(100):
        p0 = np.ones_like(x)
        p1 = x.copy()
        for n in range(1, N):
            p0, p1 = p1, ((2 * n + 1) * x * p1 - n * p0) / (n + 1)
        dp = N * (x * p1 - p0) / (x * x - 1.0)
        dx = p1 / dp
        x = x - dx
        if np.max(np.abs(dx)) < 1e-16:
            break
    p0 = np.ones_like(x)
    p1 = x.copy()
    for n in range(1, N):
        p0, p1 = p1, ((2 * n + 1) * x * p1 - n * p0) / (n + 1)
    omx2 = (1.0 - x) * (1.0 + x)
    dp = N * (p0 - x * p1) / omx2
    w = 2.0 / (omx2 * dp * dp)
    order = np.argsort(x)
    return x[order], w[order]


def chebgauss(N: int) -> tuple[np.ndarray, np.ndarray]:
    """Gauss-Chebyshev nodes ``-cos((2k+1)pi/2N)`` (ascending), weights pi/N.

    Restates ``Chebyshev.quad_points_and_weights`` (galerkin/Chebyshev.py:149-170).
    """
    k = np.arange(N)
    return np.cos(np.pi + (2 * k + 1) * np.pi / (2 * N)), np.full(N, np.pi / N)


def quad(family, N: int) -> tuple[np.ndarray, np.ndarray]:
    fam = jacobi_family(family)
    if fam is not None:
        return fam.quad(N)
    return leggauss(N) if family == "legendre" else chebgauss(N)


# --------------------------------------------------------------------------
# Jacobi family: Jacobi, ChebyshevU, Ultraspherical
# --------------------------------------------------------------------------


class JacobiFamily:
    """Q_n = g_n P_n^(a,b) on [-1, 1] under the weight (1-x)^a (1+x)^b.

    Restates galerkin/Jacobi.py:25-357 (g_n = 1), galerkin/ChebyshevU.py
    (a = b = 1/2, g_n = (n+1) / P_n(1)) and galerkin/Ultraspherical.py
    (a = b = lambda - 1/2, g_n = 1 / P_n(1)).  The polynomials come from the
    textbook recurrence (Abramowitz & Stegun 22.7.1)
        2n(n+a+b)(2n+a+b-2) P_n = (2n+a+b-1)((2n+a+b)(2n+a+b-2) x + a^2-b^2) P_{n-1}
                                  - 2(n+a-1)(n+b-1)(2n+a+b) P_{n-2},
    P_0 = 1, P_1 = (a+1) + (a+b+2)(x-1)/2, and derivatives from
    d^k P_n^(a,b) = Gamma(n+a+b+1+k) / (2^k Gamma(n+a+b+1)) P_{n-k}^(a+k,b+k)
    (galerkin/Jacobi.py:290-304).  Nodes and weights are scipy's
    ``roots_jacobi``, the routine galerkin/Jacobi.py:112-124 calls.
    """

    def __init__(self, alpha: float = 0.0, beta: float = 0.0, kind: str = "jacobi"):
        self.a, self.b, self.kind = float(alpha), float(beta), kind

    def gn(self, n) -> np.ndarray:
        n = np.asarray(n, dtype=np.float64)
        if self.kind == "jacobi":
            return np.ones_like(n)
        lg = np.vectorize(math.lgamma)
        p1 = np.exp(lg(n + self.a + 1) - math.lgamma(self.a + 1) - lg(n + 1))  # P_n^(a,a)(1)
        return (n + 1) / p1 if self.kind == "chebyshevu" else 1.0 / p1

    @staticmethod
    def raw(x, N: int, a: float, b: float) -> np.ndarray:
        """(Q, N) unnormalised P_n^(a,b)(x)."""
        x = np.asarray(x, dtype=np.float64)
        P = np.zeros((x.shape[0], N))
        if N > 0:
            P[:, 0] = 1.0
        if N > 1:
            P[:, 1] = (a + 1) + (a + b + 2) * (x - 1) / 2
        for n in range(2, N):
            c = 2 * n + a + b
            P[:, n] = ((c - 1) * (c * (c - 2) * x + a * a - b * b) * P[:, n - 1]
                       - 2 * (n + a - 1) * (n + b - 1) * c * P[:, n - 2]) / (2 * n * (n + a + b) * (c - 2))
        return P

    def vandermonde(self, x, N: int, k: int = 0) -> np.ndarray:
        x = np.asarray(x, dtype=np.float64)
        V = np.zeros((x.shape[0], N))
        if N > k:
            Pk = self.raw(x, N - k, self.a + k, self.b + k)
            for n in range(k, N):
                f = math.exp(math.lgamma(n + self.a + self.b + 1 + k) - math.lgamma(n + self.a + self.b + 1)
                             - k * math.log(2.0)) if k else 1.0
                V[:, n] = f * Pk[:, n - k]
        return V * self.gn(np.arange(N))[None, :]

    def norm_squared(self, N: int) -> np.ndarray:
        a, b = self.a, self.b
        h = np.zeros(N)
        for n in range(N):
            if n == 0:
                h[n] = 2 ** (a + b + 1) * math.exp(math.lgamma(a + 1) + math.lgamma(b + 1) - math.lgamma(a + b + 2))
            else:
                h[n] = (2 ** (a + b + 1) / (2 * n + a + b + 1)
                        * math.exp(math.lgamma(n + a + 1) + math.lgamma(n + b + 1)
                                   - math.lgamma(n + 1) - math.lgamma(n + a + b + 1)))
        return h * self.gn(np.arange(N)) ** 2

    def quad(self, N: int) -> tuple[np.ndarray, np.ndarray]:
        from scipy.special import roots_jacobi

        x, w = roots_jacobi(N, self.a, self.b)
        return np.asarray(x), np.asarray(w)

    def dirichlet_stencil(self, N: int) -> np.ndarray:
        """phi_k = Q_k + d1 Q_{k+1} + d2 Q_{k+2} with phi_k(+-1) = 0
        (get_stencil_matrix, galerkin/composite.py:765-832); the end values are
        the evaluated polynomials, solved as a 2 x 2 system per k."""
        E = self.vandermonde(np.array([-1.0, 1.0]), N)
        S = np.zeros((N - 2, N))
        for k in range(N - 2):
            d = np.linalg.solve(E[:, k + 1:k + 3], -E[:, k])
            S[k, k], S[k, k + 1], S[k, k + 2] = 1.0, d[0], d[1]
        return S


def jacobi_family(family):
    """A JacobiFamily for ``family`` ("jacobi" = (0, 0), "chebyshevu",
    "ultraspherical" = lambda 1, or an instance), None for Legendre/Chebyshev."""
    if isinstance(family, JacobiFamily):
        return family
    if family == "jacobi":
        return JacobiFamily(0.0, 0.0, "jacobi")
    if family == "chebyshevu":
        return JacobiFamily(0.5, 0.5, "chebyshevu")
    if family == "ultraspherical":
        return JacobiFamily(0.5, 0.5, "ultraspherical")
    return None


# --------------------------------------------------------------------------
# General boundary conditions (D, N, N2, N3, N4)
# --------------------------------------------------------------------------

BC_ORDER = {"D": 0, "R": 0, "N": 1, "W": 1, "N2": 2, "N3": 3, "N4": 4}


def bc_conditions(bcs: dict, df: float = 1.0) -> list[tuple]:
    """(side, derivative order[, Robin alpha]) per condition, in the
    reference's order (left before right, codes sorted;
    galerkin/composite.py:73-90).  A Robin condition R: (alpha, g) means
    u + alpha u' = g in physical units: alpha * df on the reference domain."""
    out = []
    for side, key in (("L", "left"), ("R", "right")):
        for k in sorted(bcs.get(key, {})):
            v = bcs[key][k]
            out.append((side, BC_ORDER[k], float(v[0]) * df) if k in ("R", "W") else (side, BC_ORDER[k]))
    return out


def _cond_row(family, N, cond):
    side, k = cond[0], cond[1]
    row = end_values(family, N, side, k)
    if len(cond) > 2:
        row = row + cond[2] * end_values(family, N, side, k + 1)
    return row


def end_values(family, N: int, side: str, k: int) -> np.ndarray:
    """d^k psi_j at x = -1 ("L") or +1 ("R"), j < N, from the oracle's
    derivative Vandermonde."""
    return vandermonde(family, np.array([-1.0 if side == "L" else 1.0]), N, k)[0]


def general_stencil(family, N: int, conds) -> np.ndarray:
    """(N - nb, N) stencil of phi_n = psi_n + sum_i d_i psi_{n+i} satisfying
    every condition homogeneously: per n the nb x nb system of end values
    (get_stencil_matrix, galerkin/composite.py:765-832)."""
    nb = len(conds)
    E = np.array([_cond_row(family, N, c) for c in conds])
    S = np.zeros((N - nb, N))
    for row in range(N - nb):
        d = np.linalg.solve(E[:, row + 1:row + 1 + nb], -E[:, row])
        S[row, row] = 1.0
        S[row, row + 1:row + 1 + nb] = d
    return S


def general_lift_basis(family, conds, num_derivatives: int) -> np.ndarray:
    """(nb, first + nb) lifting basis (get_bc_basis, galerkin/composite.py:835-895):
    rows of (A^-1)^T for the first offset whose boundary matrix is regular."""
    nb = len(conds)
    for first in range(num_derivatives + 2):
        A = np.array([_cond_row(family, first + nb, c)[first:] for c in conds])
        if abs(np.linalg.det(A)) > 1e-12:
            S = np.zeros((nb, first + nb))
            S[:, first:] = np.linalg.inv(A).T
            return S
    raise ValueError("no lifting basis")


# --------------------------------------------------------------------------
# Vandermonde matrices and derivative coefficients
# --------------------------------------------------------------------------


def vandermonde(family: str, x, N: int, k: int = 0) -> np.ndarray:
    """k-th derivative pseudo-Vandermonde ``V[q, n] = d^k/dx^k psi_n(x_q)``.

    Restates ``Legendre.eval_basis_functions`` (galerkin/Legendre.py:163-183),
    ``Chebyshev.eval_basis_functions`` (galerkin/Chebyshev.py:200-223) and the
    k-fold derivative of ``OrthogonalSpace.evaluate_basis_derivative``
    (galerkin/orthogonal.py:168-178), using the differentiated recurrences
    ``P^(d)_{n+1} = P^(d)_{n-1} + (2n+1) P^(d-1)_n`` (Legendre) and
    ``T^(d)_{n+1} = 2x T^(d)_n + 2d T^(d-1)_n - T^(d)_{n-1}`` (Chebyshev).
    """
    fam = jacobi_family(family)
    if fam is not None:
        return fam.vandermonde(x, N, k)
    x = np.asarray(x, dtype=np.float64)
    Q = x.shape[0]
    V = np.zeros((k + 1, Q, N))
    V[0, :, 0] = 1.0
    if N > 1:
        V[0, :, 1] = x
    for d in range(1, k + 1):
        if N > 1:
            V[d, :, 1] = 1.0 if d == 1 else 0.0
    if family == "legendre":
        for n in range(1, N - 1):
            V[0, :, n + 1] = ((2 * n + 1) * x * V[0, :, n] - n * V[0, :, n - 1]) / (n + 1)
            for d in range(1, k + 1):
                V[d, :, n + 1] = V[d, :, n - 1] + (2 * n + 1) * V[d - 1, :, n]
    elif family == "chebyshev":
        for n in range(1, N - 1):
            V[0, :, n + 1] = 2 * x * V[0, :, n] - V[0, :, n - 1]
            for d in range(1, k + 1):
                V[d, :, n + 1] = (
                    2 * x * V[d, :, n] + 2 * d * V[d - 1, :, n] - V[d, :, n - 1]
                )
    else:
        raise ValueError(family)
    return V[k]


def derivative_coeffs(family: str, c: np.ndarray, k: int = 1) -> np.ndarray:
    """Coefficients of the k-th derivative of a series (same length as c).

    Restates ``Legendre.derivative_coeffs`` (galerkin/Legendre.py:186-216) and
    ``Chebyshev.derivative_coeffs`` (galerkin/Chebyshev.py:281-315): backward
    recurrences ``b_n = (2n+1) c_{n+1} + b_{n+2}/(2n+5)*(2n+1)`` (Legendre,
    written for the raw derivative) and ``b_n = 2(n+1) c_{n+1} + b_{n+2}``,
    ``b_0 /= 2`` (Chebyshev).
    """
    c = np.asarray(c, dtype=np.float64)
    for _ in range(k):
        N = c.shape[0]
        b = np.zeros_like(c)
        if family == "legendre":
            # d/dx sum c_n P_n = sum_m b_m P_m,  b_m = (2m+1) sum_{n>m, n-m odd} c_n
            for m in range(N - 2, -1, -1):
                b[m] = (2 * m + 1) * (c[m + 1] + (b[m + 2] / (2 * m + 5) if m + 2 < N else 0.0))
        else:
            for m in range(N - 2, -1, -1):
                b[m] = 2 * (m + 1) * c[m + 1] + (b[m + 2] if m + 2 < N else 0.0)
            b[0] /= 2.0
        c = b
    return c


def norm_squared(family: str, N: int) -> np.ndarray:
    """Orthogonal-basis norms ``h_n = (psi_n, psi_n)_w``.

    Legendre ``2/(2n+1)`` (galerkin/Jacobi.py:249-251 with alpha=beta=0);
    Chebyshev ``pi, pi/2, ...`` (galerkin/Chebyshev.py:331-338).
    """
    fam = jacobi_family(family)
    if fam is not None:
        return fam.norm_squared(N)
    n = np.arange(N)
    if family == "legendre":
        return 2.0 / (2 * n + 1)
    h = np.full(N, np.pi / 2)
    h[0] = np.pi
    return h


# --------------------------------------------------------------------------
# Dirichlet stencil, closed-form matrices
# --------------------------------------------------------------------------


def dirichlet_stencil(N: int, scaling=None, family=None) -> np.ndarray:
    """Stencil ``S`` (N-2, N) of phi_k = (P_k - P_{k+2}) / s_k.

    Restates ``get_stencil_matrix`` LDRD fast case (galerkin/composite.py:786-789)
    and ``stencil_to_diamatrix`` (galerkin/composite.py:255-269); ``scaling``
    divides each stencil row (galerkin/composite.py:178-179).  A Jacobi-family
    ``family`` takes its general stencil (``JacobiFamily.dirichlet_stencil``).
    """
    fam = jacobi_family(family)
    if fam is not None:
        S = fam.dirichlet_stencil(N)
        return S if scaling is None else S / np.asarray(scaling, dtype=np.float64)[:, None]
    n = N - 2
    S = np.zeros((n, N))
    s = np.ones(n) if scaling is None else np.asarray(scaling, dtype=np.float64)
    S[np.arange(n), np.arange(n)] = 1.0 / s
    S[np.arange(n), np.arange(n) + 2] = -1.0 / s
    return S


def orth_derivative_matrix(family: str, Nt: int, Nu: int, j: int) -> np.ndarray:
    """Dense ``(psi_m, psi_n^(j))_w`` for j in (1, 2) in the orthogonal basis.

    Restates ``_dense_derivative_matrix_data`` (galerkin/Legendre.py:20-33,
    galerkin/Chebyshev.py:22-41).
    """
    r = np.arange(Nt)[:, None]
    c = np.arange(Nu)[None, :]
    off = c - r
    mask = (off >= j) & ((off - j) % 2 == 0)
    if family == "legendre":
        val = np.full((Nt, Nu), 2.0) if j == 1 else (c * (c + 1) - r * (r + 1)).astype(float)
    else:
        val = (np.pi * c + 0 * r).astype(float) if j == 1 else c * (c**2 - r**2) * np.pi / 2
    return np.where(mask, val, 0.0)


def orth_matrix(family: str, N: int, i: int, j: int) -> np.ndarray:
    """Precomputed orthogonal-basis matrix (psi^(i), psi^(j)) (galerkin/orthogonal.py:479-529)."""
    if i == 0 and j == 0:
        return np.diag(norm_squared(family, N))
    if i == 0:
        return orth_derivative_matrix(family, N, N, j)
    if j == 0:
        return orth_derivative_matrix(family, N, N, i).T
    raise ValueError("no precomputed matrix")


def quad_matrix(family: str, N: int, i: int, j: int, Nq: int | None = None,
                weight=None) -> np.ndarray:
    """Quadrature assembly ``z = (P_i^T * w) @ P_j`` in the orthogonal basis.

    Restates the quadrature branch of ``inner_bilinear``
    (galerkin/inner.py:887-900): test derivative i on rows, trial derivative j
    on columns; optional pointwise ``weight`` at the quadrature points.
    """
    Nq = N if Nq is None else Nq
    x, w = quad(family, Nq)
    if weight is not None:
        w = w * weight(x) if callable(weight) else w * weight
    Pi = vandermonde(family, x, N, i)
    Pj = vandermonde(family, x, N, j)
    return (Pi.T * w[None, :]) @ Pj


def galerkin_matrix(family: str, N: int, i: int, j: int, Nq=None, weight=None,
                    scaling=None) -> np.ndarray:
    """Dirichlet-composite Galerkin matrix ``S z S^T`` (test i, trial j).

    Restates ``Composite.apply_stencil_galerkin`` (galerkin/composite.py:290-292)
    applied to the orthogonal matrix (galerkin/inner.py:902-903).
    """
    S = dirichlet_stencil(N, scaling)
    z = quad_matrix(family, N, i, j, Nq, weight) if (Nq or weight is not None) else \
        orth_matrix(family, N, i, j)
    return S @ z @ S.T


def legendre_dirichlet_stiffness(N: int) -> np.ndarray:
    """Closed form ``(phi_k, phi_l'') = -(4k+6) delta_kl`` (galerkin/Legendre.py:439-443)."""
    k = np.arange(N - 2)
    return np.diag(-(4.0 * k + 6.0))


# --------------------------------------------------------------------------
# banded LU without pivoting (DIA solve)
# --------------------------------------------------------------------------


def banded_lu_solve(A: np.ndarray, b: np.ndarray, p: int, q: int) -> np.ndarray:
    """Solve ``A x = b`` with a no-pivot banded LU (lower bw p, upper bw q).

    Restates ``DiaMatrix.lu_factor`` no-pivot kernel (la/diamatrix.py:1936-1973)
    and ``_forward_elimination``/``_backward_substitution``
    (la/diamatrix.py:2079-2195).  ``b`` may carry trailing batch columns.
    """
    n = A.shape[0]
    LU = np.array(A, dtype=np.float64)
    for k in range(n - 1):
        piv = LU[k, k]
        if piv == 0 or not np.isfinite(piv):
            raise ValueError("zero pivot")
        for i in range(k + 1, min(n, k + p + 1)):
            LU[i, k] /= piv
            LU[i, k + 1:min(n, k + q + 1)] -= LU[i, k] * LU[k, k + 1:min(n, k + q + 1)]
    y = np.array(b, dtype=np.float64)
    for i in range(n):
        lo = max(0, i - p)
        if i > lo:
            y[i] -= LU[i, lo:i] @ y[lo:i]
    for i in range(n - 1, -1, -1):
        hi = min(n, i + q + 1)
        if hi > i + 1:
            y[i] -= LU[i, i + 1:hi] @ y[i + 1:hi]
        y[i] /= LU[i, i]
    return y


# --------------------------------------------------------------------------
# transforms and per-axis contraction
# --------------------------------------------------------------------------


def apply_along(A: np.ndarray, X: np.ndarray, axis: int) -> np.ndarray:
    """``Y = A`` applied along ``axis`` of X (la/tpmatrix.py:152-156, la/matrix.py:89-118)."""
    return np.moveaxis(np.tensordot(A, X, axes=([1], [axis])), 0, axis)


class Space1D:
    """A 1D Legendre/Chebyshev space, orthogonal or homogeneous Dirichlet.

    Mirrors ``OrthogonalSpace``/``Composite`` transforms
    (galerkin/orthogonal.py:214-277, galerkin/composite.py:205-349).
    """

    def __init__(self, family: str, N: int, dirichlet: bool = False, scaling=None):
        self.family, self.N, self.dirichlet = family, N, dirichlet
        self.S = dirichlet_stencil(N, scaling, family) if dirichlet else np.eye(N)
        self.dim = self.S.shape[0]
        self.h = norm_squared(family, N)

    def backward_matrix(self, Nq: int | None = None, k: int = 0) -> np.ndarray:
        """(Nq, dim): ``u(x_q) = sum_n c_n phi_n^(k)(x_q)`` (orthogonal.py:214-246)."""
        x, _ = quad(self.family, self.N if Nq is None else Nq)
        return vandermonde(self.family, x, self.N, k) @ self.S.T

    def scalar_product_matrix(self, Nq: int | None = None) -> np.ndarray:
        """(dim, Nq): ``(u, phi_n)_w`` by quadrature (orthogonal.py:264-277, composite.py:345-349)."""
        x, w = quad(self.family, self.N if Nq is None else Nq)
        return self.S @ (vandermonde(self.family, x, self.N).T * w[None, :])

    def mass(self) -> np.ndarray:
        """``S diag(h) S^T`` (composite.py:330-337)."""
        return self.S @ np.diag(self.h) @ self.S.T

    def forward_matrix(self, Nq: int | None = None) -> np.ndarray:
        """(dim, Nq): orthogonal ``L/h`` (orthogonal.py:256-262) or mass solve (composite.py:339-343)."""
        L = self.scalar_product_matrix(Nq)
        if not self.dirichlet:
            return L / self.h[:, None]
        return np.linalg.solve(self.mass(), L)

    def forward(self, u: np.ndarray) -> np.ndarray:
        L = self.scalar_product_matrix(u.shape[0]) @ u
        if not self.dirichlet:
            return L / self.h[:, None] if L.ndim > 1 else L / self.h
        M = self.mass()
        return banded_lu_solve(M, L, 2, 2)


def tp_backward(spaces, c, Nq=None):
    """TensorProductSpace.backward (galerkin/tensorproductspace.py:330-365)."""
    for ax, sp in enumerate(spaces):
        c = apply_along(sp.backward_matrix(None if Nq is None else Nq[ax]), c, ax + c.ndim - len(spaces))
    return c


def tp_forward(spaces, u):
    """TensorProductSpace.forward (galerkin/tensorproductspace.py:395-417)."""
    nb = u.ndim - len(spaces)
    for ax, sp in enumerate(spaces):
        u = apply_along(sp.forward_matrix(u.shape[nb + ax]), u, nb + ax)
    return u


def tp_scalar_product(spaces, u):
    """TensorProductSpace.scalar_product (galerkin/tensorproductspace.py:367-393)."""
    nb = u.ndim - len(spaces)
    for ax, sp in enumerate(spaces):
        u = apply_along(sp.scalar_product_matrix(u.shape[nb + ax]), u, nb + ax)
    return u


# --------------------------------------------------------------------------
# Kronecker and fast-diagonalisation solves
# --------------------------------------------------------------------------


def kron_matrix(terms) -> np.ndarray:
    """Dense ``sum_k s_k A_k^0 kron A_k^1 [kron A_k^2]`` (la/tpmatrix.py:1357-1402)."""
    out = None
    for s, mats in terms:
        K = mats[0]
        for m in mats[1:]:
            K = np.kron(K, m)
        out = s * K if out is None else out + s * K
    return out


def kron_solve(terms, F: np.ndarray) -> np.ndarray:
    """Exact solve via the dense Kronecker matrix (la/tpmatrix.py:1052-1086)."""
    A = kron_matrix(terms)
    return sla.solve(A, F.ravel()).reshape(F.shape)


def fd_factor(A: np.ndarray, B: np.ndarray):
    """Generalised eigendecomposition ``A V = B V diag(lam)`` for one axis.

    Restates ``tpmats_lu_factor``'s per-axis pair (la/tpmatrix.py:1160-1204):
    symmetric-definite pairs use ``eigh`` (V^T B V = I, so W = V^T); a
    non-symmetric pair (Chebyshev-Dirichlet stiffness) uses ``eig`` and
    ``W = (B V)^{-1}`` so the transform stays exact.  Returns (lam, V, W).
    """
    if np.allclose(A, A.T, rtol=1e-13, atol=1e-13 * np.abs(A).max()) and np.allclose(B, B.T):
        lam, V = sla.eigh(A, B)
        return lam, V, V.T.copy()
    lam, V = sla.eig(A, B)
    if np.max(np.abs(lam.imag)) > 1e-8 * np.max(np.abs(lam.real)):
        raise ValueError("complex eigenvalues")
    lam = lam.real
    V = V.real
    order = np.argsort(lam)
    lam, V = lam[order], V[:, order]
    W = np.linalg.inv(B @ V)
    return lam, V, W


def fd_solve(axes, terms_eig, F: np.ndarray) -> np.ndarray:
    """``U = V (W F W^T / D) V^T`` (la/tpmatrix.py:545-587).

    ``axes`` is a list of (V, W) per axis; ``terms_eig`` a list of
    (s_k, [lam_k^axis0, lam_k^axis1, ...]) with 1 standing for the B matrix.
    """
    nd = len(axes)
    Ft = F
    for i, (V, W) in enumerate(axes):
        Ft = apply_along(W, Ft, i + F.ndim - nd)
    D = np.zeros(Ft.shape[F.ndim - nd:])
    for s, lams in terms_eig:
        t = np.ones_like(D)
        for i, lam in enumerate(lams):
            shape = [1] * nd
            shape[i] = -1
            t = t * np.reshape(lam, shape)
        D = D + s * t
    U = Ft / D
    for i, (V, W) in enumerate(axes):
        U = apply_along(V, U, i + F.ndim - nd)
    return U


def helmholtz_partial_fd(n_modes_N: int, alpha: float, F: np.ndarray) -> np.ndarray:
    """Legendre-Dirichlet ``(alpha M (x) M - K (x) M - M (x) K) U = F`` by
    diagonalising axis 0 and banded-solving axis 1 (la/tpmatrix.py:1089-1233
    for the per-axis eigenpair; la/diamatrix.py:394-504 for the banded solve).
    """
    sp = Space1D("legendre", n_modes_N, dirichlet=True)
    M = sp.mass()
    K = legendre_dirichlet_stiffness(n_modes_N)
    lam, V = sla.eigh(K, M)
    G = V.T @ F
    U = np.empty_like(G)
    for i in range(G.shape[0]):
        Ai = (alpha - lam[i]) * M - K
        U[i] = banded_lu_solve(Ai, G[i], 2, 2)
    return V @ U


_PFD_CACHE: dict = {}


def helmholtz_partial_fd_lapack(N: int, alpha: float, F: np.ndarray) -> np.ndarray:
    """Same algorithm as :func:`helmholtz_partial_fd` with compiled kernels:
    per-N eigenpairs cached (as ``TPMatrices.lu_factor`` caches,
    la/tpmatrix.py:386-427), BLAS for the two transforms and LAPACK's banded
    solver (``scipy.linalg.solve_banded``) for every row.  Used as the timed
    CPU baseline; pinned to :func:`helmholtz_partial_fd` in the tests."""
    if N not in _PFD_CACHE:
        sp = Space1D("legendre", N, dirichlet=True)
        M = sp.mass()
        K = legendre_dirichlet_stiffness(N)
        lam, V = sla.eigh(K, M)
        n = N - 2
        ab_M = np.zeros((5, n))
        ab_K = np.zeros((5, n))
        for o in range(-2, 3):
            d = np.diagonal(M, o)
            dk = np.diagonal(K, o)
            if o >= 0:
                ab_M[2 - o, o:] = d
                ab_K[2 - o, o:] = dk
            else:
                ab_M[2 - o, :o] = d
                ab_K[2 - o, :o] = dk
        _PFD_CACHE[N] = (lam, V, ab_M, ab_K)
    lam, V, ab_M, ab_K = _PFD_CACHE[N]
    G = V.T @ F
    U = np.empty_like(G)
    for i in range(G.shape[0]):
        U[i] = sla.solve_banded((2, 2), (alpha - lam[i]) * ab_M - ab_K, G[i], check_finite=False)
    return V @ U


# --------------------------------------------------------------------------
# IMEX ARS443
# --------------------------------------------------------------------------

ARS443_EXPLICIT = np.array([
    [0, 0, 0, 0, 0],
    [1 / 2, 0, 0, 0, 0],
    [11 / 18, 1 / 18, 0, 0, 0],
    [5 / 6, -5 / 6, 1 / 2, 0, 0],
    [1 / 4, 7 / 4, 3 / 4, -7 / 4, 0],
])
ARS443_IMPLICIT = np.array([
    [0, 0, 0, 0, 0],
    [0, 1 / 2, 0, 0, 0],
    [0, 1 / 6, 1 / 2, 0, 0],
    [0, -1 / 2, 1 / 2, 1 / 2, 0],
    [0, 3 / 2, -3 / 2, 1 / 2, 1 / 2],
])
"""ARS(4,4,3) tableau (integrators/tableau.py:318-349)."""


def imex_ars443_step(u, dt, mass_apply, mass_solve, linear_apply, stage_solve,
                     nonlinear_sp):
    """One globally stiffly accurate ARS443 step in coefficient space.

    Restates ``IMEXRungeKutta.stage``/``step`` (integrators/imex_rk.py:55-159):
    rhs_i = M u + dt sum_j (aE_ij N_j + aI_ij L_j); stage 0 (a_00 = 0) is
    ``M^{-1} rhs``; stages 1..4 solve ``(M - dt a_ii L) U_i = rhs_i``; the last
    stage is the new state (globally stiffly accurate).
    """
    aE, aI = ARS443_EXPLICIT, ARS443_IMPLICIT
    mu = mass_apply(u)
    Ns, Ls = [], []
    U = None
    for i in range(5):
        rhs = mu.copy()
        for j in range(i):
            if aE[i, j] != 0:
                rhs = rhs + dt * aE[i, j] * Ns[j]
            if aI[i, j] != 0:
                rhs = rhs + dt * aI[i, j] * Ls[j]
        U = mass_solve(rhs) if aI[i, i] == 0 else stage_solve(rhs)
        if i < 4:
            Ns.append(nonlinear_sp(U))
            Ls.append(linear_apply(U))
    return U


def imex_step(u, dt, aE, aI, bE, bI, mass_apply, mass_solve, linear_apply, stage_solve, nonlinear_sp):
    """One step of any IMEX tableau whose implicit diagonal is 0 or one value
    (``stage_solve`` solves (M - dt a L) U = r for that value).

    Restates ``IMEXRungeKutta.step`` (integrators/imex_rk.py:96-159): stages
    as in ``imex_ars443_step``; a globally stiffly accurate tableau returns
    its last stage, any other takes M^-1 (M u + dt sum_i (bE_i N_i + bI_i L_i)).
    """
    aE, aI, bE, bI = (np.asarray(a, dtype=float) for a in (aE, aI, bE, bI))
    s = len(bI)
    gsa = np.allclose(aE[-1], bE) and np.allclose(aI[-1], bI)
    mu = mass_apply(u)
    Ns, Ls = [], []
    U = None
    for i in range(s):
        rhs = mu.copy()
        for j in range(i):
            rhs = rhs + dt * aE[i, j] * Ns[j] + dt * aI[i, j] * Ls[j]
        U = mass_solve(rhs) if aI[i, i] == 0 else stage_solve(rhs)
        Ns.append(nonlinear_sp(U))
        Ls.append(linear_apply(U))
    if gsa:
        return U
    rhs = mu.copy()
    for i in range(s):
        rhs = rhs + dt * bE[i] * Ns[i] + dt * bI[i] * Ls[i]
    return mass_solve(rhs)


# --------------------------------------------------------------------------
# Burgers pieces (Chebyshev Dirichlet, 3/2 padding)
# --------------------------------------------------------------------------


class Burgers2D:
    """2D viscous Burgers ``u_t + u (u_x + u_y) = nu Lap u`` on a Chebyshev
    Dirichlet tensor space, weak form ``(v, u_t) = nu (v, Lap u) - (v, u(u_x+u_y))``.

    Mass/linear operators follow ``BaseIntegrator`` (integrators/base.py:92-113);
    the nonlinear scalar product is evaluated on a padded mesh and truncated
    (integrators/base.py:238-248, integrators/nonlinear.py:89-133).
    """

    def __init__(self, N: int, nu: float, pad: float = 1.5):
        self.sp = Space1D("chebyshev", N, dirichlet=True)
        self.N, self.nu = N, nu
        self.Np = int(pad * N)
        S = self.sp.S
        self.M = self.sp.mass()
        self.K = S @ orth_derivative_matrix("chebyshev", N, N, 2) @ S.T  # (v, u'')
        self.B0 = self.sp.backward_matrix(self.Np)
        self.B1 = self.sp.backward_matrix(self.Np, k=1)
        self.SP = self.sp.scalar_product_matrix(self.Np)

    def mass_apply(self, U):
        return self.M @ U @ self.M.T

    def linear_apply(self, U):
        return self.nu * (self.K @ U @ self.M.T + self.M @ U @ self.K.T)

    def nonlinear_sp(self, U):
        u = self.B0 @ U @ self.B0.T
        ux = self.B1 @ U @ self.B0.T
        uy = self.B0 @ U @ self.B1.T
        return -(self.SP @ (u * (ux + uy)) @ self.SP.T)

    def solver(self, c):
        """FD solve of ``(M (x) M - c nu (K (x) M + M (x) K)) U = R``."""
        lam, V, W = fd_factor(self.K, self.M)
        terms = [(1.0, [np.ones_like(lam), np.ones_like(lam)]),
                 (-c * self.nu, [lam, np.ones_like(lam)]),
                 (-c * self.nu, [np.ones_like(lam), lam])]
        return lambda R: fd_solve([(V, W), (V, W)], terms, R)


# --------------------------------------------------------------------------
# Variable-coefficient Poisson, PCG, MLP
# --------------------------------------------------------------------------


class VarCoeffPoisson:
    """``-div(k grad u) = f`` on a Legendre-Dirichlet tensor space, k = exp(g).

    Operator ``A U = sum_d (d_d phi, k d_d phi)`` assembled by quadrature
    at the Legendre-Gauss points; the same contraction as ``assemble_multivar``'s
    ``einsum('pi,pk,qj,ql,pq->ikjl')`` (galerkin/inner.py:991-1024) applied
    matrix-free as ``TensorMatrix`` does (la/tensormatrix.py:54-56).
    """

    def __init__(self, N: int, Nq: int | None = None):
        self.sp = Space1D("legendre", N, dirichlet=True)
        self.N = N
        self.Nq = N if Nq is None else Nq
        x, w = leggauss(self.Nq)
        self.x, self.w = x, w
        self.B = self.sp.backward_matrix(self.Nq)        # (Q, n)
        self.D = self.sp.backward_matrix(self.Nq, k=1)   # (Q, n)
        self.Bw = self.B * w[:, None]
        self.Dw = self.D * w[:, None]

    def kfield(self, ghat: np.ndarray) -> np.ndarray:
        """k = exp(g) at the quadrature mesh from low-mode Legendre coefficients."""
        m = ghat.shape[0]
        V = vandermonde("legendre", self.x, m)
        return np.exp(V @ ghat @ V.T)

    def apply(self, kq: np.ndarray, U: np.ndarray) -> np.ndarray:
        ux = self.D @ U @ self.B.T
        uy = self.B @ U @ self.D.T
        return self.Dw.T @ (kq * ux) @ self.Bw + self.Bw.T @ (kq * uy) @ self.Dw

    def dense(self, kq: np.ndarray) -> np.ndarray:
        n = self.B.shape[1]
        A = np.zeros((n * n, n * n))
        E = np.zeros((n, n))
        for j in range(n * n):
            E.flat[j] = 1.0
            A[:, j] = self.apply(kq, E).ravel()
            E.flat[j] = 0.0
        return A

    def rhs(self, fhat: np.ndarray) -> np.ndarray:
        """``b = (phi_k phi_l, f)`` for orthogonal-Legendre forcing coefficients."""
        S = self.sp.S
        h = self.sp.h
        return S @ (h[:, None] * fhat * h[None, :]) @ S.T


def pcg(apply, b, precond=None, x0=None, rtol=1e-10, maxiter=5000):
    """Preconditioned conjugate gradients; stops on ||r|| <= rtol ||b||.

    Returns (x, iterations, final relative residual).
    """
    x = np.zeros_like(b) if x0 is None else x0.copy()
    r = b - apply(x)
    nb = np.linalg.norm(b)
    if np.linalg.norm(r) <= rtol * nb:
        return x, 0, np.linalg.norm(r) / nb
    z = r if precond is None else precond(r)
    p = z.copy()
    rz = np.vdot(r, z)
    for it in range(1, maxiter + 1):
        Ap = apply(p)
        a = rz / np.vdot(p, Ap)
        x = x + a * p
        r = r - a * Ap
        rn = np.linalg.norm(r)
        if rn <= rtol * nb:
            return x, it, rn / nb
        z = r if precond is None else precond(r)
        rz_new = np.vdot(r, z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, maxiter, np.linalg.norm(r) / nb


def mlp_forward(params, X):
    """tanh MLP: ``h_{l+1} = tanh(h_l W_l + b_l)``, linear last layer.

    Returns the output and the cache of layer inputs/pre-activations.
    """
    hs = [X]
    h = X
    L = len(params)
    for li, (W, b) in enumerate(params):
        z = h @ W + b
        h = z if li == L - 1 else np.tanh(z)
        hs.append(h)
    return h, hs


def mlp_backward(params, hs, dY):
    """Gradients of ``sum(dY * Y)`` w.r.t. every (W, b); returns list of (dW, db)."""
    grads = [None] * len(params)
    g = dY
    L = len(params)
    for li in range(L - 1, -1, -1):
        W, b = params[li]
        if li != L - 1:
            g = g * (1.0 - hs[li + 1] ** 2)
        grads[li] = (hs[li].T @ g, g.sum(axis=0))
        g = g @ W.T
    return grads


# --------------------------------------------------------------------------
# Inhomogeneous Dirichlet lifting (DirectSum / DirectSumTPS)
# --------------------------------------------------------------------------


def lift_basis(family: str) -> np.ndarray:
    """(2, 2) lifting basis of ``{left: D, right: D}`` in the orthogonal basis.

    Restates ``get_bc_basis`` (galerkin/composite.py:835-895): with the
    boundary matrix ``A[r, j] = psi_j(x_r)`` (x_r = -1 left, +1 right, j < 2)
    the rows of ``(A^-1)^T`` are the coefficients of the two functions that
    take the value 1 at one end and 0 at the other.
    """
    A = vandermonde(family, np.array([-1.0, 1.0]), 2)
    return np.linalg.inv(A).T


class Axis1D:
    """One axis on a true domain (a, b): ``kind`` "orth" (orthogonal basis),
    "dir" (homogeneous Dirichlet composite with row scaling, galerkin/composite.py:121-408)
    or "lift" (DirectSum of that composite and the lifting basis with numeric
    boundary values ``bvals``, galerkin/composite.py:490-638).
    Coefficient and value arrays run along their LAST axis.
    """

    def __init__(self, family: str, N: int, kind: str = "dir", domain=(-1.0, 1.0), scaling=None,
                 bvals=None):
        self.family, self.N, self.kind = family, N, kind
        self.a, self.b = float(domain[0]), float(domain[1])
        self.df = 2.0 / (self.b - self.a)
        self.S = np.eye(N) if kind == "orth" else dirichlet_stencil(N, scaling)
        self.dim = self.S.shape[0]
        self.h = norm_squared(family, N)
        self.Sb = lift_basis(family)
        self.bvals = None if bvals is None else np.asarray(bvals, dtype=np.float64)

    def ref(self, x):
        return -1.0 + (np.asarray(x, dtype=np.float64) - self.a) * self.df

    def mesh(self, Nq: int | None = None) -> np.ndarray:
        x, _ = quad(self.family, self.N if Nq is None else Nq)
        return self.a + (x + 1.0) / self.df

    def lift(self) -> np.ndarray:
        """Padded orthogonal coefficients of the lift (galerkin/composite.py:588-593)."""
        out = np.zeros(self.N)
        if self.kind == "lift":
            out[:2] = self.bvals @ self.Sb
        return out

    def to_orthogonal(self, c):
        """``c S + lift`` (galerkin/composite.py:276-279, 588-593)."""
        return np.asarray(c) @ self.S + self.lift()

    def from_orthogonal(self, a):
        """``(S S^T)^-1 S (a - lift)`` (galerkin/composite.py:281-284, 595-601)."""
        r = (np.asarray(a) - self.lift()) @ self.S.T
        return np.linalg.solve(self.S @ self.S.T, r.reshape(-1, self.dim).T).T.reshape(r.shape)

    def basis(self, x, k: int = 0) -> np.ndarray:
        """(P, N) k-th derivative of the orthogonal basis at true points x."""
        return vandermonde(self.family, self.ref(x), self.N, k) * self.df ** k

    def values(self, c, x, k: int = 0):
        """Series (k-th derivative) at true points x along the last axis."""
        return self.to_orthogonal(c) @ self.basis(x, k).T

    def orth_forward(self, u):
        """Orthogonal projection ``(V^T w u) / h`` of values at the Gauss mesh
        (galerkin/orthogonal.py:256-277; the domain factor cancels)."""
        u = np.asarray(u, dtype=np.float64)
        x, w = quad(self.family, u.shape[-1])
        return (u * w) @ vandermonde(self.family, x, self.N) / self.h

    def forward(self, u):
        """orth: projection; dir: mass solve (galerkin/composite.py:339-343);
        lift: from_orthogonal(projection) (galerkin/composite.py:628-631)."""
        if self.kind == "orth":
            return self.orth_forward(u)
        if self.kind == "lift":
            return self.from_orthogonal(self.orth_forward(u))
        u = np.asarray(u, dtype=np.float64)
        x, w = quad(self.family, u.shape[-1])
        L = (u * w) @ vandermonde(self.family, x, self.N) @ self.S.T
        M = self.S @ np.diag(self.h) @ self.S.T
        return np.linalg.solve(M, L.reshape(-1, self.dim).T).T.reshape(L.shape)

    def project(self, g):
        """project1D (galerkin/inner.py:1027-1047): forward of g on the Gauss mesh."""
        xm = self.mesh()
        vals = g(xm) if callable(g) else np.full(xm.shape, float(g))
        return self.forward(np.broadcast_to(vals, xm.shape).astype(np.float64))

    def with_bvals(self, bvals) -> "Axis1D":
        out = Axis1D.__new__(Axis1D)
        out.__dict__.update(self.__dict__)
        out.kind = "lift"
        out.bvals = np.asarray(bvals, dtype=np.float64)
        return out

    def trial_basis(self, part: str) -> np.ndarray:
        """Orthogonal coefficients (rows) of the homogeneous ("h") or lifting ("b") functions."""
        if part == "h":
            return self.S
        B = np.zeros((2, self.N))
        B[:, :2] = self.Sb
        return B


class Lifted2D:
    """DirectSumTPS restated for two axes (galerkin/tensorproductspace.py:560-851).

    ``bcs0``/``bcs1`` are the (left, right) boundary functions of the lifted
    axes: callables of the OTHER coordinate, or numbers.  ``bndvals`` maps
    (part0, part1), part in {"h", "b"}, to the fixed coefficient arrays of the
    products with a lifting factor: boundary functions projected onto the
    other axis (itself lifted by the corner values when both axes are lifted),
    and the corner values for the product of both lifting bases.
    """

    def __init__(self, ax0: Axis1D, ax1: Axis1D, bcs0=None, bcs1=None):
        self.axes = (ax0, ax1)
        self.bcs = (bcs0, bcs1)
        lifted = [ax.kind == "lift" for ax in self.axes]
        two = all(lifted)

        def at(g, z):
            return float(g(np.array([z]))[0]) if callable(g) else float(g)

        corners = {}
        if two:
            for a in (0, 1):
                o = self.axes[1 - a]
                corners[a] = [[at(g, o.a), at(g, o.b)] for g in self.bcs[a]]
        self.bndvals = {}
        parts = [("h", "b") if lf else ("h",) for lf in lifted]
        for p0 in parts[0]:
            for p1 in parts[1]:
                if (p0, p1) == ("h", "h"):
                    continue
                if (p0, p1) == ("b", "b"):
                    self.bndvals[(p0, p1)] = np.array(corners[0])
                    continue
                a = 0 if p0 == "b" else 1
                o = self.axes[1 - a]
                rows = []
                for j, g in enumerate(self.bcs[a]):
                    other = o.with_bvals(corners[a][j]) if two else _homogeneous(o)
                    rows.append(other.project(g))
                arr = np.array(rows)
                self.bndvals[(p0, p1)] = arr if a == 0 else arr.T

    def lift(self) -> np.ndarray:
        """Sum of the lifting products in the orthogonal basis (tensorproductspace.py:806-821)."""
        out = np.zeros((self.axes[0].N, self.axes[1].N))
        for (p0, p1), arr in self.bndvals.items():
            out += self.axes[0].trial_basis(p0).T @ arr @ self.axes[1].trial_basis(p1)
        return out

    def to_orthogonal(self, c):
        return self.axes[0].S.T @ c @ self.axes[1].S + self.lift()

    def from_orthogonal(self, a):
        r = a - self.lift()
        S0, S1 = self.axes[0].S, self.axes[1].S
        return np.linalg.solve(S0 @ S0.T, S0 @ r @ S1.T @ np.linalg.inv(S1 @ S1.T))

    def values(self, c, x, y, k=(0, 0)):
        """u^(k) on the grid x (axis 0) times y (axis 1) of true points."""
        return self.axes[0].basis(x, k[0]) @ self.to_orthogonal(c) @ self.axes[1].basis(y, k[1]).T

    def forward(self, u):
        """from_orthogonal of the orthogonal projection (tensorproductspace.py:778-780)."""
        a = self.axes[1].orth_forward(self.axes[0].orth_forward(u.T).T)
        return self.from_orthogonal(a)


def _homogeneous(ax: Axis1D) -> Axis1D:
    out = Axis1D.__new__(Axis1D)
    out.__dict__.update(ax.__dict__)
    if ax.kind == "lift":
        out.kind = "dir"
    return out


def _axis_matrices(ax: Axis1D, part: str, Nq: int | None = None):
    """Mass and (v, u'') matrices on the true domain between the homogeneous
    test functions and the trial functions of ``part``, by Gauss quadrature
    (the quadrature branch of inner_bilinear, galerkin/inner.py:887-909)."""
    x, w = quad(ax.family, ax.N if Nq is None else Nq)
    V = vandermonde(ax.family, x, ax.N)
    V2 = vandermonde(ax.family, x, ax.N, 2)
    T = ax.S @ (V.T * w[None, :])
    B = ax.trial_basis(part)
    M = T @ V @ B.T / ax.df
    K = T @ V2 @ B.T * ax.df
    return M, K


def lifted_helmholtz_1d(ax: Axis1D, alpha: float, f, Nq: int | None = None):
    """Solve ``(v, u'' + alpha u) = (v, f)`` for u = u_h + lift on one lifted
    axis: the lifting term moves to the right-hand side with the sign of a
    linear form (galerkin/inner.py:362-386).  Returns homogeneous coefficients."""
    x, w = quad(ax.family, ax.N if Nq is None else Nq)
    fq = f(ax.a + (x + 1.0) / ax.df) if callable(f) else np.full(x.shape, float(f))
    b = ax.S @ (vandermonde(ax.family, x, ax.N).T @ (w * fq)) / ax.df
    Mh, Kh = _axis_matrices(ax, "h", Nq)
    A = Kh + alpha * Mh
    if ax.kind == "lift":
        Mb, Kb = _axis_matrices(ax, "b", Nq)
        b = b - (Kb + alpha * Mb) @ ax.bvals
    return np.linalg.solve(A, b)


def lifted_helmholtz_2d(L: Lifted2D, alpha: float, fgrid: np.ndarray, Nq=None):
    """Solve ``(v, Lap u + alpha u) = (v, f)`` for u = u_h + lift on a
    DirectSumTPS (galerkin/inner.py:493-562: every product with a lifting
    factor is applied to its ``bndvals`` and subtracted from b).  ``fgrid``
    holds f on the Gauss mesh of both axes.  Returns (coefficients, A, b)."""
    ax0, ax1 = L.axes
    Nq = (ax0.N, ax1.N) if Nq is None else Nq
    x0, w0 = quad(ax0.family, Nq[0])
    x1, w1 = quad(ax1.family, Nq[1])
    T0 = ax0.S @ (vandermonde(ax0.family, x0, ax0.N).T * w0[None, :]) / ax0.df
    T1 = ax1.S @ (vandermonde(ax1.family, x1, ax1.N).T * w1[None, :]) / ax1.df
    b = T0 @ fgrid @ T1.T
    M0, K0 = _axis_matrices(ax0, "h", Nq[0])
    M1, K1 = _axis_matrices(ax1, "h", Nq[1])
    A = np.kron(K0, M1) + np.kron(M0, K1) + alpha * np.kron(M0, M1)
    for (p0, p1), arr in L.bndvals.items():
        m0, k0 = _axis_matrices(ax0, p0, Nq[0])
        m1, k1 = _axis_matrices(ax1, p1, Nq[1])
        b = b - (k0 @ arr @ m1.T + m0 @ arr @ k1.T + alpha * m0 @ arr @ m1.T)
    u = np.linalg.solve(A, b.ravel()).reshape(ax0.dim, ax1.dim)
    return u, A, b


# --------------------------------------------------------------------------
# Fourier basis and the per-wavenumber solver
# --------------------------------------------------------------------------


def fourier_wavenumbers(N: int, eliminate_highest_freq: bool = False) -> np.ndarray:
    """FFT-ordered wavenumbers (galerkin/Fourier.py:15-20) = numpy.fft.fftfreq(N) * N."""
    k = np.rint(np.fft.fftfreq(N) * N).astype(int)
    if eliminate_highest_freq and N % 2 == 0:
        k[N // 2] = 0
    return k


class Fourier1D:
    """Fourier transforms restated with numpy.fft (galerkin/Fourier.py:119-235).

    Coefficients along the LAST axis; the reference domain is [0, 2 pi), the
    true domain (a, b) maps onto it with df = 2 pi / (b - a).
    """

    def __init__(self, N: int, domain=(0.0, 2 * np.pi)):
        self.N = N
        self.a, self.b = float(domain[0]), float(domain[1])
        self.df = 2 * np.pi / (self.b - self.a)

    def k(self, eliminate: bool = False) -> np.ndarray:
        return fourier_wavenumbers(self.N, eliminate)

    def mesh(self, Nq: int | None = None) -> np.ndarray:
        Nq = self.N if Nq is None else Nq
        return self.a + (self.b - self.a) * np.arange(Nq) / Nq

    def backward(self, c, Nq: int | None = None, d: int = 0):
        """ifft(norm="forward") with zero padding in the middle (Fourier.py:127-149);
        derivative d by (i k df)^d on the coefficients (Fourier.py:206-219)."""
        c = np.asarray(c, dtype=np.complex128)
        if d:
            c = c * (1j * self.k(d % 2 == 1) * self.df) ** d
        Nq = self.N if Nq is None else Nq
        if Nq > self.N:
            h = self.N // 2
            pad = np.zeros(c.shape[:-1] + (Nq - self.N,), dtype=np.complex128)
            c = np.concatenate([c[..., :h], pad, c[..., h:]], axis=-1)
        return np.fft.ifft(c, axis=-1, norm="forward")

    def forward(self, u):
        """fft(norm="forward"), truncated to the space's wavenumbers (Fourier.py:166-180)."""
        u = np.asarray(u)
        out = np.fft.fft(u, axis=-1, norm="forward")
        if u.shape[-1] > self.N:
            out = out[..., self.k() % u.shape[-1]]
        return out

    def scalar_product(self, u):
        """forward * 2 pi / df (galerkin/Fourier.py:151-164)."""
        return self.forward(u) * 2 * np.pi / self.df

    def evaluate(self, x, c):
        X = (np.asarray(x, dtype=np.float64) - self.a) * self.df
        return np.exp(1j * np.multiply.outer(X, self.k())) @ np.asarray(c, dtype=np.complex128)

    def matrix(self, i: int, j: int) -> np.ndarray:
        """diag((i k)^j (-i k)^i 2 pi) df^(i+j-1) (Fourier.py:237-264, orthogonal.py:479-529)."""
        k = self.k().astype(np.complex128)
        return np.diag((1j * k) ** j * (-1j * k) ** i * 2 * np.pi * self.df ** (i + j - 1))


def wavenumber_solve(terms, F: np.ndarray, poly_axis: int) -> np.ndarray:
    """Solve ``sum_t s_t A_t^(0) (x) A_t^(1) u = F`` where every factor on the
    other axis is diagonal: one dense solve of the polynomial-axis matrix
    ``B_k = sum_t s_t d_t[k] P_t`` per wavenumber k (la/tpmatrix.py:590-1011,
    1236-1354, with numpy.linalg.solve in place of the banded LU)."""
    other = 1 - poly_axis
    n = F.shape[poly_axis]
    U = np.zeros(F.shape, dtype=np.result_type(F, np.complex128))
    for k in range(F.shape[other]):
        B = np.zeros((n, n), dtype=np.complex128)
        for s, mats in terms:
            B = B + s * np.diag(mats[other])[k] * mats[poly_axis]
        if poly_axis == 1:
            U[k, :] = np.linalg.solve(B, F[k, :])
        else:
            U[:, k] = np.linalg.solve(B, F[:, k])
    return U


def ars443_stage_coeffs():
    """Distinct implicit diagonal coefficients of ARS443 (integrators/imex_rk.py:39-48)."""
    return sorted({float(ARS443_IMPLICIT[i, i]) for i in range(5) if ARS443_IMPLICIT[i, i] != 0})


def rel_l2(a, b) -> float:
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = np.asarray(a, dtype=dt)
    b = np.asarray(b, dtype=dt)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


__all__ = [name for name in dir() if not name.startswith("_")] + ["math"]
