"""Complex exponential Fourier basis (reference: galerkin/Fourier.py).

Basis E_k(x) = exp(i k x) on the periodic reference domain [0, 2 pi), N even,
wavenumbers in FFT order.  The reference transforms with FFTs; here every
transform is a contraction with a device-built complex basis matrix
(``jfx_fourier_basis``) through the FP64 MFMA kernels, which for the sizes of
spectral-Galerkin bases (N <= a few hundred) keeps the same data path as the
polynomial axes (one kernel per axis, batched, any axis of a tensor product).
Coefficient and physical arrays of a Fourier axis are complex128.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from .. import _lib
from .. import kernels as K
from ..utils.common import Domain
from .orthogonal import MeshKind, OrthogonalSpace, as_device

TWO_PI = 2.0 * math.pi


def _fourier_wavenumbers(N: int, eliminate_highest_freq: bool = False) -> np.ndarray:
    """FFT-ordered wavenumbers (galerkin/Fourier.py:15-20)."""
    k = np.arange(N)
    k = np.where(k < (N + 1) // 2, k, k - N)
    if eliminate_highest_freq and N % 2 == 0:
        k[N // 2] = 0
    return k


class Fourier(OrthogonalSpace):
    """Fourier space of N (even) complex exponentials (galerkin/Fourier.py:23-55)."""

    family = "fourier"
    is_fourier = True

    def __init__(self, N: int, domain=None, system=None, name: str = "Fourier", fun_str: str = "E", **kw):
        if N % 2 != 0:
            raise ValueError("Fourier must use an even number of modes")
        super().__init__(N, domain=domain, system=system, name=name, fun_str=fun_str, **kw)

    @property
    def reference_domain(self) -> Domain:
        return Domain(0.0, TWO_PI)

    def wavenumbers(self, N: int | None = None, eliminate_highest_freq: bool = False) -> np.ndarray:
        return _fourier_wavenumbers(self.N if N is None else int(N), eliminate_highest_freq)

    def norm_squared(self) -> np.ndarray:
        """||E_k||^2 = 2 pi on the reference domain (galerkin/Fourier.py:202-204)."""
        return np.full(self.N, TWO_PI)

    # ------------------------------------------------------------ meshes
    def quad_points_and_weights(self, N: int | None = None):
        """Equispaced points 2 pi j / N and weights 2 pi / N (galerkin/Fourier.py:78-90)."""
        N = self.num_quad_points if N is None else int(N)
        key = ("quad", N)
        if key not in self._ops:
            x = torch.tensor(np.arange(N) * TWO_PI / N, dtype=torch.float64, device=K.device())
            w = torch.tensor(np.full(N, TWO_PI / N), dtype=torch.float64, device=K.device())
            self._ops[key] = (x, w)
        return self._ops[key]

    def mesh(self, kind: str = MeshKind.QUADRATURE, N: int | None = None) -> torch.Tensor:
        """Equispaced periodic mesh of the true domain, endpoint excluded, for
        both mesh kinds (galerkin/Fourier.py:221-235)."""
        N = self.num_quad_points if N is None else int(N)
        a, b = self.domain
        pts = a + (b - a) * np.arange(N) / N
        return torch.tensor(pts, dtype=torch.float64, device=K.device())

    # ------------------------------------------------------------ device matrices
    def _basis(self, Q: int, d: int, adjoint: bool, post: float, x: torch.Tensor | None = None) -> torch.Tensor:
        """(Q, N) (i k df)^d E_k(x_q) * post, or its conjugate transpose."""
        shape = (self.N, Q) if adjoint else (Q, self.N)
        out = torch.empty(shape, dtype=torch.complex128, device=K.device())
        xp = K._p(x) if x is not None else K._p(None)
        _lib.call("jfx_fourier_basis", xp, Q, self.N, d, float(self.domain_factor), 0 if x is not None else 1,
                  1 if adjoint else 0, float(post), K._p(torch.view_as_real(out)), K._s())
        return out

    def backward_operator(self, N: int | None = None, k: int = 0,
                          kind: str = MeshKind.QUADRATURE) -> K.ComplexAxisOperator:
        """(Q, N): the k-th derivative series on the Q-point periodic mesh; Q > N
        is the zero-padded inverse transform (galerkin/Fourier.py:119-149)."""
        Q = self.num_quad_points if N is None else int(N)
        if Q < self.N:
            raise ValueError("Fourier backward transform only supports padding, not truncation")
        key = ("bwd", Q, k)
        if key not in self._ops:
            self._ops[key] = K.ComplexAxisOperator(self._basis(Q, k, False, 1.0))
        return self._ops[key]

    def forward_operator(self, N: int | None = None) -> K.ComplexAxisOperator:
        """(N, Q): coefficients from Q >= N samples, the FFT with norm="forward"
        restricted to the space's wavenumbers (galerkin/Fourier.py:166-180)."""
        Q = self.num_quad_points if N is None else int(N)
        if Q < self.N:
            raise ValueError("Fourier forward transform only supports truncation, not padding")
        key = ("fwd", Q)
        if key not in self._ops:
            self._ops[key] = K.ComplexAxisOperator(self._basis(Q, 0, True, 1.0 / Q))
        return self._ops[key]

    def scalar_product_operator(self, N: int | None = None) -> K.ComplexAxisOperator:
        """(N, Q): (u, E_k) = forward * 2 pi / df (galerkin/Fourier.py:151-164)."""
        Q = self.num_quad_points if N is None else int(N)
        key = ("sp", Q)
        if key not in self._ops:
            self._ops[key] = K.ComplexAxisOperator(
                self._basis(Q, 0, True, TWO_PI / (Q * float(self.domain_factor))))
        return self._ops[key]

    def test_operator(self, i: int, Nq: int) -> K.ComplexAxisOperator:
        """(N, Q): b_k = sum_q w_q df^(i-1) f(x_q) conj(E_k^(i)(x_q)), the linear
        form of a (derivative of a) Fourier test function."""
        key = ("test", Nq, i)
        if key not in self._ops:
            df = float(self.domain_factor)
            self._ops[key] = K.ComplexAxisOperator(self._basis(Nq, i, True, TWO_PI / Nq / df))
        return self._ops[key]

    # ------------------------------------------------------------ transforms
    def forward(self, u, precision: str = "f64"):
        u = as_device(u)
        return K.apply_along(self.forward_operator(u.shape[-1]), u, u.ndim - 1, "f64")

    def scalar_product(self, u, precision: str = "f64"):
        u = as_device(u)
        return K.apply_along(self.scalar_product_operator(u.shape[-1]), u, u.ndim - 1, "f64")

    def evaluate(self, x, c):
        """Series at arbitrary true-domain points (galerkin/Fourier.py:56-76)."""
        shp = tuple(x.shape) if isinstance(x, torch.Tensor) else np.shape(x)
        X = self.map_reference_domain(as_device(x).reshape(-1)).contiguous()
        E = self._basis(X.shape[0], 0, False, 1.0, X)
        c = as_device(c)
        out = K.apply_along(K.ComplexAxisOperator(E), c, c.ndim - 1, "f64")
        return out.reshape(tuple(out.shape[:-1]) + tuple(shp))

    def derivative_coeffs(self, c, k: int = 0):
        """(i k)^k c (galerkin/Fourier.py:206-219), wavenumbers of the reference domain."""
        if k == 0:
            return as_device(c)
        m = self.wavenumbers(eliminate_highest_freq=k % 2 == 1).astype(np.complex128)
        f = torch.tensor((1j * m) ** k, dtype=torch.complex128, device=K.device())
        return K.apply_along(K.ComplexAxisOperator(torch.diag(f)), as_device(c), as_device(c).ndim - 1)

    # ------------------------------------------------------------ matrices
    def _matrices(self, i, trial, q=0):
        """diag((i k)^j (-i k)^i 2 pi) (galerkin/Fourier.py:237-264): real for
        even i + j, complex otherwise."""
        from ..la import DiaMatrix

        u, j = trial
        if not isinstance(u, Fourier) or q != 0:
            return None
        k = self.wavenumbers().astype(np.complex128)
        ku = u.wavenumbers().astype(np.complex128)
        n = min(self.N, u.N)
        d = (1j * ku[:n]) ** j * (-1j * k[:n]) ** i * TWO_PI
        if (i + j) % 2 == 0:
            d = d.real
        A = np.zeros((self.N, u.N), dtype=d.dtype)
        A[np.arange(n), np.arange(n)] = d
        return DiaMatrix.from_dense(A)

    def stencil_matrix(self):
        return None


__all__ = ["Fourier"]
