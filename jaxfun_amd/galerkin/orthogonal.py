"""1D orthogonal polynomial spaces on the GPU (reference: galerkin/orthogonal.py).

A space owns device-resident constant operators, built once by HIP kernels
and cached (the analogue of the reference's jit caches):

* ``backward_operator(Nq, k)``   (Nq x dim):  u^(k)(x_q) = sum_n c_n phi_n^(k)(x_q)
* ``scalar_product_operator(Nq)`` (dim x Nq):  (u, phi_n)_w by quadrature
* ``forward_operator(Nq)``        (dim x Nq):  projection (mass solve folded in)

Each is an :class:`~jaxfun_amd.kernels.AxisOperator`, applied along one axis
of a batch by the MFMA contraction kernels.
"""

from __future__ import annotations

import numpy as np
import sympy as sp
import torch

from .. import kernels as K
from ..coordinates import CartCoordSys, x as _x
from ..utils.common import Domain, lambdify


class MeshKind:
    QUADRATURE = "quadrature"
    UNIFORM = "uniform"


class OrthogonalSpace:
    """Abstract 1D orthogonal space (galerkin/orthogonal.py:36-97)."""

    is_orthogonal = True
    family: str = ""

    def __init__(self, N: int, *, domain=None, system=None, name: str = "OrthogonalSpace",
                 fun_str: str = "psi", **kw):
        self.N = int(N)
        self._num_quad_points = int(kw.get("num_quad_points", N))
        if domain is None:
            domain = self.reference_domain
        self._domain = Domain(float(domain[0]), float(domain[1]))
        self.system = system if system is not None else CartCoordSys("N", (_x,))
        self.name = name
        self.fun_str = fun_str
        self.bcs = None
        self.orthogonal = self
        self.stencil = {0: 1}
        self._ops: dict = {}

    # ------------------------------------------------------------ geometry
    @property
    def reference_domain(self) -> Domain:
        return Domain(-1.0, 1.0)

    @property
    def domain(self) -> Domain:
        return self._domain

    @property
    def domain_factor(self) -> float:
        """Reference length / true length (galerkin/orthogonal.py:316-329)."""
        a, b = self.domain
        c, d = self.reference_domain
        L, R = b - a, d - c
        return R / L if abs(L - R) > 1e-12 else 1.0

    def map_reference_domain(self, x):
        if self.domain == self.reference_domain:
            return x
        return self.reference_domain.lower + (x - self.domain.lower) * self.domain_factor

    def map_true_domain(self, X):
        if self.domain == self.reference_domain:
            return X
        return self.domain.lower + (X - self.reference_domain.lower) / self.domain_factor

    def map_expr_reference_domain(self, u):
        """u(x) rewritten with the reference coordinate X = c + (x - a) df
        (galerkin/orthogonal.py:357-373)."""
        u = sp.sympify(u)
        if not u.free_symbols:
            return u
        x = self.system.base_scalars()[0]
        a, c = self.domain.lower, self.reference_domain.lower
        return u.xreplace({x: c + (x - a) * float(self.domain_factor)})

    def map_expr_true_domain(self, u):
        """u(X) rewritten with the true coordinate x = a + (X - c) / df
        (galerkin/orthogonal.py:376-392)."""
        u = sp.sympify(u)
        if not u.free_symbols:
            return u
        x = self.system.base_scalars()[0]
        a, c = self.domain.lower, self.reference_domain.lower
        return u.xreplace({x: a + (x - c) / float(self.domain_factor)})

    @property
    def num_quad_points(self) -> int:
        return self._num_quad_points

    @property
    def dim(self) -> int:
        return self.N

    @property
    def num_dofs(self) -> int:
        return self.dim

    @property
    def dims(self) -> int:
        return 1

    @property
    def shape(self) -> tuple[int, ...]:
        return (self.num_quad_points,)

    def __len__(self) -> int:
        return 1

    def get_orthogonal(self):
        return self

    # ------------------------------------------------------------ basis data
    def norm_squared(self) -> np.ndarray:
        raise NotImplementedError

    def stencil_matrix(self) -> np.ndarray | None:
        """(dim, N) stencil; None for the orthogonal basis itself."""
        return None

    def quad_points_and_weights(self, N: int | None = None) -> tuple[torch.Tensor, torch.Tensor]:
        """Reference-domain Gauss nodes (ascending) and weights, on the device."""
        N = self.num_quad_points if N is None else int(N)
        if self.orthogonal is not self:
            return self.orthogonal.quad_points_and_weights(N)
        key = ("quad", N)
        if key not in self._ops:
            self._ops[key] = K.gauss_nodes(self.family, N)
        return self._ops[key]

    def mesh(self, kind: str = MeshKind.QUADRATURE, N: int | None = None) -> torch.Tensor:
        """Sampling mesh in the true domain (galerkin/orthogonal.py:428-444)."""
        N = self.num_quad_points if N is None else int(N)
        if kind == MeshKind.QUADRATURE:
            return self.map_true_domain(self.quad_points_and_weights(N)[0])
        a, b = self.domain
        return torch.linspace(a, b, N, dtype=torch.float64, device=K.device())

    def _points(self, kind: str, N: int | None) -> torch.Tensor:
        N = self.num_quad_points if N is None else int(N)
        if kind == MeshKind.QUADRATURE:
            return self.quad_points_and_weights(N)[0]
        return self.map_reference_domain(self.mesh(kind, N)).contiguous()

    def _vander(self, X: torch.Tensor, k: int = 0, scale: float = 1.0) -> torch.Tensor:
        """(Q, N) scale * d^k/dx^k psi_n(X) of this orthogonal basis on the device."""
        return K.vandermonde(self.family, X.contiguous(), self.N, k, scale)

    def vandermonde(self, X: torch.Tensor, k: int = 0) -> torch.Tensor:
        """Orthogonal-basis (derivative) Vandermonde at reference points (orthogonal.py:131-178)."""
        return self._vander(X, k)

    def bnd_value(self, j: int, side: str, k: int = 0, exact: bool = False):
        """d^k psi_j at the reference-domain end ("L": x = -1, "R": x = +1)
        (the ``bnd_values`` of galerkin/Jacobi.py:257-288 and
        galerkin/Chebyshev.py): Legendre is P^(0,0); Chebyshev is
        P^(-1/2,-1/2) scaled by 1 / C(j - 1/2, j), so T_j(1) = 1 and
        T_j'(1) = j^2.  An mpmath number when ``exact``."""
        import mpmath as mp

        from .Jacobi import jacobi_end_derivative

        if self.family not in ("legendre", "chebyshev"):
            raise NotImplementedError(f"boundary values of the {self.family!r} family")
        with mp.workdps(40):
            if self.family == "legendre":
                v = jacobi_end_derivative(mp.mpf(0), mp.mpf(0), mp.mpf(1), j, side, k)
            else:
                h = mp.mpf(-1) / 2
                v = jacobi_end_derivative(h, h, 1 / mp.binomial(j + h, j), j, side, k)
            return +v if exact else float(v)

    def evaluate_basis_derivative(self, X: torch.Tensor, k: int = 0) -> torch.Tensor:
        V = self.orthogonal.vandermonde(X, k)
        return self._apply_stencil_cols(V)

    def _stencil_dev(self):
        S = self.stencil_matrix()
        if S is None:
            return None
        key = "S"
        if key not in self._ops:
            self._ops[key] = K.AxisOperator(torch.tensor(S, dtype=torch.float64, device=K.device()))
        return self._ops[key]

    def _apply_stencil_cols(self, V: torch.Tensor) -> torch.Tensor:
        """V S^T: columns of an (Q, N) orthogonal matrix -> (Q, dim)."""
        S = self._stencil_dev()
        if S is None:
            return V
        return K.apply_along(S, V.contiguous(), 1, "f64")

    # ------------------------------------------------------------ device operators
    def backward_operator(self, N: int | None = None, k: int = 0,
                          kind: str = MeshKind.QUADRATURE) -> K.AxisOperator:
        """(Q, dim): series (or k-th derivative) at the mesh points
        (galerkin/orthogonal.py:214-246, galerkin/composite.py:205-218)."""
        N = self.num_quad_points if N is None else int(N)
        key = ("bwd", N, k, kind)
        if key not in self._ops:
            X = self._points(kind, N)
            V = self.orthogonal._vander(X, k, float(self.domain_factor) ** k)
            self._ops[key] = K.AxisOperator(self._apply_stencil_cols(V))
        return self._ops[key]

    def _weighted_vt(self, N: int) -> torch.Tensor:
        """(N_modes, Q): P^T diag(w) / df (galerkin/orthogonal.py:264-277, sg = 1)."""
        x, w = self.quad_points_and_weights(N)
        V = self.orthogonal._vander(x, 0)
        return K.scale_transpose(V, w, None, 1.0 / float(self.domain_factor))

    def scalar_product_operator(self, N: int | None = None) -> K.AxisOperator:
        """(dim, Q): (u, phi_n)_w (galerkin/orthogonal.py:264-277, composite.py:345-349)."""
        N = self.num_quad_points if N is None else int(N)
        key = ("sp", N)
        if key not in self._ops:
            WVt = self._weighted_vt(N)
            S = self._stencil_dev()
            self._ops[key] = K.AxisOperator(WVt if S is None else K.apply_along(S, WVt, 0, "f64"))
        return self._ops[key]

    def forward_operator(self, N: int | None = None) -> K.AxisOperator:
        """(dim, Q): orthogonal projection L / (h/df) (galerkin/orthogonal.py:256-262)."""
        N = self.num_quad_points if N is None else int(N)
        key = ("fwd", N)
        if key not in self._ops:
            x, w = self.quad_points_and_weights(N)
            V = self.orthogonal._vander(x, 0)
            invh = torch.tensor(1.0 / self.norm_squared(), dtype=torch.float64, device=x.device)
            self._ops[key] = K.AxisOperator(K.scale_transpose(V, w, invh, 1.0))
        return self._ops[key]

    # ------------------------------------------------------------ 1D transforms
    def backward(self, c, N: int | None = None, precision: str = "f64"):
        return _apply_1d(self.backward_operator(N), c, precision)

    def backward_primitive(self, c, k: int = 0, N: int | None = None, precision: str = "f64"):
        return _apply_1d(self.backward_operator(N, k), c, precision)

    def forward(self, u, precision: str = "f64"):
        return _apply_1d(self.forward_operator(_len(u)), u, precision)

    def scalar_product(self, u, precision: str = "f64"):
        return _apply_1d(self.scalar_product_operator(_len(u)), u, precision)

    def evaluate_mesh(self, c, kind: str = MeshKind.QUADRATURE, N: int | None = None):
        return _apply_1d(self.backward_operator(N, 0, kind), c, "f64")

    def evaluate(self, x, c):
        """Series values at arbitrary true-domain points x (galerkin/orthogonal.py:102-129):
        the basis evaluated at the points, contracted with the last axis of c.
        On an orthogonal space a shorter c is the series of its first len(c)
        modes (the reference's recurrence runs over len(c))."""
        shp = tuple(x.shape) if isinstance(x, torch.Tensor) else np.shape(x)
        X = self.map_reference_domain(as_device(x).reshape(-1)).contiguous()
        V = self._apply_stencil_cols(self.orthogonal._vander(X, 0))
        nc = int(np.shape(c)[-1])
        if self.is_orthogonal and nc < V.shape[1]:
            V = V[:, :nc]
        out = _apply_1d(K.AxisOperator(V.contiguous()), c, "f64")
        return out.reshape(tuple(out.shape[:-1]) + tuple(shp))

    def derivative_coeffs(self, c, k: int = 0):
        """Coefficients of d^k/dX^k (reference coordinate) of the series c, of
        the same length (Legendre.py:186-216, Chebyshev.py:281-315,
        Jacobi.py:204-247).  The reference runs a backward recurrence per
        series; here the exact map D_k = F V_k (the projection at N Gauss
        points, exact since the products have degree < 2N, of the k-th
        derivative Vandermonde) is built once on the device and applied along
        the last axis, so batches of series cost one contraction."""
        c = as_device(c)
        if k == 0:
            return c
        if not self.is_orthogonal:
            raise NotImplementedError("derivative_coeffs of an orthogonal basis (use to_orthogonal first)")
        key = ("dcoef", k)
        if key not in self._ops:
            x, _ = self.quad_points_and_weights(self.N)
            Vk = self._vander(x, k).contiguous()                                   # (N, N)
            self._ops[key] = K.AxisOperator(
                K.apply_along(self.forward_operator(self.N), Vk, 0, "f64").contiguous())
        D = self._ops[key]
        n = int(c.shape[-1])
        if n < self.N:  # a shorter series: its derivative stays in the first n modes
            D = K.AxisOperator(D.A64[:n, :n].contiguous())
        return K.apply_along(D, c, c.ndim - 1, "f64")

    def mass_matrix(self):
        from ..la import DiaMatrix

        h = self.norm_squared() / float(self.domain_factor)
        return DiaMatrix.from_dense(np.diag(h))

    def to_orthogonal(self, c):
        return c

    def from_orthogonal(self, c):
        return c

    # ------------------------------------------------------------ matrices
    def _matrices(self, i: int, trial, q: int = 0):
        return None

    def matrices(self, i: int, trial, q: int = 0, scale: float = 1.0):
        """Closed-form (test i, trial j) matrix or None (galerkin/orthogonal.py:479-529)."""
        u, j = trial
        if q != 0:
            return None
        z = self._matrices(i, trial, q)
        if z is None:
            return None
        df = float(self.domain_factor)
        return z * (scale * df ** (i + j - 1))

    def sympy_coordinate(self):
        return self.system.base_scalars()[0]

    def eval_coefficient(self, expr, N: int) -> torch.Tensor:
        """Evaluate a SymPy coefficient of x at the (true-domain) quadrature points."""
        xs = self.sympy_coordinate()
        pts = self.mesh(MeshKind.QUADRATURE, N).cpu().numpy()
        val = lambdify(xs, expr)(pts) if expr.free_symbols else float(expr) * np.ones_like(pts)
        return torch.tensor(np.broadcast_to(val, pts.shape).copy(), dtype=torch.float64,
                            device=K.device())

    def __repr__(self) -> str:
        return f"{type(self).__name__}(N={self.N}, domain={tuple(self.domain)})"


def _len(u) -> int:
    return int(u.shape[-1] if u.ndim == 1 else u.shape[-1])


def as_device(a, dtype=torch.float64) -> torch.Tensor:
    """Accept numpy arrays / torch tensors; return a contiguous device tensor.
    Complex input stays complex128 (complex data is transformed in FP64)."""
    if isinstance(a, torch.Tensor) and a.is_complex():
        t = a if a.is_cuda else a.to(K.device())
        return t.to(torch.complex128).contiguous()
    if not isinstance(a, torch.Tensor) and np.iscomplexobj(a):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.complex128, device=K.device()).contiguous()
    if isinstance(a, torch.Tensor):
        t = a if a.is_cuda else a.to(K.device())
        if t.dtype != dtype:
            t = K.cast(t.contiguous(), dtype) if t.dtype in (torch.float32, torch.float64) else t.to(dtype)
        return t.contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=K.device()).contiguous() \
        if dtype == torch.float64 else K.cast(as_device(a), dtype)


def _apply_1d(op: K.AxisOperator, c, precision: str):
    c = as_device(c, K.io_dtype(precision))
    return K.apply_along(op, c, c.ndim - 1, precision)


__all__ = ["OrthogonalSpace", "MeshKind", "as_device", "sp"]
