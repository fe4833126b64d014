"""Boundary-constrained composite bases (reference: galerkin/composite.py).

Covers Dirichlet conditions on the Legendre, Chebyshev and Jacobi-family
bases: the homogeneous stencil ``phi_k = (psi_k - psi_{k+2}) / s_k``
(galerkin/composite.py:786-789), or for the Jacobi family
``phi_k = psi_k + d1_k psi_{k+1} + d2_k psi_{k+2}`` solved from the boundary
values (galerkin/composite.py:765-832), with its transforms, and inhomogeneous values through the lifting space
``BCGeneric`` and the direct sum ``DirectSum = Composite (+) BCGeneric``
(galerkin/composite.py:411-638).  The lift is a constant coefficient vector
added on the device (``kernels.add_bcast``) to every transform.
"""

from __future__ import annotations

import copy

import numpy as np
import sympy as sp
import torch

from .. import kernels as K
from ..utils.common import lambdify, n
from .orthogonal import OrthogonalSpace


class BoundaryConditions(dict):
    """Left/right boundary conditions (galerkin/composite.py:28-118)."""

    def __init__(self, bc: dict, domain=None) -> None:
        bcs = {"left": {}, "right": {}}
        bcs.update(copy.deepcopy(dict(bc)))
        dict.__init__(self, bcs)

    def orderednames(self) -> list[str]:
        return ["L" + b for b in sorted(self["left"])] + ["R" + b for b in sorted(self["right"])]

    def orderedvals(self) -> list:
        return [self[s][k] for s in ("left", "right") for k in sorted(self[s])]

    def num_bcs(self) -> int:
        return len(self.orderedvals())

    def is_homogeneous(self) -> bool:
        return all(v == 0 for side in self.values() for v in side.values())

    def get_homogeneous(self) -> "BoundaryConditions":
        return BoundaryConditions({k: {s: 0 for s in v} for k, v in self.items()})


def family_kwargs(kw: dict) -> dict:
    """Pop the parameters of a Jacobi-family basis (alpha, beta, lambda_) from kw."""
    return {k: kw.pop(k) for k in ("alpha", "beta", "lambda_") if k in kw}


class Composite(OrthogonalSpace):
    """Basis phi_i = sum_j S_ij psi_j enforcing homogeneous Dirichlet BCs
    (galerkin/composite.py:121-408)."""

    is_orthogonal = False

    def __init__(self, N: int, orthogonal: type, bcs, domain=None, name: str = "Composite",
                 fun_str: str = "phi", system=None, stencil=None, scaling=None, **kw):
        fam = family_kwargs(kw)
        super().__init__(N, domain=domain, system=system, name=name, fun_str=fun_str, **kw)
        self.orthogonal = orthogonal(N, domain=domain, system=system, **fam)
        self.family = self.orthogonal.family
        self.bcs = BoundaryConditions(bcs)
        if "".join(self.bcs.orderednames()) != "LDRD" or not self.bcs.is_homogeneous():
            raise NotImplementedError(
                "only homogeneous Dirichlet composite bases are built on the GPU path "
                f"(got {dict(self.bcs)})")
        self.scaling = sp.S.One if scaling is None else sp.sympify(scaling)
        self._jacobi_stencil = getattr(self.orthogonal, "dirichlet_stencil", None)
        if self._jacobi_stencil is not None:
            S = self._jacobi_stencil()
            self.stencil = {0: 1 / self.scaling, 1: S[:, 1:-1].diagonal() / self.scaling,
                            2: S[:, 2:].diagonal() / self.scaling}
        else:
            self.stencil = {0: 1 / self.scaling, 2: -1 / self.scaling}
        self._mass_lu = None

    @property
    def dim(self) -> int:
        return self.N - 2

    def norm_squared(self) -> np.ndarray:
        return self.orthogonal.norm_squared()

    def scaling_values(self) -> np.ndarray:
        k = np.arange(self.dim)
        if self.scaling == sp.S.One:
            return np.ones(self.dim)
        return np.broadcast_to(lambdify(n, self.scaling)(k), (self.dim,)).astype(float)

    def stencil_matrix(self) -> np.ndarray:
        """S (dim, N) (galerkin/composite.py:255-269)."""
        m = self.dim
        s = self.scaling_values()
        if self._jacobi_stencil is not None:
            if "S_host" not in self._ops:
                self._ops["S_host"] = self._jacobi_stencil()
            return self._ops["S_host"] / s[:, None]
        S = np.zeros((m, self.N))
        S[np.arange(m), np.arange(m)] = 1.0 / s
        S[np.arange(m), np.arange(m) + 2] = -1.0 / s
        return S

    def mass_band(self) -> np.ndarray:
        """S diag(h/df) S^T as a (dim, 5) by-row band, offsets -2..2 (composite.py:330-337)."""
        S = self.stencil_matrix()
        M = S @ np.diag(self.orthogonal.norm_squared() / float(self.domain_factor)) @ S.T
        m = self.dim
        band = np.zeros((m, 5))
        for i in range(m):
            for o in range(-2, 3):
                if 0 <= i + o < m:
                    band[i, o + 2] = M[i, i + o]
        return band

    def mass_lu(self) -> torch.Tensor:
        if self._mass_lu is None:
            band = torch.tensor(self.mass_band(), dtype=torch.float64, device=K.device())
            self._mass_lu = K.band_lu_factor(band, 2, 2)
        return self._mass_lu

    def mass_matrix(self):
        from ..la import DiaMatrix

        band = self.mass_band()
        m = self.dim
        M = np.zeros((m, m))
        for i in range(m):
            for o in range(-2, 3):
                if 0 <= i + o < m:
                    M[i, i + o] = band[i, o + 2]
        return DiaMatrix.from_dense(M)

    def forward_operator(self, N: int | None = None) -> K.AxisOperator:
        """(dim, Q): (S M S^T)^-1 S P^T W, the banded mass solve of
        ``Composite.forward`` (galerkin/composite.py:339-343) folded into the
        operator by solving it on every column on the device."""
        N = self.num_quad_points if N is None else int(N)
        key = ("fwd", N)
        if key not in self._ops:
            L = self.scalar_product_operator(N).A64.clone()
            K.band_lu_solve(self.mass_lu(), 2, 2, L, 0)
            self._ops[key] = K.AxisOperator(L)
        return self._ops[key]

    def to_orthogonal(self, c):
        """c @ S (galerkin/composite.py:276-279)."""
        from .orthogonal import as_device

        c = as_device(c)
        ST = self._ops.get("ST")
        if ST is None:
            ST = self._ops["ST"] = K.AxisOperator(
                torch.tensor(self.stencil_matrix().T.copy(), dtype=torch.float64, device=K.device()))
        return K.apply_along(ST, c, c.ndim - 1, "f64")

    def from_orthogonal(self, a):
        """(S S^T)^-1 S a (galerkin/composite.py:281-284)."""
        from .orthogonal import as_device

        a = as_device(a)
        b = K.apply_along(self._stencil_dev(), a, a.ndim - 1, "f64").contiguous()
        key = "P_lu"
        if key not in self._ops:
            S = self.stencil_matrix()
            P = S @ S.T
            m = self.dim
            band = np.zeros((m, 5))
            for i in range(m):
                for o in range(-2, 3):
                    if 0 <= i + o < m:
                        band[i, o + 2] = P[i, i + o]
            self._ops[key] = K.band_lu_factor(torch.tensor(band, device=K.device()), 2, 2)
        return K.band_lu_solve(self._ops[key], 2, 2, b, b.ndim - 1)

    def get_homogeneous(self):
        return self

    def get_orthogonal(self):
        return self.orthogonal

    def __add__(self, b: "BCGeneric") -> "DirectSum":
        """Direct sum self (+) b (galerkin/composite.py:405-407)."""
        return DirectSum(self, b)


# ----------------------------------------------------------------------------- lifting
def get_bc_basis(bcs, orthogonal) -> np.ndarray:
    """Lifting basis (num_bcs, num_bcs) (galerkin/composite.py:835-895).

    Row r holds the orthogonal coefficients of the function that takes the
    value 1 in boundary condition r and 0 in the others: with the boundary
    matrix ``A[r, j] = psi_j(x_r)`` over the first num_bcs orthogonal
    functions, the rows are those of ``(A^-1)^T``.  Dirichlet conditions
    only (the conditions of the composite bases built here); for
    ``{left: D, right: D}`` on Legendre/Chebyshev the rows are
    (psi_0 -+ psi_1) / 2.  Boundary values come from the orthogonal space
    (``bnd_value``), so the Jacobi family's scalings are honoured.
    """
    bcs = bcs if isinstance(bcs, BoundaryConditions) else BoundaryConditions(bcs)
    names = bcs.orderednames()
    if any(nm[1:] != "D" for nm in names):
        raise NotImplementedError(f"lifting of {names}: only Dirichlet conditions are built")
    nb = len(names)
    A = np.array([[orthogonal.bnd_value(j, nm[0]) for j in range(nb)] for nm in names])
    return np.linalg.inv(A).T.copy()


def _as_number(v):
    e = sp.sympify(v)
    if e.free_symbols:
        raise ValueError(f"boundary value {v} depends on {e.free_symbols}; a 1D space needs numbers")
    return complex(e) if e.has(sp.I) else float(e)


class BCGeneric(Composite):
    """Space of the boundary lifting functions only (galerkin/composite.py:410-475).

    ``dim`` = number of boundary conditions, ``num_dofs`` = 0.  Its stencil
    ``S`` (num_bcs, num_bcs) comes from :func:`get_bc_basis`; the orthogonal
    space has num_bcs modes.
    """

    is_orthogonal = False

    def __init__(self, N: int, orthogonal: type, bcs, domain=None, name: str = "BCGeneric",
                 fun_str: str = "B", system=None, stencil=None, num_quad_points: int = 0, **kw):
        self.bcs = BoundaryConditions(bcs)
        nb = self.bcs.num_bcs()
        nq = int(num_quad_points) if num_quad_points else nb
        OrthogonalSpace.__init__(self, nb, domain=domain, system=system, name=name, fun_str=fun_str,
                                 num_quad_points=nq)
        self.bcs = BoundaryConditions(bcs)
        self.orthogonal = orthogonal(nb, domain=domain, system=system, num_quad_points=nq, **family_kwargs(kw))
        self.family = self.orthogonal.family
        self.S = get_bc_basis(self.bcs, self.orthogonal)
        self.scaling = sp.S.One
        self.stencil = None
        self._mass_lu = None

    @property
    def dim(self) -> int:
        return self.S.shape[0]

    @property
    def num_dofs(self) -> int:
        return 0

    def stencil_matrix(self) -> np.ndarray:
        return self.S

    def bnd_vals(self) -> np.ndarray:
        """Ordered boundary values as numbers (galerkin/composite.py:459-466)."""
        vals = [_as_number(v) for v in self.bcs.orderedvals()]
        return np.array(vals, dtype=np.complex128 if any(isinstance(v, complex) for v in vals) else np.float64)

    def get_homogeneous(self):
        raise NotImplementedError("BCGeneric does not support get_homogeneous()")


class DirectSum:
    """V = Composite (+) BCGeneric: homogeneous unknowns plus a fixed lift
    carrying the boundary values (galerkin/composite.py:490-638).

    Coefficient arrays hold the homogeneous part only (``dim`` = dim of the
    Composite).  ``to_orthogonal`` adds the lift's padded orthogonal
    coefficients; ``backward``/``evaluate_mesh``/``backward_primitive`` add
    the lift's values (or derivatives) on the requested mesh, cached per mesh;
    ``forward`` is the orthogonal projection followed by ``from_orthogonal``;
    ``scalar_product`` uses the homogeneous test functions only.  Leading
    batch dimensions are allowed; the lift is broadcast over them.
    """

    is_transient = False
    is_orthogonal = False

    def __init__(self, a: Composite, b: BCGeneric) -> None:
        if not isinstance(b, BCGeneric):
            raise TypeError("DirectSum needs a BCGeneric second summand")
        self.basespaces = (a, b)
        self.bcs = b.bcs
        self.name = a.name + "\u2295" + b.name
        self._ops: dict = {}

    def __getitem__(self, i: int):
        return self.basespaces[i]

    def __len__(self) -> int:
        return len(self.basespaces)

    def __iter__(self):
        return iter(self.basespaces)

    def __repr__(self) -> str:
        return f"DirectSum({self[0]!r}, bcs={dict(self.bcs)})"

    # ------------------------------------------------------------ delegated geometry
    @property
    def system(self):
        return self[0].system

    @property
    def N(self) -> int:
        return self[0].N

    @property
    def family(self) -> str:
        return self[0].family

    @property
    def domain(self):
        return self[0].domain

    @property
    def domain_factor(self) -> float:
        return self[0].domain_factor

    @property
    def num_quad_points(self) -> int:
        return self[0].num_quad_points

    @property
    def orthogonal(self):
        return self[0].orthogonal

    @property
    def dim(self) -> int:
        return self[0].dim

    @property
    def num_dofs(self) -> int:
        return self[0].num_dofs

    @property
    def dims(self) -> int:
        return 1

    @property
    def shape(self) -> tuple[int, ...]:
        return (self.num_quad_points,)

    def map_reference_domain(self, x):
        return self[0].map_reference_domain(x)

    def map_true_domain(self, X):
        return self[0].map_true_domain(X)

    def quad_points_and_weights(self, N: int | None = None):
        return self[0].quad_points_and_weights(N)

    def mesh(self, kind: str = "quadrature", N: int | None = None):
        return self[0].mesh(kind, N)

    def get_homogeneous(self) -> Composite:
        return self[0]

    def get_orthogonal(self):
        return self[0].get_orthogonal()

    def bnd_vals(self) -> np.ndarray:
        return self[1].bnd_vals()

    def with_system(self, system) -> "DirectSum":
        """Copy with both summands on ``system`` (device operators shared)."""
        parts = []
        for s in self.basespaces:
            c = copy.copy(s)
            c._ops = s._ops
            c.system = system
            c.orthogonal = copy.copy(s.orthogonal)
            c.orthogonal._ops = s.orthogonal._ops
            c.orthogonal.system = system
            parts.append(c)
        out = DirectSum(parts[0], parts[1])
        out._ops = self._ops
        return out

    # ------------------------------------------------------------ lift
    def lift_coefficients(self) -> torch.Tensor:
        """(N,): the lift in the orthogonal basis, ``pad(bnd_vals @ S_b)``
        (galerkin/composite.py:588-593)."""
        if "lift" not in self._ops:
            cb = self.bnd_vals() @ self[1].S
            full = np.zeros(self.N, dtype=cb.dtype)
            full[: cb.shape[0]] = cb
            self._ops["lift"] = torch.tensor(full, device=K.device())
        return self._ops["lift"]

    def lift_values(self, N: int | None = None, k: int = 0, kind: str = "quadrature") -> torch.Tensor:
        """(Q,): k-th derivative of the lift on the mesh of ``kind`` with N points."""
        N = self.num_quad_points if N is None else int(N)
        key = ("liftv", N, k, kind)
        if key not in self._ops:
            op = self.orthogonal.backward_operator(N, k, kind)
            self._ops[key] = K.apply_along(op, self.lift_coefficients().reshape(1, -1), 1,
                                           "f64").reshape(-1).contiguous()
        return self._ops[key]

    @staticmethod
    def _f64(precision: str) -> None:
        if precision != "f64":
            raise ValueError("inhomogeneous (lifted) spaces transform in f64 only")

    # ------------------------------------------------------------ transforms
    def to_orthogonal(self, c):
        """Homogeneous part mapped by the stencil plus the lift (composite.py:588-593)."""
        from .orthogonal import as_device

        return K.add_bcast(self[0].to_orthogonal(as_device(c)).contiguous(), self.lift_coefficients())

    def from_orthogonal(self, a):
        """(S S^T)^-1 S (a - lift) (galerkin/composite.py:595-601)."""
        from .orthogonal import as_device

        a = as_device(a)
        return self[0].from_orthogonal(K.add_bcast(a, self.lift_coefficients(), -1.0))

    def backward(self, c, N: int | None = None, precision: str = "f64"):
        """orthogonal.backward(to_orthogonal(c)) (composite.py:613-616), as the
        homogeneous backward plus the lift's values on the mesh."""
        self._f64(precision)
        return K.add_bcast(self[0].backward(c, N).contiguous(), self.lift_values(N))

    def backward_primitive(self, c, k: int = 0, N: int | None = None, precision: str = "f64"):
        """k-th derivative at the mesh (galerkin/composite.py:618-626)."""
        self._f64(precision)
        return K.add_bcast(self[0].backward_primitive(c, k, N).contiguous(), self.lift_values(N, k))

    def evaluate_mesh(self, c, kind: str = "quadrature", N: int | None = None):
        """Series values on a quadrature or uniform mesh (composite.py:606-611)."""
        return K.add_bcast(self[0].evaluate_mesh(c, kind, N).contiguous(), self.lift_values(N, 0, kind))

    def evaluate(self, x, c):
        """Series values at arbitrary points x of the true domain (composite.py:603-605)."""
        return self.orthogonal.evaluate(x, self.to_orthogonal(c))

    def forward(self, u, precision: str = "f64"):
        """from_orthogonal(orthogonal.forward(u)) (galerkin/composite.py:628-631)."""
        self._f64(precision)
        return self.from_orthogonal(self.orthogonal.forward(u))

    def scalar_product(self, u, precision: str = "f64"):
        """(u, phi_i) with the homogeneous test functions (composite.py:633-636)."""
        return self[0].scalar_product(u, precision)
