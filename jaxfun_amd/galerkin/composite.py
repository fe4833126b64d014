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


BC_ORDER = {"D": 0, "R": 0, "N": 1, "W": 1, "N2": 2, "N3": 3, "N4": 4}


def bc_order(key: str) -> int:
    """Derivative order of a condition code: D and R (Robin) 0, N and W
    (weighted Neumann) 1, N2..N4 2..4."""
    if key not in BC_ORDER:
        raise NotImplementedError(f"boundary condition {key!r}: D, R, N, W, N2, N3 and N4 are built")
    return BC_ORDER[key]


def _bc_value(v):
    """The value of a condition; a Robin condition is the pair (alpha, value)."""
    return v[1] if isinstance(v, (tuple, list)) else v


class BoundaryConditions(dict):
    """Left/right boundary conditions (galerkin/composite.py:28-118).

    Built from a plain dict with a domain, the values of derivative
    conditions Nk are divided by df^k (df = 2 / (b - a)), so that they are
    given in physical units and stored for the reference domain; the Robin
    condition R: (alpha, g), u + alpha u' = g, stores alpha * df for the same
    reason, and the weighted condition W: (alpha, g), u' + alpha u'' = g,
    stores (alpha * df, g / df).  A BoundaryConditions argument is taken as
    already normalised."""

    def __init__(self, bc: dict, domain=None) -> None:
        bcs = {"left": {}, "right": {}}
        bcs.update(copy.deepcopy(dict(bc)))
        dict.__init__(self, bcs)
        if domain is not None and not isinstance(bc, BoundaryConditions):
            df = 2.0 / (float(domain[1]) - float(domain[0]))
            for side in self.values():
                for k, v in side.items():
                    if k[0] == "N":
                        nd = int(k[1:]) if len(k) > 1 else 1
                        side[k] = v / df**nd
                    elif k == "R":
                        side[k] = (v[0] * df, v[1])
                    elif k == "W":
                        side[k] = (v[0] * df, v[1] / df)

    def orderednames(self) -> list[str]:
        return ["L" + b for b in sorted(self["left"])] + ["R" + b for b in sorted(self["right"])]

    def orderedvals(self) -> list:
        return [_bc_value(self[s][k]) for s in ("left", "right") for k in sorted(self[s])]

    def num_bcs(self) -> int:
        return len(self.orderedvals())

    def num_derivatives(self) -> int:
        """Sum of the derivative orders of all conditions (galerkin/composite.py:94-101)."""
        return sum(bc_order(k) for side in self.values() for k in side)

    def conditions(self) -> list[tuple[str, int, float]]:
        """(side "L"/"R", derivative order, Robin alpha or 0) of each
        condition, in ordered form."""
        out = []
        for nm in self.orderednames():
            v = self["left" if nm[0] == "L" else "right"][nm[1:]]
            out.append((nm[0], bc_order(nm[1:]), float(v[0]) if nm[1:] in ("R", "W") else 0.0))
        return out

    def is_homogeneous(self) -> bool:
        return all(_bc_value(v) == 0 for side in self.values() for v in side.values())

    def get_homogeneous(self) -> "BoundaryConditions":
        """Every value set to zero; Robin and weighted conditions keep their alpha."""
        return BoundaryConditions({k: {s: ((v[0], 0) if s in ("R", "W") else 0) for s, v in vals.items()}
                                   for k, vals in self.items()})


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
        self.bcs = BoundaryConditions(bcs)  # the stencil depends on the condition codes only
        self.nbc = self.bcs.num_bcs()
        self._conds = self.bcs.conditions()
        if self.nbc == 0 or N <= self.nbc:
            raise ValueError(f"need 0 < number of conditions < N (got {self.nbc} for N={N})")
        self._dirichlet = "".join(self.bcs.orderednames()) == "LDRD"
        self.scaling = sp.S.One if scaling is None else sp.sympify(scaling)
        if self._dirichlet and self.family in ("legendre", "chebyshev"):
            self.stencil = {0: 1 / self.scaling, 2: -1 / self.scaling}
        else:
            # per-row coefficients of each shift, before the row scaling
            S = self._general_stencil()
            self.stencil = {o: S.diagonal(o).copy() for o in range(self.nbc + 1) if np.any(S.diagonal(o) != 0)}
        self._mass_lu = None

    @property
    def dim(self) -> int:
        return self.N - self.nbc

    def _general_stencil(self) -> np.ndarray:
        """(N - nb, N) stencil of phi_k = psi_k + sum_{i=1..nb} d_i psi_{k+i}
        with every condition of ``bcs`` homogeneous (get_stencil_matrix,
        galerkin/composite.py:765-832): per k the nb x nb system of the
        conditions' end values, solved in 40-digit arithmetic; entries below
        1e-30 (the ones that vanish by symmetry) are set to zero."""
        if "S_general" in self._ops:
            return self._ops["S_general"]
        import mpmath as mp

        N, nb = self.N, self.nbc
        S = np.zeros((N - nb, N))
        with mp.workdps(40):
            f = [[_functional(self.orthogonal, j, side, k, al) for j in range(N)]
                 for side, k, al in self._conds]
            for row in range(N - nb):
                A = mp.matrix([[f[c][row + i] for i in range(1, nb + 1)] for c in range(nb)])
                b = mp.matrix([-f[c][row] for c in range(nb)])
                try:
                    d = mp.lu_solve(A, b)
                except ZeroDivisionError:
                    raise ValueError(f"the conditions {self.bcs.orderednames()} give a singular stencil row "
                                     f"{row} for this basis") from None
                S[row, row] = 1.0
                for i in range(nb):
                    S[row, row + 1 + i] = 0.0 if abs(d[i]) < mp.mpf(10) ** -30 else float(d[i])
        self._ops["S_general"] = S
        return S

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
        if not (self._dirichlet and self.family in ("legendre", "chebyshev")):
            return self._general_stencil() / s[:, None]
        S = np.zeros((m, self.N))
        S[np.arange(m), np.arange(m)] = 1.0 / s
        S[np.arange(m), np.arange(m) + 2] = -1.0 / s
        return S

    @property
    def bandwidth(self) -> int:
        """Half bandwidth w of S S^T and of the mass matrix: the stencil reaches nb modes ahead."""
        return self.nbc

    @staticmethod
    def _to_band(M: np.ndarray, w: int) -> np.ndarray:
        m = M.shape[0]
        band = np.zeros((m, 2 * w + 1))
        for i in range(m):
            for o in range(-w, w + 1):
                if 0 <= i + o < m:
                    band[i, o + w] = M[i, i + o]
        return band

    def mass_band(self) -> np.ndarray:
        """S diag(h/df) S^T as a (dim, 2w+1) by-row band, offsets -w..w (composite.py:330-337)."""
        S = self.stencil_matrix()
        M = S @ np.diag(self.orthogonal.norm_squared() / float(self.domain_factor)) @ S.T
        return self._to_band(M, self.bandwidth)

    def mass_lu(self) -> torch.Tensor:
        if self._mass_lu is None:
            band = torch.tensor(self.mass_band(), dtype=torch.float64, device=K.device())
            self._mass_lu = K.band_lu_factor(band, self.bandwidth, self.bandwidth)
        return self._mass_lu

    def mass_matrix(self):
        from ..la import DiaMatrix

        band, w = self.mass_band(), self.bandwidth
        m = self.dim
        M = np.zeros((m, m))
        for i in range(m):
            for o in range(-w, w + 1):
                if 0 <= i + o < m:
                    M[i, i + o] = band[i, o + w]
        return DiaMatrix.from_dense(M)

    def forward_operator(self, N: int | None = None) -> K.AxisOperator:
        """(dim, Q): (S M S^T)^-1 S P^T W, the banded mass solve of
        ``Composite.forward`` (galerkin/composite.py:339-343) folded into the
        operator by solving it on every column on the device."""
        N = self.num_quad_points if N is None else int(N)
        key = ("fwd", N)
        if key not in self._ops:
            L = self.scalar_product_operator(N).A64.clone()
            K.band_lu_solve(self.mass_lu(), self.bandwidth, self.bandwidth, L, 0)
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
            w = self.bandwidth
            band = self._to_band(S @ S.T, w)
            self._ops[key] = K.band_lu_factor(torch.tensor(band, device=K.device()), w, w)
        return K.band_lu_solve(self._ops[key], self.bandwidth, self.bandwidth, b, b.ndim - 1)

    def get_homogeneous(self):
        """This basis with every boundary value set to zero (galerkin/composite.py)."""
        if self.bcs.is_homogeneous():
            return self
        h = copy.copy(self)
        h.bcs = self.bcs.get_homogeneous()
        return h

    def get_orthogonal(self):
        return self.orthogonal

    def __add__(self, b: "BCGeneric") -> "DirectSum":
        """Direct sum self (+) b (galerkin/composite.py:405-407)."""
        return DirectSum(self, b)


# ----------------------------------------------------------------------------- lifting
def _functional(orthogonal, j: int, side: str, k: int, alpha: float):
    """A condition applied to psi_j: d^k psi_j at the end, plus alpha
    d^(k+1) psi_j there for a Robin (k = 0) or weighted (k = 1) condition
    (get_stencil_matrix, composite.py:806-822)."""
    import mpmath as mp

    v = mp.mpf(orthogonal.bnd_value(j, side, k, exact=True))
    if alpha:
        v += mp.mpf(alpha) * mp.mpf(orthogonal.bnd_value(j, side, k + 1, exact=True))
    return v


def get_bc_basis(bcs, orthogonal) -> np.ndarray:
    """Lifting basis (num_bcs, first + num_bcs) (galerkin/composite.py:835-895).

    Row r holds the orthogonal coefficients of the function whose condition r
    equals 1 and whose other conditions vanish: with the boundary matrix
    ``A[r, j] = d^{k_r} psi_{first+j}(x_r)`` (x_r the end of condition r,
    k_r its derivative order) the rows are those of ``(A^-1)^T``, placed in
    columns first..first+nb-1.  ``first`` is the smallest offset (up to the
    number of derivatives, plus one) whose A is invertible; for ``{left: D, right: D}``
    on Legendre/Chebyshev first = 0 and the rows are (psi_0 -+ psi_1) / 2.
    End values come from the orthogonal space (``bnd_value``), so the Jacobi
    family's scalings are honoured.
    """
    import mpmath as mp

    bcs = bcs if isinstance(bcs, BoundaryConditions) else BoundaryConditions(bcs)
    conds = bcs.conditions()
    nb = len(conds)
    with mp.workdps(40):
        # one offset beyond the reference's search, for Robin conditions whose
        # first functionals coincide (u + 2u' at x = -1 on P_0, P_1)
        for first in range(bcs.num_derivatives() + 2):
            A = mp.matrix([[_functional(orthogonal, first + j, side, k, al) for j in range(nb)]
                           for side, k, al in conds])
            try:
                singular = abs(mp.det(A)) < mp.mpf(10) ** -20
            except (ZeroDivisionError, TypeError, ValueError):  # mpmath's LU finds no pivot
                singular = True
            if singular:
                continue
            Ainv = A ** -1
            S = np.zeros((nb, first + nb))
            for r in range(nb):
                for j in range(nb):
                    S[r, first + j] = float(Ainv[j, r])
            return S
    raise ValueError(f"no lifting basis for {bcs.orderednames()}")


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
        fam = family_kwargs(kw)
        # the orthogonal space is sized once the lifting basis is known
        probe = orthogonal(nb + self.bcs.num_derivatives(), domain=domain, system=system, **fam)
        S = get_bc_basis(self.bcs, probe)
        OrthogonalSpace.__init__(self, S.shape[1], domain=domain, system=system, name=name, fun_str=fun_str,
                                 num_quad_points=nq)
        self.bcs = BoundaryConditions(bcs)
        self.orthogonal = orthogonal(S.shape[1], domain=domain, system=system, num_quad_points=nq, **fam)
        self.family = self.orthogonal.family
        self.S = S
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
