"""Boundary-constrained composite bases (reference: galerkin/composite.py).

Covers homogeneous Dirichlet conditions on Legendre / Chebyshev bases, the
stencil ``phi_k = (psi_k - psi_{k+2}) / s_k`` (galerkin/composite.py:786-789)
and its transforms.  Inhomogeneous lifting (BCGeneric / DirectSum) is not on
this build's path and raises NotImplementedError.
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


class Composite(OrthogonalSpace):
    """Basis phi_i = sum_j S_ij psi_j enforcing homogeneous Dirichlet BCs
    (galerkin/composite.py:121-408)."""

    is_orthogonal = False

    def __init__(self, N: int, orthogonal: type, bcs, domain=None, name: str = "Composite",
                 fun_str: str = "phi", system=None, stencil=None, scaling=None, **kw):
        super().__init__(N, domain=domain, system=system, name=name, fun_str=fun_str, **kw)
        self.orthogonal = orthogonal(N, domain=domain, system=system)
        self.family = self.orthogonal.family
        self.bcs = BoundaryConditions(bcs)
        if "".join(self.bcs.orderednames()) != "LDRD" or not self.bcs.is_homogeneous():
            raise NotImplementedError(
                "only homogeneous Dirichlet composite bases are built on the GPU path "
                f"(got {dict(self.bcs)})")
        self.scaling = sp.S.One if scaling is None else sp.sympify(scaling)
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
