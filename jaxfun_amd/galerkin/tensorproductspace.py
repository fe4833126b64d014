"""Tensor-product spaces (reference: galerkin/tensorproductspace.py).

Every transform is a sequence of per-axis contractions with the 1D spaces'
device operators, run by the MFMA kernels of ``libjfx``.  Leading batch
dimensions are allowed (the reference vmaps externally).  ``precision``
selects the contraction arithmetic: ``f64`` (default), ``f32``, ``bf16x3``,
``bf16x6`` (fp32 I/O) or ``ozaki3/4/5`` (fp64 I/O, fixed-point bf16 slices).
"""

from __future__ import annotations

import copy

import torch

from .. import kernels as K
from ..coordinates import CartCoordSys, default_system
from .orthogonal import MeshKind, OrthogonalSpace, as_device


class TensorProductSpace:
    """d-dimensional tensor product of 1D spaces (tensorproductspace.py:38-504)."""

    def __init__(self, basespaces, system: CartCoordSys | None = None, name: str = "TPS",
                 precision: str = "f64"):
        self.basespaces: list[OrthogonalSpace] = list(basespaces)
        self.system = default_system(len(self.basespaces)) if system is None else system
        self.name = name
        self.tensorname = "⊗".join(b.name for b in self.basespaces)
        self.precision = precision
        # fuse the last two axes into one kernel for batches of fewer planes than
        # CUs (one launch instead of two); larger batches keep every CU busy with
        # the per-axis kernels, which run closer to the MFMA / HBM roofs
        self.fused = True

    def __len__(self) -> int:
        return len(self.basespaces)

    def __iter__(self):
        return iter(self.basespaces)

    def __getitem__(self, i: int) -> OrthogonalSpace:
        return self.basespaces[i]

    @property
    def dims(self) -> int:
        return len(self)

    @property
    def is_orthogonal(self) -> bool:
        return all(s.is_orthogonal for s in self.basespaces)

    @property
    def shape(self) -> tuple[int, ...]:
        return tuple(s.num_quad_points for s in self.basespaces)

    @property
    def num_quad_points(self) -> tuple[int, ...]:
        return self.shape

    @property
    def dim(self) -> int:
        out = 1
        for s in self.basespaces:
            out *= s.dim
        return out

    @property
    def num_dofs(self) -> tuple[int, ...]:
        return tuple(s.dim for s in self.basespaces)

    def mesh(self, kind: str = MeshKind.QUADRATURE, N=None, broadcast: bool = True):
        """Per-axis meshes, broadcastable to the full grid (tensorproductspace.py:158-220)."""
        d = len(self)
        out = []
        for ax, s in enumerate(self.basespaces):
            m = s.mesh(kind, None if N is None else N[ax])
            if broadcast:
                shape = [1] * d
                shape[ax] = m.shape[0]
                m = m.reshape(shape)
            out.append(m)
        return tuple(out)

    # ------------------------------------------------------------ transforms
    def _run(self, X, ops, precision):
        prec = self.precision if precision is None else precision
        X = as_device(X, K.io_dtype(prec))
        nb = X.ndim - len(self)
        if nb < 0:
            raise ValueError(f"array of rank {X.ndim} is smaller than the space dimension {len(self)}")
        d = len(ops)
        nplanes = 1
        for n_ in X.shape[:-2]:
            nplanes *= n_
        if d >= 2 and self.fused and nplanes < K.num_cus() and K.plane_ok(ops[-2], ops[-1], prec):
            # leading axes one by one, the last two in one fused pass
            for ax, op in enumerate(ops[:-2]):
                X = K.apply_along(op, X, nb + ax, prec)
            return K.plane(ops[-2], ops[-1], X.contiguous(), prec)
        for ax, op in enumerate(ops):
            X = K.apply_along(op, X, nb + ax, prec)
        return X

    def backward(self, c, N=None, precision: str | None = None):
        """Coefficients -> values at the (possibly padded) quadrature mesh (330-365)."""
        N = tuple(s.num_quad_points if N is None or N[i] is None else N[i]
                  for i, s in enumerate(self.basespaces))
        return self._run(c, [s.backward_operator(N[i]) for i, s in enumerate(self.basespaces)], precision)

    def backward_primitive(self, c, k, N=None, precision: str | None = None):
        """Mixed derivatives d^k u at the quadrature mesh (419-460)."""
        N = tuple(s.num_quad_points if N is None or N[i] is None else N[i]
                  for i, s in enumerate(self.basespaces))
        return self._run(c, [s.backward_operator(N[i], k[i]) for i, s in enumerate(self.basespaces)],
                         precision)

    def scalar_product(self, u, precision: str | None = None):
        """(u, phi) over the quadrature mesh, truncating padded input (367-393)."""
        u = as_device(u, K.io_dtype(self.precision if precision is None else precision))
        nb = u.ndim - len(self)
        return self._run(u, [s.scalar_product_operator(u.shape[nb + i])
                             for i, s in enumerate(self.basespaces)], precision)

    def forward(self, u, precision: str | None = None):
        """Values at the quadrature mesh -> coefficients (395-417)."""
        u = as_device(u, K.io_dtype(self.precision if precision is None else precision))
        nb = u.ndim - len(self)
        return self._run(u, [s.forward_operator(u.shape[nb + i]) for i, s in enumerate(self.basespaces)],
                         precision)

    def evaluate_mesh(self, c, kind: str = MeshKind.QUADRATURE, N=None):
        """Series values on a quadrature or uniform mesh (222-261)."""
        N = tuple(s.num_quad_points if N is None or N[i] is None else N[i]
                  for i, s in enumerate(self.basespaces))
        return self._run(c, [s.backward_operator(N[i], 0, kind) for i, s in enumerate(self.basespaces)],
                         "f64")

    def to_orthogonal(self, c):
        c = as_device(c)
        nb = c.ndim - len(self)
        for ax, s in enumerate(self.basespaces):
            c = c.movedim(nb + ax, -1).contiguous()
            c = s.to_orthogonal(c).movedim(-1, nb + ax).contiguous()
        return c

    def from_orthogonal(self, a):
        a = as_device(a)
        nb = a.ndim - len(self)
        for ax, s in enumerate(self.basespaces):
            a = a.movedim(nb + ax, -1).contiguous()
            a = s.from_orthogonal(a).movedim(-1, nb + ax).contiguous()
        return a

    def get_orthogonal(self) -> "TensorProductSpace":
        return TensorProductSpace([s.get_orthogonal() for s in self.basespaces], self.system,
                                  self.name + "o", self.precision)


def TensorProduct(*basespaces, system: CartCoordSys | None = None, name: str = "T",
                  precision: str = "f64") -> TensorProductSpace:
    """Factory (tensorproductspace.py:507-557): copies the 1D spaces and
    assigns each the sub-system of its axis."""
    system = default_system(len(basespaces)) if system is None else system
    spaces = []
    for i, s in enumerate(basespaces):
        c = copy.copy(s)
        c._ops = s._ops  # device operators are shared, not copied
        c.system = system.sub_system(i)
        if hasattr(c, "orthogonal") and c.orthogonal is not s:
            c.orthogonal = copy.copy(s.orthogonal)
            c.orthogonal._ops = s.orthogonal._ops
            c.orthogonal.system = c.system
        elif getattr(s, "orthogonal", None) is s:
            c.orthogonal = c
        spaces.append(c)
    return TensorProductSpace(spaces, system, name, precision)


__all__ = ["TensorProductSpace", "TensorProduct", "torch"]
