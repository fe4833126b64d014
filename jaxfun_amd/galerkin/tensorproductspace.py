"""Tensor-product spaces (reference: galerkin/tensorproductspace.py).

Every transform is a sequence of per-axis contractions with the 1D spaces'
device operators, run by the MFMA kernels of ``libjfx``.  Leading batch
dimensions are allowed (the reference vmaps externally).  ``precision``
selects the contraction arithmetic: ``f64`` (default), ``f32``, ``bf16x3``,
``bf16x6`` (fp32 I/O), ``ozaki3/4/5`` (fp64 I/O, fixed-point bf16 slices) or
``i8x5/6/7`` (fp64 I/O, balanced int8 digit slices on int8 MFMA with exact
integer sums; i8x6 keeps a 16 x 128^3 Legendre round trip near 1e-8).
"""

from __future__ import annotations

import copy
import itertools

import numpy as np
import sympy as sp
import torch

from .. import kernels as K
from ..coordinates import CartCoordSys, default_system
from ..utils.common import lambdify
from .composite import BCGeneric, BoundaryConditions, DirectSum, bc_order
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
        # the int8 plane pass saves an FP64 HBM round trip at any batch size;
        # the others pay off for batches of fewer slabs than CUs
        if d >= 2 and self.fused and not X.is_complex() and \
                (nplanes < K.num_cus() or prec.startswith("i8x")) and K.plane_ok(ops[-2], ops[-1], prec):
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

    def map_expr_true_domain(self, u):
        """Every axis' reference coordinate mapped to its true coordinate
        (galerkin/tensorproductspace.py:210-214)."""
        for s in self.basespaces:
            u = s.map_expr_true_domain(u)
        return u

    def map_expr_reference_domain(self, u):
        """Every axis' true coordinate mapped to its reference coordinate
        (galerkin/tensorproductspace.py:216-220)."""
        for s in self.basespaces:
            u = s.map_expr_reference_domain(u)
        return u

    def _point_basis(self, ax: int, X: torch.Tensor) -> torch.Tensor:
        """(P, dim) basis of axis ax at true-domain points X."""
        s = self.basespaces[ax]
        Xr = s.map_reference_domain(X).contiguous()
        if getattr(s, "is_fourier", False):
            return s._basis(Xr.shape[0], 0, False, 1.0, Xr)
        return s._apply_stencil_cols(s.orthogonal._vander(Xr, 0)).contiguous()

    def evaluate(self, x, c):
        """Series values at scattered points (tensorproductspace.py:263-321).

        ``x`` holds one coordinate array per axis, all of one shape; ``c`` may
        carry leading batch dimensions.  The last axis is a contraction with
        its basis at the points; every other axis is a pointwise product with
        its basis at the points followed by a sum over its modes.
        """
        from ..kernels import mul_bcast

        shp = tuple(x[0].shape) if isinstance(x[0], torch.Tensor) else np.shape(x[0])
        xs = [as_device(xi).reshape(-1) for xi in x]
        d = len(self)
        B = [self._point_basis(ax, xs[ax]) for ax in range(d)]
        c = as_device(c)
        op = K.ComplexAxisOperator(B[-1]) if B[-1].is_complex() else K.AxisOperator(B[-1])
        X = K.apply_along(op, c, c.ndim - 1)
        for ax in range(d - 2, -1, -1):
            Bt = B[ax].T.contiguous()
            X = X.contiguous()
            if X.is_complex() or Bt.is_complex():
                X = K.cmul_bcast(X, Bt if Bt.is_complex() else K.to_complex(Bt))
            else:
                X = mul_bcast(X, Bt)
            ones = K.AxisOperator(torch.ones((1, Bt.shape[0]), dtype=torch.float64, device=X.device))
            X = K.apply_along(ones, X, X.ndim - 2)
            X = X.reshape(X.shape[:-2] + X.shape[-1:])
        return X.reshape(X.shape[:-1] + tuple(shp))

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


def _copy_space(s, system):
    """Copy of a 1D space on ``system``; device operators are shared, not copied."""
    if isinstance(s, DirectSum):
        return s.with_system(system)
    c = copy.copy(s)
    c._ops = s._ops
    c.system = system
    if hasattr(c, "orthogonal") and c.orthogonal is not s:
        c.orthogonal = copy.copy(s.orthogonal)
        c.orthogonal._ops = s.orthogonal._ops
        c.orthogonal.system = c.system
    elif getattr(s, "orthogonal", None) is s:
        c.orthogonal = c
    return c


def TensorProduct(*basespaces, system: CartCoordSys | None = None, name: str = "T",
                  precision: str = "f64"):
    """Factory (tensorproductspace.py:507-557): copies the 1D spaces and
    assigns each the sub-system of its axis; a product with a DirectSum
    factor is a :class:`DirectSumTPS`."""
    system = default_system(len(basespaces)) if system is None else system
    spaces = [_copy_space(s, system.sub_system(i)) for i, s in enumerate(basespaces)]
    if any(isinstance(s, DirectSum) for s in spaces):
        return DirectSumTPS(spaces, system, name, precision)
    return TensorProductSpace(spaces, system, name, precision)


# ----------------------------------------------------------------------------- lifting
def _project_1d(expr, V) -> torch.Tensor:
    """Coefficients of a SymPy expression of V's coordinate: V.forward of its
    values on V's quadrature mesh (project1D, galerkin/inner.py:1027-1047)."""
    e = sp.sympify(expr)
    pts = V.mesh().cpu().numpy()
    if e.free_symbols:
        vals = lambdify(V.system.base_scalars(), e)(pts)
    else:
        vals = complex(e) if e.has(sp.I) else float(e)
    vals = np.asarray(vals)
    dt = np.complex128 if np.iscomplexobj(vals) else np.float64
    vals = np.broadcast_to(vals.astype(dt), pts.shape).copy()
    return V.forward(torch.tensor(vals, device=K.device()))


def _lifting_bcs(bcspace: BCGeneric, values) -> BCGeneric:
    """Copy of a lifting space with new ordered values (already in
    reference-domain units)."""
    b = copy.copy(bcspace)
    names = bcspace.bcs.orderednames()
    bcs = {"left": {}, "right": {}}
    for nm, v in zip(names, values):
        bcs["left" if nm[0] == "L" else "right"][nm[1:]] = v
    b.bcs = BoundaryConditions(bcs)
    return b


class DirectSumTPS(TensorProductSpace):
    """Tensor product with DirectSum factors (galerkin/tensorproductspace.py:560-851):
    2D with one or two, 3D with one or two (and a plain or homogeneous third axis).

    The trial space splits into the tensor products of every combination of
    homogeneous and lifting summands (``tpspaces``).  Each product with
    lifting factors on the axes L carries a fixed coefficient array
    (``bndvals``): for every choice of boundary condition on the axes in L,
    the boundary function of the first axis in L, with the chosen conditions
    of the other axes in L applied to it (its value, or its k-th derivative
    for an Nk condition, at that end), projected onto the remaining axes'
    spaces — where a remaining DirectSum axis is itself lifted by that
    function's own boundary data (for one remaining axis a 1D DirectSum, for
    two a 2D DirectSumTPS).
    In 2D this is: boundary functions of the other coordinate projected onto
    the other axis (lifted by their corner values when both axes are lifted),
    and the four corner values for the product of both lifting spaces.
    Their sum, padded into the orthogonal basis, is a constant lift
    (``lift_coefficients``) that ``to_orthogonal`` adds and ``from_orthogonal``
    removes; ``backward``/``evaluate_mesh``/``backward_primitive`` add its
    values on the requested mesh (cached).  Coefficient arrays hold the
    homogeneous part only and may carry leading batch dimensions.
    """

    def __init__(self, basespaces, system: CartCoordSys | None = None, name: str = "DSTPS",
                 precision: str = "f64"):
        super().__init__(basespaces, system, name, precision)
        if precision != "f64":
            raise ValueError("inhomogeneous (lifted) tensor-product spaces transform in f64 only")
        d = len(self.basespaces)
        nds = sum(isinstance(s, DirectSum) for s in self.basespaces)
        if d > 3 or nds > 2:
            raise NotImplementedError("DirectSumTPS is built for 2D and 3D products with at most two DirectSums")
        self._ops: dict = {}
        self.tpspaces: dict[tuple, TensorProductSpace] = {}
        choices = [list(s.basespaces) if isinstance(s, DirectSum) else [s] for s in self.basespaces]
        for i, combo in enumerate(itertools.product(*choices)):
            self.tpspaces[combo] = TensorProductSpace(combo, self.system, f"{name}{i}", precision)
        self.bndvals: dict[tuple, torch.Tensor] = {}
        for combo, tps in self.tpspaces.items():
            L = [a for a, s in enumerate(combo) if isinstance(s, BCGeneric)]
            if not L:
                continue
            arr = self._boundary_coefficients(combo, L).contiguous()
            self.bndvals[combo] = arr
            tps.bnd_coefficients = arr
        self.orthogonal = self.get_orthogonal()

    def _boundary_coefficients(self, combo, L) -> torch.Tensor:
        d = len(self.basespaces)
        R = [a for a in range(d) if a not in L]
        syms = [s.system.base_scalars()[0] for s in self.basespaces]
        ends = [(s.domain.lower, s.domain.upper) for s in self.basespaces]
        nb = [combo[a].dim for a in L]
        blocks = {}
        for js in itertools.product(*[range(n) for n in nb]):
            g = sp.sympify(combo[L[0]].bcs.orderedvals()[js[0]])
            for a, j in zip(L[1:], js[1:]):
                g = self._apply_condition(g, a, combo[a].bcs.orderednames()[j], syms, ends)
            blocks[js] = self._project_remaining(g, combo, R, syms, ends)
        cplx = any(isinstance(b, complex) or (isinstance(b, torch.Tensor) and b.is_complex())
                   for b in blocks.values())
        shape = [combo[a].dim for a in range(d)]
        out = torch.zeros(shape, dtype=torch.complex128 if cplx else torch.float64, device=K.device())
        for js, b in blocks.items():
            idx = tuple(js[L.index(a)] if a in L else slice(None) for a in range(d))
            out[idx] = b if isinstance(b, torch.Tensor) else torch.tensor(b, dtype=out.dtype, device=out.device)
        return out

    def _apply_condition(self, g, a, name, syms, ends):
        """Condition ``name`` ("LD", "RN2", ...) of axis a applied to g: the
        value (D) or the k-th derivative (Nk) at that end, in reference-domain
        units (divided by df^k), as galerkin/tensorproductspace.py:636-662
        builds the projected conditions."""
        if name[1:] in ("R", "W"):
            raise NotImplementedError("Robin/weighted conditions on a lifted axis of a tensor product")
        k = bc_order(name[1:])
        z = ends[a][0 if name[0] == "L" else 1]
        e = sp.sympify(g)
        if k:
            df = 2.0 / (ends[a][1] - ends[a][0])
            e = sp.diff(e, syms[a], k) / df**k
        return e.subs(syms[a], z)

    def _project_remaining(self, g, combo, R, syms, ends):
        """Coefficients of g (a function of the coordinates of R) in the
        product of the R axes' spaces, each remaining DirectSum axis lifted by
        g's own boundary data on it (values, or derivatives for Nk conditions)."""
        if not R:
            return _num_or_expr(g)
        spaces = []
        for r in R:
            if isinstance(self.basespaces[r], DirectSum):
                names = self.basespaces[r][1].bcs.orderednames()
                vals = [_num_or_expr(self._apply_condition(g, r, nm, syms, ends)) for nm in names]
                spaces.append(DirectSum(combo[r], _lifting_bcs(self.basespaces[r][1], vals)))
            else:
                spaces.append(combo[r])
        if len(spaces) == 1:
            return _project_1d(g, spaces[0])
        sub = CartCoordSys(self.system.name + "_r", tuple(syms[r] for r in R))
        V = DirectSumTPS(spaces, sub) if any(isinstance(v, DirectSum) for v in spaces) else \
            TensorProductSpace(spaces, sub)
        grids = [v.mesh().cpu().numpy() for v in spaces]
        mesh = np.meshgrid(*grids, indexing="ij")
        e = sp.sympify(g)
        vals = lambdify(tuple(syms[r] for r in R), e)(*mesh) if e.free_symbols else _num_or_expr(e)
        vals = np.broadcast_to(np.asarray(vals), mesh[0].shape)
        dt = torch.complex128 if np.iscomplexobj(vals) else torch.float64
        return V.forward(torch.tensor(np.ascontiguousarray(vals), dtype=dt, device=K.device()))

    # ------------------------------------------------------------ parts
    def get_homogeneous(self) -> TensorProductSpace:
        key = tuple(s[0] if isinstance(s, DirectSum) else s for s in self.basespaces)
        return self.tpspaces[key]

    def get_orthogonal(self) -> TensorProductSpace:
        return TensorProductSpace([s.get_orthogonal() for s in self.basespaces], self.system,
                                  self.name + "o", self.precision)

    @property
    def dim(self) -> int:
        return self.get_homogeneous().dim

    @property
    def num_dofs(self) -> tuple[int, ...]:
        return self.get_homogeneous().num_dofs

    # ------------------------------------------------------------ lift
    def lift_coefficients(self) -> torch.Tensor:
        """(N0, N1): sum over the lifting products of their orthogonal
        coefficients, zero-padded (tensorproductspace.py:806-821)."""
        if "lift" not in self._ops:
            shape = tuple(s.N for s in self.basespaces)
            cplx = any(a.is_complex() for a in self.bndvals.values()) or \
                any(getattr(s, "is_fourier", False) for s in self.basespaces)
            tot = torch.zeros(shape, dtype=torch.complex128 if cplx else torch.float64, device=K.device())
            for combo, arr in self.bndvals.items():
                a = self.tpspaces[combo].to_orthogonal(arr)
                tot[tuple(slice(0, n) for n in a.shape)] += a
            self._ops["lift"] = tot.contiguous()
        return self._ops["lift"]

    def lift_values(self, N=None, k=None, kind: str = MeshKind.QUADRATURE) -> torch.Tensor:
        N = tuple(s.num_quad_points if N is None or N[i] is None else N[i]
                  for i, s in enumerate(self.basespaces))
        k = (0,) * len(self.basespaces) if k is None else tuple(k)
        key = ("liftv", N, k, kind)
        if key not in self._ops:
            ops = [s.orthogonal.backward_operator(N[i], k[i], kind) for i, s in enumerate(self.basespaces)]
            X = self.lift_coefficients()
            for ax, op in enumerate(ops):
                X = K.apply_along(op, X.contiguous(), ax, "f64")
            self._ops[key] = X.contiguous()
        return self._ops[key]

    @staticmethod
    def _f64(precision) -> None:
        if precision not in (None, "f64"):
            raise ValueError("inhomogeneous (lifted) tensor-product spaces transform in f64 only")

    # ------------------------------------------------------------ transforms
    def to_orthogonal(self, c):
        return K.add_bcast(self.get_homogeneous().to_orthogonal(c).contiguous(), self.lift_coefficients())

    def from_orthogonal(self, a):
        a = as_device(a)
        return self.get_homogeneous().from_orthogonal(K.add_bcast(a, self.lift_coefficients(), -1.0))

    def backward(self, c, N=None, precision: str | None = None):
        """orthogonal.backward(to_orthogonal(c)) (tensorproductspace.py:771-776),
        as the homogeneous backward plus the lift's values on the mesh."""
        self._f64(precision)
        u = self.get_homogeneous().backward(c, N).contiguous()
        return K.add_bcast(u, self.lift_values(tuple(u.shape[u.ndim - len(self):])))

    def backward_primitive(self, c, k, N=None, precision: str | None = None):
        self._f64(precision)
        u = self.get_homogeneous().backward_primitive(c, k, N).contiguous()
        return K.add_bcast(u, self.lift_values(tuple(u.shape[u.ndim - len(self):]), k))

    def evaluate_mesh(self, c, kind: str = MeshKind.QUADRATURE, N=None):
        u = self.get_homogeneous().evaluate_mesh(c, kind, N).contiguous()
        return K.add_bcast(u, self.lift_values(tuple(u.shape[u.ndim - len(self):]), None, kind))

    def evaluate(self, x, c):
        """orthogonal.evaluate(x, to_orthogonal(c)) (tensorproductspace.py:786-787)."""
        return self.orthogonal.evaluate(x, self.to_orthogonal(c))

    def forward(self, u, precision: str | None = None):
        """from_orthogonal(orthogonal.forward(u)) (tensorproductspace.py:778-780)."""
        self._f64(precision)
        return self.from_orthogonal(self.orthogonal.forward(u))

    def scalar_product(self, u, precision: str | None = None):
        raise RuntimeError("Scalar product requires homogeneous test space (call on get_homogeneous())")


def _num_or_expr(e):
    e = sp.sympify(e)
    if e.free_symbols:
        return e
    return complex(e) if e.has(sp.I) else float(e)


__all__ = ["TensorProductSpace", "TensorProduct", "DirectSumTPS", "torch"]
