"""Galerkin assembly ``inner()`` (reference: galerkin/inner.py).

The SymPy front end (splitting a weak form into separated per-axis factors,
galerkin/forms.py:241-325) runs on the host once per call.  Every array
operation runs on the device: Vandermonde matrices and Gauss nodes come from
the HIP basis kernels, quadrature matrices ``(P_i^T w) P_j`` and load vectors
``(S (x) S)(V^T W f W V)`` are MFMA contractions.  Closed-form matrices
(``use_precomputed_matrices=True``) are the reference's formulas.
"""

from __future__ import annotations

import enum

import numpy as np
import sympy as sp
import torch

from .. import kernels as K
from ..la import DiaMatrix, Matrix, TPMatrices, TPMatrix
from ..utils.common import lambdify
from .arguments import ArgumentTag, functionspace_of, get_arg
from .composite import BCGeneric, Composite
from .orthogonal import MeshKind
from .tensorproductspace import TensorProductSpace


class InnerKind(str, enum.Enum):
    BILINEAR = "bilinear"
    LINEAR = "linear"
    SYSTEM = "system"


# ----------------------------------------------------------------------------- splitting
class _Term:
    """One additive term: test derivative orders, trial derivative orders, coefficient."""

    def __init__(self, test, trial, dv, du, coeff):
        self.test, self.trial, self.dv, self.du, self.coeff = test, trial, dv, du, coeff


def _deriv_orders(f) -> dict:
    if isinstance(f, sp.Derivative):
        out: dict = {}
        for s, c in f.variable_count:
            out[s] = out.get(s, 0) + int(c)
        return out
    return {}


def split_terms(expr) -> list[_Term]:
    """Split a weak form into additive terms (galerkin/forms.py:241-325)."""
    expr = sp.expand(sp.sympify(expr).doit())
    terms = []
    for term in sp.Add.make_args(expr):
        test = trial = None
        dv, du = {}, {}
        coeff = sp.S.One
        for f in sp.Mul.make_args(term):
            tag = get_arg(f)
            if tag == ArgumentTag.TEST:
                if test is not None:
                    raise ValueError("more than one test function in a term")
                test, dv = f, _deriv_orders(f)
            elif tag == ArgumentTag.TRIAL:
                if trial is not None:
                    raise NotImplementedError("nonlinear in the trial function")
                trial, du = f, _deriv_orders(f)
            else:
                coeff = coeff * f
        if test is None:
            raise AssertionError("No TestFunction found in expression")
        terms.append(_Term(test, trial, dv, du, coeff))
    return terms


def _separate(coeff, symbols):
    """{axis: factor, 'coeff': scalar} or None when not separable."""
    if not coeff.free_symbols:
        return {"coeff": coeff}
    sep = sp.separatevars(coeff, list(symbols), dict=True, force=True)
    if sep is None:
        return None
    return sep


# ----------------------------------------------------------------------------- 1D pieces
def _quad_nodes(space, N):
    return space.orthogonal.quad_points_and_weights(N)


def _stencil_apply(space, z: torch.Tensor, axis: int) -> torch.Tensor:
    S = space._stencil_dev() if not space.is_orthogonal else None
    return z if S is None else K.apply_along(S, z.contiguous(), axis, "f64")


def bilinear_1d(v, i: int, u, j: int, coeff, Nq: int, use_precomputed: bool):
    """One separated factor (galerkin/inner.py:809-911): matrix (test row,
    trial column) of ``(psi^(i), c psi^(j))`` with stencils applied."""
    scale = 1.0
    if not coeff.free_symbols:
        scale = float(coeff)
        if use_precomputed:
            if isinstance(v, Composite):
                z = v.matrices(i, (u, j), 0, scale)
                if z is not None:
                    return z
            zo = v.orthogonal.matrices(i, (u.orthogonal, j), 0, scale)
            if zo is not None:
                Sv = v.stencil_matrix()
                Su = u.stencil_matrix()
                A = zo.todense_host()
                if Sv is not None:
                    A = Sv @ A
                if Su is not None:
                    A = A @ Su.T
                return type(zo)._combine(zo, A)
    if getattr(v, "is_fourier", False) or getattr(u, "is_fourier", False):
        raise NotImplementedError("Fourier-axis matrices are built from their closed forms only "
                                  "(constant coefficients, use_precomputed_matrices=True)")
    x, w = _quad_nodes(v, Nq)
    df = float(v.orthogonal.domain_factor)
    wq = w * (df ** (i + j - 1))
    if coeff.free_symbols:
        wq = wq * v.eval_coefficient(coeff, Nq)
    else:
        wq = wq * scale
    Pi = v.orthogonal.vandermonde(x, i)
    Pj = u.orthogonal.vandermonde(x, j)
    PiTw = K.scale_transpose(Pi.contiguous(), wq.contiguous())            # (Ni, Q)
    z = K.apply_along(K.AxisOperator(PiTw), Pj.contiguous(), 0, "f64")   # (Ni, Nj)
    z = _stencil_apply(v, z, 0)
    z = _stencil_apply(u, z, 1)
    return Matrix(z)


def test_operator(v, i: int, Nq: int) -> K.AxisOperator:
    """(dim, Q): ``b_k = sum_q w_q df^(i-1) f(x_q) phi_k^(i)(x_q)`` (inner.py:914-984);
    complex conjugated test functions on a Fourier axis."""
    if getattr(v, "is_fourier", False):
        return v.test_operator(i, Nq)
    key = ("test", Nq, i)
    if key not in v._ops:
        x, w = _quad_nodes(v, Nq)
        df = float(v.orthogonal.domain_factor)
        Pi = v.orthogonal.vandermonde(x, i)
        T = K.scale_transpose(Pi.contiguous(), w.contiguous(), None, df ** (i - 1))
        v._ops[key] = K.AxisOperator(_stencil_apply(v, T, 0).contiguous())
    return v._ops[key]


def _mesh_values(spaces, symbols, expr, Nq) -> torch.Tensor:
    grids = [s.mesh(MeshKind.QUADRATURE, Nq[i]).cpu().numpy() for i, s in enumerate(spaces)]
    mesh = np.meshgrid(*grids, indexing="ij")
    if expr.free_symbols:
        vals = lambdify(tuple(symbols), expr)(*mesh)
    else:
        vals = float(expr)
    vals = np.broadcast_to(np.asarray(vals, dtype=np.float64), mesh[0].shape).copy()
    return torch.tensor(vals, dtype=torch.float64, device=K.device())


# ----------------------------------------------------------------------------- variable coefficients
class TensorMatrix:
    """Matrix-free non-separable bilinear operator (la/tensormatrix.py:15-93).

    Each term applies ``T_0^(i0) (c * (B_0^(j0) U B_1^(j1)^T)) T_1^(i1)^T``:
    backward contractions to the quadrature mesh, a pointwise product with the
    coefficient sampled there, and test-space contractions back.
    """

    def __init__(self, terms, shape):
        self.terms = terms  # list of (coef_mesh, [(B_op, T_op) per axis])
        self.shape_dofs = shape

    def __call__(self, U, precision: str = "f64"):
        from .orthogonal import as_device

        U = as_device(U)
        nb = U.ndim - len(self.shape_dofs)
        out = None
        for coef, ops in self.terms:
            X = U
            for ax, (B, _) in enumerate(ops):
                X = K.apply_along(B, X, nb + ax, "f64")
            X = pointwise_mul(X, coef)
            for ax, (_, T) in enumerate(ops):
                X = K.apply_along(T, X, nb + ax, "f64")
            out = X if out is None else K.lincomb([(1.0, out), (1.0, X)])
        return out

    __matmul__ = __call__

    def __add__(self, other):
        return TensorMatrix(self.terms + other.terms, self.shape_dofs)


def pointwise_mul(X: torch.Tensor, C: torch.Tensor) -> torch.Tensor:
    from ..varcoeff import mul_bcast

    return mul_bcast(X, C)


# ----------------------------------------------------------------------------- inner
def _lift_of(space):
    """Fixed coefficients of a lifting trial space (1D BCGeneric: its boundary
    values; a tensor product with a lifting factor: its ``bndvals``), else None."""
    if isinstance(space, BCGeneric):
        return torch.tensor(space.bnd_vals(), dtype=torch.float64, device=K.device())
    return getattr(space, "bnd_coefficients", None)


def _coerce_kind(kind):
    if kind is None or isinstance(kind, InnerKind):
        return kind
    try:
        return InnerKind(kind)
    except ValueError as e:
        raise ValueError(f"'{kind}' is not a valid InnerKind") from e


def inner(expr, sparse: bool = False, sparse_tol: int = 1000, num_quad_points=None,
          use_precomputed_matrices: bool = True, *, kind=None):
    """Assemble bilinear and/or linear forms (galerkin/inner.py:134-201).

    Returns a Matrix/DiaMatrix (1D) or TPMatrix/TPMatrices/TensorMatrix (>1D)
    for a bilinear form, a device array for a linear form, or ``(A, b)``.
    In a system, linear terms enter ``b`` with a minus sign (``a(u,v) - L(v)``,
    galerkin/inner.py:304-305).
    """
    kind = _coerce_kind(kind)
    terms = split_terms(expr)
    V = functionspace_of(terms[0].test)
    spaces = V.basespaces if isinstance(V, TensorProductSpace) else [V]
    symbols = V.system.base_scalars()
    d = len(spaces)
    if num_quad_points is None:
        Nq = tuple(s.num_quad_points for s in spaces)
    elif isinstance(num_quad_points, int):
        Nq = (num_quad_points,) * d
    else:
        Nq = tuple(num_quad_points[i] if num_quad_points[i] is not None else spaces[i].num_quad_points
                   for i in range(d))
    has_bilinear = any(t.trial is not None for t in terms)
    Uspace = None
    amats = []       # list of (scale, [1D mats]) separable terms
    multivar = []    # TensorMatrix terms
    bvec = None

    def accumulate(b, c, X):
        return K.lincomb([(c, X)]) if b is None else K.lincomb([(1.0, b), (c, X)])

    for t in terms:
        if t.trial is not None:
            Uspace = functionspace_of(t.trial)
            uspaces = Uspace.basespaces if isinstance(Uspace, TensorProductSpace) else [Uspace]
            # a lifting part of a DirectSum / DirectSumTPS trial function has
            # fixed coefficients: its term moves to the right-hand side with the
            # sign of a linear form in a system (galerkin/inner.py:362-428,
            # 493-562, 445-490)
            lift = _lift_of(Uspace)
            sep = _separate(t.coeff, symbols)
            if sep is None:
                if d != 2:
                    raise NotImplementedError("non-separable coefficients are built for 2D spaces")
                coef = _mesh_values(spaces, symbols, t.coeff, Nq)
                ops = []
                for ax in range(d):
                    i = t.dv.get(symbols[ax], 0)
                    j = t.du.get(symbols[ax], 0)
                    ops.append((uspaces[ax].backward_operator(Nq[ax], j), test_operator(spaces[ax], i, Nq[ax])))
                if lift is not None:
                    X = TensorMatrix([(coef, ops)], tuple(s.dim for s in spaces))(lift)
                    bvec = accumulate(bvec, -1.0, X.contiguous())
                    continue
                multivar.append((coef, ops))
                continue
            c = float(sep.get("coeff", 1))
            mats = []
            for ax in range(d):
                s = symbols[ax]
                i = t.dv.get(s, 0)
                j = t.du.get(s, 0)
                cf = sp.sympify(sep.get(s, sp.S.One))
                mats.append(bilinear_1d(spaces[ax], i, uspaces[ax], j, cf, Nq[ax], use_precomputed_matrices))
            if lift is not None:
                X = lift
                for ax, m in enumerate(mats):
                    X = m.matvec(X.contiguous(), ax)
                bvec = accumulate(bvec, -c, X.contiguous())
                continue
            amats.append((c, mats))
        else:
            coef = _mesh_values(spaces, symbols, t.coeff, Nq)
            X = coef
            for ax in range(d):
                i = t.dv.get(symbols[ax], 0)
                X = K.apply_along(test_operator(spaces[ax], i, Nq[ax]), X.contiguous(), ax, "f64")
            sign = -1.0 if has_bilinear else 1.0
            bvec = accumulate(bvec, sign, X.contiguous())

    A = None
    if amats:
        if d == 1:
            tot = None
            for c, mats in amats:
                m = mats[0] * c
                m = m.tosparse(sparse_tol) if sparse else m.to_matrix()
                tot = m if tot is None else tot + m
            A = tot
        else:
            tps = []
            for c, mats in amats:
                mats = [m.tosparse(sparse_tol) if sparse else m.to_matrix() for m in mats]
                tps.append(TPMatrix(mats, c))
            A = tps[0] if len(tps) == 1 else TPMatrices(tps)
    if multivar:
        TM = TensorMatrix(multivar, tuple(s.dim for s in spaces))
        if A is not None:
            raise NotImplementedError("mixing separable and non-separable bilinear terms")
        A = TM

    if A is None and bvec is None:
        raise ValueError("No bilinear or linear forms found in expression")
    if A is None:
        result, actual = bvec, InnerKind.LINEAR
    elif bvec is None:
        result, actual = A, InnerKind.BILINEAR
    else:
        result, actual = (A, bvec), InnerKind.SYSTEM
    if kind is not None and kind != actual:
        raise ValueError(f"inner(..., kind={kind!r}) assembled {actual.value}; expected {kind}")
    return result


def project(ue, V):
    """Project a SymPy expression onto V: forward transform of its mesh values
    (galerkin/inner.py:1057-1117)."""
    spaces = V.basespaces if isinstance(V, TensorProductSpace) else [V]
    vals = _mesh_values(spaces, V.system.base_scalars(), sp.sympify(ue), tuple(s.num_quad_points for s in spaces))
    return V.forward(vals)
