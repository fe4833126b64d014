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
from .arguments import ArgumentTag, JAXFunction, functionspace_of, get_arg
from .composite import BCGeneric, Composite
from .orthogonal import MeshKind
from .tensorproductspace import TensorProductSpace


class InnerKind(str, enum.Enum):
    BILINEAR = "bilinear"
    LINEAR = "linear"
    SYSTEM = "system"


# ----------------------------------------------------------------------------- splitting
class _Term:
    """One additive term: test derivative orders, trial derivative orders,
    coefficient, and a JAXFunction factor that appears linearly (``jf``, with
    derivative orders ``dj``) in a term without a trial function."""

    def __init__(self, test, trial, dv, du, coeff, jf=None, dj=None):
        self.test, self.trial, self.dv, self.du, self.coeff = test, trial, dv, du, coeff
        self.jf, self.dj = jf, (dj or {})


def _has_jaxfunction(e) -> bool:
    return bool(sp.sympify(e).atoms(JAXFunction))


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
        jfs = []
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
            elif tag == ArgumentTag.JAXFUNC:
                jfs.append(f)
            else:
                coeff = coeff * f
        if test is None:
            raise AssertionError("No TestFunction found in expression")
        if trial is None and len(jfs) == 1 and not _has_jaxfunction(coeff):
            # linear in a function with known coefficients: A @ array
            f = jfs[0]
            terms.append(_Term(test, None, dv, {}, coeff, f.expr if isinstance(f, sp.Derivative) else f,
                               _deriv_orders(f)))
            continue
        for f in jfs:
            coeff = coeff * f
        terms.append(_Term(test, trial, dv, du, coeff))
    return terms


def _separate(coeff, symbols):
    """{axis: factor, 'coeff': scalar} or None when not separable (a
    coefficient holding a JAXFunction is evaluated on the mesh)."""
    if _has_jaxfunction(coeff):
        return None
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


def bilinear_1d(v, i: int, u, j: int, coeff, Nq: int, use_precomputed: bool, weight=None):
    """One separated factor (galerkin/inner.py:809-911): matrix (test row,
    trial column) of ``(psi^(i), c psi^(j))`` with stencils applied;
    ``weight`` (values at the Nq quadrature points) replaces c."""
    scale = 1.0
    if weight is None and not coeff.free_symbols:
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
    if weight is not None:
        if weight.is_complex():
            raise NotImplementedError("complex coefficients on a polynomial axis")
        wq = wq * weight
    elif coeff.free_symbols:
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


def _jaxfunction_values(f, symbols, Nq) -> np.ndarray:
    """Values (or derivative values) of a JAXFunction on the quadrature mesh of
    the form's test spaces (its space must share that mesh)."""
    orders = _deriv_orders(f)
    jf = f.expr if isinstance(f, sp.Derivative) else f
    V = jf.functionspace
    vsyms = tuple(V.system.base_scalars())
    k = [orders.get(s_, 0) for s_ in vsyms]
    if len(vsyms) == 1:
        vals = V.backward_primitive(jf.array, k[0], Nq[0])
    else:
        vals = V.backward_primitive(jf.array, tuple(k), tuple(Nq))
    return vals.cpu().numpy()


def _mesh_values(spaces, symbols, expr, Nq) -> torch.Tensor:
    grids = [s.mesh(MeshKind.QUADRATURE, Nq[i]).cpu().numpy() for i, s in enumerate(spaces)]
    mesh = np.meshgrid(*grids, indexing="ij")
    expr = sp.sympify(expr)
    extra_syms, extra_vals = [], []
    if _has_jaxfunction(expr):
        atoms = [a for a in expr.atoms(sp.Derivative) if isinstance(a.expr, JAXFunction)]
        atoms += list(expr.atoms(JAXFunction))
        for a in atoms:
            z = sp.Dummy()
            expr = expr.xreplace({a: z})
            extra_syms.append(z)
            extra_vals.append(_jaxfunction_values(a, symbols, Nq))
    if expr.free_symbols:
        vals = lambdify(tuple(symbols) + tuple(extra_syms), expr)(*mesh, *extra_vals)
    else:
        vals = complex(expr) if expr.has(sp.I) else float(expr)
    vals = np.asarray(vals)
    dt = np.complex128 if np.iscomplexobj(vals) else np.float64
    vals = np.broadcast_to(vals.astype(dt), mesh[0].shape).copy()
    if dt == np.complex128 and np.abs(vals.imag).max() == 0:
        vals = vals.real.copy()
        dt = np.float64
    return torch.tensor(vals, dtype=torch.complex128 if dt == np.complex128 else torch.float64,
                        device=K.device())


# ----------------------------------------------------------------------------- variable coefficients
class TensorMatrix:
    """Matrix-free non-separable bilinear operator (la/tensormatrix.py:15-93).

    Each term applies ``T_0^(i0) (c * (B_0^(j0) U B_1^(j1)^T)) T_1^(i1)^T``:
    backward contractions to the quadrature mesh, a pointwise product with the
    coefficient sampled there, and test-space contractions back.  Separable
    terms of the same form (``separable``, TPMatrix objects) are added to the
    result.  ``solve`` assembles the dense (n0 n1 x n0 n1) matrix on the
    device by applying the operator to the identity batch and factors it
    (la/tensormatrix.py:76-84).
    """

    def __init__(self, terms, shape, separable=()):
        self.terms = list(terms)  # list of (coef_mesh, [(B_op, T_op) per axis])
        self.shape_dofs = tuple(shape)
        self.separable = list(separable)
        self._lu = None

    @property
    def shape(self) -> tuple[int, int]:
        n = int(np.prod(self.shape_dofs))
        return (n, n)

    def __call__(self, U, precision: str = "f64"):
        from .orthogonal import as_device

        U = as_device(getattr(U, "array", U))
        nb = U.ndim - len(self.shape_dofs)
        out = None
        for coef, ops in self.terms:
            X = U
            for ax, (B, _) in enumerate(ops):
                X = K.apply_along(B, X.contiguous(), nb + ax, "f64")
            X = pointwise_mul(X.contiguous(), coef)
            for ax, (_, T) in enumerate(ops):
                X = K.apply_along(T, X.contiguous(), nb + ax, "f64")
            out = X if out is None else K.lincomb([(1.0, out), (1.0, X)])
        for tp in self.separable:
            X = tp(U)
            out = X if out is None else K.lincomb([(1.0, out), (1.0, X)])
        return out

    __matmul__ = __call__

    def __add__(self, other):
        from ..la import TPMatrices, TPMatrix

        if isinstance(other, TensorMatrix):
            return TensorMatrix(self.terms + other.terms, self.shape_dofs, self.separable + other.separable)
        if isinstance(other, TPMatrix):
            return TensorMatrix(self.terms, self.shape_dofs, self.separable + [other])
        if isinstance(other, TPMatrices):
            return TensorMatrix(self.terms, self.shape_dofs, self.separable + list(other.tpmats))
        return NotImplemented

    __radd__ = __add__

    def todense(self, chunk: int = 4096) -> torch.Tensor:
        """The (n0 n1 x n0 n1) matrix, column j = the operator applied to e_j
        (unit vectors applied ``chunk`` at a time as one batch)."""
        n = self.shape[0]
        cols = []
        for j0 in range(0, n, chunk):
            m = min(chunk, n - j0)
            E = torch.zeros((m, n), dtype=torch.float64, device=K.device())
            E[torch.arange(m, device=E.device), torch.arange(j0, j0 + m, device=E.device)] = 1.0
            cols.append(self(E.reshape((m,) + self.shape_dofs)).reshape(m, n))
        return torch.cat(cols, 0).T.contiguous()

    def todense_host(self) -> np.ndarray:
        return self.todense().cpu().numpy()

    @property
    def data(self) -> np.ndarray:
        """Dense (test_0, test_1, trial_0, trial_1, ...) array (la/tensormatrix.py)."""
        return self.todense_host().reshape(self.shape_dofs + self.shape_dofs)

    def solve(self, rhs, axis: int = 0):
        """Dense LU solve (la/tensormatrix.py:76-80) by the library's blocked LU
        with partial pivoting."""
        from .orthogonal import as_device

        if self._lu is None:
            self._lu = K.dense_lu_factor(self.todense())
        b = as_device(rhs)
        LU, perm = self._lu
        flat = b.reshape((-1, self.shape[0])).T.contiguous()
        x = K.dense_lu_solve(LU, perm, flat)
        return x.T.reshape(b.shape).contiguous()


def pointwise_mul(X: torch.Tensor, C: torch.Tensor) -> torch.Tensor:
    from ..kernels import mul_bcast

    return mul_bcast(X, C)


# ----------------------------------------------------------------------------- inner
def _jaxfunction_parts(jf):
    """(space, fixed coefficients) pairs that make up a JAXFunction: the
    homogeneous part with its array, and for a lifted space every lifting part
    with its boundary coefficients (the expansion of JAXFunction.doit over a
    DirectSum / DirectSumTPS, galerkin/arguments.py:785-800)."""
    from .composite import DirectSum
    from .tensorproductspace import DirectSumTPS

    V = jf.functionspace
    if isinstance(V, DirectSum):
        return [(V[0], jf.array), (V[1], torch.tensor(V[1].bnd_vals(), device=K.device()))]
    if isinstance(V, DirectSumTPS):
        out = [(V.get_homogeneous(), jf.array)]
        out += [(V.tpspaces[f], arr) for f, arr in V.bndvals.items()]
        return out
    return [(V, jf.array)]


def _lift_of(space):
    """Fixed coefficients of a lifting trial space (1D BCGeneric: its boundary
    values; a tensor product with a lifting factor: its ``bndvals``), else None."""
    if isinstance(space, BCGeneric):
        return torch.tensor(space.bnd_vals(), device=K.device())
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
            if sep is None and d == 1:
                # one axis: the coefficient's values weight the quadrature
                wq = _mesh_values(spaces, symbols, t.coeff, Nq)
                m = bilinear_1d(spaces[0], t.dv.get(symbols[0], 0), uspaces[0], t.du.get(symbols[0], 0),
                                sp.S.One, Nq[0], False, weight=wq)
                if lift is not None:
                    bvec = accumulate(bvec, -1.0, m.matvec(lift.contiguous(), 0).contiguous())
                    continue
                amats.append((1.0, [m]))
                continue
            if sep is None:
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
        elif t.jf is not None:
            # A @ array for the linear JAXFunction (its lifting parts included)
            sign = -1.0 if has_bilinear else 1.0
            sep = _separate(t.coeff, symbols)
            for P, arr in _jaxfunction_parts(t.jf):
                pspaces = P.basespaces if isinstance(P, TensorProductSpace) else [P]
                if sep is None:
                    coef = _mesh_values(spaces, symbols, t.coeff, Nq)
                    ops = [(pspaces[ax].backward_operator(Nq[ax], t.dj.get(symbols[ax], 0)),
                            test_operator(spaces[ax], t.dv.get(symbols[ax], 0), Nq[ax])) for ax in range(d)]
                    X = TensorMatrix([(coef, ops)], tuple(s_.dim for s_ in spaces))(arr)
                    bvec = accumulate(bvec, sign, X.contiguous())
                    continue
                c = complex(sep.get("coeff", 1))
                X = arr
                for ax in range(d):
                    s_ = symbols[ax]
                    m = bilinear_1d(spaces[ax], t.dv.get(s_, 0), pspaces[ax], t.dj.get(s_, 0),
                                    sp.sympify(sep.get(s_, sp.S.One)), Nq[ax], use_precomputed_matrices)
                    X = m.matvec(X.contiguous(), ax)
                if c.imag:
                    raise NotImplementedError("complex scalar coefficients")
                bvec = accumulate(bvec, sign * c.real, X.contiguous())
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
        tps = [] if A is None else (list(A.tpmats) if isinstance(A, TPMatrices) else [A])
        A = TensorMatrix(multivar, tuple(s.dim for s in spaces), tps)

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


def project1D(ue, V):
    """Project a SymPy expression of one coordinate onto a 1D space
    (galerkin/inner.py:1027-1055): the forward transform of its mesh values."""
    if isinstance(V, TensorProductSpace):
        raise ValueError("project1D takes a 1D space")
    return project(ue, V)
