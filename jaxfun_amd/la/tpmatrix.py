"""Kronecker-product operators and their solvers (reference: la/tpmatrix.py).

``TPMatrix`` applies ``scale * A0 (x) A1 [(x) A2]`` as one contraction per
axis; ``TPMatrices`` sums such terms.  ``TPMatrices.solve`` follows the
reference's dispatch (la/tpmatrix.py:386-509): all-dense factors use the dense
Kronecker LU, otherwise fast diagonalisation.  The per-axis generalised
eigenproblem is solved once on the host in FP64 (as the reference does with
``scipy.linalg.eigh``, la/tpmatrix.py:1160-1204); non-symmetric pairs (the
Chebyshev-Dirichlet stiffness) use ``scipy.linalg.eig`` with the exact
left transform ``W = (B V)^-1`` instead of ``V^T``, so the solve stays exact.
The solve itself is four (2D) or six (3D) device contractions and one
broadcast division.
"""

from __future__ import annotations

import warnings

import numpy as np
import scipy.linalg as sla
import scipy.sparse as ssp
import torch

from .. import kernels as K
from .matrix import BaseMatrix, DiaMatrix, Matrix, SolverNotApplicable


def _as_dev(x, dtype=torch.float64):
    from ..galerkin.orthogonal import as_device

    return as_device(x, dtype)


class TPMatrix:
    """Rank-d separable operator ``scale * kron(A0, A1, ...)`` (la/tpmatrix.py:46-220)."""

    is_zero = False

    def __init__(self, mats, scale=1.0, global_indices=(0, 0)):
        self.mats = list(mats)
        self.coefficient = scale
        self.global_indices = global_indices

    @property
    def dims(self) -> int:
        return len(self.mats)

    def __len__(self) -> int:
        return len(self.mats)

    @property
    def shape(self) -> tuple[int, int]:
        return (int(np.prod([m.shape[0] for m in self.mats])), int(np.prod([m.shape[1] for m in self.mats])))

    def scale(self, alpha) -> "TPMatrix":
        return TPMatrix(self.mats, self.coefficient * alpha, self.global_indices)

    def _plane_able(self, w, precision: str) -> bool:
        return self.dims == 2 and not w.is_complex() and \
            K.plane_ok(self.mats[0].operator(), self.mats[1].operator(), precision)

    def _matmul_array(self, w, precision: str = "f64", out=None, accumulate: bool = False) -> torch.Tensor:
        """A0 along axis 0, A1 along axis 1, ... (la/tpmatrix.py:152-156).  In
        2D the coefficient is applied by the fused kernel, which also adds the
        product to ``out`` when ``accumulate`` is set."""
        from .matrix import _coefficients

        w = _as_dev(_coefficients(w), K.io_dtype(precision))
        nb = w.ndim - self.dims
        if self._plane_able(w, precision):
            return K.plane(self.mats[0].operator(), self.mats[1].operator(), w.contiguous(), precision,
                           alpha=float(self.coefficient), out=out, accumulate=accumulate)
        if out is not None or accumulate:
            raise ValueError("out/accumulate need the fused 2D path")
        for i, m in enumerate(self.mats):
            w = m.matvec(w, nb + i, precision)
        if self.coefficient != 1:
            w = K.lincomb([(float(self.coefficient), w)])
        return w

    def __call__(self, u, precision: str = "f64"):
        return self._matmul_array(u, precision)

    __matmul__ = __call__

    def diagonal_or_none(self):
        diags = [m.diagonal_or_none() for m in self.mats]
        if any(d is None for d in diags):
            return None
        D = diags[0]
        for d in diags[1:]:
            D = np.multiply.outer(D, d)
        return D * float(self.coefficient)

    @property
    def is_diagonal(self) -> bool:
        return self.diagonal_or_none() is not None

    def solve(self, rhs):
        """Factor-wise solve (la/tpmatrix.py:178-213)."""
        d = self.diagonal_or_none()
        if d is not None:
            return _solve_diagonal(d, rhs)
        x = _as_dev(rhs).clone()
        nb = x.ndim - self.dims
        for i, m in enumerate(self.mats):
            x = m.solve(x, axis=nb + i)
        return K.lincomb([(1.0 / float(self.coefficient), x)])

    def lu_factor(self) -> "TPLUFactors":
        """Per-factor factorisations of this Kronecker product (la/tpmatrix.py:205-220)."""
        return TPLUFactors(self)

    def todense_host(self) -> np.ndarray:
        return tpmats_to_kron([self])

    def todense(self) -> torch.Tensor:
        return torch.tensor(self.todense_host(), device=K.device())

    def to_matrix(self) -> Matrix:
        return Matrix(self.todense_host())

    def __add__(self, other):
        if isinstance(other, TPMatrix):
            return TPMatrices([self, other])
        if isinstance(other, TPMatrices):
            return TPMatrices([self, *other.tpmats])
        return NotImplemented


class TPLUFactors:
    """Solve of ``scale * A0 (x) A1 (x) ...`` by one factor solve per axis
    (la/tpmatrix.py:223-255); the factors keep their own (banded or dense)
    factorisations."""

    def __init__(self, tp: TPMatrix) -> None:
        self.tp = tp
        self.scale = tp.coefficient
        self.shape = tuple(m.shape[0] for m in tp.mats)

    def solve(self, rhs) -> torch.Tensor:
        return self.tp.solve(rhs)


def _solve_diagonal(d: np.ndarray, rhs) -> torch.Tensor:
    x = _as_dev(rhs)
    if np.iscomplexobj(d):
        Dinv = torch.tensor(np.ascontiguousarray(1.0 / np.asarray(d)), dtype=torch.complex128, device=x.device)
        return K.cmul_bcast(x, Dinv)
    if x.is_complex():
        # the real view divides re and im by the same entry
        D2 = np.repeat(np.asarray(d, dtype=np.float64)[..., None], 2, axis=-1)
        D = torch.tensor(np.ascontiguousarray(D2), dtype=torch.float64, device=x.device)
        return torch.view_as_complex(K.div_bcast(torch.view_as_real(x), D))
    D = torch.tensor(np.ascontiguousarray(d), dtype=torch.float64, device=x.device)
    return K.div_bcast(x, D)


class TPMatricesLUFactors:
    """Fast-diagonalisation solver (la/tpmatrix.py:512-587).

    ``U = V (W F W^T / D) V^T`` (2D; one transform per axis in general) with
    ``D = sum_k s_k prod_axis lambda``.  ``precision`` selects the contraction
    arithmetic of the four/six transforms (the division follows the I/O dtype).
    """

    def __init__(self, eigvecs, lefts, per_term_eigenvalues, scales, shape):
        self.shape = tuple(shape)
        self.V = [K.AxisOperator(torch.tensor(v, device=K.device())) for v in eigvecs]
        self.W = [K.AxisOperator(torch.tensor(w, device=K.device())) for w in lefts]
        D = np.zeros(self.shape)
        for evals, s in zip(per_term_eigenvalues, scales):
            t = np.ones(self.shape)
            for i, ev in enumerate(evals):
                sh = [1] * len(self.shape)
                sh[i] = -1
                t = t * np.reshape(ev, sh)
            D = D + float(s) * t
        self.D_host = D
        self._D = {}
        self.eigvecs_host = [np.asarray(v) for v in eigvecs]
        self.lefts_host = [np.asarray(w) for w in lefts]

    def D(self, dtype) -> torch.Tensor:
        if dtype not in self._D:
            if dtype == torch.complex128:
                # D for the real view of complex data: each entry twice (re, im)
                D2 = np.repeat(self.D_host[..., None], 2, axis=-1)
                self._D[dtype] = torch.tensor(np.ascontiguousarray(D2), dtype=torch.float64, device=K.device())
            else:
                self._D[dtype] = torch.tensor(self.D_host, dtype=dtype, device=K.device())
        return self._D[dtype]

    def solve(self, rhs, precision: str = "f64") -> torch.Tensor:
        dt = K.io_dtype(precision)
        F = _as_dev(rhs, dt)
        nb = F.ndim - len(self.shape)
        if tuple(F.shape[nb:]) != self.shape:
            F = F.reshape(F.shape[:nb] + self.shape) if F.numel() % int(np.prod(self.shape)) == 0 else F
        if F.is_complex():
            for i, W in enumerate(self.W):
                F = K.apply_along(W, F, nb + i, "f64")
            U = torch.view_as_complex(K.div_bcast(torch.view_as_real(F), self.D(torch.complex128)))
            for i, V in enumerate(self.V):
                U = K.apply_along(V, U, nb + i, "f64")
            return U
        nslab = int(np.prod(F.shape[:nb])) if nb else 1
        if len(self.shape) == 2 and nslab <= K.num_cus() and K.plane_fd_ok([*self.W, *self.V], precision):
            # a few slabs: one launch, latency-bound; many slabs: the two-launch
            # path runs two workgroups per CU and streams better
            return K.plane_fd(self.W[0], self.W[1], self.V[0], self.V[1], F.contiguous(), self.D(dt), precision)
        if len(self.shape) == 2 and K.plane_ok(self.W[0], self.W[1], precision):
            F = F.contiguous()
            Ft = K.plane(self.W[0], self.W[1], F, precision, ("div", self.D(dt)))
            return K.plane(self.V[0], self.V[1], Ft, precision)
        for i, W in enumerate(self.W):
            F = K.apply_along(W, F, nb + i, precision)
        U = K.div_bcast(F, self.D(dt))
        for i, V in enumerate(self.V):
            U = K.apply_along(V, U, nb + i, precision)
        return U


class TPMatricesDenseLUFactors:
    """Dense Kronecker LU (la/tpmatrix.py:1016-1049) with partial pivoting,
    factored and solved by the library's blocked LU kernels."""

    def __init__(self, A, shape):
        self.shape = tuple(shape)
        A = A if isinstance(A, torch.Tensor) else torch.tensor(A, device=K.device())
        self.LU, self.perm = K.dense_lu_factor(A)

    def solve(self, rhs, precision: str = "f64") -> torch.Tensor:
        b = _as_dev(rhs)
        nb = b.ndim - len(self.shape)
        flat = b.reshape(b.shape[:nb] + (-1,))
        x = K.dense_lu_solve(self.LU, self.perm, flat.reshape(-1, flat.shape[-1]).T.contiguous())
        return x.T.reshape(b.shape).contiguous()


class TPMatricesBandedLUFactors:
    """LU without pivoting of the assembled Kronecker matrix in band storage
    (the ``"banded"`` and ``"rcm"`` methods of ``DiaMatrix.lu_solve``,
    la/diamatrix.py:592-770): factored by one workgroup per elimination
    step window (``jfx_band_lu_factor_wide``), solved one wave per
    right-hand side.  With ``perm`` the matrix is the reordered
    ``A[perm][:, perm]`` and right-hand sides are permuted on the way."""

    def __init__(self, band: np.ndarray, p: int, q: int, shape, perm: np.ndarray | None = None):
        self.shape = tuple(shape)
        self.n = int(np.prod(self.shape))
        self.p, self.q = int(p), int(q)
        dev = K.device()
        b = torch.tensor(np.ascontiguousarray(band), dtype=torch.float64, device=dev)
        self.lu = torch.empty_like(b)
        st = torch.zeros(1, dtype=torch.int32, device=dev)
        from .. import _lib

        _lib.call("jfx_band_lu_factor_wide", K._p(b), self.n, self.p, self.q, K._p(self.lu), K._p(st), K._s())
        k = int(st.item())
        if k != 0:
            raise ValueError(f"Zero or non-finite pivot at row {k - 1} in banded LU factorization")
        self.perm = self.inv = None
        if perm is not None:
            self.perm = torch.tensor(np.asarray(perm, dtype=np.int64), device=dev)
            inv = np.empty_like(np.asarray(perm, dtype=np.int64))
            inv[np.asarray(perm)] = np.arange(len(perm))
            self.inv = torch.tensor(inv, device=dev)

    def _solve_real(self, X: torch.Tensor) -> torch.Tensor:
        from .. import _lib

        nrhs = X.shape[0]
        if self.perm is not None:
            Y = torch.empty_like(X)
            _lib.call("jfx_gather_rows", K._p(X), K._p(self.perm), self.n, nrhs, K._p(Y), K._s())
            X = Y
        _lib.call("jfx_band_lu_solve_wide", K._p(self.lu), self.n, self.p, self.q, K._p(X), nrhs, K._s())
        if self.inv is not None:
            Y = torch.empty_like(X)
            _lib.call("jfx_gather_rows", K._p(X), K._p(self.inv), self.n, nrhs, K._p(Y), K._s())
            X = Y
        return X

    def solve(self, rhs, precision: str = "f64") -> torch.Tensor:
        b = _as_dev(rhs)
        if b.is_complex():
            bc = b.contiguous()
            re = self.solve(torch.view_as_real(bc)[..., 0].contiguous())
            im = self.solve(torch.view_as_real(bc)[..., 1].contiguous())
            return torch.complex(re, im)
        X = b.reshape(-1, self.n).clone()  # solved in place
        return self._solve_real(X).reshape(b.shape)


def _sparse_kron(tpmats) -> ssp.csr_matrix:
    """Sum of Kronecker products of the (banded) factors as a sparse matrix
    (the DiaMatrix that tpmats_to_kron forms for sparse factors,
    la/tpmatrix.py:1357-1402)."""
    out = None
    for tp in tpmats:
        Kd = ssp.csr_matrix(tp.mats[0].todense_host())
        for m in tp.mats[1:]:
            Kd = ssp.kron(Kd, ssp.csr_matrix(m.todense_host()), format="csr")
        Kd = complex(tp.coefficient).real * Kd if not np.iscomplexobj(Kd.data) else tp.coefficient * Kd
        out = Kd if out is None else out + Kd
    out = ssp.csr_matrix(out)
    out.eliminate_zeros()
    return out


def _bandwidths(A: ssp.spmatrix) -> tuple[int, int]:
    C = A.tocoo()
    if C.nnz == 0:
        return 0, 0
    d = C.col.astype(np.int64) - C.row.astype(np.int64)
    return int(max(0, -d.min())), int(max(0, d.max()))


def _band_storage(A: ssp.spmatrix, p: int, q: int) -> np.ndarray:
    C = A.tocoo()
    band = np.zeros((A.shape[0], p + q + 1))
    band[C.row, C.col - C.row + p] = C.data
    return band


class TPMatricesWavenumberSolver:
    """Per-wavenumber banded solver for Fourier x polynomial systems
    (la/tpmatrix.py:590-1011).

    Every axis but ``poly_axis`` is diagonal in every term (Fourier), so for
    each combination of the diagonal axes' indices (flattened in C order,
    ``nsys`` systems) the polynomial-axis operator is the banded matrix
    ``B_k = sum_i s_i (prod_a F_i^(a)[k_a]) P_i``.  All ``B_k`` are factored
    at once on the device (``jfx_band_lu_factor_batched``, no pivoting, one
    lane per system) and a solve is one ``jfx_band_lu_solve_batched`` launch
    along the polynomial axis (the re/im parts of complex data share a factor).
    """

    def __init__(self, poly_axis: int, shape, band: np.ndarray, p: int, q: int):
        self.poly_axis = int(poly_axis)
        self.shape = tuple(shape)
        self.p, self.q = int(p), int(q)
        self.nsys = band.shape[0]
        self.lu = K.band_lu_factor_batched(torch.tensor(np.ascontiguousarray(band), device=K.device()), p, q)

    def solve(self, rhs, precision: str = "f64") -> torch.Tensor:
        x = _as_dev(rhs).clone()
        if self.lu.is_complex() and not x.is_complex():
            x = K.to_complex(x)
        nb = x.ndim - len(self.shape)
        if tuple(x.shape[nb:]) != self.shape:
            raise ValueError(f"rhs shape {tuple(x.shape)} does not end with {self.shape}")
        return K.band_lu_solve_batched(self.lu, self.p, self.q, x, nb + self.poly_axis)


def tpmats_wavenumber_factor(A) -> TPMatricesWavenumberSolver:
    """Factor a Fourier x polynomial sum of Kronecker products per wavenumber
    (la/tpmatrix.py:1236-1354).  Raises SolverNotApplicable unless exactly one
    axis is non-diagonal; complex wavenumber weights (odd derivative orders
    on a Fourier axis) give complex bands and the complex kernels."""
    if isinstance(A, TPMatrices):
        tpmats = list(A.tpmats)
    elif isinstance(A, list):
        tpmats = A
    elif isinstance(A, TPMatrix):
        tpmats = [A]
    else:
        raise TypeError(f"tpmats_wavenumber_factor expects a list[TPMatrix] or TPMatrices, got {type(A).__name__!r}.")
    ndim = tpmats[0].dims
    diag_axes = [a for a in range(ndim) if all(tp.mats[a].diagonal_or_none() is not None for tp in tpmats)]
    poly_axes = [a for a in range(ndim) if a not in diag_axes]
    if len(poly_axes) != 1:
        raise SolverNotApplicable(
            f"tpmats_wavenumber_factor requires exactly 1 polynomial (non-diagonal) axis; "
            f"found {len(poly_axes)}: {poly_axes}. Use tpmats_lu_factor for fully-symmetric problems.")
    pa = poly_axes[0]
    shape = tuple(int(tpmats[0].mats[a].shape[0]) for a in range(ndim))
    n = shape[pa]
    # W[i, k]: scale_i * prod over the diagonal axes (C order) of their diagonals
    W = []
    for tp in tpmats:
        w = np.array([complex(tp.coefficient)])
        for a in diag_axes:
            w = np.outer(w, tp.mats[a].diagonal_or_none()).ravel()
        W.append(w)
    W = np.array(W)
    if np.abs(W.imag).max() == 0:
        W = W.real
    dense = [tp.mats[pa].todense_host() for tp in tpmats]
    nz = np.zeros((n, n), dtype=bool)
    for P in dense:
        nz |= P != 0
    rows, cols = np.nonzero(nz)
    p = int(max(0, (rows - cols).max())) if rows.size else 0
    q = int(max(0, (cols - rows).max())) if rows.size else 0
    Wd = p + q + 1
    bands = np.zeros((len(tpmats), n, Wd), dtype=np.result_type(*dense))
    for t, P in enumerate(dense):
        for o in range(-p, q + 1):
            i = np.arange(max(0, -o), min(n, n - o))
            bands[t, i, o + p] = P[i, i + o]
    band = np.einsum("tk,tnw->knw", W, bands)
    return TPMatricesWavenumberSolver(pa, shape, band, p, q)


def tpmats_to_kron(tpmats) -> np.ndarray:
    """Dense sum of Kronecker products (la/tpmatrix.py:1357-1402)."""
    out = None
    for tp in tpmats:
        Kd = tp.mats[0].todense_host()
        for m in tp.mats[1:]:
            Kd = np.kron(Kd, m.todense_host())
        Kd = float(tp.coefficient) * Kd
        out = Kd if out is None else out + Kd
    return out


def tpmats_dense_lu_factor(A) -> TPMatricesDenseLUFactors:
    """(la/tpmatrix.py:1052-1086)"""
    tpmats = [A] if isinstance(A, TPMatrix) else list(A)
    for tp in tpmats:
        for m in tp.mats:
            if not isinstance(m, Matrix):
                raise TypeError(f"tpmats_dense_lu_factor requires all factor matrices to be Matrix (dense); "
                                f"got {type(m).__name__}.")
    shape = tuple(int(m.shape[0]) for m in tpmats[0].mats)
    return TPMatricesDenseLUFactors(kron_device(tpmats), shape)


def kron_device(tpmats):
    """The dense Kronecker sum assembled on the device (``jfx_kron_sum``) for
    real factors; complex factors fall back to the host assembly."""
    from .. import _lib

    if any(m.is_complex for tp in tpmats for m in tp.mats) or any(np.iscomplexobj(tp.coefficient) for tp in tpmats):
        return tpmats_to_kron(tpmats)
    dev = K.device()
    d = len(tpmats[0].mats)
    parts = [[torch.as_tensor(m.todense_host(), dtype=torch.float64, device=dev) for m in tp.mats] for tp in tpmats]
    s = torch.tensor([float(tp.coefficient) for tp in tpmats], dtype=torch.float64, device=dev)
    ones = torch.ones(1, dtype=torch.float64, device=dev)
    # leading factors combined term by term, the last one summed over terms
    lead = []
    for mats in parts:
        Z = mats[0].contiguous()
        for m in mats[1:-1]:
            r, mm = Z.shape[0], m.shape[0]
            out = torch.empty((r * mm, r * mm), dtype=torch.float64, device=dev)
            _lib.call("jfx_kron_sum", K._p(Z), K._p(m.contiguous()), K._p(ones), 1, r, mm, K._p(out), K._s())
            Z = out
        lead.append(Z)
    X = torch.stack(lead).contiguous()
    Y = torch.stack([mats[-1] for mats in parts]).contiguous()
    r, mm = X.shape[1], Y.shape[1]
    out = torch.empty((r * mm, r * mm), dtype=torch.float64, device=dev)
    _lib.call("jfx_kron_sum", K._p(X), K._p(Y), K._p(s), len(tpmats), r, mm, K._p(out), K._s())
    return out


def _is_spd(A: np.ndarray) -> bool:
    if not np.allclose(A, A.T, rtol=1e-12, atol=1e-14 * max(1.0, np.abs(A).max())):
        return False
    try:
        np.linalg.cholesky(A)
        return True
    except np.linalg.LinAlgError:
        return False


def axis_eig(A: np.ndarray, B: np.ndarray):
    """Generalised eigenpairs ``A V = B V diag(lam)`` and the left transform W
    with ``W B V = I``: symmetric-definite -> eigh (W = V^T, la/tpmatrix.py:1173);
    otherwise eig with W = (B V)^-1."""
    sym = np.allclose(A, A.T, rtol=1e-12, atol=1e-13 * max(1.0, np.abs(A).max()))
    if sym and _is_spd(B):
        lam, V = sla.eigh(A, B)
        return lam, V, V.T.copy()
    lam, V = sla.eig(A, B)
    if np.max(np.abs(lam.imag)) > 1e-8 * max(1.0, np.max(np.abs(lam.real))):
        raise SolverNotApplicable("complex generalised eigenvalues")
    lam, V = lam.real.copy(), V.real.copy()
    order = np.argsort(lam)
    lam, V = lam[order], V[:, order]
    W = np.linalg.inv(B @ V)
    return lam, V, W


def tpmats_lu_factor(A) -> TPMatricesLUFactors:
    """Per-axis simultaneous diagonalisation (la/tpmatrix.py:1089-1233)."""
    tpmats = [A] if isinstance(A, TPMatrix) else list(A)
    ndim = tpmats[0].dims
    reps: list[np.ndarray] = []

    def rep(m) -> int:
        d = m.todense_host()
        for r, R in enumerate(reps):
            if R.shape == d.shape and np.allclose(d, R, rtol=1e-12, atol=1e-14 * max(1.0, np.abs(R).max())):
                return r
        reps.append(d)
        return len(reps) - 1

    ids = [[rep(tp.mats[i]) for i in range(ndim)] for tp in tpmats]
    eigvecs, lefts = [], []
    evals_of: dict[int, np.ndarray] = {}
    per_axis_evals: list[dict] = []
    for i in range(ndim):
        uniq = list(dict.fromkeys(ids[k][i] for k in range(len(tpmats))))
        ev: dict[int, np.ndarray] = {}
        if len(uniq) == 1:
            Am = reps[uniq[0]]
            lam, V, W = axis_eig(Am, np.eye(Am.shape[0]))
            ev[uniq[0]] = lam
        elif len(uniq) == 2:
            A0, A1 = reps[uniq[0]], reps[uniq[1]]
            if _is_spd(A1) or not _is_spd(A0):
                a, b = uniq
            else:
                b, a = uniq
            lam, V, W = axis_eig(reps[a], reps[b])
            ev[a] = lam
            ev[b] = np.ones_like(lam)
        else:
            raise SolverNotApplicable(
                f"Axis {i} has {len(uniq)} distinct factor matrices; simultaneous diagonalization "
                "requires <= 2 distinct matrices per axis.")
        eigvecs.append(V)
        lefts.append(W)
        per_axis_evals.append(ev)
    del evals_of
    per_term = [[per_axis_evals[i][ids[k][i]] for i in range(ndim)] for k in range(len(tpmats))]
    scales = [tp.coefficient for tp in tpmats]
    shape = tuple(int(tpmats[0].mats[i].shape[0]) for i in range(ndim))
    return TPMatricesLUFactors(eigvecs, lefts, per_term, scales, shape)


class TPMatrices:
    """Sum of TPMatrix terms (la/tpmatrix.py:279-509)."""

    is_zero = False

    def __init__(self, tpmats):
        self.tpmats = list(tpmats)
        self._lu_cache = None
        self._kron_cache = None

    def __len__(self) -> int:
        return len(self.tpmats)

    @property
    def shape(self):
        return self.tpmats[0].shape if self.tpmats else (0, 0)

    def _apply_array(self, u, precision: str = "f64") -> torch.Tensor:
        from .matrix import _coefficients

        w = _as_dev(_coefficients(u), K.io_dtype(precision)).contiguous()
        if all(tp._plane_able(w, precision) for tp in self.tpmats):
            # one fused launch per term, each adding into the first's output
            out = None
            for i, tp in enumerate(self.tpmats):
                out = tp._matmul_array(w, precision, out=out, accumulate=i > 0)
            return out
        parts = [tp._matmul_array(w, precision) for tp in self.tpmats]
        return parts[0] if len(parts) == 1 else K.lincomb([(1.0, p) for p in parts])

    def __call__(self, u, precision: str = "f64"):
        return self._apply_array(u, precision)

    __matmul__ = __call__

    def diagonal_or_none(self):
        tot = None
        for tp in self.tpmats:
            d = tp.diagonal_or_none()
            if d is None:
                return None
            tot = d if tot is None else tot + d
        return tot

    @property
    def is_diagonal(self) -> bool:
        return self.diagonal_or_none() is not None

    def scale(self, alpha) -> "TPMatrices":
        return TPMatrices([tp.scale(alpha) for tp in self.tpmats])

    def __add__(self, other):
        if isinstance(other, TPMatrix):
            return TPMatrices([*self.tpmats, other])
        if isinstance(other, TPMatrices):
            return TPMatrices([*self.tpmats, *other.tpmats])
        return NotImplemented

    def __sub__(self, other):
        return self + other.scale(-1.0)

    def __mul__(self, s):
        return self.scale(s)

    __rmul__ = __mul__

    def todense_host(self) -> np.ndarray:
        return tpmats_to_kron(self.tpmats)

    def todense(self) -> torch.Tensor:
        return torch.tensor(self.todense_host(), device=K.device())

    def to_matrix(self) -> Matrix:
        return Matrix(self.todense_host())

    def lu_factor(self):
        """Dispatch (la/tpmatrix.py:386-427): all-dense factors -> Kronecker LU;
        otherwise the wavenumber solver when exactly one axis is not
        diagonal, else fast diagonalisation."""
        if self._lu_cache is not None:
            return self._lu_cache
        if all(isinstance(m, Matrix) for tp in self.tpmats for m in tp.mats):
            res = tpmats_dense_lu_factor(self.tpmats)
        else:
            try:
                res = tpmats_wavenumber_factor(self.tpmats)
            except SolverNotApplicable:
                res = tpmats_lu_factor(self.tpmats)
        self._lu_cache = res
        return res

    def fd_factor(self) -> TPMatricesLUFactors:
        """Force fast diagonalisation whatever the factor storage."""
        if not isinstance(self._lu_cache, TPMatricesLUFactors):
            self._fd = getattr(self, "_fd", None) or tpmats_lu_factor(self.tpmats)
            return self._fd
        return self._lu_cache

    def solve(self, rhs, *, method: str = "auto", kron_method: str = "auto", auto_threshold: int = 100,
              precision: str = "f64"):
        """(la/tpmatrix.py:429-509): ``auto`` tries the factored path and falls
        back to the Kronecker solve on SolverNotApplicable; ``lu`` forces the
        factored path; ``kron`` the Kronecker solve; ``fd`` (an addition)
        forces fast diagonalisation."""
        d = self.diagonal_or_none()
        if d is not None:
            return _solve_diagonal(d, rhs)
        if method not in ("auto", "lu", "kron", "fd"):
            raise ValueError(f"'{method}' is not a valid TPSolveMethod")
        if method == "fd":
            return self.fd_factor().solve(rhs, precision)
        if kron_method not in ("auto", "banded", "rcm", "dense"):
            raise ValueError(f"method must be one of ['auto', 'banded', 'dense', 'rcm'], got {kron_method!r}")
        if method == "kron":
            return self._kron_solve(rhs, kron_method, auto_threshold)
        if method == "lu":
            return self.lu_factor().solve(rhs, precision)
        try:
            return self.lu_factor().solve(rhs, precision)
        except SolverNotApplicable:
            return self._kron_solve(rhs, kron_method, auto_threshold)

    def kron_factor(self, kron_method: str = "auto", auto_threshold: int = 100):
        """Factor the assembled Kronecker matrix by ``kron_method`` (cached per
        method).  All-sparse factors follow ``DiaMatrix.lu_solve``'s choice
        (la/diamatrix.py:592-770): ``banded`` (warns when p (q + 1) exceeds
        ``auto_threshold``), ``rcm`` (reverse Cuthill-McKee reordering, then
        banded), ``dense``, or ``auto`` (banded when p (q + 1) <=
        auto_threshold, else rcm when that brings it under the threshold,
        else dense, with the reference's warnings).  Any dense factor gives
        the dense Kronecker LU whatever ``kron_method`` says."""
        if self._kron_cache is None:
            self._kron_cache = {}
        shape = tuple(int(m.shape[0]) for m in self.tpmats[0].mats)
        sparse = all(isinstance(m, DiaMatrix) for tp in self.tpmats for m in tp.mats)
        method = kron_method if sparse else "dense"
        if method in self._kron_cache:
            return self._kron_cache[method]
        if method == "dense":
            f = TPMatricesDenseLUFactors(kron_device(self.tpmats), shape)
            self._kron_cache["dense"] = f
            return f
        A = _sparse_kron(self.tpmats)
        p, q = _bandwidths(A)
        if method == "banded":
            if p * (q + 1) > auto_threshold:
                warnings.warn(f"DiaMatrix.lu_solve(method='banded'): bandwidth metric p*(q+1)={p * (q + 1)} "
                              f"(p={p}, q={q}) exceeds {auto_threshold}.", stacklevel=3)
            f = TPMatricesBandedLUFactors(_band_storage(A, p, q), p, q, shape)
        elif method in ("rcm", "auto"):
            if method == "auto" and p * (q + 1) <= auto_threshold:
                f = TPMatricesBandedLUFactors(_band_storage(A, p, q), p, q, shape)
            else:
                from scipy.sparse.csgraph import reverse_cuthill_mckee

                perm = np.asarray(reverse_cuthill_mckee(A, symmetric_mode=False), dtype=np.int64)
                Ap = A[perm][:, perm]
                pp, qp = _bandwidths(Ap)
                if method == "rcm":
                    f = TPMatricesBandedLUFactors(_band_storage(Ap, pp, qp), pp, qp, shape, perm)
                elif pp * (qp + 1) > p * (q + 1):
                    warnings.warn(f"DiaMatrix.lu_solve(method='auto'): RCM reordering increased bandwidth from "
                                  f"p*(q+1)={p * (q + 1)} to p'*(q'+1)={pp * (qp + 1)}. ", stacklevel=3)
                    f = None
                elif pp * (qp + 1) <= auto_threshold:
                    f = TPMatricesBandedLUFactors(_band_storage(Ap, pp, qp), pp, qp, shape, perm)
                else:
                    warnings.warn(f"DiaMatrix.lu_solve(method='auto'): bandwidth p*(q+1)={pp * (qp + 1)} exceeds "
                                  f"threshold {auto_threshold} even after RCM reordering. ", stacklevel=3)
                    f = None
                if f is None:
                    warnings.warn("Falling back to dense solver. Consider using a larger auto_threshold or "
                                  "method='dense' to suppress this warning.", stacklevel=3)
                    f = self._kron_cache.get("dense") or TPMatricesDenseLUFactors(kron_device(self.tpmats), shape)
                    self._kron_cache["dense"] = f
        else:
            raise ValueError(f"method must be one of ['auto', 'banded', 'dense', 'rcm'], got {kron_method!r}")
        self._kron_cache[method] = f
        return f

    def _kron_solve(self, rhs, kron_method: str = "auto", auto_threshold: int = 100):
        return self.kron_factor(kron_method, auto_threshold).solve(rhs)


__all__ = ["TPMatrix", "TPMatrices", "TPMatricesLUFactors", "TPMatricesDenseLUFactors", "TPMatricesBandedLUFactors",
           "kron_device",
           "TPMatricesWavenumberSolver", "tpmats_wavenumber_factor", "DiaMatrix",
           "BaseMatrix", "tpmats_lu_factor", "tpmats_dense_lu_factor", "tpmats_to_kron", "axis_eig"]
