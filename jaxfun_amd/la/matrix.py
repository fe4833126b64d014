"""Dense and DIA (banded) 1D matrices (reference: la/matrix.py, la/diamatrix.py).

Both keep the matrix on the device in FP64 (plus a host copy for the one-time
set-up work the reference also does on the host: eigendecompositions).
``matvec(x, axis)`` applies the matrix along one axis of a batch with the
MFMA contraction kernel; ``DiaMatrix.solve`` runs the banded no-pivot LU
kernels.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import kernels as K


class SolverNotApplicable(Exception):
    """A solver strategy cannot be applied (la/matrixprotocol.py:30)."""


def _host(a) -> np.ndarray:
    """Host FP64 copy; complex data (Fourier matrices with odd derivative
    orders) stays complex128."""
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    a = np.asarray(a)
    return a.astype(np.complex128 if np.iscomplexobj(a) else np.float64)


def _coefficients(x):
    """The coefficient array of a JAXFunction, else x itself."""
    from ..galerkin.arguments import JAXFunction

    return x.array if isinstance(x, JAXFunction) else x


class BaseMatrix:
    """Common interface (la/matrixprotocol.py:68-242)."""

    is_zero = False
    __array_ufunc__ = None  # numpy defers ``x @ A`` to ``A.__rmatmul__``

    def __init__(self, data):
        self._host_data = _host(data).copy()
        self._dev = None
        self._op = None
        self._lu = None

    # ---------------------------------------------------------------- data
    @property
    def shape(self) -> tuple[int, int]:
        return self._host_data.shape

    @property
    def is_complex(self) -> bool:
        return np.iscomplexobj(self._host_data)

    @property
    def dtype(self):
        return torch.complex128 if self.is_complex else torch.float64

    @property
    def size(self) -> int:
        return int(self._host_data.size)

    def todense_host(self) -> np.ndarray:
        return self._host_data

    def todense(self) -> torch.Tensor:
        if self._dev is None:
            self._dev = torch.tensor(self._host_data, dtype=self.dtype, device=K.device())
        return self._dev

    def operator(self):
        if self._op is None:
            self._op = K.ComplexAxisOperator(self.todense()) if self.is_complex else K.AxisOperator(self.todense())
        return self._op

    def diagonal_or_none(self) -> np.ndarray | None:
        A = self._host_data
        if A.shape[0] != A.shape[1]:
            return None
        if np.count_nonzero(A - np.diag(np.diag(A))) == 0:
            return np.diag(A).copy()
        return None

    # ---------------------------------------------------------------- algebra
    def matvec(self, x, axis: int = 0, precision: str = "f64") -> torch.Tensor:
        """Apply along ``axis`` (la/matrix.py:89-118, la/diamatrix.py:227-349)."""
        from ..galerkin.orthogonal import as_device

        x = as_device(x, K.io_dtype(precision))
        return K.apply_along(self.operator(), x, axis % x.ndim, precision)

    def __matmul__(self, other):
        if isinstance(other, BaseMatrix):
            return type(self)._combine(self, self._host_data @ other._host_data, other)
        return self.matvec(_coefficients(other), 0)

    def __rmatmul__(self, other):
        """``x @ A``: A^T applied along the last axis of x."""
        from ..galerkin.orthogonal import as_device

        x = as_device(_coefficients(other))
        return K.apply_along(self.T.operator(), x, x.ndim - 1, "f64")

    @staticmethod
    def _combine(a, data, b=None):
        if isinstance(a, DiaMatrix) and (b is None or isinstance(b, DiaMatrix)):
            return DiaMatrix.from_dense(data)
        return Matrix(data)

    def __mul__(self, s):
        return type(self)._combine(self, self._host_data * float(s))

    __rmul__ = __mul__

    def __add__(self, other):
        return type(self)._combine(self, self._host_data + other._host_data, other)

    def __sub__(self, other):
        return type(self)._combine(self, self._host_data - other._host_data, other)

    def __neg__(self):
        return self * -1.0

    @property
    def T(self):
        return type(self)._combine(self, self._host_data.T.copy())

    def to_matrix(self) -> "Matrix":
        return Matrix(self._host_data)

    def pin(self, constraints) -> "PinnedMatrix":
        """Fix degrees of freedom (la/diamatrix.py:1410-1500): each (row, value)
        replaces its row by the identity row and the right-hand side entry by
        the value at solve time (negative rows count from the end)."""
        items = sorted(dict(constraints).items()) if isinstance(constraints, dict) else list(constraints)
        n = self.shape[0]
        A = self._host_data.copy()
        norm = []
        for i, val in items:
            i = int(i)
            if i < 0:
                i += n
            if not 0 <= i < n:
                raise IndexError(f"pinned row {i} out of range for a matrix of size {n}")
            A[i, :] = 0.0
            A[i, i] = 1.0
            norm.append((i, val))
        return PinnedMatrix(type(self)._combine(self, A), tuple(norm))

    def tosparse(self, tol: int = 100) -> "DiaMatrix":
        return DiaMatrix.from_dense(self._host_data, tol=tol)

    # ---------------------------------------------------------------- inspection
    @property
    def ndim(self) -> int:
        return 2

    @property
    def data(self) -> np.ndarray:
        """The dense entries (host copy kept by the matrix)."""
        return self._host_data

    @property
    def nnz(self) -> int:
        """Stored nonzero entries (la/diamatrix.py)."""
        return int(np.count_nonzero(self._host_data))

    def diagonal(self, k: int = 0) -> np.ndarray:
        """The k-th diagonal (zeros for an unstored one; empty outside the shape)."""
        return np.diagonal(self._host_data, k).copy()

    def __getitem__(self, idx):
        """Entries as of the dense matrix (basic and advanced indexing)."""
        return self._host_data[idx]

    def rmatvec(self, x, axis: int = 0, precision: str = "f64") -> torch.Tensor:
        """A^T applied along ``axis`` (la/diamatrix.py:227-349)."""
        return self.T.matvec(x, axis, precision)

    def __repr__(self) -> str:
        return f"{type(self).__name__}(shape={self.shape})"


class Matrix(BaseMatrix):
    """Dense matrix (la/matrix.py:23-214)."""

    def lu_factor(self):
        """LU with partial pivoting on the device (la/matrix.py:137-167),
        cached; raises ValueError on a zero pivot."""
        if self._lu is None:
            self._lu = K.dense_lu_factor(self.todense())
        return self._lu

    def solve(self, b, axis: int = 0):
        """Dense solve along an axis (la/matrix.py:187-214)."""
        from ..galerkin.orthogonal import as_device

        b = as_device(b)
        LU, perm = self.lu_factor()
        x = b.movedim(axis, 0).reshape(b.shape[axis], -1)
        y = K.dense_lu_solve(LU, perm, x)
        return y.reshape(b.movedim(axis, 0).shape).movedim(0, axis).contiguous()


class DiaMatrix(BaseMatrix):
    """Banded matrix in DIA form (la/diamatrix.py:101-349)."""

    def __init__(self, data, offsets=None):
        super().__init__(data)
        A = self._host_data
        if offsets is None:
            m, n = A.shape
            offs = [o for o in range(-(m - 1), n) if np.any(np.diagonal(A, o) != 0)]
            offsets = tuple(offs) if offs else (0,)
        self.offsets = tuple(int(o) for o in offsets)

    @classmethod
    def from_dense(cls, A, tol: int = 100) -> "DiaMatrix":
        """Keep diagonals with entries above ``tol`` ulps (la/diamatrix.py:158-225)."""
        A = _host(A).copy()
        thr = tol * np.spacing(max(1.0, float(np.abs(A).max()) if A.size else 1.0))
        A[np.abs(A) < thr] = 0.0
        return cls(A)

    @property
    def bandwidths(self) -> tuple[int, int]:
        return max(0, -min(self.offsets)), max(0, max(self.offsets))

    def band(self) -> np.ndarray:
        p, q = self.bandwidths
        A = self._host_data
        n = A.shape[0]
        out = np.zeros((n, p + q + 1))
        for i in range(n):
            for o in range(-p, q + 1):
                if 0 <= i + o < n:
                    out[i, o + p] = A[i, i + o]
        return out

    def lu_factor(self):
        """No-pivot banded LU on the device (la/diamatrix.py:394-504); a complex
        matrix is factored densely (complex LU with partial pivoting)."""
        if self.is_complex:
            return Matrix.lu_factor(self)
        if self._lu is None:
            if self.shape[0] != self.shape[1]:
                raise ValueError("lu_factor needs a square matrix")
            p, q = self.bandwidths
            band = torch.tensor(self.band(), dtype=torch.float64, device=K.device())
            self._lu = K.band_lu_factor(band, p, q)
        return self._lu

    def solve(self, b, axis: int = 0):
        """Banded solve along an axis (la/diamatrix.py:592-770, 1880-1925)."""
        from ..galerkin.orthogonal import as_device

        if self.is_complex:
            return Matrix.solve(self, b, axis)
        lu = self.lu_factor()
        p, q = self.bandwidths
        x = as_device(b).clone()
        return K.band_lu_solve(lu, p, q, x, axis % x.ndim)

    lu_solve = solve


def diags(diagonals, offsets=(0,), shape=None) -> "DiaMatrix":
    """DiaMatrix from its diagonals (la/diamatrix.py diags, scipy.sparse.diags
    semantics): ``diagonals[i]`` fills offset ``offsets[i]`` from its start; a
    scalar is broadcast along the diagonal; ``shape`` defaults to the square
    matrix that the first diagonal and its offset imply."""
    offsets = [int(o) for o in (offsets if isinstance(offsets, (list, tuple, np.ndarray)) else [offsets])]
    if not isinstance(diagonals, (list, tuple)):
        diagonals = [diagonals]  # one array or scalar
    elif len(offsets) == 1 and len(diagonals) != 1:
        diagonals = [np.asarray(diagonals)]  # one diagonal given as a list of numbers
    diagonals = list(diagonals)
    if len(diagonals) != len(offsets):
        raise ValueError("one diagonal per offset")
    if shape is None:
        d0 = np.atleast_1d(np.asarray(diagonals[0]))
        if d0.size == 1 and np.ndim(diagonals[0]) == 0:
            raise ValueError("shape is needed for scalar diagonals")
        n = d0.size + abs(offsets[0])
        shape = (n, n)
    m, n = shape
    cplx = any(np.iscomplexobj(d) for d in diagonals)
    A = np.zeros((m, n), dtype=np.complex128 if cplx else np.float64)
    for d, o in zip(diagonals, offsets):
        L = len(np.diagonal(A, o))
        vals = np.broadcast_to(np.asarray(d), (L,)) if np.ndim(d) == 0 else np.asarray(d)[:L]
        r0, c0 = (0, o) if o >= 0 else (-o, 0)
        A[r0 + np.arange(len(vals)), c0 + np.arange(len(vals))] = vals
    return DiaMatrix(A, offsets=tuple(offsets))


class PinnedMatrix:
    """A matrix with pinned degrees of freedom (la/pinned.py): ``solve``
    writes the pinned values into a copy of the right-hand side and solves."""

    def __init__(self, matrix: BaseMatrix, constraints):
        self.matrix = matrix
        self.constraints = tuple(constraints)

    @property
    def shape(self):
        return self.matrix.shape

    def solve(self, b, axis: int = 0):
        from ..galerkin.orthogonal import as_device

        x = as_device(_coefficients(b)).clone()
        xm = x.movedim(axis, 0)
        for i, val in self.constraints:
            xm[i] = val
        return self.matrix.solve(x, axis)

    def __matmul__(self, other):
        return self.matrix @ other
