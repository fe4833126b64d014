"""Thin torch-tensor wrappers around the libjfx C ABI.

Every function here launches a hand-written HIP kernel from ``libjfx.so`` on
the current torch stream; torch provides only device memory and the stream.
"""

from __future__ import annotations

import ctypes
import math

import torch

from . import _lib

PRECISIONS = ("f64", "f32", "bf16x3", "bf16x6", "ozaki3", "ozaki4", "ozaki5", "i8x5", "i8x6", "i8x7")
FAMILY = {"legendre": 0, "chebyshev": 1}


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("jaxfun_amd needs a ROCm GPU (gfx950); none is visible")
    _lib.load()
    return torch.device("cuda", torch.cuda.current_device())


def _p(t: torch.Tensor | None):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _s():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _req(t: torch.Tensor, dtype: torch.dtype, name: str) -> None:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA/HIP torch tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


# ---------------------------------------------------------------- basis


def gauss_nodes(family: str, N: int) -> tuple[torch.Tensor, torch.Tensor]:
    dev = device()
    x = torch.empty(N, dtype=torch.float64, device=dev)
    w = torch.empty(N, dtype=torch.float64, device=dev)
    _lib.call("jfx_gauss_nodes", FAMILY[family], N, _p(x), _p(w), _s())
    return x, w


def vandermonde(family: str, x: torch.Tensor, N: int, k: int = 0, scale: float = 1.0) -> torch.Tensor:
    _req(x, torch.float64, "x")
    V = torch.empty((x.shape[0], N), dtype=torch.float64, device=x.device)
    _lib.call("jfx_vandermonde_f64", FAMILY[family], _p(x), x.shape[0], N, k, float(scale), _p(V), _s())
    return V


def vandermonde_rec(rec: torch.Tensor, x: torch.Tensor, N: int, k: int = 0, scale: float = 1.0) -> torch.Tensor:
    """(Q, N) Vandermonde of the basis with recurrence table ``rec`` (3, N)."""
    _req(x, torch.float64, "x")
    _req(rec, torch.float64, "rec")
    if rec.numel() != 3 * N:
        raise ValueError(f"vandermonde_rec: recurrence table has {rec.numel()} entries, need {3 * N}")
    V = torch.empty((x.shape[0], N), dtype=torch.float64, device=x.device)
    _lib.call("jfx_vandermonde_rec_f64", _p(x), x.shape[0], N, k, _p(rec), float(scale), _p(V), _s())
    return V


def scale_transpose(V: torch.Tensor, rowscale: torch.Tensor | None = None,
                    colscale: torch.Tensor | None = None, s: float = 1.0) -> torch.Tensor:
    """``out[c, r] = V[r, c] * rowscale[r] * colscale[c] * s``."""
    _req(V, torch.float64, "V")
    R, C = V.shape
    out = torch.empty((C, R), dtype=torch.float64, device=V.device)
    _lib.call("jfx_scale_transpose", _p(V), R, C, _p(rowscale), _p(colscale), float(s), _p(out), _s())
    return out


# ---------------------------------------------------------------- casts / elementwise


def cast(x: torch.Tensor, dtype: torch.dtype) -> torch.Tensor:
    if x.dtype == dtype:
        return x
    x = x.contiguous()
    y = torch.empty(x.shape, dtype=dtype, device=x.device)
    if x.dtype == torch.float64 and dtype == torch.float32:
        _lib.call("jfx_cast", 0, _p(x), _p(y), x.numel(), _s())
    elif x.dtype == torch.float32 and dtype == torch.float64:
        _lib.call("jfx_cast", 1, _p(x), _p(y), x.numel(), _s())
    else:
        raise TypeError(f"unsupported cast {x.dtype} -> {dtype}")
    return y


def lincomb(terms, out: torch.Tensor | None = None, accumulate: bool = False) -> torch.Tensor:
    """``out (+)= sum c_j x_j`` for up to 8 same-shape tensors in one pass
    (real coefficients; complex tensors through their real views)."""
    terms = [(float(c), x) for c, x in terms]
    if any(x.is_complex() for _, x in terms) or (out is not None and out.is_complex()):
        terms = [(c, x if x.is_complex() else to_complex(x.contiguous())) for c, x in terms]
        ref = out if out is not None else terms[0][1]
        res = torch.empty_like(ref) if out is None else out
        lincomb([(c, torch.view_as_real(x)) for c, x in terms], torch.view_as_real(res), accumulate)
        return res
    if not terms and out is None:
        raise ValueError("lincomb needs at least one term or an output")
    ref = out if out is not None else terms[0][1]
    if out is None:
        out = torch.empty_like(ref)
    dt = {torch.float64: 0, torch.float32: 1}[out.dtype]
    for chunk_start in range(0, max(1, len(terms)), 8):
        chunk = terms[chunk_start:chunk_start + 8]
        for _, x in chunk:
            _req(x, out.dtype, "lincomb input")
            if x.shape != out.shape:
                raise ValueError("lincomb: shape mismatch")
        ptrs = (ctypes.c_void_p * 8)(*[x.data_ptr() for _, x in chunk])
        cs = (ctypes.c_double * 8)(*[c for c, _ in chunk])
        acc = 1 if (accumulate or chunk_start > 0) else 0
        _lib.call("jfx_lincomb", dt, len(chunk), ptrs, cs, _p(out), out.numel(), acc, _s())
    return out


def sum_products(terms, out: torch.Tensor | None = None) -> torch.Tensor:
    """``out = sum_t c_t * a_t * b_t`` pointwise for up to 4 (c, a, b) terms."""
    terms = [(float(c), a, b) for c, a, b in terms]
    if not 1 <= len(terms) <= 4:
        raise ValueError("sum_products takes 1 to 4 terms")
    ref = terms[0][1]
    out = torch.empty_like(ref) if out is None else out
    _req(out, out.dtype, "out")
    for _, a, b in terms:
        _req(a, out.dtype, "a")
        _req(b, out.dtype, "b")
        if a.shape != out.shape or b.shape != out.shape:
            raise ValueError(f"sum_products: shapes {tuple(a.shape)}, {tuple(b.shape)} do not match "
                             f"the output {tuple(out.shape)}")
    A = (ctypes.c_void_p * 4)(*[a.data_ptr() for _, a, _ in terms])
    B = (ctypes.c_void_p * 4)(*[b.data_ptr() for _, _, b in terms])
    C = (ctypes.c_double * 4)(*[c for c, _, _ in terms])
    dt = {torch.float64: 0, torch.float32: 1}[out.dtype]
    _lib.call("jfx_sum_products", dt, len(terms), A, B, C, _p(out), out.numel(), _s())
    return out


def mul_bcast(X: torch.Tensor, C: torch.Tensor) -> torch.Tensor:
    """``Y[b, ...] = X[b, ...] * C[...]`` (FP64, C broadcast over leading dims):
    the pointwise coefficient product of a non-separable form on the mesh."""
    _req(X, torch.float64, "X")
    _req(C, torch.float64, "C")
    if C.numel() == 0 or X.numel() % C.numel() != 0:
        raise ValueError(f"mul_bcast: {X.numel()} entries are not a whole number of {C.numel()}-entry blocks")
    Y = torch.empty_like(X)
    _lib.call("jfx_mul_bcast_f64", _p(X), _p(C), _p(Y), X.numel(), C.numel(), _s())
    return Y


def div_bcast(X: torch.Tensor, D: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """``out[b, ...] = X[b, ...] / D[...]`` with D broadcast over leading dims."""
    _req(X, X.dtype, "X")
    _req(D, X.dtype, "D")
    out = torch.empty_like(X) if out is None else out
    name = "jfx_div_bcast_f64" if X.dtype == torch.float64 else "jfx_div_bcast_f32"
    _lib.call(name, _p(X), _p(D), _p(out), X.numel(), D.numel(), _s())
    return out


def add_bcast(X: torch.Tensor, D: torch.Tensor, alpha: float = 1.0,
              out: torch.Tensor | None = None) -> torch.Tensor:
    """``out[b, ...] = X[b, ...] + alpha * D[...]`` (FP64, real or complex) with D
    broadcast over leading dims; a real X with a complex D gives a complex result."""
    if X.is_complex() or D.is_complex():
        if not X.is_complex():
            X = to_complex(X)
        if not D.is_complex():
            D = to_complex(D)
        out = torch.empty_like(X) if out is None else out
        add_bcast(torch.view_as_real(X), torch.view_as_real(D), alpha, torch.view_as_real(out))
        return out
    _req(X, torch.float64, "X")
    _req(D, torch.float64, "D")
    if tuple(X.shape[X.ndim - D.ndim:]) != tuple(D.shape):
        raise ValueError(f"add_bcast: trailing shape {tuple(X.shape)} does not end with {tuple(D.shape)}")
    out = torch.empty_like(X) if out is None else out
    if X.numel() == 0:
        return out
    _lib.call("jfx_axpy_bcast_f64", _p(X), _p(D), float(alpha), _p(out), X.numel(), D.numel(), _s())
    return out


def cmul_bcast(X: torch.Tensor, D: torch.Tensor) -> torch.Tensor:
    """``Y[b, ...] = X[b, ...] * D[...]`` (complex128) with D broadcast over leading dims."""
    if not X.is_complex():
        X = to_complex(X.contiguous())
    _req(X, torch.complex128, "X")
    _req(D, torch.complex128, "D")
    Y = torch.empty_like(X)
    _lib.call("jfx_cmul_bcast", _p(X), _p(D), _p(Y), X.numel(), D.numel(), _s())
    return Y


def to_complex(X: torch.Tensor) -> torch.Tensor:
    """Real FP64 array as complex128 (imaginary part zero) through the
    complex-combine kernel's real-input form."""
    _req(X, torch.float64, "X")
    Y = torch.empty(X.shape, dtype=torch.complex128, device=X.device)
    n = X.numel()
    if n:
        Z = torch.cat([X.reshape(1, 1, n), torch.zeros_like(X).reshape(1, 1, n)], 1)
        _lib.call("jfx_cplx_combine", _p(Z), _p(Y), 1, 1, n, 1, _s())
    return Y


# ---------------------------------------------------------------- banded LU


def band_lu_factor(band: torch.Tensor, p: int, q: int) -> torch.Tensor:
    """Factor a (n, p+q+1) by-row band matrix without pivoting.

    Raises ValueError on a zero or non-finite pivot, like
    ``DiaMatrix.lu_factor`` (la/diamatrix.py:461-471).
    """
    _req(band, torch.float64, "band")
    n = band.shape[0]
    lu = torch.empty_like(band)
    st = torch.zeros(1, dtype=torch.int32, device=band.device)
    _lib.call("jfx_band_lu_factor", _p(band), n, p, q, _p(lu), _p(st), _s())
    s = int(st.item())
    if s != 0:
        raise ValueError(f"Zero or non-finite pivot at row {s - 1} in banded LU factorization")
    return lu


def band_lu_factor_batched(band: torch.Tensor, p: int, q: int) -> torch.Tensor:
    """Factor nsys band matrices (nsys, n, p+q+1) without pivoting, one lane
    each (real FP64 or complex128)."""
    nsys, n = band.shape[0], band.shape[1]
    lu = torch.empty_like(band)
    st = torch.zeros(1, dtype=torch.int32, device=band.device)
    if band.is_complex():
        _req(band, torch.complex128, "band")
        _lib.call("jfx_band_lu_factor_batched_z", _p(band), nsys, n, p, q, _p(lu), _p(st), _s())
    else:
        _req(band, torch.float64, "band")
        _lib.call("jfx_band_lu_factor_batched", _p(band), nsys, n, p, q, _p(lu), _p(st), _s())
    s = int(st.item())
    if s != 0:
        raise ValueError(f"Zero or non-finite pivot at row {s - 1} in a batched banded LU factorization")
    return lu


BAND_SOLVE_METHODS = {"auto": 0, "lanes": 1, "wave": 2, "grid": 3}


def band_lu_solve_batched(lu: torch.Tensor, p: int, q: int, X: torch.Tensor, axis: int,
                          method: str = "auto") -> torch.Tensor:
    """Solve in place along ``axis`` of X; the systems run over the other axes
    in C order (one factor per line, the re/im pair of complex X sharing one).
    ``method`` as for ``band_lu_solve`` (real factors; complex factors use
    one lane per line)."""
    if method not in BAND_SOLVE_METHODS:
        raise ValueError(f"band solve: method must be one of {sorted(BAND_SOLVE_METHODS)}, got {method!r}")
    nsys, n = lu.shape[0], lu.shape[1]
    axis = axis % X.ndim
    if X.shape[axis] != n:
        raise ValueError(f"band solve: axis {axis} has length {X.shape[axis]}, matrices are {n}")
    if (X.numel() // n) % nsys:
        raise ValueError(f"band solve: {X.numel() // n} lines is not a multiple of {nsys} systems")
    if lu.is_complex():
        if not X.is_complex():
            raise TypeError("complex factors need a complex right-hand side")
        _req(X, torch.complex128, "X")
        P = math.prod(X.shape[:axis])
        Q = math.prod(X.shape[axis + 1:])
        _lib.call("jfx_band_lu_solve_batched_z", _p(lu), nsys, n, p, q, _p(X), P, Q, _s())
        return X
    qdiv = 1
    Xr = X
    if X.is_complex():
        _req(X, torch.complex128, "X")
        Xr = torch.view_as_real(X)
        qdiv = 2
    _req(Xr, torch.float64, "X")
    P = math.prod(X.shape[:axis])
    Q = math.prod(X.shape[axis + 1:]) * qdiv
    if method == "wave" and max(p, q) > 4:
        raise ValueError(f"band solve: the wavefront solve takes bandwidths <= 4, got p={p}, q={q}")
    _lib.call("jfx_band_lu_solve_batched_mode", _p(lu), nsys, n, p, q, _p(Xr), P, Q, qdiv, BAND_SOLVE_METHODS[method],
              _s())
    return X


def band_lu_solve(lu: torch.Tensor, p: int, q: int, X: torch.Tensor, axis: int, method: str = "auto") -> torch.Tensor:
    """Solve in place along ``axis`` of X with a factored band matrix (a complex
    X is solved through its real view: the same real factor for re and im).

    ``method``: "lanes" runs one lane per line (a sequential sweep), "wave"
    one wavefront per line (rows split over the 64 lanes, p, q <= 4), "grid"
    each line's rows over several workgroups that hand the boundary window on
    (few lines of at least 8192 rows, p, q <= 4), "auto" the grid solve where
    it applies, else the wavefront solve for few long lines and the per-lane
    sweep otherwise."""
    if method not in BAND_SOLVE_METHODS:
        raise ValueError(f"band solve: method must be one of {sorted(BAND_SOLVE_METHODS)}, got {method!r}")
    if X.is_complex():
        _req(X, torch.complex128, "X")
        band_lu_solve(lu, p, q, torch.view_as_real(X), axis % X.ndim, method)
        return X
    _req(X, torch.float64, "X")
    n = lu.shape[0]
    if X.shape[axis] != n:
        raise ValueError(f"band solve: axis {axis} has length {X.shape[axis]}, matrix is {n}")
    if method == "wave" and max(p, q) > 4:
        raise ValueError(f"band solve: the wavefront solve takes bandwidths <= 4, got p={p}, q={q}")
    P = math.prod(X.shape[:axis])
    Q = math.prod(X.shape[axis + 1:])
    if method in ("auto", "grid"):
        nbytes = _lib.load().jfx_band_lu_solve_grid_work_bytes(n, p, q, P, Q)
        if nbytes > 0:
            work = torch.empty(nbytes, dtype=torch.uint8, device=X.device)
            _lib.call("jfx_band_lu_solve_grid", _p(lu), n, p, q, _p(X), P, Q, _p(work), nbytes, _s())
            return X
        if method == "grid":
            raise ValueError(f"band solve: the grid solve takes few lines of >= 8192 rows and p, q <= 4 "
                             f"(n={n}, lines={P * Q}, p={p}, q={q})")
    _lib.call("jfx_band_lu_solve_mode", _p(lu), n, p, q, _p(X), P, Q, BAND_SOLVE_METHODS[method], _s())
    return X


# ---------------------------------------------------------------- dense LU

def dense_lu_factor(A: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """LU with partial pivoting of a square FP64 or complex128 matrix (row
    pivoting on the largest |re| + |im|), blocked over 32-column panels
    (16 complex).  Returns (LU, perm): L unit lower and U in one array, and
    row i of P A = row perm[i] of A.  Raises ValueError on a zero pivot, as
    Matrix.lu_factor does (la/matrix.py:137-167)."""
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise ValueError(f"dense LU needs a square matrix, got shape {tuple(A.shape)}")
    cplx = A.is_complex()
    _req(A, torch.complex128 if cplx else torch.float64, "A")
    n = A.shape[0]
    LU = A.contiguous().clone()
    ipiv = torch.empty(n, dtype=torch.int32, device=A.device)
    perm = torch.empty(n, dtype=torch.int32, device=A.device)
    st = torch.zeros(1, dtype=torch.int32, device=A.device)
    work = torch.empty(n * 32, dtype=torch.float64, device=A.device)  # a column-major copy of one panel
    _lib.call("jfx_dense_lu_factor", int(cplx), _p(LU), n, _p(ipiv), _p(perm), _p(st), _p(work), _s())
    s = int(st.item())
    if s != 0:
        raise ValueError(f"Matrix is singular: zero pivot in LU factorisation (column {s - 1})")
    return LU, perm


def _dense_solve_work(n: int, nrhs: int, like: torch.Tensor) -> torch.Tensor:
    """Work buffer of ``jfx_dense_lu_solve``: n x nrhs values, then
    2 ceil(n / 64) ints of block flags."""
    elt = like.element_size()
    extra = (2 * ((n + 63) // 64) * 4 + elt - 1) // elt
    return torch.empty(n * nrhs + extra, dtype=like.dtype, device=like.device)


def dense_lu_solve(LU: torch.Tensor, perm: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    """X = A^-1 B for B of shape (n,) or (n, nrhs), with (LU, perm) from
    ``dense_lu_factor``.  A complex B with a real factor is solved as its
    real and imaginary parts; a real B with a complex factor is promoted."""
    n = LU.shape[0]
    if B.shape[0] != n:
        raise ValueError(f"dense LU solve: B has {B.shape[0]} rows, the matrix is {n} x {n}")
    vec = B.ndim == 1
    B2 = B.reshape(n, -1)
    if LU.is_complex():
        B2 = (B2 if B2.is_complex() else to_complex(B2.contiguous())).contiguous()
        _req(B2, torch.complex128, "B")
        X = torch.empty_like(B2)
        W = _dense_solve_work(n, B2.shape[1], B2)
        _lib.call("jfx_dense_lu_solve", 1, _p(LU), n, _p(perm), _p(B2), _p(X), B2.shape[1], _p(W), _s())
    elif B2.is_complex():
        B2 = B2.contiguous()
        _req(B2, torch.complex128, "B")
        Br = torch.view_as_real(B2).reshape(n, -1)  # (n, 2 nrhs): re, im interleaved
        Xr = torch.empty_like(Br)
        W = _dense_solve_work(n, Br.shape[1], Br)
        _lib.call("jfx_dense_lu_solve", 0, _p(LU), n, _p(perm), _p(Br), _p(Xr), Br.shape[1], _p(W), _s())
        X = torch.view_as_complex(Xr.reshape(n, -1, 2))
    else:
        B2 = B2.contiguous()
        _req(B2, torch.float64, "B")
        X = torch.empty_like(B2)
        W = _dense_solve_work(n, B2.shape[1], B2)
        _lib.call("jfx_dense_lu_solve", 0, _p(LU), n, _p(perm), _p(B2), _p(X), B2.shape[1], _p(W), _s())
    return X.reshape(n) if vec else X


# ---------------------------------------------------------------- per-axis contraction


class AxisOperator:
    """A constant (M x K) matrix applied along one axis, prepared per precision.

    ``f64`` uses the FP64 matrix; ``f32`` an FP32 copy; ``bf16x3``/``bf16x6``
    2/3 bf16 float slices; ``ozakiS`` S fixed-point digit slices with
    power-of-two row scales; ``i8xS`` S balanced int8 digit slices with
    power-of-two row scales (FP64 data, int8 MFMA, exact integer sums).  Preparations are cached on the object, like the
    reference caches factorizations on its matrix objects.
    """

    def __init__(self, A: torch.Tensor):
        if A.dtype != torch.float64:
            A = A.to(torch.float64)
        self.A64 = A.contiguous()
        self.M, self.K = self.A64.shape
        self._prep: dict[str, tuple] = {}

    def _padded(self):
        """Zero-pad K to a multiple of 32 (the slice row stride the kernels read)."""
        Kp = (self.K + 31) // 32 * 32
        if Kp == self.K:
            return self.A64, Kp
        Ap = torch.zeros((self.M, Kp), dtype=torch.float64, device=self.A64.device)
        Ap[:, :self.K] = self.A64
        return Ap, Kp

    def prepared(self, prec: str):
        if prec in self._prep:
            return self._prep[prec]
        A = self.A64
        if prec == "f64":
            r = (A,)
        elif prec == "f32":
            r = (cast(A, torch.float32),)
        elif prec in ("bf16x3", "bf16x6"):
            S = 2 if prec == "bf16x3" else 3
            Ap, Kp = self._padded()
            sl = torch.empty((S, self.M, Kp), dtype=torch.int16, device=A.device)
            _lib.call("jfx_split_matrix", _p(Ap), self.M, Kp, 0, S, _p(sl), _p(None), _s())
            r = (sl, S)
        elif prec.startswith("ozaki"):
            S = int(prec[5:])
            Ap, Kp = self._padded()
            sl = torch.empty((S, self.M, Kp), dtype=torch.int16, device=A.device)
            rs = torch.empty(self.M, dtype=torch.float64, device=A.device)
            _lib.call("jfx_split_matrix", _p(Ap), self.M, Kp, 1, S, _p(sl), _p(rs), _s())
            r = (sl, rs, S)
        elif prec.startswith("i8x") and prec in PRECISIONS:
            S = int(prec[3:])
            Kp = (self.K + 63) // 64 * 64
            sl = torch.empty((S, self.M, Kp), dtype=torch.int8, device=A.device)
            ex = torch.empty(self.M, dtype=torch.int32, device=A.device)
            _lib.call("jfx_split_matrix_i8", _p(A), self.M, self.K, S, _p(sl), _p(ex), _s())
            r = (sl, ex, S)
        else:
            raise ValueError(f"unknown precision {prec!r}; expected one of {PRECISIONS}")
        self._prep[prec] = r
        return r


def io_dtype(prec: str) -> torch.dtype:
    return torch.float32 if prec in ("f32", "bf16x3", "bf16x6") else torch.float64


class ComplexAxisOperator:
    """A constant complex (M x K) matrix applied along one axis (FP64).

    Kept as the stacked real (2M x K) operator [Re A; Im A]: one real MFMA
    contraction over the real view of X, then ``jfx_cplx_combine`` forms
    ``Re + i Im`` of the result.
    """

    is_complex = True

    def __init__(self, A: torch.Tensor):
        A = A.to(torch.complex128)
        self.M, self.K = A.shape
        self.A = A.contiguous()
        self.stacked = AxisOperator(torch.cat([A.real, A.imag], 0).contiguous())

    def prepared(self, prec: str):
        if prec != "f64":
            raise ValueError("complex operators are applied in f64 only")
        return self.stacked.prepared(prec)


def _apply_complex(op: ComplexAxisOperator, X: torch.Tensor, axis: int) -> torch.Tensor:
    axis = axis % X.ndim
    if X.shape[axis] != op.K:
        raise ValueError(f"apply_along: axis {axis} has length {X.shape[axis]}, operator expects {op.K}")
    P = math.prod(X.shape[:axis])
    Q = math.prod(X.shape[axis + 1:])
    shape = list(X.shape)
    shape[axis] = op.M
    Y = torch.empty(shape, dtype=torch.complex128, device=X.device)
    if X.is_complex():
        _req(X, torch.complex128, "X")
        Z = apply_along(op.stacked, torch.view_as_real(X).reshape(P, op.K, 2 * Q), 1)
        _lib.call("jfx_cplx_combine", _p(Z), _p(Y), P, op.M, Q, 0, _s())
    else:
        _req(X, torch.float64, "X")
        Z = apply_along(op.stacked, X.reshape(P, op.K, Q), 1)
        _lib.call("jfx_cplx_combine", _p(Z), _p(Y), P, op.M, Q, 1, _s())
    return Y


def apply_along(op: AxisOperator, X: torch.Tensor, axis: int, prec: str = "f64",
                out: torch.Tensor | None = None) -> torch.Tensor:
    """``Y = A`` applied along ``axis`` of X: Y[p, m, q] = sum_k A[m, k] X[p, k, q].

    A complex X (FP64 only) is contracted through its real view (the re/im
    pair rides along the trailing axis); a :class:`ComplexAxisOperator`
    gives a complex result.
    """
    if getattr(op, "is_complex", False):
        if prec != "f64" or out is not None:
            raise ValueError("complex operators are applied in f64 without an out= tensor")
        return _apply_complex(op, X, axis)
    if X.is_complex():
        if prec != "f64":
            raise ValueError("complex arrays are transformed in f64 only")
        _req(X, torch.complex128, "X")
        axis = axis % X.ndim
        shape = list(X.shape)
        shape[axis] = op.M
        Y = torch.empty(shape, dtype=torch.complex128, device=X.device) if out is None else out
        apply_along(op, torch.view_as_real(X), axis, "f64", out=torch.view_as_real(Y))
        return Y
    dt = io_dtype(prec)
    _req(X, dt, "X")
    axis = axis % X.ndim
    K = X.shape[axis]
    if K != op.K:
        raise ValueError(f"apply_along: axis {axis} has length {K}, operator expects {op.K}")
    P = math.prod(X.shape[:axis])
    Q = math.prod(X.shape[axis + 1:])
    shape = list(X.shape)
    shape[axis] = op.M
    Y = torch.empty(shape, dtype=dt, device=X.device) if out is None else out
    if out is not None:
        _req(out, dt, "out")
        if list(out.shape) != shape:
            raise ValueError("apply_along: bad out shape")
    prep = op.prepared(prec)
    if prec == "f64":
        _lib.call("jfx_contract_f64", _p(prep[0]), op.M, K, _p(X), _p(Y), P, Q, _s())
    elif prec == "f32":
        _lib.call("jfx_contract_f32", _p(prep[0]), op.M, K, _p(X), _p(Y), P, Q, _s())
    elif prec in ("bf16x3", "bf16x6"):
        _lib.call("jfx_contract_bf16x", _p(prep[0]), prep[1], op.M, K, _p(X), _p(Y), P, Q, _s())
    elif prec.startswith("i8x"):
        if op.M > 128 or K > 128 or not (Q == 1 or Q % 64 == 0) or (Q == 1 and K % 2):
            raise ValueError(f"precision {prec}: needs M, K <= 128 and a trailing size of 1 (K even) or a "
                             f"multiple of 64; got M={op.M}, K={K}, trailing size {Q}")
        _lib.call("jfx_contract_i8", _p(prep[0]), _p(prep[1]), prep[2], op.M, K, _p(X), _p(Y), P, Q, _s())
    else:
        _lib.call("jfx_contract_ozaki", _p(prep[0]), _p(prep[1]), prep[2], op.M, K, _p(X), _p(Y), P, Q,
                  _s())
    return Y


PLANE_PRECISIONS = ("f64", "bf16x6", "i8x5", "i8x6")


_NUM_CUS = None


def num_cus() -> int:
    """Compute units of the current device (cached)."""
    global _NUM_CUS
    if _NUM_CUS is None:
        _NUM_CUS = torch.cuda.get_device_properties(device()).multi_processor_count
    return _NUM_CUS


def plane_ok(op0: AxisOperator, op1: AxisOperator, prec: str) -> bool:
    if getattr(op0, "is_complex", False) or getattr(op1, "is_complex", False):
        return False
    if prec.startswith("i8x") and op1.K % 2:
        return False  # the int8 plane kernel reads rows in 16-byte pairs
    return prec in PLANE_PRECISIONS and max(op0.M, op0.K, op1.M, op1.K) <= 128


def plane_fd_ok(ops, prec: str) -> bool:
    """The fused 2D solve applies: four real n x n operators, n <= 64 (FP64) or 128 (bf16x6)."""
    if prec not in ("f64", "bf16x6"):
        return False
    if any(type(o) is not AxisOperator for o in ops):
        return False
    n = ops[0].M
    if any(o.M != n or o.K != n for o in ops):
        return False
    return n <= (64 if prec == "f64" else 128)


def plane_fd(W0: AxisOperator, W1: AxisOperator, V0: AxisOperator, V1: AxisOperator, X: torch.Tensor,
             E: torch.Tensor, prec: str = "f64") -> torch.Tensor:
    """``Y[b] = V0 ((W0 X[b] W1^T) / E) V1^T`` over the last two axes in one kernel."""
    dt = io_dtype(prec)
    _req(X, dt, "X")
    _req(E, dt, "E")
    n = W0.M
    if not plane_fd_ok([W0, W1, V0, V1], prec) or tuple(X.shape[-2:]) != (n, n) or tuple(E.shape) != (n, n):
        raise ValueError("plane_fd: operators, array or divisor do not fit the fused solve")
    B = math.prod(X.shape[:-2])
    Y = torch.empty_like(X)
    if prec == "f64":
        _lib.call("jfx_plane_fd", 0, _p(W0.A64), _p(W1.A64), _p(V0.A64), _p(V1.A64), _p(None), _p(None), _p(None),
                  _p(None), _p(X), _p(Y), _p(E), B, n, _s())
    else:
        s = [o.prepared("bf16x6")[0] for o in (W0, W1, V0, V1)]
        _lib.call("jfx_plane_fd", 1, _p(None), _p(None), _p(None), _p(None), _p(s[0]), _p(s[1]), _p(s[2]), _p(s[3]),
                  _p(X), _p(Y), _p(E), B, n, _s())
    return Y


def plane(op0: AxisOperator, op1: AxisOperator, X: torch.Tensor, prec: str = "f64",
          epi: tuple | None = None, alpha: float = 1.0, out: torch.Tensor | None = None,
          accumulate: bool = False) -> torch.Tensor:
    """``Y[b] = alpha A0 X[b] A1^T`` over the last two axes of X in one kernel
    (optionally ``/ E`` or ``* E`` with ``epi=("div"|"mul", E)``); with
    ``accumulate`` the result is added to ``out``."""
    dt = io_dtype(prec)
    _req(X, dt, "X")
    if X.shape[-2] != op0.K or X.shape[-1] != op1.K:
        raise ValueError("plane: operator/array shape mismatch")
    B = math.prod(X.shape[:-2])
    shape = X.shape[:-2] + (op0.M, op1.M)
    if out is None:
        if accumulate:
            raise ValueError("plane: accumulate needs out=")
        Y = torch.empty(shape, dtype=dt, device=X.device)
    else:
        _req(out, dt, "out")
        if tuple(out.shape) != tuple(shape):
            raise ValueError("plane: bad out shape")
        if out.data_ptr() == X.data_ptr():
            raise ValueError("plane: out must not alias X")
        Y = out
    mode, E = 0, None
    if epi is not None:
        mode = {"div": 1, "mul": 2}[epi[0]]
        E = epi[1]
        _req(E, dt, "E")
        if tuple(E.shape) != (op0.M, op1.M):
            raise ValueError("plane: epilogue shape mismatch")
    acc = 1 if accumulate else 0
    if prec.startswith("i8x"):
        if mode or accumulate or alpha != 1.0:
            raise ValueError("plane: the int8 path has no epilogue, scale or accumulation")
        if not plane_ok(op0, op1, prec):
            raise ValueError(f"plane: precision {prec} needs dims <= 128 and an even last axis")
        p0, p1 = op0.prepared(prec), op1.prepared(prec)
        _lib.call("jfx_plane_i8", _p(p0[0]), _p(p0[1]), _p(p1[0]), _p(p1[1]), p0[2], _p(X), _p(Y), B, op0.M, op0.K,
                  op1.M, op1.K, _s())
        return Y
    if prec == "f64":
        _lib.call("jfx_plane_contract", 0, _p(op0.A64), _p(op1.A64), _p(None), _p(None), _p(X), _p(Y), _p(E),
                  mode, float(alpha), acc, B, op0.M, op0.K, op1.M, op1.K, _s())
    else:
        s0, s1 = op0.prepared("bf16x6")[0], op1.prepared("bf16x6")[0]
        _lib.call("jfx_plane_contract", 1, _p(None), _p(None), _p(s0), _p(s1), _p(X), _p(Y), _p(E),
                  mode, float(alpha), acc, B, op0.M, op0.K, op1.M, op1.K, _s())
    return Y
