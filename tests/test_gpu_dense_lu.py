"""Dense LU with partial pivoting (jfx_dense_lu_factor / jfx_dense_lu_solve):
the factorisation reproduces P A = L U, solves match LAPACK (numpy), pivoting
handles zero diagonals, complex matrices and mixed real/complex operands work,
and a zero pivot raises ValueError as Matrix.lu_factor does
(la/matrix.py:137-167)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _t(a):
    import torch

    return torch.tensor(np.asarray(a), device="cuda")


def _rand(rng, n, cplx=False):
    A = rng.standard_normal((n, n))
    if cplx:
        A = A + 1j * rng.standard_normal((n, n))
    return A


@pytest.mark.parametrize("n", [1, 2, 5, 16, 17, 31, 32, 33, 64, 100, 257, 1000])
def test_factor_reproduces_PA_eq_LU(n):
    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(n)
    A = _rand(rng, n)
    LU, perm = K.dense_lu_factor(_t(A))
    LU, perm = LU.cpu().numpy(), perm.cpu().numpy()
    assert sorted(perm.tolist()) == list(range(n))
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    assert np.abs(L).max() <= 1.0 + 1e-15  # partial pivoting bounds the multipliers
    assert np.abs(A[perm] - L @ U).max() <= 1e-13 * max(1.0, np.abs(A).max()) * n


@pytest.mark.parametrize("n,nrhs", [(1, 1), (7, 3), (32, 1), (33, 70), (200, 5), (1500, 2)])
def test_solve_matches_numpy(n, nrhs):
    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(100 + n)
    A = _rand(rng, n) + n ** 0.5 * np.eye(n)
    B = rng.standard_normal((n, nrhs))
    LU, perm = K.dense_lu_factor(_t(A))
    X = K.dense_lu_solve(LU, perm, _t(B)).cpu().numpy()
    Xr = np.linalg.solve(A, B)
    assert np.abs(X - Xr).max() <= 1e-11 * np.abs(Xr).max()
    # backward error of the GPU solution
    R = A @ X - B
    assert np.abs(R).max() <= 1e-13 * (np.abs(A).sum(1).max() * np.abs(X).max() + np.abs(B).max()) * max(1, n ** 0.5)
    x1 = K.dense_lu_solve(LU, perm, _t(B[:, 0])).cpu().numpy()  # a vector right-hand side
    assert x1.shape == (n,) and np.abs(x1 - Xr[:, 0]).max() <= 1e-11 * np.abs(Xr).max()


def test_pivoting_needed():
    """Zero diagonal entries: no-pivot elimination would divide by zero."""
    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(7)
    n = 50
    Q = np.eye(n)[rng.permutation(n)]
    P = Q @ np.roll(np.eye(n), 1, axis=1) @ Q.T  # a permutation without fixed points
    A = P @ np.diag(1.0 + rng.random(n)) + 0.01 * rng.standard_normal((n, n))
    A[np.arange(n), np.arange(n)] = 0.0  # every diagonal entry zero
    B = rng.standard_normal((n, 4))
    LU, perm = K.dense_lu_factor(_t(A))
    X = K.dense_lu_solve(LU, perm, _t(B)).cpu().numpy()
    assert np.allclose(A @ X, B, rtol=0, atol=1e-13)
    A2 = np.array([[0.0, 1.0], [1.0, 0.0]])
    LU, perm = K.dense_lu_factor(_t(A2))
    assert perm.cpu().numpy().tolist() == [1, 0]


@pytest.mark.parametrize("n,nrhs", [(7, 2), (40, 3), (130, 1)])
def test_complex(n, nrhs):
    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(200 + n)
    A = _rand(rng, n, cplx=True) + 2 * n ** 0.5 * np.eye(n)
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    LU, perm = K.dense_lu_factor(_t(A))
    X = K.dense_lu_solve(LU, perm, _t(B)).cpu().numpy()
    Xr = np.linalg.solve(A, B)
    assert np.abs(X - Xr).max() <= 1e-11 * np.abs(Xr).max()
    # a real right-hand side with the complex factor
    Xb = K.dense_lu_solve(LU, perm, _t(B.real.copy())).cpu().numpy()
    assert np.abs(Xb - np.linalg.solve(A, B.real)).max() <= 1e-11 * np.abs(Xr).max()


def test_real_factor_complex_rhs():
    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(9)
    n = 90
    A = _rand(rng, n) + 3 * np.eye(n)
    B = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    LU, perm = K.dense_lu_factor(_t(A))
    X = K.dense_lu_solve(LU, perm, _t(B)).cpu().numpy()
    Xr = np.linalg.solve(A, B)
    assert np.abs(X - Xr).max() <= 1e-11 * np.abs(Xr).max()


@pytest.mark.parametrize("cplx", [False, True])
def test_zero_pivot_raises(cplx):
    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(11)
    A = _rand(rng, 40, cplx) + 5 * np.eye(40)
    A[:, 3] = 0.0
    with pytest.raises(ValueError, match="column 3"):
        K.dense_lu_factor(_t(A))


def test_matrix_and_kron_paths_use_it():
    """Matrix.solve (la/matrix.py:187-214) on a dense matrix, against numpy."""
    from jaxfun_amd.la import Matrix

    rng = np.random.default_rng(12)
    A = _rand(rng, 60) + 8 * np.eye(60)
    b = rng.standard_normal((60, 5))
    x = Matrix(A).solve(_t(b), axis=0).cpu().numpy()
    assert np.abs(x - np.linalg.solve(A, b)).max() <= 1e-12 * np.abs(x).max()
    x1 = Matrix(A).solve(_t(b.T.copy()), axis=1).cpu().numpy()
    assert np.abs(x1 - np.linalg.solve(A, b).T).max() <= 1e-12 * np.abs(x).max()
    As = A.copy()
    As[:, 7] = 0.0
    with pytest.raises(ValueError):
        Matrix(As).solve(_t(b), axis=0)


@pytest.mark.parametrize("n,cplx", [(4500, False), (2100, True)])
def test_large_panels(n, cplx):
    """Panels past 4096 rows (memory-resident panel) and the 512-thread
    register panel (up to 4096 rows; complex from 2049)."""
    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(300 + n)
    A = _rand(rng, n, cplx) + 2 * n ** 0.5 * np.eye(n)
    B = rng.standard_normal((n, 2))
    LU, perm = K.dense_lu_factor(_t(A))
    X = K.dense_lu_solve(LU, perm, _t(B)).cpu().numpy()
    R = A @ X - B
    assert np.abs(R).max() <= 1e-12 * (np.abs(A).sum(1).max() * np.abs(X).max())
    Xr = np.linalg.solve(A, B)
    assert np.abs(X - Xr).max() <= 1e-10 * np.abs(Xr).max()


def test_size_limit_rejected():
    import torch

    from jaxfun_amd import _lib

    buf = torch.zeros(4, dtype=torch.float64, device="cuda")
    piv = torch.zeros(4, dtype=torch.int32, device="cuda")
    with pytest.raises(Exception, match="exceeds"):
        _lib.call("jfx_dense_lu_factor", 0, buf.data_ptr(), 20000, piv.data_ptr(), piv.data_ptr(), piv.data_ptr(),
                  buf.data_ptr() + 8, torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 129, 640, 3844])
def test_single_launch_solve_matches_blocked(n):
    """Up to 16 right-hand sides take the single-launch solve (one workgroup
    per 64-row block, device flags between blocks; templates for 1, 2, 4, 8
    and 16 columns); 17 take the per-block launches.  Both agree with LAPACK
    and with each other."""
    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(500 + n)
    A = _rand(rng, n) + n ** 0.5 * np.eye(n)
    B = rng.standard_normal((n, 17))
    LU, perm = K.dense_lu_factor(_t(A))
    Xr = np.linalg.solve(A, B)
    scale = np.abs(Xr).max()
    xb = K.dense_lu_solve(LU, perm, _t(B)).cpu().numpy()  # blocked
    x1 = K.dense_lu_solve(LU, perm, _t(B[:, 0])).cpu().numpy()
    for X, k in [(xb, 17), (x1[:, None], 1)] + [
            (K.dense_lu_solve(LU, perm, _t(B[:, :k])).cpu().numpy(), k) for k in (2, 3, 5, 8, 16)]:
        assert np.isfinite(X).all(), k
        assert np.abs(X - Xr[:, :k]).max() <= 1e-11 * scale, k
        R = A @ X - B[:, :k]
        assert np.abs(R).max() <= 1e-13 * (np.abs(A).sum(1).max() * np.abs(X).max()) * max(1, n ** 0.5), k
        assert np.abs(X - xb[:, :k]).max() <= 1e-12 * scale, k
    # repeated solves clear the work buffer's flags each time
    for _ in range(3):
        again = K.dense_lu_solve(LU, perm, _t(B[:, 0])).cpu().numpy()
        assert np.array_equal(again, x1)
