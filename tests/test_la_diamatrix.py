"""DiaMatrix / diags behaviour from the reference's tests/la/test_diamatrix.py:
host construction and inspection on the CPU, products and banded solves
along an axis on the GPU (against numpy)."""

import numpy as np
import pytest
import scipy.sparse as sps

from jaxfun_amd.la import DiaMatrix, diags


@pytest.mark.parametrize("dg,off,shape", [([np.ones(4), 2 * np.ones(3)], (0, 1), None),
                                          ([1.0, -2.0, 1.0], (-1, 0, 1), (6, 6)),
                                          ([np.arange(1.0, 4.0)], (0,), (3, 5)),
                                          (np.arange(1.0, 5.0), -1, (5, 4))])
def test_diags_vs_scipy(dg, off, shape):
    # test_diamatrix.py: test_diags_vs_scipy / test_diags_nonsquare_explicit / test_diags_scalar_broadcast
    A = diags(dg, off, shape)
    B = sps.diags(dg, off, shape=shape).toarray()
    assert A.shape == B.shape and np.array_equal(A.todense_host(), B)
    assert A.nnz == np.count_nonzero(B) and A.ndim == 2


def test_from_dense_inspection_and_transpose():
    # test_from_dense_* / test_main/super/sub_diagonal / test_unstored_diagonal_zeros / test_T_*
    rng = np.random.default_rng(0)
    B = np.triu(np.tril(rng.standard_normal((7, 9)), 2), -1)
    A = DiaMatrix.from_dense(B)
    assert A.offsets == (-1, 0, 1, 2)
    for k in (-1, 0, 1, 2, 3, -4):
        assert np.array_equal(A.diagonal(k), np.diagonal(B, k))
    assert np.array_equal(A.T.todense_host(), B.T) and A.T.offsets == (-2, -1, 0, 1)
    assert np.array_equal(A.T.T.todense_host(), B)
    assert np.array_equal(A[2, 1:4], B[2, 1:4]) and np.array_equal(A[[0, 3], [1, 4]], B[[0, 3], [1, 4]])
    assert np.array_equal((2.0 * A).todense_host(), 2 * B) and np.array_equal((-A).todense_host(), -B)
    assert np.array_equal((A + A).todense_host(), 2 * B) and np.count_nonzero((A - A).todense_host()) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("axis", [0, 1, 2, -1])
def test_matvec_and_solve_along_axis(dev, axis):
    # test_axis0_3d / test_axis1_3d / test_axis2_3d / test_negative_axis / test_solve_axis*
    rng = np.random.default_rng(1)
    n = 6
    B = np.diag(4.0 + rng.random(n)) + np.diag(rng.random(n - 1), 1) + np.diag(rng.random(n - 2), -2)
    A = DiaMatrix.from_dense(B)
    X = rng.standard_normal((n, n, n))
    Y = A.matvec(X, axis).cpu().numpy()
    ref = np.moveaxis(np.tensordot(B, X, axes=([1], [axis % 3])), 0, axis % 3)
    assert np.allclose(Y, ref, atol=1e-13)
    Z = A.solve(ref, axis=axis).cpu().numpy()
    assert np.allclose(Z, X, atol=1e-12)
    R = A.rmatvec(X, axis).cpu().numpy()
    assert np.allclose(R, np.moveaxis(np.tensordot(B.T, X, axes=([1], [axis % 3])), 0, axis % 3), atol=1e-13)


@pytest.mark.gpu
def test_matmul_forms(dev):
    # test_matmul_vector / test_rmatmul_vector / test_matmul_dia_x_matrix
    rng = np.random.default_rng(2)
    B = np.diag(rng.random(5)) + np.diag(rng.random(4), 1)
    C = rng.standard_normal((5, 5))
    A = DiaMatrix.from_dense(B)
    x = rng.standard_normal(5)
    assert np.allclose((A @ x).cpu().numpy(), B @ x, atol=1e-14)
    assert np.allclose((x @ A).cpu().numpy(), x @ B, atol=1e-14)
    from jaxfun_amd.la import Matrix

    assert np.allclose((A @ Matrix(C)).todense_host(), B @ C, atol=1e-14)
