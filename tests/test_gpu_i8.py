"""FP64 contractions on int8 MFMA (precisions i8x5/6/7): the operator split
reconstructed on the host, contractions against the FP64 oracle on every
column layout, and the batched 3D Legendre round trip at 16 x 128^3."""

import os

import numpy as np
import pytest

from oracle import spectral as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _t(a, dev):
    import torch

    return torch.tensor(np.ascontiguousarray(a), device=dev, dtype=torch.float64)


@pytest.mark.parametrize("S", [5, 6, 7])
def test_split_matrix_i8_reconstructs(dev, S):
    from jaxfun_amd.kernels import AxisOperator

    rng = np.random.default_rng(20 + S)
    A = rng.standard_normal((100, 70)) * np.logspace(-6, 2, 100)[:, None]
    A[3] = 0.0  # an all-zero row
    op = AxisOperator(_t(A, dev))
    sl, ex, s_ = op.prepared(f"i8x{S}")
    d = sl.cpu().numpy().astype(np.int64)  # (S, M, 128)
    e = ex.cpu().numpy()
    assert s_ == S and d.shape == (S, 100, 128)
    assert np.abs(d).max() <= 64 and np.all(d[:, :, 70:] == 0)
    w = 2.0 ** (-6.0 - 7.0 * np.arange(S))
    R = np.ldexp(np.tensordot(w, d, axes=1)[:, :70], e[:, None])
    rowmax = np.abs(A).max(axis=1, keepdims=True)
    assert np.all(np.abs(R - A) <= 2.0 ** (-6 - 7 * (S - 1)) * np.maximum(rowmax, 1e-300) * 2)
    assert np.all(R[3] == 0)


def _check(Y, A, X, axis, S):
    Yr = O.apply_along(A, X, axis)
    # each product is exact to ~2^(-6-7(S-1)) of the row and column maxima
    bound = 8 * A.shape[1] * 2.0 ** (-6 - 7 * (S - 1)) * np.abs(A).max() * np.abs(X).max()
    assert np.abs(Y - Yr).max() <= bound, (axis, np.abs(Y - Yr).max(), bound)
    return O.rel_l2(Y, Yr)


@pytest.mark.parametrize("S", [5, 6, 7])
@pytest.mark.parametrize("axis", [1, 2, 3])
def test_contract_i8_layouts(dev, S, axis):
    # axis 1: trailing size 128^2 (Q % 64 == 0, many tiles per p); axis 2:
    # trailing 128; axis 3: trailing 1 (64 rows p per tile)
    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(30 + axis)
    n = 128
    i = np.arange(n)
    X = rng.standard_normal((2, n, n, n)) * (1.0 + i[:, None, None] + i[None, :, None] + i[None, None, :]) ** -3.0
    A = O.Space1D("legendre", n).backward_matrix()
    op = K.AxisOperator(_t(A, dev))
    Y = K.apply_along(op, _t(X, dev), axis, f"i8x{S}").cpu().numpy()
    err = _check(Y, A, X, axis, S)
    assert err < {5: 1e-8, 6: 1e-10, 7: 1e-12}[S], err


@pytest.mark.parametrize("M,K,shape,axis", [
    (100, 128, (3, 128, 64), 1),   # M < K, Q = 64
    (128, 96, (5, 96, 192), 1),    # K not a multiple of 64 (zero-padded k-step)
    (16, 64, (70, 64), 1),         # one row block, one k-step, Q = 1 with a ragged last tile (70 rows)
    (126, 126, (3, 130, 126), 2),  # the Dirichlet size: Q = 1, K = 126, 390 rows
    (128, 128, (1, 128, 64), 1),   # a single tile
])
def test_contract_i8_shapes(dev, M, K, shape, axis):
    from jaxfun_amd import kernels as K_

    rng = np.random.default_rng(M + K)
    A = rng.standard_normal((M, K))
    X = rng.standard_normal(shape)
    X[..., 0] = 0.0  # zero columns / rows of data give zero output exactly
    op = K_.AxisOperator(_t(A, dev))
    Y = K_.apply_along(op, _t(X, dev), axis, "i8x6").cpu().numpy()
    _check(Y, A, X, axis, 6)


def test_contract_i8_rejects_unsupported(dev):
    from jaxfun_amd import kernels as K

    op = K.AxisOperator(_t(np.ones((130, 130)), dev))
    with pytest.raises(ValueError):
        K.apply_along(op, _t(np.ones((2, 130, 64)), dev), 1, "i8x6")
    op = K.AxisOperator(_t(np.ones((64, 64)), dev))
    with pytest.raises(ValueError):
        K.apply_along(op, _t(np.ones((2, 64, 30)), dev), 1, "i8x6")  # trailing 30: neither 1 nor 64k
    with pytest.raises(ValueError):
        K.apply_along(K.AxisOperator(_t(np.ones((64, 63)), dev)), _t(np.ones((5, 63)), dev), 1, "i8x6")


@pytest.mark.parametrize("S", [5, 6])
@pytest.mark.parametrize("dims,B", [((128, 128, 128, 128), 3), ((128, 128, 128, 128), 600), ((100, 70, 126, 64), 5),
                                    ((33, 64, 48, 2), 7), ((126, 126, 126, 126), 257)])
def test_plane_i8_vs_fp64(dev, S, dims, B):
    # Y = A0 X A1^T in one pass (last two axes), against FP64 and against
    # two per-axis int8 contractions
    from jaxfun_amd import kernels as K

    M0, K0, M1, K1 = dims
    rng = np.random.default_rng(M0 + K0 + M1 + K1 + B)
    A0, A1 = rng.standard_normal((M0, K0)), rng.standard_normal((M1, K1))
    X = rng.standard_normal((B, K0, K1)) * np.logspace(0, -3, K1)[None, None, :]
    ops = K.AxisOperator(_t(A0, dev)), K.AxisOperator(_t(A1, dev))
    assert K.plane_ok(*ops, f"i8x{S}")
    Y = K.plane(*ops, _t(X, dev), f"i8x{S}").cpu().numpy()
    Yr = A0 @ X @ A1.T
    err = O.rel_l2(Y, Yr)
    assert err < {5: 1e-8, 6: 1e-10}[S], err
    if K1 % 64 == 0:  # the per-axis int8 path takes a trailing size of 1 or a multiple of 64
        Y2 = K.apply_along(ops[1], K.apply_along(ops[0], _t(X, dev), 1, f"i8x{S}"), 2, f"i8x{S}").cpu().numpy()
        assert O.rel_l2(Y, Y2) < {5: 3e-8, 6: 3e-10}[S]


def test_plane_i8_rejects(dev):
    from jaxfun_amd import kernels as K

    ops = K.AxisOperator(_t(np.ones((8, 8)), dev)), K.AxisOperator(_t(np.ones((8, 7)), dev))
    assert not K.plane_ok(*ops, "i8x5")  # odd last axis
    with pytest.raises(ValueError):
        K.plane(*ops, _t(np.ones((2, 8, 7)), dev), "i8x5")
    ops = K.AxisOperator(_t(np.ones((8, 8)), dev)), K.AxisOperator(_t(np.ones((8, 8)), dev))
    assert not K.plane_ok(*ops, "i8x7")
    with pytest.raises(ValueError):
        K.plane(*ops, _t(np.ones((2, 8, 8)), dev), "i8x5", alpha=2.0)


def test_legendre3d_16x128_roundtrip_i8x6(dev):
    # the batched 3D Legendre workload (seed 1) at its own size: backward vs
    # the committed oracle samples, round trip vs the input, both <= 1e-6
    import torch

    from jaxfun_amd import Legendre, TensorProduct
    from jaxfun_amd import workloads as W

    c = W.legendre_3d_coefficients(128, 16)
    L = Legendre(128)
    T = TensorProduct(L, L, L)
    ct = torch.tensor(c, device=dev)
    u = T.backward(ct, precision="i8x6")
    fx = np.load(os.path.join(GOLDEN, "c1_legendre_3d_samples.npz"))
    idx = fx["idx"]
    got = u.cpu().numpy()[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]]
    eb = O.rel_l2(got, fx["u"])
    ch = T.forward(u, precision="i8x6")
    er = float(torch.linalg.vector_norm(ch - ct) / torch.linalg.vector_norm(ct))
    assert eb < 1e-9, eb
    assert er < 1e-7, er
    # whole slices against the oracle's backward transform
    sp1 = O.Space1D("legendre", 128)
    assert O.rel_l2(u[5].cpu().numpy(), O.tp_backward([sp1] * 3, c[5])) < 1e-9
