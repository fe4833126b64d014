"""GPU parity of the low-level kernels (through the C ABI) against the oracle."""

import numpy as np
import pytest

from oracle import spectral as O

pytestmark = pytest.mark.gpu


def _t(a, dev, dtype=None):
    import torch

    return torch.tensor(np.asarray(a), device=dev, dtype=dtype or torch.float64)


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
@pytest.mark.parametrize("N", [1, 2, 8, 64, 128, 192, 256])
def test_gauss_nodes(dev, family, N):
    from jaxfun_amd import kernels as K

    x, w = K.gauss_nodes(family, N)
    xr, wr = O.quad(family, N)
    assert np.allclose(x.cpu().numpy(), xr, atol=1e-15)
    assert np.allclose(w.cpu().numpy(), wr, rtol=1e-13)


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_vandermonde(dev, family, k):
    from jaxfun_amd import kernels as K

    x = np.linspace(-1, 1, 37)
    V = K.vandermonde(family, _t(x, dev), 40, k).cpu().numpy()
    Vr = O.vandermonde(family, x, 40, k)
    assert np.allclose(V, Vr, rtol=1e-12, atol=1e-12 * np.abs(Vr).max())


SHAPES = [  # (shape, axis, M)
    ((3, 128, 128, 128), 1, 128),
    ((2, 128, 128, 128), 3, 128),
    ((2, 128, 128, 128), 2, 128),
    ((5, 62, 62), 1, 64),
    ((7, 62, 62), 2, 62),
    ((33, 50), 0, 37),
    ((4, 256, 130), 1, 256),
    ((3, 96, 5), 1, 200),
    ((1, 17), 1, 1),
]
# shapes of the register-resident-operator kernel (M, K <= 128; Q == 1 or Q % 64 == 0):
# partial last tiles, M and K below 128, row counts that leave waves idle
RA_SHAPES = [
    ((100, 64), 1, 64),
    ((3, 60, 64), 1, 36),
    ((130, 120), 1, 124),
    ((2, 96, 128), 1, 96),
    ((5, 72), 1, 40),
    ((2, 32, 192), 1, 16),
]


@pytest.mark.parametrize("shape,axis,M", SHAPES + RA_SHAPES)
def test_contract_f64(dev, shape, axis, M):
    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(1)
    X = rng.standard_normal(shape)
    A = rng.standard_normal((M, shape[axis]))
    Y = K.apply_along(K.AxisOperator(_t(A, dev)), _t(X, dev), axis, "f64").cpu().numpy()
    Yr = O.apply_along(A, X, axis)
    assert O.rel_l2(Y, Yr) < 1e-14


# ozakiS drops the digit products of level >= S: error ~ (S+1) 2^-8S sqrt(K)
@pytest.mark.parametrize("prec,tol", [("f32", 2e-6), ("bf16x3", 2e-5), ("bf16x6", 2e-6),
                                      ("ozaki3", 1e-6), ("ozaki4", 5e-9), ("ozaki5", 5e-11)])
@pytest.mark.parametrize("shape,axis,M", SHAPES[:6] + RA_SHAPES)
def test_contract_mixed(dev, prec, tol, shape, axis, M):
    import torch

    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(2)
    X = rng.standard_normal(shape)
    A = rng.standard_normal((M, shape[axis]))
    dt = K.io_dtype(prec)
    Xd = _t(X, dev, dt)
    Y = K.apply_along(K.AxisOperator(_t(A, dev)), Xd, axis, prec).to(torch.float64).cpu().numpy()
    Yr = O.apply_along(A, Xd.to(torch.float64).cpu().numpy(), axis)
    err = O.rel_l2(Y, Yr)
    assert err < tol, (prec, err)


def test_ozaki_rejects_inexact_sizes(dev):
    from jaxfun_amd import kernels as K

    A = _t(np.ones((4, 300)), dev)
    X = _t(np.ones((2, 300)), dev)
    with pytest.raises(ValueError):
        K.apply_along(K.AxisOperator(A), X, 1, "ozaki4")


def test_band_lu(dev):
    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(3)
    n, p, q = 50, 2, 2
    A = np.zeros((n, n))
    for d in range(-p, q + 1):
        A += np.diag(rng.standard_normal(n - abs(d)), d)
    A += 10 * np.eye(n)
    band = np.zeros((n, p + q + 1))
    for i in range(n):
        for o in range(-p, q + 1):
            if 0 <= i + o < n:
                band[i, o + p] = A[i, i + o]
    lu = K.band_lu_factor(_t(band, dev), p, q)
    B = rng.standard_normal((3, n, 7))
    X = K.band_lu_solve(lu, p, q, _t(B, dev).contiguous(), 1).cpu().numpy()
    Xr = np.stack([O.banded_lu_solve(A, B[i], p, q) for i in range(3)])
    assert np.allclose(X, Xr, rtol=1e-12, atol=1e-12)
    band[0, p] = 0.0
    with pytest.raises(ValueError):
        K.band_lu_factor(_t(band, dev), p, q)


def _banded(rng, n, p, q, dominance=4.0):
    A_band = np.zeros((n, p + q + 1))
    for o in range(-p, q + 1):
        lo, hi = max(0, -o), min(n, n - o)
        A_band[lo:hi, o + p] = rng.standard_normal(hi - lo)
    A_band[:, p] += dominance * (p + q + 1) * np.sign(A_band[:, p] + 0.5)
    return A_band


def _band_to_ab(A_band, p, q):
    """The LAPACK diagonal-ordered form scipy.linalg.solve_banded takes."""
    n = A_band.shape[0]
    ab = np.zeros((p + q + 1, n))
    for o in range(-p, q + 1):
        lo, hi = max(0, -o), min(n, n - o)
        ab[q - o, lo + o:hi + o] = A_band[lo:hi, o + p]
    return ab


@pytest.mark.parametrize(
    "n,p,q,shape,axis",
    [
        (5, 2, 2, (5,), 0),  # fewer rows than lanes: most chunks empty
        (63, 1, 1, (63,), 0),
        (64, 2, 2, (64,), 0),
        (1000, 2, 2, (1000,), 0),
        (4097, 4, 4, (4097,), 0),
        (3000, 0, 2, (3000,), 0),  # upper factor only
        (3000, 3, 0, (3000,), 0),  # lower factor only
        (2001, 1, 3, (2, 2001, 3), 1),  # lines across leading and trailing axes
        (700, 2, 2, (5, 700), 1),
        (100000, 2, 2, (100000,), 0),
    ],
)
def test_band_solve_wave_vs_banded_lapack(dev, n, p, q, shape, axis):
    """One wavefront per line (rows split over the lanes, chunk maps composed by
    a wave scan) against LAPACK's banded solve and the per-lane sweep."""
    import scipy.linalg as sla

    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(n + 10 * p + q)
    A_band = _banded(rng, n, p, q)
    lu = K.band_lu_factor(_t(A_band, dev), p, q)
    B = rng.standard_normal(shape)
    Bl = np.moveaxis(B, axis, 0).reshape(n, -1)
    Xr = np.moveaxis(sla.solve_banded((p, q), _band_to_ab(A_band, p, q), Bl).reshape(
        (n,) + tuple(np.delete(shape, axis))), 0, axis)
    Xw = K.band_lu_solve(lu, p, q, _t(B, dev).contiguous(), axis, method="wave").cpu().numpy()
    Xl = K.band_lu_solve(lu, p, q, _t(B, dev).contiguous(), axis, method="lanes").cpu().numpy()
    Xa = K.band_lu_solve(lu, p, q, _t(B, dev).contiguous(), axis).cpu().numpy()
    scale = np.abs(Xr).max()
    assert np.abs(Xw - Xr).max() <= 1e-13 * scale
    assert np.abs(Xl - Xr).max() <= 1e-13 * scale
    assert np.abs(Xw - Xl).max() <= 1e-13 * scale
    same = [np.array_equal(Xa, Xw), np.array_equal(Xa, Xl)]
    if n >= 8192:  # long enough for the grid solve, which auto then takes
        Xg = K.band_lu_solve(lu, p, q, _t(B, dev).contiguous(), axis, method="grid").cpu().numpy()
        assert np.abs(Xg - Xr).max() <= 1e-13 * scale
        same.append(np.array_equal(Xa, Xg))
    assert any(same)


@pytest.mark.parametrize(
    "n,p,q,shape,axis",
    [
        (8192, 2, 2, (8192,), 0),  # two chunks
        (8193, 1, 1, (8193,), 0),  # a last chunk of one row
        (50000, 4, 4, (50000,), 0),
        (16384, 0, 2, (16384,), 0),  # upper factor only
        (16384, 3, 0, (16384,), 0),  # lower factor only
        (30000, 1, 3, (2, 30000, 3), 1),  # six lines of eight chunks
        (300000, 2, 2, (300000,), 0),  # 59 chunks of 5120 rows
        # 40 lines x 25 chunks = 1000 workgroups, beyond the GPU's resident
        # slots: waiting workgroups only wait for lower-numbered ones
        (100000, 2, 2, (100000, 40), 0),
    ],
)
def test_band_solve_grid_vs_banded_lapack(dev, n, p, q, shape, axis):
    """Each line's rows over several workgroups that hand the recurrence
    window on through device flags, against LAPACK's banded solve and the
    one-workgroup wavefront solve; "auto" takes it for these shapes."""
    import scipy.linalg as sla

    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(n + 10 * p + q + 1)
    A_band = _banded(rng, n, p, q)
    lu = K.band_lu_factor(_t(A_band, dev), p, q)
    B = rng.standard_normal(shape)
    Bl = np.moveaxis(B, axis, 0).reshape(n, -1)
    Xr = np.moveaxis(sla.solve_banded((p, q), _band_to_ab(A_band, p, q), Bl).reshape(
        (n,) + tuple(np.delete(shape, axis))), 0, axis)
    Xg = K.band_lu_solve(lu, p, q, _t(B, dev).contiguous(), axis, method="grid").cpu().numpy()
    Xw = K.band_lu_solve(lu, p, q, _t(B, dev).contiguous(), axis, method="wave").cpu().numpy()
    Xa = K.band_lu_solve(lu, p, q, _t(B, dev).contiguous(), axis).cpu().numpy()
    scale = np.abs(Xr).max()
    assert np.isfinite(Xg).all()
    assert np.abs(Xg - Xr).max() <= 1e-13 * scale
    assert np.abs(Xg - Xw).max() <= 1e-13 * scale
    assert np.array_equal(Xa, Xg)
    # repeated calls start from cleared flags
    again = K.band_lu_solve(lu, p, q, _t(B, dev).contiguous(), axis, method="grid").cpu().numpy()
    assert np.array_equal(again, Xg)


def test_band_solve_grid_rejects_unsuited_shapes(dev):
    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(3)
    lu = K.band_lu_factor(_t(_banded(rng, 1000, 2, 2), dev), 2, 2)
    with pytest.raises(ValueError, match="grid solve"):
        K.band_lu_solve(lu, 2, 2, _t(rng.standard_normal(1000), dev), 0, method="grid")


def test_band_solve_wave_legendre_mass_large(dev):
    """DiaMatrix.solve of one long vector with the Legendre-Dirichlet mass matrix
    (offsets 0, +-2) goes to the grid solve and matches LAPACK."""
    import scipy.linalg as sla

    from jaxfun_amd import kernels as K

    n = 20000
    k = np.arange(n, dtype=np.float64)
    A_band = np.zeros((n, 5))
    A_band[:, 2] = 2.0 / (2 * k + 1) + 2.0 / (2 * k + 5)
    A_band[2:, 0] = -2.0 / (2 * k[:-2] + 5)  # A[i, i-2] = M[i-2, i]
    A_band[:-2, 4] = -2.0 / (2 * k[:-2] + 5)
    lu = K.band_lu_factor(_t(A_band, dev), 2, 2)
    ab = _band_to_ab(A_band, 2, 2)

    def matvec(x):
        y = A_band[:, 2] * x
        y[2:] += A_band[2:, 0] * x[:-2]
        y[:-2] += A_band[:-2, 4] * x[2:]
        return y

    rng = np.random.default_rng(7)
    x_true = rng.standard_normal(n)
    for b in (matvec(x_true), rng.standard_normal(n)):
        x = K.band_lu_solve(lu, 2, 2, _t(b, dev).contiguous(), 0).cpu().numpy()
        xl = K.band_lu_solve(lu, 2, 2, _t(b, dev).contiguous(), 0, method="lanes").cpu().numpy()
        xr = sla.solve_banded((2, 2), ab, b)
        # backward error of a no-pivot LU of this SPD matrix: a few ulps
        absA = np.abs(A_band).sum(axis=1).max()
        for y in (x, xl, xr):
            assert np.abs(matvec(y) - b).max() <= 1e-14 * (absA * np.abs(y).max() + np.abs(b).max())
        # the mass matrix's condition number grows like n (about 1e5 here)
        assert np.linalg.norm(x - xl) <= 1e-10 * np.linalg.norm(xl)
        assert np.linalg.norm(x - xr) <= 1e-10 * np.linalg.norm(xr)
    assert np.linalg.norm(x_true - K.band_lu_solve(lu, 2, 2, _t(matvec(x_true), dev), 0).cpu().numpy()) <= 1e-10 * (
        np.linalg.norm(x_true))


@pytest.mark.parametrize("p,q", [(1, 1), (2, 2), (3, 4), (6, 5)])
def test_band_lu_factor_reports_interior_zero_pivot(dev, p, q):
    """A pivot that becomes exactly zero in the middle of the elimination (the
    2 x 2 all-ones block at rows 5-6) is reported with its row, as
    DiaMatrix.lu_factor raises (la/diamatrix.py:461-471)."""
    from jaxfun_amd import kernels as K

    n = 12
    A = 2.0 * np.eye(n)
    A[5:7, 5:7] = 1.0
    band = np.zeros((n, p + q + 1))
    for i in range(n):
        for o in range(-p, q + 1):
            if 0 <= i + o < n:
                band[i, o + p] = A[i, i + o]
    with pytest.raises(ValueError, match="row 6"):
        K.band_lu_factor(_t(band, dev), p, q)
    with pytest.raises(ValueError):
        K.band_lu_factor_batched(_t(np.stack([band, band]), dev), p, q)


def test_band_solve_wave_rejects_wide_bands(dev):
    from jaxfun_amd import kernels as K

    A_band = _banded(np.random.default_rng(1), 40, 5, 1)
    lu = K.band_lu_factor(_t(A_band, dev), 5, 1)
    x = _t(np.ones(40), dev)
    with pytest.raises(ValueError):
        K.band_lu_solve(lu, 5, 1, x, 0, method="wave")
    with pytest.raises(ValueError):
        K.band_lu_solve(lu, 5, 1, x, 0, method="scan")


def test_elementwise(dev):
    import torch

    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(4)
    a, b, c = (rng.standard_normal((5, 6)) for _ in range(3))
    y = K.lincomb([(2.0, _t(a, dev)), (-0.5, _t(b, dev))]).cpu().numpy()
    assert np.allclose(y, 2 * a - 0.5 * b)
    D = rng.standard_normal(6) + 3
    z = K.div_bcast(_t(a, dev), _t(D, dev)).cpu().numpy()
    assert np.allclose(z, a / D)
    f = K.cast(_t(c, dev), torch.float32)
    assert f.dtype == torch.float32 and np.allclose(f.cpu().numpy(), c.astype(np.float32))


@pytest.mark.parametrize("dims", [(128, 128, 128, 128), (62, 64, 62, 64), (96, 30, 17, 128), (1, 5, 7, 3),
                                  (80, 72, 64, 50), (20, 40, 100, 33)])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-14), ("bf16x6", 2e-6)])
@pytest.mark.parametrize("epi", [None, "div", "mul"])
def test_plane_contract(dev, dims, prec, tol, epi):
    import torch

    from jaxfun_amd import kernels as K

    M0, K0, M1, K1 = dims
    rng = np.random.default_rng(7)
    A0, A1 = rng.standard_normal((M0, K0)), rng.standard_normal((M1, K1))
    X = rng.standard_normal((5, K0, K1))
    E = rng.uniform(1.0, 2.0, size=(M0, M1))
    dt = K.io_dtype(prec)
    Xd = _t(X, dev, dt)
    ep = None if epi is None else (epi, _t(E, dev, dt))
    Y = K.plane(K.AxisOperator(_t(A0, dev)), K.AxisOperator(_t(A1, dev)), Xd, prec, ep)
    Xr = Xd.to(torch.float64).cpu().numpy()
    Yr = np.einsum("mk,bkl,nl->bmn", A0, Xr, A1)
    if epi == "div":
        Yr = Yr / E
    elif epi == "mul":
        Yr = Yr * E
    assert O.rel_l2(Y.to(torch.float64).cpu().numpy(), Yr) < tol


@pytest.mark.parametrize("prec,tol", [("f64", 1e-14), ("bf16x6", 2e-6)])
@pytest.mark.parametrize("dims", [(126, 126, 126, 126), (30, 100, 70, 20)])
def test_plane_alpha_accumulate(dev, prec, tol, dims):
    # Y = alpha A0 X A1^T (/ E), then += beta-free accumulation of a second term into the same out
    import torch

    from jaxfun_amd import kernels as K

    M0, K0, M1, K1 = dims
    rng = np.random.default_rng(11)
    A0, A1, B0, B1 = (rng.standard_normal(s) for s in ((M0, K0), (M1, K1), (M0, K0), (M1, K1)))
    E = rng.uniform(1.0, 2.0, size=(M0, M1))
    X = rng.standard_normal((4, K0, K1))
    dt = K.io_dtype(prec)
    Xd, Ed = _t(X, dev, dt), _t(E, dev, dt)
    ops = [K.AxisOperator(_t(a, dev)) for a in (A0, A1, B0, B1)]
    Y = K.plane(ops[0], ops[1], Xd, prec, ("div", Ed), alpha=-0.5)
    Y2 = K.plane(ops[2], ops[3], Xd, prec, alpha=3.0, out=Y, accumulate=True)
    assert Y2.data_ptr() == Y.data_ptr()
    Xr = Xd.to(torch.float64).cpu().numpy()
    Yr = -0.5 * np.einsum("mk,bkl,nl->bmn", A0, Xr, A1) / E + 3.0 * np.einsum("mk,bkl,nl->bmn", B0, Xr, B1)
    assert O.rel_l2(Y.to(torch.float64).cpu().numpy(), Yr) < tol
    with pytest.raises(ValueError):
        K.plane(ops[0], ops[1], Xd, prec, accumulate=True)


@pytest.mark.parametrize("prec,n", [("f64", 5), ("f64", 17), ("f64", 62), ("f64", 64), ("bf16x6", 30),
                                    ("bf16x6", 62), ("bf16x6", 126), ("bf16x6", 128)])
@pytest.mark.parametrize("B", [1, 3, 300])
def test_plane_fd_fused(dev, prec, n, B):
    # Y = V0 ((W0 X W1^T) / E) V1^T in one launch against the two-launch path and einsum
    import torch

    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(n + B)
    W0, W1, V0, V1 = (rng.standard_normal((n, n)) for _ in range(4))
    E = rng.uniform(1.0, 2.0, size=(n, n))
    X = rng.standard_normal((B, n, n))
    dt = K.io_dtype(prec)
    ops = [K.AxisOperator(_t(a, dev)) for a in (W0, W1, V0, V1)]
    Xd, Ed = _t(X, dev, dt), _t(E, dev, dt)
    assert K.plane_fd_ok(ops, prec)
    Y = K.plane_fd(*ops, Xd, Ed, prec)
    Y2 = K.plane(ops[2], ops[3], K.plane(ops[0], ops[1], Xd, prec, ("div", Ed)), prec)
    Xr = Xd.to(torch.float64).cpu().numpy()
    Yr = V0 @ ((W0 @ Xr @ W1.T) / E) @ V1.T
    tol = 1e-13 if prec == "f64" else 4e-6
    assert O.rel_l2(Y.to(torch.float64).cpu().numpy(), Yr) < tol
    assert O.rel_l2(Y.to(torch.float64).cpu().numpy(), Y2.to(torch.float64).cpu().numpy()) < tol


@pytest.mark.parametrize("prec,tol", [("f64", 1e-14), ("bf16x6", 2e-6)])
def test_plane_contract_many_slabs(dev, prec, tol):
    # more slabs than resident workgroups: every workgroup walks several slabs
    import torch

    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(8)
    A0, A1 = rng.standard_normal((64, 62)), rng.standard_normal((64, 62))
    X = rng.standard_normal((700, 62, 62))
    Xd = _t(X, dev, K.io_dtype(prec))
    Y = K.plane(K.AxisOperator(_t(A0, dev)), K.AxisOperator(_t(A1, dev)), Xd, prec)
    Yr = A0 @ Xd.to(torch.float64).cpu().numpy() @ A1.T
    assert O.rel_l2(Y.to(torch.float64).cpu().numpy(), Yr) < tol


@pytest.mark.parametrize("prec,tol", [("f64", 1e-14), ("f32", 2e-6), ("bf16x6", 2e-6)])
@pytest.mark.parametrize("M", [16, 126, 192, 256])
@pytest.mark.parametrize("Kc", [130, 132, 192, 256])
@pytest.mark.parametrize("shape,axis", [((3, None, 64), 1), ((2, None, 192), 1), ((None, 128), 0),
                                        ((5, 33, None), 2), ((70, None), 1)])
def test_contract_k_split(dev, prec, tol, M, Kc, shape, axis):
    # 128 < K <= 256 with Q % 64 == 0 or Q == 1 (the padded forward
    # transforms): two column halves of the operator, the second accumulating
    # into the output (f32 with K % 4 != 0 on the last axis: general kernel)
    import torch

    from jaxfun_amd import kernels as K

    shape = tuple(Kc if d is None else d for d in shape)
    rng = np.random.default_rng(M + Kc + axis)
    A = rng.standard_normal((M, Kc))
    X = rng.standard_normal(shape)
    dt = K.io_dtype(prec)
    Y = K.apply_along(K.AxisOperator(_t(A, dev)), _t(X, dev, dt), axis, prec)
    Xr = _t(X, dev, dt).to(torch.float64).cpu().numpy()
    ref = np.moveaxis(np.tensordot(A, Xr, axes=([1], [axis])), 0, axis)
    assert Y.shape == ref.shape
    assert O.rel_l2(Y.to(torch.float64).cpu().numpy(), ref) < tol


@pytest.mark.parametrize("prec,tol", [("f64", 1e-14), ("f32", 2e-6), ("bf16x6", 2e-6)])
@pytest.mark.parametrize("M", [16, 96, 128, 192])
@pytest.mark.parametrize("Kc", [10, 62, 126])
def test_contract_last_axis_even_k(dev, prec, tol, M, Kc):
    # last-axis contractions (Q == 1) with K even but not a multiple of 4 (8 for
    # bf16x6): the register-resident kernels' k-steps past K read zeros, and
    # bf16x6 reads such rows (8-byte aligned only) in pairs
    import torch

    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(M * 7 + Kc)
    A = rng.standard_normal((M, Kc))
    X = rng.standard_normal((5, 33, Kc))
    dt = K.io_dtype(prec)
    Y = K.apply_along(K.AxisOperator(_t(A, dev)), _t(X, dev, dt), 2, prec)
    ref = _t(X, dev, dt).to(torch.float64).cpu().numpy() @ A.T
    assert Y.shape == ref.shape
    assert O.rel_l2(Y.to(torch.float64).cpu().numpy(), ref) < tol


@pytest.mark.parametrize("prec,tol", [("f64", 1e-14), ("f32", 2e-6), ("bf16x6", 2e-6)])
@pytest.mark.parametrize("M", [132, 192, 256])
@pytest.mark.parametrize("shape,axis", [((3, 126, 64), 1), ((5, 7, 126), 2), ((126, 128), 0)])
def test_contract_row_split(dev, prec, tol, M, shape, axis):
    # operators with 128 < M <= 256 rows (the 3/2-padded backward transforms)
    # run as two register-resident launches over row halves
    import torch

    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(M + axis)
    A = rng.standard_normal((M, 126))
    X = rng.standard_normal(shape)
    dt = K.io_dtype(prec)
    Y = K.apply_along(K.AxisOperator(_t(A, dev)), _t(X, dev, dt), axis, prec)
    Xr = _t(X, dev, dt).to(torch.float64).cpu().numpy()
    ref = np.moveaxis(np.tensordot(A, Xr, axes=([1], [axis])), 0, axis)
    assert Y.shape == ref.shape
    assert O.rel_l2(Y.to(torch.float64).cpu().numpy(), ref) < tol


@pytest.mark.parametrize("M", [130, 192, 200, 256])
@pytest.mark.parametrize("Kc", [30, 64, 126, 128])
@pytest.mark.parametrize("shape,axis", [((70, None), 1), ((3, None, 64), 1), ((2, None, 192), 1), ((5, 33, None), 2)])
def test_contract_bf16x6_two_row_blocks(dev, M, Kc, shape, axis):
    """bf16x6 with 128 < M <= 256 and K <= 128: one launch, each wave holding
    two 16-row blocks of the operator (the 3/2-padded backward transforms)."""
    import torch

    from jaxfun_amd import kernels as K

    shape = tuple(Kc if d is None else d for d in shape)
    rng = np.random.default_rng(M + 7 * Kc + axis)
    A = rng.standard_normal((M, Kc))
    X = rng.standard_normal(shape)
    Xd = _t(X, dev, torch.float32)
    Y = K.apply_along(K.AxisOperator(_t(A, dev)), Xd, axis, "bf16x6")
    Xr = Xd.to(torch.float64).cpu().numpy()
    ref = np.moveaxis(np.tensordot(A, Xr, axes=([1], [axis])), 0, axis)
    assert Y.shape == ref.shape
    assert O.rel_l2(Y.to(torch.float64).cpu().numpy(), ref) < 2e-6


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_sum_products(dev, dtype):
    """out = sum_t c_t a_t b_t, with the shared-first-factor path
    (u u_x + u u_y reads u once) and the general path."""
    import torch

    from jaxfun_amd import kernels as K

    dt = torch.float64 if dtype == "f64" else torch.float32
    tol = 1e-14 if dtype == "f64" else 2e-6
    rng = np.random.default_rng(21)
    u, ux, uy, v = (rng.standard_normal((3, 50, 70)) for _ in range(4))
    tu, tux, tuy, tv = (_t(a, dev, dt) for a in (u, ux, uy, v))
    w = K.sum_products([(-1.0, tu, tux), (-1.0, tu, tuy)]).to(torch.float64).cpu().numpy()
    assert O.rel_l2(w, -(u * ux + u * uy)) < tol
    w = K.sum_products([(2.0, tu, tux), (-0.5, tv, tuy), (1.5, tux, tuy)]).to(torch.float64).cpu().numpy()
    assert O.rel_l2(w, 2 * u * ux - 0.5 * v * uy + 1.5 * ux * uy) < tol
    w = K.sum_products([(3.0, tu, tv)]).to(torch.float64).cpu().numpy()
    assert O.rel_l2(w, 3 * u * v) < tol


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("nin", [1, 2, 3, 5, 8, 11])
def test_lincomb_counts(dev, dtype, nin):
    """y = sum_j c_j x_j for every unrolled input count (more than 8 inputs go
    in chunks of 8 accumulating into y), with and without accumulation."""
    import torch

    from jaxfun_amd import kernels as K

    dt = torch.float64 if dtype == "f64" else torch.float32
    tol = 1e-14 if dtype == "f64" else 2e-6
    rng = np.random.default_rng(nin)
    xs = [rng.standard_normal(1001) for _ in range(nin)]
    cs = rng.standard_normal(nin)
    y = K.lincomb([(float(c), _t(x, dev, dt)) for c, x in zip(cs, xs)]).to(torch.float64).cpu().numpy()
    ref = sum(c * x for c, x in zip(cs, xs))
    assert O.rel_l2(y, ref) < tol
    y0 = rng.standard_normal(1001)
    out = _t(y0, dev, dt)
    K.lincomb([(float(c), _t(x, dev, dt)) for c, x in zip(cs, xs)], out=out, accumulate=True)
    assert O.rel_l2(out.to(torch.float64).cpu().numpy(), ref + y0) < tol


@pytest.mark.parametrize("method", ["wave", "lanes"])
def test_band_solve_complex_rhs(dev, method):
    """A complex right-hand side goes through its real view (re/im as two
    lines sharing the real factor) on either kernel."""
    import scipy.linalg as sla

    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(17)
    n, p, q = 500, 2, 2
    A_band = _banded(rng, n, p, q)
    lu = K.band_lu_factor(_t(A_band, dev), p, q)
    B = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    import torch

    X = K.band_lu_solve(lu, p, q, torch.tensor(B, device=dev), 0, method=method).cpu().numpy()
    Xr = sla.solve_banded((p, q), _band_to_ab(A_band, p, q), B)
    assert np.abs(X - Xr).max() <= 1e-13 * np.abs(Xr).max()
