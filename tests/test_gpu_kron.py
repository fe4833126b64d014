"""The Kronecker-solve path of TPMatrices.solve (reference la/tpmatrix.py:429-509
with DiaMatrix.lu_solve's method choice, la/diamatrix.py:592-770): the wide
band LU kernels against dense solves, the device Kronecker assembly against
numpy.kron, and every kron_method on operators built by inner()."""

import warnings

import numpy as np
import pytest

from oracle import spectral as O

pytestmark = pytest.mark.gpu


def _band(A, p, q):
    n = A.shape[0]
    band = np.zeros((n, p + q + 1))
    for o in range(-p, q + 1):
        i = np.arange(max(0, -o), min(n, n - o))
        band[i, o + p] = A[i, i + o]
    return band


@pytest.mark.parametrize("n,p,q", [(1, 0, 0), (7, 2, 1), (500, 3, 5), (300, 40, 17), (2000, 126, 126)])
def test_band_lu_wide_vs_dense(dev, n, p, q):
    from jaxfun_amd.la.tpmatrix import TPMatricesBandedLUFactors

    rng = np.random.default_rng(n + p + q)
    A = np.zeros((n, n))
    for o in range(-p, q + 1):
        i = np.arange(max(0, -o), min(n, n - o))
        A[i, i + o] = rng.standard_normal(len(i))
    A += np.diag(np.abs(A).sum(1) + 1.0)  # diagonally dominant: no pivoting needed
    f = TPMatricesBandedLUFactors(_band(A, p, q), p, q, (n,))
    B = rng.standard_normal((3, n))
    X = f.solve(B).cpu().numpy()
    Xr = np.linalg.solve(A, B.T).T
    assert O.rel_l2(X, Xr) < 1e-12


def test_band_lu_wide_zero_pivot(dev):
    from jaxfun_amd.la.tpmatrix import TPMatricesBandedLUFactors

    A = np.eye(6)
    A[3, 3] = 0.0
    with pytest.raises(ValueError, match="row 3"):
        TPMatricesBandedLUFactors(_band(A, 1, 1), 1, 1, (6,))


@pytest.mark.parametrize("d", [2, 3])
def test_kron_device_vs_numpy(dev, d):
    from jaxfun_amd.la.matrix import Matrix
    from jaxfun_amd.la.tpmatrix import TPMatrix, kron_device, tpmats_to_kron

    rng = np.random.default_rng(d)
    sizes = [5, 7, 3][:d]
    tps = [TPMatrix([Matrix(rng.standard_normal((m, m))) for m in sizes], s) for s in (1.5, -0.25)]
    assert O.rel_l2(kron_device(tps).cpu().numpy(), tpmats_to_kron(tps)) < 1e-15


@pytest.fixture(scope="module")
def legendre_helmholtz(dev):
    from jaxfun_amd import Div, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, TrialFunction, inner

    N = 24
    D = FunctionSpace(N, Legendre, {"left": {"D": 0}, "right": {"D": 0}})
    T = TensorProduct(D, D)
    v, u = TestFunction(T), TrialFunction(T)
    A = inner(v * (2.0 * u - Div(Grad(u))), sparse=True)
    sp1 = O.Space1D("legendre", N, True)
    M, Kd = sp1.mass(), O.legendre_dirichlet_stiffness(N)
    terms = [(2.0, [M, M]), (-1.0, [Kd, M]), (-1.0, [M, Kd])]
    return A, terms, N - 2


@pytest.mark.parametrize("kron_method,threshold,kind", [
    ("banded", 100000, "banded"), ("banded", 100, "banded"), ("rcm", 100, "rcm"), ("dense", 100, "dense"),
    ("auto", 100000, "banded"), ("auto", 100, None)])
def test_kron_methods_vs_exact(legendre_helmholtz, kron_method, threshold, kind):
    from jaxfun_amd.la.matrix import DiaMatrix
    from jaxfun_amd.la.tpmatrix import TPMatricesBandedLUFactors, TPMatricesDenseLUFactors

    A, terms, n = legendre_helmholtz
    assert all(isinstance(m, DiaMatrix) for tp in A.tpmats for m in tp.mats)
    A._kron_cache = None  # each case factors afresh
    b = np.random.default_rng(7).standard_normal((n, n))
    with warnings.catch_warnings(record=True) as rec:
        warnings.simplefilter("always")
        x = A.solve(b, method="kron", kron_method=kron_method, auto_threshold=threshold).cpu().numpy()
    assert O.rel_l2(x, O.kron_solve(terms, b)) < 1e-11
    f = A.kron_factor(kron_method, threshold)
    if kind == "banded":
        assert isinstance(f, TPMatricesBandedLUFactors) and f.perm is None
    elif kind == "rcm":
        assert isinstance(f, TPMatricesBandedLUFactors) and f.perm is not None
    elif kind == "dense":
        assert isinstance(f, TPMatricesDenseLUFactors)
    if kron_method == "banded" and threshold == 100:
        assert any("exceeds" in str(w.message) for w in rec)  # the reference warns on a wide band
    # batched and complex right-hand sides reuse the factors
    bb = np.stack([b, -2 * b])
    xb = A.solve(bb, method="kron", kron_method=kron_method, auto_threshold=threshold).cpu().numpy()
    assert O.rel_l2(xb[1], -2 * x) < 1e-12
    xc = A.solve(b + 1j * b, method="kron", kron_method=kron_method, auto_threshold=threshold).cpu().numpy()
    assert O.rel_l2(xc, (1 + 1j) * x) < 1e-12


def test_kron_method_rejects_unknown(legendre_helmholtz):
    A, _, n = legendre_helmholtz
    with pytest.raises(ValueError):
        A.solve(np.zeros((n, n)), method="kron", kron_method="cholesky")


def test_kron_banded_poisson64_vs_fd(dev):
    # the Kronecker matrix of the N = 64 Legendre Poisson operator (3844
    # unknowns, bandwidth 124 each side) by the banded path, against FD
    from jaxfun_amd import Div, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, TrialFunction, inner

    N = 64
    D = FunctionSpace(N, Legendre, {"left": {"D": 0}, "right": {"D": 0}})
    T = TensorProduct(D, D)
    v, u = TestFunction(T), TrialFunction(T)
    A = inner(v * Div(Grad(u)), sparse=True)
    b = np.random.default_rng(8).standard_normal((N - 2, N - 2))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        xk = A.solve(b, method="kron", kron_method="banded").cpu().numpy()
    xf = A.solve(b, method="fd").cpu().numpy()
    assert O.rel_l2(xk, xf) < 1e-11
    f = A.kron_factor("banded")
    assert (f.p, f.q) == (124, 124)  # K diagonal: offsets +-2 (K (x) M) and +-2 n (M (x) K)
