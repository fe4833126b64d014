"""The reference's TPMatrices solver tests on the GPU path
(tests/la/test_tpmatrices_solvers.py:85-330): factor types and dispatch,
diagonalisation, per-wavenumber and dense Kronecker solves against each
other and against manufactured solutions, with the reference's tolerances
(FP64)."""

import numpy as np
import pytest
import sympy as sp

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps
BCS = {"left": {"D": 0}, "right": {"D": 0}}
POLY = ["Legendre", "Chebyshev"]


def _np(t):
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


def _cls(name):
    import jaxfun_amd.galerkin as G

    return getattr(G, name)


def _poisson_poly2d(N, poly, sparse=True):
    from jaxfun_amd import Div, FunctionSpace, Grad, TensorProduct, TestFunction, TrialFunction, inner

    T = TensorProduct(FunctionSpace(N, _cls(poly), BCS), FunctionSpace(N, _cls(poly), BCS))
    v, u = TestFunction(T), TrialFunction(T)
    x, y = T.system.base_scalars()
    ue = (1 - x**2) * (1 - y**2)
    A, b = inner(v * Div(Grad(u)) - v * Div(Grad(ue)), sparse=sparse, kind="system")
    return T, A, b, ue


def _poisson_fourier_poly_2d(N, poly):
    from jaxfun_amd import Div, Fourier, FunctionSpace, Grad, TensorProduct, TestFunction, TrialFunction, inner

    T = TensorProduct(FunctionSpace(N, Fourier), FunctionSpace(N, _cls(poly), BCS))
    v, u = TestFunction(T), TrialFunction(T)
    x, y = T.system.base_scalars()
    ue = sp.cos(2 * x) * (1 - y**2)
    A, b = inner(v * Div(Grad(u)) - v * Div(Grad(ue)), sparse=True, kind="system")
    return T, A, b, ue


def _l2(T, uh, ue, M=40):
    x = T.system.base_scalars()
    uj = _np(T.backward(uh, N=(M,) * len(x)))
    xj = [m.cpu().numpy() for m in T.mesh(N=(M,) * len(x), broadcast=True)]
    return float(np.linalg.norm(uj - sp.lambdify(x, ue)(*xj))) / M


@pytest.mark.parametrize("poly", POLY)
def test_lu_factor_types_and_dispatch(dev, poly):
    # test_tpmatrices_solvers.py:85-116, 169-175, 191-198, 222-229, 245-270
    from jaxfun_amd.la import TPMatricesDenseLUFactors, TPMatricesLUFactors, TPMatricesWavenumberSolver, \
        tpmats_dense_lu_factor, tpmats_lu_factor, tpmats_wavenumber_factor

    _, A, b, _ = _poisson_poly2d(12, poly)
    assert isinstance(tpmats_lu_factor(A.tpmats), TPMatricesLUFactors)
    assert isinstance(tpmats_lu_factor(A.tpmats[0]), TPMatricesLUFactors)
    assert isinstance(A.lu_factor(), TPMatricesLUFactors) and A.lu_factor() is A.lu_factor()
    assert tuple(A.solve(b).shape) == tuple(b.shape)
    _, Ad, bd, _ = _poisson_poly2d(12, poly, sparse=False)
    assert isinstance(tpmats_dense_lu_factor(Ad.tpmats), TPMatricesDenseLUFactors)
    assert isinstance(Ad.lu_factor(), TPMatricesDenseLUFactors)
    with pytest.raises(TypeError):
        tpmats_dense_lu_factor(A.tpmats)
    _, Af, bf, _ = _poisson_fourier_poly_2d(12, poly)
    assert isinstance(tpmats_wavenumber_factor(Af), TPMatricesWavenumberSolver)
    assert isinstance(Af.lu_factor(), TPMatricesWavenumberSolver) and Af.lu_factor() is Af.lu_factor()
    with pytest.raises(TypeError):
        tpmats_wavenumber_factor("not a valid input")


@pytest.mark.parametrize("poly", POLY)
def test_diagonalisation_and_dense_solves(dev, poly):
    # test_tpmatrices_solvers.py:92-104, 176-188, 201-207, 273-284
    from jaxfun_amd.la import tpmats_dense_lu_factor, tpmats_lu_factor

    T, A, b, ue = _poisson_poly2d(16, poly)
    uh = tpmats_lu_factor(A.tpmats).solve(b)
    assert tuple(uh.shape) == tuple(b.shape)
    assert _l2(T, uh, ue) < 100 * EPS
    _, Ad, bd, _ = _poisson_poly2d(16, poly, sparse=False)
    ud = tpmats_dense_lu_factor(Ad.tpmats).solve(bd)
    assert _l2(T, ud, ue) < 100 * EPS
    assert np.abs(_np(uh) - _np(ud)).max() < 100 * EPS
    assert _l2(T, A.solve(b), ue) < 100 * EPS


@pytest.mark.parametrize("poly", POLY)
def test_single_term_round_trip(dev, poly):
    # test_tpmatrices_solvers.py:231-242: A0 (A0^-1 b) = b for one Kronecker term
    from jaxfun_amd.la import TPLUFactors

    _, A, b, _ = _poisson_poly2d(16, poly)
    tp = A.tpmats[0]
    assert isinstance(tp.lu_factor(), TPLUFactors)
    rhs = tp(tp.lu_factor().solve(b))
    assert np.abs(_np(rhs) - _np(b)).max() < 100 * EPS


@pytest.mark.parametrize("poly", POLY)
def test_wavenumber_solves(dev, poly):
    # test_tpmatrices_solvers.py:126-146, 287-302
    from jaxfun_amd.la import tpmats_to_kron, tpmats_wavenumber_factor

    T, A, b, ue = _poisson_fourier_poly_2d(16, poly)
    wn = tpmats_wavenumber_factor(A)
    uh = wn.solve(b)
    ref = np.linalg.solve(tpmats_to_kron(A.tpmats), _np(b).ravel()).reshape(b.shape)
    assert np.abs(_np(uh) - ref).max() < 10 * EPS * max(1.0, np.abs(ref).max())
    assert _l2(T, uh, ue) < np.sqrt(10 * EPS)
    assert _l2(T, A.solve(b), ue) < np.sqrt(10 * EPS)


def test_fourier_fourier_legendre_3d(dev):
    # test_tpmatrices_solvers.py:305-330
    from jaxfun_amd import Div, Fourier, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, \
        TrialFunction, inner
    from jaxfun_amd.la import TPMatricesWavenumberSolver

    N = 8
    T = TensorProduct(FunctionSpace(N, Fourier), FunctionSpace(N, Fourier), FunctionSpace(N, Legendre, BCS))
    v, u = TestFunction(T), TrialFunction(T)
    x, y, z = T.system.base_scalars()
    ue = sp.cos(2 * x) * sp.cos(2 * y) * (1 - z**2)
    A, b = inner(v * Div(Grad(u)) - v * Div(Grad(ue)), sparse=True, kind="system")
    assert isinstance(A.lu_factor(), TPMatricesWavenumberSolver)
    uh = A.solve(b)
    assert tuple(uh.shape) == tuple(b.shape)
    assert _l2(T, uh, ue, M=20) < np.sqrt(1000 * EPS)
