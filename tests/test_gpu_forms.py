"""Forms with known functions (JAXFunction) and non-separable coefficients
(TensorMatrix), against the oracle and the reference's own identities
(tests/galerkin/test_jaxfunction.py, the JAXFunction doctest of
galerkin/arguments.py:726-738, la/tensormatrix.py)."""

import numpy as np
import pytest
import sympy as sp

from oracle import spectral as O

pytestmark = pytest.mark.gpu

SPACES = ["legendre", "chebyshev", "chebyshevu", "jacobi", "ultraspherical"]
DOMAINS = [None, (-2.0, 2.0)]


def _np(t):
    return t.detach().cpu().numpy()


def _space(name):
    from jaxfun_amd import Chebyshev, ChebyshevU, Fourier, Jacobi, Legendre, Ultraspherical

    return {"legendre": Legendre, "chebyshev": Chebyshev, "fourier": Fourier, "chebyshevu": ChebyshevU,
            "jacobi": Jacobi, "ultraspherical": Ultraspherical}[name]


def _close(a, b, tol):
    a, b = _np(a) if hasattr(a, "detach") else np.asarray(a), _np(b) if hasattr(b, "detach") else np.asarray(b)
    return np.linalg.norm(a - b) < tol * max(1.0, np.linalg.norm(b))


def test_jaxfunction_doctest_golden(dev):
    # galerkin/arguments.py:726-738: Chebyshev(4), w = ones
    from jaxfun_amd import Chebyshev, TestFunction, TrialFunction, inner
    from jaxfun_amd.galerkin.arguments import JAXFunction

    V = Chebyshev(4, name="V")
    w = JAXFunction(np.ones(V.dim), V, name="w")
    v = TestFunction(V, name="v")
    b = _np(inner(v * w))
    assert np.allclose(b, [np.pi, np.pi / 2, np.pi / 2, np.pi / 2], atol=1e-14)
    u = TrialFunction(V, name="u")
    a = inner(u * v)
    assert np.allclose(_np(a @ w.array), b, atol=1e-14)
    assert np.allclose(_np(a @ w), b, atol=1e-14)


@pytest.mark.parametrize("space", SPACES + ["fourier"])
@pytest.mark.parametrize("domain", DOMAINS)
def test_jaxfunction_linear(dev, space, domain):
    # tests/galerkin/test_jaxfunction.py:66-101
    from jaxfun_amd import TestFunction, TrialFunction, inner
    from jaxfun_amd.galerkin.arguments import JAXFunction

    N = 8
    D = _space(space)(N, domain=domain)
    x = D.system.x
    u, v = TrialFunction(D), TestFunction(D)
    uf = JAXFunction(np.ones(N), D)
    A = inner(u * v, kind="bilinear")
    assert _close(A @ uf, inner(v * uf, kind="linear"), 1e-13)
    if space != "fourier":
        A = inner(sp.diff(u, x) * v, kind="bilinear")
        assert _close(A @ uf, inner(v * sp.diff(uf, x), kind="linear"), 1e-13)


@pytest.mark.parametrize("space", SPACES)
@pytest.mark.parametrize("domain", DOMAINS)
def test_jaxfunction_nonlinear(dev, space, domain):
    # tests/galerkin/test_jaxfunction.py:104-141
    from jaxfun_amd import TestFunction, TrialFunction, inner
    from jaxfun_amd.galerkin.arguments import JAXFunction

    for N in (8, 6):
        D = _space(space)(N, domain=domain)
        x = D.system.x
        u, v = TrialFunction(D), TestFunction(D)
        uf = JAXFunction(np.linspace(0.5, 1.5, N), D)
        A = inner(u * v * uf, kind="bilinear")
        assert _close(A @ uf, inner(v * uf**2, kind="linear"), 1e-12)
        A = inner(sp.diff(u, x) * v * sp.diff(uf, x), kind="bilinear")
        assert _close(A @ uf, inner(v * sp.diff(uf, x) ** 2, kind="linear"), 1e-12)
    # the weighted matrix against quadrature in numpy (oracle Vandermonde)
    N = 8
    D = _space(space)(N, domain=domain)
    u, v = TrialFunction(D), TestFunction(D)
    arr = np.linspace(0.5, 1.5, N)
    uf = JAXFunction(arr, D)
    xq, wq = O.quad(space, N)
    V = O.vandermonde(space, xq, N)
    df = 1.0 if domain is None else 2.0 / (domain[1] - domain[0])
    ref = (V.T * (wq * (V @ arr) / df)) @ V
    assert np.allclose(inner(u * v * uf).todense_host(), ref, atol=1e-13)


@pytest.mark.parametrize("space", SPACES)
@pytest.mark.parametrize("domain", DOMAINS)
def test_jaxfunction_directsum(dev, space, domain):
    # tests/galerkin/test_jaxfunction.py:78-88, 116-128
    from jaxfun_amd import FunctionSpace, TestFunction, TrialFunction, inner
    from jaxfun_amd.galerkin.arguments import JAXFunction

    N = 8
    D = FunctionSpace(N, _space(space), bcs={"left": {"D": 1}, "right": {"D": 1}}, domain=domain, name="D")
    u, v = TrialFunction(D), TestFunction(D)
    A, b = inner(u * v, kind="system")
    uf = JAXFunction(np.ones(D.num_dofs), D)
    b0 = _np(A @ uf) - _np(b)
    assert _close(b0, inner(v * uf, kind="linear"), 1e-13)
    A, b = inner(u * v * uf, kind="system")
    b0 = _np(A @ uf) - _np(b)
    assert _close(b0, inner(v * uf**2, kind="linear"), 1e-12)


@pytest.mark.parametrize("space", SPACES)
@pytest.mark.parametrize("domain", DOMAINS)
def test_jaxfunction_2d(dev, space, domain):
    # tests/galerkin/test_jaxfunction.py:144-216
    from jaxfun_amd import Div, Dot, Grad, TensorProduct, TestFunction, TrialFunction, inner
    from jaxfun_amd.galerkin.arguments import JAXFunction

    N = 8
    D = _space(space)(N, domain=domain)
    T = TensorProduct(D, D, name="T")
    u, v = TrialFunction(T), TestFunction(T)
    A = inner(u * v, kind="bilinear")
    arr = np.random.default_rng(0).random((N, N)) + 0.5
    uf = JAXFunction(arr, T)
    b0 = A.todense_host() @ arr.ravel()
    b1 = _np(A.mats[0] @ uf @ A.mats[1].T)
    assert np.linalg.norm(b0 - b1.ravel()) < 1e-12 * np.linalg.norm(b0)
    assert np.linalg.norm(_np(inner(uf * v, kind="linear")).ravel() - b0) < 1e-12 * np.linalg.norm(b0)
    A = inner(Div(Grad(u)) * v, kind="bilinear")
    assert _close(A @ uf, inner(v * Div(Grad(uf)), kind="linear"), 1e-12)
    n = 4
    D4 = _space(space)(n, domain=domain)
    T4 = TensorProduct(D4, D4, name="T")
    u, v = TrialFunction(T4), TestFunction(T4)
    w = JAXFunction(np.ones((n, n)), T4, name="w")
    A = inner(u * v * w, kind="bilinear")
    assert _close(A @ w, inner(v * w**2, kind="linear"), 1e-12)
    A = inner(w * Dot(Grad(v), Grad(u)), kind="bilinear")
    assert _close(A @ w, inner(0.5 * Dot(Grad(v), Grad(w**2)), kind="linear"), 1e-12)


@pytest.mark.parametrize("space", SPACES)
def test_jaxfunction_directsum_2d(dev, space):
    # tests/galerkin/test_jaxfunction.py:159-174
    from jaxfun_amd import Div, FunctionSpace, Grad, TensorProduct, TestFunction, TrialFunction, inner
    from jaxfun_amd.galerkin.arguments import JAXFunction

    N = 8
    D = FunctionSpace(N, _space(space), bcs={"left": {"D": 1}, "right": {"D": 1}}, name="D")
    T = TensorProduct(D, D, name="T")
    u, v = TrialFunction(T), TestFunction(T)
    A, b = inner(u * v, kind="system")
    w = JAXFunction(np.ones(T.num_dofs), T, name="uf")
    assert _close(_np(A @ w) - _np(b), inner(v * w, kind="linear"), 1e-12)
    C, c = inner(Div(Grad(u)) * v, kind="system")
    assert _close(_np(C @ w) - _np(c), inner(v * Div(Grad(w)), kind="linear"), 1e-12)


@pytest.mark.parametrize("space", ["legendre", "chebyshev"])
def test_tensormatrix_mixed_solve(dev, space):
    # separable + non-separable terms in one operator: -Lap u + u / (3 + x + y) = f
    from jaxfun_amd import Div, FunctionSpace, Grad, TensorProduct, TestFunction, TrialFunction, inner
    from jaxfun_amd.galerkin.inner import TensorMatrix

    N = 14
    D = FunctionSpace(N, _space(space), {"left": {"D": 0}, "right": {"D": 0}})
    T = TensorProduct(D, D)
    u, v = TrialFunction(T), TestFunction(T)
    x, y = T.system.base_scalars()
    k = 1 / (3 + x + y)
    ue = (1 - x**2) * (1 - y**2) * (1 + x * y)
    A, b = inner(-v * Div(Grad(u)) + v * k * u - v * (-Div(Grad(ue)) + k * ue), kind="system")
    assert isinstance(A, TensorMatrix) and len(A.separable) == 2
    # the dense operator against numpy quadrature with the oracle's matrices
    S = O.dirichlet_stencil(N)
    xq, wq = O.quad(space, N)
    V = O.vandermonde(space, xq, N) @ S.T
    M = O.Space1D(space, N, True).mass() if space == "legendre" else (V.T * wq) @ V
    Kx = O.galerkin_matrix(space, N, 0, 2)
    kq = 1 / (3 + np.add.outer(xq, xq))
    Tw = np.einsum("pi,pk,qj,ql,pq->ijkl", V * wq[:, None], V, V * wq[:, None], V, kq)
    dense = -(np.kron(Kx, M) + np.kron(M, Kx)) + Tw.reshape((N - 2) ** 2, (N - 2) ** 2)
    assert np.allclose(_np(A.todense()), dense, atol=1e-12 * np.abs(dense).max())
    uh = A.solve(b)
    xs = np.linspace(-1, 1, 9)
    got = O.apply_along(O.vandermonde(space, xs, N) @ S.T, O.apply_along(O.vandermonde(space, xs, N) @ S.T,
                                                                            _np(uh), 1), 0)
    ref = sp.lambdify((x, y), ue)(xs[:, None], xs[None, :])
    assert np.abs(got - ref).max() < 1e-8
    # and the solve agrees with a dense solve of the oracle's operator and load
    X, Y = np.meshgrid(xq, xq, indexing="ij")
    fq = sp.lambdify((x, y), -Div(Grad(ue)) + k * ue)(X, Y)
    bq = (V * wq[:, None]).T @ fq @ (V * wq[:, None])
    assert np.allclose(_np(b), bq, atol=1e-12 * np.abs(bq).max())
    ud = np.linalg.solve(dense, bq.ravel()).reshape(N - 2, N - 2)
    assert O.rel_l2(_np(uh), ud) < 1e-11
