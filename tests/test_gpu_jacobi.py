"""Jacobi-family spaces (Jacobi, ChebyshevU, Ultraspherical) on the GPU vs the
oracle, following the reference's own tests (tests/galerkin/test_polynomials.py:32-42,
test_legendre_jacobi_extra.py:27-37, test_forward_backward.py:20-110,
test_jaxfunction.py, test_composite.py)."""

import numpy as np
import pytest
import sympy as sp

from oracle import spectral as O

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


# (constructor kwargs, oracle family)
CASES = {
    "jacobi00": ("Jacobi", {}, O.JacobiFamily(0, 0)),
    "jacobi12": ("Jacobi", {"alpha": 1, "beta": 2}, O.JacobiFamily(1, 2)),
    "jacobi_half": ("Jacobi", {"alpha": sp.Rational(-1, 2), "beta": sp.Rational(-1, 2)}, O.JacobiFamily(-0.5, -0.5)),
    "jacobi_skew": ("Jacobi", {"alpha": sp.Rational(3, 10), "beta": sp.Rational(-7, 10)}, O.JacobiFamily(0.3, -0.7)),
    "chebyshevu": ("ChebyshevU", {}, O.jacobi_family("chebyshevu")),
    "ultraspherical": ("Ultraspherical", {}, O.jacobi_family("ultraspherical")),
    "ultraspherical2": ("Ultraspherical", {"lambda_": 2}, O.JacobiFamily(1.5, 1.5, "ultraspherical")),
    "ultraspherical0": ("Ultraspherical", {"lambda_": 0}, O.JacobiFamily(-0.5, -0.5, "ultraspherical")),
}
BASIC = ["jacobi00", "chebyshevu", "ultraspherical"]


def _cls(name):
    import jaxfun_amd.galerkin as G

    return getattr(G, name)


def _space(case, N, **kw):
    cls, fam_kw, fam = CASES[case]
    return _cls(cls)(N, **fam_kw, **kw), fam


@pytest.mark.parametrize("case", list(CASES))
def test_vandermonde_vs_oracle(dev, case):
    import torch

    from jaxfun_amd import kernels as K

    N = 40
    V, fam = _space(case, N)
    xs = np.concatenate([np.linspace(-1, 1, 17), np.random.default_rng(0).uniform(-1, 1, 31)])
    X = torch.tensor(xs, dtype=torch.float64, device=K.device())
    for k in range(4):
        got = _np(V.evaluate_basis_derivative(X, k))
        ref = fam.vandermonde(xs, N, k)
        assert np.abs(got - ref).max() < 1e-13 * max(1.0, np.abs(ref).max()), k


@pytest.mark.parametrize("case", list(CASES))
def test_quadrature_orthogonality(dev, case):
    # backward of the unit coefficient vectors, then the scalar product:
    # (Q_m, Q_n)_w = h_n delta_mn on the device's own nodes and weights
    N = 24
    V, fam = _space(case, N)
    x, w = V.quad_points_and_weights()
    xo, wo = fam.quad(N)
    assert np.allclose(_np(x), xo, atol=1e-14) and np.allclose(_np(w), wo, rtol=1e-12, atol=1e-15)
    G = _np(V.scalar_product(V.backward(np.eye(N))))
    h = fam.norm_squared(N)
    assert np.abs(G - np.diag(h)).max() < 1e-12 * h.max()
    assert np.allclose(V.norm_squared(), h, rtol=1e-13)


def test_jacobi_general_parameters(dev):
    # tests/galerkin/test_polynomials.py:32-42 and test_legendre_jacobi_extra.py:27-37
    from jaxfun_amd.galerkin import Jacobi

    J = Jacobi(5, alpha=1, beta=2)
    x = np.linspace(-1, 1, 13)
    u = _np(J.evaluate(x, np.arange(5.0)))
    assert np.isfinite(u).all()
    ref = O.JacobiFamily(1, 2).vandermonde(x, 5) @ np.arange(5.0)
    assert np.abs(u - ref).max() < 1e-13 * np.abs(ref).max()
    M = J.mass_matrix().todense_host()
    assert np.allclose(np.diag(M), J.norm_squared() / J.domain_factor, atol=1e-15)
    J3 = Jacobi(3, alpha=1, beta=0)
    for c in ([1.0], [1.0, 0.3], [1.0, 0.3, 0.2]):
        v = float(_np(J3.evaluate(np.array([0.4]), np.array(c)))[0])
        cc = np.zeros(3)
        cc[: len(c)] = c
        assert abs(v - (O.JacobiFamily(1, 0).vandermonde([0.4], 3) @ cc)[0]) < 1e-15


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("domain", [None, (-2.0, 3.0)])
def test_forward_backward(dev, case, domain):
    # tests/galerkin/test_forward_backward.py:20-47: forward(backward(project(sin x))) == project(sin x)
    from jaxfun_amd.galerkin import project1D

    D, fam = _space(case, 12, domain=domain)
    x = D.system.x
    ue = _np(project1D(sp.sin(x), D))
    uh = _np(D.forward(D.backward(ue)))
    assert np.abs(uh - ue).max() < 5e-13
    # the projection itself against the oracle (orthogonal projection L / h)
    s1 = O.Space1D(fam, 12)
    xq, _ = fam.quad(12)
    a, b = (-1.0, 1.0) if domain is None else domain
    ref = s1.forward_matrix() @ np.sin(a + (xq + 1) * (b - a) / 2)
    assert np.abs(ue - ref).max() < 1e-13


@pytest.mark.parametrize("case", list(CASES))
def test_forward_backward_composite(dev, case):
    # tests/galerkin/test_forward_backward.py:50-75 on the Dirichlet composite
    from jaxfun_amd import FunctionSpace, project
    from jaxfun_amd.galerkin import Composite

    cls, fam_kw, fam = CASES[case]
    D = FunctionSpace(12, _cls(cls), bcs={"left": {"D": 0}, "right": {"D": 0}}, **fam_kw)
    assert isinstance(D, Composite)
    x = D.system.x
    ue = _np(project(sp.sin(x * 2 * sp.pi), D))
    uh = _np(D.forward(D.backward(ue)))
    assert np.abs(uh - ue).max() < 1e-12
    # stencil and mass solve against the oracle's composite space
    s1 = O.Space1D(fam, 12, dirichlet=True)
    assert np.allclose(D.stencil_matrix(), s1.S, atol=1e-14)
    xq, _ = fam.quad(12)
    assert np.abs(ue - s1.forward_matrix() @ np.sin(2 * np.pi * xq)).max() < 1e-12
    # every basis function vanishes at both ends
    ends = _np(D.evaluate(np.array([-1.0, 1.0]), np.eye(D.dim)))
    assert np.abs(ends).max() < 1e-13


@pytest.mark.parametrize("case", BASIC + ["jacobi12"])
def test_forward_backward_2d(dev, case):
    # tests/galerkin/test_forward_backward.py:78-110
    from jaxfun_amd import TensorProduct, project

    D, fam = _space(case, 10)
    T = TensorProduct(D, D)
    x, y = T.system.base_scalars()
    ue = _np(project(sp.sin(x) * sp.sin(y), T))
    uh = _np(T.forward(T.backward(ue)))
    assert np.abs(uh - ue).max() < 1e-13
    s1 = O.Space1D(fam, 10)
    xq, _ = fam.quad(10)
    F = np.sin(xq)[:, None] * np.sin(xq)[None, :]
    assert np.abs(ue - O.tp_forward([s1, s1], F)).max() < 1e-13


@pytest.mark.parametrize("case", ["jacobi12", "chebyshevu", "ultraspherical2"])
def test_poisson2d_dirichlet(dev, case):
    # -Lap u = f with u = (1-x^2)(1-y^2) exp(x): the Galerkin operator against
    # the oracle's quadrature assembly and the solve against a dense solve
    from jaxfun_amd import Div, FunctionSpace, Grad, TensorProduct, TestFunction, TrialFunction, inner

    cls, fam_kw, fam = CASES[case]
    N = 18
    D = FunctionSpace(N, _cls(cls), {"left": {"D": 0}, "right": {"D": 0}}, **fam_kw)
    T = TensorProduct(D, D)
    v, u = TestFunction(T), TrialFunction(T)
    x, y = T.system.base_scalars()
    ue = (1 - x**2) * (1 - y**2) * sp.exp(x)
    A, b = inner(-v * Div(Grad(u)) + v * Div(Grad(ue)), kind="system")
    S = fam.dirichlet_stencil(N)
    M = S @ O.quad_matrix(fam, N, 0, 0) @ S.T
    Kx = S @ O.quad_matrix(fam, N, 0, 2) @ S.T
    dense = -(np.kron(Kx, M) + np.kron(M, Kx))
    assert np.allclose(A.todense_host(), dense, atol=1e-12 * np.abs(dense).max())
    uh = _np(A.solve(b))
    xq, wq = fam.quad(N)
    X, Y = np.meshgrid(xq, xq, indexing="ij")
    f = sp.lambdify((x, y), -Div(Grad(ue)))(X, Y)
    Vq = fam.vandermonde(xq, N) @ S.T
    bq = (Vq * wq[:, None]).T @ f @ (Vq * wq[:, None])
    assert np.allclose(_np(b), bq, atol=1e-12 * np.abs(bq).max())
    ud = np.linalg.solve(dense, bq.ravel()).reshape(N - 2, N - 2)
    assert O.rel_l2(uh, ud) < 1e-11
    xs = np.linspace(-1, 1, 7)
    Vs = fam.vandermonde(xs, N) @ S.T
    ref = sp.lambdify((x, y), ue)(xs[:, None], xs[None, :])
    assert np.abs(Vs @ uh @ Vs.T - ref).max() < 1e-7


@pytest.mark.parametrize("case", ["jacobi12", "chebyshevu", "ultraspherical"])
def test_lifted_dirichlet_1d(dev, case):
    # inhomogeneous values on a Jacobi-family space: u'' = f, u(-1) = 1, u(1) = -2
    from jaxfun_amd import FunctionSpace, TestFunction, TrialFunction, inner

    cls, fam_kw, fam = CASES[case]
    N = 16
    D = FunctionSpace(N, _cls(cls), {"left": {"D": 1}, "right": {"D": -2}}, **fam_kw)
    x = D.system.x
    ue = sp.exp(x) * (1 - x**2) - sp.Rational(1, 2) - sp.Rational(3, 2) * x
    v, u = TestFunction(D), TrialFunction(D)
    A, b = inner(v * sp.diff(u, x, 2) - v * sp.diff(ue, x, 2), kind="system")
    uh = A.solve(b)
    xs = np.linspace(-1, 1, 9)
    got = _np(D.evaluate(xs, uh))
    ref = sp.lambdify(x, ue)(xs)
    assert abs(got[0] - 1.0) < 1e-13 and abs(got[-1] + 2.0) < 1e-13
    assert np.abs(got - ref).max() < 1e-9


@pytest.mark.parametrize("case", BASIC)
def test_jaxfunction_forms(dev, case):
    # tests/galerkin/test_jaxfunction.py:66-141 with the Jacobi-family spaces
    from jaxfun_amd import TestFunction, TrialFunction, inner
    from jaxfun_amd.galerkin.arguments import JAXFunction

    N = 8
    D, fam = _space(case, N)
    x = D.system.x
    u, v = TrialFunction(D), TestFunction(D)
    uf = JAXFunction(np.linspace(0.5, 1.5, N), D)
    A = inner(u * v, kind="bilinear")
    lhs, rhs = _np(A @ uf), _np(inner(v * uf, kind="linear"))
    assert np.linalg.norm(lhs - rhs) < 1e-13 * np.linalg.norm(rhs)
    A = inner(sp.diff(u, x) * v, kind="bilinear")
    lhs, rhs = _np(A @ uf), _np(inner(v * sp.diff(uf, x), kind="linear"))
    assert np.linalg.norm(lhs - rhs) < 1e-12 * max(1.0, np.linalg.norm(rhs))
    A = inner(u * v * uf, kind="bilinear")
    lhs, rhs = _np(A @ uf), _np(inner(v * uf**2, kind="linear"))
    assert np.linalg.norm(lhs - rhs) < 1e-12 * np.linalg.norm(rhs)
    # mass matrix: closed form diag(h) and quadrature agree
    M0 = inner(u * v, kind="bilinear").todense_host()
    M1 = inner(u * v, kind="bilinear", use_precomputed_matrices=False).todense_host()
    assert np.allclose(M0, np.diag(fam.norm_squared(N)), atol=1e-14)
    assert np.allclose(M0, M1, atol=1e-13)


def test_ultraspherical_lambda0_is_chebyshev(dev):
    # galerkin/Ultraspherical.py:65-101: lambda = 0 is the Chebyshev basis
    import torch

    from jaxfun_amd import kernels as K
    from jaxfun_amd.galerkin import Ultraspherical

    U = Ultraspherical(24, lambda_=0)
    xs = np.linspace(-1, 1, 31)
    X = torch.tensor(xs, dtype=torch.float64, device=K.device())
    for k in range(3):
        assert np.abs(_np(U.evaluate_basis_derivative(X, k)) - O.vandermonde("chebyshev", xs, 24, k)).max() < 1e-11
    x, w = U.quad_points_and_weights()
    xc, wc = O.quad("chebyshev", 24)
    assert np.allclose(_np(x), xc, atol=1e-14) and np.allclose(_np(w), wc, atol=1e-14)
    assert np.allclose(U.norm_squared(), O.norm_squared("chebyshev", 24))

