"""The reference's inner() tests with known functions and padding
(tests/galerkin/test_inner.py:232-416), FP64 tolerances."""

import numpy as np
import pytest
import sympy as sp

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps
SPACES = ["Legendre", "Chebyshev"]


def _np(t):
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


def _cls(name):
    import jaxfun_amd.galerkin as G

    return getattr(G, name)


@pytest.mark.parametrize("space", SPACES)
@pytest.mark.parametrize("dims", [1, 2, 3])
def test_linear_inner(dev, space, dims):
    # test_inner.py:232-290: a JAXFunction projection of f gives the same load as f
    from jaxfun_amd import TensorProduct, TestFunction, inner
    from jaxfun_amd.galerkin.arguments import JAXFunction

    V = _cls(space)(11)
    T = V if dims == 1 else TensorProduct(*([V] * dims))
    xs = T.system.base_scalars()
    f = sp.sin(xs[0])
    if dims > 1:
        f = f * sp.cos(xs[1])
    if dims > 2:
        f = f * sp.sin(xs[2])
    v = TestFunction(T)
    a = JAXFunction(f, T, name="a")
    l0 = [_np(inner(v * fi)) for fi in (f, a)]
    assert np.allclose(l0[0], l0[1], atol=100 * EPS)
    l1 = [_np(inner(v.diff(xs[0], 1) * fi)) for fi in (f, a)]
    assert np.allclose(l1[0], l1[1], atol=(1000 if dims == 1 else 10000) * EPS)


@pytest.mark.parametrize("space", SPACES)
def test_nonlinear_inner_3d(dev, space):
    # test_inner.py:293-311
    from jaxfun_amd import TensorProduct, TestFunction, inner
    from jaxfun_amd.galerkin.arguments import JAXFunction

    V = _cls(space)(11)
    T = TensorProduct(V, V, V)
    x, y, z = T.system.base_scalars()
    f = sp.sin(x) * sp.cos(y) * sp.sin(z)
    a = JAXFunction(f, T, name="a")
    v = TestFunction(T)
    l0 = [_np(inner(v * fi)) for fi in (f**2, a**2)]
    assert np.allclose(l0[0], l0[1], atol=100 * EPS)
    l1 = [_np(inner(v.diff(y, 1) * fi)) for fi in (f**2, a**2)]
    assert np.allclose(l1[0], l1[1], atol=1000 * EPS)


@pytest.mark.parametrize("space", SPACES)
def test_bilinear_inner(dev, space):
    # test_inner.py:314-334
    from jaxfun_amd import TestFunction, TrialFunction, inner
    from jaxfun_amd.galerkin.arguments import JAXFunction

    V = _cls(space)(11)
    x = V.system.x
    v, u = TestFunction(V), TrialFunction(V)
    a = JAXFunction(sp.sin(x), V, name="a")
    m0 = [inner(v * u * fi, kind="bilinear").todense_host() for fi in (sp.sin(x), a)]
    assert np.allclose(m0[0], m0[1], atol=100 * EPS)
    m1 = [inner(v.diff(x, 1) * u * fi, kind="bilinear").todense_host() for fi in (sp.sin(x), a)]
    assert np.allclose(m1[0], m1[1], atol=1000 * EPS)


@pytest.mark.parametrize("space", SPACES)
@pytest.mark.parametrize("dims", [1, 2])
def test_inner_padding(dev, space, dims):
    # test_inner.py:337-416: with 36 modes sin(2 pi x)^2 is resolved (padding
    # changes nothing); with 8 it is not (padding changes the result, and
    # more padding than 3/2 changes nothing further)
    from jaxfun_amd import TensorProduct, TestFunction, inner
    from jaxfun_amd.galerkin.arguments import JAXFunction

    tol = np.sqrt(100 * EPS)
    for N, Np, resolved in ((36, 48, True), (8, 12, False)):
        D = _cls(space)(N)
        T = D if dims == 1 else TensorProduct(D, D)
        xs = T.system.base_scalars()
        f = sp.sin(xs[0] * 2 * sp.pi)
        if dims == 2:
            f = f * sp.sin(xs[1] * 2 * sp.pi)
        uf = JAXFunction(f, T)
        v = TestFunction(T)
        nq = Np if dims == 1 else (Np, Np)
        y0 = _np(inner(v * uf**2, num_quad_points=None, kind="linear"))
        y1 = _np(inner(v * uf**2, num_quad_points=nq, kind="linear"))
        y2 = _np(T.forward(uf.backward() ** 2))
        y3 = _np(T.forward(uf.backward(N=nq) ** 2))
        assert (np.linalg.norm(y0 - y1) < tol) == resolved
        assert (np.linalg.norm(y2 - y3) < tol) == resolved
        if not resolved:
            n2 = 16 if dims == 1 else (16, 16)
            assert np.linalg.norm(y3 - _np(T.forward(uf.backward(N=n2) ** 2))) < tol


@pytest.mark.parametrize("dims", [1, 2])
def test_inner_padding_fourier(dev, dims):
    # test_inner.py:418-453: exp(8ix) aliases to wavenumber 2 without padding
    from jaxfun_amd import Fourier, TensorProduct, TestFunction, inner
    from jaxfun_amd.galerkin.arguments import JAXFunction

    F = Fourier(6)
    T = F if dims == 1 else TensorProduct(F, F)
    xs = T.system.base_scalars()
    f = sp.exp(4j * xs[0]) if dims == 1 else sp.exp(4j * xs[0]) * sp.exp(4j * xs[1])
    uf = JAXFunction(f, T)
    v = TestFunction(T)
    nq = 9 if dims == 1 else (9, 9)
    tol = np.sqrt(100 * EPS)
    y0 = _np(inner(v * uf**2, num_quad_points=None, kind="linear"))
    y1 = _np(inner(v * uf**2, num_quad_points=nq, kind="linear"))
    idx = 2 if dims == 1 else (2, 2)
    assert np.linalg.norm(y1) < tol and abs(y0[idx] - (2 * np.pi) ** dims) < tol
    y2 = _np(T.forward(uf.backward() ** 2))
    y3 = _np(T.forward(uf.backward(N=nq) ** 2))
    assert np.linalg.norm(y3) < tol and abs(y2[idx] - 1) < tol


@pytest.mark.parametrize("space", SPACES)
@pytest.mark.parametrize("kind", ["composite", "directsum", "directsumtps"])
def test_inner_padding_lifted_resolved(dev, space, kind):
    # test_inner.py:456-516
    from jaxfun_amd import FunctionSpace, TensorProduct, TestFunction, inner
    from jaxfun_amd.coordinates import x as X, y as Y
    from jaxfun_amd.galerkin.arguments import JAXFunction

    if kind == "directsumtps":
        f = sp.cos(X * 2 * sp.pi) * sp.cos(Y * 2 * sp.pi)
        Dx = FunctionSpace(36, _cls(space), bcs={"left": {"D": f.subs(X, -1)}, "right": {"D": f.subs(X, 1)}})
        Dy = FunctionSpace(36, _cls(space), bcs={"left": {"D": f.subs(Y, -1)}, "right": {"D": f.subs(Y, 1)}})
        T = TensorProduct(Dx, Dy)
        x, y = T.system.base_scalars()
        f = f.subs({X: x, Y: y})
        nq = (48, 48)
    else:
        bcs = {"left": {"D": 0}, "right": {"D": 0}} if kind == "composite" else {"left": {"D": 1}, "right": {"D": 1}}
        T = FunctionSpace(36, _cls(space), bcs=bcs)
        x = T.system.x
        f = sp.sin(x * 2 * sp.pi) if kind == "composite" else sp.cos(x * 2 * sp.pi)
        nq = 48
    uf = JAXFunction(f, T)
    v = TestFunction(T)
    tol = np.sqrt(100 * EPS)
    y0 = _np(inner(v * uf**2, num_quad_points=None, kind="linear"))
    y1 = _np(inner(v * uf**2, num_quad_points=nq, kind="linear"))
    assert np.linalg.norm(y0 - y1) < tol
    assert np.linalg.norm(_np(T.forward(uf.backward() ** 2)) - _np(T.forward(uf.backward(N=nq) ** 2))) < tol


def test_inner_exact_poly(dev):
    # test_inner.py:519-577: x^2 P_5 P_5 needs 7 Gauss points
    from jaxfun_amd import Legendre, TensorProduct, TestFunction, TrialFunction, inner

    N = 6
    L = Legendre(N, name="L")
    x = L.system.x
    t = float(sp.integrate(x**2 * sp.legendre(N - 1, x) ** 2, (x, -1, 1)).n())
    tol = np.sqrt(100 * EPS)
    u, v = TrialFunction(L), TestFunction(L)
    A = inner(x**2 * v * u, use_precomputed_matrices=False, kind="bilinear")
    assert not abs(A[N - 1, N - 1] - t) < tol
    A = inner(x**2 * v * u, num_quad_points=7, use_precomputed_matrices=False, kind="bilinear")
    assert abs(A[N - 1, N - 1] - t) < tol
    T = TensorProduct(L, L)
    x, y = T.system.base_scalars()
    u, v = TrialFunction(T), TestFunction(T)
    B = inner((x**2 + y**2) * v * u, use_precomputed_matrices=False, kind="bilinear")
    assert not abs(B.tpmats[0].mats[0].data[N - 1, N - 1] - t) < tol
    B = inner((x**2 + y**2) * v * u, num_quad_points=(7, 7), use_precomputed_matrices=False, kind="bilinear")
    assert abs(B.tpmats[0].mats[0].data[N - 1, N - 1] - t) < tol
    assert abs(B.tpmats[1].mats[1].data[N - 1, N - 1] - t) < tol


def test_inner_exact_nonseparable_2d(dev):
    # test_inner.py:580-655: a JAXFunction coefficient and sqrt(x^2 + y^2)
    from scipy.integrate import dblquad

    from jaxfun_amd import Legendre, TensorProduct, TestFunction, TrialFunction, inner
    from jaxfun_amd.galerkin.arguments import JAXFunction

    N = 6
    T = TensorProduct(Legendre(N, name="L"), Legendre(N, name="L"))
    x, y = T.system.base_scalars()
    u, v = TrialFunction(T), TestFunction(T)
    tol = np.sqrt(100 * EPS)
    t = float(sp.integrate((x**2 + y**2) * sp.legendre(N - 1, x) ** 2 * sp.legendre(N - 1, y) ** 2,
                           (x, -1, 1), (y, -1, 1)).n())
    uf = JAXFunction(x**2 + y**2, T)
    assert not abs(inner(uf * v * u, kind="bilinear").data[-1, -1, -1, -1] - t) < tol
    assert abs(inner(uf * v * u, num_quad_points=(7, 7), kind="bilinear").data[-1, -1, -1, -1] - t) < tol
    xs, ys = sp.symbols("x,y", real=True)
    t2, _ = dblquad(sp.lambdify((xs, ys), sp.sqrt(xs**2 + ys**2) * sp.legendre(N - 1, xs) ** 2
                                * sp.legendre(N - 1, ys) ** 2), -1, 1, -1, 1, epsabs=100 * EPS)
    assert not abs(inner(sp.sqrt(x**2 + y**2) * v * u, kind="bilinear").data[-1, -1, -1, -1] - t2) < tol
    B = inner(sp.sqrt(x**2 + y**2) * v * u, num_quad_points=(32, 32), kind="bilinear")
    assert abs(B.data[-1, -1, -1, -1] - t2) < tol
