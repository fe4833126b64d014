"""General boundary conditions (D, N, N2..N4) on the GPU vs the oracle:
composite stencils, lifting, biharmonic and mixed Neumann problems, and the
reference's examples/biharmonic2D.py (tests/galerkin/test_composite.py,
galerkin/composite.py:765-895, tensorproductspace.py:560-851)."""

import numpy as np
import pytest
import sympy as sp

from oracle import spectral as O

pytestmark = pytest.mark.gpu

BCSETS = {
    "LDLNRDRN": {"left": {"D": 0, "N": 0}, "right": {"D": 0, "N": 0}},
    "LNRN": {"left": {"N": 0}, "right": {"N": 0}},
    "LDRN": {"left": {"D": 0}, "right": {"N": 0}},
    "LDLN2RD": {"left": {"D": 0, "N2": 0}, "right": {"D": 0}},
}
FAMS = {"legendre": ("Legendre", {}, "legendre"), "chebyshev": ("Chebyshev", {}, "chebyshev"),
        "jacobi12": ("Jacobi", {"alpha": 1, "beta": 2}, O.JacobiFamily(1, 2))}


def _np(t):
    return t.detach().cpu().numpy()


def _cls(name):
    import jaxfun_amd.galerkin as G

    return getattr(G, name)


@pytest.mark.parametrize("fam", list(FAMS))
@pytest.mark.parametrize("bcs", list(BCSETS))
def test_composite_stencil_and_conditions(dev, fam, bcs):
    import torch

    from jaxfun_amd import FunctionSpace
    from jaxfun_amd import kernels as K

    cls, kw, ofam = FAMS[fam]
    N = 16
    D = FunctionSpace(N, _cls(cls), BCSETS[bcs], **kw)
    conds = O.bc_conditions(BCSETS[bcs])
    S = O.general_stencil(ofam, N, conds)
    assert D.dim == N - len(conds)
    assert np.allclose(D.stencil_matrix(), S, atol=1e-12 * max(1.0, np.abs(S).max()))
    # the device basis meets every condition
    X = torch.tensor([-1.0, 1.0], dtype=torch.float64, device=K.device())
    for side, k in conds:
        V = _np(D.evaluate_basis_derivative(X, k))[0 if side == "L" else 1]
        assert np.abs(V).max() < 1e-11 * max(1.0, np.abs(_np(D.orthogonal.vandermonde(X, k))).max())
    # forward(backward(c)) = c through the banded mass solve of width nb
    c = np.random.default_rng(0).standard_normal(D.dim)
    assert O.rel_l2(_np(D.forward(D.backward(c))), c) < 1e-11


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
def test_biharmonic_1d(dev, family):
    # u'''' = f with u = u' = 0 at both ends: operator vs the oracle's
    # quadrature assembly, solution vs a dense oracle solve and vs the exact u
    from jaxfun_amd import Chebyshev, FunctionSpace, Legendre, TestFunction, TrialFunction, inner

    N = 24
    space = Legendre if family == "legendre" else Chebyshev
    D = FunctionSpace(N, space, BCSETS["LDLNRDRN"])
    x = D.system.x
    ue = (1 - x**2) ** 2 * sp.exp(x)
    v, u = TestFunction(D), TrialFunction(D)
    A, b = inner(v * sp.diff(u, x, 4) - v * sp.diff(ue, x, 4), kind="system")
    S = O.general_stencil(family, N, O.bc_conditions(BCSETS["LDLNRDRN"]))
    Ad = S @ O.quad_matrix(family, N, 0, 4) @ S.T
    assert np.allclose(A.todense_host(), Ad, atol=1e-11 * np.abs(Ad).max())
    xq, wq = O.quad(family, N)
    bq = S @ (O.vandermonde(family, xq, N).T * wq) @ sp.lambdify(x, sp.diff(ue, x, 4))(xq)
    assert np.allclose(_np(b), bq, atol=1e-12 * np.abs(bq).max())
    uh = _np(A.solve(b))
    assert O.rel_l2(uh, np.linalg.solve(Ad, bq)) < 1e-10
    xs = np.linspace(-1, 1, 11)
    assert np.abs(_np(D.evaluate(xs, uh)) - sp.lambdify(x, ue)(xs)).max() < 1e-9


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
def test_mixed_neumann_lifted_1d(dev, family):
    # u'' = f on (0, 2) with u(0) = 1 and u'(2) = -3 given in physical units
    from jaxfun_amd import Chebyshev, FunctionSpace, Legendre, TestFunction, TrialFunction, inner

    N = 20
    space = Legendre if family == "legendre" else Chebyshev
    D = FunctionSpace(N, space, {"left": {"D": 1}, "right": {"N": -3}}, domain=(0, 2))
    x = D.system.x
    ue = sp.sin(x) + 1 + (-3 - sp.cos(2)) * x
    assert abs(float(ue.subs(x, 0)) - 1) < 1e-15 and abs(float(sp.diff(ue, x).subs(x, 2)) + 3) < 1e-14
    v, u = TestFunction(D), TrialFunction(D)
    A, b = inner(v * sp.diff(u, x, 2) - v * sp.diff(ue, x, 2), kind="system")
    uh = A.solve(b)
    xs = np.linspace(0, 2, 9)
    got = _np(D.evaluate(xs, uh))
    assert np.abs(got - sp.lambdify(x, ue)(xs)).max() < 1e-12
    # the lift alone carries the boundary data
    lift = _np(D.evaluate(np.array([0.0, 2.0 - 1e-6, 2.0]), np.zeros(D.dim)))
    assert abs(lift[0] - 1.0) < 1e-13 and abs((lift[2] - lift[1]) / 1e-6 + 3.0) < 1e-5


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
def test_neumann_poisson_pinned(dev, family):
    # -u'' = f with u' = 0 at both ends: singular (constants); the reference
    # fixes the constant mode with DiaMatrix.pin (la/diamatrix.py)
    from jaxfun_amd import Chebyshev, FunctionSpace, Legendre, TestFunction, TrialFunction, inner

    N = 20
    space = Legendre if family == "legendre" else Chebyshev
    D = FunctionSpace(N, space, BCSETS["LNRN"])
    x = D.system.x
    ue = sp.cos(sp.pi * x)
    v, u = TestFunction(D), TrialFunction(D)
    A, b = inner(-v * sp.diff(u, x, 2) + v * sp.diff(ue, x, 2), kind="system")
    Ad = A.todense_host()
    assert np.linalg.matrix_rank(Ad, tol=1e-10 * np.abs(Ad).max()) == D.dim - 1
    uh = A.pin({0: 0.0}).solve(b)
    xs = np.linspace(-1, 1, 9)
    got = _np(D.evaluate(xs, uh))
    ref = sp.lambdify(x, ue)(xs)
    assert np.abs((got - got.mean()) - (ref - ref.mean())).max() < 1e-9


def test_biharmonic2d_example(dev):
    # examples/biharmonic2D.py: Legendre, scaling n + 1, D and N lifted on both
    # axes, ue = (x - x^2)^2 (x - y^2)^2; error norm / 100 on a uniform mesh
    from jaxfun_amd import Div, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, TrialFunction, inner, n
    from jaxfun_amd.coordinates import x, y

    ue = (x - x**2) ** 2 * (x - y**2) ** 2
    M = 20
    bcsx = {"left": {"D": ue.subs(x, -1), "N": ue.diff(x, 1).subs(x, -1)},
            "right": {"D": ue.subs(x, 1), "N": ue.diff(x, 1).subs(x, 1)}}
    bcsy = {"left": {"D": ue.subs(y, -1), "N": ue.diff(y, 1).subs(y, -1)},
            "right": {"D": ue.subs(y, 1), "N": ue.diff(y, 1).subs(y, 1)}}
    Dx = FunctionSpace(M, Legendre, scaling=n + 1, bcs=bcsx, name="Dx", fun_str="psi")
    Dy = FunctionSpace(M, Legendre, scaling=n + 1, bcs=bcsy, name="Dy", fun_str="phi")
    T = TensorProduct(Dx, Dy, name="T")
    v, u = TestFunction(T, name="v"), TrialFunction(T, name="u")
    A, b = inner(Div(Grad(Div(Grad(u)))) * v - Div(Grad(Div(Grad(ue)))) * v, sparse=True, kind="system")
    uh = A.solve(b)
    Nu = 100
    X, Y = [m.cpu().numpy() for m in T.mesh(kind="uniform", N=(Nu, Nu))]
    uj = _np(T.evaluate_mesh(uh, kind="uniform", N=(Nu, Nu)))
    uej = sp.lambdify((x, y), ue)(X, Y)
    assert np.linalg.norm(uj - uej) / Nu < 1e-11


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
def test_robin_lifted_1d(dev, family):
    # u'' = f on (0, 1) with u(0) + 2 u'(0) = g (Robin, physical units) and
    # u(1) = ue(1); R: (alpha, value) as in galerkin/composite.py:806-822
    import torch

    from jaxfun_amd import Chebyshev, FunctionSpace, Legendre, TestFunction, TrialFunction, inner

    N = 20
    space = Legendre if family == "legendre" else Chebyshev
    xs_ = sp.Symbol("x")
    ue0 = sp.exp(xs_) * sp.sin(2 * xs_) + 1
    g = float((ue0 + 2 * sp.diff(ue0, xs_)).subs(xs_, 0))
    D = FunctionSpace(N, space, {"left": {"R": (2.0, g)}, "right": {"D": float(ue0.subs(xs_, 1))}}, domain=(0, 1))
    x = D.system.x
    ue = ue0.subs(xs_, x)
    conds = O.bc_conditions({"left": {"R": (2.0, 0)}, "right": {"D": 0}}, df=2.0)
    assert np.allclose(D[0].stencil_matrix(), O.general_stencil(family, N, conds), atol=1e-13)
    v, u = TestFunction(D), TrialFunction(D)
    A, b = inner(v * sp.diff(u, x, 2) - v * sp.diff(ue, x, 2), kind="system")
    uh = A.solve(b)
    xs = np.linspace(0, 1, 9)
    assert np.abs(_np(D.evaluate(xs, uh)) - sp.lambdify(x, ue)(xs)).max() < 1e-11
    # the homogeneous basis meets u + 2 u' = 0 at x = 0 (alpha scaled by df = 2 on the reference side)
    from jaxfun_amd import kernels as K

    X = torch.tensor([-1.0], dtype=torch.float64, device=K.device())
    v0 = _np(D[0].evaluate_basis_derivative(X, 0))[0] + 2.0 * 2.0 * _np(D[0].evaluate_basis_derivative(X, 1))[0]
    assert np.abs(v0).max() < 1e-11


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
def test_weighted_neumann_lifted_1d(dev, family):
    # u'' = f on (0, 2) with u'(0) + 0.3 u''(0) = g (W, physical units) and u(2) = ue(2)
    # (alpha = 0.5 makes the Legendre functionals of P_2 and P_3 coincide, so
    # that stencil row is singular in the reference's construction as well)
    from jaxfun_amd import Chebyshev, FunctionSpace, Legendre, TestFunction, TrialFunction, inner

    N = 22
    space = Legendre if family == "legendre" else Chebyshev
    s_ = sp.Symbol("x")
    ue0 = sp.cos(s_) * sp.exp(s_ / 2) + s_**2
    g = float((sp.diff(ue0, s_) + 0.3 * sp.diff(ue0, s_, 2)).subs(s_, 0))
    D = FunctionSpace(N, space, {"left": {"W": (0.3, g)}, "right": {"D": float(ue0.subs(s_, 2))}}, domain=(0, 2))
    x = D.system.x
    ue = ue0.subs(s_, x)
    conds = O.bc_conditions({"left": {"W": (0.3, 0)}, "right": {"D": 0}}, df=1.0)
    assert np.allclose(D[0].stencil_matrix(), O.general_stencil(family, N, conds), atol=1e-12)
    v, u = TestFunction(D), TrialFunction(D)
    A, b = inner(v * sp.diff(u, x, 2) - v * sp.diff(ue, x, 2), kind="system")
    uh = A.solve(b)
    xs = np.linspace(0, 2, 9)
    assert np.abs(_np(D.evaluate(xs, uh)) - sp.lambdify(x, ue)(xs)).max() < 1e-10
