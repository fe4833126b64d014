"""Inhomogeneous Dirichlet lifting on the GPU (BCGeneric / DirectSum /
DirectSumTPS) against the FP64 oracle's restatement.

These follow the reference's own tests (tests/galerkin/test_composite.py:66-80,
test_additional_coverage.py:87-114, test_tensorproductspace_extra.py:215-260,
test_tensorproductspace_more.py:38-50) and examples/helmholtz2D.py.
"""

import numpy as np
import pytest
import sympy as sp

from oracle import spectral as O

pytestmark = pytest.mark.gpu

FAMILIES = ["legendre", "chebyshev"]


def _np(t):
    return t.detach().cpu().numpy()


def _space(family):
    from jaxfun_amd import Chebyshev, Legendre

    return Legendre if family == "legendre" else Chebyshev


def test_add_bcast_kernel(dev):
    import torch

    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(0)
    X = rng.standard_normal((3, 5, 7))
    D = rng.standard_normal((5, 7))
    Y = K.add_bcast(torch.tensor(X, device=dev), torch.tensor(D, device=dev), -0.75)
    # one fused multiply-add per element: within an ulp of numpy's two roundings
    assert np.allclose(_np(Y), X + (-0.75) * D[None], rtol=0, atol=1e-15)
    with pytest.raises(ValueError):
        K.add_bcast(torch.tensor(X, device=dev), torch.tensor(D[:4], device=dev).contiguous())


@pytest.mark.parametrize("M,K_", [(2, 2), (30, 2), (2, 30), (1, 2)])
def test_contraction_two_columns(dev, M, K_):
    # the lifting spaces contract with two-column operators
    import torch

    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(1)
    A = rng.standard_normal((M, K_))
    for shape, axis in (((K_,), 0), ((4, K_), 1), ((K_, 9), 0), ((3, K_, 5), 1), ((K_, 64), 0),
                        ((3, K_, 128), 1)):
        X = rng.standard_normal(shape)
        Y = _np(K.apply_along(K.AxisOperator(torch.tensor(A, device=dev)), torch.tensor(X, device=dev), axis))
        assert np.allclose(Y, O.apply_along(A, X, axis), rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("family", FAMILIES)
def test_directsum_boundary_values(dev, family):
    # tests/galerkin/test_composite.py:66-80, test_additional_coverage.py:101-114
    from jaxfun_amd.galerkin import BCGeneric, DirectSum, FunctionSpace

    FS = FunctionSpace(6, _space(family), bcs={"left": {"D": 1}, "right": {"D": 2}})
    assert isinstance(FS, DirectSum) and isinstance(FS[1], BCGeneric)
    C = FS[0]
    assert FS.dim == C.dim == 4 and FS[1].dim == 2 and FS[1].num_dofs == 0
    ends = _np(FS.evaluate(np.array([-1.0, 1.0]), np.zeros(4)))
    assert abs(ends[0] - 1.0) < 1e-15 and abs(ends[1] - 2.0) < 1e-15
    assert np.isfinite(float(_np(FS.evaluate(0.1, np.ones(4)))))
    assert _np(FS.backward(np.ones(C.dim))).shape == (C.num_quad_points,)
    H = FunctionSpace(6, _space(family), bcs={"left": {"D": 0}, "right": {"D": 0}})
    assert not isinstance(H, DirectSum) and H.bcs.is_homogeneous()


@pytest.mark.parametrize("family", FAMILIES)
def test_directsum_transforms_vs_oracle(dev, family):
    from jaxfun_amd import FunctionSpace, n

    N, dom, bv = 18, (0.0, 2.0), (0.7, -1.3)
    V = FunctionSpace(N, _space(family), bcs={"left": {"D": bv[0]}, "right": {"D": bv[1]}}, domain=dom,
                      scaling=n + 1)
    ax = O.Axis1D(family, N, "lift", dom, scaling=np.arange(N - 2) + 1.0, bvals=bv)
    c = np.random.default_rng(3).standard_normal((3, N - 2))
    assert O.rel_l2(_np(V.to_orthogonal(c)), ax.to_orthogonal(c)) < 1e-14
    assert O.rel_l2(_np(V.lift_coefficients()), ax.lift()) < 1e-15
    a = ax.to_orthogonal(c) + 1e-3 * np.random.default_rng(4).standard_normal((3, N))
    assert O.rel_l2(_np(V.from_orthogonal(a)), ax.from_orthogonal(a)) < 1e-13
    for Nq in (None, 27):
        xq = ax.mesh(Nq)
        for k in (0, 1, 2):
            got = _np(V.backward_primitive(c, k, Nq))
            assert O.rel_l2(got, ax.values(c, xq, k)) < 1e-12, (Nq, k)
    xu = np.linspace(*dom, 41)
    assert O.rel_l2(_np(V.evaluate_mesh(c, "uniform", 41)), ax.values(c, xu)) < 1e-13
    assert O.rel_l2(_np(V.evaluate(xu, c[1])), ax.values(c[1], xu)) < 1e-13
    u = ax.values(c, ax.mesh())
    assert O.rel_l2(_np(V.forward(u)), c) < 1e-12
    assert np.abs(_np(V.forward(u)) - ax.forward(u)).max() < 1e-12


@pytest.mark.parametrize("family", FAMILIES)
def test_lifted_helmholtz_1d_vs_oracle(dev, family):
    from jaxfun_amd import Div, FunctionSpace, Grad, TestFunction, TrialFunction, inner, n
    from jaxfun_amd.coordinates import x

    N, dom, alpha = 40, (0.0, 1.0), 3.0
    ue = sp.exp(sp.cos(2 * sp.pi * x))
    bcs = {"left": {"D": float(ue.subs(x, dom[0]))}, "right": {"D": float(ue.subs(x, dom[1]))}}
    D = FunctionSpace(N, _space(family), bcs=bcs, domain=dom, scaling=n + 1)
    v, u = TestFunction(D), TrialFunction(D)
    A, b = inner(v * (Div(Grad(u)) + alpha * u) - v * (Div(Grad(ue)) + alpha * ue), sparse=True, kind="system")
    uh = _np(A.solve(b))
    ax = O.Axis1D(family, N, "lift", dom, scaling=np.arange(N - 2) + 1.0,
                  bvals=[bcs["left"]["D"], bcs["right"]["D"]])
    f = sp.lambdify(x, sp.diff(ue, x, 2) + alpha * ue)
    ur = O.lifted_helmholtz_1d(ax, alpha, f)
    assert O.rel_l2(uh, ur) < 1e-11
    xj = np.linspace(*dom, 100)
    uj = _np(D.evaluate(xj, uh))
    assert np.linalg.norm(uj - sp.lambdify(x, ue)(xj)) < 1e-9


def _bc_exprs(ue, dom0, dom1):
    from jaxfun_amd.coordinates import x, y

    bx = {"left": {"D": ue.subs(x, dom0[0])}, "right": {"D": ue.subs(x, dom0[1])}}
    by = {"left": {"D": ue.subs(y, dom1[0])}, "right": {"D": ue.subs(y, dom1[1])}}
    return bx, by


def _oracle_tps(family, kinds, N, dom, ue, scaling=None):
    from jaxfun_amd.coordinates import x, y

    f = sp.lambdify((x, y), ue)
    axes = [O.Axis1D(family, N[i], kinds[i], dom[i], scaling=scaling[i] if scaling else None) for i in (0, 1)]
    bcs0 = (lambda s: f(dom[0][0], s) + 0 * s, lambda s: f(dom[0][1], s) + 0 * s) if kinds[0] == "lift" else None
    bcs1 = (lambda s: f(s, dom[1][0]) + 0 * s, lambda s: f(s, dom[1][1]) + 0 * s) if kinds[1] == "lift" else None
    return O.Lifted2D(axes[0], axes[1], bcs0, bcs1)


def _product_tps(family, kinds, N, dom, ue, scaling=None):
    from jaxfun_amd import FunctionSpace, TensorProduct

    bx, by = _bc_exprs(ue, dom[0], dom[1])
    spaces = []
    for i, (kind, bc) in enumerate(zip(kinds, (bx, by))):
        kw = {} if scaling is None else {"scaling": scaling[i]}
        if kind == "orth":
            spaces.append(FunctionSpace(N[i], _space(family), domain=dom[i]))
        elif kind == "dir":
            spaces.append(FunctionSpace(N[i], _space(family), bcs={"left": {"D": 0}, "right": {"D": 0}},
                                        domain=dom[i], **kw))
        else:
            spaces.append(FunctionSpace(N[i], _space(family), bcs=bc, domain=dom[i], **kw))
    return TensorProduct(*spaces)


KINDS_2D = [("lift", "lift"), ("lift", "dir"), ("orth", "lift"), ("lift", "orth")]


@pytest.mark.parametrize("family", FAMILIES)
@pytest.mark.parametrize("kinds", KINDS_2D)
def test_directsumtps_transforms_vs_oracle(dev, family, kinds):
    from jaxfun_amd.coordinates import x, y
    from jaxfun_amd.galerkin import DirectSumTPS

    N, dom = (14, 12), ((-1.0, 2.0), (0.0, 1.0))
    ue = sp.exp(x / 2) * sp.cos(y) + x * y**2
    T = _product_tps(family, kinds, N, dom, ue)
    L = _oracle_tps(family, kinds, N, dom, ue)
    assert isinstance(T, DirectSumTPS)
    key = {"h": 0, "b": 1}
    for (p0, p1), arr in L.bndvals.items():
        combo = tuple(T.basespaces[i][key[p]] if kinds[i] == "lift" else T.basespaces[i]
                      for i, p in enumerate((p0, p1)))
        assert O.rel_l2(_np(T.bndvals[combo]), arr) < 1e-12, (p0, p1)
    assert O.rel_l2(_np(T.lift_coefficients()), L.lift()) < 1e-12
    dims = tuple(a.dim for a in L.axes)
    c = np.random.default_rng(5).standard_normal((2,) + dims)
    X0, X1 = L.axes[0].mesh(), L.axes[1].mesh()
    ref = np.stack([L.values(ci, X0, X1) for ci in c])
    assert O.rel_l2(_np(T.backward(c)), ref) < 1e-12
    assert O.rel_l2(_np(T.to_orthogonal(c[0])), L.to_orthogonal(c[0])) < 1e-13
    # padded backward, mixed derivative and a uniform mesh
    P0, P1 = L.axes[0].mesh(21), L.axes[1].mesh(18)
    assert O.rel_l2(_np(T.backward(c[1], N=(21, 18))), L.values(c[1], P0, P1)) < 1e-12
    assert O.rel_l2(_np(T.backward_primitive(c[1], (1, 1))), L.values(c[1], X0, X1, (1, 1))) < 1e-11
    U0, U1 = np.linspace(*dom[0], 30), np.linspace(*dom[1], 25)
    assert O.rel_l2(_np(T.evaluate_mesh(c[0], "uniform", (30, 25))), L.values(c[0], U0, U1)) < 1e-12
    # forward(backward(c)) == c
    assert O.rel_l2(_np(T.forward(T.backward(c))), c) < 1e-12
    with pytest.raises(RuntimeError):
        T.scalar_product(ref[0])


def test_directsumtps_sum_of_tpspaces(dev):
    # tests/galerkin/test_tensorproductspace_extra.py:228-260: backward and
    # backward_primitive equal the sum over the split tensor products
    from jaxfun_amd import FunctionSpace, Legendre, TensorProduct

    N = (5, 6)
    F1 = FunctionSpace(N[0], Legendre, bcs={"left": {"D": 1}, "right": {"D": 2}}, domain=(-2, 2))
    F2 = FunctionSpace(N[1], Legendre, bcs={"left": {"D": 3}, "right": {"D": 4}}, domain=(-2, 2))
    T = TensorProduct(F1, F2)
    c = np.random.default_rng(0).standard_normal(T.num_dofs)
    z0 = sum(_np(v.backward_primitive(T.bndvals.get(f, c), k=(1, 1), N=N)) for f, v in T.tpspaces.items())
    assert np.linalg.norm(z0 - _np(T.backward_primitive(c, k=(1, 1), N=N))) < 1e-12
    z0 = sum(_np(v.backward(T.bndvals.get(f, c), N=N)) for f, v in T.tpspaces.items())
    assert np.linalg.norm(z0 - _np(T.backward(c, N=N))) < 1e-12
    # the homogeneous space and the orthogonal space see the same function
    O_ = T.get_orthogonal()
    assert np.linalg.norm(_np(O_.backward(T.to_orthogonal(c))) - _np(T.backward(c))) < 1e-12
    # corner values: the constant boundary functions of the first axis at the
    # second axis's ends
    corners = _np(T.bndvals[(T.basespaces[0][1], T.basespaces[1][1])])
    assert np.array_equal(corners, [[1.0, 1.0], [2.0, 2.0]])


def test_tensorproduct_forward_directsum(dev):
    # tests/galerkin/test_tensorproductspace_more.py:38-50
    from jaxfun_amd import FunctionSpace, Legendre, TensorProduct

    DS = FunctionSpace(5, Legendre, bcs={"left": {"D": 1}, "right": {"D": 2}})
    T = TensorProduct(DS, Legendre(5))
    U = np.random.default_rng(1).standard_normal((DS[0].dim, 5))
    c = _np(T.forward(T.backward(U)))
    assert c.shape == U.shape and np.allclose(c, U, atol=1e-12)


@pytest.mark.parametrize("family", FAMILIES)
@pytest.mark.parametrize("kinds", [("lift", "lift"), ("lift", "dir"), ("dir", "lift")])
@pytest.mark.parametrize("method", ["auto", "kron"])
def test_lifted_helmholtz_2d_vs_oracle(dev, family, kinds, method):
    from jaxfun_amd import Div, Grad, TestFunction, TrialFunction, inner
    from jaxfun_amd.coordinates import x, y

    N, dom, alpha = (16, 14), ((-1.0, 2.0), (0.0, 1.0)), 2.0
    # x^2 y^2 and x^3 y are not reproduced by the lift (a Boolean sum of the
    # two edge interpolations), so the homogeneous part is not zero
    ue = 1 + x * y + x**2 * y**2 - 2 * y**3 + x**3 * y
    if kinds[1] == "dir":
        ue = ue * y * (1 - y)
    if kinds[0] == "dir":
        ue = ue * (x + 1) * (2 - x)
    T = _product_tps(family, kinds, N, dom, ue)
    v, u = TestFunction(T), TrialFunction(T)
    A, b = inner(v * (Div(Grad(u)) + alpha * u) - v * (Div(Grad(ue)) + alpha * ue), sparse=True, kind="system")
    L = _oracle_tps(family, kinds, N, dom, ue)
    X0, X1 = L.axes[0].mesh(), L.axes[1].mesh()
    f = sp.lambdify((x, y), sp.diff(ue, x, 2) + sp.diff(ue, y, 2) + alpha * ue)
    ur, _, br = O.lifted_helmholtz_2d(L, alpha, f(X0[:, None], X1[None, :]) + 0 * X0[:, None])
    assert np.linalg.norm(br) > 1e-3 * np.linalg.norm(L.lift())
    # b is the difference of the load and the lifting terms: rounding relative to those
    assert O.rel_l2(_np(b), br) < 1e-11
    uh = A.solve(b, method=method)
    assert O.rel_l2(_np(uh), ur) < 1e-10
    # the polynomial manufactured solution is reproduced
    U0, U1 = np.linspace(*dom[0], 9), np.linspace(*dom[1], 8)
    uj = _np(T.evaluate_mesh(uh, "uniform", (9, 8)))
    assert np.abs(uj - sp.lambdify((x, y), ue)(U0[:, None], U1[None, :])).max() < 1e-10


@pytest.mark.parametrize("method", ["auto", "kron"])
def test_helmholtz2d_example(dev, method):
    # examples/helmholtz2D.py (M = 50, Legendre, scaling n + 1, domain (0, 1))
    from jaxfun_amd import Div, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, TrialFunction, inner, \
        n, ulp
    from jaxfun_amd.coordinates import x, y

    M = 50
    ue = sp.exp(-(x**2 + y**2))
    bcsx = {"left": {"D": ue.subs(x, 0)}, "right": {"D": ue.subs(x, 1)}}
    bcsy = {"left": {"D": ue.subs(y, 0)}, "right": {"D": ue.subs(y, 1)}}
    Dx = FunctionSpace(M, Legendre, bcs=bcsx, name="Dx", fun_str="psi", scaling=n + 1, domain=(0, 1))
    Dy = FunctionSpace(M, Legendre, bcs=bcsy, name="Dy", fun_str="phi", scaling=n + 1, domain=(0, 1))
    T = TensorProduct(Dx, Dy, name="T")
    v, u = TestFunction(T, name="v"), TrialFunction(T, name="u")
    ue = T.system.expr_psi_to_base_scalar(ue)
    A, L = inner(v * (Div(Grad(u)) + u) - v * (Div(Grad(ue)) + ue), sparse=True, kind="system")
    un = A.solve(L, method=method, kron_method="banded", auto_threshold=10000)
    N = 100
    uj = _np(T.evaluate_mesh(un, kind="uniform", N=(N, N)))
    xj = [m.cpu().numpy() for m in T.mesh(kind="uniform", N=(N, N))]
    uej = sp.lambdify((x, y), ue)(*xj)
    error = np.linalg.norm(uj - uej) / N
    assert error < ulp(1000), error
    # the right-hand side and the solution agree with the oracle's restatement;
    # the coefficients only to the system's conditioning (kappa ~ 1e7 at M = 50)
    sc = np.arange(M - 2) + 1.0
    Lo = _oracle_tps("legendre", ("lift", "lift"), (M, M), ((0.0, 1.0), (0.0, 1.0)), ue, scaling=(sc, sc))
    X0, X1 = Lo.axes[0].mesh(), Lo.axes[1].mesh()
    f = sp.lambdify((x, y), sp.diff(ue, x, 2) + sp.diff(ue, y, 2) + ue)
    ur, _, br = O.lifted_helmholtz_2d(Lo, 1.0, f(X0[:, None], X1[None, :]))
    assert O.rel_l2(_np(L), br) < 1e-11
    xu = np.linspace(0.0, 1.0, N)
    assert np.abs(uj - Lo.values(ur, xu, xu)).max() < 1e-12
    assert O.rel_l2(_np(un), ur) < 1e-7


def test_multivar_coefficient_with_lift(dev):
    # a non-separable coefficient on a lifted trial function: its lifting
    # terms are applied matrix-free to the boundary coefficients
    from jaxfun_amd import FunctionSpace, Legendre, TensorProduct, TestFunction, TrialFunction, inner
    from jaxfun_amd.coordinates import x, y

    N, dom = (12, 10), ((-1.0, 1.0), (-1.0, 1.0))
    ue = 1 + x + 2 * y + x**2 * y**2
    T = _product_tps("legendre", ("lift", "lift"), N, dom, ue)
    v, u = TestFunction(T), TrialFunction(T)
    k = 1 + x**2 * y**2 + x * y
    A, b = inner(v * k * u - v * k * ue, kind="system")
    Lo = _oracle_tps("legendre", ("lift", "lift"), N, dom, ue)
    # b = (v, k ue) - (v, k u_lift): quadrature at the Gauss mesh of both axes
    X0, X1 = Lo.axes[0].mesh(), Lo.axes[1].mesh()
    x0, w0 = O.leggauss(N[0])
    x1, w1 = O.leggauss(N[1])
    kq = sp.lambdify((x, y), k)(X0[:, None], X1[None, :])
    uq = sp.lambdify((x, y), ue)(X0[:, None], X1[None, :])
    lift_q = Lo.values(np.zeros((N[0] - 2, N[1] - 2)), X0, X1)
    T0 = Lo.axes[0].S @ (O.vandermonde("legendre", x0, N[0]).T * w0[None, :])
    T1 = Lo.axes[1].S @ (O.vandermonde("legendre", x1, N[1]).T * w1[None, :])
    br = T0 @ (kq * (uq - lift_q)) @ T1.T
    assert np.linalg.norm(br) > 1e-2
    assert O.rel_l2(_np(b), br) < 1e-12
    # ue lies in the lifted space: its homogeneous coefficients satisfy A u_h = b
    uh = Lo.forward(uq)
    assert O.rel_l2(_np(A(uh)), _np(b)) < 1e-12


@pytest.mark.parametrize("lifted", [(2,), (1, 2)])
def test_directsumtps_3d_polynomial(dev, lifted):
    # 3D products with one or two lifted axes: the boundary data of the lifted
    # axes (functions of the other two coordinates) are projected onto the
    # remaining axes, recursively lifted; a polynomial solution is reproduced
    from jaxfun_amd import Div, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, TrialFunction, inner
    from jaxfun_amd.coordinates import x, y, z
    from jaxfun_amd.galerkin import DirectSumTPS

    # y^2 z^2 is not reproduced by the lift, so the homogeneous part is not zero
    ue = (1 - x**2) * (1 + y + x * y**2 + z**2 * y**2 + x * z**3)
    if 1 not in lifted:
        ue = ue * (1 - y**2)
    syms = (x, y, z)
    spaces = []
    for ax in range(3):
        if ax in lifted:
            s_ = syms[ax]
            bcs = {"left": {"D": ue.subs(s_, -1)}, "right": {"D": ue.subs(s_, 1)}}
        else:
            bcs = {"left": {"D": 0}, "right": {"D": 0}}
        spaces.append(FunctionSpace(10, Legendre, bcs))
    T = TensorProduct(*spaces)
    assert isinstance(T, DirectSumTPS) and len(T.tpspaces) == 2 ** len(lifted)
    assert not T.lift_coefficients().is_complex()
    v, u = TestFunction(T), TrialFunction(T)
    A, b = inner(v * Div(Grad(u)) - v * Div(Grad(ue)), sparse=True, kind="system")
    uh = A.solve(b)
    uj = _np(T.backward(uh))
    xj = [_np(m) for m in T.mesh()]
    assert np.abs(uj - sp.lambdify((x, y, z), ue)(*xj)).max() < 1e-11
    assert np.abs(_np(uh)).max() > 1e-3
    assert np.abs(_np(T.forward(T.backward(uh))) - _np(uh)).max() < 1e-12


def test_directsumtps_fourier_3d(dev):
    # tests/galerkin/test_tensorproductspace_extra.py:42-71: Fourier x lifted
    # Chebyshev x lifted Chebyshev, complex boundary data
    from jaxfun_amd import Chebyshev, Div, Fourier, FunctionSpace, Grad, TensorProduct, TestFunction, \
        TrialFunction, inner, ulp
    from jaxfun_amd.coordinates import x, y, z

    N = 20
    ue = sp.sin(2 * x) * sp.exp(2 * y + z * sp.I)
    bcsy = {"left": {"D": ue.subs(y, -1)}, "right": {"D": ue.subs(y, 1)}}
    bcsz = {"left": {"D": ue.subs(z, -1)}, "right": {"D": ue.subs(z, 1)}}
    F = Fourier(N)
    Dy = FunctionSpace(N, Chebyshev, bcs=bcsy)
    Dz = FunctionSpace(N + 2, Chebyshev, bcs=bcsz)
    T = TensorProduct(F, Dy, Dz)
    v, u = TestFunction(T), TrialFunction(T)
    A, b = inner(Div(Grad(u)) * v - Div(Grad(ue)) * v, kind="system")
    uh = A.solve(b, method="lu")
    uj = _np(T.backward(uh))
    uej = sp.lambdify((x, y, z), ue)(*[_np(m) for m in T.mesh()])
    assert np.linalg.norm(uj - uej) < np.sqrt(ulp(1))
    assert np.linalg.norm(_np(T.forward(T.backward(uh))) - _np(uh)) < ulp(1000) * max(1.0, np.linalg.norm(_np(uh)))


def test_tensor_evaluate_scattered(dev):
    # TensorProductSpace.evaluate at scattered points (tensorproductspace.py:263-321)
    # equals the backward transform at the same (mesh) points; DirectSumTPS.evaluate
    # equals the sum over its split products (test_tensorproductspace_extra.py:253-258)
    from jaxfun_amd import Fourier, FunctionSpace, Legendre, TensorProduct

    rng = np.random.default_rng(7)
    T = TensorProduct(FunctionSpace(10, Legendre, {"left": {"D": 0}, "right": {"D": 0}}), Legendre(9))
    c = rng.standard_normal((2, 8, 9))
    X, Y = [_np(m) for m in T.mesh()]
    XX, YY = np.broadcast_arrays(X, Y)
    got = _np(T.evaluate([XX.ravel(), YY.ravel()], c)).reshape(2, *XX.shape)
    assert O.rel_l2(got, _np(T.backward(c))) < 1e-13
    T3 = TensorProduct(Fourier(8), Legendre(6), Legendre(5))
    c3 = rng.standard_normal((8, 6, 5)) + 1j * rng.standard_normal((8, 6, 5))
    M = [_np(m) for m in T3.mesh()]
    B = np.broadcast_arrays(*M)
    got = _np(T3.evaluate([b.ravel() for b in B], c3)).reshape(B[0].shape)
    assert O.rel_l2(got, _np(T3.backward(c3))) < 1e-13
    N = (5, 6)
    F1 = FunctionSpace(N[0], Legendre, bcs={"left": {"D": 1}, "right": {"D": 2}}, domain=(-2, 2))
    F2 = FunctionSpace(N[1], Legendre, bcs={"left": {"D": 3}, "right": {"D": 4}}, domain=(-2, 2))
    TD = TensorProduct(F1, F2)
    cd = rng.standard_normal(TD.num_dofs)
    pts = [rng.uniform(-2, 2, 13), rng.uniform(-2, 2, 13)]
    z0 = sum(_np(v.evaluate(pts, TD.bndvals.get(f, cd))) for f, v in TD.tpspaces.items())
    assert np.linalg.norm(z0 - _np(TD.evaluate(pts, cd))) < 1e-12
