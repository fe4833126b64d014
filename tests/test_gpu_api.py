"""GPU parity of the drop-in API (spaces, transforms, inner, solves) vs the oracle.

These read like the reference's own tests (tests/galerkin/test_forward_backward.py,
tests/galerkin/test_inner.py, tests/la/test_tpmatrices_solvers.py, examples/poisson2D.py)
but compare against the FP64 numpy oracle.
"""

import json
import os

import numpy as np
import pytest
import sympy as sp

from oracle import spectral as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(__file__)


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
@pytest.mark.parametrize("bc", [None, "dirichlet"])
def test_tensor_transforms_vs_oracle(dev, family, bc):
    from jaxfun_amd import Chebyshev, FunctionSpace, Legendre, TensorProduct

    N = 20
    space = Legendre if family == "legendre" else Chebyshev
    bcs = {"left": {"D": 0}, "right": {"D": 0}} if bc else None
    V = FunctionSpace(N, space, bcs)
    T = TensorProduct(V, V, V)
    s1 = O.Space1D(family, N, dirichlet=bool(bc))
    c = np.random.default_rng(0).standard_normal((2, s1.dim, s1.dim, s1.dim))
    u = T.backward(c)
    ur = O.tp_backward([s1] * 3, c)
    assert O.rel_l2(_np(u), ur) < 1e-13
    ch = T.forward(u)
    assert O.rel_l2(_np(ch), c) < 1e-12
    sp_ = T.scalar_product(u)
    assert O.rel_l2(_np(sp_), O.tp_scalar_product([s1] * 3, ur)) < 1e-13
    # padded (3/2-rule) backward and truncating forward
    T2 = TensorProduct(V, V)
    up = T2.backward(c[0, 0], N=(30, 30))
    assert O.rel_l2(_np(up), O.tp_backward([s1] * 2, c[0, 0], Nq=(30, 30))) < 1e-13
    assert O.rel_l2(_np(T2.forward(up)), c[0, 0]) < 1e-12


def test_backward_primitive_vs_oracle(dev):
    from jaxfun_amd import Chebyshev, FunctionSpace, TensorProduct

    N = 16
    V = FunctionSpace(N, Chebyshev, {"left": {"D": 0}, "right": {"D": 0}})
    T = TensorProduct(V, V)
    s1 = O.Space1D("chebyshev", N, True)
    c = np.random.default_rng(1).standard_normal((N - 2, N - 2))
    ux = T.backward_primitive(c, (1, 0), N=(24, 24))
    ref = s1.backward_matrix(24, 1) @ c @ s1.backward_matrix(24, 0).T
    assert O.rel_l2(_np(ux), ref) < 1e-13


def test_inner_golden_shenfun(dev):
    from jaxfun_amd import Chebyshev, Legendre, TestFunction, TrialFunction, inner
    from jaxfun_amd.galerkin import Composite

    g = json.load(open(os.path.join(HERE, "golden", "shenfun_inner_N11.json")))
    for name, space in (("legendre", Legendre), ("chebyshev", Chebyshev)):
        V = space(11)
        x = V.system.x
        u, v = TrialFunction(V), TestFunction(V)
        M = inner(x * v * u, sparse=True).todense_host()
        a1 = g["answer1_offdiag_x_u_v"][name]
        assert np.allclose(np.diag(M, 1), a1) and np.allclose(np.diag(M, -1), a1)
        C = Composite(11, space, {"left": {"D": 0}, "right": {"D": 0}})
        u, v = TrialFunction(C), TestFunction(C)
        M = inner(x * v * u + sp.diff(u, x) * v, sparse=False).todense_host()
        for d, vals in g["answer2_composite_x_u_v_plus_du_v"][name].items():
            assert np.allclose(np.diag(M, int(d)), vals, atol=1e-12), (name, d)


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
def test_inner_precomputed_vs_quadrature(dev, family):
    from jaxfun_amd import Chebyshev, FunctionSpace, Legendre, TestFunction, TrialFunction, inner

    space = Legendre if family == "legendre" else Chebyshev
    V = FunctionSpace(14, space, {"left": {"D": 0}, "right": {"D": 0}})
    u, v = TrialFunction(V), TestFunction(V)
    x = V.system.x
    for form, (i, j) in ((v * u, (0, 0)), (v * sp.diff(u, x, 2), (0, 2)), (sp.diff(v, x) * sp.diff(u, x), (1, 1))):
        A1 = inner(form).todense_host()
        A2 = inner(form, use_precomputed_matrices=False).todense_host()
        ref = O.galerkin_matrix(family, 14, i, j, Nq=14)
        assert np.allclose(A1, ref, atol=1e-11 * np.abs(ref).max()), (family, i, j)
        assert np.allclose(A2, ref, atol=1e-11 * np.abs(ref).max()), (family, i, j)


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
@pytest.mark.parametrize("sparse,method", [(True, "auto"), (False, "auto"), (True, "kron"), (False, "fd")])
def test_poisson2d_manufactured(dev, family, sparse, method):
    # examples/poisson2D.py: ue = (1-x^2)(1-y^2); error < ulp(1000)
    from jaxfun_amd import Chebyshev, Div, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, \
        TrialFunction, inner

    space = Legendre if family == "legendre" else Chebyshev
    D = FunctionSpace(20, space, {"left": {"D": 0}, "right": {"D": 0}})
    T = TensorProduct(D, D)
    v, u = TestFunction(T), TrialFunction(T)
    x, y = T.system.base_scalars()
    ue = (1 - x**2) * (1 - y**2)
    A, b = inner(v * Div(Grad(u)) - v * Div(Grad(ue)), sparse=sparse, kind="system")
    uh = A.solve(b, method=method)
    uj = _np(T.backward(uh))
    X, Y = [m.cpu().numpy() for m in T.mesh()]
    assert np.abs(uj - (1 - X**2) * (1 - Y**2)).max() < 1e-12


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
def test_fd_random_rhs_vs_kron(dev, family):
    from jaxfun_amd import Chebyshev, Div, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, \
        TrialFunction, inner

    space = Legendre if family == "legendre" else Chebyshev
    N = 24
    D = FunctionSpace(N, space, {"left": {"D": 0}, "right": {"D": 0}})
    T = TensorProduct(D, D)
    v, u = TestFunction(T), TrialFunction(T)
    A = inner(v * Div(Grad(u)) + 3 * v * u, sparse=True)
    F = np.random.default_rng(5).standard_normal((N - 2, N - 2))
    U = _np(A.solve(F))
    sp1, M, Kx = O.Space1D(family, N, True), None, O.galerkin_matrix(family, N, 0, 2)
    M = sp1.mass()
    Ur = O.kron_solve([(1.0, [Kx, M]), (1.0, [M, Kx]), (3.0, [M, M])], F)
    assert O.rel_l2(U, Ur) < 1e-11
    # batched right-hand sides
    Fb = np.stack([F, 2 * F])
    Ub = _np(A.solve(Fb))
    assert O.rel_l2(Ub[1], 2 * Ur) < 1e-11


def test_linear_form_vs_oracle(dev):
    from jaxfun_amd import Chebyshev, FunctionSpace, TensorProduct, TestFunction, inner

    N = 16
    D = FunctionSpace(N, Chebyshev, {"left": {"D": 0}, "right": {"D": 0}})
    T = TensorProduct(D, D)
    v = TestFunction(T)
    x, y = T.system.base_scalars()
    f = sp.exp(x) * sp.sin(y) + x * y
    b = _np(inner(v * f))
    s1 = O.Space1D("chebyshev", N, True)
    xq, _ = O.chebgauss(N)
    X, Y = np.meshgrid(xq, xq, indexing="ij")
    ref = O.tp_scalar_product([s1, s1], np.exp(X) * np.sin(Y) + X * Y)
    assert O.rel_l2(b, ref) < 1e-13


def test_1d_dirichlet_mass_solve(dev):
    from jaxfun_amd import FunctionSpace, Legendre, TestFunction, TrialFunction, inner

    V = FunctionSpace(30, Legendre, {"left": {"D": 0}, "right": {"D": 0}})
    M = inner(TestFunction(V) * TrialFunction(V), sparse=True)
    b = np.random.default_rng(2).standard_normal(28)
    x = _np(M.solve(b))
    assert np.allclose(x, np.linalg.solve(O.Space1D("legendre", 30, True).mass(), b), atol=1e-12)


def test_mixed_precision_transforms(dev):
    import torch

    from jaxfun_amd import Legendre, TensorProduct

    N = 32
    L = Legendre(N)
    T = TensorProduct(L, L, L)
    i, j, k = np.meshgrid(*(np.arange(N),) * 3, indexing="ij")
    c = np.random.default_rng(3).standard_normal((2, N, N, N)) * (1.0 + i + j + k) ** -3.0
    u64 = T.backward(c)
    for prec, tol in (("f32", 1e-6), ("bf16x6", 1e-6), ("ozaki4", 1e-8), ("ozaki5", 5e-11)):
        u = T.backward(c, precision=prec)
        assert O.rel_l2(_np(u.to(torch.float64)), _np(u64)) < tol, prec
    # the round trip amplifies the ozaki5 error (~5e-13) by ||u||/||c|| ~ N^1.5
    c2 = T.forward(u64, precision="ozaki5")
    assert O.rel_l2(_np(c2), c) < 2e-9


def test_fused_and_unfused_transforms_agree(dev):
    from jaxfun_amd import FunctionSpace, Legendre, TensorProduct

    V = FunctionSpace(40, Legendre, {"left": {"D": 0}, "right": {"D": 0}})
    T = TensorProduct(V, V, V)
    c = np.random.default_rng(9).standard_normal((2, 38, 38, 38))
    u1 = _np(T.backward(c))
    T.fused = False
    u2 = _np(T.backward(c))
    assert O.rel_l2(u1, u2) < 1e-14


def test_poisson3d_example(dev):
    # examples/poisson3D.py: Legendre-Dirichlet M = 20, scaling n + 1, method="lu"
    from jaxfun_amd import Div, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, TrialFunction, inner, \
        n, ulp

    M = 20
    D = FunctionSpace(M, Legendre, {"left": {"D": 0}, "right": {"D": 0}}, scaling=n + 1, name="D", fun_str="psi")
    T = TensorProduct(D, D, D, name="T")
    v, u = TestFunction(T), TrialFunction(T)
    x, y, z = T.system.base_scalars()
    ue = (1 - x**2) * (1 - y**2) * (1 - z**2)
    A, b = inner(v * Div(Grad(u)) - v * Div(Grad(ue)), sparse=True, kind="system")
    uh = A.solve(b, method="lu")
    uj = _np(T.evaluate_mesh(uh, kind="uniform", N=(20, 20, 20)))
    xj = [_np(m) for m in T.mesh(kind="uniform", N=(20, 20, 20))]
    uej = sp.lambdify((x, y, z), ue)(*xj)
    error = np.linalg.norm(uj - uej) / np.sqrt(T.dim)
    assert error < ulp(1000), error
