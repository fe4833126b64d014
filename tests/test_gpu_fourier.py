"""Fourier bases and the per-wavenumber solver on the GPU against the FP64
oracle (numpy.fft restatement, dense per-wavenumber solves).

These follow the reference's tests/galerkin/test_fourier.py,
tests/la/test_tpmatrices_solvers.py:50-160 and examples/poisson2D_periodic.py.
"""

import numpy as np
import pytest
import sympy as sp

from oracle import spectral as O

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("N,Q", [(8, 8), (8, 12), (16, 24), (16, 16)])
def test_fourier_basis_kernel(dev, N, Q):
    from jaxfun_amd import Fourier

    F = Fourier(N, domain=(0.0, 3.0))
    s = 2 * np.pi / 3.0
    x = np.arange(Q) * 2 * np.pi / Q
    for d in range(4):
        k = O.fourier_wavenumbers(N, d % 2 == 1).astype(float)
        E = (1j * k[None, :] * s) ** d * np.exp(1j * np.outer(x, k))
        got = _np(F._basis(Q, d, False, 0.5))
        assert np.allclose(got, 0.5 * E, rtol=0, atol=1e-13 * max(1.0, np.abs(E).max())), d
        got = _np(F._basis(Q, d, True, 2.0))
        assert np.allclose(got, 2.0 * E.conj().T, rtol=0, atol=1e-13 * max(1.0, np.abs(E).max())), d


def test_complex_operator_contraction(dev):
    import torch

    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(0)
    A = rng.standard_normal((7, 5)) + 1j * rng.standard_normal((7, 5))
    op = K.ComplexAxisOperator(torch.tensor(A, device=dev))
    for shape, axis in (((5,), 0), ((3, 5), 1), ((5, 4), 0), ((2, 5, 3), 1)):
        Xc = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        Y = _np(K.apply_along(op, torch.tensor(Xc, device=dev), axis))
        assert np.allclose(Y, O.apply_along(A, Xc, axis), atol=1e-13)
        Xr = rng.standard_normal(shape)
        Y = _np(K.apply_along(op, torch.tensor(Xr, device=dev), axis))
        assert np.allclose(Y, O.apply_along(A, Xr, axis), atol=1e-13)
    # real operator on complex data
    R = rng.standard_normal((6, 5))
    Xc = rng.standard_normal((3, 5, 2)) + 1j * rng.standard_normal((3, 5, 2))
    Y = _np(K.apply_along(K.AxisOperator(torch.tensor(R, device=dev)), torch.tensor(Xc, device=dev), 1))
    assert np.allclose(Y, O.apply_along(R, Xc, 1), atol=1e-13)


def test_fourier_truncation_and_padding(dev):
    # tests/galerkin/test_fourier.py
    from jaxfun_amd import Fourier

    F = Fourier(8)
    c = np.zeros(8, dtype=complex)
    c[1], c[2] = 1.0, 2.0
    u = F.backward(c)
    up = F.backward(c, N=12)
    assert tuple(up.shape) == (12,)
    assert np.linalg.norm(_np(F.forward(u)) - c) < 1e-13
    assert np.linalg.norm(_np(F.forward(up)) - c) < 1e-13


@pytest.mark.parametrize("dom", [None, (-1.0, 2.0)])
def test_fourier_transforms_vs_numpy_fft(dev, dom):
    from jaxfun_amd import Fourier

    N = 16
    F = Fourier(N, domain=dom)
    Fo = O.Fourier1D(N, (0.0, 2 * np.pi) if dom is None else dom)
    rng = np.random.default_rng(1)
    c = rng.standard_normal((3, N)) + 1j * rng.standard_normal((3, N))
    for Nq in (None, 24):
        assert O.rel_l2(_np(F.backward(c, Nq)), Fo.backward(c, Nq)) < 1e-13
        for k in (1, 2, 3):
            assert O.rel_l2(_np(F.backward_primitive(c, k, Nq)), Fo.backward(c, Nq, d=k)) < 1e-13, (Nq, k)
    u = rng.standard_normal((3, 24)) + 1j * rng.standard_normal((3, 24))
    assert O.rel_l2(_np(F.forward(u)), Fo.forward(u)) < 1e-13
    assert O.rel_l2(_np(F.scalar_product(u)), Fo.scalar_product(u)) < 1e-13
    ur = rng.standard_normal((2, N))
    assert O.rel_l2(_np(F.forward(ur)), Fo.forward(ur)) < 1e-13
    xs = Fo.a + (Fo.b - Fo.a) * rng.random(11)
    assert O.rel_l2(_np(F.evaluate(xs, c[0])), Fo.evaluate(xs, c[0])) < 1e-13
    assert np.allclose(_np(F.mesh()), Fo.mesh(), atol=1e-15)
    assert O.rel_l2(_np(F.forward(F.backward(c))), c) < 1e-14


def test_fourier_legendre_tensor_transforms(dev):
    from jaxfun_amd import Fourier, FunctionSpace, Legendre, TensorProduct

    N, M = 16, 14
    T = TensorProduct(Fourier(N), FunctionSpace(M, Legendre, {"left": {"D": 0}, "right": {"D": 0}}))
    Fo = O.Fourier1D(N)
    L = O.Space1D("legendre", M, dirichlet=True)
    rng = np.random.default_rng(2)
    c = rng.standard_normal((2, N, M - 2)) + 1j * rng.standard_normal((2, N, M - 2))
    ref = Fo.backward(np.swapaxes(O.apply_along(L.backward_matrix(20), c, 2), 1, 2), 24)
    ref = np.swapaxes(ref, 1, 2)
    got = _np(T.backward(c, N=(24, 20)))
    assert O.rel_l2(got, ref) < 1e-13
    u = _np(T.backward(c))
    assert O.rel_l2(_np(T.forward(u)), c) < 1e-13
    sp_ = O.apply_along(L.scalar_product_matrix(M), np.swapaxes(Fo.scalar_product(np.swapaxes(u, 1, 2)), 1, 2), 2)
    assert O.rel_l2(_np(T.scalar_product(u)), sp_) < 1e-13


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("n,method", [(11, "auto"), (11, "lanes"), (11, "wave"), (300, "wave"), (300, "lanes")])
def test_batched_band_lu_kernel(dev, axis, n, method):
    """Per-line factors (the per-wavenumber systems): one lane per line and
    wavefronts per line (each line's own factor) against numpy."""
    import torch

    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(3)
    nsys, p, q = 6, 2, 3
    A = np.zeros((nsys, n, n))
    band = np.zeros((nsys, n, p + q + 1))
    for s in range(nsys):
        for o in range(-p, q + 1):
            i = np.arange(max(0, -o), min(n, n - o))
            v = rng.standard_normal(i.size) + (8.0 if o == 0 else 0.0)
            A[s, i, i + o] = v
            band[s, i, o + p] = v
    lu = K.band_lu_factor_batched(torch.tensor(band, device=dev), p, q)
    shape = (nsys, n) if axis == 1 else (n, nsys)
    for cplx in (False, True):
        for batch in ((), (2,)):
            X = rng.standard_normal(batch + shape)
            if cplx:
                X = X + 1j * rng.standard_normal(batch + shape)
            Y = _np(K.band_lu_solve_batched(lu, p, q, torch.tensor(X, device=dev), len(batch) + axis, method))
            Xs = np.moveaxis(X, len(batch) + axis, -1).reshape(-1, nsys, n)
            ref = np.stack([[np.linalg.solve(A[s], Xs[b, s]) for s in range(nsys)] for b in range(Xs.shape[0])])
            ref = np.moveaxis(ref.reshape(np.moveaxis(X, len(batch) + axis, -1).shape), -1, len(batch) + axis)
            assert np.abs(Y - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()), (cplx, batch)
    bad = band.copy()
    bad[2, 0, p] = 0.0
    with pytest.raises(ValueError):
        K.band_lu_factor_batched(torch.tensor(bad, device=dev), p, q)


def _poisson_fourier_poly(N, poly):
    from jaxfun_amd import Chebyshev, Div, Fourier, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, \
        TrialFunction, inner

    space = Legendre if poly == "legendre" else Chebyshev
    F = FunctionSpace(N, Fourier)
    D = FunctionSpace(N, space, {"left": {"D": 0}, "right": {"D": 0}})
    T = TensorProduct(F, D)
    v, u = TestFunction(T), TrialFunction(T)
    x, y = T.system.base_scalars()
    ue = sp.cos(2 * x) * (1 - y**2)
    A, b = inner(v * Div(Grad(u)) - v * Div(Grad(ue)), sparse=True, kind="system")
    return T, A, b, ue


@pytest.mark.parametrize("poly", ["legendre", "chebyshev"])
def test_wavenumber_solver_vs_kron(dev, poly):
    # tests/la/test_tpmatrices_solvers.py:119-133
    from jaxfun_amd.la import TPMatricesWavenumberSolver, tpmats_to_kron, tpmats_wavenumber_factor

    _, A, b, _ = _poisson_fourier_poly(16, poly)
    wn = tpmats_wavenumber_factor(A)
    assert isinstance(wn, TPMatricesWavenumberSolver) and wn.poly_axis == 1
    bh = _np(b)
    ref = np.linalg.solve(tpmats_to_kron(A.tpmats), bh.ravel()).reshape(bh.shape)
    uh = _np(wn.solve(b))
    assert uh.shape == bh.shape
    assert np.abs(uh - ref).max() < 1e-13
    # the oracle's per-wavenumber restatement on the same system
    terms = [(float(tp.coefficient), [m.todense_host() for m in tp.mats]) for tp in A.tpmats]
    assert np.abs(O.wavenumber_solve(terms, bh, 1) - ref).max() < 1e-13
    # batched right-hand sides and the solve dispatch
    ub = _np(A.solve(np.stack([bh, 2 * bh])))
    assert np.abs(ub[1] - 2 * ref).max() < 1e-12
    with pytest.raises(TypeError):
        tpmats_wavenumber_factor("not a valid input")


@pytest.mark.parametrize("poly", ["legendre", "chebyshev"])
def test_fourier_poly_poisson_manufactured(dev, poly):
    # tests/la/test_tpmatrices_solvers.py:136-146
    from jaxfun_amd import ulp

    T, A, b, ue = _poisson_fourier_poly(16, poly)
    uh = A.solve(b)
    x, y = T.system.base_scalars()
    N = 40
    uj = _np(T.backward(uh, N=(N, N)))
    xj = [_np(m) for m in T.mesh(N=(N, N))]
    uej = sp.lambdify((x, y), ue)(*xj)
    assert np.linalg.norm(uj - uej) / N < np.sqrt(ulp(10))


def test_poisson2d_periodic_example(dev):
    # examples/poisson2D_periodic.py: Fourier (N = 20) x lifted Legendre (M = 80)
    from jaxfun_amd import Div, Fourier, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, TrialFunction, \
        inner, n, ulp
    from jaxfun_amd.coordinates import x, y

    ue = (1 - y**2) * sp.cos(2 * x) * sp.exp(sp.cos(sp.pi * y))
    M, N = 80, 20
    bcs = {"left": {"D": ue.subs(y, -1)}, "right": {"D": ue.subs(y, 1)}}
    D = FunctionSpace(M, Legendre, bcs, scaling=n + 1, name="D", fun_str="psi")
    F = FunctionSpace(N, Fourier, name="F", fun_str="E")
    T = TensorProduct(F, D, name="T")
    v, u = TestFunction(T, name="v"), TrialFunction(T, name="u")
    A, b = inner(v * Div(Grad(u)) - v * Div(Grad(ue)), sparse=True, kind="system")
    uh = A.solve(b)
    Nn = 100
    uj = _np(T.backward(uh, N=(2 * Nn, 2 * M)))
    xj = [_np(m) for m in T.mesh(N=(2 * Nn, 2 * M))]
    uej = sp.lambdify((x, y), ue)(*xj)
    error = np.linalg.norm(uj - uej) / Nn
    assert error < ulp(100), error


def test_fourier_lifted_legendre_poisson(dev):
    # Fourier x inhomogeneous-Dirichlet Legendre (a DirectSumTPS with a Fourier
    # axis): the boundary functions are projected onto the Fourier basis
    from jaxfun_amd import Div, Fourier, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, TrialFunction, \
        inner, n
    from jaxfun_amd.coordinates import x, y
    from jaxfun_amd.galerkin import DirectSumTPS

    ue = (1 + y**2) * sp.cos(2 * x) + y**3 * sp.sin(x) + y
    bcs = {"left": {"D": ue.subs(y, -1)}, "right": {"D": ue.subs(y, 1)}}
    D = FunctionSpace(24, Legendre, bcs, scaling=n + 1)
    F = FunctionSpace(16, Fourier)
    T = TensorProduct(F, D)
    assert isinstance(T, DirectSumTPS)
    v, u = TestFunction(T), TrialFunction(T)
    A, b = inner(v * Div(Grad(u)) - v * Div(Grad(ue)), sparse=True, kind="system")
    uh = A.solve(b)
    uj = _np(T.backward(uh, N=(32, 30)))
    xj = [_np(m) for m in T.mesh(N=(32, 30))]
    uej = sp.lambdify((x, y), ue)(*xj) + 0 * xj[0]
    assert np.abs(uj - uej).max() < 1e-12
    # forward of the values returns the homogeneous coefficients
    assert O.rel_l2(_np(T.forward(T.backward(uh))), _np(uh)) < 1e-12


def test_complex_band_and_cmul_kernels(dev):
    import torch

    from jaxfun_amd import kernels as K

    rng = np.random.default_rng(4)
    nsys, n, p, q = 5, 9, 1, 2
    A = np.zeros((nsys, n, n), dtype=complex)
    band = np.zeros((nsys, n, p + q + 1), dtype=complex)
    for s in range(nsys):
        for o in range(-p, q + 1):
            i = np.arange(max(0, -o), min(n, n - o))
            v = rng.standard_normal(i.size) + 1j * rng.standard_normal(i.size) + (6.0 if o == 0 else 0.0)
            A[s, i, i + o] = v
            band[s, i, o + p] = v
    lu = K.band_lu_factor_batched(torch.tensor(band, device=dev), p, q)
    X = rng.standard_normal((2, nsys, n)) + 1j * rng.standard_normal((2, nsys, n))
    Y = _np(K.band_lu_solve_batched(lu, p, q, torch.tensor(X, device=dev), 2))
    ref = np.stack([[np.linalg.solve(A[s], X[b, s]) for s in range(nsys)] for b in range(2)])
    assert np.abs(Y - ref).max() < 1e-12
    D = rng.standard_normal((3, 4)) + 1j * rng.standard_normal((3, 4))
    Z = rng.standard_normal((2, 3, 4)) + 1j * rng.standard_normal((2, 3, 4))
    got = _np(K.cmul_bcast(torch.tensor(Z, device=dev), torch.tensor(D, device=dev)))
    assert np.allclose(got, Z * D[None], atol=1e-15)


def test_fourier_first_derivative_1d(dev):
    # u' + u = f on a periodic interval: complex diagonal matrices
    from jaxfun_amd import Fourier, TestFunction, TrialFunction, inner

    F = Fourier(16, domain=(0.0, 3.0))
    x = F.system.x
    u, v = TrialFunction(F), TestFunction(F)
    s = 2 * sp.pi / 3
    ue = sp.sin(s * x) + sp.cos(3 * s * x)
    A, b = inner(v * (sp.diff(u, x) + u) - v * (sp.diff(ue, x) + ue), kind="system")
    assert A.is_complex
    uh = _np(A.solve(b))
    Fo = O.Fourier1D(16, (0.0, 3.0))
    ref = Fo.forward(sp.lambdify(x, ue)(Fo.mesh()))
    assert np.abs(uh - ref).max() < 1e-13


@pytest.mark.parametrize("poly", ["legendre", "chebyshev"])
def test_fourier_poly_advection_diffusion(dev, poly):
    # -Lap u + u_x = f: complex per-wavenumber weights (i k) on the Fourier axis
    from jaxfun_amd import Chebyshev, Div, Fourier, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, \
        TrialFunction, inner
    from jaxfun_amd.la import TPMatricesWavenumberSolver, tpmats_to_kron

    space = Legendre if poly == "legendre" else Chebyshev
    N = 16
    T = TensorProduct(FunctionSpace(N, Fourier), FunctionSpace(N, space, {"left": {"D": 0}, "right": {"D": 0}}))
    v, u = TestFunction(T), TrialFunction(T)
    x, y = T.system.base_scalars()
    ue = (sp.cos(2 * x) + sp.sin(x) * y) * (1 - y**2)
    A, b = inner(-v * Div(Grad(u)) + v * sp.diff(u, x) - v * (-Div(Grad(ue)) + sp.diff(ue, x)), sparse=True,
                 kind="system")
    wn = A.lu_factor()
    assert isinstance(wn, TPMatricesWavenumberSolver) and wn.lu.is_complex()
    uh = _np(A.solve(b))
    bh = _np(b)
    ref = np.linalg.solve(tpmats_to_kron(A.tpmats), bh.ravel()).reshape(bh.shape)
    assert np.abs(uh - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())
    uj = _np(T.backward(uh, N=(24, 20)))
    xj = [_np(m) for m in T.mesh(N=(24, 20))]
    assert np.abs(uj - sp.lambdify((x, y), ue)(*xj)).max() < 1e-12


def test_poisson1d_periodic_example(dev):
    # examples/poisson1D_periodic.py: complex manufactured solution, the k = 0
    # mode pinned
    from jaxfun_amd import Div, Fourier, FunctionSpace, Grad, TestFunction, TrialFunction, inner, ulp

    D = FunctionSpace(10, Fourier, name="D", fun_str="psi", domain=(-2 * np.pi, 2 * np.pi))
    v, u = TestFunction(D), TrialFunction(D)
    x = D.system.x
    ue = sp.cos(2 * x) + sp.I * sp.sin(1 * x)
    A, b = inner(v * Div(Grad(u)) - v * Div(Grad(ue)), sparse=True, kind="system")
    uh = A.pin({0: 0.0}).solve(b)
    uj = _np(D.backward(uh))
    uej = sp.lambdify(x, ue)(_np(D.mesh()))
    assert np.linalg.norm(uj - uej) < ulp(1000)
