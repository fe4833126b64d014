"""Pins for the CPU oracle: every check here is independent of the oracle's own
formulas (numpy.polynomial, sympy integrals, the reference's golden vectors,
brute-force dense solves, exact modal decay)."""

import json
import os

import numpy as np
import numpy.polynomial.chebyshev as C
import numpy.polynomial.legendre as L
import pytest
import scipy.linalg as sla
import sympy as sp

from oracle import spectral as O

HERE = os.path.dirname(__file__)


FNS = {"legendre": (L.legval, L.legder), "chebyshev": (C.chebval, C.chebder)}


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_vandermonde_vs_numpy(family, k):
    # reference tests/galerkin/test_approximations.py:48-87 compares against numpy
    N = 17
    x = np.linspace(-1, 1, 29)
    V = O.vandermonde(family, x, N, k)
    val, der = FNS[family]
    ref = np.stack([val(x, der(np.eye(N)[n], k)) for n in range(N)], axis=1)
    assert np.allclose(V, ref, rtol=1e-12, atol=1e-10 * max(1.0, np.abs(ref).max()))


@pytest.mark.parametrize("N", [1, 2, 5, 11, 64, 100, 128, 201, 256])
def test_leggauss_vs_numpy(N):
    x, w = O.leggauss(N)
    xr, wr = L.leggauss(N)
    assert np.allclose(x, xr, atol=2e-15)
    # numpy's weights lose ~1e-11 near the endpoints; mpmath pins them below
    assert np.allclose(w, wr, rtol=1e-10, atol=1e-16)
    # exact for polynomials of degree <= 2N-1
    for m in range(0, 2 * N, max(1, N // 3)):
        exact = 0.0 if m % 2 else 2.0 / (m + 1)
        assert abs(np.sum(w * x**m) - exact) < 1e-13


def test_leggauss_vs_mpmath():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 40
    for N in (64, 128):
        x, w = O.leggauss(N)
        for i in (0, 3, N // 2):
            xi = mpmath.findroot(lambda t: mpmath.legendre(N, t), x[i])
            wi = 2 / ((1 - xi**2) * mpmath.diff(lambda t: mpmath.legendre(N, t), xi) ** 2)
            assert abs(x[i] - float(xi)) < 1e-16
            assert abs(w[i] - float(wi)) < 5e-13 * float(wi)


@pytest.mark.parametrize("N", [4, 11, 64, 192])
def test_chebgauss_exactness(N):
    x, w = O.chebgauss(N)
    assert np.all(np.diff(x) > 0)
    T = C.chebvander(x, N - 1)
    G = (T.T * w) @ T
    h = np.full(N, np.pi / 2)
    h[0] = np.pi
    assert np.allclose(G, np.diag(h), atol=1e-12)


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
@pytest.mark.parametrize("k", [1, 2])
def test_derivative_coeffs(family, k):
    c = np.random.default_rng(0).standard_normal(13)
    ref = FNS[family][1](c, k)
    got = O.derivative_coeffs(family, c, k)
    assert np.allclose(got[: len(ref)], ref, rtol=1e-12, atol=1e-10)
    assert np.allclose(got[len(ref):], 0)


def test_golden_shenfun_diagonals():
    g = json.load(open(os.path.join(HERE, "golden", "shenfun_inner_N11.json")))
    N = 11
    for family in ("legendre", "chebyshev"):
        # answer1: off-diagonals of (x u, v) in the orthogonal basis
        M = O.quad_matrix(family, N, 0, 0, weight=lambda x: x)
        a1 = np.array(g["answer1_offdiag_x_u_v"][family])
        assert np.allclose(np.diag(M, 1), a1) and np.allclose(np.diag(M, -1), a1)
        # answer2: (x u + u', v) on the Dirichlet composite basis
        S = O.dirichlet_stencil(N)
        Z = S @ (O.quad_matrix(family, N, 0, 0, weight=lambda x: x)
                 + O.orth_derivative_matrix(family, N, N, 1)) @ S.T
        for d, vals in g["answer2_composite_x_u_v_plus_du_v"][family].items():
            assert np.allclose(np.diag(Z, int(d)), vals, atol=1e-12), (family, d)


def test_sympy_exact_legendre_dirichlet_mass_stiffness():
    N = 7
    x = sp.Symbol("x")
    phi = [sp.legendre(k, x) - sp.legendre(k + 2, x) for k in range(N - 2)]
    Mx = np.array([[float(sp.integrate(phi[i] * phi[j], (x, -1, 1))) for j in range(N - 2)] for i in range(N - 2)])
    Kx = np.array([[float(sp.integrate(phi[i] * sp.diff(phi[j], x, 2), (x, -1, 1))) for j in range(N - 2)]
                   for i in range(N - 2)])
    sp1 = O.Space1D("legendre", N, dirichlet=True)
    assert np.allclose(sp1.mass(), Mx, atol=1e-14)
    assert np.allclose(O.galerkin_matrix("legendre", N, 0, 2), Kx, atol=1e-13)
    assert np.allclose(O.legendre_dirichlet_stiffness(N), Kx, atol=1e-13)
    # quadrature assembly agrees with the precomputed closed form
    assert np.allclose(O.galerkin_matrix("legendre", N, 0, 2, Nq=N), Kx, atol=1e-12)


def test_sympy_exact_chebyshev_dirichlet_stiffness():
    N = 7
    x = sp.Symbol("x")
    w = 1 / sp.sqrt(1 - x**2)
    phi = [sp.chebyshevt(k, x) - sp.chebyshevt(k + 2, x) for k in range(N - 2)]
    # substitute x = cos(t) for the weighted integral
    t = sp.Symbol("t")
    def wint(f):
        return float(sp.integrate(sp.expand(f.subs(x, sp.cos(t))), (t, 0, sp.pi)))
    Kx = np.array([[wint(phi[i] * sp.diff(phi[j], x, 2)) for j in range(N - 2)] for i in range(N - 2)])
    assert np.allclose(O.galerkin_matrix("chebyshev", N, 0, 2), Kx, atol=1e-12)
    assert np.allclose(O.galerkin_matrix("chebyshev", N, 0, 2, Nq=N), Kx, atol=1e-11)
    assert np.allclose(np.tril(Kx, -1), 0, atol=1e-12)  # upper-triangular, not symmetric
    del w


def test_banded_lu_vs_dense():
    rng = np.random.default_rng(3)
    n = 40
    A = np.zeros((n, n))
    for d in range(-2, 3):
        A += np.diag(rng.standard_normal(n - abs(d)), d)
    A += 8 * np.eye(n)
    b = rng.standard_normal((n, 3))
    assert np.allclose(O.banded_lu_solve(A, b, 2, 2), np.linalg.solve(A, b), rtol=1e-12, atol=1e-12)
    A0 = A.copy()
    A0[0, 0] = 0.0
    with pytest.raises(ValueError):
        O.banded_lu_solve(A0, b, 2, 2)


def test_apply_along_vs_einsum():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((3, 5, 6, 7))
    A = rng.standard_normal((4, 6))
    assert np.allclose(O.apply_along(A, X, 2), np.einsum("mk,bikj->bimj", A, X))


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
@pytest.mark.parametrize("dirichlet", [False, True])
def test_forward_backward_identity(family, dirichlet):
    # reference tests/galerkin/test_forward_backward.py:31-104 (identity to ulp)
    N = 16
    sp1 = O.Space1D(family, N, dirichlet)
    c = np.random.default_rng(5).standard_normal((sp1.dim, sp1.dim, sp1.dim))
    u = O.tp_backward([sp1] * 3, c)
    assert np.allclose(O.tp_forward([sp1] * 3, u), c, atol=1e-12)
    # padded backward then truncating forward (3/2 rule) is still the identity
    up = O.tp_backward([sp1] * 2, c[0], Nq=(24, 24))
    assert np.allclose(O.tp_forward([sp1] * 2, up), c[0], atol=1e-12)
    if family == "chebyshev" and not dirichlet:
        x, _ = O.chebgauss(N)
        assert np.allclose(sp1.backward_matrix() @ c[0, 0], C.chebval(x, c[0, 0]), atol=1e-12)


def _poisson_terms(family, N):
    sp1 = O.Space1D(family, N, dirichlet=True)
    M = sp1.mass()
    K = O.galerkin_matrix(family, N, 0, 2)
    return sp1, M, K


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
def test_fd_vs_kron_random_rhs(family):
    N = 16
    sp1, M, K = _poisson_terms(family, N)
    F = np.random.default_rng(6).standard_normal((N - 2, N - 2))
    terms = [(1.0, [K, M]), (1.0, [M, K])]
    Uk = O.kron_solve(terms, F)
    lam, V, W = O.fd_factor(K, M)
    one = np.ones_like(lam)
    Uf = O.fd_solve([(V, W), (V, W)], [(1.0, [lam, one]), (1.0, [one, lam])], F)
    assert O.rel_l2(Uf, Uk) < 1e-12
    if family == "chebyshev":
        # the symmetric-eigh FD (what the reference does) is wrong for this
        # non-symmetric pair on a general right-hand side
        lam2, V2 = sla.eigh(K, M)
        one2 = np.ones_like(lam2)
        Ue = O.fd_solve([(V2, V2.T), (V2, V2.T)], [(1.0, [lam2, one2]), (1.0, [one2, lam2])], F)
        assert O.rel_l2(Ue, Uk) > 1e-3


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
def test_poisson_manufactured(family):
    # examples/poisson2D.py:19-43: ue = (1-x^2)(1-y^2), -Lap ue = 2(1-y^2) + 2(1-x^2)
    N = 12
    sp1, M, K = _poisson_terms(family, N)
    x, _ = O.quad(family, N)
    X, Y = np.meshgrid(x, x, indexing="ij")
    f = 2 * (1 - Y**2) + 2 * (1 - X**2)
    b = O.tp_scalar_product([sp1, sp1], f)
    # (v, Lap u) = -(v, f)  =>  (K (x) M + M (x) K) U = -b
    U = O.kron_solve([(1.0, [K, M]), (1.0, [M, K])], -b)
    u = O.tp_backward([sp1, sp1], U)
    assert np.allclose(u, (1 - X**2) * (1 - Y**2), atol=1e-12)


def test_helmholtz_partial_fd_vs_kron():
    N = 20
    sp1, M, K = _poisson_terms("legendre", N)
    F = np.random.default_rng(7).standard_normal((N - 2, N - 2))
    for alpha in (1e-2, 3.0, 1e4):
        Uk = O.kron_solve([(alpha, [M, M]), (-1.0, [K, M]), (-1.0, [M, K])], F)
        assert O.rel_l2(O.helmholtz_partial_fd(N, alpha, F), Uk) < 1e-12


def test_helmholtz_lapack_variant_matches():
    rng = np.random.default_rng(8)
    for N in (7, 40, 101):
        F = rng.standard_normal((N - 2, N - 2))
        for alpha in (0.01, 30.0):
            assert O.rel_l2(O.helmholtz_partial_fd_lapack(N, alpha, F), O.helmholtz_partial_fd(N, alpha, F)) < 1e-12


def test_ars443_tableau_consistency():
    aE, aI = O.ARS443_EXPLICIT, O.ARS443_IMPLICIT
    c = np.array([0, 1 / 2, 2 / 3, 1 / 2, 1])
    assert np.allclose(aE.sum(1), c) and np.allclose(aI.sum(1), c)
    assert O.ars443_stage_coeffs() == [0.5]


def test_imex_ars443_linear_modal_decay_third_order():
    N, nu = 12, 0.1
    bg = O.Burgers2D(N, nu)
    lam, V, W = O.fd_factor(bg.K, bg.M)
    # eigenmode U0 = v_i (x) v_j decays as exp(nu (lam_i + lam_j) t)
    order = np.argsort(np.abs(lam))
    i, j = order[0], order[1]
    U0 = np.outer(V[:, i], V[:, j])
    T = 1.0
    errs = []
    for nsteps in (10, 20, 40):
        dt = T / nsteps
        solve = bg.solver(0.5 * dt)
        U = U0.copy()
        for _ in range(nsteps):
            U = O.imex_ars443_step(U, dt, bg.mass_apply,
                                   lambda R: O.kron_solve([(1.0, [bg.M, bg.M])], R),
                                   bg.linear_apply, solve, lambda U: np.zeros_like(U))
        errs.append(O.rel_l2(U, np.exp(nu * (lam[i] + lam[j]) * T) * U0))
    assert errs[0] / errs[1] > 6 and errs[1] / errs[2] > 6, errs


def test_varcoeff_operator_vs_dense_einsum():
    N = 8
    vc = O.VarCoeffPoisson(N)
    rng = np.random.default_rng(8)
    ghat = rng.standard_normal((3, 3)) * 0.3
    kq = vc.kfield(ghat)
    # assemble_multivar einsum('pi,pk,qj,ql,pq->ikjl') (galerkin/inner.py:1016-1023)
    P0, P1 = vc.Dw, vc.D
    P2, P3 = vc.Bw, vc.B
    A1 = np.einsum("pi,pk,qj,ql,pq->ikjl", P0, P1, P2, P3, kq)
    A2 = np.einsum("pi,pk,qj,ql,pq->ikjl", vc.Bw, vc.B, vc.Dw, vc.D, kq)
    A = A1 + A2
    U = rng.standard_normal((N - 2, N - 2))
    assert np.allclose(np.einsum("ikjl,kl->ij", A, U), vc.apply(kq, U), atol=1e-12)
    # k = 1 reduces to -(K (x) M + M (x) K)
    M = O.Space1D("legendre", N, True).mass()
    K = O.legendre_dirichlet_stiffness(N)
    assert np.allclose(vc.apply(np.ones_like(kq), U), -(K @ U @ M + M @ U @ K), atol=1e-11)


def test_pcg_vs_direct():
    N = 10
    vc = O.VarCoeffPoisson(N)
    rng = np.random.default_rng(9)
    kq = vc.kfield(rng.standard_normal((4, 4)) * 0.5)
    b = rng.standard_normal((N - 2, N - 2))
    x, it, res = O.pcg(lambda U: vc.apply(kq, U), b, rtol=1e-12)
    xd = np.linalg.solve(vc.dense(kq), b.ravel()).reshape(b.shape)
    assert res <= 1e-12 and O.rel_l2(x, xd) < 1e-9 and it > 0


def test_mlp_gradients_finite_difference():
    rng = np.random.default_rng(10)
    sizes = [6, 5, 5, 4]
    params = [(rng.standard_normal((a, b)) * 0.5, rng.standard_normal(b) * 0.1)
              for a, b in zip(sizes[:-1], sizes[1:])]
    X = rng.standard_normal((3, 6))
    dY = rng.standard_normal((3, 4))
    Y, hs = O.mlp_forward(params, X)
    g = O.mlp_backward(params, hs, dY)
    eps = 1e-6
    for li in range(len(params)):
        W, b = params[li]
        for (idx, arr, gd) in [((1, 2), W, g[li][0]), ((0,), b, g[li][1])]:
            old = arr[idx]
            arr[idx] = old + eps
            yp = np.sum(dY * O.mlp_forward(params, X)[0])
            arr[idx] = old - eps
            ym = np.sum(dY * O.mlp_forward(params, X)[0])
            arr[idx] = old
            assert abs((yp - ym) / (2 * eps) - gd[idx]) < 1e-6


# ---------------------------------------------------------------- inhomogeneous Dirichlet lifting


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
def test_lift_basis_end_values(family):
    # each lifting function is 1 at its own end and 0 at the other
    val = FNS[family][0]
    S = O.lift_basis(family)
    ends = np.array([-1.0, 1.0])
    for r in range(2):
        assert np.allclose(val(ends, S[r]), np.eye(2)[r], atol=1e-15)
    assert np.allclose(S, [[0.5, -0.5], [0.5, 0.5]], atol=1e-15)


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
def test_directsum_boundary_values(family):
    # reference tests/galerkin/test_composite.py:66-80 and
    # test_additional_coverage.py:101-114: N = 6, bcs (1, 2), zero coefficients
    ax = O.Axis1D(family, 6, "lift", bvals=[1.0, 2.0])
    val = FNS[family][0]
    assert np.allclose(val(np.array([-1.0, 1.0]), ax.to_orthogonal(np.zeros(4))), [1.0, 2.0], atol=1e-15)
    # any homogeneous part vanishes at the ends
    c = np.random.default_rng(0).standard_normal(4)
    assert np.allclose(val(np.array([-1.0, 1.0]), ax.to_orthogonal(c)), [1.0, 2.0], atol=1e-14)


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
def test_lifted_helmholtz_1d_polynomial_exact(family):
    # u = 1 + 2x + x^2 - x^3 on (0, 2) lies in the discrete space: the Galerkin
    # solution reproduces it to rounding
    a, b, alpha = 0.0, 2.0, 3.0
    N = 12
    ue = np.polynomial.Polynomial([1.0, 2.0, 1.0, -1.0])
    ax = O.Axis1D(family, N, "lift", (a, b), scaling=np.arange(N - 2) + 1.0, bvals=[ue(a), ue(b)])
    c = O.lifted_helmholtz_1d(ax, alpha, lambda x: ue.deriv(2)(x) + alpha * ue(x))
    xs = np.linspace(a, b, 11)
    got = FNS[family][0](-1.0 + (xs - a) * 2.0 / (b - a), ax.to_orthogonal(c))
    assert np.abs(got - ue(xs)).max() < 1e-12
    # forward of the values on the Gauss mesh returns the coefficients
    assert np.abs(ax.forward(ax.values(c, ax.mesh())) - c).max() < 1e-12


def _poly2d():
    x, y = sp.symbols("x y")
    ue = 1 + x * y + x**2 * y - 2 * y**3 + x**3
    lap = sp.diff(ue, x, 2) + sp.diff(ue, y, 2)
    return x, y, ue, lap


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
@pytest.mark.parametrize("kinds", [("lift", "lift"), ("lift", "dir"), ("dir", "lift")])
def test_lifted_helmholtz_2d_polynomial_exact(family, kinds):
    x, y, ue, lap = _poly2d()
    dom = ((-1.0, 2.0), (0.0, 1.0))
    alpha = 2.0
    f_ue = sp.lambdify((x, y), ue)
    f_rhs = sp.lambdify((x, y), lap + alpha * ue)
    if kinds[1] == "dir":
        # homogeneous second axis: multiply by y (1 - y) so the solution vanishes there
        ue = ue * y * (1 - y)
        f_ue = sp.lambdify((x, y), ue)
        f_rhs = sp.lambdify((x, y), sp.diff(ue, x, 2) + sp.diff(ue, y, 2) + alpha * ue)
    if kinds[0] == "dir":
        ue = ue * (x + 1) * (2 - x)
        f_ue = sp.lambdify((x, y), ue)
        f_rhs = sp.lambdify((x, y), sp.diff(ue, x, 2) + sp.diff(ue, y, 2) + alpha * ue)
    ax0 = O.Axis1D(family, 14, kinds[0], dom[0])
    ax1 = O.Axis1D(family, 12, kinds[1], dom[1])
    bcs0 = (lambda s: f_ue(dom[0][0], s), lambda s: f_ue(dom[0][1], s)) if kinds[0] == "lift" else None
    bcs1 = (lambda s: f_ue(s, dom[1][0]), lambda s: f_ue(s, dom[1][1])) if kinds[1] == "lift" else None
    Lf = O.Lifted2D(ax0, ax1, bcs0, bcs1)
    X, Y = ax0.mesh(), ax1.mesh()
    F = f_rhs(X[:, None], Y[None, :]) + 0.0 * X[:, None]
    u, _, _ = O.lifted_helmholtz_2d(Lf, alpha, F)
    xs, ys = np.linspace(*dom[0], 9), np.linspace(*dom[1], 8)
    assert np.abs(Lf.values(u, xs, ys) - f_ue(xs[:, None], ys[None, :])).max() < 1e-11


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
def test_lifted_2d_forward_backward(family):
    x, y, ue, _ = _poly2d()
    f_ue = sp.lambdify((x, y), ue)
    for kinds in (("lift", "lift"), ("orth", "lift"), ("lift", "dir")):
        ax0 = O.Axis1D(family, 10, kinds[0], (-2.0, 2.0))
        ax1 = O.Axis1D(family, 9, kinds[1], (0.0, 1.0))
        bcs0 = (lambda s: f_ue(-2.0, s), lambda s: f_ue(2.0, s)) if kinds[0] == "lift" else None
        bcs1 = (lambda s: f_ue(s, 0.0), lambda s: f_ue(s, 1.0)) if kinds[1] == "lift" else None
        Lf = O.Lifted2D(ax0, ax1, bcs0, bcs1)
        c = np.random.default_rng(2).standard_normal((ax0.dim, ax1.dim))
        vals = Lf.values(c, ax0.mesh(), ax1.mesh())
        assert np.abs(Lf.forward(vals) - c).max() < 1e-12
        # the lifted function takes the (projected) boundary data on the edges
        if kinds[1] == "lift":
            edge = Lf.values(c, ax0.mesh(), np.array([0.0]))[:, 0]
            assert np.abs(edge - f_ue(ax0.mesh(), 0.0)).max() < 1e-12


def test_helmholtz2d_example_accuracy():
    # reference examples/helmholtz2D.py: ue = exp(-(x^2+y^2)) on (0, 1)^2,
    # Legendre N = 50, scaling n + 1; error / 100 < ulp(1000) on a 100^2 mesh
    x, y = sp.symbols("x y")
    ue = sp.exp(-(x**2 + y**2))
    f_ue = sp.lambdify((x, y), ue)
    f_rhs = sp.lambdify((x, y), sp.diff(ue, x, 2) + sp.diff(ue, y, 2) + ue)
    N = 50
    sc = np.arange(N - 2) + 1.0
    ax0 = O.Axis1D("legendre", N, "lift", (0.0, 1.0), scaling=sc)
    ax1 = O.Axis1D("legendre", N, "lift", (0.0, 1.0), scaling=sc)
    Lf = O.Lifted2D(ax0, ax1, (lambda s: f_ue(0.0, s), lambda s: f_ue(1.0, s)),
                    (lambda s: f_ue(s, 0.0), lambda s: f_ue(s, 1.0)))
    X, Y = ax0.mesh(), ax1.mesh()
    u, _, _ = O.lifted_helmholtz_2d(Lf, 1.0, f_rhs(X[:, None], Y[None, :]))
    xu = np.linspace(0.0, 1.0, 100)
    err = np.linalg.norm(Lf.values(u, xu, xu) - f_ue(xu[:, None], xu[None, :])) / 100
    assert err < np.spacing(1000.0), err


# ---------------------------------------------------------------- Fourier


def test_fourier_trig_polynomial_exact():
    # reference tests/galerkin/test_fourier.py: N = 8, c[1] = 1, c[2] = 2
    F = O.Fourier1D(8)
    c = np.zeros(8, dtype=complex)
    c[1], c[2] = 1.0, 2.0
    x = F.mesh()
    u = F.backward(c)
    assert np.allclose(u, np.exp(1j * x) + 2 * np.exp(2j * x), atol=1e-14)
    up = F.backward(c, 12)
    x12 = F.mesh(12)
    assert up.shape == (12,) and np.allclose(up, np.exp(1j * x12) + 2 * np.exp(2j * x12), atol=1e-14)
    assert np.abs(F.forward(u) - c).max() < 1e-14 and np.abs(F.forward(up) - c).max() < 1e-14
    # first and second derivatives of the series, on a domain of length 3
    G = O.Fourier1D(8, (0.0, 3.0))
    s = 2 * np.pi / 3.0
    xg = G.mesh()
    assert np.allclose(G.backward(c, d=1), 1j * s * np.exp(1j * s * xg) + 4j * s * np.exp(2j * s * xg), atol=1e-13)
    assert np.allclose(G.backward(c, d=2), -(s**2) * np.exp(1j * s * xg) - 8 * s**2 * np.exp(2j * s * xg), atol=1e-12)
    assert np.allclose(G.evaluate(xg, c), G.backward(c), atol=1e-14)
    # the Nyquist mode (k = -N/2) is real on the grid and dropped by odd derivatives
    cn = np.zeros(8, dtype=complex)
    cn[4] = 1.0
    assert np.allclose(F.backward(cn), np.cos(4 * x), atol=1e-14)
    assert np.abs(F.backward(cn, d=1)).max() == 0.0


def _fourier_legendre_poisson(N):
    F = O.Fourier1D(N)
    L = O.Space1D("legendre", N, dirichlet=True)
    Af = [F.matrix(0, 2), F.matrix(0, 0)]
    M = L.mass()
    Kl = O.galerkin_matrix("legendre", N, 0, 2)
    terms = [(1.0, [Af[0], M]), (1.0, [Af[1], Kl])]
    return F, L, terms


def test_wavenumber_solve_vs_kron():
    # reference tests/la/test_tpmatrices_solvers.py:126-133 (agreement with the Kronecker solve)
    N = 16
    F, L, terms = _fourier_legendre_poisson(N)
    rng = np.random.default_rng(0)
    B = rng.standard_normal((N, N - 2)) + 1j * rng.standard_normal((N, N - 2))
    # the k = 0 row of (v, u_xx) vanishes; K keeps that block non-singular
    U = O.wavenumber_solve(terms, B, 1)
    Ak = sum(s * np.kron(m[0], m[1]) for s, m in terms)
    Ur = np.linalg.solve(Ak, B.ravel()).reshape(B.shape)
    assert np.abs(U - Ur).max() < 1e-12 * max(1.0, np.abs(Ur).max())


def test_fourier_legendre_poisson_manufactured():
    # reference tests/la/test_tpmatrices_solvers.py:136-146: ue = cos(2x)(1-y^2)
    N = 16
    F, L, terms = _fourier_legendre_poisson(N)
    x = F.mesh()
    y, wy = O.leggauss(N)
    lap = (-4 * np.cos(2 * x)[:, None]) * (1 - y**2)[None, :] + np.cos(2 * x)[:, None] * (-2.0)
    b = F.scalar_product(lap.T).T
    b = (b * wy[None, :]) @ O.vandermonde("legendre", y, N) @ L.S.T
    U = O.wavenumber_solve(terms, b, 1)
    xs = np.linspace(0, 2 * np.pi, 40, endpoint=False)
    ys = np.linspace(-1, 1, 40)
    uj = F.evaluate(xs, (U @ L.S @ O.vandermonde("legendre", ys, N).T))
    ue = np.cos(2 * xs)[:, None] * (1 - ys**2)[None, :]
    assert np.linalg.norm(uj - ue) / 40 < np.sqrt(np.spacing(10.0))


# ----------------------------------------------------------------- Jacobi family
JFAMS = [(0.0, 0.0, "jacobi"), (1.0, 2.0, "jacobi"), (-0.5, -0.5, "jacobi"), (0.3, -0.7, "jacobi"),
         (0.5, 0.5, "chebyshevu"), (0.5, 0.5, "ultraspherical"), (1.5, 1.5, "ultraspherical")]


def _scipy_basis(a, b, kind, n, x):
    from scipy.special import eval_chebyu, eval_gegenbauer, eval_jacobi

    if kind == "jacobi":
        return eval_jacobi(n, a, b, x)
    if kind == "chebyshevu":
        return eval_chebyu(n, x)
    lam = a + 0.5
    return eval_gegenbauer(n, lam, x) / eval_gegenbauer(n, lam, 1.0)


@pytest.mark.parametrize("a,b,kind", JFAMS)
def test_jacobi_vandermonde_vs_scipy(a, b, kind):
    # galerkin/Jacobi.py:165-202, ChebyshevU.py, Ultraspherical.py against scipy.special
    F = O.JacobiFamily(a, b, kind)
    x = np.linspace(-1, 1, 23)
    N = 30
    V = F.vandermonde(x, N)
    R = np.stack([_scipy_basis(a, b, kind, n, x) for n in range(N)], 1)
    assert np.abs(V - R).max() < 1e-13 * max(1.0, np.abs(R).max())
    if kind == "jacobi" and a == b == 0.0:
        assert np.allclose(V, L.legvander(x, N - 1), atol=1e-14)


@pytest.mark.parametrize("a,b,kind", JFAMS)
@pytest.mark.parametrize("k", [1, 2, 3])
def test_jacobi_derivatives_vs_power_basis(a, b, kind, k):
    # each Q_n converted to the power basis from exact samples (degree n, n+1
    # Chebyshev points), differentiated there and evaluated
    F = O.JacobiFamily(a, b, kind)
    N = 9
    x = np.linspace(-0.9, 0.95, 11)
    V = F.vandermonde(x, N, k)
    for n in range(N):
        xs = np.cos(np.pi * (np.arange(n + 1) + 0.5) / (n + 1))
        p = np.polynomial.Polynomial.fit(xs, _scipy_basis(a, b, kind, n, xs), n, domain=[-1, 1], window=[-1, 1])
        ref = p.deriv(k)(x) if n >= k else np.zeros_like(x)
        assert np.abs(V[:, n] - ref).max() < 1e-10 * max(1.0, np.abs(ref).max()), n


@pytest.mark.parametrize("a,b,kind", JFAMS)
def test_jacobi_quadrature_and_norms(a, b, kind):
    from scipy.special import beta as B

    F = O.JacobiFamily(a, b, kind)
    N = 20
    x, w = F.quad(N)
    assert np.all(np.diff(x) > 0)
    # nodes are the zeros of P_N; weights integrate the weight function exactly
    assert np.abs(O.JacobiFamily(a, b).raw(x, N + 1, a, b)[:, N]).max() < 1e-12 * np.abs(
        O.JacobiFamily(a, b).raw(np.array([1.0]), N + 1, a, b)).max()
    assert abs(w.sum() - 2 ** (a + b + 1) * B(a + 1, b + 1)) < 1e-13
    # orthogonality under a rule of N + 2 points gives the norms
    xq, wq = F.quad(N + 2)
    V = F.vandermonde(xq, N)
    G = (V.T * wq) @ V
    h = F.norm_squared(N)
    assert np.abs(G - np.diag(h)).max() < 1e-12 * h.max()  # roots_jacobi weights near a (1+x)^-0.7 end
    if kind == "chebyshevu":
        assert np.allclose(h, np.pi / 2)
        # galerkin/ChebyshevU.py:84-105 closed-form nodes
        t = (np.arange(N) + 1) * np.pi / (N + 1)
        assert np.allclose(x, np.cos(t + np.pi), atol=1e-14)
        assert np.allclose(w, np.pi / (N + 1) * (1 - x**2), atol=1e-14)


@pytest.mark.parametrize("a,b,kind", JFAMS)
def test_jacobi_dirichlet_stencil(a, b, kind):
    F = O.JacobiFamily(a, b, kind)
    N = 16
    S = F.dirichlet_stencil(N)
    E = F.vandermonde(np.array([-1.0, 1.0]), N)
    assert np.abs(E @ S.T).max() < 1e-13
    k = np.arange(N - 2)
    if a == b:
        assert np.abs(np.diag(S, 1)).max() < 1e-14
    if kind == "jacobi" and a == b == 0.0:
        assert np.allclose(S, O.dirichlet_stencil(N))
    if kind == "chebyshevu":
        assert np.allclose(np.diag(S, 2), -(k + 1) / (k + 3))
    if kind == "ultraspherical":
        assert np.allclose(np.diag(S, 2), -1.0)


@pytest.mark.parametrize("a,b,kind", JFAMS)
def test_jacobi_host_tables_vs_oracle(a, b, kind):
    # the space's host set-up (recurrence table, norms, end values, stencil)
    # in exact arithmetic against the oracle; the recurrence is run here in
    # numpy exactly as the device kernel runs it
    from jaxfun_amd.galerkin import ChebyshevU, Jacobi, Ultraspherical

    N = 30
    if kind == "jacobi":
        sp_ = Jacobi(N, alpha=sp.nsimplify(a), beta=sp.nsimplify(b))
    elif kind == "chebyshevu":
        sp_ = ChebyshevU(N)
    else:
        sp_ = Ultraspherical(N, lambda_=sp.nsimplify(a + 0.5))
    F = O.JacobiFamily(a, b, kind)
    r0, r1, r2 = sp_.recurrence_table()
    x = np.linspace(-1, 1, 13)
    prev, cur = np.zeros_like(x), np.ones_like(x)
    V = np.zeros((x.size, N))
    V[:, 0] = 1.0
    for n in range(1, N):
        prev, cur = cur, r0[n] * ((x - r1[n]) * cur - (r2[n] if n >= 2 else 0.0) * prev)
        V[:, n] = cur
    R = F.vandermonde(x, N)
    assert np.abs(V - R).max() < 1e-13 * max(1.0, np.abs(R).max())
    assert np.allclose(sp_.norm_squared(), F.norm_squared(N), rtol=1e-13)
    E = F.vandermonde(np.array([-1.0, 1.0]), N)
    assert np.allclose([sp_.bnd_value(j, "L") for j in range(N)], E[0], rtol=1e-13)
    assert np.allclose([sp_.bnd_value(j, "R") for j in range(N)], E[1], rtol=1e-13)
    from jaxfun_amd.galerkin import FunctionSpace

    kw = {"alpha": sp_.alpha, "beta": sp_.beta} if kind == "jacobi" else \
        ({"lambda_": sp_.lambda_} if kind == "ultraspherical" else {})
    C = FunctionSpace(N, type(sp_), {"left": {"D": 0}, "right": {"D": 0}}, **kw)
    assert np.allclose(C.stencil_matrix(), F.dirichlet_stencil(N), atol=1e-13)


# ----------------------------------------------------------------- general boundary conditions
BCSETS = {
    "LDLNRDRN": {"left": {"D": 0, "N": 0}, "right": {"D": 0, "N": 0}},
    "LNRN": {"left": {"N": 0}, "right": {"N": 0}},
    "LDRN": {"left": {"D": 0}, "right": {"N": 0}},
    "LDLN2RD": {"left": {"D": 0, "N2": 0}, "right": {"D": 0}},
}


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
def test_general_stencil_closed_forms(family):
    # galerkin/composite.py:786-800 (LDLNRDRN) and the Chebyshev Neumann
    # stencil T_n - n^2/(n+2)^2 T_{n+2} of get_stencil_matrix's docstring
    N = 16
    S = O.general_stencil(family, N, O.bc_conditions(BCSETS["LDLNRDRN"]))
    n = np.arange(N - 4)
    if family == "legendre":
        d2, d4 = 2 * (-2 * n - 5) / (2 * n + 7), (2 * n + 3) / (2 * n + 7)
    else:
        d2, d4 = 2 * (-n - 2) / (n + 3), (n + 1) / (n + 3)
    assert np.allclose(np.diag(S, 2), d2, atol=1e-13) and np.allclose(np.diag(S, 4), d4, atol=1e-13)
    assert np.abs(np.diag(S, 1)).max() < 1e-13 and np.abs(np.diag(S, 3)).max() < 1e-13
    dd = O.bc_conditions({"left": {"D": 0}, "right": {"D": 0}})
    assert np.allclose(O.general_stencil(family, N, dd), O.dirichlet_stencil(N), atol=1e-14)
    if family == "chebyshev":
        S = O.general_stencil(family, N, O.bc_conditions(BCSETS["LNRN"]))
        k = np.arange(N - 2)
        assert np.allclose(np.diag(S, 2), -k**2 / (k + 2) ** 2, atol=1e-14)


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
@pytest.mark.parametrize("bcs", list(BCSETS))
def test_general_stencil_meets_conditions(family, bcs):
    # every basis function meets every condition, differentiated by numpy.polynomial
    val, der = FNS[family]
    N = 14
    conds = O.bc_conditions(BCSETS[bcs])
    S = O.general_stencil(family, N, conds)
    for row in S:
        for side, k in conds:
            c = der(row, k) if k else row
            assert abs(val(-1.0 if side == "L" else 1.0, c)) < 1e-13 * max(1.0, np.abs(c).sum())
    B = O.general_lift_basis(family, conds, sum(k for _, k in conds))
    for r, row in enumerate(B):
        for c_i, (side, k) in enumerate(conds):
            c = der(row, k) if k else row
            assert abs(val(-1.0 if side == "L" else 1.0, c) - (r == c_i)) < 1e-12


def test_imex_step_general_stability_function():
    # the tableau-generic IMEX step on y' = lam y (implicit part only, M = 1)
    # equals the stability function R(z) = 1 + z b^T (I - z A)^-1 1 of the
    # implicit tableau, for a stiffly accurate and a weighted-update tableau
    from jaxfun_amd.integrators import ARS222, ARS443, IMEX_SSP2_222

    lam, dt = -3.0, 0.1
    z = lam * dt
    for tab in (ARS222, ARS443, IMEX_SSP2_222):
        A, b = np.array(tab.implicit.A), np.array(tab.implicit.b)
        R = 1 + z * b @ np.linalg.solve(np.eye(len(b)) - z * A, np.ones(len(b)))
        diag = [a for a in np.diag(A) if a != 0]
        y = O.imex_step(np.array([1.0]), dt, tab.explicit.A, tab.implicit.A, tab.explicit.b, tab.implicit.b,
                        lambda u: u, lambda r: r, lambda u: lam * u, lambda r: r / (1 - dt * diag[0] * lam),
                        lambda u: 0 * u)
        assert abs(y[0] - R) < 1e-14, (tab, y, R)
    # and for ARS443 it is the dedicated restatement
    u = np.array([0.7, -0.2])
    args = (lambda v: v, lambda r: r, lambda v: -2.0 * v, lambda r: r / (1 + 2.0 * 0.5 * 0.05), lambda v: v**2)
    a = O.imex_step(u, 0.05, ARS443.explicit.A, ARS443.implicit.A, ARS443.explicit.b, ARS443.implicit.b, *args)
    assert np.allclose(a, O.imex_ars443_step(u, 0.05, *args), atol=1e-15)


@pytest.mark.parametrize("family", ["legendre", "chebyshev"])
def test_robin_stencil_and_lift_meet_conditions(family):
    # u + a u' = 0 at x = -1 and u = 0 at x = 1, checked with numpy.polynomial
    val, der = FNS[family]
    a = 0.7
    conds = O.bc_conditions({"left": {"R": (a, 0)}, "right": {"D": 0}})
    S = O.general_stencil(family, 12, conds)
    for row in S:
        assert abs(val(-1.0, row) + a * val(-1.0, der(row, 1))) < 1e-13 * max(1.0, np.abs(row).sum())
        assert abs(val(1.0, row)) < 1e-13 * max(1.0, np.abs(row).sum())
    B = O.general_lift_basis(family, O.bc_conditions({"left": {"R": (2.0, 1)}, "right": {"D": 3}}), 0)
    assert B.shape == (2, 3)  # u + 2u' at -1 coincides on P_0, P_1 for Legendre: offset 1
    for r, row in enumerate(B):
        assert abs(val(-1.0, row) + 2.0 * val(-1.0, der(row, 1)) - (r == 0)) < 1e-13
        assert abs(val(1.0, row) - (r == 1)) < 1e-13


# ---------------------------------------------------------------- Burgers nonlinear term, variable-coefficient data
#
# The padded nonlinear scalar product of the Burgers workload and the load /
# coefficient field of the variable-coefficient workload are checked against
# values computed without the oracle's matrices: exact Chebyshev-weighted
# integrals of polynomials (sympy, moments pi (a-1)!!/a!!) and numpy's own
# Legendre evaluation and quadrature.  Each check also runs on a deliberately
# wrong variant (sign, axis, dropped term, no padding, transposed data) and
# requires that variant to fail, so the check can tell the mistakes apart.


def _cheb_weighted_moment(a: int):
    """int_{-1}^{1} x^a (1 - x^2)^(-1/2) dx / pi (exact rational)."""
    if a % 2:
        return sp.Integer(0)
    return sp.factorial2(a - 1) / sp.factorial2(a)


def _burgers_nonlinear_exact(U, N):
    """-int int phi_k(x) phi_l(y) u (u_x + u_y) w(x) w(y) dx dy for
    u = sum U_ij phi_i(x) phi_j(y), phi_k = T_k - T_{k+2}, w = (1 - x^2)^(-1/2)."""
    x, y = sp.symbols("x y")
    n = N - 2
    phi = [sp.Poly(sp.chebyshevt(k, x) - sp.chebyshevt(k + 2, x), x) for k in range(n)]
    u = sum(U[i][j] * phi[i].as_expr() * phi[j].as_expr().subs(x, y) for i in range(n) for j in range(n))
    G = sp.Poly(sp.expand(u * (sp.diff(u, x) + sp.diff(u, y))), x, y)
    dx = G.degree(x) + N
    # Mk[k][a] = int x^a phi_k(x) w dx / pi
    Mk = [[sum(c * _cheb_weighted_moment(a + m[0]) for m, c in zip(phi[k].monoms(), phi[k].coeffs()))
           for a in range(dx + 1)] for k in range(n)]
    out = np.zeros((n, n))
    terms = list(zip(G.monoms(), G.coeffs()))
    for k in range(n):
        for l in range(n):
            s = sum(c * Mk[k][a] * Mk[l][b] for (a, b), c in terms)
            out[k, l] = -float(s * sp.pi**2)
    return out


def test_burgers_nonlinear_sp_vs_exact_integral():
    # integrators/base.py:238-248 + integrators/nonlinear.py:89-133: the padded
    # (3/2-rule) scalar product of u (u_x + u_y); at N = 6 the 9-point
    # Gauss-Chebyshev rule integrates the degree-15 integrand exactly
    N = 6
    rng = np.random.default_rng(31)
    Ui = rng.integers(-9, 10, size=(N - 2, N - 2))
    Ui[0, 1], Ui[1, 0] = 5, -3  # not symmetric: u(x, y) != u(y, x)
    U = [[sp.Rational(int(v), 7) for v in row] for row in Ui]
    Uf = Ui / 7.0
    exact = _burgers_nonlinear_exact(U, N)
    bg = O.Burgers2D(N, 0.1)
    got = bg.nonlinear_sp(Uf)
    assert O.rel_l2(got, exact) < 1e-13, O.rel_l2(got, exact)

    # wrong variants must fail the same check
    u = bg.B0 @ Uf @ bg.B0.T
    ux = bg.B1 @ Uf @ bg.B0.T
    uy = bg.B0 @ Uf @ bg.B1.T
    SP = bg.SP
    wrong = {
        "sign": SP @ (u * (ux + uy)) @ SP.T,
        "axis": -(SP @ (u * (uy + uy)) @ SP.T),
        "dropped u_y": -(SP @ (u * ux) @ SP.T),
        "transposed output": (-(SP @ (u * (ux + uy)) @ SP.T)).T,
        "no padding": O.Burgers2D(N, 0.1, pad=1.0).nonlinear_sp(Uf),
    }
    for name, w in wrong.items():
        assert O.rel_l2(w, exact) > 1e-3, name


def test_varcoeff_rhs_and_kfield_vs_numpy_quadrature():
    # b_kl = (phi_k phi_l, f) for f = sum fhat_ij P_i(x) P_j(y) (the linear
    # form of inner(), galerkin/inner.py:914-984), and k = exp(g) at the
    # Legendre-Gauss mesh, both from numpy.polynomial.legendre alone
    N, m = 10, 4
    vc = O.VarCoeffPoisson(N)
    rng = np.random.default_rng(32)
    fhat = rng.standard_normal((N, N))
    xq, wq = L.leggauss(3 * N)  # exact for the degree <= 4N - 2 integrand
    E = np.eye(N)
    Phi = np.stack([L.legval(xq, E[k] - E[k + 2]) for k in range(N - 2)])  # (n, Q)
    Vq = L.legvander(xq, N - 1)
    fq = Vq @ fhat @ Vq.T  # f(x_i, y_j)
    b = (Phi * wq) @ fq @ (Phi * wq).T
    got = vc.rhs(fhat)
    assert O.rel_l2(got, b) < 1e-13
    assert O.rel_l2(vc.rhs(fhat.T), b) > 1e-2  # a transposed forcing is caught

    ghat = rng.standard_normal((m, m)) * 0.5
    xg, _ = L.leggauss(N)
    Vg = L.legvander(xg, m - 1)
    kq = np.exp(Vg @ ghat @ Vg.T)
    assert O.rel_l2(vc.kfield(ghat), kq) < 1e-13
    assert O.rel_l2(vc.kfield(ghat.T), kq) > 1e-2
