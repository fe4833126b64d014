"""Where the scaled preconditioner of the variable-coefficient workload loses
its iterations (CPU, on the oracle's restatement at N = 32; the same analysis
at the workload's N = 64 is quoted in the design notes).

P_s R = (F s B) L^-1 (F s B)^T R with s = k^-1/2 clusters the spectrum of
P_s A at 1; what spread remains lies in the top-degree modes, which the
Galerkin projection of the scaling truncates.  A correction of the scaling
field's low modes (what the learned network adds), a coarse correction on
the low modes or a better initial guess cannot reach those modes; these
tests pin that claim.
"""

import numpy as np
import pytest
import scipy.linalg as sla

from jaxfun_amd import workloads as W
from oracle import spectral as O

N = 32


def _setup(seed):
    vc = O.VarCoeffPoisson(N)
    g, fh = W.varcoeff_sample(seed, N=N, gmodes=16)
    kq, b, s = vc.kfield(g), vc.rhs(fh), vc.kfield(-0.5 * g)
    Fw = vc.sp.forward_matrix(N)  # (n, Q)
    lam, V = sla.eigh(-O.legendre_dirichlet_stiffness(N), vc.sp.mass())

    def Ps(R):
        a = s * (Fw.T @ R @ Fw)
        G = V.T @ (vc.B.T @ a @ vc.B) @ V
        d = V @ (G / (lam[:, None] + lam[None, :])) @ V.T
        return Fw @ (s * (vc.B @ d @ vc.B.T)) @ Fw.T

    return vc, kq, b, Ps


def _dense(op, n):
    E = np.zeros((n, n))
    cols = []
    for j in range(n * n):
        E.flat[j] = 1.0
        cols.append(op(E).ravel())
        E.flat[j] = 0.0
    return np.array(cols).T


@pytest.fixture(scope="module")
def spectrum():
    vc, kq, b, Ps = _setup(40000)
    n = N - 2
    A = vc.dense(kq)
    P = _dense(Ps, n)
    P = 0.5 * (P + P.T)
    C = np.linalg.cholesky(P)
    ev, Y = np.linalg.eigh(C.T @ A @ C)
    return ev, C @ Y, n


def test_scaled_spectrum_clustered_with_top_mode_outliers(spectrum):
    ev, Vv, n = spectrum
    i = np.arange(n)
    deg = np.maximum(i[:, None], i[None, :]).ravel()
    assert ev.min() > 0.5 and ev.max() / ev.min() < 2.5
    assert np.mean(np.abs(ev - 1.0) < 0.05) > 0.6  # most of the spectrum sits at 1
    # the eight eigenvalues farthest from 1 live in the upper half of the degrees
    for k in np.argsort(np.abs(ev - 1.0))[::-1][:8]:
        w = Vv[:, k] ** 2
        assert w[deg >= n // 2].sum() / w.sum() > 0.8


def test_low_mode_corrections_leave_iterations_unchanged():
    vc, kq, b, Ps = _setup(40001)
    A = lambda U: vc.apply(kq, U)  # noqa: E731
    n = N - 2
    x0 = Ps(b)
    _, it_s, _ = O.pcg(A, b, Ps, x0=x0, rtol=1e-10)
    _, it_plain, _ = O.pcg(A, b, lambda R: R, rtol=1e-10)
    assert it_plain > 2 * it_s  # the scaling itself is what cuts the iterations
    # exact Galerkin coarse correction of x0 on the 8 x 8 lowest modes
    m = 8
    AZ = []
    for i in range(m):
        for j in range(m):
            E = np.zeros((n, n))
            E[i, j] = 1.0
            AZ.append(A(E)[:m, :m].ravel())
    c = np.linalg.solve(np.array(AZ).T, (b - A(x0))[:m, :m].ravel())
    x1 = x0.copy()
    x1[:m, :m] += c.reshape(m, m)
    _, it_c, _ = O.pcg(A, b, Ps, x0=x1, rtol=1e-10)
    assert abs(it_c - it_s) <= 1
    # ... although the correction removes most of the low-mode error
    u = np.linalg.solve(vc.dense(kq), b.ravel()).reshape(n, n)
    assert np.linalg.norm((u - x1)[:m, :m]) < 0.05 * np.linalg.norm((u - x0)[:m, :m])


def test_optimal_modal_diagonal_correction_gains_nothing():
    """Per-sample optimum of the richest diagonal correction a network could
    output on top of the scaling, P = S P_s S with S = diag(exp(tau)) over
    all n^2 coefficients (SPD for any tau), fitted by L-BFGS to cluster the
    spectrum of P A (minimising sum_k log(lambda_k)^2): the condition number
    moves by under 5% and PCG needs the same iterations."""
    import scipy.optimize as so

    vc, kq, b, Ps = _setup(40000)
    n = N - 2
    A = vc.dense(kq)
    P = _dense(Ps, n)
    P = 0.5 * (P + P.T)
    C = np.linalg.cholesky(P)

    def f(tau):
        Z0 = np.exp(tau)[:, None] * C
        ev, Y = np.linalg.eigh(Z0.T @ A @ Z0)
        Z = Z0 @ Y
        lg = np.log(ev)
        return np.sum(lg ** 2), (4 * Z * (A @ Z)) @ (lg / ev)

    f0 = f(np.zeros(n * n))[0]
    r = so.minimize(f, np.zeros(n * n), jac=True, method="L-BFGS-B", options={"maxiter": 300})
    assert r.fun < f0  # the fit does move the spectrum
    S = np.exp(r.x)
    ev0 = np.linalg.eigvalsh(C.T @ A @ C)
    ev1 = np.linalg.eigvalsh((S[:, None] * C).T @ A @ (S[:, None] * C))
    k0, k1 = ev0.max() / ev0.min(), ev1.max() / ev1.min()
    assert k1 > 0.95 * k0
    Aop = lambda U: vc.apply(kq, U)  # noqa: E731
    Popt = lambda R: (S * (P @ (S * R.ravel()))).reshape(n, n)  # noqa: E731
    _, it0, _ = O.pcg(Aop, b, Ps, x0=Ps(b), rtol=1e-10)
    _, it1, _ = O.pcg(Aop, b, Popt, x0=Popt(b), rtol=1e-10)
    assert it1 >= it0 - 1
