"""The committed workload fixtures (tests/golden/make_config_fixtures.py)
checked without the oracle's matrices: the Chebyshev Poisson solution must
satisfy the Galerkin equations when its Laplacian is formed by
numpy.polynomial.chebyshev differentiation, and the 3D Legendre samples must
equal direct numpy.polynomial.legendre evaluation of the series."""

import os

import numpy as np
import numpy.polynomial.chebyshev as C
import numpy.polynomial.legendre as L

from jaxfun_amd import workloads as W

HERE = os.path.join(os.path.dirname(__file__), "golden")


def test_c0_fixture_satisfies_galerkin_equations():
    N = 64
    fx = np.load(os.path.join(HERE, "c0_chebyshev_poisson_N64.npz"))
    U, u = fx["U"], fx["u"]
    assert U.shape == (N - 2, N - 2) and u.shape == (N, N)
    # orthogonal coefficients of u = sum U_kl (T_k - T_{k+2})(x) (T_l - T_{l+2})(y)
    c = np.zeros((N, N))
    c[:-2, :-2] += U
    c[2:, :-2] -= U
    c[:-2, 2:] -= U
    c[2:, 2:] += U
    lap = np.zeros((N, N))
    lap[:N - 2, :] += C.chebder(c, 2, axis=0)
    lap[:, :N - 2] += C.chebder(c, 2, axis=1)
    # (f + Lap u, phi_k phi_l)_w = 0 for every test function: the load of -Lap u
    # (exact by the orthogonality of T_n) equals the load of f
    h = W.chebyshev_norms(N)
    b = W.chebyshev_poisson_load(N)
    assert np.linalg.norm(W.dirichlet_load(-lap, h) - b) < 1e-11 * np.linalg.norm(b)
    # values on the ascending Gauss-Chebyshev mesh by numpy's evaluation
    x = np.sort(C.chebgauss(N)[0])
    assert np.allclose(u, C.chebval2d(x[:, None] + 0 * x[None, :], x[None, :] + 0 * x[:, None], c),
                       rtol=0, atol=1e-13 * np.abs(u).max())
    assert np.abs(u[0, :]).max() < 0.2 * np.abs(u).max()  # near the boundary the solution is small


def test_c1_fixture_samples_vs_numpy():
    N, B = 128, 16
    fx = np.load(os.path.join(HERE, "c1_legendre_3d_samples.npz"))
    idx, vals = fx["idx"], fx["u"]
    c = W.legendre_3d_coefficients(N, B)
    x, _ = L.leggauss(N)
    V = L.legvander(x, N - 1)
    for s in range(0, len(idx), 257):
        b, i, j, k = idx[s]
        ref = np.einsum("l,m,n,lmn->", V[i], V[j], V[k], c[b])
        assert abs(ref - vals[s]) < 1e-12 * max(1.0, abs(ref)), s
