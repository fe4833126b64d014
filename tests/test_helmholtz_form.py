"""Recognition of the Legendre-Dirichlet Helmholtz operator in a 2D
TPMatrices (the route from the reference's operator API into the ragged
batch), on host matrices: exact forms, scaled forms, and every rejection."""

import numpy as np
import pytest

from jaxfun_amd.helmholtz import helmholtz_form, legendre_dirichlet_pair
from jaxfun_amd.la.matrix import Matrix, SolverNotApplicable
from jaxfun_amd.la.tpmatrix import TPMatrices, TPMatrix
from oracle import spectral as O


def _tp(pairs):
    return TPMatrices([TPMatrix([Matrix(a), Matrix(b)], s) for s, a, b in pairs])


def test_pair_matches_oracle_closed_forms():
    for N in (6, 33, 130):
        M, Kd = legendre_dirichlet_pair(N - 2)
        assert np.allclose(M, O.Space1D("legendre", N, True).mass(), rtol=0, atol=1e-15)
        assert np.array_equal(Kd, O.legendre_dirichlet_stiffness(N))


@pytest.mark.parametrize("alpha,kappa", [(0.0, 1.0), (3.5, 1.0), (1e4, -2.0), (0.01, 0.25)])
def test_recognises_scaled_forms(alpha, kappa):
    n = 12
    M, Kd = legendre_dirichlet_pair(n)
    # kappa (alpha M(x)M - K(x)M - M(x)K), with one factor pre-combined
    A = _tp([(kappa * alpha, M, M), (-kappa, Kd, M), (-2.0 * kappa, 0.5 * M, Kd)])
    a, k, nn = helmholtz_form(A)
    assert nn == n and abs(a - alpha) <= 1e-12 * max(1.0, alpha) and abs(k - kappa) <= 1e-12 * abs(kappa)
    # a factor written as a combination: (2M + K) (x) M - 2 M(x)M ... = K(x)M
    B = _tp([(kappa * alpha, M, M), (-kappa, 2 * M + Kd, M), (2 * kappa, M, M), (-kappa, M, Kd)])
    a2, k2, _ = helmholtz_form(B)
    assert abs(a2 - alpha) <= 1e-11 * max(1.0, alpha) and abs(k2 - kappa) <= 1e-12 * abs(kappa)


def test_rejections():
    n = 10
    M, Kd = legendre_dirichlet_pair(n)
    with pytest.raises(SolverNotApplicable):  # K(x)K term
        helmholtz_form(_tp([(1.0, M, M), (-1.0, Kd, M), (-1.0, M, Kd), (1.0, Kd, Kd)]))
    with pytest.raises(SolverNotApplicable):  # unequal mixed terms
        helmholtz_form(_tp([(1.0, M, M), (-1.0, Kd, M), (-2.0, M, Kd)]))
    with pytest.raises(SolverNotApplicable):  # negative shift
        helmholtz_form(_tp([(-1.0, M, M), (-1.0, Kd, M), (-1.0, M, Kd)]))
    Mc = O.Space1D("chebyshev", n + 2, True).mass()
    with pytest.raises(SolverNotApplicable):  # another basis
        helmholtz_form(_tp([(1.0, Mc, Mc), (-1.0, Kd, Mc), (-1.0, Mc, Kd)]))
    with pytest.raises(SolverNotApplicable):  # not 2D
        helmholtz_form(TPMatrices([TPMatrix([Matrix(M), Matrix(M), Matrix(M)], 1.0)]))
