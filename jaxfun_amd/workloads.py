"""Seeded synthetic inputs of the benchmark configurations (host numpy data).

Only input generation lives here; every computation on these inputs runs in
the HIP kernels.
"""

from __future__ import annotations

import numpy as np


def forcing_coefficients(rng: np.random.Generator, N: int) -> np.ndarray:
    """f_jk ~ N(0,1) (1+j)^-2 (1+k)^-2, j, k < N (orthogonal-basis coefficients)."""
    j = np.arange(N, dtype=np.float64)
    w = (1.0 + j) ** -2.0
    return rng.standard_normal((N, N)) * np.outer(w, w)


def dirichlet_load(fhat: np.ndarray, h: np.ndarray) -> np.ndarray:
    """b = (S (x) S)(diag(h) fhat diag(h)): exact (f, phi_k phi_l) for the
    orthogonal expansion fhat and phi_k = P_k - P_{k+2}."""
    G = h[:, None] * fhat * h[None, :]
    return G[:-2, :-2] - G[2:, :-2] - G[:-2, 2:] + G[2:, 2:]


def legendre_norms(N: int) -> np.ndarray:
    return 2.0 / (2 * np.arange(N) + 1.0)


def chebyshev_norms(N: int) -> np.ndarray:
    h = np.full(N, np.pi / 2)
    h[0] = np.pi
    return h


def chebyshev_poisson_load(N: int = 64, seed: int = 0) -> np.ndarray:
    """Chebyshev-Dirichlet Poisson load vector b (N-2, N-2)."""
    rng = np.random.default_rng(seed)
    return dirichlet_load(forcing_coefficients(rng, N), chebyshev_norms(N))


def legendre_3d_coefficients(N: int = 128, B: int = 16, seed: int = 1) -> np.ndarray:
    """c_ijk ~ N(0,1) (1+i+j+k)^-3, shape (B, N, N, N), FP64."""
    rng = np.random.default_rng(seed)
    i = np.arange(N, dtype=np.float64)
    w = (1.0 + i[:, None, None] + i[None, :, None] + i[None, None, :]) ** -3.0
    out = np.empty((B, N, N, N))
    for b in range(B):
        out[b] = rng.standard_normal((N, N, N)) * w
    return out


def helmholtz_ragged_batch(nsolves: int = 4096, seed: int = 2, Nmin: int = 32, Nmax: int = 256):
    """(Ns, alphas, loads): N ~ U{Nmin..Nmax}, alpha = 10^U(-2, 4), Legendre loads."""
    rng = np.random.default_rng(seed)
    Ns = rng.integers(Nmin, Nmax + 1, size=nsolves)
    alphas = 10.0 ** rng.uniform(-2.0, 4.0, size=nsolves)
    loads = [dirichlet_load(forcing_coefficients(rng, int(N)), legendre_norms(int(N))) for N in Ns]
    return Ns, alphas, loads


def burgers_initial_conditions(N: int = 128, batch: int = 256, seed: int = 3, nmodes: int = 8):
    """Initial conditions as sums of 8 smooth Dirichlet-compatible modes
    sin(p pi (x+1)/2) sin(q pi (y+1)/2), amplitudes U(-1,1), p, q in 1..4.
    Returns a callable u0(b, X, Y) and the mode table."""
    rng = np.random.default_rng(seed)
    amp = rng.uniform(-1.0, 1.0, size=(batch, nmodes))
    pq = rng.integers(1, 5, size=(batch, nmodes, 2))

    def u0(b, X, Y):
        out = np.zeros(np.broadcast(X, Y).shape)
        for m in range(nmodes):
            out += amp[b, m] * np.sin(pq[b, m, 0] * np.pi * (X + 1) / 2) * np.sin(pq[b, m, 1] * np.pi * (Y + 1) / 2)
        return out

    return u0, amp, pq


def varcoeff_sample(seed: int, N: int = 64, gmodes: int = 16):
    """(ghat, fhat): g low-mode Legendre coefficients ~ N(0,1)(1+|m|)^-2, forcing as in the Chebyshev Poisson workload."""
    rng = np.random.default_rng(seed)
    m = np.arange(gmodes, dtype=np.float64)
    mag = np.sqrt(m[:, None] ** 2 + m[None, :] ** 2)
    ghat = rng.standard_normal((gmodes, gmodes)) * (1.0 + mag) ** -2.0
    fhat = forcing_coefficients(rng, N)
    return ghat, fhat
