"""Write the oracle fixtures of the Chebyshev Poisson (N = 64) and the batched
3D Legendre transform (16 x 128^3) workloads.

    python tests/golden/make_config_fixtures.py

Outputs (inputs are regenerated from their seeds by jaxfun_amd.workloads):

* ``c0_chebyshev_poisson_N64.npz``: ``U`` (62 x 62) -- the exact Kronecker solve
  of -Lap u = f (reference la/tpmatrix.py:1052-1086, as driven by
  examples/poisson2D.py:19-43), forcing seed 0 -- and ``u`` (64 x 64), its
  values at the Gauss-Chebyshev mesh.
* ``c1_legendre_3d_samples.npz``: ``idx`` (S x 4) random (b, i, j, k) indices
  and ``u`` (S,) the oracle's backward transform (Legendre-Gauss mesh,
  galerkin/tensorproductspace.py:330-365) of the seed-1 coefficients there.

Only tests and the benchmark read these files, as data.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from jaxfun_amd import workloads as W  # noqa: E402  (input generators only: numpy)
from oracle import spectral as O  # noqa: E402

C0_FILE = os.path.join(HERE, "c0_chebyshev_poisson_N64.npz")
C1_FILE = os.path.join(HERE, "c1_legendre_3d_samples.npz")
C1_SAMPLES = 4096


def chebyshev_poisson_oracle(N: int = 64):
    sp1 = O.Space1D("chebyshev", N, dirichlet=True)
    M = sp1.mass()
    K = O.galerkin_matrix("chebyshev", N, 0, 2)  # (phi_k, phi_l'')
    b = W.chebyshev_poisson_load(N)
    U = O.kron_solve([(1.0, [K, M]), (1.0, [M, K])], -b)  # (v, Lap u) = -(v, f)
    u = O.tp_backward([sp1, sp1], U)
    return U, u


def legendre_3d_samples(N: int = 128, B: int = 16, S: int = C1_SAMPLES):
    c = W.legendre_3d_coefficients(N, B)
    sp1 = O.Space1D("legendre", N)
    rng = np.random.default_rng(11)
    idx = np.stack([rng.integers(0, B, S), rng.integers(0, N, S), rng.integers(0, N, S),
                    rng.integers(0, N, S)], axis=1)
    vals = np.empty(S)
    for b in range(B):
        sel = np.flatnonzero(idx[:, 0] == b)
        if sel.size:
            u = O.tp_backward([sp1] * 3, c[b])
            vals[sel] = u[idx[sel, 1], idx[sel, 2], idx[sel, 3]]
    return idx.astype(np.int32), vals


def main():
    U, u = chebyshev_poisson_oracle()
    np.savez(C0_FILE, U=U, u=u)
    idx, vals = legendre_3d_samples()
    np.savez(C1_FILE, idx=idx, u=vals)
    print("wrote", C0_FILE, C1_FILE)


if __name__ == "__main__":
    main()
