"""Error budget of the split-precision MFMA products in the ragged Helmholtz
batch, emulated on the CPU: each operand of the two eigenvector transforms
is cut into 16-bit terms exactly as the kernels cut it (round to nearest
even), the products the kernels issue are summed in FP32, and the result is
compared with the oracle's FP64 partial diagonalisation.

* 2-term fp16 split of power-of-two-scaled blocks (3 products, the bench's
  headline path) and 3-term bf16 split (6 products) both stay well inside
  the 1e-6 budget;
* a 2-term bf16 split (3 products) does not, which is why it is not used.
"""

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import spectral as O


def _bf16(x):
    x = np.asarray(x, np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _acc32(terms):
    acc = np.zeros(terms[0].shape, np.float32)
    for t in terms:
        acc = (acc + t.astype(np.float32)).astype(np.float32)
    return acc.astype(np.float64)


def _split_bf16(x, nterms):
    out, r = [], np.asarray(x, np.float64)
    for _ in range(nterms):
        t = _bf16(r.astype(np.float32))
        out.append(t)
        r = r - t
    return out


def _mm_bf16(A, B, nterms):
    a = _split_bf16(A, nterms)  # the table, from FP64
    b = _split_bf16(np.asarray(B, np.float32), nterms)
    # products of level s + t < nterms, smallest first
    terms = [a[s] @ b[t] for lev in range(nterms - 1, -1, -1) for s in range(lev + 1) for t in [lev - s]]
    return _acc32(terms)


def _p2(m):
    return 14 - int(np.frexp(m)[1]) if m > 0 else 0


def _mm_f16(A, B):
    ea = _p2(np.abs(A).max())
    x = np.ldexp(A, ea)
    ah = x.astype(np.float16).astype(np.float64)
    al = (x - ah).astype(np.float16).astype(np.float64)
    Bs = np.asarray(B, np.float32)
    eb = _p2(float(np.abs(Bs).max()))
    y = np.ldexp(Bs.astype(np.float64), eb).astype(np.float32)
    bh = y.astype(np.float16).astype(np.float32)
    bl = (y - bh).astype(np.float16).astype(np.float64)
    return np.ldexp(_acc32([al @ bh.astype(np.float64), ah @ bl, ah @ bh.astype(np.float64)]), -(ea + eb))


def _solve(N, alpha, F, mm):
    sp = O.Space1D("legendre", N, dirichlet=True)
    M = sp.mass()
    K = O.legendre_dirichlet_stiffness(N)
    lam, V = sla.eigh(K, M)
    G = mm(V.T, F.astype(np.float32))
    U = np.empty_like(G)
    n = N - 2
    for i in range(n):
        Ai = (alpha - lam[i]) * M - K
        ab = np.zeros((5, n))
        for o in range(-2, 3):
            d = np.diagonal(Ai, o)
            if o >= 0:
                ab[2 - o, o:] = d
            else:
                ab[2 - o, :o] = d
        U[i] = sla.solve_banded((2, 2), ab, G[i])
    return mm(V, U.astype(np.float32)).astype(np.float32)


@pytest.fixture(scope="module")
def cases():
    from jaxfun_amd import workloads as W

    rng = np.random.default_rng(21)
    out = []
    for N, alpha in [(34, 0.05), (66, 3.0), (100, 1e4), (131, 0.5), (200, 40.0)]:
        F = W.dirichlet_load(W.forcing_coefficients(rng, N), W.legendre_norms(N))
        out.append((N, alpha, F, O.helmholtz_partial_fd_lapack(N, alpha, F)))
    return out


def _err(cases, mm):
    num = den = 0.0
    for N, alpha, F, ref in cases:
        U = _solve(N, alpha, F, mm)
        num += np.sum((U - ref) ** 2)
        den += np.sum(ref ** 2)
    return float(np.sqrt(num / den))


def test_fp16_two_term_split_meets_budget(cases):
    assert _err(cases, _mm_f16) < 4e-7


def test_bf16_three_term_split_meets_budget(cases):
    assert _err(cases, lambda A, B: _mm_bf16(A, B, 3)) < 2e-7


def test_bf16_two_term_split_misses_budget(cases):
    assert _err(cases, lambda A, B: _mm_bf16(A, B, 2)) > 1e-6
