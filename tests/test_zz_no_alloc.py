"""The allocation-free ragged Helmholtz entry point (jfx_helm_ragged_noalloc)
called through the C ABI: its output must equal the allocating entry point's
(jfx_helm_ragged, mode 1, 4 waves) bit for bit, on two calls in a row, the
second with a smaller batch, and the oracle to the mixed path's bound."""

import numpy as np
import pytest

from oracle import spectral as O

pytestmark = pytest.mark.gpu


def _batch(Ns, alphas, seed, nwaves=4):
    from jaxfun_amd import workloads as W
    from jaxfun_amd.helmholtz import HelmholtzBatch

    rng = np.random.default_rng(seed)
    loads = [W.dirichlet_load(W.forcing_coefficients(rng, int(N)), W.legendre_norms(int(N))) for N in Ns]
    hb = HelmholtzBatch(np.asarray(Ns), np.asarray(alphas), mode="mixed", nwaves=nwaves,
                        blocks_per_cu=2 if nwaves == 4 else 1)
    return hb, loads


def _noalloc(hb, F, U):
    from jaxfun_amd import _lib
    from jaxfun_amd import kernels as K

    _lib.call("jfx_helm_ragged_noalloc", 1, hb.nwaves, K._p(hb.items), hb.nitems, K._p(hb.qbounds), hb.nqueues,
              K._p(hb.Vs), K._p(hb.VTs), K._p(hb.lam), K._p(F), K._p(U), hb.nblocks, K._s())


@pytest.mark.parametrize("nwaves", [4, 8])
def test_noalloc_matches_allocating_entry_twice(dev, nwaves):
    import torch

    cases = [
        (np.array([34, 63, 128, 200, 256, 41, 97, 150]), np.array([0.05, 3.0, 500.0, 1e4, 0.01, 7.0, 80.0, 0.3]), 1),
        (np.array([40, 77, 130]), np.array([2.0, 0.02, 900.0]), 2),  # second call: a smaller batch
    ]
    for Ns, alphas, seed in cases:
        hb, loads = _batch(Ns, alphas, seed, nwaves)
        F = hb.pack(loads)
        U_ref = hb.solve(F)
        U = torch.full_like(F, float("nan"))
        _noalloc(hb, F, U)
        torch.cuda.synchronize()
        assert torch.equal(U, U_ref), (Ns, float((U - U_ref).abs().max()))
        for N, a, Fb, Ub in zip(Ns, alphas, loads, hb.unpack(U)):
            err = O.rel_l2(Ub, O.helmholtz_partial_fd(int(N), float(a), Fb))
            assert err < 1e-6, (int(N), err)


def test_noalloc_rejects_too_many_queues(dev):
    from jaxfun_amd import _lib

    lib = _lib.load()
    assert lib.jfx_helm_noalloc_max_queues() == 16
    hb, loads = _batch(np.array([34]), np.array([1.0]), 3)
    F = hb.pack(loads)
    U = F.clone()
    from jaxfun_amd import kernels as K

    rc = lib.jfx_helm_ragged_noalloc(1, 4, K._p(hb.items), hb.nitems, K._p(hb.qbounds), 17, K._p(hb.Vs),
                                     K._p(hb.VTs), K._p(hb.lam), K._p(F), K._p(U), hb.nblocks, K._s())
    assert rc == -2
    assert b"exceeds" in lib.jfx_last_error()
