"""Variable-coefficient operator, batched PCG, MLP kernels and
the learned preconditioner, against the FP64 oracle."""

import numpy as np
import pytest

from oracle import spectral as O

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch

    return torch.tensor(np.asarray(a), device=dev, dtype=torch.float64)


def test_gemm_bf16_all_transposes(dev):
    import torch

    from jaxfun_amd.learned import MLP

    rng = np.random.default_rng(0)
    for ta in (False, True):
        for tb in (False, True):
            M, N, K = 70, 45, 130
            A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
            B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
            C0 = rng.standard_normal((M, N)).astype(np.float32)
            bias = rng.standard_normal(N).astype(np.float32)
            At, Bt = (A.T if ta else A), (B.T if tb else B)
            C = torch.tensor(C0, device=dev)
            MLP.gemm(torch.tensor(A, device=dev), torch.tensor(B, device=dev), C, ta=ta, tb=tb, alpha=0.5,
                     beta=2.0, bias=torch.tensor(bias, device=dev), act=1)
            ref = np.tanh(0.5 * At.astype(np.float64) @ Bt + 2.0 * C0 + bias)
            # bf16 operands: |error| ~ 2^-8 sqrt(K) |a||b|
            assert np.abs(C.cpu().numpy() - ref).max() < 0.15, (ta, tb)
            assert O.rel_l2(C.cpu().numpy(), ref) < 2e-2


def _bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("M,N,K,offset", [
    (70, 45, 130, 0),     # K % 8 != 0: element loads
    (256, 512, 512, 0),   # an MLP layer: 16-byte fragment loads, 16 chunks over four waves
    (33, 97, 600, 0),     # 19 chunks: a second pass of the chunk loop, ragged tiles
    (512, 256, 256, 1),   # operands one element off 16-byte alignment: element loads
])
def test_gemm_bf16_vs_rounded_operands(dev, M, N, K, offset):
    # The MFMA products of bf16 operands are exact in FP32, so the result is
    # the FP64 product of the bf16-rounded operands up to FP32 accumulation
    # error, far below the operand rounding a wrong index or transpose shows.
    import torch

    from jaxfun_amd import _lib
    from jaxfun_amd import kernels as K_

    rng = np.random.default_rng(M + N + K)
    for ta in (0, 1):
        for tb in (0, 1):
            ash, bsh = ((K, M) if ta else (M, K)), ((N, K) if tb else (K, N))
            A = rng.standard_normal(ash).astype(np.float32)
            B = rng.standard_normal(bsh).astype(np.float32)
            Ad = torch.empty(A.size + offset, dtype=torch.float32, device=dev)[offset:].view(ash)
            Bd = torch.empty(B.size + offset, dtype=torch.float32, device=dev)[offset:].view(bsh)
            Ad.copy_(torch.tensor(A))
            Bd.copy_(torch.tensor(B))
            C = torch.empty((M, N), dtype=torch.float32, device=dev)
            _lib.call("jfx_gemm_bf16", ta, tb, M, N, K, 1.0, K_._p(Ad), ash[1], K_._p(Bd), bsh[1], 0.0, K_._p(C),
                      N, None, 0, K_._s())
            Ar, Br = _bf16(A), _bf16(B)
            At, Bt = (Ar.T if ta else Ar), (Br.T if tb else Br)
            ref = At @ Bt
            bound = 2.0 * K * 2.0 ** -24 * (np.abs(At) @ np.abs(Bt)) + 1e-30
            err = np.abs(C.cpu().numpy() - ref)
            assert np.all(err <= bound), (ta, tb, float((err / bound).max()))


def test_colsum_vs_numpy(dev):
    import torch

    from jaxfun_amd import _lib
    from jaxfun_amd import kernels as K_

    for rows, cols in ((256, 512), (1, 3), (1000, 130)):
        X = np.random.default_rng(rows).standard_normal((rows, cols)).astype(np.float32)
        out = torch.empty(cols, dtype=torch.float32, device=dev)
        _lib.call("jfx_colsum", K_._p(torch.tensor(X, device=dev)), K_._p(out), rows, cols, K_._s())
        ref = X.astype(np.float64).sum(0)
        assert np.all(np.abs(out.cpu().numpy() - ref) <= rows * 2.0 ** -23 * np.abs(X).sum(0)), (rows, cols)


def test_mlp_gradients_vs_oracle(dev):
    import torch

    from jaxfun_amd.learned import MLP

    net = MLP([48, 64, 64, 40], seed=3)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((32, 48)).astype(np.float32)
    dY = rng.standard_normal((32, 40)).astype(np.float32)
    Y, hs = net.forward(torch.tensor(X, device=dev))
    grads = net.backward(hs, torch.tensor(dY, device=dev))
    params = [(W.double().cpu().numpy(), b.double().cpu().numpy()) for W, b in net.params]
    Yr, hr = O.mlp_forward(params, X.astype(np.float64))
    gr = O.mlp_backward(params, hr, dY.astype(np.float64))
    assert O.rel_l2(Y.cpu().numpy(), Yr) < 2e-2
    for (dW, db), (dWr, dbr) in zip(grads, gr):
        assert O.rel_l2(dW.cpu().numpy(), dWr) < 3e-2
        assert O.rel_l2(db.cpu().numpy(), dbr) < 3e-2


def test_adam_step(dev):
    import torch

    from jaxfun_amd.learned import MLP

    net = MLP([4, 3], seed=0)
    W0 = net.params[0][0].cpu().numpy().copy()
    g = np.full_like(W0, 0.5)
    net.adam([(torch.tensor(g, device=dev), torch.zeros(3, device=dev))], lr=0.1)
    # first Adam step moves every weight by -lr * sign(g)
    assert np.allclose(net.params[0][0].cpu().numpy(), W0 - 0.1, atol=1e-6)


@pytest.fixture(scope="module")
def small_problem(dev):
    import torch

    from jaxfun_amd import workloads as W
    from jaxfun_amd.learned import VarCoeffPoisson2D

    N = 24
    op = VarCoeffPoisson2D(N, gmodes=8)
    G, Fh = zip(*[W.varcoeff_sample(100 + s, N=N, gmodes=8) for s in range(6)])
    ghat = torch.tensor(np.stack(G), device=dev)
    fhat = torch.tensor(np.stack(Fh), device=dev)
    return op, ghat, fhat, G, Fh


def test_operator_and_rhs_vs_oracle(small_problem):
    op, ghat, fhat, G, Fh = small_problem
    vc = O.VarCoeffPoisson(op.N)
    kq = op.kfield(ghat)
    b = op.rhs(fhat)
    U = np.random.default_rng(2).standard_normal((6, op.n, op.n))
    AU = op.apply(kq, _t(U, kq.device)).cpu().numpy()
    for i in range(6):
        kr = vc.kfield(G[i])
        assert O.rel_l2(kq[i].cpu().numpy(), kr) < 1e-13
        assert O.rel_l2(b[i].cpu().numpy(), vc.rhs(Fh[i])) < 1e-13
        assert O.rel_l2(AU[i], vc.apply(kr, U[i])) < 1e-12


def test_batched_pcg_vs_oracle(small_problem):
    from jaxfun_amd import learned as Lm

    op, ghat, fhat, G, Fh = small_problem
    kq = op.kfield(ghat)
    b = op.rhs(fhat)
    x, iters = Lm.pcg(lambda U: op.apply(kq, U), b, op.plain_precond(op.kbar(kq)), rtol=1e-10)
    res = Lm.true_residual(op, kq, x, b).cpu().numpy()
    assert np.all(res < 1e-9)
    vc = O.VarCoeffPoisson(op.N)
    for i in range(6):
        kr = vc.kfield(G[i])
        xd = np.linalg.solve(vc.dense(kr), vc.rhs(Fh[i]).ravel()).reshape(op.n, op.n)
        assert O.rel_l2(x[i].cpu().numpy(), xd) < 1e-8
    assert np.all(iters.cpu().numpy() > 2)


def test_scaled_preconditioner_symmetric_and_gradient(small_problem):
    import torch

    op, ghat, fhat, G, Fh = small_problem
    s = op.kfield(-0.5 * ghat)
    rng = np.random.default_rng(3)
    X = _t(rng.standard_normal((6, op.n, op.n)), s.device)
    Y = _t(rng.standard_normal((6, op.n, op.n)), s.device)
    P = op.scaled_precond(s)
    lhs = (X * P(Y)).sum().item()
    rhs = (P(X) * Y).sum().item()
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    # gradient of <w, P_s b> w.r.t. s against a central difference
    W = _t(rng.standard_normal((6, op.n, op.n)), s.device)
    z, parts = op.scaled_parts(s, X)
    gs = op.scaled_backward(s, parts, W)
    dS = _t(rng.standard_normal(tuple(s.shape)), s.device) * s
    eps = 1e-6
    fp = (W * op.scaled_precond((s + eps * dS).contiguous())(X)).sum().item()
    fm = (W * op.scaled_precond((s - eps * dS).contiguous())(X)).sum().item()
    assert abs((fp - fm) / (2 * eps) - (gs * dS).sum().item()) < 1e-6 * abs(fp)
    del torch


def test_learned_training_reduces_loss_and_pcg_converges(small_problem):
    from jaxfun_amd import learned as Lm

    op, ghat, fhat, G, Fh = small_problem
    kq = op.kfield(ghat)
    b = op.rhs(fhat)
    u, _ = Lm.pcg(lambda U: op.apply(kq, U), b, op.plain_precond(op.kbar(kq)), rtol=1e-10)
    lp = Lm.LearnedPreconditioner(op, hidden=64, layers=4)
    hist = lp.train(ghat, b, u, steps=60, batch=6, lr=3e-3)
    assert hist[-1] < hist[0]
    s = lp.field(ghat)
    P = op.scaled_precond(s)
    x, iters = Lm.pcg(lambda U: op.apply(kq, U), b, P, x0=P(b), rtol=1e-10)
    assert np.all(Lm.true_residual(op, kq, x, b).cpu().numpy() < 1e-9)
    # what the trained preconditioner gains over plain PCG (the method without it)
    _, it_plain = Lm.pcg(lambda U: op.apply(kq, U), b, op.plain_precond(op.kbar(kq)), rtol=1e-10)
    assert iters.double().mean().item() < it_plain.double().mean().item()


def test_pcg_breakdown_marked_or_raised(small_problem):
    import torch

    from jaxfun_amd import learned as Lm

    op, ghat, fhat, G, Fh = small_problem
    kq = op.kfield(ghat)
    b = op.rhs(fhat)
    b[2, 0, 0] = float("nan")
    A = lambda U: op.apply(kq, U)  # noqa: E731
    with pytest.raises(FloatingPointError):
        Lm.pcg(A, b, op.plain_precond(op.kbar(kq)), rtol=1e-10)
    x, it = Lm.pcg(A, b, op.plain_precond(op.kbar(kq)), rtol=1e-10, breakdown="mark")
    it = it.cpu().numpy()
    assert it[2] == -1 and np.all(np.delete(it, 2) > 2)
    del torch


def test_solve_learned_falls_back_to_plain_pcg(small_problem):
    from jaxfun_amd import learned as Lm

    op, ghat, fhat, G, Fh = small_problem
    kq = op.kfield(ghat)
    b = op.rhs(fhat)
    lp = Lm.LearnedPreconditioner(op, hidden=32, layers=3)
    x, it, fb = Lm.solve_learned(op, lp, kq, ghat, b)  # untrained: the classical scaling
    assert not fb.any().item() and np.all(Lm.true_residual(op, kq, x, b).cpu().numpy() < 1e-9)
    lp.net.params[-1][1].fill_(1e4)  # a network whose field overflows exp: every system breaks down
    x, it, fb = Lm.solve_learned(op, lp, kq, ghat, b)
    assert fb.all().item()
    assert np.all(Lm.true_residual(op, kq, x, b).cpu().numpy() < 1e-9)
    _, it_plain = Lm.pcg(lambda U: op.apply(kq, U), b, op.plain_precond(op.kbar(kq)), rtol=1e-10)
    assert np.array_equal(it.cpu().numpy(), it_plain.cpu().numpy())  # broken runs stop at once
