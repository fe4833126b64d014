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
