"""Variable-coefficient Poisson, batched PCG and a learned preconditioner.

Problem (one per batch entry): ``-div(k grad u) = f`` on a Legendre-Dirichlet
tensor space, ``k = exp(g)`` with g given by low-mode Legendre coefficients.
The operator is the non-separable bilinear form ``(grad v, k grad u)``
assembled matrix-free at the quadrature mesh (the reference assembles the
dense tensor with ``assemble_multivar``, galerkin/inner.py:991-1024, and
applies it as ``TensorMatrix``, la/tensormatrix.py:54-56).

Preconditioners (all SPD, so PCG stays valid):
  * plain:   P r = L^-1 r / kbar, L the constant-coefficient stiffness solved by
             fast diagonalisation, kbar the mean of k;
  * scaled:  P_s r = (F s B) L^-1 (F s B)^T r with a positive scaling field s
             on the quadrature mesh (B: coefficients -> mesh values,
             F: mesh values -> coefficients).  s = k^-1/2 is the classical
             choice; the learned preconditioner takes s = exp(sigma) with
             sigma's 16 x 16 low Legendre modes = -g/2 + D o MLP(g / D) (a
             4-layer 512-wide tanh MLP whose last layer starts at zero, so
             training starts from the classical scaling; D_ij =
             (1 + |(i, j)|)^-2 is the prior spectrum of g, which whitens the
             input and keeps the correction's field bounded: an unweighted
             correction can put coherent high modes into sigma whose sum at
             the mesh corners overflows exp), trained on (b, u) pairs that
             plain PCG generates to minimise the energy-norm error of the
             initial guess, (P_s b - u)^T A (P_s b - u) / u^T A u.  Learned PCG
             starts from x0 = P_s b.

The MLP runs on the project's BF16 MFMA GEMM (FP32 accumulation) with its own
backward kernels and Adam; the operator, preconditioners, reductions and PCG
updates are libjfx kernels.  The backward pass through P_s is written out
explicitly below (every step is linear except the two products with s).
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib
from . import kernels as K
from .galerkin import FunctionSpace, Legendre, TensorProduct, TestFunction, TrialFunction, inner
from .galerkin.inner import test_operator
from .operators import Div, Grad
from .kernels import mul_bcast


def _f32(x):
    return K.cast(x.contiguous(), torch.float32)


# ----------------------------------------------------------------------------- MLP
class MLP:
    """tanh MLP with FP32 master weights; GEMMs on BF16 MFMA (FP32 accumulate)."""

    def __init__(self, sizes, seed: int = 0):
        dev = K.device()
        rng = np.random.default_rng(seed)
        self.sizes = list(sizes)
        self.params = []
        for a, b in zip(sizes[:-1], sizes[1:]):
            lim = math.sqrt(6.0 / (a + b))  # Glorot uniform (pinns/module.py:48-49)
            W = torch.tensor(rng.uniform(-lim, lim, size=(a, b)), dtype=torch.float32, device=dev)
            bb = torch.zeros(b, dtype=torch.float32, device=dev)
            self.params.append([W, bb])
        self.m = [[torch.zeros_like(p) for p in layer] for layer in self.params]
        self.v = [[torch.zeros_like(p) for p in layer] for layer in self.params]
        self.step_count = 0

    @staticmethod
    def gemm(A, B, C=None, ta=False, tb=False, alpha=1.0, beta=0.0, bias=None, act=0):
        M = A.shape[1] if ta else A.shape[0]
        Kd = A.shape[0] if ta else A.shape[1]
        N = B.shape[0] if tb else B.shape[1]
        if C is None:
            C = torch.empty((M, N), dtype=torch.float32, device=A.device)
        _lib.call("jfx_gemm_bf16", int(ta), int(tb), M, N, Kd, float(alpha), K._p(A), A.shape[1], K._p(B),
                  B.shape[1], float(beta), K._p(C), C.shape[1], K._p(bias), act, K._s())
        return C

    def forward(self, X):
        hs = [X.contiguous()]
        h = hs[0]
        L = len(self.params)
        for li, (W, b) in enumerate(self.params):
            h = self.gemm(h, W, bias=b, act=0 if li == L - 1 else 1)
            hs.append(h)
        return h, hs

    def backward(self, hs, dY):
        grads = [None] * len(self.params)
        g = dY.contiguous()
        L = len(self.params)
        for li in range(L - 1, -1, -1):
            W, b = self.params[li]
            if li != L - 1:
                dz = torch.empty_like(g)
                _lib.call("jfx_tanh_backward", K._p(g), K._p(hs[li + 1]), K._p(dz), g.numel(), K._s())
                g = dz
            dW = self.gemm(hs[li], g, ta=True)
            db = torch.empty(W.shape[1], dtype=torch.float32, device=W.device)
            _lib.call("jfx_colsum", K._p(g), K._p(db), g.shape[0], g.shape[1], K._s())
            grads[li] = (dW, db)
            if li > 0:
                g = self.gemm(g, W, tb=True)
        return grads

    def adam(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        for (W, b), (dW, db), (mW, mb), (vW, vb) in zip(self.params, grads, self.m, self.v):
            for p, g, m, v in ((W, dW, mW, vW), (b, db, mb, vb)):
                _lib.call("jfx_adam", K._p(p), K._p(g), K._p(m), K._p(v), p.numel(), lr, b1, b2, eps,
                          self.step_count, K._s())

    def n_params(self) -> int:
        return sum(W.numel() + b.numel() for W, b in self.params)


# ----------------------------------------------------------------------------- operator
class VarCoeffPoisson2D:
    """Batched ``(grad v, k grad u)`` on Legendre-Dirichlet (N x N) modes."""

    def __init__(self, N: int = 64, gmodes: int = 16):
        self.N, self.n, self.Q, self.gm = N, N - 2, N, gmodes
        D = FunctionSpace(N, Legendre, {"left": {"D": 0}, "right": {"D": 0}})
        self.T = TensorProduct(D, D)
        S = self.T[0]
        self.B = S.backward_operator(N, 0)
        self.Dx = S.backward_operator(N, 1)
        self.TB = test_operator(S, 0, N)
        self.TD = test_operator(S, 1, N)
        self.F = S.forward_operator(N)
        self.FT = K.AxisOperator(self.F.A64.t().contiguous())
        self.BT = K.AxisOperator(self.B.A64.t().contiguous())
        x, w = S.quad_points_and_weights(N)
        self.w2 = (w[:, None] * w[None, :]).contiguous()
        self.Vg = K.AxisOperator(K.vandermonde("legendre", x, gmodes))          # (Q, gm)
        self.VgT = K.AxisOperator(self.Vg.A64.t().contiguous())                # (gm, Q)
        h = 2.0 / (2 * np.arange(N) + 1.0)
        SH = np.zeros((N - 2, N))
        SH[np.arange(N - 2), np.arange(N - 2)] = h[:N - 2]
        SH[np.arange(N - 2), np.arange(2, N)] = -h[2:]
        self.SH = K.AxisOperator(torch.tensor(SH, device=K.device()))         # S diag(h)
        v, u = TestFunction(self.T), TrialFunction(self.T)
        self.L = inner(Div(Grad(u)) * v, sparse=True).scale(-1.0).fd_factor()  # (grad v, grad u)

    def _ap(self, op, X, axis):
        return K.apply_along(op, X.contiguous(), axis, "f64")

    def kfield(self, ghat: torch.Tensor) -> torch.Tensor:
        """k = exp(g) at the (Q, Q) quadrature mesh for ghat (batch, gm, gm)."""
        g = self._ap(self.Vg, self._ap(self.Vg, ghat, 1), 2)
        k = torch.empty_like(g)
        _lib.call("jfx_exp", K._p(g), K._p(k), 1.0, g.numel(), K._s())
        return k

    def rhs(self, fhat: torch.Tensor) -> torch.Tensor:
        """b = (S (x) S)(diag(h) fhat diag(h)) for orthogonal Legendre forcing coefficients."""
        return self._ap(self.SH, self._ap(self.SH, fhat, 1), 2)

    def apply(self, kq: torch.Tensor, U: torch.Tensor) -> torch.Tensor:
        ux = self._ap(self.Dx, self._ap(self.B, U, 2), 1)
        uy = self._ap(self.B, self._ap(self.Dx, U, 2), 1)
        a = self._ap(self.TB, self._ap(self.TD, mul_bcast(ux, kq), 1), 2)
        b = self._ap(self.TD, self._ap(self.TB, mul_bcast(uy, kq), 1), 2)
        return K.lincomb([(1.0, a), (1.0, b)])

    def kbar(self, kq: torch.Tensor) -> torch.Tensor:
        out = torch.empty(kq.shape[0], dtype=torch.float64, device=kq.device)
        ones = self.w2.expand_as(kq).contiguous()
        _lib.call("jfx_batched_dot", K._p(kq), K._p(ones), K._p(out), kq.shape[0], self.Q * self.Q, K._s())
        return out / 4.0

    # ------------------------------------------------------------ preconditioners
    def plain_precond(self, kbar: torch.Tensor):
        inv = (1.0 / kbar).contiguous()

        def P(R):
            Z = self.L.solve(R)
            out = torch.empty_like(Z)
            _lib.call("jfx_scale_batch", K._p(Z), K._p(inv), K._p(out), Z.shape[0], Z[0].numel(), K._s())
            return out
        return P

    def scaled_parts(self, s: torch.Tensor, R: torch.Tensor):
        """Forward pass of P_s R, returning intermediates for the backward pass."""
        FtR = self._ap(self.FT, self._ap(self.FT, R, 1), 2)     # (b, Q, Q)
        a = mul_bcast(FtR, s)
        c = self._ap(self.BT, self._ap(self.BT, a, 1), 2)       # (b, n, n)
        d = self.L.solve(c)
        e = self._ap(self.B, self._ap(self.B, d, 1), 2)         # (b, Q, Q)
        y = mul_bcast(e, s)
        z = self._ap(self.F, self._ap(self.F, y, 1), 2)
        return z, (FtR, e)

    def scaled_precond(self, s: torch.Tensor):
        def P(R):
            return self.scaled_parts(s, R)[0]
        return P

    def scaled_backward(self, s, parts, gz):
        """d loss / d s given d loss / d z (P_s's two s-products; L symmetric)."""
        FtR, e = parts
        gy = self._ap(self.FT, self._ap(self.FT, gz, 1), 2)
        gs = mul_bcast(gy, e)
        ge = mul_bcast(gy, s)
        gd = self._ap(self.BT, self._ap(self.BT, ge, 1), 2)
        gc = self.L.solve(gd)
        ga = self._ap(self.B, self._ap(self.B, gc, 1), 2)
        return K.lincomb([(1.0, gs), (1.0, mul_bcast(ga, FtR))])

    def sigma_field(self, sig_hat: torch.Tensor) -> torch.Tensor:
        """s = exp(sigma) on the mesh from (batch, gm, gm) low-mode coefficients."""
        return self.kfield(sig_hat)


# ----------------------------------------------------------------------------- PCG
def batched_dot(x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    out = torch.empty(x.shape[0], dtype=torch.float64, device=x.device)
    _lib.call("jfx_batched_dot", K._p(x), K._p(y), K._p(out), x.shape[0], x[0].numel(), K._s())
    return out


def pcg(apply, b: torch.Tensor, precond, x0: torch.Tensor | None = None, rtol: float = 1e-10,
        maxiter: int = 2000, check_every: int = 1, snapshots=(), snap_out: dict | None = None,
        breakdown: str = "raise"):
    """Batched preconditioned CG; every system stops on ||r|| <= rtol ||b||
    (recursive residual).  Returns (x, iterations per system).  For each k in
    ``snapshots`` the iterate and residual after k iterations are stored in
    ``snap_out[k] = (x_k, r_k)`` (systems that stopped earlier keep their
    final values).  A system whose residual turns non-finite stops with
    iterations -1; ``breakdown="raise"`` then raises FloatingPointError,
    ``"mark"`` returns the marks for the caller's fallback."""
    if breakdown not in ("raise", "mark"):
        raise ValueError("breakdown must be 'raise' or 'mark'")
    B = b.shape[0]
    n = b[0].numel()
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    r = K.lincomb([(1.0, b)]) if x0 is None else K.lincomb([(1.0, b), (-1.0, apply(x))])
    bb = batched_dot(b, b)
    active = torch.ones(B, dtype=torch.int32, device=b.device)
    iters = torch.zeros(B, dtype=torch.int32, device=b.device)
    rr = batched_dot(r, r)
    _lib.call("jfx_pcg_check", K._p(rr), K._p(bb), rtol * rtol, K._p(active), K._p(iters), B, 0, K._s())
    z = precond(r)
    p = K.lincomb([(1.0, z)])
    rz = batched_dot(r, z)
    for it in range(maxiter):
        q = apply(p)
        pq = batched_dot(p, q)
        _lib.call("jfx_pcg_update_xr", K._p(x), K._p(r), K._p(p), K._p(q), K._p(rz), K._p(pq), K._p(active), B,
                  n, K._s())
        rr = batched_dot(r, r)
        _lib.call("jfx_pcg_check", K._p(rr), K._p(bb), rtol * rtol, K._p(active), K._p(iters), B, 1, K._s())
        if snap_out is not None and it + 1 in snapshots:
            snap_out[it + 1] = (x.clone(), r.clone())
        if (it + 1) % check_every == 0 and int(active.sum().item()) == 0:
            break
        z = precond(r)
        rz_new = batched_dot(r, z)
        _lib.call("jfx_pcg_update_p", K._p(p), K._p(z), K._p(rz_new), K._p(rz), K._p(active), B, n, K._s())
        rz = rz_new
    if breakdown == "raise":
        nbad = int((iters < 0).sum().item())
        if nbad:
            raise FloatingPointError(f"PCG breakdown: non-finite residual in {nbad} of {B} systems")
    return x, iters


def true_residual(op: VarCoeffPoisson2D, kq, x, b) -> torch.Tensor:
    r = K.lincomb([(1.0, b), (-1.0, op.apply(kq, x))])
    return torch.sqrt(batched_dot(r, r) / batched_dot(b, b))


def solve_learned(op: VarCoeffPoisson2D, lp: "LearnedPreconditioner", kq: torch.Tensor, ghat: torch.Tensor,
                  b: torch.Tensor, rtol: float = 1e-10, accept: float | None = None, maxiter: int = 2000):
    """Learned PCG (x0 = P_s b, preconditioner P_s with the network's field)
    with plain PCG as the fallback and the true Galerkin residual as the
    judge: systems whose learned run broke down or whose true residual
    ||b - A x|| / ||b|| exceeds ``accept`` (default 10 rtol) are solved again
    by plain PCG from zero.  Returns (x, iterations, fallback mask); the
    iterations of a fallback system count both runs."""
    accept = 10.0 * rtol if accept is None else float(accept)
    P = op.scaled_precond(lp.field(ghat))
    x, it = pcg(lambda U: op.apply(kq, U), b, P, x0=P(b), rtol=rtol, maxiter=maxiter, breakdown="mark")
    res = true_residual(op, kq, x, b)
    bad = (it < 0) | ~(res <= accept)
    if bool(bad.any().item()):
        idx = torch.nonzero(bad).flatten()
        kf, bf = kq[idx].contiguous(), b[idx].contiguous()
        xf, itf = pcg(lambda U: op.apply(kf, U), bf, op.plain_precond(op.kbar(kf)), rtol=rtol, maxiter=maxiter)
        x[idx] = xf
        it[idx] = it[idx].clamp(min=0) + itf
    return x, it, bad


# ----------------------------------------------------------------------------- learned preconditioner
class LearnedPreconditioner:
    """MLP (ghat -> correction to the low modes of log s) driving the scaled FD
    preconditioner.  ``prior`` sets the network-free part, sigma_hat = prior *
    ghat (-1/2: s = k^-1/2; 0: the network alone); ``loss`` is ``"energy"``
    (A-norm of the initial-guess error) or ``"l2"``."""

    def __init__(self, op: VarCoeffPoisson2D, hidden: int = 512, layers: int = 4, seed: int = 0,
                 prior: float = -0.5, loss: str = "energy"):
        if loss not in ("energy", "l2"):
            raise ValueError("loss must be 'energy' or 'l2'")
        gm2 = op.gm * op.gm
        self.op = op
        self.prior = float(prior)
        self.loss = loss
        self.net = MLP([gm2] + [hidden] * (layers - 1) + [gm2], seed=seed)
        self.net.params[-1][0].zero_()  # the untrained network adds nothing to the prior
        m = np.arange(op.gm, dtype=np.float64)
        D = (1.0 + np.sqrt(m[:, None] ** 2 + m[None, :] ** 2)) ** -2.0  # prior spectrum of ghat
        self.D = torch.tensor(D, device=K.device())
        self.Dinv = torch.tensor(1.0 / D, device=K.device())
        self.skipped = 0  # training steps dropped for a non-finite loss

    def sigma_hat(self, ghat: torch.Tensor):
        X = _f32(mul_bcast(ghat.contiguous(), self.Dinv).reshape(ghat.shape[0], -1))
        Y, hs = self.net.forward(X)
        corr = mul_bcast(K.cast(Y, torch.float64).reshape(ghat.shape), self.D)
        sh = K.lincomb([(1.0, corr), (self.prior, ghat.contiguous())])
        return sh, hs

    def field(self, ghat):
        sh, _ = self.sigma_hat(ghat)
        return self.op.sigma_field(sh)

    def loss_and_grad(self, ghat, b, u):
        """Mean of ||P_s b - u||^2 / ||u||^2 over the batch (in the A-norm for
        ``loss="energy"``) and its MLP gradients."""
        op = self.op
        sh, hs = self.sigma_hat(ghat)
        s = op.sigma_field(sh)
        z, parts = op.scaled_parts(s, b)
        diff = K.lincomb([(1.0, z), (-1.0, u)])
        if self.loss == "energy":
            kq = op.kfield(ghat)
            Md, Mu = op.apply(kq, diff), op.apply(kq, u)  # A symmetric: d(d^T A d)/dz = 2 A d
        else:
            Md, Mu = diff, u
        uu = batched_dot(u, Mu)
        per = batched_dot(diff, Md) / uu
        Bn = b.shape[0]
        w = (2.0 / (uu * Bn)).contiguous()
        gz = torch.empty_like(diff)
        _lib.call("jfx_scale_batch", K._p(Md), K._p(w), K._p(gz), Bn, diff[0].numel(), K._s())
        gs = op.scaled_backward(s, parts, gz)
        gsig = mul_bcast(gs, s)
        gsh = op._ap(op.VgT, op._ap(op.VgT, gsig, 1), 2)
        grads = self.net.backward(hs, _f32(mul_bcast(gsh, self.D).reshape(Bn, -1)))
        return float(per.mean().item()), grads

    def train(self, ghat, b, u, steps: int = 1000, batch: int = 256, lr: float = 1e-3, seed: int = 0,
              log=None):
        rng = np.random.default_rng(seed)
        N = ghat.shape[0]
        hist = []
        for st in range(steps):
            idx = torch.tensor(rng.choice(N, size=min(batch, N), replace=False), device=ghat.device)
            loss, grads = self.loss_and_grad(ghat[idx].contiguous(), b[idx].contiguous(), u[idx].contiguous())
            if not math.isfinite(loss):  # a batch whose loss overflowed leaves the weights alone
                self.skipped += 1
                continue
            self.net.adam(grads, lr=lr * (0.5 * (1 + math.cos(math.pi * st / steps))) + 1e-5)
            hist.append(loss)
            if log and (st % max(1, steps // 10) == 0 or st == steps - 1):
                log(f"step {st} loss {loss:.4e}")
        if not hist:
            raise FloatingPointError("training produced no finite loss")
        return hist
