"""Benchmark: ragged batch of 4096 Legendre-Dirichlet Helmholtz solves on one MI355X.

Headline (``value``): solves/s of the mixed-precision batched path (2-term
fp16 split MFMA of power-of-two-scaled blocks, FP32 accumulation and I/O,
FP64 tridiagonal solves), one persistent HIP-graph-replayed launch per step,
at <= 1e-6 relative L2 error against the plain FP64 GPU path (both measured
here).  Extra fields report the 3-term bf16 split and FP64 paths, the 3D
Legendre transform throughput per precision (16 x 128^3 round trip), the
Chebyshev-Dirichlet Poisson solve (N = 64) and the other workloads.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extras] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFS = 2500.0
FP64_PEAK_TFS = 78.6


def _events():
    import torch

    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def bench_helmholtz(steps: int, warmup: int):
    import torch

    from jaxfun_amd import workloads as W
    from jaxfun_amd.helmholtz import HelmholtzBatch, autotune

    Ns, alphas, loads = W.helmholtz_ragged_batch()
    res = {}
    outs = {}
    for mode in ("mixed16", "mixed", "fp64"):
        hb = HelmholtzBatch(Ns, alphas, mode=mode)
        F = hb.pack(loads)
        U = torch.empty_like(F)
        autotune(hb, F)
        hb.capture(F, U)
        for _ in range(warmup):
            hb.replay()
        torch.cuda.synchronize()
        per = []
        t0 = time.perf_counter()
        for _ in range(steps):
            e0, e1 = _events()
            e0.record()
            hb.replay()
            e1.record()
            per.append((e0, e1))
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        kms = float(np.mean([a.elapsed_time(b) for a, b in per]))
        res[mode] = dict(ms=wall * 1e3, kernel_ms=kms, solves_per_s=len(Ns) / wall,
                         io_bytes=hb.algorithmic_bytes(), alg_flops=hb.algorithmic_flops(),
                         mfma_flops=hb.mfma_flops(), nitems=hb.nitems, nblocks=hb.nblocks,
                         nwaves=hb.nwaves, queues=hb.nqueues)
        outs[mode] = U.to(torch.float64)
    for mode in ("mixed16", "mixed"):
        res[mode]["rel_err_vs_fp64_gpu"] = float(torch.linalg.vector_norm(outs[mode] - outs["fp64"]) /
                                                 torch.linalg.vector_norm(outs["fp64"]))
    return res, (Ns, alphas, loads)


PRECISIONS_3D = ("f64", "f32", "bf16x6", "ozaki4", "i8x5", "i8x6")


def bench_transform_3d(reps: int = 3):
    import torch

    from jaxfun_amd import Legendre, TensorProduct
    from jaxfun_amd import workloads as W

    N, B = 128, 16
    c = torch.tensor(W.legendre_3d_coefficients(N, B), device="cuda")
    L = Legendre(N)
    T = TensorProduct(L, L, L)
    u64 = T.backward(c)
    fx = np.load(os.path.join(ROOT, "tests", "golden", "c1_legendre_3d_samples.npz"))
    idx = torch.tensor(fx["idx"].astype(np.int64), device="cuda")
    uref = fx["u"]
    out = {}
    for prec in PRECISIONS_3D:
        u = T.backward(c, precision=prec)
        ch = T.forward(u, precision=prec)
        torch.cuda.synchronize()
        e0, e1 = _events()
        e0.record()
        for _ in range(reps):
            u = T.backward(c, precision=prec)
            ch = T.forward(u, precision=prec)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        elt = u.element_size()
        alg_bytes = 4.0 * N**3 * B * elt
        eb = float(torch.linalg.vector_norm(u.to(torch.float64) - u64) / torch.linalg.vector_norm(u64))
        us = u[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]].to(torch.float64).cpu().numpy()
        eo = float(np.linalg.norm(us - uref) / np.linalg.norm(uref))
        # the exact round trip is the identity: its error is measured against the input itself
        er = float(torch.linalg.vector_norm(ch.to(torch.float64) - c) / torch.linalg.vector_norm(c))
        out[prec] = dict(ms=ms, GBps=alg_bytes / ms / 1e6, TFLOPs=12.0 * N**4 * B / ms / 1e9, io_bytes_per_elt=elt,
                         rel_err_backward_vs_fp64_gpu=eb, rel_err_backward_vs_oracle_fixture=eo,
                         rel_err_roundtrip=er, **{"meets_1e-6": max(eo, er) <= 1e-6})
    ok = {p: r for p, r in out.items() if r["meets_1e-6"]}
    if ok:
        best = max(ok, key=lambda p: ok[p]["GBps"])
        out["best_at_1e-6"] = dict(precision=best, GBps=ok[best]["GBps"], ms=ok[best]["ms"],
                                   speedup_vs_f64=out["f64"]["ms"] / ok[best]["ms"])
    return out


def bench_poisson_chebyshev(reps: int = 50):
    import torch

    from jaxfun_amd import Chebyshev, Div, FunctionSpace, Grad, TensorProduct, TestFunction, TrialFunction, inner
    from jaxfun_amd import workloads as W

    N = 64
    D = FunctionSpace(N, Chebyshev, {"left": {"D": 0}, "right": {"D": 0}})
    T = TensorProduct(D, D)
    v, u = TestFunction(T), TrialFunction(T)
    A = inner(v * Div(Grad(u)), sparse=True)
    b = torch.tensor(-W.chebyshev_poisson_load(N), device="cuda")
    fd = A.fd_factor()

    def step(rhs):
        return T.backward(fd.solve(rhs))

    uj = step(b)
    fx = np.load(os.path.join(ROOT, "tests", "golden", "c0_chebyshev_poisson_N64.npz"))
    err_U = float(np.linalg.norm(fd.solve(b).cpu().numpy() - fx["U"]) / np.linalg.norm(fx["U"]))
    err_u = float(np.linalg.norm(uj.cpu().numpy() - fx["u"]) / np.linalg.norm(fx["u"]))
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step(b)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        out = step(b)
    g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.replay()
    torch.cuda.synchronize()
    lat = (time.perf_counter() - t0) / reps
    same = bool(torch.equal(out, uj))
    # batched right-hand sides: 4096 independent solves with the same operator
    bb = b.unsqueeze(0).repeat(4096, 1, 1).contiguous()
    step(bb)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        step(bb)
    torch.cuda.synchronize()
    bt = (time.perf_counter() - t0) / 5
    # the exact Kronecker solve of the same system (the reference's dense path,
    # la/tpmatrix.py:1052-1086): device assembly, LU with partial pivoting
    from jaxfun_amd.la.tpmatrix import TPMatricesDenseLUFactors, kron_device

    shape = tuple(int(m.shape[0]) for m in A.tpmats[0].mats)
    Ak = kron_device(A.tpmats)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dense = TPMatricesDenseLUFactors(Ak, shape)
    torch.cuda.synchronize()
    t_factor = time.perf_counter() - t0
    Ud = dense.solve(b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        Ud = dense.solve(b)
    torch.cuda.synchronize()
    t_solve = (time.perf_counter() - t0) / 5
    err_dense = float(np.linalg.norm(Ud.cpu().numpy() - fx["U"]) / np.linalg.norm(fx["U"]))
    return dict(single_solve_us=lat * 1e6, single_solves_per_s=1.0 / lat, graph_matches_eager=same,
                batched_4096_solves_per_s=4096 / bt,
                rel_err_coeffs_vs_oracle_kron_solve=err_U, rel_err_mesh_values_vs_oracle=err_u,
                dense_kronecker_lu={"n": int(Ak.shape[0]), "factor_ms": t_factor * 1e3, "solve_ms": t_solve * 1e3,
                                    "rel_err_vs_oracle_kron_solve": err_dense})


def bench_lifted_helmholtz(reps: int = 50, batch: int = 4096):
    """Inhomogeneous Dirichlet Helmholtz on (0, 1)^2 (Legendre M = 50, scaling
    n + 1, u = exp(-(x^2+y^2)) with its boundary values lifted), as in the
    reference's examples/helmholtz2D.py: assembly, one solve + evaluation on
    the quadrature mesh (graph replay), and a batch of right-hand sides with
    and without the lift added to the output."""
    import sympy as sp
    import torch

    from jaxfun_amd import Div, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, TrialFunction, inner, n
    from jaxfun_amd.coordinates import x, y

    M = 50
    ue = sp.exp(-(x**2 + y**2))
    bcsx = {"left": {"D": ue.subs(x, 0)}, "right": {"D": ue.subs(x, 1)}}
    bcsy = {"left": {"D": ue.subs(y, 0)}, "right": {"D": ue.subs(y, 1)}}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Dx = FunctionSpace(M, Legendre, bcs=bcsx, scaling=n + 1, domain=(0, 1))
    Dy = FunctionSpace(M, Legendre, bcs=bcsy, scaling=n + 1, domain=(0, 1))
    T = TensorProduct(Dx, Dy)
    v, u = TestFunction(T), TrialFunction(T)
    A, b = inner(v * (Div(Grad(u)) + u) - v * (Div(Grad(ue)) + ue), sparse=True, kind="system")
    fd = A.fd_factor()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0

    def step(rhs):
        return T.backward(fd.solve(rhs))

    uj = step(b)
    X, Y = [m.cpu().numpy() for m in T.mesh()]
    err = float(np.abs(uj.cpu().numpy() - np.exp(-(X**2 + Y**2))).max())
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step(b)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        out = step(b)
    g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.replay()
    torch.cuda.synchronize()
    lat = (time.perf_counter() - t0) / reps
    same = bool(torch.equal(out, uj))
    H = T.get_homogeneous()
    bb = b.unsqueeze(0).repeat(batch, 1, 1).contiguous()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / 5

    t_lift = timed(lambda: T.backward(fd.solve(bb)))
    t_hom = timed(lambda: H.backward(fd.solve(bb)))
    return dict(setup_ms=t_setup * 1e3, max_abs_err_vs_exact=err, single_solve_us=lat * 1e6,
                single_solves_per_s=1.0 / lat, graph_matches_eager=same,
                **{f"batched_{batch}_solves_per_s": batch / t_lift,
                   f"batched_{batch}_homogeneous_solves_per_s": batch / t_hom},
                lift_overhead_frac=t_lift / t_hom - 1.0)


def bench_periodic_poisson(reps: int = 50, batch: int = 1024):
    """Fourier (N = 20) x Legendre-Dirichlet (M = 80) Poisson of the reference's
    examples/poisson2D_periodic.py: the per-wavenumber banded solve plus the
    backward transform to the doubled mesh, one solve (graph replay) and a
    batch of right-hand sides."""
    import sympy as sp
    import torch

    from jaxfun_amd import Div, Fourier, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, TrialFunction, \
        inner, n
    from jaxfun_amd.coordinates import x, y

    ue = (1 - y**2) * sp.cos(2 * x) * sp.exp(sp.cos(sp.pi * y))
    M, N = 80, 20
    D = FunctionSpace(M, Legendre, {"left": {"D": ue.subs(y, -1)}, "right": {"D": ue.subs(y, 1)}}, scaling=n + 1)
    F = FunctionSpace(N, Fourier)
    T = TensorProduct(F, D)
    v, u = TestFunction(T), TrialFunction(T)
    A, b = inner(v * Div(Grad(u)) - v * Div(Grad(ue)), sparse=True, kind="system")
    wn = A.lu_factor()

    def step(rhs):
        return T.backward(wn.solve(rhs), N=(2 * N, 2 * M))

    uj = step(b)
    xj = [m.cpu().numpy() for m in T.mesh(N=(2 * N, 2 * M))]
    uej = sp.lambdify((x, y), ue)(*xj)
    err = float(np.linalg.norm(uj.cpu().numpy() - uej) / 100)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step(b)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        out = step(b)
    g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.replay()
    torch.cuda.synchronize()
    lat = (time.perf_counter() - t0) / reps
    same = bool(torch.equal(out, uj))
    bb = b.unsqueeze(0).repeat(batch, 1, 1).contiguous()
    step(bb)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        step(bb)
    torch.cuda.synchronize()
    bt = (time.perf_counter() - t0) / 5
    return dict(solver=type(wn).__name__, error_example_metric=err, single_solve_us=lat * 1e6,
                single_solves_per_s=1.0 / lat, graph_matches_eager=same,
                **{f"batched_{batch}_solves_per_s": batch / bt})


def bench_band_solve(reps: int = 20):
    """DiaMatrix.solve on few long lines (la/diamatrix.py:592-770, LUFactors.solve
    :1880-1925): the Legendre-Dirichlet mass matrix (offsets 0, +-2) of size n,
    factored once, solved for `lines` right-hand sides with each line's rows
    over several workgroups that hand the recurrence window on (`grid`, what
    `auto` takes for lines of 8192 rows and more), one workgroup per line
    (`wave`: eight wavefronts for long lines, the rows split over the lanes,
    chunk maps composed by a wave scan) and one lane per line (`lanes`: a
    sequential sweep).  Times are HIP-event means over back-to-back launches;
    the errors are relative L2 differences from the per-lane sweep."""
    import torch

    from jaxfun_amd import kernels as K

    out = {}
    gen = torch.Generator(device="cuda").manual_seed(5)
    for n, lines in ((1024, 1), (20000, 1), (20000, 16), (100000, 1), (100000, 16), (1000000, 1)):
        k = np.arange(n, dtype=np.float64)
        band = np.zeros((n, 5))
        band[:, 2] = 2.0 / (2 * k + 1) + 2.0 / (2 * k + 5)
        band[2:, 0] = -2.0 / (2 * k[:-2] + 5)
        band[:-2, 4] = -2.0 / (2 * k[:-2] + 5)
        lu = K.band_lu_factor(torch.tensor(band, device="cuda"), 2, 2)
        b = torch.randn(n, lines, dtype=torch.float64, device="cuda", generator=gen)
        r = {}
        sol = {}
        methods = ("grid", "wave", "lanes") if n >= 8192 else ("wave", "lanes")
        if n > 100000:
            methods = ("grid", "wave")  # one lane per line: ~0.4 s per solve
        for m in methods:
            K.band_lu_solve(lu, 2, 2, b.clone(), 0, method=m)  # warm-up
            nrep = reps if (m != "lanes" or n < 20000) else max(3, reps // 5)  # the slow path: fewer launches
            Xs = [b.clone() for _ in range(nrep)]  # in-place solves: one copy per launch
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for X in Xs:
                K.band_lu_solve(lu, 2, 2, X, 0, method=m)
            e1.record()
            torch.cuda.synchronize()
            r[f"{m}_us"] = e0.elapsed_time(e1) / nrep * 1e3
            sol[m] = Xs[0]
        ref = "lanes" if "lanes" in sol else "wave"
        if "lanes" in sol:
            r["speedup_wave_vs_lanes"] = r["lanes_us"] / r["wave_us"]
            r["rel_diff_wave_vs_lanes"] = float(torch.linalg.norm(sol["wave"] - sol["lanes"]) /
                                                torch.linalg.norm(sol["lanes"]))
        if "grid" in sol:
            r["speedup_grid_vs_wave"] = r["wave_us"] / r["grid_us"]
            r[f"rel_diff_grid_vs_{ref}"] = float(torch.linalg.norm(sol["grid"] - sol[ref]) /
                                                 torch.linalg.norm(sol[ref]))
        out[f"n{n}_lines{lines}"] = r
    return out


def bench_burgers(steps: int = 1000, batch: int = 256, N: int = 128):
    """IMEX ARS443 Burgers, Chebyshev-Dirichlet N x N, 3/2 padding,
    mixed precision vs FP64, drift after `steps` steps."""
    import sympy as sp
    import torch

    from jaxfun_amd import Chebyshev, Div, FunctionSpace, Grad, TensorProduct, TestFunction, TrialFunction
    from jaxfun_amd import workloads as W
    from jaxfun_amd.galerkin.arguments import t
    from jaxfun_amd.integrators import ARS443, IMEXRungeKutta

    D = FunctionSpace(N, Chebyshev, {"left": {"D": 0}, "right": {"D": 0}})
    T = TensorProduct(D, D)
    v = TestFunction(T)
    u = TrialFunction(T, transient=True)
    x, y = T.system.base_scalars()
    nu, dt = 1e-2, 1e-3
    weak = v * (u.diff(t) + u * (u.diff(x) + u.diff(y)) - nu * Div(Grad(u)))
    u0f, _, _ = W.burgers_initial_conditions(N, batch)
    X, Y = [m.cpu().numpy() for m in T.mesh()]
    U0 = T.forward(torch.tensor(np.stack([u0f(b, X, Y) for b in range(batch)]), device="cuda"))
    out = {}
    res = {}
    for prec in ("f64", "mixed"):
        integ = IMEXRungeKutta(weak, tableau=ARS443, initial=sp.S.Zero, precision=prec)
        U = U0 if prec == "f64" else U0.to(torch.float32)
        U = integ.step(U, dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps - 1):
            U = integ.step(U, dt)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        res[prec] = U.to(torch.float64)
        out[prec] = dict(s_per_step=el / (steps - 1), ic_steps_per_s=batch * (steps - 1) / el,
                         finite=bool(torch.isfinite(U).all()))
    out["drift_rel_l2_mixed_vs_f64"] = float(torch.linalg.vector_norm(res["mixed"] - res["f64"]) /
                                             torch.linalg.vector_norm(res["f64"]))
    out["config"] = dict(N=N, batch=batch, steps=steps, dt=dt, nu=nu, padding="3/2")
    return out


def bench_learned_pcg(ntrain: int = 20000, neval: int = 1024, chunk: int = 1024, train_steps: int = 1500):
    """Variable-coefficient Poisson 64^2: plain PCG vs learned PCG."""
    import torch

    from jaxfun_amd import kernels as Kn
    from jaxfun_amd import learned as Lm
    from jaxfun_amd import workloads as W

    op = Lm.VarCoeffPoisson2D(64, 16)

    def gen(seeds):
        G, Fh = zip(*[W.varcoeff_sample(int(s)) for s in seeds])
        return torch.tensor(np.stack(G), device="cuda"), torch.tensor(np.stack(Fh), device="cuda")

    # training pairs from plain PCG: (b, u) for every system, plus the Krylov
    # pairs (r_k, u - x_k) after k = 2 and 5 iterations (A (u - x_k) = r_k)
    snaps = (2, 5)
    t0 = time.perf_counter()
    gh, bs, us = [], [], []
    kry = {k: ([], []) for k in snaps}
    for s0 in range(10000, 10000 + ntrain, chunk):
        g, fh = gen(range(s0, min(s0 + chunk, 10000 + ntrain)))
        kq = op.kfield(g)
        b = op.rhs(fh)
        so = {}
        x, _ = Lm.pcg(lambda U: op.apply(kq, U), b, op.plain_precond(op.kbar(kq)), rtol=1e-10,
                      snapshots=snaps, snap_out=so)
        gh.append(g)
        bs.append(b)
        us.append(x)
        for k in snaps:
            kry[k][0].append(so[k][1])
            kry[k][1].append(Kn.lincomb([(1.0, x), (-1.0, so[k][0])]))
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    nsys = sum(t_.shape[0] for t_ in gh)
    gh = torch.cat(gh * (1 + len(snaps)))
    bs = torch.cat(bs + [r_ for k in snaps for r_ in kry[k][0]])
    us = torch.cat(us + [e_ for k in snaps for e_ in kry[k][1]])
    lp = Lm.LearnedPreconditioner(op, hidden=512, layers=4)
    t0 = time.perf_counter()
    hist = lp.train(gh, bs, us, steps=train_steps, batch=256, lr=1e-3)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    g, fh = gen(range(40000, 40000 + neval))
    kq = op.kfield(g)
    b = op.rhs(fh)
    A = lambda U: op.apply(kq, U)  # noqa: E731
    xp, itp = Lm.pcg(A, b, op.plain_precond(op.kbar(kq)), rtol=1e-10)
    s_an = op.kfield(-0.5 * g)
    xa, ita = Lm.pcg(A, b, op.scaled_precond(s_an), x0=op.scaled_precond(s_an)(b), rtol=1e-10)
    xl, itl, fb = Lm.solve_learned(op, lp, kq, g, b, rtol=1e-10)
    _, itl0 = Lm.pcg(A, b, op.scaled_precond(lp.field(g)), rtol=1e-10)
    res = {k: float(Lm.true_residual(op, kq, x_, b).max().item()) for k, x_ in
           (("plain", xp), ("analytic_sqrt_k", xa), ("learned", xl))}
    itp, ita, itl, itl0 = (t_.double().mean().item() for t_ in (itp, ita, itl, itl0))
    return dict(train_systems=nsys, train_pairs=int(gh.shape[0]), eval_systems=neval,
                mlp_params=lp.net.n_params(), loss="energy norm of P_s b - u, sigma = -g/2 + D o MLP(g / D)",
                gen_s=gen_s, train_s=train_s, train_loss_first=hist[0], train_loss_last=hist[-1],
                train_steps_skipped=lp.skipped, plain_fallbacks=int(fb.sum().item()),
                mean_iters_plain=itp, mean_iters_learned=itl, mean_iters_learned_zero_x0=itl0,
                mean_iters_scaled_sqrt_k_no_network=ita,
                iteration_ratio_plain_over_learned=itp / itl,
                iteration_ratio_no_network_over_learned=ita / itl, max_true_residual=res)


def _host_info():
    """CPU model, core count, BLAS vendor/threads and library versions of this host."""
    import platform

    import numpy
    import scipy

    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    blas = []
    try:
        import threadpoolctl

        blas = [dict(api=d.get("internal_api"), version=d.get("version"), threads=d.get("num_threads"))
                for d in threadpoolctl.threadpool_info() if d.get("user_api") == "blas"]
    except Exception:  # threadpoolctl is optional
        pass
    return dict(cpu_model=model, nproc=os.cpu_count(), sched_cores=len(os.sched_getaffinity(0)),
                blas=blas, numpy=numpy.__version__, scipy=scipy.__version__)


def cpu_baseline(data, nsample: int = 256, warmups: int = 3, reps: int = 10):
    """CPU FP64 restatement of jaxfun's algorithm (oracle.spectral's
    partial diagonalisation with BLAS transforms and LAPACK banded row
    solves) on an N-stratified sample of the same ragged batch, timed on the
    host cores in steady state: the per-N eigenpairs of every sampled N are
    built before timing (the GPU timing likewise excludes table set-up), then
    3 warm-up passes and the median of 10 timed passes over the sample."""
    from oracle import spectral as O

    Ns, alphas, loads = data
    order = np.argsort(Ns, kind="stable")
    pick = order[np.linspace(0, len(order) - 1, nsample).round().astype(int)]  # every (4096/256)-th by N
    t0 = time.perf_counter()
    for i in pick:  # table set-up, reported separately
        O.helmholtz_partial_fd_lapack(int(Ns[i]), float(alphas[i]), loads[i])
    setup_s = time.perf_counter() - t0

    def one_pass():
        for i in pick:
            O.helmholtz_partial_fd_lapack(int(Ns[i]), float(alphas[i]), loads[i])

    for _ in range(warmups):
        one_pass()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        one_pass()
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    info = _host_info()
    threads = max([b["threads"] or 1 for b in info["blas"]] or [1])
    return dict(value=nsample / med, unit="solves/s", cores=threads, kind="port",
                label="CPU FP64 restatement of jaxfun's algorithm",
                sample=f"{nsample} of the 4096 solves (seed 2) stratified by N (every 16th in N order, "
                       f"N {int(Ns[pick].min())}..{int(Ns[pick].max())}), median of {reps} passes after "
                       f"{warmups} warm-ups, per-N eigenpairs built beforehand",
                pass_s_median=med, pass_s_min=float(min(times)), table_setup_s=setup_s, host=info)


def _kernel_src_sha(name: str) -> str:
    import hashlib

    with open(os.path.join(ROOT, "jaxfun_amd", "csrc", name), "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()[:16]


def helm_roofline(mixed, kname, m):
    """Roofline of the ragged-batch kernel by the workload's algorithmic count:
    F = sum(4 n^3 + 14 n^2) flop per batch, m 16-bit MFMA products per
    logical product (6 for the 3-term bf16 split, 3 for the 2-term fp16
    split; both run at the same 2.5 PFLOP/s dense peak), bytes = F in + U out
    once (FP32).  Roofline time is max(F m / 2.5 PFLOP/s, bytes / 8 TB/s);
    the binding roof is reported."""
    ks = mixed["kernel_ms"] * 1e-3
    t_mfma = mixed["alg_flops"] * m / (BF16_PEAK_TFS * 1e12)
    t_hbm = mixed["io_bytes"] / (HBM_PEAK_GBS * 1e9)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "helm_ragged_4096_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tj = json.load(fh)
        # only valid for the kernel source it was measured on
        if tj.get("kernel") == kname and tj.get("src_sha") == _kernel_src_sha("helmholtz.hip"):
            traffic = tj["fetch_bytes"] + tj["write_bytes"]
    mfma_exec = mixed["mfma_flops"] / ks / 1e12
    if t_mfma >= t_hbm:
        achieved, peak, unit = mixed["alg_flops"] * m / ks / 1e12, BF16_PEAK_TFS, "TFLOP/s"
    else:
        achieved, peak, unit = mixed["io_bytes"] / ks / 1e9, HBM_PEAK_GBS, "GB/s"
    return {"bound": "mfma" if t_mfma >= t_hbm else "hbm", "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "traffic": traffic, "kernel": kname, "kernel_ms": mixed["kernel_ms"],
            "roof_ms": {"mfma": t_mfma * 1e3, "hbm": t_hbm * 1e3},
            "alg_flops_per_launch": mixed["alg_flops"], "products_per_flop": m,
            "alg_bytes_per_launch": mixed["io_bytes"],
            "hbm_frac": mixed["io_bytes"] / ks / 1e9 / HBM_PEAK_GBS,
            "launch": {"nwaves": mixed["nwaves"], "nblocks": mixed["nblocks"], "queues": mixed["queues"]},
            "mfma_executed_TFLOPs": mfma_exec, "mfma_executed_frac_of_bf16_peak": mfma_exec / BF16_PEAK_TFS}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skip-burgers", action="store_true", help="skip the 2D Burgers IMEX workload (1000 ARS443 steps)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy 3x (known traffic, for PMC scaling)")
    ap.add_argument("--skip-pcg", action="store_true", help="skip the variable-coefficient PCG workload (train + evaluate)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(every configuration is a single-GPU workload)", file=sys.stderr)
        sys.exit(2)
    import torch

    torch.cuda.init()
    if args.copy_calibration:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    res, data = bench_helmholtz(args.steps, args.warmup)
    mixed, bf6, fp64 = res["mixed16"], res["mixed"], res["fp64"]
    kname = "helm_compact_kernel<float, 1>"
    kname6 = "helm_compact_kernel<float, 0>" if bf6["nwaves"] == 4 else "helm_pair_kernel<float>"
    err = mixed["rel_err_vs_fp64_gpu"]
    line = {
        "metric": "solves/s + 3D transform GB/s at <=1e-6 rel err vs plain FP64 GPU; PCG iters",
        "value": mixed["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16x2-split (3 products, power-of-two block scaling)/fp32-acc "
                 "(fp64 tridiagonal recurrences, fp32-stored sweep values)",
        "data": "synthetic (seeded generators, jaxfun_amd/workloads.py)",
        "config": {"workload": "ragged batch of 4096 2D Legendre-Dirichlet Helmholtz solves, "
                               "N~U{32..256}, alpha=10^U(-2,4), seed 2", "io_dtype": "float32",
                   "parallelism": "single GPU"},
        "rel_err_vs_fp64_gpu": err,
        "meets_1e-6": err <= 1e-6,
        "fp64_path": {"solves_per_s": fp64["solves_per_s"], "ms_per_step": fp64["ms"],
                      "kernel_ms": fp64["kernel_ms"]},
        "speedup_vs_fp64_gpu": mixed["solves_per_s"] / fp64["solves_per_s"],
        "roofline": helm_roofline(mixed, kname, 3.0),
        "bf16x6_path": {"solves_per_s": bf6["solves_per_s"], "ms_per_step": bf6["ms"],
                        "kernel_ms": bf6["kernel_ms"], "rel_err_vs_fp64_gpu": bf6["rel_err_vs_fp64_gpu"],
                        "speedup_vs_fp64_gpu": bf6["solves_per_s"] / fp64["solves_per_s"],
                        "roofline": helm_roofline(bf6, kname6, 6.0)},
    }
    t0 = time.perf_counter()

    def run(key, fn):
        line[key] = fn()
        # progress on stderr (stdout carries the one JSON line)
        print(f"bench: {key} done at {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    print("bench: ragged Helmholtz batch done", file=sys.stderr, flush=True)
    if not args.no_extras:
        run("transform_3d_16x128^3", bench_transform_3d)
        run("poisson_chebyshev_N64", bench_poisson_chebyshev)
        run("lifted_helmholtz_legendre_M50", bench_lifted_helmholtz)
        run("periodic_poisson_fourier20_legendre80", bench_periodic_poisson)
        run("band_solve_legendre_mass", bench_band_solve)
    if not (args.skip_burgers or args.no_extras):
        run("burgers_imex_128x128", bench_burgers)
    if not (args.skip_pcg or args.no_extras):
        run("learned_pcg_varcoeff_64", bench_learned_pcg)
    if not args.no_cpu:
        run("cpu_baseline", lambda: cpu_baseline(data))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
