"""IMEX ARS443 Burgers (2D Chebyshev-Dirichlet, 3/2 dealiasing) vs the oracle's coefficient-space restatement."""

import numpy as np
import pytest
import sympy as sp

from oracle import spectral as O

pytestmark = pytest.mark.gpu


def _setup(N, nu, precision="f64"):
    from jaxfun_amd import Chebyshev, Div, FunctionSpace, Grad, TensorProduct, TestFunction, TrialFunction
    from jaxfun_amd.galerkin.arguments import t
    from jaxfun_amd.integrators import ARS443, IMEXRungeKutta

    D = FunctionSpace(N, Chebyshev, {"left": {"D": 0}, "right": {"D": 0}})
    T = TensorProduct(D, D)
    v = TestFunction(T)
    u = TrialFunction(T, transient=True)
    x, y = T.system.base_scalars()
    weak = v * (u.diff(t) + u * (u.diff(x) + u.diff(y)) - nu * Div(Grad(u)))
    u0 = sp.sin(sp.pi * (x + 1) / 2) * sp.sin(sp.pi * (y + 1)) * 0.7 - 0.3 * sp.sin(2 * sp.pi * (x + 1) / 2) * \
        sp.sin(sp.pi * (y + 1) / 2)
    return T, IMEXRungeKutta(weak, tableau=ARS443, initial=u0, precision=precision)


def test_burgers_steps_vs_oracle(dev):
    N, nu, dt, steps = 16, 0.1, 2e-3, 6
    T, integ = _setup(N, nu)
    U0 = integ.initial_coefficients()
    U = integ.solve(dt, steps, U0=U0).cpu().numpy()
    bg = O.Burgers2D(N, nu)
    solve = bg.solver(0.5 * dt)
    Ur = U0.cpu().numpy()
    for _ in range(steps):
        Ur = O.imex_ars443_step(Ur, dt, bg.mass_apply, lambda R: O.kron_solve([(1.0, [bg.M, bg.M])], R),
                                bg.linear_apply, solve, bg.nonlinear_sp)
    assert O.rel_l2(U, Ur) < 1e-11
    assert O.rel_l2(U, U0.cpu().numpy()) > 1e-4  # it actually evolved


def test_burgers_batched_and_mixed(dev):
    import torch

    N, nu, dt, steps = 24, 0.05, 1e-3, 10
    T, integ64 = _setup(N, nu, "f64")
    _, integm = _setup(N, nu, "mixed")
    U0 = integ64.initial_coefficients()
    B = torch.stack([U0, 0.5 * U0, -U0]).contiguous()
    U64 = integ64.solve(dt, steps, U0=B)
    Um = integm.solve(dt, steps, U0=B.to(torch.float32))
    single = integ64.solve(dt, steps, U0=(0.5 * U0).contiguous())
    assert O.rel_l2(U64[1].cpu().numpy(), single.cpu().numpy()) < 1e-13
    drift = O.rel_l2(Um.double().cpu().numpy(), U64.cpu().numpy())
    assert drift < 1e-5, drift
