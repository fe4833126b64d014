"""Diagonally implicit IMEX Runge-Kutta in coefficient space
(reference: integrators/imex_rk.py, integrators/base.py).

The weak form ``v*(u.diff(t) + N(u) - L(u))`` is split into a mass term
(the time derivative), terms linear in the trial function (assembled with
``inner`` into TPMatrix operators) and quadratic nonlinear terms
``c * v * d^a u * d^b u`` (first derivatives at most), which are evaluated on
a padded quadrature mesh (the 3/2 rule by default) and projected back with the
test space's scalar product (integrators/base.py:238-248).

Every array operation is a libjfx kernel: per-axis MFMA contractions for the
transforms, mass and linear operators, fast-diagonalisation stage solves
(non-symmetric pairs handled exactly), a pointwise sum-of-products kernel for
the nonlinear term and a fused linear-combination kernel for stage sums.
``precision='mixed'`` runs the contractions as 3-term bf16 splits with FP32
accumulation and FP32 state.
"""

from __future__ import annotations

import sympy as sp
import torch

from .. import kernels as K
from ..galerkin.arguments import ArgumentTag, functionspace_of, get_arg, t as tsym
from ..galerkin.inner import inner
from ..galerkin.orthogonal import as_device
from ..galerkin.tensorproductspace import TensorProductSpace
from ..la import TPMatrices, TPMatrix
from .tableau import ARS443, IMEXTableau


def _trial_factors(term):
    out = []
    rest = sp.S.One
    for f in sp.Mul.make_args(term):
        if isinstance(f, sp.Pow) and get_arg(f.base) == ArgumentTag.TRIAL and f.exp.is_Integer:
            out += [f.base] * int(f.exp)
        elif get_arg(f) == ArgumentTag.TRIAL:
            out.append(f)
        else:
            rest = rest * f
    return out, rest


class IMEXRungeKutta:
    def __init__(self, equation, *, tableau: IMEXTableau = ARS443, initial, time=None,
                 sparse: bool = False, sparse_tol: int = 1000, precision: str = "f64",
                 padding: float = 1.5):
        self.tableau = tableau
        self.time = time
        self.precision = {"mixed": "bf16x6"}.get(precision, precision)
        self.dt_io = K.io_dtype(self.precision)
        expr = sp.expand(sp.sympify(equation).doit())
        mass, linear, nonlin = sp.S.Zero, sp.S.Zero, []
        test = None
        for term in sp.Add.make_args(expr):
            trials, rest = _trial_factors(term)
            derivs_t = [f for f in trials if isinstance(f, sp.Derivative) and tsym in f.variables]
            for f in sp.Mul.make_args(term):
                if get_arg(f) == ArgumentTag.TEST:
                    test = f
            if derivs_t:
                if len(trials) != 1:
                    raise NotImplementedError("nonlinear time-derivative terms")
                mass += term.subs(derivs_t[0], derivs_t[0].expr)
            elif len(trials) == 1:
                linear += -term  # moved to the right-hand side
            elif len(trials) == 2:
                nonlin.append((-rest, trials))
            elif len(trials) == 0:
                raise NotImplementedError("forcing terms are not built for the IMEX path")
            else:
                raise NotImplementedError("nonlinear terms of order > 2")
        if mass == 0:
            raise ValueError("Time integrators require a first-order time derivative in the weak form")
        self.testspace = functionspace_of(test)
        if not isinstance(self.testspace, TensorProductSpace):
            raise NotImplementedError("the GPU IMEX path is built for tensor-product spaces")
        self.symbols = self.testspace.system.base_scalars()
        self.mass = self._as_tpmatrices(inner(self._strip_time(mass), sparse=sparse, sparse_tol=sparse_tol))
        self.linear = self._as_tpmatrices(inner(self._strip_time(linear), sparse=sparse, sparse_tol=sparse_tol)) \
            if linear != 0 else None
        self.nonlinear = []
        for rest, (f1, f2) in nonlin:
            c = sp.simplify(rest / test)
            if c.free_symbols:
                raise NotImplementedError("space-dependent nonlinear coefficients")
            self.nonlinear.append((float(c), self._order(f1), self._order(f2)))
        Np = tuple(int(round(padding * s.num_quad_points)) for s in self.testspace.basespaces)
        self.Np = Np
        self._stage_ops = {}
        self.initial = initial

    # -------------------------------------------------------------- set-up helpers
    @staticmethod
    def _strip_time(e):
        return sp.sympify(e).replace(lambda f: get_arg(f) == ArgumentTag.TRIAL and tsym in getattr(f, "args", ()),
                         lambda f: f.func(*[a for a in f.args if a != tsym]))

    @staticmethod
    def _as_tpmatrices(A):
        if isinstance(A, TPMatrix):
            return TPMatrices([A])
        return A

    def _order(self, f) -> tuple:
        if isinstance(f, sp.Derivative):
            d = dict((s, int(c)) for s, c in f.variable_count)
            out = tuple(d.get(s, 0) for s in self.symbols)
        else:
            out = (0,) * len(self.symbols)
        if max(out) > 1:
            raise NotImplementedError("nonlinear terms with derivatives above first order")
        return out

    def initial_coefficients(self):
        if isinstance(self.initial, sp.Expr):
            from ..galerkin.inner import project

            return K.cast(project(self.initial, self.testspace), self.dt_io)
        return as_device(self.initial, self.dt_io)

    def stage_operator(self, coef: float) -> TPMatrices:
        """M - coef * L, factored once per distinct value (imex_rk.py:39-48)."""
        key = float(coef)
        if key not in self._stage_ops:
            op = self.mass + self.linear.scale(-coef) if self.linear is not None else self.mass
            self._stage_ops[key] = op.fd_factor()
        return self._stage_ops[key]

    # -------------------------------------------------------------- operators
    def apply_mass(self, U):
        return self.mass(U, self.precision)

    def apply_linear(self, U):
        return self.linear(U, self.precision) if self.linear is not None else None

    def nonlinear_rhs_scalar_product(self, U):
        """(v, N(u)) with N evaluated on the padded mesh (integrators/base.py:238-248)."""
        if not self.nonlinear:
            return None
        T = self.testspace
        prec = self.precision
        nb = U.ndim - 2
        fields = {}
        needed = set()
        for _, o1, o2 in self.nonlinear:
            needed.add(o1)
            needed.add(o2)
        # share the axis-1 transforms between fields with the same y-derivative
        half = {}
        for (kx, ky) in sorted(needed):
            if ky not in half:
                half[ky] = K.apply_along(T[1].backward_operator(self.Np[1], ky), U, nb + 1, prec)
            fields[(kx, ky)] = K.apply_along(T[0].backward_operator(self.Np[0], kx), half[ky], nb, prec)
        # a factor shared by every term goes first: the kernel then reads it once
        terms = list(self.nonlinear)
        shared = set.intersection(*[{o1, o2} for _, o1, o2 in terms])
        if shared:
            f = min(shared)
            terms = [(c, o1, o2) if o1 == f else (c, o2, o1) for c, o1, o2 in terms]
        w = K.sum_products([(c, fields[o1], fields[o2]) for c, o1, o2 in terms])
        return T.scalar_product(w, precision=prec)

    # -------------------------------------------------------------- stepping
    def step(self, U, dt: float):
        """One step (integrators/imex_rk.py:96-159); stage 0 of a tableau with
        a_00 = 0 is U itself (M^-1 M U).  A globally stiffly accurate tableau
        returns its last stage; otherwise the update is
        M^-1 (M U + dt sum_i (bE_i N_i + bI_i L_i))."""
        tab = self.tableau
        aE, aI = tab.explicit.A, tab.implicit.A
        mu = self.apply_mass(U)
        Ns, Ls, stage = [], [], None
        last = tab.stages - 1
        for i in range(tab.stages):
            terms = [(1.0, mu)]
            for j in range(i):
                if aE[i][j] != 0.0 and Ns[j] is not None:
                    terms.append((dt * aE[i][j], Ns[j]))
                if aI[i][j] != 0.0 and Ls[j] is not None:
                    terms.append((dt * aI[i][j], Ls[j]))
            if aI[i][i] == 0.0:
                if i != 0:
                    raise NotImplementedError("explicit stages after the first")
                stage = U
            else:
                rhs = K.lincomb(terms)
                stage = self.stage_operator(dt * aI[i][i]).solve(rhs, self.precision)
            if not (i == last and tab.is_stiffly_accurate):
                Ns.append(self.nonlinear_rhs_scalar_product(stage))
                Ls.append(self.apply_linear(stage))
        if tab.is_stiffly_accurate:
            return stage
        terms = [(1.0, mu)]
        for i in range(tab.stages):
            if tab.explicit.b[i] != 0.0 and Ns[i] is not None:
                terms.append((dt * tab.explicit.b[i], Ns[i]))
            if tab.implicit.b[i] != 0.0 and Ls[i] is not None:
                terms.append((dt * tab.implicit.b[i], Ls[i]))
        return self.mass_solver().solve(K.lincomb(terms), self.precision)

    def mass_solver(self):
        """M^-1 for the final combination of a tableau that is not stiffly accurate."""
        if "mass" not in self._stage_ops:
            self._stage_ops["mass"] = self.mass.fd_factor()
        return self._stage_ops["mass"]

    def solve(self, dt: float, steps: int, U0=None):
        U = self.initial_coefficients() if U0 is None else as_device(U0, self.dt_io)
        for _ in range(steps):
            U = self.step(U, dt)
            if not bool(torch.isfinite(U).all()):
                raise FloatingPointError("non-finite state")
        return U
