"""Small shared helpers (reference: utils/common.py)."""

from __future__ import annotations

from typing import NamedTuple

import numpy as np
import sympy as sp

# Symbol used for stencil scalings, e.g. ``scaling=n + 1`` (utils/common.py).
n = sp.Symbol("n", integer=True, nonnegative=True)


class Domain(NamedTuple):
    """Physical interval (utils/common.py:98-100)."""

    lower: float
    upper: float


def ulp(x: float, dtype=np.float64) -> float:
    """Unit in the last place of x (utils/common.py:103-104)."""
    return float(np.spacing(dtype(abs(x))))


def lambdify(args, expr, modules="numpy"):
    """SymPy lambdify evaluated with numpy on host meshes (utils/common.py:155-176)."""
    return sp.lambdify(args, expr, modules=modules)
