"""Variable-coefficient helpers: pointwise products on the quadrature mesh."""

from __future__ import annotations

import torch

from . import _lib
from . import kernels as K


def mul_bcast(X: torch.Tensor, C: torch.Tensor) -> torch.Tensor:
    """``Y[b, ...] = X[b, ...] * C[...]`` (C broadcast over leading dims)."""
    K._req(X, torch.float64, "X")
    K._req(C, torch.float64, "C")
    Y = torch.empty_like(X)
    _lib.call("jfx_mul_bcast_f64", K._p(X), K._p(C), K._p(Y), X.numel(), C.numel(), K._s())
    return Y
