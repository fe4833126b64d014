"""Cartesian coordinate systems (reference: coordinates.py, Cartesian subset).

Only Cartesian systems (metric ``sg = 1``, galerkin/inner.py:266-267) are on
the path this package covers.
"""

from __future__ import annotations

import sympy as sp

x, y, z = sp.symbols("x y z", real=True)


class CartCoordSys:
    """Cartesian coordinate system over ``symbols`` (coordinates.py:102)."""

    sg = sp.S.One

    def __init__(self, name: str = "N", symbols: tuple = (x,)):
        self.name = name
        self._symbols = tuple(symbols)

    def base_scalars(self) -> tuple:
        return self._symbols

    @property
    def dims(self) -> int:
        return len(self._symbols)

    @property
    def x(self):
        return self._symbols[0]

    def sub_system(self, i: int) -> "CartCoordSys":
        return CartCoordSys(f"{self.name}{i}", (self._symbols[i],))

    def expr_psi_to_base_scalar(self, expr):
        """Identity for Cartesian systems (coordinates are the base scalars)."""
        return expr


def default_system(dims: int) -> CartCoordSys:
    return CartCoordSys("N", {1: (x,), 2: (x, y), 3: (x, y, z)}[dims])
