"""Time integrators (reference: src/jaxfun/integrators)."""

from .imex_rk import IMEXRungeKutta
from .tableau import ARS443, ButcherTableau, IMEXTableau

__all__ = ["ARS443", "ButcherTableau", "IMEXRungeKutta", "IMEXTableau"]
