"""Time integrators (reference: src/jaxfun/integrators)."""

from .imex_rk import IMEXRungeKutta
from .tableau import ARS222, ARS443, IMEX_EULER, IMEX_SSP2_222, ButcherTableau, IMEXTableau

__all__ = ["ARS222", "ARS443", "IMEX_EULER", "IMEX_SSP2_222", "ButcherTableau", "IMEXRungeKutta", "IMEXTableau"]
