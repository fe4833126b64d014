"""Butcher tableaux (reference: integrators/tableau.py)."""

from __future__ import annotations

import math
from dataclasses import dataclass


@dataclass(frozen=True)
class ButcherTableau:
    A: tuple
    b: tuple
    c: tuple


@dataclass(frozen=True)
class IMEXTableau:
    explicit: ButcherTableau
    implicit: ButcherTableau

    @property
    def stages(self) -> int:
        return len(self.implicit.b)

    @property
    def distinct_diagonal_coeffs(self) -> list[float]:
        out = []
        for i in range(self.stages):
            a = self.implicit.A[i][i]
            if a != 0.0 and a not in out:
                out.append(a)
        return out

    @property
    def is_stiffly_accurate(self) -> bool:
        return (tuple(self.explicit.A[-1]) == tuple(self.explicit.b)
                and tuple(self.implicit.A[-1]) == tuple(self.implicit.b))


# Ascher-Ruuth-Spiteri ARS(4,4,3) (integrators/tableau.py:318-349): published
# coefficients, globally stiffly accurate.
ARS443 = IMEXTableau(
    explicit=ButcherTableau(
        A=((0.0, 0.0, 0.0, 0.0, 0.0),
           (1.0 / 2.0, 0.0, 0.0, 0.0, 0.0),
           (11.0 / 18.0, 1.0 / 18.0, 0.0, 0.0, 0.0),
           (5.0 / 6.0, -5.0 / 6.0, 1.0 / 2.0, 0.0, 0.0),
           (1.0 / 4.0, 7.0 / 4.0, 3.0 / 4.0, -7.0 / 4.0, 0.0)),
        b=(1.0 / 4.0, 7.0 / 4.0, 3.0 / 4.0, -7.0 / 4.0, 0.0),
        c=(0.0, 1.0 / 2.0, 2.0 / 3.0, 1.0 / 2.0, 1.0),
    ),
    implicit=ButcherTableau(
        A=((0.0, 0.0, 0.0, 0.0, 0.0),
           (0.0, 1.0 / 2.0, 0.0, 0.0, 0.0),
           (0.0, 1.0 / 6.0, 1.0 / 2.0, 0.0, 0.0),
           (0.0, -1.0 / 2.0, 1.0 / 2.0, 1.0 / 2.0, 0.0),
           (0.0, 3.0 / 2.0, -3.0 / 2.0, 1.0 / 2.0, 1.0 / 2.0)),
        b=(0.0, 3.0 / 2.0, -3.0 / 2.0, 1.0 / 2.0, 1.0 / 2.0),
        c=(0.0, 1.0 / 2.0, 2.0 / 3.0, 1.0 / 2.0, 1.0),
    ),
)


# First-order IMEX Euler (integrators/tableau.py:129-142): forward Euler
# paired with backward Euler, globally stiffly accurate.
IMEX_EULER = IMEXTableau(
    explicit=ButcherTableau(A=((0.0, 0.0), (1.0, 0.0)), b=(1.0, 0.0), c=(0.0, 1.0)),
    implicit=ButcherTableau(A=((0.0, 0.0), (0.0, 1.0)), b=(0.0, 1.0), c=(0.0, 1.0)),
)


def _ars222() -> IMEXTableau:
    """Ascher-Ruuth-Spiteri (2,2,2) (Appl. Numer. Math. 25, 1997, sec. 2.6):
    g = 1 - 1/sqrt(2), d = 1 - 1/(2g); L-stable, globally stiffly accurate
    (integrators/tableau.py:352-385)."""
    g = 1.0 - 1.0 / math.sqrt(2.0)
    d = 1.0 - 1.0 / (2.0 * g)
    return IMEXTableau(
        explicit=ButcherTableau(A=((0.0, 0.0, 0.0), (g, 0.0, 0.0), (d, 1.0 - d, 0.0)),
                                b=(d, 1.0 - d, 0.0), c=(0.0, g, 1.0)),
        implicit=ButcherTableau(A=((0.0, 0.0, 0.0), (0.0, g, 0.0), (0.0, 1.0 - g, g)),
                                b=(0.0, 1.0 - g, g), c=(0.0, g, 1.0)),
    )


ARS222 = _ars222()


def _imex_ssp2_222() -> IMEXTableau:
    """Pareschi-Russo IMEX-SSP2(2,2,2) (J. Sci. Comput. 25, 2005), g = 1 - 1/sqrt(2):
    implicit stage 0 has a_00 = g and the update uses the weights b
    (not stiffly accurate; integrators/tableau.py:145-164)."""
    g = 1.0 - 1.0 / math.sqrt(2.0)
    return IMEXTableau(
        explicit=ButcherTableau(A=((0.0, 0.0), (1.0, 0.0)), b=(0.5, 0.5), c=(0.0, 1.0)),
        implicit=ButcherTableau(A=((g, 0.0), (1.0 - g, g)), b=(0.5, 0.5), c=(g, 1.0)),
    )


IMEX_SSP2_222 = _imex_ssp2_222()
