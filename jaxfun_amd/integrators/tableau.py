"""Butcher tableaux (reference: integrators/tableau.py)."""

from __future__ import annotations

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
