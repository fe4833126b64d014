"""CPU FP64 oracle for the spectral-Galerkin hot path.

TEST INFRASTRUCTURE ONLY.  This package restates, in plain numpy/scipy FP64,
the algorithms of jaxfun's spectral-Galerkin path (bases, quadrature,
transforms, Galerkin matrices, banded and fast-diagonalisation solves, IMEX
stepping, the variable-coefficient operator and PCG) so that the HIP product
path can be checked against it.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker (or the timed CPU baseline), never as the thing that is measured or
shipped.  The product package ``jaxfun_amd`` never imports it.

Every function cites the reference file:line it restates.  How each one is
pinned (golden vectors of the reference's own tests, numpy/mpmath/sympy
closed forms, brute-force dense solves) is listed in ``oracle/spectral.py``'s
module docstring; there are no "parity unpinned" functions in it.
"""

from . import spectral  # noqa: F401
