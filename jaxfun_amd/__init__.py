"""jaxfun_amd: MI355X-native spectral-Galerkin hot path with jaxfun's API.

Device work runs in the hand-written HIP kernels of ``libjfx.so`` (see
``include/jfx.h``); there is no CPU fallback.
"""

from . import kernels  # noqa: F401
from .coordinates import x, y, z  # noqa: F401
from .galerkin import (  # noqa: F401
    Chebyshev,
    ChebyshevU,
    Fourier,
    FunctionSpace,
    Jacobi,
    Legendre,
    TensorProduct,
    TestFunction,
    TrialFunction,
    Ultraspherical,
    inner,
    project,
)
from .operators import Div, Dot, Grad  # noqa: F401
from .utils.common import n, ulp  # noqa: F401

__version__ = "0.1.0"
