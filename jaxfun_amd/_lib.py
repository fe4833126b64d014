"""ctypes binding of ``libjfx.so``, the HIP kernel library behind every
device operation of this package.

The library is loaded lazily on first use.  There is no fallback: when the
library is missing, or no GPU is visible, every device operation raises.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjfx.so")

JFX_OK = 0
JFX_EHIP = -1
JFX_ESHAPE = -2
JFX_ESINGULAR = -3
JFX_EARG = -4

c_int = ctypes.c_int
c_ll = ctypes.c_longlong
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p

# name -> (restype, argtypes); the single source of truth for the ABI that
# include/jfx.h declares (tests check both agree).
SIGNATURES: dict[str, tuple] = {
    "jfx_version": (c_int, []),
    "jfx_last_error": (ctypes.c_char_p, []),
    "jfx_device_sync": (c_int, []),
    "jfx_gauss_nodes": (c_int, [c_int, c_int, c_vp, c_vp, c_vp]),
    "jfx_vandermonde_f64": (c_int, [c_int, c_vp, c_int, c_int, c_int, c_dbl, c_vp, c_vp]),
    "jfx_vandermonde_rec_f64": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_dbl, c_vp, c_vp]),
    "jfx_contract_f64": (c_int, [c_vp, c_int, c_int, c_vp, c_vp, c_ll, c_ll, c_vp]),
    "jfx_contract_f32": (c_int, [c_vp, c_int, c_int, c_vp, c_vp, c_ll, c_ll, c_vp]),
    "jfx_split_matrix": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp]),
    "jfx_contract_bf16x": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_vp, c_ll, c_ll, c_vp]),
    "jfx_contract_ozaki": (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_vp, c_vp, c_ll, c_ll, c_vp]),
    "jfx_split_matrix_i8": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_vp, c_vp]),
    "jfx_plane_i8": (c_int, [c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_ll, c_int, c_int, c_int, c_int, c_vp]),
    "jfx_contract_i8": (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_vp, c_vp, c_ll, c_ll, c_vp]),
    "jfx_band_lu_factor_wide": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_vp, c_vp]),
    "jfx_band_lu_solve_wide": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_ll, c_vp]),
    "jfx_kron_sum": (c_int, [c_vp, c_vp, c_vp, c_int, c_int, c_int, c_vp, c_vp]),
    "jfx_gather_rows": (c_int, [c_vp, c_vp, c_int, c_ll, c_vp, c_vp]),
    "jfx_band_lu_factor": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_vp, c_vp]),
    "jfx_band_lu_solve": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_ll, c_ll, c_vp]),
    "jfx_band_lu_solve_mode": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_ll, c_ll, c_int, c_vp]),
    "jfx_band_lu_solve_grid_work_bytes": (c_ll, [c_int, c_int, c_int, c_ll, c_ll]),
    "jfx_band_lu_solve_grid": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_ll, c_ll, c_vp, c_ll, c_vp]),
    "jfx_band_lu_solve_batched_mode": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_vp, c_ll, c_ll, c_int, c_int,
                                               c_vp]),
    "jfx_dense_lu_factor": (c_int, [c_int, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "jfx_dense_lu_solve": (c_int, [c_int, c_vp, c_int, c_vp, c_vp, c_vp, c_ll, c_vp, c_vp]),
    "jfx_div_bcast_f64": (c_int, [c_vp, c_vp, c_vp, c_ll, c_ll, c_vp]),
    "jfx_div_bcast_f32": (c_int, [c_vp, c_vp, c_vp, c_ll, c_ll, c_vp]),
    "jfx_mul_bcast_f64": (c_int, [c_vp, c_vp, c_vp, c_ll, c_ll, c_vp]),
    "jfx_axpy_bcast_f64": (c_int, [c_vp, c_vp, c_dbl, c_vp, c_ll, c_ll, c_vp]),
    "jfx_cplx_combine": (c_int, [c_vp, c_vp, c_ll, c_ll, c_ll, c_int, c_vp]),
    "jfx_fourier_basis": (c_int, [c_vp, c_int, c_int, c_int, c_dbl, c_int, c_int, c_dbl, c_vp, c_vp]),
    "jfx_band_lu_factor_batched": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp]),
    "jfx_band_lu_factor_batched_z": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp]),
    "jfx_band_lu_solve_batched_z": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_vp, c_ll, c_ll, c_vp]),
    "jfx_cmul_bcast": (c_int, [c_vp, c_vp, c_vp, c_ll, c_ll, c_vp]),
    "jfx_band_lu_solve_batched": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_vp, c_ll, c_ll, c_int, c_vp]),
    "jfx_lincomb": (c_int, [c_int, c_int, c_vp, c_vp, c_vp, c_ll, c_int, c_vp]),
    "jfx_cast": (c_int, [c_int, c_vp, c_vp, c_ll, c_vp]),
    "jfx_sum_products": (c_int, [c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_ll, c_vp]),
    "jfx_helm_item_size": (c_int, []),
    "jfx_helm_scratch_doubles_per_block": (c_int, []),
    "jfx_helm_ragged": (c_int, [c_int, c_int, c_int, c_vp, c_int, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp,
                                c_vp, c_vp, c_vp, c_vp, c_int, c_vp]),
    "jfx_helm_noalloc_max_queues": (c_int, []),
    "jfx_helm_ragged_noalloc": (c_int, [c_int, c_int, c_vp, c_int, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp,
                                        c_int, c_vp]),
    "jfx_gemm_bf16": (c_int, [c_int, c_int, c_int, c_int, c_int, ctypes.c_float, c_vp, c_int, c_vp, c_int,
                              ctypes.c_float, c_vp, c_int, c_vp, c_int, c_vp]),
    "jfx_tanh_backward": (c_int, [c_vp, c_vp, c_vp, c_ll, c_vp]),
    "jfx_colsum": (c_int, [c_vp, c_vp, c_int, c_int, c_vp]),
    "jfx_adam": (c_int, [c_vp, c_vp, c_vp, c_vp, c_ll, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                         ctypes.c_float, c_int, c_vp]),
    "jfx_batched_dot": (c_int, [c_vp, c_vp, c_vp, c_int, c_ll, c_vp]),
    "jfx_pcg_update_xr": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_ll, c_vp]),
    "jfx_pcg_update_p": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_ll, c_vp]),
    "jfx_pcg_check": (c_int, [c_vp, c_vp, c_dbl, c_vp, c_vp, c_int, c_int, c_vp]),
    "jfx_scale_batch": (c_int, [c_vp, c_vp, c_vp, c_int, c_ll, c_vp]),
    "jfx_exp": (c_int, [c_vp, c_vp, c_dbl, c_ll, c_vp]),
    "jfx_plane_contract": (c_int, [c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_dbl, c_int, c_ll,
                                   c_int, c_int, c_int, c_int, c_vp]),
    "jfx_plane_fd": (c_int, [c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_ll,
                             c_int, c_vp]),
    "jfx_scale_transpose": (c_int, [c_vp, c_int, c_int, c_vp, c_vp, c_dbl, c_vp, c_vp]),
}

_lib = None


class JfxError(RuntimeError):
    """A HIP-level failure reported by libjfx."""


def load():
    """Load ``libjfx.so`` and bind every declared entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m jaxfun_amd.build` "
            "(this package has no CPU fallback)"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> None:
    """Map a libjfx status code to the exception class the reference raises."""
    if rc == JFX_OK:
        return
    msg = load().jfx_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == JFX_ESINGULAR:
        raise ValueError(text)
    if rc in (JFX_ESHAPE, JFX_EARG):
        raise ValueError(text)
    raise JfxError(text)


def call(name: str, *args) -> None:
    rc = getattr(load(), name)(*args)
    check(rc, name)
