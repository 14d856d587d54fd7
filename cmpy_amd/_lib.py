"""ctypes binding of libcmpy_amd.so (the C ABI declared in include/cmpy_amd.h).

torch is imported before the library is loaded: torch ships its own HIP runtime
(soname libamdhip64.so.7) and loading it first makes the library bind to that
same runtime, so device pointers and streams from torch are valid here.

There is no fallback: if the shared library is missing or fails to load, every
entry point raises :class:`EDLibraryError`.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import torch  # noqa: F401  (must precede the library load, see module docstring)

LIB_PATH = Path(__file__).resolve().parent / "libcmpy_amd.so"

ED_OK = 0
ED_ERR_INVALID = -1
ED_ERR_HIP = -2
ED_ERR_NOMEM = -3
ED_ERR_NUMERIC = -4
ED_ERR_LIMIT = -5
ED_UP, ED_DN = 1, 2
ED_SIGN_REFERENCE, ED_SIGN_JORDAN_WIGNER = 0, 1


class EDLibraryError(RuntimeError):
    """The native library is missing or could not be loaded."""


class EDError(RuntimeError):
    """A C-ABI call returned a negative status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


class EDInvalidArgument(EDError, ValueError):
    pass


class EDNumericError(EDError, ArithmeticError):
    pass


# -- structures ---------------------------------------------------------------

class ModelDesc(C.Structure):
    _fields_ = [
        ("num_sites", C.c_int32), ("n_up", C.c_int32), ("n_dn", C.c_int32),
        ("num_bonds", C.c_int32), ("bonds", C.POINTER(C.c_int32)),
        ("hop", C.POINTER(C.c_double)), ("eps", C.POINTER(C.c_double)),
        ("u", C.POINTER(C.c_double)), ("width", C.c_int32),
    ]


class LanczosCfg(C.Structure):
    _fields_ = [
        ("max_iter", C.c_int32), ("min_iter", C.c_int32), ("check_every", C.c_int32),
        ("reorth", C.c_int32), ("want_vector", C.c_int32), ("fixed_iter", C.c_int32),
        ("tol", C.c_double), ("seed", C.c_uint64), ("time_matvec", C.c_int32), ("pad", C.c_int32),
    ]


class LanczosResult(C.Structure):
    _fields_ = [
        ("e0", C.c_double), ("ritz_residual", C.c_double), ("residual", C.c_double),
        ("iterations", C.c_int32), ("converged", C.c_int32), ("breakdown", C.c_int32),
        ("pad", C.c_int32), ("matvec_ms", C.c_double), ("total_ms", C.c_double),
    ]


class CfParts(C.Structure):
    _fields_ = [
        ("alpha", C.POINTER(C.c_double) * 2), ("beta", C.POINTER(C.c_double) * 2),
        ("norm2", C.c_double * 2), ("steps", C.c_int32 * 2),
    ]


class BatchDesc(C.Structure):
    _fields_ = [
        ("num_models", C.c_int64), ("num_bath", C.POINTER(C.c_int32)),
        ("n_up", C.POINTER(C.c_int32)), ("n_dn", C.POINTER(C.c_int32)),
        ("u", C.POINTER(C.c_double)), ("eps_imp", C.POINTER(C.c_double)),
        ("eps_bath", C.POINTER(C.c_double)), ("v", C.POINTER(C.c_double)),
        ("max_bath", C.c_int32), ("sign_width", C.c_int32), ("max_iter", C.c_int32),
        ("cf_steps", C.c_int32), ("check_every", C.c_int32), ("keep_basis", C.c_int32),
        ("tol", C.c_double), ("seed", C.c_uint64), ("cf_tol", C.c_double),
    ]


BATCH_MAX_BATH = 15  # ED_BATCH_MAX_BATH


vp = C.c_void_p
i32, i64, u64, dbl = C.c_int32, C.c_int64, C.c_uint64, C.c_double
P = C.POINTER

# name -> (restype, argtypes); every function declared in include/cmpy_amd.h
SIGNATURES = {
    "ed_last_error": (C.c_char_p, []),
    "ed_version": (C.c_int, []),
    "ed_device_count": (C.c_int, []),
    "ed_binom": (i64, [C.c_int, C.c_int]),
    "ed_sector_states_host": (C.c_int, [C.c_int, C.c_int, P(C.c_uint32)]),
    "ed_sector_states": (C.c_int, [C.c_int, C.c_int, vp, vp]),
    "ed_sector_rank": (C.c_int, [C.c_int, C.c_int, vp, i64, vp, vp]),
    "ed_hop_terms": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dbl, vp, vp, vp]),
    "ed_op_create": (C.c_int, [P(vp), P(ModelDesc), vp]),
    "ed_op_destroy": (C.c_int, [vp]),
    "ed_op_shape": (C.c_int, [vp, P(i64), P(i64)]),
    "ed_op_matvec": (C.c_int, [vp, vp, vp, i64, vp]),
    "ed_op_matvec_ld": (C.c_int, [vp, vp, vp, i64, i64, vp]),
    "ed_op_vector_ld": (C.c_int, [vp, P(i64)]),
    "ed_csr_matvec": (C.c_int, [vp, vp, vp, i64, vp, vp, vp]),
    "ed_op_diagonal": (C.c_int, [vp, vp, vp]),
    "ed_op_dense": (C.c_int, [vp, vp, vp]),
    "ed_op_hop_table": (C.c_int, [vp, C.c_int, P(i32), vp, vp, vp, vp]),
    "ed_apply_cdag": (C.c_int, [C.c_int] * 6 + [vp, vp, i64, vp]),
    "ed_apply_c": (C.c_int, [C.c_int] * 6 + [vp, vp, i64, vp]),
    "ed_lanczos_workspace": (C.c_int, [vp, P(LanczosCfg), P(i64)]),
    "ed_lanczos_groundstate": (C.c_int, [vp, P(LanczosCfg), vp, vp, vp, i64, P(LanczosResult), vp]),
    "ed_lanczos_coeffs_workspace": (C.c_int, [vp, i32, i32, P(i64)]),
    "ed_lanczos_coeffs": (C.c_int, [vp, vp, i32, i32, P(dbl), P(dbl), P(dbl), P(i32), vp, vp, i64, vp]),
    "ed_cf_eval": (C.c_int, [P(dbl), P(dbl), i32, dbl, dbl, i32, vp, i64, vp, i32, vp]),
    "ed_tevo_eval": (C.c_int, [P(dbl), P(dbl), i32, dbl, dbl, i32, vp, i64, vp, i32, vp]),
    "ed_pole_sum": (C.c_int, [vp, vp, i64, vp, i64, vp, i32, vp]),
    "ed_fourier_t2z": (C.c_int, [vp, vp, i64, vp, i64, vp, vp]),
    "ed_greens_cf_workspace": (C.c_int, [vp, vp, i32, i32, P(i64)]),
    "ed_greens_cf": (C.c_int, [vp, vp] + [i32] * 6 + [vp, dbl, i32, i32, vp, i64, vp, P(CfParts), vp, i64, vp]),
    "ed_greens_lehmann_t0_workspace": (C.c_int, [vp, vp, P(i64)]),
    "ed_greens_lehmann_t0": (C.c_int, [vp, vp] + [i32] * 6 + [vp, dbl, vp, i64, vp, vp, i64, vp]),
    "ed_lehmann_pairs": (C.c_int, [vp, vp, i64, vp, i64, dbl, dbl, vp, i64, vp, vp]),
    "ed_lehmann_pairs_ij": (C.c_int, [vp, vp, vp, i64, vp, i64, dbl, dbl, vp, i64, vp, vp]),
    "ed_lehmann_pairs_noalloc": (C.c_int, [vp, vp, i64, vp, i64, dbl, dbl, vp, i64, vp, vp]),
    "ed_gemm_nt": (C.c_int, [vp, vp, vp, i64, i64, i64, vp]),
    "ed_occupation": (C.c_int, [vp, i64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp]),
    "ed_occupation_states": (C.c_int, [vp, i64, vp, i64, vp, i64, C.c_int, C.c_int, vp, vp, vp]),
    "ed_dense_eigh": (C.c_int, [vp, i64, vp, vp]),
    "ed_dense_eigh_batched": (C.c_int, [vp, i64, i64, vp, vp]),
    "ed_batch_create": (C.c_int, [P(vp), P(BatchDesc), vp]),
    "ed_batch_destroy": (C.c_int, [vp]),
    "ed_batch_config": (C.c_int, [vp, i64, i32, i32, vp]),
    "ed_batch_info": (C.c_int, [vp, P(i64), P(i32), P(i32), P(dbl)]),
    "ed_batch_solve": (C.c_int, [vp, P(dbl), i32, P(dbl), P(dbl), P(i32), P(dbl), P(dbl), vp]),
    "ed_batch_backward": (C.c_int, [vp, P(dbl), P(dbl), P(dbl), P(dbl), P(i32), P(dbl), vp]),
    "ed_batch_set_bath": (C.c_int, [vp, P(dbl), P(dbl), vp]),
}

_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise EDLibraryError(
            f"{LIB_PATH} is missing: build it with `python -m cmpy_amd.build` "
            "(there is no CPU fallback)"
        )
    try:
        lib = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
    except OSError as exc:  # pragma: no cover
        raise EDLibraryError(f"cannot load {LIB_PATH}: {exc}") from exc
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def lib():
    return _load()


def last_error() -> str:
    msg = lib().ed_last_error()
    return msg.decode() if msg else ""


def check(status: int) -> None:
    if status == ED_OK:
        return
    msg = last_error()
    if status == ED_ERR_INVALID:
        raise EDInvalidArgument(status, msg)
    if status == ED_ERR_NUMERIC:
        raise EDNumericError(status, msg)
    raise EDError(status, msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


# -- device plumbing (torch provides HBM allocations and the stream) -------------

def require_gpu() -> None:
    if lib().ed_device_count() <= 0 or not torch.cuda.is_available():
        raise EDLibraryError("no HIP device is visible: the cmpy_amd engine runs on the GPU only")


def stream_ptr():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def dptr(t: torch.Tensor):
    assert t.is_cuda, "expected a device tensor"
    return C.c_void_p(t.data_ptr())


def f64p(a):
    import numpy as np

    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(P(C.c_double))


def i32p(a):
    import numpy as np

    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(P(C.c_int32))
