"""ctypes binding of ``libpynof_amd.so`` (the C ABI declared in include/pynof_amd.h).

The library is loaded from the package directory only. There is no fallback:
if it is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import pathlib
import threading

_LIB_PATH = pathlib.Path(__file__).resolve().parent / "libpynof_amd.so"
_lock = threading.Lock()
_lib = None

OK, EINVAL, EHIP, ENOMEM, ESTATE, EUNSUPPORTED = 0, 1, 2, 3, 4, 5
FP64, SPLIT, SPLIT_REFINE = 0, 1, 2
OUT_J, OUT_K, OUT_JMO, OUT_KMO, OUT_HCORE = 1, 2, 4, 8, 16
OUT_ALL = 31

PRECISIONS = {"fp64": FP64, "split": SPLIT, "split_refine": SPLIT_REFINE}
OCC_SOFTMAX, OCC_TRIGONOMETRIC = 0, 1
OCC_METHODS = {"Softmax": OCC_SOFTMAX, "Trigonometric": OCC_TRIGONOMETRIC}


class Pnof(ctypes.Structure):
    """struct pnjk_pnof: the PNOF functional and occupation parametrisation."""
    _fields_ = [(name, ctypes.c_int32) for name in
                ("ipnof", "ista", "occ_method", "high_spin", "mspin", "no1", "ndoc", "nsoc", "ncwo", "nbf5")]

# name -> (restype, argtypes)
_SIGS = {
    "pnjk_last_error": (ctypes.c_char_p, []),
    "pnjk_version": (ctypes.c_char_p, []),
    "pnjk_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "pnjk_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "pnjk_upload_b": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_double, ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.POINTER(ctypes.c_double)]),
    "pnjk_upload_b_3c": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                        ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                        ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double)]),
    "pnjk_download_b": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "pnjk_packed_ld": (ctypes.c_int64, [ctypes.c_int64, ctypes.c_int64]),
    "pnjk_upload_b_packed": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64)]),
    "pnjk_jk_build": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_uint32,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p]),
    "pnjk_jk_build_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                             ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_uint32,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p]),
    "pnjk_mo_integrals": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                         ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]),
    "pnjk_lagrangian": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                       ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pnjk_fock_rc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                    ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
    "pnjk_fock_ro": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                    ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
    "pnjk_lagrange_fock": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                          ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                          ctypes.c_void_p]),
    "pnjk_set_profiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "pnjk_stage_count": (ctypes.c_int, []),
    "pnjk_set_tile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "pnjk_stage_name": (ctypes.c_char_p, [ctypes.c_int]),
    "pnjk_stage_times": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "pnjk_group_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "pnjk_group_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "pnjk_group_build": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "pnjk_lagrange_convergence": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                 ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                                 ctypes.c_void_p]),
    "pnjk_occupation_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_void_p] * 11),
    "pnjk_pnof_coefficients": (ctypes.c_int, [ctypes.c_void_p] * 6),
    "pnjk_energy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32]
                    + [ctypes.c_void_p] * 8),
    "pnjk_footprint": (ctypes.c_uint64, [ctypes.c_void_p]),
    "pnjk_workspace_bytes": (ctypes.c_uint64, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int]),
}


class PnjkError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[pynof_amd error {code}] {msg}")
        self.code = code


def lib_path() -> pathlib.Path:
    return _LIB_PATH


def lib():
    """The loaded library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not _LIB_PATH.exists():
                raise RuntimeError(
                    f"{_LIB_PATH} is missing: build it with `python -m pynof_amd.build` "
                    "(the JK build has no CPU fallback)")
            handle = ctypes.CDLL(str(_LIB_PATH))
            for name, (res, args) in _SIGS.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def declared_symbols() -> list[str]:
    return sorted(_SIGS)


def check(rc: int) -> None:
    if rc != OK:
        msg = lib().pnjk_last_error().decode(errors="replace")
        raise PnjkError(rc, msg)
