"""ctypes binding of libqrhip.so (the C ABI declared in include/qrhip.h).

The library is loaded eagerly; a missing or unloadable library raises
ImportError, so the package never runs without its HIP kernels.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqrhip.so")

QR_OK = 0
QR_ERR_ARG = -1
QR_ERR_SIZE = -5
QR_PHASE_ROWWISE = 0
QR_PHASE_KRON = 1
QR_PREC_FP64 = 0
QR_PREC_MIXED = 1

u64 = ctypes.c_uint64
u32 = ctypes.c_uint32
i32 = ctypes.c_int32
vp = ctypes.c_void_p
u64p = ctypes.POINTER(ctypes.c_uint64)
u32p = ctypes.POINTER(ctypes.c_uint32)
i32p = ctypes.POINTER(ctypes.c_int32)
f64p = ctypes.POINTER(ctypes.c_double)
f64 = ctypes.c_double

# name -> (restype, argtypes); mirrors include/qrhip.h
SIGNATURES = {
    "qr_last_error": (ctypes.c_char_p, []),
    "qr_abi_version": (ctypes.c_int, []),
    "qr_device_alloc": (ctypes.c_int, [u64, ctypes.POINTER(vp)]),
    "qr_device_free": (ctypes.c_int, [vp]),
    "qr_copy_to_device": (ctypes.c_int, [vp, vp, u64, vp]),
    "qr_copy_to_host": (ctypes.c_int, [vp, vp, u64, vp]),
    "qr_pauli_parse": (ctypes.c_int, [ctypes.c_char_p, u32p, u32p, u32p, u64p, u64p]),
    "qr_op_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_char_p), f64p, u64, i32,
                                    ctypes.POINTER(vp)]),
    "qr_op_destroy": (ctypes.c_int, [vp]),
    "qr_op_create_device": (ctypes.c_int, [ctypes.POINTER(ctypes.c_char_p), f64p, u64, i32, vp,
                                           ctypes.POINTER(vp)]),
    "qr_op_tables": (ctypes.c_int, [vp, u64p, u32p, u64p, f64p, ctypes.POINTER(ctypes.c_uint8),
                                    u32p, i32p, u64p]),
    "qr_op_info": (ctypes.c_int, [vp, u32p, u64p, u64p, u64p]),
    "qr_op_terms": (ctypes.c_int, [vp, u64p, u64p, f64p]),
    "qr_op_row_layout": (ctypes.c_int, [vp, u64, u64p, u32p]),
    "qr_mode_parse": (ctypes.c_int, [ctypes.c_char_p, i32p, u64p, i32p]),
    "qr_csr_build": (ctypes.c_int, [vp, vp, vp, vp, vp]),
    "qr_csr_build_rows": (ctypes.c_int, [vp, u64, u64, vp, vp, vp]),
    "qr_csr_matvec": (ctypes.c_int, [u64, u64, u64, vp, vp, vp, vp, vp, vp]),
    "qr_pauli_apply": (ctypes.c_int, [vp, i32, u32, vp, vp, vp, vp]),
    "qr_pauli_apply_noalloc": (ctypes.c_int, [vp, i32, u32, vp, vp, vp, vp]),
    "qr_csr_diagonal": (ctypes.c_int, [u64, u64, u64, vp, vp, vp, vp, vp]),
    "qr_axpby": (ctypes.c_int, [u64, f64p, vp, f64p, vp, vp, vp]),
    "qr_axpy": (ctypes.c_int, [u64, f64p, vp, vp, vp, vp]),
    "qr_ax": (ctypes.c_int, [u64, f64p, vp, vp, vp]),
    "qr_precond": (ctypes.c_int, [u64, u64, u64, vp, vp, vp, vp, f64p, ctypes.c_double, vp, vp]),
    "qr_precond2": (ctypes.c_int, [u64, vp, vp, f64p, ctypes.c_double, vp, vp]),
    "qr_csr_add_count": (ctypes.c_int, [u64, f64p, u64, vp, vp, vp, f64p, u64, vp, vp, vp, vp,
                                        u64p, vp]),
    "qr_csr_add_fill": (ctypes.c_int, [u64, f64p, u64, vp, vp, vp, f64p, u64, vp, vp, vp, vp, vp,
                                       vp, vp]),
    "qr_csr_count_zeros": (ctypes.c_int, [u64, vp, f64, u64p, vp]),
    "qr_csr_eliminate_zeros": (ctypes.c_int, [u64, u64, vp, vp, vp, f64, vp, vp, vp, u64p, vp]),
    "qr_csr_write_raw": (ctypes.c_int, [ctypes.c_char_p, u64, u64, u64, vp, vp, vp, vp]),
    "qr_csr_read_raw_info": (ctypes.c_int, [ctypes.c_char_p, u64p, u64p, u64p]),
    "qr_csr_read_raw": (ctypes.c_int, [ctypes.c_char_p, u64, u64, vp, vp, vp, vp]),
    "qr_csr_write_mm": (ctypes.c_int, [ctypes.c_char_p, u64, u64, u64, vp, vp, vp, i32, vp]),
    "qr_mm_read_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(vp), u64p, u64p, u64p, vp]),
    "qr_mm_read_take": (ctypes.c_int, [vp, vp, vp, vp, vp]),
    "qr_mm_read_close": (ctypes.c_int, [vp]),
    "qr_batch_create": (ctypes.c_int, [ctypes.POINTER(vp), u32, ctypes.POINTER(vp)]),
    "qr_batch_destroy": (ctypes.c_int, [vp]),
    "qr_batch_run": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, vp, vp]),
    "qr_batch_tune": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, vp, vp]),
}


class QrTuning(ctypes.Structure):
    """qr_tuning of include/qrhip.h."""
    _fields_ = [("tiles_per_item", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32),
                ("order", ctypes.c_uint32), ("split", ctypes.c_uint32)]


class QrError(Exception):
    """Raised for every failure reported by libqrhip (the reference raises
    PyException from QrustyErr, qrusty/src/lib.rs:48-52)."""

    def __init__(self, message: str, code: int = -1):
        super().__init__(message)
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"qrusty_amd: {LIB_PATH} is not built; run `python qrusty_amd/build.py` "
            "(the HIP kernels have no CPU fallback)")
    try:
        L = ctypes.CDLL(LIB_PATH)
    except OSError as e:  # pragma: no cover - depends on the ROCm install
        raise ImportError(f"qrusty_amd: cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


lib = _load()


class QrIOError(QrError, OSError):
    """A file could not be opened, read or written (the reference raises
    OSError for these, pyqrusty/src/lib.rs:216-259)."""


QR_ERR_IO = -8
QR_MM_SYMMETRY = {"": 0, "general": 0, "symmetric": 1, "hermitian": 2, "skew-symmetric": 3}


def check(rc: int) -> None:
    if rc != QR_OK:
        msg = lib.qr_last_error().decode(errors="replace")
        raise (QrIOError if rc == QR_ERR_IO else QrError)(msg, rc)
