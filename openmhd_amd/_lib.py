"""ctypes declarations of libomhd's C ABI (include/omhd.h).

The library is built in-tree (openmhd_amd/libomhd.so, by __graft_entry__.build()
or `make -C openmhd_amd`).  There is no fallback: if it is missing, importing
the package raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("OMHD_LIB", HERE / "libomhd.so"))

OMHD_OK = 0
OMHD_ERR_PRESSURE = 1
OMHD_ERR_DENSITY = 2
OMHD_ERR_DT = 3
OMHD_ERR_DIR = 4
OMHD_ERR_TYPE = 5
OMHD_ERR_EXTENT = 6
OMHD_ERR_HIP = 7
OMHD_ERR_ARG = 8

OMHD_BC_PERIODIC = 0
OMHD_BC_MIRROR = 1

DP = C.POINTER(C.c_double)
I32 = C.c_int32


class ErrorInfo(C.Structure):
    _fields_ = [("code", C.c_int32), ("i", C.c_int32), ("j", C.c_int32), ("k", C.c_int32),
                ("value", C.c_double), ("msg", C.c_char * 160)]


class Config(C.Structure):
    _fields_ = [
        ("ndim", I32), ("nx", I32), ("ny", I32), ("nz", I32),
        ("dx", C.c_double), ("cfl", C.c_double),
        ("lm_type", I32), ("flux_type", I32), ("time_type", I32),
        ("bc_type", I32 * 6), ("bc_sign", C.c_double * 54),
        ("rows_per_wave", I32), ("resistive", I32),
        ("x0", C.c_double), ("y0", C.c_double), ("eta0", C.c_double), ("eta1", C.c_double), ("eta_w", C.c_double),
        ("reserved", I32 * 8),
    ]


class OMHDError(RuntimeError):
    """A non-zero status from libomhd, with the cell and value behind it."""

    def __init__(self, code, info: ErrorInfo | None = None, where: str = ""):
        self.code = code
        self.cell = (info.i, info.j, info.k) if info is not None else None
        self.value = info.value if info is not None else None
        msg = info.msg.decode(errors="replace") if info is not None else ""
        super().__init__(f"{where}: omhd error {code} ({msg}) at cell {self.cell} value {self.value}")


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"libomhd is not built: {LIB_PATH} missing (run __graft_entry__.build())")
    # One HIP runtime per process: PyTorch ships its own libamdhip64 with the
    # same soname (libamdhip64.so.7) as ROCm's.  Whichever is loaded first
    # serves both, and PyTorch's build does not start on ROCm's newer copy, so
    # load PyTorch's (when installed) before libomhd binds to one.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    sig = {
        "omhd_last_error": [C.POINTER(ErrorInfo)],
        "omhd_u2v": [DP, DP, I32, I32, I32],
        "omhd_limiter": [DP, DP, DP, I32, I32, I32, I32, I32],
        "omhd_flux_solver": [DP, DP, DP, C.c_double, I32, I32, I32, I32, I32],
        "omhd_glm_ss2": [DP, C.c_double, C.c_double, I32, I32, I32],
        "omhd_rk": [I32, DP, DP, DP, DP, DP, C.c_double, C.c_double, I32, I32, I32],
        "omhd_set_dt": [DP, DP, DP, DP, C.c_double, C.c_double, I32, I32, I32, I32],
        "omhd_bc_u": [DP, I32, I32, I32, C.POINTER(I32), DP],
        "omhd_bc_face": [DP, DP, I32, I32, I32, I32, I32, DP, DP],
        "omhd_flux_resistive": [DP, DP, DP, I32, I32, I32],
        "omhd_flux_resistive3d": [DP, DP, DP, I32, C.c_double, I32, I32, I32, I32],
        "omhd_create": [C.POINTER(Config), C.POINTER(C.c_void_p)],
        "omhd_upload": [C.c_void_p, DP],
        "omhd_step": [C.c_void_p, I32, DP, DP, DP, C.POINTER(I32)],
        "omhd_download": [C.c_void_p, DP, DP],
        "omhd_state_layout": [C.c_void_p, C.POINTER(DP), C.POINTER(C.c_int64), C.POINTER(I32)],
        "omhd_destroy": [C.c_void_p],
        "omhd_set_timing": [C.c_void_p, I32],
        "omhd_set_mode": [C.c_void_p, I32, I32],
        "omhd_batch_create": [C.POINTER(Config), I32, C.POINTER(C.c_void_p)],
        "omhd_batch_upload": [C.c_void_p, I32, DP],
        "omhd_batch_download": [C.c_void_p, I32, DP, DP],
        "omhd_batch_set_mode": [C.c_void_p, I32, I32],
        "omhd_batch_step": [C.c_void_p, I32, DP],
        "omhd_batch_time": [C.c_void_p, DP],
        "omhd_batch_autotune": [C.c_void_p, I32, C.POINTER(I32), C.POINTER(I32)],
        "omhd_batch_destroy": [C.c_void_p],
        "omhd_get_timing": [C.c_void_p, DP, C.POINTER(C.c_int64)],
        "omhd_step_fixed": [C.c_void_p, I32, DP, DP],
        "omhd_set_state": [C.c_void_p, DP],
        "omhd_get_state": [C.c_void_p, DP],
        "omhd_adjoint_create": [C.c_void_p, I32, I32, C.POINTER(C.c_void_p)],
        "omhd_adjoint_forward": [C.c_void_p, DP, DP, DP, DP],
        "omhd_adjoint_backward": [C.c_void_p, DP, DP],
        "omhd_adjoint_destroy": [C.c_void_p],
        "omhd_adjoint_eta_grad": [C.c_void_p, DP],
        "omhd_set_resistivity": [C.c_void_p, C.c_double, C.c_double, C.c_double],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.omhd_stream.argtypes = [C.c_void_p]
    L.omhd_stream.restype = C.c_void_p
    L.omhd_version.argtypes = []
    L.omhd_version.restype = C.c_char_p
    return L


lib = _load()


def check(rc: int, where: str = "") -> None:
    if rc != OMHD_OK:
        info = ErrorInfo()
        lib.omhd_last_error(C.byref(info))
        raise OMHDError(rc, info, where)
