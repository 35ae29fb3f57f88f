"""CPU checks of the boundary: libomhd.so loads (no GPU needed for that) and
exports every function include/omhd.h declares; the Python mirror exposes the
reference subroutine names."""
import ctypes as C
import re
from pathlib import Path

REPO = Path(__file__).resolve().parents[1]


def declared(header):
    txt = (REPO / "include" / header).read_text()
    return sorted(set(re.findall(r"\b(omhd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(REPO / "openmhd_amd" / "libomhd.so"))
    names = declared("omhd.h")
    assert len(names) >= 17
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string():
    lib = C.CDLL(str(REPO / "openmhd_amd" / "libomhd.so"))
    lib.omhd_version.restype = C.c_char_p
    assert b"gfx950" in lib.omhd_version()


def test_python_mirror_names():
    from openmhd_amd import api

    for name in ("u2v", "limiter", "flux_solver", "glm_ss2", "rk_tvd21", "rk_tvd22", "rk_std21", "rk_std22",
                 "set_dt", "bc", "bc_for_U", "bc_for_F", "bc_for_G"):
        assert callable(getattr(api, name)), name


def test_errors_without_gpu_are_reported_not_crashing():
    """Argument validation happens before any device work."""
    from openmhd_amd import OMHDError, api
    import numpy as np
    import pytest

    with pytest.raises(OMHDError) as e:
        api.limiter(np.zeros((4, 4), order="F"), np.zeros((4, 4), order="F"), np.zeros((4, 4), order="F"), 4, 4, 1, 9)
    assert e.value.code == 5  # OMHD_ERR_TYPE: the reference's 'unknown limiter' stop


def test_row_major_tensor_is_rejected():
    """A tensor that is not laid out column-major is refused before any call
    into the library (the arrays are the reference's Fortran arrays)."""
    import numpy as np
    import pytest
    import torch

    from openmhd_amd import api

    U = torch.zeros((6, 5, 9), dtype=torch.float64)  # C order: strides (45, 9, 1)
    V = torch.from_numpy(np.zeros((6, 5, 4), order="F"))
    with pytest.raises(ValueError):
        api.u2v(U, V, 6, 5)
    with pytest.raises(ValueError):
        api.u2v(torch.from_numpy(np.zeros((6, 5, 9), order="F"))[:, ::2], V, 6, 5)  # a strided view
