"""Python mirror of the reference's step subroutines, backed by libomhd.

Each function keeps the name, argument order and meaning of the Fortran
subroutine it replaces (file:line in the docstring); arrays are Fortran-ordered
float64 numpy arrays (host; staged through HBM per call) or torch CUDA tensors
laid out the same way (used in place).  Where the reference prints and calls
STOP, these raise OMHDError carrying the cell and value.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DP, I32, OMHD_BC_MIRROR, OMHD_BC_PERIODIC, check, lib

# KH wall sign table, 2D_KH/bc.f90:17-36 and :61-81: my/vy and By negated.
KH_WALL_SIGN = (1.0, -1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, 1.0)


def _fortran_strides(shape, strides) -> bool:
    """True if the element strides are those of a dense column-major array."""
    expect = 1
    for size, stride in zip(shape, strides):
        if size != 1 and stride != expect:
            return False
        expect *= size
    return True


def _ptr(a, n_expected: int | None = None):
    """Device or host pointer of a float64 array laid out in Fortran order."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):  # torch tensor
        import torch

        if a.dtype != torch.float64:
            raise TypeError("float64 tensor required")
        if not _fortran_strides(a.shape, a.stride()):
            raise ValueError("Fortran-ordered (column-major) contiguous tensor required")
        if n_expected is not None and a.numel() < n_expected:
            raise ValueError(f"tensor has {a.numel()} elements, expected {n_expected}")
        return C.cast(C.c_void_p(a.data_ptr()), DP)
    if not isinstance(a, np.ndarray) or a.dtype != np.float64:
        raise TypeError("float64 numpy array required")
    if not (a.flags.f_contiguous or (a.ndim == 1 and a.flags.c_contiguous)):
        raise ValueError("Fortran-ordered (column-major) array required")
    if n_expected is not None and a.size < n_expected:
        raise ValueError(f"array has {a.size} elements, expected {n_expected}")
    return a.ctypes.data_as(DP)


def u2v(U, V, ix, jx, kx=1):
    """u2v(U,V,ix,jx) -- 2D_basic/u2v.f90:1 (3D: 3D_basic/u2v.f90:1)."""
    n = ix * jx * kx
    check(lib.omhd_u2v(_ptr(U, 9 * n), _ptr(V, 4 * n), ix, jx, kx), "u2v")


def limiter(wk, wL, wR, ix, jx, dir, type, kx=1):
    """limiter(wk,wL,wR,ix,jx,dir,type) -- 2D_basic/limiter.f90:1."""
    n = ix * jx * kx
    check(lib.omhd_limiter(_ptr(wk, n), _ptr(wL, n), _ptr(wR, n), ix, jx, kx, dir, type), "limiter")


def flux_solver(F, VL, VR, ch, ix, jx, dir, type, kx=1):
    """flux_solver(F,VL,VR,ch,ix,jx,dir,type) -- 2D_basic/flux_solver.f90:1."""
    n = ix * jx * kx
    check(lib.omhd_flux_solver(_ptr(F, 9 * n), _ptr(VL, 9 * n), _ptr(VR, 9 * n), ch, ix, jx, kx, dir, type),
          "flux_solver")


def glm_ss2(U, ch, dt, ix, jx, kx=1):
    """glm_ss2(U,ch,dt,ix,jx) -- 2D_basic/glm_ss2.f90:1."""
    check(lib.omhd_glm_ss2(_ptr(U, 9 * ix * jx * kx), ch, dt, ix, jx, kx), "glm_ss2")


def _rk(mode, wk, wk1, wF, wG, dt, dx, ix, jx, wH, kx):
    n = 9 * ix * jx * kx
    check(lib.omhd_rk(mode, _ptr(wk, n), _ptr(wk1, n), _ptr(wF, n), _ptr(wG, n), _ptr(wH, n), dt, dx, ix, jx, kx),
          "rk")


def rk_tvd21(wk, wk1, wF, wG, dt, dx, ix, jx, wH=None, kx=1):
    """rk_tvd21(wk,wk1,wF,wG,dt,dx,ix,jx) -- 2D_basic/rk.f90:9."""
    _rk(0, wk, wk1, wF, wG, dt, dx, ix, jx, wH, kx)


def rk_tvd22(wk, wk1, wF, wG, dt, dx, ix, jx, wH=None, kx=1):
    """rk_tvd22(wk,wk1,wF,wG,dt,dx,ix,jx) -- 2D_basic/rk.f90:42."""
    _rk(1, wk, wk1, wF, wG, dt, dx, ix, jx, wH, kx)


def rk_std21(wk, wk1, wF, wG, dt, dx, ix, jx, wH=None, kx=1):
    """rk_std21(wk,wk1,wF,wG,dt,dx,ix,jx) -- 2D_basic/rk.f90:73."""
    _rk(2, wk, wk1, wF, wG, dt, dx, ix, jx, wH, kx)


def rk_std22(wk, wF, wG, dt, dx, ix, jx, wH=None, kx=1):
    """rk_std22(wk,wF,wG,dt,dx,ix,jx) -- 2D_basic/rk.f90:106."""
    _rk(3, wk, None, wF, wG, dt, dx, ix, jx, wH, kx)


def set_dt(U, V, dx, cfl, ix, jx, kx=1, ndim=2):
    """set_dt(U,V,vmax,dt,dx,cfl,ix,jx) -- 2D_basic/set_dt.f90:1.
    Returns (vmax, dt), the reference's two output arguments."""
    vmax, dt = C.c_double(), C.c_double()
    n = ix * jx * kx
    check(lib.omhd_set_dt(_ptr(U, 9 * n), _ptr(V, 4 * n), C.byref(vmax), C.byref(dt), dx, cfl, ix, jx, kx, ndim),
          "set_dt")
    return vmax.value, dt.value


def _bc_u(U, ix, jx, kx, types, signs):
    t = (I32 * 6)(*types)
    s = np.ascontiguousarray(np.asarray(signs, dtype=np.float64).reshape(54))
    check(lib.omhd_bc_u(_ptr(U, 9 * ix * jx * kx), ix, jx, kx, t, s.ctypes.data_as(DP)), "bc")


def bc(U, ix, jx, kx=1):
    """bc(U,ix,jx) -- 2D_basic/bc.f90:1 (periodic on every face)."""
    _bc_u(U, ix, jx, kx, [OMHD_BC_PERIODIC] * 6, np.ones(54))


def _kh_signs():
    s = np.ones((6, 9))
    s[2] = KH_WALL_SIGN
    s[3] = KH_WALL_SIGN
    return s


def bc_for_U(U, ix, jx):
    """bc_for_U(U,ix,jx) -- 2D_KH/bc.f90:6: periodic x, walls in y."""
    _bc_u(U, ix, jx, 1, [OMHD_BC_PERIODIC, OMHD_BC_PERIODIC, OMHD_BC_MIRROR, OMHD_BC_MIRROR, 0, 0], _kh_signs())


def _face(VL, VR, ix, jx, kx, dir, typ, slo, shi):
    n = 9 * ix * jx * kx
    a = np.ascontiguousarray(slo, dtype=np.float64)
    b = np.ascontiguousarray(shi, dtype=np.float64)
    check(lib.omhd_bc_face(_ptr(VL, n), _ptr(VR, n), ix, jx, kx, dir, typ, a.ctypes.data_as(DP),
                           b.ctypes.data_as(DP)), "bc_face")


def bc_for_F(VL, VR, ix, jx):
    """bc_for_F(VL,VR,ix,jx) -- 2D_KH/bc.f90:42 (periodic x face patch)."""
    _face(VL, VR, ix, jx, 1, 1, OMHD_BC_PERIODIC, np.ones(9), np.ones(9))


def bc_for_G(VL, VR, ix, jx):
    """bc_for_G(VL,VR,ix,jx) -- 2D_KH/bc.f90:55 (wall face patch in y)."""
    _face(VL, VR, ix, jx, 1, 2, OMHD_BC_MIRROR, KH_WALL_SIGN, KH_WALL_SIGN)


def periodic_patch(VL, VR, ix, jx, dir, kx=1):
    """The inline periodic face patch of 2D_basic/main.f90:101-102 (dir 1)
    and :119-120 (dir 2)."""
    _face(VL, VR, ix, jx, kx, dir, OMHD_BC_PERIODIC, np.ones(9), np.ones(9))
