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

import math

from ._lib import DP, I32, OMHD_BC_MIRROR, OMHD_BC_PERIODIC, check, lib

# KH wall sign table, 2D_KH/bc.f90:17-36 and :61-81: my/vy and By negated.
KH_WALL_SIGN = (1.0, -1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, 1.0)


def _negated(*slots):
    s = [1.0] * 9
    for m in slots:
        s[m] = -1.0
    return tuple(s)


# 2D_reconnection/bc.f90: slots negated in the mirror ghosts, per face
# (x-, x+, y-, y+, z-, z+): conducting walls in x (mx/vx, mz/vz, By, psi),
# a symmetry plane at the south face (my/vy, mz/vz, Bx, psi), a wall at the
# north face (my/vy, By).
RECON_SIGN = (_negated(0, 2, 6, 8), _negated(0, 2, 6, 8), _negated(1, 2, 5, 8), _negated(1, 6), _negated(),
              _negated())

# resistivity of 2D_reconnection (flux_resistive.f90:5-7): 1/RM1 at the
# X-point, 1/RM0 in the background
RECON_RM1, RECON_RM0 = 60.0, 1000.0


def _fortran_strides(shape, strides) -> bool:
    """True if the element strides are those of a dense column-major array."""
    expect = 1
    for size, stride in zip(shape, strides):
        if size != 1 and stride != expect:
            return False
        expect *= size
    return True


def torch_ready(t) -> None:
    """Wait for the work torch has queued on `t`'s device.  libomhd runs on
    streams of its own that do not wait for torch's, so a tensor that a torch
    kernel is still writing must be complete before its pointer is handed over
    (every libomhd entry point returns with its own stream drained, so the
    other direction needs nothing)."""
    if t.is_cuda:
        import torch

        torch.cuda.current_stream(t.device).synchronize()


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
        torch_ready(a)
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


def bc_for_U_recon(U, ix, jx, kx=1):
    """bc_for_U(U,ix,jx) -- 2D_reconnection/bc.f90:6: reflecting faces with
    the per-face tables of RECON_SIGN; with kx > 1 the 3D form of
    3D_reconnection_gpu/bc.cuf:7, periodic in z."""
    _bc_u(U, ix, jx, kx, [OMHD_BC_MIRROR] * 4 + [OMHD_BC_PERIODIC] * 2, RECON_SIGN)


def bc_for_F_recon(VL, VR, ix, jx, kx=1):
    """bc_for_F(VL,VR,ix,jx) -- 2D_reconnection/bc.f90:54 (x walls);
    3D_reconnection_gpu/bc.cuf:103."""
    _face(VL, VR, ix, jx, kx, 1, OMHD_BC_MIRROR, RECON_SIGN[0], RECON_SIGN[1])


def bc_for_G_recon(VL, VR, ix, jx, kx=1):
    """bc_for_G(VL,VR,ix,jx) -- 2D_reconnection/bc.f90:86 (south symmetry
    plane, north wall); 3D_reconnection_gpu/bc.cuf:147."""
    _face(VL, VR, ix, jx, kx, 2, OMHD_BC_MIRROR, RECON_SIGN[2], RECON_SIGN[3])


def bc_for_H_recon(VL, VR, ix, jx, kx):
    """bc_for_H(VL,VR,ix,jx,kx) -- 3D_reconnection_gpu/bc.cuf:191 (periodic
    z patch)."""
    _face(VL, VR, ix, jx, kx, 3, OMHD_BC_PERIODIC, np.ones(9), np.ones(9))


def eta_recon(xx: float, yy: float) -> float:
    """eta(xx,yy) -- 2D_reconnection/flux_resistive.f90:1: 1/RM0 + (RM0-RM1)/
    (RM0 RM1) cosh(min(r, 25))^-2, evaluated with the host's libm.  The
    power -2 is formed as 1/(c*c), the way the reference's compiled build
    evaluates it, so the value is bitwise the reference's."""
    c = math.cosh(min(math.sqrt(xx * xx + yy * yy), 25.0))
    return 1.0 / RECON_RM0 + (RECON_RM0 - RECON_RM1) / (RECON_RM0 * RECON_RM1) * (1.0 / (c * c))


_eta_cache: dict = {}


def eta_face_table(x, y, dx, dir, eta=eta_recon, over_dx=True):
    """The (ix,jx) Fortran-ordered table of eta at the face centres of
    direction dir: eta(x(i)+dx/2, y(j)) on x-faces, eta(x(i), y(j)+dx/2) on
    y-faces, eta(x(i), y(j)) on z-faces.  over_dx: times 1/dx, as the 2D form
    forms eta/dx (flux_resistive.f90:74, :97); without it the plain values
    the 3D form divides by dx itself (flux_resistive.cuf:47-83).  The last
    table is cached."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    key = (x.tobytes(), y.tobytes(), float(dx), int(dir), eta, bool(over_dx))
    hit = _eta_cache.get("last")
    if hit is not None and hit[0] == key:
        return hit[1]
    odx = 1.0 / dx
    t = np.empty((len(x), len(y)), order="F")
    for j, yj in enumerate(y.tolist()):
        for i, xi in enumerate(x.tolist()):
            if dir == 1:
                v = eta(xi + 0.5 * dx, yj)
            elif dir == 2:
                v = eta(xi, yj + 0.5 * dx)
            else:
                v = eta(xi, yj)
            t[i, j] = v * odx if over_dx else v
    _eta_cache["last"] = (key, t)
    return t


def flux_resistive(F, U, x, y, dx, ix, jx, dir, eta=eta_recon):
    """flux_resistive(F,U,x,y,dx,ix,jx,dir) -- 2D_reconnection/
    flux_resistive.f90:27: adds the resistive terms to F(en), F(by)/F(bx) and
    F(bz) on the faces of direction dir (dir 3: nothing).  eta(xx, yy) is the
    resistivity profile, evaluated on the host (default: that file's)."""
    if dir not in (1, 2):
        return
    n = ix * jx
    if len(x) != ix or len(y) != jx:
        raise ValueError(f"x, y have {len(x)}, {len(y)} points, expected {ix}, {jx}")
    t = eta_face_table(x, y, dx, dir, eta)
    check(lib.omhd_flux_resistive(_ptr(F, 9 * n), _ptr(U, 9 * n), t.ctypes.data_as(DP), ix, jx, dir),
          "flux_resistive")


def flux_resistive_3d(F, U, x, y, z, dx, ix, jx, kx, dir, eta=eta_recon):
    """flux_resistive(F,U,x,y,z,dx,ix,jx,kx,dir) -- 3D_reconnection_gpu/
    flux_resistive.cuf:12, as a host call: adds the resistive terms to F(en)
    and the two transverse B slots on the faces of direction dir.  eta(xx, yy)
    is the resistivity profile (default: that file's, which does not vary in
    z), evaluated on the host."""
    if dir not in (1, 2, 3):
        return
    n = ix * jx * kx
    if len(x) != ix or len(y) != jx or len(z) != kx:
        raise ValueError(f"x, y, z have {len(x)}, {len(y)}, {len(z)} points, expected {ix}, {jx}, {kx}")
    t = eta_face_table(x, y, dx, dir, eta, over_dx=False)
    check(lib.omhd_flux_resistive3d(_ptr(F, 9 * n), _ptr(U, 9 * n), t.ctypes.data_as(DP), 1, dx, ix, jx, kx, dir),
          "flux_resistive")


def set_dt_resistive(dt, dx, cfl, rm1=RECON_RM1):
    """set_dt_resistive(dt,dx,cfl) -- 2D_reconnection/flux_resistive.f90:18
    (3D_reconnection_gpu/flux_resistive.cuf:2 alike):
    returns min(dt, 0.5 cfl Rm1 dx^2), the explicit diffusion limit at the
    resistivity peak."""
    return min(dt, 0.5 * cfl * rm1 * (dx * dx))


def periodic_patch(VL, VR, ix, jx, dir, kx=1):
    """The inline periodic face patch of 2D_basic/main.f90:101-102 (dir 1)
    and :119-120 (dir 2)."""
    _face(VL, VR, ix, jx, kx, dir, OMHD_BC_PERIODIC, np.ones(9), np.ones(9))
