"""Resident-state solver: the state lives in HBM and whole TVD-RK2 steps run as
two fused kernel launches each (libomhd's omhd_create / omhd_upload /
omhd_step / omhd_download).  This is the fast path; the step sequence it
implements is 2D_basic/main.f90:56-188 (2D_KH/main.f90 with wall boundaries).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DP, I32, OMHD_BC_MIRROR, OMHD_BC_PERIODIC, Config, check, lib
from .api import KH_WALL_SIGN, _ptr


class Solver2D:
    """A 2D grid of nx*ny interior cells advanced on the GPU.

    bc: per axis ('periodic' | 'wall'); wall_sign: 9 signs (+1/-1) applied to
    the mirrored ghost (default: the KH walls of 2D_KH/bc.f90).
    """

    def __init__(self, nx, ny, dx, cfl=0.4, lm_type=2, flux_type=3, time_type=0, bc=("periodic", "periodic"),
                 wall_sign=KH_WALL_SIGN, rows_per_wave=0, store_hint=0):
        self.nx, self.ny, self.dx = nx, ny, dx
        cfg = Config()
        cfg.ndim, cfg.nx, cfg.ny, cfg.nz = 2, nx, ny, 1
        cfg.dx, cfg.cfl = dx, cfl
        cfg.lm_type, cfg.flux_type, cfg.time_type = lm_type, flux_type, time_type
        for a in range(3):
            kind = bc[a] if a < len(bc) else "periodic"
            for s in range(2):
                f = 2 * a + s
                cfg.bc_type[f] = OMHD_BC_MIRROR if kind == "wall" else OMHD_BC_PERIODIC
                for m in range(9):
                    cfg.bc_sign[9 * f + m] = wall_sign[m] if kind == "wall" else 1.0
        cfg.rows_per_wave = rows_per_wave
        cfg.store_hint = store_hint
        self.cfg = cfg
        h = C.c_void_p()
        check(lib.omhd_create(C.byref(cfg), C.byref(h)), "omhd_create")
        self.h = h
        self.t = 0.0

    # -- state transfer -----------------------------------------------------
    def upload(self, U):
        """U: (nx+2, ny+2, 9) Fortran-ordered (ghosts ignored), host or device."""
        check(lib.omhd_upload(self.h, _ptr(U, 9 * (self.nx + 2) * (self.ny + 2))), "omhd_upload")

    def download(self, with_v=False):
        ix, jx = self.nx + 2, self.ny + 2
        U = np.zeros((ix, jx, 9), order="F")
        V = np.zeros((ix, jx, 4), order="F") if with_v else None
        check(lib.omhd_download(self.h, _ptr(U), _ptr(V) if with_v else None), "omhd_download")
        return (U, V) if with_v else U

    def step(self, nsteps=1, trace=False):
        """Advance nsteps; returns (dt_trace, ch_trace) if trace else None."""
        t = C.c_double(self.t)
        dts = np.zeros(nsteps) if trace else None
        chs = np.zeros(nsteps) if trace else None
        done = I32(0)
        rc = lib.omhd_step(self.h, nsteps, C.byref(t), dts.ctypes.data_as(DP) if trace else None,
                           chs.ctypes.data_as(DP) if trace else None, C.byref(done))
        self.t = t.value
        check(rc, f"omhd_step (after {done.value} steps)")
        return (dts, chs) if trace else None

    def layout(self):
        p, plane, pitch = DP(), C.c_int64(), I32()
        check(lib.omhd_state_layout(self.h, C.byref(p), C.byref(plane), C.byref(pitch)), "omhd_state_layout")
        return C.cast(p, C.c_void_p).value, plane.value, pitch.value

    def set_timing(self, enable=True):
        check(lib.omhd_set_timing(self.h, 1 if enable else 0), "omhd_set_timing")

    def timing(self):
        """(stage_ms[2], launches[2]) accumulated since set_timing(True)."""
        ms = (C.c_double * 2)()
        n = (C.c_int64 * 2)()
        check(lib.omhd_get_timing(self.h, ms, n), "omhd_get_timing")
        return list(ms), list(n)

    def stream(self):
        return lib.omhd_stream(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib.omhd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
