"""Resident-state solver: the state lives in HBM and whole TVD-RK2 steps run as
two fused kernel launches each (libomhd's omhd_create / omhd_upload /
omhd_step / omhd_download).  This is the fast path; the step sequence it
implements is 2D_basic/main.f90:56-188 (2D_KH/main.f90 with wall boundaries).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DP, I32, OMHD_BC_MIRROR, OMHD_BC_PERIODIC, Config, OMHDError, check, lib
from .api import KH_WALL_SIGN, _ptr, torch_ready


def _cptr(a, n_expected):
    """Pointer of a C-ordered float64 array (numpy, or torch on host/device)."""
    if hasattr(a, "data_ptr"):
        import torch

        if a.dtype != torch.float64 or not a.is_contiguous():
            raise TypeError("contiguous float64 tensor required")
        if a.numel() != n_expected:
            raise ValueError(f"tensor has {a.numel()} elements, expected {n_expected}")
        torch_ready(a)
        return C.cast(C.c_void_p(a.data_ptr()), DP)
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("C-contiguous float64 numpy array required")
    if a.size != n_expected:
        raise ValueError(f"array has {a.size} elements, expected {n_expected}")
    return a.ctypes.data_as(DP)


def make_config(nx, ny, dx, nz=0, cfl=0.4, lm_type=2, flux_type=3, time_type=0, bc=("periodic", "periodic"),
                wall_sign=KH_WALL_SIGN, rows_per_wave=0, resistivity=None, origin=(0.0, 0.0)):
    """omhd_config for a grid (see Solver)."""
    cfg = Config()
    cfg.ndim = 3 if nz else 2
    cfg.nx, cfg.ny, cfg.nz = nx, ny, max(nz, 1)
    cfg.dx, cfg.cfl = dx, cfl
    cfg.lm_type, cfg.flux_type, cfg.time_type = lm_type, flux_type, time_type
    per_axis = len(wall_sign) == 3 and hasattr(wall_sign[0], "__len__")
    per_face = len(wall_sign) == 6 and hasattr(wall_sign[0], "__len__")
    for a in range(3):
        kind = bc[a] if a < len(bc) else "periodic"
        for s in range(2):
            f = 2 * a + s
            sg = wall_sign[f] if per_face else (wall_sign[a] if per_axis else wall_sign)
            cfg.bc_type[f] = OMHD_BC_MIRROR if kind == "wall" else OMHD_BC_PERIODIC
            for m in range(9):
                cfg.bc_sign[9 * f + m] = sg[m] if kind == "wall" else 1.0
    cfg.rows_per_wave = rows_per_wave
    cfg.x0, cfg.y0 = origin
    if resistivity is not None:
        cfg.resistive = 1
        cfg.eta0, cfg.eta1, cfg.eta_w = resistivity
    return cfg


class Solver:
    """A grid of nx*ny (2D) or nx*ny*nz (3D, nz > 0) interior cells advanced
    on the GPU.

    bc: per axis ('periodic' | 'wall'); wall_sign: 9 signs (+1/-1) applied to
    the mirrored ghost (default: the KH walls of 2D_KH/bc.f90), or a list of
    one 9-vector per axis, or of one per face (x-, x+, y-, y+, z-, z+; e.g.
    the 2D_reconnection/bc.f90 tables).  resistivity: (eta0, eta1, w) for the 3D resistive
    step with eta = eta0 + eta1 exp(-(x^2+y^2)/w^2) at the face centres;
    origin: (x0, y0), the lower corner of the domain.
    """

    def __init__(self, nx, ny, dx, nz=0, cfl=0.4, lm_type=2, flux_type=3, time_type=0, bc=("periodic", "periodic"),
                 wall_sign=KH_WALL_SIGN, rows_per_wave=0, resistivity=None, origin=(0.0, 0.0)):
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, dx
        self.ndim = 3 if nz else 2
        cfg = make_config(nx, ny, dx, nz, cfl, lm_type, flux_type, time_type, bc, wall_sign, rows_per_wave,
                          resistivity, origin)
        self.cfg = cfg
        h = C.c_void_p()
        check(lib.omhd_create(C.byref(cfg), C.byref(h)), "omhd_create")
        self.h = h
        self.t = 0.0

    # -- state transfer -----------------------------------------------------
    def _shape(self):
        s = (self.nx + 2, self.ny + 2)
        return s + (self.nz + 2,) if self.ndim == 3 else s

    def upload(self, U):
        """U: (nx+2, ny+2[, nz+2], 9) Fortran-ordered (ghosts ignored), host or device."""
        check(lib.omhd_upload(self.h, _ptr(U, 9 * int(np.prod(self._shape())))), "omhd_upload")

    def download(self, with_v=False):
        shp = self._shape()
        U = np.zeros(shp + (9,), order="F")
        V = np.zeros(shp + (4,), order="F") if with_v else None
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
        try:
            check(rc, f"omhd_step (after {done.value} steps)")
        except OMHDError as e:
            e.steps_done = done.value  # steps completed before the stop (the reference's stop)
            raise
        return (dts, chs) if trace else None

    def state_shape(self):
        """Shape of the interior state W: (9, [nz,] ny, nx), x fastest."""
        return (9, self.nz, self.ny, self.nx) if self.ndim == 3 else (9, self.ny, self.nx)

    def set_state(self, W):
        """Load W (state_shape(), C order: interior cells, slots of param.h)
        from host (numpy) or device (torch) memory; computes the CFL speed."""
        check(lib.omhd_set_state(self.h, _cptr(W, int(np.prod(self.state_shape())))), "omhd_set_state")

    def get_state(self, out=None):
        """The interior state as W (state_shape()); into `out` (numpy or torch,
        host or device) when given, else a new numpy array."""
        W = np.zeros(self.state_shape()) if out is None else out
        check(lib.omhd_get_state(self.h, _cptr(W, int(np.prod(self.state_shape())))), "omhd_get_state")
        return W

    def step_fixed(self, dt, ch):
        """Advance len(dt) steps with the given dt and ch per step (set_dt
        replaced by the sequence: the frozen-dt map the adjoint differentiates)."""
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        ch = np.ascontiguousarray(ch, dtype=np.float64)
        if dt.shape != ch.shape or dt.ndim != 1:
            raise ValueError("dt and ch: equal-length 1-D sequences")
        check(lib.omhd_step_fixed(self.h, len(dt), dt.ctypes.data_as(DP), ch.ctypes.data_as(DP)), "omhd_step_fixed")
        self.t += float(dt.sum())

    def set_resistivity(self, eta0, eta1, w):
        """New resistivity parameters (handles created with resistivity=...)."""
        check(lib.omhd_set_resistivity(self.h, eta0, eta1, w), "omhd_set_resistivity")

    def layout(self):
        p, plane, pitch = DP(), C.c_int64(), I32()
        check(lib.omhd_state_layout(self.h, C.byref(p), C.byref(plane), C.byref(pitch)), "omhd_state_layout")
        return C.cast(p, C.c_void_p).value, plane.value, pitch.value

    def set_timing(self, enable=True):
        check(lib.omhd_set_timing(self.h, 1 if enable else 0), "omhd_set_timing")

    def set_mode(self, device_dt=False, graph=False):
        """device_dt=True: each step's dt, CFL speed and GLM factor are computed
        on the GPU after its second stage (the GLM factor from a table of the
        host's exp values), so step(n) queues all n steps and waits once;
        graph=True also replays one captured hipGraph per step (no per-stage
        timing then).  Results, t, traces and stops equal the default host-dt
        mode bit for bit.  Ideal MHD only; needs a state."""
        check(lib.omhd_set_mode(self.h, 1 if device_dt else 0, 1 if graph else 0), "omhd_set_mode")

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


Solver2D = Solver  # the 2D form of the same class (nz = 0)


class Ensemble:
    """n independent 2D grids advanced together (omhd_batch_*): one persistent
    launch per RK stage over a tile table of all members, each member with its
    own CFL dt.  configs: a list of omhd_config (make_config(...)); members share
    the limiter and time type.  device_dt=False: dt and the GLM factor from
    the host each step; device_dt=True (+ graph=True): computed on the GPU (the
    GLM factor from a per-member table of the host's exp values), whole steps
    captured in a hipGraph.  Either way every member is bitwise equal to its
    solo Solver run."""

    def __init__(self, configs):
        self.n = len(configs)
        arr = (Config * self.n)(*configs)
        self.shapes = [(c.nx + 2, c.ny + 2) for c in configs]
        h = C.c_void_p()
        check(lib.omhd_batch_create(arr, self.n, C.byref(h)), "omhd_batch_create")
        self.h = h

    def upload(self, i, U):
        ix, jx = self.shapes[i]
        check(lib.omhd_batch_upload(self.h, i, _ptr(U, 9 * ix * jx)), "omhd_batch_upload")

    def download(self, i):
        ix, jx = self.shapes[i]
        U = np.zeros((ix, jx, 9), order="F")
        check(lib.omhd_batch_download(self.h, i, _ptr(U), None), "omhd_batch_download")
        return U

    def set_mode(self, device_dt=False, graph=False):
        check(lib.omhd_batch_set_mode(self.h, 1 if device_dt else 0, 1 if graph else 0), "omhd_batch_set_mode")

    def step(self, nsteps=1, trace=False):
        """Returns the (nsteps, n) dt trace if trace."""
        tr = np.zeros((nsteps, self.n)) if trace else None
        check(lib.omhd_batch_step(self.h, nsteps, tr.ctypes.data_as(DP) if trace else None), "omhd_batch_step")
        return tr

    def times(self):
        t = np.zeros(self.n)
        check(lib.omhd_batch_time(self.h, t.ctypes.data_as(DP)), "omhd_batch_time")
        return t

    def autotune(self, trial_steps=3):
        r, g = I32(), I32()
        check(lib.omhd_batch_autotune(self.h, trial_steps, C.byref(r), C.byref(g)), "omhd_batch_autotune")
        return r.value, g.value

    def close(self):
        if getattr(self, "h", None):
            lib.omhd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
