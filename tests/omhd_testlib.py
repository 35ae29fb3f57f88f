"""ctypes bindings used by the tests: the oracle (C restatement), the
flang-built reference routines (oracle/_ref, present only where they were
built) and random-state generators.

All arrays are numpy float64 in Fortran order with the reference's ghost layer,
e.g. U[ix, jx, 9] (2D) or U[ix, jx, kx, 9] (3D); slot order follows param.h:6-15
(mx/vx, my/vy, mz/vz, en/pr, ro, bx, by, bz, ps).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[1]
# OMHD_SANITIZED=1: the AddressSanitizer + UndefinedBehaviorSanitizer builds of
# the oracle and of the native checkers (oracle/run_sanitized.sh)
SANITIZED = os.environ.get("OMHD_SANITIZED") == "1"
ORACLE_SO = REPO / "oracle" / ("_san/liboracle.so" if SANITIZED else "liboracle.so")
NATIVE_DIR = REPO / "tests" / "native" / ("_san" if SANITIZED else "")
REF_DIR = REPO / "oracle" / "_ref"

GAMMA = 5.0 / 3.0
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)
LP = C.POINTER(C.c_long)

# KH walls, 2D_KH/bc.f90:17-36 (U) and :61-81 (VL/VR): my/vy and By negated.
KH_SIGN = np.array([1, -1, 1, 1, 1, 1, -1, 1, 1], dtype=np.float64)


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous), "need contiguous float64"
    return a.ctypes.data_as(DP)


def farr(*shape) -> np.ndarray:
    return np.zeros(shape, dtype=np.float64, order="F")


# ---------------------------------------------------------------------------
# oracle
# ---------------------------------------------------------------------------
class OrProblem(C.Structure):
    _fields_ = [
        ("ix", C.c_int), ("jx", C.c_int), ("kx", C.c_int),
        ("dx", C.c_double), ("cfl", C.c_double),
        ("lm_type", C.c_int), ("flux_type", C.c_int), ("time_type", C.c_int),
        ("bctype", C.c_int * 6),
        ("u_sgn", C.c_double * 54), ("face_sgn", C.c_double * 54),
        ("dt_fixed", C.c_double), ("ch_fixed", C.c_double),
        ("eta", DP), ("eta_max", C.c_double),
        ("res_form", C.c_int), ("rm1", C.c_double),
    ]


def make_problem(ix, jx, kx=1, dx=1.0, cfl=0.4, lm_type=2, flux_type=3, time_type=0, bc=("periodic",) * 3,
                 signs=None):
    """bc: per direction 'periodic' or 'wall' (mirror with a sign table).
    signs: per direction a 9-vector of +-1 for the mirror (default KH walls),
    or 6 9-vectors, one per face (x-, x+, y-, y+, z-, z+)."""
    p = OrProblem()
    p.ix, p.jx, p.kx = ix, jx, kx
    p.dx, p.cfl = dx, cfl
    p.lm_type, p.flux_type, p.time_type = lm_type, flux_type, time_type
    per_face = signs is not None and len(signs) == 6
    for d in range(3):
        wall = bc[d] == "wall"
        for s in range(2):
            if per_face:
                sg = np.asarray(signs[2 * d + s], dtype=np.float64)
            else:
                sg = KH_SIGN if signs is None or signs[d] is None else np.asarray(signs[d], dtype=np.float64)
            p.bctype[2 * d + s] = 1 if wall else 0
            for m in range(9):
                p.u_sgn[(2 * d + s) * 9 + m] = sg[m] if wall else 1.0
                p.face_sgn[(2 * d + s) * 9 + m] = sg[m] if wall else 1.0
    p.dt_fixed = 0.0
    p.ch_fixed = 0.0
    return p


class Oracle:
    def __init__(self, path: Path = ORACLE_SO):
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        L = self.L = C.CDLL(str(path))
        L.or_u2v.argtypes = [DP, DP, C.c_int, C.c_int, C.c_int, LP]
        L.or_limiter.argtypes = [DP, DP, DP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_hlld.argtypes = [DP, DP, C.c_double, C.c_int, DP, IP]
        L.or_flux_solver.argtypes = [DP, DP, DP, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, IP]
        L.or_glm_ss2.argtypes = [DP, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]
        L.or_glm_ss2.restype = None
        L.or_rk.argtypes = [C.c_int, DP, DP, DP, DP, DP, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]
        L.or_rk.restype = None
        L.or_set_dt.argtypes = [DP, DP, DP, DP, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, LP]
        L.or_bc_u.argtypes = [DP, C.c_int, C.c_int, C.c_int, IP, DP]
        L.or_bc_u.restype = None
        L.or_bc_face.argtypes = [DP, DP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, DP, DP]
        L.or_bc_face.restype = None
        L.or_step.argtypes = [C.POINTER(OrProblem), DP, DP, DP, LP]
        L.or_run.argtypes = [C.POINTER(OrProblem), DP, C.c_int, DP, DP, LP]
        L.or_run1d.argtypes = [DP, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                               DP, IP, DP, LP]

    @staticmethod
    def _dims(a, nvar_axis=True):
        s = a.shape[:-1] if nvar_axis else a.shape
        ix, jx = s[0], s[1]
        kx = s[2] if len(s) > 2 else 1
        return ix, jx, kx

    def u2v(self, U, V):
        ix, jx, kx = self._dims(U)
        bad = C.c_long(-1)
        rc = self.L.or_u2v(ptr(U), ptr(V), ix, jx, kx, C.byref(bad))
        return rc, bad.value

    def limiter(self, wk, wL, wR, dir, type):
        ix, jx, kx = self._dims(wk, nvar_axis=False)
        return self.L.or_limiter(ptr(wk), ptr(wL), ptr(wR), ix, jx, kx, dir, type)

    def hlld(self, L, R, ch, d):
        F = np.zeros(9)
        br = C.c_int(0)
        self.L.or_hlld(ptr(np.ascontiguousarray(L, dtype=np.float64)), ptr(np.ascontiguousarray(R, dtype=np.float64)),
                       ch, d, ptr(F), C.byref(br))
        return F, br.value

    def flux_solver(self, F, VL, VR, ch, dir, type=3, branches=None):
        ix, jx, kx = self._dims(VL)
        bp = branches.ctypes.data_as(IP) if branches is not None else None
        return self.L.or_flux_solver(ptr(F), ptr(VL), ptr(VR), ch, ix, jx, kx, dir, type, bp)

    def glm_ss2(self, U, ch, dt):
        ix, jx, kx = self._dims(U)
        self.L.or_glm_ss2(ptr(U), ch, dt, ix, jx, kx)

    def rk(self, mode, wk, wk1, F, G, H, dt, dx):
        ix, jx, kx = self._dims(wk)
        self.L.or_rk(mode, ptr(wk), ptr(wk1), ptr(F), ptr(G), ptr(H) if H is not None else None, dt, dx, ix, jx, kx)

    def set_dt(self, U, V, dx, cfl, ndim=2):
        ix, jx, kx = self._dims(U)
        vmax, dt, bad = C.c_double(), C.c_double(), C.c_long(-1)
        rc = self.L.or_set_dt(ptr(U), ptr(V), C.byref(vmax), C.byref(dt), dx, cfl, ix, jx, kx, ndim, C.byref(bad))
        return rc, vmax.value, dt.value

    def bc_u(self, U, prob):
        ix, jx, kx = self._dims(U)
        bct = (C.c_int * 6)(*prob.bctype)
        sg = np.array(prob.u_sgn[:], dtype=np.float64)
        self.L.or_bc_u(ptr(U), ix, jx, kx, bct, ptr(sg))

    def eta_faces(self, x, y, dx, eta0, eta1, w):
        """3 face arrays (x-, y-, z-faces) of eta(x,y), Fortran (ix,jx) each."""
        ix, jx = len(x), len(y)
        out = np.zeros((3, jx, ix))
        self.L.or_eta_faces.restype = C.c_double
        self.L.or_eta_faces.argtypes = [DP, DP, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        DP]
        emax = self.L.or_eta_faces(ptr(np.ascontiguousarray(x, dtype=np.float64)),
                                   ptr(np.ascontiguousarray(y, dtype=np.float64)), dx, ix, jx, eta0, eta1, w, ptr(out))
        return out, emax

    def run1d(self, U, dx, cfl=0.8, lm=1, flux=3, tt=0, tend=0.2, max_steps=29999):
        """1D shock-tube loop on U(ix,1,9) in place: (rc, t, dt trace)."""
        dts = np.zeros(max_steps)
        t, ns, bad = C.c_double(), C.c_int(), C.c_long(-1)
        rc = self.L.or_run1d(ptr(U), U.shape[0], dx, cfl, lm, flux, tt, tend, max_steps, C.byref(t), C.byref(ns),
                             ptr(dts), C.byref(bad))
        return rc, t.value, dts[:ns.value]

    def run(self, prob, U, nsteps):
        dts = np.zeros(nsteps)
        chs = np.zeros(nsteps)
        bad = C.c_long(-1)
        rc = self.L.or_run(C.byref(prob), ptr(U), nsteps, ptr(dts), ptr(chs), C.byref(bad))
        self.last_bad = bad.value  # 0-based linear cell of a failure (the reference's minloc)
        return rc, dts, chs


# ---------------------------------------------------------------------------
# flang-built reference routines (Fortran calling convention: all by reference)
# ---------------------------------------------------------------------------
def _i(v):
    return C.byref(C.c_int32(v))


def _d(v):
    return C.byref(C.c_double(v))


class Ref:
    """The reference's own subroutines, compiled unmodified from the reference
    sources by oracle/Makefile (target 'ref').  Only present where built."""

    def __init__(self, dim=2):
        self.dim = dim
        so = REF_DIR / ("libref2d.so" if dim == 2 else "libref3d.so")
        if not so.exists():
            raise FileNotFoundError(so)
        self.L = C.CDLL(str(so))
        kh = REF_DIR / "librefkh.so"
        self.KH = C.CDLL(str(kh)) if (dim == 2 and kh.exists()) else None

    def _ext(self, a, nvar_axis=True):
        s = a.shape[:-1] if nvar_axis else a.shape
        ext = [_i(x) for x in s]
        if self.dim == 3 and len(ext) == 2:
            ext.append(_i(1))
        return ext

    def u2v(self, U, V):
        self.L.u2v_(ptr(U), ptr(V), *self._ext(U))

    def limiter(self, wk, wL, wR, dir, type):
        self.L.limiter_(ptr(wk), ptr(wL), ptr(wR), *self._ext(wk, False), _i(dir), _i(type))

    def flux_solver(self, F, VL, VR, ch, dir, type=3):
        self.L.flux_solver_(ptr(F), ptr(VL), ptr(VR), _d(ch), *self._ext(VL), _i(dir), _i(type))

    def glm_ss2(self, U, ch, dt):
        self.L.glm_ss2_(ptr(U), _d(ch), _d(dt), *self._ext(U))

    def rk(self, mode, wk, wk1, F, G, H, dt, dx):
        name = ["rk_tvd21_", "rk_tvd22_", "rk_std21_", "rk_std22_"][mode]
        fn = getattr(self.L, name)
        arrs = [ptr(wk)] + ([] if mode == 3 else [ptr(wk1)]) + [ptr(F), ptr(G)]
        if self.dim == 3:
            arrs.append(ptr(H))
        fn(*arrs, _d(dt), _d(dx), *self._ext(wk))

    def set_dt(self, U, V, dx, cfl):
        vmax, dt = C.c_double(), C.c_double()
        self.L.set_dt_(ptr(U), ptr(V), C.byref(vmax), C.byref(dt), _d(dx), _d(cfl), *self._ext(U))
        return vmax.value, dt.value

    def bc(self, U):
        self.L.bc_(ptr(U), *self._ext(U))

    def bc_for_U(self, U):
        self.KH.bc_for_u_(ptr(U), *self._ext(U))

    def bc_for_F(self, VL, VR):
        self.KH.bc_for_f_(ptr(VL), ptr(VR), *self._ext(VL))

    def bc_for_G(self, VL, VR):
        self.KH.bc_for_g_(ptr(VL), ptr(VR), *self._ext(VL))


class RefRec:
    """2D_reconnection's boundary tables (bc.f90) and the 2D CPU form of the
    resistive flux (flux_resistive.f90), compiled unmodified by oracle/Makefile."""

    def __init__(self):
        so = REF_DIR / "librefrec.so"
        if not so.exists():
            raise FileNotFoundError(so)
        self.L = C.CDLL(str(so))

    def bc_for_U(self, U):
        self.L.bc_for_u_(ptr(U), _i(U.shape[0]), _i(U.shape[1]))

    def bc_for_F(self, VL, VR):
        self.L.bc_for_f_(ptr(VL), ptr(VR), _i(VL.shape[0]), _i(VL.shape[1]))

    def bc_for_G(self, VL, VR):
        self.L.bc_for_g_(ptr(VL), ptr(VR), _i(VL.shape[0]), _i(VL.shape[1]))

    def flux_resistive(self, F, U, x, y, dx, dir):
        self.L.flux_resistive_(ptr(F), ptr(U), ptr(np.ascontiguousarray(x)), ptr(np.ascontiguousarray(y)), _d(dx),
                               _i(U.shape[0]), _i(U.shape[1]), _i(dir))


# 2D_reconnection/bc.f90: slots negated in the mirror ghosts, per face
# (x-, x+, y-, y+): x conducting (mx, mz, By, psi), south symmetry plane
# (my, mz, Bx, psi), north wall (my, By)
def _neg(*slots):
    s = np.ones(9)
    s[list(slots)] = -1.0
    return s


RECON_SIGN = [_neg(0, 2, 6, 8), _neg(0, 2, 6, 8), _neg(1, 2, 5, 8), _neg(1, 6), np.ones(9), np.ones(9)]


def recon_problem(x, y, lm_type=1, flux_type=3, time_type=0, cfl=0.35):
    """An oracle problem in the form of 2D_reconnection/main.f90: the per-face
    tables of 2D_reconnection/bc.f90, the 2D CPU form of the resistive flux
    with that file's eta(x, y) on the faces, and set_dt_resistive's limit.
    dx = x(2) - x(1), as the drivers take it.  Returns (problem, tables); the
    caller keeps `tables` alive while the problem is in use."""
    from openmhd_amd.api import RECON_RM1, eta_face_table

    ix, jx = len(x), len(y)
    dx = float(x[1] - x[0])
    p = make_problem(ix, jx, dx=dx, cfl=cfl, lm_type=lm_type, flux_type=flux_type, time_type=time_type,
                     bc=("wall", "wall", "periodic"), signs=RECON_SIGN)
    tables = np.ascontiguousarray(np.stack([eta_face_table(x, y, dx, d).T for d in (1, 2)]))
    p.eta = ptr(tables)
    p.res_form = 1
    p.rm1 = RECON_RM1
    return p, tables


def recon3d_problem(x, y, nz, lm_type=1, flux_type=3, time_type=0, cfl=0.3):
    """An oracle problem in the form of 3D_reconnection_gpu/main.cuf: the
    reflecting tables of bc.cuf in x and y, periodic z, the 3D resistive flux
    with flux_resistive.cuf's eta(x, y) at the face centres (:47-83) and its
    set_dt_resistive (:2-9).  dx = x(2) - x(1).  Returns (problem, tables)."""
    from openmhd_amd.api import RECON_RM1, eta_recon

    ix, jx = len(x), len(y)
    dx = float(x[1] - x[0])
    p = make_problem(ix, jx, nz + 2, dx=dx, cfl=cfl, lm_type=lm_type, flux_type=flux_type, time_type=time_type,
                     bc=("wall", "wall", "periodic"), signs=RECON_SIGN)
    t = np.empty((3, jx, ix))
    for j in range(jx):
        for i in range(ix):
            t[0, j, i] = eta_recon(x[i] + 0.5 * dx, y[j])
            t[1, j, i] = eta_recon(x[i], y[j] + 0.5 * dx)
            t[2, j, i] = eta_recon(x[i], y[j])
    p.eta = ptr(t)
    p.res_form = 2
    p.rm1 = RECON_RM1
    return p, t


def ref_available(dim=2) -> bool:
    return (REF_DIR / ("libref2d.so" if dim == 2 else "libref3d.so")).exists()


def ref_run_2d(ref: Ref, U, nsteps, dx, cfl, lm_type, bc="periodic", time_type=0, flux_type=3):
    """One whole run sequenced like 2D_basic/main.f90:56-188 (bc='periodic') or
    2D_KH/main.f90 (bc='kh'), calling only the reference routines."""
    ix, jx = U.shape[:2]
    V = farr(ix, jx, 4)
    U1 = U.copy(order="F")
    VL, VR, F, G = (farr(ix, jx, 9) for _ in range(4))
    dts, chs = np.zeros(nsteps), np.zeros(nsteps)

    def bcU(A):
        ref.bc(A) if bc == "periodic" else ref.bc_for_U(A)

    def sweep(Ust, dir, out):
        for m in range(4):
            ref.limiter(V[:, :, m], VL[:, :, m], VR[:, :, m], dir, lm_type)
        for m in range(4, 9):
            ref.limiter(Ust[:, :, m], VL[:, :, m], VR[:, :, m], dir, lm_type)
        if dir == 1:
            if bc == "periodic":
                VR[ix - 2, :, :] = VR[0, :, :]
                VL[0, :, :] = VL[ix - 2, :, :]
            else:
                ref.bc_for_F(VL, VR)
        else:
            if bc == "periodic":
                VL[:, 0, :] = VL[:, jx - 2, :]
                VR[:, jx - 2, :] = VR[:, 0, :]
            else:
                ref.bc_for_G(VL, VR)
        ref.flux_solver(out, VL, VR, ch, dir, flux_type)

    for s in range(nsteps):
        ref.u2v(U, V)
        ch, dt = ref.set_dt(U, V, dx, cfl)
        ref.glm_ss2(U, ch, dt)
        sweep(U, 1, F)
        sweep(U, 2, G)
        ref.rk(0 if time_type == 0 else 2, U, U1, F, G, None, dt, dx)
        bcU(U1)
        ref.u2v(U1, V)
        sweep(U1, 1, F)
        sweep(U1, 2, G)
        ref.rk(1 if time_type == 0 else 3, U, U1, F, G, None, dt, dx)
        bcU(U)
        ref.glm_ss2(U, ch, dt)
        dts[s], chs[s] = dt, ch
    return dts, chs


def ref_run_3d(ref: Ref, U, nsteps, dx, cfl, lm_type, time_type=0, flux_type=3):
    """One whole periodic 3D run sequenced like 3D_basic/main.f90:57-229,
    calling only the reference routines."""
    ix, jx, kx = U.shape[:3]
    V = farr(ix, jx, kx, 4)
    U1 = U.copy(order="F")
    VL, VR, F, G, H = (farr(ix, jx, kx, 9) for _ in range(5))
    dts, chs = np.zeros(nsteps), np.zeros(nsteps)

    def sweep(Ust, dir, out):
        for m in range(4):
            ref.limiter(V[..., m], VL[..., m], VR[..., m], dir, lm_type)
        for m in range(4, 9):
            ref.limiter(Ust[..., m], VL[..., m], VR[..., m], dir, lm_type)
        if dir == 1:
            VR[ix - 2] = VR[0]
            VL[0] = VL[ix - 2]
        elif dir == 2:
            VL[:, 0] = VL[:, jx - 2]
            VR[:, jx - 2] = VR[:, 0]
        else:
            VL[:, :, 0] = VL[:, :, kx - 2]
            VR[:, :, kx - 2] = VR[:, :, 0]
        ref.flux_solver(out, VL, VR, ch, dir, flux_type)

    for s in range(nsteps):
        ref.u2v(U, V)
        ch, dt = ref.set_dt(U, V, dx, cfl)
        ref.glm_ss2(U, ch, dt)
        sweep(U, 1, F)
        sweep(U, 2, G)
        sweep(U, 3, H)
        ref.rk(0 if time_type == 0 else 2, U, U1, F, G, H, dt, dx)
        ref.bc(U1)
        ref.u2v(U1, V)
        sweep(U1, 1, F)
        sweep(U1, 2, G)
        sweep(U1, 3, H)
        ref.rk(1 if time_type == 0 else 3, U, U1, F, G, H, dt, dx)
        ref.bc(U)
        ref.glm_ss2(U, ch, dt)
        dts[s], chs[s] = dt, ch
    return dts, chs


# ---------------------------------------------------------------------------
# states
# ---------------------------------------------------------------------------
def prim_to_cons(W):
    """W[..., 9] primitive (vx,vy,vz,p,ro,bx,by,bz,ps) -> U[..., 9] conserved,
    as model.f90:46-50 computes it."""
    U = np.empty_like(W)
    ro = W[..., 4]
    v2 = W[..., 0] * W[..., 0] + W[..., 1] * W[..., 1] + W[..., 2] * W[..., 2]
    B2 = W[..., 5] * W[..., 5] + W[..., 6] * W[..., 6] + W[..., 7] * W[..., 7]
    U[..., 0:3] = ro[..., None] * W[..., 0:3]
    U[..., 3] = 0.5 * (ro * v2 + B2) + (1.0 / (GAMMA - 1)) * W[..., 3]
    U[..., 4:9] = W[..., 4:9]
    return U


def random_prim(rng, shape, vscale=1.0, bscale=1.0):
    W = np.empty(shape + (9,))
    W[..., 0:3] = rng.normal(0, vscale, shape + (3,))
    W[..., 3] = rng.uniform(0.1, 2.0, shape)
    W[..., 4] = rng.uniform(0.2, 3.0, shape)
    W[..., 5:8] = rng.normal(0, bscale, shape + (3,))
    W[..., 8] = rng.normal(0, 0.1, shape)
    return W


def smooth_state(rng, ix, jx, kx=1, amp=0.3):
    """A smooth random conserved state (Fortran order) with ghosts filled
    periodically by the caller."""
    x = np.linspace(0, 2 * np.pi, ix, endpoint=False)[:, None, None]
    y = np.linspace(0, 2 * np.pi, jx, endpoint=False)[None, :, None]
    z = np.linspace(0, 2 * np.pi, kx, endpoint=False)[None, None, :]
    W = np.empty((ix, jx, kx, 9))
    for m in range(9):
        a, b, c, ph = rng.normal(size=4)
        W[..., m] = amp * (a * np.sin(x + ph) + b * np.cos(2 * y - ph) + c * np.sin(z + 2 * ph))
    W[..., 3] = 1.0 + 0.5 * W[..., 3] ** 2
    W[..., 4] = 1.0 + 0.5 * W[..., 4] ** 2
    U = prim_to_cons(W)
    if kx == 1:
        U = U[:, :, 0, :]
    return np.asfortranarray(U)


def orszag_tang(n):
    """Orszag-Tang initial state exactly as 2D_basic/model.f90:19-50, n x n
    interior cells on [0, 2pi]^2.  Returns (U, dx) with U[n+2, n+2, 9]."""
    ix = jx = n + 2
    pi = 4.0 * np.arctan(1.0)
    dx = (2 * pi - 0.0) / float(ix - 2)
    x = np.empty(ix)
    y = np.empty(jx)
    x[0] = 0.0 - dx / 2
    for i in range(1, ix):
        x[i] = x[i - 1] + dx
    y[0] = 0.0 - dx / 2
    for j in range(1, jx):
        y[j] = y[j - 1] + dx
    U = farr(ix, jx, 9)
    X = x[:, None] * np.ones((1, jx))
    Y = np.ones((ix, 1)) * y[None, :]
    ro = GAMMA ** 2
    vx = -np.sin(Y)
    vy = np.sin(X)
    bxv = -np.sin(Y)
    byv = np.sin(2 * X)
    f1 = 1.0 / (GAMMA - 1)
    v2 = vx * vx + vy * vy + 0.0 * 0.0
    B2 = bxv * bxv + byv * byv + 0.0 * 0.0
    U[:, :, 0] = ro * vx
    U[:, :, 1] = ro * vy
    U[:, :, 2] = ro * 0.0
    U[:, :, 3] = 0.5 * (ro * v2 + B2) + f1 * GAMMA
    U[:, :, 4] = ro
    U[:, :, 5] = bxv
    U[:, :, 6] = byv
    U[:, :, 7] = 0.0
    U[:, :, 8] = 0.0
    return U, dx


class AdjCheck:
    """Host build of the adjoint building blocks (tests/native/adjcheck.cpp over
    openmhd_amd/csrc/omhd_adjoint.h): templated forward forms (double and dual
    numbers) and the hand-written reverse forms.  Built with g++ if missing."""

    def __init__(self):
        import subprocess

        d = Path(__file__).resolve().parent / "native"
        so = NATIVE_DIR / "libadjcheck.so"
        if not so.exists():
            subprocess.run(["make", "-s", "-C", str(d)] + (["sanitize"] if SANITIZED else []), check=True)
        self.L = C.CDLL(str(so))

    def __getattr__(self, name):
        return getattr(self.L, "adjc_" + name)


def cptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def soa(U):
    """Interior of a Fortran U(ix, jx, 9) as W(9, ny, nx), x fastest."""
    return np.ascontiguousarray(U[1:-1, 1:-1, :].transpose(2, 1, 0))
