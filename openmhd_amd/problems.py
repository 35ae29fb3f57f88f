"""Initial states of the benchmark configurations, generated on the host in
float64 and handed to both the GPU step and the CPU reference build (so the
two runs start from bit-identical states; transcendental functions are only
evaluated here, never on either side of the comparison).

Arrays are U(ix, jx, 9) in Fortran order with one ghost layer (ix = nx + 2),
slot order of param.h:6-15.  Conserved variables are formed from primitives
exactly as 2D_basic/model.f90:46-50 does.
"""
from __future__ import annotations

import numpy as np

GAMMA = 5.0 / 3.0  # param.h:3


def prim_to_cons_fortran(ro, vx, vy, vz, p, bx, by, bz, ps):
    """(ix,jx) primitive planes -> U(ix,jx,9), model.f90:46-50 order."""
    ix, jx = ro.shape
    U = np.empty((ix, jx, 9), order="F")
    f1 = 1.0 / (GAMMA - 1)
    v2 = vx * vx + vy * vy + vz * vz
    B2 = bx * bx + by * by + bz * bz
    U[:, :, 0] = ro * vx
    U[:, :, 1] = ro * vy
    U[:, :, 2] = ro * vz
    U[:, :, 3] = 0.5 * (ro * v2 + B2) + f1 * p
    U[:, :, 4] = ro
    U[:, :, 5] = bx
    U[:, :, 6] = by
    U[:, :, 7] = bz
    U[:, :, 8] = ps
    return U


def brio_wu(n: int = 300):
    """The Brio & Wu (1988) shock tube of 1D_basic/model.f90 (the active case),
    on [-0.5, 0.5] with n cells: rho, p, By = (1, 1, 1) for x < 0 and
    (0.125, 0.1, -1) otherwise, Bx = 0.75, v = Bz = psi = 0.  The grid is
    built the way model.f90:16-21 builds it (x(1) = -0.5 - dx/2, then
    x(i) = x(i-1) + dx).  Returns U(n+2, 1, 9) with its ghosts, dx and x."""
    ix = n + 2
    dx = (0.5 - (-0.5)) / float(ix - 2)
    x = np.empty(ix)
    x[0] = -0.5 - dx / 2
    for i in range(1, ix):
        x[i] = x[i - 1] + dx
    left = (x < 0.0)[:, None]
    z = np.zeros((ix, 1))
    ro = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    by = np.where(left, 1.0, -1.0)
    U = prim_to_cons_fortran(ro, z, z, z, p, z + 0.75, by, z, z)
    return U, dx, x


def orszag_tang(n: int):
    """Orszag-Tang vortex on [0,2pi]^2 with n x n cells, as 2D_basic/model.f90:
    rho = gamma^2, p = gamma, v = (-sin y, sin x, 0), B = (-sin y, sin 2x, 0),
    psi = 0.  Returns (U, dx); ghost cells hold the analytic values (the
    caller's first bc() overwrites them)."""
    ix = jx = n + 2
    pi = 4.0 * np.arctan(1.0)
    dx = (2 * pi - 0.0) / float(ix - 2)  # model.f90:19
    x = np.empty(ix)
    y = np.empty(jx)
    x[0] = 0.0 - dx / 2
    for i in range(1, ix):  # model.f90:22-24 (cumulative, as the reference)
        x[i] = x[i - 1] + dx
    y[0] = 0.0 - dx / 2
    for j in range(1, jx):
        y[j] = y[j - 1] + dx
    X = np.broadcast_to(x[:, None], (ix, jx))
    Y = np.broadcast_to(y[None, :], (ix, jx))
    z = np.zeros((ix, jx))
    U = prim_to_cons_fortran(np.full((ix, jx), GAMMA ** 2), -np.sin(Y), np.sin(X), z, np.full((ix, jx), GAMMA),
                             -np.sin(Y), np.sin(2 * X), z, z)
    return U, dx


# -- counter-based Gaussian noise: value depends only on (seed, i, j) ----------
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def counter_normal(seed: int, i: np.ndarray, j: np.ndarray) -> np.ndarray:
    """Standard normal deviates keyed by (seed, i, j) (Box-Muller on two
    splitmix64 uniforms); independent of how the grid is tiled."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0xD1B54A32D192ED03)) & _M64
        c = (i.astype(np.uint64) << np.uint64(32)) | j.astype(np.uint64)
        a = _splitmix64(c ^ key)
        b = _splitmix64(a ^ np.uint64(0x632BE59BD9B4E019))
    u1 = ((a >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)
    u2 = ((b >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def kelvin_helmholtz(nx: int, ny: int, seed: int = 1, delta: float = 1.0, lx: float = 40.0, u0: float = 0.5,
                     beta: float = 2.0, bx: float = 0.2, bz: float = 1.0, noise: float = 1.0e-3):
    """Kelvin-Helmholtz layer: x in [0, lx*delta) periodic, y in
    [-ly/2, ly/2] with walls (ly = ny*dx, dx = lx*delta/nx), vx = u0 tanh(y/delta),
    rho = 1, uniform p = beta |B|^2 / 2, B = (bx, 0, bz), and seeded Gaussian vy
    noise of amplitude `noise` where |y| < 2 delta.  Returns (U, dx)."""
    ix, jx = nx + 2, ny + 2
    dx = lx * delta / nx
    ly = ny * dx
    i = np.arange(ix)
    j = np.arange(jx)
    x = (i - 0.5) * dx  # cell centres, ghost at i = 0
    y = -0.5 * ly + (j - 0.5) * dx
    X = np.broadcast_to(x[:, None], (ix, jx))
    Y = np.broadcast_to(y[None, :], (ix, jx))
    I = np.broadcast_to(i[:, None], (ix, jx))
    J = np.broadcast_to(j[None, :], (ix, jx))
    p = beta * (bx * bx + bz * bz) / 2.0
    vx = u0 * np.tanh(Y / delta)
    vy = np.where(np.abs(Y) < 2.0 * delta, noise * counter_normal(seed, I, J), 0.0)
    z = np.zeros((ix, jx))
    one = np.ones((ix, jx))
    U = prim_to_cons_fortran(one, vx, vy, z, p * one, bx * one, z, bz * one, z)
    del X
    return U, dx


def harris_sheet_3d(nx: int, ny: int, nz: int, dx: float, seed: int = 5, beta_b: float = 0.2, rho0: float = 0.2,
                    lam: float = 1.0, noise: float = 1.0e-3):
    """3D Harris current sheet for the resistive reconnection configuration:
    Bx = tanh(y/lam), rho = rho0 + cosh(y/lam)^-2 / beta_b, isothermal
    p = beta_b * rho / 2 (p + B^2/2 uniform), seeded Gaussian noise of
    amplitude `noise` in all three velocity components.  The domain is
    x in [-nx dx/2, nx dx/2), y in [-ny dx/2, ny dx/2], z in [0, nz dx) (the
    sheet and the resistivity peak at x = y = 0).  Returns (U, origin) with
    U(ix,jx,kx,9) Fortran-ordered and origin = (x0, y0)."""
    ix, jx, kx = nx + 2, ny + 2, nz + 2
    x0, y0 = -0.5 * nx * dx, -0.5 * ny * dx
    y = y0 + (np.arange(jx) - 0.5) * dx
    Y = np.broadcast_to(y[None, :, None], (ix, jx, kx))
    I = np.broadcast_to(np.arange(ix)[:, None, None], (ix, jx, kx))
    J = np.broadcast_to(np.arange(jx)[None, :, None], (ix, jx, kx))
    K = np.broadcast_to(np.arange(kx)[None, None, :], (ix, jx, kx))
    ro = rho0 + np.cosh(Y / lam) ** -2 / beta_b
    p = 0.5 * beta_b * ro
    bx = np.tanh(Y / lam)
    key = (K * jx + J).astype(np.uint64)
    vx = noise * counter_normal(seed, I, key)
    vy = noise * counter_normal(seed + 1, I, key)
    vz = noise * counter_normal(seed + 2, I, key)
    z = np.zeros((ix, jx, kx))
    f1 = 1.0 / (GAMMA - 1)
    U = np.empty((ix, jx, kx, 9), order="F")
    U[..., 0] = ro * vx
    U[..., 1] = ro * vy
    U[..., 2] = ro * vz
    U[..., 3] = 0.5 * (ro * (vx * vx + vy * vy + vz * vz) + bx * bx) + f1 * p
    U[..., 4] = ro
    U[..., 5] = bx
    U[..., 6] = z
    U[..., 7] = z
    U[..., 8] = z
    return U, (x0, y0)


def harris_sheet_2d(nx: int, ny: int, lx: float = 200.0, beta: float = 0.2, b1: float = 0.03):
    """The 2D reconnection problem of 2D_reconnection/model.f90: a Harris
    sheet Bx = tanh(y), rho = 1 + cosh(y)^-2 / beta, p = beta rho / 2, v = 0,
    on x in [0, lx] and y from 0 (the south symmetry plane through the sheet),
    with the X-point perturbation Bx -= b1 y exp(-r^2/4), By += b1 x
    exp(-r^2/4) at the origin.  The grid is built the way model.f90:22-31
    builds it (dx = lx / nx, x(1) = -dx/2, then x(i) = x(i-1) + dx; y alike
    from -dx/2).  Returns (U, dx, x, y) with U(nx+2, ny+2, 9) in Fortran
    order, ghosts included (a caller fills them with the problem's
    bc_for_U before the first step)."""
    ix, jx = nx + 2, ny + 2
    dx = lx / float(ix - 2)
    x = np.empty(ix)
    y = np.empty(jx)
    x[0] = -dx / 2
    for i in range(1, ix):
        x[i] = x[i - 1] + dx
    y[0] = -dx / 2
    for j in range(1, jx):
        y[j] = y[j - 1] + dx
    X, Y = np.meshgrid(x, y, indexing="ij")
    ro = 1.0 + np.cosh(np.minimum(Y, 25.0)) ** -2 / beta
    p = 0.5 * beta * ro
    g = np.exp(-(X ** 2 + Y ** 2) / 4.0)
    bx = np.tanh(Y) - b1 * Y * g
    by = b1 * X * g
    z = np.zeros((ix, jx))
    U = prim_to_cons_fortran(ro, z, z, z, p, bx, by, z, z)
    return U, dx, x, y


def kh_ensemble(n: int, seed: int = 2026, smin: int = 64, smax: int = 1024):
    """Parameters of an ensemble of independent KH runs drawn from one master
    seed: grid sides nx, ny log-uniform in [smin, smax] (independently), shear
    speed U0 ~ U[0.2, 1.0], beta ~ U[0.5, 5], Bx ~ U[0, 0.3], and a noise seed
    per member.  Returns a list of dicts (nx, ny, u0, beta, bx, seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n):
        nx, ny = (int(round(np.exp(rng.uniform(np.log(smin), np.log(smax))))) for _ in range(2))
        out.append(dict(nx=nx, ny=ny, u0=float(rng.uniform(0.2, 1.0)), beta=float(rng.uniform(0.5, 5.0)),
                        bx=float(rng.uniform(0.0, 0.3)), seed=int(rng.integers(1, 2**31 - 1))))
    return out


def kh_member(p: dict):
    """Initial state of one ensemble member: the KH layer of kelvin_helmholtz
    on an nx x ny grid of the member's parameters (Lx = 40, y extent ny*dx)."""
    return kelvin_helmholtz(p["nx"], p["ny"], seed=p["seed"], u0=p["u0"], beta=p["beta"], bx=p["bx"])


def fourier_modes(seed: int, n_modes: int = 16, kmax: int = 8):
    """A seeded set of in-plane velocity modes for the adjoint configuration:
    integer wave vectors k with 1 <= |k| <= kmax, phases phi in [0, 2pi) and
    unit directions e in the (vx, vy) plane.  dv(x) = sum_m c_m e_m sin(k_m.x
    + phi_m) on the [0, 2pi)^2 Orszag-Tang box."""
    rng = np.random.default_rng(seed)
    ks = []
    while len(ks) < n_modes:
        k = rng.integers(-kmax, kmax + 1, 2)
        if 1 <= np.hypot(*k) <= kmax:
            ks.append(k)
    th = rng.uniform(0, 2 * np.pi, n_modes)
    return {"k": np.array(ks, dtype=np.float64), "phi": rng.uniform(0, 2 * np.pi, n_modes),
            "e": np.stack([np.cos(th), np.sin(th)], axis=1)}


def orszag_tang_perturbed(n: int, modes, coef, fixed: str = "energy"):
    """Interior Orszag-Tang state W(9, n, n) (x fastest, conserved variables)
    with the velocity perturbation of `modes` scaled by `coef` (16 values);
    numpy, or torch (then differentiable in coef).  Density and B are those of
    the vortex.  fixed="energy" keeps the total energy of the vortex (the
    kinetic energy change comes out of the pressure, so p is not uniform);
    fixed="pressure" keeps p = gamma uniform.  The uniform-pressure state sits
    on the limiters' kink for every cell (zero pressure slopes), where the
    step has direction-dependent one-sided derivatives and no gradient; the
    fixed-energy state does not."""
    dx = 2 * np.pi / n
    x = (np.arange(n) + 0.5) * dx  # cell centres of model.f90:19-24 (x(i) = -dx/2 + (i-1) dx)
    X, Y = np.meshgrid(x, x, indexing="xy")  # X[j, i] = x_i
    basis = np.stack([np.sin(k[0] * X + k[1] * Y + p) for k, p in zip(modes["k"], modes["phi"])])  # (M, n, n)
    ex, ey = modes["e"][:, 0], modes["e"][:, 1]
    fields = {"vx": -np.sin(Y), "vy": np.sin(X), "bx": -np.sin(Y), "by": np.sin(2 * X)}
    is_torch = hasattr(coef, "requires_grad")
    if is_torch:
        import torch

        def T(a):
            return torch.as_tensor(a, device=coef.device, dtype=torch.float64)

        basis, ex, ey = T(basis), T(ex), T(ey)
        fields = {k: T(v) for k, v in fields.items()}
        dvx = torch.tensordot(coef * ex, basis, dims=1)
        dvy = torch.tensordot(coef * ey, basis, dims=1)
        stack, zeros = torch.stack, torch.zeros_like
    else:
        dvx = np.tensordot(coef * ex, basis, axes=1)
        dvy = np.tensordot(coef * ey, basis, axes=1)
        stack, zeros = np.stack, np.zeros_like
    ro, p = GAMMA ** 2, GAMMA
    vx, vy = fields["vx"] + dvx, fields["vy"] + dvy
    bx, by = fields["bx"], fields["by"]
    z = zeros(vx)
    if fixed == "energy":
        v0x, v0y = fields["vx"], fields["vy"]
        en = 0.5 * (ro * (v0x * v0x + v0y * v0y) + (bx * bx + by * by)) + p / (GAMMA - 1) + z
    else:
        en = 0.5 * (ro * (vx * vx + vy * vy) + (bx * bx + by * by)) + p / (GAMMA - 1)
    return stack([ro * vx, ro * vy, z, en, z + ro, bx, by, z, z])
