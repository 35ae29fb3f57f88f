"""Pins of the oracle's 3D resistive flux in all three face directions
(or_flux_resistive, the arithmetic of 3D_reconnection_gpu/flux_resistive.cuf:86-135).

The reference has no CPU build of this loop (it is CUDA Fortran only), so the
flang vectors of the 2D CPU form pin only the in-plane terms.  What pins the
rest here:

* Cyclic symmetry.  The reference's three face branches are one expression
  under the relabelling x -> y -> z -> x (i -> j -> k, Bx -> By -> Bz): the
  z-face branch (:123-135) is the x-face branch (:89-101) with the axes
  rotated twice, term for term and in the same order.  So on a cubic grid the
  y-face flux of the rotated field, and the z-face flux of the field rotated
  twice, must equal the x-face flux of the original bit for bit.
* Closed forms with z-dependence, from Ampere's law J = curl B and the
  resistive electric field E = eta J: B = (a z, 0, Bz0) has J = (0, a, 0);
  B = (0, b z, Bz0) has J = (-b, 0, 0).  The induction equation's fluxes are
  F_x(By) = -E_z, F_x(Bz) = E_y, F_y(Bx) = E_z, F_y(Bz) = -E_x,
  F_z(Bx) = -E_y, F_z(By) = E_x, and the energy flux is the Poynting vector
  E x B.  For a linear field the face-centred differences are exact, and
  with power-of-two data every value below is exact.
* Planted slips.  Copies of the oracle source with a sign or index mistake in
  the z-face branch or in an out-of-plane k+-1 term must fail one of the
  checks above.
"""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

from omhd_testlib import ptr

ORACLE_SRC = Path(__file__).resolve().parents[1] / "oracle" / "omhd_oracle.c"
MX, MY, MZ, EN, RO, BX, BY, BZ, PS = range(9)
# rotation x -> y -> z -> x of the slots: new (x, y, z) component <- old (z, x, y) component
PERM = [MZ, MX, MY, EN, RO, BZ, BX, BY, PS]


def flux_fn(lib):
    f = lib.or_flux_resistive
    f.argtypes = [C.POINTER(C.c_double)] * 3 + [C.c_double] + [C.c_int] * 4
    f.restype = None

    def call(F, U, eta, dx, d):
        ix, jx, kx = U.shape[:3]
        f(ptr(F), ptr(U), ptr(eta), dx, ix, jx, kx, d)

    return call


def rot(A):
    """A(i,j,k,:) -> A'(i',j',k') = A(j',k',i') with the components relabelled:
    the field seen with the axes renamed x -> y -> z -> x."""
    return np.asfortranarray(A.transpose(2, 0, 1, 3)[..., PERM])


def bits_equal(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def check_cyclic(call, seed=3, n=9):
    """x-face flux of U == y-face flux of rot(U) == z-face flux of rot(rot(U)),
    bit for bit, on random B and random incoming F (uniform eta)."""
    rng = np.random.default_rng(seed)
    U = np.asfortranarray(rng.standard_normal((n, n, n, 9)))
    U[..., RO] = 1.0 + rng.random((n, n, n))
    F0 = np.asfortranarray(rng.standard_normal((n, n, n, 9)))
    eta = np.full((n, n), 0.0137)
    dx = 0.173
    F1 = F0.copy(order="F")
    call(F1, U, eta, dx, 1)
    if bits_equal(F1, F0):
        return False  # the routine must do something on random data
    G = rot(F0)
    call(G, rot(U), eta, dx, 2)
    H = rot(rot(F0))
    call(H, rot(rot(U)), eta, dx, 3)
    return bits_equal(G, rot(F1)) and bits_equal(H, rot(rot(F1)))


def check_closed_forms(call):
    """B = (a z, 0, Bz0) and B = (0, b z, Bz0): exact face fluxes on every face
    direction (see the module docstring)."""
    ix, jx, kx = 7, 6, 8
    dx, eta0, bz0 = 0.5, 0.03125, 1.0
    eta_dx = eta0 / dx
    k = np.arange(kx, dtype=np.float64)  # z in units of dx (any offset cancels)
    ok = True
    for case, coef in (("bx", 0.25), ("by", -0.5)):
        U = np.zeros((ix, jx, kx, 9), order="F")
        U[..., RO] = 1.0
        U[..., BZ] = bz0
        U[..., BX if case == "bx" else BY] = coef * k[None, None, :]
        # J * dx (the reference's JxS/JyS/JzS) and E = eta J
        Jx, Jy, Jz = (0.0, coef, 0.0) if case == "bx" else (-coef, 0.0, 0.0)
        Ex, Ey, Ez = eta_dx * Jx, eta_dx * Jy, eta_dx * Jz
        for d in (1, 2, 3):
            F = np.zeros((ix, jx, kx, 9), order="F")
            call(F, U, np.full((jx, ix), eta0), dx, d)
            lo = [1, 1, 1]
            hi = [ix - 1, jx - 1, kx - 1]  # 0-based faces of the reference's ranges
            lo[d - 1] = 0
            face = F[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
            # face averages of B (the reference's 0.5*(B(c) + B(c+1)) along d)
            sl0 = (slice(lo[0], hi[0]), slice(lo[1], hi[1]), slice(lo[2], hi[2]))
            sl1 = tuple(slice(s.start + (1 if a == d - 1 else 0), s.stop + (1 if a == d - 1 else 0))
                        for a, s in enumerate(sl0))
            Bf = [0.5 * (U[sl0 + (m,)] + U[sl1 + (m,)]) for m in (BX, BY, BZ)]
            want = np.zeros(face.shape)
            if d == 1:
                want[..., BY], want[..., BZ] = -Ez, Ey
                want[..., EN] = Ey * Bf[2] - Ez * Bf[1]
            elif d == 2:
                want[..., BX], want[..., BZ] = Ez, -Ex
                want[..., EN] = Ez * Bf[0] - Ex * Bf[2]
            else:
                want[..., BX], want[..., BY] = -Ey, Ex
                want[..., EN] = Ex * Bf[1] - Ey * Bf[0]
            ok = ok and bool(np.array_equal(face, want))
            # faces outside the reference's range are untouched
            rest = F.copy()
            rest[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = 0
            ok = ok and not rest.any()
    return ok


def test_cyclic_symmetry_bitwise(oracle):
    assert check_cyclic(flux_fn(oracle.L))
    assert check_cyclic(flux_fn(oracle.L), seed=11, n=7)


def test_closed_forms_z_dependent(oracle):
    assert check_closed_forms(flux_fn(oracle.L))


# (name, original text in omhd_oracle.c, planted replacement)
SLIPS = [
    ("z-face JyS sign", "double JyS = (B(i, j, k + 1, bx) - B(i, j, k, bx)) -",
     "double JyS = (B(i, j, k, bx) - B(i, j, k + 1, bx)) -"),
    ("z-face JxS k index", "- B(i, j - 1, k, bz) - B(i, j - 1, k + 1, bz));",
     "- B(i, j - 1, k, bz) - B(i, j - 1, k, bz));"),
    ("z-face by flux sign", "F[IX4(i, j, k, by)] = F[IX4(i, j, k, by)] + eta_dx * JxS;",
     "F[IX4(i, j, k, by)] = F[IX4(i, j, k, by)] - eta_dx * JxS;"),
    ("z-face energy uses By at k", "(JxS * (B(i, j, k, by) + B(i, j, k + 1, by))",
     "(JxS * (B(i, j, k, by) + B(i, j, k, by))"),
    ("x-face Bx(k-1) index", "- B(i, j, k - 1, bx) - B(i + 1, j, k - 1, bx));",
     "- B(i, j, k - 1, bx) - B(i + 1, j, k + 1, bx));"),
    ("y-face By(k+1) index", "0.25 * (B(i, j, k + 1, by) + B(i, j + 1, k + 1, by)",
     "0.25 * (B(i, j, k + 1, by) + B(i, j + 1, k, by)"),
]


@pytest.mark.parametrize("name,orig,bad", SLIPS, ids=[s[0] for s in SLIPS])
def test_planted_slip_fails(tmp_path, name, orig, bad):
    src = ORACLE_SRC.read_text()
    assert src.count(orig) == 1, f"pattern for '{name}' not found once in the oracle"
    mut = tmp_path / "mut.c"
    mut.write_text(src.replace(orig, bad))
    (tmp_path / "omhd_oracle.h").write_text((ORACLE_SRC.parent / "omhd_oracle.h").read_text())
    so = tmp_path / "libmut.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-ffp-contract=off", "-shared", "-o", str(so), str(mut), "-lm"],
                   check=True, capture_output=True, timeout=120)
    call = flux_fn(C.CDLL(str(so)))
    assert not (check_cyclic(call) and check_closed_forms(call)), f"planted slip '{name}' went unnoticed"
