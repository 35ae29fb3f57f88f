"""The oracle (C restatement) against the golden vectors produced by the
reference's own routines (tests/golden/gen_golden.py).  Bitwise."""
import hashlib
from pathlib import Path

import numpy as np
import pytest

from omhd_testlib import KH_SIGN, farr, make_problem, orszag_tang

G = Path(__file__).resolve().parent / "golden"


def load(name):
    return np.load(G / f"{name}.npz", allow_pickle=False)


def bits_equal(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def test_u2v(oracle):
    d = load("u2v_2d")
    V = farr(*d["V"].shape)
    rc, _ = oracle.u2v(np.asfortranarray(d["U"]), V)
    assert rc == 0 and bits_equal(V, d["V"])


@pytest.mark.parametrize("typ", range(5))
@pytest.mark.parametrize("dir", (1, 2, 3))
def test_limiter(oracle, typ, dir):
    d = load("limiter")
    wk = np.asfortranarray(d[f"wk_{typ}_{dir}"])
    wL = np.full(wk.shape, 7.0, order="F")
    wR = np.full(wk.shape, -7.0, order="F")
    assert oracle.limiter(wk, wL, wR, dir, typ) == 0
    assert bits_equal(wL, d[f"wL_{typ}_{dir}"])
    assert bits_equal(wR, d[f"wR_{typ}_{dir}"])


@pytest.mark.parametrize("dir", (1, 2, 3))
def test_hlld_every_branch(oracle, dir):
    d = load("hlld")
    VL, VR = np.asfortranarray(d[f"VL_{dir}"]), np.asfortranarray(d[f"VR_{dir}"])
    F = farr(*VL.shape)
    br = np.zeros(VL.shape[:-1], dtype=np.int32, order="F")
    assert oracle.flux_solver(F, VL, VR, float(d["ch"]), dir, 3, branches=br) == 0
    assert bits_equal(F, d[f"F_{dir}"])
    # the fixture covers all eight outcomes of flux_solver.f90:413-572
    assert set(np.unique(d[f"branch_{dir}"])) == set(range(1, 9))


def test_glm_rk_setdt_bc(oracle):
    d = load("small")
    U = np.asfortranarray(d["glm_in"]).copy(order="F")
    oracle.glm_ss2(U, 1.37, 0.0123)
    assert bits_equal(U, d["glm_out"])
    for dim, H in (("2d", None), ("3d", "rk3d_H")):
        F, Gf = d[f"rk{dim}_F"], d[f"rk{dim}_G"]
        Hh = np.asfortranarray(d[H]) if H else None
        for mode in range(4):
            A = np.asfortranarray(d[f"rk{dim}_{mode}_in_wk"]).copy(order="F")
            B = np.asfortranarray(d[f"rk{dim}_{mode}_in_wk1"]).copy(order="F")
            oracle.rk(mode, A, B, np.asfortranarray(F), np.asfortranarray(Gf), Hh, 0.0123, 0.0371)
            assert bits_equal(A, d[f"rk{dim}_{mode}_wk"]), (dim, mode)
            assert bits_equal(B, d[f"rk{dim}_{mode}_wk1"]), (dim, mode)
    rc, vmax, dt = oracle.set_dt(np.asfortranarray(d["setdt2d_U"]), np.asfortranarray(d["setdt2d_V"]), 0.0371, 0.4, 2)
    assert rc == 0 and bits_equal([vmax, dt], d["setdt2d_out"])
    rc, vmax, dt = oracle.set_dt(np.asfortranarray(d["setdt3d_U"]), np.asfortranarray(d["setdt3d_V"]), 0.0371, 0.4, 3)
    assert rc == 0 and bits_equal([vmax, dt], d["setdt3d_out"])
    A = np.asfortranarray(d["bc_in"]).copy(order="F")
    oracle.bc_u(A, make_problem(7, 6))
    assert bits_equal(A, d["bc_out"])
    A = np.asfortranarray(d["bc3d_in"]).copy(order="F")
    oracle.bc_u(A, make_problem(6, 5, 7))
    assert bits_equal(A, d["bc3d_out"])
    A = np.asfortranarray(d["bckh_in"]).copy(order="F")
    oracle.bc_u(A, make_problem(7, 6, bc=("periodic", "wall", "periodic")))
    assert bits_equal(A, d["bckh_out"])


def test_face_patch(oracle):
    d = load("small")
    import ctypes as C

    from omhd_testlib import ptr

    VL = np.asfortranarray(d["bcF_inL"]).copy(order="F")
    VR = np.asfortranarray(d["bcF_inR"]).copy(order="F")
    one = np.ones(9)
    oracle.L.or_bc_face(ptr(VL), ptr(VR), 7, 6, 1, 1, 0, ptr(one), ptr(one))
    assert bits_equal(VL, d["bcF_L"]) and bits_equal(VR, d["bcF_R"])
    s = np.ascontiguousarray(KH_SIGN)
    oracle.L.or_bc_face(ptr(VL), ptr(VR), 7, 6, 1, 2, 1, ptr(s), ptr(s))
    assert bits_equal(VL, d["bcG_L"]) and bits_equal(VR, d["bcG_R"])
    del C


@pytest.mark.parametrize("lm", range(5))
def test_every_limiter_run(oracle, lm):
    d = load("limiters_run")
    U = np.asfortranarray(d[f"U0_{lm}"]).copy(order="F")
    rc, dts, _ = oracle.run(make_problem(20, 18, dx=0.05, cfl=0.4, lm_type=lm), U, 10)
    assert rc == 0
    assert bits_equal(dts, d[f"dt_{lm}"]) and bits_equal(U, d[f"U_{lm}"])


def test_kh_walls_run(oracle):
    d = load("kh128_minmod_100")
    U = np.asfortranarray(d["U0"]).copy(order="F")
    p = make_problem(130, 130, dx=float(d["dx"]), cfl=0.4, lm_type=1, bc=("periodic", "wall", "periodic"))
    rc, dts, chs = oracle.run(p, U, 100)
    assert rc == 0
    assert bits_equal(dts, d["dt"]) and bits_equal(chs, d["ch"])
    assert bits_equal(U, d["final"])


def test_orszag_tang_1000_steps(oracle):
    """OT 200^2, MC, HLLD, GLM, TVD-RK2, CFL 0.4, 1000 steps: the
    dt sequence and the final state are bitwise those of the reference."""
    d = load("ot200_mc_1000")
    U, dx = orszag_tang(200)
    p = make_problem(202, 202, dx=dx, cfl=0.4, lm_type=2)
    oracle.bc_u(U, p)
    assert hashlib.sha256(U.tobytes(order="F")).hexdigest() == str(d["u0_sha256"])
    rc, dts, chs = oracle.run(p, U, 1000)
    assert rc == 0
    assert bits_equal(dts, d["dt"]) and bits_equal(chs, d["ch"])
    interior = np.ascontiguousarray(U[1:-1, 1:-1, :].transpose(2, 1, 0))
    assert hashlib.sha256(interior.tobytes()).hexdigest() == str(d["final_sha256"])
    from openmhd_amd.diagnostics import div_b_l2

    assert bits_equal(div_b_l2(U, dx), d["divb_l2"])  # the reference's ||div B||_2, from its own final state


def test_3d_periodic_run(oracle):
    """3D periodic (3D_basic/main.f90 sequence), MC, 4 steps, bitwise to the
    reference routines."""
    d = load("run3d")
    U = np.asfortranarray(d["U0"]).copy(order="F")
    ix, jx, kx = U.shape[:3]
    rc, dts, chs = oracle.run(make_problem(ix, jx, kx, dx=0.1, cfl=0.4, lm_type=2), U, 4)
    assert rc == 0
    assert bits_equal(dts, d["dt"]) and bits_equal(chs, d["ch"]) and bits_equal(U, d["U"])


def test_resistive_closed_form(oracle):
    """Oracle flux_resistive on B = (0, 0, a x): J_y = -a, so the x-face flux of
    Bz is eta*J_y = -eta*a and the energy flux is (E x B)_x = eta*J_y*Bz; uniform
    B gives no resistive flux (3D_reconnection_gpu/flux_resistive.cuf:89-101)."""
    import ctypes as C

    from omhd_testlib import ptr

    ix, jx, kx, dx, a = 8, 6, 5, 0.5, 0.3
    U = np.zeros((ix, jx, kx, 9), order="F")
    U[..., 4] = 1.0
    x = (np.arange(ix) - 0.5) * dx
    U[..., 7] = a * x[:, None, None]
    eta = np.full((jx, ix), 0.02)
    F = np.zeros((ix, jx, kx, 9), order="F")
    oracle.L.or_flux_resistive.argtypes = [C.POINTER(C.c_double)] * 3 + [C.c_double] + [C.c_int] * 4
    oracle.L.or_flux_resistive(ptr(F), ptr(U), ptr(eta), dx, ix, jx, kx, 1)
    inner = F[0:ix - 1, 1:jx - 1, 1:kx - 1]
    np.testing.assert_allclose(inner[..., 7], -0.02 * a, rtol=1e-14)
    bz_face = 0.5 * (U[0:ix - 1, 1:jx - 1, 1:kx - 1, 7] + U[1:ix, 1:jx - 1, 1:kx - 1, 7])
    np.testing.assert_allclose(inner[..., 3], -0.02 * a * bz_face, rtol=1e-13, atol=1e-16)
    assert np.all(inner[..., [0, 1, 2, 4, 5, 6, 8]] == 0)
    U[..., 7] = 0.7  # uniform field: no current
    U[..., 5] = -0.2
    F[...] = 0
    for d in (1, 2, 3):
        oracle.L.or_flux_resistive(ptr(F), ptr(U), ptr(eta), dx, ix, jx, kx, d)
    assert np.all(F == 0)


def test_resistivity_profile_and_dt_closed_form(oracle):
    """eta = eta0 + eta1 exp(-(x^2+y^2)/w^2) at x-faces (x + dx/2), y-faces
    (y + dx/2) and z-faces (cell centre); its peak sits on the face nearest the
    origin; the resistive dt limit is 0.5 cfl dx^2 / eta_max when it binds
    (3D_reconnection_gpu/flux_resistive.cuf:2-9)."""
    import ctypes as C

    dx, ix, jx = 0.25, 12, 10
    x = -1.5 + dx * np.arange(ix)
    y = -1.25 + dx * np.arange(jx)
    out, emax = oracle.eta_faces(x, y, dx, 1e-3, 1e-2, 2.0)
    X, Y = np.meshgrid(x, y, indexing="xy")
    for d, (ox, oy) in enumerate(((0.5 * dx, 0), (0, 0.5 * dx), (0, 0))):
        ref = 1e-3 + 1e-2 * np.exp(-((X + ox) ** 2 + (Y + oy) ** 2) / 4.0)
        np.testing.assert_allclose(out[d], ref, rtol=1e-15, atol=0)
    assert emax == out.max() and abs(emax - (1e-3 + 1e-2)) < 1e-12  # a face passes through the origin
    f = oracle.L.or_set_dt_resistive
    f.restype = C.c_double
    f.argtypes = [C.c_double] * 4
    assert f(1.0, 0.1, 0.4, 0.011) == 0.5 * 0.4 * (0.1 * 0.1) / 0.011
    assert f(1e-4, 0.1, 0.4, 0.011) == 1e-4


def test_reconnection_boundary_tables(oracle):
    """Per-face sign tables, against 2D_reconnection/bc.f90 (bitwise): x
    conducting walls, a south symmetry plane and a north wall with different
    negated slots on the two y faces; corners from x then y."""
    from omhd_testlib import RECON_SIGN, ptr

    d = load("recon")
    U = np.asfortranarray(d["bcU_in"]).copy(order="F")
    oracle.bc_u(U, make_problem(9, 7, bc=("wall", "wall", "periodic"), signs=RECON_SIGN))
    assert bits_equal(U, d["bcU_out"])
    for dr, kl, kr in ((1, "bcF_L", "bcF_R"), (2, "bcG_L", "bcG_R")):
        VL = np.asfortranarray(d["V_L"]).copy(order="F")
        VR = np.asfortranarray(d["V_R"]).copy(order="F")
        lo = np.ascontiguousarray(RECON_SIGN[2 * dr - 2])
        hi = np.ascontiguousarray(RECON_SIGN[2 * dr - 1])
        oracle.L.or_bc_face(ptr(VL), ptr(VR), 9, 7, 1, dr, 1, ptr(lo), ptr(hi))
        assert bits_equal(VL, d[kl]) and bits_equal(VR, d[kr]), dr


def test_resistive_flux_against_the_2d_reference_form(oracle):
    """The oracle's resistive flux (the 3D form of
    3D_reconnection_gpu/flux_resistive.cuf) on a z-invariant field against the
    reference's 2D CPU form (2D_reconnection/flux_resistive.f90) with that
    file's eta(x, y) profile.  The 2D form folds eta/dx into the currents, so
    the order of operations differs: agreement to rounding, not bitwise."""
    import ctypes as C

    from omhd_testlib import ptr

    d = load("recon")
    U2, x, y, dx = d["res_U"], d["res_x"], d["res_y"], float(d["res_dx"])
    ix, jx = U2.shape[:2]
    rm0, rm1 = 1000.0, 60.0
    e0, e01 = 1.0 / rm0, (rm0 - rm1) / (rm0 * rm1)

    def eta(xx, yy):
        return e0 + e01 * np.cosh(np.minimum(np.sqrt(xx ** 2 + yy ** 2), 25.0)) ** -2.0

    X, Y = np.meshgrid(x, y, indexing="xy")  # (jx, ix)
    faces = np.ascontiguousarray(np.stack([eta(X + 0.5 * dx, Y), eta(X, Y + 0.5 * dx), eta(X, Y)]))
    U3 = np.asfortranarray(np.repeat(U2[:, :, None, :], 3, axis=2))
    oracle.L.or_flux_resistive.argtypes = [C.POINTER(C.c_double)] * 3 + [C.c_double] + [C.c_int] * 4
    for dr in (1, 2):
        F0 = d[f"F0_{dr}"]
        F3 = np.asfortranarray(np.repeat(F0[:, :, None, :], 3, axis=2))
        oracle.L.or_flux_resistive(ptr(F3), ptr(U3), ptr(np.ascontiguousarray(faces[dr - 1])), dx, ix, jx, 3, dr)
        got, ref = F3[:, :, 1, :] - F0, d[f"F_{dr}"] - F0
        assert np.abs(ref).max() > 1e-6  # the flux is not trivial here
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("typ", (0, 1, 2, 4))
@pytest.mark.parametrize("dir", (1, 2, 3))
def test_other_flux_types(oracle, typ, dir):
    """LLF, HLL, HLLC and LHLLD (flux_solver.f90 types 0, 1, 2, 4) on the HLLD
    fixture's interface pairs, bitwise to the reference."""
    d0, d = load("hlld"), load("solvers")
    VL, VR = np.asfortranarray(d0[f"VL_{dir}"]), np.asfortranarray(d0[f"VR_{dir}"])
    F = farr(*VL.shape)
    assert oracle.flux_solver(F, VL, VR, float(d0["ch"]), dir, typ) == 0
    assert bits_equal(F, d[f"F_{typ}_{dir}"])


@pytest.mark.parametrize("typ", (0, 1, 2, 4))
@pytest.mark.parametrize("bc", ("periodic", "kh"))
def test_other_flux_types_runs(oracle, typ, bc):
    d = load("solvers")
    U = np.asfortranarray(d[f"run_{typ}_{bc}_U0"]).copy(order="F")
    p = make_problem(22, 20, dx=0.05, cfl=0.4, lm_type=2 if bc == "periodic" else 1, flux_type=typ,
                     bc=("periodic", "wall" if bc == "kh" else "periodic", "periodic"))
    rc, dts, chs = oracle.run(p, U, 8)
    assert rc == 0
    assert bits_equal(dts, d[f"run_{typ}_{bc}_dt"]) and bits_equal(chs, d[f"run_{typ}_{bc}_ch"])
    assert bits_equal(U, d[f"run_{typ}_{bc}_U"])


@pytest.mark.parametrize("typ", (0, 1, 2, 4))
def test_other_flux_types_3d_runs(oracle, typ):
    """3D periodic runs with LLF, HLL, HLLC and LHLLD (3D_basic routines), MC,
    4 steps: bitwise to the flang build."""
    d = load("solvers")
    U = np.asfortranarray(d[f"run3d_{typ}_U0"]).copy(order="F")
    ix, jx, kx = U.shape[:3]
    rc, dts, chs = oracle.run(make_problem(ix, jx, kx, dx=0.1, cfl=0.4, lm_type=2, flux_type=typ), U, 4)
    assert rc == 0
    assert bits_equal(dts, d[f"run3d_{typ}_dt"]) and bits_equal(chs, d[f"run3d_{typ}_ch"])
    assert bits_equal(U, d[f"run3d_{typ}_U"])
