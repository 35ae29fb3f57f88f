"""The 2D resistive reconnection problem (2D_reconnection): the oracle's 2D
CPU form of the resistive flux, the resistive dt limit, the eta(x, y) profile
of the Python mirror, and whole runs of 2D_reconnection/main.f90's loop,
against vectors from the flang build of the reference's own routines
(tests/golden/recon.npz, recon_run.npz) and, where that build is present
here, against the reference driver directly.  CPU only; the GPU side is in
tests/test_gpu_recon.py and tests/test_gpu_0_fortran.py."""
import ctypes as C
import subprocess

import numpy as np
import pytest

from omhd_testlib import DP, REF_DIR, farr, ptr, recon_problem
from test_oracle_golden import bits_equal, load

REF = REF_DIR / "main2d_recon_ref"


@pytest.fixture(scope="module")
def fr2d(oracle):
    f = oracle.L.or_flux_resistive2d
    f.argtypes = [DP, DP, DP, C.c_int, C.c_int, C.c_int]
    f.restype = None
    return f


def test_eta_profile_matches_reference_vectors():
    """eta(x, y) of 2D_reconnection/flux_resistive.f90:1-16 at 506 points
    (random, near the origin, the origin, the r = 25 cap and beyond):
    bitwise."""
    from openmhd_amd.api import eta_recon

    d = load("recon_run")
    got = np.array([eta_recon(a, b) for a, b in d["eta_pts"]])
    assert bits_equal(got, d["eta"])
    assert got.max() == pytest.approx(1.0 / 60.0, rel=1e-15) and got.min() == 1.0 / 1000.0


def test_set_dt_resistive_matches_reference_vectors(oracle):
    """set_dt_resistive (flux_resistive.f90:18-25): bitwise, both when the
    resistive limit binds and when it does not."""
    from openmhd_amd.api import RECON_RM1, set_dt_resistive

    d = load("recon_run")
    f = oracle.L.or_set_dt_resistive2d
    f.argtypes = [C.c_double] * 4
    f.restype = C.c_double
    for (dt, dx, cfl), ref in zip(d["sdr_in"], d["sdr_out"]):
        assert bits_equal(set_dt_resistive(dt, dx, cfl), ref)
        assert bits_equal(f(dt, dx, cfl, RECON_RM1), ref)
    assert (d["sdr_out"] < d["sdr_in"][:, 0]).any() and (d["sdr_out"] == d["sdr_in"][:, 0]).any()


@pytest.mark.parametrize("dir", (1, 2))
def test_flux_resistive_2d_form_matches_reference_vectors(fr2d, dir):
    """The oracle's restatement of the 2D CPU form, with the mirror's face
    tables, against flux_resistive.f90 run by flang on a smooth field near the
    resistivity peak: bitwise."""
    from openmhd_amd.api import eta_face_table

    d = load("recon")
    U, x, y, dx = np.asfortranarray(d["res_U"]), d["res_x"], d["res_y"], float(d["res_dx"])
    ix, jx = U.shape[:2]
    F = np.asfortranarray(d[f"F0_{dir}"]).copy(order="F")
    fr2d(ptr(F), ptr(U), ptr(eta_face_table(x, y, dx, dir)), ix, jx, dir)
    assert np.abs(d[f"F_{dir}"] - d[f"F0_{dir}"]).max() > 1e-6
    assert bits_equal(F, d[f"F_{dir}"])


def test_flux_resistive_2d_known_answers(fr2d):
    """What the mathematics fixes: a uniform field carries no current (F
    untouched, bitwise); By = a x with uniform Bx, Bz gives Jz = a on every
    x-face, so F(by) drops by eta a and F(en) by eta a <By> (to rounding);
    Bz = a y gives Jx = a on every y-face; only the slots of the 2D form
    change, only on the reference's face ranges; dir 3 adds nothing."""
    ix, jx, dx, a = 11, 9, 0.25, 0.7
    X, Y = np.meshgrid(dx * np.arange(ix), dx * np.arange(jx), indexing="ij")
    e = np.asfortranarray(0.01 + 0.001 * np.cos(X + Y)) / dx
    U = farr(ix, jx, 9)
    U[..., 5], U[..., 6], U[..., 7] = 0.3, -0.2, 0.5
    for dr in (1, 2, 3):
        F = np.asfortranarray(np.random.default_rng(dr).normal(size=(ix, jx, 9)))
        F0 = F.copy(order="F")
        fr2d(ptr(F), ptr(U), ptr(e), ix, jx, dr)
        assert bits_equal(F, F0), dr
    U[..., 6] = a * X
    F = farr(ix, jx, 9)
    fr2d(ptr(F), ptr(U), ptr(e), ix, jx, 1)
    eta = e * dx
    jz = slice(0, ix - 1), slice(1, jx - 1)
    np.testing.assert_allclose(F[jz + (6,)], -eta[jz] * a, rtol=1e-12)
    np.testing.assert_allclose(F[jz + (3,)], -0.5 * eta[jz] * a * (U[:-1, 1:-1, 6] + U[1:, 1:-1, 6]), rtol=1e-12,
                               atol=1e-15)
    np.testing.assert_allclose(F[jz + (7,)], 0.0, atol=1e-15)
    touched = np.zeros((ix, jx), bool)
    touched[jz] = True
    assert not F[~touched].any() and not F[..., [0, 1, 2, 4, 5, 8]].any()
    U[..., 6] = -0.2
    U[..., 7] = a * Y
    F = farr(ix, jx, 9)
    fr2d(ptr(F), ptr(U), ptr(e), ix, jx, 2)
    jy = slice(1, ix - 1), slice(0, jx - 1)
    np.testing.assert_allclose(F[jy + (7,)], -eta[jy] * a, rtol=1e-12)
    np.testing.assert_allclose(F[jy + (5,)], 0.0, atol=1e-15)


@pytest.mark.parametrize("case", ("a", "b"))
def test_oracle_run_matches_reference_driver_vectors(oracle, case):
    """The oracle's step in the form of 2D_reconnection/main.f90 against the
    project's driver linked by flang with the reference's routines (case a:
    CFL-limited; case b: dx = 0.01, limited by set_dt_resistive): dt, ch and
    the state after 8 steps, bitwise."""
    d = load("recon_run")
    lm, tt, flux, steps = (int(v) for v in d[f"{case}_case"])
    p, tab = recon_problem(d[f"{case}_x"], d[f"{case}_y"], lm_type=lm, flux_type=flux, time_type=tt)
    U = np.asfortranarray(d[f"{case}_U0"]).copy(order="F")
    rc, dts, chs = oracle.run(p, U, steps)
    assert rc == 0
    assert bits_equal(dts, d[f"{case}_dt"]) and bits_equal(chs, d[f"{case}_ch"])
    assert bits_equal(U, d[f"{case}_U"])
    if case == "b":
        assert (dts == 0.5 * 0.35 * 60.0 * (float(p.dx) ** 2)).all()


def run_recon_driver(exe, tmp_path, oracle, nx, ny, lx, steps, lm, tt, flux):
    """Run a main2d_recon build on a Harris sheet and the oracle on the same
    state; returns (driver dt, oracle dt, driver state, oracle state)."""
    from openmhd_amd import problems, snapshot

    U, _, x, y = problems.harris_sheet_2d(nx, ny, lx=lx)
    p, tab = recon_problem(x, y, lm_type=lm, flux_type=flux, time_type=tt)
    oracle.bc_u(U, p)
    fin, fout, tr = tmp_path / "in.dat", tmp_path / "out.dat", tmp_path / "trace.bin"
    snapshot.write(fin, 0.0, x, y, U)
    cmd = [str(exe), str(nx), str(ny), str(steps), "--input", str(fin), "--output", str(fout), "--trace", str(tr),
           "--lm", str(lm), "--tt", str(tt), "--flux", str(flux)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _, _, _, Ug, _ = snapshot.read(fout)
    dts = np.fromfile(tr)[:steps]
    Uo = U.copy(order="F")
    rc, dto, _ = oracle.run(p, Uo, steps)
    assert rc == 0
    return dts, dto, Ug, Uo


RUNS = [(64, 24, 12.8, 12, 1, 0, 3), (48, 20, 0.48, 10, 2, 1, 3), (40, 16, 8.0, 8, 3, 0, 1), (40, 16, 8.0, 8, 4, 1, 2),
        (40, 16, 8.0, 8, 0, 0, 0)]


@pytest.mark.skipif(not REF.exists(), reason="flang reference build absent (make -C oracle ref)")
@pytest.mark.parametrize("nx,ny,lx,steps,lm,tt,flux", RUNS)
def test_reference_recon_driver_matches_oracle(tmp_path, oracle, nx, ny, lx, steps, lm, tt, flux):
    dts, dto, Ur, Uo = run_recon_driver(REF, tmp_path, oracle, nx, ny, lx, steps, lm, tt, flux)
    assert bits_equal(dts, dto)
    assert bits_equal(Ur, Uo)
