"""The resident fused step (Solver2D / omhd_step) against the oracle and the
reference's golden runs.  Bitwise: identical dt sequence and state."""
import hashlib

import numpy as np
import pytest

from omhd_testlib import make_problem, orszag_tang, smooth_state
from test_oracle_golden import bits_equal, load

pytestmark = pytest.mark.gpu


def solver(nx, ny, dx, lm=2, tt=0, kh=False, rows=0):
    from openmhd_amd import Solver2D

    return Solver2D(nx, ny, dx, cfl=0.4, lm_type=lm, time_type=tt,
                    bc=("periodic", "wall" if kh else "periodic"), rows_per_wave=rows)


def run_both(oracle, U, nsteps, dx, lm=2, tt=0, kh=False, rows=0):
    ix, jx = U.shape[:2]
    p = make_problem(ix, jx, dx=dx, cfl=0.4, lm_type=lm, time_type=tt,
                     bc=("periodic", "wall" if kh else "periodic", "periodic"))
    oracle.bc_u(U, p)
    s = solver(ix - 2, jx - 2, dx, lm, tt, kh, rows)
    s.upload(U)
    dts, chs = s.step(nsteps, trace=True)
    Ug = s.download()
    s.close()
    Uo = U.copy(order="F")
    rc, dto, cho = oracle.run(p, Uo, nsteps)
    assert rc == 0
    return dts, chs, Ug, dto, cho, Uo


@pytest.mark.parametrize("lm", range(5))
@pytest.mark.parametrize("tt", (0, 1))
def test_every_limiter_and_time_type(oracle, lm, tt):
    rng = np.random.default_rng(lm + 10 * tt)
    U = smooth_state(rng, 75, 47)  # partial strips and chunks
    dts, chs, Ug, dto, cho, Uo = run_both(oracle, U, 6, 0.05, lm, tt)
    assert bits_equal(dts, dto) and bits_equal(chs, cho)
    assert bits_equal(Ug, Uo)


@pytest.mark.parametrize("rows", (1, 3, 16, 64))
@pytest.mark.parametrize("kh", (False, True))
def test_chunking_and_walls(oracle, rows, kh):
    rng = np.random.default_rng(rows)
    U = smooth_state(rng, 130, 70)
    dts, chs, Ug, dto, cho, Uo = run_both(oracle, U, 8, 0.05, 1, 0, kh, rows)
    assert bits_equal(dts, dto)
    assert bits_equal(Ug, Uo)


def test_golden_limiter_runs():
    d = load("limiters_run")
    for lm in range(5):
        U = np.asfortranarray(d[f"U0_{lm}"]).copy(order="F")
        s = solver(18, 16, 0.05, lm)
        s.upload(U)
        dts, _ = s.step(10, trace=True)
        Ug = s.download()
        assert bits_equal(dts, d[f"dt_{lm}"]), lm
        assert bits_equal(Ug[1:-1, 1:-1], d[f"U_{lm}"][1:-1, 1:-1]), lm
        assert bits_equal(Ug, d[f"U_{lm}"]), lm


def test_golden_kh_walls():
    d = load("kh128_minmod_100")
    U = np.asfortranarray(d["U0"]).copy(order="F")
    s = solver(128, 128, float(d["dx"]), 1, 0, kh=True)
    s.upload(U)
    dts, chs = s.step(100, trace=True)
    Ug = s.download()
    assert bits_equal(dts, d["dt"]) and bits_equal(chs, d["ch"])
    assert bits_equal(Ug, d["final"])


def test_orszag_tang_1000_steps():
    """OT 200^2, MC, 1000 steps: dt sequence and final state
    bitwise equal to the reference CPU build.  The acceptance criteria stated
    for this case are also checked as such: per-field relative error <= 1e-11
    (on the stored sample of the reference's final state) and ||div B||_2
    within 1% of the reference's."""
    d = load("ot200_mc_1000")
    U, dx = orszag_tang(200)
    from openmhd_amd import api

    api.bc(U, 202, 202)
    assert hashlib.sha256(U.tobytes(order="F")).hexdigest() == str(d["u0_sha256"])
    s = solver(200, 200, dx, 2)
    s.upload(U)
    dts, chs = s.step(1000, trace=True)
    Ug = s.download()
    assert bits_equal(dts, d["dt"]) and bits_equal(chs, d["ch"])
    interior = np.ascontiguousarray(Ug[1:-1, 1:-1, :].transpose(2, 1, 0))
    assert bits_equal(interior[:, ::7, ::7], d["sample"])
    assert hashlib.sha256(interior.tobytes()).hexdigest() == str(d["final_sha256"])
    from openmhd_amd.diagnostics import div_b_l2

    smp = interior[:, ::7, ::7]
    rel = np.abs(smp - d["sample"]).max(axis=(1, 2)) / np.maximum(np.abs(d["sample"]).max(axis=(1, 2)), 1e-300)
    divb = div_b_l2(Ug, dx)
    print(f"OT 1000 steps: max per-field relative error {rel.max():.3e}; ||div B||_2 {divb!r} "
          f"(reference {float(d['divb_l2'])!r})")
    assert (rel <= 1e-11).all()
    assert abs(divb - float(d["divb_l2"])) <= 0.01 * float(d["divb_l2"])


def test_kh_4096_two_steps_vs_oracle(oracle):
    """The KH grid of the benchmark (4096^2, minmod, walls): two full steps on the GPU
    equal the oracle's bitwise at full size."""
    from openmhd_amd.problems import kelvin_helmholtz

    U, dx = kelvin_helmholtz(4096, 4096, seed=1)
    dts, chs, Ug, dto, cho, Uo = run_both(oracle, U, 2, dx, 1, 0, kh=True)
    assert bits_equal(dts, dto)
    assert bits_equal(Ug, Uo)


def test_negative_pressure_is_reported():
    from openmhd_amd import OMHDError

    rng = np.random.default_rng(0)
    U = smooth_state(rng, 34, 34)
    U[10, 12, 3] = 0.0  # E = 0 -> p < 0 at Fortran cell (11, 13)
    s = solver(32, 32, 0.05)
    with pytest.raises(OMHDError) as e:
        s.upload(U)
    assert e.value.code == 1 and e.value.cell[:2] == (11, 13)


@pytest.mark.parametrize("rows", (5, 24))
def test_reconnection_boundary_tables(oracle, rows):
    """2D_reconnection/bc.f90's per-face tables: x conducting walls, a south
    symmetry plane and a north wall that negate different slots (the oracle's
    tables are pinned to that file's bc_for_U/F/G): dt sequence and state
    bitwise."""
    from omhd_testlib import RECON_SIGN

    from openmhd_amd import Solver

    rng = np.random.default_rng(77)
    U = smooth_state(rng, 70, 38)
    nx, ny = 68, 36
    p = make_problem(nx + 2, ny + 2, dx=0.05, cfl=0.4, lm_type=2, bc=("wall", "wall", "periodic"), signs=RECON_SIGN)
    oracle.bc_u(U, p)
    s = Solver(nx, ny, 0.05, cfl=0.4, lm_type=2, bc=("wall", "wall"), wall_sign=RECON_SIGN, rows_per_wave=rows)
    s.upload(U)
    dts, chs = s.step(6, trace=True)
    Ug = s.download()
    s.close()
    Uo = U.copy(order="F")
    rc, dto, cho = oracle.run(p, Uo, 6)
    assert rc == 0
    assert bits_equal(dts, dto) and bits_equal(chs, cho) and bits_equal(Ug, Uo)


def test_restart_from_snapshot_equals_straight_run(tmp_path):
    """Run 7 steps, write the state in the reference's snapshot format, read
    it into a new solver and run 5 more: the dt sequence and the final state
    equal a straight 12-step run bitwise (the reference restarts through the
    same files, 2D_KH_gpu/main.cuf:82-105)."""
    from openmhd_amd import Solver, problems, snapshot

    U0, dx = problems.kelvin_helmholtz(96, 64, seed=3)
    kw = dict(cfl=0.4, lm_type=1, bc=("periodic", "wall"))
    a = Solver(96, 64, dx, **kw)
    a.upload(U0)
    dts_a, _ = a.step(12, trace=True)
    Ua = a.download()
    a.close()
    b = Solver(96, 64, dx, **kw)
    b.upload(U0)
    dts_b1, _ = b.step(7, trace=True)
    Ub, Vb = b.download(with_v=True)
    f = tmp_path / "field-00007.dat"
    snapshot.write(f, b.t, snapshot.grid(96, dx, 0.0), snapshot.grid(64, dx, -32 * dx), Ub, V=Vb)
    b.close()
    t, _, _, Ur, _ = snapshot.read(f)
    c = Solver(96, 64, dx, **kw)
    c.upload(Ur)
    dts_c, _ = c.step(5, trace=True)
    Uc = c.download()
    c.close()
    assert bits_equal(np.concatenate([dts_b1, dts_c]), dts_a)
    assert bits_equal(Uc, Ua)


@pytest.mark.parametrize("typ", (0, 1, 2, 4))
@pytest.mark.parametrize("bc", ("periodic", "kh"))
def test_other_flux_types_resident(typ, bc):
    """The fused 2D step with LLF, HLL, HLLC or LHLLD: 8 steps, dt sequence and
    state bitwise to the reference's run (periodic MC; KH walls minmod)."""
    from openmhd_amd import Solver

    d = load("solvers")
    U0 = np.asfortranarray(d[f"run_{typ}_{bc}_U0"])
    s = Solver(20, 18, 0.05, cfl=0.4, lm_type=2 if bc == "periodic" else 1, flux_type=typ,
               bc=("periodic", "wall" if bc == "kh" else "periodic"), rows_per_wave=7)
    s.upload(U0)
    dts, chs = s.step(8, trace=True)
    Ug = s.download()
    s.close()
    assert bits_equal(dts, d[f"run_{typ}_{bc}_dt"]) and bits_equal(chs, d[f"run_{typ}_{bc}_ch"])
    assert bits_equal(Ug, d[f"run_{typ}_{bc}_U"])


# tiny and ragged grids: sides of 2 (less than the two-cell halo), one past a
# 60-column strip, chunks of one row, a single strip of 3 columns
@pytest.mark.parametrize("nx,ny", ((2, 2), (3, 5), (61, 2), (2, 61), (119, 7), (60, 3)))
@pytest.mark.parametrize("kh", (False, True))
def test_tiny_and_ragged_grids(oracle, nx, ny, kh):
    rng = np.random.default_rng(nx * 131 + ny)
    U = smooth_state(rng, nx + 2, ny + 2)
    dts, chs, Ug, dto, cho, Uo = run_both(oracle, U, 4, 0.05, 2, 0, kh=kh)
    assert bits_equal(dts, dto) and bits_equal(chs, cho)
    assert bits_equal(Ug, Uo)
