"""The Fortran side of the boundary: the project's reference-style drivers
(openmhd_amd/fortran/main1d.f90, main2d.f90, main2d_recon.f90, main3d.f90)
linked against the libomhd shims,
and the resident-step driver, all against the oracle on the same initial
state.  Bitwise."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from omhd_testlib import make_problem
from test_oracle_golden import bits_equal

pytestmark = pytest.mark.gpu

REPO = Path(__file__).resolve().parents[1]
FDIR = REPO / "openmhd_amd" / "fortran"


def _snapshot(tmp_path, oracle, kh, n):
    # ghosts are filled on the CPU: this module runs the drivers as child
    # processes before anything in this process has touched the GPU
    from openmhd_amd import problems, snapshot

    p = make_problem(n + 2, n + 2, bc=("periodic", "wall" if kh else "periodic", "periodic"))
    if kh:
        U, dx = problems.kelvin_helmholtz(n, n, seed=3)
        x = snapshot.grid(n, dx, 0.0)
        y = snapshot.grid(n, dx, -0.5 * n * dx)
    else:
        U, dx = problems.orszag_tang(n)
        x = snapshot.grid(n, dx, 0.0)
        y = x.copy()
    oracle.bc_u(U, p)
    f = tmp_path / "in.dat"
    snapshot.write(f, 0.0, x, y, U)
    return f, U, x[1] - x[0]


@pytest.mark.parametrize("exe", ("main2d_gpu", "main2d_resident"))
@pytest.mark.parametrize("kh", (False, True))
def test_driver_matches_oracle(tmp_path, oracle, exe, kh):
    from openmhd_amd import snapshot

    n, steps = 40, 12
    fin, U, dx = _snapshot(tmp_path, oracle, kh, n)
    out, tr = tmp_path / "out.dat", tmp_path / "trace.bin"
    cmd = [str(FDIR / exe), "--kh" if kh else "--ot", str(n), str(steps), "--input", str(fin), "--output", str(out),
           "--trace", str(tr)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _, _, _, Ug, _ = snapshot.read(out)
    dts = np.fromfile(tr)[:steps]
    p = make_problem(n + 2, n + 2, dx=dx, cfl=0.4, lm_type=1 if kh else 2,
                     bc=("periodic", "wall" if kh else "periodic", "periodic"))
    Uo = U.copy(order="F")
    rc, dto, _ = oracle.run(p, Uo, steps)
    assert rc == 0
    assert bits_equal(dts, dto)
    assert bits_equal(Ug, Uo)


# the 1D shock tube (1D_basic/main.f90) through the drop-in shims: every flux
# type, every limiter, both integrators, to t = 0.2 (about 300 steps)
@pytest.mark.parametrize("flux,lm,tt", [(3, 1, 0), (3, 2, 1), (3, 0, 0), (0, 1, 0), (1, 3, 0), (2, 4, 1), (4, 2, 0)])
def test_driver_1d_shock_tube_matches_oracle(tmp_path, oracle, flux, lm, tt):
    from openmhd_amd import problems, snapshot

    out, tr = tmp_path / "out.dat", tmp_path / "trace.bin"
    cmd = [str(FDIR / "main1d_gpu"), "--flux", str(flux), "--lm", str(lm), "--tt", str(tt), "--output", str(out),
           "--trace", str(tr)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "steps" in r.stdout, r.stdout + r.stderr
    t, _, _, Ug, Vg = snapshot.read(out)
    dts = np.fromfile(tr)
    Uo, dx, _ = problems.brio_wu(300)
    rc, to, dto = oracle.run1d(Uo, dx, 0.8, lm, flux, tt, 0.2)
    assert rc == 0
    assert len(dts) == len(dto) > 100
    assert bits_equal(dts, dto)
    assert t == to
    assert bits_equal(Ug, Uo)
    Vo = np.zeros((Uo.shape[0], 1, 4), order="F")
    assert oracle.u2v(Uo, Vo)[0] == 0
    assert bits_equal(Vg, Vo)


# the 3D driver (3D_basic/main.f90's loop) through the 3D shims
@pytest.mark.parametrize("lm,tt,flux", [(2, 0, 3), (1, 1, 3), (2, 0, 1), (4, 1, 2)])
def test_driver_3d_matches_oracle(tmp_path, oracle, lm, tt, flux):
    from test_fortran3d import run_main3d

    dts, dto, Ug, Uo = run_main3d(FDIR / "main3d_gpu", tmp_path, oracle, lm, tt, flux)
    assert bits_equal(dts, dto)
    assert bits_equal(Ug, Uo)


# the 2D reconnection driver (2D_reconnection/main.f90's loop) through the 2D
# shims and omhd_shims_recon.f90: per-face reflecting tables, eta(x, y) face
# tables, the resistive flux after each flux_solver call, set_dt_resistive
@pytest.mark.parametrize("nx,ny,lx,steps,lm,tt,flux", [(64, 24, 12.8, 12, 1, 0, 3), (48, 20, 0.48, 10, 2, 1, 3),
                                                      (40, 16, 8.0, 8, 3, 0, 1), (40, 16, 8.0, 8, 4, 1, 2),
                                                      (40, 16, 8.0, 8, 0, 0, 0), (200, 50, 40.0, 40, 1, 0, 3)])
def test_driver_recon_matches_oracle(tmp_path, oracle, nx, ny, lx, steps, lm, tt, flux):
    from test_recon2d import run_recon_driver

    dts, dto, Ug, Uo = run_recon_driver(FDIR / "main2d_recon_gpu", tmp_path, oracle, nx, ny, lx, steps, lm, tt, flux)
    assert bits_equal(dts, dto)
    assert bits_equal(Ug, Uo)


@pytest.mark.parametrize("case", ("a", "b"))
def test_driver_recon_matches_reference_vectors(tmp_path, case):
    """main2d_recon linked against the shims against the same driver linked
    by flang with the reference's routines (tests/golden/recon_run.npz)."""
    from openmhd_amd import snapshot
    from test_oracle_golden import load

    d = load("recon_run")
    lm, tt, flux, steps = (int(v) for v in d[f"{case}_case"])
    U0 = d[f"{case}_U0"]
    fin, fout, tr = tmp_path / "in.dat", tmp_path / "out.dat", tmp_path / "trace.bin"
    snapshot.write(fin, 0.0, d[f"{case}_x"], d[f"{case}_y"], np.asfortranarray(U0))
    cmd = [str(FDIR / "main2d_recon_gpu"), str(U0.shape[0] - 2), str(U0.shape[1] - 2), str(steps), "--input", str(fin),
           "--output", str(fout), "--trace", str(tr), "--lm", str(lm), "--tt", str(tt), "--flux", str(flux)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _, _, _, Ug, _ = snapshot.read(fout)
    t = np.fromfile(tr)
    assert bits_equal(t[:steps], d[f"{case}_dt"]) and bits_equal(t[steps:], d[f"{case}_ch"])
    assert bits_equal(Ug, d[f"{case}_U"])


# the 3D reconnection driver (3D_reconnection_gpu/main.cuf's loop as host
# calls) through omhd_shims3d.f90 and omhd_shims3d_recon.f90, against the
# oracle's 3D resistive step with the same tables, eta and dt limit
def run_recon3d_driver(tmp_path, oracle, n, dx, steps, lm, tt, flux, seed=11):
    from omhd_testlib import recon3d_problem

    from openmhd_amd import problems, snapshot

    nx, ny, nz = n
    U, (x0, y0) = problems.harris_sheet_3d(nx, ny, nz, dx, seed=seed, noise=1e-2)
    x = snapshot.grid(nx, dx, x0)
    y = snapshot.grid(ny, dx, y0)
    z = snapshot.grid(nz, dx, 0.0)
    p, tab = recon3d_problem(x, y, nz, lm_type=lm, flux_type=flux, time_type=tt)
    oracle.bc_u(U, p)
    fin, fout, tr = tmp_path / "in.dat", tmp_path / "out.dat", tmp_path / "trace.bin"
    snapshot.write(fin, 0.0, x, y, U, z=z)
    cmd = [str(FDIR / "main3d_recon_gpu"), *map(str, n), str(steps), "--input", str(fin), "--output", str(fout),
           "--trace", str(tr), "--lm", str(lm), "--tt", str(tt), "--flux", str(flux)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _, _, _, _, Ug, _ = snapshot.read(fout, three_d=True)
    dts = np.fromfile(tr)[:steps]
    Uo = U.copy(order="F")
    rc, dto, _ = oracle.run(p, Uo, steps)
    assert rc == 0
    return dts, dto, Ug, Uo, float(p.dx)


@pytest.mark.parametrize("n,dx,steps,lm,tt,flux", [((20, 16, 8), 0.25, 5, 1, 0, 3), ((18, 14, 6), 0.01, 4, 2, 1, 3),
                                                   ((16, 12, 6), 0.2, 4, 1, 0, 1)])
def test_driver_recon3d_matches_oracle(tmp_path, oracle, n, dx, steps, lm, tt, flux):
    dts, dto, Ug, Uo, dxg = run_recon3d_driver(tmp_path, oracle, n, dx, steps, lm, tt, flux)
    assert bits_equal(dts, dto)
    assert bits_equal(Ug, Uo)
    if dx < 0.05:  # the resistive limit sets dt (flux_resistive.cuf:8)
        assert (dts == 0.5 * 0.3 * 60.0 * (dxg * dxg)).all()


# The reference's own programs, unmodified (1D_basic and 2D_basic: main.f90,
# model.f90 and the output routine), built by oracle/Makefile with the
# reference's routines (refprog_*_cpu) and with the libomhd shims
# (refprog_*_gpu).  2D_basic is the Orszag-Tang run to t = 4 (about 840 steps,
# 41 snapshot files in the fileio.f90 format).
REFPROG = REPO / "oracle" / "_ref"


def _big_stack():  # set_dt/glm_ss2 keep automatic arrays on the stack
    import resource

    resource.setrlimit(resource.RLIMIT_STACK, (resource.getrlimit(resource.RLIMIT_STACK)[1],) * 2)


@pytest.mark.parametrize("ex", ("1d_basic", "2d_basic"))
def test_reference_program_unchanged_on_the_gpu(tmp_path, ex):
    """The same program writes the same files, byte for byte, and the same
    standard output, whether it is linked against the reference's routines
    (on the CPU) or against the drop-in shims (on the GPU)."""
    import filecmp

    cpu, gpu = REFPROG / f"refprog_{ex}_cpu", REFPROG / f"refprog_{ex}_gpu"
    if not (cpu.exists() and gpu.exists()):
        pytest.skip("reference programs not built (make -C oracle ref)")
    runs = {}
    for name, exe in (("cpu", cpu), ("gpu", gpu)):
        d = tmp_path / name
        (d / "data").mkdir(parents=True)
        print(f"reference program {ex} ({name})", flush=True)
        r = subprocess.run([str(exe)], cwd=d, capture_output=True, text=True, timeout=240, preexec_fn=_big_stack)
        assert r.returncode == 0 and "== end ==" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
        runs[name] = (r.stdout, sorted(p.name for p in (d / "data").iterdir()))
    assert runs["cpu"][0] == runs["gpu"][0]
    files = runs["cpu"][1]
    assert files == runs["gpu"][1] and len(files) >= 3
    for f in files:
        assert filecmp.cmp(tmp_path / "cpu" / "data" / f, tmp_path / "gpu" / "data" / f, shallow=False), f


def test_reference_kh_program_resident(tmp_path):
    """2D_KH's main.f90 and model.f90, unmodified, linked against the drop-in
    shims and the fileio shims (openmhd_amd/fortran/omhd_fileio.f90) and run
    with OMHD_MIRROR=1: the arrays stay in HBM from call to call and reach the
    host only at the snapshots.  The standard output and all 21 snapshot files
    of the run to t = 100 (about 8060 steps) are byte for byte those of the
    program built with the reference's routines (hashes in
    tests/golden/kh_refprog.json, from oracle/_ref/refprog_2d_kh_cpu), and the
    run takes less than a third of that serial CPU build's 142 s measured in
    the build container.  (The reference's OpenMP build of this program stops
    after two steps: flux_solver.f90 declares FL/FR private but sets their
    normal-field entries only outside the parallel region.)"""
    import hashlib
    import json
    import os
    import time

    exe = REFPROG / "refprog_2d_kh_mirror"
    if not exe.exists():
        pytest.skip("reference programs not built (make -C oracle ref)")
    gold = json.loads((REPO / "tests" / "golden" / "kh_refprog.json").read_text())
    d = tmp_path / "gpu"
    (d / "data").mkdir(parents=True)
    t0 = time.perf_counter()
    r = subprocess.run([str(exe)], cwd=d, capture_output=True, text=True, timeout=240, preexec_fn=_big_stack,
                       env=dict(os.environ, OMHD_MIRROR="1"))
    el = time.perf_counter() - t0
    print(f"2D_KH resident: {el:.1f} s", flush=True)
    assert r.returncode == 0 and "== end ==" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
    assert hashlib.sha256(r.stdout.encode()).hexdigest() == gold["stdout_sha256"]
    files = sorted(p.name for p in (d / "data").iterdir())
    assert files == sorted(gold["files"]) and len(files) == 21
    for f in files:
        assert hashlib.sha256((d / "data" / f).read_bytes()).hexdigest() == gold["files"][f], f
    assert el < 142.0 / 3
