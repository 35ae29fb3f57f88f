"""The project's 3D reference-style driver (openmhd_amd/fortran/main3d.f90,
the loop of 3D_basic/main.f90) linked with flang against the reference's own
3D_basic routines (oracle/_ref/main3d_ref) against the oracle: dt sequence
and state bitwise.  The same driver linked against the libomhd 3D shims is
checked in tests/test_gpu_0_fortran.py."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from omhd_testlib import make_problem, smooth_state
from test_oracle_golden import bits_equal

REPO = Path(__file__).resolve().parents[1]
REF3D = REPO / "oracle" / "_ref" / "main3d_ref"
CASES = [(2, 0, 3), (1, 1, 3), (2, 0, 1), (4, 1, 2)]  # (limiter, time type, flux type)


def run_main3d(exe, tmp_path, oracle, lm, tt, flux, n=(14, 12, 10), steps=4, seed=7):
    from openmhd_amd import snapshot

    dx = 0.1
    U = smooth_state(np.random.default_rng(seed), n[0] + 2, n[1] + 2, n[2] + 2)
    p = make_problem(n[0] + 2, n[1] + 2, n[2] + 2, dx=dx, cfl=0.4, lm_type=lm, time_type=tt, flux_type=flux)
    oracle.bc_u(U, p)
    fin, fout, tr = tmp_path / "in.dat", tmp_path / "out.dat", tmp_path / "trace.bin"
    snapshot.write(fin, 0.0, snapshot.grid(n[0], dx, 0.0), snapshot.grid(n[1], dx, 0.0), U,
                   z=snapshot.grid(n[2], dx, 0.0))
    cmd = [str(exe), *map(str, n), str(steps), "--input", str(fin), "--output", str(fout), "--trace", str(tr),
           "--lm", str(lm), "--tt", str(tt), "--flux", str(flux)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _, _, _, _, Ug, _ = snapshot.read(fout, three_d=True)
    dts = np.fromfile(tr)[:steps]
    Uo = U.copy(order="F")
    rc, dto, _ = oracle.run(p, Uo, steps)
    assert rc == 0
    return dts, dto, Ug, Uo


@pytest.mark.skipif(not REF3D.exists(), reason="flang reference build absent (make -C oracle ref)")
@pytest.mark.parametrize("lm,tt,flux", CASES)
def test_reference_3d_driver_matches_oracle(tmp_path, oracle, lm, tt, flux):
    dts, dto, Ur, Uo = run_main3d(REF3D, tmp_path, oracle, lm, tt, flux)
    assert bits_equal(dts, dto)
    assert bits_equal(Ur, Uo)
