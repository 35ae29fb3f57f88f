"""The 1D shock tube of 1D_basic (Brio & Wu 1988).

The oracle's 1D loop (or_run1d, 1D_basic/main.f90:56-149) is checked bitwise
against the reference's own 1D_basic routines built by flang under the
project's 1D driver (openmhd_amd/fortran/main1d.f90 -> oracle/_ref/main1d_ref),
for every flux type and several limiters and integrators.  Known answers that
do not depend on any implementation pin it further: what the conservation
form and the outflow ghosts fix while the waves are inside the domain, the
fields the coplanar problem never changes, the quiet far field, and the
mirror symmetry of a symmetric tube."""
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from openmhd_amd import problems, snapshot
from test_oracle_golden import bits_equal

REPO = Path(__file__).resolve().parents[1]
REF1D = REPO / "oracle" / "_ref" / "main1d_ref"
MX, MY, MZ, EN, RO, BX, BY, BZ, PS = range(9)
GAM = 5.0 / 3.0
need_ref = pytest.mark.skipif(not REF1D.exists(), reason="flang reference build absent (make -C oracle ref)")

# (flux type, limiter, time type): every flux type, every limiter, both integrators
CASES = [(3, 1, 0), (3, 2, 1), (3, 0, 0), (0, 1, 0), (1, 3, 0), (2, 4, 1), (4, 2, 0), (2, 1, 0)]


def run_driver(exe, tmp_path, *args):
    out, tr = tmp_path / "out.dat", tmp_path / "trace.bin"
    cmd = [str(exe), *map(str, args), "--output", str(out), "--trace", str(tr)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "steps" in r.stdout, r.stdout + r.stderr
    t, x, _, U, V = snapshot.read(out)
    return U, V, t, x, np.fromfile(tr)


def fast_speed(ro, p, bx, by):
    """Fast magnetosonic speed along x in closed form."""
    a = (GAM * p + bx * bx + by * by) / ro
    return math.sqrt(0.5 * (a + math.sqrt(a * a - 4 * GAM * p * bx * bx / (ro * ro))))


@need_ref
def test_initial_state_is_the_drivers(tmp_path):
    U, _, t, x, dts = run_driver(REF1D, tmp_path, "--steps", 1)
    U0, dx, x0 = problems.brio_wu(300)
    assert t == 0.0 and len(dts) == 0
    assert bits_equal(x, x0)
    assert bits_equal(U, U0)


@need_ref
@pytest.mark.parametrize("flux,lm,tt", CASES)
def test_oracle_matches_the_flang_build(tmp_path, oracle, flux, lm, tt):
    U, V, t, _, dts = run_driver(REF1D, tmp_path, "--flux", flux, "--lm", lm, "--tt", tt)
    Uo, dx, _ = problems.brio_wu(300)
    rc, to, dto = oracle.run1d(Uo, dx, 0.8, lm, flux, tt, 0.2)
    assert rc == 0
    assert len(dto) == len(dts) > 100
    assert bits_equal(dto, dts)
    assert to == t
    assert bits_equal(Uo, U)
    Vo = np.zeros((Uo.shape[0], 1, 4), order="F")
    assert oracle.u2v(Uo, Vo)[0] == 0
    assert bits_equal(Vo, V)


# second-order reconstructions: the first-order (flat) scheme's diffusion
# reaches the outflow faces by t = 0.1, which the conservation checks exclude
KNOWN = [c for c in CASES if c[1] > 0]


@pytest.mark.parametrize("flux,lm,tt", KNOWN)
def test_brio_wu_known_answers(oracle, flux, lm, tt):
    n, tend = 400, 0.1
    U, dx, x = problems.brio_wu(n)
    U0 = U.copy(order="F")
    rc, t, dts = oracle.run1d(U, dx, 0.8, lm, flux, tt, tend)
    assert rc == 0 and t >= tend and t - dts[-1] < tend
    inner = slice(1, n + 1)
    s = lambda A, m: float(np.sum(A[inner, 0, m]))  # noqa: E731
    # The fast fronts stay inside the domain, so the outflow faces see the
    # initial states: zero mass, energy and By flux; x-momentum flux
    # p + B^2/2 - Bx^2 on each side, whose difference is p_L - p_R = 0.9;
    # y-momentum flux -Bx By, whose difference is -2 Bx By_L = -1.5.
    for m in (RO, EN, BY):
        assert abs(s(U, m) - s(U0, m)) <= 1e-13 * max(1.0, abs(s(U0, m)))
    assert abs(s(U, MX) * dx - 0.9 * t) <= 1e-13
    assert abs(s(U, MY) * dx + 1.5 * t) <= 1e-13
    # ch = 0 and a coplanar field: Bx, psi, vz and Bz never change
    assert np.all(U[:, 0, BX] == 0.75)
    assert np.all(U[:, 0, PS] == 0.0)
    assert np.all(U[:, 0, MZ] == 0.0) and np.all(U[:, 0, BZ] == 0.0)
    # nothing moves ahead of the fast rarefactions (closed-form speeds) beyond
    # the scheme's numerical diffusion
    cl, cr = fast_speed(1.0, 1.0, 0.75, 1.0), fast_speed(0.125, 0.1, 0.75, -1.0)
    assert 1.70 < cl < 1.71 and 3.65 < cr < 3.66
    far = (x < -cl * t - 0.05) | (x > cr * t + 0.05)
    assert far.sum() > 40
    assert np.abs(U[far] - U0[far]).max() <= 1e-5
    near = (x > -cl * t + 0.02) & (x < cr * t - 0.02)
    assert np.all(np.any(U[near] != U0[near], axis=-1))
    # the solution is physical: positive density and pressure
    V = np.zeros((n + 2, 1, 4), order="F")
    assert oracle.u2v(U, V)[0] == 0
    assert U[:, 0, RO].min() > 0.1 and V[:, 0, 3].min() > 0.05


@pytest.mark.parametrize("flux", (0, 1, 2, 3, 4))
def test_symmetric_tube_stays_mirror_symmetric(oracle, flux):
    """rho = p = 1, Bx = 0.75 and By = +1 | -1: under x -> -x the equations
    keep rho, p, Bx, vy and send vx, By to minus themselves, so the solution
    keeps that symmetry; the grid is symmetric about x = 0.  A flipped index
    or sign in the x sweep breaks it at O(1)."""
    n = 200
    U, dx, x = problems.brio_wu(n)
    left = x < 0
    ro = np.ones((n + 2, 1))
    by = np.where(left, 1.0, -1.0)[:, None]
    vx = np.where(left, 0.3, -0.3)[:, None]  # colliding flows
    z = np.zeros((n + 2, 1))
    U = problems.prim_to_cons_fortran(ro, vx, z, z, ro, z + 0.75, by, z, z)
    rc, t, _ = oracle.run1d(U, dx, 0.8, 1, flux, 0, 0.1)
    assert rc == 0
    R = U[::-1]
    for m, sgn in ((RO, 1), (EN, 1), (BX, 1), (MX, -1), (BY, -1), (MY, 1)):
        d = np.abs(U[:, 0, m] - sgn * R[:, 0, m]).max()
        assert d <= 1e-12, (m, d)
    assert np.abs(U[:, 0, RO] - 1.0).max() > 0.1  # the shocks formed
