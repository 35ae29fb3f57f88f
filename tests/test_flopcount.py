"""The algorithmic floating-point operation count of the 2D step, from the
forward forms of openmhd_amd/csrc/omhd_adjoint.h run on a counting scalar
(tests/native/flopcount.cpp).  The counter is pinned against hand counts of
the reference's formulas; the per-zone-update figures are measurements."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from omhd_testlib import NATIVE_DIR

LIB = NATIVE_DIR / "libflopcount.so"
LP = C.POINTER(C.c_long)


@pytest.fixture(scope="module")
def fc():
    if not LIB.exists():
        pytest.skip("tests/native/libflopcount.so not built (make -C tests/native)")
    L = C.CDLL(str(LIB))
    L.flops_unit.argtypes = [C.c_int, C.c_int, LP]
    L.flops_unit.restype = None
    L.flops_2d_step.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int, C.c_double, LP, LP]
    L.flops_2d_step.restype = None
    return L


def unit(fc, which, lm=0):
    out = (C.c_long * 4)()
    fc.flops_unit(which, lm, out)
    return tuple(out)


def test_counter_against_hand_counts(fc):
    """u2v.f90:23-26: 3 divides (m/rho), m.v and B.B (3 mul + 2 add each), then
    p = (gam-1)(E - 0.5 (rv2 + B2)): 2 add + 2 mul.  minmod (limiter.f90:
    84-101): two differences, the sum of the two signs, one product, two
    face values.  MC (:141-158) adds the central difference and its 0.25."""
    assert unit(fc, 0) == (6, 8, 3, 0)
    assert unit(fc, 1, 0) == (0, 0, 0, 0)
    assert unit(fc, 1, 1) == (5, 1, 0, 0)
    assert unit(fc, 1, 2) == (6, 2, 0, 0)


def count_step(fc, U, lm, ch):
    """Counts per zone-update for a state U(ix, jx, 9) with ghosts: (all
    operations, those in HLLD), each as add, mul, div, sqrt; HLLD branch
    shares."""
    I = U[1:-1, 1:-1, :]
    nx, ny = I.shape[:2]
    soa = np.ascontiguousarray(I.transpose(2, 1, 0))
    out, br = (C.c_long * 8)(), (C.c_long * 9)()
    fc.flops_2d_step(soa.ctypes.data_as(C.POINTER(C.c_double)), nx, ny, lm, ch, out, br)
    z = nx * ny
    return np.array(out[:4]) / z, np.array(out[4:]) / z, np.array(br[1:]) / (2 * 2 * z)


@pytest.mark.parametrize("case", ("kh", "ot"))
def test_flops_per_zone_update(fc, case):
    """KH (minmod) and OT (MC) initial states: add, mul, div, sqrt per
    zone-update, and the share of HLLD interfaces per branch.  Everything but
    HLLD follows from the formulas and is checked exactly: per stage u2v
    (6 add, 8 mul, 3 div), 2 directions x 9 limiter calls (minmod 5+1, MC 6+2),
    the RK combine 9 x (3+1+1) in stage 1 (rk.f90:30-31) and 9 x (5 add, 2 mul)
    in stage 2 (:60-62); per step GLM's two psi scalings and the CFL speed of
    set_dt.f90:28-44 as written: 14 add, 14 mul, 2 div, 4 sqrt."""
    from openmhd_amd import problems

    if case == "kh":
        U, _ = problems.kelvin_helmholtz(256, 256, seed=1)
        lm, ch, lim = 1, 2.0, np.array([5, 1, 0, 0])
    else:
        U, _ = problems.orszag_tang(128)
        lm, ch, lim = 2, 2.5, np.array([6, 2, 0, 0])
    per, hlld, share = count_step(fc, U, lm, ch)
    u2v = np.array([6, 8, 3, 0])
    rk = 9 * np.array([4, 1, 0, 0]) + 9 * np.array([5, 2, 0, 0])
    fixed = 2 * u2v + 2 * 2 * 9 * lim + rk + np.array([0, 2, 0, 0]) + np.array([14, 14, 2, 4])
    np.testing.assert_allclose(per - hlld, fixed, rtol=0, atol=1e-9)
    total = per.sum()
    print(f"{case}: per zone-update add {per[0]:.1f} mul {per[1]:.1f} div {per[2]:.2f} sqrt {per[3]:.2f} "
          f"= {total:.1f} flops, of which HLLD {hlld.sum():.1f} ({hlld.sum() / 4:.1f} per interface: "
          f"div {hlld[2] / 4:.2f}, sqrt {hlld[3] / 4:.2f}); HLLD branch share {np.array2string(share, precision=3)}")
    assert share.sum() == pytest.approx(1.0)
