"""The 2D reconnection drop-in (omhd_flux_resistive, the reconnection
boundary tables, the Python mirror of 2D_reconnection's routines) on the GPU
against the oracle and the flang vectors of the reference's own routines:
bitwise.  The driver-level runs are in tests/test_gpu_0_fortran.py."""
import ctypes as C

import numpy as np
import pytest

from omhd_testlib import DP, farr, ptr
from test_oracle_golden import bits_equal, load

pytestmark = pytest.mark.gpu

# (ix, jx): ordinary, one-face-wide, the smallest with any face, no faces
SIZES = [(12, 10), (70, 50), (130, 3), (3, 41), (3, 3), (2, 9), (9, 2), (2, 2), (1, 5)]


@pytest.fixture(scope="module")
def lib():
    from openmhd_amd._lib import lib

    return lib


@pytest.fixture(scope="module")
def fr2d(oracle):
    f = oracle.L.or_flux_resistive2d
    f.argtypes = [DP, DP, DP, C.c_int, C.c_int, C.c_int]
    f.restype = None
    return f


def _case(rng, ix, jx):
    U = np.asfortranarray(rng.normal(size=(ix, jx, 9)))
    F = np.asfortranarray(rng.normal(size=(ix, jx, 9)))
    e = np.asfortranarray(rng.uniform(0.01, 2.0, size=(ix, jx)))
    return U, F, e


@pytest.mark.parametrize("dir", (1, 2, 3))
@pytest.mark.parametrize("ix,jx", SIZES)
def test_flux_resistive_matches_oracle(lib, fr2d, ix, jx, dir):
    """Host arrays, arbitrary face tables: F after the call equals the oracle's
    2D CPU form bitwise, including the untouched faces and slots."""
    print(f"flux_resistive host ix={ix} jx={jx} dir={dir}", flush=True)
    rng = np.random.default_rng(1000 * ix + 10 * jx + dir)
    U, F, e = _case(rng, ix, jx)
    Fo = F.copy(order="F")
    assert lib.omhd_flux_resistive(ptr(F), ptr(U), ptr(e), ix, jx, dir) == 0
    fr2d(ptr(Fo), ptr(U), ptr(e), ix, jx, dir)
    assert bits_equal(F, Fo)


@pytest.mark.parametrize("dir", (1, 2))
def test_flux_resistive_device_tensors(lib, fr2d, dir):
    """Device memory is used in place (torch tensors, Fortran order)."""
    import torch

    print(f"flux_resistive device dir={dir}", flush=True)
    rng = np.random.default_rng(31 + dir)
    ix, jx = 257, 66
    U, F, e = _case(rng, ix, jx)
    Fo = F.copy(order="F")
    fr2d(ptr(Fo), ptr(U), ptr(e), ix, jx, dir)
    dU, dF, de = (torch.from_numpy(a).to("cuda") for a in (U, F, e))
    assert dF.stride()[0] == 1
    rc = lib.omhd_flux_resistive(C.cast(C.c_void_p(dF.data_ptr()), DP), C.cast(C.c_void_p(dU.data_ptr()), DP),
                                 C.cast(C.c_void_p(de.data_ptr()), DP), ix, jx, dir)
    torch.cuda.synchronize()
    assert rc == 0
    assert bits_equal(dF.cpu().numpy(), Fo)


def test_flux_resistive_argument_errors(lib):
    """A bad direction, a null array or a non-positive extent is an error code
    with nothing launched; dir 3 is a no-op, as in the reference."""
    from openmhd_amd._lib import OMHD_ERR_ARG, OMHD_ERR_DIR, OMHD_ERR_EXTENT

    U, F, e = _case(np.random.default_rng(3), 6, 5)
    F0 = F.copy(order="F")
    assert lib.omhd_flux_resistive(ptr(F), ptr(U), ptr(e), 6, 5, 0) == OMHD_ERR_DIR
    assert lib.omhd_flux_resistive(ptr(F), ptr(U), ptr(e), 6, 5, 4) == OMHD_ERR_DIR
    assert lib.omhd_flux_resistive(ptr(F), None, ptr(e), 6, 5, 1) == OMHD_ERR_ARG
    assert lib.omhd_flux_resistive(ptr(F), ptr(U), None, 6, 5, 1) == OMHD_ERR_ARG
    assert lib.omhd_flux_resistive(ptr(F), ptr(U), ptr(e), 0, 5, 1) == OMHD_ERR_EXTENT
    assert lib.omhd_flux_resistive(ptr(F), ptr(U), ptr(e), 6, 5, 3) == 0
    assert bits_equal(F, F0)


@pytest.mark.parametrize("dir", (1, 2))
def test_api_flux_resistive_matches_reference_vectors(dir):
    """The Python mirror flux_resistive(F,U,x,y,dx,ix,jx,dir) with the
    reference's eta(x, y) against flux_resistive.f90 run by flang: bitwise."""
    from openmhd_amd import api

    d = load("recon")
    U = np.asfortranarray(d["res_U"])
    ix, jx = U.shape[:2]
    F = np.asfortranarray(d[f"F0_{dir}"]).copy(order="F")
    api.flux_resistive(F, U, d["res_x"], d["res_y"], float(d["res_dx"]), ix, jx, dir)
    assert bits_equal(F, d[f"F_{dir}"])


def test_api_recon_boundaries_match_reference_vectors():
    """bc_for_U/F/G of 2D_reconnection/bc.f90 through the mirror: bitwise."""
    from openmhd_amd import api

    d = load("recon")
    U = np.asfortranarray(d["bcU_in"]).copy(order="F")
    api.bc_for_U_recon(U, *U.shape[:2])
    assert bits_equal(U, d["bcU_out"])
    for fn, kl, kr in ((api.bc_for_F_recon, "bcF_L", "bcF_R"), (api.bc_for_G_recon, "bcG_L", "bcG_R")):
        VL = np.asfortranarray(d["V_L"]).copy(order="F")
        VR = np.asfortranarray(d["V_R"]).copy(order="F")
        fn(VL, VR, *VL.shape[:2])
        assert bits_equal(VL, d[kl]) and bits_equal(VR, d[kr])


def test_api_reconnection_step_sequence_matches_oracle(oracle):
    """Two steps of 2D_reconnection/main.f90 written out with the Python
    mirror's calls (u2v, set_dt, set_dt_resistive, glm_ss2, limiter,
    bc_for_F/G, flux_solver, flux_resistive, rk_*, bc_for_U), one host round
    trip per call, against the oracle's step: bitwise."""
    from omhd_testlib import recon_problem

    from openmhd_amd import api, problems

    nx, ny, steps = 40, 16, 2
    U, _, x, y = problems.harris_sheet_2d(nx, ny, lx=0.4)
    p, tab = recon_problem(x, y, lm_type=2, time_type=0)
    oracle.bc_u(U, p)
    Uo = U.copy(order="F")
    rc, dto, cho = oracle.run(p, Uo, steps)
    assert rc == 0
    ix, jx, dx, cfl = nx + 2, ny + 2, float(x[1] - x[0]), 0.35
    V, U1 = farr(ix, jx, 4), U.copy(order="F")
    VL, VR, F, G = (farr(ix, jx, 9) for _ in range(4))

    def sweep(Us, dir, FX):  # planes of Fortran-ordered arrays are contiguous, as V(:,:,vx) is
        for m in range(9):
            api.limiter(V[:, :, m] if m < 4 else Us[:, :, m], VL[:, :, m], VR[:, :, m], ix, jx, dir, 2)
        (api.bc_for_F_recon if dir == 1 else api.bc_for_G_recon)(VL, VR, ix, jx)
        api.flux_solver(FX, VL, VR, ch, ix, jx, dir, 3)
        api.flux_resistive(FX, Us, x, y, dx, ix, jx, dir)

    for s in range(steps):
        api.u2v(U, V, ix, jx)
        ch, dt = api.set_dt(U, V, dx, cfl, ix, jx)
        dt = api.set_dt_resistive(dt, dx, cfl)
        assert bits_equal(dt, dto[s]) and bits_equal(ch, cho[s])
        api.glm_ss2(U, ch, dt, ix, jx)
        sweep(U, 1, F)
        sweep(U, 2, G)
        api.rk_tvd21(U, U1, F, G, dt, dx, ix, jx)
        api.bc_for_U_recon(U1, ix, jx)
        api.u2v(U1, V, ix, jx)
        sweep(U1, 1, F)
        sweep(U1, 2, G)
        api.rk_tvd22(U, U1, F, G, dt, dx, ix, jx)
        api.bc_for_U_recon(U, ix, jx)
        api.glm_ss2(U, ch, dt, ix, jx)
    assert bits_equal(U, Uo)
