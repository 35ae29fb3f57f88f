"""The reconnection drop-ins on the GPU: the 2D form (omhd_flux_resistive,
2D_reconnection's boundary tables, the Python mirror of its routines) against
the oracle and the flang vectors of the reference's own routines, and the 3D
form of 3D_reconnection_gpu (omhd_flux_resistive3d, bc.cuf's tables) against
the oracle's restatement: bitwise.  The driver-level runs are in
tests/test_gpu_0_fortran.py."""
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


# ---------------------------------------------------------------------------
# the 3D form (3D_reconnection_gpu/flux_resistive.cuf) as a host-callable
# drop-in, omhd_flux_resistive3d, against the oracle's restatement
# ---------------------------------------------------------------------------
SIZES3 = [(12, 10, 8), (40, 23, 17), (3, 3, 3), (2, 7, 6), (7, 2, 6), (7, 6, 2), (3, 9, 3), (1, 4, 4)]


@pytest.fixture(scope="module")
def fr3d(oracle):
    f = oracle.L.or_flux_resistive
    f.argtypes = [DP, DP, DP, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int]
    f.restype = None
    return f


@pytest.mark.parametrize("dir", (1, 2, 3))
@pytest.mark.parametrize("ix,jx,kx", SIZES3)
def test_flux_resistive3d_matches_oracle(lib, fr3d, ix, jx, kx, dir):
    """Host arrays, a z-invariant face table of eta: bitwise to the oracle's
    3D form, untouched faces and slots included."""
    print(f"flux_resistive3d host {ix}x{jx}x{kx} dir={dir}", flush=True)
    rng = np.random.default_rng(7 * ix + 11 * jx + 13 * kx + dir)
    U = np.asfortranarray(rng.normal(size=(ix, jx, kx, 9)))
    F = np.asfortranarray(rng.normal(size=(ix, jx, kx, 9)))
    e = np.asfortranarray(rng.uniform(1e-3, 2e-2, size=(ix, jx)))
    Fo = F.copy(order="F")
    dx = 0.037
    assert lib.omhd_flux_resistive3d(ptr(F), ptr(U), ptr(e), 1, dx, ix, jx, kx, dir) == 0
    fr3d(ptr(Fo), ptr(U), ptr(e), dx, ix, jx, kx, dir)
    assert bits_equal(F, Fo)


@pytest.mark.parametrize("dir", (1, 2, 3))
def test_flux_resistive3d_full_eta_table(lib, dir):
    """eta_nz = kx reads one table per z plane: with F = 0 and plane k's
    table 2^(k mod 3) times a base table, every increment scales exactly."""
    ix, jx, kx, dx = 14, 11, 9, 0.05
    rng = np.random.default_rng(40 + dir)
    U = np.asfortranarray(rng.normal(size=(ix, jx, kx, 9)))
    e = np.asfortranarray(rng.uniform(1e-3, 2e-2, size=(ix, jx)))
    scale = 2.0 ** (np.arange(kx) % 3)
    full = np.asfortranarray(e[:, :, None] * scale[None, None, :])
    F1, Fk = farr(ix, jx, kx, 9), farr(ix, jx, kx, 9)
    assert lib.omhd_flux_resistive3d(ptr(F1), ptr(U), ptr(e), 1, dx, ix, jx, kx, dir) == 0
    assert lib.omhd_flux_resistive3d(ptr(Fk), ptr(U), ptr(full), kx, dx, ix, jx, kx, dir) == 0
    assert np.abs(F1).max() > 0
    assert bits_equal(Fk, F1 * scale[None, None, :, None])


def test_flux_resistive3d_argument_errors(lib):
    from openmhd_amd._lib import OMHD_ERR_ARG, OMHD_ERR_DIR

    U, F = farr(6, 5, 4, 9), farr(6, 5, 4, 9)
    e = farr(6, 5)
    assert lib.omhd_flux_resistive3d(ptr(F), ptr(U), ptr(e), 1, 0.1, 6, 5, 4, 0) == OMHD_ERR_DIR
    assert lib.omhd_flux_resistive3d(ptr(F), ptr(U), ptr(e), 2, 0.1, 6, 5, 4, 1) == OMHD_ERR_ARG
    assert lib.omhd_flux_resistive3d(ptr(F), None, ptr(e), 1, 0.1, 6, 5, 4, 1) == OMHD_ERR_ARG


def test_api_flux_resistive_3d_and_boundaries_match_oracle(oracle, fr3d):
    """The Python mirror's 3D reconnection routines: flux_resistive_3d with
    flux_resistive.cuf's eta(x, y) against the oracle's 3D form, and the
    bc.cuf tables (bc_for_U/F/G/H) against the oracle's, bitwise."""
    from omhd_testlib import RECON_SIGN, make_problem

    from openmhd_amd import api

    ix, jx, kx, dx = 16, 12, 7, 0.3
    x = (np.arange(ix) - ix / 2 + 0.5) * dx
    y = (np.arange(jx) - 0.5) * dx
    z = np.arange(kx) * dx
    rng = np.random.default_rng(5)
    U = np.asfortranarray(rng.normal(size=(ix, jx, kx, 9)))
    for dir in (1, 2, 3):
        F = np.asfortranarray(rng.normal(size=(ix, jx, kx, 9)))
        Fo = F.copy(order="F")
        api.flux_resistive_3d(F, U, x, y, z, dx, ix, jx, kx, dir)
        fr3d(ptr(Fo), ptr(U), ptr(api.eta_face_table(x, y, dx, dir, over_dx=False)), dx, ix, jx, kx, dir)
        assert np.abs(F - Fo).max() == 0 and bits_equal(F, Fo)
    p = make_problem(ix, jx, kx, bc=("wall", "wall", "periodic"), signs=RECON_SIGN)
    Ug, Uo = U.copy(order="F"), U.copy(order="F")
    api.bc_for_U_recon(Ug, ix, jx, kx)
    oracle.bc_u(Uo, p)
    assert bits_equal(Ug, Uo)
    for d, fn in ((1, api.bc_for_F_recon), (2, api.bc_for_G_recon), (3, api.bc_for_H_recon)):
        VL = np.asfortranarray(rng.normal(size=(ix, jx, kx, 9)))
        VR = np.asfortranarray(rng.normal(size=(ix, jx, kx, 9)))
        VLo, VRo = VL.copy(order="F"), VR.copy(order="F")
        fn(VL, VR, ix, jx, kx)
        lo = np.ascontiguousarray(RECON_SIGN[2 * d - 2])
        hi = np.ascontiguousarray(RECON_SIGN[2 * d - 1])
        oracle.L.or_bc_face(ptr(VLo), ptr(VRo), ix, jx, kx, d, 0 if d == 3 else 1, ptr(lo), ptr(hi))
        assert bits_equal(VL, VLo) and bits_equal(VR, VRo), d


def _dev_ptr(t):
    return C.cast(C.c_void_p(t.data_ptr()), DP)


@pytest.mark.parametrize("shape", [(4098, 4098, 1), (514, 258, 258)])
def test_flux_resistive_full_size_device(lib, fr2d, fr3d, shape):
    """At the benchmark grids (KH 4096^2 for the 2D form, the 3D resistive
    Harris grid 512x256x256 for the 3D form, whose state planes exceed 2^31
    bytes), on device tensors in place: bitwise to the oracle, every dir."""
    import torch

    ix, jx, kx = shape
    rng = np.random.default_rng(ix + kx)
    full = (ix, jx, 9) if kx == 1 else (ix, jx, kx, 9)
    U = np.asfortranarray(rng.standard_normal(full))
    e = np.asfortranarray(rng.uniform(1e-3, 2e-2, size=(ix, jx)))
    dU, de = torch.from_numpy(U).to("cuda"), torch.from_numpy(e).to("cuda")
    for dir in ((1, 2) if kx == 1 else (1, 2, 3)):
        print(f"flux_resistive full size {shape} dir={dir}", flush=True)
        F = np.asfortranarray(rng.standard_normal(full))
        dF = torch.from_numpy(F).to("cuda")
        if kx == 1:
            rc = lib.omhd_flux_resistive(_dev_ptr(dF), _dev_ptr(dU), _dev_ptr(de), ix, jx, dir)
            fr2d(ptr(F), ptr(U), ptr(e), ix, jx, dir)
        else:
            rc = lib.omhd_flux_resistive3d(_dev_ptr(dF), _dev_ptr(dU), _dev_ptr(de), 1, 0.0125, ix, jx, kx, dir)
            fr3d(ptr(F), ptr(U), ptr(e), 0.0125, ix, jx, kx, dir)
        torch.cuda.synchronize()
        assert rc == 0
        assert bits_equal(dF.cpu().numpy(), F)
        del dF
    del dU, de
    torch.cuda.empty_cache()
