"""Drop-in entry points (one HIP kernel per reference subroutine) against the
oracle and the golden vectors: bitwise."""
import numpy as np
import pytest

from omhd_testlib import farr, make_problem, prim_to_cons, random_prim, smooth_state
from test_oracle_golden import bits_equal, load

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from openmhd_amd import api

    return api


def test_u2v(api, oracle):
    rng = np.random.default_rng(5)
    U = np.asfortranarray(prim_to_cons(random_prim(rng, (37, 21))))
    V1, V2 = farr(37, 21, 4), farr(37, 21, 4)
    api.u2v(U, V1, 37, 21)
    oracle.u2v(U, V2)
    assert bits_equal(V1, V2)
    d = load("u2v_2d")
    V = farr(*d["V"].shape)
    api.u2v(np.asfortranarray(d["U"]), V, *d["V"].shape[:2])
    assert bits_equal(V, d["V"])


def test_u2v_reports_negative_pressure(api):
    from openmhd_amd import OMHDError

    rng = np.random.default_rng(6)
    W = random_prim(rng, (9, 8))
    W[4, 3, 3] = -0.5  # p < 0 at Fortran cell (5, 4)
    U = np.asfortranarray(prim_to_cons(W))
    with pytest.raises(OMHDError) as e:
        api.u2v(U, farr(9, 8, 4), 9, 8)
    assert e.value.code == 1 and e.value.cell[:2] == (5, 4)


@pytest.mark.parametrize("typ", range(5))
@pytest.mark.parametrize("dir", (1, 2, 3))
def test_limiter(api, typ, dir):
    d = load("limiter")
    wk = np.asfortranarray(d[f"wk_{typ}_{dir}"])
    wL = np.full(wk.shape, 7.0, order="F")
    wR = np.full(wk.shape, -7.0, order="F")
    shp = wk.shape + (1,) * (3 - wk.ndim)
    api.limiter(wk, wL, wR, shp[0], shp[1], dir, typ, kx=shp[2])
    assert bits_equal(wL, d[f"wL_{typ}_{dir}"]) and bits_equal(wR, d[f"wR_{typ}_{dir}"])


@pytest.mark.parametrize("dir", (1, 2, 3))
def test_hlld_every_branch(api, dir):
    d = load("hlld")
    VL, VR = np.asfortranarray(d[f"VL_{dir}"]), np.asfortranarray(d[f"VR_{dir}"])
    F = farr(*VL.shape)
    shp = VL.shape[:-1] + (1,) * (4 - VL.ndim)
    api.flux_solver(F, VL, VR, float(d["ch"]), shp[0], shp[1], dir, 3, kx=shp[2])
    assert bits_equal(F, d[f"F_{dir}"])


@pytest.mark.parametrize("dir", (1, 2))
def test_hlld_random(api, oracle, dir):
    rng = np.random.default_rng(10 + dir)
    for sc in (0.1, 1.0, 4.0):
        VL = np.asfortranarray(random_prim(rng, (64, 40), vscale=sc))
        VR = np.asfortranarray(random_prim(rng, (64, 40), vscale=sc))
        VL[:20, :, 5:8] *= 1e-3
        F1, F2 = farr(64, 40, 9), farr(64, 40, 9)
        api.flux_solver(F1, VL, VR, 2.1, 64, 40, dir, 3)
        oracle.flux_solver(F2, VL, VR, 2.1, dir)
        assert bits_equal(F1, F2)


def test_glm_rk_setdt_bc(api):
    d = load("small")
    U = np.asfortranarray(d["glm_in"]).copy(order="F")
    api.glm_ss2(U, 1.37, 0.0123, 10, 8)
    assert bits_equal(U, d["glm_out"])
    for dim in ("2d", "3d"):
        F = np.asfortranarray(d[f"rk{dim}_F"])
        Gf = np.asfortranarray(d[f"rk{dim}_G"])
        H = np.asfortranarray(d["rk3d_H"]) if dim == "3d" else None
        shp = F.shape[:-1] + (1,) * (4 - F.ndim)
        fns = [api.rk_tvd21, api.rk_tvd22, api.rk_std21]
        for mode in range(4):
            A = np.asfortranarray(d[f"rk{dim}_{mode}_in_wk"]).copy(order="F")
            B = np.asfortranarray(d[f"rk{dim}_{mode}_in_wk1"]).copy(order="F")
            if mode < 3:
                fns[mode](A, B, F, Gf, 0.0123, 0.0371, shp[0], shp[1], wH=H, kx=shp[2])
            else:
                api.rk_std22(A, F, Gf, 0.0123, 0.0371, shp[0], shp[1], wH=H, kx=shp[2])
            assert bits_equal(A, d[f"rk{dim}_{mode}_wk"]), (dim, mode)
            assert bits_equal(B, d[f"rk{dim}_{mode}_wk1"]), (dim, mode)
    vm, dt = api.set_dt(np.asfortranarray(d["setdt2d_U"]), np.asfortranarray(d["setdt2d_V"]), 0.0371, 0.4, 10, 8)
    assert bits_equal([vm, dt], d["setdt2d_out"])
    vm, dt = api.set_dt(np.asfortranarray(d["setdt3d_U"]), np.asfortranarray(d["setdt3d_V"]), 0.0371, 0.4, 6, 5,
                        kx=7, ndim=3)
    assert bits_equal([vm, dt], d["setdt3d_out"])
    A = np.asfortranarray(d["bc_in"]).copy(order="F")
    api.bc(A, 7, 6)
    assert bits_equal(A, d["bc_out"])
    A = np.asfortranarray(d["bc3d_in"]).copy(order="F")
    api.bc(A, 6, 5, kx=7)
    assert bits_equal(A, d["bc3d_out"])
    A = np.asfortranarray(d["bckh_in"]).copy(order="F")
    api.bc_for_U(A, 7, 6)
    assert bits_equal(A, d["bckh_out"])
    VL = np.asfortranarray(d["bcF_inL"]).copy(order="F")
    VR = np.asfortranarray(d["bcF_inR"]).copy(order="F")
    api.bc_for_F(VL, VR, 7, 6)
    assert bits_equal(VL, d["bcF_L"]) and bits_equal(VR, d["bcF_R"])
    api.bc_for_G(VL, VR, 7, 6)
    assert bits_equal(VL, d["bcG_L"]) and bits_equal(VR, d["bcG_R"])


def _dropin_run(api, U, nsteps, dx, cfl, lm, kh=False):
    """The reference step sequence (2D_basic/main.f90:56-188, 2D_KH/main.f90)
    driven call by call through the drop-in entry points."""
    ix, jx = U.shape[:2]
    V = farr(ix, jx, 4)
    U1 = U.copy(order="F")
    VL, VR, F, G = (farr(ix, jx, 9) for _ in range(4))
    dts = []

    def sweep(Ust, dir, out, ch):
        for m in range(4):
            api.limiter(V[:, :, m], VL[:, :, m], VR[:, :, m], ix, jx, dir, lm)
        for m in range(4, 9):
            api.limiter(Ust[:, :, m], VL[:, :, m], VR[:, :, m], ix, jx, dir, lm)
        if dir == 1:
            api.bc_for_F(VL, VR, ix, jx)
        elif kh:
            api.bc_for_G(VL, VR, ix, jx)
        else:
            api.periodic_patch(VL, VR, ix, jx, 2)
        api.flux_solver(out, VL, VR, ch, ix, jx, dir, 3)

    bcU = api.bc_for_U if kh else api.bc
    for _ in range(nsteps):
        api.u2v(U, V, ix, jx)
        ch, dt = api.set_dt(U, V, dx, cfl, ix, jx)
        api.glm_ss2(U, ch, dt, ix, jx)
        sweep(U, 1, F, ch)
        sweep(U, 2, G, ch)
        api.rk_tvd21(U, U1, F, G, dt, dx, ix, jx)
        bcU(U1, ix, jx)
        api.u2v(U1, V, ix, jx)
        sweep(U1, 1, F, ch)
        sweep(U1, 2, G, ch)
        api.rk_tvd22(U, U1, F, G, dt, dx, ix, jx)
        bcU(U, ix, jx)
        api.glm_ss2(U, ch, dt, ix, jx)
        dts.append(dt)
    return np.array(dts)


@pytest.mark.parametrize("kh", (False, True))
def test_dropin_whole_steps(api, oracle, kh):
    rng = np.random.default_rng(3)
    U = smooth_state(rng, 26, 22)
    p = make_problem(26, 22, dx=0.07, cfl=0.4, lm_type=2, bc=("periodic", "wall" if kh else "periodic", "periodic"))
    oracle.bc_u(U, p)
    U2 = U.copy(order="F")
    dts = _dropin_run(api, U, 5, 0.07, 0.4, 2, kh=kh)
    rc, dts2, _ = oracle.run(p, U2, 5)
    assert rc == 0 and bits_equal(dts, dts2) and bits_equal(U, U2)


@pytest.mark.parametrize("typ", (0, 1, 2, 4))
@pytest.mark.parametrize("dir", (1, 2, 3))
def test_other_flux_types_bitwise(api, typ, dir):
    """flux_solver types 0 (LLF), 1 (HLL), 2 (HLLC), 4 (LHLLD) on the GPU,
    bitwise to the reference's vectors."""
    d0, d = load("hlld"), load("solvers")
    VL, VR = np.asfortranarray(d0[f"VL_{dir}"]), np.asfortranarray(d0[f"VR_{dir}"])
    shp = VL.shape[:-1] + (1,) * (4 - VL.ndim)
    F = farr(*VL.shape)
    api.flux_solver(F, VL, VR, float(d0["ch"]), shp[0], shp[1], dir, typ, kx=shp[2])
    assert bits_equal(F, d[f"F_{typ}_{dir}"])


@pytest.mark.parametrize("dir", (1, 2, 3))
def test_device_tensors_in_place(api, dir):
    """Device memory is used in place: torch CUDA tensors in Fortran order
    give the same bits as the host-staged call and the flang fixture."""
    import torch

    d = load("hlld")
    VL, VR = np.asfortranarray(d[f"VL_{dir}"]), np.asfortranarray(d[f"VR_{dir}"])
    shp = VL.shape[:-1] + (1,) * (4 - VL.ndim)
    dev = [torch.from_numpy(a).to("cuda") for a in (VL, VR, farr(*VL.shape))]
    assert all(t.is_cuda and t.stride()[0] == 1 for t in dev)  # column-major on the device
    api.flux_solver(dev[2], dev[0], dev[1], float(d["ch"]), shp[0], shp[1], dir, 3, kx=shp[2])
    torch.cuda.synchronize()
    assert bits_equal(dev[2].cpu().numpy(), d[f"F_{dir}"])
