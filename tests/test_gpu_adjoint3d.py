"""The GPU adjoint with wall boundaries, in 3D, and with the resistive flux
(the 3D resistive reconnection physics on small grids), against the CPU reverse of the same
discrete step (tests/native/adjcheck.cpp, general form: its forward is the
oracle's bits, its reverse the exact transpose of dual-number tangents), and
the resistivity-parameter gradient against the exact tangent along d eta /
d(eta0, eta1, w) and against central differences of the GPU forward."""
import ctypes as C

import numpy as np
import pytest

from omhd_testlib import DP, AdjCheck, KH_SIGN, cptr, make_problem, smooth_state

pytestmark = pytest.mark.gpu
D = C.c_double


@pytest.fixture(scope="module")
def ac():
    return AdjCheck()


def soa3(U, nd):
    if nd == 2:
        return np.ascontiguousarray(U[1:-1, 1:-1, :].transpose(2, 1, 0))
    return np.ascontiguousarray(U[1:-1, 1:-1, 1:-1, :].transpose(3, 2, 1, 0))


def setup(oracle, nx, ny, nz, bc, lm, tt, res, seed, dx=0.1, signs=None):
    """Solver + oracle problem + CPU harness arguments for one case."""
    from openmhd_amd import Solver, snapshot

    rng = np.random.default_rng(seed)
    nd = 3 if nz else 2
    U = smooth_state(rng, nx + 2, ny + 2, (nz + 2) if nz else 1)
    p = make_problem(nx + 2, ny + 2, (nz + 2) if nz else 1, dx=dx, cfl=0.3, lm_type=lm, time_type=tt, bc=bc,
                     signs=signs)
    origin = (-0.5 * nx * dx, -0.5 * ny * dx)
    resistivity = (1e-3, 2e-2, 0.5) if res else None
    eta = None
    if res:
        x = snapshot.grid(nx, dx, origin[0])
        y = snapshot.grid(ny, dx, origin[1])
        eta, emax = oracle.eta_faces(x, y, dx, *resistivity)
        eta = np.ascontiguousarray(eta)
        p.eta = eta.ctypes.data_as(DP)
        p.eta_max = emax
    oracle.bc_u(U, p)
    s = Solver(nx, ny, dx, nz=nz, cfl=0.3, lm_type=lm, time_type=tt, bc=bc, resistivity=resistivity, origin=origin,
               **({"wall_sign": signs} if signs is not None else {}))
    s.set_state(soa3(U, nd))
    bca = (C.c_int * 3)(*[1 if b == "wall" else 0 for b in bc])
    sg = np.ascontiguousarray(np.array(p.u_sgn[:], dtype=np.float64))
    return s, U, p, soa3(U, nd), eta, bca, sg, resistivity, origin


from omhd_testlib import RECON_SIGN  # noqa: E402

CASES = [
    (12, 10, 0, ("wall", "wall", "periodic"), 2, 0, 0, 3, 2, RECON_SIGN),
    (8, 6, 5, ("wall", "wall", "periodic"), 1, 0, 1, 3, 2, RECON_SIGN),
    (12, 10, 0, ("periodic", "wall", "periodic"), 1, 0, 0, 4, 2, None),
    (12, 10, 0, ("wall", "wall", "periodic"), 2, 1, 0, 3, 3, None),
    (8, 6, 5, ("periodic",) * 3, 2, 0, 0, 3, 2, None),
    (8, 6, 5, ("periodic", "wall", "periodic"), 1, 0, 1, 3, 2, None),
    (8, 6, 6, ("wall", "wall", "wall"), 2, 1, 1, 2, 1, None),
    # large enough for the gather's visiting order to deal whole block ranges
    # to the XCDs (a block count divisible by 8) and, in 3D, 16 x 8 row tiles
    (64, 32, 0, ("periodic", "wall", "periodic"), 2, 0, 0, 2, 1, None),
    (32, 16, 16, ("periodic", "wall", "periodic"), 1, 0, 1, 2, 1, None),
]


@pytest.mark.parametrize("nx,ny,nz,bc,lm,tt,res,nsteps,every,signs", CASES)
def test_adjoint_matches_cpu_reverse(ac, oracle, nx, ny, nz, bc, lm, tt, res, nsteps, every, signs):
    from openmhd_amd import Adjoint

    s, U, p, W0, eta, bca, sg, resistivity, origin = setup(oracle, nx, ny, nz, bc, lm, tt, res, nx + 3 * res,
                                                           signs=signs)
    adj = Adjoint(s, nsteps, every)
    dts, chs = adj.forward()
    Uo = U.copy(order="F")
    rc, dto, cho = oracle.run(p, Uo, nsteps)
    assert rc == 0 and np.array_equal(dts, dto) and np.array_equal(chs, cho)
    assert np.array_equal(s.get_state().view(np.int64), soa3(Uo, 3 if nz else 2).view(np.int64))
    lamT = np.random.default_rng(7).standard_normal(W0.shape)
    lam0 = adj.backward(lamT)
    ref = np.zeros_like(W0)
    geta = np.zeros(eta.size) if res else None
    ac.L.adjc3_run_vjp(nx, ny, nz, lm, tt, D(0.1), bca, cptr(sg), cptr(eta.reshape(-1)) if res else None, nsteps,
                       cptr(dts), cptr(chs), cptr(W0), cptr(lamT), cptr(ref), cptr(geta) if res else None)
    err = np.abs(lam0 - ref).max() / np.abs(ref).max()
    assert err <= 1e-12, err
    if res:
        # dL/d(eta0, eta1, w) = geta . d eta / d(params), eta = eta0 + eta1 exp(-r^2/w^2)
        from openmhd_amd import snapshot

        e0, e1, w = resistivity
        x = snapshot.grid(nx, 0.1, origin[0])
        y = snapshot.grid(ny, 0.1, origin[1])
        ge = geta.reshape(3, ny + 2, nx + 2)
        ref_g = np.zeros(3)
        for d in range(3):
            X, Y = np.meshgrid(x + (0.05 if d == 0 else 0.0), y + (0.05 if d == 1 else 0.0), indexing="xy")
            r2 = X * X + Y * Y
            ex = np.exp(-r2 / (w * w))
            ref_g += [ge[d].sum(), (ge[d] * ex).sum(), (ge[d] * e1 * ex * 2 * r2 / w ** 3).sum()]
        g = adj.eta_grad()
        assert np.allclose(g, ref_g, rtol=1e-11, atol=1e-14 * np.abs(ref_g).max()), (g, ref_g)
        # the resistive reverse gathers (no atomics): a second pass repeats the first bit for bit
        lam0b = adj.backward(lamT)
        assert np.array_equal(lam0b.view(np.int64), lam0.view(np.int64))
        assert np.array_equal(adj.eta_grad().view(np.int64), g.view(np.int64))
    adj.close()
    s.close()


def test_eta_parameter_gradient(ac, oracle):
    """3D resistive, y walls: dL/d(eta1, w) from the adjoint equals the exact
    tangent (dual numbers along d eta/d eta1, d eta/d w) to 1e-10, and central
    differences of the GPU forward with the frozen dt/ch approach it."""
    from openmhd_amd import Adjoint, snapshot

    nx, ny, nz, nsteps = 10, 8, 6, 4
    bc = ("periodic", "wall", "periodic")
    s, U, p, W0, eta, bca, sg, (e0, e1, w), origin = setup(oracle, nx, ny, nz, bc, 2, 0, 1, 11)
    adj = Adjoint(s, nsteps, 2)
    dts, chs = adj.forward()
    lamT = np.random.default_rng(3).standard_normal(W0.shape)
    adj.backward(lamT)
    g = adj.eta_grad()
    x = snapshot.grid(nx, 0.1, origin[0])
    y = snapshot.grid(ny, 0.1, origin[1])
    de1, dw = np.zeros((3, ny + 2, nx + 2)), np.zeros((3, ny + 2, nx + 2))
    for d in range(3):
        X, Y = np.meshgrid(x + (0.05 if d == 0 else 0.0), y + (0.05 if d == 1 else 0.0), indexing="xy")
        r2 = X * X + Y * Y
        de1[d] = np.exp(-r2 / (w * w))
        dw[d] = e1 * np.exp(-r2 / (w * w)) * 2 * r2 / w ** 3
    for k, de in ((1, de1), (2, dw)):
        jv = np.zeros_like(W0)
        ac.L.adjc3_run_jvp(nx, ny, nz, 2, 0, D(0.1), bca, cptr(sg), cptr(eta.reshape(-1)),
                           cptr(np.ascontiguousarray(de.reshape(-1))), nsteps, cptr(dts), cptr(chs), cptr(W0),
                           cptr(np.zeros_like(W0)), cptr(jv))
        tl = float((lamT * jv).sum())
        assert abs(g[k] - tl) <= 1e-10 * abs(tl), (k, g[k], tl)
    # central differences in eta1 and w with the frozen sequence
    best = []
    for k in (1, 2):
        rel = []
        for eps in (1e-5, 1e-6, 1e-7):
            L = []
            for sgn in (1, -1):
                prm = [e0, e1, w]
                prm[k] += sgn * eps * prm[k]
                s.set_resistivity(*prm)
                s.set_state(W0)
                s.step_fixed(dts, chs)
                L.append(float((lamT * s.get_state()).sum()))
            fd = (L[0] - L[1]) / (2 * eps * [e0, e1, w][k])
            rel.append(abs(fd - g[k]) / abs(g[k]))
        best.append(min(rel))
    assert max(best) <= 1e-6, best
    adj.close()
    s.close()


def test_rollout_takes_eta_as_a_tensor(ac, oracle):
    """Rollout(W0, eta=tensor): the resistivity parameters are a torch input.
    At parameters other than the solver's own, torch.autograd gives eta.grad
    equal to Adjoint.eta_grad() (and W0.grad to Adjoint.backward) of the same
    run bit for bit (the resistive reverse sums in a fixed order), and equal
    to the exact tangent along d eta/d eta1 and d eta/d w to 1e-10."""
    import torch

    from openmhd_amd import Adjoint, Rollout, snapshot

    nx, ny, nz, nsteps = 10, 8, 6, 4
    bc = ("periodic", "wall", "periodic")
    s, U, p, W0, _, bca, sg, (e0, e1, w), origin = setup(oracle, nx, ny, nz, bc, 2, 0, 1, 11)
    prm = (1.5 * e0, 0.75 * e1, 0.625)  # not the solver's creation values
    lamT = np.random.default_rng(5).standard_normal(W0.shape)
    # the library's own calls
    s.set_resistivity(*prm)
    s.set_state(W0)
    adj = Adjoint(s, nsteps, 2)
    dts, chs = adj.forward()
    gW = adj.backward(lamT)
    g = adj.eta_grad()
    adj.close()
    # the same run through torch.autograd, starting from the creation values
    s.set_resistivity(e0, e1, w)
    W0t = torch.tensor(W0, device="cuda", requires_grad=True)
    etat = torch.tensor(prm, dtype=torch.float64, requires_grad=True)
    roll = Rollout(s, nsteps, 2)
    WT = roll(W0t, eta=etat)
    assert np.array_equal(roll.dt.view(np.int64), dts.view(np.int64))
    (WT * torch.tensor(lamT, device="cuda")).sum().backward()
    assert np.array_equal(etat.grad.numpy().view(np.int64), g.view(np.int64)), (etat.grad, g)
    gWt = W0t.grad.cpu().numpy()
    assert np.array_equal(gWt.view(np.int64), gW.view(np.int64))
    # exact tangents at prm
    x = snapshot.grid(nx, 0.1, origin[0])
    y = snapshot.grid(ny, 0.1, origin[1])
    eta, _ = oracle.eta_faces(x, y, 0.1, *prm)
    eta = np.ascontiguousarray(eta)
    pe0, pe1, pw = prm
    de1, dw = np.zeros((3, ny + 2, nx + 2)), np.zeros((3, ny + 2, nx + 2))
    for d in range(3):
        X, Y = np.meshgrid(x + (0.05 if d == 0 else 0.0), y + (0.05 if d == 1 else 0.0), indexing="xy")
        r2 = X * X + Y * Y
        de1[d] = np.exp(-r2 / (pw * pw))
        dw[d] = pe1 * np.exp(-r2 / (pw * pw)) * 2 * r2 / pw ** 3
    for k, de in ((1, de1), (2, dw)):
        jv = np.zeros_like(W0)
        ac.L.adjc3_run_jvp(nx, ny, nz, 2, 0, D(0.1), bca, cptr(sg), cptr(eta.reshape(-1)),
                           cptr(np.ascontiguousarray(de.reshape(-1))), nsteps, cptr(dts), cptr(chs), cptr(W0),
                           cptr(np.zeros_like(W0)), cptr(jv))
        tl = float((lamT * jv).sum())
        assert abs(etat.grad[k].item() - tl) <= 1e-10 * abs(tl), (k, etat.grad[k].item(), tl)
    s.close()
