"""The GPU adjoint (omhd_adjoint_*, the hand-written reverse kernels) against
(i) the CPU reverse of the same discrete step (host build of the reverse
forms, itself checked exactly against dual-number tangents), (ii) central
finite differences of the GPU forward with the same frozen (dt, ch), and
(iii) the torch.autograd.Function over it.  The forward pass of the adjoint
run is the product step: its states and dt sequence are the oracle's bits."""
import numpy as np
import pytest

from omhd_testlib import AdjCheck, cptr, make_problem, orszag_tang, soa

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ac():
    return AdjCheck()


def ot(n):
    U, dx = orszag_tang(n)
    return U, soa(U), dx


@pytest.mark.parametrize("lm,tt,nsteps,every", [(2, 0, 4, 2), (1, 0, 5, 3), (2, 1, 4, 4), (1, 1, 3, 1), (0, 0, 3, 2),
                                                (3, 0, 3, 2), (4, 1, 3, 2)])
def test_adjoint_matches_cpu_reverse(ac, oracle, lm, tt, nsteps, every):
    import ctypes as C

    from openmhd_amd import Adjoint, Solver

    n = 24
    U, W0, dx = ot(n)
    s = Solver(n, n, dx, lm_type=lm, time_type=tt)
    s.set_state(W0)
    adj = Adjoint(s, nsteps, every)
    dts, chs = adj.forward()
    # forward pass = the reference step (oracle bits, same dt sequence)
    p = make_problem(n + 2, n + 2, dx=dx, cfl=0.4, lm_type=lm, time_type=tt)
    oracle.bc_u(U, p)
    rc, odt, och = oracle.run(p, U, nsteps)
    assert rc == 0 and np.array_equal(dts, odt) and np.array_equal(chs, och)
    assert np.array_equal(s.get_state().view(np.int64), soa(U).view(np.int64))
    lamT = np.random.default_rng(lm + 10 * tt).standard_normal(W0.shape)
    lam0 = adj.backward(lamT)
    ref = np.zeros_like(W0)
    ac.run_vjp(n, n, lm, tt, C.c_double(dx), nsteps, cptr(dts), cptr(chs), cptr(W0), cptr(lamT), cptr(ref))
    err = np.abs(lam0 - ref).max() / np.abs(ref).max()
    assert err <= 1e-12, err
    adj.close()
    s.close()


def fd_rel(s, W0, dts, chs, lamT, grad, v, eps_list):
    """Relative differences between <grad, v> and central differences of the
    GPU forward (frozen dt/ch) at each eps."""
    an = float((grad * v).sum())
    out = []
    for eps in eps_list:
        L = []
        for sgn in (1, -1):
            s.set_state(W0 + sgn * eps * v)
            s.step_fixed(dts, chs)
            L.append(float((lamT * s.get_state()).sum()))
        out.append(abs((L[0] - L[1]) / (2 * eps) - an) / abs(an))
    return out


def smooth_field(rng, n, nmodes=3):
    x = (np.arange(n) + 0.5) * 2 * np.pi / n
    X, Y = np.meshgrid(x, x, indexing="xy")
    v = np.zeros((9, n, n))
    for m in range(9):
        for _ in range(nmodes):
            kx, ky = rng.integers(-3, 4, 2)
            v[m] += rng.standard_normal() * np.sin(kx * X + ky * Y + rng.uniform(0, 2 * np.pi))
    return v


@pytest.mark.parametrize("lm", (1, 2))
def test_adjoint_against_tangent_and_differences(ac, lm):
    """From a generic state (OT plus a smooth perturbation of every field, so
    no field sits on a limiter kink): <dL/dW0, v> of the GPU adjoint equals
    <lamT, J v> of the exact dual-number tangent of the same step to 1e-12,
    and central differences of the GPU forward approach it as eps shrinks
    (the map is piecewise smooth: a larger eps crosses limiter kinks)."""
    import ctypes as C

    from openmhd_amd import Adjoint, Solver

    n, nsteps = 32, 10
    _, W, dx = ot(n)
    W0 = W + 1e-2 * smooth_field(np.random.default_rng(9), n)
    s = Solver(n, n, dx, lm_type=lm)
    s.set_state(W0)
    adj = Adjoint(s, nsteps, 4)
    dts, chs = adj.forward()
    rng = np.random.default_rng(5)
    lamT = smooth_field(rng, n)
    grad = adj.backward(lamT)
    best = []
    for k in range(4):
        v = smooth_field(rng, n)
        jv = np.zeros_like(W0)
        ac.run_jvp(n, n, lm, 0, C.c_double(dx), nsteps, cptr(dts), cptr(chs), cptr(W0), cptr(v), cptr(jv))
        an, tl = float((grad * v).sum()), float((lamT * jv).sum())
        assert abs(an - tl) <= 1e-12 * abs(tl), (k, an, tl)
        # central differences: kink crossings shrink with eps, the rounding
        # noise of the forward map grows as 1/eps (about 1e-6 at eps 1e-9)
        best.append(min(fd_rel(s, W0, dts, chs, lamT, grad, v, (1e-7, 1e-8, 3e-9))))
    assert max(best) <= 1e-5 and float(np.median(best)) <= 1e-6, best


def test_rollout_autograd_function():
    import torch

    from openmhd_amd import Adjoint, Rollout, Solver

    n = 32
    _, W0, dx = ot(n)
    s = Solver(n, n, dx, lm_type=2)
    W = torch.tensor(W0, device="cuda", requires_grad=True)
    roll = Rollout(s, 8, every=3)
    WT = roll(W)
    target = torch.zeros_like(WT[6])
    loss = ((WT[6] - target) ** 2).sum()
    loss.backward()
    # the same gradient through the C ABI directly
    s2 = Solver(n, n, dx, lm_type=2)
    s2.set_state(W0)
    adj = Adjoint(s2, 8, 3)
    dts, _ = adj.forward()
    assert np.array_equal(dts, roll.dt)
    lamT = np.zeros_like(W0)
    lamT[6] = 2 * WT[6].detach().cpu().numpy()
    g = adj.backward(lamT)
    assert np.array_equal(W.grad.cpu().numpy(), g)
    # repeated calls with an imposed sequence
    W2 = torch.tensor(W0, device="cuda", requires_grad=True)
    WT2 = roll(W2, dt=roll.dt, ch=roll.ch)
    assert torch.equal(WT2.detach(), WT.detach())


def oracle_L(oracle, W, dts, chs, lamT, dx, lm):
    """<lamT, W_T> after the steps of the oracle (the CPU build's arithmetic,
    bitwise the reference) from the interior state W, periodic, with the
    given frozen (dt, ch) per step."""
    import ctypes as C

    from omhd_testlib import ptr

    n = W.shape[1]
    U = np.zeros((n + 2, n + 2, 9), order="F")
    U[1:-1, 1:-1, :] = W.transpose(2, 1, 0)
    p = make_problem(n + 2, n + 2, dx=dx, cfl=0.4, lm_type=lm)
    oracle.bc_u(U, p)
    dt, ch, bad = C.c_double(), C.c_double(), C.c_long(-1)
    for a, b in zip(dts, chs):
        p.dt_fixed, p.ch_fixed = float(a), float(b)
        assert oracle.L.or_step(C.byref(p), ptr(U), C.byref(dt), C.byref(ch), C.byref(bad)) == 0
    return float((lamT * soa(U)).sum())


@pytest.mark.parametrize("lm", (1, 2))
def test_adjoint_against_differences_of_the_cpu_build(oracle, lm):
    """The reference gradient by finite differences of the CPU build: central
    differences of the oracle's step (bitwise the reference's arithmetic) with
    the adjoint run's frozen (dt, ch), on the case of the test above.  They
    equal the differences of the GPU forward bit for bit (the two forward
    maps are the same function of the state), and the GPU adjoint's
    <dL/dW0, v> matches them.  How close differences can come is set by the
    forward's rounding noise over 2 eps |<dL/dW0, v>|: with a smaller
    directional derivative than here (other seeds, smaller grids) the best
    eps leaves 1e-4..1e-3, while the exact tangent check above stays at
    1e-12."""
    from openmhd_amd import Adjoint, Solver

    n, nsteps = 32, 10
    _, W, dx = ot(n)
    W0 = W + 1e-2 * smooth_field(np.random.default_rng(9), n)
    s = Solver(n, n, dx, lm_type=lm)
    s.set_state(W0)
    adj = Adjoint(s, nsteps, 4)
    dts, chs = adj.forward()
    rng = np.random.default_rng(5)
    lamT = smooth_field(rng, n)
    grad = adj.backward(lamT)
    best = []
    for k in range(4):
        v = smooth_field(rng, n)
        an = float((grad * v).sum())
        rel = []
        for eps in (1e-7, 1e-8, 3e-9):
            Lc = [oracle_L(oracle, W0 + sg * eps * v, dts, chs, lamT, dx, lm) for sg in (1, -1)]
            Lg = []
            for sg in (1, -1):
                s.set_state(W0 + sg * eps * v)
                s.step_fixed(dts, chs)
                Lg.append(float((lamT * s.get_state()).sum()))
            assert Lc == Lg, (k, eps)  # the CPU build's differences are the GPU forward's
            rel.append(abs((Lc[0] - Lc[1]) / (2 * eps) - an) / abs(an))
        best.append(min(rel))
    print(f"lm {lm}: adjoint vs central differences of the CPU build, best eps per direction: {best}")
    assert max(best) <= 1e-5 and float(np.median(best)) <= 1e-6, best
    adj.close()
    s.close()
