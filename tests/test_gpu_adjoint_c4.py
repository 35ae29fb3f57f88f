"""The 2D adjoint at its own configuration (the benchmark's adjoint case):
Orszag-Tang at 512^2 plus a velocity perturbation of 16 seeded Fourier modes
(|k| <= 8, amplitude 1e-2), 200 MC/HLLD/GLM/TVD-RK2 steps with checkpoints
every 20, loss ||By(T) - By_target||^2 with the target a forward run from a
second seeded perturbation, gradient with respect to the 16 mode
coefficients through torch.autograd (Rollout).

Checked: (1) the forward pass, dt sequence and final state, bit for bit
against the oracle's 200-step run; (2) <grad, d> against central differences
of the loss with the frozen (dt, ch) along 16 seeded directions, best eps per
direction within 1e-6; (3) the reverse pass against the exact dual-number
tangent of the same discrete map (tests/native/adjcheck.cpp) over the first
20 steps along 2 directions, within 1e-12."""
import ctypes as C

import numpy as np
import pytest

from omhd_testlib import AdjCheck, cptr, make_problem, soa
from test_oracle_golden import bits_equal

pytestmark = pytest.mark.gpu

N, NSTEPS, EVERY = 512, 200, 20


@pytest.fixture(scope="module")
def c4():
    import torch

    from openmhd_amd import Rollout, Solver, problems

    dx = 2 * np.pi / N
    s = Solver(N, N, dx, cfl=0.4, lm_type=2)
    modes, tmodes = problems.fourier_modes(1), problems.fourier_modes(2)
    tgt_coef = np.full(16, 1e-2) * np.random.default_rng(2).uniform(0.5, 1.5, 16)
    s.set_state(problems.orszag_tang_perturbed(N, tmodes, tgt_coef))
    s.step(NSTEPS)
    target = torch.tensor(s.get_state()[6], device="cuda")
    roll = Rollout(s, NSTEPS, EVERY)

    def loss_of(coef, dt=None, ch=None):
        W0 = problems.orszag_tang_perturbed(N, modes, coef)
        WT = roll(W0, dt, ch)
        return ((WT[6] - target) ** 2).sum(), W0, WT

    coef = torch.full((16,), 1e-2, dtype=torch.float64, device="cuda", requires_grad=True)
    L, W0, WT = loss_of(coef)
    L.backward()
    out = dict(s=s, dx=dx, loss_of=loss_of, coef=coef.detach(), grad=coef.grad.detach().cpu().numpy(),
               W0=W0.detach().cpu().numpy(), WT=WT.detach().cpu().numpy(), dts=roll.dt.copy(), chs=roll.ch.copy())
    yield out
    s.close()


def test_c4_forward_equals_oracle(c4, oracle):
    dx, W0 = c4["dx"], c4["W0"]
    U = np.zeros((N + 2, N + 2, 9), order="F")
    U[1:-1, 1:-1, :] = W0.transpose(2, 1, 0)
    p = make_problem(N + 2, N + 2, dx=dx, cfl=0.4, lm_type=2)
    oracle.bc_u(U, p)
    rc, odt, och = oracle.run(p, U, NSTEPS)
    assert rc == 0
    assert bits_equal(c4["dts"], odt) and bits_equal(c4["chs"], och)
    assert bits_equal(c4["WT"], soa(U))


def test_c4_gradient_against_central_differences(c4):
    """16 seeded directions in coefficient space; eps in {1e-7, 1e-8, 1e-9,
    1e-10} (the map is piecewise smooth: a larger eps crosses limiter kinks, a
    smaller one sees the forward's rounding noise); best eps per direction."""
    import torch

    rng = np.random.default_rng(16)
    best = []
    with torch.no_grad():
        for k in range(16):
            d = rng.standard_normal(16)
            d /= np.linalg.norm(d)
            dv = torch.tensor(d, device="cuda")
            an = float(c4["grad"] @ d)
            rel = []
            for eps in (1e-7, 1e-8, 1e-9, 1e-10):
                lp = float(c4["loss_of"](c4["coef"] + eps * dv, c4["dts"], c4["chs"])[0])
                lm = float(c4["loss_of"](c4["coef"] - eps * dv, c4["dts"], c4["chs"])[0])
                rel.append(abs((lp - lm) / (2 * eps) - an) / abs(an))
            best.append(min(rel))
            print(f"direction {k}: best relative difference {best[-1]:.2e}", flush=True)
    assert max(best) <= 1e-6, best


def test_c4_reverse_against_exact_tangent(c4):
    """The first 20 steps of the same run (imposed dt/ch, one checkpoint
    segment): <dL/dW0, v> from the GPU reverse equals <lamT, J v> from the
    dual-number forward along 2 seeded smooth directions, to 1e-12."""
    from openmhd_amd import Adjoint

    ac = AdjCheck()
    s, dx, W0 = c4["s"], c4["dx"], c4["W0"]
    n20 = 20
    dts, chs = c4["dts"][:n20].copy(), c4["chs"][:n20].copy()
    x = (np.arange(N) + 0.5) * dx
    X, Y = np.meshgrid(x, x, indexing="xy")
    rng = np.random.default_rng(20)

    def smooth():
        v = np.zeros((9, N, N))
        for m in range(9):
            for _ in range(3):
                kx, ky = rng.integers(-4, 5, 2)
                v[m] += rng.standard_normal() * np.sin(kx * X + ky * Y + rng.uniform(0, 2 * np.pi))
        return v

    lamT = smooth()
    s.set_state(W0)
    adj = Adjoint(s, n20, n20)
    adj.forward(dts, chs)
    grad = adj.backward(lamT)
    adj.close()
    for k in range(2):
        v = smooth()
        jv = np.zeros_like(W0)
        ac.run_jvp(N, N, 2, 0, C.c_double(dx), n20, cptr(dts), cptr(chs), cptr(np.ascontiguousarray(W0)), cptr(v),
                   cptr(jv))
        an, tl = float((grad * v).sum()), float((lamT * jv).sum())
        print(f"direction {k}: adjoint {an:.15e} tangent {tl:.15e}", flush=True)
        assert abs(an - tl) <= 1e-12 * abs(tl), (k, an, tl)
