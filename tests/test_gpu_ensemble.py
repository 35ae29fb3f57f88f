"""The batched ensemble: members of uneven sizes, each with its own dt, in one
launch per RK stage.  Host dt mode: every member bitwise equal to its solo run.
Device dt mode with graph replay: every member bitwise equal to its solo
device-dt run, and close to the host-dt run."""
import numpy as np
import pytest

from test_oracle_golden import bits_equal

pytestmark = pytest.mark.gpu


def members(n=7, seed=11, smin=24, smax=140):
    from openmhd_amd import make_config, problems

    params = problems.kh_ensemble(n, seed=seed, smin=smin, smax=smax)
    cfgs, states = [], []
    for p in params:
        U, dx = problems.kh_member(p)
        cfgs.append(make_config(p["nx"], p["ny"], dx, cfl=0.4, lm_type=1, bc=("periodic", "wall")))
        states.append(U)
    return cfgs, states


def solo(cfg, U, nsteps):
    from openmhd_amd import Solver

    s = Solver(cfg.nx, cfg.ny, cfg.dx, cfl=cfg.cfl, lm_type=cfg.lm_type, bc=("periodic", "wall"))
    s.upload(U)
    dts, _ = s.step(nsteps, trace=True)
    out = s.download()
    s.close()
    return dts, out


def test_host_dt_members_equal_solo_runs():
    from openmhd_amd import Ensemble

    cfgs, states = members()
    e = Ensemble(cfgs)
    for i, U in enumerate(states):
        e.upload(i, U)
    tr = e.step(15, trace=True)
    for i, (c, U) in enumerate(zip(cfgs, states)):
        dts, Us = solo(c, U, 15)
        assert bits_equal(tr[:, i], dts), i
        assert bits_equal(e.download(i), Us), i
    # members really do take different steps
    assert len(set(np.round(tr[0], 12))) > 1
    e.close()


def test_autotune_keeps_state_and_results():
    from openmhd_amd import Ensemble

    cfgs, states = members(n=5, seed=3)
    e = Ensemble(cfgs)
    for i, U in enumerate(states):
        e.upload(i, U)
    rows, grid = e.autotune(2)
    assert rows > 0 and grid >= 0
    tr = e.step(6, trace=True)
    for i, (c, U) in enumerate(zip(cfgs, states)):
        dts, Us = solo(c, U, 6)
        assert bits_equal(tr[:, i], dts) and bits_equal(e.download(i), Us), i
    e.close()


def test_device_dt_graph_replay():
    from openmhd_amd import Ensemble

    cfgs, states = members(n=6, seed=5)
    e = Ensemble(cfgs)
    for i, U in enumerate(states):
        e.upload(i, U)
    e.set_mode(device_dt=True, graph=True)
    for _ in range(4):  # several graph replays with host work in between
        e.step(3)
        _ = e.times()
    for i, (c, U) in enumerate(zip(cfgs, states)):
        one = Ensemble([c])  # the member's solo run in the same dt mode, eagerly
        one.upload(0, U)
        one.set_mode(device_dt=True, graph=False)
        one.step(12)
        assert bits_equal(e.download(i), one.download(0)), i
        assert bits_equal(e.times()[i], one.times()[0]), i
        one.close()
        _, Us = solo(c, U, 12)  # host-dt solo run: same up to the device exp
        Ue = e.download(i)
        rel = np.max(np.abs(Ue - Us)) / np.max(np.abs(Us))
        assert rel < 1e-10, (i, rel)
    e.close()


def test_benchmark_ensemble_members_equal_solo_runs():
    """The 256-member KH ensemble of the benchmark (sides 64-1024), host dt,
    autotuned launch: 3 steps, every member's dt sequence and state bitwise
    equal to its solo run."""
    from openmhd_amd import Ensemble

    cfgs, states = members(n=256, seed=2026, smin=64, smax=1024)
    print(f"ensemble of {len(cfgs)} members built", flush=True)
    e = Ensemble(cfgs)
    for i, U in enumerate(states):
        e.upload(i, U)
    e.autotune(2)
    tr = e.step(3, trace=True)
    for i, (c, U) in enumerate(zip(cfgs, states)):
        dts, Us = solo(c, U, 3)
        assert bits_equal(tr[:, i], dts), i
        assert bits_equal(e.download(i), Us), i
    e.close()
