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
    """Device dt with whole steps replayed from a hipGraph: every member, its
    time and its state, bitwise equal to its solo run (host dt, host libm exp
    for the GLM factor, which the device looks up in the member's table) and to
    the same member run alone in the same mode without the graph."""
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
        one = Ensemble([c])  # the member alone in the same dt mode, eagerly
        one.upload(0, U)
        one.set_mode(device_dt=True, graph=False)
        one.step(12)
        assert bits_equal(e.download(i), one.download(0)), i
        assert bits_equal(e.times()[i], one.times()[0]), i
        one.close()
        dts, Us = solo(c, U, 12)  # the plain solo run
        assert bits_equal(e.download(i), Us), i
        assert bits_equal(e.times()[i], sum_in_order(dts)), i
    e.close()


def sum_in_order(v):
    t = 0.0
    for x in v:
        t = t + float(x)
    return t


def test_benchmark_ensemble_members_equal_solo_runs():
    """The 256-member KH ensemble of the benchmark (sides 64-1024), autotuned
    launch, 3 steps, in both dt modes (host dt; device dt with hipGraph
    replay): every member's state bitwise equal to its solo run, and in host
    dt mode its dt sequence too."""
    from openmhd_amd import Ensemble

    cfgs, states = members(n=256, seed=2026, smin=64, smax=1024)
    print(f"ensemble of {len(cfgs)} members built", flush=True)
    ref = [solo(c, U, 3) for c, U in zip(cfgs, states)]
    for device_dt in (False, True):
        print(f"ensemble device_dt={device_dt}", flush=True)
        e = Ensemble(cfgs)
        for i, U in enumerate(states):
            e.upload(i, U)
        if device_dt:
            e.set_mode(device_dt=True, graph=True)
        e.autotune(2)
        tr = e.step(3, trace=not device_dt)
        for i, (dts, Us) in enumerate(ref):
            if not device_dt:
                assert bits_equal(tr[:, i], dts), i
            assert bits_equal(e.download(i), Us), (device_dt, i)
            assert bits_equal(e.times()[i], sum_in_order(dts)), (device_dt, i)
        e.close()


def test_upload_in_device_dt_mode():
    """A member (re)uploaded after the switch to device dt mode starts from its
    new state with its own dt, GLM factor and time 0: after more steps it is
    bitwise its solo run from that state, and the members left alone are
    unaffected.  Stepping before every member has a state is refused."""
    from openmhd_amd import Ensemble, OMHDError

    cfgs, states = members(n=4, seed=9)
    e = Ensemble(cfgs)
    for i, U in enumerate(states[:-1]):
        e.upload(i, U)
    with pytest.raises(OMHDError):
        e.step(1)  # the last member has no state yet
    e.upload(len(states) - 1, states[-1])
    e.set_mode(device_dt=True, graph=True)
    e.step(4)
    # member 1 restarts from member 1's state of a different ensemble draw
    # (same grid), with the speeds, dt and GLM factor of that state
    from openmhd_amd import problems

    p = problems.kh_ensemble(4, seed=9, smin=24, smax=140)[1]  # the draw members() made
    U_new, _ = problems.kh_member(dict(p, seed=p["seed"] + 1000))
    e.upload(1, U_new)
    e.step(5)
    dts, Us = solo(cfgs[1], U_new, 5)
    assert bits_equal(e.download(1), Us)
    assert bits_equal(e.times()[1], sum_in_order(dts))
    for i in (0, 2, 3):  # 9 steps since their upload
        dts, Us = solo(cfgs[i], states[i], 9)
        assert bits_equal(e.download(i), Us), i
        assert bits_equal(e.times()[i], sum_in_order(dts)), i
    e.close()
