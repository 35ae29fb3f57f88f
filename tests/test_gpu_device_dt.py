"""Device dt mode of the resident step (omhd_set_mode).

The reference computes every step's dt on the host from the reduced CFL speed
(2D_basic/main.f90:80, set_dt.f90:21-65); so does omhd_step by default, at one
host round trip per step.  In device dt mode the GPU computes dt, the CFL
speed, dt/dx and the GLM factor (from a table of the host's exp values) after
each step's second stage.  These tests hold it to the host-dt results bit for
bit: states, time, dt/ch traces, over calls of several lengths, with and
without graph replay, across mode switches and re-uploads, and at a stop
(the same code, completed-step count, cell and value)."""
import numpy as np
import pytest

from omhd_testlib import prim_to_cons, smooth_state
from test_gpu_stop import CASES as STOP_CASES, low_beta

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


def kh(n=128):
    from openmhd_amd import Solver2D
    from openmhd_amd.problems import kelvin_helmholtz

    U, dx = kelvin_helmholtz(n, n, seed=3)
    s = Solver2D(n, n, dx, cfl=0.4, lm_type=1, time_type=0, bc=("periodic", "wall"))
    s.upload(U)
    return s, U


@pytest.mark.parametrize("graph", [False, True])
def test_kh_device_dt_equals_host_dt(graph):
    ref, _ = kh()
    dts, chs = ref.step(120, trace=True)
    W, t = ref.get_state(), ref.t
    ref.close()
    s, _ = kh()
    s.set_mode(device_dt=True, graph=graph)
    d1, c1 = s.step(50, trace=True)
    s.step(30)  # untraced call in between (another captured graph)
    d3, c3 = s.step(40, trace=True)
    assert np.array_equal(bits(s.get_state()), bits(W))
    assert bits(s.t) == bits(t)
    assert np.array_equal(bits(d1), bits(dts[:50])) and np.array_equal(bits(c1), bits(chs[:50]))
    assert np.array_equal(bits(d3), bits(dts[80:])) and np.array_equal(bits(c3), bits(chs[80:]))
    s.close()


@pytest.mark.parametrize("lm,tt,graph", [(2, 0, False), (1, 1, True)])
def test_3d_device_dt_equals_host_dt(lm, tt, graph):
    from openmhd_amd import Solver

    nx, ny, nz = 24, 20, 16
    U = np.asfortranarray(smooth_state(np.random.default_rng(5), nx + 2, ny + 2, nz + 2))
    runs = []
    for device in (False, True):
        s = Solver(nx, ny, 0.1, nz=nz, cfl=0.3, lm_type=lm, time_type=tt, bc=("periodic", "wall", "periodic"))
        s.upload(U)
        if device:
            s.set_mode(device_dt=True, graph=graph)
        d, c = s.step(25, trace=True)
        runs.append((s.get_state(), s.t, d, c))
        s.close()
    (W0, t0, d0, c0), (W1, t1, d1, c1) = runs
    assert np.array_equal(bits(W1), bits(W0)) and bits(t1) == bits(t0)
    assert np.array_equal(bits(d1), bits(d0)) and np.array_equal(bits(c1), bits(c0))


def test_mode_switches_and_reupload():
    ref, U = kh(96)
    ref.step(30)
    W30, t30 = ref.get_state(), ref.t
    ref.close()
    s, _ = kh(96)
    s.step(10)
    s.set_mode(device_dt=True, graph=True)
    s.step(10)
    s.set_mode(device_dt=False)
    s.step(10)
    assert np.array_equal(bits(s.get_state()), bits(W30)) and bits(s.t) == bits(t30)
    # a new state in device mode re-seeds the device parameters from it
    s.set_mode(device_dt=True)
    s.upload(U)
    s.t = 0.0
    s.step(30)
    assert np.array_equal(bits(s.get_state()), bits(W30)) and bits(s.t) == bits(t30)
    s.close()


@pytest.mark.parametrize("nz,tt,B0,p0,v0,where", STOP_CASES)
def test_stop_equals_host_dt(nz, tt, B0, p0, v0, where):
    """Low-beta states that lose positivity on U* or on U^{n+1}: device dt mode
    stops with the host mode's code, completed steps, cell and value, and
    leaves the same state; a later call returns the same stop."""
    from openmhd_amd import OMHDError, Solver

    n = 32 if nz == 0 else 20
    U = low_beta(n, B0, p0, v0, nz)
    out = []
    for device in (False, True):
        s = Solver(n, n, 1.0 / n, nz=nz, cfl=0.4, lm_type=2, time_type=tt, bc=("periodic",) * (3 if nz else 2))
        s.upload(U)
        if device:
            s.set_mode(device_dt=True)
        with pytest.raises(OMHDError) as e:
            s.step(100)
        out.append((e.value, s.get_state(), s.t))
        if device:
            with pytest.raises(OMHDError) as e2:
                s.step(5)
            assert e2.value.code == e.value.code and e2.value.steps_done == 0
        s.close()
    (eh, Wh, th), (ed, Wd, td) = out
    print(f"host: {eh.steps_done} steps, cell {eh.cell}; device: {ed.steps_done} steps, cell {ed.cell}")
    assert (ed.code, ed.steps_done, ed.cell) == (eh.code, eh.steps_done, eh.cell)
    assert bits(ed.value) == bits(eh.value)
    assert np.array_equal(bits(Wd), bits(Wh)) and bits(td) == bits(th)


def test_dt_too_small_stops_at_the_first_step():
    """A CFL speed so large that dt < 1e-7 (set_dt.f90:55-65): both modes stop
    before the first step with the same dt in the report."""
    from openmhd_amd import OMHDError, Solver2D

    n = 32
    x = (np.arange(n + 2) - 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    z = 0 * X
    pr = 1e12 * (1 + 0.1 * np.sin(2 * np.pi * Y))  # sound speed ~ 1.3e6: dt ~ 1e-8
    U = np.asfortranarray(prim_to_cons(np.stack([z, z, z, pr, z + 1.0, z + 0.5, z, z, z], -1)))
    errs = []
    for device in (False, True):
        s = Solver2D(n, n, 1.0 / n, cfl=0.4, lm_type=2)
        s.upload(U)
        if device:
            s.set_mode(device_dt=True)
        with pytest.raises(OMHDError) as e:
            s.step(3)
        errs.append(e.value)
        s.close()
    assert errs[1].code == errs[0].code and errs[1].steps_done == errs[0].steps_done == 0
    assert bits(errs[1].value) == bits(errs[0].value)


def test_refusals_and_timing():
    from openmhd_amd import OMHDError, Solver

    s, _ = kh(64)
    with pytest.raises(OMHDError):
        s.set_mode(device_dt=False, graph=True)  # graph replay needs device dt
    s.set_mode(device_dt=True)
    for device in (True, False):  # a negative step count is refused in both modes
        s.set_mode(device_dt=device)
        with pytest.raises(OMHDError):
            s.step(-1)
    s.set_mode(device_dt=True)
    s.set_timing(True)
    s.step(6)
    ms, launches = s.timing()
    assert launches == [6, 6] and ms[0] > 0 and ms[1] > 0
    s.close()
    r = Solver(8, 6, 0.1, nz=5, cfl=0.3, lm_type=2, bc=("periodic", "wall", "periodic"),
               resistivity=(1e-3, 2e-2, 0.5))
    with pytest.raises(OMHDError):
        r.set_mode(device_dt=True)  # ideal MHD only
    r.close()


def test_adjoint_forward_in_device_dt_mode():
    """The adjoint's checkpointed forward pass calls omhd_step: in device dt
    mode its dt/ch sequence, the checkpoints and the gradient equal host dt
    mode's bit for bit; set_state re-seeds the device parameters."""
    from openmhd_amd import Adjoint

    grads = []
    for device in (False, True):
        s, U = kh(64)
        if device:
            s.set_mode(device_dt=True, graph=True)
        W0 = s.get_state()
        s.set_state(W0)  # a new state in device mode: re-seeded from its CFL speed
        adj = Adjoint(s, 12, 4)
        dts, chs = adj.forward()
        lam = adj.backward(np.random.default_rng(4).standard_normal(W0.shape))
        grads.append((dts, chs, lam, s.get_state()))
        adj.close()
        s.close()
    for a, b in zip(grads[0], grads[1]):
        assert np.array_equal(bits(a), bits(b))
