"""A run that loses positivity stops where the reference stops.

The reference's u2v prints the first cell (column-major, ghosts included) of
minimum pressure and stops (2D_basic/u2v.f90:34-43), at the start of a step
or on U* inside it.  The resident step must report the same error code, the
same number of completed steps, the same cell and, where the failing state
is the one the oracle hands back, the same pressure.  The states are low-beta
(strong By against a small pressure), which loses positivity within a few
steps: on U* with TVD-RK2, and on U^{n+1} (the next step's u2v) with the
midpoint RK2."""
import numpy as np
import pytest

from omhd_testlib import make_problem, prim_to_cons
from test_oracle_golden import bits_equal

pytestmark = pytest.mark.gpu


def low_beta(n, B0, p0, v0, nz=0):
    x = (np.arange(n + 2) - 0.5) / n
    if nz:
        X, Y, Z = np.meshgrid(x, x, (np.arange(nz + 2) - 0.5) / nz, indexing="ij")
    else:
        X, Y = np.meshgrid(x, x, indexing="ij")
        Z = 0 * X
    vx = v0 * np.sin(2 * np.pi * X) * (1 + 0.1 * np.cos(2 * np.pi * Y) + 0.05 * np.sin(2 * np.pi * Z))
    by = B0 * np.cos(2 * np.pi * X + 0.3 * np.sin(2 * np.pi * Y))
    z = 0 * X
    return np.asfortranarray(prim_to_cons(np.stack([vx, z, z, z + p0, z + 1.0, z + 0.2 * B0, by, z, z], -1)))


# (nz, time type, B0, p0, v0, where the oracle stops)
CASES = [(0, 0, 1.0, 0.03, 1.0, "ustar"), (0, 1, 1.0, 0.05, 0.5, "start"),
         (12, 0, 1.0, 0.03, 1.0, "ustar"), (12, 1, 2.0, 0.1, 0.5, "start")]


@pytest.mark.parametrize("nz,tt,B0,p0,v0,where", CASES)
def test_stop_matches_the_reference(oracle, nz, tt, B0, p0, v0, where):
    from openmhd_amd import OMHDError, Solver

    n = 32 if nz == 0 else 20
    U = low_beta(n, B0, p0, v0, nz)
    shape = U.shape[:-1]
    p = make_problem(*shape, dx=1.0 / n, cfl=0.4, lm_type=2, time_type=tt)
    oracle.bc_u(U, p)
    Uo = U.copy(order="F")
    rc, dto, _ = oracle.run(p, Uo, 100)
    assert rc == 1  # negative pressure
    done = int(np.argmax(dto == 0))
    bad = oracle.last_bad
    Vo = np.zeros(shape + (4,), order="F")
    stopped_at_start = oracle.u2v(Uo, Vo)[0] != 0  # the state handed back is the one that failed
    assert stopped_at_start == (where == "start")
    ix, jx = shape[0], shape[1]
    cell = (bad % ix + 1, bad // ix % jx + 1, bad // (ix * jx) + 1)

    s = Solver(n, n, 1.0 / n, nz=nz, cfl=0.4, lm_type=2, time_type=tt,
               bc=("periodic",) * (3 if nz else 2))
    s.upload(U)
    with pytest.raises(OMHDError) as e:
        s.step(100)
    s.close()
    print(f"oracle: {done} steps, cell {cell}; gpu: {e.value.steps_done} steps, cell {e.value.cell}")
    assert e.value.code == 1
    assert e.value.steps_done == done
    assert e.value.cell == cell
    assert e.value.value <= 0
    if stopped_at_start:
        assert bits_equal(np.float64(e.value.value), np.float64(Vo.reshape(-1, 4, order="F")[bad, 3]))
