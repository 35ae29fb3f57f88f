"""openmhd_amd.diagnostics against closed forms (CPU)."""
import numpy as np
import pytest

from openmhd_amd.diagnostics import div_b_l2, max_rel_error


def grid2d(nx, ny, dx):
    x = (np.arange(nx + 2) - 0.5) * dx
    y = (np.arange(ny + 2) - 0.5) * dx
    return np.meshgrid(x, y, indexing="ij")


def test_div_b_closed_forms():
    """B = (-sin y, sin 2x) has zero divergence and its central differences
    are exactly zero (each component is constant along its own axis);
    B = (a x, b y) has div B = a + b everywhere, so the norm is
    |a + b| sqrt(area of the cells that enter); 3D alike."""
    nx, ny, dx = 30, 20, 0.1
    X, Y = grid2d(nx, ny, dx)
    U = np.zeros((nx + 2, ny + 2, 9), order="F")
    U[..., 5], U[..., 6] = -np.sin(Y), np.sin(2 * X)
    assert div_b_l2(U, dx) == 0.0
    a, b = 0.7, -0.2
    U[..., 5], U[..., 6] = a * X, b * Y
    area = (nx - 2) * (ny - 2) * dx * dx
    assert div_b_l2(U, dx) == pytest.approx(abs(a + b) * np.sqrt(area), rel=1e-12)
    n3 = (8, 7, 6)
    g = [(np.arange(n + 2) - 0.5) * dx for n in n3]
    X3, Y3, Z3 = np.meshgrid(*g, indexing="ij")
    U3 = np.zeros(tuple(n + 2 for n in n3) + (9,), order="F")
    U3[..., 5], U3[..., 6], U3[..., 7] = 0.3 * X3, 0.4 * Y3, -0.2 * Z3
    vol = np.prod([n - 2 for n in n3]) * dx ** 3
    assert div_b_l2(U3, dx) == pytest.approx(0.5 * np.sqrt(vol), rel=1e-12)


def test_max_rel_error():
    """Per field max |dU| / max |Uref| over the interior; ghosts are ignored;
    an all-zero reference field reports the absolute error."""
    rng = np.random.default_rng(0)
    R = np.asfortranarray(rng.normal(size=(6, 5, 9)))
    R[..., 8] = 0.0
    U = R.copy(order="F")
    U[0, 0, :] += 5.0  # a ghost cell
    assert not max_rel_error(U, R).any()
    U[2, 2, 3] += 1e-9
    U[3, 1, 8] = -2e-12
    e = max_rel_error(U, R)
    assert e[3] == pytest.approx(1e-9 / np.abs(R[1:-1, 1:-1, 3]).max(), rel=1e-6)
    assert e[8] == 2e-12 and not np.delete(e, [3, 8]).any()
