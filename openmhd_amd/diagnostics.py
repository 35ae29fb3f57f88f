"""Run diagnostics, applied identically to a run of this library and a run of
the reference CPU build: the per-field relative error and the L2 norm of
div B.  The reference has no div B diagnostic of its own; the operator here is
defined once and used on both sides.

Arrays are U(ix, jx, 9) or U(ix, jx, kx, 9) in Fortran order with the
reference's one ghost layer (slot order of param.h:6-15).  Only interior cells
enter, so ghost contents never matter.
"""
from __future__ import annotations

import numpy as np


def _interior(U):
    if U.ndim == 3:
        return U[1:-1, 1:-1, :]
    return U[1:-1, 1:-1, 1:-1, :]


def max_rel_error(U, Uref) -> np.ndarray:
    """Per field: max |U - Uref| / max |Uref| over the interior cells.  A field
    that is zero in the reference reports its absolute error instead."""
    a = _interior(np.asarray(U, dtype=np.float64))
    b = _interior(np.asarray(Uref, dtype=np.float64))
    nv = a.shape[-1]
    out = np.empty(nv)
    for m in range(nv):
        d = np.abs(a[..., m] - b[..., m]).max()
        s = np.abs(b[..., m]).max()
        out[m] = d / s if s > 0 else d
    return out


def div_b_l2(U, dx, dy=None, dz=None) -> float:
    """sqrt(sum (div B)^2 dV) over the interior cells that have both
    neighbours inside the interior, with div B by second-order central
    differences of the cell-centred B: (Bx(i+1)-Bx(i-1))/(2 dx) + ...
    (the z term in 3D)."""
    dy = dx if dy is None else dy
    dz = dx if dz is None else dz
    I = _interior(np.asarray(U, dtype=np.float64))
    if I.ndim == 3:
        bx, by = I[..., 5], I[..., 6]
        div = (bx[2:, 1:-1] - bx[:-2, 1:-1]) / (2 * dx) + (by[1:-1, 2:] - by[1:-1, :-2]) / (2 * dy)
        dv = dx * dy
    else:
        bx, by, bz = I[..., 5], I[..., 6], I[..., 7]
        div = ((bx[2:, 1:-1, 1:-1] - bx[:-2, 1:-1, 1:-1]) / (2 * dx)
               + (by[1:-1, 2:, 1:-1] - by[1:-1, :-2, 1:-1]) / (2 * dy)
               + (bz[1:-1, 1:-1, 2:] - bz[1:-1, 1:-1, :-2]) / (2 * dz))
        dv = dx * dy * dz
    return float(np.sqrt(np.sum(div * div) * dv))


def reconnected_flux(W, dx, j0=None):
    """Reconnected flux of a 3D current sheet centred at y = 0: half the
    z-averaged unsigned By flux through the sheet plane,
    phi = 1/2 * sum_{i,k} |By_c(i, k)| dx / nz, with By_c the mean of the two
    rows that straddle y = 0 (rows j0-1 and j0, j0 = ny/2).  W is the
    interior state (9, nz, ny, nx), x fastest.  Returns (phi, w), where w holds
    dphi/dBy on those two rows (the derivative of |.| taken with the sign of
    By_c; as a (nz, nx) array, equal on both rows)."""
    nz, ny = W.shape[1], W.shape[2]
    j0 = ny // 2 if j0 is None else j0
    byc = 0.5 * (W[6, :, j0 - 1, :] + W[6, :, j0, :])
    phi = 0.5 * float(np.abs(byc).sum()) * dx / nz
    return phi, 0.25 * np.sign(byc) * dx / nz
