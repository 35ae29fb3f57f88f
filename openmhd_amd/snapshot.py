"""Snapshot files in the reference's format (2D_basic/fileio.f90:16-115,
3D_basic/fileio.f90): stream-unformatted, native endian,

    t:f8, ix:i4, jx:i4, [kx:i4], x(ix):f8, y(jx):f8, [z(kx):f8],
    U: 9 planes (mx,my,mz,en,ro,bx,by,bz,ps), V: 4 planes (vx,vy,vz,pr),

each plane column-major (i fastest) including the ghost layer.  The same
files feed the reference CPU build and this library, and they are what the
reference's readers (visualization/openmhd2d.py) load.
"""
from __future__ import annotations

import numpy as np


def write(path, t, x, y, U, V=None, z=None):
    U = np.asfortranarray(U, dtype=np.float64)
    ix, jx = U.shape[0], U.shape[1]
    three_d = U.ndim == 4
    if V is None:
        V = np.zeros(U.shape[:-1] + (4,), order="F")
    with open(path, "wb") as f:
        f.write(np.float64(t).tobytes())
        f.write(np.int32(ix).tobytes())
        f.write(np.int32(jx).tobytes())
        if three_d:
            f.write(np.int32(U.shape[2]).tobytes())
        f.write(np.asarray(x, dtype=np.float64).tobytes())
        f.write(np.asarray(y, dtype=np.float64).tobytes())
        if three_d:
            f.write(np.asarray(z, dtype=np.float64).tobytes())
        for m in range(9):
            f.write(np.asfortranarray(U[..., m]).tobytes(order="F"))
        for m in range(4):
            f.write(np.asfortranarray(V[..., m]).tobytes(order="F"))


def read(path, three_d=False):
    """Returns (t, x, y, [z,] U, V) with U/V Fortran-ordered."""
    with open(path, "rb") as f:
        t = np.frombuffer(f.read(8), np.float64)[0]
        ix = int(np.frombuffer(f.read(4), np.int32)[0])
        jx = int(np.frombuffer(f.read(4), np.int32)[0])
        kx = int(np.frombuffer(f.read(4), np.int32)[0]) if three_d else 1
        x = np.frombuffer(f.read(8 * ix), np.float64).copy()
        y = np.frombuffer(f.read(8 * jx), np.float64).copy()
        z = np.frombuffer(f.read(8 * kx), np.float64).copy() if three_d else None
        shp = (ix, jx, kx) if three_d else (ix, jx)
        n = ix * jx * kx
        U = np.empty(shp + (9,), order="F")
        for m in range(9):
            U[..., m] = np.frombuffer(f.read(8 * n), np.float64).reshape(shp, order="F")
        rest = f.read()
        V = None
        if len(rest) >= 4 * 8 * n:
            V = np.empty(shp + (4,), order="F")
            for m in range(4):
                V[..., m] = np.frombuffer(rest[8 * n * m:8 * n * (m + 1)], np.float64).reshape(shp, order="F")
    return (t, x, y, z, U, V) if three_d else (t, x, y, U, V)


def grid(n, dx, x0):
    """Cell-centre coordinates including ghosts: x0 - dx/2 + i*dx, i = 0..n+1
    (2D_basic/model.f90:19-28)."""
    c = np.empty(n + 2)
    c[0] = x0 - dx / 2
    for i in range(1, n + 2):
        c[i] = c[i - 1] + dx
    return c
