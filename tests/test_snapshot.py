"""Snapshot files (openmhd_amd.snapshot) against files written by the
reference's own fileio_output (2D_basic/fileio.f90:16-52, 3D_basic/fileio.f90:
16-54, committed in tests/golden/): reading gives the reference's arrays
bitwise, and writing the same arrays gives the same bytes."""
from pathlib import Path

import numpy as np
import pytest

from openmhd_amd import snapshot

G = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("dim", (2, 3))
def test_reference_snapshot_round_trip(tmp_path, dim):
    d = np.load(G / f"snapshot_ref{dim}d.npz", allow_pickle=False)
    ref = (G / f"snapshot_ref{dim}d.dat").read_bytes()
    got = snapshot.read(G / f"snapshot_ref{dim}d.dat", three_d=(dim == 3))
    t, axes, U, V = got[0], got[1:1 + dim], got[-2], got[-1]
    assert t == float(d["t"])
    for i, a in enumerate(axes):
        assert np.array_equal(a, d[f"axis{i}"])
    assert np.array_equal(U.view(np.int64), d["U"].view(np.int64))
    assert np.array_equal(V.view(np.int64), d["V"].view(np.int64))
    out = tmp_path / "s.dat"
    z = d["axis2"] if dim == 3 else None
    snapshot.write(out, float(d["t"]), d["axis0"], d["axis1"], d["U"], V=d["V"], z=z)
    assert out.read_bytes() == ref
