"""Exact tangents of the 3D resistive configuration's reconnected flux with
respect to the resistivity parameters (eta1, w), at the full 512x256x256
shape, for tests/test_gpu_c5_gradient.py.  CPU only.

The Harris sheet of openmhd_amd.problems.harris_sheet_3d (dx = 1/8, x and z
periodic, y walls, minmod, HLLD, TVD-RK2, CFL 0.4, eta = 1e-3 +
1e-2 exp(-(x^2+y^2)/4)) is advanced 4 steps by the oracle, which gives the
(dt, ch) sequence and the final state; the flux phi and its derivative w with
respect to the final By come from openmhd_amd.diagnostics.reconnected_flux.
The dual-number forward of tests/native/adjcheck.cpp (adjc3_run_jvp, the
general 3D form: walls, resistive flux; its forward is the oracle's
arithmetic) then gives dW(T)/d eta1 and dW(T)/d w exactly, and
dphi/d eta1 = <w, dBy(T)/d eta1>, likewise for w.  Writes
tests/golden/c5_tangent.npz.  Takes about 30 minutes and 50 GB of memory.

    python3 tests/golden/gen_c5_tangent.py
"""
import ctypes as C
import hashlib
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parents[1]))

from omhd_testlib import DP, AdjCheck, Oracle, cptr, make_problem  # noqa: E402

from openmhd_amd import problems, snapshot  # noqa: E402
from openmhd_amd.diagnostics import reconnected_flux  # noqa: E402

NX, NY, NZ, DX, NSTEPS = 512, 256, 256, 0.125, 4
ETA = (1e-3, 1e-2, 2.0)
BC = ("periodic", "wall", "periodic")


def main():
    t0 = time.time()
    U, origin = problems.harris_sheet_3d(NX, NY, NZ, DX)
    orc = Oracle()
    p = make_problem(NX + 2, NY + 2, NZ + 2, dx=DX, cfl=0.4, lm_type=1, time_type=0, bc=BC)
    x = snapshot.grid(NX, DX, origin[0])
    y = snapshot.grid(NY, DX, origin[1])
    eta, emax = orc.eta_faces(x, y, DX, *ETA)
    eta = np.ascontiguousarray(eta)
    p.eta = eta.ctypes.data_as(DP)
    p.eta_max = emax
    orc.bc_u(U, p)
    W0 = np.ascontiguousarray(U[1:-1, 1:-1, 1:-1, :].transpose(3, 2, 1, 0))
    rc, dts, chs = orc.run(p, U, NSTEPS)
    assert rc == 0, rc
    j0 = NY // 2
    by_rows = np.ascontiguousarray(U[1:-1, j0 - 1 + 1:j0 + 1 + 1, 1:-1, 6].transpose(2, 1, 0))  # (nz, 2, nx)
    del U
    rows = np.zeros((9, NZ, 2, NX))  # the two rows the flux reads (only By enters)
    rows[6] = by_rows
    phi, w = reconnected_flux(rows, DX, j0=1)
    print(f"oracle: {NSTEPS} steps in {time.time() - t0:.0f} s, phi = {phi!r}", flush=True)
    # d eta / d eta1 and d eta / d w at the face centres (x-faces at x + dx/2,
    # y-faces at y + dx/2, z-faces at the cell centre)
    e0, e1, ew = ETA
    de1, dw = np.zeros((3, NY + 2, NX + 2)), np.zeros((3, NY + 2, NX + 2))
    for d in range(3):
        X, Y = np.meshgrid(x + (0.5 * DX if d == 0 else 0.0), y + (0.5 * DX if d == 1 else 0.0), indexing="xy")
        r2 = X * X + Y * Y
        de1[d] = np.exp(-r2 / (ew * ew))
        dw[d] = e1 * np.exp(-r2 / (ew * ew)) * 2 * r2 / ew ** 3
    ac = AdjCheck()
    bca = (C.c_int * 3)(0, 1, 0)
    sg = np.ascontiguousarray(np.array(p.u_sgn[:], dtype=np.float64))
    dU0 = np.zeros_like(W0)
    out = {}
    for name, de in (("dphi_deta1", de1), ("dphi_dw", dw)):
        t1 = time.time()
        jv = np.zeros_like(W0)
        ac.L.adjc3_run_jvp(NX, NY, NZ, 1, 0, C.c_double(DX), bca, cptr(sg), cptr(eta.reshape(-1)),
                           cptr(np.ascontiguousarray(de.reshape(-1))), NSTEPS, cptr(dts), cptr(chs), cptr(W0),
                           cptr(dU0), cptr(jv))
        out[name] = float((w * (jv[6, :, j0 - 1, :] + jv[6, :, j0, :])).sum())
        del jv
        print(f"{name} = {out[name]!r} ({time.time() - t1:.0f} s)", flush=True)
    np.savez(HERE / "c5_tangent.npz", dt=dts, ch=chs, phi=phi, dphi_deta1=out["dphi_deta1"],
             dphi_dw=out["dphi_dw"], eta=np.array(ETA), nsteps=NSTEPS,
             by_rows_sha256=hashlib.sha256(by_rows.tobytes()).hexdigest())
    print(f"wrote c5_tangent.npz in {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
