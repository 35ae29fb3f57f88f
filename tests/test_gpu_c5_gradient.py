"""The resistivity gradient of the 3D resistive configuration at its own shape
(512x256x256, dx = 1/8, x/z periodic, y walls, minmod + HLLD + resistive flux
+ GLM + TVD-RK2, eta = eta0 + eta1 exp(-(x^2+y^2)/w^2) with (1e-3, 1e-2, 2)):
d phi / d(eta1, w) of the reconnected flux after 4 steps, from the GPU
adjoint with checkpoints every 2 steps, against exact tangents computed on
the CPU by the dual-number forward of the same discrete map
(tests/golden/gen_c5_tangent.py -> c5_tangent.npz).  Unlike central
differences at this size, the comparison has no rounding-noise floor: the
bar is 1e-10 relative.  The forward pass's dt sequence and flux must equal
the oracle's bit for bit."""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIX = Path(__file__).resolve().parent / "golden" / "c5_tangent.npz"


def test_c5_eta_gradient_against_exact_tangents():
    from openmhd_amd import Adjoint, Solver, problems
    from openmhd_amd.diagnostics import reconnected_flux

    d = np.load(FIX, allow_pickle=False)
    nsteps = int(d["nsteps"])
    eta = tuple(float(v) for v in d["eta"])
    nx, ny, nz, dx = 512, 256, 256, 0.125
    U, origin = problems.harris_sheet_3d(nx, ny, nz, dx)
    print("harris 512x256x256 built", flush=True)
    s = Solver(nx, ny, dx, nz=nz, cfl=0.4, lm_type=1, bc=("periodic", "wall", "periodic"), resistivity=eta,
               origin=origin)
    s.upload(U)
    del U
    adj = Adjoint(s, nsteps, 2)
    dts, chs = adj.forward()
    assert np.array_equal(dts.view(np.int64), d["dt"].view(np.int64))
    assert np.array_equal(chs.view(np.int64), d["ch"].view(np.int64))
    WT = s.get_state()
    phi, w = reconnected_flux(WT, dx)
    assert phi == float(d["phi"]), (phi, float(d["phi"]))
    lamT = np.zeros_like(WT)
    j0 = ny // 2
    lamT[6, :, j0 - 1, :] = w
    lamT[6, :, j0, :] = w
    del WT
    adj.backward(lamT)
    g = adj.eta_grad()
    adj.close()
    s.close()
    for k, name in ((1, "dphi_deta1"), (2, "dphi_dw")):
        ref = float(d[name])
        print(f"{name}: adjoint {g[k]:.15e} tangent {ref:.15e} rel {abs(g[k] - ref) / abs(ref):.1e}", flush=True)
        assert abs(g[k] - ref) <= 1e-10 * abs(ref), (name, g[k], ref)
