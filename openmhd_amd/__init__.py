"""openmhd_amd -- OpenMHD's finite-volume time step on MI355X (gfx950).

The compute is libomhd.so (hand-written HIP, built in-tree); this package is a
thin ctypes layer over its C ABI (include/omhd.h):

  openmhd_amd.api      the reference subroutines (u2v, limiter, flux_solver,
                       glm_ss2, rk_tvd21/22, rk_std21/22, set_dt, bc, bc_for_*,
                       and 2D_reconnection's flux_resistive, set_dt_resistive)
  openmhd_amd.solver   Solver (2D/3D): resident state, fused whole-step kernels
  openmhd_amd.problems initial states of the benchmark configurations
  openmhd_amd.diagnostics  per-field relative error and ||div B||_2 of a run
  openmhd_amd.snapshot the reference's snapshot file format (fileio.f90)
  openmhd_amd.adjoint  discrete adjoint of the 2D step, torch.autograd.Function
"""
from ._lib import LIB_PATH, OMHDError, lib  # noqa: F401  (raises ImportError if libomhd is not built)
from . import adjoint, api, problems, solver  # noqa: F401
from .adjoint import Adjoint, Rollout  # noqa: F401
from .solver import Ensemble, Solver, Solver2D, make_config  # noqa: F401

__all__ = ["api", "solver", "problems", "adjoint", "Solver", "Solver2D", "Ensemble", "Adjoint", "Rollout", "make_config",
           "OMHDError", "LIB_PATH"]
