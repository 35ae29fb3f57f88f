// omhd_internal.h -- what the libomhd translation units share beyond the C
// ABI: the resident handle and the single-stage launcher the adjoint driver
// uses to replay the forward step into its own buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/omhd.h"

namespace omhd {

// boundary type and the per-face sign masks (bit m: negate slot m in the ghost)
struct Faces {
  int bc[3];
  unsigned slo[3], shi[3];
};

}  // namespace omhd

struct omhd_handle {
  omhd_config cfg;
  int nd = 2;
  int pitch = 0;
  long long zstride = 0, plane = 0;
  double *U = nullptr;    // U^n / U^{n+1}
  double *U1 = nullptr;   // U*
  double *acc = nullptr;  // 3D: partial flux difference
  double *eta = nullptr;  // resistive face arrays
  double eta_max = 0.0;
  unsigned long long *red = nullptr;       // [0] vmax bits, [1..2] error (4 ints)
  unsigned long long *red_host = nullptr;  // pinned mirror
  hipStream_t stream = nullptr;
  double vmax = -1.0;  // CFL speed of the current state
  bool have_state = false;
  omhd::Faces faces{};
  int rows = 24;
  // optional per-stage timing with HIP events on the handle's stream
  bool timing = false;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  double stage_ms[2] = {0.0, 0.0};
  long long stage_launches[2] = {0, 0};
  // device dt mode (omhd_set_mode): each step's dt, CFL speed and GLM factor
  // are computed on the device from the stage-2 reduction, so a call of n
  // steps needs no host round trip until its end
  bool device_dt = false, use_graph = false;
  void *solo = nullptr;       // device block: step parameters, time, step count, GLM table
  void *solo_host = nullptr;  // pinned mirror of it
  double *trace_d = nullptr;  // device dt/ch trace, 2 doubles per step
  long long trace_cap = 0;
  hipGraph_t graph = nullptr;  // one captured step (stages + the dt kernel)
  hipGraphExec_t gexec = nullptr;
  bool graph_traced = false;            // whether the captured step writes the trace
  std::vector<hipEvent_t> step_ev;      // per-step timing events of device dt mode
  int halt_rc = 0;                      // a stop found on the device, returned again by later calls
  int halt_err[4] = {0, 0, 0, 0};
  double halt_dt = 0.0;
};


namespace omhd {

// One RK stage of the resident step with explicit state pointers (the handle's
// own buffers are not touched): stage 1 reads `us` (psi scaled by glm_f on
// load) and writes U* to `out`; stage 2 reads U* = `us` and U^n = `u0` and
// writes U^{n+1} to `out`.  dt and ch are the step's (frozen) values, unless
// dpar names device-side step parameters (device dt mode), which the kernels
// read instead.  Launched on the handle's stream; returns an OMHD status.
int stage_launch(omhd_handle *h, int stage, const double *us, const double *u0, double *out, double dt, double ch,
                 const void *dpar = nullptr);

// The handle's CFL speed from its current state (the upload pass).
int refresh_cfl(omhd_handle *h);

}  // namespace omhd
