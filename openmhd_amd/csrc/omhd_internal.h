// omhd_internal.h -- what the libomhd translation units share beyond the C
// ABI: the resident handle and the single-stage launcher the adjoint driver
// uses to replay the forward step into its own buffers.
#pragma once
#include <hip/hip_runtime.h>

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
};


namespace omhd {

// One RK stage of the resident step with explicit state pointers (the handle's
// own buffers are not touched): stage 1 reads `us` (psi scaled by glm_f on
// load) and writes U* to `out`; stage 2 reads U* = `us` and U^n = `u0` and
// writes U^{n+1} to `out`.  dt and ch are the step's (frozen) values.  Launched
// on the handle's stream; returns an OMHD status.
int stage_launch(omhd_handle *h, int stage, const double *us, const double *u0, double *out, double dt, double ch);

// The handle's CFL speed from its current state (the upload pass).
int refresh_cfl(omhd_handle *h);

}  // namespace omhd
