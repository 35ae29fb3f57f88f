// omhd_step.hip -- the resident 2D step: fused sweep kernels + handle API.
//
// One TVD-RK2 step (2D_basic/main.f90:56-188) is two launches:
//   stage 1: U* = U + dt/dx * (dF + dG)              (U with psi *= f on load)
//   stage 2: U' = f_psi * 0.5 * ((U + U*) + dt/dx * (dF + dG)), plus the CFL
//            speed of U' (the next step's set_dt), reduced to one max.
// where every flux difference comes from u2v -> limiter -> HLLD on the fly.
//
// Column march: one 64-lane wavefront owns a strip of 60 interior columns
// (lanes 2..61; lanes 0,1,62,63 carry the two-cell x halo) and a chunk of rows.
// It walks the chunk bottom to top, loading each row of the state exactly once
// (coalesced 512-byte row segments per variable), and keeps a three-row window
// of primitives in registers.  Per row it computes the y-slopes of the middle
// row, the y-interface flux below it, the x-slopes / x-interface fluxes of the
// oldest row with DPP lane shifts, and writes that row's update.  Boundary
// conditions are index wraps (periodic) or mirrors with sign tables, and the
// wall face states are mirrors of the inside face state, as the reference's
// face patches make them (main.f90:101-102,119-120; 2D_KH/bc.f90:42-83).
//
// State layout in HBM: 9 planes (SoA), interior cells only, rows of `pitch`
// doubles, x fastest: state[m * plane + j * pitch + i].
#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <vector>

#include "omhd_device.h"
#include "omhd_host.h"

namespace omhd {
namespace {

constexpr int STRIP = 60;  // interior columns per wavefront (64 lanes - 2x2 halo)

struct Sweep2D {
  const double *us;  // state whose fluxes are computed (stage 1: U^n, stage 2: U*)
  const double *u0;  // stage 2: U^n, the RK partner
  double *out;       // stage 1: U*, stage 2: U^{n+1} (may alias u0)
  long long plane;   // doubles between variable planes
  int pitch;         // doubles between rows
  int nx, ny;
  int nstrip, rows, ntiles;
  int bcx, bcy;                        // OMHD_BC_* for both faces of the axis
  unsigned sx_lo, sx_hi, sy_lo, sy_hi; // negated slots of mirror faces
  double dtx, glm_f, ch, ch2;
  unsigned long long *vmax_bits;       // stage 2: max speed of U^{n+1}
  int *err;                            // [code, i, j] of the first failure
};

// consecutive workgroups go round-robin to the 8 XCDs: hand each XCD a
// contiguous run of tiles so neighbouring tiles share its L2 (nb % 8 == 0)
__device__ __forceinline__ int xcd_remap(int bid, int nb) { return (bid & 7) * (nb >> 3) + (bid >> 3); }

// index c in [-2, n+1] -> interior index and the negate mask of its ghost
__device__ __forceinline__ void map_axis(int c, int n, int bc, unsigned slo, unsigned shi, int &cm, unsigned &mask) {
  cm = c;
  mask = 0;
  if (c < 0) {
    if (bc == OMHD_BC_PERIODIC) cm = c + n;
    else { cm = -1 - c; mask = slo; }
  } else if (c >= n) {
    if (bc == OMHD_BC_PERIODIC) cm = c - n;
    else { cm = 2 * n - 1 - c; mask = shi; }
  }
  cm = cm < 0 ? 0 : (cm > n - 1 ? n - 1 : cm);  // lanes past the grid: any valid cell
}

__device__ __forceinline__ double flip(double v, unsigned mask, int m) { return ((mask >> m) & 1u) ? -v : v; }

__device__ __forceinline__ void record_error(int *err, int code, int i, int j) {
  if (atomicCAS(err, 0, code) == 0) {
    err[1] = i;
    err[2] = j;
  }
}

// Per wavefront LDS: a ring of three primitive rows (rn-2, rn-1, rn) and the
// carried left face state of the y-interface above row rn-2.  Keeping these
// out of VGPRs leaves ~18 live doubles across each inlined HLLD call, so two
// wavefronts fit per SIMD.
struct SweepLds {
  double ring[3][NV][66];  // [slot][var][lane + 1]; entries 0 and 65 repeat lanes 0 and 63
  double wly[NV][64];
};

template <int LIM, int STAGE, int TT, int NT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_sweep2d(const Sweep2D p) {
  __shared__ SweepLds lds;
  const int lane = threadIdx.x;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  if (t >= p.ntiles) return;
  const int strip = t % p.nstrip, chunk = t / p.nstrip;
  const int c = strip * STRIP - 2 + lane;
  const int j0 = chunk * p.rows;
  const int hc = min(p.rows, p.ny - j0);
  int cm;
  unsigned xmask;
  map_axis(c, p.nx, p.bcx, p.sx_lo, p.sx_hi, cm, xmask);
  const bool writer = lane >= 2 && lane < 2 + STRIP && c < p.nx;
  const bool xlo_face = p.bcx == OMHD_BC_MIRROR && c == -1;
  const bool xhi_face = p.bcx == OMHD_BC_MIRROR && c == p.nx - 1;
  const double glm_us = (STAGE == 1) ? p.glm_f : 1.0;
  const long long plane = p.plane;

  double Gl[NV], nxt[NV];
  double vloc = 0.0;

  auto row_addr = [&](int rn, unsigned &mask) -> long long {
    int rm;
    unsigned ymask;
    map_axis(rn, p.ny, p.bcy, p.sy_lo, p.sy_hi, rm, ymask);
    mask = xmask ^ ymask;
    return (long long)rm * p.pitch + cm;
  };

  unsigned mask_nxt;
  {
    const long long o = row_addr(j0 - 2, mask_nxt);
#pragma unroll
    for (int m = 0; m < NV; m++) nxt[m] = p.us[m * plane + o];
  }

  for (int k = 0; k < hc + 4; ++k) {
    const int rn = j0 - 2 + k;
    const int sa = (k + 1) % 3, sb = (k + 2) % 3, sc = k % 3;  // rows rn-2, rn-1, rn
    {
      double u[NV], W[NV];
#pragma unroll
      for (int m = 0; m < NV; m++) u[m] = flip(nxt[m], mask_nxt, m);
      u[8] = glm_us * u[8];  // glm_ss2 before the sweep (main.f90:87)
      if (k + 1 < hc + 4) {  // prefetch the next row
        const long long o = row_addr(rn + 1, mask_nxt);
#pragma unroll
        for (int m = 0; m < NV; m++) nxt[m] = p.us[m * plane + o];
      }
      cons_to_prim(u, W);  // u2v
      if (!(W[3] > 0.0) || !(u[4] > 0.0)) {
        int rm;
        unsigned dummy;
        map_axis(rn, p.ny, p.bcy, p.sy_lo, p.sy_hi, rm, dummy);
        record_error(p.err, (W[3] > 0.0) ? OMHD_ERR_DENSITY : OMHD_ERR_PRESSURE, cm + 2, rm + 2);
      }
#pragma unroll
      for (int m = 0; m < NV; m++) {
        lds.ring[sc][m][lane + 1] = W[m];
        if (lane == 0) lds.ring[sc][m][0] = W[m];
        if (lane == 63) lds.ring[sc][m][65] = W[m];
      }
    }
    __syncthreads();  // one wavefront per workgroup: orders the ring writes before the neighbour reads

    // x-sweep of row r = rn-2: tq = (F(i-1) - F(i)) + G(j-1)   (rk.f90:30-31 order)
    double tq[NV];
    if (k >= 4) {
      double Lx[NV], Rx[NV], F[NV];
#pragma unroll
      for (int m = 0; m < NV; m++) {
        double wRx;
        limit<LIM>(lds.ring[sa][m][lane], lds.ring[sa][m][lane + 1], lds.ring[sa][m][lane + 2], wRx, Lx[m]);
        Rx[m] = from_next(wRx);
      }
      if (xlo_face) {
#pragma unroll
        for (int m = 0; m < NV; m++) Lx[m] = flip(Rx[m], p.sx_lo, m);
      }
      if (xhi_face) {
#pragma unroll
        for (int m = 0; m < NV; m++) Rx[m] = flip(Lx[m], p.sx_hi, m);
      }
      hlld<0>(Lx, Rx, p.ch, p.ch2, F);
#pragma unroll
      for (int m = 0; m < NV; m++) tq[m] = from_prev(F[m]) - F[m] + Gl[m];
    }

    if (k >= 2) {
      // y-slopes of row rn-1 from rows rn-2, rn-1, rn
      double Lf[NV], Rf[NV];
#pragma unroll
      for (int m = 0; m < NV; m++) {
        double wLyB;
        limit<LIM>(lds.ring[sa][m][lane + 1], lds.ring[sb][m][lane + 1], lds.ring[sc][m][lane + 1], Rf[m], wLyB);
        Lf[m] = lds.wly[m][lane];  // left state of the interface above row rn-2
        lds.wly[m][lane] = wLyB;
      }
      if (k >= 3) {
        // y-interface between rows rn-2 and rn-1
        if (p.bcy == OMHD_BC_MIRROR) {
          if (rn - 1 == 0) {  // south wall: bc_for_G, 2D_KH/bc.f90:62-70
#pragma unroll
            for (int m = 0; m < NV; m++) Lf[m] = flip(Rf[m], p.sy_lo, m);
          }
          if (rn - 1 == p.ny) {  // north wall: 2D_KH/bc.f90:73-81
#pragma unroll
            for (int m = 0; m < NV; m++) Rf[m] = flip(Lf[m], p.sy_hi, m);
          }
        }
        double G[NV];
        hlld<1>(Lf, Rf, p.ch, p.ch2, G);

        if (k >= 4 && writer) {
          const int r = rn - 2;
          const long long o = (long long)r * p.pitch + c;
          double o9[NV];
          if constexpr (STAGE == 1) {
#pragma unroll
            for (int m = 0; m < NV; m++) {
              double uu = p.us[m * plane + o];
              if (m == 8) uu = p.glm_f * uu;
              o9[m] = uu + p.dtx * (tq[m] - G[m]);  // rk_tvd21 / rk_std21
            }
          } else {
#pragma unroll
            for (int m = 0; m < NV; m++) {
              double a0 = p.u0[m * plane + o];
              if (m == 8) a0 = p.glm_f * a0;
              if constexpr (TT == 0) o9[m] = 0.5 * (a0 + p.us[m * plane + o] + p.dtx * (tq[m] - G[m]));  // rk_tvd22
              else o9[m] = a0 + p.dtx * (tq[m] - G[m]);                                                   // rk_std22
            }
            o9[8] = p.glm_f * o9[8];  // glm_ss2 after the step (main.f90:185)
          }
          if constexpr (NT) {  // streaming store: the next launch reads it, this one does not
#pragma unroll
            for (int m = 0; m < NV; m++) __builtin_nontemporal_store(o9[m], &p.out[m * plane + o]);
          } else {
#pragma unroll
            for (int m = 0; m < NV; m++) p.out[m * plane + o] = o9[m];
          }
          if constexpr (STAGE == 2) {  // next step's set_dt on U^{n+1}
            double w2[NV];
            cons_to_prim(o9, w2);
            const double vt = fastest_speed<2>(o9, w2);
            vloc = (vt > vloc || !(vt == vt)) ? vt : vloc;
          }
        }
#pragma unroll
        for (int m = 0; m < NV; m++) Gl[m] = G[m];
      }
    }
  }
  if constexpr (STAGE == 2) {
    if (!(vloc >= 0.0)) record_error(p.err, OMHD_ERR_DT, 0, 0);
    const double vm = wave_max(vloc >= 0.0 ? vloc : 0.0);
    if (lane == 0) atomic_max_nonneg(p.vmax_bits, vm);
  }
}

// CFL speed of the resident state (set_dt.f90:21-45), used once after upload
__global__ void k_cfl2d(const double *__restrict__ s, long long plane, int pitch, int nx, int ny,
                        unsigned long long *vmax_bits, int *err) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double vt = 0.0;
  if (gid < (long long)nx * ny) {
    const int i = (int)(gid % nx), j = (int)(gid / nx);
    const long long o = (long long)j * pitch + i;
    double u[NV], w[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) u[m] = s[m * plane + o];
    cons_to_prim(u, w);
    if (!(w[3] > 0.0) || !(u[4] > 0.0)) record_error(err, (w[3] > 0.0) ? OMHD_ERR_DENSITY : OMHD_ERR_PRESSURE, i + 2, j + 2);
    vt = fastest_speed<2>(u, w);
    if (!(vt >= 0.0)) {
      record_error(err, OMHD_ERR_DT, i + 2, j + 2);
      vt = 0.0;
    }
  }
  vt = wave_max(vt);
  if ((threadIdx.x & 63) == 0) atomic_max_nonneg(vmax_bits, vt);
}

// Fortran array U(ix,jx,9) (+ V(ix,jx,4)) with ghosts <-> SoA interior
__global__ void k_to_fortran(const double *__restrict__ s, long long plane, int pitch, int nx, int ny, int bcx, int bcy,
                             unsigned sxl, unsigned sxh, unsigned syl, unsigned syh, double *__restrict__ U,
                             double *__restrict__ V) {
  const int ix = nx + 2, jx = ny + 2;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)ix * jx) return;
  const int i = (int)(gid % ix), j = (int)(gid / ix);
  int im, jm;
  unsigned mx_, my_;
  map_axis(i - 1, nx, bcx, sxl, sxh, im, mx_);
  map_axis(j - 1, ny, bcy, syl, syh, jm, my_);
  const unsigned mask = mx_ ^ my_;
  const long long o = (long long)jm * pitch + im;
  const long long n = (long long)ix * jx;
  double u[NV], w[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) u[m] = flip(s[m * plane + o], mask, m);
#pragma unroll
  for (int m = 0; m < NV; m++) U[m * n + gid] = u[m];
  if (V) {
    cons_to_prim(u, w);
#pragma unroll
    for (int m = 0; m < 4; m++) V[m * n + gid] = w[m];
  }
}

__global__ void k_from_fortran(const double *__restrict__ U, double *__restrict__ s, long long plane, int pitch, int nx,
                               int ny) {
  const int ix = nx + 2, jx = ny + 2;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)nx * ny) return;
  const int i = (int)(gid % nx), j = (int)(gid / nx);
  const long long n = (long long)ix * jx;
  const long long f = (long long)(j + 1) * ix + (i + 1);
#pragma unroll
  for (int m = 0; m < NV; m++) s[m * plane + (long long)j * pitch + i] = U[m * n + f];
}

__global__ void k_zero_red(unsigned long long *red) {
  if (threadIdx.x < 3) red[threadIdx.x] = 0ull;
}

typedef void (*sweep_fn)(const Sweep2D);

template <int LIM, int NT>
void pick(int stage, int tt, sweep_fn &f) {
  if (stage == 1) f = tt == 0 ? k_sweep2d<LIM, 1, 0, NT> : k_sweep2d<LIM, 1, 1, NT>;
  else f = tt == 0 ? k_sweep2d<LIM, 2, 0, NT> : k_sweep2d<LIM, 2, 1, NT>;
}

template <int NT>
sweep_fn sweep_kernel_nt(int lim, int stage, int tt) {
  sweep_fn f = nullptr;
  switch (lim) {
    case 0: pick<0, NT>(stage, tt, f); break;
    case 1: pick<1, NT>(stage, tt, f); break;
    case 2: pick<2, NT>(stage, tt, f); break;
    case 3: pick<3, NT>(stage, tt, f); break;
    default: pick<4, NT>(stage, tt, f); break;
  }
  return f;
}

sweep_fn sweep_kernel(int lim, int stage, int tt, int nt) {
  return nt ? sweep_kernel_nt<1>(lim, stage, tt) : sweep_kernel_nt<0>(lim, stage, tt);
}

unsigned sign_mask(const double *s9) {
  unsigned m = 0;
  for (int i = 0; i < NV; i++)
    if (s9[i] < 0) m |= 1u << i;
  return m;
}

}  // namespace
}  // namespace omhd

using namespace omhd;

struct omhd_handle {
  omhd_config cfg;
  int pitch = 0;
  long long plane = 0;
  double *U = nullptr;   // U^n / U^{n+1}
  double *U1 = nullptr;  // U*
  unsigned long long *red = nullptr;  // [0] vmax bits, [1..] error (3 ints)
  unsigned long long *red_host = nullptr;  // pinned mirror
  hipStream_t stream = nullptr;
  double vmax = -1.0;  // CFL speed of the current state
  bool have_state = false;
  unsigned sxl = 0, sxh = 0, syl = 0, syh = 0;
  int rows = 32;
  int nt_store = 0;  // non-temporal stores of the stage outputs
  // optional per-stage timing with HIP events on the handle's stream
  bool timing = false;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  double stage_ms[2] = {0.0, 0.0};
  long long stage_launches[2] = {0, 0};
};

static int read_red(omhd_handle *h, double &vmax, int err[3]) {
  OMHD_HIP(hipMemcpyAsync(h->red_host, h->red, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream));
  OMHD_HIP(hipStreamSynchronize(h->stream));
  memcpy(&vmax, &h->red_host[0], sizeof(double));
  memcpy(err, &h->red_host[1], 3 * sizeof(int));
  return OMHD_OK;
}

extern "C" int omhd_create(const omhd_config *cfg, omhd_handle **out) {
  if (!cfg || !out) return set_error(OMHD_ERR_ARG, "null argument");
  if (cfg->ndim != 2) return set_error(OMHD_ERR_ARG, "the resident step is built for ndim == 2", cfg->ndim);
  if (cfg->nx < 2 || cfg->ny < 2) return set_error(OMHD_ERR_EXTENT, "need nx, ny >= 2", cfg->nx, cfg->ny);
  if (cfg->lm_type < 0 || cfg->lm_type > 4) return set_error(OMHD_ERR_TYPE, "unknown limiter", cfg->lm_type);
  if (cfg->flux_type != 3) return set_error(OMHD_ERR_TYPE, "only flux_type 3 (HLLD) is built", cfg->flux_type);
  if (cfg->time_type != 0 && cfg->time_type != 1) return set_error(OMHD_ERR_TYPE, "unknown time_type", cfg->time_type);
  if (!(cfg->dx > 0) || !(cfg->cfl > 0)) return set_error(OMHD_ERR_ARG, "dx and cfl must be positive");
  for (int a = 0; a < 2; a++)
    if (cfg->bc_type[2 * a] != cfg->bc_type[2 * a + 1] &&
        (cfg->bc_type[2 * a] == OMHD_BC_PERIODIC || cfg->bc_type[2 * a + 1] == OMHD_BC_PERIODIC))
      return set_error(OMHD_ERR_ARG, "periodic faces come in pairs", a);
  omhd_handle *h = new omhd_handle();
  h->cfg = *cfg;
  h->pitch = (cfg->nx + 7) & ~7;  // 64-byte aligned rows
  h->plane = (long long)h->pitch * cfg->ny;
  h->plane = (h->plane + 63) & ~63LL;
  h->rows = cfg->rows_per_wave > 0 ? cfg->rows_per_wave : 24;
  h->nt_store = cfg->store_hint == 1;
  h->sxl = sign_mask(cfg->bc_sign + 0);
  h->sxh = sign_mask(cfg->bc_sign + 9);
  h->syl = sign_mask(cfg->bc_sign + 18);
  h->syh = sign_mask(cfg->bc_sign + 27);
  const size_t bytes = (size_t)h->plane * NV * sizeof(double);
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void **)&h->U, bytes);
  if (e == hipSuccess) e = hipMalloc((void **)&h->U1, bytes);
  if (e == hipSuccess) e = hipMalloc((void **)&h->red, 4 * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipHostMalloc((void **)&h->red_host, 4 * sizeof(unsigned long long), hipHostMallocDefault);
  if (e != hipSuccess) {
    omhd_destroy(h);
    return set_error(OMHD_ERR_HIP, hipGetErrorString(e));
  }
  *out = h;
  return ok();
}

extern "C" int omhd_destroy(omhd_handle *h) {
  if (!h) return ok();
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->U) (void)hipFree(h->U);
  if (h->U1) (void)hipFree(h->U1);
  if (h->red) (void)hipFree(h->red);
  if (h->red_host) (void)hipHostFree(h->red_host);
  for (int i = 0; i < 3; i++)
    if (h->ev[i]) (void)hipEventDestroy(h->ev[i]);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return ok();
}

extern "C" int omhd_state_layout(omhd_handle *h, double **dev_state, int64_t *plane_stride, int32_t *row_pitch) {
  if (!h) return set_error(OMHD_ERR_ARG, "null handle");
  if (dev_state) *dev_state = h->U;
  if (plane_stride) *plane_stride = h->plane;
  if (row_pitch) *row_pitch = h->pitch;
  return ok();
}

extern "C" void *omhd_stream(omhd_handle *h) { return h ? (void *)h->stream : nullptr; }

static int cfl_pass(omhd_handle *h) {
  const int nx = h->cfg.nx, ny = h->cfg.ny;
  k_zero_red<<<1, 64, 0, h->stream>>>(h->red);
  const long long n = (long long)nx * ny;
  k_cfl2d<<<(unsigned)((n + 255) / 256), 256, 0, h->stream>>>(h->U, h->plane, h->pitch, nx, ny, h->red,
                                                              (int *)(h->red + 1));
  OMHD_HIP(hipGetLastError());
  int err[3];
  double vm;
  if (int rc = read_red(h, vm, err)) return rc;
  if (err[0]) return set_error(err[0], "invalid state", err[1], err[2], 0);
  h->vmax = vm;
  return OMHD_OK;
}

extern "C" int omhd_upload(omhd_handle *h, const double *U) {
  if (!h || !U) return set_error(OMHD_ERR_ARG, "null argument");
  const int nx = h->cfg.nx, ny = h->cfg.ny;
  const size_t n = (size_t)(nx + 2) * (ny + 2);
  Staged a;
  if (int rc = a.init(U, 9 * n, true, false, h->stream)) return rc;
  k_from_fortran<<<(unsigned)(((size_t)nx * ny + 255) / 256), 256, 0, h->stream>>>(a.dev, h->U, h->plane, h->pitch,
                                                                                  nx, ny);
  OMHD_HIP(hipGetLastError());
  a.release(h->stream);
  if (int rc = cfl_pass(h)) return rc;
  h->have_state = true;
  return ok();
}

extern "C" int omhd_download(omhd_handle *h, double *U, double *V) {
  if (!h || !U) return set_error(OMHD_ERR_ARG, "null argument");
  const int nx = h->cfg.nx, ny = h->cfg.ny;
  const size_t n = (size_t)(nx + 2) * (ny + 2);
  Staged a, b;
  if (int rc = a.init(U, 9 * n, false, true, h->stream)) return rc;
  if (V)
    if (int rc = b.init(V, 4 * n, false, true, h->stream)) return rc;
  k_to_fortran<<<(unsigned)((n + 255) / 256), 256, 0, h->stream>>>(h->U, h->plane, h->pitch, nx, ny, h->cfg.bc_type[0],
                                                                   h->cfg.bc_type[2], h->sxl, h->sxh, h->syl, h->syh,
                                                                   a.dev, V ? b.dev : nullptr);
  OMHD_HIP(hipGetLastError());
  if (int rc = a.finish(h->stream)) return rc;
  if (V)
    if (int rc = b.finish(h->stream)) return rc;
  a.release(h->stream);
  b.release(h->stream);
  OMHD_HIP(hipStreamSynchronize(h->stream));
  return ok();
}

extern "C" int omhd_step(omhd_handle *h, int32_t nsteps, double *t, double *dt_trace, double *ch_trace, int32_t *done) {
  if (done) *done = 0;
  if (!h) return set_error(OMHD_ERR_ARG, "null handle");
  if (!h->have_state) return set_error(OMHD_ERR_ARG, "omhd_upload first");
  const omhd_config &c = h->cfg;
  Sweep2D p{};
  p.plane = h->plane;
  p.pitch = h->pitch;
  p.nx = c.nx;
  p.ny = c.ny;
  p.nstrip = (c.nx + STRIP - 1) / STRIP;
  p.rows = h->rows;
  const int nchunk = (c.ny + p.rows - 1) / p.rows;
  p.ntiles = p.nstrip * nchunk;
  p.bcx = c.bc_type[0];
  p.bcy = c.bc_type[2];
  p.sx_lo = h->sxl;
  p.sx_hi = h->sxh;
  p.sy_lo = h->syl;
  p.sy_hi = h->syh;
  p.vmax_bits = h->red;
  p.err = (int *)(h->red + 1);
  const unsigned grid = (unsigned)((p.ntiles + 7) / 8 * 8);
  sweep_fn s1 = sweep_kernel(c.lm_type, 1, c.time_type, h->nt_store);
  sweep_fn s2 = sweep_kernel(c.lm_type, 2, c.time_type, h->nt_store);
  const double cr = 0.2;  // glm_ss2.f90:18
  for (int n = 0; n < nsteps; n++) {
    const double ch = h->vmax;                // set_dt returns vmax as ch (main.f90:80)
    const double dt = c.cfl * c.dx / h->vmax;  // set_dt.f90:52
    if (!(dt >= 1.0e-7)) return set_error(OMHD_ERR_DT, "dt is too small", 0, 0, 0, dt);  // set_dt.f90:55-65
    p.ch = ch;
    p.ch2 = ch * ch;
    p.glm_f = exp(-0.5 * dt * ch / cr);  // glm_ss2.f90:27, host libm
    k_zero_red<<<1, 64, 0, h->stream>>>(h->red);
    if (h->timing) OMHD_HIP(hipEventRecord(h->ev[0], h->stream));
    // stage 1: U* = U + dtx dF   (rk_tvd21: dt/dx, rk_std21: 0.5*dt/dx)
    p.us = h->U;
    p.u0 = h->U;
    p.out = h->U1;
    p.dtx = (c.time_type == 0) ? dt / c.dx : 0.5 * dt / c.dx;
    hipLaunchKernelGGL(s1, dim3(grid), dim3(64), 0, h->stream, p);
    OMHD_HIP(hipGetLastError());
    if (h->timing) OMHD_HIP(hipEventRecord(h->ev[1], h->stream));
    // stage 2: U^{n+1} in place over U^n (each cell's U^n is read only by the
    // wave that overwrites it, just before it does)
    p.us = h->U1;
    p.u0 = h->U;
    p.out = h->U;
    p.dtx = dt / c.dx;
    hipLaunchKernelGGL(s2, dim3(grid), dim3(64), 0, h->stream, p);
    OMHD_HIP(hipGetLastError());
    if (h->timing) OMHD_HIP(hipEventRecord(h->ev[2], h->stream));
    int err[3];
    double vm;
    if (int rc = read_red(h, vm, err)) return rc;
    if (err[0]) return set_error(err[0], err[0] == OMHD_ERR_PRESSURE ? "negative pressure"
                                         : err[0] == OMHD_ERR_DENSITY ? "negative density"
                                                                      : "non-finite signal speed",
                                 err[1], err[2], 0);
    h->vmax = vm;
    if (h->timing) {  // read_red synchronised the stream: both events are complete
      float a = 0.f, b = 0.f;
      OMHD_HIP(hipEventElapsedTime(&a, h->ev[0], h->ev[1]));
      OMHD_HIP(hipEventElapsedTime(&b, h->ev[1], h->ev[2]));
      h->stage_ms[0] += a;
      h->stage_ms[1] += b;
      h->stage_launches[0]++;
      h->stage_launches[1]++;
    }
    if (t) *t += dt;
    if (dt_trace) dt_trace[n] = dt;
    if (ch_trace) ch_trace[n] = ch;
    if (done) *done = n + 1;
  }
  return ok();
}

extern "C" int omhd_set_timing(omhd_handle *h, int32_t enable) {
  if (!h) return set_error(OMHD_ERR_ARG, "null handle");
  if (enable && !h->ev[0])
    for (int i = 0; i < 3; i++) OMHD_HIP(hipEventCreate(&h->ev[i]));
  h->timing = enable != 0;
  h->stage_ms[0] = h->stage_ms[1] = 0.0;
  h->stage_launches[0] = h->stage_launches[1] = 0;
  return ok();
}

extern "C" int omhd_get_timing(omhd_handle *h, double *stage_ms, int64_t *launches) {
  if (!h || !stage_ms || !launches) return set_error(OMHD_ERR_ARG, "null argument");
  for (int i = 0; i < 2; i++) {
    stage_ms[i] = h->stage_ms[i];
    launches[i] = h->stage_launches[i];
  }
  return ok();
}
