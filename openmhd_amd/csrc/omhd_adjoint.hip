// omhd_adjoint.hip -- discrete adjoint of the resident 2D step.
//
// The forward map of one step with frozen (dt, ch) (2D_basic/main.f90:80-185):
//   Ug   = G(U)                       glm_ss2: psi *= f, f = exp(-0.5 dt ch / 0.2)
//   U*   = Ug + a1 D(Ug)              rk_tvd21 (a1 = dt/dx) | rk_std21 (0.5 dt/dx)
//   U**  = c0 Ug + c1 U* + a2 D(U*)   rk_tvd22 (0.5, 0.5, 0.5 dt/dx) | rk_std22 (1, 0, dt/dx)
//   U'   = G(U**)
// where D(W) = F(i-1) - F(i) + G(j-1) - G(j) is the flux difference from
// u2v -> limiter -> HLLD.  Its reverse, given lam' = dL/dU':
//   lam** = G lam'
//   lam*  = c1 lam** + (dD/dW at U*)^T (a2 lam**)
//   lamg  = c0 lam** + lam* + (dD/dW at Ug)^T (a1 lam*)
//   lam   = G lamg
// Each (dD/dW)^T mu is four launches over the grid:
//   k_adj_prim     primitives of W (u2v)
//   k_adj_face<x>  per x-face: recompute the two limited face states and the
//                  HLLD flux, run its reverse on bF = mu(c+1) - mu(c), store the
//                  adjoints of the face states by owning cell (AL at c, AR at c+1)
//   k_adj_gather<x> per cell: the limiter reverse gathered from the three cells
//                  whose slopes read it, into bw
//   k_adj_face<y>, k_adj_gather<y, finish>: the same along y, then the u2v
//                  reverse and the RK/GLM combine above, written in place.
// Branches (limiter min/sign, HLLD regions) are re-decided from the recomputed
// values, i.e. the adjoint follows the branch the forward pass took.
//
// Checkpointing: omhd_adjoint_forward keeps the state every `every` steps;
// omhd_adjoint_backward walks the segments in reverse, replays each one with
// the resident step kernels into a tape of (U^n, U*^n), then runs the reverse
// steps of that segment.
#include <math.h>

#include <algorithm>
#include <vector>

#include "omhd_adjoint.h"
#include "omhd_host.h"
#include "omhd_internal.h"

namespace omhd {
namespace {

using adj::NV;

struct Geo {
  long long plane, zstride;
  int pitch, nx, ny, nz;
};

__device__ __forceinline__ int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

__device__ __forceinline__ bool cell_of(const Geo &g, int &i, int &j, int &k) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx * g.ny * g.nz) return false;
  i = (int)(t % g.nx);
  const long long r = t / g.nx;
  j = (int)(r % g.ny);
  k = (int)(r / g.ny);
  return true;
}

// offset of the cell `s` steps along DIR from (i, j, k), periodic
template <int DIR>
__device__ __forceinline__ long long nb(const Geo &g, int i, int j, int k, int s) {
  if (DIR == 0) i = wrap(i + s, g.nx);
  if (DIR == 1) j = wrap(j + s, g.ny);
  if (DIR == 2) k = wrap(k + s, g.nz);
  return (long long)k * g.zstride + (long long)j * g.pitch + i;
}

__global__ __launch_bounds__(256) void k_adj_prim(const Geo g, const double *__restrict__ U, double psi_f,
                                                  double *__restrict__ w) {
  int i, j, k;
  if (!cell_of(g, i, j, k)) return;
  const long long o = (long long)k * g.zstride + (long long)j * g.pitch + i;
  double u[NV], p[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) u[m] = U[m * g.plane + o];
  u[8] = psi_f * u[8];
  adj::prim_fwd<double>(u, p);
#pragma unroll
  for (int m = 0; m < NV; m++) w[m * g.plane + o] = p[m];
}

// mu = mscale * M, slot 8 also times mpsi
template <int DIR, int LIM>
__global__ __launch_bounds__(128) void k_adj_face(const Geo g, const double *__restrict__ w,
                                                  const double *__restrict__ M, double mscale, double mpsi, double ch,
                                                  double *__restrict__ AL, double *__restrict__ AR) {
  int i, j, k;
  if (!cell_of(g, i, j, k)) return;
  const long long om = nb<DIR>(g, i, j, k, -1), o0 = nb<DIR>(g, i, j, k, 0), o1 = nb<DIR>(g, i, j, k, 1),
                  o2 = nb<DIR>(g, i, j, k, 2);
  double VL[NV], VR[NV], bF[NV], bL[NV], bR[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) {
    const double *a = w + m * g.plane;
    const double am = a[om], a0 = a[o0], a1 = a[o1], a2 = a[o2];
    double gR, gL, hR, hL;
    adj::lim_fwd<double>(LIM, am, a0, a1, gR, gL);
    adj::lim_fwd<double>(LIM, a0, a1, a2, hR, hL);
    VL[m] = gL;
    VR[m] = hR;
    const double s = m == 8 ? mscale * mpsi : mscale;
    bF[m] = s * (M[m * g.plane + o1] - M[m * g.plane + o0]);  // F enters D(c) with -1, D(c+1) with +1
    bL[m] = 0.0;
    bR[m] = 0.0;
  }
  adj::hlld_vjp<DIR>(VL, VR, ch, bF, bL, bR);
#pragma unroll
  for (int m = 0; m < NV; m++) {
    AL[m * g.plane + o0] = bL[m];
    AR[m * g.plane + o1] = bR[m];
  }
}

struct Combine {
  const double *A;  // lam'
  double cA, fA;    // coefficient, and the psi factor applied to A
  const double *B;  // lam* (or null)
  double cB;
  double fO;        // psi factor applied to the result
  double *out;
};

// MODE 0: bw = contribution; 1: bw += contribution; 2: finish (u2v reverse and
// the RK/GLM combine into c.out)
template <int DIR, int LIM, int MODE>
__global__ __launch_bounds__(256) void k_adj_gather(const Geo g, const double *__restrict__ w,
                                                    const double *__restrict__ AL, const double *__restrict__ AR,
                                                    double *__restrict__ bw, const Combine c) {
  int i, j, k;
  if (!cell_of(g, i, j, k)) return;
  const long long o0 = nb<DIR>(g, i, j, k, 0);
  long long o[5];
#pragma unroll
  for (int s = 0; s < 5; s++) o[s] = nb<DIR>(g, i, j, k, s - 2);
  double b[NV], wc[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) {
    const double *a = w + m * g.plane;
    double v[5];
#pragma unroll
    for (int s = 0; s < 5; s++) v[s] = a[o[s]];
    wc[m] = v[2];
    double acc = 0.0;
    // the slope at cell q = c + s (s = -1, 0, 1) reads this cell as its
    // (wp, w0, wm) neighbour: coefficient index 1 - s
#pragma unroll
    for (int s = -1; s <= 1; s++) {
      double cR[3], cL[3];
      adj::lim_coef(LIM, v[s + 1], v[s + 2], v[s + 3], cR, cL);
      const double l = AL[m * g.plane + o[s + 2]], r = AR[m * g.plane + o[s + 2]];
      if (s == 0) acc += l + r;  // gL = w + gradL, gR = w - gradR
      acc += cL[1 - s] * l - cR[1 - s] * r;
    }
    b[m] = acc;
  }
  if (MODE == 0) {
#pragma unroll
    for (int m = 0; m < NV; m++) bw[m * g.plane + o0] = b[m];
    return;
  }
#pragma unroll
  for (int m = 0; m < NV; m++) b[m] += bw[m * g.plane + o0];
  if (MODE == 1) {
#pragma unroll
    for (int m = 0; m < NV; m++) bw[m * g.plane + o0] = b[m];
    return;
  }
  double nu[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) nu[m] = 0.0;
  adj::prim_vjp(wc, b, nu);
#pragma unroll
  for (int m = 0; m < NV; m++) {
    double r = c.cA * ((m == 8 ? c.fA : 1.0) * c.A[m * g.plane + o0]) + nu[m];
    if (c.B) r += c.cB * c.B[m * g.plane + o0];
    c.out[m * g.plane + o0] = (m == 8 ? c.fO : 1.0) * r;
  }
}

__global__ void k_lam_copy(double *__restrict__ dst, const double *__restrict__ src, const Geo g, int to_compact) {
  int i, j, k;
  if (!cell_of(g, i, j, k)) return;
  const long long n = (long long)g.nx * g.ny * g.nz;
  const long long t = ((long long)k * g.ny + j) * g.nx + i;
  const long long o = (long long)k * g.zstride + (long long)j * g.pitch + i;
  for (int m = 0; m < NV; m++) {
    if (to_compact) dst[m * n + t] = src[m * g.plane + o];
    else dst[m * g.plane + o] = src[m * n + t];
  }
}

typedef void (*face_fn)(const Geo, const double *, const double *, double, double, double, double *, double *);
typedef void (*gather_fn)(const Geo, const double *, const double *, const double *, double *, const Combine);

template <int DIR>
face_fn face_kernel(int lim) {
  switch (lim) {
    case 0: return k_adj_face<DIR, 0>;
    case 1: return k_adj_face<DIR, 1>;
    case 2: return k_adj_face<DIR, 2>;
    case 3: return k_adj_face<DIR, 3>;
    default: return k_adj_face<DIR, 4>;
  }
}

template <int DIR, int MODE>
gather_fn gather_kernel(int lim) {
  switch (lim) {
    case 0: return k_adj_gather<DIR, 0, MODE>;
    case 1: return k_adj_gather<DIR, 1, MODE>;
    case 2: return k_adj_gather<DIR, 2, MODE>;
    case 3: return k_adj_gather<DIR, 3, MODE>;
    default: return k_adj_gather<DIR, 4, MODE>;
  }
}

}  // namespace
}  // namespace omhd

using namespace omhd;

struct omhd_adjoint {
  omhd_handle *h = nullptr;
  int nsteps = 0, every = 0, nseg = 0;
  std::vector<double> dt, ch;
  size_t state = 0;            // doubles per state buffer (9 planes)
  double *ckpt = nullptr;      // nseg states: U at steps 0, every, 2 every, ...
  double *tape = nullptr;      // 2 * every states: (U^n, U*^n) of one segment
  double *work = nullptr;      // w, AL, AR, bw, lam, lam*
  bool have_forward = false;
  double *w() { return work; }
  double *AL() { return work + state; }
  double *AR() { return work + 2 * state; }
  double *bw() { return work + 3 * state; }
  double *lam() { return work + 4 * state; }
  double *lams() { return work + 5 * state; }
};

static Geo geo_of(const omhd_handle *h) {
  Geo g;
  g.plane = h->plane;
  g.zstride = h->zstride;
  g.pitch = h->pitch;
  g.nx = h->cfg.nx;
  g.ny = h->cfg.ny;
  g.nz = h->cfg.nz;
  return g;
}

extern "C" int omhd_adjoint_create(omhd_handle *h, int32_t nsteps, int32_t every, omhd_adjoint **out) {
  if (!h || !out) return set_error(OMHD_ERR_ARG, "null argument");
  if (nsteps < 1 || every < 1) return set_error(OMHD_ERR_ARG, "need nsteps >= 1 and every >= 1", nsteps, every);
  if (h->nd != 2) return set_error(OMHD_ERR_ARG, "the adjoint is built for the 2D step");
  for (int a = 0; a < 2; a++)
    if (h->faces.bc[a] != OMHD_BC_PERIODIC)
      return set_error(OMHD_ERR_ARG, "the adjoint is built for periodic boundaries", a);
  if (every > nsteps) every = nsteps;
  omhd_adjoint *a = new omhd_adjoint();
  a->h = h;
  a->nsteps = nsteps;
  a->every = every;
  a->nseg = (nsteps + every - 1) / every;
  a->state = (size_t)h->plane * NV;
  a->dt.assign(nsteps, 0.0);
  a->ch.assign(nsteps, 0.0);
  hipError_t e = hipMalloc((void **)&a->ckpt, a->state * a->nseg * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&a->tape, a->state * 2 * every * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&a->work, a->state * 6 * sizeof(double));
  if (e != hipSuccess) {
    omhd_adjoint_destroy(a);
    return set_error(OMHD_ERR_HIP, hipGetErrorString(e));
  }
  *out = a;
  return ok();
}

extern "C" int omhd_adjoint_destroy(omhd_adjoint *a) {
  if (!a) return ok();
  if (a->h && a->h->stream) (void)hipStreamSynchronize(a->h->stream);
  if (a->ckpt) (void)hipFree(a->ckpt);
  if (a->tape) (void)hipFree(a->tape);
  if (a->work) (void)hipFree(a->work);
  delete a;
  return ok();
}

extern "C" int omhd_adjoint_forward(omhd_adjoint *a, const double *dt_in, const double *ch_in, double *dt_out,
                                    double *ch_out) {
  if (!a) return set_error(OMHD_ERR_ARG, "null adjoint");
  omhd_handle *h = a->h;
  if (!h->have_state) return set_error(OMHD_ERR_ARG, "omhd_upload first");
  if ((dt_in == nullptr) != (ch_in == nullptr)) return set_error(OMHD_ERR_ARG, "give both dt and ch, or neither");
  a->have_forward = false;
  for (int s = 0; s < a->nseg; s++) {
    const int n0 = s * a->every, cnt = std::min(a->every, a->nsteps - n0);
    OMHD_HIP(hipMemcpyAsync(a->ckpt + a->state * s, h->U, a->state * sizeof(double), hipMemcpyDeviceToDevice,
                            h->stream));
    if (dt_in) {
      if (int rc = omhd_step_fixed(h, cnt, dt_in + n0, ch_in + n0)) return rc;
      for (int n = n0; n < n0 + cnt; n++) {
        a->dt[n] = dt_in[n];
        a->ch[n] = ch_in[n];
      }
    } else {
      if (int rc = omhd_step(h, cnt, nullptr, a->dt.data() + n0, a->ch.data() + n0, nullptr)) return rc;
    }
  }
  if (dt_out) memcpy(dt_out, a->dt.data(), sizeof(double) * a->nsteps);
  if (ch_out) memcpy(ch_out, a->ch.data(), sizeof(double) * a->nsteps);
  a->have_forward = true;
  return ok();
}

// nu-combine of one (dD/dW)^T mu at state W (psi of W scaled by psi_f)
static int vjp_launch(omhd_adjoint *a, const double *W, double psi_f, const double *M, double mscale, double mpsi,
                      double ch, const Combine &fin) {
  omhd_handle *h = a->h;
  const Geo g = geo_of(h);
  const int lim = h->cfg.lm_type;
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  const unsigned g256 = (unsigned)((n + 255) / 256), g128 = (unsigned)((n + 127) / 128);
  k_adj_prim<<<g256, 256, 0, h->stream>>>(g, W, psi_f, a->w());
  OMHD_HIP(hipGetLastError());
  Combine none{};
  hipLaunchKernelGGL(face_kernel<0>(lim), dim3(g128), dim3(128), 0, h->stream, g, a->w(), M, mscale, mpsi, ch, a->AL(),
                     a->AR());
  OMHD_HIP(hipGetLastError());
  hipLaunchKernelGGL((gather_kernel<0, 0>(lim)), dim3(g256), dim3(256), 0, h->stream, g, a->w(), a->AL(), a->AR(),
                     a->bw(), none);
  OMHD_HIP(hipGetLastError());
  hipLaunchKernelGGL(face_kernel<1>(lim), dim3(g128), dim3(128), 0, h->stream, g, a->w(), M, mscale, mpsi, ch, a->AL(),
                     a->AR());
  OMHD_HIP(hipGetLastError());
  hipLaunchKernelGGL((gather_kernel<1, 2>(lim)), dim3(g256), dim3(256), 0, h->stream, g, a->w(), a->AL(), a->AR(),
                     a->bw(), fin);
  OMHD_HIP(hipGetLastError());
  return OMHD_OK;
}

extern "C" int omhd_adjoint_backward(omhd_adjoint *a, const double *lam_T, double *lam_0) {
  if (!a || !lam_T || !lam_0) return set_error(OMHD_ERR_ARG, "null argument");
  if (!a->have_forward) return set_error(OMHD_ERR_ARG, "omhd_adjoint_forward first");
  omhd_handle *h = a->h;
  const omhd_config &c = h->cfg;
  const Geo g = geo_of(h);
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  const unsigned g256 = (unsigned)((n + 255) / 256);
  Staged in, out;
  if (int rc = in.init(lam_T, NV * n, true, false, h->stream)) return rc;
  k_lam_copy<<<g256, 256, 0, h->stream>>>(a->lam(), in.dev, g, 0);
  OMHD_HIP(hipGetLastError());
  in.release(h->stream);
  OMHD_HIP(hipMemsetAsync(h->red, 0, 4 * sizeof(unsigned long long), h->stream));
  const bool tvd = c.time_type == 0;
  for (int s = a->nseg - 1; s >= 0; s--) {
    const int n0 = s * a->every, cnt = std::min(a->every, a->nsteps - n0);
    double *T = a->tape;
    // replay the segment: tape[2q] = U^{n0+q}, tape[2q+1] = U*^{n0+q}
    OMHD_HIP(hipMemcpyAsync(T, a->ckpt + a->state * s, a->state * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    for (int q = 0; q < cnt; q++) {
      double *Un = T + a->state * (2 * q), *Us = T + a->state * (2 * q + 1);
      if (int rc = stage_launch(h, 1, Un, Un, Us, a->dt[n0 + q], a->ch[n0 + q])) return rc;
      if (q + 1 < cnt)
        if (int rc = stage_launch(h, 2, Us, Un, T + a->state * (2 * q + 2), a->dt[n0 + q], a->ch[n0 + q])) return rc;
    }
    for (int q = cnt - 1; q >= 0; q--) {
      const double dt = a->dt[n0 + q], ch = a->ch[n0 + q];
      const double f = exp(-0.5 * dt * ch / 0.2);  // glm_ss2.f90:27
      const double dtx = dt / c.dx;
      const double c0 = tvd ? 0.5 : 1.0, c1 = tvd ? 0.5 : 0.0, a2 = tvd ? 0.5 * dtx : dtx, a1 = tvd ? dtx : 0.5 * dtx;
      const double *Un = T + a->state * (2 * q), *Us = T + a->state * (2 * q + 1);
      // lam* = c1 G lam' + (dD/dW at U*)^T (a2 G lam')
      Combine f1{a->lam(), c1, f, nullptr, 0.0, 1.0, a->lams()};
      if (int rc = vjp_launch(a, Us, 1.0, a->lam(), a2, f, ch, f1)) return rc;
      // lam = G (c0 G lam' + lam* + (dD/dW at Ug)^T (a1 lam*))
      Combine f2{a->lam(), c0, f, a->lams(), 1.0, f, a->lam()};
      if (int rc = vjp_launch(a, Un, f, a->lams(), a1, 1.0, ch, f2)) return rc;
    }
  }
  if (int rc = out.init(lam_0, NV * n, false, true, h->stream)) return rc;
  k_lam_copy<<<g256, 256, 0, h->stream>>>(out.dev, a->lam(), g, 1);
  OMHD_HIP(hipGetLastError());
  if (int rc = out.finish(h->stream)) return rc;
  out.release(h->stream);
  OMHD_HIP(hipMemcpyAsync(h->red_host, h->red, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream));
  OMHD_HIP(hipStreamSynchronize(h->stream));
  int err[4];
  memcpy(err, &h->red_host[1], sizeof(err));
  if (err[0]) return set_error(err[0], "invalid state while replaying the forward pass", err[1], err[2], err[3]);
  return ok();
}
