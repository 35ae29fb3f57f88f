// omhd_dropin.hip -- one HIP kernel per reference subroutine, on the
// reference's own Fortran array layout (column-major, one ghost layer), behind
// the drop-in C entry points of include/omhd.h.  These exist so that a
// reference-style driver can call u2v/limiter/flux_solver/... one by one and
// get the reference's results; the fused kernels in omhd_step.hip are the
// fast path.  One lane per cell (or interface).
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include <map>
#include <mutex>
#include <vector>

#include "omhd_device.h"
#include "omhd_host.h"

namespace omhd {

omhd_error_info &tls_error() {
  static thread_local omhd_error_info e{};
  return e;
}

// ---------------------------------------------------------------------------
// Device mirrors of host arrays.  A mirror holds one contiguous host address
// range [lo, hi) in device memory.  A drop-in call on a host array inside a
// mirror works on the mirror; one that reaches outside every mirror gets a
// new mirror covering its range and every mirror it overlaps, uploaded from
// the host where no mirror held the bytes and copied from the old mirrors
// where one did.  dev_newer: a call has written the mirror since the host
// copy was last brought up to date.  All copies are synchronous: every
// drop-in entry point drains its stream before returning, so no kernel is
// pending on a mirror when the next call looks at it.
// ---------------------------------------------------------------------------
namespace {
struct Mirror {
  uintptr_t lo, hi;
  char *dev;
  bool dev_newer;
};
std::mutex mirror_mu;
std::map<uintptr_t, Mirror> mirrors;  // by lo
int mirror_state = -1;                // -1: not decided yet (OMHD_MIRROR in the environment decides)

// mirrors overlapping [lo, hi), in address order
std::vector<std::map<uintptr_t, Mirror>::iterator> overlapping(uintptr_t lo, uintptr_t hi) {
  std::vector<std::map<uintptr_t, Mirror>::iterator> out;
  auto it = mirrors.upper_bound(lo);
  if (it != mirrors.begin()) {
    auto prev = std::prev(it);
    if (prev->second.hi > lo) out.push_back(prev);
  }
  for (; it != mirrors.end() && it->second.lo < hi; ++it) out.push_back(it);
  return out;
}

int mirror_release_all(bool write_back) {
  for (auto &kv : mirrors) {
    Mirror &m = kv.second;
    if (write_back && m.dev_newer)
      OMHD_HIP(hipMemcpy((void *)m.lo, m.dev, m.hi - m.lo, hipMemcpyDeviceToHost));
    OMHD_HIP(hipFree(m.dev));
  }
  mirrors.clear();
  return OMHD_OK;
}
}  // namespace

bool mirror_on() {
  std::lock_guard<std::mutex> lock(mirror_mu);
  if (mirror_state < 0) {
    const char *e = getenv("OMHD_MIRROR");
    mirror_state = (e && atoi(e) != 0) ? 1 : 0;
  }
  return mirror_state > 0;
}

int mirror_acquire(const double *p, size_t count, bool writes, double **dev) {
  std::lock_guard<std::mutex> lock(mirror_mu);
  const uintptr_t lo = (uintptr_t)p, hi = lo + count * sizeof(double);
  auto ov = overlapping(lo, hi);
  if (ov.size() == 1 && ov[0]->second.lo <= lo && hi <= ov[0]->second.hi) {
    Mirror &m = ov[0]->second;
    *dev = (double *)(m.dev + (lo - m.lo));
    m.dev_newer = m.dev_newer || writes;
    return OMHD_OK;
  }
  uintptr_t nlo = lo, nhi = hi;
  bool newer = writes;
  for (auto &it : ov) {
    nlo = std::min(nlo, it->second.lo);
    nhi = std::max(nhi, it->second.hi);
  }
  char *d = nullptr;
  OMHD_HIP(hipMalloc((void **)&d, nhi - nlo));
  // host bytes first, then every old mirror's bytes over them
  OMHD_HIP(hipMemcpy(d, (const void *)nlo, nhi - nlo, hipMemcpyHostToDevice));
  for (auto &it : ov) {
    Mirror &m = it->second;
    OMHD_HIP(hipMemcpy(d + (m.lo - nlo), m.dev, m.hi - m.lo, hipMemcpyDeviceToDevice));
    newer = newer || m.dev_newer;
    OMHD_HIP(hipFree(m.dev));
    mirrors.erase(it);
  }
  mirrors[nlo] = Mirror{nlo, nhi, d, newer};
  *dev = (double *)(d + (lo - nlo));
  return OMHD_OK;
}

int mirror_sync(const double *p, size_t count) {
  std::lock_guard<std::mutex> lock(mirror_mu);
  const uintptr_t lo = (uintptr_t)p, hi = lo + count * sizeof(double);
  for (auto &it : overlapping(lo, hi)) {
    Mirror &m = it->second;
    if (!m.dev_newer) continue;
    // the whole mirror: host bytes outside [lo, hi) but inside it are the
    // same arrays, and the mirror holds their newest values too
    OMHD_HIP(hipMemcpy((void *)m.lo, m.dev, m.hi - m.lo, hipMemcpyDeviceToHost));
    m.dev_newer = false;
  }
  return OMHD_OK;
}

hipStream_t dropin_stream() {
  static hipStream_t s = nullptr;
  static std::once_flag once;
  std::call_once(once, [] { (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking); });
  return s;
}

namespace {

constexpr int TPB = 256;

inline unsigned nblocks(size_t n) { return (unsigned)((n + TPB - 1) / TPB); }

// ---------------------------------------------------------------------------
// u2v over all cells (u2v.f90:20-31); flags |= 1 for p <= 0, |= 2 for rho <= 0
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_u2v(const double *__restrict__ U, double *__restrict__ V, size_t n, int *flags) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double u[NV], w[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) u[m] = U[m * n + c];
  cons_to_prim(u, w);
#pragma unroll
  for (int m = 0; m < 4; m++) V[m * n + c] = w[m];
  int f = 0;
  if (w[3] <= 0) f |= 1;
  if (u[4] <= 0) f |= 2;
  if (f) atomicOr(flags, f);
}

// ---------------------------------------------------------------------------
// limiter along one direction (limiter.f90).  Lines run along `dir` with
// stride s and length nl; (a, b) enumerate the transverse ranges.
// ---------------------------------------------------------------------------
template <int T>
__global__ __launch_bounds__(256) void k_limiter(const double *__restrict__ wk, double *__restrict__ wL, double *__restrict__ wR, int nl,
                          size_t s, int na, int nb, size_t sa, size_t sb, size_t base) {
  size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t total = (size_t)nl * na * nb;
  if (gid >= total) return;
  const int l = (int)(gid % nl);  // 0-based position on the line
  const size_t r = gid / nl;
  const int a = (int)(r % na), b = (int)(r / na);
  const size_t o = base + (size_t)a * sa + (size_t)b * sb;
  if constexpr (T == 0) {  // limiter.f90:56-70
    if (l < nl - 1) {
      wL[o + l * s] = wk[o + l * s];
      wR[o + l * s] = wk[o + (l + 1) * s];
    }
  } else {
    if (l == 0) {
      wL[o] = wk[o];
    } else if (l == nl - 1) {
      wR[o + (size_t)(nl - 2) * s] = wk[o + (size_t)(nl - 1) * s];
    } else {
      double gR, gL;
      limit<T>(wk[o + (l - 1) * s], wk[o + l * s], wk[o + (l + 1) * s], gR, gL);
      wR[o + (l - 1) * s] = gR;
      wL[o + l * s] = gL;
    }
  }
}

// ---------------------------------------------------------------------------
// HLLD flux at every interface of the reference's loop range
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(256) void k_flux(double *__restrict__ F, const double *__restrict__ VL, const double *__restrict__ VR, double ch,
                       int ix, int jx, int kx, int i0, int ni, int j0, int nj, int k0, int nk, int type) {
  size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (size_t)ni * nj * nk) return;
  const int i = i0 + (int)(gid % ni);
  const size_t r = gid / ni;
  const int j = j0 + (int)(r % nj), k = k0 + (int)(r / nj);
  const size_t n = (size_t)ix * jx * kx;
  const size_t c = (size_t)i + (size_t)ix * ((size_t)j + (size_t)jx * k);
  double L[NV], R[NV], f[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) {
    L[m] = VL[m * n + c];
    R[m] = VR[m * n + c];
  }
  if (type == 3) hlld<D>(L, R, ch, ch * ch, f);
  else riemann_other<D>(type, L, R, ch, ch * ch, f);
#pragma unroll
  for (int m = 0; m < NV; m++) F[m * n + c] = f[m];
}

__global__ void k_scale(double *__restrict__ a, double f, size_t n) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) a[c] = f * a[c];
}

// ---------------------------------------------------------------------------
// RK stages (rk.f90): interior cells, all 9 variables
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rk(int mode, double *__restrict__ wk, double *__restrict__ wk1, const double *__restrict__ wF,
                     const double *__restrict__ wG, const double *__restrict__ wH, double dtx, int ix, int jx, int kx) {
  const int ni = ix - 2, nj = jx - 2, nk = wH ? kx - 2 : 1;
  size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (size_t)ni * nj * nk) return;
  const int i = 1 + (int)(gid % ni);
  const size_t r = gid / ni;
  const int j = 1 + (int)(r % nj), k = (wH ? 1 : 0) + (int)(r / nj);
  const size_t n = (size_t)ix * jx * kx;
  const size_t c = (size_t)i + (size_t)ix * ((size_t)j + (size_t)jx * k);
  const size_t sj = ix, sk = (size_t)ix * jx;
  for (int m = 0; m < NV; m++) {
    const size_t q = m * n + c;
    double d = wF[q - 1] - wF[q] + wG[q - sj] - wG[q];
    if (wH) d = d + wH[q - sk] - wH[q];
    if (mode == 0 || mode == 2) wk1[q] = wk[q] + dtx * d;
    else if (mode == 1) wk[q] = 0.5 * (wk[q] + wk1[q] + dtx * d);
    else wk[q] = wk[q] + dtx * d;
  }
}

// ---------------------------------------------------------------------------
// set_dt: per-cell fastest speed, wave max, one atomic per wave
// ---------------------------------------------------------------------------
template <int ND>
__global__ __launch_bounds__(256) void k_speed(const double *__restrict__ U, const double *__restrict__ V, int ix, int jx, int kx,
                        unsigned long long *vmax_bits, int *flags) {
  const int is = ix > 1 ? 1 : 0, js = jx > 1 ? 1 : 0, ks = kx > 1 ? 1 : 0;
  const int ni = ix > 2 ? ix - 2 : 1, nj = jx > 2 ? jx - 2 : 1, nk = kx > 2 ? kx - 2 : 1;
  size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double vt = 0.0;
  if (gid < (size_t)ni * nj * nk) {
    const int i = is + (int)(gid % ni);
    const size_t r = gid / ni;
    const int j = js + (int)(r % nj), k = ks + (int)(r / nj);
    const size_t n = (size_t)ix * jx * kx;
    const size_t c = (size_t)i + (size_t)ix * ((size_t)j + (size_t)jx * k);
    double u[NV], w[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) u[m] = U[m * n + c];
#pragma unroll
    for (int m = 0; m < 4; m++) w[m] = V[m * n + c];
    vt = fastest_speed<ND>(u, w);
    if (!(vt >= 0.0)) atomicOr(flags, 4);  // NaN
  }
  vt = wave_max(vt);
  if ((threadIdx.x & 63) == 0) atomic_max_nonneg(vmax_bits, vt);
}

// ---------------------------------------------------------------------------
// set_dt without allocation (omhd_set_dt_noalloc).  The scratch array holds
// one partial per workgroup: pass 1 has each workgroup take
// NOALLOC_CELLS_PER_GROUP interior cells and store its maximum speed and its
// flags in its own slots; pass 2, one workgroup, reduces the slots of the
// workgroups pass 1 ran.  Every slot a call reads is written earlier in the
// same call, so nothing needs clearing, and no atomics are used.
// ---------------------------------------------------------------------------
constexpr int NOALLOC_GROUPS = 32768;
constexpr int NOALLOC_CELLS_PER_THREAD = 8;
constexpr size_t NOALLOC_CELLS_PER_GROUP = (size_t)TPB * NOALLOC_CELLS_PER_THREAD;
constexpr size_t NOALLOC_MAX_CELLS = (size_t)NOALLOC_GROUPS * NOALLOC_CELLS_PER_GROUP;  // 67,108,864
// [0, G): max-speed bits per workgroup; [G, 2G): flags per workgroup;
// [2G], [2G+1]: the reduced max-speed bits and flags.  512 KiB + 16 B per GPU.
__device__ __align__(256) unsigned long long scratch[2 * NOALLOC_GROUPS + 2];

// workgroup-wide max of non-negative doubles and OR of flags (TPB threads)
__device__ __forceinline__ void group_reduce(double &v, int &f) {
  __shared__ double wv[TPB / 64];
  __shared__ int wf[TPB / 64];
  v = wave_max(v);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) f |= __shfl_xor(f, off, 64);
  if ((threadIdx.x & 63) == 0) {
    wv[threadIdx.x >> 6] = v;
    wf[threadIdx.x >> 6] = f;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < TPB / 64; w++) {
      v = dmax(v, wv[w]);
      f |= wf[w];
    }
}

template <int ND>
__global__ __launch_bounds__(256) void k_speed_part(const double *__restrict__ U, const double *__restrict__ V, int ix,
                                                    int jx, int kx, unsigned long long *part) {
  const int is = ix > 1 ? 1 : 0, js = jx > 1 ? 1 : 0, ks = kx > 1 ? 1 : 0;
  const int ni = ix > 2 ? ix - 2 : 1, nj = jx > 2 ? jx - 2 : 1, nk = kx > 2 ? kx - 2 : 1;
  const size_t total = (size_t)ni * nj * nk, n = (size_t)ix * jx * kx;
  double vt = 0.0;
  int f = 0;
  for (int q = 0; q < NOALLOC_CELLS_PER_THREAD; q++) {
    const size_t gid = (size_t)blockIdx.x * NOALLOC_CELLS_PER_GROUP + (size_t)q * TPB + threadIdx.x;
    if (gid >= total) break;
    const int i = is + (int)(gid % ni);
    const size_t r = gid / ni;
    const int j = js + (int)(r % nj), k = ks + (int)(r / nj);
    const size_t c = (size_t)i + (size_t)ix * ((size_t)j + (size_t)jx * k);
    double u[NV], w[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) u[m] = U[m * n + c];
#pragma unroll
    for (int m = 0; m < 4; m++) w[m] = V[m * n + c];
    const double c1 = fastest_speed<ND>(u, w);
    if (!(c1 >= 0.0)) f |= 4;  // NaN
    else vt = dmax(vt, c1);
  }
  group_reduce(vt, f);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = (unsigned long long)__double_as_longlong(vt);
    part[NOALLOC_GROUPS + blockIdx.x] = (unsigned long long)f;
  }
}

__global__ __launch_bounds__(256) void k_speed_final(const unsigned long long *__restrict__ part, int groups,
                                                     unsigned long long *out) {
  double vt = 0.0;
  int f = 0;
  for (int g = threadIdx.x; g < groups; g += TPB) {
    vt = dmax(vt, __longlong_as_double((long long)part[g]));
    f |= (int)part[NOALLOC_GROUPS + g];
  }
  group_reduce(vt, f);
  if (threadIdx.x == 0) {
    out[0] = (unsigned long long)__double_as_longlong(vt);
    out[1] = (unsigned long long)f;
  }
}

// ---------------------------------------------------------------------------
// ghost fill, one face pair per launch (x, then y, then z: bc.f90 order)
// ---------------------------------------------------------------------------
__global__ void k_bc_face_pair(double *__restrict__ U, int ix, int jx, int kx, int dir, int type_lo, int type_hi,
                               const double sl0, const double sl1, const double sl2, const double sl3, const double sl4,
                               const double sl5, const double sl6, const double sl7, const double sl8, const double sh0,
                               const double sh1, const double sh2, const double sh3, const double sh4, const double sh5,
                               const double sh6, const double sh7, const double sh8) {
  // transverse plane enumeration
  const int na = (dir == 1) ? jx : ix, nb = (dir == 3) ? jx : kx;
  size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (size_t)na * nb) return;
  const int a = (int)(gid % na), b = (int)(gid / na);
  const size_t n = (size_t)ix * jx * kx;
  size_t base, s;
  int nl;
  if (dir == 1) { base = (size_t)ix * ((size_t)a + (size_t)jx * b); s = 1; nl = ix; }
  else if (dir == 2) { base = (size_t)a + (size_t)ix * (size_t)jx * b; s = ix; nl = jx; }
  else { base = (size_t)a + (size_t)ix * b; s = (size_t)ix * jx; nl = kx; }
  const double sl[NV] = {sl0, sl1, sl2, sl3, sl4, sl5, sl6, sl7, sl8};
  const double sh[NV] = {sh0, sh1, sh2, sh3, sh4, sh5, sh6, sh7, sh8};
  for (int m = 0; m < NV; m++) {
    double *A = U + m * n + base;
    if (type_lo == OMHD_BC_PERIODIC) {  // bc.f90:9-14 (high side first)
      A[(size_t)(nl - 1) * s] = A[s];
      A[0] = A[(size_t)(nl - 2) * s];
    } else {
      A[0] = sl[m] * A[s];
      A[(size_t)(nl - 1) * s] = sh[m] * A[(size_t)(nl - 2) * s];
    }
  }
  (void)type_hi;
}

__global__ void k_face_patch(double *__restrict__ VL, double *__restrict__ VR, int ix, int jx, int kx, int dir,
                             int type, const double sl0, const double sl1, const double sl2, const double sl3,
                             const double sl4, const double sl5, const double sl6, const double sl7, const double sl8,
                             const double sh0, const double sh1, const double sh2, const double sh3, const double sh4,
                             const double sh5, const double sh6, const double sh7, const double sh8) {
  const int na = (dir == 1) ? jx : ix, nb = (dir == 3) ? jx : kx;
  size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (size_t)na * nb) return;
  const int a = (int)(gid % na), b = (int)(gid / na);
  const size_t n = (size_t)ix * jx * kx;
  size_t base, s;
  int nl;
  if (dir == 1) { base = (size_t)ix * ((size_t)a + (size_t)jx * b); s = 1; nl = ix; }
  else if (dir == 2) { base = (size_t)a + (size_t)ix * (size_t)jx * b; s = ix; nl = jx; }
  else { base = (size_t)a + (size_t)ix * b; s = (size_t)ix * jx; nl = kx; }
  const double sl[NV] = {sl0, sl1, sl2, sl3, sl4, sl5, sl6, sl7, sl8};
  const double sh[NV] = {sh0, sh1, sh2, sh3, sh4, sh5, sh6, sh7, sh8};
  const size_t lo = base, hi = base + (size_t)(nl - 2) * s;
  for (int m = 0; m < NV; m++) {
    double *L = VL + m * n, *R = VR + m * n;
    if (type == OMHD_BC_PERIODIC) {  // main.f90:101-102
      R[hi] = R[lo];
      L[lo] = L[hi];
    } else {  // 2D_KH/bc.f90:61-81
      L[lo] = sl[m] * R[lo];
      R[hi] = sh[m] * L[hi];
    }
  }
}

// 2D_reconnection/flux_resistive.f90:66-111, the 2D CPU form of the resistive
// flux: face currents from cell-centred B (a 2-point normal difference and a
// 4-point transverse average, Toth et al. 2008), with eta/dx at the face
// folded into them before they enter F.  One lane per face over the faces
// (i0..i0+ni-1, j0..j0+nj-1), 0-based; Fe points at F's plane en, so bx, by
// and bz are planes 2, 3 and 4 from there; B points at U's plane bx.
template <int D>
__global__ __launch_bounds__(256) void k_flux_resistive2d(double *__restrict__ Fe, const double *__restrict__ B,
                                                          const double *__restrict__ eta_dx, int ix, int i0, int j0,
                                                          int ni, int nj, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)ni * nj) return;
  const size_t c = (size_t)(i0 + (int)(t % ni)) + (size_t)ix * (size_t)(j0 + (int)(t / ni));
  const double *Bx = B, *By = B + n, *Bz = B + 2 * n;
  const double e = eta_dx[c];
  const size_t cn = c + ix;  // (i, j+1)
  if (D == 1) {              // x-face (i+1/2, j): flux_resistive.f90:74-84
    const size_t cs = c - ix;
    const double JyS = -(e * (Bz[c + 1] - Bz[c]));
    const double JzS = e * ((By[c + 1] - By[c]) - 0.25 * (((Bx[cn] + Bx[cn + 1]) - Bx[cs]) - Bx[cs + 1]));
    Fe[c] = Fe[c] + 0.5 * (JyS * (Bz[c] + Bz[c + 1]) - JzS * (By[c] + By[c + 1]));
    Fe[3 * n + c] = Fe[3 * n + c] - JzS;
    Fe[4 * n + c] = Fe[4 * n + c] + JyS;
  } else {  // y-face (i, j+1/2): :97-107
    const double JxS = e * (Bz[cn] - Bz[c]);
    const double JzS = e * (0.25 * (((By[cn + 1] + By[c + 1]) - By[cn - 1]) - By[c - 1]) - (Bx[cn] - Bx[c]));
    Fe[c] = Fe[c] + 0.5 * (JzS * (Bx[c] + Bx[cn]) - JxS * (Bz[c] + Bz[cn]));
    Fe[2 * n + c] = Fe[2 * n + c] + JzS;
    Fe[4 * n + c] = Fe[4 * n + c] - JxS;
  }
}

// 3D_reconnection_gpu/flux_resistive.cuf:86-135, the 3D form: face currents
// from cell-centred B (2-point normal difference, 4-point transverse
// averages), eta/dx formed per face as eta / dx (:83), then added to F.  One
// lane per face over (i0.., j0.., k0..) with extents (ni, nj, nk), 0-based;
// Fe points at F's plane en (bx, by, bz are planes 2, 3, 4 from there), B at
// U's plane bx; eta holds eta at the faces, (ix,jx) when it does not vary in
// z (eta_nz == 1) or (ix,jx,kx).
template <int D>
__global__ __launch_bounds__(256) void k_flux_resistive3d(double *__restrict__ Fe, const double *__restrict__ B,
                                                          const double *__restrict__ eta, int eta_nz, double dx,
                                                          int ix, int jx, int i0, int j0, int k0, int ni, int nj,
                                                          int nk, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)ni * nj * nk) return;
  const int i = i0 + (int)(t % ni);
  const size_t r = t / ni;
  const int j = j0 + (int)(r % nj), k = k0 + (int)(r / nj);
  const size_t jy = (size_t)ix, kz = (size_t)ix * jx;
  const size_t c = (size_t)i + jy * j + kz * k;
  const double *Bx = B, *By = B + n, *Bz = B + 2 * n;
  const double eta_dx = eta[(size_t)i + jy * j + (eta_nz > 1 ? kz * k : 0)] / dx;
  if (D == 1) {  // x-face (i+1/2, j, k): :89-101
    const double JyS = -(Bz[c + 1] - Bz[c]) + 0.25 * (((Bx[c + kz] + Bx[c + 1 + kz]) - Bx[c - kz]) - Bx[c + 1 - kz]);
    const double JzS = (By[c + 1] - By[c]) - 0.25 * (((Bx[c + jy] + Bx[c + 1 + jy]) - Bx[c - jy]) - Bx[c + 1 - jy]);
    Fe[c] = Fe[c] + 0.5 * eta_dx * (JyS * (Bz[c] + Bz[c + 1]) - JzS * (By[c] + By[c + 1]));
    Fe[3 * n + c] = Fe[3 * n + c] - eta_dx * JzS;
    Fe[4 * n + c] = Fe[4 * n + c] + eta_dx * JyS;
  } else if (D == 2) {  // y-face (i, j+1/2, k): :106-118
    const double JxS = (Bz[c + jy] - Bz[c]) - 0.25 * (((By[c + kz] + By[c + jy + kz]) - By[c - kz]) - By[c + jy - kz]);
    const double JzS = -(Bx[c + jy] - Bx[c]) + 0.25 * (((By[c + 1] + By[c + 1 + jy]) - By[c - 1]) - By[c - 1 + jy]);
    Fe[c] = Fe[c] + 0.5 * eta_dx * (JzS * (Bx[c] + Bx[c + jy]) - JxS * (Bz[c] + Bz[c + jy]));
    Fe[2 * n + c] = Fe[2 * n + c] + eta_dx * JzS;
    Fe[4 * n + c] = Fe[4 * n + c] - eta_dx * JxS;
  } else {  // z-face (i, j, k+1/2): :123-135
    const double JyS = (Bx[c + kz] - Bx[c]) - 0.25 * (((Bz[c + 1] + Bz[c + 1 + kz]) - Bz[c - 1]) - Bz[c - 1 + kz]);
    const double JxS = -(By[c + kz] - By[c]) + 0.25 * (((Bz[c + jy] + Bz[c + jy + kz]) - Bz[c - jy]) - Bz[c - jy + kz]);
    Fe[c] = Fe[c] + 0.5 * eta_dx * (JxS * (By[c] + By[c + kz]) - JyS * (Bx[c] + Bx[c + kz]));
    Fe[2 * n + c] = Fe[2 * n + c] - eta_dx * JyS;
    Fe[3 * n + c] = Fe[3 * n + c] + eta_dx * JxS;
  }
}

int check_ext(int ix, int jx, int kx) {
  if (ix < 1 || jx < 1 || kx < 1) return set_error(OMHD_ERR_EXTENT, "non-positive extent", ix, jx, kx);
  return OMHD_OK;
}

// 1-based (i,j,k) of a 0-based linear cell index
void unflatten(size_t c, int ix, int jx, int &i, int &j, int &k) {
  i = (int)(c % ix) + 1;
  size_t r = c / ix;
  j = (int)(r % jx) + 1;
  k = (int)(r / jx) + 1;
}

// minloc of plane[0..n) held by the caller (host or device)
size_t plane_minloc(const double *plane, size_t n, bool dev, double &val) {
  if (!dev && mirror_on() && mirror_sync(plane, n) != OMHD_OK) {
    val = 0.0;
    return 0;  // the location is diagnostic only; the caller still reports the error code
  }
  const double *h = plane;
  double *tmp = nullptr;
  if (dev) {
    tmp = (double *)malloc(n * sizeof(double));
    if (!tmp || hipMemcpy(tmp, plane, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
      free(tmp);
      val = 0.0;
      return 0;  // the location is diagnostic only; the caller still reports the error code
    }
    h = tmp;
  }
  size_t best = 0;
  for (size_t c = 1; c < n; c++)
    if (h[c] < h[best]) best = c;
  val = h[best];
  free(tmp);
  return best;
}

}  // namespace
}  // namespace omhd

using namespace omhd;

extern "C" int omhd_last_error(omhd_error_info *out) {
  if (!out) return OMHD_ERR_ARG;
  *out = tls_error();
  return OMHD_OK;
}

extern "C" int omhd_mirror_mode(int32_t on) {
  {
    std::lock_guard<std::mutex> lock(mirror_mu);
    if (!on) {
      if (int rc = mirror_release_all(true)) return rc;
    }
    mirror_state = on ? 1 : 0;
  }
  return ok();
}

extern "C" int omhd_mirror_sync(const double *host, int64_t count) {
  if (!host || count < 0) return set_error(OMHD_ERR_ARG, "null array or negative count");
  if (!mirror_on()) return ok();
  if (int rc = mirror_sync(host, (size_t)count)) return rc;
  return ok();
}

extern "C" int omhd_mirror_update(const double *host, int64_t count) {
  if (!host || count < 0) return set_error(OMHD_ERR_ARG, "null array or negative count");
  if (!mirror_on()) return ok();
  std::lock_guard<std::mutex> lock(mirror_mu);
  const uintptr_t lo = (uintptr_t)host, hi = lo + (size_t)count * sizeof(double);
  for (auto &it : overlapping(lo, hi)) {  // upload the part of each mirror that the host rewrote
    Mirror &m = it->second;
    const uintptr_t a = std::max(lo, m.lo), b = std::min(hi, m.hi);
    OMHD_HIP(hipMemcpy(m.dev + (a - m.lo), (const void *)a, b - a, hipMemcpyHostToDevice));
  }
  return ok();
}

extern "C" const char *omhd_version(void) { return "omhd 0.1 (gfx950, fp64, fused column-march sweeps)"; }

extern "C" int omhd_u2v(const double *U, double *V, int32_t ix, int32_t jx, int32_t kx) {
  if (!U || !V) return set_error(OMHD_ERR_ARG, "null array");
  if (int rc = check_ext(ix, jx, kx)) return rc;
  hipStream_t s = dropin_stream();
  const size_t n = (size_t)ix * jx * kx;
  Staged dU, dV;
  int *flags;
  if (int rc = dU.init(U, 9 * n, true, false, s)) return rc;
  if (int rc = dV.init(V, 4 * n, false, true, s)) return rc;
  OMHD_HIP(hipMallocAsync((void **)&flags, sizeof(int), s));
  OMHD_HIP(hipMemsetAsync(flags, 0, sizeof(int), s));
  k_u2v<<<nblocks(n), TPB, 0, s>>>(dU.dev, dV.dev, n, flags);
  OMHD_HIP(hipGetLastError());
  int hflags = 0;
  OMHD_HIP(hipMemcpyAsync(&hflags, flags, sizeof(int), hipMemcpyDeviceToHost, s));
  if (int rc = dV.finish(s)) return rc;
  OMHD_HIP(hipStreamSynchronize(s));
  (void)hipFreeAsync(flags, s);
  dU.release(s);
  dV.release(s);
  OMHD_HIP(hipStreamSynchronize(s));
  if (hflags & 1) {  // u2v.f90:34-38: report minloc of p
    double val;
    size_t c = plane_minloc(V + 3 * n, n, is_device_ptr(V), val);
    int i, j, k;
    unflatten(c, ix, jx, i, j, k);
    return set_error(OMHD_ERR_PRESSURE, "negative pressure", i, j, k, val);
  }
  if (hflags & 2) {  // :39-43
    double val;
    size_t c = plane_minloc(U + 4 * n, n, is_device_ptr(U), val);
    int i, j, k;
    unflatten(c, ix, jx, i, j, k);
    return set_error(OMHD_ERR_DENSITY, "negative density", i, j, k, val);
  }
  return ok();
}

extern "C" int omhd_limiter(const double *wk, double *wL, double *wR, int32_t ix, int32_t jx, int32_t kx, int32_t dir,
                            int32_t type) {
  if (!wk || !wL || !wR) return set_error(OMHD_ERR_ARG, "null array");
  if (int rc = check_ext(ix, jx, kx)) return rc;
  if (type < 0 || type > 4) return set_error(OMHD_ERR_TYPE, "unknown limiter", type);
  if (dir < 1 || dir > 3) return set_error(OMHD_ERR_DIR, "bad dir", dir);
  hipStream_t s = dropin_stream();
  const size_t n = (size_t)ix * jx * kx;
  Staged a, bl, br;
  if (int rc = a.init(wk, n, true, false, s)) return rc;
  // outputs are only partly written (limiter.f90 skips the transverse ghosts):
  // stage them in so untouched entries keep the caller's values
  if (int rc = bl.init(wL, n, true, true, s)) return rc;
  if (int rc = br.init(wR, n, true, true, s)) return rc;
  auto lo = [](int e) { return e < 2 ? e : 2; };
  auto hi = [](int e) { return e - 1 > 1 ? e - 1 : 1; };
  int nl, na, nb;
  size_t st, sa, sb, base;
  const size_t sj = ix, sk = (size_t)ix * jx;
  if (dir == 1) {
    nl = ix; st = 1;
    na = hi(jx) - lo(jx) + 1; sa = sj;
    nb = hi(kx) - lo(kx) + 1; sb = sk;
    base = (size_t)(lo(jx) - 1) * sj + (size_t)(lo(kx) - 1) * sk;
  } else if (dir == 2) {
    nl = jx; st = sj;
    na = hi(ix) - lo(ix) + 1; sa = 1;
    nb = hi(kx) - lo(kx) + 1; sb = sk;
    base = (size_t)(lo(ix) - 1) + (size_t)(lo(kx) - 1) * sk;
  } else {
    nl = kx; st = sk;
    na = hi(ix) - lo(ix) + 1; sa = 1;
    nb = hi(jx) - lo(jx) + 1; sb = sj;
    base = (size_t)(lo(ix) - 1) + (size_t)(lo(jx) - 1) * sj;
  }
  const size_t total = (size_t)nl * na * nb;
  if (nl >= 2 && total > 0) {
    switch (type) {
      case 0: k_limiter<0><<<nblocks(total), TPB, 0, s>>>(a.dev, bl.dev, br.dev, nl, st, na, nb, sa, sb, base); break;
      case 1: k_limiter<1><<<nblocks(total), TPB, 0, s>>>(a.dev, bl.dev, br.dev, nl, st, na, nb, sa, sb, base); break;
      case 2: k_limiter<2><<<nblocks(total), TPB, 0, s>>>(a.dev, bl.dev, br.dev, nl, st, na, nb, sa, sb, base); break;
      case 3: k_limiter<3><<<nblocks(total), TPB, 0, s>>>(a.dev, bl.dev, br.dev, nl, st, na, nb, sa, sb, base); break;
      default: k_limiter<4><<<nblocks(total), TPB, 0, s>>>(a.dev, bl.dev, br.dev, nl, st, na, nb, sa, sb, base); break;
    }
    OMHD_HIP(hipGetLastError());
  }
  if (int rc = bl.finish(s)) return rc;
  if (int rc = br.finish(s)) return rc;
  a.release(s);
  bl.release(s);
  br.release(s);
  OMHD_HIP(hipStreamSynchronize(s));
  return ok();
}

extern "C" int omhd_flux_solver(double *F, const double *VL, const double *VR, double ch, int32_t ix, int32_t jx,
                                int32_t kx, int32_t dir, int32_t type) {
  if (!F || !VL || !VR) return set_error(OMHD_ERR_ARG, "null array");
  if (int rc = check_ext(ix, jx, kx)) return rc;
  if (type < 0 || type > 4) return set_error(OMHD_ERR_TYPE, "unknown flux_type", type);
  if (dir < 1 || dir > 3) return set_error(OMHD_ERR_DIR, "bad dir", dir);
  hipStream_t s = dropin_stream();
  const size_t n = (size_t)ix * jx * kx;
  Staged f, l, r;
  if (int rc = f.init(F, 9 * n, false, true, s)) return rc;
  if (int rc = l.init(VL, 9 * n, true, false, s)) return rc;
  if (int rc = r.init(VR, 9 * n, true, false, s)) return rc;
  OMHD_HIP(hipMemsetAsync(f.dev, 0, 9 * n * sizeof(double), s));  // flux_solver.f90:57
  auto lo = [](int e) { return e < 2 ? e : 2; };
  auto hi = [](int e) { return e - 1 > 1 ? e - 1 : 1; };
  int i0 = lo(ix) - 1, ni = hi(ix) - lo(ix) + 1, j0 = lo(jx) - 1, nj = hi(jx) - lo(jx) + 1, k0 = lo(kx) - 1,
      nk = hi(kx) - lo(kx) + 1;
  if (dir == 1) { i0 = 0; ni = ix - 1; }
  if (dir == 2) { j0 = 0; nj = jx - 1; }
  if (dir == 3) { k0 = 0; nk = kx - 1; }
  const size_t total = (size_t)(ni > 0 ? ni : 0) * nj * nk;
  if (total > 0) {
    if (dir == 1) k_flux<0><<<nblocks(total), TPB, 0, s>>>(f.dev, l.dev, r.dev, ch, ix, jx, kx, i0, ni, j0, nj, k0, nk, type);
    if (dir == 2) k_flux<1><<<nblocks(total), TPB, 0, s>>>(f.dev, l.dev, r.dev, ch, ix, jx, kx, i0, ni, j0, nj, k0, nk, type);
    if (dir == 3) k_flux<2><<<nblocks(total), TPB, 0, s>>>(f.dev, l.dev, r.dev, ch, ix, jx, kx, i0, ni, j0, nj, k0, nk, type);
    OMHD_HIP(hipGetLastError());
  }
  if (int rc = f.finish(s)) return rc;
  f.release(s);
  l.release(s);
  r.release(s);
  OMHD_HIP(hipStreamSynchronize(s));
  return ok();
}

extern "C" int omhd_glm_ss2(double *U, double ch, double dt, int32_t ix, int32_t jx, int32_t kx) {
  if (!U) return set_error(OMHD_ERR_ARG, "null array");
  if (int rc = check_ext(ix, jx, kx)) return rc;
  const double cr = 0.2;                    // glm_ss2.f90:18
  const double f1 = exp(-0.5 * dt * ch / cr);  // :27, host libm
  hipStream_t s = dropin_stream();
  const size_t n = (size_t)ix * jx * kx;
  Staged a;
  if (int rc = a.init(U, 9 * n, true, true, s)) return rc;
  k_scale<<<nblocks(n), TPB, 0, s>>>(a.dev + 8 * n, f1, n);
  OMHD_HIP(hipGetLastError());
  if (int rc = a.finish(s)) return rc;
  a.release(s);
  OMHD_HIP(hipStreamSynchronize(s));
  return ok();
}

extern "C" int omhd_rk(int32_t mode, double *wk, double *wk1, const double *wF, const double *wG, const double *wH,
                       double dt, double dx, int32_t ix, int32_t jx, int32_t kx) {
  if (mode < 0 || mode > 3) return set_error(OMHD_ERR_TYPE, "rk mode", mode);
  if (!wk || !wF || !wG || (mode != 3 && !wk1)) return set_error(OMHD_ERR_ARG, "null array");
  if (int rc = check_ext(ix, jx, kx)) return rc;
  if (ix < 3 || jx < 3 || (wH && kx < 3)) return ok();  // empty interior
  const double dtx = (mode == 2) ? 0.5 * dt / dx : dt / dx;  // rk.f90:22,55,86,118
  hipStream_t s = dropin_stream();
  const size_t n = (size_t)ix * jx * kx;
  Staged a, b, f, g, h;
  if (int rc = a.init(wk, 9 * n, true, mode == 1 || mode == 3, s)) return rc;
  if (mode != 3)
    if (int rc = b.init(wk1, 9 * n, true, mode == 0 || mode == 2, s)) return rc;
  if (int rc = f.init(wF, 9 * n, true, false, s)) return rc;
  if (int rc = g.init(wG, 9 * n, true, false, s)) return rc;
  if (wH)
    if (int rc = h.init(wH, 9 * n, true, false, s)) return rc;
  const size_t total = (size_t)(ix - 2) * (jx - 2) * (wH ? kx - 2 : 1);
  k_rk<<<nblocks(total), TPB, 0, s>>>(mode, a.dev, b.dev, f.dev, g.dev, wH ? h.dev : nullptr, dtx, ix, jx, kx);
  OMHD_HIP(hipGetLastError());
  if (int rc = a.finish(s)) return rc;
  if (int rc = b.finish(s)) return rc;
  a.release(s);
  b.release(s);
  f.release(s);
  g.release(s);
  h.release(s);
  OMHD_HIP(hipStreamSynchronize(s));
  return ok();
}

extern "C" int omhd_set_dt(const double *U, const double *V, double *vmax, double *dt, double dx, double cfl,
                           int32_t ix, int32_t jx, int32_t kx, int32_t ndim) {
  if (!U || !V || !vmax || !dt) return set_error(OMHD_ERR_ARG, "null argument");
  if (int rc = check_ext(ix, jx, kx)) return rc;
  hipStream_t s = dropin_stream();
  const size_t n = (size_t)ix * jx * kx;
  Staged a, b;
  if (int rc = a.init(U, 9 * n, true, false, s)) return rc;
  if (int rc = b.init(V, 4 * n, true, false, s)) return rc;
  unsigned long long *red;
  OMHD_HIP(hipMallocAsync((void **)&red, 2 * sizeof(unsigned long long), s));
  OMHD_HIP(hipMemsetAsync(red, 0, 2 * sizeof(unsigned long long), s));
  const int ni = ix > 2 ? ix - 2 : 1, nj = jx > 2 ? jx - 2 : 1, nk = kx > 2 ? kx - 2 : 1;
  const size_t total = (size_t)ni * nj * nk;
  if (ndim == 3)
    k_speed<3><<<nblocks(total), TPB, 0, s>>>(a.dev, b.dev, ix, jx, kx, red, (int *)(red + 1));
  else
    k_speed<2><<<nblocks(total), TPB, 0, s>>>(a.dev, b.dev, ix, jx, kx, red, (int *)(red + 1));
  OMHD_HIP(hipGetLastError());
  unsigned long long h[2];
  OMHD_HIP(hipMemcpyAsync(h, red, sizeof h, hipMemcpyDeviceToHost, s));
  OMHD_HIP(hipStreamSynchronize(s));
  (void)hipFreeAsync(red, s);
  a.release(s);
  b.release(s);
  double vm;
  memcpy(&vm, &h[0], sizeof vm);
  *vmax = vm;
  *dt = cfl * dx / vm;  // set_dt.f90:52
  if ((int)h[1] & 4) return set_error(OMHD_ERR_DT, "non-finite signal speed", 0, 0, 0, vm);
  if (!(*dt >= 1.0e-7)) return set_error(OMHD_ERR_DT, "dt is too small", 0, 0, 0, *dt);  // :55-65
  return ok();
}

// Device address of this GPU's `scratch`, looked up once per device (the
// array's name in host code is only a stand-in for the symbol).
static int noalloc_scratch(unsigned long long **out) {
  static std::mutex mu;
  static unsigned long long *addr[64] = {};
  int dev = 0;
  OMHD_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_error(OMHD_ERR_ARG, "device index out of range", dev);
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    OMHD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(OMHD_ERR_HIP, "no device address for the set_dt scratch array");
    addr[dev] = (unsigned long long *)p;
  }
  *out = addr[dev];
  return OMHD_OK;
}

extern "C" int omhd_set_dt_noalloc(const double *U, const double *V, double *vmax, double *dt, double dx, double cfl,
                                   int32_t ix, int32_t jx, int32_t kx, int32_t ndim) {
  if (!U || !V || !vmax || !dt) return set_error(OMHD_ERR_ARG, "null argument");
  if (int rc = check_ext(ix, jx, kx)) return rc;
  if (!is_device_ptr(U) || !is_device_ptr(V))
    return set_error(OMHD_ERR_ARG, "omhd_set_dt_noalloc: U and V must be device memory (host arrays need staging)");
  const int ni = ix > 2 ? ix - 2 : 1, nj = jx > 2 ? jx - 2 : 1, nk = kx > 2 ? kx - 2 : 1;
  const size_t total = (size_t)ni * nj * nk;
  if (total > NOALLOC_MAX_CELLS)
    return set_error(OMHD_ERR_EXTENT, "omhd_set_dt_noalloc: more than 67108864 interior cells (use omhd_set_dt)", ix, jx,
                     kx);
  unsigned long long *sc = nullptr;
  if (int rc = noalloc_scratch(&sc)) return rc;
  hipStream_t s = dropin_stream();
  const unsigned groups = (unsigned)((total + NOALLOC_CELLS_PER_GROUP - 1) / NOALLOC_CELLS_PER_GROUP);
  if (ndim == 3) k_speed_part<3><<<groups, TPB, 0, s>>>(U, V, ix, jx, kx, sc);
  else k_speed_part<2><<<groups, TPB, 0, s>>>(U, V, ix, jx, kx, sc);
  OMHD_HIP(hipGetLastError());
  k_speed_final<<<1, TPB, 0, s>>>(sc, (int)groups, sc + 2 * NOALLOC_GROUPS);
  OMHD_HIP(hipGetLastError());
  unsigned long long h[2];
  OMHD_HIP(hipMemcpyAsync(h, sc + 2 * NOALLOC_GROUPS, sizeof h, hipMemcpyDeviceToHost, s));
  OMHD_HIP(hipStreamSynchronize(s));
  double vm;
  memcpy(&vm, &h[0], sizeof vm);
  *vmax = vm;
  *dt = cfl * dx / vm;  // set_dt.f90:52
  if ((int)h[1] & 4) return set_error(OMHD_ERR_DT, "non-finite signal speed", 0, 0, 0, vm);
  if (!(*dt >= 1.0e-7)) return set_error(OMHD_ERR_DT, "dt is too small", 0, 0, 0, *dt);  // :55-65
  return ok();
}

static int bc_pair(double *dev, int ix, int jx, int kx, int dir, int tlo, int thi, const double *sl, const double *sh,
                   hipStream_t s) {
  const int na = (dir == 1) ? jx : ix, nb = (dir == 3) ? jx : kx;
  const size_t total = (size_t)na * nb;
  k_bc_face_pair<<<nblocks(total), TPB, 0, s>>>(dev, ix, jx, kx, dir, tlo, thi, sl[0], sl[1], sl[2], sl[3], sl[4],
                                                 sl[5], sl[6], sl[7], sl[8], sh[0], sh[1], sh[2], sh[3], sh[4], sh[5],
                                                 sh[6], sh[7], sh[8]);
  OMHD_HIP(hipGetLastError());
  return OMHD_OK;
}

extern "C" int omhd_bc_u(double *U, int32_t ix, int32_t jx, int32_t kx, const int32_t *bc_type, const double *sign) {
  if (!U || !bc_type) return set_error(OMHD_ERR_ARG, "null argument");
  if (int rc = check_ext(ix, jx, kx)) return rc;
  static const double ones[54] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                  1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  const double *sg = sign ? sign : ones;
  for (int f = 0; f < 6; f += 2)
    if (bc_type[f] != bc_type[f + 1] && (bc_type[f] == OMHD_BC_PERIODIC || bc_type[f + 1] == OMHD_BC_PERIODIC))
      return set_error(OMHD_ERR_ARG, "periodic faces come in pairs", f);
  hipStream_t s = dropin_stream();
  const size_t n = (size_t)ix * jx * kx;
  Staged a;
  if (int rc = a.init(U, 9 * n, true, true, s)) return rc;
  if (ix > 2)
    if (int rc = bc_pair(a.dev, ix, jx, kx, 1, bc_type[0], bc_type[1], sg, sg + 9, s)) return rc;
  if (jx > 2)
    if (int rc = bc_pair(a.dev, ix, jx, kx, 2, bc_type[2], bc_type[3], sg + 18, sg + 27, s)) return rc;
  if (kx > 2)
    if (int rc = bc_pair(a.dev, ix, jx, kx, 3, bc_type[4], bc_type[5], sg + 36, sg + 45, s)) return rc;
  if (int rc = a.finish(s)) return rc;
  a.release(s);
  OMHD_HIP(hipStreamSynchronize(s));
  return ok();
}

extern "C" int omhd_bc_face(double *VL, double *VR, int32_t ix, int32_t jx, int32_t kx, int32_t dir, int32_t bc_type,
                            const double *sign_lo, const double *sign_hi) {
  if (!VL || !VR) return set_error(OMHD_ERR_ARG, "null array");
  if (int rc = check_ext(ix, jx, kx)) return rc;
  if (dir < 1 || dir > 3) return set_error(OMHD_ERR_DIR, "bad dir", dir);
  static const double ones[9] = {1, 1, 1, 1, 1, 1, 1, 1, 1};
  const double *sl = sign_lo ? sign_lo : ones, *sh = sign_hi ? sign_hi : ones;
  hipStream_t s = dropin_stream();
  const size_t n = (size_t)ix * jx * kx;
  Staged l, r;
  if (int rc = l.init(VL, 9 * n, true, true, s)) return rc;
  if (int rc = r.init(VR, 9 * n, true, true, s)) return rc;
  const int na = (dir == 1) ? jx : ix, nb = (dir == 3) ? jx : kx;
  const int nl = dir == 1 ? ix : dir == 2 ? jx : kx;
  if (nl >= 2) {
    k_face_patch<<<nblocks((size_t)na * nb), TPB, 0, s>>>(l.dev, r.dev, ix, jx, kx, dir, bc_type, sl[0], sl[1], sl[2],
                                                          sl[3], sl[4], sl[5], sl[6], sl[7], sl[8], sh[0], sh[1],
                                                          sh[2], sh[3], sh[4], sh[5], sh[6], sh[7], sh[8]);
    OMHD_HIP(hipGetLastError());
  }
  if (int rc = l.finish(s)) return rc;
  if (int rc = r.finish(s)) return rc;
  l.release(s);
  r.release(s);
  OMHD_HIP(hipStreamSynchronize(s));
  return ok();
}

extern "C" int omhd_flux_resistive(double *F, const double *U, const double *eta_dx, int32_t ix, int32_t jx,
                                   int32_t dir) {
  if (!F || !U || !eta_dx) return set_error(OMHD_ERR_ARG, "null array");
  if (int rc = check_ext(ix, jx, 1)) return rc;
  if (dir < 1 || dir > 3) return set_error(OMHD_ERR_DIR, "bad dir", dir);
  // face ranges of flux_resistive.f90:49-55 (1-based i=1..ix-1, j=2..jx-1 for
  // dir 1; i=2..ix-1, j=1..jx-1 for dir 2); dir 3 has no 2D term (:56-57)
  const int i0 = dir == 1 ? 0 : 1, j0 = dir == 1 ? 1 : 0;
  const int ni = ix - 1 - i0, nj = jx - 1 - j0;
  if (dir == 3 || ni <= 0 || nj <= 0) return ok();
  hipStream_t s = dropin_stream();
  const size_t n = (size_t)ix * jx;
  Staged f, b, e;  // only F(en..bz) and U(bx..bz) cross the bus
  if (int rc = f.init(F + 3 * n, 5 * n, true, true, s)) return rc;
  if (int rc = b.init(U + 5 * n, 3 * n, true, false, s)) return rc;
  if (int rc = e.init(eta_dx, n, true, false, s)) return rc;
  const unsigned nb = nblocks((size_t)ni * nj);
  if (dir == 1)
    k_flux_resistive2d<1><<<nb, TPB, 0, s>>>(f.dev, b.dev, e.dev, ix, i0, j0, ni, nj, n);
  else
    k_flux_resistive2d<2><<<nb, TPB, 0, s>>>(f.dev, b.dev, e.dev, ix, i0, j0, ni, nj, n);
  OMHD_HIP(hipGetLastError());
  if (int rc = f.finish(s)) return rc;
  f.release(s);
  b.release(s);
  e.release(s);
  OMHD_HIP(hipStreamSynchronize(s));
  return ok();
}

extern "C" int omhd_flux_resistive3d(double *F, const double *U, const double *eta, int32_t eta_nz, double dx,
                                     int32_t ix, int32_t jx, int32_t kx, int32_t dir) {
  if (!F || !U || !eta) return set_error(OMHD_ERR_ARG, "null array");
  if (int rc = check_ext(ix, jx, kx)) return rc;
  if (dir < 1 || dir > 3) return set_error(OMHD_ERR_DIR, "bad dir", dir);
  if (eta_nz != 1 && eta_nz != kx) return set_error(OMHD_ERR_ARG, "eta_nz must be 1 or kx", eta_nz);
  // face ranges of flux_resistive.cuf:47-68: the normal index from 1, the
  // transverse ones over the interior (1-based)
  const int i0 = dir == 1 ? 0 : 1, j0 = dir == 2 ? 0 : 1, k0 = dir == 3 ? 0 : 1;
  const int ni = ix - 1 - i0, nj = jx - 1 - j0, nk = kx - 1 - k0;
  if (ni <= 0 || nj <= 0 || nk <= 0) return ok();
  hipStream_t s = dropin_stream();
  const size_t n = (size_t)ix * jx * kx;
  Staged f, b, e;  // only F(en..bz) and U(bx..bz) cross the bus
  if (int rc = f.init(F + 3 * n, 5 * n, true, true, s)) return rc;
  if (int rc = b.init(U + 5 * n, 3 * n, true, false, s)) return rc;
  if (int rc = e.init(eta, (size_t)ix * jx * eta_nz, true, false, s)) return rc;
  const unsigned nb = nblocks((size_t)ni * nj * nk);
  if (dir == 1)
    k_flux_resistive3d<1><<<nb, TPB, 0, s>>>(f.dev, b.dev, e.dev, eta_nz, dx, ix, jx, i0, j0, k0, ni, nj, nk, n);
  else if (dir == 2)
    k_flux_resistive3d<2><<<nb, TPB, 0, s>>>(f.dev, b.dev, e.dev, eta_nz, dx, ix, jx, i0, j0, k0, ni, nj, nk, n);
  else
    k_flux_resistive3d<3><<<nb, TPB, 0, s>>>(f.dev, b.dev, e.dev, eta_nz, dx, ix, jx, i0, j0, k0, ni, nj, nk, n);
  OMHD_HIP(hipGetLastError());
  if (int rc = f.finish(s)) return rc;
  f.release(s);
  b.release(s);
  e.release(s);
  OMHD_HIP(hipStreamSynchronize(s));
  return ok();
}
