// omhd_adjoint.hip -- discrete adjoint of the resident step (2D and 3D,
// periodic and wall boundaries, optional resistive flux).
//
// The forward map of one step with frozen (dt, ch) (2D_basic/main.f90:80-185):
//   Ug   = G(U)                       glm_ss2: psi *= f, f = exp(-0.5 dt ch / 0.2)
//   U*   = Ug + a1 D(Ug)              rk_tvd21 (a1 = dt/dx) | rk_std21 (0.5 dt/dx)
//   U**  = c0 Ug + c1 U* + a2 D(U*)   rk_tvd22 (0.5, 0.5, 0.5 dt/dx) | rk_std22 (1, 0, dt/dx)
//   U'   = G(U**)
// where D(W) = F(i-1) - F(i) + G(j-1) - G(j) [+ H(k-1) - H(k)] is the flux
// difference from u2v -> limiter -> HLLD (+ the resistive flux).  Its reverse, given lam' = dL/dU':
//   lam** = G lam'
//   lam*  = c1 lam** + (dD/dW at U*)^T (a2 lam**)
//   lamg  = c0 lam** + lam* + (dD/dW at Ug)^T (a1 lam*)
//   lam   = G lamg
// Each (dD/dW)^T mu is 2 + (number of axes) launches over the grid (plus the
// low wall faces, and the resistive kernels below):
//   k_adj_prim     primitives of W (u2v)
//   k_adj_face<x>  per x-face: recompute the two limited face states and the
//                  HLLD flux, run its reverse on bF = mu(c+1) - mu(c), store the
//                  adjoints of the face states by owning cell (AL at c, AR at c+1)
//   k_adj_face<y> (and z): the same along y (z), into their own AL/AR
//   k_adj_gather   per cell: the limiter reverse gathered, per axis, from the
//                  three cells whose slopes read it; the resistive B
//                  adjoints; the u2v reverse and the RK/GLM combine above,
//                  written in place.
// Walls: a wall face's states are mirrors of the inside face state, so its
// reverse folds both state adjoints onto the inside one; the slopes of the
// two edge cells read the cell itself through its mirror ghost.
// Resistive flux, owner-gather (no atomics, so the result does not depend on
// the order threads run in): k_adj_res stores each face's current-density and
// B-sum adjoints and its eta adjoint; k_adj_res_gather sums, per cell, the
// contributions of the ten-cell stencils that read it; k_adj_eta sums each
// face column's eta adjoints over k into the face's eta entry.
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
  int pitch, nx, ny, nz, nd;
  int bc[3];               // OMHD_BC_* per axis
  unsigned slo[3], shi[3]; // slots negated in the low / high mirror ghosts
  const double *eta;       // resistive: 3 face arrays of (nx+2)*(ny+2), or null
  double dx;
};

__device__ __forceinline__ bool cell_of(const Geo &g, int &i, int &j, int &k) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx * g.ny * g.nz) return false;
  i = (int)(t % g.nx);
  const long long r = t / g.nx;
  j = (int)(r % g.ny);
  k = (int)(r / g.ny);
  return true;
}

// The same bijection of threads onto the grid's cells as cell_of, visited in
// another order, for a kernel whose 5-point stencils along y and z re-read
// neighbouring rows: blocks are dealt to the 8 XCDs in contiguous ranges (the
// dispatcher hands workgroup b to XCD b % 8, and each XCD has its own L2), and
// the rows (j, k) are walked in tiles of TY x TZ rows, TY and TZ the largest
// powers of two up to 16 and 8 that divide ny and nz, so that the rows in
// flight on one XCD are each other's neighbours along both axes.
__device__ __forceinline__ bool cell_of_tiled(const Geo &g, int &i, int &j, int &k) {
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned L = (nb % 8u == 0u) ? (b % 8u) * (nb / 8u) + b / 8u : b;
  const long long t = (long long)L * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nx * g.ny * g.nz) return false;
  i = (int)(t % g.nx);
  const long long q = t / g.nx;  // row ordinal in visiting order, < ny * nz
  const int TY = min(16, g.ny & -g.ny), TZ = min(8, g.nz & -g.nz);
  const int in = (int)(q % (TY * TZ));
  const long long T = q / (TY * TZ);
  const int nty = g.ny / TY;
  j = (int)(T % nty) * TY + in % TY;
  k = (int)(T / nty) * TZ + in / TY;
  return true;
}

// by selects, not branches: the lane-dependent tests would otherwise split
// the wavefront at every stencil position
__device__ __forceinline__ int map_ax(int c, int n, int bc, unsigned slo, unsigned shi, unsigned &mask) {
  const bool per = bc == OMHD_BC_PERIODIC, lo = c < 0, hi = c >= n;
  mask ^= (lo && !per) ? slo : 0u;
  mask ^= (hi && !per) ? shi : 0u;
  const int rlo = per ? c + n : -1 - c, rhi = per ? c - n : 2 * n - 1 - c;
  return lo ? rlo : (hi ? rhi : c);
}

// interior cell of (i, j, k) (each coordinate at most two cells outside) and
// the negate mask of its ghost image: periodic wraps, mirrors with sign tables
struct Cell {
  long long o;
  unsigned mask;
};
// MIR = false: every axis is periodic (the host's choice of kernel), so the
// map is a wrap by selects and the mask is the constant 0 (no sign flips).
__device__ __forceinline__ int wrap_ax(int c, int n) { return c < 0 ? c + n : (c >= n ? c - n : c); }
template <bool MIR = true>
__device__ __forceinline__ Cell cell_at(const Geo &g, int i, int j, int k) {
  if constexpr (!MIR) {
    i = wrap_ax(i, g.nx);
    j = wrap_ax(j, g.ny);
    if (g.nd == 3) k = wrap_ax(k, g.nz);
    return Cell{(long long)k * g.zstride + (long long)j * g.pitch + i, 0u};
  }
  unsigned mk = 0;
  i = map_ax(i, g.nx, g.bc[0], g.slo[0], g.shi[0], mk);
  j = map_ax(j, g.ny, g.bc[1], g.slo[1], g.shi[1], mk);
  if (g.nd == 3) k = map_ax(k, g.nz, g.bc[2], g.slo[2], g.shi[2], mk);
  return Cell{(long long)k * g.zstride + (long long)j * g.pitch + i, mk};
}

// The cell s steps from the interior cell (i, j, k) along axis DIR only: the
// other coordinates need no map.  MIR = false: axis DIR is periodic.
template <int DIR, bool MIR = true>
__device__ __forceinline__ Cell nbr(const Geo &g, int i, int j, int k, int s) {
  unsigned mk = 0;
  const int n = DIR == 0 ? g.nx : (DIR == 1 ? g.ny : g.nz);
  int c = (DIR == 0 ? i : (DIR == 1 ? j : k)) + s;
  if constexpr (MIR) c = map_ax(c, n, g.bc[DIR], g.slo[DIR], g.shi[DIR], mk);
  else c = wrap_ax(c, n);
  if (DIR == 0) i = c;
  else if (DIR == 1) j = c;
  else k = c;
  return Cell{(long long)k * g.zstride + (long long)j * g.pitch + i, mk};
}

__device__ __forceinline__ double flip(double v, unsigned mask, int m) {
  const long long sbit = (long long)((mask >> m) & 1u) << 63;
  return __longlong_as_double(__double_as_longlong(v) ^ sbit);
}

__device__ __forceinline__ double ld(const double *a, const Geo &g, const Cell &c, int m) {
  return flip(a[m * g.plane + c.o], c.mask, m);
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

// Resistive reverse, owner-gather form (no atomics, fixed summation order).
// Face arrays: per axis DIR, 5 planes over (nx + [DIR==0]) (ny + [DIR==1])
// (nz + [DIR==2]) faces, the DIR coordinate of the face's left cell stored
// shifted by one (left cell -1 .. n-1: the low wall face is -1).
__host__ __device__ __forceinline__ long long fplane_of(const Geo &g, int d) {
  return (long long)(g.nx + (d == 0)) * (g.ny + (d == 1)) * (g.nz + (d == 2));
}
__host__ __device__ __forceinline__ long long fbase_of(const Geo &g, int d) {
  long long b = 0;
  for (int e = 0; e < d; e++) b += 5 * fplane_of(g, e);
  return b;
}
template <int DIR>
__device__ __forceinline__ long long face_idx(const Geo &g, int i, int j, int k) {
  return ((long long)(k + (DIR == 2)) * (g.ny + (DIR == 1)) + (j + (DIR == 1))) * (g.nx + (DIR == 0)) + (i + (DIR == 0));
}

// Reverse of the resistive flux at the face with left cell (i,j,k) along DIR
// (3D_reconnection_gpu/flux_resistive.cuf:86-135), first half: the face's
// current-density and B-sum adjoints (bJ1, bS1, bJ2, bS2) and its eta adjoint,
// stored at the face.  J1 pairs with slot q1 in the energy flux and feeds slot
// q1's flux with +eta J1; J2 with q2 and -eta J2.  The B adjoints of the
// ten stencil cells are gathered by res_gather.
template <int DIR>
__device__ void resistive_face(const Geo &g, const double *__restrict__ w, int i, int j, int k, const double bF[NV],
                               double *__restrict__ F) {
  constexpr int q1 = DIR == 0 ? 7 : (DIR == 1 ? 5 : 6);  // x: bz, y: bx, z: by
  constexpr int q2 = DIR == 0 ? 6 : (DIR == 1 ? 7 : 5);  // x: by, y: bz, z: bx
  constexpr int t1 = 5 + DIR;                             // slot of J1's and J2's transverse differences
  constexpr int di = DIR == 0, dj = DIR == 1, dk = DIR == 2;
  // J1's 0.25-terms vary along axis A1, J2's along A2 (x: z and y; y: x and z; z: y and x)
  constexpr int A1 = DIR == 0 ? 2 : (DIR == 1 ? 0 : 1);
  constexpr int A2 = DIR == 0 ? 1 : (DIR == 1 ? 2 : 0);
  const int ix = g.nx + 2;
  const double e = g.eta[(long long)DIR * ix * (g.ny + 2) + (long long)(i + 1) + (long long)ix * (j + 1)] / g.dx;
  const Cell c0 = cell_at(g, i, j, k), c1 = cell_at(g, i + di, j + dj, k + dk);
  auto sh = [&](int a, int s, int base_i, int base_j, int base_k) {
    return cell_at(g, base_i + (a == 0 ? s : 0), base_j + (a == 1 ? s : 0), base_k + (a == 2 ? s : 0));
  };
  const Cell p1a = sh(A1, 1, i, j, k), p1b = sh(A1, 1, i + di, j + dj, k + dk);
  const Cell m1a = sh(A1, -1, i, j, k), m1b = sh(A1, -1, i + di, j + dj, k + dk);
  const Cell p2a = sh(A2, 1, i, j, k), p2b = sh(A2, 1, i + di, j + dj, k + dk);
  const Cell m2a = sh(A2, -1, i, j, k), m2b = sh(A2, -1, i + di, j + dj, k + dk);
  // J1 = -(q1(c+1) - q1(c)) + 0.25 (t1(+A1) terms - t1(-A1) terms)  [x: JyS, y: JzS, z: JxS]
  // J2 =  (q2(c+1) - q2(c)) - 0.25 (t1(+A2) terms - t1(-A2) terms)
  const double J1 = -(ld(w, g, c1, q1) - ld(w, g, c0, q1)) +
                    0.25 * (ld(w, g, p1a, t1) + ld(w, g, p1b, t1) - ld(w, g, m1a, t1) - ld(w, g, m1b, t1));
  const double J2 = (ld(w, g, c1, q2) - ld(w, g, c0, q2)) -
                    0.25 * (ld(w, g, p2a, t1) + ld(w, g, p2b, t1) - ld(w, g, m2a, t1) - ld(w, g, m2b, t1));
  const double S1 = ld(w, g, c0, q1) + ld(w, g, c1, q1), S2 = ld(w, g, c0, q2) + ld(w, g, c1, q2);
  // F[3] += 0.5 e (J1 S1 - J2 S2); F[q1] += e J1; F[q2] -= e J2
  const long long fp = fplane_of(g, DIR), o = face_idx<DIR>(g, i, j, k);
  F[o] = e * (0.5 * bF[3] * S1 + bF[q1]);           // bJ1
  F[fp + o] = 0.5 * e * bF[3] * J1;                 // bS1
  F[2 * fp + o] = e * (-0.5 * bF[3] * S2 - bF[q2]);  // bJ2
  F[3 * fp + o] = -0.5 * e * bF[3] * J2;            // bS2
  F[4 * fp + o] = (0.5 * bF[3] * (J1 * S1 - J2 * S2) + bF[q1] * J1 - bF[q2] * J2) / g.dx;
}

// The face coordinates f in [flo, n-1] along one axis whose stencil position
// f + s maps to cell x under cell_at's map (itself, the periodic images, or
// the mirror images with their negate masks); at most three, held in scalars
// (selected by index) so that the list stays in registers.
struct Pre {
  int f0, f1, f2, cnt;
  unsigned m0, m1, m2;
  __device__ __forceinline__ int f(int u) const { return u == 0 ? f0 : (u == 1 ? f1 : f2); }
  __device__ __forceinline__ unsigned m(int u) const { return u == 0 ? m0 : (u == 1 ? m1 : m2); }
};
__device__ __forceinline__ Pre pre_ax(int x, int s, int n, int flo, bool periodic, unsigned slo, unsigned shi) {
  Pre p{0, 0, 0, 0, 0u, 0u, 0u};
  auto add = [&](int raw, unsigned m) {
    const int fc = raw - s;
    if (fc >= flo && fc <= n - 1) {
      if (p.cnt == 0) {
        p.f0 = fc;
        p.m0 = m;
      } else if (p.cnt == 1) {
        p.f1 = fc;
        p.m1 = m;
      } else {
        p.f2 = fc;
        p.m2 = m;
      }
      p.cnt++;
    }
  };
  add(x, 0u);
  if (periodic) {
    add(x + n, 0u);
    add(x - n, 0u);
  } else {
    add(-1 - x, slo);
    add(2 * n - 1 - x, shi);
  }
  return p;
}

// Second half: the B adjoints of cell (i,j,k) from every face along DIR whose
// stencil reads it, in a fixed order: per role (the face's left cell c0 or
// right cell c1, each shifted by 0 or +-1 along A1 or A2), every face whose
// role position maps to (i,j,k), with the sign of the ghost image.
template <int DIR>
__device__ __forceinline__ void res_gather(const Geo &g, const double *__restrict__ F, int i, int j, int k,
                                           double b[3]) {
  constexpr int q1 = DIR == 0 ? 7 : (DIR == 1 ? 5 : 6);
  constexpr int q2 = DIR == 0 ? 6 : (DIR == 1 ? 7 : 5);
  constexpr int t1 = 5 + DIR;
  constexpr int A1 = DIR == 0 ? 2 : (DIR == 1 ? 0 : 1);
  constexpr int A2 = DIR == 0 ? 1 : (DIR == 1 ? 2 : 0);
  const long long fp = fplane_of(g, DIR);
  const int x[3] = {i, j, k}, n[3] = {g.nx, g.ny, g.nz};
#pragma unroll
  for (int dl = 0; dl < 2; dl++) {
#pragma unroll
    for (int r = 0; r < 5; r++) {  // 0: no shift; 1, 2: +1, -1 along A1; 3, 4: +1, -1 along A2
      int s[3] = {0, 0, 0};
      s[DIR] = dl;
      if (r == 1 || r == 2) s[A1] = r == 1 ? 1 : -1;
      if (r == 3 || r == 4) s[A2] = r == 3 ? 1 : -1;
      Pre P[3];
#pragma unroll
      for (int a = 0; a < 3; a++) {
        const bool per = g.bc[a] == OMHD_BC_PERIODIC;
        P[a] = pre_ax(x[a], s[a], n[a], (a == DIR && !per) ? -1 : 0, per, g.slo[a], g.shi[a]);
      }
      for (int u = 0; u < P[0].cnt; u++)
        for (int v = 0; v < P[1].cnt; v++)
          for (int z = 0; z < P[2].cnt; z++) {
            const long long o = face_idx<DIR>(g, P[0].f(u), P[1].f(v), P[2].f(z));
            const unsigned m = P[0].m(u) ^ P[1].m(v) ^ P[2].m(z);
            if (r == 0) {
              const double bJ1 = F[o], bS1 = F[fp + o], bJ2 = F[2 * fp + o], bS2 = F[3 * fp + o];
              b[q1 - 5] += flip(dl == 0 ? bJ1 + bS1 : -bJ1 + bS1, m, q1);
              b[q2 - 5] += flip(dl == 0 ? -bJ2 + bS2 : bJ2 + bS2, m, q2);
            } else if (r <= 2) {
              const double bJ1 = F[o];
              b[t1 - 5] += flip(r == 1 ? 0.25 * bJ1 : -0.25 * bJ1, m, t1);
            } else {
              const double bJ2 = F[2 * fp + o];
              b[t1 - 5] += flip(r == 3 ? -0.25 * bJ2 : 0.25 * bJ2, m, t1);
            }
          }
    }
  }
}

// The face-state adjoints of one face.  kind 0: interior/periodic face
// (c, c+1); 1: high wall face (c = n-1, right state = mirror of the left);
// 2: low wall face (-1, 0) (left state = mirror of the right).  mu =
// mscale * M (slot 8 also times mpsi).  Stored by owning cell: AL(c) for the
// left state of (c, c+1), AR(c) for the right state of (c-1, c); a wall
// face's two states both belong to its interior cell.
template <int DIR, int LIM, bool MIR = true>
__device__ __forceinline__ void face_vjp(const Geo &g, const double *__restrict__ w, const double *__restrict__ M,
                                         double mscale, double mpsi, double ch, double *__restrict__ AL,
                                         double *__restrict__ AR, int i, int j, int k, int kind) {
  const Cell cm = nbr<DIR, MIR>(g, i, j, k, -1), c0 = nbr<DIR, MIR>(g, i, j, k, 0), c1 = nbr<DIR, MIR>(g, i, j, k, 1);
  double VL[NV], VR[NV], bF[NV], bL[NV], bR[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) {
    const double am = ld(w, g, cm, m), a0 = ld(w, g, c0, m), a1 = ld(w, g, c1, m);
    double gR, gL;
    adj::lim_fwd<double>(LIM, am, a0, a1, gR, gL);
    const double s = m == 8 ? mscale * mpsi : mscale;
    const double mu0 = M[m * g.plane + c0.o];
    if (kind == 0) {
      const Cell c2 = nbr<DIR, MIR>(g, i, j, k, 2);
      double hR, hL;
      adj::lim_fwd<double>(LIM, a0, a1, ld(w, g, c2, m), hR, hL);
      VL[m] = gL;
      VR[m] = hR;
      bF[m] = s * (M[m * g.plane + c1.o] - mu0);  // F enters D(c) with -1, D(c+1) with +1
    } else if (kind == 1) {
      VL[m] = gL;
      VR[m] = flip(gL, g.shi[DIR], m);
      bF[m] = -(s * mu0);
    } else {
      VR[m] = gR;
      VL[m] = flip(gR, g.slo[DIR], m);
      bF[m] = s * mu0;
    }
    bL[m] = 0.0;
    bR[m] = 0.0;
  }
  adj::hlld_vjp<DIR>(VL, VR, ch, bF, bL, bR);
#pragma unroll
  for (int m = 0; m < NV; m++) {
    if (kind == 0) {
      AL[m * g.plane + c0.o] = bL[m];
      AR[m * g.plane + c1.o] = bR[m];
    } else if (kind == 1) {
      AL[m * g.plane + c0.o] = bL[m] + flip(bR[m], g.shi[DIR], m);
    } else {
      AR[m * g.plane + c0.o] = bR[m] + flip(bL[m], g.slo[DIR], m);
    }
  }
}

// one thread per cell c along DIR: the face (c, c+1), or the high wall face.
// MIR = false: axis DIR is periodic (only interior/periodic faces).
template <int DIR, int LIM, bool MIR>
__global__ __launch_bounds__(128) void k_adj_face(const Geo g, const double *__restrict__ w,
                                                  const double *__restrict__ M, double mscale, double mpsi, double ch,
                                                  double *__restrict__ AL, double *__restrict__ AR) {
  int i, j, k;
  if (!cell_of(g, i, j, k)) return;
  const int c = DIR == 0 ? i : (DIR == 1 ? j : k);
  const int n = DIR == 0 ? g.nx : (DIR == 1 ? g.ny : g.nz);
  if constexpr (MIR) {
    // a wavefront whose faces' stencils (c-1 .. c+2) all lie inside the
    // grid needs no ghost map along DIR: the periodic form, same arithmetic
    if (__all(c >= 1 && c <= n - 3)) {
      face_vjp<DIR, LIM, false>(g, w, M, mscale, mpsi, ch, AL, AR, i, j, k, 0);
      return;
    }
  }
  const bool high = MIR && g.bc[DIR] != OMHD_BC_PERIODIC && c == n - 1;
  face_vjp<DIR, LIM, MIR>(g, w, M, mscale, mpsi, ch, AL, AR, i, j, k, high ? 1 : 0);
}

// walls only: one thread per cell of the plane c = 0, the low wall face
template <int DIR, int LIM>
__global__ __launch_bounds__(128) void k_adj_face_lo(const Geo g, const double *__restrict__ w,
                                                     const double *__restrict__ M, double mscale, double mpsi,
                                                     double ch, double *__restrict__ AL, double *__restrict__ AR) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int na = DIR == 0 ? g.ny : g.nx, nb = DIR == 2 ? g.ny : g.nz;  // the two transverse extents
  if (t >= (long long)na * nb) return;
  const int u = (int)(t % na), v = (int)(t / na);
  const int i = DIR == 0 ? 0 : u, j = DIR == 0 ? u : (DIR == 1 ? 0 : v), k = DIR == 2 ? 0 : v;
  face_vjp<DIR, LIM>(g, w, M, mscale, mpsi, ch, AL, AR, i, j, k, 2);
}

// the resistive flux's reverse at every face along DIR, first half: one
// thread per cell c (the face (c, c+1), or the high wall face), plus the low
// wall face for c = 0 at a wall; bF is rebuilt from mu as the face kernels
// build it.  rf: the face arrays of all three axes.
template <int DIR>
__global__ __launch_bounds__(256) void k_adj_res(const Geo g, const double *__restrict__ w,
                                                 const double *__restrict__ M, double mscale, double mpsi,
                                                 double *__restrict__ rf) {
  int i, j, k;
  if (!cell_of(g, i, j, k)) return;
  const int c = DIR == 0 ? i : (DIR == 1 ? j : k);
  const int n = DIR == 0 ? g.nx : (DIR == 1 ? g.ny : g.nz);
  const bool wall = g.bc[DIR] != OMHD_BC_PERIODIC;
  const Cell c0 = nbr<DIR>(g, i, j, k, 0), c1 = nbr<DIR>(g, i, j, k, 1);
  double *F = rf + fbase_of(g, DIR);
  double bF[NV];
  for (int f = 0; f < ((wall && c == 0) ? 2 : 1); f++) {
    const int kind = f == 1 ? 2 : ((wall && c == n - 1) ? 1 : 0);
#pragma unroll
    for (int m = 0; m < NV; m++) {
      const double s = m == 8 ? mscale * mpsi : mscale;
      const double mu0 = M[m * g.plane + c0.o];
      bF[m] = kind == 0 ? s * (M[m * g.plane + c1.o] - mu0) : (kind == 1 ? -(s * mu0) : s * mu0);
    }
    if (kind == 2) resistive_face<DIR>(g, w, i - (DIR == 0), j - (DIR == 1), k - (DIR == 2), bF, F);
    else resistive_face<DIR>(g, w, i, j, k, bF, F);
  }
}

// res_gather for a cell one or more cells from every boundary: each role's
// position is then read by exactly one face, (i,j,k) minus the role's shift,
// with no ghost image; the same sums in the same order as res_gather.
template <int DIR>
__device__ __forceinline__ void res_gather_inner(const Geo &g, const double *__restrict__ F, int i, int j, int k,
                                                 double b[3]) {
  constexpr int q1 = DIR == 0 ? 7 : (DIR == 1 ? 5 : 6);
  constexpr int q2 = DIR == 0 ? 6 : (DIR == 1 ? 7 : 5);
  constexpr int t1 = 5 + DIR;
  constexpr int A1 = DIR == 0 ? 2 : (DIR == 1 ? 0 : 1);
  constexpr int A2 = DIR == 0 ? 1 : (DIR == 1 ? 2 : 0);
  const long long fp = fplane_of(g, DIR);
  const int x[3] = {i, j, k};
#pragma unroll
  for (int dl = 0; dl < 2; dl++) {
#pragma unroll
    for (int r = 0; r < 5; r++) {
      int s[3] = {0, 0, 0};
      s[DIR] = dl;
      if (r == 1 || r == 2) s[A1] = r == 1 ? 1 : -1;
      if (r == 3 || r == 4) s[A2] = r == 3 ? 1 : -1;
      const long long o = face_idx<DIR>(g, x[0] - s[0], x[1] - s[1], x[2] - s[2]);
      if (r == 0) {
        const double bJ1 = F[o], bS1 = F[fp + o], bJ2 = F[2 * fp + o], bS2 = F[3 * fp + o];
        b[q1 - 5] += dl == 0 ? bJ1 + bS1 : -bJ1 + bS1;
        b[q2 - 5] += dl == 0 ? -bJ2 + bS2 : bJ2 + bS2;
      } else if (r <= 2) {
        const double bJ1 = F[o];
        b[t1 - 5] += r == 1 ? 0.25 * bJ1 : -0.25 * bJ1;
      } else {
        const double bJ2 = F[2 * fp + o];
        b[t1 - 5] += r == 3 ? -0.25 * bJ2 : 0.25 * bJ2;
      }
    }
  }
}

// second half: one thread per cell, the B adjoints from the faces of all
// three axes, written to R (3 planes, slots bx, by, bz)
__global__ __launch_bounds__(256) void k_adj_res_gather(const Geo g, const double *__restrict__ rf,
                                                        double *__restrict__ R) {
  int i, j, k;
  if (!cell_of(g, i, j, k)) return;
  double b[3] = {0.0, 0.0, 0.0};
  if (i >= 1 && i <= g.nx - 2 && j >= 1 && j <= g.ny - 2 && k >= 1 && k <= g.nz - 2) {
    res_gather_inner<0>(g, rf + fbase_of(g, 0), i, j, k, b);
    res_gather_inner<1>(g, rf + fbase_of(g, 1), i, j, k, b);
    res_gather_inner<2>(g, rf + fbase_of(g, 2), i, j, k, b);
  } else {
    res_gather<0>(g, rf + fbase_of(g, 0), i, j, k, b);
    res_gather<1>(g, rf + fbase_of(g, 1), i, j, k, b);
    res_gather<2>(g, rf + fbase_of(g, 2), i, j, k, b);
  }
  const long long o = (long long)k * g.zstride + (long long)j * g.pitch + i;
#pragma unroll
  for (int m = 0; m < 3; m++) R[m * g.plane + o] = b[m];
}

// the eta adjoints: one thread per entry (DIR, i+1, j+1) of geta, the face
// (i, j) along DIR summed over k in increasing order and added to the entry
__global__ __launch_bounds__(256) void k_adj_eta(const Geo g, const double *__restrict__ rf,
                                                 double *__restrict__ geta) {
  const int ix = g.nx + 2, jx = g.ny + 2;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3LL * ix * jx) return;
  const int d = (int)(t / ((long long)ix * jx)), r = (int)(t % ((long long)ix * jx));
  const int fi = r % ix - 1, fj = r / ix - 1;
  const bool wx = g.bc[0] != OMHD_BC_PERIODIC, wy = g.bc[1] != OMHD_BC_PERIODIC, wz = g.bc[2] != OMHD_BC_PERIODIC;
  if (fi < ((d == 0 && wx) ? -1 : 0) || fi > g.nx - 1 || fj < ((d == 1 && wy) ? -1 : 0) || fj > g.ny - 1) return;
  const int fx = g.nx + (d == 0), fy = g.ny + (d == 1);
  const long long fp = fplane_of(g, d);
  const double *ge = rf + fbase_of(g, d) + 4 * fp;
  double s = 0.0;
  for (int fk = (d == 2 && wz) ? -1 : 0; fk < g.nz; fk++)
    s += ge[((long long)(fk + (d == 2)) * fy + (fj + (d == 1))) * fx + (fi + (d == 0))];
  geta[t] += s;
}

struct Combine {
  const double *A;  // lam'
  double cA, fA;    // coefficient, and the psi factor applied to A
  const double *B;  // lam* (or null)
  double cB;
  double fO;        // psi factor applied to the result
  double *out;
};

// The limiter reverse along axis D at cell (i, j, k): the three cells whose
// slopes read it (coefficient index 1 - s for the slope at q = c + s; at a
// wall, the edge cells q = 0 and q = n - 1 also read it through their mirror
// ghost).  MIR = false: axis D is periodic.  al, ar: the axis's AL/AR.
template <int D, int LIM, bool MIR>
__device__ __forceinline__ void gather_axis(const Geo &g, const double *__restrict__ w, const double *__restrict__ al,
                                            const double *__restrict__ ar, int i, int j, int k, double b[NV]) {
  const int cc = D == 0 ? i : (D == 1 ? j : k);
  const int n = D == 0 ? g.nx : (D == 1 ? g.ny : g.nz);
  Cell o[5];
#pragma unroll
  for (int s = 0; s < 5; s++) o[s] = nbr<D, MIR>(g, i, j, k, s - 2);
#pragma unroll
  for (int m = 0; m < NV; m++) {
    double v[5];
#pragma unroll
    for (int s = 0; s < 5; s++) v[s] = ld(w, g, o[s], m);
    double acc = 0.0;
#pragma unroll
    for (int s = -1; s <= 1; s++) {
      const int q = cc + s;
      if (MIR && (q < 0 || q > n - 1)) continue;
      double cR[3], cL[3];
      adj::lim_coef(LIM, v[s + 1], v[s + 2], v[s + 3], cR, cL);
      const double l = al[m * g.plane + o[s + 2].o], r = ar[m * g.plane + o[s + 2].o];
      if (s == 0) acc += l + r;  // gL = w + gradL, gR = w - gradR
      acc += cL[1 - s] * l - cR[1 - s] * r;
      if (MIR && s == 0 && q == 0) acc += flip(cL[0] * l - cR[0] * r, g.slo[D], m);
      if (MIR && s == 0 && q == n - 1) acc += flip(cL[2] * l - cR[2] * r, g.shi[D], m);
    }
    b[m] += acc;
  }
}

// One thread per cell, all axes: the limiter reverse of the three cells per
// axis whose slopes read this cell (coefficient index 1 - s for the slope at
// q = c + s; at a wall, the edge cells q = 0 and q = n - 1 also read it
// through their mirror ghost), then the resistive B adjoints, the u2v
// reverse and the RK/GLM combine into c.out.  AL/AR hold one array per axis.
template <int LIM, int ND, bool MIR>
__global__ __launch_bounds__(256) void k_adj_gather(const Geo g, const double *__restrict__ w,
                                                    const double *__restrict__ AL, const double *__restrict__ AR,
                                                    long long astride, const double *__restrict__ R, const Combine c) {
  int i, j, k;
  if (!cell_of_tiled(g, i, j, k)) return;
  const long long o0 = (long long)k * g.zstride + (long long)j * g.pitch + i;
  double b[NV], wc[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) {
    wc[m] = w[m * g.plane + o0];
    b[m] = 0.0;
  }
  // per axis, the mirror logic only where that axis is a wall and some cell
  // of the wavefront has a stencil (c-2 .. c+2) reaching past it (uniform tests)
  if (!MIR || g.bc[0] == OMHD_BC_PERIODIC || __all(i >= 2 && i <= g.nx - 3))
    gather_axis<0, LIM, false>(g, w, AL, AR, i, j, k, b);
  else gather_axis<0, LIM, true>(g, w, AL, AR, i, j, k, b);
  if (!MIR || g.bc[1] == OMHD_BC_PERIODIC || __all(j >= 2 && j <= g.ny - 3))
    gather_axis<1, LIM, false>(g, w, AL + astride, AR + astride, i, j, k, b);
  else gather_axis<1, LIM, true>(g, w, AL + astride, AR + astride, i, j, k, b);
  if constexpr (ND == 3) {
    if (!MIR || g.bc[2] == OMHD_BC_PERIODIC || __all(k >= 2 && k <= g.nz - 3))
      gather_axis<2, LIM, false>(g, w, AL + 2 * astride, AR + 2 * astride, i, j, k, b);
    else gather_axis<2, LIM, true>(g, w, AL + 2 * astride, AR + 2 * astride, i, j, k, b);
  }
  if (R) {
#pragma unroll
    for (int m = 5; m < 8; m++) b[m] += R[(m - 5) * g.plane + o0];
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
typedef void (*res_fn)(const Geo, const double *, const double *, double, double, double *);
typedef void (*gather_fn)(const Geo, const double *, const double *, const double *, long long, const double *,
                          const Combine);

// LO: the low wall faces; otherwise MIR selects the kernel for a grid with
// wall axes (true) or a fully periodic one (false)
template <int DIR, int LO, bool MIR>
face_fn face_kernel(int lim) {
  switch (lim) {
    case 0: return LO ? k_adj_face_lo<DIR, 0> : k_adj_face<DIR, 0, MIR>;
    case 1: return LO ? k_adj_face_lo<DIR, 1> : k_adj_face<DIR, 1, MIR>;
    case 2: return LO ? k_adj_face_lo<DIR, 2> : k_adj_face<DIR, 2, MIR>;
    case 3: return LO ? k_adj_face_lo<DIR, 3> : k_adj_face<DIR, 3, MIR>;
    default: return LO ? k_adj_face_lo<DIR, 4> : k_adj_face<DIR, 4, MIR>;
  }
}

template <int DIR>
face_fn face_kernel(int lo, bool mir, int lim) {
  if (lo) return face_kernel<DIR, 1, true>(lim);
  return mir ? face_kernel<DIR, 0, true>(lim) : face_kernel<DIR, 0, false>(lim);
}

face_fn face_kernel(int dir, int lo, bool mir, int lim) {
  if (dir == 0) return face_kernel<0>(lo, mir, lim);
  if (dir == 1) return face_kernel<1>(lo, mir, lim);
  return face_kernel<2>(lo, mir, lim);
}

res_fn res_kernel(int dir) { return dir == 0 ? k_adj_res<0> : (dir == 1 ? k_adj_res<1> : k_adj_res<2>); }

template <int ND, bool MIR>
gather_fn gather_kernel(int lim) {
  switch (lim) {
    case 0: return k_adj_gather<0, ND, MIR>;
    case 1: return k_adj_gather<1, ND, MIR>;
    case 2: return k_adj_gather<2, ND, MIR>;
    case 3: return k_adj_gather<3, ND, MIR>;
    default: return k_adj_gather<4, ND, MIR>;
  }
}

gather_fn gather_kernel(int nd, bool mir, int lim) {
  if (nd == 3) return mir ? gather_kernel<3, true>(lim) : gather_kernel<3, false>(lim);
  return mir ? gather_kernel<2, true>(lim) : gather_kernel<2, false>(lim);
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
  double *work = nullptr;      // w, AL[3 axes], AR[3 axes], lam, lam*
  double *R = nullptr;         // resistive: B adjoints (3 planes)
  double *rf = nullptr;        // resistive: face arrays, 5 planes per axis (fplane_of)
  double *geta = nullptr;      // resistive: adjoint of each face's eta, 3 (nx+2)(ny+2)
  std::vector<double> geta_host;
  bool have_forward = false, have_backward = false;
  double *w() { return work; }
  double *AL(int d = 0) { return work + (1 + d) * state; }
  double *AR(int d = 0) { return work + (4 + d) * state; }
  double *lam() { return work + 7 * state; }
  double *lams() { return work + 8 * state; }
  size_t neta() const { return (size_t)3 * (h->cfg.nx + 2) * (h->cfg.ny + 2); }
};

static Geo geo_of(const omhd_handle *h) {
  Geo g;
  g.plane = h->plane;
  g.zstride = h->zstride;
  g.pitch = h->pitch;
  g.nx = h->cfg.nx;
  g.ny = h->cfg.ny;
  g.nz = h->cfg.nz;
  g.nd = h->nd;
  for (int a = 0; a < 3; a++) {
    g.bc[a] = h->faces.bc[a];
    g.slo[a] = h->faces.slo[a];
    g.shi[a] = h->faces.shi[a];
  }
  g.eta = h->eta;
  g.dx = h->cfg.dx;
  return g;
}

extern "C" int omhd_adjoint_create(omhd_handle *h, int32_t nsteps, int32_t every, omhd_adjoint **out) {
  if (!h || !out) return set_error(OMHD_ERR_ARG, "null argument");
  if (nsteps < 1 || every < 1) return set_error(OMHD_ERR_ARG, "need nsteps >= 1 and every >= 1", nsteps, every);
  if (h->cfg.flux_type != 3) return set_error(OMHD_ERR_TYPE, "the adjoint is built for the HLLD flux", h->cfg.flux_type);
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
  if (e == hipSuccess) e = hipMalloc((void **)&a->work, a->state * 9 * sizeof(double));
  if (e == hipSuccess && h->cfg.resistive) e = hipMalloc((void **)&a->R, (size_t)h->plane * 3 * sizeof(double));
  if (e == hipSuccess && h->cfg.resistive) e = hipMalloc((void **)&a->geta, a->neta() * sizeof(double));
  if (e == hipSuccess && h->cfg.resistive) {
    const Geo g = geo_of(h);
    e = hipMalloc((void **)&a->rf, (size_t)fbase_of(g, 3) * sizeof(double));
  }
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
  for (double *p : {a->ckpt, a->tape, a->work, a->R, a->geta, a->rf})
    if (p) (void)hipFree(p);
  delete a;
  return ok();
}

extern "C" int omhd_adjoint_forward(omhd_adjoint *a, const double *dt_in, const double *ch_in, double *dt_out,
                                    double *ch_out) {
  if (!a) return set_error(OMHD_ERR_ARG, "null adjoint");
  omhd_handle *h = a->h;
  if (!h->have_state) return set_error(OMHD_ERR_ARG, "omhd_upload first");
  if ((dt_in == nullptr) != (ch_in == nullptr)) return set_error(OMHD_ERR_ARG, "give both dt and ch, or neither");
  a->have_forward = a->have_backward = false;
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

// (dD/dW at W)^T mu, mu = mscale * M (slot 8 also times mpsi), combined into
// fin.out (psi of W scaled by psi_f on load)
static int vjp_launch(omhd_adjoint *a, const double *W, double psi_f, const double *M, double mscale, double mpsi,
                      double ch, const Combine &fin) {
  omhd_handle *h = a->h;
  const Geo g = geo_of(h);
  const int lim = h->cfg.lm_type, res = h->cfg.resistive ? 1 : 0;
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  const unsigned g256 = (unsigned)((n + 255) / 256), g128 = (unsigned)((n + 127) / 128);
  bool mir = false;  // any wall axis
  for (int d = 0; d < g.nd; d++) mir = mir || g.bc[d] != OMHD_BC_PERIODIC;
  k_adj_prim<<<g256, 256, 0, h->stream>>>(g, W, psi_f, a->w());
  OMHD_HIP(hipGetLastError());
  for (int dir = 0; dir < g.nd; dir++) {
    // a face kernel's stencil shifts along its own axis only: the mirror
    // variant is needed where that axis is a wall
    hipLaunchKernelGGL(face_kernel(dir, 0, g.bc[dir] != OMHD_BC_PERIODIC, lim), dim3(g128), dim3(128), 0, h->stream, g,
                       a->w(), M, mscale, mpsi, ch, a->AL(dir), a->AR(dir));
    OMHD_HIP(hipGetLastError());
    if (g.bc[dir] != OMHD_BC_PERIODIC) {
      const long long nt = (long long)n / (dir == 0 ? g.nx : (dir == 1 ? g.ny : g.nz));
      hipLaunchKernelGGL(face_kernel(dir, 1, true, lim), dim3((unsigned)((nt + 127) / 128)), dim3(128), 0, h->stream,
                         g, a->w(), M, mscale, mpsi, ch, a->AL(dir), a->AR(dir));
      OMHD_HIP(hipGetLastError());
    }
    if (res) {
      hipLaunchKernelGGL(res_kernel(dir), dim3(g256), dim3(256), 0, h->stream, g, a->w(), M, mscale, mpsi, a->rf);
      OMHD_HIP(hipGetLastError());
    }
  }
  if (res) {
    k_adj_res_gather<<<g256, 256, 0, h->stream>>>(g, a->rf, a->R);
    OMHD_HIP(hipGetLastError());
    const long long ne = 3LL * (g.nx + 2) * (g.ny + 2);
    k_adj_eta<<<(unsigned)((ne + 255) / 256), 256, 0, h->stream>>>(g, a->rf, a->geta);
    OMHD_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(gather_kernel(g.nd, mir, lim), dim3(g256), dim3(256), 0, h->stream, g,
                     a->w(), a->AL(), a->AR(), (long long)a->state, res ? a->R : nullptr, fin);
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
  if (a->geta) OMHD_HIP(hipMemsetAsync(a->geta, 0, a->neta() * sizeof(double), h->stream));
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
  if (err[0])  // the low byte is the OMHD_ERR_* code; a higher bit marks the stage-2 state
    return set_error(err[0] & 0xff, "invalid state while replaying the forward pass", err[1], err[2], err[3]);
  if (a->geta) {
    a->geta_host.resize(a->neta());
    OMHD_HIP(hipMemcpy(a->geta_host.data(), a->geta, a->neta() * sizeof(double), hipMemcpyDeviceToHost));
  }
  a->have_backward = true;
  return ok();
}

extern "C" int omhd_adjoint_eta_grad(omhd_adjoint *a, double *grad) {
  if (!a || !grad) return set_error(OMHD_ERR_ARG, "null argument");
  if (!a->h->cfg.resistive) return set_error(OMHD_ERR_ARG, "the step has no resistivity");
  if (!a->have_backward) return set_error(OMHD_ERR_ARG, "omhd_adjoint_backward first");
  const omhd_config &c = a->h->cfg;
  // eta = eta0 + eta1 exp(-r^2/w^2) at the face centres, placed as the
  // forward places them (cell centres x0 - dx/2 + i dx, accumulated)
  const int ix = c.nx + 2, jx = c.ny + 2;
  std::vector<double> x(ix), y(jx);
  x[0] = c.x0 - c.dx / 2;
  for (int i = 1; i < ix; i++) x[i] = x[i - 1] + c.dx;
  y[0] = c.y0 - c.dx / 2;
  for (int j = 1; j < jx; j++) y[j] = y[j - 1] + c.dx;
  double g0 = 0, g1 = 0, gw = 0;
  for (int d = 0; d < 3; d++)
    for (int j = 0; j < jx; j++)
      for (int i = 0; i < ix; i++) {
        const double ge = a->geta_host[(size_t)d * ix * jx + (size_t)i + (size_t)ix * j];
        if (ge == 0.0) continue;
        const double xx = x[i] + (d == 0 ? 0.5 * c.dx : 0.0), yy = y[j] + (d == 1 ? 0.5 * c.dx : 0.0);
        const double r2 = xx * xx + yy * yy, w2 = c.eta_w * c.eta_w;
        const double ex = exp(-r2 / w2);
        g0 += ge;
        g1 += ge * ex;
        gw += ge * c.eta1 * ex * 2.0 * r2 / (w2 * c.eta_w);
      }
  grad[0] = g0;
  grad[1] = g1;
  grad[2] = gw;
  return ok();
}
