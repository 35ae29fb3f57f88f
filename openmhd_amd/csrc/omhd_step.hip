// omhd_step.hip -- the resident step: fused march kernels + handle API.
//
// 2D (2D_basic/main.f90:56-188): one TVD-RK2 step is two launches,
//   stage 1: U* = U + dt/dx * (dF + dG)              (U with psi *= f on load)
//   stage 2: U' = f_psi * 0.5 * ((U + U*) + dt/dx * (dF + dG)), plus the CFL
//            speed of U' (the next step's set_dt), reduced to one max,
// where every flux difference comes from u2v -> limiter -> HLLD on the fly.
// 3D (3D_basic/main.f90, 3D_reconnection_gpu/main.cuf): each stage is an XY
// pass that writes the partial difference ((F(i-1)-F(i)) + G(j-1)) - G(j) to a
// scratch state, then a Z pass that adds H(k-1) - H(k) in the reference's
// order and applies the RK combine (+ GLM, + CFL).  The resistive flux
// (flux_resistive.cuf:86-135) is added at each interface when configured.
//
// March: one 64-lane wavefront owns a strip of 60 interior columns (lanes
// 2..61; lanes 0,1,62,63 carry the two-cell x halo) and a chunk of rows of the
// march axis (y, or z), at one fixed index of the remaining axis.  It walks
// the chunk, loading each row of the state once (coalesced 512-byte segments
// per variable), keeps a three-row ring of primitives in LDS, computes the
// march-axis slopes and interface flux, and (XY passes) the x-slopes and
// x-interface fluxes of the oldest row.  Boundary conditions are index wraps
// (periodic) or mirrors with sign tables; wall face states are mirrors of the
// inside face state, as the reference's face patches make them
// (main.f90:101-102,119-120; 2D_KH/bc.f90:42-83).
//
// State layout in HBM: 9 planes (SoA), interior cells only, x fastest, rows of
// `pitch` doubles, z-planes of `zstride` doubles:
//   state[m * plane + k * zstride + j * pitch + i].
#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "omhd_device.h"
#include "omhd_host.h"
#include "omhd_internal.h"

namespace omhd {
namespace {

constexpr int STRIP = 60;  // interior columns per wavefront (64 lanes - 2x2 halo)

enum Pass { PASS_2D = 0, PASS_XY = 1, PASS_Z = 2 };

struct Sweep {
  const double *us;   // state whose fluxes are computed (stage 1: U^n, stage 2: U*)
  const double *u0;   // stage 2: U^n, the RK partner
  double *out;        // stage 1: U*, stage 2: U^{n+1} (may alias u0)
  double *acc;        // 3D: partial flux difference written by XY, read by Z
  long long plane;    // doubles between variable planes
  long long zstride;  // doubles between z-planes
  int pitch;          // doubles between rows
  int nx, ny, nz;
  int nstrip, rows, nchunk, ntiles;
  int bc[3];                  // OMHD_BC_* per axis (both faces)
  unsigned slo[3], shi[3];    // negated slots of mirror faces, per axis
  double dtx, glm_f, ch, ch2, dx;
  const double *eta;          // resistive: 3 face arrays (ix*jx each) or null
  unsigned long long *vmax_bits;  // stage 2: max speed of U^{n+1}
  int *err;                       // [code, i, j, k] of the first failure in a state read
  int *err_next;                  // stage 2, in U^{n+1}: [positivity code, non-finite speed]
  const double *dpar;  // device-side step parameters (DevPar) or null: use the fields above
  double *U, *U1;      // ensemble members: their two state buffers
  int ftype;           // flux type of the generic-solver kernels (0 LLF, 1 HLL, 2 HLLC, 4 LHLLD)
};

// Step parameters computed on the device from the reduced CFL speed (device
// dt mode: no host round trip, so whole steps can be captured in a graph).
struct DevPar {
  double ch, ch2, dtx1, dtx2, glm_f, dt;
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

// negate slot m when bit m of mask is set: a sign-bit XOR (exact, like the
// reference's U(:,1,my) = -U(:,2,my))
__device__ __forceinline__ double flip(double v, unsigned mask, int m) {
  const long long sbit = (long long)((mask >> m) & 1u) << 63;
  return __longlong_as_double(__double_as_longlong(v) ^ sbit);
}

// Failures found while a stage reads its state carry ERR_IN_USTAR when that
// state is U* (stage 2), so the host knows which buffer failed.
constexpr int ERR_IN_USTAR = 0x200;

__device__ __forceinline__ void record_error(int *err, int code, int i, int j, int k) {
  if (atomicCAS(err, 0, code) == 0) {
    err[1] = i;
    err[2] = j;
    err[3] = k;
  }
}

// Per wavefront LDS: a ring of three primitive rows (rn-2, rn-1, rn) and the
// carried left face state of the march-axis interface above row rn-2; for the
// resistive flux also the transverse B samples of each ring row and Bx of the
// row below the ring.  Keeping these out of VGPRs leaves ~18 live doubles
// across each inlined HLLD call, so two wavefronts fit per SIMD.
struct SweepLds {
  double ring[3][NV][66];  // [slot][var][lane + 1]; entries 0 and 65 repeat lanes 0 and 63
  double wly[NV][64];
};
// The resistive passes keep the out-of-plane current terms of each row
// already combined (same operations in the same order as the face formula,
// so the same bits): xo the x-face term of each ring row, mo the march-axis
// face term of the last two faces, prv the previous row's out-of-plane
// samples.  XY pass: 22976 B (7 wavefronts per CU).
struct SweepLdsRes {
  double ring[3][NV][66];
  double wly[NV][64];
  double xo[3][64];   // 0.25 (Bx(k+1)_c + Bx(k+1)_c+1 - Bx(k-1)_c - Bx(k-1)_c+1)
  double mo[2][64];   // 0.25 (By(k+1)_r-1 + By(k+1)_r - By(k-1)_r-1 - By(k-1)_r)
  double prv[2][64];  // By(k-1), By(k+1) of the previous row
  double bxm[66];     // Bx of row rn-3
};
// Z pass: no x-sweep, so the ring needs no halo entries (the Bz neighbours of
// the z-face current come over DPP); 20480 B, 8 wavefronts per CU.
struct SweepLdsResZ {
  double ring[3][NV][64];
  double wly[NV][64];
  double mo[2][64];   // 0.25 (Bz(j+1)_r-1 + Bz(j+1)_r - Bz(j-1)_r-1 - Bz(j-1)_r)
  double prv[2][64];  // Bz(j-1), Bz(j+1) of the previous row
};

// Element of variable m at this lane's column of a row: `ro` is the row's
// offset (uniform across the wavefront) and off8 the lane's byte offset in
// the row.  The uniform part stays in scalar registers; the compiler adds the
// lane's offset once per row and variable.
__device__ __forceinline__ const double &lane_at(const double *a, long long plane, int m, long long ro,
                                                 unsigned off8) {
  return *(const double *)((const char *)(a + (m * plane + ro)) + off8);
}
__device__ __forceinline__ double &lane_at(double *a, long long plane, int m, long long ro, unsigned off8) {
  return *(double *)((char *)(a + (m * plane + ro)) + off8);
}

// Store this lane's value of a 66-entry LDS row: entry lane + 1, and again
// at `edge`, the lane's second entry (0 for lane 0, 65 for lane 63, lane + 1
// itself for the others).  Two plain stores per lane instead of lane tests
// and divergent branches.
__device__ __forceinline__ int edge66(int lane) { return lane == 0 ? 0 : (lane == 63 ? 65 : lane + 1); }
__device__ __forceinline__ void put66(double *row, int lane, int edge, double v) {
  row[lane + 1] = v;
  row[edge] = v;
}

template <int RES, int PASS>
using LdsOf = typename std::conditional<
    RES != 0, typename std::conditional<PASS == PASS_Z, SweepLdsResZ, SweepLdsRes>::type, SweepLds>::type;

// One tile: strip x chunk (x slice) of one grid.
// PREF: stage 2 issues the U^n loads of the combined row early (where the
// register budget allows it without spills: the resident kernels, not the
// ensemble kernel, whose per-member parameters take more registers).
// GEN: the flux is p.ftype's (LLF, HLL, HLLC, LHLLD) instead of HLLD.
template <int LIM, int STAGE, int TT, int PASS, int RES, bool PREF = true, bool GEN = false>
__device__ __forceinline__ void march_tile(const Sweep &p, const int t, LdsOf<RES, PASS> &lds) {
  constexpr bool XS = PASS != PASS_Z;         // x-sweep of the oldest row
  constexpr bool FINAL = PASS != PASS_XY;     // RK update (else write acc)
  constexpr int MD = PASS == PASS_Z ? 2 : 1;  // march axis = direction of its HLLD
  constexpr int ND = PASS == PASS_2D ? 2 : 3;
  const int lane = threadIdx.x;
  const int edge = edge66(lane);
  constexpr int RO = (RES != 0 && PASS == PASS_Z) ? 0 : 1;  // first own entry of a ring row
  const int strip = t % p.nstrip;
  const int r2 = t / p.nstrip;
  const int chunk = r2 % p.nchunk, slice = r2 / p.nchunk;  // slice: k (XY) or j (Z)
  const int c = strip * STRIP - 2 + lane;
  const int nm = MD == 1 ? p.ny : p.nz;  // cells along the march axis
  const int j0 = chunk * p.rows;
  const int hc = min(p.rows, nm - j0);
  const long long rstride = MD == 1 ? (long long)p.pitch : p.zstride;
  const long long fixoff = PASS == PASS_Z ? (long long)slice * p.pitch : (long long)slice * p.zstride;
  int cm;
  unsigned xmask;
  map_axis(c, p.nx, p.bc[0], p.slo[0], p.shi[0], cm, xmask);
  const bool writer = lane >= 2 && lane < 2 + STRIP && c < p.nx;
  const bool xlo_face = p.bc[0] == OMHD_BC_MIRROR && c == -1;
  const bool xhi_face = p.bc[0] == OMHD_BC_MIRROR && c == p.nx - 1;
  double ch = p.ch, ch2 = p.ch2, dtx = p.dtx, glm_f = p.glm_f;
  if (p.dpar) {
    const DevPar *dp = (const DevPar *)p.dpar;
    if (!(dp->dt > 0.0)) return;  // device dt mode halted by a stop: the state stays as it was
    ch = dp->ch;
    ch2 = dp->ch2;
    dtx = STAGE == 1 ? dp->dtx1 : dp->dtx2;
    glm_f = dp->glm_f;
  }
  const double glm_us = (STAGE == 1) ? glm_f : 1.0;
  const long long plane = p.plane;
  const int ix = p.nx + 2, jx = p.ny + 2;

  double Gl[NV], nxt[NV];
  double vloc = 0.0;

  const unsigned cm8 = (unsigned)cm * 8u;  // this lane's byte offset in a row
  // offset of the first cell of (mapped) row rn, uniform across the wavefront
  auto row_off = [&](int rn, unsigned &mask) -> long long {
    int rm;
    unsigned ymask;
    map_axis(rn, nm, p.bc[MD], p.slo[MD], p.shi[MD], rm, ymask);
    mask = xmask ^ ymask;
    return fixoff + (long long)rm * rstride;
  };

  unsigned mask_nxt;
  {
    const long long ro = row_off(j0 - 2, mask_nxt);
#pragma unroll
    for (int m = 0; m < NV; m++) nxt[m] = lane_at(p.us, plane, m, ro, cm8);
  }
  // resistive XY pass: Bx and By of row r in planes k-1 and k+1 (as loaded,
  // before the ghost signs), prefetched one row ahead like the row itself
  auto aux_load = [&](int r, double a[4]) {
    int rm;
    unsigned ym;
    map_axis(r, nm, p.bc[MD], p.slo[MD], p.shi[MD], rm, ym);
#pragma unroll
    for (int s2 = 0; s2 < 2; s2++) {
      int km;
      unsigned zm;
      map_axis(slice + (s2 ? 1 : -1), p.nz, p.bc[2], p.slo[2], p.shi[2], km, zm);
      const long long ro = (long long)km * p.zstride + (long long)rm * p.pitch;
      a[s2] = lane_at(p.us, plane, 5, ro, cm8);
      a[2 + s2] = lane_at(p.us, plane, 6, ro, cm8);
    }
  };
  double auxn[4] = {0.0, 0.0, 0.0, 0.0};
  if constexpr (RES != 0 && PASS == PASS_XY) aux_load(j0 - 2, auxn);

  // resistive Z pass: the z-face eta/dx of this lane's column and the tile's
  // slice, the same for every row of the march
  double etz_dx = 0.0;
  if constexpr (RES != 0 && PASS == PASS_Z)
    etz_dx = p.eta[2 * (size_t)ix * jx + (size_t)(cm + 1) + (size_t)ix * (slice + 1)] / p.dx;
  for (int k = 0; k < hc + 4; ++k) {
    const int rn = j0 - 2 + k;
    const int sa = (k + 1) % 3, sb = (k + 2) % 3, sc = k % 3;  // rows rn-2, rn-1, rn
    // resistive XY pass: this iteration's x- and y-face eta/dx, loaded here so
    // that their latency hides behind the row's work.  The row index is
    // clamped into the table (the warm-up iterations do not use the values),
    // and lanes outside the face range read a valid entry.
    double etx_dx = 0.0, ety_dx = 0.0;
    if constexpr (RES != 0 && PASS == PASS_XY) {
      const int re = min(max(rn - 1, 0), jx - 1);
      const int cf = min(max(c + 1, 0), ix - 1);
      etx_dx = p.eta[(size_t)cf + (size_t)ix * re] / p.dx;
      ety_dx = p.eta[(size_t)ix * jx + (size_t)(cm + 1) + (size_t)ix * re] / p.dx;
    }
    {
      double u[NV], W[NV];
      if (__any(mask_nxt != 0u)) {  // a ghost row or column with negated slots (mirror faces)
#pragma unroll
        for (int m = 0; m < NV; m++) u[m] = flip(nxt[m], mask_nxt, m);
      } else {
#pragma unroll
        for (int m = 0; m < NV; m++) u[m] = nxt[m];
      }
      u[8] = glm_us * u[8];  // glm_ss2 before the sweep (main.f90:87)
      double auxc[4];
#pragma unroll
      for (int q = 0; q < 4; q++) auxc[q] = auxn[q];
      if constexpr (RES != 0 && PASS == PASS_XY) {
        if (k + 1 < hc + 4) aux_load(rn + 1, auxn);
      }
      if (k + 1 < hc + 4) {  // prefetch the next row
        unsigned mnext;
        const long long ro = row_off(rn + 1, mnext);
#pragma unroll
        for (int m = 0; m < NV; m++) nxt[m] = lane_at(p.us, plane, m, ro, cm8);
        mask_nxt = mnext;
      }
      cons_to_prim(u, W);  // u2v
      if (!(W[3] > 0.0) || !(u[4] > 0.0)) {
        int rm;
        unsigned dummy;
        map_axis(rn, nm, p.bc[MD], p.slo[MD], p.shi[MD], rm, dummy);
        const int jj = MD == 1 ? rm : slice;
        const int kk = MD == 1 ? (ND == 3 ? slice + 2 : 0) : rm + 2;
        record_error(p.err, ((W[3] > 0.0) ? OMHD_ERR_DENSITY : OMHD_ERR_PRESSURE) | (STAGE == 2 ? ERR_IN_USTAR : 0),
                     cm + 2, jj + 2, kk);
      }
      if constexpr (RES != 0) {
        // transverse B samples of this row for the resistive face currents
        int rm;
        unsigned ymask;
        map_axis(rn, nm, p.bc[MD], p.slo[MD], p.shi[MD], rm, ymask);
        if constexpr (PASS == PASS_XY) {
          put66(lds.bxm, lane, edge, lds.ring[sc][5][lane + 1]);  // Bx of row rn-3, before its slot is reused
          double bx[2], by[2];  // planes k-1 (0) and k+1 (1), ghost signs applied
#pragma unroll
          for (int s2 = 0; s2 < 2; s2++) {
            int km;
            unsigned zmask;
            map_axis(slice + (s2 ? 1 : -1), p.nz, p.bc[2], p.slo[2], p.shi[2], km, zmask);
            const unsigned mk = xmask ^ ymask ^ zmask;
            bx[s2] = flip(auxc[s2], mk, 5);
            by[s2] = flip(auxc[2 + s2], mk, 6);
          }
          // x-face (c+1/2) term of row rn: flux_resistive.cuf:89-91's k-differences of Bx
          lds.xo[sc][lane] = 0.25 * (bx[1] + from_next(bx[1]) - bx[0] - from_next(bx[0]));
          // y-face (rn-1/2) term: flux_resistive.cuf:106-108's k-differences of By
          lds.mo[k & 1][lane] = 0.25 * (lds.prv[1][lane] + by[1] - lds.prv[0][lane] - by[0]);
          lds.prv[0][lane] = by[0];
          lds.prv[1][lane] = by[1];
        } else {
          double bz[2];  // slices j-1 (0) and j+1 (1), ghost signs applied
#pragma unroll
          for (int s2 = 0; s2 < 2; s2++) {
            int jm;
            unsigned jmask;
            map_axis(slice + (s2 ? 1 : -1), p.ny, p.bc[1], p.slo[1], p.shi[1], jm, jmask);
            const unsigned mk = xmask ^ ymask ^ jmask;
            const long long ro = (long long)rm * p.zstride + (long long)jm * p.pitch;
            bz[s2] = flip(lane_at(p.us, plane, 7, ro, cm8), mk, 7);
          }
          // z-face (rn-1/2) term: flux_resistive.cuf:123-125's j-differences of Bz
          lds.mo[k & 1][lane] = 0.25 * (lds.prv[1][lane] + bz[1] - lds.prv[0][lane] - bz[0]);
          lds.prv[0][lane] = bz[0];
          lds.prv[1][lane] = bz[1];
        }
      }
#pragma unroll
      for (int m = 0; m < NV; m++) {
        if constexpr (RO == 0) lds.ring[sc][m][lane] = W[m];
        else put66(lds.ring[sc][m], lane, edge, W[m]);
      }
    }
    __syncthreads();  // one wavefront per workgroup: orders the ring writes before the neighbour reads
    // stage 2: issue the loads of U^n for the row combined at the end of this
    // iteration now, so their latency hides behind the two flux evaluations
    double a0p[NV], usp[NV];
    // TVD stage 2 folds U^n + U* (the first sum of rk_tvd22) before the
    // march-axis flux, so the U* loads go out as one batch
    constexpr bool FOLD = FINAL && STAGE == 2 && TT == 0 && PREF;
    // row rn-2, combined in this iteration: its offset (a writer lane's column c equals cm)
    const long long ro2 = fixoff + (long long)(rn - 2) * rstride;
    if constexpr (FINAL && STAGE == 2 && PREF) {
      if (k >= 4 && writer) {
#pragma unroll
        for (int m = 0; m < NV; m++) a0p[m] = lane_at(p.u0, plane, m, ro2, cm8);
      }
    }
    if constexpr (FINAL && STAGE == 1 && PREF) {
      if (k >= 4 && writer) {
#pragma unroll
        for (int m = 0; m < NV; m++) {
          // slots 4-8 (rho, B, psi) of a primitive row are the conserved
          // ones (psi with the glm factor of stage 1 applied), and the row
          // combined here is an interior row of the LDS ring: only slots
          // 0-3 are read again from memory
          if (m < 4) usp[m] = lane_at(p.us, plane, m, ro2, cm8);
          else usp[m] = lds.ring[sa][m][lane + RO];
        }
      }
    }
    // Z pass, stage 1, ideal MHD: the XY pass's partial flux difference for
    // the same row (elsewhere the registers for it spill)
    constexpr bool ZPRE = PASS == PASS_Z && PREF && RES == 0 && STAGE == 1;
    double azp[NV];
    if constexpr (ZPRE) {
      if (k >= 4 && writer) {
#pragma unroll
        for (int m = 0; m < NV; m++) azp[m] = lane_at((const double *)p.acc, plane, m, ro2, cm8);
      }
    }

    // x-sweep of row r = rn-2: tq = (F(i-1) - F(i)) + G(j-1)   (rk.f90:30-31 order)
    double tq[NV];
    if constexpr (XS) {
      if (k >= 4) {
        double Lx[NV], Rx[NV], F[NV];
        {
          // all 27 neighbour reads in flight before the first slope: left to
          // itself the scheduler issues them three at a time and waits on each
          // group (one LDS latency per variable)
          double x0[NV], x1[NV], x2[NV];
#pragma unroll
          for (int m = 0; m < NV; m++) {
            x0[m] = lds.ring[sa][m][lane];
            x1[m] = lds.ring[sa][m][lane + 1];
            x2[m] = lds.ring[sa][m][lane + 2];
          }
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int m = 0; m < NV; m++) {
            double wRx;
            limit<LIM>(x0[m], x1[m], x2[m], wRx, Lx[m]);
            Rx[m] = from_next(wRx);
          }
        }
        if (xlo_face) {
#pragma unroll
          for (int m = 0; m < NV; m++) Lx[m] = flip(Rx[m], p.slo[0], m);
        }
        if (xhi_face) {
#pragma unroll
          for (int m = 0; m < NV; m++) Rx[m] = flip(Lx[m], p.shi[0], m);
        }
        if constexpr (GEN) riemann_other<0>(p.ftype, Lx, Rx, ch, ch2, F);
        else hlld<0>(Lx, Rx, ch, ch2, F);
        if constexpr (RES != 0) {  // flux_resistive.cuf:89-101 at x-face (c+1/2, row rn-2, plane slice)
          const int l = lane + 1;
          const double *A6 = lds.ring[sa][6], *A7 = lds.ring[sa][7];
          const double *B5 = lds.ring[sb][5];
          const double JyS = -(A7[l + 1] - A7[l]) + lds.xo[sa][lane];
          const double JzS = (A6[l + 1] - A6[l]) - 0.25 * (B5[l] + B5[l + 1] - lds.bxm[l] - lds.bxm[l + 1]);
          const double eta_dx = etx_dx;
          F[3] = F[3] + 0.5 * eta_dx * (JyS * (A7[l] + A7[l + 1]) - JzS * (A6[l] + A6[l + 1]));
          F[6] = F[6] - eta_dx * JzS;
          F[7] = F[7] + eta_dx * JyS;
        }
#pragma unroll
        for (int m = 0; m < NV; m++) tq[m] = from_prev(F[m]) - F[m] + Gl[m];
      }
    }

    if (k >= 2) {
      // march-axis slopes of row rn-1 from rows rn-2, rn-1, rn
      double Lf[NV], Rf[NV];
      double y0[NV];  // row rn-2 as the ring holds it (also stage 2's fold below)
      {
        double y1[NV], y2[NV];  // likewise: every own-lane read before the first slope
#pragma unroll
        for (int m = 0; m < NV; m++) {
          y0[m] = lds.ring[sa][m][lane + RO];
          y1[m] = lds.ring[sb][m][lane + RO];
          y2[m] = lds.ring[sc][m][lane + RO];
          Lf[m] = lds.wly[m][lane];  // left state of the interface above row rn-2
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int m = 0; m < NV; m++) {
          double wLyB;
          limit<LIM>(y0[m], y1[m], y2[m], Rf[m], wLyB);
          lds.wly[m][lane] = wLyB;
        }
      }
      if constexpr (FOLD) {
        if (k >= 4 && writer) {
          double usf[NV];  // U* of row rn-2: slots 4-8 as the ring holds them (y0)
#pragma unroll
          for (int m = 0; m < NV; m++) usf[m] = m < 4 ? lane_at(p.us, plane, m, ro2, cm8) : y0[m];
#pragma unroll
          for (int m = 0; m < NV; m++) a0p[m] = (m == 8 ? glm_f * a0p[m] : a0p[m]) + usf[m];
        }
      }
      if (k >= 3) {
        // march-axis interface between rows rn-2 and rn-1
        if (p.bc[MD] == OMHD_BC_MIRROR) {
          if (rn - 1 == 0) {  // low wall: bc_for_G, 2D_KH/bc.f90:62-70
#pragma unroll
            for (int m = 0; m < NV; m++) Lf[m] = flip(Rf[m], p.slo[MD], m);
          }
          if (rn - 1 == nm) {  // high wall: 2D_KH/bc.f90:73-81
#pragma unroll
            for (int m = 0; m < NV; m++) Rf[m] = flip(Lf[m], p.shi[MD], m);
          }
        }
        double G[NV];
        if constexpr (GEN) riemann_other<MD>(p.ftype, Lf, Rf, ch, ch2, G);
        else hlld<MD>(Lf, Rf, ch, ch2, G);
        if constexpr (RES != 0) {
          const int l = lane + RO;
          const double *A5 = lds.ring[sa][5], *A6 = lds.ring[sa][6], *A7 = lds.ring[sa][7];
          const double *B5 = lds.ring[sb][5], *B6 = lds.ring[sb][6], *B7 = lds.ring[sb][7];
          if constexpr (PASS == PASS_XY) {  // flux_resistive.cuf:106-118, y-face (c, rn-2 + 1/2, slice)
            const double JxS = (B7[l] - A7[l]) - lds.mo[(k + 1) & 1][lane];
            const double JzS = -(B5[l] - A5[l]) + 0.25 * (A6[l + 1] + B6[l + 1] - A6[l - 1] - B6[l - 1]);
            const double eta_dx = ety_dx;
            G[3] = G[3] + 0.5 * eta_dx * (JzS * (A5[l] + B5[l]) - JxS * (A7[l] + B7[l]));
            G[5] = G[5] + eta_dx * JzS;
            G[7] = G[7] - eta_dx * JxS;
          } else {  // flux_resistive.cuf:123-135, z-face (c, slice, rn-2 + 1/2)
            const double a7 = A7[l], b7 = B7[l];  // column neighbours over DPP (the edge lanes write nothing)
            const double JyS = (B5[l] - A5[l]) - 0.25 * (from_next(a7) + from_next(b7) - from_prev(a7) - from_prev(b7));
            const double JxS = -(B6[l] - A6[l]) + lds.mo[(k + 1) & 1][lane];
            const double eta_dx = etz_dx;
            G[3] = G[3] + 0.5 * eta_dx * (JxS * (A6[l] + B6[l]) - JyS * (A5[l] + B5[l]));
            G[5] = G[5] - eta_dx * JyS;
            G[6] = G[6] + eta_dx * JxS;
          }
        }

        if (k >= 4 && writer) {
          double d[NV];
          if constexpr (PASS == PASS_2D) {
#pragma unroll
            for (int m = 0; m < NV; m++) d[m] = tq[m] - G[m];
          } else if constexpr (PASS == PASS_XY) {
#pragma unroll
            for (int m = 0; m < NV; m++) lane_at(p.acc, plane, m, ro2, cm8) = tq[m] - G[m];  // ((F-F)+G)-G
          } else {
#pragma unroll
            for (int m = 0; m < NV; m++) {
              double a;
              if constexpr (ZPRE) a = azp[m];
              else a = lane_at((const double *)p.acc, plane, m, ro2, cm8);
              d[m] = a + Gl[m] - G[m];  // (.. + H(k-1)) - H(k)
            }
          }
          if constexpr (FINAL) {
            double o9[NV];
            if constexpr (STAGE == 1) {
#pragma unroll
              for (int m = 0; m < NV; m++) {
                double uu;
                if constexpr (PREF) {
                  uu = usp[m];  // slot 8 from the ring: glm_f * psi already
                } else if (m >= 4) {
                  uu = lds.ring[sa][m][lane + RO];  // as above: slots 4-8 from the ring
                } else {
                  uu = lane_at(p.us, plane, m, ro2, cm8);
                }
                o9[m] = uu + dtx * d[m];  // rk_tvd21 / rk_std21
              }
            } else {
#pragma unroll
              for (int m = 0; m < NV; m++) {
                if constexpr (FOLD) {
                  o9[m] = 0.5 * (a0p[m] + dtx * d[m]);  // rk_tvd22: a0p holds U^n + U*
                } else {
                  double a0;
                  if constexpr (PREF) a0 = a0p[m];
                  else a0 = lane_at(p.u0, plane, m, ro2, cm8);
                  if (m == 8) a0 = glm_f * a0;
                  if constexpr (TT == 0) {  // rk_tvd22; U* slots 4-8 from the ring
                    const double us = m >= 4 ? lds.ring[sa][m][lane + RO] : lane_at(p.us, plane, m, ro2, cm8);
                    o9[m] = 0.5 * (a0 + us + dtx * d[m]);
                  } else {
                    o9[m] = a0 + dtx * d[m];  // rk_std22
                  }
                }
              }
              o9[8] = glm_f * o9[8];  // glm_ss2 after the step (main.f90:185)
            }
#pragma unroll
            for (int m = 0; m < NV; m++) lane_at(p.out, plane, m, ro2, cm8) = o9[m];
            if constexpr (STAGE == 2) {  // next step's u2v and set_dt on U^{n+1}
              double w2[NV];
              cons_to_prim(o9, w2);
              if (!(w2[3] > 0.0) || !(o9[4] > 0.0))
                (void)atomicCAS(p.err_next, 0, (w2[3] > 0.0) ? OMHD_ERR_DENSITY : OMHD_ERR_PRESSURE);
              const double vt = fastest_speed<ND>(o9, w2);
              vloc = (vt > vloc || !(vt == vt)) ? vt : vloc;
            }
          }
        }
#pragma unroll
        for (int m = 0; m < NV; m++) Gl[m] = G[m];
      }
    }
  }
  if constexpr (FINAL && STAGE == 2) {
    if (!(vloc >= 0.0)) (void)atomicCAS(p.err_next + 1, 0, OMHD_ERR_DT);
    const double vm = wave_max(vloc >= 0.0 ? vloc : 0.0);
    if (lane == 0) atomic_max_nonneg(p.vmax_bits, vm);
  }
}

template <int LIM, int STAGE, int TT, int PASS, int RES>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_march(const Sweep p) {
  __shared__ LdsOf<RES, PASS> lds;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  if (t >= p.ntiles) return;
  if (STAGE == 2 && p.err[0] != 0) return;  // stage 1 failed: leave U^n as it is
  march_tile<LIM, STAGE, TT, PASS, RES, (LIM != 0)>(p, t, lds);
}

// Ideal-MHD step (2D, or the 3D XY and Z passes) with one of the other flux
// types (p.ftype), same structure
template <int LIM, int STAGE, int TT, int PASS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_march_gen(const Sweep p) {
  __shared__ LdsOf<0, PASS_2D> lds;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  if (t >= p.ntiles) return;
  if (STAGE == 2 && p.err[0] != 0) return;  // stage 1 failed: leave U^n as it is
  march_tile<LIM, STAGE, TT, PASS, 0, false, true>(p, t, lds);
}

// Ensemble: one persistent launch per RK stage for every member.  The tile
// table lists (member, tile) pairs; each workgroup takes every gridDim-th entry.
template <int LIM, int STAGE, int TT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_march_ens(
    const Sweep *__restrict__ members, const int2 *__restrict__ tiles, int ntiles) {
  __shared__ LdsOf<0, PASS_2D> lds;
  for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
    const int2 tl = tiles[w];
    Sweep p = members[tl.x];
    p.us = STAGE == 1 ? p.U : p.U1;
    p.u0 = p.U;
    p.out = STAGE == 1 ? p.U1 : p.U;
    march_tile<LIM, STAGE, TT, PASS_2D, 0, false>(p, tl.y, lds);
    __syncthreads();
  }
}

// GLM factors of the device-dt mode.  With dt = (cfl*dx)/ch from the CFL
// speed ch, the argument x = ((-0.5*dt)*ch)/0.2 of glm_ss2.f90:27 lies within
// a few ulps of x0 = (-0.5*(cfl*dx))/0.2, since fl(dt*ch) is within two ulps
// of cfl*dx and the scaling by -0.5 is exact.  The host tabulates its libm's
// exp at the GLM_TAB_HALF doubles either side of x0 per member, and the
// device looks the factor up by the bit offset of its x from x0: the same
// bits as the host-dt path.
constexpr int GLM_TAB_HALF = 64, GLM_TAB = 2 * GLM_TAB_HALF + 1;

// Device dt: for each of n members (or only member `only` when it is >= 0),
// dt and the GLM factor from its reduced CFL speed (set_dt.f90:52,
// glm_ss2.f90:27, flux_resistive.cuf:8), then the speed slot is cleared for
// the next step's reduction.  The GLM factor comes from the member's table of
// host exp values (glm_tab, glm_x0: GLM_TAB doubles and the bits of x0 per
// member); an argument outside the table, which a CFL-limited dt cannot
// produce, falls back to the device's exp.
__global__ void k_device_dt(DevPar *par, unsigned long long *vmax_bits, const double *dx, const double *cfl,
                            const int *tt, const double *eta_max, const double *glm_tab, const long long *glm_x0,
                            int *err, int n, int only = -1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (only >= 0 && i != only)) return;
  const double vmax = __longlong_as_double((long long)vmax_bits[i]);
  double dt = cfl[i] * dx[i] / vmax;
  if (!(dt >= 1.0e-7)) {
    record_error(err, OMHD_ERR_DT, i, 0, 0);
    dt = 0.0;
  }
  if (eta_max && eta_max[i] > 0.0) {
    const double dr = 0.5 * cfl[i] * (dx[i] * dx[i]) / eta_max[i];
    dt = dt < dr ? dt : dr;
  }
  DevPar q;
  q.ch = vmax;
  q.ch2 = vmax * vmax;
  q.dtx1 = tt[i] == 0 ? dt / dx[i] : 0.5 * dt / dx[i];
  q.dtx2 = dt / dx[i];
  const double x = -0.5 * dt * vmax / 0.2;
  const long long off = (long long)__double_as_longlong(x) - glm_x0[i];
  q.glm_f = (off >= -GLM_TAB_HALF && off <= GLM_TAB_HALF)
                ? glm_tab[(size_t)i * GLM_TAB + (size_t)(off + GLM_TAB_HALF)]
                : exp(x);
  q.dt = dt;
  par[i] = q;
  vmax_bits[i] = 0ull;
}

// Device dt mode of one handle (omhd_set_mode).  `par` holds the parameters
// of the step about to run (dt == 0: halted by a stop); t and done are the
// time and the count of completed steps of the current call; bad_dt is a dt
// that failed set_dt's test.
struct SoloDev {
  DevPar par;
  double t;
  long long done;
  double bad_dt;
  long long glm_x0;
  double glm_tab[GLM_TAB];
};

__global__ void k_solo_begin(SoloDev *s, unsigned long long *red, double t0) {
  s->t = t0;
  s->done = 0;
  red[0] = 0ull;
}

// After the stages of a step (one thread): the bookkeeping omhd_step does on
// the host, in the same order and arithmetic.  A failure inside the step
// (err) leaves t and done as they were; a failure of the completed step's
// output (next: the following u2v or set_dt would stop) counts the step; a
// dt below set_dt's bound (set_dt.f90:55-65) stops before the next step.
// Any stop sets dt = 0, after which the stage kernels and this kernel return
// at once.  Otherwise the next step's parameters come from the reduced CFL
// speed: dt = (cfl*dx)/ch, dt/dx per stage, and the GLM factor from the table
// of host exp values around x0 (see k_device_dt).
__global__ void k_solo_dt(SoloDev *s, unsigned long long *red, double cfl, double dx, int tt, double *trace) {
  const DevPar q = s->par;
  if (!(q.dt > 0.0)) return;
  int *err = (int *)(red + 1);
  const int *nxt = (const int *)(red + 3);
  if (err[0] != 0) {
    s->par.dt = 0.0;
    return;
  }
  s->t += q.dt;
  if (trace) {
    trace[2 * s->done] = q.dt;
    trace[2 * s->done + 1] = q.ch;
  }
  s->done += 1;
  if (nxt[0] != 0 || nxt[1] != 0) {
    s->par.dt = 0.0;
    return;
  }
  const double vmax = __longlong_as_double((long long)red[0]);
  red[0] = 0ull;
  const double dt = cfl * dx / vmax;
  if (!(dt >= 1.0e-7)) {
    s->bad_dt = dt;
    s->par.dt = 0.0;
    record_error(err, OMHD_ERR_DT, 0, 0, 0);
    return;
  }
  DevPar n;
  n.ch = vmax;
  n.ch2 = vmax * vmax;
  n.dtx1 = tt == 0 ? dt / dx : 0.5 * dt / dx;
  n.dtx2 = dt / dx;
  const double x = -0.5 * dt * vmax / 0.2;
  const long long off = (long long)__double_as_longlong(x) - s->glm_x0;
  n.glm_f = (off >= -GLM_TAB_HALF && off <= GLM_TAB_HALF) ? s->glm_tab[off + GLM_TAB_HALF] : exp(x);
  n.dt = dt;
  s->par = n;
}

// CFL speed of the resident state (set_dt.f90:21-45), used once after upload
template <int ND>
__global__ void k_cfl(const double *__restrict__ s, long long plane, long long zstride, int pitch, int nx, int ny,
                      int nz, unsigned long long *vmax_bits, int *err) {
  // grid-stride: a bounded grid keeps the final atomics to one per wavefront
  // of the grid, not one per 64 cells (they all hit one address)
  const long long n = (long long)nx * ny * nz;
  double vt = 0.0;
  for (long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x; gid < n;
       gid += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(gid % nx);
    const long long r = gid / nx;
    const int j = (int)(r % ny), k = (int)(r / ny);
    const long long o = (long long)k * zstride + (long long)j * pitch + i;
    double u[NV], w[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) u[m] = s[m * plane + o];
    cons_to_prim(u, w);
    const int kk = ND == 3 ? k + 2 : 0;
    if (!(w[3] > 0.0) || !(u[4] > 0.0))
      record_error(err, (w[3] > 0.0) ? OMHD_ERR_DENSITY : OMHD_ERR_PRESSURE, i + 2, j + 2, kk);
    const double c = fastest_speed<ND>(u, w);
    if (!(c >= 0.0)) record_error(err, OMHD_ERR_DT, i + 2, j + 2, kk);
    else vt = dmax(vt, c);
  }
  vt = wave_max(vt);
  if ((threadIdx.x & 63) == 0) atomic_max_nonneg(vmax_bits, vt);
}

// Fortran array U(ix,jx[,kx],9) (+ V(..,4)) with ghosts <- SoA interior

__global__ void k_to_fortran(const double *__restrict__ s, long long plane, long long zstride, int pitch, int nx,
                             int ny, int nz, int nd, Faces f, double *__restrict__ U, double *__restrict__ V) {
  const int ix = nx + 2, jx = ny + 2, kx = nd == 3 ? nz + 2 : 1;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)ix * jx * kx;
  if (gid >= n) return;
  const int i = (int)(gid % ix);
  const long long r = gid / ix;
  const int j = (int)(r % jx), k = (int)(r / jx);
  int im, jm, km = 0;
  unsigned mx_, my_, mz_ = 0;
  map_axis(i - 1, nx, f.bc[0], f.slo[0], f.shi[0], im, mx_);
  map_axis(j - 1, ny, f.bc[1], f.slo[1], f.shi[1], jm, my_);
  if (nd == 3) map_axis(k - 1, nz, f.bc[2], f.slo[2], f.shi[2], km, mz_);
  const unsigned mask = mx_ ^ my_ ^ mz_;
  const long long o = (long long)km * zstride + (long long)jm * pitch + im;
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

__global__ void k_from_fortran(const double *__restrict__ U, double *__restrict__ s, long long plane, long long zstride,
                               int pitch, int nx, int ny, int nz, int nd) {
  const int ix = nx + 2, jx = ny + 2, kx = nd == 3 ? nz + 2 : 1;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)nx * ny * nz) return;
  const int i = (int)(gid % nx);
  const long long r = gid / nx;
  const int j = (int)(r % ny), k = (int)(r / ny);
  const long long n = (long long)ix * jx * kx;
  const long long fk = nd == 3 ? k + 1 : 0;
  const long long f = (fk * jx + (j + 1)) * ix + (i + 1);
#pragma unroll
  for (int m = 0; m < NV; m++) s[m * plane + (long long)k * zstride + (long long)j * pitch + i] = U[m * n + f];
}

__global__ void k_zero_red(unsigned long long *red) {
  if (threadIdx.x < 4) red[threadIdx.x] = 0ull;
}

typedef void (*march_fn)(const Sweep);

template <int LIM, int PASS, int RES>
march_fn pick3(int stage, int tt) {
  if (stage == 1) return tt == 0 ? k_march<LIM, 1, 0, PASS, RES> : k_march<LIM, 1, 1, PASS, RES>;
  return tt == 0 ? k_march<LIM, 2, 0, PASS, RES> : k_march<LIM, 2, 1, PASS, RES>;
}

template <int PASS, int RES>
march_fn pick2(int lim, int stage, int tt) {
  switch (lim) {
    case 0: return pick3<0, PASS, RES>(stage, tt);
    case 1: return pick3<1, PASS, RES>(stage, tt);
    case 2: return pick3<2, PASS, RES>(stage, tt);
    case 3: return pick3<3, PASS, RES>(stage, tt);
    default: return pick3<4, PASS, RES>(stage, tt);
  }
}

template <int LIM, int PASS>
march_fn pick_gen(int stage, int tt) {
  if (stage == 1) return tt == 0 ? k_march_gen<LIM, 1, 0, PASS> : k_march_gen<LIM, 1, 1, PASS>;
  return tt == 0 ? k_march_gen<LIM, 2, 0, PASS> : k_march_gen<LIM, 2, 1, PASS>;
}

template <int PASS>
march_fn pick_gen2(int lim, int stage, int tt) {
  switch (lim) {
    case 0: return pick_gen<0, PASS>(stage, tt);
    case 1: return pick_gen<1, PASS>(stage, tt);
    case 2: return pick_gen<2, PASS>(stage, tt);
    case 3: return pick_gen<3, PASS>(stage, tt);
    default: return pick_gen<4, PASS>(stage, tt);
  }
}

// gen: a flux type other than HLLD (ideal MHD only; omhd_create enforces it)
march_fn march_kernel(int pass, int res, int lim, int stage, int tt, bool gen = false) {
  if (gen) {
    if (pass == PASS_2D) return pick_gen2<PASS_2D>(lim, stage, tt);
    if (pass == PASS_XY) return pick_gen2<PASS_XY>(lim, stage, tt);
    return pick_gen2<PASS_Z>(lim, stage, tt);
  }
  if (pass == PASS_2D) return pick2<PASS_2D, 0>(lim, stage, tt);
  if (pass == PASS_XY) return res ? pick2<PASS_XY, 1>(lim, stage, tt) : pick2<PASS_XY, 0>(lim, stage, tt);
  return res ? pick2<PASS_Z, 1>(lim, stage, tt) : pick2<PASS_Z, 0>(lim, stage, tt);
}

// Chunk height when the caller leaves it to the library.  Each wavefront
// marches its chunk row after row, so once a grid has fewer than a few rounds
// of wavefronts (slots = CUs x 2 per SIMD x 4 SIMDs) the step time follows the
// chunk height, not the cell count: KH 512^2 ran 2.5x faster with 8-row chunks
// than with 28.  Such grids take the height that minimises
// ceil(tiles / slots) * (rows + 3), the 3 being the rows loaded before a chunk's
// first update (ties to the taller chunk).  Larger grids keep the measured
// heights: 28 rows in 2D (KH 4096^2), 24 in 3D (Harris 512x256x256).
int default_rows(int nx, int ny, int nz, int nd, long long slots) {
  const long long strips = (nx + STRIP - 1) / STRIP, slices = nd == 3 ? nz : 1;
  auto tiles = [&](int r) { return strips * slices * ((ny + r - 1) / r); };
  if (tiles(24) >= 4 * slots) return nd == 2 ? 28 : 24;
  int best = 28;
  long long best_cost = -1;
  for (int r : {28, 24, 20, 16, 12, 8}) {
    const long long cost = (tiles(r) + slots - 1) / slots * (r + 3);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = r;
    }
  }
  return best;
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

// red: [max speed bits | err[4] | err_next[2]]
static int read_red(omhd_handle *h, double &vmax, int err[4], int next[2] = nullptr) {
  OMHD_HIP(hipMemcpyAsync(h->red_host, h->red, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream));
  OMHD_HIP(hipStreamSynchronize(h->stream));
  memcpy(&vmax, &h->red_host[0], sizeof(double));
  memcpy(err, &h->red_host[1], 4 * sizeof(int));
  if (next) memcpy(next, &h->red_host[3], 2 * sizeof(int));
  return OMHD_OK;
}

// the face arrays of the configured resistivity eta(x,y) = eta0 + eta1 *
// exp(-(x^2+y^2)/w^2), at x-, y- and z-face centres, computed on the host with
// libm (transcendentals stay off the device so the arithmetic matches a CPU
// build); cell centres are x0 - dx/2 + i*dx accumulated as model.f90:19-28 does
static double eta_faces(const omhd_config &c, std::vector<double> &out) {
  const int ix = c.nx + 2, jx = c.ny + 2;
  std::vector<double> x(ix), y(jx);
  x[0] = c.x0 - c.dx / 2;
  for (int i = 1; i < ix; i++) x[i] = x[i - 1] + c.dx;
  y[0] = c.y0 - c.dx / 2;
  for (int j = 1; j < jx; j++) y[j] = y[j - 1] + c.dx;
  out.assign((size_t)3 * ix * jx, 0.0);
  double emax = 0.0;
  for (int d = 0; d < 3; d++)
    for (int j = 0; j < jx; j++)
      for (int i = 0; i < ix; i++) {
        const double xx = x[i] + (d == 0 ? 0.5 * c.dx : 0.0);
        const double yy = y[j] + (d == 1 ? 0.5 * c.dx : 0.0);
        const double r2 = xx * xx + yy * yy;
        const double e = c.eta0 + c.eta1 * exp(-r2 / (c.eta_w * c.eta_w));
        out[(size_t)d * ix * jx + (size_t)i + (size_t)ix * j] = e;
        if (e > emax) emax = e;
      }
  return emax;
}

extern "C" int omhd_create(const omhd_config *cfg, omhd_handle **out) {
  if (!cfg || !out) return set_error(OMHD_ERR_ARG, "null argument");
  if (cfg->ndim != 2 && cfg->ndim != 3) return set_error(OMHD_ERR_ARG, "ndim must be 2 or 3", cfg->ndim);
  const int nz = cfg->ndim == 3 ? cfg->nz : 1;
  if (cfg->nx < 2 || cfg->ny < 2 || (cfg->ndim == 3 && nz < 2))
    return set_error(OMHD_ERR_EXTENT, "need nx, ny (, nz) >= 2", cfg->nx, cfg->ny, nz);
  if (cfg->lm_type < 0 || cfg->lm_type > 4) return set_error(OMHD_ERR_TYPE, "unknown limiter", cfg->lm_type);
  if (cfg->flux_type < 0 || cfg->flux_type > 4) return set_error(OMHD_ERR_TYPE, "unknown flux_type", cfg->flux_type);
  if (cfg->flux_type != 3 && cfg->resistive)
    return set_error(OMHD_ERR_TYPE, "flux types other than HLLD are built for ideal MHD (no resistivity)",
                     cfg->flux_type);
  if (cfg->time_type != 0 && cfg->time_type != 1) return set_error(OMHD_ERR_TYPE, "unknown time_type", cfg->time_type);
  if (!(cfg->dx > 0) || !(cfg->cfl > 0)) return set_error(OMHD_ERR_ARG, "dx and cfl must be positive");
  if (cfg->resistive && cfg->ndim != 3)
    return set_error(OMHD_ERR_ARG, "the resistive flux is built for the 3D step (flux_resistive.cuf)");
  if (cfg->resistive && !(cfg->eta_w > 0)) return set_error(OMHD_ERR_ARG, "eta_w must be positive");
  for (int a = 0; a < 3; a++)
    if (cfg->bc_type[2 * a] != cfg->bc_type[2 * a + 1])
      return set_error(OMHD_ERR_ARG, "both faces of an axis share one boundary type", a);
  omhd_handle *h = new omhd_handle();
  h->cfg = *cfg;
  h->cfg.nz = nz;
  h->nd = cfg->ndim;
  h->pitch = (cfg->nx + 7) & ~7;  // 64-byte aligned rows
  h->zstride = (long long)h->pitch * cfg->ny;
  h->plane = ((h->zstride * nz) + 63) & ~63LL;
  // default chunk heights (default_rows), measured: on KH 4096^2, chunks of 16,
  // 24 or 32 rows ran about 2% slower than 20 or 28 (28 best); on the
  // 512x256x256 Harris sheet 24 was best, with 20 and 28 within 0.5%
  if (cfg->rows_per_wave > 0) {
    h->rows = cfg->rows_per_wave;
  } else {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) {
      (void)hipGetLastError();
      cus = 256;
    }
    h->rows = default_rows(cfg->nx, cfg->ny, nz, cfg->ndim, 8LL * cus);
  }
  for (int a = 0; a < 3; a++) {
    h->faces.bc[a] = cfg->bc_type[2 * a];
    h->faces.slo[a] = sign_mask(cfg->bc_sign + 18 * a);
    h->faces.shi[a] = sign_mask(cfg->bc_sign + 18 * a + 9);
  }
  const size_t bytes = (size_t)h->plane * NV * sizeof(double);
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void **)&h->U, bytes);
  if (e == hipSuccess) e = hipMalloc((void **)&h->U1, bytes);
  if (e == hipSuccess && h->nd == 3) e = hipMalloc((void **)&h->acc, bytes);
  if (e == hipSuccess) e = hipMalloc((void **)&h->red, 4 * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipHostMalloc((void **)&h->red_host, 4 * sizeof(unsigned long long), hipHostMallocDefault);
  if (e == hipSuccess && cfg->resistive) {
    std::vector<double> eh;
    h->eta_max = eta_faces(h->cfg, eh);
    e = hipMalloc((void **)&h->eta, eh.size() * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(h->eta, eh.data(), eh.size() * sizeof(double), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    omhd_destroy(h);
    return set_error(OMHD_ERR_HIP, hipGetErrorString(e));
  }
  *out = h;
  return ok();
}

extern "C" int omhd_set_resistivity(omhd_handle *h, double eta0, double eta1, double eta_w) {
  if (!h) return set_error(OMHD_ERR_ARG, "null handle");
  if (!h->cfg.resistive) return set_error(OMHD_ERR_ARG, "the handle was created without resistivity");
  if (!(eta_w > 0)) return set_error(OMHD_ERR_ARG, "eta_w must be positive");
  h->cfg.eta0 = eta0;
  h->cfg.eta1 = eta1;
  h->cfg.eta_w = eta_w;
  std::vector<double> eh;
  h->eta_max = eta_faces(h->cfg, eh);
  OMHD_HIP(hipMemcpyAsync(h->eta, eh.data(), eh.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
  OMHD_HIP(hipStreamSynchronize(h->stream));
  return ok();
}

static void solo_drop_graph(omhd_handle *h);

extern "C" int omhd_destroy(omhd_handle *h) {
  if (!h) return ok();
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->U) (void)hipFree(h->U);
  if (h->U1) (void)hipFree(h->U1);
  if (h->acc) (void)hipFree(h->acc);
  if (h->eta) (void)hipFree(h->eta);
  if (h->red) (void)hipFree(h->red);
  if (h->red_host) (void)hipHostFree(h->red_host);
  for (int i = 0; i < 3; i++)
    if (h->ev[i]) (void)hipEventDestroy(h->ev[i]);
  for (hipEvent_t e : h->step_ev) (void)hipEventDestroy(e);
  solo_drop_graph(h);
  if (h->solo) (void)hipFree(h->solo);
  if (h->solo_host) (void)hipHostFree(h->solo_host);
  if (h->trace_d) (void)hipFree(h->trace_d);
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

static size_t ncells(const omhd_handle *h) { return (size_t)h->cfg.nx * h->cfg.ny * h->cfg.nz; }
static size_t nfortran(const omhd_handle *h) {
  return (size_t)(h->cfg.nx + 2) * (h->cfg.ny + 2) * (h->nd == 3 ? h->cfg.nz + 2 : 1);
}

static int state_error(const int err[4]) {
  const char *what = err[0] == OMHD_ERR_PRESSURE ? "negative pressure"
                     : err[0] == OMHD_ERR_DENSITY ? "negative density"
                                                  : "non-finite signal speed";
  return set_error(err[0], what, err[1], err[2], err[3]);
}

// The reference's u2v stop (u2v.f90:34-43) on a state that failed its
// positivity test: over the ghost-inclusive array, prmin is the minimum of
// E(1,1,1) and every pressure, and if prmin <= 0 the report is the first
// cell (column-major) of minimum pressure with that pressure; otherwise the
// same for the density.  `state` is the resident buffer that failed.  If the
// diagnostic cannot run, the kernel's own report (code and first cell found)
// is returned instead.
static int positivity_error(omhd_handle *h, const double *state, const int err[4]) {
  const omhd_config &c = h->cfg;
  const size_t n = nfortran(h);
  double *dU = nullptr, *dV = nullptr;
  std::vector<double> p(n), ro(n);
  double e11 = 0.0;
  bool ok_ = hipMalloc((void **)&dU, 9 * n * sizeof(double)) == hipSuccess &&
             hipMalloc((void **)&dV, 4 * n * sizeof(double)) == hipSuccess;
  if (ok_) {
    k_to_fortran<<<(unsigned)((n + 255) / 256), 256, 0, h->stream>>>(state, h->plane, h->zstride, h->pitch, c.nx, c.ny,
                                                                     c.nz, h->nd, h->faces, dU, dV);
    ok_ = hipGetLastError() == hipSuccess &&
          hipMemcpyAsync(p.data(), dV + 3 * n, n * sizeof(double), hipMemcpyDeviceToHost, h->stream) == hipSuccess &&
          hipMemcpyAsync(ro.data(), dU + 4 * n, n * sizeof(double), hipMemcpyDeviceToHost, h->stream) == hipSuccess &&
          hipMemcpyAsync(&e11, dU + 3 * n, sizeof(double), hipMemcpyDeviceToHost, h->stream) == hipSuccess &&
          hipStreamSynchronize(h->stream) == hipSuccess;
  }
  if (dU) (void)hipFree(dU);
  if (dV) (void)hipFree(dV);
  if (!ok_) {
    (void)hipGetLastError();
    return state_error(err);
  }
  size_t ip = 0, ir = 0;
  for (size_t q = 1; q < n; q++) {
    if (p[q] < p[ip]) ip = q;
    if (ro[q] < ro[ir]) ir = q;
  }
  const double prmin = e11 < p[ip] ? e11 : p[ip];  // u2v.f90 starts prmin at U(1,1,en)
  const double rhomin = ro[0] < ro[ir] ? ro[0] : ro[ir];
  const int ix = c.nx + 2, jx = c.ny + 2;
  const size_t q = prmin <= 0 ? ip : ir;
  const int i = (int)(q % ix) + 1, j = (int)(q / ix % jx) + 1, k = (int)(q / ((size_t)ix * jx)) + 1;
  if (prmin <= 0) return set_error(OMHD_ERR_PRESSURE, "negative pressure", i, j, k, p[ip]);
  if (rhomin <= 0) return set_error(OMHD_ERR_DENSITY, "negative density", i, j, k, ro[ir]);
  return state_error(err);  // the kernel saw a non-positive value the arrays no longer show
}

// After a step: a failure in a state the stages read (U^n, or U* when tagged
// ERR_IN_USTAR) stops the step before it completes.
static int read_error(omhd_handle *h, const int err[4]) {
  int e[4] = {err[0] & ~ERR_IN_USTAR, err[1], err[2], err[3]};
  if (e[0] != OMHD_ERR_PRESSURE && e[0] != OMHD_ERR_DENSITY) return state_error(e);
  return positivity_error(h, (err[0] & ERR_IN_USTAR) ? h->U1 : h->U, e);
}

static void solo_drop_graph(omhd_handle *h) {
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
}

// Device dt mode: the parameters of the next step from the current CFL speed,
// exactly as step_dt and stage_launch compute them on the host, and the table
// of host exp values around x0 = (-0.5*(cfl*dx))/0.2 for the device's GLM
// factors.  A dt below set_dt's bound halts the handle: its next step call
// returns that stop, as omhd_step's would.
static int solo_seed(omhd_handle *h) {
  const omhd_config &c = h->cfg;
  SoloDev *sh = (SoloDev *)h->solo_host;
  memset(sh, 0, sizeof(SoloDev));
  const double x0 = -0.5 * (c.cfl * c.dx) / 0.2;
  memcpy(&sh->glm_x0, &x0, sizeof x0);
  for (int j = -GLM_TAB_HALF; j <= GLM_TAB_HALF; j++) {
    const long long xb = sh->glm_x0 + j;
    double x;
    memcpy(&x, &xb, sizeof x);
    sh->glm_tab[j + GLM_TAB_HALF] = exp(x);
  }
  h->halt_rc = 0;
  const double ch = h->vmax, dt = c.cfl * c.dx / ch;
  if (!(dt >= 1.0e-7)) {
    sh->par.dt = 0.0;
    h->halt_rc = OMHD_ERR_DT;
    h->halt_dt = dt;
  } else {
    sh->par.ch = ch;
    sh->par.ch2 = ch * ch;
    sh->par.dtx1 = c.time_type == 0 ? dt / c.dx : 0.5 * dt / c.dx;
    sh->par.dtx2 = dt / c.dx;
    sh->par.glm_f = exp(-0.5 * dt * ch / 0.2);  // glm_ss2.f90:27
    sh->par.dt = dt;
  }
  OMHD_HIP(hipMemcpyAsync(h->solo, sh, sizeof(SoloDev), hipMemcpyHostToDevice, h->stream));
  OMHD_HIP(hipStreamSynchronize(h->stream));
  return OMHD_OK;
}

static int cfl_pass(omhd_handle *h) {
  const omhd_config &c = h->cfg;
  k_zero_red<<<1, 64, 0, h->stream>>>(h->red);
  const size_t n = ncells(h);
  const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 2048);
  if (h->nd == 3)
    k_cfl<3><<<grid, 256, 0, h->stream>>>(h->U, h->plane, h->zstride, h->pitch, c.nx, c.ny, c.nz, h->red,
                                          (int *)(h->red + 1));
  else
    k_cfl<2><<<grid, 256, 0, h->stream>>>(h->U, h->plane, h->zstride, h->pitch, c.nx, c.ny, c.nz, h->red,
                                          (int *)(h->red + 1));
  OMHD_HIP(hipGetLastError());
  int err[4];
  double vm;
  if (int rc = read_red(h, vm, err)) return rc;
  if (err[0]) return read_error(h, err);
  h->vmax = vm;
  return h->device_dt ? solo_seed(h) : OMHD_OK;
}

extern "C" int omhd_upload(omhd_handle *h, const double *U) {
  if (!h || !U) return set_error(OMHD_ERR_ARG, "null argument");
  const omhd_config &c = h->cfg;
  Staged a;
  if (int rc = a.init(U, 9 * nfortran(h), true, false, h->stream)) return rc;
  k_from_fortran<<<(unsigned)((ncells(h) + 255) / 256), 256, 0, h->stream>>>(a.dev, h->U, h->plane, h->zstride,
                                                                             h->pitch, c.nx, c.ny, c.nz, h->nd);
  OMHD_HIP(hipGetLastError());
  a.release(h->stream);
  if (int rc = cfl_pass(h)) return rc;
  h->have_state = true;
  return ok();
}

extern "C" int omhd_download(omhd_handle *h, double *U, double *V) {
  if (!h || !U) return set_error(OMHD_ERR_ARG, "null argument");
  const omhd_config &c = h->cfg;
  const size_t n = nfortran(h);
  Staged a, b;
  if (int rc = a.init(U, 9 * n, false, true, h->stream)) return rc;
  if (V)
    if (int rc = b.init(V, 4 * n, false, true, h->stream)) return rc;
  k_to_fortran<<<(unsigned)((n + 255) / 256), 256, 0, h->stream>>>(h->U, h->plane, h->zstride, h->pitch, c.nx, c.ny,
                                                                   c.nz, h->nd, h->faces, a.dev,
                                                                   V ? b.dev : nullptr);
  OMHD_HIP(hipGetLastError());
  if (int rc = a.finish(h->stream)) return rc;
  if (V)
    if (int rc = b.finish(h->stream)) return rc;
  a.release(h->stream);
  b.release(h->stream);
  OMHD_HIP(hipStreamSynchronize(h->stream));
  return ok();
}

static Sweep base_sweep(omhd_handle *h, int pass) {
  const omhd_config &c = h->cfg;
  Sweep p{};
  p.plane = h->plane;
  p.zstride = h->zstride;
  p.pitch = h->pitch;
  p.nx = c.nx;
  p.ny = c.ny;
  p.nz = c.nz;
  p.nstrip = (c.nx + STRIP - 1) / STRIP;
  p.rows = h->rows;
  const int nm = pass == PASS_Z ? c.nz : c.ny;
  const int nslice = pass == PASS_2D ? 1 : pass == PASS_XY ? c.nz : c.ny;
  p.nchunk = (nm + p.rows - 1) / p.rows;
  p.ntiles = p.nstrip * p.nchunk * nslice;
  for (int a = 0; a < 3; a++) {
    p.bc[a] = h->faces.bc[a];
    p.slo[a] = h->faces.slo[a];
    p.shi[a] = h->faces.shi[a];
  }
  p.dx = c.dx;
  p.eta = h->eta;
  p.acc = h->acc;
  p.vmax_bits = h->red;
  p.err = (int *)(h->red + 1);
  p.err_next = (int *)(h->red + 3);
  return p;
}

static unsigned grid_of(const Sweep &p) { return (unsigned)((p.ntiles + 7) / 8 * 8); }

namespace omhd {

int stage_launch(omhd_handle *h, int stage, const double *us, const double *u0, double *out, double dt, double ch,
                 const void *dpar) {
  const omhd_config &c = h->cfg;
  const bool three = h->nd == 3;
  const int res = c.resistive ? 1 : 0;
  Sweep pa = base_sweep(h, three ? PASS_XY : PASS_2D);
  Sweep pz = three ? base_sweep(h, PASS_Z) : pa;
  const march_fn a = march_kernel(three ? PASS_XY : PASS_2D, res, c.lm_type, stage, c.time_type, c.flux_type != 3);
  const march_fn z = three ? march_kernel(PASS_Z, res, c.lm_type, stage, c.time_type, c.flux_type != 3) : nullptr;
  const double cr = 0.2;                          // glm_ss2.f90:18
  const double glm_f = exp(-0.5 * dt * ch / cr);  // glm_ss2.f90:27, host libm
  // stage 1: U* = U + dtx dF (rk_tvd21: dt/dx, rk_std21: 0.5*dt/dx); stage 2: dt/dx
  const double dtx = (stage == 1 && c.time_type != 0) ? 0.5 * dt / c.dx : dt / c.dx;
  for (Sweep *q : {&pa, &pz}) {
    q->ch = ch;
    q->ch2 = ch * ch;
    q->glm_f = glm_f;
    q->us = us;
    q->u0 = u0;
    q->out = out;
    q->dtx = dtx;
    q->ftype = c.flux_type;
    q->dpar = (const double *)dpar;
  }
  hipLaunchKernelGGL(a, dim3(grid_of(pa)), dim3(64), 0, h->stream, pa);
  OMHD_HIP(hipGetLastError());
  if (three) {
    hipLaunchKernelGGL(z, dim3(grid_of(pz)), dim3(64), 0, h->stream, pz);
    OMHD_HIP(hipGetLastError());
  }
  return OMHD_OK;
}

int refresh_cfl(omhd_handle *h) { return cfl_pass(h); }

}  // namespace omhd

// the step's dt from the current CFL speed (set_dt.f90:52, set_dt_resistive)
static int step_dt(omhd_handle *h, double &dt, double &ch) {
  const omhd_config &c = h->cfg;
  ch = h->vmax;           // set_dt returns vmax as ch (main.f90:80)
  dt = c.cfl * c.dx / h->vmax;
  if (!(dt >= 1.0e-7)) return set_error(OMHD_ERR_DT, "dt is too small", 0, 0, 0, dt);  // set_dt.f90:55-65
  if (c.resistive) {  // set_dt_resistive (flux_resistive.cuf:2-9) with the configured eta
    const double dr = 0.5 * c.cfl * (c.dx * c.dx) / h->eta_max;
    dt = dt < dr ? dt : dr;
  }
  return OMHD_OK;
}

// one step of device dt mode: the two stages read their parameters from the
// device block, then k_solo_dt prepares the next step's
static int solo_step_launch(omhd_handle *h, double *trace) {
  SoloDev *sd = (SoloDev *)h->solo;
  if (int rc = stage_launch(h, 1, h->U, h->U, h->U1, 0.0, 0.0, &sd->par)) return rc;
  if (int rc = stage_launch(h, 2, h->U1, h->U, h->U, 0.0, 0.0, &sd->par)) return rc;
  k_solo_dt<<<1, 1, 0, h->stream>>>(sd, h->red, h->cfg.cfl, h->cfg.dx, h->cfg.time_type, trace);
  OMHD_HIP(hipGetLastError());
  return OMHD_OK;
}

// the stop a halted handle returns again
static int solo_halt_error(omhd_handle *h) {
  if (h->halt_rc == OMHD_ERR_DT) return set_error(OMHD_ERR_DT, "dt is too small", 0, 0, 0, h->halt_dt);
  if (h->halt_rc == OMHD_ERR_PRESSURE || h->halt_rc == OMHD_ERR_DENSITY) return positivity_error(h, h->U, h->halt_err);
  return state_error(h->halt_err);
}

// omhd_step in device dt mode: the steps are queued without waiting, and the
// time, traces, step count and any stop are read once at the end.  The
// results are those of the host-dt loop below bit for bit, the stop included.
static int step_device(omhd_handle *h, int nsteps, double *t, double *dt_trace, double *ch_trace, int32_t *done) {
  if (h->halt_rc) return solo_halt_error(h);
  const bool tr = dt_trace || ch_trace;
  if (tr && h->trace_cap < nsteps) {
    if (h->trace_d) OMHD_HIP(hipFree(h->trace_d));
    h->trace_d = nullptr;
    h->trace_cap = 0;
    OMHD_HIP(hipMalloc((void **)&h->trace_d, 2 * (size_t)nsteps * sizeof(double)));
    h->trace_cap = nsteps;
    solo_drop_graph(h);  // the captured step names the trace buffer
  }
  double *trace = tr ? h->trace_d : nullptr;
  if (h->gexec && h->graph_traced != tr) solo_drop_graph(h);
  if (h->use_graph && !h->gexec && nsteps > 0) {
    OMHD_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int rc = solo_step_launch(h, trace);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return set_error(OMHD_ERR_HIP, hipGetErrorString(e));
    h->graph = g;
    OMHD_HIP(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
    h->graph_traced = tr;
  }
  const bool timed = h->timing && !h->use_graph;
  if (timed)
    while (h->step_ev.size() < 3 * (size_t)nsteps) {
      hipEvent_t e;
      OMHD_HIP(hipEventCreate(&e));
      h->step_ev.push_back(e);
    }
  SoloDev *sd = (SoloDev *)h->solo;
  k_solo_begin<<<1, 1, 0, h->stream>>>(sd, h->red, t ? *t : 0.0);
  OMHD_HIP(hipGetLastError());
  for (int n = 0; n < nsteps; n++) {
    if (h->use_graph) {
      OMHD_HIP(hipGraphLaunch(h->gexec, h->stream));
      continue;
    }
    if (!timed) {
      if (int rc = solo_step_launch(h, trace)) return rc;
      continue;
    }
    OMHD_HIP(hipEventRecord(h->step_ev[3 * n], h->stream));
    if (int rc = stage_launch(h, 1, h->U, h->U, h->U1, 0.0, 0.0, &sd->par)) return rc;
    OMHD_HIP(hipEventRecord(h->step_ev[3 * n + 1], h->stream));
    if (int rc = stage_launch(h, 2, h->U1, h->U, h->U, 0.0, 0.0, &sd->par)) return rc;
    OMHD_HIP(hipEventRecord(h->step_ev[3 * n + 2], h->stream));
    k_solo_dt<<<1, 1, 0, h->stream>>>(sd, h->red, h->cfg.cfl, h->cfg.dx, h->cfg.time_type, trace);
    OMHD_HIP(hipGetLastError());
  }
  SoloDev *sh = (SoloDev *)h->solo_host;
  OMHD_HIP(hipMemcpyAsync(sh, sd, sizeof(SoloDev), hipMemcpyDeviceToHost, h->stream));
  OMHD_HIP(hipMemcpyAsync(h->red_host, h->red, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream));
  std::vector<double> trh(tr ? 2 * (size_t)nsteps : 0);
  if (tr && nsteps > 0)
    OMHD_HIP(hipMemcpyAsync(trh.data(), h->trace_d, trh.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  OMHD_HIP(hipStreamSynchronize(h->stream));
  const int nd = (int)sh->done;
  if (t) *t = sh->t;
  if (done) *done = nd;
  for (int n = 0; n < nd; n++) {
    if (dt_trace) dt_trace[n] = trh[2 * n];
    if (ch_trace) ch_trace[n] = trh[2 * n + 1];
  }
  if (timed)
    for (int n = 0; n < nsteps; n++) {
      float a = 0.f, b = 0.f;
      OMHD_HIP(hipEventElapsedTime(&a, h->step_ev[3 * n], h->step_ev[3 * n + 1]));
      OMHD_HIP(hipEventElapsedTime(&b, h->step_ev[3 * n + 1], h->step_ev[3 * n + 2]));
      h->stage_ms[0] += a;
      h->stage_ms[1] += b;
      h->stage_launches[0]++;
      h->stage_launches[1]++;
    }
  if (sh->par.dt > 0.0) h->vmax = sh->par.ch;
  int err[4], next[2];
  memcpy(err, &h->red_host[1], sizeof err);
  memcpy(next, &h->red_host[3], sizeof next);
  if (err[0] == OMHD_ERR_DT) {  // the next step's dt: omhd_step stops there, at the start of that step
    h->halt_rc = OMHD_ERR_DT;
    h->halt_dt = sh->bad_dt;
    return nd == nsteps ? ok() : solo_halt_error(h);
  }
  if (err[0]) {
    h->halt_rc = err[0] & ~ERR_IN_USTAR;
    for (int i = 0; i < 4; i++) h->halt_err[i] = err[i];
    h->halt_err[0] = h->halt_rc;
    return read_error(h, err);
  }
  if (next[0] || next[1]) {
    h->halt_rc = next[0] ? next[0] : next[1];
    h->halt_err[0] = h->halt_rc;
    h->halt_err[1] = h->halt_err[2] = h->halt_err[3] = 0;
    return solo_halt_error(h);
  }
  return ok();
}

extern "C" int omhd_set_mode(omhd_handle *h, int32_t device_dt, int32_t use_graph) {
  if (!h) return set_error(OMHD_ERR_ARG, "null handle");
  if (use_graph && !device_dt) return set_error(OMHD_ERR_ARG, "graph capture needs device dt mode");
  if (device_dt && h->cfg.resistive)
    return set_error(OMHD_ERR_ARG,
                     "device dt mode is built for the ideal step: a resistivity-limited dt moves the GLM factor's "
                     "argument off the table of host exp values");
  if (device_dt && !h->have_state) return set_error(OMHD_ERR_ARG, "omhd_upload first");
  if (device_dt && !h->solo) OMHD_HIP(hipMalloc(&h->solo, sizeof(SoloDev)));
  if (device_dt && !h->solo_host) OMHD_HIP(hipHostMalloc(&h->solo_host, sizeof(SoloDev)));
  if (!use_graph || !device_dt) solo_drop_graph(h);
  const bool enter = device_dt && !h->device_dt;
  h->device_dt = device_dt != 0;
  h->use_graph = use_graph != 0;
  if (enter)
    if (int rc = solo_seed(h)) return rc;
  if (!device_dt) h->halt_rc = 0;
  return ok();
}

extern "C" int omhd_step(omhd_handle *h, int32_t nsteps, double *t, double *dt_trace, double *ch_trace, int32_t *done) {
  if (done) *done = 0;
  if (!h) return set_error(OMHD_ERR_ARG, "null handle");
  if (nsteps < 0) return set_error(OMHD_ERR_ARG, "nsteps must be >= 0", nsteps);
  if (!h->have_state) return set_error(OMHD_ERR_ARG, "omhd_upload first");
  if (h->device_dt) return step_device(h, nsteps, t, dt_trace, ch_trace, done);
  for (int n = 0; n < nsteps; n++) {
    double dt, ch;
    if (int rc = step_dt(h, dt, ch)) return rc;
    k_zero_red<<<1, 64, 0, h->stream>>>(h->red);
    if (h->timing) OMHD_HIP(hipEventRecord(h->ev[0], h->stream));
    if (int rc = stage_launch(h, 1, h->U, h->U, h->U1, dt, ch)) return rc;
    if (h->timing) OMHD_HIP(hipEventRecord(h->ev[1], h->stream));
    // stage 2: U^{n+1} in place over U^n (each cell's U^n is read only by the
    // wavefront that overwrites it, just before it does)
    if (int rc = stage_launch(h, 2, h->U1, h->U, h->U, dt, ch)) return rc;
    if (h->timing) OMHD_HIP(hipEventRecord(h->ev[2], h->stream));
    int err[4], next[2];
    double vm;
    if (int rc = read_red(h, vm, err, next)) return rc;
    if (err[0]) return read_error(h, err);
    if (next[0] || next[1]) {  // the step completed; the next u2v or set_dt stops on U^{n+1}
      if (t) *t += dt;
      if (dt_trace) dt_trace[n] = dt;
      if (ch_trace) ch_trace[n] = ch;
      if (done) *done = n + 1;
      const int e[4] = {next[0] ? next[0] : next[1], 0, 0, 0};
      return next[0] ? positivity_error(h, h->U, e) : state_error(e);
    }
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

extern "C" int omhd_step_fixed(omhd_handle *h, int32_t nsteps, const double *dt, const double *ch) {
  if (!h || (nsteps > 0 && (!dt || !ch))) return set_error(OMHD_ERR_ARG, "null argument");
  if (!h->have_state) return set_error(OMHD_ERR_ARG, "omhd_upload first");
  for (int n = 0; n < nsteps; n++)
    if (!(dt[n] > 0) || !(ch[n] >= 0)) return set_error(OMHD_ERR_ARG, "dt must be > 0 and ch >= 0", n);
  k_zero_red<<<1, 64, 0, h->stream>>>(h->red);
  for (int n = 0; n < nsteps; n++) {
    if (int rc = stage_launch(h, 1, h->U, h->U, h->U1, dt[n], ch[n])) return rc;
    if (int rc = stage_launch(h, 2, h->U1, h->U, h->U, dt[n], ch[n])) return rc;
  }
  int err[4], next[2];
  double vm;
  if (int rc = read_red(h, vm, err, next)) return rc;
  if (err[0]) return read_error(h, err);
  if (next[0] || next[1]) {
    const int e[4] = {next[0] ? next[0] : next[1], 0, 0, 0};
    return next[0] ? positivity_error(h, h->U, e) : state_error(e);
  }
  if (int rc = cfl_pass(h)) return rc;  // the speed of the final state, for a later omhd_step
  return ok();
}

__global__ void k_state_copy(double *__restrict__ dst, const double *__restrict__ src, long long plane, long long zstride,
                             int pitch, int nx, int ny, int nz, int to_compact) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)nx * ny * nz;
  if (gid >= n) return;
  const int i = (int)(gid % nx);
  const long long r = gid / nx;
  const int j = (int)(r % ny), k = (int)(r / ny);
  const long long o = (long long)k * zstride + (long long)j * pitch + i;
  for (int m = 0; m < NV; m++) {
    if (to_compact) dst[m * n + gid] = src[m * plane + o];
    else dst[m * plane + o] = src[m * n + gid];
  }
}

extern "C" int omhd_set_state(omhd_handle *h, const double *W) {
  if (!h || !W) return set_error(OMHD_ERR_ARG, "null argument");
  const omhd_config &c = h->cfg;
  Staged a;
  if (int rc = a.init(W, 9 * ncells(h), true, false, h->stream)) return rc;
  k_state_copy<<<(unsigned)((ncells(h) + 255) / 256), 256, 0, h->stream>>>(h->U, a.dev, h->plane, h->zstride,
                                                                           h->pitch, c.nx, c.ny, c.nz, 0);
  OMHD_HIP(hipGetLastError());
  a.release(h->stream);
  if (int rc = cfl_pass(h)) return rc;
  h->have_state = true;
  return ok();
}

extern "C" int omhd_get_state(omhd_handle *h, double *W) {
  if (!h || !W) return set_error(OMHD_ERR_ARG, "null argument");
  if (!h->have_state) return set_error(OMHD_ERR_ARG, "no state uploaded");
  const omhd_config &c = h->cfg;
  Staged a;
  if (int rc = a.init(W, 9 * ncells(h), false, true, h->stream)) return rc;
  k_state_copy<<<(unsigned)((ncells(h) + 255) / 256), 256, 0, h->stream>>>(a.dev, h->U, h->plane, h->zstride,
                                                                           h->pitch, c.nx, c.ny, c.nz, 1);
  OMHD_HIP(hipGetLastError());
  if (int rc = a.finish(h->stream)) return rc;
  a.release(h->stream);
  OMHD_HIP(hipStreamSynchronize(h->stream));
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

// ---------------------------------------------------------------------------
// Ensemble of independent 2D grids (batched step).
//
// Every member keeps its own SoA state; one tile table lists (member, tile)
// pairs for all members, and each RK stage is one persistent launch over it.
// Each member has its own dt: either computed on the host from its reduced CFL
// speed exactly as omhd_step does (host dt mode: every member is bitwise equal
// to its solo run), or on the device (device dt mode: no host round trip, so
// whole steps are captured in one hipGraph and replayed).
// ---------------------------------------------------------------------------
struct omhd_batch {
  int n = 0;
  std::vector<omhd_config> cfg;
  std::vector<Sweep> mem;  // host copy of the member descriptors
  Sweep *mem_d = nullptr;
  int2 *tiles_d = nullptr;
  int ntiles = 0;
  double *arena = nullptr;  // all U and U1 buffers
  size_t arena_doubles = 0;
  double *backup = nullptr;  // autotuner: state saved around trials
  unsigned long long *vmax_d = nullptr;
  int *err_d = nullptr;
  DevPar *par_d = nullptr;
  double *dx_d = nullptr, *cfl_d = nullptr, *tsum_d = nullptr;
  int *tt_d = nullptr;
  double *glm_tab_d = nullptr;     // n x GLM_TAB host exp values (device-dt GLM factors)
  long long *glm_x0_d = nullptr;   // bits of each member's x0
  std::vector<DevPar> par_h;
  std::vector<unsigned long long> vmax_h;
  std::vector<double> vmax;  // CFL speed of each member's current state
  std::vector<double> t;
  std::vector<char> uploaded;  // members given a state by omhd_batch_upload
  hipStream_t stream = nullptr;
  int rows = 24, grid = 0;
  bool device_dt = false, use_graph = false, have_state = false;  // have_state: every member uploaded
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int lim = 1, tt = 0;
};

static int batch_tiles(omhd_batch *b) {
  std::vector<int2> tl;
  for (int i = 0; i < b->n; i++) {
    Sweep &p = b->mem[i];
    p.rows = b->rows;
    p.nchunk = (p.ny + p.rows - 1) / p.rows;
    p.ntiles = p.nstrip * p.nchunk;
    for (int t = 0; t < p.ntiles; t++) tl.push_back(make_int2(i, t));
  }
  b->ntiles = (int)tl.size();
  if (b->tiles_d) OMHD_HIP(hipFree(b->tiles_d));
  OMHD_HIP(hipMalloc((void **)&b->tiles_d, tl.size() * sizeof(int2)));
  OMHD_HIP(hipMemcpy(b->tiles_d, tl.data(), tl.size() * sizeof(int2), hipMemcpyHostToDevice));
  OMHD_HIP(hipMemcpy(b->mem_d, b->mem.data(), b->n * sizeof(Sweep), hipMemcpyHostToDevice));
  if (b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  return OMHD_OK;
}

typedef void (*ens_fn)(const Sweep *, const int2 *, int);

static ens_fn ens_kernel(int lim, int stage, int tt) {
#define OMHD_ENS(L)                                                                                 \
  if (lim == L) {                                                                                   \
    if (stage == 1) return tt == 0 ? k_march_ens<L, 1, 0> : k_march_ens<L, 1, 1>;                  \
    return tt == 0 ? k_march_ens<L, 2, 0> : k_march_ens<L, 2, 1>;                                  \
  }
  OMHD_ENS(0) OMHD_ENS(1) OMHD_ENS(2) OMHD_ENS(3) OMHD_ENS(4)
#undef OMHD_ENS
  return nullptr;
}

extern "C" int omhd_batch_create(const omhd_config *cfgs, int32_t n, omhd_batch **out) {
  if (!cfgs || !out || n < 1) return set_error(OMHD_ERR_ARG, "need n >= 1 configurations");
  for (int i = 0; i < n; i++) {
    const omhd_config &c = cfgs[i];
    if (c.ndim != 2) return set_error(OMHD_ERR_ARG, "ensemble members are 2D grids", i);
    if (c.nx < 2 || c.ny < 2) return set_error(OMHD_ERR_EXTENT, "need nx, ny >= 2", c.nx, c.ny, i);
    if (c.flux_type != 3) return set_error(OMHD_ERR_TYPE, "the ensemble is built for the HLLD flux", i);
    if (c.lm_type != cfgs[0].lm_type || c.time_type != cfgs[0].time_type)
      return set_error(OMHD_ERR_ARG, "members share one limiter and one time type (one kernel per stage)", i);
    if (c.lm_type < 0 || c.lm_type > 4) return set_error(OMHD_ERR_TYPE, "unknown limiter", c.lm_type);
    if (c.resistive) return set_error(OMHD_ERR_ARG, "ensemble members are ideal MHD", i);
    for (int a = 0; a < 2; a++)
      if (c.bc_type[2 * a] != c.bc_type[2 * a + 1])
        return set_error(OMHD_ERR_ARG, "both faces of an axis share one boundary type", i);
  }
  omhd_batch *b = new omhd_batch();
  b->n = n;
  b->cfg.assign(cfgs, cfgs + n);
  b->lim = cfgs[0].lm_type;
  b->tt = cfgs[0].time_type;
  b->mem.resize(n);
  b->par_h.resize(n);
  b->vmax_h.resize(n);
  b->vmax.assign(n, -1.0);
  b->t.assign(n, 0.0);
  b->uploaded.assign(n, 0);
  std::vector<size_t> off(n);
  size_t total = 0;
  for (int i = 0; i < n; i++) {
    const omhd_config &c = cfgs[i];
    Sweep &p = b->mem[i];
    p = Sweep{};
    p.nx = c.nx;
    p.ny = c.ny;
    p.nz = 1;
    p.pitch = (c.nx + 7) & ~7;
    p.zstride = (long long)p.pitch * c.ny;
    p.plane = (p.zstride + 63) & ~63LL;
    p.nstrip = (c.nx + STRIP - 1) / STRIP;
    for (int a = 0; a < 3; a++) {
      p.bc[a] = c.bc_type[2 * a];
      p.slo[a] = sign_mask(c.bc_sign + 18 * a);
      p.shi[a] = sign_mask(c.bc_sign + 18 * a + 9);
    }
    p.dx = c.dx;
    off[i] = total;
    total += 2 * (size_t)p.plane * NV;
  }
  b->arena_doubles = total;
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void **)&b->arena, total * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&b->mem_d, n * sizeof(Sweep));
  if (e == hipSuccess) e = hipMalloc((void **)&b->vmax_d, n * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc((void **)&b->err_d, 8 * sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void **)&b->par_d, n * sizeof(DevPar));
  if (e == hipSuccess) e = hipMalloc((void **)&b->dx_d, n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&b->cfl_d, n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&b->tsum_d, n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&b->tt_d, n * sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void **)&b->glm_tab_d, (size_t)n * GLM_TAB * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&b->glm_x0_d, n * sizeof(long long));
  if (e != hipSuccess) {
    omhd_batch_destroy(b);
    return set_error(OMHD_ERR_HIP, hipGetErrorString(e));
  }
  std::vector<double> dx(n), cfl(n);
  std::vector<int> tt(n);
  for (int i = 0; i < n; i++) {
    Sweep &p = b->mem[i];
    p.U = b->arena + off[i];
    p.U1 = p.U + (size_t)p.plane * NV;
    p.vmax_bits = b->vmax_d + i;
    p.err = b->err_d;
    p.err_next = b->err_d + 4;
    p.dpar = (const double *)(b->par_d + i);
    dx[i] = cfgs[i].dx;
    cfl[i] = cfgs[i].cfl;
    tt[i] = cfgs[i].time_type;
  }
  OMHD_HIP(hipMemcpy(b->dx_d, dx.data(), n * sizeof(double), hipMemcpyHostToDevice));
  OMHD_HIP(hipMemcpy(b->cfl_d, cfl.data(), n * sizeof(double), hipMemcpyHostToDevice));
  OMHD_HIP(hipMemcpy(b->tt_d, tt.data(), n * sizeof(int), hipMemcpyHostToDevice));
  {  // the device-dt GLM factors: host libm exp around each member's x0
    std::vector<double> tab((size_t)n * GLM_TAB);
    std::vector<long long> x0b(n);
    for (int i = 0; i < n; i++) {
      const double a = cfl[i] * dx[i];
      const double x0 = -0.5 * a / 0.2;
      memcpy(&x0b[i], &x0, sizeof x0);
      for (int j = -GLM_TAB_HALF; j <= GLM_TAB_HALF; j++) {
        const long long xb = x0b[i] + j;
        double x;
        memcpy(&x, &xb, sizeof x);
        tab[(size_t)i * GLM_TAB + (size_t)(j + GLM_TAB_HALF)] = exp(x);
      }
    }
    OMHD_HIP(hipMemcpy(b->glm_tab_d, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    OMHD_HIP(hipMemcpy(b->glm_x0_d, x0b.data(), n * sizeof(long long), hipMemcpyHostToDevice));
  }
  OMHD_HIP(hipMemset(b->err_d, 0, 8 * sizeof(int)));
  OMHD_HIP(hipMemset(b->tsum_d, 0, n * sizeof(double)));
  if (int rc = batch_tiles(b)) return rc;
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  b->grid = 0;  // 0: one workgroup per tile; >0: persistent grid of that many workgroups
  *out = b;
  return ok();
}

extern "C" int omhd_batch_destroy(omhd_batch *b) {
  if (!b) return ok();
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (void *q : {(void *)b->arena, (void *)b->backup, (void *)b->mem_d, (void *)b->tiles_d, (void *)b->vmax_d,
                  (void *)b->err_d, (void *)b->par_d, (void *)b->dx_d, (void *)b->cfl_d, (void *)b->tsum_d,
                  (void *)b->tt_d, (void *)b->glm_tab_d, (void *)b->glm_x0_d})
    if (q) (void)hipFree(q);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return ok();
}

static int batch_check_err(omhd_batch *b) {
  int err[8];
  OMHD_HIP(hipMemcpyAsync(err, b->err_d, sizeof err, hipMemcpyDeviceToHost, b->stream));
  OMHD_HIP(hipStreamSynchronize(b->stream));
  if (err[0]) {
    err[0] &= ~ERR_IN_USTAR;
    return state_error(err);
  }
  if (err[4]) return set_error(err[4], err[4] == OMHD_ERR_PRESSURE ? "negative pressure" : "negative density");
  if (err[5]) return set_error(err[5], "non-finite signal speed");
  return OMHD_OK;
}

extern "C" int omhd_batch_upload(omhd_batch *b, int32_t i, const double *U) {
  if (!b || !U || i < 0 || i >= b->n) return set_error(OMHD_ERR_ARG, "bad member or null array");
  const Sweep &p = b->mem[i];
  const size_t nf = (size_t)(p.nx + 2) * (p.ny + 2);
  Staged a;
  if (int rc = a.init(U, 9 * nf, true, false, b->stream)) return rc;
  k_from_fortran<<<(unsigned)(((size_t)p.nx * p.ny + 255) / 256), 256, 0, b->stream>>>(a.dev, p.U, p.plane, p.zstride,
                                                                                        p.pitch, p.nx, p.ny, 1, 2);
  OMHD_HIP(hipGetLastError());
  a.release(b->stream);
  // first CFL speed of this member (main.f90:40)
  OMHD_HIP(hipMemsetAsync(b->err_d, 0, 8 * sizeof(int), b->stream));
  OMHD_HIP(hipMemsetAsync(b->vmax_d + i, 0, sizeof(unsigned long long), b->stream));
  k_cfl<2><<<(unsigned)(((size_t)p.nx * p.ny + 255) / 256), 256, 0, b->stream>>>(p.U, p.plane, p.zstride, p.pitch,
                                                                                  p.nx, p.ny, 1, b->vmax_d + i,
                                                                                  b->err_d);
  OMHD_HIP(hipGetLastError());
  unsigned long long v;
  OMHD_HIP(hipMemcpyAsync(&v, b->vmax_d + i, sizeof v, hipMemcpyDeviceToHost, b->stream));
  if (int rc = batch_check_err(b)) return rc;
  memcpy(&b->vmax[i], &v, sizeof(double));
  b->t[i] = 0.0;
  if (b->device_dt) {
    // device-dt mode keeps the step parameters on the device: this member's
    // next dt and GLM factor come from the speed just reduced (which also
    // clears its slot for the stage-2 reduction), and its time restarts at 0
    OMHD_HIP(hipMemsetAsync(b->tsum_d + i, 0, sizeof(double), b->stream));
    k_device_dt<<<(b->n + 63) / 64, 64, 0, b->stream>>>(b->par_d, b->vmax_d, b->dx_d, b->cfl_d, b->tt_d, nullptr,
                                                        b->glm_tab_d, b->glm_x0_d, b->err_d, b->n, i);
    OMHD_HIP(hipGetLastError());
    if (int rc = batch_check_err(b)) return rc;
  }
  b->uploaded[i] = 1;
  b->have_state = std::all_of(b->uploaded.begin(), b->uploaded.end(), [](char c) { return c != 0; });
  return ok();
}

extern "C" int omhd_batch_download(omhd_batch *b, int32_t i, double *U, double *V) {
  if (!b || !U || i < 0 || i >= b->n) return set_error(OMHD_ERR_ARG, "bad member or null array");
  const Sweep &p = b->mem[i];
  const size_t n = (size_t)(p.nx + 2) * (p.ny + 2);
  Faces f{};
  for (int a = 0; a < 3; a++) {
    f.bc[a] = p.bc[a];
    f.slo[a] = p.slo[a];
    f.shi[a] = p.shi[a];
  }
  Staged a, bb;
  if (int rc = a.init(U, 9 * n, false, true, b->stream)) return rc;
  if (V)
    if (int rc = bb.init(V, 4 * n, false, true, b->stream)) return rc;
  k_to_fortran<<<(unsigned)((n + 255) / 256), 256, 0, b->stream>>>(p.U, p.plane, p.zstride, p.pitch, p.nx, p.ny, 1, 2,
                                                                   f, a.dev, V ? bb.dev : nullptr);
  OMHD_HIP(hipGetLastError());
  if (int rc = a.finish(b->stream)) return rc;
  if (V)
    if (int rc = bb.finish(b->stream)) return rc;
  a.release(b->stream);
  bb.release(b->stream);
  OMHD_HIP(hipStreamSynchronize(b->stream));
  return ok();
}

// dt mode and graph capture.  Entering device dt mode seeds the device-side
// parameters from the current speeds.
extern "C" int omhd_batch_set_mode(omhd_batch *b, int32_t device_dt, int32_t use_graph) {
  if (!b) return set_error(OMHD_ERR_ARG, "null batch");
  if (use_graph && !device_dt) return set_error(OMHD_ERR_ARG, "graph capture needs device dt mode");
  if (device_dt && !b->device_dt) {
    std::vector<unsigned long long> bits(b->n);
    for (int i = 0; i < b->n; i++) memcpy(&bits[i], &b->vmax[i], sizeof(double));
    OMHD_HIP(hipMemcpy(b->vmax_d, bits.data(), b->n * sizeof(unsigned long long), hipMemcpyHostToDevice));
    OMHD_HIP(hipMemcpy(b->tsum_d, b->t.data(), b->n * sizeof(double), hipMemcpyHostToDevice));
    k_device_dt<<<(b->n + 63) / 64, 64, 0, b->stream>>>(b->par_d, b->vmax_d, b->dx_d, b->cfl_d, b->tt_d, nullptr,
                                                        b->glm_tab_d, b->glm_x0_d, b->err_d, b->n);
    OMHD_HIP(hipGetLastError());
    if (int rc = batch_check_err(b)) return rc;
  }
  if (!device_dt && b->device_dt) {  // back to host dt: fetch the speeds of the current states
    std::vector<DevPar> par(b->n);
    OMHD_HIP(hipMemcpy(par.data(), b->par_d, b->n * sizeof(DevPar), hipMemcpyDeviceToHost));
    OMHD_HIP(hipMemcpy(b->t.data(), b->tsum_d, b->n * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < b->n; i++) b->vmax[i] = par[i].ch;
  }
  b->device_dt = device_dt != 0;
  b->use_graph = use_graph != 0;
  return ok();
}

static int batch_launch_stages(omhd_batch *b) {
  const unsigned grid = b->grid > 0 ? (unsigned)b->grid : (unsigned)b->ntiles;
  hipLaunchKernelGGL(ens_kernel(b->lim, 1, b->tt), dim3(grid), dim3(64), 0, b->stream, b->mem_d, b->tiles_d,
                     b->ntiles);
  OMHD_HIP(hipGetLastError());
  hipLaunchKernelGGL(ens_kernel(b->lim, 2, b->tt), dim3(grid), dim3(64), 0, b->stream, b->mem_d, b->tiles_d,
                     b->ntiles);
  OMHD_HIP(hipGetLastError());
  return OMHD_OK;
}

__global__ void k_accumulate_t(double *tsum, const DevPar *par, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) tsum[i] += par[i].dt;
}

static int batch_device_step(omhd_batch *b) {
  // the parameters in par_d are those of this step; the dt kernel then
  // prepares the next step's from the speeds the stage-2 pass reduced
  k_accumulate_t<<<(b->n + 63) / 64, 64, 0, b->stream>>>(b->tsum_d, b->par_d, b->n);
  if (int rc = batch_launch_stages(b)) return rc;
  k_device_dt<<<(b->n + 63) / 64, 64, 0, b->stream>>>(b->par_d, b->vmax_d, b->dx_d, b->cfl_d, b->tt_d, nullptr,
                                                      b->glm_tab_d, b->glm_x0_d, b->err_d, b->n);
  OMHD_HIP(hipGetLastError());
  return OMHD_OK;
}

extern "C" int omhd_batch_step(omhd_batch *b, int32_t nsteps, double *dt_trace) {
  if (!b) return set_error(OMHD_ERR_ARG, "null batch");
  if (!b->have_state) return set_error(OMHD_ERR_ARG, "upload every member first");
  const double cr = 0.2;  // glm_ss2.f90:18
  if (b->device_dt) {
    if (b->use_graph && !b->gexec) {
      OMHD_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = batch_device_step(b);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc) return rc;
      if (e != hipSuccess) return set_error(OMHD_ERR_HIP, hipGetErrorString(e));
      b->graph = g;
      OMHD_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    }
    for (int s = 0; s < nsteps; s++) {
      if (b->use_graph) {
        OMHD_HIP(hipGraphLaunch(b->gexec, b->stream));
      } else if (int rc = batch_device_step(b)) {
        return rc;
      }
      if (dt_trace) {  // per-step trace costs a round trip
        OMHD_HIP(hipMemcpyAsync(b->par_h.data(), b->par_d, b->n * sizeof(DevPar), hipMemcpyDeviceToHost, b->stream));
        if (int rc = batch_check_err(b)) return rc;
        // par_d now holds the next step's parameters; the dt just taken is the
        // difference of the accumulated times
        std::vector<double> tnew(b->n);
        OMHD_HIP(hipMemcpy(tnew.data(), b->tsum_d, b->n * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < b->n; i++) dt_trace[(size_t)s * b->n + i] = tnew[i] - b->t[i];
        b->t = tnew;
      }
    }
    if (int rc = batch_check_err(b)) return rc;
    std::vector<DevPar> par(b->n);
    OMHD_HIP(hipMemcpy(par.data(), b->par_d, b->n * sizeof(DevPar), hipMemcpyDeviceToHost));
    OMHD_HIP(hipMemcpy(b->t.data(), b->tsum_d, b->n * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < b->n; i++) b->vmax[i] = par[i].ch;
    return ok();
  }
  for (int s = 0; s < nsteps; s++) {
    for (int i = 0; i < b->n; i++) {  // exactly omhd_step's host arithmetic, per member
      const omhd_config &c = b->cfg[i];
      const double ch = b->vmax[i];
      const double dt = c.cfl * c.dx / ch;
      if (!(dt >= 1.0e-7)) return set_error(OMHD_ERR_DT, "dt is too small", i, 0, 0, dt);
      DevPar &q = b->par_h[i];
      q.ch = ch;
      q.ch2 = ch * ch;
      q.dtx1 = (c.time_type == 0) ? dt / c.dx : 0.5 * dt / c.dx;
      q.dtx2 = dt / c.dx;
      q.glm_f = exp(-0.5 * dt * ch / cr);
      q.dt = dt;
    }
    OMHD_HIP(hipMemcpyAsync(b->par_d, b->par_h.data(), b->n * sizeof(DevPar), hipMemcpyHostToDevice, b->stream));
    OMHD_HIP(hipMemsetAsync(b->vmax_d, 0, b->n * sizeof(unsigned long long), b->stream));
    if (int rc = batch_launch_stages(b)) return rc;
    OMHD_HIP(hipMemcpyAsync(b->vmax_h.data(), b->vmax_d, b->n * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                            b->stream));
    if (int rc = batch_check_err(b)) return rc;
    for (int i = 0; i < b->n; i++) {
      memcpy(&b->vmax[i], &b->vmax_h[i], sizeof(double));
      b->t[i] += b->par_h[i].dt;
      if (dt_trace) dt_trace[(size_t)s * b->n + i] = b->par_h[i].dt;
    }
  }
  return ok();
}

extern "C" int omhd_batch_time(omhd_batch *b, double *t) {
  if (!b || !t) return set_error(OMHD_ERR_ARG, "null argument");
  for (int i = 0; i < b->n; i++) t[i] = b->t[i];
  return ok();
}

// Tile-shape / occupancy autotuner: times `trial_steps` steps for each
// candidate (rows per tile x launch shape), restoring the ensemble state after
// every trial, and keeps the fastest.  Returns the choice in rows/grid.
extern "C" int omhd_batch_autotune(omhd_batch *b, int32_t trial_steps, int32_t *rows_out, int32_t *grid_out) {
  if (!b || trial_steps < 1) return set_error(OMHD_ERR_ARG, "bad arguments");
  if (!b->have_state) return set_error(OMHD_ERR_ARG, "upload the members first");
  if (!b->backup) OMHD_HIP(hipMalloc((void **)&b->backup, b->arena_doubles * sizeof(double)));
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const int cand_rows[] = {8, 12, 16, 20, 24, 28, 32, 48};
  const int cand_grid[] = {0, 8 * cus, 16 * cus};  // one group per tile; persistent at 2 or 4 groups per SIMD
  const std::vector<double> vmax0 = b->vmax, t0 = b->t;
  const bool dmode = b->device_dt, gmode = b->use_graph;
  hipEvent_t e0, e1;
  OMHD_HIP(hipEventCreate(&e0));
  OMHD_HIP(hipEventCreate(&e1));
  OMHD_HIP(hipMemcpy(b->backup, b->arena, b->arena_doubles * sizeof(double), hipMemcpyDeviceToDevice));
  double best = 1e300;
  int best_rows = b->rows, best_grid = b->grid;
  for (int r : cand_rows)
    for (int g : cand_grid) {
      b->rows = r;
      b->grid = g;
      if (int rc = batch_tiles(b)) return rc;
      if (g > b->ntiles) continue;
      OMHD_HIP(hipEventRecord(e0, b->stream));
      if (int rc = omhd_batch_step(b, trial_steps, nullptr)) return rc;
      OMHD_HIP(hipEventRecord(e1, b->stream));
      OMHD_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      OMHD_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        best_rows = r;
        best_grid = g;
      }
      // restore the state the trial advanced
      OMHD_HIP(hipMemcpy(b->arena, b->backup, b->arena_doubles * sizeof(double), hipMemcpyDeviceToDevice));
      b->vmax = vmax0;
      b->t = t0;
      if (dmode) {
        b->device_dt = false;
        if (int rc = omhd_batch_set_mode(b, 1, gmode)) return rc;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->rows = best_rows;
  b->grid = best_grid;
  if (int rc = batch_tiles(b)) return rc;
  if (rows_out) *rows_out = best_rows;
  if (grid_out) *grid_out = best_grid;
  return ok();
}
