/*
 * omhd_oracle.c -- CPU restatement of OpenMHD's finite-volume step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the test oracle for the HIP library in
 * openmhd_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load it.  Nothing in the product links or calls it.
 *
 * Every routine follows the reference Fortran expression by expression, in the
 * same evaluation order (Fortran evaluates left to right; no FMA contraction:
 * build with -ffp-contract=off), so that its results are bitwise equal to the
 * flang-built reference (checked against the committed fixtures in
 * tests/golden/, which the reference's own routines produced).
 *
 * Arrays are Fortran column-major with the reference's single ghost layer:
 * U(ix,jx,kx,9), V(ix,jx,kx,4), VL/VR/F(ix,jx,kx,9).  2D is kx == 1.
 * Variable slots follow param.h:6-15 (1-based):
 *   mx/vx=1 my/vy=2 mz/vz=3 en/pr=4 ro=5 bx=6 by=7 bz=8 ps=9.
 *
 * Parity status per function (what pins it):
 *   or_u2v            flang vectors + closed form (uniform state)
 *   or_limiter        flang vectors incl. -0.0 and sign-mismatch cases
 *   or_flux_solver,   flang vectors for every HLLD branch x dir 1/2/3, and
 *   or_hlld, or_riemann  for the LLF, HLL, HLLC and LHLLD types x dir 1/2/3
 *   or_glm_ss2        flang vectors (single multiply)
 *   or_rk_*           flang vectors
 *   or_set_dt         flang vectors
 *   or_bc_*           flang vectors (periodic, KH walls)
 *   or_step / or_run  OT 200^2 MC 1000-step dt trace and final-state hash from
 *                     the flang build; KH walls 100 steps; 3D periodic 4 steps;
 *                     every flux type over 2D (periodic, KH walls) and 3D runs
 *   or_run1d          the flang build of 1D_basic's routines under the same
 *                     driver loop, bitwise for 8 (flux, limiter, integrator)
 *                     cases; known answers of the Brio-Wu tube (conservation
 *                     at the outflow faces, frozen Bx/psi/vz/Bz, mirror
 *                     symmetry of a symmetric tube)
 *   or_eta_faces,     closed form of the resistivity profile; the resistive dt
 *   or_set_dt_resistive  limit 0.5 cfl dx^2 / eta_max (flux_resistive.cuf:2-9)
 *   or_flux_resistive the reference's 3D form is CUDA Fortran only; pinned by
 *                     the flang build of its 2D CPU form on z-invariant fields
 *                     (2D_reconnection/flux_resistive.f90, which folds eta/dx
 *                     into the currents: equal to rounding, 2e-15 relative) and
 *                     by closed forms (uniform B -> no flux, linear B -> const J)
 *   or_bc_u/or_bc_face per-face sign tables: flang vectors of 2D_KH/bc.f90 and
 *                     2D_reconnection/bc.f90
 *   or_flux_resistive2d  the flang build of 2D_reconnection/flux_resistive.f90
 *   or_set_dt_resistive2d  (bitwise vectors, with that file's eta); whole runs
 *   (res_form 1)      of 2D_reconnection's loop against the project's driver
 *                     linked with flang against the reference's routines
 *   res_form 2        the 3D form above with set_dt_resistive's Rm1 limit
 *                     (3D_reconnection_gpu/flux_resistive.cuf:2-9, the same
 *                     expression as the 2D file's, pinned there)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "omhd_oracle.h"

/* param.h:3 */
static const double gam = 5.0 / 3.0;

/* 1-based Fortran indexing: A(i,j,k,m) in an (ix,jx,kx,*) array */
#define IX4(i, j, k, m) \
  ((size_t)((i) - 1) + (size_t)ix * ((size_t)((j) - 1) + (size_t)jx * ((size_t)((k) - 1) + (size_t)kx * (size_t)((m) - 1))))
#define IX3(i, j, k) ((size_t)((i) - 1) + (size_t)ix * ((size_t)((j) - 1) + (size_t)jx * (size_t)((k) - 1)))

enum { mx = 1, my = 2, mz = 3, en = 4, vx = 1, vy = 2, vz = 3, pr = 4, ro = 5, bx = 6, by = 7, bz = 8, ps = 9 };

static inline double fsign(double a, double b) { return copysign(fabs(a), b); }
static inline double fmax2(double a, double b) { return (a > b) ? a : b; }
static inline double fmin2(double a, double b) { return (a < b) ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------
 * u2v: 2D_basic/u2v.f90:1-46, 3D_basic/u2v.f90.
 * Returns 0, or OR_ERR_PRESSURE / OR_ERR_DENSITY (the reference prints the
 * minloc and stops, u2v.f90:34-43); *bad gets the 0-based linear cell index.
 * ---------------------------------------------------------------------- */
int or_u2v(const double *U, double *V, int ix, int jx, int kx, long *bad)
{
  const double f1 = gam - 1; /* u2v.f90:11 */
  double prmin = U[IX4(1, 1, 1, en)];
  double rhomin = U[IX4(1, 1, 1, ro)];
#pragma omp parallel for collapse(2) reduction(min : prmin, rhomin) schedule(static)
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        double rho = U[IX4(i, j, k, ro)];
        double v1 = U[IX4(i, j, k, mx)] / rho; /* :23 */
        double v2 = U[IX4(i, j, k, my)] / rho;
        double v3 = U[IX4(i, j, k, mz)] / rho;
        V[IX4(i, j, k, vx)] = v1;
        V[IX4(i, j, k, vy)] = v2;
        V[IX4(i, j, k, vz)] = v3;
        /* :24 dot_product(V, M), :25 dot_product(B, B) */
        double rv2 = v1 * U[IX4(i, j, k, mx)] + v2 * U[IX4(i, j, k, my)] + v3 * U[IX4(i, j, k, mz)];
        double b1 = U[IX4(i, j, k, bx)], b2 = U[IX4(i, j, k, by)], b3 = U[IX4(i, j, k, bz)];
        double B2 = b1 * b1 + b2 * b2 + b3 * b3;
        double p = f1 * (U[IX4(i, j, k, en)] - 0.5 * (rv2 + B2)); /* :26 */
        V[IX4(i, j, k, pr)] = p;
        prmin = fmin2(p, prmin);
        rhomin = fmin2(rho, rhomin);
      }
  size_t n = (size_t)ix * jx * kx;
  if (prmin <= 0) { /* :34-38: minloc of V(:,:,pr) */
    size_t best = 0;
    for (size_t c = 1; c < n; c++)
      if (V[3 * n + c] < V[3 * n + best]) best = c;
    if (bad) *bad = (long)best;
    return OR_ERR_PRESSURE;
  }
  if (rhomin <= 0) { /* :39-43 */
    size_t best = 0;
    for (size_t c = 1; c < n; c++)
      if (U[4 * n + c] < U[4 * n + best]) best = c;
    if (bad) *bad = (long)best;
    return OR_ERR_DENSITY;
  }
  return 0;
}

/* ------------------------------------------------------------------------
 * limiter: 2D_basic/limiter.f90:1-318 (dir 1,2), 3D_basic/limiter.f90 (dir 3).
 * One variable, one direction.  Along the line l=1..n (the direction's extent):
 *   wL(1) = wk(1); for l=2..n-1: grad from (wk(l-1),wk(l),wk(l+1)),
 *   wR(l-1) = wk(l) - grad, wL(l) = wk(l) + grad; wR(n-1) = wk(n).
 * Lines cover the reference's transverse ranges (min(2,ix)..max(1,ix-1) etc.).
 * ---------------------------------------------------------------------- */
static inline void lim_cell(int type, double wm, double w0, double wp, double *gR, double *gL)
{
  double gA = (w0 - wm);
  double gB = (wp - w0);
  double gC, grad;
  *gR = w0;
  *gL = w0;
  switch (type) {
  case 1: /* minmod, limiter.f90:84-98 */
    grad = (fsign(0.25, gA) + fsign(0.25, gB)) * fmin2(fabs(gA), fabs(gB));
    *gR = w0 - grad;
    *gL = w0 + grad;
    return;
  case 2: /* MC, limiter.f90:141-158 */
    gC = 0.25 * (wp - wm);
    grad = (fsign(0.5, gA) + fsign(0.5, gB)) * fmin2(fmin2(fabs(gA), fabs(gB)), fabs(gC));
    *gR = w0 - grad;
    *gL = w0 + grad;
    return;
  case 3: /* van Leer, limiter.f90:201-210 */
    if (gA * gB <= 0) grad = 0.0;
    else grad = gA * gB / (gA + gB);
    *gR = w0 - grad;
    *gL = w0 + grad;
    return;
  case 4: { /* Koren, limiter.f90:256-272; f1 = 1/6 at :22 */
    const double f1 = 1.0 / 6.0;
    gC = f1 * (2 * gA + gB);
    *gR = w0 - (fsign(0.5, gA) + fsign(0.5, gB)) * fmin2(fmin2(fabs(gA), fabs(gB)), fabs(gC));
    gC = f1 * (gA + 2 * gB);
    *gL = w0 + (fsign(0.5, gA) + fsign(0.5, gB)) * fmin2(fmin2(fabs(gA), fabs(gB)), fabs(gC));
    return;
  }
  }
}

int or_limiter(const double *wk, double *wL, double *wR, int ix, int jx, int kx, int dir, int type)
{
  if (type < 0 || type > 4) return OR_ERR_TYPE; /* limiter.f90:310-312 */
  if (dir < 1 || dir > 3) return OR_ERR_DIR;
  /* extents: line length n and stride s along dir; transverse ranges */
  int is = 1, ie = ix, js = 1, je = jx, ks = 1, ke = kx, n = 0;
  size_t s = 1;
  if (dir == 1) { js = imin(2, jx); je = imax(1, jx - 1); ks = imin(2, kx); ke = imax(1, kx - 1); n = ix; s = 1; ie = 1; }
  if (dir == 2) { is = imin(2, ix); ie = imax(1, ix - 1); ks = imin(2, kx); ke = imax(1, kx - 1); n = jx; s = (size_t)ix; je = 1; }
  if (dir == 3) { is = imin(2, ix); ie = imax(1, ix - 1); js = imin(2, jx); je = imax(1, jx - 1); n = kx; s = (size_t)ix * jx; ke = 1; }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = ks; k <= ke; k++)
    for (int j = js; j <= je; j++)
      for (int i = is; i <= ie; i++) {
        size_t b = IX3(i, j, k); /* l = 1 on this line */
        if (type == 0) { /* limiter.f90:56-70 */
          for (int l = 1; l <= n - 1; l++) {
            wL[b + (l - 1) * s] = wk[b + (l - 1) * s];
            wR[b + (l - 1) * s] = wk[b + l * s];
          }
          continue;
        }
        wL[b] = wk[b];
        for (int l = 2; l <= n - 1; l++) {
          double gR, gL;
          lim_cell(type, wk[b + (l - 2) * s], wk[b + (l - 1) * s], wk[b + l * s], &gR, &gL);
          wR[b + (l - 2) * s] = gR;
          wL[b + (l - 1) * s] = gL;
        }
        wR[b + (n - 2) * s] = wk[b + (n - 1) * s];
      }
  return 0;
}

/* ------------------------------------------------------------------------
 * HLLD flux at one interface: 2D_basic/flux_solver.f90:352-585 (type 3).
 * L and R are the 9 interface states (vx,vy,vz,p,ro,bx,by,bz,ps) [0-based
 * slots 0..8 = param.h slots 1..9].  d = dir-1 in 0..2.  F gets 9 values.
 * ---------------------------------------------------------------------- */
void or_hlld(const double *L, const double *R, double ch, int d, double *F, int *branch)
{
  const double r1 = 1.0 / (gam - 1.0); /* flux_solver.f90:38 */
  const double r2 = gam / (gam - 1.0); /* :39 */
  const double hllg_factor = 1.001;    /* :54 */
  /* flux_solver.f90:62-70 direction permutation, 0-based */
  const int vn = d % 3, vt1 = (d + 1) % 3, vt2 = (d + 2) % 3;
  const int mn = vn, mt1 = vt1, mt2 = vt2;
  const int bn = 5 + vn, bt1 = 5 + vt1, bt2 = 5 + vt2;
  const int RO = 4, PR = 3, EN = 3, PS = 8;
  double ch2 = ch * ch;
  double UL[4], UR[4], FL[9], FR[9];
  double UL1[9], UR1[9], U2[9], U_hll[9];
  double f1, f2, B2, v2, vBL, vBR, aL, aR, aM, at1, at2, ptL, ptR, pt, roL, roR, enL, enR;
  double aL1, aR1, aVL, aVR, vt1L, vt1R, vt2L, vt2R, roLs, roRs;
  for (int m = 0; m < 9; m++) { FL[m] = 0.0; FR[m] = 0.0; F[m] = 0.0; }
  int br = 0;

  /* VL --> UL  (:363-364) */
  v2 = L[0] * L[0] + L[1] * L[1] + L[2] * L[2];
  B2 = L[5] * L[5] + L[6] * L[6] + L[7] * L[7];
  vBL = L[0] * L[5] + L[1] * L[6] + L[2] * L[7];
  UL[0] = L[RO] * L[0];
  UL[1] = L[RO] * L[1];
  UL[2] = L[RO] * L[2];
  UL[EN] = 0.5 * (L[RO] * v2 + B2) + r1 * L[PR];
  /* VL --> FL  (:366-372) */
  FL[RO] = L[RO] * L[vn];
  FL[mn] = L[RO] * L[vn] * L[vn] - L[bn] * L[bn] + L[PR] + 0.5 * B2;
  FL[mt1] = L[RO] * L[vn] * L[vt1] - L[bn] * L[bt1];
  FL[mt2] = L[RO] * L[vn] * L[vt2] - L[bn] * L[bt2];
  FL[EN] = (0.5 * L[RO] * v2 + r2 * L[PR] + B2) * L[vn] - vBL * L[bn];
  FL[bt1] = L[vn] * L[bt1] - L[bn] * L[vt1];
  FL[bt2] = L[vn] * L[bt2] - L[bn] * L[vt2];
  /* fast mode^2, total pressure (:375-379) */
  f1 = gam * L[PR];
  f2 = 4 * (L[bn] * L[bn]);
  aL = ((f1 + B2) + sqrt(fmax2((f1 + B2) * (f1 + B2) - f1 * f2, 0.0))) / (2 * L[RO]);
  ptL = L[PR] + 0.5 * B2;

  /* VR --> UR, FR  (:385-401) */
  v2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
  B2 = R[5] * R[5] + R[6] * R[6] + R[7] * R[7];
  vBR = R[0] * R[5] + R[1] * R[6] + R[2] * R[7];
  UR[0] = R[RO] * R[0];
  UR[1] = R[RO] * R[1];
  UR[2] = R[RO] * R[2];
  UR[EN] = 0.5 * (R[RO] * v2 + B2) + r1 * R[PR];
  FR[RO] = R[RO] * R[vn];
  FR[mn] = R[RO] * R[vn] * R[vn] - R[bn] * R[bn] + R[PR] + 0.5 * B2;
  FR[mt1] = R[RO] * R[vn] * R[vt1] - R[bn] * R[bt1];
  FR[mt2] = R[RO] * R[vn] * R[vt2] - R[bn] * R[bt2];
  FR[EN] = (0.5 * R[RO] * v2 + r2 * R[PR] + B2) * R[vn] - vBR * R[bn];
  FR[bt1] = R[vn] * R[bt1] - R[bn] * R[vt1];
  FR[bt2] = R[vn] * R[bt2] - R[bn] * R[vt2];
  f1 = gam * R[PR];
  f2 = 4 * (R[bn] * R[bn]);
  aR = ((f1 + B2) + sqrt(fmax2((f1 + B2) * (f1 + B2) - f1 * f2, 0.0))) / (2 * R[RO]);
  ptR = R[PR] + 0.5 * B2;

  /* Riemann fan speed (:408-410) */
  f1 = sqrt(fmax2(aL, aR));
  aL = fmin2(L[vn], R[vn]) - f1;
  aR = fmax2(L[vn], R[vn]) + f1;

  if (aL >= 0) { /* :413-414 */
    for (int m = 0; m < 9; m++) F[m] = FL[m];
    br = OR_BR_FL;
  } else if (aR <= 0) { /* :416-417 */
    for (int m = 0; m < 9; m++) F[m] = FR[m];
    br = OR_BR_FR;
  } else {
    /* HLL state (:423-425) */
    f1 = 1.0 / (aR - aL);
    for (int m = 0; m < 3; m++) U_hll[m] = f1 * (aR * UR[m] - aL * UL[m] - FR[m] + FL[m]);
    for (int m = RO; m <= 7; m++) U_hll[m] = f1 * (aR * R[m] - aL * L[m] - FR[m] + FL[m]);
    /* entropy wave (:428) */
    aM = U_hll[mn] / U_hll[RO];
    /* total pressure, rho* (:431-433) */
    pt = ptL + L[RO] * (aL - L[vn]) * (aM - L[vn]);
    roL = L[RO] * (aL - L[vn]) / (aL - aM);
    roR = R[RO] * (aR - R[vn]) / (aR - aM);
    /* Alfven waves (:436-437) */
    aVL = fabs(U_hll[bn]) / sqrt(roL);
    aVR = fabs(U_hll[bn]) / sqrt(roR);

    if ((aL >= (aM - hllg_factor * aVL)) || ((aM + hllg_factor * aVR) >= aR)) {
      /* HLLC-G (:447-476) */
      f2 = aL * aR;
      F[bt1] = f1 * (aR * FL[bt1] - aL * FR[bt1] + f2 * (R[bt1] - L[bt1]));
      F[bt2] = f1 * (aR * FL[bt2] - aL * FR[bt2] + f2 * (R[bt2] - L[bt2]));
      at1 = U_hll[mt1] / U_hll[RO];
      at2 = U_hll[mt2] / U_hll[RO];
      enL = ((aL - L[vn]) * UL[EN] - ptL * L[vn] + pt * aM +
             U_hll[bn] * (vBL - aM * U_hll[bn] - at1 * U_hll[bt1] - at2 * U_hll[bt2])) / (aL - aM);
      enR = ((aR - R[vn]) * UR[EN] - ptR * R[vn] + pt * aM +
             U_hll[bn] * (vBR - aM * U_hll[bn] - at1 * U_hll[bt1] - at2 * U_hll[bt2])) / (aR - aM);
      f1 = 0.5 + fsign(0.5, aM);
      f2 = 1.0 - f1;
      F[mn] = f1 * (aL * (roL * aM - UL[mn]) + FL[mn]) + f2 * (aR * (roR * aM - UR[mn]) + FR[mn]);
      F[mt1] = f1 * (aL * (roL * at1 - UL[mt1]) + FL[mt1]) + f2 * (aR * (roR * at1 - UR[mt1]) + FR[mt1]);
      F[mt2] = f1 * (aL * (roL * at2 - UL[mt2]) + FL[mt2]) + f2 * (aR * (roR * at2 - UR[mt2]) + FR[mt2]);
      F[EN] = f1 * (aL * (enL - UL[EN]) + FL[EN]) + f2 * (aR * (enR - UR[EN]) + FR[EN]);
      F[RO] = f1 * (aL * (roL - L[RO]) + FL[RO]) + f2 * (aR * (roR - R[RO]) + FR[RO]);
      br = (aM >= 0) ? OR_BR_HLLC_L : OR_BR_HLLC_R;
    } else {
      /* L* (:482-493) */
      f1 = 1.0 / (L[RO] * (aL - L[vn]) * (aL - aM) - U_hll[bn] * U_hll[bn]);
      UL1[bt1] = L[bt1] * f1 * (L[RO] * ((aL - L[vn]) * (aL - L[vn])) - U_hll[bn] * U_hll[bn]);
      UL1[bt2] = L[bt2] * f1 * (L[RO] * ((aL - L[vn]) * (aL - L[vn])) - U_hll[bn] * U_hll[bn]);
      vt1L = L[vt1] - f1 * U_hll[bn] * L[bt1] * (aM - L[vn]);
      vt2L = L[vt2] - f1 * U_hll[bn] * L[bt2] * (aM - L[vn]);
      UL1[RO] = roL;
      UL1[EN] = ((aL - L[vn]) * UL[EN] - ptL * L[vn] + pt * aM +
                 U_hll[bn] * (vBL - aM * U_hll[bn] - vt1L * UL1[bt1] - vt2L * UL1[bt2])) / (aL - aM);
      UL1[mn] = roL * aM;
      UL1[mt1] = roL * vt1L;
      UL1[mt2] = roL * vt2L;
      /* R* (:496-507) */
      f1 = 1.0 / (R[RO] * (aR - R[vn]) * (aR - aM) - U_hll[bn] * U_hll[bn]);
      UR1[bt1] = R[bt1] * f1 * (R[RO] * ((aR - R[vn]) * (aR - R[vn])) - U_hll[bn] * U_hll[bn]);
      UR1[bt2] = R[bt2] * f1 * (R[RO] * ((aR - R[vn]) * (aR - R[vn])) - U_hll[bn] * U_hll[bn]);
      vt1R = R[vt1] - f1 * U_hll[bn] * R[bt1] * (aM - R[vn]);
      vt2R = R[vt2] - f1 * U_hll[bn] * R[bt2] * (aM - R[vn]);
      UR1[RO] = roR;
      UR1[EN] = ((aR - R[vn]) * UR[EN] - ptR * R[vn] + pt * aM +
                 U_hll[bn] * (vBR - aM * U_hll[bn] - vt1R * UR1[bt1] - vt2R * UR1[bt2])) / (aR - aM);
      UR1[mn] = roR * aM;
      UR1[mt1] = roR * vt1R;
      UR1[mt2] = roR * vt2R;
      /* rotational waves (:510-511) */
      aL1 = aM - aVL;
      aR1 = aM + aVR;
      if (aL1 >= 0) { /* F(L*) :514-519 */
        for (int m = 0; m < 4; m++) F[m] = aL * (UL1[m] - UL[m]) + FL[m];
        F[RO] = aL * (UL1[RO] - L[RO]) + FL[RO];
        F[bt1] = aL * (UL1[bt1] - L[bt1]) + FL[bt1];
        F[bt2] = aL * (UL1[bt2] - L[bt2]) + FL[bt2];
        br = OR_BR_LS;
      } else if (aR1 <= 0) { /* F(R*) :522-527 */
        for (int m = 0; m < 4; m++) F[m] = aR * (UR1[m] - UR[m]) + FR[m];
        F[RO] = aR * (UR1[RO] - R[RO]) + FR[RO];
        F[bt1] = aR * (UR1[bt1] - R[bt1]) + FR[bt1];
        F[bt2] = aR * (UR1[bt2] - R[bt2]) + FR[bt2];
        br = OR_BR_RS;
      } else { /* central states (:533-572) */
        roLs = sqrt(roL);
        roRs = sqrt(roR);
        f1 = 1.0 / (roLs + roRs);
        f2 = fsign(1.0, U_hll[bn]);
        at1 = f1 * (roLs * vt1L + roRs * vt1R + (UR1[bt1] - UL1[bt1]) * f2);
        at2 = f1 * (roLs * vt2L + roRs * vt2R + (UR1[bt2] - UL1[bt2]) * f2);
        U2[bt1] = f1 * (roLs * UR1[bt1] + roRs * UL1[bt1] + roLs * roRs * (vt1R - vt1L) * f2);
        U2[bt2] = f1 * (roLs * UR1[bt2] + roRs * UL1[bt2] + roLs * roRs * (vt2R - vt2L) * f2);
        if (aM >= 0) { /* F(L**) */
          U2[RO] = roL;
          U2[EN] = UL1[EN] - roLs * (vt1L * UL1[bt1] + vt2L * UL1[bt2] - at1 * U2[bt1] - at2 * U2[bt2]) * f2;
          U2[mn] = roL * aM;
          U2[mt1] = roL * at1;
          U2[mt2] = roL * at2;
          for (int m = 0; m < 4; m++) F[m] = aL1 * U2[m] - (aL1 - aL) * UL1[m] - aL * UL[m] + FL[m];
          F[RO] = aL1 * U2[RO] - (aL1 - aL) * UL1[RO] - aL * L[RO] + FL[RO];
          F[bt1] = aL1 * U2[bt1] - (aL1 - aL) * UL1[bt1] - aL * L[bt1] + FL[bt1];
          F[bt2] = aL1 * U2[bt2] - (aL1 - aL) * UL1[bt2] - aL * L[bt2] + FL[bt2];
          br = OR_BR_LSS;
        } else { /* F(R**) */
          U2[RO] = roR;
          U2[EN] = UR1[EN] + roRs * (vt1R * UR1[bt1] + vt2R * UR1[bt2] - at1 * U2[bt1] - at2 * U2[bt2]) * f2;
          U2[mn] = roR * aM;
          U2[mt1] = roR * at1;
          U2[mt2] = roR * at2;
          for (int m = 0; m < 4; m++) F[m] = aR1 * U2[m] - (aR1 - aR) * UR1[m] - aR * UR[m] + FR[m];
          F[RO] = aR1 * U2[RO] - (aR1 - aR) * UR1[RO] - aR * R[RO] + FR[RO];
          F[bt1] = aR1 * U2[bt1] - (aR1 - aR) * UR1[bt1] - aR * R[bt1] + FR[bt1];
          F[bt2] = aR1 * U2[bt2] - (aR1 - aR) * UR1[bt2] - aR * R[bt2] + FR[bt2];
          br = OR_BR_RSS;
        }
      }
    }
  }
  /* GLM flux pair (:580-581) */
  F[bn] = 0.5 * ((L[PS] + R[PS]) - ch * (R[bn] - L[bn]));
  F[PS] = 0.5 * (ch2 * (L[bn] + R[bn]) - ch * (R[PS] - L[PS]));
  if (branch) *branch = br;
}

/* ------------------------------------------------------------------------
 * The other flux types of flux_solver.f90 at one interface: 0 LLF (Rusanov,
 * :89-148), 1 HLL (:153-221), 2 HLLC (:226-347), 4 LHLLD (:590-789).
 * ---------------------------------------------------------------------- */

/* one side: conserved momenta/energy, physical flux, v.v, B.B, v.B
 * (flux_solver.f90:93-108 and the same statements of every type) */
static void or_side(const double *S, int vn, int vt1, int vt2, int bn, int bt1, int bt2, double *U, double *Fs,
                    double *v2o, double *B2o, double *vBo)
{
  const double r1 = 1.0 / (gam - 1.0), r2 = gam / (gam - 1.0);
  const int RO = 4, PR = 3, EN = 3;
  double v2 = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
  double B2 = S[5] * S[5] + S[6] * S[6] + S[7] * S[7];
  double vB = S[0] * S[5] + S[1] * S[6] + S[2] * S[7];
  U[0] = S[RO] * S[0];
  U[1] = S[RO] * S[1];
  U[2] = S[RO] * S[2];
  U[EN] = 0.5 * (S[RO] * v2 + B2) + r1 * S[PR];
  for (int m = 0; m < 9; m++) Fs[m] = 0.0;
  Fs[RO] = S[RO] * S[vn];
  Fs[vn] = S[RO] * S[vn] * S[vn] - S[bn] * S[bn] + S[PR] + 0.5 * B2;
  Fs[vt1] = S[RO] * S[vn] * S[vt1] - S[bn] * S[bt1];
  Fs[vt2] = S[RO] * S[vn] * S[vt2] - S[bn] * S[bt2];
  Fs[EN] = (0.5 * S[RO] * v2 + r2 * S[PR] + B2) * S[vn] - vB * S[bn];
  Fs[bt1] = S[vn] * S[bt1] - S[bn] * S[vt1];
  Fs[bt2] = S[vn] * S[bt2] - S[bn] * S[vt2];
  *v2o = v2;
  *B2o = B2;
  *vBo = vB;
}

/* fast-mode speed^2 with the max(.,0) guard (types 1-4) */
static double or_fast2(const double *S, double B2, int bn)
{
  double f1 = gam * S[3];
  double f2 = 4 * (S[bn] * S[bn]);
  return ((f1 + B2) + sqrt(fmax2((f1 + B2) * (f1 + B2) - f1 * f2, 0.0))) / (2 * S[4]);
}

void or_riemann(int type, const double *L, const double *R, double ch, int d, double *F)
{
  if (type == 3) {
    or_hlld(L, R, ch, d, F, NULL);
    return;
  }
  const int vn = d % 3, vt1 = (d + 1) % 3, vt2 = (d + 2) % 3;
  const int mn = vn, mt1 = vt1, mt2 = vt2;
  const int bn = 5 + vn, bt1 = 5 + vt1, bt2 = 5 + vt2;
  const int RO = 4, EN = 3, PS = 8;
  double UL[4], UR[4], FL[9], FR[9], v2L, B2L, vBL, v2R, B2R, vBR;
  or_side(L, vn, vt1, vt2, bn, bt1, bt2, UL, FL, &v2L, &B2L, &vBL);
  or_side(R, vn, vt1, vt2, bn, bt1, bt2, UR, FR, &v2R, &B2R, &vBR);
  for (int m = 0; m < 9; m++) F[m] = 0.0;
  if (type == 0) { /* LLF (:110-143) */
    double f1 = gam * L[3], f2 = 4 * (L[bn] * L[bn]);
    double aL = sqrt(((f1 + B2L) + sqrt((f1 + B2L) * (f1 + B2L) - f1 * f2)) / (2 * L[RO]));
    f1 = gam * R[3];
    f2 = 4 * (R[bn] * R[bn]);
    double aR = sqrt(((f1 + B2R) + sqrt((f1 + B2R) * (f1 + B2R) - f1 * f2)) / (2 * R[RO]));
    f1 = fmax2(fabs(L[vn]) + aL, fabs(R[vn]) + aR);
    for (int m = 0; m < 4; m++) F[m] = 0.5 * (FL[m] + FR[m] - f1 * (UR[m] - UL[m]));
    F[RO] = 0.5 * (FL[RO] + FR[RO] - f1 * (R[RO] - L[RO]));
    F[bt1] = 0.5 * (FL[bt1] + FR[bt1] - f1 * (R[bt1] - L[bt1]));
    F[bt2] = 0.5 * (FL[bt2] + FR[bt2] - f1 * (R[bt2] - L[bt2]));
  } else if (type == 1 || type == 2) { /* HLL (:172-216), HLLC (:247-342) */
    double aL = or_fast2(L, B2L, bn), aR = or_fast2(R, B2R, bn);
    const double ptL = L[3] + 0.5 * B2L, ptR = R[3] + 0.5 * B2R;
    double f1 = sqrt(fmax2(aL, aR));
    aL = fmin2(fmin2(L[vn], R[vn]) - f1, 0.0);
    aR = fmax2(fmax2(L[vn], R[vn]) + f1, 0.0);
    f1 = 1.0 / (aR - aL);
    double f2 = aL * aR;
    if (type == 1) {
      for (int m = 0; m < 4; m++) F[m] = f1 * (aR * FL[m] - aL * FR[m] + f2 * (UR[m] - UL[m]));
      F[RO] = f1 * (aR * FL[RO] - aL * FR[RO] + f2 * (R[RO] - L[RO]));
    } else {
      double Uh[9];
      for (int m = 0; m < 3; m++) Uh[m] = f1 * (aR * UR[m] - aL * UL[m] - FR[m] + FL[m]);
      for (int m = RO; m <= 7; m++) Uh[m] = f1 * (aR * R[m] - aL * L[m] - FR[m] + FL[m]);
      F[bt1] = f1 * (aR * FL[bt1] - aL * FR[bt1] + f2 * (R[bt1] - L[bt1]));
      F[bt2] = f1 * (aR * FL[bt2] - aL * FR[bt2] + f2 * (R[bt2] - L[bt2]));
      f2 = 1 / Uh[RO];
      const double aM = Uh[mn] * f2, at1 = Uh[mt1] * f2, at2 = Uh[mt2] * f2;
      const double pt = ptL + L[RO] * (aL - L[vn]) * (aM - L[vn]);
      const double roL = L[RO] * (aL - L[vn]) / (aL - aM);
      const double enL = ((aL - L[vn]) * UL[EN] - ptL * L[vn] + pt * aM +
                          Uh[bn] * (vBL - aM * Uh[bn] - at1 * Uh[bt1] - at2 * Uh[bt2])) / (aL - aM);
      const double roR = R[RO] * (aR - R[vn]) / (aR - aM);
      const double enR = ((aR - R[vn]) * UR[EN] - ptR * R[vn] + pt * aM +
                          Uh[bn] * (vBR - aM * Uh[bn] - at1 * Uh[bt1] - at2 * Uh[bt2])) / (aR - aM);
      f1 = 0.5 + fsign(0.5, aM);
      f2 = 1.0 - f1;
      F[mn] = f1 * (aL * (roL * aM - UL[mn]) + FL[mn]) + f2 * (aR * (roR * aM - UR[mn]) + FR[mn]);
      F[mt1] = f1 * (aL * (roL * at1 - UL[mt1]) + FL[mt1]) + f2 * (aR * (roR * at1 - UR[mt1]) + FR[mt1]);
      F[mt2] = f1 * (aL * (roL * at2 - UL[mt2]) + FL[mt2]) + f2 * (aR * (roR * at2 - UR[mt2]) + FR[mt2]);
      F[EN] = f1 * (aL * (enL - UL[EN]) + FL[EN]) + f2 * (aR * (enR - UR[EN]) + FR[EN]);
      F[RO] = f1 * (aL * (roL - L[RO]) + FL[RO]) + f2 * (aR * (roR - R[RO]) + FR[RO]);
    }
    if (type == 1) {
      F[bt1] = f1 * (aR * FL[bt1] - aL * FR[bt1] + f2 * (R[bt1] - L[bt1]));
      F[bt2] = f1 * (aR * FL[bt2] - aL * FR[bt2] + f2 * (R[bt2] - L[bt2]));
    }
  } else { /* LHLLD (:590-789) */
    double aL = or_fast2(L, B2L, bn), aR = or_fast2(R, B2R, bn);
    const double ptL = L[3] + 0.5 * B2L, ptR = R[3] + 0.5 * B2R;
    /* the second fast-mode estimate uses rho v^2 in place of gamma p */
    double f1 = L[RO] * v2L, f2 = 4 * (L[bn] * L[bn]);
    double aL1 = ((f1 + B2L) + sqrt(fmax2((f1 + B2L) * (f1 + B2L) - f1 * f2, 0.0))) / (2 * L[RO]);
    f1 = R[RO] * v2R;
    f2 = 4 * (R[bn] * R[bn]);
    double aR1 = ((f1 + B2R) + sqrt(fmax2((f1 + B2R) * (f1 + B2R) - f1 * f2, 0.0))) / (2 * R[RO]);
    f1 = fmin2(1.0, fmax2(aL1, aR1) / fmax2(aL, aR)); /* chi^2 */
    f2 = 2 * sqrt(f1) - f1;                            /* psi */
    const double psi = f2;
    f1 = sqrt(fmax2(aL, aR));
    aL = fmin2(L[vn], R[vn]) - f1;
    aR = fmax2(L[vn], R[vn]) + f1;
    if (aL >= 0) {
      for (int m = 0; m < 9; m++) F[m] = FL[m];
    } else if (aR <= 0) {
      for (int m = 0; m < 9; m++) F[m] = FR[m];
    } else {
      double Uh[9], UL1[9], UR1[9], U2[9];
      f1 = 1.0 / (aR - aL);
      for (int m = 0; m < 3; m++) Uh[m] = f1 * (aR * UR[m] - aL * UL[m] - FR[m] + FL[m]);
      for (int m = RO; m <= 7; m++) Uh[m] = f1 * (aR * R[m] - aL * L[m] - FR[m] + FL[m]);
      const double aM = Uh[mn] / Uh[RO];
      const double pt = (R[RO] * ptL * (aR - R[vn]) - L[RO] * ptR * (aL - L[vn]) +
                         psi * L[RO] * R[RO] * (aR - R[vn]) * (aL - L[vn]) * (R[vn] - L[vn])) /
                        (R[RO] * (aR - R[vn]) - L[RO] * (aL - L[vn]));
      const double roL = L[RO] * (aL - L[vn]) / (aL - aM);
      const double roR = R[RO] * (aR - R[vn]) / (aR - aM);
      const double aVL = fabs(Uh[bn]) / sqrt(roL);
      const double aVR = fabs(Uh[bn]) / sqrt(roR);
      f1 = 1.0 / (L[RO] * (aL - L[vn]) * (aL - aM) - Uh[bn] * Uh[bn]);
      UL1[bt1] = L[bt1] * f1 * (L[RO] * ((aL - L[vn]) * (aL - L[vn])) - Uh[bn] * Uh[bn]);
      UL1[bt2] = L[bt2] * f1 * (L[RO] * ((aL - L[vn]) * (aL - L[vn])) - Uh[bn] * Uh[bn]);
      const double vt1L = L[vt1] - f1 * Uh[bn] * L[bt1] * (aM - L[vn]);
      const double vt2L = L[vt2] - f1 * Uh[bn] * L[bt2] * (aM - L[vn]);
      UL1[RO] = roL;
      UL1[EN] = ((aL - L[vn]) * UL[EN] - ptL * L[vn] + pt * aM +
                 Uh[bn] * (vBL - aM * Uh[bn] - vt1L * UL1[bt1] - vt2L * UL1[bt2])) / (aL - aM);
      UL1[mn] = roL * aM;
      UL1[mt1] = roL * vt1L;
      UL1[mt2] = roL * vt2L;
      f1 = 1.0 / (R[RO] * (aR - R[vn]) * (aR - aM) - Uh[bn] * Uh[bn]);
      UR1[bt1] = R[bt1] * f1 * (R[RO] * ((aR - R[vn]) * (aR - R[vn])) - Uh[bn] * Uh[bn]);
      UR1[bt2] = R[bt2] * f1 * (R[RO] * ((aR - R[vn]) * (aR - R[vn])) - Uh[bn] * Uh[bn]);
      const double vt1R = R[vt1] - f1 * Uh[bn] * R[bt1] * (aM - R[vn]);
      const double vt2R = R[vt2] - f1 * Uh[bn] * R[bt2] * (aM - R[vn]);
      UR1[RO] = roR;
      UR1[EN] = ((aR - R[vn]) * UR[EN] - ptR * R[vn] + pt * aM +
                 Uh[bn] * (vBR - aM * Uh[bn] - vt1R * UR1[bt1] - vt2R * UR1[bt2])) / (aR - aM);
      UR1[mn] = roR * aM;
      UR1[mt1] = roR * vt1R;
      UR1[mt2] = roR * vt2R;
      aL1 = aM - aVL;
      aR1 = aM + aVR;
      if (aL1 >= 0) {
        for (int m = 0; m < 4; m++) F[m] = aL * (UL1[m] - UL[m]) + FL[m];
        F[RO] = aL * (UL1[RO] - L[RO]) + FL[RO];
        F[bt1] = aL * (UL1[bt1] - L[bt1]) + FL[bt1];
        F[bt2] = aL * (UL1[bt2] - L[bt2]) + FL[bt2];
      } else if (aR1 <= 0) {
        for (int m = 0; m < 4; m++) F[m] = aR * (UR1[m] - UR[m]) + FR[m];
        F[RO] = aR * (UR1[RO] - R[RO]) + FR[RO];
        F[bt1] = aR * (UR1[bt1] - R[bt1]) + FR[bt1];
        F[bt2] = aR * (UR1[bt2] - R[bt2]) + FR[bt2];
      } else {
        const double roLs = sqrt(roL), roRs = sqrt(roR);
        f1 = 1.0 / (roLs + roRs);
        f2 = fsign(1.0, Uh[bn]);
        const double at1 = f1 * (roLs * vt1L + roRs * vt1R + (UR1[bt1] - UL1[bt1]) * f2);
        const double at2 = f1 * (roLs * vt2L + roRs * vt2R + (UR1[bt2] - UL1[bt2]) * f2);
        U2[bt1] = f1 * (roLs * UR1[bt1] + roRs * UL1[bt1] + roLs * roRs * (vt1R - vt1L) * f2);
        U2[bt2] = f1 * (roLs * UR1[bt2] + roRs * UL1[bt2] + roLs * roRs * (vt2R - vt2L) * f2);
        if (aM >= 0) {
          U2[RO] = roL;
          U2[EN] = UL1[EN] - roLs * (vt1L * UL1[bt1] + vt2L * UL1[bt2] - at1 * U2[bt1] - at2 * U2[bt2]) * f2;
          U2[mn] = roL * aM;
          U2[mt1] = roL * at1;
          U2[mt2] = roL * at2;
          for (int m = 0; m < 4; m++) F[m] = aL1 * U2[m] - (aL1 - aL) * UL1[m] - aL * UL[m] + FL[m];
          F[RO] = aL1 * U2[RO] - (aL1 - aL) * UL1[RO] - aL * L[RO] + FL[RO];
          F[bt1] = aL1 * U2[bt1] - (aL1 - aL) * UL1[bt1] - aL * L[bt1] + FL[bt1];
          F[bt2] = aL1 * U2[bt2] - (aL1 - aL) * UL1[bt2] - aL * L[bt2] + FL[bt2];
        } else {
          U2[RO] = roR;
          U2[EN] = UR1[EN] + roRs * (vt1R * UR1[bt1] + vt2R * UR1[bt2] - at1 * U2[bt1] - at2 * U2[bt2]) * f2;
          U2[mn] = roR * aM;
          U2[mt1] = roR * at1;
          U2[mt2] = roR * at2;
          for (int m = 0; m < 4; m++) F[m] = aR1 * U2[m] - (aR1 - aR) * UR1[m] - aR * UR[m] + FR[m];
          F[RO] = aR1 * U2[RO] - (aR1 - aR) * UR1[RO] - aR * R[RO] + FR[RO];
          F[bt1] = aR1 * U2[bt1] - (aR1 - aR) * UR1[bt1] - aR * R[bt1] + FR[bt1];
          F[bt2] = aR1 * U2[bt2] - (aR1 - aR) * UR1[bt2] - aR * R[bt2] + FR[bt2];
        }
      }
    }
  }
  /* GLM flux pair, every type (:141-142 and the same lines of each type) */
  F[bn] = 0.5 * ((L[PS] + R[PS]) - ch * (R[bn] - L[bn]));
  F[PS] = 0.5 * (ch * ch * (L[bn] + R[bn]) - ch * (R[PS] - L[PS]));
}

/* flux_solver(F,VL,VR,ch,ix,jx[,kx],dir,type): flux_solver.f90:1-81 frame.
 * The CPU build zeroes F first (:57); interfaces covered: along dir 1..n-1,
 * transverse min(2,.)..max(1,.-1). */
int or_flux_solver(double *F, const double *VL, const double *VR, double ch, int ix, int jx, int kx, int dir, int type,
                   int *branches)
{
  if (dir < 1 || dir > 3) return OR_ERR_DIR;
  if (type < 0 || type > 4) return OR_ERR_TYPE;
  size_t n = (size_t)ix * jx * kx;
  memset(F, 0, sizeof(double) * 9 * n);
  int is = imin(2, ix), ie = imax(1, ix - 1), js = imin(2, jx), je = imax(1, jx - 1), ks = imin(2, kx), ke = imax(1, kx - 1);
  if (dir == 1) { is = 1; ie = ix - 1; }
  if (dir == 2) { js = 1; je = jx - 1; }
  if (dir == 3) { ks = 1; ke = kx - 1; }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = ks; k <= ke; k++)
    for (int j = js; j <= je; j++)
      for (int i = is; i <= ie; i++) {
        size_t c = IX3(i, j, k);
        double L[9], R[9], f[9];
        for (int m = 0; m < 9; m++) { L[m] = VL[c + m * n]; R[m] = VR[c + m * n]; }
        int br = 0;
        if (type == 3) or_hlld(L, R, ch, dir - 1, f, &br);
        else or_riemann(type, L, R, ch, dir - 1, f);
        for (int m = 0; m < 9; m++) F[c + m * n] = f[m];
        if (branches) branches[c] = br;
      }
  return 0;
}

/* glm_ss2: 2D_basic/glm_ss2.f90:1-39.  psi <- f*psi on the WHOLE array. */
void or_glm_ss2(double *U, double ch, double dt, int ix, int jx, int kx)
{
  const double cr = 0.2; /* :18 */
  double f1 = exp(-0.5 * dt * ch / cr);
  size_t n = (size_t)ix * jx * kx;
  double *psi = U + 8 * n;
  for (size_t c = 0; c < n; c++) psi[c] = f1 * psi[c];
}

/* rk_tvd21/22, rk_std21/22: 2D_basic/rk.f90:9-132 (wH == NULL), and
 * 3D_basic/rk.f90:9-148 (wH != NULL).  mode: 0 tvd21, 1 tvd22, 2 std21, 3 std22.
 * For std22 wk1 is unused. */
void or_rk(int mode, double *wk, double *wk1, const double *wF, const double *wG, const double *wH, double dt, double dx,
           int ix, int jx, int kx)
{
  double dtx = (mode == 2) ? 0.5 * dt / dx : dt / dx;
  int ks = 2, ke = kx - 1;
  if (!wH) { ks = 1; ke = 1; }
#pragma omp parallel for collapse(3) schedule(static)
  for (int m = 1; m <= 9; m++)
    for (int k = ks; k <= ke; k++)
      for (int j = 2; j <= jx - 1; j++)
        for (int i = 2; i <= ix - 1; i++) {
          double d;
          if (wH)
            d = wF[IX4(i - 1, j, k, m)] - wF[IX4(i, j, k, m)] + wG[IX4(i, j - 1, k, m)] - wG[IX4(i, j, k, m)] +
                wH[IX4(i, j, k - 1, m)] - wH[IX4(i, j, k, m)];
          else
            d = wF[IX4(i - 1, j, k, m)] - wF[IX4(i, j, k, m)] + wG[IX4(i, j - 1, k, m)] - wG[IX4(i, j, k, m)];
          size_t c = IX4(i, j, k, m);
          if (mode == 0 || mode == 2) wk1[c] = wk[c] + dtx * d;
          else if (mode == 1) wk[c] = 0.5 * (wk[c] + wk1[c] + dtx * d);
          else wk[c] = wk[c] + dtx * d;
        }
}

/* set_dt: 2D_basic/set_dt.f90:1-68, 3D_basic/set_dt.f90 (ndim == 3).
 * Returns OR_ERR_DT when dt < dtmin (1e-7, :7) with *bad = argmax cell. */
int or_set_dt(const double *U, const double *V, double *vmax_out, double *dt_out, double dx, double cfl, int ix, int jx,
              int kx, int ndim, long *bad)
{
  const double dtmin = 1.0e-7;
  int is = imin(2, ix), ie = imax(1, ix - 1), js = imin(2, jx), je = imax(1, jx - 1), ks = imin(2, kx), ke = imax(1, kx - 1);
  double vmax = -1.0;
  long arg = 0;
  for (int k = ks; k <= ke; k++)
    for (int j = js; j <= je; j++)
      for (int i = is; i <= ie; i++) {
        double b1 = U[IX4(i, j, k, bx)], b2 = U[IX4(i, j, k, by)], b3 = U[IX4(i, j, k, bz)];
        double B2 = b1 * b1 + b2 * b2 + b3 * b3;
        double f1 = gam * V[IX4(i, j, k, pr)];
        double rho2 = 2 * U[IX4(i, j, k, ro)];
        double f2 = 4 * f1 * (b1 * b1);
        double vfx = sqrt(((f1 + B2) + sqrt(fmax2((f1 + B2) * (f1 + B2) - f2, 0.0))) / rho2);
        f2 = 4 * f1 * (b2 * b2);
        double vfy = sqrt(((f1 + B2) + sqrt(fmax2((f1 + B2) * (f1 + B2) - f2, 0.0))) / rho2);
        double vt = fmax2(fabs(V[IX4(i, j, k, vx)]) + vfx, fabs(V[IX4(i, j, k, vy)]) + vfy);
        if (ndim == 3) {
          f2 = 4 * f1 * (b3 * b3);
          double vfz = sqrt(((f1 + B2) + sqrt(fmax2((f1 + B2) * (f1 + B2) - f2, 0.0))) / rho2);
          vt = fmax2(vt, fabs(V[IX4(i, j, k, vz)]) + vfz);
        }
        if (vt > vmax) { vmax = vt; arg = (long)IX3(i, j, k); }
      }
  double dt = cfl * dx / vmax; /* :52 */
  *vmax_out = vmax;
  *dt_out = dt;
  if (dt < dtmin) {
    if (bad) *bad = arg;
    return OR_ERR_DT;
  }
  return 0;
}

/* ------------------------------------------------------------------------
 * Boundaries.  A face is periodic (wrap) or a mirror with a sign table of
 * 9 entries (+1 copy, -1 negate), as in 2D_basic/bc.f90:9-14 (periodic) and
 * 2D_KH/bc.f90:6-38 (KH walls: my, By negated).  x faces first, then y, then
 * z, which fixes the corners the same way the reference does.
 * sign tables are indexed by param.h slot-1 for U and by the VL/VR slot for
 * interface states (vx,vy,vz,p,ro,bx,by,bz,ps).
 * ---------------------------------------------------------------------- */
void or_bc_u(double *U, int ix, int jx, int kx, const int *bctype, const double *sgn)
{
  /* bctype[2*d + s]: 0 periodic, 1 mirror; sgn[(2*d+s)*9 + m] */
  size_t n = (size_t)ix * jx * kx;
  for (int m = 0; m < 9; m++) {
    double *A = U + m * n;
    /* x */
    for (int k = 1; k <= kx; k++)
      for (int j = 1; j <= jx; j++) {
        if (bctype[0] == 0) {
          A[IX3(ix, j, k)] = A[IX3(2, j, k)];
          A[IX3(1, j, k)] = A[IX3(ix - 1, j, k)];
        } else {
          A[IX3(1, j, k)] = sgn[0 * 9 + m] * A[IX3(2, j, k)];
          A[IX3(ix, j, k)] = sgn[1 * 9 + m] * A[IX3(ix - 1, j, k)];
        }
      }
    if (jx > 1)
      for (int k = 1; k <= kx; k++)
        for (int i = 1; i <= ix; i++) {
          if (bctype[2] == 0) {
            A[IX3(i, jx, k)] = A[IX3(i, 2, k)];
            A[IX3(i, 1, k)] = A[IX3(i, jx - 1, k)];
          } else {
            A[IX3(i, 1, k)] = sgn[2 * 9 + m] * A[IX3(i, 2, k)];
            A[IX3(i, jx, k)] = sgn[3 * 9 + m] * A[IX3(i, jx - 1, k)];
          }
        }
    if (kx > 1)
      for (int j = 1; j <= jx; j++)
        for (int i = 1; i <= ix; i++) {
          if (bctype[4] == 0) {
            A[IX3(i, j, kx)] = A[IX3(i, j, 2)];
            A[IX3(i, j, 1)] = A[IX3(i, j, kx - 1)];
          } else {
            A[IX3(i, j, 1)] = sgn[4 * 9 + m] * A[IX3(i, j, 2)];
            A[IX3(i, j, kx)] = sgn[5 * 9 + m] * A[IX3(i, j, kx - 1)];
          }
        }
  }
}

/* Interface-state face patch along dir: periodic 2D_basic/main.f90:101-102,
 * 119-120; mirror 2D_KH/bc.f90:42-83 (bc_for_F / bc_for_G). */
void or_bc_face(double *VL, double *VR, int ix, int jx, int kx, int dir, int bctype, const double *sgn_lo,
                const double *sgn_hi)
{
  size_t n = (size_t)ix * jx * kx;
  for (int m = 0; m < 9; m++) {
    double *L = VL + m * n, *R = VR + m * n;
    if (dir == 1)
      for (int k = 1; k <= kx; k++)
        for (int j = 1; j <= jx; j++) {
          if (bctype == 0) {
            R[IX3(ix - 1, j, k)] = R[IX3(1, j, k)];
            L[IX3(1, j, k)] = L[IX3(ix - 1, j, k)];
          } else {
            L[IX3(1, j, k)] = sgn_lo[m] * R[IX3(1, j, k)];
            R[IX3(ix - 1, j, k)] = sgn_hi[m] * L[IX3(ix - 1, j, k)];
          }
        }
    if (dir == 2)
      for (int k = 1; k <= kx; k++)
        for (int i = 1; i <= ix; i++) {
          if (bctype == 0) {
            L[IX3(i, 1, k)] = L[IX3(i, jx - 1, k)];
            R[IX3(i, jx - 1, k)] = R[IX3(i, 1, k)];
          } else {
            L[IX3(i, 1, k)] = sgn_lo[m] * R[IX3(i, 1, k)];
            R[IX3(i, jx - 1, k)] = sgn_hi[m] * L[IX3(i, jx - 1, k)];
          }
        }
    if (dir == 3)
      for (int j = 1; j <= jx; j++)
        for (int i = 1; i <= ix; i++) {
          if (bctype == 0) {
            R[IX3(i, j, kx - 1)] = R[IX3(i, j, 1)];
            L[IX3(i, j, 1)] = L[IX3(i, j, kx - 1)];
          } else {
            L[IX3(i, j, 1)] = sgn_lo[m] * R[IX3(i, j, 1)];
            R[IX3(i, j, kx - 1)] = sgn_hi[m] * L[IX3(i, j, kx - 1)];
          }
        }
  }
}

/* ------------------------------------------------------------------------
 * Resistive flux, 3D_reconnection_gpu/flux_resistive.cuf:86-135 (arithmetic
 * of the reference's only 3D form; the 2D CPU form folds eta/dx into J and
 * orders operations differently, 2D_reconnection/flux_resistive.f90:74-84).
 * Adds eta*J terms to F(en), F(bt1), F(bt2) at the interfaces of the
 * reference's loop range.  eta is given per face as an (ix,jx) array
 * (eta(x,y) at the face centre; it is uniform in z), computed on the host.
 * ---------------------------------------------------------------------- */
void or_flux_resistive(double *F, const double *U, const double *eta_face, double dx, int ix, int jx, int kx, int dir)
{
  int is = 2, ie = ix - 1, js = 2, je = jx - 1, ks = 2, ke = kx - 1;
  if (dir == 1) { is = 1; ie = ix - 1; }
  if (dir == 2) { js = 1; je = jx - 1; }
  if (dir == 3) { ks = 1; ke = kx - 1; }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = ks; k <= ke; k++)
    for (int j = js; j <= je; j++)
      for (int i = is; i <= ie; i++) {
        const double eta_dx = eta_face[(size_t)(i - 1) + (size_t)ix * (size_t)(j - 1)] / dx;
#define B(ii, jj, kk, m) U[IX4(ii, jj, kk, m)]
        if (dir == 1) {
          double JyS = -(B(i + 1, j, k, bz) - B(i, j, k, bz)) +
                       0.25 * (B(i, j, k + 1, bx) + B(i + 1, j, k + 1, bx) - B(i, j, k - 1, bx) - B(i + 1, j, k - 1, bx));
          double JzS = (B(i + 1, j, k, by) - B(i, j, k, by)) -
                       0.25 * (B(i, j + 1, k, bx) + B(i + 1, j + 1, k, bx) - B(i, j - 1, k, bx) - B(i + 1, j - 1, k, bx));
          F[IX4(i, j, k, en)] = F[IX4(i, j, k, en)] + 0.5 * eta_dx *
                                (JyS * (B(i, j, k, bz) + B(i + 1, j, k, bz)) - JzS * (B(i, j, k, by) + B(i + 1, j, k, by)));
          F[IX4(i, j, k, by)] = F[IX4(i, j, k, by)] - eta_dx * JzS;
          F[IX4(i, j, k, bz)] = F[IX4(i, j, k, bz)] + eta_dx * JyS;
        } else if (dir == 2) {
          double JxS = (B(i, j + 1, k, bz) - B(i, j, k, bz)) -
                       0.25 * (B(i, j, k + 1, by) + B(i, j + 1, k + 1, by) - B(i, j, k - 1, by) - B(i, j + 1, k - 1, by));
          double JzS = -(B(i, j + 1, k, bx) - B(i, j, k, bx)) +
                       0.25 * (B(i + 1, j, k, by) + B(i + 1, j + 1, k, by) - B(i - 1, j, k, by) - B(i - 1, j + 1, k, by));
          F[IX4(i, j, k, en)] = F[IX4(i, j, k, en)] + 0.5 * eta_dx *
                                (JzS * (B(i, j, k, bx) + B(i, j + 1, k, bx)) - JxS * (B(i, j, k, bz) + B(i, j + 1, k, bz)));
          F[IX4(i, j, k, bx)] = F[IX4(i, j, k, bx)] + eta_dx * JzS;
          F[IX4(i, j, k, bz)] = F[IX4(i, j, k, bz)] - eta_dx * JxS;
        } else {
          double JyS = (B(i, j, k + 1, bx) - B(i, j, k, bx)) -
                       0.25 * (B(i + 1, j, k, bz) + B(i + 1, j, k + 1, bz) - B(i - 1, j, k, bz) - B(i - 1, j, k + 1, bz));
          double JxS = -(B(i, j, k + 1, by) - B(i, j, k, by)) +
                       0.25 * (B(i, j + 1, k, bz) + B(i, j + 1, k + 1, bz) - B(i, j - 1, k, bz) - B(i, j - 1, k + 1, bz));
          F[IX4(i, j, k, en)] = F[IX4(i, j, k, en)] + 0.5 * eta_dx *
                                (JxS * (B(i, j, k, by) + B(i, j, k + 1, by)) - JyS * (B(i, j, k, bx) + B(i, j, k + 1, bx)));
          F[IX4(i, j, k, bx)] = F[IX4(i, j, k, bx)] - eta_dx * JyS;
          F[IX4(i, j, k, by)] = F[IX4(i, j, k, by)] + eta_dx * JxS;
        }
#undef B
      }
}

/* Localized resistivity of the 3D reconnection configuration at the face
 * centres: eta = eta0 + eta1 * exp(-(x^2 + y^2) / w^2), x/y of the face
 * centre as 3D_reconnection_gpu/flux_resistive.cuf:48-70 places them
 * (x(i) + 0.5*dx on x-faces, y(j) + 0.5*dx on y-faces, cell centre on
 * z-faces).  eta_out gets 3 arrays of ix*jx (x-, y-, z-faces); returns the
 * maximum over faces, for set_dt_resistive. */
double or_eta_faces(const double *x, const double *y, double dx, int ix, int jx, double eta0, double eta1, double w,
                    double *eta_out)
{
  double emax = 0.0;
  for (int d = 0; d < 3; d++)
    for (int j = 0; j < jx; j++)
      for (int i = 0; i < ix; i++) {
        double xx = x[i] + (d == 0 ? 0.5 * dx : 0.0);
        double yy = y[j] + (d == 1 ? 0.5 * dx : 0.0);
        double r2 = xx * xx + yy * yy;
        double e = eta0 + eta1 * exp(-r2 / (w * w));
        eta_out[(size_t)d * ix * jx + (size_t)i + (size_t)ix * j] = e;
        if (e > emax) emax = e;
      }
  return emax;
}

/* set_dt_resistive, 3D_reconnection_gpu/flux_resistive.cuf:2-9, with the
 * configured eta: dt = min(dt, 0.5*cfl*dx^2/eta_max). */
double or_set_dt_resistive(double dt, double dx, double cfl, double eta_max)
{
  double dr = 0.5 * cfl * (dx * dx) / eta_max;
  return dt < dr ? dt : dr;
}

/* ------------------------------------------------------------------------
 * The 2D CPU form, 2D_reconnection/flux_resistive.f90:46-111: eta/dx at the
 * face comes from the caller (eta_dx, an (ix,jx) array: eta(x(i)+dx/2, y(j))/dx
 * on x-faces, :74, and eta(x(i), y(j)+dx/2)/dx on y-faces, :97) and is folded
 * into the currents before they enter the flux.  Face ranges :49-55; dir 3
 * adds nothing (:56-57).
 * ---------------------------------------------------------------------- */
void or_flux_resistive2d(double *F, const double *U, const double *eta_dx, int ix, int jx, int dir)
{
  const int kx = 1, k = 1;
  int is, js;
  if (dir == 1) { is = 1; js = 2; }
  else if (dir == 2) { is = 2; js = 1; }
  else return;
  for (int j = js; j <= jx - 1; j++)
    for (int i = is; i <= ix - 1; i++) {
      const double e = eta_dx[(size_t)(i - 1) + (size_t)ix * (size_t)(j - 1)];
#define B(ii, jj, m) U[IX4(ii, jj, k, m)]
      if (dir == 1) {
        double JyS = -(e * (B(i + 1, j, bz) - B(i, j, bz)));
        double JzS = e * ((B(i + 1, j, by) - B(i, j, by)) -
                          0.25 * (B(i, j + 1, bx) + B(i + 1, j + 1, bx) - B(i, j - 1, bx) - B(i + 1, j - 1, bx)));
        F[IX4(i, j, k, en)] = F[IX4(i, j, k, en)] +
                              0.5 * (JyS * (B(i, j, bz) + B(i + 1, j, bz)) - JzS * (B(i, j, by) + B(i + 1, j, by)));
        F[IX4(i, j, k, by)] = F[IX4(i, j, k, by)] - JzS;
        F[IX4(i, j, k, bz)] = F[IX4(i, j, k, bz)] + JyS;
      } else {
        double JxS = e * (B(i, j + 1, bz) - B(i, j, bz));
        double JzS = e * (0.25 * (B(i + 1, j + 1, by) + B(i + 1, j, by) - B(i - 1, j + 1, by) - B(i - 1, j, by)) -
                          (B(i, j + 1, bx) - B(i, j, bx)));
        F[IX4(i, j, k, en)] = F[IX4(i, j, k, en)] +
                              0.5 * (JzS * (B(i, j, bx) + B(i, j + 1, bx)) - JxS * (B(i, j, bz) + B(i, j + 1, bz)));
        F[IX4(i, j, k, bx)] = F[IX4(i, j, k, bx)] + JzS;
        F[IX4(i, j, k, bz)] = F[IX4(i, j, k, bz)] - JxS;
      }
#undef B
    }
}

/* set_dt_resistive, 2D_reconnection/flux_resistive.f90:18-25:
 * dt = min(dt, 0.5*cfl*Rm1*dx^2), Rm1 the resistivity peak's inverse. */
double or_set_dt_resistive2d(double dt, double dx, double cfl, double rm1)
{
  double dr = 0.5 * cfl * rm1 * (dx * dx);
  return dt < dr ? dt : dr;
}

/* ------------------------------------------------------------------------
 * One full step in the order of 2D_basic/main.f90:61-187 (2D) and
 * 3D_basic/main.f90 (3D):  u2v -> set_dt -> glm_ss2 -> per dimension
 * (9 x limiter -> face patch -> flux_solver) -> rk stage 1 -> bc(U1) ->
 * u2v(U1) -> same again -> rk stage 2 -> bc(U) -> glm_ss2 -> t += dt.
 * Work arrays (V, U1, VL, VR, F, G, H) are allocated per call.
 * ---------------------------------------------------------------------- */
static int sweep(const or_problem *p, const double *Ust, const double *V, double *VL, double *VR, double *FX, int dir,
                 double ch, int ix, int jx, int kx)
{
  size_t n = (size_t)ix * jx * kx;
  int rc;
  for (int m = 0; m < 4; m++)
    if ((rc = or_limiter(V + m * n, VL + m * n, VR + m * n, ix, jx, kx, dir, p->lm_type))) return rc;
  for (int m = 4; m < 9; m++)
    if ((rc = or_limiter(Ust + m * n, VL + m * n, VR + m * n, ix, jx, kx, dir, p->lm_type))) return rc;
  int d = dir - 1;
  or_bc_face(VL, VR, ix, jx, kx, dir, p->bctype[2 * d], p->face_sgn + (2 * d) * 9, p->face_sgn + (2 * d + 1) * 9);
  if ((rc = or_flux_solver(FX, VL, VR, ch, ix, jx, kx, dir, p->flux_type, NULL))) return rc;
  if (p->eta && p->res_form == 1) /* 2D_reconnection/main.f90:102,120,150,168 */
    or_flux_resistive2d(FX, Ust, p->eta + (size_t)d * ix * jx, ix, jx, dir);
  else if (p->eta)
    or_flux_resistive(FX, Ust, p->eta + (size_t)d * ix * jx, p->dx, ix, jx, kx, dir); /* main.cuf:177-285 */
  return 0;
}

int or_step(const or_problem *p, double *U, double *dt_out, double *ch_out, long *bad)
{
  int ix = p->ix, jx = p->jx, kx = p->kx, nd = (kx > 1) ? 3 : 2;
  size_t n = (size_t)ix * jx * kx;
  double *V = malloc(sizeof(double) * 4 * n);
  double *U1 = malloc(sizeof(double) * 9 * n);
  double *VL = calloc(9 * n, sizeof(double));
  double *VR = calloc(9 * n, sizeof(double));
  double *F = calloc(9 * n, sizeof(double));
  double *G = calloc(9 * n, sizeof(double));
  double *H = (nd == 3) ? calloc(9 * n, sizeof(double)) : NULL;
  double ch, dt;
  int rc = 0;
  memcpy(U1, U, sizeof(double) * 9 * n);
  if ((rc = or_u2v(U, V, ix, jx, kx, bad))) goto out;
  if (p->dt_fixed > 0) {
    dt = p->dt_fixed;
    ch = p->ch_fixed;
  } else {
    if ((rc = or_set_dt(U, V, &ch, &dt, p->dx, p->cfl, ix, jx, kx, nd, bad))) goto out;
    if (p->eta && (p->res_form == 1 || p->res_form == 2)) /* 2D_reconnection/main.f90:81, main.cuf:147 */
      dt = or_set_dt_resistive2d(dt, p->dx, p->cfl, p->rm1);
    else if (p->eta)
      dt = or_set_dt_resistive(dt, p->dx, p->cfl, p->eta_max); /* main.cuf:147 */
  }
  or_glm_ss2(U, ch, dt, ix, jx, kx);
  for (int stage = 0; stage < 2; stage++) {
    const double *Ust = stage ? U1 : U;
    if ((rc = sweep(p, Ust, V, VL, VR, F, 1, ch, ix, jx, kx))) goto out;
    if ((rc = sweep(p, Ust, V, VL, VR, G, 2, ch, ix, jx, kx))) goto out;
    if (nd == 3 && (rc = sweep(p, Ust, V, VL, VR, H, 3, ch, ix, jx, kx))) goto out;
    if (stage == 0) {
      or_rk(p->time_type == 0 ? 0 : 2, U, U1, F, G, H, dt, p->dx, ix, jx, kx);
      or_bc_u(U1, ix, jx, kx, p->bctype, p->u_sgn);
      if ((rc = or_u2v(U1, V, ix, jx, kx, bad))) goto out;
    } else {
      or_rk(p->time_type == 0 ? 1 : 3, U, U1, F, G, H, dt, p->dx, ix, jx, kx);
      or_bc_u(U, ix, jx, kx, p->bctype, p->u_sgn);
    }
  }
  or_glm_ss2(U, ch, dt, ix, jx, kx);
  *dt_out = dt;
  *ch_out = ch;
out:
  free(V); free(U1); free(VL); free(VR); free(F); free(G); free(H);
  return rc;
}

int or_run(const or_problem *p, double *U, int nsteps, double *dt_trace, double *ch_trace, long *bad)
{
  for (int s = 0; s < nsteps; s++) {
    double dt, ch;
    int rc = or_step(p, U, &dt, &ch, bad);
    if (rc) return rc;
    if (dt_trace) dt_trace[s] = dt;
    if (ch_trace) ch_trace[s] = ch;
  }
  return 0;
}

/* ------------------------------------------------------------------------
 * The 1D shock-tube loop of 1D_basic/main.f90:56-149, on U(ix, 1, 9):
 * zero-gradient ghosts after every stage (:42-43, :112-113, :142-143), the
 * CFL step of set_dt (:76), nine limiter calls in x and one flux_solver call
 * with ch = 0 (:85-98, :117-128), and the RK stages written out in the main
 * program: U* = U + dtx (F(i-1)-F(i)) then U = 0.5 (U + U* + dtx (..))
 * (time_type 0, :100-104, :130-134), or the midpoint form with 0.5 dtx and
 * U + dtx (..) (time_type 1, :105-109, :135-139).  Stops when t >= tend or
 * after max_steps steps (the reference stops at loop_max - 1 steps).
 */
static void outflow1d(double *A, int ix)
{
  for (int m = 0; m < 9; m++) {
    A[(size_t)m * ix + (ix - 1)] = A[(size_t)m * ix + (ix - 2)];
    A[(size_t)m * ix] = A[(size_t)m * ix + 1];
  }
}

static int sweep1d(const double *Us, const double *V, double *VL, double *VR, double *F, int ix, int lm, int ft)
{
  int rc;
  for (int m = 0; m < 4; m++)
    if ((rc = or_limiter(V + (size_t)m * ix, VL + (size_t)m * ix, VR + (size_t)m * ix, ix, 1, 1, 1, lm))) return rc;
  for (int m = 4; m < 9; m++)
    if ((rc = or_limiter(Us + (size_t)m * ix, VL + (size_t)m * ix, VR + (size_t)m * ix, ix, 1, 1, 1, lm))) return rc;
  return or_flux_solver(F, VL, VR, 0.0, ix, 1, 1, 1, ft, NULL);
}

int or_run1d(double *U, int ix, double dx, double cfl, int lm_type, int flux_type, int time_type, double tend,
             int max_steps, double *t_out, int *nsteps_out, double *dt_trace, long *bad)
{
  const size_t n = (size_t)ix;
  double *U1 = calloc(9 * n, sizeof(double)), *V = calloc(4 * n, sizeof(double));
  double *VL = calloc(9 * n, sizeof(double)), *VR = calloc(9 * n, sizeof(double)), *F = calloc(9 * n, sizeof(double));
  double t = 0.0, dt = 0.0, vmax = 0.0;
  int steps = 0, rc = 0;
  if (!U1 || !V || !VL || !VR || !F) { rc = -1; goto done; }
  for (;;) {
    if ((rc = or_u2v(U, V, ix, 1, 1, bad))) break;
    if (t >= tend || steps >= max_steps) break;
    if ((rc = or_set_dt(U, V, &vmax, &dt, dx, cfl, ix, 1, 1, 2, bad))) break;
    const double dtx = dt / dx;
    if ((rc = sweep1d(U, V, VL, VR, F, ix, lm_type, flux_type))) break;
    for (int m = 0; m < 9; m++)
      for (int i = 1; i < ix - 1; i++) {
        const size_t q = (size_t)m * n + i;
        if (time_type == 0)
          U1[q] = U[q] + dtx * (F[q - 1] - F[q]);
        else
          U1[q] = U[q] + 0.5 * dtx * (F[q - 1] - F[q]);
      }
    outflow1d(U1, ix);
    if ((rc = or_u2v(U1, V, ix, 1, 1, bad))) break;
    if ((rc = sweep1d(U1, V, VL, VR, F, ix, lm_type, flux_type))) break;
    for (int m = 0; m < 9; m++)
      for (int i = 1; i < ix - 1; i++) {
        const size_t q = (size_t)m * n + i;
        if (time_type == 0)
          U[q] = 0.5 * (U[q] + U1[q] + dtx * (F[q - 1] - F[q]));
        else
          U[q] = U[q] + dtx * (F[q - 1] - F[q]);
      }
    outflow1d(U, ix);
    t = t + dt;
    if (dt_trace) dt_trace[steps] = dt;
    steps++;
  }
done:
  free(U1); free(V); free(VL); free(VR); free(F);
  *t_out = t;
  *nsteps_out = steps;
  return rc;
}
