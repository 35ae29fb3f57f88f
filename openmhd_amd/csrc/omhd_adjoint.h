// omhd_adjoint.h -- the per-cell and per-interface pieces of the step and
// their hand-written reverse (vector-Jacobian) forms, in plain C++ that
// compiles for the device (included by omhd_adjoint.hip) and for the host.
//
//   prim_fwd / prim_vjp   u2v at one cell          (2D_basic/u2v.f90:23-26)
//   lim_fwd  / lim_coef   slope limiter at one cell (2D_basic/limiter.f90)
//   hlld_fwd / hlld_vjp   HLLD flux at one face     (2D_basic/flux_solver.f90:352-585)
//
// The *_fwd forms are templates over the scalar type so that the same
// statements run in double (where they give the reference's bits: they follow
// its evaluation order) and in a dual-number type (forward-mode tangents, used
// to check every reverse form exactly).  Branch decisions are taken on values
// and recorded implicitly: the reverse forms recompute them from the same
// inputs, so the adjoint differentiates the branch the forward pass took, with
// one-sided derivatives at the kinks (|x| has slope +1 at x = 0, min(a, b)
// follows b on a tie, max(a, b) follows b on a tie, sqrt(max(q, 0)) has zero
// slope where its argument is 0, and SIGN(c, x) is constant).
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define OMHD_HD __host__ __device__ inline
#else
#define OMHD_HD inline
#endif

namespace omhd {
namespace adj {

constexpr double kGam = 5.0 / 3.0;  // param.h:3
constexpr int NV = 9;               // vx vy vz p | mx my mz en ; ro bx by bz psi

OMHD_HD double val(double a) { return a; }
OMHD_HD double tsqrt(double a) { return sqrt(a); }
OMHD_HD double tabs(double a) { return a < 0 ? -a : a; }
OMHD_HD double tmin(double a, double b) { return a < b ? a : b; }
OMHD_HD double tmax(double a, double b) { return a > b ? a : b; }
// Fortran SIGN(c, x) for a constant magnitude c: copysign semantics
template <class T>
OMHD_HD double csign(double c, const T &x) {
  return copysign(c, val(x));
}
OMHD_HD double slope_abs(double x) { return x < 0 ? -1.0 : 1.0; }  // d|x|/dx as tabs takes it

// ---------------------------------------------------------------------------
// u2v at one cell: u = (mx,my,mz,en,ro,bx,by,bz,psi) -> w = (vx,vy,vz,p,ro,B,psi)
// ---------------------------------------------------------------------------
template <class T>
OMHD_HD void prim_fwd(const T u[NV], T w[NV]) {
  const T rho = u[4];
  const T v1 = u[0] / rho, v2 = u[1] / rho, v3 = u[2] / rho;
  const T rv2 = v1 * u[0] + v2 * u[1] + v3 * u[2];
  const T B2 = u[5] * u[5] + u[6] * u[6] + u[7] * u[7];
  w[0] = v1;
  w[1] = v2;
  w[2] = v3;
  w[3] = (kGam - 1) * (u[3] - 0.5 * (rv2 + B2));
  for (int m = 4; m < NV; m++) w[m] = u[m];
}

// bu += (dw/du)^T bw, from the primitive values w of the cell (v = m/rho,
// p = (gam-1)(en - 0.5 (m.v + B.B)))
OMHD_HD void prim_vjp(const double w[NV], const double bw[NV], double bu[NV]) {
  const double rho = w[4], ir = 1.0 / rho;
  const double bp = (kGam - 1) * bw[3];
  const double brv2 = -0.5 * bp;
  const double v2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double brho = bw[4] - v2 * brv2;
  for (int m = 0; m < 3; m++) {
    bu[m] += bw[m] * ir + 2.0 * w[m] * brv2;
    brho -= bw[m] * w[m] * ir;
  }
  bu[3] += bp;
  bu[4] += brho;
  for (int m = 5; m < 8; m++) bu[m] += bw[m] - w[m] * bp;  // dB2 = 2B dB, times -0.5
  bu[8] += bw[8];
}

// ---------------------------------------------------------------------------
// Slope limiter at one cell from (w[c-1], w[c], w[c+1]): gR = w0 - gradR is the
// right state of the lower face, gL = w0 + gradL the left state of the upper
// face (limiter.f90: minmod :84-101, MC :141-158, van Leer :201-210, Koren
// :256-272; type 0 copies the cell value).
// ---------------------------------------------------------------------------
template <class T>
OMHD_HD void lim_fwd(int type, const T &wm, const T &w0, const T &wp, T &gR, T &gL) {
  if (type == 0) {
    gR = w0;
    gL = w0;
  } else if (type == 1) {
    const T gA = (w0 - wm), gB = (wp - w0);
    const T grad = (csign(0.25, gA) + csign(0.25, gB)) * tmin(tabs(gA), tabs(gB));
    gR = w0 - grad;
    gL = w0 + grad;
  } else if (type == 2) {
    const T gA = (w0 - wm), gB = (wp - w0);
    const T gC = 0.25 * (wp - wm);
    const T grad = (csign(0.5, gA) + csign(0.5, gB)) * tmin(tmin(tabs(gA), tabs(gB)), tabs(gC));
    gR = w0 - grad;
    gL = w0 + grad;
  } else if (type == 3) {
    const T gA = (w0 - wm), gB = (wp - w0);
    const T grad = (val(gA) * val(gB) <= 0) ? T(0.0) : gA * gB / (gA + gB);
    gR = w0 - grad;
    gL = w0 + grad;
  } else {
    const double f1 = 1.0 / 6.0;
    const T gA = (w0 - wm), gB = (wp - w0);
    const double s = csign(0.5, gA) + csign(0.5, gB);
    const T m = tmin(tabs(gA), tabs(gB));
    T gC = f1 * (2 * gA + gB);
    gR = w0 - s * tmin(m, tabs(gC));
    gC = f1 * (gA + 2 * gB);
    gL = w0 + s * tmin(m, tabs(gC));
  }
}

// Derivatives of gradR and gradL with respect to (wm, w0, wp):
// d gradR = cR[0] dwm + cR[1] dw0 + cR[2] dwp, likewise gradL.
OMHD_HD void lim_coef(int type, double wm, double w0, double wp, double cR[3], double cL[3]) {
  double kAR = 0, kBR = 0, kAL = 0, kBL = 0;  // d grad / d gA, d gB
  const double gA = (w0 - wm), gB = (wp - w0);
  if (type == 1 || type == 2) {
    const double s = (type == 1) ? csign(0.25, gA) + csign(0.25, gB) : csign(0.5, gA) + csign(0.5, gB);
    const double aA = tabs(gA), aB = tabs(gB);
    double kA = (aA < aB) ? s * slope_abs(gA) : 0.0;
    double kB = (aA < aB) ? 0.0 : s * slope_abs(gB);
    if (type == 2) {
      const double gC = 0.25 * (wp - wm);
      const double m1 = tmin(aA, aB);
      if (!(m1 < tabs(gC))) {  // the min picked |gC| = 0.25 |gB + gA|
        const double kC = s * slope_abs(gC) * 0.25;
        kA = kC;
        kB = kC;
      }
    }
    kAR = kAL = kA;
    kBR = kBL = kB;
  } else if (type == 3) {
    if (!(gA * gB <= 0)) {
      const double S = gA + gB, iS2 = 1.0 / (S * S);
      kAR = kAL = gB * gB * iS2;
      kBR = kBL = gA * gA * iS2;
    }
  } else if (type == 4) {
    const double f1 = 1.0 / 6.0;
    const double s = csign(0.5, gA) + csign(0.5, gB);
    const double aA = tabs(gA), aB = tabs(gB);
    const double mA = (aA < aB) ? slope_abs(gA) : 0.0, mB = (aA < aB) ? 0.0 : slope_abs(gB);
    const double m = tmin(aA, aB);
    const double c1 = f1 * (2 * gA + gB), c2 = f1 * (gA + 2 * gB);
    if (m < tabs(c1)) {
      kAR = s * mA;
      kBR = s * mB;
    } else {
      kAR = s * slope_abs(c1) * 2 * f1;
      kBR = s * slope_abs(c1) * f1;
    }
    if (m < tabs(c2)) {
      kAL = s * mA;
      kBL = s * mB;
    } else {
      kAL = s * slope_abs(c2) * f1;
      kBL = s * slope_abs(c2) * 2 * f1;
    }
  }
  // gA = w0 - wm, gB = wp - w0
  cR[0] = -kAR;
  cR[1] = kAR - kBR;
  cR[2] = kBR;
  cL[0] = -kAL;
  cL[1] = kAL - kBL;
  cL[2] = kBL;
}

// ---------------------------------------------------------------------------
// HLLD flux at one interface, flux_solver.f90:352-585 (Miyoshi & Kusano 2005
// with the HLLC-G fallback of Gurski 2004).  D = dir - 1; L, R in the slot
// order (vx,vy,vz,p,ro,bx,by,bz,psi).  branch: 1 FL, 2 FR, 3/4 HLLC-G left /
// right of the contact, 5 L*, 6 R*, 7 L**, 8 R**.
// ---------------------------------------------------------------------------
template <int D, class T>
OMHD_HD void hlld_fwd(const T L[NV], const T R[NV], double ch, T F[NV], int *branch = nullptr) {
  constexpr int vn = D, vt1 = (D + 1) % 3, vt2 = (D + 2) % 3;
  constexpr int mn = vn, mt1 = vt1, mt2 = vt2;
  constexpr int bn = 5 + vn, bt1 = 5 + vt1, bt2 = 5 + vt2;
  constexpr int RO = 4, PR = 3, EN = 3, PS = 8;
  const double r1 = 1.0 / (kGam - 1.0), r2 = kGam / (kGam - 1.0), hllg = 1.001;
  T UL[4], UR[4], FL[NV], FR[NV];
  for (int m = 0; m < NV; m++) {
    FL[m] = T(0.0);
    FR[m] = T(0.0);
    F[m] = T(0.0);
  }
  int br = 0;
  T v2 = L[0] * L[0] + L[1] * L[1] + L[2] * L[2];
  T B2 = L[5] * L[5] + L[6] * L[6] + L[7] * L[7];
  const T vBL = L[0] * L[5] + L[1] * L[6] + L[2] * L[7];
  UL[0] = L[RO] * L[0];
  UL[1] = L[RO] * L[1];
  UL[2] = L[RO] * L[2];
  UL[EN] = 0.5 * (L[RO] * v2 + B2) + r1 * L[PR];
  FL[RO] = L[RO] * L[vn];
  FL[mn] = L[RO] * L[vn] * L[vn] - L[bn] * L[bn] + L[PR] + 0.5 * B2;
  FL[mt1] = L[RO] * L[vn] * L[vt1] - L[bn] * L[bt1];
  FL[mt2] = L[RO] * L[vn] * L[vt2] - L[bn] * L[bt2];
  FL[EN] = (0.5 * L[RO] * v2 + r2 * L[PR] + B2) * L[vn] - vBL * L[bn];
  FL[bt1] = L[vn] * L[bt1] - L[bn] * L[vt1];
  FL[bt2] = L[vn] * L[bt2] - L[bn] * L[vt2];
  T f1 = kGam * L[PR];
  T f2 = 4 * (L[bn] * L[bn]);
  T aL = ((f1 + B2) + tsqrt(tmax((f1 + B2) * (f1 + B2) - f1 * f2, T(0.0)))) / (2 * L[RO]);
  const T ptL = L[PR] + 0.5 * B2;

  v2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
  B2 = R[5] * R[5] + R[6] * R[6] + R[7] * R[7];
  const T vBR = R[0] * R[5] + R[1] * R[6] + R[2] * R[7];
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
  f1 = kGam * R[PR];
  f2 = 4 * (R[bn] * R[bn]);
  T aR = ((f1 + B2) + tsqrt(tmax((f1 + B2) * (f1 + B2) - f1 * f2, T(0.0)))) / (2 * R[RO]);
  const T ptR = R[PR] + 0.5 * B2;

  f1 = tsqrt(tmax(aL, aR));
  aL = tmin(L[vn], R[vn]) - f1;
  aR = tmax(L[vn], R[vn]) + f1;

  if (val(aL) >= 0) {
    for (int m = 0; m < NV; m++) F[m] = FL[m];
    br = 1;
  } else if (val(aR) <= 0) {
    for (int m = 0; m < NV; m++) F[m] = FR[m];
    br = 2;
  } else {
    T Uh[NV];
    f1 = 1.0 / (aR - aL);
    for (int m = 0; m < 3; m++) Uh[m] = f1 * (aR * UR[m] - aL * UL[m] - FR[m] + FL[m]);
    for (int m = RO; m <= 7; m++) Uh[m] = f1 * (aR * R[m] - aL * L[m] - FR[m] + FL[m]);
    const T aM = Uh[mn] / Uh[RO];
    const T pt = ptL + L[RO] * (aL - L[vn]) * (aM - L[vn]);
    const T roL = L[RO] * (aL - L[vn]) / (aL - aM);
    const T roR = R[RO] * (aR - R[vn]) / (aR - aM);
    const T aVL = tabs(Uh[bn]) / tsqrt(roL);
    const T aVR = tabs(Uh[bn]) / tsqrt(roR);
    if ((val(aL) >= val(aM - hllg * aVL)) || (val(aM + hllg * aVR) >= val(aR))) {
      f2 = aL * aR;
      F[bt1] = f1 * (aR * FL[bt1] - aL * FR[bt1] + f2 * (R[bt1] - L[bt1]));
      F[bt2] = f1 * (aR * FL[bt2] - aL * FR[bt2] + f2 * (R[bt2] - L[bt2]));
      const T at1 = Uh[mt1] / Uh[RO];
      const T at2 = Uh[mt2] / Uh[RO];
      const T enL = ((aL - L[vn]) * UL[EN] - ptL * L[vn] + pt * aM +
                     Uh[bn] * (vBL - aM * Uh[bn] - at1 * Uh[bt1] - at2 * Uh[bt2])) / (aL - aM);
      const T enR = ((aR - R[vn]) * UR[EN] - ptR * R[vn] + pt * aM +
                     Uh[bn] * (vBR - aM * Uh[bn] - at1 * Uh[bt1] - at2 * Uh[bt2])) / (aR - aM);
      const double s1 = 0.5 + csign(0.5, aM);
      const double s2 = 1.0 - s1;
      F[mn] = s1 * (aL * (roL * aM - UL[mn]) + FL[mn]) + s2 * (aR * (roR * aM - UR[mn]) + FR[mn]);
      F[mt1] = s1 * (aL * (roL * at1 - UL[mt1]) + FL[mt1]) + s2 * (aR * (roR * at1 - UR[mt1]) + FR[mt1]);
      F[mt2] = s1 * (aL * (roL * at2 - UL[mt2]) + FL[mt2]) + s2 * (aR * (roR * at2 - UR[mt2]) + FR[mt2]);
      F[EN] = s1 * (aL * (enL - UL[EN]) + FL[EN]) + s2 * (aR * (enR - UR[EN]) + FR[EN]);
      F[RO] = s1 * (aL * (roL - L[RO]) + FL[RO]) + s2 * (aR * (roR - R[RO]) + FR[RO]);
      br = (val(aM) >= 0) ? 3 : 4;
    } else {
      T UL1[NV], UR1[NV];
      f1 = 1.0 / (L[RO] * (aL - L[vn]) * (aL - aM) - Uh[bn] * Uh[bn]);
      UL1[bt1] = L[bt1] * f1 * (L[RO] * ((aL - L[vn]) * (aL - L[vn])) - Uh[bn] * Uh[bn]);
      UL1[bt2] = L[bt2] * f1 * (L[RO] * ((aL - L[vn]) * (aL - L[vn])) - Uh[bn] * Uh[bn]);
      const T vt1L = L[vt1] - f1 * Uh[bn] * L[bt1] * (aM - L[vn]);
      const T vt2L = L[vt2] - f1 * Uh[bn] * L[bt2] * (aM - L[vn]);
      UL1[RO] = roL;
      UL1[EN] = ((aL - L[vn]) * UL[EN] - ptL * L[vn] + pt * aM +
                 Uh[bn] * (vBL - aM * Uh[bn] - vt1L * UL1[bt1] - vt2L * UL1[bt2])) / (aL - aM);
      UL1[mn] = roL * aM;
      UL1[mt1] = roL * vt1L;
      UL1[mt2] = roL * vt2L;
      f1 = 1.0 / (R[RO] * (aR - R[vn]) * (aR - aM) - Uh[bn] * Uh[bn]);
      UR1[bt1] = R[bt1] * f1 * (R[RO] * ((aR - R[vn]) * (aR - R[vn])) - Uh[bn] * Uh[bn]);
      UR1[bt2] = R[bt2] * f1 * (R[RO] * ((aR - R[vn]) * (aR - R[vn])) - Uh[bn] * Uh[bn]);
      const T vt1R = R[vt1] - f1 * Uh[bn] * R[bt1] * (aM - R[vn]);
      const T vt2R = R[vt2] - f1 * Uh[bn] * R[bt2] * (aM - R[vn]);
      UR1[RO] = roR;
      UR1[EN] = ((aR - R[vn]) * UR[EN] - ptR * R[vn] + pt * aM +
                 Uh[bn] * (vBR - aM * Uh[bn] - vt1R * UR1[bt1] - vt2R * UR1[bt2])) / (aR - aM);
      UR1[mn] = roR * aM;
      UR1[mt1] = roR * vt1R;
      UR1[mt2] = roR * vt2R;
      const T aL1 = aM - aVL;
      const T aR1 = aM + aVR;
      if (val(aL1) >= 0) {
        for (int m = 0; m < 4; m++) F[m] = aL * (UL1[m] - UL[m]) + FL[m];
        F[RO] = aL * (UL1[RO] - L[RO]) + FL[RO];
        F[bt1] = aL * (UL1[bt1] - L[bt1]) + FL[bt1];
        F[bt2] = aL * (UL1[bt2] - L[bt2]) + FL[bt2];
        br = 5;
      } else if (val(aR1) <= 0) {
        for (int m = 0; m < 4; m++) F[m] = aR * (UR1[m] - UR[m]) + FR[m];
        F[RO] = aR * (UR1[RO] - R[RO]) + FR[RO];
        F[bt1] = aR * (UR1[bt1] - R[bt1]) + FR[bt1];
        F[bt2] = aR * (UR1[bt2] - R[bt2]) + FR[bt2];
        br = 6;
      } else {
        T U2[NV];
        const T roLs = tsqrt(roL);
        const T roRs = tsqrt(roR);
        f1 = 1.0 / (roLs + roRs);
        const double sB = csign(1.0, Uh[bn]);
        const T at1 = f1 * (roLs * vt1L + roRs * vt1R + (UR1[bt1] - UL1[bt1]) * sB);
        const T at2 = f1 * (roLs * vt2L + roRs * vt2R + (UR1[bt2] - UL1[bt2]) * sB);
        U2[bt1] = f1 * (roLs * UR1[bt1] + roRs * UL1[bt1] + roLs * roRs * (vt1R - vt1L) * sB);
        U2[bt2] = f1 * (roLs * UR1[bt2] + roRs * UL1[bt2] + roLs * roRs * (vt2R - vt2L) * sB);
        if (val(aM) >= 0) {
          U2[RO] = roL;
          U2[EN] = UL1[EN] - roLs * (vt1L * UL1[bt1] + vt2L * UL1[bt2] - at1 * U2[bt1] - at2 * U2[bt2]) * sB;
          U2[mn] = roL * aM;
          U2[mt1] = roL * at1;
          U2[mt2] = roL * at2;
          for (int m = 0; m < 4; m++) F[m] = aL1 * U2[m] - (aL1 - aL) * UL1[m] - aL * UL[m] + FL[m];
          F[RO] = aL1 * U2[RO] - (aL1 - aL) * UL1[RO] - aL * L[RO] + FL[RO];
          F[bt1] = aL1 * U2[bt1] - (aL1 - aL) * UL1[bt1] - aL * L[bt1] + FL[bt1];
          F[bt2] = aL1 * U2[bt2] - (aL1 - aL) * UL1[bt2] - aL * L[bt2] + FL[bt2];
          br = 7;
        } else {
          U2[RO] = roR;
          U2[EN] = UR1[EN] + roRs * (vt1R * UR1[bt1] + vt2R * UR1[bt2] - at1 * U2[bt1] - at2 * U2[bt2]) * sB;
          U2[mn] = roR * aM;
          U2[mt1] = roR * at1;
          U2[mt2] = roR * at2;
          for (int m = 0; m < 4; m++) F[m] = aR1 * U2[m] - (aR1 - aR) * UR1[m] - aR * UR[m] + FR[m];
          F[RO] = aR1 * U2[RO] - (aR1 - aR) * UR1[RO] - aR * R[RO] + FR[RO];
          F[bt1] = aR1 * U2[bt1] - (aR1 - aR) * UR1[bt1] - aR * R[bt1] + FR[bt1];
          F[bt2] = aR1 * U2[bt2] - (aR1 - aR) * UR1[bt2] - aR * R[bt2] + FR[bt2];
          br = 8;
        }
      }
    }
  }
  F[bn] = 0.5 * ((L[PS] + R[PS]) - ch * (R[bn] - L[bn]));
  F[PS] = 0.5 * (ch * ch * (L[bn] + R[bn]) - ch * (R[PS] - L[PS]));
  if (branch) *branch = br;
}

// ---------------------------------------------------------------------------
// Reverse of one side's state block of hlld_fwd: given the adjoints of the
// conserved momenta/energy U[0..3], of the physical flux FS, of the fast-mode
// speed^2 a and of the total pressure pt, accumulate into bs (the side state).
// ---------------------------------------------------------------------------
template <int D>
OMHD_HD void hlld_side_vjp(const double s[NV], const double bU[4], const double bFS[NV], double ba, double bpt,
                           double bvB, double bs[NV]) {
  constexpr int vn = D, vt1 = (D + 1) % 3, vt2 = (D + 2) % 3;
  constexpr int mn = vn, mt1 = vt1, mt2 = vt2;
  constexpr int bn = 5 + vn, bt1 = 5 + vt1, bt2 = 5 + vt2;
  constexpr int RO = 4, PR = 3, EN = 3;
  const double r1 = 1.0 / (kGam - 1.0), r2 = kGam / (kGam - 1.0);
  const double ro = s[RO], p = s[PR];
  const double v2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
  const double B2 = s[5] * s[5] + s[6] * s[6] + s[7] * s[7];
  const double vB = s[0] * s[5] + s[1] * s[6] + s[2] * s[7];
  double bv2 = 0, bB2 = 0, bro = 0, bp = 0;
  double bv[3] = {0, 0, 0}, bB[3] = {0, 0, 0};  // by velocity / field component index
  // pt = p + 0.5 B2
  bp += bpt;
  bB2 += 0.5 * bpt;
  // a = (A + sqrt(max(q, 0))) / (2 ro), A = f1 + B2, q = A*A - f1*f2
  {
    const double f1 = kGam * p, f2 = 4 * (s[bn] * s[bn]);
    const double A = f1 + B2;
    const double q = A * A - f1 * f2;
    const double sq = sqrt(q > 0 ? q : 0.0);
    const double ir = 1.0 / ro;  // reverse-only arithmetic: no correctly rounded quotients needed
    const double a = (A + sq) * (0.5 * ir);
    const double bN = 0.5 * ba * ir;
    bro -= ba * a * ir;
    double bA = bN;
    const double bq = (q > 0 && sq > 0) ? bN / (2 * sq) : 0.0;
    bA += 2 * A * bq;
    double bf1 = -f2 * bq, bf2 = -f1 * bq;
    bf1 += bA;
    bB2 += bA;
    bp += kGam * bf1;
    bB[vn] += 8 * s[bn] * bf2;
  }
  // FS[bt2] = vn*bt2 - bn*vt2 ; FS[bt1]
  bv[vn] += s[bt2] * bFS[bt2] + s[bt1] * bFS[bt1];
  bB[vt2] += s[vn] * bFS[bt2];
  bB[vn] -= s[vt2] * bFS[bt2] + s[vt1] * bFS[bt1];
  bv[vt2] -= s[bn] * bFS[bt2];
  bB[vt1] += s[vn] * bFS[bt1];
  bv[vt1] -= s[bn] * bFS[bt1];
  // FS[EN] = (0.5 ro v2 + r2 p + B2) vn - vB bn
  {
    const double E = 0.5 * ro * v2 + r2 * p + B2;
    const double bE = s[vn] * bFS[EN];
    bv[vn] += E * bFS[EN];
    bvB -= s[bn] * bFS[EN];
    bB[vn] -= vB * bFS[EN];
    bro += 0.5 * v2 * bE;
    bv2 += 0.5 * ro * bE;
    bp += r2 * bE;
    bB2 += bE;
  }
  // FS[mt] = ro vn vt - bn bt
  bro += s[vn] * s[vt2] * bFS[mt2] + s[vn] * s[vt1] * bFS[mt1];
  bv[vn] += ro * s[vt2] * bFS[mt2] + ro * s[vt1] * bFS[mt1];
  bv[vt2] += ro * s[vn] * bFS[mt2];
  bv[vt1] += ro * s[vn] * bFS[mt1];
  bB[vn] -= s[bt2] * bFS[mt2] + s[bt1] * bFS[mt1];
  bB[vt2] -= s[bn] * bFS[mt2];
  bB[vt1] -= s[bn] * bFS[mt1];
  // FS[mn] = ro vn vn - bn bn + p + 0.5 B2
  bro += s[vn] * s[vn] * bFS[mn];
  bv[vn] += 2 * ro * s[vn] * bFS[mn];
  bB[vn] -= 2 * s[bn] * bFS[mn];
  bp += bFS[mn];
  bB2 += 0.5 * bFS[mn];
  // FS[RO] = ro vn
  bro += s[vn] * bFS[RO];
  bv[vn] += ro * bFS[RO];
  // U[EN] = 0.5 (ro v2 + B2) + r1 p ; U[m] = ro v_m
  bro += 0.5 * v2 * bU[EN];
  bv2 += 0.5 * ro * bU[EN];
  bB2 += 0.5 * bU[EN];
  bp += r1 * bU[EN];
  for (int m = 0; m < 3; m++) {
    bro += s[m] * bU[m];
    bv[m] += ro * bU[m];
  }
  // vB, B2, v2
  for (int m = 0; m < 3; m++) {
    bv[m] += s[5 + m] * bvB + 2 * s[m] * bv2;
    bB[m] += s[m] * bvB + 2 * s[5 + m] * bB2;
  }
  for (int m = 0; m < 3; m++) {
    bs[m] += bv[m];
    bs[5 + m] += bB[m];
  }
  bs[PR] += bp;
  bs[RO] += bro;
}

// bL += (dF/dL)^T bF, bR += (dF/dR)^T bF at the branch hlld_fwd takes.
template <int D>
OMHD_HD void hlld_vjp(const double L[NV], const double R[NV], double ch, const double bF[NV], double bL[NV],
                      double bR[NV]) {
  constexpr int vn = D, vt1 = (D + 1) % 3, vt2 = (D + 2) % 3;
  constexpr int mn = vn, mt1 = vt1, mt2 = vt2;
  constexpr int bn = 5 + vn, bt1 = 5 + vt1, bt2 = 5 + vt2;
  constexpr int RO = 4, PR = 3, EN = 3, PS = 8;
  const double r1 = 1.0 / (kGam - 1.0), r2 = kGam / (kGam - 1.0), hllg = 1.001;
  // ---- forward recomputation (the same statements as hlld_fwd) ----
  double UL[4], UR[4], FL[NV], FR[NV];
  for (int m = 0; m < NV; m++) FL[m] = FR[m] = 0.0;
  double v2 = L[0] * L[0] + L[1] * L[1] + L[2] * L[2];
  double B2 = L[5] * L[5] + L[6] * L[6] + L[7] * L[7];
  const double vBL = L[0] * L[5] + L[1] * L[6] + L[2] * L[7];
  UL[0] = L[RO] * L[0];
  UL[1] = L[RO] * L[1];
  UL[2] = L[RO] * L[2];
  UL[EN] = 0.5 * (L[RO] * v2 + B2) + r1 * L[PR];
  FL[RO] = L[RO] * L[vn];
  FL[mn] = L[RO] * L[vn] * L[vn] - L[bn] * L[bn] + L[PR] + 0.5 * B2;
  FL[mt1] = L[RO] * L[vn] * L[vt1] - L[bn] * L[bt1];
  FL[mt2] = L[RO] * L[vn] * L[vt2] - L[bn] * L[bt2];
  FL[EN] = (0.5 * L[RO] * v2 + r2 * L[PR] + B2) * L[vn] - vBL * L[bn];
  FL[bt1] = L[vn] * L[bt1] - L[bn] * L[vt1];
  FL[bt2] = L[vn] * L[bt2] - L[bn] * L[vt2];
  double f1 = kGam * L[PR];
  double f2 = 4 * (L[bn] * L[bn]);
  const double afL = ((f1 + B2) + sqrt(tmax((f1 + B2) * (f1 + B2) - f1 * f2, 0.0))) / (2 * L[RO]);
  const double ptL = L[PR] + 0.5 * B2;
  v2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
  B2 = R[5] * R[5] + R[6] * R[6] + R[7] * R[7];
  const double vBR = R[0] * R[5] + R[1] * R[6] + R[2] * R[7];
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
  f1 = kGam * R[PR];
  f2 = 4 * (R[bn] * R[bn]);
  const double afR = ((f1 + B2) + sqrt(tmax((f1 + B2) * (f1 + B2) - f1 * f2, 0.0))) / (2 * R[RO]);
  const double ptR = R[PR] + 0.5 * B2;
  const double sa = sqrt(tmax(afL, afR));
  const double aL = tmin(L[vn], R[vn]) - sa;
  const double aR = tmax(L[vn], R[vn]) + sa;

  // ---- adjoint accumulators ----
  double bUL[4] = {0, 0, 0, 0}, bUR[4] = {0, 0, 0, 0};
  double bFL[NV], bFR[NV];
  for (int m = 0; m < NV; m++) bFL[m] = bFR[m] = 0.0;
  double baL = 0, baR = 0;      // fan speeds SL, SR
  double bptL = 0, bptR = 0, bvBL = 0, bvBR = 0;
  // GLM pair (flux_solver.f90:580-581), every branch
  bL[PS] += 0.5 * bF[bn] + 0.5 * ch * bF[PS];
  bR[PS] += 0.5 * bF[bn] - 0.5 * ch * bF[PS];
  bL[bn] += 0.5 * ch * bF[bn] + 0.5 * ch * ch * bF[PS];
  bR[bn] += -0.5 * ch * bF[bn] + 0.5 * ch * ch * bF[PS];

  if (aL >= 0) {
    for (int m = 0; m < NV; m++)
      if (m != bn && m != PS) bFL[m] += bF[m];
  } else if (aR <= 0) {
    for (int m = 0; m < NV; m++)
      if (m != bn && m != PS) bFR[m] += bF[m];
  } else {
    double Uh[NV];
    const double iw = 1.0 / (aR - aL);
    for (int m = 0; m < 3; m++) Uh[m] = iw * (aR * UR[m] - aL * UL[m] - FR[m] + FL[m]);
    for (int m = RO; m <= 7; m++) Uh[m] = iw * (aR * R[m] - aL * L[m] - FR[m] + FL[m]);
    Uh[3] = Uh[8] = 0.0;
    const double aM = Uh[mn] / Uh[RO];
    const double pt = ptL + L[RO] * (aL - L[vn]) * (aM - L[vn]);
    const double roL = L[RO] * (aL - L[vn]) / (aL - aM);
    const double roR = R[RO] * (aR - R[vn]) / (aR - aM);
    const double hb = tabs(Uh[bn]);
    const double sroL = sqrt(roL), sroR = sqrt(roR);
    const double aVL = hb / sroL;
    const double aVR = hb / sroR;
    double bUh[NV];
    for (int m = 0; m < NV; m++) bUh[m] = 0.0;
    double biw = 0, baM = 0, bpt = 0, broL = 0, broR = 0, baVL = 0, baVR = 0;
    if ((aL >= (aM - hllg * aVL)) || ((aM + hllg * aVR) >= aR)) {
      // ---- HLLC-G (flux_solver.f90:447-476) ----
      const double w2 = aL * aR;
      {
        const double Nb1 = aR * FL[bt1] - aL * FR[bt1] + w2 * (R[bt1] - L[bt1]);
        const double Nb2 = aR * FL[bt2] - aL * FR[bt2] + w2 * (R[bt2] - L[bt2]);
        biw += Nb1 * bF[bt1] + Nb2 * bF[bt2];
        const double b1 = iw * bF[bt1], b2 = iw * bF[bt2];
        baR += FL[bt1] * b1 + FL[bt2] * b2;
        bFL[bt1] += aR * b1;
        bFL[bt2] += aR * b2;
        baL -= FR[bt1] * b1 + FR[bt2] * b2;
        bFR[bt1] -= aL * b1;
        bFR[bt2] -= aL * b2;
        const double bw2 = (R[bt1] - L[bt1]) * b1 + (R[bt2] - L[bt2]) * b2;
        bR[bt1] += w2 * b1;
        bL[bt1] -= w2 * b1;
        bR[bt2] += w2 * b2;
        bL[bt2] -= w2 * b2;
        baL += aR * bw2;
        baR += aL * bw2;
      }
      const double at1 = Uh[mt1] / Uh[RO];
      const double at2 = Uh[mt2] / Uh[RO];
      double bat1 = 0, bat2 = 0;
      const bool useL = csign(0.5, aM) > 0;  // s1 = 0.5 + SIGN(0.5, aM) is 1 or 0
      // the reverse of the side the flux takes, written once and inlined once
      // per side, so every array it touches stays an ordinary register array
      auto side = [&](const double *S, const double *US, double *bUS, double *bFS, double *bS, const double aS,
                      const double roS, const double ptS, const double vBS, double &baS, double &broS,
                      double &bptS, double &bvBS) __attribute__((always_inline)) {
        const double P = vBS - aM * Uh[bn] - at1 * Uh[bt1] - at2 * Uh[bt2];
        const double Q = (aS - S[vn]) * US[EN] - ptS * S[vn] + pt * aM + Uh[bn] * P;
        const double idS = 1.0 / (aS - aM);
        const double enS = Q * idS;
        // F[mn] = aS (roS aM - US[mn]) + FS[mn]
        baS += (roS * aM - US[mn]) * bF[mn];
        broS += aS * aM * bF[mn];
        baM += aS * roS * bF[mn];
        bUS[mn] -= aS * bF[mn];
        bFS[mn] += bF[mn];
        // F[mt1], F[mt2]
        baS += (roS * at1 - US[mt1]) * bF[mt1] + (roS * at2 - US[mt2]) * bF[mt2];
        broS += aS * at1 * bF[mt1] + aS * at2 * bF[mt2];
        bat1 += aS * roS * bF[mt1];
        bat2 += aS * roS * bF[mt2];
        bUS[mt1] -= aS * bF[mt1];
        bUS[mt2] -= aS * bF[mt2];
        bFS[mt1] += bF[mt1];
        bFS[mt2] += bF[mt2];
        // F[EN] = aS (enS - US[EN]) + FS[EN]
        baS += (enS - US[EN]) * bF[EN];
        const double benS = aS * bF[EN];
        bUS[EN] -= aS * bF[EN];
        bFS[EN] += bF[EN];
        // F[RO] = aS (roS - S[RO]) + FS[RO]
        baS += (roS - S[RO]) * bF[RO];
        broS += aS * bF[RO];
        bS[RO] -= aS * bF[RO];
        bFS[RO] += bF[RO];
        // enS = Q / (aS - aM)
        {
          const double bQ = benS * idS;
          const double bd = -bQ * enS;
          baS += bd;
          baM -= bd;
          baS += US[EN] * bQ;
          bS[vn] -= US[EN] * bQ + ptS * bQ;
          bUS[EN] += (aS - S[vn]) * bQ;
          bptS -= S[vn] * bQ;
          bpt += aM * bQ;
          baM += pt * bQ;
          bUh[bn] += P * bQ;
          const double bP = Uh[bn] * bQ;
          bvBS += bP;
          baM -= Uh[bn] * bP;
          bUh[bn] -= aM * bP;
          bat1 -= Uh[bt1] * bP;
          bUh[bt1] -= at1 * bP;
          bat2 -= Uh[bt2] * bP;
          bUh[bt2] -= at2 * bP;
        }
      };
      if (useL) side(L, UL, bUL, bFL, bL, aL, roL, ptL, vBL, baL, broL, bptL, bvBL);
      else side(R, UR, bUR, bFR, bR, aR, roR, ptR, vBR, baR, broR, bptR, bvBR);
      // at = Uh[mt] / Uh[RO]
      const double iro = 1.0 / Uh[RO];
      bUh[mt1] += bat1 * iro;
      bUh[mt2] += bat2 * iro;
      bUh[RO] -= (at1 * bat1 + at2 * bat2) * iro;
    } else {
      // ---- HLLD (flux_solver.f90:482-572) ----
      double UL1[NV], UR1[NV];
      const double dL = L[RO] * (aL - L[vn]) * (aL - aM) - Uh[bn] * Uh[bn];
      const double f1L = 1.0 / dL;
      const double eL = L[RO] * ((aL - L[vn]) * (aL - L[vn])) - Uh[bn] * Uh[bn];
      UL1[bt1] = L[bt1] * f1L * eL;
      UL1[bt2] = L[bt2] * f1L * eL;
      const double vt1L = L[vt1] - f1L * Uh[bn] * L[bt1] * (aM - L[vn]);
      const double vt2L = L[vt2] - f1L * Uh[bn] * L[bt2] * (aM - L[vn]);
      UL1[RO] = roL;
      const double PL = vBL - aM * Uh[bn] - vt1L * UL1[bt1] - vt2L * UL1[bt2];
      const double QL = (aL - L[vn]) * UL[EN] - ptL * L[vn] + pt * aM + Uh[bn] * PL;
      UL1[EN] = QL / (aL - aM);
      UL1[mn] = roL * aM;
      UL1[mt1] = roL * vt1L;
      UL1[mt2] = roL * vt2L;
      const double dR = R[RO] * (aR - R[vn]) * (aR - aM) - Uh[bn] * Uh[bn];
      const double f1R = 1.0 / dR;
      const double eR = R[RO] * ((aR - R[vn]) * (aR - R[vn])) - Uh[bn] * Uh[bn];
      UR1[bt1] = R[bt1] * f1R * eR;
      UR1[bt2] = R[bt2] * f1R * eR;
      const double vt1R = R[vt1] - f1R * Uh[bn] * R[bt1] * (aM - R[vn]);
      const double vt2R = R[vt2] - f1R * Uh[bn] * R[bt2] * (aM - R[vn]);
      UR1[RO] = roR;
      const double PR_ = vBR - aM * Uh[bn] - vt1R * UR1[bt1] - vt2R * UR1[bt2];
      const double QR = (aR - R[vn]) * UR[EN] - ptR * R[vn] + pt * aM + Uh[bn] * PR_;
      UR1[EN] = QR / (aR - aM);
      UR1[mn] = roR * aM;
      UR1[mt1] = roR * vt1R;
      UR1[mt2] = roR * vt2R;
      const double aL1 = aM - aVL;
      const double aR1 = aM + aVR;
      double bUL1[NV], bUR1[NV];
      for (int m = 0; m < NV; m++) bUL1[m] = bUR1[m] = 0.0;
      double bvt1L = 0, bvt2L = 0, bvt1R = 0, bvt2R = 0, baL1 = 0, baR1 = 0;
      // indices of the flux slots a star-state branch writes: momenta + EN
      // (cons slots 0..3), RO, bt1, bt2
      if (aL1 >= 0) {
        for (int m = 0; m < 4; m++) {
          baL += (UL1[m] - UL[m]) * bF[m];
          bUL1[m] += aL * bF[m];
          bUL[m] -= aL * bF[m];
          bFL[m] += bF[m];
        }
        for (int q = 0; q < 3; q++) {
          const int m = q == 0 ? RO : (q == 1 ? bt1 : bt2);
          baL += (UL1[m] - L[m]) * bF[m];
          bUL1[m] += aL * bF[m];
          bL[m] -= aL * bF[m];
          bFL[m] += bF[m];
        }
      } else if (aR1 <= 0) {
        for (int m = 0; m < 4; m++) {
          baR += (UR1[m] - UR[m]) * bF[m];
          bUR1[m] += aR * bF[m];
          bUR[m] -= aR * bF[m];
          bFR[m] += bF[m];
        }
        for (int q = 0; q < 3; q++) {
          const int m = q == 0 ? RO : (q == 1 ? bt1 : bt2);
          baR += (UR1[m] - R[m]) * bF[m];
          bUR1[m] += aR * bF[m];
          bR[m] -= aR * bF[m];
          bFR[m] += bF[m];
        }
      } else {
        // central states (flux_solver.f90:533-572)
        double U2[NV], bU2[NV];
        for (int m = 0; m < NV; m++) bU2[m] = 0.0;
        const double roLs = sroL, roRs = sroR;
        const double fc = 1.0 / (roLs + roRs);
        const double sB = csign(1.0, Uh[bn]);
        const double Na1 = roLs * vt1L + roRs * vt1R + (UR1[bt1] - UL1[bt1]) * sB;
        const double Na2 = roLs * vt2L + roRs * vt2R + (UR1[bt2] - UL1[bt2]) * sB;
        const double at1 = fc * Na1, at2 = fc * Na2;
        const double Nb1 = roLs * UR1[bt1] + roRs * UL1[bt1] + roLs * roRs * (vt1R - vt1L) * sB;
        const double Nb2 = roLs * UR1[bt2] + roRs * UL1[bt2] + roLs * roRs * (vt2R - vt2L) * sB;
        U2[bt1] = fc * Nb1;
        U2[bt2] = fc * Nb2;
        double broLs = 0, broRs = 0, bat1 = 0, bat2 = 0;
        const bool left = aM >= 0;
        // the side whose double-star state the flux takes (L** or R**), inlined
        // once per side so every array it touches stays a register array
        auto central_side = [&](const double *S, const double *US, const double *US1, double *bUS, double *bUS1,
                                double *bFS, double *bS, const double aS, const double aS1, const double roS,
                                const double roSs, const double vt1S, const double vt2S, const double sgn,
                                double &baS, double &baS1, double &broS, double &broSs, double &bvt1S,
                                double &bvt2S) __attribute__((always_inline)) {
          U2[RO] = roS;
          const double K = vt1S * US1[bt1] + vt2S * US1[bt2] - at1 * U2[bt1] - at2 * U2[bt2];
          U2[EN] = US1[EN] + sgn * roSs * K * sB;
          U2[mn] = roS * aM;
          U2[mt1] = roS * at1;
          U2[mt2] = roS * at2;
          // F[m] = aS1 U2[m] - (aS1 - aS) US1[m] - aS XS[m] + FS[m]
          for (int m = 0; m < 4; m++) {
            baS1 += (U2[m] - US1[m]) * bF[m];
            bU2[m] += aS1 * bF[m];
            bUS1[m] -= (aS1 - aS) * bF[m];
            baS += (US1[m] - US[m]) * bF[m];
            bUS[m] -= aS * bF[m];
            bFS[m] += bF[m];
          }
          for (int q = 0; q < 3; q++) {
            const int m = q == 0 ? RO : (q == 1 ? bt1 : bt2);
            baS1 += (U2[m] - US1[m]) * bF[m];
            bU2[m] += aS1 * bF[m];
            bUS1[m] -= (aS1 - aS) * bF[m];
            baS += (US1[m] - S[m]) * bF[m];
            bS[m] -= aS * bF[m];
            bFS[m] += bF[m];
          }
          // U2[mt] = roS at, U2[mn] = roS aM, U2[RO] = roS
          broS += at2 * bU2[mt2] + at1 * bU2[mt1] + aM * bU2[mn] + bU2[RO];
          bat2 += roS * bU2[mt2];
          bat1 += roS * bU2[mt1];
          baM += roS * bU2[mn];
          // U2[EN] = US1[EN] + sgn roSs K sB
          {
            const double b = bU2[EN];
            bUS1[EN] += b;
            broSs += sgn * K * sB * b;
            const double bK = sgn * roSs * sB * b;
            bvt1S += US1[bt1] * bK;
            bUS1[bt1] += vt1S * bK;
            bvt2S += US1[bt2] * bK;
            bUS1[bt2] += vt2S * bK;
            bat1 -= U2[bt1] * bK;
            bU2[bt1] -= at1 * bK;
            bat2 -= U2[bt2] * bK;
            bU2[bt2] -= at2 * bK;
          }
        };
        // sgn: L** is UL1[EN] - roLs(...)sB, R** is UR1[EN] + roRs(...)sB
        if (left)
          central_side(L, UL, UL1, bUL, bUL1, bFL, bL, aL, aL1, roL, roLs, vt1L, vt2L, -1.0, baL, baL1, broL, broLs,
                       bvt1L, bvt2L);
        else
          central_side(R, UR, UR1, bUR, bUR1, bFR, bR, aR, aR1, roR, roRs, vt1R, vt2R, 1.0, baR, baR1, broR, broRs,
                       bvt1R, bvt2R);
        double bfc = 0;
        // U2[bt] = fc * Nb
        {
          bfc += Nb1 * bU2[bt1] + Nb2 * bU2[bt2];
          const double n1 = fc * bU2[bt1], n2 = fc * bU2[bt2];
          broLs += UR1[bt1] * n1 + UR1[bt2] * n2;
          bUR1[bt1] += roLs * n1;
          bUR1[bt2] += roLs * n2;
          broRs += UL1[bt1] * n1 + UL1[bt2] * n2;
          bUL1[bt1] += roRs * n1;
          bUL1[bt2] += roRs * n2;
          const double bP = (vt1R - vt1L) * sB * n1 + (vt2R - vt2L) * sB * n2;
          const double P = roLs * roRs;
          bvt1R += P * sB * n1;
          bvt1L -= P * sB * n1;
          bvt2R += P * sB * n2;
          bvt2L -= P * sB * n2;
          broLs += roRs * bP;
          broRs += roLs * bP;
        }
        // at = fc * Na
        {
          bfc += Na1 * bat1 + Na2 * bat2;
          const double n1 = fc * bat1, n2 = fc * bat2;
          broLs += vt1L * n1 + vt2L * n2;
          bvt1L += roLs * n1;
          bvt2L += roLs * n2;
          broRs += vt1R * n1 + vt2R * n2;
          bvt1R += roRs * n1;
          bvt2R += roRs * n2;
          bUR1[bt1] += sB * n1;
          bUL1[bt1] -= sB * n1;
          bUR1[bt2] += sB * n2;
          bUL1[bt2] -= sB * n2;
        }
        // fc = 1 / (roLs + roRs); roXs = sqrt(roX)
        {
          const double bs = -fc * fc * bfc;
          broLs += bs;
          broRs += bs;
          broL += broLs * 0.5 * sroL / roL;  // d sqrt(ro) = d ro / (2 sqrt(ro))
          broR += broRs * 0.5 * sroR / roR;
        }
      }
      // aL1 = aM - aVL, aR1 = aM + aVR
      baM += baL1 + baR1;
      baVL -= baL1;
      baVR += baR1;
      // star states, right then left (statements of flux_solver.f90:482-507)
      for (int side = 1; side >= 0; side--) {
        const bool Ls = side == 0;
        const double *S = Ls ? L : R;
        const double *US = Ls ? UL : UR;
        const double *US1 = Ls ? UL1 : UR1;
        double *bUS1 = Ls ? bUL1 : bUR1;
        double *bUS = Ls ? bUL : bUR;
        double *bS = Ls ? bL : bR;
        const double aS = Ls ? aL : aR;
        const double roS = Ls ? roL : roR;
        const double ptS = Ls ? ptL : ptR;
        const double f1S = Ls ? f1L : f1R;
        const double eS = Ls ? eL : eR;
        const double PS_ = Ls ? PL : PR_;
        const double vt1S = Ls ? vt1L : vt1R, vt2S = Ls ? vt2L : vt2R;
        double &baS = Ls ? baL : baR;
        double &broS = Ls ? broL : broR;
        double &bptS = Ls ? bptL : bptR;
        double &bvBS = Ls ? bvBL : bvBR;
        double bvt1S = Ls ? bvt1L : bvt1R, bvt2S = Ls ? bvt2L : bvt2R;
        // US1[mt] = roS vtS, US1[mn] = roS aM, US1[RO] = roS
        broS += vt2S * bUS1[mt2] + vt1S * bUS1[mt1] + aM * bUS1[mn] + bUS1[RO];
        bvt2S += roS * bUS1[mt2];
        bvt1S += roS * bUS1[mt1];
        baM += roS * bUS1[mn];
        // US1[EN] = QS / (aS - aM)
        {
          const double bQ = bUS1[EN] / (aS - aM);
          const double bd = -bQ * US1[EN];
          baS += bd;
          baM -= bd;
          baS += US[EN] * bQ;
          bS[vn] -= US[EN] * bQ + ptS * bQ;
          bUS[EN] += (aS - S[vn]) * bQ;
          bptS -= S[vn] * bQ;
          bpt += aM * bQ;
          baM += pt * bQ;
          bUh[bn] += PS_ * bQ;
          const double bP = Uh[bn] * bQ;
          bvBS += bP;
          baM -= Uh[bn] * bP;
          bUh[bn] -= aM * bP;
          bvt1S -= US1[bt1] * bP;
          bUS1[bt1] -= vt1S * bP;
          bvt2S -= US1[bt2] * bP;
          bUS1[bt2] -= vt2S * bP;
        }
        double bf1S = 0, beS = 0;
        // vtS = S[vt] - f1S Uh[bn] S[bt] (aM - S[vn])
        for (int t = 0; t < 2; t++) {
          const int vt = t ? vt2 : vt1, bt = t ? bt2 : bt1;
          const double b = t ? bvt2S : bvt1S;
          const double G = f1S * Uh[bn] * S[bt];
          bS[vt] += b;
          const double bG = -(aM - S[vn]) * b;
          baM -= G * b;
          bS[vn] += G * b;
          bf1S += Uh[bn] * S[bt] * bG;
          bUh[bn] += f1S * S[bt] * bG;
          bS[bt] += f1S * Uh[bn] * bG;
        }
        // US1[bt] = S[bt] f1S eS
        for (int t = 0; t < 2; t++) {
          const int bt = t ? bt2 : bt1;
          bS[bt] += f1S * eS * bUS1[bt];
          bf1S += S[bt] * eS * bUS1[bt];
          beS += S[bt] * f1S * bUS1[bt];
        }
        // eS = S[RO] (aS - S[vn])^2 - Uh[bn]^2
        {
          const double dd = aS - S[vn];
          bS[RO] += dd * dd * beS;
          const double bdd = 2 * S[RO] * dd * beS;
          baS += bdd;
          bS[vn] -= bdd;
          bUh[bn] -= 2 * Uh[bn] * beS;
        }
        // f1S = 1 / dS, dS = S[RO] (aS - S[vn]) (aS - aM) - Uh[bn]^2
        {
          const double bd = -f1S * f1S * bf1S;
          const double p1 = aS - S[vn], p2 = aS - aM;
          bS[RO] += p1 * p2 * bd;
          baS += S[RO] * p2 * bd + S[RO] * p1 * bd;
          bS[vn] -= S[RO] * p2 * bd;
          baM -= S[RO] * p1 * bd;
          bUh[bn] -= 2 * Uh[bn] * bd;
        }
      }
    }
    // aVS = |Uh[bn]| / sqrt(roS)
    {
      const double sg = slope_abs(Uh[bn]);
      const double irL = 1.0 / roL, irR = 1.0 / roR;
      bUh[bn] += sg * (baVL * sroL * irL + baVR * sroR * irR);  // 1/sqrt(ro) = sqrt(ro)/ro
      broL -= 0.5 * baVL * aVL * irL;
      broR -= 0.5 * baVR * aVR * irR;
    }
    // roS = S[RO] (aS - S[vn]) / (aS - aM)
    for (int side = 1; side >= 0; side--) {
      const bool Ls = side == 0;
      const double *S = Ls ? L : R;
      double *bS = Ls ? bL : bR;
      const double aS = Ls ? aL : aR;
      const double roS = Ls ? roL : roR;
      const double b = Ls ? broL : broR;
      double &baS = Ls ? baL : baR;
      const double bnum = b / (aS - aM);
      const double bden = -bnum * roS;
      bS[RO] += (aS - S[vn]) * bnum;
      baS += S[RO] * bnum + bden;
      bS[vn] -= S[RO] * bnum;
      baM -= bden;
    }
    // pt = ptL + L[RO] (aL - L[vn]) (aM - L[vn])
    {
      bptL += bpt;
      const double c1 = L[RO] * (aL - L[vn]), c2 = aM - L[vn];
      const double bc1 = c2 * bpt, bc2 = c1 * bpt;
      bL[RO] += (aL - L[vn]) * bc1;
      baL += L[RO] * bc1;
      bL[vn] -= L[RO] * bc1 + bc2;
      baM += bc2;
    }
    // aM = Uh[mn] / Uh[RO]
    {
      const double iro = 1.0 / Uh[RO];
      bUh[mn] += baM * iro;
      bUh[RO] -= baM * aM * iro;
    }
    // Uh[m] = iw (aR XR[m] - aL XL[m] - FR[m] + FL[m])
    for (int m = 0; m < 8; m++) {
      if (m == 3) continue;
      const double XR = m < 3 ? UR[m] : R[m], XL = m < 3 ? UL[m] : L[m];
      biw += (aR * XR - aL * XL - FR[m] + FL[m]) * bUh[m];
      const double bN = iw * bUh[m];
      baR += XR * bN;
      baL -= XL * bN;
      if (m < 3) {
        bUR[m] += aR * bN;
        bUL[m] -= aL * bN;
      } else {
        bR[m] += aR * bN;
        bL[m] -= aL * bN;
      }
      bFR[m] -= bN;
      bFL[m] += bN;
    }
    // iw = 1 / (aR - aL)
    {
      const double bw = -iw * iw * biw;
      baR += bw;
      baL -= bw;
    }
  }
  // fan speeds: aL = min(vnL, vnR) - sa, aR = max(vnL, vnR) + sa, sa = sqrt(max(afL, afR))
  {
    if (L[vn] < R[vn]) bL[vn] += baL;
    else bR[vn] += baL;
    if (L[vn] > R[vn]) bL[vn] += baR;
    else bR[vn] += baR;
    const double bsa = baR - baL;
    const double bm = bsa / (2 * sa);
    double bafL = 0, bafR = 0;
    if (afL > afR) bafL = bm;
    else bafR = bm;
    hlld_side_vjp<D>(L, bUL, bFL, bafL, bptL, bvBL, bL);
    hlld_side_vjp<D>(R, bUR, bFR, bafR, bptR, bvBR, bR);
  }
}

}  // namespace adj
}  // namespace omhd
