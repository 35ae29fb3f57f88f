// omhd_device.h -- per-cell / per-interface arithmetic of the OpenMHD step for
// gfx950, shared by the drop-in kernels and the fused sweep kernels.
//
// Every expression keeps the reference's evaluation order (Fortran left to
// right) and the library is compiled with -ffp-contract=off, so with IEEE
// correctly rounded divide and sqrt (the gfx950 f64 sequences are) the results
// are bitwise equal to the reference CPU build.
//
// Slot convention for an interface / primitive state (9 doubles):
//   0 vx, 1 vy, 2 vz, 3 p, 4 rho, 5 Bx, 6 By, 7 Bz, 8 psi   (param.h:6-15)
// and for a conserved state: 0 mx, 1 my, 2 mz, 3 E, 4 rho, 5..8 as above.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace omhd {

constexpr int NV = 9;
constexpr double GAM = 5.0 / 3.0;           // param.h:3
constexpr double GM1 = GAM - 1;             // u2v.f90:11
constexpr double R1 = 1.0 / (GAM - 1.0);    // flux_solver.f90:38
constexpr double R2 = GAM / (GAM - 1.0);    // flux_solver.f90:39
constexpr double HLLG_FACTOR = 1.001;       // flux_solver.f90:54

// Fortran SIGN(a,b): |a| with the sign bit of b (copysign semantics, -0 -> -)
__device__ __forceinline__ double fsign(double a, double b) { return __builtin_copysign(__builtin_fabs(a), b); }
// MIN/MAX as the reference evaluates them.  They are only applied where the
// operands are ordered (no NaN in a valid state) and where a zero's sign
// cannot reach the result (absolute values in the limiters; the fan speeds
// min/max(vL, vR) -/+ cf of flux_solver.f90:408-410), so the single-instruction
// v_min_f64 / v_max_f64 give the reference's bits.
__device__ __forceinline__ double dmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ double dmin(double a, double b) { return __builtin_fmin(a, b); }

// Square root with the bits of __builtin_sqrt in fewer instructions.  hipcc
// expands an f64 sqrt for gfx950 into: a scale of x by 2^256 when x < 2^-767,
// v_rsq_f64, one coupled Newton step for g ~ sqrt(x) and h ~ 1/(2 sqrt(x)),
// two corrections g += (x - g*g)*h, the inverse scale by 2^-128, and a select
// that returns x itself for +-0 and +inf.  On [2^-767, DBL_MAX] the scale
// factors are 2^0 and the select is not taken, so there the core below (the
// same operations on the same operands, in the same order) gives the same
// result with 8 fewer instructions.  Every other argument (zero, denormals,
// inf, NaN, negatives) takes __builtin_sqrt.  The core runs on every lane and
// the test is one branch for the whole wavefront (taken only when some lane
// is out of range, then a per-lane select), not an if/else on the exec mask.
__device__ __forceinline__ double csqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  const bool in = x >= 0x1p-767 && x <= 0x1.fffffffffffffp+1023;
  if (__builtin_expect(__any(!in), 0)) return in ? g : __builtin_sqrt(x);
  return g;
}

// ---------------------------------------------------------------------------
// Wave-level lane shifts (64-wide wavefront, DPP wave_shr/wave_shl by 1).
// from_prev: lane l receives lane l-1 (lane 0 receives 0).
// from_next: lane l receives lane l+1 (lane 63 receives 0).
// bound_ctrl set with full row and bank masks: the lane without a source is
// written 0 by the DPP move itself, so no destination needs clearing first.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double from_prev(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_mov_dpp(lo, 0x138, 0xf, 0xf, true);
  hi = __builtin_amdgcn_mov_dpp(hi, 0x138, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_next(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_mov_dpp(lo, 0x130, 0xf, 0xf, true);
  hi = __builtin_amdgcn_mov_dpp(hi, 0x130, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

// ---------------------------------------------------------------------------
// Primitive recovery of one cell: u2v.f90:23-26.  w gets (vx,vy,vz,p,rho,B,psi)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cons_to_prim(const double u[NV], double w[NV]) {
  const double rho = u[4];
  const double v1 = u[0] / rho, v2 = u[1] / rho, v3 = u[2] / rho;
  const double rv2 = v1 * u[0] + v2 * u[1] + v3 * u[2];
  const double B2 = u[5] * u[5] + u[6] * u[6] + u[7] * u[7];
  w[0] = v1;
  w[1] = v2;
  w[2] = v3;
  w[3] = GM1 * (u[3] - 0.5 * (rv2 + B2));
  w[4] = rho;
  w[5] = u[5];
  w[6] = u[6];
  w[7] = u[7];
  w[8] = u[8];
}

// ---------------------------------------------------------------------------
// Slope limiters at one cell from its two neighbours along the sweep:
// limiter.f90 type 1 (minmod :84-101), 2 (MC :141-158), 3 (van Leer :201-210),
// 4 (Koren :256-272).  gR = right state of the lower interface, gL = left
// state of the upper interface.  Type 0 (flat) is gR = gL = w0.
// ---------------------------------------------------------------------------
template <int T>
__device__ __forceinline__ void limit(double wm, double w0, double wp, double &gR, double &gL) {
  if constexpr (T == 0) {
    gR = w0;
    gL = w0;
  } else if constexpr (T == 1) {
    const double gA = (w0 - wm), gB = (wp - w0);
    const double grad = (fsign(0.25, gA) + fsign(0.25, gB)) * dmin(__builtin_fabs(gA), __builtin_fabs(gB));
    gR = w0 - grad;
    gL = w0 + grad;
  } else if constexpr (T == 2) {
    const double gA = (w0 - wm), gB = (wp - w0);
    const double gC = 0.25 * (wp - wm);
    const double grad = (fsign(0.5, gA) + fsign(0.5, gB)) *
                        dmin(dmin(__builtin_fabs(gA), __builtin_fabs(gB)), __builtin_fabs(gC));
    gR = w0 - grad;
    gL = w0 + grad;
  } else if constexpr (T == 3) {
    const double gA = (w0 - wm), gB = (wp - w0);
    const double grad = (gA * gB <= 0) ? 0.0 : gA * gB / (gA + gB);
    gR = w0 - grad;
    gL = w0 + grad;
  } else {
    const double f1 = 1.0 / 6.0;
    const double gA = (w0 - wm), gB = (wp - w0);
    const double s = fsign(0.5, gA) + fsign(0.5, gB);
    const double m = dmin(__builtin_fabs(gA), __builtin_fabs(gB));
    double gC = f1 * (2 * gA + gB);
    gR = w0 - s * dmin(m, __builtin_fabs(gC));
    gC = f1 * (gA + 2 * gB);
    gL = w0 + s * dmin(m, __builtin_fabs(gC));
  }
}

// ---------------------------------------------------------------------------
// HLLD numerical flux at one interface (Miyoshi & Kusano 2005, with the
// HLLC-G fallback of Gurski 2004), flux_solver.f90:352-585.  D = 0,1,2 is the
// sweep direction; L, R are interface states in the slot convention above.
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void hlld(const double L[NV], const double R[NV], double ch, double ch2, double F[NV]) {
  constexpr int vn = D, vt1 = (D + 1) % 3, vt2 = (D + 2) % 3;  // :62-70
  constexpr int bn = 5 + vn, bt1 = 5 + vt1, bt2 = 5 + vt2;
  constexpr int RO = 4, PR = 3, EN = 3;

  // left: conserved momenta/energy, physical flux, fast speed^2, total pressure
  const double v2L = L[0] * L[0] + L[1] * L[1] + L[2] * L[2];
  const double B2L = L[5] * L[5] + L[6] * L[6] + L[7] * L[7];
  const double vBL = L[0] * L[5] + L[1] * L[6] + L[2] * L[7];
  double UL[4], FL[NV];
  UL[0] = L[RO] * L[0];
  UL[1] = L[RO] * L[1];
  UL[2] = L[RO] * L[2];
  UL[EN] = 0.5 * (L[RO] * v2L + B2L) + R1 * L[PR];
  FL[RO] = L[RO] * L[vn];
  FL[vn] = L[RO] * L[vn] * L[vn] - L[bn] * L[bn] + L[PR] + 0.5 * B2L;
  FL[vt1] = L[RO] * L[vn] * L[vt1] - L[bn] * L[bt1];
  FL[vt2] = L[RO] * L[vn] * L[vt2] - L[bn] * L[bt2];
  FL[EN] = (0.5 * L[RO] * v2L + R2 * L[PR] + B2L) * L[vn] - vBL * L[bn];
  FL[bt1] = L[vn] * L[bt1] - L[bn] * L[vt1];
  FL[bt2] = L[vn] * L[bt2] - L[bn] * L[vt2];
  FL[bn] = 0.0;
  FL[8] = 0.0;
  double f1 = GAM * L[PR];
  double f2 = 4 * (L[bn] * L[bn]);
  const double cfL = ((f1 + B2L) + csqrt(dmax((f1 + B2L) * (f1 + B2L) - f1 * f2, 0.0))) / (2 * L[RO]);
  const double ptL = L[PR] + 0.5 * B2L;

  // right
  const double v2R = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
  const double B2R = R[5] * R[5] + R[6] * R[6] + R[7] * R[7];
  const double vBR = R[0] * R[5] + R[1] * R[6] + R[2] * R[7];
  double UR[4], FR[NV];
  UR[0] = R[RO] * R[0];
  UR[1] = R[RO] * R[1];
  UR[2] = R[RO] * R[2];
  UR[EN] = 0.5 * (R[RO] * v2R + B2R) + R1 * R[PR];
  FR[RO] = R[RO] * R[vn];
  FR[vn] = R[RO] * R[vn] * R[vn] - R[bn] * R[bn] + R[PR] + 0.5 * B2R;
  FR[vt1] = R[RO] * R[vn] * R[vt1] - R[bn] * R[bt1];
  FR[vt2] = R[RO] * R[vn] * R[vt2] - R[bn] * R[bt2];
  FR[EN] = (0.5 * R[RO] * v2R + R2 * R[PR] + B2R) * R[vn] - vBR * R[bn];
  FR[bt1] = R[vn] * R[bt1] - R[bn] * R[vt1];
  FR[bt2] = R[vn] * R[bt2] - R[bn] * R[vt2];
  FR[bn] = 0.0;
  FR[8] = 0.0;
  f1 = GAM * R[PR];
  f2 = 4 * (R[bn] * R[bn]);
  const double cfR = ((f1 + B2R) + csqrt(dmax((f1 + B2R) * (f1 + B2R) - f1 * f2, 0.0))) / (2 * R[RO]);
  const double ptR = R[PR] + 0.5 * B2R;

  // Riemann fan (:408-410): one shared fast speed
  f1 = csqrt(dmax(cfL, cfR));
  const double aL = dmin(L[vn], R[vn]) - f1;
  const double aR = dmax(L[vn], R[vn]) + f1;

  if (aL >= 0) {
#pragma unroll
    for (int m = 0; m < NV; m++) F[m] = FL[m];
  } else if (aR <= 0) {
#pragma unroll
    for (int m = 0; m < NV; m++) F[m] = FR[m];
  } else {
    // HLL average state (:423-425)
    const double fh = 1.0 / (aR - aL);
    double Uh[NV];
#pragma unroll
    for (int m = 0; m < 3; m++) Uh[m] = fh * (aR * UR[m] - aL * UL[m] - FR[m] + FL[m]);
#pragma unroll
    for (int m = RO; m <= 7; m++) Uh[m] = fh * (aR * R[m] - aL * L[m] - FR[m] + FL[m]);
    const double aM = Uh[vn] / Uh[RO];                              // :428
    const double pt = ptL + L[RO] * (aL - L[vn]) * (aM - L[vn]);    // :431
    const double roL = L[RO] * (aL - L[vn]) / (aL - aM);            // :432
    const double roR = R[RO] * (aR - R[vn]) / (aR - aM);            // :433
    const double sroL = csqrt(roL), sroR = csqrt(roR);
    const double aVL = __builtin_fabs(Uh[bn]) / sroL;              // :436
    const double aVR = __builtin_fabs(Uh[bn]) / sroR;              // :437
    const double hb = Uh[bn];

    if ((aL >= (aM - HLLG_FACTOR * aVL)) || ((aM + HLLG_FACTOR * aVR) >= aR)) {
      // HLLC-G (:447-476)
      const double g2 = aL * aR;
      F[bt1] = fh * (aR * FL[bt1] - aL * FR[bt1] + g2 * (R[bt1] - L[bt1]));
      F[bt2] = fh * (aR * FL[bt2] - aL * FR[bt2] + g2 * (R[bt2] - L[bt2]));
      const double at1 = Uh[vt1] / Uh[RO];
      const double at2 = Uh[vt2] / Uh[RO];
      const double enL = ((aL - L[vn]) * UL[EN] - ptL * L[vn] + pt * aM +
                          hb * (vBL - aM * hb - at1 * Uh[bt1] - at2 * Uh[bt2])) / (aL - aM);
      const double enR = ((aR - R[vn]) * UR[EN] - ptR * R[vn] + pt * aM +
                          hb * (vBR - aM * hb - at1 * Uh[bt1] - at2 * Uh[bt2])) / (aR - aM);
      const double wl = 0.5 + fsign(0.5, aM);
      const double wr = 1.0 - wl;
      F[vn] = wl * (aL * (roL * aM - UL[vn]) + FL[vn]) + wr * (aR * (roR * aM - UR[vn]) + FR[vn]);
      F[vt1] = wl * (aL * (roL * at1 - UL[vt1]) + FL[vt1]) + wr * (aR * (roR * at1 - UR[vt1]) + FR[vt1]);
      F[vt2] = wl * (aL * (roL * at2 - UL[vt2]) + FL[vt2]) + wr * (aR * (roR * at2 - UR[vt2]) + FR[vt2]);
      F[EN] = wl * (aL * (enL - UL[EN]) + FL[EN]) + wr * (aR * (enR - UR[EN]) + FR[EN]);
      F[RO] = wl * (aL * (roL - L[RO]) + FL[RO]) + wr * (aR * (roR - R[RO]) + FR[RO]);
    } else {
      // HLLD intermediate states L*, R* (:482-507)
      double UL1[NV], UR1[NV];
      const double hb2 = hb * hb;
      double fd = 1.0 / (L[RO] * (aL - L[vn]) * (aL - aM) - hb2);
      const double dL = aL - L[vn];
      UL1[bt1] = L[bt1] * fd * (L[RO] * (dL * dL) - hb2);
      UL1[bt2] = L[bt2] * fd * (L[RO] * (dL * dL) - hb2);
      const double vt1L = L[vt1] - fd * hb * L[bt1] * (aM - L[vn]);
      const double vt2L = L[vt2] - fd * hb * L[bt2] * (aM - L[vn]);
      UL1[RO] = roL;
      UL1[EN] = ((aL - L[vn]) * UL[EN] - ptL * L[vn] + pt * aM +
                 hb * (vBL - aM * hb - vt1L * UL1[bt1] - vt2L * UL1[bt2])) / (aL - aM);
      UL1[vn] = roL * aM;
      UL1[vt1] = roL * vt1L;
      UL1[vt2] = roL * vt2L;

      fd = 1.0 / (R[RO] * (aR - R[vn]) * (aR - aM) - hb2);
      const double dR = aR - R[vn];
      UR1[bt1] = R[bt1] * fd * (R[RO] * (dR * dR) - hb2);
      UR1[bt2] = R[bt2] * fd * (R[RO] * (dR * dR) - hb2);
      const double vt1R = R[vt1] - fd * hb * R[bt1] * (aM - R[vn]);
      const double vt2R = R[vt2] - fd * hb * R[bt2] * (aM - R[vn]);
      UR1[RO] = roR;
      UR1[EN] = ((aR - R[vn]) * UR[EN] - ptR * R[vn] + pt * aM +
                 hb * (vBR - aM * hb - vt1R * UR1[bt1] - vt2R * UR1[bt2])) / (aR - aM);
      UR1[vn] = roR * aM;
      UR1[vt1] = roR * vt1R;
      UR1[vt2] = roR * vt2R;

      const double aL1 = aM - aVL;  // rotational waves (:510-511)
      const double aR1 = aM + aVR;
      if (aL1 >= 0) {  // F(L*) (:514-519)
#pragma unroll
        for (int m = 0; m < 4; m++) F[m] = aL * (UL1[m] - UL[m]) + FL[m];
        F[RO] = aL * (UL1[RO] - L[RO]) + FL[RO];
        F[bt1] = aL * (UL1[bt1] - L[bt1]) + FL[bt1];
        F[bt2] = aL * (UL1[bt2] - L[bt2]) + FL[bt2];
      } else if (aR1 <= 0) {  // F(R*) (:522-527)
#pragma unroll
        for (int m = 0; m < 4; m++) F[m] = aR * (UR1[m] - UR[m]) + FR[m];
        F[RO] = aR * (UR1[RO] - R[RO]) + FR[RO];
        F[bt1] = aR * (UR1[bt1] - R[bt1]) + FR[bt1];
        F[bt2] = aR * (UR1[bt2] - R[bt2]) + FR[bt2];
      } else {  // central states L**, R** (:533-572)
        const double fs = 1.0 / (sroL + sroR);
        const double sg = fsign(1.0, hb);
        const double at1 = fs * (sroL * vt1L + sroR * vt1R + (UR1[bt1] - UL1[bt1]) * sg);
        const double at2 = fs * (sroL * vt2L + sroR * vt2R + (UR1[bt2] - UL1[bt2]) * sg);
        double U2[NV];
        U2[bt1] = fs * (sroL * UR1[bt1] + sroR * UL1[bt1] + sroL * sroR * (vt1R - vt1L) * sg);
        U2[bt2] = fs * (sroL * UR1[bt2] + sroR * UL1[bt2] + sroL * sroR * (vt2R - vt2L) * sg);
        if (aM >= 0) {
          U2[RO] = roL;
          U2[EN] = UL1[EN] - sroL * (vt1L * UL1[bt1] + vt2L * UL1[bt2] - at1 * U2[bt1] - at2 * U2[bt2]) * sg;
          U2[vn] = roL * aM;
          U2[vt1] = roL * at1;
          U2[vt2] = roL * at2;
#pragma unroll
          for (int m = 0; m < 4; m++) F[m] = aL1 * U2[m] - (aL1 - aL) * UL1[m] - aL * UL[m] + FL[m];
          F[RO] = aL1 * U2[RO] - (aL1 - aL) * UL1[RO] - aL * L[RO] + FL[RO];
          F[bt1] = aL1 * U2[bt1] - (aL1 - aL) * UL1[bt1] - aL * L[bt1] + FL[bt1];
          F[bt2] = aL1 * U2[bt2] - (aL1 - aL) * UL1[bt2] - aL * L[bt2] + FL[bt2];
        } else {
          U2[RO] = roR;
          U2[EN] = UR1[EN] + sroR * (vt1R * UR1[bt1] + vt2R * UR1[bt2] - at1 * U2[bt1] - at2 * U2[bt2]) * sg;
          U2[vn] = roR * aM;
          U2[vt1] = roR * at1;
          U2[vt2] = roR * at2;
#pragma unroll
          for (int m = 0; m < 4; m++) F[m] = aR1 * U2[m] - (aR1 - aR) * UR1[m] - aR * UR[m] + FR[m];
          F[RO] = aR1 * U2[RO] - (aR1 - aR) * UR1[RO] - aR * R[RO] + FR[RO];
          F[bt1] = aR1 * U2[bt1] - (aR1 - aR) * UR1[bt1] - aR * R[bt1] + FR[bt1];
          F[bt2] = aR1 * U2[bt2] - (aR1 - aR) * UR1[bt2] - aR * R[bt2] + FR[bt2];
        }
      }
    }
  }
  // GLM pair, always (:580-581)
  F[bn] = 0.5 * ((L[8] + R[8]) - ch * (R[bn] - L[bn]));
  F[8] = 0.5 * (ch2 * (L[bn] + R[bn]) - ch * (R[8] - L[8]));
}

// ---------------------------------------------------------------------------
// The other flux types of flux_solver.f90 at one interface (the type is
// uniform over a launch): 0 LLF / Rusanov (:89-148), 1 HLL (:153-221),
// 2 HLLC (:226-347), 4 LHLLD (:590-789).  Statements in the reference's order.
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void side_state(const double S[NV], double U[4], double Fs[NV], double &B2, double &vB,
                                           double &v2) {
  constexpr int vn = D, vt1 = (D + 1) % 3, vt2 = (D + 2) % 3;
  constexpr int bn = 5 + vn, bt1 = 5 + vt1, bt2 = 5 + vt2;
  constexpr int RO = 4, PR = 3, EN = 3;
  v2 = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
  B2 = S[5] * S[5] + S[6] * S[6] + S[7] * S[7];
  vB = S[0] * S[5] + S[1] * S[6] + S[2] * S[7];
  U[0] = S[RO] * S[0];
  U[1] = S[RO] * S[1];
  U[2] = S[RO] * S[2];
  U[EN] = 0.5 * (S[RO] * v2 + B2) + R1 * S[PR];
  Fs[RO] = S[RO] * S[vn];
  Fs[vn] = S[RO] * S[vn] * S[vn] - S[bn] * S[bn] + S[PR] + 0.5 * B2;
  Fs[vt1] = S[RO] * S[vn] * S[vt1] - S[bn] * S[bt1];
  Fs[vt2] = S[RO] * S[vn] * S[vt2] - S[bn] * S[bt2];
  Fs[EN] = (0.5 * S[RO] * v2 + R2 * S[PR] + B2) * S[vn] - vB * S[bn];
  Fs[bt1] = S[vn] * S[bt1] - S[bn] * S[vt1];
  Fs[bt2] = S[vn] * S[bt2] - S[bn] * S[vt2];
  Fs[bn] = 0.0;
  Fs[8] = 0.0;
}

// fast-mode speed^2 of one side, with the max(., 0) guard; with rv2 in place
// of gamma p it is LHLLD's second estimate
__device__ __forceinline__ double fast2(double f1, double B2, double bnv, double ro) {
  const double f2 = 4 * (bnv * bnv);
  return ((f1 + B2) + csqrt(dmax((f1 + B2) * (f1 + B2) - f1 * f2, 0.0))) / (2 * ro);
}

template <int D>
__device__ __forceinline__ void riemann_other(int type, const double L[NV], const double R[NV], double ch, double ch2,
                                              double F[NV]) {
  constexpr int vn = D, vt1 = (D + 1) % 3, vt2 = (D + 2) % 3;
  constexpr int bn = 5 + vn, bt1 = 5 + vt1, bt2 = 5 + vt2;
  constexpr int RO = 4, PR = 3, EN = 3;
  double UL[4], UR[4], FL[NV], FR[NV], B2L, vBL, v2L, B2R, vBR, v2R;
  side_state<D>(L, UL, FL, B2L, vBL, v2L);
  side_state<D>(R, UR, FR, B2R, vBR, v2R);
#pragma unroll
  for (int m = 0; m < NV; m++) F[m] = 0.0;
  if (type == 0) {  // LLF: no guard under the inner sqrt (:112-113)
    double f1 = GAM * L[PR], f2 = 4 * (L[bn] * L[bn]);
    const double aL = csqrt(((f1 + B2L) + csqrt((f1 + B2L) * (f1 + B2L) - f1 * f2)) / (2 * L[RO]));
    f1 = GAM * R[PR];
    f2 = 4 * (R[bn] * R[bn]);
    const double aR = csqrt(((f1 + B2R) + csqrt((f1 + B2R) * (f1 + B2R) - f1 * f2)) / (2 * R[RO]));
    f1 = dmax(__builtin_fabs(L[vn]) + aL, __builtin_fabs(R[vn]) + aR);
#pragma unroll
    for (int m = 0; m < 4; m++) F[m] = 0.5 * (FL[m] + FR[m] - f1 * (UR[m] - UL[m]));
    F[RO] = 0.5 * (FL[RO] + FR[RO] - f1 * (R[RO] - L[RO]));
    F[bt1] = 0.5 * (FL[bt1] + FR[bt1] - f1 * (R[bt1] - L[bt1]));
    F[bt2] = 0.5 * (FL[bt2] + FR[bt2] - f1 * (R[bt2] - L[bt2]));
  } else if (type == 1 || type == 2) {  // HLL, HLLC: fan speeds clamped at 0 (:207-208)
    double aL = fast2(GAM * L[PR], B2L, L[bn], L[RO]), aR = fast2(GAM * R[PR], B2R, R[bn], R[RO]);
    const double ptL = L[PR] + 0.5 * B2L, ptR = R[PR] + 0.5 * B2R;
    double f1 = csqrt(dmax(aL, aR));
    aL = dmin(dmin(L[vn], R[vn]) - f1, 0.0);
    aR = dmax(dmax(L[vn], R[vn]) + f1, 0.0);
    f1 = 1.0 / (aR - aL);
    double f2 = aL * aR;
    F[bt1] = f1 * (aR * FL[bt1] - aL * FR[bt1] + f2 * (R[bt1] - L[bt1]));
    F[bt2] = f1 * (aR * FL[bt2] - aL * FR[bt2] + f2 * (R[bt2] - L[bt2]));
    if (type == 1) {
#pragma unroll
      for (int m = 0; m < 4; m++) F[m] = f1 * (aR * FL[m] - aL * FR[m] + f2 * (UR[m] - UL[m]));
      F[RO] = f1 * (aR * FL[RO] - aL * FR[RO] + f2 * (R[RO] - L[RO]));
    } else {
      double Uh[NV];
#pragma unroll
      for (int m = 0; m < 3; m++) Uh[m] = f1 * (aR * UR[m] - aL * UL[m] - FR[m] + FL[m]);
#pragma unroll
      for (int m = RO; m <= 7; m++) Uh[m] = f1 * (aR * R[m] - aL * L[m] - FR[m] + FL[m]);
      f2 = 1 / Uh[RO];
      const double aM = Uh[vn] * f2, at1 = Uh[vt1] * f2, at2 = Uh[vt2] * f2;
      const double pt = ptL + L[RO] * (aL - L[vn]) * (aM - L[vn]);
      const double roL = L[RO] * (aL - L[vn]) / (aL - aM);
      const double enL = ((aL - L[vn]) * UL[EN] - ptL * L[vn] + pt * aM +
                          Uh[bn] * (vBL - aM * Uh[bn] - at1 * Uh[bt1] - at2 * Uh[bt2])) / (aL - aM);
      const double roR = R[RO] * (aR - R[vn]) / (aR - aM);
      const double enR = ((aR - R[vn]) * UR[EN] - ptR * R[vn] + pt * aM +
                          Uh[bn] * (vBR - aM * Uh[bn] - at1 * Uh[bt1] - at2 * Uh[bt2])) / (aR - aM);
      const double wl = 0.5 + fsign(0.5, aM);
      const double wr = 1.0 - wl;
      F[vn] = wl * (aL * (roL * aM - UL[vn]) + FL[vn]) + wr * (aR * (roR * aM - UR[vn]) + FR[vn]);
      F[vt1] = wl * (aL * (roL * at1 - UL[vt1]) + FL[vt1]) + wr * (aR * (roR * at1 - UR[vt1]) + FR[vt1]);
      F[vt2] = wl * (aL * (roL * at2 - UL[vt2]) + FL[vt2]) + wr * (aR * (roR * at2 - UR[vt2]) + FR[vt2]);
      F[EN] = wl * (aL * (enL - UL[EN]) + FL[EN]) + wr * (aR * (enR - UR[EN]) + FR[EN]);
      F[RO] = wl * (aL * (roL - L[RO]) + FL[RO]) + wr * (aR * (roR - R[RO]) + FR[RO]);
    }
  } else {  // LHLLD (type 4)
    double aL = fast2(GAM * L[PR], B2L, L[bn], L[RO]), aR = fast2(GAM * R[PR], B2R, R[bn], R[RO]);
    const double ptL = L[PR] + 0.5 * B2L, ptR = R[PR] + 0.5 * B2R;
    double aL1 = fast2(L[RO] * v2L, B2L, L[bn], L[RO]), aR1 = fast2(R[RO] * v2R, B2R, R[bn], R[RO]);
    double f1 = dmin(1.0, dmax(aL1, aR1) / dmax(aL, aR));  // chi^2
    const double psi = 2 * csqrt(f1) - f1;
    f1 = csqrt(dmax(aL, aR));
    aL = dmin(L[vn], R[vn]) - f1;
    aR = dmax(L[vn], R[vn]) + f1;
    if (aL >= 0) {
#pragma unroll
      for (int m = 0; m < NV; m++) F[m] = FL[m];
    } else if (aR <= 0) {
#pragma unroll
      for (int m = 0; m < NV; m++) F[m] = FR[m];
    } else {
      double Uh[NV], UL1[NV], UR1[NV];
      f1 = 1.0 / (aR - aL);
#pragma unroll
      for (int m = 0; m < 3; m++) Uh[m] = f1 * (aR * UR[m] - aL * UL[m] - FR[m] + FL[m]);
#pragma unroll
      for (int m = RO; m <= 7; m++) Uh[m] = f1 * (aR * R[m] - aL * L[m] - FR[m] + FL[m]);
      const double aM = Uh[vn] / Uh[RO];
      const double pt = (R[RO] * ptL * (aR - R[vn]) - L[RO] * ptR * (aL - L[vn]) +
                         psi * L[RO] * R[RO] * (aR - R[vn]) * (aL - L[vn]) * (R[vn] - L[vn])) /
                        (R[RO] * (aR - R[vn]) - L[RO] * (aL - L[vn]));
      const double roL = L[RO] * (aL - L[vn]) / (aL - aM);
      const double roR = R[RO] * (aR - R[vn]) / (aR - aM);
      const double sroL = csqrt(roL), sroR = csqrt(roR);
      const double aVL = __builtin_fabs(Uh[bn]) / sroL;
      const double aVR = __builtin_fabs(Uh[bn]) / sroR;
      const double hb = Uh[bn], hb2 = hb * hb;
      double fd = 1.0 / (L[RO] * (aL - L[vn]) * (aL - aM) - hb2);
      const double dL = aL - L[vn];
      UL1[bt1] = L[bt1] * fd * (L[RO] * (dL * dL) - hb2);
      UL1[bt2] = L[bt2] * fd * (L[RO] * (dL * dL) - hb2);
      const double vt1L = L[vt1] - fd * hb * L[bt1] * (aM - L[vn]);
      const double vt2L = L[vt2] - fd * hb * L[bt2] * (aM - L[vn]);
      UL1[RO] = roL;
      UL1[EN] = ((aL - L[vn]) * UL[EN] - ptL * L[vn] + pt * aM +
                 hb * (vBL - aM * hb - vt1L * UL1[bt1] - vt2L * UL1[bt2])) / (aL - aM);
      UL1[vn] = roL * aM;
      UL1[vt1] = roL * vt1L;
      UL1[vt2] = roL * vt2L;
      fd = 1.0 / (R[RO] * (aR - R[vn]) * (aR - aM) - hb2);
      const double dR = aR - R[vn];
      UR1[bt1] = R[bt1] * fd * (R[RO] * (dR * dR) - hb2);
      UR1[bt2] = R[bt2] * fd * (R[RO] * (dR * dR) - hb2);
      const double vt1R = R[vt1] - fd * hb * R[bt1] * (aM - R[vn]);
      const double vt2R = R[vt2] - fd * hb * R[bt2] * (aM - R[vn]);
      UR1[RO] = roR;
      UR1[EN] = ((aR - R[vn]) * UR[EN] - ptR * R[vn] + pt * aM +
                 hb * (vBR - aM * hb - vt1R * UR1[bt1] - vt2R * UR1[bt2])) / (aR - aM);
      UR1[vn] = roR * aM;
      UR1[vt1] = roR * vt1R;
      UR1[vt2] = roR * vt2R;
      aL1 = aM - aVL;
      aR1 = aM + aVR;
      if (aL1 >= 0) {
#pragma unroll
        for (int m = 0; m < 4; m++) F[m] = aL * (UL1[m] - UL[m]) + FL[m];
        F[RO] = aL * (UL1[RO] - L[RO]) + FL[RO];
        F[bt1] = aL * (UL1[bt1] - L[bt1]) + FL[bt1];
        F[bt2] = aL * (UL1[bt2] - L[bt2]) + FL[bt2];
      } else if (aR1 <= 0) {
#pragma unroll
        for (int m = 0; m < 4; m++) F[m] = aR * (UR1[m] - UR[m]) + FR[m];
        F[RO] = aR * (UR1[RO] - R[RO]) + FR[RO];
        F[bt1] = aR * (UR1[bt1] - R[bt1]) + FR[bt1];
        F[bt2] = aR * (UR1[bt2] - R[bt2]) + FR[bt2];
      } else {
        const double fs = 1.0 / (sroL + sroR);
        const double sg = fsign(1.0, hb);
        const double at1 = fs * (sroL * vt1L + sroR * vt1R + (UR1[bt1] - UL1[bt1]) * sg);
        const double at2 = fs * (sroL * vt2L + sroR * vt2R + (UR1[bt2] - UL1[bt2]) * sg);
        double U2[NV];
        U2[bt1] = fs * (sroL * UR1[bt1] + sroR * UL1[bt1] + sroL * sroR * (vt1R - vt1L) * sg);
        U2[bt2] = fs * (sroL * UR1[bt2] + sroR * UL1[bt2] + sroL * sroR * (vt2R - vt2L) * sg);
        if (aM >= 0) {
          U2[RO] = roL;
          U2[EN] = UL1[EN] - sroL * (vt1L * UL1[bt1] + vt2L * UL1[bt2] - at1 * U2[bt1] - at2 * U2[bt2]) * sg;
          U2[vn] = roL * aM;
          U2[vt1] = roL * at1;
          U2[vt2] = roL * at2;
#pragma unroll
          for (int m = 0; m < 4; m++) F[m] = aL1 * U2[m] - (aL1 - aL) * UL1[m] - aL * UL[m] + FL[m];
          F[RO] = aL1 * U2[RO] - (aL1 - aL) * UL1[RO] - aL * L[RO] + FL[RO];
          F[bt1] = aL1 * U2[bt1] - (aL1 - aL) * UL1[bt1] - aL * L[bt1] + FL[bt1];
          F[bt2] = aL1 * U2[bt2] - (aL1 - aL) * UL1[bt2] - aL * L[bt2] + FL[bt2];
        } else {
          U2[RO] = roR;
          U2[EN] = UR1[EN] + sroR * (vt1R * UR1[bt1] + vt2R * UR1[bt2] - at1 * U2[bt1] - at2 * U2[bt2]) * sg;
          U2[vn] = roR * aM;
          U2[vt1] = roR * at1;
          U2[vt2] = roR * at2;
#pragma unroll
          for (int m = 0; m < 4; m++) F[m] = aR1 * U2[m] - (aR1 - aR) * UR1[m] - aR * UR[m] + FR[m];
          F[RO] = aR1 * U2[RO] - (aR1 - aR) * UR1[RO] - aR * R[RO] + FR[RO];
          F[bt1] = aR1 * U2[bt1] - (aR1 - aR) * UR1[bt1] - aR * R[bt1] + FR[bt1];
          F[bt2] = aR1 * U2[bt2] - (aR1 - aR) * UR1[bt2] - aR * R[bt2] + FR[bt2];
        }
      }
    }
  }
  F[bn] = 0.5 * ((L[8] + R[8]) - ch * (R[bn] - L[bn]));
  F[8] = 0.5 * (ch2 * (L[bn] + R[bn]) - ch * (R[8] - L[8]));
}

// ---------------------------------------------------------------------------
// Fastest magnetosonic signal speed of one cell, set_dt.f90:27-45 (2D) and
// 3D_basic/set_dt.f90 (ND == 3).  u: conserved, w: primitives from u.
// ---------------------------------------------------------------------------
template <int ND>
__device__ __forceinline__ double fastest_speed(const double u[NV], const double w[NV]) {
  const double B2 = u[5] * u[5] + u[6] * u[6] + u[7] * u[7];
  const double f1 = GAM * w[3];
  const double ro2 = 2 * u[4];
  const double s = f1 + B2;
  double f2 = 4 * f1 * (u[5] * u[5]);
  const double vfx = csqrt((s + csqrt(dmax(s * s - f2, 0.0))) / ro2);
  f2 = 4 * f1 * (u[6] * u[6]);
  const double vfy = csqrt((s + csqrt(dmax(s * s - f2, 0.0))) / ro2);
  double vt = dmax(__builtin_fabs(w[0]) + vfx, __builtin_fabs(w[1]) + vfy);
  if constexpr (ND == 3) {
    f2 = 4 * f1 * (u[7] * u[7]);
    const double vfz = csqrt((s + csqrt(dmax(s * s - f2, 0.0))) / ro2);
    vt = dmax(vt, __builtin_fabs(w[2]) + vfz);
  }
  return vt;
}

// Monotone max for non-negative doubles through their bit patterns.
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long *p, double v) {
  atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

// Wave-wide max of a double (64 lanes) via cross-lane reduction.
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = dmax(v, __shfl_xor(v, off, 64));
  return v;
}

}  // namespace omhd
