// flopcount.cpp -- the floating-point operation count of one zone-update of
// the 2D step (u2v, 2 x 9 limiter calls, 2 HLLD interfaces and the RK combine
// per stage, two GLM scalings and the CFL speed per step).  It runs the
// forward forms of openmhd_amd/csrc/omhd_adjoint.h, which in double give the
// reference's results bit for bit, on a scalar type that counts what it does.
// Test infrastructure (measurement), not part of the library.
//
// Counted: +, - (binary), *, / and sqrt on data values, constants included
// (0.5*x is one multiply).  Not counted: negation, |x|, min/max and the
// comparisons that pick branches (selects, not arithmetic).  The forward forms
// keep two kinds of arithmetic in plain double, added here by hand: the sum of
// the two sign selections of the minmod, MC and Koren limiters (one add per
// call, limiter.f90:98,156,265) and s1 = 0.5 + sign, s2 = 1 - s1 of the HLLC-G
// branch (two adds per such interface, flux_solver.f90:447-476).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

namespace cnt {
struct Counts {
  long add = 0, mul = 0, div = 0, sqrt = 0;
};
static Counts C;

struct Cnt {
  double v;
  Cnt() : v(0) {}
  Cnt(double a) : v(a) {}  // NOLINT: implicit, constants enter as values
};
inline Cnt operator+(const Cnt &a, const Cnt &b) { C.add++; return Cnt(a.v + b.v); }
inline Cnt operator-(const Cnt &a, const Cnt &b) { C.add++; return Cnt(a.v - b.v); }
inline Cnt operator-(const Cnt &a) { return Cnt(-a.v); }
inline Cnt operator*(const Cnt &a, const Cnt &b) { C.mul++; return Cnt(a.v * b.v); }
inline Cnt operator/(const Cnt &a, const Cnt &b) { C.div++; return Cnt(a.v / b.v); }
inline Cnt operator+(double a, const Cnt &b) { C.add++; return Cnt(a + b.v); }
inline Cnt operator+(const Cnt &a, double b) { C.add++; return Cnt(a.v + b); }
inline Cnt operator-(double a, const Cnt &b) { C.add++; return Cnt(a - b.v); }
inline Cnt operator-(const Cnt &a, double b) { C.add++; return Cnt(a.v - b); }
inline Cnt operator*(double a, const Cnt &b) { C.mul++; return Cnt(a * b.v); }
inline Cnt operator*(const Cnt &a, double b) { C.mul++; return Cnt(a.v * b); }
inline Cnt operator/(double a, const Cnt &b) { C.div++; return Cnt(a / b.v); }
inline Cnt operator/(const Cnt &a, double b) { C.div++; return Cnt(a.v / b); }
inline double val(const Cnt &a) { return a.v; }
inline Cnt tsqrt(const Cnt &a) { C.sqrt++; return Cnt(sqrt(a.v)); }
inline Cnt tabs(const Cnt &a) { return Cnt(fabs(a.v)); }
inline Cnt tmin(const Cnt &a, const Cnt &b) { return a.v < b.v ? a : b; }
inline Cnt tmax(const Cnt &a, const Cnt &b) { return a.v > b.v ? a : b; }
}  // namespace cnt

#include "../../openmhd_amd/csrc/omhd_adjoint.h"

using cnt::Cnt;
using namespace omhd::adj;

namespace {

constexpr double kGamma = 5.0 / 3.0;  // param.h:3

// the per-cell fast-mode speed of set_dt.f90:28-44 (2D)
Cnt cfl_speed(const Cnt u[NV], const Cnt w[NV]) {
  const Cnt B2 = u[5] * u[5] + u[6] * u[6] + u[7] * u[7];
  const Cnt f1 = kGamma * w[3];
  Cnt f2 = 4 * f1 * (u[5] * u[5]);
  const Cnt vfx = tsqrt(((f1 + B2) + tsqrt(tmax((f1 + B2) * (f1 + B2) - f2, Cnt(0.0)))) / (2 * u[4]));
  f2 = 4 * f1 * (u[6] * u[6]);
  const Cnt vfy = tsqrt(((f1 + B2) + tsqrt(tmax((f1 + B2) * (f1 + B2) - f2, Cnt(0.0)))) / (2 * u[4]));
  return tmax(tabs(w[0]) + vfx, tabs(w[1]) + vfy);
}

bool lim_has_sign_sum(int lm) { return lm == 1 || lm == 2 || lm == 4; }

}  // namespace

extern "C" {

// One u2v and one limiter call, for pinning the counter against hand counts
// of u2v.f90:23-26 and limiter.f90.  out: add, mul, div, sqrt.
void flops_unit(int which, int lm, long out[4]) {
  cnt::C = cnt::Counts();
  Cnt u[NV] = {0.1, -0.2, 0.3, 2.0, 1.1, 0.4, -0.5, 0.6, 0.01}, w[NV];
  if (which == 0) {
    prim_fwd<Cnt>(u, w);
  } else {
    Cnt gR, gL;
    lim_fwd<Cnt>(lm, Cnt(1.0), Cnt(1.5), Cnt(1.7), gR, gL);
    if (lim_has_sign_sum(lm)) cnt::C.add++;
  }
  out[0] = cnt::C.add, out[1] = cnt::C.mul, out[2] = cnt::C.div, out[3] = cnt::C.sqrt;
}

// Operation counts of one step (both stages at the given state) summed over a
// periodic nx x ny grid; U is SoA (9, ny, nx), x fastest.  out[0..3]: add,
// mul, div, sqrt over the grid; out[4..7]: the part spent in HLLD;
// branches[1..8]: HLLD interfaces per branch.
void flops_2d_step(const double *U, int nx, int ny, int lm, double ch, long out[8], long branches[9]) {
  const size_t n = (size_t)nx * ny;
  auto at = [&](int i, int j) { return (size_t)((j % ny + ny) % ny) * nx + (size_t)((i % nx + nx) % nx); };
  cnt::C = cnt::Counts();
  cnt::Counts in_hlld;
  memset(branches, 0, 9 * sizeof(long));
  std::vector<Cnt> u(NV * n), w(NV * n), F(NV * n), G(NV * n);
  for (size_t k = 0; k < NV * n; k++) u[k] = Cnt(U[k]);
  const double dtx = 1e-3;
  for (int stage = 0; stage < 2; stage++) {
    for (size_t c = 0; c < n; c++) {  // u2v, every cell
      Cnt a[NV], b[NV];
      for (int m = 0; m < NV; m++) a[m] = u[m * n + c];
      prim_fwd<Cnt>(a, b);
      for (int m = 0; m < NV; m++) w[m * n + c] = b[m];
    }
    for (int dir = 0; dir < 2; dir++) {
      const int di = dir ? 0 : 1, dj = dir ? 1 : 0;
      std::vector<Cnt> gRv(NV * n), gLv(NV * n);
      for (int j = 0; j < ny; j++)  // one limiter call per cell and variable
        for (int i = 0; i < nx; i++)
          for (int m = 0; m < NV; m++) {
            const Cnt *a = &w[m * n];
            lim_fwd<Cnt>(lm, a[at(i - di, j - dj)], a[at(i, j)], a[at(i + di, j + dj)], gRv[m * n + at(i, j)],
                         gLv[m * n + at(i, j)]);
            if (lim_has_sign_sum(lm)) cnt::C.add++;
          }
      std::vector<Cnt> &FX = dir ? G : F;
      for (int j = 0; j < ny; j++)  // one HLLD per interface (i+1/2 or j+1/2)
        for (int i = 0; i < nx; i++) {
          Cnt L[NV], R[NV], f[NV];
          for (int m = 0; m < NV; m++) {
            L[m] = gLv[m * n + at(i, j)];
            R[m] = gRv[m * n + at(i + di, j + dj)];
          }
          int br = 0;
          const cnt::Counts before = cnt::C;
          if (dir == 0) hlld_fwd<0, Cnt>(L, R, ch, f, &br);
          else hlld_fwd<1, Cnt>(L, R, ch, f, &br);
          if (br == 3 || br == 4) cnt::C.add += 2;
          in_hlld.add += cnt::C.add - before.add;
          in_hlld.mul += cnt::C.mul - before.mul;
          in_hlld.div += cnt::C.div - before.div;
          in_hlld.sqrt += cnt::C.sqrt - before.sqrt;
          if (br >= 1 && br <= 8) branches[br]++;
          for (int m = 0; m < NV; m++) FX[m * n + at(i, j)] = f[m];
        }
    }
    for (int j = 0; j < ny; j++)  // the RK combine, rk.f90:30-31 (stage 1) and :60-62 (stage 2)
      for (int i = 0; i < nx; i++)
        for (int m = 0; m < NV; m++) {
          const size_t c = m * n + at(i, j);
          const Cnt d = ((F[m * n + at(i - 1, j)] - F[c]) + G[m * n + at(i, j - 1)]) - G[c];
          const Cnt r = stage == 0 ? u[c] + dtx * d : 0.5 * ((u[c] + u[c]) + dtx * d);
          (void)r;
        }
  }
  for (size_t c = 0; c < n; c++) {  // glm_ss2 twice (psi *= f) and the CFL speed, once per step
    Cnt psi = u[8 * n + c];
    psi = 0.99 * psi;
    psi = 0.99 * psi;
    Cnt a[NV], b[NV];
    for (int m = 0; m < NV; m++) a[m] = u[m * n + c];
    const cnt::Counts keep = cnt::C;  // set_dt reads V from the step's own u2v: not counted twice
    prim_fwd<Cnt>(a, b);
    cnt::C = keep;
    (void)cfl_speed(a, b);
  }
  out[0] = cnt::C.add, out[1] = cnt::C.mul, out[2] = cnt::C.div, out[3] = cnt::C.sqrt;
  out[4] = in_hlld.add, out[5] = in_hlld.mul, out[6] = in_hlld.div, out[7] = in_hlld.sqrt;
}

}  // extern "C"
