// adjcheck.cpp -- host build of the adjoint building blocks
// (openmhd_amd/csrc/omhd_adjoint.h) for the CPU tests.  Test infrastructure:
// it checks the hand-written reverse forms against exact forward-mode
// tangents (dual numbers) and gives a CPU reverse step to compare the GPU
// adjoint kernels with.  Grids are periodic 2D, interior only, SoA (9, ny, nx)
// with x fastest.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../openmhd_amd/csrc/omhd_adjoint.h"

namespace dual {
struct Dual {
  double v, d;
  Dual() : v(0), d(0) {}
  Dual(double a) : v(a), d(0) {}  // NOLINT: implicit, constants enter as duals
  Dual(double a, double b) : v(a), d(b) {}
};
inline Dual operator+(const Dual &a, const Dual &b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(const Dual &a, const Dual &b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator-(const Dual &a) { return Dual(-a.v, -a.d); }
inline Dual operator*(const Dual &a, const Dual &b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(const Dual &a, const Dual &b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
inline Dual operator+(double a, const Dual &b) { return Dual(a + b.v, b.d); }
inline Dual operator+(const Dual &a, double b) { return Dual(a.v + b, a.d); }
inline Dual operator-(double a, const Dual &b) { return Dual(a - b.v, -b.d); }
inline Dual operator-(const Dual &a, double b) { return Dual(a.v - b, a.d); }
inline Dual operator*(double a, const Dual &b) { return Dual(a * b.v, a * b.d); }
inline Dual operator*(const Dual &a, double b) { return Dual(a.v * b, a.d * b); }
inline Dual operator/(double a, const Dual &b) { return Dual(a / b.v, -a * b.d / (b.v * b.v)); }
inline Dual operator/(const Dual &a, double b) { return Dual(a.v / b, a.d / b); }
inline double val(const Dual &a) { return a.v; }
inline Dual tsqrt(const Dual &a) {
  const double s = sqrt(a.v);
  return Dual(s, s > 0 ? a.d / (2 * s) : 0.0);
}
inline Dual tabs(const Dual &a) { return a.v < 0 ? -a : a; }
inline Dual tmin(const Dual &a, const Dual &b) { return a.v < b.v ? a : b; }
inline Dual tmax(const Dual &a, const Dual &b) { return a.v > b.v ? a : b; }
}  // namespace dual

using dual::Dual;
using namespace omhd::adj;

namespace {

struct Grid {
  int nx, ny, lm;
  size_t n() const { return (size_t)nx * ny; }
  size_t at(int i, int j) const {
    i = (i % nx + nx) % nx;
    j = (j % ny + ny) % ny;
    return (size_t)j * nx + i;
  }
};

template <int D, class T>
void hlld_dispatch(const T L[NV], const T R[NV], double ch, T F[NV]) {
  hlld_fwd<D, T>(L, R, ch, F);
}

// flux difference D(U) = F(i-1) - F(i) + G(j-1) - G(j), evaluated left to right
// as rk.f90 does
template <class T>
void flux_diff(const Grid &g, const T *U, double ch, T *Dout) {
  const size_t n = g.n();
  std::vector<T> w(NV * n), Fx(NV * n), Fy(NV * n);
  for (size_t c = 0; c < n; c++) {
    T u[NV], p[NV];
    for (int m = 0; m < NV; m++) u[m] = U[m * n + c];
    prim_fwd<T>(u, p);
    for (int m = 0; m < NV; m++) w[m * n + c] = p[m];
  }
  for (int dir = 0; dir < 2; dir++) {
    std::vector<T> &F = dir ? Fy : Fx;
    for (int j = 0; j < g.ny; j++)
      for (int i = 0; i < g.nx; i++) {
        const int di = dir ? 0 : 1, dj = dir ? 1 : 0;
        T VL[NV], VR[NV], f[NV];
        for (int m = 0; m < NV; m++) {
          const T *a = &w[m * n];
          T gR, gL, hR, hL;
          lim_fwd<T>(g.lm, a[g.at(i - di, j - dj)], a[g.at(i, j)], a[g.at(i + di, j + dj)], gR, gL);
          lim_fwd<T>(g.lm, a[g.at(i, j)], a[g.at(i + di, j + dj)], a[g.at(i + 2 * di, j + 2 * dj)], hR, hL);
          VL[m] = gL;
          VR[m] = hR;
        }
        if (dir == 0) hlld_dispatch<0, T>(VL, VR, ch, f);
        else hlld_dispatch<1, T>(VL, VR, ch, f);
        for (int m = 0; m < NV; m++) F[m * n + g.at(i, j)] = f[m];
      }
  }
  for (int m = 0; m < NV; m++)
    for (int j = 0; j < g.ny; j++)
      for (int i = 0; i < g.nx; i++) {
        const size_t c = g.at(i, j);
        Dout[m * n + c] = Fx[m * n + g.at(i - 1, j)] - Fx[m * n + c] + Fy[m * n + g.at(i, j - 1)] - Fy[m * n + c];
      }
}

// one step: glm_ss2, stage 1, stage 2, glm_ss2 (2D_basic/main.f90:80-185 order)
template <class T>
void step(const Grid &g, int tt, double dt, double dx, double ch, double glm_f, const T *U, T *Uout, T *Ustar = nullptr,
          T *Ug = nullptr) {
  const size_t n = g.n();
  std::vector<T> u0(U, U + NV * n), u1(NV * n), d(NV * n);
  for (size_t c = 0; c < n; c++) u0[8 * n + c] = glm_f * u0[8 * n + c];
  flux_diff<T>(g, u0.data(), ch, d.data());
  const double dtx1 = tt == 0 ? dt / dx : 0.5 * dt / dx;
  for (size_t k = 0; k < NV * n; k++) u1[k] = u0[k] + dtx1 * d[k];
  flux_diff<T>(g, u1.data(), ch, d.data());
  const double dtx = dt / dx;
  for (size_t k = 0; k < NV * n; k++) Uout[k] = tt == 0 ? 0.5 * (u0[k] + u1[k] + dtx * d[k]) : u0[k] + dtx * d[k];
  for (size_t c = 0; c < n; c++) Uout[8 * n + c] = glm_f * Uout[8 * n + c];
  if (Ustar) std::copy(u1.begin(), u1.end(), Ustar);
  if (Ug) std::copy(u0.begin(), u0.end(), Ug);
}

// nu += (dD/dW)^T mu at state W (the structure of the GPU kernels: primitives,
// one reverse HLLD per face giving the adjoints of the two face states, then
// the limiter and u2v reverses gathered per cell)
void flux_diff_vjp(const Grid &g, const double *W, double ch, const double *mu, double *nu) {
  const size_t n = g.n();
  std::vector<double> w(NV * n), bw(NV * n, 0.0), AL(NV * n), AR(NV * n);
  for (size_t c = 0; c < n; c++) {
    double u[NV], p[NV];
    for (int m = 0; m < NV; m++) u[m] = W[m * n + c];
    prim_fwd<double>(u, p);
    for (int m = 0; m < NV; m++) w[m * n + c] = p[m];
  }
  for (int dir = 0; dir < 2; dir++) {
    const int di = dir ? 0 : 1, dj = dir ? 1 : 0;
    for (int j = 0; j < g.ny; j++)
      for (int i = 0; i < g.nx; i++) {
        double VL[NV], VR[NV], bF[NV], bL[NV] = {0}, bR[NV] = {0};
        for (int m = 0; m < NV; m++) {
          const double *a = &w[m * n];
          double gR, gL, hR, hL;
          lim_fwd<double>(g.lm, a[g.at(i - di, j - dj)], a[g.at(i, j)], a[g.at(i + di, j + dj)], gR, gL);
          lim_fwd<double>(g.lm, a[g.at(i, j)], a[g.at(i + di, j + dj)], a[g.at(i + 2 * di, j + 2 * dj)], hR, hL);
          VL[m] = gL;
          VR[m] = hR;
          // F of this face enters D(c) with -1 and D(c+e) with +1
          bF[m] = mu[m * n + g.at(i + di, j + dj)] - mu[m * n + g.at(i, j)];
        }
        if (dir == 0) hlld_vjp<0>(VL, VR, ch, bF, bL, bR);
        else hlld_vjp<1>(VL, VR, ch, bF, bL, bR);
        for (int m = 0; m < NV; m++) {
          AL[m * n + g.at(i, j)] = bL[m];
          AR[m * n + g.at(i + di, j + dj)] = bR[m];
        }
      }
    for (int j = 0; j < g.ny; j++)
      for (int i = 0; i < g.nx; i++)
        for (int m = 0; m < NV; m++) {
          const double *a = &w[m * n];
          const size_t c = g.at(i, j);
          double cR[3], cL[3];
          lim_coef(g.lm, a[g.at(i - di, j - dj)], a[c], a[g.at(i + di, j + dj)], cR, cL);
          const double l = AL[m * n + c], r = AR[m * n + c];
          bw[m * n + c] += l + r;  // gL = w + gradL, gR = w - gradR
          bw[m * n + g.at(i - di, j - dj)] += cL[0] * l - cR[0] * r;
          bw[m * n + c] += cL[1] * l - cR[1] * r;
          bw[m * n + g.at(i + di, j + dj)] += cL[2] * l - cR[2] * r;
        }
  }
  for (size_t c = 0; c < n; c++) {
    double p[NV], b[NV], o[NV] = {0};
    for (int m = 0; m < NV; m++) {
      p[m] = w[m * n + c];
      b[m] = bw[m * n + c];
    }
    prim_vjp(p, b, o);
    for (int m = 0; m < NV; m++) nu[m * n + c] += o[m];
  }
}

void step_vjp(const Grid &g, int tt, double dt, double dx, double ch, double glm_f, const double *U, const double *lam,
              double *lam_in) {
  const size_t n = g.n();
  std::vector<double> uo(NV * n), us(NV * n), ug(NV * n);
  step<double>(g, tt, dt, dx, ch, glm_f, U, uo.data(), us.data(), ug.data());
  const double dtx = dt / dx;
  std::vector<double> lss(lam, lam + NV * n), mu(NV * n), ls(NV * n, 0.0), lg(NV * n, 0.0);
  for (size_t c = 0; c < n; c++) lss[8 * n + c] *= glm_f;
  // stage 2: TVD U** = 0.5 (Ug + U* + dtx D(U*)); STD U** = Ug + dtx D(U*)
  const double c0 = tt == 0 ? 0.5 : 1.0, c1 = tt == 0 ? 0.5 : 0.0, a2 = tt == 0 ? 0.5 * dtx : dtx;
  for (size_t k = 0; k < NV * n; k++) {
    mu[k] = a2 * lss[k];
    ls[k] = c1 * lss[k];
    lg[k] = c0 * lss[k];
  }
  flux_diff_vjp(g, us.data(), ch, mu.data(), ls.data());
  // stage 1: U* = Ug + a1 D(Ug)
  const double a1 = tt == 0 ? dtx : 0.5 * dtx;
  for (size_t k = 0; k < NV * n; k++) {
    mu[k] = a1 * ls[k];
    lg[k] += ls[k];
  }
  flux_diff_vjp(g, ug.data(), ch, mu.data(), lg.data());
  for (size_t k = 0; k < NV * n; k++) lam_in[k] = lg[k];
  for (size_t c = 0; c < n; c++) lam_in[8 * n + c] *= glm_f;
}

}  // namespace

extern "C" {

int adjc_hlld_fwd(const double *L, const double *R, double ch, int d, double *F, int *br) {
  if (d == 0) hlld_fwd<0, double>(L, R, ch, F, br);
  else if (d == 1) hlld_fwd<1, double>(L, R, ch, F, br);
  else hlld_fwd<2, double>(L, R, ch, F, br);
  return 0;
}

int adjc_hlld_jvp(const double *L, const double *R, const double *dL, const double *dR, double ch, int d, double *dF) {
  Dual l[NV], r[NV], f[NV];
  for (int m = 0; m < NV; m++) {
    l[m] = Dual(L[m], dL[m]);
    r[m] = Dual(R[m], dR[m]);
  }
  if (d == 0) hlld_fwd<0, Dual>(l, r, ch, f);
  else if (d == 1) hlld_fwd<1, Dual>(l, r, ch, f);
  else hlld_fwd<2, Dual>(l, r, ch, f);
  for (int m = 0; m < NV; m++) dF[m] = f[m].d;
  return 0;
}

int adjc_hlld_vjp(const double *L, const double *R, double ch, int d, const double *bF, double *bL, double *bR) {
  for (int m = 0; m < NV; m++) bL[m] = bR[m] = 0.0;
  if (d == 0) hlld_vjp<0>(L, R, ch, bF, bL, bR);
  else if (d == 1) hlld_vjp<1>(L, R, ch, bF, bL, bR);
  else hlld_vjp<2>(L, R, ch, bF, bL, bR);
  return 0;
}

// limiter on a line of n values (no wrap): gR[i], gL[i] for i = 1..n-2
int adjc_lim_fwd(int type, const double *w, int n, double *gR, double *gL) {
  for (int i = 1; i < n - 1; i++) lim_fwd<double>(type, w[i - 1], w[i], w[i + 1], gR[i], gL[i]);
  return 0;
}
int adjc_lim_jvp(int type, const double *w, const double *dw, int n, double *dgR, double *dgL) {
  for (int i = 1; i < n - 1; i++) {
    Dual a(w[i - 1], dw[i - 1]), b(w[i], dw[i]), c(w[i + 1], dw[i + 1]), r, l;
    lim_fwd<Dual>(type, a, b, c, r, l);
    dgR[i] = r.d;
    dgL[i] = l.d;
  }
  return 0;
}
// bw = sum_i (dgR[i]/dw)^T bgR[i] + (dgL[i]/dw)^T bgL[i]
int adjc_lim_vjp(int type, const double *w, int n, const double *bgR, const double *bgL, double *bw) {
  for (int i = 0; i < n; i++) bw[i] = 0.0;
  for (int i = 1; i < n - 1; i++) {
    double cR[3], cL[3];
    lim_coef(type, w[i - 1], w[i], w[i + 1], cR, cL);
    bw[i] += bgR[i] + bgL[i];
    for (int o = 0; o < 3; o++) bw[i - 1 + o] += cL[o] * bgL[i] - cR[o] * bgR[i];
  }
  return 0;
}

int adjc_prim_jvp(const double *u, const double *du, double *w, double *dw) {
  Dual a[NV], b[NV];
  for (int m = 0; m < NV; m++) a[m] = Dual(u[m], du[m]);
  prim_fwd<Dual>(a, b);
  for (int m = 0; m < NV; m++) {
    w[m] = b[m].v;
    dw[m] = b[m].d;
  }
  return 0;
}
int adjc_prim_vjp(const double *u, const double *bw, double *bu) {
  double w[NV];
  prim_fwd<double>(u, w);
  for (int m = 0; m < NV; m++) bu[m] = 0.0;
  prim_vjp(w, bw, bu);
  return 0;
}

// whole steps on a periodic nx x ny grid, (dt, ch) given per step
int adjc_run_fwd(int nx, int ny, int lm, int tt, double dx, int nsteps, const double *dts, const double *chs,
                 const double *U0, double *UT) {
  Grid g{nx, ny, lm};
  std::vector<double> a(U0, U0 + NV * g.n()), b(NV * g.n());
  for (int s = 0; s < nsteps; s++) {
    const double f = exp(-0.5 * dts[s] * chs[s] / 0.2);
    step<double>(g, tt, dts[s], dx, chs[s], f, a.data(), b.data());
    a.swap(b);
  }
  memcpy(UT, a.data(), sizeof(double) * NV * g.n());
  return 0;
}

int adjc_run_jvp(int nx, int ny, int lm, int tt, double dx, int nsteps, const double *dts, const double *chs,
                 const double *U0, const double *dU0, double *dUT) {
  Grid g{nx, ny, lm};
  const size_t N = NV * g.n();
  std::vector<Dual> a(N), b(N);
  for (size_t k = 0; k < N; k++) a[k] = Dual(U0[k], dU0[k]);
  for (int s = 0; s < nsteps; s++) {
    const double f = exp(-0.5 * dts[s] * chs[s] / 0.2);
    step<Dual>(g, tt, dts[s], dx, chs[s], f, a.data(), b.data());
    a.swap(b);
  }
  for (size_t k = 0; k < N; k++) dUT[k] = a[k].d;
  return 0;
}

int adjc_run_vjp(int nx, int ny, int lm, int tt, double dx, int nsteps, const double *dts, const double *chs,
                 const double *U0, const double *lamT, double *lam0) {
  Grid g{nx, ny, lm};
  const size_t N = NV * g.n();
  std::vector<std::vector<double>> states(nsteps + 1, std::vector<double>(N));
  states[0].assign(U0, U0 + N);
  for (int s = 0; s < nsteps; s++) {
    const double f = exp(-0.5 * dts[s] * chs[s] / 0.2);
    step<double>(g, tt, dts[s], dx, chs[s], f, states[s].data(), states[s + 1].data());
  }
  std::vector<double> lam(lamT, lamT + N), prev(N);
  for (int s = nsteps - 1; s >= 0; s--) {
    const double f = exp(-0.5 * dts[s] * chs[s] / 0.2);
    step_vjp(g, tt, dts[s], dx, chs[s], f, states[s].data(), lam.data(), prev.data());
    lam.swap(prev);
  }
  memcpy(lam0, lam.data(), sizeof(double) * N);
  return 0;
}

}  // extern "C"
