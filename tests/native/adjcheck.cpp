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

// ---------------------------------------------------------------------------
// General grids: 2D or 3D, per-axis periodic or wall boundaries (mirror with a
// sign per slot, wall face states are mirrors of the inside face state), and
// the resistive flux of the 3D reconnection step.  The forward is written
// face by face; the reverse is a plain scatter (no gather), so it checks the
// GPU kernels' gather formulation independently.
// ---------------------------------------------------------------------------
namespace {

struct G3 {
  int nx, ny, nz, nd, lm;
  int bc[3];
  double slo[3][NV], shi[3][NV];
  double dx;
  size_t n() const { return (size_t)nx * ny * nz; }
  int ext(int a) const { return a == 0 ? nx : (a == 1 ? ny : nz); }
  size_t off(int i, int j, int k) const { return ((size_t)k * ny + j) * nx + i; }
  // interior index of coordinate c on axis a, and the sign of slot m there
  int map(int a, int c, int m, double &s) const {
    const int e = ext(a);
    if (c < 0) {
      if (bc[a] == 0) return c + e;
      s *= slo[a][m];
      return -1 - c;
    }
    if (c >= e) {
      if (bc[a] == 0) return c - e;
      s *= shi[a][m];
      return 2 * e - 1 - c;
    }
    return c;
  }
  size_t at(int m, int i, int j, int k, double &s) const {
    s = 1.0;
    i = map(0, i, m, s);
    j = map(1, j, m, s);
    if (nd == 3) k = map(2, k, m, s);
    return (size_t)m * n() + off(i, j, k);
  }
};

template <class T>
T get(const G3 &g, const T *w, int m, int i, int j, int k) {
  double s;
  const size_t o = g.at(m, i, j, k, s);
  return s < 0 ? -w[o] : w[o];
}

template <class T>
void flipv(const G3 &g, const double *sg, const T *a, T *b) {
  for (int m = 0; m < NV; m++) b[m] = sg[m] < 0 ? -a[m] : a[m];
}

template <int D, class T>
void hlld3(const T L[NV], const T R[NV], double ch, T F[NV]) {
  hlld_fwd<D, T>(L, R, ch, F);
}

// the face states of the face with left cell (i,j,k) along axis a (the left
// cell may be -1 and the right one e at a wall)
template <class T>
void face_states(const G3 &g, const T *w, int a, int i, int j, int k, T VL[NV], T VR[NV]) {
  const int di = a == 0, dj = a == 1, dk = a == 2;
  const int c = a == 0 ? i : (a == 1 ? j : k), e = g.ext(a);
  auto lim_at = [&](int ii, int jj, int kk, int m, T &gR, T &gL) {
    lim_fwd<T>(g.lm, get(g, w, m, ii - di, jj - dj, kk - dk), get(g, w, m, ii, jj, kk),
               get(g, w, m, ii + di, jj + dj, kk + dk), gR, gL);
  };
  if (g.bc[a] != 0 && c == -1) {
    for (int m = 0; m < NV; m++) {
      T gR, gL;
      lim_at(i + di, j + dj, k + dk, m, gR, gL);
      VR[m] = gR;
    }
    flipv(g, g.slo[a], VR, VL);
  } else if (g.bc[a] != 0 && c == e - 1) {
    for (int m = 0; m < NV; m++) {
      T gR, gL;
      lim_at(i, j, k, m, gR, gL);
      VL[m] = gL;
    }
    flipv(g, g.shi[a], VL, VR);
  } else {
    for (int m = 0; m < NV; m++) {
      T gR, gL, hR, hL;
      lim_at(i, j, k, m, gR, gL);
      lim_at(i + di, j + dj, k + dk, m, hR, hL);
      VL[m] = gL;
      VR[m] = hR;
    }
  }
}

// resistive flux at the face with left cell (i,j,k) along axis a
// (3D_reconnection_gpu/flux_resistive.cuf:86-135), added to F
template <class T>
void resistive(const G3 &g, const T *w, const T *eta, int a, int i, int j, int k, T F[NV]) {
  const int ix = g.nx + 2, jx = g.ny + 2;
  auto B = [&](int m, int ii, int jj, int kk) { return get(g, w, m, ii, jj, kk); };
  const T eta_dx = eta[(size_t)a * ix * jx + (size_t)(i + 1) + (size_t)ix * (j + 1)] / g.dx;
  if (a == 0) {
    const T JyS = -(B(7, i + 1, j, k) - B(7, i, j, k)) +
                  0.25 * (B(5, i, j, k + 1) + B(5, i + 1, j, k + 1) - B(5, i, j, k - 1) - B(5, i + 1, j, k - 1));
    const T JzS = (B(6, i + 1, j, k) - B(6, i, j, k)) -
                  0.25 * (B(5, i, j + 1, k) + B(5, i + 1, j + 1, k) - B(5, i, j - 1, k) - B(5, i + 1, j - 1, k));
    F[3] = F[3] + 0.5 * eta_dx * (JyS * (B(7, i, j, k) + B(7, i + 1, j, k)) - JzS * (B(6, i, j, k) + B(6, i + 1, j, k)));
    F[6] = F[6] - eta_dx * JzS;
    F[7] = F[7] + eta_dx * JyS;
  } else if (a == 1) {
    const T JxS = (B(7, i, j + 1, k) - B(7, i, j, k)) -
                  0.25 * (B(6, i, j, k + 1) + B(6, i, j + 1, k + 1) - B(6, i, j, k - 1) - B(6, i, j + 1, k - 1));
    const T JzS = -(B(5, i, j + 1, k) - B(5, i, j, k)) +
                  0.25 * (B(6, i + 1, j, k) + B(6, i + 1, j + 1, k) - B(6, i - 1, j, k) - B(6, i - 1, j + 1, k));
    F[3] = F[3] + 0.5 * eta_dx * (JzS * (B(5, i, j, k) + B(5, i, j + 1, k)) - JxS * (B(7, i, j, k) + B(7, i, j + 1, k)));
    F[5] = F[5] + eta_dx * JzS;
    F[7] = F[7] - eta_dx * JxS;
  } else {
    const T JyS = (B(5, i, j, k + 1) - B(5, i, j, k)) -
                  0.25 * (B(7, i + 1, j, k) + B(7, i + 1, j, k + 1) - B(7, i - 1, j, k) - B(7, i - 1, j, k + 1));
    const T JxS = -(B(6, i, j, k + 1) - B(6, i, j, k)) +
                  0.25 * (B(7, i, j + 1, k) + B(7, i, j + 1, k + 1) - B(7, i, j - 1, k) - B(7, i, j - 1, k + 1));
    F[3] = F[3] + 0.5 * eta_dx * (JxS * (B(6, i, j, k) + B(6, i, j, k + 1)) - JyS * (B(5, i, j, k) + B(5, i, j, k + 1)));
    F[5] = F[5] - eta_dx * JyS;
    F[6] = F[6] + eta_dx * JxS;
  }
}

template <class T>
void flux_diff3(const G3 &g, const T *U, double ch, const T *eta, T *Dout) {
  const size_t n = g.n();
  std::vector<T> w(NV * n);
  for (size_t c = 0; c < n; c++) {
    T u[NV], p[NV];
    for (int m = 0; m < NV; m++) u[m] = U[m * n + c];
    prim_fwd<T>(u, p);
    for (int m = 0; m < NV; m++) w[m * n + c] = p[m];
  }
  // face fluxes per axis, indexed by the left cell + 1 along that axis
  std::vector<T> Fa[3];
  for (int a = 0; a < g.nd; a++) {
    const int e = g.ext(a);
    const int ex = a == 0 ? e + 1 : g.nx, ey = a == 1 ? e + 1 : g.ny, ez = a == 2 ? e + 1 : g.nz;
    Fa[a].assign((size_t)NV * ex * ey * ez, T(0.0));
    for (int k = 0; k < ez; k++)
      for (int j = 0; j < ey; j++)
        for (int i = 0; i < ex; i++) {
          int ii = i - (a == 0), jj = j - (a == 1), kk = k - (a == 2);
          const int c = a == 0 ? ii : (a == 1 ? jj : kk);
          if (c == -1 && g.bc[a] == 0) continue;  // periodic: the face at -1/2 is the one at e-1/2
          T VL[NV], VR[NV], f[NV];
          face_states<T>(g, w.data(), a, ii, jj, kk, VL, VR);
          if (a == 0) hlld3<0, T>(VL, VR, ch, f);
          else if (a == 1) hlld3<1, T>(VL, VR, ch, f);
          else hlld3<2, T>(VL, VR, ch, f);
          if (eta) resistive<T>(g, w.data(), eta, a, ii, jj, kk, f);
          const size_t fo = ((size_t)k * ey + j) * ex + i, fn = (size_t)ex * ey * ez;
          for (int m = 0; m < NV; m++) Fa[a][m * fn + fo] = f[m];
        }
    if (g.bc[a] == 0) {  // copy face e-1/2 to -1/2
      for (int k = 0; k < ez; k++)
        for (int j = 0; j < ey; j++)
          for (int i = 0; i < ex; i++) {
            const int c = a == 0 ? i : (a == 1 ? j : k);
            if (c != 0) continue;
            const size_t fn = (size_t)ex * ey * ez, fo = ((size_t)k * ey + j) * ex + i;
            const size_t src = fo + (a == 0 ? e : (a == 1 ? (size_t)e * ex : (size_t)e * ex * ey));
            for (int m = 0; m < NV; m++) Fa[a][m * fn + fo] = Fa[a][m * fn + src];
          }
    }
  }
  for (int m = 0; m < NV; m++)
    for (int k = 0; k < g.nz; k++)
      for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) {
          auto F = [&](int a, int fi, int fj, int fk) {
            const int ex = a == 0 ? g.nx + 1 : g.nx, ey = a == 1 ? g.ny + 1 : g.ny, ez = a == 2 ? g.nz + 1 : g.nz;
            return Fa[a][(size_t)m * ex * ey * ez + ((size_t)fk * ey + fj) * ex + fi];
          };
          // face index = left cell + 1: F(i-1) is index i, F(i) is index i+1
          T d = F(0, i, j, k) - F(0, i + 1, j, k) + F(1, i, j, k) - F(1, i, j + 1, k);
          if (g.nd == 3) d = d + F(2, i, j, k) - F(2, i, j, k + 1);
          Dout[(size_t)m * g.n() + g.off(i, j, k)] = d;
        }
}

template <class T>
void step3(const G3 &g, int tt, double dt, double ch, double glm_f, const T *eta, const T *U, T *Uout,
           T *Ustar = nullptr, T *Ug = nullptr) {
  const size_t n = g.n();
  std::vector<T> u0(U, U + NV * n), u1(NV * n), d(NV * n);
  for (size_t c = 0; c < n; c++) u0[8 * n + c] = glm_f * u0[8 * n + c];
  flux_diff3<T>(g, u0.data(), ch, eta, d.data());
  const double dtx1 = tt == 0 ? dt / g.dx : 0.5 * dt / g.dx;
  for (size_t q = 0; q < NV * n; q++) u1[q] = u0[q] + dtx1 * d[q];
  flux_diff3<T>(g, u1.data(), ch, eta, d.data());
  const double dtx = dt / g.dx;
  for (size_t q = 0; q < NV * n; q++) Uout[q] = tt == 0 ? 0.5 * (u0[q] + u1[q] + dtx * d[q]) : u0[q] + dtx * d[q];
  for (size_t c = 0; c < n; c++) Uout[8 * n + c] = glm_f * Uout[8 * n + c];
  if (Ustar) std::copy(u1.begin(), u1.end(), Ustar);
  if (Ug) std::copy(u0.begin(), u0.end(), Ug);
}

// reverse of the limited face state gL (left = true) or gR at cell (i,j,k)
// along axis a: scatter b (one slot m) into bw
void lim_scatter(const G3 &g, const double *w, int a, int i, int j, int k, int m, bool left, double b, double *bw) {
  const int di = a == 0, dj = a == 1, dk = a == 2;
  double sm, s0, sp;
  const size_t om = g.at(m, i - di, j - dj, k - dk, sm), o0 = g.at(m, i, j, k, s0), op = g.at(m, i + di, j + dj, k + dk, sp);
  const double vm = sm < 0 ? -w[om] : w[om], v0 = s0 < 0 ? -w[o0] : w[o0], vp = sp < 0 ? -w[op] : w[op];
  double cR[3], cL[3];
  lim_coef(g.lm, vm, v0, vp, cR, cL);
  const double *c = left ? cL : cR;
  const double sg = left ? 1.0 : -1.0;  // gL = w0 + gradL, gR = w0 - gradR
  bw[o0] += s0 * b;
  bw[om] += sm * sg * c[0] * b;
  bw[o0] += s0 * sg * c[1] * b;
  bw[op] += sp * sg * c[2] * b;
}

void resistive_vjp(const G3 &g, const double *w, const double *eta, int a, int i, int j, int k, const double bF[NV],
                   double *bw, double *geta) {
  const int ix = g.nx + 2, jx = g.ny + 2;
  const size_t ei = (size_t)a * ix * jx + (size_t)(i + 1) + (size_t)ix * (j + 1);
  const double e = eta[ei] / g.dx;
  auto B = [&](int m, int ii, int jj, int kk) { return get(g, w, m, ii, jj, kk); };
  auto addB = [&](int m, int ii, int jj, int kk, double v) {
    double s;
    const size_t o = g.at(m, ii, jj, kk, s);
    bw[o] += s * v;
  };
  // generic form: F[en] += 0.5 e (J1 S1 - J2 S2), F[p] -= / += e J2 / J1
  int p1, p2, q1, q2;  // J1 pairs with slot p1 (sum S1 of slot p1), J2 with p2
  double J1, J2;
  int di = a == 0, dj = a == 1, dk = a == 2;
  // the first-order difference terms and the transverse 0.25 terms
  struct Term {
    int m, di, dj, dk;
    double c;
  };
  Term t1[6], t2[6];
  double sgn1, sgn2;  // F[p1] += sgn1 e J1 ; F[p2] += sgn2 e J2
  if (a == 0) {
    p1 = 7; p2 = 6; q1 = 7; q2 = 6;
    t1[0] = {7, 1, 0, 0, -1}; t1[1] = {7, 0, 0, 0, 1};
    t1[2] = {5, 0, 0, 1, 0.25}; t1[3] = {5, 1, 0, 1, 0.25}; t1[4] = {5, 0, 0, -1, -0.25}; t1[5] = {5, 1, 0, -1, -0.25};
    t2[0] = {6, 1, 0, 0, 1}; t2[1] = {6, 0, 0, 0, -1};
    t2[2] = {5, 0, 1, 0, -0.25}; t2[3] = {5, 1, 1, 0, -0.25}; t2[4] = {5, 0, -1, 0, 0.25}; t2[5] = {5, 1, -1, 0, 0.25};
    sgn1 = 1; sgn2 = -1;
  } else if (a == 1) {
    p1 = 5; p2 = 7; q1 = 5; q2 = 7;  // J1 = JzS (with bx), J2 = JxS (with bz)
    t1[0] = {5, 0, 1, 0, -1}; t1[1] = {5, 0, 0, 0, 1};
    t1[2] = {6, 1, 0, 0, 0.25}; t1[3] = {6, 1, 1, 0, 0.25}; t1[4] = {6, -1, 0, 0, -0.25}; t1[5] = {6, -1, 1, 0, -0.25};
    t2[0] = {7, 0, 1, 0, 1}; t2[1] = {7, 0, 0, 0, -1};
    t2[2] = {6, 0, 0, 1, -0.25}; t2[3] = {6, 0, 1, 1, -0.25}; t2[4] = {6, 0, 0, -1, 0.25}; t2[5] = {6, 0, 1, -1, 0.25};
    sgn1 = 1; sgn2 = -1;
  } else {
    p1 = 6; p2 = 5; q1 = 6; q2 = 5;  // J1 = JxS (with by), J2 = JyS (with bx)
    t1[0] = {6, 0, 0, 1, -1}; t1[1] = {6, 0, 0, 0, 1};
    t1[2] = {7, 0, 1, 0, 0.25}; t1[3] = {7, 0, 1, 1, 0.25}; t1[4] = {7, 0, -1, 0, -0.25}; t1[5] = {7, 0, -1, 1, -0.25};
    t2[0] = {5, 0, 0, 1, 1}; t2[1] = {5, 0, 0, 0, -1};
    t2[2] = {7, 1, 0, 0, -0.25}; t2[3] = {7, 1, 0, 1, -0.25}; t2[4] = {7, -1, 0, 0, 0.25}; t2[5] = {7, -1, 0, 1, 0.25};
    sgn1 = 1; sgn2 = -1;
  }
  J1 = 0;
  J2 = 0;
  for (int t = 0; t < 6; t++) {
    J1 += t1[t].c * B(t1[t].m, i + t1[t].di, j + t1[t].dj, k + t1[t].dk);
    J2 += t2[t].c * B(t2[t].m, i + t2[t].di, j + t2[t].dj, k + t2[t].dk);
  }
  const double S1 = B(q1, i, j, k) + B(q1, i + di, j + dj, k + dk);
  const double S2 = B(q2, i, j, k) + B(q2, i + di, j + dj, k + dk);
  // F[3] += 0.5 e (J1 S1 - J2 S2); F[p1] += sgn1 e J1; F[p2] += sgn2 e J2
  const double bJ1 = e * (0.5 * bF[3] * S1 + sgn1 * bF[p1]);
  const double bJ2 = e * (-0.5 * bF[3] * S2 + sgn2 * bF[p2]);
  for (int t = 0; t < 6; t++) {
    addB(t1[t].m, i + t1[t].di, j + t1[t].dj, k + t1[t].dk, t1[t].c * bJ1);
    addB(t2[t].m, i + t2[t].di, j + t2[t].dj, k + t2[t].dk, t2[t].c * bJ2);
  }
  const double bS1 = 0.5 * e * bF[3] * J1, bS2 = -0.5 * e * bF[3] * J2;
  addB(q1, i, j, k, bS1);
  addB(q1, i + di, j + dj, k + dk, bS1);
  addB(q2, i, j, k, bS2);
  addB(q2, i + di, j + dj, k + dk, bS2);
  if (geta) geta[ei] += (0.5 * bF[3] * (J1 * S1 - J2 * S2) + sgn1 * bF[p1] * J1 + sgn2 * bF[p2] * J2) / g.dx;
}

void flux_diff3_vjp(const G3 &g, const double *W, double ch, const double *eta, const double *mu, double *nu,
                    double *geta) {
  const size_t n = g.n();
  std::vector<double> w(NV * n), bw(NV * n, 0.0);
  for (size_t c = 0; c < n; c++) {
    double u[NV], p[NV];
    for (int m = 0; m < NV; m++) u[m] = W[m * n + c];
    prim_fwd<double>(u, p);
    for (int m = 0; m < NV; m++) w[m * n + c] = p[m];
  }
  for (int a = 0; a < g.nd; a++) {
    const int e = g.ext(a);
    const int di = a == 0, dj = a == 1, dk = a == 2;
    for (int k = 0; k < g.nz; k++)
      for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) {
          const int c = a == 0 ? i : (a == 1 ? j : k);
          // faces with left cell c (all c), and the low wall face with left cell -1
          for (int lo = 0; lo < 2; lo++) {
            if (lo && !(g.bc[a] != 0 && c == 0)) continue;
            const int ii = i - (lo ? di : 0), jj = j - (lo ? dj : 0), kk = k - (lo ? dk : 0);
            const int cl = c - lo;
            double VL[NV], VR[NV], bF[NV], bL[NV] = {0}, bR[NV] = {0};
            face_states<double>(g, w.data(), a, ii, jj, kk, VL, VR);
            const bool has_l = cl >= 0, has_r = cl + 1 < e || g.bc[a] == 0;
            for (int m = 0; m < NV; m++) {
              double mr = 0, ml = 0;
              if (has_r) {
                double s;
                mr = mu[g.at(m, ii + di, jj + dj, kk + dk, s)];  // periodic wrap only (s = 1)
              }
              if (has_l) ml = mu[(size_t)m * n + g.off(ii, jj, kk)];
              bF[m] = mr - ml;
            }
            if (a == 0) hlld_vjp<0>(VL, VR, ch, bF, bL, bR);
            else if (a == 1) hlld_vjp<1>(VL, VR, ch, bF, bL, bR);
            else hlld_vjp<2>(VL, VR, ch, bF, bL, bR);
            if (eta) resistive_vjp(g, w.data(), eta, a, ii, jj, kk, bF, bw.data(), geta);
            if (g.bc[a] != 0 && cl == -1) {  // VL = flip(VR): all onto gR of cell 0
              for (int m = 0; m < NV; m++)
                lim_scatter(g, w.data(), a, i, j, k, m, false, bR[m] + (g.slo[a][m] < 0 ? -bL[m] : bL[m]), bw.data());
            } else if (g.bc[a] != 0 && cl == e - 1) {  // VR = flip(VL): all onto gL of cell e-1
              for (int m = 0; m < NV; m++)
                lim_scatter(g, w.data(), a, ii, jj, kk, m, true, bL[m] + (g.shi[a][m] < 0 ? -bR[m] : bR[m]), bw.data());
            } else {
              for (int m = 0; m < NV; m++) {
                lim_scatter(g, w.data(), a, ii, jj, kk, m, true, bL[m], bw.data());
                lim_scatter(g, w.data(), a, ii + di, jj + dj, kk + dk, m, false, bR[m], bw.data());
              }
            }
          }
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

void step3_vjp(const G3 &g, int tt, double dt, double ch, double glm_f, const double *eta, const double *U,
               const double *lam, double *lam_in, double *geta) {
  const size_t n = g.n();
  std::vector<double> uo(NV * n), us(NV * n), ug(NV * n);
  step3<double>(g, tt, dt, ch, glm_f, eta, U, uo.data(), us.data(), ug.data());
  const double dtx = dt / g.dx;
  std::vector<double> lss(lam, lam + NV * n), mu(NV * n), ls(NV * n, 0.0), lg(NV * n, 0.0);
  for (size_t c = 0; c < n; c++) lss[8 * n + c] *= glm_f;
  const double c0 = tt == 0 ? 0.5 : 1.0, c1 = tt == 0 ? 0.5 : 0.0, a2 = tt == 0 ? 0.5 * dtx : dtx;
  for (size_t q = 0; q < NV * n; q++) {
    mu[q] = a2 * lss[q];
    ls[q] = c1 * lss[q];
    lg[q] = c0 * lss[q];
  }
  flux_diff3_vjp(g, us.data(), ch, eta, mu.data(), ls.data(), geta);
  const double a1 = tt == 0 ? dtx : 0.5 * dtx;
  for (size_t q = 0; q < NV * n; q++) {
    mu[q] = a1 * ls[q];
    lg[q] += ls[q];
  }
  flux_diff3_vjp(g, ug.data(), ch, eta, mu.data(), lg.data(), geta);
  for (size_t q = 0; q < NV * n; q++) lam_in[q] = lg[q];
  for (size_t c = 0; c < n; c++) lam_in[8 * n + c] *= glm_f;
}

G3 make_g3(int nx, int ny, int nz, int lm, double dx, const int *bc, const double *sg) {
  G3 g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz > 0 ? nz : 1;
  g.nd = nz > 0 ? 3 : 2;
  g.lm = lm;
  g.dx = dx;
  for (int a = 0; a < 3; a++) {
    g.bc[a] = bc[a];
    for (int m = 0; m < NV; m++) {
      g.slo[a][m] = sg[(2 * a) * NV + m];
      g.shi[a][m] = sg[(2 * a + 1) * NV + m];
    }
  }
  return g;
}

}  // namespace

extern "C" {

// nz = 0: 2D.  bc[3]: 0 periodic, 1 wall; sg[6*9]: signs per face (x-,x+,y-,y+,z-,z+)
// eta: null, or 3 face arrays of (nx+2)*(ny+2) (3D only)
int adjc3_run_fwd(int nx, int ny, int nz, int lm, int tt, double dx, const int *bc, const double *sg,
                  const double *eta, int nsteps, const double *dts, const double *chs, const double *U0, double *UT) {
  const G3 g = make_g3(nx, ny, nz, lm, dx, bc, sg);
  std::vector<double> a(U0, U0 + NV * g.n()), b(NV * g.n());
  for (int s = 0; s < nsteps; s++) {
    step3<double>(g, tt, dts[s], chs[s], exp(-0.5 * dts[s] * chs[s] / 0.2), eta, a.data(), b.data());
    a.swap(b);
  }
  memcpy(UT, a.data(), sizeof(double) * NV * g.n());
  return 0;
}

// tangent along (dU0, deta); deta may be null
int adjc3_run_jvp(int nx, int ny, int nz, int lm, int tt, double dx, const int *bc, const double *sg,
                  const double *eta, const double *deta, int nsteps, const double *dts, const double *chs,
                  const double *U0, const double *dU0, double *dUT) {
  const G3 g = make_g3(nx, ny, nz, lm, dx, bc, sg);
  const size_t N = NV * g.n();
  std::vector<Dual> a(N), b(N), e;
  for (size_t q = 0; q < N; q++) a[q] = Dual(U0[q], dU0[q]);
  if (eta) {
    const size_t ne = (size_t)3 * (nx + 2) * (ny + 2);
    e.resize(ne);
    for (size_t q = 0; q < ne; q++) e[q] = Dual(eta[q], deta ? deta[q] : 0.0);
  }
  for (int s = 0; s < nsteps; s++) {
    step3<Dual>(g, tt, dts[s], chs[s], exp(-0.5 * dts[s] * chs[s] / 0.2), eta ? e.data() : nullptr, a.data(),
                b.data());
    a.swap(b);
  }
  for (size_t q = 0; q < N; q++) dUT[q] = a[q].d;
  return 0;
}

// lam0 = (dU(T)/dU0)^T lamT; geta (optional, 3*(nx+2)*(ny+2), zeroed here) =
// (dU(T)/deta)^T lamT
int adjc3_run_vjp(int nx, int ny, int nz, int lm, int tt, double dx, const int *bc, const double *sg,
                  const double *eta, int nsteps, const double *dts, const double *chs, const double *U0,
                  const double *lamT, double *lam0, double *geta) {
  const G3 g = make_g3(nx, ny, nz, lm, dx, bc, sg);
  const size_t N = NV * g.n();
  if (geta) memset(geta, 0, sizeof(double) * 3 * (nx + 2) * (ny + 2));
  std::vector<std::vector<double>> st(nsteps + 1, std::vector<double>(N));
  st[0].assign(U0, U0 + N);
  for (int s = 0; s < nsteps; s++)
    step3<double>(g, tt, dts[s], chs[s], exp(-0.5 * dts[s] * chs[s] / 0.2), eta, st[s].data(), st[s + 1].data());
  std::vector<double> lam(lamT, lamT + N), prev(N);
  for (int s = nsteps - 1; s >= 0; s--) {
    step3_vjp(g, tt, dts[s], chs[s], exp(-0.5 * dts[s] * chs[s] / 0.2), eta, st[s].data(), lam.data(), prev.data(),
              eta ? geta : nullptr);
    lam.swap(prev);
  }
  memcpy(lam0, lam.data(), sizeof(double) * N);
  return 0;
}

}  // extern "C"
