/*
 * omhd_oracle.h -- CPU restatement of OpenMHD's step (TEST INFRASTRUCTURE).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 */
#ifndef OMHD_ORACLE_H
#define OMHD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  OR_OK = 0,
  OR_ERR_PRESSURE = 1,
  OR_ERR_DENSITY = 2,
  OR_ERR_DT = 3,
  OR_ERR_DIR = 4,
  OR_ERR_TYPE = 5
};

/* HLLD branch taken at an interface (flux_solver.f90:413-572) */
enum {
  OR_BR_FL = 1,
  OR_BR_FR = 2,
  OR_BR_HLLC_L = 3,
  OR_BR_HLLC_R = 4,
  OR_BR_LS = 5,
  OR_BR_RS = 6,
  OR_BR_LSS = 7,
  OR_BR_RSS = 8
};

typedef struct or_problem {
  int ix, jx, kx;      /* extents incl. one ghost layer per side; kx == 1 for 2D */
  double dx, cfl;
  int lm_type;         /* 0..4 */
  int flux_type;       /* 3 (HLLD) */
  int time_type;       /* 0 TVD-RK2, 1 RK2 */
  int bctype[6];       /* per face (x-,x+,y-,y+,z-,z+): 0 periodic, 1 mirror */
  double u_sgn[54];    /* 6 faces x 9 slots: sign applied to U ghost copies */
  double face_sgn[54]; /* 6 faces x 9 slots: sign for the VL/VR face patch */
  double dt_fixed;     /* > 0: replay mode, use dt_fixed/ch_fixed instead of set_dt */
  double ch_fixed;
  const double *eta;   /* NULL, or 3 face arrays (ix*jx each) of the resistivity */
  double eta_max;
  int res_form;        /* 0: the 3D form (eta per face, dt limit 0.5 cfl dx^2/eta_max);
                          1: the 2D CPU form of 2D_reconnection: eta holds 2 face
                          arrays of eta/dx, dt limit 0.5 cfl rm1 dx^2;
                          2: the 3D form with the dt limit 0.5 cfl rm1 dx^2
                          (3D_reconnection_gpu/flux_resistive.cuf:2-9) */
  double rm1;          /* res_form 1, 2: the Rm1 of set_dt_resistive */
} or_problem;

int or_u2v(const double *U, double *V, int ix, int jx, int kx, long *bad);
int or_limiter(const double *wk, double *wL, double *wR, int ix, int jx, int kx, int dir, int type);
void or_hlld(const double *L, const double *R, double ch, int d, double *F, int *branch);
/* flux types 0 LLF, 1 HLL, 2 HLLC, 3 HLLD, 4 LHLLD at one interface */
void or_riemann(int type, const double *L, const double *R, double ch, int d, double *F);
int or_flux_solver(double *F, const double *VL, const double *VR, double ch, int ix, int jx, int kx, int dir, int type,
                   int *branches);
void or_glm_ss2(double *U, double ch, double dt, int ix, int jx, int kx);
void or_rk(int mode, double *wk, double *wk1, const double *wF, const double *wG, const double *wH, double dt, double dx,
           int ix, int jx, int kx);
int or_set_dt(const double *U, const double *V, double *vmax, double *dt, double dx, double cfl, int ix, int jx, int kx,
              int ndim, long *bad);
void or_bc_u(double *U, int ix, int jx, int kx, const int *bctype, const double *sgn);
void or_bc_face(double *VL, double *VR, int ix, int jx, int kx, int dir, int bctype, const double *sgn_lo,
                const double *sgn_hi);
void or_flux_resistive(double *F, const double *U, const double *eta_face, double dx, int ix, int jx, int kx, int dir);
void or_flux_resistive2d(double *F, const double *U, const double *eta_dx, int ix, int jx, int dir);
double or_set_dt_resistive2d(double dt, double dx, double cfl, double rm1);
double or_eta_faces(const double *x, const double *y, double dx, int ix, int jx, double eta0, double eta1, double w,
                    double *eta_out);
double or_set_dt_resistive(double dt, double dx, double cfl, double eta_max);
int or_step(const or_problem *p, double *U, double *dt, double *ch, long *bad);
int or_run(const or_problem *p, double *U, int nsteps, double *dt_trace, double *ch_trace, long *bad);
/* 1D shock tube (1D_basic/main.f90): U(ix,1,9), outflow ghosts, ch = 0 */
int or_run1d(double *U, int ix, double dx, double cfl, int lm_type, int flux_type, int time_type, double tend,
             int max_steps, double *t_out, int *nsteps_out, double *dt_trace, long *bad);

#ifdef __cplusplus
}
#endif
#endif
