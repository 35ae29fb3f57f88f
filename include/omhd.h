/*
 * omhd.h -- C ABI of libomhd, the MI355X (gfx950) implementation of
 * OpenMHD's finite-volume time step.
 *
 * Two ways in:
 *
 *  1. Drop-in entry points, one per reference subroutine, with the reference's
 *     argument lists (Fortran column-major arrays including the single ghost
 *     layer, extents ix, jx[, kx]).  A Fortran caller reaches them through the
 *     iso_c_binding shims in openmhd_amd/fortran/ (omhd_shims.f90, 2D;
 *     omhd_shims3d.f90, 3D; omhd_shims_kh.f90, omhd_shims_recon.f90 and
 *     omhd_shims3d_recon.f90, the problem-specific routines), which keep the
 *     reference names (u2v, limiter, flux_solver, ...), so a reference
 *     main.f90 links against them unchanged.  Array pointers may be host or
 *     device memory: host arrays are staged through device buffers per call,
 *     device arrays are used in place.  Every call returns when its results
 *     are in the caller's arrays.
 *
 *  2. The resident step: a handle owns the state in HBM (structure of arrays,
 *     interior cells only) and omhd_step() advances it by whole TVD-RK2 steps
 *     with the fused sweep kernels.  This is the fast path.
 *
 * Every function returns OMHD_OK (0) or an OMHD_ERR_* code; omhd_last_error()
 * gives the cell and value behind the last failure on the calling thread.
 * The reference reports these conditions by printing and calling STOP
 * (u2v.f90:34-43, set_dt.f90:55-65, limiter.f90:310-312).
 */
#ifndef OMHD_H
#define OMHD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define OMHD_OK 0
#define OMHD_ERR_PRESSURE 1   /* p <= 0 somewhere (u2v.f90:34-38)          */
#define OMHD_ERR_DENSITY 2    /* rho <= 0 somewhere (u2v.f90:39-43)        */
#define OMHD_ERR_DT 3         /* dt < 1e-7 or not finite (set_dt.f90:55-65) */
#define OMHD_ERR_DIR 4        /* dir not in 1..3 (or 3 on a 2D grid)       */
#define OMHD_ERR_TYPE 5       /* unknown limiter / flux / time type        */
#define OMHD_ERR_EXTENT 6     /* extents too small or inconsistent         */
#define OMHD_ERR_HIP 7        /* HIP runtime failure (message in last_error) */
#define OMHD_ERR_ARG 8        /* null pointer or bad argument              */

typedef struct omhd_error_info {
  int32_t code;
  int32_t i, j, k;   /* 1-based Fortran cell index incl. ghosts, when known */
  double value;      /* offending p, rho, dt, ...                          */
  char msg[160];
} omhd_error_info;

/* Last error recorded on the calling thread (cleared by a successful call). */
int omhd_last_error(omhd_error_info *out);

/* ---- boundary description ---------------------------------------------- */
#define OMHD_BC_PERIODIC 0
#define OMHD_BC_MIRROR 1      /* ghost = mirror of the first interior cell
                                 with the slots in `sign` negated; face state
                                 = mirror of the inside face state          */

/* ---- drop-in entry points (reference subroutine interfaces) ------------- */

/* u2v(U,V,ix,jx[,kx]) -- 2D_basic/u2v.f90:1, 3D_basic/u2v.f90:1.
 * U(ix,jx,kx,9) -> V(ix,jx,kx,4) over all cells; positivity check. */
int omhd_u2v(const double *U, double *V, int32_t ix, int32_t jx, int32_t kx);

/* limiter(wk,wL,wR,ix,jx[,kx],dir,type) -- 2D_basic/limiter.f90:1,
 * 3D_basic/limiter.f90:1.  One variable plane, one direction, types 0..4. */
int omhd_limiter(const double *wk, double *wL, double *wR, int32_t ix, int32_t jx, int32_t kx, int32_t dir,
                 int32_t type);

/* flux_solver(F,VL,VR,ch,ix,jx[,kx],dir,type) -- 2D_basic/flux_solver.f90:1,
 * 3D_basic/flux_solver.f90:1.  type 0 LLF, 1 HLL, 2 HLLC, 3 HLLD, 4 LHLLD
 * (:27).  F is zeroed first, as in the CPU build (flux_solver.f90:57). */
int omhd_flux_solver(double *F, const double *VL, const double *VR, double ch, int32_t ix, int32_t jx, int32_t kx,
                     int32_t dir, int32_t type);

/* glm_ss2(U,ch,dt,ix,jx[,kx]) -- 2D_basic/glm_ss2.f90:1, 3D_basic/glm_ss2.f90:1.
 * psi *= exp(-dt*ch/(2*0.2))
 * over the whole array (ghosts included); the factor is computed on the host. */
int omhd_glm_ss2(double *U, double ch, double dt, int32_t ix, int32_t jx, int32_t kx);

/* rk_tvd21 / rk_tvd22 / rk_std21 / rk_std22 -- 2D_basic/rk.f90:9,42,73,106,
 * 3D_basic/rk.f90:9,46,81,118.  mode 0..3 in that order.  wH == NULL for 2D.
 * For rk_std22 (mode 3) wk1 is not used. */
int omhd_rk(int32_t mode, double *wk, double *wk1, const double *wF, const double *wG, const double *wH, double dt,
            double dx, int32_t ix, int32_t jx, int32_t kx);

/* set_dt(U,V,vmax,dt,dx,cfl,ix,jx[,kx]) -- 2D_basic/set_dt.f90:1,
 * 3D_basic/set_dt.f90:1.  ndim selects the 2D (x,y) or 3D (x,y,z) speeds. */
int omhd_set_dt(const double *U, const double *V, double *vmax, double *dt, double dx, double cfl, int32_t ix,
                int32_t jx, int32_t kx, int32_t ndim);

/* omhd_set_dt without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as omhd_set_dt, except that U and V must be
 * device memory (OMHD_ERR_ARG otherwise) and the grid may hold at most
 * 67108864 interior cells (OMHD_ERR_EXTENT otherwise).  The reduction uses a
 * fixed 512 KiB scratch array that the library keeps on each GPU for the
 * life of the process; calls that overlap on one GPU share it and race. */
int omhd_set_dt_noalloc(const double *U, const double *V, double *vmax, double *dt, double dx, double cfl,
                        int32_t ix, int32_t jx, int32_t kx, int32_t ndim);

/* bc(U,ix,jx[,kx]) / bc_for_U(U,ix,jx[,kx]) -- 2D_basic/bc.f90:1,
 * 2D_KH/bc.f90:6, 2D_reconnection/bc.f90:6, 3D_basic/bc.f90:1,
 * 3D_reconnection_gpu/bc.cuf:7.  bc_type[6] per face (x-,x+,y-,y+,z-,z+),
 * sign[6*9] per face and slot (+1 / -1) for OMHD_BC_MIRROR faces. */
int omhd_bc_u(double *U, int32_t ix, int32_t jx, int32_t kx, const int32_t *bc_type, const double *sign);

/* Interface-state face patch along dir: the inline periodic patch of
 * 2D_basic/main.f90:101-102,119-120 and bc_for_F/G(/H) of 2D_KH/bc.f90:42,55,
 * 2D_reconnection/bc.f90:54,86 and 3D_reconnection_gpu/bc.cuf:103,147,191. */
int omhd_bc_face(double *VL, double *VR, int32_t ix, int32_t jx, int32_t kx, int32_t dir, int32_t bc_type,
                 const double *sign_lo, const double *sign_hi);

/* flux_resistive(F,U,x,y,dx,ix,jx,dir) -- 2D_reconnection/flux_resistive.f90:27,
 * the 2D CPU form.  The caller evaluates its resistivity profile: eta_dx is an
 * (ix,jx) column-major array of eta at the face centre times 1/dx, as :74 and
 * :97 form it (eta(x(i)+dx/2, y(j))/dx for dir 1, eta(x(i), y(j)+dx/2)/dx for
 * dir 2).  Adds the resistive terms to F(en), F(by) (dir 1) or F(bx) (dir 2),
 * and F(bz) on the faces i=1..ix-1, j=2..jx-1 (dir 1) or i=2..ix-1, j=1..jx-1
 * (dir 2); dir 3 adds nothing, as in the reference.  F(ix,jx,9), U(ix,jx,9)
 * and eta_dx in host or device memory.  (set_dt_resistive, :18, is one line
 * of scalar arithmetic and stays in the caller's language.) */
int omhd_flux_resistive(double *F, const double *U, const double *eta_dx, int32_t ix, int32_t jx, int32_t dir);

/* flux_resistive(F,U,x,y,z,dx,ix,jx,kx,dir) -- 3D_reconnection_gpu/
 * flux_resistive.cuf:12, the 3D form (a CUDA Fortran kernel in the reference;
 * here callable from the host).  eta holds the resistivity at the face centres
 * of direction dir (x(i)+dx/2 on x-faces, y(j)+dx/2 on y-faces, z(k)+dx/2 on
 * z-faces, :47-68), an (ix,jx) array when it does not vary in z (eta_nz = 1,
 * as the reference's profile) or (ix,jx,kx) (eta_nz = kx); each face divides
 * it by dx (:83).  Adds the resistive terms to F(en) and the two transverse B
 * slots on the faces of :47-68 (normal index 1..n-1, transverse 2..n-1).
 * F(ix,jx,kx,9), U(ix,jx,kx,9), eta: host or device memory. */
int omhd_flux_resistive3d(double *F, const double *U, const double *eta, int32_t eta_nz, double dx, int32_t ix,
                          int32_t jx, int32_t kx, int32_t dir);

/* ---- resident step ------------------------------------------------------ */

typedef struct omhd_config {
  int32_t ndim;         /* 2 or 3                                           */
  int32_t nx, ny, nz;   /* interior cells (the reference's ix-2, jx-2, kx-2) */
  double dx;            /* cell size, dx == dy (rk.f90:4-5)                 */
  double cfl;           /* CFL number (main.f90:15)                         */
  int32_t lm_type;      /* 0 flat, 1 minmod, 2 MC, 3 van Leer, 4 Koren      */
  int32_t flux_type;    /* 0 LLF, 1 HLL, 2 HLLC, 3 HLLD, 4 LHLLD (0-2, 4: ideal MHD only) */
  int32_t time_type;    /* 0 TVD-RK2, 1 RK2 (main.f90:20-21)                */
  int32_t bc_type[6];   /* OMHD_BC_* per face: x-, x+, y-, y+, z-, z+       */
  double bc_sign[54];   /* mirror sign per face and slot (6 x 9)            */
  int32_t rows_per_wave;/* 0: default; sweep chunk height (tuning)          */
  int32_t resistive;    /* 3D: add the resistive flux of flux_resistive.cuf
                           with eta = eta0 + eta1*exp(-(x^2+y^2)/eta_w^2) at
                           the face centres, and the resistive dt limit      */
  double x0, y0;        /* lower domain corner (cell centres x0 + (i+1/2)dx) */
  double eta0, eta1, eta_w;
  int32_t reserved[8];
} omhd_config;

typedef struct omhd_handle omhd_handle;

/* Allocate the device state for cfg on the current HIP device. */
int omhd_create(const omhd_config *cfg, omhd_handle **h);

/* Load U(ix,jx[,kx],9) (Fortran layout, ghosts ignored) and compute the first
 * CFL speed, as the reference does before its loop (main.f90:38-40). */
int omhd_upload(omhd_handle *h, const double *U);

/* Advance nsteps whole steps (main.f90:56-188 minus I/O).  t (optional) is
 * advanced by each dt; dt_trace/ch_trace (optional, nsteps doubles) receive
 * each step's dt and ch.  Stops at the first error, which is returned with the
 * number of completed steps in *done (optional). */
int omhd_step(omhd_handle *h, int32_t nsteps, double *t, double *dt_trace, double *ch_trace, int32_t *done);

/* Advance nsteps steps with the given dt[n] and ch[n] instead of the CFL
 * values (the reference's loop with set_dt replaced by a fixed sequence): the
 * frozen-dt forward map the adjoint differentiates, and the replay of a
 * recorded run.  No counterpart routine in the reference. */
int omhd_step_fixed(omhd_handle *h, int32_t nsteps, const double *dt, const double *ch);

/* Copy the state back as U(ix,jx[,kx],9) with ghosts filled by the boundary
 * conditions, and optionally V(ix,jx[,kx],4) (primitives of every cell). */
int omhd_download(omhd_handle *h, double *U, double *V);

/* Device pointer of the state (SoA, 9 planes) and its layout, for callers that
 * keep the data on the GPU (the PyTorch binding, the benchmark). */
int omhd_state_layout(omhd_handle *h, double **dev_state, int64_t *plane_stride, int32_t *row_pitch);

/* Replace the resistivity parameters of a handle created with resistive = 1
 * (the face arrays and the resistive dt limit follow). */
int omhd_set_resistivity(omhd_handle *h, double eta0, double eta1, double eta_w);

/* The state as W(9, nz, ny, nx) -- interior cells only, x fastest, slots in
 * the order of param.h:6-15 -- from or to host or device memory (the layout
 * of the PyTorch binding).  omhd_set_state also computes the CFL speed. */
int omhd_set_state(omhd_handle *h, const double *W);
int omhd_get_state(omhd_handle *h, double *W);

/* HIP stream the handle's kernels run on (as void*: a hipStream_t). */
void *omhd_stream(omhd_handle *h);

int omhd_destroy(omhd_handle *h);

/* Per-stage kernel timing with HIP events recorded on the handle's stream
 * around each stage launch (no counterpart in the reference, which has no
 * timers).  Enabling resets the accumulators; stage_ms[2] / launches[2] get
 * the summed milliseconds and launch counts of stage 1 and stage 2. */
int omhd_set_timing(omhd_handle *h, int32_t enable);
int omhd_get_timing(omhd_handle *h, double *stage_ms, int64_t *launches);

/* dt mode of the resident step.  The reference computes each step's dt on
 * the host from the reduced CFL speed (main.f90:80, set_dt.f90:21-65), and
 * omhd_step does the same by default (host dt mode: one host round trip per
 * step).  In device dt mode (device_dt = 1) each step's dt, CFL speed, dt/dx
 * and GLM factor are computed on the GPU after its second stage, the GLM
 * factor looked up in a table of the host's exp values, so omhd_step queues
 * all its steps and synchronises once at the end; use_graph = 1 also replays
 * one captured hipGraph per step (per-stage timing is then not recorded).
 * Results, t, dt/ch traces, *done and every stop (positivity, dt too small)
 * are those of host dt mode bit for bit; after a stop, later omhd_step calls
 * return the same stop until a new state is loaded (omhd_upload,
 * omhd_set_state, omhd_step_fixed).  Ideal MHD only: a handle with
 * resistivity is refused (a resistivity-limited dt puts the GLM argument
 * outside the table).  Needs a state (omhd_upload first). */
int omhd_set_mode(omhd_handle *h, int32_t device_dt, int32_t use_graph);

/* ---- ensemble of independent 2D grids ------------------------------------ */
/* No counterpart routine in the reference (each reference run is one grid):
 * n independent 2D grids, each with its own extents, boundaries and CFL dt,
 * advanced together by one persistent launch per RK stage over a tile table.
 * Members share lm_type and time_type.  In host dt mode (default) each
 * member's dt and GLM factor are computed on the host exactly as omhd_step
 * computes them; in device dt mode they are computed on the GPU, the GLM
 * factor looked up in a per-member table of the host's exp values, and whole
 * steps can be captured in a hipGraph (use_graph).  In both modes every
 * member is bitwise equal to its solo run. */
typedef struct omhd_batch omhd_batch;
int omhd_batch_create(const omhd_config *cfgs, int32_t n, omhd_batch **b);
int omhd_batch_upload(omhd_batch *b, int32_t member, const double *U);
int omhd_batch_download(omhd_batch *b, int32_t member, double *U, double *V);
int omhd_batch_set_mode(omhd_batch *b, int32_t device_dt, int32_t use_graph);
/* dt_trace (optional): nsteps x n doubles, member fastest */
int omhd_batch_step(omhd_batch *b, int32_t nsteps, double *dt_trace);
int omhd_batch_time(omhd_batch *b, double *t);
/* Times trial_steps steps per candidate tile height x launch shape (one
 * workgroup per tile, or a persistent grid) and keeps the fastest; the
 * ensemble state is restored after every trial. */
int omhd_batch_autotune(omhd_batch *b, int32_t trial_steps, int32_t *rows, int32_t *grid);
int omhd_batch_destroy(omhd_batch *b);

/* ---- discrete adjoint of the step ---------------------------------------- */
/* No counterpart in the reference.  The reverse-mode derivative of nsteps
 * steps of h (frozen dt/ch sequence) with respect to the initial state, and,
 * with the resistive flux, to the resistivity parameters: hand-written reverse
 * kernels of u2v, the limiter (the branch the forward pass took), HLLD, the
 * resistive flux, the boundary maps, glm_ss2 and the RK stages.
 * omhd_adjoint_forward runs the steps from h's current state (with the CFL
 * dt, recorded into dt_out/ch_out, or with dt_in/ch_in imposed), keeping the
 * state every `every` steps; omhd_adjoint_backward maps lam_T = dL/dU(T) to
 * lam_0 = dL/dU(0), replaying each segment from its checkpoint.  lam_T and
 * lam_0 use the W(9, nz, ny, nx) layout of omhd_set_state, host or device
 * memory.  2D and 3D, periodic and wall boundaries.  After a backward pass,
 * omhd_adjoint_eta_grad gives dL/d(eta0, eta1, eta_w). */
typedef struct omhd_adjoint omhd_adjoint;
int omhd_adjoint_create(omhd_handle *h, int32_t nsteps, int32_t every, omhd_adjoint **a);
int omhd_adjoint_forward(omhd_adjoint *a, const double *dt_in, const double *ch_in, double *dt_out, double *ch_out);
int omhd_adjoint_backward(omhd_adjoint *a, const double *lam_T, double *lam_0);
int omhd_adjoint_eta_grad(omhd_adjoint *a, double *grad3);
int omhd_adjoint_destroy(omhd_adjoint *a);

/* Device mirrors of host arrays for the drop-in entry points: with mirroring
 * on, a host array passed to any entry point above keeps a device copy from
 * call to call (keyed by its host address range), so an unchanged driver that
 * touches its arrays only through these routines and the snapshot routines
 * runs with its state resident in HBM.  Host copies then go stale: only
 * omhd_mirror_sync brings them up to date, and a host write must be followed
 * by omhd_mirror_update.  The shims of fileio_output / fileio_input
 * (2D_basic/fileio.f90:16,58; openmhd_amd/fortran/omhd_fileio.f90) do both.
 * Off by default; OMHD_MIRROR=1 in the environment turns it on for the whole
 * process (drivers such as 2D_basic/main.f90, whose lines 101-102 patch VL/VR
 * on the host between calls, must run without it).  omhd_mirror_mode(0) writes
 * every mirror back to the host and frees them. */
int omhd_mirror_mode(int32_t on);
int omhd_mirror_sync(const double *host, int64_t count);
int omhd_mirror_update(const double *host, int64_t count);

/* Library version and the gfx target it was built for. */
const char *omhd_version(void);

#ifdef __cplusplus
}
#endif
#endif /* OMHD_H */
