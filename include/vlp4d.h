/*
 * vlp4d.h -- C-ABI of the MI355X vlp4d time-step library (libvlp4d.so).
 *
 * The reference (miniapps/vlp4d_mpi) has no FFI: its hot path sits behind C++
 * free functions called from main.  Each entry point below names the
 * reference function it replaces (paths relative to the reference root;
 * VMO/ = miniapps/vlp4d_mpi/openmp/).  Plain pointers and sizes only.
 *
 * Conventions
 *  - Every function returns VLP_OK (0) or a VLP_E_* code and never exits;
 *    vlp_last_error() gives the message.  The reference's exit()/abort()
 *    behaviour is reproduced by the command-line driver only.
 *  - One context = one device + one HIP stream.  Calls on a context must be
 *    serialised by the caller (as the reference's single host thread does).
 *  - 4-D host arrays are the interior N^4 of the distribution function in the
 *    reference's order: x fastest, then y, vx, vy (layout_left,
 *    lib/openmp/View.hpp:52-451), always as double on the host side.
 *  - 2-D host arrays (rho/phi, Ex, Ey) are Nx*Ny, x fastest.
 */
#ifndef VLP4D_H
#define VLP4D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  VLP_OK = 0,
  VLP_E_FILE = 1,   /* deck missing/unreadable: reference abort(), VMO/Init.cpp:18-21        */
  VLP_E_CASE = 2,   /* unknown idcase: reference abort(), VMO/Init.cpp:142-145               */
  VLP_E_CFL = 3,    /* foot outside domain: reference "Time step is too large" + exit(0),
                       VMO/Advection.hpp:18-23                                              */
  VLP_E_ARG = 4,    /* invalid argument / state                                             */
  VLP_E_HIP = 5,    /* HIP runtime error (no device, out of memory, launch failure)          */
  VLP_E_SIZE = 6    /* N_d < MMAX+2 = 10 or N_d+5 > 256 (VMO/Init.cpp:85-89,
                       VMO/Types.hpp:31), or buffer size mismatch                           */
};

/* Precision of the distribution function and spline coefficients.
 * VLP_FP64: stored and interpolated in double (the reference's arithmetic).
 * VLP_FP32: stored in float; spline recursions, interpolation weights and
 *   stencil sums in double (storage rounding is then the only FP32 error).
 * VLP_FP32_ACC: stored in float and interpolated with float weights and float
 *   accumulation.
 * Feet, the charge density, the Poisson solve and the diagnostics are double
 * in every mode. */
enum { VLP_FP64 = 64, VLP_FP32 = 32, VLP_FP32_ACC = 33 };

/* Solvers: the spline semi-Lagrangian step of miniapps/vlp4d_mpi (default) or
 * the Lagrange split-step of miniapps/vlp4d (vlp_create_lagrange). */
enum { VLP_SCHEME_SPLINE = 0, VLP_SCHEME_LAGRANGE = 1 };

/* Domain of VMO/Config.hpp:34-48 for one rank (dx derived as (max-min)/N). */
typedef struct {
  int nxmax[4];       /* points in x, y, vx, vy                         */
  double minPhy[4];
  double maxPhy[4];
  double dx[4];       /* filled by the library                          */
  int idcase;         /* 2, 5, 10, 20 (VMO/Init.cpp:129-146), or 30, 31  */
  double dt;
  int nbiter;
  int ifreq;
  int fxvx;
} vlp_config;

typedef struct vlp_ctx vlp_ctx;

/* Initial conditions beyond the reference's (its dispatch, VMO/Init.cpp:
 * 126-150, knows 2, 5, 10 and 20 and aborts on any other idcase); spline
 * solver only.  Grid point i of dimension d is minPhy_d + i dx_d.
 *   VLP_CASE_STRONG_LANDAU (30): (1 + 0.5 cos(0.5 x) + 0.5 cos(0.5 y))
 *     exp(-(vx^2 + vy^2)/2) / (2 pi)
 *   VLP_CASE_TWO_STREAM (31): (1 + 1e-3 cos(0.2 x)) exp(-vy^2/2)
 *     [exp(-(vx - 2.4)^2/2) + exp(-(vx + 2.4)^2/2)] / (4 pi)
 * The field solve treats them like any case without the rho -= 1 shift of
 * idcases 10 and 20 (the k = 0 mode is dropped by the solve either way). */
enum { VLP_CASE_STRONG_LANDAU = 30, VLP_CASE_TWO_STREAM = 31 };

/* Phase identifiers for vlp_timer(): the labels of miniapps/vlp4d_mpi/Timer.hpp:37-66. */
enum {
  VLP_T_TOTAL = 0, VLP_T_MAINLOOP, VLP_T_PACK, VLP_T_COMM, VLP_T_UNPACK,
  VLP_T_ADVEC2D, VLP_T_ADVEC4D, VLP_T_FIELD, VLP_T_ALLREDUCE, VLP_T_FOURIER,
  VLP_T_DIAG, VLP_T_SPLINE_XY, VLP_T_SPLINE_VXVY, VLP_T_COUNT
};

/* Device buffers exposed for phase-level parity checks (vlp_get_buffer). */
enum {
  VLP_BUF_F = 0,      /* f^n interior, N^4                                            */
  VLP_BUF_CXY = 1,    /* spline coeffs along x,y: x,y in [-2,N+1], vx,vy interior     */
  VLP_BUF_C4 = 2,     /* spline coeffs along all 4 dims on [-2,N+1]^4                  */
  VLP_BUF_RHO = 3,    /* after a field solve: phi (the reference overwrites rho with phi) */
  VLP_BUF_EX = 4,
  VLP_BUF_EY = 5,
  VLP_BUF_COUNT = 6
};

/* Library version string. */
const char *vlp_version(void);

/* Parse a .dat deck with the reference's semantics (VMO/Init.cpp:12-90 `import`).
 * Pure host code; needs no GPU. */
int vlp_parse_deck(const char *dat_path, vlp_config *cfg);

/* init() of VMO/Init.cpp:324-373 plus the initial field solve and diagnostics
 * of VMO/Vlp4d.cpp:29-34: allocates device state on `device`, runs the init
 * case, solves Poisson, records diagnostics at iter 0.  precision: VLP_FP64,
 * VLP_FP32 or VLP_FP32_ACC.  dat_path variant parses the deck first. */
int vlp_create(const char *dat_path, int precision, int device, vlp_ctx **out);
int vlp_create_config(const vlp_config *cfg, int precision, int device, vlp_ctx **out);
/* The Lagrange split-step solver of the non-MPI miniapp, FP64: init() and
 * the initial field solve of miniapps/vlp4d/openmp/init.cpp:287-302 and
 * vlp4d.cpp:55-58 (its own TSI20 profile, init.cpp:256-285).  vlp_step then
 * runs its onetimestep (timestep.hpp:11-55: 1-D order-5 Lagrange advections
 * x/2, y/2, field solve with that step's diagnostics, vx, vy, x/2, y/2);
 * vlp_diags fills nrj and mass (diags.cpp:13-39; nrjx/nrjy are computed as
 * well, l2norm stays 0), vlp_save writes its 3-column nrj.out
 * (diags.cpp:41-52).  The spline-only calls (phases, vlp_autotune,
 * vlp_adjoint_step, ensembles) return VLP_E_ARG on such a context. */
int vlp_create_lagrange(const vlp_config *cfg, int device, vlp_ctx **out);

/* The fxvx diagnostics of the current iteration (print_fxvx,
 * VMO/Advection.hpp:500-590; miniapps/vlp4d/openmp/advect.hpp:318-341):
 * dir/fxvx_itNNNNNN.csv = f(x, y=0, vx, vy=Nvy/2), one row per vx, and for
 * spline contexts dir/fyvy_itNNNNNN.csv = f(x=0, y, vx=Nvx/2, vy), one row
 * per vy ("%20.13le" values, ", " between them).  The spline miniapp's own
 * routine writes zeros (its guard is false at y index 0); these are the
 * sections it means.  The directory must exist. */
int vlp_write_fxvx(vlp_ctx *ctx, const char *dir);

/* finalize() of VMO/Init.cpp:375-382 (without the file output; see vlp_save). */
int vlp_destroy(vlp_ctx *ctx);

/* `nsteps` iterations of onetimestep (VMO/Timestep.hpp:34-97) followed by
 * fn.swap(fnp1) (VMO/Vlp4d.cpp:137-147).  Returns VLP_E_CFL if a foot left the
 * domain (state is then invalid; vlp_iter names the failing step).  With
 * nsteps > 1 (spline solver) step s + 1 is queued before step s's error flag
 * is read, so the GPU does not wait for the host between steps; results are
 * those of nsteps calls of one step. */
int vlp_step(vlp_ctx *ctx, int nsteps);

/* Current iteration count (the reference's `iter`). */
int vlp_iter(const vlp_ctx *ctx);

/* Copy of the parsed/derived configuration. */
int vlp_get_config(const vlp_ctx *ctx, vlp_config *cfg);

/* Diagnostic series of Diags (VMO/Diags.cpp:16-76) for iterations 0..n-1
 * (n <= iter+1).  Any pointer may be NULL. */
int vlp_diags(vlp_ctx *ctx, double *nrj, double *nrjx, double *nrjy, double *mass, double *l2norm, int n);

/* Diags::save (VMO/Diags.cpp:78-104): writes dir/nrj.out and dir/l2norm.out
 * in the reference's format for iterations 0..nbiter. */
int vlp_save(vlp_ctx *ctx, const char *dir);

/* Interior f^n as double, n = Nx*Ny*Nvx*Nvy (View::updateSelf / updateDevice,
 * lib/openmp/View.hpp:225-235).  vlp_set_f also recomputes rho, E and the
 * iteration-`iter` diagnostics from the new f, like VMO/Vlp4d.cpp:30-34. */
int vlp_get_f(vlp_ctx *ctx, double *host, size_t n);
int vlp_set_f(vlp_ctx *ctx, const double *host, size_t n);

/* Copy an internal buffer (VLP_BUF_*) to the host as double; n must equal its
 * element count (vlp_buffer_size). */
size_t vlp_buffer_size(const vlp_ctx *ctx, int which);
int vlp_get_buffer(vlp_ctx *ctx, int which, double *host, size_t n);

/* ---- phase entry points (the reference's component functions) ----------- */

/* exchangeHalo + Spline::computeCoeff_xy (VMO/Communication.cpp:541-570,
 * VMO/Spline.hpp:384-426): f^n -> CXY. */
int vlp_spline_xy(vlp_ctx *ctx);

/* Advection::advect_2D_xy with dt/2, field_rho, field_reduce, field_poisson
 * (VMO/Advection.hpp:299-406, VMO/Field.cpp:5-137): CXY -> rho, E, phi. */
int vlp_advect2d_poisson(vlp_ctx *ctx);

/* Spline::computeCoeff_vxvy (VMO/Spline.hpp:429-469): CXY -> C4. */
int vlp_spline_vxvy(vlp_ctx *ctx);

/* Advection::advect_4D (VMO/Advection.hpp:409-498) then field_rho/reduce/
 * poisson, Diags::compute and computeL2norm (VMO/Timestep.hpp:73-91):
 * C4, E -> f^{n+1}, written into F; iteration counter advanced. */
int vlp_advect4d_poisson(vlp_ctx *ctx);

/* Timer table of miniapps/vlp4d_mpi/Timer.hpp:70-75: seconds and calls for
 * VLP_T_* (GPU phases measured with HIP events). */
int vlp_timer(vlp_ctx *ctx, int which, double *seconds, int *calls);
const char *vlp_timer_label(int which);

/* Which kernel variants this context's geometry selects (bit 0: multi-lane
 * scan for the x spline, bit 1: multi-lane scan for the y/vx/vy splines,
 * bit 2: vx-blocked 4-D advection, bit 3: 4-cell blocks (else 2), bit 4: E
 * rows staged in LDS, bit 5: XCD-aware workgroup order, bit 6: advection
 * arithmetic in float, bit 7: y-streaming 4-D advection (replaces the
 * vx-blocked one), bits 8-11: its vx tile, bit 12: fused x/y spline + 2-D
 * advection kernel (replaces the separate x, y and advect2d kernels), bits
 * 13 / 14: the vx / vy spline pass uses 128-position LDS tiles, four lanes
 * per 128-point line, bit 15: both passes run as MFMA products with the line
 * operator, VLP4D_SPLINE_MFMA=1 at creation; bits 13 and 14 are then clear).
 * No reference counterpart: introspection for tests and reports. */
int vlp_kernel_variants(const vlp_ctx *ctx);

/* Message of the last error on this context (or the last global error when
 * ctx is NULL). */
const char *vlp_last_error(const vlp_ctx *ctx);

/* Autotuner: the 4-D advection's workgroup shape (vx tile, LDS plane
 * buffers) and the kernel of the vx and of the vy spline pass (256-, 128- or
 * 64-position LDS tiles, or the line kernels): every choice the geometry
 * admits is timed on the context's own buffers and the fastest is used by its
 * later steps when it beats the default choice (the preference order used
 * without tuning) by more than 2 %, else the default stays; csv_path
 * (optional) receives the timings, one row per choice with the default and
 * the chosen ones marked.  Follows the reference's tile tuner, which
 * scans the candidate tiles of every registered kernel, keeps the minimum and
 * logs a CSV (miniapps/vlp4d_mpi/kokkos/Tile_size_tuning.hpp:51-96,127-148;
 * per-kernel registration miniapps/vlp4d_mpi/kokkos/Tuning.hpp:26-65).  The
 * timed launches write F of the next step and the coefficient buffers, which
 * that step overwrites.  The VLP4D_YS_* (VLP4D_SPLINE_TILE/_WIDE) environment
 * overrides pin the advection (spline) choice. */
int vlp_autotune(vlp_ctx *ctx, const char *csv_path);
/* The tuned (vx tile, plane buffers), 0 and 0 before vlp_autotune. */
int vlp_tuned_tile(const vlp_ctx *ctx, int *vx_tile, int *plane_buffers);

/* Adjoint of one step at the context's current f (VLP_FP64 contexts):
 * grad = (d f^{n+1} / d f^n)^T g, g and grad interior arrays of n = Nx*Ny*
 * Nvx*Nvy doubles (x fastest), the derivative taken through the spline
 * coefficients, the 4-D interpolation, the characteristic feet and the field
 * of the half step.  Where a velocity foot sits exactly on the box bound
 * (the clamp of VMO/Advection.hpp:214-217 has a corner there) the branch in
 * force is differentiated.  Contributions are summed with atomic adds, so
 * repeated calls agree to round-off, not bit for bit.  The context is not
 * advanced; its coefficient, density and field buffers are left holding the
 * intermediates of the recomputed step.
 * It has no counterpart in the reference, which has no derivative code; it
 * returns VLP_E_CFL where the step from this f would. */
int vlp_adjoint_step(vlp_ctx *ctx, const double *g, double *grad, size_t n);
/* Adjoint of nsteps steps from the context's current f (VLP_FP64):
 * grad = (d f^{n+nsteps} / d f^n)^T g, by two-level checkpointing (every
 * ceil(sqrt(nsteps))-th state kept on the device, the others recomputed in
 * the reverse sweep; about 2 sqrt(nsteps) extra copies of f).  Needs iter +
 * nsteps <= nbiter; the context ends at its starting f and iteration (the
 * recomputed steps write the diagnostics a run would write).  No counterpart
 * in the reference. */
int vlp_adjoint_steps(vlp_ctx *ctx, int nsteps, const double *g, double *grad, size_t n);
/* Gradient of the field energy integrated over nsteps steps,
 * J = dt * sum_{k=1..nsteps} W(f^{n+k}),  W(f) = sum_{x,y} (Ex^2 + Ey^2) dx dy
 * of the field solved from f^{n+k} at the end of step k (the quantity whose
 * log sqrt is the nrj diagnostic, VMO/Diags.cpp:16-48, recorded by
 * VMO/Timestep.hpp:77-91), with respect to the context's current f^n
 * (VLP_FP64), by the same checkpointed reverse sweep as vlp_adjoint_steps
 * with the source dt dW/df added at every f^{n+k}; the forward sweep keeps
 * every checkpoint_every-th state (<= 0: ceil(sqrt(nsteps))) and the reverse
 * sweep recomputes the others.  *J receives J.
 * vweight NULL: grad is dJ/df^n (n = Nx*Ny*Nvx*Nvy doubles, x fastest);
 * else vweight holds Nvx*Nvy weights w(vx, vy) (vx fastest) and grad the
 * moment sum_v w(v) dJ/df^n(x, y, v) (n = Nx*Ny) -- the gradient with
 * respect to a density perturbation f^n += w(v) delta(x, y).  The context
 * ends at its starting f and iteration.  VLP_E_ARG for idcases 2 and 6, whose
 * field does not depend on f.  No counterpart in the reference. */
int vlp_adjoint_field_energy(vlp_ctx *ctx, int nsteps, int checkpoint_every, const double *vweight, double *J,
                             double *grad, size_t n);

/* The field of a caller-held distribution function, without a context: the
 * density (field_rho + field_reduce, VMO/Field.cpp:5-69, one rank), the
 * Poisson solve (field_poisson, VMO/Field.cpp:71-137 / VMO/Efield.cpp:3-106),
 * Diags::compute and computeL2norm (VMO/Diags.cpp:16-76) -- what the
 * reference's onetimestep runs after each advection (VMO/Timestep.hpp:53-67,
 * 77-91).  cfg gives the grid, box and idcase (dx is derived; dt, nbiter,
 * ifreq, fxvx are ignored but must pass the vlp_create checks).  Device
 * pointers: f holds Nx*Ny*Nvx*Nvy values, x fastest, double when f_bits is
 * 64, float when 32; phi, ex, ey receive Nx*Ny doubles (phi: the reference's
 * rho_ after the solve, i.e. the potential; rho itself for idcases 2 and 6);
 * diag receives 5 doubles: nrj, nrjx, nrjy, mass, l2norm.  stream: a
 * hipStream_t (NULL: the null stream).  Allocates its scratch (about
 * 8 * Nx*Ny*Nvy bytes) per call and returns when the results are ready. */
int vlp_field_solve(const vlp_config *cfg, const void *f, int f_bits, double *phi, double *ex, double *ey,
                    double *diag, void *stream);

/* vlp_field_solve without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight loops around a solver
 * that owns its buffers).  No counterpart in the reference.  Same arguments
 * and results; the scratch lives in a static device array of
 * VLP_FIELD_SCRATCH_DOUBLES doubles (24 MiB) that each GPU holds until the
 * process ends, and grids whose scratch exceeds it (about Nx*Ny*Nvy beyond
 * 2.6e6) return VLP_E_SIZE.  The first call on a device looks up the array's
 * address on that device; calls only enqueue work on `stream` (no
 * synchronisation).  Calls on one GPU share the array: calls that overlap
 * in time on the same GPU race. */
#define VLP_FIELD_SCRATCH_DOUBLES (3u << 20)
int vlp_field_solve_noalloc(const vlp_config *cfg, const void *f, int f_bits, double *phi, double *ex, double *ey,
                            double *diag, void *stream);

/* Ensembles: independent contexts (any grid sizes and precisions, one
 * device) advanced together, two steps of every member per launch of one
 * captured HIP graph (their kernels overlap on the GPU); an odd remainder
 * runs eagerly.  Same results as vlp_step on each member.  No reference
 * counterpart (the reference runs one solve per process); for batched
 * studies of many small solves.  A context belongs to at most one ensemble
 * at a time; vlp_destroy of a member returns VLP_E_ARG until the ensemble is
 * destroyed.  Calls on the ensemble and on its members must not overlap. */
typedef struct vlp_ensemble vlp_ensemble;
int vlp_ensemble_create(vlp_ctx *const *members, int n, vlp_ensemble **out);
/* Advance every member by nsteps (each member's iter + nsteps <= nbiter). */
int vlp_ensemble_step(vlp_ensemble *ens, int nsteps);
int vlp_ensemble_size(const vlp_ensemble *ens);
int vlp_ensemble_destroy(vlp_ensemble *ens);

#ifdef __cplusplus
}
#endif
#endif
