/*
 * vlp4d_oracle.h -- CPU oracle for the vlp4d spline time step (TEST INFRASTRUCTURE).
 *
 * This is a plain-C restatement of the one-rank semantics of the reference's
 * OpenMP backend (miniapps/vlp4d_mpi/openmp/, "VMO/" below).  It is the checker
 * for the HIP library: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Product code never calls it.
 *
 * Storage follows the reference exactly: every 4-D array has extent N_d+6 per
 * dimension with offset -3 (positions -3..N_d+2), layout_left (x fastest)
 * (VMO/Init.cpp:333-344, lib/openmp/View.hpp:52-451).  Indices are size_t
 * instead of the reference's int, so N_d > 209 does not overflow.
 *
 * Error convention: functions return VO_OK or a VO_E_* code instead of calling
 * exit()/abort() as the reference does (VMO/Advection.hpp:18-23,
 * VMO/Init.cpp:18-21,142-145).
 */
#ifndef VLP4D_ORACLE_H
#define VLP4D_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  VO_OK = 0,
  VO_E_FILE = 1,   /* deck not found            (VMO/Init.cpp:18-21)      */
  VO_E_CASE = 2,   /* unknown idcase            (VMO/Init.cpp:142-145)    */
  VO_E_CFL = 3,    /* foot outside the domain   (VMO/Advection.hpp:18-23) */
  VO_E_ARG = 4,    /* bad argument / allocation failure                  */
  VO_E_SIZE = 5    /* N_d < MMAX+2 or N_d+5 > 256 (spline buffer)        */
};

#define VO_HALO 3   /* HALO_PTS, VMO/Config.hpp:7 */
#define VO_MMAX 8   /* MMAX,     VMO/Config.hpp:8 */

/* Domain (VMO/Config.hpp:34-48) restricted to one rank. */
typedef struct {
  int nxmax[4];        /* points per dimension (x, y, vx, vy) */
  double minPhy[4];
  double maxPhy[4];
  double dx[4];        /* (max-min)/N, VMO/Init.cpp:80-82 */
  int idcase;
  double dt;
  int nbiter;
  int ifreq;
  int fxvx;
} vo_config;

/* Timer labels and order of VM/Timer.hpp:37-66. */
enum {
  VO_T_TOTAL = 0, VO_T_MAINLOOP, VO_T_PACK, VO_T_COMM, VO_T_UNPACK,
  VO_T_ADVEC2D, VO_T_ADVEC4D, VO_T_FIELD, VO_T_ALLREDUCE, VO_T_FOURIER,
  VO_T_DIAG, VO_T_SPLINE_XY, VO_T_SPLINE_VXVY, VO_T_COUNT
};

typedef struct vo_state vo_state;

/* ---- configuration ---------------------------------------------------- */
int vo_parse_deck(const char *path, vo_config *cfg);        /* VMO/Init.cpp:12-90 */
int vo_config_check(const vo_config *cfg);                   /* size limits         */
void vo_finish_config(vo_config *cfg);                       /* dx from min/max/N   */

/* ---- 4-D halo arrays --------------------------------------------------- */
size_t vo_halo_size(const vo_config *cfg);                   /* prod (N_d+6)        */
size_t vo_index(const vo_config *cfg, int ix, int iy, int ivx, int ivy);

/* ---- phases (all operate on (N+6)^4 halo arrays) ----------------------- */
int vo_init_case(const vo_config *cfg, double *fn);          /* VMO/Init.cpp:126-277, + extension cases 30, 31 */
void vo_exchange_halo(const vo_config *cfg, double *fn);     /* VMO/Communication.cpp:541-570 */
void vo_spline_xy(const vo_config *cfg, double *fn);         /* VMO/Spline.hpp:384-426 */
void vo_spline_vxvy(const vo_config *cfg, double *fn);       /* VMO/Spline.hpp:429-469 */
int vo_advect_2d(const vo_config *cfg, double *fn, double *tmp, double dt);           /* VMO/Advection.hpp:299-406 */
int vo_advect_4d(const vo_config *cfg, const double *ex, const double *ey,
                 double *fn, double *tmp_fn, double dt);                               /* VMO/Advection.hpp:409-498 */
void vo_feet(const vo_config *cfg, const double *ex, const double *ey, int ix, int iy, int ivx, int ivy,
             double dt, double xstar[4]);                                                /* VMO/Advection.hpp:166-218 */
void vo_field_rho(const vo_config *cfg, const double *fn, double *rho);               /* VMO/Field.cpp:5-55 */
void vo_field_poisson(const vo_config *cfg, double *rho, double *ex, double *ey);     /* VMO/Field.cpp:71-137 */
void vo_diags_field(const vo_config *cfg, const double *rho, const double *ex,
                    const double *ey, double out[4]);                                  /* VMO/Diags.cpp:16-48 */
double vo_l2norm(const vo_config *cfg, const double *fn);                              /* VMO/Diags.cpp:50-76 */

/* One periodic line of N values -> N+4 spline coefficients on positions -2..N+1,
 * i.e. the 1-D operator M = R*E (halo extension then prefilter). */
int vo_spline_line(int n, const double *f, double *c);

/* 2-D real FFT pair with the reference's conventions (lib/OpenMP_FFT.hpp:271-389):
 * unnormalised r2c along x then c2c along y; spectrum layout (nx/2+1) x ny,
 * complex stored as interleaved (re, im). */
void vo_rfft2(int nx, int ny, const double *in, double *out_c);
void vo_irfft2(int nx, int ny, const double *in_c, double *out);

/* ---- driver state ------------------------------------------------------- */
int vo_create(const vo_config *cfg, vo_state **out);         /* VMO/Init.cpp:324-373 + Vlp4d.cpp:29-34 */
int vo_step(vo_state *st);                                   /* VMO/Timestep.hpp:34-97 + swap */
int vo_run(vo_state *st, int nsteps);
void vo_destroy(vo_state *st);
int vo_iter(const vo_state *st);
const vo_config *vo_get_config(const vo_state *st);
double *vo_fn(vo_state *st);                                 /* current f^n (halo array) */
double *vo_field(vo_state *st, int which);                   /* 0 rho/phi, 1 ex, 2 ey */
/* Diagnostic series (length nbiter+1): which 0 nrj, 1 nrjx, 2 nrjy, 3 mass, 4 l2norm */
const double *vo_diag_series(const vo_state *st, int which);
double vo_timer_seconds(const vo_state *st, int which);
int vo_timer_calls(const vo_state *st, int which);
int vo_save(const vo_state *st, const char *dir);            /* VMO/Diags.cpp:78-104 */

/* Copy interior f^n (N^4, x fastest) out of / into a halo array. */
void vo_get_interior(const vo_config *cfg, const double *fn, double *out);
void vo_set_interior(const vo_config *cfg, double *fn, const double *in);

/* ---- Lagrange split-step vlp4d (miniapps/vlp4d/openmp, "VO/"), vlp4d_lag.c ----
 * Arrays without halo: N0*N1*N2*N3 doubles, x fastest. */
typedef struct vl_state vl_state;
void vl_lag_basis(double px, double coef[6]);                                     /* VO/advect.hpp:72-86 */
void vl_advect_xy(const vo_config *cfg, int d, const double *fn, double *fnp1, double dt);   /* :112-204, d 0|1 */
void vl_advect_v(const vo_config *cfg, int d, const double *fn, double *fnp1, const double *e,
                 double dt);                                                      /* :206-301, d 2|3 */
void vl_field_rho(const vo_config *cfg, const double *fn, double *rho);           /* VO/field.cpp:5-29 */
int vl_init_case(const vo_config *cfg, double *fn);                               /* VO/init.cpp:144-285 */
int vl_create(const vo_config *cfg, vl_state **out);   /* init + field + diag 0, VO/vlp4d.cpp:55-58 */
int vl_step(vl_state *st);                             /* onetimestep, VO/timestep.hpp:11-55 */
int vl_run(vl_state *st, int nsteps);
void vl_destroy(vl_state *st);
int vl_iter(const vl_state *st);
const double *vl_f(const vl_state *st);
void vl_set_f(vl_state *st, const double *f);
const double *vl_field(const vl_state *st, int which);        /* 0 phi, 1 ex, 2 ey (last solve) */
const double *vl_diag_series(const vl_state *st, int which);  /* 0 nrj, 1 mass (nbiter+1) */

#ifdef __cplusplus
}
#endif
#endif
