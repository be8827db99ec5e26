/*
 * vlp4d_lag.c -- CPU oracle for the Lagrange split-step vlp4d (TEST INFRASTRUCTURE).
 *
 * Plain-C restatement of the reference's OpenMP backend of the non-MPI vlp4d
 * miniapp (miniapps/vlp4d/openmp/, "VO/" below): a directional Strang split of
 * the 4-D Vlasov equation into 1-D periodic advections with the 6-point
 * (order 5) Lagrange interpolation, and the same FFT Poisson solve as the
 * spline miniapp (vo_field_poisson).  Arrays have no halo: N0*N1*N2*N3
 * doubles, x fastest (VO/init.cpp:291-292, VO/types.hpp).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "vlp4d_oracle.h"

#define LAG_OFFSET 2 /* VO/advect.hpp:67-70 (LAG_ORDER 5) */
#define LAG_PTS 6

/* The order-5 basis of VO/advect.hpp:72-86, px in [2, 3). */
void vl_lag_basis(double px, double coef[6]) {
  const double loc[] = {-1. / 24, 1. / 24., -1. / 12., 1. / 12., -1. / 24., 1. / 24.};
  const double pxm2 = px - 2.;
  const double sqrpxm2 = pxm2 * pxm2;
  const double pxm2_01 = pxm2 * (pxm2 - 1.);
  coef[0] = loc[0] * pxm2_01 * (pxm2 + 1.) * (pxm2 - 2.) * (pxm2 - 1.);
  coef[1] = loc[1] * pxm2_01 * (pxm2 - 2.) * (5 * sqrpxm2 + pxm2 - 8.);
  coef[2] = loc[2] * (pxm2 - 1.) * (pxm2 - 2.) * (pxm2 + 1.) * (5 * sqrpxm2 - 3 * pxm2 - 6.);
  coef[3] = loc[3] * pxm2 * (pxm2 + 1.) * (pxm2 - 2.) * (5 * sqrpxm2 - 7 * pxm2 - 4.);
  coef[4] = loc[4] * pxm2_01 * (pxm2 + 1.) * (5 * sqrpxm2 - 11 * pxm2 - 2.);
  coef[5] = loc[5] * pxm2_01 * pxm2 * (pxm2 + 1.) * (pxm2 - 2.);
}

/* Foot, stencil start and weights of one 1-D advection by `dep` on the
 * periodic axis [hmin, hmin + hwidth) with N points (VO/advect.hpp:137-154):
 * returns the first stencil index (may be negative; callers wrap). */
static inline int lag_stencil(double h, double dep, double hmin, double hwidth, double dh, double idh,
                              double coef[6]) {
  const double hstar = hmin + fmod(hwidth + h - dep - hmin, hwidth);
  int ipos1 = (int)floor((hstar - hmin) * idh);
  const double d_prev1 = LAG_OFFSET + idh * (hstar - (hmin + ipos1 * dh));
  ipos1 -= LAG_OFFSET;
  vl_lag_basis(d_prev1, coef);
  return ipos1;
}

#define IDX(c, ix, iy, ivx, ivy) \
  ((size_t)(ix) + (size_t)(c)->nxmax[0] * ((size_t)(iy) + (size_t)(c)->nxmax[1] * ((size_t)(ivx) + (size_t)(c)->nxmax[2] * (size_t)(ivy))))

/* advect_1D_x / _y (VO/advect.hpp:112-204): displacement dt*v of the cell's
 * own velocity along axis d (0 or 1). */
void vl_advect_xy(const vo_config *c, int d, const double *fn, double *fnp1, double dt) {
  const int N0 = c->nxmax[0], N1 = c->nxmax[1], N2 = c->nxmax[2], N3 = c->nxmax[3];
  const int N = c->nxmax[d];
  const double hmin = c->minPhy[d], hwidth = c->maxPhy[d] - c->minPhy[d], dh = c->dx[d], idh = 1. / c->dx[d];
  const int vd = d + 2;
#pragma omp parallel for collapse(2) schedule(static)
  for (int ivy = 0; ivy < N3; ivy++)
    for (int ivx = 0; ivx < N2; ivx++)
      for (int iy = 0; iy < N1; iy++)
        for (int ix = 0; ix < N0; ix++) {
          const int iv = d == 0 ? ivx : ivy;
          const double v = c->minPhy[vd] + iv * c->dx[vd];
          const int ih = d == 0 ? ix : iy;
          const double h = hmin + ih * dh;
          double coef[6];
          const int ip = lag_stencil(h, dt * v, hmin, hwidth, dh, idh, coef);
          double ftmp = 0.;
          for (int k = 0; k < LAG_PTS; k++) {
            const int j = (N + ip + k) % N;
            ftmp += coef[k] * fn[d == 0 ? IDX(c, j, iy, ivx, ivy) : IDX(c, ix, j, ivx, ivy)];
          }
          fnp1[IDX(c, ix, iy, ivx, ivy)] = ftmp;
        }
}

/* advect_1D_vx / _vy (VO/advect.hpp:206-301): displacement dt*E(x, y) along
 * velocity axis d (2 or 3), periodic in velocity too. */
void vl_advect_v(const vo_config *c, int d, const double *fn, double *fnp1, const double *e, double dt) {
  const int N0 = c->nxmax[0], N1 = c->nxmax[1], N2 = c->nxmax[2], N3 = c->nxmax[3];
  const int N = c->nxmax[d];
  const double hmin = c->minPhy[d], hwidth = c->maxPhy[d] - c->minPhy[d], dh = c->dx[d], idh = 1. / c->dx[d];
#pragma omp parallel for collapse(2) schedule(static)
  for (int ivy = 0; ivy < N3; ivy++)
    for (int ivx = 0; ivx < N2; ivx++)
      for (int iy = 0; iy < N1; iy++)
        for (int ix = 0; ix < N0; ix++) {
          const int iv = d == 2 ? ivx : ivy;
          const double v = hmin + iv * dh;
          double coef[6];
          const int ip = lag_stencil(v, dt * e[ix + (size_t)N0 * iy], hmin, hwidth, dh, idh, coef);
          double ftmp = 0.;
          for (int k = 0; k < LAG_PTS; k++) {
            const int j = (N + ip + k) % N;
            ftmp += coef[k] * fn[d == 2 ? IDX(c, ix, iy, j, ivy) : IDX(c, ix, iy, ivx, j)];
          }
          fnp1[IDX(c, ix, iy, ivx, ivy)] = ftmp;
        }
}

/* field_rho (VO/field.cpp:5-29): one running sum over (vy, vx) per point. */
void vl_field_rho(const vo_config *c, const double *fn, double *rho) {
  const int N0 = c->nxmax[0], N1 = c->nxmax[1], N2 = c->nxmax[2], N3 = c->nxmax[3];
#pragma omp parallel for
  for (int iy = 0; iy < N1; iy++)
    for (int ix = 0; ix < N0; ix++) {
      double sum = 0.;
      for (int ivy = 0; ivy < N3; ivy++)
        for (int ivx = 0; ivx < N2; ivx++) sum += fn[IDX(c, ix, iy, ivx, ivy)];
      rho[ix + (size_t)N0 * iy] = sum * c->dx[2] * c->dx[3];
    }
}

/* Initial conditions of VO/init.cpp:144-285 (no halo). */
static void cyl(const vo_config *c, double *fn, int axis) {
  const double PERIOD = 0.5 * M_PI, AMPLI = 4, cc = 0.50 * (6. / 16.), rc = 0.50 * (4. / 16.);
  for (int ivy = 0; ivy < c->nxmax[3]; ivy++)
    for (int ivx = 0; ivx < c->nxmax[2]; ivx++)
      for (int iy = 0; iy < c->nxmax[1]; iy++)
        for (int ix = 0; ix < c->nxmax[0]; ix++) {
          const double xx = axis == 0 ? c->minPhy[0] + ix * c->dx[0] : c->minPhy[1] + iy * c->dx[1];
          const double vv = axis == 0 ? c->minPhy[2] + ivx * c->dx[2] : c->minPhy[3] + ivy * c->dx[3];
          double hv = 0.0, hx = 0.0;
          if ((vv <= cc + rc) && (vv >= cc - rc)) hv = cos(PERIOD * ((vv - cc) / rc));
          else if ((vv <= -cc + rc) && (vv >= -cc - rc)) hv = -cos(PERIOD * ((vv + cc) / rc));
          if ((xx <= cc + rc) && (xx >= cc - rc)) hx = cos(PERIOD * ((xx - cc) / rc));
          else if ((xx <= -cc + rc) && (xx >= -cc - rc)) hx = -cos(PERIOD * ((xx + cc) / rc));
          fn[IDX(c, ix, iy, ivx, ivy)] = AMPLI * hx * hv;
        }
}

int vl_init_case(const vo_config *c, double *fn) {
  switch (c->idcase) {
    case 2: cyl(c, fn, 0); return VO_OK;  /* CYL02, VO/init.cpp:144-187 */
    case 5: cyl(c, fn, 1); return VO_OK;  /* CYL05, VO/init.cpp:189-232 */
    case 10:                              /* SLD10, VO/init.cpp:234-254 */
      for (int ivy = 0; ivy < c->nxmax[3]; ivy++)
        for (int ivx = 0; ivx < c->nxmax[2]; ivx++) {
          const double vy = c->minPhy[3] + ivy * c->dx[3], vx = c->minPhy[2] + ivx * c->dx[2];
          for (int iy = 0; iy < c->nxmax[1]; iy++) {
            const double y = c->minPhy[1] + iy * c->dx[1];
            for (int ix = 0; ix < c->nxmax[0]; ix++) {
              const double x = c->minPhy[0] + ix * c->dx[0];
              const double sum = vx * vx + vy * vy;
              fn[IDX(c, ix, iy, ivx, ivy)] =
                  (1. / (2 * M_PI)) * exp(-0.5 * (sum)) * (1 + 0.05 * (cos(0.5 * x) * cos(0.5 * y)));
            }
          }
        }
      return VO_OK;
    case 20: {                            /* TSI20, VO/init.cpp:256-285 */
      const double xi = 0.90, alpha = 0.05;
      for (int ivy = 0; ivy < c->nxmax[3]; ivy++)
        for (int ivx = 0; ivx < c->nxmax[2]; ivx++) {
          const double vy = c->minPhy[3] + ivy * c->dx[3], vx = c->minPhy[2] + ivx * c->dx[2];
          for (int iy = 0; iy < c->nxmax[1]; iy++)
            for (int ix = 0; ix < c->nxmax[0]; ix++) {
              const double x = c->minPhy[0] + ix * c->dx[0];
              const double sum = vx * vx + vy * vy;
              fn[IDX(c, ix, iy, ivx, ivy)] = (1 + alpha * (cos(0.5 * x))) * 1 / (2 * M_PI) *
                                             ((2 - 2 * xi) / (3 - 2 * xi)) * (1 + .5 * vx * vx / (1 - xi)) *
                                             exp(-.5 * sum);
            }
        }
      return VO_OK;
    }
    default:
      return VO_E_CASE;
  }
}

struct vl_state {
  vo_config c;
  double *fn, *fnp1, *rho, *ex, *ey;
  double *nrj, *mass; /* nbiter + 1 entries (VO/diags.cpp:4-11) */
  int iter;
};

/* field_rho, field_poisson and Diags::compute at index iter
 * (VO/vlp4d.cpp:55-58, VO/timestep.hpp:22-32, VO/diags.cpp:13-39). */
static void field_and_diag(vl_state *s) {
  vl_field_rho(&s->c, s->fn, s->rho);
  vo_field_poisson(&s->c, s->rho, s->ex, s->ey);
  double d[4];
  vo_diags_field(&s->c, s->rho, s->ex, s->ey, d);
  s->nrj[s->iter] = d[0];
  s->mass[s->iter] = d[3];
}

int vl_create(const vo_config *cfg, vl_state **out) {
  vl_state *s = (vl_state *)calloc(1, sizeof(vl_state));
  if (!s) return VO_E_ARG;
  s->c = *cfg;
  const size_t n = (size_t)cfg->nxmax[0] * cfg->nxmax[1] * cfg->nxmax[2] * cfg->nxmax[3];
  const size_t n2 = (size_t)cfg->nxmax[0] * cfg->nxmax[1];
  s->fn = (double *)calloc(n, sizeof(double));
  s->fnp1 = (double *)calloc(n, sizeof(double));
  s->rho = (double *)calloc(n2, sizeof(double));
  s->ex = (double *)calloc(n2, sizeof(double));
  s->ey = (double *)calloc(n2, sizeof(double));
  s->nrj = (double *)calloc((size_t)cfg->nbiter + 1, sizeof(double));
  s->mass = (double *)calloc((size_t)cfg->nbiter + 1, sizeof(double));
  if (!s->fn || !s->fnp1 || !s->rho || !s->ex || !s->ey || !s->nrj || !s->mass) {
    vl_destroy(s);
    return VO_E_ARG;
  }
  const int rc = vl_init_case(&s->c, s->fn);
  if (rc) {
    vl_destroy(s);
    return rc;
  }
  field_and_diag(s);
  *out = s;
  return VO_OK;
}

/* onetimestep, VO/timestep.hpp:11-55 (the result ends in fn). */
int vl_step(vl_state *s) {
  if (s->iter >= s->c.nbiter) return VO_E_ARG;
  const double dt = s->c.dt;
  s->iter++;
  vl_advect_xy(&s->c, 0, s->fn, s->fnp1, 0.5 * dt);
  vl_advect_xy(&s->c, 1, s->fnp1, s->fn, 0.5 * dt);
  field_and_diag(s);
  vl_advect_v(&s->c, 2, s->fn, s->fnp1, s->ex, dt);
  vl_advect_v(&s->c, 3, s->fnp1, s->fn, s->ey, dt);
  vl_advect_xy(&s->c, 0, s->fn, s->fnp1, 0.5 * dt);
  vl_advect_xy(&s->c, 1, s->fnp1, s->fn, 0.5 * dt);
  return VO_OK;
}

int vl_run(vl_state *s, int nsteps) {
  for (int k = 0; k < nsteps; k++) {
    const int rc = vl_step(s);
    if (rc) return rc;
  }
  return VO_OK;
}

void vl_destroy(vl_state *s) {
  if (!s) return;
  free(s->fn); free(s->fnp1); free(s->rho); free(s->ex); free(s->ey); free(s->nrj); free(s->mass);
  free(s);
}

int vl_iter(const vl_state *s) { return s->iter; }
const double *vl_f(const vl_state *s) { return s->fn; }
void vl_set_f(vl_state *s, const double *f) {
  memcpy(s->fn, f, sizeof(double) * (size_t)s->c.nxmax[0] * s->c.nxmax[1] * s->c.nxmax[2] * s->c.nxmax[3]);
}
const double *vl_field(const vl_state *s, int which) { return which == 0 ? s->rho : which == 1 ? s->ex : s->ey; }
const double *vl_diag_series(const vl_state *s, int which) { return which == 0 ? s->nrj : s->mass; }
