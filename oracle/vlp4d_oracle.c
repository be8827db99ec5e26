/*
 * vlp4d_oracle.c -- CPU oracle for the vlp4d spline time step (TEST INFRASTRUCTURE).
 *
 * A plain-C restatement of the one-rank semantics of the reference OpenMP
 * backend miniapps/vlp4d_mpi/openmp/ ("VMO/").  Every function names the
 * reference lines it follows.  It is the checker for the HIP library and the
 * timed CPU baseline ("port"); it is never part of the product path.
 *
 * Floating point: build with -O2 -ffp-contract=off and no -march, like the
 * reference's "-O3" x86-64 build (VMO/CMakeLists.txt:17-19), so that every
 * statement rounds the way the reference's does.  The one deliberate
 * difference is the FFT: FFTW3 (the reference's dependency, unpinned, absent
 * here) is replaced by a direct DFT with exactly-reduced twiddles; it agrees
 * with any exact-arithmetic DFT to ~1e-16 relative.
 *
 * Parity pinning of each function is done in tests/test_oracle_*.py (closed
 * forms, invariants, brute force); see the module docstrings there.
 */
#include "vlp4d_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* indexing helpers                                                          */
/* ------------------------------------------------------------------------ */

static inline size_t ext(const vo_config *c, int d) { return (size_t)c->nxmax[d] + 2 * VO_HALO; }

size_t vo_halo_size(const vo_config *cfg) {
  return ext(cfg, 0) * ext(cfg, 1) * ext(cfg, 2) * ext(cfg, 3);
}

/* layout_left with offset -3 in every dimension (lib/openmp/View.hpp:325-354) */
size_t vo_index(const vo_config *c, int ix, int iy, int ivx, int ivy) {
  const size_t e0 = ext(c, 0), e1 = ext(c, 1), e2 = ext(c, 2);
  return (size_t)(ix + VO_HALO) + e0 * ((size_t)(iy + VO_HALO) + e1 * ((size_t)(ivx + VO_HALO) + e2 * (size_t)(ivy + VO_HALO)));
}

#define F4(c, a, ix, iy, ivx, ivy) (a)[vo_index((c), (ix), (iy), (ivx), (ivy))]

static inline double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* A1: deck parser, VMO/Init.cpp:12-90                                       */
/* ------------------------------------------------------------------------ */

/* Advance to the character after the next ':' (VMO/Init.cpp:33-35). The
 * reference loops forever at EOF; the oracle reports VO_E_FILE instead. */
static int skip_to_colon(FILE *s) {
  int ch;
  do {
    ch = fgetc(s);
    if (ch == EOF) return -1;
  } while (ch != ':');
  return 0;
}

void vo_finish_config(vo_config *cfg) {
  for (int i = 0; i < 4; i++)
    cfg->dx[i] = (cfg->maxPhy[i] - cfg->minPhy[i]) / (double)cfg->nxmax[i];  /* VMO/Init.cpp:80-82 */
}

int vo_parse_deck(const char *path, vo_config *cfg) {
  char idcase[8] = {0};
  FILE *s = fopen(path, "r");
  if (!s) return VO_E_FILE;
  memset(cfg, 0, sizeof(*cfg));
  for (int i = 0; i < 4; i++) {                     /* VMO/Init.cpp:32-39 */
    int nx_tmp = 0;
    if (skip_to_colon(s)) goto bad;
    if (fscanf(s, " %d\n", &nx_tmp) != 1) goto bad;
    cfg->nxmax[i] = nx_tmp;
  }
  for (int i = 0; i < 4; i++) {                     /* VMO/Init.cpp:41-51 */
    if (skip_to_colon(s)) goto bad;
    if (fscanf(s, " %le\n", &cfg->minPhy[i]) != 1) goto bad;
    if (skip_to_colon(s)) goto bad;
    if (fscanf(s, " %le\n", &cfg->maxPhy[i]) != 1) goto bad;
  }
  if (skip_to_colon(s)) goto bad;                   /* VMO/Init.cpp:53-56 */
  if (!fgets(idcase, 7, s)) goto bad;
  if (skip_to_colon(s)) goto bad;                   /* VMO/Init.cpp:58-61 */
  if (fscanf(s, " %le\n", &cfg->dt) != 1) goto bad;
  if (skip_to_colon(s)) goto bad;                   /* VMO/Init.cpp:63-66 */
  if (fscanf(s, " %d\n", &cfg->nbiter) != 1) goto bad;
  if (skip_to_colon(s)) goto bad;                   /* VMO/Init.cpp:68-71 */
  if (fscanf(s, " %d\n", &cfg->ifreq) != 1) goto bad;
  if (skip_to_colon(s)) goto bad;                   /* VMO/Init.cpp:73-76 */
  if (fscanf(s, " %d\n", &cfg->fxvx) != 1) goto bad;
  cfg->idcase = atoi(idcase);                       /* VMO/Init.cpp:78 */
  fclose(s);
  vo_finish_config(cfg);
  return VO_OK;
bad:
  fclose(s);
  return VO_E_FILE;
}

/* N_d >= MMAX+2 is only a warning in the reference (VMO/Init.cpp:85-89) but
 * the alpha-sums then read outside the line; the spline buffer needs
 * N_d+5 <= 256 (VMO/Types.hpp:31, VMO/Spline.hpp:125). */
int vo_config_check(const vo_config *cfg) {
  for (int i = 0; i < 4; i++) {
    if (cfg->nxmax[i] < VO_MMAX + 2) return VO_E_SIZE;
    if (cfg->nxmax[i] + 5 > 256) return VO_E_SIZE;
  }
  if (cfg->ifreq <= 0 || cfg->nbiter < 0) return VO_E_ARG;
  return VO_OK;
}

/* ------------------------------------------------------------------------ */
/* A2: initial conditions, VMO/Init.cpp:126-277                              */
/* ------------------------------------------------------------------------ */

static void case_cyl02(const vo_config *c, double *fn) {  /* VMO/Init.cpp:152-186 */
  const double PI = M_PI, AMPLI = 4, PERIOD = 0.5 * PI;
  const double cc = 0.50 * (6. / 16.), rc = 0.50 * (4. / 16.);
  for (int ivy = 0; ivy <= c->nxmax[3] - 1; ivy++)
    for (int ivx = 0; ivx <= c->nxmax[2] - 1; ivx++) {
      const double vx = c->minPhy[2] + ivx * c->dx[2];
      for (int iy = 0; iy <= c->nxmax[1] - 1; iy++)
        for (int ix = 0; ix <= c->nxmax[0] - 1; ix++) {
          const double x = c->minPhy[0] + ix * c->dx[0];
          const double xx = x, vv = vx;
          double hv = 0.0, hx = 0.0;
          if ((vv <= cc + rc) && (vv >= cc - rc)) hv = cos(PERIOD * ((vv - cc) / rc));
          else if ((vv <= -cc + rc) && (vv >= -cc - rc)) hv = -cos(PERIOD * ((vv + cc) / rc));
          if ((xx <= cc + rc) && (xx >= cc - rc)) hx = cos(PERIOD * ((xx - cc) / rc));
          else if ((xx <= -cc + rc) && (xx >= -cc - rc)) hx = -cos(PERIOD * ((xx + cc) / rc));
          F4(c, fn, ix, iy, ivx, ivy) = (AMPLI * hx * hv);
        }
    }
}

/* The reference's CYL05 loops use '<' in three dimensions, so the last index
 * of ivx, iy and ix stays zero (VMO/Init.cpp:196-199); kept as is. */
static void case_cyl05(const vo_config *c, double *fn) {  /* VMO/Init.cpp:188-224 */
  const double PI = M_PI, AMPLI = 4, PERIOD = 0.5 * PI;
  const double cc = 0.50 * (6. / 16.), rc = 0.50 * (4. / 16.);
  for (int ivy = 0; ivy <= c->nxmax[3] - 1; ivy++)
    for (int ivx = 0; ivx < c->nxmax[2] - 1; ivx++) {
      const double vy = c->minPhy[3] + ivy * c->dx[3];
      for (int iy = 0; iy < c->nxmax[1] - 1; iy++) {
        const double y = c->minPhy[1] + iy * c->dx[1];
        for (int ix = 0; ix < c->nxmax[0] - 1; ix++) {
          const double xx = y, vv = vy;
          double hv = 0.0, hx = 0.0;
          if ((vv <= cc + rc) && (vv >= cc - rc)) hv = cos(PERIOD * ((vv - cc) / rc));
          else if ((vv <= -cc + rc) && (vv >= -cc - rc)) hv = -cos(PERIOD * ((vv + cc) / rc));
          if ((xx <= cc + rc) && (xx >= cc - rc)) hx = cos(PERIOD * ((xx - cc) / rc));
          else if ((xx <= -cc + rc) && (xx >= -cc - rc)) hx = -cos(PERIOD * ((xx + cc) / rc));
          F4(c, fn, ix, iy, ivx, ivy) = AMPLI * hx * hv;
        }
      }
    }
}

static void case_sld10(const vo_config *c, double *fn) {  /* VMO/Init.cpp:226-245 */
#pragma omp parallel for collapse(2)
  for (int ivy = 0; ivy <= c->nxmax[3] - 1; ivy++)
    for (int ivx = 0; ivx <= c->nxmax[2] - 1; ivx++) {
      const double vy = c->minPhy[3] + ivy * c->dx[3];
      const double vx = c->minPhy[2] + ivx * c->dx[2];
      for (int iy = 0; iy <= c->nxmax[1] - 1; iy++) {
        const double y = c->minPhy[1] + iy * c->dx[1];
        for (int ix = 0; ix <= c->nxmax[0] - 1; ix++) {
          const double x = c->minPhy[0] + ix * c->dx[0];
          const double sum = (vx * vx + vy * vy);
          F4(c, fn, ix, iy, ivx, ivy) = (1. / (2 * M_PI)) * exp(-0.5 * (sum)) * (1 + 0.05 * (cos(0.5 * x) * cos(0.5 * y)));
        }
      }
    }
}

static void case_tsi20(const vo_config *c, double *fn) {  /* VMO/Init.cpp:247-277 */
  const double vd = 2.4, vthx = 1.0, vthy = 0.5, alphax = 0.05, alphay = 0.25, kx = 0.2, ky = 0.2;
#pragma omp parallel for collapse(2)
  for (int ivy = 0; ivy <= c->nxmax[3] - 1; ivy++)
    for (int ivx = 0; ivx <= c->nxmax[2] - 1; ivx++) {
      const double vy = c->minPhy[3] + ivy * c->dx[3];
      const double vx = c->minPhy[2] + ivx * c->dx[2];
      for (int iy = 0; iy <= c->nxmax[1] - 1; iy++)
        for (int ix = 0; ix <= c->nxmax[0] - 1; ix++) {
          const double yy = c->minPhy[1] + iy * c->dx[1];
          const double xx = c->minPhy[0] + ix * c->dx[0];
          F4(c, fn, ix, iy, ivx, ivy) =
              1. / (4 * M_PI * vthx * vthy) *
              (1 - alphax * sin(kx * xx) - alphay * sin(ky * yy)) *
              (exp(-.5 * (pow((vx - vd) / vthx, 2) + pow(vy / vthy, 2))) +
               exp(-.5 * (pow((vx + vd) / vthx, 2) + pow(vy / vthy, 2))));
        }
    }
}

/* Extension cases with no reference counterpart (the reference's dispatch,
 * VMO/Init.cpp:126-150, knows 2, 5, 10 and 20 only), same loop form as SLD10:
 *   30  strong Landau damping: (1 + 0.5 cos(0.5 x) + 0.5 cos(0.5 y)) exp(-(vx^2 + vy^2)/2) / (2 pi)
 *   31  two-stream: (1 + 1e-3 cos(0.2 x)) exp(-vy^2/2) [exp(-(vx-2.4)^2/2) + exp(-(vx+2.4)^2/2)] / (4 pi) */
static void case_ext(const vo_config *c, double *fn) {
#pragma omp parallel for collapse(2)
  for (int ivy = 0; ivy <= c->nxmax[3] - 1; ivy++)
    for (int ivx = 0; ivx <= c->nxmax[2] - 1; ivx++) {
      const double vy = c->minPhy[3] + ivy * c->dx[3];
      const double vx = c->minPhy[2] + ivx * c->dx[2];
      for (int iy = 0; iy <= c->nxmax[1] - 1; iy++) {
        const double y = c->minPhy[1] + iy * c->dx[1];
        for (int ix = 0; ix <= c->nxmax[0] - 1; ix++) {
          const double x = c->minPhy[0] + ix * c->dx[0];
          double v;
          if (c->idcase == 30) {
            v = (1. / (2 * M_PI)) * exp(-0.5 * (vx * vx + vy * vy)) * (1 + 0.5 * cos(0.5 * x) + 0.5 * cos(0.5 * y));
          } else {
            const double a = vx - 2.4, b = vx + 2.4;
            v = (1. / (4 * M_PI)) * (1 + 1e-3 * cos(0.2 * x)) * exp(-0.5 * (vy * vy)) *
                (exp(-0.5 * (a * a)) + exp(-0.5 * (b * b)));
          }
          F4(c, fn, ix, iy, ivx, ivy) = v;
        }
      }
    }
}

/* VMO/Init.cpp:126-150: fn is expected to be zero on entry (VMO/Init.cpp:370-371). */
int vo_init_case(const vo_config *cfg, double *fn) {
  switch (cfg->idcase) {
    case 2: case_cyl02(cfg, fn); break;
    case 5: case_cyl05(cfg, fn); break;
    case 10: case_sld10(cfg, fn); break;
    case 20: case_tsi20(cfg, fn); break;
    case 30:
    case 31: case_ext(cfg, fn); break;
    default: return VO_E_CASE;
  }
  return VO_OK;
}

/* ------------------------------------------------------------------------ */
/* A4: one-rank halo exchange with the spline boundary condition             */
/*   periodic copies: VMO/Communication.hpp:713-753 (pack/local_copy/unpack)  */
/*   alpha-sums:      VMO/Communication.hpp:756-898 (boundary_condition)      */
/*   which cells:     VMO/Communication.hpp:128-444 (listBoundary), 556-614   */
/* ------------------------------------------------------------------------ */

/* For one halo cell: dims at a boundary-condition position (-3 or N+2) get
 * the truncated sum sum_{j=1..MMAX} alpha^j f(r + sign*j) with
 * r = (N + pos - sign) % N (VMO/Communication.hpp:193-196), the others the
 * periodic image (N + pos) % N (VMO/Communication.hpp:527-530).  A cell with
 * m boundary dims ends up with the m-fold sum of pass m-1; the innermost loop
 * runs over the lowest boundary dim, matching the loop nests at
 * VMO/Communication.hpp:776-783, 806-816, 839-853, 876-894. */
static double halo_value(const vo_config *c, const double *fn, const int pos[4]) {
  const double alpha = sqrt(3) - 2;
  int bdim[4], nb = 0;
  int r[4], sg[4];
  for (int d = 0; d < 4; d++) {
    const int N = c->nxmax[d];
    sg[d] = 0;
    if (pos[d] == -VO_HALO) sg[d] = -1;
    else if (pos[d] == N - 1 + VO_HALO) sg[d] = 1;
    if (sg[d] != 0) bdim[nb++] = d;
    r[d] = (N + pos[d] - sg[d]) % N;
  }
  if (nb == 0) return F4(c, fn, r[0], r[1], r[2], r[3]);
  int p[4];
  double fsum = 0.;
  if (nb == 1) {
    const int d1 = bdim[0];
    double alphap1 = alpha;
    for (int j1 = 1; j1 <= VO_MMAX; j1++) {
      memcpy(p, r, sizeof p);
      p[d1] += sg[d1] * j1;
      fsum += F4(c, fn, p[0], p[1], p[2], p[3]) * alphap1;
      alphap1 *= alpha;
    }
  } else if (nb == 2) {
    const int d1 = bdim[0], d2 = bdim[1];
    double alphap2 = alpha;
    for (int j2 = 1; j2 <= VO_MMAX; j2++) {
      double alphap1 = alpha * alphap2;
      for (int j1 = 1; j1 <= VO_MMAX; j1++) {
        memcpy(p, r, sizeof p);
        p[d1] += sg[d1] * j1;
        p[d2] += sg[d2] * j2;
        fsum += F4(c, fn, p[0], p[1], p[2], p[3]) * alphap1;
        alphap1 *= alpha;
      }
      alphap2 *= alpha;
    }
  } else if (nb == 3) {
    const int d1 = bdim[0], d2 = bdim[1], d3 = bdim[2];
    double alphap3 = alpha;
    for (int j3 = 1; j3 <= VO_MMAX; j3++) {
      double alphap2 = alpha * alphap3;
      for (int j2 = 1; j2 <= VO_MMAX; j2++) {
        double alphap1 = alpha * alphap2;
        for (int j1 = 1; j1 <= VO_MMAX; j1++) {
          memcpy(p, r, sizeof p);
          p[d1] += sg[d1] * j1;
          p[d2] += sg[d2] * j2;
          p[d3] += sg[d3] * j3;
          fsum += F4(c, fn, p[0], p[1], p[2], p[3]) * alphap1;
          alphap1 *= alpha;
        }
        alphap2 *= alpha;
      }
      alphap3 *= alpha;
    }
  } else {
    double alphap4 = alpha;
    for (int j4 = 1; j4 <= VO_MMAX; j4++) {
      double alphap3 = alpha * alphap4;
      for (int j3 = 1; j3 <= VO_MMAX; j3++) {
        double alphap2 = alpha * alphap3;
        for (int j2 = 1; j2 <= VO_MMAX; j2++) {
          double alphap1 = alpha * alphap2;
          for (int j1 = 1; j1 <= VO_MMAX; j1++) {
            fsum += F4(c, fn, r[0] + sg[0] * j1, r[1] + sg[1] * j2, r[2] + sg[2] * j3, r[3] + sg[3] * j4) * alphap1;
            alphap1 *= alpha;
          }
          alphap2 *= alpha;
        }
        alphap3 *= alpha;
      }
      alphap4 *= alpha;
    }
  }
  return fsum;
}

/* Fill every halo cell of fn from its interior (VMO/Communication.cpp:541-570).
 * All values depend on the interior only, so the order of writes is free. */
void vo_exchange_halo(const vo_config *c, double *fn) {
  const int n0 = c->nxmax[0], n1 = c->nxmax[1], n2 = c->nxmax[2], n3 = c->nxmax[3];
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int ivy = -VO_HALO; ivy < n3 + VO_HALO; ivy++)
    for (int ivx = -VO_HALO; ivx < n2 + VO_HALO; ivx++)
      for (int iy = -VO_HALO; iy < n1 + VO_HALO; iy++) {
        const int vin = (ivy >= 0 && ivy < n3 && ivx >= 0 && ivx < n2 && iy >= 0 && iy < n1);
        for (int ix = -VO_HALO; ix < n0 + VO_HALO; ix++) {
          if (vin && ix >= 0 && ix < n0) continue;
          const int pos[4] = {ix, iy, ivx, ivy};
          F4(c, fn, ix, iy, ivx, ivy) = halo_value(c, fn, pos);
        }
      }
}

/* ------------------------------------------------------------------------ */
/* A5: spline prefilter, VMO/Spline.hpp:114-199 (LONG_ENOUGH_BUFFER branch)  */
/* ------------------------------------------------------------------------ */

/* One line in place.  base[k*stride] holds position k-3, k = 0..N+5.
 * Positions -3 and N+2 carry the precomputed alpha-sums and are not written;
 * positions -2..N+1 receive the coefficients. */
static void spline_line_inplace(double *base, ptrdiff_t stride, int N, double *tmp1d) {
  const double sqrt3 = sqrt(3);
  const double alpha = sqrt3 - 2;
  const double beta = sqrt3 * (1 - alpha * alpha);
  const int istart = 1, iend = N + 4; /* positions -2 and N+1 */
  tmp1d[0] = (base[(istart - 1) * stride] + base[istart * stride]);                   /* :132 */
  for (int nn = 1; istart + nn <= iend; nn++)                                         /* :133-135 */
    tmp1d[nn] = base[(istart + nn) * stride] + alpha * tmp1d[nn - 1];
  double fnend = (base[(iend + 1) * stride] + base[iend * stride]);                   /* :144 */
  double alpha_k = alpha;
  for (int nn = 1; istart <= iend - nn; nn++) {                                       /* :146-149 */
    fnend += base[(iend - nn) * stride] * alpha_k;
    alpha_k *= alpha;
  }
  base[iend * stride] = fnend * sqrt3;                                                /* :151 */
  for (int nn = iend - 1; nn >= istart; nn--)                                         /* :153-155 */
    base[nn * stride] = beta * tmp1d[nn - istart] + alpha * base[(nn + 1) * stride];
}

/* Planes over dims (a, b) of a 4-D halo array: rows along a for every b
 * position -3..N_b+2, then columns along b for a positions -2..N_a+1
 * (VMO/Spline.hpp:119-198).  Planes cover the full halo range of the other
 * two dims. */
static void spline_planes(const vo_config *c, double *fn, int a, int b) {
  size_t st[4];
  st[0] = 1;
  for (int d = 1; d < 4; d++) st[d] = st[d - 1] * ext(c, d - 1);
  int o[2], k = 0;
  for (int d = 0; d < 4; d++)
    if (d != a && d != b) o[k++] = d;
  const int Na = c->nxmax[a], Nb = c->nxmax[b];
  const int Eo0 = (int)ext(c, o[0]), Eo1 = (int)ext(c, o[1]);
#pragma omp parallel
  {
    double tmp1d[256 + 8];
#pragma omp for collapse(2)
    for (int p1 = 0; p1 < Eo1; p1++)
      for (int p0 = 0; p0 < Eo0; p0++) {
        double *plane = fn + (size_t)p0 * st[o[0]] + (size_t)p1 * st[o[1]];
        for (int jb = 0; jb < Nb + 6; jb++)       /* rows: b = -3..N_b+2 */
          spline_line_inplace(plane + (size_t)jb * st[b], (ptrdiff_t)st[a], Na, tmp1d);
        for (int ja = 1; ja <= Na + 4; ja++)      /* cols: a = -2..N_a+1 */
          spline_line_inplace(plane + (size_t)ja * st[a], (ptrdiff_t)st[b], Nb, tmp1d);
      }
  }
}

void vo_spline_xy(const vo_config *c, double *fn) { spline_planes(c, fn, 0, 1); } /* VMO/Spline.hpp:384-426 */

/* VMO/Spline.hpp:429-469: the reference transposes (x*y) x (vx*vy) with a
 * blocked 16x16 copy (lib/OpenMP_Transpose.hpp:49-62), runs the same plane
 * routine along vx then vy, and transposes back.  The transpose only moves
 * data, so the result equals the plane routine applied in place; the oracle
 * still transposes so that the CPU baseline pays the reference's traffic. */
void vo_spline_vxvy(const vo_config *c, double *fn) {
  const size_t rows = ext(c, 0) * ext(c, 1), cols = ext(c, 2) * ext(c, 3);
  double *tr = (double *)malloc(rows * cols * sizeof(double));
  if (!tr) { spline_planes(c, fn, 2, 3); return; }
  const size_t bs = 16;
#pragma omp parallel for collapse(2)
  for (size_t j = 0; j < cols; j += bs)
    for (size_t i = 0; i < rows; i += bs)
      for (size_t cc = j; cc < j + bs && cc < cols; cc++)
        for (size_t r = i; r < i + bs && r < rows; r++) tr[cc + cols * r] = fn[r + rows * cc];
  vo_config t = *c;
  t.nxmax[0] = c->nxmax[2]; t.nxmax[1] = c->nxmax[3];
  t.nxmax[2] = c->nxmax[0]; t.nxmax[3] = c->nxmax[1];
  spline_planes(&t, tr, 0, 1);
#pragma omp parallel for collapse(2)
  for (size_t j = 0; j < rows; j += bs)
    for (size_t i = 0; i < cols; i += bs)
      for (size_t cc = j; cc < j + bs && cc < rows; cc++)
        for (size_t r = i; r < i + bs && r < cols; r++) fn[cc + rows * r] = tr[r + cols * cc];
  free(tr);
}

/* 1-D operator M = R*E on one periodic line: n values -> n+4 coefficients on
 * positions -2..n+1 (extension E as in halo_value, recursion R as above). */
int vo_spline_line(int n, const double *f, double *c) {
  if (n < VO_MMAX + 2 || n + 5 > 256) return VO_E_SIZE;
  const double alpha = sqrt(3) - 2;
  double g[256 + 8], tmp1d[256 + 8];
  for (int k = 1; k <= n + 4; k++) g[k] = f[((k - 3) % n + n) % n];
  double s = 0., ap = alpha;
  for (int j = 1; j <= VO_MMAX; j++) { s += f[n - 2 - j] * ap; ap *= alpha; }
  g[0] = s;
  s = 0.; ap = alpha;
  for (int j = 1; j <= VO_MMAX; j++) { s += f[1 + j] * ap; ap *= alpha; }
  g[n + 5] = s;
  spline_line_inplace(g, 1, n, tmp1d);
  memcpy(c, g + 1, (size_t)(n + 4) * sizeof(double));
  return VO_OK;
}

/* ------------------------------------------------------------------------ */
/* A6: 2-D half-step advection, VMO/Advection.hpp:87-164 (spline branch      */
/*     :126-163) and :299-406                                                */
/* ------------------------------------------------------------------------ */

static inline int interp_2d(const vo_config *c, const double *fn_tmp, const double xstar[2],
                            const double inv_dx[2], const double minPhy[2], const int xmin[2],
                            const int xmax[2], int ivx, int ivy, double *interp) {
  const double fx = (xstar[0] - minPhy[0]) * inv_dx[0];
  const double fy = (xstar[1] - minPhy[1]) * inv_dx[1];
  const int ix = (int)floor(fx);
  const int iy = (int)floor(fy);
  const double wx = fx - (double)ix;
  const double wy = fy - (double)iy;
  const double etax3 = (1. / 6.) * wx * wx * wx;
  const double etax0 = (1. / 6.) + 0.5 * wx * (wx - 1.) - etax3;
  const double etax2 = wx + etax0 - 2. * etax3;
  const double etax1 = 1. - etax0 - etax2 - etax3;
  const double etax[4] = {etax0, etax1, etax2, etax3};
  const double etay3 = (1. / 6.) * wy * wy * wy;
  const double etay0 = (1. / 6.) + 0.5 * wy * (wy - 1.) - etay3;
  const double etay2 = wy + etay0 - 2. * etay3;
  const double etay1 = 1. - etay0 - etay2 - etay3;
  const double etay[4] = {etay0, etay1, etay2, etay3};
  if (ix < xmin[0] - 1 || ix > xmax[0]) return 1;
  if (iy < xmin[1] - 1 || iy > xmax[1]) return 1;
  double ftmp = 0.;
  for (int jy = 0; jy <= 3; jy++) {
    double sum = 0.;
    for (int jx = 0; jx <= 3; jx++) sum += etax[jx] * F4(c, fn_tmp, ix - 1 + jx, iy - 1 + jy, ivx, ivy);
    ftmp += etay[jy] * sum;
  }
  *interp = ftmp;
  return 0;
}

int vo_advect_2d(const vo_config *c, double *fn, double *fn_tmp, double dt) {
  const double minPhyx = c->minPhy[0], minPhyy = c->minPhy[1];
  const double minPhyvx = c->minPhy[2], minPhyvy = c->minPhy[3];
  const double dx = c->dx[0], dy = c->dx[1], dvx = c->dx[2], dvy = c->dx[3];
  const double inv_dx[2] = {1. / dx, 1. / dy};
  const double minPhy[2] = {c->minPhy[0], c->minPhy[1]};
  const int xmin[2] = {0, 0};
  const int xmax[2] = {c->nxmax[0] - 1, c->nxmax[1] - 1};
  const size_t n = vo_halo_size(c);
  memcpy(fn_tmp, fn, n * sizeof(double));                   /* deep_copy, :331-332 */
  long err = 0;
#pragma omp parallel for collapse(2) reduction(+ : err)
  for (int ivy = 0; ivy < c->nxmax[3]; ivy++)
    for (int ivx = 0; ivx < c->nxmax[2]; ivx++)
      for (int iy = 0; iy < c->nxmax[1]; iy++)
        for (int ix = 0; ix < c->nxmax[0]; ix++) {
          const double x = minPhyx + ix * dx;
          const double y = minPhyy + iy * dy;
          const double vx = minPhyvx + ivx * dvx;
          const double vy = minPhyvy + ivy * dvy;
          const double depx = dt * vx;
          const double depy = dt * vy;
          const double xstar[2] = {x - depx, y - depy};
          double ftmp = 0;
          err += interp_2d(c, fn_tmp, xstar, inv_dx, minPhy, xmin, xmax, ivx, ivy, &ftmp);
          F4(c, fn, ix, iy, ivx, ivy) = ftmp;
        }
  return err ? VO_E_CFL : VO_OK;                              /* testError, :403-405 */
}

/* ------------------------------------------------------------------------ */
/* A11/A12: feet and 4-D advection, VMO/Advection.hpp:166-296, 409-498       */
/* ------------------------------------------------------------------------ */

static inline void lag3_basis(double posx, double coef[3]) {  /* :26-30 */
  coef[0] = .5 * (posx - 1.) * (posx - 2.);
  coef[1] = -1. * (posx) * (posx - 2.);
  coef[2] = .5 * (posx) * (posx - 1.);
}

static inline double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */
static inline double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */

static inline void compute_feet(double xstar[4], const double *ex, const double *ey, const double xmin[4],
                                const double rxwidth[4], const double dx[4], const double inv_dx[4],
                                const int nxmax[4], const int pos[4], double dt) {
  const int s_nxmax = nxmax[0], s_nymax = nxmax[1];
  const double x = xmin[0] + pos[0] * dx[0];
  const double y = xmin[1] + pos[1] * dx[1];
  const double vx = xmin[2] + pos[2] * dx[2];
  const double vy = xmin[3] + pos[3] * dx[3];
  double xtmp[4];
  xtmp[0] = x - 0.5 * dt * vx;
  xtmp[1] = y - 0.5 * dt * vy;
  double ftmp1 = 0., ftmp2 = 0.;
  int ipos[2];
  double coefx[2][3];
  for (int j = 0; j <= 1; j++) {                              /* :189-194 */
    xtmp[j] = inv_dx[j] * (xtmp[j] - xmin[j]);
    ipos[j] = (int)(round(xtmp[j]) - 1);
    const double posx = xtmp[j] - ipos[j];
    lag3_basis(posx, coefx[j]);
  }
  for (int k1 = 0; k1 <= 2; k1++)                            /* :199-206 */
    for (int k0 = 0; k0 <= 2; k0++) {
      const int ix = (s_nxmax + ipos[0] + k0) % s_nxmax;
      const int iy = (s_nymax + ipos[1] + k1) % s_nymax;
      ftmp1 += coefx[0][k0] * coefx[1][k1] * ex[ix + (size_t)s_nxmax * iy];
      ftmp2 += coefx[0][k0] * coefx[1][k1] * ey[ix + (size_t)s_nxmax * iy];
    }
  xtmp[2] = vx - 0.5 * dt * ftmp1;                            /* :208-211 */
  xtmp[3] = vy - 0.5 * dt * ftmp2;
  xstar[0] = x - dt * xtmp[2];                                /* :214-217 */
  xstar[1] = y - dt * xtmp[3];
  xstar[2] = dmax(dmin(vx - dt * ftmp1, xmin[2] + rxwidth[2]), xmin[2]);
  xstar[3] = dmax(dmin(vy - dt * ftmp2, xmin[3] + rxwidth[3]), xmin[3]);
}

static inline double interp_4d(const vo_config *c, const double *tmp_fn, const double xmin[4],
                               const double inv_dx[4], double xstar[4]) {  /* :221-296 */
  int ipos[4];
  double eta[4][4];
  for (int j = 0; j < 4; j++) xstar[j] = inv_dx[j] * (xstar[j] - xmin[j]);
  for (int j = 0; j < 4; j++) {
    ipos[j] = (int)floor(xstar[j]);
    const double wx = xstar[j] - ipos[j];
    const double etax3 = (1. / 6.) * wx * wx * wx;
    const double etax0 = (1. / 6.) + 0.5 * wx * (wx - 1.) - etax3;
    const double etax2 = wx + etax0 - 2. * etax3;
    const double etax1 = 1. - etax0 - etax2 - etax3;
    eta[j][0] = etax0; eta[j][1] = etax1; eta[j][2] = etax2; eta[j][3] = etax3;
  }
  /* same loop nest and rounding as the reference, with hoisted strides */
  const size_t s1 = ext(c, 0), s2 = s1 * ext(c, 1), s3 = s2 * ext(c, 2);
  const double *base = tmp_fn + vo_index(c, ipos[0] - 1, ipos[1] - 1, ipos[2] - 1, ipos[3] - 1);
  double ftmp1 = 0.;
  for (int k3 = 0; k3 <= 3; k3++)
    for (int k2 = 0; k2 <= 3; k2++) {
      double ftmp2 = 0.;
      const double *p23 = base + k2 * s2 + k3 * s3;
      for (int k1 = 0; k1 <= 3; k1++) {
        const double *p = p23 + k1 * s1;
        double ftmp3 = 0.;
        for (int k0 = 0; k0 <= 3; k0++) ftmp3 += eta[0][k0] * p[k0];
        ftmp2 += ftmp3 * eta[1][k1];
      }
      ftmp1 += ftmp2 * eta[2][k2] * eta[3][k3];
    }
  return ftmp1;
}

/* Foot of one interior cell (computeFeet with advect_4D's constants), exported
 * for tests. */
void vo_feet(const vo_config *c, const double *ex, const double *ey, int ix, int iy, int ivx, int ivy, double dt,
             double xstar[4]) {
  double rxmin[4], inv_dx[4], rxwidth[4], dx[4];
  int xmax[4];
  for (int j = 0; j < 4; j++) {
    xmax[j] = c->nxmax[j];
    rxmin[j] = c->minPhy[j];
    rxwidth[j] = c->maxPhy[j] - c->minPhy[j];
    inv_dx[j] = 1. / c->dx[j];
    dx[j] = c->dx[j];
  }
  const int pos[4] = {ix, iy, ivx, ivy};
  compute_feet(xstar, ex, ey, rxmin, rxwidth, dx, inv_dx, xmax, pos, dt);
}

/* fn holds C4 on entry; on exit its interior is f^{n+1}; tmp_fn receives a copy
 * of C4 (the reference's argument order, VMO/Timestep.hpp:74). */
int vo_advect_4d(const vo_config *c, const double *ex, const double *ey, double *fn, double *tmp_fn, double dt) {
  double rxmin[4], inv_dx[4], rxwidth[4], dx[4], locrxmindx[4], locrxmaxdx[4];
  int xmax[4];
  for (int j = 0; j < 4; j++) {                              /* :426-434 */
    xmax[j] = c->nxmax[j];
    rxmin[j] = c->minPhy[j];
    rxwidth[j] = c->maxPhy[j] - c->minPhy[j];
    inv_dx[j] = 1. / c->dx[j];
    dx[j] = c->dx[j];
    locrxmindx[j] = c->minPhy[j] + 0 * c->dx[j] - c->dx[j];
    locrxmaxdx[j] = c->minPhy[j] + (c->nxmax[j] - 1) * c->dx[j] + c->dx[j];
  }
  const size_t n = vo_halo_size(c);
  memcpy(tmp_fn, fn, n * sizeof(double));                   /* deep_copy, :437 */
  long err = 0;
#pragma omp parallel for collapse(2) reduction(+ : err)
  for (int ivy = 0; ivy < c->nxmax[3]; ivy++)
    for (int ivx = 0; ivx < c->nxmax[2]; ivx++)
      for (int iy = 0; iy < c->nxmax[1]; iy++)
        for (int ix = 0; ix < c->nxmax[0]; ix++) {
          double xstar[4];
          const int indices[4] = {ix, iy, ivx, ivy};
          compute_feet(xstar, ex, ey, rxmin, rxwidth, dx, inv_dx, xmax, indices, dt);
          int bad = 0;
          for (int j = 0; j < 4; j++) bad += (xstar[j] < locrxmindx[j] || xstar[j] > locrxmaxdx[j]);
          err += bad;
          /* The reference interpolates even at bad feet (out-of-bounds reads);
           * the oracle writes 0 there and reports the error. */
          F4(c, fn, ix, iy, ivx, ivy) = bad ? 0. : interp_4d(c, tmp_fn, rxmin, inv_dx, xstar);
        }
  return err ? VO_E_CFL : VO_OK;
}

/* ------------------------------------------------------------------------ */
/* A7/A8: density, VMO/Field.cpp:5-69 (one rank: allreduce is a copy)        */
/* ------------------------------------------------------------------------ */

void vo_field_rho(const vo_config *c, const double *fn, double *rho) {
  const double dvx = c->dx[2], dvy = c->dx[3];
#pragma omp parallel for
  for (int iy = 0; iy < c->nxmax[1]; iy++)
    for (int ix = 0; ix < c->nxmax[0]; ix++) {
      double sum = 0.;
      for (int ivy = 0; ivy < c->nxmax[3]; ivy++)
        for (int ivx = 0; ivx < c->nxmax[2]; ivx++) sum += F4(c, fn, ix, iy, ivx, ivy);
      rho[ix + (size_t)c->nxmax[0] * iy] = sum * dvx * dvy;
    }
}

/* ------------------------------------------------------------------------ */
/* A9: Poisson, VMO/Field.cpp:71-137, VMO/Efield.cpp:3-106,                  */
/*     lib/OpenMP_FFT.hpp:106-142, 271-305, 356-389                          */
/* ------------------------------------------------------------------------ */

/* exp(-2 pi i m / n) with exact values on the axes. */
static void twiddle(int n, int m, double *re, double *im) {
  m = ((m % n) + n) % n;
  if (m == 0) { *re = 1.0; *im = 0.0; return; }
  if (2 * m == n) { *re = -1.0; *im = 0.0; return; }
  if (4 * m == n) { *re = 0.0; *im = -1.0; return; }
  if (4 * m == 3 * n) { *re = 0.0; *im = 1.0; return; }
  const double ang = 2.0 * M_PI * (double)m / (double)n;
  *re = cos(ang);
  *im = -sin(ang);
}

/* forward c2c (sign -1) or backward (sign +1), unnormalised, like
 * fftw_plan_dft_1d with FFTW_FORWARD / FFTW_BACKWARD. */
static void dft_c2c(int n, const double *in, double *out, int sign, double *tw) {
  for (int k = 0; k < n; k++) {
    double sr = 0., si = 0.;
    for (int j = 0; j < n; j++) {
      const int m = (int)(((long)k * j) % n);
      const double wr = tw[2 * m], wi = sign < 0 ? tw[2 * m + 1] : -tw[2 * m + 1];
      sr += in[2 * j] * wr - in[2 * j + 1] * wi;
      si += in[2 * j] * wi + in[2 * j + 1] * wr;
    }
    out[2 * k] = sr;
    out[2 * k + 1] = si;
  }
}

/* r2c of length n -> n/2+1 bins (fftw_plan_dft_r2c_1d). */
static void dft_r2c(int n, const double *in, double *out, double *tw) {
  for (int k = 0; k <= n / 2; k++) {
    double sr = 0., si = 0.;
    for (int j = 0; j < n; j++) {
      const int m = (int)(((long)k * j) % n);
      sr += in[j] * tw[2 * m];
      si += in[j] * tw[2 * m + 1];
    }
    out[2 * k] = sr;
    out[2 * k + 1] = si;
  }
}

/* c2r of n/2+1 bins -> n reals (fftw_plan_dft_c2r_1d): Hermitian extension;
 * the imaginary parts of the DC and (even n) Nyquist bins are ignored. */
static void dft_c2r(int n, const double *in, double *out, double *tw) {
  const int nh = n / 2;
  for (int j = 0; j < n; j++) {
    double s = in[0];
    for (int k = 1; k < (n + 1) / 2; k++) {
      const int m = (int)(((long)k * j) % n);
      const double wr = tw[2 * m], wi = -tw[2 * m + 1]; /* exp(+2 pi i m/n) */
      s += 2.0 * (in[2 * k] * wr - in[2 * k + 1] * wi);
    }
    if (n % 2 == 0) s += in[2 * nh] * ((j % 2) ? -1.0 : 1.0);
    out[j] = s;
  }
}

static double *make_twiddles(int n) {
  double *tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  for (int m = 0; m < n; m++) twiddle(n, m, &tw[2 * m], &tw[2 * m + 1]);
  return tw;
}

/* rfft2_serial_left: r2c along x per row, transpose, c2c along y. */
void vo_rfft2(int nx, int ny, const double *in, double *out_c) {
  const int nxh = nx / 2 + 1;
  double *twx = make_twiddles(nx), *twy = make_twiddles(ny);
  double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)nxh * ny);
  double *line = (double *)malloc(sizeof(double) * 2 * (size_t)(nx > ny ? nx : ny) + 16);
  double *line2 = (double *)malloc(sizeof(double) * 2 * (size_t)(nx > ny ? nx : ny) + 16);
  for (int iy = 0; iy < ny; iy++) {
    dft_r2c(nx, in + (size_t)nx * iy, line, twx);
    for (int k = 0; k < nxh; k++) {                          /* buffer (ny, nxh) */
      buf[2 * ((size_t)iy + (size_t)ny * k)] = line[2 * k];
      buf[2 * ((size_t)iy + (size_t)ny * k) + 1] = line[2 * k + 1];
    }
  }
  for (int k = 0; k < nxh; k++) {
    dft_c2c(ny, buf + 2 * (size_t)ny * k, line2, -1, twy);
    for (int iy = 0; iy < ny; iy++) {
      out_c[2 * ((size_t)k + (size_t)nxh * iy)] = line2[2 * iy];
      out_c[2 * ((size_t)k + (size_t)nxh * iy) + 1] = line2[2 * iy + 1];
    }
  }
  free(line); free(line2); free(buf); free(twx); free(twy);
}

/* irfft2_serial_left: c2c backward along y per kx, c2r along x per row. */
void vo_irfft2(int nx, int ny, const double *in_c, double *out) {
  const int nxh = nx / 2 + 1;
  double *twx = make_twiddles(nx), *twy = make_twiddles(ny);
  double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)nxh * ny);
  double *line = (double *)malloc(sizeof(double) * 2 * (size_t)(nx > ny ? nx : ny) + 16);
  double *line2 = (double *)malloc(sizeof(double) * 2 * (size_t)(nx > ny ? nx : ny) + 16);
  for (int k = 0; k < nxh; k++) {
    for (int iy = 0; iy < ny; iy++) {
      line[2 * iy] = in_c[2 * ((size_t)k + (size_t)nxh * iy)];
      line[2 * iy + 1] = in_c[2 * ((size_t)k + (size_t)nxh * iy) + 1];
    }
    dft_c2c(ny, line, line2, +1, twy);
    for (int iy = 0; iy < ny; iy++) {                        /* buffer (nxh, ny) */
      buf[2 * ((size_t)k + (size_t)nxh * iy)] = line2[2 * iy];
      buf[2 * ((size_t)k + (size_t)nxh * iy) + 1] = line2[2 * iy + 1];
    }
  }
  for (int iy = 0; iy < ny; iy++) dft_c2r(nx, buf + 2 * (size_t)nxh * iy, out + (size_t)nx * iy, twx);
  free(line); free(line2); free(buf); free(twx); free(twy);
}

/* solve_poisson_fftw, VMO/Efield.cpp:54-106, with the filter of :38-43.
 * rho is overwritten with phi. */
static void solve_poisson(const vo_config *c, double *rho, double *ex, double *ey) {
  const int nx = c->nxmax[0], ny = c->nxmax[1];
  const int nxh = nx / 2 + 1, nyh = ny / 2 + 1;
  const double kx0 = 2 * M_PI / c->maxPhy[0];
  const double ky0 = 2 * M_PI / c->maxPhy[1];
  const double normcoeff = 1.0 / (double)(nx * ny);
  const size_t nc = (size_t)nxh * ny;
  double *rh = (double *)malloc(sizeof(double) * 2 * nc);
  double *eh = (double *)malloc(sizeof(double) * 2 * nc);
  double *fh = (double *)malloc(sizeof(double) * 2 * nc);
  double *filter = (double *)malloc(sizeof(double) * (size_t)nxh);
  filter[0] = 0.;                                           /* VMO/Efield.cpp:38-43 */
  for (int ix = 1; ix < nxh; ix++) {
    const double kx = ix * (2 * M_PI / c->maxPhy[0]);
    const double k2 = kx * kx;
    filter[ix] = 1. / k2;
  }
  vo_rfft2(nx, ny, rho, rh);
  for (int ix1 = 0; ix1 < nxh; ix1++) {
    const double kx = ix1 * kx0;
    for (int ix2 = 0; ix2 < ny; ix2++) {
      const size_t id = 2 * ((size_t)ix1 + (size_t)nxh * ix2);
      const double cr = rh[id], ci = rh[id + 1];
      if (ix2 == 0) {                                       /* :76-81 */
        /* (-kx*I)*rho_hat = (kx*ci, -kx*cr), then *filter, *normcoeff */
        eh[id] = ((kx * ci) * filter[ix1]) * normcoeff;
        eh[id + 1] = ((-kx * cr) * filter[ix1]) * normcoeff;
        fh[id] = 0.; fh[id + 1] = 0.;
        rh[id] = (cr * filter[ix1]) * normcoeff;
        rh[id + 1] = (ci * filter[ix1]) * normcoeff;
      } else {                                              /* :83-100 */
        const double ky = (ix2 < nyh) ? ix2 * ky0 : (ix2 - ny) * ky0;
        const double k2 = kx * kx + ky * ky;
        const double sx = kx / k2, sy = ky / k2;
        eh[id] = (sx * ci) * normcoeff;
        eh[id + 1] = (-sx * cr) * normcoeff;
        fh[id] = (sy * ci) * normcoeff;
        fh[id + 1] = (-sy * cr) * normcoeff;
        rh[id] = (cr / k2) * normcoeff;
        rh[id + 1] = (ci / k2) * normcoeff;
      }
    }
  }
  vo_irfft2(nx, ny, rh, rho);
  vo_irfft2(nx, ny, eh, ex);
  vo_irfft2(nx, ny, fh, ey);
  free(rh); free(eh); free(fh); free(filter);
}

void vo_field_poisson(const vo_config *c, double *rho, double *ex, double *ey) {
  const int nx = c->nxmax[0], ny = c->nxmax[1];
  switch (c->idcase) {
    case 2:                                                 /* VMO/Field.cpp:81-94 */
      for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++) {
          ex[ix + (size_t)nx * iy] = -(c->minPhy[0] + ix * c->dx[0]);
          ey[ix + (size_t)nx * iy] = 0.;
        }
      break;
    case 6:                                                 /* VMO/Field.cpp:95-108 */
      for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++) {
          ex[ix + (size_t)nx * iy] = -(c->minPhy[1] + iy * c->dx[1]);
          ey[ix + (size_t)nx * iy] = 0.;
        }
      break;
    case 10:
    case 20:                                                /* VMO/Field.cpp:110-126 */
      for (size_t i = 0; i < (size_t)nx * ny; i++) rho[i] -= 1.;
      solve_poisson(c, rho, ex, ey);
      break;
    default:
      solve_poisson(c, rho, ex, ey);
      break;
  }
}

/* ------------------------------------------------------------------------ */
/* A10/A13: diagnostics, VMO/Diags.cpp:16-76                                 */
/* ------------------------------------------------------------------------ */

void vo_diags_field(const vo_config *c, const double *rho, const double *ex, const double *ey, double out[4]) {
  const int nx = c->nxmax[0], ny = c->nxmax[1];
  double iter_mass = 0., it_nrj = 0., it_nrjx = 0., it_nrjy = 0.;
  for (int iy = 0; iy < ny; iy++)
    for (int ix = 0; ix < nx; ix++) {
      const double eex = ex[ix + (size_t)nx * iy];
      const double eey = ey[ix + (size_t)nx * iy];
      iter_mass += rho[ix + (size_t)nx * iy];
      it_nrj += eex * eex + eey * eey;
      it_nrjx += eex * eex;
      it_nrjy += eey * eey;
    }
  it_nrj = sqrt(it_nrj * c->dx[0] * c->dx[1]);
  it_nrj = it_nrj > 1.e-30 ? log(it_nrj) : -1.e9;
  out[0] = it_nrj;
  out[1] = sqrt(0.5 * it_nrjx * c->dx[0] * c->dx[1]);
  out[2] = sqrt(0.5 * it_nrjy * c->dx[0] * c->dx[1]);
  out[3] = iter_mass * c->dx[0] * c->dx[1];
}

double vo_l2norm(const vo_config *c, const double *fn) {
  double l2loc = 0.;
#pragma omp parallel for collapse(3) reduction(+ : l2loc)
  for (int ivy = 0; ivy < c->nxmax[3]; ivy++)
    for (int ivx = 0; ivx < c->nxmax[2]; ivx++)
      for (int iy = 0; iy < c->nxmax[1]; iy++)
        for (int ix = 0; ix < c->nxmax[0]; ix++) {
          const double v = F4(c, fn, ix, iy, ivx, ivy);
          l2loc += v * v;
        }
  return sqrt(l2loc * c->dx[0] * c->dx[1] * c->dx[2] * c->dx[3]);
}

/* ------------------------------------------------------------------------ */
/* A14: driver state, init and the time step                                 */
/* ------------------------------------------------------------------------ */

struct vo_state {
  vo_config cfg;
  double *fn, *fnp1, *tmp;
  double *rho, *ex, *ey;
  double *diag[5];
  int iter;
  double tsec[VO_T_COUNT];
  int tcalls[VO_T_COUNT];
};

static inline void tadd(vo_state *st, int k, double t0) {
  st->tsec[k] += now_s() - t0;
  st->tcalls[k]++;
}

static void field_step(vo_state *st, const double *f, int iter, int with_l2) {
  double t_a = now_s();
  vo_field_rho(&st->cfg, f, st->rho);
  tadd(st, VO_T_FIELD, t_a);
  st->tcalls[VO_T_ALLREDUCE]++;                             /* one rank: a copy */
  double t_b = now_s();
  vo_field_poisson(&st->cfg, st->rho, st->ex, st->ey);
  tadd(st, VO_T_FOURIER, t_b);
  double t_c = now_s();
  double d[4];
  vo_diags_field(&st->cfg, st->rho, st->ex, st->ey, d);
  for (int k = 0; k < 4; k++) st->diag[k][iter] = d[k];
  if (with_l2) st->diag[4][iter] = vo_l2norm(&st->cfg, f);
  tadd(st, VO_T_DIAG, t_c);
}

/* ptest of VMO/Init.cpp:279-322: code every interior cell, exchange, check
 * the two inner halo layers against the periodic image. */
static int halo_selftest(vo_state *st) {
  const vo_config *c = &st->cfg;
  const int *N = c->nxmax;
  for (int ivy = 0; ivy < N[3]; ivy++)
    for (int ivx = 0; ivx < N[2]; ivx++)
      for (int iy = 0; iy < N[1]; iy++)
        for (int ix = 0; ix < N[0]; ix++)
          F4(c, st->fn, ix, iy, ivx, ivy) = (double)((ivy + 111 * ivx) * 111 + iy) * 111 + ix;
  vo_exchange_halo(c, st->fn);
  int bad = 0;
  const int off = VO_HALO - 1;
  for (int ivy = -off; ivy < N[3] + off; ivy++)
    for (int ivx = -off; ivx < N[2] + off; ivx++) {
      const int jvy = (N[3] + ivy) % N[3], jvx = (N[2] + ivx) % N[2];
      for (int iy = -off; iy < N[1] + off; iy++) {
        const int jy = (N[1] + iy) % N[1];
        for (int ix = -off; ix < N[0] + off; ix++) {
          const int jx = (N[0] + ix) % N[0];
          const double ref = (((double)jvy + 111. * (double)jvx) * 111. + (double)jy) * 111. + (double)jx;
          if (fabs(F4(c, st->fn, ix, iy, ivx, ivy) - ref) > .1) bad++;
        }
      }
    }
  return bad ? VO_E_ARG : VO_OK;
}

int vo_create(const vo_config *cfg, vo_state **out) {
  *out = NULL;
  int rc = vo_config_check(cfg);
  if (rc) return rc;
  if (cfg->idcase != 2 && cfg->idcase != 5 && cfg->idcase != 10 && cfg->idcase != 20 && cfg->idcase != 30 &&
      cfg->idcase != 31)
    return VO_E_CASE;
  vo_state *st = (vo_state *)calloc(1, sizeof(vo_state));
  if (!st) return VO_E_ARG;
  st->cfg = *cfg;
  const size_t n = vo_halo_size(cfg);
  const size_t n2 = (size_t)cfg->nxmax[0] * cfg->nxmax[1];
  st->fn = (double *)calloc(n, sizeof(double));
  st->fnp1 = (double *)calloc(n, sizeof(double));
  st->tmp = (double *)calloc(n, sizeof(double));
  st->rho = (double *)calloc(n2, sizeof(double));
  st->ex = (double *)calloc(n2, sizeof(double));
  st->ey = (double *)calloc(n2, sizeof(double));
  for (int k = 0; k < 5; k++) st->diag[k] = (double *)calloc((size_t)cfg->nbiter + 1, sizeof(double));
  if (!st->fn || !st->fnp1 || !st->tmp || !st->rho || !st->ex || !st->ey || !st->diag[4]) {
    vo_destroy(st);
    return VO_E_ARG;
  }
  rc = halo_selftest(st);                                   /* VMO/Init.cpp:353-355 */
  if (rc) { vo_destroy(st); return rc; }
  memset(st->fn, 0, n * sizeof(double));                    /* VMO/Init.cpp:370-371 */
  rc = vo_init_case(cfg, st->fn);
  if (rc) { vo_destroy(st); return rc; }
  double t_t = now_s();
  field_step(st, st->fn, 0, 1);                             /* VMO/Vlp4d.cpp:30-34 */
  st->tsec[VO_T_TOTAL] += now_s() - t_t;
  st->iter = 0;
  *out = st;
  return VO_OK;
}

/* onetimestep (VMO/Timestep.hpp:34-97) followed by fn.swap(fnp1). */
int vo_step(vo_state *st) {
  const vo_config *c = &st->cfg;
  if (st->iter >= c->nbiter) return VO_E_ARG;
  const double t0 = now_s();
  const int iter = ++st->iter;
  int rc;
  {
    double t_p = now_s();
    vo_exchange_halo(c, st->fn);                            /* :41 */
    tadd(st, VO_T_PACK, t_p);
    st->tcalls[VO_T_COMM]++;
    st->tcalls[VO_T_UNPACK]++;
  }
  {
    double t_s = now_s();
    vo_spline_xy(c, st->fn);                                /* :43-45 */
    tadd(st, VO_T_SPLINE_XY, t_s);
  }
  memcpy(st->fnp1, st->fn, vo_halo_size(c) * sizeof(double)); /* :47 */
  {
    double t_a = now_s();
    rc = vo_advect_2d(c, st->fn, st->tmp, 0.5 * c->dt);      /* :49-51 */
    tadd(st, VO_T_ADVEC2D, t_a);
    if (rc) return rc;
  }
  /* :53-67; the half-step diagnostics are overwritten at :89-91 */
  field_step(st, st->fn, iter, 0);
  {
    double t_v = now_s();
    vo_spline_vxvy(c, st->fnp1);                            /* :69-71 */
    tadd(st, VO_T_SPLINE_VXVY, t_v);
  }
  {
    double t_b = now_s();
    rc = vo_advect_4d(c, st->ex, st->ey, st->fnp1, st->fn, c->dt); /* :73-75 */
    tadd(st, VO_T_ADVEC4D, t_b);
    if (rc) return rc;
  }
  field_step(st, st->fnp1, iter, 1);                        /* :77-96 */
  double *t = st->fn; st->fn = st->fnp1; st->fnp1 = t;      /* VMO/Vlp4d.cpp:145 */
  const double dtm = now_s() - t0;
  st->tsec[VO_T_MAINLOOP] += dtm; st->tcalls[VO_T_MAINLOOP]++;
  st->tsec[VO_T_TOTAL] += dtm;
  return VO_OK;
}

int vo_run(vo_state *st, int nsteps) {
  for (int i = 0; i < nsteps; i++) {
    const int rc = vo_step(st);
    if (rc) return rc;
  }
  return VO_OK;
}

void vo_destroy(vo_state *st) {
  if (!st) return;
  free(st->fn); free(st->fnp1); free(st->tmp);
  free(st->rho); free(st->ex); free(st->ey);
  for (int k = 0; k < 5; k++) free(st->diag[k]);
  free(st);
}

int vo_iter(const vo_state *st) { return st->iter; }
const vo_config *vo_get_config(const vo_state *st) { return &st->cfg; }
double *vo_fn(vo_state *st) { return st->fn; }
double *vo_field(vo_state *st, int which) { return which == 0 ? st->rho : which == 1 ? st->ex : st->ey; }
const double *vo_diag_series(const vo_state *st, int which) { return st->diag[which]; }
double vo_timer_seconds(const vo_state *st, int which) { return st->tsec[which]; }
int vo_timer_calls(const vo_state *st, int which) { return st->tcalls[which]; }

/* Diags::save, VMO/Diags.cpp:78-104 */
int vo_save(const vo_state *st, const char *dir) {
  char path[4096];
  const vo_config *c = &st->cfg;
  snprintf(path, sizeof path, "%s/nrj.out", dir);
  FILE *f = fopen(path, "w");
  if (!f) return VO_E_FILE;
  for (int it = 0; it <= c->nbiter; ++it)
    fprintf(f, "%17.13e %17.13e %17.13e %17.13e %17.13e\n", it * c->dt, st->diag[0][it], st->diag[1][it],
            st->diag[2][it], st->diag[3][it]);
  fclose(f);
  snprintf(path, sizeof path, "%s/l2norm.out", dir);
  f = fopen(path, "w");
  if (!f) return VO_E_FILE;
  for (int it = 0; it <= c->nbiter; ++it) fprintf(f, "%17.13e %17.13e\n", it * c->dt, st->diag[4][it]);
  fclose(f);
  return VO_OK;
}

void vo_get_interior(const vo_config *c, const double *fn, double *out) {
  const size_t n0 = c->nxmax[0], n1 = c->nxmax[1], n2 = c->nxmax[2];
#pragma omp parallel for collapse(2)
  for (int ivy = 0; ivy < c->nxmax[3]; ivy++)
    for (int ivx = 0; ivx < c->nxmax[2]; ivx++)
      for (int iy = 0; iy < c->nxmax[1]; iy++)
        for (int ix = 0; ix < c->nxmax[0]; ix++)
          out[ix + n0 * (iy + n1 * (ivx + n2 * (size_t)ivy))] = F4(c, fn, ix, iy, ivx, ivy);
}

void vo_set_interior(const vo_config *c, double *fn, const double *in) {
  const size_t n0 = c->nxmax[0], n1 = c->nxmax[1], n2 = c->nxmax[2];
#pragma omp parallel for collapse(2)
  for (int ivy = 0; ivy < c->nxmax[3]; ivy++)
    for (int ivx = 0; ivx < c->nxmax[2]; ivx++)
      for (int iy = 0; iy < c->nxmax[1]; iy++)
        for (int ix = 0; ix < c->nxmax[0]; ix++)
          F4(c, fn, ix, iy, ivx, ivy) = in[ix + n0 * (iy + n1 * (ivx + n2 * (size_t)ivy))];
}
