// vlp4d_api.cpp -- C-ABI of libvlp4d.so: context, time step orchestration,
// deck parser, diagnostics and timers (host side).
//
// The step is the reference's onetimestep (miniapps/vlp4d_mpi/openmp/
// Timestep.hpp:34-97) re-planned for one GPU as streaming sweeps over HBM,
// each a hand-written kernel (spline.hip, advect.hip, field.hip):
//   1. spline x, spline y      f     -> CXY   (halo + alpha-sums computed in-kernel)
//   2. advect 2-D + rho        CXY   -> rho partials (f_half never stored)
//   3. Poisson + E             rho   -> E, phi
//   4. spline vx, spline vy    CXY   -> C4
//   5. advect 4-D + rho + L2   C4, E -> f^{n+1}, rho partials
//   6. Poisson + diagnostics
// The reference's deep copies, transposes and halo buffers have no counterpart.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vlp4d.h"
#include "vlp4d_internal.hpp"

using vlp::FieldBufs;
using vlp::Geom;

static thread_local std::string g_last_error;

struct vlp_ctx {
  vlp_config cfg{};
  Geom g{};
  int precision = VLP_FP64;
  int scheme = VLP_SCHEME_SPLINE;  // VLP_SCHEME_LAGRANGE: the split-step Lagrange solver (lagrange.hip)
  void *lag_tab = nullptr;          // stencil tables of its fused x/2, y/2 pass
  int device = 0;
  hipStream_t stream = nullptr;
  size_t esz = 8;
  void *f = nullptr, *fnext = nullptr, *cxy = nullptr, *tv = nullptr, *c4 = nullptr;
  void *c4_alloc = nullptr;  // allocation holding c4 (c4_shift elements in)
  size_t n_f = 0, n_cxy = 0, n_tv = 0, n_c4 = 0;
  size_t n_c4s = 0;  // stored C4 (layout of c4_pitch) plus a zeroed tail
  double *rho_part = nullptr, *l2_part = nullptr;
  long n_l2 = 0;
  int xy_groups = 0;        // G of the fused xy kernel (0: separate kernels)
  long rho_outer = 0;       // density-partial planes currently in rho_part
  bool half_rho = false;    // rho_part holds the half-step density of the last spline_xy
  FieldBufs fb{};
  int *d_err = nullptr;
  int *h_err = nullptr;
  int iter = 0;
  bool broken = false;
  std::string err;
  double tsec[VLP_T_COUNT] = {};
  int tcalls[VLP_T_COUNT] = {};
  hipEvent_t ev[8] = {};
  hipEvent_t evb[8] = {};  // second set: several steps in one call alternate between ev and evb
  int ys_tvx = 0, ys_nb = 0;  // 4-D advection tile chosen by vlp_autotune (0: preference order)
  int sp_mode[2] = {vlp::SPLINE_AUTO, vlp::SPLINE_AUTO};  // vx / vy spline kernels chosen by vlp_autotune
  vlp_ensemble *owner = nullptr;  // the ensemble whose captured graph holds this context's buffers
  // MFMA form of the vx / vy spline passes (VLP4D_SPLINE_MFMA=1 at creation):
  // the line operators of lengths N2 and N3
  void *mf_m[2] = {};
  // adjoint of one step (allocated on first use)
  double *adj_c4s = nullptr, *adj_blk = nullptr, *adj_tmp = nullptr, *adj_cxy = nullptr;
  double *adj_E = nullptr, *adj_rho = nullptr, *adj_k = nullptr, *adj_g = nullptr, *adj_grad = nullptr;
  double *adj_Mt[4] = {};  // M_d^T of the line spline operator, N_d x (N_d + 4)
  double *adj_M[4] = {};   // M_d itself, (N_d + 4) x N_d row-major (the tiled contraction's operand)
  double *adj_axy = nullptr, *adj_scl = nullptr;  // reproducible scatter: per-cell (ax, ay), fixed-point scale
};

namespace {

const char *kLabels[VLP_T_COUNT] = {"total",      "MainLoop", "pack",    "comm", "unpack",
                                    "advec2D",    "advec4D",  "field",   "all_reduce",
                                    "Fourier",    "diag",     "splinecoeff_xy", "splinecoeff_vxvy"};

int set_err(vlp_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  g_last_error = msg;
  return code;
}

#define HIP_TRY(c, expr)                                                                           \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return set_err((c), VLP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
  } while (0)

// Advance past the next ':' (VMO/Init.cpp:33-35); EOF is an error here
// where the reference would spin forever.
bool skip_to_colon(FILE *s) {
  int ch;
  do {
    ch = fgetc(s);
    if (ch == EOF) return false;
  } while (ch != ':');
  return true;
}

#pragma clang fp contract(off)
void derive_dx(vlp_config *c) {
  for (int i = 0; i < 4; i++) c->dx[i] = (c->maxPhy[i] - c->minPhy[i]) / (double)c->nxmax[i];  // VMO/Init.cpp:80-82
}
#pragma clang fp contract(on)

int check_config(const vlp_config *c) {
  for (int i = 0; i < 4; i++) {
    if (c->nxmax[i] < 10) return set_err(nullptr, VLP_E_SIZE, "dimension smaller than MMAX+2 = 10");
    if (c->nxmax[i] + 5 > 256) return set_err(nullptr, VLP_E_SIZE, "dimension larger than 251 (spline line buffer)");
  }
  if (c->nbiter < 0 || c->ifreq <= 0) return set_err(nullptr, VLP_E_ARG, "nbiter < 0 or ifreq <= 0");
  if (c->idcase != 2 && c->idcase != 5 && c->idcase != 10 && c->idcase != 20 && c->idcase != VLP_CASE_STRONG_LANDAU &&
      c->idcase != VLP_CASE_TWO_STREAM)
    return set_err(nullptr, VLP_E_CASE, "Unknown test case !");
  if (!(c->dt > 0) || !std::isfinite(c->dt)) return set_err(nullptr, VLP_E_ARG, "dt must be positive");
  for (int i = 0; i < 4; i++)
    if (!(c->maxPhy[i] > c->minPhy[i])) return set_err(nullptr, VLP_E_ARG, "maxPhy must exceed minPhy");
  return VLP_OK;
}

// exp(-2 pi i m / n) with exact values on the axes, interleaved complex.
std::vector<double> twiddles(int n) {
  std::vector<double> tw(2 * (size_t)n);
  for (int m = 0; m < n; m++) {
    double re, im;
    if (m == 0) { re = 1.0; im = 0.0; }
    else if (2 * m == n) { re = -1.0; im = 0.0; }
    else if (4 * m == n) { re = 0.0; im = -1.0; }
    else if (4 * m == 3 * n) { re = 0.0; im = 1.0; }
    else {
      const double ang = 2.0 * M_PI * (double)m / (double)n;
      re = std::cos(ang);
      im = -std::sin(ang);
    }
    tw[2 * m] = re;
    tw[2 * m + 1] = im;
  }
  return tw;
}

void free_all(vlp_ctx *c) {
  void *bufs[] = {c->f,        c->fnext,    c->cxy,       c->tv,     c->c4_alloc, c->rho_part, c->l2_part,
                  c->fb.rho,   c->fb.ex,    c->fb.ey,     c->fb.exy, c->fb.spec1, c->fb.spec2, c->fb.twx,  c->fb.twy,
                  c->fb.diag_part, c->fb.rho_tmp, c->fb.diag, c->d_err, c->adj_c4s, c->adj_blk, c->adj_tmp,
                  c->adj_cxy, c->adj_E, c->adj_rho, c->adj_k, c->adj_g, c->adj_grad, c->adj_Mt[0], c->adj_Mt[1],
                  c->adj_Mt[2], c->adj_Mt[3], c->adj_M[0], c->adj_M[1], c->adj_M[2], c->adj_M[3], c->adj_axy,
                  c->adj_scl, c->lag_tab,
                  c->mf_m[0], c->mf_m[1]};
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  if (c->h_err) (void)hipHostFree(c->h_err);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : c->evb)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

template <typename T> int run_spline_xy(vlp_ctx *c) {
  const Geom &g = c->g;
  c->half_rho = false;
  if (c->xy_groups) {
    // fused: CXY and the density of the 2-D half step in one pass
    bool ok;
    if constexpr (sizeof(T) == sizeof(float)) {
      ok = c->precision == VLP_FP32_ACC
               ? vlp::launch_xy_adv2d<T, float>(g, (const T *)c->f, (T *)c->cxy, 0.5 * c->cfg.dt, c->rho_part,
                                                c->d_err, c->stream)
               : vlp::launch_xy_adv2d<T, double>(g, (const T *)c->f, (T *)c->cxy, 0.5 * c->cfg.dt, c->rho_part,
                                                 c->d_err, c->stream);
    } else {
      ok = vlp::launch_xy_adv2d<T, double>(g, (const T *)c->f, (T *)c->cxy, 0.5 * c->cfg.dt, c->rho_part, c->d_err,
                                           c->stream);
    }
    if (ok) {
      c->half_rho = true;
      c->rho_outer = (long)c->xy_groups * g.n[3];
      return VLP_OK;
    }
  }
  const long N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const long E0 = g.n[0] + 4, E1 = N1 + 4;
  T *tx = (T *)c->tv;
  vlp::launch_spline_x<T>(g, (const T *)c->f, tx, c->stream);
  // y-pass: tx (E0, N1, N2, N3) -> cxy (E0, E1, N2, N3); lanes along x
  vlp::LineMap m{};
  m.N = (int)N1;
  m.X0 = E0; m.X1 = N2; m.A = N3;
  m.is0 = 1; m.is1 = E0 * N1; m.ios = E0 * N1 * N2; m.ils = E0;
  m.os0 = 1; m.os1 = E0 * E1; m.oos = E0 * E1 * N2; m.ols = E0;
  vlp::launch_spline_line<T>(tx, (T *)c->cxy, m, c->stream);
  return VLP_OK;
}

// vx-pass: cxy (E0*E1, N2, N3) -> tv (E0*E1, E2, N3); lanes along (x, y)
vlp::LineMap vx_map(const Geom &g) {
  const long N2 = g.n[2], N3 = g.n[3];
  const long P = (long)(g.n[0] + 4) * (g.n[1] + 4), E2 = N2 + 4;
  vlp::LineMap m{};
  m.N = (int)N2;
  m.X0 = P; m.X1 = N3; m.A = 1;
  m.is0 = 1; m.is1 = P * N2; m.ios = 0; m.ils = P;
  m.os0 = 1; m.os1 = P * E2; m.oos = 0; m.ols = P;
  return m;
}

// C4 element of position (-2, -2, -2, -2), the first coefficient the vy pass
// writes, and the shift of C4 in its allocation that puts it on a 64-byte
// boundary (the pitches past x are multiples of 16 elements).
long c4_origin(const Geom &g) {
  const vlp::C4Pitch cp = vlp::c4_pitch(g.n[0], g.n[1], g.n[2]);
  return 1 + cp.s1 + cp.s2 + cp.s3;
}
long c4_shift(const Geom &g) { return (16 - c4_origin(g) % 16) % 16; }

// vy-pass: tv (E0, E1, E2, N3) -> C4 at positions -2..N+1 (its x rows hold
// exactly the E0 positions, so a tile's output is contiguous like its input)
vlp::LineMap vy_map(const Geom &g) {
  const long E0 = g.n[0] + 4, E1 = g.n[1] + 4, E2 = g.n[2] + 4;
  const vlp::C4Pitch cp = vlp::c4_pitch(g.n[0], g.n[1], g.n[2]);
  vlp::LineMap q{};
  q.N = g.n[3];
  q.X0 = E0; q.X1 = E1; q.A = E2;
  q.is0 = 1; q.is1 = E0; q.ios = E0 * E1; q.ils = E0 * E1 * E2;
  q.os0 = 1; q.os1 = cp.s1; q.oos = cp.s2; q.ols = cp.s3;
  return q;
}

template <typename T> int run_spline_vxvy(vlp_ctx *c) {
  const Geom &g = c->g;
  T *c4 = (T *)c->c4 + c4_origin(g);
  if (c->mf_m[0]) {  // matrix products on the MFMA units
    vlp::launch_spline_mfma<T>((const T *)c->cxy, (T *)c->tv, vx_map(g), (const T *)c->mf_m[0],
                               vlp::spline_mfma_kp(g.n[2]), c->stream);
    vlp::launch_spline_mfma<T>((const T *)c->tv, c4, vy_map(g), (const T *)c->mf_m[1], vlp::spline_mfma_kp(g.n[3]),
                               c->stream);
    return VLP_OK;
  }
  vlp::launch_spline_line<T>((const T *)c->cxy, (T *)c->tv, vx_map(g), c->stream, c->sp_mode[0]);
  vlp::launch_spline_line<T>((const T *)c->tv, c4, vy_map(g), c->stream, c->sp_mode[1]);
  return VLP_OK;
}

template <typename T> void run_advect2d(vlp_ctx *c) {
  if (c->half_rho) return;  // the fused xy kernel produced it
  vlp::launch_advect2d_rho<T>(c->g, (const T *)c->cxy, 0.5 * c->cfg.dt, c->precision == VLP_FP32_ACC, c->rho_part,
                              c->d_err, c->stream);
  c->rho_outer = c->g.n[3];
}

// Returns true when rho/L2 of fnext are still to be taken (run_moments_next).
template <typename T> bool run_advect4d(vlp_ctx *c) {
  const bool need = vlp::launch_advect4d<T>(c->g, (const T *)c->c4, c->fb.ex, c->fb.ey, c->fb.exy, c->cfg.dt,
                                            c->precision == VLP_FP32_ACC, (T *)c->fnext, c->rho_part, c->l2_part,
                                            c->d_err, c->stream, c->ys_tvx, c->ys_nb);
  c->rho_outer = c->g.n[3];
  c->half_rho = false;
  return need;
}

// rho and L2 partials of fnext (the reference's field_rho of fnp1 and
// computeL2norm, VMO/Timestep.hpp:77-90).
template <typename T> void run_moments_next(vlp_ctx *c) {
  vlp::launch_moments<T>(c->g, (const T *)c->fnext, c->rho_part, c->l2_part, c->stream);
  c->rho_outer = c->g.n[3];
}

template <typename T> void run_moments(vlp_ctx *c) {
  vlp::launch_moments<T>(c->g, (const T *)c->f, c->rho_part, c->l2_part, c->stream);
  c->rho_outer = c->g.n[3];
  c->half_rho = false;
}

int check_launch(vlp_ctx *c, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    c->broken = true;
    return set_err(c, VLP_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  }
  return VLP_OK;
}

// device-side error flag -> VLP_E_CFL (reference: testError, VMO/Advection.hpp:18-23)
int sync_and_check(vlp_ctx *c) {
  HIP_TRY(c, hipMemcpyAsync(c->h_err, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (*c->h_err) {
    c->broken = true;
    return set_err(c, VLP_E_CFL, "Time step is too large, exiting");
  }
  return VLP_OK;
}

int field(vlp_ctx *c, int iter, bool record, bool with_l2, int *iter_dev = nullptr) {
  vlp::launch_field(c->g, c->fb, c->rho_part, c->rho_outer, with_l2 ? c->l2_part : nullptr, c->n_l2, iter, record,
                    c->stream, iter_dev);
  return check_launch(c, "field");
}

float ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

// The kernels of one onetimestep (VMO/Timestep.hpp:34-97) on the context's
// stream, then the swap of f and fnext.  Eager (iter_dev null): HIP events
// between the phases, diagnostics at index iter + 1.  Under graph capture
// (iter_dev non-null): no events, the diagnostics index comes from the device
// counter *iter_dev (advanced by the field solve).
int enqueue_step(vlp_ctx *c, int *iter_dev, hipEvent_t *evs = nullptr) {
  const bool f64 = c->precision == VLP_FP64;
  const bool ev_on = iter_dev == nullptr;
  hipEvent_t *ev = evs ? evs : c->ev;
  if (ev_on) HIP_TRY(c, hipEventRecord(ev[0], c->stream));
  if (f64) run_spline_xy<double>(c); else run_spline_xy<float>(c);
  if (ev_on) HIP_TRY(c, hipEventRecord(ev[1], c->stream));
  if (f64) run_advect2d<double>(c); else run_advect2d<float>(c);
  if (ev_on) HIP_TRY(c, hipEventRecord(ev[2], c->stream));
  int rc = field(c, c->iter + 1, false, false);
  if (rc) return rc;
  if (ev_on) HIP_TRY(c, hipEventRecord(ev[3], c->stream));
  if (f64) run_spline_vxvy<double>(c); else run_spline_vxvy<float>(c);
  if (ev_on) HIP_TRY(c, hipEventRecord(ev[4], c->stream));
  const bool moments = f64 ? run_advect4d<double>(c) : run_advect4d<float>(c);
  if (ev_on) HIP_TRY(c, hipEventRecord(ev[5], c->stream));
  if (moments) {
    if (f64) run_moments_next<double>(c); else run_moments_next<float>(c);
  }
  rc = field(c, c->iter + 1, true, true, iter_dev);
  if (rc) return rc;
  if (ev_on) HIP_TRY(c, hipEventRecord(ev[6], c->stream));
  rc = check_launch(c, "step");
  if (rc) return rc;
  std::swap(c->f, c->fnext);
  return VLP_OK;
}

// onetimestep of the Lagrange split-step solver (VO/timestep.hpp:11-55):
// x/2, y/2 (+ density), field solve with the step's diagnostics, vx, vy, x/2,
// y/2; the result ends in f.  Timers: the x/y passes count as advec2D, the
// velocity passes as advec4D.
int lag_step_once(vlp_ctx *c) {
  hipEvent_t *ev = c->ev;
  const Geom &g = c->g;
  const double dt = c->cfg.dt;
  double *f = (double *)c->f, *t = (double *)c->fnext;
  hipStream_t s = c->stream;
  HIP_TRY(c, hipEventRecord(ev[0], s));
  // x/2 + y/2: one pass through an LDS plane and the density from a
  // streaming pass over its output, or two kernels with the density fused
  const int nch = vlp::lag_xy_chunks(g);
  const bool fused = vlp::launch_lag_xy(g, f, t, c->lag_tab, nch, s);
  if (fused) {
    std::swap(f, t);
  } else {
    vlp::launch_lag_x(g, f, t, 0.5 * dt, s);
    vlp::launch_lag_y(g, t, f, 0.5 * dt, c->rho_part, s);
  }
  HIP_TRY(c, hipEventRecord(ev[1], s));
  if (fused) vlp::launch_moments<double>(g, f, c->rho_part, c->l2_part, s);
  c->rho_outer = g.n[3];
  int rc = field(c, c->iter + 1, true, false);
  if (rc) return rc;
  HIP_TRY(c, hipEventRecord(ev[2], s));
  vlp::launch_lag_v(g, 2, f, t, c->fb.ex, dt, s);
  vlp::launch_lag_v(g, 3, t, f, c->fb.ey, dt, s);
  HIP_TRY(c, hipEventRecord(ev[3], s));
  if (vlp::launch_lag_xy(g, f, t, c->lag_tab, nch, s)) {
    std::swap(f, t);
  } else {
    vlp::launch_lag_x(g, f, t, 0.5 * dt, s);
    vlp::launch_lag_y(g, t, f, 0.5 * dt, nullptr, s);
  }
  HIP_TRY(c, hipEventRecord(ev[4], s));
  if (f != (double *)c->f) std::swap(c->f, c->fnext);  // the result is in f
  rc = check_launch(c, "step");
  if (rc) return rc;
  c->iter++;
  rc = sync_and_check(c);
  if (rc) return rc;
  c->tsec[VLP_T_ADVEC2D] += 1e-3 * (ev_ms(ev[0], ev[1]) + ev_ms(ev[3], ev[4]));
  c->tsec[VLP_T_FIELD] += 1e-3 * ev_ms(ev[1], ev[2]);
  c->tsec[VLP_T_ADVEC4D] += 1e-3 * ev_ms(ev[2], ev[3]);
  c->tcalls[VLP_T_ADVEC2D] += 2;
  c->tcalls[VLP_T_ADVEC4D]++;
  c->tcalls[VLP_T_FIELD]++;
  c->tcalls[VLP_T_FOURIER]++;
  c->tcalls[VLP_T_DIAG]++;
  return VLP_OK;
}

// One onetimestep + swap, events between phases.
// Phase timers of one spline step from its events ev[0..6].
void add_step_timers(vlp_ctx *c, hipEvent_t *ev) {
  c->tsec[VLP_T_SPLINE_XY] += 1e-3 * ev_ms(ev[0], ev[1]);
  c->tsec[VLP_T_ADVEC2D] += 1e-3 * ev_ms(ev[1], ev[2]);
  c->tsec[VLP_T_FIELD] += 1e-3 * (ev_ms(ev[2], ev[3]) + ev_ms(ev[5], ev[6]));
  c->tsec[VLP_T_SPLINE_VXVY] += 1e-3 * ev_ms(ev[3], ev[4]);
  c->tsec[VLP_T_ADVEC4D] += 1e-3 * ev_ms(ev[4], ev[5]);
  c->tcalls[VLP_T_SPLINE_XY]++;
  c->tcalls[VLP_T_ADVEC2D]++;
  c->tcalls[VLP_T_ADVEC4D]++;
  c->tcalls[VLP_T_SPLINE_VXVY]++;
  c->tcalls[VLP_T_FIELD] += 2;
  c->tcalls[VLP_T_ALLREDUCE] += 2;
  c->tcalls[VLP_T_FOURIER] += 2;
  c->tcalls[VLP_T_DIAG] += 2;
  c->tcalls[VLP_T_PACK]++;
  c->tcalls[VLP_T_COMM]++;
  c->tcalls[VLP_T_UNPACK]++;
}

int step_once(vlp_ctx *c) {
  if (c->scheme == VLP_SCHEME_LAGRANGE) return lag_step_once(c);
  int rc = enqueue_step(c, nullptr);
  if (rc) return rc;
  c->iter++;
  rc = sync_and_check(c);
  if (rc) return rc;
  add_step_timers(c, c->ev);
  return VLP_OK;
}

// Several spline steps: step s + 1 is enqueued before the host reads step s's
// error flag and events, so the GPU does not idle between steps while the
// host synchronises.  Step s uses event set s mod 2 (ev, evb) and flag slot
// s mod 2; its copy of the (sticky) device flag and a last event are queued
// behind its kernels.  A CFL error found at step s marks the context broken
// with iter at that step (the step already queued after it ran on its
// result, as the reference would not, but the context is unusable then).
int steps_pipelined(vlp_ctx *c, int nsteps) {
  // a HIP failure here leaves queued steps of unknown outcome: the context
  // is marked unusable
#define HIP_TRY_STEP(expr)                                                                  \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      c->broken = true;                                                                     \
      return set_err(c, VLP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    }                                                                                       \
  } while (0)
  hipEvent_t *sets[2] = {c->ev, c->evb};
  auto t_last = std::chrono::steady_clock::now();
  int pend = -1;
  for (int s = 0; s <= nsteps; s++) {
    const int k = s & 1;
    if (s < nsteps) {
      const int rc = enqueue_step(c, nullptr, sets[k]);
      if (rc) {
        c->broken = true;
        return rc;
      }
      c->iter++;
      HIP_TRY_STEP(hipMemcpyAsync(c->h_err + k, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY_STEP(hipEventRecord(sets[k][7], c->stream));
    }
    if (pend >= 0) {
      HIP_TRY_STEP(hipEventSynchronize(sets[pend][7]));
      if (c->h_err[pend]) {
        c->broken = true;
        if (s < nsteps) c->iter--;
        return set_err(c, VLP_E_CFL, "Time step is too large, exiting");
      }
      add_step_timers(c, sets[pend]);
      const auto now = std::chrono::steady_clock::now();
      const double dt = std::chrono::duration<double>(now - t_last).count();
      t_last = now;
      c->tsec[VLP_T_MAINLOOP] += dt;
      c->tcalls[VLP_T_MAINLOOP]++;
      c->tsec[VLP_T_TOTAL] += dt;
    }
    pend = s < nsteps ? k : -1;
  }
#undef HIP_TRY_STEP
  return VLP_OK;
}

int usable(vlp_ctx *c) {
  if (!c) return set_err(nullptr, VLP_E_ARG, "null context");
  if (c->broken) return set_err(c, VLP_E_ARG, "context unusable after an earlier error: " + c->err);
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, VLP_E_HIP, "hipSetDevice failed");
  return VLP_OK;
}

// Calls that exist for the spline solver only (phases, tile tuner, adjoint).
int spline_only(vlp_ctx *c) {
  int rc = usable(c);
  if (rc) return rc;
  if (c->scheme != VLP_SCHEME_SPLINE) return set_err(c, VLP_E_ARG, "not available for a Lagrange split-step context");
  return VLP_OK;
}

template <typename T> int upload_f(vlp_ctx *c, const double *host) {
  std::vector<T> tmp;
  const void *src = host;
  if (sizeof(T) != sizeof(double)) {
    tmp.resize(c->n_f);
    for (size_t i = 0; i < c->n_f; i++) tmp[i] = (T)host[i];
    src = tmp.data();
  }
  HIP_TRY(c, hipMemcpyAsync(c->f, src, c->n_f * sizeof(T), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return VLP_OK;
}

int download(vlp_ctx *c, const void *dev, size_t n, double *host) {
  if (c->esz == sizeof(double)) {
    HIP_TRY(c, hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> tmp(n);
    HIP_TRY(c, hipMemcpyAsync(tmp.data(), dev, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; i++) host[i] = tmp[i];
  }
  return VLP_OK;
}

// field solve + diagnostics of the current f at iteration `iter`
// (VMO/Vlp4d.cpp:30-34 for iter 0).
int refresh_from_f(vlp_ctx *c) {
  if (c->precision == VLP_FP64) run_moments<double>(c); else run_moments<float>(c);
  // the Lagrange solver records no L2 norm (VO/diags.cpp:13-39)
  int rc = field(c, c->iter, true, c->scheme != VLP_SCHEME_LAGRANGE);
  if (rc) return rc;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return VLP_OK;
}

}  // namespace

namespace vlp {
int set_error(int code, const char *msg) { return set_err(nullptr, code, msg); }
int prepare_config(vlp_config *cfg) {
  derive_dx(cfg);
  return check_config(cfg);
}
}  // namespace vlp

extern "C" {

const char *vlp_version(void) { return "vlp4d-mi355x 0.1 (gfx950)"; }

int vlp_parse_deck(const char *path, vlp_config *cfg) {
  if (!path || !cfg) return set_err(nullptr, VLP_E_ARG, "null argument");
  FILE *s = fopen(path, "r");
  if (!s) return set_err(nullptr, VLP_E_FILE, "import: error file not found");
  vlp_config c{};
  char idcase[8] = {0};
  bool ok = true;
  for (int i = 0; i < 4 && ok; i++) {  // VMO/Init.cpp:32-39
    int nx = 0;
    ok = skip_to_colon(s) && fscanf(s, " %d\n", &nx) == 1;
    c.nxmax[i] = nx;
  }
  for (int i = 0; i < 4 && ok; i++) {  // VMO/Init.cpp:41-51
    ok = skip_to_colon(s) && fscanf(s, " %le\n", &c.minPhy[i]) == 1;
    ok = ok && skip_to_colon(s) && fscanf(s, " %le\n", &c.maxPhy[i]) == 1;
  }
  ok = ok && skip_to_colon(s) && fgets(idcase, 7, s) != nullptr;   // VMO/Init.cpp:53-56
  ok = ok && skip_to_colon(s) && fscanf(s, " %le\n", &c.dt) == 1;  // :58-61
  ok = ok && skip_to_colon(s) && fscanf(s, " %d\n", &c.nbiter) == 1;
  ok = ok && skip_to_colon(s) && fscanf(s, " %d\n", &c.ifreq) == 1;
  ok = ok && skip_to_colon(s) && fscanf(s, " %d\n", &c.fxvx) == 1;
  fclose(s);
  if (!ok) return set_err(nullptr, VLP_E_FILE, "import: malformed deck");
  c.idcase = atoi(idcase);  // VMO/Init.cpp:78
  derive_dx(&c);
  *cfg = c;
  return VLP_OK;
}

int vlp_create(const char *dat_path, int precision, int device, vlp_ctx **out) {
  vlp_config cfg;
  const int rc = vlp_parse_deck(dat_path, &cfg);
  if (rc) return rc;
  return vlp_create_config(&cfg, precision, device, out);
}

static int create_ctx(const vlp_config *cfg_in, int precision, int device, int scheme, vlp_ctx **out);

int vlp_create_config(const vlp_config *cfg_in, int precision, int device, vlp_ctx **out) {
  return create_ctx(cfg_in, precision, device, VLP_SCHEME_SPLINE, out);
}

int vlp_create_lagrange(const vlp_config *cfg_in, int device, vlp_ctx **out) {
  return create_ctx(cfg_in, VLP_FP64, device, VLP_SCHEME_LAGRANGE, out);
}

static int create_ctx(const vlp_config *cfg_in, int precision, int device, int scheme, vlp_ctx **out) {
  if (!cfg_in || !out) return set_err(nullptr, VLP_E_ARG, "null argument");
  *out = nullptr;
  if (precision != VLP_FP64 && precision != VLP_FP32 && precision != VLP_FP32_ACC)
    return set_err(nullptr, VLP_E_ARG, "precision must be VLP_FP64 (64), VLP_FP32 (32) or VLP_FP32_ACC (33)");
  const bool lag = scheme == VLP_SCHEME_LAGRANGE;
  vlp_config cfg = *cfg_in;
  derive_dx(&cfg);
  int rc = check_config(&cfg);
  if (rc) return rc;
  if (lag && cfg.idcase != 2 && cfg.idcase != 5 && cfg.idcase != 10 && cfg.idcase != 20)
    return set_err(nullptr, VLP_E_CASE, "Unknown test case !");  // miniapps/vlp4d/openmp/init.cpp:124-142
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return set_err(nullptr, VLP_E_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) return set_err(nullptr, VLP_E_ARG, "device index out of range");
  vlp_ctx *c = new vlp_ctx();
  c->cfg = cfg;
  c->precision = precision;
  c->scheme = scheme;
  c->device = device;
  c->esz = precision == VLP_FP64 ? sizeof(double) : sizeof(float);
  for (int i = 0; i < 4; i++) {
    c->g.n[i] = cfg.nxmax[i];
    c->g.minPhy[i] = cfg.minPhy[i];
    c->g.maxPhy[i] = cfg.maxPhy[i];
    c->g.dx[i] = cfg.dx[i];
    c->g.inv_dx[i] = 1. / cfg.dx[i];
  }
  c->g.idcase = cfg.idcase;
  const size_t N0 = cfg.nxmax[0], N1 = cfg.nxmax[1], N2 = cfg.nxmax[2], N3 = cfg.nxmax[3];
  c->n_f = N0 * N1 * N2 * N3;
  // the Lagrange solver keeps no spline coefficients
  c->n_cxy = lag ? 0 : (N0 + 4) * (N1 + 4) * N2 * N3;
  c->n_tv = lag ? 0 : (N0 + 4) * (N1 + 4) * (N2 + 4) * N3;
  c->n_c4 = lag ? 0 : (N0 + 4) * (N1 + 4) * (N2 + 4) * (N3 + 4);
  c->n_c4s = lag ? 0 : vlp::c4_elems(N0, N1, N2, N3);
  c->n_l2 = vlp::l2_partials(c->g);
  c->xy_groups = lag ? 0 : vlp::xy_fused_groups(c->g, c->esz);
  c->rho_outer = N3;
  const size_t nxh = N0 / 2 + 1;
  auto fail = [&](int code) {
    free_all(c);
    delete c;
    return code;
  };
#define ALLOC(ptr, bytes)                                                                   \
  do {                                                                                      \
    if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess)                                  \
      return fail(set_err(nullptr, VLP_E_HIP, std::string("hipMalloc failed: ") + #ptr));  \
  } while (0)
  if (hipSetDevice(device) != hipSuccess) return fail(set_err(nullptr, VLP_E_HIP, "hipSetDevice failed"));
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(set_err(nullptr, VLP_E_HIP, "hipStreamCreate failed"));
  for (auto &e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(set_err(nullptr, VLP_E_HIP, "hipEventCreate failed"));
  for (auto &e : c->evb)
    if (hipEventCreate(&e) != hipSuccess) return fail(set_err(nullptr, VLP_E_HIP, "hipEventCreate failed"));
  ALLOC(c->f, c->n_f * c->esz);
  ALLOC(c->fnext, c->n_f * c->esz);
  if (lag) ALLOC(c->lag_tab, vlp::lag_tables_bytes(c->g));
  if (!lag) {
    ALLOC(c->cxy, c->n_cxy * c->esz);
    ALLOC(c->tv, c->n_tv * c->esz);
    // C4 starts c4_shift elements into its allocation, so that each line of
    // the vy pass's output starts on a 64-byte boundary
    ALLOC(c->c4_alloc, (c->n_c4s + 16) * c->esz);
    c->c4 = (char *)c->c4_alloc + c4_shift(c->g) * c->esz;
  }
  ALLOC(c->rho_part, N0 * N1 * N3 * (c->xy_groups > 1 ? (size_t)c->xy_groups : 1) * sizeof(double));
  ALLOC(c->l2_part, (size_t)c->n_l2 * sizeof(double));
  ALLOC(c->fb.rho, N0 * N1 * sizeof(double));
  ALLOC(c->fb.ex, N0 * N1 * sizeof(double));
  ALLOC(c->fb.ey, N0 * N1 * sizeof(double));
  ALLOC(c->fb.exy, 2 * N0 * N1 * sizeof(double));
  ALLOC(c->fb.spec1, 2 * nxh * N1 * sizeof(double));
  ALLOC(c->fb.spec2, 3 * 2 * nxh * N1 * sizeof(double));
  ALLOC(c->fb.twx, 2 * N0 * sizeof(double));
  ALLOC(c->fb.twy, 2 * N1 * sizeof(double));
  ALLOC(c->fb.diag_part, 4 * N1 * sizeof(double));
  ALLOC(c->fb.rho_tmp, (size_t)vlp::RHO_SLICES * N0 * N1 * sizeof(double));
  ALLOC(c->fb.diag, 5 * (size_t)(cfg.nbiter + 1) * sizeof(double));
  ALLOC(c->d_err, sizeof(int));
#undef ALLOC
  if (hipHostMalloc((void **)&c->h_err, 2 * sizeof(int), hipHostMallocDefault) != hipSuccess)
    return fail(set_err(nullptr, VLP_E_HIP, "hipHostMalloc failed"));
  c->fb.nbiter = cfg.nbiter;
  const std::vector<double> twx = twiddles((int)N0), twy = twiddles((int)N1);
  if (hipMemcpy(c->fb.twx, twx.data(), twx.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->fb.twy, twy.data(), twy.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(c->fb.diag, 0, 5 * (size_t)(cfg.nbiter + 1) * sizeof(double)) != hipSuccess ||
      hipMemset(c->d_err, 0, sizeof(int)) != hipSuccess || hipMemset(c->f, 0, c->n_f * c->esz) != hipSuccess ||
      (!lag && hipMemset(c->c4, 0, c->n_c4s * c->esz) != hipSuccess) ||
      hipMemset(c->fb.rho, 0, N0 * N1 * sizeof(double)) != hipSuccess ||
      hipMemset(c->fb.ex, 0, N0 * N1 * sizeof(double)) != hipSuccess ||
      hipMemset(c->fb.ey, 0, N0 * N1 * sizeof(double)) != hipSuccess ||
      hipMemset(c->fb.exy, 0, 2 * N0 * N1 * sizeof(double)) != hipSuccess)
    return fail(set_err(nullptr, VLP_E_HIP, "device initialisation failed"));
  const char *mfe = getenv("VLP4D_SPLINE_MFMA");
  if (!lag && mfe && atoi(mfe) == 1 && vlp::spline_mfma_eligible(vx_map(c->g), c->esz) &&
      vlp::spline_mfma_eligible(vy_map(c->g), c->esz)) {
    for (int k = 0; k < 2; k++) {
      const int N = c->g.n[2 + k];
      const size_t ne = vlp::spline_mfma_matrix_elems(N);
      double *work = nullptr;
      if (hipMalloc(&c->mf_m[k], ne * c->esz) != hipSuccess ||
          hipMalloc((void **)&work, ((size_t)N * N + ne) * sizeof(double)) != hipSuccess)
        return fail(set_err(nullptr, VLP_E_HIP, "hipMalloc failed (MFMA spline operator)"));
      if (precision == VLP_FP64) vlp::build_spline_mfma_matrix<double>(N, (double *)c->mf_m[k], work, c->stream);
      else vlp::build_spline_mfma_matrix<float>(N, (float *)c->mf_m[k], work, c->stream);
      const bool ok = hipStreamSynchronize(c->stream) == hipSuccess && hipGetLastError() == hipSuccess;
      (void)hipFree(work);
      if (!ok) return fail(set_err(nullptr, VLP_E_HIP, "MFMA spline operator"));
    }
  }
  // init(): VMO/Init.cpp:370-372 then VMO/Vlp4d.cpp:30-34
  const auto t0 = std::chrono::steady_clock::now();
  if (lag) {
    vlp::launch_lag_init(c->g, (double *)c->f, c->stream);  // VO/init.cpp:144-285
    vlp::launch_lag_tables(c->g, c->lag_tab, 0.5 * cfg.dt, c->stream);
  }
  else if (precision == VLP_FP64) vlp::launch_init_case<double>(c->g, (double *)c->f, c->stream);
  else vlp::launch_init_case<float>(c->g, (float *)c->f, c->stream);
  rc = refresh_from_f(c);
  if (rc) return fail(rc);
  c->tsec[VLP_T_TOTAL] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  c->tcalls[VLP_T_FIELD]++;
  c->tcalls[VLP_T_ALLREDUCE]++;
  c->tcalls[VLP_T_FOURIER]++;
  c->tcalls[VLP_T_DIAG]++;
  *out = c;
  return VLP_OK;
}

int vlp_destroy(vlp_ctx *c) {
  if (!c) return VLP_OK;
  if (c->owner)  // its buffers are baked into the ensemble's graph
    return set_err(c, VLP_E_ARG, "context is a member of a live ensemble; destroy the ensemble first");
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
  return VLP_OK;
}

int vlp_step(vlp_ctx *c, int nsteps) {
  int rc = usable(c);
  if (rc) return rc;
  if (nsteps < 0 || c->iter + nsteps > c->cfg.nbiter)
    return set_err(c, VLP_E_ARG, "step count exceeds nbiter of the configuration");
  if (c->scheme == VLP_SCHEME_SPLINE && nsteps > 1) return steps_pipelined(c, nsteps);
  for (int s = 0; s < nsteps; s++) {
    const auto t0 = std::chrono::steady_clock::now();
    rc = step_once(c);
    if (rc) return rc;
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    c->tsec[VLP_T_MAINLOOP] += dt;
    c->tcalls[VLP_T_MAINLOOP]++;
    c->tsec[VLP_T_TOTAL] += dt;
  }
  return VLP_OK;
}

int vlp_iter(const vlp_ctx *c) { return c ? c->iter : -1; }

int vlp_get_config(const vlp_ctx *c, vlp_config *cfg) {
  if (!c || !cfg) return set_err(nullptr, VLP_E_ARG, "null argument");
  *cfg = c->cfg;
  return VLP_OK;
}

int vlp_diags(vlp_ctx *c, double *nrj, double *nrjx, double *nrjy, double *mass, double *l2norm, int n) {
  int rc = usable(c);
  if (rc) return rc;
  if (n < 0 || n > c->iter + 1) return set_err(c, VLP_E_ARG, "n exceeds iter+1");
  const int nb = c->cfg.nbiter + 1;
  std::vector<double> h(5 * (size_t)nb);
  HIP_TRY(c, hipMemcpyAsync(h.data(), c->fb.diag, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  double *outs[5] = {nrj, nrjx, nrjy, mass, l2norm};
  for (int k = 0; k < 5; k++)
    if (outs[k]) memcpy(outs[k], h.data() + (size_t)k * nb, (size_t)n * sizeof(double));
  return VLP_OK;
}

int vlp_save(vlp_ctx *c, const char *dir) {
  int rc = usable(c);
  if (rc) return rc;
  const int nb = c->cfg.nbiter + 1;
  std::vector<double> h(5 * (size_t)nb);
  HIP_TRY(c, hipMemcpyAsync(h.data(), c->fb.diag, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const std::string d = dir ? dir : ".";
  if (c->scheme == VLP_SCHEME_LAGRANGE) {  // VO/diags.cpp:41-52: t, nrj, mass
    FILE *f = fopen((d + "/nrj.out").c_str(), "w");
    if (!f) return set_err(c, VLP_E_FILE, "cannot write nrj.out");
    for (int it = 0; it <= c->cfg.nbiter; ++it)
      fprintf(f, "%17.13e %17.13e %17.13e\n", it * c->cfg.dt, h[it], h[3 * nb + it]);
    fclose(f);
    return VLP_OK;
  }
  FILE *f = fopen((d + "/nrj.out").c_str(), "w");  // VMO/Diags.cpp:403-411
  if (!f) return set_err(c, VLP_E_FILE, "cannot write nrj.out");
  for (int it = 0; it <= c->cfg.nbiter; ++it)
    fprintf(f, "%17.13e %17.13e %17.13e %17.13e %17.13e\n", it * c->cfg.dt, h[it], h[nb + it], h[2 * nb + it],
            h[3 * nb + it]);
  fclose(f);
  f = fopen((d + "/l2norm.out").c_str(), "w");  // VMO/Diags.cpp:413-420
  if (!f) return set_err(c, VLP_E_FILE, "cannot write l2norm.out");
  for (int it = 0; it <= c->cfg.nbiter; ++it) fprintf(f, "%17.13e %17.13e\n", it * c->cfg.dt, h[4 * nb + it]);
  fclose(f);
  return VLP_OK;
}

// Cross sections of f for the fxvx diagnostics, in the CSV layout of
// print_fxvx (miniapps/vlp4d_mpi/openmp/Advection.hpp:500-590,
// miniapps/vlp4d/openmp/advect.hpp:318-341): one row per velocity index,
// "%20.13le, " between values.  fxvx = f(x, y = 0, vx, vy = Nvy/2); spline
// contexts also write fyvy = f(x = 0, y, vx = Nvx/2, vy).  The spline
// miniapp's own version writes zeros (its guard `nymin <= iy && iy && nymax`
// is false at iy = 0) and takes vy = Ny/2; these are the sections it means.
static int write_csv(vlp_ctx *c, const std::string &path, const std::vector<double> &v, int ncol, int nrow) {
  FILE *f = fopen(path.c_str(), "w");
  if (!f) return set_err(c, VLP_E_FILE, "cannot write " + path);
  for (int r = 0; r < nrow; r++)
    for (int k = 0; k < ncol; k++) fprintf(f, k == ncol - 1 ? "%20.13le\n" : "%20.13le, ", v[(size_t)r * ncol + k]);
  fclose(f);
  return VLP_OK;
}

int vlp_write_fxvx(vlp_ctx *c, const char *dir) {
  int rc = usable(c);
  if (rc) return rc;
  const size_t N0 = c->cfg.nxmax[0], N1 = c->cfg.nxmax[1], N2 = c->cfg.nxmax[2], N3 = c->cfg.nxmax[3];
  const size_t e = c->esz;
  const char *f = (const char *)c->f;
  std::vector<char> a(N0 * N2 * e), b(N1 * N3 * e);
  // fxvx: N2 rows of N0 contiguous values, N0*N1 apart
  HIP_TRY(c, hipMemcpy2DAsync(a.data(), N0 * e, f + (N0 * N1 * N2 * (N3 / 2)) * e, N0 * N1 * e, N0 * e, N2,
                              hipMemcpyDeviceToHost, c->stream));
  const bool fyvy = c->scheme == VLP_SCHEME_SPLINE;
  if (fyvy)  // fyvy: for each vy, N1 values N0 apart
    for (size_t ivy = 0; ivy < N3; ivy++)
      HIP_TRY(c, hipMemcpy2DAsync(b.data() + ivy * N1 * e, e, f + (N0 * N1 * (N2 / 2 + N2 * ivy)) * e, N0 * e, e, N1,
                                  hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  auto widen = [&](const std::vector<char> &raw, size_t n) {
    std::vector<double> v(n);
    for (size_t i = 0; i < n; i++)
      v[i] = e == sizeof(double) ? ((const double *)raw.data())[i] : (double)((const float *)raw.data())[i];
    return v;
  };
  const std::string d = dir ? dir : ".";
  char name[64];
  snprintf(name, sizeof name, "/fxvx_it%06d.csv", c->iter);
  rc = write_csv(c, d + name, widen(a, N0 * N2), (int)N0, (int)N2);
  if (rc || !fyvy) return rc;
  snprintf(name, sizeof name, "/fyvy_it%06d.csv", c->iter);
  return write_csv(c, d + name, widen(b, N1 * N3), (int)N1, (int)N3);
}

int vlp_get_f(vlp_ctx *c, double *host, size_t n) {
  int rc = usable(c);
  if (rc) return rc;
  if (!host || n != c->n_f) return set_err(c, VLP_E_SIZE, "f buffer size mismatch");
  return download(c, c->f, n, host);
}

int vlp_set_f(vlp_ctx *c, const double *host, size_t n) {
  int rc = usable(c);
  if (rc) return rc;
  if (!host || n != c->n_f) return set_err(c, VLP_E_SIZE, "f buffer size mismatch");
  rc = c->precision == VLP_FP64 ? upload_f<double>(c, host) : upload_f<float>(c, host);
  if (rc) return rc;
  return refresh_from_f(c);
}

size_t vlp_buffer_size(const vlp_ctx *c, int which) {
  if (!c) return 0;
  const size_t n2 = (size_t)c->cfg.nxmax[0] * c->cfg.nxmax[1];
  switch (which) {
    case VLP_BUF_F: return c->n_f;
    case VLP_BUF_CXY: return c->n_cxy;
    case VLP_BUF_C4: return c->n_c4;
    case VLP_BUF_RHO:
    case VLP_BUF_EX:
    case VLP_BUF_EY: return n2;
    default: return 0;
  }
}

int vlp_get_buffer(vlp_ctx *c, int which, double *host, size_t n) {
  int rc = usable(c);
  if (rc) return rc;
  if (!host || n != vlp_buffer_size(c, which) || n == 0) return set_err(c, VLP_E_SIZE, "buffer size mismatch");
  switch (which) {
    case VLP_BUF_F: return download(c, c->f, n, host);
    case VLP_BUF_CXY: return download(c, c->cxy, n, host);
    case VLP_BUF_C4: {
      std::vector<double> full(c->n_c4s);
      rc = download(c, c->c4, c->n_c4s, full.data());
      if (rc) return rc;
      const vlp::C4Pitch cp = vlp::c4_pitch(c->cfg.nxmax[0], c->cfg.nxmax[1], c->cfg.nxmax[2]);
      const size_t E0 = (size_t)c->cfg.nxmax[0] + 4, E1 = (size_t)c->cfg.nxmax[1] + 4,
                   E2 = (size_t)c->cfg.nxmax[2] + 4, E3 = (size_t)c->cfg.nxmax[3] + 4;
      size_t k = 0;
      for (size_t i3 = 0; i3 < E3; i3++)
        for (size_t i2 = 0; i2 < E2; i2++)
          for (size_t i1 = 0; i1 < E1; i1++)
            for (size_t i0 = 0; i0 < E0; i0++)
              host[k++] = full[(i0 + 1) + cp.s1 * (i1 + 1) + cp.s2 * (i2 + 1) + cp.s3 * (i3 + 1)];
      return VLP_OK;
    }
    default: break;
  }
  const double *src = which == VLP_BUF_RHO ? c->fb.rho : which == VLP_BUF_EX ? c->fb.ex : c->fb.ey;
  HIP_TRY(c, hipMemcpyAsync(host, src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return VLP_OK;
}

int vlp_spline_xy(vlp_ctx *c) {
  int rc = spline_only(c);
  if (rc) return rc;
  if (c->precision == VLP_FP64) run_spline_xy<double>(c); else run_spline_xy<float>(c);
  rc = check_launch(c, "spline_xy");
  if (rc) return rc;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return VLP_OK;
}

int vlp_advect2d_poisson(vlp_ctx *c) {
  int rc = spline_only(c);
  if (rc) return rc;
  if (c->iter >= c->cfg.nbiter) return set_err(c, VLP_E_ARG, "iteration count would exceed nbiter");
  if (c->precision == VLP_FP64) run_advect2d<double>(c); else run_advect2d<float>(c);
  rc = field(c, c->iter + 1, false, false);
  if (rc) return rc;
  return sync_and_check(c);
}

int vlp_spline_vxvy(vlp_ctx *c) {
  int rc = spline_only(c);
  if (rc) return rc;
  if (c->precision == VLP_FP64) run_spline_vxvy<double>(c); else run_spline_vxvy<float>(c);
  rc = check_launch(c, "spline_vxvy");
  if (rc) return rc;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return VLP_OK;
}

int vlp_advect4d_poisson(vlp_ctx *c) {
  int rc = spline_only(c);
  if (rc) return rc;
  if (c->iter >= c->cfg.nbiter) return set_err(c, VLP_E_ARG, "iteration count would exceed nbiter");
  const bool f64 = c->precision == VLP_FP64;
  const bool moments = f64 ? run_advect4d<double>(c) : run_advect4d<float>(c);
  if (moments) {
    if (f64) run_moments_next<double>(c); else run_moments_next<float>(c);
  }
  rc = field(c, c->iter + 1, true, true);
  if (rc) return rc;
  std::swap(c->f, c->fnext);
  c->iter++;
  return sync_and_check(c);
}

int vlp_timer(vlp_ctx *c, int which, double *seconds, int *calls) {
  if (!c || which < 0 || which >= VLP_T_COUNT) return set_err(c, VLP_E_ARG, "bad timer index");
  if (seconds) *seconds = c->tsec[which];
  if (calls) *calls = c->tcalls[which];
  return VLP_OK;
}

const char *vlp_timer_label(int which) { return (which >= 0 && which < VLP_T_COUNT) ? kLabels[which] : ""; }

const char *vlp_last_error(const vlp_ctx *c) { return c ? c->err.c_str() : g_last_error.c_str(); }

int vlp_kernel_variants(const vlp_ctx *c) {
  if (!c) return -1;
  if (c->scheme == VLP_SCHEME_LAGRANGE) return vlp::lag_xy_chunks(c->g) > 0 ? 1 : 0;  // bit 0: fused x/y pass
  int bits = vlp::kernel_variants(c->g, c->precision != VLP_FP64, c->precision == VLP_FP32_ACC);
  if ((bits & 128) && c->ys_tvx) bits = (bits & ~(15 << 8)) | (c->ys_tvx << 8);
  if (vlp::spline_tile_width(vx_map(c->g), c->esz, c->sp_mode[0]) == 128) bits |= 1 << 13;
  if (vlp::spline_tile_width(vy_map(c->g), c->esz, c->sp_mode[1]) == 128) bits |= 1 << 14;
  if (c->mf_m[0]) bits = (bits & ~(3 << 13)) | (1 << 15);
  return bits;
}

// Autotuner of the 4-D advection's workgroup shape and of the vx / vy spline
// kernels: every (vx tile, plane buffers) this geometry admits, and every
// distinct spline kernel choice, is launched on the context's own buffers
// (one warm-up and three timed launches, HIP events on its stream) and the
// fastest is kept for the context's later steps if it beats the preference
// order's choice by more than 2 % (else that choice stays).  Same scan-and-
// keep-the-minimum scheme as the reference's tile tuner
// (miniapps/vlp4d_mpi/kokkos/Tile_size_tuning.hpp:51-96), with its CSV log
// (:127-148).  The launches write fnext, which
// the next step overwrites, and report errors to a scratch flag.
int vlp_autotune(vlp_ctx *c, const char *csv_path) {
  int rc = spline_only(c);
  if (rc) return rc;
  HIP_TRY(c, hipSetDevice(c->device));
  // the VLP4D_YS_* / VLP4D_SPLINE_* environment overrides pin those choices
  const bool ys_env = getenv("VLP4D_YS_TVX") || getenv("VLP4D_YS_NBUF") || getenv("VLP4D_YSTREAM");
  const bool sp_env = getenv("VLP4D_SPLINE_TILE") || getenv("VLP4D_SPLINE_WIDE");
  const size_t tsz = c->precision == VLP_FP64 ? sizeof(double) : sizeof(float);
  // A choice replaces the preference order's (the incumbent) only when it is
  // faster by more than this fraction: back-to-back launches of one kernel
  // differ by about 1 %, and near-ties do not carry over to the full step.
  constexpr float kMargin = 0.02f;
  struct Cand { int tvx, nb; bool pref; float ms; };
  std::vector<Cand> cands;
  for (int tvx : {2, 4, 8})
    for (int nb : {2, 3})
      if (!ys_env && vlp::ystream_eligible(c->g, tsz, tvx, nb))
        cands.push_back({tvx, nb, vlp::ystream_preferred(c->g, tsz, tvx, nb), 0.f});
  // spline passes: each distinct kernel choice (128- or 64-position tiles,
  // line kernels) of the vx and the vy pass, on cxy -> tv and tv -> c4
  // (both overwritten by the next step)
  struct SCand { int pass, mode, width; float ms; };
  std::vector<SCand> spl;
  for (int pass = 0; pass < 2 && !c->mf_m[0] && !sp_env; pass++) {
    const vlp::LineMap m = pass ? vy_map(c->g) : vx_map(c->g);
    std::vector<int> seen;
    for (int mode : {vlp::SPLINE_AUTO, vlp::SPLINE_TILE256, vlp::SPLINE_TILE64, vlp::SPLINE_LINES}) {
      const int w = vlp::spline_tile_width(m, c->esz, mode);
      if (std::find(seen.begin(), seen.end(), w) != seen.end()) continue;
      seen.push_back(w);
      spl.push_back({pass, mode, w, 0.f});
    }
  }
  {
    hipEvent_t a = nullptr, b = nullptr;
    bool sok = hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess;
    const size_t c4off = c4_origin(c->g);
    for (auto &k : spl) {
      if (!sok) break;
      const vlp::LineMap m = k.pass ? vy_map(c->g) : vx_map(c->g);
      for (int it = 0; it < 4 && sok; it++) {
        if (it == 1) sok = hipEventRecord(a, c->stream) == hipSuccess;
        if (c->precision == VLP_FP64)
          vlp::launch_spline_line<double>(k.pass ? (const double *)c->tv : (const double *)c->cxy,
                                          k.pass ? (double *)c->c4 + c4off : (double *)c->tv, m, c->stream, k.mode);
        else
          vlp::launch_spline_line<float>(k.pass ? (const float *)c->tv : (const float *)c->cxy,
                                         k.pass ? (float *)c->c4 + c4off : (float *)c->tv, m, c->stream, k.mode);
        sok = sok && hipGetLastError() == hipSuccess;
      }
      sok = sok && hipEventRecord(b, c->stream) == hipSuccess && hipEventSynchronize(b) == hipSuccess;
      float ms = 0.f;
      sok = sok && hipEventElapsedTime(&ms, a, b) == hipSuccess;
      k.ms = ms / 3.f;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (!sok) return set_err(c, VLP_E_HIP, "autotune launch failed (spline)");
    for (int pass = 0; pass < 2; pass++) {
      const SCand *inc = nullptr, *bs = nullptr;  // SPLINE_AUTO (listed first) is the incumbent
      for (const auto &k : spl)
        if (k.pass == pass) {
          if (!inc) inc = &k;
          else if (!bs || k.ms < bs->ms) bs = &k;
        }
      if (inc) c->sp_mode[pass] = bs && bs->ms < (1.f - kMargin) * inc->ms ? bs->mode : inc->mode;
    }
  }
  const Cand *best = nullptr;
  if (!cands.empty()) {
    int *d_scratch = nullptr;
    HIP_TRY(c, hipMalloc((void **)&d_scratch, sizeof(int)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
    for (auto &k : cands) {
      if (!ok) break;
      for (int it = 0; it < 4 && ok; it++) {
        if (it == 1) ok = hipEventRecord(e0, c->stream) == hipSuccess;
        if (c->precision == VLP_FP64)
          vlp::launch_advect4d_ystream<double, double>(c->g, (const double *)c->c4, c->fb.exy, c->cfg.dt,
                                                       (double *)c->fnext, d_scratch, c->stream, k.tvx, k.nb);
        else if (c->precision == VLP_FP32)
          vlp::launch_advect4d_ystream<float, double>(c->g, (const float *)c->c4, c->fb.exy, c->cfg.dt,
                                                      (float *)c->fnext, d_scratch, c->stream, k.tvx, k.nb);
        else
          vlp::launch_advect4d_ystream<float, float>(c->g, (const float *)c->c4, c->fb.exy, c->cfg.dt,
                                                     (float *)c->fnext, d_scratch, c->stream, k.tvx, k.nb);
        ok = ok && hipGetLastError() == hipSuccess;
      }
      ok = ok && hipEventRecord(e1, c->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
      float ms = 0.f;
      ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
      k.ms = ms / 3.f;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d_scratch);
    if (!ok) return set_err(c, VLP_E_HIP, "autotune launch failed");
    const Cand *inc = &cands[0];
    for (const auto &k : cands)
      if (k.pref) inc = &k;
    best = inc;
    for (const auto &k : cands)
      if (k.ms < (1.f - kMargin) * inc->ms && k.ms < best->ms) best = &k;
    c->ys_tvx = best->tvx;
    c->ys_nb = best->nb;
  }
  if (csv_path && *csv_path) {
    FILE *f = fopen(csv_path, "w");
    if (!f) return set_err(c, VLP_E_FILE, std::string("cannot write ") + csv_path);
    fprintf(f, "kernel,nx,ny,nvx,nvy,precision,vx_tile,plane_buffers,ms_per_launch,preferred,chosen\n");
    for (const auto &k : cands)
      fprintf(f, "advect4d,%d,%d,%d,%d,%d,%d,%d,%.6f,%d,%d\n", c->g.n[0], c->g.n[1], c->g.n[2], c->g.n[3],
              c->precision, k.tvx, k.nb, k.ms, k.pref ? 1 : 0, &k == best ? 1 : 0);
    for (const auto &k : spl)
      fprintf(f, "spline_%s,%d,%d,%d,%d,%d,%d,%d,%.6f,%d,%d\n", k.pass ? "vy" : "vx", c->g.n[0], c->g.n[1],
              c->g.n[2], c->g.n[3], c->precision, k.width, 0, k.ms, k.mode == vlp::SPLINE_AUTO ? 1 : 0,
              k.mode == c->sp_mode[k.pass] ? 1 : 0);
    fclose(f);
  }
  return VLP_OK;
}

// Adjoint of one step at the context's current f (FP64): grad = (d f^{n+1} /
// d f^n)^T g, without advancing the context.  The forward phases are
// recomputed from f (spline_xy, half step, field, spline_vxvy); the adjoint
// is then C4^T I4^T g plus the foot term through E, K^T and the half step
// (adjoint.hip).  The coefficient, density and field buffers are left
// holding intermediate values; the next step recomputes them from f.
// Buffers of the adjoint, the line operators M_d^T and the field's response
// K (built on first use).
static int adjoint_prepare(vlp_ctx *c) {
  int rc = VLP_OK;
  const Geom &g = c->g;
  const long N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const long E0 = N0 + 4, E1 = N1 + 4, E2 = N2 + 4, E3 = N3 + 4;
  const size_t nblk = (size_t)E0 * E1 * E2 * E3;
  hipStream_t s = c->stream;
#define ADJ_ALLOC(p, bytes)                                                         \
  if (!(p) && hipMalloc((void **)&(p), (bytes)) != hipSuccess)                       \
    return set_err(c, VLP_E_HIP, "hipMalloc failed (adjoint buffers)");
  ADJ_ALLOC(c->adj_c4s, c->n_c4s * sizeof(double));
  ADJ_ALLOC(c->adj_blk, nblk * sizeof(double));
  ADJ_ALLOC(c->adj_tmp, nblk * sizeof(double));
  ADJ_ALLOC(c->adj_cxy, c->n_cxy * sizeof(double));
  ADJ_ALLOC(c->adj_E, 2 * N0 * N1 * sizeof(double));
  ADJ_ALLOC(c->adj_rho, N0 * N1 * sizeof(double));
  ADJ_ALLOC(c->adj_g, c->n_f * sizeof(double));
  ADJ_ALLOC(c->adj_grad, c->n_f * sizeof(double));
  ADJ_ALLOC(c->adj_axy, 2 * c->n_f * sizeof(double));
  ADJ_ALLOC(c->adj_scl, 4 * sizeof(double));
  const bool first_mt = c->adj_Mt[0] == nullptr;
  for (int d = 0; d < 4; d++) ADJ_ALLOC(c->adj_Mt[d], (size_t)g.n[d] * (g.n[d] + 4) * sizeof(double));
  for (int d = 0; d < 4; d++) ADJ_ALLOC(c->adj_M[d], (size_t)g.n[d] * (g.n[d] + 4) * sizeof(double));
#undef ADJ_ALLOC
  if (first_mt) {
    // M_d from the product spline kernel on unit vectors: line j = e_j ->
    // coefficients M[.][j], stored as Mt[j][m] (row j of M^T)
    for (int d = 0; d < 4; d++) {
      const int N = g.n[d];
      std::vector<double> id((size_t)N * N, 0.);
      for (int j = 0; j < N; j++) id[(size_t)j * N + j] = 1.;
      double *d_id = nullptr;
      HIP_TRY(c, hipMalloc((void **)&d_id, id.size() * sizeof(double)));
      HIP_TRY(c, hipMemcpy(d_id, id.data(), id.size() * sizeof(double), hipMemcpyHostToDevice));
      vlp::LineMap m{};
      m.N = N;
      m.X0 = N; m.X1 = 1; m.A = 1;
      m.is0 = N; m.is1 = 0; m.ios = 0; m.ils = 1;
      m.os0 = N + 4; m.os1 = 0; m.oos = 0; m.ols = 1;
      vlp::launch_spline_line<double>(d_id, c->adj_Mt[d], m, s);
      const hipError_t e = hipStreamSynchronize(s);
      (void)hipFree(d_id);
      if (e != hipSuccess) return set_err(c, VLP_E_HIP, "line operator");
      // and M_d row-major, transposed on the host
      std::vector<double> mt((size_t)N * (N + 4)), mm(mt.size());
      HIP_TRY(c, hipMemcpy(mt.data(), c->adj_Mt[d], mt.size() * sizeof(double), hipMemcpyDeviceToHost));
      for (int j = 0; j < N; j++)
        for (int q = 0; q < N + 4; q++) mm[(size_t)q * N + j] = mt[(size_t)j * (N + 4) + q];
      HIP_TRY(c, hipMemcpy(c->adj_M[d], mm.data(), mm.size() * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  // idcases 2 and 6 set an analytic E that does not depend on f (VMO/Field.cpp:81-108)
  const bool field_term = g.idcase != 2 && g.idcase != 6;
  if (field_term && !c->adj_k) {
    // kx, ky: E of a unit density at the origin (the field solve drops the
    // mean, so its constant shift does not enter)
    if (hipMalloc((void **)&c->adj_k, 2 * N0 * N1 * sizeof(double)) != hipSuccess)
      return set_err(c, VLP_E_HIP, "hipMalloc failed (adjoint buffers)");
    HIP_TRY(c, hipMemsetAsync(c->rho_part, 0, N0 * N1 * sizeof(double), s));
    const double unit = 1. / (g.dx[2] * g.dx[3]);
    HIP_TRY(c, hipMemcpyAsync(c->rho_part, &unit, sizeof(double), hipMemcpyHostToDevice, s));
    c->rho_outer = 1;
    rc = field(c, c->iter, false, false);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(c->adj_k, c->fb.ex, N0 * N1 * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(c->adj_k + N0 * N1, c->fb.ey, N0 * N1 * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIP_TRY(c, hipStreamSynchronize(s));
  }
  return VLP_OK;
}

// VLP4D_ADJ_TILED=0: the per-output line contraction instead of the LDS tiles.
static bool adj_tiled() {
  if (const char *e = getenv("VLP4D_ADJ_TILED")) return atoi(e) != 0;
  return true;
}

// (d f^{n+1}/d f^n)^T adj_g -> adj_grad at the context's f, on its stream.
static int adjoint_core(vlp_ctx *c) {
  const Geom &g = c->g;
  const long N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const long E0 = N0 + 4, E1 = N1 + 4, E2 = N2 + 4, E3 = N3 + 4;
  hipStream_t s = c->stream;
  const bool field_term = g.idcase != 2 && g.idcase != 6;
  // forward phases of the step from f (f itself untouched)
  run_spline_xy<double>(c);
  run_advect2d<double>(c);
  int rc = field(c, c->iter + 1, false, false);
  if (rc) return rc;
  run_spline_vxvy<double>(c);
  // reproducible form (default): fixed-point coefficient sums, E by gathers
  // (adj_tmp holds the gathers' partial sums until the contractions below)
  const bool blocked = vlp::adj_blocked();
  const vlp::AdjWork wk{c->adj_scl, c->adj_axy, c->adj_tmp};
  HIP_TRY(c, hipMemsetAsync(c->adj_c4s, 0, c->n_c4s * sizeof(double), s));
  if (!blocked) HIP_TRY(c, hipMemsetAsync(c->adj_E, 0, 2 * N0 * N1 * sizeof(double), s));
  vlp::launch_adj_advect4d(g, (const double *)c->c4, c->fb.ex, c->fb.ey, c->cfg.dt, c->adj_g, c->adj_c4s, c->adj_E,
                           field_term, s, blocked ? &wk : nullptr);
  // direct term: (M_x^T M_y^T M_vx^T M_vy^T) of the coefficient adjoint
  vlp::launch_c4_block(g, c->adj_c4s, c->adj_blk, s, blocked ? c->adj_scl : nullptr);
  const bool tiled = adj_tiled();
  vlp::launch_axis_contract(c->adj_Mt[3], (int)E3, (int)N3, c->adj_blk, c->adj_tmp, 1, E0 * E1 * E2, false, s,
                            tiled ? c->adj_M[3] : nullptr);
  vlp::launch_axis_contract(c->adj_Mt[2], (int)E2, (int)N2, c->adj_tmp, c->adj_blk, N3, E0 * E1, false, s,
                            tiled ? c->adj_M[2] : nullptr);
  vlp::launch_axis_contract(c->adj_Mt[1], (int)E1, (int)N1, c->adj_blk, c->adj_tmp, N2 * N3, E0, false, s,
                            tiled ? c->adj_M[1] : nullptr);
  vlp::launch_axis_contract(c->adj_Mt[0], (int)E0, (int)N0, c->adj_tmp, c->adj_grad, N1 * N2 * N3, 1, false, s,
                            tiled ? c->adj_M[0] : nullptr);
  if (field_term) {
    // through E: K^T, then the half step and C_xy^T
    vlp::launch_poisson_T(c->adj_E, c->adj_k, c->adj_k + N0 * N1, c->adj_rho, (int)N0, (int)N1, s);
    HIP_TRY(c, hipMemsetAsync(c->adj_cxy, 0, c->n_cxy * sizeof(double), s));
    vlp::launch_adj_advect2d(c->adj_rho, g, 0.5 * c->cfg.dt, c->adj_cxy, s);
    vlp::launch_axis_contract(c->adj_Mt[1], (int)E1, (int)N1, c->adj_cxy, c->adj_tmp, N2 * N3, E0, false, s,
                              tiled ? c->adj_M[1] : nullptr);
    vlp::launch_axis_contract(c->adj_Mt[0], (int)E0, (int)N0, c->adj_tmp, c->adj_grad, N1 * N2 * N3, 1, true, s,
                              tiled ? c->adj_M[0] : nullptr);
  }
  return check_launch(c, "adjoint");
}

static int adjoint_args(vlp_ctx *c, const double *g_host, double *grad_host, size_t n) {
  int rc = spline_only(c);
  if (rc) return rc;
  if (c->precision != VLP_FP64) return set_err(c, VLP_E_ARG, "the adjoint is implemented for VLP_FP64 contexts");
  if (!g_host || !grad_host || n != c->n_f) return set_err(c, VLP_E_SIZE, "f buffer size mismatch");
  return adjoint_prepare(c);
}

int vlp_adjoint_step(vlp_ctx *c, const double *g_host, double *grad_host, size_t n) {
  int rc = adjoint_args(c, g_host, grad_host, n);
  if (rc) return rc;
  HIP_TRY(c, hipMemcpyAsync(c->adj_g, g_host, c->n_f * sizeof(double), hipMemcpyHostToDevice, c->stream));
  rc = adjoint_core(c);
  if (rc) return rc;
  HIP_TRY(c, hipMemcpyAsync(grad_host, c->adj_grad, c->n_f * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return sync_and_check(c);  // the recomputed phases flag a foot out of the box
}

// Adds the field-energy source of the context's current f to adj_g:
// W(f) = dx dy sum (Ex^2 + Ey^2) of the field solved from f (the density,
// Poisson solve and diagnostics of VMO/Timestep.hpp:77-91), scaled by `w`,
// has dW/df(x, y, v) = dvx dvy (K^T (2 dx dy E))(x, y) for every v.
static int add_field_energy_source(vlp_ctx *c, double w) {
  const Geom &g = c->g;
  const long nxy = (long)g.n[0] * g.n[1];
  hipStream_t s = c->stream;
  run_moments<double>(c);
  int rc = field(c, c->iter, false, false);
  if (rc) return rc;
  vlp::launch_scale_pairs(c->fb.exy, w * 2. * g.dx[0] * g.dx[1], 2 * nxy, c->adj_E, s);
  vlp::launch_poisson_T(c->adj_E, c->adj_k, c->adj_k + nxy, c->adj_rho, g.n[0], g.n[1], s);
  vlp::launch_add_density_adj(c->adj_rho, g.dx[2] * g.dx[3], nxy, (long)c->n_f, c->adj_g, s);
  return check_launch(c, "field-energy source");
}

// Reverse sweep over nsteps steps from the context's f by two-level
// checkpointing -- the forward sweep keeps every C-th state (C = `every`, or
// ceil(sqrt(K)) when every <= 0); the reverse sweep recomputes each segment's states from its
// checkpoint and applies the one-step adjoint backwards through them.  Device
// memory: K/C + C copies of f.  The terminal cotangent is g_host (or
// zero when null); with `energy`, the source dt W(f^{n+k}) of every step k =
// 1..K is added as the sweep passes f^{n+k}, and *J receives
// dt sum_k W(f^{n+k}) from the forward sweep's diagnostics.  The result,
// d/df^n of <g, f^{n+K}> (+ J), is left in adj_g.  The context ends at its
// starting f and iteration, also after a failure part-way through.
static int adjoint_sweep(vlp_ctx *c, int K, int every, const double *g_host, bool energy, double *J) {
  hipStream_t s = c->stream;
  const size_t fb = c->n_f * sizeof(double);
  const int C = every > 0 ? std::min(every, K) : (int)std::ceil(std::sqrt((double)K)), nck = (K + C - 1) / C;
  std::vector<double *> ck(nck, nullptr), seg(C, nullptr);
  auto release = [&]() {
    for (double *p : ck)
      if (p) (void)hipFree(p);
    for (double *p : seg)
      if (p) (void)hipFree(p);
  };
  for (auto *v : {&ck, &seg})
    for (double *&p : *v)
      if (hipMalloc((void **)&p, fb) != hipSuccess) {
        release();
        return set_err(c, VLP_E_HIP, "hipMalloc failed (adjoint checkpoints)");
      }
  const int iter0 = c->iter;
  const double dt = c->cfg.dt;
  auto restore = [&](const double *src, int it) -> int {
    HIP_TRY(c, hipMemcpyAsync(c->f, src, fb, hipMemcpyDeviceToDevice, s));
    c->iter = it;
    return VLP_OK;
  };
  c->err.clear();
  int rc = VLP_OK;
  // forward sweep: checkpoints at 0, C, 2C, ...
  for (int k = 0; k < K && !rc; k++) {
    if (k % C == 0) {
      if (hipMemcpyAsync(ck[k / C], c->f, fb, hipMemcpyDeviceToDevice, s) != hipSuccess)
        rc = set_err(c, VLP_E_HIP, "adjoint: checkpoint copy failed");
    }
    if (!rc) rc = step_once(c);
  }
  if (!rc && energy && J) {  // W_k = 2 (nrjx_k^2 + nrjy_k^2) from the recorded diagnostics
    const int nb = c->cfg.nbiter + 1;
    std::vector<double> h(5 * (size_t)nb);
    if (hipMemcpyAsync(h.data(), c->fb.diag, h.size() * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      rc = set_err(c, VLP_E_HIP, "adjoint: diagnostics copy failed");
    double sum = 0.;
    for (int k = 1; k <= K && !rc; k++) {
      const double a = h[nb + iter0 + k], b = h[2 * (size_t)nb + iter0 + k];
      sum += 2. * (a * a + b * b);
    }
    *J = dt * sum;
  }
  if (!rc) {
    if (g_host) {
      if (hipMemcpyAsync(c->adj_g, g_host, fb, hipMemcpyHostToDevice, s) != hipSuccess)
        rc = set_err(c, VLP_E_HIP, "adjoint: copy of the cotangent failed");
    } else if (hipMemsetAsync(c->adj_g, 0, fb, s) != hipSuccess) {
      rc = set_err(c, VLP_E_HIP, "adjoint: clearing the cotangent failed");
    }
  }
  if (!rc && energy) rc = add_field_energy_source(c, dt);  // f is f^{n+K}
  // reverse sweep, segment by segment
  for (int sg = nck - 1; sg >= 0 && !rc; sg--) {
    const int k0 = sg * C, len = std::min(C, K - k0);
    rc = restore(ck[sg], iter0 + k0);
    for (int j = 0; j < len && !rc; j++) {  // the segment's states
      if (hipMemcpyAsync(seg[j], c->f, fb, hipMemcpyDeviceToDevice, s) != hipSuccess)
        rc = set_err(c, VLP_E_HIP, "adjoint: segment copy failed");
      if (!rc && j + 1 < len) rc = step_once(c);
    }
    for (int j = len - 1; j >= 0 && !rc; j--) {
      rc = restore(seg[j], iter0 + k0 + j);
      if (!rc) rc = adjoint_core(c);
      if (!rc) std::swap(c->adj_g, c->adj_grad);  // the result is the next cotangent
      if (!rc && energy && k0 + j >= 1) rc = add_field_energy_source(c, dt);  // f is f^{n+k0+j}
    }
  }
  // back to the starting state, also after a failure part-way through
  const std::string err = c->err;
  const bool restored = hipStreamSynchronize(s) == hipSuccess &&
                        hipMemcpyAsync(c->f, ck[0], fb, hipMemcpyDeviceToDevice, s) == hipSuccess &&
                        hipStreamSynchronize(s) == hipSuccess;
  c->iter = iter0;
  if (!restored) {
    c->broken = true;
    if (!rc) rc = set_err(c, VLP_E_HIP, "adjoint: restoring the starting state failed");
  } else if (rc) {
    // f and iter are the starting ones again; clear the step's error flag
    if (hipMemsetAsync(c->d_err, 0, sizeof(int), s) == hipSuccess && hipStreamSynchronize(s) == hipSuccess)
      c->broken = false;
    c->err = err;
  }
  if (!rc) rc = sync_and_check(c);
  release();
  if (rc == VLP_E_HIP && c->err.empty()) return set_err(c, VLP_E_HIP, "adjoint sweep failed");
  return rc;
}

int vlp_adjoint_steps(vlp_ctx *c, int nsteps, const double *g_host, double *grad_host, size_t n) {
  int rc = adjoint_args(c, g_host, grad_host, n);
  if (rc) return rc;
  if (nsteps < 1 || c->iter + nsteps > c->cfg.nbiter)
    return set_err(c, VLP_E_ARG, "nsteps must be at least 1 and within nbiter");
  rc = adjoint_sweep(c, nsteps, 0, g_host, false, nullptr);
  if (rc) return rc;
  HIP_TRY(c, hipMemcpyAsync(grad_host, c->adj_g, c->n_f * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return VLP_OK;
}

int vlp_adjoint_field_energy(vlp_ctx *c, int nsteps, int checkpoint_every, const double *vweight, double *J,
                             double *grad_host, size_t n) {
  int rc = spline_only(c);
  if (rc) return rc;
  if (c->precision != VLP_FP64) return set_err(c, VLP_E_ARG, "the adjoint is implemented for VLP_FP64 contexts");
  const size_t nxy = (size_t)c->cfg.nxmax[0] * c->cfg.nxmax[1];
  const size_t nv = (size_t)c->cfg.nxmax[2] * c->cfg.nxmax[3];
  if (!J || !grad_host || n != (vweight ? nxy : c->n_f)) return set_err(c, VLP_E_SIZE, "gradient buffer size mismatch");
  if (nsteps < 1 || c->iter + nsteps > c->cfg.nbiter)
    return set_err(c, VLP_E_ARG, "nsteps must be at least 1 and within nbiter");
  rc = adjoint_prepare(c);
  if (rc) return rc;
  if (!c->adj_k) return set_err(c, VLP_E_ARG, "the field energy of idcases 2 and 6 does not depend on f");
  rc = adjoint_sweep(c, nsteps, checkpoint_every, nullptr, true, J);
  if (rc) return rc;
  hipStream_t s = c->stream;
  if (!vweight) {
    HIP_TRY(c, hipMemcpyAsync(grad_host, c->adj_g, c->n_f * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    return VLP_OK;
  }
  // the velocity moment on the device; the weights and the moment sit in
  // the adjoint's gradient buffer, free after the sweep (N^4 >= nv + nxy)
  double *w = c->adj_grad, *out = c->adj_grad + nv;
  HIP_TRY(c, hipMemcpyAsync(w, vweight, nv * sizeof(double), hipMemcpyHostToDevice, s));
  vlp::launch_vmoment(c->adj_g, w, (long)nxy, (long)nv, out, s);
  rc = check_launch(c, "velocity moment");
  if (rc) return rc;
  HIP_TRY(c, hipMemcpyAsync(grad_host, out, nxy * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  return VLP_OK;
}

int vlp_tuned_tile(const vlp_ctx *c, int *vx_tile, int *plane_buffers) {
  if (!c || !vx_tile || !plane_buffers) return set_err(nullptr, VLP_E_ARG, "null argument");
  *vx_tile = c->ys_tvx;
  *plane_buffers = c->ys_nb;
  return VLP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Ensembles: independent contexts (any grids and precisions, one device)
// advanced together.  Two steps of every member are captured once into one
// HIP graph -- a fork event on the ensemble's stream, each member's kernels
// on its own stream, a join -- so that the members' kernels overlap on the
// GPU and a pair of steps of the whole ensemble is a single graph launch.
// Two steps, because each step ends by swapping f and fnext on the host: a
// pair returns every member to its buffers, and the graph stays valid.  The
// diagnostics index of each member comes from a device counter; error flags
// are sticky and are read once after the replays.
// ---------------------------------------------------------------------------
struct vlp_ensemble {
  std::vector<vlp_ctx *> m;
  int device = 0;
  hipStream_t root = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  int *d_iter = nullptr;  // one counter per member
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<void *> cap_f;  // each member's f when the graph was captured
  std::vector<int> cap_tile;  // and its tuned kernel choices (tile_key)
  double seconds = 0.;
  long steps = 0;
};

namespace {

// The tuned kernel choices of a context, as one number.
int tile_key(const vlp_ctx *c) { return ((c->ys_tvx * 16 + c->ys_nb) * 4 + c->sp_mode[0]) * 4 + c->sp_mode[1]; }

void ensemble_free(vlp_ensemble *e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  for (auto ev : e->join)
    if (ev) (void)hipEventDestroy(ev);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->d_iter) (void)hipFree(e->d_iter);
  if (e->root) (void)hipStreamDestroy(e->root);
  e->exec = nullptr;
  e->graph = nullptr;
}

int ensemble_capture(vlp_ensemble *e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
  if (hipStreamBeginCapture(e->root, hipStreamCaptureModeRelaxed) != hipSuccess)
    return set_err(nullptr, VLP_E_HIP, "hipStreamBeginCapture failed");
  int rc = VLP_OK;
  if (hipEventRecord(e->fork, e->root) != hipSuccess) rc = set_err(nullptr, VLP_E_HIP, "fork event");
  for (size_t i = 0; i < e->m.size() && !rc; i++)
    if (hipStreamWaitEvent(e->m[i]->stream, e->fork, 0) != hipSuccess) rc = set_err(nullptr, VLP_E_HIP, "fork wait");
  for (size_t i = 0; i < e->m.size() && !rc; i++) {
    rc = enqueue_step(e->m[i], e->d_iter + i);
    if (!rc) rc = enqueue_step(e->m[i], e->d_iter + i);
  }
  for (size_t i = 0; i < e->m.size() && !rc; i++)
    if (hipEventRecord(e->join[i], e->m[i]->stream) != hipSuccess ||
        hipStreamWaitEvent(e->root, e->join[i], 0) != hipSuccess)
      rc = set_err(nullptr, VLP_E_HIP, "join");
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(e->root, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ee != hipSuccess || !g) return set_err(nullptr, VLP_E_HIP, "hipStreamEndCapture failed");
  e->graph = g;
  if (hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0) != hipSuccess)
    return set_err(nullptr, VLP_E_HIP, "hipGraphInstantiate failed");
  e->cap_f.resize(e->m.size());
  e->cap_tile.resize(e->m.size());
  for (size_t i = 0; i < e->m.size(); i++) {
    e->cap_f[i] = e->m[i]->f;
    e->cap_tile[i] = tile_key(e->m[i]);
  }
  return VLP_OK;
}

}  // namespace

extern "C" {

int vlp_ensemble_create(vlp_ctx *const *members, int n, vlp_ensemble **out) {
  if (!out) return set_err(nullptr, VLP_E_ARG, "null output pointer");
  *out = nullptr;
  if (!members || n < 1) return set_err(nullptr, VLP_E_ARG, "an ensemble needs at least one member");
  for (int i = 0; i < n; i++) {
    if (!members[i]) return set_err(nullptr, VLP_E_ARG, "null member");
    if (members[i]->broken) return set_err(nullptr, VLP_E_ARG, "a member is in an error state");
    if (members[i]->device != members[0]->device) return set_err(nullptr, VLP_E_ARG, "members on different devices");
    if (members[i]->scheme != VLP_SCHEME_SPLINE)
      return set_err(nullptr, VLP_E_ARG, "ensembles take spline-solver contexts only");
    if (members[i]->owner) return set_err(nullptr, VLP_E_ARG, "a member already belongs to an ensemble");
    for (int j = 0; j < i; j++)
      if (members[j] == members[i]) return set_err(nullptr, VLP_E_ARG, "a context appears twice");
  }
  auto *e = new vlp_ensemble();
  e->m.assign(members, members + n);
  e->device = members[0]->device;
  e->join.assign(n, nullptr);
  bool ok = hipSetDevice(e->device) == hipSuccess &&
            hipStreamCreateWithFlags(&e->root, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) == hipSuccess &&
            hipMalloc((void **)&e->d_iter, n * sizeof(int)) == hipSuccess;
  for (int i = 0; i < n && ok; i++) ok = hipEventCreateWithFlags(&e->join[i], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    ensemble_free(e);
    delete e;
    return set_err(nullptr, VLP_E_HIP, "ensemble set-up failed");
  }
  for (auto *c : e->m) c->owner = e;
  *out = e;
  return VLP_OK;
}

int vlp_ensemble_step(vlp_ensemble *e, int nsteps) {
  if (!e) return set_err(nullptr, VLP_E_ARG, "null ensemble");
  if (nsteps < 0) return set_err(nullptr, VLP_E_ARG, "negative step count");
  for (auto *c : e->m) {
    if (c->broken) return set_err(nullptr, VLP_E_ARG, "a member is in an error state");
    if (c->iter + nsteps > c->cfg.nbiter) return set_err(nullptr, VLP_E_ARG, "step count exceeds a member's nbiter");
  }
  if (hipSetDevice(e->device) != hipSuccess) return set_err(nullptr, VLP_E_HIP, "hipSetDevice failed");
  const auto t0 = std::chrono::steady_clock::now();
  const int pairs = nsteps / 2;
  if (pairs > 0) {
    for (auto *c : e->m)
      if (hipStreamSynchronize(c->stream) != hipSuccess) return set_err(nullptr, VLP_E_HIP, "member stream");
    // the graph holds each member's buffers and tile as captured
    bool fresh = e->exec != nullptr;
    for (size_t i = 0; i < e->m.size() && fresh; i++)
      fresh = e->cap_f[i] == e->m[i]->f && e->cap_tile[i] == tile_key(e->m[i]);
    if (!fresh) {
      const int rc = ensemble_capture(e);
      if (rc) return rc;
    }
    std::vector<int> it(e->m.size());
    for (size_t i = 0; i < e->m.size(); i++) it[i] = e->m[i]->iter;
    if (hipMemcpy(e->d_iter, it.data(), it.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      return set_err(nullptr, VLP_E_HIP, "iteration counters");
    for (int k = 0; k < pairs; k++)
      if (hipGraphLaunch(e->exec, e->root) != hipSuccess) return set_err(nullptr, VLP_E_HIP, "hipGraphLaunch failed");
    if (hipStreamSynchronize(e->root) != hipSuccess) {
      for (auto *c : e->m) c->broken = true;  // the members' states are unknown
      return set_err(nullptr, VLP_E_HIP, "ensemble replay failed");
    }
    // every member has advanced: count the steps on all of them, then read
    // every member's error flag and report the first failure
    for (auto *c : e->m) c->iter += 2 * pairs;
    int first = VLP_OK;
    std::string msg;
    for (auto *c : e->m) {
      const int rc = sync_and_check(c);
      if (rc && !first) {
        first = rc;
        msg = c->err;
      }
    }
    if (first) return set_err(nullptr, first, msg);
  }
  if (nsteps % 2)
    for (auto *c : e->m) {
      const int rc = step_once(c);
      if (rc) return set_err(nullptr, rc, c->err);
    }
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  e->seconds += dt;
  e->steps += nsteps;
  for (auto *c : e->m) {
    c->tsec[VLP_T_MAINLOOP] += dt;
    c->tsec[VLP_T_TOTAL] += dt;
    c->tcalls[VLP_T_MAINLOOP] += nsteps;
  }
  return VLP_OK;
}

int vlp_ensemble_size(const vlp_ensemble *e) { return e ? (int)e->m.size() : -1; }

int vlp_ensemble_destroy(vlp_ensemble *e) {
  if (!e) return VLP_OK;
  (void)hipSetDevice(e->device);
  if (e->root) (void)hipStreamSynchronize(e->root);
  for (auto *c : e->m) c->owner = nullptr;
  ensemble_free(e);
  delete e;
  return VLP_OK;
}

}  // extern "C"
