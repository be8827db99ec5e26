// lagrange.hip -- the Lagrange split-step vlp4d (miniapps/vlp4d/openmp, "VO/").
//
// One step is the Strang split of VO/timestep.hpp:11-55: 1-D advections along
// x and y by dt/2 (displacement dt*v of the cell's own velocity), the density
// of the result and the FFT Poisson solve (field.hip, shared with the spline
// path), 1-D advections along vx and vy by dt (displacement dt*E(x, y)), and x,
// y by dt/2 again.  Every advection is the 6-point order-5 interpolation of
// VO/advect.hpp:72-86 on a periodic axis (velocity too).  Each pass streams f
// once (8 B read + 8 B write per cell in FP64): lanes run along x, so the six
// stencil reads of every lane are coalesced rows that stay in L2 between
// neighbouring lanes, lines or planes.  Each lane computes its stencil (the
// fmod/floor foot and the six weights) once and applies it along the axis the
// weights do not depend on.
#include <cstdlib>

#include "device_common.hpp"
#include "vlp4d_internal.hpp"

namespace vlp {

namespace {

// The order-5 basis of VO/advect.hpp:72-86, px in [2, 3).
__device__ __forceinline__ void lag5_basis(double px, double c[6]) {
  const double pxm2 = px - 2.;
  const double sq = pxm2 * pxm2;
  const double p01 = pxm2 * (pxm2 - 1.);
  c[0] = (-1. / 24) * p01 * (pxm2 + 1.) * (pxm2 - 2.) * (pxm2 - 1.);
  c[1] = (1. / 24.) * p01 * (pxm2 - 2.) * (5 * sq + pxm2 - 8.);
  c[2] = (-1. / 12.) * (pxm2 - 1.) * (pxm2 - 2.) * (pxm2 + 1.) * (5 * sq - 3 * pxm2 - 6.);
  c[3] = (1. / 12.) * pxm2 * (pxm2 + 1.) * (pxm2 - 2.) * (5 * sq - 7 * pxm2 - 4.);
  c[4] = (-1. / 24.) * p01 * (pxm2 + 1.) * (5 * sq - 11 * pxm2 - 2.);
  c[5] = (1. / 24.) * p01 * pxm2 * (pxm2 + 1.) * (pxm2 - 2.);
}

// Foot of a 1-D advection by `dep` on the periodic axis d (VO/advect.hpp:
// 137-154): the six stencil indices, already wrapped into [0, N), and weights.
__device__ __forceinline__ void lag_stencil(const Geom &g, int d, double h, double dep, int j[6], double c[6]) {
  const double hmin = g.minPhy[d], hwidth = g.maxPhy[d] - g.minPhy[d], dh = g.dx[d];
  const double idh = 1. / dh;
  const int N = g.n[d];
  const double hstar = hmin + fmod(hwidth + h - dep - hmin, hwidth);
  const int ip = (int)floor((hstar - hmin) * idh);
  lag5_basis(2. + idh * (hstar - (hmin + ip * dh)), c);
  // the reference wraps with (N + ip - 2 + k) % N; a full modulo keeps any
  // foot (even beyond one box width) in range
  int b = (ip - 2) % N;
  if (b < 0) b += N;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const int q = b + k;
    j[k] = q >= N ? q - N : q;
  }
}

}  // namespace

// Logical block of hardware block b: block b runs on XCD b % 8, so with the
// grid a multiple of 8 each XCD takes one contiguous range of logical blocks
// and the stencil neighbours of a block (the next rows, planes or lines) sit
// in the same L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned grid) {
  return (grid & 7u) ? b : (b & 7u) * (grid >> 3) + (b >> 3);
}

// advect_1D_x (VO/advect.hpp:112-158).  Lane per (x, vx, vy), x fastest; the
// weights depend on x and vx only, so each lane computes them once and runs
// down the y rows of its (vx, vy) plane.
__global__ __launch_bounds__(256) void k_lag_x(const double *__restrict__ in, double *__restrict__ out, Geom g,
                                               double dt) {
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const unsigned t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= (unsigned)N0 * N2 * N3) return;
  const int ix = (int)(t % (unsigned)N0);
  const unsigned r = t / (unsigned)N0;  // ivx + N2 * ivy
  const int ivx = (int)(r % (unsigned)N2);
  const double vx = g.minPhy[2] + ivx * g.dx[2];
  int j[6];
  double c[6];
  lag_stencil(g, 0, g.minPhy[0] + ix * g.dx[0], dt * vx, j, c);
  const double *src = in + (size_t)r * N1 * N0;
  double *dst = out + (size_t)r * N1 * N0 + ix;
#pragma unroll 2
  for (int iy = 0; iy < N1; iy++) {
    const double *row = src + (size_t)iy * N0;
    double s = 0.;
#pragma unroll
    for (int k = 0; k < 6; k++) s += c[k] * row[j[k]];
    dst[(size_t)iy * N0] = s;
  }
}

// advect_1D_y (VO/advect.hpp:160-204): lane per (x, y, vy), looping over vx
// (the weights depend on y and vy only).  With rho_part, also the velocity
// sum of the result over vx: rho_part[vy][y][x] (field_rho, VO/field.cpp:5-29).
__global__ __launch_bounds__(256) void k_lag_y(const double *__restrict__ in, double *__restrict__ out, Geom g,
                                               double dt, double *__restrict__ rho_part) {
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const unsigned t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const unsigned plane = (unsigned)N0 * N1;
  if (t >= plane * N3) return;
  const unsigned pxy = t % plane;
  const int ix = (int)(pxy % (unsigned)N0);
  const int iy = (int)(pxy / (unsigned)N0);
  const int ivy = (int)(t / plane);
  const double vy = g.minPhy[3] + ivy * g.dx[3];
  int j[6];
  double c[6];
  lag_stencil(g, 1, g.minPhy[1] + iy * g.dx[1], dt * vy, j, c);
  const double *src = in + (size_t)ivy * N2 * plane + ix;
  double *dst = out + (size_t)ivy * N2 * plane + pxy;
  double sum = 0.;
#pragma unroll 2
  for (int ivx = 0; ivx < N2; ivx++) {
    const double *p = src + (size_t)ivx * plane;
    double s = 0.;
#pragma unroll
    for (int k = 0; k < 6; k++) s += c[k] * p[j[k] * N0];
    dst[(size_t)ivx * plane] = s;
    sum += s;
  }
  if (rho_part) rho_part[t] = sum;
}

// advect_1D_vx / _vy (VO/advect.hpp:206-301): displacement dt*E(x, y) along
// velocity axis D (2 or 3).  Lane per (x, y, other velocity), x fastest; the
// lane runs along the advected axis.  The shift dt*E is the same for the
// whole line, so the foot of point iv is that of point 0 moved by iv cells:
// one set of weights, and a window of six values that slides by one load per
// output (the reference recomputes the foot per point; the two agree to
// rounding, and where rounding moves a foot across a node the interpolant,
// continuous there, gives the same value).
template <int D>
__global__ __launch_bounds__(256) void k_lag_v(const double *__restrict__ in, double *__restrict__ out, Geom g,
                                               const double *__restrict__ e, double dt) {
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const int ND = D == 2 ? N2 : N3, NO = D == 2 ? N3 : N2;  // advected axis, other axis
  const unsigned t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const unsigned plane = (unsigned)N0 * N1;
  if (t >= plane * NO) return;
  const unsigned pxy = t % plane;
  const int io = (int)(t / plane);
  int j[6];
  double c[6];
  lag_stencil(g, D, g.minPhy[D], dt * e[pxy], j, c);
  // strides (in cells) of the advected and the other velocity axes
  const size_t sD = D == 2 ? plane : (size_t)plane * N2, sO = D == 2 ? (size_t)plane * N2 : plane;
  const double *src = in + pxy + (size_t)io * sO;
  double *dst = out + pxy + (size_t)io * sO;
  double w0 = src[(size_t)j[0] * sD], w1 = src[(size_t)j[1] * sD], w2 = src[(size_t)j[2] * sD];
  double w3 = src[(size_t)j[3] * sD], w4 = src[(size_t)j[4] * sD], w5 = src[(size_t)j[5] * sD];
  int nxt = j[5];  // index of the window's last point
#pragma unroll 8
  for (int iv = 0; iv < ND; iv++) {
    double s = c[0] * w0;
    s += c[1] * w1;
    s += c[2] * w2;
    s += c[3] * w3;
    s += c[4] * w4;
    s += c[5] * w5;
    dst[(size_t)iv * sD] = s;
    nxt = nxt + 1 == ND ? 0 : nxt + 1;
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = w4;
    w4 = w5;
    w5 = src[(size_t)nxt * sD];
  }
}

// x/2 then y/2 through one LDS-resident (x, y) plane (the first and last
// pair of VO/timestep.hpp:11-55 in one pass over f).  Workgroup of 1024 lanes
// per (vy, run of G vx planes).  Per plane: stage it in LDS from registers
// (its loads were issued before the previous plane's y pass); x pass -- wave w
// owns rows w, w + 16, ...: it reads the stencils of a whole row, then writes
// the row back in place (no other wave touches it, and one wave's LDS
// operations execute in order); y pass from LDS to HBM.  The weights come
// from the context's tables (k_lag_tables): the x ones (x, vx) into registers
// once per plane -- a lane's x positions are fixed -- and the y ones (y, vy)
// as wave-uniform reads when a wave's points lie in one row; the LDS holds
// only the plane.  The plane's loads, its LDS copy and the y pass work on
// pairs of points (16 bytes per lane).  Each cell's arithmetic is that of
// k_lag_x followed by k_lag_y.
constexpr int LAG_XY_THREADS = 1024;
constexpr int LAG_XY_MAXM = 16;  // plane points per lane: planes up to 16384 points

struct LagTab {  // stencil of one point: six weights and the wrapped first index
  double w[6];
  int b;
};
static_assert(sizeof(LagTab) == 7 * sizeof(double), "LagTab is copied as 7 doubles");

__device__ __forceinline__ void lag_tab(const Geom &g, int d, double h, double dep, LagTab &t) {
  int j[6];
  lag_stencil(g, d, h, dep, j, t.w);
  t.b = j[0];
}

// The stencils of the x/2 and y/2 passes, fixed for a context: x for every
// (vx, x) -- N2 x N0 entries -- then y for every (vy, y) -- N3 x N1.
__global__ __launch_bounds__(256) void k_lag_tables(LagTab *__restrict__ tab, Geom g, double dt) {
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N2 * N0) {
    const int ivx = i / N0, ix = i % N0;
    lag_tab(g, 0, g.minPhy[0] + ix * g.dx[0], dt * (g.minPhy[2] + ivx * g.dx[2]), tab[i]);
  } else if (i < N2 * N0 + N3 * N1) {
    const int k = i - N2 * N0, ivy = k / N1, iy = k % N1;
    lag_tab(g, 1, g.minPhy[1] + iy * g.dx[1], dt * (g.minPhy[3] + ivy * g.dx[3]), tab[i]);
  }
}

template <int NC>
__global__ __launch_bounds__(LAG_XY_THREADS) void k_lag_xy(const double *__restrict__ in, double *__restrict__ out,
                                                           Geom g, const LagTab *__restrict__ tab, int G,
                                                           int nchunk) {
  extern __shared__ double pl[];  // the (x, y) plane, N0 * N1 doubles
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2];
  const int P = N0 * N1;
  const int ivy = blockIdx.x / nchunk, chunk = blockIdx.x % nchunk;
  const int v0 = chunk * G, v1 = min(N2, v0 + G);
  const int l = threadIdx.x, wave = l >> 6, lane = l & 63;
  constexpr int NW = LAG_XY_THREADS / 64;
  const LagTab *ytab = tab + (size_t)N2 * N0 + (size_t)ivy * N1;
  // points go in pairs (x, x+1), 16 bytes per lane for the plane's loads,
  // its LDS copy, the y pass's reads and its stores (N0 even, checked by the
  // launcher); with N0 a multiple of 128 a wave's 128 points of the y pass
  // lie in one row, whose weights are then wave-uniform
  constexpr int MH = LAG_XY_MAXM / 2;
  const bool row_waves2 = (N0 & 127) == 0;
  const int dix = (2 * LAG_XY_THREADS) % N0, diy = (2 * LAG_XY_THREADS) / N0;
  typedef double d2v __attribute__((ext_vector_type(2)));
  // the plane in flight: loaded during the previous plane's y pass
  d2v r[MH];
  {
    const double *src = in + ((size_t)ivy * N2 + v0) * (size_t)P;
#pragma unroll
    for (int m = 0; m < MH; m++) {
      const int p = 2 * (l + LAG_XY_THREADS * m);
      r[m] = p < P ? *(const d2v *)(src + p) : d2v{0., 0.};
    }
  }
  for (int ivx = v0; ivx < v1; ivx++) {
    const size_t off = ((size_t)ivy * N2 + ivx) * (size_t)P;
    // this lane's x stencils (x = lane + 64 c) for this vx
    LagTab xr[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const int ix = lane + 64 * c;
      if (ix < N0) xr[c] = tab[(size_t)ivx * N0 + ix];
    }
    __syncthreads();  // the previous plane's y pass has read pl
#pragma unroll
    for (int m = 0; m < MH; m++) {
      const int p = 2 * (l + LAG_XY_THREADS * m);
      if (p < P) *(d2v *)(pl + p) = r[m];
    }
    __syncthreads();
    for (int row = wave; row < N1; row += NW) {
      double *pr = pl + row * N0;
      double xv[NC];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const int ix = lane + 64 * c;
        xv[c] = 0.;
        if (ix < N0) {
          double s = 0.;
#pragma unroll
          for (int k = 0; k < 6; k++) {
            int q = xr[c].b + k;
            q = q >= N0 ? q - N0 : q;
            s += xr[c].w[k] * pr[q];
          }
          xv[c] = s;
        }
      }
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const int ix = lane + 64 * c;
        if (ix < N0) pr[ix] = xv[c];
      }
    }
    __syncthreads();
    if (ivx + 1 < v1) {  // next plane's loads, in flight during the y pass
      const double *nx = in + off + (size_t)P;
#pragma unroll
      for (int m = 0; m < MH; m++) {
        const int p = 2 * (l + LAG_XY_THREADS * m);
        if (p < P) r[m] = *(const d2v *)(nx + p);
      }
    }
    double *dst = out + off;
    int ix = (2 * l) % N0, iy = (2 * l) / N0;
#pragma unroll 4
    for (int p = 2 * l; p < P; p += 2 * LAG_XY_THREADS) {
      const int ty = row_waves2 ? __builtin_amdgcn_readfirstlane(iy) : iy;
      const LagTab &t = ytab[ty];
      d2v s = {0., 0.};
#pragma unroll
      for (int k = 0; k < 6; k++) {
        int q = t.b + k;
        q = q >= N1 ? q - N1 : q;
        const d2v v = *(const d2v *)(pl + q * N0 + ix);
        s.x += t.w[k] * v.x;
        s.y += t.w[k] * v.y;
      }
      *(d2v *)(dst + p) = s;
      ix += dix;
      iy += diy;
      if (ix >= N0) {
        ix -= N0;
        iy++;
      }
    }
  }
}

static size_t lag_xy_lds(const Geom &g) { return (size_t)g.n[0] * g.n[1] * sizeof(double); }

// vx runs per vy of the fused x/y pass (0: the plane does not fit; separate
// kernels).  About 1024 workgroups; VLP4D_LAG_FUSED=0 disables the fusion.
int lag_xy_chunks(const Geom &g) {
  if (const char *e = getenv("VLP4D_LAG_FUSED"))
    if (atoi(e) == 0) return 0;
  // rows of up to 128 points (two 64-lane chunks of x weights in registers)
  if (g.n[0] > 128 || (g.n[0] & 1) || (long)g.n[0] * g.n[1] > (long)LAG_XY_THREADS * LAG_XY_MAXM ||
      lag_xy_lds(g) > 160 * 1024)
    return 0;
  const int want = (1024 + g.n[3] - 1) / g.n[3];
  const int G = (g.n[2] + want - 1) / want;  // vx planes per workgroup
  return (g.n[2] + G - 1) / G;
}

// Initial conditions of VO/init.cpp:144-285 (TSI20 differs from the spline
// miniapp's).
__global__ __launch_bounds__(256) void k_lag_init(double *__restrict__ f, Geom g) {
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2];
  const long ncell = (long)N0 * N1 * N2 * g.n[3];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell) return;
  const int ix = (int)(t % N0), iy = (int)((t / N0) % N1), ivx = (int)((t / ((long)N0 * N1)) % N2),
            ivy = (int)(t / ((long)N0 * N1 * N2));
  const double x = g.minPhy[0] + ix * g.dx[0], y = g.minPhy[1] + iy * g.dx[1];
  const double vx = g.minPhy[2] + ivx * g.dx[2], vy = g.minPhy[3] + ivy * g.dx[3];
  double v = 0.;
  if (g.idcase == 2 || g.idcase == 5) {  // CYL02 (x, vx) / CYL05 (y, vy)
    const double PERIOD = 0.5 * M_PI, AMPLI = 4, cc = 0.50 * (6. / 16.), rc = 0.50 * (4. / 16.);
    const double xx = g.idcase == 2 ? x : y, vv = g.idcase == 2 ? vx : vy;
    double hv = 0.0, hx = 0.0;
    if ((vv <= cc + rc) && (vv >= cc - rc)) hv = cos(PERIOD * ((vv - cc) / rc));
    else if ((vv <= -cc + rc) && (vv >= -cc - rc)) hv = -cos(PERIOD * ((vv + cc) / rc));
    if ((xx <= cc + rc) && (xx >= cc - rc)) hx = cos(PERIOD * ((xx - cc) / rc));
    else if ((xx <= -cc + rc) && (xx >= -cc - rc)) hx = -cos(PERIOD * ((xx + cc) / rc));
    v = AMPLI * hx * hv;
  } else if (g.idcase == 10) {  // SLD10
    const double sum = vx * vx + vy * vy;
    v = (1. / (2 * M_PI)) * exp(-0.5 * (sum)) * (1 + 0.05 * (cos(0.5 * x) * cos(0.5 * y)));
  } else {  // TSI20
    const double xi = 0.90, alpha = 0.05;
    const double sum = vx * vx + vy * vy;
    v = (1 + alpha * (cos(0.5 * x))) * 1 / (2 * M_PI) * ((2 - 2 * xi) / (3 - 2 * xi)) * (1 + .5 * vx * vx / (1 - xi)) *
        exp(-.5 * sum);
  }
  f[t] = v;
}

static unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }

void launch_lag_init(const Geom &g, double *f, hipStream_t s) {
  const long n = (long)g.n[0] * g.n[1] * g.n[2] * g.n[3];
  hipLaunchKernelGGL(k_lag_init, dim3(blocks(n)), dim3(256), 0, s, f, g);
}

void launch_lag_x(const Geom &g, const double *in, double *out, double dt, hipStream_t s) {
  const long n = (long)g.n[0] * g.n[2] * g.n[3];
  hipLaunchKernelGGL(k_lag_x, dim3(blocks(n)), dim3(256), 0, s, in, out, g, dt);
}

size_t lag_tables_bytes(const Geom &g) {
  return ((size_t)g.n[2] * g.n[0] + (size_t)g.n[3] * g.n[1]) * sizeof(LagTab);
}

void launch_lag_tables(const Geom &g, void *tab, double dt, hipStream_t s) {
  const int n = g.n[2] * g.n[0] + g.n[3] * g.n[1];
  hipLaunchKernelGGL(k_lag_tables, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (LagTab *)tab, g, dt);
}

bool launch_lag_xy(const Geom &g, const double *in, double *out, const void *tab, int nchunk, hipStream_t s) {
  if (nchunk <= 0 || !tab) return false;
  const int G = (g.n[2] + nchunk - 1) / nchunk;
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void *)k_lag_xy<1>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipFuncSetAttribute((const void *)k_lag_xy<2>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr = true;
  }
  const dim3 grid((unsigned)(g.n[3] * nchunk)), block(LAG_XY_THREADS);
  const size_t lds = lag_xy_lds(g);
  const LagTab *t = (const LagTab *)tab;
  if (g.n[0] <= 64)  // 64-point chunks of an x row
    hipLaunchKernelGGL(k_lag_xy<1>, grid, block, lds, s, in, out, g, t, G, nchunk);
  else
    hipLaunchKernelGGL(k_lag_xy<2>, grid, block, lds, s, in, out, g, t, G, nchunk);
  return true;
}

void launch_lag_y(const Geom &g, const double *in, double *out, double dt, double *rho_part, hipStream_t s) {
  const long n = (long)g.n[0] * g.n[1] * g.n[3];
  hipLaunchKernelGGL(k_lag_y, dim3(blocks(n)), dim3(256), 0, s, in, out, g, dt, rho_part);
}

void launch_lag_v(const Geom &g, int d, const double *in, double *out, const double *e, double dt, hipStream_t s) {
  const long n = (long)g.n[0] * g.n[1] * g.n[d == 2 ? 3 : 2];
  if (d == 2) hipLaunchKernelGGL(k_lag_v<2>, dim3(blocks(n)), dim3(256), 0, s, in, out, g, e, dt);
  else hipLaunchKernelGGL(k_lag_v<3>, dim3(blocks(n)), dim3(256), 0, s, in, out, g, e, dt);
}

}  // namespace vlp
