// kernels.hip -- CDNA4 (gfx950) kernels of the vlp4d time step.
//
// Numerics follow the reference OpenMP backend (miniapps/vlp4d_mpi/openmp,
// "VMO/"); each kernel cites the lines it reproduces.  Storage type T is
// double (plain path) or float (FP32-storage path); feet, rho, the Poisson
// solve and the diagnostics are always computed in double.
#include "vlp4d_internal.hpp"

#include <cmath>
#include <vector>

namespace vlp {

// ---------------------------------------------------------------------------
// constants of the spline prefilter, VMO/Spline.hpp:21,122-123
// ---------------------------------------------------------------------------
struct SplineConsts {
  double alpha, beta, sqrt3;
};

#pragma clang fp contract(off)
static SplineConsts spline_consts() {
  SplineConsts c;
  c.sqrt3 = std::sqrt(3.0);
  c.alpha = c.sqrt3 - 2;
  c.beta = c.sqrt3 * (1 - c.alpha * c.alpha);
  return c;
}
#pragma clang fp contract(on)

// Prefilter of one periodic line, VMO/Spline.hpp:129-161 with the one-rank
// halo of VMO/Communication.hpp:756-898 computed in place of the halo cells:
//   g[-3]   = sum_{j=1..8} alpha^j f[N-2-j]      (boundary_condition, min side)
//   g[-2..N+1] = f[(i mod N)]                     (pack/unpack periodic copy)
//   g[N+2]  = sum_{j=1..8} alpha^j f[1+j]         (boundary_condition, max side)
// buf[(i+2)*st] holds f[i] (i = 0..N-1) on entry and c[i] (i = -2..N+1) on exit.
// Arithmetic in double whatever the storage type.
template <typename T>
__device__ __forceinline__ void spline_recursion(T *buf, int st, int N, const SplineConsts sc) {
  const double a = sc.alpha, b = sc.beta, s3 = sc.sqrt3;
#define FV(i) ((double)buf[((i) + 2) * st])
  double gm3 = 0., gp2 = 0., ap = a;
  for (int j = 1; j <= 8; j++) {  // MMAX = 8, VMO/Config.hpp:8
    gm3 += FV(N - 2 - j) * ap;
    gp2 += FV(1 + j) * ap;
    ap *= a;
  }
  const double f0 = FV(0), f1 = FV(1);
  // anti-causal initial value, VMO/Spline.hpp:144-149
  double fnend = gp2 + f1;
  double ak = a;
  fnend += f0 * ak;  // g[N] = f[0]
  ak *= a;
  for (int k = 2; k <= N + 1; k++) {  // g[N-1] .. g[0]
    fnend += FV(N + 1 - k) * ak;
    ak *= a;
  }
  fnend += FV(N - 1) * ak;  // g[-1]
  ak *= a;
  fnend += FV(N - 2) * ak;  // g[-2]
  // causal pass, VMO/Spline.hpp:132-135
  double cp = gm3 + FV(N - 2);
  buf[0] = (T)cp;
  cp = FV(N - 1) + a * cp;
  buf[st] = (T)cp;
  for (int p = 0; p < N; p++) {
    cp = FV(p) + a * cp;
    buf[(p + 2) * st] = (T)cp;
  }
  cp = f0 + a * cp;
  buf[(N + 2) * st] = (T)cp;
  // anti-causal pass, VMO/Spline.hpp:151-155
  double c = s3 * fnend;
  buf[(N + 3) * st] = (T)c;
  for (int p = N; p >= -2; p--) {
    c = b * (double)buf[(p + 2) * st] + a * c;
    buf[(p + 2) * st] = (T)c;
  }
#undef FV
}

// ---------------------------------------------------------------------------
// Spline along x (contiguous lines).  A 64-lane workgroup stages LINES
// consecutive lines in LDS with coalesced loads, each lane runs the
// recursion of one line, and the N+4 coefficients go back out coalesced.
// ---------------------------------------------------------------------------
template <typename T, int LINES>
__global__ __launch_bounds__(64) void k_spline_x(const T *__restrict__ f, T *__restrict__ tx, int N, long nlines,
                                                 SplineConsts sc) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const int P = (N + 4) | 1;  // odd row stride: conflict-free column access
  const long l0 = (long)blockIdx.x * LINES;
  const int nl = (int)((nlines - l0) < LINES ? (nlines - l0) : LINES);
  const T *src = f + l0 * N;
  for (int l = 0; l < nl; l++)
    for (int x = threadIdx.x; x < N; x += 64) lds[l * P + x + 2] = src[(long)l * N + x];
  __syncthreads();
  if ((int)threadIdx.x < nl) spline_recursion(lds + threadIdx.x * P, 1, N, sc);
  __syncthreads();
  const int M = N + 4;
  T *dst = tx + l0 * M;
  for (int l = 0; l < nl; l++)
    for (int x = threadIdx.x; x < M; x += 64) dst[(long)l * M + x] = lds[l * P + x];
}

template <typename T> void launch_spline_x(const Geom &g, const T *f, T *tx, hipStream_t s) {
  constexpr int LINES = 64;
  const int N = g.n[0];
  const long nlines = (long)g.n[1] * g.n[2] * g.n[3];
  const size_t lds = (size_t)LINES * ((N + 4) | 1) * sizeof(T);
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void *)k_spline_x<T, LINES>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr = true;
  }
  const long nblk = (nlines + LINES - 1) / LINES;
  hipLaunchKernelGGL((k_spline_x<T, LINES>), dim3((unsigned)nblk), dim3(64), lds, s, f, tx, N, nlines, spline_consts());
}

// ---------------------------------------------------------------------------
// Spline along a strided dimension.  One lane per line; lanes walk the
// contiguous inner index so every load/store of a wave is coalesced; the
// line lives in an LDS column private to the lane.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64) void k_spline_line(const T *__restrict__ in, T *__restrict__ out, int N, long X,
                                                    long A, long in_ls, long in_os, long out_ls, long out_os,
                                                    SplineConsts sc) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const long t = (long)blockIdx.x * 64 + threadIdx.x;
  if (t >= X * A) return;
  const long o = t / X, i = t - o * X;
  const T *src = in + i + o * in_os;
  T *buf = lds + threadIdx.x;
  for (int k = 0; k < N; k++) buf[(k + 2) * 64] = src[k * in_ls];
  spline_recursion(buf, 64, N, sc);
  T *dst = out + i + o * out_os;
  for (int k = 0; k < N + 4; k++) dst[k * out_ls] = buf[k * 64];
}

template <typename T>
void launch_spline_line(const T *in, T *out, int N, long X, long A, long in_ls, long in_os, long out_ls, long out_os,
                        hipStream_t s) {
  const size_t lds = (size_t)64 * (N + 4) * sizeof(T);
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void *)k_spline_line<T>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr = true;
  }
  const long nthr = X * A;
  hipLaunchKernelGGL(k_spline_line<T>, dim3((unsigned)((nthr + 63) / 64)), dim3(64), lds, s, in, out, N, X, A, in_ls,
                     in_os, out_ls, out_os, spline_consts());
}

// ---------------------------------------------------------------------------
// B-spline weights, VMO/Advection.hpp:131-143 / 266-274
// ---------------------------------------------------------------------------
template <typename A>
__device__ __forceinline__ void bspline_eta(A w, A eta[4]) {
  const A e3 = (A)(1. / 6.) * w * w * w;
  const A e0 = (A)(1. / 6.) + (A)0.5 * w * (w - (A)1.) - e3;
  const A e2 = w + e0 - (A)2. * e3;
  const A e1 = (A)1. - e0 - e2 - e3;
  eta[0] = e0;
  eta[1] = e1;
  eta[2] = e2;
  eta[3] = e3;
}

__device__ __forceinline__ double block_sum256(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += sh[w];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// 2-D half-step advection fused with the density reduction:
// VMO/Advection.hpp:299-406 (spline interp_2D :126-163) and VMO/Field.cpp:37-53.
// One lane per (x, y, vy), looping over vx: rho_part[vy][y][x] = sum_vx f_half,
// the same summation order as field_rho's inner loop.  f_half is not stored.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_advect2d_rho(const T *__restrict__ cxy, Geom g, double dt,
                                                      double *__restrict__ rho_part, int *__restrict__ err) {
  using A = T;
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)N0 * N1 * N3) return;
  const int ix = (int)(t % N0);
  const long r = t / N0;
  const int iy = (int)(r % N1);
  const int ivy = (int)(r / N1);
  const long E0 = N0 + 4, E1 = N1 + 4;
  const double y = g.minPhy[1] + iy * g.dx[1];
  const double vy = g.minPhy[3] + ivy * g.dx[3];
  const double depy = dt * vy;
  const double fy = ((y - depy) - g.minPhy[1]) * g.inv_dx[1];
  const int jy0 = (int)floor(fy);
  A etay[4];
  bspline_eta<A>((A)(fy - (double)jy0), etay);
  const bool ybad = (jy0 < -1 || jy0 > N1 - 1);
  const double x = g.minPhy[0] + ix * g.dx[0];
  double sum = 0.;
  int bad = 0;
  for (int ivx = 0; ivx < N2; ivx++) {
    const double vx = g.minPhy[2] + ivx * g.dx[2];
    const double depx = dt * vx;
    const double fx = ((x - depx) - g.minPhy[0]) * g.inv_dx[0];
    const int jx0 = (int)floor(fx);
    A etax[4];
    bspline_eta<A>((A)(fx - (double)jx0), etax);
    A val = 0;
    if (ybad || jx0 < -1 || jx0 > N0 - 1) {
      bad = 1;  // interp_2D returns 1 and leaves 0 (VMO/Advection.hpp:148-151)
    } else {
      // position jx0-1 sits at storage index jx0+1
      const T *p = cxy + ((long)ivx + (long)N2 * ivy) * E0 * E1 + (jx0 + 1) + E0 * (jy0 + 1);
      for (int jy = 0; jy < 4; jy++) {
        A s = 0;
        for (int jx = 0; jx < 4; jx++) s += etax[jx] * (A)p[jx + E0 * jy];
        val += etay[jy] * s;
      }
    }
    sum += (double)val;
  }
  rho_part[ix + (long)N0 * (iy + (long)N1 * ivy)] = sum;
  if (bad) atomicOr(err, 1);
}

template <typename T>
void launch_advect2d_rho(const Geom &g, const T *cxy, double dt_half, double *rho_part, int *err, hipStream_t s) {
  const long nthr = (long)g.n[0] * g.n[1] * g.n[3];
  hipLaunchKernelGGL(k_advect2d_rho<T>, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, s, cxy, g, dt_half,
                     rho_part, err);
}

// ---------------------------------------------------------------------------
// 4-D advection fused with rho and the L2 norm:
// computeFeet VMO/Advection.hpp:166-218, interp_4D :221-296, advect_4D
// :409-498, field_rho VMO/Field.cpp:37-53, computeL2norm VMO/Diags.cpp:50-76.
// Lane per (x, y, vy), loop over vx.
// ---------------------------------------------------------------------------
struct FeetConsts {
  double rxwidth[4];
  double locmin[4], locmax[4];  // VMO/Advection.hpp:432-433
};

#pragma clang fp contract(off)
static FeetConsts feet_consts(const Geom &g) {
  FeetConsts f;
  for (int j = 0; j < 4; j++) {
    f.rxwidth[j] = g.maxPhy[j] - g.minPhy[j];
    f.locmin[j] = g.minPhy[j] + 0 * g.dx[j] - g.dx[j];
    f.locmax[j] = g.minPhy[j] + (g.n[j] - 1) * g.dx[j] + g.dx[j];
  }
  return f;
}
#pragma clang fp contract(on)

template <typename T>
__global__ __launch_bounds__(256) void k_advect4d(const T *__restrict__ c4, const double *__restrict__ ex,
                                                  const double *__restrict__ ey, Geom g, FeetConsts fc, double dt,
                                                  T *__restrict__ fout, double *__restrict__ rho_part,
                                                  double *__restrict__ l2_part, int *__restrict__ err) {
  using A = T;
  __shared__ double sh[8];
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = t < (long)N0 * N1 * N3;
  double l2 = 0.;
  int bad = 0;
  if (active) {
    const int ix = (int)(t % N0);
    const long r = t / N0;
    const int iy = (int)(r % N1);
    const int ivy = (int)(r / N1);
    const long s1 = N0 + 4, s2 = s1 * (N1 + 4), s3 = s2 * (N2 + 4);
    const double x = g.minPhy[0] + ix * g.dx[0];
    const double y = g.minPhy[1] + iy * g.dx[1];
    const double vy = g.minPhy[3] + ivy * g.dx[3];
    double rho = 0.;
    for (int ivx = 0; ivx < N2; ivx++) {
      const double vx = g.minPhy[2] + ivx * g.dx[2];
      // computeFeet
      double xt0 = x - 0.5 * dt * vx;
      double xt1 = y - 0.5 * dt * vy;
      xt0 = g.inv_dx[0] * (xt0 - g.minPhy[0]);
      xt1 = g.inv_dx[1] * (xt1 - g.minPhy[1]);
      const int ip0 = (int)(round(xt0) - 1);
      const int ip1 = (int)(round(xt1) - 1);
      const double px = xt0 - ip0, py = xt1 - ip1;
      const double cx[3] = {.5 * (px - 1.) * (px - 2.), -1. * (px) * (px - 2.), .5 * (px) * (px - 1.)};
      const double cy[3] = {.5 * (py - 1.) * (py - 2.), -1. * (py) * (py - 2.), .5 * (py) * (py - 1.)};
      double ftmp1 = 0., ftmp2 = 0.;
      for (int k1 = 0; k1 <= 2; k1++) {
        // (N + i) % N of the reference, made non-negative for any i
        const int jy = ((ip1 + k1) % N1 + N1) % N1;
        for (int k0 = 0; k0 <= 2; k0++) {
          const int jx = ((ip0 + k0) % N0 + N0) % N0;
          const double cc = cx[k0] * cy[k1];
          ftmp1 += cc * ex[jx + (long)N0 * jy];
          ftmp2 += cc * ey[jx + (long)N0 * jy];
        }
      }
      const double xt2 = vx - 0.5 * dt * ftmp1;
      const double xt3 = vy - 0.5 * dt * ftmp2;
      double xs[4];
      xs[0] = x - dt * xt2;
      xs[1] = y - dt * xt3;
      {
        const double a = vx - dt * ftmp1, hi = g.minPhy[2] + fc.rxwidth[2], lo = g.minPhy[2];
        const double m = (hi < a) ? hi : a;
        xs[2] = (m < lo) ? lo : m;
      }
      {
        const double a = vy - dt * ftmp2, hi = g.minPhy[3] + fc.rxwidth[3], lo = g.minPhy[3];
        const double m = (hi < a) ? hi : a;
        xs[3] = (m < lo) ? lo : m;
      }
      // VMO/Advection.hpp:456-460; written so that a NaN foot also counts as bad
      int cbad = 0;
      for (int j = 0; j < 4; j++) cbad += !(xs[j] >= fc.locmin[j] && xs[j] <= fc.locmax[j]);
      A val = 0;
      if (cbad) {
        bad = 1;
      } else {
        int ip[4];
        A eta[4][4];
        long off[4][4];
        const long stride[4] = {1, s1, s2, s3};
        for (int j = 0; j < 4; j++) {
          const double fj = g.inv_dx[j] * (xs[j] - g.minPhy[j]);
          ip[j] = (int)floor(fj);
          bspline_eta<A>((A)(fj - ip[j]), eta[j]);
          // Stencil positions ip-1..ip+2 clamped onto the stored range -2..N+1.
          // Inside the bound check a position outside it (N+2, or -3 after
          // rounding at the lower edge) carries a weight of 0 or ~1e-16, so
          // the clamp only guarantees in-bounds reads.
          for (int k = 0; k < 4; k++) {
            int p = ip[j] - 1 + k;
            p = p < -2 ? -2 : (p > g.n[j] + 1 ? g.n[j] + 1 : p);
            off[j][k] = (long)(p + 2) * stride[j];
          }
        }
        const T *base = c4;
        const long *o0 = off[0], *o1 = off[1], *o2 = off[2], *o3 = off[3];
        A ftmp1 = 0;
        for (int k3 = 0; k3 < 4; k3++) {
          for (int k2 = 0; k2 < 4; k2++) {
            A ftmp2 = 0;
            const T *p23 = base + o3[k3] + o2[k2];
            for (int k1 = 0; k1 < 4; k1++) {
              const T *p = p23 + o1[k1];
              A ftmp3 = 0;
              for (int k0 = 0; k0 < 4; k0++) ftmp3 += eta[0][k0] * (A)p[o0[k0]];
              ftmp2 += ftmp3 * eta[1][k1];
            }
            ftmp1 += ftmp2 * eta[2][k2] * eta[3][k3];
          }
        }
        val = ftmp1;
      }
      fout[ix + (long)N0 * (iy + (long)N1 * (ivx + (long)N2 * ivy))] = (T)val;
      rho += (double)val;
      l2 += (double)val * (double)val;
    }
    rho_part[ix + (long)N0 * (iy + (long)N1 * ivy)] = rho;
  }
  const double tot = block_sum256(l2, sh);
  if (threadIdx.x == 0) l2_part[blockIdx.x] = tot;
  if (bad) atomicOr(err, 2);
}

long l2_partials(const Geom &g) { return ((long)g.n[0] * g.n[1] * g.n[3] + 255) / 256; }

template <typename T>
void launch_advect4d(const Geom &g, const T *c4, const double *ex, const double *ey, double dt, T *fout,
                     double *rho_part, double *l2_part, int *err, hipStream_t s) {
  hipLaunchKernelGGL(k_advect4d<T>, dim3((unsigned)l2_partials(g)), dim3(256), 0, s, c4, ex, ey, g, feet_consts(g), dt,
                     fout, rho_part, l2_part, err);
}

// moments of a given f: rho partials (sum over vx) and L2 partials
template <typename T>
__global__ __launch_bounds__(256) void k_moments(const T *__restrict__ f, Geom g, double *__restrict__ rho_part,
                                                 double *__restrict__ l2_part) {
  __shared__ double sh[8];
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double l2 = 0.;
  if (t < (long)N0 * N1 * N3) {
    const int ix = (int)(t % N0);
    const long r = t / N0;
    const int iy = (int)(r % N1);
    const int ivy = (int)(r / N1);
    double rho = 0.;
    for (int ivx = 0; ivx < N2; ivx++) {
      const double v = (double)f[ix + (long)N0 * (iy + (long)N1 * (ivx + (long)N2 * ivy))];
      rho += v;
      l2 += v * v;
    }
    rho_part[ix + (long)N0 * (iy + (long)N1 * ivy)] = rho;
  }
  const double tot = block_sum256(l2, sh);
  if (threadIdx.x == 0) l2_part[blockIdx.x] = tot;
}

template <typename T> void launch_moments(const Geom &g, const T *f, double *rho_part, double *l2_part, hipStream_t s) {
  hipLaunchKernelGGL(k_moments<T>, dim3((unsigned)l2_partials(g)), dim3(256), 0, s, f, g, rho_part, l2_part);
}

// ---------------------------------------------------------------------------
// Initial conditions, VMO/Init.cpp:126-277 (f must be zero on entry).
// ---------------------------------------------------------------------------
template <typename T> __global__ __launch_bounds__(256) void k_init_case(T *__restrict__ f, Geom g) {
  const long n = (long)g.n[0] * g.n[1] * g.n[2] * g.n[3];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int ix = (int)(t % g.n[0]);
  long r = t / g.n[0];
  const int iy = (int)(r % g.n[1]);
  r /= g.n[1];
  const int ivx = (int)(r % g.n[2]);
  const int ivy = (int)(r / g.n[2]);
  double val = 0.;
  switch (g.idcase) {
    case 2:
    case 5: {  // CYL02 / CYL05, VMO/Init.cpp:152-224
      const double PI = M_PI, AMPLI = 4, PERIOD = 0.5 * PI;
      const double cc = 0.50 * (6. / 16.), rc = 0.50 * (4. / 16.);
      double xx, vv;
      if (g.idcase == 2) {
        xx = g.minPhy[0] + ix * g.dx[0];
        vv = g.minPhy[2] + ivx * g.dx[2];
      } else {
        // CYL05 loops stop one short in ivx, iy, ix (VMO/Init.cpp:196-199)
        if (ivx >= g.n[2] - 1 || iy >= g.n[1] - 1 || ix >= g.n[0] - 1) return;
        xx = g.minPhy[1] + iy * g.dx[1];
        vv = g.minPhy[3] + ivy * g.dx[3];
      }
      double hv = 0.0, hx = 0.0;
      if ((vv <= cc + rc) && (vv >= cc - rc)) hv = cos(PERIOD * ((vv - cc) / rc));
      else if ((vv <= -cc + rc) && (vv >= -cc - rc)) hv = -cos(PERIOD * ((vv + cc) / rc));
      if ((xx <= cc + rc) && (xx >= cc - rc)) hx = cos(PERIOD * ((xx - cc) / rc));
      else if ((xx <= -cc + rc) && (xx >= -cc - rc)) hx = -cos(PERIOD * ((xx + cc) / rc));
      val = AMPLI * hx * hv;
      break;
    }
    case 10: {  // SLD10, VMO/Init.cpp:226-245
      const double vy = g.minPhy[3] + ivy * g.dx[3];
      const double vx = g.minPhy[2] + ivx * g.dx[2];
      const double y = g.minPhy[1] + iy * g.dx[1];
      const double x = g.minPhy[0] + ix * g.dx[0];
      const double sum = (vx * vx + vy * vy);
      val = (1. / (2 * M_PI)) * exp(-0.5 * (sum)) * (1 + 0.05 * (cos(0.5 * x) * cos(0.5 * y)));
      break;
    }
    case 20: {  // TSI20, VMO/Init.cpp:247-277
      const double vd = 2.4, vthx = 1.0, vthy = 0.5, alphax = 0.05, alphay = 0.25, kx = 0.2, ky = 0.2;
      const double vy = g.minPhy[3] + ivy * g.dx[3];
      const double vx = g.minPhy[2] + ivx * g.dx[2];
      const double yy = g.minPhy[1] + iy * g.dx[1];
      const double xx = g.minPhy[0] + ix * g.dx[0];
      const double a1 = (vx - vd) / vthx, a2 = vy / vthy, a3 = (vx + vd) / vthx;
      val = 1. / (4 * M_PI * vthx * vthy) * (1 - alphax * sin(kx * xx) - alphay * sin(ky * yy)) *
            (exp(-.5 * (a1 * a1 + a2 * a2)) + exp(-.5 * (a3 * a3 + a2 * a2)));
      break;
    }
    default:
      return;
  }
  f[t] = (T)val;
}

template <typename T> void launch_init_case(const Geom &g, T *f, hipStream_t s) {
  const long n = (long)g.n[0] * g.n[1] * g.n[2] * g.n[3];
  hipLaunchKernelGGL(k_init_case<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, f, g);
}

// ---------------------------------------------------------------------------
// Field solve, FP64.  rho = dvx*dvy*sum_vy rho_part (VMO/Field.cpp:37-53),
// rho -= 1 for idcase 10/20 (VMO/Field.cpp:110-126), then solve_poisson_fftw
// (VMO/Efield.cpp:54-106) with the FFT conventions of lib/OpenMP_FFT.hpp:
// unnormalised r2c along x, c2c along y, normalisation in k-space, inverse
// c2c along y, c2r along x (imaginary parts of DC/Nyquist ignored).  The
// transforms are direct DFTs with exactly reduced twiddle tables.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rho_rfftx(const double *__restrict__ rho_part, Geom g, FieldBufs b,
                                                   int sub1) {
  extern __shared__ double row[];
  const int N0 = g.n[0], N1 = g.n[1], N3 = g.n[3];
  const int iy = blockIdx.x;
  const int nxh = N0 / 2 + 1;
  for (int ix = threadIdx.x; ix < N0; ix += blockDim.x) {
    double s = 0.;
    for (int ivy = 0; ivy < N3; ivy++) s += rho_part[ix + (long)N0 * (iy + (long)N1 * ivy)];
    double r = s * g.dx[2] * g.dx[3];
    if (sub1) r -= 1.;
    row[ix] = r;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nxh; k += blockDim.x) {
    double sr = 0., si = 0.;
    int m = 0;
    for (int j = 0; j < N0; j++) {
      sr += row[j] * b.twx[2 * m];
      si += row[j] * b.twx[2 * m + 1];
      m += k;
      if (m >= N0) m -= N0;
    }
    b.spec1[2 * (k + (long)nxh * iy)] = sr;
    b.spec1[2 * (k + (long)nxh * iy) + 1] = si;
  }
}

__global__ __launch_bounds__(256) void k_poisson_y(Geom g, FieldBufs b) {
  extern __shared__ double sh[];
  const int N0 = g.n[0], N1 = g.n[1];
  const int nxh = N0 / 2 + 1, nyh = N1 / 2 + 1;
  const int k = blockIdx.x;
  double *col = sh;              // N1 complex
  double *hat = sh + 2 * N1;     // 3 * N1 complex: phi, ex, ey
  for (int iy = threadIdx.x; iy < N1; iy += blockDim.x) {
    col[2 * iy] = b.spec1[2 * (k + (long)nxh * iy)];
    col[2 * iy + 1] = b.spec1[2 * (k + (long)nxh * iy) + 1];
  }
  __syncthreads();
  const double kx0 = 2 * M_PI / g.maxPhy[0];
  const double ky0 = 2 * M_PI / g.maxPhy[1];
  const double normcoeff = 1.0 / (double)(N0 * N1);
  const double kx = k * kx0;
  for (int q = threadIdx.x; q < N1; q += blockDim.x) {
    double cr = 0., ci = 0.;
    int m = 0;
    for (int j = 0; j < N1; j++) {
      const double wr = b.twy[2 * m], wi = b.twy[2 * m + 1];
      cr += col[2 * j] * wr - col[2 * j + 1] * wi;
      ci += col[2 * j] * wi + col[2 * j + 1] * wr;
      m += q;
      if (m >= N1) m -= N1;
    }
    double pr, pi, er, ei, fr, fi;
    if (q == 0) {  // VMO/Efield.cpp:76-81, filter of :38-43
      const double kxf = k * (2 * M_PI / g.maxPhy[0]);
      const double filter = (k == 0) ? 0. : 1. / (kxf * kxf);
      er = ((kx * ci) * filter) * normcoeff;
      ei = ((-kx * cr) * filter) * normcoeff;
      fr = 0.;
      fi = 0.;
      pr = (cr * filter) * normcoeff;
      pi = (ci * filter) * normcoeff;
    } else {  // VMO/Efield.cpp:83-100
      const double ky = (q < nyh) ? q * ky0 : (q - N1) * ky0;
      const double k2 = kx * kx + ky * ky;
      const double sx = kx / k2, sy = ky / k2;
      er = (sx * ci) * normcoeff;
      ei = (-sx * cr) * normcoeff;
      fr = (sy * ci) * normcoeff;
      fi = (-sy * cr) * normcoeff;
      pr = (cr / k2) * normcoeff;
      pi = (ci / k2) * normcoeff;
    }
    hat[2 * q] = pr;
    hat[2 * q + 1] = pi;
    hat[2 * (N1 + q)] = er;
    hat[2 * (N1 + q) + 1] = ei;
    hat[2 * (2 * N1 + q)] = fr;
    hat[2 * (2 * N1 + q) + 1] = fi;
  }
  __syncthreads();
  for (int iy = threadIdx.x; iy < N1; iy += blockDim.x) {
    for (int fld = 0; fld < 3; fld++) {
      const double *h = hat + 2 * N1 * fld;
      double sr = 0., si = 0.;
      int m = 0;
      for (int q = 0; q < N1; q++) {
        const double wr = b.twy[2 * m], wi = -b.twy[2 * m + 1];  // exp(+2 pi i m/N)
        sr += h[2 * q] * wr - h[2 * q + 1] * wi;
        si += h[2 * q] * wi + h[2 * q + 1] * wr;
        m += iy;
        if (m >= N1) m -= N1;
      }
      double *dst = b.spec2 + 2 * ((long)nxh * N1 * fld + k + (long)nxh * iy);
      dst[0] = sr;
      dst[1] = si;
    }
  }
}

__global__ __launch_bounds__(256) void k_c2rx_diag(Geom g, FieldBufs b) {
  extern __shared__ double sh[];
  const int N0 = g.n[0], N1 = g.n[1];
  const int nxh = N0 / 2 + 1;
  const int iy = blockIdx.x;
  double *rows = sh;             // 3 * nxh complex
  double *red = sh + 6 * nxh;    // 8 doubles for reductions
  for (int i = threadIdx.x; i < 3 * nxh; i += blockDim.x) {
    const int fld = i / nxh, k = i - fld * nxh;
    const double *src = b.spec2 + 2 * ((long)nxh * N1 * fld + k + (long)nxh * iy);
    rows[2 * i] = src[0];
    rows[2 * i + 1] = src[1];
  }
  __syncthreads();
  double pm = 0., pe = 0., pex = 0., pey = 0.;
  for (int ix = threadIdx.x; ix < N0; ix += blockDim.x) {
    double out[3];
    for (int fld = 0; fld < 3; fld++) {
      const double *X = rows + 2 * nxh * fld;
      double s = X[0];
      int m = ix;
      for (int k = 1; k < (N0 + 1) / 2; k++) {
        const double wr = b.twx[2 * m], wi = -b.twx[2 * m + 1];
        s += 2.0 * (X[2 * k] * wr - X[2 * k + 1] * wi);
        m += ix;
        if (m >= N0) m -= N0;
      }
      if ((N0 & 1) == 0) s += X[2 * (N0 / 2)] * ((ix & 1) ? -1.0 : 1.0);
      out[fld] = s;
    }
    const long id = ix + (long)N0 * iy;
    b.rho[id] = out[0];
    b.ex[id] = out[1];
    b.ey[id] = out[2];
    pm += out[0];
    pe += out[1] * out[1] + out[2] * out[2];
    pex += out[1] * out[1];
    pey += out[2] * out[2];
  }
  const double s0 = block_sum256(pm, red);
  const double s1 = block_sum256(pe, red);
  const double s2 = block_sum256(pex, red);
  const double s3 = block_sum256(pey, red);
  if (threadIdx.x == 0) {
    b.diag_part[4 * iy + 0] = s0;
    b.diag_part[4 * iy + 1] = s1;
    b.diag_part[4 * iy + 2] = s2;
    b.diag_part[4 * iy + 3] = s3;
  }
}

// Analytic fields of idcase 2 and 6 (VMO/Field.cpp:81-108); rho is kept.
__global__ __launch_bounds__(256) void k_analytic_field(const double *__restrict__ rho_part, Geom g, FieldBufs b) {
  __shared__ double red[8];
  const int N0 = g.n[0], N1 = g.n[1], N3 = g.n[3];
  const int iy = blockIdx.x;
  double pm = 0., pe = 0., pex = 0., pey = 0.;
  for (int ix = threadIdx.x; ix < N0; ix += blockDim.x) {
    double s = 0.;
    for (int ivy = 0; ivy < N3; ivy++) s += rho_part[ix + (long)N0 * (iy + (long)N1 * ivy)];
    const double r = s * g.dx[2] * g.dx[3];
    const double e = (g.idcase == 2) ? -(g.minPhy[0] + ix * g.dx[0]) : -(g.minPhy[1] + iy * g.dx[1]);
    const long id = ix + (long)N0 * iy;
    b.rho[id] = r;
    b.ex[id] = e;
    b.ey[id] = 0.;
    pm += r;
    pe += e * e;
    pex += e * e;
  }
  const double s0 = block_sum256(pm, red);
  const double s1 = block_sum256(pe, red);
  const double s2 = block_sum256(pex, red);
  const double s3 = block_sum256(pey, red);
  if (threadIdx.x == 0) {
    b.diag_part[4 * iy + 0] = s0;
    b.diag_part[4 * iy + 1] = s1;
    b.diag_part[4 * iy + 2] = s2;
    b.diag_part[4 * iy + 3] = s3;
  }
}

// Diags::compute and computeL2norm, VMO/Diags.cpp:16-76.
__global__ __launch_bounds__(256) void k_diag_final(Geom g, FieldBufs b, const double *__restrict__ l2_part, long n_l2,
                                                    int iter, int with_l2) {
  __shared__ double red[8];
  const int N1 = g.n[1];
  double s[4];
  for (int c = 0; c < 4; c++) {
    double v = 0.;
    for (int iy = threadIdx.x; iy < N1; iy += blockDim.x) v += b.diag_part[4 * iy + c];
    s[c] = block_sum256(v, red);
  }
  double l2 = 0.;
  if (with_l2) {
    double v = 0.;
    for (long i = threadIdx.x; i < n_l2; i += blockDim.x) v += l2_part[i];
    l2 = block_sum256(v, red);
  }
  if (threadIdx.x == 0) {
    const int nb = b.nbiter + 1;
    double it_nrj = sqrt(s[1] * g.dx[0] * g.dx[1]);
    it_nrj = it_nrj > 1.e-30 ? log(it_nrj) : -1.e9;
    b.diag[0 * nb + iter] = it_nrj;
    b.diag[1 * nb + iter] = sqrt(0.5 * s[2] * g.dx[0] * g.dx[1]);
    b.diag[2 * nb + iter] = sqrt(0.5 * s[3] * g.dx[0] * g.dx[1]);
    b.diag[3 * nb + iter] = s[0] * g.dx[0] * g.dx[1];
    if (with_l2) b.diag[4 * nb + iter] = sqrt(l2 * g.dx[0] * g.dx[1] * g.dx[2] * g.dx[3]);
  }
}

void launch_field(const Geom &g, const FieldBufs &b, const double *rho_part, const double *l2_part, long n_l2, int iter,
                  bool record_diag, hipStream_t s) {
  const int N0 = g.n[0], N1 = g.n[1];
  const int nxh = N0 / 2 + 1;
  if (g.idcase == 2 || g.idcase == 6) {
    hipLaunchKernelGGL(k_analytic_field, dim3(N1), dim3(256), 0, s, rho_part, g, b);
  } else {
    const int sub1 = (g.idcase == 10 || g.idcase == 20);
    hipLaunchKernelGGL(k_rho_rfftx, dim3(N1), dim3(256), sizeof(double) * N0, s, rho_part, g, b, sub1);
    hipLaunchKernelGGL(k_poisson_y, dim3(nxh), dim3(256), sizeof(double) * 8 * N1, s, g, b);
    hipLaunchKernelGGL(k_c2rx_diag, dim3(N1), dim3(256), sizeof(double) * (6 * nxh + 8), s, g, b);
  }
  if (record_diag)
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, s, g, b, l2_part, n_l2, iter, l2_part != nullptr ? 1 : 0);
}

// explicit instantiations
#define VLP_INST(T)                                                                                             \
  template void launch_init_case<T>(const Geom &, T *, hipStream_t);                                            \
  template void launch_spline_x<T>(const Geom &, const T *, T *, hipStream_t);                                  \
  template void launch_spline_line<T>(const T *, T *, int, long, long, long, long, long, long, hipStream_t);    \
  template void launch_advect2d_rho<T>(const Geom &, const T *, double, double *, int *, hipStream_t);           \
  template void launch_advect4d<T>(const Geom &, const T *, const double *, const double *, double, T *,       \
                                   double *, double *, int *, hipStream_t);                                      \
  template void launch_moments<T>(const Geom &, const T *, double *, double *, hipStream_t);
VLP_INST(double)
VLP_INST(float)

}  // namespace vlp
