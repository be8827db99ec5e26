// adjoint.hip -- hand-written adjoint of one time step (FP64, gfx950).
//
// The step maps f^n to f^{n+1} = I4(C4(f^n); feet(E(f^n))) with
//   C4 = (M_vy x M_vx x M_y x M_x) f^n          (computeCoeff_xy + _vxvy,
//                                                 VMO/Spline.hpp:384-469)
//   I4 = the 4^4 B-spline interpolation at the feet (interp_4D,
//        VMO/Advection.hpp:221-296), feet from computeFeet (:166-218),
//   E  = K rho_half, rho_half = dvx dvy sum_v I2(C_xy(f^n)) (advect_2D_xy
//        :299-406, field_rho VMO/Field.cpp:37-53, Poisson :71-137).
// For a cotangent g of f^{n+1} the adjoint is
//   grad = C4^T I4^T g                                   (direct term)
//        + C_xy^T I2^T (dvx dvy (K^T lambda_E))          (through the field)
// with lambda_E the foot term: g times the derivative of I4 along the foot,
// times d foot / d E-bar (dt^2/2 for x, y; -dt for vx, vy unless clamped),
// scattered back through computeFeet's 3x3 Lagrange interpolation of E.
// M_d is the (N_d+4) x N_d spline operator of one line (the same for every
// direction of a given length), K the periodic E-from-rho operator.
#include "advect_common.hpp"

#include <cstdlib>

namespace vlp {

namespace {

// Derivative of the cubic B-spline weights bspline_eta with respect to w.
__device__ __forceinline__ void bspline_deta(double w, double d[4]) {
  const double d3 = 0.5 * w * w;
  const double d0 = 0.5 * (2. * w - 1.) - d3;
  const double d2 = 1. + d0 - 2. * d3;
  d[0] = d0;
  d[1] = -d0 - d2 - d3;
  d[2] = d2;
  d[3] = d3;
}

}  // namespace

// Per cell: the feet of the forward step, then
//  * g * w0 w1 w2 w3 added to lc4 at the 4^4 stencil (lc4 has the C4
//    storage layout, pitches c4_pitch);
//  * the foot term g * dI4/dfoot * dfoot/dEbar, scattered to lE (Ex, Ey
//    interleaved, N1 x N0) with computeFeet's Lagrange weights.
__global__ __launch_bounds__(256) void k_adj_advect4d(const double *__restrict__ c4, const double *__restrict__ ex,
                                                      const double *__restrict__ ey, Geom g, FeetConsts fc, double dt,
                                                      const double *__restrict__ gin, double *__restrict__ lc4,
                                                      double *__restrict__ lE, int field_term) {
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const long ncell = (long)N0 * N1 * N2 * N3;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell) return;
  const int ix = (int)(t % N0), iy = (int)((t / N0) % N1), ivx = (int)((t / ((long)N0 * N1)) % N2),
            ivy = (int)(t / ((long)N0 * N1 * N2));
  const double gv = gin[t];
  if (gv == 0.) return;
  const double x = g.minPhy[0] + ix * g.dx[0], y = g.minPhy[1] + iy * g.dx[1];
  const double vx = g.minPhy[2] + ivx * g.dx[2], vy = g.minPhy[3] + ivy * g.dx[3];
  const double xt1 = half_y(g, y, vy, dt);
  const int ip1 = (int)(round(xt1) - 1);
  EGlobal E{ex, ey, N0, N1, ip1};
  double xs[4];
  if (!compute_foot(g, fc, E, xt1, ip1, x, y, vx, vy, dt, xs)) return;  // the forward step reports the error
  int ip[4];
  double w[4][4], dw[4][4];
  for (int j = 0; j < 4; j++) {
    const double fj = g.inv_dx[j] * (xs[j] - g.minPhy[j]);
    ip[j] = (int)floor(fj);
    bspline_eta<double>(fj - ip[j], w[j]);
    bspline_deta(fj - ip[j], dw[j]);
  }
  const C4Pitch cp = c4_pitch(N0, N1, N2);
  const long S1 = cp.s1, S2 = cp.s2, S3 = cp.s3;
  const long base = (ip[0] + 2) + S1 * (ip[1] + 2) + S2 * (ip[2] + 2) + S3 * (long)(ip[3] + 2);
  double dI[4] = {0., 0., 0., 0.};
  for (int k3 = 0; k3 < 4; k3++)
    for (int k2 = 0; k2 < 4; k2++)
      for (int k1 = 0; k1 < 4; k1++)
        for (int k0 = 0; k0 < 4; k0++) {
          const long o = base + k0 + S1 * k1 + S2 * k2 + S3 * k3;
          atomicAdd(lc4 + o, gv * w[0][k0] * w[1][k1] * w[2][k2] * w[3][k3]);
          if (field_term) {
            const double c = c4[o];
            dI[0] += c * dw[0][k0] * w[1][k1] * w[2][k2] * w[3][k3];
            dI[1] += c * w[0][k0] * dw[1][k1] * w[2][k2] * w[3][k3];
            dI[2] += c * w[0][k0] * w[1][k1] * dw[2][k2] * w[3][k3];
            dI[3] += c * w[0][k0] * w[1][k1] * w[2][k2] * dw[3][k3];
          }
        }
  if (!field_term) return;
  // d f^{n+1} / d xs_j = dI_j / dx_j; dxs/dEbar: x, y: dt^2/2; vx, vy: -dt
  // inside the velocity box (the clamp has zero derivative outside)
  const double vhi2 = g.minPhy[2] + fc.rxwidth[2], vhi3 = g.minPhy[3] + fc.rxwidth[3];
  const double ax = gv * (dI[0] * g.inv_dx[0] * 0.5 * dt * dt +
                          ((xs[2] > g.minPhy[2] && xs[2] < vhi2) ? -dt * dI[2] * g.inv_dx[2] : 0.));
  const double ay = gv * (dI[1] * g.inv_dx[1] * 0.5 * dt * dt +
                          ((xs[3] > g.minPhy[3] && xs[3] < vhi3) ? -dt * dI[3] * g.inv_dx[3] : 0.));
  // Ebar = sum_{k0,k1} cx[k0] cy[k1] E(ip0+k0, ip1+k1), VMO/Advection.hpp:174-196
  double xt0 = x - 0.5 * dt * vx;
  xt0 = g.inv_dx[0] * (xt0 - g.minPhy[0]);
  const int ip0 = (int)(round(xt0) - 1);
  const double px = xt0 - ip0, py = xt1 - ip1;
  const double cx[3] = {.5 * (px - 1.) * (px - 2.), -1. * (px) * (px - 2.), .5 * (px) * (px - 1.)};
  const double cy[3] = {.5 * (py - 1.) * (py - 2.), -1. * (py) * (py - 2.), .5 * (py) * (py - 1.)};
  for (int k1 = 0; k1 < 3; k1++)
    for (int k0 = 0; k0 < 3; k0++) {
      const int jx = ((ip0 + k0) % N0 + N0) % N0, jy = ((ip1 + k1) % N1 + N1) % N1;
      const double cc = cx[k0] * cy[k1];
      const long e = 2 * ((long)jy * N0 + jx);
      atomicAdd(lE + e, cc * ax);
      atomicAdd(lE + e + 1, cc * ay);
    }
}

// Fixed-point form of a coefficient-adjoint contribution: x * S rounded to
// the nearest integer (|x S| < 2^51: adding 1.5 * 2^52 leaves the integer in
// the low mantissa bits).  Integer sums do not depend on the order of the
// adds, so the result is the same on every run.
__device__ __forceinline__ unsigned long long to_fixed(double x) {
  const double M = 6755399441055744.0;  // 1.5 * 2^52
  return (unsigned long long)(__double_as_longlong(x + M) - __double_as_longlong(M));
}

// The same scatter by blocks of cells, with reproducible sums.  A workgroup
// takes BX x 2 x 2 x 2 cells (x fastest, one cell per lane) and adds their
// coefficient-adjoint contributions, as 64-bit fixed-point integers (scale
// S = *scl, a power of two with |g| S < 2^50), into an LDS copy of the
// coefficient region they reach when every foot lies within one cell below
// its own (ip_d - c_d in {-1, 0}: positions c - 2 .. c + 2 per dimension);
// a contribution outside it goes straight to lc4 (read as 64-bit integers).
// The region is then added to lc4, one atomic per nonzero entry instead of
// one per cell and stencil point: (BX + 4) 6^3 / (8 BX) ~ 29 global atomics
// per cell instead of 256.  With the field term, each cell's pair
// (ax, ay) -- g times the derivative of the interpolation along the foot,
// times d foot / d E-bar -- is written to axy, for the E gathers below.
template <int BX>
__global__ __launch_bounds__(8 * BX) void k_adj_advect4d_blk(const double *__restrict__ c4,
                                                           const double *__restrict__ ex,
                                                           const double *__restrict__ ey, Geom g, FeetConsts fc,
                                                           double dt, const double *__restrict__ gin,
                                                           unsigned long long *__restrict__ lc4,
                                                           double *__restrict__ axy, const double *__restrict__ scl,
                                                           int field_term) {
  constexpr int RX = BX + 4, RY = 6, RV = 6;  // coefficient region (x, y, vx, vy)
  constexpr int NREG = RX * RY * RV * RV;
  constexpr int NT = 8 * BX;
  extern __shared__ __align__(16) unsigned char adj_smem[];
  unsigned long long *reg = (unsigned long long *)adj_smem;
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const int nbx = (N0 + BX - 1) / BX, nby = (N1 + 1) / 2, nbv = (N2 + 1) / 2;
  long b = blockIdx.x;
  const int bx = (int)(b % nbx);
  b /= nbx;
  const int by = (int)(b % nby);
  b /= nby;
  const int bvx = (int)(b % nbv), bvy = (int)(b / nbv);
  const int x0 = bx * BX, y0 = 2 * by, vx0 = 2 * bvx, vy0 = 2 * bvy;
  const int tid = threadIdx.x;
  const double S = scl[0];
  for (int i = tid; i < NREG; i += NT) reg[i] = 0ull;
  __syncthreads();
  const int lx = tid % BX, r = tid / BX;
  const int ix = x0 + lx, iy = y0 + (r & 1), ivx = vx0 + ((r >> 1) & 1), ivy = vy0 + (r >> 2);
  const C4Pitch cp = c4_pitch(N0, N1, N2);
  const long S1 = cp.s1, S2 = cp.s2, S3 = cp.s3;
  if (ix < N0 && iy < N1 && ivx < N2 && ivy < N3) {
    const long t = ix + (long)N0 * (iy + (long)N1 * (ivx + (long)N2 * ivy));
    const double gv = gin[t];
    const double x = g.minPhy[0] + ix * g.dx[0], y = g.minPhy[1] + iy * g.dx[1];
    const double vx = g.minPhy[2] + ivx * g.dx[2], vy = g.minPhy[3] + ivy * g.dx[3];
    const double xt1 = half_y(g, y, vy, dt);
    const int ip1 = (int)(round(xt1) - 1);
    EGlobal E{ex, ey, N0, N1, ip1};
    double xs[4];
    double ax = 0., ay = 0.;
    // (a zero cotangent or a foot out of the box adds nothing; the forward
    // step reports the latter)
    if (gv != 0. && compute_foot(g, fc, E, xt1, ip1, x, y, vx, vy, dt, xs)) {
      int ip[4];
      double w[4][4], dw[4][4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const double fj = g.inv_dx[j] * (xs[j] - g.minPhy[j]);
        ip[j] = (int)floor(fj);
        bspline_eta<double>(fj - ip[j], w[j]);
        bspline_deta(fj - ip[j], dw[j]);
      }
      // stencil origin in the region (positions x0 - 2 .., y0 - 2 .., ...)
      const int r0 = ip[0] - 1 - (x0 - 2), r1 = ip[1] - 1 - (y0 - 2), r2 = ip[2] - 1 - (vx0 - 2),
                r3 = ip[3] - 1 - (vy0 - 2);
      const bool inreg = r0 >= 0 && r0 + 3 < RX && r1 >= 0 && r1 + 3 < RY && r2 >= 0 && r2 + 3 < RV && r3 >= 0 &&
                         r3 + 3 < RV;
      const long base = (ip[0] + 2) + S1 * (ip[1] + 2) + S2 * (ip[2] + 2) + S3 * (long)(ip[3] + 2);
      // dI_j = sum over the stencil of c * (the weight product with dimension
      // j differentiated), contracted one dimension at a time: x (s, sd),
      // then y (u*), vx (v*), vy (dI)
      double dI[4] = {0., 0., 0., 0.};
      double gw0[4];
#pragma unroll
      for (int k0 = 0; k0 < 4; k0++) gw0[k0] = gv * S * w[0][k0];
      for (int k3 = 0; k3 < 4; k3++) {
        double v = 0., vd0 = 0., vd1 = 0., vd2 = 0.;
        for (int k2 = 0; k2 < 4; k2++) {
          double u = 0., ud0 = 0., ud1 = 0.;
          const double w23 = w[2][k2] * w[3][k3];
#pragma unroll
          for (int k1 = 0; k1 < 4; k1++) {
            const double w123 = w[1][k1] * w23;
            const int lrow = ((r3 + k3) * RV + (r2 + k2)) * RY + (r1 + k1);
            const long orow = base + S1 * k1 + S2 * k2 + S3 * k3;
#pragma unroll
            for (int k0 = 0; k0 < 4; k0++) {
              const unsigned long long q = to_fixed(gw0[k0] * w123);
              if (inreg) atomicAdd(reg + lrow * RX + r0 + k0, q);
              else atomicAdd(lc4 + orow + k0, q);
            }
            if (field_term) {
              double s0 = 0., sd = 0.;
#pragma unroll
              for (int k0 = 0; k0 < 4; k0++) {
                const double cv = c4[orow + k0];
                s0 += w[0][k0] * cv;
                sd += dw[0][k0] * cv;
              }
              u += w[1][k1] * s0;
              ud0 += w[1][k1] * sd;
              ud1 += dw[1][k1] * s0;
            }
          }
          v += w[2][k2] * u;
          vd0 += w[2][k2] * ud0;
          vd1 += w[2][k2] * ud1;
          vd2 += dw[2][k2] * u;
        }
        dI[0] += w[3][k3] * vd0;
        dI[1] += w[3][k3] * vd1;
        dI[2] += w[3][k3] * vd2;
        dI[3] += dw[3][k3] * v;
      }
      if (field_term) {
        const double vhi2 = g.minPhy[2] + fc.rxwidth[2], vhi3 = g.minPhy[3] + fc.rxwidth[3];
        ax = gv * (dI[0] * g.inv_dx[0] * 0.5 * dt * dt +
                   ((xs[2] > g.minPhy[2] && xs[2] < vhi2) ? -dt * dI[2] * g.inv_dx[2] : 0.));
        ay = gv * (dI[1] * g.inv_dx[1] * 0.5 * dt * dt +
                   ((xs[3] > g.minPhy[3] && xs[3] < vhi3) ? -dt * dI[3] * g.inv_dx[3] : 0.));
      }
    }
    if (field_term) {
      axy[2 * t] = ax;
      axy[2 * t + 1] = ay;
    }
  }
  __syncthreads();
  // the region to global memory: region position (r0, r1, r2, r3) is the
  // coefficient at position (x0 - 2 + r0, ...), storage index + 3 per
  // dimension; only entries some cell reached are nonzero
  for (int i = tid; i < NREG; i += NT) {
    const unsigned long long v = reg[i];
    if (v != 0ull) {
      const int q0 = i % RX, q1 = (i / RX) % RY, q2 = (i / (RX * RY)) % RV, q3 = i / (RX * RY * RV);
      const long o = (x0 + 1 + q0) + S1 * (y0 + 1 + q1) + S2 * (vx0 + 1 + q2) + S3 * (long)(vy0 + 1 + q3);
      atomicAdd(lc4 + o, v);
    }
  }
}

// Scale of the fixed-point coefficient adjoint: the largest |g| (as the bit
// pattern of a non-negative double, whose order is the numeric order), then
// S = 2^(50 - e) with max|g| < 2^e, and 1/S.
__global__ __launch_bounds__(256) void k_absmax_bits(const double *__restrict__ gin, long n,
                                                     unsigned long long *__restrict__ mx) {
  unsigned long long m = 0ull;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(gin[t]));
    m = b > m ? b : m;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_down(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(mx, m);
}
__global__ void k_fixed_zero(unsigned long long *__restrict__ mx) {
  if (threadIdx.x == 0) *mx = 0ull;
}
__global__ void k_fixed_scale(const unsigned long long *__restrict__ mx, double *__restrict__ scl) {
  if (threadIdx.x != 0) return;
  const double m = __longlong_as_double((long long)*mx);
  double S = 1.;
  if (m > 0. && m <= 1.7976931348623157e308) {
    int e = 0;
    (void)frexp(m, &e);  // m < 2^e
    S = ldexp(1., 50 - e);
  }
  scl[0] = S;
  scl[1] = 1. / S;
}

// E adjoint by gathers, in a fixed order.  computeFeet takes E-bar from the
// 3 x 3 nodes (ip0 + k0, ip1 + k1) (mod N0, N1) with weights cx[k0] cy[k1],
// ip0 and cx depending on (x, vx) and ip1 and cy on (y, vy).  First, for
// every (jx, y, vy): the sum over vx and over the cells x whose node
// ip0 + k0 is jx of cx[k0] (ax, ay)(x, y, vx, vy) -- ip0 = x + round(d) - 1
// with d = -dt vx / (2 dx) up to rounding, so x is one of five candidates,
// each checked exactly; then for every node (jx, jy) the same over (y, vy)
// with cy.
__global__ __launch_bounds__(256) void k_adjE_x(const double *__restrict__ axy, Geom g, double dt,
                                                double *__restrict__ part) {
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const long n = (long)N0 * N1 * N3;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int jx = (int)(t % N0), iy = (int)((t / N0) % N1), ivy = (int)(t / ((long)N0 * N1));
  double sx = 0., sy = 0.;
  for (int ivx = 0; ivx < N2; ivx++) {
    const double vx = g.minPhy[2] + ivx * g.dx[2];
    const int sh = (int)round(-0.5 * dt * vx * g.inv_dx[0]) - 1;
    for (int q = -3; q <= 1; q++) {
      const int ix = ((jx - sh + q) % N0 + N0) % N0;
      double xt0 = (g.minPhy[0] + ix * g.dx[0]) - 0.5 * dt * vx;
      xt0 = g.inv_dx[0] * (xt0 - g.minPhy[0]);
      const int ip0 = (int)(round(xt0) - 1);
      const int k0 = ((jx - ip0) % N0 + N0) % N0;
      if (k0 > 2) continue;
      const double px = xt0 - ip0;
      const double cx = k0 == 0 ? .5 * (px - 1.) * (px - 2.) : k0 == 1 ? -1. * (px) * (px - 2.) : .5 * (px) * (px - 1.);
      const long c = ix + (long)N0 * (iy + (long)N1 * (ivx + (long)N2 * ivy));
      sx += cx * axy[2 * c];
      sy += cx * axy[2 * c + 1];
    }
  }
  part[2 * t] = sx;
  part[2 * t + 1] = sy;
}
__global__ __launch_bounds__(256) void k_adjE_y(const double *__restrict__ part, Geom g, double dt,
                                                double *__restrict__ lE) {
  const int N0 = g.n[0], N1 = g.n[1], N3 = g.n[3];
  const long n = (long)N0 * N1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int jx = (int)(t % N0), jy = (int)(t / N0);
  double sx = 0., sy = 0.;
  for (int ivy = 0; ivy < N3; ivy++) {
    const double vy = g.minPhy[3] + ivy * g.dx[3];
    const int sh = (int)round(-0.5 * dt * vy * g.inv_dx[1]) - 1;
    for (int q = -3; q <= 1; q++) {
      const int iy = ((jy - sh + q) % N1 + N1) % N1;
      const double xt1 = half_y(g, g.minPhy[1] + iy * g.dx[1], vy, dt);
      const int ip1 = (int)(round(xt1) - 1);
      const int k1 = ((jy - ip1) % N1 + N1) % N1;
      if (k1 > 2) continue;
      const double py = xt1 - ip1;
      const double cy = k1 == 0 ? .5 * (py - 1.) * (py - 2.) : k1 == 1 ? -1. * (py) * (py - 2.) : .5 * (py) * (py - 1.);
      const long pidx = jx + (long)N0 * (iy + (long)N1 * ivy);
      sx += cy * part[2 * pidx];
      sy += cy * part[2 * pidx + 1];
    }
  }
  lE[2 * t] = sx;
  lE[2 * t + 1] = sy;
}

// out[o][n][i] (+)= sum_m Mt[n][m] in[o][m][i]: a line operator's transpose
// applied along one axis (Mt is M^T, Nout x Min, row-major).
__global__ __launch_bounds__(256) void k_axis_contract(const double *__restrict__ Mt, int Min, int Nout,
                                                       const double *__restrict__ in, double *__restrict__ out, long O,
                                                       long I, int accumulate) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= O * Nout * I) return;
  const long i = t % I, n = (t / I) % Nout, o = t / (I * Nout);
  const double *src = in + o * Min * I + i;
  const double *a = Mt + (long)n * Min;
  double s = 0.;
  for (int m = 0; m < Min; m++) s += a[m] * src[(long)m * I];
  if (accumulate) out[t] += s;
  else out[t] = s;
}

// The same contraction by tiles: a workgroup stages in[o][0..Min-1][64
// columns] in LDS once (columns: 64 consecutive i of one o, or, when I = 1,
// 64 consecutive o) and each wave forms 16 outputs n per pass from it, the
// operator entries M[m][n] (M = Mt^T, Min x Nout row-major) arriving as
// wave-uniform scalar loads.  Same sums in the same order as k_axis_contract.
template <bool COLO>
__global__ __launch_bounds__(256) void k_axis_contract_t(const double *__restrict__ M, int Min, int Nout,
                                                         const double *__restrict__ in, double *__restrict__ out,
                                                         long O, long I, int accumulate) {
  extern __shared__ __align__(16) double ctile[];  // [Min][64]
  const int tid = threadIdx.x, c = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  long o0, i0;
  int ncol;
  if (COLO) {
    o0 = (long)blockIdx.x * 64;
    i0 = 0;
    ncol = (int)(O - o0 < 64 ? O - o0 : 64);
  } else {
    const long nt = (I + 63) / 64;
    o0 = blockIdx.x / nt;
    i0 = (blockIdx.x % nt) * 64;
    ncol = (int)(I - i0 < 64 ? I - i0 : 64);
  }
  if (COLO) {
    for (int e = tid; e < 64 * Min; e += 256) {
      const int cc = e / Min, m = e - cc * Min;
      ctile[m * 64 + cc] = cc < ncol ? in[(o0 + cc) * Min + m] : 0.;
    }
  } else {
    const double *src = in + o0 * Min * I + i0;
    for (int e = tid; e < 64 * Min; e += 256) {
      const int m = e >> 6, cc = e & 63;
      ctile[e] = cc < ncol ? src[(long)m * I + cc] : 0.;
    }
  }
  __syncthreads();
  for (int n0 = 16 * wv; n0 < Nout; n0 += 64) {
    const int jmax = Nout - n0 < 16 ? Nout - n0 : 16;
    double acc[16];
#pragma unroll
    for (int j = 0; j < 16; j++) acc[j] = 0.;
    if (jmax == 16) {  // whole chunk: 16 consecutive entries per m, loaded together
      for (int m = 0; m < Min; m++) {
        const double x = ctile[m * 64 + c];
        const double *mr = M + (long)m * Nout + n0;
#pragma unroll
        for (int j = 0; j < 16; j++) acc[j] += mr[j] * x;
      }
    } else {
      for (int m = 0; m < Min; m++) {
        const double x = ctile[m * 64 + c];
        const double *mr = M + (long)m * Nout + n0;
#pragma unroll
        for (int j = 0; j < 16; j++)
          if (j < jmax) acc[j] += mr[j] * x;
      }
    }
    if (c < ncol) {
#pragma unroll
      for (int j = 0; j < 16; j++)
        if (j < jmax) {
          const long idx = COLO ? (o0 + c) * Nout + n0 + j : (o0 * Nout + n0 + j) * I + i0 + c;
          out[idx] = accumulate ? out[idx] + acc[j] : acc[j];
        }
    }
  }
}

// Contiguous copy of the coefficient block (positions -2..N+1 per dimension)
// out of the C4 storage layout (positions -3..N+2); with scl, the storage
// holds 64-bit fixed-point sums, converted with the factor scl[1] = 1/S.
__global__ __launch_bounds__(256) void k_c4_block(const double *__restrict__ c4s, double *__restrict__ blk, Geom g,
                                                  const double *__restrict__ scl) {
  const long E0 = g.n[0] + 4, E1 = g.n[1] + 4, E2 = g.n[2] + 4, E3 = g.n[3] + 4;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E0 * E1 * E2 * E3) return;
  const long a = t % E0, b = (t / E0) % E1, c = (t / (E0 * E1)) % E2, d = t / (E0 * E1 * E2);
  const C4Pitch cp = c4_pitch(g.n[0], g.n[1], g.n[2]);
  const long o = (a + 1) + cp.s1 * (b + 1) + cp.s2 * (c + 1) + cp.s3 * (d + 1);
  if (scl) blk[t] = (double)((const long long *)c4s)[o] * scl[1];
  else blk[t] = c4s[o];
}

// lambda_rho(j) = sum_i lE_x(i) kx(i - j) + lE_y(i) ky(i - j): the transpose
// of the periodic E-from-rho convolution, kx/ky its response to a unit
// density at the origin (indices mod N0, N1).
__global__ __launch_bounds__(256) void k_poisson_T(const double *__restrict__ lE, const double *__restrict__ kx,
                                                   const double *__restrict__ ky, double *__restrict__ lrho, int N0,
                                                   int N1) {
  // 16 outputs per workgroup, each summed over 16 slices of the source rows
  // by 16 lanes, then the slices added in a fixed order
  __shared__ double part[16][16];
  const int o = threadIdx.x & 15, p = threadIdx.x >> 4;
  const long t = (long)blockIdx.x * 16 + o;
  double s = 0.;
  if (t < (long)N0 * N1) {
    const int jx = (int)(t % N0), jy = (int)(t / N0);
    const int y0 = (int)((long)p * N1 / 16), y1 = (int)((long)(p + 1) * N1 / 16);
    for (int iy = y0; iy < y1; iy++) {
      const int dy = iy >= jy ? iy - jy : iy - jy + N1;
      const double *e = lE + 2 * (long)iy * N0;
      const double *kxr = kx + (long)dy * N0, *kyr = ky + (long)dy * N0;
      // ix = 0 .. N0-1 with dx = (ix - jx) mod N0, without a division per term
      for (int ix = 0; ix < jx; ix++) s += e[2 * ix] * kxr[ix - jx + N0] + e[2 * ix + 1] * kyr[ix - jx + N0];
      for (int ix = jx; ix < N0; ix++) s += e[2 * ix] * kxr[ix - jx] + e[2 * ix + 1] * kyr[ix - jx];
    }
  }
  part[p][o] = s;
  __syncthreads();
  if (p == 0 && t < (long)N0 * N1) {
    double r = 0.;
    for (int q = 0; q < 16; q++) r += part[q][o];
    lrho[t] = r;
  }
}

// Through the half step: f_half(x, y, v) = I2(C_xy) at (x - dt/2 vx,
// y - dt/2 vy) and rho_half = dvx dvy sum_v f_half, so lambda_{f_half} =
// dvx dvy lambda_rho(x, y), scattered by the 4x4 B-spline weights of the
// half-step foot into lcxy (the C_xy layout: positions -2..N+1 in x and y,
// then vx, vy).
__global__ __launch_bounds__(256) void k_adj_advect2d(const double *__restrict__ lrho, Geom g, double dth,
                                                      double *__restrict__ lcxy) {
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const long ncell = (long)N0 * N1 * N2 * N3;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell) return;
  const int ix = (int)(t % N0), iy = (int)((t / N0) % N1), ivx = (int)((t / ((long)N0 * N1)) % N2),
            ivy = (int)(t / ((long)N0 * N1 * N2));
  const double lv = lrho[(long)iy * N0 + ix] * g.dx[2] * g.dx[3];
  if (lv == 0.) return;
  const double x = g.minPhy[0] + ix * g.dx[0], y = g.minPhy[1] + iy * g.dx[1];
  const double vx = g.minPhy[2] + ivx * g.dx[2], vy = g.minPhy[3] + ivy * g.dx[3];
  const double fx = ((x - dth * vx) - g.minPhy[0]) * g.inv_dx[0];
  const double fy = ((y - dth * vy) - g.minPhy[1]) * g.inv_dx[1];
  const int jx0 = (int)floor(fx), jy0 = (int)floor(fy);
  if (jx0 < -1 || jx0 > N0 - 1 || jy0 < -1 || jy0 > N1 - 1) return;  // interp_2D leaves 0 there
  double ex[4], ey[4];
  bspline_eta<double>(fx - jx0, ex);
  bspline_eta<double>(fy - jy0, ey);
  const long E0 = N0 + 4, E1 = N1 + 4;
  double *pl = lcxy + ((long)ivx + (long)N2 * ivy) * E0 * E1;
  // stencil position jx0-1+k at column (position + 2)
  for (int ky = 0; ky < 4; ky++)
    for (int kx = 0; kx < 4; kx++) atomicAdd(pl + (long)(jy0 + 1 + ky) * E0 + (jx0 + 1 + kx), lv * ex[kx] * ey[ky]);
}

// The same adjoint as a gather, one lane per C_xy entry (X, Y) of a plane
// (positions -2..N+1 as columns 0..N+3): the cells whose half-step stencil
// covers X have jx0 = X - 1 - kx, kx in 0..3, and jx0(ix) = floor(ix + d)
// up to rounding, d = -dth vx / dx, so ix lies in
// [X - 5 - floor(d), X - floor(d)]; each candidate's jx0 and weights are
// recomputed exactly as the scatter (and the forward) computes them, and
// likewise in y.  No atomics: the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_adj_advect2d_g(const double *__restrict__ lrho, Geom g, double dth,
                                                        double *__restrict__ lcxy) {
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const long E0 = N0 + 4, E1 = N1 + 4;
  const long n = E0 * E1 * N2 * N3;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int X = (int)(t % E0), Y = (int)((t / E0) % E1);
  const long plane = t / (E0 * E1);
  const int ivx = (int)(plane % N2), ivy = (int)(plane / N2);
  const double vx = g.minPhy[2] + ivx * g.dx[2], vy = g.minPhy[3] + ivy * g.dx[3];
  const int fdx = (int)floor(-dth * vx * g.inv_dx[0]), fdy = (int)floor(-dth * vy * g.inv_dx[1]);
  // y candidates: weight of row Y for each, 0 where it does not reach Y
  double wy[6];
  int iyc[6];
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const int iy = Y - 5 - fdy + q;
    iyc[q] = iy;
    wy[q] = 0.;
    if (iy >= 0 && iy < N1) {
      const double y = g.minPhy[1] + iy * g.dx[1];
      const double fy = ((y - dth * vy) - g.minPhy[1]) * g.inv_dx[1];
      const int jy0 = (int)floor(fy);
      const int ky = Y - 1 - jy0;
      if (jy0 >= -1 && jy0 <= N1 - 1 && ky >= 0 && ky < 4) {
        double e[4];
        bspline_eta<double>(fy - jy0, e);
        wy[q] = e[ky];
      }
    }
  }
  double acc = 0.;
#pragma unroll
  for (int p = 0; p < 6; p++) {
    const int ix = X - 5 - fdx + p;
    if (ix < 0 || ix >= N0) continue;
    const double x = g.minPhy[0] + ix * g.dx[0];
    const double fx = ((x - dth * vx) - g.minPhy[0]) * g.inv_dx[0];
    const int jx0 = (int)floor(fx);
    const int kx = X - 1 - jx0;
    if (jx0 < -1 || jx0 > N0 - 1 || kx < 0 || kx > 3) continue;
    double e[4];
    bspline_eta<double>(fx - jx0, e);
    double sy = 0.;
#pragma unroll
    for (int q = 0; q < 6; q++)
      if (wy[q] != 0.) sy += wy[q] * lrho[(long)iyc[q] * N0 + ix];
    acc += e[kx] * sy;
  }
  lcxy[t] = acc * g.dx[2] * g.dx[3];
}

// Field-energy source: W(f) = dx dy sum (Ex^2 + Ey^2) of the field solved
// from f, so dW/dE = 2 dx dy E.  lE = scale * (Ex, Ey) pairs.
__global__ __launch_bounds__(256) void k_scale_pairs(const double *__restrict__ exy, double scale, long n,
                                                     double *__restrict__ lE) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) lE[t] = scale * exy[t];
}

// g(x, y, v) += scale * lrho(x, y) for every (vx, vy): the density's adjoint
// (rho = dvx dvy sum_v f).
__global__ __launch_bounds__(256) void k_add_density_adj(const double *__restrict__ lrho, double scale, long nxy,
                                                         long ncell, double *__restrict__ gbuf) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ncell) gbuf[t] += scale * lrho[t % nxy];
}

// out(x, y) = sum_{vx, vy} w(vx, vy) in(x, y, vx, vy), summed in a fixed order.
__global__ __launch_bounds__(256) void k_vmoment(const double *__restrict__ in, const double *__restrict__ w, long nxy,
                                                 long nv, double *__restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nxy) return;
  double s = 0.;
  for (long v = 0; v < nv; v++) s += w[v] * in[t + nxy * v];
  out[t] = s;
}

void launch_scale_pairs(const double *exy, double scale, long n, double *lE, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_pairs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, exy, scale, n, lE);
}

void launch_add_density_adj(const double *lrho, double scale, long nxy, long ncell, double *g, hipStream_t s) {
  hipLaunchKernelGGL(k_add_density_adj, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s, lrho, scale, nxy,
                     ncell, g);
}

void launch_vmoment(const double *in, const double *w, long nxy, long nv, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_vmoment, dim3((unsigned)((nxy + 255) / 256)), dim3(256), 0, s, in, w, nxy, nv, out);
}

template <int BX>
static void launch_adj_blk(const Geom &g, const double *c4, const double *ex, const double *ey, double dt,
                           const double *gin, double *lc4, double *axy, const double *scl, bool field_term,
                           hipStream_t s) {
  const size_t lds = sizeof(unsigned long long) * (size_t)(BX + 4) * 6 * 6 * 6;
  const void *fn = (const void *)k_adj_advect4d_blk<BX>;
  (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const long nb = (long)((g.n[0] + BX - 1) / BX) * ((g.n[1] + 1) / 2) * ((g.n[2] + 1) / 2) * ((g.n[3] + 1) / 2);
  hipLaunchKernelGGL(k_adj_advect4d_blk<BX>, dim3((unsigned)nb), dim3(8 * BX), lds, s, c4, ex, ey, g,
                     feet_consts(g), dt, gin, (unsigned long long *)lc4, axy, scl, field_term ? 1 : 0);
}

bool adj_blocked() {
  // VLP4D_ADJ_BLOCK=0: the per-cell scatter (floating-point atomics: one per
  // stencil point and E node) and the half-step scatter
  if (const char *e = getenv("VLP4D_ADJ_BLOCK")) return atoi(e) != 0;
  return true;
}

void launch_adj_advect4d(const Geom &g, const double *c4, const double *ex, const double *ey, double dt,
                         const double *gin, double *lc4, double *lE, bool field_term, hipStream_t s,
                         const AdjWork *wk) {
  if (wk) {
    // reproducible form: fixed-point scale from max|g|, blocked scatter,
    // then the two E gathers
    const long n = (long)g.n[0] * g.n[1] * g.n[2] * g.n[3];
    unsigned long long *mx = (unsigned long long *)(wk->scl + 2);
    hipLaunchKernelGGL(k_fixed_zero, dim3(1), dim3(64), 0, s, mx);
    hipLaunchKernelGGL(k_absmax_bits, dim3(1024), dim3(256), 0, s, gin, n, mx);
    hipLaunchKernelGGL(k_fixed_scale, dim3(1), dim3(64), 0, s, (const unsigned long long *)mx, wk->scl);
    if (g.n[0] % 64 == 0) launch_adj_blk<64>(g, c4, ex, ey, dt, gin, lc4, wk->axy, wk->scl, field_term, s);
    else launch_adj_blk<32>(g, c4, ex, ey, dt, gin, lc4, wk->axy, wk->scl, field_term, s);
    if (field_term) {
      const long np = (long)g.n[0] * g.n[1] * g.n[3], ne = (long)g.n[0] * g.n[1];
      hipLaunchKernelGGL(k_adjE_x, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, wk->axy, g, dt, wk->part);
      hipLaunchKernelGGL(k_adjE_y, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, wk->part, g, dt, lE);
    }
    return;
  }
  const long n = (long)g.n[0] * g.n[1] * g.n[2] * g.n[3];
  hipLaunchKernelGGL(k_adj_advect4d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c4, ex, ey, g,
                     feet_consts(g), dt, gin, lc4, lE, field_term ? 1 : 0);
}

void launch_axis_contract(const double *Mt, int Min, int Nout, const double *in, double *out, long O, long I,
                          bool accumulate, hipStream_t s, const double *M) {
  if (M) {  // tiled form (M = Mt^T)
    const size_t lds = (size_t)Min * 64 * sizeof(double);
    if (I == 1) {
      (void)hipFuncSetAttribute((const void *)k_axis_contract_t<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds);
      hipLaunchKernelGGL(k_axis_contract_t<true>, dim3((unsigned)((O + 63) / 64)), dim3(256), lds, s, M, Min, Nout,
                         in, out, O, I, accumulate ? 1 : 0);
    } else {
      (void)hipFuncSetAttribute((const void *)k_axis_contract_t<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds);
      hipLaunchKernelGGL(k_axis_contract_t<false>, dim3((unsigned)(O * ((I + 63) / 64))), dim3(256), lds, s, M, Min,
                         Nout, in, out, O, I, accumulate ? 1 : 0);
    }
    return;
  }
  const long n = O * Nout * I;
  hipLaunchKernelGGL(k_axis_contract, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Mt, Min, Nout, in, out, O,
                     I, accumulate ? 1 : 0);
}

void launch_c4_block(const Geom &g, const double *c4s, double *blk, hipStream_t s, const double *scl) {
  const long n = (long)(g.n[0] + 4) * (g.n[1] + 4) * (g.n[2] + 4) * (g.n[3] + 4);
  hipLaunchKernelGGL(k_c4_block, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c4s, blk, g, scl);
}

void launch_poisson_T(const double *lE, const double *kx, const double *ky, double *lrho, int N0, int N1,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_poisson_T, dim3((unsigned)((N0 * N1 + 15) / 16)), dim3(256), 0, s, lE, kx, ky, lrho, N0, N1);
}

void launch_adj_advect2d(const double *lrho, const Geom &g, double dth, double *lcxy, hipStream_t s) {
  const long n = (long)g.n[0] * g.n[1] * g.n[2] * g.n[3];
  // VLP4D_ADJ_BLOCK=0 also selects the scatter form here (lcxy zeroed by the caller)
  if (adj_blocked()) {
    const long ng = (long)(g.n[0] + 4) * (g.n[1] + 4) * g.n[2] * g.n[3];
    hipLaunchKernelGGL(k_adj_advect2d_g, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, s, lrho, g, dth, lcxy);
    return;
  }
  hipLaunchKernelGGL(k_adj_advect2d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, lrho, g, dth, lcxy);
}

}  // namespace vlp
