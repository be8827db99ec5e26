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
//    storage layout: positions -3..N+2, stride N_d + 6);
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
  const long S1 = N0 + 6, S2 = S1 * (N1 + 6), S3 = S2 * (N2 + 6);
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

// Contiguous copy of the coefficient block (positions -2..N+1 per dimension)
// out of the C4 storage layout (positions -3..N+2).
__global__ __launch_bounds__(256) void k_c4_block(const double *__restrict__ c4s, double *__restrict__ blk, Geom g) {
  const long E0 = g.n[0] + 4, E1 = g.n[1] + 4, E2 = g.n[2] + 4, E3 = g.n[3] + 4;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E0 * E1 * E2 * E3) return;
  const long a = t % E0, b = (t / E0) % E1, c = (t / (E0 * E1)) % E2, d = t / (E0 * E1 * E2);
  const long F0 = g.n[0] + 6, F1 = g.n[1] + 6, F2 = g.n[2] + 6;
  blk[t] = c4s[(a + 1) + F0 * ((b + 1) + F1 * ((c + 1) + F2 * (d + 1)))];
}

// lambda_rho(j) = sum_i lE_x(i) kx(i - j) + lE_y(i) ky(i - j): the transpose
// of the periodic E-from-rho convolution, kx/ky its response to a unit
// density at the origin (indices mod N0, N1).
__global__ __launch_bounds__(256) void k_poisson_T(const double *__restrict__ lE, const double *__restrict__ kx,
                                                   const double *__restrict__ ky, double *__restrict__ lrho, int N0,
                                                   int N1) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N0 * N1) return;
  const int jx = t % N0, jy = t / N0;
  double s = 0.;
  for (int iy = 0; iy < N1; iy++) {
    const int dy = (iy - jy + N1) % N1;
    for (int ix = 0; ix < N0; ix++) {
      const int dx = (ix - jx + N0) % N0;
      const long i = (long)iy * N0 + ix, k = (long)dy * N0 + dx;
      s += lE[2 * i] * kx[k] + lE[2 * i + 1] * ky[k];
    }
  }
  lrho[t] = s;
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

void launch_adj_advect4d(const Geom &g, const double *c4, const double *ex, const double *ey, double dt,
                         const double *gin, double *lc4, double *lE, bool field_term, hipStream_t s) {
  const long n = (long)g.n[0] * g.n[1] * g.n[2] * g.n[3];
  hipLaunchKernelGGL(k_adj_advect4d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c4, ex, ey, g,
                     feet_consts(g), dt, gin, lc4, lE, field_term ? 1 : 0);
}

void launch_axis_contract(const double *Mt, int Min, int Nout, const double *in, double *out, long O, long I,
                          bool accumulate, hipStream_t s) {
  const long n = O * Nout * I;
  hipLaunchKernelGGL(k_axis_contract, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Mt, Min, Nout, in, out, O,
                     I, accumulate ? 1 : 0);
}

void launch_c4_block(const Geom &g, const double *c4s, double *blk, hipStream_t s) {
  const long n = (long)(g.n[0] + 4) * (g.n[1] + 4) * (g.n[2] + 4) * (g.n[3] + 4);
  hipLaunchKernelGGL(k_c4_block, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c4s, blk, g);
}

void launch_poisson_T(const double *lE, const double *kx, const double *ky, double *lrho, int N0, int N1,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_poisson_T, dim3((unsigned)((N0 * N1 + 255) / 256)), dim3(256), 0, s, lE, kx, ky, lrho, N0,
                     N1);
}

void launch_adj_advect2d(const double *lrho, const Geom &g, double dth, double *lcxy, hipStream_t s) {
  const long n = (long)g.n[0] * g.n[1] * g.n[2] * g.n[3];
  hipLaunchKernelGGL(k_adj_advect2d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, lrho, g, dth, lcxy);
}

}  // namespace vlp
