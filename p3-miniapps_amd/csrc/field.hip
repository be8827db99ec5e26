// field.hip -- field solve and diagnostics kernels (FP64, gfx950).
#include "device_common.hpp"

namespace vlp {

// ---------------------------------------------------------------------------
// Field solve, FP64.  rho = dvx*dvy*sum rho_part over its outer planes
// (vy, and the vx groups of the fused xy kernel; VMO/Field.cpp:37-53),
// rho -= 1 for idcase 10/20 (VMO/Field.cpp:110-126), then solve_poisson_fftw
// (VMO/Efield.cpp:54-106) with the FFT conventions of lib/OpenMP_FFT.hpp:
// unnormalised r2c along x, c2c along y, normalisation in k-space, inverse
// c2c along y, c2r along x (imaginary parts of DC/Nyquist ignored).  The
// transforms are direct DFTs with exactly reduced twiddle tables; their loops
// are unrolled so that several terms' twiddle loads are in flight (each sum
// keeps its order).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rho_rfftx(const double *__restrict__ rho_part, long rho_outer, Geom g,
                                                   FieldBufs b, int sub1) {
  extern __shared__ double row[];
  const int N0 = g.n[0], N1 = g.n[1];
  const int iy = blockIdx.x;
  const int nxh = N0 / 2 + 1;
  for (int ix = threadIdx.x; ix < N0; ix += blockDim.x) {
    double s = 0.;
#pragma unroll 8
    for (long o = 0; o < rho_outer; o++) s += rho_part[ix + (long)N0 * (iy + (long)N1 * o)];
    double r = s * g.dx[2] * g.dx[3];
    if (sub1) r -= 1.;
    row[ix] = r;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nxh; k += blockDim.x) {
    double sr = 0., si = 0.;
    int m = 0;
#pragma unroll 8
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

// Slice s of the density partials: tmp[s][i] = sum of planes o = s, s + S, ...
__global__ __launch_bounds__(256) void k_rho_slices(const double *__restrict__ rho_part, long rho_outer, long npt,
                                                    double *__restrict__ tmp) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y, S = gridDim.y;
  if (i >= npt) return;
  double acc = 0.;
  for (long o = s; o < rho_outer; o += S) acc += rho_part[i + npt * o];
  tmp[i + npt * s] = acc;
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
#pragma unroll 8
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
    // the three fields side by side: one twiddle load per term and three
    // independent sums (each in the same order as alone)
    double sr[3] = {0., 0., 0.}, si[3] = {0., 0., 0.};
    int m = 0;
#pragma unroll 8
    for (int q = 0; q < N1; q++) {
      const double wr = b.twy[2 * m], wi = -b.twy[2 * m + 1];  // exp(+2 pi i m/N)
#pragma unroll
      for (int fld = 0; fld < 3; fld++) {
        const double *h = hat + 2 * N1 * fld;
        sr[fld] += h[2 * q] * wr - h[2 * q + 1] * wi;
        si[fld] += h[2 * q] * wi + h[2 * q + 1] * wr;
      }
      m += iy;
      if (m >= N1) m -= N1;
    }
#pragma unroll
    for (int fld = 0; fld < 3; fld++) {
      double *dst = b.spec2 + 2 * ((long)nxh * N1 * fld + k + (long)nxh * iy);
      dst[0] = sr[fld];
      dst[1] = si[fld];
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
    // the three fields side by side (one twiddle load per term, each sum in
    // the same order as alone)
    double out[3];
#pragma unroll
    for (int fld = 0; fld < 3; fld++) out[fld] = rows[2 * nxh * fld];
    int m = ix;
#pragma unroll 8
    for (int k = 1; k < (N0 + 1) / 2; k++) {
      const double wr = b.twx[2 * m], wi = -b.twx[2 * m + 1];
#pragma unroll
      for (int fld = 0; fld < 3; fld++) {
        const double *X = rows + 2 * nxh * fld;
        out[fld] += 2.0 * (X[2 * k] * wr - X[2 * k + 1] * wi);
      }
      m += ix;
      if (m >= N0) m -= N0;
    }
    if ((N0 & 1) == 0)
#pragma unroll
      for (int fld = 0; fld < 3; fld++) out[fld] += rows[2 * nxh * fld + 2 * (N0 / 2)] * ((ix & 1) ? -1.0 : 1.0);
    const long id = ix + (long)N0 * iy;
    b.rho[id] = out[0];
    b.ex[id] = out[1];
    b.ey[id] = out[2];
    b.exy[2 * id] = out[1];
    b.exy[2 * id + 1] = out[2];
    pm += out[0];
    pe += out[1] * out[1] + out[2] * out[2];
    pex += out[1] * out[1];
    pey += out[2] * out[2];
  }
  const double s0 = block_sum(pm, red);
  const double s1 = block_sum(pe, red);
  const double s2 = block_sum(pex, red);
  const double s3 = block_sum(pey, red);
  if (threadIdx.x == 0) {
    b.diag_part[4 * iy + 0] = s0;
    b.diag_part[4 * iy + 1] = s1;
    b.diag_part[4 * iy + 2] = s2;
    b.diag_part[4 * iy + 3] = s3;
  }
}

// Analytic fields of idcase 2 and 6 (VMO/Field.cpp:81-108); rho is kept.
__global__ __launch_bounds__(256) void k_analytic_field(const double *__restrict__ rho_part, long rho_outer, Geom g,
                                                        FieldBufs b) {
  __shared__ double red[8];
  const int N0 = g.n[0], N1 = g.n[1];
  const int iy = blockIdx.x;
  double pm = 0., pe = 0., pex = 0., pey = 0.;
  for (int ix = threadIdx.x; ix < N0; ix += blockDim.x) {
    double s = 0.;
#pragma unroll 8
    for (long o = 0; o < rho_outer; o++) s += rho_part[ix + (long)N0 * (iy + (long)N1 * o)];
    const double r = s * g.dx[2] * g.dx[3];
    const double e = (g.idcase == 2) ? -(g.minPhy[0] + ix * g.dx[0]) : -(g.minPhy[1] + iy * g.dx[1]);
    const long id = ix + (long)N0 * iy;
    b.rho[id] = r;
    b.ex[id] = e;
    b.ey[id] = 0.;
    b.exy[2 * id] = e;
    b.exy[2 * id + 1] = 0.;
    pm += r;
    pe += e * e;
    pex += e * e;
  }
  const double s0 = block_sum(pm, red);
  const double s1 = block_sum(pe, red);
  const double s2 = block_sum(pex, red);
  const double s3 = block_sum(pey, red);
  if (threadIdx.x == 0) {
    b.diag_part[4 * iy + 0] = s0;
    b.diag_part[4 * iy + 1] = s1;
    b.diag_part[4 * iy + 2] = s2;
    b.diag_part[4 * iy + 3] = s3;
  }
}

// Diags::compute and computeL2norm, VMO/Diags.cpp:16-76.
// The record index is iter, or -- iter_dev non-null, for launches replayed
// from a captured graph -- *iter_dev + 1, which is then stored back.
__global__ __launch_bounds__(256) void k_diag_final(Geom g, FieldBufs b, const double *__restrict__ l2_part, long n_l2,
                                                    int iter, int with_l2, int *__restrict__ iter_dev) {
  __shared__ double red[8];
  const int N1 = g.n[1];
  double s[4];
  for (int c = 0; c < 4; c++) {
    double v = 0.;
    for (int iy = threadIdx.x; iy < N1; iy += blockDim.x) v += b.diag_part[4 * iy + c];
    s[c] = block_sum(v, red);
  }
  double l2 = 0.;
  if (with_l2) {
    double v = 0.;
    // unrolled so that the independent loads are in flight together (the
    // additions keep their order)
#pragma unroll 8
    for (long i = threadIdx.x; i < n_l2; i += blockDim.x) v += l2_part[i];
    l2 = block_sum(v, red);
  }
  if (threadIdx.x == 0) {
    if (iter_dev) {
      iter = *iter_dev + 1;
      *iter_dev = iter;
    }
    const int nb = b.nbiter + 1;
    if (iter < 0 || iter >= nb) return;
    double it_nrj = sqrt(s[1] * g.dx[0] * g.dx[1]);
    it_nrj = it_nrj > 1.e-30 ? log(it_nrj) : -1.e9;
    b.diag[0 * nb + iter] = it_nrj;
    b.diag[1 * nb + iter] = sqrt(0.5 * s[2] * g.dx[0] * g.dx[1]);
    b.diag[2 * nb + iter] = sqrt(0.5 * s[3] * g.dx[0] * g.dx[1]);
    b.diag[3 * nb + iter] = s[0] * g.dx[0] * g.dx[1];
    if (with_l2) b.diag[4 * nb + iter] = sqrt(l2 * g.dx[0] * g.dx[1] * g.dx[2] * g.dx[3]);
  }
}

void launch_field(const Geom &g, const FieldBufs &b, const double *rho_part, long rho_outer, const double *l2_part,
                  long n_l2, int iter, bool record_diag, hipStream_t s, int *iter_dev) {
  const int N0 = g.n[0], N1 = g.n[1];
  const int nxh = N0 / 2 + 1;
  if (rho_outer > 64 && b.rho_tmp) {
    const long npt = (long)N0 * N1;
    hipLaunchKernelGGL(k_rho_slices, dim3((unsigned)((npt + 255) / 256), RHO_SLICES), dim3(256), 0, s, rho_part,
                       rho_outer, npt, b.rho_tmp);
    rho_part = b.rho_tmp;
    rho_outer = RHO_SLICES;
  }
  if (g.idcase == 2 || g.idcase == 6) {
    hipLaunchKernelGGL(k_analytic_field, dim3(N1), dim3(256), 0, s, rho_part, rho_outer, g, b);
  } else {
    const int sub1 = (g.idcase == 10 || g.idcase == 20);
    hipLaunchKernelGGL(k_rho_rfftx, dim3(N1), dim3(256), sizeof(double) * N0, s, rho_part, rho_outer, g, b, sub1);
    hipLaunchKernelGGL(k_poisson_y, dim3(nxh), dim3(256), sizeof(double) * 8 * N1, s, g, b);
    hipLaunchKernelGGL(k_c2rx_diag, dim3(N1), dim3(256), sizeof(double) * (6 * nxh + 8), s, g, b);
  }
  if (record_diag)
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, s, g, b, l2_part, n_l2, iter, l2_part != nullptr ? 1 : 0,
                       iter_dev);
}

}  // namespace vlp
