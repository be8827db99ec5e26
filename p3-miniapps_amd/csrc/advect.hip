// advect.hip -- advection, moments and initial-condition kernels (gfx950).
//
// Storage type T is double (plain path) or float (FP32 storage); A is the type
// of the interpolation weights and stencil sums (double, or float for float
// storage on request).  Feet and the density/L2 reductions are always double.
#include "advect_common.hpp"

#include <cstdlib>

namespace vlp {

// ---------------------------------------------------------------------------
// 2-D half-step advection fused with the density reduction:
// VMO/Advection.hpp:299-406 (spline interp_2D :126-163) and VMO/Field.cpp:37-53.
// One lane per (x, y, vy), looping over vx: rho_part[vy][y][x] = sum_vx f_half,
// the same summation order as field_rho's inner loop.  f_half is not stored.
// ---------------------------------------------------------------------------
template <typename T, typename A>
__global__ __launch_bounds__(256) void k_advect2d_rho(const T *__restrict__ cxy, Geom g, double dt,
                                                      double *__restrict__ rho_part, int *__restrict__ err) {
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
void launch_advect2d_rho(const Geom &g, const T *cxy, double dt_half, bool arith32, double *rho_part, int *err,
                         hipStream_t s) {
  const long nthr = (long)g.n[0] * g.n[1] * g.n[3];
  const dim3 grid((unsigned)((nthr + 255) / 256));
  if (arith32 && sizeof(T) == sizeof(float))
    hipLaunchKernelGGL((k_advect2d_rho<T, float>), grid, dim3(256), 0, s, cxy, g, dt_half, rho_part, err);
  else
    hipLaunchKernelGGL((k_advect2d_rho<T, double>), grid, dim3(256), 0, s, cxy, g, dt_half, rho_part, err);
}

// ---------------------------------------------------------------------------
// 4-D advection fused with rho and the L2 norm:
// computeFeet VMO/Advection.hpp:166-218, interp_4D :221-296, advect_4D
// :409-498, field_rho VMO/Field.cpp:37-53, computeL2norm VMO/Diags.cpp:50-76.
// Feet and stencil helpers live in advect_common.hpp.
// ---------------------------------------------------------------------------
// Workgroup order for the 4-D advection.  Workgroups are dealt round-robin to
// the 8 XCDs (each with its own L2); remapping gives XCD k a contiguous range
// of logical blocks, so the workgroups resident on one XCD hold neighbouring
// y rows and share most of their C4 stencil rows in that L2.  Without it the
// workgroups in flight on the whole chip hold one compact range of (y, vy)
// rows, which the shared last-level cache holds instead.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb, int remap) {
  if (!remap || (nb & 7u)) return b;
  return (b & 7u) * (nb >> 3) + (b >> 3);
}

// Lane per (x, y, vy); the lane walks vx in blocks of L cells.  In a block the
// L feet normally share the y- and vy-stencil, their vx-stencils are shifted
// by exactly one per cell and their x-stencils differ by at most one: then the
// block reads the union of its stencils once -- rows of 5 x-values over L+3
// vx-slices -- and every row feeds up to 4 cells (x offset folded into a
// zero-padded 5-weight vector).  Blocks that do not have this structure, and
// the tail block, fall back to the direct gather.
//
// ELDS (requires N0 % 64 == 0, so that every wave has one (y, vy)): the wave
// stages the three E rows its feet interpolate from in LDS once, instead of
// 18 gathered global loads per cell.
template <typename T, typename A, int L, bool ELDS>
__global__ __launch_bounds__(256) void k_advect4d_blk(const T *__restrict__ c4, const double *__restrict__ ex,
                                                      const double *__restrict__ ey, Geom g, FeetConsts fc, double dt,
                                                      int remap, T *__restrict__ fout, double *__restrict__ rho_part,
                                                      double *__restrict__ l2_part, int *__restrict__ err) {
  __shared__ double sh[8];
  extern __shared__ double esh[];  // ELDS: [wave][component][3][N0]
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x, remap);
  const long t = (long)lb * blockDim.x + threadIdx.x;
  const bool active = t < (long)N0 * N1 * N3;
  const int ix = (int)(t % N0);
  const long r = t / N0;
  const int iy = (int)(r % N1);
  const int ivy = (int)(r / N1);
  const double y = g.minPhy[1] + iy * g.dx[1];
  const double vy = g.minPhy[3] + ivy * g.dx[3];
  const double xt1 = half_y(g, y, vy, dt);
  int ip1 = (int)(round(xt1) - 1);
  ELds el;
  if (ELDS) {
    // (y, vy) and hence ip1 are uniform in the wave; a wave past the end
    // stages nothing and only joins the barrier.
    ip1 = __builtin_amdgcn_readfirstlane(ip1);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double *rx = esh + (long)w * 6 * N0;
    el.rx = rx;
    el.ry = rx + 3 * N0;
    el.N0 = N0;
    if (active) {
      for (int q = lane; q < 3 * N0; q += 64) {
        const int k1 = q / N0, jx = q - k1 * N0;
        const long id = jx + (long)N0 * (((ip1 + k1) % N1 + N1) % N1);
        rx[q] = ex[id];
        rx[3 * N0 + q] = ey[id];
      }
    }
    __syncthreads();
  }
  const EGlobal eg{ex, ey, N0, N1, ip1};
  double l2 = 0.;
  int bad = 0;
  if (active) {
    const C4Pitch cp = c4_pitch(N0, N1, N2);
    const long S1 = cp.s1, S2 = cp.s2, S3 = cp.s3;
    const double x = g.minPhy[0] + ix * g.dx[0];
    T *orow = fout + ix + (long)N0 * (iy + (long)N1 * (long)N2 * ivy);
    const long ostride = (long)N0 * N1;
    double rho = 0.;
    for (int v0 = 0; v0 < N2; v0 += L) {
      int ip[L][4];
      A wx[L];         // fractional x position; x weights built after the alignment test
      A eta[L][3][4];  // weights along y, vx, vy
      bool ok[L];
      bool fast = true;
#pragma unroll
      for (int c = 0; c < L; c++) {
        const int ivx = v0 + c;
        ok[c] = false;
        if (ivx < N2) {
          const double vx = g.minPhy[2] + ivx * g.dx[2];
          double xs[4];
          if (ELDS) ok[c] = compute_foot(g, fc, el, xt1, ip1, x, y, vx, vy, dt, xs);
          else ok[c] = compute_foot(g, fc, eg, xt1, ip1, x, y, vx, vy, dt, xs);
          if (ok[c]) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
              const double fj = g.inv_dx[j] * (xs[j] - g.minPhy[j]);
              ip[c][j] = (int)floor(fj);
              const A w = (A)(fj - ip[c][j]);
              if (j == 0) wx[c] = w;
              else bspline_eta<A>(w, eta[c][j - 1]);
            }
          }
        }
        fast = fast && ok[c];
      }
      int bx0 = 0;
      if (fast) {
        bx0 = ip[0][0];
#pragma unroll
        for (int c = 1; c < L; c++) bx0 = ip[c][0] < bx0 ? ip[c][0] : bx0;
#pragma unroll
        for (int c = 0; c < L; c++)
          fast = fast && (ip[c][0] - bx0 <= 1) && (ip[c][1] == ip[0][1]) && (ip[c][3] == ip[0][3]) &&
                 (ip[c][2] == ip[0][2] + c);
      }
      A val[L];
      if (fast) {
        A e0[L][5];
#pragma unroll
        for (int c = 0; c < L; c++) {
          A ex4[4];
          bspline_eta<A>(wx[c], ex4);
          const bool sh1 = ip[c][0] != bx0;
#pragma unroll
          for (int i = 0; i < 5; i++) {
            const A lo = (i < 4) ? ex4[i] : (A)0;
            const A hi = (i > 0) ? ex4[i - 1] : (A)0;
            e0[c][i] = sh1 ? hi : lo;
          }
          val[c] = 0;
        }
        // first stencil position bx0-1 sits at storage index bx0+2; the 5th
        // value (position bx0+3, at most N0+3) may be the next row's first
        // entry -- finite, and weighted 0 in that case.
        const T *base = c4 + (bx0 + 2) + S1 * (ip[0][1] + 2) + S2 * (ip[0][2] + 2) + S3 * (long)(ip[0][3] + 2);
#pragma unroll
        for (int j = 0; j < L + 3; j++) {
#pragma unroll
          for (int k3 = 0; k3 < 4; k3++) {
            A tk[L];
#pragma unroll
            for (int c = 0; c < L; c++) tk[c] = 0;
#pragma unroll
            for (int k1 = 0; k1 < 4; k1++) {
              T rv[5];
              load_row5(base + j * S2 + k3 * S3 + k1 * S1, rv);
#pragma unroll
              for (int c = 0; c < L; c++) {
                const int k2 = j - c;
                if (k2 >= 0 && k2 < 4) {
                  A s = 0;
#pragma unroll
                  for (int i = 0; i < 5; i++) s += e0[c][i] * (A)rv[i];
                  tk[c] += s * eta[c][0][k1];
                }
              }
            }
#pragma unroll
            for (int c = 0; c < L; c++) {
              const int k2 = j - c;
              if (k2 >= 0 && k2 < 4) val[c] += tk[c] * eta[c][1][k2] * eta[c][2][k3];
            }
          }
        }
      } else {
#pragma unroll
        for (int c = 0; c < L; c++) {
          val[c] = 0;
          if (ok[c]) {
            A e4[4][4];
            bspline_eta<A>(wx[c], e4[0]);
#pragma unroll
            for (int j = 0; j < 3; j++)
#pragma unroll
              for (int k = 0; k < 4; k++) e4[j + 1][k] = eta[c][j][k];
            val[c] = gather4d<T, A>(c4, S1, S2, S3, ip[c], e4);
          } else if (v0 + c < N2) {
            bad = 1;
          }
        }
      }
#pragma unroll
      for (int c = 0; c < L; c++) {
        if (v0 + c < N2) {
          orow[(long)(v0 + c) * ostride] = (T)val[c];
          rho += (double)val[c];
          l2 += (double)val[c] * (double)val[c];
        }
      }
    }
    rho_part[ix + (long)N0 * (iy + (long)N1 * ivy)] = rho;
  }
  const double tot = block_sum(l2, sh);
  if (threadIdx.x == 0) l2_part[lb] = tot;
  if (bad) atomicOr(err, 2);
}

long l2_partials(const Geom &g) { return ((long)g.n[0] * g.n[1] * g.n[3] + 255) / 256; }

// Tuning knobs of the 4-D advection, read from the environment at each launch
// (a getenv scan, negligible next to the kernel):
// VLP4D_ADVECT_L=2|4 cells per lane block along vx (default 2: measured at
//   128^4 it beats 4 in every precision mode -- FP64 21.3 vs 23.0 ms -- as the
//   smaller block keeps three waves per SIMD; float storage with double
//   arithmetic needs more than 256 VGPRs at 4 and runs one wave per SIMD),
// VLP4D_XCD_REMAP=1 gives each XCD a contiguous range of (y, vy) rows (off by
//   default: at 128^4 it spreads the concurrently read C4 planes over more
//   than the 256 MB last-level cache and measured ~12% slower),
// VLP4D_ELDS=0 reads E from global memory even when N0 % 64 == 0.
struct AdvectTuning {
  int L = 0, remap = 0, elds = 1;  // L = 0: default block
};
static AdvectTuning advect_tuning() {
  AdvectTuning a;
  if (const char *e = getenv("VLP4D_ADVECT_L")) {
    const int v = atoi(e);
    if (v == 2 || v == 4) a.L = v;
  }
  if (const char *e = getenv("VLP4D_XCD_REMAP")) a.remap = atoi(e) != 0;
  if (const char *e = getenv("VLP4D_ELDS")) a.elds = atoi(e) != 0;
  return a;
}
static int advect_L() {
  const int L = advect_tuning().L;
  return L ? L : 2;
}

// LDS for the staged E rows: 4 waves x 2 components x 3 rows x N0 doubles.
static bool use_elds(const Geom &g) { return advect_tuning().elds && g.n[0] % 64 == 0 && g.n[0] <= 512; }

template <typename T, typename A, int L, bool ELDS>
static void launch_blk(const Geom &g, const T *c4, const double *ex, const double *ey, double dt, T *fout,
                       double *rho_part, double *l2_part, int *err, hipStream_t s) {
  const size_t lds = ELDS ? (size_t)4 * 6 * g.n[0] * sizeof(double) : 0;
  if (ELDS && lds > 48 * 1024)
    (void)hipFuncSetAttribute((const void *)k_advect4d_blk<T, A, L, ELDS>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((k_advect4d_blk<T, A, L, ELDS>), dim3((unsigned)l2_partials(g)), dim3(256), lds, s, c4, ex, ey,
                     g, feet_consts(g), dt, advect_tuning().remap, fout, rho_part, l2_part, err);
}

template <typename T, typename A>
static void launch_advect4d_a(const Geom &g, const T *c4, const double *ex, const double *ey, double dt, T *fout,
                              double *rho_part, double *l2_part, int *err, hipStream_t s) {
  const bool el = use_elds(g);
  if (advect_L() == 2) {
    if (el) launch_blk<T, A, 2, true>(g, c4, ex, ey, dt, fout, rho_part, l2_part, err, s);
    else launch_blk<T, A, 2, false>(g, c4, ex, ey, dt, fout, rho_part, l2_part, err, s);
  } else {
    if (el) launch_blk<T, A, 4, true>(g, c4, ex, ey, dt, fout, rho_part, l2_part, err, s);
    else launch_blk<T, A, 4, false>(g, c4, ex, ey, dt, fout, rho_part, l2_part, err, s);
  }
}

// The y-streaming kernel (advect_stream.hip) when the geometry fits it --
// returns true: rho and the L2 norm are then still to be taken from fout by
// launch_moments; otherwise the vx-blocked kernel with both reductions fused.
template <>
bool launch_advect4d<double>(const Geom &g, const double *c4, const double *ex, const double *ey, const double *exy,
                             double dt, bool, double *fout, double *rho_part, double *l2_part, int *err,
                             hipStream_t s, int ys_tvx, int ys_nb) {
  if (launch_advect4d_ystream<double, double>(g, c4, exy, dt, fout, err, s, ys_tvx, ys_nb)) return true;
  launch_advect4d_a<double, double>(g, c4, ex, ey, dt, fout, rho_part, l2_part, err, s);
  return false;
}

template <>
bool launch_advect4d<float>(const Geom &g, const float *c4, const double *ex, const double *ey, const double *exy,
                            double dt, bool arith32, float *fout, double *rho_part, double *l2_part, int *err,
                            hipStream_t s, int ys_tvx, int ys_nb) {
  const bool ys = arith32 ? launch_advect4d_ystream<float, float>(g, c4, exy, dt, fout, err, s, ys_tvx, ys_nb)
                          : launch_advect4d_ystream<float, double>(g, c4, exy, dt, fout, err, s, ys_tvx, ys_nb);
  if (ys) return true;
  if (arith32) launch_advect4d_a<float, float>(g, c4, ex, ey, dt, fout, rho_part, l2_part, err, s);
  else launch_advect4d_a<float, double>(g, c4, ex, ey, dt, fout, rho_part, l2_part, err, s);
  return false;
}

// moments of a given f: rho partials (sum over vx) and L2 partials.  A
// thread takes VX consecutive x of one (y, vy) and walks vx; for VX > 1 with
// one 16-byte load per vx and four independent sums.  A block covers 256
// points (x, y, vy), so there are l2_partials(g) blocks as for the other
// producers.
template <typename T, int VX>
__global__ __launch_bounds__(256) void k_moments(const T *__restrict__ f, Geom g, double *__restrict__ rho_part,
                                                 double *__restrict__ l2_part) {
  __shared__ double sh[8];
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const long p0 = (long)blockIdx.x * 256 + (long)threadIdx.x * VX;  // first point (x, y, vy)
  const long npt = (long)N0 * N1 * N3;
  double l2 = 0.;
  if (p0 < npt) {
    const int ix = (int)(p0 % N0);
    const long r = p0 / N0;
    const int iy = (int)(r % N1);
    const int ivy = (int)(r / N1);
    const long plane = (long)N0 * N1;
    const T *src = f + ix + plane * ((long)N2 * ivy) + (long)N0 * iy;
    if constexpr (VX == 1) {  // one running sum per thread (measured fastest for double)
      double rho = 0.;
      for (int ivx = 0; ivx < N2; ivx++) {
        const double v = (double)src[plane * ivx];
        rho += v;
        l2 += v * v;
      }
      rho_part[p0] = rho;
    } else {
      double rho[4][VX], sq[4][VX];
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int j = 0; j < VX; j++) rho[u][j] = sq[u][j] = 0.;
      int ivx = 0;
      for (; ivx + 4 <= N2; ivx += 4) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
          typedef T vec_t __attribute__((ext_vector_type(VX)));
          const vec_t w = *(const vec_t *)(src + plane * (ivx + u));
#pragma unroll
          for (int j = 0; j < VX; j++) {
            const double d = (double)w[j];
            rho[u][j] += d;
            sq[u][j] += d * d;
          }
        }
      }
      for (; ivx < N2; ivx++)
#pragma unroll
        for (int j = 0; j < VX; j++) {
          const double d = (double)src[plane * ivx + j];
          rho[0][j] += d;
          sq[0][j] += d * d;
        }
#pragma unroll
      for (int j = 0; j < VX; j++) {
        rho_part[p0 + j] = (rho[0][j] + rho[1][j]) + (rho[2][j] + rho[3][j]);
        l2 += (sq[0][j] + sq[1][j]) + (sq[2][j] + sq[3][j]);
      }
    }
  }
  const double tot = block_sum(l2, sh);
  if (threadIdx.x == 0) l2_part[blockIdx.x] = tot;
}

template <typename T> void launch_moments(const Geom &g, const T *f, double *rho_part, double *l2_part, hipStream_t s) {
  // float: four consecutive x per thread (one 16-byte load per vx) where the
  // rows hold whole 16-byte chunks (measured at 128^4: 0.22 ms against 0.32
  // for one x per thread); double: one x and one running sum per thread
  // (0.39 ms against 0.42 for two x per thread and 0.41 for four sums)
  constexpr int VX = sizeof(T) == sizeof(float) ? 4 : 1;
  if (VX > 1 && g.n[0] % VX == 0)
    hipLaunchKernelGGL((k_moments<T, VX>), dim3((unsigned)l2_partials(g)), dim3(256 / VX), 0, s, f, g, rho_part,
                       l2_part);
  else
    hipLaunchKernelGGL((k_moments<T, 1>), dim3((unsigned)l2_partials(g)), dim3(256), 0, s, f, g, rho_part, l2_part);
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
    case 30:    // extension: strong Landau damping (no reference counterpart)
    case 31: {  // extension: two-stream (no reference counterpart)
      const double vy = g.minPhy[3] + ivy * g.dx[3];
      const double vx = g.minPhy[2] + ivx * g.dx[2];
      const double y = g.minPhy[1] + iy * g.dx[1];
      const double x = g.minPhy[0] + ix * g.dx[0];
      if (g.idcase == 30) {
        val = (1. / (2 * M_PI)) * exp(-0.5 * (vx * vx + vy * vy)) * (1 + 0.5 * cos(0.5 * x) + 0.5 * cos(0.5 * y));
      } else {
        const double a = vx - 2.4, b = vx + 2.4;
        val = (1. / (4 * M_PI)) * (1 + 1e-3 * cos(0.2 * x)) * exp(-0.5 * (vy * vy)) *
              (exp(-0.5 * (a * a)) + exp(-0.5 * (b * b)));
      }
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

int spline_variant_bits(const Geom &g);

int kernel_variants(const Geom &g, bool fp32_storage, bool arith32) {
  const AdvectTuning t = advect_tuning();
  const bool remap = t.remap && (l2_partials(g) % 8 == 0);
  const int L = advect_L();
  const int ys = ystream_tvx(g, fp32_storage, arith32);
  const bool xy = xy_fused_groups(g, fp32_storage ? sizeof(float) : sizeof(double)) != 0;
  return spline_variant_bits(g) | 4 | (L == 4 ? 8 : 0) | (use_elds(g) ? 16 : 0) | (remap ? 32 : 0) |
         (arith32 ? 64 : 0) | (ys ? 128 | (ys << 8) : 0) | (xy ? 4096 : 0);
}

#define VLP_INST(T)                                                                                        \
  template void launch_init_case<T>(const Geom &, T *, hipStream_t);                                       \
  template void launch_advect2d_rho<T>(const Geom &, const T *, double, bool, double *, int *, hipStream_t); \
  template void launch_moments<T>(const Geom &, const T *, double *, double *, hipStream_t);
VLP_INST(double)
VLP_INST(float)

}  // namespace vlp
