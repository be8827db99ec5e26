// xy_fused.hip -- spline prefilter along x and y, 2-D half-step advection and
// density, fused per (vx, vy) plane (gfx950).
//
// Reference steps (one rank): exchangeHalo + Spline::computeCoeff_xy
// (VMO/Communication.cpp:541-570, VMO/Spline.hpp:384-426 -> :114-199), then
// advect_2D_xy (VMO/Advection.hpp:299-406, spline interp_2D :126-163) and
// field_rho (VMO/Field.cpp:37-53).  For one (vx, vy) plane all of these touch
// only that plane, so a workgroup keeps the plane's coefficients in LDS:
//   1. x recursions: f rows straight from global memory (each lane a chunk of
//      C values, the chunked scan of spline.hip), coefficients into LDS rows;
//   2. y recursions on the LDS columns, in place;
//   3. the CXY plane (positions -2..N+1 in x and y) leaves LDS as 16-byte
//      non-temporal stores;
//   4. the 2-D advection of every (x, y) of the plane reads the stencil from
//      LDS; f at the half step is never stored, it is summed over the
//      workgroup's vx planes into density partials.
// The workgroup owns one vy and a group of N2/G consecutive vx planes; the
// density partials rho_part[grp][vy][y][x] are summed over (grp, vy) by the
// field solve.
#include "spline_scan.hpp"

#include <cstdlib>

namespace vlp {

namespace {

template <typename A> __device__ __forceinline__ void eta4(A w, A e[4]) { bspline_eta<A>(w, e); }

}  // namespace

// Block of NT = 512 lanes.  x pass: line y = r0 + tid / Px, lane p = tid % Px;
// y pass: column xc = c0 + tid / Py, lane p = tid % Py; advection: column
// x = tid % N0, rows y0 .. y0+ylen-1 (ylen <= NPT) of segment tid / N0, the
// density sums of those points in registers across the vx planes.
constexpr int XY_NT = 512, XY_NPT = 32;

template <typename T, typename A, int C>
__global__ __launch_bounds__(XY_NT) void k_xy_adv2d(const T *__restrict__ f, T *__restrict__ cxy, Geom g, double dth,
                                                   int G, double *__restrict__ rho_part, int *__restrict__ err,
                                                   SplineConsts sc, double aC, double aN3x, double aN3y) {
  extern __shared__ __align__(16) unsigned char xy_smem[];
  T *pl = (T *)xy_smem;
  const int N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const int E0 = N0 + 4, E1 = N1 + 4, R = E0 | 1;
  // y weights of the half-step foot per row (they depend on y and vy only):
  // ety[y][4], and the stencil base row jy0 in jty[y]
  A *ety = (A *)(xy_smem + (((size_t)E1 * R * sizeof(T) + 15) & ~(size_t)15));
  int *jty = (int *)(ety + 4 * N1);
  const int Px = N0 / C, Py = N1 / C;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int grp = blockIdx.x % G, ivy = blockIdx.x / G;
  const int VG = N2 / G, v0 = grp * VG;
  // advection lanes
  const int nseg = XY_NT / N0, seg = tid / N0, xi = tid - seg * N0;
  const int ylen = (N1 + nseg - 1) / nseg, y0 = seg * ylen;
  const bool aact = seg < nseg;
  double rho[XY_NPT];
#pragma unroll
  for (int k = 0; k < XY_NPT; k++) rho[k] = 0.;
  int bad = 0;
  const double vy = g.minPhy[3] + ivy * g.dx[3];
  const double xx = g.minPhy[0] + xi * g.dx[0];
  for (int y = tid; y < N1; y += XY_NT) {
    const double yy = g.minPhy[1] + y * g.dx[1];
    const double fy = ((yy - dth * vy) - g.minPhy[1]) * g.inv_dx[1];
    const int jy0 = (int)floor(fy);
    eta4<A>((A)(fy - (double)jy0), ety + 4 * y);
    jty[y] = jy0;
  }
  // (the first plane's barriers order these writes before their reads)
  // x pass lane: first plane's chunk in registers
  const int px = tid & (Px - 1), xl = tid / Px;
  for (int ivx = v0; ivx < v0 + VG; ivx++) {
    // 1. x recursions, rows 2..N1+1 of the plane
    for (int r0 = 0; r0 < N1; r0 += XY_NT / Px) {
      const int y = r0 + xl;
      double v[C];
      const T *src = f + ((long)ivx + (long)N2 * ivy) * N0 * N1 + (long)(y < N1 ? y : 0) * N0 + px * C;
#pragma unroll
      for (int i = 0; i < C; i++) v[i] = (double)src[i];
      double cm2, cm1, cN, cN1;
      scan_recursion<C>(v, px, Px, sc, aC, aN3x, cm2, cm1, cN, cN1);
      if (y < N1) {
        T *row = pl + (y + 2) * R;
#pragma unroll
        for (int i = 0; i < C; i++) row[2 + px * C + i] = (T)v[i];
        if (px == 0) {
          row[0] = (T)cm2;
          row[1] = (T)cm1;
        }
        if (px == Px - 1) {
          row[N0 + 2] = (T)cN;
          row[N0 + 3] = (T)cN1;
        }
      }
    }
    __syncthreads();
    // 2. y recursions on columns 0..E0-1 (a wave stops when its first
    // column is past the plane: its lines are whole within the wave)
    {
      const int py = tid & (Py - 1), yl = tid / Py, yl0 = (wave * 64) / Py;
      for (int c0 = 0; c0 + yl0 < E0; c0 += XY_NT / Py) {
        const int xc = c0 + yl;
        const bool act = xc < E0;
        double w[C];
        const T *col = pl + (act ? xc : 0) + (2 + py * C) * R;
#pragma unroll
        for (int i = 0; i < C; i++) w[i] = (double)col[i * R];
        double cm2, cm1, cN, cN1;
        scan_recursion<C>(w, py, Py, sc, aC, aN3y, cm2, cm1, cN, cN1);
        if (act) {
          T *cw = pl + xc;
#pragma unroll
          for (int i = 0; i < C; i++) cw[(2 + py * C + i) * R] = (T)w[i];
          if (py == 0) {
            cw[0] = (T)cm2;
            cw[R] = (T)cm1;
          }
          if (py == Py - 1) {
            cw[(N1 + 2) * R] = (T)cN;
            cw[(N1 + 3) * R] = (T)cN1;
          }
        }
      }
    }
    __syncthreads();
    // 3. store the CXY plane (positions -2..N+1 in x and y).  The plane is
    // contiguous in global memory, so it leaves as 16-byte chunks, lane after
    // lane (row and column of a chunk tracked incrementally), as
    // non-temporal stores (the plane is read again only by the vx spline
    // pass, after the whole kernel); per element when the rows do not split
    // into whole chunks.
    {
      T *dst = cxy + ((long)ivx + (long)N2 * ivy) * E0 * E1;
      constexpr int VPC = 16 / (int)sizeof(T);
      if (E0 % VPC == 0 && ((size_t)dst & 15) == 0) {
        const int cpr = E0 / VPC, nch = cpr * E1;
        const int drow = XY_NT / cpr, dcol = XY_NT - drow * cpr;
        int row = tid / cpr, col = tid - row * cpr;
        for (int k = tid; k < nch; k += XY_NT) {
          const T *s = pl + row * R + VPC * col;
          typedef T chunk_t __attribute__((ext_vector_type(VPC)));
          chunk_t v;
#pragma unroll
          for (int e = 0; e < VPC; e++) v[e] = s[e];
          __builtin_nontemporal_store(v, (chunk_t *)(dst + (long)VPC * k));
          row += drow;
          col += dcol;
          if (col >= cpr) {
            col -= cpr;
            row++;
          }
        }
      } else {
        for (int yr = wave; yr < E1; yr += XY_NT / 64)
          for (int xc = lane; xc < E0; xc += 64) dst[(long)yr * E0 + xc] = pl[yr * R + xc];
      }
    }
    // 4. 2-D advection at (x - dt/2 vx, y - dt/2 vy), interp_2D's spline branch
    if (aact) {
      const double vx = g.minPhy[2] + ivx * g.dx[2];
      const double fx = ((xx - dth * vx) - g.minPhy[0]) * g.inv_dx[0];
      const int jx0 = (int)floor(fx);
      A etax[4];
      eta4<A>((A)(fx - (double)jx0), etax);
      const bool xbad = (jx0 < -1 || jx0 > N0 - 1);
      const T *pc = pl + (xbad ? 0 : jx0 + 1);  // stencil column jx0-1 at LDS column jx0+1
      A gw[4];  // x-interpolated rows jw .. jw+3 (LDS rows)
      int jw = -1000;
      constexpr int GP = sizeof(A) == sizeof(float) ? 4 : 2;
      static_assert(XY_NPT % GP == 0, "groups tile the point range");
      // Points go in groups of GP consecutive y (four for float arithmetic,
      // two for double, whose tables take twice the registers).  The
      // segment's lanes share y, so jy0 and the group's path are uniform in a
      // wave when N0 is a multiple of 64.  When the GP stencils each move down
      // by one row from the previous point's (the usual case), the group's
      // tables and its GP new rows are all read before the arithmetic (one
      // LDS round trip per group instead of several per point); otherwise the
      // points go one at a time.
#pragma unroll
      for (int k0 = 0; k0 < XY_NPT; k0 += GP) {
        const int yg = y0 + k0;
        if (k0 < ylen && yg < N1) {
          int jr[GP];
          A ey[GP][4];
#pragma unroll
          for (int j = 0; j < GP; j++) {
            const int y = min(yg + j, N1 - 1);
            jr[j] = jty[y] + 1;  // stencil row jy0-1 at LDS row jy0+1
#pragma unroll
            for (int jy = 0; jy < 4; jy++) ey[j][jy] = ety[4 * y + jy];
          }
          const bool full = k0 + GP - 1 < ylen && yg + GP - 1 < N1;
          bool slide = full && !xbad && jr[0] == jw + 1;
#pragma unroll
          for (int j = 0; j < GP; j++) slide = slide && jr[j] >= 0 && jr[j] <= N1 && jr[j] == jr[0] + j;
          if (slide) {
            T nr[GP][4];  // rows jw+4 .. jw+3+GP
#pragma unroll
            for (int j = 0; j < GP; j++) {
              const T *p = pc + (jw + 4 + j) * R;
#pragma unroll
              for (int jx = 0; jx < 4; jx++) nr[j][jx] = p[jx];
            }
#pragma unroll
            for (int j = 0; j < GP; j++) {
              gw[0] = gw[1];
              gw[1] = gw[2];
              gw[2] = gw[3];
              A s = 0;
#pragma unroll
              for (int jx = 0; jx < 4; jx++) s += etax[jx] * (A)nr[j][jx];
              gw[3] = s;
              A val = 0;
#pragma unroll
              for (int jy = 0; jy < 4; jy++) val += ey[j][jy] * gw[jy];
              rho[k0 + j] += (double)val;
            }
            jw = jr[GP - 1];
          } else {
#pragma unroll
            for (int j = 0; j < GP; j++) {
              const int k = k0 + j, y = yg + j;
              if (k < ylen && y < N1) {
                const int jy0 = jr[j] - 1;
                A val = 0;
                if (xbad || jy0 < -1 || jy0 > N1 - 1) {
                  bad = 1;  // interp_2D returns 1 and leaves 0 (VMO/Advection.hpp:148-151)
                } else {
                  const int r = jr[j];
                  if (r == jw + 1) {
                    gw[0] = gw[1];
                    gw[1] = gw[2];
                    gw[2] = gw[3];
                    const T *p = pc + (r + 3) * R;
                    A s = 0;
#pragma unroll
                    for (int jx = 0; jx < 4; jx++) s += etax[jx] * (A)p[jx];
                    gw[3] = s;
                  } else if (r != jw) {
#pragma unroll
                    for (int jy = 0; jy < 4; jy++) {
                      const T *p = pc + (r + jy) * R;
                      A s = 0;
#pragma unroll
                      for (int jx = 0; jx < 4; jx++) s += etax[jx] * (A)p[jx];
                      gw[jy] = s;
                    }
                  }
                  jw = r;
#pragma unroll
                  for (int jy = 0; jy < 4; jy++) val += ey[j][jy] * gw[jy];
                }
                rho[k] += (double)val;
              }
            }
          }
        }
        __builtin_amdgcn_sched_barrier(0);  // one group at a time (register pressure)
      }
    }
    __syncthreads();
  }
  if (aact) {
    double *rp = rho_part + ((long)grp * N3 + ivy) * N0 * N1;
#pragma unroll
    for (int k = 0; k < XY_NPT; k++) {
      const int y = y0 + k;
      if (k < ylen && y < N1) rp[(long)y * N0 + xi] = rho[k];
    }
  }
  if (bad) atomicOr(err, 1);
}

namespace {

// LDS bytes: the plane, then the y-weight and y-row tables.
size_t xy_lds(const Geom &g, size_t tsz) {
  const size_t plane = (((size_t)(g.n[1] + 4) * ((g.n[0] + 4) | 1) * tsz) + 15) & ~(size_t)15;
  return plane + (size_t)g.n[1] * (4 * sizeof(double) + sizeof(int));
}

// C of the fused kernel for a geometry (0: not eligible): N0 = Px C and
// N1 = Py C with Px, Py powers of two, the plane fits in LDS, every lane owns
// at most XY_NPT advection points.
int xy_chunk(const Geom &g, size_t tsz) {
  const int N0 = g.n[0], N1 = g.n[1];
  for (int C : {16, 12}) {
    if (N0 % C || N1 % C) continue;
    const int Px = N0 / C, Py = N1 / C;
    if ((Px & (Px - 1)) || (Py & (Py - 1)) || Px > 16 || Py > 16) continue;
    if (N0 > XY_NT) continue;
    const int nseg = XY_NT / N0, ylen = (N1 + nseg - 1) / nseg;
    if (ylen > XY_NPT) continue;
    if (xy_lds(g, tsz) > 160 * 1024) continue;
    return C;
  }
  return 0;
}

}  // namespace

int xy_fused_groups(const Geom &g, size_t tsz) {
  if (const char *e = getenv("VLP4D_XY_FUSED"))
    if (atoi(e) == 0) return 0;
  if (!xy_chunk(g, tsz)) return 0;
  // about two rounds of workgroups over the 256 CUs (one or two per CU)
  const int per_cu = xy_lds(g, tsz) > 80 * 1024 ? 1 : 2;
  int G = 1;
  while (G * 2 <= g.n[2] && g.n[2] % (G * 2) == 0 && (long)G * g.n[3] < 256L * per_cu) G *= 2;
  return G;
}

template <typename T, typename A, int C>
static void launch_xy_t(const Geom &g, const T *f, T *cxy, double dth, int G, double *rho_part, int *err,
                        hipStream_t s) {
  const SplineConsts sc = spline_consts();
  double aC = 1.0, aN3x = 1.0, aN3y = 1.0;
  for (int i = 0; i < C; i++) aC *= sc.alpha;
  for (int i = 0; i < g.n[0] + 3; i++) aN3x *= sc.alpha;
  for (int i = 0; i < g.n[1] + 3; i++) aN3y *= sc.alpha;
  const size_t lds = xy_lds(g, sizeof(T));
  const void *fn = (const void *)k_xy_adv2d<T, A, C>;
  if (lds > 64 * 1024) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((k_xy_adv2d<T, A, C>), dim3((unsigned)(G * g.n[3])), dim3(XY_NT), lds, s, f, cxy, g, dth, G,
                     rho_part, err, sc, aC, aN3x, aN3y);
}

template <typename T, typename A>
bool launch_xy_adv2d(const Geom &g, const T *f, T *cxy, double dt_half, double *rho_part, int *err, hipStream_t s) {
  const int G = xy_fused_groups(g, sizeof(T));
  if (!G) return false;
  switch (xy_chunk(g, sizeof(T))) {
    case 16: launch_xy_t<T, A, 16>(g, f, cxy, dt_half, G, rho_part, err, s); return true;
    case 12: launch_xy_t<T, A, 12>(g, f, cxy, dt_half, G, rho_part, err, s); return true;
    default: return false;
  }
}

template bool launch_xy_adv2d<double, double>(const Geom &, const double *, double *, double, double *, int *,
                                              hipStream_t);
template bool launch_xy_adv2d<float, double>(const Geom &, const float *, float *, double, double *, int *,
                                             hipStream_t);
template bool launch_xy_adv2d<float, float>(const Geom &, const float *, float *, double, double *, int *,
                                            hipStream_t);

}  // namespace vlp
