// advect_common.hpp -- characteristic feet and stencil helpers shared by the
// 4-D advection kernels (advect.hip, advect_stream.hip).
#pragma once

#include "device_common.hpp"

namespace vlp {

// Domain limits of the foot bound check, VMO/Advection.hpp:432-433.
struct FeetConsts {
  double rxwidth[4];
  double locmin[4], locmax[4];  // VMO/Advection.hpp:432-433
};

#pragma clang fp contract(off)
inline FeetConsts feet_consts(const Geom &g) {
  FeetConsts f;
  for (int j = 0; j < 4; j++) {
    f.rxwidth[j] = g.maxPhy[j] - g.minPhy[j];
    f.locmin[j] = g.minPhy[j] + 0 * g.dx[j] - g.dx[j];
    f.locmax[j] = g.minPhy[j] + (g.n[j] - 1) * g.dx[j] + g.dx[j];
  }
  return f;
}
#pragma clang fp contract(on)

// E-field sources for compute_foot: (Ex, Ey) at stencil node (jx, k1), jx the
// wrapped x index, k1 in {0,1,2} the stencil row ip1 + k1 (wrapped).
struct EGlobal {  // straight from the N1 x N0 arrays
  const double *__restrict__ ex;
  const double *__restrict__ ey;
  int N0, N1, ip1;
  __device__ __forceinline__ void at(int jx, int k1, double &vx, double &vy) const {
    const long id = jx + (long)N0 * (((ip1 + k1) % N1 + N1) % N1);
    vx = ex[id];
    vy = ey[id];
  }
};
struct ELds {  // the three rows, staged in LDS as [3][N0] per component
  const double *rx, *ry;
  int N0;
  __device__ __forceinline__ void at(int jx, int k1, double &vx, double &vy) const {
    vx = rx[k1 * N0 + jx];
    vy = ry[k1 * N0 + jx];
  }
};

// Normalised half-step y position (y - dt/2 vy - min) / dy of computeFeet;
// it depends on (y, vy) only, so a lane computes it once for all its vx.
__device__ __forceinline__ double half_y(const Geom &g, double y, double vy, double dt) {
  const double xt1 = y - 0.5 * dt * vy;
  return g.inv_dx[1] * (xt1 - g.minPhy[1]);
}

// Characteristic foot of cell (x, y, vx, vy), computeFeet of
// VMO/Advection.hpp:166-218: half-step position, 3x3 Lagrange interpolation
// of E with periodic wrap, full-step foot, velocity clamped to the box.
// xt1 = half_y(...) and ip1 = round(xt1) - 1 come from the caller.  Returns
// false when the foot leaves [min-dx, min+N dx] in any dimension (NaN counts
// as outside).
template <typename ES>
__device__ __forceinline__ bool compute_foot(const Geom &g, const FeetConsts &fc, const ES &E, double xt1, int ip1,
                                             double x, double y, double vx, double vy, double dt, double xs[4]) {
  const int N0 = g.n[0];
  double xt0 = x - 0.5 * dt * vx;
  xt0 = g.inv_dx[0] * (xt0 - g.minPhy[0]);
  const int ip0 = (int)(round(xt0) - 1);
  const double px = xt0 - ip0, py = xt1 - ip1;
  const double cx[3] = {.5 * (px - 1.) * (px - 2.), -1. * (px) * (px - 2.), .5 * (px) * (px - 1.)};
  const double cy[3] = {.5 * (py - 1.) * (py - 2.), -1. * (py) * (py - 2.), .5 * (py) * (py - 1.)};
  double ftmp1 = 0., ftmp2 = 0.;
  for (int k1 = 0; k1 <= 2; k1++) {
    for (int k0 = 0; k0 <= 2; k0++) {
      const int jx = ((ip0 + k0) % N0 + N0) % N0;  // (N + i) % N, non-negative for any i
      const double cc = cx[k0] * cy[k1];
      double e1, e2;
      E.at(jx, k1, e1, e2);
      ftmp1 += cc * e1;
      ftmp2 += cc * e2;
    }
  }
  const double xt2 = vx - 0.5 * dt * ftmp1;
  const double xt3 = vy - 0.5 * dt * ftmp2;
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
  bool ok = true;
  for (int j = 0; j < 4; j++) ok = ok && (xs[j] >= fc.locmin[j] && xs[j] <= fc.locmax[j]);
  return ok;
}

// Unaligned vector loads of stencil rows (multi-dword global loads need only
// dword alignment).
typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ void load_row4(const double *p, double r[4]) {
  const d2u a = *(const d2u *)p, b = *(const d2u *)(p + 2);
  r[0] = a.x;
  r[1] = a.y;
  r[2] = b.x;
  r[3] = b.y;
}
__device__ __forceinline__ void load_row4(const float *p, float r[4]) {
  const f4u a = *(const f4u *)p;
  r[0] = a.x;
  r[1] = a.y;
  r[2] = a.z;
  r[3] = a.w;
}
template <typename T> __device__ __forceinline__ void load_row5(const T *p, T r[5]) {
  load_row4(p, r);
  r[4] = p[4];
}

// Direct 4^4 gather of interp_4D on the C4 layout (positions -3..N+2, pitches
// c4_pitch).  A valid foot has ip in [-1, N], so positions ip-1..ip+2 stay
// inside -3..N+2; any reading of -3 or N+2 carries weight 0 (or ~1e-16 at the
// lower edge).  Along y, vx and vy the stored value there is 0; along x that
// slot overlaps the neighbouring row, whose finite value the weight cancels.
template <typename T, typename A>
__device__ __forceinline__ A gather4d(const T *__restrict__ c4, long S1, long S2, long S3, const int ip[4],
                                      const A eta[4][4]) {
  const T *base = c4 + (ip[0] + 2) + S1 * (ip[1] + 2) + S2 * (ip[2] + 2) + S3 * (long)(ip[3] + 2);
  A ftmp1 = 0;
  for (int k3 = 0; k3 < 4; k3++)
    for (int k2 = 0; k2 < 4; k2++) {
      A ftmp2 = 0;
      const T *p23 = base + k3 * S3 + k2 * S2;
      for (int k1 = 0; k1 < 4; k1++) {
        T r[4];
        load_row4(p23 + k1 * S1, r);
        A ftmp3 = 0;
        for (int k0 = 0; k0 < 4; k0++) ftmp3 += eta[0][k0] * (A)r[k0];
        ftmp2 += ftmp3 * eta[1][k1];
      }
      ftmp1 += ftmp2 * eta[2][k2] * eta[3][k3];
    }
  return ftmp1;
}


}  // namespace vlp
