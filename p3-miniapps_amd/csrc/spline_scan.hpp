// spline_scan.hpp -- chunked-scan form of the spline prefilter recursion,
// shared by the spline kernels (spline.hip) and the fused xy kernel
// (xy_fused.hip).
#pragma once

#include "device_common.hpp"

namespace vlp {

// ---------------------------------------------------------------------------
// Multi-lane scan recursion.  A group of P consecutive lanes holds one line of
// N = P*C points, lane p the chunk [p*C, (p+1)*C) in v[].  The causal and
// anti-causal recurrences run locally per chunk, the chunk carries are
// combined by a log2(P)-step scan with multiplier alpha^C, and a second local
// pass from the exact carry produces the coefficients (so each coefficient
// sees the reference's sequence of multiply-adds from its carry on).
// C >= 10 keeps f[0..9] in lane 0 and f[N-10..N-1] in lane P-1 (the support
// of the two alpha-sums).  On return v[] holds c[p*C .. p*C+C-1]; lane 0 also
// gets c[-2], c[-1] and lane P-1 c[N], c[N+1].  Every lane of the group must
// call this together (cross-lane shuffles).
// ---------------------------------------------------------------------------
template <int C>
__device__ __forceinline__ void scan_recursion(double (&v)[C], int p, int P, const SplineConsts &sc, double aC,
                                               double aN3, double &cm2, double &cm1, double &cN, double &cN1) {
  const double a = sc.alpha, b = sc.beta, s3 = sc.sqrt3;
  double gp2 = 0., gm3 = 0., ap = a;
#pragma unroll
  for (int j = 1; j <= 8; j++) {
    gp2 += v[1 + j] * ap;      // meaningful in lane 0
    gm3 += v[C - 2 - j] * ap;  // meaningful in lane P-1
    ap *= a;
  }
  const double f0 = __shfl(v[0], 0, P), f1 = __shfl(v[1], 0, P);
  gp2 = __shfl(gp2, 0, P);
  const double fm1 = __shfl(v[C - 1], P - 1, P), fm2 = __shfl(v[C - 2], P - 1, P);
  gm3 = __shfl(gm3, P - 1, P);
  // causal
  const double cpm2 = gm3 + fm2;
  const double cpm1 = fm1 + a * cpm2;
  double cp = (p == 0) ? cpm1 : 0.;
#pragma unroll
  for (int i = 0; i < C; i++) cp = v[i] + a * cp;
  double last = cp, mult = aC;
  for (int d = 1; d < P; d <<= 1) {
    const double o = __shfl_up(last, d, P);
    if (p >= d) last += mult * o;
    mult *= mult;
  }
  double cin = __shfl_up(last, 1, P);
  if (p == 0) cin = cpm1;
  cp = cin;
#pragma unroll
  for (int i = 0; i < C; i++) {
    cp = v[i] + a * cp;
    v[i] = cp;
  }
  // lane P-1 holds cp[N-1]: positions N, N+1 and the anti-causal start, using
  // sum_{n=-2}^{N+1} alpha^(N+1-n) g[n] = cp[N+1] - alpha^(N+3) g[-3].
  const double cpN = f0 + a * cp;
  const double cpN1 = f1 + a * cpN;
  const double fnend = gp2 + (cpN1 - aN3 * gm3);
  cN1 = s3 * fnend;
  cN = b * cpN + a * cN1;
  double cr = (p == P - 1) ? cN : 0.;
#pragma unroll
  for (int i = C - 1; i >= 0; i--) cr = b * v[i] + a * cr;
  double first = cr;
  mult = aC;
  for (int d = 1; d < P; d <<= 1) {
    const double o = __shfl_down(first, d, P);
    if (p + d < P) first += mult * o;
    mult *= mult;
  }
  double cin2 = __shfl_down(first, 1, P);
  if (p == P - 1) cin2 = cN;
  cr = cin2;
#pragma unroll
  for (int i = C - 1; i >= 0; i--) {
    cr = b * v[i] + a * cr;
    v[i] = cr;
  }
  cm1 = b * cpm1 + a * cr;  // meaningful in lane 0 (cr = c[0])
  cm2 = b * cpm2 + a * cm1;
}

}  // namespace vlp
