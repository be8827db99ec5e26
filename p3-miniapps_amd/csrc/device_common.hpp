// device_common.hpp -- small device helpers shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "vlp4d_internal.hpp"

namespace vlp {

// Constants of the spline prefilter, VMO/Spline.hpp:21,122-123:
// sqrt3 = sqrt(3), alpha = sqrt3 - 2, beta = sqrt3 * (1 - alpha^2).
struct SplineConsts {
  double alpha, beta, sqrt3;
};

SplineConsts spline_consts();

// B-spline weights, VMO/Advection.hpp:131-143 / 266-274.
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

// Sum over a workgroup whose size is a multiple of 64 and at most 512;
// result valid in thread 0.  Fixed reduction order (deterministic).
__device__ __forceinline__ double block_sum(double v, double *sh) {
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

}  // namespace vlp
