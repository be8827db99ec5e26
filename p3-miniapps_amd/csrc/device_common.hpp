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

// s_waitcnt vmcnt(n) for a wave-uniform n (the count is an immediate of the
// instruction); n above 63 waits for all.
#define VLP_VMWAIT(N) \
  case N: __builtin_amdgcn_s_waitcnt(0x0F70 | ((N) & 15) | (((N) >> 4) << 14)); break;
__device__ __forceinline__ void wait_vm(int n) {
  switch (n) {
    VLP_VMWAIT(1) VLP_VMWAIT(2) VLP_VMWAIT(3) VLP_VMWAIT(4) VLP_VMWAIT(5) VLP_VMWAIT(6) VLP_VMWAIT(7)
    VLP_VMWAIT(8) VLP_VMWAIT(9) VLP_VMWAIT(10) VLP_VMWAIT(11) VLP_VMWAIT(12) VLP_VMWAIT(13) VLP_VMWAIT(14)
    VLP_VMWAIT(15) VLP_VMWAIT(16) VLP_VMWAIT(17) VLP_VMWAIT(18) VLP_VMWAIT(19) VLP_VMWAIT(20) VLP_VMWAIT(21)
    VLP_VMWAIT(22) VLP_VMWAIT(23) VLP_VMWAIT(24) VLP_VMWAIT(25) VLP_VMWAIT(26) VLP_VMWAIT(27) VLP_VMWAIT(28)
    VLP_VMWAIT(29) VLP_VMWAIT(30) VLP_VMWAIT(31) VLP_VMWAIT(32) VLP_VMWAIT(33) VLP_VMWAIT(34) VLP_VMWAIT(35)
    VLP_VMWAIT(36) VLP_VMWAIT(37) VLP_VMWAIT(38) VLP_VMWAIT(39) VLP_VMWAIT(40) VLP_VMWAIT(41) VLP_VMWAIT(42)
    VLP_VMWAIT(43) VLP_VMWAIT(44) VLP_VMWAIT(45) VLP_VMWAIT(46) VLP_VMWAIT(47) VLP_VMWAIT(48) VLP_VMWAIT(49)
    VLP_VMWAIT(50) VLP_VMWAIT(51) VLP_VMWAIT(52) VLP_VMWAIT(53) VLP_VMWAIT(54) VLP_VMWAIT(55) VLP_VMWAIT(56)
    VLP_VMWAIT(57) VLP_VMWAIT(58) VLP_VMWAIT(59) VLP_VMWAIT(60) VLP_VMWAIT(61) VLP_VMWAIT(62) VLP_VMWAIT(63)
    default: __builtin_amdgcn_s_waitcnt(0x0F70); break;
  }
}
#undef VLP_VMWAIT

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
