// field_solve.hip -- the field of a caller-held distribution function, as one
// C-ABI call on caller buffers (no context): density, Poisson solve, field
// diagnostics and L2 norm.
//
// It strings together what the reference's step does after each advection
// (VMO/Timestep.hpp:53-67, 77-91): field_rho + field_reduce (VMO/Field.cpp:
// 5-69, one rank), field_poisson (VMO/Field.cpp:71-137, VMO/Efield.cpp:
// 3-106), Diags::compute and computeL2norm (VMO/Diags.cpp:16-76), with the
// library's own kernels: k_moments (density partials per (x, y, vy) and L2
// partials), then the field kernels of field.hip.  The scratch they need --
// density and L2 partials, spectra, twiddle tables, reduction partials -- is
// carved out of one device buffer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>

#include "vlp4d_internal.hpp"

namespace vlp {
namespace {

// exp(-2 pi i m / n) with exact values on the axes, interleaved complex (the
// same table the contexts build on the host).
__global__ __launch_bounds__(256) void k_twiddles(double *tw, int n) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double re, im;
  if (m == 0) { re = 1.0; im = 0.0; }
  else if (2 * m == n) { re = -1.0; im = 0.0; }
  else if (4 * m == n) { re = 0.0; im = -1.0; }
  else if (4 * m == 3 * n) { re = 0.0; im = 1.0; }
  else {
    const double ang = 2.0 * M_PI * (double)m / (double)n;
    re = cos(ang);
    im = -sin(ang);
  }
  tw[2 * m] = re;
  tw[2 * m + 1] = im;
}

struct Layout {
  size_t rho_part, l2_part, exy, spec1, spec2, twx, twy, diag_part, rho_tmp, total;  // offsets in doubles
};

// Offsets are rounded up to 32 doubles (256 bytes).
Layout layout_of(const Geom &g) {
  const size_t N0 = g.n[0], N1 = g.n[1], N3 = g.n[3];
  const size_t nxh = N0 / 2 + 1;
  auto up = [](size_t v) { return (v + 31) / 32 * 32; };
  Layout L{};
  size_t o = 0;
  L.rho_part = o; o += up(N0 * N1 * N3);
  L.l2_part = o; o += up((size_t)l2_partials(g));
  L.exy = o; o += up(2 * N0 * N1);
  L.spec1 = o; o += up(2 * nxh * N1);
  L.spec2 = o; o += up(6 * nxh * N1);
  L.twx = o; o += up(2 * N0);
  L.twy = o; o += up(2 * N1);
  L.diag_part = o; o += up(4 * N1);
  L.rho_tmp = o; o += N3 > 64 ? up((size_t)RHO_SLICES * N0 * N1) : 0;
  L.total = o;
  return L;
}

}  // namespace

// Checks the arguments and fills g; returns VLP_OK or a recorded error.
int field_solve_geom(const vlp_config *cfg_in, const void *f, int f_bits, double *phi, double *ex, double *ey,
                     double *diag, Geom &g, size_t &scratch_doubles) {
  if (!cfg_in || !f || !phi || !ex || !ey || !diag) return set_error(VLP_E_ARG, "null argument");
  if (f_bits != 64 && f_bits != 32) return set_error(VLP_E_ARG, "f_bits must be 64 or 32");
  vlp_config cfg = *cfg_in;
  const int rc = prepare_config(&cfg);
  if (rc) return rc;
  for (int i = 0; i < 4; i++) {
    g.n[i] = cfg.nxmax[i];
    g.minPhy[i] = cfg.minPhy[i];
    g.maxPhy[i] = cfg.maxPhy[i];
    g.dx[i] = cfg.dx[i];
    g.inv_dx[i] = 1. / cfg.dx[i];
  }
  g.idcase = cfg.idcase;
  scratch_doubles = layout_of(g).total;
  return VLP_OK;
}

// Enqueues the field solve of f on s using scratch (at least
// layout_of(g).total doubles).  diag receives nrj, nrjx, nrjy, mass, l2norm.
int field_solve_enqueue(const Geom &g, const void *f, int f_bits, double *phi, double *ex, double *ey,
                        double *diag, double *scratch, hipStream_t s) {
  const Layout L = layout_of(g);
  FieldBufs b{};
  b.rho = phi;
  b.ex = ex;
  b.ey = ey;
  b.exy = scratch + L.exy;
  b.spec1 = scratch + L.spec1;
  b.spec2 = scratch + L.spec2;
  b.twx = scratch + L.twx;
  b.twy = scratch + L.twy;
  b.diag_part = scratch + L.diag_part;
  b.rho_tmp = g.n[3] > 64 ? scratch + L.rho_tmp : nullptr;
  b.diag = diag;  // nbiter 0: the five values at index 0
  b.nbiter = 0;
  hipLaunchKernelGGL(k_twiddles, dim3((unsigned)((g.n[0] + 255) / 256)), dim3(256), 0, s, b.twx, g.n[0]);
  hipLaunchKernelGGL(k_twiddles, dim3((unsigned)((g.n[1] + 255) / 256)), dim3(256), 0, s, b.twy, g.n[1]);
  double *rho_part = scratch + L.rho_part, *l2_part = scratch + L.l2_part;
  if (f_bits == 64) launch_moments<double>(g, (const double *)f, rho_part, l2_part, s);
  else launch_moments<float>(g, (const float *)f, rho_part, l2_part, s);
  launch_field(g, b, rho_part, g.n[3], l2_part, l2_partials(g), 0, true, s, nullptr);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(VLP_E_HIP, hipGetErrorString(e));
  return VLP_OK;
}

}  // namespace vlp

extern "C" {

int vlp_field_solve(const vlp_config *cfg, const void *f, int f_bits, double *phi, double *ex, double *ey,
                    double *diag, void *stream) {
  vlp::Geom g{};
  size_t nd = 0;
  int rc = vlp::field_solve_geom(cfg, f, f_bits, phi, ex, ey, diag, g, nd);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  double *scratch = nullptr;
  if (hipMalloc((void **)&scratch, nd * sizeof(double)) != hipSuccess)
    return vlp::set_error(VLP_E_HIP, "hipMalloc failed (field-solve scratch)");
  rc = vlp::field_solve_enqueue(g, f, f_bits, phi, ex, ey, diag, scratch, s);
  const hipError_t e = hipStreamSynchronize(s);
  (void)hipFree(scratch);
  if (rc) return rc;
  if (e != hipSuccess) return vlp::set_error(VLP_E_HIP, hipGetErrorString(e));
  return VLP_OK;
}

}  // extern "C"
