// field_solve_noalloc.hip -- vlp_field_solve without allocation: the same
// kernels on scratch held in a file-scope device array.
//
// For callers that cannot allocate device memory between set-up and
// tear-down (code under HIP graph capture, tight loops around a solver that
// owns its buffers): the density/L2 partials, spectra, twiddle tables and
// reduction partials of the field solve live in `scratch`, a static
// device array of VLP_FIELD_SCRATCH_DOUBLES doubles (24 MiB) that every GPU
// the library runs on holds until the process ends.  Nothing in it has to
// start at zero (every kernel writes before it reads), and it keeps the last
// call's contents.  Calls on one GPU share it: calls that overlap in time on
// the same GPU race (enqueue them on one stream, or order them).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "vlp4d_internal.hpp"

namespace {

constexpr size_t kScratchDoubles = VLP_FIELD_SCRATCH_DOUBLES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};  // device address of `scratch` on each GPU

// Device address of this GPU's `scratch`, looked up once per device with the
// device current (the array's name is a host-side stand-in, never a device
// address).  Null (error recorded) on failure.
double *scratch_on(int dev) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      vlp::set_error(VLP_E_HIP, (std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)).c_str());
      return nullptr;
    }
    g_ptr[dev] = (double *)p;
  }
  return g_ptr[dev];
}

}  // namespace

extern "C" {

int vlp_field_solve_noalloc(const vlp_config *cfg, const void *f, int f_bits, double *phi, double *ex, double *ey,
                            double *diag, void *stream) {
  vlp::Geom g{};
  size_t nd = 0;
  int rc = vlp::field_solve_geom(cfg, f, f_bits, phi, ex, ey, diag, g, nd);
  if (rc) return rc;
  if (nd > kScratchDoubles)
    return vlp::set_error(VLP_E_SIZE, "grid too large for vlp_field_solve_noalloc's scratch "
                                      "(VLP_FIELD_SCRATCH_DOUBLES; about Nx*Ny*Nvy <= 2.6e6 doubles)");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return vlp::set_error(VLP_E_HIP, "no current HIP device");
  double *ws = scratch_on(dev);
  if (!ws) return VLP_E_HIP;
  return vlp::field_solve_enqueue(g, f, f_bits, phi, ex, ey, diag, ws, (hipStream_t)stream);
}

}  // extern "C"
