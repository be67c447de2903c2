// tomo_noalloc.hip -- the non-uniform weighted SIRT update on caller-owned
// device arrays with its scratch in a fixed static device array: an entry
// point that allocates nothing, for callers that cannot allocate (code under
// HIP graph capture, tight serving loops).  It runs the kernels of
// tomo_update_weighted (launch_weighted), in the same order, so both give
// the same bits on the same inputs.
//
// Scratch layout (doubles): mean image [n2], partial squared distances
// [nshots][parts], weights [nshots] and their sum [1].  Every slot is written
// before it is read, so nothing is cleared between calls.
#include <hip/hip_runtime.h>

#include <climits>
#include <mutex>
#include <string>

#include "tomo.h"
#include "tomo_host.h"
#include "tomo_kernels.h"

using namespace tomo;
using tomo_host::set_create_error;

// One copy per GPU, held until the process ends.  Host code never passes
// this name to a kernel: it takes the device address from
// hipGetSymbolAddress (below) once per device.
static __device__ __align__(256) double scratch[TOMO_NOALLOC_SCRATCH_DOUBLES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

int fail(int status, const std::string &msg) {
  set_create_error("tomo_update_weighted_noalloc: " + msg);
  return status;
}

// Device address of this GPU's scratch array, looked up on the first call on
// the device (with that device current) and kept.
int scratch_of_current_device(double **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(TOMO_ERR_NODEV, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(TOMO_ERR_NODEV, "device index out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(TOMO_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return TOMO_OK;
}

}  // namespace

extern "C" int tomo_update_weighted_noalloc(float *d_f, const float *d_grad, int nshots, int nx, int ny,
                                            const int *mask4, double scale, const float *d_pixw,
                                            double *d_weights, void *stream) {
  if (!d_f || !d_grad || !mask4) return fail(TOMO_ERR_ARG, "null field, gradient or mask pointer");
  if (nshots < 1 || nshots > 65535 || nx < 1 || ny < 1) return fail(TOMO_ERR_ARG, "bad shot count or grid size");
  const long long n2 = (long long)nx * ny;
  if (n2 > INT_MAX / 2 || n2 * nshots > (long long)INT_MAX * 16) return fail(TOMO_ERR_ARG, "grid too large");
  const Mask m{mask4[0], mask4[1], mask4[2], mask4[3]};
  if (m.x0 < 0 || m.x1 > nx || m.x0 > m.x1 || m.y0 < 0 || m.y1 > ny || m.y0 > m.y1)
    return fail(TOMO_ERR_ARG, "mask outside the grid");
  const int parts = weighted_part_blocks((int)n2);
  const long long need = n2 + (long long)nshots * parts + nshots + 1;
  if (need > TOMO_NOALLOC_SCRATCH_DOUBLES)
    return fail(TOMO_ERR_ARG, "needs " + std::to_string(need) + " doubles of scratch (nx * ny + nshots * " +
                                  std::to_string(parts) + " + nshots + 1), above the fixed " +
                                  std::to_string((long long)TOMO_NOALLOC_SCRATCH_DOUBLES));
  double *base = nullptr;
  int rc = scratch_of_current_device(&base);
  if (rc) return rc;
  double *mean = base, *part = mean + n2, *w = part + (long long)nshots * parts;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e = launch_weighted(d_f, d_grad, nshots, nx, ny, m, scale, mean, part, parts, w, st, d_pixw);
  if (e == hipSuccess && d_weights)
    e = hipMemcpyAsync(d_weights, w, sizeof(double) * nshots, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return fail(TOMO_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
  return TOMO_OK;
}
