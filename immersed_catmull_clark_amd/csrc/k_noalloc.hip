// Fixed device workspace of icc_pcg_dev_noalloc: one file-scope array per
// GPU, sized for NA_MAX_NDOF unknowns, so that the no-allocation PCG entry
// point never calls hipMalloc.  Every process that loads the library holds
// the array on each GPU it runs on until it exits.
//
// In host code `scratch` names a host-side stand-in, not a device address;
// its device address is looked up once per device (hipGetSymbolAddress, with
// that device current) and only that address is handed to kernels and copies.
#include <hip/hip_runtime.h>

#include <mutex>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

namespace icc {

static __device__ __align__(256) double scratch[NA_SCRATCH_DOUBLES];

namespace {
constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mu;
double* na_ptr[NA_MAX_DEVICES] = {};
hipError_t na_err[NA_MAX_DEVICES] = {};
bool na_done[NA_MAX_DEVICES] = {};
}  // namespace

// Looks up this device's copy of the workspace (the device must be current);
// repeated calls return the address found the first time.
hipError_t noalloc_init(int device) {
  if (device < 0 || device >= NA_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(na_mu);
  if (!na_done[device]) {
    void* p = nullptr;
    na_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    na_ptr[device] = na_err[device] == hipSuccess ? static_cast<double*>(p) : nullptr;
    na_done[device] = true;
  }
  return na_err[device];
}

double* noalloc_scratch(int device) {
  if (device < 0 || device >= NA_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(na_mu);
  return na_done[device] ? na_ptr[device] : nullptr;
}

}  // namespace icc
