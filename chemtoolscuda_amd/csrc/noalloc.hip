// Scratch of the allocation-free entry point (ct_eval_noalloc): one
// file-scope device array per GPU, of CT_NOALLOC_SCRATCH_BYTES, that each GPU
// the library runs on holds until the process ends.  In host code the array's
// name is only a host-side stand-in; its device address is looked up once per
// device (with that device current) and kept, and only that address is ever
// handed to kernels, copies or fills.  Every call on one GPU shares that GPU's
// array: calls that overlap on one GPU race.

#include <mutex>

#include "mol.h"

namespace ct {

namespace {

static __device__ __align__(256) unsigned char scratch[CT_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

void* noalloc_scratch(int device) {
  CT_REQUIRE(device >= 0 && device < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device] == nullptr) {
    int current = 0;
    CT_HIP_CHECK(hipGetDevice(&current));
    if (current != device) CT_HIP_CHECK(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (current != device) CT_HIP_CHECK(hipSetDevice(current));
    if (e != hipSuccess || ptr == nullptr)
      throw Error(CT_E_HIP, std::string("cannot locate the allocation-free scratch: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

}  // namespace ct
