// Library-wide C ABI: version, error reporting, device handles.
#include <hip/hip_runtime.h>

#include <string>

#include "pf_internal.h"

namespace pf {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }
}  // namespace pf

using namespace pf;

extern "C" int pf_version(void) { return 100; }

extern "C" const char* pf_last_error(void) { return g_last_error.c_str(); }

extern "C" int pf_create(pf_handle** out, int device) {
  return guarded([&] {
    PF_REQUIRE(out, "null output");
    int n = 0;
    PF_HIP(hipGetDeviceCount(&n));
    PF_REQUIRE(device >= 0 && device < n, "device index out of range");
    PF_HIP(hipSetDevice(device));
    pf_handle* h = new pf_handle();
    h->device = device;
    if (hipGetDeviceProperties(&h->prop, device) != hipSuccess) {
      delete h;
      throw Error{PF_ERR_HIP, "hipGetDeviceProperties failed"};
    }
    try {
      noalloc_scratch_init(device);  // pf_filter_noalloc's per-device work array
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

extern "C" int pf_destroy(pf_handle* h) {
  if (!h) return PF_OK;
  h->partials.release();
  h->small.release();
  h->vecs.release();
  delete h;
  return PF_OK;
}
