// Host-side helpers shared by the translation units of libmonc_tvd: the
// thread-local error report behind monc_tvd_last_error and the layout of the
// FP64 metric vectors on the device.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/monc_tvd.h"
#include "tvd_faces.hpp"

namespace tvd {

// Records the message of a failure on the calling thread and returns `code`.
__attribute__((visibility("hidden"))) int fail(int code, const char* fmt, ...)
    __attribute__((format(printf, 2, 3)));

// Every device allocation of the library goes through dev_alloc, so that the
// test hook monc_tvd_test_fail_allocations can make the next n of them report
// hipErrorOutOfMemory without reaching the runtime (error paths are tested
// without exhausting a GPU).
__attribute__((visibility("hidden"))) extern int g_fail_allocs;
template <typename P>
inline hipError_t dev_alloc(P** p, size_t bytes) {
  if (__atomic_load_n(&g_fail_allocs, __ATOMIC_RELAXED) > 0 &&
      __atomic_sub_fetch(&g_fail_allocs, 1, __ATOMIC_RELAXED) >= 0) {
    *p = nullptr;
    return hipErrorOutOfMemory;
  }
  return hipMalloc((void**)p, bytes);
}

#define HIP_TRY(call)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess) return ::tvd::fail(MONC_TVD_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// Metric vectors of one domain, each of length nz, in this order in one
// device array.
enum { MET_RDZ, MET_RDZN, MET_DZ, MET_DZN, MET_TZC1, MET_TZC2, MET_TZD1, MET_TZD2, MET_N };

// The metric roles ultflx and differentiate_face_values see: for w the
// reference passes (rdzn, rdz, dz) as (rdz, rdzn, dzn) and the tzd pair
// (tvdadvection.F90:881-894, :898-904); otherwise the model's own vectors and
// the tzc pair (:739-757).
__host__ __device__ inline VMet vmet_from(const double* m, int nz, bool is_w) {
  VMet v;
  if (is_w) {
    v.rdz = m + MET_RDZN * nz;
    v.rdzn = m + MET_RDZ * nz;
    v.dzn = m + MET_DZ * nz;
    v.t1 = m + MET_TZD1 * nz;
    v.t2 = m + MET_TZD2 * nz;
  } else {
    v.rdz = m + MET_RDZ * nz;
    v.rdzn = m + MET_RDZN * nz;
    v.dzn = m + MET_DZN * nz;
    v.t1 = m + MET_TZC1 * nz;
    v.t2 = m + MET_TZC2 * nz;
  }
  return v;
}

}  // namespace tvd
