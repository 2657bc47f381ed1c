// HIP error reporting for the device-side translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "common.hpp"

namespace snc_amd {

// Reads (and so clears) HIP's last error: SNC_OK when there is none, else a
// failure naming it.
inline int hip_last(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return SNC_OK;
  return fail(SNC_E_HIP, std::string(what) + ": " + hipGetErrorName(e));
}

// Device bytes allocated while an AllocAccount is alive on this thread are
// added to its counter (graph upload, plan build, lazily built buffers).
inline thread_local int64_t* g_alloc_account = nullptr;
struct AllocAccount {
  int64_t* prev;
  explicit AllocAccount(int64_t* counter) : prev(g_alloc_account) { g_alloc_account = counter; }
  ~AllocAccount() { g_alloc_account = prev; }
};
inline void account_alloc(size_t bytes) {
  if (g_alloc_account) *g_alloc_account += (int64_t)bytes;
}

// Drops an error left by an earlier call outside this library (or by a call
// whose failure was already reported), so the launch checks of the entry point
// that follows report only its own launches.
inline void hip_clear() { (void)hipGetLastError(); }

// A failed HIP call: clears the sticky last error, so later calls do not
// report it again, and names the error.
inline int hip_fail(hipError_t e, const char* what) {
  (void)hipGetLastError();
  return fail(SNC_E_HIP, std::string(what) + ": " + hipGetErrorName(e));
}

}  // namespace snc_amd
