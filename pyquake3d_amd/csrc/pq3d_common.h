// Shared host-side helpers of the pyquake3d_amd C-ABI library: error
// reporting (status codes + a thread-local last-error message) and launch
// checks. Every extern "C" entry point returns 0 on success or a negative
// PQ3D_E* code and sets the message returned by pq3d_last_error().
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#define PQ3D_OK 0
#define PQ3D_EINVAL -1
#define PQ3D_EHIP -2
#define PQ3D_ENOTSUP -3
#define PQ3D_ERANGE -4

namespace pq3d {

void set_error(const char* fmt, ...);
void clear_error();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Grid size along x with the 2^31-1 limit checked by the caller.
inline unsigned int blocks_for(int64_t n, int per) { return (unsigned int)((n + per - 1) / per); }

}  // namespace pq3d

#define PQ3D_HIP_CHECK(expr)                                                         \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess) {                                                          \
      pq3d::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                      __LINE__);                                                     \
      return PQ3D_EHIP;                                                              \
    }                                                                                \
  } while (0)

#define PQ3D_LAUNCH_CHECK()                                                          \
  do {                                                                               \
    hipError_t _e = hipGetLastError();                                               \
    if (_e != hipSuccess) {                                                          \
      pq3d::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e),      \
                      __FILE__, __LINE__);                                           \
      return PQ3D_EHIP;                                                              \
    }                                                                                \
  } while (0)

#define PQ3D_REQUIRE(cond, ...)                                                      \
  do {                                                                               \
    if (!(cond)) {                                                                   \
      pq3d::set_error(__VA_ARGS__);                                                  \
      return PQ3D_EINVAL;                                                            \
    }                                                                                \
  } while (0)
