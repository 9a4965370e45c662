// Shared device helpers and host error plumbing for the aerial-image library.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>

#include "opticalimaging_amd.h"

namespace oi {

// ---------------------------------------------------------------- complex --
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float cnorm(float2 a) { return a.x * a.x + a.y * a.y; }
// -i * a and +i * a
__device__ __forceinline__ float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }
__device__ __forceinline__ float2 cmul_pi(float2 a) { return make_float2(-a.y, a.x); }

// Non-negative remainder.
__host__ __device__ __forceinline__ int pmod(int a, int m) {
  int r = a % m;
  return r < 0 ? r + m : r;
}

// ------------------------------------------------------------ XCD order --
// Workgroups of a 1-D grid are dispatched round-robin over the 8 XCDs
// (workgroup p on XCD p mod 8), each with its own L2.  xcd_item gives
// workgroup p the work item w such that every XCD takes one contiguous run of
// the items, in order: neighbouring items, whose edges share cache lines, run
// on one L2 at about the same time instead of each fetching those lines from
// HBM.  The grid holds xcd_grid(total) workgroups; xcd_item is false for the
// few past the last item.
constexpr int kXcdCount = 8;
__host__ __device__ __forceinline__ int xcd_grid(int total) {
  return kXcdCount * ((total + kXcdCount - 1) / kXcdCount);
}
__device__ __forceinline__ bool xcd_item(int p, int total, int& w) {
  const int per = (total + kXcdCount - 1) / kXcdCount;
  w = (p % kXcdCount) * per + p / kXcdCount;
  return w < total;
}

// ------------------------------------------------------------ host errors --
void set_error(int code, const std::string& msg);
int fail(int code, const std::string& msg);

#define OI_HIP_CHECK(expr)                                                        \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess)                                                         \
      return ::oi::fail(OI_ERR_RUNTIME, std::string(#expr) + ": " +               \
                                            hipGetErrorString(_e));               \
  } while (0)

// Checks the launch status of the kernel just enqueued.
#define OI_LAUNCH_CHECK(name)                                                     \
  do {                                                                            \
    hipError_t _e = hipGetLastError();                                            \
    if (_e != hipSuccess)                                                         \
      return ::oi::fail(OI_ERR_RUNTIME, std::string("launch ") + (name) + ": " +  \
                                            hipGetErrorString(_e));               \
  } while (0)

}  // namespace oi
