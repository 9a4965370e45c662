// Elementwise passes of the pixel-based ILT gradient step, fused:
//   oi_ilt_residual : d = 2 (I - T) and the loss sum (I - T)^2
//   oi_ilt_update   : g = dL/dm m (1 - m); theta -= lr g / rms(g) per image;
//                     optionally m = sigmoid(theta) for the next step
// Every reduction runs over a fixed split (kIltBlocks workgroups per image,
// then the partial sums in a fixed order), so results do not depend on the
// batch or on timing.  These passes stream 4-byte values: HBM-bound, float4
// loads and stores when the pointers and the image size allow.
#include "oi_common.h"
#include "oi_fastpath.h"
#include "opticalimaging_amd.h"

namespace oi {
namespace {

constexpr int kIltBlocks = 256;   // workgroups per image (enough loads in flight to stream)
constexpr int kIltThreads = 256;

typedef float f4v __attribute__((ext_vector_type(4)));

// Sum over the workgroup (4 waves of 64); the result is valid in thread 0.
__device__ __forceinline__ float block_sum(float v, float* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  float s = 0.f;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kIltThreads / 64; ++w) s += red[w];
  }
  return s;
}

// d = 2 (img - tgt); part[b][blk] = sum over the block's pixels of (img - tgt)^2
template <bool VEC>
__global__ __launch_bounds__(kIltThreads) void k_ilt_residual(const float* img, const float* tgt,
                                                             float* d, size_t npix,
                                                             float* __restrict__ part) {
  __shared__ float red[kIltThreads / 64];
  const size_t base = (size_t)blockIdx.y * npix;
  const size_t stride = (size_t)kIltBlocks * kIltThreads;
  float acc = 0.f;
  if (VEC) {
    const size_t n4 = npix / 4;
    const f4v* a4 = reinterpret_cast<const f4v*>(img + base);
    const f4v* t4 = reinterpret_cast<const f4v*>(tgt + base);
    f4v* d4 = reinterpret_cast<f4v*>(d + base);
#pragma unroll 4   // several loads in flight ahead of the running sum
    for (size_t i = blockIdx.x * (size_t)kIltThreads + threadIdx.x; i < n4; i += stride) {
      const f4v r = a4[i] - t4[i];
      acc += r.x * r.x + r.y * r.y + r.z * r.z + r.w * r.w;
      d4[i] = r * 2.f;
    }
  } else {
    for (size_t i = blockIdx.x * (size_t)kIltThreads + threadIdx.x; i < npix; i += stride) {
      const float r = img[base + i] - tgt[base + i];
      acc += r * r;
      d[base + i] = 2.f * r;
    }
  }
  const float s = block_sum(acc, red);
  if (threadIdx.x == 0) part[(size_t)blockIdx.y * kIltBlocks + blockIdx.x] = s;
}

// Sum over one wave of 64 lanes, in double, in a fixed pattern.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// out[0] = sum of n partial sums (one wave: lane l adds l, l + 64, ... in
// order, then the lanes combine in a fixed pattern)
__global__ __launch_bounds__(64) void k_ilt_total(const float* __restrict__ part, int n,
                                                  float* __restrict__ out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 64) s += part[i];
  s = wave_sum(s);
  if (threadIdx.x == 0) out[0] = (float)s;
}

__device__ __forceinline__ float chain(float gm, float m) { return gm * m * (1.f - m); }

// part[b][blk] = sum of (dL/dm m (1 - m))^2 over the block's pixels
template <bool VEC>
__global__ __launch_bounds__(kIltThreads) void k_ilt_gnorm(const float* __restrict__ g,
                                                          const float* __restrict__ m,
                                                          size_t npix, float* __restrict__ part) {
  __shared__ float red[kIltThreads / 64];
  const size_t base = (size_t)blockIdx.y * npix;
  const size_t stride = (size_t)kIltBlocks * kIltThreads;
  float acc = 0.f;
  if (VEC) {
    const size_t n4 = npix / 4;
    const f4v* g4 = reinterpret_cast<const f4v*>(g + base);
    const f4v* m4 = reinterpret_cast<const f4v*>(m + base);
#pragma unroll 4   // several loads in flight ahead of the running sum
    for (size_t i = blockIdx.x * (size_t)kIltThreads + threadIdx.x; i < n4; i += stride) {
      const f4v gv = g4[i], mv = m4[i];
      const float a = chain(gv.x, mv.x), b = chain(gv.y, mv.y);
      const float c = chain(gv.z, mv.z), e = chain(gv.w, mv.w);
      acc += a * a + b * b + c * c + e * e;
    }
  } else {
    for (size_t i = blockIdx.x * (size_t)kIltThreads + threadIdx.x; i < npix; i += stride) {
      const float a = chain(g[base + i], m[base + i]);
      acc += a * a;
    }
  }
  const float s = block_sum(acc, red);
  if (threadIdx.x == 0) part[(size_t)blockIdx.y * kIltBlocks + blockIdx.x] = s;
}

__device__ __forceinline__ float sigmoid(float x) { return 1.f / (1.f + expf(-x)); }

// theta -= (lr / rms) dL/dm m (1 - m), rms = sqrt(sum / npix) + 1e-30 of the
// image's partial sums (each workgroup adds them in the same order);
// m_next (may be m itself, or null) = sigmoid(theta) after the step
template <bool VEC>
__global__ __launch_bounds__(kIltThreads) void k_ilt_update(float* theta, const float* g,
                                                           const float* m, float lr, size_t npix,
                                                           const float* __restrict__ part,
                                                           float* m_next) {
  static_assert(kIltBlocks % 64 == 0, "whole partial sums per lane of the first wave");
  __shared__ float coef;
  if (threadIdx.x < 64) {
    double s = 0.0;                             // lane l adds l, l + 64, ... in order
    for (int i = threadIdx.x; i < kIltBlocks; i += 64) s += part[(size_t)blockIdx.y * kIltBlocks + i];
    s = wave_sum(s);
    const double rms = sqrt(s) / sqrt((double)npix) + 1e-30;
    if (threadIdx.x == 0) coef = (float)(lr / rms);
  }
  __syncthreads();
  const float c = coef;
  const size_t base = (size_t)blockIdx.y * npix;
  const size_t stride = (size_t)kIltBlocks * kIltThreads;
  if (VEC) {
    const size_t n4 = npix / 4;
    f4v* t4 = reinterpret_cast<f4v*>(theta + base);
    const f4v* g4 = reinterpret_cast<const f4v*>(g + base);
    const f4v* m4 = reinterpret_cast<const f4v*>(m + base);
    f4v* n4p = m_next ? reinterpret_cast<f4v*>(m_next + base) : nullptr;
    for (size_t i = blockIdx.x * (size_t)kIltThreads + threadIdx.x; i < n4; i += stride) {
      const f4v gv = g4[i], mv = m4[i];
      f4v t = t4[i];
      t.x -= c * chain(gv.x, mv.x);
      t.y -= c * chain(gv.y, mv.y);
      t.z -= c * chain(gv.z, mv.z);
      t.w -= c * chain(gv.w, mv.w);
      t4[i] = t;
      if (n4p) n4p[i] = (f4v){sigmoid(t.x), sigmoid(t.y), sigmoid(t.z), sigmoid(t.w)};
    }
  } else {
    for (size_t i = blockIdx.x * (size_t)kIltThreads + threadIdx.x; i < npix; i += stride) {
      const float t = theta[base + i] - c * chain(g[base + i], m[base + i]);
      theta[base + i] = t;
      if (m_next) m_next[base + i] = sigmoid(t);
    }
  }
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

}  // namespace
}  // namespace oi

namespace oi {
int ilt_sum(const float* part, int n, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_ilt_total, dim3(1), dim3(64), 0, st, part, n, out);
  OI_LAUNCH_CHECK("ilt_total");
  return OI_OK;
}
}  // namespace oi

using namespace oi;

extern "C" size_t oi_ilt_work_floats(int batch) {
  return batch > 0 ? (size_t)batch * kIltBlocks : 0;
}

extern "C" int oi_ilt_residual(const float* image, const float* target, float* resid2, float* loss,
                               int batch, size_t npix, float* work, void* stream) {
  if (batch < 0) return fail(OI_ERR_INVALID, "ilt_residual: negative batch");
  hipStream_t st = (hipStream_t)stream;
  if (batch == 0 || npix == 0) {      // nothing to compare: the loss is 0
    if (loss) {
      hipLaunchKernelGGL(k_ilt_total, dim3(1), dim3(64), 0, st, (const float*)nullptr, 0, loss);
      OI_LAUNCH_CHECK("ilt_total");
    }
    return OI_OK;
  }
  if (!image || !target || !resid2 || !work)
    return fail(OI_ERR_INVALID, "ilt_residual: null pointer");
  if (batch > 65535) return fail(OI_ERR_INVALID, "ilt_residual: batch over 65535");
  const bool vec = npix % 4 == 0 && aligned16(image) && aligned16(target) && aligned16(resid2);
  const dim3 grid(kIltBlocks, batch);
  if (vec)
    hipLaunchKernelGGL(k_ilt_residual<true>, grid, dim3(kIltThreads), 0, st, image, target, resid2,
                       npix, work);
  else
    hipLaunchKernelGGL(k_ilt_residual<false>, grid, dim3(kIltThreads), 0, st, image, target, resid2,
                       npix, work);
  OI_LAUNCH_CHECK("ilt_residual");
  if (loss) {
    hipLaunchKernelGGL(k_ilt_total, dim3(1), dim3(64), 0, st, work, batch * kIltBlocks, loss);
    OI_LAUNCH_CHECK("ilt_total");
  }
  return OI_OK;
}

extern "C" int oi_ilt_update(float* theta, const float* grad_m, const float* m, float lr,
                             int batch, size_t npix, float* work, float* m_next, void* stream) {
  if (batch < 0) return fail(OI_ERR_INVALID, "ilt_update: negative batch");
  if (batch == 0 || npix == 0) return OI_OK;
  if (!theta || !grad_m || !m || !work) return fail(OI_ERR_INVALID, "ilt_update: null pointer");
  if (batch > 65535) return fail(OI_ERR_INVALID, "ilt_update: batch over 65535");
  hipStream_t st = (hipStream_t)stream;
  const bool vec = npix % 4 == 0 && aligned16(theta) && aligned16(grad_m) && aligned16(m) &&
                   (!m_next || aligned16(m_next));
  const dim3 grid(kIltBlocks, batch);
  if (vec)
    hipLaunchKernelGGL(k_ilt_gnorm<true>, grid, dim3(kIltThreads), 0, st, grad_m, m, npix, work);
  else
    hipLaunchKernelGGL(k_ilt_gnorm<false>, grid, dim3(kIltThreads), 0, st, grad_m, m, npix, work);
  OI_LAUNCH_CHECK("ilt_gnorm");
  if (vec)
    hipLaunchKernelGGL(k_ilt_update<true>, grid, dim3(kIltThreads), 0, st, theta, grad_m, m, lr,
                       npix, work, m_next);
  else
    hipLaunchKernelGGL(k_ilt_update<false>, grid, dim3(kIltThreads), 0, st, theta, grad_m, m, lr,
                       npix, work, m_next);
  OI_LAUNCH_CHECK("ilt_update");
  return OI_OK;
}
