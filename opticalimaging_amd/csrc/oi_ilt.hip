// Elementwise passes of the pixel-based ILT gradient step, fused:
//   oi_ilt_residual : d = 2 (I - T) and the loss sum (I - T)^2
//   oi_ilt_update   : g = dL/dm m (1 - m); theta -= lr g / rms(g) per image;
//                     optionally m = sigmoid(theta) for the next step
// and the band form of the residual (the cotangent spectrum and the loss from
// the image's band, for oi_socs_band_residual_adjoint).
// Every reduction runs over a fixed split (kIltBlocks workgroups per image,
// then the partial sums in a fixed order), so results do not depend on the
// batch or on timing.  The elementwise passes stream 4-byte values:
// HBM-bound, float4 loads and stores when the pointers and the image size allow.
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

// the sigmoid the spectrum and gradient passes apply to logits (logistic)
__device__ __forceinline__ float sigmoid(float x) { return logistic(x); }

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

// theta -= (lr / rms) gt per image, rms = sqrt(sum / npix) + 1e-30 of the
// image's nparts partial sums of gt^2 (each workgroup adds them in the same
// order); gt is already the gradient with respect to theta.
template <bool VEC>
__global__ __launch_bounds__(kIltThreads) void k_ilt_update_theta(float* theta,
                                                                 const float* __restrict__ gt,
                                                                 float lr, size_t npix,
                                                                 const float* __restrict__ part,
                                                                 int nparts) {
  __shared__ float coef;
  if (threadIdx.x < 64) {
    double s = 0.0;                             // lane l adds l, l + 64, ... in order
    for (int i = threadIdx.x; i < nparts; i += 64) s += part[(size_t)blockIdx.y * nparts + i];
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
    const f4v* g4 = reinterpret_cast<const f4v*>(gt + base);
    for (size_t i = blockIdx.x * (size_t)kIltThreads + threadIdx.x; i < n4; i += stride)
      t4[i] = t4[i] - c * g4[i];
  } else {
    for (size_t i = blockIdx.x * (size_t)kIltThreads + threadIdx.x; i < npix; i += stride)
      theta[base + i] -= c * gt[base + i];
  }
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// Band form of the least-squares loss L = sum_r (I - T)^2 for band-limited
// images I(r) = sum_k S(k) e^{-2 pi i k.r / n} (n = ny nx, k in the band
// [-qr, qr] x [-qc, qc]) and T~(k) = (1/n) sum_r T(r) e^{+2 pi i k.r / n}:
//   (1/n) sum_r 2 (I - T)(r) e^{+2 pi i k.r / n} = 2 (S(k) - T~(k))  (k in the band)
//   L = n sum_band |S - T~|^2 + (sum_r T^2 - n sum_band |T~|^2)     (Parseval)
// S comes from the band Ihat[q][qr + qrm] the image synthesis reads (q >= 0;
// q < 0 is the conjugate of (-qr, -q)); the q = 0 row enters through its
// Hermitian part, which is the part a real image carries.
// The band Ihat is [q][i] (i = qr + qrm fastest) and the cotangent spectrum
// [i][j] (j = kc + qc fastest): each workgroup moves a 32 x 32 tile through
// LDS so that both are read and written along their rows.
constexpr int kBandTile = 32;

__device__ __forceinline__ float2 band_value(const float2* __restrict__ Ib, int nb, int qr, int i,
                                             int kc) {
  if (kc > 0) return Ib[(size_t)kc * nb + i];
  const int im = 2 * qr - i;                 // the order (-kr)
  if (kc < 0) {
    const float2 u = Ib[(size_t)(-kc) * nb + im];
    return make_float2(u.x, -u.y);
  }
  const float2 a = Ib[i], b = Ib[im];
  return make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
}

__global__ __launch_bounds__(kIltThreads) void k_ilt_band_cotangent(
    const float2* __restrict__ Ihat, int qr, int qc, const float2* __restrict__ tband,
    const float* __restrict__ tconst, float n, float2* __restrict__ specg,
    float* __restrict__ part) {
  constexpr int TT = kBandTile, ROWS = kIltThreads / TT;
  __shared__ float2 tile[TT][TT + 1];        // [jj][ii]
  __shared__ float red[kIltThreads / 64];
  const int img = blockIdx.z;
  const int nb = 2 * qr + 1, wgc = 2 * qc + 1;
  const int i0 = blockIdx.y * TT, j0 = blockIdx.x * TT;
  const float2* Ib = Ihat + (size_t)img * (qc + 1) * nb;
  const int lane = threadIdx.x % TT, row = threadIdx.x / TT;
  // S along i (the band's rows)
#pragma unroll
  for (int r = row; r < TT; r += ROWS) {
    const int i = i0 + lane, j = j0 + r;
    tile[r][lane] = (i < nb && j < wgc) ? band_value(Ib, nb, qr, i, j - qc) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  // 2 (S - T~) along j (the cotangent spectrum's rows)
  const size_t E = (size_t)nb * wgc;
  const float2* Tb = tband + (size_t)img * E;
  float2* Gb = specg + (size_t)img * E;
  float acc = 0.f;
#pragma unroll
  for (int r = row; r < TT; r += ROWS) {
    const int i = i0 + r, j = j0 + lane;
    if (i < nb && j < wgc) {
      const float2 s = tile[lane][r];
      const float2 t = Tb[(size_t)i * wgc + j];
      const float dx = s.x - t.x, dy = s.y - t.y;
      acc += dx * dx + dy * dy;
      Gb[(size_t)i * wgc + j] = make_float2(2.f * dx, 2.f * dy);
    }
  }
  const float sum = block_sum(acc, red);
  if (threadIdx.x == 0) {
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;
    part[(size_t)img * gridDim.x * gridDim.y + blk] = n * sum + (blk == 0 ? tconst[img] : 0.f);
  }
}

// Per-image constant sum T^2 - n sum_band |T~|^2 in float64, in two fixed
// splits (kConstBlocks partial sums of T^2 per image, then one workgroup per
// image), so the result does not depend on timing.
constexpr int kConstBlocks = 64;

__global__ __launch_bounds__(kIltThreads) void k_ilt_target_sq(const float* __restrict__ target,
                                                               size_t npix,
                                                               double* __restrict__ part) {
  __shared__ double red[kIltThreads / 64];
  const float* T = target + (size_t)blockIdx.y * npix;
  double a = 0.0;
  for (size_t i = (size_t)blockIdx.x * kIltThreads + threadIdx.x; i < npix;
       i += (size_t)kConstBlocks * kIltThreads) {
    const double v = T[i];
    a += v * v;
  }
  a = wave_sum(a);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kIltThreads / 64; ++w) s += red[w];
    part[(size_t)blockIdx.y * kConstBlocks + blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(kIltThreads) void k_ilt_target_const(
    const double* __restrict__ sq, const float2* __restrict__ tband, size_t nband, double n,
    float* __restrict__ tconst) {
  __shared__ double red[kIltThreads / 64];
  const int img = blockIdx.x;
  const float2* Tb = tband + (size_t)img * nband;
  double b = 0.0;
  for (size_t i = threadIdx.x; i < nband; i += kIltThreads) {
    const float2 u = Tb[i];
    b += (double)u.x * u.x + (double)u.y * u.y;
  }
  b = wave_sum(b);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, s = 0.0;
    for (int k = 0; k < kConstBlocks; ++k) a += sq[(size_t)img * kConstBlocks + k];
#pragma unroll
    for (int w = 0; w < kIltThreads / 64; ++w) s += red[w];
    tconst[img] = (float)(a - n * s);
  }
}

}  // namespace
}  // namespace oi

namespace oi {
int ilt_sum(const float* part, int n, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_ilt_total, dim3(1), dim3(64), 0, st, part, n, out);
  OI_LAUNCH_CHECK("ilt_total");
  return OI_OK;
}

int ilt_band_blocks(int qr, int qc) {
  const int t = kBandTile;
  return ((2 * qr + 1 + t - 1) / t) * ((2 * qc + 1 + t - 1) / t);
}

int ilt_band_cotangent(const float2* Ihat, int qr, int qc, const float2* tband,
                       const float* tconst, float n, int batch, float2* specg, float* part,
                       hipStream_t st) {
  const int t = kBandTile;
  const dim3 grid((2 * qc + 1 + t - 1) / t, (2 * qr + 1 + t - 1) / t, batch);
  hipLaunchKernelGGL(k_ilt_band_cotangent, grid, dim3(kIltThreads), 0, st, Ihat, qr, qc, tband,
                     tconst, n, specg, part);
  OI_LAUNCH_CHECK("ilt_band_cotangent");
  return OI_OK;
}

int ilt_update_theta(float* theta, const float* gt, float lr, int batch, size_t npix,
                     const float* part, int nparts, hipStream_t st) {
  const dim3 grid(kIltBlocks, batch);
  if (npix % 4 == 0 && aligned16(theta) && aligned16(gt))
    hipLaunchKernelGGL(k_ilt_update_theta<true>, grid, dim3(kIltThreads), 0, st, theta, gt, lr, npix,
                       part, nparts);
  else
    hipLaunchKernelGGL(k_ilt_update_theta<false>, grid, dim3(kIltThreads), 0, st, theta, gt, lr,
                       npix, part, nparts);
  OI_LAUNCH_CHECK("ilt_update_theta");
  return OI_OK;
}

size_t ilt_target_const_scratch(int batch) {
  return (size_t)batch * kConstBlocks * sizeof(double);
}

int ilt_target_const(const float* target, size_t npix, const float2* tband, size_t nband,
                     int batch, void* scratch, float* tconst, hipStream_t st) {
  double* sq = static_cast<double*>(scratch);
  hipLaunchKernelGGL(k_ilt_target_sq, dim3(kConstBlocks, batch), dim3(kIltThreads), 0, st, target,
                     npix, sq);
  OI_LAUNCH_CHECK("ilt_target_sq");
  hipLaunchKernelGGL(k_ilt_target_const, dim3(batch), dim3(kIltThreads), 0, st, sq, tband, nband,
                     (double)npix, tconst);
  OI_LAUNCH_CHECK("ilt_target_const");
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
