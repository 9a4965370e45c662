// Workgroup-level batched FFT in LDS: mixed radix 4/2/3/5 Stockham auto-sort.
//
// Stage with radix R after sub-transform length p (p = product of earlier
// radices): butterfly i in [0, n/R), k = i mod p, reads x_j = src[i + j n/R],
// applies twiddles w_n^{j k n/(p R)}, runs an R-point DFT and writes
// dst[(i - k) R + k + j p].  Output is in natural order after the last stage.
// Twiddles come from a table tw[m] = exp(-2 pi i m / n) computed on the host in
// float64 and rounded once; the inverse-sign transform uses conj(tw).
#pragma once

#include "oi_common.h"

namespace oi {

constexpr int kMaxStages = 16;

struct FftDesc {
  int n;
  int nst;
  unsigned char rad[kMaxStages];
  const float2* tw;  // n entries
};

// Radix plan for n = 2^a 3^b 5^c; returns false when n has another factor.
inline bool fft_factor(int n, FftDesc& d) {
  d.n = n;
  d.nst = 0;
  int m = n;
  while (m % 4 == 0 && d.nst < kMaxStages) { d.rad[d.nst++] = 4; m /= 4; }
  while (m % 2 == 0 && d.nst < kMaxStages) { d.rad[d.nst++] = 2; m /= 2; }
  while (m % 3 == 0 && d.nst < kMaxStages) { d.rad[d.nst++] = 3; m /= 3; }
  while (m % 5 == 0 && d.nst < kMaxStages) { d.rad[d.nst++] = 5; m /= 5; }
  return m == 1 && n >= 1;
}

inline bool fft_size_ok(int n) {
  FftDesc d;
  return n >= 1 && fft_factor(n, d);
}

template <bool INV>
__device__ __forceinline__ float2 twid(const float2* __restrict__ tw, int m) {
  float2 w = tw[m];
  return INV ? make_float2(w.x, -w.y) : w;
}

// In-register R-point DFT, sign -1 (forward) or +1 (INV).
template <int R, bool INV>
__device__ __forceinline__ void dft_small(float2* u);

template <>
__device__ __forceinline__ void dft_small<2, false>(float2* u) {
  float2 a = u[0], b = u[1];
  u[0] = cadd(a, b);
  u[1] = csub(a, b);
}
template <>
__device__ __forceinline__ void dft_small<2, true>(float2* u) {
  dft_small<2, false>(u);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2* u) {
  float2 s02 = cadd(u[0], u[2]), d02 = csub(u[0], u[2]);
  float2 s13 = cadd(u[1], u[3]), d13 = csub(u[1], u[3]);
  float2 rot = INV ? cmul_pi(d13) : cmul_mi(d13);  // -+ i (x1 - x3)
  u[0] = cadd(s02, s13);
  u[2] = csub(s02, s13);
  u[1] = cadd(d02, rot);
  u[3] = csub(d02, rot);
}
template <>
__device__ __forceinline__ void dft_small<4, false>(float2* u) { dft4<false>(u); }
template <>
__device__ __forceinline__ void dft_small<4, true>(float2* u) { dft4<true>(u); }

template <bool INV>
__device__ __forceinline__ void dft3(float2* u) {
  const float h = 0.86602540378443864676f;  // sqrt(3)/2
  float2 s = cadd(u[1], u[2]);
  float2 d = cscale(csub(u[1], u[2]), h);
  float2 t = make_float2(u[0].x - 0.5f * s.x, u[0].y - 0.5f * s.y);
  float2 rot = INV ? cmul_pi(d) : cmul_mi(d);
  u[0] = cadd(u[0], s);
  u[1] = cadd(t, rot);
  u[2] = csub(t, rot);
}
template <>
__device__ __forceinline__ void dft_small<3, false>(float2* u) { dft3<false>(u); }
template <>
__device__ __forceinline__ void dft_small<3, true>(float2* u) { dft3<true>(u); }

template <bool INV>
__device__ __forceinline__ void dft5(float2* u) {
  const float c1 = 0.30901699437494742410f;   // cos(2 pi / 5)
  const float c2 = -0.80901699437494742410f;  // cos(4 pi / 5)
  const float s1 = 0.95105651629515357212f;   // sin(2 pi / 5)
  const float s2 = 0.58778525229247312917f;   // sin(4 pi / 5)
  float2 a1 = cadd(u[1], u[4]), b1 = csub(u[1], u[4]);
  float2 a2 = cadd(u[2], u[3]), b2 = csub(u[2], u[3]);
  float2 x0 = u[0];
  float2 t1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
  float2 t2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
  float2 v1 = make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
  float2 v2 = make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
  float2 r1 = INV ? cmul_pi(v1) : cmul_mi(v1);
  float2 r2 = INV ? cmul_pi(v2) : cmul_mi(v2);
  u[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
  u[1] = cadd(t1, r1);
  u[4] = csub(t1, r1);
  u[2] = cadd(t2, r2);
  u[3] = csub(t2, r2);
}
template <>
__device__ __forceinline__ void dft_small<5, false>(float2* u) { dft5<false>(u); }
template <>
__device__ __forceinline__ void dft_small<5, true>(float2* u) { dft5<true>(u); }

template <int R, bool INV>
__device__ __forceinline__ void stockham_item(const float2* __restrict__ src, float2* __restrict__ dst,
                                              int i, int nq, int p, int twu,
                                              const float2* __restrict__ tw) {
  float2 u[R];
#pragma unroll
  for (int j = 0; j < R; ++j) u[j] = src[i + j * nq];
  const int k = i % p;
  const int ts = k * twu;
  if (ts != 0) {
#pragma unroll
    for (int j = 1; j < R; ++j) u[j] = cmul(u[j], twid<INV>(tw, j * ts));
  }
  dft_small<R, INV>(u);
  float2* d = dst + (i - k) * R + k;
#pragma unroll
  for (int j = 0; j < R; ++j) d[j * p] = u[j];
}

// nb transforms of length d.n, transform b at a[b*bstride ...].  All threads
// of the block must call; a must be fully written and synchronised before the
// call.  Returns the buffer (a or t) holding the result, synchronised.
template <bool INV>
__device__ float2* block_fft(float2* a, float2* t, const FftDesc& d, int nb, int bstride) {
  const int n = d.n;
  int p = 1;
  for (int st = 0; st < d.nst; ++st) {
    const int R = d.rad[st];
    const int nq = n / R;
    const int twu = n / (p * R);
    const int total = nb * nq;
    for (int w = threadIdx.x; w < total; w += blockDim.x) {
      const int bi = w / nq;
      const int i = w - bi * nq;
      const float2* s = a + bi * bstride;
      float2* o = t + bi * bstride;
      switch (R) {
        case 4: stockham_item<4, INV>(s, o, i, nq, p, twu, d.tw); break;
        case 2: stockham_item<2, INV>(s, o, i, nq, p, twu, d.tw); break;
        case 3: stockham_item<3, INV>(s, o, i, nq, p, twu, d.tw); break;
        default: stockham_item<5, INV>(s, o, i, nq, p, twu, d.tw); break;
      }
    }
    __syncthreads();
    float2* tmp = a;
    a = t;
    t = tmp;
    p *= R;
  }
  return a;
}

}  // namespace oi
