// Hermitian TCC contraction on MFMA, and the float64 Gram route.
//
// TCC[m][m'] = alpha sum_r a_r[m] conj(a_r[m']),  a_r = sqrt(w_s) G[e][s] (r = (e, s)).
// With a = x + i y over K = 2r real rows X[m] = (x, y), Z[m] = (-y, x):
//   Re TCC = X X^T,  Im TCC = X Z^T,
// two real GEMMs over K = 2r.  Operands are stored [M][Kp] (K fastest) so a
// lane's MFMA fragment is one 16-byte load.  Only tiles J >= I are computed; an
// off-diagonal tile is also written transposed and conjugated.
//
// Split-bf16 (3 terms): x = hi + lo with hi = bf16(x), lo = bf16(x - hi);
// x x' ~ hi hi' + hi lo' + lo hi' accumulated in fp32 (the lo lo' term, ~2^-16
// relative, is dropped).
#include <cmath>

#include "oi_internal.h"

namespace oi {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 128;  // output tile per workgroup (4 waves of 64 x 64)

__device__ __forceinline__ unsigned short bf16_bits(float x) {
  // round-to-nearest-even; inputs here are finite
  unsigned int u = __float_as_uint(x);
  const unsigned int lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_value(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

// Operand preparation.  mode 0: bf16 hi/lo planes [4][M][Kp] = Xh, Xl, Zh, Zl;
// mode 1: fp32 planes [2][M][Kp] = X, Z; mode 2: fp64 planes [2][M][Kp].
__global__ void k_tcc_prep(const float2* __restrict__ G, const float* __restrict__ w, int M, int nsrc,
                           int ncomp, int Kp, int mode, void* __restrict__ work) {
  const int r = ncomp * nsrc;
  const size_t total = (size_t)M * Kp;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int m = (int)(e / Kp), k = (int)(e - (size_t)m * Kp);
    float xv = 0.f, zv = 0.f;
    if (k < 2 * r) {
      const int rr = k < r ? k : k - r;
      const int s = rr % nsrc;
      const float2 a = G[(size_t)rr * M + m];
      const float sw = sqrtf(w[s]);
      const float re = a.x * sw, im = a.y * sw;
      if (k < r) {
        xv = re;
        zv = -im;
      } else {
        xv = im;
        zv = re;
      }
    }
    if (mode == 0) {
      unsigned short* P = static_cast<unsigned short*>(work);
      const unsigned short xh = bf16_bits(xv), zh = bf16_bits(zv);
      P[e] = xh;
      P[total + e] = bf16_bits(xv - bf16_value(xh));
      P[2 * total + e] = zh;
      P[3 * total + e] = bf16_bits(zv - bf16_value(zh));
    } else if (mode == 1) {
      float* P = static_cast<float*>(work);
      P[e] = xv;
      P[total + e] = zv;
    } else {
      double* P = static_cast<double*>(work);
      P[e] = (double)xv;
      P[total + e] = (double)zv;
    }
  }
}

__device__ __forceinline__ void tile_pair(int t, int nt, int& I, int& J) {
  // t enumerates pairs I <= J row by row.
  I = 0;
  int rem = t;
  while (rem >= nt - I) {
    rem -= nt - I;
    ++I;
  }
  J = I + rem;
}

constexpr int kXcds = 8;       // workgroup b runs on XCD b mod 8

// Supertile order of the split-bf16 contraction: the upper-triangle tiles are
// grouped into kSup x kSup supertiles, enumerated row by row, and supertile s
// goes to XCD s mod 8 with its tiles row by row.  A supertile's tiles run on
// one XCD at nearly the same time, so the partial cache lines where their
// output row segments meet (rows are 8 M bytes apart, not line-aligned) are
// merged in that XCD's L2 before they reach HBM, and each operand panel is
// read by kSup tiles there.  At M = 10201 the contraction took 0.42 ms in
// the earlier column order (XCD x: column tiles 8k + x), 0.37 ms with 4 x 4
// supertiles and 0.36 ms with 8 x 8.
constexpr int kSup = 8;

__host__ __device__ inline int sup_count(int si, int sj, int ntb) {
  int cmax = ntb - kSup * sj;
  cmax = cmax < kSup ? cmax : kSup;             // valid column tiles of the strip
  return si < sj ? kSup * cmax : cmax * (cmax + 1) / 2;
}

// Workgroups of the supertile-ordered grid: 8 x the largest per-XCD count.
static int sup_tile_blocks(int ntb) {
  const int ns = (ntb + kSup - 1) / kSup;
  int cnt[kXcds] = {};
  int s = 0;
  for (int si = 0; si < ns; ++si)
    for (int sj = si; sj < ns; ++sj, ++s) cnt[s % kXcds] += sup_count(si, sj, ntb);
  int most = 0;
  for (int x = 0; x < kXcds; ++x) most = cnt[x] > most ? cnt[x] : most;
  return kXcds * most;
}

__device__ __forceinline__ int floor_div8(int a) { return a >= 0 ? a / kXcds : -((-a + kXcds - 1) / kXcds); }

// Tile (I, J) of workgroup b, false past its XCD's last tile: the XCD's
// supertiles are walked a supertile row at a time (O(1) per row), then within
// the row, then the tiles inside the supertile row by row.
__device__ __forceinline__ bool sup_tile(int b, int ntb, int& I, int& J) {
  const int x = b % kXcds;
  int rem = b / kXcds;
  const int ns = (ntb + kSup - 1) / kSup;
  int s0 = 0;                                   // index of the row's first (diagonal) supertile
  for (int si = 0; si < ns; ++si) {
    const int len = ns - si;
    const int nx = floor_div8(s0 + len - 1 - x) - floor_div8(s0 - 1 - x);
    int rowt = kSup * kSup * nx;
    if (nx > 0) {
      if (s0 % kXcds == x) rowt -= kSup * kSup - sup_count(si, si, ntb);
      if (len > 1 && (s0 + len - 1) % kXcds == x) rowt -= kSup * kSup - sup_count(si, ns - 1, ntb);
    }
    if (rem < rowt) {
      for (int s = s0 + (x - s0 % kXcds + kXcds) % kXcds;; s += kXcds) {
        const int sj = si + (s - s0);
        const int n = sup_count(si, sj, ntb);
        if (rem < n) {
          int cmax = ntb - kSup * sj;
          cmax = cmax < kSup ? cmax : kSup;
          int a, c;
          if (si < sj) {
            a = rem / cmax;
            c = rem - a * cmax;
          } else {                              // diagonal supertile: rows hold c >= a
            a = 0;
            while (rem >= cmax - a) {
              rem -= cmax - a;
              ++a;
            }
            c = a + rem;
          }
          I = kSup * si + a;
          J = kSup * sj + c;
          return true;
        }
        rem -= n;
      }
    }
    rem -= rowt;
    s0 += len;
  }
  return false;
}

template <typename T, int V>
__device__ __forceinline__ void load_frag(const T* __restrict__ P, int M, int Kp, int row, int k,
                                          T* out) {
  if (row < M) {
#pragma unroll
    for (int v = 0; v < V; ++v) out[v] = P[(size_t)row * Kp + k + v];
  } else {
#pragma unroll
    for (int v = 0; v < V; ++v) out[v] = T(0);
  }
}

// Orders this wavefront's LDS accesses (a wavefront's LDS operations complete
// in order; the fences and the scheduling barrier keep the compiler from
// moving them).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kMirStride = 17;   // float2 per row of the 16 x 16 transpose buffer

// Writes a 16 x 16 accumulator fragment and, for off-diagonal tiles, its
// conjugate transpose.  Lane l holds column l&15 and rows 4(l>>4)+i for the
// fp32-accumulating forms, rows (l>>4)+4i for the fp64 form (INTERLEAVED).
// Row-major stores of the fragment are 16 consecutive entries per row; the
// transpose goes through the wave's LDS buffer (16 x 17 float2) so that it is
// written as 16 consecutive entries per row too.
template <bool INTERLEAVED, typename A>
__device__ __forceinline__ void store_frag(float2* __restrict__ tcc, int M, int r0, int c0,
                                           const A& re, const A& im, double alpha, bool mirror,
                                           float2* buf) {
  const int lane = threadIdx.x & 63;
  const int cl = lane & 15;
  const int col = c0 + cl;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int rl = INTERLEAVED ? (lane >> 4) + 4 * i : (lane >> 4) * 4 + i;
    const int row = r0 + rl;
    const float vr = (float)(alpha * (double)re[i]);
    const float vi = (float)(alpha * (double)im[i]);
    if (row < M && col < M) tcc[(size_t)row * M + col] = make_float2(vr, vi);
    if (mirror) buf[cl * kMirStride + rl] = make_float2(vr, -vi);
  }
  if (!mirror) return;   // uniform over the workgroup
  wave_lds_sync();
  const int mr = c0 + (lane >> 2);          // row of the transpose = fragment column
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int e = (lane & 3) * 4 + i;       // fragment row
    if (mr < M && r0 + e < M) tcc[(size_t)mr * M + r0 + e] = buf[(lane >> 2) * kMirStride + e];
  }
  wave_lds_sync();                          // the buffer is rewritten by the next fragment
}

// Split-bf16 contraction: workgroup tile TB x TB (4 waves in 2 x 2, each
// (TB/2)^2 as FT x FT MFMA tiles).  Panels of a k-step (32 wide) are staged in
// LDS: A rows (tile I) and X, Z rows (tile J), hi and lo each; rows padded to
// 40 halves (16-byte aligned, spread over banks).  The next k-step's panels
// are loaded into registers while this step's MFMAs run.
constexpr int kPanelK = 32, kPanelRow = 40;
constexpr int kPanels = 6;                      // Ah Al Xh Xl Zh Zl
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int TB>
struct TccTile {
  static constexpr int FT = TB / 32;                        // MFMA tiles per wave and dimension
  static constexpr int kChunks = kPanels * TB * 4 / 256;    // 16-byte panel chunks per thread
  static constexpr int kOutStride = TB + 1;
  static constexpr size_t kPanBytes = (size_t)kPanels * TB * kPanelRow * sizeof(unsigned short);
  static constexpr size_t kOutBytes = (size_t)TB * kOutStride * sizeof(float2);
  static constexpr size_t kLds = kPanBytes > kOutBytes ? kPanBytes : kOutBytes;
};

template <int TB>
__device__ __forceinline__ void tcc_panel_load(const unsigned short* __restrict__ planes, size_t plane,
                                               int M, int Kp, int rI, int rJ, int k0,
                                               u32x4 (&reg)[TccTile<TB>::kChunks]) {
#pragma unroll
  for (int q = 0; q < TccTile<TB>::kChunks; ++q) {
    const int c = threadIdx.x + 256 * q;
    const int panel = c / (TB * 4), rem = c - panel * TB * 4, row = rem >> 2, ch = rem & 3;
    // panel order: A hi (X hi rows of I), A lo, X hi (rows of J), X lo, Z hi, Z lo
    const int pl = panel < 2 ? panel : panel - 2;
    const int r = (panel < 2 ? rI : rJ) + row;
    reg[q] = r < M ? *reinterpret_cast<const u32x4*>(planes + pl * plane + (size_t)r * Kp + k0 + ch * 8)
                   : u32x4{0u, 0u, 0u, 0u};
  }
}

template <int TB>
__device__ __forceinline__ void tcc_panel_store(unsigned short* lds,
                                                const u32x4 (&reg)[TccTile<TB>::kChunks]) {
#pragma unroll
  for (int q = 0; q < TccTile<TB>::kChunks; ++q) {
    const int c = threadIdx.x + 256 * q;
    const int panel = c / (TB * 4), rem = c - panel * TB * 4, row = rem >> 2, ch = rem & 3;
    *reinterpret_cast<u32x4*>(lds + (panel * TB + row) * kPanelRow + ch * 8) = reg[q];
  }
}

// Output tile staging: the workgroup's TB x TB result goes through LDS so
// that each store instruction writes one TB-entry row segment of the TCC, and
// the mirrored tile one row of the conjugate transpose, instead of four
// 128-byte pieces of four rows.  It shares the LDS (dynamic, TccTile::kLds
// bytes) of the operand panels.
// Four waves per SIMD: the register budget (128) that lets four workgroups
// share a CU (their 33 KB of LDS each fit four times); 0.346 ms at M = 10201
// against 0.358-0.370 ms at the 132 registers the compiler takes unbounded.
template <int TB>
__global__ __launch_bounds__(256, 4) void k_tcc_bf16(const unsigned short* __restrict__ planes, int M,
                                                  int Kp, int nt, double alpha,
                                                  float2* __restrict__ tcc) {
  using TT = TccTile<TB>;
  constexpr int FT = TT::FT, kOutStride = TT::kOutStride;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  unsigned short* pan = reinterpret_cast<unsigned short*>(lds_raw);
  const size_t plane = (size_t)M * Kp;
  int I, J;
  if (!sup_tile(blockIdx.x, nt, I, J)) return;  // uniform: this XCD's share is done
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ar = (wave >> 1) * (TB / 2), bc = (wave & 1) * (TB / 2);
  const int li = lane & 15, lk = (lane >> 4) * 8;
  f32x4 accR[FT][FT], accI[FT][FT];
#pragma unroll
  for (int a = 0; a < FT; ++a)
#pragma unroll
    for (int b = 0; b < FT; ++b) {
      accR[a][b] = f32x4{0, 0, 0, 0};
      accI[a][b] = f32x4{0, 0, 0, 0};
    }
  u32x4 nxt[TT::kChunks];
  tcc_panel_load<TB>(planes, plane, M, Kp, I * TB, J * TB, 0, nxt);
  for (int k0 = 0; k0 < Kp; k0 += kPanelK) {
    __syncthreads();                            // previous step's fragment reads are done
    tcc_panel_store<TB>(pan, nxt);
    __syncthreads();
    if (k0 + kPanelK < Kp) tcc_panel_load<TB>(planes, plane, M, Kp, I * TB, J * TB, k0 + kPanelK, nxt);
    bf16x8 ah[FT], al[FT], rh[FT], rl[FT], ih[FT], il[FT];
#pragma unroll
    for (int t = 0; t < FT; ++t) {
      const int ra = ar + t * 16 + li, cb = bc + t * 16 + li;
      ah[t] = *reinterpret_cast<const bf16x8*>(pan + (0 * TB + ra) * kPanelRow + lk);
      al[t] = *reinterpret_cast<const bf16x8*>(pan + (1 * TB + ra) * kPanelRow + lk);
      rh[t] = *reinterpret_cast<const bf16x8*>(pan + (2 * TB + cb) * kPanelRow + lk);
      rl[t] = *reinterpret_cast<const bf16x8*>(pan + (3 * TB + cb) * kPanelRow + lk);
      ih[t] = *reinterpret_cast<const bf16x8*>(pan + (4 * TB + cb) * kPanelRow + lk);
      il[t] = *reinterpret_cast<const bf16x8*>(pan + (5 * TB + cb) * kPanelRow + lk);
    }
#pragma unroll
    for (int a = 0; a < FT; ++a)
#pragma unroll
      for (int b = 0; b < FT; ++b) {
        accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], rh[b], accR[a][b], 0, 0, 0);
        accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], ih[b], accI[a][b], 0, 0, 0);
        accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], rl[b], accR[a][b], 0, 0, 0);
        accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], il[b], accI[a][b], 0, 0, 0);
        accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], rh[b], accR[a][b], 0, 0, 0);
        accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], ih[b], accI[a][b], 0, 0, 0);
      }
  }
  // stage the tile (alpha applied, rounded to float32) in LDS
  float2* otile = reinterpret_cast<float2*>(lds_raw);
  __syncthreads();                              // the last step's panel reads are done
#pragma unroll
  for (int a = 0; a < FT; ++a)
#pragma unroll
    for (int b = 0; b < FT; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int rl = ar + a * 16 + (lane >> 4) * 4 + i, cl = bc + b * 16 + (lane & 15);
        otile[rl * kOutStride + cl] = make_float2((float)(alpha * (double)accR[a][b][i]),
                                                  (float)(alpha * (double)accI[a][b][i]));
      }
  __syncthreads();
  // tile (I, J): wave w writes rows w, w + 4, ...; lanes cover the row's TB entries
  for (int rl = wave; rl < TB; rl += 4) {
    const int row = I * TB + rl;
#pragma unroll
    for (int h = 0; h < TB / 64; ++h) {
      const int cl = lane + 64 * h, col = J * TB + cl;
      if (row < M && col < M) tcc[(size_t)row * M + col] = otile[rl * kOutStride + cl];
    }
  }
  if (I == J) return;                           // uniform over the workgroup
  // tile (J, I) = conjugate transpose: row J*TB + cl holds column cl of the tile
  for (int cl = wave; cl < TB; cl += 4) {
    const int row = J * TB + cl;
#pragma unroll
    for (int h = 0; h < TB / 64; ++h) {
      const int rl = lane + 64 * h, mcol = I * TB + rl;
      const float2 v = otile[rl * kOutStride + cl];
      if (row < M && mcol < M) tcc[(size_t)row * M + mcol] = make_float2(v.x, -v.y);
    }
  }
}

// Workgroup tile of the split-bf16 contraction.  At M = 10201 a 128 x 128
// tile (64 x 64 per wave, 1 KB output row segments) ran in the same time as
// 64 x 64 (0.43 against 0.42 ms): its 294 registers leave one workgroup per
// CU, and the output stream, not the MFMAs, sets the pace (the contraction
// without its stores takes 0.24 ms).
constexpr int kTileB = 64;

// fp32 MFMA 16x16x4: lane fragment is one value at k = l>>4; a 16-wide K step
// runs four MFMAs, sub-step kk using k = k0 + 4 (l>>4) + kk for both operands.
__global__ __launch_bounds__(256) void k_tcc_f32(const float* __restrict__ planes, int M, int Kp,
                                                 int nt, double alpha, float2* __restrict__ tcc) {
  const size_t plane = (size_t)M * Kp;
  const float* X = planes;
  const float* Z = planes + plane;
  int I, J;
  tile_pair(blockIdx.x, nt, I, J);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r0 = I * kTile + (wave >> 1) * 64;
  const int c0 = J * kTile + (wave & 1) * 64;
  const int li = lane & 15, lk = (lane >> 4) * 4;
  f32x4 accR[4][4], accI[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      accR[a][b] = f32x4{0, 0, 0, 0};
      accI[a][b] = f32x4{0, 0, 0, 0};
    }
  for (int k0 = 0; k0 < Kp; k0 += 16) {
    float fa[4][4], fr[4][4], fi[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      load_frag<float, 4>(X, M, Kp, r0 + t * 16 + li, k0 + lk, fa[t]);
      load_frag<float, 4>(X, M, Kp, c0 + t * 16 + li, k0 + lk, fr[t]);
      load_frag<float, 4>(Z, M, Kp, c0 + t * 16 + li, k0 + lk, fi[t]);
    }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a][kk], fr[b][kk], accR[a][b], 0, 0, 0);
          accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a][kk], fi[b][kk], accI[a][b], 0, 0, 0);
        }
  }
  const bool mirror = I != J;
  __shared__ float2 mbuf_all[4 * 16 * kMirStride];
  float2* mbuf = mbuf_all + (threadIdx.x >> 6) * 16 * kMirStride;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
      store_frag<false>(tcc, M, r0 + a * 16, c0 + b * 16, accR[a][b], accI[a][b], alpha, mirror, mbuf);
}

// fp64 MFMA 16x16x4 (64 x 32 per wave to bound registers: tile 128 x 64 per block).
__global__ __launch_bounds__(256) void k_tcc_f64(const double* __restrict__ planes, int M, int Kp,
                                                 int nt, double alpha, float2* __restrict__ tcc) {
  const size_t plane = (size_t)M * Kp;
  const double* X = planes;
  const double* Z = planes + plane;
  int I, J;
  tile_pair(blockIdx.x, nt, I, J);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r0 = I * kTile + (wave >> 1) * 64;
  const int li = lane & 15, lk = (lane >> 4) * 2;
  for (int half = 0; half < 2; ++half) {
    const int c0 = J * kTile + (wave & 1) * 64 + half * 32;
    f64x4 accR[4][2], accI[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        accR[a][b] = f64x4{0, 0, 0, 0};
        accI[a][b] = f64x4{0, 0, 0, 0};
      }
    for (int k0 = 0; k0 < Kp; k0 += 8) {
      double fa[4][2], fr[2][2], fi[2][2];
#pragma unroll
      for (int t = 0; t < 4; ++t) load_frag<double, 2>(X, M, Kp, r0 + t * 16 + li, k0 + lk, fa[t]);
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        load_frag<double, 2>(X, M, Kp, c0 + t * 16 + li, k0 + lk, fr[t]);
        load_frag<double, 2>(Z, M, Kp, c0 + t * 16 + li, k0 + lk, fi[t]);
      }
#pragma unroll
      for (int kk = 0; kk < 2; ++kk)
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            accR[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a][kk], fr[b][kk], accR[a][b], 0, 0, 0);
            accI[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a][kk], fi[b][kk], accI[a][b], 0, 0, 0);
          }
    }
    const bool mirror = I != J;
    __shared__ float2 mbuf_all[4 * 16 * kMirStride];
    float2* mbuf = mbuf_all + (threadIdx.x >> 6) * 16 * kMirStride;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
        store_frag<true>(tcc, M, r0 + a * 16, c0 + b * 16, accR[a][b], accI[a][b], alpha, mirror, mbuf);
  }
}

// H[r][r'] = sqrt(w_s w_s') sum_m conj(G[r][m]) G[r'][m], float64.  Block =
// 16 x 16 entries (upper blocks only, mirrored), m streamed through LDS.
__global__ __launch_bounds__(256) void k_gram(const float2* __restrict__ G, const double* __restrict__ w,
                                              int M, int nsrc, int n, int nb,
                                              double2* __restrict__ H) {
  __shared__ double2 sa[16][65];
  __shared__ double2 sb[16][65];
  int I, J;
  tile_pair(blockIdx.x, nb, I, J);
  const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
  const int ra = I * 16, rb = J * 16;
  double accr = 0.0, acci = 0.0;
  for (int m0 = 0; m0 < M; m0 += 64) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int rr = e >> 6, mm = e & 63;
      const int m = m0 + mm;
      float2 va = make_float2(0.f, 0.f), vb = make_float2(0.f, 0.f);
      if (m < M) {
        if (ra + rr < n) va = G[(size_t)(ra + rr) * M + m];
        if (rb + rr < n) vb = G[(size_t)(rb + rr) * M + m];
      }
      sa[rr][mm] = make_double2(va.x, va.y);
      sb[rr][mm] = make_double2(vb.x, vb.y);
    }
    __syncthreads();
#pragma unroll 8
    for (int mm = 0; mm < 64; ++mm) {
      const double2 a = sa[ti][mm], b = sb[tj][mm];
      accr += a.x * b.x + a.y * b.y;  // conj(a) b
      acci += a.x * b.y - a.y * b.x;
    }
    __syncthreads();
  }
  const int r = ra + ti, c = rb + tj;
  if (r < n && c < n) {
    const double s = sqrt(w[r % nsrc] * w[c % nsrc]);
    H[(size_t)r * n + c] = make_double2(s * accr, s * acci);
    if (I != J) H[(size_t)c * n + r] = make_double2(s * accr, -s * acci);
  }
}

// out[k][m] = sum_r sqrt(w_s) G[r][m] Q[r][k], float64 accumulation.
__global__ void k_from_gram(const float2* __restrict__ G, const double* __restrict__ w, int M, int nsrc,
                            int n, const double2* __restrict__ Q, int nk, float2* __restrict__ out) {
  const size_t total = (size_t)nk * M;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(e / M), m = (int)(e - (size_t)k * M);
    double ar = 0.0, ai = 0.0;
    for (int r = 0; r < n; ++r) {
      const float2 g = G[(size_t)r * M + m];
      const double2 q = Q[(size_t)r * nk + k];
      const double s = sqrt(w[r % nsrc]);
      ar += s * (g.x * q.x - g.y * q.y);
      ai += s * (g.x * q.y + g.y * q.x);
    }
    out[e] = make_float2((float)ar, (float)ai);
  }
}

static size_t kp_for(int nsrc, int ncomp, int precision) {
  const size_t k2 = 2ull * nsrc * ncomp;
  const size_t q = precision == OI_TCC_SPLIT_BF16 ? 32 : (precision == OI_TCC_FP32 ? 16 : 8);
  return (k2 + q - 1) / q * q;
}

}  // namespace oi

using namespace oi;

extern "C" size_t oi_tcc_work_bytes(int M, int nsrc, int ncomp, int precision) {
  if (M < 1 || nsrc < 1 || ncomp < 1) return 0;
  const size_t kp = kp_for(nsrc, ncomp, precision);
  const size_t e = (size_t)M * kp;
  if (precision == OI_TCC_SPLIT_BF16) return 4 * e * sizeof(unsigned short);
  if (precision == OI_TCC_FP32) return 2 * e * sizeof(float);
  return 2 * e * sizeof(double);
}

extern "C" int oi_tcc_build(const void* G, const float* w, int M, int nsrc, int ncomp, double alpha,
                            int precision, void* work, void* tcc, void* stream) {
  if (!G || !w || !work || !tcc || M < 1 || nsrc < 1 || ncomp < 1)
    return fail(OI_ERR_INVALID, "tcc_build: bad arguments");
  if (precision != OI_TCC_SPLIT_BF16 && precision != OI_TCC_FP32 && precision != OI_TCC_FP64)
    return fail(OI_ERR_INVALID, "tcc_build: unknown precision");
  hipStream_t st = (hipStream_t)stream;
  const int Kp = (int)kp_for(nsrc, ncomp, precision);
  const int mode = precision == OI_TCC_SPLIT_BF16 ? 0 : (precision == OI_TCC_FP32 ? 1 : 2);
  size_t tot = (size_t)M * Kp;
  int blocks = (int)((tot + 255) / 256 > 16384 ? 16384 : (tot + 255) / 256);
  hipLaunchKernelGGL(k_tcc_prep, dim3(blocks), dim3(256), 0, st, static_cast<const float2*>(G), w, M,
                     nsrc, ncomp, Kp, mode, work);
  OI_LAUNCH_CHECK("tcc_prep");
  const int nt = (M + kTile - 1) / kTile;
  const int pairs = nt * (nt + 1) / 2;
  float2* out = static_cast<float2*>(tcc);
  if (mode == 0) {
    const int ntb = (M + kTileB - 1) / kTileB;
    const int blocks = sup_tile_blocks(ntb);
    constexpr size_t lds = TccTile<kTileB>::kLds;
    if (lds > 64 * 1024)
      OI_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_tcc_bf16<kTileB>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_tcc_bf16<kTileB>, dim3(blocks), dim3(256), lds, st,
                       static_cast<const unsigned short*>(work), M, Kp, ntb, alpha, out);
  } else if (mode == 1)
    hipLaunchKernelGGL(k_tcc_f32, dim3(pairs), dim3(256), 0, st, static_cast<const float*>(work), M,
                       Kp, nt, alpha, out);
  else
    hipLaunchKernelGGL(k_tcc_f64, dim3(pairs), dim3(256), 0, st, static_cast<const double*>(work), M,
                       Kp, nt, alpha, out);
  OI_LAUNCH_CHECK("tcc_gemm");
  return OI_OK;
}

extern "C" int oi_tcc_gram(const void* G, const double* w, int M, int nsrc, int ncomp, void* H,
                           void* stream) {
  if (!G || !w || !H || M < 1 || nsrc < 1 || ncomp < 1) return fail(OI_ERR_INVALID, "tcc_gram: bad arguments");
  const int n = nsrc * ncomp;
  const int nb = (n + 15) / 16;
  hipLaunchKernelGGL(k_gram, dim3(nb * (nb + 1) / 2), dim3(256), 0, (hipStream_t)stream,
                     static_cast<const float2*>(G), w, M, nsrc, n, nb, static_cast<double2*>(H));
  OI_LAUNCH_CHECK("tcc_gram");
  return OI_OK;
}

extern "C" int oi_kernels_from_gram(const void* G, const double* w, int M, int nsrc, int ncomp,
                                    const void* Q, int nk, void* out, void* stream) {
  if (!G || !w || !Q || !out || M < 1 || nsrc < 1 || ncomp < 1 || nk < 1)
    return fail(OI_ERR_INVALID, "kernels_from_gram: bad arguments");
  const size_t total = (size_t)nk * M;
  int blocks = (int)((total + 255) / 256 > 16384 ? 16384 : (total + 255) / 256);
  hipLaunchKernelGGL(k_from_gram, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                     static_cast<const float2*>(G), w, M, nsrc, nsrc * ncomp,
                     static_cast<const double2*>(Q), nk, static_cast<float2*>(out));
  OI_LAUNCH_CHECK("kernels_from_gram");
  return OI_OK;
}
