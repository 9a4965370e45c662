// Imaging kernels with compile-time transform sizes (register FFTs, one LDS
// exchange per transform).  Same mathematics and data flow as the generic
// kernels of oi_imaging.hip; used when the image sizes and the coarse grid are
// in the size tables below.
//
//   k_fp_rows_real<P>   mask rows -> T[img][j][r] = sum_m h[r][m] e^{+2 pi i (k0c+j) m / nx}
//   k_fp_rows_c64<P>    same for one channel of a complex channels-last field
//   k_fp_cols<P>        T columns -> spec[img][c][i][j] (window rows), scaled 1/(ny nx)
//   k_fp_coherent<PL,PM> spectrum x kernels -> coarse |E|^2 -> row DFT -> R
//   k_fp_band<PL>       R columns -> band of the coarse intensity spectrum Ihat
//   k_fp_ucols<P>       Ihat columns zero-padded to ny -> U[img][q][y]
//   k_fp_urows<P>       U rows (Hermitian) -> image rows (C2R)
#include <cmath>
#include <type_traits>

#include "oi_fastpath.h"
#include "oi_fftreg.h"

namespace oi {
namespace fp {

using reg::GFFT;
using reg::pad;

constexpr int kBlock = 256;

template <class P>
constexpr int groups_per_block() {
  return kBlock / P::T;
}

// Transform scratch is allocated for every thread of the block, including the
// threads of an incomplete last group (they run the transforms' barriers with
// zeros and must not write into a live group's scratch).
template <class P>
constexpr int scratch_groups() {
  return (kBlock + P::T - 1) / P::T;
}

// ------------------------------------------------------------------ rows
// Real masks, two rows per complex transform (real parts row r0, imaginary
// parts row r0 + 1), outputs separated by Hermitian symmetry.
template <class P>
__global__ __launch_bounds__(kBlock) void k_fp_rows_real(const float* __restrict__ mask, int ny,
                                                         int k0c, int wc,
                                                         const float2* __restrict__ twt,
                                                         float2* __restrict__ T, int logits) {
  constexpr int N = P::N, TT = P::T, V = P::V, G = groups_per_block<P>();
  extern __shared__ float2 lds[];
  const int g = threadIdx.x / TT, t = threadIdx.x - g * TT;
  const bool live = g < G;
  const int gg = live ? g : 0;
  float2* scr = lds + g * P::LDS;
  const int img = blockIdx.y;
  const int r0 = 2 * (blockIdx.x * G + gg), r1 = r0 + 1;
  const float* src = mask + (size_t)img * ny * N;
  float2 v[V];
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const int x = t + TT * m;
    float a = (live && r0 < ny) ? __builtin_nontemporal_load(src + (size_t)r0 * N + x) : 0.f;
    float b = (live && r1 < ny) ? __builtin_nontemporal_load(src + (size_t)r1 * N + x) : 0.f;
    if (logits) {                    // rows past ny stay zero
      a = (live && r0 < ny) ? logistic(a) : 0.f;
      b = (live && r1 < ny) ? logistic(b) : 0.f;
    }
    v[m] = make_float2(a, b);
  }
  P::template run_live_tab<true>(v, scr, t, twt, true);
  if (live) {
#pragma unroll
    for (int m = 0; m < V; ++m) scr[pad(t + TT * m)] = v[m];
  }
  __syncthreads();
  // outputs of the block's 2G rows go through an LDS tile so that T, which is
  // column-major per image ([img][j][r]), is written in whole row runs
  constexpr int TS = 2 * G + 1;
  float2* tile = lds + scratch_groups<P>() * P::LDS;   // [wc][TS]
  if (live && r0 < ny) {
    for (int j = t; j < wc; j += TT) {
      const int k = pmod(k0c + j, N);
      const int km = k == 0 ? 0 : N - k;
      const float2 zk = scr[pad(k)], zm = scr[pad(km)];
      tile[j * TS + 2 * gg] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));      // (Z[k] + conj Z[-k]) / 2
      tile[j * TS + 2 * gg + 1] = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));  // / 2i
    }
  }
  __syncthreads();
  const int rb = 2 * blockIdx.x * G;
  float2* Tb = T + (size_t)img * wc * ny;
  for (int e = threadIdx.x; e < wc * 2 * G; e += kBlock) {
    const int j = e / (2 * G), rr = e - j * 2 * G;
    if (rb + rr < ny) Tb[(size_t)j * ny + rb + rr] = tile[j * TS + rr];
  }
}

// Row kernel for a window inside [-H T, H T] (T = P::T, H = 1 or 2; every
// image plan has U2 == 1).  After stage 1 and the exchange, thread t only
// needs outputs 0 .. H and V - H .. V - 1 of its radix-V stage-2 DFT, i.e.
// kappa = t + T m (m < H), H T and t - T (m + 1).  For H = 1 these are X[t],
// X[t + T] and X[N - T + t].  They are formed directly (one sum and H pairs of
// twiddled sums) instead of the full transform, and the Hermitian partner
// Z[-kappa] of the two packed rows comes from lane (T - t) mod T by a
// cross-lane read, so the outputs never go back through LDS.  A block walks
// kRowSets sets of 2G rows and issues the next set's loads before
// transforming the current one, so the mask reads of one set overlap the
// arithmetic of the other (the exchange scratch allows two blocks per CU).
constexpr int kRowSets = 2;

// Rows r0 and r0 + 1 (the caller guarantees both exist).  The mask is read
// once per pass: non-temporal loads keep it from displacing the pass's
// intermediates in L2 and the last-level cache, and from waiting on the
// write-back of dirty lines the previous pass's image left there.
template <class P>
__device__ __forceinline__ void rows_load_pair(float2 (&v)[P::V], const float* __restrict__ src,
                                               int r0, int t) {
  const float* a = src + (size_t)r0 * P::N + t;
#pragma unroll
  for (int m = 0; m < P::V; ++m)
    v[m] = make_float2(__builtin_nontemporal_load(a + P::T * m),
                       __builtin_nontemporal_load(a + P::N + P::T * m));
}

__device__ __forceinline__ void rows_separate(float2 zk, float2 zm, float2* dst) {
  dst[0] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));     // (Z[k] + conj Z[-k]) / 2
  dst[1] = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));    // (Z[k] - conj Z[-k]) / 2i
}

// Inverse transform (e^{+2 pi i k x / N}) of v, keeping the stage-2 outputs
// m = 0 .. H and V - H .. V - 1 of the group's thread t (X[t + T m]): stage 1
// as GFFT::run_gen, the exchange, then those outputs of the stage-2 DFT only,
// sum_j b_j e^{+2 pi i j m / V}, the pair (m, V - m) from one set of sums.
//   pos[m] = X[t + T m]        (m = 0 .. H)
//   neg[i] = X[t + T (V-1-i)]  (i = 0 .. H-1), i.e. kappa = t - T (i + 1)
// tw2: the stage-2 twiddles, w_N^{j t} at tw2[j R1 + t] (the second half of
// a twiddle table, or a copy of it in LDS).
// Every thread of the block calls (barrier).
// cos and sin of 2 pi j (p + 1) / V for H > 1 output pairs, a constexpr
// object so that no series is left for the device to evaluate (for H = 1 the
// inline root_re / root_im fold to immediates on their own)
template <int V, int H>
struct NarrowRoots {
  float c[V][H], s[V][H];
  constexpr NarrowRoots() : c(), s() {
    for (int j = 0; j < V; ++j)
      for (int p = 0; p < H; ++p) {
        c[j][p] = (float)reg::root_re((long)j * (p + 1), V);
        s[j][p] = (float)(-reg::root_im((long)j * (p + 1), V));
      }
  }
};

template <class P, int H>
__device__ __forceinline__ void narrow_inverse(const float2 (&v)[P::V], float2* scr, int t,
                                               const float2* __restrict__ tw2,
                                               float2 (&pos)[H + 1], float2 (&neg)[H]) {
  constexpr int TT = P::T, V = P::V, R1 = P::R1, U1 = P::U1;
  static_assert(P::U2 == 1 && P::R2 == V && TT == R1, "radix-V second stage");
  static_assert(H >= 1 && 2 * H < V, "output pairs inside one transform");
#pragma unroll
  for (int u = 0; u < U1; ++u) {
    float2 b[R1];
#pragma unroll
    for (int j = 0; j < R1; ++j) b[j] = v[u + j * U1];
    reg::DFT<R1, true>::run(b);
    const int i = t + TT * u;
#pragma unroll
    for (int j = 0; j < R1; ++j) scr[P::xi(i * R1 + j)] = b[j];
  }
  __syncthreads();
  float sx = 0.f, sy = 0.f, A[H], B[H], C[H], D[H];
#pragma unroll
  for (int p = 0; p < H; ++p) A[p] = B[p] = C[p] = D[p] = 0.f;
#pragma unroll
  for (int j = 0; j < V; ++j) {
    float2 b = scr[P::xi(t + j * R1)];
    if (j > 0) {
      const float2 w = tw2[j * R1 + t];             // exp(-2 pi i j t / N), used conjugated
      b = reg::mul(b, make_float2(w.x, -w.y));
    }
    sx += b.x;
    sy += b.y;
#pragma unroll
    for (int p = 0; p < H; ++p) {
      float c, s;
      if constexpr (H == 1) {
        c = (float)reg::root_re(j, V);
        s = (float)(-reg::root_im(j, V));
      } else {
        constexpr NarrowRoots<V, H> kRoots{};
        c = kRoots.c[j][p];
        s = kRoots.s[j][p];
      }
      A[p] = fmaf(b.x, c, A[p]);
      B[p] = fmaf(b.y, s, B[p]);
      C[p] = fmaf(b.x, s, C[p]);
      D[p] = fmaf(b.y, c, D[p]);
    }
  }
  pos[0] = make_float2(sx, sy);
#pragma unroll
  for (int p = 0; p < H; ++p) {
    pos[p + 1] = make_float2(A[p] - B[p], C[p] + D[p]);   // m = p + 1
    neg[p] = make_float2(A[p] + B[p], D[p] - C[p]);       // m = V - 1 - p
  }
}

// Orders kappa in [-H T, H T] of two real rows per transform.  The Hermitian
// partner Z[-kappa] of kappa = t + T m (t > 0) is output V - 1 - m of lane
// T - t, read across lanes; for t = 0 it is output V - m of the same lane.
// LDS: the exchange scratch of the G groups, which the output tile [wc][2G+1]
// reuses once every group is past its stage-2 reads, then the N stage-2
// twiddles, copied in once per block (rows_narrow_lds).
template <class P>
constexpr size_t rows_narrow_scratch(int wc) {
  return (size_t)groups_per_block<P>() * P::LDS > (size_t)wc * (2 * groups_per_block<P>() + 1)
             ? (size_t)groups_per_block<P>() * P::LDS
             : (size_t)wc * (2 * groups_per_block<P>() + 1);
}
template <class P>
constexpr size_t rows_narrow_lds(int wc) {
  return (rows_narrow_scratch<P>(wc) + P::N) * sizeof(float2);
}

template <class P, int H>
__global__ __launch_bounds__(kBlock) void k_fp_rows_real_narrow(const float* __restrict__ mask,
                                                                int ny, int k0c, int wc,
                                                                const float2* __restrict__ twt,
                                                                float2* __restrict__ T, int logits) {
  constexpr int N = P::N, TT = P::T, V = P::V, R1 = P::R1, R2 = P::R2, U1 = P::U1;
  constexpr int G = groups_per_block<P>();
  static_assert(P::U2 == 1 && R2 == V && TT == R1 && kBlock % TT == 0 && (TT & (TT - 1)) == 0 &&
                    TT <= 64,
                "narrow rows: radix-V second stage, power-of-two groups inside a wavefront");
  // ny is a multiple of 2G (checked by the launch): every set is whole.
  extern __shared__ float2 lds[];
  const int g = threadIdx.x / TT, t = threadIdx.x - g * TT;
  float2* scr = lds + g * P::LDS;
  constexpr int TS = 2 * G + 1;
  float2* tile = lds;                              // [wc][TS], over the scratch
  float2* tw2 = lds + rows_narrow_scratch<P>(wc);  // stage-2 twiddles [V][R1]
  for (int e = threadIdx.x; e < N; e += kBlock) tw2[e] = twt[N + e];
  const int img = blockIdx.y;
  const float* src = mask + (size_t)img * ny * N;
  float2* Tb = T + (size_t)img * wc * ny;
  const int lane = threadIdx.x & 63;
  const int partner = (lane & ~(TT - 1)) | ((TT - t) & (TT - 1));
  const int sets = (ny + 2 * G - 1) / (2 * G);
  float2 nxt[V];
  rows_load_pair<P>(nxt, src, 2 * (blockIdx.x * kRowSets * G + g), t);
  for (int it = 0; it < kRowSets; ++it) {
    const int set = blockIdx.x * kRowSets + it;
    if (set >= sets) break;                        // uniform over the block
    float2 v[V];
    if (logits) {
#pragma unroll
      for (int m = 0; m < V; ++m) v[m] = make_float2(logistic(nxt[m].x), logistic(nxt[m].y));
    } else {
#pragma unroll
      for (int m = 0; m < V; ++m) v[m] = nxt[m];
    }
    if (it + 1 < kRowSets && set + 1 < sets) rows_load_pair<P>(nxt, src, 2 * ((set + 1) * G + g), t);
    float2 pos[H + 1], neg[H];
    narrow_inverse<P, H>(v, scr, t, tw2, pos, neg);   // its first barrier also orders tw2
    __syncthreads();                               // every group is past its scratch reads
#pragma unroll
    for (int m = 0; m < H; ++m) {
      // kappa = t + T m; partner: lane T - t's neg[m], or (t = 0) pos[0] / neg[m - 1]
      const float2 pn = make_float2(__shfl(neg[m].x, partner), __shfl(neg[m].y, partner));
      const int j = t + TT * m - k0c;
      if (j >= 0 && j < wc)
        rows_separate(pos[m], t == 0 ? (m == 0 ? pos[0] : neg[m - 1]) : pn, tile + j * TS + 2 * g);
      // kappa = t - T (m + 1); partner: lane T - t's pos[m], or (t = 0) pos[m + 1]
      const float2 pp = make_float2(__shfl(pos[m].x, partner), __shfl(pos[m].y, partner));
      const int jn = t - TT * (m + 1) - k0c;
      if (jn >= 0 && jn < wc) rows_separate(neg[m], t == 0 ? pos[m + 1] : pp, tile + jn * TS + 2 * g);
    }
    const int jh = TT * H - k0c;                   // kappa = H T (t = 0); partner neg[H - 1]
    if (t == 0 && jh >= 0 && jh < wc) rows_separate(pos[H], neg[H - 1], tile + jh * TS + 2 * g);
    __syncthreads();
    const int rb = 2 * set * G;
    for (int e = threadIdx.x; e < wc * 2 * G; e += kBlock) {
      const int j = e / (2 * G), rr = e - j * 2 * G;
      Tb[(size_t)j * ny + rb + rr] = tile[j * TS + rr];
    }
    __syncthreads();                               // the tile (scratch) is reused
  }
}

template <class P>
__global__ __launch_bounds__(kBlock) void k_fp_rows_c64(const float2* __restrict__ field, int cin,
                                                        Chans ch, int nchan, int ny, int k0c,
                                                        int wc, const float2* __restrict__ twt,
                                                        float2* __restrict__ T) {
  constexpr int N = P::N, TT = P::T, V = P::V, G = groups_per_block<P>();
  extern __shared__ float2 lds[];
  const int g = threadIdx.x / TT, t = threadIdx.x - g * TT;
  const bool live = g < G;
  const int gg = live ? g : 0;
  float2* scr = lds + g * P::LDS;
  const int bc = blockIdx.y;
  const int img = bc / nchan, ci = bc - img * nchan;
  const int chan = ch.c[ci];
  const int r = blockIdx.x * G + gg;
  float2 v[V];
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const int x = t + TT * m;
    v[m] = (live && r < ny) ? field[(((size_t)img * ny + r) * N + x) * cin + chan] : make_float2(0.f, 0.f);
  }
  P::template run_live_tab<true>(v, scr, t, twt, true);
  if (live) {
#pragma unroll
    for (int m = 0; m < V; ++m) scr[pad(t + TT * m)] = v[m];
  }
  __syncthreads();
  constexpr int TS = G + 1;
  float2* tile = lds + scratch_groups<P>() * P::LDS;   // [wc][TS]
  if (live && r < ny)
    for (int j = t; j < wc; j += TT) tile[j * TS + gg] = scr[pad(pmod(k0c + j, N))];
  __syncthreads();
  const int rb = blockIdx.x * G;
  float2* Tb = T + (size_t)bc * wc * ny;
  for (int e = threadIdx.x; e < wc * G; e += kBlock) {
    const int j = e / G, rr = e - j * G;
    if (rb + rr < ny) Tb[(size_t)j * ny + rb + rr] = tile[j * TS + rr];
  }
}

// --------------------------------------------------------------- columns
template <class P>
__global__ __launch_bounds__(kBlock) void k_fp_cols(const float2* __restrict__ T, int k0r, int wr,
                                                    int wc, float scale,
                                                    const float2* __restrict__ twt,
                                                    float2* __restrict__ spec) {
  constexpr int N = P::N, TT = P::T, V = P::V, G = groups_per_block<P>();
  extern __shared__ float2 lds[];
  const int g = threadIdx.x / TT, t = threadIdx.x - g * TT;
  const bool live = g < G;
  const int gg = live ? g : 0;
  const int bc = blockIdx.y;
  const int j = blockIdx.x * G + gg;
  const bool col = live && j < wc;
  const float2* src = T + ((size_t)bc * wc + (col ? j : 0)) * N;   // column j, contiguous rows
  float2 v[V];
#pragma unroll
  for (int m = 0; m < V; ++m) v[m] = col ? src[t + TT * m] : make_float2(0.f, 0.f);
  P::template run_live_tab<true>(v, lds + g * P::LDS, t, twt, true);
  if (!col) return;
  float2* dst = spec + (size_t)bc * wr * wc;
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const int k = t + TT * m;
    const int kappa = k <= N - 1 - N / 2 ? k : k - N;
    const int i = kappa - k0r;
    if (i >= 0 && i < wr) dst[(size_t)i * wc + j] = make_float2(v[m].x * scale, v[m].y * scale);
  }
}

// Column kernel for a window inside [-H T, H T] (the spectrum rows k0r ..
// k0r + wr - 1): only outputs kappa = t + T m (m < H), t - T (m + 1) and H T
// are formed.
template <class P, int H>
__global__ __launch_bounds__(kBlock) void k_fp_cols_narrow(const float2* __restrict__ T, int k0r,
                                                           int wr, int wc, float scale,
                                                           const float2* __restrict__ twt,
                                                           float2* __restrict__ spec) {
  constexpr int N = P::N, TT = P::T, V = P::V, G = groups_per_block<P>();
  static_assert(kBlock % TT == 0, "whole groups");
  extern __shared__ float2 lds[];
  const int g = threadIdx.x / TT, t = threadIdx.x - g * TT;
  const int bc = blockIdx.y;
  const int j = blockIdx.x * G + g;
  const bool col = j < wc;
  const float2* src = T + ((size_t)bc * wc + (col ? j : 0)) * N;   // column j, contiguous rows
  float2 v[V];
#pragma unroll
  for (int m = 0; m < V; ++m) v[m] = col ? src[t + TT * m] : make_float2(0.f, 0.f);
  float2 pos[H + 1], neg[H];
  narrow_inverse<P, H>(v, lds + g * P::LDS, t, twt + N, pos, neg);
  if (!col) return;
  float2* dst = spec + (size_t)bc * wr * wc + j;
#pragma unroll
  for (int m = 0; m < H; ++m) {
    const int ia = t + TT * m - k0r, ib = t - TT * (m + 1) - k0r;
    if (ia >= 0 && ia < wr) dst[(size_t)ia * wc] = make_float2(pos[m].x * scale, pos[m].y * scale);
    if (ib >= 0 && ib < wr) dst[(size_t)ib * wc] = make_float2(neg[m].x * scale, neg[m].y * scale);
  }
  const int ic = TT * H - k0r;
  if (t == 0 && ic >= 0 && ic < wr) dst[(size_t)ic * wc] = make_float2(pos[H].x * scale, pos[H].y * scale);
}

// -------------------------------------------------------------- coherent
// Block (s, img, g): coarse rows t_r = s + P a (a < M = L/P), kernels g, g+G, ...
// LDS: spectrum window of the image (one channel at a time), B[a][kc] (M x wc),
// transform scratch.
template <class PL, class PM>
__global__ __launch_bounds__(kBlock) void k_fp_coherent(FpCoh A) {
  constexpr int L = PL::N, M = PM::N, P = L / M;
  constexpr int TL = PL::T, VL = PL::V, GL = groups_per_block<PL>();
  constexpr int TM = PM::T, VM = PM::V, GM = groups_per_block<PM>();
  constexpr int ROWS = (M + GL - 1) / GL;  // coarse rows per row group
  extern __shared__ float2 lds[];
  const int wr = A.wr, wc = A.wc;
  float2* specL = lds;
  float2* Bb = specL + wr * wc;
  float2* scr = Bb + M * wc;
  const int s = blockIdx.x, img = blockIdx.y, grp = blockIdx.z;
  const int tid = threadIdx.x;

  // row-stage geometry
  const int gl = tid / TL, tl = tid - gl * TL;
  const bool llive = gl < GL;
  const int gll = llive ? gl : 0;
  typename PL::Tw twl;
  PL::load_tw(twl, tl, A.tw_l);
  // column-stage geometry
  const int gm = tid / TM, tm = tid - gm * TM;
  const bool mlive = gm < GM;
  const int gmm = mlive ? gm : 0;
  typename PM::Tw twm;
  if constexpr (!std::is_same<PL, PM>::value) PM::load_tw(twm, tm, A.tw_m);

  float acc[ROWS][VL];
#pragma unroll
  for (int q = 0; q < ROWS; ++q)
#pragma unroll
    for (int m = 0; m < VL; ++m) acc[q][m] = 0.f;

  for (int c = 0; c < A.nchan; ++c) {
    const float2* sp = A.spec + ((size_t)img * A.nchan + c) * wr * wc;
    __syncthreads();
    for (int e = tid; e < wr * wc; e += kBlock) specL[e] = sp[e];
    __syncthreads();
    for (int kk = grp; kk < A.nkern; kk += A.groups) {
      const float2* kp = A.kern + (size_t)kk * wr * wc;
      // column stage: column kc handled by group gm (round-robin); M-point
      // transform of the kr fold (kr mod M) twisted by exp(-2 pi i kr s / L)
      for (int c0 = 0; c0 < wc; c0 += GM) {
        const int col = c0 + gmm;
        const bool on = mlive && col < wc;
        float2 v[VM];
#pragma unroll
        for (int m = 0; m < VM; ++m) {
          const int u = tm + TM * m;
          float2 sum = make_float2(0.f, 0.f);
          if (on) {
            for (int ir = pmod(u - A.k0r, M); ir < wr; ir += M) {
              const int kr = A.k0r + ir;
              float2 a = reg::mul(kp[ir * wc + col], specL[ir * wc + col]);
              if (P > 1) a = reg::mul(a, A.tw_l[pmod(kr * s, L)]);
              sum.x += a.x;
              sum.y += a.y;
            }
          }
          v[m] = sum;
        }
        if constexpr (std::is_same<PL, PM>::value)
          PM::template run<false>(v, scr + gm * PM::LDS, tm, twl);
        else
          PM::template run<false>(v, scr + gm * PM::LDS, tm, twm);
        if (on) {
#pragma unroll
          for (int m = 0; m < VM; ++m) Bb[(tm + TM * m) * wc + col] = v[m];
        }
      }
      __syncthreads();
      // row stage: row a handled by group gl (a = gl + GL q); L-point transform
      // of the kc fold (kc mod L); |.|^2 accumulated
#pragma unroll
      for (int q = 0; q < ROWS; ++q) {
        const int a = gll + GL * q;
        const bool on = llive && a < M;
        float2 v[VL];
#pragma unroll
        for (int m = 0; m < VL; ++m) {
          const int x = tl + TL * m;
          float2 sum = make_float2(0.f, 0.f);
          if (on)
            for (int ic = pmod(x - A.k0c, L); ic < wc; ic += L) {
              const float2 b = Bb[a * wc + ic];
              sum.x += b.x;
              sum.y += b.y;
            }
          v[m] = sum;
        }
        PL::template run<false>(v, scr + gl * PL::LDS, tl, twl);
#pragma unroll
        for (int m = 0; m < VL; ++m) acc[q][m] += v[m].x * v[m].x + v[m].y * v[m].y;
      }
      __syncthreads();
    }
  }
  // epilogue: row DFT (+ sign) of the coarse rows, orders q in [0, qc] -> R
  const int nq = A.qc + 1;
#pragma unroll
  for (int q = 0; q < ROWS; ++q) {
    const int a = gll + GL * q;
    const bool on = llive && a < M;
    float2 v[VL];
#pragma unroll
    for (int m = 0; m < VL; ++m) v[m] = make_float2(acc[q][m], 0.f);
    PL::template run<true>(v, scr + gl * PL::LDS, tl, twl);
    if (on) {
      float2* dst = A.R + ((size_t)img * A.groups + grp) * nq * L + (s + P * a);
#pragma unroll
      for (int m = 0; m < VL; ++m) {
        const int x = tl + TL * m;
        if (x < nq) dst[(size_t)x * L] = v[m];
      }
    }
  }
}

// Coherent sum, one workgroup per (image, kernel group), every coarse row and
// every window column in one pass.  Groups of T_L lanes never straddle a
// wavefront (GW = 64 / T_L groups per wave, NT / 64 waves, NT/64 * GW >= L), so
// each transform exchanges with wave-level ordering only.  All transforms work
// in place in one LDS tile [L][S]: the column stage of window column kc in tile
// column (k0c + kc) mod L, the row stage of coarse row a in tile row a.  Two
// workgroup barriers per kernel: after the column stage (rows read every
// column) and after the row stage (the next column stage writes across rows).
// The image's spectrum window sits in LDS per channel; each thread's kernel
// values are loaded one kernel ahead.  Requires wr, wc <= L.
template <int L>
constexpr int coh_stride() {
  // row stride of the tile in float2; 89 for L = 75 spreads the in-place
  // transform accesses over the 64-bit banks (fewest conflicts in a model of
  // the access pattern), padded L otherwise
  return L == 75 ? 89 : reg::padded(L);
}

template <class PL>
size_t coh_lds_bytes(int wr, int wc) {
  return ((size_t)PL::N * coh_stride<PL::N>() + (size_t)wr * wc) * sizeof(float2);
}

template <class PL, int NT>
__global__ __launch_bounds__(NT) void k_fp_coh(FpCoh A) {
  constexpr int L = PL::N, TL = PL::T, VL = PL::V;
  constexpr int S = coh_stride<L>();
  constexpr int GW = 64 / TL;
  static_assert(GW >= 1 && (NT / 64) * GW >= L, "one group per coarse row");
  static_assert(VL <= 32, "row mask holds VL bits");
  extern __shared__ float2 lds[];  // tile [L][S], then the spectrum window [wr][wc]
  const int wr = A.wr, wc = A.wc;
  float2* specL = lds + L * S;
  const int img = blockIdx.x, grp = blockIdx.y;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gw = lane / TL, t = lane - gw * TL;
  const int g = wave * GW + gw;
  const bool lane_ok = gw < GW;
  const bool live_r = lane_ok && g < L;   // coarse row g
  const bool live_c = lane_ok && g < wc;  // window column g
  typename PL::Tw tw;
  PL::load_tw(tw, t, A.tw_l);
  const int kr0 = pmod(t - A.k0r, L);     // TL m < L: kr = kr0 + TL m, less L once
  unsigned cmask = 0;                     // column-stage values inside the window
#pragma unroll
  for (int m = 0; m < VL; ++m) {
    int kr = kr0 + TL * m;
    kr = kr >= L ? kr - L : kr;
    if (live_c && kr < wr) cmask |= 1u << m;
  }
  const int xcol = live_c ? pmod(A.k0c + g, L) : 0;
  unsigned rmask = 0;
#pragma unroll
  for (int m = 0; m < VL; ++m)
    if (live_r && pmod(t + TL * m - A.k0c, L) < wc) rmask |= 1u << m;
  float2* row = lds + (live_r ? g : 0) * S;
  const int gc = live_c ? g : 0;

  float acc[VL];
#pragma unroll
  for (int m = 0; m < VL; ++m) acc[m] = 0.f;

  for (int c = 0; c < A.nchan; ++c) {
    const float2* sp = A.spec + ((size_t)img * A.nchan + c) * wr * wc;
    __syncthreads();
    for (int e = threadIdx.x; e < wr * wc; e += NT) specL[e] = sp[e];
    // this thread's kernel values of the first kernel
    float2 nxt[VL];
    {
      const float2* kp = A.kern + (size_t)grp * wr * wc;
#pragma unroll
      for (int m = 0; m < VL; ++m) {
        int kr = kr0 + TL * m;
        kr = kr >= L ? kr - L : kr;
        nxt[m] = (grp < A.nkern && ((cmask >> m) & 1u)) ? kp[kr * wc + gc] : make_float2(0.f, 0.f);
      }
    }
    __syncthreads();
    for (int kk = grp; kk < A.nkern; kk += A.groups) {
      float2 v[VL];
#pragma unroll
      for (int m = 0; m < VL; ++m) {
        int kr = kr0 + TL * m;
        kr = kr >= L ? kr - L : kr;
        v[m] = ((cmask >> m) & 1u) ? reg::mul(nxt[m], specL[kr * wc + gc]) : make_float2(0.f, 0.f);
      }
      // next kernel's values, in flight during this kernel's transforms
      const int kn = kk + A.groups;
      const float2* kp = A.kern + (size_t)(kn < A.nkern ? kn : kk) * wr * wc;
#pragma unroll
      for (int m = 0; m < VL; ++m) {
        int kr = kr0 + TL * m;
        kr = kr >= L ? kr - L : kr;
        nxt[m] = ((cmask >> m) & 1u) ? kp[kr * wc + gc] : make_float2(0.f, 0.f);
      }
      PL::template run_strided_wave<false>(v, lds + xcol, S, t, tw, live_c);
      if (live_c) {
#pragma unroll
        for (int m = 0; m < VL; ++m) lds[(t + TL * m) * S + xcol] = v[m];
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < VL; ++m) v[m] = ((rmask >> m) & 1u) ? row[t + TL * m] : make_float2(0.f, 0.f);
      reg::wave_sync();  // the group's loads precede its in-place exchange
      PL::template run_live_wave<false>(v, row, t, tw, live_r);
#pragma unroll
      for (int m = 0; m < VL; ++m) acc[m] += v[m].x * v[m].x + v[m].y * v[m].y;
      __syncthreads();
    }
  }
  // epilogue: row DFT (+ sign) of coarse row g, orders q in [0, qc] -> R[img][grp][q][g]
  const int nq = A.qc + 1;
  float2 v[VL];
#pragma unroll
  for (int m = 0; m < VL; ++m) v[m] = make_float2(acc[m], 0.f);
  PL::template run_live_wave<true>(v, row, t, tw, live_r);
  if (live_r) {
    float2* dst = A.R + ((size_t)img * A.groups + grp) * nq * L + g;
#pragma unroll
    for (int m = 0; m < VL; ++m) {
      const int x = t + TL * m;
      if (x < nq) dst[(size_t)x * L] = v[m];
    }
  }
}

// ------------------------------------------------ two-pass coherent sum
// For coarse grids too large for one workgroup (L = 256, 512).  Pass 1
// transforms every window column of every (kernel, channel) into
// B[img][kc][c][t] (column-major, t the coarse row); pass 2 gives each block G
// coarse rows, loops over all (kernel, channel) pairs reading its rows of B
// through an LDS tile, and accumulates |E|^2 in registers; its epilogue takes
// the row DFT (+ sign) to R.  Stage twiddles are read from the table at use.
template <class PL>
__global__ __launch_bounds__(kBlock) void k_fp_bcols(const float2* __restrict__ spec,
                                                     const float2* __restrict__ kern, int nk,
                                                     int nchan, int k0r, int wr, int wc,
                                                     const float2* __restrict__ twt,
                                                     float2* __restrict__ B) {
  constexpr int L = PL::N, TL = PL::T, VL = PL::V, G = groups_per_block<PL>();
  extern __shared__ float2 lds[];
  const int g = threadIdx.x / TL, t = threadIdx.x - g * TL;
  const int c = blockIdx.x * G + g;
  const int kc = blockIdx.y, img = blockIdx.z;        // kc = k * nchan + channel
  const int k = kc / nchan, ch = kc - k * nchan;
  const bool live = g < G && c < wc;
  const float2* kp = kern + (size_t)k * wr * wc;
  const float2* sp = spec + ((size_t)img * nchan + ch) * wr * wc;
  const int kr0 = pmod(t - k0r, L);
  float2 v[VL];
#pragma unroll
  for (int m = 0; m < VL; ++m) {
    int kr = kr0 + TL * m;
    kr = kr >= L ? kr - L : kr;
    // rounded product: the adjoint's k_adj_cfield_reg forms the same values
    v[m] = (live && kr < wr) ? reg::mul_rn(kp[kr * wc + c], sp[kr * wc + c]) : make_float2(0.f, 0.f);
  }
  PL::template run_live_tab<false>(v, lds + g * PL::LDS, t, twt, live);
  if (!live) return;
  float2* dst = B + (((size_t)img * nk * nchan + kc) * wc + c) * L;
#pragma unroll
  for (int m = 0; m < VL; ++m) dst[t + TL * m] = v[m];
}

// Blocks (row group, img, grp) sum the (kernel, channel) pairs kc = grp,
// grp + groups, ... into their own partial R[img][grp] (groups = 1 for the
// generic layout).
// Four waves per SIMD (at most 128 registers) up to coarse 256: four
// workgroups then share a CU, as their 34 KB of LDS allow (ILT step 0.355 ->
// 0.348 ms).  From 450 on the LDS (about 70 KB) holds two workgroups per CU
// whatever the registers, and the 30- and 32-value transforms would spill
// under 128: two waves per SIMD (up to 256 registers).
template <class PL>
constexpr int brows_min_blocks() {
  return PL::N > 256 ? 2 : 4;
}

template <class PL>
__global__ __launch_bounds__(kBlock, brows_min_blocks<PL>()) void k_fp_brows(const float2* __restrict__ B, int nkc, int k0c,
                                                     int wc, int qc, int generic_layout,
                                                     const float2* __restrict__ twt,
                                                     float2* __restrict__ R, int batch,
                                                     int groups) {
  constexpr int L = PL::N, TL = PL::T, VL = PL::V, G = groups_per_block<PL>();
  constexpr int TS = G + 1;                // tile row stride (bank spread)
  static_assert(VL <= 32, "row mask holds VL bits");
  extern __shared__ float2 lds[];
  // items (row group, image, kernel group), row groups fastest, in XCD order:
  // a column of B is split over the row groups of one (image, kernel group),
  // and where 8 G bytes do not divide the line, the lines at the edges are
  // read by two of them
  constexpr int NXR = (L + G - 1) / G;
  int w;
  if (!xcd_item(blockIdx.x, NXR * batch * groups, w)) return;   // uniform over the block
  const int xr = w % NXR, img = (w / NXR) % batch, grp = w / (NXR * batch);
  float2* tile = lds;                      // [wc][TS]
  float2* scr = lds;                       // transform scratch, over the tile once it is read
  const int g = threadIdx.x / TL, t = threadIdx.x - g * TL;
  const int r0 = xr * G;
  const int row = r0 + g;
  const bool live = g < G && row < L;
  unsigned rmask = 0;
  int cidx[VL];
#pragma unroll
  for (int m = 0; m < VL; ++m) {
    const int cc = pmod(t + TL * m - k0c, L);
    cidx[m] = cc < wc ? cc : 0;
    if (live && cc < wc) rmask |= 1u << m;
  }
  float acc[VL];
#pragma unroll
  for (int m = 0; m < VL; ++m) acc[m] = 0.f;
  for (int kc = grp; kc < nkc; kc += groups) {
    const float2* src = B + ((size_t)img * nkc + kc) * wc * L + r0;
    __syncthreads();
    for (int e = threadIdx.x; e < wc * G; e += kBlock) {
      const int cc = e / G, r = e - cc * G;
      tile[cc * TS + r] = (r0 + r < L) ? src[(size_t)cc * L + r] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    float2 v[VL];
#pragma unroll
    for (int m = 0; m < VL; ++m)
      v[m] = ((rmask >> m) & 1u) ? tile[cidx[m] * TS + (g < G ? g : 0)] : make_float2(0.f, 0.f);
    __syncthreads();                       // every row is read: the scratch reuses the tile
    PL::template run_live_tab<false>(v, scr + g * PL::LDS, t, twt, g < G);
#pragma unroll
    for (int m = 0; m < VL; ++m) acc[m] += v[m].x * v[m].x + v[m].y * v[m].y;
  }
  float2 v[VL];
#pragma unroll
  for (int m = 0; m < VL; ++m) v[m] = make_float2(acc[m], 0.f);
  PL::template run_live_tab<true>(v, scr + g * PL::LDS, t, twt, g < G);
  if (!live) return;
  const int nq = qc + 1;
#pragma unroll
  for (int m = 0; m < VL; ++m) {
    const int x = t + TL * m;
    if (x > qc) continue;
    if (generic_layout) R[((size_t)img * L + row) * nq + x] = v[m];   // [img][t][q] (one group)
    else R[(((size_t)img * groups + grp) * nq + x) * L + row] = v[m];  // [img][grp][q][t]
  }
}

// ------------------------------------------------------- band extraction
// Ihat[img][q][i] = scale * sum_t (sum_g R[img][g][q][t]) e^{+2 pi i qr t / L},
// qr = i - qr_max, i in [0, 2 qr_max + 1).
// One order q per workgroup of kBandWide threads.  The sum over the kernel
// groups (up to 256 for a batch of one image) is spread over the threads:
// S = kBandWide / L slices of the groups, each slice summed in order, then the
// slices in order, so that it never runs as `groups` dependent loads in a
// handful of threads; then one group of threads transforms the column.
constexpr int kBandWide = 256;

template <class PL>
__global__ __launch_bounds__(kBandWide) void k_fp_band(const float2* __restrict__ R, int groups,
                                                       int qc, int qrm, float scale,
                                                       const float2* __restrict__ twt,
                                                       float2* __restrict__ Ihat) {
  constexpr int L = PL::N, TL = PL::T, VL = PL::V;
  constexpr int S = kBandWide / L > 0 ? kBandWide / L : 1;   // group slices
  static_assert(TL <= kBandWide, "one transform group");
  extern __shared__ float2 lds[];
  float2* part = lds;                    // [S][L]
  float2* sum = lds + S * L;             // [L]
  float2* scr = sum + L;                 // the transform's exchange
  const int img = blockIdx.y, q = blockIdx.x;
  const int nq = qc + 1;
  const float2* src = R + ((size_t)img * groups * nq + q) * L;
  const size_t gstride = (size_t)nq * L;
  for (int i = threadIdx.x; i < S * L; i += kBandWide) {
    const int gs = i / L, t = i - gs * L;
    float2 a = make_float2(0.f, 0.f);
    int k = gs;
    // eight loads in flight, then their sums in the same order as one at a
    // time: a slice of 64 groups (256 groups of a one-image plan) is not a
    // chain of 64 dependent load latencies
    for (; k + 7 * S < groups; k += 8 * S) {
      float2 x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = src[(size_t)(k + u * S) * gstride + t];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        a.x += x[u].x;
        a.y += x[u].y;
      }
    }
    for (; k < groups; k += S) {
      const float2 x = src[(size_t)k * gstride + t];
      a.x += x.x;
      a.y += x.y;
    }
    part[i] = a;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < L; t += kBandWide) {
    float2 a = part[t];
    for (int gs = 1; gs < S; ++gs) {
      a.x += part[gs * L + t].x;
      a.y += part[gs * L + t].y;
    }
    sum[t] = a;
  }
  __syncthreads();
  const int t = threadIdx.x;
  const bool live = t < TL;
  float2 v[VL];
#pragma unroll
  for (int m = 0; m < VL; ++m) v[m] = live ? sum[t + TL * m] : make_float2(0.f, 0.f);
  PL::template run_live_tab<true>(v, scr, live ? t : 0, twt, live);
  if (!live) return;
  const int nb = 2 * qrm + 1;
  float2* dst = Ihat + ((size_t)img * nq + q) * nb;
#pragma unroll
  for (int m = 0; m < VL; ++m) {
    const int x = t + TL * m;
    const int qr = x <= qrm ? x : x - L;
    if (qr >= -qrm && qr <= qrm) dst[qr + qrm] = make_float2(v[m].x * scale, v[m].y * scale);
  }
}

// U[img][q][y] = sum_qr Ihat[img][q][qr] e^{-2 pi i qr y / ny}
// Q > 0: qrm < Q T, so only registers m < Q and m >= V - Q can hold band
// values (as in k_fp_urows).
template <class P, int Q>
__global__ __launch_bounds__(kBlock) void k_fp_ucols(const float2* __restrict__ Ihat, int qc, int qrm,
                                                     const float2* __restrict__ twt,
                                                     float2* __restrict__ U) {
  constexpr int N = P::N, TT = P::T, V = P::V, G = groups_per_block<P>();
  extern __shared__ float2 lds[];
  const int g = threadIdx.x / TT, t = threadIdx.x - g * TT;
  const bool live = g < G;
  const int gg = live ? g : 0;
  const int img = blockIdx.y;
  const int q = blockIdx.x * G + gg;
  const int nq = qc + 1;
  const bool on = live && q < nq;
  const float2* src = Ihat + ((size_t)img * nq + (on ? q : 0)) * (2 * qrm + 1);
  static_assert(Q >= 0 && 2 * Q <= V, "band registers");
  float2 v[V];
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const int y = t + TT * m;
    const int qr = y <= qrm ? y : y - N;
    const bool may = Q == 0 || m < Q || m >= V - Q;      // compile-time for Q > 0
    v[m] = (may && on && qr >= -qrm && qr <= qrm) ? src[qr + qrm] : make_float2(0.f, 0.f);
  }
  P::template run_live_tab<false>(v, lds + g * P::LDS, t, twt, true);
  if (!on) return;
  float2* dst = U + ((size_t)img * nq + q) * N;
#pragma unroll
  for (int m = 0; m < V; ++m) dst[t + TT * m] = v[m];
}

// Image rows from U (Hermitian extension over q), two rows per transform.
// Q > 0: the band has qc < Q T, so thread t's inputs x = t + T m can be
// non-zero only for m < Q or m >= V - Q; the other registers are the constant
// zero and the first-stage arithmetic on them folds away at compile time.
// MODE 1 (residual): instead of the image I, write 2 (I - target) and the
// workgroup's sum of (I - target)^2 to part[img * gridDim.x + blockIdx.x]
// (the residual of a least-squares fit to target, fused into the pass that
// produces I).  MODE 2 (chain rule): target holds logits theta; write
// I m (1 - m), m = logistic(theta), and the workgroup's sum of its squares
// (the gradient with respect to theta of a loss whose gradient with respect
// to m is I, and its norm).  MODE 0: the image.
//
// A block images rsets sets of 2G rows in turn (fp_urows_sets: up to
// kURowSets while the grid keeps at least kURowMinBlocks workgroups, so
// small batches keep their parallelism).  For Q > 0 the band tile of the
// next set (at most Q T x 2G values) is loaded into registers while the
// current set is transformed, so its latency hides behind the arithmetic
// instead of opening every block.
constexpr int kURowSets = 4;
constexpr int kURowMinBlocks = 512;

template <class P, int Q, int MODE>
__global__ __launch_bounds__(kBlock) void k_fp_urows(const float2* __restrict__ U, int ny, int qc,
                                                     const float2* __restrict__ twt,
                                                     float* __restrict__ out,
                                                     const float* __restrict__ target,
                                                     float* __restrict__ part, int rsets) {
  constexpr int N = P::N, TT = P::T, V = P::V, G = groups_per_block<P>();
  constexpr int R = 2 * G;                    // rows per set
  constexpr int TS = R + 1;                   // tile row stride (odd: lanes reading
                                              // consecutive q hit different banks)
  static_assert(Q >= 0 && 2 * Q <= V, "band registers");
  // tile values each thread prefetches (Q > 0: nq = qc + 1 <= Q T)
  constexpr int kPre = Q > 0 ? (Q * TT * R + kBlock - 1) / kBlock : 0;
  extern __shared__ float2 lds[];
  const int nq = qc + 1;
  float2* tile = lds;                         // [nq][TS]
  float2* scr = lds + nq * TS;
  const int g = threadIdx.x / TT, t = threadIdx.x - g * TT;
  const bool live = g < G;
  const int gg = live ? g : 0;
  const int img = blockIdx.y;
  const int sets = (ny + R - 1) / R;
  const int set0 = blockIdx.x * rsets;
  const int nset = min(rsets, sets - set0);
  const float2* Ub = U + (size_t)img * nq * ny;
  // tile element e = q R + r of set s: U[img][q][s R + r]
  auto fetch = [&](int set, int e) -> float2 {
    const int q = e / R, r = e - q * R, y = set * R + r;
    return (q < nq && y < ny) ? Ub[(size_t)q * ny + y] : make_float2(0.f, 0.f);
  };
  float2 pre[kPre > 0 ? kPre : 1];
  if constexpr (kPre > 0) {
#pragma unroll
    for (int i = 0; i < kPre; ++i) pre[i] = fetch(set0, threadIdx.x + kBlock * i);
  }
  float acc = 0.f;
#pragma unroll 1
  for (int s = 0; s < nset; ++s) {
    const int set = set0 + s;
    if constexpr (kPre > 0) {
#pragma unroll
      for (int i = 0; i < kPre; ++i) {
        const int e = threadIdx.x + kBlock * i;
        const int q = e / R, r = e - q * R;
        if (q < nq) tile[q * TS + r] = pre[i];
      }
    } else {
      for (int e = threadIdx.x; e < nq * R; e += kBlock) {
        const int q = e / R, r = e - q * R;
        tile[q * TS + r] = fetch(set, e);
      }
    }
    __syncthreads();
    float2 v[V];
#pragma unroll
    for (int m = 0; m < V; ++m) {
      const int x = t + TT * m;
      float2 z = make_float2(0.f, 0.f);
      const bool may = Q == 0 || m < Q || m >= V - Q;      // compile-time for Q > 0
      if (may && live && (x <= qc || x >= N - qc)) {
        const bool neg = x > qc;
        const int q = neg ? N - x : x;
        float2 u0 = tile[q * TS + 2 * gg], u1 = tile[q * TS + 2 * gg + 1];
        if (neg) {
          u0.y = -u0.y;
          u1.y = -u1.y;
        }
        z = make_float2(u0.x - u1.y, u0.y + u1.x);  // u0 + i u1
      }
      v[m] = z;
    }
    if constexpr (kPre > 0) {
      if (s + 1 < nset) {
#pragma unroll
        for (int i = 0; i < kPre; ++i) pre[i] = fetch(set + 1, threadIdx.x + kBlock * i);
      }
    }
    const int ya = set * R + 2 * gg, yb = ya + 1;
    // residual and chain forms: the target rows are loaded before the
    // transform, so their latency hides behind it (a load issued after the
    // stores would also wait for them: loads and stores share one in-order
    // counter)
    float2 tgt[MODE ? V : 1];
    if constexpr (MODE != 0) {
#pragma unroll
      for (int m = 0; m < V; ++m) {
        const int x = t + TT * m;
        tgt[m].x = live && ya < ny ? __builtin_nontemporal_load(target + ((size_t)img * ny + ya) * N + x) : 0.f;
        tgt[m].y = live && yb < ny ? __builtin_nontemporal_load(target + ((size_t)img * ny + yb) * N + x) : 0.f;
      }
    }
    P::template run_live_tab<false>(v, scr + g * P::LDS, t, twt, true);
    // the image is written once and not read by this pass: non-temporal stores
    // stream it to memory instead of leaving dirty lines in the caches
    if (live) {
#pragma unroll
      for (int m = 0; m < V; ++m) {
        const int x = t + TT * m;
        const size_t ia = ((size_t)img * ny + ya) * N + x, ib = ((size_t)img * ny + yb) * N + x;
        if (MODE == 1) {
          if (ya < ny) {
            const float r = v[m].x - tgt[m].x;
            acc += r * r;
            __builtin_nontemporal_store(2.f * r, out + ia);
          }
          if (yb < ny) {
            const float r = v[m].y - tgt[m].y;
            acc += r * r;
            __builtin_nontemporal_store(2.f * r, out + ib);
          }
        } else if (MODE == 2) {
          // the logits enter the arithmetic only here: otherwise the sigmoid
          // is scheduled right behind the loads and waits for them before the
          // transform instead of after it
          float ta = tgt[m].x, tb = tgt[m].y;
          asm volatile("" : "+v"(ta), "+v"(tb));
          if (ya < ny) {
            const float mv = logistic(ta);
            const float r = v[m].x * mv * (1.f - mv);
            acc += r * r;
            __builtin_nontemporal_store(r, out + ia);
          }
          if (yb < ny) {
            const float mv = logistic(tb);
            const float r = v[m].y * mv * (1.f - mv);
            acc += r * r;
            __builtin_nontemporal_store(r, out + ib);
          }
        } else {
          if (ya < ny) __builtin_nontemporal_store(v[m].x, out + ia);
          if (yb < ny) __builtin_nontemporal_store(v[m].y, out + ib);
        }
      }
    }
    __syncthreads();                          // the tile is rewritten by the next set
  }
  if constexpr (MODE != 0) {
    __shared__ float red[kBlock / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      float sum = 0.f;
#pragma unroll
      for (int w = 0; w < kBlock / 64; ++w) sum += red[w];
      part[(size_t)img * gridDim.x + blockIdx.x] = sum;
    }
  }
}

// ---------------------------------------------------------------- tables
// Image sizes: N = R1 x R2 (two stages), 2048 = 16 x 8 x 16 (three stages).
using Img128 = reg::GFFT<128, 8, 16>;
using Img256 = reg::GFFT<256, 16, 16>;
using Img512 = reg::GFFT<512, 16, 32>;
using Img1024 = reg::GFFT<1024, 32, 32>;
using Img2048 = reg::GFFT3<2048, 16, 8, 16>;
#define OI_IMAGE_PLANS(X) X(128, Img128) X(256, Img256) X(512, Img512) X(1024, Img1024) X(2048, Img2048)
// Coarse sizes: L = R1 x R2 and the block size of the one-pass coherent kernel
// (0: the coarse grid is too large for it; the 256-thread kernel runs).
#define OI_COARSE_PLANS(X)                                                                      \
  X(16, 4, 4, 64) X(25, 5, 5, 192) X(32, 4, 8, 128) X(36, 6, 6, 256) X(48, 4, 12, 192)         \
  X(64, 8, 8, 512) X(75, 5, 15, 448) X(81, 9, 9, 768) X(100, 10, 10, 0)                       \
  X(128, 8, 16, 0) X(144, 12, 12, 0)

template <class P>
size_t rows_lds() {
  return (size_t)scratch_groups<P>() * P::LDS * sizeof(float2);
}

template <class K>
static int allow(K kernel, size_t bytes) {
  if (bytes > kLdsLimit) return fail(OI_ERR_UNSUPPORTED, "fast path: shared memory over 160 KiB");
  if (bytes > 64 * 1024)
    OI_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return OI_OK;
}

template <class P>
static int run_rows(const FpArgs& a, hipStream_t st) {
  constexpr int G = groups_per_block<P>();
  // transform scratch + output tile [wc][2G + 1]
  const size_t lds = rows_lds<P>() + (size_t)a.wc * (2 * G + 1) * sizeof(float2);
  // narrow form: H output pairs per thread cover the window [-H T, H T]
  int narrow = 0;
  if constexpr (P::STAGES == 2) {
    if ((a.kind == OI_FIELD_REAL || a.kind == kFieldLogits) && a.ny % (2 * G) == 0) {
      for (int h = 1; h <= 4 && !narrow; ++h)
        if (2 * h < P::V && a.k0c >= -h * P::T && a.k0c + a.wc - 1 <= h * P::T) narrow = h;
    }
  }
  if (narrow) {
    if constexpr (P::STAGES == 2) {
      const int sets = (a.ny + 2 * G - 1) / (2 * G);
      dim3 grid((sets + kRowSets - 1) / kRowSets, a.batch);
      const size_t lds_n = rows_narrow_lds<P>(a.wc);
      auto go = [&](auto kernel) -> int {
        int rc = allow(kernel, lds_n);
        if (rc) return rc;
        hipLaunchKernelGGL(kernel, grid, dim3(kBlock), lds_n, st, static_cast<const float*>(a.field),
                           a.ny, a.k0c, a.wc, a.tw_x, a.T, (int)(a.kind == kFieldLogits));
        return OI_OK;
      };
      int rc = narrow == 1   ? go(k_fp_rows_real_narrow<P, 1>)
               : narrow == 2 ? go(k_fp_rows_real_narrow<P, 2>)
               : narrow == 3 ? go(k_fp_rows_real_narrow<P, 3>)
                             : go(k_fp_rows_real_narrow<P, 4>);
      if (rc) return rc;
    }
  } else if (a.kind == OI_FIELD_REAL || a.kind == kFieldLogits) {
    int rc = allow(k_fp_rows_real<P>, lds);
    if (rc) return rc;
    dim3 grid((a.ny + 2 * G - 1) / (2 * G), a.batch);
    hipLaunchKernelGGL(k_fp_rows_real<P>, grid, dim3(kBlock), lds, st,
                       static_cast<const float*>(a.field), a.ny, a.k0c, a.wc, a.tw_x, a.T,
                       (int)(a.kind == kFieldLogits));
  } else {
    int rc = allow(k_fp_rows_c64<P>, lds);
    if (rc) return rc;
    dim3 grid((a.ny + G - 1) / G, a.batch * a.nchan);
    hipLaunchKernelGGL(k_fp_rows_c64<P>, grid, dim3(kBlock), lds, st,
                       static_cast<const float2*>(a.field), a.cin, a.chans, a.nchan, a.ny, a.k0c,
                       a.wc, a.tw_x, a.T);
  }
  OI_LAUNCH_CHECK("fp_rows");
  return OI_OK;
}

template <class P>
static int run_cols(const FpArgs& a, hipStream_t st) {
  constexpr int G = groups_per_block<P>();
  const size_t lds = rows_lds<P>();
  dim3 grid((a.wc + G - 1) / G, a.batch * a.nchan);
  const float scale = 1.0f / ((float)a.ny * (float)a.nx);
  auto go = [&](auto kernel) -> int {
    int rc = allow(kernel, lds);
    if (rc) return rc;
    hipLaunchKernelGGL(kernel, grid, dim3(kBlock), lds, st, a.T, a.k0r, a.wr, a.wc, scale, a.tw_y,
                       a.spec);
    OI_LAUNCH_CHECK("fp_cols");
    return OI_OK;
  };
  if constexpr (P::STAGES == 2) {
    if (a.k0r >= -P::T && a.k0r + a.wr - 1 <= P::T) return go(k_fp_cols_narrow<P, 1>);
    if (4 < P::V && a.k0r >= -2 * P::T && a.k0r + a.wr - 1 <= 2 * P::T)
      return go(k_fp_cols_narrow<P, 2>);
    if (6 < P::V && a.k0r >= -3 * P::T && a.k0r + a.wr - 1 <= 3 * P::T)
      return go(k_fp_cols_narrow<P, 3>);
    if (8 < P::V && a.k0r >= -4 * P::T && a.k0r + a.wr - 1 <= 4 * P::T)
      return go(k_fp_cols_narrow<P, 4>);
  }
  return go(k_fp_cols<P>);
}

template <class P>
static int run_ucols(const FpArgs& a, hipStream_t st) {
  constexpr int G = groups_per_block<P>();
  const size_t lds = rows_lds<P>();
  dim3 grid((a.qc + 1 + G - 1) / G, a.batch);
  auto go = [&](auto kernel) -> int {
    int rc = allow(kernel, lds);
    if (rc) return rc;
    hipLaunchKernelGGL(kernel, grid, dim3(kBlock), lds, st, a.Ihat, a.qc, a.qr, a.tw_y, a.U);
    OI_LAUNCH_CHECK("fp_ucols");
    return OI_OK;
  };
  const int q = (a.qr + 1 + P::T - 1) / P::T;
  if (q <= 2 && 2 * 2 <= P::V) return go(k_fp_ucols<P, 2>);
  if (q <= 3 && 2 * 3 <= P::V) return go(k_fp_ucols<P, 3>);
  if (q <= 4 && 2 * 4 < P::V) return go(k_fp_ucols<P, 4>);
  return go(k_fp_ucols<P, 0>);
}

template <class P>
static int run_urows(const FpArgs& a, hipStream_t st) {
  constexpr int G = groups_per_block<P>();
  const size_t lds = rows_lds<P>() + (size_t)(a.qc + 1) * (2 * G + 1) * sizeof(float2);
  const int rsets = fp_urows_sets(a.ny, P::N, a.batch);
  dim3 grid(fp_urows_blocks(a.ny, P::N, a.batch), a.batch);
  auto go = [&](auto kernel, const float* side) -> int {
    int rc = allow(kernel, lds);
    if (rc) return rc;
    hipLaunchKernelGGL(kernel, grid, dim3(kBlock), lds, st, a.U, a.ny, a.qc, a.tw_x, a.out, side,
                       a.part, rsets);
    OI_LAUNCH_CHECK("fp_urows");
    return OI_OK;
  };
  // registers that can hold band values: Q = ceil((qc + 1) / T) at each end
  const int q = (a.qc + 1 + P::T - 1) / P::T;
  if (a.target) {
    if (q <= 2 && 2 * 2 <= P::V) return go(k_fp_urows<P, 2, 1>, a.target);
    if (q <= 3 && 2 * 3 <= P::V) return go(k_fp_urows<P, 3, 1>, a.target);
    if (q <= 4 && 2 * 4 < P::V) return go(k_fp_urows<P, 4, 1>, a.target);
    return go(k_fp_urows<P, 0, 1>, a.target);
  }
  if (a.theta) {
    if (q <= 2 && 2 * 2 <= P::V) return go(k_fp_urows<P, 2, 2>, a.theta);
    if (q <= 3 && 2 * 3 <= P::V) return go(k_fp_urows<P, 3, 2>, a.theta);
    if (q <= 4 && 2 * 4 < P::V) return go(k_fp_urows<P, 4, 2>, a.theta);
    return go(k_fp_urows<P, 0, 2>, a.theta);
  }
  if (q <= 2 && 2 * 2 <= P::V) return go(k_fp_urows<P, 2, 0>, nullptr);
  if (q <= 3 && 2 * 3 <= P::V) return go(k_fp_urows<P, 3, 0>, nullptr);
  if (q <= 4 && 2 * 4 < P::V) return go(k_fp_urows<P, 4, 0>, nullptr);
  return go(k_fp_urows<P, 0, 0>, nullptr);
}

// The one-pass kernel serves coarse sizes with a block size in the table when
// the window folds onto the coarse grid without collisions.
template <class PL, int NT>
static bool one_pass(const FpArgs& a) {
  return NT > 0 && a.wr <= PL::N && a.wc <= PL::N;
}

template <class PL, int NT>
static size_t coherent_lds(const FpArgs& a) {
  if (one_pass<PL, NT>(a)) return coh_lds_bytes<PL>(a.wr, a.wc);
  const size_t scr = (size_t)scratch_groups<PL>() * PL::LDS;
  return ((size_t)a.wr * a.wc + (size_t)PL::N * a.wc + scr) * sizeof(float2);
}

template <class PL, int NT>
static int run_coherent(const FpArgs& a, const FpCoh& c, hipStream_t st) {
  const size_t lds = coherent_lds<PL, NT>(a);
  if constexpr (NT > 0) {
    if (one_pass<PL, NT>(a)) {
      int rc = allow(k_fp_coh<PL, NT>, lds);
      if (rc) return rc;
      dim3 grid(a.batch, c.groups);
      hipLaunchKernelGGL((k_fp_coh<PL, NT>), grid, dim3(NT), lds, st, c);
      OI_LAUNCH_CHECK("fp_coh");
      return OI_OK;
    }
  }
  int rc = allow(k_fp_coherent<PL, PL>, lds);
  if (rc) return rc;
  dim3 grid(1, a.batch, c.groups);
  hipLaunchKernelGGL((k_fp_coherent<PL, PL>), grid, dim3(kBlock), lds, st, c);
  OI_LAUNCH_CHECK("fp_coherent");
  return OI_OK;
}

template <class PL>
static int run_band(const FpArgs& a, int groups, hipStream_t st) {
  constexpr int L = PL::N, S = kBandWide / L > 0 ? kBandWide / L : 1;
  const size_t lds = ((size_t)S * L + L + PL::LDS) * sizeof(float2);
  int rc = allow(k_fp_band<PL>, lds);
  if (rc) return rc;
  const float scale = 1.0f / ((float)PL::N * (float)PL::N);
  hipLaunchKernelGGL(k_fp_band<PL>, dim3(a.qc + 1, a.batch), dim3(kBandWide), lds, st, a.R, groups,
                     a.qc, a.qr, scale, a.tw_l, a.Ihat);
  OI_LAUNCH_CHECK("fp_band");
  return OI_OK;
}

// Large coarse grids of the two-pass coherent sum.
#define OI_TWO_PASS_PLANS(X) X(144, 12, 12) X(256, 16, 16) X(450, 15, 30) X(512, 16, 32)

template <class PL>
static size_t two_pass_rows_lds(int wc) {
  // the row tile and the transform scratch share the LDS (k_fp_brows)
  const size_t tile = (size_t)wc * (groups_per_block<PL>() + 1);
  const size_t scr = (size_t)scratch_groups<PL>() * PL::LDS;
  return (tile > scr ? tile : scr) * sizeof(float2);
}

template <class PL>
static int run_two_pass(const FpArgs& a, const FpCoh& c, float2* B, int generic_layout,
                        hipStream_t st) {
  constexpr int G = groups_per_block<PL>();
  const size_t lds1 = rows_lds<PL>();
  int rc = allow(k_fp_bcols<PL>, lds1);
  if (rc) return rc;
  dim3 g1((a.wc + G - 1) / G, c.nkern * c.nchan, a.batch);
  hipLaunchKernelGGL(k_fp_bcols<PL>, g1, dim3(kBlock), lds1, st, c.spec, c.kern, c.nkern, c.nchan,
                     a.k0r, a.wr, a.wc, a.tw_l, B);
  OI_LAUNCH_CHECK("fp_bcols");
  const size_t lds2 = two_pass_rows_lds<PL>(a.wc);
  rc = allow(k_fp_brows<PL>, lds2);
  if (rc) return rc;
  if (generic_layout && c.groups != 1) return fail(OI_ERR_INVALID, "two-pass path: one group in the generic layout");
  const int items = ((PL::N + G - 1) / G) * a.batch * c.groups;
  hipLaunchKernelGGL(k_fp_brows<PL>, dim3(xcd_grid(items)), dim3(kBlock), lds2, st, B,
                     c.nkern * c.nchan, a.k0c, a.wc, a.qc, generic_layout, a.tw_l, c.R, a.batch,
                     c.groups);
  OI_LAUNCH_CHECK("fp_brows");
  return OI_OK;
}

}  // namespace fp

// ------------------------------------------------------------ dispatch
bool fp_two_pass_size_ok(int l) {
#define X(NN, A, B) if (l == NN) return true;
  OI_TWO_PASS_PLANS(X)
#undef X
  return false;
}

int fp_two_pass_choose(int need) {
  int best = 0;
#define X(NN, A, B) if (NN >= need && (best == 0 || NN < best)) best = NN;
  OI_TWO_PASS_PLANS(X)
#undef X
  return best;
}

size_t fp_two_pass_lds(int L, int wc) {
#define X(NN, A, B) if (L == NN) return fp::two_pass_rows_lds<reg::GFFT<NN, A, B>>(wc);
  OI_TWO_PASS_PLANS(X)
#undef X
  return (size_t)-1;
}

int fp_two_pass(const FpArgs& a, const FpCoh& c, float2* B, int generic_layout, hipStream_t st) {
  int rc = OI_ERR_UNSUPPORTED;
#define X(NN, A, B_) if (a.L == NN) rc = fp::run_two_pass<reg::GFFT<NN, A, B_>>(a, c, B, generic_layout, st);
  OI_TWO_PASS_PLANS(X)
#undef X
  if (rc == OI_ERR_UNSUPPORTED) return fail(rc, "two-pass path: unsupported coarse size");
  return rc;
}

// Band extraction + column and row transforms to the image (stages 4, 5) for
// partial sums R[img][groups][q][t] on an L x L coarse grid.
int fp_band(const FpArgs& a, int groups, hipStream_t st) {
  int rc = OI_ERR_UNSUPPORTED;
#define X(NN, A, B, NT) if (a.L == NN) rc = fp::run_band<reg::GFFT<NN, A, B>>(a, groups, st);
  OI_COARSE_PLANS(X)
#undef X
#define X(NN, A, B) \
  if (rc == OI_ERR_UNSUPPORTED && a.L == NN) rc = fp::run_band<reg::GFFT<NN, A, B>>(a, groups, st);
  OI_TWO_PASS_PLANS(X)
#undef X
  if (rc) return rc == OI_ERR_UNSUPPORTED ? fail(rc, "fast path: unsupported coarse size") : rc;
  return OI_OK;
}

int fp_interp(const FpArgs& a, int groups, hipStream_t st) {
  int rc = fp_band(a, groups, st);
  if (rc) return rc;
  rc = OI_ERR_UNSUPPORTED;
#define X(NN, PT) if (a.ny == NN) rc = fp::run_ucols<fp::PT>(a, st);
  OI_IMAGE_PLANS(X)
#undef X
  if (rc) return rc;
  if (a.mark) a.mark(a.mark_ctx, 4);
  rc = OI_ERR_UNSUPPORTED;
#define X(NN, PT) if (a.nx == NN) rc = fp::run_urows<fp::PT>(a, st);
  OI_IMAGE_PLANS(X)
#undef X
  if (rc) return rc;
  if (a.mark) a.mark(a.mark_ctx, 5);
  return OI_OK;
}
// The image from a band Ihat[img][q][2qr+1] (q in [0, qc]), without the
// band kernel: the interpolation column and row transforms only.
int fp_band_image(const FpArgs& a, hipStream_t st) {
  int rc = OI_ERR_UNSUPPORTED;
#define X(NN, PT) if (a.ny == NN) rc = fp::run_ucols<fp::PT>(a, st);
  OI_IMAGE_PLANS(X)
#undef X
  if (rc) return rc == OI_ERR_UNSUPPORTED ? fail(rc, "fast path: unsupported ny") : rc;
  rc = OI_ERR_UNSUPPORTED;
#define X(NN, PT) if (a.nx == NN) rc = fp::run_urows<fp::PT>(a, st);
  OI_IMAGE_PLANS(X)
#undef X
  if (rc) return rc == OI_ERR_UNSUPPORTED ? fail(rc, "fast path: unsupported nx") : rc;
  return OI_OK;
}

int fp_urows_sets(int ny, int nx, int batch) {
#define X(NN, PT)                                                                            \
  if (nx == NN) {                                                                          \
    const int sets = (ny + 2 * fp::groups_per_block<fp::PT>() - 1) / (2 * fp::groups_per_block<fp::PT>()); \
    const long long r = (long long)sets * (batch > 0 ? batch : 1) / fp::kURowMinBlocks;    \
    return r < 1 ? 1 : (r > fp::kURowSets ? fp::kURowSets : (int)r);                       \
  }
  OI_IMAGE_PLANS(X)
#undef X
  return 1;
}

int fp_urows_blocks(int ny, int nx, int batch) {
#define X(NN, PT)                                                                            \
  if (nx == NN) {                                                                          \
    const int sets = (ny + 2 * fp::groups_per_block<fp::PT>() - 1) / (2 * fp::groups_per_block<fp::PT>()); \
    const int r = fp_urows_sets(ny, nx, batch);                                            \
    return (sets + r - 1) / r;                                                             \
  }
  OI_IMAGE_PLANS(X)
#undef X
  return 0;
}

bool fp_stage_radices(int n, int* r1, int* r2) {
#define X(NN, PT)                                                                            \
  if (n == NN && fp::PT::STAGES == 2) {                                                    \
    *r1 = fp::PT::R1;                                                                      \
    *r2 = fp::PT::R2;                                                                      \
    return true;                                                                           \
  }
  OI_IMAGE_PLANS(X)
#undef X
#define X(NN, A, B, NT)                                                                      \
  if (n == NN) {                                                                           \
    *r1 = A;                                                                               \
    *r2 = B;                                                                               \
    return true;                                                                           \
  }
  OI_COARSE_PLANS(X)
#undef X
#define X(NN, A, B)                                                                          \
  if (n == NN) {                                                                           \
    *r1 = A;                                                                               \
    *r2 = B;                                                                               \
    return true;                                                                           \
  }
  OI_TWO_PASS_PLANS(X)
#undef X
  return false;
}

bool fp_image_size_ok(int n) {
#define X(NN, PT) if (n == NN) return true;
  OI_IMAGE_PLANS(X)
#undef X
  return false;
}

bool fp_coarse_size_ok(int l) {
#define X(NN, A, B, NT) if (l == NN) return true;
  OI_COARSE_PLANS(X)
#undef X
  return false;
}

int fp_coarse_choose(int need) {
  int best = 0;
#define X(NN, A, B, NT) if (NN >= need && (best == 0 || NN < best)) best = NN;
  OI_COARSE_PLANS(X)
#undef X
  return best;
}

size_t fp_coherent_lds(int L, int wr, int wc) {
  FpArgs a{};
  a.wr = wr;
  a.wc = wc;
#define X(NN, A, B, NT) if (L == NN) return fp::coherent_lds<reg::GFFT<NN, A, B>, NT>(a);
  OI_COARSE_PLANS(X)
#undef X
  return (size_t)-1;
}

bool fp_one_pass_ok(int L) {
#define X(NN, A, B, NT) if (L == NN) return NT > 0;
  OI_COARSE_PLANS(X)
#undef X
  return false;
}

int fp_coherent_rows_fit(int L) {
  // accumulator registers per thread of the 256-thread kernel: ceil(L / groups) * V
#define X(NN, A, B, NT)                                                                  \
  if (L == NN) {                                                                         \
    using P = reg::GFFT<NN, A, B>;                                                       \
    constexpr int G = fp::groups_per_block<P>();                                         \
    return (NT > 0) || ((NN + G - 1) / G) * P::V <= 96;                                  \
  }
  OI_COARSE_PLANS(X)
#undef X
  return 0;
}

int fp_spectrum(const FpArgs& a, hipStream_t st) {
  int rc = OI_ERR_UNSUPPORTED;
#define X(NN, PT) if (a.nx == NN) rc = fp::run_rows<fp::PT>(a, st);
  OI_IMAGE_PLANS(X)
#undef X
  if (rc) return rc == OI_ERR_UNSUPPORTED ? fail(rc, "fast path: unsupported nx") : rc;
  if (a.mark) a.mark(a.mark_ctx, 1);
  rc = OI_ERR_UNSUPPORTED;
#define X(NN, PT) if (a.ny == NN) rc = fp::run_cols<fp::PT>(a, st);
  OI_IMAGE_PLANS(X)
#undef X
  if (rc) return rc == OI_ERR_UNSUPPORTED ? fail(rc, "fast path: unsupported ny") : rc;
  return OI_OK;
}

static int fp_coherent(const FpArgs& a, const FpCoh& c, hipStream_t st) {
  int rc = OI_ERR_UNSUPPORTED;
#define X(NN, A, B, NT) if (a.L == NN) rc = fp::run_coherent<reg::GFFT<NN, A, B>, NT>(a, c, st);
  OI_COARSE_PLANS(X)
#undef X
  if (rc) return rc == OI_ERR_UNSUPPORTED ? fail(rc, "fast path: unsupported coarse size") : rc;
  return OI_OK;
}

int fp_image(const FpArgs& a, const FpCoh& c, hipStream_t st) {
  int rc = fp_coherent(a, c, st);
  if (rc) return rc;
  if (a.mark) a.mark(a.mark_ctx, 3);
  return fp_interp(a, c.groups, st);
}

int fp_image_band(const FpArgs& a, const FpCoh& c, hipStream_t st) {
  int rc = fp_coherent(a, c, st);
  if (rc) return rc;
  return fp_band(a, c.groups, st);
}

}  // namespace oi
