// Register-resident FFTs with compile-time sizes.
//
// DFT<N, INV>: an N-point DFT on N values held by one thread, fully unrolled
// mixed-radix (first factor 4, 2, 3 or 5): Good-Thomas prime-factor splits
// where the two factors are coprime, Cooley-Tukey otherwise.  Twiddle factors are
// compile-time constants: exp(-+2 pi i m / N) is evaluated by constexpr series
// in float64 and rounded once, so no table is read and nothing spills.
//
// GFFT<N, R1, R2>: an N = R1 R2 point transform shared by T = N / V threads
// (V = max(R1, R2), the smaller radix divides the larger).  Thread t holds
// x[t + T m], m in [0, V), on entry and X[t + T m] on exit (Stockham order).
// Stage 1 (radix R1) runs on the thread's own values; its outputs pass once
// through a padded LDS scratch of the group; stage 2 (radix R2) applies the
// twiddles w_N^{j (i mod R1)} (held in registers, loaded once per thread) and
// lands back in registers.  R2 = 1 means a single-thread transform (T = 1).
#pragma once

#include "oi_common.h"

namespace oi {
namespace reg {

constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double series_sin(double x) {
  double x2 = x * x, term = x, s = x;
  for (int k = 1; k < 16; ++k) {
    term *= -x2 / ((2.0 * k) * (2.0 * k + 1.0));
    s += term;
  }
  return s;
}
constexpr double series_cos(double x) {
  double x2 = x * x, term = 1.0, s = 1.0;
  for (int k = 1; k < 16; ++k) {
    term *= -x2 / ((2.0 * k - 1.0) * (2.0 * k));
    s += term;
  }
  return s;
}
// exp(-2 pi i m / n): argument reduced to [-pi, pi] in integer arithmetic.
constexpr double root_re(long m, long n) {
  m %= n;
  if (m < 0) m += n;
  return series_cos(2.0 * kPi * (double)(2 * m <= n ? m : m - n) / (double)n);
}
constexpr double root_im(long m, long n) {
  m %= n;
  if (m < 0) m += n;
  return -series_sin(2.0 * kPi * (double)(2 * m <= n ? m : m - n) / (double)n);
}

template <int N>
struct FirstFactor {
  static constexpr int value = (N % 4 == 0 && N != 4) ? 4 : (N % 2 == 0 ? 2 : (N % 3 == 0 ? 3 : (N % 5 == 0 ? 5 : N)));
};

__device__ __forceinline__ float2 mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// The complex product with each multiply and add rounded on its own (no
// fused multiply-add), so that two kernels forming the same product get the
// same bits whatever the compiler fuses elsewhere in them.
__device__ __forceinline__ float2 mul_rn(float2 a, float2 b) {
#pragma clang fp contract(off)
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int N, bool INV>
struct DFT;

template <bool INV>
struct DFT<1, INV> {
  __device__ __forceinline__ static void run(float2*) {}
};

template <bool INV>
struct DFT<2, INV> {
  __device__ __forceinline__ static void run(float2* u) {
    const float2 a = u[0], b = u[1];
    u[0] = make_float2(a.x + b.x, a.y + b.y);
    u[1] = make_float2(a.x - b.x, a.y - b.y);
  }
};

template <bool INV>
struct DFT<3, INV> {
  __device__ __forceinline__ static void run(float2* u) {
    const float h = 0.86602540378443864676f;
    const float2 s = make_float2(u[1].x + u[2].x, u[1].y + u[2].y);
    const float2 d = make_float2((u[1].x - u[2].x) * h, (u[1].y - u[2].y) * h);
    const float2 t = make_float2(u[0].x - 0.5f * s.x, u[0].y - 0.5f * s.y);
    const float2 r = INV ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
    u[0] = make_float2(u[0].x + s.x, u[0].y + s.y);
    u[1] = make_float2(t.x + r.x, t.y + r.y);
    u[2] = make_float2(t.x - r.x, t.y - r.y);
  }
};

template <bool INV>
struct DFT<4, INV> {
  __device__ __forceinline__ static void run(float2* u) {
    const float2 s02 = make_float2(u[0].x + u[2].x, u[0].y + u[2].y);
    const float2 d02 = make_float2(u[0].x - u[2].x, u[0].y - u[2].y);
    const float2 s13 = make_float2(u[1].x + u[3].x, u[1].y + u[3].y);
    const float2 d13 = make_float2(u[1].x - u[3].x, u[1].y - u[3].y);
    const float2 r = INV ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
    u[0] = make_float2(s02.x + s13.x, s02.y + s13.y);
    u[2] = make_float2(s02.x - s13.x, s02.y - s13.y);
    u[1] = make_float2(d02.x + r.x, d02.y + r.y);
    u[3] = make_float2(d02.x - r.x, d02.y - r.y);
  }
};

template <bool INV>
struct DFT<5, INV> {
  __device__ __forceinline__ static void run(float2* u) {
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    const float2 a1 = make_float2(u[1].x + u[4].x, u[1].y + u[4].y);
    const float2 b1 = make_float2(u[1].x - u[4].x, u[1].y - u[4].y);
    const float2 a2 = make_float2(u[2].x + u[3].x, u[2].y + u[3].y);
    const float2 b2 = make_float2(u[2].x - u[3].x, u[2].y - u[3].y);
    const float2 x0 = u[0];
    const float2 t1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    const float2 t2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    const float2 v1 = make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
    const float2 v2 = make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
    const float2 r1 = INV ? make_float2(-v1.y, v1.x) : make_float2(v1.y, -v1.x);
    const float2 r2 = INV ? make_float2(-v2.y, v2.x) : make_float2(v2.y, -v2.x);
    u[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
    u[1] = make_float2(t1.x + r1.x, t1.y + r1.y);
    u[4] = make_float2(t1.x - r1.x, t1.y - r1.y);
    u[2] = make_float2(t2.x + r2.x, t2.y + r2.y);
    u[3] = make_float2(t2.x - r2.x, t2.y - r2.y);
  }
};

constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }
// multiplicative inverse of a modulo m (gcd(a, m) = 1), by search
constexpr int inv_mod(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 0;
}

// N = F M.  Coprime F, M (6, 10, 12, 15, 20, 30, 36, 60): Good-Thomas
// prime-factor form, no twiddle factors (Chinese-remainder index maps):
//   n = (M f + F m) mod N,  k = (k1 M (M^-1 mod F) + k2 F (F^-1 mod M)) mod N,
//   X[k] = sum_m w_M^{m k2} sum_f w_F^{f k1} x[n].
// Otherwise Cooley-Tukey, n = M f + m, k = k1 + F k2:
//   Y[m][k1] = sum_f x[M f + m] w_F^{f k1};  Y[m][k1] *= w_N^{m k1};
//   X[k1 + F k2] = sum_m Y[m][k1] w_M^{m k2}.
template <int N, bool INV>
struct DFT {
  __device__ __forceinline__ static void run(float2* x) {
    constexpr int F = FirstFactor<N>::value;
    constexpr int M = N / F;
    static_assert(F > 1 && F <= 5 && M * F == N, "DFT size must be a product of 2, 3 and 5");
    float2 y[N];
    if constexpr (gcd_c(F, M) == 1) {
      constexpr int A = inv_mod(M % F, F), B = inv_mod(F % M, M);
#pragma unroll
      for (int m = 0; m < M; ++m) {
        float2 t[F];
#pragma unroll
        for (int f = 0; f < F; ++f) t[f] = x[(M * f + F * m) % N];
        DFT<F, INV>::run(t);
#pragma unroll
        for (int k1 = 0; k1 < F; ++k1) y[m * F + k1] = t[k1];
      }
#pragma unroll
      for (int k1 = 0; k1 < F; ++k1) {
        float2 t[M];
#pragma unroll
        for (int m = 0; m < M; ++m) t[m] = y[m * F + k1];
        DFT<M, INV>::run(t);
#pragma unroll
        for (int k2 = 0; k2 < M; ++k2) x[(k1 * M * A + k2 * F * B) % N] = t[k2];
      }
      return;
    }
#pragma unroll
    for (int m = 0; m < M; ++m) {
      float2 t[F];
#pragma unroll
      for (int f = 0; f < F; ++f) t[f] = x[M * f + m];
      DFT<F, INV>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < F; ++k1) {
        const int e = m * k1;   // twiddle w_N^e, 0 <= e < N
        if (e == 0) {
          y[m * F + k1] = t[k1];
        } else if (2 * e == N) {          // -1
          y[m * F + k1] = make_float2(-t[k1].x, -t[k1].y);
        } else if (4 * e == N || 4 * e == 3 * N) {
          // -i (forward, e = N/4) or +i: a swap and a sign, exact, no multiply
          const bool minus_i = (4 * e == N) != INV;
          y[m * F + k1] = minus_i ? make_float2(t[k1].y, -t[k1].x) : make_float2(-t[k1].y, t[k1].x);
        } else {
          const float wr = (float)root_re(m * k1, N);
          const float wi = (float)(INV ? -root_im(m * k1, N) : root_im(m * k1, N));
          y[m * F + k1] = mul(t[k1], make_float2(wr, wi));
        }
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < F; ++k1) {
      float2 t[M];
#pragma unroll
      for (int m = 0; m < M; ++m) t[m] = y[m * F + k1];
      DFT<M, INV>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < M; ++k2) x[k1 + F * k2] = t[k2];
    }
  }
};

// Orders this wavefront's LDS accesses before and after the call (no
// workgroup barrier): fences at wavefront scope around a scheduling barrier.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Padded LDS index: one spare slot every 32 entries breaks power-of-two strides.
__device__ __forceinline__ int pad(int x) { return x + (x >> 5); }
__host__ __device__ constexpr int padded(int n) { return n + (n >> 5) + 1; }

template <int N_, int R1_, int R2_>
struct GFFT {
  static constexpr int N = N_;
  static constexpr int R1 = R1_;
  static constexpr int R2 = R2_;
  static constexpr int V = R2 == 1 ? N : (R1 > R2 ? R1 : R2);
  static constexpr int T = N / V;
  static constexpr int U1 = V / R1;                 // stage-1 butterflies per thread
  static constexpr int U2 = R2 == 1 ? 1 : V / R2;   // stage-2 butterflies per thread
  // Exchange index: sizes that are multiples of 32 take one spare slot every
  // 32 entries (power-of-two strides would meet in one bank); other sizes use
  // the plain index, which keeps every LDS address a per-thread base plus an
  // immediate offset.
  static constexpr bool PAD = (N % 32) == 0;
  __device__ __forceinline__ static int xi(int i) { return PAD ? pad(i) : i; }
  static constexpr int LDS = R2 == 1 ? 0 : padded(N);  // scratch (float2) per group
  static constexpr int STAGES = 2;
  static_assert(R2 == 1 ? R1 == N : (R1 * R2 == N && (V % R1) == 0 && (V % R2) == 0),
                "GFFT: N = R1 R2 with the smaller radix dividing the larger");

  // Stage-2 twiddles of thread t: w_N^{j (i mod R1)}, i = t + T u.
  struct Tw {
    float2 w[U2][R2 > 1 ? R2 - 1 : 1];
  };

  // table: exp(-2 pi i m / N), m in [0, N) (float64 values rounded once).
  __device__ __forceinline__ static void load_tw(Tw& tw, int t, const float2* __restrict__ table) {
    if constexpr (R2 > 1) {
#pragma unroll
      for (int u = 0; u < U2; ++u) {
        const int k = (t + T * u) % R1;
#pragma unroll
        for (int j = 1; j < R2; ++j) tw.w[u][j - 1] = table[j * k];
      }
    }
  }

  // v: x[t + T m] in, X[t + T m] out.  lds: this group's scratch (LDS entries);
  // every thread of the block must call (barriers inside) when R2 > 1.
  template <bool INV>
  __device__ __forceinline__ static void run(float2 (&v)[V], float2* lds, int t, const Tw& tw) {
    run_impl<INV>(v, [lds](int i) -> float2& { return lds[xi(i)]; }, t, tw, true);
  }

  // The same transform exchanging through base[i * stride], i in [0, N) (no
  // padding); threads with live == false take part in the barriers only.
  template <bool INV>
  __device__ __forceinline__ static void run_strided(float2 (&v)[V], float2* base, int stride, int t,
                                                     const Tw& tw, bool live) {
    run_impl<INV>(v, [base, stride](int i) -> float2& { return base[i * stride]; }, t, tw, live);
  }

  // Exchange through a padded contiguous region (pad(i) < padded(N)), live-gated.
  template <bool INV>
  __device__ __forceinline__ static void run_live(float2 (&v)[V], float2* lds, int t, const Tw& tw,
                                                  bool live) {
    run_impl<INV>(v, [lds](int i) -> float2& { return lds[xi(i)]; }, t, tw, live);
  }

  // Twiddles read from a table at use instead of registers.  The table holds
  // 2N entries: table[m] = exp(-2 pi i m / N) for m < N, then the stage-2
  // twiddles w_N^{j k} at table[N + j R1 + k] (j < R2, k < R1), the same
  // values laid out so that the consecutive k of a group's lanes read
  // consecutive entries instead of a stride-j gather.
  struct TabTw {
    const float2* table;
    __device__ __forceinline__ float2 get(int u, int j, int t) const {
      return table[N + j * R1 + (t + T * u) % R1];
    }
  };
  struct RegTw {
    const Tw& tw;
    __device__ __forceinline__ float2 get(int u, int j, int) const { return tw.w[u][j - 1]; }
  };

  template <bool INV>
  __device__ __forceinline__ static void run_strided_tab(float2 (&v)[V], float2* base, int stride,
                                                         int t, const float2* table, bool live) {
    run_gen<INV>(v, [base, stride](int i) -> float2& { return base[i * stride]; }, t,
                 TabTw{table}, live);
  }
  template <bool INV>
  __device__ __forceinline__ static void run_live_tab(float2 (&v)[V], float2* lds, int t,
                                                      const float2* table, bool live) {
    run_gen<INV>(v, [lds](int i) -> float2& { return lds[xi(i)]; }, t, TabTw{table}, live);
  }

  template <bool INV, class Addr>
  __device__ __forceinline__ static void run_impl(float2 (&v)[V], Addr at, int t, const Tw& tw,
                                                  bool live) {
    run_gen<INV>(v, at, t, RegTw{tw}, live);
  }

  // Exchange within one wavefront (the group's lanes and its scratch belong to
  // a single wave): LDS operations of a wave complete in order, so a compiler
  // barrier with wavefront-scope fences replaces the workgroup barrier.
  template <bool INV>
  __device__ __forceinline__ static void run_strided_wave(float2 (&v)[V], float2* base, int stride,
                                                          int t, const Tw& tw, bool live) {
    run_gen<INV, true>(v, [base, stride](int i) -> float2& { return base[i * stride]; }, t,
                       RegTw{tw}, live);
  }
  template <bool INV>
  __device__ __forceinline__ static void run_live_wave(float2 (&v)[V], float2* lds, int t,
                                                       const Tw& tw, bool live) {
    run_gen<INV, true>(v, [lds](int i) -> float2& { return lds[xi(i)]; }, t, RegTw{tw}, live);
  }

  template <bool INV, bool WAVE = false, class Addr, class TW>
  __device__ __forceinline__ static void run_gen(float2 (&v)[V], Addr at, int t, const TW& tw,
                                                 bool live) {
    if constexpr (R2 == 1) {
      DFT<N, INV>::run(v);
    } else {
#pragma unroll
      for (int u = 0; u < U1; ++u) {
        float2 b[R1];
#pragma unroll
        for (int j = 0; j < R1; ++j) b[j] = v[u + j * U1];
        DFT<R1, INV>::run(b);
        const int i = t + T * u;
        if (live) {
#pragma unroll
          for (int j = 0; j < R1; ++j) at(i * R1 + j) = b[j];
        }
      }
      sync<WAVE>();
#pragma unroll
      for (int u = 0; u < U2; ++u) {
        const int i = t + T * u;
        float2 b[R2];
#pragma unroll
        for (int j = 0; j < R2; ++j) b[j] = at(i + j * (N / R2));
#pragma unroll
        for (int j = 1; j < R2; ++j) {
          const float2 w = tw.get(u, j, t);
          b[j] = mul(b[j], INV ? make_float2(w.x, -w.y) : w);
        }
        DFT<R2, INV>::run(b);
#pragma unroll
        for (int j = 0; j < R2; ++j) v[u + j * U2] = b[j];
      }
      sync<WAVE>();
    }
  }

  template <bool WAVE>
  __device__ __forceinline__ static void sync() {
    if constexpr (WAVE) {
      wave_sync();
    } else {
      __syncthreads();
    }
  }
};

// GFFT3<N, R1, R2, R3>: an N = R1 R2 R3 point transform with three radix
// stages and two LDS exchanges, for sizes whose two-stage form would hold too
// many values per thread (2048 = 16 x 8 x 16: 16 values per thread, T = 128
// threads per transform instead of 64 values per thread).  Same interface as
// GFFT for the table-twiddle path: thread t holds x[t + T m] on entry and
// X[t + T m] on exit.
//
// With n = i + M j (M = N / R1) and k = k1 + R1 k', and then i = a + R3 b,
// k' = c + R2 d:
//   stage 1 (registers): Y[i][k1] = sum_j x[i + M j] w_R1^{j k1}, stored at i R1 + k1;
//   stage 2: Z[k1][a][c] = w_M^{a c} sum_b w_R2^{b c} w_N^{(a + R3 b) k1} Y[a + R3 b][k1]
//            for p = k1 + R1 a, read at p + (N / R2) b, stored at a R1 R2 + k1 + R1 c;
//   stage 3 (p3 = k1 + R1 c): X[p3 + (N / R3) d] = sum_a w_R3^{a d} Z[k1][a][c].
// Every thread of the block calls (three workgroup barriers inside).
template <int N_, int R1_, int R2_, int R3_>
struct GFFT3 {
  static constexpr int N = N_;
  static constexpr int R1 = R1_, R2 = R2_, R3 = R3_;
  static constexpr int V = R1 > R2 ? (R1 > R3 ? R1 : R3) : (R2 > R3 ? R2 : R3);
  static constexpr int T = N / V;
  static constexpr int U1 = V / R1, U2 = V / R2, U3 = V / R3;
  static constexpr bool PAD = (N % 32) == 0;
  static constexpr int LDS = padded(N);
  static constexpr int STAGES = 3;
  static_assert(R1 * R2 * R3 == N && V % R1 == 0 && V % R2 == 0 && V % R3 == 0,
                "GFFT3: N = R1 R2 R3 with every radix dividing the largest");
  __device__ __forceinline__ static int xi(int i) { return PAD ? pad(i) : i; }

  // table[m] = exp(-2 pi i m / N), m in [0, N); INV uses it conjugated.
  template <bool INV>
  __device__ __forceinline__ static float2 tw(const float2* __restrict__ table, int m) {
    const float2 w = table[m];
    return INV ? make_float2(w.x, -w.y) : w;
  }

  template <bool INV>
  __device__ __forceinline__ static void run_live_tab(float2 (&v)[V], float2* lds, int t,
                                                      const float2* __restrict__ table, bool live) {
#pragma unroll
    for (int u = 0; u < U1; ++u) {
      float2 b[R1];
#pragma unroll
      for (int j = 0; j < R1; ++j) b[j] = v[u + j * U1];
      DFT<R1, INV>::run(b);
      const int i = t + T * u;
      if (live) {
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) lds[xi(i * R1 + k1)] = b[k1];
      }
    }
    __syncthreads();
    float2 z[U2][R2];
#pragma unroll
    for (int u = 0; u < U2; ++u) {
      const int p = t + T * u, k1 = p % R1, a = p / R1;
#pragma unroll
      for (int bb = 0; bb < R2; ++bb) {
        const float2 x = lds[xi(p + (N / R2) * bb)];
        z[u][bb] = mul(x, tw<INV>(table, (a + R3 * bb) * k1));
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U2; ++u) {
      const int p = t + T * u, k1 = p % R1, a = p / R1;
      DFT<R2, INV>::run(z[u]);
      if (live) {
#pragma unroll
        for (int c = 0; c < R2; ++c)
          lds[xi(a * R1 * R2 + k1 + R1 * c)] = mul(z[u][c], tw<INV>(table, R1 * a * c));
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U3; ++u) {
      const int p3 = t + T * u;
      float2 b[R3];
#pragma unroll
      for (int aa = 0; aa < R3; ++aa) b[aa] = lds[xi(p3 + (N / R3) * aa)];
      DFT<R3, INV>::run(b);
#pragma unroll
      for (int d = 0; d < R3; ++d) v[u + U3 * d] = b[d];
    }
    __syncthreads();
  }
};

}  // namespace reg
}  // namespace oi
