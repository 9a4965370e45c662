// Adjoint of the SOCS image with respect to a real mask.
//
// Forward (one image, real mask h on an ny x nx grid, window kappa in W):
//   s(kappa)  = (1/(ny nx)) sum_m h[m] exp(+2 pi i kappa.m / N)
//   E_k[j]    = sum_kappa phi_k(kappa) s(kappa) exp(-2 pi i kappa.j / N)
//   I[j]      = sum_k |E_k[j]|^2
// For a real image cotangent g = dL/dI:
//   dL/dh[m] = (2/(ny nx)) Re sum_kappa exp(+2 pi i kappa.m / N) Q(kappa),
//   Q(kappa) = sum_k phi_k(kappa) A_k(kappa),
//   A_k(kappa) = sum_j g[j] conj(E_k[j]) exp(-2 pi i kappa.j / N)
//              = sum_kappa' conj(phi_k s)(kappa') ghat(kappa - kappa'),
// so A_k needs the spectrum of g only for |nu| <= q = (band - 1) per axis.
// With gamma[t] = sum_{|nu|<=q} s_g(nu) exp(-2 pi i nu.t / L) (s_g the scaled
// spectrum of g, as s above) and e_k[t] the coherent field on the coarse L
// grid (the forward's), A_k(kappa) = (ny nx / L^2) sum_t gamma[t] conj(e_k[t])
// exp(-2 pi i kappa.t / L) exactly when L >= 2 band - 1 (no alias reaches a
// window index).  For kernel sets that hold every eigenpair of their TCC the
// same holds with band = one pupil: aliased terms carry TCC(kappa, kappa')
// with |kappa - kappa'| >= band, which is zero.  On a direct plan (L = N)
// gamma = g.  Fields are recomputed per kernel, never stored from the forward.
//
// Launches (all transforms in LDS, Stockham, any 2-3-5-smooth length):
//   gamma : spectrum of g on [-q, q]^2 (the generic spectrum kernels), then
//           k_adj_cfield + k_adj_rfield (real part) on the coarse grid
//   k_adj_cfield  column stage of e_k:  Ec[b][k][t][kc]
//   k_adj_rfield  row stage of e_k, times gamma, conjugated, forward row DFT
//                 kept at window columns (in place over Ec)
//   k_adj_cwin    forward column DFT kept at window rows, times phi_k -> Ak
//   k_adj_ksum    Q = sum_k Ak
//   k_adj_gcols / k_adj_grows  window -> image grid, real part, two rows per
//                 complex transform (Hermitian packing)
// On image sizes with register transforms the last three are replaced by
// k_adj_fold (sum over k and the Hermitian part, in the band layout of the
// forward interpolation) and the forward's k_fp_ucols / k_fp_urows.
#include "oi_adjoint.h"
#include "oi_fastpath.h"
#include "oi_fftreg.h"

namespace oi {

static int lds_items(int n, int cap) {
  int v = 4096 / (n > 0 ? n : 1);
  if (v > cap) v = cap;
  return v < 1 ? 1 : v;
}

template <typename K>
static int allow_lds_adj(K kernel, size_t bytes) {
  if (bytes > kLdsLimit) return fail(OI_ERR_UNSUPPORTED, "adjoint: shared-memory request exceeds 160 KiB");
  if (bytes > 64 * 1024)
    OI_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return OI_OK;
}

// Ec[b][k][t][kc] = sum_ir A(ir, kc) exp(-2 pi i (k0r + ir) t / lr),
// A = kern[k][ir][kc] spec[b][ir][kc] (kern == nullptr: A = spec).
__global__ __launch_bounds__(kThreads) void k_adj_cfield(const float2* __restrict__ spec,
                                                         const float2* __restrict__ kern, int nk,
                                                         int k0r, int wr, int wc, int lr, int ncol,
                                                         FftDesc flr, float2* __restrict__ Ec) {
  extern __shared__ float2 lds[];
  float2* a = lds;
  float2* t = lds + ncol * lr;
  const int j0 = blockIdx.x * ncol, k = blockIdx.y, b = blockIdx.z;
  const float2* sb = spec + (size_t)b * wr * wc;
  const float2* kb = kern ? kern + (size_t)k * wr * wc : nullptr;
  for (int e = threadIdx.x; e < ncol * lr; e += blockDim.x) {
    const int u = e / ncol, col = e - u * ncol;
    const int kc = j0 + col;
    float2 sum = make_float2(0.f, 0.f);
    if (kc < wc)
      for (int ir = pmod(u - k0r, lr); ir < wr; ir += lr) {
        float2 v = sb[(size_t)ir * wc + kc];
        if (kb) v = cmul(v, kb[(size_t)ir * wc + kc]);
        sum = cadd(sum, v);
      }
    a[col * lr + u] = sum;
  }
  __syncthreads();
  const float2* z = block_fft<false>(a, t, flr, ncol, lr);
  float2* out = Ec + ((size_t)b * nk + k) * lr * wc;
  for (int e = threadIdx.x; e < ncol * lr; e += blockDim.x) {
    const int tr = e / ncol, col = e - tr * ncol;
    if (j0 + col < wc) out[(size_t)tr * wc + j0 + col] = z[col * lr + tr];
  }
}

// Row stage.  Rows r of Ec[b][k] (length wc, index kc -> kappa = k0c + kc):
//   e[x] = sum_kc Ec[r][kc] exp(-2 pi i (k0c + kc) x / lc)
// gamma_in == nullptr: gamma_out[b][r][x] = Re e[x] (nk must be 1).
// Otherwise p[x] = gamma_in[b][r][x] conj(e[x]) and, in place,
//   Ec[r][kc] = sum_x p[x] exp(-2 pi i (k0c + kc) x / lc).
__global__ __launch_bounds__(kThreads) void k_adj_rfield(float2* __restrict__ Ec, int nk, int k0c,
                                                         int wc, int lr, int lc, int nrow,
                                                         FftDesc flc,
                                                         const float* __restrict__ gamma_in,
                                                         float* __restrict__ gamma_out) {
  extern __shared__ float2 lds[];
  float2* a = lds;
  float2* t = lds + nrow * lc;
  const int r0 = blockIdx.x * nrow, k = blockIdx.y, b = blockIdx.z;
  float2* Eb = Ec + ((size_t)b * nk + k) * lr * wc;
  for (int e = threadIdx.x; e < nrow * lc; e += blockDim.x) {
    const int i = e / lc, x = e - i * lc;
    const int r = r0 + i;
    float2 sum = make_float2(0.f, 0.f);
    if (r < lr)
      for (int kc = pmod(x - k0c, lc); kc < wc; kc += lc) sum = cadd(sum, Eb[(size_t)r * wc + kc]);
    a[e] = sum;
  }
  __syncthreads();
  float2* z = block_fft<false>(a, t, flc, nrow, lc);
  if (!gamma_in) {
    for (int e = threadIdx.x; e < nrow * lc; e += blockDim.x) {
      const int i = e / lc, x = e - i * lc;
      const int r = r0 + i;
      if (r < lr) gamma_out[((size_t)b * lr + r) * lc + x] = z[e].x;
    }
    return;
  }
  float2* o = (z == a) ? t : a;
  const float* gb = gamma_in + (size_t)b * lr * lc;
  for (int e = threadIdx.x; e < nrow * lc; e += blockDim.x) {
    const int i = e / lc, x = e - i * lc;
    const int r = r0 + i;
    const float g = r < lr ? gb[(size_t)r * lc + x] : 0.f;
    o[e] = make_float2(g * z[e].x, -g * z[e].y);
  }
  __syncthreads();
  const float2* y = block_fft<false>(o, z, flc, nrow, lc);
  for (int e = threadIdx.x; e < nrow * wc; e += blockDim.x) {
    const int i = e / wc, kc = e - i * wc;
    const int r = r0 + i;
    if (r < lr) Eb[(size_t)r * wc + kc] = y[i * lc + pmod(k0c + kc, lc)];
  }
}

// Ak[b][k][kr][kc] = phi_k(kr, kc) sum_t P[t][kc] exp(-2 pi i (k0r + kr) t / lr).
__global__ __launch_bounds__(kThreads) void k_adj_cwin(const float2* __restrict__ P,
                                                       const float2* __restrict__ kern, int nk,
                                                       int k0r, int wr, int wc, int lr, int ncol,
                                                       FftDesc flr, float2* __restrict__ Ak) {
  extern __shared__ float2 lds[];
  float2* a = lds;
  float2* t = lds + ncol * lr;
  const int j0 = blockIdx.x * ncol, k = blockIdx.y, b = blockIdx.z;
  const float2* pb = P + ((size_t)b * nk + k) * lr * wc;
  for (int e = threadIdx.x; e < ncol * lr; e += blockDim.x) {
    const int u = e / ncol, col = e - u * ncol;
    a[col * lr + u] = (j0 + col < wc) ? pb[(size_t)u * wc + j0 + col] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const float2* z = block_fft<false>(a, t, flr, ncol, lr);
  const float2* kb = kern + (size_t)k * wr * wc;
  float2* out = Ak + ((size_t)b * nk + k) * wr * wc;
  for (int e = threadIdx.x; e < wr * ncol; e += blockDim.x) {
    const int kr = e / ncol, col = e - kr * ncol;
    const int kc = j0 + col;
    if (kc < wc) out[(size_t)kr * wc + kc] = cmul(kb[(size_t)kr * wc + kc], z[col * lr + pmod(k0r + kr, lr)]);
  }
}

// ---------------------------------------------------- register transforms
// The three coarse-grid stages above with compile-time transform sizes (the
// coarse sizes of the fast imaging path): a group of P::T lanes holds one
// coarse row or window column, x[t + T m] in registers, one LDS exchange per
// transform.  Same arithmetic as the LDS kernels; window indices fold onto
// the coarse grid without collisions (wr <= lr, wc <= lc).
constexpr int kRegBlock = 256;

template <class P>
__global__ __launch_bounds__(kRegBlock) void k_adj_cfield_reg(const float2* __restrict__ spec,
                                                              const float2* __restrict__ kern,
                                                              int nk, int k0r, int wr, int wc,
                                                              const float2* __restrict__ twt,
                                                              float2* __restrict__ Ec) {
  constexpr int L = P::N, T = P::T, V = P::V, G = kRegBlock / T;
  extern __shared__ float2 lds[];
  const int g = threadIdx.x / T, t = threadIdx.x - g * T;
  const int k = blockIdx.y, b = blockIdx.z;
  const int kc = blockIdx.x * G + g;
  const bool on = g < G && kc < wc;
  const float2* sb = spec + (size_t)b * wr * wc + (on ? kc : 0);
  const float2* kb = kern ? kern + (size_t)k * wr * wc + (on ? kc : 0) : nullptr;
  float2 v[V];
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const int ir = pmod(t + T * m - k0r, L);
    float2 a = make_float2(0.f, 0.f);
    if (on && ir < wr) {
      a = sb[(size_t)ir * wc];
      // kernel x spectrum as the forward's k_fp_bcols forms it, so that Ec
      // holds the same bits as its column buffer
      if (kb) a = reg::mul_rn(kb[(size_t)ir * wc], a);
    }
    v[m] = a;
  }
  P::template run_live_tab<false>(v, lds + (g < G ? g : 0) * P::LDS, t, twt, g < G);
  if (!on) return;
  float2* out = Ec + ((size_t)b * nk + k) * L * wc + kc;
#pragma unroll
  for (int m = 0; m < V; ++m) out[(size_t)(t + T * m) * wc] = v[m];
}

template <class P>
__global__ __launch_bounds__(kRegBlock) void k_adj_rfield_reg(float2* __restrict__ Ec, int nk,
                                                              int k0c, int wc, int lr,
                                                              const float2* __restrict__ twt,
                                                              const float* __restrict__ gamma_in,
                                                              float* __restrict__ gamma_out) {
  constexpr int L = P::N, T = P::T, V = P::V, G = kRegBlock / T;
  extern __shared__ float2 lds[];
  const int g = threadIdx.x / T, t = threadIdx.x - g * T;
  const int k = blockIdx.y, b = blockIdx.z;
  const int r = blockIdx.x * G + g;
  const bool on = g < G && r < lr;
  float2* scr = lds + (g < G ? g : 0) * P::LDS;
  float2* row = Ec + (((size_t)b * nk + k) * lr + (on ? r : 0)) * wc;
  float2 v[V];
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const int kc = pmod(t + T * m - k0c, L);
    v[m] = (on && kc < wc) ? row[kc] : make_float2(0.f, 0.f);
  }
  P::template run_live_tab<false>(v, scr, t, twt, g < G);
  if (!gamma_in) {
    if (on) {
      float* go = gamma_out + ((size_t)b * lr + r) * L;
#pragma unroll
      for (int m = 0; m < V; ++m) go[t + T * m] = v[m].x;
    }
    return;
  }
  const float* gi = gamma_in + ((size_t)b * lr + (on ? r : 0)) * L;
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const float gm = on ? gi[t + T * m] : 0.f;
    v[m] = make_float2(gm * v[m].x, -gm * v[m].y);
  }
  P::template run_live_tab<false>(v, scr, t, twt, g < G);
  if (!on) return;
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const int kc = pmod(t + T * m - k0c, L);
    if (kc < wc) row[kc] = v[m];
  }
}

// k_adj_rfield_reg (with gamma) taking the column stage from the forward
// pass's two-pass buffer Bc[b][k][kc][t] (k_fp_bcols: the same values as
// k_adj_cfield_reg's Ec, column-major) instead of recomputing it.  A block
// stages rows r0 .. r0 + G - 1 of every window column through LDS (one
// contiguous run of G values per column), runs the row stage and writes the
// result to Ec[b][k][r][kc] as k_adj_rfield_reg does.  The staging tile and
// the transform scratch share the LDS (the larger of the two).
template <class P>
__global__ __launch_bounds__(kRegBlock) void k_adj_rfield_bcols(const float2* __restrict__ Bc,
                                                                float2* __restrict__ Ec, int nk,
                                                                int k0c, int wc, int lr,
                                                                const float2* __restrict__ twt,
                                                                const float* __restrict__ gamma_in) {
  constexpr int L = P::N, T = P::T, V = P::V, G = kRegBlock / T;
  constexpr int TS = G + 1;
  extern __shared__ float2 lds[];
  float2* tile = lds;                              // [wc][TS], then the scratch over it
  float2* scr = lds;
  const int g = threadIdx.x / T, t = threadIdx.x - g * T;
  const int gg = g < G ? g : 0;
  const int k = blockIdx.y, b = blockIdx.z;
  const int r0 = blockIdx.x * G;
  const int r = r0 + gg;
  const bool on = g < G && r < lr;
  const float2* src = Bc + ((size_t)b * nk + k) * wc * lr + r0;
  for (int e = threadIdx.x; e < wc * G; e += kRegBlock) {
    const int kc = e / G, rr = e - kc * G;
    tile[kc * TS + rr] = r0 + rr < lr ? src[(size_t)kc * lr + rr] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  float2 v[V];
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const int kc = pmod(t + T * m - k0c, L);
    v[m] = (on && kc < wc) ? tile[kc * TS + gg] : make_float2(0.f, 0.f);
  }
  __syncthreads();                                 // the tile is read: the scratch reuses it
  float2* sg = scr + (size_t)gg * P::LDS;
  P::template run_live_tab<false>(v, sg, t, twt, g < G);
  const float* gi = gamma_in + ((size_t)b * lr + (on ? r : 0)) * L;
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const float gm = on ? gi[t + T * m] : 0.f;
    v[m] = make_float2(gm * v[m].x, -gm * v[m].y);
  }
  P::template run_live_tab<false>(v, sg, t, twt, g < G);
  if (!on) return;
  float2* row = Ec + (((size_t)b * nk + k) * lr + r) * wc;
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const int kc = pmod(t + T * m - k0c, L);
    if (kc < wc) row[kc] = v[m];
  }
}

template <class P>
__global__ __launch_bounds__(kRegBlock) void k_adj_cwin_reg(const float2* __restrict__ Pc,
                                                            const float2* __restrict__ kern, int nk,
                                                            int k0r, int wr, int wc,
                                                            const float2* __restrict__ twt,
                                                            float2* __restrict__ Ak, int batch) {
  constexpr int L = P::N, T = P::T, V = P::V, G = kRegBlock / T;
  extern __shared__ float2 lds[];
  // items (column block, kernel, image), column blocks fastest, in XCD order:
  // a row of Pc is split over the column blocks of one (image, kernel), and
  // the lines at their edges are read by two of them
  const int nxb = (wc + G - 1) / G;
  int w;
  if (!xcd_item(blockIdx.x, nxb * nk * batch, w)) return;   // uniform over the block
  const int xb = w % nxb, k = (w / nxb) % nk, b = w / (nxb * nk);
  const int g = threadIdx.x / T, t = threadIdx.x - g * T;
  const int kc = xb * G + g;
  const bool on = g < G && kc < wc;
  const float2* pb = Pc + ((size_t)b * nk + k) * L * wc + (on ? kc : 0);
  float2 v[V];
#pragma unroll
  for (int m = 0; m < V; ++m) v[m] = on ? pb[(size_t)(t + T * m) * wc] : make_float2(0.f, 0.f);
  P::template run_live_tab<false>(v, lds + (g < G ? g : 0) * P::LDS, t, twt, g < G);
  if (!on) return;
  const float2* kb = kern + (size_t)k * wr * wc + kc;
  float2* out = Ak + ((size_t)b * nk + k) * wr * wc + kc;
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const int kr = pmod(t + T * m - k0r, L);
    if (kr < wr) out[(size_t)kr * wc] = cmul(kb[(size_t)kr * wc], v[m]);
  }
}

// Coarse sizes with register transforms: L = R1 x R2 (the smaller radix
// dividing the larger), as the fast imaging path's coarse grids.
#define OI_ADJ_PLANS(X)                                                                        \
  X(16, 4, 4) X(25, 5, 5) X(32, 4, 8) X(36, 6, 6) X(48, 4, 12) X(64, 8, 8) X(75, 5, 15)       \
  X(81, 9, 9) X(100, 10, 10) X(128, 8, 16) X(144, 12, 12) X(256, 16, 16) X(450, 15, 30)     \
  X(512, 16, 32)

static bool adj_reg_size(int L) {
#define X(NN, A, B) if (L == NN) return true;
  OI_ADJ_PLANS(X)
#undef X
  return false;
}

template <class P>
static size_t adj_reg_lds() {
  return (size_t)(kRegBlock / P::T) * P::LDS * sizeof(float2);
}

// Column stage (kern == nullptr: gamma's spectrum) with register transforms.
static int adj_cfield_reg(int L, dim3 grid_y_z, const float2* spec, const float2* kern, int nk,
                          int k0r, int wr, int wc, const float2* twt, float2* Ec, hipStream_t st) {
#define X(NN, A, B)                                                                            \
  if (L == NN) {                                                                               \
    using P = reg::GFFT<NN, A, B>;                                                             \
    constexpr int G = kRegBlock / P::T;                                                        \
    int rc = allow_lds_adj(k_adj_cfield_reg<P>, adj_reg_lds<P>());                             \
    if (rc) return rc;                                                                         \
    dim3 grid((wc + G - 1) / G, grid_y_z.y, grid_y_z.z);                                       \
    hipLaunchKernelGGL(k_adj_cfield_reg<P>, grid, dim3(kRegBlock), adj_reg_lds<P>(), st, spec, \
                       kern, nk, k0r, wr, wc, twt, Ec);                                        \
    OI_LAUNCH_CHECK("adj_cfield_reg");                                                         \
    return OI_OK;                                                                              \
  }
  OI_ADJ_PLANS(X)
#undef X
  return fail(OI_ERR_UNSUPPORTED, "adjoint: no register transform for this coarse size");
}

static int adj_rfield_reg(int L, dim3 grid_y_z, float2* Ec, int nk, int k0c, int wc, int lr,
                          const float2* twt, const float* gamma_in, float* gamma_out,
                          hipStream_t st) {
#define X(NN, A, B)                                                                            \
  if (L == NN) {                                                                               \
    using P = reg::GFFT<NN, A, B>;                                                             \
    constexpr int G = kRegBlock / P::T;                                                        \
    int rc = allow_lds_adj(k_adj_rfield_reg<P>, adj_reg_lds<P>());                             \
    if (rc) return rc;                                                                         \
    dim3 grid((lr + G - 1) / G, grid_y_z.y, grid_y_z.z);                                       \
    hipLaunchKernelGGL(k_adj_rfield_reg<P>, grid, dim3(kRegBlock), adj_reg_lds<P>(), st, Ec,   \
                       nk, k0c, wc, lr, twt, gamma_in, gamma_out);                             \
    OI_LAUNCH_CHECK("adj_rfield_reg");                                                         \
    return OI_OK;                                                                              \
  }
  OI_ADJ_PLANS(X)
#undef X
  return fail(OI_ERR_UNSUPPORTED, "adjoint: no register transform for this coarse size");
}

// Sizes of the forward's two-pass column buffer.
#define OI_ADJ_BCOLS_PLANS(X) X(144, 12, 12) X(256, 16, 16) X(450, 15, 30) X(512, 16, 32)

static bool adj_bcols_size(int L) {
#define X(NN, A, B) if (L == NN) return true;
  OI_ADJ_BCOLS_PLANS(X)
#undef X
  return false;
}

static int adj_rfield_bcols(int L, dim3 grid_y_z, const float2* Bc, float2* Ec, int nk, int k0c,
                            int wc, int lr, const float2* twt, const float* gamma_in, hipStream_t st) {
#define X(NN, A, B)                                                                            \
  if (L == NN) {                                                                               \
    using P = reg::GFFT<NN, A, B>;                                                             \
    constexpr int G = kRegBlock / P::T;                                                        \
    const size_t tile = (size_t)wc * (G + 1), scr = (size_t)G * P::LDS;                        \
    const size_t lds = (tile > scr ? tile : scr) * sizeof(float2);                             \
    int rc = allow_lds_adj(k_adj_rfield_bcols<P>, lds);                                        \
    if (rc) return rc;                                                                         \
    dim3 grid((lr + G - 1) / G, grid_y_z.y, grid_y_z.z);                                       \
    hipLaunchKernelGGL(k_adj_rfield_bcols<P>, grid, dim3(kRegBlock), lds, st, Bc, Ec, nk, k0c, \
                       wc, lr, twt, gamma_in);                                                 \
    OI_LAUNCH_CHECK("adj_rfield_bcols");                                                       \
    return OI_OK;                                                                              \
  }
  OI_ADJ_BCOLS_PLANS(X)
#undef X
  return fail(OI_ERR_UNSUPPORTED, "adjoint: no two-pass column buffer for this coarse size");
}

static int adj_cwin_reg(int L, dim3 grid_y_z, const float2* Pc, const float2* kern, int nk,
                        int k0r, int wr, int wc, const float2* twt, float2* Ak, hipStream_t st) {
#define X(NN, A, B)                                                                            \
  if (L == NN) {                                                                               \
    using P = reg::GFFT<NN, A, B>;                                                             \
    constexpr int G = kRegBlock / P::T;                                                        \
    int rc = allow_lds_adj(k_adj_cwin_reg<P>, adj_reg_lds<P>());                               \
    if (rc) return rc;                                                                         \
    const int items = ((wc + G - 1) / G) * (int)grid_y_z.y * (int)grid_y_z.z;                  \
    hipLaunchKernelGGL(k_adj_cwin_reg<P>, dim3(xcd_grid(items)), dim3(kRegBlock),              \
                       adj_reg_lds<P>(), st, Pc, kern, nk, k0r, wr, wc, twt, Ak,               \
                       (int)grid_y_z.z);                                                       \
    OI_LAUNCH_CHECK("adj_cwin_reg");                                                           \
    return OI_OK;                                                                              \
  }
  OI_ADJ_PLANS(X)
#undef X
  return fail(OI_ERR_UNSUPPORTED, "adjoint: no register transform for this coarse size");
}

// Q[b][i] = sum_k Ak[b][k][i], i over the window.
__global__ void k_adj_ksum(const float2* __restrict__ Ak, int nk, int per, int batch,
                           float2* __restrict__ Q) {
  const size_t n = (size_t)batch * per;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t b = e / per, i = e - b * per;
    const float2* src = Ak + b * nk * per + i;
    float2 s = make_float2(0.f, 0.f);
    for (int k = 0; k < nk; ++k) s = cadd(s, src[(size_t)k * per]);
    Q[e] = s;
  }
}

// The gradient's spectrum folded for the forward interpolation kernels
// (k_fp_ucols / k_fp_urows, which sum F[q][qr] e^{-2 pi i (qr y / ny + q x / nx)}
// over the Hermitian extension F[-q][-qr] = conj F[q][qr]):
//   grad = s Re sum Qw(kr, kc) e^{+2 pi i (kr y / ny + kc x / nx)}
//   F[b][q][qr + qrm] = s/2 (conj Qw(qr, q) + Qw(-qr, -q)),  q in [0, qcm], |qr| <= qrm,
// with Qw(kr, kc) = sum_k Ak[b][k] at window orders (kr, kc), zero outside.
__global__ void k_adj_fold(const float2* __restrict__ Ak, int nk, int k0r, int wr, int k0c, int wc,
                           int qrm, int qcm, float s, int batch, float2* __restrict__ F) {
  const int nb = 2 * qrm + 1, nq = qcm + 1;
  const size_t per = (size_t)nq * nb, n = (size_t)batch * per, win = (size_t)wr * wc;
  // q runs fastest over the threads, so both window reads walk along a row
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t b = e / per, rem = e - b * per;
    const int i = (int)(rem / nq), q = (int)(rem - (size_t)i * nq), qr = i - qrm;
    const float2* ab = Ak + b * nk * win;
    float2 a = make_float2(0.f, 0.f), c = make_float2(0.f, 0.f);
    int r = qr - k0r, cc = q - k0c;
    if (r >= 0 && r < wr && cc >= 0 && cc < wc) {
#pragma unroll 4
      for (int k = 0; k < nk; ++k) a = cadd(a, ab[k * win + (size_t)r * wc + cc]);
    }
    r = -qr - k0r;
    cc = -q - k0c;
    if (r >= 0 && r < wr && cc >= 0 && cc < wc) {
#pragma unroll 4
      for (int k = 0; k < nk; ++k) c = cadd(c, ab[k * win + (size_t)r * wc + cc]);
    }
    F[(b * nq + q) * nb + i] = make_float2(0.5f * s * (a.x + c.x), 0.5f * s * (c.y - a.y));
  }
}

// Z[b][kc][y] = sum_kr Q[b][kr][kc] exp(+2 pi i (k0r + kr) y / ny)   (wr <= ny)
__global__ __launch_bounds__(kThreads) void k_adj_gcols(const float2* __restrict__ Q, int k0r,
                                                        int wr, int wc, int ny, int ncol,
                                                        FftDesc fy, float2* __restrict__ Z) {
  extern __shared__ float2 lds[];
  float2* a = lds;
  float2* t = lds + ncol * ny;
  const int j0 = blockIdx.x * ncol, b = blockIdx.y;
  for (int e = threadIdx.x; e < ncol * ny; e += blockDim.x) a[e] = make_float2(0.f, 0.f);
  __syncthreads();
  const float2* qb = Q + (size_t)b * wr * wc;
  for (int e = threadIdx.x; e < wr * ncol; e += blockDim.x) {
    const int kr = e / ncol, col = e - kr * ncol;
    if (j0 + col < wc) a[col * ny + pmod(k0r + kr, ny)] = qb[(size_t)kr * wc + j0 + col];
  }
  __syncthreads();
  const float2* z = block_fft<true>(a, t, fy, ncol, ny);
  float2* zb = Z + (size_t)b * wc * ny;
  for (int e = threadIdx.x; e < ncol * ny; e += blockDim.x) {
    const int col = e / ny, y = e - col * ny;
    if (j0 + col < wc) zb[(size_t)(j0 + col) * ny + y] = z[e];
  }
}

// grad[b][y][x] = scale Re sum_kc Z[b][kc][y] exp(+2 pi i (k0c + kc) x / nx)   (wc <= nx)
// Rows y1 = 2 p, y2 = 2 p + 1 share one transform: with h(kappa) =
// (z(kappa) + conj z(-kappa)) / 2 (Hermitian, transform = Re of z's),
// transform(h1 + i h2) = Re1 + i Re2.
__global__ __launch_bounds__(kThreads) void k_adj_grows(const float2* __restrict__ Z, int k0c,
                                                        int wc, int ny, int nx, int npair,
                                                        FftDesc fx, float scale,
                                                        float* __restrict__ grad) {
  extern __shared__ float2 lds[];
  float2* z1 = lds;                  // [npair][nx]
  float2* z2 = lds + npair * nx;     // [npair][nx]
  float2* v = lds + 2 * npair * nx;  // [npair][nx]
  const int p0 = blockIdx.x * npair, b = blockIdx.y;
  for (int e = threadIdx.x; e < 2 * npair * nx; e += blockDim.x) lds[e] = make_float2(0.f, 0.f);
  __syncthreads();
  const float2* zb = Z + (size_t)b * wc * ny;
  for (int e = threadIdx.x; e < npair * 2 * wc; e += blockDim.x) {
    const int kc = e / (2 * npair), r = e - kc * 2 * npair;
    const int i = r >> 1, y = 2 * (p0 + i) + (r & 1);
    if (y >= ny) continue;
    float2* dst = (r & 1) ? z2 : z1;
    dst[i * nx + pmod(k0c + kc, nx)] = zb[(size_t)kc * ny + y];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < npair * nx; e += blockDim.x) {
    const int i = e / nx, m = e - i * nx;
    const int mm = i * nx + (m == 0 ? 0 : nx - m);
    const float2 a1 = z1[e], c1 = z1[mm], a2 = z2[e], c2 = z2[mm];
    const float2 h1 = make_float2(0.5f * (a1.x + c1.x), 0.5f * (a1.y - c1.y));
    const float2 h2 = make_float2(0.5f * (a2.x + c2.x), 0.5f * (a2.y - c2.y));
    v[e] = make_float2(h1.x - h2.y, h1.y + h2.x);
  }
  __syncthreads();
  const float2* o = block_fft<true>(v, z1, fx, npair, nx);
  for (int e = threadIdx.x; e < npair * nx; e += blockDim.x) {
    const int i = e / nx, x = e - i * nx;
    const int y1 = 2 * (p0 + i), y2 = y1 + 1;
    float* gb = grad + (size_t)b * ny * nx;
    if (y1 < ny) gb[(size_t)y1 * nx + x] = scale * o[e].x;
    if (y2 < ny) gb[(size_t)y2 * nx + x] = scale * o[e].y;
  }
}

// ------------------------------------------------------------- launchers
// largest |order| of a window [k0, k0 + w - 1]
static int adj_half(int k0, int w) {
  const int a = -k0, b = k0 + w - 1;
  return a > b ? a : b;
}

size_t adjoint_work_bytes(const AdjGeom& g, int batch, int nkern) {
  const size_t B = (size_t)batch, K = (size_t)nkern;
  const size_t gq = g.direct ? 0 : (size_t)(2 * g.qr + 1) * (2 * g.qc + 1);
  size_t n = 0;
  n += g.direct ? 0 : B * g.ny * (2 * g.qc + 1) * sizeof(float2);  // Tg
  n += B * gq * sizeof(float2);                                      // spec_g
  n += g.direct ? 0 : B * g.lr * (2 * g.qc + 1) * sizeof(float2);   // Ecg
  n += g.direct ? 0 : B * g.lr * g.lc * sizeof(float);              // gamma
  n += B * K * g.lr * g.wc * sizeof(float2);                         // Ec / P
  n += B * K * g.wr * g.wc * sizeof(float2);                         // Ak
  n += B * g.wr * g.wc * sizeof(float2);                             // Q
  n += B * g.wc * g.ny * sizeof(float2);                             // Z
  const size_t qrm = adj_half(g.k0r, g.wr), qcm = adj_half(g.k0c, g.wc);
  n += B * (qcm + 1) * (2 * qrm + 1) * sizeof(float2);               // F
  n += B * (qcm + 1) * g.ny * sizeof(float2);                        // U
  return n + 10 * 256;
}

static char* carve(char*& p, size_t bytes) {
  char* r = p;
  p += (bytes + 255) & ~(size_t)255;
  return r;
}

float2* adjoint_cotangent_spectrum(const AdjGeom& g, int batch, void* work) {
  if (g.direct) return nullptr;
  char* w = static_cast<char*>(work);
  (void)carve(w, (size_t)batch * g.ny * (2 * g.qc + 1) * sizeof(float2));   // Tg
  return (float2*)w;
}

int adjoint_band_spectrum(const AdjGeom& g, const float* img, int batch, void* work, float2* out,
                          hipStream_t st) {
  if (g.direct) return fail(OI_ERR_INVALID, "adjoint: no band spectrum on a direct plan");
  const int wgc = 2 * g.qc + 1, wgr = 2 * g.qr + 1;
  float2* Tg = static_cast<float2*>(work);     // the first region of the work buffer
  if (fp_image_size_ok(g.ny) && fp_image_size_ok(g.nx)) {
    // register transforms of the image sizes (same spectrum, scale and layout)
    FpArgs a{};
    a.field = img;
    a.kind = OI_FIELD_REAL;
    a.cin = 1;
    a.nchan = 1;
    a.chans.c[0] = 0;
    a.batch = batch;
    a.ny = g.ny;
    a.nx = g.nx;
    a.k0r = -g.qr;
    a.wr = wgr;
    a.k0c = -g.qc;
    a.wc = wgc;
    a.tw_x = g.fx.tw;
    a.tw_y = g.fy.tw;
    a.T = Tg;
    a.spec = out;
    return fp_spectrum(a, st);
  }
  const int chan0 = 0;
  int rc = launch_spec_rows(img, OI_FIELD_REAL, 1, &chan0, 1, batch, g.ny, g.nx, g.fx, -g.qc, wgc,
                            Tg, st);
  if (rc) return rc;
  return launch_spec_cols(Tg, batch, g.ny, g.nx, g.fy, -g.qr, wgr, wgc, out, st);
}

bool adjoint_chain_ok(const AdjGeom& g) {
  const int qrm = adj_half(g.k0r, g.wr), qcm = adj_half(g.k0c, g.wc);
  return fp_image_size_ok(g.ny) && fp_image_size_ok(g.nx) && 2 * qrm + 1 <= g.ny &&
         2 * qcm + 1 <= g.nx;
}

int launch_adjoint(const AdjGeom& g, const float* gI, const float2* spec, int batch,
                   const float2* kern, int nkern, void* work, float* grad, hipStream_t st,
                   const float2* bcols, const float* chain_theta, float* chain_part) {
  if (!gI && g.direct) return fail(OI_ERR_INVALID, "adjoint: a direct plan needs the cotangent image");
  if (chain_theta && (!chain_part || !adjoint_chain_ok(g)))
    return fail(OI_ERR_UNSUPPORTED, "adjoint: the chain-rule form needs the compile-time-size image pass");
  const size_t B = (size_t)batch, K = (size_t)nkern;
  char* w = static_cast<char*>(work);
  const int wgc = 2 * g.qc + 1, wgr = 2 * g.qr + 1;
  if (!g.direct) (void)carve(w, B * g.ny * wgc * sizeof(float2));   // Tg
  float2* specg = g.direct ? nullptr : (float2*)carve(w, B * wgr * wgc * sizeof(float2));
  float2* Ecg = g.direct ? nullptr : (float2*)carve(w, B * g.lr * wgc * sizeof(float2));
  float* gamma = g.direct ? nullptr : (float*)carve(w, B * g.lr * g.lc * sizeof(float));
  float2* Ec = (float2*)carve(w, B * K * g.lr * g.wc * sizeof(float2));
  float2* Ak = (float2*)carve(w, B * K * g.wr * g.wc * sizeof(float2));
  float2* Q = (float2*)carve(w, B * g.wr * g.wc * sizeof(float2));
  float2* Z = (float2*)carve(w, B * g.wc * g.ny * sizeof(float2));
  const int qrm = adj_half(g.k0r, g.wr), qcm = adj_half(g.k0c, g.wc);
  float2* F = (float2*)carve(w, B * (qcm + 1) * (2 * qrm + 1) * sizeof(float2));
  float2* U = (float2*)carve(w, B * (qcm + 1) * g.ny * sizeof(float2));

  // register transforms on the coarse grid when both sizes are tabled and
  // the window (and gamma's band) fold onto it without collisions
  const bool reg = !g.direct && adj_reg_size(g.lr) && adj_reg_size(g.lc) && g.wr <= g.lr &&
                   g.wc <= g.lc && 2 * g.qr + 1 <= g.lr && 2 * g.qc + 1 <= g.lc;
  const int ncol_r = lds_items(g.lr, 64);
  const int nrow_c = lds_items(g.lc, 16);
  const size_t lds_r = 2ull * ncol_r * g.lr * sizeof(float2);
  const size_t lds_c = 2ull * nrow_c * g.lc * sizeof(float2);
  int rc;
  if ((rc = allow_lds_adj(k_adj_cfield, lds_r))) return rc;
  if ((rc = allow_lds_adj(k_adj_rfield, lds_c))) return rc;
  if ((rc = allow_lds_adj(k_adj_cwin, lds_r))) return rc;

  if (!g.direct) {
    // gamma: the cotangent's spectrum on [-q, q]^2 taken to the coarse grid
    // (null gI: the caller left that spectrum in the work buffer)
    if (gI) {
      rc = adjoint_band_spectrum(g, gI, batch, work, specg, st);
      if (rc) return rc;
    }
    if (reg) {
      if ((rc = adj_cfield_reg(g.lr, dim3(1, 1, batch), specg, nullptr, 1, -g.qr, wgr, wgc,
                               g.flr.tw, Ecg, st)))
        return rc;
      if ((rc = adj_rfield_reg(g.lc, dim3(1, 1, batch), Ecg, 1, -g.qc, wgc, g.lr, g.flc.tw,
                               nullptr, gamma, st)))
        return rc;
    } else {
      dim3 g1((wgc + ncol_r - 1) / ncol_r, 1, batch);
      hipLaunchKernelGGL(k_adj_cfield, g1, dim3(kThreads), lds_r, st, specg, nullptr, 1, -g.qr,
                         wgr, wgc, g.lr, ncol_r, g.flr, Ecg);
      OI_LAUNCH_CHECK("adj_cfield(gamma)");
      dim3 g2((g.lr + nrow_c - 1) / nrow_c, 1, batch);
      hipLaunchKernelGGL(k_adj_rfield, g2, dim3(kThreads), lds_c, st, Ecg, 1, -g.qc, wgc, g.lr,
                         g.lc, nrow_c, g.flc, nullptr, gamma);
      OI_LAUNCH_CHECK("adj_rfield(gamma)");
    }
  }
  const float* gam = g.direct ? gI : gamma;

  if (reg && bcols && !g.direct && g.lr == g.lc && adj_bcols_size(g.lr)) {
    // the column stage is the forward pass's column buffer
    if ((rc = adj_rfield_bcols(g.lc, dim3(1, nkern, batch), bcols, Ec, nkern, g.k0c, g.wc, g.lr,
                               g.flc.tw, gam, st)))
      return rc;
    if ((rc = adj_cwin_reg(g.lr, dim3(1, nkern, batch), Ec, kern, nkern, g.k0r, g.wr, g.wc,
                           g.flr.tw, Ak, st)))
      return rc;
  } else if (reg) {
    if ((rc = adj_cfield_reg(g.lr, dim3(1, nkern, batch), spec, kern, nkern, g.k0r, g.wr, g.wc,
                             g.flr.tw, Ec, st)))
      return rc;
    if ((rc = adj_rfield_reg(g.lc, dim3(1, nkern, batch), Ec, nkern, g.k0c, g.wc, g.lr, g.flc.tw,
                             gam, nullptr, st)))
      return rc;
    if ((rc = adj_cwin_reg(g.lr, dim3(1, nkern, batch), Ec, kern, nkern, g.k0r, g.wr, g.wc,
                           g.flr.tw, Ak, st)))
      return rc;
  } else {
    dim3 g3((g.wc + ncol_r - 1) / ncol_r, nkern, batch);
    hipLaunchKernelGGL(k_adj_cfield, g3, dim3(kThreads), lds_r, st, spec, kern, nkern, g.k0r, g.wr,
                       g.wc, g.lr, ncol_r, g.flr, Ec);
    OI_LAUNCH_CHECK("adj_cfield");
    dim3 g4((g.lr + nrow_c - 1) / nrow_c, nkern, batch);
    hipLaunchKernelGGL(k_adj_rfield, g4, dim3(kThreads), lds_c, st, Ec, nkern, g.k0c, g.wc, g.lr,
                       g.lc, nrow_c, g.flc, gam, nullptr);
    OI_LAUNCH_CHECK("adj_rfield");
    hipLaunchKernelGGL(k_adj_cwin, g3, dim3(kThreads), lds_r, st, Ec, kern, nkern, g.k0r, g.wr,
                       g.wc, g.lr, ncol_r, g.flr, Ak);
    OI_LAUNCH_CHECK("adj_cwin");
  }
  const float scale = 2.0f / ((float)g.lr * (float)g.lc);
  if (adjoint_chain_ok(g)) {
    // register transforms of the image sizes: the kernel sum folded into the
    // band layout of the forward interpolation, then its two kernels
    const size_t nf = B * (qcm + 1) * (2 * qrm + 1);
    const size_t nblk = (nf + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(k_adj_fold, dim3((unsigned)(nblk < 4096 ? nblk : 4096)), dim3(kThreads), 0,
                       st, Ak, nkern, g.k0r, g.wr, g.k0c, g.wc, qrm, qcm, scale, batch, F);
    OI_LAUNCH_CHECK("adj_fold");
    FpArgs a{};
    a.batch = batch;
    a.ny = g.ny;
    a.nx = g.nx;
    a.qr = qrm;
    a.qc = qcm;
    a.tw_x = g.fx.tw;
    a.tw_y = g.fy.tw;
    a.Ihat = F;
    a.U = U;
    a.out = grad;
    a.theta = chain_theta;
    a.part = chain_part;
    return fp_band_image(a, st);
  }
  const int per = g.wr * g.wc;
  const int nblk = (int)((B * per + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_adj_ksum, dim3(nblk < 4096 ? nblk : 4096), dim3(kThreads), 0, st, Ak, nkern,
                     per, batch, Q);
  OI_LAUNCH_CHECK("adj_ksum");

  const int ncol_y = lds_items(g.ny, 16);
  const size_t lds_y = 2ull * ncol_y * g.ny * sizeof(float2);
  if ((rc = allow_lds_adj(k_adj_gcols, lds_y))) return rc;
  dim3 g5((g.wc + ncol_y - 1) / ncol_y, batch);
  hipLaunchKernelGGL(k_adj_gcols, g5, dim3(kThreads), lds_y, st, Q, g.k0r, g.wr, g.wc, g.ny, ncol_y,
                     g.fy, Z);
  OI_LAUNCH_CHECK("adj_gcols");
  int npair = 2048 / g.nx;
  npair = npair < 1 ? 1 : (npair > 8 ? 8 : npair);
  const size_t lds_x = 3ull * npair * g.nx * sizeof(float2);
  if ((rc = allow_lds_adj(k_adj_grows, lds_x))) return rc;
  const int pairs = (g.ny + 1) / 2;
  dim3 g6((pairs + npair - 1) / npair, batch);
  hipLaunchKernelGGL(k_adj_grows, g6, dim3(kThreads), lds_x, st, Z, g.k0c, g.wc, g.ny, g.nx, npair,
                     g.fx, scale, grad);
  OI_LAUNCH_CHECK("adj_grows");
  return OI_OK;
}

}  // namespace oi
