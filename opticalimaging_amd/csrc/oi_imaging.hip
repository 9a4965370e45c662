// Imaging kernels: windowed spectrum, coarse-grid coherent sum, Fourier
// interpolation to the image grid, full centred spectrum, kernel windows.
//
// The image of a coherent sum  I(j) = sum_k |E_k(j)|^2,
//     E_k(j) = sum_kappa A_k(kappa) exp(-2 pi i kappa.j / N),
// is a trigonometric polynomial whose spectrum lies within +-(w-1) when every
// A_k is supported on w consecutive indices.  It is therefore computed exactly
// on any coarse grid L >= 2w-1 and carried to the N-point image grid by
// zero-padded Fourier interpolation.  Each stage below is one launch:
//   k_spec_rows / k_spec_cols   mask -> windowed spectrum (pruned 2-D DFT)
//   k_coherent                  kernels x spectrum -> coarse |E|^2 rows,
//                               followed by the row DFT of the coarse image
//   k_interp_cols / _rows       coarse spectrum -> image (C2R, N points)
#include "oi_fastpath.h"

namespace oi {

static size_t fft_lds(int nb, int n) { return 2ull * nb * n * sizeof(float2); }

// Transforms per workgroup so that two ping-pong buffers take about 64 KiB.
static int per_block(int n) {
  int v = 4096 / (n > 0 ? n : 1);
  return v < 1 ? 1 : (v > 16 ? 16 : v);
}

template <typename K>
static int allow_lds(K kernel, size_t bytes) {
  if (bytes > kLdsLimit) return fail(OI_ERR_UNSUPPORTED, "shared-memory request exceeds 160 KiB");
  if (bytes > 64 * 1024) {
    OI_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  }
  return OI_OK;
}

// ------------------------------------------------------------------------
// Row pass of the windowed spectrum, real masks: two rows packed per complex
// transform.  T[img][j][r] = sum_m h[r][m] exp(+2 pi i (k0c+j) m / nx).
// ------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_spec_rows_real(const float* __restrict__ mask, int ny,
                                                             int nx, int nfft, FftDesc fx, int k0c,
                                                             int wc, float2* __restrict__ T) {
  extern __shared__ float2 lds[];
  float2* a = lds;
  float2* t = lds + nfft * nx;
  const int img = blockIdx.y;
  const int r0 = blockIdx.x * 2 * nfft;
  const float* src = mask + (size_t)img * ny * nx;
  for (int e = threadIdx.x; e < nfft * nx; e += blockDim.x) {
    const int f = e / nx, m = e - f * nx;
    const int ra = r0 + 2 * f, rb = ra + 1;
    const float va = ra < ny ? src[(size_t)ra * nx + m] : 0.f;
    const float vb = rb < ny ? src[(size_t)rb * nx + m] : 0.f;
    a[e] = make_float2(va, vb);
  }
  __syncthreads();
  const float2* z = block_fft<true>(a, t, fx, nfft, nx);
  float2* Tb = T + (size_t)img * wc * ny;
  const int rows = 2 * nfft;
  for (int e = threadIdx.x; e < wc * rows; e += blockDim.x) {
    const int j = e / rows, rr = e - j * rows;
    const int r = r0 + rr;
    if (r >= ny) continue;
    const int f = rr >> 1;
    const int k = pmod(k0c + j, nx);
    const int km = k == 0 ? 0 : nx - k;
    const float2 zk = z[f * nx + k], zm = z[f * nx + km];
    float2 v;
    if ((rr & 1) == 0)
      v = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));   // (Z[k] + conj Z[-k]) / 2
    else
      v = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));  // (Z[k] - conj Z[-k]) / 2i
    Tb[(size_t)j * ny + r] = v;
  }
}

// Row pass, complex channels-last fields [img][ny][nx][cin].
__global__ __launch_bounds__(kThreads) void k_spec_rows_c64(const float2* __restrict__ field, int cin,
                                                            Chans ch, int nchan, int ny, int nx,
                                                            int nfft, FftDesc fx, int k0c, int wc,
                                                            float2* __restrict__ T) {
  extern __shared__ float2 lds[];
  float2* a = lds;
  float2* t = lds + nfft * nx;
  const int bc = blockIdx.y;
  const int img = bc / nchan, ci = bc - img * nchan;
  const int chan = ch.c[ci];
  const int r0 = blockIdx.x * nfft;
  for (int e = threadIdx.x; e < nfft * nx; e += blockDim.x) {
    const int f = e / nx, m = e - f * nx;
    const int r = r0 + f;
    a[e] = r < ny ? field[(((size_t)img * ny + r) * nx + m) * cin + chan] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const float2* z = block_fft<true>(a, t, fx, nfft, nx);
  float2* Tb = T + (size_t)bc * wc * ny;
  for (int e = threadIdx.x; e < wc * nfft; e += blockDim.x) {
    const int j = e / nfft, f = e - j * nfft;
    const int r = r0 + f;
    if (r >= ny) continue;
    Tb[(size_t)j * ny + r] = z[f * nx + pmod(k0c + j, nx)];
  }
}

// Column pass: spec[bc][i][j] = scale * sum_r T[bc][j][r] exp(+2 pi i (k0r+i) r / ny).
__global__ __launch_bounds__(kThreads) void k_spec_cols(const float2* __restrict__ T, int ny, int ncol,
                                                        FftDesc fy, int k0r, int wr, int wc,
                                                        float scale, float2* __restrict__ spec) {
  extern __shared__ float2 lds[];
  float2* a = lds;
  float2* t = lds + ncol * ny;
  const int bc = blockIdx.y;
  const int j0 = blockIdx.x * ncol;
  const float2* Tb = T + (size_t)bc * wc * ny;
  for (int e = threadIdx.x; e < ncol * ny; e += blockDim.x) {
    const int j = e / ny, r = e - j * ny;
    a[e] = (j0 + j) < wc ? Tb[(size_t)(j0 + j) * ny + r] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const float2* z = block_fft<true>(a, t, fy, ncol, ny);
  float2* sb = spec + (size_t)bc * wr * wc;
  for (int e = threadIdx.x; e < wr * ncol; e += blockDim.x) {
    const int i = e / ncol, j = e - i * ncol;
    if (j0 + j >= wc) continue;
    sb[(size_t)i * wc + j0 + j] = cscale(z[j * ny + pmod(k0r + i, ny)], scale);
  }
}

int launch_spec_rows(const void* field, int kind, int cin, const int* chans, int nchan, int batch,
                     int ny, int nx, const FftDesc& fx, int k0c, int wc, float2* T,
                     hipStream_t st) {
  const int nfft = per_block(nx);
  const size_t lds = fft_lds(nfft, nx);
  if (kind == OI_FIELD_REAL) {
    int rc = allow_lds(k_spec_rows_real, lds);
    if (rc) return rc;
    dim3 grid((ny + 2 * nfft - 1) / (2 * nfft), batch);
    hipLaunchKernelGGL(k_spec_rows_real, grid, dim3(kThreads), lds, st,
                       static_cast<const float*>(field), ny, nx, nfft, fx, k0c, wc, T);
    OI_LAUNCH_CHECK("spec_rows_real");
  } else {
    Chans c{};
    for (int i = 0; i < nchan; ++i) c.c[i] = chans[i];
    int rc = allow_lds(k_spec_rows_c64, lds);
    if (rc) return rc;
    dim3 grid((ny + nfft - 1) / nfft, batch * nchan);
    hipLaunchKernelGGL(k_spec_rows_c64, grid, dim3(kThreads), lds, st,
                       static_cast<const float2*>(field), cin, c, nchan, ny, nx, nfft, fx, k0c,
                       wc, T);
    OI_LAUNCH_CHECK("spec_rows_c64");
  }
  return OI_OK;
}

int launch_spec_cols(const float2* T, int nbc, int ny, int nx, const FftDesc& fy, int k0r, int wr,
                     int wc, float2* spec, hipStream_t st) {
  const int ncol = per_block(ny);
  const size_t lds = fft_lds(ncol, ny);
  int rc = allow_lds(k_spec_cols, lds);
  if (rc) return rc;
  dim3 grid((wc + ncol - 1) / ncol, nbc);
  const float scale = 1.0f / ((float)ny * (float)nx);
  hipLaunchKernelGGL(k_spec_cols, grid, dim3(kThreads), lds, st, T, ny, ncol, fy, k0r, wr, wc,
                     scale, spec);
  OI_LAUNCH_CHECK("spec_cols");
  return OI_OK;
}

// ------------------------------------------------------------------------
// Coarse-grid coherent sum.  Block (s, img, g) owns coarse rows t = s + P a,
// a in [0, rows), and kernels g, g+G, ...  For each kernel and channel:
//   column stage  B[a][kc] = sum_kr A(kr,kc) exp(-2 pi i kr (P a + s) / lr)
//                 as a fold of kr mod rows (after the phase twist
//                 exp(-2 pi i kr s / lr)) and a rows-point DFT;
//   row stage     E[a][t] = sum_kc B[a][kc] exp(-2 pi i kc t / lc), kc folded
//                 mod lc, an lc-point DFT;
//   |E|^2 accumulated in registers.
// Epilogue: direct mode writes the coarse rows (= image rows); otherwise the
// rows' DFT (+ sign), orders q in [0, qc], goes to R.
// ------------------------------------------------------------------------
size_t coherent_lds_bytes(int wc, int rows, int lc) {
  return (2ull * wc * rows + 2ull * rows * lc) * sizeof(float2);
}

__global__ __launch_bounds__(kThreads) void k_coherent(CohArgs A) {
  extern __shared__ float2 lds[];
  const int rows = A.rows, lc = A.lc, wc = A.wc, wr = A.wr;
  float2* colA = lds;
  float2* colB = colA + wc * rows;
  float2* rowA = colB + wc * rows;
  float2* rowB = rowA + rows * lc;
  const int s = blockIdx.x, img = blockIdx.y, g = blockIdx.z;
  const int nrl = rows * lc;
  float acc[kAccMax];
#pragma unroll
  for (int i = 0; i < kAccMax; ++i) acc[i] = 0.f;

  for (int kk = g; kk < A.nkern; kk += A.groups) {
    const float2* kp = A.kern + (size_t)kk * wr * wc;
    for (int c = 0; c < A.nchan; ++c) {
      const float2* sp = A.spec + ((size_t)img * A.nchan + c) * wr * wc;
      for (int e = threadIdx.x; e < wc * rows; e += blockDim.x) {
        const int u = e / wc, cc = e - u * wc;
        float2 sum = make_float2(0.f, 0.f);
        for (int ir = pmod(u - A.k0r, rows); ir < wr; ir += rows) {
          const int kr = A.k0r + ir;
          const float2 v = cmul(kp[ir * wc + cc], sp[ir * wc + cc]);
          sum = cadd(sum, cmul(v, A.tw_lr[pmod(kr * s, A.lr)]));
        }
        colA[cc * rows + u] = sum;
      }
      __syncthreads();
      const float2* col = block_fft<false>(colA, colB, A.frows, wc, rows);
      for (int e = threadIdx.x; e < nrl; e += blockDim.x) {
        const int a = e / lc, slot = e - a * lc;
        float2 sum = make_float2(0.f, 0.f);
        for (int ic = pmod(slot - A.k0c, lc); ic < wc; ic += lc) sum = cadd(sum, col[ic * rows + a]);
        rowA[e] = sum;
      }
      __syncthreads();
      const float2* row = block_fft<false>(rowA, rowB, A.flc, rows, lc);
#pragma unroll
      for (int i = 0; i < kAccMax; ++i) {
        const int e = threadIdx.x + i * kThreads;
        if (e < nrl) acc[i] += cnorm(row[e]);
      }
    }
  }

  if (A.direct) {
    float* dst = A.Ic + ((size_t)img * A.groups + g) * A.lr * lc;
#pragma unroll
    for (int i = 0; i < kAccMax; ++i) {
      const int e = threadIdx.x + i * kThreads;
      if (e < nrl) {
        const int a = e / lc, t = e - a * lc;
        dst[(size_t)(s + A.phases * a) * lc + t] = acc[i];
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < kAccMax; ++i) {
      const int e = threadIdx.x + i * kThreads;
      if (e < nrl) rowA[e] = make_float2(acc[i], 0.f);
    }
    __syncthreads();
    const float2* row = block_fft<true>(rowA, rowB, A.flc, rows, lc);
    const int nq = A.qc + 1;
    float2* dst = A.R + ((size_t)img * A.groups + g) * A.lr * nq;
    for (int e = threadIdx.x; e < rows * nq; e += blockDim.x) {
      const int a = e / nq, q = e - a * nq;
      dst[(size_t)(s + A.phases * a) * nq + q] = row[a * lc + q];
    }
  }
}

int launch_coherent(const CohArgs& a, int batch, hipStream_t st) {
  const size_t lds = coherent_lds_bytes(a.wc, a.rows, a.lc);
  int rc = allow_lds(k_coherent, lds);
  if (rc) return rc;
  if ((size_t)a.rows * a.lc > (size_t)kThreads * kAccMax)
    return fail(OI_ERR_INVALID, "coherent: rows*lc exceeds the per-block accumulator");
  dim3 grid(a.phases, batch, a.groups);
  hipLaunchKernelGGL(k_coherent, grid, dim3(kThreads), lds, st, a);
  OI_LAUNCH_CHECK("coherent");
  return OI_OK;
}

// ------------------------------------------------------------------------
// Fourier interpolation.  Columns: Ihat(q_r, q_c) = scale * sum_t R[t][q_c]
// exp(+2 pi i q_r t / lr) for |q_r| <= qr, zero-padded to ny and transformed
// back: U[y][q_c] = sum_qr Ihat exp(-2 pi i q_r y / ny).
// ------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_interp_cols(InterpArgs A, int ncol) {
  extern __shared__ float2 lds[];
  const int L = A.lr > A.ny ? A.lr : A.ny;
  float2* bufA = lds;
  float2* bufB = lds + ncol * L;
  const int img = blockIdx.y, q0 = blockIdx.x * ncol;
  const int nq = A.qc + 1;
  for (int e = threadIdx.x; e < ncol * A.lr; e += blockDim.x) {
    const int j = e / A.lr, t = e - j * A.lr;
    const int q = q0 + j;
    float2 sum = make_float2(0.f, 0.f);
    if (q < nq)
      for (int g = 0; g < A.groups; ++g)
        sum = cadd(sum, A.R[(((size_t)img * A.groups + g) * A.lr + t) * nq + q]);
    bufA[j * L + t] = sum;
  }
  __syncthreads();
  float2* z = block_fft<true>(bufA, bufB, A.flr, ncol, L);
  float2* o = (z == bufA) ? bufB : bufA;
  for (int e = threadIdx.x; e < ncol * A.ny; e += blockDim.x) {
    const int j = e / A.ny, y = e - j * A.ny;
    float2 v = make_float2(0.f, 0.f);
    if (y <= A.qr)
      v = cscale(z[j * L + y], A.scale);
    else if (y >= A.ny - A.qr)
      v = cscale(z[j * L + (y - A.ny + A.lr)], A.scale);
    o[j * L + y] = v;
  }
  __syncthreads();
  const float2* w = block_fft<false>(o, z, A.fny, ncol, L);
  for (int e = threadIdx.x; e < ncol * A.ny; e += blockDim.x) {
    const int y = e / ncol, j = e - y * ncol;
    const int q = q0 + j;
    if (q < nq) A.U[((size_t)img * A.ny + y) * nq + q] = w[j * L + y];
  }
}

// Rows: two image rows per complex transform (Hermitian extension of U).
__global__ __launch_bounds__(kThreads) void k_interp_rows(InterpArgs A, int nfft) {
  extern __shared__ float2 lds[];
  float2* a = lds;
  float2* t = lds + nfft * A.nx;
  const int img = blockIdx.y;
  const int ybase = blockIdx.x * 2 * nfft;
  const int nq = A.qc + 1;
  const int nx = A.nx, ny = A.ny;
  for (int e = threadIdx.x; e < nfft * nx; e += blockDim.x) {
    const int f = e / nx, x = e - f * nx;
    const int y0 = ybase + 2 * f, y1 = y0 + 1;
    const float2* U0 = A.U + ((size_t)img * ny + y0) * nq;
    const float2* U1 = U0 + nq;
    float2 z = make_float2(0.f, 0.f);
    if (x <= A.qc || x >= nx - A.qc) {
      const bool neg = x > A.qc;
      const int q = neg ? nx - x : x;
      float2 u0 = y0 < ny ? U0[q] : make_float2(0.f, 0.f);
      float2 u1 = y1 < ny ? U1[q] : make_float2(0.f, 0.f);
      if (neg) {
        u0 = cconj(u0);
        u1 = cconj(u1);
      }
      z = make_float2(u0.x - u1.y, u0.y + u1.x);  // u0 + i u1
    }
    a[e] = z;
  }
  __syncthreads();
  const float2* z = block_fft<false>(a, t, A.fnx, nfft, nx);
  for (int e = threadIdx.x; e < nfft * nx; e += blockDim.x) {
    const int f = e / nx, x = e - f * nx;
    const int y0 = ybase + 2 * f, y1 = y0 + 1;
    if (y0 < ny) A.out[((size_t)img * ny + y0) * nx + x] = z[e].x;
    if (y1 < ny) A.out[((size_t)img * ny + y1) * nx + x] = z[e].y;
  }
}

int launch_interp(const InterpArgs& a, int batch, hipStream_t st, StageMark mark, void* ctx) {
  const int L = a.lr > a.ny ? a.lr : a.ny;
  const int ncol = per_block(L);
  size_t lds = fft_lds(ncol, L);
  int rc = allow_lds(k_interp_cols, lds);
  if (rc) return rc;
  dim3 g1((a.qc + 1 + ncol - 1) / ncol, batch);
  hipLaunchKernelGGL(k_interp_cols, g1, dim3(kThreads), lds, st, a, ncol);
  OI_LAUNCH_CHECK("interp_cols");
  if (mark) mark(ctx, 4);
  const int nfft = per_block(a.nx);
  lds = fft_lds(nfft, a.nx);
  rc = allow_lds(k_interp_rows, lds);
  if (rc) return rc;
  dim3 g2((a.ny + 2 * nfft - 1) / (2 * nfft), batch);
  hipLaunchKernelGGL(k_interp_rows, g2, dim3(kThreads), lds, st, a, nfft);
  OI_LAUNCH_CHECK("interp_rows");
  if (mark) mark(ctx, 5);
  return OI_OK;
}

__global__ void k_group_sum(const float* __restrict__ Ic, int groups, size_t per_image,
                            float* __restrict__ out, size_t total) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t img = e / per_image, r = e - img * per_image;
    float s = 0.f;
    for (int g = 0; g < groups; ++g) s += Ic[(img * groups + g) * per_image + r];
    out[e] = s;
  }
}

int launch_group_sum(const float* Ic, int batch, int groups, size_t per_image, float* out,
                     hipStream_t st) {
  const size_t total = (size_t)batch * per_image;
  const int blocks = (int)((total + kThreads - 1) / kThreads > 65535 ? 65535
                                                                    : (total + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_group_sum, dim3(blocks), dim3(kThreads), 0, st, Ic, groups, per_image, out,
                     total);
  OI_LAUNCH_CHECK("group_sum");
  return OI_OK;
}

// ------------------------------------------------------------------------
// Full centred spectrum (Scatter.py:20-23): work[bc][r][jc] holds the row
// transform of the ifftshifted field with the column fftshift folded into the
// store; the column kernel finishes the 2-D transform and folds the row shift.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_full_rows(const void* __restrict__ field, int kind, int cin,
                                                        Chans ch, int nchan, int ny, int nx, int nfft,
                                                        FftDesc fx, float2* __restrict__ work) {
  extern __shared__ float2 lds[];
  float2* a = lds;
  float2* t = lds + nfft * nx;
  const int bc = blockIdx.y;
  const int img = bc / nchan, ci = bc - img * nchan;
  const int r0 = blockIdx.x * nfft;
  const int hy = ny / 2, hx = nx / 2;
  for (int e = threadIdx.x; e < nfft * nx; e += blockDim.x) {
    const int f = e / nx, n = e - f * nx;
    const int r = r0 + f;
    float2 v = make_float2(0.f, 0.f);
    if (r < ny) {
      const int rs = (r + hy) % ny, cs = (n + hx) % nx;
      if (kind == OI_FIELD_REAL)
        v = make_float2(static_cast<const float*>(field)[((size_t)img * ny + rs) * nx + cs], 0.f);
      else
        v = static_cast<const float2*>(field)[(((size_t)img * ny + rs) * nx + cs) * cin + ch.c[ci]];
    }
    a[e] = v;
  }
  __syncthreads();
  const float2* z = block_fft<true>(a, t, fx, nfft, nx);
  for (int e = threadIdx.x; e < nfft * nx; e += blockDim.x) {
    const int f = e / nx, jc = e - f * nx;
    const int r = r0 + f;
    if (r < ny) work[((size_t)bc * ny + r) * nx + jc] = z[f * nx + pmod(jc - hx, nx)];
  }
}

__global__ __launch_bounds__(kThreads) void k_full_cols(const float2* __restrict__ work, int nchan,
                                                        int ny, int nx, int ncol, FftDesc fy,
                                                        float scale, float2* __restrict__ hk) {
  extern __shared__ float2 lds[];
  float2* a = lds;
  float2* t = lds + ncol * ny;
  const int bc = blockIdx.y;
  const int img = bc / nchan, ci = bc - img * nchan;
  const int c0 = blockIdx.x * ncol;
  const int hy = ny / 2;
  for (int e = threadIdx.x; e < ncol * ny; e += blockDim.x) {
    const int r = e / ncol, j = e - r * ncol;
    const int c = c0 + j;
    a[j * ny + r] = c < nx ? work[((size_t)bc * ny + r) * nx + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const float2* z = block_fft<true>(a, t, fy, ncol, ny);
  for (int e = threadIdx.x; e < ncol * ny; e += blockDim.x) {
    const int jr = e / ncol, j = e - jr * ncol;
    const int c = c0 + j;
    if (c < nx)
      hk[(((size_t)img * ny + jr) * nx + c) * nchan + ci] = cscale(z[j * ny + pmod(jr - hy, ny)], scale);
  }
}

int launch_full_rows(const void* field, int kind, int cin, const int* chans, int nchan, int batch,
                     int ny, int nx, const FftDesc& fx, float2* work, hipStream_t st) {
  Chans c{};
  for (int i = 0; i < nchan; ++i) c.c[i] = kind == OI_FIELD_REAL ? 0 : chans[i];
  const int nfft = per_block(nx);
  const size_t lds = fft_lds(nfft, nx);
  int rc = allow_lds(k_full_rows, lds);
  if (rc) return rc;
  dim3 grid((ny + nfft - 1) / nfft, batch * nchan);
  hipLaunchKernelGGL(k_full_rows, grid, dim3(kThreads), lds, st, field, kind, cin, c, nchan, ny, nx,
                     nfft, fx, work);
  OI_LAUNCH_CHECK("full_rows");
  return OI_OK;
}

int launch_full_cols(const float2* work, int nchan, int batch, int ny, int nx, const FftDesc& fy,
                     float scale, float2* hk, hipStream_t st) {
  int ncol = per_block(ny);
  const size_t lds = fft_lds(ncol, ny);
  int rc = allow_lds(k_full_cols, lds);
  if (rc) return rc;
  dim3 grid((nx + ncol - 1) / ncol, batch * nchan);
  hipLaunchKernelGGL(k_full_cols, grid, dim3(kThreads), lds, st, work, nchan, ny, nx, ncol, fy,
                     scale, hk);
  OI_LAUNCH_CHECK("full_cols");
  return OI_OK;
}

// ------------------------------------------------------------------------
// Dense SOCS kernels [ny][nx][nk] <-> windows [nk][wr][wc].
// ------------------------------------------------------------------------
__global__ void k_kernel_support(const float2* __restrict__ dense, int ny, int nx, int nk,
                                 int* __restrict__ rowflag, int* __restrict__ colflag) {
  const size_t total = (size_t)ny * nx;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * blockDim.x) {
    const float2* p = dense + idx * nk;
    bool nz = false;
    for (int k = 0; k < nk && !nz; ++k) nz = (p[k].x != 0.f) || (p[k].y != 0.f);
    if (nz) {
      const int r = (int)(idx / nx), c = (int)(idx - (size_t)r * nx);
      if (rowflag[r] == 0) atomicOr(&rowflag[r], 1);
      if (colflag[c] == 0) atomicOr(&colflag[c], 1);
    }
  }
}

__global__ void k_kernel_window(const float2* __restrict__ dense, int ny, int nx, int nk, int k0r,
                                int wr, int k0c, int wc, float2* __restrict__ out) {
  const size_t total = (size_t)nk * wr * wc;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(e / ((size_t)wr * wc));
    const int rem = (int)(e - (size_t)k * wr * wc);
    const int i = rem / wc, j = rem - i * wc;
    const int r = pmod(k0r + i, ny), c = pmod(k0c + j, nx);
    out[e] = dense[((size_t)r * nx + c) * nk + k];
  }
}

__global__ void k_kernel_dense(const float2* __restrict__ win, int nk, int k0r, int wr, int k0c,
                               int wc, int ny, int nx, float2* __restrict__ dense) {
  const size_t total = (size_t)ny * nx * nk;
  const int hy = ny / 2, hx = nx / 2;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t pix = e / nk;
    const int k = (int)(e - pix * nk);
    const int r = (int)(pix / nx), c = (int)(pix - (size_t)r * nx);
    const int kr = pmod(r + hy, ny) - hy, kc = pmod(c + hx, nx) - hx;  // centred index of (r, c)
    const int i = kr - k0r, j = kc - k0c;
    float2 v = make_float2(0.f, 0.f);
    if (i >= 0 && i < wr && j >= 0 && j < wc) v = win[((size_t)k * wr + i) * wc + j];
    dense[e] = v;
  }
}

static int grid_for(size_t total) {
  size_t b = (total + kThreads - 1) / kThreads;
  if (b > 16384) b = 16384;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace oi

using namespace oi;

extern "C" int oi_kernel_support(const void* dense, int ny, int nx, int nk, int* rowflag,
                                 int* colflag, void* stream) {
  if (!dense || !rowflag || !colflag || ny <= 0 || nx <= 0 || nk <= 0)
    return fail(OI_ERR_INVALID, "kernel_support: bad arguments");
  hipLaunchKernelGGL(k_kernel_support, dim3(grid_for((size_t)ny * nx)), dim3(kThreads), 0,
                     (hipStream_t)stream, static_cast<const float2*>(dense), ny, nx, nk, rowflag,
                     colflag);
  OI_LAUNCH_CHECK("kernel_support");
  return OI_OK;
}

extern "C" int oi_kernel_window(const void* dense, int ny, int nx, int nk, int k0r, int wr, int k0c,
                                int wc, void* out, void* stream) {
  if (!dense || !out || ny <= 0 || nx <= 0 || nk <= 0 || wr <= 0 || wc <= 0 || wr > ny || wc > nx)
    return fail(OI_ERR_INVALID, "kernel_window: bad arguments");
  hipLaunchKernelGGL(k_kernel_window, dim3(grid_for((size_t)nk * wr * wc)), dim3(kThreads), 0,
                     (hipStream_t)stream, static_cast<const float2*>(dense), ny, nx, nk, k0r, wr,
                     k0c, wc, static_cast<float2*>(out));
  OI_LAUNCH_CHECK("kernel_window");
  return OI_OK;
}

extern "C" int oi_kernel_dense(const void* win, int nk, int k0r, int wr, int k0c, int wc, int ny,
                               int nx, void* dense, void* stream) {
  if (!win || !dense || ny <= 0 || nx <= 0 || nk <= 0 || wr <= 0 || wc <= 0 || wr > ny || wc > nx ||
      k0r < -(ny / 2) || k0r + wr - 1 > ny - 1 - ny / 2 || k0c < -(nx / 2) ||
      k0c + wc - 1 > nx - 1 - nx / 2)
    return fail(OI_ERR_INVALID, "kernel_dense: window outside the centred index range");
  hipLaunchKernelGGL(k_kernel_dense, dim3(grid_for((size_t)ny * nx * nk)), dim3(kThreads), 0,
                     (hipStream_t)stream, static_cast<const float2*>(win), nk, k0r, wr, k0c, wc, ny,
                     nx, static_cast<float2*>(dense));
  OI_LAUNCH_CHECK("kernel_dense");
  return OI_OK;
}

// ------------------------------------------------------------ coded masks --
// Masks sent as one byte per pixel: out[i] = levels[codes[i]] (codes past the
// table give 0).  Sixteen codes per thread and iteration (one 16-byte load,
// four 16-byte stores); the table sits in LDS.
namespace {
struct Levels {
  float v[256];
};

__global__ __launch_bounds__(256) void k_decode_u8(const uint8_t* __restrict__ codes, size_t n,
                                                   Levels lv, int nl, float* __restrict__ out) {
  __shared__ float lut[256];
  lut[threadIdx.x] = (int)threadIdx.x < nl ? lv.v[threadIdx.x] : 0.f;
  __syncthreads();
  const size_t n16 = n / 16;
  const uint4* c16 = reinterpret_cast<const uint4*>(codes);
  float4* o4 = reinterpret_cast<float4*>(out);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) {
    const uint4 c = c16[i];
    const uint32_t w[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
    for (int q = 0; q < 4; ++q)
      o4[4 * i + q] = make_float4(lut[w[q] & 0xff], lut[(w[q] >> 8) & 0xff], lut[(w[q] >> 16) & 0xff],
                                  lut[w[q] >> 24]);
  }
  if (blockIdx.x == 0) {
    for (size_t i = n16 * 16 + threadIdx.x; i < n; i += blockDim.x) out[i] = lut[codes[i]];
  }
}

// One bit per pixel: each thread expands 4 bytes (32 pixels, bit 7 of a byte
// first) into eight 16-byte stores.
__global__ __launch_bounds__(256) void k_decode_bits(const uint8_t* __restrict__ bits, size_t n,
                                                     float v0, float v1, float* __restrict__ out) {
  const size_t n32 = n / 32;
  const uint32_t* b4 = reinterpret_cast<const uint32_t*>(bits);
  float4* o4 = reinterpret_cast<float4*>(out);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n32; i += stride) {
    const uint32_t w = b4[i];
#pragma unroll
    for (int byte = 0; byte < 4; ++byte) {
      const uint32_t c = (w >> (8 * byte)) & 0xffu;     // little-endian words: byte order kept
      o4[8 * i + 2 * byte] = make_float4((c & 0x80u) ? v1 : v0, (c & 0x40u) ? v1 : v0,
                                         (c & 0x20u) ? v1 : v0, (c & 0x10u) ? v1 : v0);
      o4[8 * i + 2 * byte + 1] = make_float4((c & 0x08u) ? v1 : v0, (c & 0x04u) ? v1 : v0,
                                             (c & 0x02u) ? v1 : v0, (c & 0x01u) ? v1 : v0);
    }
  }
  if (blockIdx.x == 0) {
    for (size_t i = n32 * 32 + threadIdx.x; i < n; i += blockDim.x)
      out[i] = ((bits[i >> 3] >> (7 - (i & 7))) & 1u) ? v1 : v0;
  }
}
}  // namespace

extern "C" int oi_mask_decode_bits(const void* bits, size_t count, const float* levels, float* out,
                                   void* stream) {
  if (!bits || !out || !levels) return fail(OI_ERR_INVALID, "mask_decode_bits: null pointer");
  if ((reinterpret_cast<uintptr_t>(bits) & 3) || (reinterpret_cast<uintptr_t>(out) & 15))
    return fail(OI_ERR_INVALID, "mask_decode_bits: bits must be 4-byte and out 16-byte aligned");
  if (count == 0) return OI_OK;
  const size_t n32 = count / 32;
  size_t blocks = (n32 + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_decode_bits, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     static_cast<const uint8_t*>(bits), count, levels[0], levels[1], out);
  OI_LAUNCH_CHECK("mask_decode_bits");
  return OI_OK;
}

extern "C" int oi_mask_decode_u8(const void* codes, size_t count, const float* levels, int nlevels,
                                 float* out, void* stream) {
  if (!codes || !out || !levels || nlevels < 1 || nlevels > 256)
    return fail(OI_ERR_INVALID, "mask_decode_u8: null pointer or level count outside 1..256");
  if ((reinterpret_cast<uintptr_t>(codes) & 15) || (reinterpret_cast<uintptr_t>(out) & 15))
    return fail(OI_ERR_INVALID, "mask_decode_u8: codes and out must be 16-byte aligned");
  if (count == 0) return OI_OK;
  Levels lv;
  for (int i = 0; i < 256; ++i) lv.v[i] = i < nlevels ? levels[i] : 0.f;
  const size_t n16 = count / 16;
  size_t blocks = (n16 + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_decode_u8, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     static_cast<const uint8_t*>(codes), count, lv, nlevels, out);
  OI_LAUNCH_CHECK("mask_decode_u8");
  return OI_OK;
}

