// RCWA near-field producer: the elementwise and gather stages of the layer
// and field computations (float64 complex).  The dense algebra between them
// (products, solves, inverses of 2N x 2N matrices) runs in rocBLAS/rocSOLVER
// through the host layer.
//
// Index convention: Fourier order (ox, oy), ox in [-Ox, Ox], oy in [-Oy, Oy],
// flattened as i = (ox + Ox) * (2 Oy + 1) + (oy + Oy) (rcwa.py:340-347, the
// 'ij' meshgrid of rcwa.py:356-361).  Matrices are row-major c128.
#include <cmath>

#include "oi_internal.h"

namespace oi {

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 zdiv(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// sqrt on the branch with Im >= 0 (rcwa.py:866-867: conj when Im < 0), via
// the principal root sqrt(r) e^{i phi/2}.
__device__ __forceinline__ double2 zsqrt_upper(double2 z) {
  const double r = hypot(z.x, z.y);
  double re = sqrt(0.5 * (r + z.x));
  double im = sqrt(0.5 * (r - z.x));
  if (z.y < 0) im = -im;  // principal root
  if (im < 0) im = -im;   // conj when the imaginary part is negative
  return make_double2(re, im);
}

// E[i][j] = F[(ox_i - ox_j) mod n0][(oy_i - oy_j) mod n1] (rcwa.py:435-450,
// F = fft2(eps) / (n0 n1) already scaled).
__global__ void k_rcwa_convmat(const double2* __restrict__ F, int n0, int n1, int Ox, int Oy,
                               double2* __restrict__ E) {
  const int ny = 2 * Oy + 1;
  const int N = (2 * Ox + 1) * ny;
  const size_t total = (size_t)N * N;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / N), j = (int)(e - (size_t)i * N);
    const int dx = i / ny - j / ny, dy = i % ny - j % ny;
    E[e] = F[(size_t)pmod(dx, n0) * n1 + pmod(dy, n1)];
  }
}

// P, Q of a layer (rcwa.py:470-497) from its convolution matrix E and
// EiK = [E^-1 Kx | E^-1 Ky] (N x 2N), with Kx, Ky diagonal:
//   P = [[Kx EiKy, I - Kx EiKx], [Ky EiKy - I, -Ky EiKx]]
//   Q = [[-Kx Ky, Kx Kx - E], [E - Ky Ky, Kx Ky]]
__global__ void k_rcwa_pq(const double2* __restrict__ E, const double2* __restrict__ EiK,
                          const double* __restrict__ kx, const double* __restrict__ ky, int N,
                          double2* __restrict__ P, double2* __restrict__ Q) {
  const size_t n2 = 2 * (size_t)N;
  const size_t total = n2 * n2;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / n2), c = (int)(e - (size_t)r * n2);
    const int i = r % N, j = c % N;
    const bool top = r < N, left = c < N;
    const double2 eix = EiK[(size_t)i * n2 + j];      // (E^-1 Kx)[i][j]
    const double2 eiy = EiK[(size_t)i * n2 + N + j];  // (E^-1 Ky)[i][j]
    const double ki = top ? kx[i] : ky[i];
    const double d = (i == j) ? 1.0 : 0.0;
    double2 p;
    if (top && left)
      p = make_double2(ki * eiy.x, ki * eiy.y);
    else if (top)
      p = make_double2(d - ki * eix.x, -ki * eix.y);
    else if (left)
      p = make_double2(ki * eiy.x - d, ki * eiy.y);
    else
      p = make_double2(-ki * eix.x, -ki * eix.y);
    P[e] = p;
    const double2 ee = E[(size_t)i * N + j];
    const double kxy = (i == j) ? kx[i] * ky[i] : 0.0;
    const double kxx = (i == j) ? kx[i] * kx[i] : 0.0;
    const double kyy = (i == j) ? ky[i] * ky[i] : 0.0;
    double2 q;
    if (top && left)
      q = make_double2(-kxy, 0.0);
    else if (top)
      q = make_double2(kxx - ee.x, -ee.y);
    else if (left)
      q = make_double2(ee.x - kyy, ee.y);
    else
      q = make_double2(kxy, 0.0);
    Q[e] = q;
  }
}

// Mode matrix of a homogeneous medium (get_V, rcwa.py:865-872):
//   kz = sqrt(eps - kx^2 - ky^2) (Im >= 0),
//   V = [[-Kx Ky / kz, -(kz + Ky^2 / kz)], [kz + Kx^2 / kz, Kx Ky / kz]].
__global__ void k_rcwa_modes(double2 eps, const double* __restrict__ kx,
                             const double* __restrict__ ky, int N, double2* __restrict__ V,
                             double2* __restrict__ kz) {
  const size_t n2 = 2 * (size_t)N;
  const size_t total = n2 * n2;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / n2), c = (int)(e - (size_t)r * n2);
    const int i = r % N, j = c % N;
    double2 v = make_double2(0.0, 0.0);
    if (i == j) {
      const double2 k = zsqrt_upper(make_double2(eps.x - kx[i] * kx[i] - ky[i] * ky[i], eps.y));
      const double2 one = make_double2(1.0, 0.0);
      const double2 inv = zdiv(one, k);
      const bool top = r < N, left = c < N;
      if (top && left)
        v = make_double2(-kx[i] * ky[i] * inv.x, -kx[i] * ky[i] * inv.y);
      else if (top)
        v = make_double2(-(k.x + ky[i] * ky[i] * inv.x), -(k.y + ky[i] * ky[i] * inv.y));
      else if (left)
        v = make_double2(k.x + kx[i] * kx[i] * inv.x, k.y + kx[i] * kx[i] * inv.y);
      else
        v = make_double2(kx[i] * ky[i] * inv.x, kx[i] * ky[i] * inv.y);
      if (r == c && top && kz) kz[i] = k;
    }
    V[e] = v;
  }
}

// Field synthesis on an x-y grid (rcwa.py:306-315):
//   out[x][y][c] = sum_o amp[c][o] exp(i omega (kx_o x + ky_o y)),  c < 3,
// with kx_o = kxo[o / noy], ky_o = kyo[o % noy].  The phase factors are
// separable: the block stages exp(i omega kx x) for its x and exp(i omega
// ky y) for its y in LDS once; each thread then sums over the orders.
constexpr int kFieldTile = 16;
__global__ __launch_bounds__(256) void k_rcwa_fields(const double2* __restrict__ amp, int nox, int noy,
                                                     const double* __restrict__ kxo,
                                                     const double* __restrict__ kyo, double omega,
                                                     const double* __restrict__ xs, int nx,
                                                     const double* __restrict__ ys, int ny,
                                                     double2* __restrict__ out) {
  extern __shared__ double2 sm[];
  double2* px = sm;                              // [nox][kFieldTile]
  double2* py = sm + (size_t)nox * kFieldTile;   // [noy][kFieldTile]
  const int x0 = blockIdx.x * kFieldTile, y0 = blockIdx.y * kFieldTile;
  for (int e = threadIdx.x; e < nox * kFieldTile; e += blockDim.x) {
    const int o = e / kFieldTile, t = e - o * kFieldTile;
    const double a = x0 + t < nx ? omega * kxo[o] * xs[x0 + t] : 0.0;
    px[e] = make_double2(cos(a), sin(a));
  }
  for (int e = threadIdx.x; e < noy * kFieldTile; e += blockDim.x) {
    const int o = e / kFieldTile, t = e - o * kFieldTile;
    const double a = y0 + t < ny ? omega * kyo[o] * ys[y0 + t] : 0.0;
    py[e] = make_double2(cos(a), sin(a));
  }
  __syncthreads();
  const int tx = threadIdx.x / kFieldTile, ty = threadIdx.x - tx * kFieldTile;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= nx || y >= ny) return;
  const int N = nox * noy;
  double2 acc[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
  for (int a = 0; a < nox; ++a) {
    const double2 fx = px[a * kFieldTile + tx];
    double2 row[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    for (int b = 0; b < noy; ++b) {
      const double2 fy = py[b * kFieldTile + ty];
      const int o = a * noy + b;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double2 v = amp[(size_t)c * N + o];
        row[c].x += v.x * fy.x - v.y * fy.y;
        row[c].y += v.x * fy.y + v.y * fy.x;
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      acc[c].x += row[c].x * fx.x - row[c].y * fx.y;
      acc[c].y += row[c].x * fx.y + row[c].y * fx.x;
    }
  }
  double2* dst = out + ((size_t)x * ny + y) * 3;
#pragma unroll
  for (int c = 0; c < 3; ++c) dst[c] = acc[c];
}

static int grid_for(size_t total) {
  size_t b = (total + 255) / 256;
  if (b > 8192) b = 8192;
  return b < 1 ? 1 : (int)b;
}

}  // namespace oi

using namespace oi;

extern "C" int oi_rcwa_convmat(const void* F, int n0, int n1, int ox, int oy, void* E,
                               void* stream) {
  if (!F || !E || n0 < 1 || n1 < 1 || ox < 0 || oy < 0)
    return fail(OI_ERR_INVALID, "rcwa_convmat: bad arguments");
  const size_t N = (size_t)(2 * ox + 1) * (2 * oy + 1);
  hipLaunchKernelGGL(k_rcwa_convmat, dim3(grid_for(N * N)), dim3(256), 0, (hipStream_t)stream,
                     static_cast<const double2*>(F), n0, n1, ox, oy, static_cast<double2*>(E));
  OI_LAUNCH_CHECK("rcwa_convmat");
  return OI_OK;
}

extern "C" int oi_rcwa_pq(const void* E, const void* EiK, const double* kx, const double* ky, int N,
                          void* P, void* Q, void* stream) {
  if (!E || !EiK || !kx || !ky || !P || !Q || N < 1)
    return fail(OI_ERR_INVALID, "rcwa_pq: bad arguments");
  const size_t n2 = 2 * (size_t)N;
  hipLaunchKernelGGL(k_rcwa_pq, dim3(grid_for(n2 * n2)), dim3(256), 0, (hipStream_t)stream,
                     static_cast<const double2*>(E), static_cast<const double2*>(EiK), kx, ky, N,
                     static_cast<double2*>(P), static_cast<double2*>(Q));
  OI_LAUNCH_CHECK("rcwa_pq");
  return OI_OK;
}

extern "C" int oi_rcwa_modes(double eps_re, double eps_im, const double* kx, const double* ky, int N,
                             void* V, void* kz, void* stream) {
  if (!kx || !ky || !V || N < 1) return fail(OI_ERR_INVALID, "rcwa_modes: bad arguments");
  const size_t n2 = 2 * (size_t)N;
  hipLaunchKernelGGL(k_rcwa_modes, dim3(grid_for(n2 * n2)), dim3(256), 0, (hipStream_t)stream,
                     make_double2(eps_re, eps_im), kx, ky, N, static_cast<double2*>(V),
                     static_cast<double2*>(kz));
  OI_LAUNCH_CHECK("rcwa_modes");
  return OI_OK;
}

extern "C" int oi_rcwa_fields(const void* amp, int nox, int noy, const double* kxo,
                              const double* kyo, double omega, const double* xs, int nx,
                              const double* ys, int ny, void* out, void* stream) {
  if (!amp || !kxo || !kyo || !xs || !ys || !out || nox < 1 || noy < 1 || nx < 1 || ny < 1)
    return fail(OI_ERR_INVALID, "rcwa_fields: bad arguments");
  const size_t lds = (size_t)(nox + noy) * kFieldTile * sizeof(double2);
  if (lds > 64 * 1024) return fail(OI_ERR_UNSUPPORTED, "rcwa_fields: more than 255 orders per axis");
  dim3 grid((nx + kFieldTile - 1) / kFieldTile, (ny + kFieldTile - 1) / kFieldTile);
  hipLaunchKernelGGL(k_rcwa_fields, grid, dim3(kFieldTile * kFieldTile), lds, (hipStream_t)stream,
                     static_cast<const double2*>(amp), nox, noy, kxo, kyo, omega, xs, nx, ys, ny,
                     static_cast<double2*>(out));
  OI_LAUNCH_CHECK("rcwa_fields");
  return OI_OK;
}
