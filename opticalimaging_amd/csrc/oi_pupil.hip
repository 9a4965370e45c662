// Pupil functions: Fringe Zernike wavefront, characteristic (polarisation)
// matrix and shifted-pupil kernels on a spectrum window, in float64.
//
// The 37 Fringe coefficients are folded on the host into harmonic polynomials
//   W(rho, theta) = sum_{m=0..5} A_m(rho) cos(m theta) + B_m(rho) sin(m theta),
// A_m, B_m of degree <= 12 in rho, from the factorial sum of the Zernike radial
// polynomials R_n^m; cos/sin(m theta) come from (f, g)/rho by the Chebyshev
// recursion, so no atan2 is needed on the device.
#include <climits>
#include <cmath>
#include <cstring>
#include <string>

#include "oi_internal.h"

namespace oi {

struct ZernPoly {
  double a[6][13];  // cos (and m = 0) terms
  double b[6][13];  // sin terms
};

// Fringe order Z1..Z37: (n, m, kind) with kind 0 symmetric, 1 cos, -1 sin.
static const int kFringe[37][3] = {
    {0, 0, 0},  {1, 1, 1},  {1, 1, -1}, {2, 0, 0},  {2, 2, 1},  {2, 2, -1}, {3, 1, 1},
    {3, 1, -1}, {4, 0, 0},  {3, 3, 1},  {3, 3, -1}, {4, 2, 1},  {4, 2, -1}, {5, 1, 1},
    {5, 1, -1}, {6, 0, 0},  {4, 4, 1},  {4, 4, -1}, {5, 3, 1},  {5, 3, -1}, {6, 2, 1},
    {6, 2, -1}, {7, 1, 1},  {7, 1, -1}, {8, 0, 0},  {5, 5, 1},  {5, 5, -1}, {6, 4, 1},
    {6, 4, -1}, {7, 3, 1},  {7, 3, -1}, {8, 2, 1},  {8, 2, -1}, {9, 1, 1},  {9, 1, -1},
    {10, 0, 0}, {12, 0, 0}};

static double fact(int n) {
  double r = 1.0;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}

static ZernPoly fold_zernike(const double* coeffs) {
  ZernPoly z;
  std::memset(&z, 0, sizeof(z));
  for (int t = 0; t < 37; ++t) {
    const double c = coeffs[t];
    if (c == 0.0) continue;
    const int n = kFringe[t][0], m = kFringe[t][1], kind = kFringe[t][2];
    for (int s = 0; s <= (n - m) / 2; ++s) {
      const double r = ((s % 2) ? -1.0 : 1.0) * fact(n - s) /
                       (fact(s) * fact((n + m) / 2 - s) * fact((n - m) / 2 - s));
      const int p = n - 2 * s;
      if (kind == -1)
        z.b[m][p] += c * r;
      else
        z.a[m][p] += c * r;
    }
  }
  return z;
}

// Wavefront in waves at (rho, cos theta, sin theta).
__device__ double zern_eval(const ZernPoly& z, double rho, double ct, double st) {
  double rp[13];
  rp[0] = 1.0;
#pragma unroll
  for (int k = 1; k < 13; ++k) rp[k] = rp[k - 1] * rho;
  double w = 0.0;
  double cm = 1.0, sm = 0.0;  // cos(m theta), sin(m theta)
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    double pa = 0.0, pb = 0.0;
#pragma unroll
    for (int k = 0; k < 13; ++k) {
      pa += z.a[m][k] * rp[k];
      pb += z.b[m][k] * rp[k];
    }
    w += pa * cm + pb * sm;
    const double cn = cm * ct - sm * st;
    const double sn = sm * ct + cm * st;
    cm = cn;
    sm = sn;
  }
  return w;
}

// Foci per launch of the multi-focus pupil kernel (carried in its arguments).
constexpr int kMaxFoci = 32;

struct PupilParams {
  ZernPoly z;
  double lam, na, mag, n, focus, pr, pc;
  double foci[kMaxFoci];  // the launch's foci (foci[0] = focus for one focus)
  int vector;
};

// The focus-independent factors of the pupil at normalised frequency (f, g),
// rho^2 = f^2 + g^2 <= 1: obliquity, twice the wavefront W in waves, and
// twice the axial wavenumber sqrt(n^2 - (NA/M)^2 rho^2).
struct PupilBase {
  double ob, w2, kz2;
};

__device__ PupilBase pupil_base(const PupilParams& P, double f, double g, double rho2) {
  const double rho = sqrt(rho2);
  double ct = 1.0, st = 0.0;
  if (rho > 0.0) {
    ct = f / rho;
    st = g / rho;
  }
  const double nm = P.na / P.mag;
  PupilBase b;
  // x^(1/4) as two square roots (correctly rounded each) instead of pow's
  // log and exp
  b.ob = sqrt(sqrt((1.0 - nm * nm * rho2) / (1.0 - P.na * P.na * rho2)));
  b.w2 = 2.0 * zern_eval(P.z, rho, ct, st);
  b.kz2 = 2.0 * sqrt(P.n * P.n - nm * nm * rho2);
  return b;
}

// obliquity x exp(2 pi i W) x defocus at focus z (CalculateAerialImage.py:64-67):
// phase / pi = 2 W + 2 sqrt(n^2 - (NA/M)^2 rho^2) z / lambda
__device__ __forceinline__ double2 pupil_at(const PupilBase& b, double z, double lam) {
  const double ph = b.w2 + b.kz2 * z / lam;
  double s, c;
  sincospi(ph, &s, &c);
  return make_double2(b.ob * c, b.ob * s);
}


__device__ __forceinline__ double2 dmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 ddiv(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ __forceinline__ double2 dcsqrt(double x) {
  return x >= 0.0 ? make_double2(sqrt(x), 0.0) : make_double2(0.0, sqrt(-x));
}

// Characteristic matrix entries (CalculateCharacteristicMatrix.py:3-16).
__device__ void char_matrix(double f, double g, double na, double n, double2* M6) {
  const double a = (na / n) * f, b = (na / n) * g;
  const double2 gam = dcsqrt(1.0 - a * a - b * b);
  const double2 den = make_double2(1.0 + gam.x, gam.y);
  const double2 ia = ddiv(make_double2(1.0, 0.0), den);
  M6[0] = make_double2(1.0 - a * a * ia.x, -a * a * ia.y);  // Mxx
  M6[1] = make_double2(-a * b * ia.x, -a * b * ia.y);       // Myx
  M6[2] = M6[1];                                            // Mxy
  M6[3] = make_double2(1.0 - b * b * ia.x, -b * b * ia.y);  // Myy
  M6[4] = make_double2(-a, 0.0);                            // Mxz
  M6[5] = make_double2(-b, 0.0);                            // Myz
}

// Grid (window blocks, source points): workgroup (x, y) takes window entries
// x*256 .. x*256+255 of source points y, y + gridDim.y, ..., so the point's
// values are uniform over the workgroup and no 64-bit division is needed.
// The focus-independent factors (obliquity, wavefront, polarisation) are
// formed once per entry and the nfoci kernel sets written from them:
// G[f][c][s][i][j] at the focus P.foci[f], f < nfoci <= kMaxFoci.
__global__ __launch_bounds__(256) void k_pupil(PupilParams P, int nfoci,
                                               const double* __restrict__ sx,
                                               const double* __restrict__ sy,
                                               const double* __restrict__ scale,
                                               const double2* __restrict__ px,
                                               const double2* __restrict__ py, int nsrc, int k0r,
                                               int wr, int k0c, int wc, float2* __restrict__ G) {
  const int per = wr * wc;
  const int rem = blockIdx.x * blockDim.x + threadIdx.x;
  if (rem >= per) return;
  const int ncomp = P.vector ? 3 : 1;
  const size_t set = (size_t)ncomp * nsrc * per;   // one focus's kernels
  const int i = rem / wc, j = rem - i * wc;
  for (int s = blockIdx.y; s < nsrc; s += gridDim.y) {
    const double f = (double)(k0r + i) / P.pr + sx[s];
    const double g = (double)(k0c + j) / P.pc + sy[s];
    const double rho2 = f * f + g * g;
    float2* out = G + (size_t)s * per + rem;
    if (!(rho2 <= 1.0)) {
      for (int z = 0; z < nfoci; ++z)
        for (int c = 0; c < ncomp; ++c) out[z * set + (size_t)c * nsrc * per] = make_float2(0.f, 0.f);
      continue;
    }
    const PupilBase b = pupil_base(P, f, g, rho2);
    const double sc = scale[s];
    double2 pol[3] = {make_double2(1.0, 0.0), make_double2(0, 0), make_double2(0, 0)};
    if (P.vector) {
      double2 M6[6];
      char_matrix(f, g, P.na, P.n, M6);
      const double2 ex = px[s], ey = py[s];
      for (int c = 0; c < 3; ++c) {
        // pol_c = px M_xc + py M_yc with (M_xc, M_yc) = (M6[0],M6[1]), (M6[2],M6[3]), (M6[4],M6[5])
        const double2 t0 = dmul(ex, M6[2 * c]);
        const double2 t1 = dmul(ey, M6[2 * c + 1]);
        pol[c] = make_double2(t0.x + t1.x, t0.y + t1.y);
      }
    }
    for (int z = 0; z < nfoci; ++z) {
      const double2 p = pupil_at(b, P.foci[z], P.lam);
      for (int c = 0; c < ncomp; ++c) {
        const double2 q = P.vector ? dmul(pol[c], p) : p;
        out[z * set + (size_t)c * nsrc * per] = make_float2((float)(q.x * sc), (float)(q.y * sc));
      }
    }
  }
}

__global__ void k_zernike(const float* __restrict__ rho, const float* __restrict__ theta, int n,
                          ZernPoly z, float* __restrict__ out) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    double s, c;
    sincos((double)theta[e], &s, &c);
    out[e] = (float)zern_eval(z, (double)rho[e], c, s);
  }
}

__global__ void k_char_matrix(const float* __restrict__ f, const float* __restrict__ g, int n,
                              double na, double idx, float2* __restrict__ out) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    double2 M6[6];
    char_matrix((double)f[e], (double)g[e], na, idx, M6);
    for (int c = 0; c < 6; ++c) out[(size_t)c * n + e] = make_float2((float)M6[c].x, (float)M6[c].y);
  }
}

static int blocks_for(size_t total) {
  size_t b = (total + 255) / 256;
  if (b > 8192) b = 8192;
  return b < 1 ? 1 : (int)b;
}

}  // namespace oi

using namespace oi;

static int launch_pupil(const char* who, const oi_optics* opt, const double* foci, int nfoci,
                        const double* sx, const double* sy, const double* scale, const double* px,
                        const double* py, int nsrc, int k0r, int wr, int k0c, int wc, void* G,
                        void* stream) {
  const std::string w(who);
  if (!opt || !sx || !sy || !scale || !G || nsrc < 1 || wr < 1 || wc < 1)
    return fail(OI_ERR_INVALID, w + ": bad arguments");
  if (opt->vector && (!px || !py)) return fail(OI_ERR_INVALID, w + ": vector mode needs px, py");
  if (!(opt->pitch_r > 0) || !(opt->pitch_c > 0) || !(opt->wavelength > 0))
    return fail(OI_ERR_INVALID, w + ": pitches and wavelength must be positive");
  if ((size_t)wr * wc > (size_t)INT_MAX - 256) return fail(OI_ERR_INVALID, w + ": window too large");
  PupilParams P;
  P.z = fold_zernike(opt->zernike);
  P.lam = opt->wavelength;
  P.na = opt->na;
  P.mag = opt->magnification;
  P.n = opt->index_image;
  P.focus = opt->focus;
  P.pr = opt->pitch_r;
  P.pc = opt->pitch_c;
  P.vector = opt->vector ? 1 : 0;
  const dim3 grid((unsigned)((wr * wc + 255) / 256), (unsigned)(nsrc < 65535 ? nsrc : 65535));
  const size_t set = (size_t)(P.vector ? 3 : 1) * nsrc * wr * wc;   // one focus's kernels
  for (int z0 = 0; z0 < nfoci; z0 += kMaxFoci) {
    const int nz = nfoci - z0 < kMaxFoci ? nfoci - z0 : kMaxFoci;
    for (int z = 0; z < kMaxFoci; ++z) P.foci[z] = z < nz ? (foci ? foci[z0 + z] : opt->focus) : 0.0;
    hipLaunchKernelGGL(k_pupil, grid, dim3(256), 0, (hipStream_t)stream, P, nz, sx, sy, scale,
                       reinterpret_cast<const double2*>(px), reinterpret_cast<const double2*>(py),
                       nsrc, k0r, wr, k0c, wc, static_cast<float2*>(G) + (size_t)z0 * set);
    OI_LAUNCH_CHECK("pupil_kernels");
  }
  return OI_OK;
}

extern "C" int oi_pupil_kernels(const oi_optics* opt, const double* sx, const double* sy,
                                const double* scale, const double* px, const double* py, int nsrc,
                                int k0r, int wr, int k0c, int wc, void* G, void* stream) {
  return launch_pupil("pupil_kernels", opt, nullptr, 1, sx, sy, scale, px, py, nsrc, k0r, wr, k0c,
                      wc, G, stream);
}

extern "C" int oi_pupil_kernels_foci(const oi_optics* opt, const double* foci, int nfoci,
                                     const double* sx, const double* sy, const double* scale,
                                     const double* px, const double* py, int nsrc, int k0r, int wr,
                                     int k0c, int wc, void* G, void* stream) {
  if (!foci || nfoci < 1) return fail(OI_ERR_INVALID, "pupil_kernels_foci: foci must hold nfoci >= 1 values");
  return launch_pupil("pupil_kernels_foci", opt, foci, nfoci, sx, sy, scale, px, py, nsrc, k0r, wr,
                      k0c, wc, G, stream);
}

extern "C" int oi_zernike(const float* rho, const float* theta, int n, const double* coeffs,
                          float* out, void* stream) {
  if (!rho || !theta || !coeffs || !out || n < 0) return fail(OI_ERR_INVALID, "zernike: bad arguments");
  if (n == 0) return OI_OK;
  ZernPoly z = fold_zernike(coeffs);
  hipLaunchKernelGGL(k_zernike, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, rho, theta,
                     n, z, out);
  OI_LAUNCH_CHECK("zernike");
  return OI_OK;
}

extern "C" int oi_char_matrix(const float* f, const float* g, int n, double na, double index_image,
                              void* out, void* stream) {
  if (!f || !g || !out || n < 0) return fail(OI_ERR_INVALID, "char_matrix: bad arguments");
  if (n == 0) return OI_OK;
  hipLaunchKernelGGL(k_char_matrix, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, f, g, n,
                     na, index_image, static_cast<float2*>(out));
  OI_LAUNCH_CHECK("char_matrix");
  return OI_OK;
}
