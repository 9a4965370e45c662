// Float64 Hermitian eigensolver: parallel two-sided cyclic Jacobi.
//
// Indices are paired by the round-robin (circle) schedule: m = n rounded up to
// even, m-1 rounds per sweep, m/2 disjoint pairs per round (index n, when n is
// odd, is a dummy).  A round computes all m/2 rotations from the current
// matrix, then updates every 2x2 block (pair i, pair j) as U_i^H A_ij U_j and
// the eigenvector columns as V U_j.  For pair (p, q) with c = A_pq = |c| e^{i phi},
//   tau = (A_qq - A_pp) / (2|c|),  t = sign(tau) / (|tau| + sqrt(1 + tau^2)),
//   cs = 1/sqrt(1 + t^2), sn = t cs,
//   U = [[cs, sn], [-sn e^{-i phi}, cs e^{-i phi}]]
// zeroes A_pq (phase removal followed by the real Jacobi rotation).  Sweeps
// stop when the off-diagonal Frobenius norm falls below tol times the diagonal
// norm.
#include <cfloat>
#include <cmath>
#include <string>

#include "oi_internal.h"

namespace oi {

struct Rot {
  double c, s, er, ei;  // cos, sin, e^{i phi}
};

__device__ __forceinline__ void pair_of(int i, int s, int m, int& p, int& q) {
  if (i == 0) {
    p = 0;
    q = 1 + pmod(s - 1, m - 1);
  } else {
    p = 1 + (s + i - 1) % (m - 1);
    q = 1 + (s + m - 2 - i) % (m - 1);
  }
}

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

__global__ void k_eye(double2* __restrict__ V, int n) {
  const size_t total = (size_t)n * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / n), c = (int)(e - (size_t)r * n);
    V[e] = make_double2(r == c ? 1.0 : 0.0, 0.0);
  }
}

__global__ void k_rot(const double2* __restrict__ A, int n, int m, int s, Rot* __restrict__ rot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m / 2) return;
  int p, q;
  pair_of(i, s, m, p, q);
  Rot r{1.0, 0.0, 1.0, 0.0};
  if (p < n && q < n) {
    const double app = A[(size_t)p * n + p].x, aqq = A[(size_t)q * n + q].x;
    const double2 apq = A[(size_t)p * n + q];
    const double mag = hypot(apq.x, apq.y);
    if (mag > 0.0) {
      const double tau = (aqq - app) / (2.0 * mag);
      const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
      r.c = 1.0 / sqrt(1.0 + t * t);
      r.s = t * r.c;
      r.er = apq.x / mag;
      r.ei = apq.y / mag;
    }
  }
  rot[i] = r;
}

// Applies one round: items [0, h*h) are 2x2 blocks of A, items [h*h, h*h + n*h)
// are (row, pair) updates of V.  h = m/2.
__global__ void k_apply(double2* __restrict__ A, double2* __restrict__ V, int n, int m, int s,
                        const Rot* __restrict__ rot) {
  const int h = m / 2;
  const size_t nblk = (size_t)h * h;
  const size_t total = nblk + (size_t)n * h;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    if (e < nblk) {
      const int i = (int)(e / h), j = (int)(e - (size_t)i * h);
      int pi, qi, pj, qj;
      pair_of(i, s, m, pi, qi);
      pair_of(j, s, m, pj, qj);
      const Rot ri = rot[i], rj = rot[j];
      const int R[2] = {pi, qi}, C[2] = {pj, qj};
      double2 X[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          X[a][b] = (R[a] < n && C[b] < n) ? A[(size_t)R[a] * n + C[b]] : make_double2(0.0, 0.0);
      // T = X U_j with U_j = [[c, s], [-s conj(e), c conj(e)]]
      const double2 ej = zconj(make_double2(rj.er, rj.ei));
      double2 T[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        T[a][0] = zadd(zscale(X[a][0], rj.c), zscale(zmul(X[a][1], ej), -rj.s));
        T[a][1] = zadd(zscale(X[a][0], rj.s), zscale(zmul(X[a][1], ej), rj.c));
      }
      // Y = U_i^H T with U_i^H = [[c, -s e], [s, c e]]
      const double2 ei = make_double2(ri.er, ri.ei);
      double2 Y[2][2];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        Y[0][b] = zadd(zscale(T[0][b], ri.c), zscale(zmul(T[1][b], ei), -ri.s));
        Y[1][b] = zadd(zscale(T[0][b], ri.s), zscale(zmul(T[1][b], ei), ri.c));
      }
      if (i == j) {
        Y[0][1] = make_double2(0.0, 0.0);
        Y[1][0] = make_double2(0.0, 0.0);
        Y[0][0].y = 0.0;
        Y[1][1].y = 0.0;
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          if (R[a] < n && C[b] < n) A[(size_t)R[a] * n + C[b]] = Y[a][b];
    } else {
      const size_t f = e - nblk;
      const int k = (int)(f / h), j = (int)(f - (size_t)k * h);
      int pj, qj;
      pair_of(j, s, m, pj, qj);
      if (pj >= n || qj >= n) continue;
      const Rot rj = rot[j];
      const double2 ej = zconj(make_double2(rj.er, rj.ei));
      const double2 vp = V[(size_t)k * n + pj], vq = V[(size_t)k * n + qj];
      const double2 vqe = zmul(vq, ej);
      V[(size_t)k * n + pj] = zadd(zscale(vp, rj.c), zscale(vqe, -rj.s));
      V[(size_t)k * n + qj] = zadd(zscale(vp, rj.s), zscale(vqe, rj.c));
    }
  }
}

// out[0] = sum_{p != q} |A_pq|^2, out[1] = sum_p A_pp^2 (single block).
__global__ __launch_bounds__(1024) void k_offnorm(const double2* __restrict__ A, int n,
                                                  double* __restrict__ out) {
  __shared__ double so[1024], sd[1024];
  double off = 0.0, dg = 0.0;
  const size_t total = (size_t)n * n;
  for (size_t e = threadIdx.x; e < total; e += blockDim.x) {
    const int r = (int)(e / n), c = (int)(e - (size_t)r * n);
    const double2 v = A[e];
    const double a2 = v.x * v.x + v.y * v.y;
    if (r == c)
      dg += a2;
    else
      off += a2;
  }
  so[threadIdx.x] = off;
  sd[threadIdx.x] = dg;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      so[threadIdx.x] += so[threadIdx.x + w];
      sd[threadIdx.x] += sd[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = so[0];
    out[1] = sd[0];
  }
}

__global__ void k_diag(const double2* __restrict__ A, int n, double* __restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = A[(size_t)i * n + i].x;
}

static int blocks_for(size_t total) {
  size_t b = (total + 255) / 256;
  if (b > 4096) b = 4096;
  return b < 1 ? 1 : (int)b;
}

}  // namespace oi

using namespace oi;

extern "C" size_t oi_eigh_work_bytes(int n) {
  if (n < 1) return 0;
  const int m = n + (n & 1);
  return (size_t)(m / 2) * sizeof(Rot) + 2 * sizeof(double);
}

extern "C" int oi_eigh_hermitian(void* A, int n, double* w, void* V, int max_sweeps, double tol,
                                 void* work, int* sweeps_out, void* stream) {
  if (!A || !w || !V || !work || n < 1 || max_sweeps < 0)
    return fail(OI_ERR_INVALID, "eigh_hermitian: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  double2* a = static_cast<double2*>(A);
  double2* v = static_cast<double2*>(V);
  const int m = n + (n & 1);
  Rot* rot = static_cast<Rot*>(work);
  double* norms = reinterpret_cast<double*>(static_cast<char*>(work) + (size_t)(m / 2) * sizeof(Rot));
  hipLaunchKernelGGL(k_eye, dim3(blocks_for((size_t)n * n)), dim3(256), 0, st, v, n);
  OI_LAUNCH_CHECK("eigh_eye");
  int sweep = 0;
  if (n > 1) {
    const int h = m / 2;
    const int rb = (h + 255) / 256;
    const int ab = blocks_for((size_t)h * h + (size_t)n * h);
    // A sweep count that runs out still passes when the off-diagonal norm is
    // at the rounding level of an n x n rotation sequence.
    const double acc = tol > 4.0 * n * DBL_EPSILON ? tol : 4.0 * n * DBL_EPSILON;
    for (;; ++sweep) {
      hipLaunchKernelGGL(k_offnorm, dim3(1), dim3(1024), 0, st, a, n, norms);
      OI_LAUNCH_CHECK("eigh_offnorm");
      double host[2];
      OI_HIP_CHECK(hipMemcpyAsync(host, norms, sizeof(host), hipMemcpyDeviceToHost, st));
      OI_HIP_CHECK(hipStreamSynchronize(st));
      if (!std::isfinite(host[0]) || !std::isfinite(host[1]))
        return fail(OI_ERR_RUNTIME, "eigh_hermitian: matrix norm is not finite");
      if (!(host[0] > tol * tol * host[1])) break;
      if (sweep == max_sweeps) {
        if (host[0] > acc * acc * host[1])
          return fail(OI_ERR_RUNTIME, "eigh_hermitian: no convergence in " +
                                          std::to_string(max_sweeps) + " sweeps");
        break;
      }
      for (int s = 0; s < m - 1; ++s) {
        hipLaunchKernelGGL(k_rot, dim3(rb), dim3(256), 0, st, a, n, m, s, rot);
        hipLaunchKernelGGL(k_apply, dim3(ab), dim3(256), 0, st, a, v, n, m, s, rot);
      }
      OI_LAUNCH_CHECK("eigh_round");
    }
  }
  hipLaunchKernelGGL(k_diag, dim3((n + 255) / 256), dim3(256), 0, st, a, n, w);
  OI_LAUNCH_CHECK("eigh_diag");
  if (sweeps_out) *sweeps_out = sweep;
  return OI_OK;
}
