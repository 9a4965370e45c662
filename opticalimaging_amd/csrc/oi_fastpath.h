// Fast-path (compile-time transform sizes) imaging launches.
#pragma once

#include "oi_internal.h"

namespace oi {

struct Chans {
  int c[4];
};

// Shared launch arguments of the fast path.
// Internal field kind of the fast path: a real field given as logits, the
// transformed field being sigmoid(field) (the ILT step's mask parametrisation).
constexpr int kFieldLogits = 16;
// sigmoid from the hardware exponential and reciprocal (a few ulp; 0 and 1
// at the ends), cheap enough to apply to every pixel a pass streams
__device__ __forceinline__ float logistic(float x) {
  return __builtin_amdgcn_rcpf(1.f + __expf(-x));
}

struct FpArgs {
  const void* field;
  int kind, cin, nchan;
  Chans chans;
  int batch, ny, nx;
  int k0r, wr, k0c, wc;
  int L;          // square coarse grid
  int qr, qc;     // band half-widths
  const float2* tw_x;  // exp(-2 pi i m / nx)
  const float2* tw_y;  // exp(-2 pi i m / ny)
  const float2* tw_l;  // exp(-2 pi i m / L)
  float2* T;      // [B][C][wc][ny]
  float2* spec;   // [B][C][wr][wc]
  float2* R;      // [B][G][L][qc+1]
  float2* Ihat;   // [B][qc+1][2qr+1]
  float2* U;      // [B][qc+1][ny]
  float* out;     // [B][ny][nx]
  const float* target;  // non-null: out = 2 (I - target), part gets the sums of (I - target)^2
  const float* theta;   // non-null (target null): out = I m (1 - m), m = logistic(theta),
                        // part gets the sums of out^2
  float* part;          // [B][fp_urows_blocks(ny, nx, B)]
  StageMark mark;
  void* mark_ctx;
};

struct FpCoh {
  const float2* spec;
  const float2* kern;
  int nchan, nkern;
  int k0r, wr, k0c, wc;
  int groups;
  int qc;
  const float2* tw_l;  // exp(-2 pi i m / L) (stage twiddles and phase twist)
  const float2* tw_m;  // column-stage transform table
  float2* R;
};

bool fp_image_size_ok(int n);
bool fp_coarse_size_ok(int l);
// Radices (R1, R2) of the two-stage register transform of length n, false
// when n has none.  Twiddle tables of length n carry 2n entries: [0, n) holds
// exp(-2 pi i m / n) and [n, 2n) the stage-2 twiddles of that transform,
// entry n + j R1 + k = exp(-2 pi i j k / n) (j < R2, k < R1), so that the
// lanes of a transform group read consecutive entries.
bool fp_stage_radices(int n, int* r1, int* r2);
int fp_coarse_choose(int need);          // smallest table size >= need, 0 if none
size_t fp_coherent_lds(int L, int wr, int wc);
int fp_coherent_rows_fit(int L);
bool fp_one_pass_ok(int L);             // the one-pass coherent kernel serves L
int fp_spectrum(const FpArgs& a, hipStream_t st);  // marks stage 1 between its launches
int fp_image(const FpArgs& a, const FpCoh& c, hipStream_t st);  // marks stages 3, 4, 5
int fp_interp(const FpArgs& a, int groups, hipStream_t st);      // marks stages 4, 5
// the one-pass coherent sum and the band Ihat of the image, without the image
int fp_image_band(const FpArgs& a, const FpCoh& c, hipStream_t st);
// the band Ihat from partial sums R (the first kernel of fp_interp)
int fp_band(const FpArgs& a, int groups, hipStream_t st);
// interpolation column and row transforms of a band already in a.Ihat
int fp_band_image(const FpArgs& a, hipStream_t st);
// row sets per workgroup and workgroups per image of the image-row pass for a
// batch (its residual form writes one partial sum per workgroup)
int fp_urows_sets(int ny, int nx, int batch);
int fp_urows_blocks(int ny, int nx, int batch);
// out[0] = sum of part[0 .. n) in a fixed order (oi_ilt.hip)
int ilt_sum(const float* part, int n, float* out, hipStream_t st);
// Band form of the least-squares loss (oi_ilt.hip).  Ihat [batch][qc + 1][2 qr + 1]
// is the image band (the layout k_fp_band writes); tband [batch][2 qr + 1][2 qc + 1]
// the targets' band spectrum and tconst[b] = sum T^2 - n sum_band |T~|^2.
// Writes the cotangent spectrum 2 (S - T~) to specg (tband's layout) and
// ilt_band_blocks(qr, qc) partial sums of the loss per image to part.
int ilt_band_blocks(int qr, int qc);
int ilt_band_cotangent(const float2* Ihat, int qr, int qc, const float2* tband,
                       const float* tconst, float n, int batch, float2* specg, float* part,
                       hipStream_t st);
// theta -= lr gt / rms(gt) per image, the rms from nparts partial sums of
// gt^2 per image (part [batch][nparts], added in a fixed order)
int ilt_update_theta(float* theta, const float* gt, float lr, int batch, size_t npix,
                     const float* part, int nparts, hipStream_t st);
// tconst[b] = sum_j T[b][j]^2 - n sum_band |tband[b]|^2 (accumulated in
// float64); scratch: ilt_target_const_scratch(batch) bytes of device memory
size_t ilt_target_const_scratch(int batch);
int ilt_target_const(const float* target, size_t npix, const float2* tband, size_t nband,
                     int batch, void* scratch, float* tconst, hipStream_t st);
// Two-pass coherent sum for coarse grids L in {256, 512}: B holds
// batch * nkern * nchan * wc * L c64; R in [img][q][t] (generic_layout = 0) or
// [img][t][q] (1).
bool fp_two_pass_size_ok(int l);
int fp_two_pass_choose(int need);
size_t fp_two_pass_lds(int L, int wc);
int fp_two_pass(const FpArgs& a, const FpCoh& c, float2* B, int generic_layout, hipStream_t st);

}  // namespace oi
