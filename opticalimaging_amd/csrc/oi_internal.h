// Internal declarations shared by the library's translation units.
#pragma once

#include "oi_common.h"
#include "oi_fft.h"

namespace oi {

constexpr int kThreads = 256;       // threads per workgroup for all imaging kernels
constexpr int kAccMax = 24;         // coarse-intensity values held per thread
constexpr size_t kLdsLimit = 160 * 1024;

// Arguments of the coarse-grid coherent-sum kernel.
struct CohArgs {
  const float2* spec;  // [B][nchan][wr][wc]
  const float2* kern;  // [nkern][wr][wc]
  int nchan, nkern;
  int k0r, wr, k0c, wc;
  int lr, lc, rows, phases;
  int groups;
  int qc;
  int direct;
  FftDesc frows, flc;
  const float2* tw_lr;  // exp(-2 pi i m / lr)
  float2* R;            // non-direct: [B][groups][lr][qc+1]
  float* Ic;            // direct: [B][groups][lr][lc]
};

struct InterpArgs {
  const float2* R;  // [B][groups][lr][qc+1]
  int groups;
  int lr, lc, ny, nx, qr, qc;
  float scale;      // 1 / (lr lc)
  FftDesc flr, fny, fnx;
  float2* U;        // [B][ny][qc+1]
  float* out;       // [B][ny][nx]
};

// Launchers (return OI_OK or a negative status; errors recorded).
int launch_spec_rows(const void* field, int kind, int cin, const int* chans, int nchan,
                     int batch, int ny, int nx, const FftDesc& fx, int k0c, int wc,
                     float2* T, hipStream_t st);
int launch_spec_cols(const float2* T, int nbc, int ny, int nx, const FftDesc& fy, int k0r,
                     int wr, int wc, float2* spec, hipStream_t st);
int launch_coherent(const CohArgs& a, int batch, hipStream_t st);
// mark(ctx, stage) is called after each of the two launches (stages 4, 5).
typedef void (*StageMark)(void* ctx, int stage);
int launch_interp(const InterpArgs& a, int batch, hipStream_t st, StageMark mark, void* ctx);
int launch_group_sum(const float* Ic, int batch, int groups, size_t per_image, float* out,
                     hipStream_t st);
int launch_full_rows(const void* field, int kind, int cin, const int* chans, int nchan,
                     int batch, int ny, int nx, const FftDesc& fx, float2* work, hipStream_t st);
int launch_full_cols(const float2* work, int nchan, int batch, int ny, int nx,
                     const FftDesc& fy, float scale, float2* hk, hipStream_t st);

size_t coherent_lds_bytes(int wc, int rows, int lc);

// Twiddle tables for a set of sizes, uploaded to one device allocation.
struct TwiddlePool {
  float2* dev = nullptr;
  size_t bytes = 0;
};
// sizes: list of FFT lengths; descs receive factor plans and table pointers.
int make_twiddles(const int* sizes, int count, FftDesc* descs, TwiddlePool& pool);
void free_twiddles(TwiddlePool& pool);

}  // namespace oi
