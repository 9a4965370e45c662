// Adjoint (mask gradient) of the SOCS image: internal interface.
#pragma once

#include "oi_internal.h"

namespace oi {

struct AdjGeom {
  int ny, nx;
  int k0r, wr, k0c, wc;
  int lr, lc, qr, qc;
  int direct;
  FftDesc fx, fy, flr, flc;
};

size_t adjoint_work_bytes(const AdjGeom& g, int batch, int nkern);

// Where launch_adjoint keeps the cotangent's band spectrum in `work`:
// c64 [batch][2 qr + 1][2 qc + 1], element [b][i][j] at frequency
// (i - qr, j - qc) (null on direct plans).
float2* adjoint_cotangent_spectrum(const AdjGeom& g, int batch, void* work);
// out[b][i][j] = (1 / (ny nx)) sum_{y,x} img[b][y][x] e^{+2 pi i ((i - qr) y / ny + (j - qc) x / nx)}
// for real images [batch][ny][nx]; the first region of `work` (an adjoint
// work buffer) is its scratch, so out must not lie there.
int adjoint_band_spectrum(const AdjGeom& g, const float* img, int batch, void* work, float2* out,
                          hipStream_t st);

// grad[b] = d/d mask[b] of sum_j gI[b][j] I[b][j], I the SOCS image of the real
// masks whose windowed spectrum is spec (c64 [batch][wr][wc]).  bcols (or
// null): the two-pass column buffer [batch][nkern][wc][lr] that the forward
// pass of the same masks and kernels left (k_fp_bcols); when given on a
// square register-size coarse grid, the adjoint reads its column stage from
// there instead of recomputing it (same values).
// gI null: the cotangent's band spectrum (the layout adjoint_band_spectrum
// writes) is already at adjoint_cotangent_spectrum(g, batch, work); not on
// direct plans, which take the cotangent image itself.
// chain_theta non-null (needs adjoint_chain_ok): the masks are
// m = logistic(chain_theta) and grad receives the gradient with respect to
// theta, grad m (1 - m), with its per-workgroup sums of squares in chain_part
// ([batch][fp_urows_blocks(ny, nx, batch)]).
bool adjoint_chain_ok(const AdjGeom& g);
int launch_adjoint(const AdjGeom& g, const float* gI, const float2* spec, int batch,
                   const float2* kern, int nkern, void* work, float* grad, hipStream_t st,
                   const float2* bcols = nullptr, const float* chain_theta = nullptr,
                   float* chain_part = nullptr);

}  // namespace oi
