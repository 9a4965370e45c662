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

// grad[b] = d/d mask[b] of sum_j gI[b][j] I[b][j], I the SOCS image of the real
// masks whose windowed spectrum is spec (c64 [batch][wr][wc]).  bcols (or
// null): the two-pass column buffer [batch][nkern][wc][lr] that the forward
// pass of the same masks and kernels left (k_fp_bcols); when given on a
// square register-size coarse grid, the adjoint reads its column stage from
// there instead of recomputing it (same values).
int launch_adjoint(const AdjGeom& g, const float* gI, const float2* spec, int batch,
                   const float2* kern, int nkern, void* work, float* grad, hipStream_t st,
                   const float2* bcols = nullptr);

}  // namespace oi
