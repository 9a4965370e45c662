/*
 * C ABI of the MI355X aerial-image library (libopticalimaging_amd).
 *
 * The reference (gary21978/OpticalImaging) is pure Python over torch and has no
 * FFI: its imaging entry points are Python functions taking parameter objects.
 * Each entry point below replaces the numerical core of one of them; the
 * Python package `opticalimaging_amd` keeps the reference signatures and calls
 * these through ctypes with device pointers of torch tensors.
 *
 * Conventions
 *   - Every function returns an int status: OI_OK (0) or a negative code.
 *     OI_ERR_INVALID maps to Python ValueError, the others to RuntimeError.
 *     oi_last_error() returns the message of the last failure on the calling
 *     thread.
 *   - All array arguments are caller-owned device pointers.  Complex arrays are
 *     interleaved float32 (re, im) pairs ("c64").  Layouts are row-major with
 *     the last index fastest.
 *   - Plans are library-owned; they hold twiddle tables and scratch sized at
 *     creation.  Calls that take a plan do not allocate device memory.
 *   - Work is enqueued on the given HIP stream; no call synchronises the device
 *     unless its comment says so.  A plan is used by one host thread at a time.
 *   - No C++ exception crosses this boundary.
 *
 * Frequency conventions (shared by all entry points)
 *   A grid of ny x nx samples with pitches Period_Y, Period_X has spectrum
 *   index kappa = (kr, kc), kr in [-(ny/2), ny-1-ny/2] (integer division), and
 *   normalised frequency (kr / p_y, kc / p_x) with p = Period * NA / lambda.
 *   Rows go with Period_Y and with the source X coordinate
 *   (CalculateAerialImage.py:31,55; CalculateTCCMatrix.py:12,23).
 *   A "window" is the rectangle kr in [k0r, k0r+wr), kc in [k0c, k0c+wc).
 *   The windowed spectrum of a field h is
 *       spec(kappa) = (1/(ny nx)) sum_m h[m] exp(+2 pi i kappa.m / n)
 *   i.e. Scatter.CalculateSpectrum's Hk (Scatter.py:20-23) without its p_x p_y
 *   factor and its fftshift phase.
 */
#ifndef OPTICALIMAGING_AMD_H
#define OPTICALIMAGING_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OI_OK 0
#define OI_ERR_INVALID (-1)     /* bad argument or configuration -> ValueError */
#define OI_ERR_RUNTIME (-2)     /* HIP runtime failure -> RuntimeError         */
#define OI_ERR_UNSUPPORTED (-3) /* valid request this build cannot serve       */
#define OI_ERR_NOMEM (-4)       /* device allocation failed                    */

/* Field encodings accepted by the spectrum entry points. */
#define OI_FIELD_REAL 0         /* float32 [batch][ny][nx], one channel (thin mask)  */
#define OI_FIELD_C64_CHLAST 1   /* c64 [batch][ny][nx][cin] (ScatterField layout)    */

/* TCC contraction precisions. */
#define OI_TCC_SPLIT_BF16 0     /* 3-term split-bf16 on MFMA, fp32 accumulation      */
#define OI_TCC_FP32 1           /* fp32 MFMA                                         */
#define OI_TCC_FP64 2           /* fp64 MFMA                                         */

/* Last error message of the calling thread ("" if none). */
const char* oi_last_error(void);

/* Library version, major*10000 + minor*100 + patch. */
int oi_version(void);

/* ------------------------------------------------------------------ plan --
 * Geometry of one imaging problem.  Inputs to oi_plan_create: ny, nx, the
 * spectrum window, nchan (channels carried per image), max_batch, max_groups
 * (kernels are split into this many partial sums, fixed for the plan so a
 * result never depends on the batch it is computed in; 0 = choose),
 * max_kernels (the largest kernel count any call on the plan passes; every
 * buffer that grows with it is allocated here, and calls with more kernels
 * fail with OI_ERR_INVALID; 0 = 256) and optionally lr, lc (coarse grid;
 * 0 = choose).
 * Outputs filled by the library: max_groups (when 0), lr, lc, rows (coarse rows per workgroup),
 * direct (1 when the coarse grid equals the image grid), qr, qc (half-widths
 * of the intensity band carried from the coarse to the image grid), path
 * (1 when the image sizes and a square coarse grid have compile-time-size
 * kernels; then lr = lc).
 *
 * The coarse grid must satisfy lr >= 2*band_r - 1 where band_r is the widest
 * row extent of any single coherent kernel's support (and likewise for
 * columns); oi_plan_create takes band_r = wr, band_c = wc unless band_r/band_c
 * are set (> 0), which the Abbe path uses (one shifted pupil per kernel). */
typedef struct oi_geometry {
  int ny, nx;
  int k0r, wr, k0c, wc;
  int band_r, band_c;
  int nchan;
  int max_batch;
  int max_groups;
  int lr, lc;
  int rows;
  int direct;
  int qr, qc;
  int path;       /* 0: generic kernels; 1: compile-time-size kernels, one
                     pass over the coarse grid; 2: compile-time-size kernels,
                     two-pass coherent sum (coarse 144, 256 or 512)          */
  int max_kernels;
} oi_geometry;

typedef struct oi_plan oi_plan;

int oi_plan_create(oi_geometry* geo, oi_plan** out);
int oi_plan_destroy(oi_plan* plan);
/* Device bytes held by the plan (twiddles + scratch). */
size_t oi_plan_bytes(const oi_plan* plan);

/* Per-stage timing.  With slots > 0 every later imaging call on the plan
 * records HIP events on its stream at the boundaries of its five stages
 * (spectrum rows | spectrum columns | coherent sum | interpolation columns |
 * interpolation rows; on the compile-time-size path the interpolation-column
 * stage is two launches, band extraction and column transform), keeping the
 * last `slots` calls; slots = 0 turns it off.
 * oi_plan_read_timing waits for the recorded events and returns, per stage,
 * the summed milliseconds (stage_ms[5]) over *ncalls recorded calls.        */
int oi_plan_set_timing(oi_plan* plan, int slots);
int oi_plan_read_timing(oi_plan* plan, double* stage_ms, int* ncalls);

/* Which channels of a field carry anything: for npix pixels of cin channels
 * (OI_FIELD_REAL float32 [npix][cin] or OI_FIELD_C64_CHLAST c64
 * [npix][cin], cin <= 16), ORs into the device word *mask bit 2c when
 * channel c holds a non-zero value and bit 2c + 1 when it holds a non-zero
 * imaginary part (*mask must be zero on entry).  Lets a caller skip the
 * channels that contribute nothing to an image before planning, with one
 * pass over the field instead of a reduction per channel; the reference
 * transforms every channel (CalculateAerialImage.py:76-87).              */
int oi_field_flags(const void* field, int kind, long long npix, int cin, unsigned int* mask,
                   void* stream);

/* Windowed spectrum of a batch of fields.
 * Replaces the spectrum step of Scatter.CalculateSpectrum (Scatter.py:10-23)
 * as used inside CalculateAerialImage.py:27 and CalculateAerialImage_SOCS.py:4.
 *   field : OI_FIELD_REAL float32 [batch][ny][nx]  or
 *           OI_FIELD_C64_CHLAST c64 [batch][ny][nx][cin]
 *   chans : for C64_CHLAST, the nchan channel indices to transform (< cin);
 *           ignored for REAL (nchan must be 1)
 *   spec  : out, c64 [batch][nchan][wr][wc]                                  */
int oi_spectrum_window(oi_plan* plan, const void* field, int kind, int cin,
                       const int* chans, int batch, void* spec, void* stream);

/* Coherent-sum image: out[b][y][x] = sum_c sum_k |sum_kappa
 *   kern[k](kappa) spec[b][c](kappa) exp(-2 pi i kappa.(y,x)/(ny,nx))|^2.
 * With kern = SOCS kernels this is CalculateAerialImage_SOCS
 * (CalculateAerialImage_SOCS.py:5-10); with kern = weighted shifted pupils it
 * is the source sum of CalculateAbbeImage (CalculateAerialImage.py:46-92).
 *   spec : c64 [batch][nchan][wr][wc] (from oi_spectrum_window)
 *   kern : c64 [nkern][wr][wc]
 *   out  : float32 [batch][ny][nx]                                            */
int oi_coherent_image(oi_plan* plan, const void* spec, int batch,
                      const void* kern, int nkern, float* out, void* stream);

/* oi_spectrum_window followed by oi_coherent_image; spec lives in the plan.
 * Replaces CalculateAerialImage_SOCS (CalculateAerialImage_SOCS.py:3-11) for a
 * batch of masks.                                                            */
int oi_socs_image(oi_plan* plan, const void* field, int kind, int cin,
                  const int* chans, int batch, const void* kern, int nkern,
                  float* out, void* stream);

/* No-allocation form of oi_socs_image.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It takes
 * the geometry inputs of oi_plan_create instead of a plan (max_batch and
 * max_kernels are ignored: kernel groups are chosen as for a plan of
 * `batch` images, so each image is summed as oi_socs_image sums it with such
 * a plan) and calls no hipMalloc, hipFree or other allocating routine.
 * Twiddle tables and scratch live in one static device array of
 * OI_NOALLOC_SCRATCH_BYTES per GPU, held from module load until the process
 * ends; the batch runs in passes of as many images as fit.  A geometry whose
 * single image needs more than that fails with OI_ERR_INVALID.  Every call
 * on one GPU shares the array: calls that overlap on one GPU (several
 * streams or host threads) race and give undefined images.  Enqueued on
 * `stream`; no synchronisation.                                              */
#define OI_NOALLOC_SCRATCH_BYTES (32u << 20)
int oi_socs_image_noalloc(const oi_geometry* geo, const void* field, int kind, int cin,
                          const int* chans, int batch, const void* kern, int nkern,
                          float* out, void* stream);

/* Full centred spectrum Hk = scale * fftshift(ifft2(ifftshift(h))) of one
 * batch of fields, written channels-last c64 [batch][ny][nx][nchan].
 * Replaces Scatter.CalculateSpectrum's Hk (Scatter.py:20-23) with
 * scale = p_x p_y.  work: c64 scratch of batch*nchan*ny*nx elements.
 * Creates and frees its own twiddle tables (not for hot loops).            */
int oi_spectrum_full(const void* field, int kind, int cin, const int* chans,
                     int nchan, int batch, int ny, int nx, float scale,
                     void* work, void* hk, void* stream);

/* ------------------------------------------------------ SOCS kernel sets --
 * Support of dense SOCS kernels c64 [ny][nx][nk] (DecomposeTCC_SOCS.py:24-35
 * layout).  rowflag[ny], colflag[nx] (int32, zeroed by the caller) receive 1
 * where any kernel value is non-zero.                                        */
int oi_kernel_support(const void* dense, int ny, int nx, int nk, int* rowflag,
                      int* colflag, void* stream);

/* Window of dense kernels: out c64 [nk][wr][wc], out[k][i][j] =
 * dense[(k0r+i) mod ny][(k0c+j) mod nx][k], i.e. kernel value at kappa in the
 * layout CalculateAerialImage_SOCS.py:6 multiplies with ifftshift(Hk).      */
int oi_kernel_window(const void* dense, int ny, int nx, int nk, int k0r, int wr,
                     int k0c, int wc, void* out, void* stream);

/* Dense kernels from windowed ones (inverse of oi_kernel_window; the rest of
 * the dense array is zeroed).  Writes c64 [ny][nx][nk].                     */
int oi_kernel_dense(const void* win, int nk, int k0r, int wr, int k0c, int wc,
                    int ny, int nx, void* dense, void* stream);

/* Masks sent as one byte per pixel (serving): out[i] = levels[codes[i]] for
 * i < count, float32; codes >= nlevels give 0.  levels is a host array of
 * nlevels <= 256 floats (copied into the launch); codes and out are device
 * pointers aligned to 16 bytes.  No counterpart in the reference, whose masks
 * are complex ScatterField tensors converted at use (Scatter.py:13-14); a
 * two-level attenuated-PSM clip crosses PCIe in a quarter of its float32
 * bytes this way.                                                           */
int oi_mask_decode_u8(const void* codes, size_t count, const float* levels, int nlevels,
                      float* out, void* stream);
/* Two-level masks sent as one bit per pixel (serving): out[i] = levels[b_i]
 * for i < count, with b_i = bit 7 - (i mod 8) of byte i / 8 of `bits` (the
 * order of numpy.packbits; a clip packed row by row has rows of n / 8 bytes
 * when n is a multiple of 8).  levels is a host array of 2 floats; bits is a
 * device pointer aligned to 4 bytes, out one aligned to 16.  No counterpart
 * in the reference; a binary clip crosses PCIe in 1/32 of its float32 bytes. */
int oi_mask_decode_bits(const void* bits, size_t count, const float* levels, float* out,
                        void* stream);

/* ----------------------------------------------------------- pupil / TCC --
 * Optics of one focus plane (Projection.py:5-10, Source.py:18).            */
typedef struct oi_optics {
  double wavelength;      /* nm                                            */
  double na;              /* numerical aperture                            */
  double magnification;   /* reduction ratio M                             */
  double index_image;     /* n                                             */
  double focus;           /* nm                                            */
  double zernike[37];     /* Fringe Z1..Z37, waves                         */
  double pitch_r, pitch_c;/* p_y, p_x: Period * NA / wavelength            */
  int vector;             /* 0 scalar, 1 vector (3 components)             */
} oi_optics;

/* Shifted-pupil kernels on a window.
 * G[e][s][i][j] = scale[s] * P(kr/p_r + sx[s], kc/p_c + sy[s]) * pol_e,
 * kr = k0r+i, kc = k0c+j, zero outside rho <= 1 (rho^2 decided in float64).
 * P = obliquity * exp(2 pi i W) * defocus (CalculateAerialImage.py:62-67,
 * CalculateTCCMatrix.py:25-38); pol_e = px[s] M_xe + py[s] M_ye in vector mode
 * (CalculateAerialImage.py:68-74, CalculateTCCMatrix.py:40-46), 1 in scalar.
 *   sx, sy, scale : float64 [nsrc];  px, py : c128 [nsrc] (vector mode only)
 *   G : c64 [ncomp][nsrc][wr][wc], ncomp = 3 (vector) or 1 (scalar).
 * Evaluated in float64, rounded to c64.                                      */
int oi_pupil_kernels(const oi_optics* opt, const double* sx, const double* sy,
                     const double* scale, const double* px, const double* py,
                     int nsrc, int k0r, int wr, int k0c, int wc, void* G,
                     void* stream);

/* The same kernels at nfoci foci in one pass (the pupil factors of
 * CalculateAerialImage.py:64-74, which the reference re-evaluates in every
 * pass of its focus loop, CalculateAerialImage.py:30-90): the obliquity,
 * wavefront and polarisation of each entry are evaluated once and the
 * defocus phase per focus.  foci: float64 [nfoci] in host memory (opt->focus
 * is not read).  G : c64 [nfoci][ncomp][nsrc][wr][wc]; G[f] equals what
 * oi_pupil_kernels writes with opt->focus = foci[f], bit for bit.           */
int oi_pupil_kernels_foci(const oi_optics* opt, const double* foci, int nfoci,
                          const double* sx, const double* sy, const double* scale,
                          const double* px, const double* py, int nsrc, int k0r,
                          int wr, int k0c, int wc, void* G, void* stream);

/* Hermitian TCC contraction (GetTCCMatrix, CalculateTCCMatrix.py:58-76, summed
 * over components as at :48-51 and scaled as at :55):
 *   tcc[m][m'] = alpha * sum_e sum_s w[s] G[e][s][m] conj(G[e][s][m'])
 *   G : c64 [ncomp][nsrc][M]; w : float32 [nsrc]; tcc : c64 [M][M].
 * precision: OI_TCC_SPLIT_BF16 | OI_TCC_FP32 | OI_TCC_FP64.
 * work: scratch of oi_tcc_work_bytes(M, nsrc, ncomp, precision) bytes.       */
size_t oi_tcc_work_bytes(int M, int nsrc, int ncomp, int precision);
int oi_tcc_build(const void* G, const float* w, int M, int nsrc, int ncomp,
                 double alpha, int precision, void* work, void* tcc,
                 void* stream);

/* Gram matrix of B = [sqrt(w) G_e^T]_e in float64:
 *   H[(e,s)][(e',s')] = sqrt(w_s w_s') sum_m conj(G[e][s][m]) G[e'][s'][m]
 *   G : c64 [ncomp][nsrc][M]; H : c128 [ncomp*nsrc][ncomp*nsrc].
 * Non-zero eigenpairs of H are those of the TCC (DecomposeTCC_SOCS.py:7).   */
int oi_tcc_gram(const void* G, const double* w, int M, int nsrc, int ncomp,
                void* H, void* stream);

/* Kernels from Gram eigenvectors: out[k][m] = sum_(e,s) sqrt(w_s) G[e][s][m]
 * Q[(e,s)][k], float64 accumulation.  Q : c128 [ncomp*nsrc][nk] (columns are
 * eigenvectors, already scaled as wanted); out : c64 [nk][M].              */
int oi_kernels_from_gram(const void* G, const double* w, int M, int nsrc,
                         int ncomp, const void* Q, int nk, void* out,
                         void* stream);

/* Hermitian eigensolver, float64 (parallel cyclic Jacobi).  Replaces the
 * torch.svd of DecomposeTCC_SOCS.py:7 (for a Hermitian PSD TCC the singular
 * values and left vectors are |eigenvalues| and eigenvectors).
 *   A : c128 [n][n] Hermitian, overwritten (diagonalised)
 *   w : float64 [n] eigenvalues (unsorted, in V's column order)
 *   V : c128 [n][n] eigenvectors as columns
 *   work : oi_eigh_work_bytes(n) bytes; sweeps_out : host int or NULL.
 * Synchronises the stream once per sweep (convergence test: off-diagonal
 * Frobenius norm <= tol times the matrix's).  Returns OI_ERR_RUNTIME when a
 * norm is not finite, or when max_sweeps sweeps leave the off-diagonal norm
 * above max(tol, 4 n DBL_EPSILON) times the matrix's.                      */
size_t oi_eigh_work_bytes(int n);
int oi_eigh_hermitian(void* A, int n, double* w, void* V, int max_sweeps, double tol,
                      void* work, int* sweeps_out, void* stream);

/* -------------------------------------------------------------- elementwise
 * Fringe Zernike wavefront W(rho, theta) (Projection.py:13-128, Orientation
 * already applied to coeffs).  rho, theta, out : float32 [n].              */
int oi_zernike(const float* rho, const float* theta, int n, const double* coeffs,
               float* out, void* stream);

/* Characteristic matrix (CalculateCharacteristicMatrix.py:3-16).
 * f, g : float32 [n]; out : c64 [6][n] = Mxx, Myx, Mxy, Myy, Mxz, Myz.      */
int oi_char_matrix(const float* f, const float* g, int n, double na,
                   double index_image, void* out, void* stream);

/* ------------------------------------------------------------- adjoint --
 * Mask gradient of the SOCS image (no reference counterpart: the reference
 * differentiates CalculateAerialImage_SOCS, CalculateAerialImage_SOCS.py:3-11,
 * only through torch autograd).  For real masks (plan with nchan = 1):
 *   grad_mask[b][y][x] = d/d mask[b][y][x]  sum_j grad_image[b][j] I[b][j],
 * I = oi_socs_image(mask[b]) with the same kernels.  Coherent fields are
 * recomputed per kernel on the plan's coarse grid; nothing is kept from the
 * forward call.  mask, grad_image, grad_mask : float32 [batch][ny][nx];
 * work : device scratch of oi_socs_adjoint_work_bytes(plan, batch, nkern)
 * bytes (contents need no initialisation).                                  */
size_t oi_socs_adjoint_work_bytes(const oi_plan* plan, int batch, int nkern);
int oi_socs_adjoint(oi_plan* plan, const float* mask, const float* grad_image, int batch,
                    const void* kern, int nkern, void* work, float* grad_mask, void* stream);

/* ------------------------------------------------------------------ ILT --
 * The elementwise passes of one normalised gradient step of the pixel-based
 * ILT objective sum_b ||I(m_b) - T_b||^2, m = sigmoid(theta), around
 * oi_socs_image and oi_socs_adjoint (no reference counterpart: the reference
 * has no mask-optimisation loop).  All arrays are float32 [batch][npix]
 * device arrays; work holds oi_ilt_work_floats(batch) floats (no
 * initialisation needed; one call at a time per work array).  Sums run over
 * a fixed split of each image, so results do not depend on the batch.
 *
 * oi_ilt_residual: resid2 = 2 (image - target) (resid2 may be image);
 *   loss (device float, may be null) = sum over the batch of (image - target)^2.
 * oi_ilt_update: g = grad_m m (1 - m); per image b,
 *   theta_b -= lr g_b / (sqrt(sum g_b^2 / npix) + 1e-30);
 *   m_next (may be m, or null) = sigmoid(theta) after the step.            */
size_t oi_ilt_work_floats(int batch);
/* oi_socs_residual: the SOCS image of real masks (as oi_socs_image with
 * kind OI_FIELD_REAL) fused with oi_ilt_residual: resid2 = 2 (I - target),
 * loss (device float, may be null) = sum over the batch of (I - target)^2,
 * computed in the pass that produces I (I itself is not written).  Plans
 * whose image sizes have no compile-time transforms (oi_geometry.path 0)
 * return OI_ERR_UNSUPPORTED before anything is launched: use oi_socs_image
 * and oi_ilt_residual there.  Calls on one plan share its scratch.         */
int oi_socs_residual(oi_plan* plan, const float* mask, int batch, const void* kern, int nkern,
                     const float* target, float* resid2, float* loss, void* stream);
/* oi_socs_residual followed by oi_socs_adjoint with grad_image = resid2:
 * the gradient step's forward and backward passes sharing the mask spectrum
 * (one spectrum pass instead of two) and, on two-pass plans (path 2), the
 * kernels' column-stage fields of the forward pass.  work as for
 * oi_socs_adjoint.  resid2 and loss equal oi_socs_residual's bit for bit;
 * grad_mask equals oi_socs_adjoint's to float32 rounding.
 * OI_ERR_UNSUPPORTED as oi_socs_residual (nothing launched).              */
int oi_socs_residual_adjoint(oi_plan* plan, const float* mask, int batch, const void* kern,
                             int nkern, const float* target, float* resid2, float* loss,
                             void* work, float* grad_mask, void* stream);
/* The same gradient step without the image.  The SOCS image is band-limited,
 * I(r) = sum_k S(k) e^{-2 pi i k.r / n} over the plan's band
 * [-qr, qr] x [-qc, qc] (n = ny nx), so with the targets' band spectrum
 * T~(k) = (1/n) sum_r T(r) e^{+2 pi i k.r / n} the cotangent's spectrum on
 * that band is 2 (S - T~) and, by Parseval,
 *   loss = n sum_band |S - T~|^2 + (sum_r T^2 - n sum_band |T~|^2);
 * the adjoint needs nothing else.  The image's row and column transforms,
 * the residual image and the cotangent's spectrum pass are not run.
 *
 * oi_ilt_target_band: tband (device, oi_ilt_target_band_bytes(plan, batch)
 *   bytes) = the band spectrum and the per-image constant of target[batch],
 *   once per set of targets; work is an oi_socs_adjoint work buffer (its
 *   first region is the transform scratch).
 * oi_socs_band_residual_adjoint: grad_mask and loss (device float, may be
 *   null) of the step for mask[batch] against the targets tband was made
 *   from (same batch); equal to oi_socs_residual_adjoint's to float32
 *   rounding (the image band enters directly instead of through the rounded
 *   image).  Plans with oi_geometry.direct set or without the residual pass
 *   return OI_ERR_UNSUPPORTED (oi_ilt_target_band_bytes gives 0) before
 *   anything is launched.  Calls on one plan share its scratch.             */
size_t oi_ilt_target_band_bytes(const oi_plan* plan, int batch);
int oi_ilt_target_band(oi_plan* plan, const float* target, int batch, void* work, void* tband,
                       void* stream);
int oi_socs_band_residual_adjoint(oi_plan* plan, const float* mask, int batch, const void* kern,
                                  int nkern, const void* tband, float* loss, void* work,
                                  float* grad_mask, void* stream);
/* oi_ilt_band_step: one whole normalised gradient step on the logits,
 *   m = sigmoid(theta) [batch][ny][nx] (device, updated in place): the loss
 *   (device float, may be null) and dL/dm as oi_socs_band_residual_adjoint,
 *   grad_theta = dL/dm m (1 - m) (device [batch][ny][nx], written; not
 *   theta itself), then per
 *   image theta -= lr grad_theta / (sqrt(sum grad_theta^2 / npix) + 1e-30),
 *   as oi_ilt_update.  m is never stored: the spectrum pass and the
 *   gradient's last pass take the sigmoid of theta as they load it, and the
 *   norm's partial sums come from the pass that writes grad_theta.  Equal to
 *   oi_socs_band_residual_adjoint followed by oi_ilt_update to float32
 *   rounding.  OI_ERR_UNSUPPORTED as oi_socs_band_residual_adjoint.        */
int oi_ilt_band_step(oi_plan* plan, float* theta, int batch, const void* kern, int nkern,
                     const void* tband, float lr, float* loss, void* work, float* grad_theta,
                     void* stream);
int oi_ilt_residual(const float* image, const float* target, float* resid2, float* loss,
                    int batch, size_t npix, float* work, void* stream);
int oi_ilt_update(float* theta, const float* grad_m, const float* m, float lr, int batch,
                  size_t npix, float* work, float* m_next, void* stream);

/* ----------------------------------------------------------------- RCWA --
 * Stages of the RCWA near-field producer (rcwa.py:4-872) that are not dense
 * linear algebra; the products, solves and inverses between them run in
 * rocBLAS/rocSOLVER (opticalimaging_amd/rcwa.py).  Fourier orders
 * (ox, oy), |ox| <= Ox, |oy| <= Oy, are flattened as
 * i = (ox + Ox)(2 Oy + 1) + (oy + Oy) (rcwa.py:340-347, 356-361);
 * N = (2 Ox + 1)(2 Oy + 1).  Arrays are c128 (interleaved float64 pairs),
 * row-major; kx, ky are the N normalised wave-vector components (float64).
 *
 * oi_rcwa_convmat: convolution matrix of a permittivity grid,
 *   E[i][j] = F[(ox_i - ox_j) mod n0][(oy_i - oy_j) mod n1], F = fft2(eps) / (n0 n1)
 *   already computed (c128 [n0][n1]); E : c128 [N][N].  Replaces the gather
 *   of rcwa.py:435-450.                                                      */
int oi_rcwa_convmat(const void* F, int n0, int n1, int ox, int oy, void* E, void* stream);

/* oi_rcwa_pq: P and Q of a layer (rcwa.py:470-497) from E (c128 [N][N]) and
 *   EiK = [E^-1 Kx | E^-1 Ky] (c128 [N][2N]); P, Q : c128 [2N][2N].           */
int oi_rcwa_pq(const void* E, const void* EiK, const double* kx, const double* ky, int N,
               void* P, void* Q, void* stream);

/* oi_rcwa_modes: mode matrix V (c128 [2N][2N]) and kz (c128 [N], may be
 *   null) of a homogeneous medium of permittivity eps (get_V,
 *   rcwa.py:865-872; kz on the branch with Im >= 0).                         */
int oi_rcwa_modes(double eps_re, double eps_im, const double* kx, const double* ky, int N,
                  void* V, void* kz, void* stream);

/* oi_rcwa_fields: field synthesis on an x-y grid (rcwa.py:306-315),
 *   out[x][y][c] = sum_o amp[c][o] exp(i omega (kxo[o / noy] xs[x] + kyo[o % noy] ys[y])),
 *   amp : c128 [3][nox noy] (Ex, Ey, Ez order amplitudes), kxo : float64
 *   [nox], kyo : float64 [noy], xs : float64 [nx], ys : float64 [ny];
 *   out : c128 [nx][ny][3] (the ScatterField layout).  nox + noy <= 256.     */
int oi_rcwa_fields(const void* amp, int nox, int noy, const double* kxo, const double* kyo,
                   double omega, const double* xs, int nx, const double* ys, int ny, void* out,
                   void* stream);

#ifdef __cplusplus
}
#endif

#endif /* OPTICALIMAGING_AMD_H */
