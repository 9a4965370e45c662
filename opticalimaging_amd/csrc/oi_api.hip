// C ABI: error state, twiddle tables, imaging plans and the imaging entry
// points (spectrum window, coherent image, SOCS image, full spectrum).
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "oi_adjoint.h"
#include "oi_fastpath.h"

namespace oi {

static thread_local std::string g_last_error;

void set_error(int code, const std::string& msg) {
  (void)code;
  g_last_error = msg;
}

int fail(int code, const std::string& msg) {
  set_error(code, msg);
  return code;
}

int make_twiddles(const int* sizes, int count, FftDesc* descs, TwiddlePool& pool) {
  std::vector<int> uniq;
  for (int i = 0; i < count; ++i) {
    if (!fft_factor(sizes[i], descs[i]))
      return fail(OI_ERR_UNSUPPORTED, "FFT length " + std::to_string(sizes[i]) +
                                          " is not a product of 2, 3 and 5");
    bool seen = false;
    for (int u : uniq) seen |= (u == sizes[i]);
    if (!seen) uniq.push_back(sizes[i]);
  }
  // each length n takes 2n entries (the layout fp_stage_radices describes)
  size_t total = 0;
  for (int u : uniq) total += 2 * (size_t)u;
  std::vector<float2> host(total);
  std::vector<size_t> off(uniq.size());
  size_t pos = 0;
  for (size_t k = 0; k < uniq.size(); ++k) {
    off[k] = pos;
    const int n = uniq[k];
    for (int m = 0; m < n; ++m) {
      // exp(-2 pi i m / n) in float64, rounded once.
      const double ang = 2.0 * M_PI * (double)m / (double)n;
      host[pos + m] = make_float2((float)std::cos(ang), (float)-std::sin(ang));
    }
    int r1 = 0, r2 = 0;
    const bool two = fp_stage_radices(n, &r1, &r2);
    for (int e = 0; e < n; ++e)
      host[pos + n + e] = two ? host[pos + (e / r1) * (e % r1)] : make_float2(0.f, 0.f);
    pos += 2 * (size_t)n;
  }
  float2* dev = nullptr;
  if (hipMalloc(&dev, total * sizeof(float2)) != hipSuccess)
    return fail(OI_ERR_NOMEM, "twiddle table allocation failed");
  if (hipMemcpy(dev, host.data(), total * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(dev);
    return fail(OI_ERR_RUNTIME, "twiddle table upload failed");
  }
  for (int i = 0; i < count; ++i)
    for (size_t k = 0; k < uniq.size(); ++k)
      if (uniq[k] == sizes[i]) descs[i].tw = dev + off[k];
  pool.dev = dev;
  pool.bytes = total * sizeof(float2);
  return OI_OK;
}

void free_twiddles(TwiddlePool& pool) {
  if (pool.dev) (void)hipFree(pool.dev);
  pool.dev = nullptr;
  pool.bytes = 0;
}

static bool smooth235(int n) {
  if (n < 1) return false;
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

// Smallest 2,3,5-smooth multiple of 8 that is >= need; n when that reaches n.
static int choose_coarse(int need, int n) {
  if (need >= n) return n;
  int L = need < 8 ? 8 : need;
  while (!(smooth235(L) && L % 8 == 0)) ++L;
  return L >= n ? n : L;
}

}  // namespace oi

using namespace oi;

// Stage boundaries recorded per call when timing is on: before spectrum rows,
// after rows, after columns, after the coherent sum, after the interpolation
// columns, after the interpolation rows.
constexpr int kStages = 5;

struct oi_plan {
  oi_geometry g;
  FftDesc fx, fy, flr, flc, frows;
  TwiddlePool pool;
  float2* T = nullptr;
  float2* spec = nullptr;
  float2* R = nullptr;
  float2* U = nullptr;
  float* Ic = nullptr;
  float2* Ihat = nullptr;
  float* part = nullptr;    // residual sums of the image-row pass, [B][ny]
  float2* Bbuf = nullptr;   // two-pass column fields of max_batch x max_kernels kernels
  size_t bytes = 0;
  std::vector<hipEvent_t> ev;  // slots * (kStages + 1)
  int slots = 0, next = 0, recorded = 0;
};

static void mark(oi_plan* p, int stage, hipStream_t st) {
  if (p->slots == 0) return;
  (void)hipEventRecord(p->ev[(size_t)p->next * (kStages + 1) + stage], st);
}

static void mark_done(oi_plan* p) {
  if (p->slots == 0) return;
  p->next = (p->next + 1) % p->slots;
  if (p->recorded < p->slots) ++p->recorded;
}

extern "C" const char* oi_last_error(void) { return g_last_error.c_str(); }

extern "C" int oi_version(void) { return 10000; }

static void plan_free(oi_plan* p) {
  if (!p) return;
  for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
  for (void* ptr : {(void*)p->T, (void*)p->spec, (void*)p->R, (void*)p->U, (void*)p->Ic,
                    (void*)p->Ihat, (void*)p->part, (void*)p->Bbuf})
    if (ptr) (void)hipFree(ptr);
  free_twiddles(p->pool);
  delete p;
}

// Validates a geometry and fills the library-chosen fields (coarse grid,
// path, rows, kernel groups, band half-widths); allocates nothing.
static int plan_geometry(const oi_geometry& in, oi_geometry& g) {
  g = in;
  if (g.ny <= 0 || g.nx <= 0 || g.wr <= 0 || g.wc <= 0 || g.nchan <= 0 || g.nchan > 3 ||
      g.max_batch <= 0)
    return fail(OI_ERR_INVALID, "plan_create: sizes must be positive (nchan <= 3)");
  if (g.wr > g.ny || g.wc > g.nx)
    return fail(OI_ERR_INVALID, "plan_create: spectrum window larger than the grid");
  if (g.max_kernels < 0) return fail(OI_ERR_INVALID, "plan_create: max_kernels must be >= 0");
  if (g.max_kernels == 0) g.max_kernels = 256;
  if (g.k0r < -(g.ny / 2) || g.k0r + g.wr - 1 > g.ny - 1 - g.ny / 2 || g.k0c < -(g.nx / 2) ||
      g.k0c + g.wc - 1 > g.nx - 1 - g.nx / 2)
    return fail(OI_ERR_INVALID, "plan_create: window outside the centred index range");
  if (!fft_size_ok(g.ny) || !fft_size_ok(g.nx))
    return fail(OI_ERR_UNSUPPORTED, "plan_create: grid sizes must be products of 2, 3 and 5");
  const int band_r = g.band_r > 0 ? g.band_r : g.wr;
  const int band_c = g.band_c > 0 ? g.band_c : g.wc;
  g.band_r = band_r;
  g.band_c = band_c;
  int lr = g.lr > 0 ? g.lr : choose_coarse(2 * band_r - 1, g.ny);
  int lc = g.lc > 0 ? g.lc : choose_coarse(2 * band_c - 1, g.nx);
  if (lr > g.ny || lc > g.nx) return fail(OI_ERR_INVALID, "plan_create: coarse grid exceeds the image grid");
  // A coarse grid equal to the image grid computes the image directly (this
  // also covers bands wider than the grid, where the image itself aliases).
  g.direct = (lr == g.ny || lc == g.nx) ? 1 : 0;
  if (!g.direct && (lr < 2 * band_r - 1 || lc < 2 * band_c - 1))
    return fail(OI_ERR_INVALID, "plan_create: coarse grid smaller than 2*band-1");
  if (g.direct) {
    lr = g.ny;
    lc = g.nx;
  }
  if (!fft_size_ok(lr) || !fft_size_ok(lc))
    return fail(OI_ERR_INVALID, "plan_create: coarse sizes must be products of 2, 3 and 5");
  g.lr = lr;
  g.lc = lc;
  g.qr = g.direct ? 0 : band_r - 1;
  g.qc = g.direct ? 0 : band_c - 1;
  // Compile-time-size path: both image sizes tabled and a tabled square coarse
  // grid L >= 2 band - 1 in both directions, smaller than the image.
  g.path = 0;
  if (!g.direct && in.lr <= 0 && in.lc <= 0 && fp_image_size_ok(g.ny) && fp_image_size_ok(g.nx)) {
    const int need = (2 * band_r - 1) > (2 * band_c - 1) ? 2 * band_r - 1 : 2 * band_c - 1;
    const int L = fp_coarse_choose(need);
    if (L > 0 && L < g.ny && L < g.nx && fp_coherent_rows_fit(L) &&
        fp_coherent_lds(L, g.wr, g.wc) <= kLdsLimit &&
        (fp_one_pass_ok(L) || !fp_two_pass_size_ok(L))) {
      g.path = 1;
      g.lr = g.lc = lr = lc = L;
    }
  }
  // Two-pass path: coarse grids beyond one workgroup's reach (L = 256, 512)
  // when the window folds onto them without collisions.
  if (g.path == 0 && !g.direct && in.lr <= 0 && in.lc <= 0) {
    const int need = (2 * band_r - 1) > (2 * band_c - 1) ? 2 * band_r - 1 : 2 * band_c - 1;
    // coarse grids past one workgroup's reach, and those the one-pass kernel
    // does not serve when the two-pass table has them (144)
    const int L1 = fp_coarse_choose(need);
    const int L = need > 144 ? fp_two_pass_choose(need)
                             : (L1 > 0 && !fp_one_pass_ok(L1) && fp_two_pass_size_ok(L1) ? L1 : 0);
    if (L > 0 && L < g.ny && L < g.nx && g.wr <= L && g.wc <= L &&
        fp_two_pass_lds(L, g.wc) <= kLdsLimit) {
      g.path = 2;
      g.lr = g.lc = lr = lc = L;
      if (g.max_groups <= 0) {
        // kernel groups of the row pass: about 1024 blocks (16 coarse rows
        // each) over the batch, at most 8 partial sums per image
        const int blocks = (L / 16) * g.max_batch;
        int gr = (1024 + blocks - 1) / blocks;
        g.max_groups = gr < 1 ? 1 : (gr > 8 ? 8 : gr);
      }
    }
  }
  // Coarse rows per workgroup: largest of 16, 8, 4, 2, 1 dividing lr whose
  // accumulator and shared-memory footprint fit.
  int rows = g.path ? lr : 0;
  for (int cand : {16, 8, 4, 2, 1}) {
    if (g.path) break;
    if (lr % cand) continue;
    if ((size_t)cand * lc > (size_t)kThreads * kAccMax) continue;
    if (coherent_lds_bytes(g.wc, cand, lc) > 120 * 1024) continue;
    rows = cand;
    break;
  }
  if (!rows) return fail(OI_ERR_UNSUPPORTED, "plan_create: no workgroup shape fits this geometry");
  g.rows = rows;  // the compile-time path keeps every coarse row in one block
  // Kernel groups: fixed per plan so an image's summation order never depends
  // on the batch it is computed in.  0 = enough groups for ~512 workgroups.
  if (g.max_groups <= 0) {
    // about one workgroup per CU for the compile-time-size kernels (one
    // large block per CU; more partial sums cost more than they hide), about
    // 512 workgroups for the generic kernels
    const int phases = lr / g.rows;
    const int target = g.path == 1 ? 256 : 512;
    // one image fills the CUs with up to 256 one-pass groups (its kernels
    // over the groups); the generic kernels stop at 64
    const int cap = g.path == 1 ? 256 : 64;
    int gr = (target + phases * g.max_batch - 1) / (phases * g.max_batch);
    g.max_groups = gr < 1 ? 1 : (gr > cap ? cap : gr);
  }
  return OI_OK;
}

// Device scratch of a plan for max_batch images: byte size of each buffer.
struct ScratchLayout {
  size_t T, spec, R, U, Ic, Ihat, part, Bbuf;
  size_t total() const { return T + spec + R + U + Ic + Ihat + part + Bbuf; }
};

static ScratchLayout scratch_layout(const oi_geometry& g, size_t B) {
  const size_t C = (size_t)g.nchan, G = (size_t)g.max_groups;
  const size_t lr = (size_t)g.lr, lc = (size_t)g.lc;
  ScratchLayout s;
  s.T = B * C * g.wc * g.ny * sizeof(float2);
  s.spec = B * C * g.wr * g.wc * sizeof(float2);
  s.R = g.direct ? 0 : B * G * lr * (size_t)(g.qc + 1) * sizeof(float2);
  s.U = g.direct ? 0 : B * (size_t)g.ny * (g.qc + 1) * sizeof(float2);
  s.Ic = (g.direct && G > 1) ? B * G * lr * lc * sizeof(float) : 0;
  s.Ihat = g.path ? B * (size_t)(g.qc + 1) * (2 * g.qr + 1) * sizeof(float2) : 0;
  s.part = g.path ? B * (size_t)g.ny * sizeof(float) : 0;
  s.Bbuf = g.path == 2 ? B * (size_t)g.max_kernels * C * g.wc * lr * sizeof(float2) : 0;
  return s;
}

extern "C" int oi_plan_create(oi_geometry* geo, oi_plan** out) {
  if (!geo || !out) return fail(OI_ERR_INVALID, "plan_create: null argument");
  *out = nullptr;
  oi_geometry g;
  int rc = plan_geometry(*geo, g);
  if (rc) return rc;
  const int lr = g.lr, lc = g.lc, rows = g.rows;
  oi_plan* p = new (std::nothrow) oi_plan();
  if (!p) return fail(OI_ERR_NOMEM, "plan_create: host allocation failed");
  p->g = g;
  int sizes[5] = {g.nx, g.ny, lr, lc, rows};
  FftDesc d[5];
  rc = make_twiddles(sizes, 5, d, p->pool);
  if (rc) {
    plan_free(p);
    return rc;
  }
  p->fx = d[0];
  p->fy = d[1];
  p->flr = d[2];
  p->flc = d[3];
  p->frows = d[4];
  const ScratchLayout sl = scratch_layout(g, (size_t)g.max_batch);
  struct Want {
    void** ptr;
    size_t bytes;
  } want[] = {
      {(void**)&p->T, sl.T},       {(void**)&p->spec, sl.spec}, {(void**)&p->R, sl.R},
      {(void**)&p->U, sl.U},       {(void**)&p->Ic, sl.Ic},     {(void**)&p->Ihat, sl.Ihat},
      {(void**)&p->part, sl.part}, {(void**)&p->Bbuf, sl.Bbuf},
  };
  p->bytes = p->pool.bytes;
  for (auto& w : want) {
    if (!w.bytes) continue;
    if (hipMalloc(w.ptr, w.bytes) != hipSuccess) {
      plan_free(p);
      return fail(OI_ERR_NOMEM, "plan_create: scratch allocation failed");
    }
    p->bytes += w.bytes;
  }
  *geo = g;
  *out = p;
  return OI_OK;
}

extern "C" int oi_plan_destroy(oi_plan* plan) {
  plan_free(plan);
  return OI_OK;
}

extern "C" size_t oi_plan_bytes(const oi_plan* plan) { return plan ? plan->bytes : 0; }

static int check_field(const oi_plan* p, const void* field, int kind, int cin, const int* chans,
                       int batch) {
  if (!p || !field) return fail(OI_ERR_INVALID, "null plan or field");
  if (batch < 1 || batch > p->g.max_batch) return fail(OI_ERR_INVALID, "batch outside [1, max_batch]");
  if (kind == OI_FIELD_REAL) {
    if (p->g.nchan != 1) return fail(OI_ERR_INVALID, "real masks carry one channel");
  } else if (kind == OI_FIELD_C64_CHLAST) {
    if (!chans || cin < 1) return fail(OI_ERR_INVALID, "complex fields need cin and chans");
    for (int i = 0; i < p->g.nchan; ++i)
      if (chans[i] < 0 || chans[i] >= cin) return fail(OI_ERR_INVALID, "channel index out of range");
  } else {
    return fail(OI_ERR_INVALID, "unknown field kind");
  }
  return OI_OK;
}

static void mark_cb(void* ctx, int stage);

struct MarkCtx {
  oi_plan* p;
  hipStream_t st;
};

static thread_local MarkCtx g_mark_ctx;

static void mark_cb(void* ctx, int stage) {
  MarkCtx* x = static_cast<MarkCtx*>(ctx);
  mark(x->p, stage, x->st);
}

static FpArgs fast_args(oi_plan* p, int batch, hipStream_t st) {
  const oi_geometry& g = p->g;
  FpArgs a{};
  a.nchan = g.nchan;
  a.batch = batch;
  a.ny = g.ny;
  a.nx = g.nx;
  a.k0r = g.k0r;
  a.wr = g.wr;
  a.k0c = g.k0c;
  a.wc = g.wc;
  a.L = g.lr;
  a.qr = g.qr;
  a.qc = g.qc;
  a.tw_x = p->fx.tw;
  a.tw_y = p->fy.tw;
  a.tw_l = p->flr.tw;
  a.T = p->T;
  a.spec = p->spec;
  a.R = p->R;
  a.Ihat = p->Ihat;
  a.U = p->U;
  g_mark_ctx = MarkCtx{p, st};
  a.mark = mark_cb;
  a.mark_ctx = &g_mark_ctx;
  return a;
}

static int spectrum_impl(oi_plan* p, const void* field, int kind, int cin, const int* chans,
                         int batch, void* spec, hipStream_t st);

// One pass over the field; each workgroup ORs its lanes' bits together and
// one lane per workgroup updates the mask (a few hundred atomics on one word
// instead of one per wave).
constexpr int kFlagBlocks = 256;

__global__ __launch_bounds__(256) void k_field_flags(const void* __restrict__ field, int cplx,
                                                     long long npix, int cin,
                                                     unsigned int* __restrict__ mask) {
  __shared__ unsigned int red[4];
  unsigned int bits = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += stride) {
    for (int c = 0; c < cin; ++c) {
      if (cplx) {
        const float2 v = static_cast<const float2*>(field)[p * cin + c];
        if (v.x != 0.f || v.y != 0.f) bits |= 1u << (2 * c);
        if (v.y != 0.f) bits |= 2u << (2 * c);
      } else if (static_cast<const float*>(field)[p * cin + c] != 0.f) {
        bits |= 1u << (2 * c);
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) bits |= __shfl_xor(bits, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int all = red[0] | red[1] | red[2] | red[3];
    if (all) atomicOr(mask, all);
  }
}

extern "C" int oi_field_flags(const void* field, int kind, long long npix, int cin, unsigned int* mask,
                              void* stream) {
  if (!field || !mask || npix < 0 || cin < 1 || cin > 16)
    return fail(OI_ERR_INVALID, "field_flags: bad arguments (cin must be in [1, 16])");
  if (kind != OI_FIELD_REAL && kind != OI_FIELD_C64_CHLAST)
    return fail(OI_ERR_INVALID, "field_flags: unknown field kind");
  if (npix == 0) return OI_OK;
  long long blocks = (npix + 255) / 256;
  if (blocks > kFlagBlocks) blocks = kFlagBlocks;
  hipLaunchKernelGGL(k_field_flags, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, field,
                     kind == OI_FIELD_C64_CHLAST ? 1 : 0, npix, cin, mask);
  OI_LAUNCH_CHECK("field_flags");
  return OI_OK;
}

extern "C" int oi_spectrum_window(oi_plan* p, const void* field, int kind, int cin,
                                  const int* chans, int batch, void* spec, void* stream) {
  int rc = check_field(p, field, kind, cin, chans, batch);
  if (rc) return rc;
  if (!spec) return fail(OI_ERR_INVALID, "null spec");
  return spectrum_impl(p, field, kind, cin, chans, batch, spec, (hipStream_t)stream);
}

static int spectrum_impl(oi_plan* p, const void* field, int kind, int cin, const int* chans,
                         int batch, void* spec, hipStream_t st) {
  int rc;
  const oi_geometry& g = p->g;
  mark(p, 0, st);
  if (g.path == 1 || (g.path == 2 && fp_image_size_ok(g.ny) && fp_image_size_ok(g.nx))) {
    FpArgs a = fast_args(p, batch, st);
    a.field = field;
    a.kind = kind;
    a.cin = cin;
    for (int i = 0; i < g.nchan; ++i) a.chans.c[i] = kind == OI_FIELD_REAL ? 0 : chans[i];
    a.spec = static_cast<float2*>(spec);
    rc = fp_spectrum(a, st);
    mark(p, 2, st);
    return rc;
  }
  rc = launch_spec_rows(field, kind, cin, chans, g.nchan, batch, g.ny, g.nx, p->fx, g.k0c, g.wc,
                        p->T, st);
  if (rc) return rc;
  mark(p, 1, st);
  rc = launch_spec_cols(p->T, batch * g.nchan, g.ny, g.nx, p->fy, g.k0r, g.wr, g.wc,
                        static_cast<float2*>(spec), st);
  mark(p, 2, st);
  return rc;
}

static int coherent_impl(oi_plan* p, const void* spec, int batch, const void* kern, int nkern,
                         float* out, const float* target, hipStream_t st, bool band_only = false);

extern "C" int oi_coherent_image(oi_plan* p, const void* spec, int batch, const void* kern,
                                 int nkern, float* out, void* stream) {
  return coherent_impl(p, spec, batch, kern, nkern, out, nullptr, (hipStream_t)stream);
}

// The residual form needs the compile-time-size image kernels (their row pass
// writes 2 (I - target) and its partial sums).
static bool residual_fused(const oi_plan* p) {
  const oi_geometry& g = p->g;
  return (g.path == 1 || g.path == 2) && fp_image_size_ok(g.ny) && fp_image_size_ok(g.nx) &&
         fp_urows_blocks(g.ny, g.nx, 1) <= g.ny;
}

// target non-null: out = 2 (I - target), partial sums of (I - target)^2 in p->part.
// band_only (plans with the residual pass): the band p->Ihat of the image, no image.
static int coherent_impl(oi_plan* p, const void* spec, int batch, const void* kern, int nkern,
                         float* out, const float* target, hipStream_t st, bool band_only) {
  if (!p || !spec || !kern || (!out && !band_only))
    return fail(OI_ERR_INVALID, "coherent_image: null argument");
  if ((target || band_only) && !residual_fused(p))
    return fail(OI_ERR_UNSUPPORTED, "socs_residual: plan without the compile-time-size image pass");
  if (batch < 1 || batch > p->g.max_batch) return fail(OI_ERR_INVALID, "batch outside [1, max_batch]");
  if (nkern < 1) return fail(OI_ERR_INVALID, "coherent_image: nkern must be >= 1");
  if (nkern > p->g.max_kernels)
    return fail(OI_ERR_INVALID, "coherent_image: nkern exceeds the plan's max_kernels (" +
                                    std::to_string(p->g.max_kernels) + ")");
  const oi_geometry& g = p->g;
  const int phases = g.lr / g.rows;
  int groups = g.max_groups < nkern ? g.max_groups : nkern;
  if (g.path == 2) {
    FpArgs a = fast_args(p, batch, st);
    a.out = out;
    a.target = target;
    a.part = p->part;
    FpCoh c{};
    c.spec = static_cast<const float2*>(spec);
    c.kern = static_cast<const float2*>(kern);
    c.nchan = g.nchan;
    c.nkern = nkern;
    c.k0r = g.k0r;
    c.wr = g.wr;
    c.k0c = g.k0c;
    c.wc = g.wc;
    const bool fast_io = fp_image_size_ok(g.ny) && fp_image_size_ok(g.nx);
    const int nkc = nkern * g.nchan;
    c.groups = fast_io ? (groups < nkc ? groups : nkc) : 1;
    c.qc = g.qc;
    c.tw_l = p->flr.tw;
    c.tw_m = p->flr.tw;
    c.R = p->R;
    int rc = fp_two_pass(a, c, p->Bbuf, fast_io ? 0 : 1, st);
    if (rc) return rc;
    if (band_only) return fp_band(a, c.groups, st);
    mark(p, 3, st);
    if (fast_io) {
      rc = fp_interp(a, c.groups, st);
    } else {
      InterpArgs ia{};
      ia.R = p->R;
      ia.groups = 1;
      ia.lr = g.lr;
      ia.lc = g.lc;
      ia.ny = g.ny;
      ia.nx = g.nx;
      ia.qr = g.qr;
      ia.qc = g.qc;
      ia.scale = 1.0f / ((float)g.lr * (float)g.lc);
      ia.flr = p->flr;
      ia.fny = p->fy;
      ia.fnx = p->fx;
      ia.U = p->U;
      ia.out = out;
      rc = launch_interp(ia, batch, st, mark_cb, &g_mark_ctx);
    }
    mark_done(p);
    return rc;
  }
  if (g.path) {
    FpArgs a = fast_args(p, batch, st);
    a.out = out;
    a.target = target;
    a.part = p->part;
    FpCoh c{};
    c.spec = static_cast<const float2*>(spec);
    c.kern = static_cast<const float2*>(kern);
    c.nchan = g.nchan;
    c.nkern = nkern;
    c.k0r = g.k0r;
    c.wr = g.wr;
    c.k0c = g.k0c;
    c.wc = g.wc;
    c.groups = groups;
    c.qc = g.qc;
    c.tw_l = p->flr.tw;
    c.tw_m = p->flr.tw;
    c.R = p->R;
    if (band_only) return fp_image_band(a, c, st);
    int rc = fp_image(a, c, st);
    mark_done(p);
    return rc;
  }
  CohArgs a{};
  a.spec = static_cast<const float2*>(spec);
  a.kern = static_cast<const float2*>(kern);
  a.nchan = g.nchan;
  a.nkern = nkern;
  a.k0r = g.k0r;
  a.wr = g.wr;
  a.k0c = g.k0c;
  a.wc = g.wc;
  a.lr = g.lr;
  a.lc = g.lc;
  a.rows = g.rows;
  a.phases = phases;
  a.groups = groups;
  a.qc = g.qc;
  a.direct = g.direct;
  a.frows = p->frows;
  a.flc = p->flc;
  a.tw_lr = p->flr.tw;
  a.R = p->R;
  a.Ic = g.direct ? (groups == 1 ? out : p->Ic) : nullptr;
  int rc = launch_coherent(a, batch, st);
  if (rc) return rc;
  mark(p, 3, st);
  if (g.direct) {
    if (groups > 1) rc = launch_group_sum(p->Ic, batch, groups, (size_t)g.ny * g.nx, out, st);
    mark(p, 4, st);
    mark(p, 5, st);
    mark_done(p);
    return rc;
  }
  InterpArgs ia{};
  ia.R = p->R;
  ia.groups = groups;
  ia.lr = g.lr;
  ia.lc = g.lc;
  ia.ny = g.ny;
  ia.nx = g.nx;
  ia.qr = g.qr;
  ia.qc = g.qc;
  ia.scale = 1.0f / ((float)g.lr * (float)g.lc);
  ia.flr = p->flr;
  ia.fny = p->fy;
  ia.fnx = p->fx;
  ia.U = p->U;
  ia.out = out;
  struct Ctx {
    oi_plan* p;
    hipStream_t st;
  } ctx{p, st};
  rc = launch_interp(ia, batch, st, [](void* c, int stage) {
    Ctx* x = static_cast<Ctx*>(c);
    mark(x->p, stage, x->st);
  }, &ctx);
  mark_done(p);
  return rc;
}

extern "C" int oi_plan_set_timing(oi_plan* p, int slots) {
  if (!p || slots < 0) return fail(OI_ERR_INVALID, "plan_set_timing: bad arguments");
  for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
  p->ev.clear();
  p->slots = 0;
  p->next = 0;
  p->recorded = 0;
  if (slots == 0) return OI_OK;
  p->ev.resize((size_t)slots * (kStages + 1));
  for (auto& e : p->ev) OI_HIP_CHECK(hipEventCreate(&e));
  p->slots = slots;
  return OI_OK;
}

extern "C" int oi_plan_read_timing(oi_plan* p, double* stage_ms, int* ncalls) {
  if (!p || !stage_ms || !ncalls) return fail(OI_ERR_INVALID, "plan_read_timing: bad arguments");
  for (int k = 0; k < kStages; ++k) stage_ms[k] = 0.0;
  *ncalls = p->recorded;
  for (int c = 0; c < p->recorded; ++c) {
    hipEvent_t* e = &p->ev[(size_t)c * (kStages + 1)];
    OI_HIP_CHECK(hipEventSynchronize(e[kStages]));
    for (int k = 0; k < kStages; ++k) {
      float ms = 0.f;
      OI_HIP_CHECK(hipEventElapsedTime(&ms, e[k], e[k + 1]));
      stage_ms[k] += ms;
    }
  }
  return OI_OK;
}

extern "C" int oi_socs_image(oi_plan* p, const void* field, int kind, int cin, const int* chans,
                             int batch, const void* kern, int nkern, float* out, void* stream) {
  int rc = oi_spectrum_window(p, field, kind, cin, chans, batch, p ? p->spec : nullptr, stream);
  if (rc) return rc;
  return oi_coherent_image(p, p->spec, batch, kern, nkern, out, stream);
}

extern "C" int oi_socs_residual(oi_plan* p, const float* mask, int batch, const void* kern,
                                int nkern, const float* target, float* resid2, float* loss,
                                void* stream) {
  if (!p || !mask || !kern || !target || !resid2)
    return fail(OI_ERR_INVALID, "socs_residual: null argument");
  if (p->g.nchan != 1) return fail(OI_ERR_INVALID, "socs_residual: real masks need a one-channel plan");
  if (!residual_fused(p))
    return fail(OI_ERR_UNSUPPORTED, "socs_residual: plan without the compile-time-size image pass");
  const int chan0 = 0;
  int rc = oi_spectrum_window(p, mask, OI_FIELD_REAL, 1, &chan0, batch, p->spec, stream);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  rc = coherent_impl(p, p->spec, batch, kern, nkern, resid2, target, st);
  if (rc || !loss) return rc;
  return ilt_sum(p->part, batch * fp_urows_blocks(p->g.ny, p->g.nx, batch), loss, st);
}

extern "C" int oi_spectrum_full(const void* field, int kind, int cin, const int* chans, int nchan,
                                int batch, int ny, int nx, float scale, void* work, void* hk,
                                void* stream) {
  if (!field || !work || !hk || batch < 1 || ny < 1 || nx < 1 || nchan < 1 || nchan > 3)
    return fail(OI_ERR_INVALID, "spectrum_full: bad arguments");
  if (kind == OI_FIELD_REAL) {
    if (nchan != 1) return fail(OI_ERR_INVALID, "spectrum_full: real fields carry one channel");
  } else if (kind == OI_FIELD_C64_CHLAST) {
    if (!chans) return fail(OI_ERR_INVALID, "spectrum_full: chans required");
    for (int i = 0; i < nchan; ++i)
      if (chans[i] < 0 || chans[i] >= cin) return fail(OI_ERR_INVALID, "spectrum_full: bad channel");
  } else {
    return fail(OI_ERR_INVALID, "spectrum_full: unknown field kind");
  }
  int sizes[2] = {nx, ny};
  FftDesc d[2];
  TwiddlePool pool;
  int rc = make_twiddles(sizes, 2, d, pool);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const float s = scale / ((float)ny * (float)nx);
  rc = launch_full_rows(field, kind, cin, chans, nchan, batch, ny, nx, d[0],
                        static_cast<float2*>(work), st);
  if (!rc)
    rc = launch_full_cols(static_cast<const float2*>(work), nchan, batch, ny, nx, d[1], s,
                          static_cast<float2*>(hk), st);
  // The tables must outlive the kernels that read them.
  if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = fail(OI_ERR_RUNTIME, "spectrum_full: sync failed");
  free_twiddles(pool);
  return rc;
}

// ------------------------------------------------------------- adjoint
static AdjGeom adj_geom(const oi_plan* p) {
  const oi_geometry& g = p->g;
  AdjGeom a{};
  a.ny = g.ny;
  a.nx = g.nx;
  a.k0r = g.k0r;
  a.wr = g.wr;
  a.k0c = g.k0c;
  a.wc = g.wc;
  a.lr = g.lr;
  a.lc = g.lc;
  a.qr = g.qr;
  a.qc = g.qc;
  a.direct = g.direct;
  a.fx = p->fx;
  a.fy = p->fy;
  a.flr = p->flr;
  a.flc = p->flc;
  return a;
}

extern "C" int oi_socs_residual_adjoint(oi_plan* p, const float* mask, int batch, const void* kern,
                                        int nkern, const float* target, float* resid2, float* loss,
                                        void* work, float* grad_mask, void* stream) {
  if (!p || !work || !grad_mask) return fail(OI_ERR_INVALID, "socs_residual_adjoint: null argument");
  int rc = oi_socs_residual(p, mask, batch, kern, nkern, target, resid2, loss, stream);
  if (rc) return rc;
  // the residual pass left this batch's mask spectrum in p->spec and, on the
  // two-pass path, its kernels' column stage in p->Bbuf
  const float2* bcols = p->g.path == 2 ? p->Bbuf : nullptr;
  return launch_adjoint(adj_geom(p), resid2, p->spec, batch, static_cast<const float2*>(kern), nkern,
                        work, grad_mask, (hipStream_t)stream, bcols);
}

// The band form (plans with the residual pass, not direct ones): the image's
// band and the targets' band spectrum give the cotangent spectrum and the loss
// without the image (oi_ilt.hip states the identities).
static bool band_residual_ok(const oi_plan* p) {
  // the loss's partial sums go to p->part (ny per image)
  return p->g.nchan == 1 && !p->g.direct && residual_fused(p) &&
         ilt_band_blocks(p->g.qr, p->g.qc) <= p->g.ny;
}

static size_t band_values(const oi_plan* p) {
  return (size_t)(2 * p->g.qr + 1) * (2 * p->g.qc + 1);
}

// the target band buffer: c64 [batch][2 qr + 1][2 qc + 1], then float [batch]
// (the constants) at the next 256-byte boundary
static size_t band_const_offset(const oi_plan* p, int batch) {
  return ((size_t)batch * band_values(p) * sizeof(float2) + 255) & ~(size_t)255;
}

extern "C" size_t oi_ilt_target_band_bytes(const oi_plan* p, int batch) {
  if (!p || batch < 1 || !band_residual_ok(p)) return 0;
  const size_t B = (size_t)batch;
  return band_const_offset(p, batch) + B * sizeof(float);
}

extern "C" int oi_ilt_target_band(oi_plan* p, const float* target, int batch, void* work,
                                  void* tband, void* stream) {
  if (!p || !target || !work || !tband) return fail(OI_ERR_INVALID, "ilt_target_band: null argument");
  if (!band_residual_ok(p))
    return fail(OI_ERR_UNSUPPORTED, "ilt_target_band: plan without the band form of the residual");
  if (batch < 1 || batch > p->g.max_batch) return fail(OI_ERR_INVALID, "batch outside [1, max_batch]");
  hipStream_t st = (hipStream_t)stream;
  const AdjGeom ag = adj_geom(p);
  float2* band = static_cast<float2*>(tband);
  int rc = adjoint_band_spectrum(ag, target, batch, work, band, st);
  if (rc) return rc;
  float* tconst = reinterpret_cast<float*>(static_cast<char*>(tband) + band_const_offset(p, batch));
  // the band spectrum is done with the work buffer's transform scratch (same
  // stream), which then holds the constants' partial sums
  return ilt_target_const(target, (size_t)p->g.ny * p->g.nx, band, band_values(p), batch, work,
                          tconst, st);
}

// The band form of one gradient step for real masks `field` (logits when
// `logits`): spectrum, coherent sum up to the image band, cotangent spectrum
// and loss from the band, adjoint.  With logits the adjoint's last pass
// writes the gradient with respect to the logits and its norm's partial sums
// (in p->part, which the loss sums have left by then: same stream).
static int band_step_impl(oi_plan* p, const float* field, bool logits, int batch, const void* kern,
                          int nkern, const void* tband, float* loss, void* work, float* grad,
                          hipStream_t st) {
  const int chan0 = 0;
  const int slots = p->slots;
  p->slots = 0;  // stage timing covers the imaging calls only
  int rc = spectrum_impl(p, field, logits ? kFieldLogits : OI_FIELD_REAL, 1, &chan0, batch,
                         p->spec, st);
  if (!rc) rc = coherent_impl(p, p->spec, batch, kern, nkern, nullptr, nullptr, st, true);
  p->slots = slots;
  if (rc) return rc;
  const AdjGeom ag = adj_geom(p);
  const float2* band = static_cast<const float2*>(tband);
  const float* tconst = reinterpret_cast<const float*>(static_cast<const char*>(tband) +
                                                       band_const_offset(p, batch));
  rc = ilt_band_cotangent(p->Ihat, p->g.qr, p->g.qc, band, tconst,
                          (float)p->g.ny * (float)p->g.nx, batch,
                          adjoint_cotangent_spectrum(ag, batch, work), p->part, st);
  if (rc) return rc;
  if (loss && (rc = ilt_sum(p->part, batch * ilt_band_blocks(p->g.qr, p->g.qc), loss, st)))
    return rc;
  const float2* bcols = p->g.path == 2 ? p->Bbuf : nullptr;
  return launch_adjoint(ag, nullptr, p->spec, batch, static_cast<const float2*>(kern), nkern, work,
                        grad, st, bcols, logits ? field : nullptr, logits ? p->part : nullptr);
}

static int band_step_check(const oi_plan* p, int batch, int nkern, const char* who) {
  if (!band_residual_ok(p))
    return fail(OI_ERR_UNSUPPORTED, std::string(who) + ": plan without the band form of the residual");
  if (batch < 1 || batch > p->g.max_batch) return fail(OI_ERR_INVALID, "batch outside [1, max_batch]");
  if (nkern < 1 || nkern > p->g.max_kernels)
    return fail(OI_ERR_INVALID, std::string(who) + ": nkern outside [1, max_kernels]");
  return OI_OK;
}

extern "C" int oi_socs_band_residual_adjoint(oi_plan* p, const float* mask, int batch,
                                             const void* kern, int nkern, const void* tband,
                                             float* loss, void* work, float* grad_mask,
                                             void* stream) {
  if (!p || !mask || !kern || !tband || !work || !grad_mask)
    return fail(OI_ERR_INVALID, "socs_band_residual_adjoint: null argument");
  int rc = band_step_check(p, batch, nkern, "socs_band_residual_adjoint");
  if (rc) return rc;
  return band_step_impl(p, mask, false, batch, kern, nkern, tband, loss, work, grad_mask,
                        (hipStream_t)stream);
}

extern "C" int oi_ilt_band_step(oi_plan* p, float* theta, int batch, const void* kern, int nkern,
                                const void* tband, float lr, float* loss, void* work,
                                float* grad_theta, void* stream) {
  if (!p || !theta || !kern || !tband || !work || !grad_theta)
    return fail(OI_ERR_INVALID, "ilt_band_step: null argument");
  if (grad_theta == theta) return fail(OI_ERR_INVALID, "ilt_band_step: grad_theta must not be theta");
  int rc = band_step_check(p, batch, nkern, "ilt_band_step");
  if (rc) return rc;
  if (!adjoint_chain_ok(adj_geom(p)))
    return fail(OI_ERR_UNSUPPORTED, "ilt_band_step: plan without the band form of the residual");
  hipStream_t st = (hipStream_t)stream;
  rc = band_step_impl(p, theta, true, batch, kern, nkern, tband, loss, work, grad_theta, st);
  if (rc) return rc;
  return ilt_update_theta(theta, grad_theta, lr, batch, (size_t)p->g.ny * p->g.nx, p->part,
                          fp_urows_blocks(p->g.ny, p->g.nx, batch), st);
}

extern "C" size_t oi_socs_adjoint_work_bytes(const oi_plan* p, int batch, int nkern) {
  if (!p || batch < 1 || nkern < 1) return 0;
  return adjoint_work_bytes(adj_geom(p), batch, nkern);
}

extern "C" int oi_socs_adjoint(oi_plan* p, const float* mask, const float* grad_image, int batch,
                               const void* kern, int nkern, void* work, float* grad_mask,
                               void* stream) {
  if (!p || !mask || !grad_image || !kern || !work || !grad_mask)
    return fail(OI_ERR_INVALID, "socs_adjoint: null argument");
  if (p->g.nchan != 1) return fail(OI_ERR_INVALID, "socs_adjoint: the plan must carry one channel (real masks)");
  if (batch < 1 || batch > p->g.max_batch) return fail(OI_ERR_INVALID, "batch outside [1, max_batch]");
  if (nkern < 1) return fail(OI_ERR_INVALID, "socs_adjoint: nkern must be >= 1");
  if (nkern > p->g.max_kernels)
    return fail(OI_ERR_INVALID, "socs_adjoint: nkern exceeds the plan's max_kernels");
  hipStream_t st = (hipStream_t)stream;
  const int chan0 = 0;
  const int slots = p->slots;
  p->slots = 0;  // stage timing covers the forward imaging calls only
  int rc = spectrum_impl(p, mask, OI_FIELD_REAL, 1, &chan0, batch, p->spec, st);
  p->slots = slots;
  if (rc) return rc;
  return launch_adjoint(adj_geom(p), grad_image, p->spec, batch, static_cast<const float2*>(kern),
                        nkern, work, grad_mask, st);
}

// --------------------------------------------- no-allocation entry point
// oi_socs_image_noalloc runs the same launches as oi_socs_image with the
// plan's twiddle tables and scratch carved from `scratch`, a fixed device
// array that every GPU holds from module load until the process ends.  Host
// code never passes `scratch` itself: its device address on the current GPU
// comes from hipGetSymbolAddress, taken once per device on the first call.
// Calls on one GPU share the array, so calls that overlap on one GPU race.
// No buffer in it needs to start at zero: every launch writes what it reads.
namespace {
constexpr size_t kNoAllocScratch = (size_t)OI_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
constexpr size_t kAlign = 256;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

// The 2n-entry tables of make_twiddles computed on the device: exp(-2 pi i
// m / n) in float64, rounded once, for m < n, then the stage-2 entries
// n + j r1 + k = the value at m = j k (r1 = 0: no two-stage transform, zeros).
__global__ void k_twiddle_table(float2* __restrict__ tw, int n, int r1) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * n) return;
  int m = e;
  if (e >= n) {
    if (r1 == 0) {
      tw[e] = make_float2(0.f, 0.f);
      return;
    }
    m = ((e - n) / r1) * ((e - n) % r1);
  }
  const double ang = 2.0 * M_PI * (double)m / (double)n;
  tw[e] = make_float2((float)cos(ang), (float)-sin(ang));
}

// Device address of `scratch` on the current GPU (looked up once per device).
static int scratch_of_current_device(unsigned char** out) {
  int dev = -1;
  OI_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(OI_ERR_UNSUPPORTED, "socs_image_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(OI_ERR_RUNTIME, std::string("socs_image_noalloc: hipGetSymbolAddress failed: ") +
                                      hipGetErrorString(e));
    g_scratch_dev[dev] = ptr;
  }
  *out = static_cast<unsigned char*>(g_scratch_dev[dev]);
  return OI_OK;
}

static size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

static size_t aligned_total(const ScratchLayout& s) {
  return align_up(s.T) + align_up(s.spec) + align_up(s.R) + align_up(s.U) + align_up(s.Ic) +
         align_up(s.Ihat) + align_up(s.part) + align_up(s.Bbuf);
}

extern "C" int oi_socs_image_noalloc(const oi_geometry* geo, const void* field, int kind, int cin,
                                     const int* chans, int batch, const void* kern, int nkern,
                                     float* out, void* stream) {
  if (!geo || !field || !kern || !out) return fail(OI_ERR_INVALID, "socs_image_noalloc: null argument");
  if (batch < 1) return fail(OI_ERR_INVALID, "socs_image_noalloc: batch must be >= 1");
  if (nkern < 1) return fail(OI_ERR_INVALID, "socs_image_noalloc: nkern must be >= 1");
  // Kernel groups are chosen as for a plan of `batch` images and `nkern`
  // kernels, so each image is summed exactly as oi_socs_image sums it there.
  oi_geometry in = *geo;
  in.max_batch = batch;
  in.max_kernels = nkern;
  oi_geometry g;
  int rc = plan_geometry(in, g);
  if (rc) return rc;
  // twiddle tables first, one per distinct length
  const int sizes[5] = {g.nx, g.ny, g.lr, g.lc, g.rows};
  FftDesc d[5];
  size_t tw_off[5];
  size_t used = 0;
  for (int i = 0; i < 5; ++i) {
    if (!fft_factor(sizes[i], d[i]))
      return fail(OI_ERR_UNSUPPORTED, "socs_image_noalloc: FFT length " + std::to_string(sizes[i]) +
                                          " is not a product of 2, 3 and 5");
    tw_off[i] = used;
    for (int j = 0; j < i; ++j)
      if (sizes[j] == sizes[i]) tw_off[i] = tw_off[j];
    if (tw_off[i] == used) used += align_up(2 * (size_t)sizes[i] * sizeof(float2));
  }
  const size_t per_image = aligned_total(scratch_layout(g, 1));
  if (used + per_image > kNoAllocScratch)
    return fail(OI_ERR_INVALID, "socs_image_noalloc: one image needs " +
                                    std::to_string(used + per_image) + " bytes of scratch, beyond the " +
                                    std::to_string(kNoAllocScratch) + " bytes this entry point holds");
  // images per pass: align_up(c x) <= c align_up(x), so c images fit in c
  // one-image layouts
  const size_t fit = (kNoAllocScratch - used) / per_image;
  const int chunk = fit < (size_t)batch ? (int)fit : batch;
  unsigned char* base = nullptr;
  rc = scratch_of_current_device(&base);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  oi_plan pl;  // on the stack: pointers into `scratch`, nothing owned
  pl.g = g;
  pl.g.max_batch = chunk;
  for (int i = 0; i < 5; ++i) {
    d[i].tw = reinterpret_cast<const float2*>(base + tw_off[i]);
    bool first = true;
    for (int j = 0; j < i; ++j) first &= (sizes[j] != sizes[i]);
    if (first) {
      int r1 = 0, r2 = 0;
      if (!fp_stage_radices(sizes[i], &r1, &r2)) r1 = 0;
      hipLaunchKernelGGL(k_twiddle_table, dim3((2 * sizes[i] + 255) / 256), dim3(256), 0, st,
                         reinterpret_cast<float2*>(base + tw_off[i]), sizes[i], r1);
      OI_LAUNCH_CHECK("twiddle_table");
    }
  }
  pl.fx = d[0];
  pl.fy = d[1];
  pl.flr = d[2];
  pl.flc = d[3];
  pl.frows = d[4];
  const ScratchLayout sl = scratch_layout(pl.g, (size_t)chunk);
  unsigned char* cur = base + used;
  auto carve = [&cur](size_t bytes) -> unsigned char* {
    unsigned char* p = bytes ? cur : nullptr;
    cur += align_up(bytes);
    return p;
  };
  pl.T = reinterpret_cast<float2*>(carve(sl.T));
  pl.spec = reinterpret_cast<float2*>(carve(sl.spec));
  pl.R = reinterpret_cast<float2*>(carve(sl.R));
  pl.U = reinterpret_cast<float2*>(carve(sl.U));
  pl.Ic = reinterpret_cast<float*>(carve(sl.Ic));
  pl.Ihat = reinterpret_cast<float2*>(carve(sl.Ihat));
  pl.part = reinterpret_cast<float*>(carve(sl.part));
  pl.Bbuf = reinterpret_cast<float2*>(carve(sl.Bbuf));
  if ((size_t)(cur - base) > kNoAllocScratch)
    return fail(OI_ERR_RUNTIME, "socs_image_noalloc: scratch layout overflow");
  const size_t pix = (size_t)g.ny * g.nx;
  const size_t in_stride = kind == OI_FIELD_REAL ? pix * sizeof(float)
                                                 : pix * (size_t)(cin > 0 ? cin : 1) * sizeof(float2);
  for (int b0 = 0; b0 < batch; b0 += chunk) {
    const int nb = batch - b0 < chunk ? batch - b0 : chunk;
    const void* fb = static_cast<const unsigned char*>(field) + (size_t)b0 * in_stride;
    rc = check_field(&pl, fb, kind, cin, chans, nb);
    if (rc) return rc;
    rc = spectrum_impl(&pl, fb, kind, cin, chans, nb, pl.spec, st);
    if (rc) return rc;
    rc = coherent_impl(&pl, pl.spec, nb, kern, nkern, out + (size_t)b0 * pix, nullptr, st);
    if (rc) return rc;
  }
  return OI_OK;
}
