// C ABI of libcnpc_hip.so: context (stream, rocFFT plan cache, scratch) and
// the thin wrappers that validate arguments and launch the kernels of
// paint.hip, spectral.hip and contract.hip.  See include/cnpc.h.
#include <rocfft/rocfft.h>

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/cnpc.h"
#include "cnpc_common.h"

namespace cnpc {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

// Static scratch of the allocation-free entry point
// (cnpc_contract_zgemm_noalloc): the FP64 partial tiles and the line
// twiddles.  Each device holds its own copy until the process ends; host
// code uses its device address from hipGetSymbolAddress (g_scratch_ptr),
// never the array's name.
static __device__ __align__(256) unsigned char scratch[CNPC_NOALLOC_SCRATCH_BYTES];
static void* g_scratch_ptr[CNPC_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

__global__ void zero_kernel(uint4* p, int64_t n16, unsigned char* tail, int ntail) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0, 0, 0, 0);
  if (blockIdx.x == 0 && (int)threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

__global__ void absmax_kernel(const float* __restrict__ x, int64_t n, double* __restrict__ partial) {
  float m = 0.f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(x[i]));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

}  // namespace cnpc

using namespace cnpc;

struct FftPlan {
  rocfft_plan plan = nullptr;
  rocfft_execution_info info = nullptr;
  void* work = nullptr;
  size_t work_bytes = 0;
};

struct cnpc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // key: (n, batch, prec, direction, dims)
  std::map<std::tuple<int, int, int, int, int>, FftPlan> plans;
  void* scratch = nullptr;  // reduction partials
  size_t scratch_bytes = 0;
  void* synth = nullptr;    // shell-synthesis cube, bins and twiddles
  size_t synth_bytes = 0;
  void* pwork = nullptr;    // tiled paint: keys, tile counts, sorted particles
  size_t pwork_bytes = 0;
  void* tw400 = nullptr;    // line twiddles of the fixed 400^3 c2r (made once)
  // recorded graphs hold the addresses of the work areas and plans: while
  // holds > 0 a work area that grows keeps its old allocation here (freed
  // when the last hold is released) and the plan cache is not cleared
  int holds = 0;
  std::vector<void*> retired;
};

#define CHECK_HIP(expr)                                                                        \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) {                                                                    \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                            \
      return 2;                                                                                \
    }                                                                                          \
  } while (0)

#define REQUIRE(cond, msg)     \
  do {                         \
    if (!(cond)) {             \
      set_error(msg);          \
      return 1;                \
    }                          \
  } while (0)

// A work area cannot grow while the context's stream is being captured into
// a graph (the allocation and the synchronisation before the free are not
// stream work); the same calls made once before the capture size it.
static int check_not_capturing(cnpc_ctx* ctx, const char* what) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(ctx->stream, &cs) != hipSuccess) {
    (void)hipGetLastError();  // a stream the query does not cover is not recording
    cs = hipStreamCaptureStatusNone;
  }
  REQUIRE(cs == hipStreamCaptureStatusNone,
          std::string(what) + " while the stream is being captured into a graph; make the same calls once before "
          "the capture");
  return 0;
}

// Frees a work area that is being replaced, or keeps it while recorded
// graphs hold the context's addresses.
static int release_area(cnpc_ctx* ctx, void*& p) {
  if (!p) return 0;
  if (ctx->holds > 0) {
    ctx->retired.push_back(p);
  } else {
    CHECK_HIP(hipStreamSynchronize(ctx->stream));
    CHECK_HIP(hipFree(p));
  }
  p = nullptr;
  return 0;
}

static int ensure_scratch(cnpc_ctx* ctx, size_t bytes) {
  if (ctx->scratch_bytes >= bytes) return 0;
  if (int rc = check_not_capturing(ctx, "the context's scratch area must grow")) return rc;
  if (int rc = release_area(ctx, ctx->scratch)) return rc;
  ctx->scratch_bytes = 0;
  CHECK_HIP(hipMalloc(&ctx->scratch, bytes));
  ctx->scratch_bytes = bytes;
  return 0;
}

static int ensure_synth(cnpc_ctx* ctx, size_t bytes) {
  if (ctx->synth_bytes >= bytes) return 0;
  if (int rc = check_not_capturing(ctx, "the context's synthesis area must grow")) return rc;
  if (int rc = release_area(ctx, ctx->synth)) return rc;
  ctx->synth_bytes = 0;
  CHECK_HIP(hipMalloc(&ctx->synth, bytes));
  ctx->synth_bytes = bytes;
  return 0;
}

static int ensure_pwork(cnpc_ctx* ctx, size_t bytes) {
  if (ctx->pwork_bytes >= bytes) return 0;
  if (int rc = check_not_capturing(ctx, "the context's paint area must grow")) return rc;
  if (int rc = release_area(ctx, ctx->pwork)) return rc;
  ctx->pwork_bytes = 0;
  CHECK_HIP(hipMalloc(&ctx->pwork, bytes));
  ctx->pwork_bytes = bytes;
  return 0;
}

// dims = 3: N^3 transforms; dims = 2: N^2 planes (the (y, z) pass of the
// shell synthesis).  Default packed layouts.
static int get_plan(cnpc_ctx* ctx, int n, int batch, int prec, int dir, FftPlan** out, int dims = 3) {
  auto key = std::make_tuple(n, batch, prec, dir, dims);
  auto it = ctx->plans.find(key);
  if (it != ctx->plans.end()) {
    *out = &it->second;
    return 0;
  }
  if (int rc = check_not_capturing(ctx, "a new FFT plan must be created")) return rc;
  static std::once_flag once;
  std::call_once(once, [] { rocfft_setup(); });
  FftPlan p;
  size_t lengths[3] = {(size_t)n, (size_t)n, (size_t)n};
  // dir: 0 real forward, 1 real inverse, 2 complex forward, 3 complex inverse
  const rocfft_transform_type tt = dir == 0   ? rocfft_transform_type_real_forward
                                   : dir == 1 ? rocfft_transform_type_real_inverse
                                   : dir == 2 ? rocfft_transform_type_complex_forward
                                              : rocfft_transform_type_complex_inverse;
  rocfft_status s = rocfft_plan_create(
      &p.plan, rocfft_placement_notinplace, tt,
      prec == 0 ? rocfft_precision_double : rocfft_precision_single, (size_t)dims, lengths, (size_t)batch, nullptr);
  REQUIRE(s == rocfft_status_success, "rocfft_plan_create failed (status " + std::to_string((int)s) + ")");
  s = rocfft_plan_get_work_buffer_size(p.plan, &p.work_bytes);
  REQUIRE(s == rocfft_status_success, "rocfft_plan_get_work_buffer_size failed");
  s = rocfft_execution_info_create(&p.info);
  REQUIRE(s == rocfft_status_success, "rocfft_execution_info_create failed");
  if (p.work_bytes) {
    CHECK_HIP(hipMalloc(&p.work, p.work_bytes));
    s = rocfft_execution_info_set_work_buffer(p.info, p.work, p.work_bytes);
    REQUIRE(s == rocfft_status_success, "rocfft_execution_info_set_work_buffer failed");
  }
  auto res = ctx->plans.emplace(key, p);
  *out = &res.first->second;
  return 0;
}

static void fill_edges(BinEdges& E, const double* edges, int nbins) {
  std::memset(&E, 0, sizeof(E));
  E.nbins = nbins;
  for (int i = 0; i <= nbins; ++i) E.e[i] = edges[i];
  E.e0 = edges[0];
  const double de = (edges[nbins] - edges[0]) / nbins;
  E.inv_de = de > 0 ? 1.0 / de : 0.0;
}

static void to_spec(PolySpec& d, const cnpc_poly& s) {
  static_assert(sizeof(PolySpec) == sizeof(cnpc_poly), "PolySpec layout");
  std::memcpy(&d, &s, sizeof(PolySpec));
}

extern "C" {

const char* cnpc_last_error(void) { return g_err.c_str(); }
int cnpc_version(void) { return 1; }

int cnpc_ctx_create(cnpc_ctx** out, int device) {
  REQUIRE(out != nullptr, "cnpc_ctx_create: null out pointer");
  int ndev = 0;
  CHECK_HIP(hipGetDeviceCount(&ndev));
  REQUIRE(device >= 0 && device < ndev, "cnpc_ctx_create: invalid device " + std::to_string(device));
  CHECK_HIP(hipSetDevice(device));
  REQUIRE(device < CNPC_MAX_DEVICES, "cnpc_ctx_create: device index beyond the library's limit");
  {  // once per device, with the device current: the static scratch's address
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
      void* p = nullptr;
      CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      REQUIRE(p != nullptr, "cnpc_ctx_create: no device address for the static scratch");
      g_scratch_ptr[device] = p;
    }
  }
  cnpc_ctx* c = new cnpc_ctx();
  c->device = device;
  *out = c;
  return 0;
}

int cnpc_ctx_clear_plans(cnpc_ctx* ctx) {
  REQUIRE(ctx, "null context");
  REQUIRE(ctx->holds == 0, "cnpc_ctx_clear_plans: recorded graphs still hold the context's plans");
  CHECK_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& kv : ctx->plans) {
    if (kv.second.plan) rocfft_plan_destroy(kv.second.plan);
    if (kv.second.info) rocfft_execution_info_destroy(kv.second.info);
    if (kv.second.work) (void)hipFree(kv.second.work);
  }
  ctx->plans.clear();
  return 0;
}

int cnpc_ctx_destroy(cnpc_ctx* ctx) {
  if (!ctx) return 0;
  ctx->holds = 0;
  int rc = cnpc_ctx_clear_plans(ctx);
  for (void* p : ctx->retired) (void)hipFree(p);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  if (ctx->pwork) (void)hipFree(ctx->pwork);
  if (ctx->synth) (void)hipFree(ctx->synth);
  if (ctx->tw400) (void)hipFree(ctx->tw400);
  delete ctx;
  return rc;
}

int cnpc_ctx_set_stream(cnpc_ctx* ctx, void* stream) {
  REQUIRE(ctx, "null context");
  ctx->stream = (hipStream_t)stream;
  return 0;
}

int cnpc_ctx_hold(cnpc_ctx* ctx, int delta) {
  REQUIRE(ctx && (delta == 1 || delta == -1), "cnpc_ctx_hold: delta must be +1 or -1");
  REQUIRE(ctx->holds + delta >= 0, "cnpc_ctx_hold: no hold to release");
  ctx->holds += delta;
  if (ctx->holds == 0 && !ctx->retired.empty()) {
    CHECK_HIP(hipStreamSynchronize(ctx->stream));
    for (void* p : ctx->retired) CHECK_HIP(hipFree(p));
    ctx->retired.clear();
  }
  return 0;
}

int cnpc_ctx_reserve(cnpc_ctx* ctx, int64_t bytes) {
  REQUIRE(ctx && bytes >= 0, "cnpc_ctx_reserve: bad arguments");
  return ensure_scratch(ctx, (size_t)bytes);
}

int cnpc_sync(cnpc_ctx* ctx) {
  REQUIRE(ctx, "null context");
  CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int cnpc_zero(cnpc_ctx* ctx, void* ptr, int64_t nbytes) {
  REQUIRE(ctx && (ptr || nbytes == 0) && nbytes >= 0, "cnpc_zero: bad arguments");
  if (nbytes == 0) return 0;
  REQUIRE(((uintptr_t)ptr & 15) == 0, "cnpc_zero: pointer must be 16-byte aligned");
  const int64_t n16 = nbytes / 16;
  const int ntail = (int)(nbytes - n16 * 16);
  hipLaunchKernelGGL(zero_kernel, dim3(grid_for(n16 > 0 ? n16 : 1, 256, 1 << 14)), dim3(256), 0, ctx->stream,
                     (uint4*)ptr, n16, (unsigned char*)ptr + n16 * 16, ntail);
  CHECK_HIP(hipGetLastError());
  return 0;
}

int cnpc_paint(cnpc_ctx* ctx, const double* pos, const double* w, int64_t np, int n, double boxsize, int window,
               int interlaced, const double* center, int prec, void* mesh0, void* mesh1) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 2 && n % 2 == 0 && n <= 4096, "cnpc_paint: nmesh must be even and in [2, 4096]");
  REQUIRE(window >= 1 && window <= 4, "cnpc_paint: window must be 1..4 (ngp, cic, tsc, pcs)");
  REQUIRE(prec == 0 || prec == 1, "cnpc_paint: prec must be 0 (fp64) or 1 (fp32)");
  REQUIRE(boxsize > 0, "cnpc_paint: boxsize must be positive");
  REQUIRE(np >= 0 && (np == 0 || (pos && w)), "cnpc_paint: null particle arrays");
  REQUIRE(mesh0 && (!interlaced || mesh1), "cnpc_paint: null mesh");
  if (paint_tiled_ok(n) && np < ((int64_t)1 << 31)) {  // sorted, LDS-tiled, writes every cell
    if (int rc = ensure_pwork(ctx, paint_tiled_work_bytes(np, n, 1))) return rc;
    CHECK_HIP(launch_paint_tiled(pos, w, nullptr, 1, np, np, n, boxsize, window, center, interlaced ? 2 : 1, prec,
                                 mesh0, mesh1, ctx->pwork, ctx->stream));
    return 0;
  }
  const int64_t bytes = (int64_t)n * n * n * (prec == 0 ? 8 : 4);
  if (int rc = cnpc_zero(ctx, mesh0, bytes)) return rc;
  if (interlaced)
    if (int rc = cnpc_zero(ctx, mesh1, bytes)) return rc;
  CHECK_HIP(launch_paint(pos, w, np, n, boxsize, window, center, interlaced ? 2 : 1, prec, mesh0, mesh1, ctx->stream));
  return 0;
}

int cnpc_paint_batch(cnpc_ctx* ctx, const double* pos, const double* w, const int64_t* offsets, int nbatch,
                     int64_t total_count, int64_t max_count, int n, double boxsize, int window, int interlaced, const double* center, int prec,
                     void* mesh0, void* mesh1) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 2 && n % 2 == 0 && n <= 4096, "cnpc_paint_batch: nmesh must be even and in [2, 4096]");
  REQUIRE(window >= 1 && window <= 4, "cnpc_paint_batch: window must be 1..4 (ngp, cic, tsc, pcs)");
  REQUIRE(prec == 0 || prec == 1, "cnpc_paint_batch: prec must be 0 (fp64) or 1 (fp32)");
  REQUIRE(boxsize > 0, "cnpc_paint_batch: boxsize must be positive");
  REQUIRE(nbatch >= 1 && nbatch <= 65535 && max_count >= 0, "cnpc_paint_batch: bad batch size");
  REQUIRE(offsets && mesh0 && (!interlaced || mesh1), "cnpc_paint_batch: null offsets or meshes");
  REQUIRE(max_count == 0 || (pos && w), "cnpc_paint_batch: null particle arrays");
  REQUIRE(total_count >= max_count, "cnpc_paint_batch: total_count below max_count");
  if (paint_tiled_ok(n) && total_count >= 0 && total_count < ((int64_t)1 << 31) &&
      (int64_t)nbatch * (n / 8) * (n / 8) * (n / 8) * 8 < ((int64_t)1 << 31)) {
    // sorted, LDS-tiled, every cell of every mesh written once
    if (int rc = ensure_pwork(ctx, paint_tiled_work_bytes(total_count, n, nbatch))) return rc;
    CHECK_HIP(launch_paint_tiled(pos, w, offsets, nbatch, total_count, max_count, n, boxsize, window, center,
                                 interlaced ? 2 : 1, prec, mesh0, interlaced ? mesh1 : nullptr, ctx->pwork,
                                 ctx->stream));
    return 0;
  }
  const int64_t bytes = (int64_t)nbatch * n * n * n * (prec == 0 ? 8 : 4);
  if (int rc = cnpc_zero(ctx, mesh0, bytes)) return rc;
  if (interlaced)
    if (int rc = cnpc_zero(ctx, mesh1, bytes)) return rc;
  CHECK_HIP(launch_paint_batch(pos, w, offsets, nbatch, max_count, n, boxsize, window, center, interlaced ? 2 : 1,
                               prec, mesh0, interlaced ? mesh1 : nullptr, ctx->stream));
  return 0;
}

static int run_fft(cnpc_ctx* ctx, int n, int batch, int prec, int dir, void* in, void* out, int dims = 3) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 2 && n % 2 == 0 && batch >= 1, "fft: bad size");
  REQUIRE(prec == 0 || prec == 1, "fft: prec must be 0 or 1");
  REQUIRE(in && out && in != out, "fft: null or aliased buffers (out-of-place only)");
  if (dir == 1 && dims == 3 && n == 400 && prec == 1) {
    // FP32 400^3 c2r: three passes of the fixed 20 x 20 line plan
    if (!ctx->tw400) {
      if (int rc = check_not_capturing(ctx, "the first 400^3 c2r (twiddle table)")) return rc;
      CHECK_HIP(hipMalloc(&ctx->tw400, 400 * sizeof(float) * 2));
      CHECK_HIP(launch_twiddles(400, ctx->tw400, ctx->stream));
    }
    CHECK_HIP(launch_c2r400(batch, in, out, ctx->tw400, ctx->stream));
    return 0;
  }
  FftPlan* p = nullptr;
  if (int rc = get_plan(ctx, n, batch, prec, dir, &p, dims)) return rc;
  rocfft_status s = rocfft_execution_info_set_stream(p->info, ctx->stream);
  REQUIRE(s == rocfft_status_success, "rocfft_execution_info_set_stream failed");
  void* ib[1] = {in};
  void* ob[1] = {out};
  s = rocfft_execute(p->plan, ib, ob, p->info);
  REQUIRE(s == rocfft_status_success, "rocfft_execute failed (status " + std::to_string((int)s) + ")");
  return 0;
}

int cnpc_fft_r2c(cnpc_ctx* ctx, int n, int batch, int prec, void* in, void* out) {
  return run_fft(ctx, n, batch, prec, 0, in, out);
}
int cnpc_fft_c2r(cnpc_ctx* ctx, int n, int batch, int prec, void* in, void* out) {
  return run_fft(ctx, n, batch, prec, 1, in, out);
}

int cnpc_prepare_spectrum(cnpc_ctx* ctx, int n, double boxsize, int prec, const void* c1, const void* c2,
                          int interlaced, int comp, int window, double scale, double sub0, int zero_k0, void* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 2 && n % 2 == 0 && boxsize > 0, "prepare: bad mesh");
  REQUIRE(prec == 0 || prec == 1, "prepare: prec must be 0 or 1");
  REQUIRE(comp >= 0 && comp <= 2 && window >= 1 && window <= 4, "prepare: bad compensation / window");
  REQUIRE(c1 && out && (!interlaced || c2), "prepare: null buffers");
  PrepareArgs a;
  a.n = n; a.interlaced = interlaced ? 1 : 0; a.comp = comp; a.p = window;
  a.kf = 2.0 * M_PI / boxsize; a.H = boxsize / n; a.scale = scale; a.sub0 = sub0;
  a.zero0 = zero_k0 ? 1 : 0; a.pad = 0;
  // per-axis tables: n doubles (compensation) + n complex doubles (phase)
  if (int rc = ensure_scratch(ctx, (size_t)(n + 2) * 24)) return rc;
  CHECK_HIP(launch_prepare(prec, a, c1, c2, out, ctx->scratch, ctx->stream));
  return 0;
}

int cnpc_bin_spectrum(cnpc_ctx* ctx, int n, double boxsize, int prec, const void* A, const void* B,
                      const double* edges, int nbins, int npoly, const cnpc_poly* polys, double ab_scale,
                      int noise_window, double noise_amp, double* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 2 && n % 2 == 0 && boxsize > 0, "bin: bad mesh");
  REQUIRE(nbins >= 1 && nbins <= CNPC_MAX_BINS, "bin: nbins must be in [1, 256]");
  REQUIRE(npoly >= 0 && npoly <= CNPC_MAX_POLY, "bin: npoly must be in [0, 9]");
  REQUIRE(edges && out, "bin: null edges / out");
  REQUIRE(npoly == 0 || (A && B && polys), "bin: null spectra / polys");
  REQUIRE(noise_window >= 0 && noise_window <= 4, "bin: bad noise window");
  for (int i = 0; i < nbins; ++i) REQUIRE(edges[i + 1] > edges[i], "bin: edges must increase");
  BinEdges E;
  fill_edges(E, edges, nbins);
  PolySet P;
  std::memset(&P, 0, sizeof(P));
  for (int q = 0; q < npoly; ++q) {
    REQUIRE(polys[q].deg >= 0 && polys[q].deg < CNPC_ZDEG && polys[q].m >= 0, "bin: bad polynomial");
    to_spec(P.s[q], polys[q]);
  }
  BinArgs a;
  a.n = n; a.npoly = npoly; a.noise_p = noise_window; a.conj_b = 1;
  a.kf = 2.0 * M_PI / boxsize; a.ab_scale = ab_scale; a.noise_amp = noise_amp;
  if (int rc = cnpc_zero(ctx, out, (int64_t)(2 + 2 * npoly) * nbins * 8)) return rc;
  CHECK_HIP(launch_bin(prec, a, E, P, A, B, out, ctx->stream));
  return 0;
}

int cnpc_cross_spectrum(cnpc_ctx* ctx, int n, int prec, const void* A, const void* B, double ab_scale,
                        int noise_window, double noise_amp, void* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 2 && n % 2 == 0, "cross_spectrum: bad mesh");
  REQUIRE(prec == 0 || prec == 1, "cross_spectrum: prec must be 0 or 1");
  REQUIRE(A && B && out, "cross_spectrum: null buffers");
  REQUIRE(noise_window >= 1 && noise_window <= 4, "cross_spectrum: noise window must be 1..4");
  BinArgs a;
  a.n = n; a.npoly = 0; a.noise_p = noise_window; a.conj_b = 1;
  a.kf = 0.0; a.ab_scale = ab_scale; a.noise_amp = noise_amp;
  CHECK_HIP(launch_cross_spectrum(prec, a, A, B, out, ctx->stream));
  return 0;
}

// mixed-radix factorisation of n, odd radices first (their Stockham writes
// stride by an odd count and stay off the LDS bank conflicts); 0 if n has a
// prime factor above 5
static int factor_radices(int n, SynthArgs& a) {
  int m = n, k = 0;
  for (int r : {5, 3})
    while (m % r == 0 && k < CNPC_MAX_RADICES) { a.rad[k++] = r; m /= r; }
  while (m % 4 == 0 && k < CNPC_MAX_RADICES) { a.rad[k++] = 4; m /= 4; }
  while (m % 2 == 0 && k < CNPC_MAX_RADICES) { a.rad[k++] = 2; m /= 2; }
  a.nrad = k;
  return m == 1;
}

// every mode with |k| < kmax has |n_d| <= R
static int synth_radius(double kmax, double boxsize) { return (int)std::ceil(kmax / (2.0 * M_PI / boxsize)); }

// per-shell radii of a synthesis call: shell s holds |k| < edges[bin0+s+1],
// so |n_d| <= ceil(edges / kf) (one more for rounding), at most R
static void shell_radii(SynthArgs& a, const double* edges, double boxsize) {
  for (int s = 0; s < a.nshell; ++s) {
    const int r = synth_radius(edges[a.bin0 + s + 1], boxsize) + 1;
    a.rs[s] = r < a.R ? r : a.R;
  }
}

int cnpc_shell_synth(cnpc_ctx* ctx, int n, double boxsize, int prec, const void* D, const double* edges, int nbins,
                     int bin0, int nshell, const cnpc_poly* weight, void* work, void* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 2 && n % 2 == 0 && boxsize > 0, "shell_synth: bad mesh");
  REQUIRE(prec == 0 || prec == 1, "shell_synth: prec must be 0 or 1");
  REQUIRE(nbins >= 1 && nbins <= 127 && edges, "shell_synth: nbins must be in [1, 127]");
  REQUIRE(bin0 >= 0 && nshell >= 1 && bin0 + nshell <= nbins, "shell_synth: shell range outside the bins");
  REQUIRE(work && out && weight && work != out, "shell_synth: null or aliased buffers");
  REQUIRE(weight->deg >= 0 && weight->deg < CNPC_ZDEG && weight->m >= 0 && weight->part >= 0 && weight->part <= 3,
          "shell_synth: bad weight polynomial");
  for (int i = 0; i < nbins; ++i) REQUIRE(edges[i + 1] > edges[i] && edges[i] >= 0, "shell_synth: bad edges");
  const double kf = 2.0 * M_PI / boxsize;
  const int R = synth_radius(edges[nbins], boxsize);
  REQUIRE(R < n / 2, "shell_synth: k_max reaches the Nyquist index (needs edges[nbins] < pi n / L)");
  SynthArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = n; a.R = R; a.bin0 = bin0; a.nshell = nshell; a.kf = kf;
  REQUIRE(nshell <= CNPC_SYNTH_MAXS, "shell_synth: too many shells in one call");
  shell_radii(a, edges, boxsize);
  REQUIRE(factor_radices(n, a), "shell_synth: nmesh must factor into 2, 3 and 5");
  if (prec == 1) merge_radices(a);  // the FP32 passes take the composite radices
  REQUIRE(synth_lines_per_block(prec, a) >= 1, "shell_synth: mesh too large for the LDS line buffers");
  const size_t cs = prec == 0 ? 16 : 8;
  const size_t P = 2 * (size_t)R + 1, Q = (size_t)R + 1, cube = P * Q * P;
  // work holds U (nshell n P Q) and T (nshell n Q n); its size is nshell
  // half spectra, nshell n n (n/2+1)
  REQUIRE(P * Q + Q * (size_t)n <= (size_t)n * (n / 2 + 1),
          "shell_synth: k_max too large for the work area (needs k_max below a quarter of the grid Nyquist)");
  const size_t vbytes = (cube * cs + 255) / 256 * 256, bbytes = (cube + 255) / 256 * 256;
  if (int rc = ensure_synth(ctx, vbytes + bbytes + (size_t)n * cs)) return rc;
  BinEdges E;
  fill_edges(E, edges, nbins);
  PolySpec W;
  to_spec(W, *weight);
  char* base = (char*)ctx->synth;
  CHECK_HIP(launch_synth(prec, a, E, W, D, base, (signed char*)(base + vbytes), base + vbytes + bbytes, work, out,
                         ctx->stream));
  return 0;
}

int cnpc_shell_synth_half(cnpc_ctx* ctx, int n, double boxsize, const void* D, const double* edges, int nbins,
                          int bin0, int nshell, const cnpc_poly* weight, void* work, void* out, int nslots, int slot0,
                          int* kz_rows, int reuse_cube) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 8 && n % 8 == 0 && boxsize > 0, "shell_synth_half: nmesh must be a multiple of 8");
  REQUIRE(nbins >= 1 && nbins <= 127 && edges, "shell_synth_half: nbins must be in [1, 127]");
  REQUIRE(bin0 >= 0 && nshell >= 1 && bin0 + nshell <= nbins, "shell_synth_half: shell range outside the bins");
  REQUIRE(slot0 >= 0 && nslots >= 1 && slot0 + nshell <= nslots, "shell_synth_half: slots outside the layout");
  REQUIRE(work && out && weight && work != out, "shell_synth_half: null or aliased buffers");
  REQUIRE(weight->deg >= 0 && weight->deg < CNPC_ZDEG && weight->m >= 0 && weight->part >= 0 && weight->part <= 3,
          "shell_synth_half: bad weight polynomial");
  for (int i = 0; i < nbins; ++i) REQUIRE(edges[i + 1] > edges[i] && edges[i] >= 0, "shell_synth_half: bad edges");
  const int R = synth_radius(edges[nbins], boxsize);
  REQUIRE(R < n / 2, "shell_synth_half: k_max reaches the Nyquist index (needs edges[nbins] < pi n / L)");
  SynthArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = n; a.R = R; a.bin0 = bin0; a.nshell = nshell; a.kf = 2.0 * M_PI / boxsize;
  REQUIRE(nshell <= CNPC_SYNTH_MAXS, "shell_synth_half: too many shells in one call");
  shell_radii(a, edges, boxsize);
  if (kz_rows)
    for (int s = 0; s < nshell; ++s) kz_rows[s] = a.rs[s] + 1;
  REQUIRE(factor_radices(n, a), "shell_synth_half: nmesh must factor into 2, 3 and 5");
  merge_radices(a);
  REQUIRE(synth_lines_per_block(1, a) >= 1, "shell_synth_half: mesh too large for the LDS line buffers");
  const size_t P = 2 * (size_t)R + 1, Q = (size_t)R + 1, cube = P * Q * P;
  // work holds U (nshell n P Q complex; P rounded up to whole blocks of 16
  // lines), at most nshell half spectra
  REQUIRE((P + 15) / 16 * 16 * Q <= (size_t)n * (n / 2 + 1), "shell_synth_half: k_max too large for the work area");
  const size_t vbytes = (cube * 8 + 255) / 256 * 256, bbytes = (cube + 255) / 256 * 256;
  REQUIRE(!reuse_cube || ctx->synth_bytes >= vbytes + bbytes + (size_t)n * 8,
          "shell_synth_half: reuse_cube without a previous cube of this shape");
  if (int rc = ensure_synth(ctx, vbytes + bbytes + (size_t)n * 8)) return rc;
  BinEdges E;
  fill_edges(E, edges, nbins);
  PolySpec W;
  to_spec(W, *weight);
  char* base = (char*)ctx->synth;
  CHECK_HIP(launch_synth_half(a, E, W, D, base, (signed char*)(base + vbytes), base + vbytes + bbytes, work, out,
                              nslots, slot0, reuse_cube ? 1 : 0, ctx->stream));
  return 0;
}

// fixed == nullptr: partials and twiddles in the context's scratch area
// (grown on demand); else in the CNPC_NOALLOC_SCRATCH_BYTES at fixed (no
// allocation: nblk is capped to what fits, sizes beyond it are refused).
static int contract_zgemm_impl(cnpc_ctx* ctx, int n, double boxsize, double kmax, const float* G, const void* HA,
                               int na, int sa_slots, const int* kz_a, const void* HB, int nb, int sb_slots,
                               const int* kz_b, int nblk, double* C, double* C_bb, char* fixed) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 8 && n % 8 == 0 && boxsize > 0 && kmax > 0, "contract_zgemm: nmesh must be a multiple of 8");
  REQUIRE(G && HA && C, "contract_zgemm: null buffers");
  REQUIRE(!C_bb || (HB && HB != HA), "contract_zgemm: C_bb needs a B different from A");
  REQUIRE(!C_bb || (C_bb != C), "contract_zgemm: C_bb must not alias C");
  REQUIRE(na >= 1 && na <= 32 && sa_slots >= na, "contract_zgemm: na must be in [1, 32] and fit the slots of HA");
  const int same = HB == nullptr || HB == HA;
  if (same) {
    REQUIRE(HB == nullptr || (nb == na && sb_slots == sa_slots), "contract_zgemm: B == A needs the same rows");
    nb = na;
    sb_slots = sa_slots;
  } else {
    REQUIRE(nb >= 1 && nb <= 32 && sb_slots >= nb, "contract_zgemm: nb must be in [1, 32] and fit the slots of HB");
  }
  REQUIRE((((uintptr_t)G | (uintptr_t)HA | (uintptr_t)(same ? HA : HB)) & 15) == 0,
          "contract_zgemm: buffers must be 16-byte aligned");
  const int R = synth_radius(kmax, boxsize);
  REQUIRE(R < n / 2, "contract_zgemm: k_max reaches the Nyquist index");
  SynthArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = n; a.R = R;
  REQUIRE(factor_radices(n, a), "contract_zgemm: nmesh must factor into 2, 3 and 5");
  merge_radices(a);
  int dev = 0, lds_max = 0;
  CHECK_HIP(hipGetDevice(&dev));
  CHECK_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  REQUIRE(zgemm_supported(a, na, nb, same, (size_t)lds_max),
          "contract_zgemm: grid too large for the LDS line buffers");
  const size_t per_blk = (size_t)1024 * 8 * (C_bb ? 2 : 1);  // FP64 partials (two sets with C_bb)
  const size_t twb = ((size_t)n * 8 + 255) / 256 * 256;       // line twiddles
  const bool auto_blk = nblk <= 0;
  if (auto_blk) {  // as many workgroups as stay resident
    CHECK_HIP(hipDeviceGetAttribute(&nblk, hipDeviceAttributeMultiprocessorCount, dev));
    nblk *= zgemm_blocks_per_cu(same);
  }
  if (nblk > 4096) nblk = 4096;
  char* base = nullptr;
  if (fixed) {
    REQUIRE(twb + per_blk <= CNPC_NOALLOC_SCRATCH_BYTES, "contract_zgemm_noalloc: nmesh " + std::to_string(n) +
                                                             " needs more than the static scratch");
    const int fit = (int)((CNPC_NOALLOC_SCRATCH_BYTES - twb) / per_blk);
    if (auto_blk && nblk > fit) nblk = fit;
    REQUIRE(nblk <= fit, "contract_zgemm_noalloc: " + std::to_string(nblk) + " workgroups need " +
                             std::to_string(twb + nblk * per_blk) + " bytes of scratch, more than the static " +
                             std::to_string(CNPC_NOALLOC_SCRATCH_BYTES) + " (at most " + std::to_string(fit) +
                             " workgroups" + (C_bb ? " with C_bb)" : ")"));
    base = fixed;
  } else {
    if (int rc = ensure_scratch(ctx, nblk * per_blk + twb)) return rc;
    base = (char*)ctx->scratch;
  }
  const size_t pbytes = nblk * per_blk;
  CHECK_HIP(launch_twiddles(n, base + pbytes, ctx->stream));
  int qa[32], qb[32];
  for (int s = 0; s < 32; ++s) {
    qa[s] = (kz_a && s < na) ? kz_a[s] : R + 1;
    qb[s] = (!same && kz_b && s < nb) ? kz_b[s] : (same ? qa[s] : R + 1);
    REQUIRE(qa[s] >= 0 && qa[s] <= R + 1 && qb[s] >= 0 && qb[s] <= R + 1,
            "contract_zgemm: kz rows per slot must be in [0, R + 1]");
  }
  CHECK_HIP(launch_zgemm(a, R + 1, HA, na, sa_slots, qa, same ? HA : HB, nb, sb_slots, qb, same, G, base + pbytes,
                         nblk, (double*)base, C, C_bb, ctx->stream));
  return 0;
}

int cnpc_contract_zgemm(cnpc_ctx* ctx, int n, double boxsize, double kmax, const float* G, const void* HA, int na,
                        int sa_slots, const int* kz_a, const void* HB, int nb, int sb_slots, const int* kz_b,
                        int nblk, double* C, double* C_bb) {
  return contract_zgemm_impl(ctx, n, boxsize, kmax, G, HA, na, sa_slots, kz_a, HB, nb, sb_slots, kz_b, nblk, C, C_bb,
                             nullptr);
}

int cnpc_contract_zgemm_noalloc(cnpc_ctx* ctx, int n, double boxsize, double kmax, const float* G, const void* HA,
                                int na, int sa_slots, const int* kz_a, const void* HB, int nb, int sb_slots,
                                const int* kz_b, int nblk, double* C, double* C_bb) {
  REQUIRE(ctx, "null context");
  REQUIRE(ctx->device >= 0 && ctx->device < CNPC_MAX_DEVICES && g_scratch_ptr[ctx->device],
          "contract_zgemm_noalloc: the static scratch of this device was not set up (cnpc_ctx_create)");
  return contract_zgemm_impl(ctx, n, boxsize, kmax, G, HA, na, sa_slots, kz_a, HB, nb, sb_slots, kz_b, nblk, C, C_bb,
                             (char*)g_scratch_ptr[ctx->device]);
}

int cnpc_contract_zgemm_supported(cnpc_ctx* ctx, int n, int na, int nb, int same, int* ok) {
  REQUIRE(ctx && ok, "contract_zgemm_supported: null arguments");
  *ok = 0;
  if (n < 8 || n % 8 != 0) return 0;
  SynthArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = n;
  if (!factor_radices(n, a)) return 0;
  merge_radices(a);
  int dev = 0, lds_max = 0;
  CHECK_HIP(hipGetDevice(&dev));
  CHECK_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  *ok = zgemm_supported(a, na, same ? na : nb, same ? 1 : 0, (size_t)lds_max);
  return 0;
}

int cnpc_resample_spectrum(cnpc_ctx* ctx, int n, int m, int prec, const void* in, int nmax, void* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 2 && n % 2 == 0 && m >= 2 && m % 2 == 0, "resample: meshes must be even");
  REQUIRE(prec == 0 || prec == 1, "resample: prec must be 0 or 1");
  REQUIRE(nmax >= 0 && nmax < m / 2 && nmax < n / 2, "resample: nmax must be below both Nyquist indices");
  REQUIRE(in && out && in != out, "resample: null or aliased buffers");
  CHECK_HIP(launch_resample(prec, n, m, nmax, in, out, ctx->stream));
  return 0;
}

int cnpc_shell_filter(cnpc_ctx* ctx, int n, double boxsize, int prec, const void* D, const double* edges, int nbins,
                      int bin0, int nshell, const cnpc_poly* weight, void* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 2 && n % 2 == 0 && boxsize > 0, "shell_filter: bad mesh");
  REQUIRE(prec == 0 || prec == 1, "shell_filter: prec must be 0 or 1");
  REQUIRE(nbins >= 1 && nbins <= CNPC_MAX_BINS && edges, "shell_filter: bad edges");
  REQUIRE(bin0 >= 0 && nshell >= 1 && bin0 + nshell <= nbins, "shell_filter: shell range outside the bins");
  REQUIRE(out && weight, "shell_filter: null buffers");
  REQUIRE(weight->deg >= 0 && weight->deg < CNPC_ZDEG && weight->m >= 0, "shell_filter: bad weight polynomial");
  BinEdges E;
  fill_edges(E, edges, nbins);
  PolySpec W;
  to_spec(W, *weight);
  ShellArgs a;
  a.n = n; a.bin0 = bin0; a.nshell = nshell; a.pad = 0;
  a.kf = 2.0 * M_PI / boxsize;
  a.stride = (int64_t)n * n * (n / 2 + 1);
  CHECK_HIP(launch_shell_filter(prec, a, E, W, D, out, ctx->stream));
  return 0;
}

int cnpc_contract_pairs(cnpc_ctx* ctx, int prec, int64_t ncell, const void* G, const void* A, const void* B,
                        int npairs, const int32_t* ia, const int32_t* ib, double* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(prec == 0 || prec == 1, "contract_pairs: prec must be 0 or 1");
  REQUIRE(ncell > 0 && npairs >= 0 && G && A && B && (npairs == 0 || (ia && ib && out)),
          "contract_pairs: bad arguments");
  if (npairs == 0) return 0;
  for (int p = 0; p < npairs; ++p) REQUIRE(ia[p] >= 0 && ib[p] >= 0, "contract_pairs: negative shell index");
  if (int rc = check_not_capturing(ctx, "contract_pairs reads its sums back to the host and cannot run")) return rc;
  const int nblocks = 2048;
  if (int rc = ensure_scratch(ctx, (size_t)(nblocks + npairs) * 8)) return rc;
  double* partial = (double*)ctx->scratch;
  double* dout = partial + nblocks;
  const size_t es = prec == 0 ? 8 : 4;
  for (int p = 0; p < npairs; ++p) {
    const char* a = (const char*)A + (size_t)ia[p] * ncell * es;
    const char* b = (const char*)B + (size_t)ib[p] * ncell * es;
    CHECK_HIP(launch_triple_dot(prec, G, a, b, ncell, partial, nblocks, dout + p, ctx->stream));
  }
  CHECK_HIP(hipMemcpyAsync(out, dout, (size_t)npairs * 8, hipMemcpyDeviceToHost, ctx->stream));
  CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int cnpc_contract_gemm(cnpc_ctx* ctx, int mode, int64_t ncell, const void* G, const void* A, int na, const void* B,
                       int nb, int same, float sa, float sb, int nblk, double* C) {
  REQUIRE(ctx, "null context");
  REQUIRE(mode >= 0 && mode <= 3, "contract_gemm: mode must be 0 (bf16x3), 1 (fp16x3), 2 (fp64) or 3 (fp8x6)");
  REQUIRE(ncell > 0 && na >= 1 && nb >= 1 && G && A && B && C, "contract_gemm: bad arguments");
  REQUIRE(ncell % 64 == 0, "contract_gemm: ncell must be a multiple of 64");
  REQUIRE(!same || (A == B && na == nb), "contract_gemm: same requires A == B and na == nb");
  REQUIRE(sa > 0 && sb > 0, "contract_gemm: prescales must be positive");
  const size_t es = mode == 2 ? 8 : 4;
  const uintptr_t mis = (uintptr_t)G | (uintptr_t)A | (uintptr_t)B;
  REQUIRE((mis & 15) == 0, "contract_gemm: buffers must be 16-byte aligned");
  (void)es;
  if (nblk <= 0) nblk = 2048;
  if (nblk > 65536) nblk = 65536;
  if (int rc = ensure_scratch(ctx, (size_t)nblk * 1024 * 8)) return rc;
  CHECK_HIP(launch_gemm(mode, G, A, ncell, na, B, ncell, nb, same ? 1 : 0, ncell, sa, sb, nblk,
                        (double*)ctx->scratch, C, ctx->stream));
  return 0;
}

int cnpc_absmax_f32(cnpc_ctx* ctx, const float* x, int64_t n, double* out) {
  REQUIRE(ctx && x && out && n > 0, "absmax: bad arguments");
  if (int rc = check_not_capturing(ctx, "absmax reads its result back to the host and cannot run")) return rc;
  const int nblocks = 1024;
  if (int rc = ensure_scratch(ctx, (size_t)nblocks * 8)) return rc;
  hipLaunchKernelGGL(absmax_kernel, dim3(nblocks), dim3(256), 0, ctx->stream, x, n, (double*)ctx->scratch);
  CHECK_HIP(hipGetLastError());
  std::vector<double> h(nblocks);
  CHECK_HIP(hipMemcpyAsync(h.data(), ctx->scratch, nblocks * 8, hipMemcpyDeviceToHost, ctx->stream));
  CHECK_HIP(hipStreamSynchronize(ctx->stream));
  double m = 0;
  for (double v : h) m = v > m ? v : m;
  *out = m;
  return 0;
}

// ---------------------------------------------------------------- survey

int cnpc_fft_c2c(cnpc_ctx* ctx, int n, int batch, int prec, int dir, void* in, void* out) {
  REQUIRE(dir == -1 || dir == 1, "fft_c2c: dir must be -1 (forward) or +1 (inverse)");
  return run_fft(ctx, n, batch, prec, dir == -1 ? 2 : 3, in, out);
}

int cnpc_full_spectrum(cnpc_ctx* ctx, int n, double boxsize, int prec, const void* c1, const void* c2,
                       int interlaced, int comp, int window, double scale, void* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 2 && n % 2 == 0 && boxsize > 0, "full_spectrum: bad mesh");
  REQUIRE(prec == 0 || prec == 1, "full_spectrum: prec must be 0 or 1");
  REQUIRE(comp >= 0 && comp <= 2 && window >= 1 && window <= 4, "full_spectrum: bad compensation / window");
  REQUIRE(c1 && out && (!interlaced || c2), "full_spectrum: null buffers");
  REQUIRE(out != c1 && out != c2, "full_spectrum: out must not alias the half spectra");
  FullSpecArgs a;
  a.n = n; a.interlaced = interlaced ? 1 : 0; a.comp = comp; a.p = window;
  a.kf = 2.0 * M_PI / boxsize; a.H = boxsize / n; a.scale = scale;
  CHECK_HIP(launch_full_spectrum(prec, a, c1, c2, out, ctx->stream));
  return 0;
}

static int check_poly(const cnpc_poly* w, const char* who) {
  REQUIRE(w->deg >= 0 && w->deg < CNPC_ZDEG && w->m >= 0 && w->part >= 0 && w->part <= 3,
          std::string(who) + ": bad weight polynomial");
  return 0;
}

int cnpc_xweight(cnpc_ctx* ctx, int n, double boxsize, int prec, const double* boxcenter, const cnpc_poly* w,
                 const void* F, void* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 2 && n % 2 == 0 && boxsize > 0, "xweight: bad mesh");
  REQUIRE(prec == 0 || prec == 1, "xweight: prec must be 0 or 1");
  REQUIRE(F && out && w && boxcenter, "xweight: null arguments");
  if (int rc = check_poly(w, "xweight")) return rc;
  XWeightArgs a;
  a.n = n; a.pad = 0; a.H = boxsize / n;
  for (int d = 0; d < 3; ++d) a.off[d] = 0.5 * boxsize / n + boxcenter[d];
  PolySpec W;
  to_spec(W, *w);
  CHECK_HIP(launch_xweight(prec, a, W, F, out, ctx->stream));
  return 0;
}

int cnpc_kcombine(cnpc_ctx* ctx, int n, double boxsize, int prec, const void* A, const void* B, int opb, int comp,
                  int window, int shot, double s_re, double s_im, const cnpc_poly* w, double c_re, double c_im,
                  int accumulate, void* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 2 && n % 2 == 0 && boxsize > 0, "kcombine: bad mesh");
  REQUIRE(prec == 0 || prec == 1, "kcombine: prec must be 0 or 1");
  REQUIRE(opb >= 0 && opb <= 2 && comp >= 0 && comp <= 2 && window >= 1 && window <= 4, "kcombine: bad options");
  REQUIRE(A && out && (opb == 0 || B), "kcombine: null buffers");
  REQUIRE(B != out || opb == 0 || B == A, "kcombine: B must not alias out");
  KCombineArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = n; a.opb = opb; a.comp = comp; a.p = window; a.shot = shot ? 1 : 0; a.poly = w ? 1 : 0;
  a.acc = accumulate ? 1 : 0; a.kf = 2.0 * M_PI / boxsize;
  a.s_re = s_re; a.s_im = s_im; a.c_re = c_re; a.c_im = c_im;
  PolySpec P;
  std::memset(&P, 0, sizeof(P));
  if (w) {
    if (int rc = check_poly(w, "kcombine")) return rc;
    to_spec(P, *w);
  }
  CHECK_HIP(launch_kcombine(prec, a, P, A, B, out, ctx->stream));
  return 0;
}

int cnpc_full_bin(cnpc_ctx* ctx, int n, double boxsize, int prec, const void* A, const void* B, int opb,
                  const double* edges, int nbins, double* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 2 && n % 2 == 0 && boxsize > 0, "full_bin: bad mesh");
  REQUIRE(prec == 0 || prec == 1, "full_bin: prec must be 0 or 1");
  REQUIRE(nbins >= 1 && nbins <= CNPC_MAX_BINS, "full_bin: nbins must be in [1, 256]");
  REQUIRE(opb >= 0 && opb <= 2 && A && out && edges && (opb == 0 || B), "full_bin: bad arguments");
  for (int i = 0; i < nbins; ++i) REQUIRE(edges[i + 1] > edges[i], "full_bin: edges must increase");
  BinEdges E;
  fill_edges(E, edges, nbins);
  if (int rc = cnpc_zero(ctx, out, (int64_t)2 * nbins * 8)) return rc;
  CHECK_HIP(launch_full_bin(prec, n, opb, 2.0 * M_PI / boxsize, E, A, B, out, ctx->stream));
  return 0;
}

int cnpc_triple_sum(cnpc_ctx* ctx, int prec, int64_t count, const void* A, const void* B, const void* C, double* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(prec == 0 || prec == 1, "triple_sum: prec must be 0 or 1");
  REQUIRE(count >= 1 && A && out, "triple_sum: bad arguments");
  if (int rc = ensure_scratch(ctx, (size_t)triple_sum_blocks(count) * 16)) return rc;
  CHECK_HIP(launch_triple_sum(prec, count, A, B, C, ctx->scratch, out, ctx->stream));
  return 0;
}

int cnpc_split_complex(cnpc_ctx* ctx, int prec, int64_t count, const void* src, void* re, void* im) {
  REQUIRE(ctx, "null context");
  REQUIRE(prec == 0 || prec == 1, "split_complex: prec must be 0 or 1");
  REQUIRE(count >= 0 && src && re && im && re != im, "split_complex: bad arguments");
  if (count == 0) return 0;
  CHECK_HIP(launch_split(prec, count, src, re, im, ctx->stream));
  return 0;
}

int cnpc_resample_full(cnpc_ctx* ctx, int n, int m, int prec, const void* in, int nmax, void* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 2 && m >= 2 && n % 2 == 0 && m % 2 == 0, "resample_full: bad grids");
  REQUIRE(prec == 0 || prec == 1, "resample_full: prec must be 0 or 1");
  REQUIRE(nmax >= 0 && 2 * nmax < m && 2 * nmax < n, "resample_full: nmax must be below m/2 and n/2");
  REQUIRE(in && out && in != out, "resample_full: null or aliased buffers");
  CHECK_HIP(launch_resample_full(prec, n, m, nmax, in, out, ctx->stream));
  return 0;
}

int cnpc_reflect_conj(cnpc_ctx* ctx, int n, int prec, const void* A, double c1, double c2, void* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(n >= 2 && n % 2 == 0, "reflect_conj: bad mesh");
  REQUIRE(prec == 0 || prec == 1, "reflect_conj: prec must be 0 or 1");
  REQUIRE(A && out && A != out, "reflect_conj: null or aliased buffers");
  CHECK_HIP(launch_reflect(prec, n, c1, c2, A, out, ctx->stream));
  return 0;
}

int cnpc_particle_sums(cnpc_ctx* ctx, int64_t np, const double* pos, const double* W, const double* Wfkp,
                       const double* NZ, int npoly, const cnpc_poly* polys, double* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(np >= 0 && out && npoly >= 0 && npoly <= CNPC_MAX_POLY, "particle_sums: bad arguments");
  REQUIRE(np == 0 || (W && Wfkp && (npoly == 0 || pos)), "particle_sums: null arrays");
  REQUIRE(npoly == 0 || polys, "particle_sums: null polys");
  const int nv = 5 + 2 * npoly;
  if (np == 0) return cnpc_zero(ctx, out, (int64_t)nv * 8);
  PolySet P;
  std::memset(&P, 0, sizeof(P));
  for (int q = 0; q < npoly; ++q) {
    if (int rc = check_poly(&polys[q], "particle_sums")) return rc;
    to_spec(P.s[q], polys[q]);
  }
  if (int rc = ensure_scratch(ctx, (size_t)particle_sums_blocks(np) * nv * 8)) return rc;
  CHECK_HIP(launch_particle_sums(np, pos, W, Wfkp, NZ, P, npoly, (double*)ctx->scratch, out, ctx->stream));
  return 0;
}

int cnpc_pair_weights(cnpc_ctx* ctx, int64_t np, const double* W, const double* Wfkp, int power, double scale,
                      double* out) {
  REQUIRE(ctx, "null context");
  REQUIRE(np >= 0 && (power == 1 || power == 2), "pair_weights: bad arguments");
  REQUIRE(np == 0 || (W && out), "pair_weights: null arrays");
  if (np == 0) return 0;
  CHECK_HIP(launch_pair_weights(np, W, Wfkp, power, scale, out, ctx->stream));
  return 0;
}

// Gauss-Legendre nodes/weights on [-1, 1] by Newton iteration on P_n.
static void gauss_legendre(int n, double* t, double* w) {
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double x = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = x;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (x * p1 - p0) / (x * x - 1.0);
      const double dx = p1 / dp;
      x -= dx;
      if (std::fabs(dx) < 1e-16) break;
    }
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; ++k) {
      const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (x * p1 - p0) / (x * x - 1.0);
    t[i] = -x;
    t[n - 1 - i] = x;
    w[i] = w[n - 1 - i] = 2.0 / ((1.0 - x * x) * dp * dp);
  }
}

int cnpc_radecz_to_xyz(cnpc_ctx* ctx, int64_t np, const double* ra, const double* dec, const double* z, double omega_m,
                       double* pos) {
  REQUIRE(ctx, "null context");
  REQUIRE(np >= 0 && omega_m > 0.0 && omega_m <= 1.0, "radecz_to_xyz: bad arguments (0 < omega_m <= 1)");
  REQUIRE(np == 0 || (ra && dec && z && pos), "radecz_to_xyz: null arrays");
  if (np == 0) return 0;
  static RadeczArgs base;
  static std::once_flag once;
  std::call_once(once, [] {
    std::memset(&base, 0, sizeof(base));
    base.nodes = 48;
    gauss_legendre(base.nodes, base.t, base.w);
  });
  RadeczArgs a = base;
  a.om = omega_m;
  a.ode = 1.0 - omega_m;
  a.dh = 299792.458 / 100.0;
  CHECK_HIP(launch_radecz(np, ra, dec, z, a, pos, ctx->stream));
  return 0;
}

int cnpc_radecz_to_xyz_flrw(cnpc_ctx* ctx, int64_t np, const double* ra, const double* dec, const double* z,
                            const double* cosmo, double dh, double* pos) {
  REQUIRE(ctx, "null context");
  REQUIRE(cosmo, "radecz_to_xyz_flrw: null cosmology");
  const int nm = (int)cosmo[5];
  REQUIRE(np >= 0 && cosmo[0] > 0.0 && cosmo[1] >= 0.0 && cosmo[3] >= 0.0 && cosmo[4] >= 0.0 && nm >= 0 && nm <= 3 &&
              (double)nm == cosmo[5] && dh > 0.0,
          "radecz_to_xyz_flrw: bad arguments (Omega_m > 0, Omega_gamma >= 0, 0..3 massive species, dh > 0)");
  for (int m = 0; m < nm; ++m) REQUIRE(cosmo[6 + m] >= 0.0, "radecz_to_xyz_flrw: negative neutrino mass");
  REQUIRE(np == 0 || (ra && dec && z && pos), "radecz_to_xyz_flrw: null arrays");
  if (np == 0) return 0;
  RadeczArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nodes = 48;
  gauss_legendre(a.nodes, a.t, a.w);
  a.om = cosmo[0];
  a.og = cosmo[1];
  a.ode = cosmo[2];
  a.nupre = cosmo[3];
  a.nless = cosmo[4];
  a.nmassive = nm;
  for (int m = 0; m < nm; ++m) a.y[m] = cosmo[6 + m];
  a.dh = dh;
  CHECK_HIP(launch_radecz(np, ra, dec, z, a, pos, ctx->stream));
  return 0;
}

}  // extern "C"
