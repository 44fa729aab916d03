/* cnpc.h -- C ABI of libcnpc_hip.so, the MI355X (gfx950) mesh-estimator
 * library behind cosmonpc_amd.run_stats().
 *
 * The reference (CosmoNPC v0.2.0) has no native or FFI boundary: its hot path
 * is Python over pmesh/numpy.  Each entry point below replaces one numeric
 * step of that Python path; the comment names the reference code it stands
 * for (paths relative to the reference's src/cosmonpc/).
 *
 * Conventions
 *  - Every function returns 0 on success, a nonzero status otherwise;
 *    cnpc_last_error() returns the message of the last failure on the calling
 *    thread.  Nothing throws across the boundary.
 *  - Device buffers are caller-owned raw device pointers (e.g. torch
 *    tensor.data_ptr()).  Host arrays (bin edges, coefficient tables) are
 *    plain host pointers and are copied into kernel arguments.
 *  - prec: 0 = FP64 data, 1 = FP32 data.
 *  - Meshes are C order N*N*N; half spectra are N*N*(N/2+1) complex
 *    interleaved (the rocFFT real-forward layout).  Frequencies are wrapped:
 *    n = i for i < N/2 and i - N otherwise.
 *  - All work is enqueued on the context's stream (cnpc_ctx_set_stream); a
 *    context is not thread-safe (one context per host thread / stream).
 *  - Calls do not allocate device memory except the first use of an FFT
 *    shape (plan + work area, cached in the context) and growth of the
 *    context's reduction scratch (cnpc_ctx_reserve pre-sizes it).
 */
#ifndef CNPC_H
#define CNPC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cnpc_ctx cnpc_ctx;

/* Polynomial weight of a unit vector n (khat or xhat):
 *   w(n) = pre * part( (x+iy)^m * sum_{j<=deg} zc[j] (r.n)^j ),  w(0) = k0
 * part: 0 the value, 1 its real part, 2 its imaginary part, 3 its conjugate.
 * Y_lm(n) and L_l(r.n) are of this form (Y_lm: r = zhat). */
typedef struct cnpc_poly {
  int32_t m;
  int32_t deg;
  int32_t part;
  int32_t pad;
  double zc[9];
  double pre_re, pre_im;
  double k0_re, k0_im;
  double rx, ry, rz;
} cnpc_poly;

const char* cnpc_last_error(void);
int cnpc_version(void);

/* Context: stream, cached rocFFT plans, reduction scratch.  Stands for the
 * pmesh ParticleMesh object the reference builds per paint
 * (mesh_generator.py:189-191). */
int cnpc_ctx_create(cnpc_ctx** out, int device);
int cnpc_ctx_destroy(cnpc_ctx* ctx);
int cnpc_ctx_set_stream(cnpc_ctx* ctx, void* hip_stream);
int cnpc_ctx_reserve(cnpc_ctx* ctx, int64_t scratch_bytes);
int cnpc_ctx_clear_plans(cnpc_ctx* ctx);

/* Paint weighted particles onto mesh0 (and the +0.5-cell shifted copy onto
 * mesh1 when interlaced).  Zero-fills the meshes first.  window: 1 ngp,
 * 2 cic, 3 tsc, 4 pcs.  boxcenter may be NULL.
 * Replaces pm.paint in pm_painter (mesh_generator.py:194-205). */
int cnpc_paint(cnpc_ctx* ctx, const double* pos_xyz, const double* weight, int64_t np, int nmesh, double boxsize,
               int window, int interlaced, const double* boxcenter, int prec, void* mesh0, void* mesh1);

/* Unnormalised 3-D real-to-half-spectrum forward FFT and its inverse, batch
 * meshes at a stride of one mesh.  The c2r input is overwritten.
 * Replace pmesh r2c / c2r (all call sites; e.g. mesh_generator.py:204-205,
 * :219; clustering_estimator.py:602, :815, :845, :938-969). */
int cnpc_fft_r2c(cnpc_ctx* ctx, int nmesh, int batch, int prec, void* in, void* out);
int cnpc_fft_c2r(cnpc_ctx* ctx, int nmesh, int batch, int prec, void* in, void* out);

/* out = comp(k) * (scale * combine(c1, c2) - sub0 * [k==0]), optionally
 * clearing k = 0.  combine = c1 (interlaced == 0) or c1/2 + c2/2 e^{i kH/2}.
 * comp: 0 none, 1 1/sinc^p, 2 1/sqrt(C_p) with p = window order.
 * Replaces the interlace combine (mesh_generator.py:207-218), the density
 * recovery (clustering_estimator.py:808) and cfield.apply(Compensate*)
 * (math_evaluator.py:206-275, :980-1021). out may alias c1. */
int cnpc_prepare_spectrum(cnpc_ctx* ctx, int nmesh, double boxsize, int prec, const void* c1, const void* c2,
                          int interlaced, int comp, int window, double scale, double sub0, int zero_k0, void* out);

/* Radial binning over bins [edges[i], edges[i+1]) of
 *   f(k) = ab_scale * a(k) conj(b(k)) - noise_amp * C_p/sinc^{2p}(k)
 * times each of npoly polynomial weights, with half-spectrum mode weights.
 * out (device, FP64, length (2 + 2*npoly)*nbins) receives
 * [count, sum|k|, Re sums (npoly x nbins), Im sums (npoly x nbins)].
 * a/b may be NULL when npoly == 0 (counts only).
 * Replaces get_kbin_count / radial_binning (math_evaluator.py:452-539) and
 * the binned products of clustering_estimator.py:638-693, :1405-1434. */
int cnpc_bin_spectrum(cnpc_ctx* ctx, int nmesh, double boxsize, int prec, const void* a, const void* b,
                      const double* edges, int nbins, int npoly, const cnpc_poly* polys, double ab_scale,
                      int noise_window, double noise_amp, double* out);

/* out = ab_scale * a(k) conj(b(k)) - noise_amp * C_p/sinc^{2p}(k) on the half
 * spectrum (noise_window p = 1..4).  Replaces the S3 source field of
 * clustering_estimator.py:1197-1211 (before its c2r); out may alias a. */
int cnpc_cross_spectrum(cnpc_ctx* ctx, int nmesh, int prec, const void* a, const void* b, double ab_scale,
                        int noise_window, double noise_amp, void* out);

/* Copy the modes |n_x|, |n_y|, n_z <= nmax (nmax < m/2, n/2) of an n-grid half
 * spectrum into an m-grid half spectrum, zero elsewhere.  A band-limited
 * triple product sum_x G F_i F_j is the same on any grid m >= 4R+2 when
 * the shells F have |n_d| <= R and G is cut at 2R, up to the factor (n/m)^3,
 * so the bispectrum contraction of clustering_estimator.py:972 can run on
 * the smaller grid.  No reference counterpart (pmesh keeps one grid). */
int cnpc_resample_spectrum(cnpc_ctx* ctx, int n, int m, int prec, const void* in, int nmax, void* out);

/* Shell filter: for s in [0, nshell): out[s] = w(khat) * delta(k) *
 * [edges[bin0+s] <= |k| < edges[bin0+s+1]] (full half spectra, stride
 * N*N*(N/2+1) complex); delta_k may be NULL (weight alone).  Replaces
 * (ylm_weighted_cfield * mask_i) of clustering_estimator.py:898-902,
 * :934-970 and the S3 shells conj(Y) * mask of :1213-1217. */
int cnpc_shell_filter(cnpc_ctx* ctx, int nmesh, double boxsize, int prec, const void* delta_k, const double* edges,
                      int nbins, int bin0, int nshell, const cnpc_poly* weight, void* out);

/* Pruned shell synthesis: out[s] = c2r(w(khat) delta(k) 1[k in bin bin0+s])
 * for s < nshell, the same real shells as cnpc_shell_filter followed by
 * cnpc_fft_c2r, computed as a sparse x-direction DFT over the modes inside
 * |n_d| <= R = ceil(edges[nbins] L / 2 pi) followed by a batched 2-D c2r
 * over (y, z).  work: nshell*n*n*(n/2+1) complex; out: nshell*n^3 real;
 * delta_k may be NULL (weight alone); nbins <= 127.  Replaces the shell IFFTs
 * c2r(ylm_weighted_cfield * mask_i) of clustering_estimator.py:934-970. */
int cnpc_shell_synth(cnpc_ctx* ctx, int nmesh, double boxsize, int prec, const void* delta_k, const double* edges,
                     int nbins, int bin0, int nshell, const cnpc_poly* weight, void* work, void* out);

/* Pairwise plain contraction: out[p] = sum_x G A_{ia[p]} B_{ib[p]}, one full
 * streaming pass per pair (the np.sum(G_00 * F_i * F_j) of
 * clustering_estimator.py:972, :1053).  Shell s of A is at A + s*ncell. out is
 * a host array of npairs doubles; the call synchronises the stream. */
int cnpc_contract_pairs(cnpc_ctx* ctx, int prec, int64_t ncell, const void* G, const void* A, const void* B,
                        int npairs, const int32_t* ia, const int32_t* ib, double* out);

/* GEMM contraction C[a][b] = sum_x A_a G B_b for a < na, b < nb (device FP64
 * C, row-major na x nb).  mode: 0 split-bf16 x3, 1 split-fp16 x3 (FP32
 * storage, sa/sb power-of-two prescales), 2 FP64 MFMA (FP64 storage).
 * same != 0 declares B == A (na == nb).  nblk: workgroups (0 = default).
 * Same sum as cnpc_contract_pairs for all pairs at once; the whole block
 * loop of clustering_estimator.py:911-981. */
int cnpc_contract_gemm(cnpc_ctx* ctx, int mode, int64_t ncell, const void* G, const void* A, int na, const void* B,
                       int nb, int same, float sa, float sb, int nblk, double* C);

/* Max |x| over n values (device FP32 in, host double out; synchronises).
 * Used to pick the power-of-two prescales of the fp16 split. */
int cnpc_absmax_f32(cnpc_ctx* ctx, const float* x, int64_t n, double* out);

/* Zero-fill n bytes of device memory with a kernel (safe under stream capture). */
int cnpc_zero(cnpc_ctx* ctx, void* ptr, int64_t nbytes);

/* Device synchronisation of the context stream. */
int cnpc_sync(cnpc_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
