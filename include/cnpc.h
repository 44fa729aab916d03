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
 *  - Graph recording: the enqueue-only calls (no host read-back) may be
 *    captured into a hipGraph on the context's stream once the same calls
 *    have run eagerly; a call that would create a plan or grow a work area,
 *    or that reads back to the host (cnpc_contract_pairs, cnpc_absmax_f32),
 *    fails while the stream is capturing.
 */
#ifndef CNPC_H
#define CNPC_H
#include <stddef.h>
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
/* delta = +1 / -1: take / release a hold on the context's work areas and
 * plans for a recorded graph that uses their addresses.  While held, a work
 * area that must grow keeps its old allocation alive (freed when the last
 * hold is released) and cnpc_ctx_clear_plans fails.  No counterpart in the
 * reference (pmesh has no graphs); replays must stay ordered with the
 * context's other work (same stream). */
int cnpc_ctx_hold(cnpc_ctx* ctx, int delta);

/* Paint weighted particles onto mesh0 (and the +0.5-cell shifted copy onto
 * mesh1 when interlaced); every cell of the meshes is overwritten.  window:
 * 1 ngp, 2 cic, 3 tsc, 4 pcs.  boxcenter may be NULL.  For nmesh % 8 == 0
 * (16..4096) the particles are sorted by tile (16^3 cells when nmesh % 16
 * == 0, else 8^3) and each tile is summed in LDS in 64-bit fixed point and
 * written once (no global atomics, bitwise reproducible; work memory grows
 * with np in the context); other meshes use per-particle atomic scatter.
 * Replaces pm.paint in pm_painter (mesh_generator.py:194-205). */
int cnpc_paint(cnpc_ctx* ctx, const double* pos_xyz, const double* weight, int64_t np, int nmesh, double boxsize,
               int window, int interlaced, const double* boxcenter, int prec, void* mesh0, void* mesh1);

/* Paint nbatch catalogs at once: catalog b is particles [offsets[b],
 * offsets[b+1]) of the concatenated arrays (offsets: device int64, nbatch +
 * 1 entries) and lands in mesh b of the mesh0 (and mesh1) batches, nbatch *
 * nmesh^3 values each.  total_count >= offsets[nbatch] and max_count >= the
 * largest catalog are upper bounds: they size the launches and the work
 * memory, while the device offsets decide what is painted, so the call can
 * be recorded once in a graph and replayed with other catalogs within the
 * same bounds.  For nmesh % 8 == 0 every catalog takes the sorted,
 * LDS-tiled path of cnpc_paint in the same launches (every cell written
 * once, no global atomics); other meshes are zero-filled and painted by
 * per-particle atomic scatter.  The multi-mock form of pm_painter
 * (mesh_generator.py:172-224) for runs over many catalogs. */
int cnpc_paint_batch(cnpc_ctx* ctx, const double* pos_xyz, const double* weight, const int64_t* offsets, int nbatch,
                     int64_t total_count, int64_t max_count, int nmesh, double boxsize, int window, int interlaced,
                     const double* boxcenter, int prec, void* mesh0, void* mesh1);

/* Unnormalised 3-D real-to-half-spectrum forward FFT and its inverse, batch
 * meshes at a stride of one mesh.  The c2r input is overwritten.  The FP32
 * c2r of nmesh 400 (the contraction grid of the 512^3 bispectrum) runs on a
 * fixed 20 x 20 line plan in three passes; the rest through rocFFT plans.
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
 * cnpc_fft_c2r, computed by x, y and z FFT passes restricted to the modes
 * inside |n_d| <= R = ceil(edges[nbins] L / 2 pi).  work:
 * nshell*n*n*(n/2+1) complex; out: nshell*n^3 real; delta_k may be NULL
 * (weight alone); nbins <= 127; n must factor into 2, 3 and 5.  Replaces the
 * shell IFFTs c2r(ylm_weighted_cfield * mask_i) of
 * clustering_estimator.py:934-970. */
int cnpc_shell_synth(cnpc_ctx* ctx, int nmesh, double boxsize, int prec, const void* delta_k, const double* edges,
                     int nbins, int bin0, int nshell, const cnpc_poly* weight, void* work, void* out);

/* The x and y passes of cnpc_shell_synth (FP32 only): shells bin0 ..
 * bin0+nshell-1 into slots slot0 .. slot0+nshell-1 of the half-synthesised
 * layout (complex64), the input of cnpc_contract_zgemm, with R = ceil(
 * edges[nbins] / (2 pi / boxsize)) and Q = R+1: for nmesh == 400 and 4Q <=
 * nmesh the blocked out[x][nslots][ceil(Q/8)][y/2][kz%8][y%2]
 * (n*nslots*ceil(Q/8)*n*8 entries, kz block kz/8 on the third axis), else
 * out[x][y/2][nslots][Q][y%2] (n*(n/2)*nslots*Q*2 entries).  A shell's modes have
 * kz < kz_rows[s] (its own radius): only those rows of its slot are written.
 * kz_rows: host array of nshell ints, or NULL.  work: nshell half spectra.
 * reuse_cube != 0 skips rebuilding the weighted, binned mode cube: the caller
 * guarantees the same delta_k contents, weight, edges and grid as the
 * previous synthesis call on this context.  n must be a multiple of 8.
 * Same shells as clustering_estimator.py:934-970 before their z transform. */
int cnpc_shell_synth_half(cnpc_ctx* ctx, int nmesh, double boxsize, const void* delta_k, const double* edges,
                          int nbins, int bin0, int nshell, const cnpc_poly* weight, void* work, void* out,
                          int nslots, int slot0, int* kz_rows, int reuse_cube);

/* Pairwise plain contraction: out[p] = sum_x G A_{ia[p]} B_{ib[p]}, one full
 * streaming pass per pair (the np.sum(G_00 * F_i * F_j) of
 * clustering_estimator.py:972, :1053).  Shell s of A is at A + s*ncell. out is
 * a host array of npairs doubles; the call synchronises the stream. */
int cnpc_contract_pairs(cnpc_ctx* ctx, int prec, int64_t ncell, const void* G, const void* A, const void* B,
                        int npairs, const int32_t* ia, const int32_t* ib, double* out);

/* GEMM contraction C[a][b] = sum_x A_a G B_b for a < na, b < nb (device FP64
 * C, row-major na x nb).  mode: 0 split-bf16 x3, 1 split-fp16 x3, 3 split-fp8
 * (three e4m3 pieces, six products; prescaled |values| <= 2^7) (FP32
 * storage, sa/sb power-of-two prescales), 2 FP64 MFMA (FP64 storage).
 * same != 0 declares B == A (na == nb).  nblk: workgroups (0 = default).
 * Same sum as cnpc_contract_pairs for all pairs at once; the whole block
 * loop of clustering_estimator.py:911-981. */
int cnpc_contract_gemm(cnpc_ctx* ctx, int mode, int64_t ncell, const void* G, const void* A, int na, const void* B,
                       int nb, int same, float sa, float sb, int nblk, double* C);

/* Fused z synthesis + split-bf16 contraction: C[a][b] = sum_x A_a G B_b
 * (device FP64 C, row-major na x nb) where A_a, B_b are the shells held in
 * slots 0 .. na-1 of HA and 0 .. nb-1 of HB in the layout of
 * cnpc_shell_synth_half (sa_slots / sb_slots slots per tile; kmax =
 * edges[nbins] of that call), G a real FP32 n^3 mesh.  kz_a / kz_b: the
 * kz_rows that call reported for each slot (host arrays; NULL = all R+1
 * rows).  HB NULL (or HA) means B == A.  na, nb <= 32.  Equals
 * cnpc_shell_synth followed by cnpc_contract_gemm mode 0 without the real
 * shells in HBM; the block loop of clustering_estimator.py:911-981.
 * C_bb (NULL = not computed; needs B != A): also C_bb[b][c] = sum_x B_b G
 * B_c (nb x nb) from the same z transforms, the signal matrix of a second
 * configuration whose legs are both B (the [0,0,0] configuration beside
 * [l,0,l], which the reference computes in a separate run). */
int cnpc_contract_zgemm(cnpc_ctx* ctx, int nmesh, double boxsize, double kmax, const float* G, const void* HA, int na,
                        int sa_slots, const int* kz_a, const void* HB, int nb, int sb_slots, const int* kz_b,
                        int nblk, double* C, double* C_bb);

/* Size of the static per-device scratch of cnpc_contract_zgemm_noalloc:
 * 8 KiB of FP64 partials per workgroup (16 KiB with C_bb) plus the line
 * twiddles (nmesh * 8 bytes, rounded up to 256). */
#define CNPC_NOALLOC_SCRATCH_BYTES ((size_t)(8u << 20) + (size_t)(64u << 10))

/* cnpc_contract_zgemm without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as cnpc_contract_zgemm; the partial tiles
 * and twiddles live in a static device array of CNPC_NOALLOC_SCRATCH_BYTES
 * per device (set up by cnpc_ctx_create) instead of the context's scratch
 * area.  nblk <= 0 takes as many workgroups as fit that array; an explicit
 * nblk that does not fit, or a grid whose twiddles do not, is refused (status
 * 1, message in cnpc_last_error).  Every call on a device shares its array:
 * calls that overlap on one device (two streams, two host threads) race. */
int cnpc_contract_zgemm_noalloc(cnpc_ctx* ctx, int nmesh, double boxsize, double kmax, const float* G, const void* HA,
                                int na, int sa_slots, const int* kz_a, const void* HB, int nb, int sb_slots,
                                const int* kz_b, int nblk, double* C, double* C_bb);

/* *ok = 1 when cnpc_contract_zgemm accepts this grid and these row counts
 * on the current device (LDS line buffers and FFT staging), else 0. */
int cnpc_contract_zgemm_supported(cnpc_ctx* ctx, int nmesh, int na, int nb, int same, int* ok);

/* Max |x| over n values (device FP32 in, host double out; synchronises).
 * Used to pick the power-of-two prescales of the fp16 split. */
int cnpc_absmax_f32(cnpc_ctx* ctx, const float* x, int64_t n, double* out);

/* Zero-fill n bytes of device memory with a kernel (safe under stream capture). */
int cnpc_zero(cnpc_ctx* ctx, void* ptr, int64_t nbytes);

/* Device synchronisation of the context stream. */
int cnpc_sync(cnpc_ctx* ctx);

/* ------------------------------------------------------------------ survey
 * Survey fields live on full N^3 complex meshes ("full"), index
 * (ix*N + iy)*N + iz, as the reference keeps them on complex128 c2c meshes. */

/* Unnormalised 3-D complex-to-complex FFT of full meshes, out of place.
 * dir: -1 forward (e^{-ik.x}), +1 inverse.  Replaces pmesh r2c / c2r of the
 * survey estimators' complex fields (clustering_estimator.py:3116-3126,
 * :1711-1717, :2257-2269, :2375-2381). */
int cnpc_fft_c2c(cnpc_ctx* ctx, int nmesh, int batch, int prec, int dir, void* in, void* out);

/* Full spectrum of a painted field from its half spectra:
 * out(k) = scale * comp(k) * combine(c1, c2)(k) on the whole N^3 grid, with
 * combine as in cnpc_prepare_spectrum.  Gives r2c(rfield) * scale, with the
 * non-Hermitian Nyquist planes of the interlaced field intact
 * (mesh_generator.py:207-219 followed by clustering_estimator.py:334-343). */
int cnpc_full_spectrum(cnpc_ctx* ctx, int nmesh, double boxsize, int prec, const void* c1, const void* c2,
                       int interlaced, int comp, int window, double scale, void* out);

/* out(x) = F(x) * w(xhat) on a full mesh, xhat the unit vector of the cell
 * centre x_wrapped + H/2 + boxcenter.  out may alias F.
 * Replaces rfield * Y_lm(xgrid) (get_xgrid math_evaluator.py:25-32;
 * clustering_estimator.py:3116, :3120, :1710, :2256, :2374). */
int cnpc_xweight(cnpc_ctx* ctx, int nmesh, double boxsize, int prec, const double* boxcenter, const cnpc_poly* w,
                 const void* F, void* out);

/* out(k) (+)= c * w(khat) * [comp(k) * (A op B)(k) - s * C_bk(k)] on a full
 * spectrum.  opb: 0 A, 1 A*B, 2 A*conj(B) (B may alias A); comp codes as in
 * cnpc_prepare_spectrum; shot != 0 subtracts s times the bk shot-noise
 * window prod_d C_p/sinc^{2p}; w may be NULL (w = 1); accumulate != 0 adds
 * to out.  out may alias A.  Replaces the Y_lm(kgrid) weighting, the
 * compensation .apply() calls and the S_LM subtraction of
 * clustering_estimator.py:3117-3134, :1711-1714, :2259-2294, :2376-2396,
 * :2541-2564. */
int cnpc_kcombine(cnpc_ctx* ctx, int nmesh, double boxsize, int prec, const void* A, const void* B, int opb, int comp,
                  int window, int shot, double s_re, double s_im, const cnpc_poly* w, double c_re, double c_im,
                  int accumulate, void* out);

/* Per-bin sums of (A op B)(k) over the full grid into the device FP64 array
 * out = [re(nbins), im(nbins)].  opb as in cnpc_kcombine.
 * Replaces radial_binning on complex c2c fields (math_evaluator.py:504-539;
 * clustering_estimator.py:456, :2306, :2398). */
int cnpc_full_bin(cnpc_ctx* ctx, int nmesh, double boxsize, int prec, const void* A, const void* B, int opb,
                  const double* edges, int nbins, double* out);

/* (re, im) of sum_x A(x)*B(x)*C(x) over count complex values (B, C may be
 * NULL) into the device FP64 array out[2]; FP64, deterministic order.
 * Replaces the mixed-mesh / mesh normalisation sums
 * (mesh_generator.py:499-504, :828, :920-922, :940-987). */
int cnpc_triple_sum(cnpc_ctx* ctx, int prec, int64_t count, const void* A, const void* B, const void* C, double* out);

/* Split count complex values into real and imaginary arrays (same precision). */
int cnpc_split_complex(cnpc_ctx* ctx, int prec, int64_t count, const void* src, void* re, void* im);

/* Full spectrum on N^3 -> its |n_d| <= nmax modes on an M^3 full grid
 * (M > 2 nmax, other modes zero).  Spectral truncation used by the reduced
 * contraction grid; no reference counterpart (the reference contracts on N^3). */
int cnpc_resample_full(cnpc_ctx* ctx, int n, int m, int prec, const void* in, int nmax, void* out);

/* out(k) = c1 * A(k) + c2 * conj(A(-k)) on a full spectrum, -k the wrapped
 * index inverse on every axis (out must not alias A).  Replaces
 * G + conj(space_inversion_transposed_complex(G)) of the survey pk fast
 * estimators (clustering_estimator.py:429-444; math_evaluator.py:806-977). */
int cnpc_reflect_conj(cnpc_ctx* ctx, int nmesh, int prec, const void* A, double c1, double c2, void* out);

/* FP64 catalog sums over np objects into the device FP64 array out:
 *   out[0] sum W*Wfkp, out[1] sum W, out[2] sum W*NZ*Wfkp^2,
 *   out[3] sum (W*Wfkp)^2, out[4] sum NZ^2 W^2 Wfkp^3, then for each of the
 *   npoly weights (re, im) of sum (W*Wfkp)^3 * w(pos/|pos|).
 * NZ may be NULL (1).  Replaces compute_particle_sums_pk / _bk
 * (mesh_generator.py:1131-1268). */
int cnpc_particle_sums(cnpc_ctx* ctx, int64_t np, const double* pos_xyz, const double* W, const double* Wfkp,
                       const double* NZ, int npoly, const cnpc_poly* polys, double* out);

/* out[i] = scale * (W[i] * Wfkp[i])^power, power 1 or 2 (Wfkp may be NULL).
 * The paint weights of mesh_generator.py:362-382, :1084-1120. */
int cnpc_pair_weights(cnpc_ctx* ctx, int64_t np, const double* W, const double* Wfkp, int power, double scale,
                      double* out);

/* RA/DEC (degrees) and redshift -> comoving Cartesian positions (np x 3,
 * Mpc/h) in flat LCDM without radiation (h only sets units: Mpc/h).
 * Replaces ra_dec_z_to_xyz (catalog_processor.py:589-645). */
int cnpc_radecz_to_xyz(cnpc_ctx* ctx, int64_t np, const double* ra, const double* dec, const double* z, double omega_m,
                       double* pos_xyz);

/* The same in a flat FLRW cosmology with photons and neutrinos (massless
 * and massive), the default Planck18 branch of ra_dec_z_to_xyz
 * (catalog_processor.py:609-632: astropy's Planck18 comoving distance in Mpc
 * times the config's h).  cosmo[9] = {Omega_m, Omega_gamma, Omega_de,
 * nu_prefac, n_massless_nu, n_massive_nu (0..3), y_1, y_2, y_3}: E(z)^2 =
 * (1+z)^3 (Omega_m + Omega_gamma (1 + nu(z)) (1+z)) + Omega_de with
 * nu(z) = nu_prefac (sum_i f(y_i / (1+z)) + n_massless_nu),
 * f(y) = (1 + (0.3173 y)^1.83)^(1/1.83) (y_i = m_i c^2 / (k_B T_nu0));
 * dh = c / H0 in the output unit (Mpc/h: c / H0[Mpc] times h). */
int cnpc_radecz_to_xyz_flrw(cnpc_ctx* ctx, int64_t np, const double* ra, const double* dec, const double* z,
                            const double* cosmo, double dh, double* pos_xyz);

#ifdef __cplusplus
}
#endif
#endif
