// Shared declarations of the cnpc HIP library (gfx950).
//
// Layouts in HBM (all C order, x slowest, z fastest):
//   real mesh       : N*N*N scalars (double or float)
//   half spectrum   : N*N*(N/2+1) complex (double2 / float2), the rocFFT
//                     real-forward output; index (ix*N + iy)*(N/2+1) + iz
//   shell stack     : nshell consecutive real meshes, shell s at s*N^3
// Frequencies follow the wrapped convention n = i for i < N/2, i - N for
// i >= N/2 (so the last z plane of the half spectrum is n_z = -N/2).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#define CNPC_MAX_BINS 256
#define CNPC_MAX_POLY 9
#define CNPC_ZDEG 9

namespace cnpc {

void set_error(const std::string& msg);

// Complex polynomial weight evaluated at a unit vector n = (x, y, z):
//   w(n) = pre * part(P(n)),  part = P (0), Re P (1), Im P (2), conj P (3)
//   P(n) = (x + i y)^m * sum_j zc[j] * t^j,   t = rx*x + ry*y + rz*z
// At the zero vector (the k = 0 mode, where khat is set to 0) the weight is
// the fixed value (k0_re, k0_im).  Spherical harmonics Y_lm (m >= 0, axis z)
// and Legendre polynomials L_l(r . n) are both of this form.
struct PolySpec {
  int m;
  int deg;
  int part;
  int pad;
  double zc[CNPC_ZDEG];
  double pre_re, pre_im;
  double k0_re, k0_im;
  double rx, ry, rz;
};

struct BinEdges {
  int nbins;
  int pad;
  double e0;
  double inv_de;
  double e[CNPC_MAX_BINS + 1];
};

__device__ __forceinline__ int wrap_freq(int i, int n) { return i >= (n >> 1) ? i - n : i; }

// |k| exactly as the reference forms it: sqrt((0 + kx^2) + ky^2) + kz^2) with
// k_d = n_d * (2 pi / L); no fused multiply-add so bin membership is the
// same bit pattern as numpy's.
__device__ __forceinline__ double knorm_of(int nx, int ny, int nz, double kf, double* kx, double* ky, double* kz) {
  double a = __dmul_rn((double)nx, kf);
  double b = __dmul_rn((double)ny, kf);
  double c = __dmul_rn((double)nz, kf);
  *kx = a; *ky = b; *kz = c;
  double s = __dadd_rn(__dadd_rn(__dmul_rn(a, a), __dmul_rn(b, b)), __dmul_rn(c, c));
  return __dsqrt_rn(s);
}

// bin b with e[b] <= k < e[b+1], or -1 (numpy searchsorted side='right' - 1
// restricted to [e[0], e[nbins]) as in math_evaluator.radial_binning).
__device__ __forceinline__ int bin_of(const BinEdges& E, double k) {
  if (!(k >= E.e[0]) || !(k < E.e[E.nbins])) return -1;
  int b = (int)floor((k - E.e0) * E.inv_de);
  if (b < 0) b = 0;
  if (b > E.nbins - 1) b = E.nbins - 1;
  while (b > 0 && k < E.e[b]) --b;
  while (b < E.nbins - 1 && k >= E.e[b + 1]) ++b;
  return b;
}

__device__ __forceinline__ double2 poly_eval(const PolySpec& s, double x, double y, double z, bool zero_vec) {
  if (zero_vec) return make_double2(s.k0_re, s.k0_im);
  double t = s.rx * x + s.ry * y + s.rz * z;
  double p = s.zc[s.deg];
  for (int j = s.deg - 1; j >= 0; --j) p = p * t + s.zc[j];
  double re = p, im = 0.0;
  for (int j = 0; j < s.m; ++j) {  // multiply by (x + i y)
    double nr = re * x - im * y;
    double ni = re * y + im * x;
    re = nr; im = ni;
  }
  if (s.part == 1) {
    im = 0.0;
  } else if (s.part == 2) {
    re = im;
    im = 0.0;
  } else if (s.part == 3) {
    im = -im;
  }
  return make_double2(s.pre_re * re - s.pre_im * im, s.pre_re * im + s.pre_im * re);
}

// Window compensations on the circular frequency w_d = 2 pi n_d / N.
// comp: 0 none, 1 divide by prod_d sinc(w_d/2)^p  (CompensateNGP..PCS,
//       math_evaluator.py:206-248), 2 divide by prod_d sqrt(C_p(w_d))
//       (Compensate*Shotnoise, math_evaluator.py:251-275).
__device__ __forceinline__ double alias_c(int p, double s) {
  if (p == 2) return 1.0 - 2.0 / 3 * s;
  if (p == 3) return 1.0 - s + 2.0 / 15 * s * s;
  if (p == 4) return 1.0 - 4.0 / 3.0 * s + 2.0 / 5.0 * s * s - 4.0 / 315.0 * s * s * s;
  return 1.0;
}

__device__ __forceinline__ double sinc_half(double w) {  // np.sinc(0.5 w / pi)
  if (w == 0.0) return 1.0;
  double h = 0.5 * w;
  return sin(h) / h;
}

__device__ __forceinline__ double comp_factor(int comp, int p, int nx, int ny, int nz, int n) {
  if (comp == 0) return 1.0;
  const double c = 2.0 * M_PI / n;
  double w[3] = {nx * c, ny * c, nz * c};
  double f = 1.0;
  for (int d = 0; d < 3; ++d) {
    if (comp == 1) {
      double t = sinc_half(w[d]);
      double tp = t;
      for (int i = 1; i < p; ++i) tp *= t;
      f /= tp;
    } else if (p > 1) {
      double s = sin(0.5 * w[d]);
      f /= sqrt(alias_c(p, s * s));
    }
  }
  return f;
}

// C_p(w) / sinc(w/2)^{2p} product (Compensate_bk_noise_*, math_evaluator.py:278-309)
__device__ __forceinline__ double bk_noise_factor(int p, int nx, int ny, int nz, int n) {
  if (p <= 1) return 1.0;
  const double c = 2.0 * M_PI / n;
  double w[3] = {nx * c, ny * c, nz * c};
  double f = 1.0;
  for (int d = 0; d < 3; ++d) {
    double s = sin(0.5 * w[d]);
    double t = sinc_half(w[d]);
    double tp = t;
    for (int i = 1; i < p; ++i) tp *= t;
    f *= alias_c(p, s * s) / (tp * tp);
  }
  return f;
}

template <typename T> struct cplx;
template <> struct cplx<double> { typedef double2 type; };
template <> struct cplx<float> { typedef float2 type; };

inline unsigned grid_for(int64_t n, int block, int max_blocks = 8192) {
  int64_t g = (n + block - 1) / block;
  if (g > max_blocks) g = max_blocks;
  if (g < 1) g = 1;
  return (unsigned)g;
}


// ---------------------------------------------------------------- kernel args
// comp: 0 none, 1 divide by prod_d sinc(w_d/2)^p, 2 divide by prod_d sqrt(C_p(w_d))
struct PrepareArgs {
  int n;
  int interlaced;
  int comp;
  int p;
  double kf;      // 2 pi / L
  double H;       // L / N
  double scale;   // multiplies the (combined) spectrum
  double sub0;    // subtracted from the k = 0 mode after scaling
  int zero0;      // clear the k = 0 mode
  int pad;
};

struct BinArgs {
  int n;
  int npoly;
  int noise_p;        // window order of the bk noise term
  int conj_b;         // 1: f = a * conj(b);  0: f = a * b
  double kf;
  double ab_scale;    // multiplies a*conj(b)
  double noise_amp;   // f -= noise_amp * C_p/sinc^{2p}
};

struct PolySet {
  PolySpec s[CNPC_MAX_POLY];
};

struct ShellArgs {
  int n;
  int bin0;
  int nshell;
  int pad;
  double kf;
  int64_t stride;  // complex elements between consecutive output shells
};

#define CNPC_MAX_RADICES 24

#define CNPC_SYNTH_MAXS 128
#define CNPC_MAX_DEVICES 64  // per-device launch state kept by the library

// Layout of the half-synthesised shells (x and y transformed, kz rows 0 ..
// Q-1 with Q = R + 1).  Tile-major, H[x][y/2][slot][kz][y%2] (complex),
// one contiguous run per (x, y pair) for the fused contraction; or, for the
// fixed 20 x 20 plans when 4Q <= n, blocked, H[x][slot][kz/8][y/2][kz%8][y%2]:
// the y pass then writes one contiguous run of n x 8 values per (x, slot,
// kz block), and the contraction reads rows y, y+1 of a (slot, kz block)
// as 128 contiguous bytes.
__host__ __device__ inline int half_blocked(int n, int Q) { return n == 400 && 4 * Q <= n ? 1 : 0; }

struct SynthArgs {
  int n;        // grid
  int R;        // |n_d| <= R holds every shell mode
  int bin0;     // first shell
  int nshell;   // shells in this call
  double kf;
  int nrad;     // mixed-radix factorisation of n (radices 2, 3, 4, 5)
  int rad[CNPC_MAX_RADICES];
  int rs[CNPC_SYNTH_MAXS];  // per shell of the call: its modes have |n_d| <= rs[s] <= R
};

// ---------------------------------------------------------------- survey (full grids)
struct FullSpecArgs {
  int n;
  int interlaced;
  int comp;
  int p;
  double kf;
  double H;
  double scale;
};

struct XWeightArgs {
  int n;
  int pad;
  double H;
  double off[3];  // H/2 + boxcenter
};

// out (+)= c * w(khat) * [comp(k) * (A op B) - s * C_bk(k)]
struct KCombineArgs {
  int n;
  int opb;   // 0: A, 1: A*B, 2: A*conj(B)
  int comp;  // comp_factor code
  int p;     // window order
  int shot;  // subtract s * bk_noise_factor
  int poly;  // multiply by the polynomial weight of khat
  int acc;   // accumulate into out
  int pad;
  double kf;
  double s_re, s_im;
  double c_re, c_im;
};

#define CNPC_GL_MAX 64
struct RadeczArgs {
  int nodes;
  int nmassive;   // massive neutrino species (0..3)
  double om;      // Omega_m
  double og;      // Omega_gamma (0: no radiation)
  double ode;     // Omega_de
  double nupre;   // 7/8 (4/11)^(4/3) N_eff / N_species
  double nless;   // massless neutrino species
  double y[3];    // m c^2 / (k_B T_nu0) of the massive species
  double dh;      // c / H0 in the output unit
  double t[CNPC_GL_MAX];
  double w[CNPC_GL_MAX];
};

#define CNPC_PSUM_MAX (5 + 2 * CNPC_MAX_POLY)

hipError_t launch_full_spectrum(int prec, const FullSpecArgs& a, const void* c1, const void* c2, void* out,
                                hipStream_t st);
hipError_t launch_xweight(int prec, const XWeightArgs& a, const PolySpec& W, const void* F, void* out, hipStream_t st);
hipError_t launch_kcombine(int prec, const KCombineArgs& a, const PolySpec& P, const void* A, const void* B, void* out,
                           hipStream_t st);
hipError_t launch_full_bin(int prec, int n, int opb, double kf, const BinEdges& E, const void* A, const void* B,
                           double* out, hipStream_t st);
int triple_sum_blocks(int64_t count);
hipError_t launch_triple_sum(int prec, int64_t count, const void* A, const void* B, const void* C, void* partial,
                             double* out, hipStream_t st);
hipError_t launch_split(int prec, int64_t count, const void* src, void* re, void* im, hipStream_t st);
hipError_t launch_resample_full(int prec, int n, int m, int nmax, const void* in, void* out, hipStream_t st);
hipError_t launch_reflect(int prec, int n, double c1, double c2, const void* A, void* out, hipStream_t st);
int particle_sums_blocks(int64_t np);
hipError_t launch_particle_sums(int64_t np, const double* pos, const double* W, const double* WF, const double* NZ,
                                const PolySet& P, int npoly, double* partial, double* out, hipStream_t st);
hipError_t launch_pair_weights(int64_t np, const double* W, const double* WF, int power, double scale, double* out,
                               hipStream_t st);
hipError_t launch_radecz(int64_t np, const double* ra, const double* dec, const double* z, const RadeczArgs& a,
                         double* pos, hipStream_t st);

// merge adjacent radices into the composite ones (16, 20, 25) the wide LDS
// FFT passes handle (fewer passes, fewer barriers)
inline void merge_radices(SynthArgs& a) {
  int k = 0;
  for (int s = 0; s < a.nrad; ++s) {
    const int m = (s + 1 < a.nrad) ? a.rad[s] * a.rad[s + 1] : 0;
    if (m == 16 || m == 20 || m == 25) {
      a.rad[k++] = m;
      ++s;
    } else {
      a.rad[k++] = a.rad[s];
    }
  }
  a.nrad = k;
}

inline int wide_radix(int R) { return R == 16 || R == 20 || R == 25; }

int synth_lines_per_block(int prec, const SynthArgs& a);
hipError_t launch_synth_half(const SynthArgs& a, const BinEdges& E, const PolySpec& W, const void* D, void* V,
                             signed char* bins, void* tw, void* work, void* out, int nslots, int slot0,
                             int reuse_cube, hipStream_t st);
hipError_t launch_twiddles(int n, void* tw, hipStream_t st);
hipError_t launch_c2r400(int batch, void* in, void* out, const void* tw, hipStream_t st);
size_t zgemm_lds_bytes(int n, int nlines);
int zgemm_supported(const SynthArgs& a, int na, int nb, int same, size_t lds_limit);
int zgemm_blocks_per_cu(int same);
hipError_t launch_zgemm(const SynthArgs& a, int Q, const void* HA, int na, int sa_slots, const int* qa,
                        const void* HB, int nb, int sb_slots, const int* qb, int same, const float* G, const void* tw,
                        int nblk, double* partial, double* C, double* Cbb, hipStream_t st);
hipError_t launch_gemm_finalize(const double* partial, int nblk, int a0, int b0, int na, int nb, double inv_scale,
                                double* C, int ldc, hipStream_t st);
hipError_t launch_synth(int prec, const SynthArgs& a, const BinEdges& E, const PolySpec& W, const void* D, void* V,
                        signed char* bins, void* tw, void* work, void* out, hipStream_t st);

int paint_tiled_ok(int n);
size_t paint_tiled_work_bytes(int64_t np, int n, int nbatch);
hipError_t launch_paint_tiled(const double* pos, const double* w, const int64_t* off, int nbatch, int64_t np,
                              int64_t np_max, int n, double boxsize, int order, const double* center, int ncopy,
                              int prec, void* m0, void* m1, void* work, hipStream_t st);
hipError_t launch_paint(const double* pos, const double* w, int64_t np, int n, double boxsize, int order,
                        const double* center, int ncopy, int prec, void* m0, void* m1, hipStream_t st);
hipError_t launch_paint_batch(const double* pos, const double* w, const int64_t* off, int nb, int64_t np_max, int n,
                              double boxsize, int order, const double* center, int ncopy, int prec, void* m0, void* m1,
                              hipStream_t st);
hipError_t launch_prepare(int prec, const PrepareArgs& a, const void* c1, const void* c2, void* out, void* tables,
                          hipStream_t st);
hipError_t launch_bin(int prec, const BinArgs& a, const BinEdges& E, const PolySet& P, const void* A, const void* B,
                      double* out, hipStream_t st);
hipError_t launch_shell_filter(int prec, const ShellArgs& a, const BinEdges& E, const PolySpec& W, const void* D,
                               void* out, hipStream_t st);
hipError_t launch_cross_spectrum(int prec, const BinArgs& a, const void* A, const void* B, void* out, hipStream_t st);
hipError_t launch_resample(int prec, int n, int m, int nmax, const void* in, void* out, hipStream_t st);
hipError_t launch_triple_dot(int prec, const void* G, const void* A, const void* B, int64_t n, double* partial,
                             int nblocks, double* out, hipStream_t st);
hipError_t launch_gemm(int mode, const void* G, const void* A, int64_t lda, int na, const void* B, int64_t ldb, int nb,
                       int same, int64_t K, float sa, float sb, int nblk, double* partial, double* C, hipStream_t st);

}  // namespace cnpc
