// Full-grid kernels of the survey-geometry estimators.
//
// Survey fields are kept as full N^3 complex meshes (x-space and k-space),
// as the reference keeps them on pmesh complex128 c2c meshes: the interlaced
// field c = c1/2 + c2/2 exp(+i k.H/2) is not Hermitian on the Nyquist planes,
// so its inverse transform has an imaginary part, and the x-space Y_lm
// products of the survey estimators carry that part into every k bin.
// Full complex meshes reproduce it exactly.  Layout: index (ix*N + iy)*N + iz,
// x (or k_x) slowest, wrapped frequencies as in cnpc_common.h.
//
//   full_spectrum  half spectra c1 [, c2] -> full N^3 spectrum of the painted
//                  (interlaced) field, scaled and compensated
//   xweight        F(x) * w(xhat), xhat of cell centres around the observer
//   kcombine       out (+)= coef * w(khat) * [comp(k) * A op B - s * C_bk(k)]
//   full_bin       per-bin sums of A op B over the full grid
//   triple_sum     sum_x A*B*C (complex, FP64 partials)
//   split          complex mesh -> real and imaginary meshes
//   resample_full  full spectrum on N^3 -> its |n_d| <= nmax modes on M^3
//   reflect        c1 A(k) + c2 conj(A(-k))
//   particle_sums  FP64 catalog sums (alpha, I22, I33, N0, S_LM)
//   pair_weights   (W * W_fkp)^p * scale per object
//   radecz         RA/DEC/Z -> comoving Cartesian positions (flat LCDM)
#include "cnpc_common.h"

namespace cnpc {

template <typename T>
using C2 = typename cplx<T>::type;

__device__ __forceinline__ void full_index(int64_t idx, int n, int* ix, int* iy, int* iz) {
  *iz = (int)(idx % n);
  const int64_t r = idx / n;
  *iy = (int)(r % n);
  *ix = (int)(r / n);
}

// Value of a real field's spectrum at full-grid (ix, iy, iz) from its
// rocFFT half spectrum: c(k) for iz <= N/2, conj(c(-k)) above.
template <typename T>
__device__ __forceinline__ double2 half_at(const C2<T>* __restrict__ h, int n, int ix, int iy, int iz) {
  const int nzh = n / 2 + 1;
  if (iz <= n / 2) {
    const C2<T> v = h[((int64_t)ix * n + iy) * nzh + iz];
    return make_double2((double)v.x, (double)v.y);
  }
  const int jx = ix == 0 ? 0 : n - ix, jy = iy == 0 ? 0 : n - iy, jz = n - iz;
  const C2<T> v = h[((int64_t)jx * n + jy) * nzh + jz];
  return make_double2((double)v.x, -(double)v.y);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ------------------------------------------------------------- full spectrum
template <typename T>
__global__ void __launch_bounds__(256) full_spectrum_kernel(FullSpecArgs a, const C2<T>* __restrict__ c1,
                                                            const C2<T>* __restrict__ c2, C2<T>* __restrict__ out) {
  const int n = a.n;
  const int64_t total = (int64_t)n * n * n;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    int ix, iy, iz;
    full_index(idx, n, &ix, &iy, &iz);
    const int nx = wrap_freq(ix, n), ny = wrap_freq(iy, n), nz = wrap_freq(iz, n);
    double2 v = half_at<T>(c1, n, ix, iy, iz);
    if (a.interlaced) {
      // c = c1/2 + c2/2 * exp(+i/2 sum_d k_d H)   (mesh_generator.py:212-218)
      const double kH = (double)nx * a.kf * a.H + (double)ny * a.kf * a.H + (double)nz * a.kf * a.H;
      double sn, cs;
      sincos(0.5 * kH, &sn, &cs);
      const double2 w = half_at<T>(c2, n, ix, iy, iz);
      v.x = v.x * 0.5 + 0.5 * (w.x * cs - w.y * sn);
      v.y = v.y * 0.5 + 0.5 * (w.x * sn + w.y * cs);
    }
    const double f = a.scale * comp_factor(a.comp, a.p, nx, ny, nz, n);
    C2<T> o;
    o.x = (T)(v.x * f);
    o.y = (T)(v.y * f);
    out[idx] = o;
  }
}

// ------------------------------------------------------------- x weight
// xhat = (x_wrapped + H/2 + boxcenter) / |.|   (math_evaluator.py:25-32)
template <typename T>
__global__ void __launch_bounds__(256) xweight_kernel(XWeightArgs a, PolySpec W, const C2<T>* __restrict__ F,
                                                      C2<T>* __restrict__ out) {
  const int n = a.n;
  const int64_t total = (int64_t)n * n * n;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    int ix, iy, iz;
    full_index(idx, n, &ix, &iy, &iz);
    const double x = __dadd_rn(__dmul_rn((double)wrap_freq(ix, n), a.H), a.off[0]);
    const double y = __dadd_rn(__dmul_rn((double)wrap_freq(iy, n), a.H), a.off[1]);
    const double z = __dadd_rn(__dmul_rn((double)wrap_freq(iz, n), a.H), a.off[2]);
    const double r = __dsqrt_rn(__dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)), __dmul_rn(z, z)));
    const double2 w = poly_eval(W, x / r, y / r, z / r, false);
    const C2<T> f = F[idx];
    const double2 v = cmul(make_double2((double)f.x, (double)f.y), w);
    C2<T> o;
    o.x = (T)v.x;
    o.y = (T)v.y;
    out[idx] = o;
  }
}

// ------------------------------------------------------------- k combine
template <typename T>
__global__ void __launch_bounds__(256) kcombine_kernel(KCombineArgs a, PolySpec P, const C2<T>* __restrict__ A,
                                                       const C2<T>* __restrict__ B, C2<T>* out) {
  const int n = a.n;
  const int64_t total = (int64_t)n * n * n;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    int ix, iy, iz;
    full_index(idx, n, &ix, &iy, &iz);
    const int nx = wrap_freq(ix, n), ny = wrap_freq(iy, n), nz = wrap_freq(iz, n);
    const C2<T> av = A[idx];
    double2 v = make_double2((double)av.x, (double)av.y);
    if (a.opb) {
      const C2<T> bv = B[idx];
      const double2 b = make_double2((double)bv.x, a.opb == 2 ? -(double)bv.y : (double)bv.y);
      v = cmul(v, b);
    }
    if (a.comp) {
      const double f = comp_factor(a.comp, a.p, nx, ny, nz, n);
      v.x *= f;
      v.y *= f;
    }
    if (a.shot) {
      const double g = bk_noise_factor(a.p, nx, ny, nz, n);
      v.x -= a.s_re * g;
      v.y -= a.s_im * g;
    }
    if (a.poly) {
      double kx, ky, kz;
      const double k = knorm_of(nx, ny, nz, a.kf, &kx, &ky, &kz);
      const bool k0 = (nx == 0 && ny == 0 && nz == 0);
      const double2 w = k0 ? poly_eval(P, 0.0, 0.0, 0.0, true) : poly_eval(P, kx / k, ky / k, kz / k, false);
      v = cmul(v, w);
    }
    v = cmul(make_double2(a.c_re, a.c_im), v);
    if (a.acc) {
      const C2<T> o0 = out[idx];
      v.x += (double)o0.x;
      v.y += (double)o0.y;
    }
    C2<T> o;
    o.x = (T)v.x;
    o.y = (T)v.y;
    out[idx] = o;
  }
}

// ------------------------------------------------------------- full-grid binning
// out (FP64, caller-zeroed): [re(nb), im(nb)] of sum_bin A op B
template <typename T>
__global__ void __launch_bounds__(256) full_bin_kernel(int n, int opb, double kf, BinEdges E,
                                                       const C2<T>* __restrict__ A, const C2<T>* __restrict__ B,
                                                       double* __restrict__ out) {
  extern __shared__ double lds[];
  const int nb = E.nbins;
  for (int i = threadIdx.x; i < 2 * nb; i += blockDim.x) lds[i] = 0.0;
  __syncthreads();
  const int64_t total = (int64_t)n * n * n;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    int ix, iy, iz;
    full_index(idx, n, &ix, &iy, &iz);
    const int nx = wrap_freq(ix, n), ny = wrap_freq(iy, n), nz = wrap_freq(iz, n);
    double kx, ky, kz;
    const int b = bin_of(E, knorm_of(nx, ny, nz, kf, &kx, &ky, &kz));
    if (b < 0) continue;
    const C2<T> av = A[idx];
    const double ar = (double)av.x, ai = (double)av.y;
    double re = ar, im = ai;
    if (opb) {
      const C2<T> bv = B[idx];
      const double br = (double)bv.x, bi = opb == 2 ? -(double)bv.y : (double)bv.y;
      // unfused products: A * conj(A) has an exactly zero imaginary part
      re = __dsub_rn(__dmul_rn(ar, br), __dmul_rn(ai, bi));
      im = __dadd_rn(__dmul_rn(ar, bi), __dmul_rn(ai, br));
    }
    atomicAdd(&lds[b], re);
    atomicAdd(&lds[nb + b], im);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 2 * nb; i += blockDim.x)
    if (lds[i] != 0.0) unsafeAtomicAdd(&out[i], lds[i]);
}

// ------------------------------------------------------------- triple sum
// partial[block] = (re, im) of sum_x A * B * C over the block's cells; B, C
// may be null (factor 1).  Deterministic: fixed grid-stride order and a
// fixed-order final reduction.
template <typename T>
__global__ void __launch_bounds__(256) triple_sum_kernel(int64_t count, const C2<T>* __restrict__ A,
                                                         const C2<T>* __restrict__ B, const C2<T>* __restrict__ C,
                                                         double2* __restrict__ partial) {
  double sr = 0.0, si = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const C2<T> av = A[i];
    double2 v = make_double2((double)av.x, (double)av.y);
    if (B) {
      const C2<T> bv = B[i];
      v = cmul(v, make_double2((double)bv.x, (double)bv.y));
    }
    if (C) {
      const C2<T> cv = C[i];
      v = cmul(v, make_double2((double)cv.x, (double)cv.y));
    }
    sr += v.x;
    si += v.y;
  }
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_down(sr, off, 64);
    si += __shfl_down(si, off, 64);
  }
  __shared__ double red[2][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wv] = sr;
    red[1][wv] = si;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    partial[blockIdx.x] = make_double2((red[0][0] + red[0][1]) + (red[0][2] + red[0][3]),
                                       (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]));
}

__global__ void __launch_bounds__(64) reduce_pairs_kernel(const double2* __restrict__ partial, int np,
                                                          double* __restrict__ out) {
  double sr = 0.0, si = 0.0;
  for (int i = threadIdx.x; i < np; i += 64) {
    sr += partial[i].x;
    si += partial[i].y;
  }
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_down(sr, off, 64);
    si += __shfl_down(si, off, 64);
  }
  if (threadIdx.x == 0) {
    out[0] = sr;
    out[1] = si;
  }
}

// ------------------------------------------------------------- split / resample
template <typename T>
__global__ void __launch_bounds__(256) split_kernel(int64_t count, const C2<T>* __restrict__ src, T* __restrict__ re,
                                                    T* __restrict__ im) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const C2<T> v = src[i];
    re[i] = v.x;
    im[i] = v.y;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) resample_full_kernel(int n, int m, int nmax, const C2<T>* __restrict__ in,
                                                            C2<T>* __restrict__ out) {
  const int64_t total = (int64_t)m * m * m;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    int jx, jy, jz;
    full_index(idx, m, &jx, &jy, &jz);
    const int fx = wrap_freq(jx, m), fy = wrap_freq(jy, m), fz = wrap_freq(jz, m);
    C2<T> o;
    o.x = (T)0;
    o.y = (T)0;
    if (abs(fx) <= nmax && abs(fy) <= nmax && abs(fz) <= nmax) {
      const int ix = fx < 0 ? fx + n : fx, iy = fy < 0 ? fy + n : fy, iz = fz < 0 ? fz + n : fz;
      o = in[((int64_t)ix * n + iy) * n + iz];
    }
    out[idx] = o;
  }
}

// ------------------------------------------------------------- reflection
// out(k) = c1 * A(k) + c2 * conj(A(-k)), -k the index map i -> (-i) mod N on
// every axis (space_inversion_transposed_complex, math_evaluator.py:806-977).
template <typename T>
__global__ void __launch_bounds__(256) reflect_kernel(int n, double c1, double c2, const C2<T>* __restrict__ A,
                                                      C2<T>* __restrict__ out) {
  const int64_t total = (int64_t)n * n * n;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    int ix, iy, iz;
    full_index(idx, n, &ix, &iy, &iz);
    const int jx = ix == 0 ? 0 : n - ix, jy = iy == 0 ? 0 : n - iy, jz = iz == 0 ? 0 : n - iz;
    const C2<T> a = A[idx];
    const C2<T> b = A[((int64_t)jx * n + jy) * n + jz];
    C2<T> o;
    o.x = (T)(c1 * (double)a.x + c2 * (double)b.x);
    o.y = (T)(c1 * (double)a.y - c2 * (double)b.y);
    out[idx] = o;
  }
}

// ------------------------------------------------------------- catalog sums
// Per-object terms (mesh_generator.py:1131-1268):
//   s0 W*Wfkp, s1 W, s2 W*NZ*Wfkp*Wfkp, s3 (W*Wfkp)^2, s4 NZ^2*W^2*Wfkp^3,
//   then (re, im) of (W*Wfkp)^3 * w_M(xhat) for each of the npoly weights
//   (the conjugated Racah y_LM, M = 0..L).
__global__ void __launch_bounds__(256) particle_sums_kernel(int64_t np, const double* __restrict__ pos,
                                                            const double* __restrict__ W,
                                                            const double* __restrict__ WF,
                                                            const double* __restrict__ NZ, PolySet P, int npoly,
                                                            double* __restrict__ partial) {
  double v[CNPC_PSUM_MAX];
#pragma unroll
  for (int j = 0; j < CNPC_PSUM_MAX; ++j) v[j] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
    const double w = W[i], f = WF[i], nz = NZ ? NZ[i] : 1.0;
    const double wf = w * f;
    v[0] += wf;
    v[1] += w;
    v[2] += ((w * nz) * f) * f;
    v[3] += wf * wf;
    v[4] += ((nz * nz) * (w * w)) * (f * f * f);
    if (npoly) {
      const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
      const double r = __dsqrt_rn(__dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)), __dmul_rn(z, z)));
      const double w3 = wf * wf * wf;
#pragma unroll
      for (int j = 0; j < CNPC_MAX_POLY; ++j) {
        if (j < npoly) {
          const double2 yv = poly_eval(P.s[j], x / r, y / r, z / r, false);
          v[5 + 2 * j] += w3 * yv.x;
          v[6 + 2 * j] += w3 * yv.y;
        }
      }
    }
  }
  __shared__ double red[4][CNPC_PSUM_MAX];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < CNPC_PSUM_MAX; ++j) {
    double s = v[j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) red[wv][j] = s;
  }
  __syncthreads();
  const int nv = 5 + 2 * npoly;
  if ((int)threadIdx.x < nv) {
    const int j = threadIdx.x;
    partial[(int64_t)blockIdx.x * nv + j] = (red[0][j] + red[1][j]) + (red[2][j] + red[3][j]);
  }
}

__global__ void __launch_bounds__(64) reduce_cols_kernel(const double* __restrict__ partial, int nblk, int nv,
                                                         double* __restrict__ out) {
  const int j = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 64) s += partial[(int64_t)b * nv + j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (threadIdx.x == 0) out[j] = s;
}

__global__ void __launch_bounds__(256) pair_weights_kernel(int64_t np, const double* __restrict__ W,
                                                           const double* __restrict__ WF, int power, double scale,
                                                           double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
    double w = W[i] * (WF ? WF[i] : 1.0);
    if (power == 2) w = w * w;
    out[i] = w * scale;
  }
}

// ------------------------------------------------------------- RA/DEC/Z
// chi(z) = dh * int_0^z dz'/E(z') by Gauss-Legendre quadrature on [0, z];
// x = chi cos(dec) cos(ra), y = chi cos(dec) sin(ra), z = chi sin(dec)
// (catalog_processor.py:589-645).  Flat FLRW: E^2 = q^3 (Om + Or(q) q) + Ode
// with q = 1 + z and Or = Og (1 + nu(q)), nu = nupre (sum_i f(y_i / q) +
// nless), f(y) = (1 + (0.3173 y)^1.83)^(1/1.83) (the massive-neutrino
// density fit of Komatsu et al. 2011, ApJS 192, 18, eq. 26, as astropy
// evaluates it); Og = 0 is flat LCDM without radiation.
__global__ void __launch_bounds__(256) radecz_kernel(int64_t np, const double* __restrict__ ra,
                                                     const double* __restrict__ dec, const double* __restrict__ zz,
                                                     RadeczArgs a, double* __restrict__ pos) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
    const double z = zz[i];
    const double half = 0.5 * z;
    double s = 0.0;
    for (int j = 0; j < a.nodes; ++j) {
      const double zp = half * (a.t[j] + 1.0);
      const double q = 1.0 + zp;
      double e2 = a.om * (q * q * q) + a.ode;
      if (a.og > 0.0) {
        double nu = a.nless;
        for (int m = 0; m < a.nmassive; ++m) nu += pow(1.0 + pow(0.3173 * (a.y[m] / q), 1.83), 0.54644808743);
        e2 += a.og * (1.0 + a.nupre * nu) * (q * q * q * q);
      }
      s += a.w[j] / sqrt(e2);
    }
    const double chi = a.dh * (half * s);
    const double r = ra[i] * (M_PI / 180.0), d = dec[i] * (M_PI / 180.0);
    double sr, cr, sd, cd;
    sincos(r, &sr, &cr);
    sincos(d, &sd, &cd);
    pos[3 * i] = chi * cd * cr;
    pos[3 * i + 1] = chi * cd * sr;
    pos[3 * i + 2] = chi * sd;
  }
}

// ------------------------------------------------------------- launchers
static dim3 grid_n(int64_t total) { return dim3(grid_for(total, 256, 1 << 16)); }

hipError_t launch_full_spectrum(int prec, const FullSpecArgs& a, const void* c1, const void* c2, void* out,
                                hipStream_t st) {
  const int64_t total = (int64_t)a.n * a.n * a.n;
  if (prec == 0)
    hipLaunchKernelGGL(full_spectrum_kernel<double>, grid_n(total), dim3(256), 0, st, a, (const double2*)c1,
                       (const double2*)c2, (double2*)out);
  else
    hipLaunchKernelGGL(full_spectrum_kernel<float>, grid_n(total), dim3(256), 0, st, a, (const float2*)c1,
                       (const float2*)c2, (float2*)out);
  return hipGetLastError();
}

hipError_t launch_xweight(int prec, const XWeightArgs& a, const PolySpec& W, const void* F, void* out,
                          hipStream_t st) {
  const int64_t total = (int64_t)a.n * a.n * a.n;
  if (prec == 0)
    hipLaunchKernelGGL(xweight_kernel<double>, grid_n(total), dim3(256), 0, st, a, W, (const double2*)F,
                       (double2*)out);
  else
    hipLaunchKernelGGL(xweight_kernel<float>, grid_n(total), dim3(256), 0, st, a, W, (const float2*)F,
                       (float2*)out);
  return hipGetLastError();
}

hipError_t launch_kcombine(int prec, const KCombineArgs& a, const PolySpec& P, const void* A, const void* B, void* out,
                           hipStream_t st) {
  const int64_t total = (int64_t)a.n * a.n * a.n;
  if (prec == 0)
    hipLaunchKernelGGL(kcombine_kernel<double>, grid_n(total), dim3(256), 0, st, a, P, (const double2*)A,
                       (const double2*)B, (double2*)out);
  else
    hipLaunchKernelGGL(kcombine_kernel<float>, grid_n(total), dim3(256), 0, st, a, P, (const float2*)A,
                       (const float2*)B, (float2*)out);
  return hipGetLastError();
}

hipError_t launch_full_bin(int prec, int n, int opb, double kf, const BinEdges& E, const void* A, const void* B,
                           double* out, hipStream_t st) {
  const int64_t total = (int64_t)n * n * n;
  const size_t lds = 2 * E.nbins * sizeof(double);
  dim3 g(grid_for(total, 256, 2048));
  if (prec == 0)
    hipLaunchKernelGGL(full_bin_kernel<double>, g, dim3(256), lds, st, n, opb, kf, E, (const double2*)A,
                       (const double2*)B, out);
  else
    hipLaunchKernelGGL(full_bin_kernel<float>, g, dim3(256), lds, st, n, opb, kf, E, (const float2*)A,
                       (const float2*)B, out);
  return hipGetLastError();
}

int triple_sum_blocks(int64_t count) { return (int)grid_for(count, 256, 1024); }

hipError_t launch_triple_sum(int prec, int64_t count, const void* A, const void* B, const void* C, void* partial,
                             double* out, hipStream_t st) {
  const int nblk = triple_sum_blocks(count);
  if (prec == 0)
    hipLaunchKernelGGL(triple_sum_kernel<double>, dim3(nblk), dim3(256), 0, st, count, (const double2*)A,
                       (const double2*)B, (const double2*)C, (double2*)partial);
  else
    hipLaunchKernelGGL(triple_sum_kernel<float>, dim3(nblk), dim3(256), 0, st, count, (const float2*)A,
                       (const float2*)B, (const float2*)C, (double2*)partial);
  hipLaunchKernelGGL(reduce_pairs_kernel, dim3(1), dim3(64), 0, st, (const double2*)partial, nblk, out);
  return hipGetLastError();
}

hipError_t launch_split(int prec, int64_t count, const void* src, void* re, void* im, hipStream_t st) {
  if (prec == 0)
    hipLaunchKernelGGL(split_kernel<double>, grid_n(count), dim3(256), 0, st, count, (const double2*)src, (double*)re,
                       (double*)im);
  else
    hipLaunchKernelGGL(split_kernel<float>, grid_n(count), dim3(256), 0, st, count, (const float2*)src, (float*)re,
                       (float*)im);
  return hipGetLastError();
}

hipError_t launch_resample_full(int prec, int n, int m, int nmax, const void* in, void* out, hipStream_t st) {
  const int64_t total = (int64_t)m * m * m;
  if (prec == 0)
    hipLaunchKernelGGL(resample_full_kernel<double>, grid_n(total), dim3(256), 0, st, n, m, nmax,
                       (const double2*)in, (double2*)out);
  else
    hipLaunchKernelGGL(resample_full_kernel<float>, grid_n(total), dim3(256), 0, st, n, m, nmax,
                       (const float2*)in, (float2*)out);
  return hipGetLastError();
}

hipError_t launch_reflect(int prec, int n, double c1, double c2, const void* A, void* out, hipStream_t st) {
  const int64_t total = (int64_t)n * n * n;
  if (prec == 0)
    hipLaunchKernelGGL(reflect_kernel<double>, grid_n(total), dim3(256), 0, st, n, c1, c2, (const double2*)A,
                       (double2*)out);
  else
    hipLaunchKernelGGL(reflect_kernel<float>, grid_n(total), dim3(256), 0, st, n, c1, c2, (const float2*)A,
                       (float2*)out);
  return hipGetLastError();
}

int particle_sums_blocks(int64_t np) { return (int)grid_for(np, 256, 512); }

hipError_t launch_particle_sums(int64_t np, const double* pos, const double* W, const double* WF, const double* NZ,
                                const PolySet& P, int npoly, double* partial, double* out, hipStream_t st) {
  const int nblk = particle_sums_blocks(np);
  const int nv = 5 + 2 * npoly;
  hipLaunchKernelGGL(particle_sums_kernel, dim3(nblk), dim3(256), 0, st, np, pos, W, WF, NZ, P, npoly, partial);
  hipLaunchKernelGGL(reduce_cols_kernel, dim3(nv), dim3(64), 0, st, (const double*)partial, nblk, nv, out);
  return hipGetLastError();
}

hipError_t launch_pair_weights(int64_t np, const double* W, const double* WF, int power, double scale, double* out,
                               hipStream_t st) {
  hipLaunchKernelGGL(pair_weights_kernel, grid_n(np), dim3(256), 0, st, np, W, WF, power, scale, out);
  return hipGetLastError();
}

hipError_t launch_radecz(int64_t np, const double* ra, const double* dec, const double* z, const RadeczArgs& a,
                         double* pos, hipStream_t st) {
  hipLaunchKernelGGL(radecz_kernel, grid_n(np), dim3(256), 0, st, np, ra, dec, z, a, pos);
  return hipGetLastError();
}

}  // namespace cnpc
