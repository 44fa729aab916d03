// Half-spectrum kernels: interlace combine + window compensation, radial
// binning (mode counts, sum |k|, binned cross spectra times polynomial
// weights), and the fused shell filter + Ylm weight that feeds the shell
// inverse FFTs.
//
// All kernels walk the N*N*(N/2+1) half spectrum of a real mesh.  A mode with
// 0 < iz < N/2 stands for itself and its mirror -k, so sums over the full
// grid (which the reference forms on complex128 c2c meshes) become weighted
// sums with weight 2 for those modes and 1 on the iz = 0 and iz = N/2 planes.
// Every quantity binned by the reference is either f(-k) = conj(f(k)) or
// f(-k) = -conj(f(k)); the host picks the real or the imaginary total.
#include "cnpc_common.h"

namespace cnpc {

// ------------------------------------------------------------- prepare
// Interlace combine, scaling and the compensation comp_factor() codes.

// One axis' share of the compensation and of the interlacing phase, for
// every index i (wrapped frequency n_d): comp(n_d) with comp_factor =
// comp(nx) comp(ny) comp(nz), and e^{i/2 n_d kf H}, the factors of the
// phase e^{+i/2 sum_d k_d H} (mesh_generator.py:212-218).
__global__ void prep_axis_kernel(PrepareArgs a, double* __restrict__ comp1d, double2* __restrict__ ph1d) {
  const int n = a.n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int nd = wrap_freq(i, n);
    comp1d[i] = comp_factor(a.comp, a.p, nd, 0, 0, n);
    double sn, cs;
    sincos(0.5 * ((double)nd * a.kf * a.H), &sn, &cs);
    ph1d[i] = make_double2(cs, sn);
  }
}

// one workgroup per (ix, iy) row of the half spectrum; the per-axis factors
// come from the tables above (cached), so the kernel streams at HBM rate
template <typename T>
__global__ void __launch_bounds__(256) prepare_kernel(PrepareArgs a, const double* __restrict__ comp1d,
                                                      const double2* __restrict__ ph1d,
                                                      const typename cplx<T>::type* __restrict__ c1,
                                                      const typename cplx<T>::type* __restrict__ c2,
                                                      typename cplx<T>::type* __restrict__ out) {
  const int n = a.n, nzh = n / 2 + 1;
  const int64_t row = blockIdx.x;
  const int ix = (int)(row / n), iy = (int)(row - (int64_t)ix * n);
  const double cxy = comp1d[ix] * comp1d[iy];
  const double2 px = ph1d[ix], py = ph1d[iy];
  const double pxr = px.x * py.x - px.y * py.y, pxi = px.x * py.y + px.y * py.x;
  const bool row0 = (ix == 0 && iy == 0);
  for (int iz = threadIdx.x; iz < nzh; iz += blockDim.x) {
    const int64_t idx = row * nzh + iz;
    double re = (double)c1[idx].x, im = (double)c1[idx].y;
    if (a.interlaced) {
      // c = c1/2 + c2/2 * exp(+i/2 sum_d k_d H)   (mesh_generator.py:212-218)
      const double2 pz = ph1d[iz];
      const double cs = pxr * pz.x - pxi * pz.y, sn = pxr * pz.y + pxi * pz.x;
      const double r2 = (double)c2[idx].x, i2 = (double)c2[idx].y;
      re = re * 0.5 + 0.5 * (r2 * cs - i2 * sn);
      im = im * 0.5 + 0.5 * (r2 * sn + i2 * cs);
    }
    re *= a.scale;
    im *= a.scale;
    const bool k0 = row0 && iz == 0;
    if (k0) re -= a.sub0;
    const double f = cxy * comp1d[iz];
    re *= f;
    im *= f;
    if (k0 && a.zero0) { re = 0.0; im = 0.0; }
    typename cplx<T>::type o;
    o.x = (T)re;
    o.y = (T)im;
    out[idx] = o;
  }
}

hipError_t launch_prepare(int prec, const PrepareArgs& a, const void* c1, const void* c2, void* out, void* tables,
                          hipStream_t st) {
  double* comp1d = (double*)tables;
  double2* ph1d = (double2*)(comp1d + ((a.n + 1) & ~1));
  hipLaunchKernelGGL(prep_axis_kernel, dim3((a.n + 255) / 256), dim3(256), 0, st, a, comp1d, ph1d);
  const int64_t rows = (int64_t)a.n * a.n;
  if (prec == 0)
    hipLaunchKernelGGL(prepare_kernel<double>, dim3((unsigned)rows), dim3(256), 0, st, a, comp1d, ph1d,
                       (const double2*)c1, (const double2*)c2, (double2*)out);
  else
    hipLaunchKernelGGL(prepare_kernel<float>, dim3((unsigned)rows), dim3(256), 0, st, a, comp1d, ph1d,
                       (const float2*)c1, (const float2*)c2, (float2*)out);
  return hipGetLastError();
}

// ------------------------------------------------------------- binning
// out layout (FP64, caller-zeroed): [count(nb), ksum(nb), re(npoly*nb), im(npoly*nb)]


template <typename T>
__global__ void __launch_bounds__(256) bin_kernel(BinArgs a, BinEdges E, PolySet P,
                                                  const typename cplx<T>::type* __restrict__ A,
                                                  const typename cplx<T>::type* __restrict__ B, double* __restrict__ out) {
  extern __shared__ double lds[];  // (2 + 2*npoly) * nbins
  const int nb = E.nbins;
  const int nacc = (2 + 2 * a.npoly) * nb;
  for (int i = threadIdx.x; i < nacc; i += blockDim.x) lds[i] = 0.0;
  __syncthreads();
  const int n = a.n, nzh = n / 2 + 1;
  const int64_t total = (int64_t)n * n * nzh;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int iz = (int)(idx % nzh);
    const int64_t r = idx / nzh;
    const int iy = (int)(r % n), ix = (int)(r / n);
    const int nx = wrap_freq(ix, n), ny = wrap_freq(iy, n), nz = wrap_freq(iz, n);
    double kx, ky, kz;
    const double k = knorm_of(nx, ny, nz, a.kf, &kx, &ky, &kz);
    const int b = bin_of(E, k);
    if (b < 0) continue;
    const double wmode = (iz == 0 || iz == n / 2) ? 1.0 : 2.0;
    atomicAdd(&lds[b], wmode);
    atomicAdd(&lds[nb + b], wmode * k);
    if (a.npoly == 0) continue;
    const double ar = (double)A[idx].x, ai = (double)A[idx].y;
    const double br = (double)B[idx].x, bi = (double)B[idx].y;
    double fr, fi;
    if (a.conj_b) { fr = ar * br + ai * bi; fi = ai * br - ar * bi; }
    else { fr = ar * br - ai * bi; fi = ar * bi + ai * br; }
    fr *= a.ab_scale;
    fi *= a.ab_scale;
    if (a.noise_p > 0 || a.noise_amp != 0.0) fr -= a.noise_amp * bk_noise_factor(a.noise_p, nx, ny, nz, n);
    const bool zv = (k == 0.0);
    const double inv = zv ? 0.0 : 1.0 / k;
    for (int q = 0; q < a.npoly; ++q) {
      const double2 w = poly_eval(P.s[q], kx * inv, ky * inv, kz * inv, zv);
      atomicAdd(&lds[(2 + q) * nb + b], wmode * (fr * w.x - fi * w.y));
      atomicAdd(&lds[(2 + a.npoly + q) * nb + b], wmode * (fr * w.y + fi * w.x));
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nacc; i += blockDim.x) {
    const double v = lds[i];
    if (v != 0.0) unsafeAtomicAdd(&out[i], v);
  }
}

hipError_t launch_bin(int prec, const BinArgs& a, const BinEdges& E, const PolySet& P, const void* A, const void* B,
                      double* out, hipStream_t st) {
  const int64_t total = (int64_t)a.n * a.n * (a.n / 2 + 1);
  const size_t lds = (size_t)(2 + 2 * a.npoly) * E.nbins * sizeof(double);
  dim3 blk(256), grd(grid_for(total, 256, 2048));
  if (prec == 0)
    hipLaunchKernelGGL(bin_kernel<double>, grd, blk, lds, st, a, E, P, (const double2*)A, (const double2*)B, out);
  else
    hipLaunchKernelGGL(bin_kernel<float>, grd, blk, lds, st, a, E, P, (const float2*)A, (const float2*)B, out);
  return hipGetLastError();
}

// ------------------------------------------------------------- shell filter
// out[s] = w(khat) * delta(k) * [bin(k) == bin0 + s], s = 0..nshell-1, each a
// full half spectrum (zeros outside the shell) ready for a c2r transform:
// the filtered-field product of clustering_estimator.py:934-970 with the
// Y_lm weight of :898-902 folded in.

template <typename T>
__global__ void __launch_bounds__(256) shell_filter_kernel(ShellArgs a, BinEdges E, PolySpec W,
                                                           const typename cplx<T>::type* __restrict__ D,
                                                           typename cplx<T>::type* __restrict__ out) {
  const int n = a.n, nzh = n / 2 + 1;
  const int64_t total = (int64_t)n * n * nzh;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int iz = (int)(idx % nzh);
    const int64_t r = idx / nzh;
    const int iy = (int)(r % n), ix = (int)(r / n);
    const int nx = wrap_freq(ix, n), ny = wrap_freq(iy, n), nz = wrap_freq(iz, n);
    double kx, ky, kz;
    const double k = knorm_of(nx, ny, nz, a.kf, &kx, &ky, &kz);
    const int b = bin_of(E, k) - a.bin0;
    typename cplx<T>::type v;
    v.x = (T)0;
    v.y = (T)0;
    if (b >= 0 && b < a.nshell) {
      const bool zv = (k == 0.0);
      const double inv = zv ? 0.0 : 1.0 / k;
      const double2 w = poly_eval(W, kx * inv, ky * inv, kz * inv, zv);
      // a null spectrum filters the weight alone (the S3 shells of
      // clustering_estimator.py:1213-1217)
      const double dr = D ? (double)D[idx].x : 1.0, di = D ? (double)D[idx].y : 0.0;
      v.x = (T)(dr * w.x - di * w.y);
      v.y = (T)(dr * w.y + di * w.x);
    }
    for (int s = 0; s < a.nshell; ++s) {
      typename cplx<T>::type o;
      o.x = (s == b) ? v.x : (T)0;
      o.y = (s == b) ? v.y : (T)0;
      out[s * a.stride + idx] = o;
    }
  }
}

// ------------------------------------------------------------- resample
// Copies the modes with |n_x|, |n_y|, n_z <= nmax of an N-grid half spectrum
// into an M-grid half spectrum (zeros elsewhere).  With nmax < M/2 every
// copied mode keeps its physical k, so an M-point c2r evaluates the same
// band-limited field on the coarser M-grid.
template <typename T>
__global__ void __launch_bounds__(256) resample_kernel(int n, int m, int nmax, const typename cplx<T>::type* __restrict__ in,
                                                       typename cplx<T>::type* __restrict__ out) {
  const int mzh = m / 2 + 1, nzh = n / 2 + 1;
  const int64_t total = (int64_t)m * m * mzh;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int iz = (int)(idx % mzh);
    const int64_t r = idx / mzh;
    const int iy = (int)(r % m), ix = (int)(r / m);
    const int nx = wrap_freq(ix, m), ny = wrap_freq(iy, m), nz = wrap_freq(iz, m);
    typename cplx<T>::type v;
    v.x = (T)0;
    v.y = (T)0;
    if (nx <= nmax && nx >= -nmax && ny <= nmax && ny >= -nmax && nz >= 0 && nz <= nmax) {
      const int jx = nx < 0 ? nx + n : nx, jy = ny < 0 ? ny + n : ny;
      v = in[((int64_t)jx * n + jy) * nzh + nz];
    }
    out[idx] = v;
  }
}

hipError_t launch_resample(int prec, int n, int m, int nmax, const void* in, void* out, hipStream_t st) {
  const int64_t total = (int64_t)m * m * (m / 2 + 1);
  dim3 blk(256), grd(grid_for(total, 256, 1 << 16));
  if (prec == 0)
    hipLaunchKernelGGL(resample_kernel<double>, grd, blk, 0, st, n, m, nmax, (const double2*)in, (double2*)out);
  else
    hipLaunchKernelGGL(resample_kernel<float>, grd, blk, 0, st, n, m, nmax, (const float2*)in, (float2*)out);
  return hipGetLastError();
}

// ------------------------------------------------------------- cross spectrum
// out = ab_scale * a conj(b) - noise_amp * C_p/sinc^{2p}(k): the S3 source
// spectrum of clustering_estimator.py:1197-1211 (before its c2r).
template <typename T>
__global__ void __launch_bounds__(256) cross_spectrum_kernel(BinArgs a, const typename cplx<T>::type* __restrict__ A,
                                                             const typename cplx<T>::type* __restrict__ B,
                                                             typename cplx<T>::type* __restrict__ out) {
  const int n = a.n, nzh = n / 2 + 1;
  const int64_t total = (int64_t)n * n * nzh;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int iz = (int)(idx % nzh);
    const int64_t r = idx / nzh;
    const int iy = (int)(r % n), ix = (int)(r / n);
    const int nx = wrap_freq(ix, n), ny = wrap_freq(iy, n), nz = wrap_freq(iz, n);
    const double ar = (double)A[idx].x, ai = (double)A[idx].y;
    const double br = (double)B[idx].x, bi = (double)B[idx].y;
    double fr = (ar * br + ai * bi) * a.ab_scale;
    double fi = (ai * br - ar * bi) * a.ab_scale;
    if (a.noise_amp != 0.0) fr -= a.noise_amp * bk_noise_factor(a.noise_p, nx, ny, nz, n);
    typename cplx<T>::type o;
    o.x = (T)fr;
    o.y = (T)fi;
    out[idx] = o;
  }
}

hipError_t launch_cross_spectrum(int prec, const BinArgs& a, const void* A, const void* B, void* out, hipStream_t st) {
  const int64_t total = (int64_t)a.n * a.n * (a.n / 2 + 1);
  dim3 blk(256), grd(grid_for(total, 256, 1 << 16));
  if (prec == 0)
    hipLaunchKernelGGL(cross_spectrum_kernel<double>, grd, blk, 0, st, a, (const double2*)A, (const double2*)B,
                       (double2*)out);
  else
    hipLaunchKernelGGL(cross_spectrum_kernel<float>, grd, blk, 0, st, a, (const float2*)A, (const float2*)B,
                       (float2*)out);
  return hipGetLastError();
}

hipError_t launch_shell_filter(int prec, const ShellArgs& a, const BinEdges& E, const PolySpec& W, const void* D,
                               void* out, hipStream_t st) {
  const int64_t total = (int64_t)a.n * a.n * (a.n / 2 + 1);
  dim3 blk(256), grd(grid_for(total, 256, 1 << 16));
  if (prec == 0)
    hipLaunchKernelGGL(shell_filter_kernel<double>, grd, blk, 0, st, a, E, W, (const double2*)D, (double2*)out);
  else
    hipLaunchKernelGGL(shell_filter_kernel<float>, grd, blk, 0, st, a, E, W, (const float2*)D, (float2*)out);
  return hipGetLastError();
}

}  // namespace cnpc
