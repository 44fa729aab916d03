// Pruned shell synthesis: the inverse 3-D FFT of a shell-filtered spectrum
// split into a sparse x-direction DFT (this file) and a batched 2-D c2r over
// (y, z) (rocFFT, cnpc_api.hip).
//
// A shell touches only modes with |k| < k_max, i.e. |n_d| <= R with
// R = ceil(k_max / kf) well below N/2.  The full 3-D c2r of the reference
// (one N^3 IFFT per shell, clustering_estimator.py:934-970) spends most of
// its passes on zeros.  Here:
//
//  1. cube_kernel: V[ky][kz][kx] = w(khat) delta(k) for |n_x|,|n_y| <= R,
//     0 <= n_z <= R, plus the bin of every mode -- (2R+1)^2 (R+1) entries,
//     tens of MB, shared by all shells of a call;
//  2. xsynth_kernel: for every shell s, plane x and line (ky, kz),
//       P_s(x, ky, kz) = sum_{kx in shell s on the line} V e^{+2 pi i kx x / N},
//     written as the half-spectrum planes [s][x][iy][iz] (zeros outside the
//     R-box) that a batched 2-D c2r turns into the real shells [s][x][y][z].
//
// On a line (ky, kz) the bin is non-decreasing in |kx|, so each shell covers
// one |kx| interval [a, b] (both signs).  Twiddles come from an LDS table
// indexed by (kx * x) mod N, advanced incrementally.
#include "cnpc_common.h"

namespace cnpc {

template <typename T>
__global__ void __launch_bounds__(256) cube_kernel(SynthArgs a, BinEdges E, PolySpec W,
                                                   const typename cplx<T>::type* __restrict__ D,
                                                   typename cplx<T>::type* __restrict__ V, signed char* __restrict__ bins) {
  const int R = a.R, P = 2 * R + 1, n = a.n, nzh = n / 2 + 1;
  const int64_t total = (int64_t)P * (R + 1) * P;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
    const int kxi = (int)(idx % P);
    const int64_t r = idx / P;
    const int nz = (int)(r % (R + 1));
    const int kyi = (int)(r / (R + 1));
    const int nx = kxi - R, ny = kyi - R;
    double kx, ky, kz;
    const double k = knorm_of(nx, ny, nz, a.kf, &kx, &ky, &kz);
    const int b = bin_of(E, k);
    typename cplx<T>::type v;
    v.x = (T)0;
    v.y = (T)0;
    if (b >= 0) {
      const bool zv = (k == 0.0);
      const double inv = zv ? 0.0 : 1.0 / k;
      const double2 w = poly_eval(W, kx * inv, ky * inv, kz * inv, zv);
      double dr = 1.0, di = 0.0;
      if (D) {
        const int jx = nx < 0 ? nx + n : nx, jy = ny < 0 ? ny + n : ny;
        const typename cplx<T>::type d = D[((int64_t)jx * n + jy) * nzh + nz];
        dr = (double)d.x;
        di = (double)d.y;
      }
      v.x = (T)(dr * w.x - di * w.y);
      v.y = (T)(dr * w.y + di * w.x);
    }
    V[idx] = v;
    bins[idx] = (signed char)(b < 0 ? -1 : b);
  }
}

// one workgroup per (iy, block of ZB consecutive iz); threads = ZB lanes x (256/ZB) x-rows
template <typename T, int ZB>
__global__ void __launch_bounds__(256) xsynth_kernel(SynthArgs a, const typename cplx<T>::type* __restrict__ V,
                                                     const signed char* __restrict__ bins,
                                                     const typename cplx<T>::type* __restrict__ twg,
                                                     typename cplx<T>::type* __restrict__ out) {
  typedef typename cplx<T>::type C;
  const int n = a.n, nzh = n / 2 + 1, R = a.R, P = 2 * R + 1;
  const int iy = blockIdx.x;
  const int iz0 = blockIdx.y * ZB;
  const int ny = wrap_freq(iy, n);
  const int lane = threadIdx.x % ZB, xrow = threadIdx.x / ZB;
  constexpr int XR = 256 / ZB;
  const int iz = iz0 + lane;
  const int64_t plane = (int64_t)n * nzh;
  const bool active_line = (ny >= -R && ny <= R);
  const bool any_k = active_line && iz0 <= R;
  if (!any_k) {  // outside the R-box: zeros for every shell and plane
    if (iz < nzh) {
      C z;
      z.x = (T)0;
      z.y = (T)0;
      for (int s = 0; s < a.nshell; ++s)
        for (int x = xrow; x < n; x += XR) out[((int64_t)s * n + x) * plane + (int64_t)iy * nzh + iz] = z;
    }
    return;
  }
  extern __shared__ unsigned char smem[];
  C* tw = (C*)smem;                    // n twiddles e^{+2 pi i j / n}
  C* lv = tw + n;                      // ZB lines x (R+1) values, kx >= 0 then kx < 0 mirrored
  C* lvn = lv + ZB * (R + 1);
  int* ab = (int*)(lvn + ZB * (R + 1));  // [ZB][nshell][2]
  for (int j = threadIdx.x; j < n; j += blockDim.x) tw[j] = twg[j];
  const int kyi = ny + R;
  for (int t = threadIdx.x; t < ZB * (R + 1); t += blockDim.x) {
    const int l = t / (R + 1), kx = t % (R + 1);
    const int nz = iz0 + l;
    C vp, vn;
    vp.x = vn.x = (T)0;
    vp.y = vn.y = (T)0;
    if (nz <= R) {
      const int64_t base = ((int64_t)kyi * (R + 1) + nz) * P;
      vp = V[base + R + kx];
      vn = V[base + R - kx];
    }
    lv[l * (R + 1) + kx] = vp;
    lvn[l * (R + 1) + kx] = vn;
  }
  // |kx| interval of every shell on every line of the block
  for (int t = threadIdx.x; t < ZB * a.nshell; t += blockDim.x) {
    const int l = t / a.nshell, s = t % a.nshell;
    const int nz = iz0 + l;
    int lo = 1, hi = 0;  // empty
    if (nz <= R) {
      const int64_t base = ((int64_t)kyi * (R + 1) + nz) * P + R;
      const int want = a.bin0 + s;
      for (int kx = 0; kx <= R; ++kx) {
        if (bins[base + kx] == want) {
          if (lo > hi) lo = kx;
          hi = kx;
        }
      }
    }
    ab[(l * a.nshell + s) * 2 + 0] = lo;
    ab[(l * a.nshell + s) * 2 + 1] = hi;
  }
  __syncthreads();
  if (iz >= nzh) return;
  const C* myv = lv + lane * (R + 1);
  const C* myvn = lvn + lane * (R + 1);
  for (int s = 0; s < a.nshell; ++s) {
    const int lo = ab[(lane * a.nshell + s) * 2 + 0], hi = ab[(lane * a.nshell + s) * 2 + 1];
    for (int x = xrow; x < n; x += XR) {
      T ar = (T)0, ai = (T)0;
      if (lo <= hi) {
        int j = (int)(((int64_t)lo * x) % n);  // (kx * x) mod n
        for (int kx = lo; kx <= hi; ++kx) {
          const C t = tw[j];
          const C v = myv[kx];
          ar += v.x * t.x - v.y * t.y;
          ai += v.x * t.y + v.y * t.x;
          if (kx != 0) {  // -kx: conj twiddle
            const C u = myvn[kx];
            ar += u.x * t.x + u.y * t.y;
            ai += u.y * t.x - u.x * t.y;
          }
          j += x;
          if (j >= n) j -= n;
        }
      }
      C o;
      o.x = ar;
      o.y = ai;
      out[((int64_t)s * n + x) * plane + (int64_t)iy * nzh + iz] = o;
    }
  }
}

template <typename T>
__global__ void twiddle_kernel(int n, typename cplx<T>::type* tw) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    double s, c;
    sincospi(2.0 * (double)j / (double)n, &s, &c);
    tw[j].x = (T)c;
    tw[j].y = (T)s;
  }
}

size_t xsynth_lds_bytes(int prec, int n, int R, int nshell, int zb) {
  const size_t cs = prec == 0 ? 16 : 8;
  return cs * (size_t)n + 2 * cs * (size_t)zb * (R + 1) + sizeof(int) * 2 * (size_t)zb * nshell;
}

template <typename T, int ZB>
static void launch_x(const SynthArgs& a, const void* V, const signed char* bins, const void* tw, void* out, size_t lds,
                     hipStream_t st) {
  typedef typename cplx<T>::type C;
  dim3 grd(a.n, (a.n / 2 + 1 + ZB - 1) / ZB);
  hipLaunchKernelGGL((xsynth_kernel<T, ZB>), grd, dim3(256), lds, st, a, (const C*)V, bins, (const C*)tw, (C*)out);
}

// zb: lines (iz values) per workgroup, 8 / 16 / 32 (chosen by the caller so
// the LDS footprint xsynth_lds_bytes fits)
hipError_t launch_synth(int prec, const SynthArgs& a, const BinEdges& E, const PolySpec& W, const void* D, void* V,
                        signed char* bins, void* tw, void* out, int zb, hipStream_t st) {
  const int R = a.R, P = 2 * R + 1, n = a.n;
  const int64_t cube = (int64_t)P * (R + 1) * P;
  dim3 blk(256);
  const size_t lds = xsynth_lds_bytes(prec, n, R, a.nshell, zb);
  if (prec == 0) {
    hipLaunchKernelGGL(twiddle_kernel<double>, dim3((n + 255) / 256), blk, 0, st, n, (double2*)tw);
    hipLaunchKernelGGL(cube_kernel<double>, dim3(grid_for(cube, 256, 1 << 14)), blk, 0, st, a, E, W, (const double2*)D,
                       (double2*)V, bins);
    if (zb == 32) launch_x<double, 32>(a, V, bins, tw, out, lds, st);
    else if (zb == 16) launch_x<double, 16>(a, V, bins, tw, out, lds, st);
    else launch_x<double, 8>(a, V, bins, tw, out, lds, st);
  } else {
    hipLaunchKernelGGL(twiddle_kernel<float>, dim3((n + 255) / 256), blk, 0, st, n, (float2*)tw);
    hipLaunchKernelGGL(cube_kernel<float>, dim3(grid_for(cube, 256, 1 << 14)), blk, 0, st, a, E, W, (const float2*)D,
                       (float2*)V, bins);
    if (zb == 32) launch_x<float, 32>(a, V, bins, tw, out, lds, st);
    else if (zb == 16) launch_x<float, 16>(a, V, bins, tw, out, lds, st);
    else launch_x<float, 8>(a, V, bins, tw, out, lds, st);
  }
  return hipGetLastError();
}

}  // namespace cnpc
