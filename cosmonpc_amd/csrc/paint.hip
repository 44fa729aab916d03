// Particle painting (mass assignment) onto periodic N^3 meshes.
//
// Follows pm_painter (mesh_generator.py:172-224): B-spline windows of order
// 1..4 (NGP/CIC/TSC/PCS) with mesh points at integer mesh coordinates
// u = (x - boxcenter) * N / L (+0.5 for the interlacing copy,
// pm.affine.shift(0.5) at mesh_generator.py:201), periodic wrap.
//
// One thread per particle paints both the plain and the shifted copy, so each
// particle's 32 bytes are read once.  Cell index and fractional offset always
// come from the FP64 position (a float32 position in a 2000 Mpc/h box would
// carry ~3e-5 cell of error); weights are formed in FP64 and rounded to the
// mesh type for the atomic add.
#include "cnpc_common.h"

namespace cnpc {

template <int ORDER>
__device__ __forceinline__ int window(double u, double* w) {
  if (ORDER == 1) {
    double i0 = floor(u + 0.5);
    w[0] = 1.0;
    return (int)i0;
  } else if (ORDER == 2) {
    double i0 = floor(u);
    double f = u - i0;
    w[0] = 1.0 - f;
    w[1] = f;
    return (int)i0;
  } else if (ORDER == 3) {
    double i1 = floor(u + 0.5);
    double d = u - i1;
    w[0] = 0.5 * (0.5 - d) * (0.5 - d);
    w[1] = 0.75 - d * d;
    w[2] = 0.5 * (0.5 + d) * (0.5 + d);
    return (int)i1 - 1;
  } else {
    double i1 = floor(u);
    double f = u - i1;
    double g = 1.0 - f;
    w[0] = g * g * g / 6.0;
    w[1] = (4.0 - 6.0 * f * f + 3.0 * f * f * f) / 6.0;
    w[2] = (4.0 - 6.0 * g * g + 3.0 * g * g * g) / 6.0;
    w[3] = f * f * f / 6.0;
    return (int)i1 - 1;
  }
}

__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }

template <typename T, int ORDER>
__device__ __forceinline__ void paint_one(const double* __restrict__ pos, const double* __restrict__ wgt, int64_t p,
                                          int n, double scale, double cx, double cy, double cz, int ncopy,
                                          T* __restrict__ mesh0, T* __restrict__ mesh1) {
  const int64_t n2 = (int64_t)n * n;
  const double x = pos[3 * p + 0] - cx;
  const double y = pos[3 * p + 1] - cy;
  const double z = pos[3 * p + 2] - cz;
  const double m = wgt[p];
  for (int c = 0; c < ncopy; ++c) {
    const double sh = 0.5 * c;
    double wx[ORDER], wy[ORDER], wz[ORDER];
    // u = x * N/L + shift, rounded like numpy (no fused multiply-add)
    int ax = window<ORDER>(__dadd_rn(__dmul_rn(x, scale), sh), wx);
    int ay = window<ORDER>(__dadd_rn(__dmul_rn(y, scale), sh), wy);
    int az = window<ORDER>(__dadd_rn(__dmul_rn(z, scale), sh), wz);
    T* mesh = c == 0 ? mesh0 : mesh1;
    ax %= n; if (ax < 0) ax += n;
    ay %= n; if (ay < 0) ay += n;
    az %= n; if (az < 0) az += n;
#pragma unroll
    for (int i = 0; i < ORDER; ++i) {
      int ix = ax + i; if (ix >= n) ix -= n;
      const double mx = m * wx[i];
#pragma unroll
      for (int j = 0; j < ORDER; ++j) {
        int iy = ay + j; if (iy >= n) iy -= n;
        const double mxy = mx * wy[j];
        T* row = mesh + ix * n2 + (int64_t)iy * n;
#pragma unroll
        for (int k = 0; k < ORDER; ++k) {
          int iz = az + k; if (iz >= n) iz -= n;
          atomic_add(row + iz, (T)(mxy * wz[k]));
        }
      }
    }
  }
}

template <typename T, int ORDER>
__global__ void __launch_bounds__(256) paint_kernel(const double* __restrict__ pos, const double* __restrict__ wgt,
                                                    int64_t np, int n, double scale, double cx, double cy, double cz,
                                                    int ncopy, T* __restrict__ mesh0, T* __restrict__ mesh1) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x)
    paint_one<T, ORDER>(pos, wgt, p, n, scale, cx, cy, cz, ncopy, mesh0, mesh1);
}

// Many catalogs in one launch: catalog b owns particles [off[b], off[b+1])
// of the concatenated arrays and paints into mesh b of each batch
// (blockIdx.y = b; the x-blocks stride over that catalog).
template <typename T, int ORDER>
__global__ void __launch_bounds__(256) paint_batch_kernel(const double* __restrict__ pos,
                                                          const double* __restrict__ wgt,
                                                          const int64_t* __restrict__ off, int n, double scale,
                                                          double cx, double cy, double cz, int ncopy,
                                                          T* __restrict__ mesh0, T* __restrict__ mesh1) {
  const int b = blockIdx.y;
  const int64_t n3 = (int64_t)n * n * n;
  const int64_t p0 = off[b], p1 = off[b + 1];
  T* m0 = mesh0 + b * n3;
  T* m1 = mesh1 ? mesh1 + b * n3 : nullptr;
  for (int64_t p = p0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < p1; p += (int64_t)gridDim.x * blockDim.x)
    paint_one<T, ORDER>(pos, wgt, p, n, scale, cx, cy, cz, ncopy, m0, m1);
}

template <typename T>
hipError_t launch_paint_t(const double* pos, const double* w, int64_t np, int n, double boxsize, int order,
                          const double* center, int ncopy, T* m0, T* m1, hipStream_t st) {
  const double scale = (double)n / boxsize;
  const double cx = center ? center[0] : 0.0, cy = center ? center[1] : 0.0, cz = center ? center[2] : 0.0;
  dim3 blk(256), grd(grid_for(np, 256, 1 << 16));
  if (np == 0) return hipSuccess;
  switch (order) {
    case 1: hipLaunchKernelGGL((paint_kernel<T, 1>), grd, blk, 0, st, pos, w, np, n, scale, cx, cy, cz, ncopy, m0, m1); break;
    case 2: hipLaunchKernelGGL((paint_kernel<T, 2>), grd, blk, 0, st, pos, w, np, n, scale, cx, cy, cz, ncopy, m0, m1); break;
    case 3: hipLaunchKernelGGL((paint_kernel<T, 3>), grd, blk, 0, st, pos, w, np, n, scale, cx, cy, cz, ncopy, m0, m1); break;
    case 4: hipLaunchKernelGGL((paint_kernel<T, 4>), grd, blk, 0, st, pos, w, np, n, scale, cx, cy, cz, ncopy, m0, m1); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_paint_batch_t(const double* pos, const double* w, const int64_t* off, int nb, int64_t np_max, int n,
                                double boxsize, int order, const double* center, int ncopy, T* m0, T* m1,
                                hipStream_t st) {
  const double scale = (double)n / boxsize;
  const double cx = center ? center[0] : 0.0, cy = center ? center[1] : 0.0, cz = center ? center[2] : 0.0;
  if (nb == 0 || np_max == 0) return hipSuccess;
  dim3 blk(256), grd(grid_for(np_max, 256, (1 << 16) / nb + 1), nb);
  switch (order) {
    case 1: hipLaunchKernelGGL((paint_batch_kernel<T, 1>), grd, blk, 0, st, pos, w, off, n, scale, cx, cy, cz, ncopy, m0, m1); break;
    case 2: hipLaunchKernelGGL((paint_batch_kernel<T, 2>), grd, blk, 0, st, pos, w, off, n, scale, cx, cy, cz, ncopy, m0, m1); break;
    case 3: hipLaunchKernelGGL((paint_batch_kernel<T, 3>), grd, blk, 0, st, pos, w, off, n, scale, cx, cy, cz, ncopy, m0, m1); break;
    case 4: hipLaunchKernelGGL((paint_batch_kernel<T, 4>), grd, blk, 0, st, pos, w, off, n, scale, cx, cy, cz, ncopy, m0, m1); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_paint_batch(const double* pos, const double* w, const int64_t* off, int nb, int64_t np_max, int n,
                              double boxsize, int order, const double* center, int ncopy, int prec, void* m0, void* m1,
                              hipStream_t st) {
  if (prec == 0)
    return launch_paint_batch_t<double>(pos, w, off, nb, np_max, n, boxsize, order, center, ncopy, (double*)m0,
                                        (double*)m1, st);
  return launch_paint_batch_t<float>(pos, w, off, nb, np_max, n, boxsize, order, center, ncopy, (float*)m0,
                                     (float*)m1, st);
}

hipError_t launch_paint(const double* pos, const double* w, int64_t np, int n, double boxsize, int order,
                        const double* center, int ncopy, int prec, void* m0, void* m1, hipStream_t st) {
  if (prec == 0)
    return launch_paint_t<double>(pos, w, np, n, boxsize, order, center, ncopy, (double*)m0, (double*)m1, st);
  return launch_paint_t<float>(pos, w, np, n, boxsize, order, center, ncopy, (float*)m0, (float*)m1, st);
}

}  // namespace cnpc
