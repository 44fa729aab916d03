// Particle painting (mass assignment) onto periodic N^3 meshes.
//
// Follows pm_painter (mesh_generator.py:172-224): B-spline windows of order
// 1..4 (NGP/CIC/TSC/PCS) with mesh points at integer mesh coordinates
// u = (x - boxcenter) * N / L (+0.5 for the interlacing copy,
// pm.affine.shift(0.5) at mesh_generator.py:201), periodic wrap.
//
// Meshes with n % 8 == 0 (16 <= n <= 4096) take the sorted, LDS-tiled
// owner-computes path below, for one catalog or a batch (bitwise
// reproducible: fixed-point tile sums); the others take the scatter path:
// one thread per particle paints both the plain and the shifted copy with
// global atomics, so each particle's 32 bytes are read once.  Cell index and
// fractional offset always come from the FP64 position (a float32 position
// in a 2000 Mpc/h box would carry ~3e-5 cell of error); weights are formed
// in FP64.
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

__global__ void fill_int_kernel(int* __restrict__ p, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
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

// ---------------------------------------------------------------- tiled paint
// Owner-computes painting (n % 8 == 0, 16 <= n <= 4096), for one catalog or
// a batch of catalogs (catalog b -> mesh b).  Tiles are PT^3 cells (PT = 16
// when n % 16 == 0, else 8).  A particle belongs to the tile of its window's
// first cell (plain copy); its class says across which upper tile faces its
// windows reach (the shifted copy starts at most one cell later and a
// window spans at most 4 cells, so at most into the next tile).  A counting
// sort orders the particles by (catalog, tile, class).  One workgroup per
// (catalog, tile) then accumulates in LDS the window weights falling inside
// its tile, from its own particles and from the particles of the seven lower
// neighbour tiles whose class reaches into it (a lower x neighbour
// contributes classes with the x bit, ...), and writes the tile once with
// plain stores (also the zero fill).  No global atomics on the mesh, and a
// particle is read by 1 + (tiles its windows reach) workgroups: about 1.9
// reads per particle at PT = 16 instead of 8.  Window weights are formed
// in FP64 per axis, rounded to the mesh type and multiplied in it; the tile
// sums them in 64-bit fixed point (below), so the result does not depend on
// the order of the additions and the mesh is bitwise reproducible.
constexpr int PAINT_TILE_THREADS = 256;

// first window cell of both copies along one axis, relative to the mesh
template <int ORDER>
__device__ __forceinline__ void first_cells(double u, int n, int& a0, int& a1) {
  double w[ORDER];
  a0 = window<ORDER>(__dadd_rn(u, 0.0), w);
  a1 = window<ORDER>(__dadd_rn(u, 0.5), w);
  a0 %= n; if (a0 < 0) a0 += n;
  a1 %= n; if (a1 < 0) a1 += n;
}

// key = (b * ntiles + tile) * 8 + class, class bits (x, y, z) = 4, 2, 1.
// Catalog b = blockIdx.y owns particles [off[b], off[b + 1]) (off == nullptr:
// one catalog of np particles).
template <int ORDER, int PT>
__global__ void __launch_bounds__(256) paint_key_kernel(const double* __restrict__ pos, const int64_t* __restrict__ off,
                                                        int64_t np, int n, double scale, double cx, double cy,
                                                        double cz, int nt, int ncopy, int* __restrict__ key,
                                                        int* __restrict__ count) {
  const int b = blockIdx.y;
  const int64_t p0 = off ? off[b] : 0, p1 = off ? off[b + 1] : np;
  const double c[3] = {cx, cy, cz};
  for (int64_t p = p0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < p1; p += (int64_t)gridDim.x * blockDim.x) {
    int t[3], cls = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      int a0, a1;
      first_cells<ORDER>(__dmul_rn(pos[3 * p + d] - c[d], scale), n, a0, a1);
      t[d] = a0 / PT;
      // last cell of either copy, relative to the tile (a1 is a0 or a0 + 1,
      // modulo n)
      const int r = a0 - t[d] * PT + (ncopy > 1 && a1 != a0 ? 1 : 0) + ORDER - 1;
      cls = 2 * cls + (r >= PT ? 1 : 0);
    }
    const int k = ((b * nt + t[0]) * nt + t[1]) * nt + t[2];
    const int kk = k * 8 + cls;
    key[p] = kk;
    atomicAdd(count + kk, 1);
  }
}

// exclusive prefix sum of count[0 .. m) into start[0 .. m] in three steps:
// each workgroup scans 1024 counts (4 per thread) and leaves its total in
// bsum; one workgroup scans the totals; each workgroup adds its offset
constexpr int SCAN_TILE = 1024;

__global__ void __launch_bounds__(256) paint_scan_tiles_kernel(const int* __restrict__ count, int m,
                                                               int* __restrict__ start, int* __restrict__ bsum) {
  __shared__ int part[256];
  const int t = threadIdx.x, b = blockIdx.x;
  const int i0 = b * SCAN_TILE + 4 * t;
  int v[4], s = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    v[j] = (i0 + j < m) ? count[i0 + j] : 0;
    s += v[j];
  }
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int x = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += x;
    __syncthreads();
  }
  int run = part[t] - s;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (i0 + j < m) start[i0 + j] = run;
    run += v[j];
  }
  if (t == 255) bsum[b] = part[255];
}

// in-place exclusive scan of bsum[0 .. nb) (one workgroup of 1024 threads,
// contiguous chunks); bsum[nb] = total
__global__ void __launch_bounds__(1024) paint_scan_top_kernel(int* __restrict__ bsum, int nb) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int chunk = (nb + 1023) / 1024;
  const int b = t * chunk, e = min(nb, b + chunk);
  int s = 0;
  for (int i = b; i < e; ++i) s += bsum[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int x = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += x;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int i = b; i < e; ++i) {
    const int c = bsum[i];
    bsum[i] = run;
    run += c;
  }
  if (t == 1023) bsum[nb] = part[1023];
}

__global__ void __launch_bounds__(256) paint_scan_add_kernel(int* __restrict__ start, int m,
                                                             const int* __restrict__ bsum, int nb) {
  const int b = blockIdx.x, off = bsum[b];
  for (int j = threadIdx.x; j < SCAN_TILE; j += blockDim.x) {
    const int i = b * SCAN_TILE + j;
    if (i < m) start[i] += off;
  }
  if (b == 0 && threadIdx.x == 0) start[m] = bsum[nb];
}

// off != nullptr: the particles are [0, off[nbatch]) (a batch of catalogs;
// np is then an upper bound that sizes the grid)
__global__ void __launch_bounds__(256) paint_scatter_kernel(const double* __restrict__ pos,
                                                            const double* __restrict__ w, int64_t np,
                                                            const int64_t* __restrict__ off, int nbatch,
                                                            const int* __restrict__ key,
                                                            const int* __restrict__ start, int* __restrict__ cursor,
                                                            double* __restrict__ spos, double* __restrict__ sw) {
  const int64_t npe = off ? off[nbatch] : np;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npe; p += (int64_t)gridDim.x * blockDim.x) {
    const int k = key[p];
    const int64_t q = (int64_t)start[k] + atomicAdd(cursor + k, 1);
    spos[3 * q + 0] = pos[3 * p + 0];
    spos[3 * q + 1] = pos[3 * p + 1];
    spos[3 * q + 2] = pos[3 * p + 2];
    sw[q] = w[p];
  }
}

// The particle ranges a tile reads: for source tile offset s = (dx, dy, dz)
// (bits 4, 2, 1 of s: the source is the lower neighbour along those axes)
// the classes that contain every bit of s.  Runs of consecutive classes.
struct PaintSeg {
  int n;
  int first[14], len[14];  // source offset s, first class, number of classes (per run)
  int src[14];
};

__host__ __device__ constexpr PaintSeg paint_segments() {
  PaintSeg g{};
  for (int s = 0; s < 8; ++s) {
    int c = 0;
    while (c < 8) {
      if ((c & s) != s) {
        ++c;
        continue;
      }
      int e = c;
      while (e + 1 < 8 && ((e + 1) & s) == s) ++e;
      g.src[g.n] = s;
      g.first[g.n] = c;
      g.len[g.n] = e - c + 1;
      ++g.n;
      c = e + 1;
    }
  }
  return g;
}

constexpr PaintSeg kPaintSeg = paint_segments();
static_assert(kPaintSeg.n == 14, "27 neighbour/class combinations form 14 runs");

// Lane layout of the accumulation step: KL lanes per particle (one per z
// offset of the window), PG particles per wave instruction; WS = stride of a
// particle's staged values (odd, so lane-strided stores hit distinct banks).
template <int ORDER>
struct TileLanes {
  static constexpr int KL = ORDER == 1 ? 1 : ORDER == 2 ? 2 : 4;
  static constexpr int PG = 64 / KL;
  static constexpr int WS = (3 * ORDER) | 1;
};

// The tile sums in 64-bit fixed point: contribution v becomes round(v * 2^S)
// and the LDS additions are integer (LDS integer atomics run at full rate;
// floating-point LDS atomics do not), exact and order-independent, so the
// painted mesh is bitwise reproducible.  Per tile, S = OFF - e with
// 2^e > (particles read) * max|w| >= any cell's |sum|: FP64 meshes keep
// OFF = 61 (resolution below 2^-61 of that bound), FP32 meshes OFF = 50,
// which lets a contribution convert with one add of 1.5 * 2^52.
template <typename T>
struct TileFixed;
template <>
struct TileFixed<float> {
  static constexpr int OFF = 50;
  // round(a * b) for a = mx * wz (exact in FP64) and b = wy: one FMA onto
  // 1.5 * 2^52 leaves the integer in the low mantissa bits
  __device__ static long long cvt(double a, double b) {
    return __double_as_longlong(fma(a, b, 6755399441055744.0)) - 0x4338000000000000LL;
  }
};
template <>
struct TileFixed<double> {
  static constexpr int OFF = 61;
  __device__ static long long cvt(double a, double b) { return __double2ll_rn(a * b); }
};

// window<ORDER> with the PCS sixths as a multiplication by 1/6 (within an
// ulp of the division)
template <int ORDER>
__device__ __forceinline__ int window_tile(double u, double* w) {
  if (ORDER != 4) return window<ORDER>(u, w);
  constexpr double sixth = 1.0 / 6.0;
  const double i1 = floor(u);
  const double f = u - i1;
  const double g = 1.0 - f;
  w[0] = g * g * g * sixth;
  w[1] = (4.0 - 6.0 * f * f + 3.0 * f * f * f) * sixth;
  w[2] = (4.0 - 6.0 * g * g + 3.0 * g * g * g) * sixth;
  w[3] = f * f * f * sixth;
  return (int)i1 - 1;
}

// One workgroup per (catalog, tile); the copies are painted one after the
// other through one accumulator.  Each wave takes 64 particles at a time:
// one lane per particle forms the FP64 windows of the copy and stages the
// x weights times mass times 2^S, the y and z weights (all rounded to T) and
// the window's position in the tile; then KL lanes per particle, one per z
// offset, add the ORDER^2 products of the particle's (x, y) cells.
template <typename T, int ORDER, int PT>
__global__ void __launch_bounds__(PAINT_TILE_THREADS) paint_tile_kernel(
    const double* __restrict__ spos, const double* __restrict__ sw, const int* __restrict__ start, int n, int nt,
    double scale, double cx, double cy, double cz, int ncopy, T* __restrict__ mesh0, T* __restrict__ mesh1) {
  using L = TileLanes<ORDER>;
  constexpr int PT3 = PT * PT * PT, NW = PAINT_TILE_THREADS / 64;
  constexpr int DEAD = 0xFFFFFF;  // staged position of an empty slot: outside every tile
  __shared__ unsigned long long acc[PT3];
  __shared__ T stw[NW][64 * L::WS];
  __shared__ int stp[NW][64];
  __shared__ int seg_start[14], seg_count[14];
  __shared__ double wred[NW];
  const int nt3 = nt * nt * nt;
  // consecutive workgroups land on different XCDs: give each XCD a
  // contiguous run of tiles, so neighbouring tiles share an L2
  int bid = blockIdx.x;
  if (gridDim.x % 8 == 0) bid = (bid & 7) * (gridDim.x >> 3) + (bid >> 3);
  const int b = bid / nt3, tile = bid - b * nt3;
  const int tz = tile % nt, ty = (tile / nt) % nt, tx = tile / (nt * nt);
  const int x0 = tx * PT, y0 = ty * PT, z0 = tz * PT;
  for (int i = threadIdx.x; i < PT3; i += blockDim.x) acc[i] = 0ull;
  // the particle ranges to read (runs of consecutive classes, one source tile each)
  if (threadIdx.x < 14) {
    const int q = threadIdx.x;
    int s = 0, first = 0, len = 0;
#pragma unroll
    for (int j = 0; j < 14; ++j)
      if (j == q) {
        s = kPaintSeg.src[j];
        first = kPaintSeg.first[j];
        len = kPaintSeg.len[j];
      }
    const int sx = (s & 4) ? (tx == 0 ? nt - 1 : tx - 1) : tx;
    const int sy = (s & 2) ? (ty == 0 ? nt - 1 : ty - 1) : ty;
    const int sz = (s & 1) ? (tz == 0 ? nt - 1 : tz - 1) : tz;
    const int k0 = (((b * nt + sx) * nt + sy) * nt + sz) * 8 + first;
    const int s0 = start[k0];
    seg_start[q] = s0;
    seg_count[q] = start[k0 + len] - s0;
  }
  __syncthreads();
  int sst[14], scnt[14], total = 0;
#pragma unroll
  for (int q = 0; q < 14; ++q) {
    sst[q] = seg_start[q];
    scnt[q] = seg_count[q];
    total += scnt[q];
  }
  auto particle = [&](int v) -> int64_t {
    int64_t p = -1;
    int run = 0;
#pragma unroll
    for (int q = 0; q < 14; ++q) {
      if (v >= run && v < run + scnt[q]) p = (int64_t)sst[q] + (v - run);
      run += scnt[q];
    }
    return p;
  };
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // the tile's scale: max |w| over the particles read
  double wmax = 0.0;
  for (int v = threadIdx.x; v < total; v += blockDim.x) wmax = fmax(wmax, fabs(sw[particle(v)]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) wmax = fmax(wmax, __shfl_xor(wmax, o));
  if (lane == 0) wred[wv] = wmax;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NW; ++q) wmax = fmax(wmax, wred[q]);
  double bound = (double)total * wmax;
  if (!(bound <= 1e300)) bound = 1e300;
  int e = 0;
  frexp(bound, &e);
  const int S = TileFixed<T>::OFF - e;
  const int kk = lane % L::KL, slot = lane / L::KL;
  const int kw = kk < ORDER ? kk : ORDER - 1;
  T* const st = stw[wv];
  int* const sp = stp[wv];
  const int64_t n3 = (int64_t)n * n * n;
  // the wave's first 64 particles stay in registers for the second copy
  const int64_t p_first = particle(wv * 64 + lane);
  double xf = 0.0, yf = 0.0, zf = 0.0, mf = 0.0;
  if (p_first >= 0) {
    xf = spos[3 * p_first + 0] - cx;
    yf = spos[3 * p_first + 1] - cy;
    zf = spos[3 * p_first + 2] - cz;
    mf = sw[p_first];
  }
  for (int c = 0; c < ncopy; ++c) {
    const double sh = 0.5 * c;
    for (int base = wv * 64; base < total; base += NW * 64) {
      int64_t p = p_first;
      double x = xf, y = yf, z = zf, m = mf;
      if (base != wv * 64) {
        p = particle(base + lane);
        if (p >= 0) {
          x = spos[3 * p + 0] - cx;
          y = spos[3 * p + 1] - cy;
          z = spos[3 * p + 2] - cz;
          m = sw[p];
        }
      }
      if (p >= 0) {
        double wx[ORDER], wy[ORDER], wz[ORDER];
        int r[3];
        r[0] = window_tile<ORDER>(__dadd_rn(__dmul_rn(x, scale), sh), wx);
        r[1] = window_tile<ORDER>(__dadd_rn(__dmul_rn(y, scale), sh), wy);
        r[2] = window_tile<ORDER>(__dadd_rn(__dmul_rn(z, scale), sh), wz);
        const int o[3] = {x0, y0, z0};
        int pk = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          // first window cell relative to the tile, in [-4, PT]
          int rr = r[d] % n;
          if (rr < 0) rr += n;
          rr -= o[d];
          if (rr < 0) rr += n;
          if (rr >= n - 4) rr -= n;
          pk |= (rr + 4) << (8 * d);
        }
        T* w = st + lane * L::WS;
#pragma unroll
        for (int i = 0; i < ORDER; ++i) {
          w[i] = (T)ldexp(m * wx[i], S);
          w[ORDER + i] = (T)wy[i];
          w[2 * ORDER + i] = (T)wz[i];
        }
        sp[lane] = pk;
      } else {
        sp[lane] = DEAD;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int cnt = min(64, total - base);
      for (int g = 0; g < cnt; g += L::PG) {
        const int q = g + slot;
        const int pk = sp[q];
        const T* w = st + q * L::WS;
        const double wz = (double)w[2 * ORDER + kw];
        double mxz[ORDER], wy[ORDER];
#pragma unroll
        for (int i = 0; i < ORDER; ++i) {
          mxz[i] = (double)w[i] * wz;
          wy[i] = (double)w[ORDER + i];
        }
        const int rx = (pk & 255) - 4, ry = ((pk >> 8) & 255) - 4, iz = ((pk >> 16) & 255) - 4 + kk;
        const bool zok = kk < ORDER && (unsigned)iz < (unsigned)PT;
        unsigned long long* a = acc + (rx * PT + ry) * PT + iz;
#pragma unroll
        for (int i = 0; i < ORDER; ++i) {
#pragma unroll
          for (int j = 0; j < ORDER; ++j) {
            if (zok && (unsigned)(rx + i) < (unsigned)PT && (unsigned)(ry + j) < (unsigned)PT)
              atomicAdd(a + (i * PT + j) * PT, (unsigned long long)TileFixed<T>::cvt(mxz[i], wy[j]));
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __syncthreads();
    // write the tile (XP x planes per pass) and clear the accumulator
    constexpr int XP = PAINT_TILE_THREADS / (PT * PT);
    static_assert(XP >= 1 && PT % XP == 0 && XP * PT * PT == PAINT_TILE_THREADS, "whole x planes per pass");
    const int rem = threadIdx.x % (PT * PT), ix0 = threadIdx.x / (PT * PT);
    T* mp = (c == 0 ? mesh0 : mesh1) + b * n3 + ((int64_t)(x0 + ix0) * n + y0 + rem / PT) * n + z0 + rem % PT;
    const int64_t step = (int64_t)XP * n * n;
#pragma unroll 4
    for (int ix = ix0; ix < PT; ix += XP, mp += step) {
      const int t = ix * PT * PT + rem;
      *mp = (T)ldexp((double)(long long)acc[t], -S);
      acc[t] = 0ull;
    }
    __syncthreads();
  }
}

static int paint_pt(int n) { return n % 16 == 0 && n >= 32 ? 16 : 8; }

int paint_tiled_ok(int n) { return n % 8 == 0 && n >= 16 && n <= 4096; }

size_t paint_tiled_work_bytes(int64_t np, int n, int nbatch) {
  const int64_t nt = n / paint_pt(n), m = (int64_t)nbatch * nt * nt * nt * 8, nb = (m + SCAN_TILE - 1) / SCAN_TILE;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  return al(np * 4) + 3 * al((m + 1) * 4) + al((nb + 1) * 4) + al(np * 24) + al(np * 8);
}

template <typename T, int ORDER, int PT>
static void launch_tiles(const double* spos, const double* sw, const int* start, int n, int nt, int nbatch,
                         double scale, double cx, double cy, double cz, int ncopy, T* m0, T* m1, hipStream_t st) {
  hipLaunchKernelGGL((paint_tile_kernel<T, ORDER, PT>), dim3(nbatch * nt * nt * nt), dim3(PAINT_TILE_THREADS), 0, st,
                     spos, sw, start, n, nt, scale, cx, cy, cz, ncopy, m0, m1);
}

template <int ORDER, int PT>
static void launch_keys(const double* pos, const int64_t* off, int64_t np, int64_t np_max, int nbatch, int n,
                        double scale, double cx, double cy, double cz, int nt, int ncopy, int* key, int* count,
                        hipStream_t st) {
  const dim3 grd(grid_for(np_max > 0 ? np_max : 1, 256, (1 << 16) / nbatch + 1), nbatch);
  hipLaunchKernelGGL((paint_key_kernel<ORDER, PT>), grd, dim3(256), 0, st, pos, off, np, n, scale, cx, cy, cz, nt,
                     ncopy, key, count);
}

// off == nullptr: one catalog of np particles; else nbatch catalogs, catalog
// b = particles [off[b], off[b + 1]) (device array), np = their total,
// np_max = the largest.
template <typename T>
static hipError_t launch_paint_tiled_t(const double* pos, const double* w, const int64_t* off, int nbatch, int64_t np,
                                       int64_t np_max, int n, double boxsize, int order, const double* center,
                                       int ncopy, T* m0, T* m1, void* work, hipStream_t st) {
  const double scale = (double)n / boxsize;
  const double cx = center ? center[0] : 0.0, cy = center ? center[1] : 0.0, cz = center ? center[2] : 0.0;
  const int PT = paint_pt(n), nt = n / PT;
  const int64_t m64 = (int64_t)nbatch * nt * nt * nt * 8;
  if (m64 >= ((int64_t)1 << 31) || np >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
  const int m = (int)m64;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  char* base = (char*)work;
  int* key = (int*)base;
  base += al(np * 4);
  int* count = (int*)base;
  base += al((size_t)(m + 1) * 4);
  int* cursor = (int*)base;
  base += al((size_t)(m + 1) * 4);
  int* start = (int*)base;
  base += al((size_t)(m + 1) * 4);
  const int nb = (m + SCAN_TILE - 1) / SCAN_TILE;
  int* bsum = (int*)base;
  base += al((size_t)(nb + 1) * 4);
  double* spos = (double*)base;
  base += al(np * 24);
  double* sw = (double*)base;
  dim3 blk(256), grd(grid_for(np > 0 ? np : 1, 256, 1 << 16));
  hipLaunchKernelGGL(fill_int_kernel, dim3(grid_for(2 * (int64_t)(m + 1), 256, 1 << 14)), blk, 0, st, count,
                     2 * (int64_t)al((size_t)(m + 1) * 4) / 4);
  if (np > 0) {
#define CNPC_KEYS(O)                                                                                            \
  (PT == 16 ? launch_keys<O, 16>(pos, off, np, np_max, nbatch, n, scale, cx, cy, cz, nt, ncopy, key, count, st) \
            : launch_keys<O, 8>(pos, off, np, np_max, nbatch, n, scale, cx, cy, cz, nt, ncopy, key, count, st))
    switch (order) {
      case 1: CNPC_KEYS(1); break;
      case 2: CNPC_KEYS(2); break;
      case 3: CNPC_KEYS(3); break;
      case 4: CNPC_KEYS(4); break;
      default: return hipErrorInvalidValue;
    }
#undef CNPC_KEYS
  }
  hipLaunchKernelGGL(paint_scan_tiles_kernel, dim3(nb), dim3(256), 0, st, count, m, start, bsum);
  hipLaunchKernelGGL(paint_scan_top_kernel, dim3(1), dim3(1024), 0, st, bsum, nb);
  hipLaunchKernelGGL(paint_scan_add_kernel, dim3(nb), dim3(256), 0, st, start, m, bsum, nb);
  if (np > 0)
    hipLaunchKernelGGL(paint_scatter_kernel, grd, blk, 0, st, pos, w, np, off, nbatch, key, start, cursor, spos, sw);
#define CNPC_TILES(O)                                                                                              \
  (PT == 16 ? launch_tiles<T, O, 16>(spos, sw, start, n, nt, nbatch, scale, cx, cy, cz, ncopy, m0, m1, st)         \
            : launch_tiles<T, O, 8>(spos, sw, start, n, nt, nbatch, scale, cx, cy, cz, ncopy, m0, m1, st))
  switch (order) {
    case 1: CNPC_TILES(1); break;
    case 2: CNPC_TILES(2); break;
    case 3: CNPC_TILES(3); break;
    case 4: CNPC_TILES(4); break;
    default: return hipErrorInvalidValue;
  }
#undef CNPC_TILES
  return hipGetLastError();
}

hipError_t launch_paint_tiled(const double* pos, const double* w, const int64_t* off, int nbatch, int64_t np,
                              int64_t np_max, int n, double boxsize, int order, const double* center, int ncopy,
                              int prec, void* m0, void* m1, void* work, hipStream_t st) {
  if (prec == 0)
    return launch_paint_tiled_t<double>(pos, w, off, nbatch, np, np_max, n, boxsize, order, center, ncopy,
                                        (double*)m0, (double*)m1, work, st);
  return launch_paint_tiled_t<float>(pos, w, off, nbatch, np, np_max, n, boxsize, order, center, ncopy, (float*)m0,
                                     (float*)m1, work, st);
}

hipError_t launch_paint(const double* pos, const double* w, int64_t np, int n, double boxsize, int order,
                        const double* center, int ncopy, int prec, void* m0, void* m1, hipStream_t st) {
  if (prec == 0)
    return launch_paint_t<double>(pos, w, np, n, boxsize, order, center, ncopy, (double*)m0, (double*)m1, st);
  return launch_paint_t<float>(pos, w, np, n, boxsize, order, center, ncopy, (float*)m0, (float*)m1, st);
}

}  // namespace cnpc
