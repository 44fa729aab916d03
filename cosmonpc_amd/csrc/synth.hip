// Pruned shell synthesis: the inverse 3-D FFT of shell-filtered spectra,
// written for spectra that vanish outside the box |n_d| <= R.
//
// A shell touches only modes with |k| < k_max, i.e. |n_d| <= R with
// R = ceil(k_max / kf) well below N/2.  The full 3-D c2r of the reference
// (one N^3 IFFT per shell, clustering_estimator.py:934-970) spends most of
// its passes on zeros.  Here, for a batch of shells:
//
//  1. cube_kernel   V[ky][kz][kx] = w(khat) delta(k) for |n_x|,|n_y| <= R,
//                   0 <= n_z <= R, plus the bin of every mode: (2R+1)^2 (R+1)
//                   entries shared by all shells of the call;
//  2. xsynth_kernel sparse x-DFT: U[s][x][ky][kz] = sum_{kx in shell s on the
//                   line (ky, kz)} V e^{+2 pi i kx x / N}.  On a line the bin is
//                   non-decreasing in |kx|, so a shell is one |kx| interval;
//  3. ysynth_kernel y-FFT (length N, mixed-radix Stockham in LDS) of the
//                   2R+1 non-zero ky of every (s, x, kz) column:
//                   T[s][x][kz][y];
//  4. zsynth_kernel z-c2r of the R+1 non-zero kz of every (s, x, y) row,
//                   two rows per complex FFT (real part = row y, imaginary
//                   part = row y+1), written as the real shells [s][x][y][z].
//
// All transforms are the unnormalised inverse e^{+i k.x} of the reference's
// pmesh c2r.  Twiddles come from one table W[j] = e^{2 pi i j / N} in LDS.
#include "cnpc_common.h"

namespace cnpc {

template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// ------------------------------------------------------------------ cube
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

// ------------------------------------------------------------------ x pass
// one workgroup per (ky, block of ZB consecutive kz); threads = ZB lanes x
// (256 / ZB) x-rows.  Output U[s][x][ky][kz], ky index = ny + R.
template <typename T, int ZB>
__global__ void __launch_bounds__(256) xsynth_kernel(SynthArgs a, const typename cplx<T>::type* __restrict__ V,
                                                     const signed char* __restrict__ bins,
                                                     const typename cplx<T>::type* __restrict__ twg,
                                                     typename cplx<T>::type* __restrict__ U) {
  typedef typename cplx<T>::type C;
  const int n = a.n, R = a.R, P = 2 * R + 1, Q = R + 1;
  const int kyi = blockIdx.x;
  const int kz0 = blockIdx.y * ZB;
  const int lane = threadIdx.x % ZB, xrow = threadIdx.x / ZB;
  constexpr int XR = 256 / ZB;
  const int kz = kz0 + lane;
  extern __shared__ unsigned char smem[];
  C* tw = (C*)smem;                      // n twiddles
  C* lv = tw + n;                        // ZB lines x Q values at +kx
  C* lvn = lv + ZB * Q;                  // ... and at -kx
  int* ab = (int*)(lvn + ZB * Q);        // [ZB][nshell][2] |kx| interval per shell
  for (int j = threadIdx.x; j < n; j += blockDim.x) tw[j] = twg[j];
  for (int t = threadIdx.x; t < ZB * Q; t += blockDim.x) {
    const int l = t / Q, kx = t % Q;
    const int z = kz0 + l;
    C vp, vn;
    vp.x = vn.x = (T)0;
    vp.y = vn.y = (T)0;
    if (z < Q) {
      const int64_t base = ((int64_t)kyi * Q + z) * P;
      vp = V[base + R + kx];
      vn = V[base + R - kx];
    }
    lv[l * Q + kx] = vp;
    lvn[l * Q + kx] = vn;
  }
  for (int t = threadIdx.x; t < ZB * a.nshell; t += blockDim.x) {
    const int l = t / a.nshell, s = t % a.nshell;
    const int z = kz0 + l;
    int lo = 1, hi = 0;  // empty
    if (z < Q) {
      const int64_t base = ((int64_t)kyi * Q + z) * P + R;
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
  if (kz >= Q) return;
  const C* myv = lv + lane * Q;
  const C* myvn = lvn + lane * Q;
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
          if (kx != 0) {  // -kx: conjugate twiddle
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
      U[(((int64_t)s * n + x) * P + kyi) * Q + kz] = o;
    }
  }
}

// ------------------------------------------------------------------ LDS FFT
// One pass of the mixed-radix Stockham inverse FFT (e^{+2 pi i}) over nl
// lines of length n: thread item i in [0, n/R) of each line reads
// src[i + r n/R], multiplies by e^{2 pi i r k/(p R)} (k = i mod p), does an
// R-point DFT and writes dst[(i/p) p R + k + r p].
template <int RR, typename C>
__device__ __forceinline__ void stockham_pass(const C* __restrict__ src, C* __restrict__ dst, const C* __restrict__ W,
                                              int n, int nl, int p) {
  const int Tn = n / RR;
  const int pR = p * RR;
  const int tstep = n / pR;
  const int dstep = n / RR;
  for (int t = threadIdx.x; t < nl * Tn; t += blockDim.x) {
    const int l = t / Tn, i = t - l * Tn;
    const int k = i % p;
    const C* s = src + l * n;
    C u[RR];
#pragma unroll
    for (int r = 0; r < RR; ++r) {
      u[r] = s[i + r * Tn];
      if (r) u[r] = cmul(u[r], W[r * k * tstep]);
    }
    C* d = dst + l * n + (i / p) * pR + k;
#pragma unroll
    for (int b = 0; b < RR; ++b) {
      C v = u[0];
#pragma unroll
      for (int q = 1; q < RR; ++q) {
        const C w = W[((q * b) % RR) * dstep];
        const C m = cmul(u[q], w);
        v.x += m.x;
        v.y += m.y;
      }
      d[b * p] = v;
    }
  }
}

// runs all passes; returns the buffer that holds the result
template <typename C>
__device__ C* stockham_lds(C* b0, C* b1, const C* W, int n, int nl, const SynthArgs& a) {
  C* src = b0;
  C* dst = b1;
  int p = 1;
  for (int s = 0; s < a.nrad; ++s) {
    const int R = a.rad[s];
    if (R == 4) stockham_pass<4>(src, dst, W, n, nl, p);
    else if (R == 2) stockham_pass<2>(src, dst, W, n, nl, p);
    else if (R == 3) stockham_pass<3>(src, dst, W, n, nl, p);
    else if (R == 5) stockham_pass<5>(src, dst, W, n, nl, p);
    else stockham_pass<7>(src, dst, W, n, nl, p);
    __syncthreads();
    C* t = src;
    src = dst;
    dst = t;
    p *= R;
  }
  return src;
}

// ------------------------------------------------------------------ y pass
// one workgroup per (plane s*n + x, block of NL kz columns)
template <typename T>
__global__ void __launch_bounds__(256) ysynth_kernel(SynthArgs a, int NL, const typename cplx<T>::type* __restrict__ U,
                                                     const typename cplx<T>::type* __restrict__ twg,
                                                     typename cplx<T>::type* __restrict__ Tm) {
  typedef typename cplx<T>::type C;
  const int n = a.n, R = a.R, P = 2 * R + 1, Q = R + 1;
  const int64_t plane = blockIdx.x;
  const int kz0 = blockIdx.y * NL;
  extern __shared__ unsigned char smem[];
  C* W = (C*)smem;
  C* b0 = W + n;
  C* b1 = b0 + NL * n;
  for (int j = threadIdx.x; j < n; j += blockDim.x) W[j] = twg[j];
  for (int j = threadIdx.x; j < NL * n; j += blockDim.x) {
    b0[j].x = (T)0;
    b0[j].y = (T)0;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NL * P; t += blockDim.x) {
    const int l = t % NL, kyi = t / NL;
    const int kz = kz0 + l;
    if (kz < Q) {
      int y = kyi - R;
      if (y < 0) y += n;
      b0[l * n + y] = U[(plane * P + kyi) * Q + kz];
    }
  }
  __syncthreads();
  const C* res = stockham_lds(b0, b1, W, n, NL, a);
  for (int t = threadIdx.x; t < NL * n; t += blockDim.x) {
    const int l = t / n, y = t - l * n;
    const int kz = kz0 + l;
    if (kz < Q) Tm[(plane * Q + kz) * n + y] = res[l * n + y];
  }
}

// ------------------------------------------------------------------ z pass
// one workgroup per (plane s*n + x, block of 2*NL rows); line l carries rows
// y0+2l (real part) and y0+2l+1 (imaginary part)
template <typename T>
__global__ void __launch_bounds__(256) zsynth_kernel(SynthArgs a, int NL, const typename cplx<T>::type* __restrict__ Tm,
                                                     const typename cplx<T>::type* __restrict__ twg,
                                                     T* __restrict__ out) {
  typedef typename cplx<T>::type C;
  const int n = a.n, Q = a.R + 1;
  const int64_t plane = blockIdx.x;
  const int y0 = blockIdx.y * 2 * NL;
  extern __shared__ unsigned char smem[];
  C* W = (C*)smem;
  C* b0 = W + n;
  C* b1 = b0 + NL * n;
  for (int j = threadIdx.x; j < n; j += blockDim.x) W[j] = twg[j];
  for (int j = threadIdx.x; j < NL * n; j += blockDim.x) {
    b0[j].x = (T)0;
    b0[j].y = (T)0;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NL * Q; t += blockDim.x) {
    const int l = t % NL, kz = t / NL;
    const int ya = y0 + 2 * l, yb = ya + 1;
    C A, B;
    A.x = A.y = B.x = B.y = (T)0;
    const int64_t row = (plane * Q + kz) * n;
    if (ya < n) A = Tm[row + ya];
    if (yb < n) B = Tm[row + yb];
    C* L = b0 + l * n;
    if (kz == 0) {  // c2r reads only the real part of the zero mode
      L[0].x = A.x;
      L[0].y = B.x;
    } else {        // w[k] = A + iB,  w[n-k] = conj(A) + i conj(B)
      L[kz].x = A.x - B.y;
      L[kz].y = A.y + B.x;
      L[n - kz].x = A.x + B.y;
      L[n - kz].y = B.x - A.y;
    }
  }
  __syncthreads();
  const C* res = stockham_lds(b0, b1, W, n, NL, a);
  for (int t = threadIdx.x; t < 2 * NL * n; t += blockDim.x) {
    const int r = t / n, z = t - r * n;
    const int y = y0 + r;
    if (y < n) {
      const C v = res[(r >> 1) * n + z];
      out[(plane * n + y) * n + z] = (r & 1) ? v.y : v.x;
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

// lines per y / z workgroup so that twiddles + two line buffers fit 64 KB
int synth_lines_per_block(int prec, int n) {
  const size_t cs = prec == 0 ? 16 : 8;
  const size_t cap = 64 * 1024 / cs;
  if (cap <= (size_t)n) return 0;
  int nl = (int)((cap - n) / (2 * (size_t)n));
  return nl > 16 ? 16 : nl;
}

template <typename T, int ZB>
static void launch_x(const SynthArgs& a, const void* V, const signed char* bins, const void* tw, void* U, size_t lds,
                     hipStream_t st) {
  typedef typename cplx<T>::type C;
  const int Q = a.R + 1;
  dim3 grd(2 * a.R + 1, (Q + ZB - 1) / ZB);
  hipLaunchKernelGGL((xsynth_kernel<T, ZB>), grd, dim3(256), lds, st, a, (const C*)V, bins, (const C*)tw, (C*)U);
}

template <typename T>
static hipError_t launch_synth_t(const SynthArgs& a, const BinEdges& E, const PolySpec& Wp, const void* D, void* V,
                                 signed char* bins, void* tw, void* work, void* out, int zb, hipStream_t st) {
  typedef typename cplx<T>::type C;
  const int prec = sizeof(T) == 8 ? 0 : 1;
  const int R = a.R, P = 2 * R + 1, Q = R + 1, n = a.n;
  const int64_t cube = (int64_t)P * Q * P;
  C* U = (C*)work;
  C* Tm = U + (int64_t)a.nshell * n * P * Q;
  dim3 blk(256);
  hipLaunchKernelGGL(twiddle_kernel<T>, dim3((n + 255) / 256), blk, 0, st, n, (C*)tw);
  hipLaunchKernelGGL(cube_kernel<T>, dim3(grid_for(cube, 256, 1 << 14)), blk, 0, st, a, E, Wp, (const C*)D, (C*)V,
                     bins);
  const size_t lx = xsynth_lds_bytes(prec, n, R, a.nshell, zb);
  if (zb == 32) launch_x<T, 32>(a, V, bins, tw, U, lx, st);
  else if (zb == 16) launch_x<T, 16>(a, V, bins, tw, U, lx, st);
  else launch_x<T, 8>(a, V, bins, tw, U, lx, st);
  const int NL = synth_lines_per_block(prec, n);
  const size_t lyz = sizeof(C) * ((size_t)n + 2 * (size_t)NL * n);
  const int64_t planes = (int64_t)a.nshell * n;
  hipLaunchKernelGGL(ysynth_kernel<T>, dim3((unsigned)planes, (Q + NL - 1) / NL), blk, lyz, st, a, NL, (const C*)U,
                     (const C*)tw, Tm);
  hipLaunchKernelGGL(zsynth_kernel<T>, dim3((unsigned)planes, (n + 2 * NL - 1) / (2 * NL)), blk, lyz, st, a, NL,
                     (const C*)Tm, (const C*)tw, (T*)out);
  return hipGetLastError();
}

hipError_t launch_synth(int prec, const SynthArgs& a, const BinEdges& E, const PolySpec& W, const void* D, void* V,
                        signed char* bins, void* tw, void* work, void* out, int zb, hipStream_t st) {
  if (prec == 0) return launch_synth_t<double>(a, E, W, D, V, bins, tw, work, out, zb, st);
  return launch_synth_t<float>(a, E, W, D, V, bins, tw, work, out, zb, st);
}

}  // namespace cnpc
