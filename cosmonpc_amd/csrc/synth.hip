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
// (256 / ZB) x-rows.  Output U[s][x][ky][kz], ky index = ny + R.  The lines'
// values and bins are staged in LDS first (coalesced), then every shell's
// |kx| interval is found by scanning the staged bins.
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
  signed char* lb = (signed char*)(ab + 2 * ZB * a.nshell);  // [ZB][Q] bins at +kx
  for (int j = threadIdx.x; j < n; j += blockDim.x) tw[j] = twg[j];
  for (int t = threadIdx.x; t < ZB * Q; t += blockDim.x) {
    const int l = t / Q, kx = t % Q;
    const int z = kz0 + l;
    C vp, vn;
    vp.x = vn.x = (T)0;
    vp.y = vn.y = (T)0;
    signed char b = -1;
    if (z < Q) {
      const int64_t base = ((int64_t)kyi * Q + z) * P;
      vp = V[base + R + kx];
      vn = V[base + R - kx];
      b = bins[base + R + kx];
    }
    lv[l * Q + kx] = vp;
    lvn[l * Q + kx] = vn;
    lb[l * Q + kx] = b;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < ZB * a.nshell; t += blockDim.x) {
    const int l = t / a.nshell, s = t % a.nshell;
    const signed char want = (signed char)(a.bin0 + s);
    const signed char* bl = lb + l * Q;
    int lo = 1, hi = 0;  // empty
    for (int kx = 0; kx < Q; ++kx) {
      if (bl[kx] == want) {
        if (lo > hi) lo = kx;
        hi = kx;
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
// cos / sin of 2 pi j / R for the small DFTs
template <int RR> struct Roots;
template <> struct Roots<2> {
  static constexpr double c[2] = {1.0, -1.0};
  static constexpr double s[2] = {0.0, 0.0};
};
template <> struct Roots<3> {
  static constexpr double c[3] = {1.0, -0.5, -0.5};
  static constexpr double s[3] = {0.0, 0.86602540378443865, -0.86602540378443865};
};
template <> struct Roots<4> {
  static constexpr double c[4] = {1.0, 0.0, -1.0, 0.0};
  static constexpr double s[4] = {0.0, 1.0, 0.0, -1.0};
};
template <> struct Roots<5> {
  static constexpr double c[5] = {1.0, 0.30901699437494742, -0.80901699437494742, -0.80901699437494742,
                                  0.30901699437494742};
  static constexpr double s[5] = {0.0, 0.95105651629515357, 0.58778525229247313, -0.58778525229247313,
                                  -0.95105651629515357};
};
// One in-place pass of the mixed-radix Stockham inverse FFT (e^{+2 pi i})
// over nl lines of length n in LDS.  Butterfly i in [0, n/R) of a line reads
// buf[i + r n/R], multiplies by e^{2 pi i r k/(p R)} (k = i mod p), does an
// R-point DFT and writes buf[(i/p) p R + k + r p].  All reads are staged in
// registers before the barrier, so one buffer suffices; a thread owns at most
// MAXB butterflies (the host sizes nl accordingly).
template <int RR, int MAXB, typename C>
__device__ __forceinline__ void stockham_pass(C* __restrict__ buf, const C* __restrict__ W, int n, int nl, int p) {
  typedef decltype(C().x) T;
  const int Tn = n / RR;
  const int pR = p * RR;
  const int tstep = n / pR;
  const int total = nl * Tn;
  C u[MAXB][RR];
  int dsto[MAXB];
#pragma unroll
  for (int it = 0; it < MAXB; ++it) {
    const int t = threadIdx.x + it * blockDim.x;
    dsto[it] = -1;
    if (t < total) {
      const int l = t / Tn, i = t - l * Tn;
      const int k = i % p;
      const C* s = buf + l * n;
#pragma unroll
      for (int r = 0; r < RR; ++r) {
        u[it][r] = s[i + r * Tn];
        if (r) u[it][r] = cmul(u[it][r], W[r * k * tstep]);
      }
      dsto[it] = l * n + (i / p) * pR + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < MAXB; ++it) {
    if (dsto[it] >= 0) {
      C* d = buf + dsto[it];
#pragma unroll
      for (int b = 0; b < RR; ++b) {
        C v = u[it][0];
#pragma unroll
        for (int q = 1; q < RR; ++q) {
          const int j = (q * b) % RR;
          const T wc = (T)Roots<RR>::c[j], ws = (T)Roots<RR>::s[j];
          v.x += u[it][q].x * wc - u[it][q].y * ws;
          v.y += u[it][q].x * ws + u[it][q].y * wc;
        }
        d[b * p] = v;
      }
    }
  }
  __syncthreads();
}

template <int MAXB, typename C>
__device__ void stockham_lds(C* buf, const C* W, int n, int nl, const SynthArgs& a) {
  int p = 1;
  for (int s = 0; s < a.nrad; ++s) {
    const int R = a.rad[s];
    if (R == 4) stockham_pass<4, MAXB>(buf, W, n, nl, p);
    else if (R == 2) stockham_pass<2, MAXB>(buf, W, n, nl, p);
    else if (R == 3) stockham_pass<3, MAXB>(buf, W, n, nl, p);
    else stockham_pass<5, MAXB>(buf, W, n, nl, p);
    p *= R;
  }
}

// butterflies per thread and pass (bounds the registers of the staged pass)
template <typename T> struct MaxB { static constexpr int v = 4; };

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
  C* buf = W + n;
  for (int j = threadIdx.x; j < n; j += blockDim.x) W[j] = twg[j];
  for (int j = threadIdx.x; j < NL * n; j += blockDim.x) {
    buf[j].x = (T)0;
    buf[j].y = (T)0;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NL * P; t += blockDim.x) {
    const int l = t % NL, kyi = t / NL;
    const int kz = kz0 + l;
    if (kz < Q) {
      int y = kyi - R;
      if (y < 0) y += n;
      buf[l * n + y] = U[(plane * P + kyi) * Q + kz];
    }
  }
  __syncthreads();
  stockham_lds<MaxB<T>::v>(buf, W, n, NL, a);
  for (int t = threadIdx.x; t < NL * n; t += blockDim.x) {
    const int l = t / n, y = t - l * n;
    const int kz = kz0 + l;
    if (kz < Q) Tm[(plane * Q + kz) * n + y] = buf[l * n + y];
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
  C* buf = W + n;
  for (int j = threadIdx.x; j < n; j += blockDim.x) W[j] = twg[j];
  for (int j = threadIdx.x; j < NL * n; j += blockDim.x) {
    buf[j].x = (T)0;
    buf[j].y = (T)0;
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
    C* L = buf + l * n;
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
  stockham_lds<MaxB<T>::v>(buf, W, n, NL, a);
  for (int t = threadIdx.x; t < 2 * NL * n; t += blockDim.x) {
    const int r = t / n, z = t - r * n;
    const int y = y0 + r;
    if (y < n) {
      const C v = buf[(r >> 1) * n + z];
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
  return cs * (size_t)n + 2 * cs * (size_t)zb * (R + 1) + sizeof(int) * 2 * (size_t)zb * nshell +
         (size_t)zb * (R + 1);
}

// lines per y / z workgroup: twiddles + one line buffer within 48 KB, and at
// most MAXB butterflies per thread in every pass (smallest radix rmin)
int synth_lines_per_block(int prec, int n, int rmin) {
  const size_t cs = prec == 0 ? 16 : 8;
  const size_t cap = 48 * 1024 / cs;
  if (cap <= (size_t)n) return 0;
  int nl = (int)((cap - n) / (size_t)n);
  const int maxb = 4;  // MaxB<T>::v
  const int nb = maxb * 256 * rmin / n;
  if (nl > nb) nl = nb;
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
  int rmin = 8;
  for (int i = 0; i < a.nrad; ++i) rmin = a.rad[i] < rmin ? a.rad[i] : rmin;
  const int NL = synth_lines_per_block(prec, n, rmin);
  const size_t lyz = sizeof(C) * ((size_t)n + (size_t)NL * n);
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
