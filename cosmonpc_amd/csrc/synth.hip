// Pruned shell synthesis: the inverse 3-D FFT of shell-filtered spectra,
// written for spectra that vanish outside the box |n_d| <= R.
//
// A shell touches only modes with |k| < k_max, i.e. |n_d| <= R with
// R = ceil(k_max / kf) well below N/2.  The full 3-D c2r of the reference
// (one N^3 IFFT per shell, clustering_estimator.py:934-970) spends most of
// its passes on zeros.  Here, for a batch of shells (P = 2R+1, Q = R+1):
//
//  1. cube_kernel  V[kz][ky][kx] = w(khat) delta(k) for |n_x|,|n_y| <= R,
//                  0 <= n_z <= R, plus the bin of every mode; shared by all
//                  shells of the call;
//  2. xfft_kernel  x-FFT (length N) of every (kz, ky) line, masked to shell
//                  s: U[s][x][kz][ky] (the fixed 20 x 20 plans write
//                  U[s][kz][ky block][x][ky % block], one contiguous run per
//                  workgroup);
//  3. yfft_kernel  y-FFT of the P non-zero ky of every (s, x, kz) line
//                  (contiguous in U): T[s][x][kz][y];
//  4. zfft_kernel  z-c2r of the Q non-zero kz of every (s, x, y) row, two rows
//                  per complex FFT (real part = row y, imaginary part = row
//                  y+1), written as the real shells [s][x][y][z].
//
// The FFTs are in-place mixed-radix Stockham passes in LDS over NL lines per
// workgroup.  Lines are padded to N+1 entries so that accesses striding over
// lines spread over the LDS banks, and the odd radices run first, where the
// Stockham writes stride by the radix.  All transforms are the unnormalised
// inverse e^{+i k.x} of the reference's pmesh c2r.  Each workgroup lays the
// twiddles of all passes out pass by pass in LDS (conflict-free reads).
#include "lds_fft.h"

namespace cnpc {

// threads per workgroup of the x, y and z passes
#ifndef CNPC_SYNTH_THREADS
#define CNPC_SYNTH_THREADS 256
#endif
constexpr int SYNTH_THREADS = CNPC_SYNTH_THREADS;

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
    const int kyi = (int)(r % P);
    const int nz = (int)(r / P);
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
// one workgroup per (block of NL ky lines, kz); line l = ky index ky0 + l
template <typename T>
__global__ void __launch_bounds__(SYNTH_THREADS) xfft_kernel(SynthArgs a, int NL, const typename cplx<T>::type* __restrict__ V,
                                                   const signed char* __restrict__ bins,
                                                   const typename cplx<T>::type* __restrict__ twg,
                                                   typename cplx<T>::type* __restrict__ U) {
  typedef typename cplx<T>::type C;
  const int n = a.n, R = a.R, P = 2 * R + 1, Q = R + 1, ls = n + 1;
  const int ky0 = blockIdx.x * NL, kz = blockIdx.y;
  const int nl = min(NL, P - ky0);
  extern __shared__ unsigned char smem[];
  C* W = (C*)smem;
  C* buf = W + n;
  build_pass_twiddles(W, twg, n, a);
  const float inv_n = 1.0f / (float)n, inv_nl = 1.0f / (float)nl;
  const int64_t lbase = ((int64_t)kz * P + ky0) * P;
  for (int s = 0; s < a.nshell; ++s) {
    const signed char want = (signed char)(a.bin0 + s);
    // lines with kz or every ky of the block beyond the shell's radius are
    // zero for this shell: skipped (the y pass never reads them)
    const int rs = a.rs[s];
    if (kz > rs || ky0 - R > rs || ky0 + nl - 1 - R < -rs) continue;
    // loads issued UNR at a time before the LDS writes (latency hiding)
    for (int t0 = threadIdx.x; t0 < nl * n; t0 += UNR * blockDim.x) {
      C v[UNR];
      signed char bb[UNR];
      int dst[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int t = t0 + u * blockDim.x;
        dst[u] = -1;
        bb[u] = -1;
        v[u].x = (T)0;
        v[u].y = (T)0;
        if (t < nl * n) {
          const int l = fdiv(t, n, inv_n), pos = t - l * n;
          const int f = pos < (n >> 1) ? pos : pos - n;
          dst[u] = l * ls + pos;
          if (f >= -R && f <= R) {
            const int64_t j = lbase + (int64_t)l * P + f + R;
            bb[u] = bins[j];
            v[u] = V[j];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        if (dst[u] >= 0) {
          C o = v[u];
          if (bb[u] != want) o.x = o.y = (T)0;
          buf[dst[u]] = o;
        }
      }
    }
    __syncthreads();
    stockham_lds<MaxB<T>::v, sizeof(T) == 4>(buf, W, n, ls, nl, a);
    // U[s][x][kz][ky]: consecutive threads take consecutive ky of one x
    for (int t = threadIdx.x; t < nl * n; t += blockDim.x) {
      const int x = fdiv(t, nl, inv_nl), l = t - x * nl;
      U[(((int64_t)s * n + x) * Q + kz) * P + ky0 + l] = buf[l * ls + x];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ y pass
// one workgroup per (plane s*n + x, block of NL kz lines).  nslots == 0: the
// output is T[s][x][kz][y] for the z pass; nslots > 0: the half-synthesised
// layout H[x][y/2][slot0 + s][kz][y%2] of the fused z-synthesis contraction
// (each (x, y pair) tile one contiguous run over all slots).
template <typename T>
__global__ void __launch_bounds__(SYNTH_THREADS) yfft_kernel(SynthArgs a, int NL, const typename cplx<T>::type* __restrict__ U,
                                                   const typename cplx<T>::type* __restrict__ twg,
                                                   typename cplx<T>::type* __restrict__ Tm, int nslots, int slot0) {
  typedef typename cplx<T>::type C;
  const int n = a.n, R = a.R, P = 2 * R + 1, Q = R + 1, ls = n + 1;
  const int64_t plane = blockIdx.x;
  const int kz0 = blockIdx.y * NL;
  const int nl = min(NL, Q - kz0);
  // kz lines beyond the shell's radius are zero: not computed, not written
  // (the z pass and the fused contraction read kz <= rs only)
  const int rs = a.rs[(int)(plane / n)];
  if (kz0 > rs) return;
  extern __shared__ unsigned char smem[];
  C* W = (C*)smem;
  C* buf = W + n;
  build_pass_twiddles(W, twg, n, a);
  const float inv_n = 1.0f / (float)n;
  const C* src = U + (plane * Q + kz0) * P;
  for (int t0 = threadIdx.x; t0 < nl * n; t0 += UNR * blockDim.x) {
    C v[UNR];
    int dst[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int t = t0 + u * blockDim.x;
      dst[u] = -1;
      v[u].x = (T)0;
      v[u].y = (T)0;
      if (t < nl * n) {
        const int l = fdiv(t, n, inv_n), y = t - l * n;
        const int f = y < (n >> 1) ? y : y - n;
        dst[u] = l * ls + y;
        if (f >= -rs && f <= rs && kz0 + l <= rs) v[u] = src[l * P + f + R];
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u)
      if (dst[u] >= 0) buf[dst[u]] = v[u];
  }
  __syncthreads();
  stockham_lds<MaxB<T>::v, sizeof(T) == 4>(buf, W, n, ls, nl, a);
  if (nslots == 0) {
    C* dst = Tm + (plane * Q + kz0) * n;
    for (int t = threadIdx.x; t < nl * n; t += blockDim.x) {
      const int l = fdiv(t, n, inv_n), y = t - l * n;
      dst[(int64_t)l * n + y] = buf[l * ls + y];
    }
  } else {
    // consecutive threads: (y%2, l) fastest, so each y pair writes 2 nl
    // consecutive entries
    const int s = (int)(plane / n), x = (int)(plane - (int64_t)s * n);
    const float inv_nl = 1.0f / (float)nl;
    const int64_t xbase = (int64_t)x * (n >> 1);
    for (int t = threadIdx.x; t < nl * n; t += blockDim.x) {
      const int j = t & 1, r = t >> 1;
      const int yp = fdiv(r, nl, inv_nl), l = r - yp * nl;
      Tm[(((xbase + yp) * nslots + slot0 + s) * Q + kz0 + l) * 2 + j] = buf[l * ls + 2 * yp + j];
    }
  }
}

// ------------------------------------------------------------------ z pass
// one workgroup per (plane s*n + x, block of 2*NL rows); line l carries rows
// y0+2l (real part) and y0+2l+1 (imaginary part)
template <typename T>
__global__ void __launch_bounds__(SYNTH_THREADS) zfft_kernel(SynthArgs a, int NL, const typename cplx<T>::type* __restrict__ Tm,
                                                   const typename cplx<T>::type* __restrict__ twg,
                                                   T* __restrict__ out) {
  typedef typename cplx<T>::type C;
  const int n = a.n, R = a.R, Q = R + 1, ls = n + 1;
  const int64_t plane = blockIdx.x;
  const int y0 = blockIdx.y * 2 * NL;
  const int nrows = min(2 * NL, n - y0);
  const int nl = (nrows + 1) >> 1;
  extern __shared__ unsigned char smem[];
  C* W = (C*)smem;
  C* buf = W + n;
  build_pass_twiddles(W, twg, n, a);
  // this shell's modes have kz <= rs: zero the modes rs+1 .. n-rs-1 of every
  // line (the rest is written below)
  const int rs = a.rs[(int)(plane / n)], qs = rs + 1;
  const int nz0 = n - 2 * rs - 1;
  const float inv_z = 1.0f / (float)(nz0 > 0 ? nz0 : 1), inv_nl = 1.0f / (float)nl;
  for (int t = threadIdx.x; t < nl * nz0; t += blockDim.x) {
    const int l = fdiv(t, nz0, inv_z), j = t - l * nz0;
    C z;
    z.x = (T)0;
    z.y = (T)0;
    buf[l * ls + qs + j] = z;
  }
  constexpr int UZ = UNR / 2;
  for (int t0 = threadIdx.x; t0 < nl * qs; t0 += UZ * blockDim.x) {
    C A[UZ], B[UZ];
    int kzs[UZ], loff[UZ];
#pragma unroll
    for (int u = 0; u < UZ; ++u) {
      const int t = t0 + u * blockDim.x;
      kzs[u] = -1;
      loff[u] = 0;
      A[u].x = A[u].y = B[u].x = B[u].y = (T)0;
      if (t < nl * qs) {
        const int kz = fdiv(t, nl, inv_nl), l = t - kz * nl;
        const int ya = y0 + 2 * l, yb = ya + 1;
        const C* row = Tm + (plane * Q + kz) * n;
        A[u] = row[ya];
        if (yb < n) B[u] = row[yb];
        kzs[u] = kz;
        loff[u] = l * ls;
      }
    }
#pragma unroll
    for (int u = 0; u < UZ; ++u) {
      const int kz = kzs[u];
      if (kz < 0) continue;
      C* L = buf + loff[u];
      if (kz == 0) {  // c2r reads only the real part of the zero mode
        L[0].x = A[u].x;
        L[0].y = B[u].x;
      } else {        // w[k] = A + iB,  w[n-k] = conj(A) + i conj(B)
        L[kz].x = A[u].x - B[u].y;
        L[kz].y = A[u].y + B[u].x;
        L[n - kz].x = A[u].x + B[u].y;
        L[n - kz].y = B[u].x - A[u].y;
      }
    }
  }
  __syncthreads();
  stockham_lds<MaxB<T>::v, sizeof(T) == 4>(buf, W, n, ls, nl, a);
  const float inv_n = 1.0f / (float)n;
  T* dst = out + (plane * n + y0) * n;
  for (int t = threadIdx.x; t < nrows * n; t += blockDim.x) {
    const int r = fdiv(t, n, inv_n), z = t - r * n;
    const C v = buf[(r >> 1) * ls + z];
    dst[(int64_t)r * n + z] = (r & 1) ? v.y : v.x;
  }
}


// ------------------------------------------------------------ fixed plans
// The x and y passes of the half synthesis for grids n = N1 * N2 with the
// two-pass in-place FFT of lds_fft.h (FixedPlan): inputs are scattered to
// their padded positions, outputs gathered back in natural order, so U and
// H are the same as with the generic passes.  NL lines per workgroup, line
// pitch LSC = KP + 2 complex.
template <int N1, int N2>
struct SynthFixed {
  typedef FixedPlan<N1, N2> P;
  static constexpr int LSC = P::KP + 2;
};

template <int N1, int N2, int THREADS, int NL>
__global__ void __launch_bounds__(THREADS) xfft_fixed_kernel(SynthArgs a, const float2* __restrict__ V,
                                                             const signed char* __restrict__ bins,
                                                             const float2* __restrict__ twg, float2* __restrict__ U) {
  typedef FixedPlan<N1, N2> F;
  constexpr int n = F::n, LSC = SynthFixed<N1, N2>::LSC;
  const int R = a.R, P = 2 * R + 1, Q = R + 1;
  const int ky0 = blockIdx.x * NL, kz = blockIdx.y;
  const int nl = min(NL, P - ky0);
  extern __shared__ unsigned char smem[];
  float2* T1 = (float2*)smem;
  float2* buf = T1 + n;
  const int64_t lbase = ((int64_t)kz * P + ky0) * P;
  // lines with kz or every ky of the block beyond a shell's radius are zero
  // for that shell: skipped (the y pass never reads them)
  auto live = [&](int s) {
    const int rs = a.rs[s];
    return !(kz > rs || ky0 - R > rs || ky0 + nl - 1 - R < -rs);
  };
  bool any = false;
  for (int s = 0; s < a.nshell; ++s) any = any || live(s);
  if (!any) return;
  // the block's cube lines (values and bins) are read once, into registers,
  // and masked to each shell from there
  constexpr int KPT = (n + THREADS - 1) / THREADS;
  float2 cv[NL][KPT];
  int cb[NL][KPT];
#pragma unroll
  for (int l = 0; l < NL; ++l)
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      const int k = threadIdx.x + i * THREADS;
      const int f = k < (n >> 1) ? k : k - n;
      cv[l][i] = make_float2(0.f, 0.f);
      cb[l][i] = -1;
      if (l < nl && k < n && f >= -R && f <= R) {
        const int64_t j = lbase + (int64_t)l * P + f + R;
        cb[l][i] = bins[j];
        cv[l][i] = V[j];
      }
    }
  // the twiddle table's loads after the cube's (in-order completion: one
  // wait covers both)
  fixed_twiddles<N1, N2>(T1, twg);
  for (int s = 0; s < a.nshell; ++s) {
    if (!live(s)) continue;
    const int want = a.bin0 + s;
#pragma unroll
    for (int l = 0; l < NL; ++l)
#pragma unroll
      for (int i = 0; i < KPT; ++i) {
        const int k = threadIdx.x + i * THREADS;
        if (l < nl && k < n) buf[l * LSC + F::ipos(k)] = cb[l][i] == want ? cv[l][i] : make_float2(0.f, 0.f);
      }
    __syncthreads();
    fixed_fft_inplace<N1, N2, THREADS>(buf, T1, nl, LSC, R + 1);
    __syncthreads();
    // U[s][kz][ky block][x][ky % NL] (the fixed plans' layout): one
    // contiguous run of n NL entries per workgroup and shell
    // (two lines of one x per 16-byte store)
    static_assert(NL % 2 == 0, "the x pass stores line pairs");
    float4* dst = (float4*)(U + (((int64_t)s * Q + kz) * gridDim.x + blockIdx.x) * n * NL);
    for (int t = threadIdx.x; t < (NL / 2) * n; t += THREADS) {
      const int x = t / (NL / 2), l = 2 * (t - x * (NL / 2));
      if (l < nl) {
        const float2 a = buf[l * LSC + F::opos(x)];
        const float2 b = l + 1 < nl ? buf[(l + 1) * LSC + F::opos(x)] : make_float2(0.f, 0.f);
        stream_store(dst + t, a.x, a.y, b.x, b.y);
      }
    }
    __syncthreads();
  }
}

// y pass into the half-synthesised layout (tile-major or blocked, as
// half_blocked() chooses)
template <int N1, int N2, int THREADS, int NL, int NLX>
__global__ void __launch_bounds__(THREADS) yfft_fixed_kernel(SynthArgs a, const float2* __restrict__ U,
                                                             const float2* __restrict__ twg, float2* __restrict__ H,
                                                             int nslots, int slot0) {
  typedef FixedPlan<N1, N2> F;
  constexpr int n = F::n, LSC = SynthFixed<N1, N2>::LSC;
  const int R = a.R, P = 2 * R + 1, Q = R + 1;
  // grid (kz blocks, planes): the kz blocks of a plane, which write
  // adjacent segments of H, are dispatched together
  const int64_t plane = blockIdx.y;
  const int kz0 = blockIdx.x * NL;
  const int s = (int)(plane / n), x = (int)(plane - (int64_t)s * n);
  // kz lines beyond the shell's radius are zero: not computed, not written
  // (the contraction reads kz <= rs only)
  const int rs = a.rs[s];
  if (kz0 > rs) return;
  const int nl = min(NL, min(Q, rs + 1) - kz0);
  extern __shared__ unsigned char smem[];
  float2* T1 = (float2*)smem;
  float2* buf = T1 + n;
  // mode k of every line into its padded position: thread t takes k = t, t +
  // THREADS, ... of every line; line l is U[s][kz0 + l][ky / NLX][x][ky %
  // NLX] (the x pass's layout, NLX lines per block).  All loads are issued
  // before the twiddle table's (in-order completion: one wait covers both)
  constexpr int KPT = (n + THREADS - 1) / THREADS;
  const int nbx = (P + NLX - 1) / NLX;
  const int64_t lstride = (int64_t)nbx * n * NLX;
  const float2* src = U + (((int64_t)s * Q + kz0) * nbx * n + x) * NLX;
  int off[KPT];
  bool ok[KPT];
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    const int k = threadIdx.x + i * THREADS;
    const int f = k < (n >> 1) ? k : k - n, ky = f + R;
    ok[i] = k < n && f >= -rs && f <= rs;
    off[i] = ok[i] ? (ky / NLX) * n * NLX + ky % NLX : 0;
  }
  float2 v[NL][KPT];
#pragma unroll
  for (int l = 0; l < NL; ++l)
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      v[l][i] = make_float2(0.f, 0.f);
      if (l < nl && ok[i]) v[l][i] = src[l * lstride + off[i]];
    }
  fixed_twiddles<N1, N2>(T1, twg);
#pragma unroll
  for (int l = 0; l < NL; ++l)
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      const int k = threadIdx.x + i * THREADS;
      if (l < nl && k < n) buf[l * LSC + F::ipos(k)] = v[l][i];
    }
  __syncthreads();
  fixed_fft_inplace<N1, N2, THREADS>(buf, T1, nl, LSC, rs + 1);
  __syncthreads();
  // each y pair writes 2 nl consecutive entries, both y of a line as one
  // 16-byte streaming store: thread t -> pair t / NL, line t % NL.
  // Tile-major H[x][y/2][slot][kz][y%2]: pair yp at + yp nslots Q float4;
  // blocked H[x][slot][kz/8][y/2][kz%8][y%2]: this workgroup's (x, slot, kz
  // block) is one contiguous run, pair yp at + 8 yp float4
  static_assert(NL == 8, "the blocked layout holds kz blocks of NL = 8 lines");
  const bool blocked = half_blocked(n, Q);
  float4* dst = (float4*)(H + (blocked ? (((int64_t)x * nslots + slot0 + s) * ((Q + 7) / 8) + blockIdx.x) * n * 8
                                       : (((int64_t)x * (n >> 1) * nslots + slot0 + s) * Q + kz0) * 2));
  const int64_t ystride = blocked ? NL : (int64_t)nslots * Q;  // float4 per y pair
  for (int t = threadIdx.x; t < NL * (n >> 1); t += THREADS) {
    const int yp = t / NL, l = t - yp * NL;
    if (l < nl) {
      const float2 a = buf[l * LSC + F::opos(2 * yp)], b = buf[l * LSC + F::opos(2 * yp + 1)];
      stream_store(dst + yp * ystride + l, a.x, a.y, b.x, b.y);
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

// LDS bytes per FFT workgroup: the budget trades lines per workgroup (longer
// coalesced segments) against workgroups per CU (latency hiding)
#ifndef CNPC_SYNTH_LDS_BUDGET
#define CNPC_SYNTH_LDS_BUDGET (40 * 1024)
#endif

// lines per workgroup: twiddles + NL padded lines within the LDS budget, at
// most 16 lines, and at most MAXB butterflies per thread in every pass
int synth_lines_per_block(int prec, const SynthArgs& a) {
  const int n = a.n;
  const size_t cs = prec == 0 ? 16 : 8;
  const size_t budget = CNPC_SYNTH_LDS_BUDGET;
  if (budget <= (size_t)n * cs + (size_t)(n + 1) * cs) return 0;
  int nl = (int)((budget - (size_t)n * cs) / ((size_t)(n + 1) * cs));
  const int maxb = prec == 0 ? MaxB<double>::v : MaxB<float>::v;
  for (int s = 0; s < a.nrad; ++s) {
    // nl * n / R butterflies <= maxb * threads (one per thread for the composite radices)
    const int lim = (wide_radix(a.rad[s]) ? 1 : maxb) * SYNTH_THREADS * a.rad[s] / n;
    if (nl > lim) nl = lim;
  }
  const int cap = 16 * SYNTH_THREADS / 256;
  return nl > cap ? cap : nl;
}

// lines per workgroup of the fixed-plan passes (400 = 20 x 20): twiddles +
// NL padded lines within the LDS budget
#ifndef CNPC_SYNTH_FIXED_NL
#define CNPC_SYNTH_FIXED_NL 8
#endif
#ifndef CNPC_SYNTH_FIXED_THREADS
#define CNPC_SYNTH_FIXED_THREADS 256
#endif
// the x pass's own shape (lines per workgroup, threads)
#ifndef CNPC_SYNTH_FIXED_NL_X
#define CNPC_SYNTH_FIXED_NL_X CNPC_SYNTH_FIXED_NL
#endif
#ifndef CNPC_SYNTH_FIXED_THREADS_X
#define CNPC_SYNTH_FIXED_THREADS_X CNPC_SYNTH_FIXED_THREADS
#endif

template <typename T>
static hipError_t launch_synth_t(const SynthArgs& a, const BinEdges& E, const PolySpec& Wp, const void* D, void* V,
                                 signed char* bins, void* tw, void* work, void* out, hipStream_t st) {
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
  const int NL = synth_lines_per_block(prec, a);
  const size_t lds = sizeof(C) * ((size_t)n + (size_t)NL * (n + 1));
  const int64_t planes = (int64_t)a.nshell * n;
  const dim3 fblk(SYNTH_THREADS);
  hipLaunchKernelGGL(xfft_kernel<T>, dim3((P + NL - 1) / NL, Q), fblk, lds, st, a, NL, (const C*)V, bins,
                     (const C*)tw, U);
  hipLaunchKernelGGL(yfft_kernel<T>, dim3((unsigned)planes, (Q + NL - 1) / NL), fblk, lds, st, a, NL, (const C*)U,
                     (const C*)tw, Tm, 0, 0);
  hipLaunchKernelGGL(zfft_kernel<T>, dim3((unsigned)planes, (n + 2 * NL - 1) / (2 * NL)), fblk, lds, st, a, NL,
                     (const C*)Tm, (const C*)tw, (T*)out);
  return hipGetLastError();
}

hipError_t launch_synth(int prec, const SynthArgs& a, const BinEdges& E, const PolySpec& W, const void* D, void* V,
                        signed char* bins, void* tw, void* work, void* out, hipStream_t st) {
  if (prec == 0) return launch_synth_t<double>(a, E, W, D, V, bins, tw, work, out, st);
  return launch_synth_t<float>(a, E, W, D, V, bins, tw, work, out, st);
}

hipError_t launch_twiddles(int n, void* tw, hipStream_t st) {
  hipLaunchKernelGGL(twiddle_kernel<float>, dim3((n + 255) / 256), dim3(256), 0, st, n, (float2*)tw);
  return hipGetLastError();
}

// x and y passes only (FP32), into slots slot0 .. slot0 + nshell - 1 of the
// half-synthesised layout
hipError_t launch_synth_half(const SynthArgs& a, const BinEdges& E, const PolySpec& Wp, const void* D, void* V,
                             signed char* bins, void* tw, void* work, void* out, int nslots, int slot0,
                             int reuse_cube, hipStream_t st) {
  typedef float2 C;
  const int R = a.R, P = 2 * R + 1, Q = R + 1, n = a.n;
  const int64_t cube = (int64_t)P * Q * P;
  dim3 blk(256);
  if (!reuse_cube) {  // twiddles, w(khat) delta(k) and bins of the cube: shared by every shell
    hipLaunchKernelGGL(twiddle_kernel<float>, dim3((n + 255) / 256), blk, 0, st, n, (C*)tw);
    hipLaunchKernelGGL(cube_kernel<float>, dim3(grid_for(cube, 256, 1 << 14)), blk, 0, st, a, E, Wp, (const C*)D,
                       (C*)V, bins);
  }
  const int64_t planes = (int64_t)a.nshell * n;
  const dim3 fblk(SYNTH_THREADS);
  if (n == 400) {  // fixed plan 400 = 20 x 20
    constexpr int NL = CNPC_SYNTH_FIXED_NL, FT = CNPC_SYNTH_FIXED_THREADS;
    constexpr int NLX = CNPC_SYNTH_FIXED_NL_X, FTX = CNPC_SYNTH_FIXED_THREADS_X;
    const size_t ldsx = sizeof(C) * ((size_t)n + (size_t)NLX * SynthFixed<20, 20>::LSC);
    const size_t lds = sizeof(C) * ((size_t)n + (size_t)NL * SynthFixed<20, 20>::LSC);
    hipLaunchKernelGGL((xfft_fixed_kernel<20, 20, FTX, NLX>), dim3((P + NLX - 1) / NLX, Q), dim3(FTX), ldsx, st, a,
                       (const C*)V, bins, (const C*)tw, (C*)work);
    hipLaunchKernelGGL((yfft_fixed_kernel<20, 20, FT, NL, NLX>), dim3((Q + NL - 1) / NL, (unsigned)planes), dim3(FT),
                       lds, st, a, (const C*)work, (const C*)tw, (C*)out, nslots, slot0);
    return hipGetLastError();
  }
  const int NL = synth_lines_per_block(1, a);
  const size_t lds = sizeof(C) * ((size_t)n + (size_t)NL * (n + 1));
  hipLaunchKernelGGL(xfft_kernel<float>, dim3((P + NL - 1) / NL, Q), fblk, lds, st, a, NL, (const C*)V, bins,
                     (const C*)tw, (C*)work);
  hipLaunchKernelGGL(yfft_kernel<float>, dim3((unsigned)planes, (Q + NL - 1) / NL), fblk, lds, st, a, NL,
                     (const C*)work, (const C*)tw, (C*)out, nslots, slot0);
  return hipGetLastError();
}

}  // namespace cnpc
