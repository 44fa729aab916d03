// In-place mixed-radix Stockham inverse FFTs over lines held in LDS, shared
// by the shell synthesis passes and the fused z-synthesis contraction.
//
// The transforms are the unnormalised inverse e^{+i k.x} of the reference's
// pmesh c2r.  Lines are padded to n+1 entries so that accesses striding over
// lines spread over the LDS banks; the twiddles of all passes sit in one LDS
// table of n - 1 entries (build_pass_twiddles).
#pragma once
#include "cnpc_common.h"

namespace cnpc {

template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// floor(a / b) for 0 <= a < 2^22 and 1 <= b <= 2^12 from a float reciprocal,
// corrected to the exact quotient (the float estimate is within one).
__device__ __forceinline__ int fdiv(int a, int b, float inv_b) {
  int q = (int)(((float)a + 0.5f) * inv_b);
  if (q * b > a) --q;
  else if ((q + 1) * b <= a) ++q;
  return q;
}

// ------------------------------------------------------------------ LDS FFT
// R-point inverse DFTs v[b] = sum_q u[q] e^{+2 pi i q b / R} in registers,
// written out for R = 2, 3, 4, 5 with the fewest multiplications
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) {
  C r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}
template <typename C>
__device__ __forceinline__ C csub(C a, C b) {
  C r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}
// a + i s d  and  a - i s d
template <typename C, typename T>
__device__ __forceinline__ void pm_i(C a, C d, T s, C& plus, C& minus) {
  plus.x = a.x - s * d.y;
  plus.y = a.y + s * d.x;
  minus.x = a.x + s * d.y;
  minus.y = a.y - s * d.x;
}

template <int RR> struct SmallDft;
template <> struct SmallDft<2> {
  template <typename C> __device__ __forceinline__ static void run(C (&u)[2]) {
    const C a = u[0], b = u[1];
    u[0] = cadd(a, b);
    u[1] = csub(a, b);
  }
};
template <> struct SmallDft<3> {
  template <typename C> __device__ __forceinline__ static void run(C (&u)[3]) {
    typedef decltype(C().x) T;
    const T s = (T)0.86602540378443865;
    const C t = cadd(u[1], u[2]), d = csub(u[1], u[2]);
    C m;
    m.x = u[0].x - (T)0.5 * t.x;
    m.y = u[0].y - (T)0.5 * t.y;
    u[0] = cadd(u[0], t);
    pm_i(m, d, s, u[1], u[2]);
  }
};
template <> struct SmallDft<4> {
  template <typename C> __device__ __forceinline__ static void run(C (&u)[4]) {
    typedef decltype(C().x) T;
    const C a = cadd(u[0], u[2]), b = csub(u[0], u[2]);
    const C c = cadd(u[1], u[3]), d = csub(u[1], u[3]);
    u[0] = cadd(a, c);
    u[2] = csub(a, c);
    pm_i(b, d, (T)1, u[1], u[3]);
  }
};
template <> struct SmallDft<5> {
  template <typename C> __device__ __forceinline__ static void run(C (&u)[5]) {
    typedef decltype(C().x) T;
    const T c1 = (T)0.30901699437494742, c2 = (T)-0.80901699437494742;
    const T s1 = (T)0.95105651629515357, s2 = (T)0.58778525229247313;
    const C t1 = cadd(u[1], u[4]), t2 = cadd(u[2], u[3]);
    const C d1 = csub(u[1], u[4]), d2 = csub(u[2], u[3]);
    C a1, a2, b1, b2;
    a1.x = u[0].x + c1 * t1.x + c2 * t2.x;
    a1.y = u[0].y + c1 * t1.y + c2 * t2.y;
    a2.x = u[0].x + c2 * t1.x + c1 * t2.x;
    a2.y = u[0].y + c2 * t1.y + c1 * t2.y;
    b1.x = s1 * d1.x + s2 * d2.x;
    b1.y = s1 * d1.y + s2 * d2.y;
    b2.x = s2 * d1.x - s1 * d2.x;
    b2.y = s2 * d1.y - s1 * d2.y;
    u[0] = cadd(u[0], cadd(t1, t2));
    pm_i(a1, b1, (T)1, u[1], u[4]);
    pm_i(a2, b2, (T)1, u[2], u[3]);
  }
};

// cos / sin of 2 pi j / N at compile time (range-reduced Taylor series,
// double precision), for the internal twiddles of the composite DFTs
constexpr double ct_pi = 3.14159265358979323846;
constexpr double ct_sin_red(double x) {  // |x| <= pi/4
  double t = x, s = x;
  for (int k = 1; k < 14; ++k) {
    t *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    s += t;
  }
  return s;
}
constexpr double ct_cos_red(double x) {
  double t = 1.0, s = 1.0;
  for (int k = 1; k < 14; ++k) {
    t *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    s += t;
  }
  return s;
}
// e^{2 pi i j / N}: octant reduction on j / N
constexpr double ct_cos(int j, int N) {
  const int m = ((j % N) + N) % N;
  const int o = (8 * m) / N;  // octant
  const double x = 2.0 * ct_pi * m / N;
  return o == 0 || o == 7 ? ct_cos_red(o == 0 ? x : x - 2.0 * ct_pi)
       : o == 1 || o == 2 ? -ct_sin_red(x - 0.5 * ct_pi)
       : o == 3 || o == 4 ? -ct_cos_red(x - ct_pi)
                          : ct_sin_red(x - 1.5 * ct_pi);
}
constexpr double ct_sin(int j, int N) {
  const int m = ((j % N) + N) % N;
  const int o = (8 * m) / N;
  const double x = 2.0 * ct_pi * m / N;
  return o == 0 || o == 7 ? ct_sin_red(o == 0 ? x : x - 2.0 * ct_pi)
       : o == 1 || o == 2 ? ct_cos_red(x - 0.5 * ct_pi)
       : o == 3 || o == 4 ? -ct_sin_red(x - ct_pi)
                          : -ct_cos_red(x - 1.5 * ct_pi);
}

// e^{2 pi i j / N} for j < N as compile-time tables
template <int N> struct TwTable {
  double c[N], s[N];
  constexpr TwTable() : c(), s() {
    for (int j = 0; j < N; ++j) {
      c[j] = ct_cos(j, N);
      s[j] = ct_sin(j, N);
    }
  }
};
template <int N> struct TwConst { static constexpr TwTable<N> t{}; };

// Composite DFT of N = N1 N2 points in registers (Cooley-Tukey): with
// q = N2 n1 + n2 and k = k1 + N1 k2, X[k] = sum_n2 w_N^{n2 k1} w_N2^{n2 k2}
// sum_n1 u[q] w_N1^{n1 k1}.
template <int N1, int N2> struct CompositeDft {
  template <typename C> __device__ __forceinline__ static void run(C (&u)[N1 * N2]) {
    typedef decltype(C().x) T;
    C y[N2][N1];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
      C a[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) a[n1] = u[N2 * n1 + n2];
      SmallDft<N1>::run(a);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) {
        if (n2 * k1 == 0) {
          y[n2][k1] = a[k1];
        } else {
          C w;
          w.x = (T)TwConst<N1 * N2>::t.c[n2 * k1];
          w.y = (T)TwConst<N1 * N2>::t.s[n2 * k1];
          y[n2][k1] = cmul(a[k1], w);
        }
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      C b[N2];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) b[n2] = y[n2][k1];
      SmallDft<N2>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) u[k1 + N1 * k2] = b[k2];
    }
  }
};
template <> struct SmallDft<16> : CompositeDft<4, 4> {};

// u i^m (m a compile-time constant after unrolling)
template <typename C>
__device__ __forceinline__ C mul_ipow(C u, int m) {
  C r;
  switch (m & 3) {
    case 0: r = u; break;
    case 1: r.x = -u.y; r.y = u.x; break;
    case 2: r.x = -u.x; r.y = -u.y; break;
    default: r.x = u.y; r.y = -u.x; break;
  }
  return r;
}

// 20-point DFT by the prime-factor (Good-Thomas) algorithm, 20 = 4 x 5 with
// no twiddles: input q = (5 n1 + 4 n2) mod 20, output k = (5 k1 + 16 k2)
// mod 20 (k1 = k mod 4, k2 = k mod 5), X[k] = sum_n2 w5^{n2 k2} sum_n1
// w4^{n1 k1} u[q].  ZMID: the inputs q = 5 .. 14 are zero, and each 4-point
// DFT has two nonzero inputs (sums of two terms times powers of i).
template <bool ZMID>
struct Pfa20 {
  template <typename C> __device__ __forceinline__ static void run(C (&u)[20]) {
    C y[5][4];
#pragma unroll
    for (int n2 = 0; n2 < 5; ++n2) {
      if (!ZMID) {
        C a[4];
#pragma unroll
        for (int n1 = 0; n1 < 4; ++n1) a[n1] = u[(5 * n1 + 4 * n2) % 20];
        SmallDft<4>::run(a);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = a[k1];
      } else {
        int p0 = -1, p1 = -1;
#pragma unroll
        for (int n1 = 0; n1 < 4; ++n1) {
          const int q = (5 * n1 + 4 * n2) % 20;
          if (q < 5 || q > 14) {
            if (p0 < 0) p0 = n1;
            else p1 = n1;
          }
        }
        const C a0 = u[(5 * p0 + 4 * n2) % 20], a1 = u[(5 * p1 + 4 * n2) % 20];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = cadd(mul_ipow(a0, p0 * k1), mul_ipow(a1, p1 * k1));
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      C b[5];
#pragma unroll
      for (int n2 = 0; n2 < 5; ++n2) b[n2] = y[n2][k1];
      SmallDft<5>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < 5; ++k2) u[(5 * k1 + 16 * k2) % 20] = b[k2];
    }
  }
};
template <> struct SmallDft<20> : Pfa20<false> {};
template <> struct SmallDft<25> : CompositeDft<5, 5> {};

// Twiddles of every pass in pass order, pass s (radix R, stride p) at offset
// off_s = sum_{t<s} (R_t - 1) p_t: entry (r - 1) p + k = e^{2 pi i r k/(p R)}
// for r in [1, R), k in [0, p).  Consecutive butterflies (consecutive k)
// read consecutive entries, so the twiddle reads are free of bank conflicts.
// The table has n - 1 entries.
template <typename C>
__device__ void build_pass_twiddles(C* __restrict__ Wt, const C* __restrict__ twg, int n, const SynthArgs& a) {
  int off = 0, p = 1;
  for (int s = 0; s < a.nrad; ++s) {
    const int R = a.rad[s], cnt = (R - 1) * p, tstep = n / (p * R);
    for (int j = threadIdx.x; j < cnt; j += blockDim.x) {
      const int r = j / p + 1, k = j - (r - 1) * p;
      Wt[off + j] = twg[r * k * tstep];
    }
    off += cnt;
    p *= R;
  }
}

// One in-place pass of the mixed-radix Stockham inverse FFT (e^{+2 pi i})
// over nl lines of length n, line l at buf + l*ls.  Butterfly i in [0, n/R)
// reads buf[i + r n/R], multiplies by e^{2 pi i r k/(p R)} (k = i mod p;
// none in the first pass, p = 1), does an R-point DFT and writes
// buf[(i/p) p R + k + r p].  All reads are staged in registers before the
// barrier, so one buffer suffices; a thread owns at most MAXB butterflies
// (the host sizes nl accordingly).  Tw: this pass's twiddles.
// PLANAR: write the pass's output as two planes per line, real parts at
// floats [0, n) and imaginary parts at floats [ls, ls + n) of the line's
// 2 ls floats (the consumer then reads runs of one part with vector loads).
template <int RR, int MAXB, bool PLANAR = false, typename C>
__device__ __forceinline__ void stockham_pass(C* __restrict__ buf, const C* __restrict__ Tw, int n, int ls, int nl,
                                              int p) {
  const int Tn = n / RR;
  const int pR = p * RR;
  const int total = nl * Tn;
  const float inv_Tn = 1.0f / (float)Tn, inv_p = 1.0f / (float)p;
  C u[MAXB][RR];
  int dsto[MAXB], dline[MAXB];
#pragma unroll
  for (int it = 0; it < MAXB; ++it) {
    const int t = threadIdx.x + it * blockDim.x;
    dsto[it] = -1;
    dline[it] = 0;
    if (t < total) {
      const int l = fdiv(t, Tn, inv_Tn), i = t - l * Tn;
      const int q = fdiv(i, p, inv_p), k = i - q * p;
      const C* s = buf + l * ls;
#pragma unroll
      for (int r = 0; r < RR; ++r) u[it][r] = s[i + r * Tn];
      if (p > 1) {
#pragma unroll
        for (int r = 1; r < RR; ++r) u[it][r] = cmul(u[it][r], Tw[(r - 1) * p + k]);
      }
      dsto[it] = q * pR + k;
      dline[it] = l * ls;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < MAXB; ++it) {
    if (dsto[it] >= 0) {
      SmallDft<RR>::run(u[it]);
      if (PLANAR) {
        typedef decltype(C().x) T;
        T* f = (T*)buf + 2 * dline[it] + dsto[it];
#pragma unroll
        for (int b = 0; b < RR; ++b) {
          f[b * p] = u[it][b].x;
          f[ls + b * p] = u[it][b].y;
        }
      } else {
        C* d = buf + dline[it] + dsto[it];
#pragma unroll
        for (int b = 0; b < RR; ++b) d[b * p] = u[it][b];
      }
    }
  }
  __syncthreads();
}

template <int MAXB, typename C>
__device__ __forceinline__ void stockham_single(C* buf, const C* Tw, int n, int ls, int nl, int R, int p) {
  if (R == 4) stockham_pass<4, MAXB>(buf, Tw, n, ls, nl, p);
  else if (R == 2) stockham_pass<2, MAXB>(buf, Tw, n, ls, nl, p);
  else if (R == 3) stockham_pass<3, MAXB>(buf, Tw, n, ls, nl, p);
  else stockham_pass<5, MAXB>(buf, Tw, n, ls, nl, p);
}

// the same with the composite radices 16, 20 and 25 as well (one such
// butterfly per thread: the host checks nl * n / R <= blockDim.x)
template <int MAXB, typename C>
__device__ __forceinline__ void stockham_single_wide(C* buf, const C* Tw, int n, int ls, int nl, int R, int p) {
  if (R == 16) stockham_pass<16, 1>(buf, Tw, n, ls, nl, p);
  else if (R == 20) stockham_pass<20, 1>(buf, Tw, n, ls, nl, p);
  else if (R == 25) stockham_pass<25, 1>(buf, Tw, n, ls, nl, p);
  else stockham_single<MAXB>(buf, Tw, n, ls, nl, R, p);
}

// a last pass writing planar output (see stockham_pass)
template <int MAXB, typename C>
__device__ __forceinline__ void stockham_last_planar(C* buf, const C* Tw, int n, int ls, int nl, int R, int p) {
  if (R == 16) stockham_pass<16, 1, true>(buf, Tw, n, ls, nl, p);
  else if (R == 20) stockham_pass<20, 1, true>(buf, Tw, n, ls, nl, p);
  else if (R == 25) stockham_pass<25, 1, true>(buf, Tw, n, ls, nl, p);
  else if (R == 4) stockham_pass<4, MAXB, true>(buf, Tw, n, ls, nl, p);
  else if (R == 2) stockham_pass<2, MAXB, true>(buf, Tw, n, ls, nl, p);
  else if (R == 3) stockham_pass<3, MAXB, true>(buf, Tw, n, ls, nl, p);
  else stockham_pass<5, MAXB, true>(buf, Tw, n, ls, nl, p);
}

// Whole inverse FFT of nl lines (composite radices admitted), the last pass
// writing planar lines.
template <int MAXB, typename C>
__device__ void stockham_lds_planar(C* buf, const C* Wt, int n, int ls, int nl, const SynthArgs& a) {
  int p = 1, off = 0;
  for (int s = 0; s < a.nrad; ++s) {
    if (s + 1 == a.nrad)
      stockham_last_planar<MAXB>(buf, Wt + off, n, ls, nl, a.rad[s], p);
    else
      stockham_single_wide<MAXB>(buf, Wt + off, n, ls, nl, a.rad[s], p);
    off += (a.rad[s] - 1) * p;
    p *= a.rad[s];
  }
}


// Whole inverse FFT of nl lines; Wt from build_pass_twiddles.  WIDE admits
// the composite radices in a.rad.
template <int MAXB, bool WIDE = false, typename C>
__device__ void stockham_lds(C* buf, const C* Wt, int n, int ls, int nl, const SynthArgs& a) {
  int p = 1, off = 0;
  for (int s = 0; s < a.nrad; ++s) {
    if (WIDE)
      stockham_single_wide<MAXB>(buf, Wt + off, n, ls, nl, a.rad[s], p);
    else
      stockham_single<MAXB>(buf, Wt + off, n, ls, nl, a.rad[s], p);
    off += (a.rad[s] - 1) * p;
    p *= a.rad[s];
  }
}

// ------------------------------------------------------------ fixed plans
// n = N1 * N2 with both factors radices of SmallDft (e.g. 400 = 20 x 20): an
// inverse FFT as two in-place passes over a line of N1 rows of RW complex
// (RW = N2 + 1 by default, N2 for unpadded rows).  Input mode k = k2 + N2 k1
// sits at position k2 + RW k1.
//   pass 1, butterfly (line, k2): the N1-point DFT over k1 (stride RW),
//     times e^{2 pi i k2 z1 / n}, written back in place (row z1);
//   pass 2, butterfly (line, z1): the N2-point DFT over the contiguous row
//     z1, written back in place: output z = z1 + N1 z2 at position z2 + RW z1.
// Each butterfly reads and writes only its own positions, so a pass needs no
// barrier between its reads and writes; padded rows start RW (odd for even
// N2) entries apart, which spreads lanes that walk the rows of one line over
// the LDS banks.  The pad column (k2 = N2) is never read.  Consumers either
// read the output in this order (the fused contraction, with G staged alike)
// or gather it by opos(z).
template <int N1, int N2, int RWP = N2 + 1>
struct FixedPlan {
  static constexpr int n = N1 * N2;
  static constexpr int RW = RWP;
  static constexpr int KP = N1 * RW;  // positions per line
  __device__ static int ipos(int k) { return (k % N2) + RW * (k / N2); }  // input mode k
  __device__ static int opos(int z) { return (z / N1) + RW * (z % N1); }  // output z
};

// twiddles of pass 1: T1[z1 N2 + k2] = e^{2 pi i k2 z1 / n}, from the table
// tw[j] = e^{2 pi i j / n}
template <int N1, int N2, typename C>
__device__ __forceinline__ void fixed_twiddles(C* __restrict__ T1, const C* __restrict__ tw) {
  constexpr int n = N1 * N2;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const int z1 = j / N2, k2 = j - z1 * N2;
    T1[j] = tw[(k2 * z1) % n];
  }
}

// Streaming 16-byte store (nontemporal): for outputs written once for a
// later kernel; scattered pieces written this way reach HBM faster than
// through the cache hierarchy.
__device__ __forceinline__ void stream_store(float4* p, float a, float b, float c, float d) {
  typedef float f32x4s __attribute__((ext_vector_type(4)));
  const f32x4s v = {a, b, c, d};
  __builtin_nontemporal_store(v, (f32x4s*)p);
}

// Workgroup barrier that orders LDS only: waits for this wave's LDS
// operations, not for its global loads, so loads issued earlier stay in
// flight across it (a __syncthreads() fence would wait for them too).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// First-pass butterfly (line at col - k2, column k2) of a fixed plan: the
// N1-point DFT over the column's inputs (stride RW) times e^{2 pi i k2 z1 /
// n}, in place.  qz > 0: the modes qz .. n - qz are zero and not read
// (whatever the positions hold; for 20 x 20 plans with qz <= n / 4 the
// 20-point DFT also skips their arithmetic); qz <= 0: every mode is read.
// BANDZ (20 x 20, qz <= n / 4): the caller has written zeros at the modes qz
// .. n/4 - 1 and 3n/4 .. n - qz, so the butterfly reads its ten band inputs
// (k1 = 0 .. 4, 15 .. 19) without per-mode conditions.
#ifndef FFT_TW_GROUP
#define FFT_TW_GROUP 10
#endif
template <int N1, int N2, bool BANDZ, int RW, typename C>
__device__ __forceinline__ void fixed_pass1(C* __restrict__ col, const C* __restrict__ T1, int k2, int qz) {
  constexpr int n = N1 * N2;
  C u[N1];
  bool pruned = false;
  if constexpr (N1 == 20) {
    if (qz > 0 && 4 * qz <= n) {
      // modes k2 + N2 k1 with k1 = 5 .. 14 lie in [qz, n - qz]: zero, not read
      pruned = true;
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) {
        const int k = k2 + N2 * k1;
        if (k1 >= 5 && k1 <= 14) {
          u[k1].x = 0;
          u[k1].y = 0;
        } else if (!BANDZ && k >= qz && k <= n - qz) {
          u[k1].x = 0;
          u[k1].y = 0;
        } else {
          u[k1] = col[RW * k1];
        }
      }
      Pfa20<true>::run(u);
    }
  }
  if (!pruned) {
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      const int k = k2 + N2 * k1;
      if (qz > 0 && k >= qz && k <= n - qz) {
        u[k1].x = 0;
        u[k1].y = 0;
      } else {
        u[k1] = col[RW * k1];
      }
    }
    SmallDft<N1>::run(u);
  }
  // the write-back reads its twiddles in groups of TWG before the group's
  // stores, so it waits once per group (read between the stores, each
  // twiddle would wait for its own LDS round trip); groups bound the
  // registers the twiddles hold
  col[0] = u[0];
  constexpr int TWG = FFT_TW_GROUP;
#pragma unroll
  for (int z0 = 1; z0 < N1; z0 += TWG) {
    C tw[TWG];
#pragma unroll
    for (int j = 0; j < TWG; ++j)
      if (z0 + j < N1) tw[j] = T1[(z0 + j) * N2 + k2];
#pragma unroll
    for (int j = 0; j < TWG; ++j)
      if (z0 + j < N1) col[RW * (z0 + j)] = cmul(u[z0 + j], tw[j]);
  }
}

// Second-pass butterfly: the N2-point DFT over the contiguous row z1 of a
// line (row = line + RW z1), in place (output z = z1 + N1 z2 at position z2 +
// RW z1)
template <int N2, typename C>
__device__ __forceinline__ void fixed_pass2(C* __restrict__ row) {
  C u[N2];
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) u[k2] = row[k2];
  SmallDft<N2>::run(u);
#pragma unroll
  for (int z2 = 0; z2 < N2; ++z2) row[z2] = u[z2];
}

// Both passes over nl lines of pitch lsc (complex); the caller synchronises
// before (inputs written) and after (outputs read).  Thread t takes line t %
// nl of butterfly t / nl: lanes of a wave walk the lines first, whose pitch
// spreads them over the LDS banks.
template <int N1, int N2, int THREADS, bool BANDZ = false, int RW = N2 + 1, typename C>
__device__ __forceinline__ void fixed_fft_inplace(C* __restrict__ buf, const C* __restrict__ T1, int nl, int lsc,
                                                  int qz = 0) {
  const float inv_nl = 1.0f / (float)nl;
  for (int t = threadIdx.x; t < nl * N2; t += THREADS) {
    const int k2 = fdiv(t, nl, inv_nl), l = t - k2 * nl;
    fixed_pass1<N1, N2, BANDZ, RW>(buf + l * lsc + k2, T1, k2, qz);
  }
  lds_barrier();
  for (int t = threadIdx.x; t < nl * N1; t += THREADS) {
    const int z1 = fdiv(t, nl, inv_nl), l = t - z1 * nl;
    fixed_pass2<N2>(buf + l * lsc + RW * z1);
  }
}

// The same with at most one butterfly per thread and pass, its place
// computed once by the caller (FixedSlots): for launches that run the same
// line count round after round.
struct FixedSlots {
  int c1, k2;  // first pass: column offset (line * lsc + k2) and k2, c1 < 0: none
  int r2;      // second pass: row offset (line * lsc + RW z1), < 0: none
};
template <int N1, int N2, int RW>
__device__ __forceinline__ FixedSlots fixed_slots(int nl, int lsc) {
  FixedSlots s;
  const int t = threadIdx.x;
  s.k2 = t / nl;
  s.c1 = t < nl * N2 ? (t - s.k2 * nl) * lsc + s.k2 : -1;
  const int z1 = t / nl;
  s.r2 = t < nl * N1 ? (t - z1 * nl) * lsc + RW * z1 : -1;
  return s;
}
template <int N1, int N2, bool BANDZ, int RW, typename C>
__device__ __forceinline__ void fixed_fft_slots(C* __restrict__ buf, const C* __restrict__ T1, const FixedSlots& s,
                                                int qz) {
  if (s.c1 >= 0) fixed_pass1<N1, N2, BANDZ, RW>(buf + s.c1, T1, s.k2, qz);
  lds_barrier();
  if (s.r2 >= 0) fixed_pass2<N2>(buf + s.r2);
}

// butterflies per thread of a pass (bounds the registers of the staged pass)
template <typename T> struct MaxB { static constexpr int v = 4; };

// global loads a thread issues before writing them to LDS
constexpr int UNR = 8;

}  // namespace cnpc
