// In-place mixed-radix Stockham inverse FFTs over lines held in LDS, shared
// by the shell synthesis passes and the fused z-synthesis contraction.
//
// The transforms are the unnormalised inverse e^{+i k.x} of the reference's
// pmesh c2r.  Lines are padded to n+1 entries so that accesses striding over
// lines spread over the LDS banks; twiddles come from one table
// W[j] = e^{2 pi i j / n} in LDS.
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
// over nl lines of length n, line l at buf + l*ls.  Butterfly i in [0, n/R)
// reads buf[i + r n/R], multiplies by e^{2 pi i r k/(p R)} (k = i mod p), does
// an R-point DFT and writes buf[(i/p) p R + k + r p].  All reads are staged in
// registers before the barrier, so one buffer suffices; a thread owns at most
// MAXB butterflies (the host sizes nl accordingly).
template <int RR, int MAXB, typename C>
__device__ __forceinline__ void stockham_pass(C* __restrict__ buf, const C* __restrict__ W, int n, int ls, int nl,
                                              int p) {
  typedef decltype(C().x) T;
  const int Tn = n / RR;
  const int pR = p * RR;
  const int tstep = n / pR;
  const int total = nl * Tn;
  const float inv_Tn = 1.0f / (float)Tn, inv_p = 1.0f / (float)p;
  C u[MAXB][RR];
  int dsto[MAXB];
#pragma unroll
  for (int it = 0; it < MAXB; ++it) {
    const int t = threadIdx.x + it * blockDim.x;
    dsto[it] = -1;
    if (t < total) {
      const int l = fdiv(t, Tn, inv_Tn), i = t - l * Tn;
      const int q = fdiv(i, p, inv_p), k = i - q * p;
      const C* s = buf + l * ls;
#pragma unroll
      for (int r = 0; r < RR; ++r) {
        u[it][r] = s[i + r * Tn];
        if (r) u[it][r] = cmul(u[it][r], W[r * k * tstep]);
      }
      dsto[it] = l * ls + q * pR + k;
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
__device__ __forceinline__ void stockham_single(C* buf, const C* W, int n, int ls, int nl, int R, int p) {
  if (R == 4) stockham_pass<4, MAXB>(buf, W, n, ls, nl, p);
  else if (R == 2) stockham_pass<2, MAXB>(buf, W, n, ls, nl, p);
  else if (R == 3) stockham_pass<3, MAXB>(buf, W, n, ls, nl, p);
  else stockham_pass<5, MAXB>(buf, W, n, ls, nl, p);
}

// Whole inverse FFT of nl lines.
template <int MAXB, typename C>
__device__ void stockham_lds(C* buf, const C* W, int n, int ls, int nl, const SynthArgs& a) {
  int p = 1;
  for (int s = 0; s < a.nrad; ++s) {
    stockham_single<MAXB>(buf, W, n, ls, nl, a.rad[s], p);
    p *= a.rad[s];
  }
}

// butterflies per thread of a pass (bounds the registers of the staged pass)
template <typename T> struct MaxB { static constexpr int v = 4; };

// global loads a thread issues before writing them to LDS
constexpr int UNR = 8;

}  // namespace cnpc
