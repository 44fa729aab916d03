// 3-D complex-to-real inverse FFT of a 400^3 half spectrum (FP32) on the
// fixed 20 x 20 line plan of lds_fft.h, in place of the general library
// transform for the one grid the fused bispectrum path contracts on (the
// G_00 = c2r(delta) of clustering_estimator.py:845-851 after the resample to
// the contraction grid).  Same result as cnpc_fft_c2r: out(x) = sum_k S(k)
// e^{+2 pi i k.x / n} over the Hermitian completion of S, unnormalised.
//
// Three passes, each a read and a write of the spectrum:
//   1. y: c2c along y of every (x, kz) line, in place; a workgroup takes 16
//      consecutive kz of one x plane (128-byte pieces per y);
//   2. x: the same along x, in place;
//   3. z: c2r of every row, two rows per complex FFT (row y in the real part,
//      row y+1 in the imaginary part: L(k) = A(k) + i B(k) and L(n-k) =
//      conj A(k) + i conj B(k), the kz = 0 and n/2 terms taking the real
//      parts of A and B, as the c2r of a row does); a workgroup takes 16
//      row pairs, which are contiguous in the spectrum and in the output.
// Each pass's 20-point DFTs are one per thread (16 lines x 20 butterflies =
// 320 threads).

#include "cnpc_common.h"
#include "lds_fft.h"

namespace cnpc {

namespace {
constexpr int C4N = 400, C4H = C4N / 2 + 1;  // grid, kz per half-spectrum row
typedef FixedPlan<20, 20> C4P;
constexpr int C4LSC = C4P::KP + 2;  // line pitch (complex)
constexpr int C4NL = 16;            // row pairs per workgroup (z pass)
constexpr int C4T = C4NL * 20;      // threads (z pass)
#ifndef CNPC_C2R400_LINES
#define CNPC_C2R400_LINES 16
#endif
constexpr int C4L = CNPC_C2R400_LINES;  // kz lines per workgroup (y and x passes)
constexpr size_t c4_lds(int lines) { return sizeof(float2) * ((size_t)C4N + (size_t)lines * C4LSC); }
}  // namespace

// c2c pass over the lines kz0 .. kz0 + NL - 1 of block (blockIdx.y,
// blockIdx.z): position p of line l at base + p * pstride + l
template <int NL>
__global__ void __launch_bounds__(NL * 20) c2r400_lines_kernel(float2* __restrict__ S, int64_t pstride,
                                                               int64_t bstride, int64_t mstride,
                                                               const float2* __restrict__ twg) {
  constexpr int T = NL * 20;
  extern __shared__ unsigned char smem[];
  float2* T1 = (float2*)smem;
  float2* buf = T1 + C4N;
  const int kz0 = blockIdx.x * NL, nl = min(NL, C4H - kz0);
  float2* base = S + blockIdx.z * mstride + blockIdx.y * bstride + kz0;
  // item i of thread t: position t / NL + 20 i, line t % NL (NL lanes read
  // one position's lines as one contiguous piece)
  constexpr int IPT = C4N / 20;
  const int l = threadIdx.x % NL, p0 = threadIdx.x / NL;
  float2 v[IPT];
#pragma unroll
  for (int i = 0; i < IPT; ++i) v[i] = l < nl ? base[(p0 + 20 * i) * pstride + l] : make_float2(0.f, 0.f);
  fixed_twiddles<20, 20>(T1, twg);
  if (l < nl) {
#pragma unroll
    for (int i = 0; i < IPT; ++i) buf[l * C4LSC + C4P::ipos(p0 + 20 * i)] = v[i];
  }
  __syncthreads();
  fixed_fft_inplace<20, 20, T>(buf, T1, nl, C4LSC);
  __syncthreads();
  if (l < nl) {
#pragma unroll
    for (int i = 0; i < IPT; ++i) base[(p0 + 20 * i) * pstride + l] = buf[l * C4LSC + C4P::opos(p0 + 20 * i)];
  }
}

// z pass: row pairs q0 .. q0 + 15 (q = x * n/2 + y/2) of spectrum
// blockIdx.z into real rows
__global__ void __launch_bounds__(C4T) c2r400_rows_kernel(const float2* __restrict__ S, float* __restrict__ out,
                                                          int64_t min_stride, int64_t mout_stride,
                                                          const float2* __restrict__ twg) {
  extern __shared__ unsigned char smem[];
  float2* T1 = (float2*)smem;
  float2* buf = T1 + C4N;
  const int64_t q0 = (int64_t)blockIdx.x * C4NL;
  const float2* src = S + blockIdx.z * min_stride + q0 * 2 * C4H;
  float* dst = out + blockIdx.z * mout_stride + q0 * 2 * C4N;
  // item t: line j = t / 201, mode k = t % 201 of both rows
  constexpr int ITEMS = C4NL * C4H, IPT = (ITEMS + C4T - 1) / C4T;
  float2 a[IPT], b[IPT];
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const int t = threadIdx.x + i * C4T;
    const int j = t / C4H, k = t - j * C4H;
    if (t < ITEMS) {
      a[i] = src[(2 * j) * C4H + k];
      b[i] = src[(2 * j + 1) * C4H + k];
    }
  }
  fixed_twiddles<20, 20>(T1, twg);
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const int t = threadIdx.x + i * C4T;
    const int j = t / C4H, k = t - j * C4H;
    if (t < ITEMS) {
      float2* line = buf + j * C4LSC;
      if (k == 0 || k == C4N / 2) {
        line[C4P::ipos(k)] = make_float2(a[i].x, b[i].x);
      } else {
        line[C4P::ipos(k)] = make_float2(a[i].x - b[i].y, a[i].y + b[i].x);
        line[C4P::ipos(C4N - k)] = make_float2(a[i].x + b[i].y, b[i].x - a[i].y);
      }
    }
  }
  __syncthreads();
  fixed_fft_inplace<20, 20, C4T>(buf, T1, C4NL, C4LSC);
  __syncthreads();
  // item t: line j = t / 100, z = 4 (t % 100) .. + 3 of both rows
  for (int t = threadIdx.x; t < C4NL * (C4N / 4); t += C4T) {
    const int j = t / (C4N / 4), z = 4 * (t - j * (C4N / 4));
    const float2* line = buf + j * C4LSC;
    const float2 v0 = line[C4P::opos(z)], v1 = line[C4P::opos(z + 1)];
    const float2 v2 = line[C4P::opos(z + 2)], v3 = line[C4P::opos(z + 3)];
    stream_store((float4*)(dst + (2 * j) * C4N + z), v0.x, v1.x, v2.x, v3.x);
    stream_store((float4*)(dst + (2 * j + 1) * C4N + z), v0.y, v1.y, v2.y, v3.y);
  }
}

// dynamic LDS beyond 64 KiB needs the kernel's attribute raised, once per
// device
template <int NL>
static hipError_t c4_allow_lds() {
  constexpr size_t lds = c4_lds(NL);
  if (lds <= 65536) return hipSuccess;
  static size_t allowed[CNPC_MAX_DEVICES] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= CNPC_MAX_DEVICES) return hipErrorInvalidDevice;
  if (allowed[dev] < lds) {
    e = hipFuncSetAttribute((const void*)c2r400_lines_kernel<NL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
    if (e != hipSuccess) return e;
    allowed[dev] = lds;
  }
  return hipSuccess;
}

// in: batch half spectra 400 x 400 x 201 (overwritten); out: batch real
// 400^3 meshes; tw: e^{2 pi i j / 400}, j < 400
hipError_t launch_c2r400(int batch, void* in, void* out, const void* tw, hipStream_t st) {
  if (hipError_t e = c4_allow_lds<C4L>()) return e;
  const int64_t plane = (int64_t)C4N * C4N * C4H;
  const dim3 grid((C4H + C4L - 1) / C4L, C4N, batch);
  float2* S = (float2*)in;
  const float2* T = (const float2*)tw;
  hipLaunchKernelGGL(c2r400_lines_kernel<C4L>, grid, dim3(C4L * 20), c4_lds(C4L), st, S, (int64_t)C4H,
                     (int64_t)C4N * C4H, plane, T);
  hipLaunchKernelGGL(c2r400_lines_kernel<C4L>, grid, dim3(C4L * 20), c4_lds(C4L), st, S, (int64_t)C4N * C4H,
                     (int64_t)C4H, plane, T);
  static_assert((C4N * C4N / 2) % C4NL == 0, "row pairs per workgroup must divide the pairs");
  hipLaunchKernelGGL(c2r400_rows_kernel, dim3(C4N * C4N / 2 / C4NL, 1, batch), dim3(C4T), c4_lds(C4NL), st,
                     (const float2*)S, (float*)out, plane, (int64_t)C4N * C4N * C4N, T);
  return hipGetLastError();
}

}  // namespace cnpc
