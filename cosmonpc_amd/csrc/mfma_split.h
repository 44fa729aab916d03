// MFMA helpers shared by the split-precision contractions: operand vector
// types, the hi + lo split of FP32 values and the fixed-order reduction of
// the per-wave 32 x 32 accumulators of a workgroup.
#pragma once
#include "cnpc_common.h"

namespace cnpc {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// Sums the accumulators of waves w_lo .. w_hi-1 (all by default) into one
// FP64 32 x 32 tile, wave by wave in a fixed order, and stores it to out.
// Every thread of the workgroup must call it (it synchronises).
__device__ __forceinline__ void block_reduce_store(const float (&acc)[2][2][4], double* out, int w_lo = 0,
                                                   int w_hi = 1 << 20) {
  // acc[bi][bj][j] holds C[16*bi + 4*(lane/16) + j][16*bj + lane%16]
  __shared__ double red[32 * 32];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < 1024; i += blockDim.x) red[i] = 0.0;
  __syncthreads();
  const int nw = (int)(blockDim.x >> 6) < w_hi ? (int)(blockDim.x >> 6) : w_hi;
  for (int ww = w_lo; ww < nw; ++ww) {
    if (w == ww) {
      for (int bi = 0; bi < 2; ++bi)
        for (int bj = 0; bj < 2; ++bj)
          for (int j = 0; j < 4; ++j) {
            const int r = 16 * bi + 4 * (lane >> 4) + j, c = 16 * bj + (lane & 15);
            red[r * 32 + c] += (double)acc[bi][bj][j];
          }
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < 1024; i += blockDim.x) out[i] = red[i];
  __syncthreads();  // red may be reused by a second call
}

// The same for FP64 per-wave accumulators.
__device__ __forceinline__ void block_reduce_store64(const double (&acc)[2][2][4], double* out, int w_lo = 0,
                                                     int w_hi = 1 << 20) {
  __shared__ double red64[32 * 32];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < 1024; i += blockDim.x) red64[i] = 0.0;
  __syncthreads();
  const int nw = (int)(blockDim.x >> 6) < w_hi ? (int)(blockDim.x >> 6) : w_hi;
  for (int ww = w_lo; ww < nw; ++ww) {
    if (w == ww) {
      for (int bi = 0; bi < 2; ++bi)
        for (int bj = 0; bj < 2; ++bj)
          for (int j = 0; j < 4; ++j) {
            const int r = 16 * bi + 4 * (lane >> 4) + j, c = 16 * bj + (lane & 15);
            red64[r * 32 + c] += acc[bi][bj][j];
          }
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < 1024; i += blockDim.x) out[i] = red64[i];
  __syncthreads();
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2v __attribute__((ext_vector_type(2)));
typedef unsigned u32x4v __attribute__((ext_vector_type(4)));

// two FP32 values rounded to nearest-even bf16 in one register (element 0 in
// the low half): one v_cvt_pk_bf16_f32
__device__ __forceinline__ unsigned pack_bf16x2(float a, float b) {
  const f32x2v v = {a, b};
  return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
}

// hi = bf16(v), lo = bf16(v - hi) for 8 values, two per conversion: the FP32
// value of each hi comes from its packed bits by a shift or a mask, so a
// value costs half a conversion for hi, one bit operation, one subtraction
// and half a conversion for lo.
__device__ __forceinline__ void split8_bf16(const float (&v)[8], bf16x8& hi, bf16x8& lo) {
  u32x4v h, l;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const unsigned hp = pack_bf16x2(v[2 * p], v[2 * p + 1]);
    const float h0 = __uint_as_float(hp << 16), h1 = __uint_as_float(hp & 0xffff0000u);
    h[p] = hp;
    l[p] = pack_bf16x2(v[2 * p] - h0, v[2 * p + 1] - h1);
  }
  hi = __builtin_bit_cast(bf16x8, h);
  lo = __builtin_bit_cast(bf16x8, l);
}

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2v __attribute__((ext_vector_type(2)));

// split8_bf16 for 4 values (the operands of a 16x16x16 product)
__device__ __forceinline__ void split4_bf16(const float (&v)[4], bf16x4& hi, bf16x4& lo) {
  u32x2v h, l;
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const unsigned hp = pack_bf16x2(v[2 * p], v[2 * p + 1]);
    const float h0 = __uint_as_float(hp << 16), h1 = __uint_as_float(hp & 0xffff0000u);
    h[p] = hp;
    l[p] = pack_bf16x2(v[2 * p] - h0, v[2 * p + 1] - h1);
  }
  hi = __builtin_bit_cast(bf16x4, h);
  lo = __builtin_bit_cast(bf16x4, l);
}

template <bool F16>
__device__ __forceinline__ void split8(const float (&v)[8], bf16x8& hi, bf16x8& lo, f16x8& hh, f16x8& ll) {
  if (F16) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      _Float16 h = (_Float16)v[j];
      hh[j] = h;
      ll[j] = (_Float16)(v[j] - (float)h);
    }
  } else {
    split8_bf16(v, hi, lo);
  }
}

}  // namespace cnpc
