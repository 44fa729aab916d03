// Fused z-synthesis + split-bf16 contraction  S_ab = sum_x A_a(x) G(x) B_b(x).
//
// The shells A_a, B_b are the real inverse FFTs of shell-filtered spectra
// (clustering_estimator.py:934-970) and S_ab is the reference's
// np.sum(G_00 * F_i * F_j) (:972).  The shell synthesis leaves them half
// done, x and y transformed, in H[x][y/2][slot][kz][y%2] for the R + 1
// non-zero kz.  Writing the real shells (N^3 FP32 each) and reading them back
// in the GEMM is most of the HBM traffic of the unfused path; here the z pass
// runs in LDS right before the MFMAs and the real shells never reach HBM.
//
// A workgroup loops over tiles (x, y pair), one contiguous run of H over all
// slots, and stages the tile's two G rows in LDS.  Then, per FFT round:
//   1. two real shells share one complex line (two real transforms per
//      complex FFT): line l = P + iQ with, along z, the Hermitian half
//      spectra p, q of the two shells expanded as w[kz] = p + iq,
//      w[n-kz] = conj(p) + i conj(q).  For B == A the pairs are (A_2l,
//      A_2l+1) and one round holds both rows; for B != A they are (A_l, B_l)
//      and each row is a round;
//   2. the lines get the in-place Stockham inverse FFT in LDS (real part =
//      first shell, imaginary part = second shell, z = 0 .. n-1);
//   3. each row's n values of K are cut in chunks of 32 (the last one padded
//      with zeros); wave w takes chunks w, w + nw, ...: its lanes read 8
//      consecutive values of their A and B rows, form B * G, split hi + lo in
//      bf16 and issue the three 16x16x32 MFMAs of each of the 2 x 2 output
//      blocks (FP32 accumulate).
// With B != A the same lines also give S'_bc = sum_x B_b G B_c (BB): the
// first half of the waves accumulates A^T G B, the second half B^T G B over
// the same chunks, so a configuration whose legs are both B (e.g. [0,0,0]
// beside [2,0,2]) costs no FFT of its own.
// Slot s holds kz rows below q[s] only (its own shell radius); the rest of
// its line is zero.  The MFMA phase reads its operands without masks: rows
// without a shell read zero lines (lines L .. 31 for B != A, line L for
// B == A), and the G rows are zero beyond z = n, so the padded part of the
// last chunk (whatever the lines hold there, all finite) adds nothing.
// The per-wave accumulators are summed in a fixed order into one FP64
// 32 x 32 partial per workgroup, and the partials in a fixed order by the
// finalize kernel (deterministic).
#include "lds_fft.h"
#include "mfma_split.h"

namespace cnpc {

// 12 waves, 3 per SIMD: the FFT staging and the MFMA operands fit the 168
// registers of that occupancy without spills; one workgroup per CU (about
// 100 KB of LDS at n = 400)
#ifndef ZG_THREADS_DEF
#define ZG_THREADS_DEF 768
#endif
#ifndef ZG_MAXB_DEF
#define ZG_MAXB_DEF 4
#endif
constexpr int ZG_THREADS = ZG_THREADS_DEF;
constexpr int ZG_MAXB = ZG_MAXB_DEF;  // butterflies per thread and FFT pass

// G row stride in LDS: the full width of a row's 32-wide chunks
__host__ __device__ inline int zg_gstride(int n) { return ((n + 31) >> 5) << 5; }
// floats of zeros after the lines: the unmasked reads of the last line's
// last chunk reach at most 32 - 8 floats past its second plane
constexpr int ZG_TAIL_PAD = 32;

struct ZGemmArgs {
  int n, Q;                 // grid, non-zero kz per line (R + 1)
  int Qi;                   // fixed plan: fetched modes per line (Q, or n/4 + 1 when 4Q <= n)
  int blocked;              // half shells in the blocked layout (half_blocked)
  int nkb;                  // blocked layout: kz blocks of 8 per slot
  int na, nb;               // rows of A and B (<= 32)
  int sa_slots, sb_slots;   // slots per tile in HA / HB
  int same;                 // B is A
  int nlines;               // complex lines per row
  int ntiles;               // n * n / 2
  int qa[32], qb[32];       // kz rows held per slot (the rest is zero)
  SynthArgs fft;            // n and its radices
  const float4* HA;
  const float4* HB;
  const float* G;           // [x][y][z]
  const float2* tw;         // e^{2 pi i j / n}
  double* partial;          // [gridDim.x][32 * 32]
  double* partial_bb;       // [gridDim.x][32 * 32] of B^T G B (BB launches)
};

// 8 consecutive z values of one part of one planar line (two 16-byte LDS
// loads)
__device__ __forceinline__ void zg_row8u(const float2* buf, int ls, int line, int part, int z0, float (&v)[8]) {
  const float4* p = (const float4*)((const float*)buf + 2 * line * ls + part * ls + z0);
  const float4 a = p[0], b = p[1];
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

// One round = the lines of one FFT and the MFMA chunks they feed.  B == A
// (BR): one round per tile, both rows (line r * Lh + l holds row r of
// shells 2l, 2l+1), 2 * nchr chunks.  B != A: one round per row (line l
// holds A_l and B_l), nchr chunks; BB also accumulates B^T G B.
template <int THREADS, bool BR, bool BB>
__global__ void __launch_bounds__(THREADS) zgemm_bf16x3_kernel(ZGemmArgs g) {
  // lines of n + 4 complex: 16-byte aligned planes for the planar last pass
  const int n = g.n, Q = g.Q, R = Q - 1, ls = n + 4, half = n >> 1, L = g.nlines;
  const int gls = zg_gstride(n);        // G row stride: the chunks' full width
  extern __shared__ unsigned char smem[];
  float2* W = (float2*)smem;
  float* Gs = (float*)(W + n);          // G rows y, y + 1 of the tile (zero beyond n)
  float2* buf = (float2*)(Gs + 2 * gls);
  build_pass_twiddles(W, g.tw, n, g.fft);
  {
    // zero once: the G row tails, the lines beyond L (zero rows), the float
    // range [2n, 2 ls) of every line (never written by the FFT) and the
    // tail pad after the lines
    const int lalloc = BR ? L + 1 : 32;
    for (int t = threadIdx.x; t < 2 * (gls - n); t += blockDim.x)
      Gs[(t / (gls - n)) * gls + n + t % (gls - n)] = 0.f;
    float* fb = (float*)buf;
    for (int t = threadIdx.x; t < (lalloc - L) * 2 * ls + ZG_TAIL_PAD; t += blockDim.x) fb[2 * L * ls + t] = 0.f;
    for (int t = threadIdx.x; t < L * 2 * (ls - n); t += blockDim.x) {
      const int l = t / (2 * (ls - n)), j = t - l * 2 * (ls - n);
      fb[2 * l * ls + 2 * n + j] = 0.f;
    }
  }
  __shared__ int qa[32], qb[32];  // kz rows per slot (read by lane-varying index)
  if (threadIdx.x < 32) {
    qa[threadIdx.x] = g.qa[threadIdx.x];
    qb[threadIdx.x] = g.qb[threadIdx.x];
  }
  __syncthreads();  // every wave reads qa / qb in its first fill loop
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  // BB: waves [0, nw/2) take A^T G B, waves [nw/2, nw) B^T G B (wave-uniform)
  const bool wbb = BB && w >= (nw >> 1);
  const int wq = BB ? (wbb ? w - (nw >> 1) : w) : w, nwq = BB ? (nw >> 1) : nw;
  const int rr = lane & 15, kg = lane >> 4;
  const int nchr = (n + 31) >> 5;      // chunks per row
  const int Lh = (g.na + 1) >> 1;      // lines per row when B == A
  const int nz0 = n - 2 * R - 1;
  const float inv_z = 1.0f / (float)(nz0 > 0 ? nz0 : 1), inv_q = 1.0f / (float)Q;
  float acc[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[i][j][q] = 0.f;
  f16x8 fdh, fdl;  // unused fp16 outputs of split8
  const int rounds = BR ? 1 : 2;  // BR: B == A, both rows in one round
  for (int64_t tile = blockIdx.x; tile < g.ntiles; tile += gridDim.x) {
    const int x = (int)(tile / half), yp = (int)(tile - (int64_t)x * half);
    const float4* TA = g.HA + tile * g.sa_slots * Q;
    const float4* TB = g.HB + tile * g.sb_slots * Q;
    for (int r0 = 0; r0 < rounds; ++r0) {
      // modes Q .. n-Q are zero (the previous FFT overwrote them)
      for (int t = threadIdx.x; t < L * nz0; t += blockDim.x) {
        const int l = fdiv(t, nz0, inv_z), j = t - l * nz0;
        buf[l * ls + Q + j] = make_float2(0.f, 0.f);
      }
      if (r0 == 0) {
        const float* Grow = g.G + ((int64_t)x * n + 2 * yp) * n;
        for (int t = threadIdx.x; t < 2 * n; t += blockDim.x) Gs[(t >= n ? gls - n : 0) + t] = Grow[t];
      }
      if (BR) {
        // (l, kz): shells 2l, 2l+1; both rows -> lines l and Lh + l (BR),
        // else row r0 -> line l
        for (int t = threadIdx.x; t < Lh * Q; t += blockDim.x) {
          const int l = fdiv(t, Q, inv_q), kz = t - l * Q;
          const int s0 = 2 * l, s1 = s0 + 1;
          float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
          if (s0 < g.na && kz < qa[s0]) a = TA[s0 * Q + kz];
          if (s1 < g.na && kz < qa[s1]) b = TA[s1 * Q + kz];
#pragma unroll
          for (int rb = 0; rb < (BR ? 2 : 1); ++rb) {
            const int r = BR ? rb : r0;
            const float2 p = r ? make_float2(a.z, a.w) : make_float2(a.x, a.y);
            const float2 q = r ? make_float2(b.z, b.w) : make_float2(b.x, b.y);
            float2* Ln = buf + ((BR ? r * Lh : 0) + l) * ls;
            if (kz == 0) {  // c2r reads only the real part of the zero mode
              Ln[0] = make_float2(p.x, q.x);
            } else {
              Ln[kz] = make_float2(p.x - q.y, p.y + q.x);
              Ln[n - kz] = make_float2(p.x + q.y, q.x - p.y);
            }
          }
        }
      } else {
        for (int t = threadIdx.x; t < L * Q; t += blockDim.x) {
          const int l = fdiv(t, Q, inv_q), kz = t - l * Q;
          float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
          if (l < g.na && kz < qa[l]) a = TA[l * Q + kz];
          if (l < g.nb && kz < qb[l]) b = TB[l * Q + kz];
          const float2 p = r0 ? make_float2(a.z, a.w) : make_float2(a.x, a.y);
          const float2 q = r0 ? make_float2(b.z, b.w) : make_float2(b.x, b.y);
          float2* Ln = buf + l * ls;
          if (kz == 0) {
            Ln[0] = make_float2(p.x, q.x);
          } else {
            Ln[kz] = make_float2(p.x - q.y, p.y + q.x);
            Ln[n - kz] = make_float2(p.x + q.y, q.x - p.y);
          }
        }
      }
      __syncthreads();
      stockham_lds_planar<ZG_MAXB>(buf, W, n, ls, L, g.fft);
      const int nchunk = BR ? 2 * nchr : nchr;
      for (int c = wq; c < nchunk; c += nwq) {
        const int r = BR ? (c >= nchr ? 1 : 0) : r0;
        const int z0 = 32 * (c - ((BR && r) ? nchr : 0)) + 8 * kg;
        bf16x8 ah[2], al[2], bh[2], bl[2];
        if (BR) {  // B == A: row i is part i & 1 of line r Lh + i / 2; rows past the lines read the zero line L
          float vg[8];
          {
            const float4* gp = (const float4*)(Gs + r * gls + z0);
            const float4 g0 = gp[0], g1 = gp[1];
            vg[0] = g0.x; vg[1] = g0.y; vg[2] = g0.z; vg[3] = g0.w;
            vg[4] = g1.x; vg[5] = g1.y; vg[6] = g1.z; vg[7] = g1.w;
          }
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int i = 16 * h + rr;
            float a[8], b[8];
            const int line = (i >> 1) < Lh ? r * Lh + (i >> 1) : L;
            zg_row8u(buf, ls, line, i & 1, z0, a);
#pragma unroll
            for (int j = 0; j < 8; ++j) b[j] = a[j] * vg[j];
            split8<false>(a, ah[h], al[h], fdh, fdl);
            split8<false>(b, bh[h], bl[h], fdh, fdl);
          }
        } else {  // B != A: line i is row i, no masks (zero rows, zero G tail)
          float vg[8];
          {
            const float4* gp = (const float4*)(Gs + r * gls + z0);
            const float4 g0 = gp[0], g1 = gp[1];
            vg[0] = g0.x; vg[1] = g0.y; vg[2] = g0.z; vg[3] = g0.w;
            vg[4] = g1.x; vg[5] = g1.y; vg[6] = g1.z; vg[7] = g1.w;
          }
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int i = 16 * h + rr;
            float a[8], b[8];
            zg_row8u(buf, ls, i, wbb ? 1 : 0, z0, a);
            if (wbb) {
#pragma unroll
              for (int j = 0; j < 8; ++j) b[j] = a[j] * vg[j];
            } else {
              zg_row8u(buf, ls, i, 1, z0, b);
#pragma unroll
              for (int j = 0; j < 8; ++j) b[j] *= vg[j];
            }
            split8<false>(a, ah[h], al[h], fdh, fdl);
            split8<false>(b, bh[h], bl[h], fdh, fdl);
          }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            f32x4 cacc = {acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]};
            cacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], cacc, 0, 0, 0);
            cacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], cacc, 0, 0, 0);
            cacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], cacc, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[i][j][q] = cacc[q];
          }
      }
      __syncthreads();  // the next round overwrites buf (and Gs)
    }
  }
  if (BB) {
    block_reduce_store(acc, g.partial + (int64_t)blockIdx.x * 1024, 0, nw >> 1);
    block_reduce_store(acc, g.partial_bb + (int64_t)blockIdx.x * 1024, nw >> 1, nw);
  } else {
    block_reduce_store(acc, g.partial + (int64_t)blockIdx.x * 1024);
  }
}

// ---------------------------------------------------------------- fixed plan
// Grids n = N1 * N2 with both factors composite radices (e.g. 400 = 20 x
// 20): the line FFT runs as two in-place passes, and the contraction sums
// over z in the order the FFT leaves it (G is staged in the same order), so
// no pass sorts its output and no pass needs a barrier between its reads and
// its writes.
//
// A line holds N1 unpadded rows of RW = N2 complex (every position is an
// output, so the MFMA phase has no pad columns to skip): input mode k = k2 +
// N2 k1 sits at position k2 + RW k1.
//   pass 1, butterfly (line, k2): the N1-point inverse DFT over k1 (stride
//     RW), times e^{2 pi i k2 z1 / n}, written back in place (row z1);
//   pass 2, butterfly (line, z1): the N2-point inverse DFT over the
//     contiguous row z1, written back in place: z = z1 + N1 z2 at position
//     z2 + RW z1 (real part = first shell, imaginary part = second shell).
// G is zero past the KP = N1 RW positions, so the MFMA phase reads its
// operands without masks (the last chunk's overhang holds finite values:
// the line pitch's two spare positions, the next line, or the zeroed tail).
#ifndef ZGF_THREADS_DEF
#define ZGF_THREADS_DEF 768
#endif
constexpr int ZGF_THREADS = ZGF_THREADS_DEF;  // threads of the fixed-plan kernels

template <int N1, int N2>
struct ZgFixed : FixedPlan<N1, N2, N2> {
  typedef FixedPlan<N1, N2, N2> B;
  static constexpr int NCH = (B::KP + 31) / 32;    // 32-wide chunks per row
  static constexpr int NFULL = B::KP / 32;         // chunks of 32 valid positions
  static constexpr int KT = B::KP - 32 * NFULL;    // positions of the last, partial chunk
  static constexpr int GLS = NCH * 32;             // G row stride in LDS
  // line pitch (complex): at least two spare positions, and odd (5 mod 32):
  // the lanes of an FFT pass walk the lines of one position, and an odd
  // pitch in 8-byte units puts them on all the LDS banks (an even pitch
  // leaves half the banks unused and doubles the cycles of every access of
  // both passes).  Lines are then 8-byte aligned only, so the MFMA phase
  // reads them 8 bytes at a time
  static constexpr int LSC = (B::KP + 2 - 5 + 31) / 32 * 32 + 5;
  static constexpr int DUMP = B::KP + 1;           // spare position of a line (never a mode)
  static constexpr int TAIL = 2 * (GLS - B::KP);   // floats read past the last line
  __device__ static int pos(int k) { return B::ipos(k); }
};

template <int N1, int N2>
size_t zgemm_fixed_lds_bytes(int nlines_alloc) {
  typedef ZgFixed<N1, N2> F;
  return sizeof(float2) * (size_t)F::n + sizeof(float) * 2 * (size_t)F::GLS +
         sizeof(float2) * (size_t)nlines_alloc * F::LSC + sizeof(float) * (size_t)F::TAIL;
}

// 8 consecutive positions of an interleaved line (8-byte aligned): real and
// imaginary parts
__device__ __forceinline__ void zgf_row8(const float2* p, float (&re)[8], float (&im)[8]) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const float2 v = p[j];
    re[j] = v.x;
    im[j] = v.y;
  }
}

// 4 consecutive positions of an interleaved line (8-byte aligned)
__device__ __forceinline__ void zgf_row4(const float2* p, float (&re)[4], float (&im)[4]) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float2 v = p[j];
    re[j] = v.x;
    im[j] = v.y;
  }
}

// Two-output launch, the 16 positions z0 .. z0 + 15 past the last full
// chunk, output block (TI, TJ) only: C += A_i G B_j and C_bb += B_i G B_j
// for i in block TI, j in block TJ by 16x16x16 products (lane group kg
// takes positions z0 + 4 kg .. + 3)
template <int TI, int TJ>
__device__ __forceinline__ void zgf_tail_block(const float2* buf, const float* Gz, int lsc, int rr, int kg,
                                               float (&acc)[2][2][4], float (&acc2)[2][2][4]) {
  const float4 gq = *(const float4*)(Gz + 4 * kg);
  const float vg[4] = {gq.x, gq.y, gq.z, gq.w};
  float re[4], im[4], b[4];
  zgf_row4(buf + (16 * TJ + rr) * lsc + 4 * kg, re, im);
#pragma unroll
  for (int j = 0; j < 4; ++j) b[j] = im[j] * vg[j];
  bf16x4 gh, gl, ah, al, bh, bl;
  split4_bf16(b, gh, gl);
  zgf_row4(buf + (16 * TI + rr) * lsc + 4 * kg, re, im);
  split4_bf16(re, ah, al);
  split4_bf16(im, bh, bl);
  f32x4 c1 = {acc[TI][TJ][0], acc[TI][TJ][1], acc[TI][TJ][2], acc[TI][TJ][3]};
  f32x4 c2 = {acc2[TI][TJ][0], acc2[TI][TJ][1], acc2[TI][TJ][2], acc2[TI][TJ][3]};
  c1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, gh, c1, 0, 0, 0);
  c2 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bh, gh, c2, 0, 0, 0);
  c1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, gl, c1, 0, 0, 0);
  c2 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bh, gl, c2, 0, 0, 0);
  c1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, gh, c1, 0, 0, 0);
  c2 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bl, gh, c2, 0, 0, 0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    acc[TI][TJ][q] = c1[q];
    acc2[TI][TJ][q] = c2[q];
  }
}

// The operands of one round are fetched into registers, MAXF items per
// thread: item t = threadIdx.x + it * THREADS is line l = t / Q, mode kz =
// t % Q.  B == A: the float4 of shells 2l and 2l + 1 (both rows of the
// tile); B != A: the float4 of A_l and of B_l (both rows, the second kept
// for the tile's second round).  What an item reads and
// where it lands in LDS is the same every round, so it is computed once:
//   lds word: float2 index of mode kz in its line (bits 0-13), of mode n - kz
//     (bits 16-29; for kz = 0 the spare position KP + 1 of the line, which
//     the passes never read) and kz == 0 (bit 31); items past the round's
//     lines land on the spare position of line 0 (they carry zeros);
//   fetch word: byte offset of the item in the tile's run of HA (and HB),
//     or in the x plane for the blocked layout (bits 0-29), bit 30 set when
//     the A slot holds no row kz, bit 31 when the B slot holds none: masking
//     with ZG_A_MASK / ZG_B_MASK leaves the offset, or an offset beyond the
//     run, which the buffer load answers with zeros.
template <bool BR, int MAXF>
struct ZgfFetch {
  float4 a[MAXF], b[MAXF];  // rows y0, y0 + 1 of each item
};
constexpr unsigned ZG_A_MASK = 0x7fffffffu, ZG_B_MASK = 0xbfffffffu;
constexpr int ZG_BUF_FLAGS = 0x00020000;  // buffer resource word 3 (gfx9 raw buffer)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t zg_rsrc(const void* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, ZG_BUF_FLAGS);
}

template <int THREADS, bool BR, int MAXF, int N1, int N2>
__device__ __forceinline__ void zgf_items(const ZGemmArgs& g, const int* qa, const int* qb, int Qi, int total,
                                          unsigned (&lw)[MAXF], unsigned (&fw)[MAXF]) {
  typedef ZgFixed<N1, N2> P;
  constexpr int n = N1 * N2, LSC = P::LSC;
  const int Q = g.Q;
#pragma unroll
  for (int it = 0; it < MAXF; ++it) {
    const int t = threadIdx.x + it * THREADS;
    const int l = t / Qi, kz = t - l * Qi;
    const bool ok = t < total;
    const unsigned l1 = ok ? (unsigned)(l * LSC + P::ipos(kz)) : (unsigned)P::DUMP;
    const unsigned l2 = ok ? (unsigned)(l * LSC + (kz ? P::ipos(n - kz) : P::DUMP)) : (unsigned)P::DUMP;
    lw[it] = l1 | (l2 << 16) | (ok && kz == 0 ? 0x80000000u : 0u);
    bool va, vb;
    int slot;
    if constexpr (BR) {  // A: shell 2l, "B": shell 2l + 1, both in HA
      slot = 2 * l;
      va = ok && 2 * l < g.na && kz < qa[2 * l];
      vb = ok && 2 * l + 1 < g.na && kz < qa[2 * l + 1];
    } else {
      slot = l;
      va = ok && l < g.na && kz < qa[l];
      vb = ok && l < g.nb && kz < qb[l];
    }
    // tile-major: (slot, kz) float4 of both rows; blocked: (slot, kz / 8)
    // runs of n / 2 row pairs of 8 float4, pair yp at + 128 yp bytes (added
    // per round)
    const unsigned off = !ok ? 0u
                         : g.blocked ? (unsigned)(((slot * g.nkb + kz / 8) * n) * 64 + (kz % 8) * 16)
                                     : (unsigned)((slot * Q + kz) * 16);
    fw[it] = off | (va ? 0u : 0x40000000u) | (vb ? 0u : 0x80000000u);
  }
}

template <int THREADS, bool BR, int MAXF>
__device__ __forceinline__ void zgf_issue(ZgfFetch<BR, MAXF>& f, const ZGemmArgs& g, const unsigned (&fw)[MAXF],
                                          int64_t tile) {
  const int Q = g.Q;
  if (g.blocked) {
    // blocked layout: the records are the tile's x plane, the tile's row
    // pair yp = y0 / 2 a scalar offset
    const int n = g.n, half = n >> 1;
    const int64_t x = tile / half;
    const int y0 = 2 * (int)(tile - x * half);
    const int sb = g.nkb * n * 64;  // bytes per slot
    const __amdgpu_buffer_rsrc_t ra = zg_rsrc((const char*)g.HA + x * g.sa_slots * (int64_t)sb, g.sa_slots * sb);
    const __amdgpu_buffer_rsrc_t rb = zg_rsrc((const char*)g.HB + x * g.sb_slots * (int64_t)sb, g.sb_slots * sb);
#pragma unroll
    for (int it = 0; it < MAXF; ++it) {
      // B == A: shells 2l and 2l + 1 sit one slot apart
      f.a[it] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(ra, (int)(fw[it] & ZG_A_MASK),
                                                                                 64 * y0, 0));
      f.b[it] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(BR ? ra : rb,
                                                                                 (int)(fw[it] & ZG_B_MASK),
                                                                                 (BR ? sb : 0) + 64 * y0, 0));
    }
    return;
  }
  const __amdgpu_buffer_rsrc_t ra = zg_rsrc(g.HA + tile * g.sa_slots * Q, g.sa_slots * Q * 16);
  const __amdgpu_buffer_rsrc_t rb = zg_rsrc(g.HB + tile * g.sb_slots * Q, g.sb_slots * Q * 16);
#pragma unroll
  for (int it = 0; it < MAXF; ++it) {
    // B == A: shells 2l and 2l + 1 sit Q float4 apart
    f.a[it] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(ra, (int)(fw[it] & ZG_A_MASK), 0, 0));
    f.b[it] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(BR ? ra : rb, (int)(fw[it] & ZG_B_MASK),
                                                                               BR ? Q * 16 : 0, 0));
  }
}

template <int THREADS, bool BR, bool BB, int N1, int N2, int MAXF>
__global__ void __launch_bounds__(THREADS) zgemm_fixed_kernel(ZGemmArgs g) {
  typedef ZgFixed<N1, N2> F;
  constexpr int n = F::n, RW = F::RW, GLS = F::GLS, NCH = F::NCH, LSC = F::LSC;
  const int Q = g.Q, half = n >> 1, L = g.nlines;
  extern __shared__ unsigned char smem[];
  float2* T1 = (float2*)smem;            // e^{2 pi i k2 z1 / n} at z1 N2 + k2
  float* Gs = (float*)(T1 + n);          // G of the tile's rows in FFT order (zero at pads and past KP)
  float2* buf = (float2*)(Gs + 2 * GLS); // the lines, LSC complex each
  const int lalloc = BR ? L + 1 : 32;
  fixed_twiddles<N1, N2>(T1, g.tw);
  for (int t = threadIdx.x; t < 2 * GLS; t += THREADS) Gs[t] = 0.f;
  {
    float* fb = (float*)buf;
    for (int t = threadIdx.x; t < lalloc * 2 * LSC + F::TAIL; t += THREADS) fb[t] = 0.f;
  }
  __shared__ int qa[32], qb[32];
  if (threadIdx.x < 32) {
    qa[threadIdx.x] = g.qa[threadIdx.x];
    qb[threadIdx.x] = g.qb[threadIdx.x];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int nw = THREADS / 64;
  const int rr = lane & 15, kg = lane >> 4;
  const int Lh = (g.na + 1) >> 1;
  float acc[2][2][4], acc2[2][2][4];  // acc2: B^T G B (BB only)
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[i][j][q] = acc2[i][j][q] = 0.f;
  constexpr int rounds = BR ? 1 : 2;
  // items per line: the modes below n/4 when the band fits (zeros past each
  // shell's radius, which the first FFT pass then reads unconditionally),
  // else the Q rows
  const int Qi = g.Qi;
  const int total = BR ? Lh * Qi : L * Qi;  // fetched items per round
  // Rounds (tile, r0) in order, tile = blockIdx.x, blockIdx.x + gridDim.x,
  // ...: the operands of the next round (lines and G rows) are loaded into
  // registers while this round's FFT and MFMAs run (the barriers order LDS
  // only, so the loads stay in flight across them).
  constexpr int GPT = ((BR ? 2 : 1) * n + THREADS - 1) / THREADS;  // G values per thread and round
  unsigned lw[MAXF], fw[MAXF];
  zgf_items<THREADS, BR, MAXF, N1, N2>(g, qa, qb, Qi, total, lw, fw);
  // each thread's butterfly of either FFT pass, the same every round (the
  // host admits at most THREADS / N2 lines)
  const FixedSlots fs = fixed_slots<N1, N2, RW>(L, LSC);
  ZgfFetch<BR, MAXF> pf;
  float gv[GPT];
  // B != A: one fetch per tile holds both rows; the round of row y0 + 1
  // fills from the registers the round of row y0 left
  auto issue_round = [&](int64_t tl, int rr) {
    if (BR || rr == 0) zgf_issue<THREADS, BR, MAXF>(pf, g, fw, tl);
    const int xx = (int)(tl / half), yy = (int)(tl - (int64_t)xx * half);
    const float* Grow = g.G + ((int64_t)xx * n + 2 * yy) * n + (BR ? 0 : rr * n);
#pragma unroll
    for (int i = 0; i < GPT; ++i) {
      const int t = threadIdx.x + i * THREADS;
      gv[i] = t < (BR ? 2 : 1) * n ? Grow[t] : 0.f;
    }
  };
  int64_t tile = blockIdx.x;
  int r0 = 0;
  if (tile < g.ntiles) issue_round(tile, 0);
  while (tile < g.ntiles) {
    {
      // the round's G row(s) in FFT order: z = z1 + N1 z2 at z2 + RW z1
#pragma unroll
      for (int i = 0; i < GPT; ++i) {
        const int t = threadIdx.x + i * THREADS;
        if (t < (BR ? 2 : 1) * n) {
          const int r = BR ? (t >= n ? 1 : 0) : r0, z = t - (BR && t >= n ? n : 0);
          const int z2 = z / N1, z1 = z - z2 * N1;
          Gs[r * GLS + z2 + RW * z1] = gv[i];
        }
      }
      // the round's lines from the fetched registers: mode kz = p + i q
      // and mode n - kz = conj(p) + i conj(q) of the line (the zero mode
      // keeps the real parts only)
#pragma unroll
      for (int it = 0; it < MAXF; ++it) {
        const unsigned wd = lw[it];
        const bool z0 = (int)wd < 0;
        float2* L1 = buf + (wd & 0x3fffu);
        float2* L2 = buf + ((wd >> 16) & 0x3fffu);
        if constexpr (BR) {
          float4 a = pf.a[it], b = pf.b[it];
          if (z0) a.y = a.w = b.y = b.w = 0.f;
#pragma unroll
          for (int r = 0; r < 2; ++r) {
            const float2 p = r ? make_float2(a.z, a.w) : make_float2(a.x, a.y);
            const float2 q = r ? make_float2(b.z, b.w) : make_float2(b.x, b.y);
            L1[r * Lh * LSC] = make_float2(p.x - q.y, p.y + q.x);
            L2[r * Lh * LSC] = make_float2(p.x + q.y, q.x - p.y);
          }
        } else {
          const float4 a = pf.a[it], b = pf.b[it];
          float2 p = r0 ? make_float2(a.z, a.w) : make_float2(a.x, a.y);
          float2 q = r0 ? make_float2(b.z, b.w) : make_float2(b.x, b.y);
          if (z0) p.y = q.y = 0.f;
          L1[0] = make_float2(p.x - q.y, p.y + q.x);
          L2[0] = make_float2(p.x + q.y, q.x - p.y);
        }
      }
    }
    // the next round's loads
    int64_t ntile = tile;
    int nr0 = r0 + 1;
    if (nr0 == rounds) {
      nr0 = 0;
      ntile += gridDim.x;
    }
    if (ntile < g.ntiles) issue_round(ntile, nr0);
    lds_barrier();
    // the line FFTs, in place (z = z1 + N1 z2 at position z2 + RW z1)
    // (modes Q .. n - Q are zero: pass 1 does not read their positions)
    if (Qi > Q)
      fixed_fft_slots<N1, N2, true, RW>(buf, T1, fs, Q);
    else
      fixed_fft_slots<N1, N2, false, RW>(buf, T1, fs, Q);
    lds_barrier();
    constexpr int nchunk = BR ? 2 * NCH : NCH;
    if constexpr (BB) {
      // every wave takes chunks c = w, w + nw, ... of 32 positions and feeds
      // both outputs: line i holds A_i (real part) and B_i (imaginary part);
      // B_j G is split once for both, A_i and B_i once each.  The last 16
      // positions are four 16x16x16 output blocks, one on each of the last
      // four waves (one per SIMD), so no wave takes a second chunk
      static_assert(F::KT == 16 && F::NFULL <= nw && nw >= 4, "the two-output plan: 32 k + 16 positions");
      for (int c = w; c < F::NFULL; c += nw) {
        const int z0 = 32 * c + 8 * kg;
        float vg[8];
        {
          const float4* gp = (const float4*)(Gs + r0 * GLS + z0);
          const float4 g0 = gp[0], g1 = gp[1];
          vg[0] = g0.x; vg[1] = g0.y; vg[2] = g0.z; vg[3] = g0.w;
          vg[4] = g1.x; vg[5] = g1.y; vg[6] = g1.z; vg[7] = g1.w;
        }
        bf16x8 gh[2], gl[2], ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          float re[8], im[8], b[8];
          zgf_row8(buf + (16 * h + rr) * LSC + z0, re, im);
#pragma unroll
          for (int j = 0; j < 8; ++j) b[j] = im[j] * vg[j];
          split8_bf16(b, gh[h], gl[h]);
          split8_bf16(re, ah[h], al[h]);
          split8_bf16(im, bh[h], bl[h]);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            f32x4 c1 = {acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]};
            f32x4 c2 = {acc2[i][j][0], acc2[i][j][1], acc2[i][j][2], acc2[i][j][3]};
            c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], gh[j], c1, 0, 0, 0);
            c2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[i], gh[j], c2, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], gl[j], c1, 0, 0, 0);
            c2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[i], gl[j], c2, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], gh[j], c1, 0, 0, 0);
            c2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[i], gh[j], c2, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              acc[i][j][q] = c1[q];
              acc2[i][j][q] = c2[q];
            }
          }
      }
      {
        const float* Gz = Gs + r0 * GLS + 32 * F::NFULL;
        const float2* bz = buf + 32 * F::NFULL;
        switch (w - (nw - 4)) {
          case 0: zgf_tail_block<0, 0>(bz, Gz, LSC, rr, kg, acc, acc2); break;
          case 1: zgf_tail_block<0, 1>(bz, Gz, LSC, rr, kg, acc, acc2); break;
          case 2: zgf_tail_block<1, 0>(bz, Gz, LSC, rr, kg, acc, acc2); break;
          case 3: zgf_tail_block<1, 1>(bz, Gz, LSC, rr, kg, acc, acc2); break;
          default: break;
        }
      }
    } else {
    for (int c = w; c < nchunk; c += nw) {
      const int r = BR ? (c >= NCH ? 1 : 0) : r0;
      const int z0 = 32 * (c - ((BR && r) ? NCH : 0)) + 8 * kg;
      bf16x8 ah[2], al[2], bh[2], bl[2];
      float vg[8];
      {
        const float4* gp = (const float4*)(Gs + r * GLS + z0);
        const float4 g0 = gp[0], g1 = gp[1];
        vg[0] = g0.x; vg[1] = g0.y; vg[2] = g0.z; vg[3] = g0.w;
        vg[4] = g1.x; vg[5] = g1.y; vg[6] = g1.z; vg[7] = g1.w;
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = 16 * h + rr;
        float re[8], im[8], a[8], b[8];
        if constexpr (BR) {  // row i: part i & 1 of line r Lh + i / 2 (rows past the lines: the zero line L)
          const int line = (i >> 1) < Lh ? r * Lh + (i >> 1) : L;
          zgf_row8(buf + line * LSC + z0, re, im);
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            a[j] = (i & 1) ? im[j] : re[j];
            b[j] = a[j] * vg[j];
          }
        } else {  // line i: A_i real, B_i imaginary
          zgf_row8(buf + i * LSC + z0, re, im);
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            a[j] = re[j];
            b[j] = im[j] * vg[j];
          }
        }
        split8_bf16(a, ah[h], al[h]);
        split8_bf16(b, bh[h], bl[h]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          f32x4 cacc = {acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]};
          cacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], cacc, 0, 0, 0);
          cacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], cacc, 0, 0, 0);
          cacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], cacc, 0, 0, 0);
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[i][j][q] = cacc[q];
        }
    }
    }
    lds_barrier();  // the next round overwrites the lines and G
    tile = ntile;
    r0 = nr0;
  }
  __syncthreads();
  if (BB) {
    block_reduce_store(acc, g.partial + (int64_t)blockIdx.x * 1024);
    block_reduce_store(acc2, g.partial_bb + (int64_t)blockIdx.x * 1024);
  } else {
    block_reduce_store(acc, g.partial + (int64_t)blockIdx.x * 1024);
  }
}

// Launch shapes.  B == A (kernel BR = true): 2 * ceil(na/2) lines per round
// (both rows) and one zero line.  B != A: max(na, nb) lines per round, 32
// lines of LDS (the rows beyond stay zero).  768 threads, one workgroup per
// CU.
static int zg_lines(int na, int nb, int same) { return same ? 2 * ((na + 1) / 2) : (na > nb ? na : nb); }
static int zg_lines_alloc(int L, int same) { return same ? L + 1 : 32; }
static int zg_threads(int) { return ZG_THREADS; }

// twiddles, the tile's two G rows (stride zg_gstride) and the lines (plus
// the tail pad that the unmasked reads of the last chunk may touch)
size_t zgemm_lds_bytes(int n, int nlines) {
  return sizeof(float2) * (size_t)n + sizeof(float) * 2 * (size_t)zg_gstride(n) +
         sizeof(float2) * (size_t)nlines * (n + 4) + sizeof(float) * ZG_TAIL_PAD;
}

// host check: every pass of the line FFT fits the staged registers and the
// line buffers (plus the 8 KB reduction tile) fit the LDS
int zgemm_supported(const SynthArgs& a, int na, int nb, int same, size_t lds_limit) {
  const int n = a.n;
  if (n % 8 != 0 || na < 1 || na > 32 || nb < 1 || nb > 32) return 0;
  const int L = zg_lines(na, nb, same);
  for (int s = 0; s < a.nrad; ++s)
    if ((int64_t)L * (n / a.rad[s]) > (int64_t)(wide_radix(a.rad[s]) ? 1 : ZG_MAXB) * zg_threads(same)) return 0;
  return zgemm_lds_bytes(n, zg_lines_alloc(L, same)) + 32 * 32 * sizeof(double) <= lds_limit ? 1 : 0;
}

// default workgroups per CU (one is resident at a time): four keep each
// wave's FP32 sums a quarter as long as one, at the same launch time
int zgemm_blocks_per_cu(int) { return 4; }

template <int T, bool BR, bool BB = false>
static hipError_t zg_launch(const ZGemmArgs& g, int nblk, size_t lds, hipStream_t st) {
  // dynamic LDS beyond the default 64 KB must be allowed per kernel and
  // device; the allowance only grows, so repeated launches (and launches
  // recorded into a graph after an eager one) make no attribute call
  static size_t allowed[CNPC_MAX_DEVICES] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= CNPC_MAX_DEVICES) return hipErrorInvalidDevice;
  if (lds > allowed[dev]) {
    e = hipFuncSetAttribute((const void*)zgemm_bf16x3_kernel<T, BR, BB>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    allowed[dev] = lds;
  }
  hipLaunchKernelGGL((zgemm_bf16x3_kernel<T, BR, BB>), dim3(nblk), dim3(T), lds, st, g);
  return hipGetLastError();
}

template <int T, bool BR, bool BB, int N1, int N2, int MAXF>
static hipError_t zgf_launch_f(const ZGemmArgs& g, int nblk, size_t lds, hipStream_t st) {
  static size_t allowed[CNPC_MAX_DEVICES] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= CNPC_MAX_DEVICES) return hipErrorInvalidDevice;
  if (lds > allowed[dev]) {
    e = hipFuncSetAttribute((const void*)zgemm_fixed_kernel<T, BR, BB, N1, N2, MAXF>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    allowed[dev] = lds;
  }
  hipLaunchKernelGGL((zgemm_fixed_kernel<T, BR, BB, N1, N2, MAXF>), dim3(nblk), dim3(T), lds, st, g);
  return hipGetLastError();
}

// items a thread fetches per round: B == A, ceil(na / 2) Q float4 pairs;
// B != A, max(na, nb) Q float2 pairs (at most 4 per thread)
static int zgf_items_per_line(int n, int Q) { return 4 * Q <= n ? n / 4 + 1 : Q; }
static int zgf_maxf(int Qi, int na, int nb, int same) {
  const int total = same ? ((na + 1) / 2) * Qi : (na > nb ? na : nb) * Qi;
  return total <= 2 * ZGF_THREADS ? 2 : total <= 4 * ZGF_THREADS ? 4 : total <= 8 * ZGF_THREADS ? 8 : 0;
}

template <int T, bool BR, bool BB, int N1, int N2>
static hipError_t zgf_launch(const ZGemmArgs& g, int nblk, size_t lds, hipStream_t st) {
  switch (zgf_maxf(g.Qi, g.na, g.nb, BR ? 1 : 0)) {
    case 2: return zgf_launch_f<T, BR, BB, N1, N2, 2>(g, nblk, lds, st);
    case 4: return zgf_launch_f<T, BR, BB, N1, N2, 4>(g, nblk, lds, st);
    case 8: return zgf_launch_f<T, BR, BB, N1, N2, 8>(g, nblk, lds, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_zgemm(const SynthArgs& a, int Q, const void* HA, int na, int sa_slots, const int* qa,
                        const void* HB, int nb, int sb_slots, const int* qb, int same, const float* G, const void* tw,
                        int nblk, double* partial, double* C, double* Cbb, hipStream_t st) {
  if (Cbb && same) return hipErrorInvalidValue;
  ZGemmArgs g;
  g.n = a.n; g.Q = Q; g.na = na; g.nb = same ? na : nb;
  g.sa_slots = sa_slots; g.sb_slots = same ? sa_slots : sb_slots;
  g.same = same;
  g.nlines = zg_lines(na, g.nb, same);
  g.ntiles = a.n * (a.n / 2);
  for (int s = 0; s < 32; ++s) {
    g.qa[s] = qa[s];
    g.qb[s] = qb[s];
  }
  g.fft = a;
  g.HA = (const float4*)HA;
  g.HB = (const float4*)(same ? HA : HB);
  g.G = G;
  g.tw = (const float2*)tw;
  g.partial = partial;
  g.partial_bb = partial + (int64_t)nblk * 1024;  // used by BB launches only
  hipError_t e;
  g.Qi = zgf_items_per_line(a.n, Q);
  g.blocked = half_blocked(a.n, Q);
  g.nkb = (Q + 7) / 8;
  // fixed plan 400 = 20 x 20 (the only reader of the blocked layout); its
  // fetch words hold offsets into one record (a tile's run, or an x plane
  // when blocked) below 2^30
  const int64_t slot_bytes = g.blocked ? (int64_t)g.nkb * a.n * 64 : (int64_t)Q * 16;
  const bool fixed = a.n == 400 && zgf_maxf(g.Qi, na, g.nb, same) > 0 &&
                     (int64_t)(g.sa_slots > g.sb_slots ? g.sa_slots : g.sb_slots) * slot_bytes < (1 << 30);
  if (g.blocked && !fixed) return hipErrorInvalidValue;
  // the fixed plan's FFT passes take one butterfly per thread
  if (fixed && g.nlines * 20 > ZGF_THREADS) return hipErrorInvalidValue;
  if (fixed) {
    const size_t lds = zgemm_fixed_lds_bytes<20, 20>(zg_lines_alloc(g.nlines, same));
    if (same)
      e = zgf_launch<ZGF_THREADS, true, false, 20, 20>(g, nblk, lds, st);
    else if (Cbb)
      e = zgf_launch<ZGF_THREADS, false, true, 20, 20>(g, nblk, lds, st);
    else
      e = zgf_launch<ZGF_THREADS, false, false, 20, 20>(g, nblk, lds, st);
  } else {
    const size_t lds = zgemm_lds_bytes(a.n, zg_lines_alloc(g.nlines, same));
    if (same)
      e = zg_launch<ZG_THREADS, true>(g, nblk, lds, st);
    else if (Cbb)
      e = zg_launch<ZG_THREADS, false, true>(g, nblk, lds, st);
    else
      e = zg_launch<ZG_THREADS, false>(g, nblk, lds, st);
  }
  if (e != hipSuccess) return e;
  e = launch_gemm_finalize(partial, nblk, 0, 0, na, g.nb, 1.0, C, g.nb, st);
  if (e != hipSuccess || !Cbb) return e;
  return launch_gemm_finalize(g.partial_bb, nblk, 0, 0, g.nb, g.nb, 1.0, Cbb, g.nb, st);
}

}  // namespace cnpc
