// Dense complex linear algebra for the bemret chain on gfx950.
//
// - k_cgemm: complex GEMM on MFMA (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32),
//   64x64 output tile per 256-thread workgroup, 2x2 waves of 32x32, LDS-staged
//   16-deep k slices with register prefetch of the next slice.
// - k_getf2_coop: cooperative (grid-synchronised) LU of a tall panel with
//   partial pivoting; the panel stays in LDS across the grid and each column
//   costs one grid barrier (candidate rows are published with their values).
// - recursive LU / triangular solves: all O(n^3) work lands in k_cgemm.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include "mnp_linalg.hpp"
#include "mnp_ludiag.hpp"

namespace cg = cooperative_groups;

namespace mnp {

// ------------------------------------------------------------ MFMA traits
template <class T> struct MF;
// Both 16x16x4 variants take A[l % 16][l / 16] and B[l / 16][l % 16] from
// lane l.  Their accumulators differ: result r of lane l is
// D[l / 16 + 4 r][l % 16] for f64 and D[4 (l / 16) + r][l % 16] for f32.
template <> struct MF<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct MF<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};


// ------------------------------------------------------------------ GEMM
// C = alpha A B + beta C on MFMA (operand layouts: MF above).  A complex
// product is four real MFMA products (re re, -im im -> re; re im, im re -> im).
//
// Tile: 2 x 2 waves, each wave WM x WN blocks of 16 x 16, so the workgroup
// computes (32 WM) x (32 WN); k advances in slices of 16.  The LDS holds two
// slices: while the MFMAs consume one, the next slice (prefetched into
// registers during the previous step) is stored into the other, so each
// step needs a single barrier.
constexpr int GBK = 16;

// THREE_M selects the 3-multiplication product: P1 = Ar Br, P2 = Ai Bi,
// P3 = (Ar + Ai)(Br + Bi), Cr = P1 - P2, Ci = P3 - P1 - P2 (3 MFMA per 4).
template <class T, int WM, int WN, bool THREE_M>
__global__ void __launch_bounds__(256) k_cgemm(int64_t M, int64_t N, int64_t K, cplx<T> alpha,
                                               const cplx<T>* A, int64_t lda,
                                               const cplx<T>* B, int64_t ldb,
                                               cplx<T> beta, int beta_zero, cplx<T>* C,
                                               int64_t ldc, int64_t sa, int64_t sb, int64_t sc) {
  using acc_t = typename MF<T>::acc_t;
  // batch index blockIdx.z: operands at constant element strides
  A += blockIdx.z * sa;
  B += blockIdx.z * sb;
  C += blockIdx.z * sc;
  constexpr int BM = 32 * WM, BN = 32 * WN, PAD = 1;
  constexpr int LA = BM * GBK / 256, LB = BN * GBK / 256;  // elements per thread
  __shared__ T sAr[2][GBK][BM + PAD], sAi[2][GBK][BM + PAD];
  __shared__ T sBr[2][GBK][BN + PAD], sBi[2][GBK][BN + PAD];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid & 1, wn = wid >> 1;
  // Workgroups are dispatched round-robin over the 8 XCDs (each with its own
  // L2).  Renumber them so that each XCD gets a contiguous range of tiles,
  // and walk that range in groups of 8 tile rows so neighbouring workgroups
  // share A row panels and B column panels in the XCD's L2.
  int64_t tx, ty;
  {
    const int64_t gx = gridDim.x, gy = gridDim.y, tot = gx * gy;
    const int64_t L = blockIdx.x + gx * (int64_t)blockIdx.y;
    const int64_t per = tot / 8, rem = tot % 8, xcd = L % 8, idx = L / 8;
    const int64_t nl = xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
    constexpr int64_t G = 8;
    const int64_t group = nl / (G * gy), first = group * G;
    const int64_t gsz = min(G, gx - first), r = nl % (G * gy);
    tx = first + r % gsz;
    ty = r / gsz;
  }
  const int64_t m0 = tx * BM, n0 = ty * BN;

  acc_t cr[WM][WN], ci[WM][WN], c3[THREE_M ? WM : 1][THREE_M ? WN : 1];
#pragma unroll
  for (int a = 0; a < WM; ++a)
#pragma unroll
    for (int b = 0; b < WN; ++b) {
      cr[a][b] = acc_t{0, 0, 0, 0};
      ci[a][b] = acc_t{0, 0, 0, 0};
      if constexpr (THREE_M) c3[a][b] = acc_t{0, 0, 0, 0};
    }
  // Loads read clamped (always valid) addresses without branches, so that
  // they stay in flight during the MFMAs; entries outside the operands are
  // zeroed when the slice is stored.  A predicated load would be compiled to
  // a branch that waits for it at once.
  cplx<T> ra[LA], rb[LB];
  auto loadA = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < LA; ++e) {
      const int idx = tid + 256 * e, m = idx % BM, k = idx / BM;
      const int64_t gm = min(m0 + m, M - 1), gk = min(k0 + k, K - 1);
      ra[e] = A[gk * lda + gm];
    }
  };
  auto loadB = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < LB; ++e) {
      const int idx = tid + 256 * e, k = idx % GBK, n = idx / GBK;
      const int64_t gk = min(k0 + k, K - 1), gn = min(n0 + n, N - 1);
      rb[e] = B[gn * ldb + gk];
    }
  };
  auto store = [&](int buf, int64_t k0) {
#pragma unroll
    for (int e = 0; e < LA; ++e) {
      const int idx = tid + 256 * e, m = idx % BM, k = idx / BM;
      const cplx<T> v = (m0 + m < M && k0 + k < K) ? ra[e] : czero<T>();
      sAr[buf][k][m] = v.x;
      sAi[buf][k][m] = v.y;
    }
#pragma unroll
    for (int e = 0; e < LB; ++e) {
      const int idx = tid + 256 * e, k = idx % GBK, n = idx / GBK;
      const cplx<T> v = (k0 + k < K && n0 + n < N) ? rb[e] : czero<T>();
      sBr[buf][k][n] = v.x;
      sBi[buf][k][n] = v.y;
    }
  };
  if (K > 0) {
    loadA(0);
    loadB(0);
    store(0, 0);
  }
  int buf = 0;
  for (int64_t k0 = 0; k0 < K; k0 += GBK) {
    __syncthreads();  // slice `buf` is stored; slice `buf ^ 1` is no longer read
    const bool more = k0 + GBK < K;
    if (more) {
      loadA(k0 + GBK);
      loadB(k0 + GBK);
    }
#pragma unroll
    for (int kk = 0; kk < GBK / 4; ++kk) {
      const int kr = kk * 4 + (lane >> 4), cl = lane & 15;
      T ar[WM], ai[WM], nai[WM], br[WN], bi[WN];
#pragma unroll
      for (int a = 0; a < WM; ++a) {
        ar[a] = sAr[buf][kr][wm * 16 * WM + a * 16 + cl];
        ai[a] = sAi[buf][kr][wm * 16 * WM + a * 16 + cl];
        nai[a] = -ai[a];
      }
#pragma unroll
      for (int b = 0; b < WN; ++b) {
        br[b] = sBr[buf][kr][wn * 16 * WN + b * 16 + cl];
        bi[b] = sBi[buf][kr][wn * 16 * WN + b * 16 + cl];
      }
      if constexpr (THREE_M) {
        // cr accumulates P1, ci accumulates P2, c3 accumulates P3
#pragma unroll
        for (int a = 0; a < WM; ++a)
#pragma unroll
          for (int b = 0; b < WN; ++b) {
            cr[a][b] = MF<T>::mma(ar[a], br[b], cr[a][b]);
            ci[a][b] = MF<T>::mma(ai[a], bi[b], ci[a][b]);
            c3[a][b] = MF<T>::mma(ar[a] + ai[a], br[b] + bi[b], c3[a][b]);
          }
      } else {
#pragma unroll
        for (int a = 0; a < WM; ++a)
#pragma unroll
          for (int b = 0; b < WN; ++b) {
            cr[a][b] = MF<T>::mma(ar[a], br[b], cr[a][b]);
            ci[a][b] = MF<T>::mma(ar[a], bi[b], ci[a][b]);
          }
#pragma unroll
        for (int a = 0; a < WM; ++a)
#pragma unroll
          for (int b = 0; b < WN; ++b) {
            cr[a][b] = MF<T>::mma(nai[a], bi[b], cr[a][b]);
            ci[a][b] = MF<T>::mma(ai[a], br[b], ci[a][b]);
          }
      }
    }
    if (more) store(buf ^ 1, k0 + GBK);
    buf ^= 1;
  }
#pragma unroll
  for (int a = 0; a < WM; ++a)
#pragma unroll
    for (int b = 0; b < WN; ++b) {
      const int64_t col = n0 + wn * 16 * WN + b * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm * 16 * WM + a * 16 + MF<T>::row(lane, r);
        if (row < M && col < N) {
          cplx<T> v;
          if constexpr (THREE_M)
            v = mk<T>(cr[a][b][r] - ci[a][b][r], c3[a][b][r] - cr[a][b][r] - ci[a][b][r]);
          else
            v = mk<T>(cr[a][b][r], ci[a][b][r]);
          cplx<T> o = cmul<T>(alpha, v);
          cplx<T>* pc = C + col * ldc + row;
          if (!beta_zero) o = cadd<T>(o, cmul<T>(beta, *pc));
          *pc = o;
        }
      }
    }
}

typedef __attribute__((address_space(3))) void* lds_ptr_t;
template <int S> using stage_t = std::integral_constant<int, S>;

// ------------------------------------------ FP32 3M GEMM, direct-to-LDS loads
// The 3M product of k_cgemm for complex64 operands, with the slices streamed
// from HBM straight into LDS by global_load_lds_dword (lane l of a wave
// writes bytes 4 l .. 4 l + 3 of the wave's 256-byte LDS destination), two
// stages deep, in the operands' own interleaved (re, im) order: a fragment
// read is one ds_read_b64, and no registers or VALU carry the slices.
//
// A stage: the 16 k-rows of the tile's BM rows, row kr at sA[kr][0 .. BM), one
// complex of padding per row.  Wave w loads rows 32 w .. 32 w + 31 of every
// k-row: lane l fetches component l & 1 of row m0 + 32 w + l / 2 (clamped to
// M - 1) in column k0 + kr (clamped to K - 1).
// B stage: column c of the tile at sB[c][0 .. 16); one wave load covers
// columns 2 j, 2 j + 1, lane l fetching component l & 1 of k = (l % 32 / 2 + c)
// mod 16 into position l % 32 / 2.  Element (k, c) thus sits at position
// (k - c) mod 16, which spreads the 16 columns of a fragment read over the
// banks.  Rows, columns and k beyond the operands read clamped (valid)
// addresses; the k-rows of A beyond K are zeroed in LDS before the last
// slice's MFMAs, which cancels whatever B holds there.
constexpr int FBK = 16;
template <int WM, int WN>
__global__ void __launch_bounds__(256) k_cgemm_f32lds(int64_t M, int64_t N, int64_t K, float2 alpha,
                                                      const float* __restrict__ A, int64_t lda,
                                                      const float* __restrict__ B, int64_t ldb, float2 beta,
                                                      int beta_zero, float2* C, int64_t ldc) {
  using acc_t = MF<float>::acc_t;
  constexpr int BM = 32 * WM, BN = 32 * WN;
  constexpr int NB = BN / 2 / 4;  // B wave loads per wave and slice
  static_assert(BM == 128 && BN % 8 == 0, "four 32-row A wave loads per k-row, B columns in pairs per wave");
  __shared__ float2 sA0[FBK][BM + 1], sA1[FBK][BM + 1];
  __shared__ float2 sB0[BN][FBK], sB1[BN][FBK];
  auto stA = [&](auto st) -> float2(&)[FBK][BM + 1] {
    if constexpr (decltype(st)::value == 0) return sA0; else return sA1;
  };
  auto stB = [&](auto st) -> float2(&)[BN][FBK] {
    if constexpr (decltype(st)::value == 0) return sB0; else return sB1;
  };
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid & 1, wn = wid >> 1;
  int64_t tx, ty;
  {  // XCD-aware tile order, as in k_cgemm
    const int64_t gx = gridDim.x, gy = gridDim.y, tot = gx * gy;
    const int64_t L = blockIdx.x + gx * (int64_t)blockIdx.y;
    const int64_t per = tot / 8, rem = tot % 8, xcd = L % 8, idx = L / 8;
    const int64_t nl = xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
    constexpr int64_t G = 8;
    const int64_t group = nl / (G * gy), first = group * G;
    const int64_t gsz = min(G, gx - first), r = nl % (G * gy);
    tx = first + r % gsz;
    ty = r / gsz;
  }
  const int64_t m0 = tx * BM, n0 = ty * BN;
  // per-lane element offsets (floats) of the wave loads; B's relative to its
  // tile's first column (below 2^31: checked by the launcher)
  const int64_t aoff = min(m0 + 32 * wid + (lane >> 1), M - 1) * 2 + (lane & 1);
  const float* Bt = B + n0 * ldb * 2;
  const int kb = ((lane & 31) >> 1) + 2 * wid + (lane >> 5);  // k of load j: (kb + 8 j) mod 16
  int boff[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int c = 2 * (wid + 4 * j) + (lane >> 5);
    boff[j] = (int)((min(n0 + c, N - 1) - n0) * ldb * 2) + (lane & 1);
  }
  acc_t cr[WM][WN], ci[WM][WN], c3[WM][WN];
#pragma unroll
  for (int a = 0; a < WM; ++a)
#pragma unroll
    for (int b = 0; b < WN; ++b) {
      cr[a][b] = acc_t{0, 0, 0, 0};
      ci[a][b] = acc_t{0, 0, 0, 0};
      c3[a][b] = acc_t{0, 0, 0, 0};
    }
  auto issue = [&](auto st, int64_t k0) {
    auto& dA = stA(st);
    auto& dB = stB(st);
    const bool full = k0 + FBK <= K;
#pragma unroll
    for (int kr = 0; kr < FBK; ++kr) {
      const int64_t gk = min(k0 + kr, K - 1);
      __builtin_amdgcn_global_load_lds((const void*)(A + gk * lda * 2 + aoff), (lds_ptr_t)&dA[kr][32 * wid], 4, 0,
                                       0);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int kj = (kb + 8 * j) & 15;
      const int64_t gk = full ? k0 + kj : min(k0 + kj, K - 1);
      __builtin_amdgcn_global_load_lds((const void*)(Bt + boff[j] + gk * 2), (lds_ptr_t)&dB[2 * (wid + 4 * j)][0], 4,
                                       0, 0);
    }
  };
  const int r16 = lane & 15, kq = lane >> 4;
  auto compute = [&](auto st) {
    auto& xA = stA(st);
    auto& xB = stB(st);
#pragma unroll
    for (int kk = 0; kk < FBK / 4; ++kk) {
      const int kr = kk * 4 + kq;
      float ar[WM], ai[WM], br[WN], bi[WN];
#pragma unroll
      for (int a = 0; a < WM; ++a) {
        const float2 v = xA[kr][wm * 16 * WM + a * 16 + r16];
        ar[a] = v.x;
        ai[a] = v.y;
      }
#pragma unroll
      for (int b = 0; b < WN; ++b) {
        const int col = wn * 16 * WN + b * 16 + r16;
        const float2 v = xB[col][(kr - col) & 15];
        br[b] = v.x;
        bi[b] = v.y;
      }
#pragma unroll
      for (int a = 0; a < WM; ++a)
#pragma unroll
        for (int b = 0; b < WN; ++b) {
          cr[a][b] = MF<float>::mma(ar[a], br[b], cr[a][b]);
          ci[a][b] = MF<float>::mma(ai[a], bi[b], ci[a][b]);
          c3[a][b] = MF<float>::mma(ar[a] + ai[a], br[b] + bi[b], c3[a][b]);
        }
    }
  };
  // slice starting at k0 lives in stage (k0 / FBK) % 2
  auto step = [&](auto st, int64_t k0) {
    __builtin_amdgcn_s_waitcnt(0);  // this wave's loads of the slice have landed
    __syncthreads();                // all waves' have; nobody reads the other stage any more
    if (k0 + FBK < K) issue(stage_t<1 - decltype(st)::value>{}, k0 + FBK);
    if (k0 + FBK > K) {             // last, partial slice: zero A's k-rows beyond K
      auto& xA = stA(st);
      const int kv = (int)(K - k0);
      for (int e = tid; e < (FBK - kv) * BM; e += 256) xA[kv + e / BM][e % BM] = make_float2(0.f, 0.f);
      __syncthreads();
    }
    compute(st);
  };
  if (K > 0) issue(stage_t<0>{}, 0);
  for (int64_t k0 = 0; k0 < K; k0 += 2 * FBK) {
    step(stage_t<0>{}, k0);
    if (k0 + FBK < K) step(stage_t<1>{}, k0 + FBK);
  }
#pragma unroll
  for (int a = 0; a < WM; ++a)
#pragma unroll
    for (int b = 0; b < WN; ++b) {
      const int64_t col = n0 + wn * 16 * WN + b * 16 + r16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm * 16 * WM + a * 16 + MF<float>::row(lane, r);
        if (row < M && col < N) {
          const float2 v = make_float2(cr[a][b][r] - ci[a][b][r], c3[a][b][r] - cr[a][b][r] - ci[a][b][r]);
          float2 o = cmul<float>(alpha, v);
          float2* pc = C + col * ldc + row;
          if (!beta_zero) o = cadd<float>(o, cmul<float>(beta, *pc));
          *pc = o;
        }
      }
    }
}

// ------------------------------------------------ split-bf16 complex GEMM
// C = alpha A B + beta C for complex64 operands on bf16 MFMA.  Each FP32
// value x is held as a pair of bf16, hi = rn(x) and lo = rn(x - hi), so
// x = hi + lo to about 2^-17 |x|, and each real product is formed as
// Xh Yh + Xh Yl + Xl Yh (the dropped Xl Yl is below 2^-16 |X||Y|) by
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  The complex product takes
// four real ones (Cr = Ar Br - Ai Bi, Ci = Ar Bi + Ai Br): 12 MFMAs per
// 16 x 16 x 32 complex block, whose 196,608 bf16 flops stand for 65,536
// complex-algorithmic ones.
//
// The operands are split once, before the product, by one HBM pass each
// (k_split) into four planes (re hi, re lo, im hi, im lo).  A
// plane of an operand with R rows (A: its rows; B: its columns) is an array
// of 16-byte chunks [KS][Rp][4]: chunk (s, r, g) holds the 8 values
// k = 32 s + 8 g .. 32 s + 8 g + 7 of row r, zero outside the operand; Rp is
// R rounded up to the tile and KS = ceil(K / 32).  A 32-deep slice of 16
// rows is thus 1 KB of contiguous memory.
//
// Operand layout of v_mfma_f32_16x16x32_bf16: lane l holds 8 consecutive-k
// values A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16]; the result
// of lane l is D[4 (l / 16) + r][l % 16], r = 0..3.  A and B use the same
// k slots, so the sum does not depend on the hardware's order of k inside a
// lane.  In LDS a row (or column) holds the four chunks of a slice; chunk g
// of row m sits at slot g ^ lds_swz(m % 16), which keeps the fragment reads
// free of bank conflicts.
//
// The product streams slices from HBM straight into LDS with
// global_load_lds_dwordx4 (lane l of a wave writes bytes 16 l .. 16 l + 15 of
// the wave's LDS destination), NS stages deep: while the MFMAs consume one
// stage, the loads of the next NS - 1 slices are in flight.  A wave load covers
// 16 rows of one plane, its lane l fetching chunk g = (l % 4) ^ swizzle of
// row l / 4, so that the swizzle above is applied by the fetch itself.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
// 16-byte chunk as a native vector (HIP's uint4 is a union-based struct,
// which keeps per-thread arrays of it out of registers)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void split_pair(float a, float b, uint32_t& hi, uint32_t& lo) {
  const bf16x2_t h = {(__bf16)a, (__bf16)b};
  hi = __builtin_bit_cast(uint32_t, h);
  const float ah = __builtin_bit_cast(float, hi << 16), bh = __builtin_bit_cast(float, hi & 0xffff0000u);
  const bf16x2_t l = {(__bf16)(a - ah), (__bf16)(b - bh)};
  lo = __builtin_bit_cast(uint32_t, l);
}

// 8 complex values -> chunks re hi, re lo, im hi, im lo
__device__ __forceinline__ void split8(const float2* v, u32x4* out) {
  uint32_t h[4], l[4], ih[4], il[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    split_pair(v[2 * q].x, v[2 * q + 1].x, h[q], l[q]);
    split_pair(v[2 * q].y, v[2 * q + 1].y, ih[q], il[q]);
  }
  out[0] = u32x4{h[0], h[1], h[2], h[3]};
  out[1] = u32x4{l[0], l[1], l[2], l[3]};
  out[2] = u32x4{ih[0], ih[1], ih[2], ih[3]};
  out[3] = u32x4{il[0], il[1], il[2], il[3]};
}

__device__ __forceinline__ f32x4_t mfma_bf16(u32x4 a, u32x4 b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                 __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
}

// Planes of A (R x K, column-major) and of B (K x R, column-major; its
// columns are the plane's rows).  A wave covers 16 rows of one 32-deep slice:
// lane l takes row l / 4 and k-group l % 4, so that each of its four plane
// stores writes 1 KB of contiguous memory.  Reads are clamped to the operand
// (always valid, no branches) and zeroed outside it.
template <bool B_OPERAND>
__global__ void __launch_bounds__(256) k_split(int64_t R, int64_t K, const float2* X, int64_t ld, int64_t Rp,
                                               int64_t KS, u32x4* P) {
  const int64_t r = blockIdx.x * (int64_t)64 + (threadIdx.x >> 2), sl = blockIdx.y;
  const int g = threadIdx.x & 3;
  if (r >= Rp) return;
  const int64_t k0 = 32 * sl + 8 * g, rc = min(r, R - 1);
  float2 v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int64_t k = min(k0 + j, K - 1);
    v[j] = B_OPERAND ? X[rc * ld + k] : X[k * ld + rc];
  }
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if (r >= R || k0 + j >= K) v[j] = make_float2(0.f, 0.f);
  u32x4 c[4];
  split8(v, c);
#pragma unroll
  for (int p = 0; p < 4; ++p) P[((p * KS + sl) * Rp + r) * 4 + g] = c[p];
}


// LDS slot of k-group g of row r: g ^ lds_swz(r % 16).  A fragment read
// (ds_read_b128: lane l takes row l % 16, k-group l / 16) then hits distinct
// banks both when 8 lanes are served per cycle from 32 banks (rows r, r + 2,
// r + 4, r + 6 get distinct slots) and when 16 lanes are served from 64 banks
// (rows r, r + 4, r + 8, r + 12 do).
__device__ __forceinline__ int lds_swz(int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

// Each stage is its own __shared__ array and the main loop is unrolled over
// the stages, so every LDS access names its stage at compile time: the
// compiler then sees that the fragment reads of one stage do not alias the
// loads landing in another and does not make them wait for those loads.
template <int WM, int WN, int NS>
__global__ void __launch_bounds__(256) k_cgemm_sbf16(int64_t M, int64_t N, int64_t KS, float2 alpha,
                                                     const u32x4* PA, int64_t Mp, const u32x4* PB,
                                                     int64_t Np, float2 beta, int beta_zero, float2* C,
                                                     int64_t ldc) {
  static_assert(NS == 2 || NS == 3, "two or three stages");
  constexpr int BM = 32 * WM, BN = 32 * WN;
  constexpr int LA = 4 * BM / 16, LB = 4 * BN / 16;  // wave loads per slice (4 planes x 16-row groups)
  static_assert(LA % 4 == 0 && LB % 4 == 0, "wave loads must split evenly over the 4 waves");
  // loads per wave and stage; waiting for the oldest stage leaves the loads
  // of the NS - 2 younger ones outstanding (s_waitcnt: vmcnt in bits 3:0 and
  // 15:14, expcnt 6:4 and lgkmcnt 11:8 left at their maxima)
  constexpr int PER = LA / 4 + LB / 4, OUT = (NS - 2) * PER;
  static_assert(OUT < 64, "vmcnt holds at most 63 outstanding loads");
  constexpr int WAIT_OLDEST = (OUT & 15) | ((OUT >> 4) << 14) | (7 << 4) | (15 << 8);
  __shared__ u32x4 sA0[4][BM][4], sA1[4][BM][4], sA2[NS > 2 ? 4 : 1][BM][4];
  __shared__ u32x4 sB0[4][BN][4], sB1[4][BN][4], sB2[NS > 2 ? 4 : 1][BN][4];
  auto stA = [&](auto st) -> u32x4(&)[4][BM][4] {
    if constexpr (decltype(st)::value == 0) return sA0;
    else if constexpr (decltype(st)::value == 1) return sA1;
    else return reinterpret_cast<u32x4(&)[4][BM][4]>(sA2);  // full size when NS == 3, the only use
  };
  auto stB = [&](auto st) -> u32x4(&)[4][BN][4] {
    if constexpr (decltype(st)::value == 0) return sB0;
    else if constexpr (decltype(st)::value == 1) return sB1;
    else return reinterpret_cast<u32x4(&)[4][BN][4]>(sB2);
  };
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid & 1, wn = wid >> 1;
  int64_t tx, ty;
  {  // XCD-aware tile order, as in k_cgemm
    const int64_t gx = gridDim.x, gy = gridDim.y, tot = gx * gy;
    const int64_t L = blockIdx.x + gx * (int64_t)blockIdx.y;
    const int64_t per = tot / 8, rem = tot % 8, xcd = L % 8, idx = L / 8;
    const int64_t nl = xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
    constexpr int64_t G = 4;
    const int64_t group = nl / (G * gy), first = group * G;
    const int64_t gsz = min(G, gx - first), r = nl % (G * gy);
    tx = first + r % gsz;
    ty = r / gsz;
  }
  const int64_t m0 = tx * BM, n0 = ty * BN;

  f32x4_t cr[WM][WN], ci[WM][WN];
#pragma unroll
  for (int a = 0; a < WM; ++a)
#pragma unroll
    for (int b = 0; b < WN; ++b) {
      cr[a][b] = f32x4_t{0, 0, 0, 0};
      ci[a][b] = f32x4_t{0, 0, 0, 0};
    }
  // lane part of a wave load: row lane / 4 of the 16-row group, chunk
  // (lane % 4) ^ swz(row) (groups start at multiples of 16)
  const int lrow = lane >> 2, lchunk = (lane & 3) ^ lds_swz(lrow);
  auto issue = [&](auto st, int64_t sl) {
    auto& dA = stA(st);
    auto& dB = stB(st);
#pragma unroll
    for (int i = 0; i < LA / 4; ++i) {
      const int w = wid + 4 * i, p = w / (BM / 16), rg = w % (BM / 16);
      const u32x4* g = PA + ((p * KS + sl) * Mp + m0 + 16 * rg + lrow) * 4 + lchunk;
      __builtin_amdgcn_global_load_lds((const void*)g, (lds_ptr_t)&dA[p][16 * rg][0], 16, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < LB / 4; ++i) {
      const int w = wid + 4 * i, p = w / (BN / 16), rg = w % (BN / 16);
      const u32x4* g = PB + ((p * KS + sl) * Np + n0 + 16 * rg + lrow) * 4 + lchunk;
      __builtin_amdgcn_global_load_lds((const void*)g, (lds_ptr_t)&dB[p][16 * rg][0], 16, 0, 0);
    }
  };
  // fragment slot of this lane: k-group lane / 16 of rows (columns) base + lane % 16
  const int r16 = lane & 15, slot = (lane >> 4) ^ lds_swz(r16);
  const uint32_t sgn = 0x80008000u;
  auto compute = [&](auto st) {
    auto& xA = stA(st);
    auto& xB = stB(st);
    u32x4 brh[WN], brl[WN], bih[WN], bil[WN];
#pragma unroll
    for (int b = 0; b < WN; ++b) {
      const int col = wn * 16 * WN + b * 16 + r16;
      brh[b] = xB[0][col][slot];
      brl[b] = xB[1][col][slot];
      bih[b] = xB[2][col][slot];
      bil[b] = xB[3][col][slot];
    }
#pragma unroll
    for (int a = 0; a < WM; ++a) {
      const int row = wm * 16 * WM + a * 16 + r16;
      const u32x4 arh = xA[0][row][slot], arl = xA[1][row][slot];
      const u32x4 aih = xA[2][row][slot], ail = xA[3][row][slot];
      const u32x4 nih = aih ^ sgn, nil = ail ^ sgn;
#pragma unroll
      for (int b = 0; b < WN; ++b) {
        cr[a][b] = mfma_bf16(arh, brh[b], cr[a][b]);
        ci[a][b] = mfma_bf16(arh, bih[b], ci[a][b]);
        cr[a][b] = mfma_bf16(arh, brl[b], cr[a][b]);
        ci[a][b] = mfma_bf16(arh, bil[b], ci[a][b]);
        cr[a][b] = mfma_bf16(arl, brh[b], cr[a][b]);
        ci[a][b] = mfma_bf16(arl, bih[b], ci[a][b]);
        cr[a][b] = mfma_bf16(nih, bih[b], cr[a][b]);
        ci[a][b] = mfma_bf16(aih, brh[b], ci[a][b]);
        cr[a][b] = mfma_bf16(nih, bil[b], cr[a][b]);
        ci[a][b] = mfma_bf16(aih, brl[b], ci[a][b]);
        cr[a][b] = mfma_bf16(nil, bih[b], cr[a][b]);
        ci[a][b] = mfma_bf16(ail, brh[b], ci[a][b]);
      }
    }
  };
  // slice sl lives in stage sl % NS
  auto step = [&](auto st, int64_t sl) {
    constexpr int S = decltype(st)::value;
    // this wave's loads of slice sl have landed (near the end, where fewer
    // younger loads are outstanding, wait for all); after the barrier all
    // waves' have, and no wave still reads the stage of slice sl - 1
    if (sl + NS - 2 < KS)
      __builtin_amdgcn_s_waitcnt(WAIT_OLDEST);
    else
      __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (sl + NS - 1 < KS) issue(stage_t<(S + NS - 1) % NS>{}, sl + NS - 1);
    compute(st);
  };
  // prologue: slices 0 .. NS - 2 in flight
  if (KS > 0) issue(stage_t<0>{}, 0);
  if constexpr (NS == 3)
    if (KS > 1) issue(stage_t<1>{}, 1);
  for (int64_t sl = 0; sl < KS; sl += NS) {
    step(stage_t<0>{}, sl);
    if (sl + 1 < KS) step(stage_t<1>{}, sl + 1);
    if constexpr (NS == 3)
      if (sl + 2 < KS) step(stage_t<2>{}, sl + 2);
  }
#pragma unroll
  for (int a = 0; a < WM; ++a)
#pragma unroll
    for (int b = 0; b < WN; ++b) {
      const int64_t col = n0 + wn * 16 * WN + b * 16 + r16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm * 16 * WM + a * 16 + 4 * (lane >> 4) + r;
        if (row < M && col < N) {
          const float2 v = make_float2(cr[a][b][r], ci[a][b][r]);
          float2 o = cmul<float>(alpha, v);
          float2* pc = C + col * ldc + row;
          if (!beta_zero) o = cadd<float>(o, cmul<float>(beta, *pc));
          *pc = o;
        }
      }
    }
}

// Split planes of the operands of the current product (reused by the next;
// launches on one stream are ordered).
// Per thread (concurrent tasks run their products on their own streams).
// Never destroyed: a static or thread-local destructor would call hipFree
// while the process exits, after the HIP runtime or a profiler attached to it
// may already be torn down; the device memory goes with the process.
static Workspace& split_ws() {
  static thread_local Workspace* ws = new Workspace;
  return *ws;
}

template <int WM, int WN, int NS>
static void gemm_sbf16(int64_t m, int64_t n, int64_t k, float2 al, MatView<float> A, MatView<float> B,
                       float2 be, bool bz, MatView<float> C) {
  constexpr int BM = 32 * WM, BN = 32 * WN;
  const int64_t Mp = cdiv_up(m, BM) * BM, Np = cdiv_up(n, BN) * BN, KS = cdiv_up(k, 32);
  u32x4 *PA = nullptr, *PB = nullptr;
  if (KS > 0) {
    MNP_CHECK(KS <= 65535, MNP_ERR_ARG, "gemm: k too large for the split-bf16 product");
    PA = (u32x4*)split_ws().get(0, sizeof(u32x4) * 16 * (size_t)(KS * Mp));
    PB = (u32x4*)split_ws().get(1, sizeof(u32x4) * 16 * (size_t)(KS * Np));
    hipLaunchKernelGGL(k_split<false>, dim3((unsigned)(Mp / 64), (unsigned)KS), dim3(256), 0, stream(), m, k, A.p,
                       A.ld, Mp, KS, PA);
    hipLaunchKernelGGL(k_split<true>, dim3((unsigned)(Np / 64), (unsigned)KS), dim3(256), 0, stream(), n, k, B.p,
                       B.ld, Np, KS, PB);
    MNP_HIP(hipGetLastError());
  }
  const int64_t gx = Mp / BM, ny = Np / BN;
  for (int64_t y0 = 0; y0 < ny; y0 += 65535) {
    const int64_t gy = std::min<int64_t>(65535, ny - y0);
    const int64_t noff = y0 * BN;
    const int64_t ncols = std::min<int64_t>(n - noff, gy * BN);
    TimerScope ts(MNP_TIMER_GEMM_SBF16, 8.0 * (double)m * (double)ncols * (double)k,
                  (double)sizeof(float2) * ((double)m * k + (double)k * ncols + (double)m * ncols * (bz ? 1.0 : 2.0)));
    hipLaunchKernelGGL((k_cgemm_sbf16<WM, WN, NS>), dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, stream(), m,
                       n - noff, KS, al, PA, Mp, PB ? PB + noff * 4 : PB, Np, be, (int)bz, C.p + noff * C.ld, C.ld);
  }
  MNP_HIP(hipGetLastError());
}

// ------------------------------------------------ tall-skinny products
// C = alpha A B + beta C with few right-hand sides (n <= 16): the refinement
// residual (the FP64 operator times n x 4 npol vectors), the chain's products
// with excitation vectors and the updates of the triangular solves for few
// right-hand sides.  Such a product streams A once and is bound by HBM; MFMA
// tiles would use a quarter of their width or less, on too few workgroups.
// One thread per row of A (coalesced column reads), 16-deep slices of B
// broadcast from LDS, k split over blockIdx.y into partial sums that
// k_skinny_sum adds in a fixed order, so results do not depend on timing.
constexpr int SKT = 16;  // k per LDS slice
constexpr int SNR = 16;  // most right-hand sides

template <class T, int NR>
__global__ void __launch_bounds__(256) k_skinny(int64_t m, int n, int64_t k, int64_t kc, const cplx<T>* A,
                                                int64_t lda, const cplx<T>* B, int64_t ldb, cplx<T>* W,
                                                cplx<T> alpha, cplx<T> beta, int beta_zero, cplx<T>* C,
                                                int64_t ldc, int direct) {
  __shared__ cplx<T> sb[SKT][NR];
  const int64_t row = blockIdx.x * (int64_t)256 + threadIdx.x;
  const int64_t kb = blockIdx.y * kc, ke = min(k, kb + kc);
  const cplx<T>* a_row = A + min(row, m - 1);  // clamped: loads stay valid, unused rows are not stored
  T ar[NR], ai[NR];
#pragma unroll
  for (int j = 0; j < NR; ++j) ar[j] = ai[j] = T(0);
  for (int64_t k0 = kb; k0 < ke; k0 += SKT) {
    __syncthreads();  // the previous slice is no longer read
    for (int e = threadIdx.x; e < SKT * NR; e += 256) {
      const int kk = e % SKT, j = e / SKT;
      sb[kk][j] = (k0 + kk < ke && j < n) ? B[j * ldb + k0 + kk] : czero<T>();
    }
    cplx<T> a[SKT];
#pragma unroll
    for (int kk = 0; kk < SKT; ++kk) a[kk] = a_row[min(k0 + kk, ke - 1) * lda];
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < SKT; ++kk) {
      const cplx<T> x = k0 + kk < ke ? a[kk] : czero<T>();
#pragma unroll
      for (int j = 0; j < NR; ++j) {
        const cplx<T> b = sb[kk][j];
        ar[j] = fma(x.x, b.x, ar[j]);
        ar[j] = fma(-x.y, b.y, ar[j]);
        ai[j] = fma(x.x, b.y, ai[j]);
        ai[j] = fma(x.y, b.x, ai[j]);
      }
    }
  }
  if (row >= m) return;
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    if (j >= n) break;
    const cplx<T> v = mk<T>(ar[j], ai[j]);
    if (direct) {
      cplx<T>* pc = C + j * ldc + row;
      cplx<T> o = cmul<T>(alpha, v);
      if (!beta_zero) o = cadd<T>(o, cmul<T>(beta, *pc));
      *pc = o;
    } else {
      W[(blockIdx.y * (int64_t)n + j) * m + row] = v;
    }
  }
}

// C = alpha (sum over the ks partial products W[s] (n x m each)) + beta C
template <class T>
__global__ void k_skinny_sum(int64_t m, int n, int64_t ks, const cplx<T>* W, cplx<T> alpha, cplx<T> beta,
                             int beta_zero, cplx<T>* C, int64_t ldc) {
  const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (row >= m) return;
  cplx<T> s = czero<T>();
#pragma unroll 8
  for (int64_t q = 0; q < ks; ++q) s = cadd<T>(s, W[(q * n + j) * m + row]);
  cplx<T>* pc = C + j * ldc + row;
  cplx<T> o = cmul<T>(alpha, s);
  if (!beta_zero) o = cadd<T>(o, cmul<T>(beta, *pc));
  *pc = o;
}

template <class T>
static Workspace& skinny_ws() {
  static thread_local Workspace* ws = new Workspace;  // never destroyed, as split_ws()
  return *ws;
}

template <class T>
static void skinny(int64_t m, int64_t n, int64_t k, cplx<T> al, MatView<T> A, MatView<T> B, cplx<T> be, bool bz,
                   MatView<T> C) {
  // about 4,096 workgroups: row blocks x k chunks of at least 64 (each
  // workgroup's chain of dependent slices stays short when m is small)
  const int64_t mt = cdiv_up(m, 256);
  int64_t ks = std::max<int64_t>(1, std::min<int64_t>(cdiv_up(4096, mt), k / 64));
  const int64_t kc = cdiv_up(cdiv_up(k, ks), SKT) * SKT;
  ks = cdiv_up(k, kc);
  MNP_CHECK(ks <= 65535, MNP_ERR_INTERNAL, "skinny: too many k chunks");
  cplx<T>* W = ks > 1 ? (cplx<T>*)skinny_ws<T>().get(0, sizeof(cplx<T>) * (size_t)(ks * n * m)) : nullptr;
  const dim3 grid((unsigned)mt, (unsigned)ks);
  if (n <= 8)
    hipLaunchKernelGGL((k_skinny<T, 8>), grid, dim3(256), 0, stream(), m, (int)n, k, kc, A.p, A.ld, B.p, B.ld, W,
                       al, be, (int)bz, C.p, C.ld, (int)(ks == 1));
  else
    hipLaunchKernelGGL((k_skinny<T, SNR>), grid, dim3(256), 0, stream(), m, (int)n, k, kc, A.p, A.ld, B.p, B.ld,
                       W, al, be, (int)bz, C.p, C.ld, (int)(ks == 1));
  if (ks > 1)
    hipLaunchKernelGGL(k_skinny_sum<T>, dim3((unsigned)cdiv_up(m, 256), (unsigned)n), dim3(256), 0, stream(), m,
                       (int)n, ks, W, al, be, (int)bz, C.p, C.ld);
  MNP_HIP(hipGetLastError());
}

// Products of complex64 matrices on the split-bf16 kernel when set (the
// MNP_SOLVER_MIXED_BF16 chain sets it around its factorisations).
static thread_local int g_split_bf16 = 0;
void set_split_bf16(int on) { g_split_bf16 = on; }
int split_bf16() { return g_split_bf16; }

template <class T, int WM, int WN, bool THREE_M = false>
static void gemm_launch(int64_t m, int64_t n, int64_t k, cplx<T> al, MatView<T> A, MatView<T> B,
                        cplx<T> be, bool bz, MatView<T> C, int64_t batch = 1, int64_t sa = 0,
                        int64_t sb = 0, int64_t sc = 0) {
  constexpr int BM = 32 * WM, BN = 32 * WN;
  const int64_t gx = cdiv_up(m, BM);
  const int64_t ny = cdiv_up(n, BN);
  for (int64_t y0 = 0; y0 < ny; y0 += 65535) {
    const int64_t gy = std::min<int64_t>(65535, ny - y0);
    const int64_t noff = y0 * BN;
    const int64_t ncols = std::min<int64_t>(n - noff, gy * BN);
    // Only the large-tile variant (the chain's dominant kernel) is timed:
    // events around each of the recursion's tens of thousands of small
    // launches would cost more host time than the launches themselves.
    // work = the complex product's 8 m n k real flops, whichever algorithm runs it
    TimerScope ts(WM == 4 ? (sizeof(T) == 8 ? MNP_TIMER_GEMM_FP64 : MNP_TIMER_GEMM_FP32) : -1,
                  8.0 * (double)m * (double)ncols * (double)k * (double)batch,
                  (double)sizeof(cplx<T>) * (double)batch *
                      ((double)m * k + (double)k * ncols + (double)m * ncols * (bz ? 1.0 : 2.0)));
    hipLaunchKernelGGL((k_cgemm<T, WM, WN, THREE_M>), dim3((unsigned)gx, (unsigned)gy, (unsigned)batch), dim3(256), 0,
                       stream(), m, n - noff, k, al, A.p, A.ld, B.p + noff * B.ld, B.ld, be, (int)bz,
                       C.p + noff * C.ld, C.ld, sa, sb, sc);
  }
  MNP_HIP(hipGetLastError());
}

template <int WM, int WN>
static void gemm_f32lds(int64_t m, int64_t n, int64_t k, float2 al, MatView<float> A, MatView<float> B,
                        float2 be, bool bz, MatView<float> C) {
  constexpr int BM = 32 * WM, BN = 32 * WN;
  const int64_t gx = cdiv_up(m, BM), ny = cdiv_up(n, BN);
  for (int64_t y0 = 0; y0 < ny; y0 += 65535) {
    const int64_t gy = std::min<int64_t>(65535, ny - y0);
    const int64_t noff = y0 * BN;
    const int64_t ncols = std::min<int64_t>(n - noff, gy * BN);
    TimerScope ts(MNP_TIMER_GEMM_FP32, 8.0 * (double)m * (double)ncols * (double)k,
                  8.0 * ((double)m * k + (double)k * ncols + (double)m * ncols * (bz ? 1.0 : 2.0)));
    hipLaunchKernelGGL((k_cgemm_f32lds<WM, WN>), dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, stream(), m,
                       n - noff, k, al, (const float*)A.p, A.ld, (const float*)(B.p + noff * B.ld), B.ld, be, (int)bz,
                       C.p + noff * C.ld, C.ld);
  }
  MNP_HIP(hipGetLastError());
}

template <class T>
void gemm(int64_t m, int64_t n, int64_t k, double2 alpha, MatView<T> A, MatView<T> B, double2 beta,
          MatView<T> C) {
  if (m <= 0 || n <= 0) return;
  const bool bz = beta.x == 0 && beta.y == 0;
  const cplx<T> al = mk<T>((T)alpha.x, (T)alpha.y), be = mk<T>((T)beta.x, (T)beta.y);
  if (k <= 0) {
    MNP_CHECK(bz, MNP_ERR_INTERNAL, "gemm: k == 0 with beta != 0 unsupported");
    k = 0;
  }
  // Few right-hand sides: a stream over A, bound by HBM.
  if (n <= SNR && m >= 64 && k >= 256) {
    skinny<T>(m, n, k, al, A, B, be, bz, C);
    return;
  }
  // Large products use the 128 x 128 tile; thin ones keep enough workgroups.
  const int64_t big_tiles = cdiv_up(m, 128) * cdiv_up(n, 128);
  if (m >= 128 && n >= 128 && k >= 64 && big_tiles >= 256) {
    if constexpr (sizeof(T) == 8)
      gemm_launch<T, 4, 2>(m, n, k, al, A, B, be, bz, C);  // 128 x 64: accumulators fit without spills
    else if (g_split_bf16)
      gemm_sbf16<4, 4, 2>(m, n, k, al, A, B, be, bz, C);
    else if (B.ld < ((int64_t)1 << 31) / (2 * 96))
      gemm_f32lds<4, 3>(m, n, k, al, A, B, be, bz, C);  // 3M on the FP32 chain
    else
      gemm_launch<T, 4, 3, true>(m, n, k, al, A, B, be, bz, C);
  } else
    gemm_launch<T, 2, 2>(m, n, k, al, A, B, be, bz, C);
}

// batch products C_z = alpha A_z B_z + beta C_z, X_z = X + z * sx (64 x 64 tiles)
template <class T>
void gemm_batched(int64_t m, int64_t n, int64_t k, double2 alpha, MatView<T> A, int64_t sa, MatView<T> B,
                  int64_t sb, double2 beta, MatView<T> C, int64_t sc, int64_t batch) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  MNP_CHECK(k > 0 && batch < 65536, MNP_ERR_INTERNAL, "gemm_batched: bad sizes");
  const bool bz = beta.x == 0 && beta.y == 0;
  gemm_launch<T, 2, 2>(m, n, k, mk<T>((T)alpha.x, (T)alpha.y), A, B, mk<T>((T)beta.x, (T)beta.y), bz, C,
                       batch, sa, sb, sc);
}

// Products of a batch of equal shapes at element strides (operand z at
// X + z * sx), with the output tile chosen by the caller (the batched
// solver's tuner): GEMM_TILE_64 is 64 x 64 (4M), GEMM_TILE_128x64 128 x 64
// (4M), GEMM_TILE_128x96_3M 128 x 96 on the 3-multiplication form (complex64
// only; complex128 takes 128 x 64).  C may alias A when one output tile spans
// all of C's columns and A's rows are C's rows (or alias B when one tile
// spans all of C's rows and B's columns are C's columns): every workgroup
// reads the aliased rows or columns it writes, all of them before its
// epilogue, and no other workgroup reads them.
template <class T>
void gemm_strided(int64_t m, int64_t n, int64_t k, double2 alpha, MatView<T> A, int64_t sa, MatView<T> B,
                  int64_t sb, double2 beta, MatView<T> C, int64_t sc, int64_t batch, int tile) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  MNP_CHECK(k > 0 && batch < 65536, MNP_ERR_INTERNAL, "gemm_strided: bad sizes");
  const bool bz = beta.x == 0 && beta.y == 0;
  const cplx<T> al = mk<T>((T)alpha.x, (T)alpha.y), be = mk<T>((T)beta.x, (T)beta.y);
  if (tile == GEMM_TILE_64)
    gemm_launch<T, 2, 2>(m, n, k, al, A, B, be, bz, C, batch, sa, sb, sc);
  else if (tile == GEMM_TILE_128x96_3M && sizeof(T) == 4)
    gemm_launch<T, 4, 3, true>(m, n, k, al, A, B, be, bz, C, batch, sa, sb, sc);
  else
    gemm_launch<T, 4, 2>(m, n, k, al, A, B, be, bz, C, batch, sa, sb, sc);
}
template void gemm_strided<float>(int64_t, int64_t, int64_t, double2, MatView<float>, int64_t, MatView<float>,
                                  int64_t, double2, MatView<float>, int64_t, int64_t, int);
template void gemm_strided<double>(int64_t, int64_t, int64_t, double2, MatView<double>, int64_t, MatView<double>,
                                   int64_t, double2, MatView<double>, int64_t, int64_t, int);

// ------------------------------------------------------- cooperative panel
// LU with partial pivoting (pivot = first row of largest |Re|+|Im|, the
// izamax rule of LAPACK getf2) of an m x nb panel.  Workgroup b keeps rows
// [b*chunk, (b+1)*chunk) in LDS for the whole factorisation.
template <class T>
__global__ void __launch_bounds__(256) k_getf2_coop(int64_t m, int nb, cplx<T>* A, int64_t lda,
                                                    int32_t* ipiv, int* info, int info_off,
                                                    int64_t chunk, T* cval, int32_t* cidx,
                                                    cplx<T>* crow, cplx<T>* rowj) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* P = (cplx<T>*)smem_raw;  // chunk x nb, row-major
  cplx<T>* urow = P + chunk * nb;   // pivot row
  __shared__ T rv[256];
  __shared__ int64_t ri[256];
  __shared__ int rw[256];
  cg::grid_group grid = cg::this_grid();
  const int tid = threadIdx.x;
  const int64_t b = blockIdx.x, nblk = gridDim.x;
  const int64_t r0 = b * chunk, r1 = min(m, r0 + chunk), nr = r1 > r0 ? r1 - r0 : 0;
  for (int64_t idx = tid; idx < nr * nb; idx += 256) {
    const int64_t r = idx % nr, c = idx / nr;
    P[r * nb + c] = A[c * lda + r0 + r];
  }
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    const int par = j & 1;
    // local candidate
    T best = T(-1);
    int64_t bi = -1;
    for (int64_t r = tid; r < nr; r += 256) {
      if (r0 + r >= j) {
        const T v = cabs1<T>(P[r * nb + j]);
        if (v > best) {
          best = v;
          bi = r0 + r;
        }
      }
    }
    rv[tid] = best;
    ri[tid] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) {
        const T v2 = rv[tid + s];
        const int64_t i2 = ri[tid + s];
        if (i2 >= 0 && (v2 > rv[tid] || (v2 == rv[tid] && (ri[tid] < 0 || i2 < ri[tid])))) {
          rv[tid] = v2;
          ri[tid] = i2;
        }
      }
      __syncthreads();
    }
    if (tid == 0) {
      cval[par * nblk + b] = rv[0];
      cidx[par * nblk + b] = (int32_t)ri[0];
    }
    if (ri[0] >= 0)
      for (int c = tid; c < nb; c += 256) crow[(par * nblk + b) * nb + c] = P[(ri[0] - r0) * nb + c];
    if (j >= r0 && j < r1)
      for (int c = tid; c < nb; c += 256) rowj[par * nb + c] = P[(j - r0) * nb + c];
    __syncthreads();
    grid.sync();
    // global pivot: every workgroup reduces the candidates redundantly
    best = T(-1);
    bi = -1;
    int bw = -1;
    for (int64_t w = tid; w < nblk; w += 256) {
      const T v = cval[par * nblk + w];
      const int64_t id = cidx[par * nblk + w];
      if (id >= 0 && (v > best || (v == best && (bi < 0 || id < bi)))) {
        best = v;
        bi = id;
        bw = (int)w;
      }
    }
    rv[tid] = best;
    ri[tid] = bi;
    rw[tid] = bw;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) {
        const T v2 = rv[tid + s];
        const int64_t i2 = ri[tid + s];
        if (i2 >= 0 && (v2 > rv[tid] || (v2 == rv[tid] && (ri[tid] < 0 || i2 < ri[tid])))) {
          rv[tid] = v2;
          ri[tid] = i2;
          rw[tid] = rw[tid + s];
        }
      }
      __syncthreads();
    }
    const int64_t p = ri[0];
    const int w = rw[0];
    for (int c = tid; c < nb; c += 256) urow[c] = crow[(par * nblk + w) * nb + c];
    __syncthreads();
    if (j >= r0 && j < r1)
      for (int c = tid; c < nb; c += 256) P[(j - r0) * nb + c] = urow[c];
    if (p != j && p >= r0 && p < r1)
      for (int c = tid; c < nb; c += 256) P[(p - r0) * nb + c] = rowj[par * nb + c];
    if (b == 0 && tid == 0) {
      ipiv[j] = (int32_t)p;
      if (urow[j].x == 0 && urow[j].y == 0 && *info == 0) *info = info_off + j + 1;
    }
    __syncthreads();
    const cplx<T> piv = urow[j];
    if (piv.x != 0 || piv.y != 0) {
      for (int64_t r = tid; r < nr; r += 256)
        if (r0 + r > j) P[r * nb + j] = cdiv<T>(P[r * nb + j], piv);
      __syncthreads();
      const int nc = nb - j - 1;
      for (int64_t idx = tid; idx < nr * nc; idx += 256) {
        const int64_t r = idx % nr;
        const int c = j + 1 + (int)(idx / nr);
        if (r0 + r > j) P[r * nb + c] = csub<T>(P[r * nb + c], cmul<T>(P[r * nb + j], urow[c]));
      }
    }
    __syncthreads();
  }
  for (int64_t idx = tid; idx < nr * nb; idx += 256) {
    const int64_t r = idx % nr, c = idx / nr;
    A[c * lda + r0 + r] = P[r * nb + c];
  }
}

struct DevInfo {
  int ncu = 0;
  int lds_max = 0;
};
static DevInfo dev_info() {
  int d;
  MNP_HIP(hipGetDevice(&d));
  DevInfo di;
  MNP_HIP(hipDeviceGetAttribute(&di.ncu, hipDeviceAttributeMultiprocessorCount, d));
  MNP_HIP(hipDeviceGetAttribute(&di.lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, d));
  return di;
}

template <class T>
static void getf2_leaf(int64_t m, int nb, MatView<T> A, int32_t* ipiv, int* info, int info_off,
                       Workspace& ws) {
  static DevInfo di = dev_info();
  const size_t esz = sizeof(cplx<T>);
  const int64_t lds_budget = 64 * 1024;
  int64_t rows_max = (lds_budget - (int64_t)(nb * esz)) / (int64_t)(nb * esz);
  MNP_CHECK(rows_max >= 1, MNP_ERR_INTERNAL, "getf2: panel too wide");
  int64_t P = std::max<int64_t>(cdiv_up(m, rows_max), std::min<int64_t>(di.ncu, cdiv_up(m, 32)));
  int64_t chunk = cdiv_up(m, P);
  P = cdiv_up(m, chunk);
  const size_t lds = (size_t)(chunk * nb + nb) * esz;
  int occ = 0;
  MNP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_getf2_coop<T>, 256, lds));
  MNP_CHECK((int64_t)occ * di.ncu >= P, MNP_ERR_INTERNAL,
            "getf2: panel of " + std::to_string(m) + " rows does not fit co-resident workgroups");
  // scratch: cval[2P], cidx[2P], crow[2P*nb], rowj[2*nb]
  const size_t off_idx = 2 * P * sizeof(T);
  const size_t off_row = ((off_idx + 2 * P * sizeof(int32_t)) + 15) / 16 * 16;
  const size_t off_rj = off_row + 2 * P * nb * esz;
  char* base = (char*)ws.get(4, off_rj + 2 * nb * esz);
  T* cval = (T*)base;
  int32_t* cidx = (int32_t*)(base + off_idx);
  cplx<T>* crow = (cplx<T>*)(base + off_row);
  cplx<T>* rowj = (cplx<T>*)(base + off_rj);
  cplx<T>* Ap = A.p;
  int64_t lda = A.ld;
  void* args[] = {&m, &nb, &Ap, &lda, &ipiv, &info, &info_off, &chunk, &cval, &cidx, &crow, &rowj};
  TimerScope ts(MNP_TIMER_PANEL, 0.0);
  MNP_HIP(hipLaunchCooperativeKernel((const void*)k_getf2_coop<T>, dim3((unsigned)P), dim3(256), args,
                                     lds, stream()));
}

// ------------------------------------------------------------ permutations
// Applies the swaps s <-> ipiv[s] (s in [k0, k1), rows in [rlo, rhi)) to an
// identity permutation held in LDS by one thread, then lists the rows whose
// source changed: dst[t] = r, src[t] = sigma(r).  Optionally writes sigma.
__global__ void __launch_bounds__(256) k_perm(int64_t k0, int64_t k1, int64_t rlo, int64_t rhi,
                                              const int32_t* __restrict__ ipiv, int32_t* dst,
                                              int32_t* src, int32_t* count, int32_t* sigma) {
  extern __shared__ int32_t perm[];
  const int64_t len = rhi - rlo;
  for (int64_t r = threadIdx.x; r < len; r += blockDim.x) perm[r] = (int32_t)(rlo + r);
  if (threadIdx.x == 0) *count = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int64_t s = k0; s < k1; ++s) {
      const int64_t p = ipiv[s];
      const int32_t t = perm[s - rlo];
      perm[s - rlo] = perm[p - rlo];
      perm[p - rlo] = t;
    }
  }
  __syncthreads();
  for (int64_t r = threadIdx.x; r < len; r += blockDim.x) {
    const int32_t v = perm[r];
    if (sigma) sigma[r] = v;
    if (v != rlo + r) {
      const int t = atomicAdd(count, 1);
      dst[t] = (int32_t)(rlo + r);
      src[t] = v;
    }
  }
}

template <class T>
__global__ void k_rows_gather(int64_t maxcnt, const int32_t* count, const int32_t* src,
                              const cplx<T>* A, int64_t lda, int64_t col0, cplx<T>* tmp) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t c = blockIdx.y;
  if (t >= *count) return;
  tmp[c * maxcnt + t] = A[(col0 + c) * lda + src[t]];
}
template <class T>
__global__ void k_rows_scatter(int64_t maxcnt, const int32_t* count, const int32_t* dst, cplx<T>* A,
                               int64_t lda, int64_t col0, const cplx<T>* tmp) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t c = blockIdx.y;
  if (t >= *count) return;
  A[(col0 + c) * lda + dst[t]] = tmp[c * maxcnt + t];
}

static void launch_perm(int64_t k0, int64_t k1, int64_t rlo, int64_t rhi, const int32_t* ipiv,
                        int32_t* dst, int32_t* src, int32_t* count, int32_t* sigma) {
  const size_t lds = sizeof(int32_t) * (size_t)(rhi - rlo);
  static std::once_flag attr;
  std::call_once(attr, [] {
    (void)hipFuncSetAttribute((const void*)k_perm, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  });
  MNP_CHECK(lds <= 160 * 1024, MNP_ERR_ARG, "permutation longer than 40960 rows");
  hipLaunchKernelGGL(k_perm, dim3(1), dim3(256), lds, stream(), k0, k1, rlo, rhi, ipiv, dst, src,
                     count, sigma);
  MNP_HIP(hipGetLastError());
}

// Row interchanges s <-> ipiv[s], s in [k0, k1), on columns [c0, c1) of A,
// all touched rows within [k0, m).
template <class T>
static void laswp(MatView<T> A, int64_t m, int64_t c0, int64_t c1, int64_t k0, int64_t k1,
                  const int32_t* ipiv, Workspace& ws) {
  if (c1 <= c0 || k1 <= k0) return;
  const int64_t maxcnt = std::min<int64_t>(2 * (k1 - k0), m - k0);
  int32_t* ib = (int32_t*)ws.get(0, sizeof(int32_t) * (size_t)(2 * maxcnt + 4));
  int32_t* dst = ib;
  int32_t* src = ib + maxcnt;
  int32_t* cnt = ib + 2 * maxcnt;
  launch_perm(k0, k1, k0, m, ipiv, dst, src, cnt, nullptr);
  const int64_t ncols = c1 - c0;
  const size_t budget = (size_t)256 << 20;
  int64_t cchunk = std::max<int64_t>(1, (int64_t)(budget / (sizeof(cplx<T>) * (size_t)maxcnt)));
  cchunk = std::min<int64_t>({cchunk, ncols, 65535});
  cplx<T>* tmp = (cplx<T>*)ws.get(1, sizeof(cplx<T>) * (size_t)(maxcnt * cchunk));
  for (int64_t c = c0; c < c1; c += cchunk) {
    const int64_t nc = std::min(cchunk, c1 - c);
    dim3 grid((unsigned)cdiv_up(maxcnt, 256), (unsigned)nc);
    hipLaunchKernelGGL(k_rows_gather<T>, grid, dim3(256), 0, stream(), maxcnt, cnt, src, A.p, A.ld, c,
                       tmp);
    hipLaunchKernelGGL(k_rows_scatter<T>, grid, dim3(256), 0, stream(), maxcnt, cnt, dst, A.p, A.ld, c,
                       tmp);
  }
  MNP_HIP(hipGetLastError());
}

void perm_from_pivots(int64_t n, const int32_t* ipiv, int32_t* sigma, Workspace& ws) {
  int32_t* ib = (int32_t*)ws.get(0, sizeof(int32_t) * (size_t)(2 * n + 4));
  launch_perm(0, n, 0, n, ipiv, ib, ib + n, ib + 2 * n, sigma);
}

__global__ void k_add_offset(int64_t n, int32_t* ipiv, int32_t off) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) ipiv[i] += off;
}

// ----------------------------------------------------- small helpers
template <class T>
__global__ void k_copy(int64_t m, int64_t n, const cplx<T>* A, int64_t lda, cplx<T>* B, int64_t ldb) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (i < m && j < n) B[j * ldb + i] = A[j * lda + i];
}
template <class Ts, class Td>
__global__ void k_convert(int64_t m, int64_t n, const cplx<Ts>* A, int64_t lda, cplx<Td>* B,
                          int64_t ldb) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (i < m && j < n) {
    const cplx<Ts> v = A[j * lda + i];
    B[j * ldb + i] = mk<Td>((Td)v.x, (Td)v.y);
  }
}
static dim3 grid2(int64_t m, int64_t n) {
  return dim3((unsigned)cdiv_up(m, 256), (unsigned)std::min<int64_t>(n, 65535),
              (unsigned)cdiv_up(n, 65535));
}
template <class T>
void copy_mat(int64_t m, int64_t n, MatView<T> A, MatView<T> B) {
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL(k_copy<T>, grid2(m, n), dim3(256), 0, stream(), m, n, A.p, A.ld, B.p, B.ld);
  MNP_HIP(hipGetLastError());
}
template <class Ts, class Td>
void convert_mat(int64_t m, int64_t n, MatView<Ts> A, MatView<Td> B) {
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL((k_convert<Ts, Td>), grid2(m, n), dim3(256), 0, stream(), m, n, A.p, A.ld, B.p,
                     B.ld);
  MNP_HIP(hipGetLastError());
}

// ------------------------------------------- diagonal-block inverses
// The triangular solves multiply by the inverses of the 32 x 32 diagonal
// blocks of L (unit lower) and U.  Block b covers rows/columns
// [32 b, 32 b + 32) and is stored column-major (ld 32) at inv + 1024 b;
// a partial last block is padded with the identity.
constexpr int TBN = 256;  // diagonal blocks of the blocked triangular solves

// One workgroup per block: threads 0..31 build column j of L^{-1},
// threads 32..63 column j of U^{-1}.
template <class T>
__global__ void __launch_bounds__(64) k_trtri_blocks(int64_t n, const cplx<T>* LU, int64_t ld,
                                                     cplx<T>* Linv, cplx<T>* Uinv) {
  __shared__ cplx<T> sA[TB][TB + 1];
  __shared__ cplx<T> sL[TB][TB + 1];
  __shared__ cplx<T> sU[TB][TB + 1];
  const int64_t b = blockIdx.x;
  const int nb = (int)min((int64_t)TB, n - b * TB);
  const cplx<T>* A = LU + (b * TB) * ld + b * TB;
  for (int idx = threadIdx.x; idx < TB * TB; idx += 64) {
    const int r = idx % TB, c = idx / TB;
    sA[r][c] = (r < nb && c < nb) ? A[c * ld + r] : (r == c ? mk<T>(1, 0) : czero<T>());
  }
  __syncthreads();
  const int j = threadIdx.x & 31;
  if (threadIdx.x < 32) {
    for (int i = 0; i < TB; ++i) {
      cplx<T> s = i == j ? mk<T>(1, 0) : czero<T>();
      if (i > j && i < nb)
        for (int k = j; k < i; ++k) s = csub<T>(s, cmul<T>(sA[i][k], sL[k][j]));
      sL[i][j] = i < j ? czero<T>() : s;
    }
  } else {
    for (int i = TB - 1; i >= 0; --i) {
      if (i > j) {
        sU[i][j] = czero<T>();
        continue;
      }
      cplx<T> s = i == j ? mk<T>(1, 0) : czero<T>();
      for (int k = i + 1; k <= j; ++k) s = csub<T>(s, cmul<T>(sA[i][k], sU[k][j]));
      sU[i][j] = cdiv<T>(s, sA[i][i]);
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < TB * TB; idx += 64) {
    const int r = idx % TB, c = idx / TB;
    if (Linv) Linv[b * TB2 + idx] = sL[r][c];
    if (Uinv) Uinv[b * TB2 + idx] = sU[r][c];
  }
}

template <class T>
void block_inverses(int64_t n, MatView<T> LU, cplx<T>* Linv, cplx<T>* Uinv) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_trtri_blocks<T>, dim3((unsigned)cdiv_up(n, TB)), dim3(64), 0, stream(), n, LU.p,
                     LU.ld, Linv, Uinv);
  MNP_HIP(hipGetLastError());
}

static const double2 ONE = {1.0, 0.0}, MONE = {-1.0, 0.0}, ZERO = {0.0, 0.0};

// ---------------------------------------- 256-block inverses of L and U
// The inverses of the 256 x 256 diagonal blocks of a factor are built from
// its 32 x 32 ones by three doublings (32 -> 64 -> 128 -> 256), each a
// batch over all blocks of that size:
//   lower: [[A^-1, 0], [-B^-1 L21 A^-1, B^-1]],  upper: [[A^-1, -A^-1 U12 B^-1], [0, B^-1]].
// k_inv_assemble copies the two diagonal quadrants and zeroes the empty
// one; two batched GEMMs fill the coupling quadrant.
template <class T>
__global__ void k_inv_assemble(int s, const cplx<T>* in, cplx<T>* out, int lower) {
  const int64_t i = blockIdx.y;  // block of size 2s
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t s2 = 2 * (int64_t)s;
  if (e >= s2 * s2) return;
  const int r = (int)(e % s2), c = (int)(e / s2);
  cplx<T>* o = out + i * s2 * s2;
  if (r < s && c < s)
    o[e] = in[(2 * i) * s * s + (int64_t)c * s + r];
  else if (r >= s && c >= s)
    o[e] = in[(2 * i + 1) * s * s + (int64_t)(c - s) * s + (r - s)];
  else if ((lower && r < s) || (!lower && r >= s))
    o[e] = czero<T>();
}

template <class T>
void block_inverses_nb(int64_t n, MatView<T> LU, const cplx<T>* Linv, const cplx<T>* Uinv, cplx<T>* LinvNB,
                       cplx<T>* UinvNB, Workspace& ws) {
  const int64_t nb = n / TBN;
  if (nb == 0) return;
  const int64_t full = nb * TBN;
  // ping-pong buffers (full x TBN entries each) and the coupling temporaries
  cplx<T>* buf[2] = {(cplx<T>*)ws.get(3, sizeof(cplx<T>) * (size_t)(full * TBN)),
                     (cplx<T>*)ws.get(6, sizeof(cplx<T>) * (size_t)(full * TBN))};
  cplx<T>* tmp = (cplx<T>*)ws.get(7, sizeof(cplx<T>) * (size_t)(full * TBN / 2));
  for (int lower = 1; lower >= 0; --lower) {
    const cplx<T>* in = lower ? Linv : Uinv;
    for (int64_t sz = TB; sz < TBN; sz *= 2) {
      const int64_t s2 = 2 * sz, nbl = full / s2, ss = sz * sz;
      cplx<T>* out = (s2 == TBN) ? (lower ? LinvNB : UinvNB) : buf[(sz / TB) % 2];
      hipLaunchKernelGGL(k_inv_assemble<T>, dim3((unsigned)cdiv_up(s2 * s2, 256), (unsigned)nbl), dim3(256), 0,
                         stream(), (int)sz, in, out, lower);
      MNP_HIP(hipGetLastError());
      const int64_t sdiag = s2 * (LU.ld + 1);  // stride between consecutive 2s diagonal blocks of LU
      if (lower) {
        // tmp_i = L21_i A_i^-1 ;  X_ll = -B_i^-1 tmp_i
        gemm_batched<T>(sz, sz, sz, ONE, MatView<T>{LU.at(sz, 0), LU.ld}, sdiag,
                        MatView<T>{const_cast<cplx<T>*>(in), sz}, 2 * ss, ZERO, MatView<T>{tmp, sz}, ss, nbl);
        gemm_batched<T>(sz, sz, sz, MONE, MatView<T>{const_cast<cplx<T>*>(in) + ss, sz}, 2 * ss,
                        MatView<T>{tmp, sz}, ss, ZERO, MatView<T>{out + sz, s2}, s2 * s2, nbl);
      } else {
        // tmp_i = U12_i B_i^-1 ;  X_ur = -A_i^-1 tmp_i
        gemm_batched<T>(sz, sz, sz, ONE, MatView<T>{LU.at(0, sz), LU.ld}, sdiag,
                        MatView<T>{const_cast<cplx<T>*>(in) + ss, sz}, 2 * ss, ZERO, MatView<T>{tmp, sz}, ss,
                        nbl);
        gemm_batched<T>(sz, sz, sz, MONE, MatView<T>{const_cast<cplx<T>*>(in), sz}, 2 * ss,
                        MatView<T>{tmp, sz}, ss, ZERO, MatView<T>{out + sz * s2, s2}, s2 * s2, nbl);
      }
      in = out;
    }
  }
}

static int64_t split(int64_t n, int64_t leaf) {
  int64_t n1 = (n / 2) / leaf * leaf;
  return n1 < leaf ? leaf : n1;
}

// ----------------------------------------- fused diagonal-block solve
// B <- L^{-1} B for one diagonal block of a unit lower factor (nb <= 256
// rows; L at the block's origin, below the diagonal; Linv the inverses of its
// 32 x 32 diagonal blocks), in one launch: workgroup = CB columns of B, held
// in LDS for the whole solve.  For each 32-row block i in turn:
//   x_i = Linv_i b_i,  then  b_j -= L_ji x_i for the rows below,
// as 16 x 16 MFMA tiles (complex products as four real ones) spread over the
// four waves, with a barrier after each stage.  This replaces the recursion's
// 32-row leaves and bottom updates (about 15 dependent launches per 256 rows).
// Rows past nb are zero in LDS and never stored; the padded part of the last
// inverse block multiplies them only.
template <class T, int CB>
__global__ void __launch_bounds__(256) k_trsm_llu_blk(int nb, int64_t nrhs, const cplx<T>* L, int64_t ldl,
                                                      const cplx<T>* Linv, cplx<T>* B, int64_t ldb) {
  using acc_t = typename MF<T>::acc_t;
  constexpr int NCT = CB / 16;  // column tiles
  __shared__ T sr[TBN][CB + 1], si[TBN][CB + 1];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int l16 = lane & 15, lk = lane >> 4;
  const int64_t c0 = (int64_t)blockIdx.x * CB;
  for (int idx = tid; idx < TBN * CB; idx += 256) {
    const int r = idx % TBN, c = idx / TBN;
    cplx<T> v = czero<T>();
    if (r < nb && c0 + c < nrhs) v = B[(c0 + c) * ldb + r];
    sr[r][c] = v.x;
    si[r][c] = v.y;
  }
  __syncthreads();
  // A operand of one 16-row tile (rows from arow, the 32 columns from acol of
  // L or Linv); every stage issues all of its loads before its first MFMA
  auto load_a = [&](const cplx<T>* A, int64_t lda, int arow, int acol, int rows_valid, cplx<T> (&a)[TB / 4]) {
    const int row = min(arow + l16, rows_valid - 1);  // clamped: valid address, zeroed below
#pragma unroll
    for (int s = 0; s < TB / 4; ++s) a[s] = A[(int64_t)(acol + 4 * s + lk) * lda + row];
    if (arow + l16 >= rows_valid)
#pragma unroll
      for (int s = 0; s < TB / 4; ++s) a[s] = czero<T>();
  };
  // (cr, ci) += a x_i(16 columns from 16 ct), x_i at LDS rows xrow ..
  auto mma_tile = [&](const cplx<T> (&a)[TB / 4], int xrow, int ct, acc_t& cr, acc_t& ci) {
#pragma unroll
    for (int s = 0; s < TB / 4; ++s) {
      const int k = xrow + 4 * s + lk, col = 16 * ct + l16;
      const T br = sr[k][col], bi = si[k][col];
      cr = MF<T>::mma(a[s].x, br, cr);
      cr = MF<T>::mma(-a[s].y, bi, cr);
      ci = MF<T>::mma(a[s].x, bi, ci);
      ci = MF<T>::mma(a[s].y, br, ci);
    }
  };
  constexpr int MAXR = (TBN - TB) / 16 / 4;  // row tiles below a block per wave (at most 14 over 4 waves)
  const int nblk = (nb + TB - 1) / TB;
  for (int i = 0; i < nblk; ++i) {
    const int rbeg = TB * (i + 1), nrt = rbeg < nb ? (nb - rbeg + 15) / 16 : 0;
    // loads: this wave's diagonal tile (x_i = Linv_i b_i, 2 x NCT tiles, one
    // per wave) and its row tiles below (rt = wid + 4 q)
    const bool dw = wid < 2 * NCT;
    const int drt = wid % 2, dct = wid / 2;
    cplx<T> ad[TB / 4], au[MAXR + 1][TB / 4];
    if (dw) load_a(Linv + (int64_t)i * TB2, TB, 16 * drt, 0, TB, ad);
#pragma unroll
    for (int q = 0; q <= MAXR; ++q)
      if (wid + 4 * q < nrt) load_a(L, ldl, rbeg + 16 * (wid + 4 * q), TB * i, nb, au[q]);
    acc_t xr = acc_t{0, 0, 0, 0}, xi = acc_t{0, 0, 0, 0};
    if (dw) mma_tile(ad, TB * i, dct, xr, xi);
    __syncthreads();  // b_i is read by every diagonal tile before it is replaced
    if (dw)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = TB * i + 16 * drt + MF<T>::row(lane, r);
        sr[row][16 * dct + l16] = xr[r];
        si[row][16 * dct + l16] = xi[r];
      }
    __syncthreads();
    // b_j -= L_ji x_i for the rows below the block
    if (nrt) {
#pragma unroll
      for (int q = 0; q <= MAXR; ++q) {
        const int rt = wid + 4 * q;
        if (rt < nrt)
#pragma unroll
          for (int ct = 0; ct < NCT; ++ct) {
            acc_t ur = acc_t{0, 0, 0, 0}, ui = acc_t{0, 0, 0, 0};
            mma_tile(au[q], TB * i, ct, ur, ui);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = rbeg + 16 * rt + MF<T>::row(lane, r);
              sr[row][16 * ct + l16] -= ur[r];
              si[row][16 * ct + l16] -= ui[r];
            }
          }
      }
      __syncthreads();
    }
  }
  for (int idx = tid; idx < TBN * CB; idx += 256) {
    const int r = idx % TBN, c = idx / TBN;
    if (r < nb && c0 + c < nrhs) B[(c0 + c) * ldb + r] = mk<T>(sr[r][c], si[r][c]);
  }
}

template <class T>
static void trsm_llu_blk(int64_t nb, int64_t nrhs, MatView<T> LU, const cplx<T>* Linv, MatView<T> B) {
  constexpr int CB = sizeof(T) == 4 ? 32 : 16;
  hipLaunchKernelGGL((k_trsm_llu_blk<T, CB>), dim3((unsigned)cdiv_up(nrhs, CB)), dim3(256), 0, stream(), (int)nb,
                     nrhs, LU.p, LU.ld, Linv, B.p, B.ld);
  MNP_HIP(hipGetLastError());
}

// Recursive triangular solves.  Every split point is a multiple of 32 from
// the factor's origin, so each leaf is one diagonal block and is applied as
// an in-place product with its inverse (the GEMM tile covers the whole
// 32-row or 32-column leaf, so each workgroup reads its operand block before
// it overwrites it).
// B <- L^{-1} B
template <class T>
void trsm_llu(int64_t n, int64_t nrhs, MatView<T> LU, const cplx<T>* Linv, MatView<T> B) {
  if (n <= 0 || nrhs <= 0) return;
  if (n <= TB) {
    gemm<T>(n, nrhs, n, ONE, MatView<T>{const_cast<cplx<T>*>(Linv), TB}, B, ZERO, B);
    return;
  }
  if (n <= TBN) {  // one diagonal block of up to 256 rows: one fused launch
    trsm_llu_blk<T>(n, nrhs, LU, Linv, B);
    return;
  }
  // above one block, split at multiples of 256 so that the leaves are whole blocks
  const int64_t n1 = std::max<int64_t>(TBN, (n / 2) / TBN * TBN), n2 = n - n1;
  trsm_llu<T>(n1, nrhs, LU, Linv, B);
  gemm<T>(n2, nrhs, n1, MONE, MatView<T>{LU.at(n1, 0), LU.ld}, B, ONE, MatView<T>{B.at(n1, 0), B.ld});
  trsm_llu<T>(n2, nrhs, MatView<T>{LU.at(n1, n1), LU.ld}, Linv + (n1 / TB) * TB2,
              MatView<T>{B.at(n1, 0), B.ld});
}
// B <- U^{-1} B
template <class T>
void trsm_lun(int64_t n, int64_t nrhs, MatView<T> LU, const cplx<T>* Uinv, MatView<T> B) {
  if (n <= 0 || nrhs <= 0) return;
  if (n <= TB) {
    gemm<T>(n, nrhs, n, ONE, MatView<T>{const_cast<cplx<T>*>(Uinv), TB}, B, ZERO, B);
    return;
  }
  const int64_t n1 = split(n, TB), n2 = n - n1;
  trsm_lun<T>(n2, nrhs, MatView<T>{LU.at(n1, n1), LU.ld}, Uinv + (n1 / TB) * TB2,
              MatView<T>{B.at(n1, 0), B.ld});
  gemm<T>(n1, nrhs, n2, MONE, MatView<T>{LU.at(0, n1), LU.ld}, MatView<T>{B.at(n1, 0), B.ld}, ONE, B);
  trsm_lun<T>(n1, nrhs, LU, Uinv, B);
}
// B <- B U^{-1}   (B m x n)
template <class T>
void trsm_run(int64_t m, int64_t n, MatView<T> LU, const cplx<T>* Uinv, MatView<T> B) {
  if (n <= 0 || m <= 0) return;
  if (n <= TB) {
    gemm<T>(m, n, n, ONE, B, MatView<T>{const_cast<cplx<T>*>(Uinv), TB}, ZERO, B);
    return;
  }
  const int64_t n1 = split(n, TB), n2 = n - n1;
  trsm_run<T>(m, n1, LU, Uinv, B);
  gemm<T>(m, n2, n1, MONE, B, MatView<T>{LU.at(0, n1), LU.ld}, ONE, MatView<T>{B.at(0, n1), B.ld});
  trsm_run<T>(m, n2, MatView<T>{LU.at(n1, n1), LU.ld}, Uinv + (n1 / TB) * TB2,
              MatView<T>{B.at(0, n1), B.ld});
}
// B <- B L^{-1}   (B m x n, L unit lower)
template <class T>
void trsm_rlu(int64_t m, int64_t n, MatView<T> LU, const cplx<T>* Linv, MatView<T> B) {
  if (n <= 0 || m <= 0) return;
  if (n <= TB) {
    gemm<T>(m, n, n, ONE, B, MatView<T>{const_cast<cplx<T>*>(Linv), TB}, ZERO, B);
    return;
  }
  const int64_t n1 = split(n, TB), n2 = n - n1;
  trsm_rlu<T>(m, n2, MatView<T>{LU.at(n1, n1), LU.ld}, Linv + (n1 / TB) * TB2,
              MatView<T>{B.at(0, n1), B.ld});
  gemm<T>(m, n1, n2, MONE, MatView<T>{B.at(0, n1), B.ld}, MatView<T>{LU.at(n1, 0), LU.ld}, ONE, B);
  trsm_rlu<T>(m, n1, LU, Linv, B);
}

// Blocked triangular solves with the 256-block inverses: per block one
// product with the block's inverse (into W, copied back) and one update of
// the remaining rows or columns, so every product has an inner dimension of
// 256.  A trailing partial block uses the recursive solves.
// B <- L^{-1} B
template <class T>
void trsm_llu_nb(int64_t n, int64_t nrhs, const LU<T>& f, MatView<T> B, Workspace& ws) {
  const int64_t nb = n / TBN;
  cplx<T>* W = (cplx<T>*)ws.get(2, sizeof(cplx<T>) * (size_t)(TBN * nrhs));
  const MatView<T> Wv{W, TBN};
  for (int64_t i = 0; i < nb; ++i) {
    const int64_t r0 = i * TBN;
    gemm<T>(TBN, nrhs, TBN, ONE, MatView<T>{f.LinvNB + i * TBN * TBN, TBN}, MatView<T>{B.at(r0, 0), B.ld}, ZERO,
            Wv);
    copy_mat<T>(TBN, nrhs, Wv, MatView<T>{B.at(r0, 0), B.ld});
    if (r0 + TBN < n)
      gemm<T>(n - r0 - TBN, nrhs, TBN, MONE, MatView<T>{f.A.at(r0 + TBN, r0), f.A.ld}, MatView<T>{B.at(r0, 0), B.ld},
              ONE, MatView<T>{B.at(r0 + TBN, 0), B.ld});
  }
  const int64_t t0 = nb * TBN;
  if (t0 < n)
    trsm_llu<T>(n - t0, nrhs, MatView<T>{f.A.at(t0, t0), f.A.ld}, f.Linv + (t0 / TB) * TB2,
                MatView<T>{B.at(t0, 0), B.ld});
}
// B <- U^{-1} B
template <class T>
void trsm_lun_nb(int64_t n, int64_t nrhs, const LU<T>& f, MatView<T> B, Workspace& ws) {
  const int64_t nb = n / TBN, t0 = nb * TBN;
  if (t0 < n) {
    trsm_lun<T>(n - t0, nrhs, MatView<T>{f.A.at(t0, t0), f.A.ld}, f.Uinv + (t0 / TB) * TB2,
                MatView<T>{B.at(t0, 0), B.ld});
    if (t0 > 0)
      gemm<T>(t0, nrhs, n - t0, MONE, MatView<T>{f.A.at(0, t0), f.A.ld}, MatView<T>{B.at(t0, 0), B.ld}, ONE, B);
  }
  cplx<T>* W = (cplx<T>*)ws.get(2, sizeof(cplx<T>) * (size_t)(TBN * nrhs));
  const MatView<T> Wv{W, TBN};
  for (int64_t i = nb - 1; i >= 0; --i) {
    const int64_t r0 = i * TBN;
    gemm<T>(TBN, nrhs, TBN, ONE, MatView<T>{f.UinvNB + i * TBN * TBN, TBN}, MatView<T>{B.at(r0, 0), B.ld}, ZERO,
            Wv);
    copy_mat<T>(TBN, nrhs, Wv, MatView<T>{B.at(r0, 0), B.ld});
    if (r0 > 0)
      gemm<T>(r0, nrhs, TBN, MONE, MatView<T>{f.A.at(0, r0), f.A.ld}, MatView<T>{B.at(r0, 0), B.ld}, ONE, B);
  }
}
// B <- B U^{-1}   (B m x n)
template <class T>
void trsm_run_nb(int64_t m, int64_t n, const LU<T>& f, MatView<T> B, Workspace& ws) {
  const int64_t nb = n / TBN;
  cplx<T>* W = (cplx<T>*)ws.get(2, sizeof(cplx<T>) * (size_t)(m * TBN));
  const MatView<T> Wv{W, m};
  for (int64_t j = 0; j < nb; ++j) {
    const int64_t c0 = j * TBN;
    gemm<T>(m, TBN, TBN, ONE, MatView<T>{B.at(0, c0), B.ld}, MatView<T>{f.UinvNB + j * TBN * TBN, TBN}, ZERO, Wv);
    copy_mat<T>(m, TBN, Wv, MatView<T>{B.at(0, c0), B.ld});
    if (c0 + TBN < n)
      gemm<T>(m, n - c0 - TBN, TBN, MONE, MatView<T>{B.at(0, c0), B.ld}, MatView<T>{f.A.at(c0, c0 + TBN), f.A.ld},
              ONE, MatView<T>{B.at(0, c0 + TBN), B.ld});
  }
  const int64_t t0 = nb * TBN;
  if (t0 < n)
    trsm_run<T>(m, n - t0, MatView<T>{f.A.at(t0, t0), f.A.ld}, f.Uinv + (t0 / TB) * TB2,
                MatView<T>{B.at(0, t0), B.ld});
}
// B <- B L^{-1}   (B m x n, L unit lower)
template <class T>
void trsm_rlu_nb(int64_t m, int64_t n, const LU<T>& f, MatView<T> B, Workspace& ws) {
  const int64_t nb = n / TBN, t0 = nb * TBN;
  if (t0 < n) {
    trsm_rlu<T>(m, n - t0, MatView<T>{f.A.at(t0, t0), f.A.ld}, f.Linv + (t0 / TB) * TB2,
                MatView<T>{B.at(0, t0), B.ld});
    if (t0 > 0)
      gemm<T>(m, t0, n - t0, MONE, MatView<T>{B.at(0, t0), B.ld}, MatView<T>{f.A.at(t0, 0), f.A.ld}, ONE, B);
  }
  cplx<T>* W = (cplx<T>*)ws.get(2, sizeof(cplx<T>) * (size_t)(m * TBN));
  const MatView<T> Wv{W, m};
  for (int64_t j = nb - 1; j >= 0; --j) {
    const int64_t c0 = j * TBN;
    gemm<T>(m, TBN, TBN, ONE, MatView<T>{B.at(0, c0), B.ld}, MatView<T>{f.LinvNB + j * TBN * TBN, TBN}, ZERO, Wv);
    copy_mat<T>(m, TBN, Wv, MatView<T>{B.at(0, c0), B.ld});
    if (c0 > 0)
      gemm<T>(m, c0, TBN, MONE, MatView<T>{B.at(0, c0), B.ld}, MatView<T>{f.A.at(c0, 0), f.A.ld}, ONE, B);
  }
}

// Recursive solves whose leaves are the 256-row (or 256-column) diagonal
// blocks, applied through their stored inverses (one product into W, copied
// back); splits fall on multiples of 256 from the factor's origin, and a
// trailing partial block uses the 32-leaf recursion.  For many right-hand
// sides this keeps the large inner dimensions of the recursion and removes
// its three lowest levels.
template <class T>
void trsm_llu_h(int64_t n, int64_t nrhs, MatView<T> LU, const cplx<T>* Linv, const cplx<T>* LinvNB, MatView<T> B,
                Workspace& ws) {
  if (n <= 0 || nrhs <= 0) return;
  if (n < TBN) return trsm_llu<T>(n, nrhs, LU, Linv, B);
  if (n == TBN) {
    cplx<T>* W = (cplx<T>*)ws.get(2, sizeof(cplx<T>) * (size_t)(TBN * nrhs));
    gemm<T>(TBN, nrhs, TBN, ONE, MatView<T>{const_cast<cplx<T>*>(LinvNB), TBN}, B, ZERO, MatView<T>{W, TBN});
    copy_mat<T>(TBN, nrhs, MatView<T>{W, TBN}, B);
    return;
  }
  const int64_t n1 = std::max<int64_t>(TBN, (n / 2) / TBN * TBN), n2 = n - n1;
  trsm_llu_h<T>(n1, nrhs, LU, Linv, LinvNB, B, ws);
  gemm<T>(n2, nrhs, n1, MONE, MatView<T>{LU.at(n1, 0), LU.ld}, B, ONE, MatView<T>{B.at(n1, 0), B.ld});
  trsm_llu_h<T>(n2, nrhs, MatView<T>{LU.at(n1, n1), LU.ld}, Linv + (n1 / TB) * TB2, LinvNB + (n1 / TBN) * TBN * TBN,
                MatView<T>{B.at(n1, 0), B.ld}, ws);
}
template <class T>
void trsm_lun_h(int64_t n, int64_t nrhs, MatView<T> LU, const cplx<T>* Uinv, const cplx<T>* UinvNB, MatView<T> B,
                Workspace& ws) {
  if (n <= 0 || nrhs <= 0) return;
  if (n < TBN) return trsm_lun<T>(n, nrhs, LU, Uinv, B);
  if (n == TBN) {
    cplx<T>* W = (cplx<T>*)ws.get(2, sizeof(cplx<T>) * (size_t)(TBN * nrhs));
    gemm<T>(TBN, nrhs, TBN, ONE, MatView<T>{const_cast<cplx<T>*>(UinvNB), TBN}, B, ZERO, MatView<T>{W, TBN});
    copy_mat<T>(TBN, nrhs, MatView<T>{W, TBN}, B);
    return;
  }
  const int64_t n1 = std::max<int64_t>(TBN, (n / 2) / TBN * TBN), n2 = n - n1;
  trsm_lun_h<T>(n2, nrhs, MatView<T>{LU.at(n1, n1), LU.ld}, Uinv + (n1 / TB) * TB2, UinvNB + (n1 / TBN) * TBN * TBN,
                MatView<T>{B.at(n1, 0), B.ld}, ws);
  gemm<T>(n1, nrhs, n2, MONE, MatView<T>{LU.at(0, n1), LU.ld}, MatView<T>{B.at(n1, 0), B.ld}, ONE, B);
  trsm_lun_h<T>(n1, nrhs, LU, Uinv, UinvNB, B, ws);
}
template <class T>
void trsm_run_h(int64_t m, int64_t n, MatView<T> LU, const cplx<T>* Uinv, const cplx<T>* UinvNB, MatView<T> B,
                Workspace& ws) {
  if (n <= 0 || m <= 0) return;
  if (n < TBN) return trsm_run<T>(m, n, LU, Uinv, B);
  if (n == TBN) {
    cplx<T>* W = (cplx<T>*)ws.get(2, sizeof(cplx<T>) * (size_t)(m * TBN));
    gemm<T>(m, TBN, TBN, ONE, B, MatView<T>{const_cast<cplx<T>*>(UinvNB), TBN}, ZERO, MatView<T>{W, m});
    copy_mat<T>(m, TBN, MatView<T>{W, m}, B);
    return;
  }
  const int64_t n1 = std::max<int64_t>(TBN, (n / 2) / TBN * TBN), n2 = n - n1;
  trsm_run_h<T>(m, n1, LU, Uinv, UinvNB, B, ws);
  gemm<T>(m, n2, n1, MONE, B, MatView<T>{LU.at(0, n1), LU.ld}, ONE, MatView<T>{B.at(0, n1), B.ld});
  trsm_run_h<T>(m, n2, MatView<T>{LU.at(n1, n1), LU.ld}, Uinv + (n1 / TB) * TB2, UinvNB + (n1 / TBN) * TBN * TBN,
                MatView<T>{B.at(0, n1), B.ld}, ws);
}
template <class T>
void trsm_rlu_h(int64_t m, int64_t n, MatView<T> LU, const cplx<T>* Linv, const cplx<T>* LinvNB, MatView<T> B,
                Workspace& ws) {
  if (n <= 0 || m <= 0) return;
  if (n < TBN) return trsm_rlu<T>(m, n, LU, Linv, B);
  if (n == TBN) {
    cplx<T>* W = (cplx<T>*)ws.get(2, sizeof(cplx<T>) * (size_t)(m * TBN));
    gemm<T>(m, TBN, TBN, ONE, B, MatView<T>{const_cast<cplx<T>*>(LinvNB), TBN}, ZERO, MatView<T>{W, m});
    copy_mat<T>(m, TBN, MatView<T>{W, m}, B);
    return;
  }
  const int64_t n1 = std::max<int64_t>(TBN, (n / 2) / TBN * TBN), n2 = n - n1;
  trsm_rlu_h<T>(m, n2, MatView<T>{LU.at(n1, n1), LU.ld}, Linv + (n1 / TB) * TB2, LinvNB + (n1 / TBN) * TBN * TBN,
                MatView<T>{B.at(0, n1), B.ld}, ws);
  gemm<T>(m, n1, n2, MONE, MatView<T>{B.at(0, n1), B.ld}, MatView<T>{LU.at(n1, 0), LU.ld}, ONE, B);
  trsm_rlu_h<T>(m, n1, LU, Linv, LinvNB, B, ws);
}

// ------------------------------------------------ unpivoted panel leaf
// LU without pivoting of the nb x nb diagonal block (nb <= 32) of a complex64
// factor, which also writes the block's L^{-1} and U^{-1}; the block is padded
// to 32 x 32 with the identity, which the elimination leaves unchanged.  Three
// waves:
// - wave 0 factors the block in registers, lane r holding row r; the pivot
//   row is broadcast with readlane at each (unrolled) step, so the
//   elimination needs no barriers.  Multipliers are row entries times the
//   reciprocal of the pivot, as LAPACK's getf2 scales them.
// - wave 1 then forms L^{-1} and wave 2 U^{-1}, concurrently, each lane one
//   row, by substitution with unit triangles: U = D U~ with U~ unit upper,
//   U^{-1} = U~^{-1} D^{-1}.
template <class T>
__global__ void __launch_bounds__(192) k_lu_nopiv_diag(int nb, cplx<T>* A, int64_t lda, int* info,
                                                       int info_off, cplx<T>* Linv, cplx<T>* Uinv) {
  lu_nopiv_diag_body<T>(nb, A, lda, info, info_off, Linv, Uinv);
}
template <class T>
__global__ void __launch_bounds__(256) k_lu_nopiv_diag_lds(int nb, cplx<T>* A, int64_t lda, int* info,
                                                       int info_off, cplx<T>* Linv, cplx<T>* Uinv) {
  lu_nopiv_diag_lds_body<T>(nb, A, lda, info, info_off, Linv, Uinv);
}

// Partial pivoting would keep the diagonal of column j iff no entry below
// it has a larger |Re|+|Im| at step j; in the unpivoted factors that entry
// is l_ij u_jj (ties go to the diagonal, the lower index), so one pass over
// the finished L and U decides it.  Thread (row i, chunk of 64 columns);
// sets *flag when some column would swap.
constexpr int VERIFY_COLS = 64;
template <class T>
__global__ void k_verify_nopiv(int64_t m, int64_t n, const cplx<T>* A, int64_t lda, int* flag) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j0 = (int64_t)blockIdx.y * VERIFY_COLS;
  if (i >= m) return;
  const int64_t j1 = min(min(n, i), j0 + VERIFY_COLS);
  bool bad = false;
  for (int64_t j = j0; j < j1; ++j) {
    const cplx<T> u = A[j * lda + j];
    const cplx<T> v = cmul<T>(A[j * lda + i], u);
    bad |= cabs1<T>(v) > cabs1<T>(u);
  }
  if (bad) atomicOr(flag, 1);
}

__global__ void k_iota(int64_t n, int32_t* ipiv, int32_t off) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) ipiv[i] = (int32_t)(off + i);
}

// ------------------------------------------------------------ recursive LU
struct LUState {
  int mode;
  int* info;  // device
  int* flag;  // device: partial pivoting would have swapped (mode PIVOT_NONE)
};

template <class T>
static void getrf_rec(int64_t m, int64_t n, MatView<T> A, int32_t* ipiv, int info_off,
                      cplx<T>* Linv, cplx<T>* Uinv, const LUState& st, Workspace& ws) {
  if (n <= TB) {
    if (st.mode == PIVOT_PARTIAL) {
      getf2_leaf<T>(m, (int)n, A, ipiv, st.info, info_off, ws);
      block_inverses<T>(n, A, Linv, Uinv);
    } else {
      {
        TimerScope ts(MNP_TIMER_PANEL, 0.0);
        if constexpr (sizeof(T) == 4)
          hipLaunchKernelGGL(k_lu_nopiv_diag<T>, dim3(1), dim3(192), 0, stream(), (int)n, A.p, A.ld, st.info,
                             info_off, Linv, Uinv);
        else
          hipLaunchKernelGGL(k_lu_nopiv_diag_lds<T>, dim3(1), dim3(256), 0, stream(), (int)n, A.p, A.ld,
                             st.info, info_off, Linv, Uinv);
      }
      if (m > n) {
        MatView<T> A21{A.at(n, 0), A.ld};
        gemm<T>(m - n, n, n, ONE, A21, MatView<T>{Uinv, TB}, ZERO, A21);  // L21 = A21 U11^{-1}
      }
      MNP_HIP(hipGetLastError());
    }
    return;
  }
  const int64_t n1 = split(n, TB), n2 = n - n1;
  const bool piv = st.mode == PIVOT_PARTIAL;
  getrf_rec<T>(m, n1, A, ipiv, info_off, Linv, Uinv, st, ws);
  if (piv) laswp<T>(A, m, n1, n, 0, n1, ipiv, ws);
  trsm_llu<T>(n1, n2, A, Linv, MatView<T>{A.at(0, n1), A.ld});
  gemm<T>(m - n1, n2, n1, MONE, MatView<T>{A.at(n1, 0), A.ld}, MatView<T>{A.at(0, n1), A.ld}, ONE,
          MatView<T>{A.at(n1, n1), A.ld});
  getrf_rec<T>(m - n1, n2, MatView<T>{A.at(n1, n1), A.ld}, ipiv + n1, info_off + (int)n1,
               Linv + (n1 / TB) * TB2, Uinv + (n1 / TB) * TB2, st, ws);
  if (piv) {
    hipLaunchKernelGGL(k_add_offset, dim3((unsigned)cdiv_up(n2, 256)), dim3(256), 0, stream(), n2,
                       ipiv + n1, (int32_t)n1);
    laswp<T>(A, m, 0, n1, n1, n, ipiv, ws);
  }
}

template <class T>
int getrf(int64_t m, int64_t n, MatView<T> A, int32_t* ipiv, cplx<T>* Linv, cplx<T>* Uinv, int mode,
          int* would_swap, Workspace& ws) {
  MNP_CHECK(m >= n && n > 0, MNP_ERR_ARG, "getrf: need m >= n > 0");
  int* d = (int*)ws.get(5, 2 * sizeof(int));
  MNP_HIP(hipMemsetAsync(d, 0, 2 * sizeof(int), stream()));
  LUState st{mode, d, d + 1};
  getrf_rec<T>(m, n, A, ipiv, 0, Linv, Uinv, st, ws);
  if (mode == PIVOT_NONE) {  // identity pivots; one pass for the pivoting check
    hipLaunchKernelGGL(k_iota, dim3((unsigned)cdiv_up(n, 256)), dim3(256), 0, stream(), n, ipiv, 0);
    const dim3 grid((unsigned)cdiv_up(m, 256), (unsigned)cdiv_up(n, VERIFY_COLS));
    hipLaunchKernelGGL(k_verify_nopiv<T>, grid, dim3(256), 0, stream(), m, n, A.p, A.ld, st.flag);
    MNP_HIP(hipGetLastError());
  }
  int h[2] = {0, 0};
  MNP_HIP(hipMemcpyAsync(h, d, 2 * sizeof(int), hipMemcpyDeviceToHost, stream()));
  MNP_HIP(hipStreamSynchronize(stream()));
  if (would_swap) *would_swap = h[1];
  return h[0];
}

template <class T>
__global__ void k_gather_rows(int64_t n, int64_t ncols, const int32_t* sigma, const cplx<T>* A,
                              int64_t lda, cplx<T>* B, int64_t ldb) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (i < n && j < ncols) B[j * ldb + i] = A[j * lda + sigma[i]];
}
template <class T>
void gather_rows(int64_t n, int64_t ncols, const int32_t* sigma, MatView<T> A, MatView<T> B) {
  if (n <= 0 || ncols <= 0) return;
  hipLaunchKernelGGL(k_gather_rows<T>, grid2(n, ncols), dim3(256), 0, stream(), n, ncols, sigma, A.p,
                     A.ld, B.p, B.ld);
  MNP_HIP(hipGetLastError());
}

// X[:, sigma[r]] = W[:, r]
template <class T>
__global__ void k_scatter_cols(int64_t m, int64_t n, const int32_t* sigma, const cplx<T>* W, int64_t ldw,
                               cplx<T>* X, int64_t ldx) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t r = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (i < m && r < n) X[(int64_t)sigma[r] * ldx + i] = W[r * ldw + i];
}

template <class T>
void getrs(int64_t n, int64_t nrhs, const LU<T>& f, MatView<T> B, Workspace& ws) {
  if (f.sigma) {
    cplx<T>* W = (cplx<T>*)ws.get(1, sizeof(cplx<T>) * (size_t)(n * nrhs));
    MatView<T> Wv{W, n};
    gather_rows<T>(n, nrhs, f.sigma, B, Wv);
    copy_mat<T>(n, nrhs, Wv, B);
  }
  if (f.LinvNB && n >= 2 * TBN && nrhs <= TBN) {  // few right-hand sides: blocked, 3 launches per block
    trsm_llu_nb<T>(n, nrhs, f, B, ws);
    trsm_lun_nb<T>(n, nrhs, f, B, ws);
  } else if (f.LinvNB && n >= 2 * TBN) {
    trsm_llu_h<T>(n, nrhs, f.A, f.Linv, f.LinvNB, B, ws);
    trsm_lun_h<T>(n, nrhs, f.A, f.Uinv, f.UinvNB, B, ws);
  } else {
    trsm_llu<T>(n, nrhs, f.A, f.Linv, B);
    trsm_lun<T>(n, nrhs, f.A, f.Uinv, B);
  }
}

template <class T>
void getrs_right(int64_t m, int64_t n, const LU<T>& f, MatView<T> B, Workspace& ws) {
  if (f.LinvNB && n >= 2 * TBN && m <= TBN) {
    trsm_run_nb<T>(m, n, f, B, ws);
    trsm_rlu_nb<T>(m, n, f, B, ws);
  } else if (f.LinvNB && n >= 2 * TBN) {
    trsm_run_h<T>(m, n, f.A, f.Uinv, f.UinvNB, B, ws);
    trsm_rlu_h<T>(m, n, f.A, f.Linv, f.LinvNB, B, ws);
  } else {
    trsm_run<T>(m, n, f.A, f.Uinv, B);
    trsm_rlu<T>(m, n, f.A, f.Linv, B);
  }
  if (f.sigma) {
    cplx<T>* W = (cplx<T>*)ws.get(1, sizeof(cplx<T>) * (size_t)(m * n));
    copy_mat<T>(m, n, B, MatView<T>{W, m});
    hipLaunchKernelGGL(k_scatter_cols<T>, grid2(m, n), dim3(256), 0, stream(), m, n, f.sigma, W, m, B.p,
                       B.ld);
    MNP_HIP(hipGetLastError());
  }
}

template <class T>
__global__ void k_perm_identity(int64_t n, const int32_t* sigma, cplx<T>* X, int64_t ldx) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (i < n && j < n) X[j * ldx + i] = (sigma ? sigma[i] : i) == j ? mk<T>(1, 0) : czero<T>();
}

template <class T>
void getri(int64_t n, const LU<T>& f, MatView<T> X, Workspace& ws) {
  hipLaunchKernelGGL(k_perm_identity<T>, grid2(n, n), dim3(256), 0, stream(), n, f.sigma, X.p, X.ld);
  MNP_HIP(hipGetLastError());
  if (f.LinvNB && n >= 2 * TBN) {
    trsm_llu_h<T>(n, n, f.A, f.Linv, f.LinvNB, X, ws);
    trsm_lun_h<T>(n, n, f.A, f.Uinv, f.UinvNB, X, ws);
  } else {
    trsm_llu<T>(n, n, f.A, f.Linv, X);
    trsm_lun<T>(n, n, f.A, f.Uinv, X);
  }
}

// ---------------------------------------------------- instantiations
#define MNP_INST(T)                                                                              \
  template void gemm<T>(int64_t, int64_t, int64_t, double2, MatView<T>, MatView<T>, double2,     \
                        MatView<T>);                                                             \
  template int getrf<T>(int64_t, int64_t, MatView<T>, int32_t*, cplx<T>*, cplx<T>*, int, int*,   \
                        Workspace&);                                                             \
  template void block_inverses<T>(int64_t, MatView<T>, cplx<T>*, cplx<T>*);                      \
  template void block_inverses_nb<T>(int64_t, MatView<T>, const cplx<T>*, const cplx<T>*, cplx<T>*, \
                                     cplx<T>*, Workspace&);                                       \
  template void gather_rows<T>(int64_t, int64_t, const int32_t*, MatView<T>, MatView<T>);        \
  template void getrs<T>(int64_t, int64_t, const LU<T>&, MatView<T>, Workspace&);                \
  template void getrs_right<T>(int64_t, int64_t, const LU<T>&, MatView<T>, Workspace&);          \
  template void getri<T>(int64_t, const LU<T>&, MatView<T>, Workspace&);                         \
  template void copy_mat<T>(int64_t, int64_t, MatView<T>, MatView<T>);
MNP_INST(double)
MNP_INST(float)
template void convert_mat<double, float>(int64_t, int64_t, MatView<double>, MatView<float>);
template void convert_mat<float, double>(int64_t, int64_t, MatView<float>, MatView<double>);

}  // namespace mnp

// ============================================================== C ABI
using namespace mnp;

namespace {
// Process-lifetime scratch of the ABI's building blocks; never destroyed (no
// hipFree from a static destructor while the process exits).
Workspace& abi_ws() {
  static Workspace* ws = new Workspace;
  return *ws;
}
Workspace& abi_inv() {
  static Workspace* ws = new Workspace;
  return *ws;
}
// Factors handed over through the ABI: permutation and block inverses are
// rebuilt from (LU, ipiv).
template <class T>
static LU<T> abi_factors(int64_t n, const void* A, int64_t lda, const int32_t* ipiv) {
  int32_t* sigma = (int32_t*)abi_inv().get(1, sizeof(int32_t) * n);
  perm_from_pivots(n, ipiv, sigma, abi_ws());
  cplx<T>* inv = (cplx<T>*)abi_inv().get(0, 2 * sizeof(cplx<T>) * inv_blocks(n));
  MatView<T> f{(cplx<T>*)A, lda};
  block_inverses<T>(n, f, inv, inv + inv_blocks(n));
  return LU<T>{f, inv, inv + inv_blocks(n), sigma};
}

template <class F>
int by_precision(int precision, F&& f) {
  return guarded([&] {
    MNP_CHECK(precision == MNP_FP64 || precision == MNP_FP32, MNP_ERR_ARG, "unknown precision");
    if (precision == MNP_FP64)
      f(double());
    else
      f(float());
  });
}
}  // namespace

extern "C" {

int mnp_gemm(int precision, int64_t m, int64_t n, int64_t k, mnp_complex alpha, const void* A,
             int64_t lda, const void* B, int64_t ldb, mnp_complex beta, void* C, int64_t ldc) {
  if (precision == MNP_FP32_SPLIT_BF16)
    return guarded([&] {
      MNP_CHECK(m >= 0 && n >= 0 && k >= 0 && lda >= std::max<int64_t>(1, m) &&
                    ldb >= std::max<int64_t>(1, k) && ldc >= std::max<int64_t>(1, m),
                MNP_ERR_ARG, "gemm: bad sizes");
      MNP_CHECK(m == 0 || n == 0 || (A && B && C), MNP_ERR_ARG, "gemm: NULL operand");
      if (m == 0 || n == 0) return;
      const bool bz = beta.re == 0 && beta.im == 0;
      MNP_CHECK(k > 0 || bz, MNP_ERR_ARG, "gemm: k == 0 with beta != 0 unsupported");
      gemm_sbf16<4, 4, 2>(m, n, k, make_float2((float)alpha.re, (float)alpha.im),
                                 MatView<float>{(float2*)A, lda}, MatView<float>{(float2*)B, ldb},
                                 make_float2((float)beta.re, (float)beta.im), bz,
                                 MatView<float>{(float2*)C, ldc});
    });
  return by_precision(precision, [&](auto t) {
    using T = decltype(t);
    MNP_CHECK(m >= 0 && n >= 0 && k >= 0 && lda >= std::max<int64_t>(1, m) &&
                  ldb >= std::max<int64_t>(1, k) && ldc >= std::max<int64_t>(1, m),
              MNP_ERR_ARG, "gemm: bad sizes");
    gemm<T>(m, n, k, to_d2(alpha), MatView<T>{(cplx<T>*)A, lda}, MatView<T>{(cplx<T>*)B, ldb},
            to_d2(beta), MatView<T>{(cplx<T>*)C, ldc});
  });
}

int mnp_getrf(int precision, int64_t m, int64_t n, void* A, int64_t lda, int32_t* ipiv, int* info) {
  return by_precision(precision, [&](auto t) {
    using T = decltype(t);
    MNP_CHECK(A && ipiv && m >= n && n > 0 && lda >= m, MNP_ERR_ARG, "getrf: bad arguments");
    cplx<T>* inv = (cplx<T>*)abi_inv().get(0, 2 * sizeof(cplx<T>) * inv_blocks(n));
    const int r = getrf<T>(m, n, MatView<T>{(cplx<T>*)A, lda}, ipiv, inv, inv + inv_blocks(n),
                           PIVOT_PARTIAL, nullptr, abi_ws());
    if (info) *info = r;
  });
}

int mnp_getrf_nopiv(int precision, int64_t m, int64_t n, void* A, int64_t lda, int* info,
                    int* would_swap) {
  return by_precision(precision, [&](auto t) {
    using T = decltype(t);
    MNP_CHECK(A && m >= n && n > 0 && lda >= m, MNP_ERR_ARG, "getrf_nopiv: bad arguments");
    cplx<T>* inv = (cplx<T>*)abi_inv().get(0, 2 * sizeof(cplx<T>) * inv_blocks(n));
    int32_t* ipiv = (int32_t*)abi_inv().get(1, sizeof(int32_t) * n);
    int sw = 0;
    const int r = getrf<T>(m, n, MatView<T>{(cplx<T>*)A, lda}, ipiv, inv, inv + inv_blocks(n),
                           PIVOT_NONE, &sw, abi_ws());
    if (info) *info = r;
    if (would_swap) *would_swap = sw;
  });
}

int mnp_getrs(int precision, int64_t n, int64_t nrhs, const void* LUp, int64_t lda, const int32_t* ipiv,
              void* B, int64_t ldb) {
  return by_precision(precision, [&](auto t) {
    using T = decltype(t);
    MNP_CHECK(LUp && ipiv && B && n > 0 && nrhs >= 0 && lda >= n && ldb >= n, MNP_ERR_ARG,
              "getrs: bad arguments");
    getrs<T>(n, nrhs, abi_factors<T>(n, LUp, lda, ipiv), MatView<T>{(cplx<T>*)B, ldb}, abi_ws());
  });
}

int mnp_getrs_right(int precision, int64_t m, int64_t n, const void* LUp, int64_t lda,
                    const int32_t* ipiv, void* B, int64_t ldb) {
  return by_precision(precision, [&](auto t) {
    using T = decltype(t);
    MNP_CHECK(LUp && ipiv && B && n > 0 && m >= 0 && lda >= n && ldb >= m, MNP_ERR_ARG,
              "getrs_right: bad arguments");
    getrs_right<T>(m, n, abi_factors<T>(n, LUp, lda, ipiv), MatView<T>{(cplx<T>*)B, ldb}, abi_ws());
  });
}

int mnp_getri(int precision, int64_t n, const void* LUp, int64_t lda, const int32_t* ipiv, void* X,
              int64_t ldx) {
  return by_precision(precision, [&](auto t) {
    using T = decltype(t);
    MNP_CHECK(LUp && ipiv && X && n > 0 && lda >= n && ldx >= n, MNP_ERR_ARG, "getri: bad arguments");
    getri<T>(n, abi_factors<T>(n, LUp, lda, ipiv), MatView<T>{(cplx<T>*)X, ldx}, abi_ws());
  });
}

}  // extern "C"
