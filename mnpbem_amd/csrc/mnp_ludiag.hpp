// Unpivoted LU of one diagonal block (nb <= 32) with the inverses of its
// triangular factors, as device functions shared by the dense chain
// (linalg.hip) and the batched solver (batch.hip).  Each is the body of a
// workgroup: lu_nopiv_diag_body for complex64 with 192 threads,
// lu_nopiv_diag_lds_body for complex128 with 256 threads.
#pragma once

#include <type_traits>

#include "mnp_common.hpp"

namespace mnp {

constexpr int TB = 32;
constexpr int64_t TB2 = TB * TB;

template <class T> __device__ __forceinline__ cplx<T> czero() { return mk<T>(T(0), T(0)); }

template <class T> __device__ __forceinline__ T bcast(T v, int l);
template <> __device__ __forceinline__ float bcast<float>(float v, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
template <> __device__ __forceinline__ double bcast<double>(double v, int l) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
template <class T> __device__ __forceinline__ cplx<T> bcastc(cplx<T> v, int l) {
  return mk<T>(bcast<T>(v.x, l), bcast<T>(v.y, l));
}

// Compile-time loop: f(std::integral_constant<int, i>) for i = B .. E - 1,
// so that register arrays are indexed by constants.
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

template <class T>
__device__ __forceinline__ void lu_nopiv_diag_body(int nb, cplx<T>* A, int64_t lda, int* info, int info_off,
                                                   cplx<T>* Linv, cplx<T>* Uinv) {
  __shared__ cplx<T> sP[TB][TB + 1];  // the block, row r at sP[r]; then its factors
  __shared__ cplx<T> sX[TB][TB + 1];  // L^{-1}
  __shared__ cplx<T> sY[TB][TB + 1];  // U^{-1}
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int r = lane & (TB - 1);  // lanes 32..63 repeat rows 0..31 and store nothing
  const cplx<T> one = mk<T>(1, 0);
  for (int idx = tid; idx < TB * TB; idx += 192) {
    const int rr = idx % TB, c = idx / TB;
    sP[rr][c] = (rr < nb && c < nb) ? A[c * lda + rr] : mk<T>(rr == c ? T(1) : T(0), T(0));
  }
  __syncthreads();
  if (wid == 0) {
    cplx<T> p[TB];
    static_for<0, TB>([&](auto C) { p[C] = sP[r][C]; });
    int first_zero = -1;
    static_for<0, TB>([&](auto J) {
      constexpr int j = J;
      const cplx<T> piv = bcastc<T>(p[j], j);
      // a zero pivot (uniform) is reported; its column is left as it is
      const bool zp = piv.x == 0 && piv.y == 0;
      if (zp && first_zero < 0) first_zero = j;
      const cplx<T> inv = cdiv<T>(one, zp ? one : piv);
      const T below = (r > j && !zp) ? T(1) : T(0);  // multiplier of this row (0 on and above)
      const cplx<T> l = cscale<T>(cmul<T>(p[j], inv), below);
      p[j] = (r > j && !zp) ? l : p[j];
      static_for<j + 1, TB>([&](auto C) { p[C] = csub<T>(p[C], cmul<T>(l, bcastc<T>(p[C], j))); });
    });
    if (lane < TB) static_for<0, TB>([&](auto C) { sP[r][C] = p[C]; });
    if (lane == 0 && first_zero >= 0 && first_zero < nb && *info == 0) *info = info_off + first_zero + 1;
  }
  __syncthreads();
  if (wid == 1) {
    // row r of L^{-1}: row k is final at step k and is eliminated from the rows below
    cplx<T> x[TB];
    static_for<0, TB>([&](auto C) { x[C] = mk<T>(C == r ? T(1) : T(0), T(0)); });
    static_for<0, TB - 1>([&](auto K) {
      constexpr int k = K;
      const cplx<T> lrk = cscale<T>(sP[r][k], r > k ? T(1) : T(0));
      static_for<0, k + 1>([&](auto C) { x[C] = csub<T>(x[C], cmul<T>(lrk, bcastc<T>(x[C], k))); });
    });
    if (lane < TB) static_for<0, TB>([&](auto C) { sX[r][C] = x[C]; });
  } else if (wid == 2) {
    // row r of U~^{-1} (U~ = D^{-1} U): row k is final at step k (descending)
    const cplx<T> dinv = cdiv<T>(one, sP[r][r]);
    cplx<T> y[TB];
    static_for<0, TB>([&](auto C) { y[C] = mk<T>(C == r ? T(1) : T(0), T(0)); });
    static_for<1, TB>([&](auto KK) {
      constexpr int k = TB - KK;  // TB - 1 .. 1
      const cplx<T> urk = cscale<T>(cmul<T>(sP[r][k], dinv), r < k ? T(1) : T(0));
      static_for<k, TB>([&](auto C) { y[C] = csub<T>(y[C], cmul<T>(urk, bcastc<T>(y[C], k))); });
    });
    static_for<0, TB>([&](auto C) { y[C] = cmul<T>(y[C], bcastc<T>(dinv, C)); });  // columns times 1/u_cc
    if (lane < TB) static_for<0, TB>([&](auto C) { sY[r][C] = y[C]; });
  }
  __syncthreads();
  for (int idx = tid; idx < TB * TB; idx += 192) {
    const int rr = idx % TB, c = idx / TB;
    if (rr < nb && c < nb) A[c * lda + rr] = sP[rr][c];
    Linv[idx] = sX[rr][c];
    Uinv[idx] = sY[rr][c];
  }
}

// The same for complex128, in LDS by one workgroup (the register form
// above would need more than 256 VGPRs per lane for FP64 rows): 32 steps of
// elimination over the 1,024 entries, then L^{-1} and U^{-1} by elimination
// on the identity.
template <class T>
__device__ __forceinline__ void lu_nopiv_diag_lds_body(int nb, cplx<T>* A, int64_t lda, int* info, int info_off,
                                                       cplx<T>* Linv, cplx<T>* Uinv) {
  __shared__ cplx<T> P[TB][TB + 1];
  __shared__ cplx<T> sL[TB][TB + 1];
  __shared__ cplx<T> sU[TB][TB + 1];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < TB * TB; idx += 256) {
    const int r = idx % TB, c = idx / TB;
    P[r][c] = (r < nb && c < nb) ? A[c * lda + r] : (r == c ? mk<T>(1, 0) : czero<T>());
  }
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    const cplx<T> piv = P[j][j];
    const bool ok = piv.x != 0 || piv.y != 0;
    if (!ok && tid == 0 && *info == 0) *info = info_off + j + 1;
    if (ok && tid > j && tid < nb) P[tid][j] = cdiv<T>(P[tid][j], piv);
    __syncthreads();
    if (ok)
      for (int idx = tid; idx < TB * TB; idx += 256) {
        const int r = idx % TB, c = idx / TB;
        if (r > j && c > j && r < nb && c < nb) P[r][c] = csub<T>(P[r][c], cmul<T>(P[r][j], P[j][c]));
      }
    __syncthreads();
  }
  // L^{-1} and U^{-1} by elimination on the identity, all 1024 entries per
  // step: X <- L^{-1} I forward (row k is final at step k), Y <- U^{-1} I
  // backward (row k scaled by 1/u_kk, then eliminated from the rows above).
  // The padding rows/columns (>= nb) of P hold the identity.
  for (int idx = tid; idx < TB * TB; idx += 256) {
    const int r = idx % TB, c = idx / TB;
    sL[r][c] = r == c ? mk<T>(1, 0) : czero<T>();
    sU[r][c] = r == c ? mk<T>(1, 0) : czero<T>();
  }
  __syncthreads();
  for (int k = 0; k < TB; ++k) {
    for (int idx = tid; idx < TB * TB; idx += 256) {
      const int r = idx % TB, c = idx / TB;
      if (r > k && c <= k) sL[r][c] = csub<T>(sL[r][c], cmul<T>(P[r][k], sL[k][c]));
    }
    __syncthreads();
  }
  for (int k = TB - 1; k >= 0; --k) {
    const cplx<T> inv = cdiv<T>(mk<T>(1, 0), P[k][k]);
    if (tid < TB && tid >= k) sU[k][tid] = cmul<T>(sU[k][tid], inv);
    __syncthreads();
    for (int idx = tid; idx < TB * TB; idx += 256) {
      const int r = idx % TB, c = idx / TB;
      if (r < k && c >= k) sU[r][c] = csub<T>(sU[r][c], cmul<T>(P[r][k], sU[k][c]));
    }
    __syncthreads();
  }
  for (int idx = tid; idx < TB * TB; idx += 256) {
    const int r = idx % TB, c = idx / TB;
    if (r < nb && c < nb) A[c * lda + r] = P[r][c];
    Linv[idx] = sL[r][c];
    Uinv[idx] = sU[r][c];
  }
}

}  // namespace mnp
