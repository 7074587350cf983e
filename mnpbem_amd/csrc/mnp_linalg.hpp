// Dense complex linear algebra on gfx950 (column-major, interleaved complex).
#pragma once

#include "mnp_greenret.hpp"

namespace mnp {

// Grow-only device scratch owned by a solver handle.
struct Workspace {
  void* buf[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  size_t cap[6] = {0, 0, 0, 0, 0, 0};
  void* get(int slot, size_t bytes) {
    if (bytes > cap[slot]) {
      if (buf[slot]) MNP_HIP(hipFree(buf[slot]));
      buf[slot] = nullptr;
      MNP_HIP(hipMalloc(&buf[slot], bytes));
      cap[slot] = bytes;
    }
    return buf[slot];
  }
  ~Workspace() {
    for (void* p : buf)
      if (p) (void)hipFree(p);
  }
};

// Matrix view: column-major, leading dimension ld.
template <class T>
struct MatView {
  cplx<T>* p;
  int64_t ld;
  __host__ __device__ cplx<T>* at(int64_t i, int64_t j) const { return p + j * ld + i; }
};

// C = alpha * A * B + beta * C  (A m x k, B k x n); beta == 0 ignores C's contents.
template <class T>
void gemm(int64_t m, int64_t n, int64_t k, double2 alpha, MatView<T> A, MatView<T> B,
          double2 beta, MatView<T> C);

// In-place LU with partial pivoting of an m x n matrix (m >= n), LAPACK
// convention: swaps applied to rows produce L U; ipiv[j] = absolute row
// (0-based) swapped with row j.  Returns the first zero pivot + 1, or 0.
template <class T>
int getrf(int64_t m, int64_t n, MatView<T> A, int32_t* ipiv, Workspace& ws);

// sigma (device, n) with (S A)[r] = A[sigma[r]] for the swaps ipiv[0..n).
void perm_from_pivots(int64_t n, const int32_t* ipiv, int32_t* sigma, Workspace& ws);

// Row gather B[r, :] = A[sigma[r], :] for r < n over ncols columns (A != B).
template <class T>
void gather_rows(int64_t n, int64_t ncols, const int32_t* sigma, MatView<T> A, MatView<T> B);

// Triangular solves with the factors of getrf (n x n), in place on B:
//   lower unit from the left  : B <- L^{-1} B        (B n x nrhs)
//   upper from the left       : B <- U^{-1} B
//   upper from the right      : B <- B U^{-1}        (B m x n)
//   lower unit from the right : B <- B L^{-1}
template <class T>
void trsm_llu(int64_t n, int64_t nrhs, MatView<T> LU, MatView<T> B, Workspace& ws);
template <class T>
void trsm_lun(int64_t n, int64_t nrhs, MatView<T> LU, MatView<T> B, Workspace& ws);
template <class T>
void trsm_run(int64_t m, int64_t n, MatView<T> LU, MatView<T> B, Workspace& ws);
template <class T>
void trsm_rlu(int64_t m, int64_t n, MatView<T> LU, MatView<T> B, Workspace& ws);

// B <- A^{-1} B using getrf factors and sigma (B n x nrhs).
template <class T>
void getrs(int64_t n, int64_t nrhs, MatView<T> LU, const int32_t* sigma, MatView<T> B,
           Workspace& ws);
// X <- B A^{-1} (B m x n, in place).
template <class T>
void getrs_right(int64_t m, int64_t n, MatView<T> LU, const int32_t* sigma, MatView<T> B,
                 Workspace& ws);
// X = A^{-1} (n x n) from getrf factors and sigma.
template <class T>
void getri(int64_t n, MatView<T> LU, const int32_t* sigma, MatView<T> X, Workspace& ws);

// Elementwise helpers.
template <class T>
void copy_mat(int64_t m, int64_t n, MatView<T> A, MatView<T> B);  // B = A
template <class Ts, class Td>
void convert_mat(int64_t m, int64_t n, MatView<Ts> A, MatView<Td> B);

}  // namespace mnp
