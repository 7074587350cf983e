// Dense complex linear algebra on gfx950 (column-major, interleaved complex).
#pragma once

#include "mnp_greenret.hpp"

namespace mnp {

// Grow-only device scratch owned by a solver handle.
struct Workspace {
  static constexpr int NSLOT = 8;
  void* buf[NSLOT] = {};
  size_t cap[NSLOT] = {};
  void* get(int slot, size_t bytes) {
    MNP_CHECK(slot >= 0 && slot < NSLOT, MNP_ERR_INTERNAL, "workspace slot out of range");
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

// Batched products of equal shapes at element strides (operand z at X + z sx)
// with a chosen output tile; see linalg.hip for when C may alias A or B.
enum { GEMM_TILE_64 = 0, GEMM_TILE_128x64 = 1, GEMM_TILE_128x96_3M = 2, GEMM_NTILES = 3 };
template <class T>
void gemm_strided(int64_t m, int64_t n, int64_t k, double2 alpha, MatView<T> A, int64_t sa, MatView<T> B,
                  int64_t sb, double2 beta, MatView<T> C, int64_t sc, int64_t batch, int tile);

// While set, large complex64 products (the 128 x 96 tile class) run on the
// split-bf16 kernel: each FP32 operand as hi + lo bf16, three bf16 products
// per real product, FP32 accumulation (about 2^-16 relative per product).
void set_split_bf16(int on);
int split_bf16();
struct SplitBf16Scope {
  int prev;
  explicit SplitBf16Scope(int on) : prev(split_bf16()) { set_split_bf16(on); }
  ~SplitBf16Scope() { set_split_bf16(prev); }
};

enum { PIVOT_PARTIAL = 0, PIVOT_NONE = 1 };

// In-place LU of an m x n matrix (m >= n) in LAPACK convention: the swaps
// ipiv[j] (0-based row exchanged with row j) applied to the rows give L U.
//   PIVOT_PARTIAL: partial pivoting on the first largest |Re|+|Im| (LAPACK getf2).
//   PIVOT_NONE:    no row exchanges (ipiv[j] = j); *would_swap is set when
//                  partial pivoting would have exchanged a row, i.e. when the
//                  factors differ from PIVOT_PARTIAL's.
// Also writes the inverses of the 32 x 32 diagonal blocks of L and U
// (ceil(n/32) blocks of 1024 entries each).  Returns the first zero pivot
// + 1, or 0.
template <class T>
int getrf(int64_t m, int64_t n, MatView<T> A, int32_t* ipiv, cplx<T>* Linv, cplx<T>* Uinv,
          int mode, int* would_swap, Workspace& ws);
// Diagonal-block inverses of existing factors.
template <class T>
void block_inverses(int64_t n, MatView<T> LU, cplx<T>* Linv, cplx<T>* Uinv);
// Inverses of the full 256 x 256 diagonal blocks of L and U (n / 256 blocks,
// 256 x 256 column-major each) from the 32 x 32 ones.
template <class T>
void block_inverses_nb(int64_t n, MatView<T> LU, const cplx<T>* Linv, const cplx<T>* Uinv, cplx<T>* LinvNB,
                       cplx<T>* UinvNB, Workspace& ws);
inline int64_t inv_blocks_nb(int64_t n) { return (n / 256) * 256 * 256; }  // entries per 256-block array

// Factors of an n x n matrix.  sigma (device, n) gives (S A)[r] = A[sigma[r]]
// for the pivoting permutation S, or is NULL when no rows were exchanged.
template <class T>
struct LU {
  MatView<T> A;
  cplx<T>* Linv;
  cplx<T>* Uinv;
  const int32_t* sigma;
  cplx<T>* LinvNB = nullptr;  // inverses of the 256 x 256 diagonal blocks (block_inverses_nb), or null
  cplx<T>* UinvNB = nullptr;
};

// sigma from the swaps ipiv[0..n).
void perm_from_pivots(int64_t n, const int32_t* ipiv, int32_t* sigma, Workspace& ws);

// Row gather B[r, :] = A[sigma[r], :] for r < n over ncols columns (A != B).
template <class T>
void gather_rows(int64_t n, int64_t ncols, const int32_t* sigma, MatView<T> A, MatView<T> B);

// B <- A^{-1} B (B n x nrhs); B <- B A^{-1} (B m x n); X = A^{-1}.
template <class T>
void getrs(int64_t n, int64_t nrhs, const LU<T>& f, MatView<T> B, Workspace& ws);
template <class T>
void getrs_right(int64_t m, int64_t n, const LU<T>& f, MatView<T> B, Workspace& ws);
template <class T>
void getri(int64_t n, const LU<T>& f, MatView<T> X, Workspace& ws);

// Elementwise helpers.
template <class T>
void copy_mat(int64_t m, int64_t n, MatView<T> A, MatView<T> B);  // B = A
template <class Ts, class Td>
void convert_mat(int64_t m, int64_t n, MatView<Ts> A, MatView<Td> B);

inline int64_t inv_blocks(int64_t n) { return (n + 31) / 32 * 1024; }  // entries per inverse array

}  // namespace mnp
