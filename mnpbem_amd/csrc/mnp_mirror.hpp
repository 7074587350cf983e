// Mirror symmetry helpers shared by the retarded and quasistatic solvers with
// mirror symmetry (MNPBEM Greenfun/@compgreenretmirror, @compgreenstatmirror).
//
// The full particle is J = 2 or 4 copies of the n-face part p, copy j in
// faces j n .. (j + 1) n - 1.  For symmetry value s (row s of the J x J
// table of +-1, Particles/@comparticlemirror/private/init.m:29-37) a matrix
// of the part is M_s = sum_j tab[s][j] M(p, copy j) (compgreenretmirror/
// eval.m:23-42, compgreenstatmirror/eval.m:18-40).
#pragma once

#include "mnp_greenret.hpp"

namespace mnp {

struct SymTab {
  int8_t t[4][4];
};
struct SymOut {
  OutPair o[4];
};

// out_s[i, k] = sum_j tab[s][j] B[i, j n + k]  (B n x J n, leading dimension n)
template <int J>
__global__ void __launch_bounds__(256) k_sym_contract(int64_t n, SymTab tab, const double2* __restrict__ B,
                                                      SymOut out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t k = blockIdx.y;
  if (i >= n) return;
  double2 v[J];
#pragma unroll
  for (int j = 0; j < J; ++j) v[j] = B[(j * n + k) * n + i];
#pragma unroll
  for (int s = 0; s < J; ++s) {
    double2 acc = mk<double>(0, 0);
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const double t = tab.t[s][j];
      acc.x += t * v[j].x;
      acc.y += t * v[j].y;
    }
    out.o[s].store(k * n + i, acc);
  }
}

// Launches k_sym_contract<J> on the current stream (n <= 65535: one grid row per column).
inline void sym_contract(int J, int64_t n, const SymTab& tab, const double2* B, const OutPair* outs) {
  SymOut o{};
  for (int s = 0; s < J; ++s) o.o[s] = outs[s];
  const dim3 grid((unsigned)cdiv_up(n, 256), (unsigned)n);
  if (J == 2)
    hipLaunchKernelGGL(k_sym_contract<2>, grid, dim3(256), 0, stream(), n, tab, B, o);
  else
    hipLaunchKernelGGL(k_sym_contract<4>, grid, dim3(256), 0, stream(), n, tab, B, o);
  MNP_HIP(hipGetLastError());
}

// Checks a J x J symmetry table (row-major): entries +-1, column 0 all +1
// (copy 0 is the part itself), rows mutually orthogonal; copies it into tab.
inline void check_symtable(const char* who, int J, const int* table, SymTab& tab) {
  const std::string w(who);
  MNP_CHECK(J == 2 || J == 4, MNP_ERR_ARG, w + ": nsym must be 2 or 4");
  for (int s = 0; s < J; ++s) {
    MNP_CHECK(table[s * J] == 1, MNP_ERR_ARG, w + ": symmetry table column 0 must be +1");
    for (int t = 0; t < J; ++t) {
      int dot = 0;
      for (int j = 0; j < J; ++j) {
        const int a = table[s * J + j], b = table[t * J + j];
        MNP_CHECK(a == 1 || a == -1, MNP_ERR_ARG, w + ": symmetry table entries must be +-1");
        dot += a * b;
      }
      MNP_CHECK(dot == (s == t ? J : 0), MNP_ERR_ARG, w + ": symmetry table rows not orthogonal");
    }
  }
  for (int s = 0; s < J; ++s)
    for (int j = 0; j < J; ++j) tab.t[s][j] = (int8_t)table[s * J + j];
}

}  // namespace mnp
