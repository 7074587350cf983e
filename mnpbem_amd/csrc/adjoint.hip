// Adjoint of the dense retarded BEM solve on gfx950: the transposed solve
// lam = A^{-T} c with the forward factors, and the gradient of the weighted
// plane-wave extinction with respect to the vertex positions and to the
// inside dielectric function.
//
// No counterpart in the reference, which has no derivatives; this
// differentiates the operator of BEM/@bemretiter/private/afun.m:6-35 that
// BEM/@bemret/private/initmat.m:5-72 and mldivide.m:11-49 invert.  With
// A x = b and J = c^T x = pol . e(dir) (@planewaveret/extinction.m:12-16,
// @spectrumret/farfield.m:59-84), a real parameter t moves J by
//   dJ/dt = dc/dt^T x + lam^T (db/dt - dA/dt x),   A^T lam = c.
// lam^T dA x is a sum over the elements of G(k1), G(k2), F(k1), F(k2) and the
// normals: element (r, c) of each matrix carries the weight W_rc = u_r . x_c
// of a lam-vector u and an x-vector, so its derivative with respect to the
// geometry is taken where the element is evaluated (eval1.m's closed form,
// the Taylor quadrature of greenret/private/init.m:42-177, the polar self
// rule, and the closed-surface sum rule of compgreenret/private/
// initclosed.m:19-89), and nothing of dA is stored.
//
// Transposed factors: A = S^T L U gives A^T = (U^T D^-1)(D L^T) S with D the
// diagonal of U, so the factors of A^T follow from the forward ones by one
// elementwise pass; their block inverses are rebuilt, and the permutation S
// is applied after the triangular solves.
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "mnp_bemret.hpp"
#include "mnp_chain.hpp"
#include "mnp_linalg.hpp"

namespace mnp {

// ------------------------------------------------------------ transposes
// B (n x m, ld ldb) = A^T (A m x n, ld lda); 32 x 32 tiles through LDS.
template <class T>
__global__ void __launch_bounds__(256) k_transpose(int64_t m, int64_t n, const cplx<T>* __restrict__ A,
                                                   int64_t lda, cplx<T>* __restrict__ B, int64_t ldb) {
  __shared__ cplx<T> tile[32][33];
  const int64_t r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;  // tile of A: rows r0.., cols c0..
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {
    const int64_t r = r0 + tx, c = c0 + yy;
    if (r < m && c < n) tile[yy][tx] = A[c * lda + r];
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int64_t i = c0 + tx, j = r0 + yy;  // B(i, j) = A(j, i)
    if (i < n && j < m) B[j * ldb + i] = tile[tx][yy];
  }
}
template <class T>
void transpose(int64_t m, int64_t n, const cplx<T>* A, int64_t lda, cplx<T>* B, int64_t ldb) {
  hipLaunchKernelGGL(k_transpose<T>, dim3((unsigned)cdiv_up(m, 32), (unsigned)cdiv_up(n, 32)), dim3(256), 0, stream(),
                     m, n, A, lda, B, ldb);
  MNP_HIP(hipGetLastError());
}

// Factors of A^T from the in-place factors L\U of A (n x n, ld n):
//   B(i, j) = U(j, i) / U(j, j) (i > j),  U(i, i) L(j, i) (i < j),  U(i, i) (i == j).
template <class T>
__global__ void __launch_bounds__(256) k_transpose_lu(int64_t n, const cplx<T>* __restrict__ A,
                                                      cplx<T>* __restrict__ B) {
  __shared__ cplx<T> tile[32][33];
  const int64_t r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {
    const int64_t r = r0 + tx, c = c0 + yy;
    if (r < n && c < n) tile[yy][tx] = A[c * n + r];
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int64_t i = c0 + tx, j = r0 + yy;  // B(i, j) from A(j, i)
    if (i >= n || j >= n) continue;
    const cplx<T> a = tile[tx][yy];
    cplx<T> v;
    if (i > j)
      v = cdiv<T>(a, A[j * n + j]);
    else if (i < j)
      v = cmul<T>(A[i * n + i], a);
    else
      v = a;
    B[j * n + i] = v;
  }
}

// B[sigma[r], :] = W[r, :] (rows of n x ncols arrays, ld n)
template <class T>
__global__ void k_scatter_rows(int64_t n, int64_t ncols, const int32_t* sigma, const cplx<T>* W, cplx<T>* B) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i < n && j < ncols) B[j * n + sigma[i]] = W[j * n + i];
}

// ------------------------------------------------------ transposed chain
// Sigma1^T, Deltai^T and the factors of G1^T, G2^T, Sigma^T (slots 0, 1, 2)
// of a built chain, in its precision.
template <class T>
struct TChain {
  int64_t n;
  cplx<T>*S1T = nullptr, *DiT = nullptr, *F[3] = {nullptr, nullptr, nullptr};
  cplx<T>*inv = nullptr, *invnb = nullptr;
  LU<T> f[3];
  const int32_t* sig[3] = {nullptr, nullptr, nullptr};
  Workspace ws, ws2;
  explicit TChain(int64_t n_) : n(n_) {
    const int64_t nn = n * n;
    S1T = dev_alloc<cplx<T>>(nn);
    DiT = dev_alloc<cplx<T>>(nn);
    for (auto& p : F) p = dev_alloc<cplx<T>>(nn);
    inv = dev_alloc<cplx<T>>(6 * inv_blocks(n));
    invnb = dev_alloc<cplx<T>>(6 * inv_blocks_nb(n));
  }
  ~TChain() {
    void* ps[] = {S1T, DiT, F[0], F[1], F[2], inv, invnb};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
  TChain(const TChain&) = delete;
  TChain& operator=(const TChain&) = delete;

  void build(const Chain<T>& c) {
    transpose<T>(n, n, c.S1, n, S1T, n);
    transpose<T>(n, n, c.Di, n, DiT, n);
    const int src[3] = {0, 1, 3};  // G1, G2, Sigma of the forward chain
    for (int s = 0; s < 3; ++s) {
      const LU<T>& fw = c.f[src[s]];
      MNP_CHECK(fw.A.p && fw.A.ld == n, MNP_ERR_STATE, "adjoint: forward factors missing");
      hipLaunchKernelGGL(k_transpose_lu<T>, dim3((unsigned)cdiv_up(n, 32), (unsigned)cdiv_up(n, 32)), dim3(256), 0,
                         stream(), n, fw.A.p, F[s]);
      MNP_HIP(hipGetLastError());
      cplx<T>* Linv = inv + (2 * s) * inv_blocks(n);
      cplx<T>* Uinv = Linv + inv_blocks(n);
      cplx<T>* LinvNB = invnb + (2 * s) * inv_blocks_nb(n);
      cplx<T>* UinvNB = LinvNB + inv_blocks_nb(n);
      const MatView<T> v{F[s], n};
      block_inverses<T>(n, v, Linv, Uinv);
      block_inverses_nb<T>(n, v, Linv, Uinv, LinvNB, UinvNB, ws);
      f[s] = LU<T>{v, Linv, Uinv, nullptr, LinvNB, UinvNB};
      sig[s] = fw.sigma;
    }
  }
  // B <- (factor s)^{-T} B for B n x nrhs (ld n)
  void solve(int s, int64_t nrhs, cplx<T>* B) {
    getrs<T>(n, nrhs, f[s], MatView<T>{B, n}, ws);
    if (sig[s]) {
      cplx<T>* W = (cplx<T>*)ws2.get(0, sizeof(cplx<T>) * (size_t)(n * nrhs));
      copy_mat<T>(n, nrhs, MatView<T>{B, n}, MatView<T>{W, n});
      hipLaunchKernelGGL(k_scatter_rows<T>, dim3((unsigned)cdiv_up(n, 256), (unsigned)nrhs), dim3(256), 0, stream(),
                         n, nrhs, sig[s], W, B);
      MNP_HIP(hipGetLastError());
    }
  }
};

// Steps of the transposed mldivide (chain_apply read backwards); vectors are
// n x (.. p) blocks with components fastest (c + 3 q), lam in the layout of b.
// phibar = V1s, abar = V1h, S = V1s + V2s, H = V1h + V2h
template <class T>
__global__ void k_mldT1(int64_t n, int64_t p, const cplx<T>* V1, const cplx<T>* V2, cplx<T>* lam, cplx<T>* S,
                        cplx<T>* H) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t col = blockIdx.y;  // 0 .. 4p
  if (i >= n) return;
  const cplx<T> a = V1[col * n + i], b = V2[col * n + i];
  lam[col * n + i] = a;  // phi (col < p) and a (p .. 4p) occupy the same columns of b
  if (col < p)
    S[col * n + i] = cadd<T>(a, b);
  else
    H[(col - p) * n + i] = cadd<T>(a, b);
}
// S_q += ikdL sum_c n_c U_(c+3q)
template <class T>
__global__ void k_mldT2(int64_t n, int64_t p, cplx<T>* S, const cplx<T>* U, const double* nvec, double2 ikdL) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  double2 s = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) {
    const cplx<T> u = U[(c + 3 * q) * n + i];
    s = cadd<double>(s, cscale<double>(mk<double>(u.x, u.y), nvec[c * n + i]));
  }
  const double2 t = cmul<double>(ikdL, s);
  const cplx<T> v = S[q * n + i];
  S[q * n + i] = mk<T>((T)(v.x + t.x), (T)(v.y + t.y));
}
// De-bar = S ; Tb_(c+3q) = ikdL n_c S_q
template <class T>
__global__ void k_mldT3(int64_t n, int64_t p, const cplx<T>* S, cplx<T>* lam, cplx<T>* Tb, const double* nvec,
                        double2 ikdL) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const cplx<T> s = S[q * n + i];
  lam[(4 * p + q) * n + i] = s;
  const double2 t = cmul<double>(ikdL, mk<double>(s.x, s.y));
  for (int c = 0; c < 3; ++c) {
    const double nc = nvec[c * n + i];
    Tb[(c + 3 * q) * n + i] = mk<T>((T)(nc * t.x), (T)(nc * t.y));
  }
}
// alpha-bar = U ; Ya = -U ; phibar += ikL1 n . U ; Yp = -L1 De-bar ; abar += ikL1 n De-bar
template <class T>
__global__ void k_mldT4(int64_t n, int64_t p, const cplx<T>* U, cplx<T>* lam, cplx<T>* Ya, cplx<T>* Yp,
                        const double* nvec, double2 ikL1, double2 L1) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const cplx<T> de = lam[(4 * p + q) * n + i];
  const double2 deb = mk<double>(de.x, de.y);
  const double2 ikde = cmul<double>(ikL1, deb);
  double2 nu = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) {
    const int64_t col = c + 3 * q;
    const double nc = nvec[c * n + i];
    const cplx<T> u = U[col * n + i];
    lam[(5 * p + col) * n + i] = u;
    Ya[col * n + i] = mk<T>(-u.x, -u.y);
    nu = cadd<double>(nu, cscale<double>(mk<double>(u.x, u.y), nc));
    const cplx<T> a = lam[(p + col) * n + i];
    lam[(p + col) * n + i] = mk<T>((T)(a.x + nc * ikde.x), (T)(a.y + nc * ikde.y));
  }
  const double2 t = cmul<double>(ikL1, nu);
  const cplx<T> ph = lam[q * n + i];
  lam[q * n + i] = mk<T>((T)(ph.x + t.x), (T)(ph.y + t.y));
  const double2 y = cmul<double>(L1, deb);
  Yp[q * n + i] = mk<T>((T)(-y.x), (T)(-y.y));
}

// lam = M^{-T} c for the chain M (precision T): c in the layout of x, lam in
// the layout of b (n x 8p each).  tmp holds 22 p n entries.
template <class T>
void chain_apply_T(TChain<T>& t, const Chain<T>& c, const GreenRet& g, int64_t p, const cplx<T>* cin, cplx<T>* lam,
                   cplx<T>* tmp) {
  const int64_t n = t.n;
  const double2 L1 = c.eps1, L2 = c.eps2;
  const double2 ik = mk<double>(0, c.k0);
  const double2 ikdL = cmul<double>(ik, csub<double>(L1, L2)), ikL1 = cmul<double>(ik, L1);
  cplx<T>* V1 = tmp;            // 4p
  cplx<T>* V2 = V1 + 4 * p * n;  // 4p
  cplx<T>* S = V2 + 4 * p * n;   // p
  cplx<T>* H = S + p * n;        // 3p
  cplx<T>* U = H + 3 * p * n;    // 3p
  cplx<T>* Tb = U + 3 * p * n;   // 3p
  cplx<T>* Ya = Tb + 3 * p * n;  // 3p
  cplx<T>* Yp = Ya + 3 * p * n;  // p
  const double2 one = {1, 0}, zero = {0, 0};
  copy_mat<T>(n, 4 * p, MatView<T>{const_cast<cplx<T>*>(cin), n}, MatView<T>{V1, n});
  copy_mat<T>(n, 4 * p, MatView<T>{const_cast<cplx<T>*>(cin) + 4 * p * n, n}, MatView<T>{V2, n});
  t.solve(0, 4 * p, V1);  // G1^{-T}
  t.solve(1, 4 * p, V2);  // G2^{-T}
  const dim3 g1((unsigned)cdiv_up(n, 256), (unsigned)p);
  hipLaunchKernelGGL(k_mldT1<T>, dim3((unsigned)cdiv_up(n, 256), (unsigned)(4 * p)), dim3(256), 0, stream(), n, p, V1,
                     V2, lam, S, H);
  MNP_HIP(hipGetLastError());
  gemm<T>(n, 3 * p, n, one, MatView<T>{t.DiT, n}, MatView<T>{H, n}, zero, MatView<T>{U, n});
  hipLaunchKernelGGL(k_mldT2<T>, g1, dim3(256), 0, stream(), n, p, S, U, g.geo.nvec, ikdL);
  MNP_HIP(hipGetLastError());
  t.solve(2, p, S);  // Sigma^{-T}
  hipLaunchKernelGGL(k_mldT3<T>, g1, dim3(256), 0, stream(), n, p, S, lam, Tb, g.geo.nvec, ikdL);
  MNP_HIP(hipGetLastError());
  gemm<T>(n, 3 * p, n, one, MatView<T>{t.DiT, n}, MatView<T>{Tb, n}, one, MatView<T>{U, n});
  hipLaunchKernelGGL(k_mldT4<T>, g1, dim3(256), 0, stream(), n, p, U, lam, Ya, Yp, g.geo.nvec, ikL1, L1);
  MNP_HIP(hipGetLastError());
  gemm<T>(n, 3 * p, n, one, MatView<T>{t.S1T, n}, MatView<T>{Ya, n}, one, MatView<T>{lam + p * n, n});
  gemm<T>(n, p, n, one, MatView<T>{t.S1T, n}, MatView<T>{Yp, n}, one, MatView<T>{lam, n});
}

// ---------------------------------------------- transposed exact operator
// From lam (layout of b):  U1 = [lphi - i k0 e1 n . lal | la_c - i k0 e1 n_c lDe],
// W1 = [e1 lDe | lal], and U2, W2 with e2 (each n x 4p).  Then
//   A^T lam = [G1^T U1 + H1^T W1 | -(G2^T U2 + H2^T W2)]  (layout of x).
__global__ void k_adj_uw(int64_t n, int64_t p, const double2* lam, double2* U1, double2* W1, double2* U2,
                         double2* W2, const double* nvec, double k0, double2 e1, double2 e2) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const double2 ik = mk<double>(0, k0);
  const double2 lphi = lam[q * n + i], lde = lam[(4 * p + q) * n + i];
  double2 nlal = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) nlal = cadd<double>(nlal, cscale<double>(lam[(5 * p + c + 3 * q) * n + i], nvec[c * n + i]));
  const double2 ik1 = cmul<double>(ik, e1), ik2 = cmul<double>(ik, e2);
  U1[q * n + i] = csub<double>(lphi, cmul<double>(ik1, nlal));
  U2[q * n + i] = csub<double>(lphi, cmul<double>(ik2, nlal));
  W1[q * n + i] = cmul<double>(e1, lde);
  W2[q * n + i] = cmul<double>(e2, lde);
  for (int c = 0; c < 3; ++c) {
    const int64_t col = p + c + 3 * q;
    const double nc = nvec[c * n + i];
    const double2 la = lam[(p + c + 3 * q) * n + i], lal = lam[(5 * p + c + 3 * q) * n + i];
    U1[col * n + i] = csub<double>(la, cscale<double>(cmul<double>(ik1, lde), nc));
    U2[col * n + i] = csub<double>(la, cscale<double>(cmul<double>(ik2, lde), nc));
    W1[col * n + i] = lal;
    W2[col * n + i] = lal;
  }
}
// r = c - A^T lam from the four products (n x 4p each)
__global__ void k_adj_res(int64_t n, int64_t p, const double2* c, const double2* P1, const double2* Q1,
                          const double2* P2, const double2* Q2, double2* r) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t col = blockIdx.y;  // 0 .. 4p
  if (i >= n) return;
  const int64_t a = col * n + i, b = (4 * p + col) * n + i;
  r[a] = csub<double>(c[a], cadd<double>(P1[a], Q1[a]));
  r[b] = cadd<double>(c[b], cadd<double>(P2[a], Q2[a]));
}

__global__ void k_adj_axpy(int64_t len, double2* x, const double2* d) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) x[i] = cadd<double>(x[i], d[i]);
}

// ------------------------------------------------------------ adjoint state
struct Adjoint {
  int64_t n;
  uint64_t epoch = ~(uint64_t)0;
  std::unique_ptr<TChain<double>> t64;
  std::unique_ptr<TChain<float>> t32;
  double2 *G1T = nullptr, *G2T = nullptr, *H1T = nullptr, *H2T = nullptr;  // FP64 operator^T (mixed)
  Workspace vec, res, grad, big;
  // refined pairs by row (CSR) and faces by vertex (CSR), built once
  int64_t* rptr = nullptr;
  int32_t* rlist = nullptr;
  int64_t* vptr = nullptr;
  int64_t* vlist = nullptr;
  int* flag = nullptr;
  explicit Adjoint(int64_t n_) : n(n_) {}
  ~Adjoint() {
    void* ps[] = {G1T, G2T, H1T, H2T, rptr, rlist, vptr, vlist, flag};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
};
void adjoint_free(Adjoint* a) { delete a; }

void BemRet::solve_transposed(int64_t p, const double2* c, double2* lam) {
  MNP_CHECK(ready, MNP_ERR_STATE, "bemret_solve_transposed: initmat not called");
  MNP_CHECK(!g->compound(), MNP_ERR_STATE, "bemret_solve_transposed: not for compound particles");
  MNP_CHECK(p > 0 && p <= 64, MNP_ERR_ARG, "bemret_solve_transposed: npol must be in 1..64");
  if (!adj) adj = new Adjoint(n);
  Adjoint& a = *adj;
  const int64_t nn = n * n, len = 8 * p * n;
  if (a.epoch != epoch) {
    if (solver == MNP_SOLVER_FP64) {
      if (!a.t64) a.t64.reset(new TChain<double>(n));
      a.t64->build(*c64);
    } else {
      if (!a.t32) a.t32.reset(new TChain<float>(n));
      a.t32->build(*c32);
      double2** dst[4] = {&a.G1T, &a.G2T, &a.H1T, &a.H2T};
      const double2* src[4] = {G1, G2, H1, H2};
      for (int s = 0; s < 4; ++s) {
        if (!*dst[s]) *dst[s] = dev_alloc<double2>(nn);
        transpose<double>(n, n, src[s], n, *dst[s], n);
      }
    }
    a.epoch = epoch;
  }
  last_iter = 0;
  last_relres = 0;
  if (solver == MNP_SOLVER_FP64) {
    double2* tmp = (double2*)a.vec.get(0, sizeof(double2) * (size_t)(22 * p * n));
    chain_apply_T<double>(*a.t64, *c64, *g, p, c, lam, tmp);
    return;
  }
  // mixed: the FP32 transposed chain refined with the exact FP64 A^T
  float2* c32b = (float2*)a.vec.get(1, sizeof(float2) * len);
  float2* l32 = (float2*)a.vec.get(2, sizeof(float2) * len);
  float2* tmp = (float2*)a.vec.get(3, sizeof(float2) * (size_t)(22 * p * n));
  double2* work = (double2*)a.res.get(0, sizeof(double2) * (size_t)(40 * p * n));
  double2 *U1 = work, *W1 = U1 + 4 * p * n, *U2 = W1 + 4 * p * n, *W2 = U2 + 4 * p * n;
  double2 *P1 = W2 + 4 * p * n, *Q1 = P1 + 4 * p * n, *P2 = Q1 + 4 * p * n, *Q2 = P2 + 4 * p * n;
  double2* rr = Q2 + 4 * p * n;  // 8p
  auto precond = [&](const double2* src, double2* dst) {
    convert_mat<double, float>(n, 8 * p, MatView<double>{const_cast<double2*>(src), n}, MatView<float>{c32b, n});
    chain_apply_T<float>(*a.t32, *c32, *g, p, c32b, l32, tmp);
    convert_mat<float, double>(n, 8 * p, MatView<float>{l32, n}, MatView<double>{dst, n});
  };
  precond(c, lam);
  const double cn = norm2(c, len);
  const double2 one = {1, 0}, zero = {0, 0};
  const double k0 = 2 * M_PI / enei;
  const dim3 gq((unsigned)cdiv_up(n, 256), (unsigned)p), g4((unsigned)cdiv_up(n, 256), (unsigned)(4 * p));
  for (int it = 0;; ++it) {
    hipLaunchKernelGGL(k_adj_uw, gq, dim3(256), 0, stream(), n, p, lam, U1, W1, U2, W2, g->geo.nvec, k0, eps1, eps2);
    MNP_HIP(hipGetLastError());
    gemm<double>(n, 4 * p, n, one, MatView<double>{a.G1T, n}, MatView<double>{U1, n}, zero, MatView<double>{P1, n});
    gemm<double>(n, 4 * p, n, one, MatView<double>{a.H1T, n}, MatView<double>{W1, n}, zero, MatView<double>{Q1, n});
    gemm<double>(n, 4 * p, n, one, MatView<double>{a.G2T, n}, MatView<double>{U2, n}, zero, MatView<double>{P2, n});
    gemm<double>(n, 4 * p, n, one, MatView<double>{a.H2T, n}, MatView<double>{W2, n}, zero, MatView<double>{Q2, n});
    hipLaunchKernelGGL(k_adj_res, g4, dim3(256), 0, stream(), n, p, c, P1, Q1, P2, Q2, rr);
    MNP_HIP(hipGetLastError());
    const double rn = norm2(rr, len);
    last_relres = cn > 0 ? rn / cn : 0;
    last_iter = it;
    if (!(last_relres > tol) || it >= maxit) break;
    precond(rr, P1);  // correction into P1 .. (n x 8p fits in P1, Q1)
    hipLaunchKernelGGL(k_adj_axpy, dim3((unsigned)cdiv_up(len, 256)), dim3(256), 0, stream(), len, lam, P1);
    MNP_HIP(hipGetLastError());
  }
  MNP_CHECK(!(last_relres > tol), MNP_ERR_NOCONV,
            "bemret_solve_transposed: refinement stalled at relative residual " + sci(last_relres));
}


// ------------------------------------------------------ extinction gradient
// All gradient kernels accumulate real numbers: the imaginary parts of the
// contributions to dJ (J = sum_q w_q pol_q . e_q(dir_q), the weights folded
// into c and hence into lam); the vertex gather scales by 4 pi / Re k.

__device__ inline double2 ikc(double2 k) { return mk<double>(-k.y, k.x); }  // i k
__device__ inline double imag_of(double2 a, double2 b) { return a.x * b.y + a.y * b.x; }  // Im(a b)

// E_m(z), E_{m-1}(z), E_{m-2}(z) of E_j(z) = sum_{t=0}^{j} z^t / t!  (E_{-1} = E_{-2} = 0)
__device__ inline void expsums(double2 z, int m, double2& Em, double2& Em1, double2& Em2) {
  double2 t = mk<double>(1, 0), s = mk<double>(1, 0);
  double2 s1 = mk<double>(0, 0), s2 = mk<double>(0, 0);
  for (int j = 1; j <= m; ++j) {
    s2 = s1;
    s1 = s;
    t = cscale<double>(cmul<double>(t, z), 1.0 / j);
    s = cadd<double>(s, t);
  }
  Em = s;
  Em1 = m >= 1 ? s1 : mk<double>(0, 0);
  Em2 = m >= 2 ? s2 : mk<double>(0, 0);
}

// Unrefined pairs (eval1.m:24-40 and :110-124): per pair (r, c), with R =
// pos_r - pos_c, d = |R|, for (k, WG, WH) in {(k1, WG1, WH1), (k2, WG2, WH2)}:
//   G = a_c e^{ikd}/d,  F = (n_r . R) f a_c,  f = (ik - 1/d) e^{ikd}/d^2,
//   dG/dR = G (ik - 1/d) R/d,  dF/dR = f a_c n_r + (n_r . R) f' a_c R/d,
//   f' = e^{ikd}/d^2 (3/d^2 - 3ik/d - k^2),  dG/da_c = G/a_c, dF/dn_r = f a_c R,
// and the static sum-rule term T = area_r F0(r, c)/area_c = -area_r n_r.R/d^3
// with weight Im(Omega_c), Omega_c = WH1_cc + WH2_cc.
// Grid: x = tiles of 64 rows (one per lane), y = tiles of 256 columns (64
// per wave).  Column terms are reduced over the lanes per column and
// written per row tile; row terms are combined over the four waves and
// written per column tile.
constexpr int ADJ_CT = 256;
constexpr int ADJ_MAXP = 8;  // polarisations per gradient call
__global__ void __launch_bounds__(256) k_adj_bulk(int64_t n, Geo geo, double rel, double absc,
                                                  const double2* __restrict__ WG1, const double2* __restrict__ WH1,
                                                  const double2* __restrict__ WG2, const double2* __restrict__ WH2,
                                                  double2 k1, double2 k2, int64_t nrt, int64_t nct,
                                                  double* __restrict__ rowpart, double* __restrict__ colpart) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t r = blockIdx.x * 64 + lane;
  const int64_t c0 = blockIdx.y * (int64_t)ADJ_CT + wave * 64;
  const bool rv = r < n;
  double pr[3] = {0, 0, 0}, nr[3] = {0, 0, 0}, ar = 0;
  if (rv)
    for (int q = 0; q < 3; ++q) {
      pr[q] = geo.pos[q * n + r];
      nr[q] = geo.nvec[q * n + r];
    }
  if (rv) ar = geo.area[r];
  double racc[7] = {0, 0, 0, 0, 0, 0, 0};  // gpos_r (3), gn_r (3), garea_r
  const int64_t c1 = c0 + 64 < n ? c0 + 64 : n;
  for (int64_t c = c0; c < c1; ++c) {
    double cacc[4] = {0, 0, 0, 0};  // gpos_c (3), garea_c
    if (rv && refine_kind(geo, n, r, c, rel, absc) == 0) {
      const double R[3] = {pr[0] - geo.pos[c], pr[1] - geo.pos[n + c], pr[2] - geo.pos[2 * n + c]};
      const double d = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
      const double id = 1.0 / d, id2 = id * id;
      const double ac = geo.area[c];
      const double nR = nr[0] * R[0] + nr[1] * R[1] + nr[2] * R[2];
      const int64_t e = c * n + r;
      double sR = 0, sN = 0, sA = 0;  // -Im of the coefficients of R/d, of n_r, of 1/a_c
      for (int kk = 0; kk < 2; ++kk) {
        const double2 k = kk ? k2 : k1;
        const double2 WG = kk ? WG2[e] : WG1[e], WH = kk ? WH2[e] : WH1[e];
        const double2 ph = cexp_ikd(k, d);
        const double2 ik = ikc(k);
        const double2 G = cscale<double>(ph, ac * id);
        const double2 ikd = mk<double>(ik.x - id, ik.y);  // ik - 1/d
        const double2 f = cscale<double>(cmul<double>(ikd, ph), id2);
        const double2 k2c = cmul<double>(k, k);
        const double2 fpc = mk<double>(3 * id2 - 3 * id * ik.x - k2c.x, -3 * id * ik.y - k2c.y);
        const double2 fp = cscale<double>(cmul<double>(ph, fpc), id2);
        // R/d coefficient: WG G (ik - 1/d) + WH (n.R) f' a_c
        sR -= imag_of(WG, cmul<double>(G, ikd)) + imag_of(WH, cscale<double>(fp, nR * ac));
        sN -= imag_of(WH, cscale<double>(f, ac));  // n_r coefficient (and R coefficient of gn_r)
        sA -= imag_of(WG, cscale<double>(ph, id)) + imag_of(WH, cscale<double>(f, nR));
      }
      const double2 Om = cadd<double>(WH1[c * n + c], WH2[c * n + c]);
      const double io = Om.y;
      const double id3 = id2 * id, id5 = id3 * id2;
      for (int q = 0; q < 3; ++q) {
        const double dT = -ar * (nr[q] * id3 - 3 * nR * R[q] * id5);
        const double V = sR * R[q] * id + sN * nr[q] + io * dT;
        racc[q] += V;
        cacc[q] -= V;
        racc[3 + q] += sN * R[q] - io * ar * R[q] * id3;
      }
      racc[6] += -io * nR * id3;
      cacc[3] += sA;
    }
    for (int q = 0; q < 4; ++q) {
      double v = cacc[q];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0) colpart[((int64_t)q * nrt + blockIdx.x) * n + c] = v;
    }
  }
  __shared__ double sh[4][7][64];
  for (int q = 0; q < 7; ++q) sh[wave][q][lane] = racc[q];
  __syncthreads();
  if (wave == 0 && rv)
    for (int q = 0; q < 7; ++q)
      rowpart[((int64_t)q * nct + blockIdx.y) * n + r] = sh[0][q][lane] + sh[1][q][lane] + sh[2][q][lane] + sh[3][q][lane];
}

// Refined pairs, one thread each.  Near field (kind 1, init.m:116-176 as
// evaluated by eval1.m): over the 28 points y_q of source face c with
// weights w_q = wbar_q qarea_c, xv = pos_r - y_q, r_q = |xv|, r0 = |pos_r - pos_c|:
//   G = e^{ik r0} sum_q w_q E_m(ik(r_q - r0)) / r_q,
//   F = e^{ik r0} sum_q w_q (n_r . xv) [-E_m / r_q^3 + ik E_{m-1} / r_q^2]
// (the Taylor tables summed in closed form), differentiated in r_q, r0,
// xv, n_r and qarea_c, and the refined sum-rule term
// T = -area_r sum_q wbar_q n_r . xv / r_q^3 with weight Im(Omega_c).
// Self terms (kind 2, r == c, init.m:44-95): G = sum_q w_q E_m(ik r_q) / r_q
// over the polar points with w_q = wbar_q area_c; F_cc is -2 pi - fsum_c
// (the normal part of the polar f vanishes for a flat face), whose
// derivative the sum-rule terms of the other pairs carry.
// prow[7][nnz]: gpos_r, gn_r, garea_r; pcol[10][nnz]: vertices of face c
// (a * 3 + coordinate) and garea_c.
__global__ void __launch_bounds__(128) k_adj_refined(int64_t nnz, int64_t n, Geo geo, Pairs pr, int order,
                                                     const double* __restrict__ q18, const double* __restrict__ qpol,
                                                     int npolq, const double2* __restrict__ WG1,
                                                     const double2* __restrict__ WH1, const double2* __restrict__ WG2,
                                                     const double2* __restrict__ WH2, double2 k1, double2 k2,
                                                     double* __restrict__ prow, double* __restrict__ pcol, int* flag) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= nnz) return;
  const int64_t r = pr.row[p], c = pr.col[p];
  double gpos[3] = {0, 0, 0}, gn[3] = {0, 0, 0}, gar = 0, gvert[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, gac = 0;
  double vc[3][3], pc[3], prr[3], nr[3];
  for (int q = 0; q < 3; ++q) {
    pc[q] = geo.pos[q * n + c];
    prr[q] = geo.pos[q * n + r];
    nr[q] = geo.nvec[q * n + r];
    for (int a = 0; a < 3; ++a) vc[a][q] = geo.vtx[(a * 3 + q) * n + c];
  }
  const int64_t e = c * n + r;
  const double2 WG[2] = {WG1[e], WG2[e]}, WH[2] = {WH1[e], WH2[e]};
  const double2 ks[2] = {k1, k2};
  if (pr.kind[p] == 2) {
    if (r != c) {
      atomicOr(flag, 1);  // coincident centroids of distinct faces: not differentiated
      return;
    }
    const double ac = geo.area[c];
    double2 sA = mk<double>(0, 0);
    for (int q = 0; q < npolq; ++q) {
      const double x = qpol[q], y = qpol[MNP_MAX_POLAR + q], z = 1 - x - y, w = qpol[2 * MNP_MAX_POLAR + q] * ac;
      const double bary[3] = {x, y, z};
      double xv[3];
      for (int t = 0; t < 3; ++t) xv[t] = pc[t] - (x * vc[0][t] + y * vc[1][t] + z * vc[2][t]);
      const double rq = sqrt(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]);
      double al = 0;
      for (int kk = 0; kk < 2; ++kk) {
        const double2 k = ks[kk], ik = ikc(k);
        double2 Em, Em1, Em2;
        expsums(cscale<double>(ik, rq), order, Em, Em1, Em2);
        const double2 phe = cexp_ikd(k, 2.220446049250313e-16);
        const double2 dG = cscale<double>(cmul<double>(phe, csub<double>(cscale<double>(cmul<double>(ik, Em1), 1.0 / rq),
                                                                        cscale<double>(Em, 1.0 / (rq * rq)))),
                                          w);
        al -= imag_of(WG[kk], dG) / rq;
        sA = cadd<double>(sA, cmul<double>(WG[kk], cscale<double>(cmul<double>(phe, Em), w / (rq * ac))));
      }
      for (int a = 0; a < 3; ++a)
        for (int t = 0; t < 3; ++t) gvert[a][t] += (1.0 / 3 - bary[a]) * al * xv[t];
    }
    gac = -sA.y;
  } else {
    const double ar = geo.area[r], qa = geo.qarea[c];
    const double R0[3] = {prr[0] - pc[0], prr[1] - pc[1], prr[2] - pc[2]};
    const double r0 = sqrt(R0[0] * R0[0] + R0[1] * R0[1] + R0[2] * R0[2]);
    const double2 Om = cadd<double>(WH1[c * n + c], WH2[c * n + c]);
    const double io = Om.y;
    double2 ph[2];
    for (int kk = 0; kk < 2; ++kk) ph[kk] = cexp_ikd(ks[kk], r0);
    double2 s0 = mk<double>(0, 0), sA = mk<double>(0, 0);
    for (int q = 0; q < 28; ++q) {
      const double x = q18[q], y = q18[28 + q], z = 1 - x - y, wb = q18[56 + q], w = wb * qa;
      const double bary[3] = {x, y, z};
      double xv[3];
      for (int t = 0; t < 3; ++t) xv[t] = prr[t] - (x * vc[0][t] + y * vc[1][t] + z * vc[2][t]);
      const double rq = sqrt(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]);
      const double ir = 1.0 / rq, ir2 = ir * ir, ir3 = ir2 * ir;
      const double nx = nr[0] * xv[0] + nr[1] * xv[1] + nr[2] * xv[2];
      double al = 0, be = 0;  // coefficients of xv and n_r
      for (int kk = 0; kk < 2; ++kk) {
        const double2 k = ks[kk], ik = ikc(k), kq = cmul<double>(k, k);
        double2 Em, Em1, Em2;
        expsums(cscale<double>(ik, rq - r0), order, Em, Em1, Em2);
        const double2 pw = cscale<double>(ph[kk], w);
        const double2 ikE1 = cmul<double>(ik, Em1);
        // dG/dr_q, U, U_r, U_r0
        const double2 dG = cmul<double>(pw, csub<double>(cscale<double>(ikE1, ir), cscale<double>(Em, ir2)));
        const double2 U = csub<double>(cscale<double>(ikE1, ir2), cscale<double>(Em, ir3));
        const double2 Ur = csub<double>(csub<double>(cscale<double>(Em, 3 * ir3 * ir), cscale<double>(ikE1, 3 * ir3)),
                                        cscale<double>(cmul<double>(kq, Em2), ir2));
        const double2 Ur0 = cadd<double>(cscale<double>(ikE1, ir3), cscale<double>(cmul<double>(kq, Em2), ir2));
        const double2 Fq = cscale<double>(cmul<double>(pw, U), nx);
        al -= (imag_of(WG[kk], dG) + imag_of(WH[kk], cscale<double>(cmul<double>(pw, Ur), nx))) * ir;
        be -= imag_of(WH[kk], cmul<double>(pw, U));
        const double2 dGr0 = cscale<double>(cmul<double>(cmul<double>(ik, pw), csub<double>(Em, Em1)), ir);
        const double2 dFr0 = cadd<double>(cmul<double>(ik, Fq), cscale<double>(cmul<double>(pw, Ur0), nx));
        s0 = cadd<double>(s0, cadd<double>(cmul<double>(WG[kk], dGr0), cmul<double>(WH[kk], dFr0)));
        sA = cadd<double>(sA, cscale<double>(cadd<double>(cmul<double>(WG[kk], cscale<double>(cmul<double>(pw, Em), ir)),
                                                          cmul<double>(WH[kk], Fq)),
                                             1.0 / qa));
      }
      // sum-rule term of this point: T_q = -area_r wbar n_r.xv / r^3
      const double ir5 = ir3 * ir2;
      for (int t = 0; t < 3; ++t) {
        const double dT = -ar * wb * (nr[t] * ir3 - 3 * nx * xv[t] * ir5);
        const double cx = al * xv[t] + be * nr[t] + io * dT;
        gpos[t] += cx;
        for (int a = 0; a < 3; ++a) gvert[a][t] -= bary[a] * cx;
        gn[t] += be * xv[t] - io * ar * wb * xv[t] * ir3;
      }
      gar += io * (-wb * nx * ir3);  // d T_q / d area_r
    }
    const double c0 = -s0.y;
    for (int t = 0; t < 3; ++t) {
      const double v = c0 * R0[t] / r0;
      gpos[t] += v;
      for (int a = 0; a < 3; ++a) gvert[a][t] -= v / 3;
    }
    gac = -sA.y;
  }
  for (int t = 0; t < 3; ++t) {
    prow[t * nnz + p] = gpos[t];
    prow[(3 + t) * nnz + p] = gn[t];
  }
  prow[6 * nnz + p] = gar;
  for (int a = 0; a < 3; ++a)
    for (int t = 0; t < 3; ++t) pcol[(a * 3 + t) * nnz + p] = gvert[a][t];
  pcol[9 * nnz + p] = gac;
}


struct PolDirW {
  double pol[ADJ_MAXP][3], dir[ADJ_MAXP][3], w[ADJ_MAXP];
};

// Terms local to face i (one thread each), loc[7][n] = gpos, gn, garea:
//   far field (c):  J_iq = w_q i k0 e^{-ik dir.pos_i} area_i pol . h2_iq,
//                   dJ/dpos = -ik dir J, dJ/darea = J / area;
//   plane wave (b): lam^T b_i = e^{ik pos.dir}/(i k0) [la.pol - i k0 e2 lDe n.pol
//                   + ik (n.dir) lal.pol]  (potential.m:72-76, excitation.m);
//   normals in A:   -lam^T dA x through n_i in the De and alpha rows of afun.m:
//                   d/dn_c: i k0 [lDe (e1 G1h1_c - e2 G2h2_c) + lal_c (e1 G1s1 - e2 G2s2)].
__global__ void k_adj_local(int64_t n, int64_t p, const double2* __restrict__ lam, const double2* __restrict__ x,
                            const double2* __restrict__ GX1, const double2* __restrict__ GX2, Geo geo, PolDirW pd,
                            double k0, double2 k, double2 e1, double2 e2, double* __restrict__ loc) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double pos[3], nv[3];
  for (int t = 0; t < 3; ++t) {
    pos[t] = geo.pos[t * n + i];
    nv[t] = geo.nvec[t * n + i];
  }
  const double area = geo.area[i];
  double gpos[3] = {0, 0, 0}, gn[3] = {0, 0, 0}, garea = 0;
  const double2 ik = ikc(k), ik0 = mk<double>(0, k0);
  const double2 ik0e2 = cmul<double>(ik0, e2);
  for (int64_t q = 0; q < p; ++q) {
    const double* pol = pd.pol[q];
    const double* dir = pd.dir[q];
    double pdd = 0, npol = 0, ndir = 0;
    for (int t = 0; t < 3; ++t) {
      pdd += pos[t] * dir[t];
      npol += nv[t] * pol[t];
      ndir += nv[t] * dir[t];
    }
    // far field
    double2 hp = mk<double>(0, 0), lap = mk<double>(0, 0), lalp = mk<double>(0, 0);
    for (int t = 0; t < 3; ++t) {
      hp = cadd<double>(hp, cscale<double>(x[(5 * p + t + 3 * q) * n + i], pol[t]));
      lap = cadd<double>(lap, cscale<double>(lam[(p + t + 3 * q) * n + i], pol[t]));
      lalp = cadd<double>(lalp, cscale<double>(lam[(5 * p + t + 3 * q) * n + i], pol[t]));
    }
    const double2 phf = cscale<double>(cexp_ikd(mk<double>(-k.x, -k.y), pdd), area);
    const double2 J = cscale<double>(cmul<double>(ik0, cmul<double>(phf, hp)), pd.w[q]);
    const double2 mikJ = cmul<double>(mk<double>(-ik.x, -ik.y), J);
    for (int t = 0; t < 3; ++t) gpos[t] += mikJ.y * dir[t];
    garea += J.y / area;
    // plane wave
    const double2 phe = cdiv<double>(cexp_ikd(k, pdd), ik0);
    const double2 lde = lam[(4 * p + q) * n + i];
    const double2 Q = cadd<double>(csub<double>(lap, cscale<double>(cmul<double>(ik0e2, lde), npol)),
                                   cscale<double>(cmul<double>(ik, lalp), ndir));
    const double gq = imag_of(cmul<double>(ik, phe), Q);
    const double an = -imag_of(phe, cmul<double>(ik0e2, lde));  // coefficient of pol in gn
    const double bn = imag_of(phe, cmul<double>(ik, lalp));     // coefficient of dir in gn
    for (int t = 0; t < 3; ++t) {
      gpos[t] += gq * dir[t];
      gn[t] += an * pol[t] + bn * dir[t];
    }
    // normals of the operator
    const double2 gs1 = GX1[q * n + i], gs2 = GX2[q * n + i];
    const double2 ws = csub<double>(cmul<double>(e1, gs1), cmul<double>(e2, gs2));
    for (int t = 0; t < 3; ++t) {
      const int64_t col = p + t + 3 * q;
      const double2 wh = csub<double>(cmul<double>(e1, GX1[col * n + i]), cmul<double>(e2, GX2[col * n + i]));
      const double2 v = cadd<double>(cmul<double>(lde, wh), cmul<double>(lam[(5 * p + t + 3 * q) * n + i], ws));
      gn[t] += imag_of(ik0, v);
    }
  }
  for (int t = 0; t < 3; ++t) {
    loc[t * n + i] = gpos[t];
    loc[(3 + t) * n + i] = gn[t];
  }
  loc[6 * n + i] = garea;
}

// Per face: the pair, refined and local terms summed, then carried to its
// three vertices through pos = (v1 + v2 + v3)/3, nvec = N/|N|, area = |N|/2
// with N = (v1 - v2) x (v3 - v2) (norm_flat.m); fg[9][n] (vertex a, coordinate t).
__global__ void k_adj_face(int64_t n, int64_t nnz, Geo geo, const int64_t* __restrict__ off,
                           const int64_t* __restrict__ rptr, const int32_t* __restrict__ rlist, int64_t nrt,
                           int64_t nct, const double* __restrict__ rowpart, const double* __restrict__ colpart,
                           const double* __restrict__ prow, const double* __restrict__ pcol,
                           const double* __restrict__ loc, double* __restrict__ fg) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double gp[3], gn[3], ga, gv[9];
  for (int t = 0; t < 3; ++t) {
    gp[t] = loc[t * n + i];
    gn[t] = loc[(3 + t) * n + i];
  }
  ga = loc[6 * n + i];
  for (int64_t b = 0; b < nct; ++b) {
    for (int t = 0; t < 3; ++t) {
      gp[t] += rowpart[(t * nct + b) * n + i];
      gn[t] += rowpart[((3 + t) * nct + b) * n + i];
    }
    ga += rowpart[(6 * nct + b) * n + i];
  }
  for (int64_t b = 0; b < nrt; ++b) {
    for (int t = 0; t < 3; ++t) gp[t] += colpart[(t * nrt + b) * n + i];
    ga += colpart[(3 * nrt + b) * n + i];
  }
  for (int t = 0; t < 9; ++t) gv[t] = 0;
  for (int64_t pp = off[i]; pp < off[i + 1]; ++pp) {  // pairs with source face i
    for (int t = 0; t < 9; ++t) gv[t] += pcol[t * nnz + pp];
    ga += pcol[9 * nnz + pp];
  }
  for (int64_t s = rptr[i]; s < rptr[i + 1]; ++s) {  // pairs with observation face i
    const int64_t pp = rlist[s];
    for (int t = 0; t < 3; ++t) {
      gp[t] += prow[t * nnz + pp];
      gn[t] += prow[(3 + t) * nnz + pp];
    }
    ga += prow[6 * nnz + pp];
  }
  double v[3][3], nu[3];
  for (int a = 0; a < 3; ++a)
    for (int t = 0; t < 3; ++t) v[a][t] = geo.vtx[(a * 3 + t) * n + i];
  for (int t = 0; t < 3; ++t) nu[t] = geo.nvec[t * n + i];
  const double Nn = 2 * geo.area[i];
  const double gnn = gn[0] * nu[0] + gn[1] * nu[1] + gn[2] * nu[2];
  double cv[3], e1[3], e2[3];
  for (int t = 0; t < 3; ++t) {
    cv[t] = (gn[t] - gnn * nu[t]) / Nn + 0.5 * ga * nu[t];
    e1[t] = v[0][t] - v[1][t];
    e2[t] = v[2][t] - v[1][t];
  }
  const double d1[3] = {e2[1] * cv[2] - e2[2] * cv[1], e2[2] * cv[0] - e2[0] * cv[2], e2[0] * cv[1] - e2[1] * cv[0]};
  const double d3[3] = {cv[1] * e1[2] - cv[2] * e1[1], cv[2] * e1[0] - cv[0] * e1[2], cv[0] * e1[1] - cv[1] * e1[0]};
  for (int t = 0; t < 3; ++t) {
    const double pc = gp[t] / 3;
    fg[(0 * 3 + t) * n + i] = gv[0 * 3 + t] + pc + d1[t];
    fg[(1 * 3 + t) * n + i] = gv[1 * 3 + t] + pc - d1[t] - d3[t];
    fg[(2 * 3 + t) * n + i] = gv[2 * 3 + t] + pc + d3[t];
  }
}

// dverts[v + t nv] = scale sum over the (face, corner) of vertex v of fg.
__global__ void k_adj_vertex(int64_t nv, int64_t n, const int64_t* __restrict__ vptr,
                             const int64_t* __restrict__ vlist, const double* __restrict__ fg, double scale,
                             double* __restrict__ dv) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= nv) return;
  double s[3] = {0, 0, 0};
  for (int64_t k = vptr[v]; k < vptr[v + 1]; ++k) {
    const int64_t f = vlist[k] / 3, a = vlist[k] % 3;
    for (int t = 0; t < 3; ++t) s[t] += fg[(a * 3 + t) * n + f];
  }
  for (int t = 0; t < 3; ++t) dv[t * nv + v] = scale * s[t];
}

// c of the weighted extinction (layout of x): only the h2 rows,
// c = w_q i k0 e^{-ik dir_q.pos_j} area_j pol_q.
__global__ void k_adj_c(int64_t n, int64_t p, Geo geo, PolDirW pd, double k0, double2 k, double2* c) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t col = blockIdx.y;  // 0 .. 8p
  if (i >= n) return;
  double2 v = mk<double>(0, 0);
  if (col >= 5 * p) {
    const int64_t q = (col - 5 * p) / 3, t = (col - 5 * p) % 3;
    double pdd = 0;
    for (int u = 0; u < 3; ++u) pdd += geo.pos[u * n + i] * pd.dir[q][u];
    const double2 ph = cscale<double>(cexp_ikd(mk<double>(-k.x, -k.y), pdd), geo.area[i] * pd.w[q] * pd.pol[q][t]);
    v = cmul<double>(mk<double>(0, k0), ph);
  }
  c[col * n + i] = v;
}

// sum_i a_i b_i (complex, unconjugated), partial sums per workgroup
__global__ void __launch_bounds__(256) k_adj_dot(int64_t len, const double2* a, const double2* b, double2* part) {
  __shared__ double2 red[256];
  double2 s = mk<double>(0, 0);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
    s = cadd<double>(s, cmul<double>(a[i], b[i]));
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] = cadd<double>(red[threadIdx.x], red[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Gradient of sum_q w_q ext_q with respect to the vertices (host nv x 3)
// and to eps1 (deps1 = {d/dRe eps1, d/dIm eps1}) at the current wavelength,
// from the solution x of the plane wave (pol, dir, host p x 3) from outside.
static void extinction_gradient(BemRet& b, int64_t p, const double* pol, const double* dir, const double* w,
                                const double2* sig1, const double2* sig2, const double2* h1, const double2* h2,
                                double* dverts, double* deps1) {
  const GreenRet& g = *b.g;
  const int64_t n = b.n, nn = n * n, nnz = g.nnz, nv = g.nverts;
  MNP_CHECK(b.ready, MNP_ERR_STATE, "extinction_gradient: initmat not called");
  MNP_CHECK(p >= 1 && p <= ADJ_MAXP, MNP_ERR_ARG, "extinction_gradient: npol must be in 1..8");
  MNP_CHECK(n <= 65535, MNP_ERR_ARG, "extinction_gradient: at most 65535 faces");
  PolDirW pd{};
  for (int64_t q = 0; q < p; ++q) {
    double s = 0;
    for (int t = 0; t < 3; ++t) {
      pd.pol[q][t] = pol[t * p + q];
      pd.dir[q][t] = dir[t * p + q];
      s += pd.pol[q][t] * pd.dir[q][t];
    }
    MNP_CHECK(s == 0, MNP_ERR_ARG, "extinction_gradient: polarisation not perpendicular to direction");
    pd.w[q] = w ? w[q] : 1.0;
  }
  if (!b.adj) b.adj = new Adjoint(n);
  Adjoint& a = *b.adj;
  const double k0 = 2 * M_PI / b.enei;
  const double2 k1 = wavenumber(k0, b.eps1), k2 = wavenumber(k0, b.eps2);
  const int64_t len = 8 * p * n;
  const dim3 gq((unsigned)cdiv_up(n, 256), (unsigned)p), g8((unsigned)cdiv_up(n, 256), (unsigned)(8 * p));
  // forward solution, functional, adjoint solution
  double2* x = (double2*)a.grad.get(0, sizeof(double2) * len);
  double2* c = (double2*)a.grad.get(1, sizeof(double2) * len);
  double2* lam = (double2*)a.grad.get(2, sizeof(double2) * len);
  launch_pack(n, p, sig1, h1, sig2, h2, x);
  hipLaunchKernelGGL(k_adj_c, g8, dim3(256), 0, stream(), n, p, g.geo, pd, k0, k2, c);
  MNP_HIP(hipGetLastError());
  b.solve_transposed(p, c, lam);
  // the FP64 operator: kept by the mixed solver, assembled again for the FP64 one
  const double2 *G1m = b.G1, *G2m = b.G2, *H1m = b.H1;
  if (!G1m) {
    double2* t1 = (double2*)a.big.get(0, sizeof(double2) * nn);
    double2* t2 = (double2*)a.big.get(1, sizeof(double2) * nn);
    double2* t3 = (double2*)a.big.get(2, sizeof(double2) * nn);
    g.assemble(k1, 2 * M_PI, OutPair{t1, nullptr}, OutPair{t3, nullptr}, n);
    g.assemble(k2, -2 * M_PI, OutPair{t2, nullptr}, OutPair{}, n);
    G1m = t1;
    G2m = t2;
    H1m = t3;
  }
  const double2 one = {1, 0}, zero = {0, 0}, mone = {-1, 0};
  double2* prod = (double2*)a.grad.get(3, sizeof(double2) * (size_t)(36 * p * n));
  double2 *GX1 = prod, *GX2 = GX1 + 4 * p * n, *HX1 = GX2 + 4 * p * n, *dGX = HX1 + 4 * p * n,
          *dHX = dGX + 4 * p * n, *reps = dHX + 4 * p * n;  // reps: 8p
  double2 *X1T = reps + 8 * p * n, *X2T = X1T + 4 * p * n;  // 4p x n each
  const MatView<double> X1{x, n}, X2{x + 4 * p * n, n};
  gemm<double>(n, 4 * p, n, one, MatView<double>{const_cast<double2*>(G1m), n}, X1, zero, MatView<double>{GX1, n});
  gemm<double>(n, 4 * p, n, one, MatView<double>{const_cast<double2*>(G2m), n}, X2, zero, MatView<double>{GX2, n});
  // d/d eps1: D = lam^T r with r = -(dA/deps1) x (dk1/deps1 = k1 / (2 eps1))
  if (deps1) {
    double2* dG = (double2*)a.big.get(3, sizeof(double2) * nn);
    double2* dH = (double2*)a.big.get(4, sizeof(double2) * nn);
    g.assemble_dk(k1, OutPair{dG, nullptr}, OutPair{dH, nullptr}, n);
    const double2 dk = cdiv<double>(k1, cscale<double>(b.eps1, 2.0));
    gemm<double>(n, 4 * p, n, one, MatView<double>{const_cast<double2*>(H1m), n}, X1, zero, MatView<double>{HX1, n});
    gemm<double>(n, 4 * p, n, dk, MatView<double>{dG, n}, X1, zero, MatView<double>{dGX, n});
    gemm<double>(n, 4 * p, n, dk, MatView<double>{dH, n}, X1, zero, MatView<double>{dHX, n});
    launch_deps1(n, p, GX1, HX1, dGX, dHX, reps, g.geo.nvec, k0, b.eps1);
    const int nb = 256;
    double2* part = (double2*)a.grad.get(4, sizeof(double2) * nb);
    hipLaunchKernelGGL(k_adj_dot, dim3(nb), dim3(256), 0, stream(), len, lam, reps, part);
    MNP_HIP(hipGetLastError());
    std::vector<double2> hp(nb);
    MNP_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double2) * nb, hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipStreamSynchronize(stream()));
    double2 D = mk<double>(0, 0);
    for (auto& v : hp) D = cadd<double>(D, v);
    const double sc = 4 * M_PI / k2.x;
    deps1[0] = sc * D.y;  // d/dRe eps1 = Im D
    deps1[1] = sc * D.x;  // d/dIm eps1 = Im(i D) = Re D
  }
  if (!dverts) return;
  // element weights W = u x^T of the four matrices (n x n each)
  double2* U = (double2*)a.grad.get(5, sizeof(double2) * (size_t)(16 * p * n));
  double2 *U1 = U, *W1 = U1 + 4 * p * n, *U2 = W1 + 4 * p * n, *W2 = U2 + 4 * p * n;
  hipLaunchKernelGGL(k_adj_uw, gq, dim3(256), 0, stream(), n, p, lam, U1, W1, U2, W2, g.geo.nvec, k0, b.eps1, b.eps2);
  MNP_HIP(hipGetLastError());
  transpose<double>(n, 4 * p, x, n, X1T, 4 * p);
  transpose<double>(n, 4 * p, x + 4 * p * n, n, X2T, 4 * p);
  double2* WG1 = (double2*)a.big.get(5, sizeof(double2) * nn);
  double2* WH1 = (double2*)a.big.get(6, sizeof(double2) * nn);
  double2* WG2 = (double2*)a.big.get(7, sizeof(double2) * nn);
  double2* WH2 = (double2*)a.res.get(1, sizeof(double2) * nn);
  gemm<double>(n, n, 4 * p, one, MatView<double>{U1, n}, MatView<double>{X1T, 4 * p}, zero, MatView<double>{WG1, n});
  gemm<double>(n, n, 4 * p, one, MatView<double>{W1, n}, MatView<double>{X1T, 4 * p}, zero, MatView<double>{WH1, n});
  gemm<double>(n, n, 4 * p, mone, MatView<double>{U2, n}, MatView<double>{X2T, 4 * p}, zero, MatView<double>{WG2, n});
  gemm<double>(n, n, 4 * p, mone, MatView<double>{W2, n}, MatView<double>{X2T, 4 * p}, zero, MatView<double>{WH2, n});
  // index lists, once per mesh
  if (!a.rptr) {
    std::vector<int32_t> rows(nnz);
    if (nnz) MNP_HIP(hipMemcpy(rows.data(), g.pairs.row, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
    std::vector<int64_t> rp(n + 1, 0);
    for (int32_t r : rows) ++rp[r + 1];
    for (int64_t i = 0; i < n; ++i) rp[i + 1] += rp[i];
    std::vector<int32_t> rl(nnz > 0 ? nnz : 1);
    std::vector<int64_t> fill(rp.begin(), rp.end() - 1);
    for (int64_t q = 0; q < nnz; ++q) rl[fill[rows[q]]++] = (int32_t)q;
    std::vector<int64_t> vp(nv + 1, 0);
    for (int64_t t = 0; t < 3 * n; ++t) ++vp[g.faces_h[t] + 1];
    for (int64_t v = 0; v < nv; ++v) vp[v + 1] += vp[v];
    std::vector<int64_t> vl(3 * n), vf(vp.begin(), vp.end() - 1);
    for (int64_t corner = 0; corner < 3; ++corner)
      for (int64_t f = 0; f < n; ++f) vl[vf[g.faces_h[corner * n + f]]++] = f * 3 + corner;
    a.rptr = dev_alloc<int64_t>(n + 1);
    a.rlist = dev_alloc<int32_t>((int64_t)rl.size());
    a.vptr = dev_alloc<int64_t>(nv + 1);
    a.vlist = dev_alloc<int64_t>(3 * n);
    a.flag = dev_alloc<int>(1);
    MNP_HIP(hipMemcpy(a.rptr, rp.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
    MNP_HIP(hipMemcpy(a.rlist, rl.data(), sizeof(int32_t) * rl.size(), hipMemcpyHostToDevice));
    MNP_HIP(hipMemcpy(a.vptr, vp.data(), sizeof(int64_t) * (nv + 1), hipMemcpyHostToDevice));
    MNP_HIP(hipMemcpy(a.vlist, vl.data(), sizeof(int64_t) * 3 * n, hipMemcpyHostToDevice));
  }
  const int64_t nrt = cdiv_up(n, 64), nct = cdiv_up(n, ADJ_CT);
  double* rowpart = (double*)a.grad.get(6, sizeof(double) * (size_t)(7 * nct * n));
  double* colpart = (double*)a.grad.get(7, sizeof(double) * (size_t)(4 * nrt * n));
  double* prow = (double*)a.res.get(2, sizeof(double) * (size_t)(7 * (nnz > 0 ? nnz : 1)));
  double* pcol = (double*)a.res.get(3, sizeof(double) * (size_t)(10 * (nnz > 0 ? nnz : 1)));
  double* loc = (double*)a.res.get(4, sizeof(double) * (size_t)(7 * n));
  double* fg = (double*)a.res.get(5, sizeof(double) * (size_t)(9 * n));
  double* dv = (double*)a.res.get(6, sizeof(double) * (size_t)(3 * nv));
  hipLaunchKernelGGL(k_adj_bulk, dim3((unsigned)nrt, (unsigned)nct), dim3(256), 0, stream(), n, g.geo,
                     g.op.rel_cutoff, g.op.abs_cutoff, WG1, WH1, WG2, WH2, k1, k2, nrt, nct, rowpart, colpart);
  MNP_HIP(hipGetLastError());
  if (nnz) {
    MNP_HIP(hipMemsetAsync(a.flag, 0, sizeof(int), stream()));
    hipLaunchKernelGGL(k_adj_refined, dim3((unsigned)cdiv_up(nnz, 128)), dim3(128), 0, stream(), nnz, n, g.geo,
                       g.pairs, g.order, g.q18, g.qpol, g.npolq, WG1, WH1, WG2, WH2, k1, k2, prow, pcol, a.flag);
    MNP_HIP(hipGetLastError());
    int hflag = 0;
    MNP_HIP(hipMemcpyAsync(&hflag, a.flag, sizeof(int), hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_CHECK(hflag == 0, MNP_ERR_ARG, "extinction_gradient: distinct faces with coincident centroids");
  }
  hipLaunchKernelGGL(k_adj_local, dim3((unsigned)cdiv_up(n, 256)), dim3(256), 0, stream(), n, p, lam, x, GX1, GX2,
                     g.geo, pd, k0, k2, b.eps1, b.eps2, loc);
  hipLaunchKernelGGL(k_adj_face, dim3((unsigned)cdiv_up(n, 256)), dim3(256), 0, stream(), n, nnz, g.geo, g.off, a.rptr,
                     a.rlist, nrt, nct, rowpart, colpart, prow, pcol, loc, fg);
  hipLaunchKernelGGL(k_adj_vertex, dim3((unsigned)cdiv_up(nv, 256)), dim3(256), 0, stream(), nv, n, a.vptr, a.vlist, fg,
                     4 * M_PI / k2.x, dv);
  MNP_HIP(hipGetLastError());
  MNP_HIP(hipMemcpyAsync(dverts, dv, sizeof(double) * 3 * nv, hipMemcpyDeviceToHost, stream()));
  MNP_HIP(hipStreamSynchronize(stream()));
}

}  // namespace mnp

// ============================================================== C ABI
using namespace mnp;

extern "C" {

int mnp_bemret_solve_transposed(mnp_bemret_t bem, int64_t npol, const void* c, void* lam) {
  return guarded([&] {
    MNP_CHECK(bem && c && lam, MNP_ERR_ARG, "bemret_solve_transposed: NULL argument");
    MNP_CHECK(c != lam, MNP_ERR_ARG, "bemret_solve_transposed: c and lam must differ");
    bem->b->solve_transposed(npol, (const double2*)c, (double2*)lam);
  });
}

int mnp_bemret_extinction_gradient(mnp_bemret_t bem, int64_t npol, const double* pol, const double* dir,
                                   const double* weights, const void* sig1, const void* sig2, const void* h1,
                                   const void* h2, double* dverts, double* deps1) {
  return guarded([&] {
    MNP_CHECK(bem && pol && dir && sig1 && sig2 && h1 && h2, MNP_ERR_ARG, "extinction_gradient: NULL argument");
    extinction_gradient(*bem->b, npol, pol, dir, weights, (const double2*)sig1, (const double2*)sig2,
                        (const double2*)h1, (const double2*)h2, dverts, deps1);
  });
}

}  // extern "C"
