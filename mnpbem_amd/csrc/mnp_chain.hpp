// The dense retarded BEM chain shared by the dense solvers (BEM/@bemret and,
// per symmetry value, BEM/@bemretmirror): chain state, its factorisation and
// mldivide on the factors.  See bemret.hip for the algebra.
#pragma once

#include <cmath>
#include <functional>
#include <string>
#include <vector>

#include "mnp_greenret.hpp"
#include "mnp_linalg.hpp"

namespace mnp {

// k = k0 sqrt(eps), principal root (Material/@epsdrude/subsref.m:24)
inline double2 wavenumber(double k0, double2 e) {
  const double r = std::hypot(e.x, e.y);
  double sr = std::sqrt(0.5 * (r + e.x)), si = std::sqrt(0.5 * (r - e.x));
  if (e.y < 0) si = -si;
  return mk<double>(k0 * sr, k0 * si);
}

// Host-side helpers of the dense solvers (bemret.hip).
double dev_norm2(int64_t len, const double2* v, Workspace& ws, int slot);  // |v|, synchronises
void launch_excitation(int64_t n, int64_t p, const double2* const in[8], double2* b, const double* nvec,
                       double k, double2 e1, double2 e2);
void launch_residual(int64_t n, int64_t p, const double2* b, const double2* GX1, const double2* GX2,
                     const double2* HX1, const double2* HX2, double2* r, const double* nvec, double k,
                     double2 e1, double2 e2);
void launch_unpack(int64_t n, int64_t p, const double2* x, double2* sig1, double2* h1, double2* sig2,
                   double2* h2);

// ---------------------------------------------------------- chain state
template <class T>
struct Chain {
  int64_t n = 0;
  bool exact_pivoting = true;  // FP64: factors must equal partial pivoting's
  cplx<T>*G1 = nullptr, *G2 = nullptr, *S1 = nullptr, *S2 = nullptr, *D = nullptr, *Di = nullptr;
  cplx<T>* backup = nullptr;   // copy of a matrix while it is factorised without pivoting
  cplx<T>* backup2 = nullptr;  // the same for the G2 branch, factorised concurrently with G1
  cplx<T>* inv = nullptr;      // 32-block inverses: 4 factors x (L, U)
  cplx<T>* invnb = nullptr;    // 256-block inverses: 4 factors x (L, U)
  int32_t *piv = nullptr, *piv2 = nullptr, *sig[4] = {nullptr, nullptr, nullptr, nullptr};
  LU<T> f[4];                  // G1, G2, Sigma1 - Sigma2, Sigma
  int fallbacks = 0;           // factorisations redone with partial pivoting
  bool ok_g1 = false, ok_g2 = false, ok_d = false;  // chain_sector_try's progress
  bool sigma_formed = false, ok_s = false;          // chain_build_try's
  // chain_explicit: G1^{-1}, G2^{-1} and Sigma^{-1} as matrices (over
  // backup, backup2 and D), so that chain_apply multiplies instead of
  // solving; cleared by every rebuild
  bool explicit_ok = false;
  double2 eps1{}, eps2{};
  double k0 = 0;
  // Compound particles (BEM/@bemret/private/initmat.m:19-42): eps1, eps2 per
  // unknown (device, n) instead of the scalars; with nested media (some
  // inside medium is some outside medium) also the full matrices L1 = G1
  // eps1 G1^{-1}, L2 = G2 eps2 G2^{-1} and two work matrices (enable_full_L).
  const double2* e1v = nullptr;
  const double2* e2v = nullptr;
  cplx<T>*L1m = nullptr, *L2m = nullptr, *W1 = nullptr, *W2 = nullptr;
  Workspace ws, ws2;  // ws2: the G2 branch

  // sigma_only: only the buffers of the Sigma factorisation (S2 holds Sigma,
  // factor slot 3), for Sigma matrices assembled from other chains' parts.
  explicit Chain(int64_t n_, bool exact, bool sigma_only = false) : n(n_), exact_pivoting(exact) {
    const int64_t nn = n * n;
    if (!sigma_only) {
      G1 = dev_alloc<cplx<T>>(nn);
      G2 = dev_alloc<cplx<T>>(nn);
      S1 = dev_alloc<cplx<T>>(nn);
      D = dev_alloc<cplx<T>>(nn);
      Di = dev_alloc<cplx<T>>(nn);
      backup2 = dev_alloc<cplx<T>>(nn);
      piv2 = dev_alloc<int32_t>(n);
    }
    S2 = dev_alloc<cplx<T>>(nn);
    backup = dev_alloc<cplx<T>>(nn);
    inv = dev_alloc<cplx<T>>(8 * inv_blocks(n));
    invnb = dev_alloc<cplx<T>>(8 * inv_blocks_nb(n));
    piv = dev_alloc<int32_t>(n);
    for (auto& q : sig) q = dev_alloc<int32_t>(n);
  }
  void enable_full_L() {
    if (L1m) return;
    const int64_t nn = n * n;
    L1m = dev_alloc<cplx<T>>(nn);
    L2m = dev_alloc<cplx<T>>(nn);
    W1 = dev_alloc<cplx<T>>(nn);
    W2 = dev_alloc<cplx<T>>(nn);
  }
  ~Chain() {
    void* ps[] = {G1, G2, S1, S2, D, Di, backup, backup2, inv, invnb, piv, piv2, sig[0], sig[1], sig[2], sig[3],
                  L1m, L2m, W1, W2};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
  MatView<T> v(cplx<T>* p) const { return MatView<T>{p, n}; }

  // Factorises A in place as factor slot `s` without row exchanges (scratch
  // bk, pv, w).  Returns false, with A restored, when partial pivoting would
  // have exchanged rows and exact pivoting is required, or on a zero pivot.
  bool factor_try(int s, cplx<T>* A, cplx<T>* bk, int32_t* pv, Workspace& w) {
    cplx<T>* Linv = inv + (2 * s) * inv_blocks(n);
    cplx<T>* Uinv = Linv + inv_blocks(n);
    copy_mat<T>(n, n, v(A), v(bk));
    int swap = 0;
    const int info = getrf<T>(n, n, v(A), pv, Linv, Uinv, PIVOT_NONE, &swap, w);
    if (info != 0 || (swap && exact_pivoting)) {
      copy_mat<T>(n, n, v(bk), v(A));
      return false;
    }
    cplx<T>* LinvNB = invnb + (2 * s) * inv_blocks_nb(n);
    cplx<T>* UinvNB = LinvNB + inv_blocks_nb(n);
    block_inverses_nb<T>(n, v(A), Linv, Uinv, LinvNB, UinvNB, w);
    f[s] = LU<T>{v(A), Linv, Uinv, nullptr, LinvNB, UinvNB};
    return true;
  }
  // Partial pivoting after factor_try declined.
  const LU<T>& factor_pivoted(int s, cplx<T>* A, const char* what) {
    ++fallbacks;
    cplx<T>* Linv = inv + (2 * s) * inv_blocks(n);
    cplx<T>* Uinv = Linv + inv_blocks(n);
    const int info = getrf<T>(n, n, v(A), piv, Linv, Uinv, PIVOT_PARTIAL, nullptr, ws);
    MNP_CHECK(info == 0, MNP_ERR_SINGULAR,
              std::string("bemret: singular ") + what + " (zero pivot " + std::to_string(info) + ")");
    perm_from_pivots(n, piv, sig[s], ws);
    cplx<T>* LinvNB = invnb + (2 * s) * inv_blocks_nb(n);
    cplx<T>* UinvNB = LinvNB + inv_blocks_nb(n);
    block_inverses_nb<T>(n, v(A), Linv, Uinv, LinvNB, UinvNB, ws);
    f[s] = LU<T>{v(A), Linv, Uinv, sig[s], LinvNB, UinvNB};
    return f[s];
  }
  // Factorises A in place as factor slot `s`: without row exchanges first
  // (no grid-wide pivot search), with partial pivoting when factor_try
  // declines.
  const LU<T>& factor(int s, cplx<T>* A, const char* what) {
    if (factor_try(s, A, backup, piv, ws)) return f[s];
    return factor_pivoted(s, A, what);
  }
};

// Sigma = e1 S1 - e2 S2 + c (Di o n n^T), written over S2.
template <class T>
__global__ void k_sigma(int64_t n, const cplx<T>* S1, cplx<T>* S2, const cplx<T>* Di,
                        const double* nvec, double2 e1, double2 e2, double2 c) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (i >= n || j >= n) return;
  const int64_t idx = j * n + i;
  const double nn = nvec[i] * nvec[j] + nvec[n + i] * nvec[n + j] + nvec[2 * n + i] * nvec[2 * n + j];
  const cplx<T> s1 = S1[idx], s2 = S2[idx], di = Di[idx];
  const double2 a = cmul<double>(e1, mk<double>(s1.x, s1.y));
  const double2 b = cmul<double>(e2, mk<double>(s2.x, s2.y));
  const double2 d = cmul<double>(c, mk<double>(di.x * nn, di.y * nn));
  S2[idx] = mk<T>((T)(a.x - b.x + d.x), (T)(a.y - b.y + d.y));
}
// Compound, diagonal L: Sigma = Sigma1 diag(e1) - Sigma2 diag(e2) + c (diag(L)
// Deltai o n n^T) diag(L), L = e1 - e2 per unknown (initmat.m:55-56), over S2.
template <class T>
__global__ void k_sigma_v(int64_t n, const cplx<T>* S1, cplx<T>* S2, const cplx<T>* Di, const double* nvec,
                          const double2* e1, const double2* e2, double c) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (i >= n || j >= n) return;
  const int64_t idx = j * n + i;
  const double nn = nvec[i] * nvec[j] + nvec[n + i] * nvec[n + j] + nvec[2 * n + i] * nvec[2 * n + j];
  const cplx<T> s1 = S1[idx], s2 = S2[idx], di = Di[idx];
  const double2 a = cmul<double>(e1[j], mk<double>(s1.x, s1.y));
  const double2 b = cmul<double>(e2[j], mk<double>(s2.x, s2.y));
  const double2 Li = csub<double>(e1[i], e2[i]), Lj = csub<double>(e1[j], e2[j]);
  const double2 d = cmul<double>(cmul<double>(Li, Lj), mk<double>(c * di.x * nn, c * di.y * nn));
  S2[idx] = mk<T>((T)(a.x - b.x + d.x), (T)(a.y - b.y + d.y));
}
// B[:, j] = A[:, j] e[j]  (n x n);  A o= n n^T
template <class T>
__global__ void k_scale_cols(int64_t n, const cplx<T>* A, const double2* e, cplx<T>* B) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (i >= n || j >= n) return;
  const cplx<T> a = A[j * n + i];
  const double2 v = cmul<double>(mk<double>(a.x, a.y), e[j]);
  B[j * n + i] = mk<T>((T)v.x, (T)v.y);
}
template <class T>
__global__ void k_scale_nn(int64_t n, cplx<T>* A, const double* nvec) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (i >= n || j >= n) return;
  const double nn = nvec[i] * nvec[j] + nvec[n + i] * nvec[n + j] + nvec[2 * n + i] * nvec[2 * n + j];
  const cplx<T> a = A[j * n + i];
  A[j * n + i] = mk<T>((T)(a.x * nn), (T)(a.y * nn));
}
template <class T>
__global__ void k_sub(int64_t n, const cplx<T>* A, const cplx<T>* B, cplx<T>* C) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (i >= n || j >= n) return;
  const int64_t idx = j * n + i;
  C[idx] = csub<T>(A[idx], B[idx]);
}
inline dim3 gridnn(int64_t m, int64_t n) {
  return dim3((unsigned)cdiv_up(m, 256), (unsigned)std::min<int64_t>(n, 65535),
              (unsigned)cdiv_up(n, 65535));
}

// Factorises the parts of the chain that precede Sigma: LU(G1), Sigma1 =
// H1 G1^{-1} (over H1 in S1), LU(G2), Sigma2 = H2 G2^{-1} (over H2 in S2),
// LU(Sigma1 - Sigma2) and Deltai.  On entry G1, G2, S1 (= H1), S2 (= H2)
// hold the assembled matrices.
// The parts of the chain that precede Sigma: LU(G1), Sigma1 = H1 G1^{-1}
// (over H1 in S1), LU(G2), Sigma2 = H2 G2^{-1} (over H2 in S2), LU(Sigma1 -
// Sigma2) and Deltai.  On entry G1, G2, S1 (= H1), S2 (= H2) hold the
// assembled matrices.
//
// chain_sector_try does what needs no partial pivoting (no grid-wide pivot
// search, no cooperative launch), so that several chains may run it at once
// on their own streams; it returns false when a factorisation must be redone
// with partial pivoting, which chain_sector_finish then does, alone.
template <class T>
bool chain_sector_try(Chain<T>& c) {
  const int64_t n = c.n;
  c.fallbacks = 0;
  c.explicit_ok = false;
  c.ok_g1 = c.ok_g2 = c.ok_d = false;
  if (c.L1m) {  // G1 diag(eps1), G2 diag(eps2), solved below into L1, L2
    hipLaunchKernelGGL(k_scale_cols<T>, gridnn(n, n), dim3(256), 0, stream(), n, c.G1, c.e1v, c.L1m);
    hipLaunchKernelGGL(k_scale_cols<T>, gridnn(n, n), dim3(256), 0, stream(), n, c.G2, c.e2v, c.L2m);
    MNP_HIP(hipGetLastError());
  }
  c.ok_g1 = c.factor_try(0, c.G1, c.backup, c.piv, c.ws);
  if (c.ok_g1) {
    getrs_right<T>(n, n, c.f[0], c.v(c.S1), c.ws);                // Sigma1 = H1 G1^{-1}
    if (c.L1m) getrs_right<T>(n, n, c.f[0], c.v(c.L1m), c.ws);   // L1 = G1 eps1 G1^{-1}
  }
  c.ok_g2 = c.factor_try(1, c.G2, c.backup2, c.piv2, c.ws2);
  if (c.ok_g2) {
    getrs_right<T>(n, n, c.f[1], c.v(c.S2), c.ws2);               // Sigma2 = H2 G2^{-1}
    if (c.L2m) getrs_right<T>(n, n, c.f[1], c.v(c.L2m), c.ws2);  // L2 = G2 eps2 G2^{-1}
  }
  if (!c.ok_g1 || !c.ok_g2) return false;
  hipLaunchKernelGGL(k_sub<T>, gridnn(n, n), dim3(256), 0, stream(), n, c.S1, c.S2, c.D);
  MNP_HIP(hipGetLastError());
  c.ok_d = c.factor_try(2, c.D, c.backup, c.piv, c.ws);
  if (c.ok_d) getri<T>(n, c.f[2], c.v(c.Di), c.ws);  // Deltai
  return c.ok_d;
}
template <class T>
void chain_sector_finish(Chain<T>& c) {
  const int64_t n = c.n;
  if (c.ok_d) return;
  const bool d_formed = c.ok_g1 && c.ok_g2;
  if (!c.ok_g1) {
    getrs_right<T>(n, n, c.factor_pivoted(0, c.G1, "G1"), c.v(c.S1), c.ws);
    if (c.L1m) getrs_right<T>(n, n, c.f[0], c.v(c.L1m), c.ws);
  }
  if (!c.ok_g2) {
    getrs_right<T>(n, n, c.factor_pivoted(1, c.G2, "G2"), c.v(c.S2), c.ws);
    if (c.L2m) getrs_right<T>(n, n, c.f[1], c.v(c.L2m), c.ws);
  }
  if (!d_formed) {
    hipLaunchKernelGGL(k_sub<T>, gridnn(n, n), dim3(256), 0, stream(), n, c.S1, c.S2, c.D);
    MNP_HIP(hipGetLastError());
    getri<T>(n, c.factor(2, c.D, "Sigma1 - Sigma2"), c.v(c.Di), c.ws);
  } else {
    getri<T>(n, c.factor_pivoted(2, c.D, "Sigma1 - Sigma2"), c.v(c.Di), c.ws);
  }
  c.ok_d = true;
}
template <class T>
void chain_sector(Chain<T>& c) {
  if (!chain_sector_try(c)) chain_sector_finish(c);
}

// Sigma = eps1 Sigma1 - eps2 Sigma2 + k^2 (eps1 - eps2)^2 (Deltai o n n^T), over S2
// nvec: n x 3 normals of the chain's unknowns (the faces, or aggregates of faces)
template <class T>
void chain_sigma(Chain<T>& c, const double* nvec) {
  const int64_t n = c.n;
  if (c.L1m) {
    // nested media: Sigma = Sigma1 L1 - Sigma2 L2 + k^2 ((L Deltai) o n n^T) L, L = L1 - L2
    const double2 one = {1, 0}, mone = {-1, 0}, zero = {0, 0};
    const double2 kk = {c.k0 * c.k0, 0}, mkk = {-c.k0 * c.k0, 0};
    gemm<T>(n, n, n, one, c.v(c.S1), c.v(c.L1m), zero, c.v(c.W1));
    gemm<T>(n, n, n, mone, c.v(c.S2), c.v(c.L2m), one, c.v(c.W1));
    gemm<T>(n, n, n, one, c.v(c.L1m), c.v(c.Di), zero, c.v(c.W2));
    gemm<T>(n, n, n, mone, c.v(c.L2m), c.v(c.Di), one, c.v(c.W2));
    hipLaunchKernelGGL(k_scale_nn<T>, gridnn(n, n), dim3(256), 0, stream(), n, c.W2, nvec);
    MNP_HIP(hipGetLastError());
    gemm<T>(n, n, n, kk, c.v(c.W2), c.v(c.L1m), one, c.v(c.W1));
    gemm<T>(n, n, n, mkk, c.v(c.W2), c.v(c.L2m), one, c.v(c.W1));
    copy_mat<T>(n, n, c.v(c.W1), c.v(c.S2));
    c.sigma_formed = true;
    return;
  }
  if (c.e1v) {
    hipLaunchKernelGGL(k_sigma_v<T>, gridnn(n, n), dim3(256), 0, stream(), n, c.S1, c.S2, c.Di, nvec, c.e1v, c.e2v,
                       c.k0 * c.k0);
    MNP_HIP(hipGetLastError());
    c.sigma_formed = true;
    return;
  }
  const double2 L = csub<double>(c.eps1, c.eps2);
  const double2 cc = cscale<double>(cmul<double>(L, L), c.k0 * c.k0);
  hipLaunchKernelGGL(k_sigma<T>, gridnn(n, n), dim3(256), 0, stream(), n, c.S1, c.S2, c.Di,
                     nvec, c.eps1, c.eps2, cc);
  MNP_HIP(hipGetLastError());
  c.sigma_formed = true;
}
// Factorises the chain (chain_sector, then Sigma over S2 and its LU), split
// like chain_sector into a part without partial pivoting and a finish.
template <class T>
bool chain_build_try(Chain<T>& c, const double* nvec) {
  c.sigma_formed = c.ok_s = false;
  if (!chain_sector_try(c)) return false;
  chain_sigma(c, nvec);
  c.ok_s = c.factor_try(3, c.S2, c.backup, c.piv, c.ws);
  return c.ok_s;
}
template <class T>
void chain_build_finish(Chain<T>& c, const double* nvec) {
  if (c.ok_s) return;
  if (!c.sigma_formed) {
    chain_sector_finish(c);
    chain_sigma(c, nvec);
    c.factor(3, c.S2, "Sigma");
  } else {
    c.factor_pivoted(3, c.S2, "Sigma");
  }
  c.ok_s = true;
}
template <class T>
void chain_build(Chain<T>& c, const double* nvec) {
  if (!chain_build_try(c, nvec)) chain_build_finish(c, nvec);
}
// After chain_build: the inverses of G1, G2 and Sigma as matrices, for a
// chain applied many times to few vectors (a preconditioner), where the
// triangular solves of chain_apply would be dozens of small dependent
// launches each.  Overwrites the factorisation of Sigma1 - Sigma2 (D, no
// longer needed once Deltai is formed) and the backups.
template <class T>
void chain_explicit(Chain<T>& c) {
  MNP_CHECK(c.ok_s && c.G1 && c.D, MNP_ERR_STATE, "chain_explicit: chain not built");
  MNP_CHECK(!c.e1v, MNP_ERR_STATE, "chain_explicit: not for compound particles");
  getri<T>(c.n, c.f[0], c.v(c.backup), c.ws);
  getri<T>(c.n, c.f[1], c.v(c.backup2), c.ws);
  getri<T>(c.n, c.f[3], c.v(c.D), c.ws);
  c.explicit_ok = true;
}
template <class T>
bool chain_build_try(Chain<T>& c, const GreenRet& g) { return chain_build_try(c, g.geo.nvec); }
template <class T>
void chain_build_finish(Chain<T>& c, const GreenRet& g) { chain_build_finish(c, g.geo.nvec); }
template <class T>
void chain_build(Chain<T>& c, const GreenRet& g) { chain_build(c, g.geo.nvec); }

// ------------------------------------------------------------- mldivide
// Vectors are n x (8 p) blocks.  b = [phi (p) | a (3p) | De (p) | alpha (3p)],
// x = [sig1 (p) | h1 (3p) | sig2 (p) | h2 (3p)]; vector columns are ordered
// component-fastest (c + 3 q), as Matlab's n x 3 x npol arrays.

// alpha <- alpha - Y_a + i k L1 n (x) phi ;  De <- De - L1 Y_phi + i k L1 n . a
template <class T>
__global__ void k_mld1(int64_t n, int64_t p, const cplx<T>* b, const cplx<T>* Y, cplx<T>* alpha,
                       cplx<T>* De, const double* nvec, double k, double2 L1) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const cplx<T>* phi = b;
  const cplx<T>* a = b + n * p;
  const cplx<T>* De0 = b + 4 * n * p;
  const cplx<T>* al0 = b + 5 * n * p;
  const double2 ikL1 = cmul<double>(mk<double>(0, k), L1);
  const cplx<T> ph = phi[q * n + i];
  const double2 ikL1phi = cmul<double>(ikL1, mk<double>(ph.x, ph.y));
  double2 na = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) {
    const int64_t col = c + 3 * q;
    const double nc = nvec[c * n + i];
    const cplx<T> ya = Y[col * n + i], a0 = al0[col * n + i], ac = a[col * n + i];
    alpha[col * n + i] = mk<T>((T)(a0.x - ya.x + nc * ikL1phi.x), (T)(a0.y - ya.y + nc * ikL1phi.y));
    na.x += nc * ac.x;
    na.y += nc * ac.y;
  }
  const cplx<T> yp = Y[(3 * p + q) * n + i], d0 = De0[q * n + i];
  const double2 l1y = cmul<double>(L1, mk<double>(yp.x, yp.y));
  const double2 t = cmul<double>(ikL1, na);
  De[q * n + i] = mk<T>((T)(d0.x - l1y.x + t.x), (T)(d0.y - l1y.y + t.y));
}
// rhs = De + i k (L1 - L2) n . t
template <class T>
__global__ void k_mld2(int64_t n, int64_t p, const cplx<T>* De, const cplx<T>* t, cplx<T>* rhs,
                       const double* nvec, double2 ikdL) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  double2 s = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) {
    const cplx<T> v = t[(c + 3 * q) * n + i];
    s.x += nvec[c * n + i] * v.x;
    s.y += nvec[c * n + i] * v.y;
  }
  const double2 u = cmul<double>(ikdL, s);
  const cplx<T> d = De[q * n + i];
  rhs[q * n + i] = mk<T>((T)(d.x + u.x), (T)(d.y + u.y));
}
// u = i k (L1 - L2) n (x) sig2 + alpha
template <class T>
__global__ void k_mld3(int64_t n, int64_t p, const cplx<T>* sig2, const cplx<T>* alpha, cplx<T>* u,
                       const double* nvec, double2 ikdL) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const cplx<T> s = sig2[q * n + i];
  const double2 w = cmul<double>(ikdL, mk<double>(s.x, s.y));
  for (int c = 0; c < 3; ++c) {
    const int64_t col = c + 3 * q;
    const cplx<T> a = alpha[col * n + i];
    const double nc = nvec[c * n + i];
    u[col * n + i] = mk<T>((T)(a.x + nc * w.x), (T)(a.y + nc * w.y));
  }
}
// V1 = [sig2 + phi | h2 + a] ; V2 = [sig2 | h2]   (n x 4p each)
template <class T>
__global__ void k_mld4(int64_t n, int64_t p, const cplx<T>* b, const cplx<T>* sig2,
                       const cplx<T>* h2, cplx<T>* V1, cplx<T>* V2) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t col = blockIdx.y;  // 0 .. 4p
  if (i >= n) return;
  const cplx<T> src = col < p ? sig2[col * n + i] : h2[(col - p) * n + i];
  const cplx<T> add = b[col * n + i];  // phi then a
  V1[col * n + i] = cadd<T>(src, add);
  V2[col * n + i] = src;
}

// Compound mldivide steps (mldivide.m:29-40 with L1, L2 applied beforehand):
// alpha <- alpha - Y_a + i k n (x) P, De <- De - Y_phi + i k n . A1 with P = L1 phi,
// A1 = L1 a; rhs = De + i k n . tL; u = alpha + i k n (x) sL.
template <class T>
__global__ void k_mldg1(int64_t n, int64_t p, const cplx<T>* b, const cplx<T>* Y, const cplx<T>* P,
                        const cplx<T>* A1, cplx<T>* alpha, cplx<T>* De, const double* nvec, double k) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const cplx<T>* De0 = b + 4 * n * p;
  const cplx<T>* al0 = b + 5 * n * p;
  const cplx<T> pp = P[q * n + i];
  const double2 ikP = mk<double>(-k * pp.y, k * pp.x);
  double2 na = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) {
    const int64_t col = c + 3 * q;
    const double nc = nvec[c * n + i];
    const cplx<T> ya = Y[col * n + i], a0 = al0[col * n + i], ac = A1[col * n + i];
    alpha[col * n + i] = mk<T>((T)(a0.x - ya.x + nc * ikP.x), (T)(a0.y - ya.y + nc * ikP.y));
    na.x += nc * ac.x;
    na.y += nc * ac.y;
  }
  const cplx<T> yp = Y[(3 * p + q) * n + i], d0 = De0[q * n + i];
  De[q * n + i] = mk<T>((T)(d0.x - yp.x - k * na.y), (T)(d0.y - yp.y + k * na.x));
}
template <class T>
__global__ void k_mldg2(int64_t n, int64_t p, const cplx<T>* De, const cplx<T>* tL, cplx<T>* rhs, const double* nvec,
                        double k) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  double2 s = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) {
    const cplx<T> v = tL[(c + 3 * q) * n + i];
    s.x += nvec[c * n + i] * v.x;
    s.y += nvec[c * n + i] * v.y;
  }
  const cplx<T> d = De[q * n + i];
  rhs[q * n + i] = mk<T>((T)(d.x - k * s.y), (T)(d.y + k * s.x));
}
template <class T>
__global__ void k_mldg3(int64_t n, int64_t p, const cplx<T>* sL, const cplx<T>* alpha, cplx<T>* u, const double* nvec,
                        double k) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const cplx<T> s = sL[q * n + i];
  for (int c = 0; c < 3; ++c) {
    const int64_t col = c + 3 * q;
    const cplx<T> a = alpha[col * n + i];
    const double nc = nvec[c * n + i];
    u[col * n + i] = mk<T>((T)(a.x - nc * k * s.y), (T)(a.y + nc * k * s.x));
  }
}
// out = diag(e) X or diag(e1 - e2) X (X n x ncols)
template <class T>
__global__ void k_rowscale(int64_t n, int64_t ncols, const cplx<T>* X, const double2* e1, const double2* e2,
                           cplx<T>* out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t col = blockIdx.y;
  if (i >= n) return;
  const double2 e = e2 ? csub<double>(e1[i], e2[i]) : e1[i];
  const cplx<T> v = X[col * n + i];
  const double2 r = cmul<double>(e, mk<double>(v.x, v.y));
  out[col * n + i] = mk<T>((T)r.x, (T)r.y);
}

// out = L1 X (which = 1) or (L1 - L2) X (which = 0) for a compound chain:
// row scalings with diagonal L, products with the full matrices.
template <class T>
void chain_apply_L(Chain<T>& c, int which, int64_t ncols, const cplx<T>* X, cplx<T>* out) {
  const int64_t n = c.n;
  if (c.L1m) {
    const double2 one = {1, 0}, mone = {-1, 0}, zero = {0, 0};
    MatView<T> Xv{const_cast<cplx<T>*>(X), n}, Ov{out, n};
    gemm<T>(n, ncols, n, one, c.v(c.L1m), Xv, zero, Ov);
    if (which == 0) gemm<T>(n, ncols, n, mone, c.v(c.L2m), Xv, one, Ov);
    return;
  }
  hipLaunchKernelGGL(k_rowscale<T>, dim3((unsigned)cdiv_up(n, 256), (unsigned)ncols), dim3(256), 0, stream(), n,
                     ncols, X, c.e1v, which == 0 ? c.e2v : (const double2*)nullptr, out);
  MNP_HIP(hipGetLastError());
}

// mldivide (mldivide.m:11-49) for a compound chain (eps per unknown, L
// diagonal or full); tmp holds n x 26p.
template <class T>
void chain_apply_compound(Chain<T>& c, const double* nvec, int64_t p, const cplx<T>* b, cplx<T>* x, cplx<T>* tmp) {
  const int64_t n = c.n, pn = p * n;
  const double k = c.k0;
  // tmp: Y (4p) | alpha (3p) | De (p) | t (3p) | s2 (p) | u (3p) | h2 (3p) | P (p) | A1 (3p) | tL (3p) | sL (p)
  cplx<T>* Y = tmp;
  cplx<T>* alpha = Y + 4 * pn;
  cplx<T>* De = alpha + 3 * pn;
  cplx<T>* t = De + pn;
  cplx<T>* s2 = t + 3 * pn;
  cplx<T>* u = s2 + pn;
  cplx<T>* h2 = u + 3 * pn;
  cplx<T>* P = h2 + 3 * pn;
  cplx<T>* A1 = P + pn;
  cplx<T>* tL = A1 + 3 * pn;
  cplx<T>* sL = tL + 3 * pn;
  const double2 one = {1, 0}, zero = {0, 0};
  chain_apply_L<T>(c, 1, p, b, P);            // L1 phi
  chain_apply_L<T>(c, 1, 3 * p, b + pn, A1);  // L1 a
  gemm<T>(n, 3 * p, n, one, c.v(c.S1), MatView<T>{const_cast<cplx<T>*>(b) + pn, n}, zero, MatView<T>{Y, n});
  gemm<T>(n, p, n, one, c.v(c.S1), MatView<T>{P, n}, zero, MatView<T>{Y + 3 * pn, n});
  const dim3 gp((unsigned)cdiv_up(n, 256), (unsigned)p);
  hipLaunchKernelGGL(k_mldg1<T>, gp, dim3(256), 0, stream(), n, p, b, Y, P, A1, alpha, De, nvec, k);
  gemm<T>(n, 3 * p, n, one, c.v(c.Di), MatView<T>{alpha, n}, zero, MatView<T>{t, n});
  chain_apply_L<T>(c, 0, 3 * p, t, tL);       // (L1 - L2) Deltai alpha
  hipLaunchKernelGGL(k_mldg2<T>, gp, dim3(256), 0, stream(), n, p, De, tL, s2, nvec, k);
  getrs<T>(n, p, c.f[3], MatView<T>{s2, n}, c.ws);  // sig2 = Sigma^{-1} rhs
  chain_apply_L<T>(c, 0, p, s2, sL);          // (L1 - L2) sig2
  hipLaunchKernelGGL(k_mldg3<T>, gp, dim3(256), 0, stream(), n, p, sL, alpha, u, nvec, k);
  gemm<T>(n, 3 * p, n, one, c.v(c.Di), MatView<T>{u, n}, zero, MatView<T>{h2, n});
  hipLaunchKernelGGL(k_mld4<T>, dim3((unsigned)cdiv_up(n, 256), (unsigned)(4 * p)), dim3(256), 0, stream(), n, p, b,
                     s2, h2, x, x + 4 * pn);
  MNP_HIP(hipGetLastError());
  getrs<T>(n, 4 * p, c.f[0], MatView<T>{x, n}, c.ws);
  getrs<T>(n, 4 * p, c.f[1], MatView<T>{x + 4 * pn, n}, c.ws);
}

// x = M^{-1} b for the chain c (precision T), b and x n x 8p in T.
template <class T>
void chain_apply(Chain<T>& c, const double* nvec, int64_t p, const cplx<T>* b, cplx<T>* x, cplx<T>* tmp) {
  if (c.e1v) {
    chain_apply_compound<T>(c, nvec, p, b, x, tmp);
    return;
  }
  const int64_t n = c.n;
  const double k = c.k0;
  const double2 L1 = c.eps1, L2 = c.eps2;
  const double2 ikdL = cmul<double>(mk<double>(0, k), csub<double>(L1, L2));
  // tmp layout: Y (4p) | alpha (3p) | De (p) | t (3p) | rhs/sig2 (p) | u (3p) | h2 (3p)
  cplx<T>* Y = tmp;
  cplx<T>* alpha = Y + 4 * p * n;
  cplx<T>* De = alpha + 3 * p * n;
  cplx<T>* t = De + p * n;
  cplx<T>* s2 = t + 3 * p * n;
  cplx<T>* u = s2 + p * n;
  cplx<T>* h2 = u + 3 * p * n;
  const double2 one = {1, 0}, zero = {0, 0};
  // Y = Sigma1 [a | phi]: a occupies b columns p..4p, phi columns 0..p
  gemm<T>(n, 3 * p, n, one, c.v(c.S1), MatView<T>{const_cast<cplx<T>*>(b) + n * p, n}, zero,
          MatView<T>{Y, n});
  gemm<T>(n, p, n, one, c.v(c.S1), MatView<T>{const_cast<cplx<T>*>(b), n}, zero,
          MatView<T>{Y + 3 * p * n, n});
  const dim3 gp((unsigned)cdiv_up(n, 256), (unsigned)p);
  hipLaunchKernelGGL(k_mld1<T>, gp, dim3(256), 0, stream(), n, p, b, Y, alpha, De, nvec, k, L1);
  gemm<T>(n, 3 * p, n, one, c.v(c.Di), MatView<T>{alpha, n}, zero, MatView<T>{t, n});
  if (c.explicit_ok) {
    // Y is free after k_mld1, alpha | De (4p columns) after k_mld3
    hipLaunchKernelGGL(k_mld2<T>, gp, dim3(256), 0, stream(), n, p, De, t, Y, nvec, ikdL);
    gemm<T>(n, p, n, one, c.v(c.D), MatView<T>{Y, n}, zero, MatView<T>{s2, n});  // Sigma^{-1} rhs
    hipLaunchKernelGGL(k_mld3<T>, gp, dim3(256), 0, stream(), n, p, s2, alpha, u, nvec, ikdL);
    gemm<T>(n, 3 * p, n, one, c.v(c.Di), MatView<T>{u, n}, zero, MatView<T>{h2, n});
    hipLaunchKernelGGL(k_mld4<T>, dim3((unsigned)cdiv_up(n, 256), (unsigned)(4 * p)), dim3(256), 0,
                       stream(), n, p, b, s2, h2, Y, alpha);
    MNP_HIP(hipGetLastError());
    gemm<T>(n, 4 * p, n, one, c.v(c.backup), MatView<T>{Y, n}, zero, MatView<T>{x, n});
    gemm<T>(n, 4 * p, n, one, c.v(c.backup2), MatView<T>{alpha, n}, zero, MatView<T>{x + 4 * p * n, n});
    return;
  }
  hipLaunchKernelGGL(k_mld2<T>, gp, dim3(256), 0, stream(), n, p, De, t, s2, nvec, ikdL);
  getrs<T>(n, p, c.f[3], MatView<T>{s2, n}, c.ws);  // sig2 = Sigma^{-1} rhs
  hipLaunchKernelGGL(k_mld3<T>, gp, dim3(256), 0, stream(), n, p, s2, alpha, u, nvec, ikdL);
  gemm<T>(n, 3 * p, n, one, c.v(c.Di), MatView<T>{u, n}, zero, MatView<T>{h2, n});
  // V1 -> x[0:4p] (sig1 | h1), V2 -> x[4p:8p] (sig2 | h2)
  hipLaunchKernelGGL(k_mld4<T>, dim3((unsigned)cdiv_up(n, 256), (unsigned)(4 * p)), dim3(256), 0,
                     stream(), n, p, b, s2, h2, x, x + 4 * p * n);
  MNP_HIP(hipGetLastError());
  getrs<T>(n, 4 * p, c.f[0], MatView<T>{x, n}, c.ws);
  getrs<T>(n, 4 * p, c.f[1], MatView<T>{x + 4 * p * n, n}, c.ws);
}
template <class T>
void chain_apply(Chain<T>& c, const GreenRet& g, int64_t p, const cplx<T>* b, cplx<T>* x, cplx<T>* tmp) {
  chain_apply(c, g.geo.nvec, p, b, x, tmp);
}

}  // namespace mnp
