// Iterative retarded BEM solver with H-matrices (MNPBEM BEM/@bemretiter).
//
// initmat (bemretiter/private/initmat.m:5-31): G1 = G(k1), G2 = G(k2),
// H1 = F(k1) + 2 pi, H2 = F(k2) - 2 pi as H-matrices (full blocks with
// refinement + ACA blocks, hmatrix.hip), plus the preconditioner.
// solve (bemretiter/solve.m, bemiter/solve.m): GMRES without restart on
// the exact operator afun (private/afun.m:6-35, four H-matvecs),
// preconditioned by mfun (private/mfun.m:10-52) from the right in the
// flexible form, so that the minimised residual is |b - A x|.
//
// Preconditioner: the reference compresses the whole bemret chain into
// H-matrices of rank 4 (private/initprecond.m, H-LU and H-inverse), so that
// it couples the whole surface.  Here, two levels:
//  - smoother: the same chain applied exactly on diagonal blocks of at most
//    32 consecutive cluster indices inside each leaf cluster (block Jacobi):
//    G1, G2, Sigma1, Deltai, Sigma of every block are formed and factorised
//    in LDS by one workgroup, in complex FP32;
//  - coarse level: the bemret chain on the aggregated operators P^T G1 P,
//    P^T G2 P, P^T H1 P, P^T H2 P (P the indicator of the H-matrix row
//    tiles, <= 32 faces each; summed from the stored blocks and factors) with
//    the tiles' mean normals, on the FP32 MFMA chain of the dense solver.
// mfun applies the coarse correction and then the smoother to what remains:
//   z = P y + B^{-1} (v - A P y),  y = C^{-1} P^T v.
// precond 1 keeps the smoother alone.  The preconditioner changes the
// iteration count only; the converged solution solves the same equations.
//
// Substrate (bemretlayeriter, BEM/@bemretlayeriter/private/initmat.m,
// afun.m:1-47): G2 and H2 become the five components p, ss, hs, sh, hh of
// the outside Green function with the interface (Greenfun/+aca/
// @compgreenretlayer/private/eval2.m: direct + reflected for p, ss, hh,
// reflected only for hs, sh), each an H-matrix of its own; afun takes the
// structured products.  The preconditioner uses the p components in place
// of G2 and H2 (the free-space chain with the parallel reflection).
#include <complex>
#include <memory>
#include <vector>

#include "mnp_chain.hpp"
#include "mnp_gmres.hpp"
#include "mnp_layertab.hpp"
#include "mnp_hmatrix.hpp"

namespace mnp {

void launch_excitation(int64_t n, int64_t p, const double2* const in[8], double2* b, const double* nvec,
                       double k, double2 e1, double2 e2);
void launch_residual(int64_t n, int64_t p, const double2* b, const double2* GX1, const double2* GX2,
                     const double2* HX1, const double2* HX2, double2* r, const double* nvec, double k,
                     double2 e1, double2 e2);
void launch_unpack(int64_t n, int64_t p, const double2* x, double2* sig1, double2* h1, double2* sig2,
                   double2* h2);

constexpr int PB = 32;  // largest leaf cluster handled by the block preconditioner
using cf = float2;

// ------------------------------------------------ small dense LDS algebra
// All matrices are PB x PB row-major in LDS, nb <= PB active rows/cols.
__device__ void lds_lu(cf (*M)[PB + 1], int nb) {  // in place, no pivoting
  for (int j = 0; j < nb; ++j) {
    const cf piv = M[j][j];
    for (int i = j + 1 + (int)threadIdx.x; i < nb; i += blockDim.x) M[i][j] = cdiv<float>(M[i][j], piv);
    __syncthreads();
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
      const int i = e / nb, c = e % nb;
      if (i > j && c > j) M[i][c] = csub<float>(M[i][c], cmul<float>(M[i][j], M[j][c]));
    }
    __syncthreads();
  }
}
// X <- X (L U)^{-1}  (rows of X independent; X nb x nb)
__device__ void lds_rsolve(cf (*X)[PB + 1], cf (*LU)[PB + 1], int nb) {
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    for (int j = 0; j < nb; ++j) {  // X U = Y
      cf s = X[i][j];
      for (int k = 0; k < j; ++k) s = csub<float>(s, cmul<float>(X[i][k], LU[k][j]));
      X[i][j] = cdiv<float>(s, LU[j][j]);
    }
    for (int j = nb - 1; j >= 0; --j) {  // Z L = X
      cf s = X[i][j];
      for (int k = j + 1; k < nb; ++k) s = csub<float>(s, cmul<float>(X[i][k], LU[k][j]));
      X[i][j] = s;
    }
  }
  __syncthreads();
}
// Out <- (L U)^{-1}, LU from lds_lu: one thread per column of the identity.
__device__ void lds_inverse(cf (*Out)[PB + 1], cf (*LU)[PB + 1], int nb) {
  for (int e = threadIdx.x; e < PB * PB; e += blockDim.x) {
    const int i = e % PB, j = e / PB;
    Out[i][j] = i == j ? make_float2(1.f, 0.f) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < nb; c += blockDim.x) {
    for (int i = 0; i < nb; ++i) {
      cf s = Out[i][c];
      for (int k = 0; k < i; ++k) s = csub<float>(s, cmul<float>(LU[i][k], Out[k][c]));
      Out[i][c] = s;
    }
    for (int i = nb - 1; i >= 0; --i) {
      cf s = Out[i][c];
      for (int k = i + 1; k < nb; ++k) s = csub<float>(s, cmul<float>(LU[i][k], Out[k][c]));
      Out[i][c] = cdiv<float>(s, LU[i][i]);
    }
  }
  __syncthreads();
}

// Per preconditioner block: G1^{-1}, G2^{-1}, Sigma1, Deltai, Sigma^{-1}
// (5 x PB x PB, FP32), the bemret chain (initmat.m) on a diagonal block.
// pb[5 b ..] = {first cluster index, size nb <= PB, offset in its leaf,
// leaf size m, value offset of the leaf's diagonal full block}.
template <class TS>
__global__ void __launch_bounds__(256) k_bjprec_build(const int64_t* pb, const cplx<TS>* G1v, const cplx<TS>* G2v,
                                                      const cplx<TS>* H1v, const cplx<TS>* H2v, const int32_t* perm,
                                                      const double* nvec, int64_t n, double2 e1, double2 e2,
                                                      double2 cc, float2* out) {
  __shared__ cf A[PB][PB + 1], B[PB][PB + 1], S1[PB][PB + 1], S2[PB][PB + 1], D[PB][PB + 1], X[PB][PB + 1];
  const int64_t l = blockIdx.x;
  const int r0 = (int)pb[5 * l], nb = (int)pb[5 * l + 1], off = (int)pb[5 * l + 2];
  const int64_t m = pb[5 * l + 3], vo = pb[5 * l + 4];
  const int tid = threadIdx.x;
  auto ld = [&](const cplx<TS>* v, int i, int j) {
    const cplx<TS> z = v[vo + (int64_t)(off + j) * m + off + i];
    return make_float2((float)z.x, (float)z.y);
  };
  for (int e = tid; e < PB * PB; e += 256) {
    const int i = e % PB, j = e / PB;
    const bool in = i < nb && j < nb;
    const cf id = i == j ? make_float2(1.f, 0.f) : make_float2(0.f, 0.f);
    A[i][j] = in ? ld(G1v, i, j) : id;
    B[i][j] = in ? ld(G2v, i, j) : id;
    S1[i][j] = in ? ld(H1v, i, j) : id;
    S2[i][j] = in ? ld(H2v, i, j) : id;
  }
  __syncthreads();
  lds_lu(A, nb);
  lds_lu(B, nb);
  lds_rsolve(S1, A, nb);  // Sigma1 = H1 G1^{-1}
  lds_rsolve(S2, B, nb);  // Sigma2 = H2 G2^{-1}
  for (int e = tid; e < PB * PB; e += 256) {
    const int i = e % PB, j = e / PB;
    D[i][j] = csub<float>(S1[i][j], S2[i][j]);
    X[i][j] = i == j ? make_float2(1.f, 0.f) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_lu(D, nb);
  lds_inverse(X, D, nb);  // Deltai
  // Sigma = e1 Sigma1 - e2 Sigma2 + cc (Deltai o n n^T), into S2
  for (int e = tid; e < nb * nb; e += 256) {
    const int i = e % nb, j = e / nb;
    const int64_t fi = perm[r0 + i], fj = perm[r0 + j];
    const double nn = nvec[fi] * nvec[fj] + nvec[n + fi] * nvec[n + fj] + nvec[2 * n + fi] * nvec[2 * n + fj];
    const double2 a = cmul<double>(e1, mk<double>(S1[i][j].x, S1[i][j].y));
    const double2 b = cmul<double>(e2, mk<double>(S2[i][j].x, S2[i][j].y));
    const double2 d = cmul<double>(cc, mk<double>(X[i][j].x * nn, X[i][j].y * nn));
    S2[i][j] = make_float2((float)(a.x - b.x + d.x), (float)(a.y - b.y + d.y));
  }
  __syncthreads();
  lds_lu(S2, nb);
  float2* o = out + l * 5 * PB * PB;
  auto put = [&](int slot, cf (*M)[PB + 1]) {
    for (int e = tid; e < PB * PB; e += 256) o[slot * PB * PB + e] = M[e / PB][e % PB];
    __syncthreads();
  };
  lds_inverse(D, S2, nb);  // Sigma^{-1}
  put(4, D);
  put(2, S1);
  put(3, X);
  lds_inverse(S2, A, nb);  // G1^{-1}
  put(0, S2);
  lds_inverse(S2, B, nb);  // G2^{-1}
  put(1, S2);
}

// x = M^{-1} b on one preconditioner block's rows (mfun.m:27-52 with the
// block matrices).  b, x: n x 8p FP64 in face order,
// [phi | a | De | alpha] -> [sig1 | h1 | sig2 | h2].
__global__ void __launch_bounds__(256) k_bjprec_apply(const int64_t* pb, const int32_t* perm,
                                                      const float2* fac, const double* nvec, int64_t n, int p,
                                                      double k, double2 L1, double2 L2, const double2* b,
                                                      double2* x) {
  __shared__ cf M[PB][PB + 1];
  __shared__ cf V[PB][2 * PB + 1];   // up to 8p <= 64 vector columns (p <= 8)
  __shared__ cf W[PB][2 * PB + 1];
  const int64_t l = blockIdx.x;
  const int r0 = (int)pb[5 * l], nb = (int)pb[5 * l + 1];
  const float2* F = fac + l * 5 * PB * PB;
  const int tid = threadIdx.x, nc = 8 * p;
  auto loadM = [&](int which) {
    __syncthreads();
    for (int e = tid; e < PB * PB; e += 256) M[e / PB][e % PB] = F[which * PB * PB + e];
    __syncthreads();
  };
  // V = local b (rows of this leaf)
  for (int e = tid; e < nb * nc; e += 256) {
    const int i = e % nb, c = e / nb;
    const double2 v = b[(int64_t)c * n + perm[r0 + i]];
    V[i][c] = make_float2((float)v.x, (float)v.y);
  }
  const double2 ik = mk<double>(0, k);
  const double2 ikL1 = cmul<double>(ik, L1), ikdL = cmul<double>(ik, csub<double>(L1, L2));
  auto nv = [&](int i, int c) { return (float)nvec[(int64_t)c * n + perm[r0 + i]]; };
  auto mul = [&](const cf z, double2 s) {
    const double2 r = cmul<double>(mk<double>(z.x, z.y), s);
    return make_float2((float)r.x, (float)r.y);
  };
  // W[:, 0:4p] = Sigma1 [a | phi]  (a = V cols p..4p, phi = V cols 0..p)
  loadM(2);
  for (int e = tid; e < nb * 4 * p; e += 256) {
    const int i = e % nb, c = e / nb;
    const int src = c < 3 * p ? p + c : c - 3 * p;
    cf s = make_float2(0.f, 0.f);
    for (int kk = 0; kk < nb; ++kk) s = cadd<float>(s, cmul<float>(M[i][kk], V[kk][src]));
    W[i][c] = s;
  }
  __syncthreads();
  // alpha' = alpha - Ya + ik L1 n phi (V cols 5p..8p), De' = De - L1 Yphi + ik L1 n.a (V col 4p+q)
  for (int e = tid; e < nb * p; e += 256) {
    const int i = e % nb, q = e / nb;
    const cf phi = V[i][q];
    const cf w = mul(phi, ikL1);
    cf na = make_float2(0.f, 0.f);
    for (int c = 0; c < 3; ++c) {
      const int col = c + 3 * q;
      const float n_c = nv(i, c);
      V[i][5 * p + col] = cadd<float>(csub<float>(V[i][5 * p + col], W[i][col]), make_float2(n_c * w.x, n_c * w.y));
      na = cadd<float>(na, make_float2(n_c * V[i][p + col].x, n_c * V[i][p + col].y));
    }
    V[i][4 * p + q] = cadd<float>(csub<float>(V[i][4 * p + q], mul(W[i][3 * p + q], L1)), mul(na, ikL1));
  }
  // t = Deltai alpha' -> W[:, 0:3p]
  loadM(3);
  for (int e = tid; e < nb * 3 * p; e += 256) {
    const int i = e % nb, c = e / nb;
    cf s = make_float2(0.f, 0.f);
    for (int kk = 0; kk < nb; ++kk) s = cadd<float>(s, cmul<float>(M[i][kk], V[kk][5 * p + c]));
    W[i][c] = s;
  }
  __syncthreads();
  // rhs = De' + ik (L1-L2) n.t -> W[:, 3p+q]
  for (int e = tid; e < nb * p; e += 256) {
    const int i = e % nb, q = e / nb;
    cf s = make_float2(0.f, 0.f);
    for (int c = 0; c < 3; ++c) {
      const float n_c = nv(i, c);
      s = cadd<float>(s, make_float2(n_c * W[i][c + 3 * q].x, n_c * W[i][c + 3 * q].y));
    }
    W[i][3 * p + q] = cadd<float>(V[i][4 * p + q], mul(s, ikdL));
  }
  loadM(4);
  // sig2 = Sigma^{-1} rhs -> W[:, 7p .. 8p]
  for (int e = tid; e < nb * p; e += 256) {
    const int i = e % nb, c = e / nb;
    cf s = make_float2(0.f, 0.f);
    for (int kk = 0; kk < nb; ++kk) s = cadd<float>(s, cmul<float>(M[i][kk], W[kk][3 * p + c]));
    W[i][7 * p + c] = s;
  }
  __syncthreads();
  // u = ik (L1-L2) n sig2 + alpha' -> W[:, 4p .. 7p]
  for (int e = tid; e < nb * p; e += 256) {
    const int i = e % nb, q = e / nb;
    const cf w = mul(W[i][7 * p + q], ikdL);
    for (int c = 0; c < 3; ++c) {
      const float n_c = nv(i, c);
      W[i][4 * p + c + 3 * q] = cadd<float>(V[i][5 * p + c + 3 * q], make_float2(n_c * w.x, n_c * w.y));
    }
  }
  loadM(3);
  // h2 = Deltai u -> W[:, 0:3p]  (t no longer needed)
  for (int e = tid; e < nb * 3 * p; e += 256) {
    const int i = e % nb, c = e / nb;
    cf s = make_float2(0.f, 0.f);
    for (int kk = 0; kk < nb; ++kk) s = cadd<float>(s, cmul<float>(M[i][kk], W[kk][4 * p + c]));
    W[i][c] = s;
  }
  __syncthreads();
  // V1 = [sig2 + phi | h2 + a] -> V cols 0..4p ; V2 = [sig2 | h2] -> V cols 4p..8p
  for (int e = tid; e < nb * 4 * p; e += 256) {
    const int i = e % nb, c = e / nb;
    const cf src = c < p ? W[i][7 * p + c] : W[i][c - p];
    V[i][c] = cadd<float>(src, V[i][c]);
    V[i][4 * p + c] = src;
  }
  __syncthreads();
  // G1^{-1} on V cols 0..4p, G2^{-1} on V cols 4p..8p -> W
  loadM(0);
  for (int e = tid; e < nb * 4 * p; e += 256) {
    const int i = e % nb, c = e / nb;
    cf s = make_float2(0.f, 0.f);
    for (int kk = 0; kk < nb; ++kk) s = cadd<float>(s, cmul<float>(M[i][kk], V[kk][c]));
    W[i][c] = s;
  }
  loadM(1);
  for (int e = tid; e < nb * 4 * p; e += 256) {
    const int i = e % nb, c = 4 * p + e / nb;
    cf s = make_float2(0.f, 0.f);
    for (int kk = 0; kk < nb; ++kk) s = cadd<float>(s, cmul<float>(M[i][kk], V[kk][c]));
    W[i][c] = s;
  }
  __syncthreads();
  for (int e = tid; e < nb * nc; e += 256) {
    const int i = e % nb, c = e / nb;
    x[(int64_t)c * n + perm[r0 + i]] = mk<double>(W[i][c].x, W[i][c].y);
  }
}

// ----------------------------------------------------- coarse level
// y[s, q] = sum over the cluster indices c of tile s of v[perm[c], q]
__global__ void k_restrict(int64_t n, int64_t nc, int64_t ncol, const int32_t* perm, const int32_t* tilebeg,
                           const double2* v, float2* y) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (s >= nc) return;
  double2 acc = mk<double>(0, 0);
  for (int c = tilebeg[s]; c < tilebeg[s + 1]; ++c) acc = cadd<double>(acc, v[q * n + perm[c]]);
  y[q * nc + s] = make_float2((float)acc.x, (float)acc.y);
}
// z[perm[c], q] = y[tile of c, q]
__global__ void k_prolong(int64_t n, int64_t nc, const int32_t* perm, const int32_t* rowtile, const float2* y,
                          double2* z) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (c >= n) return;
  const float2 v = y[q * nc + rowtile[c]];
  z[q * n + perm[c]] = mk<double>(v.x, v.y);
}
// Coarse operator of aggregates of `merge` consecutive tiles from the tile
// one: Ca[a, b] = sum over the tiles s of a and t of b of C[s, t], in FP32
// for the chain (nc tiles, na aggregates).
__global__ void k_coarse_merge(int64_t nc, int64_t na, int merge, const double2* C, float2* Ca) {
  const int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.y;
  if (a >= na) return;
  double2 acc = mk<double>(0, 0);
  for (int64_t t = b * merge; t < min(nc, (b + 1) * merge); ++t)
    for (int64_t s = a * merge; s < min(nc, (a + 1) * merge); ++s) acc = cadd<double>(acc, C[t * nc + s]);
  Ca[b * na + a] = make_float2((float)acc.x, (float)acc.y);
}
// r = v - w ;  z += d
__global__ void k_vsub(int64_t len, const double2* v, const double2* w, double2* r) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) r[i] = csub<double>(v[i], w[i]);
}
__global__ void k_vadd(int64_t len, double2* z, const double2* d) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) z[i] = cadd<double>(z[i], d[i]);
}
__global__ void k_c64_to_c32(int64_t len, const double2* a, float2* b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) b[i] = make_float2((float)a[i].x, (float)a[i].y);
}

// ------------------------------------------------------------ solver
// X2 = [sig2 (p) | h2perp (p) | h2par (2p, c + 2q)] from x = [sig1 | h1 | sig2
// | h2] (n x 8p, h columns c + 3q), the inputs of the substrate components.
__global__ void k_layit_gather(int64_t n, int64_t p, const double2* x, double2* X2) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  X2[q * n + i] = x[(4 * p + q) * n + i];
  X2[(p + q) * n + i] = x[(5 * p + 2 + 3 * q) * n + i];
  for (int c = 0; c < 2; ++c) X2[(2 * p + c + 2 * q) * n + i] = x[(5 * p + c + 3 * q) * n + i];
}

// y = A x with the substrate (bemretlayeriter/private/afun.m:3-46) from the
// products GX1 = G1 [s1 | h1], HX1 = H1 [s1 | h1] (n x 4p) and, for G2 and H2,
// Css = C.ss s2, Chs = C.hs s2, Csh = C.sh h2perp, Chh = C.hh h2perp (p
// columns each) and Cp = C.p h2par (2p); layout of y as b: phi | a | De |
// alpha.
__global__ void k_layit_apply(int64_t n, int64_t p, const double2* GX1, const double2* HX1, const double2* Gss,
                              const double2* Ghs, const double2* Gsh, const double2* Ghh, const double2* Gp,
                              const double2* Hss, const double2* Hhs, const double2* Hsh, const double2* Hhh,
                              const double2* Hp, const double* nvec, double k, double2 e1, double2 e2, double2* y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const double2 ik = mk<double>(0, k);
  const int64_t s = q * n + i;
  const double2 Gs1 = GX1[s], Hs1 = HX1[s];
  const double2 Gs2 = cadd<double>(Gss[s], Gsh[s]), Hs2 = cadd<double>(Hss[s], Hsh[s]);
  const double2 Gperp2 = cadd<double>(Ghh[s], Ghs[s]), Hperp2 = cadd<double>(Hhh[s], Hhs[s]);
  y[s] = csub<double>(Gs1, Gs2);
  const double2 w = csub<double>(cmul<double>(e1, Gs1), cmul<double>(e2, Gs2));  // for outer(n, G sig, eps)
  double2 nGh = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) {
    const int64_t col = c + 3 * q;
    const double nc = nvec[c * n + i];
    const double2 Gh1 = GX1[(p + col) * n + i], Hh1 = HX1[(p + col) * n + i];
    const double2 Gh2 = c < 2 ? Gp[(c + 2 * q) * n + i] : Gperp2;
    const double2 Hh2 = c < 2 ? Hp[(c + 2 * q) * n + i] : Hperp2;
    y[(p + col) * n + i] = csub<double>(Gh1, Gh2);
    y[(5 * p + col) * n + i] = csub<double>(csub<double>(Hh1, Hh2), cmul<double>(ik, cscale<double>(w, nc)));
    nGh = cadd<double>(nGh, cscale<double>(csub<double>(cmul<double>(e1, Gh1), cmul<double>(e2, Gh2)), nc));
  }
  y[(4 * p + q) * n + i] =
      csub<double>(csub<double>(cmul<double>(e1, Hs1), cmul<double>(e2, Hs2)), cmul<double>(ik, nGh));
}

struct BemRetIter {
  const GreenRet* g;
  int64_t n;
  std::unique_ptr<HMat> G1, G2, H1, H2;
  float2* prec = nullptr;  // npb x 5 x PB x PB
  int64_t* d_pb = nullptr;  // npb x 5 block descriptors (k_bjprec_build)
  int64_t npb = 0;
  int precision;
  int prec_mode = 2;  // 0 none, 1 block Jacobi, 2 two levels (coarse chain + block Jacobi)
  std::unique_ptr<Chain<float>> coarse;  // the chain on aggregates of the H-matrices' row tiles
  double* cnvec = nullptr;               // mean normals of the aggregates, na x 3
  int64_t nc = 0;                        // row tiles
  int merge = 1;                         // row tiles per aggregate
  int64_t na = 0;                        // aggregates
  int32_t* d_aggbeg = nullptr;           // first cluster index of each aggregate (na + 1)
  int32_t* d_rowagg = nullptr;           // aggregate of each cluster index (n)
  Workspace cws;
  double tol = 1e-6;
  int maxit = 100;
  double enei = 0, k0 = 0;
  double2 eps1{}, eps2{};
  bool ready = false;
  int last_iter = 0, last_flag = 0;
  double last_relres = 0, last_true_relres = 0;
  Workspace ws, hws, mws[4];  // mws: one per product of the batched H-matrix products
  // substrate: G2, H2 hold the p components, G2x, H2x the ss, hs, sh, hh ones
  GreenRetLayer* gl = nullptr;
  std::unique_ptr<HMat> G2x[4], H2x[4];
  double2 eps_down{};

  BemRetIter(const GreenRet* g_, int prec_, const int64_t* perm, int64_t nfull, const int64_t* full,
             int64_t nlow, const int64_t* low, double htol, int kmax, GreenRetLayer* gl_ = nullptr)
      : g(g_), n(g_->n), precision(prec_), gl(gl_) {
    if (gl)
      for (int c = 0; c < 4; ++c) {
        G2x[c].reset(new HMat(g, prec_, perm, nfull, full, nlow, low, htol, kmax));
        H2x[c].reset(new HMat(g, prec_, perm, nfull, full, nlow, low, htol, kmax));
      }
    G1.reset(new HMat(g, prec_, perm, nfull, full, nlow, low, htol, kmax));
    G2.reset(new HMat(g, prec_, perm, nfull, full, nlow, low, htol, kmax));
    H1.reset(new HMat(g, prec_, perm, nfull, full, nlow, low, htol, kmax));
    H2.reset(new HMat(g, prec_, perm, nfull, full, nlow, low, htol, kmax));
    // preconditioner blocks: each leaf's diagonal block cut into near-equal
    // runs of at most PB consecutive cluster indices
    std::vector<int64_t> pb;
    std::vector<int64_t> voff(G1->nfull + 1);
    for (int64_t b = 0; b < G1->nfull; ++b)
      voff[b + 1] = voff[b] + (int64_t)(G1->fblk_h[4 * b + 1] - G1->fblk_h[4 * b]) *
                                  (G1->fblk_h[4 * b + 3] - G1->fblk_h[4 * b + 2]);
    for (int64_t l = 0; l < G1->nleaf; ++l) {
      MNP_CHECK(G1->diag_h[l] >= 0, MNP_ERR_ARG, "bemretiter: a leaf cluster has no diagonal full block");
      const int64_t a = G1->lrow_h[2 * l], m = G1->lrow_h[2 * l + 1] - a;
      const int64_t nch = (m + PB - 1) / PB;
      for (int64_t c = 0; c < nch; ++c) {
        const int64_t o0 = c * m / nch, o1 = (c + 1) * m / nch;
        pb.insert(pb.end(), {a + o0, o1 - o0, o0, m, voff[G1->diag_h[l]]});
      }
    }
    npb = (int64_t)pb.size() / 5;
    d_pb = dev_alloc<int64_t>((int64_t)pb.size());
    MNP_HIP(hipMemcpy(d_pb, pb.data(), sizeof(int64_t) * pb.size(), hipMemcpyHostToDevice));
    prec = dev_alloc<float2>(npb * 5 * PB * PB);
    nc = G1->ntile;
    nv_h.resize(3 * n);
    perm_h.assign(perm, perm + n);
    MNP_HIP(hipMemcpy(nv_h.data(), g->geo.nvec, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    set_merge(2);  // measured on the 81,920-face sphere: 14 iterations instead of 10, 10% less time
  }
  std::vector<double> nv_h;
  std::vector<int64_t> perm_h;
  // Coarse level: aggregates of `m` consecutive row tiles (consecutive
  // cluster indices, so spatially compact) and their mean normals.
  void set_merge(int m) {
    MNP_CHECK(m >= 1 && m <= 64, MNP_ERR_ARG, "bemretiter: 1..64 tiles per coarse aggregate");
    if (m == merge && cnvec) return;
    merge = m;
    na = (nc + m - 1) / m;
    std::vector<double> cn(3 * na, 0.0);
    std::vector<int32_t> beg(na + 1), rowagg(n);
    for (int64_t a = 0; a < na; ++a) {
      const int64_t c0 = G1->tilebeg_h[a * m], c1 = G1->tilebeg_h[std::min<int64_t>(nc, (a + 1) * m)];
      beg[a] = (int32_t)c0;
      beg[a + 1] = (int32_t)c1;
      for (int64_t c = c0; c < c1; ++c) {
        rowagg[c] = (int32_t)a;
        for (int d = 0; d < 3; ++d) cn[d * na + a] += nv_h[d * n + perm_h[c]];
      }
      for (int d = 0; d < 3; ++d) cn[d * na + a] /= (double)(c1 - c0);
    }
    for (void* q : {(void*)cnvec, (void*)d_aggbeg, (void*)d_rowagg})
      if (q) MNP_HIP(hipFree(q));
    cnvec = dev_alloc<double>(3 * na);
    d_aggbeg = dev_alloc<int32_t>(na + 1);
    d_rowagg = dev_alloc<int32_t>(n);
    MNP_HIP(hipMemcpy(cnvec, cn.data(), sizeof(double) * 3 * na, hipMemcpyHostToDevice));
    MNP_HIP(hipMemcpy(d_aggbeg, beg.data(), sizeof(int32_t) * (na + 1), hipMemcpyHostToDevice));
    MNP_HIP(hipMemcpy(d_rowagg, rowagg.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    coarse.reset();
    ready = false;
  }
  ~BemRetIter() {
    for (void* q : {(void*)prec, (void*)d_pb, (void*)cnvec, (void*)d_aggbeg, (void*)d_rowagg})
      if (q) (void)hipFree(q);
  }

  // The coarse chain of this wavelength from the four filled H-matrices.
  void build_coarse(double2 e1, double2 e2) {
    if (!coarse) coarse.reset(new Chain<float>(na, false));
    Chain<float>& c = *coarse;
    c.eps1 = e1;
    c.eps2 = e2;
    c.k0 = k0;
    const int64_t nn = nc * nc;
    double2* C = (double2*)cws.get(2, sizeof(double2) * (size_t)nn);
    const HMat* M[4] = {G1.get(), G2.get(), H1.get(), H2.get()};
    float2* dst[4] = {c.G1, c.G2, c.S1, c.S2};
    for (int i = 0; i < 4; ++i) {
      M[i]->coarsen(C, cws);
      if (merge == 1)
        hipLaunchKernelGGL(k_c64_to_c32, dim3((unsigned)cdiv_up(nn, 256)), dim3(256), 0, stream(), nn, C, dst[i]);
      else
        hipLaunchKernelGGL(k_coarse_merge, dim3((unsigned)cdiv_up(na, 256), (unsigned)na), dim3(256), 0, stream(), nc,
                           na, merge, C, dst[i]);
      MNP_HIP(hipGetLastError());
    }
    chain_build<float>(c, cnvec);
    chain_explicit<float>(c);  // applied once per GMRES iteration: products, not solves
  }

  void initmat(double enei_, double2 e1, double2 e2) {
    ready = false;  // until the preconditioner is built
    enei = enei_;
    eps1 = e1;
    eps2 = e2;
    k0 = 2 * M_PI / enei;
    auto ksq = [&](double2 e) {
      const double r = std::hypot(e.x, e.y);
      double sr = std::sqrt(0.5 * (r + e.x)), si = std::sqrt(0.5 * (r - e.x));
      if (e.y < 0) si = -si;
      return mk<double>(k0 * sr, k0 * si);
    };
    const double2 k1 = ksq(e1), k2 = ksq(e2);
    G1->fill(k1, MNP_KEY_G);
    H1->fill(k1, MNP_KEY_H1);
    if (gl) {  // components p, ss, hs, sh, hh; the direct part in p, ss, hh
      const LayerPairs P = greenretlayer_pairs(gl, enei, e2, eps_down);
      G2->fill_layer(k2, MNP_KEY_G, 0, true, P);
      H2->fill_layer(k2, MNP_KEY_H2, 0, true, P);
      for (int c = 1; c < 5; ++c) {
        const bool direct = c == 1 || c == 4;
        G2x[c - 1]->fill_layer(k2, MNP_KEY_G, c, direct, P);
        H2x[c - 1]->fill_layer(k2, MNP_KEY_H2, c, direct, P);
      }
      greenretlayer_check(gl);
    } else {
      G2->fill(k2, MNP_KEY_G);
      H2->fill(k2, MNP_KEY_H2);
    }
    const double2 L = csub<double>(e1, e2);
    const double2 cc = cscale<double>(cmul<double>(L, L), k0 * k0);
    if (precision == MNP_FP64)
      hipLaunchKernelGGL(k_bjprec_build<double>, dim3((unsigned)npb), dim3(256), 0, stream(), d_pb,
                         (const double2*)G1->fvals, (const double2*)G2->fvals, (const double2*)H1->fvals,
                         (const double2*)H2->fvals, G1->perm, g->geo.nvec, n, e1, e2, cc, prec);
    else
      hipLaunchKernelGGL(k_bjprec_build<float>, dim3((unsigned)npb), dim3(256), 0, stream(), d_pb,
                         (const float2*)G1->fvals, (const float2*)G2->fvals, (const float2*)H1->fvals,
                         (const float2*)H2->fvals, G1->perm, g->geo.nvec, n, e1, e2, cc, prec);
    MNP_HIP(hipGetLastError());
    if (prec_mode == 2) build_coarse(e1, e2);
    ready = true;
  }

  // y = A x with the substrate: 12 H-matrix products in three batches
  void afun_layer(int64_t p, const double2* x, double2* y, double2* prod) {
    double2* X2 = prod + 20 * p * n;  // n x 4p, after GX1, HX1 (4p each) and the components (6p each)
    double2 *GX1 = prod, *HX1 = prod + 4 * p * n;
    double2* Gc = prod + 8 * p * n;  // ss, hs, sh, hh (p each), p (2p)
    double2* Hc = Gc + 6 * p * n;
    hipLaunchKernelGGL(k_layit_gather, dim3((unsigned)cdiv_up(n, 256), (unsigned)p), dim3(256), 0, stream(), n, p, x,
                       X2);
    MNP_HIP(hipGetLastError());
    const double2 *s2 = X2, *hperp = X2 + p * n, *hpar = X2 + 2 * p * n;
    Workspace* w[4] = {&mws[0], &mws[1], &mws[2], &mws[3]};
    {
      const HMat* A[4] = {G1.get(), H1.get(), G2.get(), H2.get()};
      const double2* xs[4] = {x, x, hpar, hpar};
      double2* ys[4] = {GX1, HX1, Gc + 4 * p * n, Hc + 4 * p * n};
      // G1, H1 take 4p columns, the p components 2p: two batches
      HMat::matvec_batch(2, A, 4 * p, xs, ys, w);
      HMat::matvec_batch(2, A + 2, 2 * p, xs + 2, ys + 2, w);
    }
    for (int side = 0; side < 2; ++side) {
      std::unique_ptr<HMat>* C = side ? H2x : G2x;
      double2* out = side ? Hc : Gc;
      const HMat* A[4] = {C[0].get(), C[1].get(), C[2].get(), C[3].get()};  // ss, hs, sh, hh
      const double2* xs[4] = {s2, s2, hperp, hperp};
      double2* ys[4] = {out, out + p * n, out + 2 * p * n, out + 3 * p * n};
      HMat::matvec_batch(4, A, p, xs, ys, w);
    }
    hipLaunchKernelGGL(k_layit_apply, dim3((unsigned)cdiv_up(n, 256), (unsigned)p), dim3(256), 0, stream(), n, p, GX1,
                       HX1, Gc, Gc + p * n, Gc + 2 * p * n, Gc + 3 * p * n, Gc + 4 * p * n, Hc, Hc + p * n,
                       Hc + 2 * p * n, Hc + 3 * p * n, Hc + 4 * p * n, g->geo.nvec, k0, eps1, eps2, y);
    MNP_HIP(hipGetLastError());
  }

  // y = A x (afun.m), x, y n x 8p
  void afun(int64_t p, const double2* x, double2* y, double2* prod) {
    if (gl) {
      afun_layer(p, x, y, prod);
      return;
    }
    double2 *GX1 = prod, *GX2 = prod + 4 * p * n, *HX1 = prod + 8 * p * n, *HX2 = prod + 12 * p * n;
    const HMat* A[4] = {G1.get(), G2.get(), H1.get(), H2.get()};
    const double2* xs[4] = {x, x + 4 * p * n, x, x + 4 * p * n};
    double2* ys[4] = {GX1, GX2, HX1, HX2};
    Workspace* w[4] = {&mws[0], &mws[1], &mws[2], &mws[3]};
    HMat::matvec_batch(4, A, 4 * p, xs, ys, w);
    launch_residual(n, p, nullptr, GX1, GX2, HX1, HX2, y, g->geo.nvec, k0, eps1, eps2);  // y = -A x
    const int64_t len = 8 * p * n;
    scale_vec(len, y, -1.0);
  }
  void bj_apply(int64_t p, const double2* b, double2* x) {
    hipLaunchKernelGGL(k_bjprec_apply, dim3((unsigned)npb), dim3(256), 0, stream(), d_pb, G1->perm, prec,
                       g->geo.nvec, n, (int)p, k0, eps1, eps2, b, x);
    MNP_HIP(hipGetLastError());
  }
  void mfun(int64_t p, const double2* b, double2* x, double2* prod) {
    const int64_t len = 8 * p * n;
    if (prec_mode == 0) {
      MNP_HIP(hipMemcpyAsync(x, b, sizeof(double2) * len, hipMemcpyDeviceToDevice, stream()));
      return;
    }
    if (prec_mode == 1) {
      bj_apply(p, b, x);
      return;
    }
    // x = P y + B^{-1} (b - A P y), y = C^{-1} P^T b
    float2* yb = (float2*)cws.get(3, sizeof(float2) * (size_t)(8 * p * na));
    float2* yx = (float2*)cws.get(4, sizeof(float2) * (size_t)(8 * p * na));
    float2* ytmp = (float2*)cws.get(5, sizeof(float2) * (size_t)(18 * p * na));
    double2* w = (double2*)cws.get(6, sizeof(double2) * (size_t)len);
    double2* d = (double2*)cws.get(7, sizeof(double2) * (size_t)len);
    hipLaunchKernelGGL(k_restrict, dim3((unsigned)cdiv_up(na, 256), (unsigned)(8 * p)), dim3(256), 0, stream(), n,
                       na, (int64_t)(8 * p), G1->perm, d_aggbeg, b, yb);
    MNP_HIP(hipGetLastError());
    chain_apply<float>(*coarse, cnvec, p, yb, yx, ytmp);
    hipLaunchKernelGGL(k_prolong, dim3((unsigned)cdiv_up(n, 256), (unsigned)(8 * p)), dim3(256), 0, stream(), n, na,
                       G1->perm, d_rowagg, yx, x);
    MNP_HIP(hipGetLastError());
    afun(p, x, w, prod);
    const unsigned gl = (unsigned)cdiv_up(len, 256);
    hipLaunchKernelGGL(k_vsub, dim3(gl), dim3(256), 0, stream(), len, b, w, w);
    bj_apply(p, w, d);
    hipLaunchKernelGGL(k_vadd, dim3(gl), dim3(256), 0, stream(), len, x, d);
    MNP_HIP(hipGetLastError());
  }

  void solve(int64_t p, const double2* const in[8], double2* sig1, double2* sig2, double2* h1, double2* h2) {
    MNP_CHECK(ready, MNP_ERR_STATE, "bemretiter_solve: initmat not called");
    MNP_CHECK(p >= 1 && p <= 8, MNP_ERR_ARG, "bemretiter_solve: npol must be in 1..8");
    const int64_t len = 8 * p * n;
    double2* b = (double2*)ws.get(0, sizeof(double2) * len);
    double2* x = (double2*)ws.get(6, sizeof(double2) * len);
    double2* prod = (double2*)hws.get(7, sizeof(double2) * (gl ? 24 : 16) * p * n);
    launch_excitation(n, p, in, b, g->geo.nvec, k0, eps1, eps2);
    const GmresOut o = fgmres(
        len, b, x, tol, maxit, [&](const double2* v, double2* w) { afun(p, v, w, prod); },
        [&](const double2* v, double2* z) { mfun(p, v, z, prod); }, ws);
    last_iter = o.iter;
    last_flag = o.flag;
    last_relres = o.relres;
    last_true_relres = o.true_relres;
    launch_unpack(n, p, x, sig1, h1, sig2, h2);
  }
};

}  // namespace mnp

// ============================================================== C ABI
using namespace mnp;

struct mnp_bemretiter_s {
  BemRetIter* b;
};

extern "C" {

int mnp_bemretiter_create(mnp_greenret_t g, int precision, const int64_t* perm, int64_t nfull,
                          const int64_t* full, int64_t nlow, const int64_t* low, double htol, int kmax,
                          mnp_bemretiter_t* out) {
  return guarded([&] {
    MNP_CHECK(!g || !g->g->compound(), MNP_ERR_ARG, "bemretiter_create: compound particles take the dense solver (mnp_bemret_*)");
    MNP_CHECK(g && perm && out, MNP_ERR_ARG, "bemretiter_create: NULL argument");
    auto* h = new mnp_bemretiter_s{nullptr};
    try {
      h->b = new BemRetIter(g->g, precision, perm, nfull, full, nlow, low, htol, kmax);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int mnp_bemretiter_create_layer(mnp_greenret_t g, struct mnp_greenretlayer_s* gl, int precision, const int64_t* perm,
                                int64_t nfull, const int64_t* full, int64_t nlow, const int64_t* low, double htol,
                                int kmax, mnp_bemretiter_t* out) {
  return guarded([&] {
    MNP_CHECK(g && gl && perm && out, MNP_ERR_ARG, "bemretiter_create_layer: NULL argument");
    MNP_CHECK(!g->g->compound(), MNP_ERR_ARG, "bemretiter_create_layer: compound particles are not supported");
    auto* h = new mnp_bemretiter_s{nullptr};
    try {
      h->b = new BemRetIter(g->g, precision, perm, nfull, full, nlow, low, htol, kmax, greenretlayer_of(gl));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int mnp_bemretiter_initmat_layer(mnp_bemretiter_t it, double enei, mnp_complex eps_in, mnp_complex eps_up,
                                 mnp_complex eps_down) {
  return guarded([&] {
    MNP_CHECK(it && enei > 0, MNP_ERR_ARG, "bemretiter_initmat_layer: bad argument");
    MNP_CHECK(it->b->gl, MNP_ERR_STATE, "bemretiter_initmat_layer: the solver was not created with a substrate");
    it->b->eps_down = to_d2(eps_down);
    it->b->initmat(enei, to_d2(eps_in), to_d2(eps_up));
  });
}

int mnp_bemretiter_destroy(mnp_bemretiter_t it) {
  return guarded([&] {
    if (!it) return;
    delete it->b;
    delete it;
  });
}

int mnp_bemretiter_set_options(mnp_bemretiter_t it, double tol, int maxit, int precond) {
  return guarded([&] {
    MNP_CHECK(it && tol > 0 && maxit >= 1, MNP_ERR_ARG, "bemretiter_set_options: bad argument");
    it->b->tol = tol;
    it->b->maxit = maxit;
    MNP_CHECK(precond >= 0 && precond <= 2, MNP_ERR_ARG, "bemretiter_set_options: precond must be 0, 1 or 2");
    it->b->prec_mode = precond;
  });
}

int mnp_bemretiter_set_coarse(mnp_bemretiter_t it, int tiles_per_aggregate) {
  return guarded([&] {
    MNP_CHECK(it, MNP_ERR_ARG, "bemretiter_set_coarse: NULL handle");
    it->b->set_merge(tiles_per_aggregate);
  });
}

int mnp_bemretiter_initmat(mnp_bemretiter_t it, double enei, mnp_complex eps1, mnp_complex eps2) {
  return guarded([&] {
    MNP_CHECK(it && enei > 0, MNP_ERR_ARG, "bemretiter_initmat: bad argument");
    MNP_CHECK(!it->b->gl, MNP_ERR_STATE, "bemretiter_initmat: a substrate solver takes mnp_bemretiter_initmat_layer");
    it->b->initmat(enei, to_d2(eps1), to_d2(eps2));
  });
}

int mnp_bemretiter_solve(mnp_bemretiter_t it, int64_t npol, const void* phi1, const void* phi1p,
                         const void* a1, const void* a1p, const void* phi2, const void* phi2p,
                         const void* a2, const void* a2p, void* sig1, void* sig2, void* h1, void* h2) {
  return guarded([&] {
    MNP_CHECK(it, MNP_ERR_ARG, "bemretiter_solve: NULL handle");
    const double2* in[8] = {(const double2*)phi1, (const double2*)phi1p, (const double2*)a1,
                            (const double2*)a1p,  (const double2*)phi2,  (const double2*)phi2p,
                            (const double2*)a2,   (const double2*)a2p};
    it->b->solve(npol, in, (double2*)sig1, (double2*)sig2, (double2*)h1, (double2*)h2);
  });
}

int mnp_bemretiter_stats(mnp_bemretiter_t it, int* flag, double* relres, int* iter, double* true_relres,
                         double* compression) {
  return guarded([&] {
    MNP_CHECK(it, MNP_ERR_ARG, "bemretiter_stats: NULL handle");
    const BemRetIter& B = *it->b;
    if (flag) *flag = B.last_flag;
    if (relres) *relres = B.last_relres;
    if (iter) *iter = B.last_iter;
    if (true_relres) *true_relres = B.last_true_relres;
    if (compression)
      *compression = B.G1->filled ? (double)(B.G1->nfullvals + B.G1->nlowvals) / ((double)B.n * B.n) : 0.0;
  });
}

}  // extern "C"
