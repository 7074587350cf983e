// Batched dense retarded BEM solves on gfx950 (MNPBEM BEM/@bemret over many
// particles and wavelengths, as in the spectra loops of
// Demo/planewave/ret/demospecret1.m:34-42).
//
// A batch holds `cap` members of one face count n, each a (particle,
// wavelength) pair with its own mesh and dielectric functions.  Every stage
// runs for all members in one launch per step (grid z = problem):
//   - assembly of G(k1), G(k2), F(k1) + 2 pi, F(k2) - 2 pi for the 2 cap
//     (member, side) problems, each reading its mesh from a descriptor, into
//     the chain's matrices and an FP64 copy (the exact operator);
//   - the bemret chain (BEM/@bemret/private/initmat.m:5-72) by blocked
//     right-looking algorithms without row exchanges: LU(G1), LU(G2) and
//     Sigma_s = H_s G_s^{-1} as 2 cap problems, D = Sigma1 - Sigma2, LU(D),
//     Deltai = D^{-1}, Sigma and LU(Sigma) as cap problems; panels of NB
//     columns factored by 32-column steps (a diagonal-block kernel and up to
//     three products), then one product of inner dimension NB for the
//     trailing matrix, all on the strided batched MFMA GEMM;
//   - mldivide (BEM/@bemret/mldivide.m:11-49) with blocked triangular solves;
//   - refinement with the exact FP64 operator (BEM/@bemretiter/private/
//     afun.m:6-35): x <- x + M^{-1} (b - A x) for a fixed number of sweeps
//     (the chain in complex FP32 for MNP_SOLVER_MIXED, 3 sweeps; in complex
//     FP64 for MNP_SOLVER_FP64, 1 sweep), then the relative residual per
//     member;
//   - far fields of the outside charges for the plane-wave directions and the
//     pinfty normals (@spectrumret/farfield.m:59-84).
// The pipeline is captured once as a hipGraph and replayed per batch
// (parameters and mesh descriptors are read from device memory, written
// before each replay).  A member with a zero pivot or a residual above 1e-12
// is reported, not solved; the caller solves it with the dense solver.
// Without row exchanges the FP64 factors are not LAPACK's whenever partial
// pivoting would exchange rows (the dense FP64 solver then refactors); the
// refinement sweep with the exact operator makes the solution agree with the
// pivoted one to the tolerance instead.
//
// The panel width NB and the tile of the large products (64 x 64, 128 x 64,
// or 128 x 96 on the 3-multiplication form for complex64) are chosen per
// (n, cap, solver) by timing the captured pipeline with each candidate once
// (BVARIANTS, tuned_variant).
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "mnp_chain.hpp"
#include "mnp_linalg.hpp"
#include "mnp_ludiag.hpp"

namespace mnp {

void farfield_host_sums(int64_t npol, const double* pol, int64_t nd, int64_t npinf, const double2* he,
                        const double2* hh, const double* pn, const double* pa, double2 k, double nb, double* ext,
                        double* sca, double* abs_);

// Mesh of one member: device arrays of a GreenRet.
struct BGeo {
  const double* pos;
  const double* nvec;
  const double* area;
  const int32_t* prow;
  const int32_t* pcol;
  const int8_t* pkind;
  const double* gtab;
  const double* ftab;
  int64_t nnz;
  int order;
};
// Wavelength-dependent parameters of one member.
struct BPrm {
  double k0;
  double2 k1, k2;  // inside, outside wavenumbers
  double2 e1, e2;  // inside, outside dielectric functions
  double2 c;       // k0^2 (e1 - e2)^2
};

// ------------------------------------------------------------ assembly
// Problem z = 2 b + s: s = 0 writes G(k1) and F(k1) + 2 pi delta, s = 1
// G(k2) and F(k2) - 2 pi delta of member b (greenret/private/eval1.m:24-40,
// :110-124), into G + z n^2, H + z n^2 (and the FP64 copies when Gd, Hd are
// set); refined elements follow in k_bassemble_refined.
template <class T>
__global__ void __launch_bounds__(256) k_bassemble(int64_t n, const BGeo* __restrict__ geos,
                                                   const BPrm* __restrict__ prm, cplx<T>* G, cplx<T>* H,
                                                   double2* Gd, double2* Hd) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  const int z = blockIdx.z, b = z >> 1, s = z & 1;
  if (i >= n) return;
  const double* pos = geos[b].pos;
  const double* nvec = geos[b].nvec;
  const double aj = geos[b].area[j];
  const double2 k = s ? prm[b].k2 : prm[b].k1;
  const double dx = pos[i] - pos[j], dy = pos[n + i] - pos[n + j], dz = pos[2 * n + i] - pos[2 * n + j];
  const double d = fmax(sqrt(dx * dx + dy * dy + dz * dz), 2.220446049250313e-16);
  const double2 ph = cexp_ikd(k, d);
  const double gr = (1.0 / d) * aj;
  const double2 g = mk<double>(gr * ph.x, gr * ph.y);
  const double id2 = 1.0 / (d * d);
  const double2 fk = mk<double>((-k.y - 1.0 / d) * id2, k.x * id2);
  const double in = dx * nvec[i] + dy * nvec[n + i] + dz * nvec[2 * n + i];
  double2 F = cmul<double>(mk<double>(fk.x * in * aj, fk.y * in * aj), ph);
  if (i == j) F.x += s ? -2 * M_PI : 2 * M_PI;
  const int64_t idx = (int64_t)z * n * n + j * n + i;
  G[idx] = mk<T>((T)g.x, (T)g.y);
  H[idx] = mk<T>((T)F.x, (T)F.y);
  if (Gd) {
    Gd[idx] = g;
    Hd[idx] = F;
  }
}

// Refined pairs of each member (Taylor tables of greenret/private/init.m):
// G = (sum_o g_o (ik)^o) e^{ikd}, F = (n_i . sum_o f_o (ik)^o) e^{ikd} (+- 2 pi
// on self pairs); a grid-stride loop covers any member's pair count.
template <class T>
__global__ void __launch_bounds__(256) k_bassemble_refined(int64_t n, const BGeo* __restrict__ geos,
                                                           const BPrm* __restrict__ prm, cplx<T>* G, cplx<T>* H,
                                                           double2* Gd, double2* Hd) {
  const int z = blockIdx.z, b = z >> 1, s = z & 1;
  const BGeo& g = geos[b];
  const int64_t nnz = g.nnz;
  const int order = g.order;
  const double2 k = s ? prm[b].k2 : prm[b].k1;
  const double sgn2pi = s ? -2 * M_PI : 2 * M_PI;
  const double2 ik = mk<double>(-k.y, k.x);
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = g.prow[p], j = g.pcol[p];
    const double dx = g.pos[i] - g.pos[j], dy = g.pos[n + i] - g.pos[n + j], dz = g.pos[2 * n + i] - g.pos[2 * n + j];
    const double d = fmax(sqrt(dx * dx + dy * dy + dz * dz), 2.220446049250313e-16);
    const double2 ph = cexp_ikd(k, d);
    const double n0 = g.nvec[i], n1 = g.nvec[n + i], n2 = g.nvec[2 * n + i];
    double2 pw = mk<double>(1, 0), Gs = mk<double>(0, 0), Fs = mk<double>(0, 0);
    for (int o = 0; o <= order; ++o) {
      const double go = g.gtab[o * nnz + p];
      const double fo = n0 * g.ftab[(o * 3 + 0) * nnz + p] + n1 * g.ftab[(o * 3 + 1) * nnz + p] +
                        n2 * g.ftab[(o * 3 + 2) * nnz + p];
      Gs = cadd<double>(Gs, cscale<double>(pw, go));
      Fs = cadd<double>(Fs, cscale<double>(pw, fo));
      pw = cmul<double>(pw, ik);
    }
    const double2 gv = cmul<double>(Gs, ph);
    double2 hv = cmul<double>(Fs, ph);
    if (g.pkind[p] == 2) hv.x += sgn2pi;
    const int64_t idx = (int64_t)z * n * n + j * n + i;
    G[idx] = mk<T>((T)gv.x, (T)gv.y);
    H[idx] = mk<T>((T)hv.x, (T)hv.y);
    if (Gd) {
      Gd[idx] = gv;
      Hd[idx] = hv;
    }
  }
}

// ------------------------------------------------------ factorisations
// Diagonal block of problem z = blockIdx.x: A + z sa (ld lda), status word
// info[z * istride] (first zero pivot + 1), inverses at inv + z sinv (L^{-1})
// and inv + z sinv + uoff (U^{-1}).
__global__ void __launch_bounds__(192) k_blu_diag_f(int nb, float2* A, int64_t lda, int64_t sa, int* info,
                                                    int istride, int info_off, float2* inv, int64_t sinv,
                                                    int64_t uoff) {
  const int64_t z = blockIdx.x;
  lu_nopiv_diag_body<float>(nb, A + z * sa, lda, info + z * istride, info_off, inv + z * sinv,
                            inv + z * sinv + uoff);
}
__global__ void __launch_bounds__(256) k_blu_diag_d(int nb, double2* A, int64_t lda, int64_t sa, int* info,
                                                    int istride, int info_off, double2* inv, int64_t sinv,
                                                    int64_t uoff) {
  const int64_t z = blockIdx.x;
  lu_nopiv_diag_lds_body<double>(nb, A + z * sa, lda, info + z * istride, info_off, inv + z * sinv,
                                 inv + z * sinv + uoff);
}

__global__ void k_bclear(int64_t len, int* a) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) a[i] = 0;
}

// D_b = S_{2b} - S_{2b+1};  X_b = I
template <class T>
__global__ void k_bsub(int64_t n, const cplx<T>* S, cplx<T>* D) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y, b = blockIdx.z, nn = n * n;
  if (i >= n) return;
  D[b * nn + j * n + i] = csub<T>(S[2 * b * nn + j * n + i], S[(2 * b + 1) * nn + j * n + i]);
}
template <class T>
__global__ void k_bident(int64_t n, cplx<T>* X) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y, b = blockIdx.z;
  if (i >= n) return;
  X[b * n * n + j * n + i] = mk<T>(i == j ? T(1) : T(0), T(0));
}
// Sigma = e1 Sigma1 - e2 Sigma2 + k0^2 (e1 - e2)^2 (Deltai o n n^T) over
// S_{2b+1} (initmat.m:55-56 with scalar L1 = e1, L2 = e2)
template <class T>
__global__ void k_bsigma(int64_t n, cplx<T>* S, const cplx<T>* Di, const BGeo* __restrict__ geos,
                         const BPrm* __restrict__ prm) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y, b = blockIdx.z, nn = n * n;
  if (i >= n) return;
  const double* nv = geos[b].nvec;
  const double nnv = nv[i] * nv[j] + nv[n + i] * nv[n + j] + nv[2 * n + i] * nv[2 * n + j];
  const BPrm& q = prm[b];
  const cplx<T> s1 = S[2 * b * nn + j * n + i], s2 = S[(2 * b + 1) * nn + j * n + i], di = Di[b * nn + j * n + i];
  const double2 a = cmul<double>(q.e1, mk<double>(s1.x, s1.y));
  const double2 c = cmul<double>(q.e2, mk<double>(s2.x, s2.y));
  const double2 d = cmul<double>(q.c, mk<double>(di.x * nnv, di.y * nnv));
  S[(2 * b + 1) * nn + j * n + i] = mk<T>((T)(a.x - c.x + d.x), (T)(a.y - c.y + d.y));
}

// ------------------------------------------------------------ vectors
// Member b's vectors are contiguous blocks: b, x of n x 8p (phi | a | De |
// alpha ; sig1 | h1 | sig2 | h2, components fastest), the chain's
// temporaries of n x 18p.
//
// b of a plane wave in the outside medium (planewaveret/potential.m:47-88,
// excitation.m:21-29 with phi = 0): a = a2, De = -i k0 e2 n . a2, alpha = a2p,
// a2 = e^{i k2 pos.dir} pol / (i k0), a2p = i k2 (n . dir) a2.
__global__ void k_bexcite(int64_t n, int64_t p, const double* __restrict__ pol, const double* __restrict__ dir,
                          const BGeo* __restrict__ geos, const BPrm* __restrict__ prm, double2* bv) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y, b = blockIdx.z;
  if (i >= n) return;
  const double* pos = geos[b].pos;
  const double* nvec = geos[b].nvec;
  const BPrm& pr = prm[b];
  double2* B = bv + b * 8 * p * n;
  double pd = 0, nd = 0;
  for (int c = 0; c < 3; ++c) {
    pd += pos[c * n + i] * dir[c * p + q];
    nd += nvec[c * n + i] * dir[c * p + q];
  }
  const double2 ph = cdiv<double>(cexp_ikd(pr.k2, pd), mk<double>(0, pr.k0));
  const double2 php = cmul<double>(cmul<double>(mk<double>(0, nd), pr.k2), ph);
  double na = 0;
  double2 nav = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) {
    const double pc = pol[c * p + q];
    const int64_t col = c + 3 * q;
    B[(p + col) * n + i] = cscale<double>(ph, pc);
    B[(5 * p + col) * n + i] = cscale<double>(php, pc);
    na += nvec[c * n + i] * pc;
  }
  nav = cscale<double>(ph, na);
  B[q * n + i] = mk<double>(0, 0);
  // De = -i k0 e2 (n . a2)
  const double2 t = cmul<double>(pr.e2, nav);
  B[(4 * p + q) * n + i] = mk<double>(pr.k0 * t.y, -pr.k0 * t.x);
}

// mldivide steps (mnp_chain.hpp's k_mld1..4 with the member's L1 = e1, L2 = e2)
template <class T>
__global__ void k_bmld1(int64_t n, int64_t p, const cplx<T>* bv, const cplx<T>* tmp, cplx<T>* tmpw,
                        const BGeo* __restrict__ geos, const BPrm* __restrict__ prm) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y, b = blockIdx.z;
  if (i >= n) return;
  const cplx<T>* B = bv + b * 8 * p * n;
  const cplx<T>* Y = tmp + b * 18 * p * n;
  cplx<T>* alpha = tmpw + b * 18 * p * n + 4 * p * n;
  cplx<T>* De = alpha + 3 * p * n;
  const double* nvec = geos[b].nvec;
  const double2 L1 = prm[b].e1;
  const double2 ikL1 = cmul<double>(mk<double>(0, prm[b].k0), L1);
  const cplx<T> ph = B[q * n + i];
  const double2 ikL1phi = cmul<double>(ikL1, mk<double>(ph.x, ph.y));
  double2 na = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) {
    const int64_t col = c + 3 * q;
    const double nc = nvec[c * n + i];
    const cplx<T> ya = Y[col * n + i], a0 = B[(5 * p + col) * n + i], ac = B[(p + col) * n + i];
    alpha[col * n + i] = mk<T>((T)(a0.x - ya.x + nc * ikL1phi.x), (T)(a0.y - ya.y + nc * ikL1phi.y));
    na.x += nc * ac.x;
    na.y += nc * ac.y;
  }
  const cplx<T> yp = Y[(3 * p + q) * n + i], d0 = B[(4 * p + q) * n + i];
  const double2 l1y = cmul<double>(L1, mk<double>(yp.x, yp.y));
  const double2 t = cmul<double>(ikL1, na);
  De[q * n + i] = mk<T>((T)(d0.x - l1y.x + t.x), (T)(d0.y - l1y.y + t.y));
}
// rhs = De + i k0 (L1 - L2) n . t   (into the s2 slot)
template <class T>
__global__ void k_bmld2(int64_t n, int64_t p, cplx<T>* tmp, const BGeo* __restrict__ geos,
                        const BPrm* __restrict__ prm) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y, b = blockIdx.z;
  if (i >= n) return;
  cplx<T>* base = tmp + b * 18 * p * n;
  const cplx<T>* De = base + 7 * p * n;
  const cplx<T>* t = base + 8 * p * n;
  cplx<T>* rhs = base + 11 * p * n;
  const double* nvec = geos[b].nvec;
  const double2 ikdL = cmul<double>(mk<double>(0, prm[b].k0), csub<double>(prm[b].e1, prm[b].e2));
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
// u = i k0 (L1 - L2) n (x) sig2 + alpha
template <class T>
__global__ void k_bmld3(int64_t n, int64_t p, cplx<T>* tmp, const BGeo* __restrict__ geos,
                        const BPrm* __restrict__ prm) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y, b = blockIdx.z;
  if (i >= n) return;
  cplx<T>* base = tmp + b * 18 * p * n;
  const cplx<T>* alpha = base + 4 * p * n;
  const cplx<T>* sig2 = base + 11 * p * n;
  cplx<T>* u = base + 12 * p * n;
  const double* nvec = geos[b].nvec;
  const double2 ikdL = cmul<double>(mk<double>(0, prm[b].k0), csub<double>(prm[b].e1, prm[b].e2));
  const cplx<T> s = sig2[q * n + i];
  const double2 w = cmul<double>(ikdL, mk<double>(s.x, s.y));
  for (int c = 0; c < 3; ++c) {
    const int64_t col = c + 3 * q;
    const cplx<T> a = alpha[col * n + i];
    const double nc = nvec[c * n + i];
    u[col * n + i] = mk<T>((T)(a.x + nc * w.x), (T)(a.y + nc * w.y));
  }
}
// x = [sig2 + phi | h2 + a | sig2 | h2]
template <class T>
__global__ void k_bmld4(int64_t n, int64_t p, const cplx<T>* bv, const cplx<T>* tmp, cplx<T>* xv) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t col = blockIdx.y, b = blockIdx.z;  // col 0 .. 4p
  if (i >= n) return;
  const cplx<T>* base = tmp + b * 18 * p * n;
  const cplx<T> src = col < p ? base[11 * p * n + col * n + i] : base[15 * p * n + (col - p) * n + i];
  const cplx<T> add = bv[b * 8 * p * n + col * n + i];
  cplx<T>* X = xv + b * 8 * p * n;
  X[col * n + i] = cadd<T>(src, add);
  X[(4 * p + col) * n + i] = src;
}

// r = b - A x with GX_{2b+s} = G_s x_s, HX_{2b+s} = H_s x_s (afun.m)
__global__ void k_bresidual(int64_t n, int64_t p, const double2* bv, const double2* GX, const double2* HX,
                            double2* rv, const BGeo* __restrict__ geos, const BPrm* __restrict__ prm) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y, b = blockIdx.z;
  if (i >= n) return;
  const int64_t np = 4 * p * n;
  const double2* B = bv + b * 8 * p * n;
  double2* r = rv + b * 8 * p * n;
  const double2 *GX1 = GX + 2 * b * np, *GX2 = GX1 + np, *HX1 = HX + 2 * b * np, *HX2 = HX1 + np;
  const double* nvec = geos[b].nvec;
  const double2 e1 = prm[b].e1, e2 = prm[b].e2, ik = mk<double>(0, prm[b].k0);
  const double2 Gs1 = GX1[q * n + i], Gs2 = GX2[q * n + i], Hs1 = HX1[q * n + i], Hs2 = HX2[q * n + i];
  r[q * n + i] = csub<double>(B[q * n + i], csub<double>(Gs1, Gs2));
  const double2 w = csub<double>(cmul<double>(e1, Gs1), cmul<double>(e2, Gs2));
  double2 nGh = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) {
    const int64_t col = c + 3 * q;
    const double nc = nvec[c * n + i];
    const double2 Gh1 = GX1[(p + col) * n + i], Gh2 = GX2[(p + col) * n + i];
    const double2 Hh1 = HX1[(p + col) * n + i], Hh2 = HX2[(p + col) * n + i];
    r[(p + col) * n + i] = csub<double>(B[(p + col) * n + i], csub<double>(Gh1, Gh2));
    const double2 al = csub<double>(csub<double>(Hh1, Hh2), cmul<double>(ik, cscale<double>(w, nc)));
    r[(5 * p + col) * n + i] = csub<double>(B[(5 * p + col) * n + i], al);
    nGh = cadd<double>(nGh, cscale<double>(csub<double>(cmul<double>(e1, Gh1), cmul<double>(e2, Gh2)), nc));
  }
  const double2 De =
      csub<double>(csub<double>(cmul<double>(e1, Hs1), cmul<double>(e2, Hs2)), cmul<double>(ik, nGh));
  r[(4 * p + q) * n + i] = csub<double>(B[(4 * p + q) * n + i], De);
}

// |r_b| / |b_b| per member, one workgroup per member, fixed summation order
__global__ void __launch_bounds__(256) k_brelres(int64_t len, const double2* r, const double2* bv, double* out) {
  const int64_t b = blockIdx.x;
  double sr = 0, sb = 0;
  for (int64_t i = threadIdx.x; i < len; i += blockDim.x) {
    const double2 u = r[b * len + i], v = bv[b * len + i];
    sr += u.x * u.x + u.y * u.y;
    sb += v.x * v.x + v.y * v.y;
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = sr;
  red[1][threadIdx.x] = sb;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b] = red[1][0] > 0 ? sqrt(red[0][0] / red[1][0]) : 0.0;
}

template <class Ts, class Td>
__global__ void k_bconvert(int64_t len, const cplx<Ts>* a, cplx<Td>* b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) b[i] = mk<Td>((Td)a[i].x, (Td)a[i].y);
}
__global__ void k_baxpy(int64_t len, double2* x, const double2* d) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) x[i] = cadd<double>(x[i], d[i]);
}

// Far fields of member b's outside charges sig2, h2 (x columns 4p .. 8p) in
// the ndir directions dir (3 x ndir, column-major): workgroup (d, b),
// e = i k0 sum ph h2 - i k2 dir sum ph sig2, h = i k2 dir x sum ph h2,
// ph = e^{-i k2 dir . pos} area (@spectrumret/farfield.m:59-84).
constexpr int BFF_MAXP = 8;
__global__ void __launch_bounds__(256) k_bfarfield(int64_t n, int64_t p, int64_t ndir, const double* __restrict__ dir,
                                                   const BGeo* __restrict__ geos, const BPrm* __restrict__ prm,
                                                   const double2* xv, double2* ev, double2* hv) {
  const int64_t d = blockIdx.x, b = blockIdx.z;
  const double dx = dir[d], dy = dir[ndir + d], dz = dir[2 * ndir + d];
  const double* pos = geos[b].pos;
  const double* area = geos[b].area;
  const double2 k = prm[b].k2;
  const double k0 = prm[b].k0;
  const double2* sig2 = xv + b * 8 * p * n + 4 * p * n;
  const double2* h2 = sig2 + p * n;
  double2 acc[BFF_MAXP][4];
  for (int q = 0; q < BFF_MAXP; ++q)
    for (int c = 0; c < 4; ++c) acc[q][c] = mk<double>(0, 0);
  const double2 mk_ = mk<double>(-k.x, -k.y);
  for (int64_t j = threadIdx.x; j < n; j += blockDim.x) {
    const double pd = dx * pos[j] + dy * pos[n + j] + dz * pos[2 * n + j];
    const double2 ph = cscale<double>(cexp_ikd(mk_, pd), area[j]);
    for (int q = 0; q < p && q < BFF_MAXP; ++q) {
      for (int c = 0; c < 3; ++c) acc[q][c] = cadd<double>(acc[q][c], cmul<double>(ph, h2[(c + 3 * q) * n + j]));
      acc[q][3] = cadd<double>(acc[q][3], cmul<double>(ph, sig2[q * n + j]));
    }
  }
  __shared__ double2 red[256];
  for (int q = 0; q < p && q < BFF_MAXP; ++q)
    for (int c = 0; c < 4; ++c) {
      red[threadIdx.x] = acc[q][c];
      __syncthreads();
      for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = cadd<double>(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
      }
      if (threadIdx.x == 0) acc[q][c] = red[0];
      __syncthreads();
    }
  if (threadIdx.x == 0) {
    double2* e = ev + b * ndir * 3 * p;
    double2* h = hv + b * ndir * 3 * p;
    const double2 ik0 = mk<double>(0, k0), ikk = mk<double>(-k.y, k.x);
    const double dv[3] = {dx, dy, dz};
    for (int q = 0; q < p && q < BFF_MAXP; ++q) {
      const double2 s = acc[q][3];
      for (int c = 0; c < 3; ++c)
        e[d + ndir * (c + 3 * q)] =
            csub<double>(cmul<double>(ik0, acc[q][c]), cscale<double>(cmul<double>(ikk, s), dv[c]));
      const double2 A0 = acc[q][0], A1 = acc[q][1], A2 = acc[q][2];
      h[d + ndir * (0 + 3 * q)] = cmul<double>(ikk, csub<double>(cscale<double>(A2, dy), cscale<double>(A1, dz)));
      h[d + ndir * (1 + 3 * q)] = cmul<double>(ikk, csub<double>(cscale<double>(A0, dz), cscale<double>(A2, dx)));
      h[d + ndir * (2 + 3 * q)] = cmul<double>(ikk, csub<double>(cscale<double>(A1, dx), cscale<double>(A0, dy)));
    }
  }
}

// ------------------------------------------------- blocked algorithms
static const double2 B_ONE = {1, 0}, B_MONE = {-1, 0}, B_ZERO = {0, 0};

// Tile of a product: the tuned tile for large ones, 64 x 64 otherwise.
static int big_tile(int64_t m, int64_t n, int tile) { return (m >= 128 && n >= 128) ? tile : GEMM_TILE_64; }

inline dim3 bgrid(int64_t n, int64_t y, int64_t z) {
  return dim3((unsigned)cdiv_up(n, 256), (unsigned)y, (unsigned)z);
}

// In-place LU without row exchanges of P problems A + z sa (n x n, ld n),
// right-looking over panels of NB columns (NB a multiple of 32): the panel
// is factored by 32-column steps (the diagonal block and its L^{-1}, U^{-1}
// into inv + z sinv, U^{-1} at + uoff; L21 = A21 U11^{-1}, in place since one
// output tile spans the 32 columns; the update of the panel's own columns),
// then the panel's block row U12 = L11^{-1} A12 by forward substitution and
// the trailing update A22 -= L21 U12 with inner dimension NB.
template <class T>
static void blu(int64_t n, int64_t P, cplx<T>* A, int64_t sa, cplx<T>* inv, int64_t sinv, int* info, int istride,
                int tile, int NB) {
  const int64_t uoff = inv_blocks(n);
  auto at = [&](int64_t i, int64_t j) { return MatView<T>{A + i + j * n, n}; };
  for (int64_t J0 = 0; J0 < n; J0 += NB) {
    const int64_t W = std::min<int64_t>(NB, n - J0);
    for (int64_t j0 = J0; j0 < J0 + W; j0 += TB) {
      const int nb = (int)std::min<int64_t>(TB, J0 + W - j0);
      cplx<T>* Li = inv + (j0 / TB) * TB2;
      if constexpr (sizeof(T) == 4)
        hipLaunchKernelGGL(k_blu_diag_f, dim3((unsigned)P), dim3(192), 0, stream(), nb, A + j0 + j0 * n, n, sa, info,
                           istride, (int)j0, Li, sinv, uoff);
      else
        hipLaunchKernelGGL(k_blu_diag_d, dim3((unsigned)P), dim3(256), 0, stream(), nb, A + j0 + j0 * n, n, sa, info,
                           istride, (int)j0, Li, sinv, uoff);
      MNP_HIP(hipGetLastError());
      const int64_t rr = n - j0 - nb, rc = J0 + W - j0 - nb;
      if (rr <= 0) continue;
      gemm_strided<T>(rr, nb, nb, B_ONE, at(j0 + nb, j0), sa, MatView<T>{Li + uoff, TB}, sinv, B_ZERO,
                      at(j0 + nb, j0), sa, P, GEMM_TILE_64);
      if (rc <= 0) continue;
      gemm_strided<T>(nb, rc, nb, B_ONE, MatView<T>{Li, TB}, sinv, at(j0, j0 + nb), sa, B_ZERO, at(j0, j0 + nb), sa,
                      P, GEMM_TILE_64);
      gemm_strided<T>(rr, rc, nb, B_MONE, at(j0 + nb, j0), sa, at(j0, j0 + nb), sa, B_ONE, at(j0 + nb, j0 + nb), sa,
                      P, big_tile(rr, rc, tile));
    }
    const int64_t R = n - J0 - W;
    if (R <= 0) continue;
    for (int64_t j0 = J0; j0 < J0 + W; j0 += TB) {
      const int nb = (int)std::min<int64_t>(TB, J0 + W - j0);
      gemm_strided<T>(nb, R, nb, B_ONE, MatView<T>{inv + (j0 / TB) * TB2, TB}, sinv, at(j0, J0 + W), sa, B_ZERO,
                      at(j0, J0 + W), sa, P, GEMM_TILE_64);
      const int64_t rp = J0 + W - j0 - nb;
      if (rp > 0)
        gemm_strided<T>(rp, R, nb, B_MONE, at(j0 + nb, j0), sa, at(j0, J0 + W), sa, B_ONE, at(j0 + nb, J0 + W), sa, P,
                        big_tile(rp, R, tile));
    }
    gemm_strided<T>(R, R, W, B_MONE, at(J0 + W, J0), sa, at(J0, J0 + W), sa, B_ONE, at(J0 + W, J0 + W), sa, P,
                    big_tile(R, R, tile));
  }
}

// X <- X A^{-1} for P problems (X m x n at X + z sx, ld ldx; A = L U at
// LU + z slu with its block inverses), over panels of NB columns: X U = B by
// panels ascending (32-column steps inside a panel, then the update of the
// columns right of it with inner dimension NB), then X L = Y by panels
// descending.
template <class T>
static void bright(int64_t m, int64_t n, int64_t P, cplx<T>* X, int64_t ldx, int64_t sx, cplx<T>* LU, int64_t slu,
                   cplx<T>* inv, int64_t sinv, int tile, int NB) {
  const int64_t uoff = inv_blocks(n);
  auto xv = [&](int64_t j) { return MatView<T>{X + j * ldx, ldx}; };
  auto lu = [&](int64_t i, int64_t j) { return MatView<T>{LU + i + j * n, n}; };
  for (int64_t J0 = 0; J0 < n; J0 += NB) {
    const int64_t W = std::min<int64_t>(NB, n - J0);
    for (int64_t j0 = J0; j0 < J0 + W; j0 += TB) {
      const int nb = (int)std::min<int64_t>(TB, J0 + W - j0);
      gemm_strided<T>(m, nb, nb, B_ONE, xv(j0), sx, MatView<T>{inv + uoff + (j0 / TB) * TB2, TB}, sinv, B_ZERO,
                      xv(j0), sx, P, GEMM_TILE_64);
      const int64_t rc = J0 + W - j0 - nb;
      if (rc > 0)
        gemm_strided<T>(m, rc, nb, B_MONE, xv(j0), sx, lu(j0, j0 + nb), slu, B_ONE, xv(j0 + nb), sx, P,
                        big_tile(m, rc, tile));
    }
    const int64_t R = n - J0 - W;
    if (R > 0)
      gemm_strided<T>(m, R, W, B_MONE, xv(J0), sx, lu(J0, J0 + W), slu, B_ONE, xv(J0 + W), sx, P,
                      big_tile(m, R, tile));
  }
  const int64_t npan = cdiv_up(n, NB);
  for (int64_t pj = npan - 1; pj >= 0; --pj) {
    const int64_t J0 = pj * NB, W = std::min<int64_t>(NB, n - J0);
    for (int64_t j0 = J0 + (W - 1) / TB * TB; j0 >= J0; j0 -= TB) {
      const int nb = (int)std::min<int64_t>(TB, J0 + W - j0);
      gemm_strided<T>(m, nb, nb, B_ONE, xv(j0), sx, MatView<T>{inv + (j0 / TB) * TB2, TB}, sinv, B_ZERO, xv(j0), sx,
                      P, GEMM_TILE_64);
      const int64_t rc = j0 - J0;
      if (rc > 0)
        gemm_strided<T>(m, rc, nb, B_MONE, xv(j0), sx, lu(j0, J0), slu, B_ONE, xv(J0), sx, P, big_tile(m, rc, tile));
    }
    if (J0 > 0)
      gemm_strided<T>(m, J0, W, B_MONE, xv(J0), sx, lu(J0, 0), slu, B_ONE, xv(0), sx, P, big_tile(m, J0, tile));
  }
}

// X <- A^{-1} X for P problems (X n x q): L y = b ascending, U x = y descending.
template <class T>
static void bleft(int64_t n, int64_t q, int64_t P, cplx<T>* X, int64_t ldx, int64_t sx, cplx<T>* LU, int64_t slu,
                  cplx<T>* inv, int64_t sinv) {
  const int64_t uoff = inv_blocks(n), nbk = cdiv_up(n, TB);
  for (int64_t jb = 0; jb < nbk; ++jb) {
    const int64_t j0 = jb * TB;
    const int nb = (int)std::min<int64_t>(TB, n - j0);
    MatView<T> Xj{X + j0, ldx};
    gemm_strided<T>(nb, q, nb, B_ONE, MatView<T>{inv + jb * TB2, TB}, sinv, Xj, sx, B_ZERO, Xj, sx, P, GEMM_TILE_64);
    const int64_t r = n - j0 - nb;
    if (r > 0)
      gemm_strided<T>(r, q, nb, B_MONE, MatView<T>{LU + (j0 + nb) + j0 * n, n}, slu, Xj, sx, B_ONE,
                      MatView<T>{X + j0 + nb, ldx}, sx, P, GEMM_TILE_64);
  }
  for (int64_t jb = nbk - 1; jb >= 0; --jb) {
    const int64_t j0 = jb * TB;
    const int nb = (int)std::min<int64_t>(TB, n - j0);
    MatView<T> Xj{X + j0, ldx};
    gemm_strided<T>(nb, q, nb, B_ONE, MatView<T>{inv + uoff + jb * TB2, TB}, sinv, Xj, sx, B_ZERO, Xj, sx, P,
                    GEMM_TILE_64);
    if (j0 > 0)
      gemm_strided<T>(j0, q, nb, B_MONE, MatView<T>{LU + j0 * n, n}, slu, Xj, sx, B_ONE, MatView<T>{X, ldx}, sx, P,
                      GEMM_TILE_64);
  }
}

// ------------------------------------------------------------ the batch
// Tuning variants: panel width of the LU and right solves, and the tile of
// their large products.
struct BVariant {
  int nb, tile;
};
constexpr int BNVAR = 6;
constexpr BVariant BVARIANTS[BNVAR] = {{32, GEMM_TILE_64},      {128, GEMM_TILE_64},       {128, GEMM_TILE_128x64},
                                       {128, GEMM_TILE_128x96_3M}, {256, GEMM_TILE_128x64}, {256, GEMM_TILE_128x96_3M}};

struct BemRetBatch {
  int64_t n, cap, p, npinf, ndir;
  int solver;
  int variant = -1;   // BVARIANTS index; -1: tune at the first run
  int sweeps;         // refinement sweeps
  double tol = 1e-12;
  bool use_graph = true;
  double var_ms[BNVAR] = {0, 0, 0, 0, 0, 0};
  // device state
  void* mats = nullptr;                  // Gs | Ss | Dm | Dim (chain precision)
  void* invs = nullptr;                  // inv | inv2
  double2 *Gd = nullptr, *Hd = nullptr;  // the exact FP64 operator
  double2 *bv = nullptr, *xv = nullptr, *rv = nullptr, *GX = nullptr, *HX = nullptr;
  void *bT = nullptr, *xT = nullptr, *tmpT = nullptr;
  int* status = nullptr;  // 4 cap zero-pivot words
  double* relres = nullptr;
  BGeo* geos = nullptr;
  BPrm* prm = nullptr;
  double* dirs = nullptr;  // 3 x ndir
  double* polv = nullptr;  // 3 x p (pol), then 3 x p (dir)
  double2 *ef = nullptr, *hf = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_var = -2;
  size_t graph_nodes = 0;

  BemRetBatch(int64_t n_, int64_t cap_, int solver_, int64_t p_, int64_t npinf_)
      : n(n_), cap(cap_), p(p_), npinf(npinf_), ndir(p_ + npinf_), solver(solver_),
        sweeps(solver_ == MNP_SOLVER_FP64 ? 1 : 3) {
    const int64_t nn = n * n, es = solver == MNP_SOLVER_FP64 ? 16 : 8;
    mats = dev_alloc<char>((int64_t)6 * cap * nn * es);
    invs = dev_alloc<char>((int64_t)4 * cap * 2 * inv_blocks(n) * es);
    Gd = dev_alloc<double2>(2 * cap * nn);
    Hd = dev_alloc<double2>(2 * cap * nn);
    GX = dev_alloc<double2>(2 * cap * 4 * p * n);
    HX = dev_alloc<double2>(2 * cap * 4 * p * n);
    rv = dev_alloc<double2>(cap * 8 * p * n);
    if (solver != MNP_SOLVER_FP64) {
      bT = dev_alloc<float2>(cap * 8 * p * n);
      xT = dev_alloc<float2>(cap * 8 * p * n);
      tmpT = dev_alloc<float2>(cap * 18 * p * n);
    } else {
      tmpT = dev_alloc<double2>(cap * 18 * p * n);
    }
    bv = dev_alloc<double2>(cap * 8 * p * n);
    xv = dev_alloc<double2>(cap * 8 * p * n);
    status = dev_alloc<int>(4 * cap);
    relres = dev_alloc<double>(cap);
    geos = dev_alloc<BGeo>(cap);
    prm = dev_alloc<BPrm>(cap);
    dirs = dev_alloc<double>(3 * ndir);
    polv = dev_alloc<double>(6 * p);
    ef = dev_alloc<double2>(cap * ndir * 3 * p);
    hf = dev_alloc<double2>(cap * ndir * 3 * p);
  }
  ~BemRetBatch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    void* ps[] = {mats, invs, Gd, Hd, bv, xv, rv, GX, HX, bT, xT, tmpT, status, relres, geos, prm, dirs, polv, ef, hf};
    for (void* q : ps)
      if (q) (void)hipFree(q);
  }
  bool variant_ok(int v) const {
    return v >= 0 && v < BNVAR && !(solver == MNP_SOLVER_FP64 && BVARIANTS[v].tile == GEMM_TILE_128x96_3M);
  }

  // The chain in precision T for all members (initmat.m:5-72): factors in
  // Gs (G1, G2), Ss (Sigma1; Sigma over Sigma2), Dm (Sigma1 - Sigma2), Dim
  // (Deltai), their block inverses in inv, inv2; zero pivots in status.
  template <class T>
  void chain(const BVariant& v) {
    const int64_t nn = n * n, uoff = inv_blocks(n), B = cap;
    cplx<T>* Gs = (cplx<T>*)mats;
    cplx<T>* Ss = Gs + 2 * B * nn;
    cplx<T>* Dm = Ss + 2 * B * nn;
    cplx<T>* Dim = Dm + B * nn;
    cplx<T>* inv = (cplx<T>*)invs;
    cplx<T>* inv2 = inv + 2 * B * 2 * uoff;
    int* zp = status;
    blu<T>(n, 2 * B, Gs, nn, inv, 2 * uoff, zp, 1, v.tile, v.nb);            // LU(G1), LU(G2)
    bright<T>(n, n, 2 * B, Ss, n, nn, Gs, nn, inv, 2 * uoff, v.tile, v.nb);  // Sigma_s = H_s G_s^{-1}
    hipLaunchKernelGGL(k_bsub<T>, bgrid(n, n, B), dim3(256), 0, stream(), n, Ss, Dm);
    blu<T>(n, B, Dm, nn, inv2, 4 * uoff, zp + 2 * B, 2, v.tile, v.nb);       // LU(Sigma1 - Sigma2)
    hipLaunchKernelGGL(k_bident<T>, bgrid(n, n, B), dim3(256), 0, stream(), n, Dim);
    bright<T>(n, n, B, Dim, n, nn, Dm, nn, inv2, 4 * uoff, v.tile, v.nb);    // Deltai
    hipLaunchKernelGGL(k_bsigma<T>, bgrid(n, n, B), dim3(256), 0, stream(), n, Ss, Dim, geos, prm);
    blu<T>(n, B, Ss + nn, 2 * nn, inv2 + 2 * uoff, 4 * uoff, zp + 2 * B + 1, 2, v.tile, v.nb);  // LU(Sigma)
    MNP_HIP(hipGetLastError());
  }
  // mldivide (mldivide.m:11-49) in precision T: x = M^{-1} b for all members
  // (b, x: cap x n x 8p contiguous).
  template <class T>
  void apply(const cplx<T>* b, cplx<T>* x) {
    const int64_t nn = n * n, uoff = inv_blocks(n), B = cap, pn = p * n;
    cplx<T>* Gs = (cplx<T>*)mats;
    cplx<T>* Ss = Gs + 2 * B * nn;
    cplx<T>* Dim = Ss + 3 * B * nn;
    cplx<T>* inv = (cplx<T>*)invs;
    cplx<T>* inv2 = inv + 2 * B * 2 * uoff;
    cplx<T>* tmp = (cplx<T>*)tmpT;
    const int64_t st = 18 * pn, sv = 8 * pn;
    cplx<T>* bb = const_cast<cplx<T>*>(b);
    // Y = Sigma1 [a | phi]
    gemm_strided<T>(n, 3 * p, n, B_ONE, MatView<T>{Ss, n}, 2 * nn, MatView<T>{bb + pn, n}, sv, B_ZERO,
                    MatView<T>{tmp, n}, st, B, GEMM_TILE_64);
    gemm_strided<T>(n, p, n, B_ONE, MatView<T>{Ss, n}, 2 * nn, MatView<T>{bb, n}, sv, B_ZERO,
                    MatView<T>{tmp + 3 * pn, n}, st, B, GEMM_TILE_64);
    const dim3 gp = bgrid(n, p, B);
    hipLaunchKernelGGL(k_bmld1<T>, gp, dim3(256), 0, stream(), n, p, b, tmp, tmp, geos, prm);
    // t = Deltai alpha
    gemm_strided<T>(n, 3 * p, n, B_ONE, MatView<T>{Dim, n}, nn, MatView<T>{tmp + 4 * pn, n}, st, B_ZERO,
                    MatView<T>{tmp + 8 * pn, n}, st, B, GEMM_TILE_64);
    hipLaunchKernelGGL(k_bmld2<T>, gp, dim3(256), 0, stream(), n, p, tmp, geos, prm);
    bleft<T>(n, p, B, tmp + 11 * pn, n, st, Ss + nn, 2 * nn, inv2 + 2 * uoff, 4 * uoff);  // sig2 = Sigma^{-1} rhs
    hipLaunchKernelGGL(k_bmld3<T>, gp, dim3(256), 0, stream(), n, p, tmp, geos, prm);
    gemm_strided<T>(n, 3 * p, n, B_ONE, MatView<T>{Dim, n}, nn, MatView<T>{tmp + 12 * pn, n}, st, B_ZERO,
                    MatView<T>{tmp + 15 * pn, n}, st, B, GEMM_TILE_64);  // h2 = Deltai u
    hipLaunchKernelGGL(k_bmld4<T>, bgrid(n, 4 * p, B), dim3(256), 0, stream(), n, p, b, tmp, x);
    MNP_HIP(hipGetLastError());
    bleft<T>(n, 4 * p, 2 * B, x, n, 4 * pn, Gs, nn, inv, 2 * uoff);  // G1^{-1} [...], G2^{-1} [...]
  }

  // The whole pipeline for all members, on stream(): assembly (chain matrices
  // and the FP64 operator), the chain, x = M^{-1} b, `sweeps` refinement
  // sweeps x += M^{-1} (b - A x) with the FP64 operator (afun.m), the final
  // relative residual per member, far fields.
  void record(const BVariant& v) {
    const int64_t nn = n * n, B = cap, len = B * 8 * p * n;
    hipLaunchKernelGGL(k_bclear, dim3((unsigned)cdiv_up(4 * cap, 256)), dim3(256), 0, stream(), 4 * cap, status);
    const dim3 ga((unsigned)cdiv_up(n, 256), (unsigned)n, (unsigned)(2 * B));
    const dim3 gr(512u, 1u, (unsigned)(2 * B));
    const bool fp64 = solver == MNP_SOLVER_FP64;
    if (fp64) {
      double2* Gs = (double2*)mats;
      hipLaunchKernelGGL(k_bassemble<double>, ga, dim3(256), 0, stream(), n, geos, prm, Gs, Gs + 2 * B * nn, Gd, Hd);
      hipLaunchKernelGGL(k_bassemble_refined<double>, gr, dim3(256), 0, stream(), n, geos, prm, Gs, Gs + 2 * B * nn,
                         Gd, Hd);
    } else {
      float2* Gs = (float2*)mats;
      hipLaunchKernelGGL(k_bassemble<float>, ga, dim3(256), 0, stream(), n, geos, prm, Gs, Gs + 2 * B * nn, Gd, Hd);
      hipLaunchKernelGGL(k_bassemble_refined<float>, gr, dim3(256), 0, stream(), n, geos, prm, Gs, Gs + 2 * B * nn,
                         Gd, Hd);
    }
    MNP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_bexcite, bgrid(n, p, B), dim3(256), 0, stream(), n, p, polv, polv + 3 * p, geos, prm, bv);
    MNP_HIP(hipGetLastError());
    const unsigned gl = (unsigned)cdiv_up(len, 256);
    if (fp64)
      chain<double>(v);
    else
      chain<float>(v);
    auto precond = [&](const double2* src, double2* dst) {
      if (fp64) {
        apply<double>(src, dst);
      } else {
        float2 *b32 = (float2*)bT, *x32 = (float2*)xT;
        hipLaunchKernelGGL((k_bconvert<double, float>), dim3(gl), dim3(256), 0, stream(), len, src, b32);
        apply<float>(b32, x32);
        hipLaunchKernelGGL((k_bconvert<float, double>), dim3(gl), dim3(256), 0, stream(), len, x32, dst);
      }
    };
    auto residual = [&]() {
      const int64_t np = 4 * p * n;
      gemm_strided<double>(n, 4 * p, n, B_ONE, MatView<double>{Gd, n}, nn, MatView<double>{xv, n}, np, B_ZERO,
                           MatView<double>{GX, n}, np, 2 * B, GEMM_TILE_64);
      gemm_strided<double>(n, 4 * p, n, B_ONE, MatView<double>{Hd, n}, nn, MatView<double>{xv, n}, np, B_ZERO,
                           MatView<double>{HX, n}, np, 2 * B, GEMM_TILE_64);
      hipLaunchKernelGGL(k_bresidual, bgrid(n, p, B), dim3(256), 0, stream(), n, p, bv, GX, HX, rv, geos, prm);
    };
    precond(bv, xv);
    for (int s = 0; s < sweeps; ++s) {
      residual();
      precond(rv, GX);  // correction (cap x n x 8p) into GX
      hipLaunchKernelGGL(k_baxpy, dim3(gl), dim3(256), 0, stream(), len, xv, GX);
    }
    residual();
    hipLaunchKernelGGL(k_brelres, dim3((unsigned)B), dim3(256), 0, stream(), 8 * p * n, rv, bv, relres);
    MNP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_bfarfield, dim3((unsigned)ndir, 1u, (unsigned)B), dim3(256), 0, stream(), n, p, ndir, dirs,
                       geos, prm, xv, ef, hf);
    MNP_HIP(hipGetLastError());
  }

  void capture(int var) {
    if (exec && graph_var == var) return;
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    graph_var = -2;
    hipStream_t cs;
    MNP_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    const hipStream_t prev = thread_stream();
    set_thread_stream(cs);
    pause_timers(true);
    hipError_t err = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (err == hipSuccess) {
      try {
        record(BVARIANTS[var]);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(cs, &g);
        if (g) (void)hipGraphDestroy(g);
        set_thread_stream(prev);
        pause_timers(false);
        (void)hipStreamDestroy(cs);
        throw;
      }
      err = hipStreamEndCapture(cs, &graph);
    }
    set_thread_stream(prev);
    pause_timers(false);
    (void)hipStreamDestroy(cs);
    MNP_HIP(err);
    MNP_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    MNP_HIP(hipGraphGetNodes(graph, nullptr, &graph_nodes));
    graph_var = var;
  }

  // Runs the pipeline (graph replay or direct launches) on stream().
  void launch(int var) {
    if (use_graph) {
      capture(var);
      MNP_HIP(hipGraphLaunch(exec, stream()));
    } else {
      record(BVARIANTS[var]);
    }
  }

  // Variant: each candidate timed on one run of the pipeline (after one
  // untimed run that captures it), the fastest kept for (n, cap, solver).
  int tuned_variant() {
    if (variant >= 0) return variant;
    static std::mutex mu;
    static std::map<std::tuple<int64_t, int64_t, int>, int> table;
    {
      std::lock_guard<std::mutex> lk(mu);
      auto it = table.find({n, cap, solver});
      if (it != table.end()) return it->second;
    }
    hipEvent_t e0, e1;
    MNP_HIP(hipEventCreate(&e0));
    MNP_HIP(hipEventCreate(&e1));
    int best = 0;
    float best_ms = 1e30f;
    for (int c = 0; c < BNVAR; ++c) {
      if (!variant_ok(c) || (c > 0 && BVARIANTS[c].nb >= 2 * n)) continue;  // panels wider than the matrix repeat another variant
      launch(c);
      MNP_HIP(hipEventRecord(e0, stream()));
      launch(c);
      MNP_HIP(hipEventRecord(e1, stream()));
      MNP_HIP(hipEventSynchronize(e1));
      float ms = 0;
      MNP_HIP(hipEventElapsedTime(&ms, e0, e1));
      var_ms[c] = ms;
      if (ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::lock_guard<std::mutex> lk(mu);
    table[{n, cap, solver}] = best;
    return best;
  }
};

}  // namespace mnp

// ============================================================== C ABI
using namespace mnp;

struct mnp_bemret_batch_s {
  mnp::BemRetBatch* b;
};

extern "C" {

int mnp_bemret_batch_create(int64_t nfaces, int64_t capacity, int solver, int64_t npol, int64_t npinf,
                            mnp_bemret_batch_t* out) {
  return guarded([&] {
    MNP_CHECK(out, MNP_ERR_ARG, "bemret_batch_create: NULL argument");
    *out = nullptr;
    MNP_CHECK(nfaces >= 1 && nfaces <= 65535, MNP_ERR_ARG, "bemret_batch_create: faces must be in 1..65535");
    MNP_CHECK(capacity >= 1 && 2 * capacity < 65536, MNP_ERR_ARG, "bemret_batch_create: bad capacity");
    MNP_CHECK(npol >= 1 && npol <= BFF_MAXP && npinf >= 0 && npol + npinf < 65536, MNP_ERR_ARG,
              "bemret_batch_create: npol must be in 1..8");
    MNP_CHECK(solver == MNP_SOLVER_FP64 || solver == MNP_SOLVER_MIXED, MNP_ERR_ARG,
              "bemret_batch_create: solver must be MNP_SOLVER_FP64 or MNP_SOLVER_MIXED");
    *out = new mnp_bemret_batch_s{new BemRetBatch(nfaces, capacity, solver, npol, npinf)};
  });
}

int mnp_bemret_batch_destroy(mnp_bemret_batch_t b) {
  return guarded([&] {
    if (!b) return;
    delete b->b;
    delete b;
  });
}

int mnp_bemret_batch_set_options(mnp_bemret_batch_t b, int use_graph, int variant, int sweeps) {
  return guarded([&] {
    MNP_CHECK(b, MNP_ERR_ARG, "bemret_batch_set_options: NULL handle");
    BemRetBatch& B = *b->b;
    MNP_CHECK((variant == -1 || B.variant_ok(variant)) && sweeps >= -1 && sweeps <= 20, MNP_ERR_ARG,
              "bemret_batch_set_options: bad variant or sweeps");
    B.use_graph = use_graph != 0;
    if (sweeps < 0) sweeps = B.solver == MNP_SOLVER_FP64 ? 1 : 3;
    if (variant != B.variant || sweeps != B.sweeps) B.graph_var = -2;  // capture again
    B.variant = variant;
    B.sweeps = sweeps;
  });
}

int mnp_bemret_batch_cross_sections(mnp_bemret_batch_t bh, int64_t count, const mnp_greenret_t* greens,
                                    const double* enei, const mnp_complex* eps_in, const mnp_complex* eps_out,
                                    const double* pol, const double* dir, const void* pinfty_nvec,
                                    const void* pinfty_area, double* ext, double* sca, double* abs_, int* status,
                                    double* relres) {
  return guarded([&] {
    MNP_CHECK(bh && greens && enei && eps_in && eps_out && pol && dir && ext && sca && abs_ && status, MNP_ERR_ARG,
              "bemret_batch_cross_sections: NULL argument");
    BemRetBatch& B = *bh->b;
    MNP_CHECK(count >= 1 && count <= B.cap, MNP_ERR_ARG, "bemret_batch_cross_sections: count must be in 1..capacity");
    MNP_CHECK(B.npinf == 0 || (pinfty_nvec && pinfty_area), MNP_ERR_ARG,
              "bemret_batch_cross_sections: pinfty surface missing");
    const int64_t n = B.n, p = B.p, cap = B.cap, ndir = B.ndir;
    std::vector<BGeo> hg(cap);
    std::vector<BPrm> hp(cap);
    for (int64_t i = 0; i < cap; ++i) {
      const int64_t m = std::min(i, count - 1);  // unused slots repeat the last member
      MNP_CHECK(greens[m] && greens[m]->g->n == n, MNP_ERR_ARG,
                "bemret_batch_cross_sections: every member needs a mesh of the batch's face count");
      MNP_CHECK(enei[m] > 0, MNP_ERR_ARG, "bemret_batch_cross_sections: wavelengths must be positive");
      const GreenRet& G = *greens[m]->g;
      MNP_CHECK(!G.compound(), MNP_ERR_ARG, "bemret_batch_cross_sections: compound particles take the dense solver");
      hg[i] = BGeo{G.geo.pos, G.geo.nvec, G.geo.area, G.pairs.row, G.pairs.col, G.pairs.kind, G.gtab, G.ftab,
                   G.nnz, G.order};
      const double k0 = 2 * M_PI / enei[m];
      const double2 e1 = to_d2(eps_in[m]), e2 = to_d2(eps_out[m]);
      const double2 L = csub<double>(e1, e2);
      hp[i] = BPrm{k0, wavenumber(k0, e1), wavenumber(k0, e2), e1, e2, cscale<double>(cmul<double>(L, L), k0 * k0)};
    }
    std::vector<double> hd(3 * ndir), hpol(6 * p);
    for (int c = 0; c < 3; ++c)
      for (int64_t q = 0; q < p; ++q) {
        hd[c * ndir + q] = dir[c * p + q];
        hpol[c * p + q] = pol[c * p + q];
        hpol[3 * p + c * p + q] = dir[c * p + q];
      }
    std::vector<double> pn(3 * B.npinf), pa(B.npinf);
    if (B.npinf) {
      MNP_HIP(hipMemcpyAsync(pn.data(), pinfty_nvec, sizeof(double) * 3 * B.npinf, hipMemcpyDeviceToHost, stream()));
      MNP_HIP(hipMemcpyAsync(pa.data(), pinfty_area, sizeof(double) * B.npinf, hipMemcpyDeviceToHost, stream()));
      MNP_HIP(hipStreamSynchronize(stream()));
      for (int c = 0; c < 3; ++c)
        for (int64_t d = 0; d < B.npinf; ++d) hd[c * ndir + p + d] = pn[c * B.npinf + d];
    }
    MNP_HIP(hipMemcpyAsync(B.geos, hg.data(), sizeof(BGeo) * cap, hipMemcpyHostToDevice, stream()));
    MNP_HIP(hipMemcpyAsync(B.prm, hp.data(), sizeof(BPrm) * cap, hipMemcpyHostToDevice, stream()));
    MNP_HIP(hipMemcpyAsync(B.dirs, hd.data(), sizeof(double) * 3 * ndir, hipMemcpyHostToDevice, stream()));
    MNP_HIP(hipMemcpyAsync(B.polv, hpol.data(), sizeof(double) * 6 * p, hipMemcpyHostToDevice, stream()));
    B.launch(B.tuned_variant());
    std::vector<double2> he(cap * ndir * 3 * p), hh(cap * ndir * 3 * p);
    std::vector<int> st(4 * cap);
    std::vector<double> rr(cap, 0.0);
    MNP_HIP(hipMemcpyAsync(he.data(), B.ef, sizeof(double2) * he.size(), hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipMemcpyAsync(hh.data(), B.hf, sizeof(double2) * hh.size(), hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipMemcpyAsync(st.data(), B.status, sizeof(int) * 4 * cap, hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipMemcpyAsync(rr.data(), B.relres, sizeof(double) * cap, hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipStreamSynchronize(stream()));
    for (int64_t m = 0; m < count; ++m) {
      // zero pivots: G1, G2 at 2m, 2m + 1; Sigma1 - Sigma2, Sigma at 2 cap + 2m (+1)
      const bool zp = st[2 * m] || st[2 * m + 1] || st[2 * cap + 2 * m] || st[2 * cap + 2 * m + 1];
      status[m] = zp ? 1 : (!(rr[m] <= B.tol) ? 3 : 0);
      if (relres) relres[m] = rr[m];
      const double2 k = hp[m].k2;
      const double nb = std::sqrt(std::hypot(eps_out[m].re, eps_out[m].im));
      farfield_host_sums(p, pol, ndir, B.npinf, he.data() + m * ndir * 3 * p, hh.data() + m * ndir * 3 * p, pn.data(),
                         pa.data(), k, nb, ext + m * p, sca + m * p, abs_ + m * p);
    }
  });
}

int mnp_bemret_batch_info(mnp_bemret_batch_t b, int* variant, double* variant_ms, int64_t* graph_nodes) {
  return guarded([&] {
    MNP_CHECK(b, MNP_ERR_ARG, "bemret_batch_info: NULL handle");
    BemRetBatch& B = *b->b;
    if (variant) *variant = B.graph_var >= 0 ? B.graph_var : B.variant;
    if (variant_ms)
      for (int c = 0; c < BNVAR; ++c) variant_ms[c] = B.var_ms[c];
    if (graph_nodes) *graph_nodes = (int64_t)B.graph_nodes;
  });
}

}  // extern "C"
