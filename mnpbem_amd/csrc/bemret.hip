// Dense retarded BEM solver (MNPBEM BEM/@bemret) on gfx950.
//
// initmat (BEM/@bemret/private/initmat.m:5-72) builds, per wavelength,
//   G1 = G(k1), G2 = G(k2), H1 = F(k1) + 2 pi, H2 = F(k2) - 2 pi,
//   Sigma1 = H1 G1^{-1}, Sigma2 = H2 G2^{-1}, Deltai = (Sigma1 - Sigma2)^{-1},
//   Sigma = eps1 Sigma1 - eps2 Sigma2 + k^2 (eps1 - eps2)^2 (Deltai o n n^T).
// Instead of the reference's explicit inverses of G1, G2 and Sigma, the
// chain keeps their LU factors (G1^{-1} and G2^{-1} are only ever applied)
// and forms Sigma1, Sigma2 by right triangular solves; only Deltai, whose
// entries enter Sigma, is inverted explicitly.
//
// MNP_SOLVER_FP64 runs the chain and mldivide (mldivide.m:11-49) in complex
// FP64.  MNP_SOLVER_MIXED runs the chain in complex FP32 and uses it as the
// preconditioner mfun of the exact operator afun (BEM/@bemretiter/private/
// afun.m:6-35, mfun.m:10-52) applied with the FP64 matrices:
//   x <- x + M^{-1} (b - A x)  until |b - A x| / |b| < tol.
// MNP_SOLVER_MIXED_BF16 is MNP_SOLVER_MIXED with the chain's large complex64
// products on the split-bf16 kernel (a less exact preconditioner; the
// refinement against the FP64 operator sets the final accuracy).
#include <algorithm>
#include <cmath>
#include <functional>
#include <memory>
#include <mutex>
#include <vector>

#include "mnp_bemret.hpp"
#include "mnp_chain.hpp"
#include "mnp_linalg.hpp"

namespace mnp {

// b = [phi | a | De | alpha] from the excitation potentials (excitation.m:21-29).
__global__ void k_excitation(int64_t n, int64_t p, const double2* phi1, const double2* phi1p,
                             const double2* a1, const double2* a1p, const double2* phi2,
                             const double2* phi2p, const double2* a2, const double2* a2p,
                             double2* b, const double* nvec, double k, double2 e1_, double2 e2_,
                             const double2* e1v, const double2* e2v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const double2 e1 = e1v ? e1v[i] : e1_, e2 = e2v ? e2v[i] : e2_;  // per face for a compound
  const double2 z = mk<double>(0, 0);
  auto ld = [&](const double2* v, int64_t col) { return v ? v[col * n + i] : z; };
  const double2 p1 = ld(phi1, q), p2 = ld(phi2, q);
  const double2 ik = mk<double>(0, k);
  b[q * n + i] = csub<double>(p2, p1);
  double2 na1 = z, na2 = z;
  const double2 e2p2 = cmul<double>(e2, p2), e1p1 = cmul<double>(e1, p1);
  for (int c = 0; c < 3; ++c) {
    const int64_t col = c + 3 * q;
    const double nc = nvec[c * n + i];
    const double2 va1 = ld(a1, col), va2 = ld(a2, col);
    b[(p + col) * n + i] = csub<double>(va2, va1);
    const double2 w = csub<double>(cscale<double>(e2p2, nc), cscale<double>(e1p1, nc));
    b[(5 * p + col) * n + i] = csub<double>(csub<double>(ld(a2p, col), ld(a1p, col)), cmul<double>(ik, w));
    na1 = cadd<double>(na1, cscale<double>(va1, nc));
    na2 = cadd<double>(na2, cscale<double>(va2, nc));
  }
  const double2 t = csub<double>(cmul<double>(e2, na2), cmul<double>(e1, na1));
  b[(4 * p + q) * n + i] = csub<double>(
      csub<double>(cmul<double>(e2, ld(phi2p, q)), cmul<double>(e1, ld(phi1p, q))), cmul<double>(ik, t));
}

// r = b - A x with the exact operator (afun.m): GX1 = G1 [sig1|h1],
// GX2 = G2 [sig2|h2], HX1 = H1 [sig1|h1], HX2 = H2 [sig2|h2]  (each n x 4p).
// With b == NULL the kernel writes -A x instead.
__global__ void k_residual(int64_t n, int64_t p, const double2* b_, const double2* GX1,
                           const double2* GX2, const double2* HX1, const double2* HX2, double2* r,
                           const double* nvec, double k, double2 e1_, double2 e2_, const double2* e1v,
                           const double2* e2v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const double2 e1 = e1v ? e1v[i] : e1_, e2 = e2v ? e2v[i] : e2_;  // per face for a compound
  const double2 zero = mk<double>(0, 0);
  auto b = [&](int64_t idx) { return b_ ? b_[idx] : zero; };
  const double2 ik = mk<double>(0, k);
  const double2 Gs1 = GX1[q * n + i], Gs2 = GX2[q * n + i];
  const double2 Hs1 = HX1[q * n + i], Hs2 = HX2[q * n + i];
  // phi
  r[q * n + i] = csub<double>(b(q * n + i), csub<double>(Gs1, Gs2));
  const double2 w = csub<double>(cmul<double>(e1, Gs1), cmul<double>(e2, Gs2));
  double2 nGh = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) {
    const int64_t col = c + 3 * q;
    const double nc = nvec[c * n + i];
    const double2 Gh1 = GX1[(p + col) * n + i], Gh2 = GX2[(p + col) * n + i];
    const double2 Hh1 = HX1[(p + col) * n + i], Hh2 = HX2[(p + col) * n + i];
    r[(p + col) * n + i] = csub<double>(b((p + col) * n + i), csub<double>(Gh1, Gh2));
    const double2 al = csub<double>(csub<double>(Hh1, Hh2), cmul<double>(ik, cscale<double>(w, nc)));
    r[(5 * p + col) * n + i] = csub<double>(b((5 * p + col) * n + i), al);
    nGh = cadd<double>(nGh, cscale<double>(csub<double>(cmul<double>(e1, Gh1), cmul<double>(e2, Gh2)), nc));
  }
  const double2 De = csub<double>(csub<double>(cmul<double>(e1, Hs1), cmul<double>(e2, Hs2)),
                                  cmul<double>(ik, nGh));
  r[(4 * p + q) * n + i] = csub<double>(b((4 * p + q) * n + i), De);
}

// Nested media (full L1, L2): the operator the chain inverts puts eps on the
// charges, r = b - A x with
//   phi = G1 s1 - G2 s2,  a = G1 h1 - G2 h2,
//   alpha = H1 h1 - H2 h2 - i k n (G1 (e1 s1) - G2 (e2 s2)),
//   De = H1 (e1 s1) - H2 (e2 s2) - i k n . (G1 (e1 h1) - G2 (e2 h2))
// from GX = G [s | h], GE = G [e s | e h], HM = H [e s | h] (each n x 4p).
__global__ void k_residual_nested(int64_t n, int64_t p, const double2* b, const double2* GX1, const double2* GX2,
                                  const double2* GE1, const double2* GE2, const double2* HM1, const double2* HM2,
                                  double2* r, const double* nvec, double k) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const double2 ik = mk<double>(0, k);
  r[q * n + i] = csub<double>(b[q * n + i], csub<double>(GX1[q * n + i], GX2[q * n + i]));
  const double2 ges = csub<double>(GE1[q * n + i], GE2[q * n + i]);
  double2 nge = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) {
    const int64_t col = c + 3 * q;
    const double nc = nvec[c * n + i];
    r[(p + col) * n + i] =
        csub<double>(b[(p + col) * n + i], csub<double>(GX1[(p + col) * n + i], GX2[(p + col) * n + i]));
    const double2 al = csub<double>(csub<double>(HM1[(p + col) * n + i], HM2[(p + col) * n + i]),
                                    cmul<double>(ik, cscale<double>(ges, nc)));
    r[(5 * p + col) * n + i] = csub<double>(b[(5 * p + col) * n + i], al);
    nge = cadd<double>(nge, cscale<double>(csub<double>(GE1[(p + col) * n + i], GE2[(p + col) * n + i]), nc));
  }
  const double2 De = csub<double>(csub<double>(HM1[q * n + i], HM2[q * n + i]), cmul<double>(ik, nge));
  r[(4 * p + q) * n + i] = csub<double>(b[(4 * p + q) * n + i], De);
}
// xe = [e1 s1 | e1 h1 | e2 s2 | e2 h2], xm = [e1 s1 | h1 | e2 s2 | h2] from x (n x 8p)
__global__ void k_scale_x(int64_t n, int64_t p, const double2* x, const double2* e1, const double2* e2, double2* xe,
                          double2* xm) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t col = blockIdx.y;  // 0 .. 8p
  if (i >= n) return;
  const int64_t side = col / (4 * p), within = col % (4 * p);
  const double2 e = side ? e2[i] : e1[i];
  const double2 v = x[col * n + i], ev = cmul<double>(e, v);
  xe[col * n + i] = ev;
  xm[col * n + i] = within < p ? ev : v;
}

__global__ void k_norm2(int64_t len, const double2* v, double* out) {
  __shared__ double red[256];
  double s = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * blockDim.x)
    s += v[i].x * v[i].x + v[i].y * v[i].y;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}
__global__ void k_axpy(int64_t len, double2* x, const double2* d) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) x[i] = cadd<double>(x[i], d[i]);
}
__global__ void k_unpack(int64_t n, int64_t p, const double2* x, double2* sig1, double2* h1,
                         double2* sig2, double2* h2) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t col = blockIdx.y;  // 0 .. 8p
  if (i >= n) return;
  const double2 v = x[col * n + i];
  if (col < p) {
    if (sig1) sig1[col * n + i] = v;
  } else if (col < 4 * p) {
    if (h1) h1[(col - p) * n + i] = v;
  } else if (col < 5 * p) {
    if (sig2) sig2[(col - 4 * p) * n + i] = v;
  } else {
    if (h2) h2[(col - 5 * p) * n + i] = v;
  }
}

__global__ void k_pack(int64_t n, int64_t p, const double2* sig1, const double2* h1, const double2* sig2,
                       const double2* h2, double2* x) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t col = blockIdx.y;  // 0 .. 8p
  if (i >= n) return;
  double2 v;
  if (col < p)
    v = sig1[col * n + i];
  else if (col < 4 * p)
    v = h1[(col - p) * n + i];
  else if (col < 5 * p)
    v = sig2[(col - 4 * p) * n + i];
  else
    v = h2[(col - 5 * p) * n + i];
  x[col * n + i] = v;
}

// r = -(dA/deps1) x in the layout of b (see k_residual for A): with
// GX = G1 [sig1 | h1], HX = H1 [sig1 | h1] and dGX, dHX their derivatives
// along eps1 (through k1),
//   phi:   dGs1                a:  dGh1
//   alpha: dHh1 - i k0 n (Gs1 + eps1 dGs1)
//   De:    Hs1 + eps1 dHs1 - i k0 n . (Gh1 + eps1 dGh1)
__global__ void k_deps1(int64_t n, int64_t p, const double2* GX, const double2* HX, const double2* dGX,
                        const double2* dHX, double2* r, const double* nvec, double k0, double2 e1) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const double2 ik = mk<double>(0, k0);
  const double2 Gs = GX[q * n + i], dGs = dGX[q * n + i];
  const double2 Hs = HX[q * n + i], dHs = dHX[q * n + i];
  r[q * n + i] = mk<double>(-dGs.x, -dGs.y);
  const double2 ts = cadd<double>(Gs, cmul<double>(e1, dGs));  // Gs1 + eps1 dGs1
  double2 nGh = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) {
    const int64_t col = c + 3 * q;
    const double nc = nvec[c * n + i];
    const double2 Gh = GX[(p + col) * n + i], dGh = dGX[(p + col) * n + i], dHh = dHX[(p + col) * n + i];
    r[(p + col) * n + i] = mk<double>(-dGh.x, -dGh.y);
    const double2 al = csub<double>(dHh, cmul<double>(ik, cscale<double>(ts, nc)));
    r[(5 * p + col) * n + i] = mk<double>(-al.x, -al.y);
    nGh = cadd<double>(nGh, cscale<double>(cadd<double>(Gh, cmul<double>(e1, dGh)), nc));
  }
  const double2 De = csub<double>(cadd<double>(Hs, cmul<double>(e1, dHs)), cmul<double>(ik, nGh));
  r[(4 * p + q) * n + i] = mk<double>(-De.x, -De.y);
}

double dev_norm2(int64_t len, const double2* v, Workspace& ws, int slot) {
  const int nb = 512;
  double* part = (double*)ws.get(slot, sizeof(double) * nb);
  hipLaunchKernelGGL(k_norm2, dim3(nb), dim3(256), 0, stream(), len, v, part);
  MNP_HIP(hipGetLastError());
  std::vector<double> h(nb);
  MNP_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, stream()));
  MNP_HIP(hipStreamSynchronize(stream()));
  double s = 0;
  for (double v2 : h) s += v2;
  return std::sqrt(s);
}

void launch_excitation(int64_t n, int64_t p, const double2* const in[8], double2* b, const double* nvec,
                       double k, double2 e1, double2 e2) {
  hipLaunchKernelGGL(k_excitation, dim3((unsigned)cdiv_up(n, 256), (unsigned)p), dim3(256), 0, stream(), n, p,
                     in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], b, nvec, k, e1, e2,
                     (const double2*)nullptr, (const double2*)nullptr);
  MNP_HIP(hipGetLastError());
}
void launch_residual(int64_t n, int64_t p, const double2* b, const double2* GX1, const double2* GX2,
                     const double2* HX1, const double2* HX2, double2* r, const double* nvec, double k,
                     double2 e1, double2 e2) {
  hipLaunchKernelGGL(k_residual, dim3((unsigned)cdiv_up(n, 256), (unsigned)p), dim3(256), 0, stream(), n, p, b,
                     GX1, GX2, HX1, HX2, r, nvec, k, e1, e2, (const double2*)nullptr, (const double2*)nullptr);
  MNP_HIP(hipGetLastError());
}
void launch_unpack(int64_t n, int64_t p, const double2* x, double2* sig1, double2* h1, double2* sig2,
                   double2* h2) {
  hipLaunchKernelGGL(k_unpack, dim3((unsigned)cdiv_up(n, 256), (unsigned)(8 * p)), dim3(256), 0, stream(), n, p, x,
                     sig1, h1, sig2, h2);
  MNP_HIP(hipGetLastError());
}

void launch_pack(int64_t n, int64_t p, const double2* sig1, const double2* h1, const double2* sig2,
                 const double2* h2, double2* x) {
  hipLaunchKernelGGL(k_pack, dim3((unsigned)cdiv_up(n, 256), (unsigned)(8 * p)), dim3(256), 0, stream(), n, p,
                     sig1, h1, sig2, h2, x);
  MNP_HIP(hipGetLastError());
}
void launch_deps1(int64_t n, int64_t p, const double2* GX, const double2* HX, const double2* dGX,
                  const double2* dHX, double2* r, const double* nvec, double k0, double2 e1) {
  hipLaunchKernelGGL(k_deps1, dim3((unsigned)cdiv_up(n, 256), (unsigned)p), dim3(256), 0, stream(), n, p, GX, HX,
                     dGX, dHX, r, nvec, k0, e1);
  MNP_HIP(hipGetLastError());
}

BemRet::BemRet(const GreenRet* g_, int solver_) : g(g_), solver(solver_), n(g_->n) {
  if (solver == MNP_SOLVER_FP64) {
    c64.reset(new Chain<double>(n, true));
  } else {
    c32.reset(new Chain<float>(n, false));
    G1 = dev_alloc<double2>(n * n);
    G2 = dev_alloc<double2>(n * n);
    H1 = dev_alloc<double2>(n * n);
    H2 = dev_alloc<double2>(n * n);
  }
}
BemRet::~BemRet() {
  adjoint_free(adj);
  void* ps[] = {G1, G2, H1, H2, e1v, e2v};
  for (void* p : ps)
    if (p) (void)hipFree(p);
}

void BemRet::initmat(double enei_, double2 e1, double2 e2) {
  if (!initmat_try(enei_, e1, e2)) initmat_finish();
}

// Compound particle (initmat.m:5-72 with eps per face and the connectivity of
// compgreenret/private/eval1.m): epsm[m - 1] is medium m's dielectric function.
void BemRet::initmat_media(double enei_, const std::vector<double2>& epsm) {
  MNP_CHECK(g->compound(), MNP_ERR_STATE, "bemret_initmat_media: the Green function is not a compound");
  MNP_CHECK((int)epsm.size() >= g->nmed, MNP_ERR_ARG, "bemret_initmat_media: a dielectric function per medium");
  ready = false;
  ++epoch;
  enei = enei_;
  eps1 = eps2 = mk<double>(0, 0);
  const double k0 = 2 * M_PI / enei;
  KTab kt{};
  for (int m = 0; m < g->nmed; ++m) kt.k[m] = wavenumber(k0, epsm[m]);
  std::vector<double2> h1(n), h2(n);
  for (int64_t i = 0; i < n; ++i) {
    h1[i] = epsm[g->fin_h[i] - 1];
    h2[i] = epsm[g->fout_h[i] - 1];
  }
  if (!e1v) {
    e1v = dev_alloc<double2>(n);
    e2v = dev_alloc<double2>(n);
  }
  MNP_HIP(hipMemcpyAsync(e1v, h1.data(), sizeof(double2) * n, hipMemcpyHostToDevice, stream()));
  MNP_HIP(hipMemcpyAsync(e2v, h2.data(), sizeof(double2) * n, hipMemcpyHostToDevice, stream()));
  MNP_HIP(hipStreamSynchronize(stream()));
  bool ok;
  if (solver == MNP_SOLVER_FP64) {
    Chain<double>& c = *c64;
    c.e1v = e1v;
    c.e2v = e2v;
    c.k0 = k0;
    if (g->con12) c.enable_full_L();
    g->assemble_side(1, kt, OutPair{c.G1, nullptr}, OutPair{c.S1, nullptr}, n);
    g->assemble_side(2, kt, OutPair{c.G2, nullptr}, OutPair{c.S2, nullptr}, n);
    ok = chain_build_try<double>(c, *g);
  } else {
    Chain<float>& c = *c32;
    c.e1v = e1v;
    c.e2v = e2v;
    c.k0 = k0;
    if (g->con12) c.enable_full_L();
    g->assemble_side(1, kt, OutPair{G1, c.G1}, OutPair{H1, c.S1}, n);
    g->assemble_side(2, kt, OutPair{G2, c.G2}, OutPair{H2, c.S2}, n);
    SplitBf16Scope split(solver == MNP_SOLVER_MIXED_BF16);
    ok = chain_build_try<float>(c, *g);
  }
  ready = ok;
  if (!ok) initmat_finish();
}

bool BemRet::initmat_try(double enei_, double2 e1, double2 e2) {
  MNP_CHECK(!g->compound(), MNP_ERR_STATE, "bemret_initmat: a compound takes mnp_bemret_initmat_media");
  ready = false;
  ++epoch;
  enei = enei_;
  eps1 = e1;
  eps2 = e2;
  const double k0 = 2 * M_PI / enei;
  const double2 k1 = wavenumber(k0, e1), k2 = wavenumber(k0, e2);
  bool ok;
  if (solver == MNP_SOLVER_FP64) {
    Chain<double>& c = *c64;
    c.eps1 = e1;
    c.eps2 = e2;
    c.k0 = k0;
    g->assemble(k1, 2 * M_PI, OutPair{c.G1, nullptr}, OutPair{c.S1, nullptr}, n);
    g->assemble(k2, -2 * M_PI, OutPair{c.G2, nullptr}, OutPair{c.S2, nullptr}, n);
    ok = chain_build_try<double>(c, *g);
  } else {
    Chain<float>& c = *c32;
    c.eps1 = e1;
    c.eps2 = e2;
    c.k0 = k0;
    g->assemble(k1, 2 * M_PI, OutPair{G1, c.G1}, OutPair{H1, c.S1}, n);
    g->assemble(k2, -2 * M_PI, OutPair{G2, c.G2}, OutPair{H2, c.S2}, n);
    SplitBf16Scope split(solver == MNP_SOLVER_MIXED_BF16);
    ok = chain_build_try<float>(c, *g);
  }
  ready = ok;
  return ok;
}
void BemRet::initmat_finish() {
  if (solver == MNP_SOLVER_FP64) {
    chain_build_finish<double>(*c64, *g);
  } else {
    SplitBf16Scope split(solver == MNP_SOLVER_MIXED_BF16);
    chain_build_finish<float>(*c32, *g);
  }
  ready = true;
}

double BemRet::norm2(const double2* v, int64_t len) { return dev_norm2(len, v, vws, 5); }

void BemRet::solve(int64_t p, const double2* const in[8], double2* sig1, double2* sig2, double2* h1,
                   double2* h2) {
  MNP_CHECK(ready, MNP_ERR_STATE, "bemret_solve: initmat not called");
  MNP_CHECK(p > 0 && p <= 64, MNP_ERR_ARG, "bemret_solve: npol must be in 1..64");
  const int64_t len = 8 * p * n;
  const double k0 = 2 * M_PI / enei;
  double2* b = (double2*)vws.get(0, sizeof(double2) * len);
  double2* x = (double2*)vws.get(1, sizeof(double2) * len);
  hipLaunchKernelGGL(k_excitation, dim3((unsigned)cdiv_up(n, 256), (unsigned)p), dim3(256), 0,
                     stream(), n, p, in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], b,
                     g->geo.nvec, k0, eps1, eps2, (const double2*)e1v, (const double2*)e2v);
  MNP_HIP(hipGetLastError());
  solve_b(p, b, x);
  hipLaunchKernelGGL(k_unpack, dim3((unsigned)cdiv_up(n, 256), (unsigned)(8 * p)), dim3(256), 0,
                     stream(), n, p, x, sig1, h1, sig2, h2);
  MNP_HIP(hipGetLastError());
}

// x = A^{-1} b for b, x n x 8p (layouts of chain_apply); b is not modified
void BemRet::solve_b(int64_t p, const double2* b, double2* x) {
  const int64_t len = 8 * p * n;
  const double k0 = 2 * M_PI / enei;
  const size_t tmp_elems = (size_t)((g->compound() ? 26 : 18) * p * n);
  last_iter = 0;
  last_relres = 0;
  if (solver == MNP_SOLVER_FP64) {
    double2* tmp = (double2*)vws.get(2, sizeof(double2) * tmp_elems);
    chain_apply<double>(*c64, *g, p, b, x, tmp);
  } else {
    Chain<float>& c = *c32;
    float2* b32 = (float2*)vws.get(2, sizeof(float2) * len);
    float2* x32 = (float2*)vws.get(3, sizeof(float2) * len);
    float2* tmp = (float2*)vws.get(4, sizeof(float2) * tmp_elems);
    // residual buffers: 4 products (n x 4p) + r (n x 8p), FP64; nested media
    // also G [eps s | eps h] (2 x n x 4p) and the scaled vectors (2 x n x 8p)
    const bool nested = g->compound() && g->con12;
    double2* work = (double2*)res_ws.get(0, sizeof(double2) * (size_t)((nested ? 48 : 24) * p * n));
    double2 *GX1 = work, *GX2 = work + 4 * p * n, *HX1 = work + 8 * p * n, *HX2 = work + 12 * p * n,
            *rr = work + 16 * p * n;
    double2 *GE1 = work + 24 * p * n, *GE2 = work + 28 * p * n, *xe = work + 32 * p * n, *xm = work + 40 * p * n;
    auto precond = [&](const double2* src, double2* dst) {
      convert_mat<double, float>(n, 8 * p, MatView<double>{const_cast<double2*>(src), n},
                                 MatView<float>{b32, n});
      chain_apply<float>(c, *g, p, b32, x32, tmp);
      convert_mat<float, double>(n, 8 * p, MatView<float>{x32, n}, MatView<double>{dst, n});
    };
    precond(b, x);
    const double bn = norm2(b, len);
    const double2 one = {1, 0}, zero = {0, 0};
    for (int it = 0;; ++it) {
      gemm<double>(n, 4 * p, n, one, MatView<double>{G1, n}, MatView<double>{x, n}, zero,
                   MatView<double>{GX1, n});
      gemm<double>(n, 4 * p, n, one, MatView<double>{G2, n}, MatView<double>{x + 4 * p * n, n}, zero,
                   MatView<double>{GX2, n});
      if (nested) {
        // the operator the chain inverts with full L1, L2 (eps on the charges)
        hipLaunchKernelGGL(k_scale_x, dim3((unsigned)cdiv_up(n, 256), (unsigned)(8 * p)), dim3(256), 0, stream(), n,
                           p, x, (const double2*)e1v, (const double2*)e2v, xe, xm);
        gemm<double>(n, 4 * p, n, one, MatView<double>{G1, n}, MatView<double>{xe, n}, zero,
                     MatView<double>{GE1, n});
        gemm<double>(n, 4 * p, n, one, MatView<double>{G2, n}, MatView<double>{xe + 4 * p * n, n}, zero,
                     MatView<double>{GE2, n});
        gemm<double>(n, 4 * p, n, one, MatView<double>{H1, n}, MatView<double>{xm, n}, zero,
                     MatView<double>{HX1, n});
        gemm<double>(n, 4 * p, n, one, MatView<double>{H2, n}, MatView<double>{xm + 4 * p * n, n}, zero,
                     MatView<double>{HX2, n});
        hipLaunchKernelGGL(k_residual_nested, dim3((unsigned)cdiv_up(n, 256), (unsigned)p), dim3(256), 0, stream(),
                           n, p, b, GX1, GX2, GE1, GE2, HX1, HX2, rr, g->geo.nvec, k0);
      } else {
        gemm<double>(n, 4 * p, n, one, MatView<double>{H1, n}, MatView<double>{x, n}, zero,
                     MatView<double>{HX1, n});
        gemm<double>(n, 4 * p, n, one, MatView<double>{H2, n}, MatView<double>{x + 4 * p * n, n}, zero,
                     MatView<double>{HX2, n});
        hipLaunchKernelGGL(k_residual, dim3((unsigned)cdiv_up(n, 256), (unsigned)p), dim3(256), 0,
                           stream(), n, p, b, GX1, GX2, HX1, HX2, rr, g->geo.nvec, k0, eps1, eps2,
                           (const double2*)e1v, (const double2*)e2v);
      }
      MNP_HIP(hipGetLastError());
      const double rn = norm2(rr, len);
      last_relres = bn > 0 ? rn / bn : 0;
      last_iter = it;
      if (!(last_relres > tol) || it >= maxit) break;
      precond(rr, GX1);  // correction into GX1 (n x 8p fits in GX1..GX2)
      hipLaunchKernelGGL(k_axpy, dim3((unsigned)cdiv_up(len, 256)), dim3(256), 0, stream(), len, x,
                         GX1);
      MNP_HIP(hipGetLastError());
    }
    MNP_CHECK(!(last_relres > tol), MNP_ERR_NOCONV,
              "bemret_solve: refinement stalled at relative residual " + sci(last_relres));
  }
}

// Derivative of the solution x = [sig1 | h1 | sig2 | h2] with respect to
// the inside dielectric function eps1 at the current wavelength (complex
// derivative; A(eps1) x = b with b from the outside excitation):
//   A dx = -(dA/deps1) x,
// solved with the factors of the forward solve.  eps1 enters A through
// k1 = k0 sqrt(eps1) in G1, H1 (dk1/deps1 = k1 / (2 eps1)) and explicitly
// in the alpha and De rows.
void BemRet::tangent_eps1(int64_t p, const double2* sig1, const double2* sig2, const double2* h1,
                          const double2* h2, double2* dsig1, double2* dsig2, double2* dh1, double2* dh2) {
  MNP_CHECK(ready, MNP_ERR_STATE, "bemret_tangent: initmat not called");
  MNP_CHECK(!g->compound(), MNP_ERR_STATE, "bemret_tangent: not for compound particles");
  MNP_CHECK(p > 0 && p <= 64, MNP_ERR_ARG, "bemret_tangent: npol must be in 1..64");
  const int64_t len = 8 * p * n, nn = n * n;
  const double k0 = 2 * M_PI / enei;
  const double2 k1 = wavenumber(k0, eps1);
  const double2 dk = cdiv<double>(k1, cscale<double>(eps1, 2.0));  // dk1 / deps1
  double2* x = (double2*)tws.get(0, sizeof(double2) * len);
  double2* rhs = (double2*)tws.get(1, sizeof(double2) * len);
  double2* dx = (double2*)tws.get(2, sizeof(double2) * len);
  double2* prod = (double2*)tws.get(3, sizeof(double2) * 16 * p * n);  // G1 x1, H1 x1, dG1 x1, dH1 x1
  double2* dG = (double2*)tws.get(4, sizeof(double2) * nn);
  double2* dH = (double2*)tws.get(5, sizeof(double2) * nn);
  hipLaunchKernelGGL(k_pack, dim3((unsigned)cdiv_up(n, 256), (unsigned)(8 * p)), dim3(256), 0, stream(), n, p,
                     sig1, h1, sig2, h2, x);
  MNP_HIP(hipGetLastError());
  // G1, H1 themselves: kept by the mixed solver, re-assembled for the FP64 one
  const double2 *G1m = G1, *H1m = H1;
  if (!G1m) {
    double2* tg = (double2*)tws.get(6, sizeof(double2) * nn);
    double2* th = (double2*)tws.get(7, sizeof(double2) * nn);
    g->assemble(k1, 2 * M_PI, OutPair{tg, nullptr}, OutPair{th, nullptr}, n);
    G1m = tg;
    H1m = th;
  }
  g->assemble_dk(k1, OutPair{dG, nullptr}, OutPair{dH, nullptr}, n);
  const double2 one = {1, 0}, zero = {0, 0};
  const MatView<double> X1{x, n};
  double2 *GX = prod, *HX = prod + 4 * p * n, *dGX = prod + 8 * p * n, *dHX = prod + 12 * p * n;
  gemm<double>(n, 4 * p, n, one, MatView<double>{const_cast<double2*>(G1m), n}, X1, zero, MatView<double>{GX, n});
  gemm<double>(n, 4 * p, n, one, MatView<double>{const_cast<double2*>(H1m), n}, X1, zero, MatView<double>{HX, n});
  gemm<double>(n, 4 * p, n, dk, MatView<double>{dG, n}, X1, zero, MatView<double>{dGX, n});
  gemm<double>(n, 4 * p, n, dk, MatView<double>{dH, n}, X1, zero, MatView<double>{dHX, n});
  hipLaunchKernelGGL(k_deps1, dim3((unsigned)cdiv_up(n, 256), (unsigned)p), dim3(256), 0, stream(), n, p, GX, HX,
                     dGX, dHX, rhs, g->geo.nvec, k0, eps1);
  MNP_HIP(hipGetLastError());
  solve_b(p, rhs, dx);
  hipLaunchKernelGGL(k_unpack, dim3((unsigned)cdiv_up(n, 256), (unsigned)(8 * p)), dim3(256), 0, stream(), n, p,
                     dx, dsig1, dh1, dsig2, dh2);
  MNP_HIP(hipGetLastError());
}

// ---------------------------------------------------- excitation / output
// a2 = e^{i k pos.dir}/(i k0) pol ; a2p = (i k n.dir) a2  (planewaveret/potential.m:72-76)
__global__ void k_planewave(int64_t n, int64_t p, const double* pos, const double* nvec,
                            const double* pol, const double* dir, double2 k, double k0, double2* a,
                            double2* ap) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  double pd = 0, nd = 0;
  for (int c = 0; c < 3; ++c) {
    pd += pos[c * n + i] * dir[c * p + q];
    nd += nvec[c * n + i] * dir[c * p + q];
  }
  const double2 ph0 = cexp_ikd(k, pd);  // exp(i k pd) with pd possibly negative
  const double2 ph = cdiv<double>(ph0, mk<double>(0, k0));
  const double2 ikn = cmul<double>(mk<double>(0, nd), k);
  const double2 php = cmul<double>(ikn, ph);
  for (int c = 0; c < 3; ++c) {
    const double pc = pol[c * p + q];
    a[(c + 3 * q) * n + i] = cscale<double>(ph, pc);
    ap[(c + 3 * q) * n + i] = cscale<double>(php, pc);
  }
}

// Far field (spectrumret/farfield.m:59-84) of the outside charges/currents,
// one workgroup per direction:
//   e = i k0 sum_j ph_j h2_j - i k dir sum_j ph_j sig2_j,
//   h = i k dir x sum_j ph_j h2_j,   ph_j = e^{-i k dir.pos_j} area_j.
constexpr int FF_MAXP = 8;
__global__ void __launch_bounds__(256) k_farfield(int64_t n, int64_t p, int64_t ndir, const double* dir,
                                                  const double* pos, const double* area,
                                                  const double2* sig2, const double2* h2, double2 k,
                                                  double k0, double2* e, double2* h) {
  const int64_t d = blockIdx.x;
  const double dx = dir[d], dy = dir[ndir + d], dz = dir[2 * ndir + d];
  double2 acc[FF_MAXP][4];
  for (int q = 0; q < FF_MAXP; ++q)
    for (int c = 0; c < 4; ++c) acc[q][c] = mk<double>(0, 0);
  const double2 mk_ = mk<double>(-k.x, -k.y);
  for (int64_t j = threadIdx.x; j < n; j += blockDim.x) {
    const double pd = dx * pos[j] + dy * pos[n + j] + dz * pos[2 * n + j];
    const double2 ph = cscale<double>(cexp_ikd(mk_, pd), area[j]);
    for (int q = 0; q < p && q < FF_MAXP; ++q) {
      for (int c = 0; c < 3; ++c) acc[q][c] = cadd<double>(acc[q][c], cmul<double>(ph, h2[(c + 3 * q) * n + j]));
      acc[q][3] = cadd<double>(acc[q][3], cmul<double>(ph, sig2[q * n + j]));
    }
  }
  __shared__ double2 red[256];
  for (int q = 0; q < p && q < FF_MAXP; ++q)
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
    const double2 ik0 = mk<double>(0, k0), ik = mk<double>(-k.y, k.x);
    const double dv[3] = {dx, dy, dz};
    for (int q = 0; q < p && q < FF_MAXP; ++q) {
      const double2 s = acc[q][3];
      for (int c = 0; c < 3; ++c) {
        const double2 ev = csub<double>(cmul<double>(ik0, acc[q][c]), cscale<double>(cmul<double>(ik, s), dv[c]));
        e[d + ndir * (c + 3 * q)] = ev;
      }
      // dir x A
      const double2 A0 = acc[q][0], A1 = acc[q][1], A2 = acc[q][2];
      const double2 cx = csub<double>(cscale<double>(A2, dy), cscale<double>(A1, dz));
      const double2 cy = csub<double>(cscale<double>(A0, dz), cscale<double>(A2, dx));
      const double2 cz = csub<double>(cscale<double>(A1, dx), cscale<double>(A0, dy));
      h[d + ndir * (0 + 3 * q)] = cmul<double>(ik, cx);
      h[d + ndir * (1 + 3 * q)] = cmul<double>(ik, cy);
      h[d + ndir * (2 + 3 * q)] = cmul<double>(ik, cz);
    }
  }
}

// Cross sections from the far fields of the npol propagation directions
// (first npol of the nd directions) and of the npinf pinfty normals pn with
// areas pa (he, hh: nd x 3 x npol), in FP64:
//   ext = 4 pi / Re k Im(pol . e(dir_q))            (planewaveret/extinction.m:12-16)
//   sca = sum area 0.5 Re(n . (e x conj h)) / (0.5 nb)  (scattering.m:20-29)
// and abs = ext - sca; any output may be NULL.
void farfield_host_sums(int64_t npol, const double* pol, int64_t nd, int64_t npinf, const double2* he,
                        const double2* hh, const double* pn, const double* pa, double2 k, double nb, double* ext,
                        double* sca, double* abs_) {
  for (int64_t q = 0; q < npol; ++q) {
    double2 pe = mk<double>(0, 0);
    for (int c = 0; c < 3; ++c) pe = cadd<double>(pe, cscale<double>(he[q + nd * (c + 3 * q)], pol[c * npol + q]));
    const double e_val = 4 * M_PI / k.x * pe.y;
    double s = 0;
    for (int64_t d = 0; d < npinf; ++d) {
      const int64_t dd = npol + d;
      double2 E[3], H[3];
      for (int c = 0; c < 3; ++c) {
        E[c] = he[dd + nd * (c + 3 * q)];
        H[c] = hh[dd + nd * (c + 3 * q)];
        H[c].y = -H[c].y;
      }
      const double2 x = csub<double>(cmul<double>(E[1], H[2]), cmul<double>(E[2], H[1]));
      const double2 y = csub<double>(cmul<double>(E[2], H[0]), cmul<double>(E[0], H[2]));
      const double2 z = csub<double>(cmul<double>(E[0], H[1]), cmul<double>(E[1], H[0]));
      const double ds = 0.5 * (pn[d] * x.x + pn[npinf + d] * y.x + pn[2 * npinf + d] * z.x);
      s += pa[d] * ds;
    }
    const double s_val = s / (0.5 * nb);
    if (ext) ext[q] = e_val;
    if (sca) sca[q] = s_val;
    if (abs_) abs_[q] = e_val - s_val;
  }
}

static double2 medium_k(double enei, mnp_complex em) {
  const double k0 = 2 * M_PI / enei;
  const double r = std::hypot(em.re, em.im);
  double sr = std::sqrt(0.5 * (r + em.re)), si = std::sqrt(0.5 * (r - em.re));
  if (em.im < 0) si = -si;
  return mk<double>(k0 * sr, k0 * si);
}

void farfield(const GreenRet& g, double enei, mnp_complex em, int64_t ndir, const double* dir,
                     int64_t p, const double2* sig2, const double2* h2, double2* e, double2* h) {
  MNP_CHECK(p >= 1 && p <= FF_MAXP, MNP_ERR_ARG, "farfield: npol must be in 1..8");
  const double2 k = medium_k(enei, em);
  hipLaunchKernelGGL(k_farfield, dim3((unsigned)ndir), dim3(256), 0, stream(), g.n, p, ndir, dir,
                     g.geo.pos, g.geo.area, sig2, h2, k, 2 * M_PI / enei, e, h);
  MNP_HIP(hipGetLastError());
}

// ------------------------------------- cross sections without allocation
// Scratch of mnp_planewave_cross_sections_noalloc, one copy per GPU for the
// life of the process: the ndir = npol + npinf directions (3 x ndir
// doubles), the far fields e and h (ndir x 3 x npol complex each) and the
// per-polarisation results (ext, sca).  The sizes below are the limits the
// entry point accepts.
constexpr int64_t NA_MAX_DIRS = 4096;
constexpr int64_t NA_DIR_BYTES = 3 * NA_MAX_DIRS * sizeof(double);
constexpr int64_t NA_FIELD_BYTES = NA_MAX_DIRS * 3 * FF_MAXP * sizeof(double2);
constexpr int64_t NA_RES_BYTES = 2 * FF_MAXP * sizeof(double);
constexpr int64_t NA_SCRATCH_BYTES = NA_DIR_BYTES + 2 * NA_FIELD_BYTES + NA_RES_BYTES;
static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

// Device address of this GPU's copy of `scratch`: looked up with
// hipGetSymbolAddress once per device, on the first call with that device
// current (the array's host-side name is not a device address).
static void* noalloc_scratch() {
  static std::mutex mu;
  static void* ptr[256] = {};
  int dev = 0;
  MNP_HIP(hipGetDevice(&dev));
  MNP_CHECK(dev >= 0 && dev < 256, MNP_ERR_HIP, "cross_sections_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    MNP_CHECK(e == hipSuccess && p, MNP_ERR_HIP,
              std::string("cross_sections_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    ptr[dev] = p;
  }
  return ptr[dev];
}

struct PolDirs {
  double pol[FF_MAXP][3];
};

// One workgroup per polarisation q, from the far fields of k_farfield at the
// directions [dir_0 .. dir_{npol-1} | pinfty normals]:
//   ext_q = 4 pi / Re k Im(pol_q . e(dir_q))              (extinction.m:12-16)
//   sca_q = sum_d area_d 0.5 Re(n_d . (e x conj h)) / (0.5 nb)  (scattering.m)
// res[2q] = ext_q, res[2q + 1] = sca_q.
__global__ void __launch_bounds__(256) k_xsec_reduce(int64_t p, int64_t npinf, const double2* __restrict__ ef,
                                                     const double2* __restrict__ hf, const double* __restrict__ pn,
                                                     const double* __restrict__ pa, PolDirs pd, double kre, double nb,
                                                     double* res) {
  const int64_t q = blockIdx.x;
  const int64_t nd = p + npinf;
  double s = 0;
  for (int64_t d = threadIdx.x; d < npinf; d += blockDim.x) {
    const int64_t dd = p + d;
    double2 E[3], H[3];
    for (int c = 0; c < 3; ++c) {
      E[c] = ef[dd + nd * (c + 3 * q)];
      H[c] = hf[dd + nd * (c + 3 * q)];
      H[c].y = -H[c].y;
    }
    const double2 x = csub<double>(cmul<double>(E[1], H[2]), cmul<double>(E[2], H[1]));
    const double2 y = csub<double>(cmul<double>(E[2], H[0]), cmul<double>(E[0], H[2]));
    const double2 z = csub<double>(cmul<double>(E[0], H[1]), cmul<double>(E[1], H[0]));
    s += pa[d] * (0.5 * (pn[d] * x.x + pn[npinf + d] * y.x + pn[2 * npinf + d] * z.x));
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double pe = 0;
    for (int c = 0; c < 3; ++c) pe += ef[q + nd * (c + 3 * q)].y * pd.pol[q][c];
    res[2 * q] = 4 * M_PI / kre * pe;
    res[2 * q + 1] = red[0] / (0.5 * nb);
  }
}

}  // namespace mnp

// ============================================================== C ABI
using namespace mnp;


extern "C" {

int mnp_bemret_create(mnp_greenret_t g, int solver, mnp_bemret_t* out) {
  return guarded([&] {
    MNP_CHECK(g && out, MNP_ERR_ARG, "bemret_create: NULL argument");
    MNP_CHECK(solver == MNP_SOLVER_FP64 || solver == MNP_SOLVER_MIXED || solver == MNP_SOLVER_MIXED_BF16,
              MNP_ERR_ARG, "bemret_create: unknown solver");
    auto* h = new mnp_bemret_s{nullptr};
    try {
      h->b = new BemRet(g->g, solver);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int mnp_bemret_destroy(mnp_bemret_t bem) {
  return guarded([&] {
    if (!bem) return;
    delete bem->b;
    delete bem;
  });
}

int mnp_bemret_initmat(mnp_bemret_t bem, double enei, mnp_complex eps1, mnp_complex eps2) {
  return guarded([&] {
    MNP_CHECK(bem, MNP_ERR_ARG, "bemret_initmat: NULL handle");
    MNP_CHECK(enei > 0, MNP_ERR_ARG, "bemret_initmat: wavelength must be positive");
    bem->b->initmat(enei, to_d2(eps1), to_d2(eps2));
  });
}

int mnp_bemret_initmat_media(mnp_bemret_t bem, double enei, int32_t nmed, const mnp_complex* eps) {
  return guarded([&] {
    MNP_CHECK(bem && eps, MNP_ERR_ARG, "bemret_initmat_media: NULL argument");
    MNP_CHECK(enei > 0, MNP_ERR_ARG, "bemret_initmat_media: wavelength must be positive");
    MNP_CHECK(nmed >= 1 && nmed <= MNP_MAX_MED, MNP_ERR_ARG, "bemret_initmat_media: 1..16 media");
    std::vector<double2> e(nmed);
    for (int m = 0; m < nmed; ++m) e[m] = to_d2(eps[m]);
    bem->b->initmat_media(enei, e);
  });
}

int mnp_bemret_initmat_many(int count, const mnp_bemret_t* bems, const double* enei, const mnp_complex* eps1,
                            const mnp_complex* eps2) {
  return guarded([&] {
    MNP_CHECK(count >= 0 && (count == 0 || (bems && enei && eps1 && eps2)), MNP_ERR_ARG,
              "bemret_initmat_many: bad argument");
    for (int i = 0; i < count; ++i) {
      MNP_CHECK(bems[i] && enei[i] > 0, MNP_ERR_ARG, "bemret_initmat_many: NULL handle or bad wavelength");
      for (int j = 0; j < i; ++j)
        MNP_CHECK(bems[j] != bems[i], MNP_ERR_ARG, "bemret_initmat_many: a handle appears twice");
    }
    // groups of four on concurrent streams; pivoted refactorisations (a
    // cooperative launch) afterwards, one solver at a time
    for (int i0 = 0; i0 < count; i0 += 4) {
      const int i1 = std::min(count, i0 + 4);
      std::vector<char> ok(i1 - i0, 0);
      std::vector<std::function<void()>> tasks;
      for (int i = i0; i < i1; ++i)
        tasks.push_back([&, i] { ok[i - i0] = bems[i]->b->initmat_try(enei[i], to_d2(eps1[i]), to_d2(eps2[i])); });
      run_concurrent(tasks);
      for (int i = i0; i < i1; ++i)
        if (!ok[i - i0]) bems[i]->b->initmat_finish();
    }
  });
}

int mnp_bemret_solve(mnp_bemret_t bem, int64_t npol, const void* phi1, const void* phi1p,
                     const void* a1, const void* a1p, const void* phi2, const void* phi2p,
                     const void* a2, const void* a2p, void* sig1, void* sig2, void* h1, void* h2) {
  return guarded([&] {
    MNP_CHECK(bem, MNP_ERR_ARG, "bemret_solve: NULL handle");
    const double2* in[8] = {(const double2*)phi1, (const double2*)phi1p, (const double2*)a1,
                            (const double2*)a1p,  (const double2*)phi2,  (const double2*)phi2p,
                            (const double2*)a2,   (const double2*)a2p};
    bem->b->solve(npol, in, (double2*)sig1, (double2*)sig2, (double2*)h1, (double2*)h2);
  });
}

int mnp_bemret_stats(mnp_bemret_t bem, int* iterations, double* relres) {
  return guarded([&] {
    MNP_CHECK(bem, MNP_ERR_ARG, "bemret_stats: NULL handle");
    if (iterations) *iterations = bem->b->last_iter;
    if (relres) *relres = bem->b->last_relres;
  });
}

int mnp_bemret_tangent_eps1(mnp_bemret_t bem, int64_t npol, const void* sig1, const void* sig2, const void* h1,
                            const void* h2, void* dsig1, void* dsig2, void* dh1, void* dh2) {
  return guarded([&] {
    MNP_CHECK(bem && sig1 && sig2 && h1 && h2 && dsig1 && dsig2 && dh1 && dh2, MNP_ERR_ARG,
              "bemret_tangent_eps1: NULL argument");
    bem->b->tangent_eps1(npol, (const double2*)sig1, (const double2*)sig2, (const double2*)h1, (const double2*)h2,
                         (double2*)dsig1, (double2*)dsig2, (double2*)dh1, (double2*)dh2);
  });
}

int mnp_bemret_set_refinement(mnp_bemret_t bem, double tol, int maxit) {
  return guarded([&] {
    MNP_CHECK(bem && tol > 0 && maxit >= 0, MNP_ERR_ARG, "bemret_set_refinement: bad argument");
    bem->b->tol = tol;
    bem->b->maxit = maxit;
  });
}

int mnp_planewave_potential(mnp_greenret_t g, double enei, mnp_complex eps_medium, int64_t npol,
                            const double* pol, const double* dir, void* a2, void* a2p) {
  return guarded([&] {
    MNP_CHECK(g && pol && dir && a2 && a2p, MNP_ERR_ARG, "planewave_potential: NULL argument");
    MNP_CHECK(npol >= 1 && npol <= 64, MNP_ERR_ARG, "planewave_potential: npol out of range");
    for (int64_t q = 0; q < npol; ++q) {
      double s = 0;
      for (int c = 0; c < 3; ++c) s += pol[c * npol + q] * dir[c * npol + q];
      MNP_CHECK(s == 0, MNP_ERR_ARG, "planewave_potential: polarisation not perpendicular to direction");
    }
    const GreenRet& G = *g->g;
    double* d = dev_alloc<double>(6 * npol);
    MNP_HIP(hipMemcpyAsync(d, pol, sizeof(double) * 3 * npol, hipMemcpyHostToDevice, stream()));
    MNP_HIP(hipMemcpyAsync(d + 3 * npol, dir, sizeof(double) * 3 * npol, hipMemcpyHostToDevice, stream()));
    hipLaunchKernelGGL(k_planewave, dim3((unsigned)cdiv_up(G.n, 256), (unsigned)npol), dim3(256), 0,
                       stream(), G.n, npol, G.geo.pos, G.geo.nvec, d, d + 3 * npol,
                       medium_k(enei, eps_medium), 2 * M_PI / enei, (double2*)a2, (double2*)a2p);
    MNP_HIP(hipGetLastError());
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_HIP(hipFree(d));
  });
}

int mnp_farfield(mnp_greenret_t g, double enei, mnp_complex eps_medium, int64_t ndir, const void* dir,
                 int64_t npol, const void* sig2, const void* h2, void* e, void* h) {
  return guarded([&] {
    MNP_CHECK(g && dir && sig2 && h2 && e && h, MNP_ERR_ARG, "farfield: NULL argument");
    farfield(*g->g, enei, eps_medium, ndir, (const double*)dir, npol, (const double2*)sig2,
             (const double2*)h2, (double2*)e, (double2*)h);
  });
}

int mnp_planewave_cross_sections(mnp_greenret_t g, double enei, mnp_complex eps_medium, int64_t npol,
                                 const double* pol, const double* dir, int64_t npinf,
                                 const void* pinfty_nvec, const void* pinfty_area, const void* sig2,
                                 const void* h2, double* ext, double* sca, double* abs_) {
  return guarded([&] {
    MNP_CHECK(g && pol && dir && pinfty_nvec && pinfty_area && sig2 && h2, MNP_ERR_ARG,
              "cross_sections: NULL argument");
    MNP_CHECK(npol >= 1 && npol <= FF_MAXP, MNP_ERR_ARG, "cross_sections: npol must be in 1..8");
    const GreenRet& G = *g->g;
    const int64_t nd = npol + npinf;
    // directions: the npol propagation directions, then the pinfty normals
    double* ddir = dev_alloc<double>(3 * nd);
    for (int c = 0; c < 3; ++c) {
      MNP_HIP(hipMemcpyAsync(ddir + c * nd, dir + c * npol, sizeof(double) * npol, hipMemcpyHostToDevice,
                             stream()));
      MNP_HIP(hipMemcpyAsync(ddir + c * nd + npol, (const double*)pinfty_nvec + c * npinf,
                             sizeof(double) * npinf, hipMemcpyDeviceToDevice, stream()));
    }
    double2* ef = dev_alloc<double2>(nd * 3 * npol);
    double2* hf = dev_alloc<double2>(nd * 3 * npol);
    farfield(G, enei, eps_medium, nd, ddir, npol, (const double2*)sig2, (const double2*)h2, ef, hf);
    std::vector<double2> he(nd * 3 * npol), hh(nd * 3 * npol);
    std::vector<double> pn(3 * npinf), pa(npinf);
    MNP_HIP(hipMemcpyAsync(he.data(), ef, sizeof(double2) * he.size(), hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipMemcpyAsync(hh.data(), hf, sizeof(double2) * hh.size(), hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipMemcpyAsync(pn.data(), pinfty_nvec, sizeof(double) * 3 * npinf, hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipMemcpyAsync(pa.data(), pinfty_area, sizeof(double) * npinf, hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_HIP(hipFree(ddir));
    MNP_HIP(hipFree(ef));
    MNP_HIP(hipFree(hf));
    const double2 k = medium_k(enei, eps_medium);
    const double nb = std::sqrt(std::hypot(eps_medium.re, eps_medium.im));  // |sqrt(eps)| (lossless medium)
    farfield_host_sums(npol, pol, nd, npinf, he.data(), hh.data(), pn.data(), pa.data(), k, nb, ext, sca, abs_);
  });
}

int mnp_planewave_cross_sections_noalloc(mnp_greenret_t g, double enei, mnp_complex eps_medium, int64_t npol,
                                         const double* pol, const double* dir, int64_t npinf,
                                         const void* pinfty_nvec, const void* pinfty_area, const void* sig2,
                                         const void* h2, double* ext, double* sca, double* abs_) {
  return guarded([&] {
    MNP_CHECK(g && pol && dir && pinfty_nvec && pinfty_area && sig2 && h2, MNP_ERR_ARG,
              "cross_sections_noalloc: NULL argument");
    MNP_CHECK(npol >= 1 && npol <= FF_MAXP, MNP_ERR_ARG, "cross_sections_noalloc: npol must be in 1..8");
    MNP_CHECK(npinf >= 1 && npol + npinf <= NA_MAX_DIRS, MNP_ERR_ARG,
              "cross_sections_noalloc: npol + npinf must be at most " + std::to_string(NA_MAX_DIRS) +
                  " (the size of the library's scratch)");
    const GreenRet& G = *g->g;
    unsigned char* base = (unsigned char*)noalloc_scratch();
    // every region is written in full before it is read: nothing to clear
    double* ddir = (double*)base;
    double2* ef = (double2*)(base + NA_DIR_BYTES);
    double2* hf = (double2*)(base + NA_DIR_BYTES + NA_FIELD_BYTES);
    double* res = (double*)(base + NA_DIR_BYTES + 2 * NA_FIELD_BYTES);
    const int64_t nd = npol + npinf;
    for (int c = 0; c < 3; ++c) {
      MNP_HIP(hipMemcpyAsync(ddir + c * nd, dir + c * npol, sizeof(double) * npol, hipMemcpyHostToDevice,
                             stream()));
      MNP_HIP(hipMemcpyAsync(ddir + c * nd + npol, (const double*)pinfty_nvec + c * npinf,
                             sizeof(double) * npinf, hipMemcpyDeviceToDevice, stream()));
    }
    farfield(G, enei, eps_medium, nd, ddir, npol, (const double2*)sig2, (const double2*)h2, ef, hf);
    PolDirs pd{};
    for (int64_t q = 0; q < npol; ++q)
      for (int c = 0; c < 3; ++c) pd.pol[q][c] = pol[c * npol + q];
    const double2 k = medium_k(enei, eps_medium);
    const double nb = std::sqrt(std::hypot(eps_medium.re, eps_medium.im));  // |sqrt(eps)| (lossless medium)
    hipLaunchKernelGGL(k_xsec_reduce, dim3((unsigned)npol), dim3(256), 0, stream(), npol, npinf, ef, hf,
                       (const double*)pinfty_nvec, (const double*)pinfty_area, pd, k.x, nb, res);
    MNP_HIP(hipGetLastError());
    double h[2 * FF_MAXP];
    MNP_HIP(hipMemcpyAsync(h, res, sizeof(double) * 2 * npol, hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipStreamSynchronize(stream()));
    for (int64_t q = 0; q < npol; ++q) {
      if (ext) ext[q] = h[2 * q];
      if (sca) sca[q] = h[2 * q + 1];
      if (abs_) abs_[q] = h[2 * q] - h[2 * q + 1];
    }
  });
}

}  // extern "C"
