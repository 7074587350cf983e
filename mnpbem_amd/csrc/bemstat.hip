// Quasistatic BEM solver and plane wave on gfx950 (MNPBEM BEM/@bemstat,
// Simulation/static/@planewavestat).
//
// sig = -(Lambda I + F)^{-1} phi' with Lambda = 2 pi (eps1 + eps2)/(eps1 - eps2)
// (bemstat/subsref.m:31-40, mldivide.m:13-15).  F is the static surface
// derivative with refinement and the closed-surface correction
// (compgreenstat/init.m:24-55), which is the retarded F at k = 0: it is
// assembled once per mesh by GreenRet::assemble in FP64.  Per wavelength the
// shifted matrix is factorised (LU without row exchanges first, partial
// pivoting when the FP64 factors would differ or on a zero pivot), in FP64
// (the drop-in) or in complex FP32 with FP64 iterative refinement against
// F x + Lambda x.
#include <complex>
#include <memory>
#include <vector>

#include "mnp_linalg.hpp"
#include "mnp_mirror.hpp"

namespace mnp {

// M = F + lam I in precision T
template <class T>
__global__ void k_shift_diag(int64_t n, const double2* F, double2 lam, cplx<T>* M) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (i >= n || j >= n) return;
  double2 v = F[j * n + i];
  if (i == j) v = cadd<double>(v, lam);
  M[j * n + i] = mk<T>((T)v.x, (T)v.y);
}
// y <- a * y (complex scalar), len entries
__global__ void k_cscal(int64_t len, double2* y, double2 a) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) y[i] = cmul<double>(a, y[i]);
}
// r = -phip - lam x  (the F x part is added by a GEMM afterwards)
__global__ void k_stat_res0(int64_t len, const double2* phip, const double2* x, double2 lam, double2* r) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= len) return;
  const double2 lx = cmul<double>(lam, x[i]);
  r[i] = mk<double>(-phip[i].x - lx.x, -phip[i].y - lx.y);
}
__global__ void k_add(int64_t len, double2* x, const double2* d) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) x[i] = cadd<double>(x[i], d[i]);
}
// partial sums of |v|^2 over chunks
__global__ void k_sumsq(int64_t len, const double2* v, double* part) {
  double s = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
    s += v[i].x * v[i].x + v[i].y * v[i].y;
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// phi' = -nvec . pol  (planewavestat/potential.m:13); pol device npol x 3
__global__ void k_pwstat(int64_t n, int64_t npol, const double* nvec, const double* pol, double2* phip) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const double v =
      nvec[i] * pol[q] + nvec[n + i] * pol[npol + q] + nvec[2 * n + i] * pol[2 * npol + q];
  phip[q * n + i] = mk<double>(-v, 0.0);
}
// dipole moment partial sums: part[(q * nchunk + c) * 3 + d] = sum_i area_i pos_id sig_iq
__global__ void k_dipmoment(int64_t n, const double* pos, const double* area, const double2* sig,
                            double2* part) {
  const int64_t c = blockIdx.x, q = blockIdx.y, nchunk = gridDim.x;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t i = c * blockDim.x + threadIdx.x; i < n; i += nchunk * blockDim.x) {
    const double2 v = sig[q * n + i];
    for (int d = 0; d < 3; ++d) {
      const double w = area[i] * pos[d * n + i];
      s[2 * d] += w * v.x;
      s[2 * d + 1] += w * v.y;
    }
  }
  __shared__ double red[6][256];
  for (int t = 0; t < 6; ++t) red[t][threadIdx.x] = s[t];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int t = 0; t < 6; ++t) red[t][threadIdx.x] += red[t][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 3)
    part[(q * nchunk + c) * 3 + threadIdx.x] = mk<double>(red[2 * threadIdx.x][0], red[2 * threadIdx.x + 1][0]);
}

template <class T>
struct StatFactor {
  int64_t n;
  bool exact;
  cplx<T>*M = nullptr, *backup = nullptr, *inv = nullptr, *invnb = nullptr;
  int32_t *piv = nullptr, *sig = nullptr;
  LU<T> f{};
  int fallbacks = 0;
  Workspace ws;
  StatFactor(int64_t n_, bool exact_) : n(n_), exact(exact_) {
    M = dev_alloc<cplx<T>>(n * n);
    backup = dev_alloc<cplx<T>>(n * n);
    inv = dev_alloc<cplx<T>>(2 * inv_blocks(n));
    invnb = dev_alloc<cplx<T>>(2 * inv_blocks_nb(n));
    piv = dev_alloc<int32_t>(n);
    sig = dev_alloc<int32_t>(n);
  }
  ~StatFactor() {
    void* ps[] = {M, backup, inv, invnb, piv, sig};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
  StatFactor(const StatFactor&) = delete;
  StatFactor& operator=(const StatFactor&) = delete;
  // M = F + lam I, factorised
  void build(const double2* F, double2 lam) {
    const MatView<T> Mv{M, n};
    hipLaunchKernelGGL(k_shift_diag<T>, dim3((unsigned)cdiv_up(n, 256), (unsigned)std::min<int64_t>(n, 65535),
                                             (unsigned)cdiv_up(n, 65535)),
                       dim3(256), 0, stream(), n, F, lam, M);
    MNP_HIP(hipGetLastError());
    copy_mat<T>(n, n, Mv, MatView<T>{backup, n});
    cplx<T>*Linv = inv, *Uinv = inv + inv_blocks(n);
    int swap = 0;
    int info = getrf<T>(n, n, Mv, piv, Linv, Uinv, PIVOT_NONE, &swap, ws);
    const int32_t* sigma = nullptr;
    if (info != 0 || (swap && exact)) {
      ++fallbacks;
      copy_mat<T>(n, n, MatView<T>{backup, n}, Mv);
      info = getrf<T>(n, n, Mv, piv, Linv, Uinv, PIVOT_PARTIAL, nullptr, ws);
      perm_from_pivots(n, piv, sig, ws);
      sigma = sig;
    }
    MNP_CHECK(info == 0, MNP_ERR_SINGULAR, "bemstat: singular Lambda + F (zero pivot " + std::to_string(info) + ")");
    cplx<T>*LinvNB = invnb, *UinvNB = invnb + inv_blocks_nb(n);
    block_inverses_nb<T>(n, Mv, Linv, Uinv, LinvNB, UinvNB, ws);
    f = LU<T>{Mv, Linv, Uinv, sigma, LinvNB, UinvNB};
  }
};

// F = F0 + f2 Fr  (compgreenstatlayer/eval.m:33-38 for a particle above the interface)
__global__ void k_axpby2(int64_t len, const double2* F0, const double2* Fr, double2 f2, double2* F) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) F[i] = cadd<double>(F0[i], cmul<double>(f2, Fr[i]));
}
// phi' = -nvec . e, e complex (npol x 3 in constant-size argument)
constexpr int MNP_STAT_MAXPOL = 64;
struct FieldArg {
  double2 e[MNP_STAT_MAXPOL][3];
};
__global__ void k_pwstat_field(int64_t n, const double* nvec, FieldArg a, double2* phip) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  double2 v = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) v = cadd<double>(v, cscale<double>(a.e[q][c], nvec[c * n + i]));
  phip[q * n + i] = mk<double>(-v.x, -v.y);
}

struct BemStat {
  const GreenRet* g;
  int64_t n;
  int solver;
  double2* F = nullptr;  // static surface derivative, FP64, n x n
  // substrate: F = F0 + f2 Fr per wavelength (F0 direct, Fr the reflected part)
  double2 *F0 = nullptr, *Fr = nullptr;
  std::unique_ptr<StatFactor<double>> f64;
  std::unique_ptr<StatFactor<float>> f32;
  double2 lam{};
  bool ready = false;
  double refine_tol = 1e-12;
  int refine_maxit = 10;
  int last_sweeps = 0;
  double last_relres = 0;
  Workspace vws;

  BemStat(const GreenRet* g_, int solver_) : g(g_), n(g_->n), solver(solver_) {
    F = dev_alloc<double2>(n * n);
    g->assemble(mk<double>(0, 0), 0.0, OutPair{}, OutPair{F, nullptr}, n);
    if (solver == MNP_SOLVER_FP64)
      f64.reset(new StatFactor<double>(n, true));
    else
      f32.reset(new StatFactor<float>(n, false));
  }
  // Solver for a given FP64 surface derivative F (n x n, taken over: freed
  // with the solver); the quasistatic mirror solver makes one per symmetry value.
  BemStat(int64_t n_, double2* F_owned, int solver_) : g(nullptr), n(n_), solver(solver_), F(F_owned) {
    if (solver == MNP_SOLVER_FP64)
      f64.reset(new StatFactor<double>(n, true));
    else
      f32.reset(new StatFactor<float>(n, false));
  }
  // Quasistatic solver above a substrate (BEM/@bemstatlayer): gimg is the
  // Green function of [particle; its image in the interface] without the
  // closed-surface correction; the reflected part is its rows of the
  // particle against the image's columns.
  BemStat(const GreenRet* g_, const GreenRet* gimg, int solver_) : BemStat(g_, solver_) {
    F0 = F;
    F = dev_alloc<double2>(n * n);
    Fr = dev_alloc<double2>(n * n);
    double2* B = dev_alloc<double2>(n * gimg->n);
    try {
      gimg->assemble_rows(mk<double>(0, 0), 0.0, OutPair{}, OutPair{B, nullptr}, n, n);
      MNP_HIP(hipMemcpyAsync(Fr, B + n * n, sizeof(double2) * n * n, hipMemcpyDeviceToDevice, stream()));
      MNP_HIP(hipStreamSynchronize(stream()));
    } catch (...) {
      (void)hipFree(B);
      throw;
    }
    MNP_HIP(hipFree(B));
  }
  ~BemStat() {
    if (F) (void)hipFree(F);
    if (F0) (void)hipFree(F0);
    if (Fr) (void)hipFree(Fr);
  }

  // F = F0 + f2 Fr, then initmat
  void initmat_layer(double2 e1, double2 e2, double2 f2) {
    MNP_CHECK(Fr, MNP_ERR_STATE, "bemstat: not created for a layer structure");
    const int64_t len = n * n;
    hipLaunchKernelGGL(k_axpby2, dim3((unsigned)cdiv_up(len, 256)), dim3(256), 0, stream(), len, F0, Fr, f2, F);
    MNP_HIP(hipGetLastError());
    initmat(e1, e2);
  }

  void initmat(double2 e1, double2 e2) {
    ready = false;  // until the factorisation is complete
    // Lambda = 2 pi (eps1 + eps2) / (eps1 - eps2)
    lam = cscale<double>(cdiv<double>(cadd<double>(e1, e2), csub<double>(e1, e2)), 2 * M_PI);
    if (f64) {
      f64->build(F, lam);
    } else {
      SplitBf16Scope split(solver == MNP_SOLVER_MIXED_BF16);
      f32->build(F, lam);
    }
    ready = true;
  }

  double norm(const double2* v, int64_t len) {
    const int nb = 256;
    double* part = (double*)vws.get(5, sizeof(double) * nb);
    hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(256), 0, stream(), len, v, part);
    std::vector<double> h(nb);
    MNP_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipStreamSynchronize(stream()));
    double s = 0;
    for (double x : h) s += x;
    return std::sqrt(s);
  }

  // sig = -(Lambda + F)^{-1} phip, n x p
  void solve(int64_t p, const double2* phip, double2* sig) {
    MNP_CHECK(ready, MNP_ERR_STATE, "bemstat_solve: initmat not called");
    MNP_CHECK(p >= 1, MNP_ERR_ARG, "bemstat_solve: npol < 1");
    const int64_t len = n * p;
    const unsigned gl = (unsigned)cdiv_up(len, 256);
    if (f64) {
      copy_mat<double>(n, p, MatView<double>{const_cast<double2*>(phip), n}, MatView<double>{sig, n});
      getrs<double>(n, p, f64->f, MatView<double>{sig, n}, f64->ws);
      hipLaunchKernelGGL(k_cscal, dim3(gl), dim3(256), 0, stream(), len, sig, mk<double>(-1, 0));
      MNP_HIP(hipGetLastError());
      last_sweeps = 0;
      last_relres = 0;
      return;
    }
    // mixed: x0 = M32^{-1}(-phip), then x += M32^{-1}(-phip - (F + Lambda) x)
    double2* r = (double2*)vws.get(0, sizeof(double2) * len);
    float2* r32 = (float2*)vws.get(1, sizeof(float2) * len);
    const double bn = norm(phip, len);
    MNP_HIP(hipMemsetAsync(sig, 0, sizeof(double2) * len, stream()));
    int it = 0;
    double rel = 1.0;
    for (; it <= refine_maxit; ++it) {
      hipLaunchKernelGGL(k_stat_res0, dim3(gl), dim3(256), 0, stream(), len, phip, sig, lam, r);
      gemm<double>(n, p, n, mk<double>(-1, 0), MatView<double>{F, n}, MatView<double>{sig, n}, mk<double>(1, 0),
                   MatView<double>{r, n});
      rel = bn > 0 ? norm(r, len) / bn : 0.0;
      if (rel <= refine_tol) break;
      if (it == refine_maxit) break;
      convert_mat<double, float>(n, p, MatView<double>{r, n}, MatView<float>{r32, n});
      getrs<float>(n, p, f32->f, MatView<float>{r32, n}, f32->ws);
      convert_mat<float, double>(n, p, MatView<float>{r32, n}, MatView<double>{r, n});
      hipLaunchKernelGGL(k_add, dim3(gl), dim3(256), 0, stream(), len, sig, r);
      MNP_HIP(hipGetLastError());
    }
    last_sweeps = it;
    last_relres = rel;
    MNP_CHECK(rel <= refine_tol, MNP_ERR_NOCONV,
              "bemstat_solve: refinement stopped at relative residual " + std::to_string(rel));
  }
};

// Quasistatic solver with mirror symmetry (BEM/@bemstatmirror): one solver
// per symmetry value s with F_s = sum_j tab[s][j] F(p, copy j)
// (compgreenstatmirror/eval.m:18-40) and mat_s = -(Lambda + F_s)^{-1}
// (bemstatmirror/subsref.m:29-37).  g is the full particle (J n faces, copy
// j in faces j n ..); the n x J n rows of the part are assembled once per
// mesh in FP64 and contracted into the J matrices.  A surface charge goes
// through the solver of the symmetry value of its potential
// (bemstatmirror/mldivide.m:14-22).
struct BemStatMirror {
  int J;
  int64_t n;
  std::vector<std::unique_ptr<BemStat>> s;

  BemStatMirror(const GreenRet* g, int solver, int J_, const int* table) : J(J_), n(g->n / J_) {
    SymTab tab{};
    check_symtable("bemstatmirror_create", J, table, tab);
    double2* B = dev_alloc<double2>(n * g->n);
    std::vector<double2*> Fs(J, nullptr);
    try {
      g->assemble_rows(mk<double>(0, 0), 0.0, OutPair{}, OutPair{B, nullptr}, n, n);
      OutPair outs[4];
      for (int k = 0; k < J; ++k) {
        Fs[k] = dev_alloc<double2>(n * n);
        outs[k] = OutPair{Fs[k], nullptr};
      }
      sym_contract(J, n, tab, B, outs);
      MNP_HIP(hipStreamSynchronize(stream()));
      for (int k = 0; k < J; ++k) {
        s.emplace_back(new BemStat(n, Fs[k], solver));
        Fs[k] = nullptr;  // owned by s[k] now
      }
    } catch (...) {
      (void)hipFree(B);
      for (double2* p : Fs)
        if (p) (void)hipFree(p);
      throw;
    }
    MNP_HIP(hipFree(B));
  }
  void initmat(double2 e1, double2 e2) {
    for (auto& b : s) b->ready = false;
    for (auto& b : s) b->initmat(e1, e2);
  }
};

}  // namespace mnp

// ============================================================== C ABI
using namespace mnp;

struct mnp_bemstat_s {
  BemStat* b;
};
struct mnp_bemstatmirror_s {
  BemStatMirror* b;
};

extern "C" {

int mnp_bemstat_create(mnp_greenret_t g, int solver, mnp_bemstat_t* out) {
  return guarded([&] {
    MNP_CHECK(!g || !g->g->compound(), MNP_ERR_ARG, "bemstat_create: compound particles take the dense solver (mnp_bemret_*)");
    MNP_CHECK(g && out, MNP_ERR_ARG, "bemstat_create: NULL argument");
    MNP_CHECK(solver == MNP_SOLVER_FP64 || solver == MNP_SOLVER_MIXED || solver == MNP_SOLVER_MIXED_BF16, MNP_ERR_ARG,
              "bemstat_create: unknown solver");
    auto* h = new mnp_bemstat_s{nullptr};
    try {
      h->b = new BemStat(g->g, solver);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int mnp_bemstatlayer_create(mnp_greenret_t g, mnp_greenret_t gimage, int solver, mnp_bemstat_t* out) {
  return guarded([&] {
    MNP_CHECK(!g || !g->g->compound(), MNP_ERR_ARG, "bemstatlayer_create: compound particles take the dense solver (mnp_bemret_*)");
    MNP_CHECK(g && gimage && out, MNP_ERR_ARG, "bemstatlayer_create: NULL argument");
    MNP_CHECK(solver == MNP_SOLVER_FP64 || solver == MNP_SOLVER_MIXED || solver == MNP_SOLVER_MIXED_BF16, MNP_ERR_ARG,
              "bemstatlayer_create: unknown solver");
    MNP_CHECK(gimage->g->n == 2 * g->g->n, MNP_ERR_ARG,
              "bemstatlayer_create: the image Green function must hold the particle and its image");
    MNP_CHECK(!gimage->g->op.closed, MNP_ERR_ARG,
              "bemstatlayer_create: the image Green function must be built without the closed-surface correction");
    auto* h = new mnp_bemstat_s{nullptr};
    try {
      h->b = new BemStat(g->g, gimage->g, solver);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int mnp_bemstatlayer_initmat(mnp_bemstat_t b, double enei, mnp_complex eps1, mnp_complex eps2, mnp_complex eps_up,
                             mnp_complex eps_down) {
  return guarded([&] {
    MNP_CHECK(b && enei > 0, MNP_ERR_ARG, "bemstatlayer_initmat: bad argument");
    MNP_CHECK(eps1.re != eps2.re || eps1.im != eps2.im, MNP_ERR_ARG, "bemstatlayer_initmat: eps1 == eps2");
    const double2 eu = to_d2(eps_up), ed = to_d2(eps_down);
    const double2 den = cadd<double>(ed, eu);
    MNP_CHECK(den.x != 0 || den.y != 0, MNP_ERR_ARG, "bemstatlayer_initmat: eps_up + eps_down == 0");
    // f2 = -(eps2 - eps1) / (eps2 + eps1) of the layer (compgreenstatlayer/eval.m:18)
    const double2 f2 = cscale<double>(cdiv<double>(csub<double>(ed, eu), den), -1.0);
    b->b->initmat_layer(to_d2(eps1), to_d2(eps2), f2);
  });
}

int mnp_planewavestatlayer_potential(mnp_greenret_t g, int64_t npol, const mnp_complex* e, void* phip) {
  return guarded([&] {
    MNP_CHECK(g && e && phip && npol >= 1 && npol <= MNP_STAT_MAXPOL, MNP_ERR_ARG,
              "planewavestatlayer_potential: bad argument");
    const GreenRet& G = *g->g;
    FieldArg a{};
    for (int64_t q = 0; q < npol; ++q)
      for (int c = 0; c < 3; ++c) a.e[q][c] = to_d2(e[q * 3 + c]);
    hipLaunchKernelGGL(k_pwstat_field, dim3((unsigned)cdiv_up(G.n, 256), (unsigned)npol), dim3(256), 0, stream(),
                       G.n, G.geo.nvec, a, (double2*)phip);
    MNP_HIP(hipGetLastError());
  });
}

int mnp_stat_dipole_moment(mnp_greenret_t g, int64_t npol, const void* sig, mnp_complex* dip) {
  return guarded([&] {
    MNP_CHECK(g && sig && dip && npol >= 1 && npol < 65536, MNP_ERR_ARG, "stat_dipole_moment: bad argument");
    const GreenRet& G = *g->g;
    const int nchunk = 64;
    double2* part = dev_alloc<double2>(npol * nchunk * 3);
    hipLaunchKernelGGL(k_dipmoment, dim3(nchunk, (unsigned)npol), dim3(256), 0, stream(), G.n, G.geo.pos,
                       G.geo.area, (const double2*)sig, part);
    MNP_HIP(hipGetLastError());
    std::vector<double2> h(npol * nchunk * 3);
    MNP_HIP(hipMemcpyAsync(h.data(), part, sizeof(double2) * h.size(), hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_HIP(hipFree(part));
    for (int64_t q = 0; q < npol; ++q)
      for (int d = 0; d < 3; ++d) {
        double re = 0, im = 0;
        for (int c = 0; c < nchunk; ++c) {
          re += h[(q * nchunk + c) * 3 + d].x;
          im += h[(q * nchunk + c) * 3 + d].y;
        }
        dip[q * 3 + d] = mnp_complex{re, im};
      }
  });
}

int mnp_bemstatmirror_create(mnp_greenret_t g, int solver, int nsym, const int* symtable, mnp_bemstatmirror_t* out) {
  return guarded([&] {
    MNP_CHECK(!g || !g->g->compound(), MNP_ERR_ARG, "bemstatmirror_create: compound particles take the dense solver (mnp_bemret_*)");
    MNP_CHECK(g && symtable && out, MNP_ERR_ARG, "bemstatmirror_create: NULL argument");
    MNP_CHECK(solver == MNP_SOLVER_FP64 || solver == MNP_SOLVER_MIXED || solver == MNP_SOLVER_MIXED_BF16,
              MNP_ERR_ARG, "bemstatmirror_create: unknown solver");
    MNP_CHECK(nsym == 2 || nsym == 4, MNP_ERR_ARG, "bemstatmirror_create: nsym must be 2 or 4");
    MNP_CHECK(g->g->n % nsym == 0, MNP_ERR_ARG, "bemstatmirror_create: faces not a multiple of nsym");
    MNP_CHECK(g->g->n <= 65535, MNP_ERR_ARG, "bemstatmirror_create: at most 65535 faces in the full particle");
    auto* h = new mnp_bemstatmirror_s{nullptr};
    try {
      h->b = new BemStatMirror(g->g, solver, nsym, symtable);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int mnp_bemstatmirror_destroy(mnp_bemstatmirror_t b) {
  return guarded([&] {
    if (!b) return;
    delete b->b;
    delete b;
  });
}

int mnp_bemstatmirror_initmat(mnp_bemstatmirror_t b, double enei, mnp_complex eps1, mnp_complex eps2) {
  return guarded([&] {
    MNP_CHECK(b && enei > 0, MNP_ERR_ARG, "bemstatmirror_initmat: bad argument");
    MNP_CHECK(eps1.re != eps2.re || eps1.im != eps2.im, MNP_ERR_ARG, "bemstatmirror_initmat: eps1 == eps2");
    b->b->initmat(to_d2(eps1), to_d2(eps2));
  });
}

int mnp_bemstatmirror_solve(mnp_bemstatmirror_t b, int sym, int64_t npol, const void* phip, void* sig) {
  return guarded([&] {
    MNP_CHECK(b && phip && sig, MNP_ERR_ARG, "bemstatmirror_solve: NULL argument");
    MNP_CHECK(sym >= 0 && sym < b->b->J, MNP_ERR_ARG, "bemstatmirror_solve: symmetry index out of range");
    b->b->s[sym]->solve(npol, (const double2*)phip, (double2*)sig);
  });
}

int mnp_bemstatmirror_stats(mnp_bemstatmirror_t b, int sym, int* sweeps, double* relres) {
  return guarded([&] {
    MNP_CHECK(b, MNP_ERR_ARG, "bemstatmirror_stats: NULL handle");
    MNP_CHECK(sym >= 0 && sym < b->b->J, MNP_ERR_ARG, "bemstatmirror_stats: symmetry index out of range");
    if (sweeps) *sweeps = b->b->s[sym]->last_sweeps;
    if (relres) *relres = b->b->s[sym]->last_relres;
  });
}

int mnp_bemstatmirror_set_refinement(mnp_bemstatmirror_t b, double tol, int maxit) {
  return guarded([&] {
    MNP_CHECK(b && tol > 0 && maxit >= 0, MNP_ERR_ARG, "bemstatmirror_set_refinement: bad argument");
    for (auto& s : b->b->s) {
      s->refine_tol = tol;
      s->refine_maxit = maxit;
    }
  });
}

int mnp_bemstat_destroy(mnp_bemstat_t b) {
  return guarded([&] {
    if (!b) return;
    delete b->b;
    delete b;
  });
}

int mnp_bemstat_initmat(mnp_bemstat_t b, double enei, mnp_complex eps1, mnp_complex eps2) {
  return guarded([&] {
    MNP_CHECK(b && enei > 0, MNP_ERR_ARG, "bemstat_initmat: bad argument");
    MNP_CHECK(eps1.re != eps2.re || eps1.im != eps2.im, MNP_ERR_ARG, "bemstat_initmat: eps1 == eps2");
    b->b->initmat(to_d2(eps1), to_d2(eps2));
  });
}

int mnp_bemstat_solve(mnp_bemstat_t b, int64_t npol, const void* phip, void* sig) {
  return guarded([&] {
    MNP_CHECK(b && phip && sig, MNP_ERR_ARG, "bemstat_solve: NULL argument");
    b->b->solve(npol, (const double2*)phip, (double2*)sig);
  });
}

int mnp_bemstat_stats(mnp_bemstat_t b, int* sweeps, double* relres) {
  return guarded([&] {
    MNP_CHECK(b, MNP_ERR_ARG, "bemstat_stats: NULL handle");
    if (sweeps) *sweeps = b->b->last_sweeps;
    if (relres) *relres = b->b->last_relres;
  });
}

int mnp_planewavestat_potential(mnp_greenret_t g, int64_t npol, const double* pol, void* phip) {
  return guarded([&] {
    MNP_CHECK(g && pol && phip && npol >= 1 && npol < 65536, MNP_ERR_ARG, "planewavestat_potential: bad argument");
    const GreenRet& G = *g->g;
    double* dpol = dev_alloc<double>(3 * npol);
    MNP_HIP(hipMemcpyAsync(dpol, pol, sizeof(double) * 3 * npol, hipMemcpyHostToDevice, stream()));
    hipLaunchKernelGGL(k_pwstat, dim3((unsigned)cdiv_up(G.n, 256), (unsigned)npol), dim3(256), 0, stream(), G.n,
                       npol, G.geo.nvec, dpol, (double2*)phip);
    MNP_HIP(hipGetLastError());
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_HIP(hipFree(dpol));
  });
}

int mnp_planewavestat_cross_sections(mnp_greenret_t g, double enei, mnp_complex eps_medium, int64_t npol,
                                     const double* pol, const void* sig, double* ext, double* sca,
                                     double* abs_) {
  return guarded([&] {
    MNP_CHECK(g && pol && sig && npol >= 1 && npol < 65536 && enei > 0, MNP_ERR_ARG,
              "planewavestat_cross_sections: bad argument");
    const GreenRet& G = *g->g;
    const int nchunk = 64;
    double2* part = dev_alloc<double2>(npol * nchunk * 3);
    hipLaunchKernelGGL(k_dipmoment, dim3(nchunk, (unsigned)npol), dim3(256), 0, stream(), G.n, G.geo.pos,
                       G.geo.area, (const double2*)sig, part);
    MNP_HIP(hipGetLastError());
    std::vector<double2> h(npol * nchunk * 3);
    MNP_HIP(hipMemcpyAsync(h.data(), part, sizeof(double2) * h.size(), hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_HIP(hipFree(part));
    using C = std::complex<double>;
    const C k = 2 * M_PI / enei * std::sqrt(C(eps_medium.re, eps_medium.im));
    for (int64_t q = 0; q < npol; ++q) {
      C dip[3] = {0, 0, 0};
      for (int c = 0; c < nchunk; ++c)
        for (int d = 0; d < 3; ++d) dip[d] += C(h[(q * nchunk + c) * 3 + d].x, h[(q * nchunk + c) * 3 + d].y);
      double a2 = 0;
      C pd = 0;
      for (int d = 0; d < 3; ++d) {
        a2 += std::norm(dip[d]);
        pd += pol[d * npol + q] * dip[d];
      }
      const C s = 8 * M_PI / 3 * k * k * k * k * a2;  // @planewavestat/scattering.m:14-18
      const C a = 4 * M_PI * k * pd.imag();          // absorption.m:14-19
      if (sca) sca[q] = s.real();
      if (abs_) abs_[q] = a.real();
      if (ext) ext[q] = (s + a).real();              // extinction.m
    }
  });
}

}  // extern "C"
