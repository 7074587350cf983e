// Dense retarded BEM solver with mirror symmetry (MNPBEM BEM/@bemretmirror)
// on gfx950.
//
// The particle is the n-face part p of a closed surface that J = 2 ('x',
// 'y') or 4 ('xy') mirror copies of p complete; the Green function is the
// full particle's, with copy j in faces j n .. (j + 1) n - 1 (the order of
// Particles/@comparticlemirror/private/init.m:42-54).  For symmetry value s
// (row s of the J x J table of +-1, Particles/@comparticlemirror/private/
// init.m:29-37) the n x n matrices are
//   G_s = sum_j tab[s][j] G(p, copy j)
// (Greenfun/@compgreenretmirror/eval.m:23-42): each element of the n x J n
// block is assembled once and contracted into all J of them.  The chain of
// BEM/@bemretmirror/private/initmat.m:25-66 runs per symmetry value (LU(G1),
// Sigma1, LU(G2), Sigma2, Deltai), and Sigma of initsigmai.m:19-30 per
// combination (x, y, z) of the symmetry values of the excitation's Cartesian
// components, cached until the next initmat.  mldivide.m:25-87 runs on these
// factors with each vector component through its own symmetry value.
//
// MNP_SOLVER_FP64 runs everything in FP64; MNP_SOLVER_MIXED (_BF16) runs the
// chains in complex FP32 and refines against the exact FP64 operator, as the
// solver without symmetry does.
#include <cmath>
#include <map>
#include <memory>
#include <vector>

#include "mnp_chain.hpp"
#include "mnp_mirror.hpp"

namespace mnp {

// Sigma = e1 S1_z - e2 S2_z + c sum_comp (Di_comp o n_comp n_comp^T)
// (initsigmai.m:19-30 with scalar L1 = eps1, L2 = eps2).
template <class T>
__global__ void k_sigma3(int64_t n, const cplx<T>* S1, const cplx<T>* S2, const cplx<T>* Dx, const cplx<T>* Dy,
                         const cplx<T>* Dz, const double* nvec, double2 e1, double2 e2, double2 c, cplx<T>* out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (i >= n || j >= n) return;
  const int64_t idx = j * n + i;
  const cplx<T> s1 = S1[idx], s2 = S2[idx], dx = Dx[idx], dy = Dy[idx], dz = Dz[idx];
  const double nx = nvec[i] * nvec[j], ny = nvec[n + i] * nvec[n + j], nz = nvec[2 * n + i] * nvec[2 * n + j];
  const double2 d = mk<double>(nx * dx.x + ny * dy.x + nz * dz.x, nx * dx.y + ny * dy.y + nz * dz.y);
  const double2 a = cmul<double>(e1, mk<double>(s1.x, s1.y));
  const double2 b = cmul<double>(e2, mk<double>(s2.x, s2.y));
  const double2 e = cmul<double>(c, d);
  out[idx] = mk<T>((T)(a.x - b.x + e.x), (T)(a.y - b.y + e.y));
}

// x = M^{-1} b with per-component symmetry values: cs[0..2] are the chains
// of the x, y, z components (the scalars go with z), sg holds LU(Sigma) of
// that combination.  b, x are n x 8p as in chain_apply; tmp n x 18p.
template <class T>
void mirror_apply(Chain<T>* const cs[3], Chain<T>& sg, const double* nvec, int64_t n, int64_t p, const cplx<T>* b,
                  cplx<T>* x, cplx<T>* tmp) {
  const double k = sg.k0;
  const double2 L1 = sg.eps1, L2 = sg.eps2;
  const double2 ikdL = cmul<double>(mk<double>(0, k), csub<double>(L1, L2));
  cplx<T>* Y = tmp;
  cplx<T>* alpha = Y + 4 * p * n;
  cplx<T>* De = alpha + 3 * p * n;
  cplx<T>* t = De + p * n;
  cplx<T>* s2 = t + 3 * p * n;
  cplx<T>* u = s2 + p * n;
  cplx<T>* h2 = u + 3 * p * n;
  const double2 one = {1, 0}, zero = {0, 0};
  // component c of an n x 3p vector block (columns c + 3 q)
  auto comp = [&](const cplx<T>* base, int c) { return MatView<T>{const_cast<cplx<T>*>(base) + c * n, 3 * n}; };
  Chain<T>& cz = *cs[2];
  for (int c = 0; c < 3; ++c)  // Y_a = Sigma1 a per component
    gemm<T>(n, p, n, one, cs[c]->v(cs[c]->S1), comp(b + n * p, c), zero, comp(Y, c));
  gemm<T>(n, p, n, one, cz.v(cz.S1), MatView<T>{const_cast<cplx<T>*>(b), n}, zero, MatView<T>{Y + 3 * p * n, n});
  const dim3 gp((unsigned)cdiv_up(n, 256), (unsigned)p);
  hipLaunchKernelGGL(k_mld1<T>, gp, dim3(256), 0, stream(), n, p, b, Y, alpha, De, nvec, k, L1);
  for (int c = 0; c < 3; ++c) gemm<T>(n, p, n, one, cs[c]->v(cs[c]->Di), comp(alpha, c), zero, comp(t, c));
  hipLaunchKernelGGL(k_mld2<T>, gp, dim3(256), 0, stream(), n, p, De, t, s2, nvec, ikdL);
  MNP_HIP(hipGetLastError());
  getrs<T>(n, p, sg.f[3], MatView<T>{s2, n}, sg.ws);  // sig2 = Sigma^{-1} rhs
  hipLaunchKernelGGL(k_mld3<T>, gp, dim3(256), 0, stream(), n, p, s2, alpha, u, nvec, ikdL);
  MNP_HIP(hipGetLastError());
  for (int c = 0; c < 3; ++c) gemm<T>(n, p, n, one, cs[c]->v(cs[c]->Di), comp(u, c), zero, comp(h2, c));
  hipLaunchKernelGGL(k_mld4<T>, dim3((unsigned)cdiv_up(n, 256), (unsigned)(4 * p)), dim3(256), 0, stream(), n, p,
                     b, s2, h2, x, x + 4 * p * n);
  MNP_HIP(hipGetLastError());
  // sig1 = G1_z^{-1} (sig2 + phi), h1 = G1^{-1} (h2 + a); sig2, h2 through G2
  getrs<T>(n, p, cz.f[0], MatView<T>{x, n}, cz.ws);
  for (int c = 0; c < 3; ++c) getrs<T>(n, p, cs[c]->f[0], comp(x + n * p, c), cs[c]->ws);
  getrs<T>(n, p, cz.f[1], MatView<T>{x + 4 * p * n, n}, cz.ws);
  for (int c = 0; c < 3; ++c) getrs<T>(n, p, cs[c]->f[1], comp(x + 5 * p * n, c), cs[c]->ws);
}

struct BemRetMirror {
  const GreenRet* g;  // full particle
  int solver, J;
  int64_t n;          // faces of the symmetry part
  SymTab tab{};
  double* nvec = nullptr;  // normals of the symmetry part, n x 3
  std::vector<std::unique_ptr<Chain<double>>> s64;
  std::vector<std::unique_ptr<Chain<float>>> s32;
  std::map<int, std::unique_ptr<Chain<double>>> sig64;  // Sigma per (x, y, z) combination
  std::map<int, std::unique_ptr<Chain<float>>> sig32;
  std::vector<double2*> op;  // mixed: FP64 G1, G2, H1, H2 per symmetry value
  double2 *BG = nullptr, *BH = nullptr;  // n x J n assembly blocks
  bool ready = false;
  double enei = 0, k0 = 0;
  double2 eps1{}, eps2{};
  double tol = 1e-12;
  int maxit = 10;
  int last_iter = 0;
  double last_relres = 0;
  Workspace vws, res_ws;

  BemRetMirror(const GreenRet* g_, int solver_, int J_, const int* table)
      : g(g_), solver(solver_), J(J_), n(g_->n / J_) {
    check_symtable("bemretmirror_create", J, table, tab);
    nvec = dev_alloc<double>(3 * n);
    for (int c = 0; c < 3; ++c)
      MNP_HIP(hipMemcpy(nvec + c * n, g->geo.nvec + c * g->n, sizeof(double) * n, hipMemcpyDeviceToDevice));
    BG = dev_alloc<double2>(n * g->n);
    BH = dev_alloc<double2>(n * g->n);
    for (int s = 0; s < J; ++s) {
      if (solver == MNP_SOLVER_FP64) {
        s64.emplace_back(new Chain<double>(n, true));
      } else {
        s32.emplace_back(new Chain<float>(n, false));
        for (int w = 0; w < 4; ++w) op.push_back(dev_alloc<double2>(n * n));
      }
    }
  }
  ~BemRetMirror() {
    for (double2* p : op) (void)hipFree(p);
    void* ps[] = {nvec, BG, BH};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
  double2* opm(int s, int w) const { return op[4 * s + w]; }  // w: 0 G1, 1 G2, 2 H1, 3 H2

  void contract(const double2* B, const OutPair* outs) { sym_contract(J, n, tab, B, outs); }

  void initmat(double enei_, double2 e1, double2 e2) {
    ready = false;  // until every chain is complete
    enei = enei_;
    eps1 = e1;
    eps2 = e2;
    k0 = 2 * M_PI / enei;
    sig64.clear();
    sig32.clear();
    const double2 kk[2] = {wavenumber(k0, e1), wavenumber(k0, e2)};
    for (int side = 0; side < 2; ++side) {
      g->assemble_rows(kk[side], side == 0 ? 2 * M_PI : -2 * M_PI, OutPair{BG, nullptr}, OutPair{BH, nullptr}, n, n);
      OutPair og[4], oh[4];
      for (int s = 0; s < J; ++s) {
        if (solver == MNP_SOLVER_FP64) {
          Chain<double>& c = *s64[s];
          og[s] = OutPair{side == 0 ? c.G1 : c.G2, nullptr};
          oh[s] = OutPair{side == 0 ? c.S1 : c.S2, nullptr};
        } else {
          Chain<float>& c = *s32[s];
          og[s] = OutPair{opm(s, side), side == 0 ? c.G1 : c.G2};
          oh[s] = OutPair{opm(s, 2 + side), side == 0 ? c.S1 : c.S2};
        }
      }
      contract(BG, og);
      contract(BH, oh);
    }
    // the J symmetry chains are independent: one task (and stream) each;
    // their large products are small enough (n / J) that overlapping the
    // chains' latency-bound kernels pays
    SplitBf16Scope split(solver == MNP_SOLVER_MIXED_BF16);
    // (factorisations that need partial pivoting are finished afterwards, one
    // chain at a time: the pivoted panel is a cooperative launch)
    std::vector<std::function<void()>> tasks;
    for (int s = 0; s < J; ++s) {
      if (solver == MNP_SOLVER_FP64) {
        Chain<double>& c = *s64[s];
        c.eps1 = e1, c.eps2 = e2, c.k0 = k0;
        tasks.push_back([&c] { chain_sector_try<double>(c); });
      } else {
        Chain<float>& c = *s32[s];
        c.eps1 = e1, c.eps2 = e2, c.k0 = k0;
        tasks.push_back([&c] { chain_sector_try<float>(c); });
      }
    }
    run_concurrent(tasks);
    for (int s = 0; s < J; ++s) {
      if (solver == MNP_SOLVER_FP64)
        chain_sector_finish<double>(*s64[s]);
      else
        chain_sector_finish<float>(*s32[s]);
    }
    ready = true;
  }

  template <class T>
  Chain<T>& sigma(std::vector<std::unique_ptr<Chain<T>>>& sec, std::map<int, std::unique_ptr<Chain<T>>>& cache,
                  const int sx[3]) {
    const int key = (sx[0] * 4 + sx[1]) * 4 + sx[2];
    auto it = cache.find(key);
    if (it != cache.end()) return *it->second;
    std::unique_ptr<Chain<T>> c(new Chain<T>(n, solver == MNP_SOLVER_FP64, true));
    c->eps1 = eps1, c->eps2 = eps2, c->k0 = k0;
    const double2 L = csub<double>(eps1, eps2);
    const double2 cc = cscale<double>(cmul<double>(L, L), k0 * k0);
    Chain<T>& z = *sec[sx[2]];
    hipLaunchKernelGGL(k_sigma3<T>, gridnn(n, n), dim3(256), 0, stream(), n, z.S1, z.S2, sec[sx[0]]->Di,
                       sec[sx[1]]->Di, z.Di, nvec, eps1, eps2, cc, c->S2);
    MNP_HIP(hipGetLastError());
    SplitBf16Scope split(solver == MNP_SOLVER_MIXED_BF16);
    c->factor(3, c->S2, "Sigma");
    return *(cache[key] = std::move(c));
  }

  // Y = A [s | h] for the operator matrix w of the symmetry values sx
  // (scalar column block through z, component c through sx[c]); X, Y n x 4p.
  void opmul(int w, const int sx[3], int64_t p, const double2* X, double2* Y) {
    const double2 one = {1, 0}, zero = {0, 0};
    gemm<double>(n, p, n, one, MatView<double>{opm(sx[2], w), n}, MatView<double>{const_cast<double2*>(X), n}, zero,
                 MatView<double>{Y, n});
    for (int c = 0; c < 3; ++c)
      gemm<double>(n, p, n, one, MatView<double>{opm(sx[c], w), n},
                   MatView<double>{const_cast<double2*>(X) + (p + c) * n, 3 * n}, zero,
                   MatView<double>{Y + (p + c) * n, 3 * n});
  }

  void solve(const int sx[3], int64_t p, const double2* const in[8], double2* sig1, double2* sig2, double2* h1,
             double2* h2) {
    MNP_CHECK(ready, MNP_ERR_STATE, "bemretmirror_solve: initmat not called");
    MNP_CHECK(p > 0 && p <= 64, MNP_ERR_ARG, "bemretmirror_solve: npol must be in 1..64");
    for (int c = 0; c < 3; ++c)
      MNP_CHECK(sx[c] >= 0 && sx[c] < J, MNP_ERR_ARG, "bemretmirror_solve: symmetry index out of range");
    const int64_t len = 8 * p * n;
    double2* b = (double2*)vws.get(0, sizeof(double2) * len);
    double2* x = (double2*)vws.get(1, sizeof(double2) * len);
    launch_excitation(n, p, in, b, nvec, k0, eps1, eps2);
    const size_t tmp_elems = (size_t)(18 * p * n);
    last_iter = 0;
    last_relres = 0;
    if (solver == MNP_SOLVER_FP64) {
      Chain<double>& sg = sigma<double>(s64, sig64, sx);
      Chain<double>* cs[3] = {s64[sx[0]].get(), s64[sx[1]].get(), s64[sx[2]].get()};
      double2* tmp = (double2*)vws.get(2, sizeof(double2) * tmp_elems);
      mirror_apply<double>(cs, sg, nvec, n, p, b, x, tmp);
    } else {
      Chain<float>& sg = sigma<float>(s32, sig32, sx);
      Chain<float>* cs[3] = {s32[sx[0]].get(), s32[sx[1]].get(), s32[sx[2]].get()};
      float2* b32 = (float2*)vws.get(2, sizeof(float2) * len);
      float2* x32 = (float2*)vws.get(3, sizeof(float2) * len);
      float2* tmp = (float2*)vws.get(4, sizeof(float2) * tmp_elems);
      double2* work = (double2*)res_ws.get(0, sizeof(double2) * (size_t)(24 * p * n));
      double2 *GX1 = work, *GX2 = work + 4 * p * n, *HX1 = work + 8 * p * n, *HX2 = work + 12 * p * n,
              *rr = work + 16 * p * n;
      auto precond = [&](const double2* src, double2* dst) {
        convert_mat<double, float>(n, 8 * p, MatView<double>{const_cast<double2*>(src), n}, MatView<float>{b32, n});
        mirror_apply<float>(cs, sg, nvec, n, p, b32, x32, tmp);
        convert_mat<float, double>(n, 8 * p, MatView<float>{x32, n}, MatView<double>{dst, n});
      };
      {
        SplitBf16Scope split(solver == MNP_SOLVER_MIXED_BF16);
        precond(b, x);
      }
      const double bn = dev_norm2(len, b, vws, 5);
      for (int it = 0;; ++it) {
        opmul(0, sx, p, x, GX1);
        opmul(1, sx, p, x + 4 * p * n, GX2);
        opmul(2, sx, p, x, HX1);
        opmul(3, sx, p, x + 4 * p * n, HX2);
        launch_residual(n, p, b, GX1, GX2, HX1, HX2, rr, nvec, k0, eps1, eps2);
        const double rn = dev_norm2(len, rr, vws, 5);
        last_relres = bn > 0 ? rn / bn : 0;
        last_iter = it;
        if (!(last_relres > tol) || it >= maxit) break;
        {
          SplitBf16Scope split(solver == MNP_SOLVER_MIXED_BF16);
          precond(rr, GX1);
        }
        add_into(len, x, GX1);
      }
      MNP_CHECK(!(last_relres > tol), MNP_ERR_NOCONV,
                "bemretmirror_solve: refinement stalled at relative residual " + sci(last_relres));
    }
    launch_unpack(n, p, x, sig1, h1, sig2, h2);
  }

  static void add_into(int64_t len, double2* x, const double2* d);
};

__global__ void k_add_into(int64_t len, double2* x, const double2* d) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) x[i] = cadd<double>(x[i], d[i]);
}
void BemRetMirror::add_into(int64_t len, double2* x, const double2* d) {
  hipLaunchKernelGGL(k_add_into, dim3((unsigned)cdiv_up(len, 256)), dim3(256), 0, stream(), len, x, d);
  MNP_HIP(hipGetLastError());
}

// Values on the full particle from the solutions of two excitations
// (planewaveretmirror/full.m:22-64 with compstructmirror/expand.m:22-46):
//   out[j n + i, c, o] = w[o][0] s1[j][c] in1[i, c, src[o]] + w[o][1] s2[j][c] in2[i, c, src[o]].
constexpr int MNP_MIRROR_MAXOUT = 64;
struct FullArgs {
  double w[MNP_MIRROR_MAXOUT][2];
  int src[MNP_MIRROR_MAXOUT];
  int8_t s[2][4][3];
};
__global__ void k_mirror_full(int64_t n, int nsym, int ncomp, FullArgs a, const double2* in1, const double2* in2,
                              double2* out) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t col = blockIdx.y;  // c + ncomp o
  const int64_t N = nsym * n;
  if (r >= N) return;
  const int j = (int)(r / n), c = (int)(col % ncomp), o = (int)(col / ncomp);
  const int64_t i = r % n;
  const int64_t ic = (c + ncomp * (int64_t)a.src[o]) * n + i;
  const double w0 = a.w[o][0] * a.s[0][j][c];
  double2 v = mk<double>(w0 * in1[ic].x, w0 * in1[ic].y);
  if (in2) {
    const double w1 = a.w[o][1] * a.s[1][j][c];
    v.x += w1 * in2[ic].x;
    v.y += w1 * in2[ic].y;
  }
  out[col * N + r] = v;
}

}  // namespace mnp

// ============================================================== C ABI
using namespace mnp;

struct mnp_bemretmirror_s {
  BemRetMirror* b;
};

extern "C" {

int mnp_bemretmirror_create(mnp_greenret_t g, int solver, int nsym, const int* symtable, mnp_bemretmirror_t* out) {
  return guarded([&] {
    MNP_CHECK(!g || !g->g->compound(), MNP_ERR_ARG, "bemretmirror_create: compound particles take the dense solver (mnp_bemret_*)");
    MNP_CHECK(g && symtable && out, MNP_ERR_ARG, "bemretmirror_create: NULL argument");
    MNP_CHECK(solver == MNP_SOLVER_FP64 || solver == MNP_SOLVER_MIXED || solver == MNP_SOLVER_MIXED_BF16,
              MNP_ERR_ARG, "bemretmirror_create: unknown solver");
    MNP_CHECK(nsym == 2 || nsym == 4, MNP_ERR_ARG, "bemretmirror_create: nsym must be 2 or 4");
    MNP_CHECK(g->g->n % nsym == 0, MNP_ERR_ARG, "bemretmirror_create: faces not a multiple of nsym");
    MNP_CHECK(g->g->n <= 65535, MNP_ERR_ARG, "bemretmirror_create: at most 65535 faces in the full particle");
    auto* h = new mnp_bemretmirror_s{nullptr};
    try {
      h->b = new BemRetMirror(g->g, solver, nsym, symtable);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int mnp_bemretmirror_destroy(mnp_bemretmirror_t bem) {
  return guarded([&] {
    if (!bem) return;
    delete bem->b;
    delete bem;
  });
}

int mnp_bemretmirror_initmat(mnp_bemretmirror_t bem, double enei, mnp_complex eps1, mnp_complex eps2) {
  return guarded([&] {
    MNP_CHECK(bem, MNP_ERR_ARG, "bemretmirror_initmat: NULL handle");
    MNP_CHECK(enei > 0, MNP_ERR_ARG, "bemretmirror_initmat: wavelength must be positive");
    bem->b->initmat(enei, to_d2(eps1), to_d2(eps2));
  });
}

int mnp_bemretmirror_solve(mnp_bemretmirror_t bem, int symx, int symy, int symz, int64_t npol, const void* phi1,
                           const void* phi1p, const void* a1, const void* a1p, const void* phi2, const void* phi2p,
                           const void* a2, const void* a2p, void* sig1, void* sig2, void* h1, void* h2) {
  return guarded([&] {
    MNP_CHECK(bem, MNP_ERR_ARG, "bemretmirror_solve: NULL handle");
    const double2* in[8] = {(const double2*)phi1, (const double2*)phi1p, (const double2*)a1, (const double2*)a1p,
                            (const double2*)phi2, (const double2*)phi2p, (const double2*)a2, (const double2*)a2p};
    const int sx[3] = {symx, symy, symz};
    bem->b->solve(sx, npol, in, (double2*)sig1, (double2*)sig2, (double2*)h1, (double2*)h2);
  });
}

int mnp_bemretmirror_stats(mnp_bemretmirror_t bem, int* iterations, double* relres) {
  return guarded([&] {
    MNP_CHECK(bem, MNP_ERR_ARG, "bemretmirror_stats: NULL handle");
    if (iterations) *iterations = bem->b->last_iter;
    if (relres) *relres = bem->b->last_relres;
  });
}

int mnp_bemretmirror_set_refinement(mnp_bemretmirror_t bem, double tol, int maxit) {
  return guarded([&] {
    MNP_CHECK(bem && tol > 0 && maxit >= 0, MNP_ERR_ARG, "bemretmirror_set_refinement: bad argument");
    bem->b->tol = tol;
    bem->b->maxit = maxit;
  });
}

int mnp_mirror_full(int64_t n, int nsym, int ncomp, int64_t nsrc, int64_t nout, const int* sign1, const int* sign2,
                    const double* w, const int* src, const void* in1, const void* in2, void* out) {
  return guarded([&] {
    MNP_CHECK(sign1 && w && src && in1 && out && (sign2 || !in2), MNP_ERR_ARG, "mirror_full: NULL argument");
    MNP_CHECK(n >= 1 && (nsym == 2 || nsym == 4) && (ncomp == 1 || ncomp == 3) && nsrc >= 1 && nout >= 1 &&
                  nout <= MNP_MIRROR_MAXOUT,
              MNP_ERR_ARG, "mirror_full: bad size");
    FullArgs a{};
    for (int j = 0; j < nsym; ++j)
      for (int c = 0; c < ncomp; ++c) {
        a.s[0][j][c] = (int8_t)sign1[j * ncomp + c];
        a.s[1][j][c] = (int8_t)(sign2 ? sign2[j * ncomp + c] : 0);
      }
    for (int64_t o = 0; o < nout; ++o) {
      MNP_CHECK(src[o] >= 0 && src[o] < nsrc, MNP_ERR_ARG, "mirror_full: source column out of range");
      a.src[o] = src[o];
      a.w[o][0] = w[2 * o];
      a.w[o][1] = w[2 * o + 1];
    }
    hipLaunchKernelGGL(k_mirror_full, dim3((unsigned)cdiv_up(nsym * n, 256), (unsigned)(ncomp * nout)), dim3(256), 0,
                       stream(), n, nsym, ncomp, a, (const double2*)in1, (const double2*)in2, (double2*)out);
    MNP_HIP(hipGetLastError());
  });
}

}  // extern "C"
