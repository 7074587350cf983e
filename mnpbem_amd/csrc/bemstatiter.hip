// Iterative quasistatic BEM solver (MNPBEM BEM/@bemstatiter) on gfx950.
//
// sig = -(Lambda + F)^{-1} phip with F, the surface derivative of the static
// Green function, held as an H-matrix (the reference fills it with
// aca.compgreenstat and evaluates 'F', private/init.m:11-13; here the
// retarded fill at k = 0, whose kernels reduce to the static ones).  The
// reference solves afun(x) = phip with afun(x) = -(F x + Lambda x)
// (private/afun.m:7, solve.m:17-20) by GMRES; here the same operator goes
// through the flexible GMRES of the retarded iterative solver, preconditioned
// by the exact inverses of -(Lambda + F) on diagonal blocks of at most 32
// consecutive cluster indices of each leaf (the reference uses an H-matrix LU
// or a dense inverse of -Lambda - F, private/initmat.m:22-30).
#include <cmath>
#include <memory>
#include <vector>

#include "mnp_gmres.hpp"
#include "mnp_hmatrix.hpp"

namespace mnp {

constexpr int SPB = 32;  // preconditioner block size

// Inverse of -(lam I + F_bb) for each block (descriptor pb[5 b ..] = {first
// cluster index, size nb, offset in its leaf, leaf size m, value offset of the
// leaf's diagonal full block}) by Gauss-Jordan elimination with partial
// pivoting in LDS, in FP64; out[b] is SPB x SPB column-major.
template <class TS>
__global__ void __launch_bounds__(256) k_statprec_build(const int64_t* pb, const cplx<TS>* Fv, double2 lam,
                                                        double2* out) {
  __shared__ double2 W[SPB][2 * SPB + 1];
  __shared__ double2 colk[SPB];
  __shared__ int piv;
  const int64_t b = blockIdx.x;
  const int nb = (int)pb[5 * b + 1], off = (int)pb[5 * b + 2], m = (int)pb[5 * b + 3];
  const int64_t vo = pb[5 * b + 4];
  const int tid = threadIdx.x;
  for (int e = tid; e < SPB * 2 * SPB; e += 256) {
    const int i = e % SPB, j = e / SPB;
    double2 v = mk<double>(0, 0);
    if (j < SPB) {
      if (i < nb && j < nb) {
        const cplx<TS> f = Fv[vo + (int64_t)(off + j) * m + (off + i)];
        v = mk<double>(-(double)f.x, -(double)f.y);
        if (i == j) v = csub<double>(v, lam);
      } else if (i == j) {
        v = mk<double>(1, 0);  // padding: identity
      }
    } else {
      v = mk<double>(i == j - SPB ? 1.0 : 0.0, 0);
    }
    W[i][j] = v;
  }
  __syncthreads();
  for (int k = 0; k < SPB; ++k) {
    if (tid == 0) {
      int p = k;
      double best = cabs1<double>(W[k][k]);
      for (int i = k + 1; i < SPB; ++i) {
        const double a = cabs1<double>(W[i][k]);
        if (a > best) {
          best = a;
          p = i;
        }
      }
      piv = p;
    }
    __syncthreads();
    const int p = piv;
    if (p != k)
      for (int j = tid; j < 2 * SPB; j += 256) {
        const double2 t = W[k][j];
        W[k][j] = W[p][j];
        W[p][j] = t;
      }
    __syncthreads();
    const double2 inv = cdiv<double>(mk<double>(1, 0), W[k][k]);
    __syncthreads();
    for (int j = tid; j < 2 * SPB; j += 256) W[k][j] = cmul<double>(W[k][j], inv);
    if (tid < SPB) colk[tid] = W[tid][k];  // multipliers, saved before column k changes
    __syncthreads();
    for (int e = tid; e < SPB * 2 * SPB; e += 256) {
      const int i = e % SPB, j = e / SPB;
      if (i != k) W[i][j] = csub<double>(W[i][j], cmul<double>(colk[i], W[k][j]));
    }
    __syncthreads();
  }
  for (int e = tid; e < SPB * SPB; e += 256) {
    const int i = e % SPB, j = e / SPB;
    out[b * SPB * SPB + e] = W[i][SPB + j];
  }
}

// z = M^{-1} v on every block (v, z face order, n x p column-major)
__global__ void __launch_bounds__(256) k_statprec_apply(const int64_t* pb, const int32_t* perm, const double2* inv,
                                                        int64_t n, int p, const double2* v, double2* z) {
  __shared__ double2 vb[SPB][MNP_HMV_MAXRHS];
  const int64_t b = blockIdx.x;
  const int64_t r0 = pb[5 * b];
  const int nb = (int)pb[5 * b + 1];
  const int tid = threadIdx.x;
  for (int e = tid; e < nb * p; e += 256) {
    const int j = e % nb, q = e / nb;
    vb[j][q] = v[(int64_t)q * n + perm[r0 + j]];
  }
  __syncthreads();
  for (int e = tid; e < nb * p; e += 256) {
    const int i = e % nb, q = e / nb;
    const double2* Mi = inv + b * SPB * SPB + i;
    double2 s = mk<double>(0, 0);
    for (int j = 0; j < nb; ++j) s = cadd<double>(s, cmul<double>(Mi[(int64_t)j * SPB], vb[j][q]));
    z[(int64_t)q * n + perm[r0 + i]] = s;
  }
}

// y <- -(y + lam x)
__global__ void k_stat_afun(int64_t len, const double2* x, double2* y, double2 lam) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) {
    const double2 v = cadd<double>(y[i], cmul<double>(lam, x[i]));
    y[i] = mk<double>(-v.x, -v.y);
  }
}

struct BemStatIter {
  const GreenRet* g;
  int64_t n;
  std::unique_ptr<HMat> F;
  double2* prec = nullptr;  // npb x SPB x SPB
  int64_t* d_pb = nullptr;  // npb x 5 block descriptors
  int64_t npb = 0;
  bool use_prec = true;
  double tol = 1e-6;
  int maxit = 100;
  double2 lam{};
  bool filled = false, ready = false;
  GmresOut last;
  Workspace ws, hws;

  BemStatIter(const GreenRet* g_, int prec_, const int64_t* perm, int64_t nfull, const int64_t* full, int64_t nlow,
              const int64_t* low, double htol, int kmax)
      : g(g_), n(g_->n) {
    F.reset(new HMat(g, prec_, perm, nfull, full, nlow, low, htol, kmax));
    // preconditioner blocks: each leaf's diagonal block cut into near-equal
    // runs of at most SPB consecutive cluster indices
    std::vector<int64_t> pb, voff(F->nfull + 1, 0);
    for (int64_t b = 0; b < F->nfull; ++b)
      voff[b + 1] = voff[b] + (int64_t)(F->fblk_h[4 * b + 1] - F->fblk_h[4 * b]) *
                                  (F->fblk_h[4 * b + 3] - F->fblk_h[4 * b + 2]);
    for (int64_t l = 0; l < F->nleaf; ++l) {
      MNP_CHECK(F->diag_h[l] >= 0, MNP_ERR_ARG, "bemstatiter: a leaf cluster has no diagonal full block");
      const int64_t a = F->lrow_h[2 * l], m = F->lrow_h[2 * l + 1] - a;
      const int64_t nch = (m + SPB - 1) / SPB;
      for (int64_t c = 0; c < nch; ++c) {
        const int64_t o0 = c * m / nch, o1 = (c + 1) * m / nch;
        pb.insert(pb.end(), {a + o0, o1 - o0, o0, m, voff[F->diag_h[l]]});
      }
    }
    npb = (int64_t)pb.size() / 5;
    d_pb = dev_alloc<int64_t>((int64_t)pb.size());
    MNP_HIP(hipMemcpy(d_pb, pb.data(), sizeof(int64_t) * pb.size(), hipMemcpyHostToDevice));
    prec = dev_alloc<double2>(npb * SPB * SPB);
  }
  ~BemStatIter() {
    if (prec) (void)hipFree(prec);
    if (d_pb) (void)hipFree(d_pb);
  }

  // F does not depend on the wavelength: filled once; Lambda per wavelength
  void initmat(double2 e1, double2 e2) {
    ready = false;  // until the preconditioner is built
    if (!filled) {
      F->fill(mk<double>(0, 0), MNP_KEY_F);
      filled = true;
    }
    lam = cscale<double>(cdiv<double>(cadd<double>(e1, e2), csub<double>(e1, e2)), 2 * M_PI);
    if (F->prec == MNP_FP64)
      hipLaunchKernelGGL(k_statprec_build<double>, dim3((unsigned)npb), dim3(256), 0, stream(), d_pb,
                         (const double2*)F->fvals, lam, prec);
    else
      hipLaunchKernelGGL(k_statprec_build<float>, dim3((unsigned)npb), dim3(256), 0, stream(), d_pb,
                         (const float2*)F->fvals, lam, prec);
    MNP_HIP(hipGetLastError());
    ready = true;
  }

  void solve(int64_t p, const double2* phip, double2* sig) {
    MNP_CHECK(ready, MNP_ERR_STATE, "bemstatiter_solve: initmat not called");
    MNP_CHECK(p >= 1 && p <= MNP_HMV_MAXRHS, MNP_ERR_ARG, "bemstatiter_solve: npol must be in 1..8");
    const int64_t len = n * p;
    const unsigned gl = (unsigned)cdiv_up(len, 256);
    auto afun = [&](const double2* x, double2* y) {
      F->matvec(p, x, y, hws);
      hipLaunchKernelGGL(k_stat_afun, dim3(gl), dim3(256), 0, stream(), len, x, y, lam);
      MNP_HIP(hipGetLastError());
    };
    auto mfun = [&](const double2* v, double2* z) {
      if (!use_prec) {
        MNP_HIP(hipMemcpyAsync(z, v, sizeof(double2) * len, hipMemcpyDeviceToDevice, stream()));
        return;
      }
      hipLaunchKernelGGL(k_statprec_apply, dim3((unsigned)npb), dim3(256), 0, stream(), d_pb, F->perm, prec, n,
                         (int)p, v, z);
      MNP_HIP(hipGetLastError());
    };
    last = fgmres(len, phip, sig, tol, maxit, afun, mfun, ws);
  }
};

}  // namespace mnp

// ============================================================== C ABI
using namespace mnp;

struct mnp_bemstatiter_s {
  BemStatIter* b;
};

extern "C" {

int mnp_bemstatiter_create(mnp_greenret_t g, int precision, const int64_t* perm, int64_t nfull,
                           const int64_t* full, int64_t nlow, const int64_t* low, double htol, int kmax,
                           mnp_bemstatiter_t* out) {
  return guarded([&] {
    MNP_CHECK(!g || !g->g->compound(), MNP_ERR_ARG, "bemstatiter_create: compound particles take the dense solver (mnp_bemret_*)");
    MNP_CHECK(g && perm && out, MNP_ERR_ARG, "bemstatiter_create: NULL argument");
    auto* h = new mnp_bemstatiter_s{nullptr};
    try {
      h->b = new BemStatIter(g->g, precision, perm, nfull, full, nlow, low, htol, kmax);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int mnp_bemstatiter_destroy(mnp_bemstatiter_t it) {
  return guarded([&] {
    if (!it) return;
    delete it->b;
    delete it;
  });
}

int mnp_bemstatiter_set_options(mnp_bemstatiter_t it, double tol, int maxit, int precond) {
  return guarded([&] {
    MNP_CHECK(it && tol > 0 && maxit >= 1, MNP_ERR_ARG, "bemstatiter_set_options: bad argument");
    it->b->tol = tol;
    it->b->maxit = maxit;
    it->b->use_prec = precond != 0;
  });
}

int mnp_bemstatiter_initmat(mnp_bemstatiter_t it, double enei, mnp_complex eps1, mnp_complex eps2) {
  return guarded([&] {
    MNP_CHECK(it && enei > 0, MNP_ERR_ARG, "bemstatiter_initmat: bad argument");
    it->b->initmat(to_d2(eps1), to_d2(eps2));
  });
}

int mnp_bemstatiter_solve(mnp_bemstatiter_t it, int64_t npol, const void* phip, void* sig) {
  return guarded([&] {
    MNP_CHECK(it && phip && sig, MNP_ERR_ARG, "bemstatiter_solve: NULL argument");
    it->b->solve(npol, (const double2*)phip, (double2*)sig);
  });
}

int mnp_bemstatiter_stats(mnp_bemstatiter_t it, int* flag, double* relres, int* iter, double* true_relres,
                          double* compression) {
  return guarded([&] {
    MNP_CHECK(it, MNP_ERR_ARG, "bemstatiter_stats: NULL handle");
    const BemStatIter& B = *it->b;
    if (flag) *flag = B.last.flag;
    if (relres) *relres = B.last.relres;
    if (iter) *iter = B.last.iter;
    if (true_relres) *true_relres = B.last.true_relres;
    if (compression)
      *compression = B.F->filled ? (double)(B.F->nfullvals + B.F->nlowvals) / ((double)B.n * B.n) : 0.0;
  });
}

}  // extern "C"
