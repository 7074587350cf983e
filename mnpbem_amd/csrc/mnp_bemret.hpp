// The dense retarded solver's state (BEM/@bemret), shared by bemret.hip
// (initmat, solve, tangent) and adjoint.hip (transposed solve, extinction
// gradient).  Member functions are defined out of line in those files.
#pragma once

#include <memory>
#include <vector>

#include "mnp_chain.hpp"
#include "mnp_linalg.hpp"

namespace mnp {

// x = [sig1 | h1 | sig2 | h2] (n x 8p) from the compstruct fields.
void launch_pack(int64_t n, int64_t p, const double2* sig1, const double2* h1, const double2* sig2,
                 const double2* h2, double2* x);
// r = -(dA/deps1) x in the layout of b, from GX = G1 [sig1 | h1], HX = H1 [sig1 | h1]
// and dGX, dHX their derivatives along eps1.
void launch_deps1(int64_t n, int64_t p, const double2* GX, const double2* HX, const double2* dGX,
                  const double2* dHX, double2* r, const double* nvec, double k0, double2 e1);

struct Adjoint;                  // transposed chain and gradient scratch (adjoint.hip)
void adjoint_free(Adjoint* a);   // releases it

struct BemRet {
  const GreenRet* g;
  int solver;
  int64_t n;
  bool ready = false;
  double enei = 0;
  double2 eps1{}, eps2{};
  std::unique_ptr<Chain<double>> c64;
  std::unique_ptr<Chain<float>> c32;
  double2 *G1 = nullptr, *G2 = nullptr, *H1 = nullptr, *H2 = nullptr;  // FP64 operator (mixed)
  double2 *e1v = nullptr, *e2v = nullptr;  // eps inside / outside per face (compound)
  Workspace vws;  // vectors
  Workspace tws;  // tangent solve
  Workspace res_ws;
  double tol = 1e-12;
  int maxit = 10;
  int last_iter = 0;
  double last_relres = 0;
  Adjoint* adj = nullptr;  // built on the first transposed solve after each initmat
  uint64_t epoch = 0;      // counts initmat calls (the transposed chain's validity)

  BemRet(const GreenRet* g_, int solver_);
  ~BemRet();
  BemRet(const BemRet&) = delete;
  BemRet& operator=(const BemRet&) = delete;

  void initmat(double enei_, double2 e1, double2 e2);
  // Assembly and the factorisations that need no partial pivoting (safe on a
  // worker stream); false when initmat_finish must complete the chain.
  bool initmat_try(double enei_, double2 e1, double2 e2);
  void initmat_finish();
  // compound particle: epsm[m - 1] the dielectric function of medium m
  void initmat_media(double enei_, const std::vector<double2>& epsm);
  double norm2(const double2* v, int64_t len);
  void solve(int64_t p, const double2* const in[8], double2* sig1, double2* sig2, double2* h1, double2* h2);
  // x = A^{-1} b for b, x n x 8p (layouts of chain_apply); b is not modified
  void solve_b(int64_t p, const double2* b, double2* x);
  void tangent_eps1(int64_t p, const double2* sig1, const double2* sig2, const double2* h1, const double2* h2,
                    double2* dsig1, double2* dsig2, double2* dh1, double2* dh2);

  // adjoint.hip: lam = A^{-T} c (plain transpose) for c in the layout of x
  // and lam in the layout of b, n x 8p each; the same refinement as solve_b.
  void solve_transposed(int64_t p, const double2* c, double2* lam);
};

}  // namespace mnp

// Opaque handle behind mnp_bemret_t.
struct mnp_bemret_s {
  mnp::BemRet* b;
};
