// Flexible GMRES for the iterative BEM solvers (bemretiter, bemstatiter).
#pragma once
#include <functional>

#include "mnp_linalg.hpp"

namespace mnp {

struct GmresOut {
  int iter = 0;             // iterations taken
  int flag = 0;             // 1 when relres > tol at the end
  double relres = 0;        // GMRES's residual estimate / |b|
  double true_relres = 0;   // |b - A x| / |b|, recomputed at the end
};

// Solves A x = b (vectors of len complex FP64 entries) by GMRES without
// restart, right-preconditioned in the flexible form (z_j = M^{-1} v_j kept,
// x = Z y): the residual it minimises is the true |b - A x|, also when M^{-1}
// is applied inexactly, and convergence is |b - A x| <= tol |b| as Matlab's
// gmres reports relres (BEM/@bemiter/solve.m:14-45 calls gmres with
// restart = [], tol, maxit, mfun).  afun(v, w): w = A v; mfun(v, z): z =
// M^{-1} v.  Uses slots 1, 2, 3, 4, 5 and 7 of ws.
GmresOut fgmres(int64_t len, const double2* b, double2* x, double tol, int maxit,
                const std::function<void(const double2*, double2*)>& afun,
                const std::function<void(const double2*, double2*)>& mfun, Workspace& ws);

// v <- a v (len complex FP64 entries), on the library's stream
void scale_vec(int64_t len, double2* v, double a);

}  // namespace mnp
