// Flexible GMRES shared by the iterative BEM solvers (see mnp_gmres.hpp).
#include <cmath>
#include <complex>
#include <vector>

#include "mnp_gmres.hpp"

namespace mnp {

// ------------------------------------------------------- GMRES kernels
// part[c + nchunk * j] = sum over chunk c of conj(V_j) w
__global__ void __launch_bounds__(256) k_mdot(int64_t len, const double2* V, int64_t ldv, const double2* w,
                                              double2* part) {
  const int64_t c = blockIdx.x, j = blockIdx.y, nchunk = gridDim.x;
  const double2* v = V + j * ldv;
  double sr = 0, si = 0;
  for (int64_t i = c * blockDim.x + threadIdx.x; i < len; i += nchunk * blockDim.x) {
    const double2 a = v[i], b = w[i];
    sr += a.x * b.x + a.y * b.y;
    si += a.x * b.y - a.y * b.x;
  }
  __shared__ double rr[256], ri[256];
  rr[threadIdx.x] = sr;
  ri[threadIdx.x] = si;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      rr[threadIdx.x] += rr[threadIdx.x + s];
      ri[threadIdx.x] += ri[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[c + nchunk * j] = mk<double>(rr[0], ri[0]);
}
// w <- w - sum_j V_j h_j  (h on the device, nj vectors)
__global__ void k_msub(int64_t len, const double2* V, int64_t ldv, const double2* h, int nj, double2* w) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= len) return;
  double2 s = w[i];
  for (int j = 0; j < nj; ++j) s = csub<double>(s, cmul<double>(V[j * ldv + i], h[j]));
  w[i] = s;
}
// x = sum_j V_j y_j
__global__ void k_mcomb(int64_t len, const double2* V, int64_t ldv, const double2* y, int nj, double2* x) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= len) return;
  double2 s = mk<double>(0, 0);
  for (int j = 0; j < nj; ++j) s = cadd<double>(s, cmul<double>(V[j * ldv + i], y[j]));
  x[i] = s;
}
__global__ void k_scale2(int64_t len, double2* v, double a) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) v[i] = cscale<double>(v[i], a);
}

void scale_vec(int64_t len, double2* v, double a) {
  hipLaunchKernelGGL(k_scale2, dim3((unsigned)cdiv_up(len, 256)), dim3(256), 0, stream(), len, v, a);
  MNP_HIP(hipGetLastError());
}

// conj(V_j) . w for j < nj (V: nj vectors of len at stride len)
static std::vector<std::complex<double>> mdot(int64_t len, const double2* V, int nj, const double2* w,
                                              Workspace& ws) {
  const int nchunk = 128;
  double2* part = (double2*)ws.get(3, sizeof(double2) * nchunk * (size_t)nj);
  hipLaunchKernelGGL(k_mdot, dim3(nchunk, nj), dim3(256), 0, stream(), len, V, len, w, part);
  std::vector<double2> h(nchunk * (size_t)nj);
  MNP_HIP(hipMemcpyAsync(h.data(), part, sizeof(double2) * h.size(), hipMemcpyDeviceToHost, stream()));
  MNP_HIP(hipStreamSynchronize(stream()));
  std::vector<std::complex<double>> r(nj);
  for (int j = 0; j < nj; ++j)
    for (int c = 0; c < nchunk; ++c)
      r[j] += std::complex<double>(h[c + (size_t)nchunk * j].x, h[c + (size_t)nchunk * j].y);
  return r;
}

GmresOut fgmres(int64_t len, const double2* b, double2* x, double tol, int maxit,
                const std::function<void(const double2*, double2*)>& afun,
                const std::function<void(const double2*, double2*)>& mfun, Workspace& ws) {
  const int m = std::max(1, (int)std::min<int64_t>(maxit, len));
  double2* V = (double2*)ws.get(1, sizeof(double2) * len * (m + 1));
  double2* Z = (double2*)ws.get(7, sizeof(double2) * len * m);
  double2* Ax = (double2*)ws.get(2, sizeof(double2) * len);
  double2* w = (double2*)ws.get(4, sizeof(double2) * len);
  double2* hbuf = (double2*)ws.get(5, sizeof(double2) * (m + 2));
  auto norm = [&](const double2* v) { return std::sqrt(std::max(0.0, mdot(len, v, 1, v, ws)[0].real())); };
  const unsigned gl = (unsigned)cdiv_up(len, 256);
  const double beta = norm(b);  // r0 = b (x0 = 0)
  using C = std::complex<double>;
  std::vector<std::vector<C>> Hm(m + 1, std::vector<C>(m, 0.0));
  std::vector<C> cs(m), sn(m), gv(m + 1, 0.0);
  gv[0] = beta;
  int it = 0;
  double res = beta;
  GmresOut out;
  if (beta > 0) {
    MNP_HIP(hipMemcpyAsync(V, b, sizeof(double2) * len, hipMemcpyDeviceToDevice, stream()));
    hipLaunchKernelGGL(k_scale2, dim3(gl), dim3(256), 0, stream(), len, V, 1.0 / beta);
    for (it = 0; it < m; ++it) {
      double2* zj = Z + (int64_t)it * len;
      mfun(V + (int64_t)it * len, zj);
      afun(zj, w);
      std::vector<C> h(it + 2, 0.0);
      for (int pass = 0; pass < 2; ++pass) {  // classical Gram-Schmidt, twice
        std::vector<C> d = mdot(len, V, it + 1, w, ws);
        std::vector<double2> dh(it + 1);
        for (int j = 0; j <= it; ++j) {
          h[j] += d[j];
          dh[j] = mk<double>(d[j].real(), d[j].imag());
        }
        MNP_HIP(hipMemcpyAsync(hbuf, dh.data(), sizeof(double2) * (it + 1), hipMemcpyHostToDevice, stream()));
        hipLaunchKernelGGL(k_msub, dim3(gl), dim3(256), 0, stream(), len, V, len, hbuf, it + 1, w);
        MNP_HIP(hipGetLastError());
      }
      const double hn = norm(w);
      h[it + 1] = hn;
      // Givens rotations on column it of the Hessenberg matrix
      for (int j = 0; j < it; ++j) {
        const C t = std::conj(cs[j]) * h[j] + std::conj(sn[j]) * h[j + 1];
        h[j + 1] = -sn[j] * h[j] + cs[j] * h[j + 1];
        h[j] = t;
      }
      const double a = std::abs(h[it]), bb = std::abs(h[it + 1]);
      const double rr = std::hypot(a, bb);
      cs[it] = rr > 0 ? h[it] / rr : C(1.0);
      sn[it] = rr > 0 ? h[it + 1] / rr : C(0.0);
      h[it] = rr;
      h[it + 1] = 0.0;
      gv[it + 1] = -sn[it] * gv[it];
      gv[it] = std::conj(cs[it]) * gv[it];
      for (int j = 0; j <= it; ++j) Hm[j][it] = h[j];
      res = std::abs(gv[it + 1]);
      if (res <= tol * beta || hn == 0 || rr == 0) {
        ++it;
        break;
      }
      if (it + 1 < m) {
        MNP_HIP(hipMemcpyAsync(V + (int64_t)(it + 1) * len, w, sizeof(double2) * len, hipMemcpyDeviceToDevice,
                               stream()));
        hipLaunchKernelGGL(k_scale2, dim3(gl), dim3(256), 0, stream(), len, V + (int64_t)(it + 1) * len, 1.0 / hn);
      }
    }
    if (it > m) it = m;
    // y = R^{-1} g (columns with a zero diagonal, from a breakdown, are dropped)
    int kk = it;
    while (kk > 0 && Hm[kk - 1][kk - 1] == C(0.0)) --kk;
    std::vector<C> y(kk, 0.0);
    for (int i = kk - 1; i >= 0; --i) {
      C s = gv[i];
      for (int j = i + 1; j < kk; ++j) s -= Hm[i][j] * y[j];
      y[i] = s / Hm[i][i];
    }
    std::vector<double2> dy(std::max(kk, 1));
    for (int j = 0; j < kk; ++j) dy[j] = mk<double>(y[j].real(), y[j].imag());
    if (kk) {
      MNP_HIP(hipMemcpyAsync(hbuf, dy.data(), sizeof(double2) * kk, hipMemcpyHostToDevice, stream()));
      hipLaunchKernelGGL(k_mcomb, dim3(gl), dim3(256), 0, stream(), len, Z, len, hbuf, kk, x);
    } else {
      MNP_HIP(hipMemsetAsync(x, 0, sizeof(double2) * len, stream()));
    }
    MNP_HIP(hipGetLastError());
  } else {
    MNP_HIP(hipMemsetAsync(x, 0, sizeof(double2) * len, stream()));
  }
  out.iter = it;
  // true relative residual |b - A x| / |b|
  afun(x, Ax);
  double2* r = V;  // reuse the first basis vector
  MNP_HIP(hipMemcpyAsync(r, b, sizeof(double2) * len, hipMemcpyDeviceToDevice, stream()));
  {
    std::vector<double2> one{mk<double>(1, 0)};
    MNP_HIP(hipMemcpyAsync(hbuf, one.data(), sizeof(double2), hipMemcpyHostToDevice, stream()));
    hipLaunchKernelGGL(k_msub, dim3(gl), dim3(256), 0, stream(), len, Ax, len, hbuf, 1, r);
  }
  out.relres = beta > 0 ? res / beta : 0;
  out.true_relres = beta > 0 ? norm(r) / beta : 0;
  out.flag = out.relres > tol ? 1 : 0;
  return out;
}

}  // namespace mnp
