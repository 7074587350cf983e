// Retarded particle above a substrate (one planar interface at z0, the
// particle in the upper medium), MNPBEM's bemretlayer path:
//
//   * the reflected Green function as Sommerfeld integrals over k_parallel
//     on the path of Particles/@layerstructure/green.m:39-127 (semi-ellipse
//     to 2 k1max, then the real axis when Z >= r / ratio, else a Hankel
//     split up and down from 2 k1max; intbessel.m, inthankel.m), with the
//     coefficients of reflectionsubs.m:27-60.  One wavefront per table point
//     integrates all 15 functions (G, dG/dr, dG/dz of the components p, ss,
//     hs, sh, hh) by composite Gauss-Legendre panels in FP64; Bessel and
//     Hankel functions of complex argument by their power series (|z| < 2)
//     and Watson's integral with generalized Gauss-Laguerre nodes;
//   * the normalised tables d G, d^3 Fr / r, d^3 Fz / zmin on the (r, z1)
//     grid of Greenfun/@greentablayer (norm.m), saved at several wavelengths
//     and interpolated linearly in between (eval.m:24-50);
//   * the reflected matrices of the particle with itself
//     (Greenfun/@greenretlayer/private/initrefl2.m): bilinear interpolation
//     (private/interp2.m, Misc/@igrid2/finterp.m) at the centroids, the
//     shape-function refinement of near pairs (init2.m:23-99) and the polar
//     quadrature of diagonal elements (initrefl2.m:39-99);
//   * the chain of BEM/@bemretlayer/private/initmat.m:17-95 and mldivide.m
//     for unique eps inside and outside, in complex FP64 on the library's
//     GEMM and LU kernels;
//   * the plane wave's potentials at the faces (Simulation/retarded/
//     @planewaveretlayer/potential.m with Particles/@layerstructure/
//     fresnel.m) and the far field with the interface (@spectrumretlayer/
//     farfield.m).

#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <type_traits>
#include <vector>

#include "mnp_chain.hpp"
#include "mnp_common.hpp"
#include "mnp_greenret.hpp"
#include "mnp_layertab.hpp"
#include "mnp_linalg.hpp"

namespace mnp {
namespace {

constexpr int NGL = 16;             // Gauss-Legendre nodes per panel
constexpr int NLAG = 32;            // generalized Gauss-Laguerre nodes of Watson's integral
constexpr int PANELS_ELLIPSE = 16;  // panels of the semi-ellipse leg
constexpr double TAIL_DECAY = 40.0; // tail legs end at exp(-40)

__constant__ double c_gl[2][NGL];        // nodes on [0, 1], weights
__constant__ double c_lag[2][2][NLAG];   // alpha = -1/2, +1/2: nodes, weights / Gamma(alpha + 1)

// Gauss rule of a Jacobi matrix (diagonal a, off-diagonal b[k] coupling k-1
// and k, total mass mu0): eigenvalues by Sturm-sequence bisection, weights
// mu0 / sum_k q_k(x)^2 from the three-term recurrence (Christoffel).
void jacobi_rule(int n, const std::vector<double>& a, const std::vector<double>& b, double mu0, double* x,
                 double* w) {
  double lo = 1e300, hi = -1e300;
  for (int k = 0; k < n; ++k) {
    const double rad = (k > 0 ? std::fabs(b[k]) : 0) + (k + 1 < n ? std::fabs(b[k + 1]) : 0);
    lo = std::min(lo, a[k] - rad);
    hi = std::max(hi, a[k] + rad);
  }
  auto count_below = [&](double t) {
    int c = 0;
    double d = a[0] - t;
    if (d < 0) ++c;
    for (int k = 1; k < n; ++k) {
      if (d == 0) d = 1e-300;
      d = (a[k] - t) - b[k] * b[k] / d;
      if (d < 0) ++c;
    }
    return c;
  };
  for (int k = 0; k < n; ++k) {
    double l = lo, h = hi;
    for (int it = 0; it < 200; ++it) {
      const double m = 0.5 * (l + h);
      if (m <= l || m >= h) break;
      if (count_below(m) > k)
        h = m;
      else
        l = m;
    }
    const double t = 0.5 * (l + h);
    double qm = 0, q = 1, s = 1;
    for (int j = 0; j + 1 < n; ++j) {
      const double qn = ((t - a[j]) * q - (j > 0 ? b[j] * qm : 0.0)) / b[j + 1];
      qm = q;
      q = qn;
      s += q * q;
    }
    x[k] = t;
    w[k] = mu0 / s;
  }
}

// The rules into constant memory, once per device.
void upload_rules() {
  static std::mutex mu;
  static std::vector<int> devs;
  std::lock_guard<std::mutex> lock(mu);
  int dev = 0;
  MNP_HIP(hipGetDevice(&dev));
  if (std::find(devs.begin(), devs.end(), dev) != devs.end()) return;
  std::vector<double> a(NLAG), b(NLAG + 1, 0.0);
  double gl[2][NGL];
  {
    std::vector<double> al(NGL, 0.0), bl(NGL + 1, 0.0);
    for (int k = 1; k < NGL; ++k) bl[k] = k / std::sqrt(4.0 * k * k - 1.0);
    double x[NGL], w[NGL];
    jacobi_rule(NGL, al, bl, 2.0, x, w);
    for (int k = 0; k < NGL; ++k) {
      gl[0][k] = 0.5 * (x[k] + 1);
      gl[1][k] = 0.5 * w[k];
    }
  }
  double lag[2][2][NLAG];
  for (int s = 0; s < 2; ++s) {
    const double alpha = s == 0 ? -0.5 : 0.5;
    for (int k = 0; k < NLAG; ++k) a[k] = 2.0 * k + alpha + 1;
    for (int k = 1; k < NLAG; ++k) b[k] = std::sqrt(k * (k + alpha));
    jacobi_rule(NLAG, a, b, 1.0, lag[s][0], lag[s][1]);
  }
  MNP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gl), gl, sizeof gl));
  MNP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_lag), lag, sizeof lag));
  devs.push_back(dev);
}

// ------------------------------------------------------ complex helpers
__host__ __device__ inline double2 csqrt_p(double2 z) {  // principal branch, signed zeros respected
  const double r = hypot(z.x, z.y);
  const double t = sqrt(0.5 * (r + fabs(z.x)));
  if (t == 0) return mk<double>(0, z.y);
  if (z.x >= 0) return mk<double>(t, z.y / (2 * t));
  return mk<double>(fabs(z.y) / (2 * t), copysign(t, z.y));
}
__device__ inline double2 cexp_(double2 z) {
  double s, c;
  sincos(z.y, &s, &c);
  const double m = exp(z.x);
  return mk<double>(m * c, m * s);
}
__device__ inline double2 clog_(double2 z) { return mk<double>(log(hypot(z.x, z.y)), atan2(z.y, z.x)); }
__device__ inline double2 conj_(double2 z) { return mk<double>(z.x, -z.y); }
__device__ inline double2 mul_i(double2 z) { return mk<double>(-z.y, z.x); }
__device__ inline double2 cfma(double2 acc, double2 a, double2 b) { return cadd<double>(acc, cmul<double>(a, b)); }

// kz = sqrt(k^2 - kpar^2) times sign(Im kz + 1e-10) (reflectionsubs.m:24-25)
__device__ inline double2 kz_of(double2 k, double2 kp) {
  const double2 kz = csqrt_p(csub<double>(cmul<double>(k, k), cmul<double>(kp, kp)));
  const double s = kz.y + 1e-10;
  return s > 0 ? kz : (s < 0 ? mk<double>(-kz.x, -kz.y) : mk<double>(0, 0));
}

// ------------------------------------------------------ Bessel functions
// J0, J1, Y0, Y1 by their power series, for |z| < 2.
__device__ void bessel_series(double2 z, double2& J0, double2& J1, double2& Y0, double2& Y1) {
  const double EG = 0.5772156649015329;
  const double2 q = cscale<double>(cmul<double>(z, z), 0.25);
  const double2 mq = mk<double>(-q.x, -q.y);
  double2 t0 = mk<double>(1, 0), t1 = mk<double>(1, 0);
  double2 s0 = mk<double>(0, 0), s1 = mk<double>(0, 0), y0 = mk<double>(0, 0), y1 = mk<double>(0, 0);
  double H = 0;
  for (int m = 0; m < 22; ++m) {
    if (m > 0) {
      t0 = cscale<double>(cmul<double>(t0, mq), 1.0 / ((double)m * m));
      t1 = cscale<double>(cmul<double>(t1, mq), 1.0 / ((double)m * (m + 1)));
      H += 1.0 / m;
    }
    s0 = cadd<double>(s0, t0);
    s1 = cadd<double>(s1, t1);
    y0 = csub<double>(y0, cscale<double>(t0, H));
    y1 = cadd<double>(y1, cscale<double>(t1, 2 * (-EG + H) + 1.0 / (m + 1)));
  }
  const double2 zh = cscale<double>(z, 0.5);
  J0 = s0;
  J1 = cmul<double>(s1, zh);
  const double2 L = clog_(zh);
  Y0 = cscale<double>(cadd<double>(cmul<double>(cadd<double>(L, mk<double>(EG, 0)), J0), y0), 2 / M_PI);
  Y1 = csub<double>(csub<double>(cscale<double>(cmul<double>(L, J1), 2 / M_PI),
                                 cdiv<double>(mk<double>(2 / M_PI, 0), z)),
                    cscale<double>(cmul<double>(zh, y1), 1 / M_PI));
}

// Hankel functions of the first kind H0, H1 for |z| >= 2, -pi/2 < arg z < 3 pi/2,
// by Watson's integral H_n(z) = sqrt(2 / (pi z)) exp(i (z - n pi/2 - pi/4)) /
// Gamma(n + 1/2) int_0^inf exp(-u) u^(n - 1/2) (1 + i u / (2 z))^(n - 1/2) du.
__device__ void hankel1_large(double2 z, double2& H0, double2& H1) {
  const double2 pre = cmul<double>(csqrt_p(cdiv<double>(mk<double>(2 / M_PI, 0), z)),
                                   cexp_(mul_i(mk<double>(z.x - 0.25 * M_PI, z.y))));
  const double2 iz2 = cdiv<double>(mk<double>(0, 0.5), z);  // i / (2 z)
  double2 a0 = mk<double>(0, 0), a1 = mk<double>(0, 0);
  for (int k = 0; k < NLAG; ++k) {
    const double2 w0 = cadd<double>(mk<double>(1, 0), cscale<double>(iz2, c_lag[0][0][k]));
    const double2 w1 = cadd<double>(mk<double>(1, 0), cscale<double>(iz2, c_lag[1][0][k]));
    a0 = cadd<double>(a0, cscale<double>(cdiv<double>(mk<double>(1, 0), csqrt_p(w0)), c_lag[0][1][k]));
    a1 = cadd<double>(a1, cscale<double>(csqrt_p(w1), c_lag[1][1][k]));
  }
  H0 = cmul<double>(pre, a0);
  H1 = mk<double>(0, 0);
  const double2 t = cmul<double>(pre, a1);
  H1 = mk<double>(t.y, -t.x);  // times exp(-i pi / 2)
}

__device__ void hankel1(double2 z, double2& H0, double2& H1) {
  if (hypot(z.x, z.y) < 2) {
    double2 J0, J1, Y0, Y1;
    bessel_series(z, J0, J1, Y0, Y1);
    H0 = cadd<double>(J0, mul_i(Y0));
    H1 = cadd<double>(J1, mul_i(Y1));
  } else {
    hankel1_large(z, H0, H1);
  }
}

// J0, J1 for complex z near the real axis: series below |z| = 2, else
// (H1(z) + conj(H1(conj z))) / 2.
__device__ void bessel_j(double2 z, double2& J0, double2& J1) {
  if (hypot(z.x, z.y) < 2) {
    double2 Y0, Y1;
    bessel_series(z, J0, J1, Y0, Y1);
    return;
  }
  double2 a0, a1, b0, b1;
  hankel1_large(z, a0, a1);
  hankel1_large(conj_(z), b0, b1);
  J0 = cscale<double>(cadd<double>(a0, conj_(b0)), 0.5);
  J1 = cscale<double>(cadd<double>(a1, conj_(b1)), 0.5);
}

// ------------------------------------------------------ table kernel
struct LayPrm {
  double k0;
  double2 e1, e2, k1, k2;  // upper and lower media
  double k1max, semi, ratio;
};

// Coefficient [0, 0] (both points above) of every component at kpar
// (reflectionsubs.m:27-60), order p, ss, hs, sh, hh; returns k1z.
__device__ double2 coeff_upper(const LayPrm& P, double2 kp, double2 c[5]) {
  const double2 k1z = kz_of(P.k1, kp), k2z = kz_of(P.k2, kp);
  const double2 sum = cadd<double>(k1z, k2z), dif = csub<double>(k1z, k2z);
  const double2 D = cmul<double>(cadd<double>(k2z, k1z), cadd<double>(cmul<double>(P.e1, k2z), cmul<double>(P.e2, k1z)));
  const double2 e1k1 = cmul<double>(P.e1, k1z), e2k1 = cmul<double>(P.e2, k1z), e1k2 = cmul<double>(P.e1, k2z);
  c[0] = cdiv<double>(dif, cadd<double>(k2z, k1z));
  c[1] = cdiv<double>(cmul<double>(sum, csub<double>(csub<double>(cscale<double>(e1k1, 2), e2k1), e1k2)), D);
  const double2 de = csub<double>(P.e2, P.e1);
  c[2] = cdiv<double>(cscale<double>(cmul<double>(cmul<double>(de, P.e1), k1z), -2 * P.k0), D);
  c[3] = cdiv<double>(cscale<double>(cmul<double>(de, k1z), -2 * P.k0), D);
  c[4] = cdiv<double>(cmul<double>(dif, cadd<double>(csub<double>(cscale<double>(e1k1, 2), e2k1), e1k2)), D);
  return k1z;
}

// Adds w times the integrands of G, Fr, Fz at kpar (intbessel.m:45-58).
__device__ void add_integrand(const LayPrm& P, double Z, double2 kp, double2 w, double2 b0, double2 b1,
                              double2 acc[NQ]) {
  double2 c[5];
  const double2 k1z = coeff_upper(P, kp, c);
  const double2 ph = cexp_(mul_i(cscale<double>(k1z, Z)));
  const double2 base = cdiv<double>(kp, k1z);                        // kpar / k1z
  const double2 wg = cmul<double>(w, mul_i(cmul<double>(b0, base)));  // i J0 kpar / k1z
  const double2 wr = cmul<double>(w, mul_i(cmul<double>(b1, cmul<double>(mk<double>(-kp.x, -kp.y), base))));
  const double2 wz = cmul<double>(w, cmul<double>(mk<double>(-b0.x, -b0.y), kp));
  for (int q = 0; q < 5; ++q) {
    const double2 R = cmul<double>(c[q], ph);
    acc[q] = cfma(acc[q], wg, R);
    acc[5 + q] = cfma(acc[5 + q], wr, R);
    acc[10 + q] = cfma(acc[10 + q], wz, R);
  }
}

// One wavefront per table point (r_i, Z_j): the integrals of the three legs
// and the normalised values d G, d^3 Fr / r, d^3 Fz / Z at out[(i nz + j) 15 + q].
__global__ void __launch_bounds__(64) k_layer_table(int64_t nr, int64_t nz, const double* __restrict__ rg,
                                                    const double* __restrict__ Zt, LayPrm P,
                                                    double2* __restrict__ out) {
  const int64_t pt = blockIdx.x;
  const int64_t i = pt / nz, j = pt % nz;
  const double r = rg[i], Z = Zt[j];
  const int lane = threadIdx.x;
  double2 acc[NQ];
  for (int q = 0; q < NQ; ++q) acc[q] = mk<double>(0, 0);
  // semi-ellipse kpar = k1max (1 - cos t - i semi sin t), t in [0, pi]
  for (int g = lane; g < PANELS_ELLIPSE * NGL; g += 64) {
    const int p = g / NGL, q = g % NGL;
    const double t = (p + c_gl[0][q]) * M_PI / PANELS_ELLIPSE;
    const double wt = c_gl[1][q] * M_PI / PANELS_ELLIPSE;
    const double st = sin(t), ct = cos(t);
    const double2 kp = mk<double>(P.k1max * (1 - ct), -P.k1max * P.semi * st);
    const double2 w = mk<double>(P.k1max * st * wt, -P.k1max * P.semi * ct * wt);
    double2 J0, J1;
    bessel_j(cscale<double>(kp, r), J0, J1);
    add_integrand(P, Z, kp, w, J0, J1, acc);
  }
  // tail: the real axis (Z >= r / ratio) or the Hankel leg, panels of width
  // k1max / 4 doubling up to the decay length
  const bool real_leg = Z >= r / P.ratio;
  const double scale = 1.0 / (real_leg ? Z : r);
  const double L = TAIL_DECAY * scale;
  double e = 0, h = 0.25 * P.k1max;
  const int sub = lane / NGL, q = lane % NGL;
  for (int p = 0; e < L; ++p) {
    const double e1 = fmin(e + fmin(h, scale), L);
    if ((p & 3) == sub) {
      const double hp = e1 - e;
      const double s = e + hp * c_gl[0][q], ws = hp * c_gl[1][q];
      if (real_leg) {
        const double2 kp = mk<double>(2 * P.k1max + s, 0);
        double2 J0, J1;
        bessel_j(mk<double>(kp.x * r, 0), J0, J1);
        add_integrand(P, Z, kp, mk<double>(ws, 0), J0, J1, acc);
      } else {
        // 0.5 H1(kpar r) f(kpar) dkpar + 0.5 H2(kpar* r) f(kpar*) dkpar*
        const double2 kp = mk<double>(2 * P.k1max, s);
        double2 H0, H1;
        hankel1(cscale<double>(kp, r), H0, H1);
        add_integrand(P, Z, kp, mk<double>(0, 0.5 * ws), H0, H1, acc);
        add_integrand(P, Z, conj_(kp), mk<double>(0, -0.5 * ws), conj_(H0), conj_(H1), acc);
      }
    }
    e = e1;
    h *= 2;
  }
  for (int qq = 0; qq < NQ; ++qq)
    for (int off = 32; off > 0; off >>= 1) {
      acc[qq].x += __shfl_xor(acc[qq].x, off);
      acc[qq].y += __shfl_xor(acc[qq].y, off);
    }
  if (lane == 0) {
    const double d = sqrt(r * r + Z * Z);
    const double fr = d * d * d / r, fz = d * d * d / Z;
    double2* o = out + pt * NQ;
    for (int c = 0; c < 5; ++c) {
      o[c] = cscale<double>(acc[c], d);
      o[5 + c] = cscale<double>(acc[5 + c], fr);
      o[10 + c] = cscale<double>(acc[10 + c], fz);
    }
  }
}

// out = (1 - t) a + t b
__global__ void k_lerp(int64_t len, const double2* a, const double2* b, double t, double2* out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) out[i] = mk<double>((1 - t) * a[i].x + t * b[i].x, (1 - t) * a[i].y + t * b[i].y);
}

// ------------------------------------------------------ interpolation
// Reflected G and F = nvec . Gp at the centroids, all pairs
// (initrefl2.m:21-37): G[c n^2 + j n + i], F likewise.
__global__ void __launch_bounds__(256) k_layer_assemble(int64_t n, Geo geo, TabView T, double2* __restrict__ G,
                                                        double2* __restrict__ F) {
#pragma clang fp contract(off)
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int64_t i = idx % n, j = idx / n;
  const double x = geo.pos[i] - geo.pos[j], y = geo.pos[n + i] - geo.pos[n + j];
  const double r = sqrt(x * x + y * y);
  double2 o[NQ];
  tab_interp(T, r, geo.pos[2 * n + i], geo.pos[2 * n + j], o);
  const double a = geo.area[j], rr = fmax(r, 1e-10);
  const double nx = geo.nvec[i], ny = geo.nvec[n + i], nz = geo.nvec[2 * n + i];
  const int64_t nn = n * n;
  for (int c = 0; c < 5; ++c) {
    G[c * nn + idx] = cscale<double>(o[c], a);
    const double2 gx = cscale<double>(cscale<double>(cscale<double>(o[5 + c], x), 1.0 / rr), a);
    const double2 gy = cscale<double>(cscale<double>(cscale<double>(o[5 + c], y), 1.0 / rr), a);
    const double2 gz = cscale<double>(o[10 + c], a);
    F[c * nn + idx] = mk<double>(nx * gx.x + ny * gy.x + nz * gz.x, nx * gx.y + ny * gy.y + nz * gz.y);
  }
}

// Refined near pairs, weights once per geometry (init2.m:23-99): per pair p
// (row i, face j) and vertex a, W[(p 3 + a) 4 + {g, f1, f2, fz}], the
// vertex's clamped radius R0[p 3 + a] and height Z2[p 3 + a], and the row's
// height Z1[p].
__device__ inline void lay_dist(const LayGeo& L, const double* p1, const double* p2, double& r, double& z,
                                double& d, double& x, double& y) {
#pragma clang fp contract(off)
  x = p1[0] - p2[0];
  y = p1[1] - p2[1];
  z = round_z(fabs(p1[2] - L.z0), L.z0, L.zmin) + round_z(fabs(p2[2] - L.z0), L.z0, L.zmin);
  r = fmax(L.rmin, sqrt(x * x + y * y));
  d = sqrt(r * r + z * z);
}

// rpos: the rows' positions (nr x 3, column-major): the faces' centroids,
// or points for compgreenretlayer(pt, p).
__global__ void __launch_bounds__(256) k_layer_refine_init(int64_t np, int64_t n, Geo geo, LayGeo L,
                                                           const double* __restrict__ rpos, int64_t nr,
                                                           const int32_t* __restrict__ rows,
                                                           const int32_t* __restrict__ cols,
                                                           const double* __restrict__ q18, double* W, double* R0,
                                                           double* Z2, double* Z1) {
#pragma clang fp contract(off)
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int64_t i = rows[p], j = cols[p];
  const double pi3[3] = {rpos[i], rpos[nr + i], rpos[2 * nr + i]};
  double v[3][3];
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) v[a][c] = geo.vtx[(a * 3 + c) * n + j];
  double r0[3], z0[3], d0[3];
  for (int a = 0; a < 3; ++a) {
    double x, y;
    lay_dist(L, pi3, v[a], r0[a], z0[a], d0[a], x, y);
  }
  double acc[3][4] = {};
  const double qa = geo.qarea[j];
  for (int q = 0; q < 28; ++q) {
    const double xq = q18[q], yq = q18[28 + q], wq = q18[56 + q] * qa, zq = 1 - xq - yq;
    double pt[3];
    for (int c = 0; c < 3; ++c) pt[c] = xq * v[0][c] + yq * v[1][c] + zq * v[2][c];
    double r, z, d, x, y;
    lay_dist(L, pi3, pt, r, z, d, x, y);
    const double d3 = d * d * d;
    const double s[3] = {xq, yq, zq};
    for (int a = 0; a < 3; ++a) {
      const double fr0 = d0[a] * d0[a] * d0[a] / r0[a] * s[a];
      const double fz0 = d0[a] * d0[a] * d0[a] / z0[a] * s[a];
      acc[a][0] += d0[a] * s[a] / d * wq;
      acc[a][1] += fr0 * x / d3 * wq;
      acc[a][2] += fr0 * y / d3 * wq;
      acc[a][3] += fz0 * z / d3 * wq;
    }
  }
  for (int a = 0; a < 3; ++a) {
    for (int k = 0; k < 4; ++k) W[(p * 3 + a) * 4 + k] = acc[a][k];
    R0[p * 3 + a] = r0[a];
    Z2[p * 3 + a] = round_z(v[a][2], L.z0, L.zmin);
  }
  Z1[p] = round_z(pi3[2], L.z0, L.zmin);
}

__global__ void __launch_bounds__(256) k_layer_refine_apply(int64_t np, int64_t n, Geo geo, TabView T,
                                                            const int32_t* __restrict__ rows,
                                                            const int32_t* __restrict__ cols,
                                                            const double* __restrict__ W,
                                                            const double* __restrict__ R0,
                                                            const double* __restrict__ Z2,
                                                            const double* __restrict__ Z1, double2* G, double2* F,
                                                            int compact) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int64_t i = rows[p], j = cols[p];
  double2 g[5], gx[5], gy[5], gz[5];
  for (int c = 0; c < 5; ++c) g[c] = gx[c] = gy[c] = gz[c] = mk<double>(0, 0);
  for (int a = 0; a < 3; ++a) {
    double2 o[NQ];
    tab_interp(T, R0[p * 3 + a], Z1[p], Z2[p * 3 + a], o);
    const double* w = W + (p * 3 + a) * 4;
    for (int c = 0; c < 5; ++c) {
      g[c] = cadd<double>(g[c], cscale<double>(o[c], w[0]));
      gx[c] = cadd<double>(gx[c], cscale<double>(o[5 + c], w[1]));
      gy[c] = cadd<double>(gy[c], cscale<double>(o[5 + c], w[2]));
      gz[c] = cadd<double>(gz[c], cscale<double>(o[10 + c], w[3]));
    }
  }
  const double nx = geo.nvec[i], ny = geo.nvec[n + i], nz = geo.nvec[2 * n + i];
  // dense 5 x n x n, or compact 5 x np (compact != 0)
  const int64_t nn = compact ? np : n * n, idx = compact ? p : j * n + i;
  for (int c = 0; c < 5; ++c) {
    G[c * nn + idx] = g[c];
    F[c * nn + idx] = mk<double>(nx * gx[c].x + ny * gy[c].x + nz * gz[c].x, nx * gx[c].y + ny * gy[c].y + nz * gz[c].y);
  }
}

// Diagonal elements with the polar rule (initrefl2.m:39-99), faces id[0..nd).
__global__ void __launch_bounds__(64) k_layer_diag(int64_t nd, int64_t n, Geo geo, TabView T,
                                                   const int32_t* __restrict__ id, const double* __restrict__ qpol,
                                                   int npol, double2* G, double2* F, int compact) {
#pragma clang fp contract(off)
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nd) return;
  const int64_t f = id[t];
  double v[3][3];
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) v[a][c] = geo.vtx[(a * 3 + c) * n + f];
  const double px = geo.pos[f], py = geo.pos[n + f], pz = geo.pos[2 * n + f], area = geo.area[f];
  double2 g[5], gx[5], gy[5], gz[5];
  for (int c = 0; c < 5; ++c) g[c] = gx[c] = gy[c] = gz[c] = mk<double>(0, 0);
  for (int q = 0; q < npol; ++q) {
    const double xq = qpol[q], yq = qpol[MNP_MAX_POLAR + q], zq = 1 - xq - yq;
    const double w = qpol[2 * MNP_MAX_POLAR + q] * area;
    double pt[3];
    for (int c = 0; c < 3; ++c) pt[c] = xq * v[0][c] + yq * v[1][c] + zq * v[2][c];
    const double x = pt[0] - px, y = pt[1] - py;
    const double r = sqrt(x * x + y * y);
    double2 o[NQ];
    const double rc = fmax(tab_interp(T, r, pt[2], pz, o), 1e-10);
    for (int c = 0; c < 5; ++c) {
      g[c] = cadd<double>(g[c], cscale<double>(o[c], w));
      gx[c] = cadd<double>(gx[c], cscale<double>(cscale<double>(cscale<double>(o[5 + c], x), 1.0 / rc), w));
      gy[c] = cadd<double>(gy[c], cscale<double>(cscale<double>(cscale<double>(o[5 + c], y), 1.0 / rc), w));
      gz[c] = cadd<double>(gz[c], cscale<double>(o[10 + c], w));
    }
  }
  const double nx = geo.nvec[f], ny = geo.nvec[n + f], nz = geo.nvec[2 * n + f];
  const int64_t nn = compact ? nd : n * n, idx = compact ? t : f * n + f;  // dense or compact 5 x nd
  for (int c = 0; c < 5; ++c) {
    G[c * nn + idx] = g[c];
    F[c * nn + idx] = mk<double>(nx * gx[c].x + ny * gy[c].x + nz * gz[c].x, nx * gx[c].y + ny * gy[c].y + nz * gz[c].y);
  }
}


// Reflected potentials and fields at points above the interface from the
// outside charges and currents (compgreenretlayer(pt, p): potential.m,
// field.m with private/matmul2.m, matmul3.m), one workgroup of 256 per
// (point, column q), its threads over the faces, summed in the workgroup:
// every face by the table at the centroids, the refined pairs of point
// i (rowptr[i] .. rowptr[i + 1], faces increasing) from their vertex
// weights.  phi (npt x ncol), a, e, h (npt x 3 x ncol): phi = Gss s + Gsh hz,
// a = [Gp hx, Gp hy, Ghh hz + Ghs s], e = i k0 a - (Gp_ss s + Gp_sh hz),
// h = curl of the structured product.
__global__ void __launch_bounds__(256) k_pt_layer(int64_t npt, int64_t ncol, const double* __restrict__ pts,
                                                  int64_t n, Geo geo, TabView T, const int64_t* __restrict__ rowptr,
                                                  const int32_t* __restrict__ pcol, const double* __restrict__ W,
                                                  const double* __restrict__ R0, const double* __restrict__ Z2,
                                                  const double* __restrict__ Z1, double k0,
                                                  const double2* __restrict__ sig2, const double2* __restrict__ h2,
                                                  double2* phi, double2* a, double2* e, double2* h) {
#pragma clang fp contract(off)
  const int64_t i = blockIdx.x % npt, q = blockIdx.x / npt;
  const double px = pts[i], py = pts[npt + i], pz = pts[2 * npt + i];
  double2 ph = mk<double>(0, 0), av[3], ev[3], hv[3];
  for (int c = 0; c < 3; ++c) av[c] = ev[c] = hv[c] = mk<double>(0, 0);
  const int64_t p0 = rowptr[i], pe = rowptr[i + 1];
  for (int64_t j = threadIdx.x; j < n; j += blockDim.x) {
    double2 g[5], gp[5][3];
    int64_t pr = -1;  // refined pair (i, j): binary search of the point's sorted face list
    for (int64_t lo = p0, hi = pe - 1; lo <= hi;) {
      const int64_t mid = (lo + hi) >> 1;
      if (pcol[mid] == j) {
        pr = mid;
        break;
      }
      if (pcol[mid] < j)
        lo = mid + 1;
      else
        hi = mid - 1;
    }
    if (pr >= 0) {  // refined pair: sum over the face's vertices
      for (int c = 0; c < 5; ++c) g[c] = gp[c][0] = gp[c][1] = gp[c][2] = mk<double>(0, 0);
      for (int v = 0; v < 3; ++v) {
        double2 o[NQ];
        tab_interp(T, R0[pr * 3 + v], Z1[pr], Z2[pr * 3 + v], o);
        const double* w = W + (pr * 3 + v) * 4;
        for (int c = 0; c < 5; ++c) {
          g[c] = cadd<double>(g[c], cscale<double>(o[c], w[0]));
          gp[c][0] = cadd<double>(gp[c][0], cscale<double>(o[5 + c], w[1]));
          gp[c][1] = cadd<double>(gp[c][1], cscale<double>(o[5 + c], w[2]));
          gp[c][2] = cadd<double>(gp[c][2], cscale<double>(o[10 + c], w[3]));
        }
      }
    } else {
      const double x = px - geo.pos[j], y = py - geo.pos[n + j];
      const double r = sqrt(x * x + y * y);
      double2 o[NQ];
      tab_interp(T, r, pz, geo.pos[2 * n + j], o);
      const double ar = geo.area[j], rr = fmax(r, 1e-10);
      for (int c = 0; c < 5; ++c) {
        g[c] = cscale<double>(o[c], ar);
        gp[c][0] = cscale<double>(cscale<double>(cscale<double>(o[5 + c], x), 1.0 / rr), ar);
        gp[c][1] = cscale<double>(cscale<double>(cscale<double>(o[5 + c], y), 1.0 / rr), ar);
        gp[c][2] = cscale<double>(o[10 + c], ar);
      }
    }
    // components p 0, ss 1, hs 2, sh 3, hh 4
    const double2 s = sig2[q * n + j];
    const double2 hx = h2[(0 + 3 * q) * n + j], hy = h2[(1 + 3 * q) * n + j], hz = h2[(2 + 3 * q) * n + j];
    ph = cfma(cfma(ph, g[1], s), g[3], hz);
    av[0] = cfma(av[0], g[0], hx);
    av[1] = cfma(av[1], g[0], hy);
    av[2] = cfma(cfma(av[2], g[4], hz), g[2], s);
    double2 M[3][3];
    for (int c = 0; c < 3; ++c) {
      ev[c] = csub<double>(ev[c], cadd<double>(cmul<double>(gp[1][c], s), cmul<double>(gp[3][c], hz)));
      M[c][0] = cmul<double>(gp[0][c], hx);
      M[c][1] = cmul<double>(gp[0][c], hy);
      M[c][2] = cadd<double>(cmul<double>(gp[4][c], hz), cmul<double>(gp[2][c], s));
    }
    hv[0] = cadd<double>(hv[0], csub<double>(M[1][2], M[2][1]));
    hv[1] = cadd<double>(hv[1], csub<double>(M[2][0], M[0][2]));
    hv[2] = cadd<double>(hv[2], csub<double>(M[0][1], M[1][0]));
  }
  // sums over the workgroup: xor butterflies in each wave, then the four
  // wave partials in order
  __shared__ double2 part[4][10];
  auto wred = [&](double2 v, int u) {
    for (int o = 32; o > 0; o >>= 1) {
      v.x += __shfl_xor(v.x, o);
      v.y += __shfl_xor(v.y, o);
    }
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6][u] = v;
  };
  wred(ph, 0);
  for (int c = 0; c < 3; ++c) {
    wred(av[c], 1 + c);
    wred(ev[c], 4 + c);
    wred(hv[c], 7 + c);
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double2 r[10];
  for (int u = 0; u < 10; ++u)
    r[u] = cadd<double>(cadd<double>(part[0][u], part[1][u]), cadd<double>(part[2][u], part[3][u]));
  phi[q * npt + i] = r[0];
  for (int c = 0; c < 3; ++c) {
    const int64_t id = (c + 3 * q) * npt + i;
    a[id] = r[1 + c];
    e[id] = cadd<double>(mul_i(cscale<double>(r[1 + c], k0)), r[4 + c]);
    h[id] = r[7 + c];
  }
}

// ------------------------------------------------------ plane wave and far field
// Coefficient of component set (p, hh, sh, hs) for the observer in medium
// o and the source in medium s (1 = above), reflectionsubs.m:27-60,
// without phases; returns kz of media 1 and 2 in kz[0..1].
struct Coef4 {
  double2 p, hh, sh, hs;
};
__device__ Coef4 coeff_entry(double k0, double2 e1, double2 e2, double2 k1z, double2 k2z, int o, int s) {
  const double2 one = mk<double>(1, 0);
  const double2 rr = cdiv<double>(csub<double>(k1z, k2z), cadd<double>(k2z, k1z));
  const double2 D = cmul<double>(cadd<double>(k2z, k1z), cadd<double>(cmul<double>(e1, k2z), cmul<double>(e2, k1z)));
  const double2 e1k1 = cmul<double>(e1, k1z), e2k1 = cmul<double>(e2, k1z), e1k2 = cmul<double>(e1, k2z),
                e2k2 = cmul<double>(e2, k2z);
  const double2 hh1 = cdiv<double>(cmul<double>(csub<double>(k1z, k2z), cadd<double>(csub<double>(cscale<double>(e1k1, 2), e2k1), e1k2)), D);
  const double2 hh2 = cdiv<double>(cmul<double>(csub<double>(k2z, k1z), cadd<double>(csub<double>(cscale<double>(e2k2, 2), e1k2), e2k1)), D);
  const double2 sh1 = cdiv<double>(cscale<double>(cmul<double>(csub<double>(e2, e1), k1z), -2 * k0), D);
  const double2 sh2 = cdiv<double>(cscale<double>(cmul<double>(csub<double>(e1, e2), k2z), -2 * k0), D);
  const double2 hs1 = cdiv<double>(cscale<double>(cmul<double>(cmul<double>(csub<double>(e2, e1), e1), k1z), -2 * k0), D);
  const double2 hs2 = cdiv<double>(cscale<double>(cmul<double>(cmul<double>(csub<double>(e1, e2), e2), k2z), -2 * k0), D);
  const double2 q12 = cdiv<double>(k1z, k2z), q21 = cdiv<double>(k2z, k1z);
  Coef4 c;
  if (o == 1 && s == 1) {
    c.p = rr; c.hh = hh1; c.sh = sh1; c.hs = hs1;
  } else if (o == 1 && s == 2) {  // mat1 / mat2 entry (1, 2)
    c.p = cadd<double>(one, rr);
    c.hh = cmul<double>(q12, cadd<double>(hh2, one));
    c.sh = cmul<double>(mk<double>(-q12.x, -q12.y), sh2);
    c.hs = cmul<double>(mk<double>(-q12.x, -q12.y), hs2);
  } else if (o == 2 && s == 1) {  // entry (2, 1)
    c.p = csub<double>(one, rr);
    c.hh = cmul<double>(q21, cadd<double>(hh1, one));
    c.sh = cmul<double>(q21, sh1);
    c.hs = cmul<double>(q21, hs1);
  } else {
    c.p = mk<double>(-rr.x, -rr.y);
    c.hh = hh2;
    c.sh = mk<double>(-sh2.x, -sh2.y);
    c.hs = mk<double>(-hs2.x, -hs2.y);
  }
  return c;
}

// kz with fresnel.m's 1e-10 i offset before the sign
__device__ inline double2 kz_off(double2 k, double2 kp) {
  double2 kz = csqrt_p(csub<double>(cmul<double>(k, k), cmul<double>(kp, kp)));
  kz.y += 1e-10;
  return kz.y > 0 ? kz : (kz.y < 0 ? mk<double>(-kz.x, -kz.y) : mk<double>(0, 0));
}

// Outside potentials of the plane waves at the faces (potential.m:20-105,
// every face above the interface): one thread per (face, wave); the normal
// derivative differentiates the phase factors exactly.
__global__ void k_pw_layer(int64_t n, int64_t npol, const double* __restrict__ pos, const double* __restrict__ nvec,
                           const double* __restrict__ pd, double k0, double2 e1, double2 e2, double z0,
                           double2* phi, double2* phip, double2* a, double2* ap) {
#pragma clang fp contract(off)
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * npol) return;
  const int64_t i = t % n, q = t / n;
  const double pol[3] = {pd[q], pd[npol + q], pd[2 * npol + q]};
  const double dir[3] = {pd[3 * npol + q], pd[4 * npol + q], pd[5 * npol + q]};
  const int medium = dir[2] < 0 ? 1 : 2;
  const double z2 = medium == 1 ? z0 + 1e-10 : z0 - 1e-10;
  const double2 nb = csqrt_p(medium == 1 ? e1 : e2);
  const double2 kp = cscale<double>(nb, k0 * sqrt(dir[0] * dir[0] + dir[1] * dir[1]));
  const double2 k1 = cscale<double>(csqrt_p(e1), k0), k2 = cscale<double>(csqrt_p(e2), k0);
  const double2 k1z = kz_of(k1, kp), k2z = kz_of(k2, kp);
  const Coef4 c = coeff_entry(k0, e1, e2, k1z, k2z, 1, medium);
  const double x = pos[i], y = pos[n + i], z = pos[2 * n + i];
  const double2 g1 = cexp_(mul_i(cscale<double>(k1z, fabs(z - z0))));
  const double2 g2 = cexp_(mul_i(cscale<double>(medium == 1 ? k1z : k2z, fabs(z2 - z0))));
  const double2 ratio = cdiv<double>(medium == 1 ? kz_off(k1, kp) : kz_off(k2, kp), kz_off(k1, kp));
  const double2 ph = cmul<double>(cmul<double>(g1, g2), ratio);
  const double inn = x * dir[0] + y * dir[1] + z2 * dir[2];
  const double2 fac = cdiv<double>(cexp_(mul_i(cscale<double>(nb, k0 * inn))), mk<double>(0, k0));
  const double nx = nvec[i], ny = nvec[n + i], nz = nvec[2 * n + i];
  const double2 dph = mul_i(cadd<double>(cscale<double>(nb, k0 * (nx * dir[0] + ny * dir[1])), cscale<double>(k1z, nz)));
  double2 av[3];
  const double2 rp = cmul<double>(c.p, ph);
  av[0] = cmul<double>(cscale<double>(fac, pol[0]), rp);
  av[1] = cmul<double>(cscale<double>(fac, pol[1]), rp);
  av[2] = cmul<double>(cscale<double>(fac, pol[2]), cmul<double>(c.hh, ph));
  const double2 pv = cmul<double>(cscale<double>(fac, pol[2]), cmul<double>(c.sh, ph));
  double2 apv[3];
  for (int cc = 0; cc < 3; ++cc) apv[cc] = cmul<double>(av[cc], dph);
  if (medium == 1) {
    const double2 f0 = cdiv<double>(cexp_(mul_i(cscale<double>(nb, k0 * (x * dir[0] + y * dir[1] + z * dir[2])))),
                                    mk<double>(0, k0));
    const double2 d0 = mul_i(cscale<double>(nb, k0 * (nx * dir[0] + ny * dir[1] + nz * dir[2])));
    for (int cc = 0; cc < 3; ++cc) {
      const double2 a0 = cscale<double>(f0, pol[cc]);
      av[cc] = cadd<double>(av[cc], a0);
      apv[cc] = cadd<double>(apv[cc], cmul<double>(a0, d0));
    }
  }
  phi[q * n + i] = pv;
  phip[q * n + i] = cmul<double>(pv, dph);
  for (int cc = 0; cc < 3; ++cc) {
    a[(cc + 3 * q) * n + i] = av[cc];
    ap[(cc + 3 * q) * n + i] = apv[cc];
  }
}

constexpr int LFF_MAXP = 8;

// Far field with the interface (@spectrumretlayer/farfield.m:20-104), one
// workgroup per direction: reflected (upper half space) or transmitted
// (lower) part of the outside charges and currents, through a lossless
// medium only, plus the direct far field above; e transverse, h = k / k0
// dir x e.  e, h at d + ndir (c + 3 q).
__global__ void __launch_bounds__(256) k_farfield_layer(int64_t n, int64_t p, int64_t ndir, const double* dir,
                                                        const double* pos, const double* area, const double2* sig2,
                                                        const double2* h2, double k0, double2 e1, double2 e2,
                                                        double z0, double2* e, double2* h) {
  const int64_t d = blockIdx.x;
  const double dx = dir[d], dy = dir[ndir + d], dz = dir[2 * ndir + d];
  const bool up = dz >= 0;
  const double2 k1 = cscale<double>(csqrt_p(e1), k0), k2 = cscale<double>(csqrt_p(e2), k0);
  const double2 km = up ? k1 : k2;
  const bool refl = km.y == 0;
  const double2 kp = mk<double>(km.x * sqrt(1 - dz * dz), 0);
  const double2 k1z = kz_of(k1, kp), k2z = kz_of(k2, kp);
  const Coef4 c = coeff_entry(k0, e1, e2, k1z, k2z, up ? 1 : 2, 1);
  const double2 g1 = cexp_(mul_i(cscale<double>(up ? k1z : k2z, 0.0)));
  double2 acc[LFF_MAXP][7];
  for (int q = 0; q < LFF_MAXP; ++q)
    for (int cc = 0; cc < 7; ++cc) acc[q][cc] = mk<double>(0, 0);
  for (int64_t j = threadIdx.x; j < n; j += blockDim.x) {
    const double x = pos[j], y = pos[n + j], z = pos[2 * n + j];
    if (refl) {
      const double dist = x * dx + y * dy + z0 * dz;
      const double2 phase = cscale<double>(cexp_(mul_i(cscale<double>(km, -dist))), area[j]);
      const double2 g = cmul<double>(g1, cexp_(mul_i(cscale<double>(k1z, fabs(z - z0)))));
      const double2 wp = cmul<double>(cmul<double>(c.p, g), phase), wh = cmul<double>(cmul<double>(c.hh, g), phase),
                    ws = cmul<double>(cmul<double>(c.hs, g), phase);
      for (int q = 0; q < p; ++q) {
        acc[q][0] = cfma(acc[q][0], wp, h2[(0 + 3 * q) * n + j]);
        acc[q][1] = cfma(acc[q][1], wp, h2[(1 + 3 * q) * n + j]);
        acc[q][2] = cfma(cfma(acc[q][2], wh, h2[(2 + 3 * q) * n + j]), ws, sig2[q * n + j]);
      }
    }
    if (up) {
      const double pd = dx * x + dy * y + dz * z;
      const double2 ph = cscale<double>(cexp_ikd(mk<double>(-k1.x, -k1.y), pd), area[j]);
      for (int q = 0; q < p; ++q) {
        for (int cc = 0; cc < 3; ++cc) acc[q][3 + cc] = cfma(acc[q][3 + cc], ph, h2[(cc + 3 * q) * n + j]);
        acc[q][6] = cfma(acc[q][6], ph, sig2[q * n + j]);
      }
    }
  }
  __shared__ double2 red[256];
  for (int q = 0; q < p; ++q)
    for (int cc = 0; cc < 7; ++cc) {
      red[threadIdx.x] = acc[q][cc];
      __syncthreads();
      for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = cadd<double>(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
      }
      if (threadIdx.x == 0) acc[q][cc] = red[0];
      __syncthreads();
    }
  if (threadIdx.x == 0) {
    const double2 ik0 = mk<double>(0, k0), ik1 = mul_i(k1);
    const double dv[3] = {dx, dy, dz};
    const double kr = km.x / k0, ki = km.y / k0;
    for (int q = 0; q < p; ++q) {
      double2 ev[3];
      for (int cc = 0; cc < 3; ++cc) {
        ev[cc] = cmul<double>(ik0, acc[q][cc]);
        if (up)
          ev[cc] = cadd<double>(ev[cc], csub<double>(cmul<double>(ik0, acc[q][3 + cc]),
                                                     cscale<double>(cmul<double>(ik1, acc[q][6]), dv[cc])));
      }
      const double2 de = cadd<double>(cadd<double>(cscale<double>(ev[0], dx), cscale<double>(ev[1], dy)),
                                      cscale<double>(ev[2], dz));
      for (int cc = 0; cc < 3; ++cc) ev[cc] = csub<double>(ev[cc], cscale<double>(de, dv[cc]));
      const double2 cx = csub<double>(cscale<double>(ev[2], dy), cscale<double>(ev[1], dz));
      const double2 cy = csub<double>(cscale<double>(ev[0], dz), cscale<double>(ev[2], dx));
      const double2 cz = csub<double>(cscale<double>(ev[1], dx), cscale<double>(ev[0], dy));
      const double2 kk = mk<double>(kr, ki);
      for (int cc = 0; cc < 3; ++cc) e[d + ndir * (cc + 3 * q)] = ev[cc];
      h[d + ndir * (0 + 3 * q)] = cmul<double>(kk, cx);
      h[d + ndir * (1 + 3 * q)] = cmul<double>(kk, cy);
      h[d + ndir * (2 + 3 * q)] = cmul<double>(kk, cz);
    }
  }
}

// ------------------------------------------------------ dipoles above the interface
// greenderiv.m:32-57's base values of component c for one observer-dipole
// pair: the interpolated G, Fr, F1 (d/dz at the observer) and forward
// differences with eta = 1e-6 of the interpolated table in r (Frr, Fr1)
// and in the dipole height (F2, Fr2, F12).
struct GdBase {
  double2 G, Fr, F1, Frr, Fr1, F2, Fr2, F12;
};
__device__ inline double2 fdiff(double2 a, double2 b, double eta) {
  return mk<double>((a.x - b.x) / eta, (a.y - b.y) / eta);
}
__device__ GdBase gd_base(const TabView& T, double r, double z1, double z2, int c) {
  const double eta = 1e-6;
  GdBase b;
  double2 t, fr, fz, g;
  tab_interp_c(T, r, z1, z2, c, 0, b.G, t, t);
  tab_interp_c(T, r, z1, z2, c, 1, t, b.Fr, b.F1);
  tab_interp_c(T, r + eta, z1, z2, c, 1, t, fr, fz);
  b.Frr = fdiff(fr, b.Fr, eta);
  b.Fr1 = fdiff(fz, b.F1, eta);
  tab_interp_c(T, r, z1, z2 + eta, c, 0, g, t, t);
  tab_interp_c(T, r, z1, z2 + eta, c, 1, t, fr, fz);
  b.F2 = fdiff(g, b.G, eta);
  b.Fr2 = fdiff(fr, b.Fr, eta);
  b.F12 = fdiff(fz, b.F1, eta);
  return b;
}
// F{a, b} of greenderiv.m:58-96: a = 2, 3, 4 differentiates in x, y, z at
// the observer, b = 2, 3, 4 in x', y', z' at the dipole, 1 neither; the
// mixed F{2, 3} = F{3, 2} as greenderiv.m:73-78 writes it.
__device__ inline double2 gd_F(const GdBase& B, int a, int b, double ux, double uy, double r) {
  auto sc = [](double2 z, double s) { return mk<double>(z.x * s, z.y * s); };
  auto sub2 = [](double2 p, double2 q) { return mk<double>(p.x - q.x, p.y - q.y); };
  switch (10 * a + b) {
    case 11: return B.G;
    case 12: return sc(B.Fr, -ux);
    case 13: return sc(B.Fr, -uy);
    case 21: return sc(B.Fr, ux);
    case 31: return sc(B.Fr, uy);
    case 41: return B.F1;
    case 14: return B.F2;
    case 22: return sub2(sc(sc(B.Fr, -uy * uy), 1.0 / r), sc(B.Frr, ux * ux));
    case 23:
    case 32: return sub2(sc(sc(B.Fr, -ux * uy), 1.0 / r), sc(B.Frr, ux * uy));
    case 33: return sub2(sc(sc(B.Fr, -ux * ux), 1.0 / r), sc(B.Frr, uy * uy));
    case 24: return sc(B.Fr2, ux);
    case 34: return sc(B.Fr2, uy);
    case 42: return sc(B.Fr1, -ux);
    case 43: return sc(B.Fr1, -uy);
    default: return B.F12;  // 44
  }
}

// Reflected part of the dipoles' potentials at the faces (MODE 0,
// @dipoleretlayer/potential.m:29-96, added to o0 = phi, o1 = phip, o2 = a,
// o3 = ap) or their reflected field at points (MODE 1, field.m:45-128 with
// 'refl', written to o2 = e, o3 = h): one thread per (observer, dipole
// position), the moments dip (npt x 3 x ndip) and dip / eps of the dipoles'
// medium (ieps = 1 / eps); column ipt + npt idip.  shift moves the dipoles
// along x (the self-interaction of greenderiv.m:19-21).
template <int MODE>
__global__ void __launch_bounds__(256) k_dip_layer(int64_t no, const double* __restrict__ obs,
                                                   const double* __restrict__ onv, int64_t npt,
                                                   const double* __restrict__ pts, double shift, int64_t ndip,
                                                   const double* __restrict__ dip, double2 ieps, double k0, TabView T,
                                                   double2* o0, double2* o1, double2* o2, double2* o3) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= no * npt) return;
  const int64_t i = t % no, j = t / no;
  const double x = obs[i] - (pts[j] + shift), y = obs[no + i] - pts[npt + j];
  const double z1 = obs[2 * no + i], z2 = pts[2 * npt + j];
  double r = sqrt(x * x + y * y);
  r = r < T.rmin ? T.rmin : r;
  const double ux = x / r, uy = y / r;
  const double2 mik = mk<double>(0, -k0);
  auto F = [&](const GdBase& B, int a, int b) { return gd_F(B, a, b, ux, uy, r); };
  auto dot3 = [](double2 a, double2 b, double2 c, double2 d2x, double2 d2y, double2 d2z) {
    return cadd<double>(cadd<double>(cmul<double>(a, d2x), cmul<double>(b, d2y)), cmul<double>(c, d2z));
  };
  auto rs = [](double2 z, double s) { return mk<double>(z.x * s, z.y * s); };
  if constexpr (MODE == 0) {
    const double n1 = onv[i], n2 = onv[no + i], n3 = onv[2 * no + i];
    auto D = [&](const GdBase& B, int b) {
      return cadd<double>(cadd<double>(rs(F(B, 2, b), n1), rs(F(B, 3, b), n2)), rs(F(B, 4, b), n3));
    };
    GdBase B = gd_base(T, r, z1, z2, 0);  // p
    const double2 Gp = B.G, Dp = D(B, 1);
    B = gd_base(T, r, z1, z2, 1);  // ss
    const double2 S2 = F(B, 1, 2), S3 = F(B, 1, 3), S4 = F(B, 1, 4), DS2 = D(B, 2), DS3 = D(B, 3), DS4 = D(B, 4);
    B = gd_base(T, r, z1, z2, 2);  // hs
    const double2 H2 = F(B, 1, 2), H3 = F(B, 1, 3), H4 = F(B, 1, 4), DH2 = D(B, 2), DH3 = D(B, 3), DH4 = D(B, 4);
    B = gd_base(T, r, z1, z2, 3);  // sh
    const double2 Gsh = B.G, Dsh = D(B, 1);
    B = gd_base(T, r, z1, z2, 4);  // hh
    const double2 Ghh = B.G, Dhh = D(B, 1);
    for (int64_t q = 0; q < ndip; ++q) {
      const double dx = dip[j + npt * (3 * q)], dy = dip[j + npt * (1 + 3 * q)], dz = dip[j + npt * (2 + 3 * q)];
      const double2 ex = rs(ieps, dx), ey = rs(ieps, dy), ez = rs(ieps, dz);
      const int64_t col = j + npt * q;
      const double2 phi = cadd<double>(dot3(S2, S3, S4, ex, ey, ez), cmul<double>(mik, rs(Gsh, dz)));
      const double2 phip = cadd<double>(dot3(DS2, DS3, DS4, ex, ey, ez), cmul<double>(mik, rs(Dsh, dz)));
      const double2 a[3] = {cmul<double>(mik, rs(Gp, dx)), cmul<double>(mik, rs(Gp, dy)),
                            cadd<double>(cmul<double>(mik, rs(Ghh, dz)), dot3(H2, H3, H4, ex, ey, ez))};
      // the hh term of a2p(:, 3) with the y component, potential.m:82
      const double2 ap[3] = {cmul<double>(mik, rs(Dp, dx)), cmul<double>(mik, rs(Dp, dy)),
                             cadd<double>(cmul<double>(mik, rs(Dhh, dy)), dot3(DH2, DH3, DH4, ex, ey, ez))};
      o0[i + no * col] = cadd<double>(o0[i + no * col], phi);
      o1[i + no * col] = cadd<double>(o1[i + no * col], phip);
      for (int c = 0; c < 3; ++c) {
        o2[i + no * (c + 3 * col)] = cadd<double>(o2[i + no * (c + 3 * col)], a[c]);
        o3[i + no * (c + 3 * col)] = cadd<double>(o3[i + no * (c + 3 * col)], ap[c]);
      }
    }
  } else {
    GdBase B = gd_base(T, r, z1, z2, 0);  // p
    const double2 Gp = B.G, P41 = F(B, 4, 1), P21 = F(B, 2, 1), P31 = F(B, 3, 1);
    B = gd_base(T, r, z1, z2, 1);  // ss
    const double2 S22 = F(B, 2, 2), S23 = F(B, 2, 3), S24 = F(B, 2, 4), S32 = F(B, 3, 2), S33 = F(B, 3, 3),
                  S34 = F(B, 3, 4), S42 = F(B, 4, 2), S43 = F(B, 4, 3), S44 = F(B, 4, 4);
    B = gd_base(T, r, z1, z2, 2);  // hs
    const double2 H12 = F(B, 1, 2), H13 = F(B, 1, 3), H14 = F(B, 1, 4), H32 = F(B, 3, 2), H33 = F(B, 3, 3),
                  H34 = F(B, 3, 4), H42 = F(B, 4, 2), H43 = F(B, 4, 3), H44 = F(B, 4, 4);
    B = gd_base(T, r, z1, z2, 3);  // sh
    const double2 Sh21 = F(B, 2, 1), Sh31 = F(B, 3, 1), Sh14 = F(B, 1, 4);
    B = gd_base(T, r, z1, z2, 4);  // hh
    const double2 Ghh = B.G, Hh31 = F(B, 3, 1), Hh41 = F(B, 4, 1);
    const double2 ik0 = mk<double>(0, k0);
    for (int64_t q = 0; q < ndip; ++q) {
      const double dx = dip[j + npt * (3 * q)], dy = dip[j + npt * (1 + 3 * q)], dz = dip[j + npt * (2 + 3 * q)];
      const double2 ex = rs(ieps, dx), ey = rs(ieps, dy), ez = rs(ieps, dz);
      const int64_t col = j + npt * q;
      const double2 a1 = cmul<double>(mik, rs(Gp, dx)), a2 = cmul<double>(mik, rs(Gp, dy));
      const double2 a3 = cadd<double>(cmul<double>(mik, rs(Ghh, dz)), dot3(H12, H13, H14, ex, ey, ez));
      const double2 a23 = cadd<double>(cmul<double>(mik, rs(Hh31, dz)), dot3(H32, H33, H34, ex, ey, ez));
      const double2 a32 = cmul<double>(mik, rs(P41, dy)), a31 = cmul<double>(mik, rs(P41, dx));
      const double2 a13 = cadd<double>(cmul<double>(mik, rs(Hh41, dz)), dot3(H42, H43, H44, ex, ey, ez));
      const double2 a12 = cmul<double>(mik, rs(P21, dy)), a21 = cmul<double>(mik, rs(P31, dx));
      const double2 phi1 = cadd<double>(dot3(S22, S23, S24, ex, ey, ez), cmul<double>(mik, rs(Sh21, dz)));
      const double2 phi2 = cadd<double>(dot3(S32, S33, S34, ex, ey, ez), cmul<double>(mik, rs(Sh31, dz)));
      const double2 phi3 = cadd<double>(dot3(S42, S43, S44, ex, ey, ez), cmul<double>(mik, rs(Sh14, dz)));
      const double2 e[3] = {csub<double>(cmul<double>(ik0, a1), phi1), csub<double>(cmul<double>(ik0, a2), phi2),
                            csub<double>(cmul<double>(ik0, a3), phi3)};
      const double2 h[3] = {csub<double>(a32, a23), csub<double>(a13, a31), csub<double>(a21, a12)};
      for (int c = 0; c < 3; ++c) {
        o2[i + no * (c + 3 * col)] = e[c];
        o3[i + no * (c + 3 * col)] = h[c];
      }
    }
  }
}

// Far field of dipoles above the interface in the directions dir (device
// ndir x 3), @dipoleretlayer/farfield.m:21-177: one thread per (direction,
// dipole position).  Reflected (upward) or transmitted (downward) part,
// only through a lossless medium, with the gradient at the dipole by
// forward differences (eta 1e-6); the dipole's own far field in the upper
// medium upward (dipoleret/farfield.m:40-55); transverse part of e and
// h = k / k0 dir x e.  e, h device out ndir x 3 x (npt ndip).
__global__ void __launch_bounds__(256) k_dip_layer_ff(int64_t ndir, const double* __restrict__ dir, int64_t npt,
                                                      const double* __restrict__ pts, int64_t ndip,
                                                      const double* __restrict__ dip, double k0, double2 e1,
                                                      double2 e2, double z0, double2* e, double2* h) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ndir * npt) return;
  const int64_t d = t % ndir, j = t / ndir;
  const double dv[3] = {dir[d], dir[ndir + d], dir[2 * ndir + d]};
  const bool up = dv[2] >= 0;
  const double2 k1 = cscale<double>(csqrt_p(e1), k0), k2 = cscale<double>(csqrt_p(e2), k0);
  const double2 km = up ? k1 : k2;
  const double eta = 1e-6;
  const double px = pts[j], py = pts[npt + j], pz = pts[2 * npt + j];
  double2 rp[4], rh[4], rs_[4];
  const bool refl = km.y == 0;
  if (refl) {
    const double2 kp = mk<double>(km.x * sqrt(1 - dv[2] * dv[2]), 0);
    const double2 k1z = kz_of(k1, kp), k2z = kz_of(k2, kp);
    const Coef4 c = coeff_entry(k0, e1, e2, k1z, k2z, up ? 1 : 2, 1);
    for (int m = 0; m < 4; ++m) {
      const double x = px + (m == 1 ? eta : 0), y = py + (m == 2 ? eta : 0), z = pz + (m == 3 ? eta : 0);
      const double2 g2 = cexp_(mul_i(cscale<double>(k1z, fabs(z - z0))));
      const double dist = x * dv[0] + y * dv[1] + z0 * dv[2];
      const double2 ph = cmul<double>(g2, cexp_(mul_i(cscale<double>(km, -dist))));
      rp[m] = cmul<double>(c.p, ph);
      rh[m] = cmul<double>(c.hh, ph);
      rs_[m] = cmul<double>(c.hs, ph);
    }
    for (int m = 1; m < 4; ++m) rs_[m] = fdiff(rs_[m], rs_[0], eta);
  }
  const double2 ie1 = cdiv<double>(mk<double>(1, 0), e1);
  const double2 mik = mk<double>(0, -k0), ik0 = mk<double>(0, k0);
  for (int64_t q = 0; q < ndip; ++q) {
    const double dm[3] = {dip[j + npt * (3 * q)], dip[j + npt * (1 + 3 * q)], dip[j + npt * (2 + 3 * q)]};
    double2 ev[3] = {mk<double>(0, 0), mk<double>(0, 0), mk<double>(0, 0)};
    if (refl) {
      const double2 a0 = cmul<double>(mik, cscale<double>(rp[0], dm[0]));
      const double2 a1 = cmul<double>(mik, cscale<double>(rp[0], dm[1]));
      double2 a2 = cmul<double>(mik, cscale<double>(rh[0], dm[2]));
      for (int m = 1; m < 4; ++m) a2 = cadd<double>(a2, cmul<double>(rs_[m], cscale<double>(ie1, dm[m - 1])));
      ev[0] = cmul<double>(ik0, a0);
      ev[1] = cmul<double>(ik0, a1);
      ev[2] = cmul<double>(ik0, a2);
    }
    if (up) {  // k^2 ((dir x p) x dir) exp(-i k dir . pos) / eps
      const double2 g = cexp_ikd(mk<double>(-k1.x, -k1.y), dv[0] * px + dv[1] * py + dv[2] * pz);
      const double hx = dv[1] * dm[2] - dv[2] * dm[1], hy = dv[2] * dm[0] - dv[0] * dm[2],
                   hz = dv[0] * dm[1] - dv[1] * dm[0];
      const double ee[3] = {hy * dv[2] - hz * dv[1], hz * dv[0] - hx * dv[2], hx * dv[1] - hy * dv[0]};
      const double2 f = cmul<double>(cmul<double>(cmul<double>(k1, k1), g), ie1);
      for (int c = 0; c < 3; ++c) ev[c] = cadd<double>(ev[c], cscale<double>(f, ee[c]));
    }
    const double2 de =
        cadd<double>(cadd<double>(cscale<double>(ev[0], dv[0]), cscale<double>(ev[1], dv[1])), cscale<double>(ev[2], dv[2]));
    for (int c = 0; c < 3; ++c) ev[c] = csub<double>(ev[c], cscale<double>(de, dv[c]));
    const double2 kk = cdiv<double>(km, mk<double>(k0, 0));
    const double2 hv[3] = {csub<double>(cscale<double>(ev[2], dv[1]), cscale<double>(ev[1], dv[2])),
                           csub<double>(cscale<double>(ev[0], dv[2]), cscale<double>(ev[2], dv[0])),
                           csub<double>(cscale<double>(ev[1], dv[0]), cscale<double>(ev[0], dv[1]))};
    const int64_t col = j + npt * q;
    for (int c = 0; c < 3; ++c) {
      e[d + ndir * (c + 3 * col)] = ev[c];
      h[d + ndir * (c + 3 * col)] = cmul<double>(kk, hv[c]);
    }
  }
}

// ------------------------------------------------------ chain helpers
// (templated on the chain's precision T; scalars arrive as double2)
template <class T>
__host__ __device__ inline cplx<T> tc(double2 z) {
  return mk<T>((T)z.x, (T)z.y);
}
// out = alpha x + beta y
template <class T>
__global__ void k_zaxpby(int64_t len, double2 al_, const cplx<T>* x, double2 be_, const cplx<T>* y, cplx<T>* out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const cplx<T> al = tc<T>(al_), be = tc<T>(be_);
  if (i < len) out[i] = cadd<T>(cmul<T>(al, x[i]), cmul<T>(be, y[i]));
}
// W = e1 S1 + c (Gamma .* npar npar')
template <class T>
__global__ void k_lay_w(int64_t n, const double* nvec, double2 e1_, const cplx<T>* S1, double2 c_, const cplx<T>* Gam,
                        cplx<T>* W) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int64_t i = idx % n, j = idx / n;
  const T nn = (T)(nvec[i] * nvec[j] + nvec[n + i] * nvec[n + j]);
  W[idx] = cadd<T>(cmul<T>(tc<T>(e1_), S1[idx]), cscale<T>(cmul<T>(tc<T>(c_), Gam[idx]), nn));
}
// block of M (ld 2n): Mb(i, j) += -al H(i, j) - be nz_i Gc(i, j)
template <class T>
__global__ void k_lay_mblock(int64_t n, int64_t ld, cplx<T>* Mb, double2 al_, const cplx<T>* H, double2 be_,
                             const double* nz, const cplx<T>* Gc) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int64_t i = idx % n, j = idx / n;
  cplx<T>& m = Mb[j * ld + i];
  m = csub<T>(csub<T>(m, cmul<T>(tc<T>(al_), H[idx])), cscale<T>(cmul<T>(tc<T>(be_), Gc[idx]), (T)nz[i]));
}
// excitation.m:17-29 with scalar eps1 (inside) and eps2 (outside), FP64
__global__ void k_lay_exc(int64_t n, int64_t p, const double2* phi1, const double2* phi1p, const double2* a1,
                          const double2* a1p, const double2* phi2, const double2* phi2p, const double2* a2,
                          const double2* a2p, const double* nvec, double k, double2 e1, double2 e2, double2* phi,
                          double2* a, double2* alpha, double2* De) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * p) return;
  const int64_t i = t % n, q = t / n;
  auto ld = [&](const double2* v, int64_t idx) { return v ? v[idx] : mk<double>(0, 0); };
  const double2 f1 = ld(phi1, t), f2 = ld(phi2, t);
  phi[t] = csub<double>(f2, f1);
  const double2 ik = mk<double>(0, k);
  double2 na1 = mk<double>(0, 0), na2 = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) {
    const int64_t id = (c + 3 * q) * n + i;
    const double nc = nvec[c * n + i];
    const double2 A1 = ld(a1, id), A2 = ld(a2, id);
    a[id] = csub<double>(A2, A1);
    na1 = cadd<double>(na1, cscale<double>(A1, nc));
    na2 = cadd<double>(na2, cscale<double>(A2, nc));
    const double2 s = csub<double>(cmul<double>(f2, e2), cmul<double>(f1, e1));
    alpha[id] = csub<double>(csub<double>(ld(a2p, id), ld(a1p, id)), cscale<double>(cmul<double>(ik, s), nc));
  }
  De[t] = csub<double>(csub<double>(cmul<double>(e2, ld(phi2p, t)), cmul<double>(e1, ld(phi1p, t))),
                       cmul<double>(ik, csub<double>(cmul<double>(e2, na2), cmul<double>(e1, na1))));
}
// alpha <- alpha - T + c n phi  (T = Sigma1 a)
template <class T>
__global__ void k_lay_alpha(int64_t n, int64_t p, cplx<T>* alpha, const cplx<T>* Tm, const cplx<T>* phi,
                            const double* nvec, double2 c_) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * p) return;
  const int64_t i = t % n, q = t / n;
  const cplx<T> cp = cmul<T>(tc<T>(c_), phi[t]);
  for (int cc = 0; cc < 3; ++cc) {
    const int64_t id = (cc + 3 * q) * n + i;
    alpha[id] = cadd<T>(csub<T>(alpha[id], Tm[id]), cscale<T>(cp, (T)nvec[cc * n + i]));
  }
}
// rhs = [De - e1 T2 + c1 n.a + c2 npar.T3 ; alpha_z]  (2n x p, ld 2n)
template <class T>
__global__ void k_lay_rhs(int64_t n, int64_t p, const cplx<T>* De, const cplx<T>* T2, const cplx<T>* a,
                          const cplx<T>* T3, const cplx<T>* alpha, const double* nvec, double2 e1_, double2 c1_,
                          double2 c2_, cplx<T>* rhs) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * p) return;
  const int64_t i = t % n, q = t / n;
  cplx<T> na = mk<T>(0, 0), nt = mk<T>(0, 0);
  for (int cc = 0; cc < 3; ++cc) {
    const int64_t id = (cc + 3 * q) * n + i;
    na = cadd<T>(na, cscale<T>(a[id], (T)nvec[cc * n + i]));
    if (cc < 2) nt = cadd<T>(nt, cscale<T>(T3[id], (T)nvec[cc * n + i]));
  }
  rhs[q * 2 * n + i] = cadd<T>(cadd<T>(csub<T>(De[t], cmul<T>(tc<T>(e1_), T2[t])), cmul<T>(tc<T>(c1_), na)),
                               cmul<T>(tc<T>(c2_), nt));
  rhs[q * 2 * n + n + i] = alpha[(2 + 3 * q) * n + i];
}
// V = alpha_par + c npar U, V_z = 0
template <class T>
__global__ void k_lay_v(int64_t n, int64_t p, const cplx<T>* alpha, const cplx<T>* U, const double* nvec, double2 c_,
                        cplx<T>* V) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * p) return;
  const int64_t i = t % n, q = t / n;
  const cplx<T> cu = cmul<T>(tc<T>(c_), U[t]);
  for (int cc = 0; cc < 2; ++cc) {
    const int64_t id = (cc + 3 * q) * n + i;
    V[id] = cadd<T>(alpha[id], cscale<T>(cu, (T)nvec[cc * n + i]));
  }
  V[(2 + 3 * q) * n + i] = mk<T>(0, 0);
}
// h (n x 3 x p): z-component from src (column q at src + q ld)
template <class T>
__global__ void k_lay_setz(int64_t n, int64_t p, const cplx<T>* src, int64_t ld, cplx<T>* h) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * p) return;
  const int64_t i = t % n, q = t / n;
  h[(2 + 3 * q) * n + i] = src[q * ld + i];
}
// sig2 (n x p) from the first n rows of xi (ld 2n)
template <class T>
__global__ void k_lay_copy_rows(int64_t n, int64_t p, const cplx<T>* src, int64_t ld, cplx<T>* dst) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * p) return;
  dst[t] = src[(t / n) * ld + t % n];
}
// Residual r = b - A x of the operator the chain inverts, FP64, from the
// products: Gs1 = G1 s1, Gh1 = G1 h1, Hs1 = H1 s1, Hh1 = H1 h1, A2 = [G2p h2x,
// G2p h2y, G2hh h2z] with X = G2hs s2 still to add to its z, Y = G2sh s2,
// phi2 = G2ss s2 + G2sh h2z, dA2 = [H2p h2x, H2p h2y, H2hs s2 + H2hh h2z],
// dphi2 = H2ss s2 + H2sh h2z.  b and r hold (phi, a, alpha, De) blocks.
struct LayProd {
  const double2 *Gs1, *Gh1, *Hs1, *Hh1, *A2, *X, *Y, *phi2, *dA2, *dphi2;
};
__global__ void k_lay_resid(int64_t n, int64_t p, LayProd P, const double2* b, double2* r, const double* nvec,
                            double k, double2 e1, double2 e2) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * p) return;
  const int64_t i = t % n, q = t / n;
  const int64_t np1 = n * p, np3 = 3 * n * p;
  const double2 ik = mk<double>(0, k);
  const double2 *bphi = b, *ba = b + np1, *balpha = ba + np3, *bDe = balpha + np3;
  double2 *rphi = r, *ra = r + np1, *ralpha = ra + np3, *rDe = ralpha + np3;
  const double2 gs1 = P.Gs1[t], ph2 = P.phi2[t];
  rphi[t] = csub<double>(bphi[t], csub<double>(gs1, ph2));
  const double2 s = csub<double>(cmul<double>(e1, gs1), cmul<double>(e2, ph2));
  double2 nin = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) {
    const int64_t id = (c + 3 * q) * n + i;
    const double nc = nvec[c * n + i];
    const double2 A = c < 2 ? P.A2[id] : cadd<double>(P.A2[id], P.X[t]);
    ra[id] = csub<double>(ba[id], csub<double>(P.Gh1[id], A));
    const double2 al = csub<double>(csub<double>(P.Hh1[id], P.dA2[id]), cscale<double>(cmul<double>(ik, s), nc));
    ralpha[id] = csub<double>(balpha[id], al);
    nin = cadd<double>(nin, cscale<double>(csub<double>(cmul<double>(e1, P.Gh1[id]), cmul<double>(e2, A)), nc));
  }
  const double2 corr = cscale<double>(cmul<double>(cmul<double>(ik, csub<double>(e1, e2)), csub<double>(P.X[t], P.Y[t])),
                                      nvec[2 * n + i]);
  const double2 De = cadd<double>(csub<double>(csub<double>(cmul<double>(e1, P.Hs1[t]), cmul<double>(e2, P.dphi2[t])),
                                               cmul<double>(ik, nin)),
                                  corr);
  rDe[t] = csub<double>(bDe[t], De);
}

inline unsigned nblk(int64_t len, int bs = 256) { return (unsigned)cdiv_up(len, bs); }

}  // namespace

// ------------------------------------------------------ host objects
struct LayerTab {
  mnp_layer_options op;
  int64_t nr = 0, nz = 0;
  double z2 = 0;
  double *r = nullptr, *z1 = nullptr, *Zt = nullptr;
  double2* cur = nullptr;                 // table at the current wavelength
  double2* saved = nullptr;               // tables of set(), nsaved x size
  std::vector<double> enei;               // wavelengths of the saved tables
  int* err = nullptr;
  double cur_enei = -1;
  double2 cur_e1{}, cur_e2{};

  int64_t size() const { return nr * nz * NQ; }
  LayerTab(const mnp_layer_options& o, int64_t nr_, const double* r_, int64_t nz_, const double* z1_, double z2_)
      : op(o), nr(nr_), nz(nz_), z2(z2_) {
    MNP_CHECK(nr >= 2 && nz >= 2, MNP_ERR_ARG, "layertab: the grid needs at least 2 x 2 points");
    for (int64_t i = 0; i + 1 < nr; ++i)
      MNP_CHECK(r_[i] < r_[i + 1] && r_[0] > 0, MNP_ERR_ARG, "layertab: radii must be positive and increasing");
    for (int64_t j = 0; j + 1 < nz; ++j)
      MNP_CHECK(z1_[j] < z1_[j + 1], MNP_ERR_ARG, "layertab: heights must be increasing");
    MNP_CHECK(z1_[0] > op.z0 && z2 > op.z0, MNP_ERR_ARG, "layertab: the table must lie above the interface");
    upload_rules();
    std::vector<double> zt(nz);
    for (int64_t j = 0; j < nz; ++j) zt[j] = std::fabs(z1_[j] - op.z0) + std::fabs(z2 - op.z0);
    r = dev_alloc<double>(nr);
    z1 = dev_alloc<double>(nz);
    Zt = dev_alloc<double>(nz);
    cur = dev_alloc<double2>(size());
    err = dev_alloc<int>(1);
    MNP_HIP(hipMemcpy(r, r_, sizeof(double) * nr, hipMemcpyHostToDevice));
    MNP_HIP(hipMemcpy(z1, z1_, sizeof(double) * nz, hipMemcpyHostToDevice));
    MNP_HIP(hipMemcpy(Zt, zt.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
  }
  ~LayerTab() {
    void* ps[] = {r, z1, Zt, cur, saved, err};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
  // green.m for the whole grid at one wavelength into out
  void compute(double enei_, double2 e1, double2 e2, double2* out) const {
    LayPrm P;
    P.k0 = 2 * M_PI / enei_;
    P.e1 = e1;
    P.e2 = e2;
    P.k1 = wavenumber(P.k0, e1);
    P.k2 = wavenumber(P.k0, e2);
    P.k1max = std::max(P.k1.x, P.k2.x) + P.k0;
    P.semi = op.semi;
    P.ratio = op.ratio;
    hipLaunchKernelGGL(k_layer_table, dim3((unsigned)(nr * nz)), dim3(64), 0, stream(), nr, nz, r, Zt, P, out);
    MNP_HIP(hipGetLastError());
  }
  void set(int64_t ne, const double* en, const mnp_complex* e1, const mnp_complex* e2) {
    MNP_CHECK(ne >= 1, MNP_ERR_ARG, "layertab set: no wavelengths");
    for (int64_t i = 0; i + 1 < ne; ++i)
      MNP_CHECK(en[i] < en[i + 1], MNP_ERR_ARG, "layertab set: wavelengths must be increasing");
    if (saved) MNP_HIP(hipFree(saved));
    saved = dev_alloc<double2>(ne * size());
    for (int64_t i = 0; i < ne; ++i) compute(en[i], to_d2(e1[i]), to_d2(e2[i]), saved + i * size());
    enei.assign(en, en + ne);
    cur_enei = -1;
  }
  // eval.m:17-50: the table at enei
  const double2* at(double enei_, double2 e1, double2 e2) {
    if (enei_ == cur_enei && e1.x == cur_e1.x && e1.y == cur_e1.y && e2.x == cur_e2.x && e2.y == cur_e2.y) return cur;
    if (enei.empty()) {
      compute(enei_, e1, e2, cur);
    } else {
      MNP_CHECK(enei_ >= enei.front() && enei_ <= enei.back(), MNP_ERR_ARG,
                "greentablayer: wavelength outside the tabulated range");
      if (enei.size() == 1) {
        MNP_HIP(hipMemcpyAsync(cur, saved, sizeof(double2) * size(), hipMemcpyDeviceToDevice, stream()));
      } else {
        size_t i = std::upper_bound(enei.begin(), enei.end(), enei_) - enei.begin();
        i = std::min(i, enei.size() - 1) - 1;
        const double t = (enei_ - enei[i]) / (enei[i + 1] - enei[i]);
        hipLaunchKernelGGL(k_lerp, dim3(nblk(size())), dim3(256), 0, stream(), size(), saved + i * size(),
                           saved + (i + 1) * size(), t, cur);
        MNP_HIP(hipGetLastError());
      }
    }
    cur_enei = enei_;
    cur_e1 = e1;
    cur_e2 = e2;
    return cur;
  }
  TabView view(const double2* v) const {
    TabView T;
    T.nr = nr;
    T.nz = nz;
    T.r = r;
    T.z1 = z1;
    T.v = v;
    T.z0 = op.z0;
    T.rmin = op.rmin;
    T.zmin = op.zmin;
    T.err = err;
    return T;
  }
};

struct GreenRetLayer {
  const GreenRet* g;
  LayerTab* tab;
  int64_t np = 0, nd = 0;
  int32_t *rows = nullptr, *cols = nullptr, *id = nullptr;
  double *W = nullptr, *R0 = nullptr, *Z2 = nullptr, *Z1 = nullptr;

  GreenRetLayer(const GreenRet* g_, LayerTab* t_) : g(g_), tab(t_) {
    MNP_CHECK(!g->compound(), MNP_ERR_ARG, "greenretlayer: compound particles are not supported");
    const int64_t n = g->n;
    std::vector<double> pos(3 * n), rad(n);
    MNP_HIP(hipMemcpy(pos.data(), g->geo.pos, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    MNP_HIP(hipMemcpy(rad.data(), g->geo.rad, sizeof(double) * n, hipMemcpyDeviceToHost));
    const double z0 = tab->op.z0;
    for (int64_t i = 0; i < n; ++i)
      MNP_CHECK(pos[2 * n + i] - z0 >= tab->op.ztol, MNP_ERR_ARG,
                "greenretlayer: every face must lie above the interface (faces in it are not supported)");
    // refinematrixlayer.m:24-56 (column-major order of find)
    std::vector<int32_t> hr, hc, hd;
    std::vector<double> a2(n);
    for (int64_t i = 0; i < n; ++i) a2[i] = pos[i] * pos[i] + pos[n + i] * pos[n + i];
    const double rel = g->op.rel_cutoff, absc = g->op.abs_cutoff;
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < n; ++i) {
        double r2 = a2[i] + a2[j] - 2 * (pos[i] * pos[j] + pos[n + i] * pos[n + j]);
        if (r2 < 1e-10) r2 = 0;
        const double z = std::fabs(pos[2 * n + i] - z0) + std::fabs(pos[2 * n + j] - z0);
        const double d = std::sqrt(r2 + z * z);
        const double d2 = d - rad[j];
        if (!(d2 < absc || d2 / rad[i] < rel)) continue;
        if (i == j) {
          hd.push_back((int32_t)i);
        } else {
          hr.push_back((int32_t)i);
          hc.push_back((int32_t)j);
        }
      }
    np = (int64_t)hr.size();
    nd = (int64_t)hd.size();
    rows = dev_alloc<int32_t>(np);
    cols = dev_alloc<int32_t>(np);
    id = dev_alloc<int32_t>(nd);
    W = dev_alloc<double>(12 * np);
    R0 = dev_alloc<double>(3 * np);
    Z2 = dev_alloc<double>(3 * np);
    Z1 = dev_alloc<double>(np);
    if (np) {
      MNP_HIP(hipMemcpy(rows, hr.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice));
      MNP_HIP(hipMemcpy(cols, hc.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice));
      LayGeo L{z0, tab->op.rmin, tab->op.zmin};
      hipLaunchKernelGGL(k_layer_refine_init, dim3(nblk(np)), dim3(256), 0, stream(), np, n, g->geo, L, g->geo.pos, n,
                         rows, cols, g->q18, W, R0, Z2, Z1);
      MNP_HIP(hipGetLastError());
    }
    if (nd) MNP_HIP(hipMemcpy(id, hd.data(), sizeof(int32_t) * nd, hipMemcpyHostToDevice));
    // column offsets of the pairs (found column by column, rows increasing)
    // and the diagonal element of each face
    std::vector<int64_t> co(n + 1, 0);
    for (int64_t p = 0; p < np; ++p) ++co[hc[p] + 1];
    for (int64_t j = 0; j < n; ++j) co[j + 1] += co[j];
    std::vector<int32_t> dm(n, -1);
    for (int64_t t = 0; t < nd; ++t) dm[hd[t]] = (int32_t)t;
    coff = dev_alloc<int64_t>(n + 1);
    dmap = dev_alloc<int32_t>(n);
    MNP_HIP(hipMemcpy(coff, co.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
    MNP_HIP(hipMemcpy(dmap, dm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    MNP_HIP(hipStreamSynchronize(stream()));
  }
  int64_t* coff = nullptr;
  int32_t* dmap = nullptr;
  // Refined values of the pairs and diagonal elements alone (component-major,
  // 5 x np and 5 x nd), with the table at enei: the near-field part of the
  // reflected matrices for H-matrix storage.
  double2 *pG = nullptr, *pF = nullptr, *dG = nullptr, *dF = nullptr;
  LayerPairs pairs(double enei, double2 e1, double2 e2) {
    const int64_t n = g->n;
    const double2* tv = tab->at(enei, e1, e2);
    const TabView T = tab->view(tv);
    if (!pG) {
      pG = dev_alloc<double2>(5 * std::max<int64_t>(np, 1));
      pF = dev_alloc<double2>(5 * std::max<int64_t>(np, 1));
      dG = dev_alloc<double2>(5 * std::max<int64_t>(nd, 1));
      dF = dev_alloc<double2>(5 * std::max<int64_t>(nd, 1));
    }
    MNP_HIP(hipMemsetAsync(tab->err, 0, sizeof(int), stream()));
    if (np) {
      hipLaunchKernelGGL(k_layer_refine_apply, dim3(nblk(np)), dim3(256), 0, stream(), np, n, g->geo, T, rows, cols,
                         W, R0, Z2, Z1, pG, pF, 1);
      MNP_HIP(hipGetLastError());
    }
    if (nd) {
      hipLaunchKernelGGL(k_layer_diag, dim3(nblk(nd, 64)), dim3(64), 0, stream(), nd, n, g->geo, T, id, g->qpol,
                         g->npolq, dG, dF, 1);
      MNP_HIP(hipGetLastError());
    }
    return LayerPairs{T, np, nd, rows, cols, id, pG, pF, dG, dF, coff, dmap};
  }
  void check() {
    int herr = 0;
    MNP_HIP(hipMemcpyAsync(&herr, tab->err, sizeof(int), hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_CHECK(herr == 0, MNP_ERR_ARG, "greentablayer: a face pair lies outside the table (tabspace of this particle?)");
  }
  ~GreenRetLayer() {
    void* ps[] = {rows, cols, id, W, R0, Z2, Z1, pG, pF, dG, dF, coff, dmap};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
  // Reflected G and F (5 x n x n each) at enei.
  void eval(double enei, double2 e1, double2 e2, double2* G, double2* F) {
    const int64_t n = g->n;
    const double2* tv = tab->at(enei, e1, e2);
    const TabView T = tab->view(tv);
    MNP_HIP(hipMemsetAsync(tab->err, 0, sizeof(int), stream()));
    hipLaunchKernelGGL(k_layer_assemble, dim3(nblk(n * n)), dim3(256), 0, stream(), n, g->geo, T, G, F);
    MNP_HIP(hipGetLastError());
    if (np) {
      hipLaunchKernelGGL(k_layer_refine_apply, dim3(nblk(np)), dim3(256), 0, stream(), np, n, g->geo, T, rows, cols,
                         W, R0, Z2, Z1, G, F, 0);
      MNP_HIP(hipGetLastError());
    }
    if (nd) {
      hipLaunchKernelGGL(k_layer_diag, dim3(nblk(nd, 64)), dim3(64), 0, stream(), nd, n, g->geo, T, id, g->qpol,
                         g->npolq, G, F, 0);
      MNP_HIP(hipGetLastError());
    }
    int herr = 0;
    MNP_HIP(hipMemcpyAsync(&herr, tab->err, sizeof(int), hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_CHECK(herr == 0, MNP_ERR_ARG, "greentablayer: a face pair lies outside the table (tabspace of this particle?)");
  }
};

LayerPairs greenretlayer_pairs(GreenRetLayer* gl, double enei, double2 e1, double2 e2) {
  return gl->pairs(enei, e1, e2);
}
void greenretlayer_check(GreenRetLayer* gl) { gl->check(); }

// Reflected Green function between points above the interface and a
// particle (greenretlayer with p1 a compoint): the near pairs of
// refinematrixlayer.m (divided by the face radius, as a compoint has none)
// with their shape-function weights, per point in face order.
struct PointLayer {
  const GreenRet* g;
  LayerTab* tab;
  int64_t npt = 0, np = 0;
  double* pts = nullptr;  // npt x 3 column-major
  int64_t* rowptr = nullptr;
  int32_t *rows = nullptr, *cols = nullptr;
  double *W = nullptr, *R0 = nullptr, *Z2 = nullptr, *Z1 = nullptr;

  PointLayer(const GreenRet* g_, LayerTab* t_, int64_t npt_, const double* pts_h) : g(g_), tab(t_), npt(npt_) {
    MNP_CHECK(!g->compound(), MNP_ERR_ARG, "compgreenretlayer: compound particles are not supported");
    MNP_CHECK(npt >= 1, MNP_ERR_ARG, "compgreenretlayer: no points");
    const int64_t n = g->n;
    const double z0 = tab->op.z0;
    for (int64_t i = 0; i < npt; ++i)
      MNP_CHECK(pts_h[2 * npt + i] > z0, MNP_ERR_ARG, "compgreenretlayer: every point must lie above the interface");
    std::vector<double> pos(3 * n), rad(n);
    MNP_HIP(hipMemcpy(pos.data(), g->geo.pos, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    MNP_HIP(hipMemcpy(rad.data(), g->geo.rad, sizeof(double) * n, hipMemcpyDeviceToHost));
    std::vector<double> b2(n);
    for (int64_t j = 0; j < n; ++j) b2[j] = pos[j] * pos[j] + pos[n + j] * pos[n + j];
    const double rel = g->op.rel_cutoff, absc = g->op.abs_cutoff;
    std::vector<int64_t> rp(npt + 1, 0);
    std::vector<int32_t> hr, hc;
    for (int64_t i = 0; i < npt; ++i) {
      const double xi = pts_h[i], yi = pts_h[npt + i], zi = std::fabs(pts_h[2 * npt + i] - z0);
      const double a2 = xi * xi + yi * yi;
      for (int64_t j = 0; j < n; ++j) {
        double r2 = a2 + b2[j] - 2 * (xi * pos[j] + yi * pos[n + j]);
        if (r2 < 1e-10) r2 = 0;
        const double z = zi + std::fabs(pos[2 * n + j] - z0);
        const double d2 = std::sqrt(r2 + z * z) - rad[j];
        if (d2 < absc || d2 / rad[j] < rel) {
          hr.push_back((int32_t)i);
          hc.push_back((int32_t)j);
        }
      }
      rp[i + 1] = (int64_t)hr.size();
    }
    np = (int64_t)hr.size();
    pts = dev_alloc<double>(3 * npt);
    rowptr = dev_alloc<int64_t>(npt + 1);
    rows = dev_alloc<int32_t>(np);
    cols = dev_alloc<int32_t>(np);
    W = dev_alloc<double>(12 * np);
    R0 = dev_alloc<double>(3 * np);
    Z2 = dev_alloc<double>(3 * np);
    Z1 = dev_alloc<double>(np);
    MNP_HIP(hipMemcpy(pts, pts_h, sizeof(double) * 3 * npt, hipMemcpyHostToDevice));
    MNP_HIP(hipMemcpy(rowptr, rp.data(), sizeof(int64_t) * (npt + 1), hipMemcpyHostToDevice));
    if (np) {
      MNP_HIP(hipMemcpy(rows, hr.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice));
      MNP_HIP(hipMemcpy(cols, hc.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice));
      LayGeo L{z0, tab->op.rmin, tab->op.zmin};
      hipLaunchKernelGGL(k_layer_refine_init, dim3(nblk(np)), dim3(256), 0, stream(), np, n, g->geo, L, pts, npt, rows,
                         cols, g->q18, W, R0, Z2, Z1);
      MNP_HIP(hipGetLastError());
    }
    MNP_HIP(hipStreamSynchronize(stream()));
  }
  ~PointLayer() {
    void* ps[] = {pts, rowptr, rows, cols, W, R0, Z2, Z1};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
  void fields(double enei, double2 e1, double2 e2, int64_t ncol, const double2* sig2, const double2* h2, double2* phi,
              double2* a, double2* e, double2* h) {
    const double2* tv = tab->at(enei, e1, e2);
    const TabView T = tab->view(tv);
    MNP_HIP(hipMemsetAsync(tab->err, 0, sizeof(int), stream()));
    hipLaunchKernelGGL(k_pt_layer, dim3((unsigned)(npt * ncol)), dim3(256), 0, stream(), npt, ncol, pts, g->n, g->geo, T,
                       rowptr, cols, W, R0, Z2, Z1, 2 * M_PI / enei, sig2, h2, phi, a, e, h);
    MNP_HIP(hipGetLastError());
    int herr = 0;
    MNP_HIP(hipMemcpyAsync(&herr, tab->err, sizeof(int), hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_CHECK(herr == 0, MNP_ERR_ARG, "greentablayer: a point-face pair lies outside the table (tabspace with the points?)");
  }
};

// The bemretlayer chain (initmat.m:17-95, mldivide.m:31-76) in precision T
// for unique eps1 (inside) and eps2 (outside = upper medium): L1 = eps1,
// L2p = eps2, so Gammapar = i k (eps1 - eps2) Gamma .* (npar npar').  G1 and
// G2.p stay factorised (their inverses are applied by triangular solves),
// Gamma is formed explicitly for its elementwise product; each LU first runs
// without row exchanges and is redone with partial pivoting when that would
// have exchanged a row, so the factors are partial pivoting's.
template <class T>
struct LayerChain {
  using C = cplx<T>;
  struct Fac {
    int32_t* sig = nullptr;
    C* inv = nullptr;
    LU<T> f{};
  };
  int64_t n;
  const double* nvec;
  C *G1, *F2, *S1, *Gam, *Tm, *M, *Mb;
  C* G2c;            // the five G2 components in T (the owner's FP64 ones when T is double)
  bool own_g2 = false;
  int32_t* piv = nullptr;
  Fac f1, f2, fg, fm;
  Workspace ws, wsv;  // factorisations; the vectors of apply
  double k = 0;
  double2 e1{}, e2{};
  int fallbacks = 0;  // factorisations redone with partial pivoting

  LayerChain(int64_t n_, const double* nvec_) : n(n_), nvec(nvec_) {
    const int64_t nn = n * n;
    G1 = dev_alloc<C>(nn);
    F2 = dev_alloc<C>(nn);
    S1 = dev_alloc<C>(nn);
    Gam = dev_alloc<C>(nn);
    Tm = dev_alloc<C>(2 * nn);
    M = dev_alloc<C>(4 * nn);
    Mb = dev_alloc<C>(4 * nn);
    piv = dev_alloc<int32_t>(2 * n);
    for (Fac* f : {&f1, &f2, &fg}) {
      f->sig = dev_alloc<int32_t>(n);
      f->inv = dev_alloc<C>(2 * inv_blocks(n));
    }
    fm.sig = dev_alloc<int32_t>(2 * n);
    fm.inv = dev_alloc<C>(2 * inv_blocks(2 * n));
    if (!std::is_same<T, double>::value) {
      G2c = dev_alloc<C>(5 * nn);
      own_g2 = true;
    }
  }
  ~LayerChain() {
    void* ps[] = {G1, F2, S1, Gam, Tm, M, Mb, piv, f1.sig, f1.inv, f2.sig, f2.inv, fg.sig, fg.inv, fm.sig, fm.inv,
                  own_g2 ? (void*)G2c : nullptr};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
  C* G2(int c) const { return G2c + c * n * n; }
  MatView<T> v(C* p, int64_t ld = -1) const { return MatView<T>{p, ld < 0 ? n : ld}; }
  // dst (T) = src (FP64), m x m
  void load(const double2* src, C* dst, int64_t m) {
    if constexpr (std::is_same<T, double>::value)
      copy_mat<double>(m, m, MatView<double>{(double2*)src, m}, MatView<double>{dst, m});
    else
      convert_mat<double, float>(m, m, MatView<double>{(double2*)src, m}, MatView<float>{dst, m});
  }

  // LU of the m x m matrix A (ld m) in place, backup: m x m scratch.
  void factor(int64_t m, C* A, C* backup, Fac& F, const char* what) {
    MatView<T> a{A, m};
    C *Li = F.inv, *Ui = F.inv + inv_blocks(m);
    copy_mat<T>(m, m, a, MatView<T>{backup, m});
    int swap = 0;
    int info = getrf<T>(m, m, a, piv, Li, Ui, PIVOT_NONE, &swap, ws);
    const int32_t* sig = nullptr;
    if (info != 0 || swap) {
      ++fallbacks;
      copy_mat<T>(m, m, MatView<T>{backup, m}, a);
      info = getrf<T>(m, m, a, piv, Li, Ui, PIVOT_PARTIAL, nullptr, ws);
      MNP_CHECK(info == 0, MNP_ERR_SINGULAR, std::string("bemretlayer: singular ") + what);
      perm_from_pivots(m, piv, F.sig, ws);
      sig = F.sig;
    }
    F.f = LU<T>{a, Li, Ui, sig};
  }

  // From the FP64 G1, H1 and the structured G2c, H2c (each 5 x n x n).
  void build(double k0, double2 eps1, double2 eps2, const double2* G1d, const double2* H1d, double2* G2d,
             const double2* H2d) {
    const int64_t nn = n * n;
    k = k0;
    e1 = eps1;
    e2 = eps2;
    if constexpr (std::is_same<T, double>::value)
      G2c = G2d;
    else
      for (int c = 0; c < 5; ++c) load(G2d + c * nn, G2(c), n);
    const double2 one = mk<double>(1, 0), zero = mk<double>(0, 0);
    // G1 and G2.p factorised; Sigma1 = H1 G1^{-1}, Sigma1 - Sigma2p = Sigma1 - H2p G2p^{-1}
    load(G1d, G1, n);
    factor(n, G1, Mb, f1, "G1");
    copy_mat<T>(n, n, v(G2(0)), v(F2));
    factor(n, F2, Mb, f2, "G2.p");
    load(H1d, S1, n);
    getrs_right<T>(n, n, f1.f, v(S1), ws);
    load(H2d, Tm, n);  // H2.p
    getrs_right<T>(n, n, f2.f, v(Tm), ws);
    hipLaunchKernelGGL(k_zaxpby<T>, dim3(nblk(nn)), dim3(256), 0, stream(), nn, one, S1, mk<double>(-1, 0), Tm, Tm);
    MNP_HIP(hipGetLastError());
    // Gamma = (Sigma1 - Sigma2p)^{-1}
    factor(n, Tm, Mb, fg, "Sigma1 - Sigma2p");
    getri<T>(n, fg.f, v(Gam), ws);
    // W = eps1 Sigma1 + k^2 (eps1 - eps2)^2 Gamma .* (npar npar')
    const double2 de = csub<double>(e1, e2);
    const double2 de2k = cscale<double>(cmul<double>(de, de), k * k);
    hipLaunchKernelGGL(k_lay_w<T>, dim3(nblk(nn)), dim3(256), 0, stream(), n, nvec, e1, S1, de2k, Gam, Tm);
    MNP_HIP(hipGetLastError());
    // M = [W G2ss - eps2 H2ss - ik de nz G2sh,  W G2sh - eps2 H2sh - ik de nz G2hh;
    //      S1 G2hs - H2hs - ik de nz G2ss,      S1 G2hh - H2hh - ik de nz G2sh]
    const int64_t ld = 2 * n;
    const double2 ikde = cmul<double>(mk<double>(0, k), de);
    struct Blk {
      C* A;
      int gc, hc, zc;
      double2 al;
      int64_t off;
    };
    const Blk blks[4] = {{Tm, 1, 1, 3, e2, 0},
                         {Tm, 3, 3, 4, e2, n * ld},
                         {S1, 2, 2, 1, one, n},
                         {S1, 4, 4, 3, one, n * ld + n}};
    C* Hc = Tm + nn;  // H2 component in T
    for (const Blk& b : blks) {
      gemm<T>(n, n, n, one, v(b.A), v(G2(b.gc)), zero, v(M + b.off, ld));
      load(H2d + b.hc * nn, Hc, n);
      hipLaunchKernelGGL(k_lay_mblock<T>, dim3(nblk(nn)), dim3(256), 0, stream(), n, ld, M + b.off, b.al, Hc, ikde,
                         nvec + 2 * n, G2(b.zc));
      MNP_HIP(hipGetLastError());
    }
    factor(ld, M, Mb, fm, "response matrix");
  }

  // mldivide.m:31-76 from phi, a, alpha, De (T, consumed) into sig1, sig2, h1, h2 (T).
  void apply(int64_t p, C* phi, C* a, C* alpha, C* De, C* sig1, C* sig2, C* h1, C* h2) {
    const int64_t np1 = n * p, np3 = 3 * n * p;
    C* buf = (C*)wsv.get(0, sizeof(C) * (size_t)(2 * np1 + 3 * np3 + 2 * np1));
    C *T2 = buf, *U = T2 + np1, *T1 = U + np1, *T3 = T1 + np3, *V = T3 + np3, *xi = V + np3;
    const double2 one = mk<double>(1, 0), zero = mk<double>(0, 0);
    const unsigned b1 = nblk(np1);
    // alpha = alpha - Sigma1 a + ik eps1 n phi
    gemm<T>(n, 3 * p, n, one, v(S1), v(a), zero, v(T1));
    hipLaunchKernelGGL(k_lay_alpha<T>, dim3(b1), dim3(256), 0, stream(), n, p, alpha, T1, phi, nvec,
                       cmul<double>(mk<double>(0, k), e1));
    // De = De - eps1 Sigma1 phi + ik eps1 n.a + ik (eps1 - eps2) npar.(Gamma alpha); rhs = [De; alpha_z]
    gemm<T>(n, p, n, one, v(S1), v(phi), zero, v(T2));
    gemm<T>(n, 3 * p, n, one, v(Gam), v(alpha), zero, v(T3));
    const double2 de = csub<double>(e1, e2);
    hipLaunchKernelGGL(k_lay_rhs<T>, dim3(b1), dim3(256), 0, stream(), n, p, De, T2, a, T3, alpha, nvec, e1,
                       cmul<double>(mk<double>(0, k), e1), cmul<double>(mk<double>(0, k), de), xi);
    MNP_HIP(hipGetLastError());
    getrs<T>(2 * n, p, fm.f, v(xi, 2 * n), ws);
    // sig2, h2z = xi; U = G2ss sig2 + G2sh h2z
    gemm<T>(n, p, n, one, v(G2(1)), v(xi, 2 * n), zero, v(U));
    gemm<T>(n, p, n, one, v(G2(3)), v(xi + n, 2 * n), one, v(U));
    // h2par = G2p^{-1} Gamma (alpha_par + ik (eps1 - eps2) npar U)
    hipLaunchKernelGGL(k_lay_v<T>, dim3(b1), dim3(256), 0, stream(), n, p, alpha, U, nvec,
                       cmul<double>(mk<double>(0, k), de), V);
    gemm<T>(n, 3 * p, n, one, v(Gam), v(V), zero, v(h2));
    getrs<T>(n, 3 * p, f2.f, v(h2), ws);  // z-components stay zero
    // h1 = G1^{-1} ([G2p h2par; G2hs sig2 + G2hh h2z] + a)
    gemm<T>(n, 3 * p, n, one, v(G2(0)), v(h2), zero, v(h1));
    gemm<T>(n, p, n, one, v(G2(2)), v(xi, 2 * n), zero, v(h1 + 2 * n, 3 * n));
    gemm<T>(n, p, n, one, v(G2(4)), v(xi + n, 2 * n), one, v(h1 + 2 * n, 3 * n));
    hipLaunchKernelGGL(k_zaxpby<T>, dim3(nblk(np3)), dim3(256), 0, stream(), np3, one, h1, one, a, h1);
    getrs<T>(n, 3 * p, f1.f, v(h1), ws);
    hipLaunchKernelGGL(k_lay_setz<T>, dim3(b1), dim3(256), 0, stream(), n, p, xi + n, 2 * n, h2);
    hipLaunchKernelGGL(k_lay_copy_rows<T>, dim3(b1), dim3(256), 0, stream(), n, p, xi, 2 * n, sig2);
    // sig1 = G1^{-1} (U + phi)
    hipLaunchKernelGGL(k_zaxpby<T>, dim3(b1), dim3(256), 0, stream(), np1, one, U, one, phi, sig1);
    getrs<T>(n, p, f1.f, v(sig1), ws);
    MNP_HIP(hipGetLastError());
  }
};

// Dense solver above the substrate: the FP64 operator (G1, H1 and the
// structured G2, H2 with the reflected parts), and either the FP64 chain
// (MNP_SOLVER_FP64, the plain drop-in) or the complex FP32 chain refined
// against the FP64 operator (MNP_SOLVER_MIXED): x += chain(b - A x) until
// |b - A x| <= tol |b|, A the operator the FP64 chain inverts.
struct BemRetLayer {
  const GreenRet* g;
  GreenRetLayer* gl;
  int64_t n;
  int solver;
  double2 *G1, *H1, *G2c, *H2c;  // FP64 operator
  std::unique_ptr<LayerChain<double>> c64;
  std::unique_ptr<LayerChain<float>> c32;
  Workspace wsr;  // refinement vectors
  double k = 0, enei = -1;
  double2 e1{}, e2{};
  bool ready = false;
  double tol = 1e-12;
  int maxit = 10, last_it = 0;
  double last_relres = 0;

  BemRetLayer(const GreenRet* g_, GreenRetLayer* gl_, int solver_) : g(g_), gl(gl_), n(g_->n), solver(solver_) {
    MNP_CHECK(gl->g == g, MNP_ERR_ARG, "bemretlayer: the layer Green function belongs to another particle");
    MNP_CHECK(solver == MNP_SOLVER_FP64 || solver == MNP_SOLVER_MIXED, MNP_ERR_ARG,
              "bemretlayer: solver must be MNP_SOLVER_FP64 or MNP_SOLVER_MIXED");
    const int64_t nn = n * n;
    G1 = dev_alloc<double2>(nn);
    H1 = dev_alloc<double2>(nn);
    G2c = dev_alloc<double2>(5 * nn);
    H2c = dev_alloc<double2>(5 * nn);
    if (solver == MNP_SOLVER_FP64)
      c64.reset(new LayerChain<double>(n, g->geo.nvec));
    else
      c32.reset(new LayerChain<float>(n, g->geo.nvec));
  }
  ~BemRetLayer() {
    void* ps[] = {G1, H1, G2c, H2c};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
  int fallbacks() const { return c64 ? c64->fallbacks : c32->fallbacks; }
  double2* G2(int c) const { return G2c + c * n * n; }
  double2* H2(int c) const { return H2c + c * n * n; }
  MatView<double> v(double2* p, int64_t ld = -1) const { return MatView<double>{p, ld < 0 ? n : ld}; }

  void initmat(double enei_, double2 eps_in, double2 eps_up, double2 eps_down) {
    const int64_t nn = n * n;
    enei = enei_;
    e1 = eps_in;
    e2 = eps_up;
    const double k0 = 2 * M_PI / enei;
    k = k0;
    ready = false;
    g->assemble(wavenumber(k0, eps_in), 2 * M_PI, OutPair{G1, nullptr}, OutPair{H1, nullptr}, n);
    // reflected parts into G2c, H2c; the direct outside G, H added to p, ss, hh
    // (compgreenretlayer/private/eval1.m:37-45)
    gl->eval(enei, eps_up, eps_down, G2c, H2c);
    double2* Tg = (double2*)wsr.get(0, sizeof(double2) * (size_t)(2 * nn));
    g->assemble(wavenumber(k0, eps_up), -2 * M_PI, OutPair{Tg, nullptr}, OutPair{Tg + nn, nullptr}, n);
    const double2 one = mk<double>(1, 0);
    for (int c : {0, 1, 4}) {
      hipLaunchKernelGGL(k_zaxpby<double>, dim3(nblk(nn)), dim3(256), 0, stream(), nn, one, G2(c), one, Tg, G2(c));
      hipLaunchKernelGGL(k_zaxpby<double>, dim3(nblk(nn)), dim3(256), 0, stream(), nn, one, H2(c), one, Tg + nn,
                         H2(c));
    }
    MNP_HIP(hipGetLastError());
    if (c64)
      c64->build(k0, e1, e2, G1, H1, G2c, H2c);
    else
      c32->build(k0, e1, e2, G1, H1, G2c, H2c);
    ready = true;
  }

  // r = b - A x (all FP64; x = sig1 | sig2 | h1 | h2 buffers, b and r packed phi | a | alpha | De)
  void residual(int64_t p, const double2* b, const double2* s1, const double2* s2, const double2* h1,
                const double2* h2, double2* r) {
    const int64_t np1 = n * p, np3 = 3 * np1;
    double2* buf = (double2*)wsr.get(1, sizeof(double2) * (size_t)(7 * np1 + 4 * np3));
    double2 *Gs1 = buf, *Hs1 = Gs1 + np1, *X = Hs1 + np1, *Y = X + np1, *phi2 = Y + np1, *dphi2 = phi2 + np1,
            *Gh1 = dphi2 + np1 + np1, *Hh1 = Gh1 + np3, *A2 = Hh1 + np3, *dA2 = A2 + np3;
    const double2 one = mk<double>(1, 0), zero = mk<double>(0, 0);
    const MatView<double> S1v{(double2*)s1, n}, S2v{(double2*)s2, n}, H1v{(double2*)h1, n};
    const MatView<double> hx{(double2*)h2, 3 * n}, hy{(double2*)h2 + n, 3 * n}, hz{(double2*)h2 + 2 * n, 3 * n};
    gemm<double>(n, p, n, one, v(G1), S1v, zero, v(Gs1));
    gemm<double>(n, 3 * p, n, one, v(G1), H1v, zero, v(Gh1));
    gemm<double>(n, p, n, one, v(H1), S1v, zero, v(Hs1));
    gemm<double>(n, 3 * p, n, one, v(H1), H1v, zero, v(Hh1));
    gemm<double>(n, p, n, one, v(G2(0)), hx, zero, v(A2, 3 * n));
    gemm<double>(n, p, n, one, v(G2(0)), hy, zero, v(A2 + n, 3 * n));
    gemm<double>(n, p, n, one, v(G2(4)), hz, zero, v(A2 + 2 * n, 3 * n));
    gemm<double>(n, p, n, one, v(G2(2)), S2v, zero, v(X));
    gemm<double>(n, p, n, one, v(G2(3)), S2v, zero, v(Y));
    gemm<double>(n, p, n, one, v(G2(1)), S2v, zero, v(phi2));
    gemm<double>(n, p, n, one, v(G2(3)), hz, one, v(phi2));
    gemm<double>(n, p, n, one, v(H2(0)), hx, zero, v(dA2, 3 * n));
    gemm<double>(n, p, n, one, v(H2(0)), hy, zero, v(dA2 + n, 3 * n));
    gemm<double>(n, p, n, one, v(H2(2)), S2v, zero, v(dA2 + 2 * n, 3 * n));
    gemm<double>(n, p, n, one, v(H2(4)), hz, one, v(dA2 + 2 * n, 3 * n));
    gemm<double>(n, p, n, one, v(H2(1)), S2v, zero, v(dphi2));
    gemm<double>(n, p, n, one, v(H2(3)), hz, one, v(dphi2));
    LayProd P{Gs1, Gh1, Hs1, Hh1, A2, X, Y, phi2, dA2, dphi2};
    hipLaunchKernelGGL(k_lay_resid, dim3(nblk(np1)), dim3(256), 0, stream(), n, p, P, b, r, g->geo.nvec, k, e1, e2);
    MNP_HIP(hipGetLastError());
  }

  void solve(int64_t p, const double2* const in[8], double2* sig1, double2* sig2, double2* h1, double2* h2) {
    MNP_CHECK(ready, MNP_ERR_STATE, "bemretlayer: solve before initmat");
    const int64_t np1 = n * p, np3 = 3 * np1, nb = 2 * np1 + 2 * np3;
    double2* b = (double2*)wsr.get(2, sizeof(double2) * (size_t)nb);
    double2 *phi = b, *a = b + np1, *alpha = a + np3, *De = alpha + np3;
    hipLaunchKernelGGL(k_lay_exc, dim3(nblk(np1)), dim3(256), 0, stream(), n, p, in[0], in[1], in[2], in[3], in[4],
                       in[5], in[6], in[7], g->geo.nvec, k, e1, e2, phi, a, alpha, De);
    MNP_HIP(hipGetLastError());
    last_it = 0;
    last_relres = 0;
    if (c64) {
      c64->apply(p, phi, a, alpha, De, sig1, sig2, h1, h2);  // consumes b
      return;
    }
    // mixed: FP32 chain applications refined with the FP64 operator
    float2* f = (float2*)wsr.get(3, sizeof(float2) * (size_t)(2 * nb));
    float2* fb = f;            // packed right-hand side in FP32
    float2* fx = f + nb;       // sig1 | sig2 | h1 | h2 in FP32
    double2* r = (double2*)wsr.get(4, sizeof(double2) * (size_t)nb);
    double2* dx = (double2*)wsr.get(5, sizeof(double2) * (size_t)nb);
    auto to32 = [&](const double2* src, float2* dst) {
      convert_mat<double, float>(nb, 1, MatView<double>{(double2*)src, nb}, MatView<float>{dst, nb});
    };
    auto chain = [&](const double2* rhs, double2* out) {  // out = sig1 | sig2 | h1 | h2 (FP64)
      to32(rhs, fb);
      c32->apply(p, fb, fb + np1, fb + np1 + np3, fb + np1 + 2 * np3, fx, fx + np1, fx + 2 * np1, fx + 2 * np1 + np3);
      convert_mat<float, double>(nb, 1, MatView<float>{fx, nb}, MatView<double>{out, nb});
    };
    double2* x = dx;  // first solution directly into the outputs
    chain(b, r);
    MNP_HIP(hipMemcpyAsync(sig1, r, sizeof(double2) * np1, hipMemcpyDeviceToDevice, stream()));
    MNP_HIP(hipMemcpyAsync(sig2, r + np1, sizeof(double2) * np1, hipMemcpyDeviceToDevice, stream()));
    MNP_HIP(hipMemcpyAsync(h1, r + 2 * np1, sizeof(double2) * np3, hipMemcpyDeviceToDevice, stream()));
    MNP_HIP(hipMemcpyAsync(h2, r + 2 * np1 + np3, sizeof(double2) * np3, hipMemcpyDeviceToDevice, stream()));
    const double nrm_b = dev_norm2(nb, b, wsr, 6);
    for (int it = 0;; ++it) {
      residual(p, b, sig1, sig2, h1, h2, r);
      last_relres = nrm_b > 0 ? dev_norm2(nb, r, wsr, 6) / nrm_b : 0.0;
      last_it = it;
      if (last_relres <= tol) break;
      MNP_CHECK(it < maxit, MNP_ERR_NOCONV,
                "bemretlayer: refinement stopped at relative residual " + sci(last_relres));
      chain(r, x);
      const double2 one = mk<double>(1, 0);
      double2* outs[4] = {sig1, sig2, h1, h2};
      const int64_t off[4] = {0, np1, 2 * np1, 2 * np1 + np3}, len[4] = {np1, np1, np3, np3};
      for (int i = 0; i < 4; ++i)
        hipLaunchKernelGGL(k_zaxpby<double>, dim3(nblk(len[i])), dim3(256), 0, stream(), len[i], one, outs[i], one,
                           x + off[i], outs[i]);
      MNP_HIP(hipGetLastError());
    }
  }
};

}  // namespace mnp

namespace mnp {
// Uploads the dipoles (host pts npt x 3, dip npt x 3 x ndip) and runs
// k_dip_layer<MODE> against the table at enei; raises when a pair fell
// outside the table.
template <int MODE>
void run_dip_layer(LayerTab& tab, double enei, double2 e1, double2 e2, int64_t no, const double* obs,
                   const double* onv, int64_t npt, const double* pts, double shift, int64_t ndip, const double* dip,
                   double2* o0, double2* o1, double2* o2, double2* o3) {
  for (int64_t i = 0; i < npt; ++i)
    MNP_CHECK(pts[2 * npt + i] > tab.op.z0, MNP_ERR_ARG, "dipolelayer: the dipoles must lie above the interface");
  double* dp = dev_alloc<double>(3 * npt + 3 * npt * ndip);
  MNP_HIP(hipMemcpy(dp, pts, sizeof(double) * 3 * npt, hipMemcpyHostToDevice));
  MNP_HIP(hipMemcpy(dp + 3 * npt, dip, sizeof(double) * 3 * npt * ndip, hipMemcpyHostToDevice));
  const TabView T = tab.view(tab.at(enei, e1, e2));
  MNP_HIP(hipMemsetAsync(tab.err, 0, sizeof(int), stream()));
  const double2 ieps = cdiv<double>(mk<double>(1, 0), e1);
  hipLaunchKernelGGL(k_dip_layer<MODE>, dim3(nblk(no * npt)), dim3(256), 0, stream(), no, obs, onv, npt, dp, shift,
                     ndip, dp + 3 * npt, ieps, 2 * M_PI / enei, T, o0, o1, o2, o3);
  MNP_HIP(hipGetLastError());
  int herr = 0;
  MNP_HIP(hipMemcpyAsync(&herr, tab.err, sizeof(int), hipMemcpyDeviceToHost, stream()));
  MNP_HIP(hipStreamSynchronize(stream()));
  MNP_HIP(hipFree(dp));
  MNP_CHECK(herr == 0, MNP_ERR_ARG, "greentablayer: a dipole pair lies outside the table (tabspace with the dipoles?)");
}
}  // namespace mnp

struct mnp_layertab_s {
  mnp::LayerTab* t;
};
struct mnp_greenretlayer_s {
  mnp::GreenRetLayer* gl;
};
namespace mnp {
GreenRetLayer* greenretlayer_of(mnp_greenretlayer_t h) { return h ? h->gl : nullptr; }
}  // namespace mnp
struct mnp_bemretlayer_s {
  mnp::BemRetLayer* b;
};
struct mnp_pointlayer_s {
  mnp::PointLayer* pl;
};

using namespace mnp;

extern "C" {

void mnp_layer_options_default(mnp_layer_options* op) {
  if (!op) return;
  op->z0 = 0;
  op->ztol = 2e-2;
  op->rmin = 1e-2;
  op->zmin = 1e-2;
  op->semi = 0.1;
  op->ratio = 2;
}

int mnp_layertab_create(const mnp_layer_options* op, int64_t nr, const double* r, int64_t nz, const double* z1,
                        double z2, mnp_layertab_t* out) {
  return guarded([&] {
    MNP_CHECK(op && r && z1 && out, MNP_ERR_ARG, "layertab_create: NULL argument");
    auto* t = new LayerTab(*op, nr, r, nz, z1, z2);
    *out = new mnp_layertab_s{t};
  });
}
int mnp_layertab_destroy(mnp_layertab_t t) {
  return guarded([&] {
    if (!t) return;
    delete t->t;
    delete t;
  });
}
int mnp_layertab_set(mnp_layertab_t t, int64_t nenei, const double* enei, const mnp_complex* eps_up,
                     const mnp_complex* eps_down) {
  return guarded([&] {
    MNP_CHECK(t && enei && eps_up && eps_down, MNP_ERR_ARG, "layertab_set: NULL argument");
    t->t->set(nenei, enei, eps_up, eps_down);
    MNP_HIP(hipStreamSynchronize(stream()));
  });
}
int mnp_layertab_eval(mnp_layertab_t t, double enei, mnp_complex eps_up, mnp_complex eps_down, void* out) {
  return guarded([&] {
    MNP_CHECK(t, MNP_ERR_ARG, "layertab_eval: NULL argument");
    const double2* v = t->t->at(enei, to_d2(eps_up), to_d2(eps_down));
    if (out)
      MNP_HIP(hipMemcpyAsync(out, v, sizeof(double2) * t->t->size(), hipMemcpyDeviceToDevice, stream()));
    MNP_HIP(hipStreamSynchronize(stream()));
  });
}
int mnp_greenretlayer_create(mnp_greenret_t g, mnp_layertab_t tab, mnp_greenretlayer_t* out) {
  return guarded([&] {
    MNP_CHECK(g && tab && out, MNP_ERR_ARG, "greenretlayer_create: NULL argument");
    auto* gl = new GreenRetLayer(g->g, tab->t);
    *out = new mnp_greenretlayer_s{gl};
  });
}
int mnp_greenretlayer_destroy(mnp_greenretlayer_t gl) {
  return guarded([&] {
    if (!gl) return;
    delete gl->gl;
    delete gl;
  });
}
int mnp_greenretlayer_refined(mnp_greenretlayer_t gl, int64_t* npairs, int64_t* ndiag) {
  return guarded([&] {
    MNP_CHECK(gl && npairs && ndiag, MNP_ERR_ARG, "greenretlayer_refined: NULL argument");
    *npairs = gl->gl->np;
    *ndiag = gl->gl->nd;
  });
}
int mnp_greenretlayer_eval(mnp_greenretlayer_t gl, double enei, mnp_complex eps_up, mnp_complex eps_down, void* G,
                           void* F) {
  return guarded([&] {
    MNP_CHECK(gl && G && F, MNP_ERR_ARG, "greenretlayer_eval: NULL argument");
    gl->gl->eval(enei, to_d2(eps_up), to_d2(eps_down), (double2*)G, (double2*)F);
  });
}
int mnp_bemretlayer_create(mnp_greenret_t g, mnp_greenretlayer_t gl, int solver, mnp_bemretlayer_t* out) {
  return guarded([&] {
    MNP_CHECK(g && gl && out, MNP_ERR_ARG, "bemretlayer_create: NULL argument");
    auto* b = new BemRetLayer(g->g, gl->gl, solver);
    *out = new mnp_bemretlayer_s{b};
  });
}
int mnp_bemretlayer_destroy(mnp_bemretlayer_t b) {
  return guarded([&] {
    if (!b) return;
    delete b->b;
    delete b;
  });
}
int mnp_bemretlayer_initmat(mnp_bemretlayer_t b, double enei, mnp_complex eps_in, mnp_complex eps_up,
                            mnp_complex eps_down) {
  return guarded([&] {
    MNP_CHECK(b, MNP_ERR_ARG, "bemretlayer_initmat: NULL argument");
    b->b->initmat(enei, to_d2(eps_in), to_d2(eps_up), to_d2(eps_down));
  });
}
int mnp_bemretlayer_solve(mnp_bemretlayer_t b, int64_t npol, const void* phi1, const void* phi1p, const void* a1,
                          const void* a1p, const void* phi2, const void* phi2p, const void* a2, const void* a2p,
                          void* sig1, void* sig2, void* h1, void* h2) {
  return guarded([&] {
    MNP_CHECK(b && sig1 && sig2 && h1 && h2, MNP_ERR_ARG, "bemretlayer_solve: NULL argument");
    MNP_CHECK(npol >= 1, MNP_ERR_ARG, "bemretlayer_solve: npol must be positive");
    const double2* in[8] = {(const double2*)phi1, (const double2*)phi1p, (const double2*)a1, (const double2*)a1p,
                            (const double2*)phi2, (const double2*)phi2p, (const double2*)a2, (const double2*)a2p};
    b->b->solve(npol, in, (double2*)sig1, (double2*)sig2, (double2*)h1, (double2*)h2);
  });
}
int mnp_bemretlayer_stats(mnp_bemretlayer_t b, int* fallbacks, int* iterations, double* relres) {
  return guarded([&] {
    MNP_CHECK(b && fallbacks && iterations && relres, MNP_ERR_ARG, "bemretlayer_stats: NULL argument");
    *fallbacks = b->b->fallbacks();
    *iterations = b->b->last_it;
    *relres = b->b->last_relres;
  });
}
int mnp_pointlayer_create(mnp_greenret_t g, mnp_layertab_t tab, int64_t npt, const double* pts,
                          mnp_pointlayer_t* out) {
  return guarded([&] {
    MNP_CHECK(g && tab && pts && out, MNP_ERR_ARG, "pointlayer_create: NULL argument");
    auto* pl = new PointLayer(g->g, tab->t, npt, pts);
    *out = new mnp_pointlayer_s{pl};
  });
}
int mnp_pointlayer_destroy(mnp_pointlayer_t pl) {
  return guarded([&] {
    if (!pl) return;
    delete pl->pl;
    delete pl;
  });
}
int mnp_pointlayer_fields(mnp_pointlayer_t pl, double enei, mnp_complex eps_up, mnp_complex eps_down, int64_t ncol,
                          const void* sig2, const void* h2, void* phi, void* a, void* e, void* h) {
  return guarded([&] {
    MNP_CHECK(pl && sig2 && h2 && phi && a && e && h, MNP_ERR_ARG, "pointlayer_fields: NULL argument");
    MNP_CHECK(ncol >= 1, MNP_ERR_ARG, "pointlayer_fields: ncol must be positive");
    pl->pl->fields(enei, to_d2(eps_up), to_d2(eps_down), ncol, (const double2*)sig2, (const double2*)h2,
                   (double2*)phi, (double2*)a, (double2*)e, (double2*)h);
  });
}
int mnp_planewaveretlayer_potential(mnp_greenret_t g, const mnp_layer_options* op, double enei, mnp_complex eps_up,
                                    mnp_complex eps_down, int64_t npol, const double* pol, const double* dir,
                                    void* phi2, void* phi2p, void* a2, void* a2p) {
  return guarded([&] {
    MNP_CHECK(g && op && pol && dir && phi2 && phi2p && a2 && a2p, MNP_ERR_ARG,
              "planewaveretlayer_potential: NULL argument");
    MNP_CHECK(npol >= 1, MNP_ERR_ARG, "planewaveretlayer_potential: npol must be positive");
    const int64_t n = g->g->n;
    std::vector<double> pd(6 * npol);
    for (int64_t i = 0; i < 3 * npol; ++i) {
      pd[i] = pol[i];
      pd[3 * npol + i] = dir[i];
    }
    double* d = dev_alloc<double>(6 * npol);
    MNP_HIP(hipMemcpy(d, pd.data(), sizeof(double) * 6 * npol, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_pw_layer, dim3(nblk(n * npol)), dim3(256), 0, stream(), n, npol, g->g->geo.pos,
                       g->g->geo.nvec, d, 2 * M_PI / enei, to_d2(eps_up), to_d2(eps_down), op->z0, (double2*)phi2,
                       (double2*)phi2p, (double2*)a2, (double2*)a2p);
    MNP_HIP(hipGetLastError());
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_HIP(hipFree(d));
  });
}
int mnp_dipolelayer_potential(mnp_greenret_t g, mnp_layertab_t tab, double enei, mnp_complex eps_up,
                              mnp_complex eps_down, int64_t npt, const double* pts, int64_t ndip, const double* dip,
                              void* phi, void* phip, void* a, void* ap) {
  return guarded([&] {
    MNP_CHECK(g && tab && pts && dip && phi && phip && a && ap, MNP_ERR_ARG, "dipolelayer_potential: NULL argument");
    MNP_CHECK(npt >= 1 && ndip >= 1, MNP_ERR_ARG, "dipolelayer_potential: npt and ndip must be positive");
    const GreenRet& G = *g->g;
    run_dip_layer<0>(*tab->t, enei, to_d2(eps_up), to_d2(eps_down), G.n, G.geo.pos, G.geo.nvec, npt, pts, 0.0, ndip,
                     dip, (double2*)phi, (double2*)phip, (double2*)a, (double2*)ap);
  });
}
int mnp_dipolelayer_field(mnp_layertab_t tab, double enei, mnp_complex eps_up, mnp_complex eps_down, int64_t nobs,
                          const double* obs, int64_t npt, const double* pts, int64_t ndip, const double* dip, void* e,
                          void* h) {
  return guarded([&] {
    MNP_CHECK(tab && obs && pts && dip && e && h, MNP_ERR_ARG, "dipolelayer_field: NULL argument");
    MNP_CHECK(nobs >= 1 && npt >= 1 && ndip >= 1, MNP_ERR_ARG, "dipolelayer_field: sizes must be positive");
    bool same = nobs == npt;
    for (int64_t i = 0; same && i < 3 * npt; ++i) same = obs[i] == pts[i];
    double* d = dev_alloc<double>(3 * nobs);
    MNP_HIP(hipMemcpy(d, obs, sizeof(double) * 3 * nobs, hipMemcpyHostToDevice));
    run_dip_layer<1>(*tab->t, enei, to_d2(eps_up), to_d2(eps_down), nobs, d, nullptr, npt, pts,
                     same ? tab->t->op.rmin : 0.0, ndip, dip, nullptr, nullptr, (double2*)e, (double2*)h);
    MNP_HIP(hipFree(d));
  });
}
int mnp_dipolelayer_farfield(const mnp_layer_options* op, double enei, mnp_complex eps_up, mnp_complex eps_down,
                             int64_t ndir, const void* dir, int64_t npt, const double* pts, int64_t ndip,
                             const double* dip, void* e, void* h) {
  return guarded([&] {
    MNP_CHECK(op && dir && pts && dip && e && h, MNP_ERR_ARG, "dipolelayer_farfield: NULL argument");
    MNP_CHECK(ndir >= 1 && npt >= 1 && ndip >= 1, MNP_ERR_ARG, "dipolelayer_farfield: sizes must be positive");
    for (int64_t i = 0; i < npt; ++i)
      MNP_CHECK(pts[2 * npt + i] > op->z0, MNP_ERR_ARG, "dipolelayer_farfield: the dipoles must lie above the interface");
    double* dp = dev_alloc<double>(3 * npt + 3 * npt * ndip);
    MNP_HIP(hipMemcpy(dp, pts, sizeof(double) * 3 * npt, hipMemcpyHostToDevice));
    MNP_HIP(hipMemcpy(dp + 3 * npt, dip, sizeof(double) * 3 * npt * ndip, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_dip_layer_ff, dim3(nblk(ndir * npt)), dim3(256), 0, stream(), ndir, (const double*)dir, npt,
                       dp, ndip, dp + 3 * npt, 2 * M_PI / enei, to_d2(eps_up), to_d2(eps_down), op->z0, (double2*)e,
                       (double2*)h);
    MNP_HIP(hipGetLastError());
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_HIP(hipFree(dp));
  });
}
int mnp_farfield_layer(mnp_greenret_t g, const mnp_layer_options* op, double enei, mnp_complex eps_up,
                       mnp_complex eps_down, int64_t ndir, const void* dir, int64_t npol, const void* sig2,
                       const void* h2, void* e, void* h) {
  return guarded([&] {
    MNP_CHECK(g && op && dir && sig2 && h2 && e && h, MNP_ERR_ARG, "farfield_layer: NULL argument");
    MNP_CHECK(npol >= 1 && npol <= LFF_MAXP, MNP_ERR_ARG, "farfield_layer: npol must be in 1..8");
    const GreenRet& G = *g->g;
    hipLaunchKernelGGL(k_farfield_layer, dim3((unsigned)ndir), dim3(256), 0, stream(), G.n, npol, ndir,
                       (const double*)dir, G.geo.pos, G.geo.area, (const double2*)sig2, (const double2*)h2,
                       2 * M_PI / enei, to_d2(eps_up), to_d2(eps_down), op->z0, (double2*)e, (double2*)h);
    MNP_HIP(hipGetLastError());
    MNP_HIP(hipStreamSynchronize(stream()));
  });
}

}  // extern "C"
