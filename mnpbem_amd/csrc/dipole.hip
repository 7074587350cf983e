// Dipole excitation, fields at points and decay rates on gfx950 (MNPBEM
// Simulation/retarded/@dipoleret, Greenfun/@compgreenret/field.m).
//
//  - k_dip_potential: phi, phi', A, A' of every (position, orientation)
//    column on every face, one thread per face and column
//    (@dipoleret/potential.m:64-122);
//  - k_pt_field: E = i k0 G h - Gp sig and H = Gp x h at points, one
//    workgroup per (point, column) reducing over the faces; point-face pairs
//    within the refinement cutoff (refinematrix.m:21-68, the point taking the
//    face radius) use the rule-18 Taylor form of greenret/private/init.m
//    (shared with the face-face near field);
//  - decay rates: the induced field at each dipole, the far fields of the
//    surface charges plus the dipole's own far field on the pinfty normals,
//    Poynting sums and normalisation in FP64 on the host
//    (@dipoleret/decayrate.m:13-62, farfield.m:40-55, scattering.m).
#include <complex>
#include <vector>

#include "mnp_linalg.hpp"
#include "mnp_nearfield.hpp"

namespace mnp {

void farfield(const GreenRet& g, double enei, mnp_complex em, int64_t ndir, const double* dir, int64_t p,
              const double2* sig2, const double2* h2, double2* e, double2* h);
constexpr int FF_CHUNK = 8;  // columns per far-field launch

__constant__ double c_q18p[84];  // rule 18: x(28) | y(28) | w(28)

static void upload_rule18() {
  double q[84];
  rule18(q, q + 28, q + 56);
  MNP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_q18p), q, sizeof(q)));
}

static double2 wavenumber(double enei, mnp_complex eps) {
  const std::complex<double> k = 2 * M_PI / enei * std::sqrt(std::complex<double>(eps.re, eps.im));
  return mk<double>(k.real(), k.imag());
}

// ------------------------------------------------------------ potentials
// pts: npt x 3, dip: npt x 3 x ndip (column-major); column q = ipt + npt idip.
__global__ void __launch_bounds__(256) k_dip_potential(int64_t n, const double* __restrict__ pos,
                                                       const double* __restrict__ nvec, int64_t npt,
                                                       const double* __restrict__ pts, const double* __restrict__ dip,
                                                       double2 k, double2 eps, double k0, double2* phi, double2* phip,
                                                       double2* a, double2* ap) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const int64_t ipt = q % npt;
  double u[3], nv[3], dp[3];
  for (int c = 0; c < 3; ++c) {
    u[c] = pos[c * n + i] - pts[c * npt + ipt];
    nv[c] = nvec[c * n + i];
    dp[c] = dip[ipt + npt * (c + 3 * (q / npt))];
  }
  const double r = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (int c = 0; c < 3; ++c) u[c] /= r;
  const double2 G = cscale<double>(cexp_ikd(k, r), 1.0 / r);
  const double2 F = cmul<double>(mk<double>(-k.y - 1.0 / r, k.x), G);  // (ik - 1/r) G
  const double en = nv[0] * u[0] + nv[1] * u[1] + nv[2] * u[2];
  const double ep = u[0] * dp[0] + u[1] * dp[1] + u[2] * dp[2];
  const double np = nv[0] * dp[0] + nv[1] * dp[1] + nv[2] * dp[2];
  const double2 ieps = cdiv<double>(mk<double>(1, 0), eps);
  // phi = -ep F / eps
  phi[q * n + i] = cscale<double>(cmul<double>(F, ieps), -ep);
  // phi' = (np - 3 en ep)/r^2 (1 - i k r) G / eps + k^2 ep en G / eps
  const double2 one_ikr = mk<double>(1 + k.y * r, -k.x * r);
  const double2 t1 = cscale<double>(cmul<double>(one_ikr, G), (np - 3 * en * ep) / (r * r));
  const double2 t2 = cscale<double>(cmul<double>(cmul<double>(k, k), G), ep * en);
  phip[q * n + i] = cmul<double>(cadd<double>(t1, t2), ieps);
  // A = -i k0 dip G, A' = -i k0 dip en F
  for (int c = 0; c < 3; ++c) {
    const double s = -k0 * dp[c];
    a[(c + 3 * q) * n + i] = mk<double>(-s * G.y, s * G.x);
    ap[(c + 3 * q) * n + i] = mk<double>(-s * en * F.y, s * en * F.x);
  }
}

// ------------------------------------------------------- fields at points
// One workgroup per (point, column); diag != 0 evaluates column q at point
// q % npt only (grid.x = 1) and writes e/hf as 3 x ncol.
// POT: the potentials phi = G sig (into e, npt x ncol) and A = G h (into hf,
// npt x 3 x ncol) instead of the fields (compgreenret/potential.m:30-40).
template <bool POT>
__global__ void __launch_bounds__(256) k_pt_field(int64_t n, Geo geo, int64_t npt, const double* __restrict__ pts,
                                                  int64_t ncol, const double2* __restrict__ sig,
                                                  const double2* __restrict__ h, double2 k, double k0, int order,
                                                  double rel, double absc, int diag, double2* e, double2* hf) {
  const int64_t q = blockIdx.y;
  const int64_t ipt = diag ? q % npt : blockIdx.x;
  double x[3];
  for (int c = 0; c < 3; ++c) x[c] = pts[c * npt + ipt];
  double fact[MNP_MAX_ORDER + 2];
  fact[0] = 1;
  for (int o = 1; o <= order + 1; ++o) fact[o] = fact[o - 1] * o;
  const double2 ik = mk<double>(-k.y, k.x);
  double acc[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // e (3 complex), h (3 complex)
  for (int64_t j = threadIdx.x; j < n; j += blockDim.x) {
    double xv[3];
    for (int c = 0; c < 3; ++c) xv[c] = x[c] - geo.pos[c * n + j];
    const double d = sqrt(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]);
    const double d2 = d - geo.rad[j];
    const double2 ph = cexp_ikd(k, d);
    double2 G, Gp[3];
    if (d2 < absc || d2 / geo.rad[j] < rel) {
      double g[MNP_MAX_ORDER + 1], f[3][MNP_MAX_ORDER + 1];
      for (int o = 0; o <= order; ++o) g[o] = f[0][o] = f[1][o] = f[2][o] = 0;
      near_taylor(x, geo, n, j, order, fact, c_q18p, g, f);
      double2 pw = mk<double>(1, 0), gs = mk<double>(0, 0), fs[3] = {gs, gs, gs};
      for (int o = 0; o <= order; ++o) {
        gs = cadd<double>(gs, cscale<double>(pw, g[o]));
        for (int c = 0; c < 3; ++c) fs[c] = cadd<double>(fs[c], cscale<double>(pw, f[c][o]));
        pw = cmul<double>(pw, ik);
      }
      G = cmul<double>(gs, ph);
      for (int c = 0; c < 3; ++c) Gp[c] = cmul<double>(fs[c], ph);
    } else {
      const double ar = geo.area[j];
      G = cscale<double>(ph, ar / d);
      const double2 fk = cmul<double>(mk<double>((-k.y - 1.0 / d) / (d * d), k.x / (d * d)), ph);
      for (int c = 0; c < 3; ++c) Gp[c] = cscale<double>(fk, xv[c] * ar);
    }
    const double2 s = sig[q * n + j];
    double2 hj[3] = {{0, 0}, {0, 0}, {0, 0}};  // h == nullptr: no currents
    if (h)
      for (int c = 0; c < 3; ++c) hj[c] = h[(c + 3 * q) * n + j];
    if constexpr (POT) {
      const double2 v = cmul<double>(G, s);
      acc[0] += v.x;
      acc[1] += v.y;
      for (int c = 0; c < 3; ++c) {
        const double2 a = cmul<double>(G, hj[c]);
        acc[2 + 2 * c] += a.x;
        acc[3 + 2 * c] += a.y;
      }
      continue;
    }
    const double2 ik0G = cmul<double>(mk<double>(0, k0), G);
    for (int c = 0; c < 3; ++c) {
      const double2 v = csub<double>(cmul<double>(ik0G, hj[c]), cmul<double>(Gp[c], s));
      acc[2 * c] += v.x;
      acc[2 * c + 1] += v.y;
    }
    const double2 cx = csub<double>(cmul<double>(Gp[1], hj[2]), cmul<double>(Gp[2], hj[1]));
    const double2 cy = csub<double>(cmul<double>(Gp[2], hj[0]), cmul<double>(Gp[0], hj[2]));
    const double2 cz = csub<double>(cmul<double>(Gp[0], hj[1]), cmul<double>(Gp[1], hj[0]));
    acc[6] += cx.x;
    acc[7] += cx.y;
    acc[8] += cy.x;
    acc[9] += cy.y;
    acc[10] += cz.x;
    acc[11] += cz.y;
  }
  __shared__ double red[4][12];
  for (int t = 0; t < 12; ++t)
    for (int s = 32; s > 0; s >>= 1) acc[t] += __shfl_xor(acc[t], s);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int t = 0; t < 12; ++t) red[w][t] = acc[t];
  __syncthreads();
  if (POT && threadIdx.x < 8) {
    const int t = threadIdx.x;
    const double v = red[0][t] + red[1][t] + red[2][t] + red[3][t];
    // t 0, 1: phi; t 2..7: A component (t - 2) / 2
    double2* out = t < 2 ? e : hf;
    const int64_t idx = t < 2 ? ipt + npt * q : ipt + npt * ((t - 2) / 2 + 3 * q);
    if ((t & 1) == 0)
      out[idx].x = v;
    else
      out[idx].y = v;
  }
  if (!POT && threadIdx.x < 12) {
    const int t = threadIdx.x;
    const double v = red[0][t] + red[1][t] + red[2][t] + red[3][t];
    const int c = (t % 6) / 2, re = t & 1;
    double2* out = t < 6 ? e : hf;
    const int64_t idx = diag ? c + 3 * q : ipt + npt * (c + 3 * q);
    if (re == 0)
      out[idx].x = v;
    else
      out[idx].y = v;
  }
}

// ------------------------------------------------- quasistatic dipoles
// phi' = -nvec . E with E = w (3 u (u . p) - p) / r^3 (u = r / |r|, r = face -
// point) of dipole p at each point (Simulation/static/@dipolestat/field.m:
// 50-78, potential.m:17) plus, when wimg is set, the image dipole at the
// point mirrored at z = z0 with p_z negated (@dipolestatlayer/init.m:35-39,
// field.m:23-28).  w, wimg: per point (1 / eps of the point's medium, q1 /
// eps of the image point's medium).  pts npt x 3, dip npt x 3 x ndip.
__device__ __forceinline__ double2 dipstat_nE(const double r[3], const double nv[3], const double pz_sign,
                                              const double* dp, double2 w) {
  const double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  const double d = sqrt(d2), inv3 = 1.0 / (d2 * d);
  const double u[3] = {r[0] / d, r[1] / d, r[2] / d};
  const double p[3] = {dp[0], dp[1], pz_sign * dp[2]};
  const double up = u[0] * p[0] + u[1] * p[1] + u[2] * p[2];
  double ne = 0;
  for (int c = 0; c < 3; ++c) ne += nv[c] * (3 * u[c] * up - p[c]);
  ne *= inv3;
  return mk<double>(w.x * ne, w.y * ne);
}

__global__ void __launch_bounds__(256) k_dipstat_phip(int64_t n, const double* __restrict__ pos,
                                                      const double* __restrict__ nvec, int64_t npt,
                                                      const double* __restrict__ pts, const double* __restrict__ dip,
                                                      const double2* __restrict__ w, double z0,
                                                      const double2* __restrict__ wimg, double2* phip) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const int64_t ipt = q % npt, idip = q / npt;
  double nv[3], r[3], dp[3];
  for (int c = 0; c < 3; ++c) {
    nv[c] = nvec[c * n + i];
    r[c] = pos[c * n + i] - pts[c * npt + ipt];
    dp[c] = dip[ipt + npt * (c + 3 * idip)];
  }
  double2 v = dipstat_nE(r, nv, 1.0, dp, w[ipt]);
  if (wimg) {
    r[2] = pos[2 * n + i] - (2 * z0 - pts[2 * npt + ipt]);
    v = cadd<double>(v, dipstat_nE(r, nv, -1.0, dp, wimg[ipt]));
  }
  phip[q * n + i] = mk<double>(-v.x, -v.y);
}

static void pt_field(const GreenRet& G, double enei, mnp_complex eps, int64_t npt, const double* dpts, int64_t ncol,
                     const double2* sig, const double2* h, int diag, double2* e, double2* hf) {
  upload_rule18();
  const dim3 grid(diag ? 1u : (unsigned)npt, (unsigned)ncol);
  hipLaunchKernelGGL(k_pt_field<false>, grid, dim3(256), 0, stream(), G.n, G.geo, npt, dpts, ncol, sig, h,
                     wavenumber(enei, eps), 2 * M_PI / enei, G.order, G.op.rel_cutoff, G.op.abs_cutoff, diag, e,
                     hf);
  MNP_HIP(hipGetLastError());
}

// host npt x 3 (or npt x 3 x ndip) doubles -> device
static double* upload(const double* h, int64_t count) {
  double* d = dev_alloc<double>(count);
  MNP_HIP(hipMemcpyAsync(d, h, sizeof(double) * count, hipMemcpyHostToDevice, stream()));
  return d;
}

}  // namespace mnp

// ============================================================== C ABI
using namespace mnp;

extern "C" {

int mnp_dipole_potential(mnp_greenret_t g, double enei, mnp_complex eps_medium, int64_t npt, const double* pts,
                         int64_t ndip, const double* dip, void* phi, void* phip, void* a, void* ap) {
  return guarded([&] {
    MNP_CHECK(g && pts && dip && phi && phip && a && ap, MNP_ERR_ARG, "dipole_potential: NULL argument");
    MNP_CHECK(npt >= 1 && ndip >= 1 && enei > 0, MNP_ERR_ARG, "dipole_potential: bad sizes");
    const GreenRet& G = *g->g;
    double* dpts = upload(pts, 3 * npt);
    double* ddip = upload(dip, 3 * npt * ndip);
    const int64_t npol = npt * ndip;
    MNP_CHECK(npol < 65536, MNP_ERR_ARG, "dipole_potential: too many columns");
    const dim3 grid((unsigned)cdiv_up(G.n, 256), (unsigned)npol);
    hipLaunchKernelGGL(k_dip_potential, grid, dim3(256), 0, stream(), G.n, G.geo.pos, G.geo.nvec, npt, dpts, ddip,
                       wavenumber(enei, eps_medium), mk<double>(eps_medium.re, eps_medium.im), 2 * M_PI / enei,
                       (double2*)phi, (double2*)phip, (double2*)a, (double2*)ap);
    MNP_HIP(hipGetLastError());
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_HIP(hipFree(dpts));
    MNP_HIP(hipFree(ddip));
  });
}

int mnp_field_points(mnp_greenret_t g, double enei, mnp_complex eps_medium, int64_t npt, const double* pts,
                     int64_t ncol, const void* sig, const void* h, void* e, void* hfield) {
  return guarded([&] {
    MNP_CHECK(g && pts && sig && e && hfield, MNP_ERR_ARG, "field_points: NULL argument");
    MNP_CHECK(npt >= 1 && ncol >= 1 && enei > 0, MNP_ERR_ARG, "field_points: bad sizes");
    MNP_CHECK(npt < (1LL << 31) && ncol < 65536, MNP_ERR_ARG, "field_points: too many points or columns");
    const GreenRet& G = *g->g;
    double* dpts = upload(pts, 3 * npt);
    pt_field(G, enei, eps_medium, npt, dpts, ncol, (const double2*)sig, (const double2*)h, 0, (double2*)e,
             (double2*)hfield);
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_HIP(hipFree(dpts));
  });
}

int mnp_potential_points(mnp_greenret_t g, double enei, mnp_complex eps_medium, int64_t npt, const double* pts,
                         int64_t ncol, const void* sig, const void* h, void* phi, void* a) {
  return guarded([&] {
    MNP_CHECK(g && pts && sig && phi && a, MNP_ERR_ARG, "potential_points: NULL argument");
    MNP_CHECK(npt >= 1 && ncol >= 1 && enei > 0, MNP_ERR_ARG, "potential_points: bad sizes");
    MNP_CHECK(npt < (1LL << 31) && ncol < 65536, MNP_ERR_ARG, "potential_points: too many points or columns");
    const GreenRet& G = *g->g;
    double* dpts = upload(pts, 3 * npt);
    upload_rule18();
    hipLaunchKernelGGL(k_pt_field<true>, dim3((unsigned)npt, (unsigned)ncol), dim3(256), 0, stream(), G.n, G.geo,
                       npt, dpts, ncol, (const double2*)sig, (const double2*)h, wavenumber(enei, eps_medium),
                       2 * M_PI / enei, G.order, G.op.rel_cutoff, G.op.abs_cutoff, 0, (double2*)phi, (double2*)a);
    MNP_HIP(hipGetLastError());
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_HIP(hipFree(dpts));
  });
}

int mnp_dipolestat_potential(mnp_greenret_t g, int64_t npt, const double* pts, int64_t ndip, const double* dip,
                             const mnp_complex* w, double z0, const mnp_complex* wimage, void* phip) {
  return guarded([&] {
    MNP_CHECK(g && pts && dip && w && phip, MNP_ERR_ARG, "dipolestat_potential: NULL argument");
    MNP_CHECK(npt >= 1 && ndip >= 1 && npt * ndip < 65536, MNP_ERR_ARG, "dipolestat_potential: bad sizes");
    const GreenRet& G = *g->g;
    double* dpts = upload(pts, 3 * npt);
    double* ddip = upload(dip, 3 * npt * ndip);
    double2* dw = dev_alloc<double2>(2 * npt);
    MNP_HIP(hipMemcpyAsync(dw, w, sizeof(double2) * npt, hipMemcpyHostToDevice, stream()));
    if (wimage) MNP_HIP(hipMemcpyAsync(dw + npt, wimage, sizeof(double2) * npt, hipMemcpyHostToDevice, stream()));
    const dim3 grid((unsigned)cdiv_up(G.n, 256), (unsigned)(npt * ndip));
    hipLaunchKernelGGL(k_dipstat_phip, grid, dim3(256), 0, stream(), G.n, G.geo.pos, G.geo.nvec, npt, dpts, ddip, dw,
                       z0, wimage ? dw + npt : nullptr, (double2*)phip);
    MNP_HIP(hipGetLastError());
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_HIP(hipFree(dpts));
    MNP_HIP(hipFree(ddip));
    MNP_HIP(hipFree(dw));
  });
}

int mnp_stat_field_own(mnp_greenret_t g, int64_t npt, const double* pts, int64_t ncol, const void* sig,
                       mnp_complex* e) {
  return guarded([&] {
    MNP_CHECK(g && pts && sig && e, MNP_ERR_ARG, "stat_field_own: NULL argument");
    MNP_CHECK(npt >= 1 && ncol >= 1 && ncol < 65536, MNP_ERR_ARG, "stat_field_own: bad sizes");
    const GreenRet& G = *g->g;
    double* dpts = upload(pts, 3 * npt);
    double2* ed = dev_alloc<double2>(3 * ncol);
    double2* hd = dev_alloc<double2>(3 * ncol);
    // no currents (h = nullptr) and k = 0 leave E = -Gp sig
    pt_field(G, 1.0, mnp_complex{0, 0}, npt, dpts, ncol, (const double2*)sig, nullptr, 1, ed, hd);
    MNP_HIP(hipMemcpyAsync(e, ed, sizeof(double2) * 3 * ncol, hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_HIP(hipFree(dpts));
    for (double2* p : {ed, hd}) MNP_HIP(hipFree(p));
  });
}

int mnp_dipole_decayrate(mnp_greenret_t g, double enei, mnp_complex eps_medium, int64_t npt, const double* pts,
                         int64_t ndip, const double* dip, int64_t npinf, const void* pinfty_nvec,
                         const void* pinfty_area, const void* sig2, const void* h2, double* tot, double* rad,
                         double* rad0) {
  return guarded([&] {
    MNP_CHECK(g && pts && dip && pinfty_nvec && pinfty_area && sig2 && h2 && tot && rad, MNP_ERR_ARG,
              "dipole_decayrate: NULL argument");
    MNP_CHECK(npt >= 1 && ndip >= 1 && npinf >= 1 && enei > 0, MNP_ERR_ARG, "dipole_decayrate: bad sizes");
    const GreenRet& G = *g->g;
    const int64_t ncol = npt * ndip;
    MNP_CHECK(ncol < 65536, MNP_ERR_ARG, "dipole_decayrate: too many columns");
    // induced field at each dipole from its own column
    double* dpts = upload(pts, 3 * npt);
    double2* ed = dev_alloc<double2>(3 * ncol);
    double2* hd = dev_alloc<double2>(3 * ncol);
    pt_field(G, enei, eps_medium, npt, dpts, ncol, (const double2*)sig2, (const double2*)h2, 1, ed, hd);
    std::vector<double2> eind(3 * ncol);
    MNP_HIP(hipMemcpyAsync(eind.data(), ed, sizeof(double2) * 3 * ncol, hipMemcpyDeviceToHost, stream()));
    // far fields of the surface charges on the pinfty normals, FF_CHUNK columns at a time
    double2* ef = dev_alloc<double2>(npinf * 3 * FF_CHUNK);
    double2* hf = dev_alloc<double2>(npinf * 3 * FF_CHUNK);
    std::vector<double2> efar(npinf * 3 * ncol), hfar(npinf * 3 * ncol);
    for (int64_t q0 = 0; q0 < ncol; q0 += FF_CHUNK) {
      const int64_t nq = std::min<int64_t>(FF_CHUNK, ncol - q0);
      farfield(G, enei, eps_medium, npinf, (const double*)pinfty_nvec, nq, (const double2*)sig2 + q0 * G.n,
               (const double2*)h2 + 3 * q0 * G.n, ef, hf);
      MNP_HIP(hipMemcpyAsync(efar.data() + npinf * 3 * q0, ef, sizeof(double2) * npinf * 3 * nq,
                             hipMemcpyDeviceToHost, stream()));
      MNP_HIP(hipMemcpyAsync(hfar.data() + npinf * 3 * q0, hf, sizeof(double2) * npinf * 3 * nq,
                             hipMemcpyDeviceToHost, stream()));
      MNP_HIP(hipStreamSynchronize(stream()));
    }
    std::vector<double> pn(3 * npinf), pa(npinf);
    MNP_HIP(hipMemcpyAsync(pn.data(), pinfty_nvec, sizeof(double) * 3 * npinf, hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipMemcpyAsync(pa.data(), pinfty_area, sizeof(double) * npinf, hipMemcpyDeviceToHost, stream()));
    MNP_HIP(hipStreamSynchronize(stream()));
    MNP_HIP(hipFree(dpts));
    MNP_HIP(hipFree(ed));
    MNP_HIP(hipFree(hd));
    MNP_HIP(hipFree(ef));
    MNP_HIP(hipFree(hf));
    using C = std::complex<double>;
    const C eps(eps_medium.re, eps_medium.im), nb = std::sqrt(eps);
    const double k0 = 2 * M_PI / enei;
    const C k = k0 * nb;
    const double gamma = 4.0 / 3.0 * k0 * k0 * k0;
    for (int64_t q = 0; q < ncol; ++q) {
      const int64_t ipt = q % npt, idip = q / npt;
      double p[3], x[3];
      for (int c = 0; c < 3; ++c) {
        p[c] = dip[ipt + npt * (c + 3 * idip)];
        x[c] = pts[c * npt + ipt];
      }
      // total rate: 1 + Im(E_ind . p) / (0.5 nb gamma)
      C ep = 0;
      for (int c = 0; c < 3; ++c) ep += C(eind[c + 3 * q].x, eind[c + 3 * q].y) * p[c];
      tot[q] = std::real(1.0 + ep.imag() / (0.5 * nb * gamma));
      // radiated power of the surface far field plus the dipole's own
      double power = 0;
      for (int64_t d = 0; d < npinf; ++d) {
        const double dir[3] = {pn[d], pn[npinf + d], pn[2 * npinf + d]};
        const C gph = std::exp(-C(0, 1) * k * (dir[0] * x[0] + dir[1] * x[1] + dir[2] * x[2]));
        C hh[3] = {C(dir[1] * p[2] - dir[2] * p[1]) * gph, C(dir[2] * p[0] - dir[0] * p[2]) * gph,
                   C(dir[0] * p[1] - dir[1] * p[0]) * gph};
        C ee[3] = {hh[1] * dir[2] - hh[2] * dir[1], hh[2] * dir[0] - hh[0] * dir[2],
                   hh[0] * dir[1] - hh[1] * dir[0]};
        C E[3], H[3];
        for (int c = 0; c < 3; ++c) {
          const int64_t idx = d + npinf * (c + 3 * q);
          E[c] = C(efar[idx].x, efar[idx].y) + k * k * ee[c] / eps;
          H[c] = C(hfar[idx].x, hfar[idx].y) + k * k * hh[c] / nb;
        }
        const C sx = E[1] * std::conj(H[2]) - E[2] * std::conj(H[1]);
        const C sy = E[2] * std::conj(H[0]) - E[0] * std::conj(H[2]);
        const C sz = E[0] * std::conj(H[1]) - E[1] * std::conj(H[0]);
        power += pa[d] * 0.5 * (dir[0] * sx.real() + dir[1] * sy.real() + dir[2] * sz.real());
      }
      rad[q] = std::real(power / (2 * M_PI * k0) / (0.5 * nb * gamma));
      if (rad0) rad0[q] = std::real(nb * gamma);
    }
  });
}

}  // extern "C"
