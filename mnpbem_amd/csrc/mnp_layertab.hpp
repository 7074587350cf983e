// Tables of the reflected Green function of one interface (MNPBEM
// Greenfun/@greentablayer) as seen by the kernels: the normalised table on
// its (r, z1) grid and its bilinear interpolation (interp2.m, finterp.m),
// shared by the dense substrate path and the H-matrices of the iterative
// substrate solver.
#pragma once

#include "mnp_common.hpp"

namespace mnp {

constexpr int NQ = 15;  // G, Fr, Fz x (p, ss, hs, sh, hh)

// ------------------------------------------------------ interpolation
struct TabView {
  int64_t nr, nz;
  const double* r;
  const double* z1;
  const double2* v;  // (i nz + j) 15 + q
  double z0, rmin, zmin;
  int* err;          // set when a point lies outside the table
};

__device__ inline double round_z(double z, double z0, double zmin) {  // round.m
  const double d = fabs(z - z0);
  if (d <= zmin) {
    const double s = z > z0 ? 1.0 : (z < z0 ? -1.0 : 0.0);
    return z0 + s * zmin;
  }
  return z;
}

__device__ inline int64_t upper_index(const double* x, int64_t n, double v) {  // #{x <= v}
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t m = (lo + hi) / 2;
    if (x[m] <= v)
      lo = m + 1;
    else
      hi = m;
  }
  return lo;
}

// interp2.m:17-44: G, Fr, Fz at (r, z1, z2), and the clamped r.
__device__ inline double tab_interp(const TabView& T, double r, double z1, double z2, double2 o[NQ]) {
#pragma clang fp contract(off)
  r = fmax(T.rmin, r);
  z1 = round_z(z1, T.z0, T.zmin);
  z2 = round_z(z2, T.z0, T.zmin);
  const double z = z1 + fabs(round_z(z2, T.z0, T.zmin) - T.z0);
  const double zm = fabs(z - T.z0);
  const double d = sqrt(r * r + zm * zm);
  int64_t ix = upper_index(T.r, T.nr, r), iy = upper_index(T.z1, T.nz, z);
  if (ix == 0 || iy == 0 || r > T.r[T.nr - 1] || z > T.z1[T.nz - 1]) {
    atomicOr(T.err, 1);
    ix = ix == 0 ? 1 : ix;
    iy = iy == 0 ? 1 : iy;
  }
  ix = (ix < T.nr - 1 ? ix : T.nr - 1) - 1;
  iy = (iy < T.nz - 1 ? iy : T.nz - 1) - 1;
  const double xx = (r - T.r[ix]) / (T.r[ix + 1] - T.r[ix]);
  const double yy = (z - T.z1[iy]) / (T.z1[iy + 1] - T.z1[iy]);
  const double w0 = (1 - xx) * (1 - yy), w1 = xx * (1 - yy), w2 = (1 - xx) * yy, w3 = xx * yy;
  const double2* v00 = T.v + (ix * T.nz + iy) * NQ;
  const double2* v10 = v00 + T.nz * NQ;
  const double2* v01 = v00 + NQ;
  const double2* v11 = v10 + NQ;
  const double d3 = d * d * d;
  for (int q = 0; q < NQ; ++q) {
    const double fx = w0 * v00[q].x + w1 * v10[q].x + w2 * v01[q].x + w3 * v11[q].x;
    const double fy = w0 * v00[q].y + w1 * v10[q].y + w2 * v01[q].y + w3 * v11[q].y;
    if (q < 5)
      o[q] = mk<double>(fx / d, fy / d);
    else if (q < 10)
      o[q] = mk<double>(fx * r / d3, fy * r / d3);
    else
      o[q] = mk<double>(fx * zm / d3, fy * zm / d3);
  }
  return r;
}

struct LayGeo {
  double z0, rmin, zmin;
};

// tab_interp restricted to component c (p, ss, hs, sh, hh = 0..4): g, and
// for F (wantF) the radial and z parts fr, fz; returns the clamped r.
__device__ inline double tab_interp_c(const TabView& T, double r, double z1, double z2, int c, int wantF, double2& g,
                                      double2& fr, double2& fz) {
#pragma clang fp contract(off)
  r = fmax(T.rmin, r);
  z1 = round_z(z1, T.z0, T.zmin);
  z2 = round_z(z2, T.z0, T.zmin);
  const double z = z1 + fabs(round_z(z2, T.z0, T.zmin) - T.z0);
  const double zm = fabs(z - T.z0);
  const double d = sqrt(r * r + zm * zm);
  int64_t ix = upper_index(T.r, T.nr, r), iy = upper_index(T.z1, T.nz, z);
  if (ix == 0 || iy == 0 || r > T.r[T.nr - 1] || z > T.z1[T.nz - 1]) {
    atomicOr(T.err, 1);
    ix = ix == 0 ? 1 : ix;
    iy = iy == 0 ? 1 : iy;
  }
  ix = (ix < T.nr - 1 ? ix : T.nr - 1) - 1;
  iy = (iy < T.nz - 1 ? iy : T.nz - 1) - 1;
  const double xx = (r - T.r[ix]) / (T.r[ix + 1] - T.r[ix]);
  const double yy = (z - T.z1[iy]) / (T.z1[iy + 1] - T.z1[iy]);
  const double w0 = (1 - xx) * (1 - yy), w1 = xx * (1 - yy), w2 = (1 - xx) * yy, w3 = xx * yy;
  const double2* v00 = T.v + (ix * T.nz + iy) * NQ;
  const double2* v10 = v00 + T.nz * NQ;
  const double2* v01 = v00 + NQ;
  const double2* v11 = v10 + NQ;
  auto at = [&](int q) {
    return mk<double>(w0 * v00[q].x + w1 * v10[q].x + w2 * v01[q].x + w3 * v11[q].x,
                      w0 * v00[q].y + w1 * v10[q].y + w2 * v01[q].y + w3 * v11[q].y);
  };
  if (!wantF) {
    const double2 f = at(c);
    g = mk<double>(f.x / d, f.y / d);
  } else {
    const double d3 = d * d * d;
    const double2 a = at(5 + c), b = at(10 + c);
    fr = mk<double>(a.x * r / d3, a.y * r / d3);
    fz = mk<double>(b.x * zm / d3, b.y * zm / d3);
  }
  return r;
}

// Reflected G (isF = 0) or F = n_i . Gp (isF = 1) of component c between the
// centroids of faces i (row) and j (column), times the column's area: one
// entry of the reflected matrices at the centroids (initrefl2.m:21-37).
// pos, nvec (n x 3 column-major) and area in the caller's ordering.
__device__ inline double2 refl_entry(const TabView& T, const double* pos, const double* nvec, const double* area,
                                     int64_t n, int64_t i, int64_t j, int c, int isF) {
#pragma clang fp contract(off)
  const double x = pos[i] - pos[j], y = pos[n + i] - pos[n + j];
  const double r = sqrt(x * x + y * y);
  double2 g = mk<double>(0, 0), fr = g, fz = g;
  tab_interp_c(T, r, pos[2 * n + i], pos[2 * n + j], c, isF, g, fr, fz);
  const double a = area[j];
  if (!isF) return cscale<double>(g, a);
  const double rr = fmax(r, 1e-10);
  const double2 gx = cscale<double>(cscale<double>(cscale<double>(fr, x), 1.0 / rr), a);
  const double2 gy = cscale<double>(cscale<double>(cscale<double>(fr, y), 1.0 / rr), a);
  const double2 gz = cscale<double>(fz, a);
  const double nx = nvec[i], ny = nvec[n + i], nz = nvec[2 * n + i];
  return mk<double>(nx * gx.x + ny * gy.x + nz * gz.x, nx * gx.y + ny * gy.y + nz * gz.y);
}

// The reflected Green function of one particle at one wavelength as the
// H-matrices of the iterative substrate solver use it: the table, and the
// refined values (initrefl2.m) of its near pairs (rows, cols: face indices)
// and polar-rule diagonal elements (faces id), component-major:
// pG[c np + p], pF[c np + p], dG[c nd + t], dF[c nd + t].
struct LayerPairs {
  TabView T;
  int64_t np, nd;
  const int32_t *rows, *cols, *id;
  const double2 *pG, *pF, *dG, *dF;
  const int64_t* coff;  // pairs of column j: [coff[j], coff[j + 1]), rows increasing
  const int32_t* dmap;  // diagonal element of face i: dmap[i] (-1: none)
};
struct GreenRetLayer;  // layer.hip
// Tables at enei and the refined values of gl's pairs (device buffers owned
// by gl, valid until its next call).
LayerPairs greenretlayer_pairs(GreenRetLayer* gl, double enei, double2 e1, double2 e2);
// Raises if a reflected entry evaluated since the last greenretlayer_pairs
// call fell outside the table.
void greenretlayer_check(GreenRetLayer* gl);
GreenRetLayer* greenretlayer_of(mnp_greenretlayer_t h);

}  // namespace mnp
