// Device-resident retarded Green function of one closed particle.
#pragma once

#include "mnp_common.hpp"

namespace mnp {

constexpr int MNP_MAX_ORDER = 8;
constexpr int MNP_MAX_POLAR = 3 * 16 * 16;

hipStream_t stream();
// This thread's stream override (NULL: the library's stream), e.g. a stream
// under capture; and whether this thread's TimerScopes record nothing (event
// records have no place in a captured graph).
hipStream_t thread_stream();
void set_thread_stream(hipStream_t s);
void pause_timers(bool on);

template <class T>
T* dev_alloc(int64_t count) {
  T* p = nullptr;
  if (count <= 0) count = 1;
  MNP_HIP(hipMalloc((void**)&p, sizeof(T) * (size_t)count));
  return p;
}

// Face geometry, structure of arrays, column-major n x 3 (x, y, z blocks).
struct Geo {
  double* pos = nullptr;    // centroids
  double* nvec = nullptr;   // unit outer normals
  double* t1 = nullptr;     // unit tangent (v1 - v2)/|v1 - v2|
  double* t2 = nullptr;     // nvec x t1
  double* vtx = nullptr;    // vertices, (a*3 + c)*n + i for vertex a, coordinate c
  double* area = nullptr;   // 0.5 |(v1-v2) x (v3-v2)|
  double* qarea = nullptr;  // 0.5 |(v2-v1) x (v3-v1)| (quadrature weights)
  double* rad = nullptr;    // max centroid-vertex distance
};

// Refined (row, col) pairs in column-major order.
struct Pairs {
  int32_t* row = nullptr;
  int32_t* col = nullptr;
  int8_t* kind = nullptr;   // 2 = self (coincident centroids), 1 = near field
  int64_t* diag = nullptr;  // index of pair (i, i), or -1
};

// Assembly output: an FP64 matrix and/or an FP32 mirror of the same values.
struct OutPair {
  double2* d = nullptr;
  float2* f = nullptr;
  __host__ __device__ bool any() const { return d || f; }
  __device__ void store(int64_t idx, double2 v) const {
    if (d) d[idx] = v;
    if (f) f[idx] = make_float2((float)v.x, (float)v.y);
  }
};

// Refinement predicate of Greenfun/+green/refinematrix.m:51-63 with the
// distance of Misc/+misc/pdist2.m (|a|^2+|b|^2-2a.b, d^2 < 1e-10 -> 0):
// returns 2 for coincident centroids, 1 for a near-field pair, else 0.
__device__ __forceinline__ int refine_kind(const Geo& geo, int64_t n, int64_t i, int64_t j,
                                           double rel, double absc) {
#pragma clang fp contract(off)
  const double xi = geo.pos[i], yi = geo.pos[n + i], zi = geo.pos[2 * n + i];
  const double xj = geo.pos[j], yj = geo.pos[n + j], zj = geo.pos[2 * n + j];
  double d2 = (xi * xi + yi * yi + zi * zi) + (xj * xj + yj * yj + zj * zj) -
              2 * (xi * xj + yi * yj + zi * zj);
  if (d2 < 1e-10) d2 = 0;
  const double d = sqrt(d2);
  if (d == 0) return 2;
  const double dd = d - geo.rad[j];
  return (dd < absc || dd / geo.rad[i] < rel) ? 1 : 0;
}

// Several parts with their own media (Particles/@comparticle): faces of part
// q are [part_off[q], part_off[q + 1]); inout[2 q], inout[2 q + 1] its inside
// and outside media (1-based); bit r of cmask[q] set when part r belongs to
// the closed surface of part q (Particles/@comparticle/closed.m; 0: part q
// has no closed surface).
struct CompoundSpec {
  int nparts;
  const int64_t* part_off;
  const int32_t* inout;
  const uint32_t* cmask;
};
constexpr int MNP_MAX_MED = 16;    // media of a compound
constexpr int MNP_MAX_PARTS = 32;  // parts of a compound (bits of cmask)
// Wavenumber of each medium, by value into the assembly kernels.
struct KTab {
  double2 k[MNP_MAX_MED];
};

struct GreenRet {
  int64_t n = 0;
  int64_t nnz = 0;
  int order = 5;
  int npolq = 0;
  mnp_green_options op;
  Geo geo;
  Pairs pairs;
  int64_t* off = nullptr;  // column offsets into the pair list (n + 1)
  double* gtab = nullptr;  // nnz x (order + 1)
  double* ftab = nullptr;  // nnz x 3 x (order + 1)
  // mesh connectivity and the quadrature rules in global memory (the
  // __constant__ copies are private to greenret.hip's code object)
  int64_t nverts = 0;
  std::vector<int64_t> faces_h;  // host n x 3, column-major, 0-based
  double* q18 = nullptr;         // device [x(28) | y(28) | w(28)]
  double* qpol = nullptr;        // device [x | y | w], MNP_MAX_POLAR each

  // compound particle: device per-face media (1-based) and part, the closed
  // surface of each part; all null for one particle
  int nparts = 1;
  int nmed = 0;                    // largest medium index of a compound
  bool con12 = false;              // some part's inside medium is some part's outside medium
  int32_t* fin = nullptr;
  int32_t* fout = nullptr;
  int32_t* fpart = nullptr;
  uint32_t* cmask = nullptr;
  std::vector<int32_t> fin_h, fout_h;
  bool compound() const { return fin != nullptr; }

  // like: take the refined pairs of that Green function (same faces) instead
  // of selecting them for this mesh; cs: the parts of a compound
  GreenRet(const double* verts, int64_t nv, const int64_t* faces, int64_t nf,
           const mnp_green_options& op, const GreenRet* like = nullptr, const CompoundSpec* cs = nullptr);
  ~GreenRet();
  GreenRet(const GreenRet&) = delete;
  GreenRet& operator=(const GreenRet&) = delete;

  // G(k) into g and F(k) + sgn2pi * I into h (either may be empty).
  void assemble(double2 k, double sgn2pi, OutPair g, OutPair h, int64_t ld) const;
  // Rows 0 .. nrows-1 of the same matrices (nrows x n, leading dimension ld).
  void assemble_rows(double2 k, double sgn2pi, OutPair g, OutPair h, int64_t nrows, int64_t ld) const;
  // dG/dk into g and dF/dk (= dH/dk) into h, n x n.
  void assemble_dk(double2 k, OutPair g, OutPair h, int64_t ld) const;
  // Compound: side 1 writes G1 = g{1,1} - g{2,1} and H1 = F1 + 2 pi, side 2
  // G2 = g{2,2} - g{1,2} and H2 = F2 - 2 pi (BEM/@bemret/private/initmat.m:
  // 27-31), kt.k[m - 1] the wavenumber of medium m.
  void assemble_side(int side, const KTab& kt, OutPair g, OutPair h, int64_t ld) const;
};

}  // namespace mnp

// Opaque handle behind mnp_greenret_t.
struct mnp_greenret_s {
  mnp::GreenRet* g;
};
