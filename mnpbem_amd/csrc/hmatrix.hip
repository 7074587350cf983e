// H-matrices of the retarded Green function on gfx950 (MNPBEM Greenfun/+aca,
// Greenfun/hmatrices, mex/hmatgreenret.cpp, mex/hlib).
//
// The block structure (cluster order, full leaf blocks, admissible blocks)
// comes from the host, as in the reference where Matlab builds the tree and
// hands it to the MEX kernels.  On the GPU:
//  - k_hfull: full blocks element-wise, with the refined near-field/self
//    entries (greenret eval2) and +-2 pi on coincident centroids (H1/H2);
//  - k_aca:   one wave (blocks with m + n <= 320, four per workgroup) or one
//    workgroup per admissible block runs the partially pivoted ACA of
//    mex/hlib/aca.cpp:158-218 on the collocation kernel of
//    mex/acagreen/acagreen.cpp:79-127, in FP64, rank <= min(kmax, m, n);
//  - k_hmv_cols / k_hmv_rows: y = H x with blocks and factors stored in
//    FP64 or FP32 and FP64 accumulation: t = R^T x over the column groups
//    (the R_b^T of all low-rank blocks of one column cluster side by side,
//    copied once per fill into contiguous slices, k_hmv_pack_cols), then
//    one workgroup per row tile
//    (<= 32 rows of a leaf) streams the tile's packed row panel (its slices
//    of the full blocks and its L_b segments, copied once per fill into one
//    contiguous column-major panel, k_hmv_pack) against the rows of x and t
//    its columns name (deterministic, no atomics).
#include <algorithm>
#include <cmath>
#include <vector>

#include "mnp_hmatrix.hpp"
#include "mnp_layertab.hpp"

namespace mnp {

// ------------------------------------------------------------ kernels
// Collocation kernel between cluster indices i (row) and j (column).
__device__ __forceinline__ double2 colloc(const double* cpos, const double* cnvec, const double* carea,
                                          int64_t n, int64_t i, int64_t j, double2 k, int isF) {
  const double x = cpos[i] - cpos[j], y = cpos[n + i] - cpos[n + j], z = cpos[2 * n + i] - cpos[2 * n + j];
  const double d = sqrt(x * x + y * y + z * z);
  const double2 fac = cexp_ikd(k, d);
  const double a = carea[j];
  if (!isF) return mk<double>(fac.x / d * a, fac.y / d * a);
  const double in = x * cnvec[i] + y * cnvec[n + i] + z * cnvec[2 * n + i];
  const double2 t = mk<double>(-k.y - 1.0 / d, k.x);  // i k - 1/d
  const double2 v = cmul<double>(t, fac);
  const double s = in / (d * d) * a;
  return mk<double>(v.x * s, v.y * s);
}

// The reflected part of a substrate H-matrix (Greenfun/+aca/
// @compgreenretlayer/private/eval2.m): component comp of the tabulated
// reflected G or F, the refined values of its near pairs (pv, column
// offsets coff over rows) and polar diagonal elements (dv by dmap), and
// whether the direct part is kept (p, ss, hh) or not (hs, sh).
struct LayerFill {
  int on, direct, comp;
  TabView T;
  const int64_t* coff;
  const int32_t* rows;
  const int32_t* dmap;
  const double2* pv;  // refined pair values of this component, np
  const double2* dv;  // refined diagonal values of this component, nd
};

// Entries of admissible blocks for ACA, (i, j) in cluster order: the direct
// collocation kernel, and for a substrate the reflected component at the
// centroids (admissible blocks hold no refined pairs).
struct EntDirect {
  const double *cpos, *cnvec, *carea;
  int64_t n;
  double2 wav;
  int isF;
  __device__ double2 operator()(int64_t i, int64_t j) const { return colloc(cpos, cnvec, carea, n, i, j, wav, isF); }
};
struct EntLayer {
  EntDirect d;
  int direct, comp;
  TabView T;
  __device__ double2 operator()(int64_t i, int64_t j) const {
    const double2 r = refl_entry(T, d.cpos, d.cnvec, d.carea, d.n, i, j, comp, d.isF);
    return direct ? cadd<double>(d(i, j), r) : r;
  }
};

// Full blocks: entry (i, j) of block b -> faces fi = perm[r0+i], fj = perm[c0+j];
// with a substrate (L.on) the reflected part of component L.comp is added:
// the refined value of a near pair or polar diagonal element, else the
// centroid value (initrefl2.m), and the direct part is kept only where it
// belongs (L.direct).
template <class TS>
__global__ void __launch_bounds__(256) k_hfull(int64_t nb, const int32_t* blk, const int64_t* voff,
                                               const int32_t* perm, int64_t n, Geo geo, Pairs pr,
                                               const int64_t* off, int64_t nnz, int order,
                                               const double* gt, const double* ft, double2 k,
                                               int key, LayerFill L, cplx<TS>* vals) {
  const int64_t b = blockIdx.x;
  const int r0 = blk[4 * b], r1 = blk[4 * b + 1], c0 = blk[4 * b + 2], c1 = blk[4 * b + 3];
  const int m = r1 - r0, w = c1 - c0;
  cplx<TS>* out = vals + voff[b];
  const double2 ik = mk<double>(-k.y, k.x);
  // refined pair of (fi, fj) in a column's sorted row list, or -1
  auto find = [](const int32_t* rows, int64_t lo, int64_t hi, int64_t fi) -> int64_t {
    while (lo <= hi) {
      const int64_t mid = (lo + hi) >> 1;
      const int32_t rr = rows[mid];
      if (rr == fi) return mid;
      if (rr < fi)
        lo = mid + 1;
      else
        hi = mid - 1;
    }
    return -1;
  };
  for (int64_t e = threadIdx.x; e < (int64_t)m * w; e += blockDim.x) {
    const int i = (int)(e % m), j = (int)(e / m);
    const int64_t fi = perm[r0 + i], fj = perm[c0 + j];
    double2 v = mk<double>(0, 0);
    if (!L.on || L.direct) {
      const double dx = geo.pos[fi] - geo.pos[fj], dy = geo.pos[n + fi] - geo.pos[n + fj],
                   dz = geo.pos[2 * n + fi] - geo.pos[2 * n + fj];
      const double d = fmax(sqrt(dx * dx + dy * dy + dz * dz), 2.220446049250313e-16);
      const double2 ph = cexp_ikd(k, d);
      const int64_t p = find(pr.row, off[fj], off[fj + 1] - 1, fi);  // rows of column fj sorted
      if (p >= 0) {
        double2 pw = mk<double>(1, 0), s = mk<double>(0, 0);
        for (int o = 0; o <= order; ++o) {
          double c;
          if (key == MNP_KEY_G)
            c = gt[o * nnz + p];
          else
            c = geo.nvec[fi] * ft[(o * 3 + 0) * nnz + p] + geo.nvec[n + fi] * ft[(o * 3 + 1) * nnz + p] +
                geo.nvec[2 * n + fi] * ft[(o * 3 + 2) * nnz + p];
          s = cadd<double>(s, cscale<double>(pw, c));
          pw = cmul<double>(pw, ik);
        }
        v = cmul<double>(s, ph);
      } else if (key == MNP_KEY_G) {
        const double g = (1.0 / d) * geo.area[fj];
        v = mk<double>(g * ph.x, g * ph.y);
      } else {
        const double id2 = 1.0 / (d * d);
        const double2 fk = mk<double>((-k.y - 1.0 / d) * id2, k.x * id2);
        const double in = dx * geo.nvec[fi] + dy * geo.nvec[n + fi] + dz * geo.nvec[2 * n + fi];
        v = cmul<double>(mk<double>(fk.x * in * geo.area[fj], fk.y * in * geo.area[fj]), ph);
      }
      if (fi == fj) v.x += key == MNP_KEY_H1 ? 2 * M_PI : key == MNP_KEY_H2 ? -2 * M_PI : 0.0;
    }
    if (L.on) {
      const int64_t q = find(L.rows, L.coff[fj], L.coff[fj + 1] - 1, fi);
      double2 r;
      if (q >= 0)
        r = L.pv[q];
      else if (fi == fj && L.dmap[fi] >= 0)
        r = L.dv[L.dmap[fi]];
      else
        r = refl_entry(L.T, geo.pos, geo.nvec, geo.area, n, fi, fj, L.comp, key != MNP_KEY_G);
      v = cadd<double>(v, r);
    }
    out[(int64_t)j * m + i] = mk<TS>((TS)v.x, (TS)v.y);
  }
}

// ACA pivots: the first largest |Re| + |Im| (lowest index on ties, as
// izamax); idx < 0 means none yet.  z is the entry itself.
struct Piv {
  double v;
  int idx;
  double2 z;
};
__device__ __forceinline__ void piv_merge(Piv& a, double v, int idx, double2 z) {
  if (idx >= 0 && (a.idx < 0 || v > a.v || (v == a.v && idx < a.idx))) {
    a.v = v;
    a.idx = idx;
    a.z = z;
  }
}
__device__ __forceinline__ void piv_take(Piv& a, double v, int idx, double2 z) {
  if (v > a.v) {  // ascending indices per thread: strict keeps the first
    a.v = v;
    a.idx = idx;
    a.z = z;
  }
}

// Orders the memory operations of a group of G threads: G = 64 is one wave
// (program order already holds between its lanes; this only stops the
// compiler from moving loads and stores across), G = 256 the workgroup.
template <int G>
__device__ __forceinline__ void gsync() {
  if constexpr (G == 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Sums s[0..NS) and merges the pivot p over the group; every thread gets
// the same result (xor butterflies are symmetric, the four wave partials
// are read in one order).  For G = 256 the caller guarantees that no thread
// still reads red/pred from an earlier call (a gsync lies between calls).
template <int G, int NS>
__device__ __forceinline__ void greduce(double (&s)[NS], Piv& p, double* red, Piv* pred) {
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int q = 0; q < NS; ++q) s[q] += __shfl_xor(s[q], o);
    const double v = __shfl_xor(p.v, o);
    const int idx = __shfl_xor(p.idx, o);
    const double zx = __shfl_xor(p.z.x, o), zy = __shfl_xor(p.z.y, o);
    piv_merge(p, v, idx, mk<double>(zx, zy));
  }
  if constexpr (G == 256) {
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
      for (int q = 0; q < NS; ++q) red[4 * q + wv] = s[q];
      pred[wv] = p;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NS; ++q) s[q] = red[4 * q] + red[4 * q + 1] + red[4 * q + 2] + red[4 * q + 3];
    p = pred[0];
    for (int u = 1; u < 4; ++u) piv_merge(p, pred[u].v, pred[u].idx, pred[u].z);
  }
}

// Buffer loads of one complex value at byte offset voff + soff of a buffer
// resource: offsets past its size read zero (the hardware bounds check).
template <class TS>
__device__ __forceinline__ cplx<TS> bload(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ cplx<float> bload<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  return mk<float>(__uint_as_float(v[0]), __uint_as_float(v[1]));
}
template <>
__device__ __forceinline__ cplx<double> bload<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  return mk<double>(__hiloint2double((int)v[1], (int)v[0]), __hiloint2double((int)v[3], (int)v[2]));
}
// Two consecutive complex values (one 16-byte load for FP32 storage).
template <class TS>
__device__ __forceinline__ void bload2(__amdgpu_buffer_rsrc_t r, int voff, int soff, cplx<TS>& c0, cplx<TS>& c1);
template <>
__device__ __forceinline__ void bload2<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff, cplx<float>& c0,
                                              cplx<float>& c1) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  c0 = mk<float>(__uint_as_float(v[0]), __uint_as_float(v[1]));
  c1 = mk<float>(__uint_as_float(v[2]), __uint_as_float(v[3]));
}
template <>
__device__ __forceinline__ void bload2<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff, cplx<double>& c0,
                                               cplx<double>& c1) {
  c0 = bload<double>(r, voff, soff);
  c1 = bload<double>(r, voff + 16, soff);
}
constexpr int BUF_WORD3 = 0x00020000;  // raw (untyped) buffer access on gfx9

// v - sum_{l < k} P(c, l) coef[l], subtracted in order of l: column l of P
// at byte offset l ldb of the buffer resource rs, entry c at byte offset
// voff (the columns are this thread's own earlier results, coef is shared).
// The loads go out eight at a time with the column offsets in scalar
// registers; the last group of eight reads column k - 1 in place of the
// columns past it and skips their terms, so no load waits alone.
__device__ __forceinline__ double2 aca_residual(double2 v, __amdgpu_buffer_rsrc_t rs, int voff, int ldb, int k,
                                                const double2* coef) {
  for (int l = 0; l < k; l += 8) {
    double2 t[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) t[j] = bload<double>(rs, voff, min(l + j, k - 1) * ldb);
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (l + j < k) v = csub<double>(v, cmul<double>(t[j], coef[l + j]));
  }
  return v;
}

// Partially pivoted ACA of admissible block b (rows [r0, r1), cols [c0, c1)
// in cluster order) by a group of G threads, thread tid of the group.
// Scratch: A (m x kb) then B (n x kb) at soff[b]; rank[b] on exit.  Rows
// tid + G t are this thread's; the pivot rows already taken are kept as
// bits t of used0/used1 (m <= 128 G, checked on the host).
template <int G, class Ent>
__device__ __forceinline__ void aca_block(int b, int tid, const int32_t* blk, const int64_t* soff, const Ent& E,
                                          double tol, int kmax, double2* scratch, int32_t* rank, double2* coef,
                                          double* red, Piv* pred) {
  b = __builtin_amdgcn_readfirstlane(b);
  const int r0 = blk[4 * b], r1 = blk[4 * b + 1], c0 = blk[4 * b + 2], c1 = blk[4 * b + 3];
  const int m = r1 - r0, w = c1 - c0;
  const int kb = min(kmax, min(m, w));
  double2* A = scratch + soff[b];
  double2* B = A + (int64_t)m * kb;
  const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc((void*)A, (short)0, m * kb * 16, BUF_WORD3);
  const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc((void*)B, (short)0, w * kb * 16, BUF_WORD3);
  uint64_t used0 = 0, used1 = 0;  // bits t < 64, 64 <= t < 128
  int r = 0, k = 0;
  double nsum = 0;
  for (k = 0; k < kb; ++k) {
    // row r of the residual -> B(:, k), with A(r, l) as coefficients; the
    // cutoff looks at the row of the matrix itself
    for (int l = tid; l < k; l += G) coef[l] = A[(int64_t)l * m + r];
    gsync<G>();
    double s1[1] = {0};
    Piv pc{-1.0, -1, mk<double>(0, 0)};
    for (int c = tid; c < w; c += G) {
      double2 v = E(r0 + r, c0 + c);
      s1[0] += v.x * v.x + v.y * v.y;
      v = aca_residual(v, rb, 16 * c, 16 * w, k, coef);
      B[(int64_t)k * w + c] = v;
      piv_take(pc, cabs1<double>(v), c, v);
    }
    greduce<G, 1>(s1, pc, red, pred);
    if (sqrt(s1[0]) < 1e-10) {  // zero row: the rank counts it with A(:, k) = 0
      for (int i = tid; i < m; i += G) A[(int64_t)k * m + i] = mk<double>(0, 0);
      break;
    }
    const int cp = pc.idx;
    const double2 scale = cdiv<double>(mk<double>(1, 0), pc.z);
    // column cp of the residual -> A(:, k), with the scaled B(cp, l) as
    // coefficients; their loads go out with those of the scaling pass
    double2 cf0 = mk<double>(0, 0), cf1 = mk<double>(0, 0);
    if (tid < k) cf0 = B[(int64_t)tid * w + cp];
    if (tid + G < k) cf1 = B[(int64_t)(tid + G) * w + cp];
    double s2[2] = {0, 0};  // |A(:, k)|^2, |B(:, k)|^2
    for (int c = tid; c < w; c += G) {
      const double2 v = cmul<double>(B[(int64_t)k * w + c], scale);
      B[(int64_t)k * w + c] = v;
      s2[1] += v.x * v.x + v.y * v.y;
    }
    if (tid < k) coef[tid] = cf0;
    if (tid + G < k) coef[tid + G] = cf1;
    gsync<G>();
    if (r % G == tid) {
      const int tr = r / G;
      if (tr < 64)
        used0 |= 1ull << tr;
      else
        used1 |= 1ull << (tr - 64);
    }
    Piv pr{-1.0, -1, mk<double>(0, 0)};
    int t = 0;
    for (int i = tid; i < m; i += G, ++t) {
      double2 v = E(r0 + i, c0 + cp);
      v = aca_residual(v, ra, 16 * i, 16 * m, k, coef);
      A[(int64_t)k * m + i] = v;
      s2[0] += v.x * v.x + v.y * v.y;
      // next pivot row: izamax of the column with the used rows set to zero
      const bool u = ((t < 64 ? used0 >> t : used1 >> (t - 64)) & 1ull) != 0;
      piv_take(pr, u ? 0.0 : cabs1<double>(v), i, v);
    }
    greduce<G, 2>(s2, pr, red, pred);
    r = pr.idx;
    const double nk = sqrt(s2[0]) * sqrt(s2[1]);
    if (nk < tol * nsum || k + 1 >= m) break;
    nsum = sqrt(nsum * nsum + nk * nk);
  }
  if (tid == 0) rank[b] = min(k + 1, kb);
}

// One ACA work item per workgroup of 256 threads: item[1] == -2 marks one
// large block for the whole workgroup (item[0]), otherwise item[0..3] are up
// to four small blocks (-1: none), one per wave.
template <class Ent>
__global__ void __launch_bounds__(256) k_aca(const int32_t* items, const int32_t* blk, const int64_t* soff, Ent E,
                                             double tol, int kmax, double2* scratch, int32_t* rank) {
  __shared__ double2 coef[4][128];
  __shared__ double red[8];
  __shared__ Piv pred[4];
  const int32_t* it = items + 4 * (int64_t)blockIdx.x;
  if (it[1] == -2) {
    aca_block<256>(it[0], threadIdx.x, blk, soff, E, tol, kmax, scratch, rank, coef[0], red, pred);
    return;
  }
  const int wv = threadIdx.x >> 6;
  const int32_t b = it[wv];
  if (b >= 0)
    aca_block<64>(b, threadIdx.x & 63, blk, soff, E, tol, kmax, scratch, rank, coef[wv], nullptr, nullptr);
}

// Compacts the ACA scratch of block b = order[blockIdx.x] into its factors
// (FP32 or FP64): L (m x r) as consecutive row-tile segments, each (rows of
// the tile) x r column-major, so that the product kernel reads one
// contiguous segment per tile; then R^T (r x n, column-major: the r entries
// of each column index contiguous), transposed through LDS in chunks of JC
// columns so that both the reads of B and the writes of R^T are contiguous.
// rowtile[c] = tile of cluster row c, tilebeg[t] = first row.
template <class TS>
__global__ void __launch_bounds__(256) k_aca_store(const int32_t* order, const int32_t* blk, const int64_t* soff,
                                                   const double2* scratch, const int32_t* rank, int kmax,
                                                   const int64_t* loff, const int32_t* rowtile,
                                                   const int32_t* tilebeg, cplx<TS>* pool) {
  constexpr int JC = 32768 / (128 * (int)sizeof(cplx<TS>));  // columns per chunk (r <= 128)
  __shared__ cplx<TS> tile[JC * 128];
  const int64_t b = order[blockIdx.x];
  const int r0 = blk[4 * b];
  const int m = blk[4 * b + 1] - r0, w = blk[4 * b + 3] - blk[4 * b + 2];
  const int kb = min(kmax, min(m, w));
  const int r = rank[b];
  const double2* A = scratch + soff[b];
  const double2* B = A + (int64_t)m * kb;
  cplx<TS>* L = pool + loff[b];
  cplx<TS>* RT = L + (int64_t)m * r;
  for (int64_t e = threadIdx.x; e < (int64_t)m * r; e += blockDim.x) {
    const int i = (int)(e % m), l = (int)(e / m);
    const int row = r0 + i, t = rowtile[row], tb = tilebeg[t], ms = tilebeg[t + 1] - tb;
    L[(int64_t)(tb - r0) * r + (int64_t)l * ms + (row - tb)] = mk<TS>((TS)A[e].x, (TS)A[e].y);
  }
  for (int j0 = 0; j0 < w; j0 += JC) {
    const int jc = min(JC, w - j0);
    __syncthreads();
    for (int e = threadIdx.x; e < r * JC; e += blockDim.x) {
      const int l = e / JC, jj = e % JC;
      if (jj < jc) {
        const double2 v = B[(int64_t)l * w + j0 + jj];
        tile[jj * r + l] = mk<TS>((TS)v.x, (TS)v.y);
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < jc * r; e += blockDim.x) RT[(int64_t)j0 * r + e] = tile[e];
  }
}

// Products of one batched launch (up to HMV_MAXB H-matrices of one size and
// storage precision, e.g. G1, G2, H1, H2 of the iterative solver): per
// product m its plan and buffers; work units (first-phase items, partial
// sums, row tiles) of product m are [beg[m], beg[m + 1]) of the launch.
constexpr int HMV_MAXB = 4;
struct HmvBatch {
  int nm;
  int64_t n;
  int64_t ibeg[HMV_MAXB + 1], mbeg[HMV_MAXB + 1], tbeg[HMV_MAXB + 1];
  const int32_t* perm[HMV_MAXB];
  const double2* x[HMV_MAXB];  // face order, n x nrhs
  double2* y[HMV_MAXB];        // face order, n x nrhs
  double2* xt[HMV_MAXB];       // [x in cluster order; t], row stride nrhs
  double2* tpart[HMV_MAXB];
  const int64_t* ioff[HMV_MAXB];
  const int32_t* imeta[HMV_MAXB];
  const int32_t* multi[HMV_MAXB];
  const void* rpanel[HMV_MAXB];
  const int32_t* tilebeg[HMV_MAXB];
  const int64_t* tcol[HMV_MAXB];
  const int64_t* poff[HMV_MAXB];
  const int32_t* cidx[HMV_MAXB];
  const void* panel[HMV_MAXB];
  int64_t trow0[HMV_MAXB];  // first row of t in xt (= n)
};

// Product of work unit u (wave-uniform): the last m with beg[m] <= u.
__device__ __forceinline__ int batch_of(const int64_t* beg, int nm, int64_t u) {
  int m = 0;
  while (m + 1 < nm && u >= beg[m + 1]) ++m;
  return m;
}

// x (face order, FP64, n x nrhs) -> xt (cluster order, right-hand sides
// interleaved: xt[c * nrhs + q]); the rows of t follow in the same buffer.
// blockIdx.z = product.
__global__ void k_to_cluster(HmvBatch B, int nrhs) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  const int m = blockIdx.z;
  if (i < B.n) B.xt[m][i * nrhs + q] = B.x[m][q * B.n + B.perm[m][i]];
}

constexpr int HMV_UNROLL = 8;  // independent loads in flight per lane
constexpr int HMV_KC = 64;      // right-hand-side rows staged in LDS per wave

// The product kernels are instantiated for NR = 4 or 8 right-hand sides (the
// pass's nrhs <= NR): LDS and accumulators are sized by NR, and the columns
// q >= nrhs are zero in LDS, so the inner loops carry no per-column branches.
// Matrix and vector loads read clamped, always valid addresses and are
// zeroed where they fall outside; a predicated load would be compiled to a
// branch that waits for it at once.

// Stages rows [c0, c0 + kc) of the interleaved right-hand sides B (row j at
// B[j * nrhs], FP64 or FP32) into this wave's LDS buffer as FP64 rows of NR
// (zero beyond nrhs), all loads in flight before the first store; LDS
// operations of one wave complete in order, so the wave's later reads see
// these writes.
template <int NR, class TB>
__device__ __forceinline__ void stage_rows(double2* bs, const TB* B, int c0, int kc, int nrhs, int lane) {
  const TB* src = B + (int64_t)c0 * nrhs;
  constexpr int PER = HMV_KC * NR / 64;
  double2 v[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = lane + 64 * i, row = e / NR, q = e % NR;
    const TB x = src[min(row, kc - 1) * nrhs + min(q, nrhs - 1)];
    v[i] = (row < kc && q < nrhs) ? mk<double>(x.x, x.y) : mk<double>(0, 0);
  }
#pragma unroll
  for (int i = 0; i < PER; ++i) bs[lane + 64 * i] = v[i];
  __builtin_amdgcn_wave_barrier();
}

// acc0[q] += a0 * bs[q], acc1[q] += a1 * bs[q]: one LDS read of bs serves two rows
template <class TS, int NR>
__device__ __forceinline__ void lds_fma2(double2 (&acc0)[NR], double2 (&acc1)[NR], const cplx<TS> a0,
                                         const cplx<TS> a1, const double2* bs) {
  const double ar = a0.x, ai = a0.y, br = a1.x, bi = a1.y;
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const double2 b = bs[q];
    acc0[q].x = fma(-ai, b.y, fma(ar, b.x, acc0[q].x));
    acc0[q].y = fma(ai, b.x, fma(ar, b.y, acc0[q].y));
    acc1[q].x = fma(-bi, b.y, fma(br, b.x, acc1[q].x));
    acc1[q].y = fma(bi, b.x, fma(br, b.y, acc1[q].y));
  }
}

// acc0 += A0(:, 0:K) * B(0:K, :), acc1 += A1(:, 0:K) * B(0:K, :) for this
// lane's two rows: A's columns at stride lda (A0, A1 already offset to the
// rows), B's rows staged HMV_KC at a time; each staged value read from LDS
// feeds both rows.  Lanes take the columns j = j0 + js t; the first batch of
// A loads is issued before the staging, so the two round trips overlap.
template <class TS, class TB, int NR>
__device__ __forceinline__ void col_product2(double2 (&acc0)[NR], double2 (&acc1)[NR], double2* bs,
                                             const cplx<TS>* A0, const cplx<TS>* A1, int64_t lda, int K,
                                             const TB* B, int nrhs, int lane, int j0, int js) {
  constexpr int U = HMV_UNROLL / 2;
  const cplx<TS> z = mk<TS>(TS(0), TS(0));
  for (int k0 = 0; k0 < K; k0 += HMV_KC) {
    const int kc = min(HMV_KC, K - k0);
    const cplx<TS>* Ak0 = A0 + (int64_t)k0 * lda;
    const cplx<TS>* Ak1 = A1 + (int64_t)k0 * lda;
    cplx<TS> a0[U], a1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t o = (int64_t)min(j0 + u * js, kc - 1) * lda;
      a0[u] = Ak0[o];
      a1[u] = Ak1[o];
    }
    stage_rows<NR>(bs, B, k0, kc, nrhs, lane);
    const int step = U * js, nit = (kc + step - 1) / step;  // the same for every lane
    for (int i = 0; i < nit; ++i) {
      const int jb = j0 + i * step;
      if (i > 0)
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int64_t o = (int64_t)min(jb + u * js, kc - 1) * lda;
          a0[u] = Ak0[o];
          a1[u] = Ak1[o];
        }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = jb + u * js;
        const bool in = j < kc;
        lds_fma2<TS, NR>(acc0, acc1, in ? a0[u] : z, in ? a1[u] : z, bs + min(j, kc - 1) * NR);
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// First phase, t = R^T x over the column groups (the low-rank blocks of
// one column cluster; their factors R_b^T side by side, so that the x rows
// of the cluster are read once for all of them): the group's factors are
// cut into slices of <= 64, each stored (columns of the cluster) x nf
// row-major and contiguous (k_hmv_pack_cols), and t rows follow the groups'
// factors in order.  One wave per item = (slice, consecutive columns, about
// 16K entries): meta[4 i ..] = {nf, number of columns nj, first row of x, t
// row of the slice's first factor or -(1 + partial-sum offset)}, off[i] =
// the first panel entry.  lane = (l, jj): factors 2 l and 2 l + 1 (one
// 16-byte load for FP32 storage), columns jj, jj + 2, ...; the item is summed in chunks of CCH columns with the
// software pipeline of k_hmv_rows (the next chunk's panel values and x rows
// in flight while the current one is summed; the item as a buffer, so
// reads past its end return zero).  Factors >= nf read other columns of
// the slice and are not stored.  The partial sums over jj are combined by
// a shuffle.  Items of a slice with several items write partial sums (row
// stride nrhs) that k_hmv_cols_sum adds in item order.
// Workgroup b -> work unit, each of the eight XCDs (to which workgroups are
// dispatched round-robin) taking a contiguous range of units, so that
// neighbouring units, which read the same right-hand-side rows, share an
// L2.  The grid has 8 ceil(nunits / 8) workgroups; units >= nunits exit.
__device__ __forceinline__ int64_t xcd_unit(int64_t b, int64_t nunits) {
  const int64_t per = (nunits + 7) / 8;
  return (b % 8) * per + b / 8;
}

constexpr int CCH = 16;  // columns per chunk of the first phase
template <class TS, int NR>
__global__ void __launch_bounds__(64) k_hmv_cols(HmvBatch B, int nrhs) {
  constexpr int U = CCH / 2;          // columns per lane per chunk
  constexpr int PER = CCH * NR / 64;  // x values per lane per chunk
  __shared__ double2 bs[CCH * NR];
  const int64_t u0 = xcd_unit(blockIdx.x, B.ibeg[B.nm]);  // one wave per workgroup: items end independently
  if (u0 >= B.ibeg[B.nm]) return;
  const int mb = batch_of(B.ibeg, B.nm, u0);
  const int64_t it = u0 - B.ibeg[mb];
  const int64_t* off = B.ioff[mb];
  const int32_t* meta = B.imeta[mb];
  const cplx<TS>* rpanel = (const cplx<TS>*)B.rpanel[mb];
  const double2* xc = B.xt[mb];
  double2* t = B.xt[mb] + B.trow0[mb] * nrhs;
  double2* tpart = B.tpart[mb];
  const int lane = threadIdx.x & 63, la = lane & 31, jj = lane >> 5;
  const int nf = __builtin_amdgcn_readfirstlane(meta[4 * it]), nj = __builtin_amdgcn_readfirstlane(meta[4 * it + 1]);
  const int xr = meta[4 * it + 2], dst = meta[4 * it + 3];
  const int64_t o64 = off[it];
  const cplx<TS>* RT = rpanel + (((int64_t)__builtin_amdgcn_readfirstlane((int)(o64 >> 32)) << 32) |
                                 (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)o64));
  constexpr int ES = (int)sizeof(cplx<TS>);
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)RT, (short)0, nj * nf * ES, BUF_WORD3);
  int ou[U];
#pragma unroll
  for (int u = 0; u < U; ++u) ou[u] = ((jj + 2 * u) * nf + 2 * la) * ES;
  const double2* X = xc + (int64_t)xr * nrhs;
  double2 acc0[NR], acc1[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) acc0[q] = acc1[q] = mk<double>(0, 0);
  auto load_panel = [&](int c0, cplx<TS> (&a0)[U], cplx<TS> (&a1)[U]) {
    const int so = __builtin_amdgcn_readfirstlane(c0 * nf * ES);
#pragma unroll
    for (int u = 0; u < U; ++u) bload2<TS>(rs, ou[u], so, a0[u], a1[u]);
  };
  auto load_x = [&](int c0, double2 (&v)[PER]) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = lane + 64 * i, row = e / NR, q = e % NR;
      const double2 x = X[(int64_t)min(c0 + row, nj - 1) * nrhs + min(q, nrhs - 1)];
      v[i] = (c0 + row < nj && q < nrhs) ? x : mk<double>(0, 0);
    }
  };
  auto stage = [&](int c0, const cplx<TS> (&a0)[U], const cplx<TS> (&a1)[U], const double2 (&v)[PER],
                   cplx<TS> (&b0)[U], cplx<TS> (&b1)[U], double2 (&w)[PER]) {
#pragma unroll
    for (int i = 0; i < PER; ++i) bs[lane + 64 * i] = v[i];
    load_x(c0 + CCH, w);
    load_panel(c0 + CCH, b0, b1);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < U; ++u) lds_fma2<TS, NR>(acc0, acc1, a0[u], a1[u], bs + (jj + 2 * u) * NR);
    __builtin_amdgcn_wave_barrier();
  };
  {
    cplx<TS> a0[U], a1[U], b0[U], b1[U];
    double2 v[PER], w[PER];
    int c0 = 0;
    load_x(c0, v);
    load_panel(c0, a0, a1);
#pragma unroll 1
    for (;;) {
      stage(c0, a0, a1, v, b0, b1, w);
      c0 += CCH;
      if (c0 >= nj) break;
      stage(c0, b0, b1, w, a0, a1, v);
      c0 += CCH;
      if (c0 >= nj) break;
    }
  }
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    acc0[q].x += __shfl_xor(acc0[q].x, 32);
    acc0[q].y += __shfl_xor(acc0[q].y, 32);
    acc1[q].x += __shfl_xor(acc1[q].x, 32);
    acc1[q].y += __shfl_xor(acc1[q].y, 32);
  }
  double2* d = dst >= 0 ? t + (int64_t)dst * nrhs : tpart + (-(int64_t)dst - 1);
  if (jj == 0)
#pragma unroll
    for (int q = 0; q < NR; ++q)
      if (q < nrhs) {
        if (2 * la < nf) d[(int64_t)(2 * la) * nrhs + q] = acc0[q];
        if (2 * la + 1 < nf) d[(int64_t)(2 * la + 1) * nrhs + q] = acc1[q];
      }
}

// t rows of a slice with several items = the sum of their partial sums, in
// item order.  multi[4 i ..] = {t row, nf, first partial-sum offset, number
// of items}; the items' partial sums lie nf * MNP_HMV_MAXRHS apart.
__global__ void k_hmv_cols_sum(HmvBatch B, int nrhs) {
  const int64_t u0 = blockIdx.x;
  if (u0 >= B.mbeg[B.nm]) return;
  const int mb = batch_of(B.mbeg, B.nm, u0);
  const int64_t i = u0 - B.mbeg[mb];
  const int32_t* multi = B.multi[mb];
  const double2* tpart = B.tpart[mb];
  double2* t = B.xt[mb] + B.trow0[mb] * nrhs;
  const int tr = multi[4 * i], nf = multi[4 * i + 1], p0 = multi[4 * i + 2], ni = multi[4 * i + 3];
  const int len = nf * nrhs;
  for (int e = threadIdx.x; e < len; e += blockDim.x) {
    double2 v = mk<double>(0, 0);
    for (int k0 = 0; k0 < ni; k0 += 8) {  // eight loads in flight, summed in item order
      double2 p[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        p[u] = tpart[p0 + (int64_t)min(k0 + u, ni - 1) * nf * MNP_HMV_MAXRHS + e];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (k0 + u < ni) v = cadd<double>(v, p[u]);
    }
    t[(int64_t)tr * nrhs + e] = v;
  }
}

// rowblk[trow[b] + l] = b for the factors l < r_b of every low-rank block.
__global__ void k_hmv_rowblk(int64_t nlow, const int64_t* trow, const int32_t* rank, int32_t* rowblk) {
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= nlow) return;
  for (int l = 0; l < rank[b]; ++l) rowblk[trow[b] + l] = (int32_t)b;
}

// Column-group slices, built once per fill: slice i (sl[4 i ..] = {t row of
// its first factor, nf, first column of the cluster, number of columns},
// panel entries from soff[i]) holds entry (j, f) = R^T of the block owning t
// row (first + f), at column j, factor (first + f - trow[b]).  One
// workgroup per slice; the panel is written contiguously.
template <class TS>
__global__ void __launch_bounds__(256) k_hmv_pack_cols(const int32_t* sl, const int64_t* soff, const int32_t* rowblk,
                                                       const int64_t* trow, const int32_t* blk, const int32_t* rank,
                                                       const int64_t* loff, const cplx<TS>* pool, cplx<TS>* rpanel) {
  __shared__ int64_t src[64];
  __shared__ int stride[64];
  const int64_t i = blockIdx.x;
  const int t0 = sl[4 * i], nf = sl[4 * i + 1], w = sl[4 * i + 3];
  cplx<TS>* P = rpanel + soff[i];
  if (threadIdx.x < nf) {  // factor f: R^T column of its block, stride r
    const int f = threadIdx.x, b = rowblk[t0 + f], r = rank[b];
    const int m = blk[4 * b + 1] - blk[4 * b];
    src[f] = loff[b] + (int64_t)m * r + (t0 + f - trow[b]);
    stride[f] = r;
  }
  __syncthreads();
  for (int64_t e = threadIdx.x; e < (int64_t)nf * w; e += blockDim.x) {
    const int j = (int)(e / nf), f = (int)(e % nf);
    P[e] = pool[src[f] + (int64_t)j * stride[f]];
  }
}

// Row panels, built once per fill: tile T's panel is (rows of T) x K_T,
// column-major with leading dimension mt (contiguous), at poff[T]; its
// columns are, in descriptor order, the tile's rows of each full block
// (column j of the block) and the tile's L_b segment of each low-rank block
// (factor l).  cidx[tcol[T] + k] is the row of the right-hand-side buffer
// that column k multiplies: the column of x (rows 0 .. n) or n + the row of
// t.  Descriptors: A at aoff (rows at stride lda, K columns), boff >= 0 the
// first column of x, boff < 0 the first row of t as -boff - 1.  One
// workgroup per tile; source and panel reads and writes are contiguous per
// column.
template <class TS>
__global__ void __launch_bounds__(256) k_hmv_pack(const int32_t* tilebeg, const int32_t* tdbeg, const int64_t* taoff,
                                                  const int64_t* tboff, const int32_t* tlda, const int32_t* tK,
                                                  const int64_t* tcol, const int64_t* poff, const cplx<TS>* fvals,
                                                  const cplx<TS>* pool, int64_t n, cplx<TS>* panel,
                                                  int32_t* cidx) {
  const int tile = blockIdx.x;
  const int mt = tilebeg[tile + 1] - tilebeg[tile];
  cplx<TS>* P = panel + poff[tile];
  int32_t* C = cidx + tcol[tile];
  int64_t col = 0;
  const int s1 = tdbeg[tile + 1];
  for (int s = tdbeg[tile]; s < s1; ++s) {
    const int64_t aoff = taoff[s], boff = tboff[s];
    const int lda = tlda[s], K = tK[s];
    const cplx<TS>* src = (boff >= 0 ? fvals : pool) + aoff;
    for (int e = threadIdx.x; e < mt * K; e += blockDim.x) {
      const int i = e % mt, j = e / mt;
      P[(col + j) * mt + i] = src[(int64_t)j * lda + i];
    }
    const int64_t b0 = boff >= 0 ? boff : n + (-boff - 1);
    for (int j = threadIdx.x; j < K; j += blockDim.x) C[col + j] = (int32_t)(b0 + j);
    col += K;
  }
}

// Second phase: one workgroup per row tile (<= 32 rows of one leaf) streams
// its row panel against the rows of xt = [x; t] its columns name, and
// writes the tile's rows of y in face order (y[q n + perm[row]]).  The four
// waves take chunks of RKC columns round-robin.  Software pipeline per
// wave: while chunk c is summed, the panel values of chunk c + 1 and its
// right-hand-side rows (gathered by the indices loaded one chunk earlier)
// are in flight, with the indices of chunk c + 2.  lane = (rp, kk): rows
// 2 rp and 2 rp + 1 (one 16-byte load for FP32 storage), kk splitting the
// columns four ways, so one load instruction reads four 32-row column
// segments and each LDS value read serves two rows.
constexpr int RKC = 32;  // columns per chunk of the second phase

template <class TS, int NR>
__global__ void __launch_bounds__(256) k_hmv_rows(HmvBatch B, int nrhs) {
  constexpr int PER = RKC * NR / 64;  // right-hand-side values per lane per chunk
  constexpr int CPL = RKC / 4;        // columns per lane per chunk
  __shared__ double2 red[4][32][NR];
  __shared__ double2 bsh[4][RKC * NR];
  const int64_t u0 = xcd_unit(blockIdx.x, B.tbeg[B.nm]);
  if (u0 >= B.tbeg[B.nm]) return;
  const int mb = batch_of(B.tbeg, B.nm, u0);
  const int64_t tile = u0 - B.tbeg[mb];
  const int32_t* tilebeg = B.tilebeg[mb];
  const int64_t* tcol = B.tcol[mb];
  const int64_t* poff = B.poff[mb];
  const int32_t* cidx = B.cidx[mb];
  const cplx<TS>* panel = (const cplx<TS>*)B.panel[mb];
  const double2* xt = B.xt[mb];
  const int32_t* perm = B.perm[mb];
  const int64_t n = B.n;
  double2* y = B.y[mb];
  const int tb = tilebeg[tile], mt = tilebeg[tile + 1] - tb;
  const int64_t cb = tcol[tile];
  // wave-uniform by construction; read through readfirstlane so that the
  // buffer resource below lives in scalar registers
  const int K = __builtin_amdgcn_readfirstlane((int)(tcol[tile + 1] - cb));
  const int64_t po = poff[tile];
  const cplx<TS>* A = panel + (((int64_t)__builtin_amdgcn_readfirstlane((int)(po >> 32)) << 32) |
                               (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)po));
  const int32_t* idx = cidx + cb;
  const int lane = threadIdx.x & 63, rp = lane & 15, kk = lane >> 4;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int mtu = __builtin_amdgcn_readfirstlane(mt);
  double2* bs = bsh[wv];
  double2 acc0[NR], acc1[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) acc0[q] = acc1[q] = mk<double>(0, 0);
  // the tile's panel as a buffer: columns past K read zero; rows past the
  // tile read the next column (or zero) and are not stored
  constexpr int ES = (int)sizeof(cplx<TS>);
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)A, (short)0, K * mtu * ES, BUF_WORD3);
  int ou[CPL];
#pragma unroll
  for (int u = 0; u < CPL; ++u) ou[u] = ((kk + 4 * u) * mtu + 2 * rp) * ES;
  // the tile's column sources as a buffer too: indices past K read zero
  const __amdgpu_buffer_rsrc_t ri = __builtin_amdgcn_make_buffer_rsrc((void*)idx, (short)0, K * 4, BUF_WORD3);
  int oi[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) oi[i] = ((lane + 64 * i) / NR) * 4;
  // every load below is unconditional (out-of-range columns read zero and
  // gather row 0), so the loop body is straight-line code
  auto load_ids = [&](int k0, int (&id)[PER]) {
    const int so = __builtin_amdgcn_readfirstlane(k0 * 4);
#pragma unroll
    for (int i = 0; i < PER; ++i) id[i] = (int)__builtin_amdgcn_raw_buffer_load_b32(ri, oi[i], so, 0);
  };
  auto load_panel = [&](int k0, cplx<TS> (&a0)[CPL], cplx<TS> (&a1)[CPL]) {
    const int so = __builtin_amdgcn_readfirstlane(k0 * mtu * ES);
#pragma unroll
    for (int u = 0; u < CPL; ++u) bload2<TS>(rs, ou[u], so, a0[u], a1[u]);
  };
  auto load_rhs = [&](int k0, const int (&id)[PER], double2 (&v)[PER]) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = lane + 64 * i, row = e / NR, q = e % NR;
      const double2 x = xt[(int64_t)id[i] * nrhs + min(q, nrhs - 1)];
      v[i] = (k0 + row < K && q < nrhs) ? x : mk<double>(0, 0);
    }
  };
  const int step = 4 * RKC;
  // one pipeline stage: stage chunk k0's right-hand sides in LDS; put in
  // flight the indices of chunk k0 + 2 step, the right-hand sides of chunk
  // k0 + step (by the indices loaded one stage earlier) and its panel
  // values, in the order in which the next stage waits for them; sum chunk
  // k0.  The register
  // sets alternate between stages, so no copy waits on a load in flight.
  auto stage = [&](int k0, const cplx<TS> (&a0)[CPL], const cplx<TS> (&a1)[CPL], const double2 (&v)[PER],
                   cplx<TS> (&b0)[CPL], cplx<TS> (&b1)[CPL], double2 (&w)[PER], const int (&idc)[PER],
                   int (&idf)[PER]) {
#pragma unroll
    for (int i = 0; i < PER; ++i) bs[lane + 64 * i] = v[i];
    load_ids(k0 + 2 * step, idf);
    load_rhs(k0 + step, idc, w);
    load_panel(k0 + step, b0, b1);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < CPL; ++u) lds_fma2<TS, NR>(acc0, acc1, a0[u], a1[u], bs + (kk + 4 * u) * NR);
    __builtin_amdgcn_wave_barrier();
  };
  int k0 = wv * RKC;
  if (k0 < K) {
    int id0[PER], id1[PER];
    cplx<TS> a0[CPL], a1[CPL], b0[CPL], b1[CPL];
    double2 v[PER], w[PER];
    load_ids(k0, id0);
    load_ids(k0 + step, id1);
    load_rhs(k0, id0, v);
    load_panel(k0, a0, a1);
#pragma unroll 1
    for (;;) {
      stage(k0, a0, a1, v, b0, b1, w, id1, id0);
      k0 += step;
      if (k0 >= K) break;
      stage(k0, b0, b1, w, a0, a1, v, id0, id1);
      k0 += step;
      if (k0 >= K) break;
    }
  }
#pragma unroll
  for (int q = 0; q < NR; ++q)
    for (int sh = 16; sh < 64; sh <<= 1) {
      acc0[q].x += __shfl_xor(acc0[q].x, sh);
      acc0[q].y += __shfl_xor(acc0[q].y, sh);
      acc1[q].x += __shfl_xor(acc1[q].x, sh);
      acc1[q].y += __shfl_xor(acc1[q].y, sh);
    }
  if (kk == 0)
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      red[wv][2 * rp][q] = acc0[q];
      red[wv][2 * rp + 1][q] = acc1[q];
    }
  __syncthreads();
  for (int e = threadIdx.x; e < mt * nrhs; e += blockDim.x) {
    const int i = e / nrhs, q = e % nrhs;
    double2 v = red[0][i][q];
    for (int w2 = 1; w2 < 4; ++w2) v = cadd<double>(v, red[w2][i][q]);
    y[(int64_t)q * n + perm[tb + i]] = v;
  }
}

// ---------------------------------------------------------------- host
HMat::HMat(const GreenRet* g_, int prec_, const int64_t* perm_h, int64_t nfull_, const int64_t* full_h,
           int64_t nlow_, const int64_t* low_h, double htol_, int kmax_)
    : g(g_), n(g_->n), nfull(nfull_), nlow(nlow_), htol(htol_), kmax(kmax_), prec(prec_) {
  MNP_CHECK(prec == MNP_FP64 || prec == MNP_FP32, MNP_ERR_ARG, "hmat: precision must be MNP_FP64 or MNP_FP32");
  MNP_CHECK(kmax >= 1 && kmax <= 128, MNP_ERR_ARG, "hmat: kmax must be in 1..128");
  std::vector<int32_t> perm32(n), seen(n, 0);
  for (int64_t i = 0; i < n; ++i) {
    MNP_CHECK(perm_h[i] >= 0 && perm_h[i] < n && !seen[perm_h[i]], MNP_ERR_ARG, "hmat: perm is not a permutation");
    seen[perm_h[i]] = 1;
    perm32[i] = (int32_t)perm_h[i];
  }
  // blocks (column-major 4 x nb on input) -> int32 rows of 4; coverage check
  auto pack = [&](int64_t nb, const int64_t* src, std::vector<int32_t>& dst) {
    dst.resize(4 * nb);
    for (int64_t b = 0; b < nb; ++b)
      for (int c = 0; c < 4; ++c) {
        const int64_t v = src[4 * b + c];
        MNP_CHECK(v >= 0 && v <= n, MNP_ERR_ARG, "hmat: block range out of bounds");
        dst[4 * b + c] = (int32_t)v;
      }
    for (int64_t b = 0; b < nb; ++b)
      MNP_CHECK(dst[4 * b] < dst[4 * b + 1] && dst[4 * b + 2] < dst[4 * b + 3], MNP_ERR_ARG, "hmat: empty block");
  };
  pack(nfull, full_h, fblk_h);
  pack(nlow, low_h, lblk_h);
  double area = 0;
  for (int64_t b = 0; b < nfull; ++b) area += (double)(fblk_h[4 * b + 1] - fblk_h[4 * b]) * (fblk_h[4 * b + 3] - fblk_h[4 * b + 2]);
  for (int64_t b = 0; b < nlow; ++b) area += (double)(lblk_h[4 * b + 1] - lblk_h[4 * b]) * (lblk_h[4 * b + 3] - lblk_h[4 * b + 2]);
  MNP_CHECK(area == (double)n * (double)n, MNP_ERR_ARG, "hmat: blocks do not tile the matrix");
  // full-block value offsets
  std::vector<int64_t> voff(nfull + 1, 0);
  for (int64_t b = 0; b < nfull; ++b)
    voff[b + 1] = voff[b] + (int64_t)(fblk_h[4 * b + 1] - fblk_h[4 * b]) * (fblk_h[4 * b + 3] - fblk_h[4 * b + 2]);
  nfullvals = voff[nfull];
  // leaf row clusters = row ranges of full blocks (leaf x leaf); every row lies in one
  std::vector<std::pair<int32_t, int32_t>> leaves;
  for (int64_t b = 0; b < nfull; ++b) leaves.emplace_back(fblk_h[4 * b], fblk_h[4 * b + 1]);
  std::sort(leaves.begin(), leaves.end());
  leaves.erase(std::unique(leaves.begin(), leaves.end()), leaves.end());
  nleaf = (int64_t)leaves.size();
  std::vector<int64_t> lrow(2 * nleaf);
  std::vector<int32_t> rowleaf(n, -1);
  for (int64_t l = 0; l < nleaf; ++l) {
    lrow[2 * l] = leaves[l].first;
    lrow[2 * l + 1] = leaves[l].second;
    for (int32_t i = leaves[l].first; i < leaves[l].second; ++i) {
      MNP_CHECK(rowleaf[i] < 0, MNP_ERR_ARG, "hmat: overlapping leaf row clusters");
      rowleaf[i] = (int32_t)l;
    }
  }
  for (int64_t i = 0; i < n; ++i) MNP_CHECK(rowleaf[i] >= 0, MNP_ERR_ARG, "hmat: row not in a leaf cluster");
  // per-leaf entry lists (full block ids >= 0, low-rank -(id+1)); diagonal blocks
  std::vector<std::vector<int32_t>> lists(nleaf);
  diag_h.assign(nleaf, -1);
  for (int64_t b = 0; b < nfull; ++b) {
    const int32_t l = rowleaf[fblk_h[4 * b]];
    lists[l].push_back((int32_t)b);
    if (fblk_h[4 * b] == fblk_h[4 * b + 2] && fblk_h[4 * b + 1] == fblk_h[4 * b + 3]) diag_h[l] = (int32_t)b;
  }
  for (int64_t b = 0; b < nlow; ++b) {
    const int32_t la = rowleaf[lblk_h[4 * b]], lb = rowleaf[lblk_h[4 * b + 1] - 1];
    for (int32_t l = la; l <= lb; ++l) lists[l].push_back((int32_t)(-b - 1));
  }
  std::vector<int64_t> loffs(nleaf + 1, 0);
  std::vector<int32_t> lentry;
  for (int64_t l = 0; l < nleaf; ++l) {
    lentry.insert(lentry.end(), lists[l].begin(), lists[l].end());
    loffs[l + 1] = (int64_t)lentry.size();
  }
  lrow_h = lrow;
  // row tiles: each leaf cut into near-equal runs of at most 32 rows
  std::vector<int32_t> tilebeg{0}, tileleaf, rowtile(n);
  for (int64_t l = 0; l < nleaf; ++l) {
    const int64_t a = lrow[2 * l], m = lrow[2 * l + 1] - a, nt = (m + 31) / 32;
    MNP_CHECK(a == tilebeg.back(), MNP_ERR_INTERNAL, "hmat: leaf clusters are not contiguous");
    for (int64_t c = 0; c < nt; ++c) {
      const int64_t t0 = a + c * m / nt, t1 = a + (c + 1) * m / nt;
      for (int64_t i = t0; i < t1; ++i) rowtile[i] = (int32_t)tileleaf.size();
      tilebeg.push_back((int32_t)t1);
      tileleaf.push_back((int32_t)l);
    }
  }
  ntile = (int64_t)tileleaf.size();
  rowtile_h = rowtile;
  tilebeg_h = tilebeg;
  tileleaf_h = tileleaf;
  lentry_h = lentry;
  loffs_h = loffs;
  voff_h = voff;
  // device copies
  auto up = [&](auto& vec, auto*& dptr) {
    using E = typename std::remove_reference<decltype(vec)>::type::value_type;
    dptr = dev_alloc<E>((int64_t)vec.size());
    if (!vec.empty()) MNP_HIP(hipMemcpy(dptr, vec.data(), sizeof(E) * vec.size(), hipMemcpyHostToDevice));
  };
  up(perm32, perm);
  up(fblk_h, fblk);
  up(lblk_h, lblk);
  up(voff, d_voff);
  up(lrow, d_lrow);
  up(tilebeg, d_tilebeg);
  up(tileleaf, d_tileleaf);
  up(rowtile, d_rowtile);
  up(loffs, d_loffs);
  up(lentry, d_lentry);
  // cluster-ordered geometry for ACA
  std::vector<double> pos(3 * n), nv(3 * n), ar(n), cp(3 * n), cn(3 * n), ca(n);
  MNP_HIP(hipMemcpy(pos.data(), g->geo.pos, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  MNP_HIP(hipMemcpy(nv.data(), g->geo.nvec, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  MNP_HIP(hipMemcpy(ar.data(), g->geo.area, sizeof(double) * n, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) {
      cp[c * n + i] = pos[c * n + perm32[i]];
      cn[c * n + i] = nv[c * n + perm32[i]];
    }
    ca[i] = ar[perm32[i]];
  }
  up(cp, cpos);
  up(cn, cnvec);
  up(ca, carea);
  rank = dev_alloc<int32_t>(nlow);
  d_loff = dev_alloc<int64_t>(nlow + 1);
  d_toff = dev_alloc<int64_t>(nlow + 1);
}

HMat::~HMat() {
  void* ps[] = {d_ioff, d_imeta, d_multi, d_slice, d_soff, d_rowblk, rpanel, d_tdbeg, d_taoff, d_tboff, d_tlda, d_tK,
                perm, fblk, lblk, d_voff, d_lrow, d_loffs, d_lentry, d_tilebeg, d_tileleaf, d_rowtile,
                cpos, cnvec, carea, rank,
                d_loff, d_toff, fvals, pool, d_aca_items, d_aca_order, d_aca_soff, d_dsel, d_dlb, d_droff,
                d_citem, d_ccoff, d_tcol, d_poff, d_cidx, panel};
  for (void* p : ps)
    if (p) (void)hipFree(p);
}

void HMat::fill(double2 k, int key_) {
  if (prec == MNP_FP64)
    fill_impl<double>(k, key_, nullptr);
  else
    fill_impl<float>(k, key_, nullptr);
}

void HMat::fill_layer(double2 k, int key_, int comp, bool direct, const LayerPairs& P) {
  MNP_CHECK(comp >= 0 && comp < 5, MNP_ERR_ARG, "hmat: substrate component must be 0..4");
  const bool isF = key_ != MNP_KEY_G;
  LayerFill L{1, direct ? 1 : 0, comp, P.T, P.coff, P.rows, P.dmap,
              (isF ? P.pF : P.pG) + (int64_t)comp * P.np, (isF ? P.dF : P.dG) + (int64_t)comp * P.nd};
  if (prec == MNP_FP64)
    fill_impl<double>(k, key_, &L);
  else
    fill_impl<float>(k, key_, &L);
}

// ACA scratch shared by all H-matrices; never destroyed (no hipFree from a
// static destructor while the process exits).
static Workspace& aca_ws() {
  static Workspace* ws = new Workspace;
  return *ws;
}

template <class TS>
void HMat::fill_impl(double2 k, int key_, const LayerFill* lay) {
  key = key_;
  hipStream_t s = stream();
  const size_t es = sizeof(cplx<TS>);
  // full blocks
  if (!fvals) fvals = dev_alloc<cplx<TS>>(nfullvals);
  if (nfull) {
    const GreenRet& G = *g;
    TimerScope ts(MNP_TIMER_ASSEMBLE, (double)es * (double)nfullvals, (double)es * (double)nfullvals);
    const LayerFill off{};
    hipLaunchKernelGGL(k_hfull<TS>, dim3((unsigned)nfull), dim3(256), 0, s, nfull, fblk, d_voff, perm, n, G.geo,
                       G.pairs, G.off, G.nnz, G.order, G.gtab, G.ftab, k, key, lay ? *lay : off, (cplx<TS>*)fvals);
    MNP_HIP(hipGetLastError());
  }
  // ACA over the work plan, in batches whose FP64 scratch fits the budget
  const int isF = key == MNP_KEY_G ? 0 : 1;
  std::vector<int32_t> rank_h(nlow, 0);
  std::vector<int64_t> loff_h(nlow + 1, 0);
  if (nlow) {
    if (aca_icut.empty()) aca_plan();
    // scratch shared by every H-matrix fill (grown on demand, kept until
    // the process ends): a fill per operator and wavelength would otherwise
    // pay an allocation and a free of several GiB each
    double2* scratch = (double2*)aca_ws().get(0, sizeof(double2) * (size_t)aca_maxbatch);
    int64_t pos = 0;  // pool entries assigned so far (blocks in work order)
    for (size_t c = 0; c + 1 < aca_icut.size(); ++c) {
      const int64_t i0 = aca_icut[c], i1 = aca_icut[c + 1], o0 = aca_bcut[c], o1 = aca_bcut[c + 1];
      {
        TimerScope ts(MNP_TIMER_ACA, 0.0);
        const EntDirect ed{cpos, cnvec, carea, n, k, isF};
        if (lay)
          hipLaunchKernelGGL(k_aca<EntLayer>, dim3((unsigned)(i1 - i0)), dim3(256), 0, s, d_aca_items + 4 * i0, lblk,
                             d_aca_soff, EntLayer{ed, lay->direct, lay->comp, lay->T}, htol, kmax, scratch, rank);
        else
          hipLaunchKernelGGL(k_aca<EntDirect>, dim3((unsigned)(i1 - i0)), dim3(256), 0, s, d_aca_items + 4 * i0, lblk,
                             d_aca_soff, ed, htol, kmax, scratch, rank);
        MNP_HIP(hipGetLastError());
      }
      MNP_HIP(hipMemcpyAsync(rank_h.data(), rank, sizeof(int32_t) * nlow, hipMemcpyDeviceToHost, s));
      MNP_HIP(hipStreamSynchronize(s));
      const int64_t pos0 = pos;
      for (int64_t o = o0; o < o1; ++o) {
        const int32_t b = aca_order_h[o];
        const int64_t m = lblk_h[4 * b + 1] - lblk_h[4 * b], w = lblk_h[4 * b + 3] - lblk_h[4 * b + 2];
        loff_h[b] = pos;
        pos += (m + w) * rank_h[b];
      }
      if (pos > poolcap) {  // grow the pool, keeping earlier batches
        const int64_t cap = std::max<int64_t>(pos + pos / 2, 1 << 20);
        void* np = dev_alloc<cplx<TS>>(cap);
        if (pool && pos0) MNP_HIP(hipMemcpyAsync(np, pool, es * pos0, hipMemcpyDeviceToDevice, s));
        MNP_HIP(hipStreamSynchronize(s));
        if (pool) MNP_HIP(hipFree(pool));
        pool = np;
        poolcap = cap;
      }
      loff_h[nlow] = pos;
      MNP_HIP(hipMemcpy(d_loff, loff_h.data(), sizeof(int64_t) * (nlow + 1), hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_aca_store<TS>, dim3((unsigned)(o1 - o0)), dim3(256), 0, s, d_aca_order + o0, lblk,
                         d_aca_soff, scratch, rank, kmax, d_loff, d_rowtile, d_tilebeg, (cplx<TS>*)pool);
      MNP_HIP(hipGetLastError());  // a next batch reuses the scratch after this store, in stream order
    }
  }
  nlowvals = loff_h[nlow];
  maxrank = 0;
  for (int32_t r : rank_h) maxrank = std::max(maxrank, (int)r);
  rank_host = rank_h;
  loff_host = loff_h;
  build_plan();
  if (nslice) {
    hipLaunchKernelGGL(k_hmv_rowblk, dim3((unsigned)cdiv_up(nlow, 256)), dim3(256), 0, s, nlow, d_toff, rank,
                       d_rowblk);
    hipLaunchKernelGGL(k_hmv_pack_cols<TS>, dim3((unsigned)nslice), dim3(256), 0, s, d_slice, d_soff, d_rowblk,
                       d_toff, lblk, rank, d_loff, (const cplx<TS>*)pool, (cplx<TS>*)rpanel);
  }
  if (ntile)
    hipLaunchKernelGGL(k_hmv_pack<TS>, dim3((unsigned)ntile), dim3(256), 0, s, d_tilebeg, d_tdbeg, d_taoff, d_tboff,
                       d_tlda, d_tK, d_tcol, d_poff, (const cplx<TS>*)fvals, (const cplx<TS>*)pool, n,
                       (cplx<TS>*)panel, d_cidx);
  MNP_HIP(hipGetLastError());
  filled = true;
}

// ACA work plan: admissible blocks with m + w <= 320 run one per wave (four
// per workgroup), larger ones one per workgroup of 256 threads; work items
// in decreasing estimated time (the largest start first, so that the launch
// does not end on a few long blocks), grouped into batches whose FP64 scratch
// (m + w) min(kmax, m, w) per block fits a budget of a quarter of the free
// device memory (at least 8 GiB).
void HMat::aca_plan() {
  struct Item {
    double cost;
    int32_t b[4];
  };
  std::vector<Item> big, small;
  std::vector<std::pair<double, int32_t>> sm;
  std::vector<int64_t> sz(nlow);
  for (int64_t b = 0; b < nlow; ++b) {
    const int64_t m = lblk_h[4 * b + 1] - lblk_h[4 * b], w = lblk_h[4 * b + 3] - lblk_h[4 * b + 2];
    const int64_t kb = std::min<int64_t>(kmax, std::min(m, w));
    sz[b] = ((m + w) * kb + 31) / 32 * 32;
    const bool wave = m + w <= 320;
    const int G = wave ? 64 : 256;
    MNP_CHECK(m <= 128 * (int64_t)G, MNP_ERR_ARG, "hmat: admissible block with more than 32768 rows");
    // steps x (elements per thread x (kernel + residual loads) + reductions)
    const double steps = (double)std::min<int64_t>(kb, 24);
    const double cost = steps * ((double)((m + G - 1) / G + (w + G - 1) / G) * (1.0 + steps / 16.0) + 4.0);
    if (wave)
      sm.emplace_back(cost, (int32_t)b);
    else
      big.push_back({cost, {(int32_t)b, -2, -1, -1}});
  }
  std::stable_sort(sm.begin(), sm.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
  for (size_t i = 0; i < sm.size(); i += 4) {
    Item it{sm[i].first, {-1, -1, -1, -1}};
    for (size_t j = 0; j < 4 && i + j < sm.size(); ++j) it.b[j] = sm[i + j].second;
    small.push_back(it);
  }
  std::vector<Item> items(big);
  items.insert(items.end(), small.begin(), small.end());
  std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.cost > y.cost; });
  size_t fr = 0, tot = 0;
  MNP_HIP(hipMemGetInfo(&fr, &tot));
  const int64_t budget = std::max<int64_t>((int64_t)1 << 29, (int64_t)(fr / 4 / sizeof(double2)));
  aca_items_h.clear();
  aca_order_h.clear();
  aca_icut.assign(1, 0);
  aca_bcut.assign(1, 0);
  aca_soff_h.assign(nlow, 0);
  aca_maxbatch = 0;
  int64_t acc = 0;
  for (size_t i = 0; i < items.size(); ++i) {
    int64_t isz = 0;
    for (int j = 0; j < 4; ++j)
      if (items[i].b[j] >= 0) isz += sz[items[i].b[j]];
    if (acc + isz > budget && (int64_t)i > aca_icut.back()) {
      aca_icut.push_back((int64_t)i);
      aca_bcut.push_back((int64_t)aca_order_h.size());
      acc = 0;
    }
    for (int j = 0; j < 4; ++j) {
      const int32_t b = items[i].b[j];
      aca_items_h.push_back(b);
      if (b >= 0) {
        aca_order_h.push_back(b);
        aca_soff_h[b] = acc;
        acc += sz[b];
      }
    }
    aca_maxbatch = std::max(aca_maxbatch, acc);
  }
  aca_icut.push_back((int64_t)items.size());
  aca_bcut.push_back((int64_t)aca_order_h.size());
  for (void** p : {(void**)&d_aca_items, (void**)&d_aca_order, (void**)&d_aca_soff})
    if (*p) {
      MNP_HIP(hipFree(*p));
      *p = nullptr;
    }
  d_aca_items = dev_alloc<int32_t>((int64_t)aca_items_h.size());
  d_aca_order = dev_alloc<int32_t>((int64_t)aca_order_h.size());
  d_aca_soff = dev_alloc<int64_t>(nlow);
  MNP_HIP(hipMemcpy(d_aca_items, aca_items_h.data(), sizeof(int32_t) * aca_items_h.size(), hipMemcpyHostToDevice));
  MNP_HIP(hipMemcpy(d_aca_order, aca_order_h.data(), sizeof(int32_t) * aca_order_h.size(), hipMemcpyHostToDevice));
  MNP_HIP(hipMemcpy(d_aca_soff, aca_soff_h.data(), sizeof(int64_t) * nlow, hipMemcpyHostToDevice));
}

// Product plan.  The structure (first-phase items, the row tiles'
// descriptor lists, the full-block descriptors) is fixed by the block tree
// and built once; the low-rank descriptors (factor offset, t offset, rank)
// and the first-phase partial-sum offsets follow the ranks of each fill.
// Low-rank descriptor q of the list: entry dsel[q] of the descriptor arrays,
// block dlb[q], first row of its tile relative to the block droff[q].
__global__ void k_plan_low(int64_t nq, const int32_t* dsel, const int32_t* dlb, const int32_t* droff,
                           const int32_t* rank, const int64_t* loff, const int64_t* toff, int64_t* taoff,
                           int64_t* tboff, int32_t* tK) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const int64_t d = dsel[q];
  const int32_t lb = dlb[q], r = rank[lb];
  taoff[d] = loff[lb] + (int64_t)droff[q] * r;
  tboff[d] = -toff[lb] - 1;
  tK[d] = r;
}

void HMat::build_plan_static() {
  // column groups: the low-rank blocks of each column cluster, in block order
  {
    std::vector<int32_t> order(nlow);
    for (int64_t b = 0; b < nlow; ++b) order[b] = (int32_t)b;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
      return std::make_pair(lblk_h[4 * a + 2], lblk_h[4 * a + 3]) < std::make_pair(lblk_h[4 * b + 2], lblk_h[4 * b + 3]);
    });
    grp_beg_h.assign(1, 0);
    grp_blk_h = order;
    for (int64_t i = 1; i <= nlow; ++i)
      if (i == nlow || lblk_h[4 * order[i] + 2] != lblk_h[4 * order[i - 1] + 2] ||
          lblk_h[4 * order[i] + 3] != lblk_h[4 * order[i - 1] + 3])
        grp_beg_h.push_back((int32_t)i);
    if (nlow == 0) grp_beg_h.assign(1, 0);
  }
  std::vector<int32_t> tdbeg{0}, tlda, tK, dsel, dlb, droff;
  std::vector<int64_t> taoff, tboff;
  tkfull_h.assign(ntile, 0);
  dtile_h.clear();
  for (int64_t t = 0; t < ntile; ++t) {
    const int32_t tb = tilebeg_h[t], mt = tilebeg_h[t + 1] - tb, leaf = tileleaf_h[t];
    for (int64_t s = loffs_h[leaf]; s < loffs_h[leaf + 1]; ++s) {
      const int32_t ent = lentry_h[s];
      if (ent >= 0) {
        const int32_t* bk = &fblk_h[4 * ent];
        taoff.push_back(voff_h[ent] + (tb - bk[0]));
        tboff.push_back(bk[2]);
        tlda.push_back(bk[1] - bk[0]);
        tK.push_back(bk[3] - bk[2]);
        tkfull_h[t] += bk[3] - bk[2];
      } else {
        const int32_t lb = -ent - 1;
        dtile_h.push_back((int32_t)t);
        dsel.push_back((int32_t)taoff.size());
        dlb.push_back(lb);
        droff.push_back(tb - lblk_h[4 * lb]);
        taoff.push_back(0);  // set per fill by k_plan_low
        tboff.push_back(0);
        tlda.push_back(mt);
        tK.push_back(0);
      }
    }
    tdbeg.push_back((int32_t)taoff.size());
  }
  nlowdesc = (int64_t)dsel.size();
  auto up = [&](auto& vec, auto*& dptr) {
    using E = typename std::remove_reference<decltype(vec)>::type::value_type;
    dptr = dev_alloc<E>((int64_t)vec.size());
    if (!vec.empty()) MNP_HIP(hipMemcpy(dptr, vec.data(), sizeof(E) * vec.size(), hipMemcpyHostToDevice));
  };
  up(tdbeg, d_tdbeg);
  up(taoff, d_taoff);
  up(tboff, d_tboff);
  up(tlda, d_tlda);
  up(tK, d_tK);
  up(dsel, d_dsel);
  up(dlb, d_dlb);
  lowdesc_block_h = dlb;
  up(droff, d_droff);
  d_tcol = dev_alloc<int64_t>(ntile + 1);
  d_poff = dev_alloc<int64_t>(ntile + 1);
}

static inline int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// Rank-dependent part of the plan, on the stream after the fill's uploads
// of loff and toff: the low-rank descriptors, the partial-sum offsets of the
// first phase, the row panels' columns and offsets (panels and column
// sources grown on demand).
void HMat::build_plan() {
  if (!d_tdbeg) build_plan_static();
  hipStream_t s = stream();
  // t rows in column-group order; slices of <= 64 factors per group; items
  // of <= jch columns per slice, with partial sums where a slice has several
  toff_h.assign(nlow + 1, 0);
  slice_h.clear();
  soff_h.assign(1, 0);
  ioff_h.clear();
  imeta_h.clear();
  multi_h.clear();
  {
    int64_t tr = 0, tp = 0;
    const int64_t ngrp = (int64_t)grp_beg_h.size() - 1;
    for (int64_t g = 0; g < ngrp; ++g) {
      const int64_t t0 = tr;
      for (int32_t i = grp_beg_h[g]; i < grp_beg_h[g + 1]; ++i) {
        const int32_t b = grp_blk_h[i];
        toff_h[b] = tr;
        tr += rank_host[b];
      }
      if (tr == t0) continue;
      const int32_t b0 = grp_blk_h[grp_beg_h[g]];
      const int32_t c0 = lblk_h[4 * b0 + 2], w = lblk_h[4 * b0 + 3] - c0;
      for (int64_t f0 = t0; f0 < tr; f0 += 64) {
        const int32_t nf = (int32_t)std::min<int64_t>(64, tr - f0);
        const int64_t base = soff_h.back();
        slice_h.insert(slice_h.end(), {(int32_t)f0, nf, c0, w});
        soff_h.push_back(align_up(base + (int64_t)nf * w, 32));  // slices start on 256-byte boundaries
        const int32_t jch = std::max(64, std::min(512, (16384 / nf) / 64 * 64));
        const int32_t nj = (w + jch - 1) / jch;
        if (nj > 1) multi_h.insert(multi_h.end(), {(int32_t)f0, nf, (int32_t)tp, nj});
        for (int32_t c = 0; c < nj; ++c) {
          const int32_t j0 = (int32_t)((int64_t)c * w / nj), j1 = (int32_t)((int64_t)(c + 1) * w / nj);
          ioff_h.push_back(base + (int64_t)j0 * nf);
          imeta_h.insert(imeta_h.end(), {nf, j1 - j0, c0 + j0, nj > 1 ? (int32_t)(-1 - tp) : (int32_t)f0});
          if (nj > 1) tp += (int64_t)nf * MNP_HMV_MAXRHS;
        }
      }
    }
    toff_h[nlow] = tr;
    trows = tr;
    tpartlen = tp;
    MNP_CHECK(tp < ((int64_t)1 << 31), MNP_ERR_ARG, "hmat: too many partial sums");
  }
  nslice = (int64_t)slice_h.size() / 4;
  nitems = (int64_t)ioff_h.size();
  nmulti = (int64_t)multi_h.size() / 4;
  auto grow = [&](auto*& ptr, int64_t& cap, int64_t need) {
    using E = typename std::remove_reference<decltype(*ptr)>::type;
    if (need <= cap) return;
    if (ptr) MNP_HIP(hipFree(ptr));
    cap = need + need / 8;
    ptr = dev_alloc<E>(cap);
  };
  grow(d_slice, slicecap, 4 * std::max<int64_t>(nslice, 1));
  grow(d_soff, soffcap, nslice + 1);
  grow(d_ioff, ioffcap, std::max<int64_t>(nitems, 1));
  grow(d_imeta, imetacap, 4 * std::max<int64_t>(nitems, 1));
  grow(d_multi, multicap, 4 * std::max<int64_t>(nmulti, 1));
  grow(d_rowblk, rowblkcap, std::max<int64_t>(trows, 1));
  if (soff_h.back() > rpanelcap) {
    if (rpanel) MNP_HIP(hipFree(rpanel));
    rpanelcap = soff_h.back() + soff_h.back() / 8;
    rpanel = prec == MNP_FP64 ? (void*)dev_alloc<double2>(rpanelcap) : (void*)dev_alloc<float2>(rpanelcap);
  }
  MNP_HIP(hipMemcpyAsync(d_toff, toff_h.data(), sizeof(int64_t) * (nlow + 1), hipMemcpyHostToDevice, s));
  if (nslice) {
    MNP_HIP(hipMemcpyAsync(d_slice, slice_h.data(), sizeof(int32_t) * slice_h.size(), hipMemcpyHostToDevice, s));
    MNP_HIP(hipMemcpyAsync(d_soff, soff_h.data(), sizeof(int64_t) * soff_h.size(), hipMemcpyHostToDevice, s));
    MNP_HIP(hipMemcpyAsync(d_ioff, ioff_h.data(), sizeof(int64_t) * ioff_h.size(), hipMemcpyHostToDevice, s));
    MNP_HIP(hipMemcpyAsync(d_imeta, imeta_h.data(), sizeof(int32_t) * imeta_h.size(), hipMemcpyHostToDevice, s));
  }
  if (nmulti)
    MNP_HIP(hipMemcpyAsync(d_multi, multi_h.data(), sizeof(int32_t) * multi_h.size(), hipMemcpyHostToDevice, s));
  tcol_h.assign(ntile + 1, 0);
  poff_h.assign(ntile + 1, 0);
  {
    std::vector<int64_t> kt(tkfull_h);
    int64_t q = 0;
    for (int64_t t = 0; t < ntile; ++t) {
      // low-rank descriptors of tile t are consecutive in dtile_h
      for (; q < (int64_t)dtile_h.size() && dtile_h[q] == t; ++q) kt[t] += rank_host[lowdesc_block_h[q]];
      const int64_t mt = tilebeg_h[t + 1] - tilebeg_h[t];
      tcol_h[t + 1] = tcol_h[t] + kt[t];
      poff_h[t + 1] = align_up(poff_h[t] + mt * kt[t], 32);  // panels start on 256-byte boundaries
    }
  }
  if (poff_h[ntile] > panelcap) {
    if (panel) MNP_HIP(hipFree(panel));
    panelcap = poff_h[ntile] + poff_h[ntile] / 8;
    panel = prec == MNP_FP64 ? (void*)dev_alloc<double2>(panelcap) : (void*)dev_alloc<float2>(panelcap);
  }
  if (tcol_h[ntile] > cidxcap) {
    if (d_cidx) MNP_HIP(hipFree(d_cidx));
    cidxcap = tcol_h[ntile] + tcol_h[ntile] / 8;
    d_cidx = dev_alloc<int32_t>(cidxcap);
  }
  MNP_CHECK(n + trows < ((int64_t)1 << 31), MNP_ERR_ARG, "hmat: too many rows for the column sources");
  if (nlowdesc)
    hipLaunchKernelGGL(k_plan_low, dim3((unsigned)cdiv_up(nlowdesc, 256)), dim3(256), 0, s, nlowdesc, d_dsel, d_dlb,
                       d_droff, rank, d_loff, d_toff, d_taoff, d_tboff, d_tK);
  MNP_HIP(hipGetLastError());
  MNP_HIP(hipMemcpyAsync(d_tcol, tcol_h.data(), sizeof(int64_t) * (ntile + 1), hipMemcpyHostToDevice, s));
  MNP_HIP(hipMemcpyAsync(d_poff, poff_h.data(), sizeof(int64_t) * (ntile + 1), hipMemcpyHostToDevice, s));
  MNP_HIP(hipStreamSynchronize(s));  // the host vectors stay valid until the copies are done
}

void HMat::matvec(int64_t nrhs, const double2* x, double2* y, Workspace& ws) const {
  const HMat* self = this;
  Workspace* w = &ws;
  matvec_batch(1, &self, nrhs, &x, &y, &w);
}

void HMat::matvec_batch(int nm, const HMat* const* A, int64_t nrhs, const double2* const* x, double2* const* y,
                        Workspace* const* ws) {
  MNP_CHECK(nm >= 1 && nm <= HMV_MAXB, MNP_ERR_ARG, "hmat: 1 to 4 products per batch");
  for (int m = 0; m < nm; ++m) {
    MNP_CHECK(A[m]->filled, MNP_ERR_STATE, "hmat: matvec before fill");
    MNP_CHECK(A[m]->n == A[0]->n && A[m]->prec == A[0]->prec, MNP_ERR_ARG,
              "hmat: batched products need one size and one storage precision");
  }
  MNP_CHECK(nrhs >= 1, MNP_ERR_ARG, "hmat: nrhs < 1");
  if (A[0]->prec == MNP_FP64)
    matvec_batch_impl<double>(nm, A, nrhs, x, y, ws);
  else
    matvec_batch_impl<float>(nm, A, nrhs, x, y, ws);
}

template <class TS>
void HMat::matvec_batch_impl(int nm, const HMat* const* A, int64_t nrhs, const double2* const* x,
                             double2* const* y, Workspace* const* ws) {
  hipStream_t s = stream();
  const int64_t n = A[0]->n;
  double vals = 0;
  for (int m = 0; m < nm; ++m) vals += (double)(A[m]->nfullvals + A[m]->nlowvals);
  const double mvbytes = (double)sizeof(cplx<TS>) * vals * (double)((nrhs + MNP_HMV_MAXRHS - 1) / MNP_HMV_MAXRHS);
  TimerScope ts(MNP_TIMER_HMATVEC, mvbytes, mvbytes);
  for (int64_t q0 = 0; q0 < nrhs; q0 += MNP_HMV_MAXRHS) {
    const int nq = (int)std::min<int64_t>(MNP_HMV_MAXRHS, nrhs - q0);
    HmvBatch B{};
    B.nm = nm;
    B.n = n;
    for (int m = 0; m < nm; ++m) {
      const HMat& H = *A[m];
      // xt = [x in cluster order; t], row stride nq
      B.xt[m] = (double2*)ws[m]->get(0, sizeof(double2) * (n + std::max<int64_t>(H.trows, 1)) * nq);
      B.tpart[m] = (double2*)ws[m]->get(3, sizeof(double2) * std::max<int64_t>(H.tpartlen, 1));
      B.trow0[m] = n;
      B.perm[m] = H.perm;
      B.x[m] = x[m] + q0 * n;
      B.y[m] = y[m] + q0 * n;
      B.ioff[m] = H.d_ioff;
      B.imeta[m] = H.d_imeta;
      B.multi[m] = H.d_multi;
      B.rpanel[m] = H.rpanel;
      B.tilebeg[m] = H.d_tilebeg;
      B.tcol[m] = H.d_tcol;
      B.poff[m] = H.d_poff;
      B.cidx[m] = H.d_cidx;
      B.panel[m] = H.panel;
      B.ibeg[m + 1] = B.ibeg[m] + H.nitems;
      B.mbeg[m + 1] = B.mbeg[m] + H.nmulti;
      B.tbeg[m + 1] = B.tbeg[m] + H.ntile;
    }
    hipLaunchKernelGGL(k_to_cluster, dim3((unsigned)cdiv_up(n, 256), (unsigned)nq, (unsigned)nm), dim3(256), 0, s, B,
                       nq);
    if (const int64_t ni = B.ibeg[nm]) {
      const dim3 g((unsigned)(8 * cdiv_up(ni, 8)));
      if (nq <= 4)
        hipLaunchKernelGGL((k_hmv_cols<TS, 4>), g, dim3(64), 0, s, B, nq);
      else
        hipLaunchKernelGGL((k_hmv_cols<TS, MNP_HMV_MAXRHS>), g, dim3(64), 0, s, B, nq);
    }
    if (const int64_t nmu = B.mbeg[nm]) hipLaunchKernelGGL(k_hmv_cols_sum, dim3((unsigned)nmu), dim3(128), 0, s, B, nq);
    const dim3 gt((unsigned)(8 * cdiv_up(B.tbeg[nm], 8)));
    if (nq <= 4)
      hipLaunchKernelGGL((k_hmv_rows<TS, 4>), gt, dim3(256), 0, s, B, nq);
    else
      hipLaunchKernelGGL((k_hmv_rows<TS, MNP_HMV_MAXRHS>), gt, dim3(256), 0, s, B, nq);
    MNP_HIP(hipGetLastError());
  }
}

// ------------------------------------------------------ coarse operator
// Full blocks: one workgroup per block, the sum of its entries over each
// (row tile, column tile) pair (with leaves of at most 32 rows, one pair).
template <class TS>
__global__ void __launch_bounds__(256) k_coarse_full(const int32_t* fblk, const int64_t* voff, const int32_t* rowtile,
                                                     const int32_t* tilebeg, const cplx<TS>* fvals, int64_t nc,
                                                     double2* C) {
  const int64_t b = blockIdx.x;
  const int r0 = fblk[4 * b], r1 = fblk[4 * b + 1], c0 = fblk[4 * b + 2], c1 = fblk[4 * b + 3];
  const int m = r1 - r0;
  const cplx<TS>* V = fvals + voff[b];
  __shared__ double2 red[256];
  for (int s = rowtile[r0]; s <= rowtile[r1 - 1]; ++s)
    for (int t = rowtile[c0]; t <= rowtile[c1 - 1]; ++t) {
      const int i0 = tilebeg[s] - r0, i1 = tilebeg[s + 1] - r0, j0 = tilebeg[t] - c0, j1 = tilebeg[t + 1] - c0;
      const int mi = i1 - i0, cnt = mi * (j1 - j0);
      double2 acc = mk<double>(0, 0);
      for (int e = threadIdx.x; e < cnt; e += 256) {
        const cplx<TS> v = V[(int64_t)(j0 + e / mi) * m + i0 + e % mi];
        acc.x += v.x;
        acc.y += v.y;
      }
      red[threadIdx.x] = acc;
      __syncthreads();
      for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] = cadd<double>(red[threadIdx.x], red[threadIdx.x + st]);
        __syncthreads();
      }
      if (threadIdx.x == 0) C[(int64_t)t * nc + s] = red[0];
      __syncthreads();
    }
}

// Low-rank blocks L R^T in two flattened passes, so that large blocks spread
// over many workgroups: (1) one wave per (block, tile) item sums the tile's
// rows of L (or of R) into Lagg_s (or Ragg_t), r entries, lane l < r taking
// factor l; (2) one thread per coarse entry (s, t) of every block forms
// C(s, t) = Lagg_s . Ragg_t.  item[2 i ..] = {block, tile index within the
// block's rows (>= 0) or -1 - index within its columns}; entries of block b are [coff[b], coff[b + 1]) in column-major order of
// its ns x nt coarse block, with its aggregates at aggoff[b] (ns r of L
// then nt r of R).
template <class TS>
__global__ void __launch_bounds__(256) k_coarse_agg(int64_t nitems, const int32_t* item, const int64_t* aggoff,
                                                    const int32_t* lblk, const int32_t* rank, const int64_t* loff,
                                                    const int32_t* rowtile, const int32_t* tilebeg,
                                                    const cplx<TS>* pool, double2* agg) {
  const int64_t it = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (it >= nitems) return;
  const int lane = threadIdx.x & 63;
  const int b = item[2 * it], sel = item[2 * it + 1];
  const int r0 = lblk[4 * b], m = lblk[4 * b + 1] - r0, c0 = lblk[4 * b + 2];
  const int ns = rowtile[lblk[4 * b + 1] - 1] - rowtile[r0] + 1;
  const int r = rank[b];
  const cplx<TS>* L = pool + loff[b];
  double2* out = agg + aggoff[b] + (int64_t)(sel >= 0 ? sel : ns + (-1 - sel)) * r;
  if (sel >= 0) {
    const int s = rowtile[r0] + sel;
    const int a = tilebeg[s] - r0, mt = tilebeg[s + 1] - tilebeg[s];
    const cplx<TS>* seg = L + (int64_t)a * r;  // the tile's segment, mt x r column-major
    for (int l = lane; l < r; l += 64) {
      double2 acc = mk<double>(0, 0);
      for (int i = 0; i < mt; ++i) {
        const cplx<TS> v = seg[(int64_t)l * mt + i];
        acc.x += v.x;
        acc.y += v.y;
      }
      out[l] = acc;
    }
  } else {
    const int t = rowtile[c0] + (-1 - sel);
    const cplx<TS>* RT = L + (int64_t)m * r;
    for (int l = lane; l < r; l += 64) {
      double2 acc = mk<double>(0, 0);
      for (int j = tilebeg[t]; j < tilebeg[t + 1]; ++j) {
        const cplx<TS> v = RT[(int64_t)(j - c0) * r + l];
        acc.x += v.x;
        acc.y += v.y;
      }
      out[l] = acc;
    }
  }
}

__global__ void __launch_bounds__(256) k_coarse_dot(int64_t nlow, int64_t ntot, const int64_t* coff,
                                                    const int64_t* aggoff, const int32_t* lblk, const int32_t* rank,
                                                    const int32_t* rowtile, const double2* agg, int64_t nc,
                                                    double2* C) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= ntot) return;
  int64_t lo = 0, hi = nlow;  // block b with coff[b] <= e < coff[b + 1]
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if (coff[mid] <= e)
      lo = mid;
    else
      hi = mid;
  }
  const int64_t b = lo;
  const int r0 = lblk[4 * b], r1 = lblk[4 * b + 1], c0 = lblk[4 * b + 2];
  const int s0 = rowtile[r0], ns = rowtile[r1 - 1] - s0 + 1, t0 = rowtile[c0];
  const int64_t q = e - coff[b];
  const int si = (int)(q % ns), ti = (int)(q / ns);
  const int r = rank[b];
  const double2* La = agg + aggoff[b] + (int64_t)si * r;
  const double2* Ra = agg + aggoff[b] + (int64_t)ns * r + (int64_t)ti * r;
  double2 acc = mk<double>(0, 0);
  for (int l = 0; l < r; ++l) acc = cadd<double>(acc, cmul<double>(La[l], Ra[l]));
  C[(int64_t)(t0 + ti) * nc + s0 + si] = acc;
}

void HMat::coarsen(double2* C, Workspace& ws) const {
  if (prec == MNP_FP64)
    coarsen_impl<double>(C, ws);
  else
    coarsen_impl<float>(C, ws);
}

template <class TS>
void HMat::coarsen_impl(double2* C, Workspace& ws) const {
  MNP_CHECK(filled, MNP_ERR_STATE, "hmat: coarsen before fill");
  hipStream_t s = stream();
  if (nfull)
    hipLaunchKernelGGL(k_coarse_full<TS>, dim3((unsigned)nfull), dim3(256), 0, s, fblk, d_voff, d_rowtile, d_tilebeg,
                       (const cplx<TS>*)fvals, ntile, C);
  if (nlow) {
    if (!d_citem) {  // items and entry offsets depend on the block structure only
      std::vector<int64_t> coff(nlow + 1, 0);
      std::vector<int32_t> item;
      for (int64_t b = 0; b < nlow; ++b) {
        const int64_t ns = rowtile_h[lblk_h[4 * b + 1] - 1] - rowtile_h[lblk_h[4 * b]] + 1;
        const int64_t nt = rowtile_h[lblk_h[4 * b + 3] - 1] - rowtile_h[lblk_h[4 * b + 2]] + 1;
        coff[b + 1] = coff[b] + ns * nt;
        for (int64_t i = 0; i < ns + nt; ++i) {
          item.push_back((int32_t)b);
          item.push_back(i < ns ? (int32_t)i : (int32_t)(-1 - (i - ns)));
        }
      }
      ncitem = (int64_t)item.size() / 2;
      ncentry = coff[nlow];
      d_citem = dev_alloc<int32_t>(2 * ncitem);
      d_ccoff = dev_alloc<int64_t>(nlow + 1);
      MNP_HIP(hipMemcpy(d_citem, item.data(), sizeof(int32_t) * 2 * ncitem, hipMemcpyHostToDevice));
      MNP_HIP(hipMemcpy(d_ccoff, coff.data(), sizeof(int64_t) * (nlow + 1), hipMemcpyHostToDevice));
    }
    // aggregate offsets from this fill's ranks
    std::vector<int64_t> aggoff(nlow + 1, 0);
    for (int64_t b = 0; b < nlow; ++b) {
      const int64_t ns = rowtile_h[lblk_h[4 * b + 1] - 1] - rowtile_h[lblk_h[4 * b]] + 1;
      const int64_t nt = rowtile_h[lblk_h[4 * b + 3] - 1] - rowtile_h[lblk_h[4 * b + 2]] + 1;
      aggoff[b + 1] = aggoff[b] + (ns + nt) * rank_host[b];
    }
    int64_t* d_aggoff = (int64_t*)ws.get(0, sizeof(int64_t) * (nlow + 1));
    double2* agg = (double2*)ws.get(1, sizeof(double2) * (size_t)std::max<int64_t>(aggoff[nlow], 1));
    MNP_HIP(hipMemcpyAsync(d_aggoff, aggoff.data(), sizeof(int64_t) * (nlow + 1), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_coarse_agg<TS>, dim3((unsigned)cdiv_up(ncitem, 4)), dim3(256), 0, s, ncitem, d_citem,
                       d_aggoff, lblk, rank, d_loff, d_rowtile, d_tilebeg, (const cplx<TS>*)pool, agg);
    hipLaunchKernelGGL(k_coarse_dot, dim3((unsigned)cdiv_up(ncentry, 256)), dim3(256), 0, s, nlow, ncentry, d_ccoff,
                       d_aggoff, lblk, rank, d_rowtile, agg, ntile, C);
    MNP_HIP(hipStreamSynchronize(s));  // aggoff is host memory of this call
  }
  MNP_HIP(hipGetLastError());
}

}  // namespace mnp

// ============================================================== C ABI
using namespace mnp;

struct mnp_hmat_s {
  HMat* h;
  Workspace ws;
};

extern "C" {

int mnp_hmat_create(mnp_greenret_t g, int precision, const int64_t* perm, int64_t nfull, const int64_t* full,
                    int64_t nlow, const int64_t* low, double htol, int kmax, mnp_hmat_t* out) {
  return guarded([&] {
    MNP_CHECK(!g || !g->g->compound(), MNP_ERR_ARG, "hmat_create: compound particles take the dense solver (mnp_bemret_*)");
    MNP_CHECK(g && perm && out && (full || !nfull) && (low || !nlow), MNP_ERR_ARG, "hmat_create: NULL argument");
    MNP_CHECK(htol > 0, MNP_ERR_ARG, "hmat_create: htol must be positive");
    auto* h = new mnp_hmat_s{nullptr, {}};
    try {
      h->h = new HMat(g->g, precision, perm, nfull, full, nlow, low, htol, kmax);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int mnp_hmat_destroy(mnp_hmat_t h) {
  return guarded([&] {
    if (!h) return;
    delete h->h;
    delete h;
  });
}

int mnp_hmat_fill(mnp_hmat_t h, mnp_complex k, int key) {
  return guarded([&] {
    MNP_CHECK(h, MNP_ERR_ARG, "hmat_fill: NULL handle");
    MNP_CHECK(key >= MNP_KEY_G && key <= MNP_KEY_H2, MNP_ERR_ARG, "hmat_fill: bad key");
    h->h->fill(to_d2(k), key);
  });
}

int mnp_hmat_fill_layer(mnp_hmat_t h, mnp_complex k, int key, int comp, int direct, struct mnp_greenretlayer_s* gl,
                        double enei, mnp_complex eps_up, mnp_complex eps_down) {
  return guarded([&] {
    MNP_CHECK(h && gl, MNP_ERR_ARG, "hmat_fill_layer: NULL handle");
    MNP_CHECK(key == MNP_KEY_G || key == MNP_KEY_H2, MNP_ERR_ARG, "hmat_fill_layer: key must be MNP_KEY_G or MNP_KEY_H2");
    GreenRetLayer* L = greenretlayer_of(gl);
    const LayerPairs P = greenretlayer_pairs(L, enei, to_d2(eps_up), to_d2(eps_down));
    h->h->fill_layer(to_d2(k), key, comp, direct != 0, P);
    greenretlayer_check(L);
  });
}

int mnp_hmat_matvec(mnp_hmat_t h, int64_t nrhs, const void* x, void* y) {
  return guarded([&] {
    MNP_CHECK(h && x && y, MNP_ERR_ARG, "hmat_matvec: NULL argument");
    h->h->matvec(nrhs, (const double2*)x, (double2*)y, h->ws);
  });
}

int mnp_hmat_stats(mnp_hmat_t h, int64_t* full_entries, int64_t* lowrank_entries, int* max_rank,
                   double* compression) {
  return guarded([&] {
    MNP_CHECK(h, MNP_ERR_ARG, "hmat_stats: NULL handle");
    const HMat& H = *h->h;
    if (full_entries) *full_entries = H.nfullvals;
    if (lowrank_entries) *lowrank_entries = H.nlowvals;
    if (max_rank) *max_rank = H.maxrank;
    if (compression) *compression = (double)(H.nfullvals + H.nlowvals) / ((double)H.n * (double)H.n);
  });
}

int mnp_hmat_coarsen(mnp_hmat_t h, int64_t* ntile, int64_t* tilebeg, void* C) {
  return guarded([&] {
    MNP_CHECK(h && ntile, MNP_ERR_ARG, "hmat_coarsen: NULL argument");
    const HMat& H = *h->h;
    MNP_CHECK(H.filled, MNP_ERR_STATE, "hmat_coarsen: not filled");
    *ntile = H.ntile;
    if (tilebeg)
      for (int64_t t = 0; t <= H.ntile; ++t) tilebeg[t] = H.tilebeg_h[t];
    if (C) {
      H.coarsen((double2*)C, h->ws);
      MNP_HIP(hipStreamSynchronize(stream()));
    }
  });
}

int mnp_hmat_ranks(mnp_hmat_t h, int32_t* ranks) {
  return guarded([&] {
    MNP_CHECK(h && ranks, MNP_ERR_ARG, "hmat_ranks: NULL argument");
    MNP_CHECK(h->h->filled, MNP_ERR_STATE, "hmat_ranks: not filled");
    std::copy(h->h->rank_host.begin(), h->h->rank_host.end(), ranks);
  });
}

}  // extern "C"
