// H-matrix of the retarded Green function (full leaf blocks + ACA blocks).
#pragma once

#include <vector>

#include "mnp_linalg.hpp"

namespace mnp {

struct LayerPairs;  // mnp_layertab.hpp
struct LayerFill;   // hmatrix.hip

constexpr int MNP_HMV_MAXRHS = 8;  // right-hand sides per H-matvec pass

struct HMat {
  const GreenRet* g;
  int64_t n, nfull, nlow;
  double htol;
  int kmax;
  int prec;  // MNP_FP64 or MNP_FP32 storage of blocks and factors
  int key = -1;
  bool filled = false;
  // block structure: [row_begin, row_end, col_begin, col_end) in cluster order
  std::vector<int32_t> fblk_h, lblk_h, diag_h;
  std::vector<int64_t> lrow_h;
  std::vector<int32_t> rank_host;
  int64_t nleaf = 0, ntile = 0, nfullvals = 0, nlowvals = 0, poolcap = 0, trows = 0;
  int maxrank = 0;
  int32_t *perm = nullptr, *fblk = nullptr, *lblk = nullptr, *rank = nullptr, *d_lentry = nullptr;
  int32_t *d_tilebeg = nullptr, *d_tileleaf = nullptr, *d_rowtile = nullptr;  // row tiles of <= 32 rows
  // d_toff: first row of each low-rank block's t_b in the product's t
  int64_t *d_voff = nullptr, *d_lrow = nullptr, *d_loffs = nullptr, *d_loff = nullptr, *d_toff = nullptr;
  double *cpos = nullptr, *cnvec = nullptr, *carea = nullptr;
  // product plan: row-tile descriptors (static); column groups (static) and
  // their slices and first-phase items (per fill)
  std::vector<int32_t> lentry_h, tilebeg_h, tileleaf_h, rowtile_h;
  std::vector<int64_t> loffs_h, voff_h;
  int32_t *d_tdbeg = nullptr, *d_tlda = nullptr, *d_tK = nullptr;
  int64_t *d_taoff = nullptr, *d_tboff = nullptr;
  std::vector<int32_t> grp_beg_h, grp_blk_h;  // blocks of each column group (CSR)
  std::vector<int32_t> slice_h, imeta_h, multi_h;
  std::vector<int64_t> toff_h, soff_h, ioff_h;
  int64_t nslice = 0, nitems = 0, nmulti = 0, tpartlen = 0;
  int64_t slicecap = 0, soffcap = 0, ioffcap = 0, imetacap = 0, multicap = 0, rowblkcap = 0, rpanelcap = 0;
  int32_t *d_slice = nullptr, *d_imeta = nullptr, *d_multi = nullptr, *d_rowblk = nullptr;
  int64_t *d_soff = nullptr, *d_ioff = nullptr;
  void* rpanel = nullptr;  // column-group slices of R^T (k_hmv_pack_cols)
  int32_t *d_dsel = nullptr, *d_dlb = nullptr, *d_droff = nullptr;  // low-rank descriptors (k_plan_low)
  int64_t nlowdesc = 0;
  // row panels (k_hmv_pack, per fill): per tile the columns of its full
  // blocks (static) and the tile of each low-rank descriptor; per fill the
  // tiles' first columns and panel offsets, the column sources, the panels
  std::vector<int64_t> tkfull_h, tcol_h, poff_h;
  std::vector<int32_t> dtile_h, lowdesc_block_h;
  int64_t *d_tcol = nullptr, *d_poff = nullptr;
  int32_t* d_cidx = nullptr;
  void* panel = nullptr;
  int64_t panelcap = 0, cidxcap = 0;
  // coarsening of low-rank blocks (coarsen): (block, tile) items and the
  // blocks' first coarse entries, built at the first call
  mutable int32_t* d_citem = nullptr;
  mutable int64_t* d_ccoff = nullptr;
  mutable int64_t ncitem = 0, ncentry = 0;
  void* fvals = nullptr;  // full blocks, column-major per block, complex (prec)
  void* pool = nullptr;   // per low-rank block: L (m x r, row-tile segments) then R^T (r x n), complex (prec)

  HMat(const GreenRet* g, int prec, const int64_t* perm, int64_t nfull, const int64_t* full, int64_t nlow,
       const int64_t* low, double htol, int kmax);
  ~HMat();
  HMat(const HMat&) = delete;
  HMat& operator=(const HMat&) = delete;
  // Fills the operator `key` (MNP_KEY_*) at wavenumber k.
  void fill(double2 k, int key);
  // Fills component comp (p, ss, hs, sh, hh = 0..4) of the substrate
  // operator `key` (MNP_KEY_G or MNP_KEY_H2) at the outside wavenumber k:
  // the reflected part from P (tables and refined near-field values of this
  // wavelength) plus, if `direct`, the direct part.
  void fill_layer(double2 k, int key, int comp, bool direct, const LayerPairs& P);
  // y = H x, x and y complex FP64 in face order (n x nrhs).
  void matvec(int64_t nrhs, const double2* x, double2* y, Workspace& ws) const;
  // y[m] = A[m] x[m] for m < nm <= 4 H-matrices of one size and storage
  // precision, each phase one launch for all of them; ws[m] per product.
  static void matvec_batch(int nm, const HMat* const* A, int64_t nrhs, const double2* const* x, double2* const* y,
                           Workspace* const* ws);
  // C = P^T H P (ntile x ntile, column-major, complex FP64) for P the
  // indicator of the row tiles (<= 32 consecutive cluster indices each):
  // entry (s, t) sums the operator over the rows of tile s and the columns
  // of tile t, from the stored full blocks and low-rank factors.
  void coarsen(double2* C, Workspace& ws) const;

 private:
  void build_plan_static();
  void build_plan();
  void aca_plan();
  template <class TS>
  void fill_impl(double2 k, int key, const LayerFill* lay);
  template <class TS>
  static void matvec_batch_impl(int nm, const HMat* const* A, int64_t nrhs, const double2* const* x,
                                double2* const* y, Workspace* const* ws);
  template <class TS>
  void coarsen_impl(double2* C, Workspace& ws) const;
  std::vector<int64_t> loff_host;  // low-rank factor offsets of the last fill
  // ACA work plan (aca_plan, at the first fill): 4 block ids per workgroup
  // item, blocks in work order, per-block scratch offsets relative to their
  // batch, batch boundaries over items and over blocks
  std::vector<int32_t> aca_items_h, aca_order_h;
  std::vector<int64_t> aca_soff_h, aca_icut, aca_bcut;
  int64_t aca_maxbatch = 0;
  int32_t *d_aca_items = nullptr, *d_aca_order = nullptr;
  int64_t* d_aca_soff = nullptr;
};

}  // namespace mnp


