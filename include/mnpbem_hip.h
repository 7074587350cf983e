/*
 * mnpbem_hip.h - C ABI of the MI355X (gfx950) retarded-BEM library
 * (libmnpbem_hip.so).
 *
 * Drop-in boundary for MNPBEM's full-Maxwell (retarded) BEM hot path.  The
 * Matlab classes keep their calling conventions; a MEX adapter (or the
 * Python host layer in mnpbem_amd/) converts Matlab arrays and calls these
 * functions.  Each entry point names the reference interface it replaces.
 *
 * Conventions
 *   - Every function returns MNP_OK (0) or an MNP_ERR_* code; the message of
 *     the last failure on the calling thread is mnp_last_error().  Nothing
 *     here calls exit().
 *   - Arrays are column-major (Matlab/LAPACK layout).  Geometry arrays of n
 *     faces are n x 3 (x column, then y, then z).
 *   - Complex arrays are interleaved (re, im) pairs: complex128 for
 *     precision MNP_FP64, complex64 for MNP_FP32.
 *   - Pointers documented as "device" address GPU memory (HBM) of the current
 *     device; "host" pointers address CPU memory.  Work is queued on the
 *     stream set with mnp_set_stream (default: the null stream).
 *   - Face indices are 0-based (MEX adapters subtract 1 from Matlab faces).
 */
#ifndef MNPBEM_HIP_H
#define MNPBEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double re, im;
} mnp_complex;

enum {
  MNP_OK = 0,
  MNP_ERR_ARG = 1,       /* invalid argument (sizes, NULL pointers, options)   */
  MNP_ERR_HIP = 2,       /* HIP runtime failure                                */
  MNP_ERR_SINGULAR = 3,  /* exactly singular pivot in an LU factorisation      */
  MNP_ERR_NOCONV = 4,    /* iterative refinement did not reach its tolerance   */
  MNP_ERR_INTERNAL = 5,
  MNP_ERR_STATE = 6      /* call out of order (e.g. solve before initmat)      */
};

enum {
  MNP_FP64 = 0,
  MNP_FP32 = 1,
  /* mnp_gemm only: complex64 operands, products from hi/lo bf16 splits on
   * bf16 MFMA with FP32 accumulation (the kernel of MNP_SOLVER_MIXED_BF16) */
  MNP_FP32_SPLIT_BF16 = 2
};
enum { MNP_KEY_G = 0, MNP_KEY_F = 1, MNP_KEY_H1 = 2, MNP_KEY_H2 = 3 };

/* Solver variants of mnp_bemret_create. */
enum {
  MNP_SOLVER_FP64 = 0,  /* complex FP64 chain: the plain drop-in of bemret      */
  MNP_SOLVER_MIXED = 1, /* complex FP32 chain as preconditioner + FP64 refinement
                           with the exact operator (bemretiter, precond 'full') */
  MNP_SOLVER_MIXED_BF16 = 2 /* as MNP_SOLVER_MIXED, with the chain's large
                           complex64 products formed from hi/lo bf16 splits of
                           the FP32 operands on bf16 MFMA (about 2^-16 relative
                           per product); the refinement is unchanged */
};

/* ------------------------------------------------------------- library */
const char* mnp_last_error(void);
int mnp_version(void);
/* Selects the HIP device used by subsequent calls. */
int mnp_set_device(int device);
/* Stream for all queued work (a hipStream_t passed as void*; NULL = null stream). */
int mnp_set_stream(void* stream);
void* mnp_get_stream(void);
int mnp_synchronize(void);

/* Kernel timers, the counterpart of the tic/toc timer map of
 * mex/hlib/hoptions.h:65-72.  When enabled, every launch of the listed
 * kernel families is bracketed by HIP events on the library stream;
 * mnp_timers_read synchronises and sums the elapsed times.  work is the
 * algorithmic work of the family: real flops (gemm: 8 m n k per complex
 * product) or bytes written (assembly). */
enum {
  MNP_TIMER_GEMM_FP64 = 0, /* complex GEMM, complex128, large-tile launches      */
  MNP_TIMER_GEMM_FP32 = 1, /* complex GEMM, complex64, large-tile launches       */
  MNP_TIMER_PANEL = 2,     /* cooperative LU panel                              */
  MNP_TIMER_ASSEMBLE = 3,  /* dense G/H assembly and H-matrix full blocks        */
  MNP_TIMER_HMATVEC = 4,   /* H-matrix times matrix (bytes of stored blocks)     */
  MNP_TIMER_ACA = 5,       /* ACA of admissible blocks                           */
  MNP_TIMER_GEMM_SBF16 = 6, /* complex64 GEMM on split-bf16 MFMA, large tiles    */
  MNP_TIMER_COUNT = 7
};
typedef struct {
  int64_t launches[MNP_TIMER_COUNT];
  double ms[MNP_TIMER_COUNT];
  double work[MNP_TIMER_COUNT];   /* algorithmic flops (GEMM) or bytes (others)   */
  double bytes[MNP_TIMER_COUNT];  /* algorithmic global-memory bytes              */
} mnp_timers;
/* on != 0 enables and resets the timers; 0 disables them. */
int mnp_timers_enable(int on);
int mnp_timers_read(mnp_timers* out);

/* ------------------------------------------------ Green function (greenret)
 * Replaces compgreenret(p, p, op) for one closed particle of flat triangles:
 *   Greenfun/@compgreenret/private/init.m:4-31 -> @greenret/private/init.m
 *   (refinematrix, self/near-field Taylor tables, deriv 'cart') and
 *   @compgreenret/private/initclosed.m (closed-surface sum rule).
 * Everything, including the refinement tables, is computed on the GPU.
 */
typedef struct mnp_greenret_s* mnp_greenret_t;

typedef struct {
  int order;          /* order of the exp(ikr) expansion (bemoptions: 5)      */
  double rel_cutoff;  /* RelCutoff (bemoptions: 3)                            */
  double abs_cutoff;  /* AbsCutoff (0)                                        */
  int npol_rho;       /* polar self-integration rule, radial points - 1 (7)   */
  int npol_phi;       /* polar self-integration rule, angular points - 1 (5)  */
  int closed;         /* 1: closed surface, apply the sum rule (comparticle   */
                      /*    closed argument)                                  */
} mnp_green_options;

/* Fills op with bemoptions('sim','ret') defaults (Misc/bemoptions.m:18-24,
 * Misc/integration/@quadface/private/init.m:8-15). */
void mnp_green_options_default(mnp_green_options* op);

/* verts: host nverts x 3 (nm); faces: host nfaces x 3, 0-based vertex indices,
 * normal (v1-v2)x(v3-v2) pointing out of the particle (norm_flat.m:36-38). */
int mnp_greenret_create(const double* verts, int64_t nverts, const int64_t* faces,
                        int64_t nfaces, const mnp_green_options* op, mnp_greenret_t* out);
/* As mnp_greenret_create, with the refined pairs (refinematrix.m's
 * selection) of `like`, a Green function of the same faces, instead of a
 * selection for this mesh.  No counterpart in the reference; for moving
 * meshes whose elements should stay continuous in the vertices (finite
 * differences, shape optimisation): the selection is piecewise constant in
 * the vertices and otherwise makes G and H jump where a pair crosses the
 * cutoff. */
int mnp_greenret_create_like(const double* verts, int64_t nverts, const int64_t* faces,
                             int64_t nfaces, const mnp_green_options* op, mnp_greenret_t like,
                             mnp_greenret_t* out);
/* compgreenret(p, p, op) for a compound of nparts parts (replaces
 * Greenfun/@compgreenret/private/init.m:4-31 with Particles/@compound/
 * connect.m and initclosed.m:19-89): faces of part q are
 * [part_offsets[q], part_offsets[q + 1]) of the concatenated mesh; inout
 * (nparts x 2, column-major, 1-based media <= 16) the inside and outside
 * medium of each part; closed_mask[q] has bit r set when part r belongs to
 * the closed surface of part q (0: no closed-surface correction for part
 * q's faces), as comparticle's closed arguments give it.  At most 32 parts.
 * Solvers on it take mnp_bemret_initmat_media. */
int mnp_greenret_create_compound(const double* verts, int64_t nverts, const int64_t* faces,
                                 int64_t nfaces, int32_t nparts, const int64_t* part_offsets,
                                 const int32_t* inout, const uint32_t* closed_mask,
                                 const mnp_green_options* op, mnp_greenret_t* out);
int mnp_greenret_destroy(mnp_greenret_t g);
/* Number of faces and of refined (self + near-field) matrix elements. */
int mnp_greenret_size(mnp_greenret_t g, int64_t* nfaces, int64_t* nrefined);
/* Copies face geometry to host arrays (each may be NULL): pos, nvec, tvec1,
 * tvec2 are n x 3, area n (Particles/@particle/private/norm_flat.m). */
int mnp_greenret_geometry(mnp_greenret_t g, double* pos, double* nvec, double* tvec1,
                          double* tvec2, double* area);
/* Copies the refinement tables to host arrays (each may be NULL):
 * rows/cols (nrefined, column-major order as Matlab find(ir)), gtab
 * nrefined x (order+1), ftab nrefined x 3 x (order+1)  (greenret.g, .f). */
int mnp_greenret_refined(mnp_greenret_t g, int64_t* rows, int64_t* cols, double* gtab,
                         double* ftab);
/* g{i,j}.G/.F/.H1/.H2(k) as one dense n x n matrix (device out, leading
 * dimension ld >= n) at wavenumber k (Greenfun/@greenret/private/eval1.m). */
int mnp_greenret_eval(mnp_greenret_t g, mnp_complex k, int key, int precision, void* out,
                      int64_t ld);

/* -------------------------------------------------- dense solver (bemret)
 * Replaces bemsolver(p, op) -> bemret (BEM/@bemret): initmat builds the
 * chain for one wavelength (private/initmat.m:5-72), solve applies it
 * (mldivide.m:11-49).  With MNP_SOLVER_MIXED the chain is built in complex
 * FP32 and used as preconditioner of the exact operator
 * (BEM/@bemretiter/private/afun.m, mfun.m) with FP64 residuals until the
 * relative residual is below refine_tol.
 */
typedef struct mnp_bemret_s* mnp_bemret_t;

int mnp_bemret_create(mnp_greenret_t g, int solver, mnp_bemret_t* out);
int mnp_bemret_destroy(mnp_bemret_t bem);
/* eps1: dielectric function inside the particle, eps2 outside, at enei (nm). */
int mnp_bemret_initmat(mnp_bemret_t bem, double enei, mnp_complex eps1, mnp_complex eps2);
/* Excitation potentials (device, complex128; NULL means zero): phi* are
 * n x npol, a* are n x 3 x npol.  Outputs (device, complex128): sig1, sig2
 * n x npol, h1, h2 n x 3 x npol (compstruct fields of mldivide.m:48-49). */
int mnp_bemret_solve(mnp_bemret_t bem, int64_t npol, const void* phi1, const void* phi1p,
                     const void* a1, const void* a1p, const void* phi2, const void* phi2p,
                     const void* a2, const void* a2p, void* sig1, void* sig2, void* h1,
                     void* h2);
/* bem(enei) for a solver on a compound Green function
 * (mnp_greenret_create_compound): eps[m - 1] is the dielectric function of
 * medium m (1 <= m <= nmed), each face taking those of its part's inside and
 * outside media.  Replaces BEM/@bemret/private/initmat.m:5-72 with eps1,
 * eps2 per face (lines 19-24) and, when some part's inside medium is some
 * part's outside medium (con{1,2} ~= 0), the full L1 = G1 eps1 G1^{-1},
 * L2 = G2 eps2 G2^{-1} (lines 34-42). */
int mnp_bemret_initmat_media(mnp_bemret_t bem, double enei, int32_t nmed, const mnp_complex* eps);
/* Refinement statistics of the last solve (MIXED): sweeps and final relative residual. */
int mnp_bemret_stats(mnp_bemret_t bem, int* iterations, double* relres);
/* Refinement settings for MNP_SOLVER_MIXED (defaults: 1e-12, 10). */
int mnp_bemret_set_refinement(mnp_bemret_t bem, double tol, int maxit);
/* mnp_bemret_initmat for `count` solvers (distinct handles, e.g. one per
 * particle of a sweep), each at its own wavelength and dielectric functions:
 * the assemblies and factorisations run four at a time on concurrent worker
 * streams (after the work the caller's stream holds), the caller's stream
 * then waits for them.  No counterpart in the reference; it batches the
 * per-particle subsref '()' of BEM/@bemret/subsref.m:23-26. */
int mnp_bemret_initmat_many(int count, const mnp_bemret_t* bems, const double* enei,
                            const mnp_complex* eps1, const mnp_complex* eps2);
/* ------------------------------------------------- batched dense solves
 * Many (particle, wavelength) members of one face count solved together:
 * plane-wave cross sections of each member (the spectra loops of
 * Demo/planewave/ret/demospecret1.m:34-42 over bemsolver(p, op) \
 * planewave(...)(p, enei) and ext/sca/abs, run for many particles).  Every
 * stage handles all members in one launch per step (assembly of
 * Greenfun/@greenret/private/eval1.m, the chain of BEM/@bemret/private/
 * initmat.m:5-72 by blocked LU and triangular solves without row exchanges,
 * mldivide.m:11-49, and refinement sweeps with the exact FP64 operator of
 * BEM/@bemretiter/private/afun.m:6-35), captured once as a hipGraph and
 * replayed per batch.  capacity: members per launch; npol <= 8 plane waves
 * in the outside medium; npinf: size of the pinfty surface for the
 * scattering cross section (0: sca not computed).  solver: MNP_SOLVER_FP64
 * (FP64 chain, 1 sweep) or MNP_SOLVER_MIXED (complex FP32 chain, 3 sweeps).
 * No counterpart in the reference. */
typedef struct mnp_bemret_batch_s* mnp_bemret_batch_t;
int mnp_bemret_batch_create(int64_t nfaces, int64_t capacity, int solver, int64_t npol,
                            int64_t npinf, mnp_bemret_batch_t* out);
int mnp_bemret_batch_destroy(mnp_bemret_batch_t batch);
/* use_graph: capture and replay (1, default) or launch directly (0);
 * variant: panel width and product tile, 0: 32 / 64 x 64, 1: 128 / 64 x 64,
 * 2: 128 / 128 x 64, 3: 128 / 128 x 96 (3-multiplication form, MIXED only),
 * 4: 256 / 128 x 64, 5: 256 / 128 x 96 (MIXED only); -1 (default): the
 * fastest, each candidate timed once per (faces, capacity, solver); sweeps:
 * refinement sweeps (-1: the solver's default). */
int mnp_bemret_batch_set_options(mnp_bemret_batch_t batch, int use_graph, int variant, int sweeps);
/* Cross sections of count <= capacity members: member m has the Green
 * function greens[m] (nfaces faces), wavelength enei[m], eps_in[m] inside and
 * eps_out[m] outside (the medium of the plane waves); pol, dir host 3 x npol
 * (column-major, pol perpendicular to dir) shared by all members; pinfty
 * normals and areas on the device (npinf x 3, npinf).  Out (host): ext, sca,
 * abs count x npol (member fastest in npol), status[m] (0 solved; 1 zero
 * pivot; 3 relative residual above 1e-12 after the sweeps); relres[m] (may
 * be NULL).  A member with status != 0 has no valid cross sections: solve it
 * with the dense solver. */
int mnp_bemret_batch_cross_sections(mnp_bemret_batch_t batch, int64_t count,
                                    const mnp_greenret_t* greens, const double* enei,
                                    const mnp_complex* eps_in, const mnp_complex* eps_out,
                                    const double* pol, const double* dir, const void* pinfty_nvec,
                                    const void* pinfty_area, double* ext, double* sca, double* abs,
                                    int* status, double* relres);
/* The variant in use, the tuner's timing of the pipeline per variant (ms, 6
 * entries; 0 when not timed) and the captured graph's node count. */
int mnp_bemret_batch_info(mnp_bemret_batch_t batch, int* variant, double* variant_ms,
                          int64_t* graph_nodes);

/* Sensitivity of a solution to the inside dielectric function (no counterpart
 * in the reference; it differentiates the operator that bemretiter/private/
 * afun.m:6-35 applies and initmat.m:5-72 inverts).  Given the solution
 * sig1, sig2, h1, h2 of mnp_bemret_solve at the current wavelength for an
 * excitation from outside (phi1, a1 and their derivatives zero), returns its
 * complex derivative with respect to eps1: A dx = -(dA/deps1) x, with eps1
 * entering A through k1 = k0 sqrt(eps1) in G1, H1 (their k-derivatives are
 * assembled, refined terms included) and explicitly in the alpha and De
 * rows.  The solve reuses the forward factors (and, for MIXED, the FP64
 * refinement).  Layouts as mnp_bemret_solve's outputs. */
int mnp_bemret_tangent_eps1(mnp_bemret_t bem, int64_t npol, const void* sig1, const void* sig2,
                            const void* h1, const void* h2, void* dsig1, void* dsig2, void* dh1,
                            void* dh2);

/* Transposed solve with the forward factors (no counterpart in the
 * reference; the adjoint of mldivide.m:11-49 for the operator of
 * bemretiter/private/afun.m:6-35): lam = A^{-T} c with the plain transpose.
 * c (device, complex128, n x 8 npol) is laid out as the solution
 * [sig1 | h1 | sig2 | h2] (columns q, then c + 3 q for vectors, as
 * mnp_bemret_solve's outputs stacked); lam (device, n x 8 npol) as the
 * excitation [phi | a | De | alpha] of excitation.m:21-29.  FP64 solvers
 * apply the transposed chain (factors of A^T formed from the forward LU
 * factors, the inverses of Sigma1 and Deltai transposed); MIXED refines the
 * FP32 transposed chain with the exact FP64 A^T to the refinement tolerance
 * (mnp_bemret_stats reports it).  Valid until the next initmat. */
int mnp_bemret_solve_transposed(mnp_bemret_t bem, int64_t npol, const void* c, void* lam);
/* Gradient of the weighted plane-wave extinction sum_q w_q sigma_ext,q
 * (@planewaveret/extinction.m:12-16 with the far field of
 * @spectrumret/farfield.m:59-84) with respect to the vertex positions and the
 * inside dielectric function, by one transposed solve (no counterpart in the
 * reference).  sig1..h2: the solution of mnp_bemret_solve at the current
 * wavelength for the plane wave pol, dir (host npol x 3, npol <= 8, pol
 * perpendicular to dir) from outside, a2/a2p of mnp_planewave_potential;
 * weights: host npol (NULL: all 1).  dverts: host nverts x 3 (the mesh of
 * mnp_greenret_create), d/d vertex coordinates, including the derivatives of
 * the refined near-field and self integrals (greenret/private/init.m:42-177)
 * and of the closed-surface sum rule (compgreenret/private/initclosed.m:
 * 19-89); the refinement selection is held fixed.  deps1: host 2,
 * (d/d Re eps1, d/d Im eps1).  Either output may be NULL. */
int mnp_bemret_extinction_gradient(mnp_bemret_t bem, int64_t npol, const double* pol, const double* dir,
                                   const double* weights, const void* sig1, const void* sig2,
                                   const void* h1, const void* h2, double* dverts, double* deps1);

/* ------------------------------------------------ dense building blocks
 * Replace the LAPACK/BLAS calls behind Matlab's inv and mtimes in
 * BEM/@bemret/private/initmat.m:27-72 (zgetrf/zgetri/zgemm) and in
 * mex/hlib/basemat.cpp:105-120,224-240.  Device arrays, column-major,
 * complex128 (MNP_FP64) or complex64 (MNP_FP32).
 */
/* C = alpha A B + beta C, A m x k, B k x n. */
int mnp_gemm(int precision, int64_t m, int64_t n, int64_t k, mnp_complex alpha, const void* A,
             int64_t lda, const void* B, int64_t ldb, mnp_complex beta, void* C, int64_t ldc);
/* In-place LU with partial pivoting of an m x n matrix (m >= n); ipiv
 * (device int32, n) holds 0-based row interchanges.  *info (host) is 0 or
 * the 1-based column of the first exactly zero pivot. */
int mnp_getrf(int precision, int64_t m, int64_t n, void* A, int64_t lda, int32_t* ipiv, int* info);
/* LU without row exchanges (the bemret chain's first attempt).  *would_swap
 * (host) is set to 1 when partial pivoting would have exchanged rows, i.e.
 * when the factors differ from mnp_getrf's; otherwise they are the same
 * factors with ipiv = 0, 1, ..., n-1. */
int mnp_getrf_nopiv(int precision, int64_t m, int64_t n, void* A, int64_t lda, int* info,
                    int* would_swap);
/* B <- A^{-1} B (B n x nrhs) from mnp_getrf factors. */
int mnp_getrs(int precision, int64_t n, int64_t nrhs, const void* LU, int64_t lda,
              const int32_t* ipiv, void* B, int64_t ldb);
/* B <- B A^{-1} (B m x n) from mnp_getrf factors. */
int mnp_getrs_right(int precision, int64_t m, int64_t n, const void* LU, int64_t lda,
                    const int32_t* ipiv, void* B, int64_t ldb);
/* X = A^{-1} (n x n) from mnp_getrf factors. */
int mnp_getri(int precision, int64_t n, const void* LU, int64_t lda, const int32_t* ipiv, void* X,
              int64_t ldx);

/* ------------------------------------------------------ H-matrices (ACA)
 * Replaces hmatrix(tree) filled by aca.compgreenret (Greenfun/+aca/
 * @compgreenret/eval.m:16-62): full leaf blocks element-wise with
 * refinement (@hmatrix/fillval.m, @greenret/private/eval2.m), admissible
 * blocks by partially pivoted ACA of the collocation kernel (MEX
 * hmatgreenret, mex/hlib/aca.cpp:158-218, mex/acagreen/acagreen.cpp:79-149),
 * and the H-matrix times matrix product (MEX hmatmul1, mex/hlib/hmatrix.h:
 * 107-116).  The cluster tree and block tree are built by the caller
 * (@clustertree, admissibility.m), as in the reference.
 */
typedef struct mnp_hmat_s* mnp_hmat_t;
/* precision (MNP_FP64 or MNP_FP32) is the storage of the full blocks and
 * low-rank factors; ACA runs in FP64 either way and products accumulate in
 * FP64.  perm[c] = face index of cluster index c (n entries); full and low are
 * 4 x nblocks (column-major) [row_begin, row_end, col_begin, col_end) in
 * cluster indices; the blocks must tile the n x n matrix and the full blocks
 * must include the diagonal leaf blocks.  htol and kmax are the ACA
 * tolerance and maximum rank (hoptions; kmax <= 128). */
int mnp_hmat_create(mnp_greenret_t g, int precision, const int64_t* perm, int64_t nfull,
                    const int64_t* full, int64_t nlow, const int64_t* low, double htol, int kmax,
                    mnp_hmat_t* out);
int mnp_hmat_destroy(mnp_hmat_t h);
/* Fills G, F, H1 (F + 2 pi) or H2 (F - 2 pi) at wavenumber k (key MNP_KEY_*). */
int mnp_hmat_fill(mnp_hmat_t h, mnp_complex k, int key);
/* Component comp (p, ss, hs, sh, hh = 0..4) of the outside Green function
 * above a substrate as an H-matrix (Greenfun/+aca/@compgreenretlayer/
 * private/eval2.m for g{2,2}): the tabulated reflected G (key MNP_KEY_G) or
 * F (MNP_KEY_H2) of gl at enei, with the refined near pairs in the full
 * blocks, plus, if direct != 0, the direct part at wavenumber k (with the
 * -2 pi of H2); ACA on the sum of the two in the admissible blocks. */
int mnp_hmat_fill_layer(mnp_hmat_t h, mnp_complex k, int key, int comp, int direct, struct mnp_greenretlayer_s* gl,
                        double enei, mnp_complex eps_up, mnp_complex eps_down);
/* y = H x (device, complex128, face order, n x nrhs). */
int mnp_hmat_matvec(mnp_hmat_t h, int64_t nrhs, const void* x, void* y);
/* Stored entries of full and low-rank blocks, largest rank and compression
 * (stored entries / n^2) of the last fill (hmatrix/compression.m). */
int mnp_hmat_stats(mnp_hmat_t h, int64_t* full_entries, int64_t* lowrank_entries, int* max_rank,
                   double* compression);
/* Ranks of the low-rank blocks (host, nlow entries) of the last fill. */
int mnp_hmat_ranks(mnp_hmat_t h, int32_t* ranks);
/* Coarse operator P^T H P of the last fill over the row tiles of the
 * product plan (each leaf cluster cut into near-equal runs of at most 32
 * consecutive cluster indices): *ntile (host) their number; tilebeg (host,
 * ntile + 1, may be NULL) their first cluster indices; C (device, may be
 * NULL) ntile x ntile complex128, column-major, entry (s, t) the sum of H
 * over the rows of tile s and the columns of tile t, summed from the stored
 * full blocks and low-rank factors.  No counterpart in the reference: the
 * coarse level of the iterative solver's preconditioner. */
int mnp_hmat_coarsen(mnp_hmat_t h, int64_t* ntile, int64_t* tilebeg, void* C);

/* ------------------------------------------ iterative solver (bemretiter)
 * Replaces bemsolver(p, op) with op.iter set -> bemretiter (BEM/@bemretiter):
 * initmat fills G1, G2, H1, H2 as H-matrices (private/initmat.m:5-31) and
 * the preconditioner; solve runs GMRES without restart on the exact
 * operator afun (private/afun.m:6-35), preconditioned by mfun
 * (bemiter/solve.m:33-36, private/mfun.m) from the right in the flexible
 * form, and stops when |b - A x| <= tol |b|.  In place of the reference's
 * rank-4 H-matrix chain (private/initprecond.m), the preconditioner couples
 * the surface on two levels: the bemret chain on the H-matrices aggregated
 * over tiles of at most 32 faces, then the bemret chain applied exactly on
 * diagonal blocks of at most 32 faces (see mnp_bemretiter_set_options).
 * Precision and block arguments as for mnp_hmat_create. */
typedef struct mnp_bemretiter_s* mnp_bemretiter_t;
int mnp_bemretiter_create(mnp_greenret_t g, int precision, const int64_t* perm, int64_t nfull,
                          const int64_t* full, int64_t nlow, const int64_t* low, double htol,
                          int kmax, mnp_bemretiter_t* out);
int mnp_bemretiter_destroy(mnp_bemretiter_t it);
/* GMRES tolerance and iteration limit (bemiter.options: 1e-6, 100) and the
 * preconditioner: 0 none, 1 the block chain alone, 2 (default) two levels:
 * the bemret chain on the H-matrices aggregated over their row tiles (P^T G1
 * P, P^T G2 P, P^T H1 P, P^T H2 P, tiles' mean normals; FP32 MFMA chain)
 * followed by the block chain on what remains, z = P y + B^{-1}(v - A P y),
 * y = C^{-1} P^T v. */
int mnp_bemretiter_set_options(mnp_bemretiter_t it, double tol, int maxit, int precond);
/* Coarse level of precond 2: aggregates of tiles_per_aggregate consecutive
 * row tiles (1..64; default 2).  Fewer, larger
 * aggregates make the coarse chain cheaper and GMRES take more iterations.
 * No counterpart in the reference (its preconditioner is the H-matrix chain
 * of private/initprecond.m). */
int mnp_bemretiter_set_coarse(mnp_bemretiter_t it, int tiles_per_aggregate);
int mnp_bemretiter_initmat(mnp_bemretiter_t it, double enei, mnp_complex eps1, mnp_complex eps2);
/* Iterative solver above a substrate (bemsolver with op.layer and iterative
 * options -> BEM/@bemretlayeriter: private/initmat.m:1-42, afun.m:1-47,
 * excitation.m, solve.m; the H-matrices of the components p, ss, hs, sh, hh
 * of Greenfun/+aca/@compgreenretlayer/private/eval2.m, lowrank1.m,
 * lowrank2.m with mex/hmatgreentab1.cpp, hmatgreentab2.cpp,
 * acagreen/greentab.cpp): gl the particle's reflected Green function
 * (mnp_greenretlayer_create, its table covering the particle); the other
 * arguments as mnp_bemretiter_create.  ACA runs on the sum of the direct and
 * the tabulated reflected kernel where the reference compresses the two
 * separately and truncates their concatenation; the preconditioner is that
 * of mnp_bemretiter_set_options with the p components in place of G2, H2.
 * Options, solve, stats and destroy as for mnp_bemretiter_*. */
int mnp_bemretiter_create_layer(mnp_greenret_t g, struct mnp_greenretlayer_s* gl, int precision, const int64_t* perm,
                                int64_t nfull, const int64_t* full, int64_t nlow, const int64_t* low,
                                double htol, int kmax, mnp_bemretiter_t* out);
/* eps_in inside the particle, eps_up (the particle's outside) above and
 * eps_down below the interface (bemretlayeriter/private/initmat.m). */
int mnp_bemretiter_initmat_layer(mnp_bemretiter_t it, double enei, mnp_complex eps_in, mnp_complex eps_up,
                                 mnp_complex eps_down);
/* Same arguments as mnp_bemret_solve (npol <= 8). */
int mnp_bemretiter_solve(mnp_bemretiter_t it, int64_t npol, const void* phi1, const void* phi1p,
                         const void* a1, const void* a1p, const void* phi2, const void* phi2p,
                         const void* a2, const void* a2p, void* sig1, void* sig2, void* h1,
                         void* h2);
/* Last solve: flag (0 converged, 1 not), GMRES relative residual estimate,
 * iterations, relative residual |b - A x| / |b| recomputed from x;
 * compression of G1 (bemiter/setiter.m, hinfo.m). */
int mnp_bemretiter_stats(mnp_bemretiter_t it, int* flag, double* relres, int* iter,
                         double* true_relres, double* compression);

/* ----------------------------- iterative quasistatic solver (bemstatiter)
 * Replaces bemsolver(p, op) with op.sim = 'stat' and op.iter set ->
 * bemstatiter (BEM/@bemstatiter): F of the static Green function as an
 * H-matrix (private/init.m:11-13, filled once: it does not depend on the
 * wavelength), Lambda = 2 pi (eps1 + eps2) / (eps1 - eps2) per wavelength
 * (private/initmat.m:12), and GMRES on afun(x) = -(F x + Lambda x)
 * (private/afun.m:7, solve.m:17-20) with right-hand side phip, so that the
 * solution is sig.  Flexible right preconditioning by the exact inverses of
 * -(Lambda + F) on diagonal blocks of at most 32 faces, in place of the
 * reference's H-matrix LU or dense inverse (private/initmat.m:22-30).
 * Precision and block arguments as for mnp_hmat_create. */
typedef struct mnp_bemstatiter_s* mnp_bemstatiter_t;
int mnp_bemstatiter_create(mnp_greenret_t g, int precision, const int64_t* perm, int64_t nfull,
                           const int64_t* full, int64_t nlow, const int64_t* low, double htol,
                           int kmax, mnp_bemstatiter_t* out);
int mnp_bemstatiter_destroy(mnp_bemstatiter_t it);
/* GMRES tolerance and iteration limit (bemiter.options: 1e-6, 100); precond 0 disables it. */
int mnp_bemstatiter_set_options(mnp_bemstatiter_t it, double tol, int maxit, int precond);
int mnp_bemstatiter_initmat(mnp_bemstatiter_t it, double enei, mnp_complex eps1, mnp_complex eps2);
/* sig (n x npol, device, complex128) from phip (npol <= 8). */
int mnp_bemstatiter_solve(mnp_bemstatiter_t it, int64_t npol, const void* phip, void* sig);
/* As mnp_bemretiter_stats; compression of F. */
int mnp_bemstatiter_stats(mnp_bemstatiter_t it, int* flag, double* relres, int* iter,
                          double* true_relres, double* compression);

/* --------------------------------------------- plane wave and observables
 * Simulation/retarded/@planewaveret/potential.m:47-88 for a particle whose
 * outside is the embedding medium: a2 and a2p (device out, complex128,
 * n x 3 x npol).  pol, dir: host npol x 3, pol perpendicular to dir. */
int mnp_planewave_potential(mnp_greenret_t g, double enei, mnp_complex eps_medium,
                            int64_t npol, const double* pol, const double* dir, void* a2,
                            void* a2p);
/* Far fields of sig2/h2 in directions dir (device ndir x 3 doubles):
 * e, h device out ndir x 3 x npol complex128 (@spectrumret/farfield.m). */
int mnp_farfield(mnp_greenret_t g, double enei, mnp_complex eps_medium, int64_t ndir,
                 const void* dir, int64_t npol, const void* sig2, const void* h2, void* e,
                 void* h);
/* Extinction (optical theorem), scattering (Poynting flux over the surface
 * with normals pinfty_nvec, device npinf x 3, and areas pinfty_area, device
 * npinf) and absorption cross sections, host out (npol each), FP64
 * reductions (@planewaveret/extinction.m, scattering.m, absorption.m). */
int mnp_planewave_cross_sections(mnp_greenret_t g, double enei, mnp_complex eps_medium,
                                 int64_t npol, const double* pol, const double* dir,
                                 int64_t npinf, const void* pinfty_nvec,
                                 const void* pinfty_area, const void* sig2, const void* h2,
                                 double* ext, double* sca, double* abs_);
/* mnp_planewave_cross_sections without any device allocation: same
 * arguments, results and FP64 arithmetic (the scattering sum reduced on the
 * GPU in a fixed order instead of on the host).  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its directions, far fields and sums live in a file-scope device array of
 * 3.2 MB per GPU, held until the process ends, so npol + npinf must be at
 * most 4096 (MNP_ERR_ARG otherwise).  Calls that overlap on one GPU (several
 * host threads or streams) race on that array; the call synchronises the
 * library stream before it returns. */
int mnp_planewave_cross_sections_noalloc(mnp_greenret_t g, double enei, mnp_complex eps_medium,
                                         int64_t npol, const double* pol, const double* dir,
                                         int64_t npinf, const void* pinfty_nvec,
                                         const void* pinfty_area, const void* sig2, const void* h2,
                                         double* ext, double* sca, double* abs_);

/* ------------------------------------------------- mirror symmetry solver
 * Replaces bemsolver(p, op) for a comparticlemirror p (op.sym = 'x', 'y' or
 * 'xy') -> bemretmirror (BEM/@bemretmirror).  g is the Green function of the
 * full particle (Particles/@comparticlemirror/full, init.m:42-54): nsym = 2
 * or 4 copies of the n-face symmetry part, copy j in faces j n .. (j+1) n - 1,
 * copy 0 the part itself.  symtable is nsym x nsym row-major, +-1, row s the
 * values of symmetry s on the copies (comparticlemirror/private/init.m:29-37).
 * initmat assembles, per symmetry value s, G_s = sum_j symtable[s][j]
 * G(part, copy j) (Greenfun/@compgreenretmirror/eval.m:23-42) for G1, G2,
 * H1, H2 and factorises the chain of bemretmirror/private/initmat.m:25-66
 * per s.  solve takes one excitation of the symmetry part (fields as for
 * mnp_bemret_solve, n rows) with the symmetry indices symx, symy, symz
 * (0-based rows of symtable) of its x, y and z components (scalars go with
 * z; mldivide.m:30-33), factorises Sigma of that combination once per
 * wavelength (initsigmai.m:19-30) and returns sig1, sig2, h1, h2 on the
 * symmetry part (mldivide.m:25-87).  solver as for mnp_bemret_create. */
typedef struct mnp_bemretmirror_s* mnp_bemretmirror_t;
int mnp_bemretmirror_create(mnp_greenret_t g, int solver, int nsym, const int* symtable,
                            mnp_bemretmirror_t* out);
int mnp_bemretmirror_destroy(mnp_bemretmirror_t bem);
int mnp_bemretmirror_initmat(mnp_bemretmirror_t bem, double enei, mnp_complex eps1,
                             mnp_complex eps2);
int mnp_bemretmirror_solve(mnp_bemretmirror_t bem, int symx, int symy, int symz, int64_t npol,
                           const void* phi1, const void* phi1p, const void* a1, const void* a1p,
                           const void* phi2, const void* phi2p, const void* a2, const void* a2p,
                           void* sig1, void* sig2, void* h1, void* h2);
/* Refinement sweeps and final relative residual of the last mixed solve. */
int mnp_bemretmirror_stats(mnp_bemretmirror_t bem, int* iterations, double* relres);
int mnp_bemretmirror_set_refinement(mnp_bemretmirror_t bem, double tol, int maxit);
/* Values on the full particle from two solutions on the symmetry part
 * (Simulation/retarded/@planewaveretmirror/full.m:22-64 over
 * Particles/@compstructmirror/expand.m:22-46):
 *   out[j n + i, c, o] = w[2o] sign1[j][c] in1[i, c, src[o]]
 *                      + w[2o+1] sign2[j][c] in2[i, c, src[o]],
 * in1/in2 device n x ncomp x nsrc (in2 may be NULL), out device
 * (nsym n) x ncomp x nout, ncomp 1 (scalars) or 3 (vectors), sign1/sign2
 * host nsym x ncomp (+-1: the symmetry value of component c on copy j), w
 * host nout x 2, src host nout (nout <= 64). */
int mnp_mirror_full(int64_t n, int nsym, int ncomp, int64_t nsrc, int64_t nout, const int* sign1,
                    const int* sign2, const double* w, const int* src, const void* in1,
                    const void* in2, void* out);

/* ---------------------------------------------------- quasistatic solver
 * Replaces bemsolver(p, op) with op.sim = 'stat' -> bemstat (BEM/@bemstat):
 * F is the static surface derivative with refinement and the closed-surface
 * correction (Greenfun/@compgreenstat/init.m:24-55, @greenstat/private/
 * init.m), equal to the retarded F at k = 0 and assembled once at create;
 * initmat forms and factorises Lambda I + F with Lambda = 2 pi (eps1 + eps2)
 * / (eps1 - eps2) (bemstat/subsref.m:31-40); solve returns
 * sig = -(Lambda I + F)^{-1} phip (mldivide.m:13-15).  solver MNP_SOLVER_FP64,
 * MNP_SOLVER_MIXED (complex FP32 factors, FP64 refinement to 1e-12) or
 * MNP_SOLVER_MIXED_BF16 (the FP32 factorisation's large products on the
 * split-bf16 kernel, the same refinement). */
typedef struct mnp_bemstat_s* mnp_bemstat_t;
int mnp_bemstat_create(mnp_greenret_t g, int solver, mnp_bemstat_t* out);
int mnp_bemstat_destroy(mnp_bemstat_t b);
int mnp_bemstat_initmat(mnp_bemstat_t b, double enei, mnp_complex eps1, mnp_complex eps2);
/* phip, sig: device complex128 n x npol. */
int mnp_bemstat_solve(mnp_bemstat_t b, int64_t npol, const void* phip, void* sig);
/* Refinement sweeps and final relative residual of the last mixed solve. */
int mnp_bemstat_stats(mnp_bemstat_t b, int* sweeps, double* relres);
/* phi' = -nvec . pol (Simulation/static/@planewavestat/potential.m:13):
 * pol host npol x 3, phip device out n x npol complex128. */
int mnp_planewavestat_potential(mnp_greenret_t g, int64_t npol, const double* pol, void* phip);
/* Cross sections from the dipole moment p = sum area pos sig (FP64 device
 * sums): sca = 8 pi / 3 k^4 |p|^2, abs = 4 pi k Im(pol . p), ext = sca + abs
 * (@planewavestat/scattering.m:14-18, absorption.m:14-19, extinction.m);
 * host out, npol each. */
int mnp_planewavestat_cross_sections(mnp_greenret_t g, double enei, mnp_complex eps_medium,
                                     int64_t npol, const double* pol, const void* sig, double* ext,
                                     double* sca, double* abs_);

/* ---------------------------------------- quasistatic solver, substrate
 * Replaces bemsolver(p, op) with op.sim = 'stat' and op.layer (one
 * interface) -> bemstatlayer (BEM/@bemstatlayer) for a particle above the
 * interface (no face in it).  g is the particle's Green function (with the
 * closed-surface correction); gimage the Green function of the particle
 * followed by its image, the particle flipped at the interface plane
 * (compgreenstatlayer/init.m:21-24: shift, flip(.., 3), shift back), built
 * with closed = 0.  The reflected F is gimage's rows of the particle against
 * the image's columns (compgreenstat(p1, p2r), init.m:26-29), assembled once;
 * initmat forms F = F_direct + f2 F_reflected with f2 = -(eps_down - eps_up)
 * / (eps_down + eps_up) (compgreenstatlayer/eval.m:16-38) and factorises
 * eps1 H1 - eps2 H2 = (eps1 - eps2)(Lambda I + F) (bemstatlayer/subsref.m:
 * 28-32).  The handle is a mnp_bemstat_t: solve, stats and destroy are
 * mnp_bemstat_solve (sig = mat phip, mldivide.m:13), _stats, _destroy. */
int mnp_bemstatlayer_create(mnp_greenret_t g, mnp_greenret_t gimage, int solver, mnp_bemstat_t* out);
int mnp_bemstatlayer_initmat(mnp_bemstat_t bem, double enei, mnp_complex eps1, mnp_complex eps2,
                             mnp_complex eps_up, mnp_complex eps_down);
/* phi' = -nvec . e for the field e (host npol x 3, npol <= 64) at the faces
 * (@planewavestatlayer/potential.m:13-14 with the field of field.m for faces
 * above the interface). */
int mnp_planewavestatlayer_potential(mnp_greenret_t g, int64_t npol, const mnp_complex* e, void* phip);
/* Dipole moment sum_i area_i pos_i sig_i per column (host npol x 3), FP64
 * (@planewavestatlayer/extinction.m:13, scattering.m:12). */
int mnp_stat_dipole_moment(mnp_greenret_t g, int64_t npol, const void* sig, mnp_complex* dip);

/* ---------------------------------------------- retarded substrate
 * One planar interface at z0 between an upper medium (eps_up) and a lower
 * one (eps_down), particles above it: Particles/@layerstructure with one z
 * (options of layerstructure.options: ztol, rmin, zmin, semi, ratio),
 * Greenfun/@greentablayer, @greenretlayer, @compgreenretlayer,
 * BEM/@bemretlayer, Simulation/retarded/@planewaveretlayer and
 * @spectrumretlayer.  Faces lying in the interface are not supported. */
typedef struct {
  double z0, ztol, rmin, zmin, semi, ratio;
} mnp_layer_options;
/* layerstructure.options defaults (Particles/@layerstructure/options.m:22-29), z0 = 0. */
void mnp_layer_options_default(mnp_layer_options* op);

/* Table of the reflected Green function for sources at height z2 (host r:
 * nr increasing radii, z1: nz increasing observer heights, all above z0)
 * (Greenfun/@greentablayer/greentablayer.m, private/init.m:8-17). */
typedef struct mnp_layertab_s* mnp_layertab_t;
int mnp_layertab_create(const mnp_layer_options* op, int64_t nr, const double* r, int64_t nz,
                        const double* z1, double z2, mnp_layertab_t* out);
int mnp_layertab_destroy(mnp_layertab_t t);
/* Tables at the wavelengths enei (host, increasing; eps_up/eps_down there):
 * greentablayer/set.m.  The Sommerfeld integrals of Particles/@layerstructure/
 * green.m:22-127 run on the GPU, one wavefront per grid point. */
int mnp_layertab_set(mnp_layertab_t t, int64_t nenei, const double* enei, const mnp_complex* eps_up,
                     const mnp_complex* eps_down);
/* The table at enei (greentablayer/eval.m:17-50: new when nothing is set,
 * else linear in wavelength between the set tables): out (device, may be
 * NULL) nr x nz x 15 complex128 at (i nz + j) 15 + q, q = 0..4 d G,
 * 5..9 d^3 Fr / r, 10..14 d^3 Fz / zmin of the components p, ss, hs, sh, hh
 * (greentablayer/norm.m:12-27). */
int mnp_layertab_eval(mnp_layertab_t t, double enei, mnp_complex eps_up, mnp_complex eps_down,
                      void* out);

/* Reflected Green function of a particle with itself on a table
 * (Greenfun/@greenretlayer/greenretlayer.m, private/init.m, init2.m): the
 * near pairs of green.refinematrixlayer (RelCutoff/AbsCutoff of g's options)
 * get their shape-function weights once here. */
typedef struct mnp_greenretlayer_s* mnp_greenretlayer_t;
int mnp_greenretlayer_create(mnp_greenret_t g, mnp_layertab_t tab, mnp_greenretlayer_t* out);
int mnp_greenretlayer_destroy(mnp_greenretlayer_t gl);
/* Refined off-diagonal pairs and diagonal elements (host out). */
int mnp_greenretlayer_refined(mnp_greenretlayer_t gl, int64_t* npairs, int64_t* ndiag);
/* Reflected G and F = nvec . Gp at enei (greenretlayer/private/initrefl2.m):
 * G, F device out 5 x n x n complex128 (component c at c n^2), components
 * p, ss, hs, sh, hh.  Fails when a pair lies outside the table. */
int mnp_greenretlayer_eval(mnp_greenretlayer_t gl, double enei, mnp_complex eps_up,
                           mnp_complex eps_down, void* G, void* F);

/* Dense retarded solver above the substrate (BEM/@bemretlayer:
 * private/initmat.m:17-95 with unique eps inside and outside, mldivide.m);
 * the outside medium is the upper one.  solver: MNP_SOLVER_FP64 (the
 * complex FP64 chain) or MNP_SOLVER_MIXED (the chain in complex FP32,
 * refined with the FP64 operator the FP64 chain inverts until the relative
 * residual is below 1e-12, at most 10 sweeps).  Arguments of solve as
 * mnp_bemret_solve. */
typedef struct mnp_bemretlayer_s* mnp_bemretlayer_t;
int mnp_bemretlayer_create(mnp_greenret_t g, mnp_greenretlayer_t gl, int solver, mnp_bemretlayer_t* out);
int mnp_bemretlayer_destroy(mnp_bemretlayer_t bem);
int mnp_bemretlayer_initmat(mnp_bemretlayer_t bem, double enei, mnp_complex eps_in, mnp_complex eps_up,
                            mnp_complex eps_down);
int mnp_bemretlayer_solve(mnp_bemretlayer_t bem, int64_t npol, const void* phi1, const void* phi1p,
                          const void* a1, const void* a1p, const void* phi2, const void* phi2p,
                          const void* a2, const void* a2p, void* sig1, void* sig2, void* h1, void* h2);
/* LU factorisations so far that partial pivoting had to redo (each is tried
 * without row exchanges first; the factors are partial pivoting's either
 * way); refinement sweeps and final relative residual of the last solve
 * (MIXED; 0 and 0 for FP64). */
int mnp_bemretlayer_stats(mnp_bemretlayer_t bem, int* fallbacks, int* iterations, double* relres);
/* Reflected potentials and fields at points above the interface
 * (compgreenretlayer(pt, p) for a compoint in the upper medium:
 * Greenfun/@compgreenretlayer/potential.m, field.m with private/matmul2.m,
 * matmul3.m; greenretlayer with p1 the points, near pairs by
 * refinematrixlayer.m).  pts host npt x 3; fields: sig2 (n x ncol) and h2
 * (n x 3 x ncol) the outside charges and currents (device); phi (npt x
 * ncol), a, e, h (npt x 3 x ncol) device out, the reflected parts only (the
 * direct ones are mnp_potential_points / mnp_field_points). */
typedef struct mnp_pointlayer_s* mnp_pointlayer_t;
int mnp_pointlayer_create(mnp_greenret_t g, mnp_layertab_t tab, int64_t npt, const double* pts,
                          mnp_pointlayer_t* out);
int mnp_pointlayer_destroy(mnp_pointlayer_t pl);
int mnp_pointlayer_fields(mnp_pointlayer_t pl, double enei, mnp_complex eps_up, mnp_complex eps_down,
                          int64_t ncol, const void* sig2, const void* h2, void* phi, void* a, void* e, void* h);
/* Outside potentials of plane waves above the substrate at the faces
 * (Simulation/retarded/@planewaveretlayer/potential.m:20-105 with
 * Particles/@layerstructure/fresnel.m; the normal derivative of the phase
 * factors is exact where the reference differences with a 1e-8 step):
 * pol, dir host npol x 3; phi2, phi2p device n x npol, a2, a2p n x 3 x npol. */
int mnp_planewaveretlayer_potential(mnp_greenret_t g, const mnp_layer_options* op, double enei,
                                    mnp_complex eps_up, mnp_complex eps_down, int64_t npol,
                                    const double* pol, const double* dir, void* phi2, void* phi2p,
                                    void* a2, void* a2p);
/* Far fields with the interface (Simulation/retarded/@spectrumretlayer/
 * farfield.m:20-104): directions dir device ndir x 3; e, h device out
 * ndir x 3 x npol (npol <= 8). */
int mnp_farfield_layer(mnp_greenret_t g, const mnp_layer_options* op, double enei, mnp_complex eps_up,
                       mnp_complex eps_down, int64_t ndir, const void* dir, int64_t npol,
                       const void* sig2, const void* h2, void* e, void* h);

/* Dipoles above the interface (Simulation/retarded/@dipoleretlayer) at pts
 * (host npt x 3, above z0, in the upper medium) with moments dip (host
 * npt x 3 x ndip), columns ipt + npt idip; tab must cover the face-dipole
 * and dipole-dipole pairs (tabspace with the dipoles as sources).
 * Reflected part of the potentials at the faces, ADDED to phi, phip
 * (device n x npol) and a, ap (device n x 3 x npol), e.g. after
 * mnp_dipole_potential with eps_up: replaces @dipoleretlayer/potential.m:
 * 29-96 with private/greenderiv.m:15-99 (forward differences eta = 1e-6 of
 * the interpolated table; the hh term of ap(:, 3) takes the moment's y
 * component, as potential.m:82 does). */
int mnp_dipolelayer_potential(mnp_greenret_t g, mnp_layertab_t tab, double enei, mnp_complex eps_up,
                              mnp_complex eps_down, int64_t npt, const double* pts, int64_t ndip, const double* dip,
                              void* phi, void* phip, void* a, void* ap);
/* Reflected field of the dipoles at the points obs (host nobs x 3): e, h
 * device out nobs x 3 x npol (@dipoleretlayer/field.m:45-128 with 'refl';
 * obs equal to pts moves the dipoles by rmin along x, greenderiv.m:19-21). */
int mnp_dipolelayer_field(mnp_layertab_t tab, double enei, mnp_complex eps_up, mnp_complex eps_down, int64_t nobs,
                          const double* obs, int64_t npt, const double* pts, int64_t ndip, const double* dip, void* e,
                          void* h);
/* Far field of the dipoles in the directions dir (device ndir x 3): e, h
 * device out ndir x 3 x npol, upward direct plus reflected, downward
 * transmitted (@dipoleretlayer/farfield.m:21-177). */
int mnp_dipolelayer_farfield(const mnp_layer_options* op, double enei, mnp_complex eps_up, mnp_complex eps_down,
                             int64_t ndir, const void* dir, int64_t npt, const double* pts, int64_t ndip,
                             const double* dip, void* e, void* h);

/* ---------------------------------------------- quasistatic dipoles
 * phi' = -nvec . E at the faces (device out n x npt ndip, column ipt + npt
 * idip) of dipoles dip (host npt x 3 x ndip) at pts (host npt x 3), E = w
 * (3 u (u . p) - p) / r^3 with w[ipt] = 1 / eps of the point's medium
 * (Simulation/static/@dipolestat/field.m:50-78, potential.m:17).  With
 * wimage set (host npt), the image dipole at the point mirrored at z = z0
 * with p_z negated adds wimage[ipt] = q1 / eps_image times its field
 * (@dipolestatlayer/init.m:35-39, field.m:23-28). */
int mnp_dipolestat_potential(mnp_greenret_t g, int64_t npt, const double* pts, int64_t ndip,
                             const double* dip, const mnp_complex* w, double z0,
                             const mnp_complex* wimage, void* phip);
/* Quasistatic field E = -Gp sig at point ipt = q % npt from column q of sig
 * (device n x ncol), host out ncol x 3: @compgreenstat/field.m:14-25 for
 * points off the surface, with the refinement of greenstat/private/init.m
 * for point-face pairs within the cutoff (the retarded point field at
 * k = 0).  With g the Green function of [particle; image] and sig padded
 * by n zero rows first, this is the reflected part of @compgreenstatlayer's
 * field (before the factor f2). */
int mnp_stat_field_own(mnp_greenret_t g, int64_t npt, const double* pts, int64_t ncol, const void* sig,
                       mnp_complex* e);

/* ------------------------------------ quasistatic solver, mirror symmetry
 * Replaces bemsolver(p, op) with op.sim = 'stat' and p a comparticlemirror
 * -> bemstatmirror (BEM/@bemstatmirror).  g is the Green function of the
 * full particle (nsym n faces, copy j in faces j n ..; closed-surface
 * correction as for the full particle), symtable the nsym x nsym table of
 * +-1 (row-major, Particles/@comparticlemirror/private/init.m:29-37).  F_s =
 * sum_j symtable[s][j] F(p, copy j) is assembled once
 * (compgreenstatmirror/eval.m:18-40); initmat factorises Lambda I + F_s for
 * every s (bemstatmirror/subsref.m:29-37); solve returns sig = -(Lambda I +
 * F_sym)^{-1} phip for the symmetry index sym of the potential's
 * z-component value (mldivide.m:14-22).  phip, sig: device complex128
 * n x npol.  Solvers and refinement as mnp_bemstat_*. */
typedef struct mnp_bemstatmirror_s* mnp_bemstatmirror_t;
int mnp_bemstatmirror_create(mnp_greenret_t g, int solver, int nsym, const int* symtable,
                             mnp_bemstatmirror_t* out);
int mnp_bemstatmirror_destroy(mnp_bemstatmirror_t bem);
int mnp_bemstatmirror_initmat(mnp_bemstatmirror_t bem, double enei, mnp_complex eps1, mnp_complex eps2);
int mnp_bemstatmirror_solve(mnp_bemstatmirror_t bem, int sym, int64_t npol, const void* phip, void* sig);
int mnp_bemstatmirror_stats(mnp_bemstatmirror_t bem, int sym, int* sweeps, double* relres);
int mnp_bemstatmirror_set_refinement(mnp_bemstatmirror_t bem, double tol, int maxit);


/* ------------------------------------------------ dipoles and point fields
 * Dipoles at points pts (host npt x 3, column-major) with moments dip (host
 * npt x 3 x ndip); column q = ipt + npt * idip of the outputs, as the
 * reference's (n, npt, ndip) arrays.  eps_medium is the permittivity of the
 * medium holding the points.
 *
 * Replaces @dipoleret/potential.m:64-122 for the side of the particle that
 * faces the dipoles: phi, phip device out n x npol, a, ap device out
 * n x 3 x npol (complex128); pass them as phi1.. or phi2.. to the solver. */
int mnp_dipole_potential(mnp_greenret_t g, double enei, mnp_complex eps_medium, int64_t npt,
                         const double* pts, int64_t ndip, const double* dip, void* phi, void* phip,
                         void* a, void* ap);
/* Fields at points from the surface charges sig (device n x ncol) and
 * currents h (device n x 3 x ncol) of the particle side facing the points:
 * e = i k0 G h - Gp sig, hfield = Gp x h, device out npt x 3 x ncol
 * (@compgreenret/field.m:14-40 with compgreenret(pt, p)).  Point-face pairs
 * within the refinement cutoff of the greenret options (refinematrix.m,
 * the point taking the face radius) use the rule-18 Taylor form
 * (greenret/private/init.m:96-177).  Points must not coincide with
 * quadrature points of a face.  h may be NULL (no currents); with
 * eps_medium = 0 (k = 0) e is the quasistatic field -Gp sig
 * (@compgreenstat/field.m). */
int mnp_field_points(mnp_greenret_t g, double enei, mnp_complex eps_medium, int64_t npt,
                     const double* pts, int64_t ncol, const void* sig, const void* h, void* e,
                     void* hfield);
/* Potentials at points from the surface charges sig (device n x ncol) and
 * currents h (device n x 3 x ncol, or NULL) of the particle side facing the
 * points: phi = G sig, device out npt x ncol, and a = G h, device out
 * npt x 3 x ncol (zero when h is NULL), with the wavenumber of eps_medium.
 * Replaces @compgreenret/potential.m:30-40 with compgreenret(pt, p) (its
 * phi and a; points carry no normals, so no surface derivatives); the same
 * near-field refinement as mnp_field_points. */
int mnp_potential_points(mnp_greenret_t g, double enei, mnp_complex eps_medium, int64_t npt,
                         const double* pts, int64_t ncol, const void* sig, const void* h, void* phi,
                         void* a);
/* Total and radiative decay rates in units of the free-space rate and the
 * free-space rate nb 4/3 k0^3 (host out npt x ndip) for dipoles outside the
 * particle, from sig2/h2 of the solve (device n x npol, n x 3 x npol):
 * @dipoleret/decayrate.m:13-62 with the radiated power of the surface far
 * field plus the dipole far field over the pinfty surface (device
 * npinf x 3 normals, npinf areas; farfield.m:40-55, scattering.m).  The
 * medium must be lossless (the reference warns otherwise); real parts are
 * returned.  rad0 may be NULL. */
int mnp_dipole_decayrate(mnp_greenret_t g, double enei, mnp_complex eps_medium, int64_t npt,
                         const double* pts, int64_t ndip, const double* dip, int64_t npinf,
                         const void* pinfty_nvec, const void* pinfty_area, const void* sig2,
                         const void* h2, double* tot, double* rad, double* rad0);

#ifdef __cplusplus
}
#endif

#endif /* MNPBEM_HIP_H */
