/*
 * tritok_bem.h -- C-ABI of the MI355X (gfx950) boundary-element library
 * libtritok_amd.so.
 *
 * Drop-in boundary for the hot path of the Tritok velocity-vorticity BEM
 * solver.  A Fortran host binds every function below through ISO_C_BINDING
 * (fortran/tritok_amd_bind.f90); arrays are passed exactly as the reference
 * holds them: column-major, index arrays 1-based, INTEGER = int32,
 * REAL(8) = double.  No function takes or returns a framework type.
 *
 * Status codes: 0 = success, negative = error; tt_last_error() returns the
 * message of the last error on the calling thread.
 *
 * Each function names the reference routine it replaces (file:line in the
 * reference's src/).
 */
#ifndef TRITOK_BEM_H
#define TRITOK_BEM_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Plain views of the arrays the reference keeps in TYPE(meshType)
 * (inc_types.f90:77-212).  The library never frees them. */
typedef struct tt_mesh_view {
  int32_t nnodes, nbnodes, nbelem, nicell, nq;
  const double *x;    /* x(nnodes,3)                                   */
  const int32_t *ibc; /* ibc(nbelem,9)  boundary elements              */
  const int32_t *idc; /* idc(nicell,27) cells                          */
  const int32_t *ibf; /* ibf(nicell,6,4) flux nodes of cell faces      */
  const double *xq;   /* xq(nq,3)       flux node coordinates          */
  const int32_t *flpm;/* flpm(nicell,6) flux sign of a cell face       */
  const int32_t *gbn; /* gbn(nbnodes)   boundary node -> node          */
  const int32_t *lbn; /* lbn(nnodes)    node -> boundary node, 0 inside*/
  const double *nx, *ny, *nz; /* source point normals (nbnodes)        */
} tt_mesh_view;

/* TYPE(GaussType) (inc_types.f90:367-372) as GaussPosWeights fills it
 * (lib_int.f:2673-3173) plus GAUS kmBr kmBs kmDr kmDs and IIDI of tri.inp
 * (lib_IO.f:322-325): rule indices 1..9 (N = 2,4,6,8,16,32,40,48,64), the
 * number of sub-cells per direction for singular cells, and the table
 * gi(220), ome(220) with 1-based ranges ng1(9)..ng2(9) per rule.  A NULL
 * table selects Gauss-Legendre values computed by the library; a Fortran host
 * passes gp%gi, gp%ome, gp%ng1, gp%ng2 to reproduce its own table. */
typedef struct tt_quad {
  int32_t kmBr, kmBs, kmDr, kmDs, iiDiv;
  const double *gi, *ome;
  const int32_t *ng1, *ng2;
} tt_quad;

const char *tt_last_error(void);
/* library version and the device it runs on (name, CU count, HBM bytes) */
int tt_device_info(char *name, int32_t name_len, int32_t *num_cu, int64_t *hbm_bytes);

/* ------------------------------------------------------------------------ */
/* Integral assembly.                                                        */

/* Replaces FMATkmdcSteady (lib_fmat.f:3-446) as called by corMin_kmdcSteady
 * (lib_int.f:1254-1340): all per-cell macro matrices for every cell, with the
 * free-term, Ht row-sum and linear-field corrections.  Outputs are host arrays,
 * column-major: H,B,Abdx,Abdy,Abdz,Dx,Dy,Dz (nicell*51, 27), G (nicell*51, 24),
 * Htx,Hty,Htz (nicell*27, 27); ont[11] as written to tri.macro.int.
 * copy = COPY keyword (replicate cell 1). */
int tt_macro_integrals(const tt_mesh_view *m, const tt_quad *q, int32_t copy, double *H, double *G,
                       double *B, double *Abdx, double *Abdy, double *Abdz, double *Htx,
                       double *Hty, double *Htz, double *Dx, double *Dy, double *Dz, double ont[11]);

/* Replaces FMATmodHelm (lib_mHelm.f:2-483) as called by corMin_modHelm
 * (lib_mHelm.f:1400-1502): the same matrices as tt_macro_integrals for the
 * modified-Helmholtz equation lap u - mu^2 u = f, kernel exp(-mu r)/(4 pi r)
 * (INTEDc27modHelm :1637-2033, INTEBc9modHelm :2034-2348).  Each source
 * point has mu = sqrt(beta / diff): diff (nnodes) = mHdiff for node sources,
 * qdiff (nq) = mHqDiff for flux sources; beta = inp%beta.  The free term is
 * c = -sum H + mu^2 sum B and the linear-field corrections use H - mu^2 B.
 * *mumax (may be NULL) = the largest mu, as written to tri.modHelm.macro.int.
 * Errors: diff or qdiff <= 0 (with beta > 0), beta < 0. */
int tt_macro_integrals_mh(const tt_mesh_view *m, const tt_quad *q, int32_t copy, const double *diff,
                          const double *qdiff, double beta, double *H, double *G, double *B, double *Abdx,
                          double *Abdy, double *Abdz, double *Htx, double *Hty, double *Htz, double *Dx,
                          double *Dy, double *Dz, double ont[11], double *mumax);

/* Measurement hook (no reference counterpart): device time (ms) of the
 * calling thread's last macro assembly kernel (tt_macro_integrals,
 * tt_macro_integrals_mh, tt_case_create) and its fundamental-solution
 * evaluations over cells (domain) and faces (boundary). */
int tt_last_integral_stats(double *ms, int64_t *domain_points, int64_t *boundary_points);

/* Replaces FMATsdkm (lib_sdkm.f:1286-1530) as called by sdKm_corInt
 * (lib_sdkm.f:2377-2449) for boundary rows row0..row1-1 (0-based; the
 * reference's zac:kon).  Host outputs with leading dimension row1-row0:
 * H,Htx,Hty,Htz (rows, nbnodes), Dx,Dy,Dz (rows, nnodes); ont[10] as written
 * to tri.sdkm.int.N. */
int tt_sd_integrals(const tt_mesh_view *m, const tt_quad *q, int32_t row0, int32_t row1, double *H,
                    double *Htx, double *Hty, double *Htz, double *Dx, double *Dy, double *Dz,
                    double ont[10]);

/* ------------------------------------------------------------------------ */
/* Sparse least squares (the reference's CSR, 1-based).                      */

typedef struct tt_lsqr tt_lsqr; /* device-resident scaled system */

/* Replaces prelsqr2 (lib_Solve_v11.f:4501-4535): uploads A (m x n, CSR
 * ia(m+1), ja(nnz), a(nnz)), computes q_j = 1/||A(:,j)|| and scales the device
 * copy in place.  q (n) is returned to the host when non-NULL, and a is scaled
 * in place on the host too when scale_host != 0 (the reference's semantics). */
int tt_lsqr_setup(int32_t m, int32_t n, int32_t nnz, const int32_t *ia, const int32_t *ja,
                  double *a, double *q, int32_t scale_host, tt_lsqr **out);
/* Replaces slvlsqr2 + LSQR (lib_Solve_v11.f:4401-4500, :3128-3887): hot start
 * from x, atol = eps4, btol = eps4*||b||/||b - A x0||, conlim = 1e8, damp = 0.
 * Returns istop (0..5) in *istop and the iteration count in *nit. */
int tt_lsqr_solve(tt_lsqr *s, const double *b, double *x, int32_t maxit, float eps4, int32_t *nit,
                  int32_t *istop);
/* No counterpart in the reference: tt_lsqr_solve for callers that cannot
 * allocate device memory (code under HIP stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  b_dev (m) and x_dev (n) are device
 * pointers on the handle's device; x_dev holds the hot start on entry and the
 * solution on exit.  The work vectors (u, v, w, the correction, partial sums,
 * LSQR state) live in a library-owned scratch array of 32 MiB per device,
 * held until the process ends: systems with m + 3n above about 4.19 million
 * are rejected with an error.  Every call on one device shares that scratch,
 * so calls that overlap on one GPU race.  Same kernels and results as
 * tt_lsqr_solve; nothing is allocated or freed. */
int tt_lsqr_solve_noalloc(tt_lsqr *s, const double *b_dev, double *x_dev, int32_t maxit, float eps4, int32_t *nit,
                          int32_t *istop);
void tt_lsqr_free(tt_lsqr *s);

/* Many independent systems in one persistent launch: prelsqr2 for each at
 * setup, slvlsqr2 for all of them per solve call (no reference counterpart
 * for the batching: the reference solves its systems one at a time).  System
 * k has m[k] rows, n[k] unknowns and nnz[k] entries; ia holds the systems'
 * 1-based row pointers one after another (m[k]+1 each), ja/a their entries,
 * q receives the column scalings (sum of n), and with scale_host the scaled
 * values are written back to a.  A solve takes b (sum of m) and the hot start
 * x (sum of n, overwritten by the solutions), the reference's maxit and REAL(4)
 * tolerance for every system, and returns each system's iteration count and
 * istop; *ms (may be NULL) = device time of the launch. */
typedef struct tt_lsqr_batch tt_lsqr_batch;
int tt_lsqr_batch_setup(int32_t nsys, const int32_t *m, const int32_t *n, const int64_t *nnz,
                        const int32_t *ia, const int32_t *ja, double *a, double *q, int32_t scale_host,
                        tt_lsqr_batch **out);
int tt_lsqr_batch_solve(tt_lsqr_batch *h, const double *b, double *x, int32_t maxit, float eps4, int32_t *nit,
                        int32_t *istop, double *ms);
/* Gang width of every system (no reference counterpart; a tuning and test
 * hook): 0 = automatic (a system's share of the batch's work in units of one
 * workgroup's fair share, 1..64), else that width for all.  *max_width (may
 * be NULL) = the widest gang of the new schedule.  Results depend on the
 * width only through the order of the parallel sums. */
int tt_lsqr_batch_set_width(tt_lsqr_batch *h, int32_t width, int32_t *max_width);
/* Values the batched LSQR products read (no reference counterpart; the
 * reference computes in FP64): TT_PREC_FP64 (default) or TT_PREC_FP32, FP32
 * copies of the scaled values (made on first use) multiplied into FP64
 * vectors with FP64 sums.  The hot-start residual b - A x0, the LSQR scalars,
 * norms and stopping tests stay FP64, so a solve differs from the FP64 one by
 * the values' rounding (about 6e-8 relative per entry) times the system's
 * conditioning. */
int tt_lsqr_batch_set_precision(tt_lsqr_batch *h, int32_t prec);
/* Autotuner of the gang widths (no reference counterpart): times one full
 * solve of b from the hot start x (unchanged) with the automatic widths
 * scaled by 0.5, 1, 2 and 4 (ms[4], may be NULL) and keeps the fastest
 * schedule; *best_scale (may be NULL) = its scale.  Results depend on the
 * widths only through the order of the parallel sums. */
int tt_lsqr_batch_tune(tt_lsqr_batch *h, const double *b, const double *x, int32_t maxit, float eps4,
                       double *best_scale, double ms[4]);
void tt_lsqr_batch_free(tt_lsqr_batch *h);

/* Many independent subdomain systems solved on the cell-block format (no
 * reference counterpart for the batching; per system prelsqr2 + slvlsqr2,
 * lib_Solve_v11.f:4401-4535).  The systems are given as for
 * tt_lsqr_batch_setup (1-based CSR, one after another), with rows in groups
 * of rows_per_cell (51: vorticity/energy systems, 27: kinematics systems)
 * whose entries lie in at most rows_per_cell distinct columns per group --
 * the structure of the reference's subdomain systems; anything else is
 * rejected.  Each system is stored as per-cell dense blocks of its scaled
 * matrix and solved with the fused cell-block LSQR; systems are solved in
 * groups of up to 32 of similar size, every iteration of a group one launch
 * of each LSQR kernel for all its systems, each system with its own scalars
 * and stop (its iterates do not depend on the group).  A solve takes b (sum
 * of m) and the hot start x (sum of n, overwritten by the solutions), the
 * reference's maxit and REAL(4) tolerance, and returns each system's
 * iteration count and istop; *ms (may be NULL) = device time of the solve. */
typedef struct tt_block_batch tt_block_batch;
int tt_block_batch_setup(int32_t nsys, const int32_t *m, const int32_t *n, const int64_t *nnz,
                         const int32_t *ia, const int32_t *ja, const double *a, int32_t rows_per_cell,
                         tt_block_batch **out);
/* Blocks the iterations read (no reference counterpart): TT_PREC_FP64
 * (default), TT_PREC_FP32 or TT_PREC_FP16 copies of the scaled blocks with
 * FP32 products; the hot-start residual, vectors, scalars and stopping tests
 * stay FP64. */
int tt_block_batch_set_precision(tt_block_batch *h, int32_t prec);
int tt_block_batch_solve(tt_block_batch *h, const double *b, double *x, int32_t maxit, float eps4, int32_t *nit,
                         int32_t *istop, double *ms);
void tt_block_batch_free(tt_block_batch *h);

/* Shared blocks in batched solves (no reference counterpart; a tuning and
 * test hook): when systems of one batched solve (tt_block_batch_solve,
 * tt_case_picard_step_batch, the batched server) have FP32 or FP16 blocks,
 * column maps and gather tables that are equal bit for bit -- the subdomain
 * systems of cases on one mesh whose matrices differ only below the blocks'
 * precision, compared on the device once per pair -- they read one copy of
 * the blocks, and a cell-kernel launch whose systems all do interleaves their
 * workgroups so that each block is fetched once into an XCD's L2 for all of
 * them.  Each system's arithmetic is unchanged, so are its results.  on = 0
 * turns this off for the process (default on).  The statistics count the
 * interleaved launches since the library was loaded and the systems they
 * carried. */
int tt_lsqr_set_block_sharing(int32_t on);
int tt_lsqr_block_sharing_stats(int64_t *launches, int64_t *systems);

/* Replaces CRSxV (lib_util.f:1351-1372): y = A x on the device for a CSR
 * matrix given by host arrays (m rows, n columns). */
int tt_csr_mv(int32_t m, int32_t n, int32_t nnz, const int32_t *ia, const int32_t *ja,
              const double *a, const double *x, double *y);

/* ------------------------------------------------------------------------ */
/* Compressed matrices: cross approximation (lib_ACA.f), stretched-Haar      */
/* wavelets (lib_WT.f) and block matrices of compressed parts (lib_ctree.f). */
/* Matrices are host arrays, column-major; results stay on the device behind */
/* a handle; vectors are host arrays.                                       */

typedef struct tt_aca tt_aca;
/* Replaces AcaDo (lib_fm.f:2003-2028): type 1 aca_CrossAprox (lib_ACA.f:126-181,
 * rank INT(eps*min(m,n))), type 2 acaj_CrossAprox / ACA_AB (:182-225, :372-510,
 * stop when the EPS_ACA_NORM estimate <= eps), type 3 aca_RankCrossAprox
 * (:24-125, stop when getRMSnorm of A vec against a (b vec) < eps; vec
 * required -- the reference draws it with RANDOM_NUMBER).  With vec, *norm is
 * that getRMSnorm for the final rank (NaN without vec).  *rank = acaMat%rank. */
int tt_aca_compress(int32_t m, int32_t n, const double *a, int32_t type, double eps, const double *vec,
                    tt_aca **out, int32_t *rank, double *norm);
/* acaMat%a (m x rank) and acaMat%b (rank x n), column-major. */
int tt_aca_factors(const tt_aca *h, double *a, double *b);
/* aca_abxv (lib_ACA.f:226-249): y = a (b x). */
int tt_aca_mv(tt_aca *h, const double *x, double *y);
/* aca_abxv on host factors. */
int tt_aca_abxv(int32_t m, int32_t n, int32_t rank, const double *a, const double *b, const double *x,
                double *y);
void tt_aca_free(tt_aca *h);

typedef struct tt_wt tt_wt;
/* Replaces Che_WcomMatrix (lib_WT.f:148-210): W A W^T (Che_WaWT :428-579),
 * entries below kappa * mean|.| zeroed (Che_AbsPovp :719-740, Che_Zanemari
 * :741-777), CSR form (Che_FormCRS :355-401).  *nnz = stored entries. */
int tt_wt_compress(int32_t m, int32_t n, const double *a, double kappa, tt_wt **out, int32_t *nnz);
/* crsMatrixType zvr(m+1), c(nnz), v(nnz) (1-based); returns the reference's
 * count m*n - (zeroed entries). */
int tt_wt_csr(const tt_wt *h, int32_t *zvr, int32_t *c, double *v);
/* Che_WTAWx (lib_WT.f:297-329): y = W^T (C (W x)). */
int tt_wt_mv(tt_wt *h, const double *x, double *y);
/* Che_WTAWx on a host CSR. */
int tt_wt_wtawx(int32_t nrow, int32_t ncol, int32_t nnz, const int32_t *zvr, const int32_t *c, const double *v,
                const double *x, double *y);
void tt_wt_free(tt_wt *h);

typedef struct tt_hmat tt_hmat;
/* Replaces cTree_FillMatrices (lib_ctree.f:2-96) for a partition given as the
 * reference's HmatrixType parts (inc_types.f90:494-520): part p has rows
 * rnodes[rptr[p]..rptr[p+1]) and columns cnodes[cptr[p]..cptr[p+1]) (1-based,
 * rptr[0] = cptr[0] = 0); admissible parts (admiss[p] = 1) are compressed
 * with ctype 2 (wavelet, wt_kappa) or 3 (cross approximation of aca_type with
 * aca_eps; type 3 draws its test vector from a fixed hash), the others kept
 * dense.  a is nrow x ncol with leading dimension lda. */
int tt_hmat_create_parts(int32_t nrow, int32_t ncol, const double *a, int64_t lda, int32_t nparts,
                         const int32_t *rptr, const int32_t *rnodes, const int32_t *cptr, const int32_t *cnodes,
                         const int32_t *admiss, int32_t ctype, int32_t aca_type, double aca_eps, double wt_kappa,
                         tt_hmat **out);
/* The same with the library's own partition (the reference reads its cluster
 * trees from a .ctree file that no input here provides): row and column
 * points xr (nrow x 3), xc (ncol x 3) are split by geometric bisection into
 * clusters of at most `leaf` points, and a block is admissible when
 * min(diam_r, diam_c) < eta * dist (cTree_Admis, lib_ctree.f:575-605, with
 * bounding-box diameters and distances). */
int tt_hmat_create(int32_t nrow, int32_t ncol, const double *a, int64_t lda, const double *xr, const double *xc,
                   int32_t leaf, double eta, int32_t ctype, int32_t aca_type, double aca_eps, double wt_kappa,
                   tt_hmat **out);
/* The partition tt_hmat_create uses, computed on the host (no device call):
 * call with NULL arrays for the sizes, then with arrays of those sizes
 * (rptr/cptr nparts+1, node lists, admiss nparts) for the parts, which
 * tt_hmat_create_parts accepts. */
int tt_hmat_partition(int32_t nrow, const double *xr, int32_t ncol, const double *xc, int32_t leaf, double eta,
                      int32_t *nparts, int64_t *rnodes_len, int64_t *cnodes_len, int32_t *rptr, int32_t *rnodes,
                      int32_t *cptr, int32_t *cnodes, int32_t *admiss);
/* Sizes: parts, lengths of the node lists, doubles stored, largest rank. */
int tt_hmat_info(const tt_hmat *h, int32_t *nparts, int64_t *rnodes_len, int64_t *cnodes_len, int64_t *stored,
                 int32_t *max_rank);
/* The partition: rptr/cptr (nparts+1), node lists, ctype (1 dense, 2, 3), rank. */
int tt_hmat_parts(const tt_hmat *h, int32_t *rptr, int32_t *rnodes, int32_t *cptr, int32_t *cnodes,
                  int32_t *ctype, int32_t *rank);
/* Stored data of one part: dense nr x nc; ACA a (nr x rank) then b (rank x nc);
 * wavelet CSR values into d and zvr(nr+1) then c into iv.  Returns the count
 * of doubles. */
int tt_hmat_part_data(const tt_hmat *h, int32_t part, double *d, int32_t *iv);
/* cTree_MxV (lib_ctree.f:99-164) for nrhs = 1..4 columns: y(nrow x nrhs) =
 * H x(ncol x nrhs), parts summed in part order. */
int tt_hmat_mv(tt_hmat *h, int32_t nrhs, const double *x, double *y);
void tt_hmat_free(tt_hmat *h);

/* ------------------------------------------------------------------------ */
/* Dense single-domain kinematics.                                          */

/* Replaces dgefa (lib_Solve_v11.f:787-904) as used by FormPREfm pret=2:
 * partial-pivot LU of a(n,n) in place (host array) in LINPACK's storage
 * (negated multipliers, each column's multipliers in the row order of its own
 * elimination step), pivots ipvt(n) as doubles.  Returns 0, or k > 0 when
 * U(k,k) = 0 (dgefa's info), or < 0 on error. */
int tt_dense_lu(int32_t n, double *a, double *ipvt);
/* Replaces dgesl job=0 (lib_Solve_v11.f:905-1037): solves in place with the
 * factor of tt_dense_lu (or of dgefa). */
int tt_dense_lu_solve(int32_t n, const double *a, const double *ipvt, double *b);

typedef struct tt_sd tt_sd; /* device-resident single-domain kinematics */
/* Replaces sdKm_corInt (lib_sdkm.f:2377-2449) for all boundary rows,
 * sdKm_sysm (lib_sdkm.f:1171-1210) and FormPREfm pret=2 (lib_Solve_v11.f:
 * 403-454) for the three equations: the seven dense matrices, the three
 * system matrices and their factorisations stay on the device.  kode(nbnodes,3)
 * as ReadBic sets it; the mesh view must carry the source point normals. */
int tt_sd_create(const tt_mesh_view *m, const tt_quad *q, const int32_t *kode, tt_sd **out);
/* Replaces sdkm_solve (lib_sdkm.f:23-121) for eqn = 1, 2, 3 in turn:
 * boundary vorticities from vel(nnodes,3) and vort(nnodes,3), relaxed with
 * urBw on rows with kode 0/3 and written back into vort. */
int tt_sd_solve(tt_sd *h, const double *vel, double *vort, const double urBw[3]);
/* Mixed-precision solve of one kinematics equation (eqn 1..3; no reference
 * counterpart -- the reference solves with dgesl in FP64, lib_Solve_v11.f:
 * 905-1037): A_eqn X = B for nrhs = 1..4 right-hand sides (B, X host arrays
 * nbnodes x nrhs, column-major) by FP64 iterative refinement, X <- X + M (B -
 * A X) until ||B - A X||_F <= tol ||B||_F or maxit corrections.  M is the
 * inverse held in prec: TT_PREC_FP64 (FP64 products), TT_PREC_FP32, _BF16 or
 * _FP16 (a copy with a power-of-two scale per row, applied on the matrix
 * cores with FP32 accumulation to the residual scaled per column by a power
 * of two; made on first use and kept).  *iters = corrections, *relres = the
 * final relative residual, ms (may be NULL) = {device ms of the solve, mean
 * device ms of one correction product}. */
int tt_sd_refine(tt_sd *h, int32_t eqn, int32_t prec, int32_t nrhs, const double *B, double *X, double tol,
                 int32_t maxit, int32_t *iters, double *relres, double ms[2]);
/* The AcaCompSmat step of the compressed kinematics (lib_fm.f:2029-2086,
 * drivers AcaDo :2003-2028, cTree_FillMatrices lib_ctree.f:2-96) on a matrix
 * that stays on the device: which = 0..6 selects H, Htx, Hty, Htz (nbnodes x
 * nbnodes) or Dx, Dy, Dz (nbnodes x nnodes); the rows and columns are split
 * by the library's geometric bisection of the boundary nodes (and, for D, of
 * all nodes) into clusters of at most `leaf` points, admissible blocks (as
 * tt_hmat_create) are cross-approximated with aca_type and aca_eps, the rest
 * kept dense.  The result is a tt_hmat (tt_hmat_mv, tt_hmat_info). */
int tt_sd_hmat(tt_sd *h, int32_t which, int32_t leaf, double eta, int32_t aca_type, double aca_eps,
               tt_hmat **out);
/* y = M x for one of the seven device-resident matrices (which as in
 * tt_sd_hmat; no reference counterpart: the dense product the compressed
 * forms are checked against). */
int tt_sd_matvec(tt_sd *h, int32_t which, const double *x, double *y);
void tt_sd_free(tt_sd *h);

/* ------------------------------------------------------------------------ */
/* Single-domain modified-Helmholtz operator (EQNS imH 3-7, main.f90:556-777
 * and the solves of :1345-1395): the diffusion-convection equation with
 * variable diffusivity solved with the modified-Helmholtz fundamental
 * solution over the whole domain.  Rows: U rows are the nodes (source x(i),
 * mu = sqrt(beta/diff(i))); Q rows the boundary flux points, element-major
 * (row 4e + f, source xq(ibcf(4e + f)), mu = sqrt(beta/qdiff(ibcf))).  The
 * ten matrices of each set (k = 0..9: H, G, B, Bp, Abdx, Abdy, Abdz, Dx, Dy,
 * Dz) stay on the device, column-major with the row count as leading
 * dimension; columns: boundary nodes (H, Abd), boundary flux points (G),
 * nodes (B, Bp, D). */
typedef struct tt_mhsd tt_mhsd;

/* corMin_SDkin_mHfs (lib_fm.f:1646-1745) without the file: FMATsdmHfs
 * (lib_fm.f:1183-1645, integrals with free terms and linear-field
 * corrections) and MakeBprime (lib_fm.f:800-858).  ibcf: 4 nbelem flux-node
 * ids (tt_host_mh_codes); diff(nnodes), qdiff(nq): diffusivities; beta > 0.
 * ont[5] = ontH, ontG, ontDx, ontDy, ontDz; *mumax (may be NULL). */
int tt_mhsd_create(const tt_mesh_view *m, const tt_quad *q, const int32_t *ibcf, const double *diff,
                   const double *qdiff, double beta, tt_mhsd **out, double ont[5], double *mumax);
/* The same operator from matrices already computed (the reference's cache
 * branch of corMin_SDkin_mHfs, ReadmHfullIntDisk lib_fm.f:1746-1820, e.g.
 * from tt_read_mhsd_int): U[k] nnodes rows, Q[k] 4 nbelem rows, column-major
 * in the order H, G, B, Bp, Abdx, Abdy, Abdz, Dx, Dy, Dz. */
int tt_mhsd_create_from(const tt_mesh_view *m, const int32_t *ibcf, const double *diff,
                        const double *qdiff, double beta, const double *const U[10],
                        const double *const Q[10], tt_mhsd **out);
/* Copies matrix k (0..9) of row set `set` (0: U, 1: Q) into dst. */
int tt_mhsd_matrix(tt_mhsd *h, int32_t set, int32_t k, double *dst);
/* fmmHSysM (lib_fm.f:1061-1181) and FormPREfm(pret 2) (lib_Solve_v11.f:
 * 403-451): the system matrix (nunk x nunk) and right-hand-side matrix
 * (nunk x nknown) of the codes kode(nnodes), qkode(4 nbelem) as
 * tt_host_mh_codes returns them, factorised on the device (LU and its
 * inverse).  Errors: codes that do not number 1..nunk / 1..nknown once each,
 * a singular system. */
int tt_mhsd_system(tt_mhsd *h, const int32_t *kode, const int32_t *qkode, int32_t nunk, int32_t nknown);
/* The right-hand-side matrices (B, Bp, Dx, Dy, Dz, Abdx, Abdy, Abdz of both
 * sets) for the solves (main.f90:609-777): mode 0 dense (imH 3, 4); 1
 * wavelet, WaveletCompSmat with kappa (imH 5); 2 cross approximation,
 * AcaCompSmat with aca_type and aca_eps (imH 6); 3 block matrices with the
 * library's own geometric partition (leaf, eta) and blocks of ct_type 1
 * dense, 2 wavelet, 3 cross approximation (imH 7; the reference reads its
 * cluster tree from a file instead).  *ratio (may be NULL) = stored /
 * full entries. */
int tt_mhsd_compress(tt_mhsd *h, int32_t mode, int32_t aca_type, double aca_eps, double wt_kappa,
                     int32_t ct_type, double ct_eta, int32_t leaf, double *ratio);
/* One solve, fmSolveModHelmDC / DCmm / DCwawt / DCaca / DChmat (lib_fm.f:
 * 2-1059) by the current mode: r = RHS b plus the equation rows of B (-rhsv)
 * + Bp u + D_c u (v_c + grad diff_c) - Abd_c u v_c, x = A^-1 r (SolvEQNfm
 * with the LU), then u(i) = (1 - urmh) u(i) + urmh x on the unknown nodes and
 * likewise q on the unknown flux points.  u(nnodes), qu(4 nbelem) in place;
 * velocity, dgrad (nnodes x 3) and rhsv(nnodes) host arrays. */
int tt_mhsd_solve(tt_mhsd *h, double *u, double *qu, const double *velocity, const double *rhsv,
                  const double *dgrad, double urmh);
/* Device times in ms of the last integrals, system, compression and solve,
 * the compression ratio and the number of unknowns (any may be NULL). */
int tt_mhsd_stats(const tt_mhsd *h, double ms[4], double *ratio, int32_t *nunk);
void tt_mhsd_free(tt_mhsd *h);

/* ------------------------------------------------------------------------ */
/* Whole-case driver: the reference's main.f90 setup (:237-1084) and the
 * steady Picard loop (:1200-1589) with matrices resident in HBM.            */

typedef struct tt_case tt_case;

/* Case inputs as the reference's host holds them after ReadBic
 * (lib_bic.f:50-502), DetKMunk3 (lib_km.f:676-727) and FillKobc
 * (lib_mesh.f:230-276). */
typedef struct tt_case_view {
  tt_mesh_view mesh;
  tt_quad quad;
  int32_t copy;
  const int32_t *kode;   /* kode(nbnodes,3)                      */
  const int32_t *kobc;   /* kobc(nbnodes) (may be NULL if no dv/dn=0) */
  const int32_t *wkode;  /* wkode(nnodes,3), wqkode(nq,3)        */
  const int32_t *wqkode;
  const int32_t *Tkode;  /* Tkode(nnodes), Tqkode(nq)            */
  const int32_t *Tqkode;
  int32_t nunk[3], nbk[3], Tnunk, Tnb;
  const int32_t *kmu;    /* kmp%u(nnodes) */
  int32_t nsmcol, nrhscol;
  int32_t iDT;
  double ReN, RaN, PrN, gx, gy, gz, tstep;
  double urBw[3], urDv[3], urDw[3], urDT;
  double eps;
  int32_t nnlit, lsqs_maxit, iDLSE;
  float lsqs_eps, dlse_r, dlse_epsth;
  int32_t iBw;           /* EQNS iBw: 1 single-domain kinematics, 0 none (main.f90:410,1207) */
} tt_case_view;

/* precision of the iteration matrices (see tt_case_set_precision) */
enum { TT_PREC_FP64 = 0, TT_PREC_FP32 = 1, TT_PREC_BF16 = 2, TT_PREC_FP16 = 3 };

/* Builds everything the Picard loop needs on the device: macro and sdkm
 * integrals, the three factorizations, all least-squares systems, and the
 * fields (initial values vel(nn,3), vort(nn,3), qvort(nq,3), temp(nn),
 * qtemp(nq)).  prec as tt_case_set_precision: with anything but TT_PREC_FP64
 * the least-squares bidiagonalisation products run on reduced-precision
 * copies of the scaled blocks (FP16 for TT_PREC_FP16, FP32 otherwise; FP32
 * arithmetic) while the initial residual of every solve (b - A x0, the hot
 * start), the LSQR scalars and norms, all vectors and the solution update
 * stay FP64, so the Picard loop refines the reduced-precision corrections
 * against the FP64 equations.  refine_tol is reserved (pass 0). */
int tt_case_create(const tt_case_view *c, const double *vel, const double *vort,
                   const double *qvort, const double *temp, const double *qtemp, int32_t prec,
                   double refine_tol, tt_case **out);
/* Restarts the iteration from new initial fields on the resident matrices
 * (the state main.f90:1145-1199 sets up before the first Picard step). */
int tt_case_reset(tt_case *c, const double *vel, const double *vort, const double *qvort,
                  const double *temp, const double *qtemp);
/* Switches the precision of a resident case.  TT_PREC_FP64: FP64 throughout.
 * TT_PREC_FP32, TT_PREC_BF16: FP32 least-squares blocks; TT_PREC_FP16: FP16
 * least-squares blocks (the scaled blocks' columns have unit norm); in all
 * three the products accumulate in FP32 while the residual anchors, scalars and
 * vectors stay FP64.  The dense kinematics products (the seven matrices of
 * sdkm_rhs and the three inverse system matrices) run on copies in that
 * format with a power-of-two scale per row, taken on the change of the
 * fields since the last FP64 anchor step; an anchor step (FP64 products
 * and solves) runs every anchor period (default 16) Picard iterations and after
 * a reset, so the loop converges to the FP64 fixed point.  Copies are made on
 * first use. */
int tt_case_set_precision(tt_case *c, int32_t prec);
/* Anchor period of the reduced-precision kinematics (>= 1; 1 = every step FP64). */
int tt_case_set_anchor_period(tt_case *c, int32_t period);
/* Grid of the least-squares cell kernel (no reference counterpart; a tuning
 * and test hook): grid > 0 uses that many blocks for every system and block
 * precision, 0 returns to the autotuner's choice.  Results do not depend on
 * it (the partial sums are per fixed group of cells). */
int tt_case_set_grid(tt_case *c, int32_t grid);
/* Threads per workgroup of the least-squares unknown kernel (no reference
 * counterpart; a tuning and test hook): 256, 512 or 1024 for every system
 * and block precision, 0 = leave it to the autotuner (which only picks it
 * when it also picks the grid).  Results do not depend on it (the partial
 * sums are per fixed group of 256 unknowns). */
int tt_case_set_unknown_threads(tt_case *c, int32_t threads);
/* How the least-squares iterations are issued (no reference counterpart; a
 * tuning and test hook): 0 (default) launches every kernel on the case's
 * stream, 1 replays chunks of 2..64 iterations from HIP graphs captured once
 * per system.  The kernels and their arguments are the same, so are the
 * results. */
int tt_case_set_graphs(tt_case *c, int32_t on);
/* Measurement hooks (no reference counterpart): time every stride-th launch
 * of the least-squares block kernel with HIP events on the case's stream
 * (stride 0 = off; resets the statistics), and read the average duration,
 * timed and total launch counts per system shape (rows = 51: vorticity and
 * energy systems, rows = 27: kinematics systems). */
int tt_case_probe(tt_case *c, int32_t stride);
int tt_case_probe_stats(tt_case *c, int32_t rows, double *avg_ms, int64_t *sampled,
                        int64_t *launches);
/* Sizes of a least-squares system (sys 0..2 vorticity x,y,z, 3 energy, 4
 * kinematics): info = {rows m, unknowns n, nonzeros, cells, rows per cell,
 * columns per cell, then the cell-kernel grid the autotuner chose for FP64,
 * FP32 and FP16 blocks (0: not tuned yet)}. */
int tt_case_system_info(tt_case *c, int32_t sys, int64_t info[9]);
/* What the autotuner picked for a system (no reference counterpart), per
 * block precision FP64, FP32, FP16 (0: not tuned yet): the cell kernel's
 * grid and the unknown kernel's threads per workgroup (256, 512 or 1024).
 * Neither choice changes a result. */
int tt_case_system_tuning(tt_case *c, int32_t sys, int32_t grid[3], int32_t threads[3]);
/* One Picard iteration (main.f90:1200-1589): err[9] = the nine relative
 * changes (tri.sta), nit[12] = solver iterations (tri.ite).  Returns 1 while
 * the loop should continue, 0 when converged or at nnlit, < 0 on error --
 * including a diverged iteration (a non-finite relative change, where the
 * reference would stop as if converged), and a vorticity or energy LSQR solve
 * that stops with istop > 3 (WarnErr level 4 in the reference).  A kinematics
 * solve that stops with istop > 3 keeps its updated velocities and the step
 * goes on, as lib_km.f:274 does. */
int tt_case_picard_step(tt_case *c, double err[9], int32_t nit[12]);
/* One Picard iteration of each of ncase cases (1..16) at once: every LSQR
 * iteration of a phase runs the cell kernels of all the cases' systems in one
 * launch and their unknown kernels in another (no reference counterpart: the
 * reference runs one case per process; for a serving loop that batches the
 * Krylov iterations of its active cases).  Each case's result is bit for bit
 * that of tt_case_picard_step on it alone.  err (ncase x 9), nit (ncase x 12)
 * and status (ncase) receive per-case results as tt_case_picard_step's:
 * status 1 = continue, 0 = converged / at nnlit / already finished,
 * -1 = failed (tt_last_error names the first failed case); a failed case does
 * not stop the others.  The cases must iterate with the same block precision;
 * the batch runs on the first unfinished case's stream, after the others'
 * streams are drained.  Returns 0, or -1 if the batch could not run. */
int tt_case_picard_step_batch(tt_case **cases, int32_t ncase, double *err, int32_t *nit, int32_t *status);
/* Runs Picard iterations until convergence; returns the iteration count. */
int tt_case_run(tt_case *c, int32_t max_iter, double *sta, int32_t *ite);
/* Copies the current fields to host arrays (any may be NULL). */
int tt_case_fields(tt_case *c, double *vel, double *vort, double *qvort, double *temp,
                   double *qtemp);
/* Device time of the last Picard step split by phase (ms): [0] boundary
 * vorticities (sdkm products and solves), [1] domain velocities (KM), [2]
 * temperature, [3] domain vorticities (VTE), [4] relative changes, [5] unused,
 * [6] whole step; and setup times (ms): [7] macro integrals, [8] sdkm
 * integrals, [9] LU factorization and inversion, [10] system build. */
int tt_case_timings(tt_case *c, double t[11]);
/* WrstF (lib_IO.f:604-685) for the case's current state: velocity,
 * vorticity, qvorticity, temp, qtemp; the previous Picard iterate as the pnls
 * fields; the initial fields of the (single) time step as the pts and ppts
 * fields; zero DC fields (not on this path). */
int tt_case_write_restart(tt_case *c, const char *path, int32_t tstep, int32_t glnnlit, double rtime);
/* HBM bytes held by the case on the device. */
int64_t tt_case_device_bytes(tt_case *c);
void tt_case_free(tt_case *c);

/* ------------------------------------------------------------------------ */
/* Host-side preprocessing (the reference's Fortran host side; these run on
 * the CPU and produce the arrays above).                                    */

typedef struct tt_host tt_host; /* mesh + macro mesh + boundary conditions */

/* ReadMeshGid3D (lib_mesh.f:279-504) incl. CheckLinearGeometry and
 * SourcePointNormals, then GenMakroMesh (lib_mesh.f:865-982). */
int tt_host_read_geo(const char *path, tt_host **out);
/* The same from arrays x(nn,3), ibc(ne,9), idc(nc,27), bewn(ne). */
int tt_host_from_arrays(int32_t nn, int32_t nb, int32_t ne, int32_t nc, const double *x,
                        const int32_t *ibc, const int32_t *idc, const int32_t *bewn,
                        tt_host **out);
/* Structured unit-cube mesh of n^3 27-node cells in the GiD conventions of
 * the reference's mesh/kva*.geo (walls 1:z=0 2:y=0 3:x=1 4:y=1 5:x=0 6:z=1). */
int tt_host_cube(int32_t n, tt_host **out);
/* ReadBic (lib_bic.f:50-502) from the text of a .bic file, DetKMunk3 and
 * FillKobc. */
int tt_host_read_bic(tt_host *h, const char *bic_text);
/* ReadInputFile (lib_IO.f:2-359) keywords that the hot path uses. */
int tt_host_read_inp(tt_host *h, const char *inp_text);
/* Fills a case view pointing into h (valid while h lives). */
int tt_host_case_view(tt_host *h, tt_case_view *v);
/* Initial fields after ReadBic (pointers into h). */
int tt_host_fields(tt_host *h, const double **vel, const double **vort, const double **qvort,
                   const double **temp, const double **qtemp);
/* Modified-Helmholtz codes of the single-domain form after ReadBic's MHEL
 * block (lib_bic.f:366-395) and its numbering (:463-480): kode(nn) as
 * mesh%mHkode and qkode(4 nbelem) as mesh%mHqfmkode (> 0: known value
 * number, < 0: minus the unknown number), the prescribed values u(nn) and
 * q(4 nbelem), the flux-node ids ibcf (4 nbelem, 1-based), the wall of
 * each boundary node iwn(nb) and of each boundary element bewn(nbelem).
 * The boundary flux points are element-major: point f of element e is entry
 * 4 e + f.  Pointers into h. */
int tt_host_mh_codes(tt_host *h, const int32_t **kode, const int32_t **qkode, const double **u,
                     const double **q, const int32_t **ibcf, const int32_t **iwn,
                     const int32_t **bewn, int32_t *nunk, int32_t *nknown);
/* Writes the mesh as a GiD .geo file the reference can read. */
int tt_host_write_geo(tt_host *h, const char *path);
void tt_host_free(tt_host *h);

/* ------------------------------------------------------------------------ */
/* The reference's files (host only, no device call): integral caches,
 * restart and iteration history, so that matrices and fields computed here
 * drop into an unmodified Fortran run and back.  Binary files are Fortran
 * unformatted sequential files (4-byte record lengths around each record),
 * vectors in records of 8192 doubles as wrvec writes them (lib_int.f:
 * 3231-3252), matrices column by column (WrSMat :3174-3186).               */

/* WriteMacroIntDisk (lib_int.f:1341-1391): tri.macro.int from the arrays
 * tt_macro_integrals returns (nicell*51 or nicell*27 rows, column-major),
 * INTversion 1 and the header nnodes, nbnodes. */
int tt_write_macro_int(const char *path, int32_t nicell, int32_t nnodes, int32_t nbnodes, const double *H,
                       const double *G, const double *B, const double *Abdx, const double *Abdy,
                       const double *Abdz, const double *Htx, const double *Hty, const double *Htz,
                       const double *Dx, const double *Dy, const double *Dz, const double ont[11]);
/* ReadMacroIntDisk (lib_int.f:1393-1441); *nnodes, *nbnodes from the header
 * (may be NULL). */
int tt_read_macro_int(const char *path, int32_t nicell, int32_t *nnodes, int32_t *nbnodes, double *H, double *G,
                      double *B, double *Abdx, double *Abdy, double *Abdz, double *Htx, double *Hty, double *Htz,
                      double *Dx, double *Dy, double *Dz, double ont[11]);
/* WritemHfullIntDisk (lib_fm.f:1822-1895): tri.modHelm.full.int from the
 * matrices tt_mhsd_matrix returns (U[k]: nnodes rows, Q[k]: nbnpof rows, the
 * order H, G, B, Bp, Abdx, Abdy, Abdz, Dx, Dy, Dz); header version, nnodes,
 * nbnodes, beta, sumdiff (the sum of diff in node order, the reference's
 * cache key), mumax; ont[5] last. */
int tt_write_mhsd_int(const char *path, int32_t version, int32_t nnodes, int32_t nbnodes, int32_t nbnpof,
                      double beta, double sumdiff, double mumax, const double *const U[10],
                      const double *const Q[10], const double ont[5]);
/* ReadmHfullIntDisk (lib_fm.f:1746-1820): hdr_out = version, nnodes, nbnodes;
 * hd_out = beta, sumdiff, mumax (CheckmHfullIntFile's key, :1897-1916). */
int tt_read_mhsd_int(const char *path, int32_t nbnpof, int32_t hdr_out[3], double hd_out[3],
                     double *const U[10], double *const Q[10], double ont[5]);
/* WriteSDIntDisk (lib_sdkm.f:2641-2676): tri.sdkm.int.N for `rows` boundary
 * rows (the reference's env%nmn) from the arrays tt_sd_integrals returns;
 * header nnodes, nbnodes, nproc (the number of rank files; 1 for one GPU). */
int tt_write_sdkm_int(const char *path, int32_t nnodes, int32_t nbnodes, int32_t nproc, int32_t rows,
                      const double *H, const double *Htx, const double *Hty, const double *Htz, const double *Dx,
                      const double *Dy, const double *Dz, const double ont[10]);
/* ReadSDIntDisk (lib_sdkm.f:2678-2719); hdr_out[3] (may be NULL) = nnodes,
 * nbnodes, nproc of the header (CheckSDIntFile, :2603-2620). */
int tt_read_sdkm_int(const char *path, int32_t rows, int32_t hdr_out[3], double *H, double *Htx, double *Hty,
                     double *Htz, double *Dx, double *Dy, double *Dz, double ont[10]);
/* WrstF (lib_IO.f:604-685), restart version 4: cnt%tstep, cnt%glnnlit,
 * cnt%rtime and the seventeen fields in WrstF's order: velocity,
 * pnlsvelocity, vorticity, ptsvorticity, pptsvorticity, pnlsvorticity
 * (nnodes,3); qvorticity (nq,3); temp, ptstemp, pptstemp, pnlstemp (nnodes);
 * qtemp (nq); dc, ptsdc, pptsdc, pnlsdc (nnodes); qdc (nq). */
int tt_write_restart(const char *path, int32_t nnodes, int32_t nq, int32_t tstep, int32_t glnnlit, double rtime,
                     const double *const fields[17]);
/* RrstF (lib_IO.f:505-602): versions 3 (the first twelve fields) and 4. */
int tt_read_restart(const char *path, int32_t nnodes, int32_t nq, int32_t *version, int32_t *tstep,
                    int32_t *glnnlit, double *rtime, double *const fields[17]);
/* tri.sta and tri.ite (either path may be NULL): the header lines of
 * lib_IO.f:1468-1485 (id_line in place of the program's name, version and
 * date; NULL gives a default) and per row the lines of CheckConvergencevw
 * (lib_util.f:1141-1146): tstep, tsnnlit, err (nrows x 9, row-major: the nine
 * relative changes of tt_case_picard_step, G10.5) and nit (nrows x 12,
 * row-major, I5). */
int tt_write_sta_ite(const char *sta_path, const char *ite_path, const char *id_line, int32_t nrows,
                     const int32_t *tstep, const int32_t *tsnnlit, const double *err, const int32_t *nit);

/* ------------------------------------------------------------------------ */
/* Request loop over whole cases (no reference counterpart: the reference runs
 * one case per process).  Worker threads, one per slot, admit submitted cases
 * in arrival order as device memory frees up (first fit against the budget).
 * A worker keeps a batch of up to max_batch active cases of one kind (same
 * mesh size and precision), steps them together with
 * tt_case_picard_step_batch, retires the finished ones (converged fields kept
 * on the host, device memory released) and refills the batch from the queue
 * between steps.  A case's results do not depend on its batch.  Batches of
 * different slots run concurrently on the device.                           */

typedef struct tt_server tt_server;
typedef struct tt_server_info {
  int32_t status;            /* 0 converged, 1 stopped at max_iter, -1 failed (error) */
  int32_t picard_iterations;
  int64_t lsqr_iterations;   /* all least-squares iterations of the case */
  double final_change;       /* largest relative change of the last Picard step */
  double t_submit, t_start, t_setup, t_end; /* s since tt_server_create: queued,
                                               admitted, built, finished */
  int64_t estimated_bytes, device_bytes;
  char error[256];
  int32_t batch_peak;        /* most cases that shared one of its Picard steps */
} tt_server_info;

/* slots concurrent cases; mem_budget device bytes for admission (0: 80 % of
 * the free memory); anchor_period of reduced-precision cases (0: 16);
 * max_iter caps the Picard steps per case (0: the case's own RELA limit). */
int tt_server_create(int32_t slots, int64_t mem_budget, int32_t anchor_period, int32_t max_iter,
                     tt_server **out);
/* As tt_server_create with batches of up to max_batch (1..16) cases per slot
 * (tt_server_create: one case per slot). */
int tt_server_create_batched(int32_t slots, int32_t max_batch, int64_t mem_budget, int32_t anchor_period,
                             int32_t max_iter, tt_server **out);
/* Queues the case held by h (which must live until the request finishes) in
 * precision prec (TT_PREC_*); *id identifies the request. */
int tt_server_submit(tt_server *s, tt_host *h, int32_t prec, int64_t *id);
/* Blocks until request id finishes; fills info. */
int tt_server_wait(tt_server *s, int64_t id, tt_server_info *info);
/* Converged fields of a finished request (arrays sized as tt_case_fields). */
int tt_server_fields(tt_server *s, int64_t id, double *vel, double *vort, double *qvort, double *temp,
                     double *qtemp);
/* Forgets a finished request. */
int tt_server_release(tt_server *s, int64_t id);
/* Queue length, device bytes admitted, budget. */
int tt_server_load(tt_server *s, int32_t *queued, int64_t *bytes_in_use, int64_t *budget);
/* Finishes the queued requests, stops the workers, frees the server. */
void tt_server_free(tt_server *s);

#ifdef __cplusplus
}
#endif
#endif
