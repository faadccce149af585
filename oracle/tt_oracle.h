/*
 * tt_oracle.h -- CPU restatement of the Tritok boundary-element hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP library
 * (tritok_amd) is compared against, and the timed CPU baseline.  Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it; the
 * product never links or calls it.
 *
 * Every routine follows one reference routine operation by operation (same loop
 * order, same expression association) so that results agree with the reference
 * binary to rounding.  Citations are file:line into the reference's src/.
 *
 * Array conventions (the reference's, so a Fortran host can pass its arrays
 * unchanged): column-major, index arrays 1-based.
 *   x[nn*3]            node coordinates  x(nnodes,3)
 *   idc[nc*27]         cell connectivity idc(nicell,27)
 *   ibc[ne*9]          boundary elements ibc(nbelem,9)
 *   ibf[nc*6*4]        cell face flux nodes ibf(nicell,6,4)
 *   xq[nq*3]           flux node coordinates xq(nq,3)
 *   flpm[nc*6]         flux sign per cell face flpm(nicell,6)
 *   gbn[nb], lbn[nn]   boundary node list / inverse (0 = interior)
 *
 * Pinning (how each function is checked in tests/test_oracle.py):
 *   or_gauss                 Golub-Welsch closed form (numpy leggauss) and the
 *                            polynomial-exactness property.
 *   or_shape27/9/dl34        Kronecker property at nodes, partition of unity,
 *                            exact integrals of products (lib_int.f:3461 21/4).
 *   or_inted27, or_inteb9    reference binary matrices (tests/golden npz files),
 *                            rigid-body / linear-field identities, brute force.
 *   or_macro_integrals       reference tri.macro.int (dc2, nc2, jit2, dc8).
 *   or_macro_integrals_mh    reference tri.modHelm.macro.int (uniform and
 *                            variable mu); mu = 0 equals or_macro_integrals
 *                            exactly; closed-form modified-Helmholtz fields.
 *   or_sd_integrals          reference tri.sdkm.int.1 (dc2, jit2, dc8 rows).
 *   or_aca                   brute-force cross approximation written in the
 *                            tests; exact recovery of rank-r matrices; the
 *                            Frobenius error bound at the stopping rank.
 *   or_wt_*                  the stretched-Haar matrix built explicitly in the
 *                            tests (orthonormal rows at n = 2^k, exact
 *                            round trip at kappa = 0 for any size).
 *   or_hmat_mxv              the dense product of the assembled matrix.
 *   or_lsqr2, or_dgefa/dgesl scipy lsqr / LAPACK getrf on the same systems; the
 *                            reference's per-solve iteration counts (tri.ite).
 *   or_picard                reference converged fields (tri.rst) and per-
 *                            iteration histories (tri.sta / tri.ite).
 */
#ifndef TT_ORACLE_H
#define TT_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Gauss-Legendre rule by reference index 1..9 -> N = 2,4,6,8,16,32,40,48,64
 * (GaussPosWeights, lib_int.f:2673-3173), computed.  Returns N, fills gi[N],
 * ome[N]. */
int or_gauss(int rule, double *gi, double *ome);

/* Shape functions (lib_interpol.f): 27-node cell (cshape27 :352-405),
 * 9-node face (cshape9 :597-619), discontinuous 4-node flux (dl34shape4
 * :571-592). */
void or_shape27(double xi, double eta, double zeta, double *fi);
void or_shape9(double xi, double eta, double *fi);
void or_dl34(double xi, double eta, double *fi);

/* Quadrature configuration: GAUS kmBr kmBs kmDr kmDs and IIDI, and
 * optionally the table of TYPE(GaussType) (gi(220), ome(220), ng1(9), ng2(9));
 * with a NULL table the rules come from or_gauss. */
typedef struct {
  int kmBr, kmBs, kmDr, kmDs, iiDiv;
  const double *gi, *ome;
  const int *ng1, *ng2;
} or_quad;

/* INTEDc27dcLap (lib_int.f:4-398): integrals over one 27-node cell with 8-node
 * geometry.  corners[8][3] in the reference's argument order x1..x8 =
 * idc(6,7,2,3,5,8,1,4).  isrc 0 = regular, 1..27 node, 28..51 flux point.
 * Outputs be/dxe/dye/dze[27].  Returns the number of kernel evaluations. */
long or_inted27(const or_quad *q, const double p[3], const double corners[8][3], int isrc,
                double *be, double *dxe, double *dye, double *dze);

/* INTEBc9kmdclap (lib_int.f:1462-1779) when macro != 0, INTEBc9sdkm
 * (lib_sdkm.f:2082-2376, no subdivision, no ge/ad) when macro == 0.
 * corners[4][3] = ibc(1,3,5,7).  isrc 0 regular, 1..9 node, 10..13 flux.
 * Outputs he, adx, ady, adz, heyz, hezx, hexy [9] and ge[4] (macro only;
 * pass NULL for unused). Returns the number of kernel evaluations. */
long or_inteb9(const or_quad *q, int macro, const double p[3], const double corners[4][3], int isrc,
               double minedge, double *he, double *ge, double *adx, double *ady, double *adz,
               double *heyz, double *hezx, double *hexy);

/* FMATkmdcSteady (lib_fmat.f:3-446): all per-cell macro matrices, column-major.
 * Rows of H,G,B,Abd*,D* = nc*51, rows of Ht* = nc*27.  ont[11] in the order of
 * the reference's cache file.  copy as the COPY keyword.  Cells are processed in
 * parallel with OpenMP when nthreads > 1. */
void or_macro_integrals(const or_quad *q, int nn, int nc, int nq, const double *x,
                        const int *idc, const int *ibf, const double *xq, const int *flpm,
                        int copy, double *H, double *G, double *B, double *Abdx, double *Abdy,
                        double *Abdz, double *Htx, double *Hty, double *Htz, double *Dx,
                        double *Dy, double *Dz, double ont[11], int nthreads);

/* Modified-Helmholtz variants (lib_mHelm.f): INTEDc27modHelm (:1637-2033)
 * and INTEBc9modHelm (:2034-2348) -- kernel exp(-mu r)/(4 pi r), gradient
 * scaled by (1 + mu r); FMATmodHelm (:2-483) with mu = sqrt(beta/diff) of each
 * source, diff(nn) at nodes and qdiff(nq) at flux nodes; *mumax = the largest
 * mu used.  mu = 0 (beta = 0) gives the Laplace results bit for bit. */
long or_inted27_mh(const or_quad *q, const double p[3], const double corners[8][3], int isrc,
                   double mu, double *be, double *dxe, double *dye, double *dze);
long or_inteb9_mh(const or_quad *q, const double p[3], const double corners[4][3], int isrc,
                  double minedge, double mu, double *he, double *ge, double *adx, double *ady,
                  double *adz, double *heyz, double *hezx, double *hexy);
void or_macro_integrals_mh(const or_quad *q, int nn, int nc, int nq, const double *x,
                           const int *idc, const int *ibf, const double *xq, const int *flpm,
                           int copy, const double *diff, const double *qdiff, double beta,
                           double *H, double *G, double *B, double *Abdx, double *Abdy,
                           double *Abdz, double *Htx, double *Hty, double *Htz, double *Dx,
                           double *Dy, double *Dz, double ont[11], double *mumax, int nthreads);

/* Kernel evaluations (quadrature points) of the last or_macro_integrals /
 * or_macro_integrals_mh call: over cells and over faces. */
void or_last_points(long *domain, long *boundary);

/* FMATsdkm (lib_sdkm.f:1286-1530) for boundary rows row0..row1-1 (0-based),
 * matrices column-major with leading dimension ld (>= row1-row0):
 * H,Htx,Hty,Htz (ld x nb), Dx,Dy,Dz (ld x nn).  ont[10] as the cache file. */
void or_sd_integrals(const or_quad *q, int nn, int nb, int ne, int nc, const double *x,
                     const int *ibc, const int *idc, const int *gbn, const int *lbn, int row0,
                     int row1, int ld, double *H, double *Htx, double *Hty, double *Htz,
                     double *Dx, double *Dy, double *Dz, double ont[10], int nthreads);

/* ---- single-domain modified Helmholtz (EQNS imH 3-7, lib_fm.f) ---------- */

/* FMATsdmHfs (lib_fm.f:1183-1645).  U rows: nodes; Q rows: boundary flux
 * points element-major (row 4e + f, source xq(ibcf(e, f))).  U[10], Q[10]:
 * H, G, B, Bp, Abdx, Abdy, Abdz, Dx, Dy, Dz, column-major with ld = nn / 4 ne
 * (Bp is left zero: or_mhsd_bprime).  ibcf: ne x 4 column-major (1-based flux
 * node ids).  ont = ontH, ontG, ontDx, ontDy, ontDz. */
void or_mhsd_integrals(const or_quad *q, int nn, int nb, int ne, int nc, int nq, const double *x,
                       const int *ibc, const int *idc, const int *ibf, const double *xq,
                       const int *ibcf, const int *gbn, const int *lbn, const double *diff,
                       const double *qdiff, double beta, double **U, double **Q, double ont[5],
                       double *mumax, int nthreads);
/* MakeBprime (lib_fm.f:800-858) */
void or_mhsd_bprime(int nn, int ne, const int *ibcf, const double *diff, const double *qdiff,
                    double beta, const double *UB, double *UBp, const double *QB, double *QBp);
/* fmmHSysM (lib_fm.f:1061-1181): kode(nn), qkode(4 ne) as mesh%mHkode /
 * mHqfmkode; A (nunk x nunk), RHS (nunk x nknown) column-major. */
void or_mhsd_sysm(int nn, int nb, int ne, const int *ibcf, const int *lbn, const int *kode,
                  const int *qkode, int nunk, int nknown, const double *diff, const double *qdiff,
                  const double *UH, const double *UG, const double *QH, const double *QG, double *A,
                  double *RHS);
/* fmSolveModHelmDC's right-hand side (lib_fm.f:905-1016); velocity and dgrad
 * nn x 3 column-major. */
void or_mhsd_rhs(int nn, int nb, int ne, const int *ibcf, const int *gbn, const int *kode,
                 const int *qkode, int nunk, int nknown, const double *diff, const double *qdiff,
                 double beta, const double *u, const double *qu, const double *velocity,
                 const double *rhsv, const double *dgrad, double **U, double **Q, const double *RHS,
                 double *r);
/* the product vectors and right-hand side of the matrix-vector variants
 * (fmSolveModHelmDCaca / wawt / hmat / mm) */
void or_mhsd_vectors(int nn, int nb, const int *gbn, const double *u, const double *velocity,
                     const double *rhsv, const double *dgrad, double **v);
void or_mhsd_rhs_mv(int nn, int ne, const int *kode, const int *qkode, int nunk, int nknown,
                    const double *u, const double *qu, const double *RHS, double **yu, double **yq,
                    double *r);

/* ---- compressed matrices (lib_ACA.f, lib_WT.f, lib_ctree.f) ------------ */

/* Cross approximation of R (m x n, column-major, overwritten by the residual):
 * type 1 aca_CrossAprox (rank INT(eps*min(m,n))), type 2 ACA_AB (stop when the
 * EPS_ACA_NORM estimate <= eps), type 3 aca_RankCrossAprox (stop when the
 * getRMSnorm test with vec < eps; vec required).  a (m x rank) and b (rank x n)
 * need room for rank = min(m,n).  *norm (if vec) = getRMSnorm of A vec against
 * a(b vec).  Returns the rank. */
int or_aca(int type, int m, int n, double *R, double eps, const double *vec, double *a, double *b,
           double *norm);
/* aca_abxv: y = a (b x) */
void or_aca_abxv(int m, int n, int rank, const double *a, const double *b, const double *x, double *y);

/* Stretched Haar wavelets: DvaNaK; Che_WxV and Che_WTxV in place on a vector
 * of length n; Che_WaWT in place on A (m x n); Che_AbsPovp + Che_Zanemari with
 * threshold kappa*mean|A| (returns the threshold, *nnz as the reference
 * counts it); Che_FormCRS (1-based zvr[m+1], c, v; returns the entries
 * written); Che_WTAWx: y = W^T (C (W x)). */
int or_dvanak(int n);
void or_wt_wxv(int n, double *r);
void or_wt_wtxv(int n, double *v);
void or_wt_wawt(int m, int n, double *A);
double or_wt_threshold(int m, int n, double *A, double kappa, int *nnz);
int or_wt_form_csr(int m, int n, const double *A, int *zvr, int *c, double *v);
void or_wt_wtawx(int nrow, int ncol, const int *zvr, const int *c, const double *v, const double *x,
                 double *y);

/* cTree_MxV: y = sum over parts of scatter(op_p gather(x)).  Part p has rows
 * rnodes[rptr[p]..rptr[p+1]) and columns cnodes[cptr[p]..) (1-based), ctype 1
 * (dense, nr x nc column-major at dpool+doff[p]), 2 (wavelet CSR: zvr and c at
 * ipool+ioff[p], values at dpool+doff[p]) or 3 (ACA: a then b at
 * dpool+doff[p], rank[p]).  Parts are summed in order. */
void or_hmat_mxv(int nrow, int nparts, const int *rptr, const int *rnodes, const int *cptr, const int *cnodes,
                 const int *ctype, const int *rank, const long *doff, const double *dpool, const long *ioff,
                 const int *ipool, const double *x, double *y);

/* ---- linear algebra (lib_Solve_v11.f, lib_Blas.f) ---------------------- */

/* DNRM2 (lib_Blas.f:263-322), scaled two-norm. */
double or_dnrm2(int n, const double *x);

/* prelsqr2 (lib_Solve_v11.f:4501-4535): q_j = 1/||A(:,j)||, A <- A diag(q).
 * CSR is 1-based (ia[m+1], ja[nnz]). */
void or_prelsqr2(int m, int n, const int *ia, const int *ja, double *a, double *q);

/* slvlsqr2 + LSQR (lib_Solve_v11.f:4401-4500, :3128-3887).  eps4 is the
 * single-precision tolerance the reference passes.  x holds the hot start on
 * entry and the solution on exit.  Returns istop; *nit = iterations. */
int or_slvlsqr2(int m, int n, const int *ia, const int *ja, const double *a, const double *q,
                const double *b, double *x, int maxit, float eps4, int *nit);

/* LINPACK dgefa / dgesl (lib_Solve_v11.f:787-1037), column-major a(lda,n),
 * pivots stored as doubles (ipvt is REAL(8) in the reference). */
int or_dgefa(double *a, int lda, int n, double *ipvt);
void or_dgesl(const double *a, int lda, int n, const double *ipvt, double *b);

/* ---- the Picard driver ------------------------------------------------- */

/* A complete steady case after preprocessing (mesh, boundary conditions,
 * numbering), as the reference's main.f90 holds it before the time loop.  All
 * arrays follow the conventions above; codes follow lib_bic.f:411-488
 * (negative = unknown number, positive = known number). */
typedef struct {
  int nn, nb, ne, nc, nq;
  const double *x;
  const int *idc, *ibc, *ibf, *flpm, *gbn, *lbn;
  const double *xq;
  const double *nx, *ny, *nz;   /* source point normals, SourcePointNormals */
  const int *kode;              /* kode(nb,3) */
  const int *kobc;              /* kobc(nb), nearest node for dv/dn=0 */
  const int *wkode, *wqkode;    /* wkode(nn,3), wqkode(nq,3) */
  const int *Tkode, *Tqkode;    /* Tkode(nn), Tqkode(nq) */
  int nunk[3], nbk[3], Tnunk, Tnb;
  const int *kmu;               /* kmp%u(nn) */
  int nsmcol, nrhscol;
  /* physics and numerics (tri.inp) */
  int iDT;                      /* 0 or 1 */
  double ReN, RaN, PrN, gx, gy, gz, tstep;
  double urBw[3], urDv[3], urDw[3], urDT;
  double eps;                   /* RELA eps */
  int nnlit, lsqs_maxit, iDLSE;
  float lsqs_eps, dlse_r, dlse_epsth;
  or_quad quad;
  int copy;
  int iBw;                      /* 1: single-domain kinematics, 0: none */
} or_case;

/* Fields in/out: velocity(nn,3), vorticity(nn,3), qvorticity(nq,3), temp(nn),
 * qtemp(nq), initialised as after ReadBic.  History out (may be NULL):
 * sta[maxhist*9] relative changes, ite[maxhist*12] solver iterations.
 * Returns the number of Picard iterations run. */
int or_picard(const or_case *c, double *velocity, double *vorticity, double *qvorticity,
              double *temp, double *qtemp, int maxhist, double *sta, int *ite, int nthreads,
              double *seconds_per_iter);

/* Pieces of the Picard step, exported for unit tests and the CPU baseline.
 * sdkm_rhs (lib_sdkm.f:349-519) for equation eqn (1..3) over rows r0..r1-1. */
void or_sdkm_rhs(int eqn, int nn, int nb, const int *lbn, const int *gbn, const int *kode,
                 const double *nx, const double *ny, const double *nz, const double *H,
                 const double *Htx, const double *Hty, const double *Htz, const double *Dx,
                 const double *Dy, const double *Dz, int ld, const double *vel, const double *vort,
                 double *dfi, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
