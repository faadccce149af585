/*
 * or_compress.c -- CPU restatement of the reference's compressed matrix
 * formats and products.  TEST INFRASTRUCTURE ONLY (the oracle): see tt_oracle.h.
 *
 * Follows:
 *   aca_CrossAprox (type 1)       lib_ACA.f:126-181
 *   ACA_AB / ADAPTIVE_CROSS_APPROXIMATION2 / EPS_ACA_NORM / APPLAY_ACA (type 2)
 *                                 lib_ACA.f:372-510
 *   aca_RankCrossAprox (type 3)   lib_ACA.f:24-125
 *   FindLargestValue / CopyCR2ab / SubsApp   lib_ACA.f:343-367, 312-342, 287-311
 *   aca_abxv                      lib_ACA.f:226-249
 *   getRMSnorm                    lib_util.f:36-56
 *   Che_WaWT, Che_AbsPovp, Che_Zanemari, Che_FormCRS
 *                                 lib_WT.f:428-579, 719-740, 741-777, 355-401
 *   Che_WxV, Che_CRSmatMul, Che_WTxV, Che_WTAWx
 *                                 lib_WT.f:580-658, 330-354, 659-718, 297-329
 *   DvaNaK                        lib_WT.f:780-797
 *   cTree_MxV                     lib_ctree.f:99-164
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "tt_oracle.h"

#define R_(i, j) R[(i) + (size_t)m * (j)]

/* getRMSnorm (lib_util.f:36-56) */
static double rms_norm(int n, const double *v1, const double *v2) {
  double s = 0.0, s2 = 0.0;
  for (int i = 0; i < n; ++i) {
    s = s + (v1[i] - v2[i]) * (v1[i] - v2[i]);
    s2 = s2 + v1[i] * v1[i];
  }
  return s2 > 0.0 ? sqrt(s / s2) : sqrt(s);
}

void or_aca_abxv(int m, int n, int rank, const double *a, const double *b, const double *x, double *y) {
  /* tmpv = b x, y = a tmpv (a(m,rank), b(rank,n) column-major) */
  double *t = (double *)calloc(rank > 0 ? rank : 1, sizeof(double));
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < rank; ++k) t[k] = t[k] + b[k + (size_t)rank * j] * x[j];
  for (int i = 0; i < m; ++i) y[i] = 0.0;
  for (int k = 0; k < rank; ++k)
    for (int i = 0; i < m; ++i) y[i] = y[i] + a[i + (size_t)m * k] * t[k];
  free(t);
}

/* y = A x of the full matrix (MATMUL(FullMat, vec)) */
static void full_mv(int m, int n, const double *A, const double *x, double *y) {
  for (int i = 0; i < m; ++i) y[i] = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) y[i] = y[i] + A[i + (size_t)m * j] * x[j];
}

/* one cross step of types 1 and 3: largest |R| in column-major scan order
 * (first maximum wins), a(:,k) = R(:,lcol)/delta, b(k,:) = R(lrow,:),
 * R -= a(:,k) b(k,:) */
static void cross_step_cm(int m, int n, double *R, double *a, double *b, int ldb, int k) {
  double maxv = -1.0;
  int lrow = 0, lcol = 0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i)
      if (fabs(R_(i, j)) > maxv) {
        maxv = fabs(R_(i, j));
        lrow = i;
        lcol = j;
      }
  double delta = R_(lrow, lcol);
  for (int i = 0; i < m; ++i) a[i + (size_t)m * k] = delta != 0.0 ? R_(i, lcol) / delta : R_(i, lcol);
  for (int j = 0; j < n; ++j) b[k + (size_t)ldb * j] = R_(lrow, j);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) R_(i, j) = R_(i, j) - a[i + (size_t)m * k] * b[k + (size_t)ldb * j];
}

int or_aca(int type, int m, int n, double *R, double eps, const double *vec, double *a, double *b, double *norm) {
  const int mn = m < n ? m : n;
  double *A0 = NULL, *rez1 = NULL, *rez2 = NULL;
  if (vec) {
    A0 = (double *)malloc(sizeof(double) * (size_t)m * n);
    memcpy(A0, R, sizeof(double) * (size_t)m * n);
    rez1 = (double *)malloc(sizeof(double) * m);
    rez2 = (double *)malloc(sizeof(double) * m);
    full_mv(m, n, A0, vec, rez1);
  }
  int rank = 0;
  if (type == 1) {
    /* fixed rank INT(eps*min(m,n)); b is (rank, n) */
    rank = (int)(eps * (double)mn);
    for (int k = 0; k < rank; ++k) cross_step_cm(m, n, R, a, b, rank, k);
  } else if (type == 3) {
    /* grow the rank until the random-vector test passes; b is (min(m,n), n)
     * while growing, compacted to (rank, n) at the end */
    double *bb = (double *)calloc((size_t)mn * n, sizeof(double));
    rank = mn;
    for (int k = 0; k < mn; ++k) {
      cross_step_cm(m, n, R, a, bb, mn, k);
      or_aca_abxv(m, n, mn, a, bb, vec, rez2); /* a, bb hold zeros beyond k */
      if (rms_norm(m, rez1, rez2) < eps) {
        rank = k + 1;
        break;
      }
    }
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < rank; ++k) b[k + (size_t)rank * j] = bb[k + (size_t)mn * j];
    free(bb);
  } else {
    /* type 2, ACA_AB: Frobenius-type stopping norm of EPS_ACA_NORM; the
     * vectors are stored at descending slots of (M, RANK) work arrays */
    const int RK = mn;
    double *WA = (double *)malloc(sizeof(double) * (size_t)m * RK);
    double *WB = (double *)malloc(sizeof(double) * (size_t)RK * n);
    double snorm = 0.0, nrm = 1.0;
    int counter = RK;
    while (counter >= 1 && nrm > eps) {
      /* APPLAY_ACA: largest |R| in row-major scan order (first wins) */
      double maxs = -1.0;
      int iz = 0, jz = 0;
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j)
          if (fabs(R_(i, j)) > maxs) {
            iz = i;
            jz = j;
            maxs = fabs(R_(i, j));
          }
      double sigma = 1.0 / R_(iz, jz);
      double *av = WA + (size_t)m * (counter - 1);
      double *bv = (double *)malloc(sizeof(double) * n);
      for (int i = 0; i < m; ++i) av[i] = R_(i, jz);
      for (int j = 0; j < n; ++j) bv[j] = R_(iz, j) * sigma;
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) R_(i, j) = R_(i, j) - av[i] * bv[j];
      /* EPS_ACA_NORM */
      double s1 = 0.0, s2 = 0.0, s3 = 0.0;
      for (int j = 0; j < m; ++j) s1 = s1 + av[j] * av[j];
      for (int j = 0; j < n; ++j) s2 = s2 + bv[j] * bv[j];
      for (int j = 0; j < n; ++j) WB[(counter - 1) + (size_t)RK * j] = bv[j];
      for (int k = counter - 1; k < RK; ++k)
        for (int j = 0; j < m; ++j) s3 = s3 + WA[j + (size_t)m * k] * av[j];
      for (int j = 0; j < n; ++j)
        for (int k = counter - 1; k < RK; ++k) s3 = s3 + WB[k + (size_t)RK * j] * bv[j];
      snorm = snorm + s3 + s1 * s2;
      nrm = sqrt(s1) * sqrt(s2) / sqrt(snorm);
      free(bv);
      ++rank;
      --counter;
    }
    /* FILL_AB on the way back: the k-th vector computed becomes column k */
    for (int k = 0; k < rank; ++k) {
      int slot = RK - 1 - k;
      memcpy(a + (size_t)m * k, WA + (size_t)m * slot, sizeof(double) * m);
      for (int j = 0; j < n; ++j) b[k + (size_t)rank * j] = WB[slot + (size_t)RK * j];
    }
    free(WA);
    free(WB);
  }
  if (vec && norm) {
    or_aca_abxv(m, n, rank, a, b, vec, rez2);
    *norm = rms_norm(m, rez1, rez2);
  }
  free(A0);
  free(rez1);
  free(rez2);
  return rank;
}
#undef R_

/* ------------------------------------------------------------------------ */
/* stretched Haar wavelets                                                   */

int or_dvanak(int n) {
  if (n == 1) return 0;
  for (int i = 0; i <= 100; ++i)
    if ((1L << i) >= n) return i - 1;
  return -1;
}

/* forward transform of one sequence (Che_WxV body on a gathered copy):
 * r (length nv, stride st) is padded to n = 2^(kmax+1) -- dup != 0 repeats
 * each entry from podv on (row transform and Che_WxV), dup == 0 follows it
 * with a zero (column transform) -- and replaced by its first nv coefficients */
static void haar_fwd(double *r, int nv, size_t st, int dup) {
  int kmax = or_dvanak(nv);
  int n = 1 << (kmax + 1);
  double *v = (double *)malloc(sizeof(double) * n);
  double norma = 1.0 / sqrt((double)n);
  if (n != nv) {
    int podv = 2 * (nv - (1 << kmax)) + 1;
    for (int ii = 1; ii <= podv - 1; ++ii) v[ii - 1] = r[(ii - 1) * st];
    int ii = podv - 1, jj = ii;
    do {
      ++ii;
      ++jj;
      v[ii - 1] = r[(jj - 1) * st];
      ++ii;
      v[ii - 1] = dup ? r[(jj - 1) * st] : 0.0;
    } while (ii < n);
  } else {
    for (int ii = 1; ii <= n; ++ii) v[ii - 1] = r[(ii - 1) * st];
  }
  for (int ii = 1; ii <= nv; ++ii) {
    double acc = 0.0;
    if (ii == 1) {
      for (int i = 1; i <= n; ++i) acc = acc + v[i - 1] * norma;
    } else {
      int k = or_dvanak(ii);
      int l = ii - (1 << k);
      int d = 1 << (kmax - k);
      int psp = (2 * l - 2) * d + 1, psr = (2 * l - 1) * d, pzg = 2 * l * d;
      double valcek = pow(2.0, (double)k / 2) * norma;
      for (int i = psp; i <= psr; ++i) acc = acc + valcek * v[i - 1];
      for (int i = psr + 1; i <= pzg; ++i) acc = acc - valcek * v[i - 1];
    }
    r[(ii - 1) * st] = acc;
  }
  free(v);
}

void or_wt_wxv(int n, double *r) { haar_fwd(r, n, 1, 1); }

void or_wt_wtxv(int nvek, double *v) {
  int kmax = or_dvanak(nvek);
  int n = 1 << (kmax + 1);
  double *r = (double *)malloc(sizeof(double) * n);
  double norma = 1.0 / sqrt((double)n);
  for (int ii = 1; ii <= n; ++ii) {
    double acc = v[0] * norma;
    for (int k = 0; k <= kmax; ++k) {
      double valcek = pow(2.0, (double)k / 2) * norma;
      int d = 1 << (kmax - k);
      /* iks=1+int((real(ii)-0.01)/real(DvaNaKmaxmk)), single precision */
      float t = ((float)ii - 0.01f) / (float)d;
      int iks = 1 + (int)t;
      if (iks % 2 == 0) {
        int jj = (1 << k) + iks / 2;
        if (jj <= nvek) acc = acc - valcek * v[jj - 1];
      } else {
        int jj = (1 << k) + (iks + 1) / 2;
        if (jj <= nvek) acc = acc + valcek * v[jj - 1];
      }
    }
    r[ii - 1] = acc;
  }
  if (n != nvek) {
    int podv = 2 * (nvek - (1 << kmax)) + 1;
    for (int ii = 1; ii <= podv - 1; ++ii) v[ii - 1] = r[ii - 1];
    int ii = podv - 1, jj = ii;
    do {
      ii += 2;
      ++jj;
      v[jj - 1] = r[ii - 1];
    } while (ii < n);
  } else {
    memcpy(v, r, sizeof(double) * n);
  }
  free(r);
}

void or_wt_wawt(int m, int n, double *A) {
  for (int kk = 0; kk < n; ++kk) haar_fwd(A + (size_t)m * kk, m, 1, 1); /* columns: W a */
  for (int kk = 0; kk < m; ++kk) haar_fwd(A + kk, n, (size_t)m, 0);     /* rows */
}

double or_wt_threshold(int m, int n, double *A, double kappa, int *nnz) {
  /* Che_AbsPovp */
  double povp = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) povp = povp + fabs(A[j + (size_t)m * i]);
  povp = povp / (float)m / (float)n; /* povp/real(nVR)/real(nST) */
  double tresh = kappa * povp;
  /* Che_Zanemari */
  int nz = 0;
  for (int i = 2; i <= m; ++i) {
    int allzero = 1;
    for (int j = 2; j <= n - 1; ++j) {
      double *p = &A[(i - 1) + (size_t)m * (j - 1)];
      if (fabs(*p) < tresh) {
        *p = 0.0;
        ++nz;
      } else {
        allzero = 0;
      }
    }
    double *p = &A[(i - 1) + (size_t)m * (n - 1)];
    if (!allzero) {
      if (fabs(*p) < tresh) {
        *p = 0.0;
        ++nz;
      }
    } else if (*p == 0.0) {
      *p = 1.0e-14;
    }
  }
  if (nnz) *nnz = m * n - nz;
  return tresh;
}

int or_wt_form_csr(int m, int n, const double *A, int *zvr, int *c, double *v) {
  /* Che_FormCRS (every row assumed to hold a non-zero) */
  int st = 0, stvr = 0, ibu = 0;
  for (int i = 0; i <= m; ++i) zvr[i] = 0;
  for (int i = 1; i <= m; ++i)
    for (int j = 1; j <= n; ++j) {
      double a = A[(i - 1) + (size_t)m * (j - 1)];
      if (a != 0.0) {
        ++st;
        v[st - 1] = a;
        c[st - 1] = j;
        if (i != ibu) {
          ++stvr;
          ibu = i;
          zvr[stvr - 1] = st;
        }
      }
    }
  zvr[stvr] = st + 1;
  return st;
}

void or_wt_wtawx(int nrow, int ncol, const int *zvr, const int *c, const double *v, const double *x, double *y) {
  double *vx = (double *)malloc(sizeof(double) * ncol);
  memcpy(vx, x, sizeof(double) * ncol);
  or_wt_wxv(ncol, vx);
  for (int r = 0; r < nrow; ++r) {
    double acc = 0.0;
    for (int k = zvr[r]; k <= zvr[r + 1] - 1; ++k) acc = acc + v[k - 1] * vx[c[k - 1] - 1];
    y[r] = acc;
  }
  or_wt_wtxv(nrow, y);
  free(vx);
}

/* ------------------------------------------------------------------------ */
void or_hmat_mxv(int nrow, int nparts, const int *rptr, const int *rnodes, const int *cptr, const int *cnodes,
                 const int *ctype, const int *rank, const long *doff, const double *dpool, const long *ioff,
                 const int *ipool, const double *x, double *y) {
  for (int i = 0; i < nrow; ++i) y[i] = 0.0;
  for (int p = 0; p < nparts; ++p) {
    int nr = rptr[p + 1] - rptr[p], nc = cptr[p + 1] - cptr[p];
    double *vv = (double *)malloc(sizeof(double) * (nc > 0 ? nc : 1));
    double *rr = (double *)malloc(sizeof(double) * (nr > 0 ? nr : 1));
    for (int ic = 0; ic < nc; ++ic) vv[ic] = x[cnodes[cptr[p] + ic] - 1];
    const double *d = dpool + doff[p];
    if (ctype[p] == 1) {
      full_mv(nr, nc, d, vv, rr);
    } else if (ctype[p] == 2) {
      const int *zvr = ipool + ioff[p], *cc = zvr + nr + 1;
      or_wt_wtawx(nr, nc, zvr, cc, d, vv, rr);
    } else {
      or_aca_abxv(nr, nc, rank[p], d, d + (size_t)nr * rank[p], vv, rr);
    }
    for (int ir = 0; ir < nr; ++ir) y[rnodes[rptr[p] + ir] - 1] = y[rnodes[rptr[p] + ir] - 1] + rr[ir];
    free(vv);
    free(rr);
  }
}
