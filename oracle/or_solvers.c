/*
 * or_solvers.c -- CPU restatement of the reference's linear algebra.
 * TEST INFRASTRUCTURE ONLY (the oracle): see tt_oracle.h.
 *
 *   DNRM2                    lib_Blas.f:263-322 (Hammarling's scaled sum of squares)
 *   prelsqr2, aprod 3/4      lib_Solve_v11.f:4501-4535, :3967-4022
 *   slvlsqr2                 lib_Solve_v11.f:4401-4500 (hot start, column scaling)
 *   LSQR                     lib_Solve_v11.f:3128-3887 (Paige & Saunders, damp = 0,
 *                            wantse = .false.; printing removed)
 *   aprod 1/2 (matvec1, tmatvec2)  lib_Solve_v11.f:3914-3966
 *   d2norm                   lib_Solve_v11.f:4023-4069
 *   dgefa / dgesl (LINPACK)  lib_Solve_v11.f:787-1037
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "tt_oracle.h"

double or_dnrm2(int n, const double *x) {
  if (n < 1) return 0.0;
  if (n == 1) return fabs(x[0]);
  double scale = 0.0, ssq = 1.0;
  for (int i = 0; i < n; ++i) {
    if (x[i] != 0.0) {
      double a = fabs(x[i]);
      if (scale < a) {
        double r = scale / a;
        ssq = 1.0 + ssq * (r * r);
        scale = a;
      } else {
        double r = a / scale;
        ssq = ssq + r * r;
      }
    }
  }
  return scale * sqrt(ssq);
}

static double d2norm(double a, double b) {
  double scale = fabs(a) + fabs(b);
  if (scale == 0.0) return 0.0;
  double sa = a / scale, sb = b / scale;
  return scale * sqrt(sa * sa + sb * sb);
}

/* y = y + A x (matvec1: per-row sum then added) */
static void matvec1(int m, const int *ia, const int *ja, const double *a, const double *u, double *v) {
  for (int i = 0; i < m; ++i) {
    double sum = 0.0;
    for (int j = ia[i] - 1; j < ia[i + 1] - 1; ++j) sum = sum + a[j] * u[ja[j] - 1];
    v[i] = v[i] + sum;
  }
}

/* x = x + A^T y (tmatvec2: scatter in row order) */
static void tmatvec2(int m, const int *ia, const int *ja, const double *a, const double *u, double *v) {
  for (int j = 0; j < m; ++j)
    for (int i = ia[j] - 1; i < ia[j + 1] - 1; ++i) {
      int ii = ja[i] - 1;
      v[ii] = v[ii] + a[i] * u[j];
    }
}

void or_prelsqr2(int m, int n, const int *ia, const int *ja, double *a, double *q) {
  for (int i = 0; i < n; ++i) q[i] = 0.0;
  for (int j = 0; j < m; ++j)
    for (int i = ia[j] - 1; i < ia[j + 1] - 1; ++i) {
      int ii = ja[i] - 1;
      q[ii] = q[ii] + a[i] * a[i];
    }
  for (int i = 0; i < n; ++i) q[i] = 1.0 / sqrt(q[i]);
  for (int i = 0; i < m; ++i)
    for (int ij = ia[i] - 1; ij < ia[i + 1] - 1; ++ij) a[ij] = a[ij] * q[ja[ij] - 1];
}

static void dscal(int n, double da, double *x) {
  for (int i = 0; i < n; ++i) x[i] = da * x[i];
}

/* LSQR with damp = 0; u holds b on entry (destroyed).  Returns istop. */
static int lsqr(int m, int n, const int *ia, const int *ja, const double *a, double *u, double *x, double *v,
                double *w, double atol, double btol, double conlim, int itnlim, int *itn_out) {
  const double zero = 0.0, one = 1.0;
  int itn = 0, istop = 0;
  double ctol = conlim > zero ? one / conlim : zero;
  double anorm = zero, acond = zero, dnorm = zero, res2 = zero, psi = zero;
  double xnorm = zero, xnorm1 = zero, cs2 = -one, sn2 = zero, z = zero;
  double alpha, beta, bnorm, rnorm, arnorm, rhobar, phibar;
  for (int i = 0; i < n; ++i) v[i] = x[i] = zero;
  alpha = zero;
  beta = or_dnrm2(m, u);
  if (beta > zero) {
    dscal(m, one / beta, u);
    tmatvec2(m, ia, ja, a, u, v);
    alpha = or_dnrm2(n, v);
  }
  if (alpha > zero) {
    dscal(n, one / alpha, v);
    memcpy(w, v, sizeof(double) * n);
  }
  arnorm = alpha * beta;
  if (arnorm == zero) {
    *itn_out = 0;
    return 0;
  }
  rhobar = alpha;
  phibar = beta;
  bnorm = beta;
  rnorm = beta;
  for (;;) {
    itn = itn + 1;
    dscal(m, -alpha, u);
    matvec1(m, ia, ja, a, v, u);
    beta = or_dnrm2(m, u);
    double temp = d2norm(alpha, beta);
    temp = d2norm(temp, zero);
    anorm = d2norm(anorm, temp);
    if (beta > zero) {
      dscal(m, one / beta, u);
      dscal(n, -beta, v);
      tmatvec2(m, ia, ja, a, u, v);
      alpha = or_dnrm2(n, v);
      if (alpha > zero) dscal(n, one / alpha, v);
    }
    double rhbar1 = rhobar;
    double rho = d2norm(rhbar1, beta);
    double cs = rhbar1 / rho;
    double sn = beta / rho;
    double theta = sn * alpha;
    rhobar = -cs * alpha;
    double phi = cs * phibar;
    phibar = sn * phibar;
    double tau = sn * phi;
    double t1 = phi / rho;
    double t2 = -theta / rho;
    double t3 = one / rho;
    double dknorm = zero;
    for (int i = 0; i < n; ++i) {
      double t = w[i];
      x[i] = t1 * t + x[i];
      w[i] = t2 * t + v[i];
      double tt = t3 * t;
      dknorm = tt * tt + dknorm;
    }
    dknorm = sqrt(dknorm);
    dnorm = d2norm(dnorm, dknorm);
    double delta = sn2 * rho;
    double gambar = -cs2 * rho;
    double rhs = phi - delta * z;
    double zbar = rhs / gambar;
    xnorm = d2norm(xnorm1, zbar);
    double gamma = d2norm(gambar, theta);
    cs2 = gambar / gamma;
    sn2 = theta / gamma;
    z = rhs / gamma;
    xnorm1 = d2norm(xnorm1, z);
    acond = anorm * dnorm;
    res2 = d2norm(res2, psi);
    rnorm = d2norm(res2, phibar);
    arnorm = alpha * fabs(tau);
    double test1 = rnorm / bnorm;
    double test2 = zero;
    if (rnorm > zero) test2 = arnorm / (anorm * rnorm);
    double test3 = one / acond;
    t1 = test1 / (one + anorm * xnorm / bnorm);
    double rtol = btol + atol * anorm * xnorm / bnorm;
    t3 = one + test3;
    t2 = one + test2;
    t1 = one + t1;
    if (itn >= itnlim) istop = 5;
    if (t3 <= one) istop = 4;
    if (t2 <= one) istop = 2;
    if (t1 <= one) istop = 1;
    if (test3 <= ctol) istop = 4;
    if (test2 <= atol) istop = 2;
    if (test1 <= rtol) istop = 1;
    if (istop != 0) break;
  }
  *itn_out = itn;
  return istop;
}

int or_slvlsqr2(int m, int n, const int *ia, const int *ja, const double *a, const double *q, const double *b,
                double *x, int maxit, float eps4, int *nit) {
  double *x0 = (double *)malloc(sizeof(double) * n), *wm = (double *)calloc(m, sizeof(double));
  double *u = (double *)malloc(sizeof(double) * m), *v = (double *)malloc(sizeof(double) * n);
  double *w = (double *)malloc(sizeof(double) * n);
  double conlim = 1.0e8, eps = (double)eps4;
  if (maxit == 0) maxit = 10 * (m + n + 50);
  memcpy(x0, x, sizeof(double) * n);
  for (int i = 0; i < n; ++i) x[i] = x[i] / q[i];
  matvec1(m, ia, ja, a, x, wm);
  memcpy(u, b, sizeof(double) * m);
  for (int i = 0; i < m; ++i) u[i] = u[i] + (-1.0) * wm[i];
  double nrm_r = or_dnrm2(m, u), nrm_b = or_dnrm2(m, b);
  double atol = eps, btol = eps * nrm_b / nrm_r;
  int istop = lsqr(m, n, ia, ja, a, u, x, v, w, atol, btol, conlim, maxit, nit);
  for (int i = 0; i < n; ++i) x[i] = q[i] * x[i];
  for (int i = 0; i < n; ++i) x[i] = x[i] + 1.0 * x0[i];
  free(x0);
  free(wm);
  free(u);
  free(v);
  free(w);
  return istop;
}

/* idamax: first index of the largest |dx(i)| (1-based) */
static int idamax(int n, const double *dx) {
  if (n < 1) return 0;
  int im = 1;
  double dmax = fabs(dx[0]);
  for (int i = 2; i <= n; ++i)
    if (fabs(dx[i - 1]) > dmax) {
      im = i;
      dmax = fabs(dx[i - 1]);
    }
  return im;
}

int or_dgefa(double *a, int lda, int n, double *ipvt) {
#define A(i, j) a[((i)-1) + (size_t)lda * ((j)-1)]
  int info = 0;
  for (int k = 1; k <= n - 1; ++k) {
    int l = idamax(n - k + 1, &A(k, k)) + k - 1;
    ipvt[k - 1] = (double)l;
    if (A(l, k) == 0.0) {
      info = k;
      continue;
    }
    if (l != k) {
      double t = A(l, k);
      A(l, k) = A(k, k);
      A(k, k) = t;
    }
    double t = -1.0 / A(k, k);
    for (int i = k + 1; i <= n; ++i) A(i, k) = t * A(i, k);
    for (int j = k + 1; j <= n; ++j) {
      double tj = A(l, j);
      if (l != k) {
        A(l, j) = A(k, j);
        A(k, j) = tj;
      }
      for (int i = k + 1; i <= n; ++i) A(i, j) = A(i, j) + tj * A(i, k);
    }
  }
  ipvt[n - 1] = (double)n;
  if (A(n, n) == 0.0) info = n;
  return info;
}

void or_dgesl(const double *a, int lda, int n, const double *ipvt, double *b) {
  for (int k = 1; k <= n - 1; ++k) {
    int l = (int)ipvt[k - 1];
    double t = b[l - 1];
    if (l != k) {
      b[l - 1] = b[k - 1];
      b[k - 1] = t;
    }
    for (int i = k + 1; i <= n; ++i) b[i - 1] = b[i - 1] + t * A(i, k);
  }
  for (int kb = 1; kb <= n; ++kb) {
    int k = n + 1 - kb;
    b[k - 1] = b[k - 1] / A(k, k);
    double t = -b[k - 1];
    for (int i = 1; i <= k - 1; ++i) b[i - 1] = b[i - 1] + t * A(i, k);
  }
#undef A
}
