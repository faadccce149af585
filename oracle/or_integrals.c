/*
 * or_integrals.c -- CPU restatement of the reference's integral assembly.
 * TEST INFRASTRUCTURE ONLY (the oracle): see tt_oracle.h.
 *
 * Follows, operation by operation:
 *   GaussPosWeights  lib_int.f:2673-3173 (values computed here, see or_gauss)
 *   cshape27/cshape9/dl34shape4/xieta34  lib_interpol.f:352-405,597-619,571-592,538-568
 *   INTEDc27dcLap    lib_int.f:4-398
 *   INTEBc9kmdclap   lib_int.f:1462-1779,  INTEBc9sdkm lib_sdkm.f:2082-2376
 *   idc2ibc          lib_interpol.f:663-700, FindMinEdge lib_int.f:2450-2488
 *   FMATkmdcSteady   lib_fmat.f:3-446
 *   FMATsdkm         lib_sdkm.f:1286-1530
 *   INTEDc27modHelm, INTEBc9modHelm, FMATmodHelm  lib_mHelm.f:1637-2033, 2034-2348, 2-483
 */
#include "tt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* --------------------------------------------------------------------------
 * Gauss-Legendre nodes/weights.  The reference types the tables by hand
 * (lib_int.f:2680-3170); here they are computed by Newton iteration on the
 * Legendre recurrence in long double, which agrees with the typed values to
 * the last digit the reference prints (checked against numpy.leggauss).
 * ------------------------------------------------------------------------ */
static const int k_rule_n[9] = {2, 4, 6, 8, 16, 32, 40, 48, 64};

int or_gauss(int rule, double *gi, double *ome) {
  if (rule < 1 || rule > 9) return 0;
  int n = k_rule_n[rule - 1];
  for (int i = 0; i < (n + 1) / 2; ++i) {
    long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
    long double pp = 0;
    for (int it = 0; it < 100; ++it) {
      long double p1 = 1, p2 = 0;
      for (int j = 1; j <= n; ++j) {
        long double p3 = p2;
        p2 = p1;
        p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1);
      long double z1 = z;
      z = z1 - p1 / pp;
      if (fabsl(z - z1) < 1e-19L) break;
    }
    long double w = 2 / ((1 - z * z) * pp * pp);
    gi[i] = (double)(-z);
    gi[n - 1 - i] = (double)z;
    ome[i] = ome[n - 1 - i] = (double)w;
  }
  return n;
}

/* rule `rule` of the configuration: from the host's table when given */
static int quad_rule(const or_quad *q, int rule, double *gi, double *ome) {
  if (q->gi && q->ome && q->ng1 && q->ng2) {
    int a = q->ng1[rule - 1], b = q->ng2[rule - 1];
    for (int k = a; k <= b; ++k) {
      gi[k - a] = q->gi[k - 1];
      ome[k - a] = q->ome[k - 1];
    }
    return b - a + 1;
  }
  return or_gauss(rule, gi, ome);
}

/* -------------------------------------------------------------------------- */
void or_shape27(double xi, double eta, double zeta, double *fi) {
  const double ena8 = 1.0 / 8.0, ena4 = 1.0 / 4.0, ena2 = 1.0 / 2.0;
  double epz = 1.0 + zeta, emz = 1.0 - zeta, epe = 1.0 + eta, eme = 1.0 - eta;
  double epx = 1.0 + xi, emx = 1.0 - xi;
  fi[0] = -ena8 * zeta * epz * eta * epe * emx * xi;
  fi[1] = ena8 * emz * zeta * eta * epe * emx * xi;
  fi[2] = -ena8 * emz * zeta * eta * epe * xi * epx;
  fi[3] = ena8 * zeta * epz * eta * epe * xi * epx;
  fi[4] = ena8 * zeta * epz * eme * eta * emx * xi;
  fi[5] = -ena8 * emz * zeta * eme * eta * emx * xi;
  fi[6] = ena8 * emz * zeta * eme * eta * xi * epx;
  fi[7] = -ena8 * zeta * epz * eme * eta * xi * epx;
  fi[8] = -ena4 * emz * epz * eta * epe * emx * xi;
  fi[9] = -ena4 * emz * zeta * eta * epe * emx * epx;
  fi[10] = ena4 * emz * epz * eta * epe * xi * epx;
  fi[11] = ena4 * zeta * epz * eta * epe * emx * epx;
  fi[12] = -ena4 * zeta * epz * eme * epe * emx * xi;
  fi[13] = ena4 * emz * zeta * eme * epe * emx * xi;
  fi[14] = -ena4 * emz * zeta * eme * epe * xi * epx;
  fi[15] = ena4 * zeta * epz * eme * epe * xi * epx;
  fi[16] = ena4 * emz * epz * eme * eta * emx * xi;
  fi[17] = ena4 * emz * zeta * eme * eta * emx * epx;
  fi[18] = -ena4 * emz * epz * eme * eta * xi * epx;
  fi[19] = -ena4 * zeta * epz * eme * eta * emx * epx;
  fi[20] = ena2 * emz * epz * eta * epe * emx * epx;
  fi[21] = -ena2 * emz * epz * eme * epe * emx * xi;
  fi[22] = -ena2 * emz * zeta * eme * epe * emx * epx;
  fi[23] = ena2 * emz * epz * eme * epe * xi * epx;
  fi[24] = ena2 * zeta * epz * eme * epe * emx * epx;
  fi[25] = -ena2 * emz * epz * eme * eta * emx * epx;
  fi[26] = emz * epz * eme * epe * emx * epx;
}

void or_shape9(double xsi, double eta, double *fi) {
  double x2 = xsi * xsi, e2 = eta * eta;
  fi[0] = 0.25 * (xsi - x2) * (eta - e2);
  fi[1] = 0.50 * (1.00 - x2) * (e2 - eta);
  fi[2] = 0.25 * (xsi + x2) * (e2 - eta);
  fi[3] = 0.50 * (xsi + x2) * (1.00 - e2);
  fi[4] = 0.25 * (xsi + x2) * (eta + e2);
  fi[5] = 0.50 * (1.00 - x2) * (e2 + eta);
  fi[6] = 0.25 * (x2 - xsi) * (eta + e2);
  fi[7] = 0.50 * (xsi - x2) * (e2 - 1.00);
  fi[8] = (1.00 - x2) * (1.00 - e2);
}

void or_dl34(double xsi, double eta, double *fi) {
  const double sd = 4.0 / 9.0, tc = 3.0 / 4.0;
  fi[0] = sd * (xsi - tc) * (eta - tc);
  fi[1] = -sd * (xsi + tc) * (eta - tc);
  fi[2] = sd * (xsi + tc) * (eta + tc);
  fi[3] = -sd * (xsi - tc) * (eta + tc);
}

/* Source point positions in [0,1]^3 for isrc 1..51 (INTEDc27dcLap DATA,
 * lib_int.f:50-67): 27 cell nodes then 6 faces x 4 flux points. */
static const double k_ksi51[51] = {
    0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.5, 1.0, 0.5, 0.0, 0.0, 1.0, 1.0, 0.0,
    0.5, 1.0, 0.5, 0.5, 0.0, 0.5, 1.0, 0.5, 0.5, 0.5, 0.125, 0.125, 0.875, 0.875, 0.125, 0.875,
    0.875, 0.125, 1.0, 1.0, 1.0, 1.0, 0.875, 0.125, 0.125, 0.875, 0.0, 0.0, 0.0, 0.0, 0.125,
    0.875, 0.875, 0.125};
static const double k_eta51[51] = {
    1.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 0.5, 0.5, 0.5, 0.5, 0.0,
    0.0, 0.0, 0.0, 1.0, 0.5, 0.5, 0.5, 0.5, 0.0, 0.5, 0.125, 0.875, 0.875, 0.125, 0.0, 0.0,
    0.0, 0.0, 0.125, 0.875, 0.875, 0.125, 1.0, 1.0, 1.0, 1.0, 0.875, 0.125, 0.125, 0.875,
    0.125, 0.125, 0.875, 0.875};
static const double k_zet51[51] = {
    1.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 0.5, 0.0, 0.5, 1.0, 1.0, 0.0, 0.0, 1.0, 0.5,
    0.0, 0.5, 1.0, 0.5, 0.5, 0.0, 0.5, 1.0, 0.5, 0.5, 0.0, 0.0, 0.0, 0.0, 0.125, 0.125,
    0.875, 0.875, 0.125, 0.125, 0.875, 0.875, 0.125, 0.125, 0.875, 0.875, 0.125, 0.125,
    0.875, 0.875, 1.0, 1.0, 1.0, 1.0};

/* face source point positions in [0,1]^2 (INTEBc9kmdclap DATA, lib_int.f:1490-1491) */
static const double k_ksi13[13] = {0.0, 0.5, 1.0, 1.0, 1.0, 0.5, 0.0, 0.0, 0.5, 0.125, 0.875, 0.875, 0.125};
static const double k_eta13[13] = {0.0, 0.0, 0.0, 0.5, 1.0, 1.0, 1.0, 0.5, 0.5, 0.125, 0.125, 0.875, 0.875};

/* Kernel factors at distance ra: fung = u*, gfac multiplies the gradient.
 * Laplace (mu = 0): u* = 1/(4 pi r).  Modified Helmholtz (INTEDc27modHelm
 * lib_mHelm.f:1905-1910, INTEBc9modHelm :2227-2233): u* = exp(-mu r)/(4 pi r)
 * and the gradient carries (1 + mu r).  With mu = 0 both factors are exactly
 * 1, so the two kernels coincide bit for bit. */
static inline double kern_g(double PI, double ra, double mu) {
  return mu != 0.0 ? 0.25 / (PI * ra) * exp(-mu * ra) : 0.25 / (PI * ra);
}
static inline double kern_f(double ra, double mu) { return mu != 0.0 ? 1 + mu * ra : 1.0; }

static long inted27_k(const or_quad *q, const double p[3], const double c[8][3], int isrc, double mu,
                      double *be, double *dxe, double *dye, double *dze) {
  double gis[64], oms[64], gir[64], omr[64];
  int ns = quad_rule(q, q->kmDs, gis, oms);
  int nr = quad_rule(q, q->kmDr, gir, omr);
  const double PI = 2.0 * asin(1.0), PI2 = PI * 2.0;
  long npts = 0;
  for (int k = 0; k < 27; ++k) be[k] = dxe[k] = dye[k] = dze[k] = 0.0;
  int nd = isrc != 0 ? q->iiDiv : 1;
  for (int ix = 1; ix <= nd; ++ix)
    for (int iy = 1; iy <= nd; ++iy)
      for (int iz = 1; iz <= nd; ++iz) {
        double a = -1.0 + (ix - 1) * 2.0 / nd, b = -1.0 + (ix)*2.0 / nd;
        double cc = -1.0 + (iy - 1) * 2.0 / nd, d = -1.0 + (iy)*2.0 / nd;
        double e = -1.0 + (iz - 1) * 2.0 / nd, f = -1.0 + (iz)*2.0 / nd;
        double dex = 0.125 * (b - a) * (d - cc) * (f - e);
        int ising = 0;
        double xx1 = 0, yy1 = 0, zz1 = 0;
        if (isrc != 0) {
          xx1 = k_ksi51[isrc - 1];
          xx1 = -1.0 + 2.0 * xx1;
          xx1 = (xx1 - a) / (b - a);
          yy1 = k_eta51[isrc - 1];
          yy1 = -1.0 + 2.0 * yy1;
          yy1 = (yy1 - cc) / (d - cc);
          zz1 = k_zet51[isrc - 1];
          zz1 = -1.0 + 2.0 * zz1;
          zz1 = (zz1 - e) / (f - e);
          ising = (xx1 >= 0.0 && xx1 <= 1.0 && yy1 >= 0.0 && yy1 <= 1.0 && zz1 >= 0.0 && zz1 <= 1.0);
        }
        if (ising) {
          for (int L = 1; L <= 4; ++L) {
            double aa0, gama, ffi0, ffi1;
            if (L == 1) {
              if (1.0 - xx1 == 0.0) continue;
              aa0 = sqrt(yy1 * yy1 + (1.0 - xx1) * (1.0 - xx1));
              gama = acos(yy1 / aa0);
              ffi0 = 1.5 * PI + acos(yy1 / aa0);
              ffi1 = PI2 + atan((1.0 - yy1) / (1.0 - xx1));
            } else if (L == 2) {
              if (1.0 - yy1 == 0.0) continue;
              aa0 = sqrt((1.0 - yy1) * (1.0 - yy1) + (1.0 - xx1) * (1.0 - xx1));
              gama = acos((1.0 - xx1) / aa0);
              ffi0 = asin((1.0 - yy1) / aa0);
              ffi1 = 0.5 * PI + atan(xx1 / (1.0 - yy1));
            } else if (L == 3) {
              if (xx1 == 0.0) continue;
              aa0 = sqrt(xx1 * xx1 + (1.0 - yy1) * (1.0 - yy1));
              gama = acos((1.0 - yy1) / aa0);
              ffi0 = 0.5 * PI + acos((1.0 - yy1) / aa0);
              ffi1 = PI + atan(yy1 / xx1);
            } else {
              if (yy1 == 0.0) continue;
              aa0 = sqrt(yy1 * yy1 + xx1 * xx1);
              gama = acos(xx1 / aa0);
              ffi0 = PI + acos(xx1 / aa0);
              ffi1 = 1.5 * PI + atan((1.0 - xx1) / yy1);
            }
            for (int K = 0; K < ns; ++K) {
              double ffi = (ffi1 - ffi0) * gis[K] / 2.0 + (ffi1 + ffi0) / 2.0;
              double ro0 = aa0 * sin(gama) / sin(ffi - ffi0 + gama);
              for (int N = 1; N <= 3; ++N) {
                double aa1, eps, th0, th1, rr1;
                if (N == 1) {
                  aa1 = 1.0 - zz1;
                  eps = PI / 2.0;
                  th0 = 0.0;
                  th1 = acos(aa1 / sqrt(aa1 * aa1 + ro0 * ro0));
                } else if (N == 2) {
                  aa1 = sqrt(ro0 * ro0 + (1.0 - zz1) * (1.0 - zz1));
                  eps = acos((1.0 - zz1) / aa1);
                  th0 = eps;
                  rr1 = fmin(1.0, ro0 / sqrt(zz1 * zz1 + ro0 * ro0));
                  th1 = 0.5 * PI + acos(rr1);
                } else {
                  aa1 = sqrt(ro0 * ro0 + zz1 * zz1);
                  rr1 = fmin(1.0, ro0 / aa1);
                  eps = acos(rr1);
                  th0 = 0.5 * PI + eps;
                  th1 = PI;
                }
                for (int I = 0; I < ns; ++I) {
                  double th = (th1 - th0) * gis[I] / 2.0 + (th1 + th0) / 2.0;
                  double ro1 = aa1 * sin(eps) / sin(th - th0 + eps);
                  for (int J = 0; J < ns; ++J) {
                    double ro = (ro1 * gis[J] + ro1) / 2.0;
                    double eti = 2.0 * (xx1 + ro * sin(th) * cos(ffi)) - 1.0;
                    double etj = 2.0 * (yy1 + ro * sin(th) * sin(ffi)) - 1.0;
                    double etk = 2.0 * (zz1 + ro * cos(th)) - 1.0;
                    eti = a + 0.5 * (eti + 1.0) * (b - a);
                    etj = cc + 0.5 * (etj + 1.0) * (d - cc);
                    etk = e + 0.5 * (etk + 1.0) * (f - e);
                    /* weights as AJAC=1.953125D-3*AJAC*OME(I)*OME(J)*OME(K)*(ffi1-ffi0)*
                     * (th1-th0)*ro1*ro**2*SIN(th); AJAC=dex*AJAC (lib_int.f:249-250) */
                    {
                      double e1m = 1.0 - eti, e1p = 1.0 + eti, e2m = 1.0 - etj, e2p = 1.0 + etj,
                             e3m = 1.0 - etk, e3p = 1.0 + etk;
                      double fig[8];
                      fig[0] = 0.125 * e1m * e2m * e3m;
                      fig[1] = 0.125 * e1p * e2m * e3m;
                      fig[2] = 0.125 * e1m * e2p * e3m;
                      fig[3] = 0.125 * e1p * e2p * e3m;
                      fig[4] = 0.125 * e1m * e2m * e3p;
                      fig[5] = 0.125 * e1p * e2m * e3p;
                      fig[6] = 0.125 * e1m * e2p * e3p;
                      fig[7] = 0.125 * e1p * e2p * e3p;
                      double xc[3];
                      for (int dd = 0; dd < 3; ++dd)
                        xc[dd] = fig[0] * c[0][dd] + fig[1] * c[1][dd] + fig[2] * c[2][dd] +
                                 fig[3] * c[3][dd] + fig[4] * c[4][dd] + fig[5] * c[5][dd] +
                                 fig[6] * c[6][dd] + fig[7] * c[7][dd];
                      double aj[3][3];
                      for (int dd = 0; dd < 3; ++dd) {
                        aj[0][dd] = -e2m * e3m * c[0][dd] + e2m * e3m * c[1][dd] - e2p * e3m * c[2][dd] +
                                    e2p * e3m * c[3][dd] - e2m * e3p * c[4][dd] + e2m * e3p * c[5][dd] -
                                    e2p * e3p * c[6][dd] + e2p * e3p * c[7][dd];
                        aj[1][dd] = -e1m * e3m * c[0][dd] - e1p * e3m * c[1][dd] + e1m * e3m * c[2][dd] +
                                    e1p * e3m * c[3][dd] - e1m * e3p * c[4][dd] - e1p * e3p * c[5][dd] +
                                    e1m * e3p * c[6][dd] + e1p * e3p * c[7][dd];
                        aj[2][dd] = -e1m * e2m * c[0][dd] - e1p * e2m * c[1][dd] - e1m * e2p * c[2][dd] -
                                    e1p * e2p * c[3][dd] + e1m * e2m * c[4][dd] + e1p * e2m * c[5][dd] +
                                    e1m * e2p * c[6][dd] + e1p * e2p * c[7][dd];
                      }
                      double ajac = (aj[0][0] * aj[1][1] * aj[2][2] + aj[0][1] * aj[1][2] * aj[2][0] +
                                     aj[0][2] * aj[1][0] * aj[2][1] - aj[0][0] * aj[1][2] * aj[2][1] -
                                     aj[0][1] * aj[1][0] * aj[2][2] - aj[0][2] * aj[1][1] * aj[2][0]);
                      ajac = 1.953125e-3 * ajac * oms[I] * oms[J] * oms[K] * (ffi1 - ffi0) * (th1 - th0) *
                             ro1 * (ro * ro) * sin(th);
                      ajac = dex * ajac;
                      double rx = p[0] - xc[0], ry = p[1] - xc[1], rz = p[2] - xc[2];
                      double ra = sqrt(rx * rx + ry * ry + rz * rz);
                      ++npts;
                      if (ra == 0.0) continue;
                      if (ra < 1.0e-12) ra = 1.0e-12;
                      double ira2 = 1.0 / (ra * ra);
                      double fung = kern_g(PI, ra, mu), gf = kern_f(ra, mu);
                      double funx = (p[0] - xc[0]) * ira2 * fung * gf;
                      double funy = (p[1] - xc[1]) * ira2 * fung * gf;
                      double funz = (p[2] - xc[2]) * ira2 * fung * gf;
                      double f27[27];
                      or_shape27(eti, etj, etk, f27);
                      for (int kk = 0; kk < 27; ++kk) {
                        double dom = ajac * f27[kk];
                        dxe[kk] = dxe[kk] + funx * dom;
                        dye[kk] = dye[kk] + funy * dom;
                        dze[kk] = dze[kk] + funz * dom;
                        be[kk] = be[kk] + fung * dom;
                      }
                    }
                  }
                }
              }
            }
          }
        } else {
          for (int K = 0; K < nr; ++K) {
            double gik = e + 0.5 * (gir[K] + 1.0) * (f - e);
            for (int J = 0; J < nr; ++J) {
              double gij = cc + 0.5 * (gir[J] + 1.0) * (d - cc);
              for (int I = 0; I < nr; ++I) {
                double gii = a + 0.5 * (gir[I] + 1.0) * (b - a);
                /* AJAC=1.953125D-3*AJAC*OME(I)*OME(J)*OME(K)*dex (lib_int.f:353) */
                double e1m = 1.0 - gii, e1p = 1.0 + gii, e2m = 1.0 - gij, e2p = 1.0 + gij,
                       e3m = 1.0 - gik, e3p = 1.0 + gik;
                double fig[8];
                fig[0] = 0.125 * e1m * e2m * e3m;
                fig[1] = 0.125 * e1p * e2m * e3m;
                fig[2] = 0.125 * e1m * e2p * e3m;
                fig[3] = 0.125 * e1p * e2p * e3m;
                fig[4] = 0.125 * e1m * e2m * e3p;
                fig[5] = 0.125 * e1p * e2m * e3p;
                fig[6] = 0.125 * e1m * e2p * e3p;
                fig[7] = 0.125 * e1p * e2p * e3p;
                double xc[3];
                for (int dd = 0; dd < 3; ++dd)
                  xc[dd] = fig[0] * c[0][dd] + fig[1] * c[1][dd] + fig[2] * c[2][dd] + fig[3] * c[3][dd] +
                           fig[4] * c[4][dd] + fig[5] * c[5][dd] + fig[6] * c[6][dd] + fig[7] * c[7][dd];
                double aj[3][3];
                for (int dd = 0; dd < 3; ++dd) {
                  aj[0][dd] = -e2m * e3m * c[0][dd] + e2m * e3m * c[1][dd] - e2p * e3m * c[2][dd] +
                              e2p * e3m * c[3][dd] - e2m * e3p * c[4][dd] + e2m * e3p * c[5][dd] -
                              e2p * e3p * c[6][dd] + e2p * e3p * c[7][dd];
                  aj[1][dd] = -e1m * e3m * c[0][dd] - e1p * e3m * c[1][dd] + e1m * e3m * c[2][dd] +
                              e1p * e3m * c[3][dd] - e1m * e3p * c[4][dd] - e1p * e3p * c[5][dd] +
                              e1m * e3p * c[6][dd] + e1p * e3p * c[7][dd];
                  aj[2][dd] = -e1m * e2m * c[0][dd] - e1p * e2m * c[1][dd] - e1m * e2p * c[2][dd] -
                              e1p * e2p * c[3][dd] + e1m * e2m * c[4][dd] + e1p * e2m * c[5][dd] +
                              e1m * e2p * c[6][dd] + e1p * e2p * c[7][dd];
                }
                double ajac = (aj[0][0] * aj[1][1] * aj[2][2] + aj[0][1] * aj[1][2] * aj[2][0] +
                               aj[0][2] * aj[1][0] * aj[2][1] - aj[0][0] * aj[1][2] * aj[2][1] -
                               aj[0][1] * aj[1][0] * aj[2][2] - aj[0][2] * aj[1][1] * aj[2][0]);
                ajac = 1.953125e-3 * ajac * omr[I] * omr[J] * omr[K] * dex;
                double rx = p[0] - xc[0], ry = p[1] - xc[1], rz = p[2] - xc[2];
                double ra = sqrt(rx * rx + ry * ry + rz * rz);
                ++npts;
                if (ra < 1.0e-12) ra = 1.0e-12;
                double ira2 = 1.0 / (ra * ra);
                double fung = kern_g(PI, ra, mu), gf = kern_f(ra, mu);
                double funx = (p[0] - xc[0]) * ira2 * fung * gf;
                double funy = (p[1] - xc[1]) * ira2 * fung * gf;
                double funz = (p[2] - xc[2]) * ira2 * fung * gf;
                double f27[27];
                or_shape27(gii, gij, gik, f27);
                for (int kk = 0; kk < 27; ++kk) {
                  double dom = ajac * f27[kk];
                  dxe[kk] = dxe[kk] + funx * dom;
                  dye[kk] = dye[kk] + funy * dom;
                  dze[kk] = dze[kk] + funz * dom;
                  be[kk] = be[kk] + fung * dom;
                }
              }
            }
          }
        }
      }
  return npts;
}

long or_inted27(const or_quad *q, const double p[3], const double c[8][3], int isrc, double *be,
                double *dxe, double *dye, double *dze) {
  return inted27_k(q, p, c, isrc, 0.0, be, dxe, dye, dze);
}

long or_inted27_mh(const or_quad *q, const double p[3], const double c[8][3], int isrc, double mu,
                   double *be, double *dxe, double *dye, double *dze) {
  return inted27_k(q, p, c, isrc, mu, be, dxe, dye, dze);
}

/* one boundary-face kernel evaluation (shared by the singular and regular
 * branches of INTEBc9kmdclap; the two only differ in how eti/etj/weight are
 * formed) */
static inline void bnd_point(int macro, const double p[3], const double c[4][3], double eti, double etj,
                             double wpre, double mu, double *he, double *ge, double *adx, double *ady,
                             double *adz, double *heyz, double *hezx, double *hexy) {
  const double PI = 2.0 * asin(1.0);
  double e1m = 1.0 - eti, e1p = 1.0 + eti, e2m = 1.0 - etj, e2p = 1.0 + etj;
  double f4[4] = {0.25 * e1m * e2m, 0.25 * e1p * e2m, 0.25 * e1p * e2p, 0.25 * e1m * e2p};
  double xc0 = f4[0] * c[0][0] + f4[1] * c[1][0] + f4[2] * c[2][0] + f4[3] * c[3][0];
  double yc0 = f4[0] * c[0][1] + f4[1] * c[1][1] + f4[2] * c[2][1] + f4[3] * c[3][1];
  double zc0 = f4[0] * c[0][2] + f4[1] * c[1][2] + f4[2] * c[2][2] + f4[3] * c[3][2];
  double xks = 0.25 * (-e2m * c[0][0] + e2m * c[1][0] + e2p * c[2][0] - e2p * c[3][0]);
  double yks = 0.25 * (-e2m * c[0][1] + e2m * c[1][1] + e2p * c[2][1] - e2p * c[3][1]);
  double zks = 0.25 * (-e2m * c[0][2] + e2m * c[1][2] + e2p * c[2][2] - e2p * c[3][2]);
  double xet = 0.25 * (-e1m * c[0][0] - e1p * c[1][0] + e1p * c[2][0] + e1m * c[3][0]);
  double yet = 0.25 * (-e1m * c[0][1] - e1p * c[1][1] + e1p * c[2][1] + e1m * c[3][1]);
  double zet = 0.25 * (-e1m * c[0][2] - e1p * c[1][2] + e1p * c[2][2] + e1m * c[3][2]);
  double anx = yks * zet - yet * zks;
  double any = xet * zks - xks * zet;
  double anz = xks * yet - xet * yks;
  double ajac = 1.0 / sqrt(anx * anx + any * any + anz * anz);
  double anx1 = anx * ajac, any1 = any * ajac, anz1 = anz * ajac;
  ajac = wpre / ajac;
  double rx = p[0] - xc0, ry = p[1] - yc0, rz = p[2] - zc0;
  double ra = sqrt(rx * rx + ry * ry + rz * rz);
  double ira2 = 1.0 / (ra * ra);
  double fung = kern_g(PI, ra, mu), gf = kern_f(ra, mu);
  double funx = (p[0] - xc0) * ira2 * fung * gf;
  double funy = (p[1] - yc0) * ira2 * fung * gf;
  double funz = (p[2] - zc0) * ira2 * fung * gf;
  double funh = funx * anx1 + funy * any1 + funz * anz1;
  double f9[9];
  or_shape9(eti, etj, f9);
  for (int s = 0; s < 9; ++s) {
    double dg = ajac * f9[s];
    he[s] = he[s] + funh * dg;
    if (macro) {
      adx[s] = adx[s] + fung * anx1 * dg;
      ady[s] = ady[s] + fung * any1 * dg;
      adz[s] = adz[s] + fung * anz1 * dg;
    }
    hexy[s] = hexy[s] - (funx * any1 - funy * anx1) * dg;
    heyz[s] = heyz[s] - (funy * anz1 - funz * any1) * dg;
    hezx[s] = hezx[s] - (funz * anx1 - funx * anz1) * dg;
  }
  if (macro) {
    double fg[4];
    or_dl34(eti, etj, fg);
    for (int s = 0; s < 4; ++s) {
      double dg = ajac * fg[s];
      ge[s] = ge[s] + fung * dg;
    }
  }
}

static long inteb9_k(const or_quad *q, int macro, const double p[3], const double c[4][3], int isrc,
                     double minedge, double mu, double *he, double *ge, double *adx, double *ady, double *adz,
                     double *heyz, double *hezx, double *hexy) {
  double gis[64], oms[64], gir[64], omr[64];
  int ns = quad_rule(q, q->kmBs, gis, oms);
  int nr = quad_rule(q, q->kmBr, gir, omr);
  const double PI = 2.0 * asin(1.0), PI2 = 2.0 * PI;
  double dadx[9], dady[9], dadz[9], dge[4];
  if (!macro) { adx = dadx; ady = dady; adz = dadz; ge = dge; }
  for (int s = 0; s < 9; ++s) he[s] = adx[s] = ady[s] = adz[s] = hexy[s] = heyz[s] = hezx[s] = 0.0;
  for (int s = 0; s < 4; ++s) ge[s] = 0.0;
  long npts = 0;
  int ndx = 1, ndy = 1;
  if (macro) {
    double d1 = sqrt((c[0][0] - c[1][0]) * (c[0][0] - c[1][0]) + (c[0][1] - c[1][1]) * (c[0][1] - c[1][1]) +
                     (c[0][2] - c[1][2]) * (c[0][2] - c[1][2]));
    double d2 = sqrt((c[1][0] - c[2][0]) * (c[1][0] - c[2][0]) + (c[1][1] - c[2][1]) * (c[1][1] - c[2][1]) +
                     (c[1][2] - c[2][2]) * (c[1][2] - c[2][2]));
    double d3 = sqrt((c[2][0] - c[3][0]) * (c[2][0] - c[3][0]) + (c[2][1] - c[3][1]) * (c[2][1] - c[3][1]) +
                     (c[2][2] - c[3][2]) * (c[2][2] - c[3][2]));
    double d4 = sqrt((c[3][0] - c[0][0]) * (c[3][0] - c[0][0]) + (c[3][1] - c[0][1]) * (c[3][1] - c[0][1]) +
                     (c[3][2] - c[0][2]) * (c[3][2] - c[0][2]));
    double d13 = fmax(d1, d3), d24 = fmax(d2, d4);
    int t1 = (int)(d13 / minedge + 0.5), t2 = (int)(d24 / minedge + 0.5);
    ndx = q->iiDiv > t1 ? q->iiDiv : t1;
    ndy = q->iiDiv > t2 ? q->iiDiv : t2;
  }
  for (int ix = 1; ix <= ndx; ++ix)
    for (int iy = 1; iy <= ndy; ++iy) {
      double a = -1.0 + (ix - 1) * 2.0 / ndx, b = -1.0 + (ix)*2.0 / ndx;
      double cc = -1.0 + (iy - 1) * 2.0 / ndy, d = -1.0 + (iy)*2.0 / ndy;
      double dex = 0.25 * (b - a) * (d - cc);
      int ising = 0;
      double xx1 = 0, yy1 = 0;
      if (isrc != 0) {
        xx1 = k_ksi13[isrc - 1];
        xx1 = -1.0 + 2.0 * xx1;
        xx1 = (xx1 - a) / (b - a);
        yy1 = k_eta13[isrc - 1];
        yy1 = -1.0 + 2.0 * yy1;
        yy1 = (yy1 - cc) / (d - cc);
        ising = (xx1 >= 0.0 && xx1 <= 1.0 && yy1 >= 0.0 && yy1 <= 1.0);
      }
      if (ising) {
        for (int K = 1; K <= 4; ++K) {
          double al, fii, th0, th1;
          if (K == 1) {
            if (1.0 - xx1 == 0.0) continue;
            al = sqrt(yy1 * yy1 + (1.0 - xx1) * (1.0 - xx1));
            fii = acos(yy1 / al);
            th0 = 1.5 * PI + acos(yy1 / al);
            th1 = PI2 + atan((1.0 - yy1) / (1.0 - xx1));
          } else if (K == 2) {
            if (1.0 - yy1 == 0.0) continue;
            al = sqrt((1.0 - yy1) * (1.0 - yy1) + (1.0 - xx1) * (1.0 - xx1));
            fii = acos((1.0 - xx1) / al);
            th0 = asin((1.0 - yy1) / al);
            th1 = 0.5 * PI + atan(xx1 / (1.0 - yy1));
          } else if (K == 3) {
            if (xx1 == 0.0) continue;
            al = sqrt(xx1 * xx1 + (1.0 - yy1) * (1.0 - yy1));
            fii = acos((1.0 - yy1) / al);
            th0 = 0.5 * PI + acos((1.0 - yy1) / al);
            th1 = PI + atan(yy1 / xx1);
          } else {
            if (yy1 == 0.0) continue;
            al = sqrt(yy1 * yy1 + xx1 * xx1);
            fii = acos(xx1 / al);
            th0 = PI + acos(xx1 / al);
            th1 = 1.5 * PI + atan((1.0 - xx1) / yy1);
          }
          for (int I = 0; I < ns; ++I)
            for (int J = 0; J < ns; ++J) {
              double th = (th1 - th0) * gis[I] / 2.0 + (th1 + th0) / 2.0;
              double roth = al * sin(fii) / sin(th - th0 + fii);
              double ro = roth * gis[J] / 2.0 + roth / 2.0;
              double eti = 2.0 * (xx1 + ro * cos(th)) - 1.0;
              double etj = 2.0 * (yy1 + ro * sin(th)) - 1.0;
              eti = a + 0.5 * (eti + 1.0) * (b - a);
              etj = cc + 0.5 * (etj + 1.0) * (d - cc);
              /* AJAC=dex*(TH1-TH0)*ROTH*RO*OME(I)*OME(J)/AJAC (lib_int.f:1645) */
              double wpre = dex * (th1 - th0) * roth * ro * oms[I] * oms[J];
              bnd_point(macro, p, c, eti, etj, wpre, mu, he, ge, adx, ady, adz, heyz, hezx, hexy);
              ++npts;
            }
        }
      } else {
        for (int I = 0; I < nr; ++I) {
          double gii = a + 0.5 * (gir[I] + 1.0) * (b - a);
          for (int J = 0; J < nr; ++J) {
            double gij = cc + 0.5 * (gir[J] + 1.0) * (d - cc);
            /* AJAC=dex*OME(I)*OME(J)/AJAC (lib_int.f:1733) */
            double wpre = dex * omr[I] * omr[J];
            bnd_point(macro, p, c, gii, gij, wpre, mu, he, ge, adx, ady, adz, heyz, hezx, hexy);
            ++npts;
          }
        }
      }
    }
  return npts;
}

long or_inteb9(const or_quad *q, int macro, const double p[3], const double c[4][3], int isrc,
               double minedge, double *he, double *ge, double *adx, double *ady, double *adz,
               double *heyz, double *hezx, double *hexy) {
  return inteb9_k(q, macro, p, c, isrc, minedge, 0.0, he, ge, adx, ady, adz, heyz, hezx, hexy);
}

long or_inteb9_mh(const or_quad *q, const double p[3], const double c[4][3], int isrc, double minedge,
                  double mu, double *he, double *ge, double *adx, double *ady, double *adz, double *heyz,
                  double *hezx, double *hexy) {
  return inteb9_k(q, 1, p, c, isrc, minedge, mu, he, ge, adx, ady, adz, heyz, hezx, hexy);
}

/* idc2ibc local permutations (lib_interpol.f:683-688), 1-based local nodes */
static const int k_face9[6][9] = {{6, 14, 2, 10, 3, 15, 7, 18, 23}, {6, 18, 7, 19, 8, 20, 5, 17, 26},
                                  {7, 15, 3, 11, 4, 16, 8, 19, 24}, {3, 10, 2, 9, 1, 12, 4, 11, 21},
                                  {2, 14, 6, 17, 5, 13, 1, 9, 22},  {5, 20, 8, 16, 4, 12, 1, 13, 25}};
/* cell geometry corners, reference order x1..x8 = idc(6,7,2,3,5,8,1,4) */
static const int k_geo8[8] = {6, 7, 2, 3, 5, 8, 1, 4};

#define IDC(ic, k) (idc[(ic) + (size_t)nc * ((k)-1)])
#define XC(n, d) (x[((n)-1) + (size_t)nn * (d)])

static void cell_corners(int nn, int nc, const double *x, const int *idc, int ic, double c[8][3]) {
  for (int k = 0; k < 8; ++k)
    for (int d = 0; d < 3; ++d) c[k][d] = XC(IDC(ic, k_geo8[k]), d);
}

static int find_col(int nc, const int *idc, int ic, int node) {
  for (int k = 1; k <= 27; ++k)
    if (IDC(ic, k) == node) return k;
  return -1;
}

/* kernel evaluations of the last macro assembly: domain, boundary */
static long g_points[2];

void or_last_points(long *domain, long *boundary) {
  *domain = g_points[0];
  *boundary = g_points[1];
}

/* FMATkmdcSteady (mun == NULL) and FMATmodHelm (mun/muq: mu of every node /
 * flux node as source).  The two differ in the kernel, in the free term
 * (c = -sum H + mu^2 sum B, lib_mHelm.f:343-357) and in the linear-field
 * correction ((H - mu^2 B) x, :421-443); FMATmodHelm applies its corrections
 * after the domain integrals (:318-458), FMATkmdcSteady right after the
 * boundary integrals -- the results do not depend on that order. */
static void macro_k(const or_quad *q, int nn, int nc, int nq, const double *x, const int *idc,
                    const int *ibf, const double *xq, const int *flpm, int copy, const double *mun,
                    const double *muq, double *H, double *G, double *B, double *Abdx, double *Abdy,
                    double *Abdz, double *Htx, double *Hty, double *Htz, double *Dx, double *Dy, double *Dz,
                    double ont[11], double *mumax, int nthreads) {
  const int helm = mun != NULL;
  const size_t m1 = (size_t)nc * 51, m2 = (size_t)nc * 27;
#define M1(A, r, c) A[(r) + m1 * (c)]
#define M2(A, r, c) A[(r) + m2 * (c)]
  memset(H, 0, sizeof(double) * m1 * 27);
  memset(G, 0, sizeof(double) * m1 * 24);
  memset(B, 0, sizeof(double) * m1 * 27);
  memset(Abdx, 0, sizeof(double) * m1 * 27);
  memset(Abdy, 0, sizeof(double) * m1 * 27);
  memset(Abdz, 0, sizeof(double) * m1 * 27);
  memset(Dx, 0, sizeof(double) * m1 * 27);
  memset(Dy, 0, sizeof(double) * m1 * 27);
  memset(Dz, 0, sizeof(double) * m1 * 27);
  memset(Htx, 0, sizeof(double) * m2 * 27);
  memset(Hty, 0, sizeof(double) * m2 * 27);
  memset(Htz, 0, sizeof(double) * m2 * 27);
  int ncomp = copy ? 1 : nc;
  /* per-cell accuracy estimates: [ontH, ontG, Htx, Hty, Htz, Dx, Dy, Dz, HtyDx, HtzDy, HtxDz] */
  double *ontc = (double *)calloc((size_t)ncomp * 12, sizeof(double));
  long *cpts = (long *)calloc((size_t)ncomp * 2, sizeof(long));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int ic = 0; ic < ncomp; ++ic) {
    double *oc = ontc + (size_t)ic * 12;
    /* ibc for this cell and the shortest edge */
    int ibcl[6][9];
    for (int s = 0; s < 6; ++s)
      for (int k = 0; k < 9; ++k) ibcl[s][k] = IDC(ic, k_face9[s][k]);
    double minedge = 1.0e20;
    for (int s = 0; s < 6; ++s) {
      int cn[5] = {ibcl[s][0], ibcl[s][2], ibcl[s][4], ibcl[s][6], ibcl[s][0]};
      for (int e = 0; e < 4; ++e) {
        double dx = XC(cn[e], 0) - XC(cn[e + 1], 0), dy = XC(cn[e], 1) - XC(cn[e + 1], 1),
               dz = XC(cn[e], 2) - XC(cn[e + 1], 2);
        double el = sqrt(dx * dx + dy * dy + dz * dz);
        if (el < minedge) minedge = el;
      }
    }
    double he[9], ge[4], abx[9], aby[9], abz[9], htx[9], hty[9], htz[9];
    /* ---- boundary integrals, u sources (lib_fmat.f:137-215) ---- */
    for (int it = 1; it <= 27; ++it) {
      int dn = IDC(ic, it);
      double p[3] = {XC(dn, 0), XC(dn, 1), XC(dn, 2)};
      size_t i = (size_t)ic * 51 + (it - 1), ii = (size_t)ic * 27 + (it - 1);
      double mu = helm ? mun[dn - 1] : 0.0;
      if (mu > oc[11]) oc[11] = mu;
      for (int je = 0; je < 6; ++je) {
        double cf[4][3];
        for (int k = 0; k < 4; ++k)
          for (int d = 0; d < 3; ++d) cf[k][d] = XC(ibcl[je][2 * k], d);
        int isrc = 0;
        for (int s = 0; s < 9; ++s)
          if (dn == ibcl[je][s]) isrc = s + 1;
        cpts[2 * ic + 1] += inteb9_k(q, 1, p, cf, isrc, minedge, mu, he, ge, abx, aby, abz, htx, hty, htz);
        for (int s = 0; s < 9; ++s) {
          int col = find_col(nc, idc, ic, ibcl[je][s]) - 1;
          M1(H, i, col) = M1(H, i, col) + he[s];
          M1(Abdx, i, col) = M1(Abdx, i, col) + abx[s];
          M1(Abdy, i, col) = M1(Abdy, i, col) + aby[s];
          M1(Abdz, i, col) = M1(Abdz, i, col) + abz[s];
          M2(Htx, ii, col) = M2(Htx, ii, col) - htx[s];
          M2(Hty, ii, col) = M2(Hty, ii, col) - hty[s];
          M2(Htz, ii, col) = M2(Htz, ii, col) - htz[s];
        }
        for (int s = 0; s < 4; ++s) {
          int col = je * 4 + s;
          M1(G, i, col) = M1(G, i, col) + ge[s] * (double)flpm[ic + (size_t)nc * je];
        }
      }
      if (helm) continue; /* free term after the domain integrals */
      double cfree = 0.0;
      for (int j = 0; j < 27; ++j) cfree = cfree - M1(H, i, j);
      int col = find_col(nc, idc, ic, dn) - 1;
      M1(H, i, col) = M1(H, i, col) + cfree;
      double *Ht[3] = {Htx, Hty, Htz};
      for (int t = 0; t < 3; ++t) {
        double vs = 0.0;
        for (int j = 0; j < 27; ++j) vs = vs + M2(Ht[t], ii, j);
        if (fabs(vs) > oc[2 + t]) oc[2 + t] = fabs(vs);
        M2(Ht[t], ii, it - 1) = M2(Ht[t], ii, it - 1) - vs;
      }
      if (it == 27) oc[0] = fabs(1.0 - cfree);
    }
    /* ---- boundary integrals, q sources (lib_fmat.f:219-298) ---- */
    for (int its = 0; its < 6; ++its)
      for (int itf = 1; itf <= 4; ++itf) {
        int dn = ibf[ic + (size_t)nc * (its + 6 * (itf - 1))];
        double p[3] = {xq[(dn - 1)], xq[(dn - 1) + (size_t)nq], xq[(dn - 1) + 2 * (size_t)nq]};
        size_t i = (size_t)ic * 51 + 27 + its * 4 + (itf - 1);
        double mu = helm ? muq[dn - 1] : 0.0;
        if (mu > oc[11]) oc[11] = mu;
        for (int je = 0; je < 6; ++je) {
          double cf[4][3];
          for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) cf[k][d] = XC(ibcl[je][2 * k], d);
          int isrc = (its == je) ? itf + 9 : 0;
          cpts[2 * ic + 1] += inteb9_k(q, 1, p, cf, isrc, minedge, mu, he, ge, abx, aby, abz, htx, hty, htz);
          for (int s = 0; s < 9; ++s) {
            int col = find_col(nc, idc, ic, ibcl[je][s]) - 1;
            M1(H, i, col) = M1(H, i, col) + he[s];
            M1(Abdx, i, col) = M1(Abdx, i, col) + abx[s];
            M1(Abdy, i, col) = M1(Abdy, i, col) + aby[s];
            M1(Abdz, i, col) = M1(Abdz, i, col) + abz[s];
          }
          for (int s = 0; s < 4; ++s) {
            int col = je * 4 + s;
            M1(G, i, col) = M1(G, i, col) + ge[s] * (double)flpm[ic + (size_t)nc * je];
          }
        }
        if (helm) continue;
        double cfree = 0.0;
        for (int j = 0; j < 27; ++j) cfree = cfree - M1(H, i, j);
        oc[1] += fabs(0.5 - cfree);
        const double tc = 3.0 / 4.0;
        double xi = (itf == 1 || itf == 4) ? -tc : tc;
        double eta = (itf <= 2) ? -tc : tc;
        double fi[9];
        or_shape9(xi, eta, fi);
        for (int s = 0; s < 9; ++s) {
          int col = find_col(nc, idc, ic, ibcl[its][s]) - 1;
          M1(H, i, col) = M1(H, i, col) + cfree * fi[s];
        }
      }
    /* ---- domain integrals (lib_fmat.f:303-420) ---- */
    double cc8[8][3];
    cell_corners(nn, nc, x, idc, ic, cc8);
    double be[27], dxe[27], dye[27], dze[27];
    for (int it = 1; it <= 27; ++it) {
      int dn = IDC(ic, it);
      double p[3] = {XC(dn, 0), XC(dn, 1), XC(dn, 2)};
      size_t i = (size_t)ic * 51 + (it - 1), ii = (size_t)ic * 27 + (it - 1);
      cpts[2 * ic] += inted27_k(q, p, cc8, it, helm ? mun[dn - 1] : 0.0, be, dxe, dye, dze);
      for (int s = 0; s < 27; ++s) {
        M1(B, i, s) = be[s];
        M1(Dx, i, s) = dxe[s];
        M1(Dy, i, s) = dye[s];
        M1(Dz, i, s) = dze[s];
      }
      if (helm) continue; /* corrections after all domain integrals */
      /* linear-field corrections: (coordinate used, D matrix, ont slot) */
      double *Dm[3] = {Dz, Dx, Dy};
      int crd[3] = {2, 0, 1}, slot[3] = {7, 5, 6};
      for (int t = 0; t < 3; ++t) {
        double hv = 0.0, vs = 0.0;
        for (int j = 0; j < 27; ++j) {
          hv = hv + M1(H, i, j) * XC(IDC(ic, j + 1), crd[t]);
          vs = vs + M1(Dm[t], i, j);
        }
        if (fabs(hv - vs) > oc[slot[t]]) oc[slot[t]] = fabs(hv - vs);
        M1(Dm[t], i, it - 1) = M1(Dm[t], i, it - 1) + hv - vs;
      }
      double hxy = 0.0, hyz = 0.0, hzx = 0.0;
      for (int j = 0; j < 27; ++j) {
        int nj = IDC(ic, j + 1);
        hxy = hxy + M2(Htx, ii, j) * XC(nj, 1) + M1(Dz, i, j);
        hyz = hyz + M2(Hty, ii, j) * XC(nj, 2) + M1(Dx, i, j);
        hzx = hzx + M2(Htz, ii, j) * XC(nj, 0) + M1(Dy, i, j);
      }
      if (fabs(hyz) > oc[8]) oc[8] = fabs(hyz);
      if (fabs(hzx) > oc[9]) oc[9] = fabs(hzx);
      if (fabs(hxy) > oc[10]) oc[10] = fabs(hxy);
    }
    for (int its = 0; its < 6; ++its)
      for (int itf = 1; itf <= 4; ++itf) {
        int dn = ibf[ic + (size_t)nc * (its + 6 * (itf - 1))];
        double p[3] = {xq[(dn - 1)], xq[(dn - 1) + (size_t)nq], xq[(dn - 1) + 2 * (size_t)nq]};
        size_t i = (size_t)ic * 51 + 27 + its * 4 + (itf - 1);
        cpts[2 * ic] += inted27_k(q, p, cc8, 27 + its * 4 + itf, helm ? muq[dn - 1] : 0.0, be, dxe, dye, dze);
        for (int s = 0; s < 27; ++s) {
          M1(B, i, s) = be[s];
          M1(Dx, i, s) = dxe[s];
          M1(Dy, i, s) = dye[s];
          M1(Dz, i, s) = dze[s];
        }
      }
    if (helm) {
      /* ---- free terms and Ht row sums (lib_mHelm.f:318-404) ---- */
      double cfree = 0.0;
      for (int it = 1; it <= 27; ++it) {
        int dn = IDC(ic, it);
        double mu = mun[dn - 1];
        size_t i = (size_t)ic * 51 + (it - 1), ii = (size_t)ic * 27 + (it - 1);
        double vH = 0.0, vB = 0.0;
        for (int j = 0; j < 27; ++j) vH = vH + M1(H, i, j);
        for (int j = 0; j < 27; ++j) vB = vB + M1(B, i, j);
        cfree = -vH + mu * mu * vB;
        int col = find_col(nc, idc, ic, dn) - 1;
        M1(H, i, col) = M1(H, i, col) + cfree;
        double *Ht[3] = {Htx, Hty, Htz};
        for (int t = 0; t < 3; ++t) {
          double vs = 0.0;
          for (int j = 0; j < 27; ++j) vs = vs + M2(Ht[t], ii, j);
          if (fabs(vs) > oc[2 + t]) oc[2 + t] = fabs(vs);
          M2(Ht[t], ii, it - 1) = M2(Ht[t], ii, it - 1) - vs;
        }
      }
      oc[0] = fabs(1.0 - cfree);
      for (int its = 0; its < 6; ++its)
        for (int itf = 1; itf <= 4; ++itf) {
          int dn = ibf[ic + (size_t)nc * (its + 6 * (itf - 1))];
          double mu = muq[dn - 1];
          size_t i = (size_t)ic * 51 + 27 + its * 4 + (itf - 1);
          double vH = 0.0, vB = 0.0;
          for (int j = 0; j < 27; ++j) vH = vH + M1(H, i, j);
          for (int j = 0; j < 27; ++j) vB = vB + M1(B, i, j);
          double c = -vH + mu * mu * vB;
          oc[1] += fabs(0.5 - c);
          const double tc = 3.0 / 4.0;
          double fi[9];
          or_shape9((itf == 1 || itf == 4) ? -tc : tc, (itf <= 2) ? -tc : tc, fi);
          for (int s = 0; s < 9; ++s) {
            int col = find_col(nc, idc, ic, ibcl[its][s]) - 1;
            M1(H, i, col) = M1(H, i, col) + c * fi[s];
          }
        }
      /* ---- linear-field corrections with H - mu^2 B (lib_mHelm.f:409-458) ---- */
      for (int it = 1; it <= 27; ++it) {
        int dn = IDC(ic, it);
        double mu = mun[dn - 1];
        size_t i = (size_t)ic * 51 + (it - 1), ii = (size_t)ic * 27 + (it - 1);
        double *Dm[3] = {Dz, Dx, Dy};
        int crd[3] = {2, 0, 1}, slot[3] = {7, 5, 6};
        for (int t = 0; t < 3; ++t) {
          double hv = 0.0, vs = 0.0;
          for (int j = 0; j < 27; ++j) {
            hv = hv + (M1(H, i, j) - mu * mu * M1(B, i, j)) * XC(IDC(ic, j + 1), crd[t]);
            vs = vs + M1(Dm[t], i, j);
          }
          if (fabs(hv - vs) > oc[slot[t]]) oc[slot[t]] = fabs(hv - vs);
          M1(Dm[t], i, it - 1) = M1(Dm[t], i, it - 1) + hv - vs;
        }
        double hxy = 0.0, hyz = 0.0, hzx = 0.0;
        for (int j = 0; j < 27; ++j) {
          int nj = IDC(ic, j + 1);
          hxy = hxy + M2(Htx, ii, j) * XC(nj, 1) + M1(Dz, i, j);
          hyz = hyz + M2(Hty, ii, j) * XC(nj, 2) + M1(Dx, i, j);
          hzx = hzx + M2(Htz, ii, j) * XC(nj, 0) + M1(Dy, i, j);
        }
        if (fabs(hyz) > oc[8]) oc[8] = fabs(hyz);
        if (fabs(hzx) > oc[9]) oc[9] = fabs(hzx);
        if (fabs(hxy) > oc[10]) oc[10] = fabs(hxy);
      }
    }
  }
  /* copy=1: replicate cell 1 (lib_fmat.f:89-120) */
  if (copy)
    for (int ic = 1; ic < nc; ++ic) {
      for (int it = 0; it < 51; ++it) {
        size_t i = (size_t)ic * 51 + it;
        for (int j = 0; j < 27; ++j) {
          M1(H, i, j) = M1(H, it, j);
          M1(B, i, j) = M1(B, it, j);
          M1(Abdx, i, j) = M1(Abdx, it, j);
          M1(Abdy, i, j) = M1(Abdy, it, j);
          M1(Abdz, i, j) = M1(Abdz, it, j);
          M1(Dx, i, j) = M1(Dx, it, j);
          M1(Dy, i, j) = M1(Dy, it, j);
          M1(Dz, i, j) = M1(Dz, it, j);
        }
        for (int je = 0; je < 6; ++je)
          for (int s = 0; s < 4; ++s) {
            int col = je * 4 + s;
            M1(G, i, col) = M1(G, it, col) * (double)flpm[ic + (size_t)nc * je] * (double)flpm[(size_t)nc * je];
          }
      }
      for (int it = 0; it < 27; ++it) {
        size_t i = (size_t)ic * 27 + it;
        for (int j = 0; j < 27; ++j) {
          M2(Htx, i, j) = M2(Htx, it, j);
          M2(Hty, i, j) = M2(Hty, it, j);
          M2(Htz, i, j) = M2(Htz, it, j);
        }
      }
    }
  /* reduce the accuracy estimates in cell order (sums for H/G, max else) */
  double o[12] = {0};
  for (int ic = 0; ic < ncomp; ++ic) {
    double *oc = ontc + (size_t)ic * 12;
    o[0] += oc[0];
    o[1] += oc[1];
    for (int t = 2; t < 12; ++t)
      if (oc[t] > o[t]) o[t] = oc[t];
  }
  if (copy) {
    o[1] = o[1] / 24;
  } else {
    o[1] = o[1] / nc / 24;
    o[0] = o[0] / nc;
  }
  /* file order: ontH,ontG,ontHtx,ontHty,ontHtz,ontDx,ontDy,ontDz,ontHtyDx,ontHtzDy,ontHtxDz */
  for (int t = 0; t < 11; ++t) ont[t] = o[t];
  if (mumax) *mumax = o[11];
  g_points[0] = g_points[1] = 0;
  for (int ic = 0; ic < ncomp; ++ic) g_points[0] += cpts[2 * ic], g_points[1] += cpts[2 * ic + 1];
  free(cpts);
  free(ontc);
#undef M1
#undef M2
}

void or_macro_integrals(const or_quad *q, int nn, int nc, int nq, const double *x, const int *idc,
                        const int *ibf, const double *xq, const int *flpm, int copy, double *H,
                        double *G, double *B, double *Abdx, double *Abdy, double *Abdz, double *Htx,
                        double *Hty, double *Htz, double *Dx, double *Dy, double *Dz, double ont[11],
                        int nthreads) {
  macro_k(q, nn, nc, nq, x, idc, ibf, xq, flpm, copy, NULL, NULL, H, G, B, Abdx, Abdy, Abdz, Htx, Hty, Htz, Dx,
          Dy, Dz, ont, NULL, nthreads);
}

void or_macro_integrals_mh(const or_quad *q, int nn, int nc, int nq, const double *x, const int *idc,
                           const int *ibf, const double *xq, const int *flpm, int copy, const double *diff,
                           const double *qdiff, double beta, double *H, double *G, double *B, double *Abdx,
                           double *Abdy, double *Abdz, double *Htx, double *Hty, double *Htz, double *Dx,
                           double *Dy, double *Dz, double ont[11], double *mumax, int nthreads) {
  /* mu = SQRT(inp%beta/mHdiff(dn)) per source (lib_mHelm.f:139,197,279,303) */
  double *mun = (double *)malloc(sizeof(double) * (size_t)nn);
  double *muq = (double *)malloc(sizeof(double) * (size_t)nq);
  for (int k = 0; k < nn; ++k) mun[k] = sqrt(beta / diff[k]);
  for (int k = 0; k < nq; ++k) muq[k] = sqrt(beta / qdiff[k]);
  macro_k(q, nn, nc, nq, x, idc, ibf, xq, flpm, copy, mun, muq, H, G, B, Abdx, Abdy, Abdz, Htx, Hty, Htz, Dx, Dy,
          Dz, ont, mumax, nthreads);
  free(mun);
  free(muq);
}

/* -------------------------------------------------------------------------- */
void or_sd_integrals(const or_quad *q, int nn, int nb, int ne, int nc, const double *x, const int *ibc,
                     const int *idc, const int *gbn, const int *lbn, int row0, int row1, int ld,
                     double *H, double *Htx, double *Hty, double *Htz, double *Dx, double *Dy,
                     double *Dz, double ont[10], int nthreads) {
#define SM(A, r, c) A[(r) + (size_t)ld * (c)]
#define IBC(e, k) ibc[(e) + (size_t)ne * ((k)-1)]
  int nr = row1 - row0;
  for (size_t c = 0; c < (size_t)nb; ++c)
    for (int r = 0; r < nr; ++r) SM(H, r, c) = SM(Htx, r, c) = SM(Hty, r, c) = SM(Htz, r, c) = 0.0;
  for (size_t c = 0; c < (size_t)nn; ++c)
    for (int r = 0; r < nr; ++r) SM(Dx, r, c) = SM(Dy, r, c) = SM(Dz, r, c) = 0.0;
  double *ontr = (double *)calloc((size_t)nr * 10, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int r = 0; r < nr; ++r) {
    int i = row0 + r; /* 0-based boundary index */
    double *oc = ontr + (size_t)r * 10;
    int dn = gbn[i];
    double p[3] = {XC(dn, 0), XC(dn, 1), XC(dn, 2)};
    double he[9], htx[9], hty[9], htz[9];
    for (int je = 0; je < ne; ++je) {
      double cf[4][3];
      for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) cf[k][d] = XC(IBC(je, 2 * k + 1), d);
      int isrc = 0;
      for (int s = 1; s <= 9; ++s)
        if (dn == IBC(je, s)) isrc = s;
      or_inteb9(q, 0, p, cf, isrc, 0.0, he, NULL, NULL, NULL, NULL, htx, hty, htz);
      for (int s = 1; s <= 9; ++s) {
        int j = lbn[IBC(je, s) - 1] - 1;
        SM(H, r, j) = SM(H, r, j) + he[s - 1];
        SM(Htx, r, j) = SM(Htx, r, j) - htx[s - 1];
        SM(Hty, r, j) = SM(Hty, r, j) - hty[s - 1];
        SM(Htz, r, j) = SM(Htz, r, j) - htz[s - 1];
      }
    }
    double cfree = 0.0;
    for (int j = 0; j < nb; ++j) cfree = cfree - SM(H, r, j);
    oc[0] = cfree;
    SM(H, r, i) = SM(H, r, i) + cfree;
    double *Ht[3] = {Htx, Hty, Htz};
    for (int t = 0; t < 3; ++t) {
      double vs = 0.0;
      for (int j = 0; j < nb; ++j) vs = vs + SM(Ht[t], r, j);
      oc[1 + t] = fabs(vs);
      SM(Ht[t], r, i) = SM(Ht[t], r, i) - vs;
    }
    double be[27], dxe[27], dye[27], dze[27];
    for (int ic = 0; ic < nc; ++ic) {
      double cc8[8][3];
      cell_corners(nn, nc, x, idc, ic, cc8);
      int isrc = 0;
      for (int s = 1; s <= 27; ++s)
        if (dn == IDC(ic, s)) isrc = s;
      or_inted27(q, p, cc8, isrc, be, dxe, dye, dze);
      for (int s = 1; s <= 27; ++s) {
        int j = IDC(ic, s) - 1;
        SM(Dx, r, j) = SM(Dx, r, j) + dxe[s - 1];
        SM(Dy, r, j) = SM(Dy, r, j) + dye[s - 1];
        SM(Dz, r, j) = SM(Dz, r, j) + dze[s - 1];
      }
    }
    double *Dm[3] = {Dz, Dx, Dy};
    int crd[3] = {2, 0, 1}, slot[3] = {6, 4, 5};
    for (int t = 0; t < 3; ++t) {
      double hv = 0.0, vs = 0.0;
      for (int s = 0; s < nb; ++s) hv = hv + SM(H, r, s) * XC(gbn[s], crd[t]);
      for (int j = 0; j < nn; ++j) vs = vs + SM(Dm[t], r, j);
      oc[slot[t]] = fabs(hv - vs);
      SM(Dm[t], r, dn - 1) = SM(Dm[t], r, dn - 1) + hv - vs;
    }
    double v1 = 0, v2 = 0, v3 = 0, hxy = 0, hyz = 0, hzx = 0;
    for (int j = 0; j < nn; ++j) {
      v1 = v1 + SM(Dz, r, j);
      v2 = v2 + SM(Dx, r, j);
      v3 = v3 + SM(Dy, r, j);
    }
    for (int j = 0; j < nb; ++j) {
      hxy = hxy + SM(Htx, r, j) * XC(gbn[j], 1);
      hyz = hyz + SM(Hty, r, j) * XC(gbn[j], 2);
      hzx = hzx + SM(Htz, r, j) * XC(gbn[j], 0);
    }
    oc[7] = fabs(hxy + v1); /* HtxDz */
    oc[8] = fabs(hyz + v2); /* HtyDx */
    oc[9] = fabs(hzx + v3); /* HtzDy */
  }
  /* order in the cache file: ontH,ontHtx,ontHty,ontHtz,ontDx,ontDy,ontDz,ontHtxDz,ontHtyDx,ontHtzDy */
  double o[10] = {0};
  for (int r = 0; r < nr; ++r) {
    double *oc = ontr + (size_t)r * 10;
    if (oc[0] > o[0]) o[0] = oc[0];
    for (int t = 1; t < 10; ++t)
      if (oc[t] > o[t]) o[t] = oc[t];
  }
  o[0] = fabs(o[0] - 0.5);
  for (int t = 0; t < 10; ++t) ont[t] = o[t];
  free(ontr);
#undef SM
#undef IBC
}
