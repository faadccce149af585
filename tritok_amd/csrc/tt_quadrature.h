// tt_quadrature.h -- device-side quadrature: shape functions, the Laplace
// kernels and the per-element integration loops of the reference
// (lib_int.f INTEDc27dcLap :4-398, INTEBc9kmdclap :1462-1779,
// lib_sdkm.f INTEBc9sdkm :2082-2376).  The loops keep the reference's
// subdivision rules and singular transforms; sums run in the same point
// order per output.
#pragma once
#include <hip/hip_runtime.h>

namespace tt {

constexpr int kMaxRule = 64;

// Quadrature rules as used by one integration call: regular and singular rule
// of the boundary (B) and domain (D) integrals, and the sub-cell count.
struct QuadDev {
  double gbr[kMaxRule], obr[kMaxRule];  // boundary regular
  double gbs[kMaxRule], obs[kMaxRule];  // boundary singular
  double gdr[kMaxRule], odr[kMaxRule];  // domain regular
  double gds[kMaxRule], ods[kMaxRule];  // domain singular
  int nbr, nbs, ndr, nds, iiDiv;
};

// source point positions in [0,1]^3 for isrc 1..51 (27 nodes, 24 flux points)
__constant__ static const double kKsi51[51] = {
    0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.5, 1.0, 0.5, 0.0, 0.0, 1.0, 1.0, 0.0,
    0.5, 1.0, 0.5, 0.5, 0.0, 0.5, 1.0, 0.5, 0.5, 0.5, 0.125, 0.125, 0.875, 0.875, 0.125, 0.875,
    0.875, 0.125, 1.0, 1.0, 1.0, 1.0, 0.875, 0.125, 0.125, 0.875, 0.0, 0.0, 0.0, 0.0, 0.125,
    0.875, 0.875, 0.125};
__constant__ static const double kEta51[51] = {
    1.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 0.5, 0.5, 0.5, 0.5, 0.0,
    0.0, 0.0, 0.0, 1.0, 0.5, 0.5, 0.5, 0.5, 0.0, 0.5, 0.125, 0.875, 0.875, 0.125, 0.0, 0.0,
    0.0, 0.0, 0.125, 0.875, 0.875, 0.125, 1.0, 1.0, 1.0, 1.0, 0.875, 0.125, 0.125, 0.875,
    0.125, 0.125, 0.875, 0.875};
__constant__ static const double kZet51[51] = {
    1.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 0.5, 0.0, 0.5, 1.0, 1.0, 0.0, 0.0, 1.0, 0.5,
    0.0, 0.5, 1.0, 0.5, 0.5, 0.0, 0.5, 1.0, 0.5, 0.5, 0.0, 0.0, 0.0, 0.0, 0.125, 0.125,
    0.875, 0.875, 0.125, 0.125, 0.875, 0.875, 0.125, 0.125, 0.875, 0.875, 0.125, 0.125,
    0.875, 0.875, 1.0, 1.0, 1.0, 1.0};
__constant__ static const double kKsi13[13] = {0.0, 0.5, 1.0, 1.0, 1.0, 0.5, 0.0, 0.0, 0.5, 0.125, 0.875, 0.875, 0.125};
__constant__ static const double kEta13[13] = {0.0, 0.0, 0.0, 0.5, 1.0, 1.0, 1.0, 0.5, 0.5, 0.125, 0.125, 0.875, 0.875};

__device__ __forceinline__ void shape27(double xi, double eta, double zeta, double *fi) {
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

__device__ __forceinline__ void shape9(double xsi, double eta, double *fi) {
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

__device__ __forceinline__ void dl34(double xsi, double eta, double *fi) {
  const double sd = 4.0 / 9.0, tc = 3.0 / 4.0;
  fi[0] = sd * (xsi - tc) * (eta - tc);
  fi[1] = -sd * (xsi + tc) * (eta - tc);
  fi[2] = sd * (xsi + tc) * (eta + tc);
  fi[3] = -sd * (xsi - tc) * (eta + tc);
}

// Trilinear map of the 8-node cell geometry (corners in the reference's
// argument order x1..x8): position and Jacobian determinant at (e1,e2,e3).
// detj carries the reference's factor 512 (its derivatives omit 1/8).
__device__ __forceinline__ void cell_map(const double (&c)[8][3], double eti, double etj, double etk, double *xc,
                                         double *detj) {
  double e1m = 1.0 - eti, e1p = 1.0 + eti, e2m = 1.0 - etj, e2p = 1.0 + etj, e3m = 1.0 - etk, e3p = 1.0 + etk;
  double fig[8] = {0.125 * e1m * e2m * e3m, 0.125 * e1p * e2m * e3m, 0.125 * e1m * e2p * e3m, 0.125 * e1p * e2p * e3m,
                   0.125 * e1m * e2m * e3p, 0.125 * e1p * e2m * e3p, 0.125 * e1m * e2p * e3p, 0.125 * e1p * e2p * e3p};
  double aj[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    xc[d] = fig[0] * c[0][d] + fig[1] * c[1][d] + fig[2] * c[2][d] + fig[3] * c[3][d] + fig[4] * c[4][d] +
            fig[5] * c[5][d] + fig[6] * c[6][d] + fig[7] * c[7][d];
    aj[0][d] = -e2m * e3m * c[0][d] + e2m * e3m * c[1][d] - e2p * e3m * c[2][d] + e2p * e3m * c[3][d] -
               e2m * e3p * c[4][d] + e2m * e3p * c[5][d] - e2p * e3p * c[6][d] + e2p * e3p * c[7][d];
    aj[1][d] = -e1m * e3m * c[0][d] - e1p * e3m * c[1][d] + e1m * e3m * c[2][d] + e1p * e3m * c[3][d] -
               e1m * e3p * c[4][d] - e1p * e3p * c[5][d] + e1m * e3p * c[6][d] + e1p * e3p * c[7][d];
    aj[2][d] = -e1m * e2m * c[0][d] - e1p * e2m * c[1][d] - e1m * e2p * c[2][d] - e1p * e2p * c[3][d] +
               e1m * e2m * c[4][d] + e1p * e2m * c[5][d] + e1m * e2p * c[6][d] + e1p * e2p * c[7][d];
  }
  *detj = (aj[0][0] * aj[1][1] * aj[2][2] + aj[0][1] * aj[1][2] * aj[2][0] + aj[0][2] * aj[1][0] * aj[2][1] -
           aj[0][0] * aj[1][2] * aj[2][1] - aj[0][1] * aj[1][0] * aj[2][2] - aj[0][2] * aj[1][1] * aj[2][0]);
}

// Fundamental-solution factors at distance ra: u* and the factor of its
// gradient.  Laplace: u* = 1/(4 pi r), factor 1.  Modified Helmholtz
// (INTEDc27modHelm lib_mHelm.f:1905-1910, INTEBc9modHelm :2227-2233):
// u* = exp(-mu r)/(4 pi r), gradient factor (1 + mu r).
template <bool Helm>
__device__ __forceinline__ void kernel_factors(double ra, double mu, double &fung, double &gf) {
  const double PI = 3.141592653589793;
  fung = 0.25 / (PI * ra);
  gf = 1.0;
  if (Helm) {
    fung = fung * exp(-mu * ra);
    gf = 1 + mu * ra;
  }
}

// Accumulator of the 27-node domain integrals.  NC = 4 also keeps be (u*),
// NC = 3 keeps only the gradient components (single-domain kinematics).
// Helm: modified-Helmholtz kernel with the source's mu.
template <int NC, bool Helm = false>
struct DomAcc {
  double a[NC][27];
  double mu = 0.0;
  int npts = 0;  // kernel evaluations (measurement)
  __device__ void zero() {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int k = 0; k < 27; ++k) a[c][k] = 0.0;
  }
  // add one quadrature point: weight w (incl. Jacobian), kernel terms
  __device__ __forceinline__ void add(const double *p, const double *xc, double ajac, double eti, double etj,
                                      double etk, bool skip_zero) {
    ++npts;
    double rx = p[0] - xc[0], ry = p[1] - xc[1], rz = p[2] - xc[2];
    double ra = sqrt(rx * rx + ry * ry + rz * rz);
    if (skip_zero && ra == 0.0) return;
    if (ra < 1.0e-12) ra = 1.0e-12;
    double ira2 = 1.0 / (ra * ra);
    double fung, gf;
    kernel_factors<Helm>(ra, mu, fung, gf);
    double f[4];
    f[0] = rx * ira2 * fung * gf;
    f[1] = ry * ira2 * fung * gf;
    f[2] = rz * ira2 * fung * gf;
    f[3] = fung;
    double f27[27];
    shape27(eti, etj, etk, f27);
#pragma unroll
    for (int k = 0; k < 27; ++k) {
      double dom = ajac * f27[k];
#pragma unroll
      for (int c = 0; c < NC; ++c) a[c][k] += f[c] * dom;
    }
  }
};

// Domain integral of one (source, cell) pair: INTEDc27dcLap (INTEDc27modHelm
// with a Helm accumulator).
template <class Acc>
__device__ void integrate_cell(const QuadDev &q, const double *p, const double (&c)[8][3], int isrc, Acc &acc) {
  const double PI = 3.141592653589793, PI2 = 2.0 * PI;
  int nd = isrc != 0 ? q.iiDiv : 1;
  for (int ix = 1; ix <= nd; ++ix)
    for (int iy = 1; iy <= nd; ++iy)
      for (int iz = 1; iz <= nd; ++iz) {
        double a = -1.0 + (ix - 1) * 2.0 / nd, b = -1.0 + ix * 2.0 / nd;
        double cc = -1.0 + (iy - 1) * 2.0 / nd, d = -1.0 + iy * 2.0 / nd;
        double e = -1.0 + (iz - 1) * 2.0 / nd, f = -1.0 + iz * 2.0 / nd;
        double dex = 0.125 * (b - a) * (d - cc) * (f - e);
        bool ising = false;
        double xx1 = 0, yy1 = 0, zz1 = 0;
        if (isrc != 0) {
          xx1 = ((-1.0 + 2.0 * kKsi51[isrc - 1]) - a) / (b - a);
          yy1 = ((-1.0 + 2.0 * kEta51[isrc - 1]) - cc) / (d - cc);
          zz1 = ((-1.0 + 2.0 * kZet51[isrc - 1]) - e) / (f - e);
          ising = xx1 >= 0.0 && xx1 <= 1.0 && yy1 >= 0.0 && yy1 <= 1.0 && zz1 >= 0.0 && zz1 <= 1.0;
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
            for (int K = 0; K < q.nds; ++K) {
              double ffi = (ffi1 - ffi0) * q.gds[K] / 2.0 + (ffi1 + ffi0) / 2.0;
              double ro0 = aa0 * sin(gama) / sin(ffi - ffi0 + gama);
              double sf = sin(ffi), cf = cos(ffi);
              for (int N = 1; N <= 3; ++N) {
                double aa1, eps, th0, th1;
                if (N == 1) {
                  aa1 = 1.0 - zz1;
                  eps = PI / 2.0;
                  th0 = 0.0;
                  th1 = acos(aa1 / sqrt(aa1 * aa1 + ro0 * ro0));
                } else if (N == 2) {
                  aa1 = sqrt(ro0 * ro0 + (1.0 - zz1) * (1.0 - zz1));
                  eps = acos((1.0 - zz1) / aa1);
                  th0 = eps;
                  double rr1 = fmin(1.0, ro0 / sqrt(zz1 * zz1 + ro0 * ro0));
                  th1 = 0.5 * PI + acos(rr1);
                } else {
                  aa1 = sqrt(ro0 * ro0 + zz1 * zz1);
                  double rr1 = fmin(1.0, ro0 / aa1);
                  eps = acos(rr1);
                  th0 = 0.5 * PI + eps;
                  th1 = PI;
                }
                for (int I = 0; I < q.nds; ++I) {
                  double th = (th1 - th0) * q.gds[I] / 2.0 + (th1 + th0) / 2.0;
                  double ro1 = aa1 * sin(eps) / sin(th - th0 + eps);
                  double sth = sin(th), cth = cos(th);
                  double wki = q.ods[I] * q.ods[K];
                  for (int J = 0; J < q.nds; ++J) {
                    double ro = (ro1 * q.gds[J] + ro1) / 2.0;
                    double eti = 2.0 * (xx1 + ro * sth * cf) - 1.0;
                    double etj = 2.0 * (yy1 + ro * sth * sf) - 1.0;
                    double etk = 2.0 * (zz1 + ro * cth) - 1.0;
                    eti = a + 0.5 * (eti + 1.0) * (b - a);
                    etj = cc + 0.5 * (etj + 1.0) * (d - cc);
                    etk = e + 0.5 * (etk + 1.0) * (f - e);
                    double xc[3], dj;
                    cell_map(c, eti, etj, etk, xc, &dj);
                    double ajac = 1.953125e-3 * dj * q.ods[J] * wki * (ffi1 - ffi0) * (th1 - th0) * ro1 * (ro * ro) * sth;
                    ajac = dex * ajac;
                    acc.add(p, xc, ajac, eti, etj, etk, true);
                  }
                }
              }
            }
          }
        } else {
          for (int K = 0; K < q.ndr; ++K) {
            double gik = e + 0.5 * (q.gdr[K] + 1.0) * (f - e);
            for (int J = 0; J < q.ndr; ++J) {
              double gij = cc + 0.5 * (q.gdr[J] + 1.0) * (d - cc);
              double wjk = q.odr[J] * q.odr[K] * dex;
              for (int I = 0; I < q.ndr; ++I) {
                double gii = a + 0.5 * (q.gdr[I] + 1.0) * (b - a);
                double xc[3], dj;
                cell_map(c, gii, gij, gik, xc, &dj);
                double ajac = 1.953125e-3 * dj * q.odr[I] * wjk;
                acc.add(p, xc, ajac, gii, gij, gik, false);
              }
            }
          }
        }
      }
}

// ---------------------------------------------------------------------------
// Wavefront-parallel domain integral of one (source, cell) pair.  The same
// quadrature points as integrate_cell, taken 64 at a time: in the first phase
// every lane evaluates one point (geometry, Jacobian, kernel) and stages in
// LDS the nine products l_a(xi) l_b(eta) of its 1-D quadratic Lagrange factors
// and the 3 NC products of its NC weighted kernel values with l_c(zeta); in
// the second phase lane k of each half-wave (k < 27) accumulates shape
// function k's NC sums over that half's 32 points, one fused multiply-add per
// sum and point.  Each lane holds NC accumulators instead of 27 NC, so nothing
// spills.  The 27-node shape functions are the products l_a(xi) l_b(eta)
// l_c(zeta) of the quadratic Lagrange factors (a, b, c the node's position in
// {-1, 0, 1}, cshape27's node order), equal to cshape27 to rounding.
// the three 1-D quadratic Lagrange factors at t (nodes -1, 0, 1)
__device__ __forceinline__ void lagrange3(double t, double *l) {
  l[0] = 0.5 * t * (t - 1.0);
  l[1] = (1.0 - t) * (1.0 + t);
  l[2] = 0.5 * t * (t + 1.0);
}

__constant__ static const signed char kNode27[27][3] = {
    {-1, 1, 1},  {-1, 1, -1}, {1, 1, -1},  {1, 1, 1},  {-1, -1, 1}, {-1, -1, -1}, {1, -1, -1},
    {1, -1, 1},  {-1, 1, 0},  {0, 1, -1},  {1, 1, 0},  {0, 1, 1},   {-1, 0, 1},   {-1, 0, -1},
    {1, 0, -1},  {1, 0, 1},   {-1, -1, 0}, {0, -1, -1}, {1, -1, 0}, {0, -1, 1},   {0, 1, 0},
    {-1, 0, 0},  {0, 0, -1},  {1, 0, 0},   {0, 0, 1},  {0, -1, 0},  {0, 0, 0}};

// entries of the singular transform's table: (K, N, I) triples of one chunk
// of K values (see WaveCell::run)
constexpr int kTrigTab = 192;
// at most kTrigTab / 3 K values per chunk (one (N, I) grid of at least 3 entries each)
constexpr int kTrigK = kTrigTab / 3;
// staging rows: point 32 h + j of a stage sits in column 34 h + j of a row of
// 68 doubles, so that the rows one read instruction touches (nine products
// and two half-waves, or three zeta factors and two half-waves) spread over
// the 64 LDS banks; two consecutive points are one 16-byte read
constexpr int kStRow = 68, kStHalf = 34;

// LDS doubles one wavefront needs for WaveCell<NC> (16-byte aligned)
template <int NC>
constexpr int wave_cell_lds() {
  return (9 + 3 * NC) * kStRow + kTrigTab * 4 + 2 * kTrigK;
}

template <int NC, bool Helm>
struct WaveCell {
  double acc[NC];  // lane k < 27 of the lower half-wave: the sums of shape function k
  long npts = 0;   // points this lane evaluated

  __device__ __forceinline__ void stage(const double *p, const double (&c)[8][3], double mu, bool ok, double eti,
                                        double etj, double etk, double wpre, bool skip_zero, double *lds,
                                        int ax, int ay, int az) {
    const int lane = threadIdx.x & 63;
    double f[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) f[q] = 0.0;
    if (ok) {
      ++npts;
      double xc[3], dj;
      cell_map(c, eti, etj, etk, xc, &dj);
      const double ajac = wpre * dj;
      const double rx = p[0] - xc[0], ry = p[1] - xc[1], rz = p[2] - xc[2];
      double ra = sqrt(rx * rx + ry * ry + rz * rz);
      if (!(skip_zero && ra == 0.0)) {
        if (ra < 1.0e-12) ra = 1.0e-12;
        const double ira2 = 1.0 / (ra * ra);
        double fung, gf;
        kernel_factors<Helm>(ra, mu, fung, gf);
        f[0] = rx * ira2 * fung * gf * ajac;
        f[1] = ry * ira2 * fung * gf * ajac;
        f[2] = rz * ira2 * fung * gf * ajac;
        if (NC == 4) f[NC - 1] = fung * ajac;
      }
    }
    double lx[3], ly[3], lz[3];
    lagrange3(eti, lx);
    lagrange3(etj, ly);
    lagrange3(etk, lz);
    const int half = lane >> 5;
    double *P = lds + half * kStHalf, *G = P + 9 * kStRow;
    const int col = lane & 31;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) P[(3 * a + b) * kStRow + col] = lx[a] * ly[b];
#pragma unroll
    for (int q = 0; q < NC; ++q)
#pragma unroll
      for (int c = 0; c < 3; ++c) G[(3 * q + c) * kStRow + col] = f[q] * lz[c];
    __syncthreads();
    const double2 *Pr = reinterpret_cast<const double2 *>(P + (3 * ax + ay) * kStRow);
    const double2 *Gr = reinterpret_cast<const double2 *>(G + az * kStRow);
#pragma unroll 4
    for (int j = 0; j < 16; ++j) {
      const double2 pv = Pr[j];
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        const double2 gv = Gr[q * (3 * kStRow / 2) + j];
        acc[q] += pv.x * gv.x;
        acc[q] += pv.y * gv.y;
      }
    }
    __syncthreads();
  }

  // the integral over the cell; on return lanes 0..26 hold the sums of shape
  // functions 0..26 (both halves combined).  regular = false leaves out the
  // sub-cubes that do not hold the source (integrated elsewhere, see
  // macro_reg_kernel).
  __device__ void run(const QuadDev &q, const double *p, const double (&c)[8][3], int isrc, double mu, double *lds,
                      bool regular = true) {
    const double PI = 3.141592653589793, PI2 = 2.0 * PI;
    const int lane = threadIdx.x & 63, k = lane & 31;
    const int ax = k < 27 ? kNode27[k][0] + 1 : 1, ay = k < 27 ? kNode27[k][1] + 1 : 1,
              az = k < 27 ? kNode27[k][2] + 1 : 1;
#pragma unroll
    for (int qq = 0; qq < NC; ++qq) acc[qq] = 0.0;
    const int nd = isrc != 0 ? q.iiDiv : 1;
    for (int ix = 1; ix <= nd; ++ix)
      for (int iy = 1; iy <= nd; ++iy)
        for (int iz = 1; iz <= nd; ++iz) {
          const double a = -1.0 + (ix - 1) * 2.0 / nd, b = -1.0 + ix * 2.0 / nd;
          const double cc = -1.0 + (iy - 1) * 2.0 / nd, d = -1.0 + iy * 2.0 / nd;
          const double e = -1.0 + (iz - 1) * 2.0 / nd, f = -1.0 + iz * 2.0 / nd;
          const double dex = 0.125 * (b - a) * (d - cc) * (f - e);
          bool ising = false;
          double xx1 = 0, yy1 = 0, zz1 = 0;
          if (isrc != 0) {
            xx1 = ((-1.0 + 2.0 * kKsi51[isrc - 1]) - a) / (b - a);
            yy1 = ((-1.0 + 2.0 * kEta51[isrc - 1]) - cc) / (d - cc);
            zz1 = ((-1.0 + 2.0 * kZet51[isrc - 1]) - e) / (f - e);
            ising = xx1 >= 0.0 && xx1 <= 1.0 && yy1 >= 0.0 && yy1 <= 1.0 && zz1 >= 0.0 && zz1 <= 1.0;
          }
          if (ising) {
            for (int Ls = 1; Ls <= 4; ++Ls) {
              double aa0, gama, ffi0, ffi1;
              if (Ls == 1) {
                if (1.0 - xx1 == 0.0) continue;
                aa0 = sqrt(yy1 * yy1 + (1.0 - xx1) * (1.0 - xx1));
                gama = acos(yy1 / aa0);
                ffi0 = 1.5 * PI + acos(yy1 / aa0);
                ffi1 = PI2 + atan((1.0 - yy1) / (1.0 - xx1));
              } else if (Ls == 2) {
                if (1.0 - yy1 == 0.0) continue;
                aa0 = sqrt((1.0 - yy1) * (1.0 - yy1) + (1.0 - xx1) * (1.0 - xx1));
                gama = acos((1.0 - xx1) / aa0);
                ffi0 = asin((1.0 - yy1) / aa0);
                ffi1 = 0.5 * PI + atan(xx1 / (1.0 - yy1));
              } else if (Ls == 3) {
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
              // The angles of the transform depend on K (ffi, ro0), on (K, N)
              // (the theta range) and on (K, N, I) (theta, ro1), not on J: a
              // pre-pass evaluates them once per distinct value, as many
              // (K, N, I) triples at a time as the table holds, and the
              // stages read them back (the same expressions, so the same bits
              // as evaluating them in every lane).
              const int nds = q.nds, kb = kTrigTab / (3 * nds) > 0 ? kTrigTab / (3 * nds) : 1;
              double *tab = lds + (9 + 3 * NC) * kStRow, *ktab = tab + 4 * kTrigTab;
              for (int K0 = 0; K0 < nds; K0 += kb) {
                const int K1 = K0 + kb < nds ? K0 + kb : nds;
                const int ne = (K1 - K0) * 3 * nds;
                for (int e0 = 0; e0 < ne; e0 += 64) {
                  const int te = e0 + lane;
                  if (te < ne && te < kTrigTab) {
                    const int K = K0 + te / (3 * nds), N = (te / nds) % 3 + 1, I = te % nds;
                    const double ffi = (ffi1 - ffi0) * q.gds[K] / 2.0 + (ffi1 + ffi0) / 2.0;
                    const double ro0 = aa0 * sin(gama) / sin(ffi - ffi0 + gama);
                    double aa1, eps, th0, th1;
                    if (N == 1) {
                      aa1 = 1.0 - zz1;
                      eps = PI / 2.0;
                      th0 = 0.0;
                      th1 = acos(aa1 / sqrt(aa1 * aa1 + ro0 * ro0));
                    } else if (N == 2) {
                      aa1 = sqrt(ro0 * ro0 + (1.0 - zz1) * (1.0 - zz1));
                      eps = acos((1.0 - zz1) / aa1);
                      th0 = eps;
                      const double rr1 = fmin(1.0, ro0 / sqrt(zz1 * zz1 + ro0 * ro0));
                      th1 = 0.5 * PI + acos(rr1);
                    } else {
                      aa1 = sqrt(ro0 * ro0 + zz1 * zz1);
                      const double rr1 = fmin(1.0, ro0 / aa1);
                      eps = acos(rr1);
                      th0 = 0.5 * PI + eps;
                      th1 = PI;
                    }
                    const double th = (th1 - th0) * q.gds[I] / 2.0 + (th1 + th0) / 2.0;
                    tab[te] = aa1 * sin(eps) / sin(th - th0 + eps);  // ro1
                    tab[kTrigTab + te] = sin(th);
                    tab[2 * kTrigTab + te] = cos(th);
                    tab[3 * kTrigTab + te] = th1 - th0;
                    if (N == 1 && I == 0) ktab[K - K0] = sin(ffi), ktab[kTrigK + K - K0] = cos(ffi);
                  }
                }
                __syncthreads();
                for (int K = K0; K < K1; ++K) {
                  const double sf = ktab[K - K0], cf = ktab[kTrigK + K - K0];
                  for (int N = 1; N <= 3; ++N) {
                    const int np = nds * nds;
                    for (int base = 0; base < np; base += 64) {
                      const int t = base + lane;
                      const bool ok = t < np;
                      const int I = ok ? t / nds : 0, J = ok ? t % nds : 0;
                      const int te = ((K - K0) * 3 + (N - 1)) * nds + I;
                      const double ro1 = tab[te], sth = tab[kTrigTab + te], cth = tab[2 * kTrigTab + te];
                      const double dth = tab[3 * kTrigTab + te];
                      const double ro = (ro1 * q.gds[J] + ro1) / 2.0;
                      double eti = 2.0 * (xx1 + ro * sth * cf) - 1.0;
                      double etj = 2.0 * (yy1 + ro * sth * sf) - 1.0;
                      double etk = 2.0 * (zz1 + ro * cth) - 1.0;
                      eti = a + 0.5 * (eti + 1.0) * (b - a);
                      etj = cc + 0.5 * (etj + 1.0) * (d - cc);
                      etk = e + 0.5 * (etk + 1.0) * (f - e);
                      const double wpre = dex * 1.953125e-3 * q.ods[J] * (q.ods[I] * q.ods[K]) * (ffi1 - ffi0) *
                                          dth * ro1 * (ro * ro) * sth;
                      stage(p, c, mu, ok, eti, etj, etk, wpre, true, lds, ax, ay, az);
                    }
                  }
                }
              }
            }
          } else if (regular) {
            const int n1 = q.ndr, np = n1 * n1 * n1;
            for (int base = 0; base < np; base += 64) {
              const int t = base + lane;
              const bool ok = t < np;
              const int I = ok ? t % n1 : 0, J = ok ? (t / n1) % n1 : 0, K = ok ? t / (n1 * n1) : 0;
              const double gii = a + 0.5 * (q.gdr[I] + 1.0) * (b - a);
              const double gij = cc + 0.5 * (q.gdr[J] + 1.0) * (d - cc);
              const double gik = e + 0.5 * (q.gdr[K] + 1.0) * (f - e);
              const double wpre = 1.953125e-3 * q.odr[I] * (q.odr[J] * q.odr[K] * dex);
              stage(p, c, mu, ok, gii, gij, gik, wpre, false, lds, ax, ay, az);
            }
          }
        }
#pragma unroll
    for (int qq = 0; qq < NC; ++qq) acc[qq] += __shfl_xor(acc[qq], 32, 64);
  }
};

// Accumulator of the 9-node face integrals; Macro also keeps ge and ad*.
template <bool Macro, bool Helm = false>
struct BndAcc {
  double he[9], hexy[9], heyz[9], hezx[9];
  double ad[Macro ? 3 : 1][Macro ? 9 : 1];
  double ge[Macro ? 4 : 1];
  double mu = 0.0;
  int npts = 0;  // kernel evaluations (measurement)
  __device__ void zero() {
#pragma unroll
    for (int s = 0; s < 9; ++s) he[s] = hexy[s] = heyz[s] = hezx[s] = 0.0;
    if (Macro) {
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int s = 0; s < 9; ++s) ad[c][s] = 0.0;
#pragma unroll
      for (int s = 0; s < 4; ++s) ge[s] = 0.0;
    }
  }
  __device__ __forceinline__ void add(const double *p, const double (&c)[4][3], double eti, double etj,
                                      double wpre) {
    ++npts;
    double e1m = 1.0 - eti, e1p = 1.0 + eti, e2m = 1.0 - etj, e2p = 1.0 + etj;
    double f4[4] = {0.25 * e1m * e2m, 0.25 * e1p * e2m, 0.25 * e1p * e2p, 0.25 * e1m * e2p};
    double xc[3], ks[3], et[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      xc[d] = f4[0] * c[0][d] + f4[1] * c[1][d] + f4[2] * c[2][d] + f4[3] * c[3][d];
      ks[d] = 0.25 * (-e2m * c[0][d] + e2m * c[1][d] + e2p * c[2][d] - e2p * c[3][d]);
      et[d] = 0.25 * (-e1m * c[0][d] - e1p * c[1][d] + e1p * c[2][d] + e1m * c[3][d]);
    }
    double anx = ks[1] * et[2] - et[1] * ks[2];
    double any = et[0] * ks[2] - ks[0] * et[2];
    double anz = ks[0] * et[1] - et[0] * ks[1];
    double ajac = 1.0 / sqrt(anx * anx + any * any + anz * anz);
    double n1 = anx * ajac, n2 = any * ajac, n3 = anz * ajac;
    ajac = wpre / ajac;
    double rx = p[0] - xc[0], ry = p[1] - xc[1], rz = p[2] - xc[2];
    double ra = sqrt(rx * rx + ry * ry + rz * rz);
    double ira2 = 1.0 / (ra * ra);
    double fung, gf;
    kernel_factors<Helm>(ra, mu, fung, gf);
    double funx = rx * ira2 * fung * gf, funy = ry * ira2 * fung * gf, funz = rz * ira2 * fung * gf;
    double funh = funx * n1 + funy * n2 + funz * n3;
    double cxy = funx * n2 - funy * n1, cyz = funy * n3 - funz * n2, czx = funz * n1 - funx * n3;
    double f9[9];
    shape9(eti, etj, f9);
#pragma unroll
    for (int s = 0; s < 9; ++s) {
      double dg = ajac * f9[s];
      he[s] += funh * dg;
      hexy[s] -= cxy * dg;
      heyz[s] -= cyz * dg;
      hezx[s] -= czx * dg;
      if (Macro) {
        ad[0][s] += fung * n1 * dg;
        ad[1][s] += fung * n2 * dg;
        ad[2][s] += fung * n3 * dg;
      }
    }
    if (Macro) {
      double fg[4];
      dl34(eti, etj, fg);
#pragma unroll
      for (int s = 0; s < 4; ++s) ge[s] += fung * (ajac * fg[s]);
    }
  }
};

// Face integral of one (source, face) pair: INTEBc9kmdclap (Macro),
// INTEBc9modHelm (Macro, Helm accumulator) or INTEBc9sdkm (no subdivision).
template <bool Macro, class Acc>
__device__ void integrate_face(const QuadDev &q, const double *p, const double (&c)[4][3], int isrc, double minedge,
                               Acc &acc) {
  const double PI = 3.141592653589793, PI2 = 2.0 * PI;
  int ndx = 1, ndy = 1;
  if (Macro) {
    double dd[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int f = (e + 1) & 3;
      double dx = c[e][0] - c[f][0], dy = c[e][1] - c[f][1], dz = c[e][2] - c[f][2];
      dd[e] = sqrt(dx * dx + dy * dy + dz * dz);
    }
    int t1 = (int)(fmax(dd[0], dd[2]) / minedge + 0.5), t2 = (int)(fmax(dd[1], dd[3]) / minedge + 0.5);
    ndx = q.iiDiv > t1 ? q.iiDiv : t1;
    ndy = q.iiDiv > t2 ? q.iiDiv : t2;
  }
  for (int ix = 1; ix <= ndx; ++ix)
    for (int iy = 1; iy <= ndy; ++iy) {
      double a = -1.0 + (ix - 1) * 2.0 / ndx, b = -1.0 + ix * 2.0 / ndx;
      double cc = -1.0 + (iy - 1) * 2.0 / ndy, d = -1.0 + iy * 2.0 / ndy;
      double dex = 0.25 * (b - a) * (d - cc);
      bool ising = false;
      double xx1 = 0, yy1 = 0;
      if (isrc != 0) {
        xx1 = ((-1.0 + 2.0 * kKsi13[isrc - 1]) - a) / (b - a);
        yy1 = ((-1.0 + 2.0 * kEta13[isrc - 1]) - cc) / (d - cc);
        ising = xx1 >= 0.0 && xx1 <= 1.0 && yy1 >= 0.0 && yy1 <= 1.0;
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
          for (int I = 0; I < q.nbs; ++I) {
            double th = (th1 - th0) * q.gbs[I] / 2.0 + (th1 + th0) / 2.0;
            double roth = al * sin(fii) / sin(th - th0 + fii);
            double cth = cos(th), sth = sin(th);
            for (int J = 0; J < q.nbs; ++J) {
              double ro = roth * q.gbs[J] / 2.0 + roth / 2.0;
              double eti = 2.0 * (xx1 + ro * cth) - 1.0;
              double etj = 2.0 * (yy1 + ro * sth) - 1.0;
              eti = a + 0.5 * (eti + 1.0) * (b - a);
              etj = cc + 0.5 * (etj + 1.0) * (d - cc);
              acc.add(p, c, eti, etj, dex * (th1 - th0) * roth * ro * q.obs[I] * q.obs[J]);
            }
          }
        }
      } else {
        for (int I = 0; I < q.nbr; ++I) {
          double gii = a + 0.5 * (q.gbr[I] + 1.0) * (b - a);
          for (int J = 0; J < q.nbr; ++J) {
            double gij = cc + 0.5 * (q.gbr[J] + 1.0) * (d - cc);
            acc.add(p, c, gii, gij, dex * q.obr[I] * q.obr[J]);
          }
        }
      }
    }
}

}  // namespace tt
