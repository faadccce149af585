/* Oracle (test infrastructure only; nothing in the library links or calls this
 * file): single-domain modified-Helmholtz path of the reference (EQNS imH 3-7).
 *
 *   or_mhsd_integrals  FMATsdmHfs            lib_fm.f:1183-1645
 *   or_mhsd_bprime     MakeBprime            lib_fm.f:800-858
 *   or_mhsd_sysm       fmmHSysM              lib_fm.f:1061-1181
 *   or_mhsd_rhs        fmSolveModHelmDC's right-hand side, lib_fm.f:905-1026
 *   or_mhsd_rhs_mv     the matrix-vector form of fmSolveModHelmDCaca/wawt/
 *                      chmat/mm (lib_fm.f:2-400 ...): eight products per row set
 *
 * Rows: U rows are the nodes 1..nn (source x(it), mu = sqrt(beta/diff(it)));
 * Q rows the boundary flux points, element-major: row (e, f) = 4 e + f with
 * source xq(ibcf(e, f)) and mu = sqrt(beta/qdiff(ibcf(e, f))).  Matrix order
 * as the reference's integral file: H, G, B, Bp, Abdx, Abdy, Abdz, Dx, Dy, Dz;
 * column-major, ld = the row count (nn for U, nf = 4 ne for Q).
 *
 * Pins: the reference's own tri.modHelm.full.int (WritemHfullIntDisk,
 * lib_fm.f:1822-1895) of the mhsd cases in tests/golden; the solve chain
 * against the reference's mH field (tri.dat) and RMS error (tri.mH-rms.dat)
 * of the same runs. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "tt_oracle.h"

#define IBC(e, k) (ibc[(e) + (size_t)ne * ((k)-1)])
#define IDC(c, k) (idc[(c) + (size_t)nc * ((k)-1)])
#define IBF(c, s, f) (ibf[(c) + (size_t)nc * (((s)-1) + 6 * ((f)-1))])
#define IBCF(e, f) (ibcf[(e) + (size_t)ne * ((f)-1)])
#define X(n, d) (x[((n)-1) + (size_t)nn * (d)])
#define XQ(n, d) (xq[((n)-1) + (size_t)nq * (d)])

static const int k_geo8m[8] = {6, 7, 2, 3, 5, 8, 1, 4};

/* FindMinEdgeOneElem (lib_fm.f:1920-1951): shortest of the four corner-to-corner edges */
static double min_edge_elem(int nn, int ne, const double *x, const int *ibc, int e) {
  const int k[5] = {1, 3, 5, 7, 1};
  double m = 1.0e20;
  for (int s = 0; s < 4; ++s) {
    const int a = IBC(e, k[s]), b = IBC(e, k[s + 1]);
    const double dx = X(a, 0) - X(b, 0), dy = X(a, 1) - X(b, 1), dz = X(a, 2) - X(b, 2);
    const double l = sqrt(dx * dx + dy * dy + dz * dz);
    if (l < m) m = l;
  }
  return m;
}

/* The integrals of one row over every boundary element and every cell
 * (lib_fm.f:1283-1378 for U rows, :1503-1603 for Q rows), added into the
 * row's entries in element and cell order. */
static void row_integrals(const or_quad *q, int nn, int nb, int ne, int nc, int nq, const double *x, const int *ibc,
                          const int *idc, const int *ibf, const int *ibcf, const int *lbn, const double p[3],
                          double mu, int unode, int qnode, size_t r, size_t ld, double **M) {
  double he[9], ge[4], abx[9], aby[9], abz[9], heyz[9], hezx[9], hexy[9];
  double be[27], dxe[27], dye[27], dze[27];
  for (int e = 0; e < ne; ++e) {
    double c[4][3];
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) c[k][d] = X(IBC(e, 2 * k + 1), d);
    int isrc = 0;
    if (unode) {
      for (int s = 1; s <= 9; ++s)
        if (IBC(e, s) == unode) isrc = s;
    } else {
      for (int s = 1; s <= 4; ++s)
        if (IBCF(e, s) == qnode) isrc = s + 9;
    }
    const double minedge = min_edge_elem(nn, ne, x, ibc, e);
    or_inteb9_mh(q, p, c, isrc, minedge, mu, he, ge, abx, aby, abz, heyz, hezx, hexy);
    for (int s = 1; s <= 9; ++s) {
      const size_t j = (size_t)lbn[IBC(e, s) - 1] - 1;
      M[0][r + ld * j] += he[s - 1];
      M[4][r + ld * j] += abx[s - 1];
      M[5][r + ld * j] += aby[s - 1];
      M[6][r + ld * j] += abz[s - 1];
    }
    for (int s = 0; s < 4; ++s) M[1][r + ld * ((size_t)e * 4 + s)] += ge[s];
  }
  for (int ic = 0; ic < nc; ++ic) {
    double c[8][3];
    for (int k = 0; k < 8; ++k)
      for (int d = 0; d < 3; ++d) c[k][d] = X(IDC(ic, k_geo8m[k]), d);
    int isrc = 0;
    if (unode) {
      for (int s = 1; s <= 27; ++s)
        if (IDC(ic, s) == unode) isrc = s;
    } else {
      for (int fc = 1; fc <= 6; ++fc)
        for (int s = 1; s <= 4; ++s)
          if (IBF(ic, fc, s) == qnode) isrc = 27 + (fc - 1) * 4 + s;
    }
    or_inted27_mh(q, p, c, isrc, mu, be, dxe, dye, dze);
    for (int s = 1; s <= 27; ++s) {
      const size_t j = (size_t)IDC(ic, s) - 1;
      M[2][r + ld * j] += be[s - 1];
      M[7][r + ld * j] += dxe[s - 1];
      M[8][r + ld * j] += dye[s - 1];
      M[9][r + ld * j] += dze[s - 1];
    }
  }
}

void or_mhsd_integrals(const or_quad *q, int nn, int nb, int ne, int nc, int nq, const double *x, const int *ibc,
                       const int *idc, const int *ibf, const double *xq, const int *ibcf, const int *gbn,
                       const int *lbn, const double *diff, const double *qdiff, double beta, double **U, double **Q,
                       double ont[5], double *mumax, int nthreads) {
  const size_t nf = (size_t)ne * 4;
  const size_t ucols[10] = {nb, nf, nn, nn, nb, nb, nb, nn, nn, nn};
  for (int k = 0; k < 10; ++k) {
    memset(U[k], 0, sizeof(double) * nn * ucols[k]);
    memset(Q[k], 0, sizeof(double) * nf * ucols[k]);
  }
  double oH = 0.0, oG = 0.0, oDx = 0.0, oDy = 0.0, oDz = 0.0, mm = 0.0;
  /* U source points (lib_fm.f:1280-1496) */
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(max : oH, oDx, oDy, oDz, mm)
  for (int it = 1; it <= nn; ++it) {
    const size_t r = (size_t)it - 1, ld = nn;
    const double mu = sqrt(beta / diff[it - 1]);
    if (mu > mm) mm = mu;
    const double p[3] = {X(it, 0), X(it, 1), X(it, 2)};
    row_integrals(q, nn, nb, ne, nc, nq, x, ibc, idc, ibf, ibcf, lbn, p, mu, it, 0, r, ld, U);
    /* free term for u = 1, q = 0, f = -mu^2 (:1381-1395) */
    double vH = 0.0, vB = 0.0;
    for (int j = 0; j < nb; ++j) vH = vH + U[0][r + ld * j];
    for (int j = 0; j < nn; ++j) vB = vB + U[2][r + ld * j];
    const double c = -vH + mu * mu * vB;
    if (lbn[it - 1] > 0) {
      U[0][r + ld * ((size_t)lbn[it - 1] - 1)] += c;
      if (it == 1 && fabs(0.125 - c) > oH) oH = fabs(0.125 - c);
    }
    /* linear fields (z, 0, 0), (0, x, 0), (0, 0, y) (:1397-1444): D_z, D_x, D_y in that order */
    const int crd[3] = {2, 0, 1}, mat[3] = {9, 7, 8};
    for (int t = 0; t < 3; ++t) {
      double wH = 0.0, wB = 0.0, wD = 0.0;
      for (int j = 0; j < nb; ++j) wH = wH + U[0][r + ld * j] * X(gbn[j], crd[t]);
      for (int j = 0; j < nn; ++j) {
        wB = wB + U[2][r + ld * j] * X(j + 1, crd[t]);
        wD = wD + U[mat[t]][r + ld * j];
      }
      double vs = wH - mu * mu * wB - wD;
      if (lbn[it - 1] == 0) vs = vs + X(it, crd[t]);
      const double a = fabs(vs);
      if (t == 0 && a > oDz) oDz = a;
      if (t == 1 && a > oDx) oDx = a;
      if (t == 2 && a > oDy) oDy = a;
      U[mat[t]][r + ld * r] = U[mat[t]][r + ld * r] + vs;
    }
  }
  /* Q source points (lib_fm.f:1502-1637) */
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(max : oG, mm)
  for (int e = 0; e < ne; ++e)
    for (int f = 1; f <= 4; ++f) {
      const int it = IBCF(e, f);
      const size_t r = (size_t)e * 4 + (f - 1), ld = nf;
      const double mu = sqrt(beta / qdiff[it - 1]);
      if (mu > mm) mm = mu;
      const double p[3] = {XQ(it, 0), XQ(it, 1), XQ(it, 2)};
      row_integrals(q, nn, nb, ne, nc, nq, x, ibc, idc, ibf, ibcf, lbn, p, mu, 0, it, r, ld, Q);
      double vH = 0.0, vB = 0.0;
      for (int j = 0; j < nb; ++j) vH = vH + Q[0][r + ld * j];
      for (int j = 0; j < nn; ++j) vB = vB + Q[2][r + ld * j];
      const double c = -vH + mu * mu * vB;
      if (fabs(0.5 - c) > oG) oG = fabs(0.5 - c);
      /* c spread over the element's nine nodes by phi_9 at the flux point (xieta34, :1624-1634) */
      const double tc = 3.0 / 4.0;
      const double xi = (f == 1 || f == 4) ? -tc : tc, eta = (f <= 2) ? -tc : tc;
      double fi[9];
      or_shape9(xi, eta, fi);
      for (int s = 1; s <= 9; ++s) {
        const size_t j = (size_t)lbn[IBC(e, s) - 1] - 1;
        Q[0][r + ld * j] = Q[0][r + ld * j] + c * fi[s - 1];
      }
    }
  ont[0] = oH, ont[1] = oG, ont[2] = oDx, ont[3] = oDy, ont[4] = oDz;
  *mumax = mm;
}

void or_mhsd_bprime(int nn, int ne, const int *ibcf, const double *diff, const double *qdiff, double beta,
                    const double *UB, double *UBp, const double *QB, double *QBp) {
  const size_t nf = (size_t)ne * 4;
  for (int row = 0; row < nn; ++row) {
    const double mu = sqrt(beta / diff[row]);
    for (int col = 0; col < nn; ++col) {
      const double mun = sqrt(beta / diff[col]), t = mu / mun;
      UBp[row + (size_t)nn * col] = UB[row + (size_t)nn * col] * beta * (t * t - 1.0);
    }
  }
  for (int e = 0; e < ne; ++e)
    for (int f = 1; f <= 4; ++f) {
      const size_t row = (size_t)e * 4 + f - 1;
      const double mu = sqrt(beta / qdiff[IBCF(e, f) - 1]);
      for (int col = 0; col < nn; ++col) {
        const double mun = sqrt(beta / diff[col]), t = mu / mun;
        QBp[row + nf * col] = QB[row + nf * col] * beta * (t * t - 1.0);
      }
    }
}

/* fmmHSysM: A (nunk x nunk) and RHS (nunk x nknown), column-major. */
void or_mhsd_sysm(int nn, int nb, int ne, const int *ibcf, const int *lbn, const int *kode, const int *qkode,
                  int nunk, int nknown, const double *diff, const double *qdiff, const double *UH, const double *UG,
                  const double *QH, const double *QG, double *A, double *RHS) {
  const size_t nf = (size_t)ne * 4;
  for (size_t k = 0; k < (size_t)nunk * nunk; ++k) A[k] = -99999.999;
  for (size_t k = 0; k < (size_t)nunk * nknown; ++k) RHS[k] = -99999.999;
  int ieq = 0;
  for (int pass = 0; pass < 2; ++pass) {
    const size_t nrow = pass == 0 ? (size_t)nn : nf;
    const double *H = pass == 0 ? UH : QH, *G = pass == 0 ? UG : QG;
    for (size_t row = 0; row < nrow; ++row) {
      if ((pass == 0 ? kode[row] : qkode[row]) >= 0) continue;
      const size_t e = (size_t)ieq++;
      for (int j = 0; j < nn; ++j) {
        double val = 0.0;
        if (lbn[j] != 0) val = pass == 0 ? val + diff[j] * H[row + nrow * (lbn[j] - 1)]
                                        : diff[j] * H[row + nrow * (lbn[j] - 1)];
        const int nuq = kode[j];
        if (nuq < 0) A[e + (size_t)nunk * (-nuq - 1)] = val;
        else RHS[e + (size_t)nunk * (nuq - 1)] = -val;
      }
      for (int ii = 0; ii < ne; ++ii)
        for (int jj = 1; jj <= 4; ++jj) {
          const int j = IBCF(ii, jj);
          const size_t col = (size_t)ii * 4 + jj - 1;
          const int nuq = qkode[col];
          if (nuq < 0) A[e + (size_t)nunk * (-nuq - 1)] = -G[row + nrow * col] * qdiff[j - 1];
          else RHS[e + (size_t)nunk * (nuq - 1)] = G[row + nrow * col] * qdiff[j - 1];
        }
      if (pass == 0 && lbn[row] == 0) A[e + (size_t)nunk * e] = A[e + (size_t)nunk * e] + 1.0 * diff[row];
    }
  }
  (void)nb;
}

/* b (known values) and r = RHS b of fmSolveModHelmDC (lib_fm.f:905-925) */
static void known_and_rhsb(int nn, int nf, const int *kode, const int *qkode, int nunk, int nknown,
                           const double *u, const double *qu, const double *RHS, double *r) {
  double *b = malloc(sizeof(double) * (nknown > 0 ? nknown : 1));
  for (int i = 0; i < nn; ++i)
    if (kode[i] > 0) b[kode[i] - 1] = u[i];
  for (int i = 0; i < nf; ++i)
    if (qkode[i] > 0) b[qkode[i] - 1] = qu[i];
  for (int i = 0; i < nunk; ++i) {
    double s = 0.0;
    for (int k = 0; k < nknown; ++k) s = s + RHS[i + (size_t)nunk * k] * b[k];
    r[i] = s;
  }
  free(b);
}

/* fmSolveModHelmDC's right-hand side r (nunk): RHS b, then per equation row
 * the domain and boundary terms in column order (lib_fm.f:929-1016). */
void or_mhsd_rhs(int nn, int nb, int ne, const int *ibcf, const int *gbn, const int *kode, const int *qkode,
                 int nunk, int nknown, const double *diff, const double *qdiff, double beta, const double *u,
                 const double *qu, const double *velocity, const double *rhsv, const double *dgrad, double **U,
                 double **Q, const double *RHS, double *r) {
  const int nf = ne * 4;
  known_and_rhsb(nn, nf, kode, qkode, nunk, nknown, u, qu, RHS, r);
  int ieq = 0;
  for (int pass = 0; pass < 2; ++pass) {
    const size_t nrow = pass == 0 ? (size_t)nn : (size_t)nf;
    double **M = pass == 0 ? U : Q;
    for (size_t row = 0; row < nrow; ++row) {
      if ((pass == 0 ? kode[row] : qkode[row]) >= 0) continue;
      const int e = ieq++;
      const double mu = pass == 0 ? sqrt(beta / diff[row])
                                  : sqrt(beta / qdiff[IBCF((int)(row / 4), (int)(row % 4) + 1) - 1]);
      double s = r[e];
      for (int col = 0; col < nn; ++col) {
        const double mun = sqrt(beta / diff[col]), t = mu / mun;
        const size_t k = row + nrow * col;
        s = s - M[2][k] * rhsv[col] -
            u[col] * (-M[7][k] * (velocity[col] + dgrad[col]) - M[8][k] * (velocity[col + nn] + dgrad[col + nn]) -
                      M[9][k] * (velocity[col + 2 * nn] + dgrad[col + 2 * nn]) - M[2][k] * beta * (t * t - 1.0));
      }
      for (int col = 0; col < nb; ++col) {
        const int node = gbn[col] - 1;
        const size_t k = row + nrow * col;
        s = s - u[node] * (M[4][k] * velocity[node] + M[5][k] * velocity[node + nn] +
                           M[6][k] * velocity[node + 2 * nn]);
      }
      r[e] = s;
    }
  }
}

/* The product vectors of the matrix-vector variants (lib_fm.f:269-301):
 * v[0..4] (nn): -rhsv, u, u (v + grad diff)_x,y,z; v[5..7] (nb): -u v_x,y,z on the boundary. */
void or_mhsd_vectors(int nn, int nb, const int *gbn, const double *u, const double *velocity, const double *rhsv,
                     const double *dgrad, double **v) {
  for (int col = 0; col < nn; ++col) {
    v[0][col] = -rhsv[col];
    v[1][col] = u[col];
    v[2][col] = u[col] * (velocity[col] + dgrad[col]);
    v[3][col] = u[col] * (velocity[col + nn] + dgrad[col + nn]);
    v[4][col] = u[col] * (velocity[col + 2 * nn] + dgrad[col + 2 * nn]);
  }
  for (int col = 0; col < nb; ++col) {
    const int node = gbn[col] - 1;
    v[5][col] = -u[node] * velocity[node];
    v[6][col] = -u[node] * velocity[node + nn];
    v[7][col] = -u[node] * velocity[node + 2 * nn];
  }
}

/* r of the matrix-vector variants: RHS b plus, per equation row, the eight
 * products y[k][row] (B, Bp, Dx, Dy, Dz, Abdx, Abdy, Abdz) added in that
 * order (lib_fm.f:322-361). */
void or_mhsd_rhs_mv(int nn, int ne, const int *kode, const int *qkode, int nunk, int nknown, const double *u,
                    const double *qu, const double *RHS, double **yu, double **yq, double *r) {
  const int nf = ne * 4;
  known_and_rhsb(nn, nf, kode, qkode, nunk, nknown, u, qu, RHS, r);
  int ieq = 0;
  for (int row = 0; row < nn; ++row)
    if (kode[row] < 0) {
      r[ieq] = r[ieq] + yu[0][row] + yu[1][row] + yu[2][row] + yu[3][row] + yu[4][row] + yu[5][row] + yu[6][row] +
               yu[7][row];
      ++ieq;
    }
  for (int row = 0; row < nf; ++row)
    if (qkode[row] < 0) {
      r[ieq] = r[ieq] + yq[0][row] + yq[1][row] + yq[2][row] + yq[3][row] + yq[4][row] + yq[5][row] + yq[6][row] +
               yq[7][row];
      ++ieq;
    }
}
