/*
 * or_picard.c -- CPU restatement of the reference's steady Picard iteration.
 * TEST INFRASTRUCTURE ONLY (the oracle): see tt_oracle.h.
 *
 *   setup order       main.f90:410-513 (sdkm), :791-875 (macro, VTE CSR), :956-963 (T),
 *                     :1057-1061 (KM)
 *   time/Picard loop  main.f90:1145-1589
 *   sdKm_sysm         lib_sdkm.f:1171-1210,  sdkm_solve/rhs/distunk lib_sdkm.f:23-121,349-519,1075-1096
 *   SolvEQNfm slvt=0  lib_Solve_v11.f:267-402 (zero rhs shortcut, dgesl)
 *   sMat2crsSysRhsB_count  lib_solvFDdaeVD.f:1307-1474
 *   sMat2crsSysRhsB_w_fill lib_solvFDwTE.f:593-752, _Temp_fill lib_solvFDdae.f:289-371
 *   sMat2crsSysRhsKM3 lib_km.f:561-675, solv_kmMacro3/rhsKM3/SetUpXknv3 lib_km.f:214-415
 *   SolveFDwTE        lib_solvFDwTE.f:190-332, BouyancyRHS :505-592
 *   SolveFDdae        lib_solvFDdae.f:168-288
 *   CRSxV             lib_util.f:1351-1372
 *   CheckConvergencevw lib_util.f:1048-1150, DLSE lib_util.f:443-475
 *   SetTimeScheme     lib_util.f:2-32 (scheme 1)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "tt_oracle.h"

typedef struct {
  int neq, nnz;
  int *i, *j;  /* 1-based */
  double *v;
  int *d;      /* fill map */
} csr;

static void csr_free(csr *m) {
  free(m->i);
  free(m->j);
  free(m->v);
  free(m->d);
  memset(m, 0, sizeof(*m));
}

static void crsxv(const csr *m, const double *vek, double *rez) {
  for (int i = 0; i < m->neq; ++i) {
    rez[i] = 0.0;
    for (int j = m->i[i] - 1; j < m->i[i + 1] - 1; ++j) rez[i] = rez[i] + m->v[j] * vek[m->j[j] - 1];
  }
}

/* sMat2crsSysRhsB_count: pattern with fill maps; rows = nc*51 */
static void csr_count(const or_case *c, const int *wkode, const int *wqkode, int nunk, int nb, csr *sys,
                      csr *rhs) {
  int nc = c->nc, neq = nc * 51;
  sys->neq = rhs->neq = neq;
  sys->nnz = rhs->nnz = 0;
  for (int ic = 0; ic < nc; ++ic)
    for (int it = 0; it < 51; ++it) {
      for (int col = 0; col < 27; ++col) {
        int nuq = wkode[c->idc[ic + (size_t)nc * col] - 1];
        if (nuq < 0) sys->nnz++; else rhs->nnz++;
      }
      for (int s = 0; s < 6; ++s)
        for (int jj = 0; jj < 4; ++jj) {
          int nuq = wqkode[c->ibf[ic + (size_t)nc * (s + 6 * jj)] - 1];
          if (nuq < 0) sys->nnz++; else rhs->nnz++;
        }
    }
  sys->i = malloc(sizeof(int) * (neq + 1));
  rhs->i = malloc(sizeof(int) * (neq + 1));
  sys->j = malloc(sizeof(int) * sys->nnz);
  rhs->j = malloc(sizeof(int) * rhs->nnz);
  sys->d = malloc(sizeof(int) * sys->nnz);
  rhs->d = malloc(sizeof(int) * rhs->nnz);
  sys->v = calloc(sys->nnz, sizeof(double));
  rhs->v = calloc(rhs->nnz, sizeof(double));
  int *frs = calloc(nunk + 1, sizeof(int)), *frr = calloc(nb + 1, sizeof(int));
  int ks = 0, kr = 0, is = 0, ir = 0;
  for (int ic = 0; ic < nc; ++ic)
    for (int it = 0; it < 51; ++it) {
      int row = ic * 51 + it;
      int used_s[51], used_r[51], nus = 0, nur = 0;
      for (int col = 0; col < 27; ++col) {
        int nuq = wkode[c->idc[ic + (size_t)nc * col] - 1];
        if (nuq < 0) { frs[-nuq] = ++is; used_s[nus++] = -nuq; }
        else { frr[nuq] = ++ir; used_r[nur++] = nuq; }
      }
      for (int s = 0; s < 6; ++s)
        for (int jj = 0; jj < 4; ++jj) {
          int nuq = wqkode[c->ibf[ic + (size_t)nc * (s + 6 * jj)] - 1];
          if (nuq < 0) { frs[-nuq] = ++is; used_s[nus++] = -nuq; }
          else { frr[nuq] = ++ir; used_r[nur++] = nuq; }
        }
      /* the reference scans the dense row 1..nunk: columns come out sorted */
      for (int a = 1; a < nus; ++a)
        for (int b = a; b > 0 && used_s[b - 1] > used_s[b]; --b) { int t = used_s[b]; used_s[b] = used_s[b - 1]; used_s[b - 1] = t; }
      for (int a = 1; a < nur; ++a)
        for (int b = a; b > 0 && used_r[b - 1] > used_r[b]; --b) { int t = used_r[b]; used_r[b] = used_r[b - 1]; used_r[b - 1] = t; }
      sys->i[row] = ks + 1;
      rhs->i[row] = kr + 1;
      for (int a = 0; a < nus; ++a) {
        ++ks;
        sys->d[frs[used_s[a]] - 1] = ks;
        sys->j[ks - 1] = used_s[a];
        frs[used_s[a]] = 0;
      }
      for (int a = 0; a < nur; ++a) {
        ++kr;
        rhs->d[frr[used_r[a]] - 1] = kr;
        rhs->j[kr - 1] = used_r[a];
        frr[used_r[a]] = 0;
      }
    }
  sys->i[neq] = ks + 1;
  rhs->i[neq] = kr + 1;
  free(frs);
  free(frr);
}

/* sMat2crsSysRhsB_w_fill / _Temp_fill: H + B*beta*ren on u columns, -G on q columns */
static void csr_fill(const or_case *c, const int *wkode, const int *wqkode, const double *H, const double *G,
                     const double *B, double beta, double ren, csr *sys, csr *rhs) {
  int nc = c->nc;
  size_t m1 = (size_t)nc * 51;
  int is = 0, ir = 0;
  for (int ic = 0; ic < nc; ++ic)
    for (int it = 0; it < 51; ++it) {
      size_t row = (size_t)ic * 51 + it;
      for (int col = 0; col < 27; ++col) {
        int j = c->idc[ic + (size_t)nc * col];
        double val = H[row + m1 * col] + B[row + m1 * col] * beta * ren;
        int nuq = wkode[j - 1];
        if (nuq < 0) sys->v[sys->d[is++] - 1] = val;
        else rhs->v[rhs->d[ir++] - 1] = -val;
      }
      for (int s = 0; s < 6; ++s)
        for (int jj = 0; jj < 4; ++jj) {
          int col = s * 4 + jj;
          int j = c->ibf[ic + (size_t)nc * (s + 6 * jj)];
          int nuq = wqkode[j - 1];
          if (nuq < 0) sys->v[sys->d[is++] - 1] = -G[row + m1 * col];
          else rhs->v[rhs->d[ir++] - 1] = G[row + m1 * col];
        }
    }
}

/* sMat2crsSysRhsKM3: u-source rows only, exact zeros dropped */
static void csr_km(const or_case *c, const double *H, csr *sys, csr *rhs) {
  int nc = c->nc, neq = nc * 27;
  size_t m1 = (size_t)nc * 51;
  sys->neq = rhs->neq = neq;
  sys->nnz = rhs->nnz = 0;
  for (int ic = 0; ic < nc; ++ic)
    for (int it = 0; it < 27; ++it) {
      size_t row = (size_t)ic * 51 + it;
      for (int col = 0; col < 27; ++col) {
        int nuq = c->kmu[c->idc[ic + (size_t)nc * col] - 1];
        if (H[row + m1 * col] != 0.0) { if (nuq < 0) sys->nnz++; else rhs->nnz++; }
      }
    }
  sys->i = malloc(sizeof(int) * (neq + 1));
  rhs->i = malloc(sizeof(int) * (neq + 1));
  sys->j = malloc(sizeof(int) * (sys->nnz + 1));
  rhs->j = malloc(sizeof(int) * (rhs->nnz + 1));
  sys->v = malloc(sizeof(double) * (sys->nnz + 1));
  rhs->v = malloc(sizeof(double) * (rhs->nnz + 1));
  sys->d = rhs->d = NULL;
  int ks = 0, kr = 0, crow = 0;
  for (int ic = 0; ic < nc; ++ic)
    for (int it = 0; it < 27; ++it) {
      size_t row = (size_t)ic * 51 + it;
      int cs[27], cr[27], ns = 0, nr = 0;
      double vs[27], vr[27];
      for (int col = 0; col < 27; ++col) {
        int nuq = c->kmu[c->idc[ic + (size_t)nc * col] - 1];
        double h = H[row + m1 * col];
        if (nuq < 0) {
          int k;  /* FRsys(ABS(nuq)) = smatH: later column overwrites */
          for (k = 0; k < ns && cs[k] != -nuq; ++k) {}
          if (k == ns) { cs[ns] = -nuq; vs[ns++] = h; } else vs[k] = h;
        } else {
          int k;
          for (k = 0; k < nr && cr[k] != nuq; ++k) {}
          if (k == nr) { cr[nr] = nuq; vr[nr++] = -h; } else vr[k] = -h;
        }
      }
      for (int a = 1; a < ns; ++a)
        for (int b = a; b > 0 && cs[b - 1] > cs[b]; --b) {
          int t = cs[b]; cs[b] = cs[b - 1]; cs[b - 1] = t;
          double tv = vs[b]; vs[b] = vs[b - 1]; vs[b - 1] = tv;
        }
      for (int a = 1; a < nr; ++a)
        for (int b = a; b > 0 && cr[b - 1] > cr[b]; --b) {
          int t = cr[b]; cr[b] = cr[b - 1]; cr[b - 1] = t;
          double tv = vr[b]; vr[b] = vr[b - 1]; vr[b - 1] = tv;
        }
      sys->i[crow] = ks + 1;
      rhs->i[crow] = kr + 1;
      for (int a = 0; a < ns; ++a)
        if (vs[a] != 0.0) { sys->v[ks] = vs[a]; sys->j[ks] = cs[a]; ++ks; }
      for (int a = 0; a < nr; ++a)
        if (vr[a] != 0.0) { rhs->v[kr] = vr[a]; rhs->j[kr] = cr[a]; ++kr; }
      ++crow;
    }
  sys->i[crow] = ks + 1;
  rhs->i[crow] = kr + 1;
  sys->nnz = ks;
  rhs->nnz = kr;
}

void or_sdkm_rhs(int eqn, int nn, int nb, const int *lbn, const int *gbn, const int *kode, const double *nx,
                 const double *ny, const double *nz, const double *H, const double *Htx, const double *Hty,
                 const double *Htz, const double *Dx, const double *Dy, const double *Dz, int ld,
                 const double *vel, const double *vort, double *dfi, int nthreads) {
#define V(k, d) vel[(k) + (size_t)nn * (d)]
#define W(k, d) vort[(k) + (size_t)nn * (d)]
#define S(A, i, j) A[(i) + (size_t)ld * (j)]
  (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int i = 0; i < nb; ++i) {
    double d = 0.0;
    for (int k = 0; k < nn; ++k) {
      int j = lbn[k];
      if (j != 0) {
        --j;
        if (eqn == 1)
          d = d + S(H, i, j) * (ny[i] * V(k, 2) - nz[i] * V(k, 1)) - S(Htx, i, j) * (ny[i] * V(k, 1) + nz[i] * V(k, 2)) +
              V(k, 0) * (S(Hty, i, j) * ny[i] + S(Htz, i, j) * nz[i]) +
              S(Dx, i, k) * (+nx[i] * W(k, 0) + ny[i] * W(k, 1) + nz[i] * W(k, 2));
        else if (eqn == 2)
          d = d + S(H, i, j) * (nz[i] * V(k, 0) - nx[i] * V(k, 2)) - S(Hty, i, j) * (nx[i] * V(k, 0) + nz[i] * V(k, 2)) +
              V(k, 1) * (S(Htx, i, j) * nx[i] + S(Htz, i, j) * nz[i]) +
              S(Dy, i, k) * (nx[i] * W(k, 0) + ny[i] * W(k, 1) + nz[i] * W(k, 2));
        else
          d = d + S(H, i, j) * (nx[i] * V(k, 1) - ny[i] * V(k, 0)) - S(Htz, i, j) * (nx[i] * V(k, 0) + ny[i] * V(k, 1)) +
              V(k, 2) * (S(Htx, i, j) * nx[i] + S(Hty, i, j) * ny[i]) +
              S(Dz, i, k) * (nx[i] * W(k, 0) + ny[i] * W(k, 1) + nz[i] * W(k, 2));
      } else {
        if (eqn == 1)
          d = d + S(Dx, i, k) * (ny[i] * W(k, 1) + nz[i] * W(k, 2)) - W(k, 0) * (S(Dy, i, k) * ny[i] + S(Dz, i, k) * nz[i]);
        else if (eqn == 2)
          d = d + S(Dy, i, k) * (nx[i] * W(k, 0) + nz[i] * W(k, 2)) - W(k, 1) * (S(Dx, i, k) * nx[i] + S(Dz, i, k) * nz[i]);
        else
          d = d + S(Dz, i, k) * (nx[i] * W(k, 0) + ny[i] * W(k, 1)) - W(k, 2) * (S(Dx, i, k) * nx[i] + S(Dy, i, k) * ny[i]);
      }
    }
    dfi[i] = d;
  }
  for (int i = 0; i < nb; ++i) {
    int kd = kode[i + (size_t)nb * (eqn - 1)];
    if (kd == 1 || kd == 2) dfi[i] = W(gbn[i] - 1, eqn - 1);
  }
#undef V
#undef W
#undef S
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int or_picard(const or_case *c, double *vel, double *vort, double *qvort, double *temp, double *qtemp, int maxhist,
              double *sta, int *ite, int nthreads, double *seconds_per_iter) {
  int nn = c->nn, nb = c->nb, nc = c->nc, nq = c->nq;
  size_t m1 = (size_t)nc * 51, m2 = (size_t)nc * 27;
  /* ---- macro integrals ---- */
  double *mH = malloc(sizeof(double) * m1 * 27), *mG = malloc(sizeof(double) * m1 * 24);
  double *mB = malloc(sizeof(double) * m1 * 27), *mAx = malloc(sizeof(double) * m1 * 27);
  double *mAy = malloc(sizeof(double) * m1 * 27), *mAz = malloc(sizeof(double) * m1 * 27);
  double *mHtx = malloc(sizeof(double) * m2 * 27), *mHty = malloc(sizeof(double) * m2 * 27);
  double *mHtz = malloc(sizeof(double) * m2 * 27), *mDx = malloc(sizeof(double) * m1 * 27);
  double *mDy = malloc(sizeof(double) * m1 * 27), *mDz = malloc(sizeof(double) * m1 * 27);
  double ontm[11], onts[10];
  or_macro_integrals(&c->quad, nn, nc, nq, c->x, c->idc, c->ibf, c->xq, c->flpm, c->copy, mH, mG, mB, mAx, mAy,
                     mAz, mHtx, mHty, mHtz, mDx, mDy, mDz, ontm, nthreads);
  /* ---- single-domain kinematics ---- */
  /* (main.f90:410: only with EQNS iBw != 0) */
  const int kin = c->iBw == 1;
  const size_t nbb = kin ? (size_t)nb : 0;
  double *sH = malloc(sizeof(double) * nbb * nb + 8), *sHtx = malloc(sizeof(double) * nbb * nb + 8);
  double *sHty = malloc(sizeof(double) * nbb * nb + 8), *sHtz = malloc(sizeof(double) * nbb * nb + 8);
  double *sDx = malloc(sizeof(double) * nbb * nn + 8), *sDy = malloc(sizeof(double) * nbb * nn + 8);
  double *sDz = malloc(sizeof(double) * nbb * nn + 8);
  if (kin)
    or_sd_integrals(&c->quad, nn, nb, c->ne, nc, c->x, c->ibc, c->idc, c->gbn, c->lbn, 0, nb, nb, sH, sHtx, sHty,
                  sHtz, sDx, sDy, sDz, onts, nthreads);
  double *sA[3], *piv[3];
  for (int e = 0; e < 3; ++e) {
    sA[e] = malloc(sizeof(double) * nbb * nb + 8);
    piv[e] = malloc(sizeof(double) * nb);
    if (!kin) continue;
    for (int j = 0; j < nb; ++j) {
      int dn = c->gbn[j] - 1;
      for (int i = 0; i < nb; ++i) {
        int kd = c->kode[i + (size_t)nb * e];
        double v;
        if (kd == 0 || kd == 3)
          v = c->nx[i] * sDx[i + (size_t)nb * dn] + c->ny[i] * sDy[i + (size_t)nb * dn] + c->nz[i] * sDz[i + (size_t)nb * dn];
        else
          v = (i != j) ? 0.0 : 1.0;
        sA[e][i + (size_t)nb * j] = v;
      }
    }
    or_dgefa(sA[e], nb, nb, piv[e]);
  }
  /* ---- time scheme (TimeScheme 1) and CSR systems ---- */
  double beta = 1.50 / c->tstep, beta2 = 2.00 / c->tstep, beta3 = 0.50 / c->tstep;
  double ren = c->ReN;
  if (c->iDT > 0) ren = 1.0 / c->PrN;
  csr vs[3], vr[3], ts_, tr_, ks_, kr_;
  double *vq[3], *tq = NULL, *kq;
  for (int e = 0; e < 3; ++e) {
    csr_count(c, c->wkode + (size_t)nn * e, c->wqkode + (size_t)nq * e, c->nunk[e], c->nbk[e], &vs[e], &vr[e]);
    csr_fill(c, c->wkode + (size_t)nn * e, c->wqkode + (size_t)nq * e, mH, mG, mB, beta, ren / 1.0, &vs[e], &vr[e]);
    vq[e] = malloc(sizeof(double) * c->nunk[e]);
    or_prelsqr2(vs[e].neq, c->nunk[e], vs[e].i, vs[e].j, vs[e].v, vq[e]);
  }
  if (c->iDT == 1) {
    csr_count(c, c->Tkode, c->Tqkode, c->Tnunk, c->Tnb, &ts_, &tr_);
    csr_fill(c, c->Tkode, c->Tqkode, mH, mG, mB, beta, ren * c->PrN / 1.0, &ts_, &tr_);
    tq = malloc(sizeof(double) * c->Tnunk);
    or_prelsqr2(ts_.neq, c->Tnunk, ts_.i, ts_.j, ts_.v, tq);
  }
  csr_km(c, mH, &ks_, &kr_);
  kq = malloc(sizeof(double) * c->nsmcol);
  or_prelsqr2(ks_.neq, c->nsmcol, ks_.i, ks_.j, ks_.v, kq);
  /* ---- state ---- */
  double *pnlsv = malloc(sizeof(double) * nn * 3), *pnlsw = malloc(sizeof(double) * nn * 3);
  double *ptsw = malloc(sizeof(double) * nn * 3), *pptsw = malloc(sizeof(double) * nn * 3);
  double *rhsv = malloc(sizeof(double) * nn * 3);
  double *pnlsT = malloc(sizeof(double) * nn), *ptsT = malloc(sizeof(double) * nn);
  double *pptsT = malloc(sizeof(double) * nn), *Ttime = malloc(sizeof(double) * nn);
  memcpy(pptsw, vort, sizeof(double) * nn * 3);
  memcpy(ptsw, vort, sizeof(double) * nn * 3);
  memcpy(pnlsw, vort, sizeof(double) * nn * 3);
  memcpy(pnlsv, vel, sizeof(double) * nn * 3);
  memcpy(pptsT, temp, sizeof(double) * nn);
  memcpy(ptsT, temp, sizeof(double) * nn);
  memcpy(pnlsT, temp, sizeof(double) * nn);
  /* time step 1 (TIME tstep 1 1) */
  float dlse[9];
  for (int k = 0; k < 9; ++k) dlse[k] = c->lsqs_eps;
  memcpy(pptsw, ptsw, sizeof(double) * nn * 3);
  memcpy(ptsw, vort, sizeof(double) * nn * 3);
  for (size_t k = 0; k < (size_t)nn * 3; ++k) rhsv[k] = -ptsw[k] * beta2 + pptsw[k] * beta3;
  if (c->iDT > 0) {
    memcpy(pptsT, ptsT, sizeof(double) * nn);
    memcpy(ptsT, temp, sizeof(double) * nn);
    for (int k = 0; k < nn; ++k) Ttime[k] = -ptsT[k] * beta2 + pptsT[k] * beta3;
  }
  double *dfi = malloc(sizeof(double) * nb);
  size_t rmax = m1 > (size_t)nn ? m1 : (size_t)nn;
  double *r = malloc(sizeof(double) * rmax), *xs = malloc(sizeof(double) * (rmax + nq));
  double *kb = malloc(sizeof(double) * (nn + nq + 1));
  int it = 0;
  double t0 = now_s();
  for (;;) {
    ++it;
    int nit[12] = {0};
    /* ---- boundary vorticities: sdkm_solve eqn 1..3 (main.f90:1207, iBw = 1) ---- */
    for (int e = 0; e < 3 && kin; ++e) {
      or_sdkm_rhs(e + 1, nn, nb, c->lbn, c->gbn, c->kode, c->nx, c->ny, c->nz, sH, sHtx, sHty, sHtz, sDx, sDy, sDz,
                  nb, vel, vort, dfi, nthreads);
      if (or_dnrm2(nb, dfi) == 0.0) {
        for (int i = 0; i < nb; ++i) dfi[i] = 0.0;
      } else {
        or_dgesl(sA[e], nb, nb, piv[e], dfi);
      }
      for (int i = 0; i < nb; ++i) {
        int kd = c->kode[i + (size_t)nb * e];
        int j = c->gbn[i] - 1;
        if (kd == 0 || kd == 3) vort[j + (size_t)nn * e] = vort[j + (size_t)nn * e] * (1.0 - c->urBw[e]) + c->urBw[e] * dfi[i];
      }
    }
    /* ---- domain velocities: solv_kmMacro3 eqn 1..3 ---- */
    for (int e = 0; e < 3; ++e) {
      for (int i = 0; i < nn; ++i) {
        int u = c->kmu[i];
        if (u < 0) xs[-u - 1] = vel[i + (size_t)nn * e];
        else kb[u - 1] = vel[i + (size_t)nn * e];
      }
      crsxv(&kr_, kb, r);
      int ii = 0;
      for (int ic = 0; ic < nc; ++ic)
        for (int t = 0; t < 27; ++t) {
          size_t i = (size_t)ic * 51 + t;
          for (int j = 0; j < 27; ++j) {
            int node = c->idc[ic + (size_t)nc * j] - 1;
#define VV(d) vel[node + (size_t)nn * (d)]
#define WW(d) vort[node + (size_t)nn * (d)]
            if (e == 0)
              r[ii] = r[ii] - mHtz[ii + m2 * j] * VV(1) + mHty[ii + m2 * j] * VV(2) + mDz[i + m1 * j] * WW(1) -
                      mDy[i + m1 * j] * WW(2);
            else if (e == 1)
              r[ii] = r[ii] + mHtz[ii + m2 * j] * VV(0) - mHtx[ii + m2 * j] * VV(2) - mDz[i + m1 * j] * WW(0) +
                      mDx[i + m1 * j] * WW(2);
            else
              r[ii] = r[ii] - mHty[ii + m2 * j] * VV(0) + mHtx[ii + m2 * j] * VV(1) + mDy[i + m1 * j] * WW(0) -
                      mDx[i + m1 * j] * WW(1);
          }
          ++ii;
        }
      or_slvlsqr2(ks_.neq, c->nsmcol, ks_.i, ks_.j, ks_.v, kq, r, xs, c->lsqs_maxit, dlse[e], &nit[3 + e]);
      for (int i = 0; i < nn; ++i) {
        int u = c->kmu[i];
        if (u < 0) vel[i + (size_t)nn * e] = (1.0 - c->urDv[e]) * vel[i + (size_t)nn * e] + c->urDv[e] * xs[-u - 1];
      }
      for (int i = 0; i < nb; ++i) {
        int kd = c->kode[i + (size_t)nb * e];
        if (kd == 2 || kd == 3) vel[c->gbn[i] - 1 + (size_t)nn * e] = vel[c->kobc[i] - 1 + (size_t)nn * e];
      }
    }
    /* ---- temperature: SolveFDdae ---- */
    if (c->iDT == 1) {
      double enazalfa = ren * c->PrN / 1.0;
      for (int i = 0; i < nn; ++i) {
        if (c->Tkode[i] < 0) xs[-c->Tkode[i] - 1] = temp[i];
        else kb[c->Tkode[i] - 1] = temp[i];
      }
      for (int i = 0; i < nq; ++i) {
        if (c->Tqkode[i] < 0) xs[-c->Tqkode[i] - 1] = qtemp[i];
        else kb[c->Tqkode[i] - 1] = qtemp[i];
      }
      crsxv(&tr_, kb, r);
      for (int j = 0; j < 27; ++j) {
        size_t i = 0;
        for (int ic = 0; ic < nc; ++ic)
          for (int t = 0; t < 51; ++t, ++i) {
            int node = c->idc[ic + (size_t)nc * j] - 1;
            double u = temp[node];
            r[i] = r[i] - mB[i + m1 * j] * Ttime[node] * enazalfa -
                   enazalfa * (+mAx[i + m1 * j] * VV(0) * u + mAy[i + m1 * j] * VV(1) * u + mAz[i + m1 * j] * VV(2) * u -
                               mDx[i + m1 * j] * VV(0) * u - mDy[i + m1 * j] * VV(1) * u - mDz[i + m1 * j] * VV(2) * u);
          }
      }
      or_slvlsqr2(ts_.neq, c->Tnunk, ts_.i, ts_.j, ts_.v, tq, r, xs, c->lsqs_maxit, dlse[6], &nit[9]);
      for (int i = 0; i < nn; ++i)
        if (c->Tkode[i] < 0) temp[i] = (1.0 - c->urDT) * temp[i] + c->urDT * xs[-c->Tkode[i] - 1];
      for (int i = 0; i < nq; ++i)
        if (c->Tqkode[i] < 0) qtemp[i] = xs[-c->Tqkode[i] - 1];
    }
    /* ---- domain vorticities: SolveFDwTE eqn 1..3 ---- */
    for (int e = 0; e < 3; ++e) {
      const int *wk = c->wkode + (size_t)nn * e, *wq = c->wqkode + (size_t)nq * e;
      for (int i = 0; i < nn; ++i) {
        if (wk[i] < 0) xs[-wk[i] - 1] = vort[i + (size_t)nn * e];
        else kb[wk[i] - 1] = vort[i + (size_t)nn * e];
      }
      for (int i = 0; i < nq; ++i) {
        if (wq[i] < 0) xs[-wq[i] - 1] = qvort[i + (size_t)nq * e];
        else kb[wq[i] - 1] = qvort[i + (size_t)nq * e];
      }
      crsxv(&vr[e], kb, r);
      double rn = ren / 1.0;
      for (int j = 0; j < 27; ++j) {
        size_t i = 0;
        for (int ic = 0; ic < nc; ++ic)
          for (int t = 0; t < 51; ++t, ++i) {
            int node = c->idc[ic + (size_t)nc * j] - 1;
            double we = WW(e), ve = VV(e);
            r[i] = r[i] - mB[i + m1 * j] * rhsv[node + (size_t)nn * e] * ren / 1.0 -
                   rn * (+mAx[i + m1 * j] * (VV(0) * we - ve * WW(0)) + mAy[i + m1 * j] * (VV(1) * we - ve * WW(1)) +
                         mAz[i + m1 * j] * (VV(2) * we - ve * WW(2)) - mDx[i + m1 * j] * (VV(0) * we - ve * WW(0)) -
                         mDy[i + m1 * j] * (VV(1) * we - ve * WW(1)) - mDz[i + m1 * j] * (VV(2) * we - ve * WW(2)));
          }
      }
      if (c->iDT > 0) {
        /* BouyancyRHS with dc = 0, Rro = 1e20, kbf = 0, Ram = 0, nanofluid factors 1 */
        double fac = c->RaN / c->PrN / ren * 1.0 / 1.0, facm = 0.0 / c->PrN / ren * 1.0 / 1.0;
        const double Rro = 1.0e20;
        for (int j = 0; j < 27; ++j) {
          size_t i = 0;
          for (int ic = 0; ic < nc; ++ic)
            for (int t = 0; t < 51; ++t, ++i) {
              int node = c->idc[ic + (size_t)nc * j] - 1;
              double tt = temp[node] - 0.0 / Rro;
              double ax = mAx[i + m1 * j] - mDx[i + m1 * j], ay = mAy[i + m1 * j] - mDy[i + m1 * j],
                     az = mAz[i + m1 * j] - mDz[i + m1 * j];
              if (e == 0)
                r[i] = r[i] - fac * (+c->gz * ay * tt - c->gy * az * tt) + facm * (+ay * 0.0 - az * 0.0);
              else if (e == 1)
                r[i] = r[i] - fac * (-c->gz * ax * tt + c->gx * az * tt) + facm * (+ax * 0.0 - az * 0.0);
              else
                r[i] = r[i] - fac * (+c->gy * ax * tt - c->gx * ay * tt) + facm * (+ax * 0.0 - ay * 0.0);
            }
        }
      }
      or_slvlsqr2(vs[e].neq, c->nunk[e], vs[e].i, vs[e].j, vs[e].v, vq[e], r, xs, c->lsqs_maxit, dlse[3 + e],
                  &nit[6 + e]);
      for (int i = 0; i < nn; ++i)
        if (wk[i] < 0)
          vort[i + (size_t)nn * e] = (1.0 - c->urDw[e]) * vort[i + (size_t)nn * e] + c->urDw[e] * xs[-wk[i] - 1];
      for (int i = 0; i < nq; ++i)
        if (wq[i] < 0) qvort[i + (size_t)nq * e] = xs[-wq[i] - 1];
    }
#undef VV
#undef WW
    /* ---- CheckConvergencevw ---- */
    double rnd[9] = {0}, rno[9] = {0}, err[9];
    for (int i = 0; i < nn; ++i) {
      double dif[9];
      for (int d = 0; d < 3; ++d) {
        dif[d] = vel[i + (size_t)nn * d] - pnlsv[i + (size_t)nn * d];
        dif[3 + d] = vort[i + (size_t)nn * d] - pnlsw[i + (size_t)nn * d];
      }
      dif[6] = temp[i] - pnlsT[i];
      dif[7] = dif[8] = 0.0;
      for (int j = 0; j < 9; ++j) rnd[j] = rnd[j] + dif[j] * dif[j];
      for (int d = 0; d < 3; ++d) {
        rno[d] = rno[d] + vel[i + (size_t)nn * d] * vel[i + (size_t)nn * d];
        rno[3 + d] = rno[3 + d] + vort[i + (size_t)nn * d] * vort[i + (size_t)nn * d];
      }
      rno[6] = rno[6] + temp[i] * temp[i];
      for (int d = 0; d < 3; ++d) {
        pnlsv[i + (size_t)nn * d] = vel[i + (size_t)nn * d];
        pnlsw[i + (size_t)nn * d] = vort[i + (size_t)nn * d];
      }
      pnlsT[i] = temp[i];
    }
    for (int j = 0; j < 9; ++j) err[j] = 9.999999999;
    double mx = fmax(fmax(rno[0], rno[1]), rno[2]);
    if (mx > 1.0e-15)
      for (int j = 0; j < 3; ++j) err[j] = sqrt(rnd[j] / mx);
    mx = fmax(fmax(rno[3], rno[4]), rno[5]);
    if (mx > 1.0e-15)
      for (int j = 3; j < 6; ++j) err[j] = sqrt(rnd[j] / mx);
    for (int j = 6; j < 9; ++j)
      if (rno[j] > 1.0e-15) err[j] = sqrt(rnd[j] / rno[j]);
    for (int j = 0; j < 9; ++j) {
      if (rnd[j] == 0 && rno[j] == 0.0) err[j] = 0.0;
      if (err[j] > 9.999999999) err[j] = 9.999999999;
    }
    if (sta && it <= maxhist) memcpy(sta + (size_t)(it - 1) * 9, err, sizeof(err));
    if (ite && it <= maxhist) memcpy(ite + (size_t)(it - 1) * 12, nit, sizeof(nit));
    /* ---- DLSE ---- */
    if (c->iDLSE == 1 && it > 1) {
      if (it == 2)
        for (int k = 0; k < 9; ++k) dlse[k] = c->dlse_epsth;
      for (int k = 0; k < 8; ++k) {
        float old = dlse[k];
        if (err[k] / (double)c->dlse_r < (double)c->dlse_epsth && err[k] != 0.0)
          dlse[k] = (float)(err[k] / (double)c->dlse_r);
        else
          dlse[k] = c->dlse_epsth;
        if (old < dlse[k]) dlse[k] = old;
        if (dlse[k] < c->lsqs_eps) dlse[k] = c->lsqs_eps;
      }
    }
    int more = 0;
    for (int j = 0; j < 9; ++j)
      if (err[j] > c->eps) more = 1;
    if (!(it < c->nnlit && more)) break;
  }
  if (seconds_per_iter) *seconds_per_iter = (now_s() - t0) / it;
  /* ---- cleanup ---- */
  free(mH); free(mG); free(mB); free(mAx); free(mAy); free(mAz); free(mHtx); free(mHty); free(mHtz);
  free(mDx); free(mDy); free(mDz); free(sH); free(sHtx); free(sHty); free(sHtz); free(sDx); free(sDy); free(sDz);
  for (int e = 0; e < 3; ++e) { free(sA[e]); free(piv[e]); csr_free(&vs[e]); csr_free(&vr[e]); free(vq[e]); }
  if (c->iDT == 1) { csr_free(&ts_); csr_free(&tr_); free(tq); }
  csr_free(&ks_); csr_free(&kr_); free(kq);
  free(pnlsv); free(pnlsw); free(ptsw); free(pptsw); free(rhsv); free(pnlsT); free(ptsT); free(pptsT); free(Ttime);
  free(dfi); free(r); free(xs); free(kb);
  return it;
}
