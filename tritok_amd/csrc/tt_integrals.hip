// tt_integrals.hip -- integral assembly on gfx950.
//
// Macro matrices (FMATkmdcSteady, lib_fmat.f:3-446): one thread per
// (cell, source) row; lanes of a wavefront are consecutive cells with the same
// source index, so every lane follows the same subdivision and singular path
// on a structured mesh.  The thread owns its matrix rows and applies the
// free-term, Ht row-sum and linear-field corrections itself.
//
// Single-domain kinematics (FMATsdkm, lib_sdkm.f:1286-1530): one thread per
// boundary row; the wavefront walks the boundary elements / cells in order,
// so element and cell geometry loads are uniform and every column update is a
// coalesced 64-row segment of the column-major matrices.  Cells that contain
// the row's own node (singular, spherical transform) are taken out of that
// walk and integrated by a second kernel, one thread per (row, cell) pair,
// then added in cell order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "tritok_bem.h"
#include "tt_integrals.h"

namespace tt {

// idc2ibc local node permutations (lib_interpol.f:683-688), 1-based
__constant__ static const int kFace9[6][9] = {{6, 14, 2, 10, 3, 15, 7, 18, 23}, {6, 18, 7, 19, 8, 20, 5, 17, 26},
                                              {7, 15, 3, 11, 4, 16, 8, 19, 24}, {3, 10, 2, 9, 1, 12, 4, 11, 21},
                                              {2, 14, 6, 17, 5, 13, 1, 9, 22},  {5, 20, 8, 16, 4, 12, 1, 13, 25}};
__constant__ static const int kGeo8[8] = {6, 7, 2, 3, 5, 8, 1, 4};

// ---------------------------------------------------------------------------
// host helpers
int gauss_legendre(int rule, double *gi, double *ome) {
  static const int kn[9] = {2, 4, 6, 8, 16, 32, 40, 48, 64};
  if (rule < 1 || rule > 9) return 0;
  int n = kn[rule - 1];
  for (int i = 0; i < (n + 1) / 2; ++i) {
    long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L)), pp = 0;
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

static int rule_from(const tt_quad &q, int rule, double *gi, double *ome) {
  if (rule < 1 || rule > 9) throw std::runtime_error("Gauss rule index outside 1..9");
  if (q.gi && q.ome && q.ng1 && q.ng2) {
    int a = q.ng1[rule - 1], b = q.ng2[rule - 1];
    if (a < 1 || b > 220 || b - a + 1 > kMaxRule || b < a) throw std::runtime_error("bad Gauss table ranges");
    for (int k = a; k <= b; ++k) gi[k - a] = q.gi[k - 1], ome[k - a] = q.ome[k - 1];
    return b - a + 1;
  }
  return gauss_legendre(rule, gi, ome);
}

void build_quad(const tt_quad &q, QuadDev &o) {
  std::memset(&o, 0, sizeof(o));
  o.nbr = rule_from(q, q.kmBr, o.gbr, o.obr);
  o.nbs = rule_from(q, q.kmBs, o.gbs, o.obs);
  o.ndr = rule_from(q, q.kmDr, o.gdr, o.odr);
  o.nds = rule_from(q, q.kmDs, o.gds, o.ods);
  if (q.iiDiv < 1 || q.iiDiv > 8) throw std::runtime_error("IIDI must be in 1..8");
  o.iiDiv = q.iiDiv;
}

void upload_mesh(const tt_mesh_view &v, DevMesh &d, hipStream_t s) {
  d.nn = v.nnodes, d.nb = v.nbnodes, d.ne = v.nbelem, d.nc = v.nicell, d.nq = v.nq;
  if (d.nn <= 0 || d.nc <= 0 || d.ne <= 0 || d.nb <= 0 || d.nq <= 0) throw std::runtime_error("empty mesh");
  d.x.upload(v.x, (size_t)d.nn * 3, s);
  d.xq.upload(v.xq, (size_t)d.nq * 3, s);
  d.idc.upload(v.idc, (size_t)d.nc * 27, s);
  d.ibc.upload(v.ibc, (size_t)d.ne * 9, s);
  d.ibf.upload(v.ibf, (size_t)d.nc * 24, s);
  d.flpm.upload(v.flpm, (size_t)d.nc * 6, s);
  d.gbn.upload(v.gbn, (size_t)d.nb, s);
  d.lbn.upload(v.lbn, (size_t)d.nn, s);
  if (v.nx) {
    d.nx.upload(v.nx, d.nb, s);
    d.ny.upload(v.ny, d.nb, s);
    d.nz.upload(v.nz, d.nb, s);
  }
}

MeshPtrs ptrs(const DevMesh &d) {
  return MeshPtrs{d.nn, d.nb, d.ne, d.nc, d.nq, d.x.p, d.xq.p, d.idc.p, d.ibc.p, d.ibf.p, d.flpm.p, d.gbn.p, d.lbn.p};
}

// ---------------------------------------------------------------------------
// macro matrices

// device time and kernel evaluations of the calling thread's last macro
// assembly (measurement hook)
struct MacroStats {
  float ms;
  long long dom_pts, bnd_pts;
};
static thread_local MacroStats g_last_macro{0.f, 0, 0};

struct MacroOut {
  double *H, *G, *B, *Ax, *Ay, *Az, *Htx, *Hty, *Htz, *Dx, *Dy, *Dz;
  double *ont;  // per row: 11 accuracy terms
};

// Helm: FMATmodHelm (lib_mHelm.f:2-483) with mu of each node / flux node in
// mun / muq; otherwise FMATkmdcSteady.  The free term, Ht and linear-field
// corrections follow all integrals of the row (the modified-Helmholtz free
// term needs the row's B).
template <bool Helm>
__global__ void __launch_bounds__(64) macro_kernel(MeshPtrs m, const QuadDev *__restrict__ qp, int ncomp,
                                                   const double *__restrict__ mun, const double *__restrict__ muq,
                                                   MacroOut o, unsigned long long *__restrict__ npts) {
  const int ic = blockIdx.x * 64 + threadIdx.x;
  const int it = blockIdx.y;  // source 0..50
  if (ic >= ncomp) return;
  const QuadDev &q = *qp;
  const size_t nc = m.nc, m1 = nc * 51, m2 = nc * 27;
  const size_t i = (size_t)ic * 51 + it, ii = (size_t)ic * 27 + it;
  const bool usrc = it < 27;
  auto X = [&](int node, int d) { return m.x[(node - 1) + (size_t)m.nn * d]; };
  auto IDC = [&](int k) { return m.idc[ic + nc * (k - 1)]; };
  double p[3], mu = 0.0;
  int its = 0, itf = 0;
  if (usrc) {
    int dn = IDC(it + 1);
    p[0] = X(dn, 0), p[1] = X(dn, 1), p[2] = X(dn, 2);
    if (Helm) mu = mun[dn - 1];
  } else {
    its = (it - 27) / 4, itf = (it - 27) % 4 + 1;
    int dn = m.ibf[ic + nc * (its + 6 * (itf - 1))];
    p[0] = m.xq[dn - 1], p[1] = m.xq[dn - 1 + m.nq], p[2] = m.xq[dn - 1 + 2 * (size_t)m.nq];
    if (Helm) mu = muq[dn - 1];
  }
  // shortest edge of the cell's faces (FindMinEdge)
  double minedge = 1.0e20;
  for (int s = 0; s < 6; ++s)
    for (int e = 0; e < 4; ++e) {
      int a = IDC(kFace9[s][2 * e]), b = IDC(kFace9[s][(2 * e + 2) % 8]);
      double dx = X(a, 0) - X(b, 0), dy = X(a, 1) - X(b, 1), dz = X(a, 2) - X(b, 2);
      minedge = fmin(minedge, sqrt(dx * dx + dy * dy + dz * dz));
    }
  double *ont = o.ont + i * 12;
  for (int k = 0; k < 11; ++k) ont[k] = 0.0;
  ont[11] = mu;
  // ---- boundary integrals over the six faces
  int pts_b = 0;
  for (int je = 0; je < 6; ++je) {
    double c[4][3];
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) c[k][d] = X(IDC(kFace9[je][2 * k]), d);
    int isrc = 0;
    if (usrc) {
      for (int s = 0; s < 9; ++s)
        if (kFace9[je][s] == it + 1) isrc = s + 1;
    } else if (its == je) {
      isrc = itf + 9;
    }
    BndAcc<true, Helm> acc;
    acc.zero();
    acc.mu = mu;
    integrate_face<true>(q, p, c, isrc, minedge, acc);
    pts_b += acc.npts;
    for (int s = 0; s < 9; ++s) {
      size_t col = kFace9[je][s] - 1;
      o.H[i + m1 * col] += acc.he[s];
      o.Ax[i + m1 * col] += acc.ad[0][s];
      o.Ay[i + m1 * col] += acc.ad[1][s];
      o.Az[i + m1 * col] += acc.ad[2][s];
      if (usrc) {
        o.Htx[ii + m2 * col] -= acc.heyz[s];
        o.Hty[ii + m2 * col] -= acc.hezx[s];
        o.Htz[ii + m2 * col] -= acc.hexy[s];
      }
    }
    double sg = (double)m.flpm[ic + nc * je];
    for (int s = 0; s < 4; ++s) o.G[i + m1 * (je * 4 + s)] += acc.ge[s] * sg;
  }
  // ---- domain integrals
  double cc[8][3];
  for (int k = 0; k < 8; ++k)
    for (int d = 0; d < 3; ++d) cc[k][d] = X(IDC(kGeo8[k]), d);
  DomAcc<4, Helm> dacc;
  dacc.zero();
  dacc.mu = mu;
  integrate_cell(q, p, cc, it + 1, dacc);
  if (npts) {
    atomicAdd(&npts[0], (unsigned long long)dacc.npts);
    atomicAdd(&npts[1], (unsigned long long)pts_b);
  }
  for (int s = 0; s < 27; ++s) {
    o.Dx[i + m1 * s] = dacc.a[0][s];
    o.Dy[i + m1 * s] = dacc.a[1][s];
    o.Dz[i + m1 * s] = dacc.a[2][s];
    o.B[i + m1 * s] = dacc.a[3][s];
  }
  // ---- free term (rigid body u = 1; modified Helmholtz: f = -mu^2) and Ht row sums
  double cfree = 0.0;
  if (Helm) {
    double vH = 0.0, vB = 0.0;
    for (int j = 0; j < 27; ++j) vH += o.H[i + m1 * j];
    for (int j = 0; j < 27; ++j) vB += dacc.a[3][j];
    cfree = -vH + mu * mu * vB;
  } else {
    for (int j = 0; j < 27; ++j) cfree -= o.H[i + m1 * j];
  }
  if (usrc) {
    o.H[i + m1 * it] += cfree;
    double *Ht[3] = {o.Htx, o.Hty, o.Htz};
    for (int t = 0; t < 3; ++t) {
      double vs = 0.0;
      for (int j = 0; j < 27; ++j) vs += Ht[t][ii + m2 * j];
      ont[2 + t] = fabs(vs);
      Ht[t][ii + m2 * it] -= vs;
    }
    if (it == 26) ont[0] = fabs(1.0 - cfree);
  } else {
    ont[1] = fabs(0.5 - cfree);
    const double tc = 3.0 / 4.0;
    double fi[9];
    shape9((itf == 1 || itf == 4) ? -tc : tc, itf <= 2 ? -tc : tc, fi);
    for (int s = 0; s < 9; ++s) o.H[i + m1 * (kFace9[its][s] - 1)] += cfree * fi[s];
  }
  if (usrc) {
    // linear fields v = (z,0,0), (0,x,0), (0,0,y): sum_j (H - mu^2 B)_ij x_j = sum_j D_ij
    double *Dm[3] = {o.Dz, o.Dx, o.Dy};
    const int crd[3] = {2, 0, 1}, slot[3] = {7, 5, 6};
    for (int t = 0; t < 3; ++t) {
      double hv = 0.0, vs = 0.0;
      for (int j = 0; j < 27; ++j) {
        double h = Helm ? o.H[i + m1 * j] - mu * mu * o.B[i + m1 * j] : o.H[i + m1 * j];
        hv += h * X(IDC(j + 1), crd[t]);
        vs += Dm[t][i + m1 * j];
      }
      ont[slot[t]] = fabs(hv - vs);
      Dm[t][i + m1 * it] = Dm[t][i + m1 * it] + hv - vs;
    }
    double hxy = 0, hyz = 0, hzx = 0;
    for (int j = 0; j < 27; ++j) {
      int nj = IDC(j + 1);
      hxy += o.Htx[ii + m2 * j] * X(nj, 1) + o.Dz[i + m1 * j];
      hyz += o.Hty[ii + m2 * j] * X(nj, 2) + o.Dx[i + m1 * j];
      hzx += o.Htz[ii + m2 * j] * X(nj, 0) + o.Dy[i + m1 * j];
    }
    ont[8] = fabs(hyz), ont[9] = fabs(hzx), ont[10] = fabs(hxy);
  }
}

// COPY=1 (lib_fmat.f:89-120): rows of cell 1 replicated, G with the flux sign
__global__ void macro_copy_kernel(int nc, const int *__restrict__ flpm, MacroOut o) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t m1 = (size_t)nc * 51, m2 = (size_t)nc * 27;
  if (r >= m1) return;
  size_t ic = r / 51, it = r % 51;
  if (ic == 0) return;
  for (int j = 0; j < 27; ++j) {
    o.H[r + m1 * j] = o.H[it + m1 * j];
    o.B[r + m1 * j] = o.B[it + m1 * j];
    o.Ax[r + m1 * j] = o.Ax[it + m1 * j];
    o.Ay[r + m1 * j] = o.Ay[it + m1 * j];
    o.Az[r + m1 * j] = o.Az[it + m1 * j];
    o.Dx[r + m1 * j] = o.Dx[it + m1 * j];
    o.Dy[r + m1 * j] = o.Dy[it + m1 * j];
    o.Dz[r + m1 * j] = o.Dz[it + m1 * j];
  }
  for (int je = 0; je < 6; ++je)
    for (int s = 0; s < 4; ++s)
      o.G[r + m1 * (je * 4 + s)] = o.G[it + m1 * (je * 4 + s)] * (double)flpm[ic + (size_t)nc * je] * (double)flpm[nc * je];
  if (it < 27) {
    size_t rr = ic * 27 + it;
    for (int j = 0; j < 27; ++j) {
      o.Htx[rr + m2 * j] = o.Htx[it + m2 * j];
      o.Hty[rr + m2 * j] = o.Hty[it + m2 * j];
      o.Htz[rr + m2 * j] = o.Htz[it + m2 * j];
    }
  }
}

void macro_integrals_dev(const DevMesh &m, const QuadDev &qh, int copy, MacroDev &o, double ont[11], hipStream_t s,
                         const double *mun, const double *muq, double *mumax) {
  const size_t nc = m.nc, m1 = nc * 51, m2 = nc * 27;
  o.nc = m.nc;
  DBuf<double> *mats1[] = {&o.H, &o.B, &o.Ax, &o.Ay, &o.Az, &o.Dx, &o.Dy, &o.Dz};
  for (auto *b : mats1) b->alloc(m1 * 27);
  o.G.alloc(m1 * 24);
  o.Htx.alloc(m2 * 27), o.Hty.alloc(m2 * 27), o.Htz.alloc(m2 * 27);
  for (auto *b : mats1) b->zero(s);
  o.G.zero(s), o.Htx.zero(s), o.Hty.zero(s), o.Htz.zero(s);
  DBuf<QuadDev> dq(1);
  dq.upload(&qh, 1, s);
  int ncomp = copy ? 1 : m.nc;
  DBuf<double> ontd((size_t)ncomp * 51 * 12);
  MacroOut out{o.H.p, o.G.p, o.B.p, o.Ax.p, o.Ay.p, o.Az.p, o.Htx.p, o.Hty.p, o.Htz.p, o.Dx.p, o.Dy.p, o.Dz.p, ontd.p};
  DBuf<unsigned long long> cnt(2);
  cnt.zero(s);
  hipEvent_t e0, e1;
  TT_HIP(hipEventCreate(&e0));
  TT_HIP(hipEventCreate(&e1));
  TT_HIP(hipEventRecord(e0, s));
  if (mun && muq)
    hipLaunchKernelGGL(macro_kernel<true>, dim3(ceil_div(ncomp, 64), 51), dim3(64), 0, s, ptrs(m), dq.p, ncomp, mun,
                       muq, out, cnt.p);
  else
    hipLaunchKernelGGL(macro_kernel<false>, dim3(ceil_div(ncomp, 64), 51), dim3(64), 0, s, ptrs(m), dq.p, ncomp,
                       nullptr, nullptr, out, cnt.p);
  TT_LAUNCH_CHECK();
  TT_HIP(hipEventRecord(e1, s));
  if (copy && m.nc > 1) {
    hipLaunchKernelGGL(macro_copy_kernel, dim3(ceil_div(m1, 256)), dim3(256), 0, s, m.nc, m.flpm.p, out);
    TT_LAUNCH_CHECK();
  }
  std::vector<double> oh(ontd.n);
  ontd.download(oh.data(), oh.size(), s);
  unsigned long long hc[2] = {0, 0};
  cnt.download(hc, 2, s);
  TT_HIP(hipStreamSynchronize(s));
  float ms = 0.f;
  TT_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0), (void)hipEventDestroy(e1);
  g_last_macro = MacroStats{ms, (long long)hc[0], (long long)hc[1]};
  // accumulate as FMATkmdcSteady / FMATmodHelm: sums for ontH/ontG, maxima
  // otherwise (slot 11: the largest mu)
  double r[12] = {0};
  for (int ic = 0; ic < ncomp; ++ic)
    for (int it = 0; it < 51; ++it) {
      const double *v = &oh[((size_t)ic * 51 + it) * 12];
      r[0] += v[0];
      r[1] += v[1];
      for (int t = 2; t < 12; ++t) r[t] = std::max(r[t], v[t]);
    }
  if (copy) {
    r[1] /= 24;
  } else {
    r[1] = r[1] / m.nc / 24;
    r[0] /= m.nc;
  }
  std::memcpy(ont, r, sizeof(double) * 11);
  if (mumax) *mumax = r[11];
}

// ---------------------------------------------------------------------------
// single-domain kinematics
struct SdOut {
  int ld;
  double *H, *Htx, *Hty, *Htz, *Dx, *Dy, *Dz;
};

__global__ void __launch_bounds__(64) sd_bnd_kernel(MeshPtrs m, const QuadDev *__restrict__ qp, int row0, int nrows,
                                                    SdOut o) {
  const int r = blockIdx.x * 64 + threadIdx.x;
  const bool active = r < nrows;
  const int i = active ? row0 + r : row0;
  const int dn = m.gbn[i];
  const double p[3] = {m.x[dn - 1], m.x[dn - 1 + m.nn], m.x[dn - 1 + 2 * (size_t)m.nn]};
  const size_t ne = m.ne, ld = o.ld;
  for (int je = 0; je < m.ne; ++je) {
    double c[4][3];
    int nodes[9];
    for (int k = 0; k < 9; ++k) nodes[k] = m.ibc[je + ne * k];
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) c[k][d] = m.x[nodes[2 * k] - 1 + (size_t)m.nn * d];
    int isrc = 0;
    for (int s = 0; s < 9; ++s)
      if (nodes[s] == dn) isrc = s + 1;
    BndAcc<false> acc;
    acc.zero();
    integrate_face<false>(*qp, p, c, isrc, 0.0, acc);
    if (!active) continue;
    for (int s = 0; s < 9; ++s) {
      size_t j = m.lbn[nodes[s] - 1] - 1;
      o.H[r + ld * j] += acc.he[s];
      o.Htx[r + ld * j] -= acc.heyz[s];
      o.Hty[r + ld * j] -= acc.hezx[s];
      o.Htz[r + ld * j] -= acc.hexy[s];
    }
  }
}

__global__ void __launch_bounds__(64) sd_dom_kernel(MeshPtrs m, const QuadDev *__restrict__ qp, int row0, int nrows,
                                                    SdOut o) {
  const int r = blockIdx.x * 64 + threadIdx.x;
  const bool active = r < nrows;
  const int i = active ? row0 + r : row0;
  const int dn = m.gbn[i];
  const double p[3] = {m.x[dn - 1], m.x[dn - 1 + m.nn], m.x[dn - 1 + 2 * (size_t)m.nn]};
  const size_t nc = m.nc, ld = o.ld;
  for (int ic = 0; ic < m.nc; ++ic) {
    int nodes[27];
    bool sing = false;
    for (int k = 0; k < 27; ++k) {
      nodes[k] = m.idc[ic + nc * k];
      sing |= nodes[k] == dn;
    }
    double c[8][3];
    for (int k = 0; k < 8; ++k)
      for (int d = 0; d < 3; ++d) c[k][d] = m.x[nodes[kGeo8[k] - 1] - 1 + (size_t)m.nn * d];
    if (!active || sing) continue;  // singular cells: sd_dom_singular_kernel
    DomAcc<3> acc;
    acc.zero();
    integrate_cell(*qp, p, c, 0, acc);
    for (int s = 0; s < 27; ++s) {
      size_t j = nodes[s] - 1;
      o.Dx[r + ld * j] += acc.a[0][s];
      o.Dy[r + ld * j] += acc.a[1][s];
      o.Dz[r + ld * j] += acc.a[2][s];
    }
  }
}

// one thread per singular (row, cell) pair -> staging[pair][3][27]
__global__ void __launch_bounds__(64) sd_dom_singular_kernel(MeshPtrs m, const QuadDev *__restrict__ qp,
                                                             const int *__restrict__ prow, const int *__restrict__ pcell,
                                                             const int *__restrict__ psrc, int npairs,
                                                             double *__restrict__ staging) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  if (k >= npairs) return;
  const size_t nc = m.nc;
  int i = prow[k], ic = pcell[k];
  int dn = m.gbn[i];
  const double p[3] = {m.x[dn - 1], m.x[dn - 1 + m.nn], m.x[dn - 1 + 2 * (size_t)m.nn]};
  double c[8][3];
  for (int t = 0; t < 8; ++t)
    for (int d = 0; d < 3; ++d) c[t][d] = m.x[m.idc[ic + nc * (kGeo8[t] - 1)] - 1 + (size_t)m.nn * d];
  DomAcc<3> acc;
  acc.zero();
  integrate_cell(*qp, p, c, psrc[k], acc);
  for (int t = 0; t < 3; ++t)
    for (int s = 0; s < 27; ++s) staging[(size_t)k * 81 + t * 27 + s] = acc.a[t][s];
}

// adds each row's singular pairs (cell order) into D
__global__ void sd_dom_singular_apply(MeshPtrs m, int row0, int nrows, const int *__restrict__ off,
                                      const int *__restrict__ pcell, const double *__restrict__ staging, SdOut o) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const size_t nc = m.nc, ld = o.ld;
  for (int k = off[r]; k < off[r + 1]; ++k) {
    int ic = pcell[k];
    for (int s = 0; s < 27; ++s) {
      size_t j = m.idc[ic + nc * s] - 1;
      o.Dx[r + ld * j] += staging[(size_t)k * 81 + s];
      o.Dy[r + ld * j] += staging[(size_t)k * 81 + 27 + s];
      o.Dz[r + ld * j] += staging[(size_t)k * 81 + 54 + s];
    }
  }
}

// free term, Ht row sums and linear-field corrections (lib_sdkm.f:1387-1522)
__global__ void sd_corr_kernel(MeshPtrs m, int row0, int nrows, SdOut o, double *__restrict__ ont) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const size_t ld = o.ld, nb = m.nb, nn = m.nn;
  const int i = row0 + r, dn = m.gbn[i];
  auto X = [&](int node, int d) { return m.x[(node - 1) + nn * d]; };
  double *on = ont + (size_t)r * 10;
  double cfree = 0.0;
  for (size_t j = 0; j < nb; ++j) cfree -= o.H[r + ld * j];
  on[0] = cfree;
  o.H[r + ld * i] += cfree;
  double *Ht[3] = {o.Htx, o.Hty, o.Htz};
  for (int t = 0; t < 3; ++t) {
    double vs = 0.0;
    for (size_t j = 0; j < nb; ++j) vs += Ht[t][r + ld * j];
    on[1 + t] = fabs(vs);
    Ht[t][r + ld * i] -= vs;
  }
  double *Dm[3] = {o.Dz, o.Dx, o.Dy};
  const int crd[3] = {2, 0, 1}, slot[3] = {6, 4, 5};
  for (int t = 0; t < 3; ++t) {
    double hv = 0.0, vs = 0.0;
    for (size_t s = 0; s < nb; ++s) hv += o.H[r + ld * s] * X(m.gbn[s], crd[t]);
    for (size_t j = 0; j < nn; ++j) vs += Dm[t][r + ld * j];
    on[slot[t]] = fabs(hv - vs);
    Dm[t][r + ld * (dn - 1)] = Dm[t][r + ld * (dn - 1)] + hv - vs;
  }
  double v1 = 0, v2 = 0, v3 = 0, hxy = 0, hyz = 0, hzx = 0;
  for (size_t j = 0; j < nn; ++j) v1 += o.Dz[r + ld * j], v2 += o.Dx[r + ld * j], v3 += o.Dy[r + ld * j];
  for (size_t j = 0; j < nb; ++j) {
    int g = m.gbn[j];
    hxy += o.Htx[r + ld * j] * X(g, 1);
    hyz += o.Hty[r + ld * j] * X(g, 2);
    hzx += o.Htz[r + ld * j] * X(g, 0);
  }
  on[7] = fabs(hxy + v1), on[8] = fabs(hyz + v2), on[9] = fabs(hzx + v3);
}

void sd_integrals_dev(const DevMesh &m, const tt_mesh_view &hv, const QuadDev &qh, int row0, int row1, SdDev &o,
                      double ont[10], hipStream_t s) {
  if (row0 < 0 || row1 > m.nb || row1 <= row0) throw std::runtime_error("row range outside 0..nbnodes");
  const int nr = row1 - row0;
  o.row0 = row0, o.row1 = row1;
  DBuf<double> *hm[] = {&o.H, &o.Htx, &o.Hty, &o.Htz}, *dm[] = {&o.Dx, &o.Dy, &o.Dz};
  for (auto *b : hm) b->alloc((size_t)nr * m.nb), b->zero(s);
  for (auto *b : dm) b->alloc((size_t)nr * m.nn), b->zero(s);
  DBuf<QuadDev> dq(1);
  dq.upload(&qh, 1, s);
  SdOut out{nr, o.H.p, o.Htx.p, o.Hty.p, o.Htz.p, o.Dx.p, o.Dy.p, o.Dz.p};
  hipLaunchKernelGGL(sd_bnd_kernel, dim3(ceil_div(nr, 64)), dim3(64), 0, s, ptrs(m), dq.p, row0, nr, out);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(sd_dom_kernel, dim3(ceil_div(nr, 64)), dim3(64), 0, s, ptrs(m), dq.p, row0, nr, out);
  TT_LAUNCH_CHECK();
  // singular (row, cell) pairs: cells containing the row's node, in cell order
  std::vector<std::vector<int>> node_cells(m.nn);
  for (int ic = 0; ic < m.nc; ++ic)
    for (int k = 0; k < 27; ++k) {
      int nd = hv.idc[ic + (size_t)m.nc * k] - 1;
      if (node_cells[nd].empty() || node_cells[nd].back() != ic) node_cells[nd].push_back(ic);
    }
  std::vector<int> prow, pcell, psrc, off(nr + 1, 0);
  for (int r = 0; r < nr; ++r) {
    int dn = hv.gbn[row0 + r];
    for (int ic : node_cells[dn - 1]) {
      int isrc = 0;
      for (int k = 0; k < 27; ++k)
        if (hv.idc[ic + (size_t)m.nc * k] == dn) isrc = k + 1;
      prow.push_back(row0 + r), pcell.push_back(ic), psrc.push_back(isrc);
    }
    off[r + 1] = (int)prow.size();
  }
  int np = (int)prow.size();
  DBuf<int> dprow, dpcell, dpsrc, doff;
  dprow.upload(prow, s), dpcell.upload(pcell, s), dpsrc.upload(psrc, s), doff.upload(off, s);
  DBuf<double> staging((size_t)np * 81);
  if (np) {
    hipLaunchKernelGGL(sd_dom_singular_kernel, dim3(ceil_div(np, 64)), dim3(64), 0, s, ptrs(m), dq.p, dprow.p,
                       dpcell.p, dpsrc.p, np, staging.p);
    TT_LAUNCH_CHECK();
    hipLaunchKernelGGL(sd_dom_singular_apply, dim3(ceil_div(nr, 64)), dim3(64), 0, s, ptrs(m), row0, nr, doff.p,
                       dpcell.p, staging.p, out);
    TT_LAUNCH_CHECK();
  }
  DBuf<double> ontd((size_t)nr * 10);
  hipLaunchKernelGGL(sd_corr_kernel, dim3(ceil_div(nr, 64)), dim3(64), 0, s, ptrs(m), row0, nr, out, ontd.p);
  TT_LAUNCH_CHECK();
  std::vector<double> oh(ontd.n);
  ontd.download(oh.data(), oh.size(), s);
  TT_HIP(hipStreamSynchronize(s));
  double r[10] = {0};
  for (int k = 0; k < nr; ++k)
    for (int t = 0; t < 10; ++t) r[t] = std::max(r[t], oh[(size_t)k * 10 + t]);
  r[0] = std::fabs(r[0] - 0.5);
  std::memcpy(ont, r, sizeof(r));
}

}  // namespace tt

// ---------------------------------------------------------------------------
// C-ABI
using namespace tt;

extern "C" int tt_macro_integrals(const tt_mesh_view *mv, const tt_quad *q, int32_t copy, double *H, double *G,
                                  double *B, double *Abdx, double *Abdy, double *Abdz, double *Htx, double *Hty,
                                  double *Htz, double *Dx, double *Dy, double *Dz, double ont[11]) {
  return guarded("tt_macro_integrals", [&] {
    if (!mv || !q) throw std::runtime_error("NULL argument");
    QuadDev qh;
    build_quad(*q, qh);
    DevMesh dm;
    upload_mesh(*mv, dm, 0);
    MacroDev md;
    macro_integrals_dev(dm, qh, copy, md, ont, 0);
    size_t m1 = (size_t)mv->nicell * 51 * 27, m2 = (size_t)mv->nicell * 27 * 27;
    md.H.download(H, m1), md.B.download(B, m1), md.Ax.download(Abdx, m1), md.Ay.download(Abdy, m1);
    md.Az.download(Abdz, m1), md.Dx.download(Dx, m1), md.Dy.download(Dy, m1), md.Dz.download(Dz, m1);
    md.G.download(G, (size_t)mv->nicell * 51 * 24);
    md.Htx.download(Htx, m2), md.Hty.download(Hty, m2), md.Htz.download(Htz, m2);
    TT_HIP(hipDeviceSynchronize());
    return 0;
  });
}

extern "C" int tt_macro_integrals_mh(const tt_mesh_view *mv, const tt_quad *q, int32_t copy, const double *diff,
                                     const double *qdiff, double beta, double *H, double *G, double *B, double *Abdx,
                                     double *Abdy, double *Abdz, double *Htx, double *Hty, double *Htz, double *Dx,
                                     double *Dy, double *Dz, double ont[11], double *mumax) {
  return guarded("tt_macro_integrals_mh", [&] {
    if (!mv || !q || !diff || !qdiff) throw std::runtime_error("NULL argument");
    if (!(beta >= 0.0)) throw std::runtime_error("beta must be >= 0");
    // mu = SQRT(inp%beta/mHdiff(dn)) of each source (lib_mHelm.f:139,197,279,303)
    std::vector<double> mun(mv->nnodes), muq(mv->nq);
    for (int k = 0; k < mv->nnodes; ++k) mun[k] = std::sqrt(beta / diff[k]);
    for (int k = 0; k < mv->nq; ++k) muq[k] = std::sqrt(beta / qdiff[k]);
    for (double v : mun)
      if (!std::isfinite(v)) throw std::runtime_error("diff must be positive (mu = sqrt(beta/diff) not finite)");
    for (double v : muq)
      if (!std::isfinite(v)) throw std::runtime_error("qdiff must be positive (mu = sqrt(beta/qdiff) not finite)");
    QuadDev qh;
    build_quad(*q, qh);
    DevMesh dm;
    upload_mesh(*mv, dm, 0);
    DBuf<double> dmun, dmuq;
    dmun.upload(mun.data(), mun.size(), 0);
    dmuq.upload(muq.data(), muq.size(), 0);
    MacroDev md;
    double mm = 0.0;
    macro_integrals_dev(dm, qh, copy, md, ont, 0, dmun.p, dmuq.p, &mm);
    if (mumax) *mumax = mm;
    size_t m1 = (size_t)mv->nicell * 51 * 27, m2 = (size_t)mv->nicell * 27 * 27;
    md.H.download(H, m1), md.B.download(B, m1), md.Ax.download(Abdx, m1), md.Ay.download(Abdy, m1);
    md.Az.download(Abdz, m1), md.Dx.download(Dx, m1), md.Dy.download(Dy, m1), md.Dz.download(Dz, m1);
    md.G.download(G, (size_t)mv->nicell * 51 * 24);
    md.Htx.download(Htx, m2), md.Hty.download(Hty, m2), md.Htz.download(Htz, m2);
    TT_HIP(hipDeviceSynchronize());
    return 0;
  });
}

extern "C" int tt_last_integral_stats(double *ms, int64_t *domain_points, int64_t *boundary_points) {
  if (ms) *ms = g_last_macro.ms;
  if (domain_points) *domain_points = g_last_macro.dom_pts;
  if (boundary_points) *boundary_points = g_last_macro.bnd_pts;
  return 0;
}

extern "C" int tt_sd_integrals(const tt_mesh_view *mv, const tt_quad *q, int32_t row0, int32_t row1, double *H,
                               double *Htx, double *Hty, double *Htz, double *Dx, double *Dy, double *Dz,
                               double ont[10]) {
  return guarded("tt_sd_integrals", [&] {
    if (!mv || !q) throw std::runtime_error("NULL argument");
    QuadDev qh;
    build_quad(*q, qh);
    DevMesh dm;
    upload_mesh(*mv, dm, 0);
    SdDev sd;
    sd_integrals_dev(dm, *mv, qh, row0, row1, sd, ont, 0);
    size_t nr = row1 - row0;
    sd.H.download(H, nr * mv->nbnodes), sd.Htx.download(Htx, nr * mv->nbnodes);
    sd.Hty.download(Hty, nr * mv->nbnodes), sd.Htz.download(Htz, nr * mv->nbnodes);
    sd.Dx.download(Dx, nr * mv->nnodes), sd.Dy.download(Dy, nr * mv->nnodes), sd.Dz.download(Dz, nr * mv->nnodes);
    TT_HIP(hipDeviceSynchronize());
    return 0;
  });
}
