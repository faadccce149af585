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

// Macro matrices in three launches, one (cell, source) row at a time:
//   macro_bnd_kernel  one thread per row: the boundary integrals over the
//                     cell's six faces (H, G, Abd, Ht before corrections);
//   macro_dom_kernel  one wavefront per row: the domain integral (B, D);
//   macro_corr_kernel one thread per row: free term, Ht row sums and
//                     linear-field corrections, accuracy estimates.
// Helm: FMATmodHelm (lib_mHelm.f:2-483) with mu of each node / flux node in
// mun / muq; otherwise FMATkmdcSteady (lib_fmat.f:3-446).
template <bool Helm>
__global__ void __launch_bounds__(64) macro_bnd_kernel(MeshPtrs m, const QuadDev *__restrict__ qp, int ncomp,
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
  if (npts) atomicAdd(&npts[1], (unsigned long long)pts_b);
}

// the source point of row (ic, it) and its mu
template <bool Helm>
__device__ __forceinline__ void macro_source(const MeshPtrs &m, int ic, int it, const double *__restrict__ mun,
                                             const double *__restrict__ muq, double *p, double &mu) {
  const size_t nc = m.nc;
  mu = 0.0;
  if (it < 27) {
    const int dn = m.idc[ic + nc * it];
    p[0] = m.x[dn - 1], p[1] = m.x[dn - 1 + m.nn], p[2] = m.x[dn - 1 + 2 * (size_t)m.nn];
    if (Helm) mu = mun[dn - 1];
  } else {
    const int its = (it - 27) / 4, itf = (it - 27) % 4 + 1;
    const int dn = m.ibf[ic + nc * (its + 6 * (itf - 1))];
    p[0] = m.xq[dn - 1], p[1] = m.xq[dn - 1 + m.nq], p[2] = m.xq[dn - 1 + 2 * (size_t)m.nq];
    if (Helm) mu = muq[dn - 1];
  }
}

template <bool Helm>
__global__ void __launch_bounds__(64) macro_dom_kernel(MeshPtrs m, const QuadDev *__restrict__ qp,
                                                       const double *__restrict__ mun, const double *__restrict__ muq,
                                                       MacroOut o, unsigned long long *__restrict__ npts) {
  __shared__ __align__(16) double lds[wave_cell_lds<4>()];
  const int ic = blockIdx.x, it = blockIdx.y;
  const size_t nc = m.nc, m1 = nc * 51, i = (size_t)ic * 51 + it;
  double p[3], mu;
  macro_source<Helm>(m, ic, it, mun, muq, p, mu);
  double cc[8][3];
  for (int k = 0; k < 8; ++k) {
    const int node = m.idc[ic + nc * (kGeo8[k] - 1)];
    for (int d = 0; d < 3; ++d) cc[k][d] = m.x[(node - 1) + (size_t)m.nn * d];
  }
  WaveCell<4, Helm> w;
  w.run(*qp, p, cc, it + 1, mu, lds, false);  // the regular sub-cubes: macro_reg_kernel
  const int lane = threadIdx.x;
  if (lane < 27) {
    o.Dx[i + m1 * lane] = w.acc[0];
    o.Dy[i + m1 * lane] = w.acc[1];
    o.Dz[i + m1 * lane] = w.acc[2];
    o.B[i + m1 * lane] = w.acc[3];
  }
  if (npts) {
    long n = w.npts;
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    if (lane == 0) atomicAdd(&npts[0], (unsigned long long)n);
  }
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

// The domain integral over the sub-cubes that do not hold the source, for all
// 51 sources of a cell at once (one workgroup per cell, four wavefronts of 16
// sources).  These sub-cubes' Gauss points do not depend on the source: their
// positions and weighted Jacobians are computed once per cell into LDS, and
// the contraction sum_points shape(point) * kernel(source, point) runs on the
// FP64 matrix cores (16x16x4: shapes x points x sources), with the kernel
// values of each point evaluated as WaveCell does.  A source on or inside a
// sub-cube's closure is integrated over that sub-cube by macro_dom_kernel;
// this kernel adds the rest to its sums (so it runs after it).
constexpr int kRegChunk = 512;  // Gauss points staged per pass
template <bool Helm>
__global__ void __launch_bounds__(256) macro_reg_kernel(MeshPtrs m, const QuadDev *__restrict__ qp,
                                                        const double *__restrict__ mun, const double *__restrict__ muq,
                                                        MacroOut o, unsigned long long *__restrict__ npts) {
  __shared__ double4 geo[kRegChunk];  // x, y, z, weight * Jacobian
  __shared__ double4 loc[kRegChunk];  // local coordinates of the point (w: 0 for padding)
  const QuadDev &q = *qp;
  const int ic = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, n = lane & 15, kq = lane >> 4;
  const int it = 16 * w + n;  // this lane's source row (kernel operand)
  const bool valid = it < 51;
  const size_t nc = m.nc, m1 = nc * 51, i = (size_t)ic * 51 + (valid ? it : 0);
  double p[3] = {0.0, 0.0, 0.0}, mu = 0.0;
  if (valid) macro_source<Helm>(m, ic, it, mun, muq, p, mu);
  double cc[8][3];
  for (int k = 0; k < 8; ++k) {
    const int node = m.idc[ic + nc * (kGeo8[k] - 1)];
    for (int d = 0; d < 3; ++d) cc[k][d] = m.x[(node - 1) + (size_t)m.nn * d];
  }
  // the shape functions of this lane's operand rows 16t + n (node positions)
  int sa[2], sb[2], sc[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int sf = 16 * t + n;
    sa[t] = sf < 27 ? kNode27[sf][0] + 1 : 1, sb[t] = sf < 27 ? kNode27[sf][1] + 1 : 1;
    sc[t] = sf < 27 ? kNode27[sf][2] + 1 : 1;
  }
  dbl4 acc[4][2];
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c][0] = acc[c][1] = dbl4{0.0, 0.0, 0.0, 0.0};
  long cnt = 0;
  const int nd = q.iiDiv, n1 = q.ndr, np = n1 * n1 * n1;
  for (int ix = 1; ix <= nd; ++ix)
    for (int iy = 1; iy <= nd; ++iy)
      for (int iz = 1; iz <= nd; ++iz) {
        const double a = -1.0 + (ix - 1) * 2.0 / nd, b = -1.0 + ix * 2.0 / nd;
        const double c0 = -1.0 + (iy - 1) * 2.0 / nd, d = -1.0 + iy * 2.0 / nd;
        const double e = -1.0 + (iz - 1) * 2.0 / nd, f = -1.0 + iz * 2.0 / nd;
        const double dex = 0.125 * (b - a) * (d - c0) * (f - e);
        // as WaveCell::run decides: the source's position in the sub-cube
        bool use = valid;
        if (valid) {
          const double xx1 = ((-1.0 + 2.0 * kKsi51[it]) - a) / (b - a);
          const double yy1 = ((-1.0 + 2.0 * kEta51[it]) - c0) / (d - c0);
          const double zz1 = ((-1.0 + 2.0 * kZet51[it]) - e) / (f - e);
          use = !(xx1 >= 0.0 && xx1 <= 1.0 && yy1 >= 0.0 && yy1 <= 1.0 && zz1 >= 0.0 && zz1 <= 1.0);
        }
        if (use && kq == 0) cnt += np;
        for (int base = 0; base < np; base += kRegChunk) {
          const int len = min(kRegChunk, np - base), len4 = (len + 3) / 4 * 4;
          __syncthreads();  // the previous chunk's reads are done
          for (int j = threadIdx.x; j < len4; j += 256) {
            double4 g{0.0, 0.0, 0.0, 0.0}, l{0.0, 0.0, 0.0, 0.0};
            if (j < len) {
              const int t = base + j;
              const int I = t % n1, J = (t / n1) % n1, K = t / (n1 * n1);
              const double gii = a + 0.5 * (q.gdr[I] + 1.0) * (b - a);
              const double gij = c0 + 0.5 * (q.gdr[J] + 1.0) * (d - c0);
              const double gik = e + 0.5 * (q.gdr[K] + 1.0) * (f - e);
              const double wpre = 1.953125e-3 * q.odr[I] * (q.odr[J] * q.odr[K] * dex);
              double xc[3], dj;
              cell_map(cc, gii, gij, gik, xc, &dj);
              g = double4{xc[0], xc[1], xc[2], wpre * dj};
              l = double4{gii, gij, gik, 1.0};
            }
            geo[j] = g, loc[j] = l;
          }
          __syncthreads();
          for (int k0 = 0; k0 < len4; k0 += 4) {
            const double4 g = geo[k0 + kq], l = loc[k0 + kq];
            double lx[3], ly[3], lz[3];
            lagrange3(l.x, lx);
            lagrange3(l.y, ly);
            lagrange3(l.z, lz);
            double ph[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) ph[t] = lx[sa[t]] * ly[sb[t]] * lz[sc[t]];
            double fq[4] = {0.0, 0.0, 0.0, 0.0};
            if (use && l.w != 0.0) {
              const double rx = p[0] - g.x, ry = p[1] - g.y, rz = p[2] - g.z;
              double ra = sqrt(rx * rx + ry * ry + rz * rz);
              if (ra < 1.0e-12) ra = 1.0e-12;
              const double ira2 = 1.0 / (ra * ra);
              double fung, gf;
              kernel_factors<Helm>(ra, mu, fung, gf);
              fq[0] = rx * ira2 * fung * gf * g.w;
              fq[1] = ry * ira2 * fung * gf * g.w;
              fq[2] = rz * ira2 * fung * gf * g.w;
              fq[3] = fung * g.w;
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              acc[c][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ph[0], fq[c], acc[c][0], 0, 0, 0);
              acc[c][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ph[1], fq[c], acc[c][1], 0, 0, 0);
            }
          }
        }
      }
  // register r of tile t in lane l: shape function 16t + 4r + l/16, source l % 16
  if (valid) {
    double *out[4] = {o.Dx, o.Dy, o.Dz, o.B};
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int sf = 16 * t + 4 * r + kq;
        if (sf >= 27) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) out[c][i + m1 * sf] += acc[c][t][r];
      }
  }
  if (npts) {
    long v = cnt;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) atomicAdd(&npts[0], (unsigned long long)v);
  }
}

template <bool Helm>
__global__ void __launch_bounds__(64) macro_corr_kernel(MeshPtrs m, int ncomp, const double *__restrict__ mun,
                                                        const double *__restrict__ muq, MacroOut o) {
  const int ic = blockIdx.x * 64 + threadIdx.x;
  const int it = blockIdx.y;  // source 0..50
  if (ic >= ncomp) return;
  const size_t nc = m.nc, m1 = nc * 51, m2 = nc * 27;
  const size_t i = (size_t)ic * 51 + it, ii = (size_t)ic * 27 + it;
  const bool usrc = it < 27;
  auto X = [&](int node, int d) { return m.x[(node - 1) + (size_t)m.nn * d]; };
  auto IDC = [&](int k) { return m.idc[ic + nc * (k - 1)]; };
  double p[3], mu;
  macro_source<Helm>(m, ic, it, mun, muq, p, mu);
  const int its = usrc ? 0 : (it - 27) / 4, itf = usrc ? 0 : (it - 27) % 4 + 1;
  double *ont = o.ont + i * 12;
  for (int k = 0; k < 11; ++k) ont[k] = 0.0;
  ont[11] = mu;
  // ---- free term (rigid body u = 1; modified Helmholtz: f = -mu^2) and Ht row sums
  double cfree = 0.0;
  if (Helm) {
    double vH = 0.0, vB = 0.0;
    for (int j = 0; j < 27; ++j) vH += o.H[i + m1 * j];
    for (int j = 0; j < 27; ++j) vB += o.B[i + m1 * j];
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
  if (mun && muq) {
    hipLaunchKernelGGL(macro_bnd_kernel<true>, dim3(ceil_div(ncomp, 64), 51), dim3(64), 0, s, ptrs(m), dq.p, ncomp,
                       mun, muq, out, cnt.p);
    TT_LAUNCH_CHECK();
    hipLaunchKernelGGL(macro_dom_kernel<true>, dim3(ncomp, 51), dim3(64), 0, s, ptrs(m), dq.p, mun, muq, out, cnt.p);
    TT_LAUNCH_CHECK();
    hipLaunchKernelGGL(macro_reg_kernel<true>, dim3(ncomp), dim3(256), 0, s, ptrs(m), dq.p, mun, muq, out, cnt.p);
    TT_LAUNCH_CHECK();
    hipLaunchKernelGGL(macro_corr_kernel<true>, dim3(ceil_div(ncomp, 64), 51), dim3(64), 0, s, ptrs(m), ncomp, mun,
                       muq, out);
  } else {
    hipLaunchKernelGGL(macro_bnd_kernel<false>, dim3(ceil_div(ncomp, 64), 51), dim3(64), 0, s, ptrs(m), dq.p, ncomp,
                       nullptr, nullptr, out, cnt.p);
    TT_LAUNCH_CHECK();
    hipLaunchKernelGGL(macro_dom_kernel<false>, dim3(ncomp, 51), dim3(64), 0, s, ptrs(m), dq.p, nullptr, nullptr, out,
                       cnt.p);
    TT_LAUNCH_CHECK();
    hipLaunchKernelGGL(macro_reg_kernel<false>, dim3(ncomp), dim3(256), 0, s, ptrs(m), dq.p, nullptr, nullptr, out,
                       cnt.p);
    TT_LAUNCH_CHECK();
    hipLaunchKernelGGL(macro_corr_kernel<false>, dim3(ceil_div(ncomp, 64), 51), dim3(64), 0, s, ptrs(m), ncomp,
                       nullptr, nullptr, out);
  }
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

// Boundary integrals of the single-domain rows (H and the Ht cross products,
// the boundary part of FMATsdkm, lib_sdkm.f:1286-1530), spread over (row,
// face) pairs: one launch per colour of boundary faces that share no node,
// one thread per pair (rows fastest, so the column-major output is written
// coalesced).  Within a launch
// no two threads add into the same (row, node) entry; the colours run in a
// fixed order, so the sums are deterministic (they add a row's faces in colour
// order rather than face order: a rounding-level difference).
__global__ void __launch_bounds__(256) sd_bnd_pairs_kernel(MeshPtrs m, const QuadDev *__restrict__ qp, int row0,
                                                           int nrows, const int *__restrict__ faces, int nf,
                                                           SdOut o) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)nrows * nf) return;
  const int r = (int)(t % nrows), je = faces[t / nrows];
  const int dn = m.gbn[row0 + r];
  const double p[3] = {m.x[dn - 1], m.x[dn - 1 + m.nn], m.x[dn - 1 + 2 * (size_t)m.nn]};
  const size_t ne = m.ne, ld = o.ld;
  double c[4][3];
  int nodes[9];
  for (int k = 0; k < 9; ++k) nodes[k] = m.ibc[je + ne * k];
  for (int k = 0; k < 4; ++k)
    for (int d = 0; d < 3; ++d) c[k][d] = m.x[nodes[2 * k] - 1 + (size_t)m.nn * d];
  int isrc = 0;
  for (int q = 0; q < 9; ++q)
    if (nodes[q] == dn) isrc = q + 1;
  BndAcc<false> acc;
  acc.zero();
  integrate_face<false>(*qp, p, c, isrc, 0.0, acc);
  for (int q = 0; q < 9; ++q) {
    const size_t j = m.lbn[nodes[q] - 1] - 1;
    o.H[r + ld * j] += acc.he[q];
    o.Htx[r + ld * j] -= acc.heyz[q];
    o.Hty[r + ld * j] -= acc.hezx[q];
    o.Htz[r + ld * j] -= acc.hexy[q];
  }
}

// ---- regular domain integrals of the single-domain rows on the FP64 matrix cores.
// The regular rule (INTEDc27dcLap with the node outside the cell) uses the same
// reference-cell points in every cell, so the 27 x 3 integrals of a row over a
// cell are sum_p phi_s(p) K_d(row, p), with phi the 27 shape functions at the
// points and K_d = r_d / (4 pi r^3) w_p |J(p)| (the kernel of DomAcc<3>).  Per
// cell that is a (32 x P) * (P x rows) product per component: the shape table
// is the A operand and the kernel values, one (row, point) pair per lane, are
// the B operand of v_mfma_f64_16x16x4f64.  Cells of one colour share no node,
// so the waves of one launch add into disjoint entries of D.

// geo[(c*Pp + p)*4 + {0,1,2,3}] = position and weight w_p |J| of point p in
// cell c (Pp: the point count padded to a multiple of 16; padding points have
// weight 0); phi[p*32 + s] = shape function s at
// point p (0 for s >= 27 and padding points).
__global__ void sd_geom_kernel(MeshPtrs m, const QuadDev *__restrict__ qp, int Pp, double *__restrict__ geo,
                               double *__restrict__ phi) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)m.nc * Pp) return;
  const int c = (int)(t / Pp), p = (int)(t % Pp);
  const QuadDev &q = *qp;
  const int n1 = q.ndr, P = n1 * n1 * n1;
  const size_t nc = m.nc;
  double cc[8][3];
  for (int k = 0; k < 8; ++k) {
    const int node = m.idc[c + nc * (kGeo8[k] - 1)];
    for (int d = 0; d < 3; ++d) cc[k][d] = m.x[node - 1 + (size_t)m.nn * d];
  }
  double *g = geo + t * 4;
  double f27[27];
  if (p < P) {
    const int I = p % n1, J = (p / n1) % n1, K = p / (n1 * n1);
    // one sub-cell: a = -1, b = 1 in every direction, dex = 1
    const double gii = -1.0 + 0.5 * (q.gdr[I] + 1.0) * 2.0;
    const double gij = -1.0 + 0.5 * (q.gdr[J] + 1.0) * 2.0;
    const double gik = -1.0 + 0.5 * (q.gdr[K] + 1.0) * 2.0;
    double xc[3], dj;
    cell_map(cc, gii, gij, gik, xc, &dj);
    g[0] = xc[0], g[1] = xc[1], g[2] = xc[2];
    g[3] = 1.953125e-3 * dj * q.odr[I] * (q.odr[J] * q.odr[K]);
    if (c == 0) shape27(gii, gij, gik, f27);
  } else {
    g[0] = cc[0][0], g[1] = cc[0][1], g[2] = cc[0][2], g[3] = 0.0;
    for (int s = 0; s < 27; ++s) f27[s] = 0.0;
  }
  if (c == 0)
    for (int s = 0; s < 32; ++s) phi[(size_t)p * 32 + s] = s < 27 ? f27[s] : 0.0;
}

// Block: 4 wavefronts, 16 rows each; blockIdx.y: a chunk of `chunk` cells of
// the colour list `cells`.  Lane l evaluates row l%16 at point k0 + l/16.  The
// row's own cells (singular) get zero weight here.
// NP: 16-point passes per cell (Pp / 16) when known at compile time, so the
// products of a cell form one unrolled chain with the accumulators in place;
// NP = 0 loops over Pp / 16 passes.
template <int NP>
__global__ void __launch_bounds__(256) sd_dom_mfma_kernel(MeshPtrs m, int row0, int nrows, int Pp,
                                                          const double *__restrict__ geo,
                                                          const double *__restrict__ phi,
                                                          const int *__restrict__ cells, int ncells, int chunk,
                                                          const unsigned char *__restrict__ touches, SdOut o) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n = lane & 15, kq = lane >> 4;
  const int r = blockIdx.x * 64 + w * 16 + n;
  const bool valid = r < nrows;
  const int dn = valid ? m.gbn[row0 + r] : 0;
  const size_t nn = m.nn, nc = m.nc, ld = o.ld;
  const double px = valid ? m.x[dn - 1] : 0.0, py = valid ? m.x[dn - 1 + nn] : 0.0,
               pz = valid ? m.x[dn - 1 + 2 * nn] : 0.0;
  double *D[3] = {o.Dx, o.Dy, o.Dz};
  const double k4pi = 0.25 / 3.141592653589793;
  const int c0 = blockIdx.y * chunk, c1 = min(ncells, c0 + chunk);
  // lane k < 27 holds node k of the cell; the next cell's ids are loaded one
  // cell ahead
  int cnext = c0 < c1 ? cells[c0] : 0;
  int nidnext = (c0 < c1 && lane < 27) ? m.idc[cnext + nc * lane] : 0;
  for (int ci = c0; ci < c1; ++ci) {
    const int c = cnext, nid = nidnext;
    if (ci + 1 < c1) {
      cnext = cells[ci + 1];
      nidnext = lane < 27 ? m.idc[cnext + nc * lane] : 0;
    }
    bool use = valid;
    if (touches[c])
      for (int k = 0; k < 27; ++k) use = use && __builtin_amdgcn_readlane(nid, k) != dn;
    // the entries of D this lane adds to (register i of tile t: shape
    // function 16t + 4i + kq, row n), read before the products
    size_t jo[2][4];
    double dv[3][2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int sf = 16 * t + 4 * i + kq;
        const int node = __shfl(nid, sf < 27 ? sf : 0, 64);
        jo[t][i] = (size_t)r + ld * (size_t)(node - 1);
#pragma unroll
        for (int d = 0; d < 3; ++d) dv[d][t][i] = (valid && sf < 27) ? D[d][jo[t][i]] : 0.0;
      }
    dbl4 acc[3][2];
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[d][0] = acc[d][1] = dbl4{0.0, 0.0, 0.0, 0.0};
    const double *g = geo + (size_t)c * Pp * 4;
    // 16 points (four k-steps) per pass: all loads of the pass first, then the
    // kernel values, then the 24 products
    const int npass = NP > 0 ? NP : Pp / 16;
    constexpr int kUnroll = NP > 0 ? NP : 1;
#pragma unroll kUnroll
    for (int ps = 0; ps < npass; ++ps) {
      const int k0 = 16 * ps;
      double gx[4], gy[4], gz[4], gw[4], a0[4], a1[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int p = k0 + 4 * q + kq;
        const double4 v = *reinterpret_cast<const double4 *>(g + p * 4);
        gx[q] = v.x, gy[q] = v.y, gz[q] = v.z, gw[q] = v.w;
        a0[q] = phi[p * 32 + n], a1[q] = phi[p * 32 + 16 + n];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double rx = px - gx[q], ry = py - gy[q], rz = pz - gz[q];
        const double r2 = fmax(rx * rx + ry * ry + rz * rz, 1.0e-24);
        // 1/r: hardware estimate and two Newton steps
        double y = __builtin_amdgcn_rsq(r2);
        y = fma(0.5 * y, fma(-r2 * y, y, 1.0), y);
        y = fma(0.5 * y, fma(-r2 * y, y, 1.0), y);
        const double f = use ? gw[q] * k4pi * (y * y * y) : 0.0;
        const double b0 = rx * f, b1 = ry * f, b2 = rz * f;
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[q], b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[q], b0, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[q], b1, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[q], b1, acc[1][1], 0, 0, 0);
        acc[2][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[q], b2, acc[2][0], 0, 0, 0);
        acc[2][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[q], b2, acc[2][1], 0, 0, 0);
      }
    }
    // register i of lane l holds C[4i + l/16][l%16] (the f64 16x16x4 output
    // layout): shape function 16t + 4i + kq, row n
    if (!valid) continue;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (16 * t + 4 * i + kq >= 27) continue;
#pragma unroll
        for (int d = 0; d < 3; ++d) D[d][jo[t][i]] = dv[d][t][i] + acc[d][t][i];
      }
  }
}

// one thread per singular (row, cell) pair -> staging[pair][3][27]
__global__ void __launch_bounds__(64) sd_dom_singular_kernel(MeshPtrs m, const QuadDev *__restrict__ qp,
                                                             const int *__restrict__ prow, const int *__restrict__ pcell,
                                                             const int *__restrict__ psrc, int npairs,
                                                             double *__restrict__ staging) {
  // one wavefront per (row, cell) pair: the pair's quadrature points are
  // evaluated 64 at a time (WaveCell, as for the macro domain integrals)
  __shared__ __align__(16) double lds[wave_cell_lds<3>()];
  const int k = blockIdx.x;
  if (k >= npairs) return;
  const size_t nc = m.nc;
  const int i = prow[k], ic = pcell[k];
  const int dn = m.gbn[i];
  const double p[3] = {m.x[dn - 1], m.x[dn - 1 + m.nn], m.x[dn - 1 + 2 * (size_t)m.nn]};
  double c[8][3];
  for (int t = 0; t < 8; ++t)
    for (int d = 0; d < 3; ++d) c[t][d] = m.x[m.idc[ic + nc * (kGeo8[t] - 1)] - 1 + (size_t)m.nn * d];
  WaveCell<3, false> w;
  w.run(*qp, p, c, psrc[k], 0.0, lds);
  const int lane = threadIdx.x;
  if (lane < 27)
    for (int t = 0; t < 3; ++t) staging[(size_t)k * 81 + t * 27 + lane] = w.acc[t];
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

// Free term, Ht row sums and linear-field corrections (lib_sdkm.f:1387-1522)
// in two passes.  sd_sums_kernel: per row and column slice (blockIdx.y of
// kSdSplit) the partial sums of the uncorrected rows:
//   0 sum H, 1..3 sum Ht_t, 4..6 sum H(s) x_c(s) (c = z, x, y),
//   7..9 sum D_t (t = z, x, y), 10..12 sum Htx y, Hty z, Htz x.
// sd_apply_kernel: one thread per row adds the slices in slice order and
// applies the corrections; the sums of the corrected rows that the estimates
// use follow from the uncorrected sums and the corrections.
constexpr int kSdSplit = 64, kSdSums = 13;

__global__ void __launch_bounds__(64) sd_sums_kernel(MeshPtrs m, int nrows, SdOut o, double *__restrict__ part) {
  const int r = blockIdx.x * 64 + threadIdx.x;
  if (r >= nrows) return;
  const size_t ld = o.ld, nb = m.nb, nn = m.nn;
  const int sp = blockIdx.y;
  auto X = [&](int node, int d) { return m.x[(node - 1) + nn * d]; };
  double v[kSdSums];
  for (int k = 0; k < kSdSums; ++k) v[k] = 0.0;
  const size_t b0 = nb * sp / kSdSplit, b1 = nb * (sp + 1) / kSdSplit;
  for (size_t j = b0; j < b1; ++j) {
    const int g = m.gbn[j];
    const double h = o.H[r + ld * j], tx = o.Htx[r + ld * j], ty = o.Hty[r + ld * j], tz = o.Htz[r + ld * j];
    v[0] += h;
    v[1] += tx, v[2] += ty, v[3] += tz;
    v[4] += h * X(g, 2), v[5] += h * X(g, 0), v[6] += h * X(g, 1);
    v[10] += tx * X(g, 1), v[11] += ty * X(g, 2), v[12] += tz * X(g, 0);
  }
  const size_t n0 = nn * sp / kSdSplit, n1 = nn * (sp + 1) / kSdSplit;
  for (size_t j = n0; j < n1; ++j) v[7] += o.Dz[r + ld * j], v[8] += o.Dx[r + ld * j], v[9] += o.Dy[r + ld * j];
  double *out = part + ((size_t)sp * nrows + r) * kSdSums;
  for (int k = 0; k < kSdSums; ++k) out[k] = v[k];
}

__global__ void sd_apply_kernel(MeshPtrs m, int row0, int nrows, SdOut o, const double *__restrict__ part,
                                double *__restrict__ ont) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const size_t ld = o.ld, nn = m.nn;
  const int i = row0 + r, dn = m.gbn[i];
  auto X = [&](int node, int d) { return m.x[(node - 1) + nn * d]; };
  double v[kSdSums];
  for (int k = 0; k < kSdSums; ++k) v[k] = 0.0;
  for (int sp = 0; sp < kSdSplit; ++sp) {
    const double *pp = part + ((size_t)sp * nrows + r) * kSdSums;
    for (int k = 0; k < kSdSums; ++k) v[k] += pp[k];
  }
  double *on = ont + (size_t)r * 10;
  // free term: rigid-body field u = 1
  const double cfree = -v[0];
  on[0] = cfree;
  o.H[r + ld * i] += cfree;
  double *Ht[3] = {o.Htx, o.Hty, o.Htz};
  for (int t = 0; t < 3; ++t) {
    on[1 + t] = fabs(v[1 + t]);
    Ht[t][r + ld * i] -= v[1 + t];
  }
  // linear fields (z,0,0), (0,x,0), (0,0,y): sum_s H_is x_c(s) = sum_j D_ij
  double *Dm[3] = {o.Dz, o.Dx, o.Dy};
  const int crd[3] = {2, 0, 1}, slot[3] = {6, 4, 5};
  double hv[3];
  for (int t = 0; t < 3; ++t) {
    hv[t] = v[4 + t] + cfree * X(dn, crd[t]);
    const double dlt = hv[t] - v[7 + t];
    on[slot[t]] = fabs(dlt);
    Dm[t][r + ld * (dn - 1)] = Dm[t][r + ld * (dn - 1)] + dlt;
  }
  // curl estimates on the corrected rows: sum_j D_t,ij is hv[t] after the correction
  const double hxy = v[10] - v[1] * X(dn, 1), hyz = v[11] - v[2] * X(dn, 2), hzx = v[12] - v[3] * X(dn, 0);
  on[7] = fabs(hxy + hv[0]), on[8] = fabs(hyz + hv[1]), on[9] = fabs(hzx + hv[2]);
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
  // boundary faces: colour them so that no two of one colour share a node
  // (greedy, first free colour), then one launch per colour over (row, face)
  {
    const int ne = m.ne;
    std::vector<uint64_t> used(m.nn, 0);
    std::vector<int> colour(ne);
    int ncol = 0;
    for (int je = 0; je < ne; ++je) {
      uint64_t mask = 0;
      for (int k = 0; k < 9; ++k) mask |= used[hv.ibc[je + (size_t)ne * k] - 1];
      if (~mask == 0) throw std::runtime_error("face colouring needs more than 64 colours");
      const int col = __builtin_ctzll(~mask);
      for (int k = 0; k < 9; ++k) used[hv.ibc[je + (size_t)ne * k] - 1] |= 1ull << col;
      colour[je] = col, ncol = std::max(ncol, col + 1);
    }
    std::vector<int> fstart(ncol + 1, 0), flist(ne);
    for (int je = 0; je < ne; ++je) ++fstart[colour[je] + 1];
    for (int k = 0; k < ncol; ++k) fstart[k + 1] += fstart[k];
    {
      std::vector<int> fill(fstart.begin(), fstart.end() - 1);
      for (int je = 0; je < ne; ++je) flist[fill[colour[je]]++] = je;
    }
    DBuf<int> dfaces;
    dfaces.upload(flist, s);
    for (int k = 0; k < ncol; ++k) {
      const int nf = fstart[k + 1] - fstart[k];
      if (nf == 0) continue;
      hipLaunchKernelGGL(sd_bnd_pairs_kernel, dim3(ceil_div((size_t)nr * nf, 256)), dim3(256), 0, s, ptrs(m), dq.p,
                         row0, nr, dfaces.p + fstart[k], nf, out);
    }
    TT_LAUNCH_CHECK();
    TT_HIP(hipStreamSynchronize(s));  // dfaces is released on return
  }
  // regular cells: colour the cells so that no two of one colour share a node
  // (greedy, first free colour), then one product launch per colour
  {
    const int n1 = qh.ndr, Pp = (n1 * n1 * n1 + 15) / 16 * 16;
    std::vector<uint64_t> used(m.nn, 0);
    std::vector<int> colour(m.nc);
    std::vector<unsigned char> touches(m.nc, 0);
    int ncol = 0;
    for (int ic = 0; ic < m.nc; ++ic) {
      uint64_t mask = 0;
      for (int k = 0; k < 27; ++k) {
        const int nd = hv.idc[ic + (size_t)m.nc * k] - 1;
        mask |= used[nd];
        touches[ic] |= hv.lbn[nd] != 0;
      }
      if (~mask == 0) throw std::runtime_error("cell colouring needs more than 64 colours");
      const int col = __builtin_ctzll(~mask);
      colour[ic] = col, ncol = std::max(ncol, col + 1);
      for (int k = 0; k < 27; ++k) used[hv.idc[ic + (size_t)m.nc * k] - 1] |= 1ull << col;
    }
    std::vector<int> cstart(ncol + 1, 0), clist(m.nc);
    for (int ic = 0; ic < m.nc; ++ic) ++cstart[colour[ic] + 1];
    for (int k = 0; k < ncol; ++k) cstart[k + 1] += cstart[k];
    {
      std::vector<int> fill(cstart.begin(), cstart.end() - 1);
      for (int ic = 0; ic < m.nc; ++ic) clist[fill[colour[ic]]++] = ic;
    }
    DBuf<double> geo((size_t)m.nc * Pp * 4), phi((size_t)Pp * 32);
    DBuf<int> dcl;
    DBuf<unsigned char> dtouch;
    dcl.upload(clist, s), dtouch.upload(touches, s);
    hipLaunchKernelGGL(sd_geom_kernel, dim3(ceil_div((int)((size_t)m.nc * Pp), 256)), dim3(256), 0, s, ptrs(m), dq.p,
                       Pp, geo.p, phi.p);
    TT_LAUNCH_CHECK();
    const int bx = ceil_div(nr, 64);
    for (int k = 0; k < ncol; ++k) {
      const int cnt = cstart[k + 1] - cstart[k];
      const int chunk = std::max(1, (int)(((int64_t)cnt * bx + 4095) / 4096));
      const dim3 grid(bx, ceil_div(cnt, chunk));
      auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, ptrs(m), row0, nr, Pp, geo.p, phi.p, dcl.p + cstart[k], cnt,
                           chunk, dtouch.p, out);
      };
      switch (Pp / 16) {
        case 1: launch(sd_dom_mfma_kernel<1>); break;
        case 2: launch(sd_dom_mfma_kernel<2>); break;
        case 4: launch(sd_dom_mfma_kernel<4>); break;
        case 8: launch(sd_dom_mfma_kernel<8>); break;
        case 14: launch(sd_dom_mfma_kernel<14>); break;
        default: launch(sd_dom_mfma_kernel<0>); break;
      }
      TT_LAUNCH_CHECK();
    }
    TT_HIP(hipStreamSynchronize(s));
  }
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
    hipLaunchKernelGGL(sd_dom_singular_kernel, dim3(np), dim3(64), 0, s, ptrs(m), dq.p, dprow.p,
                       dpcell.p, dpsrc.p, np, staging.p);
    TT_LAUNCH_CHECK();
    hipLaunchKernelGGL(sd_dom_singular_apply, dim3(ceil_div(nr, 64)), dim3(64), 0, s, ptrs(m), row0, nr, doff.p,
                       dpcell.p, staging.p, out);
    TT_LAUNCH_CHECK();
  }
  DBuf<double> ontd((size_t)nr * 10), part((size_t)kSdSplit * nr * kSdSums);
  hipLaunchKernelGGL(sd_sums_kernel, dim3(ceil_div(nr, 64), kSdSplit), dim3(64), 0, s, ptrs(m), nr, out, part.p);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(sd_apply_kernel, dim3(ceil_div(nr, 64)), dim3(64), 0, s, ptrs(m), row0, nr, out, part.p,
                     ontd.p);
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
