// tt_mhsd.hip -- the single-domain modified-Helmholtz path of the reference
// (EQNS imH 3-7: main.f90:556-777 and the solves of :1345-1395, lib_fm.f) on
// gfx950.
//
// Integrals (FMATsdmHfs, lib_fm.f:1183-1645).  Rows are the nodes (U rows,
// source x(it), mu = sqrt(beta/diff(it))) and the boundary flux points (Q
// rows, element-major: row 4e + f, source xq(ibcf(e, f)), mu from qdiff).
// Columns: boundary nodes (H, Abdx/y/z), boundary flux points (G) and nodes
// (B, Bp, Dx/y/z); every matrix column-major with the row count as leading
// dimension, resident in HBM.
//   * faces: one thread per (row, element) of one element colour (elements
//     of a colour share no node, so the adds into H and Abd do not collide);
//     the subdivided polar quadrature of INTEBc9modHelm (tt_quadrature.h);
//   * regular cells: the reference-cell points are the same in every cell, so
//     the 27 x 4 sums of a row over a cell are a (32 x P) * (P x rows) product
//     per output on the FP64 matrix cores (v_mfma_f64_16x16x4f64), one cell
//     colour per launch, the row's own cells weighted zero;
//   * singular (row, cell) pairs (the source is a node or flux point of the
//     cell): one wavefront per pair through WaveCell, staged, then added per
//     row in cell order;
//   * free terms and linear-field corrections (lib_fm.f:1379-1444, 1606-1634)
//     from per-row column sums taken in slices and added in slice order;
//   * MakeBprime (lib_fm.f:800-858) elementwise.
// System (fmmHSysM, lib_fm.f:1061-1181): one thread per (equation, column)
// writes A and RHS; FormPREfm's LU (pret 2) is rocSOLVER getrf with the
// explicit inverse (getri), so each solve is one streaming product.
// Solve (fmSolveModHelmDC and its matrix-vector variants): r = RHS b plus the
// eight right-hand-side products of the equation rows -- dense: one fused
// pass over [B Bp Dx Dy Dz Abdx Abdy Abdz]; compressed: the wavelet, cross
// approximation or block-matrix product of each matrix (tt_compress) -- then
// x = A^-1 r and the under-relaxed update of u and q.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "tt_compress.h"
#include "tt_integrals.h"
#include "tt_sums.h"

namespace tt {
namespace {

__constant__ const int kGeo8m[8] = {6, 7, 2, 3, 5, 8, 1, 4};
constexpr int kMats = 10;  // H, G, B, Bp, Abdx, Abdy, Abdz, Dx, Dy, Dz
enum { MH = 0, MG, MB, MBp, MAx, MAy, MAz, MDx, MDy, MDz };
constexpr int kMaxSing = 16;  // cells around one node or flux node
constexpr int kSplit = 64, kSums = 11;
constexpr int kRows = 64;

// One row set (U: nodes, Q: boundary flux points) as kernel argument.
struct RowSet {
  int n;                     // rows = leading dimension of the matrices
  const double4 *src;        // x, y, z, mu of each row's source point
  const int *unode, *qnode;  // 1-based node / flux-node id of the source (U / Q rows; other NULL)
  double *M[kMats];
};

// ---- boundary integrals: INTEBc9modHelm per (row, element)
__global__ void __launch_bounds__(64) mhsd_bnd_kernel(MeshPtrs m, const QuadDev *__restrict__ qp, RowSet rs,
                                                      const int *__restrict__ elems, int nel,
                                                      const int *__restrict__ ibcf4,
                                                      const double *__restrict__ minedge) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)rs.n * nel) return;
  const int r = (int)(t % rs.n), je = elems[t / rs.n];
  const size_t ne = m.ne, ld = rs.n;
  int nodes[9];
  for (int k = 0; k < 9; ++k) nodes[k] = m.ibc[je + ne * k];
  double c[4][3];
  for (int k = 0; k < 4; ++k)
    for (int d = 0; d < 3; ++d) c[k][d] = m.x[nodes[2 * k] - 1 + (size_t)m.nn * d];
  const double4 s = rs.src[r];
  const double p[3] = {s.x, s.y, s.z};
  int isrc = 0;
  if (rs.unode) {
    const int un = rs.unode[r];
    for (int k = 0; k < 9; ++k)
      if (nodes[k] == un) isrc = k + 1;
  } else {
    const int qn = rs.qnode[r];
    for (int k = 0; k < 4; ++k)
      if (ibcf4[(size_t)je * 4 + k] == qn) isrc = k + 10;
  }
  BndAcc<true, true> acc;
  acc.zero();
  acc.mu = s.w;
  integrate_face<true>(*qp, p, c, isrc, minedge[je], acc);
  for (int k = 0; k < 9; ++k) {
    const size_t j = (size_t)(m.lbn[nodes[k] - 1] - 1);
    rs.M[MH][r + ld * j] += acc.he[k];
    rs.M[MAx][r + ld * j] += acc.ad[0][k];
    rs.M[MAy][r + ld * j] += acc.ad[1][k];
    rs.M[MAz][r + ld * j] += acc.ad[2][k];
  }
  for (int k = 0; k < 4; ++k) rs.M[MG][r + ld * ((size_t)je * 4 + k)] += acc.ge[k];
}

// ---- regular cells on the FP64 matrix cores
typedef double dbl4 __attribute__((ext_vector_type(4)));

// geo[(c*Pp + p)*4 + {0,1,2,3}] = position and weight w_p |J| of point p in
// cell c (padding points weight 0); phi[p*32 + s] = shape function s at point p
__global__ void mhsd_geom_kernel(MeshPtrs m, const QuadDev *__restrict__ qp, int Pp, double *__restrict__ geo,
                                 double *__restrict__ phi) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)m.nc * Pp) return;
  const int c = (int)(t / Pp), p = (int)(t % Pp);
  const QuadDev &q = *qp;
  const int n1 = q.ndr, P = n1 * n1 * n1;
  const size_t nc = m.nc;
  double cc[8][3];
  for (int k = 0; k < 8; ++k) {
    const int node = m.idc[c + nc * (kGeo8m[k] - 1)];
    for (int d = 0; d < 3; ++d) cc[k][d] = m.x[node - 1 + (size_t)m.nn * d];
  }
  double *g = geo + t * 4;
  double f27[27];
  if (p < P) {
    const int I = p % n1, J = (p / n1) % n1, K = p / (n1 * n1);
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

// Block: 4 wavefronts of 16 rows; blockIdx.y: a chunk of cells of one colour.
// Lane l evaluates the kernel of row l%16 at point k0 + l/16: the B operand of
// the products, the shape table the A operand.  Outputs B (u*), Dx, Dy, Dz
// (r_d / r^2 u* (1 + mu r)), as DomAcc<4, true>.
template <int NP>
__global__ void __launch_bounds__(256) mhsd_dom_mfma_kernel(MeshPtrs m, RowSet rs, const int *__restrict__ sing,
                                                            int Pp, const double *__restrict__ geo,
                                                            const double *__restrict__ phi,
                                                            const int *__restrict__ cells, int ncells, int chunk) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n = lane & 15, kq = lane >> 4;
  const int r = blockIdx.x * 64 + w * 16 + n;
  const bool valid = r < rs.n;
  const double4 s = valid ? rs.src[r] : make_double4(0.0, 0.0, 0.0, 0.0);
  const size_t nc = m.nc, ld = rs.n;
  double *out[4] = {rs.M[MB], rs.M[MDx], rs.M[MDy], rs.M[MDz]};
  const double k4pi = 0.25 / 3.141592653589793;
  const int c0 = blockIdx.y * chunk, c1 = min(ncells, c0 + chunk);
  // the row's own cells (singular pairs, integrated elsewhere)
  int sc[kMaxSing];
#pragma unroll
  for (int k = 0; k < kMaxSing; ++k) sc[k] = valid ? sing[(size_t)r * kMaxSing + k] : -1;
  for (int ci = c0; ci < c1; ++ci) {
    const int c = cells[ci];
    const int nid = lane < 27 ? m.idc[c + nc * lane] : 0;
    bool use = valid;
#pragma unroll
    for (int k = 0; k < kMaxSing; ++k) use = use && sc[k] != c;
    dbl4 acc[4][2];
#pragma unroll
    for (int o = 0; o < 4; ++o) acc[o][0] = acc[o][1] = dbl4{0.0, 0.0, 0.0, 0.0};
    const double *g = geo + (size_t)c * Pp * 4;
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
        const double rx = s.x - gx[q], ry = s.y - gy[q], rz = s.z - gz[q];
        const double r2 = fmax(rx * rx + ry * ry + rz * rz, 1.0e-24);
        // 1/r: hardware estimate and two Newton steps
        double y = __builtin_amdgcn_rsq(r2);
        y = fma(0.5 * y, fma(-r2 * y, y, 1.0), y);
        y = fma(0.5 * y, fma(-r2 * y, y, 1.0), y);
        const double ra = r2 * y;
        const double fung = k4pi * y * exp(-s.w * ra);
        const double fb = use ? gw[q] * fung : 0.0;
        const double fd = use ? gw[q] * (fung * (1.0 + s.w * ra)) * (y * y) : 0.0;
        const double b[4] = {fb, rx * fd, ry * fd, rz * fd};
#pragma unroll
        for (int o = 0; o < 4; ++o) {
          acc[o][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[q], b[o], acc[o][0], 0, 0, 0);
          acc[o][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[q], b[o], acc[o][1], 0, 0, 0);
        }
      }
    }
    // register i of lane l holds C[4i + l/16][l%16]: shape function 16t + 4i + kq, row n
    int node[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int sf = 16 * t + 4 * i + kq;
        node[t][i] = __shfl(nid, sf < 27 ? sf : 0, 64);
      }
    if (!valid) continue;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (16 * t + 4 * i + kq >= 27) continue;
        const size_t jo = (size_t)r + ld * (size_t)(node[t][i] - 1);
#pragma unroll
        for (int o = 0; o < 4; ++o) out[o][jo] += acc[o][t][i];
      }
  }
}

// ---- singular pairs: one wavefront per (row, cell) -> staging[pair][4][27]
__global__ void __launch_bounds__(64) mhsd_dom_singular_kernel(MeshPtrs m, const QuadDev *__restrict__ qp, RowSet rs,
                                                               const int *__restrict__ prow,
                                                               const int *__restrict__ pcell,
                                                               const int *__restrict__ psrc, int npairs,
                                                               double *__restrict__ staging) {
  __shared__ __align__(16) double lds[wave_cell_lds<4>()];
  const int k = blockIdx.x;
  if (k >= npairs) return;
  const size_t nc = m.nc;
  const int r = prow[k], ic = pcell[k];
  const double4 s = rs.src[r];
  const double p[3] = {s.x, s.y, s.z};
  double c[8][3];
  for (int t = 0; t < 8; ++t)
    for (int d = 0; d < 3; ++d) c[t][d] = m.x[m.idc[ic + nc * (kGeo8m[t] - 1)] - 1 + (size_t)m.nn * d];
  WaveCell<4, true> w;
  w.run(*qp, p, c, psrc[k], s.w, lds);
  const int lane = threadIdx.x;
  if (lane < 27)
    for (int t = 0; t < 4; ++t) staging[(size_t)k * 108 + t * 27 + lane] = w.acc[t];
}

// adds each row's singular pairs in cell order: B, Dx, Dy, Dz
__global__ void mhsd_singular_apply(MeshPtrs m, RowSet rs, const int *__restrict__ off,
                                    const int *__restrict__ pcell, const double *__restrict__ staging) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rs.n) return;
  const size_t nc = m.nc, ld = rs.n;
  for (int k = off[r]; k < off[r + 1]; ++k) {
    const int ic = pcell[k];
    const double *st = staging + (size_t)k * 108;
    for (int s = 0; s < 27; ++s) {
      const size_t j = (size_t)m.idc[ic + nc * s] - 1;
      rs.M[MB][r + ld * j] += st[81 + s];
      rs.M[MDx][r + ld * j] += st[s];
      rs.M[MDy][r + ld * j] += st[27 + s];
      rs.M[MDz][r + ld * j] += st[54 + s];
    }
  }
}

// ---- free terms and linear-field corrections
// partial sums of one column slice: 0 sum H, 1 sum B, 2..4 sum H x_c(gbn),
// 5..7 sum B x_c, 8..10 sum D_c (c = z, x, y)
__global__ void __launch_bounds__(64) mhsd_sums_kernel(MeshPtrs m, RowSet rs, int full, double *__restrict__ part) {
  const int r = blockIdx.x * 64 + threadIdx.x;
  if (r >= rs.n) return;
  const size_t ld = rs.n, nb = m.nb, nn = m.nn;
  const int sp = blockIdx.y;
  const int crd[3] = {2, 0, 1}, mat[3] = {MDz, MDx, MDy};
  double v[kSums];
  for (int k = 0; k < kSums; ++k) v[k] = 0.0;
  const size_t b0 = nb * sp / kSplit, b1 = nb * (sp + 1) / kSplit;
  for (size_t j = b0; j < b1; ++j) {
    const double h = rs.M[MH][r + ld * j];
    v[0] += h;
    if (full) {
      const int g = m.gbn[j] - 1;
      for (int t = 0; t < 3; ++t) v[2 + t] += h * m.x[g + nn * crd[t]];
    }
  }
  const size_t n0 = nn * sp / kSplit, n1 = nn * (sp + 1) / kSplit;
  for (size_t j = n0; j < n1; ++j) {
    const double b = rs.M[MB][r + ld * j];
    v[1] += b;
    if (full)
      for (int t = 0; t < 3; ++t) {
        v[5 + t] += b * m.x[j + nn * crd[t]];
        v[8 + t] += rs.M[mat[t]][r + ld * j];
      }
  }
  double *o = part + ((size_t)sp * rs.n + r) * kSums;
  for (int k = 0; k < kSums; ++k) o[k] = v[k];
}

// U rows: free term c = -sum H + mu^2 sum B on the node's own H column,
// then D_c(it, it) += sum H x_c - mu^2 sum B x_c - sum D_c (+ x_c(it) inside);
// ont[r*4 + {0,1,2,3}] = |0.125 - c| (row 0 on the boundary), |vs| for z, x, y
__global__ void mhsd_apply_u_kernel(MeshPtrs m, RowSet rs, const double *__restrict__ part, double *__restrict__ ont) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rs.n) return;
  const size_t ld = rs.n, nn = m.nn;
  double v[kSums];
  for (int k = 0; k < kSums; ++k) v[k] = 0.0;
  for (int sp = 0; sp < kSplit; ++sp) {
    const double *pp = part + ((size_t)sp * rs.n + r) * kSums;
    for (int k = 0; k < kSums; ++k) v[k] += pp[k];
  }
  const double mu = rs.src[r].w;
  const double c = -v[0] + mu * mu * v[1];
  const int lb = m.lbn[r];
  double *o = ont + (size_t)r * 4;
  o[0] = 0.0;
  if (lb > 0) {
    rs.M[MH][r + ld * (size_t)(lb - 1)] += c;
    if (r == 0) o[0] = fabs(0.125 - c);
  }
  const int crd[3] = {2, 0, 1}, mat[3] = {MDz, MDx, MDy};
  for (int t = 0; t < 3; ++t) {
    const double xr = m.x[r + nn * crd[t]];
    const double wh = lb > 0 ? v[2 + t] + c * xr : v[2 + t];
    double vs = wh - mu * mu * v[5 + t] - v[8 + t];
    if (lb == 0) vs = vs + xr;
    o[1 + t] = fabs(vs);
    rs.M[mat[t]][r + ld * (size_t)r] += vs;
  }
}

// Q rows: c = -sum H + mu^2 sum B spread over the element's nine nodes by
// phi_9 at the flux point; ont[r] = |0.5 - c|
__global__ void mhsd_apply_q_kernel(MeshPtrs m, RowSet rs, const double *__restrict__ part, double *__restrict__ ont) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rs.n) return;
  const size_t ld = rs.n, ne = m.ne;
  double v0 = 0.0, v1 = 0.0;
  for (int sp = 0; sp < kSplit; ++sp) {
    const double *pp = part + ((size_t)sp * rs.n + r) * kSums;
    v0 += pp[0], v1 += pp[1];
  }
  const double mu = rs.src[r].w;
  const double c = -v0 + mu * mu * v1;
  ont[r] = fabs(0.5 - c);
  const int e = r / 4, f = r % 4 + 1;
  const double tc = 3.0 / 4.0;
  double fi[9];
  shape9((f == 1 || f == 4) ? -tc : tc, f <= 2 ? -tc : tc, fi);
  for (int s = 0; s < 9; ++s) {
    const size_t j = (size_t)(m.lbn[m.ibc[e + ne * s] - 1] - 1);
    rs.M[MH][r + ld * j] += c * fi[s];
  }
}

// MakeBprime: Bp = B beta ((mu_row / mu_col)^2 - 1)
__global__ void mhsd_bprime_kernel(RowSet rs, int nn, const double *__restrict__ mucol, double beta) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)rs.n * nn) return;
  const int r = (int)(t % rs.n), col = (int)(t / rs.n);
  const double q = rs.src[r].w / mucol[col];
  rs.M[MBp][t] = rs.M[MB][t] * beta * (q * q - 1.0);
}

// ---- fmmHSysM: thread (e, c): equation e, column c over the nn nodes then the
// nf boundary flux points
__global__ void mhsd_sysm_kernel(int nunk, int nn, int nf, const int2 *__restrict__ eqrow, RowSet U, RowSet Q,
                                 const int *__restrict__ kode, const int *__restrict__ qkode,
                                 const int *__restrict__ lbn, const int *__restrict__ ibcf4,
                                 const double *__restrict__ diff, const double *__restrict__ qdiff,
                                 double *__restrict__ A, double *__restrict__ RHS) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nunk * (nn + nf)) return;
  const int e = (int)(t % nunk), c = (int)(t / nunk);
  const int2 er = eqrow[e];
  const RowSet &rs = er.x == 0 ? U : Q;
  const size_t ld = rs.n, row = er.y;
  if (c < nn) {
    double val = 0.0;
    if (lbn[c] != 0) val = diff[c] * rs.M[MH][row + ld * (size_t)(lbn[c] - 1)];
    if (er.x == 0 && c == (int)row && lbn[c] == 0) val = val + diff[c];  // c(xi) = 1 inside
    const int nuq = kode[c];
    if (nuq < 0) A[e + (size_t)nunk * (-nuq - 1)] = val;
    else RHS[e + (size_t)nunk * (nuq - 1)] = -val;
  } else {
    const int fp = c - nn, j = ibcf4[fp];
    const double g = rs.M[MG][row + ld * (size_t)fp];
    const int nuq = qkode[fp];
    if (nuq < 0) A[e + (size_t)nunk * (-nuq - 1)] = -g * qdiff[j - 1];
    else RHS[e + (size_t)nunk * (nuq - 1)] = g * qdiff[j - 1];
  }
}

// ---- solve
// known values b and the product vectors v0..v7 (lib_fm.f:269-301)
__global__ void mhsd_vectors_kernel(int nn, int nb, int nf, const int *__restrict__ kode,
                                    const int *__restrict__ qkode, const int *__restrict__ gbn,
                                    const double *__restrict__ u, const double *__restrict__ qu,
                                    const double *__restrict__ vel, const double *__restrict__ rhsv,
                                    const double *__restrict__ dg, double *__restrict__ b, double *__restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nn) {
    if (kode[i] > 0) b[kode[i] - 1] = u[i];
    const double ui = u[i];
    v[i] = -rhsv[i];
    v[(size_t)nn + i] = ui;
    for (int d = 0; d < 3; ++d) v[(size_t)nn * (2 + d) + i] = ui * (vel[i + (size_t)nn * d] + dg[i + (size_t)nn * d]);
  }
  if (i < nf && qkode[i] > 0) b[qkode[i] - 1] = qu[i];
  if (i < nb) {
    const int node = gbn[i] - 1;
    for (int d = 0; d < 3; ++d) v[(size_t)nn * 5 + (size_t)nb * d + i] = -u[node] * vel[node + (size_t)nn * d];
  }
}

// Products of up to eight column-major matrices with their vectors, summed
// per row (rows[i] of each, or row i), over one slice of the concatenated
// columns: part[ks * nr + i].
struct MV8 {
  const double *M[8];
  const double *v[8];
  int ncol[8];
  int nmat;
};

__global__ void __launch_bounds__(kRows) mhsd_mv_kernel(int nr, const int *__restrict__ rows, long ld, MV8 mv,
                                                        int total, int ks_cnt, double *__restrict__ part) {
  const int i = blockIdx.x * kRows + threadIdx.x, ks = blockIdx.y;
  if (i >= nr) return;
  const size_t row = rows ? (size_t)rows[i] : (size_t)i;
  const int chunk = (total + ks_cnt - 1) / ks_cnt, j0 = ks * chunk, j1 = min(total, j0 + chunk);
  double acc = 0.0;
  int base = 0;
  for (int k = 0; k < mv.nmat; ++k) {
    const int lo = max(j0, base), hi = min(j1, base + mv.ncol[k]);
    const double *__restrict__ M = mv.M[k];
    const double *__restrict__ v = mv.v[k];
#pragma unroll 8
    for (int j = lo; j < hi; ++j) acc += M[row + ld * (size_t)(j - base)] * v[j - base];
    base += mv.ncol[k];
  }
  part[(size_t)ks * nr + i] = acc;
}

// out[i] = base[i] (if given) + sum of the slices in order
__global__ void mhsd_reduce_kernel(int nr, int ks_cnt, const double *__restrict__ part,
                                   const double *__restrict__ base, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  const double s = slice_sum(part + i, (size_t)nr, ks_cnt);
  out[i] = base ? base[i] + s : s;
}

// r[e] = rhsb[e] + y0[row] + ... + y7[row] in the reference's order
// (rB, rBp, rDx, rDy, rDz, rAx, rAy, rAz)
__global__ void mhsd_combine_kernel(int ne, const int *__restrict__ rows, int nrow, const double *__restrict__ y,
                                    const double *__restrict__ rhsb, double *__restrict__ r) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const size_t row = rows[e];
  double s = rhsb[e];
  for (int k = 0; k < 8; ++k) s = s + y[(size_t)nrow * k + row];
  r[e] = s;
}

// under-relaxed update of the unknown u and q
__global__ void mhsd_relax_kernel(int nn, int nf, const int *__restrict__ kode, const int *__restrict__ qkode,
                                  double ur, const double *__restrict__ x, double *__restrict__ u,
                                  double *__restrict__ qu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nn && kode[i] < 0) u[i] = (1.0 - ur) * u[i] + ur * x[-kode[i] - 1];
  if (i < nf && qkode[i] < 0) qu[i] = (1.0 - ur) * qu[i] + ur * x[-qkode[i] - 1];
}

// column slices for about 1,024 wavefronts over the chip: with eight columns'
// loads in flight per wavefront that is the ~4 MB in flight these streams run
// fastest at (8^3 solve 0.35 ms at 1,024 wavefronts, 0.48 ms at 8,192)
int split_cols(int nr, int ncols) {
  const int rows = (nr + kRows - 1) / kRows;
  int ks = (1024 + rows - 1) / rows;
  return std::max(1, std::min(ks, std::max(1, ncols / 64)));
}

// greedy colouring of items (elements or cells) so that no two of a colour
// share a node; returns the items grouped by colour and the group starts
void colour_items(int nitems, int per, const int *ids, size_t stride, int nn, std::vector<int> &list,
                  std::vector<int> &start) {
  std::vector<uint64_t> used(nn, 0);
  std::vector<int> colour(nitems);
  int ncol = 0;
  for (int it = 0; it < nitems; ++it) {
    uint64_t mask = 0;
    for (int k = 0; k < per; ++k) mask |= used[ids[it + stride * k] - 1];
    if (~mask == 0) throw std::runtime_error("colouring needs more than 64 colours");
    const int col = __builtin_ctzll(~mask);
    for (int k = 0; k < per; ++k) used[ids[it + stride * k] - 1] |= 1ull << col;
    colour[it] = col, ncol = std::max(ncol, col + 1);
  }
  start.assign(ncol + 1, 0);
  for (int it = 0; it < nitems; ++it) ++start[colour[it] + 1];
  for (int k = 0; k < ncol; ++k) start[k + 1] += start[k];
  list.resize(nitems);
  std::vector<int> fill(start.begin(), start.end() - 1);
  for (int it = 0; it < nitems; ++it) list[fill[colour[it]]++] = it;
}

}  // namespace

// ---------------------------------------------------------------------------
// The operator: matrices, system and compressed forms of one mesh.
struct MhsdComp {  // compressed form of one right-hand-side matrix
  AcaDev aca;
  WtDev wt;
  tt_hmat *hm = nullptr;
  MhsdComp() = default;
  MhsdComp(const MhsdComp &) = delete;
  MhsdComp &operator=(const MhsdComp &) = delete;
  ~MhsdComp() {
    if (hm) tt_hmat_free(hm);
  }
};

}  // namespace tt

struct tt_mhsd {
  hipStream_t s = nullptr;
  tt::DevMesh m;
  tt::QuadDev qh;
  tt::DBuf<tt::QuadDev> dq;
  int nn = 0, nb = 0, ne = 0, nc = 0, nq = 0, nf = 0;
  double beta = 0.0, mumax = 0.0;
  std::vector<double> hx, hxq;  // host coordinates (block-matrix partitions)
  std::vector<int> hgbn, hibcf;
  tt::DBuf<double> diff, qdiff, mucol;
  tt::DBuf<double4> srcU, srcQ;
  tt::DBuf<int> unode, qnode, ibcf4;
  tt::DBuf<double> U[tt::kMats], Q[tt::kMats];
  // system
  int nunk = 0, nknown = 0, nequ = 0, neqq = 0;
  tt::DBuf<int> kode, qkode, rowsU, rowsQ;
  tt::DBuf<double> Ainv, RHS;
  // compressed right-hand-side matrices (mode 1..3), [0..7] U then [8..15] Q
  int mode = 0;
  std::vector<std::unique_ptr<tt::MhsdComp>> comp;
  double ratio = 1.0;
  // solve work
  tt::DBuf<double> du, dqu, dvel, drhsv, ddg, b, v, rhsb, r, x, y, part;
  float ms_integrals = 0.f, ms_system = 0.f, ms_compress = 0.f, ms_solve = 0.f;
  ~tt_mhsd() {
    comp.clear();
    if (s) (void)hipStreamDestroy(s);
  }
  tt::RowSet rowset(int set) {
    tt::RowSet r{};
    r.n = set == 0 ? nn : nf;
    r.src = set == 0 ? srcU.p : srcQ.p;
    r.unode = set == 0 ? unode.p : nullptr;
    r.qnode = set == 0 ? nullptr : qnode.p;
    for (int k = 0; k < tt::kMats; ++k) r.M[k] = set == 0 ? U[k].p : Q[k].p;
    return r;
  }
  size_t cols(int k) const {
    switch (k) {
      case tt::MH: case tt::MAx: case tt::MAy: case tt::MAz: return nb;
      case tt::MG: return nf;
      default: return nn;
    }
  }
};

namespace tt {
namespace {

struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t s;
  explicit Timer(hipStream_t st) : s(st) {
    TT_HIP(hipEventCreate(&a));
    TT_HIP(hipEventCreate(&b));
    TT_HIP(hipEventRecord(a, s));
  }
  float stop() {
    TT_HIP(hipEventRecord(b, s));
    TT_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    TT_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
  }
  ~Timer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

void integrals(tt_mhsd &h, const tt_mesh_view &v, double ont[5]) {
  hipStream_t s = h.s;
  const int nn = h.nn, nb = h.nb, ne = h.ne, nc = h.nc, nf = h.nf;
  for (int k = 0; k < kMats; ++k) {
    h.U[k].alloc((size_t)nn * h.cols(k)), h.U[k].zero(s);
    h.Q[k].alloc((size_t)nf * h.cols(k)), h.Q[k].zero(s);
  }
  RowSet sets[2] = {h.rowset(0), h.rowset(1)};
  // ---- faces
  {
    std::vector<double> me(ne);
    const int kc[5] = {1, 3, 5, 7, 1};
    for (int e = 0; e < ne; ++e) {  // FindMinEdgeOneElem
      double mn = 1.0e20;
      for (int k = 0; k < 4; ++k) {
        const int a = v.ibc[e + (size_t)ne * (kc[k] - 1)] - 1, b = v.ibc[e + (size_t)ne * (kc[k + 1] - 1)] - 1;
        const double dx = v.x[a] - v.x[b], dy = v.x[a + nn] - v.x[b + nn],
                     dz = v.x[a + 2 * (size_t)nn] - v.x[b + 2 * (size_t)nn];
        const double l = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (l < mn) mn = l;
      }
      me[e] = mn;
    }
    std::vector<int> list, start;
    colour_items(ne, 9, v.ibc, ne, nn, list, start);
    DBuf<double> dme;
    DBuf<int> dl;
    dme.upload(me, s), dl.upload(list, s);
    for (int set = 0; set < 2; ++set)
      for (size_t k = 0; k + 1 < start.size(); ++k) {
        const int cnt = start[k + 1] - start[k];
        if (!cnt) continue;
        hipLaunchKernelGGL(mhsd_bnd_kernel, dim3(ceil_div((size_t)sets[set].n * cnt, 64)), dim3(64), 0, s, ptrs(h.m),
                           h.dq.p, sets[set], dl.p + start[k], cnt, h.ibcf4.p, dme.p);
        TT_LAUNCH_CHECK();
      }
    TT_HIP(hipStreamSynchronize(s));
  }
  // ---- singular pairs: (row, cell) with the source on the cell
  std::vector<std::vector<std::pair<int, int>>> ncell(nn), fcell(h.nq);  // (cell, isrc)
  for (int ic = 0; ic < nc; ++ic) {
    for (int k = 0; k < 27; ++k) ncell[v.idc[ic + (size_t)nc * k] - 1].push_back({ic, k + 1});
    for (int fc = 0; fc < 6; ++fc)
      for (int k = 0; k < 4; ++k) fcell[v.ibf[ic + (size_t)nc * (fc + 6 * k)] - 1].push_back({ic, 27 + fc * 4 + k + 1});
  }
  for (int set = 0; set < 2; ++set) {
    const int nr = sets[set].n;
    std::vector<int> prow, pcell, psrc, off(nr + 1, 0), sing((size_t)nr * kMaxSing, -1);
    for (int r = 0; r < nr; ++r) {
      const auto &lst = set == 0 ? ncell[r] : fcell[h.hibcf[r] - 1];
      if ((int)lst.size() > kMaxSing) throw std::runtime_error("more than 16 cells around a source point");
      for (size_t k = 0; k < lst.size(); ++k) {
        prow.push_back(r), pcell.push_back(lst[k].first), psrc.push_back(lst[k].second);
        sing[(size_t)r * kMaxSing + k] = lst[k].first;
      }
      off[r + 1] = (int)prow.size();
    }
    // regular cells on the matrix cores, one cell colour per launch
    const int n1 = h.qh.ndr, Pp = (n1 * n1 * n1 + 15) / 16 * 16;
    std::vector<int> clist, cstart;
    colour_items(nc, 27, v.idc, nc, nn, clist, cstart);
    DBuf<double> geo((size_t)nc * Pp * 4), phi((size_t)Pp * 32);
    DBuf<int> dcl, dsing;
    dcl.upload(clist, s), dsing.upload(sing, s);
    hipLaunchKernelGGL(mhsd_geom_kernel, dim3(ceil_div((size_t)nc * Pp, 256)), dim3(256), 0, s, ptrs(h.m), h.dq.p, Pp,
                       geo.p, phi.p);
    TT_LAUNCH_CHECK();
    const int bx = ceil_div(nr, 64);
    for (size_t k = 0; k + 1 < cstart.size(); ++k) {
      const int cnt = cstart[k + 1] - cstart[k];
      if (!cnt) continue;
      const int chunk = std::max(1, (int)(((int64_t)cnt * bx + 4095) / 4096));
      const dim3 grid(bx, ceil_div(cnt, chunk));
      auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, ptrs(h.m), sets[set], dsing.p, Pp, geo.p, phi.p,
                           dcl.p + cstart[k], cnt, chunk);
      };
      switch (Pp / 16) {
        case 1: launch(mhsd_dom_mfma_kernel<1>); break;
        case 2: launch(mhsd_dom_mfma_kernel<2>); break;
        case 4: launch(mhsd_dom_mfma_kernel<4>); break;
        default: launch(mhsd_dom_mfma_kernel<0>); break;
      }
      TT_LAUNCH_CHECK();
    }
    const int np = (int)prow.size();
    if (np) {
      DBuf<int> dprow, dpcell, dpsrc, doff;
      dprow.upload(prow, s), dpcell.upload(pcell, s), dpsrc.upload(psrc, s), doff.upload(off, s);
      DBuf<double> staging((size_t)np * 108);
      hipLaunchKernelGGL(mhsd_dom_singular_kernel, dim3(np), dim3(64), 0, s, ptrs(h.m), h.dq.p, sets[set], dprow.p,
                         dpcell.p, dpsrc.p, np, staging.p);
      TT_LAUNCH_CHECK();
      hipLaunchKernelGGL(mhsd_singular_apply, dim3(ceil_div(nr, 64)), dim3(64), 0, s, ptrs(h.m), sets[set], doff.p,
                         dpcell.p, staging.p);
      TT_LAUNCH_CHECK();
      TT_HIP(hipStreamSynchronize(s));
    }
    TT_HIP(hipStreamSynchronize(s));
  }
  // ---- free terms and corrections, then B'
  double o[5] = {0, 0, 0, 0, 0};
  for (int set = 0; set < 2; ++set) {
    const int nr = sets[set].n;
    DBuf<double> part((size_t)kSplit * nr * kSums), dont((size_t)nr * 4);
    hipLaunchKernelGGL(mhsd_sums_kernel, dim3(ceil_div(nr, 64), kSplit), dim3(64), 0, s, ptrs(h.m), sets[set],
                       set == 0 ? 1 : 0, part.p);
    TT_LAUNCH_CHECK();
    if (set == 0)
      hipLaunchKernelGGL(mhsd_apply_u_kernel, dim3(ceil_div(nr, 64)), dim3(64), 0, s, ptrs(h.m), sets[set], part.p,
                         dont.p);
    else
      hipLaunchKernelGGL(mhsd_apply_q_kernel, dim3(ceil_div(nr, 64)), dim3(64), 0, s, ptrs(h.m), sets[set], part.p,
                         dont.p);
    TT_LAUNCH_CHECK();
    hipLaunchKernelGGL(mhsd_bprime_kernel, dim3(ceil_div((size_t)nr * nn, 256)), dim3(256), 0, s, sets[set], nn,
                       h.mucol.p, h.beta);
    TT_LAUNCH_CHECK();
    std::vector<double> ho((size_t)nr * 4);
    dont.download(ho.data(), set == 0 ? (size_t)nr * 4 : (size_t)nr, s);
    TT_HIP(hipStreamSynchronize(s));
    if (set == 0) {
      for (int r = 0; r < nr; ++r) {
        o[0] = std::max(o[0], ho[(size_t)r * 4]);
        o[4] = std::max(o[4], ho[(size_t)r * 4 + 1]);  // z
        o[2] = std::max(o[2], ho[(size_t)r * 4 + 2]);  // x
        o[3] = std::max(o[3], ho[(size_t)r * 4 + 3]);  // y
      }
    } else {
      for (int r = 0; r < nr; ++r) o[1] = std::max(o[1], ho[r]);
    }
  }
  std::memcpy(ont, o, sizeof(o));
  (void)nb;
}

void gemv_rows(tt_mhsd &h, int nr, const int *rows, long ld, const MV8 &mv, const double *base, double *out) {
  int total = 0;
  for (int k = 0; k < mv.nmat; ++k) total += mv.ncol[k];
  const int ks = split_cols(nr, total);
  if (h.part.n < (size_t)ks * nr) throw std::runtime_error("product partials exceed their buffer");
  hipLaunchKernelGGL(mhsd_mv_kernel, dim3(ceil_div(nr, kRows), ks), dim3(kRows), 0, h.s, nr, rows, ld, mv, total, ks,
                     h.part.p);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(mhsd_reduce_kernel, dim3(ceil_div(nr, 256)), dim3(256), 0, h.s, nr, ks, h.part.p, base, out);
  TT_LAUNCH_CHECK();
}

MV8 one(const double *M, int ncol, const double *v) {
  MV8 mv{};
  mv.M[0] = M, mv.v[0] = v, mv.ncol[0] = ncol, mv.nmat = 1;
  return mv;
}

// the eight right-hand-side matrices of a row set and the offset of their vectors in h.v
void rhs_mats(tt_mhsd &h, int set, const double *M[8], int ncol[8], const double *vec[8]) {
  const int ids[8] = {MB, MBp, MDx, MDy, MDz, MAx, MAy, MAz};
  for (int k = 0; k < 8; ++k) {
    M[k] = set == 0 ? h.U[ids[k]].p : h.Q[ids[k]].p;
    ncol[k] = k < 5 ? h.nn : h.nb;
    vec[k] = k < 5 ? h.v.p + (size_t)h.nn * k : h.v.p + (size_t)h.nn * 5 + (size_t)h.nb * (k - 5);
  }
}

}  // namespace
}  // namespace tt

using namespace tt;

namespace {
// mesh, sources and mu of the rows (the reference's expressions, evaluated once per row)
std::unique_ptr<tt_mhsd> mhsd_setup(const tt_mesh_view *mv, const int32_t *ibcf, const double *diff,
                                    const double *qdiff, double beta, double *mumax) {
  if (!(beta > 0.0)) throw std::runtime_error("beta must be positive");
  auto h = std::make_unique<tt_mhsd>();
  TT_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
  hipStream_t s = h->s;
  h->nn = mv->nnodes, h->nb = mv->nbnodes, h->ne = mv->nbelem, h->nc = mv->nicell, h->nq = mv->nq;
  h->nf = 4 * h->ne;
  h->beta = beta;
  upload_mesh(*mv, h->m, s);
  {
    const int nn = h->nn, nf = h->nf;
    h->hx.assign(mv->x, mv->x + (size_t)nn * 3);
    h->hxq.assign(mv->xq, mv->xq + (size_t)h->nq * 3);
    h->hgbn.assign(mv->gbn, mv->gbn + h->nb);
    h->hibcf.assign(ibcf, ibcf + nf);
    for (int r = 0; r < nf; ++r)
      if (ibcf[r] < 1 || ibcf[r] > h->nq) throw std::runtime_error("flux-node id outside 1..nq");
    // sources and mu (the reference's expressions, evaluated once per row)
    std::vector<double4> su(nn), sq(nf);
    std::vector<double> mc(nn);
    std::vector<int> un(nn);
    double mm = 0.0;
    for (int i = 0; i < nn; ++i) {
      if (!(diff[i] > 0.0)) throw std::runtime_error("diffusivity must be positive");
      const double mu = std::sqrt(beta / diff[i]);
      su[i] = make_double4(mv->x[i], mv->x[i + nn], mv->x[i + 2 * (size_t)nn], mu);
      mc[i] = mu, un[i] = i + 1, mm = std::max(mm, mu);
    }
    for (int r = 0; r < nf; ++r) {
      const int it = ibcf[r] - 1;
      if (!(qdiff[it] > 0.0)) throw std::runtime_error("flux-node diffusivity must be positive");
      const double mu = std::sqrt(beta / qdiff[it]);
      sq[r] = make_double4(mv->xq[it], mv->xq[it + h->nq], mv->xq[it + 2 * (size_t)h->nq], mu);
      mm = std::max(mm, mu);
    }
    h->mumax = mm;
    if (mumax) *mumax = mm;
    h->srcU.upload(su, s), h->srcQ.upload(sq, s), h->mucol.upload(mc, s);
    h->unode.upload(un, s);
    h->qnode.upload(ibcf, nf, s);
    h->ibcf4.upload(ibcf, nf, s);
    h->diff.upload(diff, nn, s), h->qdiff.upload(qdiff, h->nq, s);
    TT_HIP(hipStreamSynchronize(s));  // the host staging vectors end here
  }
  return h;
}
}  // namespace

extern "C" int tt_mhsd_create(const tt_mesh_view *mv, const tt_quad *q, const int32_t *ibcf, const double *diff,
                              const double *qdiff, double beta, tt_mhsd **out, double ont[5], double *mumax) {
  return guarded("tt_mhsd_create", [&] {
    if (!mv || !q || !ibcf || !diff || !qdiff || !out || !ont) throw std::runtime_error("NULL argument");
    auto h = mhsd_setup(mv, ibcf, diff, qdiff, beta, mumax);
    build_quad(*q, h->qh);
    h->dq.upload(&h->qh, 1, h->s);
    Timer t(h->s);
    integrals(*h, *mv, ont);
    h->ms_integrals = t.stop();
    *out = h.release();
    return 0;
  });
}

extern "C" int tt_mhsd_create_from(const tt_mesh_view *mv, const int32_t *ibcf, const double *diff,
                                   const double *qdiff, double beta, const double *const U[10],
                                   const double *const Q[10], tt_mhsd **out) {
  return guarded("tt_mhsd_create_from", [&] {
    if (!mv || !ibcf || !diff || !qdiff || !U || !Q || !out) throw std::runtime_error("NULL argument");
    auto h = mhsd_setup(mv, ibcf, diff, qdiff, beta, nullptr);
    for (int k = 0; k < kMats; ++k) {
      if (!U[k] || !Q[k]) throw std::runtime_error("NULL matrix");
      h->U[k].upload(U[k], (size_t)h->nn * h->cols(k), h->s);
      h->Q[k].upload(Q[k], (size_t)h->nf * h->cols(k), h->s);
    }
    TT_HIP(hipStreamSynchronize(h->s));
    *out = h.release();
    return 0;
  });
}

extern "C" int tt_mhsd_matrix(tt_mhsd *h, int32_t set, int32_t k, double *dst) {
  return guarded("tt_mhsd_matrix", [&] {
    if (!h || !dst || set < 0 || set > 1 || k < 0 || k >= kMats) throw std::runtime_error("bad argument");
    DBuf<double> &b = set == 0 ? h->U[k] : h->Q[k];
    b.download(dst, b.n, h->s);
    TT_HIP(hipStreamSynchronize(h->s));
    return 0;
  });
}

extern "C" int tt_mhsd_system(tt_mhsd *h, const int32_t *kode, const int32_t *qkode, int32_t nunk, int32_t nknown) {
  return guarded("tt_mhsd_system", [&] {
    if (!h || !kode || !qkode) throw std::runtime_error("NULL argument");
    const int nn = h->nn, nf = h->nf;
    // the codes must number the unknowns 1..nunk and the known values 1..nknown once each
    std::vector<int> seen_u(std::max(nunk, 1), 0), seen_k(std::max(nknown, 1), 0);
    std::vector<int> ru, rq;
    std::vector<int2> eq;
    auto take = [&](int c) {
      if (c < 0) {
        if (-c > nunk || seen_u[-c - 1]++) throw std::runtime_error("unknown numbering is not 1..nunk once each");
      } else if (c > 0) {
        if (c > nknown || seen_k[c - 1]++) throw std::runtime_error("known numbering is not 1..nknown once each");
      } else {
        throw std::runtime_error("zero code");
      }
    };
    for (int i = 0; i < nn; ++i) {
      take(kode[i]);
      if (kode[i] < 0) {
        if (-kode[i] - 1 != (int)eq.size()) throw std::runtime_error("unknown nodes must be numbered in node order");
        ru.push_back(i), eq.push_back(make_int2(0, i));
      }
    }
    for (int i = 0; i < nf; ++i) {
      take(qkode[i]);
      if (qkode[i] < 0) {
        if (-qkode[i] - 1 != (int)eq.size()) throw std::runtime_error("unknown flux points must follow the nodes");
        rq.push_back(i), eq.push_back(make_int2(1, i));
      }
    }
    if ((int)eq.size() != nunk) throw std::runtime_error("nunk does not match the codes");
    hipStream_t s = h->s;
    h->nunk = nunk, h->nknown = nknown, h->nequ = (int)ru.size(), h->neqq = (int)rq.size();
    h->kode.upload(kode, nn, s), h->qkode.upload(qkode, nf, s);
    h->rowsU.upload(ru.empty() ? std::vector<int>{0} : ru, s);
    h->rowsQ.upload(rq.empty() ? std::vector<int>{0} : rq, s);
    DBuf<int2> deq;
    deq.upload(eq, s);
    h->Ainv.alloc((size_t)nunk * nunk);
    h->RHS.alloc((size_t)nunk * std::max(nknown, 1));
    Timer t(s);
    hipLaunchKernelGGL(mhsd_sysm_kernel, dim3(ceil_div((size_t)nunk * (nn + nf), 256)), dim3(256), 0, s, nunk, nn, nf,
                       deq.p, h->rowset(0), h->rowset(1), h->kode.p, h->qkode.p, h->m.lbn.p, h->ibcf4.p, h->diff.p,
                       h->qdiff.p, h->Ainv.p, h->RHS.p);
    TT_LAUNCH_CHECK();
    rocblas_handle bh;
    if (rocblas_create_handle(&bh) != rocblas_status_success) throw std::runtime_error("rocblas_create_handle failed");
    rocblas_set_stream(bh, s);
    DBuf<int> ipiv(nunk), info(1);
    const bool ok = rocsolver_dgetrf(bh, nunk, nunk, h->Ainv.p, nunk, ipiv.p, info.p) == rocblas_status_success;
    int inf = 0;
    if (ok) info.download(&inf, 1, s);
    TT_HIP(hipStreamSynchronize(s));
    bool ok2 = false;
    if (ok && inf == 0) {
      ok2 = rocsolver_dgetri(bh, nunk, h->Ainv.p, nunk, ipiv.p, info.p) == rocblas_status_success;
      TT_HIP(hipStreamSynchronize(s));
    }
    rocblas_destroy_handle(bh);
    if (!ok || !ok2) throw std::runtime_error("rocSOLVER factorisation failed");
    if (inf != 0) throw std::runtime_error("system matrix is singular");
    h->ms_system = t.stop();
    // solve work
    h->du.alloc(nn), h->dqu.alloc(nf), h->dvel.alloc((size_t)nn * 3), h->drhsv.alloc(nn), h->ddg.alloc((size_t)nn * 3);
    h->b.alloc(std::max(nknown, 1)), h->v.alloc((size_t)nn * 5 + (size_t)h->nb * 3);
    h->rhsb.alloc(nunk), h->r.alloc(nunk), h->x.alloc(nunk), h->y.alloc((size_t)8 * std::max(nn, nf));
    // column-slice partials of the products (split_cols keeps slices x rows below this)
    h->part.alloc((size_t)kRows * (8192 + ceil_div(std::max({nunk, nn, nf}), kRows)));
    return 0;
  });
}

extern "C" int tt_mhsd_compress(tt_mhsd *h, int32_t mode, int32_t aca_type, double aca_eps, double wt_kappa,
                                int32_t ct_type, double ct_eta, int32_t leaf, double *ratio) {
  return guarded("tt_mhsd_compress", [&] {
    if (!h) throw std::runtime_error("NULL argument");
    if (mode < 0 || mode > 3) throw std::runtime_error("mode must be 0 (dense), 1 (wavelet), 2 (ACA) or 3 (blocks)");
    hipStream_t s = h->s;
    h->comp.clear();
    h->mode = mode;
    h->ratio = 1.0;
    if (mode == 0) {
      if (ratio) *ratio = 1.0;
      return 0;
    }
    Timer t(s);
    const int ids[8] = {MB, MBp, MDx, MDy, MDz, MAx, MAy, MAz};
    double stored = 0.0, full = 0.0;
    std::mt19937_64 rng(20251203);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    // row points: nodes (U) and flux points (Q); column points: nodes or boundary nodes
    std::vector<double> pu(h->hx), pq((size_t)h->nf * 3), pb((size_t)h->nb * 3);
    for (int r = 0; r < h->nf; ++r)
      for (int d = 0; d < 3; ++d) pq[r + (size_t)h->nf * d] = h->hxq[h->hibcf[r] - 1 + (size_t)h->nq * d];
    for (int j = 0; j < h->nb; ++j)
      for (int d = 0; d < 3; ++d) pb[j + (size_t)h->nb * d] = h->hx[h->hgbn[j] - 1 + (size_t)h->nn * d];
    for (int set = 0; set < 2; ++set)
      for (int k = 0; k < 8; ++k) {
        const int m = set == 0 ? h->nn : h->nf, n = k < 5 ? h->nn : h->nb;
        const double *A = set == 0 ? h->U[ids[k]].p : h->Q[ids[k]].p;
        auto c = std::make_unique<MhsdComp>();
        full += (double)m * n;
        if (mode == 1) {
          wt_compress_dev(A, m, m, n, wt_kappa, c->wt, s);
          stored += (double)c->wt.nnz;
        } else if (mode == 2) {
          std::vector<double> vec(n);
          for (auto &e : vec) e = uni(rng);
          DBuf<double> R((size_t)m * n);
          TT_HIP(hipMemcpyAsync(R.p, A, sizeof(double) * (size_t)m * n, hipMemcpyDeviceToDevice, s));
          double nrm = 0.0;
          aca_compress_dev(R.p, m, n, aca_type, aca_eps, vec.data(), c->aca, &nrm, s);
          stored += (double)m * c->aca.rank + (double)c->aca.rank * n;
        } else {
          TT_HIP(hipStreamSynchronize(s));
          if (ct_type < 1 || ct_type > 3) throw std::runtime_error("block type must be 1 (dense), 2 or 3");
          c->hm = hmat_from_device(m, n, A, m, set == 0 ? pu.data() : pq.data(), k < 5 ? pu.data() : pb.data(), leaf,
                                   ct_eta, ct_type == 1 ? 1 : ct_type == 2 ? 2 : 3, aca_type, aca_eps, wt_kappa);
          stored += (double)hmat_stored_entries(c->hm);
        }
        h->comp.push_back(std::move(c));
      }
    TT_HIP(hipStreamSynchronize(s));
    h->ms_compress = t.stop();
    h->ratio = stored / full;
    if (ratio) *ratio = h->ratio;
    return 0;
  });
}

extern "C" int tt_mhsd_solve(tt_mhsd *h, double *u, double *qu, const double *velocity, const double *rhsv,
                             const double *dgrad, double urmh) {
  return guarded("tt_mhsd_solve", [&] {
    if (!h || !u || !qu || !velocity || !rhsv || !dgrad) throw std::runtime_error("NULL argument");
    if (!h->nunk) throw std::runtime_error("tt_mhsd_system has not been called");
    hipStream_t s = h->s;
    const int nn = h->nn, nb = h->nb, nf = h->nf, nunk = h->nunk;
    h->du.upload(u, nn, s), h->dqu.upload(qu, nf, s);
    h->dvel.upload(velocity, (size_t)nn * 3, s), h->drhsv.upload(rhsv, nn, s), h->ddg.upload(dgrad, (size_t)nn * 3, s);
    Timer t(s);
    hipLaunchKernelGGL(mhsd_vectors_kernel, dim3(ceil_div(std::max(nn, nf), 256)), dim3(256), 0, s, nn, nb, nf,
                       h->kode.p, h->qkode.p, h->m.gbn.p, h->du.p, h->dqu.p, h->dvel.p, h->drhsv.p, h->ddg.p, h->b.p,
                       h->v.p);
    TT_LAUNCH_CHECK();
    // RHS b
    if (h->nknown > 0)
      gemv_rows(*h, nunk, nullptr, nunk, one(h->RHS.p, h->nknown, h->b.p), nullptr, h->rhsb.p);
    else
      h->rhsb.zero(s);
    if (h->mode == 0) {
      // one fused pass over the eight matrices of the equation rows, added to RHS b
      for (int set = 0; set < 2; ++set) {
        const int ne = set == 0 ? h->nequ : h->neqq;
        if (!ne) continue;
        MV8 mv{};
        rhs_mats(*h, set, mv.M, mv.ncol, mv.v);
        mv.nmat = 8;
        const int e0 = set == 0 ? 0 : h->nequ;
        gemv_rows(*h, ne, set == 0 ? h->rowsU.p : h->rowsQ.p, set == 0 ? nn : nf, mv, h->rhsb.p + e0, h->r.p + e0);
      }
    } else {
      // the products of the compressed matrices, then the per-row sums in the reference's order
      for (int set = 0; set < 2; ++set) {
        const int nr = set == 0 ? nn : nf, ne = set == 0 ? h->nequ : h->neqq;
        const double *M[8], *vec[8];
        int ncol[8];
        rhs_mats(*h, set, M, ncol, vec);
        for (int k = 0; k < 8; ++k) {
          MhsdComp &c = *h->comp[(size_t)set * 8 + k];
          double *yk = h->y.p + (size_t)nr * k;
          if (h->mode == 1) wt_mv_dev(c.wt, vec[k], yk, s);
          else if (h->mode == 2) aca_mv_dev(c.aca, vec[k], yk, s);
          else hmat_apply(c.hm, vec[k], yk, s);
        }
        if (!ne) continue;
        const int e0 = set == 0 ? 0 : h->nequ;
        hipLaunchKernelGGL(mhsd_combine_kernel, dim3(ceil_div(ne, 256)), dim3(256), 0, s, ne,
                           set == 0 ? h->rowsU.p : h->rowsQ.p, nr, h->y.p, h->rhsb.p + e0, h->r.p + e0);
        TT_LAUNCH_CHECK();
      }
    }
    // x = A^-1 r, then the relaxed update
    gemv_rows(*h, nunk, nullptr, nunk, one(h->Ainv.p, nunk, h->r.p), nullptr, h->x.p);
    hipLaunchKernelGGL(mhsd_relax_kernel, dim3(ceil_div(std::max(nn, nf), 256)), dim3(256), 0, s, nn, nf, h->kode.p,
                       h->qkode.p, urmh, h->x.p, h->du.p, h->dqu.p);
    TT_LAUNCH_CHECK();
    h->ms_solve = t.stop();
    h->du.download(u, nn, s), h->dqu.download(qu, nf, s);
    TT_HIP(hipStreamSynchronize(s));
    return 0;
  });
}

extern "C" int tt_mhsd_stats(const tt_mhsd *h, double ms[4], double *ratio, int32_t *nunk) {
  return guarded("tt_mhsd_stats", [&] {
    if (!h) throw std::runtime_error("NULL argument");
    if (ms) ms[0] = h->ms_integrals, ms[1] = h->ms_system, ms[2] = h->ms_compress, ms[3] = h->ms_solve;
    if (ratio) *ratio = h->ratio;
    if (nunk) *nunk = h->nunk;
    return 0;
  });
}

extern "C" void tt_mhsd_free(tt_mhsd *h) {
  if (!h) return;
  if (h->s) (void)hipStreamSynchronize(h->s);
  delete h;
}
