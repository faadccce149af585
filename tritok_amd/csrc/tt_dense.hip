// tt_dense.hip -- dense single-domain kinematics on gfx950.
//
// Setup: A_e(i,j) = n(i) . D(i, gbn(j)) on rows with kode 0/3, identity rows
// otherwise (sdKm_sysm, lib_sdkm.f:1171-1210); LU with partial pivoting
// (rocSOLVER getrf, the same first-maximum pivot rule as LINPACK dgefa,
// lib_Solve_v11.f:787-904) and the explicit inverse (getri), so that every
// solve of the Picard loop is one HBM-streaming matrix-vector product.
//
// Per Picard iteration (sdkm_solve x, y, z; lib_sdkm.f:23-121,349-519): the
// reference reads all seven matrices once per equation.  Here they are read
// once per iteration: [H Htx Hty Htz] against the three boundary velocity
// components (12 sums per row) and [Dx Dy Dz] against the three vorticity
// components split into boundary and domain columns (18 sums per row).  Each
// equation's right-hand side is a per-row combination of these sums.  The
// Gauss-Seidel order of the reference (y uses the new boundary w_x, z the new
// w_x and w_y) is kept by correcting the boundary sums with D * dw after each
// solve.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <vector>

#include "tt_dense.h"

namespace tt {

namespace {

constexpr int kRowsPerBlock = 64;

__global__ void build_sys_kernel(int nb, int e, const int *__restrict__ kode, const double *__restrict__ nx,
                                 const double *__restrict__ ny, const double *__restrict__ nz,
                                 const int *__restrict__ gbn, const double *__restrict__ Dx,
                                 const double *__restrict__ Dy, const double *__restrict__ Dz, double *__restrict__ A) {
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nb * nb) return;
  int i = (int)(idx % nb), j = (int)(idx / nb);
  int kd = kode[i + (size_t)nb * e];
  double v;
  if (kd == 0 || kd == 3) {
    size_t dn = gbn[j] - 1;
    v = nx[i] * Dx[i + (size_t)nb * dn] + ny[i] * Dy[i + (size_t)nb * dn] + nz[i] * Dz[i + (size_t)nb * dn];
  } else {
    v = i == j ? 1.0 : 0.0;
  }
  A[idx] = v;
}

// [H Htx Hty Htz](i, j) * vel(gbn(j), c): 12 sums per row, split over column
// chunks (blockIdx.y), partials part[(ks*12 + o)*nb + i]
__global__ void __launch_bounds__(kRowsPerBlock) prod_h_kernel(int nb, int nn, int ks_cnt, const double *__restrict__ H,
                                                               const double *__restrict__ Htx,
                                                               const double *__restrict__ Hty,
                                                               const double *__restrict__ Htz,
                                                               const double *__restrict__ vel,
                                                               const int *__restrict__ gbn, double *__restrict__ part) {
  const int i = blockIdx.x * kRowsPerBlock + threadIdx.x, ks = blockIdx.y;
  const int chunk = (nb + ks_cnt - 1) / ks_cnt, j0 = ks * chunk, j1 = min(nb, j0 + chunk);
  double acc[12];
#pragma unroll
  for (int o = 0; o < 12; ++o) acc[o] = 0.0;
  if (i < nb) {
    for (int j = j0; j < j1; ++j) {
      const size_t g = gbn[j] - 1;
      const double v0 = vel[g], v1 = vel[g + nn], v2 = vel[g + 2 * (size_t)nn];
      const size_t a = i + (size_t)nb * j;
      const double m[4] = {H[a], Htx[a], Hty[a], Htz[a]};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        acc[q * 3 + 0] += m[q] * v0;
        acc[q * 3 + 1] += m[q] * v1;
        acc[q * 3 + 2] += m[q] * v2;
      }
    }
#pragma unroll
    for (int o = 0; o < 12; ++o) part[((size_t)ks * 12 + o) * nb + i] = acc[o];
  }
}

// [Dx Dy Dz](i, k) * vort(k, c), k split into boundary (0) / domain (1):
// 18 sums per row, output index m*6 + c*2 + dom
__global__ void __launch_bounds__(kRowsPerBlock) prod_d_kernel(int nb, int nn, int ks_cnt, const double *__restrict__ Dx,
                                                               const double *__restrict__ Dy,
                                                               const double *__restrict__ Dz,
                                                               const double *__restrict__ vort,
                                                               const int *__restrict__ isbnd, double *__restrict__ part) {
  const int i = blockIdx.x * kRowsPerBlock + threadIdx.x, ks = blockIdx.y;
  const int chunk = (nn + ks_cnt - 1) / ks_cnt, k0 = ks * chunk, k1 = min(nn, k0 + chunk);
  double acc[18];
#pragma unroll
  for (int o = 0; o < 18; ++o) acc[o] = 0.0;
  if (i < nb) {
    for (int k = k0; k < k1; ++k) {
      const double w0 = vort[k], w1 = vort[k + nn], w2 = vort[k + 2 * (size_t)nn];
      const size_t a = i + (size_t)nb * k;
      const double m[3] = {Dx[a], Dy[a], Dz[a]};
      const int dom = isbnd[k] ? 0 : 1;
      if (dom == 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          acc[q * 6 + 0] += m[q] * w0;
          acc[q * 6 + 2] += m[q] * w1;
          acc[q * 6 + 4] += m[q] * w2;
        }
      } else {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          acc[q * 6 + 1] += m[q] * w0;
          acc[q * 6 + 3] += m[q] * w1;
          acc[q * 6 + 5] += m[q] * w2;
        }
      }
    }
#pragma unroll
    for (int o = 0; o < 18; ++o) part[((size_t)ks * 18 + o) * nb + i] = acc[o];
  }
}

// out[o*nb+i] (+)= sum over chunks of part, in chunk order
__global__ void reduce_parts_kernel(int nb, int no, int ks_cnt, const double *__restrict__ part, double *__restrict__ out,
                                    int accumulate) {
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nb * no) return;
  int o = (int)(idx / nb), i = (int)(idx % nb);
  double s = 0.0;
  for (int k = 0; k < ks_cnt; ++k) s += part[((size_t)k * no + o) * nb + i];
  out[idx] = accumulate ? out[idx] + s : s;
}

// right-hand side of equation e (sdkm_rhs) from the fused sums
__global__ void combine_kernel(int nb, int nn, int e, const double *__restrict__ YH, const double *__restrict__ YD,
                               const double *__restrict__ nxp, const double *__restrict__ nyp,
                               const double *__restrict__ nzp, const int *__restrict__ kode,
                               const int *__restrict__ gbn, const double *__restrict__ vort, double *__restrict__ dfi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const double nx = nxp[i], ny = nyp[i], nz = nzp[i];
  auto Hs = [&](int q, int c) { return YH[(size_t)(q * 3 + c) * nb + i]; };
  auto DB = [&](int q, int c) { return YD[(size_t)(q * 6 + c * 2) * nb + i]; };
  auto DD = [&](int q, int c) { return YD[(size_t)(q * 6 + c * 2 + 1) * nb + i]; };
  double d;
  if (e == 0)
    d = (ny * Hs(0, 2) - nz * Hs(0, 1)) - (ny * Hs(1, 1) + nz * Hs(1, 2)) + (ny * Hs(2, 0) + nz * Hs(3, 0)) +
        (nx * DB(0, 0) + ny * DB(0, 1) + nz * DB(0, 2)) + (ny * DD(0, 1) + nz * DD(0, 2)) -
        (ny * DD(1, 0) + nz * DD(2, 0));
  else if (e == 1)
    d = (nz * Hs(0, 0) - nx * Hs(0, 2)) - (nx * Hs(2, 0) + nz * Hs(2, 2)) + (nx * Hs(1, 1) + nz * Hs(3, 1)) +
        (nx * DB(1, 0) + ny * DB(1, 1) + nz * DB(1, 2)) + (nx * DD(1, 0) + nz * DD(1, 2)) -
        (nx * DD(0, 1) + nz * DD(2, 1));
  else
    d = (nx * Hs(0, 1) - ny * Hs(0, 0)) - (nx * Hs(3, 0) + ny * Hs(3, 1)) + (nx * Hs(1, 2) + ny * Hs(2, 2)) +
        (nx * DB(2, 0) + ny * DB(2, 1) + nz * DB(2, 2)) + (nx * DD(2, 0) + ny * DD(2, 1)) -
        (nx * DD(0, 2) + ny * DD(1, 2));
  int kd = kode[i + (size_t)nb * e];
  if (kd == 1 || kd == 2) d = vort[(gbn[i] - 1) + (size_t)nn * e];
  dfi[i] = d;
}

// y = M * x for a column-major nb x nb M, split over column chunks
__global__ void __launch_bounds__(kRowsPerBlock) gemv_kernel(int nb, int ks_cnt, const double *__restrict__ M,
                                                             const double *__restrict__ x, double *__restrict__ part) {
  const int i = blockIdx.x * kRowsPerBlock + threadIdx.x, ks = blockIdx.y;
  const int chunk = (nb + ks_cnt - 1) / ks_cnt, j0 = ks * chunk, j1 = min(nb, j0 + chunk);
  if (i >= nb) return;
  double acc = 0.0;
  for (int j = j0; j < j1; ++j) acc += M[i + (size_t)nb * j] * x[j];
  part[(size_t)ks * nb + i] = acc;
}

// sdkm_distunk: relaxed update of rows with kode 0/3; dw = new - old
__global__ void relax_kernel(int nb, int nn, int e, double ur, const double *__restrict__ xs,
                             const int *__restrict__ kode, const int *__restrict__ gbn, double *__restrict__ vort,
                             double *__restrict__ dw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  int kd = kode[i + (size_t)nb * e];
  double d = 0.0;
  if (kd == 0 || kd == 3) {
    size_t a = (gbn[i] - 1) + (size_t)nn * e;
    double o = vort[a];
    double nw = o * (1.0 - ur) + ur * xs[i];
    vort[a] = nw;
    d = nw - o;
  }
  dw[i] = d;
}

// boundary sums of later equations after a boundary-vorticity update:
// part[(ks*nm + t)*nb + i] = sum_j D_{m_t}(i, gbn(j)) * dw(j)
__global__ void __launch_bounds__(kRowsPerBlock) dcorr_kernel(int nb, int ks_cnt, int nm, const double *__restrict__ D0,
                                                              const double *__restrict__ D1,
                                                              const int *__restrict__ gbn,
                                                              const double *__restrict__ dw, double *__restrict__ part) {
  const int i = blockIdx.x * kRowsPerBlock + threadIdx.x, ks = blockIdx.y;
  const int chunk = (nb + ks_cnt - 1) / ks_cnt, j0 = ks * chunk, j1 = min(nb, j0 + chunk);
  if (i >= nb) return;
  double a0 = 0.0, a1 = 0.0;
  for (int j = j0; j < j1; ++j) {
    const double d = dw[j];
    if (d == 0.0) continue;  // uniform across the wavefront
    const size_t col = (size_t)(gbn[j] - 1) * nb;
    a0 += D0[i + col] * d;
    if (nm > 1) a1 += D1[i + col] * d;
  }
  part[((size_t)ks * nm + 0) * nb + i] = a0;
  if (nm > 1) part[((size_t)ks * nm + 1) * nb + i] = a1;
}

// adds correction sums into YD slots
__global__ void add_slots_kernel(int nb, int nm, int ks_cnt, const double *__restrict__ part, const int s0, const int s1,
                                 double *__restrict__ YD) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  for (int t = 0; t < nm; ++t) {
    double s = 0.0;
    for (int k = 0; k < ks_cnt; ++k) s += part[((size_t)k * nm + t) * nb + i];
    int slot = t == 0 ? s0 : s1;
    YD[(size_t)slot * nb + i] += s;
  }
}

int split_for(int nb, int ncols) {
  // enough wavefronts to keep HBM busy (~8 per CU on 256 CUs)
  int rows = (nb + kRowsPerBlock - 1) / kRowsPerBlock;
  int ks = (2048 + rows - 1) / rows;
  ks = std::max(1, std::min(ks, std::max(1, ncols / 64)));
  return ks;
}

}  // namespace

void dense_setup(DenseKin &d, const DevMesh &m, const int *kode_host, SdPrec prec, hipStream_t s, double *ms_lu) {
  const int nb = m.nb, nn = m.nn;
  d.nb = nb, d.nn = nn, d.prec = prec;
  d.kode.upload(kode_host, (size_t)nb * 3, s);
  std::vector<int> isb(nn, 0);
  std::vector<int> lbn(nn);
  TT_HIP(hipMemcpyAsync(lbn.data(), m.lbn.p, sizeof(int) * nn, hipMemcpyDeviceToHost, s));
  TT_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < nn; ++k) isb[k] = lbn[k];
  d.isbnd.upload(isb, s);
  rocblas_handle h;
  if (rocblas_create_handle(&h) != rocblas_status_success) throw std::runtime_error("rocblas_create_handle failed");
  rocblas_set_stream(h, s);
  DBuf<int> ipiv(nb), info(1);
  hipEvent_t e0, e1;
  TT_HIP(hipEventCreate(&e0));
  TT_HIP(hipEventCreate(&e1));
  TT_HIP(hipEventRecord(e0, s));
  for (int e = 0; e < 3; ++e) {
    d.Ainv[e].alloc((size_t)nb * nb);
    hipLaunchKernelGGL(build_sys_kernel, dim3(ceil_div((size_t)nb * nb, 256)), dim3(256), 0, s, nb, e, d.kode.p,
                       m.nx.p, m.ny.p, m.nz.p, m.gbn.p, d.mats.Dx.p, d.mats.Dy.p, d.mats.Dz.p, d.Ainv[e].p);
    TT_LAUNCH_CHECK();
    if (rocsolver_dgetrf(h, nb, nb, d.Ainv[e].p, nb, ipiv.p, info.p) != rocblas_status_success)
      throw std::runtime_error("rocsolver_dgetrf failed");
    int inf = 0;
    info.download(&inf, 1, s);
    TT_HIP(hipStreamSynchronize(s));
    if (inf != 0) throw std::runtime_error("single-domain system matrix is singular");
    if (rocsolver_dgetri(h, nb, d.Ainv[e].p, nb, ipiv.p, info.p) != rocblas_status_success)
      throw std::runtime_error("rocsolver_dgetri failed");
  }
  TT_HIP(hipEventRecord(e1, s));
  TT_HIP(hipEventSynchronize(e1));
  float ms = 0;
  TT_HIP(hipEventElapsedTime(&ms, e0, e1));
  if (ms_lu) *ms_lu = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  rocblas_destroy_handle(h);
  d.YH.alloc((size_t)12 * nb);
  d.YD.alloc((size_t)18 * nb);
  int ksd = split_for(nb, nn), ksh = split_for(nb, nb);
  d.part.alloc((size_t)std::max(ksd * 18, ksh * 12) * nb);
  d.dfi.alloc(nb), d.xs.alloc(nb), d.dw.alloc(nb);
}

void dense_bw_step(DenseKin &d, const DevMesh &m, const double *vel, double *vort, const double urBw[3],
                   hipStream_t s) {
  const int nb = d.nb, nn = d.nn;
  const unsigned rb = ceil_div(nb, kRowsPerBlock);
  const int ksh = split_for(nb, nb), ksd = split_for(nb, nn);
  const SdDev &M = d.mats;
  hipLaunchKernelGGL(prod_h_kernel, dim3(rb, ksh), dim3(kRowsPerBlock), 0, s, nb, nn, ksh, M.H.p, M.Htx.p, M.Hty.p,
                     M.Htz.p, vel, m.gbn.p, d.part.p);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(reduce_parts_kernel, dim3(ceil_div((size_t)12 * nb, 256)), dim3(256), 0, s, nb, 12, ksh, d.part.p,
                     d.YH.p, 0);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(prod_d_kernel, dim3(rb, ksd), dim3(kRowsPerBlock), 0, s, nb, nn, ksd, M.Dx.p, M.Dy.p, M.Dz.p,
                     vort, d.isbnd.p, d.part.p);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(reduce_parts_kernel, dim3(ceil_div((size_t)18 * nb, 256)), dim3(256), 0, s, nb, 18, ksd,
                     d.part.p, d.YD.p, 0);
  TT_LAUNCH_CHECK();
  const double *Dm[3] = {M.Dx.p, M.Dy.p, M.Dz.p};
  for (int e = 0; e < 3; ++e) {
    hipLaunchKernelGGL(combine_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, nb, nn, e, d.YH.p, d.YD.p, m.nx.p,
                       m.ny.p, m.nz.p, d.kode.p, m.gbn.p, vort, d.dfi.p);
    hipLaunchKernelGGL(gemv_kernel, dim3(rb, ksh), dim3(kRowsPerBlock), 0, s, nb, ksh, d.Ainv[e].p, d.dfi.p, d.part.p);
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, nb, 1, ksh, d.part.p, d.xs.p, 0);
    hipLaunchKernelGGL(relax_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, nb, nn, e, urBw[e], d.xs.p, d.kode.p,
                       m.gbn.p, vort, d.dw.p);
    TT_LAUNCH_CHECK();
    if (e < 2) {
      // later equations e2 > e use sum_{boundary k} D_{e2}(i,k) w_e(k): slot e2*6 + e*2
      int nm = 2 - e;
      const double *D0 = Dm[e + 1], *D1 = nm > 1 ? Dm[e + 2] : nullptr;
      hipLaunchKernelGGL(dcorr_kernel, dim3(rb, ksh), dim3(kRowsPerBlock), 0, s, nb, ksh, nm, D0, D1, m.gbn.p, d.dw.p,
                         d.part.p);
      hipLaunchKernelGGL(add_slots_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, nb, nm, ksh, d.part.p,
                         (e + 1) * 6 + e * 2, (e + 2) * 6 + e * 2, d.YD.p);
      TT_LAUNCH_CHECK();
    }
  }
}

}  // namespace tt
