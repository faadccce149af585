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
#include "tt_sums.h"

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

// one column's three products into the boundary (D = 0) or domain (D = 1)
// slots q*6 + c*2 + D
template <int D>
__device__ __forceinline__ void prod_d_add(double (&acc)[18], const double (&m)[3], const double (&w)[3]) {
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[q * 6 + c * 2 + D] += m[q] * w[c];
}

// [Dx Dy Dz](i, k) * vort(k, c), k split into boundary (0) / domain (1):
// 18 sums per row, output index m*6 + c*2 + dom; four columns' loads issued
// together, added in column order
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
    int k = k0;
    for (; k + 4 <= k1; k += 4) {
      double m[4][3], w[4][3];
      int bnd[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const size_t a = i + (size_t)nb * (k + u);
        m[u][0] = Dx[a], m[u][1] = Dy[a], m[u][2] = Dz[a];
        w[u][0] = vort[k + u], w[u][1] = vort[k + u + nn], w[u][2] = vort[k + u + 2 * (size_t)nn];
        bnd[u] = isbnd[k + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (bnd[u])  // uniform across the wavefront
          prod_d_add<0>(acc, m[u], w[u]);
        else
          prod_d_add<1>(acc, m[u], w[u]);
      }
    }
    for (; k < k1; ++k) {
      const size_t a = i + (size_t)nb * k;
      const double m[3] = {Dx[a], Dy[a], Dz[a]};
      const double w[3] = {vort[k], vort[k + nn], vort[k + 2 * (size_t)nn]};
      if (isbnd[k])
        prod_d_add<0>(acc, m, w);
      else
        prod_d_add<1>(acc, m, w);
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
  const double s = slice_sum(part + (size_t)o * nb + i, (size_t)no * nb, ks_cnt);
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
  int j = j0;
  // eight columns' loads issued together, added in column order
  for (; j + 8 <= j1; j += 8) {
    double m[8], xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) m[u] = M[i + (size_t)nb * (j + u)], xv[u] = x[j + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += m[u] * xv[u];
  }
  for (; j < j1; ++j) acc += M[i + (size_t)nb * j] * x[j];
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
  int j = j0;
  // columns with a change (uniform across the wavefront), four columns'
  // loads issued together, added in column order
  for (; j + 4 <= j1; j += 4) {
    double d[4], m0[4], m1[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      d[u] = dw[j + u];
      m0[u] = m1[u] = 0.0;
      if (d[u] != 0.0) {
        const size_t col = (size_t)(gbn[j + u] - 1) * nb;
        m0[u] = D0[i + col];
        if (nm > 1) m1[u] = D1[i + col];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (d[u] != 0.0) {
        a0 += m0[u] * d[u];
        if (nm > 1) a1 += m1[u] * d[u];
      }
  }
  for (; j < j1; ++j) {
    const double d = dw[j];
    if (d == 0.0) continue;
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
    const double s = slice_sum(part + (size_t)t * nb + i, (size_t)nm * nb, ks_cnt);
    int slot = t == 0 ? s0 : s1;
    YD[(size_t)slot * nb + i] += s;
  }
}

// ---------------------------------------------------------------------------
// Reduced-precision storage: value = stored * scale(row), scale a power of two
// chosen so that the row's largest magnitude lies in [0.5, 1).

enum class Lo { F32, F16, BF16 };

__device__ __forceinline__ float bf16_to_f(uint16_t u) { return __uint_as_float((uint32_t)u << 16); }
__device__ __forceinline__ uint16_t f_to_bf16(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ uint16_t f_to_f16(float f) {
  _Float16 h = (_Float16)f;
  return *reinterpret_cast<uint16_t *>(&h);
}
__device__ __forceinline__ float f16_to_f(uint16_t u) { return (float)*reinterpret_cast<_Float16 *>(&u); }

// four consecutive rows of one column
template <Lo L>
struct Load4;
template <>
struct Load4<Lo::F32> {
  using T = float;
  __device__ static float4 at(const float *p) { return *reinterpret_cast<const float4 *>(p); }
};
template <>
struct Load4<Lo::F16> {
  using T = uint16_t;
  __device__ static float4 at(const uint16_t *p) {
    const ushort4 v = *reinterpret_cast<const ushort4 *>(p);
    return make_float4(f16_to_f(v.x), f16_to_f(v.y), f16_to_f(v.z), f16_to_f(v.w));
  }
};
template <>
struct Load4<Lo::BF16> {
  using T = uint16_t;
  __device__ static float4 at(const uint16_t *p) {
    const ushort4 v = *reinterpret_cast<const ushort4 *>(p);
    return make_float4(bf16_to_f(v.x), bf16_to_f(v.y), bf16_to_f(v.z), bf16_to_f(v.w));
  }
};

constexpr int kLoRows = 4;                    // rows per thread of the product kernels
constexpr int kLoBlockRows = 64 * kLoRows;    // rows per workgroup (and per stored row block)

// offset of (row i0, column j) in a copy with nc columns (i0 a multiple of
// kLoRows: the thread's four rows follow)
__device__ __forceinline__ size_t lo_at(int i0, size_t j, int nc) {
  return ((size_t)(i0 / kLoBlockRows) * nc + j) * kLoBlockRows + i0 % kLoBlockRows;
}

// largest |M(i, j)| over a chunk of columns, per row
__global__ void rowmax_kernel(int nr, int nc, int ks_cnt, const double *__restrict__ M, float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, ks = blockIdx.y;
  if (i >= nr) return;
  const int chunk = (nc + ks_cnt - 1) / ks_cnt, j0 = ks * chunk, j1 = min(nc, j0 + chunk);
  double m = 0.0;
  for (int j = j0; j < j1; ++j) m = fmax(m, fabs(M[i + (size_t)nr * j]));
  out[(size_t)ks * nr + i] = (float)m;
}
__global__ void rowscale_kernel(int nr, int ks_cnt, const float *__restrict__ part, float *__restrict__ scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  float m = 0.f;
  for (int k = 0; k < ks_cnt; ++k) m = fmaxf(m, part[(size_t)k * nr + i]);
  int e = 0;
  if (m > 0.f) frexpf(m, &e);  // m = f * 2^e, f in [0.5, 1)
  scale[i] = ldexpf(1.f, e);   // stored = value / scale
}
// Reduced-precision copies are stored in row blocks of kLoBlockRows rows:
// block b holds rows b*kLoBlockRows.. of every column, column after column,
// so that a workgroup of the product kernels (one row block, a range of
// columns) streams one contiguous piece of memory.
template <Lo L>
__global__ void to_low_kernel(int nr, int nc, int ldl, const double *__restrict__ M, const float *__restrict__ scale,
                              typename Load4<L>::T *__restrict__ out) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)ldl * nc) return;
  const size_t bsz = (size_t)kLoBlockRows * nc;
  const int i = (int)(k / bsz) * kLoBlockRows + (int)(k % kLoBlockRows);
  const size_t j = k % bsz / kLoBlockRows;
  const float v = i < nr ? (float)(M[i + (size_t)nr * j] / (double)scale[i]) : 0.f;
  if constexpr (L == Lo::F32)
    out[k] = v;
  else if constexpr (L == Lo::F16)
    out[k] = f_to_f16(v);
  else
    out[k] = f_to_bf16(v);
}

// the change of a field since the anchor step in FP32: out[k] = (f0, f1, f2,
// flag) for node gather[k] (or k), flag = 1 for a domain node (isbnd == 0)
__global__ void delta4_kernel(int n, int nn, const int *__restrict__ gather, const double *__restrict__ f,
                              const double *__restrict__ fa, const int *__restrict__ isbnd, float4 *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const size_t g = gather ? (size_t)(gather[k] - 1) : (size_t)k;
  out[k] = make_float4((float)(f[g] - fa[g]), (float)(f[g + nn] - fa[g + nn]),
                       (float)(f[g + 2 * (size_t)nn] - fa[g + 2 * (size_t)nn]),
                       isbnd ? (isbnd[g] ? 0.f : 1.f) : 0.f);
}


// delta of [H Htx Hty Htz] products: sum_j Hq(i,j) * dv(j, c), dv = the change of
// the velocity at boundary node gbn(j) (delta4_kernel); four columns per step,
// their loads issued together
template <Lo L>
__global__ void __launch_bounds__(64) lprod_h_kernel(int nb, int ldl, int ks_cnt,
                                                     const typename Load4<L>::T *__restrict__ H0,
                                                     const typename Load4<L>::T *__restrict__ H1,
                                                     const typename Load4<L>::T *__restrict__ H2,
                                                     const typename Load4<L>::T *__restrict__ H3,
                                                     const float *__restrict__ rscale, const float4 *__restrict__ dv,
                                                     double *__restrict__ part) {
  const int i0 = blockIdx.x * kLoBlockRows + threadIdx.x * kLoRows, ks = blockIdx.y;
  const int chunk = (nb + ks_cnt - 1) / ks_cnt, j0 = ks * chunk, j1 = min(nb, j0 + chunk);
  float acc[4][3][4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[q][c][r] = 0.f;
  if (i0 >= nb) return;
  const typename Load4<L>::T *Hq[4] = {H0, H1, H2, H3};
  auto step = [&](const float4 d, const float4 (&m)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float mv[4] = {m[q].x, m[q].y, m[q].z, m[q].w};
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        acc[q][0][r] += mv[r] * d.x;
        acc[q][1][r] += mv[r] * d.y;
        acc[q][2][r] += mv[r] * d.z;
      }
    }
  };
  int j = j0;
  for (; j + 4 <= j1; j += 4) {
    float4 d[4], m[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      d[u] = dv[j + u];
      const size_t o = lo_at(i0, j + u, nb);
#pragma unroll
      for (int q = 0; q < 4; ++q) m[u][q] = Load4<L>::at(Hq[q] + o);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) step(d[u], m[u]);
  }
  for (; j < j1; ++j) {
    float4 m[4];
    const size_t o = lo_at(i0, j, nb);
#pragma unroll
    for (int q = 0; q < 4; ++q) m[q] = Load4<L>::at(Hq[q] + o);
    step(dv[j], m);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + r;
    if (i >= nb) break;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        part[((size_t)ks * 12 + q * 3 + c) * nb + i] = (double)acc[q][c][r] * (double)rscale[(size_t)q * ldl + i];
  }
}

// delta of [Dx Dy Dz] products split into boundary (0) / domain (1) columns:
// slot m*6 + c*2 + dom; dw(k) = (change of the vorticity at node k, domain
// flag) from delta4_kernel; four columns per step, loads issued together
template <Lo L>
__global__ void __launch_bounds__(64) lprod_d_kernel(int nb, int nn, int ldl, int ks_cnt,
                                                     const typename Load4<L>::T *__restrict__ D0,
                                                     const typename Load4<L>::T *__restrict__ D1,
                                                     const typename Load4<L>::T *__restrict__ D2,
                                                     const float *__restrict__ rscale, const float4 *__restrict__ dw,
                                                     double *__restrict__ part) {
  const int i0 = blockIdx.x * kLoBlockRows + threadIdx.x * kLoRows, ks = blockIdx.y;
  const int chunk = (nn + ks_cnt - 1) / ks_cnt, k0 = ks * chunk, k1 = min(nn, k0 + chunk);
  float acc[3][6][4];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int o = 0; o < 6; ++o)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[q][o][r] = 0.f;
  if (i0 >= nb) return;
  const typename Load4<L>::T *Dq[3] = {D0, D1, D2};
  // boundary columns add into slots 0, 2, 4, domain columns into 1, 3, 5
#define TT_LPROD_D_ADD(W, M, S0)                                    \
  _Pragma("unroll") for (int q = 0; q < 3; ++q) {                   \
    const float mv[4] = {(M)[q].x, (M)[q].y, (M)[q].z, (M)[q].w};   \
    _Pragma("unroll") for (int r = 0; r < 4; ++r) {                 \
      acc[q][(S0)][r] += mv[r] * (W).x;                             \
      acc[q][(S0) + 2][r] += mv[r] * (W).y;                         \
      acc[q][(S0) + 4][r] += mv[r] * (W).z;                         \
    }                                                               \
  }
  // the branch is uniform across the wavefront (one column per step)
#define TT_LPROD_D_STEP(W, M)   \
  if ((W).w == 0.f) {           \
    TT_LPROD_D_ADD(W, M, 0)     \
  } else {                      \
    TT_LPROD_D_ADD(W, M, 1)     \
  }
  int k = k0;
  for (; k + 4 <= k1; k += 4) {
    float4 w[4], m[4][3];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      w[u] = dw[k + u];
      const size_t o = lo_at(i0, k + u, nn);
#pragma unroll
      for (int q = 0; q < 3; ++q) m[u][q] = Load4<L>::at(Dq[q] + o);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) TT_LPROD_D_STEP(w[u], m[u])
  }
  for (; k < k1; ++k) {
    float4 m[3];
    const size_t o = lo_at(i0, k, nn);
#pragma unroll
    for (int q = 0; q < 3; ++q) m[q] = Load4<L>::at(Dq[q] + o);
    const float4 w = dw[k];
    TT_LPROD_D_STEP(w, m)
  }
#undef TT_LPROD_D_STEP
#undef TT_LPROD_D_ADD
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + r;
    if (i >= nb) break;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double sc = rscale[(size_t)(4 + q) * ldl + i];
#pragma unroll
      for (int o = 0; o < 6; ++o) part[((size_t)ks * 18 + q * 6 + o) * nb + i] = (double)acc[q][o][r] * sc;
    }
  }
}

// partials of M_lo * (x - x_a) for an nb x nb matrix (x_a may be null)
template <Lo L>
__global__ void __launch_bounds__(64) lgemv_kernel(int nb, int ldl, int ks_cnt, const typename Load4<L>::T *__restrict__ M,
                                                   const float *__restrict__ scale, const double *__restrict__ x,
                                                   const double *__restrict__ x_a, double *__restrict__ part) {
  const int i0 = blockIdx.x * kLoBlockRows + threadIdx.x * kLoRows, ks = blockIdx.y;
  const int chunk = (nb + ks_cnt - 1) / ks_cnt, j0 = ks * chunk, j1 = min(nb, j0 + chunk);
  if (i0 >= nb) return;
  float a[4] = {0.f, 0.f, 0.f, 0.f};
  // eight columns' loads in flight per lane
#pragma unroll 8
  for (int j = j0; j < j1; ++j) {
    const float d = (float)(x_a ? x[j] - x_a[j] : x[j]);
    const float4 m = Load4<L>::at(M + lo_at(i0, j, nb));
    a[0] += m.x * d, a[1] += m.y * d, a[2] += m.z * d, a[3] += m.w * d;
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
    if (i0 + r < nb) part[(size_t)ks * nb + i0 + r] = (double)a[r] * (double)scale[i0 + r];
}

// partials of sum_j D_lo(i, gbn(j)) * dw(j) for one or two D matrices
template <Lo L>
__global__ void __launch_bounds__(64) ldcorr_kernel(int nb, int nn, int ks_cnt, int nm,
                                                    const typename Load4<L>::T *__restrict__ D0,
                                                    const typename Load4<L>::T *__restrict__ D1,
                                                    const float *__restrict__ s0, const float *__restrict__ s1,
                                                    const int *__restrict__ gbn, const double *__restrict__ dw,
                                                    double *__restrict__ part) {
  const int i0 = blockIdx.x * kLoBlockRows + threadIdx.x * kLoRows, ks = blockIdx.y;
  const int chunk = (nb + ks_cnt - 1) / ks_cnt, j0 = ks * chunk, j1 = min(nb, j0 + chunk);
  if (i0 >= nb) return;
  float a0[4] = {0.f, 0.f, 0.f, 0.f}, a1[4] = {0.f, 0.f, 0.f, 0.f};
  // eight columns' loads in flight per lane (a zero change adds zero: no
  // branch, so the loop unrolls)
#pragma unroll 8
  for (int j = j0; j < j1; ++j) {
    const float d = (float)dw[j];
    const size_t col = lo_at(i0, (size_t)(gbn[j] - 1), nn);
    const float4 m = Load4<L>::at(D0 + col);
    a0[0] += m.x * d, a0[1] += m.y * d, a0[2] += m.z * d, a0[3] += m.w * d;
    if (nm > 1) {
      const float4 n = Load4<L>::at(D1 + col);
      a1[0] += n.x * d, a1[1] += n.y * d, a1[2] += n.z * d, a1[3] += n.w * d;
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + r;
    if (i >= nb) break;
    part[((size_t)ks * nm + 0) * nb + i] = (double)a0[r] * (double)s0[i];
    if (nm > 1) part[((size_t)ks * nm + 1) * nb + i] = (double)a1[r] * (double)s1[i];
  }
}

// out[o*nb+i] = base[o*nb+i] + sum over chunks of part
__global__ void reduce_base_kernel(int nb, int no, int ks_cnt, const double *__restrict__ part,
                                   const double *__restrict__ base, double *__restrict__ out) {
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nb * no) return;
  int o = (int)(idx / nb), i = (int)(idx % nb);
  const double s = slice_sum(part + (size_t)o * nb + i, (size_t)no * nb, ks_cnt);
  out[idx] = base[idx] + s;
}

__global__ void copy_kernel(size_t n, const double *__restrict__ a, double *__restrict__ b) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) b[k] = a[k];
}

int lo_split(int nb, int ncols) {
  int rows = (nb + kLoBlockRows - 1) / kLoBlockRows;
  int ks = (2048 + rows - 1) / rows;
  return std::max(1, std::min(ks, std::max(1, ncols / 64)));
}

int split_for(int nb, int ncols, int waves = 2048) {
  // enough wavefronts to keep HBM busy (~8 per CU on 256 CUs)
  int rows = (nb + kRowsPerBlock - 1) / kRowsPerBlock;
  int ks = (waves + rows - 1) / rows;
  ks = std::max(1, std::min(ks, std::max(1, ncols / 64)));
  return ks;
}
// prod_d_kernel: about 6 MB of loads in flight over the chip (1,024
// wavefronts of four columns x three matrices) streams fastest; twice that
// is ~15% slower (9,602 x 68,921: 5.8 against 4.9 TB/s; 13,826 x 117,649:
// 5.7-5.9 against 5.2; 1,538 x 4,913: 5.5 against 5.0)
int split_d(int nb, int nn) { return split_for(nb, nn, 1024); }

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
  // the three equations' systems factorised and inverted together (one
  // batched call each: every panel step launches once for all three)
  DBuf<int> ipiv((size_t)3 * nb), info(3);
  DBuf<double *> aptr(3);
  hipEvent_t e0, e1;
  TT_HIP(hipEventCreate(&e0));
  TT_HIP(hipEventCreate(&e1));
  TT_HIP(hipEventRecord(e0, s));
  double *ap[3];
  for (int e = 0; e < 3; ++e) {
    d.Ainv[e].alloc((size_t)nb * nb);
    hipLaunchKernelGGL(build_sys_kernel, dim3(ceil_div((size_t)nb * nb, 256)), dim3(256), 0, s, nb, e, d.kode.p,
                       m.nx.p, m.ny.p, m.nz.p, m.gbn.p, d.mats.Dx.p, d.mats.Dy.p, d.mats.Dz.p, d.Ainv[e].p);
    TT_LAUNCH_CHECK();
    ap[e] = d.Ainv[e].p;
  }
  aptr.upload(ap, 3, s);
  if (rocsolver_dgetrf_batched(h, nb, nb, aptr.p, nb, ipiv.p, nb, info.p, 3) != rocblas_status_success)
    throw std::runtime_error("rocsolver_dgetrf_batched failed");
  int inf[3] = {0, 0, 0};
  info.download(inf, 3, s);
  TT_HIP(hipStreamSynchronize(s));
  if (inf[0] != 0 || inf[1] != 0 || inf[2] != 0) throw std::runtime_error("single-domain system matrix is singular");
  if (rocsolver_dgetri_batched(h, nb, aptr.p, nb, ipiv.p, nb, info.p, 3) != rocblas_status_success)
    throw std::runtime_error("rocsolver_dgetri_batched failed");
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
  int ksd = split_d(nb, nn), ksh = split_for(nb, nb);
  d.part.alloc((size_t)std::max(ksd * 18, ksh * 12) * nb);
  d.dfi.alloc(nb), d.xs.alloc(nb), d.dw.alloc(nb);
}

namespace {

// FP64 step; with save = true it also records the anchor (fields, products,
// right-hand sides and solutions of the three equations)
void step_fp64(DenseKin &d, const DevMesh &m, const double *vel, double *vort, const double urBw[3], hipStream_t s,
               bool save) {
  const int nb = d.nb, nn = d.nn;
  const unsigned rb = ceil_div(nb, kRowsPerBlock);
  const int ksh = split_for(nb, nb), ksd = split_d(nb, nn);
  const SdDev &M = d.mats;
  if (save) {
    hipLaunchKernelGGL(copy_kernel, dim3(ceil_div((size_t)nn * 3, 256)), dim3(256), 0, s, (size_t)nn * 3, vel,
                       d.vel_a.p);
    hipLaunchKernelGGL(copy_kernel, dim3(ceil_div((size_t)nn * 3, 256)), dim3(256), 0, s, (size_t)nn * 3, vort,
                       d.vort_a.p);
  }
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
  if (save) {
    hipLaunchKernelGGL(copy_kernel, dim3(ceil_div((size_t)12 * nb, 256)), dim3(256), 0, s, (size_t)12 * nb, d.YH.p,
                       d.YH_a.p);
    hipLaunchKernelGGL(copy_kernel, dim3(ceil_div((size_t)18 * nb, 256)), dim3(256), 0, s, (size_t)18 * nb, d.YD.p,
                       d.YD_a.p);
  }
  const double *Dm[3] = {M.Dx.p, M.Dy.p, M.Dz.p};
  for (int e = 0; e < 3; ++e) {
    hipLaunchKernelGGL(combine_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, nb, nn, e, d.YH.p, d.YD.p, m.nx.p,
                       m.ny.p, m.nz.p, d.kode.p, m.gbn.p, vort, d.dfi.p);
    hipLaunchKernelGGL(gemv_kernel, dim3(rb, ksh), dim3(kRowsPerBlock), 0, s, nb, ksh, d.Ainv[e].p, d.dfi.p, d.part.p);
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, nb, 1, ksh, d.part.p, d.xs.p, 0);
    if (save) {
      hipLaunchKernelGGL(copy_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, (size_t)nb, d.dfi.p,
                         d.dfi_a.p + (size_t)nb * e);
      hipLaunchKernelGGL(copy_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, (size_t)nb, d.xs.p,
                         d.sol_a.p + (size_t)nb * e);
    }
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

// delta step on the reduced-precision copies, relative to the last anchor
template <Lo L>
void step_delta(DenseKin &d, const DevMesh &m, const double *vel, double *vort, const double urBw[3], hipStream_t s) {
  using T = typename Load4<L>::T;
  auto lo = [&](int k) -> const T * {
    if constexpr (L == Lo::F32)
      return d.lo32[k].p;
    else
      return d.lo16[k].p;
  };
  const int nb = d.nb, nn = d.nn, ldl = d.ldl;
  const unsigned rb = ceil_div(nb, kLoBlockRows);
  const int ksh = lo_split(nb, nb), ksd = lo_split(nb, nn);
  const float *sc = d.rscale.p;
  hipLaunchKernelGGL(delta4_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, nb, nn, m.gbn.p, vel, d.vel_a.p,
                     (const int *)nullptr, d.dv4.p);
  hipLaunchKernelGGL(delta4_kernel, dim3(ceil_div(nn, 256)), dim3(256), 0, s, nn, nn, (const int *)nullptr, vort,
                     d.vort_a.p, d.isbnd.p, d.dw4.p);
  hipLaunchKernelGGL((lprod_h_kernel<L>), dim3(rb, ksh), dim3(64), 0, s, nb, ldl, ksh, lo(0), lo(1), lo(2), lo(3), sc,
                     d.dv4.p, d.part.p);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(reduce_base_kernel, dim3(ceil_div((size_t)12 * nb, 256)), dim3(256), 0, s, nb, 12, ksh, d.part.p,
                     d.YH_a.p, d.YH.p);
  hipLaunchKernelGGL((lprod_d_kernel<L>), dim3(rb, ksd), dim3(64), 0, s, nb, nn, ldl, ksd, lo(4), lo(5), lo(6), sc,
                     d.dw4.p, d.part.p);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(reduce_base_kernel, dim3(ceil_div((size_t)18 * nb, 256)), dim3(256), 0, s, nb, 18, ksd, d.part.p,
                     d.YD_a.p, d.YD.p);
  for (int e = 0; e < 3; ++e) {
    hipLaunchKernelGGL(combine_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, nb, nn, e, d.YH.p, d.YD.p, m.nx.p,
                       m.ny.p, m.nz.p, d.kode.p, m.gbn.p, vort, d.dfi.p);
    hipLaunchKernelGGL((lgemv_kernel<L>), dim3(rb, ksh), dim3(64), 0, s, nb, ldl, ksh, lo(7 + e),
                       sc + (size_t)(7 + e) * ldl, (const double *)d.dfi.p, (const double *)(d.dfi_a.p + (size_t)nb * e),
                       d.part.p);
    hipLaunchKernelGGL(reduce_base_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, nb, 1, ksh, d.part.p,
                       d.sol_a.p + (size_t)nb * e, d.xs.p);
    hipLaunchKernelGGL(relax_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, nb, nn, e, urBw[e], d.xs.p, d.kode.p,
                       m.gbn.p, vort, d.dw.p);
    TT_LAUNCH_CHECK();
    if (e < 2) {
      const int nm = 2 - e;
      hipLaunchKernelGGL((ldcorr_kernel<L>), dim3(rb, ksh), dim3(64), 0, s, nb, nn, ksh, nm, lo(4 + e + 1),
                         nm > 1 ? lo(4 + e + 2) : lo(4 + e + 1), sc + (size_t)(4 + e + 1) * ldl,
                         sc + (size_t)(4 + (nm > 1 ? e + 2 : e + 1)) * ldl, m.gbn.p, (const double *)d.dw.p, d.part.p);
      hipLaunchKernelGGL(add_slots_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, nb, nm, ksh, d.part.p,
                         (e + 1) * 6 + e * 2, (e + 2) * 6 + e * 2, d.YD.p);
      TT_LAUNCH_CHECK();
    }
  }
}

}  // namespace

void dense_bw_step(DenseKin &d, const DevMesh &m, const double *vel, double *vort, const double urBw[3],
                   hipStream_t s) {
  if (d.prec == SdPrec::F64) {
    step_fp64(d, m, vel, vort, urBw, s, false);
    return;
  }
  if (d.since_anchor == 0)
    step_fp64(d, m, vel, vort, urBw, s, true);
  else if (d.prec == SdPrec::F32)
    step_delta<Lo::F32>(d, m, vel, vort, urBw, s);
  else if (d.prec == SdPrec::F16)
    step_delta<Lo::F16>(d, m, vel, vort, urBw, s);
  else
    step_delta<Lo::BF16>(d, m, vel, vort, urBw, s);
  d.since_anchor = (d.since_anchor + 1) % std::max(1, d.anchor_period);
}

void dense_set_precision(DenseKin &d, SdPrec prec, hipStream_t s) {
  d.since_anchor = 0;
  if (prec == d.prec) return;
  for (auto &b : d.lo16) b.release();
  for (auto &b : d.lo32) b.release();
  d.prec = prec;
  if (prec == SdPrec::F64) return;
  const int nb = d.nb, nn = d.nn;
  d.ldl = (nb + kLoBlockRows - 1) / kLoBlockRows * kLoBlockRows;  // whole row blocks
  d.rscale.alloc((size_t)10 * d.ldl);
  const SdDev &M = d.mats;
  const double *mats[10] = {M.H.p, M.Htx.p, M.Hty.p, M.Htz.p, M.Dx.p, M.Dy.p, M.Dz.p, d.Ainv[0].p, d.Ainv[1].p,
                            d.Ainv[2].p};
  const int ncol[10] = {nb, nb, nb, nb, nn, nn, nn, nb, nb, nb};
  const int ks = 64;
  DBuf<float> mx((size_t)ks * nb);
  for (int k = 0; k < 10; ++k) {
    const size_t total = (size_t)d.ldl * ncol[k];
    hipLaunchKernelGGL(rowmax_kernel, dim3(ceil_div(nb, 256), ks), dim3(256), 0, s, nb, ncol[k], ks, mats[k], mx.p);
    hipLaunchKernelGGL(rowscale_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, nb, ks, mx.p,
                       d.rscale.p + (size_t)k * d.ldl);
    if (prec == SdPrec::F32) {
      d.lo32[k].alloc(total);
      hipLaunchKernelGGL(to_low_kernel<Lo::F32>, dim3(ceil_div(total, 256)), dim3(256), 0, s, nb, ncol[k], d.ldl,
                         mats[k], d.rscale.p + (size_t)k * d.ldl, d.lo32[k].p);
    } else {
      d.lo16[k].alloc(total);
      if (prec == SdPrec::F16)
        hipLaunchKernelGGL(to_low_kernel<Lo::F16>, dim3(ceil_div(total, 256)), dim3(256), 0, s, nb, ncol[k], d.ldl,
                           mats[k], d.rscale.p + (size_t)k * d.ldl, d.lo16[k].p);
      else
        hipLaunchKernelGGL(to_low_kernel<Lo::BF16>, dim3(ceil_div(total, 256)), dim3(256), 0, s, nb, ncol[k], d.ldl,
                           mats[k], d.rscale.p + (size_t)k * d.ldl, d.lo16[k].p);
    }
    TT_LAUNCH_CHECK();
  }
  TT_HIP(hipStreamSynchronize(s));  // mx is released on return
  d.YH_a.alloc((size_t)12 * nb), d.YD_a.alloc((size_t)18 * nb);
  d.dfi_a.alloc((size_t)3 * nb), d.sol_a.alloc((size_t)3 * nb);
  d.vel_a.alloc((size_t)3 * nn), d.vort_a.alloc((size_t)3 * nn);
  d.dv4.alloc(nb), d.dw4.alloc(nn);
  const int need = std::max({lo_split(nb, nn) * 18, lo_split(nb, nb) * 12, split_d(nb, nn) * 18,
                             split_for(nb, nb) * 12});
  if ((size_t)need * nb > d.part.n) d.part.alloc((size_t)need * nb);
}

// ---------------------------------------------------------------------------
// Mixed-precision solve of one equation's system with FP64 iterative
// refinement: X <- X + M (B - A X) until ||B - A X||_F <= tol ||B||_F, with A
// the FP64 system matrix and M its inverse held in FP32, BF16 or FP16 with a
// power-of-two scale per row, applied to up to four right-hand sides on the
// matrix cores with FP32 accumulation (the right-hand sides scaled per column
// by a power of two into the same format).

namespace {

typedef _Float16 mx_h8 __attribute__((ext_vector_type(8)));
typedef __bf16 mx_b8 __attribute__((ext_vector_type(8)));
typedef float mx_f4 __attribute__((ext_vector_type(4)));

// k per tile: 32 for 16-bit entries (one v_mfma_f32_16x16x32_{f16,bf16}, eight
// per lane), 16 for FP32 (four v_mfma_f32_16x16x4_f32 steps, four per lane)
template <Lo L>
constexpr int mx_kt() {
  return L == Lo::F32 ? 16 : 32;
}

template <Lo L>
__device__ __forceinline__ typename Load4<L>::T mx_conv(float x) {
  if constexpr (L == Lo::F32)
    return x;
  else if constexpr (L == Lo::F16)
    return f_to_f16(x);
  else
    return f_to_bf16(x);
}

// Tile-major copy of M / rowscale: tile (ti, tk), lane l, entry j holds row
// ti*16 + l%16 and column
//   16-bit: tk*32 + 8*(l/16) + j (j < 8), the A operand of 16x16x32;
//   FP32:   tk*16 + 4*j + l/16  (j < 4), the A operands of four 16x16x4 steps;
// zero outside the n x n matrix.  A wavefront reads a tile as 64 consecutive
// 16-byte pieces.
template <Lo L>
__global__ void __launch_bounds__(64) mx_tiles_kernel(int n, int nkt, const double *__restrict__ M,
                                                      const float *__restrict__ scale,
                                                      typename Load4<L>::T *__restrict__ out) {
  constexpr int KT = mx_kt<L>(), PER = KT / 4;
  const int ti = blockIdx.x, tk = blockIdx.y, l = threadIdx.x;
  const int row = ti * 16 + (l & 15);
  const double sc = row < n ? (double)scale[row] : 1.0;
  typename Load4<L>::T v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int k = L == Lo::F32 ? tk * KT + 4 * j + (l >> 4) : tk * KT + 8 * (l >> 4) + j;
    const float x = (row < n && k < n) ? (float)(M[row + (size_t)n * k] / sc) : 0.f;
    v[j] = mx_conv<L>(x);
  }
  typename Load4<L>::T *o = out + ((size_t)(ti * nkt + tk) * 64 + l) * PER;
#pragma unroll
  for (int j = 0; j < PER; ++j) o[j] = v[j];
}

// per right-hand side c: cscale[c] = 2^e >= max |R(:, c)| (1 for a zero column)
__global__ void __launch_bounds__(256) mx_colscale_kernel(int n, const double *__restrict__ R,
                                                          double *__restrict__ cscale) {
  __shared__ double sh[4];
  const int c = blockIdx.x;
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, fabs(R[i + (size_t)n * c]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    int e = 0;
    if (m > 0.0) frexp(m, &e);
    cscale[c] = m > 0.0 ? ldexp(1.0, e) : 1.0;
  }
}

// R / cscale in the B-operand layout: 16-bit column-major with kp = nkt*32
// rows (lane reads eight consecutive k); FP32 as [c][tk][l/16][j] with k =
// tk*16 + 4j + l/16 (lane reads four)
template <Lo L>
__global__ void mx_rhs_kernel(int n, int nkt, int nrhs, const double *__restrict__ R,
                              const double *__restrict__ cscale, typename Load4<L>::T *__restrict__ out) {
  constexpr int KT = mx_kt<L>();
  const size_t kp = (size_t)nkt * KT;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kp * nrhs) return;
  const int c = (int)(t / kp), p = (int)(t % kp);
  int k;
  if constexpr (L == Lo::F32) {
    const int tk = p / 16, r = p % 16, g = r / 4, j = r % 4;
    k = tk * 16 + 4 * j + g;
  } else {
    k = p;
  }
  const float x = k < n ? (float)(R[k + (size_t)n * c] / cscale[c]) : 0.f;
  out[t] = mx_conv<L>(x);
}

// part[ks][c][row] = sum over this slice's k tiles of (M/s)(row, k) (R/t)(k, c).
// Block: four wavefronts, one 16-row strip each; blockIdx.y: the k slice.
// Output register i of lane l: row 4*(l/16) + i of the strip, column l%16
// (the 16x16 FP32-accumulator layout); columns >= nrhs carry zeros.
template <Lo L>
__global__ void __launch_bounds__(256) mx_apply_kernel(int n, int nkt, int nrhs, int kchunk, int npad,
                                                       const typename Load4<L>::T *__restrict__ tiles,
                                                       const typename Load4<L>::T *__restrict__ rlow,
                                                       float *__restrict__ part) {
  constexpr int KT = mx_kt<L>(), PER = KT / 4;
  const int l = threadIdx.x & 63, ti = blockIdx.x * 4 + (threadIdx.x >> 6), ks = blockIdx.y;
  if (ti * 16 >= n) return;
  const int tk0 = ks * kchunk, tk1 = min(nkt, tk0 + kchunk);
  const int col = l & 15, g = l >> 4;
  const bool hasc = col < nrhs;
  mx_f4 acc = {0.f, 0.f, 0.f, 0.f};
  const size_t kp = (size_t)nkt * KT;
  auto load = [&](int tk, uint4 &a, uint4 &b) {
    a = *reinterpret_cast<const uint4 *>(tiles + ((size_t)(ti * nkt + tk) * 64 + l) * PER);
    b = uint4{0u, 0u, 0u, 0u};
    if (hasc) {
      const size_t off = L == Lo::F32 ? ((size_t)col * nkt + tk) * 16 + g * 4 : (size_t)col * kp + tk * KT + 8 * g;
      b = *reinterpret_cast<const uint4 *>(rlow + off);
    }
  };
  auto mma = [&](const uint4 &a, const uint4 &b) {
    if constexpr (L == Lo::F16) {
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(mx_h8, a), __builtin_bit_cast(mx_h8, b), acc,
                                                   0, 0, 0);
    } else if constexpr (L == Lo::BF16) {
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mx_b8, a), __builtin_bit_cast(mx_b8, b), acc,
                                                    0, 0, 0);
    } else {
      const mx_f4 af = __builtin_bit_cast(mx_f4, a), bf = __builtin_bit_cast(mx_f4, b);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[j], bf[j], acc, 0, 0, 0);
    }
  };
  // four tiles per step, their loads issued together
  int tk = tk0;
  for (; tk + 4 <= tk1; tk += 4) {
    uint4 a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) load(tk + q, a[q], b[q]);
#pragma unroll
    for (int q = 0; q < 4; ++q) mma(a[q], b[q]);
  }
  for (; tk < tk1; ++tk) {
    uint4 a, b;
    load(tk, a, b);
    mma(a, b);
  }
  if (!hasc) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = ti * 16 + 4 * g + i;
    part[((size_t)ks * nrhs + col) * npad + row] = acc[i];
  }
}

// X(i, c) += rowscale(i) * cscale(c) * sum over slices (in slice order)
__global__ void mx_update_kernel(int n, int nrhs, int ks_cnt, int npad, const float *__restrict__ part,
                                 const float *__restrict__ rscale, const double *__restrict__ cscale,
                                 double *__restrict__ X) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * nrhs) return;
  const int i = (int)(t % n), c = (int)(t / n);
  const double s = slice_sum(part + (size_t)c * npad + i, (size_t)nrhs * npad, ks_cnt);
  X[t] += s * (double)rscale[i] * cscale[c];
}

// part[ks][c][i] = sum over column slice ks of M(i, j) X(j, c), c < nrhs <= 4
// (FP64, column-major M read once for all right-hand sides)
__global__ void __launch_bounds__(kRowsPerBlock) gemm4_kernel(int n, int ks_cnt, int nrhs, const double *__restrict__ M,
                                                              const double *__restrict__ X, double *__restrict__ part) {
  const int i = blockIdx.x * kRowsPerBlock + threadIdx.x, ks = blockIdx.y;
  const int chunk = (n + ks_cnt - 1) / ks_cnt, j0 = ks * chunk, j1 = min(n, j0 + chunk);
  if (i >= n) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  int j = j0;
  // four columns' loads issued together, added in column order
  for (; j + 4 <= j1; j += 4) {
    double m[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) m[u] = M[i + (size_t)n * (j + u)];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (c < nrhs) acc[c] += m[u] * X[j + u + (size_t)n * c];
  }
  for (; j < j1; ++j) {
    const double m = M[i + (size_t)n * j];
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (c < nrhs) acc[c] += m * X[j + (size_t)n * c];
  }
  for (int c = 0; c < nrhs; ++c) part[((size_t)ks * nrhs + c) * n + i] = acc[c];
}

// out = base + sign * (sum of the slices), element-wise (out may alias base)
__global__ void combine4_kernel(int n, int nrhs, int ks_cnt, const double *__restrict__ part, const double *base,
                                double sign, double *out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * nrhs) return;
  const int i = (int)(t % n), c = (int)(t / n);
  const double s = slice_sum(part + (size_t)c * n + i, (size_t)nrhs * n, ks_cnt);
  out[t] = base[t] + sign * s;
}

struct MxBuf {
  Lo lo = Lo::F32;
  int nkt = 0;
  DBuf<uint16_t> t16;
  DBuf<float> t32, rscale;
};

}  // namespace

struct DenseMx {
  MxBuf m[3][3];  // [equation][FP32, BF16, FP16]
};

void dense_mx_free(DenseMx *p) { delete p; }

RefineResult dense_refine(DenseKin &d, const DevMesh &mesh, int e, SdPrec prec, int nrhs, const double *B,
                          double *X, double tol, int maxit, hipStream_t s) {
  const int n = d.nb;
  if (e < 0 || e > 2) throw std::runtime_error("equation must be 1..3");
  if (nrhs < 1 || nrhs > 4) throw std::runtime_error("nrhs must be 1..4");
  if (!d.mx) d.mx = new DenseMx();
  rocblas_handle h;
  if (rocblas_create_handle(&h) != rocblas_status_success) throw std::runtime_error("rocblas_create_handle failed");
  rocblas_set_stream(h, s);
  rocblas_set_pointer_mode(h, rocblas_pointer_mode_host);
  // the FP64 system matrix A_e (rebuilt; the resident copy holds its inverse)
  DBuf<double> A((size_t)n * n);
  hipLaunchKernelGGL(build_sys_kernel, dim3(ceil_div((size_t)n * n, 256)), dim3(256), 0, s, n, e, d.kode.p,
                     mesh.nx.p, mesh.ny.p, mesh.nz.p, mesh.gbn.p, d.mats.Dx.p, d.mats.Dy.p, d.mats.Dz.p, A.p);
  TT_LAUNCH_CHECK();
  const Lo lo = prec == SdPrec::F32 ? Lo::F32 : prec == SdPrec::BF16 ? Lo::BF16 : Lo::F16;
  const int KT = lo == Lo::F32 ? 16 : 32, nkt = (int)ceil_div(n, KT), nst = (int)ceil_div(n, 16), npad = nst * 16;
  MxBuf *mb = nullptr;
  if (prec != SdPrec::F64) {
    mb = &d.mx->m[e][(int)prec - 1];
    if (mb->nkt == 0) {  // tile-major copy of the inverse, made on first use
      const int ks = split_for(n, n);
      DBuf<float> pm((size_t)ks * n);
      mb->rscale.alloc(n);
      hipLaunchKernelGGL(rowmax_kernel, dim3(ceil_div(n, 256), ks), dim3(256), 0, s, n, n, ks, d.Ainv[e].p, pm.p);
      hipLaunchKernelGGL(rowscale_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, ks, pm.p, mb->rscale.p);
      const size_t tot = (size_t)nst * nkt * 64 * (KT / 4);
      if (lo == Lo::F32) {
        mb->t32.alloc(tot);
        hipLaunchKernelGGL(mx_tiles_kernel<Lo::F32>, dim3(nst, nkt), dim3(64), 0, s, n, nkt, d.Ainv[e].p,
                           mb->rscale.p, mb->t32.p);
      } else {
        mb->t16.alloc(tot);
        if (lo == Lo::F16)
          hipLaunchKernelGGL(mx_tiles_kernel<Lo::F16>, dim3(nst, nkt), dim3(64), 0, s, n, nkt, d.Ainv[e].p,
                             mb->rscale.p, mb->t16.p);
        else
          hipLaunchKernelGGL(mx_tiles_kernel<Lo::BF16>, dim3(nst, nkt), dim3(64), 0, s, n, nkt, d.Ainv[e].p,
                             mb->rscale.p, mb->t16.p);
      }
      TT_LAUNCH_CHECK();
      TT_HIP(hipStreamSynchronize(s));  // pm is released on return
      mb->lo = lo, mb->nkt = nkt;
    }
  }
  // slices of k for about 1,024 wavefronts (one per SIMD; four tiles' loads
  // in flight each): at 13,826 rows the 16-bit products stream 4.0 TB/s there
  // against 3.5-3.7 at 4,096 wavefronts
  const int ksl = std::max(1, std::min(nkt, (int)ceil_div(1024, nst)));
  const int kchunk = (int)ceil_div(nkt, ksl), ks_cnt = (int)ceil_div(nkt, kchunk);
  DBuf<double> R((size_t)n * nrhs), cs(nrhs);
  // FP64 products (residual; the FP64 correction) in column slices: about
  // four wavefronts per SIMD, four columns' loads in flight each
  const int ks64 = std::max(1, std::min(n / 64, (int)ceil_div(4096, ceil_div(n, kRowsPerBlock))));
  DBuf<double> p64((size_t)ks64 * nrhs * n);
  DBuf<float> part((size_t)ks_cnt * nrhs * npad);
  DBuf<uint16_t> r16;
  DBuf<float> r32;
  if (lo == Lo::F32)
    r32.alloc((size_t)nkt * KT * nrhs);
  else
    r16.alloc((size_t)nkt * KT * nrhs);
  TT_HIP(hipMemsetAsync(X, 0, sizeof(double) * n * nrhs, s));
  TT_HIP(hipMemcpyAsync(R.p, B, sizeof(double) * n * nrhs, hipMemcpyDeviceToDevice, s));
  double nb2 = 0.0;
  if (rocblas_dnrm2(h, n * nrhs, B, 1, &nb2) != rocblas_status_success) throw std::runtime_error("rocblas_dnrm2");
  RefineResult res;
  hipEvent_t ev[4];
  for (auto &x : ev) TT_HIP(hipEventCreate(&x));
  TT_HIP(hipEventRecord(ev[0], s));
  double apply_ms = 0.0;
  int applies = 0;
  for (int it = 0; it < maxit; ++it) {
    // correction X += M R
    TT_HIP(hipEventRecord(ev[2], s));
    if (prec == SdPrec::F64) {
      hipLaunchKernelGGL(gemm4_kernel, dim3(ceil_div(n, kRowsPerBlock), ks64), dim3(kRowsPerBlock), 0, s, n, ks64,
                         nrhs, d.Ainv[e].p, R.p, p64.p);
      hipLaunchKernelGGL(combine4_kernel, dim3(ceil_div((size_t)n * nrhs, 256)), dim3(256), 0, s, n, nrhs, ks64,
                         p64.p, X, 1.0, X);
      TT_LAUNCH_CHECK();
    } else {
      hipLaunchKernelGGL(mx_colscale_kernel, dim3(nrhs), dim3(256), 0, s, n, R.p, cs.p);
      const size_t nr = (size_t)nkt * KT * nrhs;
      const dim3 grid(ceil_div(nst, 4), ks_cnt);
      if (lo == Lo::F32) {
        hipLaunchKernelGGL(mx_rhs_kernel<Lo::F32>, dim3(ceil_div(nr, 256)), dim3(256), 0, s, n, nkt, nrhs, R.p, cs.p,
                           r32.p);
        hipLaunchKernelGGL(mx_apply_kernel<Lo::F32>, grid, dim3(256), 0, s, n, nkt, nrhs, kchunk, npad, mb->t32.p,
                           r32.p, part.p);
      } else if (lo == Lo::F16) {
        hipLaunchKernelGGL(mx_rhs_kernel<Lo::F16>, dim3(ceil_div(nr, 256)), dim3(256), 0, s, n, nkt, nrhs, R.p, cs.p,
                           r16.p);
        hipLaunchKernelGGL(mx_apply_kernel<Lo::F16>, grid, dim3(256), 0, s, n, nkt, nrhs, kchunk, npad, mb->t16.p,
                           r16.p, part.p);
      } else {
        hipLaunchKernelGGL(mx_rhs_kernel<Lo::BF16>, dim3(ceil_div(nr, 256)), dim3(256), 0, s, n, nkt, nrhs, R.p, cs.p,
                           r16.p);
        hipLaunchKernelGGL(mx_apply_kernel<Lo::BF16>, grid, dim3(256), 0, s, n, nkt, nrhs, kchunk, npad, mb->t16.p,
                           r16.p, part.p);
      }
      hipLaunchKernelGGL(mx_update_kernel, dim3(ceil_div((size_t)n * nrhs, 256)), dim3(256), 0, s, n, nrhs, ks_cnt,
                         npad, part.p, mb->rscale.p, cs.p, X);
      TT_LAUNCH_CHECK();
    }
    TT_HIP(hipEventRecord(ev[3], s));
    // FP64 residual R = B - A X and its norm
    hipLaunchKernelGGL(gemm4_kernel, dim3(ceil_div(n, kRowsPerBlock), ks64), dim3(kRowsPerBlock), 0, s, n, ks64, nrhs,
                       A.p, X, p64.p);
    hipLaunchKernelGGL(combine4_kernel, dim3(ceil_div((size_t)n * nrhs, 256)), dim3(256), 0, s, n, nrhs, ks64, p64.p,
                       B, -1.0, R.p);
    TT_LAUNCH_CHECK();
    double nr2 = 0.0;
    if (rocblas_dnrm2(h, n * nrhs, R.p, 1, &nr2) != rocblas_status_success) throw std::runtime_error("rocblas_dnrm2");
    float ms = 0.f;
    TT_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
    apply_ms += ms, ++applies;
    res.iters = it + 1;
    res.relres = nb2 > 0.0 ? nr2 / nb2 : 0.0;
    if (!(res.relres > tol)) break;  // converged (or not finite: reported as is)
  }
  TT_HIP(hipEventRecord(ev[1], s));
  TT_HIP(hipEventSynchronize(ev[1]));
  float tot = 0.f;
  TT_HIP(hipEventElapsedTime(&tot, ev[0], ev[1]));
  res.ms = tot;
  res.ms_apply = applies ? apply_ms / applies : 0.0;
  for (auto &x : ev) (void)hipEventDestroy(x);
  rocblas_destroy_handle(h);
  return res;
}

}  // namespace tt
