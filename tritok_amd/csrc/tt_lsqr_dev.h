// tt_lsqr_dev.h -- device pieces shared by the LSQR solvers (block and CSR
// storage): deterministic reductions and the scalar recurrences of the
// reference LSQR (lib_Solve_v11.f:3283-3420, damp = 0), arranged so that the
// kernel of iteration k finishes iteration k-1 (see tt_lsqr.hip).
#pragma once
#include "tt_lsqr.h"

namespace tt {
namespace lsqr_dev {

__device__ __forceinline__ double d2norm(double a, double b) {
  double scale = fabs(a) + fabs(b);
  if (scale == 0.0) return 0.0;
  double sa = a / scale, sb = b / scale;
  return scale * sqrt(sa * sa + sb * sb);
}

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b on XCD
// b % 8), and each XCD has its own L2.  xcd_logical numbers the workgroups of
// [lo, hi) so that those of one XCD take consecutive indices: a kernel that
// hands out consecutive cells or unknowns by this index keeps each XCD's
// gathers within one region of the vector it gathers from.  A bijection of
// [lo, hi) onto [0, hi - lo) for any dispatch order; only the locality assumes
// the round-robin.
constexpr int kXcd = 8;
__device__ __forceinline__ int xcd_logical(int b, int lo, int hi) {
  const int N = hi - lo, q = N / kXcd, rem = N % kXcd, p = b - lo;
  const int dx = p % kXcd;  // position of b's XCD in the order starting at lo's
  return dx * q + min(dx, rem) + p / kXcd;
}

// sum over the 64 lanes of a wavefront (butterfly; every lane gets the total)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block of NT threads (multiple of 64): wave sums combined in wave order
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) r += sh[i];
  __syncthreads();
  return r;
}

// sum of p[0..n) with lane i taking i, i+NT, ...; 16 loads per lane in
// flight at a time; deterministic
template <int NT>
__device__ __forceinline__ double reduce_partials(const double *__restrict__ p, int n, double *sh) {
  double acc = 0.0;
  for (int base = 0; base < n; base += 16 * NT) {
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int i = base + q * NT + (int)threadIdx.x;
      v[q] = i < n ? p[i] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) acc += v[q];
  }
  return block_sum<NT>(acc, sh);
}

// Scalar recurrences run by the cell kernel of iteration k (k >= 1) from the
// state after iteration k-1's cell kernel (in) and the unknown kernel's sums:
// nvs = ||v_k||^2 (unnormalised), dks = ||t3 w||^2 of iteration k-2's update.
__device__ __forceinline__ void scalar_update(int k, const LsqrState &in, double nvs, double dks, LsqrState &S) {
  S = in;
  if (in.done) return;
  if (k == 1) {
    // initialisation: alpha_1 and the exit when A^T b = 0 (phibar, bnorm and
    // rnorm were set to beta_1 by the unknown kernel of iteration 0)
    const double a1 = sqrt(nvs);
    S.alpha = a1;
    S.rhobar = a1;
    S.arnorm = a1 * in.beta;
    if (S.arnorm == 0.0) S.done = 1, S.istop = 0, S.itn = 0;
    return;
  }
  const double alpha = in.alpha, beta = in.beta;  // alpha_{k-1}, beta_k
  if (k >= 3) {
    // finish iteration k-2 (whose x/w update the unknown kernel of iteration k-1 applied)
    const double dknorm = sqrt(dks);
    S.dnorm = d2norm(S.dnorm, dknorm);
    const double rho = S.pending_rho, phi = S.pending_phi, theta = S.pending_theta, tau = S.pending_tau;
    const double delta = S.sn2 * rho, gambar = -S.cs2 * rho;
    const double rhs = phi - delta * S.z;
    const double zbar = rhs / gambar;
    S.xnorm = d2norm(S.xnorm1, zbar);
    const double gamma = d2norm(gambar, theta);
    S.cs2 = gambar / gamma;
    S.sn2 = theta / gamma;
    S.z = rhs / gamma;
    S.xnorm1 = d2norm(S.xnorm1, S.z);
    S.acond = S.anorm * S.dnorm;
    S.res2 = d2norm(S.res2, 0.0);
    S.rnorm = d2norm(S.res2, S.phibar);
    S.arnorm = alpha * fabs(tau);
    const double test1 = S.rnorm / S.bnorm;
    double test2 = 0.0;
    if (S.rnorm > 0.0) test2 = S.arnorm / (S.anorm * S.rnorm);
    const double test3 = 1.0 / S.acond;
    double tt1 = test1 / (1.0 + S.anorm * S.xnorm / S.bnorm);
    const double rtol = S.btol + S.atol * S.anorm * S.xnorm / S.bnorm;
    const double tt3 = 1.0 + test3, tt2 = 1.0 + test2;
    tt1 = 1.0 + tt1;
    int istop = 0;
    const int itn = k - 2;
    if (itn >= S.itnlim) istop = 5;
    if (tt3 <= 1.0) istop = 4;
    if (tt2 <= 1.0) istop = 2;
    if (tt1 <= 1.0) istop = 1;
    if (test3 <= S.ctol) istop = 4;
    if (test2 <= S.atol) istop = 2;
    if (test1 <= rtol) istop = 1;
    if (istop != 0) {
      S.done = 1, S.istop = istop, S.itn = itn;
      return;
    }
  }
  // rest of iteration k-1: anorm with alpha_{k-1}, beta_k; alpha_k; plane rotation
  const double anew = beta > 0.0 ? sqrt(nvs) : alpha;  // the reference keeps alpha when beta = 0
  double temp = d2norm(alpha, beta);
  temp = d2norm(temp, 0.0);
  S.anorm = d2norm(S.anorm, temp);
  const double rhbar1 = S.rhobar;
  const double rho = d2norm(rhbar1, beta);
  const double cs = rhbar1 / rho, sn = beta / rho;
  const double theta = sn * anew;
  S.rhobar = -cs * anew;
  const double phi = cs * S.phibar;
  S.phibar = sn * S.phibar;
  const double tau = sn * phi;
  S.t1 = phi / rho;
  S.t2 = -theta / rho;
  S.t3 = 1.0 / rho;
  S.pending_rho = rho, S.pending_phi = phi, S.pending_theta = theta, S.pending_tau = tau;
  S.alpha = anew;
  S.itn = k - 1;
}

}  // namespace lsqr_dev
}  // namespace tt
