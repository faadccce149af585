// tt_lsqr_batch.hip -- many independent least-squares systems solved in one
// persistent launch: slvlsqr2 (lib_Solve_v11.f:4401-4500) with the column
// scaling of prelsqr2 (:4501-4535) and the LSQR iteration (:3128-3887,
// damp = 0) for every system of a batch.
//
// Each workgroup owns whole systems and runs their iterations to the end with
// workgroup barriers only (no host round trip, no grid synchronisation): the
// systems are taken in order of decreasing size, workgroup g solving systems
// g, g + G, g + 2G, ...  Products use 16 lanes per CSR row and per CSC column;
// the vectors of all systems are slices of batch-wide arrays.  The scalar
// recurrences are evaluated by every thread from the same reduced sums, so a
// solve is bit-for-bit repeatable.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "tritok_bem.h"
#include "tt_hip.h"

namespace tt {
namespace {

constexpr int kNT = 1024;  // threads per workgroup
constexpr int kLanes = 16;  // lanes per row / column

struct SysDesc {
  int m, n;
  long rp;   // offset of the system's m+1 row pointers (0-based entries, system-local)
  long cp;   // offset of its n+1 column pointers
  long e0;   // offset of its entries (CSR and CSC)
  long mo;   // offset of its rows in the batch vectors
  long no;   // offset of its unknowns
};

__device__ __forceinline__ double d2n(double a, double b) {
  const double scale = fabs(a) + fabs(b);
  if (scale == 0.0) return 0.0;
  const double sa = a / scale, sb = b / scale;
  return scale * sqrt(sa * sa + sb * sb);
}

__device__ __forceinline__ double lane_group_sum(double v) {
#pragma unroll
  for (int o = kLanes / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ double wg_sum(double v, double *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < kNT / 64; ++i) r += sh[i];
  __syncthreads();
  return r;
}

// u <- A v + c u (c may be 0), returns ||u||^2 (rows over groups of 16 lanes)
__device__ double rows_apply(const SysDesc &d, const int *__restrict__ rp, const int *__restrict__ ja,
                             const double *__restrict__ a, const double *__restrict__ v, double c,
                             double *__restrict__ u, double *sh) {
  const int g = threadIdx.x / kLanes, lane = threadIdx.x % kLanes, G = kNT / kLanes;
  double acc = 0.0;
  for (int r = g; r < d.m; r += G) {
    double s = 0.0;
    for (int e = rp[r] + lane; e < rp[r + 1]; e += kLanes) s += a[e] * v[ja[e]];
    s = lane_group_sum(s);
    const double un = c != 0.0 ? s + c * u[r] : s;
    if (lane == 0) {
      u[r] = un;
      acc += un * un;
    }
  }
  return wg_sum(acc, sh);
}

// v <- A^T u + c v, returns ||v||^2 (columns over groups of 16 lanes, CSC)
__device__ double cols_apply(const SysDesc &d, const int *__restrict__ cp, const int *__restrict__ ri,
                             const double *__restrict__ at, const double *__restrict__ u, double c,
                             double *__restrict__ v, double *sh) {
  const int g = threadIdx.x / kLanes, lane = threadIdx.x % kLanes, G = kNT / kLanes;
  double acc = 0.0;
  for (int j = g; j < d.n; j += G) {
    double s = 0.0;
    for (int e = cp[j] + lane; e < cp[j + 1]; e += kLanes) s += at[e] * u[ri[e]];
    s = lane_group_sum(s);
    const double vn = c != 0.0 ? s + c * v[j] : s;
    if (lane == 0) {
      v[j] = vn;
      acc += vn * vn;
    }
  }
  return wg_sum(acc, sh);
}

__device__ void scale_vec(int len, double f, double *__restrict__ x) {
  for (int i = threadIdx.x; i < len; i += kNT) x[i] *= f;
  __syncthreads();
}

__global__ void __launch_bounds__(kNT) lsqr_batch_kernel(int nsys, const int *__restrict__ order,
                                                         const SysDesc *__restrict__ sys, const int *__restrict__ rpp,
                                                         const int *__restrict__ jap, const double *__restrict__ ap,
                                                         const int *__restrict__ cpp, const int *__restrict__ rip,
                                                         const double *__restrict__ atp, const double *__restrict__ qp,
                                                         const double *__restrict__ bp, double *__restrict__ xp,
                                                         double *__restrict__ up, double *__restrict__ vp,
                                                         double *__restrict__ wp, double *__restrict__ dxp, int maxit,
                                                         double eps, int *__restrict__ nit_out,
                                                         int *__restrict__ istop_out) {
  __shared__ double sh[kNT / 64];
  for (int oi = blockIdx.x; oi < nsys; oi += gridDim.x) {
    const int s = order[oi];
    const SysDesc d = sys[s];
    const int *rp = rpp + d.rp, *ja = jap + d.e0, *cp = cpp + d.cp, *ri = rip + d.e0;
    const double *a = ap + d.e0, *at = atp + d.e0, *q = qp + d.no, *b = bp + d.mo;
    double *x = xp + d.no, *u = up + d.mo, *v = vp + d.no, *w = wp + d.no, *dx = dxp + d.no;
    const int itnlim = maxit > 0 ? maxit : 10 * (d.m + d.n + 50);
    // hot start: u = b - A (x0 / q)
    for (int j = threadIdx.x; j < d.n; j += kNT) v[j] = x[j] / q[j];
    __syncthreads();
    rows_apply(d, rp, ja, a, v, 0.0, u, sh);
    double bb = 0.0, rr = 0.0;
    for (int i = threadIdx.x; i < d.m; i += kNT) {
      const double r = b[i] - u[i];
      u[i] = r;
      rr += r * r;
      bb += b[i] * b[i];
    }
    const double nrm_r = sqrt(wg_sum(rr, sh)), nrm_b = sqrt(wg_sum(bb, sh));
    const double atol = eps, btol = eps * nrm_b / nrm_r, ctol = 1.0 / 1.0e8;
    // LSQR on the correction (x-space scaled by q)
    for (int j = threadIdx.x; j < d.n; j += kNT) v[j] = 0.0, dx[j] = 0.0;
    __syncthreads();
    double alpha = 0.0, beta = nrm_r;
    int itn = 0, istop = 0;
    if (beta > 0.0) {
      scale_vec(d.m, 1.0 / beta, u);
      alpha = sqrt(cols_apply(d, cp, ri, at, u, 0.0, v, sh));
    }
    if (alpha > 0.0) {
      scale_vec(d.n, 1.0 / alpha, v);
      for (int j = threadIdx.x; j < d.n; j += kNT) w[j] = v[j];
      __syncthreads();
    }
    if (alpha * beta != 0.0) {
      double anorm = 0.0, acond = 0.0, dnorm = 0.0, res2 = 0.0, xnorm1 = 0.0, cs2 = -1.0, sn2 = 0.0, z = 0.0;
      double rhobar = alpha, phibar = beta;
      const double bnorm = beta;
      for (;;) {
        ++itn;
        beta = sqrt(rows_apply(d, rp, ja, a, v, -alpha, u, sh));
        const double temp = d2n(d2n(alpha, beta), 0.0);
        anorm = d2n(anorm, temp);
        if (beta > 0.0) {
          scale_vec(d.m, 1.0 / beta, u);
          alpha = sqrt(cols_apply(d, cp, ri, at, u, -beta, v, sh));
          if (alpha > 0.0) scale_vec(d.n, 1.0 / alpha, v);
        }
        const double rhbar1 = rhobar;
        const double rho = d2n(rhbar1, beta);
        const double cs = rhbar1 / rho, sn = beta / rho;
        const double theta = sn * alpha;
        rhobar = -cs * alpha;
        const double phi = cs * phibar;
        phibar = sn * phibar;
        const double tau = sn * phi;
        const double t1 = phi / rho, t2 = -theta / rho, t3 = 1.0 / rho;
        double dk = 0.0;
        for (int j = threadIdx.x; j < d.n; j += kNT) {
          const double t = w[j];
          dx[j] = t1 * t + dx[j];
          w[j] = t2 * t + v[j];
          const double tt = t3 * t;
          dk += tt * tt;
        }
        const double dknorm = sqrt(wg_sum(dk, sh));
        dnorm = d2n(dnorm, dknorm);
        const double delta = sn2 * rho, gambar = -cs2 * rho;
        const double rhs = phi - delta * z;
        const double zbar = rhs / gambar;
        const double xnorm = d2n(xnorm1, zbar);
        const double gamma = d2n(gambar, theta);
        cs2 = gambar / gamma;
        sn2 = theta / gamma;
        z = rhs / gamma;
        xnorm1 = d2n(xnorm1, z);
        acond = anorm * dnorm;
        res2 = d2n(res2, 0.0);
        const double rnorm = d2n(res2, phibar);
        const double arnorm = alpha * fabs(tau);
        const double test1 = rnorm / bnorm;
        const double test2 = rnorm > 0.0 ? arnorm / (anorm * rnorm) : 0.0;
        const double test3 = 1.0 / acond;
        const double tt1 = 1.0 + test1 / (1.0 + anorm * xnorm / bnorm);
        const double rtol = btol + atol * anorm * xnorm / bnorm;
        if (itn >= itnlim) istop = 5;
        if (1.0 + test3 <= 1.0) istop = 4;
        if (1.0 + test2 <= 1.0) istop = 2;
        if (tt1 <= 1.0) istop = 1;
        if (test3 <= ctol) istop = 4;
        if (test2 <= atol) istop = 2;
        if (test1 <= rtol) istop = 1;
        if (istop != 0) break;
      }
    }
    // x = q dx + x0
    for (int j = threadIdx.x; j < d.n; j += kNT) x[j] = q[j] * dx[j] + x[j];
    if (threadIdx.x == 0) nit_out[s] = itn, istop_out[s] = istop;
    __syncthreads();
  }
}

}  // namespace
}  // namespace tt

using namespace tt;

struct tt_lsqr_batch {
  int nsys = 0;
  long mtot = 0, ntot = 0, etot = 0;
  std::vector<SysDesc> desc;
  DBuf<SysDesc> dsys;
  DBuf<int> order, rp, ja, cp, ri;
  DBuf<double> a, at, q, b, x, u, v, w, dx;
  DBuf<int> nit, istop;
  hipStream_t s = nullptr;
  ~tt_lsqr_batch() {
    if (s) (void)hipStreamDestroy(s);
  }
};

extern "C" int tt_lsqr_batch_setup(int32_t nsys, const int32_t *m, const int32_t *n, const int64_t *nnz,
                                   const int32_t *ia, const int32_t *ja, double *a, double *q, int32_t scale_host,
                                   tt_lsqr_batch **out) {
  return guarded("tt_lsqr_batch_setup", [&] {
    if (nsys < 1 || !m || !n || !nnz || !ia || !ja || !a || !q || !out) throw std::runtime_error("bad argument");
    auto h = std::make_unique<tt_lsqr_batch>();
    TT_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
    h->nsys = nsys;
    h->desc.resize(nsys);
    long mo = 0, no = 0, eo = 0, rpo = 0, cpo = 0;
    for (int k = 0; k < nsys; ++k) {
      if (m[k] < 1 || n[k] < 1 || nnz[k] < 1) throw std::runtime_error("empty system");
      h->desc[k] = SysDesc{m[k], n[k], rpo, cpo, eo, mo, no};
      mo += m[k], no += n[k], eo += nnz[k], rpo += m[k] + 1, cpo += n[k] + 1;
    }
    h->mtot = mo, h->ntot = no, h->etot = eo;
    // 0-based system-local CSR and its CSC copy (entries of a column in row order)
    std::vector<int> rp(rpo), jz(eo), cp(cpo), ri(eo);
    std::vector<double> at(eo);
    long ia_off = 0;
    for (int k = 0; k < nsys; ++k) {
      const SysDesc &d = h->desc[k];
      const int32_t *iak = ia + ia_off;
      if (iak[0] != 1 || iak[d.m] - 1 != nnz[k]) throw std::runtime_error("row pointers do not describe nnz entries");
      for (int r = 0; r <= d.m; ++r) rp[d.rp + r] = iak[r] - 1;
      std::vector<int> cnt(d.n + 1, 0);
      for (long e = 0; e < nnz[k]; ++e) {
        const int c = ja[d.e0 + e] - 1;
        if (c < 0 || c >= d.n) throw std::runtime_error("column index outside 1..n");
        jz[d.e0 + e] = c;
        ++cnt[c + 1];
      }
      for (int c = 0; c < d.n; ++c) cnt[c + 1] += cnt[c];
      for (int c = 0; c <= d.n; ++c) cp[d.cp + c] = cnt[c];
      for (int r = 0; r < d.m; ++r)
        for (int e = rp[d.rp + r]; e < rp[d.rp + r + 1]; ++e) {
          const int c = jz[d.e0 + e];
          const int pos = cnt[c]++;
          ri[d.e0 + pos] = r;
          at[d.e0 + pos] = a[d.e0 + e];
        }
      ia_off += d.m + 1;
    }
    // prelsqr2: q_j = 1/||A(:,j)|| summed in row order, A <- A diag(q)
    std::vector<double> hq(no, 0.0), as(eo);
    for (int k = 0; k < nsys; ++k) {
      const SysDesc &d = h->desc[k];
      for (int r = 0; r < d.m; ++r)
        for (int e = rp[d.rp + r]; e < rp[d.rp + r + 1]; ++e) {
          const double v = a[d.e0 + e];
          hq[d.no + jz[d.e0 + e]] += v * v;
        }
      for (int j = 0; j < d.n; ++j) hq[d.no + j] = 1.0 / std::sqrt(hq[d.no + j]);
      for (long e = 0; e < nnz[k]; ++e) as[d.e0 + e] = a[d.e0 + e] * hq[d.no + jz[d.e0 + e]];
      std::vector<int> fill(cp.begin() + d.cp, cp.begin() + d.cp + d.n);
      for (int r = 0; r < d.m; ++r)
        for (int e = rp[d.rp + r]; e < rp[d.rp + r + 1]; ++e) at[d.e0 + fill[jz[d.e0 + e]]++] = as[d.e0 + e];
    }
    h->dsys.upload(h->desc, h->s);
    h->rp.upload(rp, h->s), h->ja.upload(jz, h->s), h->cp.upload(cp, h->s), h->ri.upload(ri, h->s);
    h->a.upload(as, h->s);
    h->at.upload(at, h->s);
    h->q.upload(hq, h->s);
    std::memcpy(q, hq.data(), sizeof(double) * no);
    if (scale_host) std::memcpy(a, as.data(), sizeof(double) * eo);
    // largest systems first
    std::vector<int> ord(nsys);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return nnz[x] > nnz[y]; });
    h->order.upload(ord, h->s);
    h->b.alloc(mo), h->u.alloc(mo), h->x.alloc(no), h->v.alloc(no), h->w.alloc(no), h->dx.alloc(no);
    h->nit.alloc(nsys), h->istop.alloc(nsys);
    TT_HIP(hipStreamSynchronize(h->s));
    *out = h.release();
    return 0;
  });
}

extern "C" int tt_lsqr_batch_solve(tt_lsqr_batch *h, const double *b, double *x, int32_t maxit, float eps4,
                                   int32_t *nit, int32_t *istop, double *ms) {
  return guarded("tt_lsqr_batch_solve", [&] {
    if (!h || !b || !x) throw std::runtime_error("NULL argument");
    h->b.upload(b, h->mtot, h->s);
    h->x.upload(x, h->ntot, h->s);
    int dev = 0, cus = 0;
    TT_HIP(hipGetDevice(&dev));
    TT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = std::min(h->nsys, 2 * cus);
    hipEvent_t e0, e1;
    TT_HIP(hipEventCreate(&e0));
    TT_HIP(hipEventCreate(&e1));
    TT_HIP(hipEventRecord(e0, h->s));
    hipLaunchKernelGGL(lsqr_batch_kernel, dim3(grid), dim3(kNT), 0, h->s, h->nsys, h->order.p, h->dsys.p, h->rp.p,
                       h->ja.p, h->a.p, h->cp.p, h->ri.p, h->at.p, h->q.p, h->b.p, h->x.p, h->u.p, h->v.p, h->w.p,
                       h->dx.p, maxit, (double)eps4, h->nit.p, h->istop.p);
    TT_LAUNCH_CHECK();
    TT_HIP(hipEventRecord(e1, h->s));
    h->x.download(x, h->ntot, h->s);
    if (nit) h->nit.download(nit, h->nsys, h->s);
    if (istop) h->istop.download(istop, h->nsys, h->s);
    TT_HIP(hipStreamSynchronize(h->s));
    float t = 0.f;
    TT_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0), (void)hipEventDestroy(e1);
    if (ms) *ms = t;
    return 0;
  });
}

extern "C" void tt_lsqr_batch_free(tt_lsqr_batch *h) { delete h; }
