// tt_lsqr_batch.hip -- many independent least-squares systems solved in one
// persistent launch: slvlsqr2 (lib_Solve_v11.f:4401-4500) with the column
// scaling of prelsqr2 (:4501-4535) and the LSQR iteration (:3128-3887,
// damp = 0) for every system of a batch.
//
// The launch is sized to the workgroups resident on the device at once.  The
// host assigns every system to a gang of W workgroups (W = 1 for most; large
// systems get wider gangs so that all workgroups finish together): systems in
// order of decreasing work, each to the W workgroups that fall free first
// (list scheduling).  Every workgroup runs its assignments in that global
// order, so the members of a gang meet at its barriers without any cycle of
// waits.  Inside a gang, rows, columns and vector entries are interleaved over
// the members; a gang barrier is an agent-scope release/acquire counter (the
// members are co-resident by construction) with a bounded wait that flags an
// error instead of hanging; gang sums add the members' partials in member
// order, so every member obtains the same scalars and a solve is bit-for-bit
// repeatable for a given schedule.  Width 1 is a plain workgroup with
// workgroup barriers only.  Products use 16 lanes per CSR row and per CSC
// column; the vectors of all systems are slices of batch-wide arrays.
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

// one assignment of a workgroup: system, gang slot, this member's rank and the
// gang's width
struct Assign {
  int sys, gang, rank, width;
};

constexpr long kSpinLimit = 1L << 23;  // bounded wait at a gang barrier (seconds)

// The gang a workgroup works in for the current system.
struct Gang {
  int rank = 0, width = 1;
  unsigned *cnt = nullptr;  // arrival counter (zero at launch)
  double *part = nullptr;   // 2 x 64 partial sums
  unsigned epoch = 0;       // barriers passed
  int red = 0;              // reductions done (selects the partial buffer)
  int *err = nullptr;

  __device__ void sync() {
    if (width == 1) {
      __syncthreads();
      return;
    }
    // every thread's writes to the gang's vectors made visible device-wide
    // (the members may sit on different XCDs), then one arrival per workgroup
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    ++epoch;
    if (threadIdx.x == 0) {
      __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned target = epoch * (unsigned)width;
      long spins = 0;
      while (__hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > kSpinLimit) {
          *err = 1;
          break;
        }
      }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // and the other members' writes seen
  }
  // sum over the gang's workgroups, the same value in every member
  __device__ double sum(double v, double *sh) {
    const double w = wg_sum(v, sh);
    if (width == 1) return w;
    double *buf = part + 64 * (red++ & 1);
    if (threadIdx.x == 0) buf[rank] = w;
    sync();
    double r = 0.0;
    for (int i = 0; i < width; ++i) r += buf[i];
    return r;
  }
};

// u <- A v + c u (c may be 0) on this member's rows, returns ||u||^2 over the
// gang; VT = the stored values' type (FP32 values are widened, products and
// sums in FP64)
template <typename VT>
__device__ double rows_apply(Gang &gg, const SysDesc &d, const int *__restrict__ rp, const int *__restrict__ ja,
                             const VT *__restrict__ a, const double *__restrict__ v, double c,
                             double *__restrict__ u, double *sh) {
  const int G = kNT / kLanes, g = gg.rank * G + threadIdx.x / kLanes, lane = threadIdx.x % kLanes;
  const int stride = gg.width * G;
  double acc = 0.0;
  for (int r = g; r < d.m; r += stride) {
    double s = 0.0;
    for (int e = rp[r] + lane; e < rp[r + 1]; e += kLanes) s += (double)a[e] * v[ja[e]];
    s = lane_group_sum(s);
    const double un = c != 0.0 ? s + c * u[r] : s;
    if (lane == 0) {
      u[r] = un;
      acc += un * un;
    }
  }
  return gg.sum(acc, sh);
}

// v <- A^T u + c v on this member's columns (CSC), returns ||v||^2 over the gang
template <typename VT>
__device__ double cols_apply(Gang &gg, const SysDesc &d, const int *__restrict__ cp, const int *__restrict__ ri,
                             const VT *__restrict__ at, const double *__restrict__ u, double c,
                             double *__restrict__ v, double *sh) {
  const int G = kNT / kLanes, g = gg.rank * G + threadIdx.x / kLanes, lane = threadIdx.x % kLanes;
  const int stride = gg.width * G;
  double acc = 0.0;
  for (int j = g; j < d.n; j += stride) {
    double s = 0.0;
    for (int e = cp[j] + lane; e < cp[j + 1]; e += kLanes) s += (double)at[e] * u[ri[e]];
    s = lane_group_sum(s);
    const double vn = c != 0.0 ? s + c * v[j] : s;
    if (lane == 0) {
      v[j] = vn;
      acc += vn * vn;
    }
  }
  return gg.sum(acc, sh);
}

// Element-wise work over the gang: entry i by thread i mod (width * kNT) of
// member (i / kNT) mod width.  Every phase that writes entries another thread
// reads ends at a barrier; an entry scaled and then used within one phase is
// handled by the same thread in both.
#define TT_OWN(gg, i, len) \
  for (int i = (gg).rank * kNT + (int)threadIdx.x; i < (len); i += (gg).width * kNT)

__device__ void scale_own(Gang &gg, int len, double f, double *__restrict__ x) {
  TT_OWN(gg, i, len) x[i] *= f;
}

// VT: type of the values the LSQR products read (double, or float copies of
// the scaled values); the hot-start residual b - A x0 always uses the FP64
// values (ap), and every vector, scalar and sum is FP64.
template <typename VT>
__global__ void __launch_bounds__(kNT) lsqr_batch_kernel(const int *__restrict__ aptr, const Assign *__restrict__ asg,
                                                         const SysDesc *__restrict__ sys, const int *__restrict__ rpp,
                                                         const int *__restrict__ jap, const double *__restrict__ ap,
                                                         const VT *__restrict__ alo, const int *__restrict__ cpp,
                                                         const int *__restrict__ rip, const VT *__restrict__ atlo,
                                                         const double *__restrict__ qp,
                                                         const double *__restrict__ bp, double *__restrict__ xp,
                                                         double *__restrict__ up, double *__restrict__ vp,
                                                         double *__restrict__ wp, double *__restrict__ dxp, int maxit,
                                                         double eps, unsigned *__restrict__ gcnt,
                                                         double *__restrict__ gpart, int *__restrict__ err,
                                                         int *__restrict__ nit_out, int *__restrict__ istop_out) {
  __shared__ double sh[kNT / 64];
  for (int ai = aptr[blockIdx.x]; ai < aptr[blockIdx.x + 1]; ++ai) {
    const Assign as = asg[ai];
    Gang gg;
    gg.rank = as.rank, gg.width = as.width, gg.err = err;
    gg.cnt = gcnt + as.gang, gg.part = gpart + (size_t)as.gang * 128;
    const int s = as.sys;
    const SysDesc d = sys[s];
    const int *rp = rpp + d.rp, *ja = jap + d.e0, *cp = cpp + d.cp, *ri = rip + d.e0;
    const double *a = ap + d.e0, *q = qp + d.no, *b = bp + d.mo;
    const VT *al = alo + d.e0, *at = atlo + d.e0;
    double *x = xp + d.no, *u = up + d.mo, *v = vp + d.no, *w = wp + d.no, *dx = dxp + d.no;
    const int itnlim = maxit > 0 ? maxit : 10 * (d.m + d.n + 50);
    // hot start: u = b - A (x0 / q)
    TT_OWN(gg, j, d.n) v[j] = x[j] / q[j];
    gg.sync();
    rows_apply(gg, d, rp, ja, a, v, 0.0, u, sh);
    double bb = 0.0, rr = 0.0;
    TT_OWN(gg, i, d.m) {
      const double r = b[i] - u[i];
      u[i] = r;
      rr += r * r;
      bb += b[i] * b[i];
    }
    const double nrm_r = sqrt(gg.sum(rr, sh)), nrm_b = sqrt(gg.sum(bb, sh));
    const double atol = eps, btol = eps * nrm_b / nrm_r, ctol = 1.0 / 1.0e8;
    // LSQR on the correction (x-space scaled by q)
    TT_OWN(gg, j, d.n) v[j] = 0.0, dx[j] = 0.0;
    double alpha = 0.0, beta = nrm_r;
    int itn = 0, istop = 0;
    if (beta > 0.0) {
      scale_own(gg, d.m, 1.0 / beta, u);
      gg.sync();
      alpha = sqrt(cols_apply(gg, d, cp, ri, at, u, 0.0, v, sh));
    }
    if (alpha > 0.0) {
      scale_own(gg, d.n, 1.0 / alpha, v);
      TT_OWN(gg, j, d.n) w[j] = v[j];
    }
    gg.sync();
    if (alpha * beta != 0.0) {
      double anorm = 0.0, acond = 0.0, dnorm = 0.0, res2 = 0.0, xnorm1 = 0.0, cs2 = -1.0, sn2 = 0.0, z = 0.0;
      double rhobar = alpha, phibar = beta;
      const double bnorm = beta;
      for (;;) {
        ++itn;
        beta = sqrt(rows_apply(gg, d, rp, ja, al, v, -alpha, u, sh));
        const double temp = d2n(d2n(alpha, beta), 0.0);
        anorm = d2n(anorm, temp);
        if (beta > 0.0) {
          scale_own(gg, d.m, 1.0 / beta, u);
          gg.sync();
          alpha = sqrt(cols_apply(gg, d, cp, ri, at, u, -beta, v, sh));
          if (alpha > 0.0) scale_own(gg, d.n, 1.0 / alpha, v);
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
        TT_OWN(gg, j, d.n) {
          const double t = w[j];
          dx[j] = t1 * t + dx[j];
          w[j] = t2 * t + v[j];
          const double tt = t3 * t;
          dk += tt * tt;
        }
        // the sum is also the barrier after which every member reads the new v
        const double dknorm = sqrt(gg.sum(dk, sh));
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
        if (*err) break;  // a gang barrier timed out: abandon (the host reports it)
      }
    }
    // x = q dx + x0
    TT_OWN(gg, j, d.n) x[j] = q[j] * dx[j] + x[j];
    if (threadIdx.x == 0 && gg.rank == 0) nit_out[s] = itn, istop_out[s] = istop;
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
  std::vector<double> work;  // bytes one iteration of each system moves (scheduling weight)
  DBuf<SysDesc> dsys;
  DBuf<int> rp, ja, cp, ri;
  DBuf<double> a, at, q, b, x, u, v, w, dx;
  DBuf<float> a32, at32;  // FP32 copies of the scaled values (TT_PREC_FP32)
  int prec = TT_PREC_FP64;
  DBuf<int> nit, istop;
  // schedule: per workgroup its assignments aptr[g]..aptr[g+1] in global order
  int grid = 0, force_width = 0, max_width = 0;
  double width_scale = 1.0;  // automatic widths are scaled by this (tt_lsqr_batch_tune)
  DBuf<int> aptr;
  DBuf<Assign> asg;
  DBuf<unsigned> gcnt;
  DBuf<double> gpart;
  DBuf<int> err;
  hipStream_t s = nullptr;
  ~tt_lsqr_batch() {
    if (s) (void)hipStreamDestroy(s);
  }
};

namespace {

// Workgroups of lsqr_batch_kernel resident on the device at once (the fewer
// of the FP64 and FP32 instantiations, so a schedule fits either).
int resident_grid() {
  int dev = 0, cus = 0, p64 = 0, p32 = 0;
  TT_HIP(hipGetDevice(&dev));
  TT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  TT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&p64, lsqr_batch_kernel<double>, kNT, 0));
  TT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&p32, lsqr_batch_kernel<float>, kNT, 0));
  return std::max(1, std::min(p64, p32)) * cus;
}

__global__ void narrow_kernel(long n, const double *__restrict__ a, float *__restrict__ b) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) b[k] = (float)a[k];
}

// List scheduling of the systems onto gangs of workgroups: systems in
// decreasing work, each to the `width` workgroups that fall free first.
// width = the system's share of the batch's work in units of one workgroup's
// fair share (1 .. 64), or force_width.
void make_schedule(tt_lsqr_batch &h) {
  const int G = h.grid;
  double total = 0.0;
  for (double w : h.work) total += w;
  const double share = total / G;
  std::vector<int> ord(h.nsys);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return h.work[x] > h.work[y]; });
  std::vector<double> free_at(G, 0.0);
  std::vector<std::vector<Assign>> lists(G);
  std::vector<int> idx(G);
  h.max_width = 1;
  for (int s : ord) {
    int W = h.force_width > 0 ? h.force_width : (int)std::floor(h.width_scale * h.work[s] / share);
    W = std::max(1, std::min({W, 64, G}));
    // the W workgroups that fall free first (ties: lower index)
    std::iota(idx.begin(), idx.end(), 0);
    std::partial_sort(idx.begin(), idx.begin() + W, idx.end(), [&](int x, int y) {
      return free_at[x] < free_at[y] || (free_at[x] == free_at[y] && x < y);
    });
    std::sort(idx.begin(), idx.begin() + W);
    double start = 0.0;
    for (int k = 0; k < W; ++k) start = std::max(start, free_at[idx[k]]);
    const double end = start + h.work[s] / W * (W > 1 ? 1.1 : 1.0);
    for (int k = 0; k < W; ++k) {
      free_at[idx[k]] = end;
      lists[idx[k]].push_back(Assign{s, s, k, W});
    }
    h.max_width = std::max(h.max_width, W);
  }
  std::vector<int> aptr(G + 1, 0);
  std::vector<Assign> flat;
  for (int g = 0; g < G; ++g) {
    flat.insert(flat.end(), lists[g].begin(), lists[g].end());
    aptr[g + 1] = (int)flat.size();
  }
  h.grid = G;
  h.aptr.upload(aptr, h.s);
  h.asg.upload(flat, h.s);
  TT_HIP(hipStreamSynchronize(h.s));
}

}  // namespace

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
    // scheduling weight: the bytes of one iteration (both products, the vectors)
    h->work.resize(nsys);
    for (int k = 0; k < nsys; ++k)
      h->work[k] = 2.0 * (12.0 * nnz[k] + 4.0 * (m[k] + 1)) + 8.0 * (3.0 * m[k] + 6.0 * n[k]);
    h->grid = resident_grid();
    make_schedule(*h);
    h->b.alloc(mo), h->u.alloc(mo), h->x.alloc(no), h->v.alloc(no), h->w.alloc(no), h->dx.alloc(no);
    h->nit.alloc(nsys), h->istop.alloc(nsys);
    h->gcnt.alloc(nsys), h->gpart.alloc((size_t)nsys * 128), h->err.alloc(1);
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
    h->gcnt.zero(h->s), h->err.zero(h->s);
    hipEvent_t e0, e1;
    TT_HIP(hipEventCreate(&e0));
    TT_HIP(hipEventCreate(&e1));
    TT_HIP(hipEventRecord(e0, h->s));
    // grid <= the workgroups resident at once: the members of a gang run together
    if (h->prec == TT_PREC_FP32)
      hipLaunchKernelGGL(lsqr_batch_kernel<float>, dim3(h->grid), dim3(kNT), 0, h->s, h->aptr.p, h->asg.p, h->dsys.p,
                         h->rp.p, h->ja.p, h->a.p, h->a32.p, h->cp.p, h->ri.p, h->at32.p, h->q.p, h->b.p, h->x.p,
                         h->u.p, h->v.p, h->w.p, h->dx.p, maxit, (double)eps4, h->gcnt.p, h->gpart.p, h->err.p,
                         h->nit.p, h->istop.p);
    else
      hipLaunchKernelGGL(lsqr_batch_kernel<double>, dim3(h->grid), dim3(kNT), 0, h->s, h->aptr.p, h->asg.p, h->dsys.p,
                         h->rp.p, h->ja.p, h->a.p, h->a.p, h->cp.p, h->ri.p, h->at.p, h->q.p, h->b.p, h->x.p, h->u.p,
                         h->v.p, h->w.p, h->dx.p, maxit, (double)eps4, h->gcnt.p, h->gpart.p, h->err.p, h->nit.p,
                         h->istop.p);
    TT_LAUNCH_CHECK();
    TT_HIP(hipEventRecord(e1, h->s));
    h->x.download(x, h->ntot, h->s);
    if (nit) h->nit.download(nit, h->nsys, h->s);
    if (istop) h->istop.download(istop, h->nsys, h->s);
    int herr = 0;
    h->err.download(&herr, 1, h->s);
    TT_HIP(hipStreamSynchronize(h->s));
    if (herr) throw std::runtime_error("a gang barrier timed out (workgroups of one gang not resident together)");
    float t = 0.f;
    TT_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0), (void)hipEventDestroy(e1);
    if (ms) *ms = t;
    return 0;
  });
}

extern "C" int tt_lsqr_batch_set_width(tt_lsqr_batch *h, int32_t width, int32_t *max_width) {
  return guarded("tt_lsqr_batch_set_width", [&] {
    if (!h || width < 0 || width > 64) throw std::runtime_error("width must be 0 (automatic) .. 64");
    h->force_width = width;
    h->grid = resident_grid();
    make_schedule(*h);
    if (max_width) *max_width = h->max_width;
    return 0;
  });
}

extern "C" int tt_lsqr_batch_tune(tt_lsqr_batch *h, const double *b, const double *x, int32_t maxit, float eps4,
                                  double *best_scale, double ms_out[4]) {
  return guarded("tt_lsqr_batch_tune", [&] {
    if (!h || !b || !x) throw std::runtime_error("NULL argument");
    // candidate scalings of the automatic gang widths, each timed on a full
    // solve of the given systems (the hot start x is not changed); the
    // fastest schedule stays
    const double cand[4] = {0.5, 1.0, 2.0, 4.0};
    std::vector<double> xw(h->ntot);
    std::vector<int32_t> nit(h->nsys), ist(h->nsys);
    double best = 1e300, bs = 1.0;
    const int fw = h->force_width;
    h->force_width = 0;
    for (int c = 0; c < 4; ++c) {
      h->width_scale = cand[c];
      h->grid = resident_grid();
      make_schedule(*h);
      double t = 0.0;
      std::memcpy(xw.data(), x, sizeof(double) * h->ntot);
      if (tt_lsqr_batch_solve(h, b, xw.data(), maxit, eps4, nit.data(), ist.data(), &t) < 0)
        throw std::runtime_error(tt_last_error());
      if (ms_out) ms_out[c] = t;
      if (t < best) best = t, bs = cand[c];
    }
    h->width_scale = bs;
    h->force_width = fw;
    h->grid = resident_grid();
    make_schedule(*h);
    if (best_scale) *best_scale = bs;
    return 0;
  });
}

extern "C" int tt_lsqr_batch_set_precision(tt_lsqr_batch *h, int32_t prec) {
  return guarded("tt_lsqr_batch_set_precision", [&] {
    if (!h) throw std::runtime_error("NULL handle");
    if (prec != TT_PREC_FP64 && prec != TT_PREC_FP32)
      throw std::runtime_error("batched solves take TT_PREC_FP64 or TT_PREC_FP32 values");
    if (prec == TT_PREC_FP32 && !h->a32.p) {
      h->a32.alloc(h->etot), h->at32.alloc(h->etot);
      const unsigned g = ceil_div(h->etot, 256);
      hipLaunchKernelGGL(narrow_kernel, dim3(g), dim3(256), 0, h->s, h->etot, h->a.p, h->a32.p);
      hipLaunchKernelGGL(narrow_kernel, dim3(g), dim3(256), 0, h->s, h->etot, h->at.p, h->at32.p);
      TT_LAUNCH_CHECK();
      TT_HIP(hipStreamSynchronize(h->s));
    }
    h->prec = prec;
    return 0;
  });
}

extern "C" void tt_lsqr_batch_free(tt_lsqr_batch *h) { delete h; }
