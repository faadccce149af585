// tt_csr.hip -- the reference's sparse interfaces on the device: CSR x vector
// (CRSxV, lib_util.f:1351-1372) and LSQR on a general CSR system with the
// column scaling of prelsqr2 and the hot start of slvlsqr2
// (lib_Solve_v11.f:4401-4535, LSQR :3128-3887).
//
// The LSQR iteration has the structure of the block solver (tt_lsqr.hip):
//   row kernel of iteration k: block 0 runs the scalar recurrences; the other
//     blocks compute u' = A v / alpha - alpha u / beta, 16 lanes per row,
//     grid-stride over rows, and write one partial of ||u'||^2 per block;
//   column kernel of iteration k: beta from the row partials, A^T u' from a
//     column-major (CSC) copy made at setup, the x/w update of iteration k-1,
//     v' and its partial sums; block 0 stores beta.
// No atomics, no device-scope fences; a solve is bit-for-bit repeatable.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "tritok_bem.h"
#include "tt_hip.h"
#include "tt_lsqr_dev.h"

namespace tt {
namespace {

using namespace lsqr_dev;

constexpr int kTPB = 256;             // threads per block
constexpr int kLPR = 16;              // lanes per row (or column)
constexpr int kRPB = kTPB / kLPR;     // rows per block and pass

__device__ __forceinline__ double group_sum16(double v) {
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// y = A x (0-based CSR)
__global__ void __launch_bounds__(kTPB) csr_mv_kernel(int m, const int *__restrict__ ia, const int *__restrict__ ja,
                                                      const double *__restrict__ a, const double *__restrict__ x,
                                                      double *__restrict__ y) {
  const int row = blockIdx.x * kRPB + threadIdx.x / kLPR, lane = threadIdx.x % kLPR;
  double s = 0.0;
  if (row < m)
    for (int e = ia[row] + lane; e < ia[row + 1]; e += kLPR) s += a[e] * x[ja[e]];
  s = group_sum16(s);
  if (row < m && lane == 0) y[row] = s;
}

// q_j = 1 / ||A(:,j)|| summed in row order (the CSC copy keeps it), A <- A diag(q)
__global__ void colscale_kernel(int n, const int *__restrict__ cp, const double *__restrict__ at,
                                double *__restrict__ q) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double s = 0.0;
  for (int e = cp[j]; e < cp[j + 1]; ++e) s = s + __dmul_rn(at[e], at[e]);  // unfused, as prelsqr2
  q[j] = 1.0 / sqrt(s);
}
__global__ void scale_vals_kernel(int nnz, const int *__restrict__ idx, const double *__restrict__ q,
                                  double *__restrict__ v) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nnz) v[e] = v[e] * q[idx[e]];
}
__global__ void scale_cols_kernel(int n, const int *__restrict__ cp, const double *__restrict__ q,
                                  double *__restrict__ at) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  for (int e = cp[j]; e < cp[j + 1]; ++e) at[e] = at[e] * q[j];
}

// r0 = b - A (x0 / q) per row; partials of ||r0||^2 and ||b||^2 per block
__global__ void __launch_bounds__(kTPB) csr_r0_kernel(int m, const int *__restrict__ ia, const int *__restrict__ ja,
                                                      const double *__restrict__ a, const double *__restrict__ q,
                                                      const double *__restrict__ x0, const double *__restrict__ b,
                                                      double *__restrict__ u, double *__restrict__ pa1,
                                                      double *__restrict__ pa2) {
  __shared__ double sh[kTPB / 64];
  const int lane = threadIdx.x % kLPR, G = gridDim.x;
  double acc1 = 0.0, acc2 = 0.0;
  for (int r0 = blockIdx.x * kRPB; r0 < m; r0 += G * kRPB) {
    const int row = r0 + threadIdx.x / kLPR;
    double s = 0.0;
    if (row < m)
      for (int e = ia[row] + lane; e < ia[row + 1]; e += kLPR) s += a[e] * (x0[ja[e]] / q[ja[e]]);
    s = group_sum16(s);
    if (row < m && lane == 0) {
      const double bb = b[row], r = bb - s;
      u[row] = r;
      acc1 += r * r, acc2 += bb * bb;
    }
  }
  const double s1 = block_sum<kTPB>(acc1, sh), s2 = block_sum<kTPB>(acc2, sh);
  if (threadIdx.x == 0) pa1[blockIdx.x] = s1, pa2[blockIdx.x] = s2;
}

// iteration k >= 1, rows: block 0 scalars, blocks 1.. u' = A v / alpha - alpha u / beta
__global__ void __launch_bounds__(kTPB) csr_row_kernel(int k, int m, const int *__restrict__ ia,
                                                       const int *__restrict__ ja, const double *__restrict__ a,
                                                       const double *__restrict__ vp, double *__restrict__ u,
                                                       double *__restrict__ pa1, const double *__restrict__ pb1,
                                                       const double *__restrict__ pb2, int nb,
                                                       const LsqrState *__restrict__ sin,
                                                       LsqrState *__restrict__ sout) {
  __shared__ double sh[kTPB / 64];
  if (blockIdx.x == 0) {
    const double nvs = reduce_partials<kTPB>(pb1, nb, sh);
    const double dks = k >= 3 ? reduce_partials<kTPB>(pb2, nb, sh) : 0.0;
    if (threadIdx.x == 0) {
      LsqrState o;
      scalar_update(k, *sin, nvs, dks, o);
      *sout = o;
    }
    return;
  }
  if (sin->done) return;
  const double beta = sin->beta, alpha_prev = sin->alpha;
  const double nvs = reduce_partials<kTPB>(pb1, nb, sh);
  const double alpha = (k == 1 || beta > 0.0) ? sqrt(nvs) : alpha_prev;
  const int lane = threadIdx.x % kLPR, G = gridDim.x - 1;
  double acc = 0.0;
  for (int r0 = (blockIdx.x - 1) * kRPB; r0 < m; r0 += G * kRPB) {
    const int row = r0 + threadIdx.x / kLPR;
    double s = 0.0;
    if (row < m)
      for (int e = ia[row] + lane; e < ia[row + 1]; e += kLPR) {
        const double vj = vp[ja[e]];
        s += a[e] * (alpha > 0.0 ? vj / alpha : vj);
      }
    s = group_sum16(s);
    if (row < m && lane == 0) {
      const double uo = u[row], uold = beta > 0.0 ? uo / beta : uo;
      const double un = -alpha * uold + s;
      u[row] = un;
      acc += un * un;
    }
  }
  const double s1 = block_sum<kTPB>(acc, sh);
  if (threadIdx.x == 0) pa1[blockIdx.x - 1] = s1;
}

// iteration k (0: initialisation), columns
__global__ void __launch_bounds__(kTPB) csr_col_kernel(int n, int k, const int *__restrict__ cp,
                                                       const int *__restrict__ ci, const double *__restrict__ at,
                                                       const double *__restrict__ u, double *__restrict__ vp,
                                                       double *__restrict__ w, double *__restrict__ x,
                                                       const double *__restrict__ pa1, const double *__restrict__ pa2,
                                                       int na, double *__restrict__ pb1, double *__restrict__ pb2,
                                                       LsqrState *st) {
  __shared__ double sh[kTPB / 64];
  if (k >= 1 && st->done) return;
  const double beta = sqrt(reduce_partials<kTPB>(pa1, na, sh));  // beta_{k+1}
  const double alpha = st->alpha, t1 = st->t1, t2 = st->t2, t3 = st->t3;
  const int lane = threadIdx.x % kLPR, G = gridDim.x;
  double nv = 0.0, dk = 0.0;
  for (int c0 = blockIdx.x * kRPB; c0 < n; c0 += G * kRPB) {
    const int j = c0 + threadIdx.x / kLPR;
    double s = 0.0;
    if (j < n)
      for (int e = cp[j] + lane; e < cp[j + 1]; e += kLPR) s += at[e] * u[ci[e]];
    s = group_sum16(s);
    if (j < n && lane == 0) {
      const double vold = vp[j];
      double vk = 0.0;
      if (k >= 1) {
        vk = alpha > 0.0 ? vold / alpha : vold;
        if (k == 1) {
          w[j] = vk;
          x[j] = 0.0;
        } else {
          const double wo = w[j];
          x[j] = t1 * wo + x[j];
          w[j] = t2 * wo + vk;
          const double tt = t3 * wo;
          dk += tt * tt;
        }
      }
      double nvp;
      if (beta > 0.0)
        nvp = s / beta - beta * vk;
      else
        nvp = k == 0 ? 0.0 : vold;
      vp[j] = nvp;
      nv += nvp * nvp;
    }
  }
  const double s1 = block_sum<kTPB>(nv, sh), s2 = block_sum<kTPB>(dk, sh);
  if (threadIdx.x == 0) pb1[blockIdx.x] = s1, pb2[blockIdx.x] = s2;
  if (blockIdx.x == 0) {
    double nbv = 0.0;
    if (k == 0) nbv = sqrt(reduce_partials<kTPB>(pa2, na, sh));
    if (threadIdx.x == 0) {
      st->beta = beta;
      if (k == 0) {
        st->nrm_r = beta, st->nrm_b = nbv;
        st->btol = st->atol * nbv / beta;
        st->phibar = beta, st->bnorm = beta, st->rnorm = beta;
      }
    }
  }
}

__global__ void csr_finish_kernel(int n, const double *__restrict__ q, const double *__restrict__ dx,
                                  double *__restrict__ x) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) x[j] = q[j] * dx[j] + x[j];
}

// 1-based CSR from the host -> 0-based device CSR and CSC (entries of a
// column in row order)
void upload_csr(int m, int n, int nnz, const int32_t *ia, const int32_t *ja, std::vector<int> &ia0,
                std::vector<int> &ja0, std::vector<int> &cp, std::vector<int> &ci, std::vector<int> &perm) {
  if (m <= 0 || n <= 0 || nnz < 0) throw std::runtime_error("bad CSR dimensions");
  if (ia[0] != 1 || ia[m] - 1 != nnz) throw std::runtime_error("row pointers must run from 1 to nnz + 1");
  ia0.resize(m + 1);
  ja0.resize(nnz);
  for (int i = 0; i <= m; ++i) ia0[i] = ia[i] - 1;
  for (int i = 0; i < m; ++i)
    if (ia0[i + 1] < ia0[i]) throw std::runtime_error("row pointers must not decrease");
  for (int e = 0; e < nnz; ++e) {
    ja0[e] = ja[e] - 1;
    if (ja0[e] < 0 || ja0[e] >= n) throw std::runtime_error("column index outside 1..n");
  }
  cp.assign(n + 1, 0);
  for (int e = 0; e < nnz; ++e) cp[ja0[e] + 1]++;
  for (int j = 0; j < n; ++j) cp[j + 1] += cp[j];
  std::vector<int> pos(cp.begin(), cp.end() - 1);
  ci.resize(nnz);
  perm.resize(nnz);
  for (int i = 0; i < m; ++i)
    for (int e = ia0[i]; e < ia0[i + 1]; ++e) {
      const int p = pos[ja0[e]]++;
      ci[p] = i;
      perm[p] = e;  // CSC entry p holds CSR entry e
    }
}

__global__ void gather_kernel(int nnz, const int *__restrict__ perm, const double *__restrict__ a,
                              double *__restrict__ at) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nnz) at[p] = a[perm[p]];
}

}  // namespace
}  // namespace tt

using namespace tt;

struct tt_lsqr {
  int m = 0, n = 0, nnz = 0, grow = 0, gcol = 0, dev = 0;
  hipStream_t s = nullptr;
  DBuf<int> ia, ja, cp, ci;
  DBuf<double> a, at, q;
  DBuf<double> u, v, w, x, dx, b, pa1, pa2, pb1, pb2;
  DBuf<LsqrState> state;
  LsqrState *host_state = nullptr;
  int last_nit = 0;
  ~tt_lsqr() {
    if (host_state) (void)hipHostFree(host_state);
    if (s) (void)hipStreamDestroy(s);
  }
};

extern "C" int tt_csr_mv(int32_t m, int32_t n, int32_t nnz, const int32_t *ia, const int32_t *ja, const double *a,
                         const double *x, double *y) {
  return guarded("tt_csr_mv", [&] {
    if (!ia || !ja || !a || !x || !y) throw std::runtime_error("NULL argument");
    std::vector<int> ia0, ja0, cp, ci, perm;
    upload_csr(m, n, nnz, ia, ja, ia0, ja0, cp, ci, perm);
    DBuf<int> dia, dja;
    DBuf<double> da, dx, dy(m);
    dia.upload(ia0), dja.upload(ja0), da.upload(a, nnz), dx.upload(x, n);
    hipLaunchKernelGGL(csr_mv_kernel, dim3(ceil_div(m, kRPB)), dim3(kTPB), 0, 0, m, dia.p, dja.p, da.p, dx.p, dy.p);
    TT_LAUNCH_CHECK();
    dy.download(y, m);
    TT_HIP(hipDeviceSynchronize());
    return 0;
  });
}

extern "C" int tt_lsqr_setup(int32_t m, int32_t n, int32_t nnz, const int32_t *ia, const int32_t *ja, double *a,
                             double *q, int32_t scale_host, tt_lsqr **out) {
  return guarded("tt_lsqr_setup", [&] {
    if (!ia || !ja || !a || !out) throw std::runtime_error("NULL argument");
    auto s = std::make_unique<tt_lsqr>();
    s->m = m, s->n = n, s->nnz = nnz;
    TT_HIP(hipStreamCreateWithFlags(&s->s, hipStreamNonBlocking));
    std::vector<int> ia0, ja0, cp, ci, perm;
    upload_csr(m, n, nnz, ia, ja, ia0, ja0, cp, ci, perm);
    hipStream_t st = s->s;
    s->ia.upload(ia0, st), s->ja.upload(ja0, st), s->cp.upload(cp, st), s->ci.upload(ci, st);
    s->a.upload(a, nnz, st);
    DBuf<int> dperm;
    dperm.upload(perm, st);
    s->at.alloc(nnz);
    s->q.alloc(n);
    if (nnz) {
      hipLaunchKernelGGL(gather_kernel, dim3(ceil_div(nnz, 256)), dim3(256), 0, st, nnz, dperm.p, s->a.p, s->at.p);
      TT_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(colscale_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, s->cp.p, s->at.p, s->q.p);
    if (nnz) {
      hipLaunchKernelGGL(scale_vals_kernel, dim3(ceil_div(nnz, 256)), dim3(256), 0, st, nnz, s->ja.p, s->q.p, s->a.p);
      hipLaunchKernelGGL(scale_cols_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, s->cp.p, s->q.p, s->at.p);
    }
    TT_LAUNCH_CHECK();
    if (q) s->q.download(q, n, st);
    if (scale_host) s->a.download(a, nnz, st);
    int dev = 0, cus = 0;
    TT_HIP(hipGetDevice(&dev));
    TT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    s->dev = dev;
    s->grow = std::max(1, std::min((int)ceil_div(m, kRPB), 8 * cus));
    s->gcol = std::max(1, std::min((int)ceil_div(n, kRPB), 8 * cus));
    s->u.alloc(m), s->b.alloc(m), s->v.alloc(n), s->w.alloc(n), s->x.alloc(n), s->dx.alloc(n);
    s->pa1.alloc(std::max(s->grow, 1)), s->pa2.alloc(std::max(s->grow, 1));
    s->pb1.alloc(s->gcol), s->pb2.alloc(s->gcol);
    s->state.alloc(2);
    TT_HIP(hipHostMalloc((void **)&s->host_state, sizeof(LsqrState), hipHostMallocDefault));
    TT_HIP(hipStreamSynchronize(st));
    *out = s.release();
    return 0;
  });
}

namespace {

// Work vectors of one solve: u (m), v, w, dx (n), the row and column partials
// and the two state copies.
struct CsrWork {
  double *u, *v, *w, *dx, *pa1, *pa2, *pb1, *pb2;
  LsqrState *state;
};

// slvlsqr2's initial state (both copies): atol = eps, conlim = 1e8, itnlim
__global__ void csr_init_state_kernel(LsqrState *st, double atol, int itnlim) {
  const int i = threadIdx.x;
  if (i >= 2) return;
  LsqrState o{};
  o.cs2 = -1.0;
  o.atol = atol;
  o.ctol = 1.0 / 1.0e8;
  o.itnlim = itnlim;
  st[i] = o;
}

__global__ void csr_zero_kernel(int n, double *__restrict__ p) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) p[j] = 0.0;
}

// The solve of slvlsqr2 (lib_Solve_v11.f:4401-4500) on device b and x (x is
// the hot start on entry and the solution on exit).  The state is read back
// into *rb (host memory) once per batch of iterations.
void csr_solve(tt_lsqr *s, hipStream_t st, const double *b, double *x, const CsrWork &wk, int32_t maxit, float eps4,
               LsqrState *rb, int32_t *nit, int32_t *istop) {
  const int m = s->m, n = s->n;
  const int itnlim = maxit == 0 ? 10 * (m + n + 50) : maxit;
  hipLaunchKernelGGL(csr_init_state_kernel, dim3(1), dim3(64), 0, st, wk.state, (double)eps4, itnlim);
  hipLaunchKernelGGL(csr_zero_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, wk.dx);
  hipLaunchKernelGGL(csr_r0_kernel, dim3(s->grow), dim3(kTPB), 0, st, m, s->ia.p, s->ja.p, s->a.p, s->q.p, x, b,
                     wk.u, wk.pa1, wk.pa2);
  hipLaunchKernelGGL(csr_col_kernel, dim3(s->gcol), dim3(kTPB), 0, st, n, 0, s->cp.p, s->ci.p, s->at.p, wk.u, wk.v,
                     wk.w, wk.dx, wk.pa1, wk.pa2, s->grow, wk.pb1, wk.pb2, wk.state);
  TT_LAUNCH_CHECK();
  int k = 1, batch = s->last_nit > 0 ? s->last_nit + 2 : 8;
  for (;;) {
    for (int i = 0; i < batch; ++i, ++k) {
      hipLaunchKernelGGL(csr_row_kernel, dim3(s->grow + 1), dim3(kTPB), 0, st, k, m, s->ia.p, s->ja.p, s->a.p, wk.v,
                         wk.u, wk.pa1, wk.pb1, wk.pb2, s->gcol, wk.state + ((k - 1) & 1), wk.state + (k & 1));
      hipLaunchKernelGGL(csr_col_kernel, dim3(s->gcol), dim3(kTPB), 0, st, n, k, s->cp.p, s->ci.p, s->at.p, wk.u,
                         wk.v, wk.w, wk.dx, wk.pa1, wk.pa2, s->grow, wk.pb1, wk.pb2, wk.state + (k & 1));
    }
    TT_LAUNCH_CHECK();
    TT_HIP(hipMemcpyAsync(rb, wk.state + ((k - 1) & 1), sizeof(LsqrState), hipMemcpyDeviceToHost, st));
    TT_HIP(hipStreamSynchronize(st));
    if (rb->done) break;
    if (k > itnlim + 3) throw std::runtime_error("LSQR did not signal termination");
    batch = std::max(4, std::min(64, k / 4));
  }
  if (nit) *nit = rb->itn;
  if (istop) *istop = rb->istop;
  s->last_nit = rb->itn;
  hipLaunchKernelGGL(csr_finish_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, s->q.p, wk.dx, x);
  TT_LAUNCH_CHECK();
}

// Scratch of the allocation-free solve: one array per device, held until the
// process ends.  In host code `scratch` only names the symbol; its device
// address comes from hipGetSymbolAddress, once per device (scratch_base).
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr size_t kScratchDoubles = kScratchBytes / sizeof(double);
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_base[kMaxDevices] = {};

// device address of `scratch` on the current device
double *scratch_base() {
  int dev = 0;
  TT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index outside the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_base[dev]) {
    void *p = nullptr;
    TT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_base[dev] = static_cast<double *>(p);
  }
  return g_scratch_base[dev];
}

// makes dev current for the lifetime of the guard
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    TT_HIP(hipGetDevice(&prev));
    if (prev != dev) TT_HIP(hipSetDevice(dev));
  }
  ~DeviceGuard() {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

}  // namespace

extern "C" int tt_lsqr_solve(tt_lsqr *s, const double *b, double *x, int32_t maxit, float eps4, int32_t *nit,
                             int32_t *istop) {
  return guarded("tt_lsqr_solve", [&] {
    if (!s || !b || !x) throw std::runtime_error("NULL argument");
    hipStream_t st = s->s;
    s->b.upload(b, s->m, st);
    s->x.upload(x, s->n, st);
    const CsrWork wk{s->u.p, s->v.p, s->w.p, s->dx.p, s->pa1.p, s->pa2.p, s->pb1.p, s->pb2.p, s->state.p};
    csr_solve(s, st, s->b.p, s->x.p, wk, maxit, eps4, s->host_state, nit, istop);
    s->x.download(x, s->n, st);
    TT_HIP(hipStreamSynchronize(st));
    return 0;
  });
}

extern "C" int tt_lsqr_solve_noalloc(tt_lsqr *s, const double *b_dev, double *x_dev, int32_t maxit, float eps4,
                                     int32_t *nit, int32_t *istop) {
  return guarded("tt_lsqr_solve_noalloc", [&] {
    if (!s || !b_dev || !x_dev) throw std::runtime_error("NULL argument");
    DeviceGuard dg(s->dev);
    double *base = scratch_base();
    // carve the work vectors out of the scratch, each on a 256-byte boundary
    const size_t align = 256 / sizeof(double);
    auto up = [&](size_t c) { return (c + align - 1) / align * align; };
    const size_t n_state = up((2 * sizeof(LsqrState) + sizeof(double) - 1) / sizeof(double));
    const size_t need = n_state + up(s->m) + 3 * up(s->n) + 2 * up(s->grow) + 2 * up(s->gcol);
    if (need > kScratchDoubles) {
      char msg[256];
      std::snprintf(msg, sizeof(msg), "system needs %zu bytes of work space, the scratch holds %zu (m + 3n too large)",
                    need * sizeof(double), kScratchBytes);
      throw std::runtime_error(msg);
    }
    CsrWork wk;
    size_t off = 0;
    wk.state = reinterpret_cast<LsqrState *>(base);
    off += n_state;
    auto take = [&](size_t c) {
      double *p = base + off;
      off += up(c);
      return p;
    };
    wk.u = take(s->m), wk.v = take(s->n), wk.w = take(s->n), wk.dx = take(s->n);
    wk.pa1 = take(s->grow), wk.pa2 = take(s->grow), wk.pb1 = take(s->gcol), wk.pb2 = take(s->gcol);
    LsqrState rb;
    csr_solve(s, s->s, b_dev, x_dev, wk, maxit, eps4, &rb, nit, istop);
    TT_HIP(hipStreamSynchronize(s->s));
    return 0;
  });
}

extern "C" void tt_lsqr_free(tt_lsqr *s) { delete s; }
