// tt_dense_abi.hip -- C-ABI of the dense single-domain kinematics:
//   tt_dense_lu / tt_dense_lu_solve: the LINPACK pair dgefa / dgesl
//     (lib_Solve_v11.f:787-1037) on host arrays, computed on the device and
//     returned in LINPACK's storage (negated multipliers, each column's
//     multipliers in the row order of its own elimination step, pivots as
//     REAL(8));
//   tt_sd_create / tt_sd_solve / tt_sd_free: the device-resident form the
//     Picard loop uses -- integrals (sdKm_corInt, lib_sdkm.f:2377-2449),
//     system matrices and factorisation (sdKm_sysm + FormPREfm,
//     lib_sdkm.f:1171-1210, lib_Solve_v11.f:403-454) once, then one
//     boundary-vorticity update for x, y, z per call (sdkm_solve,
//     lib_sdkm.f:23-121).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "tritok_bem.h"
#include "tt_compress.h"
#include "tt_dense.h"
#include "tt_integrals.h"

namespace tt {
namespace {

// dgesl job = 0 on LINPACK storage: one workgroup walks the n elimination
// steps (interchange, then an axpy with the column's multipliers), then the
// back substitution, column by column (coalesced column reads).
__global__ void __launch_bounds__(1024) linpack_solve_kernel(int n, const double *__restrict__ a,
                                                             const double *__restrict__ ipvt, double *__restrict__ b) {
  __shared__ double t_sh;
  const int tid = threadIdx.x;
  for (int k = 0; k < n - 1; ++k) {
    if (tid == 0) {
      const int l = (int)ipvt[k] - 1;
      const double t = b[l];
      if (l != k) b[l] = b[k], b[k] = t;
      t_sh = t;
    }
    __syncthreads();
    const double t = t_sh;
    const double *col = a + (size_t)n * k;
    for (int i = k + 1 + tid; i < n; i += 1024) b[i] = b[i] + t * col[i];
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; --k) {
    if (tid == 0) {
      b[k] = b[k] / a[k + (size_t)n * k];
      t_sh = -b[k];
    }
    __syncthreads();
    const double t = t_sh;
    const double *col = a + (size_t)n * k;
    for (int i = tid; i < k; i += 1024) b[i] = b[i] + t * col[i];
    __syncthreads();
  }
}

}  // namespace
}  // namespace tt

using namespace tt;

extern "C" int tt_dense_lu(int32_t n, double *a, double *ipvt) {
  return guarded("tt_dense_lu", [&]() -> int {
    if (n <= 0 || !a || !ipvt) throw std::runtime_error("bad argument");
    DBuf<double> da;
    da.upload(a, (size_t)n * n);
    DBuf<int> piv(n), info(1);
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success) throw std::runtime_error("rocblas_create_handle failed");
    const rocblas_status st = rocsolver_dgetrf(h, n, n, da.p, n, piv.p, info.p);
    rocblas_destroy_handle(h);
    if (st != rocblas_status_success) throw std::runtime_error("rocsolver_dgetrf failed");
    std::vector<int> hp(n);
    int hinfo = 0;
    piv.download(hp.data(), n);
    info.download(&hinfo, 1);
    da.download(a, (size_t)n * n);
    TT_HIP(hipDeviceSynchronize());
    // LAPACK applies every later interchange to the earlier multiplier
    // columns; LINPACK leaves them in the order of their own step and stores
    // them negated.  Undo the later interchanges column by column.
    for (int k = 0; k < n - 1; ++k) {
      double *col = a + (size_t)n * k;
      for (int j = n - 2; j > k; --j) {
        const int l = hp[j] - 1;
        if (l != j) std::swap(col[j], col[l]);
      }
      for (int i = k + 1; i < n; ++i) col[i] = -col[i];
    }
    for (int k = 0; k < n; ++k) ipvt[k] = (double)hp[k];
    ipvt[n - 1] = (double)n;
    return hinfo;  // 0, or k when U(k,k) = 0 (as dgefa's info)
  });
}

extern "C" int tt_dense_lu_solve(int32_t n, const double *a, const double *ipvt, double *b) {
  return guarded("tt_dense_lu_solve", [&] {
    if (n <= 0 || !a || !ipvt || !b) throw std::runtime_error("bad argument");
    DBuf<double> da, dp, db;
    da.upload(a, (size_t)n * n), dp.upload(ipvt, n), db.upload(b, n);
    hipLaunchKernelGGL(linpack_solve_kernel, dim3(1), dim3(1024), 0, 0, n, da.p, dp.p, db.p);
    TT_LAUNCH_CHECK();
    db.download(b, n);
    TT_HIP(hipDeviceSynchronize());
    return 0;
  });
}

struct tt_sd {
  hipStream_t s = nullptr;
  DevMesh mesh;
  DenseKin dk;
  int nn = 0, nb = 0;
  DBuf<double> vel, vort;
  std::vector<double> xb, xa;  // boundary-node and node coordinates (n x 3, column-major)
  ~tt_sd() {
    if (s) (void)hipStreamDestroy(s);
  }
};

extern "C" int tt_sd_create(const tt_mesh_view *m, const tt_quad *q, const int32_t *kode, tt_sd **out) {
  return guarded("tt_sd_create", [&] {
    if (!m || !q || !kode || !out) throw std::runtime_error("NULL argument");
    if (!m->nx || !m->ny || !m->nz) throw std::runtime_error("the mesh view needs source point normals");
    auto h = std::make_unique<tt_sd>();
    TT_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
    upload_mesh(*m, h->mesh, h->s);
    QuadDev qh;
    build_quad(*q, qh);
    double ont[10];
    sd_integrals_dev(h->mesh, *m, qh, 0, m->nbnodes, h->dk.mats, ont, h->s);
    double ms = 0;
    dense_setup(h->dk, h->mesh, kode, SdPrec::F64, h->s, &ms);
    h->nn = m->nnodes, h->nb = m->nbnodes;
    h->xa.assign(m->x, m->x + (size_t)m->nnodes * 3);
    h->xb.resize((size_t)m->nbnodes * 3);
    for (int i = 0; i < m->nbnodes; ++i)
      for (int d = 0; d < 3; ++d)
        h->xb[i + (size_t)m->nbnodes * d] = m->x[(m->gbn[i] - 1) + (size_t)m->nnodes * d];
    h->vel.alloc((size_t)m->nnodes * 3);
    h->vort.alloc((size_t)m->nnodes * 3);
    TT_HIP(hipStreamSynchronize(h->s));
    *out = h.release();
    return 0;
  });
}

extern "C" int tt_sd_solve(tt_sd *h, const double *vel, double *vort, const double urBw[3]) {
  return guarded("tt_sd_solve", [&] {
    if (!h || !vel || !vort || !urBw) throw std::runtime_error("NULL argument");
    const size_t n3 = (size_t)h->nn * 3;
    h->vel.upload(vel, n3, h->s);
    h->vort.upload(vort, n3, h->s);
    dense_bw_step(h->dk, h->mesh, h->vel.p, h->vort.p, urBw, h->s);
    h->vort.download(vort, n3, h->s);
    TT_HIP(hipStreamSynchronize(h->s));
    return 0;
  });
}

extern "C" int tt_sd_refine(tt_sd *h, int32_t eqn, int32_t prec, int32_t nrhs, const double *B, double *X,
                            double tol, int32_t maxit, int32_t *iters, double *relres, double ms[2]) {
  return guarded("tt_sd_refine", [&] {
    if (!h || !B || !X) throw std::runtime_error("NULL argument");
    if (eqn < 1 || eqn > 3) throw std::runtime_error("eqn must be 1..3");
    if (prec < TT_PREC_FP64 || prec > TT_PREC_FP16) throw std::runtime_error("unknown precision");
    if (nrhs < 1 || nrhs > 4) throw std::runtime_error("nrhs must be 1..4");
    if (!(tol >= 0.0) || maxit < 1) throw std::runtime_error("tol must be >= 0 and maxit >= 1");
    const size_t len = (size_t)h->dk.nb * nrhs;
    DBuf<double> b, x(len);
    b.upload(B, len, h->s);
    const SdPrec sp = prec == TT_PREC_FP64 ? SdPrec::F64
                      : prec == TT_PREC_FP32 ? SdPrec::F32
                      : prec == TT_PREC_BF16 ? SdPrec::BF16
                                             : SdPrec::F16;
    const RefineResult r = dense_refine(h->dk, h->mesh, eqn - 1, sp, nrhs, b.p, x.p, tol, maxit, h->s);
    x.download(X, len, h->s);
    TT_HIP(hipStreamSynchronize(h->s));
    if (iters) *iters = r.iters;
    if (relres) *relres = r.relres;
    if (ms) ms[0] = r.ms, ms[1] = r.ms_apply;
    return 0;
  });
}

extern "C" void tt_sd_free(tt_sd *h) { delete h; }

namespace {
// y = M x for a column-major nr x nc device matrix: one thread per row
__global__ void dense_mv_kernel(int nr, int nc, const double *__restrict__ M, const double *__restrict__ x,
                                double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  double s = 0.0;
  for (int j = 0; j < nc; ++j) s += M[i + (size_t)nr * j] * x[j];
  y[i] = s;
}
const double *sd_matrix(tt_sd *h, int which, int &ncol) {
  const SdDev &m = h->dk.mats;
  const double *p[7] = {m.H.p, m.Htx.p, m.Hty.p, m.Htz.p, m.Dx.p, m.Dy.p, m.Dz.p};
  if (which < 0 || which > 6) throw std::runtime_error("which must be 0..6 (H, Htx, Hty, Htz, Dx, Dy, Dz)");
  ncol = which < 4 ? h->nb : h->nn;
  return p[which];
}
}  // namespace

extern "C" int tt_sd_hmat(tt_sd *h, int32_t which, int32_t leaf, double eta, int32_t aca_type, double aca_eps,
                          tt_hmat **out) {
  return guarded("tt_sd_hmat", [&] {
    if (!h || !out) throw std::runtime_error("NULL argument");
    if (aca_type < 1 || aca_type > 3) throw std::runtime_error("aca_type must be 1, 2 or 3");
    int ncol = 0;
    const double *A = sd_matrix(h, which, ncol);
    TT_HIP(hipStreamSynchronize(h->s));
    *out = hmat_from_device(h->nb, ncol, A, h->nb, h->xb.data(), which < 4 ? h->xb.data() : h->xa.data(), leaf, eta,
                            3, aca_type, aca_eps, 0.0);
    return 0;
  });
}

extern "C" int tt_sd_matvec(tt_sd *h, int32_t which, const double *x, double *y) {
  return guarded("tt_sd_matvec", [&] {
    if (!h || !x || !y) throw std::runtime_error("NULL argument");
    int ncol = 0;
    const double *A = sd_matrix(h, which, ncol);
    DBuf<double> dx, dy(h->nb);
    dx.upload(x, ncol, h->s);
    hipLaunchKernelGGL(dense_mv_kernel, dim3(ceil_div(h->nb, 256)), dim3(256), 0, h->s, h->nb, ncol, A, dx.p, dy.p);
    TT_LAUNCH_CHECK();
    dy.download(y, h->nb, h->s);
    TT_HIP(hipStreamSynchronize(h->s));
    return 0;
  });
}
