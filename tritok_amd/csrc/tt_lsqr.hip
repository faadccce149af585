// tt_lsqr.hip -- LSQR on the device for the block-structured subdomain systems.
//
// Per iteration two kernels, no host synchronisation:
//   lsqr_A (one wavefront per cell): u' <- A v - alpha u, ||u'||^2 partial,
//           P <- A^T u' per cell (the block is read once, from LDS, for both);
//           the last block to finish reduces ||u'||^2 -> beta.
//   lsqr_B (one thread per unknown): finishes the x/w update of the previous
//           iteration, gathers A^T u' from the cells of its column,
//           v' <- A^T u / beta - beta v, ||v'||^2 partial; the last block
//           reduces alpha, runs the previous iteration's stopping tests and
//           the new plane rotation.
// Every reduction uses a fixed assignment of terms to threads, so a solve is
// bit-for-bit repeatable.  The scalar recurrences are those of the reference
// LSQR (lib_Solve_v11.f:3283-3420) with damp = 0.
#include <cmath>
#include <cstring>

#include "tt_lsqr.h"

namespace tt {

namespace {

__device__ __forceinline__ double d2norm(double a, double b) {
  double scale = fabs(a) + fabs(b);
  if (scale == 0.0) return 0.0;
  double sa = a / scale, sb = b / scale;
  return scale * sqrt(sa * sa + sb * sb);
}

// deterministic block-wide sum (blockDim power of two, <= 256)
template <int NT>
__device__ double block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int o = NT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// Signals block completion; returns true in the last block (after an acquire
// fence, so the other blocks' partials are visible).
__device__ bool last_block(unsigned *counter, bool *flag) {
  __threadfence();
  if (threadIdx.x == 0) {
    unsigned t = atomicAdd(counter, 1u);
    *flag = (t == gridDim.x - 1);
  }
  __syncthreads();
  bool last = *flag;
  if (last) __threadfence();
  return last;
}

template <int NT>
__device__ double reduce_partials(const double *p, int n, double *sh) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) acc += p[i];
  return block_sum<NT>(acc, sh);
}

// ---------------------------------------------------------------------------
// setup kernels
__global__ void colnorm_kernel(int n, int R, int L, const int *__restrict__ tptr, const int *__restrict__ tidx,
                               const double *__restrict__ Afull, double *__restrict__ q) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double s = 0.0;
  for (int e = tptr[j]; e < tptr[j + 1]; ++e) {
    int cl = tidx[e], c = cl / L, l = cl % L;
    const double *blk = Afull + (size_t)c * R * L;
    for (int r = 0; r < R; ++r) s += blk[r * L + l] * blk[r * L + l];
  }
  q[j] = 1.0 / sqrt(s);
}

__global__ void scale_kernel(size_t total, int R, int L, const int *__restrict__ colmap, const double *__restrict__ q,
                             const double *__restrict__ Afull, double *__restrict__ A, float *__restrict__ A32) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  size_t c = k / ((size_t)R * L);
  int l = (int)(k % L);
  int j = colmap[c * L + l];
  double v = j >= 0 ? Afull[k] * q[j] : 0.0;
  A[k] = v;
  if (A32) A32[k] = (float)v;
}

// ---------------------------------------------------------------------------
// r0 = b - A (x0/q), P = A^T r0; last block: norms and tolerances
template <typename T>
__global__ void __launch_bounds__(64) lsqr_r0(int R, int L, const T *__restrict__ A, const int *__restrict__ colmap,
                                              const double *__restrict__ q, const double *__restrict__ x0,
                                              const double *__restrict__ b, double *__restrict__ u,
                                              double *__restrict__ P, double *__restrict__ part1,
                                              double *__restrict__ part2, LsqrState *st) {
  extern __shared__ unsigned char lds_raw[];
  T *blk = reinterpret_cast<T *>(lds_raw);
  __shared__ double xs[64], rr[64], sh[64];
  __shared__ bool flag;
  const int c = blockIdx.x, t = threadIdx.x;
  const T *Ab = A + (size_t)c * R * L;
  for (int k = t; k < R * L; k += 64) blk[k] = Ab[k];
  if (t < L) {
    int j = colmap[(size_t)c * L + t];
    xs[t] = j >= 0 ? x0[j] / q[j] : 0.0;
  }
  __syncthreads();
  double r = 0.0, bb = 0.0;
  if (t < R) {
    double s = 0.0;
    for (int l = 0; l < L; ++l) s += (double)blk[t * L + l] * xs[l];
    bb = b[(size_t)c * R + t];
    r = bb - s;
    u[(size_t)c * R + t] = r;
    rr[t] = r;
  }
  __syncthreads();
  if (t < L) {
    double s = 0.0;
    for (int i = 0; i < R; ++i) s += (double)blk[i * L + t] * rr[i];
    P[(size_t)c * L + t] = s;
  }
  double s1 = block_sum<64>(r * r, sh), s2 = block_sum<64>(bb * bb, sh);
  if (t == 0) part1[c] = s1, part2[c] = s2;
  if (!last_block(&st->counter, &flag)) return;
  double nr = sqrt(reduce_partials<64>(part1, gridDim.x, sh));
  double nb = sqrt(reduce_partials<64>(part2, gridDim.x, sh));
  if (t == 0) {
    st->counter = 0;
    st->nrm_r = nr, st->nrm_b = nb;
    st->btol = st->atol * nb / nr;
    st->beta = nr;
  }
}

// ---------------------------------------------------------------------------
// one iteration, cell part: u' = A v - alpha u ; P = A^T u'
template <typename T>
__global__ void __launch_bounds__(64) lsqr_A(int R, int L, const T *__restrict__ A, const int *__restrict__ colmap,
                                             const double *__restrict__ vp, double *__restrict__ u,
                                             double *__restrict__ P, double *__restrict__ part1, LsqrState *st) {
  if (st->done) return;
  extern __shared__ unsigned char lds_raw[];
  T *blk = reinterpret_cast<T *>(lds_raw);
  __shared__ double vs[64], uu[64], sh[64];
  __shared__ bool flag;
  const int c = blockIdx.x, t = threadIdx.x;
  const double alpha = st->alpha, beta = st->beta;
  const T *Ab = A + (size_t)c * R * L;
  for (int k = t; k < R * L; k += 64) blk[k] = Ab[k];
  if (t < L) {
    int j = colmap[(size_t)c * L + t];
    vs[t] = j >= 0 ? (alpha > 0.0 ? vp[j] / alpha : vp[j]) : 0.0;
  }
  __syncthreads();
  double un = 0.0;
  if (t < R) {
    double s = 0.0;
    for (int l = 0; l < L; ++l) s += (double)blk[t * L + l] * vs[l];
    double uo = u[(size_t)c * R + t];
    if (beta > 0.0) uo = uo / beta;
    un = -alpha * uo + s;
    u[(size_t)c * R + t] = un;
    uu[t] = un;
  }
  __syncthreads();
  if (t < L) {
    double s = 0.0;
    for (int i = 0; i < R; ++i) s += (double)blk[i * L + t] * uu[i];
    P[(size_t)c * L + t] = s;
  }
  double s1 = block_sum<64>(un * un, sh);
  if (t == 0) part1[c] = s1;
  if (!last_block(&st->counter, &flag)) return;
  double nb = sqrt(reduce_partials<64>(part1, gridDim.x, sh));
  if (t == 0) {
    st->counter = 0;
    st->beta = nb;  // beta_{k+1}
  }
}

// ---------------------------------------------------------------------------
// one iteration, unknown part.  k = iteration number (0: initialisation)
__global__ void __launch_bounds__(256) lsqr_B(int n, int k, const int *__restrict__ tptr, const int *__restrict__ tidx,
                                              const double *__restrict__ P, double *__restrict__ vp,
                                              double *__restrict__ w, double *__restrict__ x,
                                              double *__restrict__ part1, double *__restrict__ part2, LsqrState *st) {
  if (st->done) return;
  __shared__ double sh[256];
  __shared__ bool flag;
  const int j = blockIdx.x * 256 + threadIdx.x;
  const double alpha = st->alpha, beta = st->beta;
  const double t1 = st->t1, t2 = st->t2, t3 = st->t3;
  double nv = 0.0, dk = 0.0;
  if (j < n) {
    double vk = 0.0;
    if (k >= 1) {
      vk = alpha > 0.0 ? vp[j] / alpha : vp[j];
      if (k == 1) {
        w[j] = vk;
        x[j] = 0.0;
      } else {
        double wo = w[j];
        x[j] = t1 * wo + x[j];
        w[j] = t2 * wo + vk;
        double tt = t3 * wo;
        dk = tt * tt;
      }
    }
    double s = 0.0;
    for (int e = tptr[j]; e < tptr[j + 1]; ++e) s += P[tidx[e]];
    double nvp;
    if (beta > 0.0)
      nvp = s / beta - beta * vk;
    else
      nvp = k == 0 ? 0.0 : vp[j];  // the reference keeps v and alpha when beta = 0
    vp[j] = nvp;
    nv = nvp * nvp;
  }
  double s1 = block_sum<256>(nv, sh), s2 = block_sum<256>(dk, sh);
  if (threadIdx.x == 0) part1[blockIdx.x] = s1, part2[blockIdx.x] = s2;
  if (!last_block(&st->counter, &flag)) return;
  double nvs = reduce_partials<256>(part1, gridDim.x, sh);
  double dks = reduce_partials<256>(part2, gridDim.x, sh);
  if (threadIdx.x != 0) return;
  st->counter = 0;
  LsqrState &S = *st;
  if (k == 0) {
    // initialisation: alpha_1, and the reference's exit when A^T b = 0
    double a1 = sqrt(nvs);
    S.alpha = a1;
    S.rhobar = a1;
    S.phibar = S.beta;
    S.bnorm = S.beta;
    S.rnorm = S.beta;
    S.arnorm = a1 * S.beta;
    if (S.arnorm == 0.0) S.done = 1, S.istop = 0, S.itn = 0;
    return;
  }
  if (k >= 2) {
    // finish iteration k-1 (whose x/w update this kernel applied)
    double dknorm = sqrt(dks);
    S.dnorm = d2norm(S.dnorm, dknorm);
    double rho = S.pending_rho, phi = S.pending_phi, theta = S.pending_theta, tau = S.pending_tau;
    double delta = S.sn2 * rho, gambar = -S.cs2 * rho;
    double rhs = phi - delta * S.z;
    double zbar = rhs / gambar;
    S.xnorm = d2norm(S.xnorm1, zbar);
    double gamma = d2norm(gambar, theta);
    S.cs2 = gambar / gamma;
    S.sn2 = theta / gamma;
    S.z = rhs / gamma;
    S.xnorm1 = d2norm(S.xnorm1, S.z);
    S.acond = S.anorm * S.dnorm;
    S.res2 = d2norm(S.res2, 0.0);
    S.rnorm = d2norm(S.res2, S.phibar);
    S.arnorm = alpha * fabs(tau);
    double test1 = S.rnorm / S.bnorm, test2 = 0.0;
    if (S.rnorm > 0.0) test2 = S.arnorm / (S.anorm * S.rnorm);
    double test3 = 1.0 / S.acond;
    double tt1 = test1 / (1.0 + S.anorm * S.xnorm / S.bnorm);
    double rtol = S.btol + S.atol * S.anorm * S.xnorm / S.bnorm;
    double tt3 = 1.0 + test3, tt2 = 1.0 + test2;
    tt1 = 1.0 + tt1;
    int istop = 0, itn = k - 1;
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
  // iteration k: anorm with alpha_k, beta_{k+1}; then alpha_{k+1} and the rotation
  double anew = beta > 0.0 ? sqrt(nvs) : alpha;
  double temp = d2norm(alpha, beta);
  temp = d2norm(temp, 0.0);
  S.anorm = d2norm(S.anorm, temp);
  double rhbar1 = S.rhobar;
  double rho = d2norm(rhbar1, beta);
  double cs = rhbar1 / rho, sn = beta / rho;
  double theta = sn * anew;
  S.rhobar = -cs * anew;
  double phi = cs * S.phibar;
  S.phibar = sn * S.phibar;
  double tau = sn * phi;
  S.t1 = phi / rho;
  S.t2 = -theta / rho;
  S.t3 = 1.0 / rho;
  S.pending_rho = rho, S.pending_phi = phi, S.pending_theta = theta, S.pending_tau = tau;
  S.alpha = anew;
  S.itn = k;
}

__global__ void to32_kernel(size_t n, const double *__restrict__ a, float *__restrict__ b) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) b[k] = (float)a[k];
}

__global__ void finish_kernel(int n, const double *__restrict__ q, const double *__restrict__ dx,
                              const double *__restrict__ x0, double *__restrict__ x) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  x[j] = q[j] * dx[j] + x0[j];
}

}  // namespace

// ---------------------------------------------------------------------------
void block_setup(BlockSys &s, int ncell, int R, int L, int n, const std::vector<int> &colmap, const double *Afull,
                 bool want32, hipStream_t st) {
  if (R > 64 || L > 64) throw std::runtime_error("block LSQR supports at most 64x64 blocks");
  s.ncell = ncell, s.R = R, s.L = L, s.n = n;
  // transpose gather lists in cell order
  std::vector<int> cnt(n + 1, 0), idx;
  for (size_t k = 0; k < colmap.size(); ++k)
    if (colmap[k] >= 0) cnt[colmap[k] + 1]++;
  for (int j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
  for (int j = 0; j < n; ++j)
    if (cnt[j + 1] == cnt[j]) throw std::runtime_error("an unknown appears in no equation");
  idx.resize(cnt[n]);
  std::vector<int> pos(cnt.begin(), cnt.end() - 1);
  for (size_t k = 0; k < colmap.size(); ++k)
    if (colmap[k] >= 0) idx[pos[colmap[k]]++] = (int)k;
  s.colmap.upload(colmap, st);
  s.tptr.upload(cnt, st);
  s.tidx.upload(idx, st);
  s.q.alloc(n);
  size_t total = (size_t)ncell * R * L;
  s.A.alloc(total);
  s.has32 = want32;
  if (want32) s.A32.alloc(total);
  hipLaunchKernelGGL(colnorm_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, R, L, s.tptr.p, s.tidx.p, Afull,
                     s.q.p);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(scale_kernel, dim3(ceil_div(total, 256)), dim3(256), 0, st, total, R, L, s.colmap.p, s.q.p,
                     Afull, s.A.p, want32 ? s.A32.p : nullptr);
  TT_LAUNCH_CHECK();
}

void block_make32(BlockSys &s, hipStream_t st) {
  if (s.has32) return;
  size_t total = (size_t)s.ncell * s.R * s.L;
  s.A32.alloc(total);
  hipLaunchKernelGGL(to32_kernel, dim3(ceil_div(total, 256)), dim3(256), 0, st, total, s.A.p, s.A32.p);
  TT_LAUNCH_CHECK();
  s.has32 = true;
}

LsqrProbe::~LsqrProbe() {
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

namespace {
// records the start event of a sampled launch; returns the pair index or -1
long probe_begin(LsqrProbe *p, hipStream_t st) {
  if (!p || p->stride <= 0) return -1;
  unsigned long long k = p->seen++;
  if (k % (unsigned long long)p->stride) return -1;
  if (2 * (p->used + 1) > p->pool.size()) {
    for (int i = 0; i < 64; ++i) {
      hipEvent_t e;
      TT_HIP(hipEventCreate(&e));
      p->pool.push_back(e);
    }
  }
  long idx = (long)p->used++;
  TT_HIP(hipEventRecord(p->pool[2 * idx], st));
  return idx;
}
void probe_end(LsqrProbe *p, long idx, hipStream_t st) {
  if (idx >= 0) TT_HIP(hipEventRecord(p->pool[2 * idx + 1], st));
}
// after the stream has been synchronised
void probe_collect(LsqrProbe *p) {
  if (!p) return;
  for (size_t i = 0; i < p->used; ++i) {
    float ms = 0.f;
    TT_HIP(hipEventElapsedTime(&ms, p->pool[2 * i], p->pool[2 * i + 1]));
    p->total_ms += ms;
    p->sampled++;
  }
  p->used = 0;
}
}  // namespace

void LsqrWork::ensure(int m_, int n_, int ncell_, int L_) {
  if (m_ > m) u.alloc(m_), m = m_;
  if (n_ > n) v.alloc(n_), w.alloc(n_), x.alloc(n_), n = n_;
  if ((size_t)ncell_ * L_ > P.n) P.alloc((size_t)ncell_ * L_);
  size_t np = std::max<size_t>((size_t)ncell_, ceil_div(n_, 256)) + 1;
  if (np > part1.n) part1.alloc(np), part2.alloc(np), part3.alloc(np);
  ncell = std::max(ncell, ncell_), L = std::max(L, L_);
  if (!state.p) {
    state.alloc(1);
    TT_HIP(hipHostMalloc((void **)&host_state, sizeof(LsqrState), hipHostMallocDefault));
  }
}

LsqrWork::~LsqrWork() {
  if (host_state) (void)hipHostFree(host_state);
}

LsqrResult block_lsqr(BlockSys &s, LsqrWork &wk, const double *b, double *x, int maxit, float eps4, bool use32,
                      hipStream_t st, LsqrProbe *probe) {
  const int m = s.ncell * s.R, n = s.n;
  wk.ensure(m, n, s.ncell, s.L);
  if (use32 && !s.has32) throw std::runtime_error("single-precision blocks not built");
  LsqrState init;
  std::memset(&init, 0, sizeof(init));
  init.cs2 = -1.0;
  init.atol = (double)eps4;
  init.ctol = 1.0 / 1.0e8;
  init.itnlim = maxit == 0 ? 10 * (m + n + 50) : maxit;
  TT_HIP(hipMemcpyAsync(wk.state.p, &init, sizeof(init), hipMemcpyHostToDevice, st));
  const unsigned gc = s.ncell, gn = ceil_div(n, 256);
  TT_HIP(hipMemsetAsync(wk.x.p, 0, sizeof(double) * n, st));
  const size_t lds64 = sizeof(double) * s.R * s.L, lds32 = sizeof(float) * s.R * s.L;
  hipLaunchKernelGGL(lsqr_r0<double>, dim3(gc), dim3(64), lds64, st, s.R, s.L, s.A.p, s.colmap.p, s.q.p, x, b, wk.u.p,
                     wk.P.p, wk.part1.p, wk.part2.p, wk.state.p);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(lsqr_B, dim3(gn), dim3(256), 0, st, n, 0, s.tptr.p, s.tidx.p, wk.P.p, wk.v.p, wk.w.p, wk.x.p,
                     wk.part1.p, wk.part2.p, wk.state.p);
  TT_LAUNCH_CHECK();
  // the previous solve's count predicts this one's (the Picard loop changes
  // only the right-hand side): termination is seen after nit + 1 launch pairs
  int k = 1, batch = s.last_nit > 0 ? s.last_nit + 1 : 4;
  LsqrResult res;
  for (;;) {
    for (int b2 = 0; b2 < batch; ++b2, ++k) {
      long pi = probe_begin(probe, st);
      if (use32)
        hipLaunchKernelGGL(lsqr_A<float>, dim3(gc), dim3(64), lds32, st, s.R, s.L, s.A32.p, s.colmap.p, wk.v.p, wk.u.p,
                           wk.P.p, wk.part1.p, wk.state.p);
      else
        hipLaunchKernelGGL(lsqr_A<double>, dim3(gc), dim3(64), lds64, st, s.R, s.L, s.A.p, s.colmap.p, wk.v.p, wk.u.p,
                           wk.P.p, wk.part1.p, wk.state.p);
      probe_end(probe, pi, st);
      hipLaunchKernelGGL(lsqr_B, dim3(gn), dim3(256), 0, st, n, k, s.tptr.p, s.tidx.p, wk.P.p, wk.v.p, wk.w.p, wk.x.p,
                         wk.part1.p, wk.part2.p, wk.state.p);
    }
    TT_LAUNCH_CHECK();
    TT_HIP(hipMemcpyAsync(wk.host_state, wk.state.p, sizeof(LsqrState), hipMemcpyDeviceToHost, st));
    TT_HIP(hipStreamSynchronize(st));
    probe_collect(probe);
    if (wk.host_state->done) {
      res.nit = wk.host_state->itn;
      res.istop = wk.host_state->istop;
      break;
    }
    if (k > init.itnlim + 2) throw std::runtime_error("LSQR did not signal termination");
    batch = k < 16 ? 4 : std::min(std::max(4, k / 4), 64);
  }
  s.last_nit = res.nit;
  // x = q * dx + x0 (x0 is still in x: the kernels only wrote wk.x)
  hipLaunchKernelGGL(finish_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, s.q.p, wk.x.p, x, x);
  TT_LAUNCH_CHECK();
  return res;
}

}  // namespace tt
