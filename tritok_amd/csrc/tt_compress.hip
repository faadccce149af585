// tt_compress.hip -- compressed operators on gfx950: cross approximation
// (lib_ACA.f), stretched-Haar wavelet compression (lib_WT.f) and block
// matrices of dense / wavelet / cross-approximated parts (lib_ctree.f).
//
// The reference's arithmetic is sequential; the kernels here keep its order
// wherever a result feeds a decision (pivot search, thresholds) or is cheap
// to keep: every wavelet coefficient, CSR row and dense row is summed by one
// thread in the reference's order with unfused multiply and add, and pivot
// searches break ties towards the reference's scan order.  Norm estimates
// (stopping tests) are parallel reductions and may differ in the last bits.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "tritok_bem.h"
#include "tt_compress.h"

// This file is compiled with -ffp-contract=off: products and sums round
// separately, as the reference computes them, so that pivots and thresholds
// see the same values as the CPU restatement.

namespace tt {

static int dvanak_host(int n) {
  if (n == 1) return 0;
  for (int i = 0; i <= 62; ++i)
    if ((1L << i) >= n) return i - 1;
  return -1;
}

HaarTab haar_tab(int nv) {
  if (nv < 1 || nv > (1 << 29)) throw std::runtime_error("wavelet length outside 1..2^29");
  HaarTab t{};
  t.nv = nv;
  t.kmax = dvanak_host(nv);
  t.N = 1 << (t.kmax + 1);
  t.podv = 2 * (nv - (1 << t.kmax)) + 1;
  t.norma = 1.0 / std::sqrt((double)t.N);
  for (int k = 0; k <= t.kmax; ++k) t.valcek[k] = std::pow(2.0, (double)k / 2) * t.norma;
  return t;
}

namespace {

constexpr int kNT = 256;

__device__ __forceinline__ int dvanak(int n) {
  if (n == 1) return 0;
  return 31 - __clz(n - 1);  // ceil(log2 n) - 1 for n >= 2
}

// element i (1-based) of the padded sequence: entries from podv on are
// repeated (dup) or followed by a zero (Che_WaWT column step)
template <class Get>
__device__ __forceinline__ double haar_pad(const HaarTab &t, int i, bool dup, Get get) {
  if (t.N == t.nv || i <= t.podv - 1) return get(i);
  int u = i - (t.podv - 1);
  int jj = t.podv - 1 + (u + 1) / 2;
  return (u & 1) ? get(jj) : (dup ? get(jj) : 0.0);
}

// coefficient ii (1-based) of the forward transform (Che_WxV, Che_WaWT)
template <class Get>
__device__ double haar_coef(const HaarTab &t, int ii, bool dup, Get get) {
  double acc = 0.0;
  if (ii == 1) {
    for (int i = 1; i <= t.N; ++i) acc = __dadd_rn(acc, __dmul_rn(haar_pad(t, i, dup, get), t.norma));
  } else {
    const int k = dvanak(ii), l = ii - (1 << k), d = 1 << (t.kmax - k);
    const int psp = (2 * l - 2) * d + 1, psr = (2 * l - 1) * d, pzg = 2 * l * d;
    const double vc = t.valcek[k];
    for (int i = psp; i <= psr; ++i) acc = __dadd_rn(acc, __dmul_rn(vc, haar_pad(t, i, dup, get)));
    for (int i = psr + 1; i <= pzg; ++i) acc = __dsub_rn(acc, __dmul_rn(vc, haar_pad(t, i, dup, get)));
  }
  return acc;
}

// entry j (1-based, 1..nv) of the inverse transform followed by the
// compaction of duplicated entries (Che_WTxV); get(jj) = coefficient jj
template <class Get>
__device__ double haar_inv(const HaarTab &t, int j, Get get) {
  const int ii = (t.N == t.nv || j <= t.podv - 1) ? j : 2 * j - t.podv + 1;
  double acc = __dmul_rn(get(1), t.norma);
  for (int k = 0; k <= t.kmax; ++k) {
    const double vc = t.valcek[k];
    const int d = 1 << (t.kmax - k);
    // iks=1+int((real(ii)-0.01)/real(DvaNaKmaxmk)) in single precision
    const float tf = __fdiv_rn(__fsub_rn((float)ii, 0.01f), (float)d);
    const int iks = 1 + (int)tf;
    if ((iks & 1) == 0) {
      const int jj = (1 << k) + iks / 2;
      if (jj <= t.nv) acc = __dsub_rn(acc, __dmul_rn(vc, get(jj)));
    } else {
      const int jj = (1 << k) + (iks + 1) / 2;
      if (jj <= t.nv) acc = __dadd_rn(acc, __dmul_rn(vc, get(jj)));
    }
  }
  return acc;
}

// (|value|, scan index) pairs: the larger value wins, equal values go to the
// smaller index (the first maximum in the reference's scan order)
__device__ __forceinline__ bool better(double v, long k, double bv, long bk) {
  return v > bv || (v == bv && k < bk);
}

// (row, column) of the search key of the largest |R|; a search that found no
// ordered value (an all-NaN residual) leaves the sentinel key, which maps to
// (0, 0) so that every later access stays inside the matrix
__device__ __forceinline__ void pivot_of(long key, bool rowmajor, int m, int n, int &i, int &j) {
  if (key < 0 || key >= (long)m * n) key = 0;
  i = rowmajor ? (int)(key / n) : (int)(key % m);
  j = rowmajor ? (int)(key % n) : (int)(key / m);
}

template <int NT>
__device__ void block_argmax(double &v, long &k, double *sv, long *sk) {
  const int tid = threadIdx.x;
  sv[tid] = v, sk[tid] = k;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s && better(sv[tid + s], sk[tid + s], sv[tid], sk[tid])) sv[tid] = sv[tid + s], sk[tid] = sk[tid + s];
    __syncthreads();
  }
  v = sv[0], k = sk[0];
  __syncthreads();
}

template <int NT>
__device__ double block_sum(double v, double *sd) {
  const int tid = threadIdx.x;
  sd[tid] = v;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s) sd[tid] += sd[tid + s];
    __syncthreads();
  }
  const double r = sd[0];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// wavelet compression kernels
__global__ void __launch_bounds__(kNT) wt_cols_kernel(const double *__restrict__ A, long lda, int m, HaarTab t,
                                                      bool dup, double *__restrict__ out, long ldo) {
  const int kk = blockIdx.x;
  const int ii = blockIdx.y * kNT + threadIdx.x + 1;
  if (ii > m) return;
  const double *col = A + (size_t)lda * kk;
  out[(ii - 1) + (size_t)ldo * kk] = haar_coef(t, ii, dup, [&](int j) { return col[j - 1]; });
}

__global__ void transpose_kernel(const double *__restrict__ in, long ldi, int rows, int cols, double *__restrict__ out,
                                 long ldo) {
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  for (int y = threadIdx.y; y < 32; y += 8) {
    const int r = bx + threadIdx.x, c = by + y;
    if (r < rows && c < cols) tile[y][threadIdx.x] = in[r + (size_t)ldi * c];
  }
  __syncthreads();
  for (int y = threadIdx.y; y < 32; y += 8) {
    const int c = by + threadIdx.x, r = bx + y;
    if (r < rows && c < cols) out[c + (size_t)ldo * r] = tile[threadIdx.x][y];
  }
}

__global__ void abs_sum_kernel(const double *__restrict__ a, size_t n, double *__restrict__ part) {
  __shared__ double sd[kNT];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (size_t)gridDim.x * kNT) s += fabs(a[i]);
  s = block_sum<kNT>(s, sd);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Che_Zanemari on row i (= column i of the transposed matrix Tt, n x m) by
// one wavefront; counts the entries Che_FormCRS will keep (the non-zeros).
__global__ void wt_threshold_kernel(double *__restrict__ Tt, int m, int n, double tresh, int *__restrict__ cnt,
                                    int *__restrict__ nz_total) {
  const int row = blockIdx.x * 4 + threadIdx.x / 64;  // 4 wavefronts per block
  const int lane = threadIdx.x & 63;
  if (row >= m) return;
  double *r = Tt + (size_t)n * row;
  int nz = 0, kept = 0, any = 0;
  if (row == 0) {  // the scaling row is kept whole
    for (int j = lane; j < n; j += 64) kept += r[j] != 0.0;
  } else {
    for (int j = 1 + lane; j <= n - 2; j += 64) {  // columns 2..n-1
      const double v = r[j];
      if (fabs(v) < tresh) {
        r[j] = 0.0;
        ++nz;
      } else {
        any = 1;
        kept += v != 0.0;
      }
    }
    any = __any(any);
    if (lane == 0) {
      if (n >= 2) kept += r[0] != 0.0;  // column 1 is kept
      double v = r[n - 1];              // column n: may go only if another entry stays
      if (any) {
        if (fabs(v) < tresh) {
          v = 0.0;
          ++nz;
        }
      } else if (v == 0.0) {
        v = 1.0e-14;
      }
      r[n - 1] = v;
      kept += v != 0.0;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    kept += __shfl_xor(kept, o);
    nz += __shfl_xor(nz, o);
  }
  if (lane == 0) {
    cnt[row] = kept;
    atomicAdd(nz_total, nz);
  }
}

// Che_FormCRS: entries of row `row` in column order (one wavefront per row)
__global__ void wt_fill_kernel(const double *__restrict__ Tt, int m, int n, const int *__restrict__ zvr,
                               int *__restrict__ c, double *__restrict__ v) {
  const int row = blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (row >= m) return;
  const double *r = Tt + (size_t)n * row;
  int pos = zvr[row] - 1;
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int j = j0 + lane;
    const double a = j < n ? r[j] : 0.0;
    const bool keep = a != 0.0;
    const unsigned long long mask = __ballot(keep);
    const int before = __popcll(mask & ((1ull << lane) - 1ull));
    if (keep) {
      c[pos + before] = j + 1;
      v[pos + before] = a;
    }
    pos += __popcll(mask);
  }
}



// the same coefficients with one wavefront per coefficient: its range split
// over the lanes, the lanes' sums added by a butterfly (lane 0's result)
__global__ void __launch_bounds__(kNT) wxv_wave_kernel(const double *__restrict__ x, HaarTab t,
                                                       double *__restrict__ xw) {
  const int ii = blockIdx.x * (kNT / 64) + (int)(threadIdx.x / 64) + 1, lane = threadIdx.x & 63;
  if (ii > t.nv) return;
  auto get = [&](int j) { return x[j - 1]; };
  double acc = 0.0;
  if (ii == 1) {
    for (int i = 1 + lane; i <= t.N; i += 64) acc += haar_pad(t, i, true, get) * t.norma;
  } else {
    const int k = dvanak(ii), l = ii - (1 << k), d = 1 << (t.kmax - k);
    const int psp = (2 * l - 2) * d + 1, psr = (2 * l - 1) * d, pzg = 2 * l * d;
    const double vc = t.valcek[k];
    for (int i = psp + lane; i <= psr; i += 64) acc += vc * haar_pad(t, i, true, get);
    for (int i = psr + 1 + lane; i <= pzg; i += 64) acc -= vc * haar_pad(t, i, true, get);
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) xw[ii - 1] = acc;
}

// 1-based CSR product with one wavefront per row (lane-strided entries)
__global__ void __launch_bounds__(kNT) csr_wave_kernel(int m, const int *__restrict__ zvr, const int *__restrict__ c,
                                                       const double *__restrict__ v, const double *__restrict__ x,
                                                       double *__restrict__ y) {
  const int r = blockIdx.x * (kNT / 64) + (int)(threadIdx.x / 64), lane = threadIdx.x & 63;
  if (r >= m) return;
  double acc = 0.0;
  for (int k = zvr[r] - 1 + lane; k < zvr[r + 1] - 1; k += 64) acc += v[k] * x[c[k] - 1];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) y[r] = acc;
}

// y = A x for a column-major A (lda) in column slices: part[ks * rows + i]
__global__ void __launch_bounds__(64) cm_mv_part_kernel(const double *__restrict__ A, long lda, int rows, int cols,
                                                        const double *__restrict__ x, int ks_cnt,
                                                        double *__restrict__ part) {
  const int i = blockIdx.x * 64 + threadIdx.x, ks = blockIdx.y;
  if (i >= rows) return;
  const int chunk = (cols + ks_cnt - 1) / ks_cnt, j0 = ks * chunk, j1 = min(cols, j0 + chunk);
  double acc = 0.0;
#pragma unroll 8
  for (int j = j0; j < j1; ++j) acc += A[i + (size_t)lda * j] * x[j];
  part[(size_t)ks * rows + i] = acc;
}

__global__ void part_sum_kernel(int rows, int ks_cnt, const double *__restrict__ part, double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  double s = 0.0;
  for (int k = 0; k < ks_cnt; ++k) s += part[(size_t)k * rows + i];
  y[i] = s;
}

int cm_split(int rows, int cols) {
  if (rows < 1 || cols < 1) return 1;
  const int rb = (rows + 63) / 64;
  return std::max(1, std::min((4096 + rb - 1) / rb, std::max(1, cols / 32)));
}

__global__ void __launch_bounds__(kNT) wtxv_kernel(const double *__restrict__ yw, HaarTab t, double *__restrict__ y) {
  const int j = blockIdx.x * kNT + threadIdx.x + 1;
  if (j <= t.nv) y[j - 1] = haar_inv(t, j, [&](int jj) { return yw[jj - 1]; });
}

// ---------------------------------------------------------------------------
// cross approximation, one matrix over the whole device (standalone entry)
struct AcaState {
  int pi, pj;
  double delta, snorm, norm, rms;
};

// optional residual update R -= a b (unfused), then the largest |R| per block
template <bool RowMajor>
__global__ void __launch_bounds__(kNT) aca_sweep_kernel(double *__restrict__ R, int m, int n,
                                                        const double *__restrict__ ak, const double *__restrict__ bk,
                                                        double *__restrict__ pv, long *__restrict__ pk) {
  __shared__ double sv[kNT];
  __shared__ long sk[kNT];
  double bv = -1.0;
  long bkey = 0x7fffffffffffffffL;
  const size_t total = (size_t)m * n;
  for (size_t L = (size_t)blockIdx.x * kNT + threadIdx.x; L < total; L += (size_t)gridDim.x * kNT) {
    const int i = (int)(L % m), j = (int)(L / m);
    double r = R[L];
    if (ak) {
      r = __dsub_rn(r, __dmul_rn(ak[i], bk[j]));
      R[L] = r;
    }
    const long key = RowMajor ? (long)i * n + j : (long)L;
    if (better(fabs(r), key, bv, bkey)) bv = fabs(r), bkey = key;
  }
  block_argmax<kNT>(bv, bkey, sv, sk);
  if (threadIdx.x == 0) pv[blockIdx.x] = bv, pk[blockIdx.x] = bkey;
}

template <bool RowMajor>
__global__ void __launch_bounds__(kNT) aca_pick_kernel(const double *__restrict__ R, int m, int n, int np,
                                                       const double *__restrict__ pv, const long *__restrict__ pk,
                                                       AcaState *st) {
  __shared__ double sv[kNT];
  __shared__ long sk[kNT];
  double bv = -1.0;
  long bkey = 0x7fffffffffffffffL;
  for (int p = threadIdx.x; p < np; p += kNT)
    if (better(pv[p], pk[p], bv, bkey)) bv = pv[p], bkey = pk[p];
  block_argmax<kNT>(bv, bkey, sv, sk);
  if (threadIdx.x == 0) {
    int i = 0, j = 0;
    pivot_of(bkey, RowMajor, m, n, i, j);
    st->pi = i, st->pj = j;
    st->delta = R[i + (size_t)m * j];
  }
}

// a_k = column of the pivot, b_k = row (types 1, 3: a scaled by 1/delta;
// type 2: b scaled by sigma = 1/delta); b vectors stored contiguously
__global__ void aca_extract_kernel(const double *__restrict__ R, int m, int n, int type, const AcaState *st,
                                   double *__restrict__ ak, double *__restrict__ bk) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int pi = st->pi, pj = st->pj;
  const double delta = st->delta;
  if (t < m) {
    const double c = R[t + (size_t)m * pj];
    ak[t] = (type == 2 || delta == 0.0) ? c : __ddiv_rn(c, delta);
  }
  if (t < n) {
    const double r = R[pi + (size_t)m * t];
    bk[t] = type == 2 ? __dmul_rn(r, __ddiv_rn(1.0, delta)) : r;
  }
}

// dot products of the newest pair with every stored one (EPS_ACA_NORM): block
// (l, 0) = a_l . a_k, block (l, 1) = b_l . b_k
__global__ void __launch_bounds__(kNT) aca_dots_kernel(const double *__restrict__ A, const double *__restrict__ B,
                                                       int m, int n, int k, double *__restrict__ dots) {
  __shared__ double sd[kNT];
  const int l = blockIdx.x, w = blockIdx.y;
  const double *u = w == 0 ? A + (size_t)m * l : B + (size_t)n * l;
  const double *v = w == 0 ? A + (size_t)m * k : B + (size_t)n * k;
  const int len = w == 0 ? m : n;
  double s = 0.0;
  for (int i = threadIdx.x; i < len; i += kNT) s += u[i] * v[i];
  s = block_sum<kNT>(s, sd);
  if (threadIdx.x == 0) dots[w * (k + 1) + l] = s;
}

__global__ void aca_norm_kernel(const double *__restrict__ dots, int k, AcaState *st) {
  if (threadIdx.x != 0) return;
  const double s1 = dots[k], s2 = dots[(k + 1) + k];
  double s3 = 0.0;
  for (int l = k; l >= 0; --l) s3 += dots[l];
  for (int l = k; l >= 0; --l) s3 += dots[(k + 1) + l];
  st->snorm = st->snorm + s3 + s1 * s2;
  st->norm = sqrt(s1) * sqrt(s2) / sqrt(st->snorm);
}

// type 3: t_k = b_k . vec, rez2 += a_k t_k, rms of rez1 - rez2
__global__ void __launch_bounds__(kNT) aca_rms_kernel(const double *__restrict__ ak, const double *__restrict__ bk,
                                                      int m, int n, const double *__restrict__ vec,
                                                      const double *__restrict__ rez1, double *__restrict__ rez2,
                                                      AcaState *st) {
  __shared__ double sd[kNT];
  __shared__ double tk;
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += kNT) s += bk[j] * vec[j];
  s = block_sum<kNT>(s, sd);
  if (threadIdx.x == 0) tk = s;
  __syncthreads();
  double e = 0.0, f = 0.0;
  for (int i = threadIdx.x; i < m; i += kNT) {
    const double r = __dadd_rn(rez2[i], __dmul_rn(ak[i], tk));
    rez2[i] = r;
    e += (rez1[i] - r) * (rez1[i] - r);
    f += rez1[i] * rez1[i];
  }
  e = block_sum<kNT>(e, sd);
  f = block_sum<kNT>(f, sd);
  if (threadIdx.x == 0) st->rms = f > 0.0 ? sqrt(e / f) : sqrt(e);
}

__global__ void __launch_bounds__(kNT) dense_mv_kernel(const double *__restrict__ A, long lda, int m, int n,
                                                       const double *__restrict__ x, double *__restrict__ y) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i >= m) return;
  double acc = 0.0;
  for (int j = 0; j < n; ++j) acc = __dadd_rn(acc, __dmul_rn(A[i + (size_t)lda * j], x[j]));
  y[i] = acc;
}

// b vectors (rank x n, vector-contiguous) -> column-major (rank x n)
__global__ void aca_bpack_kernel(const double *__restrict__ Bv, int rank, int n, double *__restrict__ b) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)rank * n) return;
  const int k = (int)(t % rank), j = (int)(t / rank);
  b[t] = Bv[(size_t)k * n + j];
}



// ---------------------------------------------------------------------------
// block matrices: per-part cross approximation (one workgroup per part)
struct AcaJob {
  int nr, nc, part;
  long r_off, a_off, b_off, v_off;
};

// deterministic test vector of the type-3 stopping rule for part p
__device__ __forceinline__ double test_vec(int p, int j) {
  unsigned h = (unsigned)p * 0x9E3779B1u ^ (unsigned)j * 0x85EBCA77u;
  h ^= h >> 15, h *= 0x2C1B3C6Du, h ^= h >> 12, h *= 0x297A2D39u, h ^= h >> 15;
  return (double)(h >> 8) * (1.0 / 16777216.0);
}

__global__ void __launch_bounds__(kNT) aca_parts_kernel(const AcaJob *__restrict__ jobs, int type, double eps,
                                                        double *__restrict__ Rp, double *__restrict__ Ap,
                                                        double *__restrict__ Bp, double *__restrict__ Vp,
                                                        int *__restrict__ ranks) {
  __shared__ double sv[kNT];
  __shared__ long sk[kNT];
  __shared__ int s_pi, s_pj, s_stop;
  __shared__ double s_snorm;
  const AcaJob jb = jobs[blockIdx.x];
  const int nr = jb.nr, nc = jb.nc, mn = nr < nc ? nr : nc, tid = threadIdx.x;
  double *R = Rp + jb.r_off, *A = Ap + jb.a_off, *B = Bp + jb.b_off;
  double *rez1 = Vp + jb.v_off, *rez2 = rez1 + nr;
  const bool rowmajor = type == 2;
  const int kfix = type == 1 ? (int)(eps * (double)mn) : mn;
  const size_t total = (size_t)nr * nc;
  if (type == 3)
    for (int i = tid; i < nr; i += kNT) {
      double acc = 0.0;
      for (int j = 0; j < nc; ++j) acc = __dadd_rn(acc, __dmul_rn(R[i + (size_t)nr * j], test_vec(jb.part, j)));
      rez1[i] = acc;
      rez2[i] = 0.0;
    }
  // first pivot
  double bv = -1.0;
  long bkey = 0x7fffffffffffffffL;
  for (size_t L = tid; L < total; L += kNT) {
    const int i = (int)(L % nr), j = (int)(L / nr);
    const long key = rowmajor ? (long)i * nc + j : (long)L;
    if (better(fabs(R[L]), key, bv, bkey)) bv = fabs(R[L]), bkey = key;
  }
  block_argmax<kNT>(bv, bkey, sv, sk);
  if (tid == 0) {
    pivot_of(bkey, rowmajor, nr, nc, s_pi, s_pj);
    s_snorm = 0.0;
  }
  __syncthreads();
  int rank = 0;
  for (int k = 0; k < kfix; ++k) {
    const int pi = s_pi, pj = s_pj;
    const double delta = R[pi + (size_t)nr * pj];
    // type 2 divides by the pivot: a zero pivot means the residual is zero
    // (it is the largest |R|), and the approximation ends without a new term
    if (type == 2 && delta == 0.0) break;
    double *ak = A + (size_t)nr * k, *bk = B + (size_t)nc * k;
    for (int i = tid; i < nr; i += kNT) {
      const double c = R[i + (size_t)nr * pj];
      ak[i] = (type == 2 || delta == 0.0) ? c : __ddiv_rn(c, delta);
    }
    for (int j = tid; j < nc; j += kNT) {
      const double r = R[pi + (size_t)nr * j];
      bk[j] = type == 2 ? __dmul_rn(r, __ddiv_rn(1.0, delta)) : r;
    }
    __syncthreads();
    rank = k + 1;
    bool stop = rank >= kfix;
    if (type == 2) {
      double s1 = 0.0, s2 = 0.0, s3 = 0.0;
      for (int i = tid; i < nr; i += kNT) {
        s1 += ak[i] * ak[i];
        for (int l = 0; l <= k; ++l) s3 += A[i + (size_t)nr * l] * ak[i];
      }
      for (int j = tid; j < nc; j += kNT) {
        s2 += bk[j] * bk[j];
        for (int l = 0; l <= k; ++l) s3 += B[j + (size_t)nc * l] * bk[j];
      }
      s1 = block_sum<kNT>(s1, sv);
      s2 = block_sum<kNT>(s2, sv);
      s3 = block_sum<kNT>(s3, sv);
      if (tid == 0) {
        s_snorm = s_snorm + s3 + s1 * s2;
        s_stop = !(sqrt(s1) * sqrt(s2) / sqrt(s_snorm) > eps);  // continue only while NORM > EPS
      }
      __syncthreads();
      stop = stop || s_stop;
    } else if (type == 3) {
      double t = 0.0;
      for (int j = tid; j < nc; j += kNT) t += bk[j] * test_vec(jb.part, j);
      t = block_sum<kNT>(t, sv);
      double e = 0.0, f = 0.0;
      for (int i = tid; i < nr; i += kNT) {
        const double r = __dadd_rn(rez2[i], __dmul_rn(ak[i], t));
        rez2[i] = r;
        e += (rez1[i] - r) * (rez1[i] - r);
        f += rez1[i] * rez1[i];
      }
      e = block_sum<kNT>(e, sv);
      f = block_sum<kNT>(f, sv);
      if (tid == 0) s_stop = (f > 0.0 ? sqrt(e / f) : sqrt(e)) < eps;
      __syncthreads();
      stop = stop || s_stop;
    }
    if (stop) break;
    bv = -1.0, bkey = 0x7fffffffffffffffL;
    for (size_t L = tid; L < total; L += kNT) {
      const int i = (int)(L % nr), j = (int)(L / nr);
      const double r = __dsub_rn(R[L], __dmul_rn(ak[i], bk[j]));
      R[L] = r;
      const long key = rowmajor ? (long)i * nc + j : (long)L;
      if (better(fabs(r), key, bv, bkey)) bv = fabs(r), bkey = key;
    }
    block_argmax<kNT>(bv, bkey, sv, sk);
    if (tid == 0) pivot_of(bkey, rowmajor, nr, nc, s_pi, s_pj);
    __syncthreads();
  }
  if (tid == 0) ranks[blockIdx.x] = rank;
}

// gather part blocks A(rows, cols) into per-part column-major storage
__global__ void __launch_bounds__(kNT) gather_parts_kernel(const HPart *__restrict__ parts, const int *__restrict__ rn,
                                                           const int *__restrict__ cn, const double *__restrict__ A,
                                                           long lda, const long *__restrict__ dst_off,
                                                           double *__restrict__ dst) {
  const HPart p = parts[blockIdx.x];
  const long o = dst_off[blockIdx.x];
  if (o < 0) return;
  const size_t total = (size_t)p.nr * p.nc;
  for (size_t L = threadIdx.x; L < total; L += kNT) {
    const int i = (int)(L % p.nr), j = (int)(L / p.nr);
    dst[o + L] = A[(size_t)(rn[p.roff + i] - 1) + (size_t)lda * (cn[p.coff + j] - 1)];
  }
}

// cross-approximation factors into the part pool: a (nr x rank), b (rank x nc)
__global__ void __launch_bounds__(kNT) aca_pack_parts_kernel(const HPart *__restrict__ parts,
                                                             const AcaJob *__restrict__ jobs,
                                                             const double *__restrict__ Ap,
                                                             const double *__restrict__ Bp, double *__restrict__ dpool) {
  const AcaJob jb = jobs[blockIdx.x];
  const HPart p = parts[jb.part];
  const int r = p.rank;
  double *a = dpool + p.doff, *b = a + (size_t)p.nr * r;
  for (size_t L = threadIdx.x; L < (size_t)p.nr * r; L += kNT) a[L] = Ap[jb.a_off + L];
  for (size_t L = threadIdx.x; L < (size_t)r * p.nc; L += kNT) {
    const int k = (int)(L % r), j = (int)(L / r);
    b[L] = Bp[jb.b_off + (size_t)p.nc * k + j];
  }
}

// product of every part with the gathered columns of X (nrhs <= 4): part
// results to prod[(poff + row) * 4 + rhs]
__global__ void __launch_bounds__(kNT) hmat_parts_kernel(const HPart *__restrict__ parts, const int *__restrict__ cn,
                                                         const double *__restrict__ dpool,
                                                         const int *__restrict__ ipool,
                                                         const HaarTab *__restrict__ tabs, double *__restrict__ wscr,
                                                         const double *__restrict__ X, long ldx, int nrhs,
                                                         double *__restrict__ prod) {
  const HPart p = parts[blockIdx.x];
  const int tid = threadIdx.x;
  const double *d = dpool + p.doff;
  const int *cols = cn + p.coff;
  auto xval = [&](int j, int r) { return X[(size_t)(cols[j] - 1) + (size_t)ldx * r]; };
  if (p.ctype == 1) {
    for (int o = tid; o < p.nr * nrhs; o += kNT) {
      const int i = o % p.nr, r = o / p.nr;
      double acc = 0.0;
      for (int j = 0; j < p.nc; ++j) acc = __dadd_rn(acc, __dmul_rn(d[i + (size_t)p.nr * j], xval(j, r)));
      prod[(p.poff + i) * 4 + r] = acc;
    }
  } else if (p.ctype == 3) {
    double *t = wscr + p.soff;
    const int rk = p.rank;
    const double *a = d, *b = d + (size_t)p.nr * rk;
    for (int o = tid; o < rk * nrhs; o += kNT) {
      const int k = o % rk, r = o / rk;
      double acc = 0.0;
      for (int j = 0; j < p.nc; ++j) acc = __dadd_rn(acc, __dmul_rn(b[k + (size_t)rk * j], xval(j, r)));
      t[k + rk * r] = acc;
    }
    __syncthreads();
    for (int o = tid; o < p.nr * nrhs; o += kNT) {
      const int i = o % p.nr, r = o / p.nr;
      double acc = 0.0;
      for (int k = 0; k < rk; ++k) acc = __dadd_rn(acc, __dmul_rn(a[i + (size_t)p.nr * k], t[k + rk * r]));
      prod[(p.poff + i) * 4 + r] = acc;
    }
  } else {
    const HaarTab &tc = tabs[2 * blockIdx.x], &tr = tabs[2 * blockIdx.x + 1];
    double *xw = wscr + p.soff, *yw = xw + (size_t)p.nc * nrhs;
    const int *zvr = ipool + p.ioff, *cc = zvr + p.nr + 1;
    for (int o = tid; o < p.nc * nrhs; o += kNT) {
      const int ii = o % p.nc + 1, r = o / p.nc;
      xw[(ii - 1) + (size_t)p.nc * r] = haar_coef(tc, ii, true, [&](int j) { return xval(j - 1, r); });
    }
    __syncthreads();
    for (int o = tid; o < p.nr * nrhs; o += kNT) {
      const int i = o % p.nr, r = o / p.nr;
      double acc = 0.0;
      for (int k = zvr[i]; k <= zvr[i + 1] - 1; ++k)
        acc = __dadd_rn(acc, __dmul_rn(d[k - 1], xw[(cc[k - 1] - 1) + (size_t)p.nc * r]));
      yw[i + (size_t)p.nr * r] = acc;
    }
    __syncthreads();
    for (int o = tid; o < p.nr * nrhs; o += kNT) {
      const int j = o % p.nr + 1, r = o / p.nr;
      prod[(p.poff + j - 1) * 4 + r] = haar_inv(tr, j, [&](int jj) { return yw[(jj - 1) + (size_t)p.nr * r]; });
    }
  }
}

// y(row) = sum over the parts holding the row, in part order (cTree_MxV)
__global__ void __launch_bounds__(kNT) hmat_rows_kernel(int nrow, const int *__restrict__ rowptr,
                                                        const int *__restrict__ rowent, const double *__restrict__ prod,
                                                        int nrhs, double *__restrict__ Y, long ldy) {
  const int r = blockIdx.x * kNT + threadIdx.x;
  if (r >= nrow) return;
  for (int c = 0; c < nrhs; ++c) {
    double acc = 0.0;
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) acc = __dadd_rn(acc, prod[(size_t)rowent[e] * 4 + c]);
    Y[r + (size_t)ldy * c] = acc;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
int wt_compress_dev(const double *A, long lda, int m, int n, double kappa, WtDev &w, hipStream_t s) {
  if (m < 1 || n < 1) throw std::runtime_error("empty matrix");
  const HaarTab tc = haar_tab(m), tr = haar_tab(n);
  w.m = m, w.n = n;
  DBuf<double> T1((size_t)m * n), Tt((size_t)m * n);
  // W on every column (entries of the padded column repeated), as Che_WaWT's first step
  hipLaunchKernelGGL(wt_cols_kernel, dim3(n, ceil_div(m, kNT)), dim3(kNT), 0, s, A, lda, m, tc, true, T1.p, (long)m);
  TT_LAUNCH_CHECK();
  // rows become columns: transform them with zero insertion into Tt's columns
  DBuf<double> T1t((size_t)m * n);
  hipLaunchKernelGGL(transpose_kernel, dim3(ceil_div(m, 32), ceil_div(n, 32)), dim3(32, 8), 0, s, T1.p, (long)m, m, n,
                     T1t.p, (long)n);
  TT_LAUNCH_CHECK();
  T1.release();
  hipLaunchKernelGGL(wt_cols_kernel, dim3(m, ceil_div(n, kNT)), dim3(kNT), 0, s, T1t.p, (long)n, n, tr, false, Tt.p,
                     (long)n);
  TT_LAUNCH_CHECK();
  T1t.release();
  // Che_AbsPovp: mean |a|
  const int nb = 1024;
  DBuf<double> part(nb);
  hipLaunchKernelGGL(abs_sum_kernel, dim3(nb), dim3(kNT), 0, s, Tt.p, (size_t)m * n, part.p);
  TT_LAUNCH_CHECK();
  std::vector<double> hp(nb);
  part.download(hp.data(), nb, s);
  TT_HIP(hipStreamSynchronize(s));
  double povp = 0.0;
  for (double v : hp) povp += v;
  povp = povp / (float)m / (float)n;
  const double tresh = kappa * povp;
  DBuf<int> cnt(m), nz(1);
  nz.zero(s);
  hipLaunchKernelGGL(wt_threshold_kernel, dim3(ceil_div(m, 4)), dim3(256), 0, s, Tt.p, m, n, tresh, cnt.p, nz.p);
  TT_LAUNCH_CHECK();
  std::vector<int> hc(m);
  int hnz = 0;
  cnt.download(hc.data(), m, s);
  nz.download(&hnz, 1, s);
  TT_HIP(hipStreamSynchronize(s));
  std::vector<int> zvr(m + 1);
  zvr[0] = 1;
  for (int i = 0; i < m; ++i) zvr[i + 1] = zvr[i] + hc[i];
  w.nnz = zvr[m] - 1;
  w.zvr.upload(zvr, s);
  w.c.alloc(std::max(w.nnz, 1));
  w.v.alloc(std::max(w.nnz, 1));
  hipLaunchKernelGGL(wt_fill_kernel, dim3(ceil_div(m, 4)), dim3(256), 0, s, Tt.p, m, n, w.zvr.p, w.c.p, w.v.p);
  TT_LAUNCH_CHECK();
  w.xw.alloc(n), w.yw.alloc(m);
  TT_HIP(hipStreamSynchronize(s));
  return m * n - hnz;
}

void wt_mv_dev(WtDev &w, const double *x, double *y, hipStream_t s) {
  const HaarTab tc = haar_tab(w.n), tr = haar_tab(w.m);
  hipLaunchKernelGGL(wxv_wave_kernel, dim3(ceil_div(w.n, kNT / 64)), dim3(kNT), 0, s, x, tc, w.xw.p);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(csr_wave_kernel, dim3(ceil_div(w.m, kNT / 64)), dim3(kNT), 0, s, w.m, w.zvr.p, w.c.p, w.v.p,
                     w.xw.p, w.yw.p);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(wtxv_kernel, dim3(ceil_div(w.m, kNT)), dim3(kNT), 0, s, w.yw.p, tr, y);
  TT_LAUNCH_CHECK();
}

void aca_product_buffers(AcaDev &h) {
  h.t.alloc(std::max(h.rank, 1));
  h.part.alloc(std::max({(size_t)cm_split(h.rank, h.n) * h.rank, (size_t)cm_split(h.m, h.rank) * h.m, (size_t)1}));
}

int aca_compress_dev(double *R, int m, int n, int type, double eps, const double *vec_host, AcaDev &out,
                     double *norm, hipStream_t s) {
  if (type < 1 || type > 3) throw std::runtime_error("ACA type must be 1, 2 or 3");
  if (type == 3 && !vec_host) throw std::runtime_error("ACA type 3 needs the test vector");
  const int mn = std::min(m, n);
  const int kfix = type == 1 ? (int)(eps * (double)mn) : mn;
  if (kfix < 0) throw std::runtime_error("ACA type 1 needs eps >= 0");
  DBuf<double> A((size_t)m * std::max(kfix, 1)), Bv((size_t)n * std::max(kfix, 1));
  DBuf<double> dvec, rez1, rez2, dots(2 * (size_t)std::max(kfix, 1));
  if (vec_host) {
    dvec.upload(vec_host, n, s);
    rez1.alloc(m), rez2.alloc(m);
    hipLaunchKernelGGL(dense_mv_kernel, dim3(ceil_div(m, kNT)), dim3(kNT), 0, s, R, (long)m, m, n, dvec.p, rez1.p);
    TT_LAUNCH_CHECK();
    rez2.zero(s);
  }
  const int np = 1024;
  DBuf<double> pv(np);
  DBuf<long> pk(np);
  DBuf<AcaState> st(1);
  AcaState h0{};
  st.upload(&h0, 1, s);
  auto sweep = [&](const double *ak, const double *bk) {
    if (type == 2) {
      hipLaunchKernelGGL(aca_sweep_kernel<true>, dim3(np), dim3(kNT), 0, s, R, m, n, ak, bk, pv.p, pk.p);
      TT_LAUNCH_CHECK();
      hipLaunchKernelGGL(aca_pick_kernel<true>, dim3(1), dim3(kNT), 0, s, R, m, n, np, pv.p, pk.p, st.p);
    } else {
      hipLaunchKernelGGL(aca_sweep_kernel<false>, dim3(np), dim3(kNT), 0, s, R, m, n, ak, bk, pv.p, pk.p);
      TT_LAUNCH_CHECK();
      hipLaunchKernelGGL(aca_pick_kernel<false>, dim3(1), dim3(kNT), 0, s, R, m, n, np, pv.p, pk.p, st.p);
    }
    TT_LAUNCH_CHECK();
  };
  int rank = 0;
  if (kfix > 0) sweep(nullptr, nullptr);
  for (int k = 0; k < kfix; ++k) {
    double *ak = A.p + (size_t)m * k, *bk = Bv.p + (size_t)n * k;
    if (type == 2) {  // a zero pivot: the residual is zero, nothing more to approximate
      AcaState h;
      st.download(&h, 1, s);
      TT_HIP(hipStreamSynchronize(s));
      if (h.delta == 0.0) break;
    }
    hipLaunchKernelGGL(aca_extract_kernel, dim3(ceil_div(std::max(m, n), kNT)), dim3(kNT), 0, s, R, m, n, type, st.p,
                       ak, bk);
    TT_LAUNCH_CHECK();
    rank = k + 1;
    bool stop = rank >= kfix;
    if (type == 2) {
      hipLaunchKernelGGL(aca_dots_kernel, dim3(k + 1, 2), dim3(kNT), 0, s, A.p, Bv.p, m, n, k, dots.p);
      TT_LAUNCH_CHECK();
      hipLaunchKernelGGL(aca_norm_kernel, dim3(1), dim3(64), 0, s, dots.p, k, st.p);
      TT_LAUNCH_CHECK();
    } else if (type == 3) {
      hipLaunchKernelGGL(aca_rms_kernel, dim3(1), dim3(kNT), 0, s, ak, bk, m, n, dvec.p, rez1.p, rez2.p, st.p);
      TT_LAUNCH_CHECK();
    }
    if (type != 1) {
      AcaState h;
      st.download(&h, 1, s);
      TT_HIP(hipStreamSynchronize(s));
      stop = stop || (type == 2 ? !(h.norm > eps) : h.rms < eps);  // type 2 goes on only while NORM > EPS
    }
    if (stop) break;
    sweep(ak, bk);
  }
  out.m = m, out.n = n, out.rank = rank;
  out.a.alloc((size_t)m * std::max(rank, 1));
  out.b.alloc((size_t)n * std::max(rank, 1));
  out.t.alloc(std::max(rank, 1));
  aca_product_buffers(out);
  if (rank > 0) {
    TT_HIP(hipMemcpyAsync(out.a.p, A.p, sizeof(double) * (size_t)m * rank, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(aca_bpack_kernel, dim3(ceil_div((size_t)rank * n, kNT)), dim3(kNT), 0, s, Bv.p, rank, n,
                       out.b.p);
    TT_LAUNCH_CHECK();
  }
  if (norm) {
    *norm = NAN;
    if (vec_host) {
      // getRMSnorm of A vec (rez1) against a (b vec)
      DBuf<double> y(m);
      aca_mv_dev(out, dvec.p, y.p, s);
      std::vector<double> r1(m), r2(m);
      rez1.download(r1.data(), m, s);
      y.download(r2.data(), m, s);
      TT_HIP(hipStreamSynchronize(s));
      double e = 0.0, f = 0.0;
      for (int i = 0; i < m; ++i) e += (r1[i] - r2[i]) * (r1[i] - r2[i]), f += r1[i] * r1[i];
      *norm = f > 0.0 ? std::sqrt(e / f) : std::sqrt(e);
    }
  }
  TT_HIP(hipStreamSynchronize(s));
  return rank;
}

void aca_mv_dev(AcaDev &h, const double *x, double *y, hipStream_t s) {
  if (h.rank == 0) {
    TT_HIP(hipMemsetAsync(y, 0, sizeof(double) * h.m, s));
    return;
  }
  // t = b x (rank x n), then y = a t (m x rank): column-slice products
  const int k1 = cm_split(h.rank, h.n), k2 = cm_split(h.m, h.rank);
  const size_t need = std::max((size_t)k1 * h.rank, (size_t)k2 * h.m);
  if (h.part.n < need) throw std::runtime_error("ACA product partials exceed their buffer");
  hipLaunchKernelGGL(cm_mv_part_kernel, dim3(ceil_div(h.rank, 64), k1), dim3(64), 0, s, h.b.p, (long)h.rank, h.rank,
                     h.n, x, k1, h.part.p);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(part_sum_kernel, dim3(ceil_div(h.rank, kNT)), dim3(kNT), 0, s, h.rank, k1, h.part.p, h.t.p);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(cm_mv_part_kernel, dim3(ceil_div(h.m, 64), k2), dim3(64), 0, s, h.a.p, (long)h.m, h.m, h.rank,
                     h.t.p, k2, h.part.p);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(part_sum_kernel, dim3(ceil_div(h.m, kNT)), dim3(kNT), 0, s, h.m, k2, h.part.p, y);
  TT_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// block partition by geometric bisection (no reference counterpart: the
// reference reads its cluster trees from an external .ctree file)
namespace {
struct Cluster {
  int begin, end;
  double lo[3], hi[3];
  int child[2] = {-1, -1};
};

void build_clusters(const double *x, long ld, std::vector<int> &idx, int begin, int end, int leaf,
                    std::vector<Cluster> &out, int self) {
  Cluster &c = out[self];
  c.begin = begin, c.end = end;
  for (int d = 0; d < 3; ++d) c.lo[d] = 1e300, c.hi[d] = -1e300;
  for (int k = begin; k < end; ++k)
    for (int d = 0; d < 3; ++d) {
      const double v = x[idx[k] + ld * d];
      c.lo[d] = std::min(c.lo[d], v), c.hi[d] = std::max(c.hi[d], v);
    }
  if (end - begin <= leaf) return;
  int ax = 0;
  for (int d = 1; d < 3; ++d)
    if (c.hi[d] - c.lo[d] > c.hi[ax] - c.lo[ax]) ax = d;
  const int mid = begin + (end - begin) / 2;
  std::nth_element(idx.begin() + begin, idx.begin() + mid, idx.begin() + end, [&](int a, int b) {
    const double va = x[a + ld * ax], vb = x[b + ld * ax];
    return va < vb || (va == vb && a < b);
  });
  const int c0 = (int)out.size();
  out.push_back(Cluster{});
  out.push_back(Cluster{});
  out[self].child[0] = c0, out[self].child[1] = c0 + 1;
  build_clusters(x, ld, idx, begin, mid, leaf, out, c0);
  build_clusters(x, ld, idx, mid, end, leaf, out, c0 + 1);
}

double diam(const Cluster &c) {
  double s = 0;
  for (int d = 0; d < 3; ++d) s += (c.hi[d] - c.lo[d]) * (c.hi[d] - c.lo[d]);
  return std::sqrt(s);
}

double box_dist(const Cluster &a, const Cluster &b) {
  double s = 0;
  for (int d = 0; d < 3; ++d) {
    const double g = std::max(0.0, std::max(a.lo[d] - b.hi[d], b.lo[d] - a.hi[d]));
    s += g * g;
  }
  return std::sqrt(s);
}

struct Partition {
  std::vector<int> rptr{0}, cptr{0}, rn, cn, admiss;
};

void block_parts(const std::vector<Cluster> &tr, const std::vector<int> &ir, int a, const std::vector<Cluster> &tc,
                 const std::vector<int> &ic, int b, double eta, Partition &P) {
  const Cluster &A = tr[a], &B = tc[b];
  const bool adm = std::min(diam(A), diam(B)) < eta * box_dist(A, B);  // cTree_Admis criterion
  const bool la = A.child[0] < 0, lb = B.child[0] < 0;
  if (adm || (la && lb)) {
    for (int k = A.begin; k < A.end; ++k) P.rn.push_back(ir[k] + 1);
    for (int k = B.begin; k < B.end; ++k) P.cn.push_back(ic[k] + 1);
    P.rptr.push_back((int)P.rn.size());
    P.cptr.push_back((int)P.cn.size());
    P.admiss.push_back(adm ? 1 : 0);
    return;
  }
  if (la) {
    for (int c : B.child) block_parts(tr, ir, a, tc, ic, c, eta, P);
  } else if (lb) {
    for (int c : A.child) block_parts(tr, ir, c, tc, ic, b, eta, P);
  } else {
    for (int ca : A.child)
      for (int cb : B.child) block_parts(tr, ir, ca, tc, ic, cb, eta, P);
  }
}

Partition make_partition(int nrow, const double *xr, long ldr, int ncol, const double *xc, long ldc, int leaf,
                         double eta) {
  std::vector<int> ir(nrow), ic(ncol);
  std::iota(ir.begin(), ir.end(), 0), std::iota(ic.begin(), ic.end(), 0);
  std::vector<Cluster> tr(1), tc(1);
  build_clusters(xr, ldr, ir, 0, nrow, leaf, tr, 0);
  build_clusters(xc, ldc, ic, 0, ncol, leaf, tc, 0);
  Partition P;
  block_parts(tr, ir, 0, tc, ic, 0, eta, P);
  return P;
}
}  // namespace

// cTree_FillMatrices: copy every part of the device matrix A (lda) and
// compress the admissible ones (ctype 2 wavelet, 3 cross approximation).
void hmat_build_dev(HmatDev &H, int nrow, int ncol, const double *A, long lda, int nparts, const int *rptr,
                    const int *rnodes, const int *cptr, const int *cnodes, const int *admiss, int ctype, int aca_type,
                    double aca_eps, double wt_kappa, hipStream_t s) {
  if (ctype < 1 || ctype > 3) throw std::runtime_error("compression type must be 1, 2 or 3");
  H.nrow = nrow, H.ncol = ncol, H.nparts = nparts;
  H.hparts.assign(nparts, HPart{});
  H.hrn.assign(rnodes, rnodes + rptr[nparts]);
  H.hcn.assign(cnodes, cnodes + cptr[nparts]);
  for (int v : H.hrn)
    if (v < 1 || v > nrow) throw std::runtime_error("row node outside 1..nrow");
  for (int v : H.hcn)
    if (v < 1 || v > ncol) throw std::runtime_error("column node outside 1..ncol");
  long poff = 0;
  for (int p = 0; p < nparts; ++p) {
    HPart &q = H.hparts[p];
    q.nr = rptr[p + 1] - rptr[p], q.nc = cptr[p + 1] - cptr[p];
    if (q.nr < 1 || q.nc < 1) throw std::runtime_error("empty matrix part");
    q.roff = rptr[p], q.coff = cptr[p];
    q.ctype = admiss[p] ? ctype : 1;
    q.poff = poff;
    poff += q.nr;
  }
  H.rn.upload(H.hrn, s);
  H.cn.upload(H.hcn, s);
  DBuf<HPart> dparts(nparts);
  dparts.upload(H.hparts, s);
  // gather: dense parts straight into a dense pool, compressed ones into a work pool
  std::vector<long> dense_off(nparts, -1), work_off(nparts, -1);
  long nd = 0, nw = 0;
  for (int p = 0; p < nparts; ++p) {
    const HPart &q = H.hparts[p];
    if (q.ctype == 1)
      dense_off[p] = nd, nd += (long)q.nr * q.nc;
    else
      work_off[p] = nw, nw += (long)q.nr * q.nc;
  }
  DBuf<double> dense(std::max(nd, 1L)), work(std::max(nw, 1L));
  DBuf<long> doffs(nparts);
  doffs.upload(dense_off, s);
  hipLaunchKernelGGL(gather_parts_kernel, dim3(nparts), dim3(kNT), 0, s, dparts.p, H.rn.p, H.cn.p, A, lda, doffs.p,
                     dense.p);
  TT_LAUNCH_CHECK();
  if (nw > 0) {
    doffs.upload(work_off, s);
    hipLaunchKernelGGL(gather_parts_kernel, dim3(nparts), dim3(kNT), 0, s, dparts.p, H.rn.p, H.cn.p, A, lda, doffs.p,
                       work.p);
    TT_LAUNCH_CHECK();
  }
  // cross approximation of every admissible part, one workgroup each
  std::vector<AcaJob> jobs;
  long na = 0, nb = 0, nv = 0;
  for (int p = 0; p < nparts; ++p) {
    const HPart &q = H.hparts[p];
    if (q.ctype != 3) continue;
    const int mn = std::min(q.nr, q.nc);
    const int kcap = aca_type == 1 ? std::max(0, (int)(aca_eps * (double)mn)) : mn;
    jobs.push_back(AcaJob{q.nr, q.nc, p, work_off[p], na, nb, nv});
    na += (long)q.nr * std::max(kcap, 1);
    nb += (long)q.nc * std::max(kcap, 1);
    nv += aca_type == 3 ? 2L * q.nr : 0;
  }
  std::vector<int> ranks(jobs.size());
  DBuf<double> Ap, Bp;
  DBuf<AcaJob> djobs;
  if (!jobs.empty()) {
    Ap.alloc(na), Bp.alloc(nb);
    DBuf<double> Vp(std::max(nv, 1L));
    djobs.upload(jobs, s);
    DBuf<int> dr(jobs.size());
    hipLaunchKernelGGL(aca_parts_kernel, dim3((unsigned)jobs.size()), dim3(kNT), 0, s, djobs.p, aca_type, aca_eps,
                       work.p, Ap.p, Bp.p, Vp.p, dr.p);
    TT_LAUNCH_CHECK();
    dr.download(ranks.data(), ranks.size(), s);
    TT_HIP(hipStreamSynchronize(s));
  }
  // wavelet compression of each admissible part
  std::vector<std::unique_ptr<WtDev>> wts(nparts);
  for (int p = 0; p < nparts; ++p) {
    const HPart &q = H.hparts[p];
    if (q.ctype != 2) continue;
    wts[p] = std::make_unique<WtDev>();
    wt_compress_dev(work.p + work_off[p], q.nr, q.nr, q.nc, wt_kappa, *wts[p], s);
  }
  // final pools
  long dsz = 0, isz = 0, ssz = 0;
  H.max_rank = 0;
  for (size_t k = 0; k < jobs.size(); ++k) H.hparts[jobs[k].part].rank = ranks[k];
  for (int p = 0; p < nparts; ++p) {
    HPart &q = H.hparts[p];
    q.doff = dsz, q.ioff = isz, q.soff = ssz;
    if (q.ctype == 1) {
      dsz += (long)q.nr * q.nc;
    } else if (q.ctype == 3) {
      dsz += (long)(q.nr + q.nc) * q.rank;
      ssz += 4L * std::max(q.rank, 1);
      H.max_rank = std::max(H.max_rank, q.rank);
    } else {
      dsz += wts[p]->nnz;
      isz += q.nr + 1 + wts[p]->nnz;
      ssz += 4L * (q.nr + q.nc);
    }
  }
  H.stored = dsz;
  H.dpool.alloc(std::max(dsz, 1L));
  H.ipool.alloc(std::max(isz, 1L));
  H.wscr.alloc(std::max(ssz, 1L));
  H.prod.alloc(std::max(poff, 1L) * 4);
  H.parts.upload(H.hparts, s);
  for (int p = 0; p < nparts; ++p) {
    const HPart &q = H.hparts[p];
    if (q.ctype == 1) {
      TT_HIP(hipMemcpyAsync(H.dpool.p + q.doff, dense.p + dense_off[p], sizeof(double) * q.nr * q.nc,
                            hipMemcpyDeviceToDevice, s));
    } else if (q.ctype == 2) {
      const WtDev &w = *wts[p];
      TT_HIP(hipMemcpyAsync(H.dpool.p + q.doff, w.v.p, sizeof(double) * w.nnz, hipMemcpyDeviceToDevice, s));
      TT_HIP(hipMemcpyAsync(H.ipool.p + q.ioff, w.zvr.p, sizeof(int) * (q.nr + 1), hipMemcpyDeviceToDevice, s));
      TT_HIP(hipMemcpyAsync(H.ipool.p + q.ioff + q.nr + 1, w.c.p, sizeof(int) * w.nnz, hipMemcpyDeviceToDevice, s));
    }
  }
  if (!jobs.empty()) {
    hipLaunchKernelGGL(aca_pack_parts_kernel, dim3((unsigned)jobs.size()), dim3(kNT), 0, s, H.parts.p, djobs.p, Ap.p,
                       Bp.p, H.dpool.p);
    TT_LAUNCH_CHECK();
  }
  std::vector<HaarTab> tabs(2 * (size_t)nparts);
  for (int p = 0; p < nparts; ++p)
    if (H.hparts[p].ctype == 2) tabs[2 * p] = haar_tab(H.hparts[p].nc), tabs[2 * p + 1] = haar_tab(H.hparts[p].nr);
  H.tabs.upload(tabs, s);
  // rows' contributions in part order
  std::vector<int> rowptr(nrow + 1, 0), rowent(poff);
  for (int p = 0; p < nparts; ++p)
    for (int i = 0; i < H.hparts[p].nr; ++i) ++rowptr[H.hrn[H.hparts[p].roff + i]];
  for (int r = 0; r < nrow; ++r) rowptr[r + 1] += rowptr[r];
  std::vector<int> fill(rowptr.begin(), rowptr.end() - 1);
  for (int p = 0; p < nparts; ++p)
    for (int i = 0; i < H.hparts[p].nr; ++i) {
      const int r = H.hrn[H.hparts[p].roff + i] - 1;
      rowent[fill[r]++] = (int)(H.hparts[p].poff + i);
    }
  H.rowptr.upload(rowptr, s);
  H.rowent.upload(rowent.empty() ? std::vector<int>{0} : rowent, s);
  TT_HIP(hipStreamSynchronize(s));
}

void hmat_mv_dev(HmatDev &H, int nrhs, const double *X, long ldx, double *Y, long ldy, hipStream_t s) {
  if (nrhs < 1 || nrhs > 4) throw std::runtime_error("1 to 4 right-hand sides");
  hipLaunchKernelGGL(hmat_parts_kernel, dim3(H.nparts), dim3(kNT), 0, s, H.parts.p, H.cn.p, H.dpool.p, H.ipool.p,
                     H.tabs.p, H.wscr.p, X, ldx, nrhs, H.prod.p);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(hmat_rows_kernel, dim3(ceil_div(H.nrow, kNT)), dim3(kNT), 0, s, H.nrow, H.rowptr.p, H.rowent.p,
                     H.prod.p, nrhs, Y, ldy);
  TT_LAUNCH_CHECK();
}

}  // namespace tt

using namespace tt;

// ---------------------------------------------------------------------------
// C-ABI
struct tt_aca {
  hipStream_t s = nullptr;
  AcaDev d;
  DBuf<double> x, y;
  ~tt_aca() {
    if (s) (void)hipStreamDestroy(s);
  }
};

struct tt_wt {
  hipStream_t s = nullptr;
  WtDev d;
  int nnz_ref = 0;
  DBuf<double> x, y;
  ~tt_wt() {
    if (s) (void)hipStreamDestroy(s);
  }
};

struct tt_hmat {
  hipStream_t s = nullptr;
  HmatDev d;
  DBuf<double> x, y;
  ~tt_hmat() {
    if (s) (void)hipStreamDestroy(s);
  }
};

extern "C" int tt_aca_compress(int32_t m, int32_t n, const double *full, int32_t type, double eps, const double *vec,
                               tt_aca **out, int32_t *rank, double *norm) {
  return guarded("tt_aca_compress", [&] {
    if (!full || !out) throw std::runtime_error("NULL argument");
    if (m < 1 || n < 1) throw std::runtime_error("empty matrix");
    auto h = std::make_unique<tt_aca>();
    TT_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
    DBuf<double> R;
    R.upload(full, (size_t)m * n, h->s);
    const int r = aca_compress_dev(R.p, m, n, type, eps, vec, h->d, norm, h->s);
    h->x.alloc(n), h->y.alloc(m);
    if (rank) *rank = r;
    *out = h.release();
    return 0;
  });
}

extern "C" int tt_aca_factors(const tt_aca *h, double *a, double *b) {
  return guarded("tt_aca_factors", [&] {
    if (!h || !a || !b) throw std::runtime_error("NULL argument");
    h->d.a.download(a, (size_t)h->d.m * h->d.rank, h->s);
    h->d.b.download(b, (size_t)h->d.n * h->d.rank, h->s);
    TT_HIP(hipStreamSynchronize(h->s));
    return 0;
  });
}

extern "C" int tt_aca_mv(tt_aca *h, const double *x, double *y) {
  return guarded("tt_aca_mv", [&] {
    if (!h || !x || !y) throw std::runtime_error("NULL argument");
    h->x.upload(x, h->d.n, h->s);
    aca_mv_dev(h->d, h->x.p, h->y.p, h->s);
    h->y.download(y, h->d.m, h->s);
    TT_HIP(hipStreamSynchronize(h->s));
    return 0;
  });
}

extern "C" int tt_aca_abxv(int32_t m, int32_t n, int32_t rank, const double *a, const double *b, const double *x,
                           double *y) {
  return guarded("tt_aca_abxv", [&] {
    if (!a || !b || !x || !y) throw std::runtime_error("NULL argument");
    if (m < 1 || n < 1 || rank < 0) throw std::runtime_error("bad sizes");
    AcaDev d;
    d.m = m, d.n = n, d.rank = rank;
    d.a.upload(a, (size_t)m * rank), d.b.upload(b, (size_t)rank * n);
    aca_product_buffers(d);
    DBuf<double> dx, dy(m);
    dx.upload(x, n);
    aca_mv_dev(d, dx.p, dy.p, 0);
    dy.download(y, m);
    TT_HIP(hipDeviceSynchronize());
    return 0;
  });
}

extern "C" void tt_aca_free(tt_aca *h) { delete h; }

extern "C" int tt_wt_compress(int32_t m, int32_t n, const double *full, double kappa, tt_wt **out, int32_t *nnz) {
  return guarded("tt_wt_compress", [&] {
    if (!full || !out) throw std::runtime_error("NULL argument");
    if (m < 1 || n < 1) throw std::runtime_error("empty matrix");
    if (!(kappa >= 0.0)) throw std::runtime_error("kappa must be >= 0");
    auto h = std::make_unique<tt_wt>();
    TT_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
    DBuf<double> A;
    A.upload(full, (size_t)m * n, h->s);
    h->nnz_ref = wt_compress_dev(A.p, m, m, n, kappa, h->d, h->s);
    h->x.alloc(n), h->y.alloc(m);
    if (nnz) *nnz = h->d.nnz;
    *out = h.release();
    return 0;
  });
}

extern "C" int tt_wt_csr(const tt_wt *h, int32_t *zvr, int32_t *c, double *v) {
  return guarded("tt_wt_csr", [&] {
    if (!h || !zvr || !c || !v) throw std::runtime_error("NULL argument");
    h->d.zvr.download(zvr, h->d.m + 1, h->s);
    h->d.c.download(c, h->d.nnz, h->s);
    h->d.v.download(v, h->d.nnz, h->s);
    TT_HIP(hipStreamSynchronize(h->s));
    return h->nnz_ref;
  });
}

extern "C" int tt_wt_mv(tt_wt *h, const double *x, double *y) {
  return guarded("tt_wt_mv", [&] {
    if (!h || !x || !y) throw std::runtime_error("NULL argument");
    h->x.upload(x, h->d.n, h->s);
    wt_mv_dev(h->d, h->x.p, h->y.p, h->s);
    h->y.download(y, h->d.m, h->s);
    TT_HIP(hipStreamSynchronize(h->s));
    return 0;
  });
}

extern "C" int tt_wt_wtawx(int32_t nrow, int32_t ncol, int32_t nnz, const int32_t *zvr, const int32_t *c,
                           const double *v, const double *x, double *y) {
  return guarded("tt_wt_wtawx", [&] {
    if (!zvr || !c || !v || !x || !y) throw std::runtime_error("NULL argument");
    if (nrow < 1 || ncol < 1 || nnz < 0) throw std::runtime_error("bad sizes");
    if (zvr[0] != 1 || zvr[nrow] - 1 != nnz) throw std::runtime_error("zvr does not describe nnz entries");
    for (int k = 0; k < nnz; ++k)
      if (c[k] < 1 || c[k] > ncol) throw std::runtime_error("column index outside 1..ncol");
    WtDev w;
    w.m = nrow, w.n = ncol, w.nnz = nnz;
    w.zvr.upload(zvr, nrow + 1), w.c.upload(c, std::max(nnz, 1)), w.v.upload(v, std::max(nnz, 1));
    w.xw.alloc(ncol), w.yw.alloc(nrow);
    DBuf<double> dx, dy(nrow);
    dx.upload(x, ncol);
    wt_mv_dev(w, dx.p, dy.p, 0);
    dy.download(y, nrow);
    TT_HIP(hipDeviceSynchronize());
    return 0;
  });
}

extern "C" void tt_wt_free(tt_wt *h) { delete h; }

static tt_hmat *hmat_from(int32_t nrow, int32_t ncol, const double *full, int64_t lda, int32_t nparts,
                          const int32_t *rptr, const int32_t *rn, const int32_t *cptr, const int32_t *cn,
                          const int32_t *admiss, int32_t ctype, int32_t aca_type, double aca_eps, double wt_kappa) {
  auto h = std::make_unique<tt_hmat>();
  TT_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
  DBuf<double> A((size_t)lda * ncol);
  TT_HIP(hipMemcpyAsync(A.p, full, sizeof(double) * (size_t)lda * ncol, hipMemcpyHostToDevice, h->s));
  hmat_build_dev(h->d, nrow, ncol, A.p, lda, nparts, rptr, rn, cptr, cn, admiss, ctype, aca_type, aca_eps, wt_kappa,
                 h->s);
  h->x.alloc((size_t)ncol * 4), h->y.alloc((size_t)nrow * 4);
  return h.release();
}

namespace tt {
// A block matrix of a device-resident matrix A (nrow x ncol, leading
// dimension lda, left unchanged) with the library's own partition of the row
// points xr and column points xc (host arrays, n x 3 column-major).
tt_hmat *hmat_from_device(int nrow, int ncol, const double *A, long lda, const double *xr, const double *xc,
                          int leaf, double eta, int ctype, int aca_type, double aca_eps, double wt_kappa) {
  if (nrow < 1 || ncol < 1 || lda < nrow || leaf < 1 || !(eta > 0.0)) throw std::runtime_error("bad sizes");
  Partition P = make_partition(nrow, xr, nrow, ncol, xc, ncol, leaf, eta);
  auto h = std::make_unique<tt_hmat>();
  TT_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
  hmat_build_dev(h->d, nrow, ncol, A, lda, (int)P.admiss.size(), P.rptr.data(), P.rn.data(), P.cptr.data(),
                 P.cn.data(), P.admiss.data(), ctype, aca_type, aca_eps, wt_kappa, h->s);
  h->x.alloc((size_t)ncol * 4), h->y.alloc((size_t)nrow * 4);
  return h.release();
}

void hmat_apply(tt_hmat *h, const double *x, double *y, hipStream_t s) {
  hmat_mv_dev(h->d, 1, x, (long)h->d.ncol, y, (long)h->d.nrow, s);
}

long hmat_stored_entries(const tt_hmat *h) { return h->d.stored; }
}  // namespace tt

extern "C" int tt_hmat_create(int32_t nrow, int32_t ncol, const double *full, int64_t lda, const double *xr,
                              const double *xc, int32_t leaf, double eta, int32_t ctype, int32_t aca_type,
                              double aca_eps, double wt_kappa, tt_hmat **out) {
  return guarded("tt_hmat_create", [&] {
    if (!full || !xr || !xc || !out) throw std::runtime_error("NULL argument");
    if (nrow < 1 || ncol < 1 || lda < nrow || leaf < 1 || !(eta > 0.0)) throw std::runtime_error("bad sizes");
    Partition P = make_partition(nrow, xr, nrow, ncol, xc, ncol, leaf, eta);
    *out = hmat_from(nrow, ncol, full, lda, (int)P.admiss.size(), P.rptr.data(), P.rn.data(), P.cptr.data(),
                     P.cn.data(), P.admiss.data(), ctype, aca_type, aca_eps, wt_kappa);
    return 0;
  });
}

extern "C" int tt_hmat_partition(int32_t nrow, const double *xr, int32_t ncol, const double *xc, int32_t leaf,
                                 double eta, int32_t *nparts, int64_t *rnodes_len, int64_t *cnodes_len,
                                 int32_t *rptr, int32_t *rnodes, int32_t *cptr, int32_t *cnodes, int32_t *admiss) {
  return guarded("tt_hmat_partition", [&] {
    if (!xr || !xc) throw std::runtime_error("NULL argument");
    if (nrow < 1 || ncol < 1 || leaf < 1 || !(eta > 0.0)) throw std::runtime_error("bad sizes");
    const Partition P = make_partition(nrow, xr, nrow, ncol, xc, ncol, leaf, eta);
    if (nparts) *nparts = (int32_t)P.admiss.size();
    if (rnodes_len) *rnodes_len = (int64_t)P.rn.size();
    if (cnodes_len) *cnodes_len = (int64_t)P.cn.size();
    if (rptr) std::copy(P.rptr.begin(), P.rptr.end(), rptr);
    if (cptr) std::copy(P.cptr.begin(), P.cptr.end(), cptr);
    if (rnodes) std::copy(P.rn.begin(), P.rn.end(), rnodes);
    if (cnodes) std::copy(P.cn.begin(), P.cn.end(), cnodes);
    if (admiss) std::copy(P.admiss.begin(), P.admiss.end(), admiss);
    return 0;
  });
}

extern "C" int tt_hmat_create_parts(int32_t nrow, int32_t ncol, const double *full, int64_t lda, int32_t nparts,
                                    const int32_t *rptr, const int32_t *rnodes, const int32_t *cptr,
                                    const int32_t *cnodes, const int32_t *admiss, int32_t ctype, int32_t aca_type,
                                    double aca_eps, double wt_kappa, tt_hmat **out) {
  return guarded("tt_hmat_create_parts", [&] {
    if (!full || !rptr || !rnodes || !cptr || !cnodes || !admiss || !out) throw std::runtime_error("NULL argument");
    if (nrow < 1 || ncol < 1 || lda < nrow || nparts < 1) throw std::runtime_error("bad sizes");
    if (rptr[0] != 0 || cptr[0] != 0) throw std::runtime_error("rptr/cptr must start at 0");
    *out = hmat_from(nrow, ncol, full, lda, nparts, rptr, rnodes, cptr, cnodes, admiss, ctype, aca_type, aca_eps,
                     wt_kappa);
    return 0;
  });
}

extern "C" int tt_hmat_info(const tt_hmat *h, int32_t *nparts, int64_t *rnodes_len, int64_t *cnodes_len,
                            int64_t *stored, int32_t *max_rank) {
  return guarded("tt_hmat_info", [&] {
    if (!h) throw std::runtime_error("NULL argument");
    if (nparts) *nparts = h->d.nparts;
    if (rnodes_len) *rnodes_len = (int64_t)h->d.hrn.size();
    if (cnodes_len) *cnodes_len = (int64_t)h->d.hcn.size();
    if (stored) *stored = h->d.stored;
    if (max_rank) *max_rank = h->d.max_rank;
    return 0;
  });
}

extern "C" int tt_hmat_parts(const tt_hmat *h, int32_t *rptr, int32_t *rnodes, int32_t *cptr, int32_t *cnodes,
                             int32_t *ctype, int32_t *rank) {
  return guarded("tt_hmat_parts", [&] {
    if (!h) throw std::runtime_error("NULL argument");
    const auto &P = h->d.hparts;
    for (int p = 0; p < h->d.nparts; ++p) {
      if (rptr) rptr[p] = (int)P[p].roff, rptr[p + 1] = (int)(P[p].roff + P[p].nr);
      if (cptr) cptr[p] = (int)P[p].coff, cptr[p + 1] = (int)(P[p].coff + P[p].nc);
      if (ctype) ctype[p] = P[p].ctype;
      if (rank) rank[p] = P[p].rank;
    }
    if (rnodes) std::copy(h->d.hrn.begin(), h->d.hrn.end(), rnodes);
    if (cnodes) std::copy(h->d.hcn.begin(), h->d.hcn.end(), cnodes);
    return 0;
  });
}

extern "C" int tt_hmat_part_data(const tt_hmat *h, int32_t part, double *d, int32_t *iv) {
  return guarded("tt_hmat_part_data", [&] {
    if (!h || part < 0 || part >= h->d.nparts) throw std::runtime_error("bad part");
    const HPart &q = h->d.hparts[part];
    long nd = 0;
    if (q.ctype == 1) {
      nd = (long)q.nr * q.nc;
    } else if (q.ctype == 3) {
      nd = (long)(q.nr + q.nc) * q.rank;
    } else {
      int z = 0;
      TT_HIP(hipMemcpy(&z, h->d.ipool.p + q.ioff + q.nr, sizeof(int), hipMemcpyDeviceToHost));
      nd = z - 1;
    }
    if (d && nd > 0) TT_HIP(hipMemcpy(d, h->d.dpool.p + q.doff, sizeof(double) * nd, hipMemcpyDeviceToHost));
    if (iv && q.ctype == 2)
      TT_HIP(hipMemcpy(iv, h->d.ipool.p + q.ioff, sizeof(int) * (q.nr + 1 + nd), hipMemcpyDeviceToHost));
    return (int)nd;
  });
}

extern "C" int tt_hmat_mv(tt_hmat *h, int32_t nrhs, const double *x, double *y) {
  return guarded("tt_hmat_mv", [&] {
    if (!h || !x || !y) throw std::runtime_error("NULL argument");
    if (nrhs < 1 || nrhs > 4) throw std::runtime_error("1 to 4 right-hand sides");
    const size_t nc = h->d.ncol, nr = h->d.nrow;
    h->x.upload(x, nc * nrhs, h->s);
    hmat_mv_dev(h->d, nrhs, h->x.p, (long)nc, h->y.p, (long)nr, h->s);
    h->y.download(y, nr * nrhs, h->s);
    TT_HIP(hipStreamSynchronize(h->s));
    return 0;
  });
}

extern "C" void tt_hmat_free(tt_hmat *h) { delete h; }
