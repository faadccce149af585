// tt_lsqr.hip -- LSQR on the device for the block-structured subdomain systems.
//
// Per iteration two kernels, no host synchronisation, no atomics and no
// device-scope fences:
//   lsqr_A (one wavefront per cell): reduces the unknown kernel's partial
//           sums (the same terms in the same order in every block, so every
//           block obtains the same alpha), runs the scalar recurrences
//           (stopping tests of iteration k-2, plane rotation of iteration
//           k-1), then u' <- A v - alpha u, ||u'||^2 partial and P <- A^T u'
//           for its cell.  The block is read once from HBM with 16-byte
//           loads, staged in LDS and used for both products (FP64 arithmetic on FP64 blocks; FP32 arithmetic with two
//           accumulators on the FP32 blocks of the mixed mode, whose solves are
//           anchored by an FP64 initial residual).  Block 0 stores the scalars.
//   lsqr_B (one thread per unknown): reduces the cell partials to beta,
//           finishes the x/w update of the previous iteration, gathers
//           A^T u' from the cells of its column (fixed-width index table,
//           loads issued together), v' <- A^T u / beta - beta v, and writes
//           the partial sums of ||v'||^2 and of the step norm.
// Results are therefore bit-for-bit repeatable.  The scalar recurrences are
// those of the reference LSQR (lib_Solve_v11.f:3283-3420) with damp = 0; they
// run one kernel later than in a sequential code, so termination at
// iteration k is seen by the cell kernel of iteration k+2, after the unknown
// kernel of iteration k+1 has applied the x update of iteration k and
// nothing more.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include <hip/hip_ext.h>

#include "tt_lsqr.h"
#include "tt_lsqr_dev.h"

namespace tt {

namespace {

using namespace lsqr_dev;


// value of v in lane `lane` of the wavefront, as a scalar (no LDS traffic)
__device__ __forceinline__ double lane_value(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <typename T>
struct Vec16;
template <>
struct Vec16<double> {
  using type = double2;
};
template <>
struct Vec16<float> {
  using type = float4;
};
template <>
struct Vec16<_Float16> {
  using type = uint4;  // eight halves, moved as bits
};

// padded block length of an R x L block in element type T (16-byte multiple;
// 16-bit blocks have rows padded to row_len, a multiple of four halves)
template <typename T, int L>
constexpr int row_len() {
  return sizeof(T) == 2 ? (L + 3) / 4 * 4 : L;
}
template <typename T, int R, int L>
constexpr int block_len() {
  return sizeof(T) == 2 ? (R * row_len<T, L>() + 7) / 8 * 8 : (R * L + 3) / 4 * 4;
}

// Loads one padded block into LDS: all 16-byte loads of a lane are issued
// before the first LDS store.
template <typename T, int S>
__device__ __forceinline__ void stage_block(const T *__restrict__ src, T *blk) {
  using V = typename Vec16<T>::type;
  constexpr int NV = S * (int)sizeof(T) / 16, PER = (NV + 63) / 64;
  const V *s = reinterpret_cast<const V *>(src);
  V *d = reinterpret_cast<V *>(blk);
  V reg[PER];
  const int t = threadIdx.x;
#pragma unroll
  for (int i = 0; i < PER; ++i) reg[i] = (t + 64 * i < NV) ? s[t + 64 * i] : V{};
#pragma unroll
  for (int i = 0; i < PER; ++i)
    if (t + 64 * i < NV) d[t + 64 * i] = reg[i];
}

// A block held in registers between its HBM load and its LDS store, so that
// a wavefront streams the next cell's block while it computes on this one.
template <typename T, int S>
struct RegBlock {
  using V = typename Vec16<T>::type;
  static constexpr int NV = S * (int)sizeof(T) / 16, PER = (NV + 63) / 64;
  V r[PER];
  __device__ __forceinline__ void load(const T *__restrict__ src) {
    const V *s = reinterpret_cast<const V *>(src);
    const int t = threadIdx.x & 63;  // lane: one wavefront moves one block
#pragma unroll
    for (int i = 0; i < PER; ++i) r[i] = (t + 64 * i < NV) ? s[t + 64 * i] : V{};
  }
  __device__ __forceinline__ void store(T *blk) const {
    V *d = reinterpret_cast<V *>(blk);
    const int t = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < PER; ++i)
      if (t + 64 * i < NV) d[t + 64 * i] = r[i];
  }
};

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

// scaled, padded blocks from the unpadded full blocks (known columns -> 0)
__global__ void scale_kernel(size_t total, int R, int L, int S, const int *__restrict__ colmap,
                             const double *__restrict__ q, const double *__restrict__ Afull, double *__restrict__ A,
                             float *__restrict__ A32) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  size_t c = k / S;
  int e = (int)(k % S);
  double v = 0.0;
  if (e < R * L) {
    int l = e % L;
    int j = colmap[c * L + l];
    if (j >= 0) v = Afull[c * R * L + e] * q[j];
  }
  A[k] = v;
  if (A32) A32[k] = (float)v;
}

__global__ void to32_kernel(size_t n, const double *__restrict__ a, float *__restrict__ b) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) b[k] = (float)a[k];
}

// FP64 blocks (stride S, rows of L) -> FP16 blocks (rows padded to LP,
// stride S16 >= R*LP), padding zero
__global__ void to16_kernel(size_t total16, int R, int L, int LP, int S, int S16, const double *__restrict__ a,
                            _Float16 *__restrict__ b) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total16) return;
  const size_t c = k / S16;
  const int e = (int)(k % S16), r = e / LP, l = e % LP;
  b[k] = (r < R && l < L) ? (_Float16)a[c * S + (size_t)r * L + l] : (_Float16)0.0f;
}

constexpr int kNB = 256;  // threads per block of the unknown kernel
// cells per partial sum of the cell kernel: at most 4096 partials per launch,
// a fixed function of the system so that the sums do not depend on the grid
// (FP64 blocks, one wavefront per workgroup); with nw wavefronts per workgroup
// (FP32 and FP16 blocks) a multiple of nw cells and at most 1024 partials
constexpr int kNW = 4;  // wavefronts per workgroup of the reduced-precision cell kernel
inline int cell_group(int ncell, int nw) {
  if (nw == 1) return std::max(1, (ncell + 4095) / 4096);
  const int per = std::max(1, (ncell + 1023) / 1024);
  return (per + nw - 1) / nw * nw;
}
inline int cell_waves(int prec) { return prec == kBlocks64 ? 1 : kNW; }

// The arguments of one system's cell kernel and unknown kernel.  A kernel
// body takes them with its workgroup index and workgroup count, so that one
// launch can run the iterations of several systems (workgroup ranges per
// system, see Jobs) with the arithmetic of one launch per system.
struct CellArgs {
  int ncell, cgroup, nb;
  const void *A;  // blocks in the kernel's element type
  const int *colmap;
  const double *vp, *pb1, *pb2;
  double *u, *P, *pa1;
  const LsqrState *sin;
  LsqrState *sout;
};
struct R0Args {  // initial residual of one system
  const double *A, *q, *x0, *b;
  const int *colmap;
  double *u, *P, *pa1, *pa2;
  int ncell;
};
struct InitArgs {  // the state copies and dx = 0 of one system's solve
  LsqrState *st;
  double *x;
  double atol;
  int n, itnlim;
};
struct FinArgs {  // x = q * dx + x0 of one system
  const double *q, *dx, *x0;
  double *x;
  int n;
};
struct UnkArgs {
  int n, na;
  const int *gtab;
  const void *P;  // partials in the kernel's type
  double *vp, *w, *x;
  const double *pa1, *pa2;
  double *pb1, *pb2;
  LsqrState *st;
};
// systems of one batched launch: system i runs on workgroups [first[i], first[i+1])
constexpr int kMaxJobs = 32;
template <typename Args>
struct Jobs {
  Args a[kMaxJobs];
  int first[kMaxJobs + 1];
  int n;
  // il > 0 (cell kernels of systems that read one shared block array, each
  // with gs workgroups): workgroups interleaved so that the systems'
  // workgroups of one logical index l are dispatched together on one XCD
  // (l % 8): b = 8*il*(l/8) + 8*i + l%8 for system i
  int il = 0, gs = 0;
};
// the system of workgroup b (uniform over the workgroup)
template <typename Args>
__device__ __forceinline__ int job_of(const Jobs<Args> &J, int b) {
  int i = 0;
  while (i + 1 < J.n && b >= J.first[i + 1]) ++i;
  return i;
}

// (the pointers are parameters of their own so that they keep their
// no-alias property when the body is inlined into a launch wrapper)
// ---------------------------------------------------------------------------
// r0 = b - A (x0/q), P = A^T r0; per-cell partials of ||r0||^2 and ||b||^2
template <int R, int L>
__device__ __forceinline__ void cell_r0(const double *__restrict__ A, const int *__restrict__ colmap,
                                        const double *__restrict__ q, const double *__restrict__ x0,
                                        const double *__restrict__ b, double *__restrict__ u, double *__restrict__ P,
                                        double *__restrict__ pa1, double *__restrict__ pa2, const int c) {
  constexpr int S = (R * L + 3) / 4 * 4;
  __shared__ __align__(16) double blk[S];
  __shared__ double xs[64], rr[64];
  const int t = threadIdx.x;
  stage_block<double, S>(A + (size_t)c * S, blk);
  if (t < L) {
    int j = colmap[(size_t)c * L + t];
    xs[t] = j >= 0 ? x0[j] / q[j] : 0.0;
  }
  const double bb = t < R ? b[(size_t)c * R + t] : 0.0;
  __syncthreads();
  double r = 0.0;
  if (t < R) {
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < L; ++l) s += blk[t * L + l] * xs[l];
    r = bb - s;
    u[(size_t)c * R + t] = r;
    rr[t] = r;
  }
  __syncthreads();
  if (t < L) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < R; ++i) s += blk[i * L + t] * rr[i];
    P[(size_t)c * L + t] = s;
  }
  const double s1 = wave_sum(r * r), s2 = wave_sum(bb * bb);
  if (t == 0) pa1[c] = s1, pa2[c] = s2;
}
template <int R, int L>
__global__ void __launch_bounds__(64) lsqr_r0(R0Args a) {
  cell_r0<R, L>(a.A, a.colmap, a.q, a.x0, a.b, a.u, a.P, a.pa1, a.pa2, blockIdx.x);
}
template <int R, int L>
__global__ void __launch_bounds__(64) lsqr_r0_multi(Jobs<R0Args> J) {
  const int i = job_of(J, blockIdx.x);
  const R0Args &a = J.a[i];
  cell_r0<R, L>(a.A, a.colmap, a.q, a.x0, a.b, a.u, a.P, a.pa1, a.pa2, blockIdx.x - J.first[i]);
}
// the solves' starting states (zero but for the constants the recurrences
// start from) and dx = 0
__global__ void init_multi_kernel(Jobs<InitArgs> J) {
  const int i = job_of(J, blockIdx.x);
  const InitArgs &a = J.a[i];
  const int j = (blockIdx.x - J.first[i]) * blockDim.x + threadIdx.x;
  if (j < a.n) a.x[j] = 0.0;
  if (blockIdx.x == J.first[i] && threadIdx.x < 2) {
    LsqrState s;
    memset(&s, 0, sizeof(s));
    s.cs2 = -1.0;
    s.atol = a.atol;
    s.ctol = 1.0 / 1.0e8;
    s.itnlim = a.itnlim;
    a.st[threadIdx.x] = s;
  }
}
__global__ void finish_multi_kernel(Jobs<FinArgs> J) {
  const int i = job_of(J, blockIdx.x);
  const FinArgs &a = J.a[i];
  const int j = (blockIdx.x - J.first[i]) * blockDim.x + threadIdx.x;
  if (j < a.n) a.x[j] = a.q[j] * a.dx[j] + a.x0[j];
}

// ---------------------------------------------------------------------------
// iteration k, cell part.  Block 0 runs the scalar recurrences (stopping
// tests of iteration k-2, rotation of iteration k-1) and stores the state;
// blocks 1..G do u' = A v - alpha u and P = A^T u' for a contiguous range of
// about ncell/G cells (G sized to the blocks resident on the device at once;
// the ranges of one XCD adjacent, see xcd_logical), prefetching the
// next cell's block into registers while computing on the current one.  The
// cell blocks need only alpha_k = ||v_k||, which every block reduces from the
// unknown kernel's partials in the same order; when block 0 finds that the
// iteration stopped, the cell blocks' work of this launch is simply unused
// (the unknown kernel that follows sees the stop and leaves x unchanged).
template <typename T, int R, int L>
__device__ __forceinline__ void cell_A(const int ncell, const int cgroup, const T *__restrict__ A,
                                       const int *__restrict__ colmap, const double *__restrict__ vp,
                                       double *__restrict__ u, double *__restrict__ P, double *__restrict__ pa1,
                                       const double *__restrict__ pb1, const double *__restrict__ pb2, const int nb,
                                       const LsqrState *__restrict__ sin, LsqrState *__restrict__ sout,
                                       const int bid, const int nblk) {
  constexpr int S = block_len<T, R, L>();
  __shared__ __align__(16) T blk[S];
  __shared__ double vs[64], uu[64], sh[1];
  __shared__ __align__(16) float vsf[64], uuf[64];
  constexpr bool f32 = sizeof(T) <= 4;  // FP32 / FP16 blocks: FP32 products, two accumulators
  const int t = threadIdx.x;
  // the iteration index travels with the scalars, so a launch takes no
  // per-iteration argument (a captured chunk replays unchanged)
  const int k = sin->kseq + 1;
  if (bid == 0) {
    const double nvs = reduce_partials<64>(pb1, nb, sh);
    const double dks = k >= 3 ? reduce_partials<64>(pb2, nb, sh) : 0.0;
    if (t == 0) {
      LsqrState o;
      scalar_update(k, *sin, nvs, dks, o);
      o.kseq = k;
      *sout = o;
    }
    return;
  }
  if (sin->done) return;
  // this block's cells: a contiguous range of whole cell groups (consecutive
  // ranges on one XCD); one partial sum of ||u'||^2 per group, so the sums
  // do not depend on the grid size
  const int G = nblk - 1, lb = xcd_logical(bid, 1, G + 1);
  const int ngroup = (ncell + cgroup - 1) / cgroup;
  const int g0 = (int)((long long)lb * ngroup / G), g1 = (int)((long long)(lb + 1) * ngroup / G);
  const int cend = min(ncell, g1 * cgroup);
  int c = g0 * cgroup;
  if (c >= cend) return;
  // column maps run one cell ahead of the v gathers that need them, so the
  // gather of cell c+1 is issued before cell c is computed
  const int j = t < L ? colmap[(size_t)c * L + t] : -1;
  int jn = (c + 1 < cend && t < L) ? colmap[(size_t)(c + 1) * L + t] : -1;
  double uo = t < R ? u[(size_t)c * R + t] : 0.0;
  RegBlock<T, S> rb;
  rb.load(A + (size_t)c * S);
  double vj = j >= 0 ? vp[j] : 0.0;
  const double beta = sin->beta, alpha_prev = sin->alpha;
  const double nvs = reduce_partials<64>(pb1, nb, sh);
  // alpha_k as the recurrences define it (kept from k-1 when beta_k = 0)
  const double alpha = (k == 1 || beta > 0.0) ? sqrt(nvs) : alpha_prev;
  double acc = 0.0;
  for (;;) {
    rb.store(blk);
    {
      const double vn = t < L ? (alpha > 0.0 ? vj / alpha : vj) : 0.0;
      if constexpr (f32)
        vsf[t] = (float)vn;
      else
        vs[t] = vn;
    }
    __syncthreads();
    const int cn = c + 1;
    double vjn = 0.0, uon = 0.0;
    int jnn = -1;
    if (cn < cend) {
      vjn = jn >= 0 ? vp[jn] : 0.0;
      uon = t < R ? u[(size_t)cn * R + t] : 0.0;
      rb.load(A + (size_t)cn * S);
      if (cn + 1 < cend && t < L) jnn = colmap[(size_t)(cn + 1) * L + t];
    }
    double un = 0.0;
    {
      double s = 0.0;
      const int row = t < R ? t : R - 1;
      if constexpr (sizeof(T) == 2) {
        // row of LP halves (zero padded), four per 8-byte read; v padded with zeros
        constexpr int LP = row_len<T, L>();
        const uint2 *rp = reinterpret_cast<const uint2 *>(blk + row * LP);
        float f0 = 0.f, f1 = 0.f;
#pragma unroll
        for (int k = 0; k < LP / 4; ++k) {
          const uint2 cw = rp[k];
          const float4 va = *reinterpret_cast<const float4 *>(vsf + 4 * k);
          const _Float16 *h = reinterpret_cast<const _Float16 *>(&cw);
          f0 += (float)h[0] * va.x;
          f1 += (float)h[1] * va.y;
          f0 += (float)h[2] * va.z;
          f1 += (float)h[3] * va.w;
        }
        s = (double)f0 + (double)f1;
      } else if constexpr (f32) {
        float f0 = 0.f, f1 = 0.f;
#pragma unroll
        for (int l = 0; l + 1 < L; l += 2) {
          f0 += (float)blk[row * L + l] * vsf[l];
          f1 += (float)blk[row * L + l + 1] * vsf[l + 1];
        }
        if (L & 1) f0 += (float)blk[row * L + L - 1] * vsf[L - 1];
        s = (double)f0 + (double)f1;
      } else {
#pragma unroll
        for (int l = 0; l < L; ++l) s += (double)blk[row * L + l] * vs[l];
      }
      const double uold = beta > 0.0 ? uo / beta : uo;
      if (t < R) {
        un = -alpha * uold + s;
        u[(size_t)c * R + t] = un;
      }
    }
    if constexpr (f32)
      uuf[t] = (float)un;
    else
      uu[t] = un;
    __syncthreads();
    if constexpr (sizeof(T) == 2) {
      // lane t: columns 2p, 2p+1 (p = t mod 32) over one half of the rows
      // (t / 32), four-byte reads; the halves are added across the wave
      constexpr int LP = row_len<T, L>(), NPAIR = (L + 1) / 2, RH = (R + 1) / 2;
      const int p = t & 31, hf = t >> 5, i0 = hf ? RH : 0;
      const int pc = p < NPAIR ? p : 0;
      float s0 = 0.f, s1 = 0.f;
#pragma unroll
      for (int k = 0; k < RH; ++k) {
        const int i = i0 + k;
        if (i < R) {
          const unsigned int pw = *reinterpret_cast<const unsigned int *>(blk + i * LP + 2 * pc);
          const _Float16 *h = reinterpret_cast<const _Float16 *>(&pw);
          const float ui = uuf[i];
          s0 += (float)h[0] * ui;
          s1 += (float)h[1] * ui;
        }
      }
      s0 += __shfl_xor(s0, 32, 64);
      s1 += __shfl_xor(s1, 32, 64);
      if (hf == 0 && p < NPAIR) {
        float *Pf = reinterpret_cast<float *>(P);
        Pf[(size_t)c * L + 2 * p] = s0;
        if (2 * p + 1 < L) Pf[(size_t)c * L + 2 * p + 1] = s1;
      }
    } else {
      double s = 0.0;
      const int col = t < L ? t : L - 1;
      if constexpr (f32) {
        float f0 = 0.f, f1 = 0.f;
#pragma unroll
        for (int i = 0; i + 1 < R; i += 2) {
          f0 += (float)blk[i * L + col] * uuf[i];
          f1 += (float)blk[(i + 1) * L + col] * uuf[i + 1];
        }
        if (R & 1) f0 += (float)blk[(R - 1) * L + col] * uuf[R - 1];
        s = (double)f0 + (double)f1;
      } else {
#pragma unroll
        for (int i = 0; i < R; ++i) s += (double)blk[i * L + col] * uu[i];
      }
      if (t < L) P[(size_t)c * L + t] = s;
    }
    acc += un * un;
    if (cn % cgroup == 0 || cn >= cend) {  // the group ends with cell c
      const double s1 = wave_sum(acc);
      if (t == 0) pa1[c / cgroup] = s1;
      acc = 0.0;
    }
    if (cn >= cend) break;
    vj = vjn, jn = jnn, uo = uon, c = cn;
    __syncthreads();  // blk, vs and uu are rewritten next
  }
}

#define TT_CELL_ARGS(T, a)                                                                                       \
  (a).ncell, (a).cgroup, static_cast<const T *>((a).A), (a).colmap, (a).vp, (a).u, (a).P, (a).pa1, (a).pb1, \
      (a).pb2, (a).nb, (a).sin, (a).sout
template <typename T, int R, int L>
__global__ void __launch_bounds__(64) lsqr_A(CellArgs a) {
  cell_A<T, R, L>(TT_CELL_ARGS(T, a), blockIdx.x, gridDim.x);
}
template <typename T, int R, int L>
__global__ void __launch_bounds__(64) lsqr_A_multi(Jobs<CellArgs> J) {
  const int i = job_of(J, blockIdx.x);
  cell_A<T, R, L>(TT_CELL_ARGS(T, J.a[i]), blockIdx.x - J.first[i], J.first[i + 1] - J.first[i]);
}

// orders a wavefront's LDS stores before its own later LDS loads by other
// lanes (a wavefront's LDS operations complete in order; this keeps the
// compiler from moving them across)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One cell of the reduced-precision cell kernel, its block and v already in
// this wavefront's LDS: u' = A v - alpha u (stored, returned) and A^T u'
// (stored to P).
// v entry of column l, held by lane l, as a wavefront-uniform value
__device__ __forceinline__ float lane_float(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// On 16-bit blocks the v entries come from the lanes' own vf (lane l: column
// l) through the scalar unit rather than LDS broadcasts; on FP32 blocks from
// vsf in LDS.
template <typename T, int R, int L>
__device__ __forceinline__ double aw_compute(int t, int c, const T *blk, const float *vsf, float vf, float *uuf,
                                             double uo, double alpha, double beta, double *__restrict__ u,
                                             double *__restrict__ P) {
  double un = 0.0;
  {
    double s = 0.0;
    const int row = t < R ? t : R - 1;
    if constexpr (sizeof(T) == 2) {
      constexpr int LP = row_len<T, L>();
      const uint2 *rp = reinterpret_cast<const uint2 *>(blk + row * LP);
      float f0 = 0.f, f1 = 0.f;
#pragma unroll
      for (int q = 0; q < LP / 4; ++q) {
        const uint2 cw = rp[q];
        const _Float16 *h = reinterpret_cast<const _Float16 *>(&cw);
        f0 += (float)h[0] * lane_float(vf, 4 * q);
        f1 += (float)h[1] * lane_float(vf, 4 * q + 1);
        f0 += (float)h[2] * lane_float(vf, 4 * q + 2);
        f1 += (float)h[3] * lane_float(vf, 4 * q + 3);
      }
      s = (double)f0 + (double)f1;
    } else {
      float f0 = 0.f, f1 = 0.f;
#pragma unroll
      for (int l = 0; l + 1 < L; l += 2) {
        f0 += (float)blk[row * L + l] * vsf[l];
        f1 += (float)blk[row * L + l + 1] * vsf[l + 1];
      }
      if (L & 1) f0 += (float)blk[row * L + L - 1] * vsf[L - 1];
      s = (double)f0 + (double)f1;
    }
    const double uold = beta > 0.0 ? uo / beta : uo;
    if (t < R) {
      un = -alpha * uold + s;
      u[(size_t)c * R + t] = un;
    }
  }
  uuf[t] = (float)un;
  wave_sync();
  if constexpr (sizeof(T) == 2) {
    constexpr int LP = row_len<T, L>(), NPAIR = (L + 1) / 2, RH = (R + 1) / 2;
    const int p = t & 31, hf = t >> 5, i0 = hf ? RH : 0;
    const int pc = p < NPAIR ? p : 0;
    float s0 = 0.f, s1 = 0.f;
#pragma unroll
    for (int q = 0; q < RH; ++q) {
      const int i = i0 + q;
      if (i < R) {
        const unsigned int pw = *reinterpret_cast<const unsigned int *>(blk + i * LP + 2 * pc);
        const _Float16 *h = reinterpret_cast<const _Float16 *>(&pw);
        const float ui = uuf[i];
        s0 += (float)h[0] * ui;
        s1 += (float)h[1] * ui;
      }
    }
    s0 += __shfl_xor(s0, 32, 64);
    s1 += __shfl_xor(s1, 32, 64);
    if (hf == 0 && p < NPAIR) {
      float *Pf = reinterpret_cast<float *>(P);
      Pf[(size_t)c * L + 2 * p] = s0;
      if (2 * p + 1 < L) Pf[(size_t)c * L + 2 * p + 1] = s1;
    }
  } else {
    const int col = t < L ? t : L - 1;
    float f0 = 0.f, f1 = 0.f;
#pragma unroll
    for (int i = 0; i + 1 < R; i += 2) {
      f0 += (float)blk[i * L + col] * uuf[i];
      f1 += (float)blk[(i + 1) * L + col] * uuf[i + 1];
    }
    if (R & 1) f0 += (float)blk[(R - 1) * L + col] * uuf[R - 1];
    if (t < L) P[(size_t)c * L + t] = (double)f0 + (double)f1;
  }
  return un;
}

// The same cell work on reduced-precision blocks with NW wavefronts per
// workgroup.  A workgroup takes a contiguous range of whole groups of cgroup
// cells (a multiple of NW); wavefront w takes cells cb + w, cb + w + NW, ...
// of each group, prefetching its next cell as above.  A group's ||u'||^2
// partial is the wavefronts' sums added in wavefront order, so it depends on
// neither the grid nor the timing; groups are NW times larger than with one
// wavefront per workgroup, so the unknown kernel reduces NW times fewer
// partials, and alpha is reduced once per workgroup.  LDS staging is per
// wavefront and synchronised within it; the workgroup barrier sits at group
// ends, which every wavefront reaches.
template <typename T, int R, int L, int NW>
__device__ __forceinline__ void cell_Aw(const int ncell, const int cgroup, const T *__restrict__ A,
                                        const int *__restrict__ colmap, const double *__restrict__ vp,
                                        double *__restrict__ u, double *__restrict__ P, double *__restrict__ pa1,
                                        const double *__restrict__ pb1, const double *__restrict__ pb2, const int nb,
                                        const LsqrState *__restrict__ sin, LsqrState *__restrict__ sout,
                                        const int bid, const int nblk) {
  static_assert(sizeof(T) <= 4, "multi-wavefront cell kernel: FP32 or FP16 blocks");
  constexpr int S = block_len<T, R, L>();
  __shared__ __align__(16) T blk_w[NW][S];
  __shared__ __align__(16) float vsf_w[NW][64], uuf_w[NW][64];
  __shared__ double sh[NW], gp[NW];
  // the wavefront index, known uniform (so cell indices and their branches
  // stay scalar)
  const int t = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  T *blk = blk_w[w];
  float *vsf = vsf_w[w], *uuf = uuf_w[w];
  const int k = sin->kseq + 1;
  if (bid == 0) {
    const double nvs = reduce_partials<64 * NW>(pb1, nb, sh);
    const double dks = k >= 3 ? reduce_partials<64 * NW>(pb2, nb, sh) : 0.0;
    if (threadIdx.x == 0) {
      LsqrState o;
      scalar_update(k, *sin, nvs, dks, o);
      o.kseq = k;
      *sout = o;
    }
    return;
  }
  if (sin->done) return;
  const int G = nblk - 1, lb = xcd_logical(bid, 1, G + 1);
  const int ngroup = (ncell + cgroup - 1) / cgroup;
  const int g0 = (int)((long long)lb * ngroup / G), g1 = (int)((long long)(lb + 1) * ngroup / G);
  if (g0 >= g1) return;
  auto gend = [&](int g) { return min(ncell, (g + 1) * cgroup); };
  // this wavefront's cell after c, or -1 (only the last group can be short)
  auto next_cell = [&](int c) -> int {
    const int g = c / cgroup;
    if (c + NW < gend(g)) return c + NW;
    if (g + 1 >= g1) return -1;
    const int cn = (g + 1) * cgroup + w;
    return cn < gend(g + 1) ? cn : -1;
  };
  int c = g0 * cgroup + w;
  if (c >= gend(g0)) c = -1;
  int cn = c >= 0 ? next_cell(c) : -1;
  // the wavefronts' ||u'||^2 sums of group g, added in wavefront order
  auto group_partial = [&](int g, double acc) {
    const double ws = wave_sum(acc);
    if (t == 0) gp[w] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
      double sg = gp[0];
#pragma unroll
      for (int i = 1; i < NW; ++i) sg += gp[i];
      pa1[g] = sg;
    }
    __syncthreads();
  };
  {
    int jn = -1;
    double uo = 0.0, vj = 0.0;
    RegBlock<T, S> rb;
    if (c >= 0) {
      const int j = t < L ? colmap[(size_t)c * L + t] : -1;
      jn = (cn >= 0 && t < L) ? colmap[(size_t)cn * L + t] : -1;
      uo = t < R ? u[(size_t)c * R + t] : 0.0;
      rb.load(A + (size_t)c * S);
      vj = j >= 0 ? vp[j] : 0.0;
    }
    const double beta = sin->beta, alpha_prev = sin->alpha;
    const double nvs = reduce_partials<64 * NW>(pb1, nb, sh);
    const double alpha = (k == 1 || beta > 0.0) ? sqrt(nvs) : alpha_prev;
    for (int g = g0; g < g1; ++g) {
      double acc = 0.0;
      while (c >= 0 && c < gend(g)) {
        rb.store(blk);
        const float vf = (float)(t < L ? (alpha > 0.0 ? vj / alpha : vj) : 0.0);
        if constexpr (sizeof(T) != 2) vsf[t] = vf;
        wave_sync();
        const int cnn = cn >= 0 ? next_cell(cn) : -1;
        double vjn = 0.0, uon = 0.0;
        int jnn = -1;
        if (cn >= 0) {
          vjn = jn >= 0 ? vp[jn] : 0.0;
          uon = t < R ? u[(size_t)cn * R + t] : 0.0;
          rb.load(A + (size_t)cn * S);
          if (cnn >= 0 && t < L) jnn = colmap[(size_t)cnn * L + t];
        }
        const double un = aw_compute<T, R, L>(t, c, blk, vsf, vf, uuf, uo, alpha, beta, u, P);
        acc += un * un;
        c = cn, cn = cnn, vj = vjn, jn = jnn, uo = uon;
        wave_sync();  // blk, vsf and uuf are rewritten next
      }
      group_partial(g, acc);
    }
  }
}

template <typename T, int R, int L, int NW>
__global__ void __launch_bounds__(64 * NW, 4) lsqr_Aw(CellArgs a) {
  cell_Aw<T, R, L, NW>(TT_CELL_ARGS(T, a), blockIdx.x, gridDim.x);
}
template <typename T, int R, int L, int NW>
__global__ void __launch_bounds__(64 * NW, 4) lsqr_Aw_multi(Jobs<CellArgs> J) {
  if (J.il > 0) {
    const int b = blockIdx.x, i = (b / kXcd) % J.il, l = kXcd * (b / (kXcd * J.il)) + b % kXcd;
    if (l >= J.gs) return;
    cell_Aw<T, R, L, NW>(TT_CELL_ARGS(T, J.a[i]), l, J.gs);
    return;
  }
  const int i = job_of(J, blockIdx.x);
  cell_Aw<T, R, L, NW>(TT_CELL_ARGS(T, J.a[i]), blockIdx.x - J.first[i], J.first[i + 1] - J.first[i]);
}

// byte-wise comparison of two device arrays of 32-bit words: *flag = 1 where any word differs
__global__ void differ_kernel(size_t nw, const uint32_t *__restrict__ x, const uint32_t *__restrict__ y,
                              int *__restrict__ flag) {
  int d = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += (size_t)gridDim.x * blockDim.x)
    d |= x[i] != y[i];
  if (d) *flag = 1;
}
__global__ void set_int_kernel(int *p, int v) { *p = v; }

// ---------------------------------------------------------------------------
// iteration k, unknown part (k = 0: initialisation after r0)
//
// The unknown kernel runs NB = 256, 512 or 1024 threads per workgroup (picked
// by block_tune) with the same bits for any NB: its partial sums are per fixed
// group of kNB = 256 unknowns, and the reductions of the cell kernel's
// partials are done by the first kNB threads in the order of a kNB-thread
// workgroup (the other threads add exact zeros).

// reduce_partials<kNB> computed in a workgroup of NB >= kNB threads
template <int NB>
__device__ __forceinline__ double reduce_partials_nb(const double *__restrict__ p, int n, double *sh) {
  double acc = 0.0;
  if ((int)threadIdx.x < kNB)
    for (int base = 0; base < n; base += 16 * kNB) {
      double v[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int i = base + q * kNB + (int)threadIdx.x;
        v[q] = i < n ? p[i] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 16; ++q) acc += v[q];
    }
  return block_sum<NB>(acc, sh);
}
// sum over each group of kNB consecutive threads (wave sums added in wave
// order, as block_sum<kNB>); every thread gets its group's sum
template <int NB>
__device__ __forceinline__ double group_sum_nb(double v, double *sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  const int g0 = (threadIdx.x / kNB) * (kNB / 64);
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < kNB / 64; ++i) r += sh[g0 + i];
  __syncthreads();
  return r;
}

template <int W, typename PT, int NB>
__device__ __forceinline__ void unk_B(const int n, const int *__restrict__ gtab, const PT *__restrict__ P,
                                      double *__restrict__ vp, double *__restrict__ w, double *__restrict__ x,
                                      const double *__restrict__ pa1, const double *__restrict__ pa2, const int na,
                                      double *__restrict__ pb1, double *__restrict__ pb2, LsqrState *st,
                                      const int bid, const int nblk) {
  static_assert(NB % kNB == 0, "whole groups of kNB unknowns per workgroup");
  __shared__ double sh[NB / 64];
  const int jb = xcd_logical(bid, 0, nblk) * NB;  // first unknown of the workgroup
  const int j = jb + threadIdx.x;
  // the index table and the unknown's own vectors do not depend on the
  // state: they are requested together with it
  // (threads past the last unknown repeat its loads and store nothing, and
  // padding entries of the table gather entry 0 and discard it: no load sits
  // behind a branch, so each wait counts exactly the loads it needs)
  const int jj = j < n ? j : n - 1;
  int idx[W];
  {
    const int4 *g = reinterpret_cast<const int4 *>(gtab + (size_t)jj * W);
#pragma unroll
    for (int q = 0; q < W / 4; ++q) {
      int4 v = g[q];
      idx[4 * q] = v.x, idx[4 * q + 1] = v.y, idx[4 * q + 2] = v.z, idx[4 * q + 3] = v.w;
    }
  }
  const double vold = vp[jj], wo = w[jj], xo = x[jj];
  const int k = st->kseq;  // written by this iteration's cell kernel (0: the initial copy)
  if (k >= 1 && st->done) return;
  // the column's gathers are issued before the partial-sum reduction, so
  // their latency overlaps it
  double pv[W];
#pragma unroll
  for (int e = 0; e < W; ++e) {
    const PT pe = P[idx[e] >= 0 ? idx[e] : 0];
    pv[e] = idx[e] >= 0 ? (double)pe : 0.0;
  }
  const double beta = sqrt(reduce_partials_nb<NB>(pa1, na, sh));  // beta_{k+1}
  const double alpha = st->alpha, t1 = st->t1, t2 = st->t2, t3 = st->t3;
  double nv = 0.0, dk = 0.0;
  if (j < n) {
    double vk = 0.0;
    if (k >= 1) {
      vk = alpha > 0.0 ? vold / alpha : vold;
      if (k == 1) {
        w[j] = vk;
        x[j] = 0.0;
      } else {
        // x/w update of iteration k-1 (rotation computed by this iteration's cell kernel)
        x[j] = t1 * wo + xo;
        w[j] = t2 * wo + vk;
        const double tt = t3 * wo;
        dk = tt * tt;
      }
    }
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < W; ++e)
      if (idx[e] >= 0) s += pv[e];
    double nvp;
    if (beta > 0.0)
      nvp = s / beta - beta * vk;
    else
      nvp = k == 0 ? 0.0 : vold;  // the reference keeps v and alpha when beta = 0
    vp[j] = nvp;
    nv = nvp * nvp;
  }
  const double s1 = group_sum_nb<NB>(nv, sh), s2 = group_sum_nb<NB>(dk, sh);
  const int grp = (jb + (int)threadIdx.x) / kNB, ngrp = (n + kNB - 1) / kNB;
  if (threadIdx.x % kNB == 0 && grp < ngrp) pb1[grp] = s1, pb2[grp] = s2;
  if (bid == 0) {
    double nbv = 0.0;
    if (k == 0) nbv = sqrt(reduce_partials_nb<NB>(pa2, na, sh));
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

#define TT_UNK_ARGS(PT, a) \
  (a).n, (a).gtab, static_cast<const PT *>((a).P), (a).vp, (a).w, (a).x, (a).pa1, (a).pa2, (a).na, (a).pb1, (a).pb2, (a).st
template <int W, typename PT, int NB>
__global__ void __launch_bounds__(NB) lsqr_B(UnkArgs a) {
  unk_B<W, PT, NB>(TT_UNK_ARGS(PT, a), blockIdx.x, gridDim.x);
}
template <int W, typename PT, int NB>
__global__ void __launch_bounds__(NB) lsqr_B_multi(Jobs<UnkArgs> J) {
  const int i = job_of(J, blockIdx.x);
  unk_B<W, PT, NB>(TT_UNK_ARGS(PT, J.a[i]), blockIdx.x - J.first[i], J.first[i + 1] - J.first[i]);
}

__global__ void finish_kernel(int n, const double *__restrict__ q, const double *__restrict__ dx,
                              const double *__restrict__ x0, double *__restrict__ x) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  x[j] = q[j] * dx[j] + x0[j];
}

// launch helpers over the supported block shapes and gather widths
R0Args r0_args(BlockSys &s, LsqrWork &wk, const double *b, const double *x) {
  R0Args a;
  a.A = s.A.p, a.q = s.q.p, a.x0 = x, a.b = b, a.colmap = s.colmap.p;
  a.u = wk.u.p, a.P = wk.P.p, a.pa1 = wk.pa1.p, a.pa2 = wk.pa2.p;
  a.ncell = s.ncell;
  return a;
}
template <int R, int L>
void launch_r0(BlockSys &s, LsqrWork &wk, const double *b, const double *x, hipStream_t st) {
  hipLaunchKernelGGL((lsqr_r0<R, L>), dim3(s.ncell), dim3(64), 0, st, r0_args(s, wk, b, x));
}
// the cell kernel of a block type: one wavefront per workgroup on FP64
// blocks (their LDS staging limits residency), kNW on FP32/FP16 blocks
template <typename T, int R, int L>
const void *cell_kernel() {
  if constexpr (sizeof(T) == 8)
    return reinterpret_cast<const void *>(lsqr_A<T, R, L>);
  else
    return reinterpret_cast<const void *>(lsqr_Aw<T, R, L, kNW>);
}
template <typename T>
constexpr int cell_threads() {
  return sizeof(T) == 8 ? 64 : 64 * kNW;
}
// resident workgroups of a cell-kernel instantiation on the whole device,
// computed once per device and published under a lock (serving threads build
// cases concurrently)
template <typename T, int R, int L>
int persistent_grid(int ngroup) {
  static std::mutex mu;
  static std::map<int, int> resident;  // device -> workgroups resident on the whole device
  int dev = 0;
  TT_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto it = resident.find(dev);
  if (it == resident.end()) {
    int cus = 0, per_cu = 0, lds_per_cu = 0;
    TT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    TT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, cell_kernel<T, R, L>(), cell_threads<T>(), 0));
    hipFuncAttributes fa;
    TT_HIP(hipFuncGetAttributes(&fa, cell_kernel<T, R, L>()));
    TT_HIP(hipDeviceGetAttribute(&lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    // static LDS rounded to the 512-byte allocation granule
    const size_t lds = (fa.sharedSizeBytes + 511) / 512 * 512;
    if (lds > 0 && lds_per_cu > 0) per_cu = std::min<int>(per_cu, (int)(lds_per_cu / lds));
    if (per_cu < 1) per_cu = 1;
    it = resident.emplace(dev, per_cu * cus).first;
  }
  return std::min(ngroup, it->second);
}
// the cell kernel's arguments and grid (workgroups, the scalar one included)
// for iteration k of system s
template <typename T>
CellArgs cell_args(BlockSys &s, LsqrWork &wk, int k, const T *A, int &grid) {
  constexpr int NW = sizeof(T) == 8 ? 1 : kNW;
  const int cg = cell_group(s.ncell, NW);
  const int ngroup = ceil_div(s.ncell, cg);
  const int tuned = s.grid_a[sizeof(T) == 8 ? 0 : sizeof(T) == 4 ? 1 : 2];
  int g = 0;
  if (s.R == 51)
    g = std::min(ngroup, tuned > 0 ? tuned : persistent_grid<T, 51, 51>(ngroup));
  else
    g = std::min(ngroup, tuned > 0 ? tuned : persistent_grid<T, 27, 27>(ngroup));
  wk.na = ngroup;
  grid = g + 1;
  CellArgs a;
  a.ncell = s.ncell, a.cgroup = cg, a.nb = (int)ceil_div(s.n, kNB);
  a.A = A, a.colmap = s.colmap.p;
  a.vp = wk.v.p, a.pb1 = wk.pb1.p, a.pb2 = wk.pb2.p;
  a.u = wk.u.p, a.P = wk.P.p, a.pa1 = wk.pa1.p;
  a.sin = wk.state.p + ((k - 1) & 1);
  a.sout = wk.state.p + (k & 1);
  return a;
}
// e0/e1 non-null: the dispatch itself records its start and end (sampled timing)
template <typename T, int R, int L>
void launch_A(BlockSys &s, LsqrWork &wk, int k, const T *A, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
  constexpr int NW = sizeof(T) == 8 ? 1 : kNW;
  int grid = 0;
  const CellArgs a = cell_args<T>(s, wk, k, A, grid);
  if constexpr (NW == 1) {
    if (e0)
      hipExtLaunchKernelGGL((lsqr_A<T, R, L>), dim3(grid), dim3(64), 0, st, e0, e1, 0, a);
    else
      hipLaunchKernelGGL((lsqr_A<T, R, L>), dim3(grid), dim3(64), 0, st, a);
  } else {
    if (e0)
      hipExtLaunchKernelGGL((lsqr_Aw<T, R, L, NW>), dim3(grid), dim3(64 * NW), 0, st, e0, e1, 0, a);
    else
      hipLaunchKernelGGL((lsqr_Aw<T, R, L, NW>), dim3(grid), dim3(64 * NW), 0, st, a);
  }
}
// the unknown kernel's arguments for iteration k (k = 0: the initial
// residual's FP64 partials; later the cell kernel's partials in its block
// precision's type)
UnkArgs unk_args(BlockSys &s, LsqrWork &wk, int k) {
  UnkArgs a;
  a.n = s.n, a.na = k == 0 ? s.ncell : wk.na;
  a.gtab = s.gtab.p, a.P = wk.P.p;
  a.vp = wk.v.p, a.w = wk.w.p, a.x = wk.x.p;
  a.pa1 = wk.pa1.p, a.pa2 = wk.pa2.p;
  a.pb1 = wk.pb1.p, a.pb2 = wk.pb2.p;
  a.st = wk.state.p + (k & 1);
  return a;
}
template <int W>
void launch_B(BlockSys &s, LsqrWork &wk, int k, int prec, hipStream_t st) {
  const int nbt = s.nb_b[prec] > 0 ? s.nb_b[prec] : kNB;  // threads per workgroup (tuned)
  const bool half = k >= 1 && prec == kBlocks16;
  const UnkArgs a = unk_args(s, wk, k);
  auto go = [&](auto nbc) {
    constexpr int NB = decltype(nbc)::value;
    if (half)
      hipLaunchKernelGGL((lsqr_B<W, float, NB>), dim3(ceil_div(s.n, NB)), dim3(NB), 0, st, a);
    else
      hipLaunchKernelGGL((lsqr_B<W, double, NB>), dim3(ceil_div(s.n, NB)), dim3(NB), 0, st, a);
  };
  if (nbt == 1024)
    go(std::integral_constant<int, 1024>());
  else if (nbt == 512)
    go(std::integral_constant<int, 512>());
  else
    go(std::integral_constant<int, 256>());
}

void r0(BlockSys &s, LsqrWork &wk, const double *b, const double *x, hipStream_t st) {
  if (s.R == 51)
    launch_r0<51, 51>(s, wk, b, x, st);
  else
    launch_r0<27, 27>(s, wk, b, x, st);
}
void stepA(BlockSys &s, LsqrWork &wk, int k, int prec, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
  if (s.R == 51) {
    if (prec == kBlocks16)
      launch_A<_Float16, 51, 51>(s, wk, k, (const _Float16 *)s.A16.p, st, e0, e1);
    else if (prec == kBlocks32)
      launch_A<float, 51, 51>(s, wk, k, (const float *)s.A32.p, st, e0, e1);
    else
      launch_A<double, 51, 51>(s, wk, k, (const double *)s.A.p, st, e0, e1);
  } else {
    if (prec == kBlocks16)
      launch_A<_Float16, 27, 27>(s, wk, k, (const _Float16 *)s.A16.p, st, e0, e1);
    else if (prec == kBlocks32)
      launch_A<float, 27, 27>(s, wk, k, (const float *)s.A32.p, st, e0, e1);
    else
      launch_A<double, 27, 27>(s, wk, k, (const double *)s.A.p, st, e0, e1);
  }
}
void stepB(BlockSys &s, LsqrWork &wk, int k, int prec, hipStream_t st) {
  switch (s.gw) {
    case 4: launch_B<4>(s, wk, k, prec, st); break;
    case 8: launch_B<8>(s, wk, k, prec, st); break;
    case 12: launch_B<12>(s, wk, k, prec, st); break;
    default: launch_B<16>(s, wk, k, prec, st); break;
  }
}

std::atomic<int> g_sharing{1};
std::atomic<long long> g_shared_launches{0}, g_shared_systems{0};

// ---- several systems per launch (block_lsqr_multi): the same kernel bodies
// on workgroup ranges, one range per system
template <typename T, int R, int L>
void launch_A_multi(const std::vector<LsqrJob *> &act, int k, int prec, hipStream_t st) {
  constexpr int NW = sizeof(T) == 8 ? 1 : kNW;
  Jobs<CellArgs> J;
  J.n = (int)act.size();
  J.first[0] = 0;
  for (int i = 0; i < J.n; ++i) {
    BlockSys &s = *act[i]->s;
    const BlockSys &sa = act[i]->share_sys ? *act[i]->share_sys : s;  // equal blocks: read one copy
    const T *A = prec == kBlocks16   ? (const T *)sa.A16.p
                 : prec == kBlocks32 ? (const T *)sa.A32.p
                                     : (const T *)s.A.p;
    int grid = 0;
    J.a[i] = cell_args<T>(s, *act[i]->wk, k, A, grid);
    J.first[i + 1] = J.first[i] + grid;
  }
  if constexpr (NW == 1) {
    hipLaunchKernelGGL((lsqr_A_multi<T, R, L>), dim3(J.first[J.n]), dim3(64), 0, st, J);
  } else {
    // systems reading the same blocks: one launch per such class with their
    // workgroups interleaved at the class's first system's grid, so that each
    // block is fetched once into an XCD's L2 for all of them; the other
    // systems in one launch of workgroup ranges
    std::vector<std::vector<int>> cls;
    for (int i = 0; i < J.n; ++i) {
      size_t c = 0;
      while (c < cls.size() && J.a[cls[c][0]].A != J.a[i].A) ++c;
      if (c == cls.size()) cls.emplace_back();
      cls[c].push_back(i);
    }
    Jobs<CellArgs> rest;
    rest.n = 0;
    rest.first[0] = 0;
    for (const auto &c : cls) {
      if (c.size() == 1) {
        const int i = c[0];
        rest.a[rest.n] = J.a[i];
        rest.first[rest.n + 1] = rest.first[rest.n] + (J.first[i + 1] - J.first[i]);
        ++rest.n;
        continue;
      }
      Jobs<CellArgs> I;
      I.n = (int)c.size();
      for (int q = 0; q < I.n; ++q) I.a[q] = J.a[c[q]];
      I.il = I.n;
      I.gs = J.first[c[0] + 1] - J.first[c[0]];  // results do not depend on the grid
      const int gp = (I.gs + kXcd - 1) / kXcd * kXcd;
      hipLaunchKernelGGL((lsqr_Aw_multi<T, R, L, NW>), dim3(gp * I.n), dim3(64 * NW), 0, st, I);
      g_shared_launches += 1;
      g_shared_systems += I.n;
    }
    if (rest.n > 0)
      hipLaunchKernelGGL((lsqr_Aw_multi<T, R, L, NW>), dim3(rest.first[rest.n]), dim3(64 * NW), 0, st, rest);
  }
}
template <int R, int L>
void stepA_multi_rl(const std::vector<LsqrJob *> &act, int k, int prec, hipStream_t st) {
  if (prec == kBlocks16)
    launch_A_multi<_Float16, R, L>(act, k, prec, st);
  else if (prec == kBlocks32)
    launch_A_multi<float, R, L>(act, k, prec, st);
  else
    launch_A_multi<double, R, L>(act, k, prec, st);
}
void stepA_multi(const std::vector<LsqrJob *> &act, int k, int prec, hipStream_t st) {
  if (act[0]->s->R == 51)
    stepA_multi_rl<51, 51>(act, k, prec, st);
  else
    stepA_multi_rl<27, 27>(act, k, prec, st);
}
// the unknown kernels of systems with gather width W, kNB threads per
// workgroup (the results do not depend on the width of a workgroup)
template <int W>
void launch_B_multi(const std::vector<LsqrJob *> &act, int k, int prec, hipStream_t st) {
  Jobs<UnkArgs> J;
  J.n = 0;
  J.first[0] = 0;
  for (LsqrJob *j : act) {
    if (j->s->gw != W) continue;
    J.a[J.n] = unk_args(*j->s, *j->wk, k);
    J.first[J.n + 1] = J.first[J.n] + (int)ceil_div(j->s->n, kNB);
    ++J.n;
  }
  if (J.n == 0) return;
  if (k >= 1 && prec == kBlocks16)
    hipLaunchKernelGGL((lsqr_B_multi<W, float, kNB>), dim3(J.first[J.n]), dim3(kNB), 0, st, J);
  else
    hipLaunchKernelGGL((lsqr_B_multi<W, double, kNB>), dim3(J.first[J.n]), dim3(kNB), 0, st, J);
}
void stepB_multi(const std::vector<LsqrJob *> &act, int k, int prec, hipStream_t st) {
  launch_B_multi<4>(act, k, prec, st);
  launch_B_multi<8>(act, k, prec, st);
  launch_B_multi<12>(act, k, prec, st);
  launch_B_multi<16>(act, k, prec, st);
}

// events of a sampled launch of iteration k (recorded by the dispatch); -1 if not sampled
long probe_begin(LsqrProbe *p, int k) {
  if (!p || p->stride <= 0) return -1;
  unsigned long long n = p->seen++;
  if (n % (unsigned long long)p->stride) return -1;
  if (2 * (p->used + 1) > p->pool.size()) {
    for (int i = 0; i < 64; ++i) {
      hipEvent_t e;
      TT_HIP(hipEventCreate(&e));
      p->pool.push_back(e);
    }
  }
  long idx = (long)p->used++;
  if (p->iter.size() < p->used) p->iter.resize(p->used);
  p->iter[idx] = k;
  return idx;
}
// after the solve: keeps the launches that did the cell work (iterations
// 1..nit+1; later launches of the batch return at once)
void probe_collect(LsqrProbe *p, int nit) {
  if (!p) return;
  for (size_t i = 0; i < p->used; ++i) {
    if (p->iter[i] > nit + 1) continue;
    float ms = 0.f;
    TT_HIP(hipEventElapsedTime(&ms, p->pool[2 * i], p->pool[2 * i + 1]));
    p->total_ms += ms;
    p->sampled++;
  }
  p->used = 0;
}

std::atomic<unsigned> g_work_gen{0};
std::atomic<unsigned long long> g_sys_uid{0};

// whether two systems' blocks of precision prec (FP32 or FP16), column maps
// and gather tables are equal bit for bit; compared on the device once per
// pair of systems
bool lo_blocks_equal(BlockSys &x, BlockSys &y, int prec, hipStream_t st) {
  if (x.uid == y.uid) return true;
  for (const auto &e : x.twins[prec])
    if (e.first == y.uid) return e.second;
  bool eq = x.ncell == y.ncell && x.R == y.R && x.L == y.L && x.n == y.n && x.gw == y.gw &&
            (prec == kBlocks16 ? x.has16 && y.has16 : x.has32 && y.has32);
  if (eq) {
    DBuf<int> flag;
    flag.alloc(1);
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, st, flag.p, 0);
    auto cmp = [&](const void *p, const void *q, size_t bytes) {
      const size_t nw = bytes / 4;
      const unsigned blocks = (unsigned)std::min<size_t>(2048, std::max<size_t>(1, ceil_div(nw, 256)));
      hipLaunchKernelGGL(differ_kernel, dim3(blocks), dim3(256), 0, st, nw, static_cast<const uint32_t *>(p),
                         static_cast<const uint32_t *>(q), flag.p);
    };
    if (prec == kBlocks16)
      cmp(x.A16.p, y.A16.p, (size_t)x.ncell * block_stride16(x.R, x.L) * sizeof(_Float16));
    else
      cmp(x.A32.p, y.A32.p, (size_t)x.ncell * x.stride * sizeof(float));
    cmp(x.colmap.p, y.colmap.p, (size_t)x.ncell * x.L * sizeof(int));
    cmp(x.gtab.p, y.gtab.p, (size_t)x.n * x.gw * sizeof(int));
    TT_LAUNCH_CHECK();
    int h = 1;
    TT_HIP(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    TT_HIP(hipStreamSynchronize(st));
    eq = h == 0;
  }
  x.twins[prec].push_back({y.uid, eq});
  y.twins[prec].push_back({x.uid, eq});
  return eq;
}

// records 2^(lvl+1) iterations starting at an odd index (the state copies
// alternate with the parity of the index, so such a chunk replays at any odd
// index); nothing runs during the capture
hipGraphExec_t capture_chunk(BlockSys &s, LsqrWork &wk, int prec, int iters, hipStream_t st) {
  hipGraph_t g = nullptr;
  TT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  for (int i = 1; i <= iters; ++i) {
    stepA(s, wk, i, prec, st, nullptr, nullptr);
    stepB(s, wk, i, prec, st);
  }
  const hipError_t le = hipGetLastError();
  const hipError_t ee = hipStreamEndCapture(st, &g);
  if (le != hipSuccess || ee != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    ::tt::hip_check(le != hipSuccess ? le : ee, "LSQR iteration capture", __FILE__, __LINE__);
  }
  hipGraphExec_t ex = nullptr;
  const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  ::tt::hip_check(ie, "LSQR iteration graph instantiation", __FILE__, __LINE__);
  return ex;
}

}  // namespace

LsqrGraphs::LsqrGraphs(LsqrGraphs &&o) noexcept : wk(o.wk), gen(o.gen), grid(o.grid), nb(o.nb) {
  for (int i = 0; i < kLevels; ++i) exec[i] = o.exec[i], o.exec[i] = nullptr;
  o.wk = nullptr;
}
LsqrGraphs &LsqrGraphs::operator=(LsqrGraphs &&o) noexcept {
  if (this != &o) {
    reset();
    wk = o.wk, gen = o.gen, grid = o.grid, nb = o.nb;
    for (int i = 0; i < kLevels; ++i) exec[i] = o.exec[i], o.exec[i] = nullptr;
    o.wk = nullptr;
  }
  return *this;
}
void LsqrGraphs::reset() {
  for (auto &e : exec)
    if (e) (void)hipGraphExecDestroy(e), e = nullptr;
  wk = nullptr, gen = 0, grid = 0, nb = 0;
}

// ---------------------------------------------------------------------------
void block_setup(BlockSys &s, int ncell, int R, int L, int n, const std::vector<int> &colmap, const double *Afull,
                 bool want32, hipStream_t st) {
  if (!((R == 51 && L == 51) || (R == 27 && L == 27)))
    throw std::runtime_error("block LSQR supports 51x51 and 27x27 cell blocks");
  s.ncell = ncell, s.R = R, s.L = L, s.n = n, s.stride = block_stride(R, L);
  s.uid = ++g_sys_uid;
  for (auto &t : s.twins) t.clear();
  // transpose gather lists in cell order
  std::vector<int> cnt(n + 1, 0), idx;
  for (size_t k = 0; k < colmap.size(); ++k)
    if (colmap[k] >= 0) cnt[colmap[k] + 1]++;
  int wmax = 0;
  for (int j = 0; j < n; ++j) {
    if (cnt[j + 1] == 0) throw std::runtime_error("an unknown appears in no equation");
    wmax = std::max(wmax, cnt[j + 1]);
  }
  if (wmax > 16) throw std::runtime_error("an unknown is shared by more than 16 cells");
  for (int j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
  idx.resize(cnt[n]);
  std::vector<int> pos(cnt.begin(), cnt.end() - 1);
  for (size_t k = 0; k < colmap.size(); ++k)
    if (colmap[k] >= 0) idx[pos[colmap[k]]++] = (int)k;
  s.gw = (wmax + 3) / 4 * 4;
  std::vector<int> g((size_t)n * s.gw, -1);
  for (int j = 0; j < n; ++j)
    for (int e = cnt[j]; e < cnt[j + 1]; ++e) g[(size_t)j * s.gw + (e - cnt[j])] = idx[e];
  s.colmap.upload(colmap, st);
  s.tptr.upload(cnt, st);
  s.tidx.upload(idx, st);
  s.gtab.upload(g, st);
  s.q.alloc(n);
  size_t total = (size_t)ncell * s.stride;
  s.A.alloc(total);
  s.has32 = want32;
  if (want32) s.A32.alloc(total);
  hipLaunchKernelGGL(colnorm_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, R, L, s.tptr.p, s.tidx.p, Afull,
                     s.q.p);
  TT_LAUNCH_CHECK();
  hipLaunchKernelGGL(scale_kernel, dim3(ceil_div(total, 256)), dim3(256), 0, st, total, R, L, s.stride, s.colmap.p,
                     s.q.p, Afull, s.A.p, want32 ? s.A32.p : nullptr);
  TT_LAUNCH_CHECK();
  TT_HIP(hipStreamSynchronize(st));  // host vectors above are released on return
}

void BlockSys::SolveHistory::add(int it) {
  nit[head] = it;
  head = (head + 1) % kLen;
  n = std::min(n + 1, kLen);
}

// The count of the previous solve, or of the one before it when that lag
// has predicted the recent solves better: under the Picard loop the counts
// of a system often alternate between two levels from step to step (the
// right-hand side's change alternates), which a lag of two follows exactly.
int BlockSys::SolveHistory::predict() const {
  if (n == 0) return -1;
  auto at = [&](int i) { return nit[(head - 1 - i + 2 * kLen) % kLen]; };  // i = 0: most recent
  if (n < 3) return at(0);
  const int e1 = std::abs(at(0) - at(1)) + (n >= 4 ? std::abs(at(1) - at(2)) : 0);
  const int e2 = std::abs(at(0) - at(2)) + (n >= 4 ? std::abs(at(1) - at(3)) : 0);
  return e2 < e1 ? at(1) : at(0);
}

void block_make16(BlockSys &s, hipStream_t st) {
  if (s.has16) return;
  const int S16 = block_stride16(s.R, s.L);
  const size_t total16 = (size_t)s.ncell * S16;
  s.A16.alloc(total16);
  hipLaunchKernelGGL(to16_kernel, dim3(ceil_div(total16, 256)), dim3(256), 0, st, total16, s.R, s.L, block_row16(s.L),
                     s.stride, S16, s.A.p, s.A16.p);
  TT_LAUNCH_CHECK();
  s.has16 = true;
}

void block_make32(BlockSys &s, hipStream_t st) {
  if (s.has32) return;
  size_t total = (size_t)s.ncell * s.stride;
  s.A32.alloc(total);
  hipLaunchKernelGGL(to32_kernel, dim3(ceil_div(total, 256)), dim3(256), 0, st, total, s.A.p, s.A32.p);
  TT_LAUNCH_CHECK();
  s.has32 = true;
}

LsqrProbe::~LsqrProbe() {
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

void LsqrWork::ensure(int m_, int n_, int ncell_, int L_) {
  bool moved = false;
  if (m_ > m) u.alloc(m_), m = m_, moved = true;
  if (n_ > n) v.alloc(n_), w.alloc(n_), x.alloc(n_), n = n_, moved = true;
  if ((size_t)ncell_ * L_ > P.n) P.alloc((size_t)ncell_ * L_), moved = true;
  if ((size_t)ncell_ > pa1.n) pa1.alloc(ncell_), pa2.alloc(ncell_), moved = true;
  const size_t nb = ceil_div(n_, kNB);
  if (nb > pb1.n) pb1.alloc(nb), pb2.alloc(nb), moved = true;
  ncell = std::max(ncell, ncell_), L = std::max(L, L_);
  if (!state.p) {
    state.alloc(2);
    TT_HIP(hipHostMalloc((void **)&host_state, sizeof(LsqrState), hipHostMallocDefault));
    moved = true;
  }
  // process-wide numbering: a later work area at the same address never
  // matches graphs captured with an earlier one
  if (moved) gen = ++g_work_gen;
}

LsqrWork::~LsqrWork() {
  if (host_state) (void)hipHostFree(host_state);
}

namespace {

__global__ void fill_kernel(int n, double v, double *__restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// cell-kernel grid chosen per (rows per cell, cells, block precision), kept
// for the process: systems of the same shape tune once
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, std::pair<int, int>> g_tuned;  // -> (cell grid, unknown-kernel threads)

}  // namespace

// Autotuning of the cell kernel's grid: a scratch LSQR on this system (b = 1,
// zero tolerance, its own work buffers) times a few iterations at each
// candidate grid -- the blocks resident on the device and its halvings, down
// to 1/8 -- twice each, and keeps the fastest.  The cell partial sums are per
// fixed group of cells, so the grid does not change any result.
void block_tune(BlockSys &s, int prec, hipStream_t st) {
  if (s.grid_a[prec] > 0) return;
  const int forced_nb = s.nb_b[prec];  // set by tt_case_set_unknown_threads: kept
  const auto key = std::make_tuple(s.R, s.ncell, prec);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
      s.grid_a[prec] = it->second.first;
      s.nb_b[prec] = forced_nb > 0 ? forced_nb : it->second.second;
      return;
    }
  }
  const int m = s.ncell * s.R, n = s.n, ngroup = ceil_div(s.ncell, cell_group(s.ncell, cell_waves(prec)));
  int resident = 0;
  if (s.R == 51)
    resident = prec == kBlocks16 ? persistent_grid<_Float16, 51, 51>(ngroup)
               : prec == kBlocks32 ? persistent_grid<float, 51, 51>(ngroup)
                                   : persistent_grid<double, 51, 51>(ngroup);
  else
    resident = prec == kBlocks16 ? persistent_grid<_Float16, 27, 27>(ngroup)
               : prec == kBlocks32 ? persistent_grid<float, 27, 27>(ngroup)
                                   : persistent_grid<double, 27, 27>(ngroup);
  std::vector<int> cand;
  for (int d = 1; d <= 8; d *= 2) {
    const int g = std::min(ngroup, std::max(1, resident / d));
    if (std::find(cand.begin(), cand.end(), g) == cand.end()) cand.push_back(g);
  }
  LsqrWork tw;
  tw.ensure(m, n, s.ncell, s.L);
  DBuf<double> b(m), x(n);
  hipLaunchKernelGGL(fill_kernel, dim3(ceil_div(m, 256)), dim3(256), 0, st, m, 1.0, b.p);
  hipLaunchKernelGGL(fill_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, 0.0, x.p);
  LsqrState init;
  std::memset(&init, 0, sizeof(init));
  init.cs2 = -1.0;
  init.ctol = 0.0;
  init.itnlim = 1 << 30;
  LsqrState init2[2] = {init, init};
  TT_HIP(hipMemcpyAsync(tw.state.p, init2, sizeof(init2), hipMemcpyHostToDevice, st));
  TT_HIP(hipMemsetAsync(tw.x.p, 0, sizeof(double) * n, st));
  r0(s, tw, b.p, x.p, st);
  stepB(s, tw, 0, prec, st);
  constexpr int kIts = 6, kReps = 2;
  std::vector<hipEvent_t> ev(2 * cand.size() * kReps);
  for (auto &e : ev) TT_HIP(hipEventCreate(&e));
  int k = 1;
  for (int rep = 0; rep < kReps; ++rep)
    for (size_t ci = 0; ci < cand.size(); ++ci) {
      s.grid_a[prec] = cand[ci];
      TT_HIP(hipEventRecord(ev[2 * (rep * cand.size() + ci)], st));
      for (int i = 0; i < kIts; ++i, ++k) {
        stepA(s, tw, k, prec, st, nullptr, nullptr);
        stepB(s, tw, k, prec, st);
      }
      TT_HIP(hipEventRecord(ev[2 * (rep * cand.size() + ci) + 1], st));
    }
  TT_LAUNCH_CHECK();
  TT_HIP(hipMemcpyAsync(tw.host_state, tw.state.p + ((k - 1) & 1), sizeof(LsqrState), hipMemcpyDeviceToHost, st));
  TT_HIP(hipStreamSynchronize(st));
  int best = 0;
  float best_ms = 1e30f;
  for (size_t ci = 0; ci < cand.size(); ++ci) {
    float t = 1e30f;
    for (int rep = 0; rep < kReps; ++rep) {
      float ms = 0.f;
      TT_HIP(hipEventElapsedTime(&ms, ev[2 * (rep * cand.size() + ci)], ev[2 * (rep * cand.size() + ci) + 1]));
      t = std::min(t, ms);
    }
    if (t < best_ms) best_ms = t, best = cand[ci];
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  // a scratch solve that stopped early timed empty launches: keep the default
  s.grid_a[prec] = tw.host_state->done ? resident : best;
  // then the unknown kernel's workgroup size on that grid (results do not
  // depend on it: its partial sums are per fixed group of unknowns)
  const int nbc[3] = {256, 512, 1024};
  std::vector<hipEvent_t> ev2(2 * 3 * kReps);
  for (auto &e : ev2) TT_HIP(hipEventCreate(&e));
  for (int rep = 0; rep < kReps; ++rep)
    for (int ci = 0; ci < 3; ++ci) {
      s.nb_b[prec] = nbc[ci];
      TT_HIP(hipEventRecord(ev2[2 * (rep * 3 + ci)], st));
      for (int i = 0; i < kIts; ++i, ++k) {
        stepA(s, tw, k, prec, st, nullptr, nullptr);
        stepB(s, tw, k, prec, st);
      }
      TT_HIP(hipEventRecord(ev2[2 * (rep * 3 + ci) + 1], st));
    }
  TT_LAUNCH_CHECK();
  TT_HIP(hipMemcpyAsync(tw.host_state, tw.state.p + ((k - 1) & 1), sizeof(LsqrState), hipMemcpyDeviceToHost, st));
  TT_HIP(hipStreamSynchronize(st));
  int best_nb = kNB;
  float best_nb_ms = 1e30f;
  for (int ci = 0; ci < 3; ++ci) {
    float t = 1e30f;
    for (int rep = 0; rep < kReps; ++rep) {
      float ms = 0.f;
      TT_HIP(hipEventElapsedTime(&ms, ev2[2 * (rep * 3 + ci)], ev2[2 * (rep * 3 + ci) + 1]));
      t = std::min(t, ms);
    }
    if (t < best_nb_ms) best_nb_ms = t, best_nb = nbc[ci];
  }
  for (auto &e : ev2) (void)hipEventDestroy(e);
  s.nb_b[prec] = tw.host_state->done ? kNB : best_nb;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[key] = std::make_pair(s.grid_a[prec], s.nb_b[prec]);
  if (forced_nb > 0) s.nb_b[prec] = forced_nb;
}

LsqrResult block_lsqr(BlockSys &s, LsqrWork &wk, const double *b, double *x, int maxit, float eps4, int prec,
                      hipStream_t st, LsqrProbe *probe, bool graphs, int slot) {
  const int m = s.ncell * s.R, n = s.n;
  if (slot < 0 || slot > 2) throw std::runtime_error("solve history slot must be 0..2");
  if (prec == kBlocks32 && !s.has32) throw std::runtime_error("single-precision blocks not built");
  if (prec == kBlocks16 && !s.has16) throw std::runtime_error("half-precision blocks not built");
  block_tune(s, prec, st);
  wk.ensure(m, n, s.ncell, s.L);
  LsqrState init;
  std::memset(&init, 0, sizeof(init));
  init.cs2 = -1.0;
  init.atol = (double)eps4;
  init.ctol = 1.0 / 1.0e8;
  init.itnlim = maxit == 0 ? 10 * (m + n + 50) : maxit;
  LsqrState init2[2] = {init, init};
  TT_HIP(hipMemcpyAsync(wk.state.p, init2, sizeof(init2), hipMemcpyHostToDevice, st));
  TT_HIP(hipMemsetAsync(wk.x.p, 0, sizeof(double) * n, st));
  r0(s, wk, b, x, st);
  TT_LAUNCH_CHECK();
  stepB(s, wk, 0, prec, st);
  TT_LAUNCH_CHECK();
  // recent solves of this slot predict the count (the Picard loop changes
  // only the right-hand side and, through the dynamic tolerance, eps)
  int k = 1, batch = 8;
  // termination after nit iterations shows two launches later; two more
  // cover a small rise (an undershoot costs a host round trip, an idle
  // launch a few microseconds)
  const int pred = s.hist[slot].predict();
  if (pred >= 0) batch = pred + 4;
  const int more = std::max(4, pred / 8);
  int extra = 0;  // batches after the first
  LsqrGraphs &G = s.graphs[prec];
  if (graphs && (G.wk != &wk || G.gen != wk.gen || G.grid != s.grid_a[prec] || G.nb != s.nb_b[prec])) {
    G.reset();
    G.wk = &wk, G.gen = wk.gen, G.grid = s.grid_a[prec], G.nb = s.nb_b[prec];
  }
  // one iteration on the stream (sampled launches record their own events)
  auto one = [&](long pi) {
    stepA(s, wk, k, prec, st, pi >= 0 ? probe->pool[2 * pi] : nullptr, pi >= 0 ? probe->pool[2 * pi + 1] : nullptr);
    stepB(s, wk, k, prec, st);
    ++k;
  };
  // replays `pending` (even) iterations from k as chunks of 64, 32, .., 2
  auto replay = [&](int pending) {
    for (int lvl = LsqrGraphs::kLevels - 1; pending > 0; --lvl) {
      const int len = 2 << lvl;
      while (pending >= len) {
        if (!G.exec[lvl]) G.exec[lvl] = capture_chunk(s, wk, prec, len, st);
        TT_HIP(hipGraphLaunch(G.exec[lvl], st));
        k += len, pending -= len;
      }
    }
  };
  LsqrResult res;
  for (;;) {
    if (!graphs) {
      for (int b2 = 0; b2 < batch; ++b2) one(probe_begin(probe, k));
    } else {
      // even counts from odd indices; a pair holding a sampled cell-kernel
      // launch runs on the stream
      int pending = 0;
      for (int b2 = 0; b2 < batch; b2 += 2) {
        const long p0 = probe_begin(probe, k + pending), p1 = probe_begin(probe, k + pending + 1);
        if (p0 < 0 && p1 < 0) {
          pending += 2;
          continue;
        }
        replay(pending);
        pending = 0;
        one(p0);
        one(p1);
      }
      replay(pending);
    }
    TT_LAUNCH_CHECK();
    TT_HIP(hipMemcpyAsync(wk.host_state, wk.state.p + ((k - 1) & 1), sizeof(LsqrState), hipMemcpyDeviceToHost, st));
    TT_HIP(hipStreamSynchronize(st));
    if (wk.host_state->done) {
      res.nit = wk.host_state->itn;
      res.istop = wk.host_state->istop;
      break;
    }
    if (k > init.itnlim + 3) throw std::runtime_error("LSQR did not signal termination");
    // after a predicted first batch the remainder is usually short: small
    // batches, doubling; without a prediction a quarter of the count so far
    batch = pred >= 0 ? std::min(64, more << std::min(extra, 4)) : std::max(more, std::min(64, k / 4));
    ++extra;
  }
  probe_collect(probe, res.nit);
  s.hist[slot].add(res.nit);
  // x = q * dx + x0 (x0 is still in x: the kernels only wrote wk.x)
  hipLaunchKernelGGL(finish_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, s.q.p, wk.x.p, x, x);
  TT_LAUNCH_CHECK();
  return res;
}

void block_lsqr_multi(std::vector<LsqrJob> &jobs, int prec, hipStream_t st) {
  if (jobs.empty()) return;
  if (jobs.size() > (size_t)kMaxJobs) throw std::runtime_error("too many systems for one batched LSQR");
  const int R = jobs[0].s->R;
  int itnlim_max = 0, pred = 0;
  for (LsqrJob &j : jobs) {
    BlockSys &s = *j.s;
    LsqrWork &wk = *j.wk;
    if (s.R != R) throw std::runtime_error("batched LSQR systems must share the block shape");
    if (j.slot < 0 || j.slot > 2) throw std::runtime_error("solve history slot must be 0..2");
    if (prec == kBlocks32 && !s.has32) throw std::runtime_error("single-precision blocks not built");
    if (prec == kBlocks16 && !s.has16) throw std::runtime_error("half-precision blocks not built");
    for (const LsqrJob &o : jobs)
      if (&o != &j && o.wk == j.wk) throw std::runtime_error("batched LSQR systems need separate work areas");
    const int m = s.ncell * s.R, n = s.n;
    block_tune(s, prec, st);
    wk.ensure(m, n, s.ncell, s.L);
    itnlim_max = std::max(itnlim_max, j.maxit == 0 ? 10 * (m + n + 50) : j.maxit);
    // the batch runs until its slowest system stops: the largest prediction
    const int p = s.hist[j.slot].predict();
    pred = (p < 0 || pred < 0) ? -1 : std::max(pred, p);
  }
  // systems with bit-identical reduced-precision blocks share their cell-kernel launches
  for (size_t i = 0; i < jobs.size(); ++i) {
    jobs[i].share_sys = nullptr;
    if (prec == kBlocks64 || !g_sharing.load()) continue;
    for (size_t q = 0; q < i; ++q)
      if (!jobs[q].share_sys && lo_blocks_equal(*jobs[i].s, *jobs[q].s, prec, st)) {
        jobs[i].share_sys = jobs[q].s;
        break;
      }
  }
  std::vector<LsqrJob *> act;
  for (LsqrJob &j : jobs) act.push_back(&j);
  // starting states (as block_lsqr's), initial residuals, iteration 0's
  // unknown kernels: one launch each for all systems
  {
    Jobs<InitArgs> I;
    Jobs<R0Args> Z;
    I.n = Z.n = (int)jobs.size();
    I.first[0] = Z.first[0] = 0;
    for (int i = 0; i < I.n; ++i) {
      LsqrJob &j = jobs[i];
      BlockSys &s = *j.s;
      const int m = s.ncell * s.R, n = s.n;
      I.a[i] = InitArgs{j.wk->state.p, j.wk->x.p, (double)j.eps4, n, j.maxit == 0 ? 10 * (m + n + 50) : j.maxit};
      I.first[i + 1] = I.first[i] + (int)ceil_div(n, 256);
      Z.a[i] = r0_args(s, *j.wk, j.b, j.x);
      Z.first[i + 1] = Z.first[i] + s.ncell;
    }
    hipLaunchKernelGGL(init_multi_kernel, dim3(I.first[I.n]), dim3(256), 0, st, I);
    if (R == 51)
      hipLaunchKernelGGL((lsqr_r0_multi<51, 51>), dim3(Z.first[Z.n]), dim3(64), 0, st, Z);
    else
      hipLaunchKernelGGL((lsqr_r0_multi<27, 27>), dim3(Z.first[Z.n]), dim3(64), 0, st, Z);
    TT_LAUNCH_CHECK();
    stepB_multi(act, 0, prec, st);
    TT_LAUNCH_CHECK();
  }
  int k = 1, batch = pred >= 0 ? pred + 4 : 8, extra = 0;
  const int more = std::max(4, pred / 8);
  for (;;) {
    for (int b2 = 0; b2 < batch; ++b2) {
      stepA_multi(act, k, prec, st);
      stepB_multi(act, k, prec, st);
      ++k;
    }
    TT_LAUNCH_CHECK();
    for (LsqrJob *j : act)
      TT_HIP(hipMemcpyAsync(j->wk->host_state, j->wk->state.p + ((k - 1) & 1), sizeof(LsqrState),
                            hipMemcpyDeviceToHost, st));
    TT_HIP(hipStreamSynchronize(st));
    std::vector<LsqrJob *> still;
    Jobs<FinArgs> F;
    F.n = 0;
    F.first[0] = 0;
    for (LsqrJob *j : act) {
      const LsqrState &hs = *j->wk->host_state;
      if (!hs.done) {
        still.push_back(j);
        continue;
      }
      // later launches return at once for a stopped system: its x is final
      j->res.nit = hs.itn;
      j->res.istop = hs.istop;
      j->s->hist[j->slot].add(j->res.nit);
      F.a[F.n] = FinArgs{j->s->q.p, j->wk->x.p, j->x, j->x, j->s->n};
      F.first[F.n + 1] = F.first[F.n] + (int)ceil_div(j->s->n, 256);
      ++F.n;
    }
    if (F.n > 0) {
      hipLaunchKernelGGL(finish_multi_kernel, dim3(F.first[F.n]), dim3(256), 0, st, F);
      TT_LAUNCH_CHECK();
    }
    act.swap(still);
    if (act.empty()) break;
    if (k > itnlim_max + 3) throw std::runtime_error("LSQR did not signal termination");
    batch = pred >= 0 ? std::min(64, more << std::min(extra, 4)) : std::max(more, std::min(64, k / 4));
    ++extra;
  }
}

void block_sharing(int on) { g_sharing = on ? 1 : 0; }
void block_sharing_stats(long long *launches, long long *systems) {
  if (launches) *launches = g_shared_launches.load();
  if (systems) *systems = g_shared_systems.load();
}

}  // namespace tt
