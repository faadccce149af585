// tt_lsqr.h -- LSQR (Paige & Saunders, as in lib_Solve_v11.f:3128-3887 with
// the hot start and column scaling of slvlsqr2/prelsqr2 :4401-4535) on the
// device, for two storage formats:
//
//  * BlockSys: the subdomain systems.  Every cell contributes R equations that
//    involve only its L local variables, so A is a list of dense R x L blocks
//    with a per-cell column map (local column -> unknown, -1 = known).  One
//    pass over a block yields both A*v and A^T*(A*v - alpha*u) for that cell.
//  * CsrSys: a general CSR matrix (the C-ABI entry points), with a transposed
//    copy for A^T*u.
#pragma once
#include <utility>
#include <vector>

#include "tt_hip.h"

namespace tt {

// LSQR scalar state (device memory, two copies): the cell kernel of
// iteration k reads copy (k-1)&1 and its block 0 writes copy k&1; the
// unknown kernel of iteration k reads copy k&1 and its block 0 stores
// beta_{k+1} there for the next cell kernel.  No kernel writes a copy that
// another block of the same launch reads.
struct LsqrState {
  double alpha, beta;                 // alpha_k and beta_k (beta_{k+1} after the unknown kernel)
  double rhobar, phibar, anorm, dnorm, xnorm, xnorm1, cs2, sn2, z, res2, rnorm, arnorm, bnorm, acond;
  double t1, t2, t3;                  // update coefficients of the last rotation
  double atol, btol, ctol, nrm_b, nrm_r;
  double pending_phi, pending_rho, pending_theta, pending_tau;  // rotation awaiting its test
  int itn, istop, done, itnlim;
  int kseq;                           // iteration index of the cell kernel that wrote this copy (0 = initial)
};

// Block storage: per cell an R x L row-major block, padded to `stride`
// elements (a multiple of 4) so that every block starts on a 16-byte boundary
// in both precisions and is read with 16-byte vector loads.
inline int block_stride(int R, int L) { return (R * L + 3) / 4 * 4; }
// 16-bit blocks: rows padded to a multiple of 4 entries (8 bytes), so that a
// row is read with 8-byte loads, and the block to a multiple of 8 entries
// (16 bytes) for 16-byte block loads; the padding holds zeros
inline int block_row16(int L) { return (L + 3) / 4 * 4; }
inline int block_stride16(int R, int L) { return (R * block_row16(L) + 7) / 8 * 8; }

// precision of the blocks an LSQR solve iterates with
enum BlockPrec { kBlocks64 = 0, kBlocks32 = 1, kBlocks16 = 2 };

struct LsqrWork;

// Captured iteration chunks of one system for one work area and block
// precision: exec[i] runs 2^(i+1) iterations (cell kernel + unknown kernel
// each) starting at an odd iteration index.  Move-only; destroys its graphs.
struct LsqrGraphs {
  static constexpr int kLevels = 6;  // chunks of 2 .. 64 iterations
  const LsqrWork *wk = nullptr;      // work area the graphs were captured with
  unsigned gen = 0;                  // its allocation generation at capture
  int grid = 0;                      // cell-kernel grid at capture
  int nb = 0;                        // unknown-kernel threads per workgroup at capture
  hipGraphExec_t exec[kLevels] = {};
  LsqrGraphs() = default;
  LsqrGraphs(const LsqrGraphs &) = delete;
  LsqrGraphs &operator=(const LsqrGraphs &) = delete;
  LsqrGraphs(LsqrGraphs &&o) noexcept;
  LsqrGraphs &operator=(LsqrGraphs &&o) noexcept;
  void reset();
  ~LsqrGraphs() { reset(); }
};

struct BlockSys {
  int ncell = 0, R = 0, L = 0, n = 0, stride = 0;
  DBuf<double> A;         // scaled blocks: ncell*stride
  DBuf<float> A32;        // the same in single precision (mixed-precision solves)
  DBuf<_Float16> A16;     // the same in half precision, rows padded to block_row16
  DBuf<int> colmap;       // ncell*L
  DBuf<int> tptr, tidx;   // for unknown j: entries tidx[tptr[j]..tptr[j+1]) = cell*L + l
  int gw = 0;             // gather width: max cells per unknown, rounded up to 4
  DBuf<int> gtab;         // n*gw: the entries of tidx per unknown in cell order, -1 padded
  DBuf<double> q;         // column scaling
  bool has32 = false, has16 = false;
  int grid_a[3] = {0, 0, 0};  // cell-kernel blocks per block precision (BlockPrec), 0 = resident blocks
  int nb_b[3] = {0, 0, 0};    // unknown-kernel threads per workgroup per block precision, 0 = 256
  // recent solves per right-hand-side slot (a system solved for several
  // field components keeps one history per component): their iteration
  // counts size the first launch batch of the next solve
  struct SolveHistory {
    static constexpr int kLen = 4;
    int nit[kLen] = {};
    int n = 0, head = 0;  // entries held; next entry to write
    void add(int it);
    int predict() const;  // -1 without history
  };
  SolveHistory hist[3];
  LsqrGraphs graphs[3];   // captured iteration chunks per block precision
  // identity for batched solves: systems whose reduced-precision blocks,
  // column maps and gather tables are equal bit for bit share one block
  // stream per launch (block_lsqr_multi); comparisons made so far, per block
  // precision: (other system's uid, equal)
  unsigned long long uid = 0;
  std::vector<std::pair<unsigned long long, bool>> twins[3];
};

// Builds the solver structure from the full unscaled blocks on the device
// (Afull: ncell*R*L, including known columns), scaling unknown columns by
// q_j = 1/||A(:,j)|| (prelsqr2).  colmap given on the host.
void block_setup(BlockSys &s, int ncell, int R, int L, int n, const std::vector<int> &colmap, const double *Afull,
                 bool want32, hipStream_t st);

struct LsqrWork {
  int m = 0, n = 0, ncell = 0, L = 0;
  DBuf<double> u, v, w, x, P;  // P: FP32 values (first half) while a solve runs on FP16 blocks
  DBuf<double> pa1, pa2;    // partial sums of the cell kernels (per cell for r0, per block after)
  int na = 0;               // partials written by the last cell kernel
  DBuf<double> pb1, pb2;    // per-block partial sums of the unknown kernel
  DBuf<LsqrState> state;    // two copies
  LsqrState *host_state = nullptr;  // pinned
  unsigned gen = 0;         // bumped whenever a buffer above is reallocated
  void ensure(int m, int n, int ncell, int L);
  ~LsqrWork();
};

// Samples the duration of every stride-th lsqr_A launch with a pair of HIP
// events on the solve's stream (stride 0 = off).
struct LsqrProbe {
  int stride = 0;
  unsigned long long seen = 0;     // lsqr_A launches issued while enabled
  unsigned long long sampled = 0;  // launches timed
  double total_ms = 0.0;           // summed duration of the timed launches
  std::vector<hipEvent_t> pool;    // event pool, pairs (start, stop)
  std::vector<int> iter;           // iteration index of each recorded pair
  size_t used = 0;                 // pairs recorded and not yet read
  ~LsqrProbe();
};

struct LsqrResult {
  int nit = 0, istop = 0;
};

// slvlsqr2 on a block system: b (m = ncell*R) on the device, x holds the hot
// start on entry and the solution on exit (device).  prec (BlockPrec) selects
// the blocks the iterations read (FP64; FP32 or FP16 with FP32 arithmetic);
// the initial residual b - A*x0 always uses the double-precision blocks.
// graphs: the iterations are replayed from captured chunks (s.graphs) rather
// than launched one kernel at a time; the kernels and their arguments are the
// same either way, and so are the results.
// slot (0..2) selects the solve history that sizes the launch batches.
LsqrResult block_lsqr(BlockSys &s, LsqrWork &wk, const double *b, double *x, int maxit, float eps4, int prec,
                      hipStream_t st, LsqrProbe *probe = nullptr, bool graphs = true, int slot = 0);

// One system of a batched solve (block_lsqr_multi): as block_lsqr's
// arguments; res is filled in.
struct LsqrJob {
  BlockSys *s = nullptr;
  const BlockSys *share_sys = nullptr;  // set by block_lsqr_multi: an earlier system with the same blocks
  LsqrWork *wk = nullptr;  // a work area of its own
  const double *b = nullptr;
  double *x = nullptr;
  int maxit = 0;
  float eps4 = 0.f;
  int slot = 0;
  LsqrResult res;
};

// block_lsqr on up to 32 systems of one block shape at once, all on stream
// st: every iteration runs the cell kernels of the systems still iterating in
// one launch and their unknown kernels in one launch per gather width.  The
// kernel bodies, grids per system and reductions are those of block_lsqr, so
// each system's iterates, stop and result are bit for bit those of a solve on
// its own.
// Systems whose FP32 or FP16 blocks, column maps and gather tables compare
// equal bit for bit read one copy of the blocks; when every system of a
// cell-kernel launch does, their workgroups are interleaved so that the
// systems' workgroups on the same cells run together on one XCD (each block
// fetched once into that L2).  Per-system arithmetic is unchanged.  on = 0
// turns this off (process-wide); the counters report the interleaved
// launches and the systems they carried.
void block_lsqr_multi(std::vector<LsqrJob> &jobs, int prec, hipStream_t st);
void block_sharing(int on);
void block_sharing_stats(long long *launches, long long *systems);

// Picks the cell kernel's grid for block precision prec by timing candidate
// grids on a scratch solve (once per system shape and precision in a process).
void block_tune(BlockSys &s, int prec, hipStream_t st);

// single- and half-precision copies of the scaled blocks (built on demand)
void block_make32(BlockSys &s, hipStream_t st);
void block_make16(BlockSys &s, hipStream_t st);

}  // namespace tt
