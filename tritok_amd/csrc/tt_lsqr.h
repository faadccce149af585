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
#include <vector>

#include "tt_hip.h"

namespace tt {

// LSQR scalar state, shared by the kernels of one solve (device memory).
struct LsqrState {
  double alpha, beta;                 // current alpha_k, beta_{k+1}
  double rhobar, phibar, anorm, dnorm, xnorm, xnorm1, cs2, sn2, z, res2, rnorm, arnorm, bnorm, acond;
  double t1, t2, t3;                  // update coefficients of the last rotation
  double atol, btol, ctol, nrm_b, nrm_r;
  double pending_phi, pending_rho, pending_theta, pending_tau;  // rotation awaiting its test
  int itn, istop, done, itnlim;
  unsigned counter;                   // last-block counter of the current kernel
  int pad;
};

struct BlockSys {
  int ncell = 0, R = 0, L = 0, n = 0;
  DBuf<double> A;         // scaled blocks, row-major per cell: ncell*R*L
  DBuf<float> A32;        // the same in single precision (mixed-precision solves)
  DBuf<int> colmap;       // ncell*L
  DBuf<int> tptr, tidx;   // for unknown j: entries tidx[tptr[j]..tptr[j+1]) = cell*L + l
  DBuf<double> q;         // column scaling
  bool has32 = false;
  int last_nit = 0;       // iterations of the previous solve (sizes the first launch batch)
};

// Builds the solver structure from the full unscaled blocks on the device
// (Afull: ncell*R*L, including known columns), scaling unknown columns by
// q_j = 1/||A(:,j)|| (prelsqr2).  colmap given on the host.
void block_setup(BlockSys &s, int ncell, int R, int L, int n, const std::vector<int> &colmap, const double *Afull,
                 bool want32, hipStream_t st);

struct LsqrWork {
  int m = 0, n = 0, ncell = 0, L = 0;
  DBuf<double> u, v, w, x, P, part1, part2, part3;
  DBuf<LsqrState> state;
  LsqrState *host_state = nullptr;  // pinned
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
  size_t used = 0;                 // pairs recorded and not yet read
  ~LsqrProbe();
};

struct LsqrResult {
  int nit = 0, istop = 0;
};

// slvlsqr2 on a block system: b (m = ncell*R) on the device, x holds the hot
// start on entry and the solution on exit (device).  use32 selects the
// single-precision blocks for the iterations; the initial residual
// b - A*x0 always uses the double-precision blocks.
LsqrResult block_lsqr(BlockSys &s, LsqrWork &wk, const double *b, double *x, int maxit, float eps4, bool use32,
                      hipStream_t st, LsqrProbe *probe = nullptr);

// single-precision copy of the scaled blocks (built on demand)
void block_make32(BlockSys &s, hipStream_t st);

}  // namespace tt
