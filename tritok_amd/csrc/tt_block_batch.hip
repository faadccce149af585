// tt_block_batch.hip -- many independent subdomain least-squares systems
// solved on the cell-block format: slvlsqr2 (lib_Solve_v11.f:4401-4500) with
// prelsqr2's column scaling (:4501-4535) for every system of a batch, each
// system held as per-cell dense blocks (tt_lsqr.h) rather than CSR.
//
// The subdomain systems' rows come in groups of R (a cell's equations) that
// involve only the cell's own (at most L = R) unknowns.  The blocks hold each
// matrix entry once and the fused cell kernel reads a block once for both
// A v and A^T u', where a CSR solver reads values and indices for A and again
// (a transposed copy) for A^T.  Systems are solved in groups of up to 32 of
// similar size: every LSQR iteration of a group is one cell-kernel launch and
// one unknown-kernel launch for all its systems (block_lsqr_multi), each
// system with its own scalars, stop and work area, so a system's iterates do
// not depend on the group it runs in.
#include <algorithm>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <vector>

#include "tritok_bem.h"
#include "tt_lsqr.h"

using namespace tt;

struct tt_block_batch {
  int R = 0, prec = kBlocks64;
  std::vector<std::unique_ptr<BlockSys>> sys;
  std::vector<std::unique_ptr<LsqrWork>> work;
  std::vector<int> m, n;
  std::vector<size_t> moff, noff;  // offsets of the systems' rows and unknowns
  std::vector<int> order;          // systems by decreasing row count
  DBuf<double> b, x;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~tt_block_batch() {
    if (s) (void)hipStreamSynchronize(s);
    sys.clear();
    work.clear();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
  }
};

namespace {

constexpr int kGroup = 32;  // systems per launch (block_lsqr_multi's limit)

// The cell blocks of one CSR system whose rows come in groups of R sharing
// their columns: per cell the distinct columns of its rows in order of first
// appearance (colmap, 0-based unknowns, -1 padded to L) and the R x L block
// (row-major, zero where a row has no entry).
void csr_to_blocks(int m, int n, int R, int L, const int32_t *ia, const int32_t *ja, const double *a,
                   std::vector<int> &colmap, std::vector<double> &full) {
  if (m % R != 0) throw std::runtime_error("rows are not a whole number of cells");
  const int nc = m / R;
  colmap.assign((size_t)nc * L, -1);
  full.assign((size_t)nc * R * L, 0.0);
  std::vector<int> slot(n, -1);
  for (int c = 0; c < nc; ++c) {
    int used = 0;
    int *cm = colmap.data() + (size_t)c * L;
    for (int r = 0; r < R; ++r) {
      const int row = c * R + r;
      for (int k = ia[row] - 1; k < ia[row + 1] - 1; ++k) {
        const int j = ja[k] - 1;
        if (j < 0 || j >= n) throw std::runtime_error("column index out of range");
        if (slot[j] < 0) {
          if (used == L) throw std::runtime_error("a cell's rows use more columns than the block holds");
          slot[j] = used;
          cm[used++] = j;
        }
        full[((size_t)c * R + r) * L + slot[j]] += a[k];
      }
    }
    for (int l = 0; l < used; ++l) slot[cm[l]] = -1;
  }
}

}  // namespace

extern "C" int tt_block_batch_setup(int32_t nsys, const int32_t *m, const int32_t *n, const int64_t *nnz,
                                    const int32_t *ia, const int32_t *ja, const double *a, int32_t rows_per_cell,
                                    tt_block_batch **out) {
  return guarded("tt_block_batch_setup", [&] {
    if (!m || !n || !nnz || !ia || !ja || !a || !out) throw std::runtime_error("NULL argument");
    if (nsys < 1) throw std::runtime_error("nsys must be positive");
    if (rows_per_cell != 51 && rows_per_cell != 27) throw std::runtime_error("rows_per_cell must be 51 or 27");
    auto h = std::make_unique<tt_block_batch>();
    h->R = rows_per_cell;
    const int R = rows_per_cell, L = rows_per_cell;
    TT_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
    TT_HIP(hipEventCreate(&h->e0));
    TT_HIP(hipEventCreate(&h->e1));
    h->moff.assign(nsys + 1, 0);
    h->noff.assign(nsys + 1, 0);
    size_t ipos = 0, epos = 0;
    std::vector<int> colmap;
    std::vector<double> full;
    for (int k = 0; k < nsys; ++k) {
      if (m[k] < 1 || n[k] < 1 || nnz[k] < 1) throw std::runtime_error("empty system");
      if (ia[ipos] != 1 || ia[ipos + m[k]] - 1 != nnz[k]) throw std::runtime_error("row pointers inconsistent");
      csr_to_blocks(m[k], n[k], R, L, ia + ipos, ja + epos, a + epos, colmap, full);
      auto bs = std::make_unique<BlockSys>();
      DBuf<double> af;
      af.upload(full, h->s);
      block_setup(*bs, m[k] / R, R, L, n[k], colmap, af.p, false, h->s);  // synchronises: af may go
      h->sys.push_back(std::move(bs));
      h->work.push_back(std::make_unique<LsqrWork>());
      h->m.push_back(m[k]), h->n.push_back(n[k]);
      h->moff[k + 1] = h->moff[k] + m[k], h->noff[k + 1] = h->noff[k] + n[k];
      ipos += (size_t)m[k] + 1, epos += (size_t)nnz[k];
    }
    h->order.resize(nsys);
    std::iota(h->order.begin(), h->order.end(), 0);
    std::stable_sort(h->order.begin(), h->order.end(), [&](int p, int q) { return h->m[p] > h->m[q]; });
    h->b.alloc(h->moff[nsys]);
    h->x.alloc(h->noff[nsys]);
    *out = h.release();
    return 0;
  });
}

extern "C" int tt_block_batch_set_precision(tt_block_batch *h, int32_t prec) {
  return guarded("tt_block_batch_set_precision", [&] {
    if (!h) throw std::runtime_error("NULL batch");
    if (prec != TT_PREC_FP64 && prec != TT_PREC_FP32 && prec != TT_PREC_FP16)
      throw std::runtime_error("precision must be TT_PREC_FP64, TT_PREC_FP32 or TT_PREC_FP16");
    const int bp = prec == TT_PREC_FP64 ? kBlocks64 : prec == TT_PREC_FP32 ? kBlocks32 : kBlocks16;
    for (auto &s : h->sys) {
      if (bp == kBlocks32) block_make32(*s, h->s);
      if (bp == kBlocks16) block_make16(*s, h->s);
    }
    TT_HIP(hipStreamSynchronize(h->s));
    h->prec = bp;
    return 0;
  });
}

extern "C" int tt_block_batch_solve(tt_block_batch *h, const double *b, double *x, int32_t maxit, float eps4,
                                    int32_t *nit, int32_t *istop, double *ms) {
  return guarded("tt_block_batch_solve", [&] {
    if (!h || !b || !x || !nit || !istop) throw std::runtime_error("NULL argument");
    const int ns = (int)h->sys.size();
    h->b.upload(b, h->moff[ns], h->s);
    h->x.upload(x, h->noff[ns], h->s);
    std::vector<LsqrJob> jobs;
    TT_HIP(hipEventRecord(h->e0, h->s));
    for (int g0 = 0; g0 < ns; g0 += kGroup) {
      jobs.clear();
      for (int i = g0; i < std::min(ns, g0 + kGroup); ++i) {
        const int k = h->order[i];
        LsqrJob j;
        j.s = h->sys[k].get(), j.wk = h->work[k].get();
        j.b = h->b.p + h->moff[k], j.x = h->x.p + h->noff[k];
        j.maxit = maxit, j.eps4 = eps4, j.slot = 0;
        jobs.push_back(j);
      }
      block_lsqr_multi(jobs, h->prec, h->s);
      for (size_t q = 0; q < jobs.size(); ++q) {
        const int k = h->order[g0 + (int)q];
        nit[k] = jobs[q].res.nit, istop[k] = jobs[q].res.istop;
      }
    }
    TT_HIP(hipEventRecord(h->e1, h->s));
    h->x.download(x, h->noff[ns], h->s);
    TT_HIP(hipStreamSynchronize(h->s));
    if (ms) {
      float f = 0.f;
      TT_HIP(hipEventElapsedTime(&f, h->e0, h->e1));
      *ms = f;
    }
    return 0;
  });
}

extern "C" void tt_block_batch_free(tt_block_batch *h) { delete h; }

extern "C" int tt_lsqr_set_block_sharing(int32_t on) {
  return guarded("tt_lsqr_set_block_sharing", [&] {
    if (on < 0 || on > 1) throw std::runtime_error("on must be 0 or 1");
    block_sharing(on);
    return 0;
  });
}

extern "C" int tt_lsqr_block_sharing_stats(int64_t *launches, int64_t *systems) {
  return guarded("tt_lsqr_block_sharing_stats", [&] {
    if (!launches || !systems) throw std::runtime_error("NULL argument");
    long long l = 0, n = 0;
    block_sharing_stats(&l, &n);
    *launches = l, *systems = n;
    return 0;
  });
}
