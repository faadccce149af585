// tt_compress.h -- device-resident compressed operators: cross approximation
// (lib_ACA.f), stretched-Haar wavelet compression (lib_WT.f) and the block
// (H-) matrix that combines them (lib_ctree.f).
#pragma once
#include <vector>

#include "tt_hip.h"

namespace tt {

// Stretched Haar transform of one sequence of length nv: padded length
// N = 2^(kmax+1), first duplicated entry podv, and the level factors
// norma = 1/sqrt(N), valcek[k] = 2^(k/2) * norma (host-computed so the device
// uses exactly the reference's values).
struct HaarTab {
  int nv, kmax, N, podv;
  double norma;
  double valcek[32];
};
HaarTab haar_tab(int nv);

// Wavelet-compressed matrix: CSR (1-based, as Che_FormCRS) of W A W^T.
struct WtDev {
  int m = 0, n = 0, nnz = 0;
  DBuf<int> zvr, c;
  DBuf<double> v;
  DBuf<double> xw, yw;  // scratch of one product
};
// Che_WaWT + Che_AbsPovp + Che_Zanemari + Che_FormCRS of the device matrix
// A (m x n, leading dimension lda); returns the reference's nnz count.
int wt_compress_dev(const double *A, long lda, int m, int n, double kappa, WtDev &out, hipStream_t s);
// Che_WTAWx on device vectors.
void wt_mv_dev(WtDev &w, const double *x, double *y, hipStream_t s);

// Cross approximation of the device matrix R (m x n, column-major,
// overwritten): factors a (m x rank) and b (rank x n) column-major.
struct AcaDev {
  int m = 0, n = 0, rank = 0;
  DBuf<double> a, b, t;
  DBuf<double> part;  // column-slice partial sums of the two products
};
int aca_compress_dev(double *R, int m, int n, int type, double eps, const double *vec_host, AcaDev &out,
                     double *norm, hipStream_t s);
void aca_mv_dev(AcaDev &h, const double *x, double *y, hipStream_t s);
// the product's work buffers for factors set by hand (m, n, rank, a, b)
void aca_product_buffers(AcaDev &h);

// Block matrix: parts with row / column node lists (1-based), each dense,
// wavelet (ctype 2) or cross-approximated (ctype 3).
struct HPart {
  int nr, nc, ctype, rank;
  long roff, coff;  // into the row / column node lists
  long doff;        // dense values, or ACA a then b, or wavelet CSR values
  long ioff;        // wavelet zvr then c
  long poff;        // product scratch (nr * nrhs)
  long soff;        // wavelet scratch (padded row + column lengths)
};
struct HmatDev {
  int nrow = 0, ncol = 0, nparts = 0, max_rank = 0;
  std::vector<HPart> hparts;
  std::vector<int> hrn, hcn;  // node lists on the host (export)
  DBuf<HPart> parts;
  DBuf<int> rn, cn, ipool, rowptr, rowent;
  DBuf<double> dpool, prod, wscr;
  DBuf<HaarTab> tabs;  // per part: column table then row table (wavelet parts)
  long stored = 0;     // doubles held by the compressed form
};

}  // namespace tt

struct tt_hmat;
namespace tt {
// tt_hmat_create's block matrix built from a matrix already on the device
// (A: nrow x ncol, leading dimension lda; xr, xc: host point coordinates,
// n x 3 column-major); the caller's stream work on A must be complete.
tt_hmat *hmat_from_device(int nrow, int ncol, const double *A, long lda, const double *xr, const double *xc,
                          int leaf, double eta, int ctype, int aca_type, double aca_eps, double wt_kappa);
// y = H x for device vectors x (ncol) and y (nrow) on stream s.
void hmat_apply(tt_hmat *h, const double *x, double *y, hipStream_t s);
// doubles the compressed form holds
long hmat_stored_entries(const tt_hmat *h);
}  // namespace tt
