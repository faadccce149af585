// tt_dense.h -- dense single-domain kinematics on the device: system matrices
// (sdKm_sysm, lib_sdkm.f:1171-1210), their factorisation (FormPREfm pret=2 ->
// dgefa), and the boundary-vorticity step of one Picard iteration (sdkm_solve
// for x, y, z: lib_sdkm.f:23-121, sdkm_rhs :349-519, sdkm_distunk
// :1075-1096) with fused products.
#pragma once
#include "tt_hip.h"
#include "tt_integrals.h"

namespace tt {

// Storage precision of the dense kinematics matrices for the products.
enum class SdPrec { F64 = 0, F32 = 1, BF16 = 2, F16 = 3 };

struct DenseMx;  // tile-major reduced-precision inverses (dense_refine)
void dense_mx_free(DenseMx *p);

struct DenseKin {
  int nb = 0, nn = 0;
  SdDev mats;                 // H, Htx, Hty, Htz (nb x nb), Dx, Dy, Dz (nb x nn), FP64
  DBuf<double> Ainv[3];       // inverse of the system matrix of each equation
  // per-iteration work
  DBuf<double> YH, YD, part, dfi, xs, dw;
  DBuf<int> kode;             // nb x 3
  DBuf<int> isbnd;            // nn: boundary index + 1 or 0
  // Reduced-precision form: the ten matrices (order H, Htx, Hty, Htz, Dx, Dy,
  // Dz, Ainv x, y, z) stored in row blocks of 256 rows (each block column
  // after column, so a product workgroup streams contiguous memory; ldl = nb
  // rounded up to 256) and a power-of-two scale per row (10 x ldl floats);
  // products are taken on the change of the fields since an FP64 anchor step.
  SdPrec prec = SdPrec::F64;
  int ldl = 0;
  DBuf<uint16_t> lo16[10];
  DBuf<float> lo32[10];
  DBuf<float> rscale;
  int anchor_period = 16;     // FP64 step every anchor_period steps
  int since_anchor = 0;       // 0: the next step is an anchor step
  DBuf<double> YH_a, YD_a, dfi_a, sol_a, vel_a, vort_a;
  DBuf<float4> dv4, dw4;      // FP32 changes of the fields since the anchor step
  DenseMx *mx = nullptr;      // made by dense_refine on first use
  DenseKin() = default;
  DenseKin(const DenseKin &) = delete;
  DenseKin &operator=(const DenseKin &) = delete;
  ~DenseKin() { dense_mx_free(mx); }
};

struct RefineResult {
  int iters = 0;
  double relres = 0.0;   // ||B - A X||_F / ||B||_F after the last iteration
  double ms = 0.0;       // device time of the refinement
  double ms_apply = 0.0;  // mean device time of one correction product M R
};

// Solves A_e X = B (e = 0, 1, 2; B, X device arrays n x nrhs, nrhs <= 4) by
// FP64 iterative refinement with the inverse applied in prec (F64: rocBLAS
// FP64; F32, BF16, F16: tile-major copies with per-row power-of-two scales
// on the matrix cores, FP32 accumulation).
RefineResult dense_refine(DenseKin &d, const DevMesh &mesh, int e, SdPrec prec, int nrhs, const double *B,
                          double *X, double tol, int maxit, hipStream_t s);

// Builds A_e, factorises and inverts it (rocSOLVER getrf/getri).
void dense_setup(DenseKin &d, const DevMesh &m, const int *kode_host, SdPrec prec, hipStream_t s, double *ms_lu);

// Boundary-vorticity step for eqn 1..3 in order: vel (nn,3) read, vort (nn,3)
// boundary entries updated with relaxation urBw.  In a reduced precision the
// step is an FP64 anchor step every anchor_period calls and a delta step
// otherwise.
void dense_bw_step(DenseKin &d, const DevMesh &m, const double *vel, double *vort, const double urBw[3],
                   hipStream_t s);

// Builds the reduced-precision copies (F32, F16 or BF16; F64 releases them).
void dense_set_precision(DenseKin &d, SdPrec prec, hipStream_t s);

}  // namespace tt
