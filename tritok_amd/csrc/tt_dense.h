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

struct DenseKin {
  int nb = 0, nn = 0;
  SdDev mats;                 // H, Htx, Hty, Htz (nb x nb), Dx, Dy, Dz (nb x nn), FP64
  DBuf<double> Ainv[3];       // inverse of the system matrix of each equation
  // low-precision copies for the product (per-row power-of-two scale)
  SdPrec prec = SdPrec::F64;
  DBuf<uint16_t> lowH[4], lowD[3];  // bf16/fp16 payloads
  DBuf<float> f32H[4], f32D[3];
  DBuf<float> rscale;               // 7 x nb row scales (2^k)
  // per-iteration work
  DBuf<double> YH, YD, part, dfi, xs, dw, vold;
  DBuf<int> kode;             // nb x 3
  DBuf<int> isbnd;            // nn: boundary index + 1 or 0
  // anchored delta products for reduced precision
  DBuf<double> vanchor, wanchor;  // fields the current products correspond to
  bool anchored = false;
  int since_anchor = 0;
};

// Builds A_e, factorises and inverts it (rocSOLVER getrf/getri).
void dense_setup(DenseKin &d, const DevMesh &m, const int *kode_host, SdPrec prec, hipStream_t s, double *ms_lu);

// Boundary-vorticity step for eqn 1..3 in order: vel (nn,3) read, vort (nn,3)
// boundary entries updated with relaxation urBw.
void dense_bw_step(DenseKin &d, const DevMesh &m, const double *vel, double *vort, const double urBw[3],
                   hipStream_t s);

}  // namespace tt
