// tt_integrals.h -- device integral assembly (macro per-cell matrices and the
// dense single-domain kinematics matrices), used by the C-ABI entry points and
// by the whole-case driver.
#pragma once
#include "tritok_bem.h"
#include "tt_hip.h"
#include "tt_quadrature.h"

namespace tt {

// Mesh arrays on the device (reference layouts, 1-based indices).
struct DevMesh {
  int nn = 0, nb = 0, ne = 0, nc = 0, nq = 0;
  DBuf<double> x, xq, nx, ny, nz;
  DBuf<int> idc, ibc, ibf, flpm, gbn, lbn;
};

// Raw pointers of a DevMesh for kernel arguments.
struct MeshPtrs {
  int nn, nb, ne, nc, nq;
  const double *x, *xq;
  const int *idc, *ibc, *ibf, *flpm, *gbn, *lbn;
};

void upload_mesh(const tt_mesh_view &v, DevMesh &d, hipStream_t s);
MeshPtrs ptrs(const DevMesh &d);

// Host-side quadrature rules from the C-ABI description (the host's table
// when given, computed Gauss-Legendre values otherwise).
void build_quad(const tt_quad &q, QuadDev &out);
int gauss_legendre(int rule, double *gi, double *ome);

// Macro matrices, column-major, on the device.
struct MacroDev {
  int nc = 0;
  DBuf<double> H, G, B, Ax, Ay, Az, Htx, Hty, Htz, Dx, Dy, Dz;
};
// mun/muq (device, nnodes / nq): modified-Helmholtz mu of every node / flux
// node as source (FMATmodHelm); NULL for the Laplace kernel (FMATkmdcSteady).
void macro_integrals_dev(const DevMesh &m, const QuadDev &qh, int copy, MacroDev &out, double ont[11],
                         hipStream_t s, const double *mun = nullptr, const double *muq = nullptr,
                         double *mumax = nullptr);

// Single-domain kinematics rows row0..row1-1, column-major with ld = rows.
struct SdDev {
  int row0 = 0, row1 = 0;
  DBuf<double> H, Htx, Hty, Htz, Dx, Dy, Dz;
};
void sd_integrals_dev(const DevMesh &m, const tt_mesh_view &hv, const QuadDev &qh, int row0, int row1, SdDev &out,
                      double ont[10], hipStream_t s);

}  // namespace tt
