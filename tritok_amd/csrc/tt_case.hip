// tt_case.hip -- whole-case driver: the reference's setup (main.f90:410-1061)
// and steady Picard loop (main.f90:1200-1589) with every matrix resident in
// HBM.  Order per iteration, as the reference:
//   boundary vorticities x,y,z (dense kinematics, tt_dense.hip)
//   domain velocities x,y,z    (KM systems, solv_kmMacro3, lib_km.f:214-303)
//   temperature                (SolveFDdae, lib_solvFDdae.f:168-288)
//   domain vorticities x,y,z   (SolveFDwTE, lib_solvFDwTE.f:190-332)
//   relative changes           (CheckConvergencevw, lib_util.f:1048-1150)
//   dynamic LSQR tolerance     (DLSE, lib_util.f:443-475)
// Subdomain systems are stored as per-cell blocks (tt_lsqr.h).  The known
// columns of a block are the reference's rhsm CSR (values negated), so the
// right-hand side b = rhsm*known + (matrix terms) is formed per cell row.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "tritok_bem.h"
#include "tt_dense.h"
#include "tt_integrals.h"
#include "tt_lsqr.h"

using namespace tt;

namespace {

// one subdomain least-squares system (VTE component, energy, or kinematics)
struct SubSys {
  int R = 0, L = 0, n = 0, ncell = 0;
  BlockSys bs;
  const double *Afull = nullptr;  // unscaled blocks incl. known columns (shared)
  DBuf<int> colfield;             // ncell*L: node index (l < 27) or flux index (l >= 27), 0-based
  DBuf<int> colknown;             // ncell*L: 1 if the column's value is known
  DBuf<int> unk_field;            // n: field index (node) or nn + flux index
  long long nnz = 0;              // structural nonzeros of the unknown columns
};

__global__ void afull_kernel(int nc, int R, int L, const double *__restrict__ H, const double *__restrict__ G,
                             const double *__restrict__ B, double beta, double ren, double *__restrict__ Af) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t total = (size_t)nc * R * L;
  if (k >= total) return;
  size_t c = k / ((size_t)R * L);
  int r = (int)(k / L % R), l = (int)(k % L);
  size_t m1 = (size_t)nc * 51, row = c * 51 + r;
  // sMat2crsSysRhsB_w_fill / _Temp_fill: H + B*beta*ren on node columns, -G on flux columns
  Af[k] = l < 27 ? H[row + m1 * l] + B[row + m1 * l] * beta * ren : -G[row + m1 * (l - 27)];
}

// x0 gather: x[u] = field value of unknown u
__global__ void gather_kernel(int n, int nn, const int *__restrict__ uf, const double *__restrict__ fnode,
                              const double *__restrict__ fflux, double *__restrict__ x) {
  int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  int f = uf[u];
  x[u] = f < nn ? fnode[f] : fflux[f - nn];
}

// solution scatter with relaxation of node values (fluxes copied)
__global__ void scatter_kernel(int n, int nn, double ur, const int *__restrict__ uf, const double *__restrict__ x,
                               double *__restrict__ fnode, double *__restrict__ fflux) {
  int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  int f = uf[u];
  if (f < nn)
    fnode[f] = (1.0 - ur) * fnode[f] + ur * x[u];
  else
    fflux[f - nn] = x[u];
}

// Known-column part of one cell row: the sum over the known columns of
// (-Afull) * value. The cell's known values are staged in LDS by its
// workgroup (thread l < L stages column l, with its known flag); every row of
// the cell reads them from there.
__device__ __forceinline__ void stage_known(int c, int L, const int *__restrict__ colknown,
                                            const int *__restrict__ colfield, const double *__restrict__ fnode,
                                            const double *__restrict__ fflux, double *kval, int *kflag) {
  const int l = threadIdx.x;
  if (l < L) {
    const int kn = colknown[(size_t)c * L + l];
    kflag[l] = kn;
    kval[l] = kn ? (l < 27 ? fnode[colfield[(size_t)c * L + l]] : fflux[colfield[(size_t)c * L + l]]) : 0.0;
  }
}
__device__ __forceinline__ double known_part(int c, int r, int R, int L, const double *__restrict__ Af,
                                             const double *kval, const int *kflag) {
  const double *a = Af + ((size_t)c * R + r) * L;
  double s = 0.0;
  for (int l = 0; l < L; ++l)
    if (kflag[l]) s = s + (-a[l]) * kval[l];
  return s;
}

// Ex, Ey, Ez: Ax - Dx, Ay - Dy, Az - Dz (the convection and buoyancy terms
// use only these differences; the case forms them once at setup, in place
// of Ax, Ay, Az)
struct MacroP {
  const double *B, *Ex, *Ey, *Ez, *Dx, *Dy, *Dz, *Htx, *Hty, *Htz;
};

__global__ void adv_diff_kernel(size_t n, double *__restrict__ a, const double *__restrict__ d) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = a[i] - d[i];
}

// SolveFDwTE right-hand side (lib_solvFDwTE.f:245-285) incl. BouyancyRHS; one
// 64-thread workgroup per cell: the cell's 27 nodal values and its known
// columns are staged in LDS once, each of the 51 rows then streams only its
// matrix entries
__global__ void __launch_bounds__(64) vte_rhs_kernel(int nc, int nn, int e, MacroP mp, const int *__restrict__ idc,
                                                     const double *__restrict__ Af, const int *__restrict__ colknown,
                                                     const int *__restrict__ colfield, const double *__restrict__ vel,
                                                     const double *__restrict__ vort,
                                                     const double *__restrict__ qvort_e,
                                                     const double *__restrict__ rhsv_e,
                                                     const double *__restrict__ temp, double ren, int buoy,
                                                     double fac, double gx, double gy, double gz,
                                                     double *__restrict__ r) {
  __shared__ double sv[27][8];  // V0 V1 V2 W0 W1 W2 rhsv_e temp
  __shared__ double kval[64];
  __shared__ int kflag[64];
  const int c = blockIdx.x, it = threadIdx.x;
  if (it < 27) {
    const size_t node = idc[c + (size_t)nc * it] - 1;
    for (int d = 0; d < 3; ++d) sv[it][d] = vel[node + (size_t)nn * d], sv[it][3 + d] = vort[node + (size_t)nn * d];
    sv[it][6] = rhsv_e[node];
    sv[it][7] = buoy ? temp[node] : 0.0;
  }
  stage_known(c, 51, colknown, colfield, vort + (size_t)nn * e, qvort_e, kval, kflag);
  __syncthreads();
  if (it >= 51) return;
  const size_t m1 = (size_t)nc * 51, i = (size_t)c * 51 + it;
  double v = known_part(c, it, 51, 51, Af, kval, kflag);
  const double rn = ren / 1.0;
  // one pass over the matrices: the buoyancy terms (BouyancyRHS, added after
  // the convection sum as there) are summed beside the convection terms and
  // subtracted at the end (the reference subtracts them one at a time: a
  // rounding-level difference)
  double bsum = 0.0;
  const double Rro = 1.0e20;
  for (int j = 0; j < 27; ++j) {
    const double V0 = sv[j][0], V1 = sv[j][1], V2 = sv[j][2], W0 = sv[j][3], W1 = sv[j][4], W2 = sv[j][5];
    const double we = sv[j][3 + e], ve = sv[j][e];
    const size_t a = i + m1 * j;
    const double ax = mp.Ex[a], ay = mp.Ey[a], az = mp.Ez[a];
    v = v - mp.B[a] * sv[j][6] * ren / 1.0 -
        rn * (+ax * (V0 * we - ve * W0) + ay * (V1 * we - ve * W1) + az * (V2 * we - ve * W2));
    const double tt = buoy ? sv[j][7] - 0.0 / Rro : 0.0;
    if (e == 0)
      bsum += fac * (+gz * ay * tt - gy * az * tt);
    else if (e == 1)
      bsum += fac * (-gz * ax * tt + gx * az * tt);
    else
      bsum += fac * (+gy * ax * tt - gx * ay * tt);
  }
  if (buoy) v = v - bsum;
  r[i] = v;
}

// SolveFDdae right-hand side (lib_solvFDdae.f:209-254); one 64-thread
// workgroup per cell with the nodal and known values staged in LDS
__global__ void __launch_bounds__(64) t_rhs_kernel(int nc, int nn, MacroP mp, const int *__restrict__ idc,
                                                   const double *__restrict__ Af, const int *__restrict__ colknown,
                                                   const int *__restrict__ colfield, const double *__restrict__ vel,
                                                   const double *__restrict__ temp, const double *__restrict__ qtemp,
                                                   const double *__restrict__ ttime, double enazalfa,
                                                   double *__restrict__ r) {
  __shared__ double sv[27][5];  // V0 V1 V2 temp ttime
  __shared__ double kval[64];
  __shared__ int kflag[64];
  const int c = blockIdx.x, it = threadIdx.x;
  if (it < 27) {
    const size_t node = idc[c + (size_t)nc * it] - 1;
    for (int d = 0; d < 3; ++d) sv[it][d] = vel[node + (size_t)nn * d];
    sv[it][3] = temp[node];
    sv[it][4] = ttime[node];
  }
  stage_known(c, 51, colknown, colfield, temp, qtemp, kval, kflag);
  __syncthreads();
  if (it >= 51) return;
  const size_t m1 = (size_t)nc * 51, i = (size_t)c * 51 + it;
  double v = known_part(c, it, 51, 51, Af, kval, kflag);
  for (int j = 0; j < 27; ++j) {
    const double V0 = sv[j][0], V1 = sv[j][1], V2 = sv[j][2], u = sv[j][3];
    const size_t a = i + m1 * j;
    v = v - mp.B[a] * sv[j][4] * enazalfa - enazalfa * (+mp.Ex[a] * V0 * u + mp.Ey[a] * V1 * u + mp.Ez[a] * V2 * u);
  }
  r[i] = v;
}

// rhsKM3 (lib_km.f:304-389) incl. the rhsm*knv part
// Two cells per 64-thread workgroup (lanes 0-26 and 32-58, one row each); each
// cell's nodal velocities and vorticities and its known columns are staged in
// LDS once, the rows then stream only their matrix entries
__global__ void __launch_bounds__(64) km_rhs_kernel(int nc, int nn, int e, MacroP mp, const int *__restrict__ idc,
                                                    const double *__restrict__ Af, const int *__restrict__ colknown,
                                                    const int *__restrict__ colfield, const double *__restrict__ vel,
                                                    const double *__restrict__ vort, double *__restrict__ r) {
  __shared__ double sv[2][27][6];  // V0 V1 V2 W0 W1 W2
  __shared__ double kval[2][32];
  __shared__ int kflag[2][32];
  const int h = threadIdx.x >> 5, it = threadIdx.x & 31, c = blockIdx.x * 2 + h;
  const bool row = it < 27 && c < nc;
  if (row) {
    const size_t node = idc[c + (size_t)nc * it] - 1;
    for (int d = 0; d < 3; ++d) sv[h][it][d] = vel[node + (size_t)nn * d], sv[h][it][3 + d] = vort[node + (size_t)nn * d];
    const int kn = colknown[(size_t)c * 27 + it];
    kflag[h][it] = kn;
    kval[h][it] = kn ? vel[(size_t)nn * e + colfield[(size_t)c * 27 + it]] : 0.0;
  }
  __syncthreads();
  if (!row) return;
  const size_t m1 = (size_t)nc * 51, m2 = (size_t)nc * 27, ii = (size_t)c * 27 + it, i = (size_t)c * 51 + it;
  double v = known_part(c, it, 27, 27, Af, kval[h], kflag[h]);
  for (int j = 0; j < 27; ++j) {
    const size_t a2 = ii + m2 * j, a1 = i + m1 * j;
    const double V0 = sv[h][j][0], V1 = sv[h][j][1], V2 = sv[h][j][2];
    const double W0 = sv[h][j][3], W1 = sv[h][j][4], W2 = sv[h][j][5];
    if (e == 0)
      v = v - mp.Htz[a2] * V1 + mp.Hty[a2] * V2 + mp.Dz[a1] * W1 - mp.Dy[a1] * W2;
    else if (e == 1)
      v = v + mp.Htz[a2] * V0 - mp.Htx[a2] * V2 - mp.Dz[a1] * W0 + mp.Dx[a1] * W2;
    else
      v = v - mp.Hty[a2] * V0 + mp.Htx[a2] * V1 + mp.Dy[a1] * W0 - mp.Dx[a1] * W1;
  }
  r[ii] = v;
}

__global__ void km_block_kernel(int nc, const double *__restrict__ H, double *__restrict__ Af) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)nc * 27 * 27) return;
  size_t c = k / 729;
  int r = (int)(k / 27 % 27), l = (int)(k % 27);
  Af[k] = H[c * 51 + r + (size_t)nc * 51 * l];
}

// dv/dn = 0 walls: velocity(gbn(i)) = velocity(kobc(i)) for kode 2/3
__global__ void dvdn_kernel(int nb, int nn, int e, const int *__restrict__ kode, const int *__restrict__ gbn,
                            const int *__restrict__ kobc, double *__restrict__ vel) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  int kd = kode[i + (size_t)nb * e];
  if (kd == 2 || kd == 3) vel[gbn[i] - 1 + (size_t)nn * e] = vel[kobc[i] - 1 + (size_t)nn * e];
}

// CheckConvergencevw partial sums: 7 differences and 7 norms per block; pnls <- current
__global__ void __launch_bounds__(256) conv_kernel(int nn, const double *__restrict__ vel, const double *__restrict__ vort,
                                                   const double *__restrict__ temp, double *__restrict__ pv,
                                                   double *__restrict__ pw, double *__restrict__ pt,
                                                   double *__restrict__ part) {
  __shared__ double sh[256];
  int i = blockIdx.x * 256 + threadIdx.x;
  double a[14];
  for (int k = 0; k < 14; ++k) a[k] = 0.0;
  if (i < nn) {
    for (int d = 0; d < 3; ++d) {
      double v = vel[i + (size_t)nn * d], w = vort[i + (size_t)nn * d];
      double dv = v - pv[i + (size_t)nn * d], dwv = w - pw[i + (size_t)nn * d];
      a[d] = dv * dv, a[3 + d] = dwv * dwv;
      a[7 + d] = v * v, a[10 + d] = w * w;
      pv[i + (size_t)nn * d] = v, pw[i + (size_t)nn * d] = w;
    }
    double t = temp[i], dt = t - pt[i];
    a[6] = dt * dt, a[13] = t * t;
    pt[i] = t;
  }
  for (int k = 0; k < 14; ++k) {
    sh[threadIdx.x] = a[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * 14 + k] = sh[0];
    __syncthreads();
  }
}

__global__ void time_rhs_kernel(size_t n, double beta2, double beta3, const double *__restrict__ f,
                                double *__restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = -f[i] * beta2 + f[i] * beta3;
}

}  // namespace

// ---------------------------------------------------------------------------
struct tt_case {
  hipStream_t s = nullptr;
  tt_case_view cv{};
  std::vector<int> kode_h, kobc_h;
  DevMesh mesh;
  QuadDev qh{};
  MacroDev macro;
  DenseKin dense;
  DBuf<double> Af_vte, Af_t, Af_km;
  SubSys vte[3], tsys, km;
  LsqrWork work;
  DBuf<double> vel, vort, qvort, temp, qtemp, pv, pw, pt, rhsv, ttime, b, x, cpart;
  DBuf<int> kode, kobc;
  int lsqr_prec = kBlocks64;  // blocks the least-squares iterations read (BlockPrec)
  float dlse[9];
  LsqrProbe probe51, probe27;  // lsqr_A samples on 51-row (VTE, T) and 27-row (KM) systems
  bool graphs = false;         // LSQR iterations replayed from captured graphs (else stream launches)
  double beta = 0, beta2 = 0, beta3 = 0, ren = 0;
  int it = 0;
  bool finished = false;
  int64_t km_solver_warnings = 0;  // kinematics solves that ended with istop > 3 (the case continues)
  std::vector<double> vort0, temp0;  // fields at the start of the time step (restart pts/ppts fields)
  double t_phase[11] = {0};
  hipEvent_t join = nullptr;  // orders this case's stream against a batch's stream
};

namespace {

void build_subsys(tt_case &c, SubSys &S, int R, int L, const double *Af, const int *wk, const int *wq, int n,
                  bool want32) {
  const auto &v = c.cv;
  const int nc = v.mesh.nicell, nn = v.mesh.nnodes;
  S.R = R, S.L = L, S.n = n, S.ncell = nc, S.Afull = Af;
  std::vector<int> colmap((size_t)nc * L), cf((size_t)nc * L), ck((size_t)nc * L), uf(n, -1);
  for (int ic = 0; ic < nc; ++ic)
    for (int l = 0; l < L; ++l) {
      int code, field;
      if (l < 27) {
        field = v.mesh.idc[ic + (size_t)nc * l] - 1;
        code = wk[field];
      } else {
        int side = (l - 27) / 4, j = (l - 27) % 4;
        field = v.mesh.ibf[ic + (size_t)nc * (side + 6 * j)] - 1;
        code = wq[field];
      }
      size_t k = (size_t)ic * L + l;
      cf[k] = field;
      ck[k] = code > 0;
      colmap[k] = code < 0 ? -code - 1 : -1;
      if (code < 0) {
        S.nnz += R;
        if (-code - 1 >= n) throw std::runtime_error("unknown number out of range");
        uf[-code - 1] = l < 27 ? field : nn + field;
      }
    }
  for (int u = 0; u < n; ++u)
    if (uf[u] < 0) throw std::runtime_error("unknown without a field position");
  S.colfield.upload(cf, c.s);
  S.colknown.upload(ck, c.s);
  S.unk_field.upload(uf, c.s);
  block_setup(S.bs, nc, R, L, n, colmap, Af, want32, c.s);
}

MacroP macro_ptrs(const MacroDev &m) {
  return MacroP{m.B.p, m.Ax.p, m.Ay.p, m.Az.p, m.Dx.p, m.Dy.p, m.Dz.p, m.Htx.p, m.Hty.p, m.Htz.p};
}

struct EvTimer {
  hipEvent_t a, b;
  hipStream_t s;
  explicit EvTimer(hipStream_t st) : s(st) {
    TT_HIP(hipEventCreate(&a));
    TT_HIP(hipEventCreate(&b));
    TT_HIP(hipEventRecord(a, s));
  }
  double stop() {
    TT_HIP(hipEventRecord(b, s));
    TT_HIP(hipEventSynchronize(b));
    float ms = 0;
    TT_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
  }
};

}  // namespace

// ---------------------------------------------------------------------------
namespace {
// initial fields and iteration state (main.f90:1145-1199 for the first time step)
void reset_fields(tt_case &cr, const double *vel, const double *vort, const double *qvort, const double *temp,
                  const double *qtemp) {
  tt_case *c = &cr;
  const auto &m = c->cv.mesh;
  const int nn = m.nnodes, nq = m.nq;
  hipStream_t s = c->s;
  {
    if (!vel || !vort || !qvort || !temp || !qtemp) throw std::runtime_error("NULL field");
    c->vort0.assign(vort, vort + (size_t)nn * 3);
    c->temp0.assign(temp, temp + nn);
    c->vel.upload(vel, (size_t)nn * 3, s);
    c->vort.upload(vort, (size_t)nn * 3, s);
    c->qvort.upload(qvort, (size_t)nq * 3, s);
    c->temp.upload(temp, nn, s);
    c->qtemp.upload(qtemp, nq, s);
    c->pv.upload(vel, (size_t)nn * 3, s);
    c->pw.upload(vort, (size_t)nn * 3, s);
    c->pt.upload(temp, nn, s);
    // time-derivative terms of the first (only) time step
    if (!c->rhsv.p) c->rhsv.alloc((size_t)nn * 3);
    if (!c->ttime.p) c->ttime.alloc(nn);
    hipLaunchKernelGGL(time_rhs_kernel, dim3(ceil_div((size_t)nn * 3, 256)), dim3(256), 0, s, (size_t)nn * 3,
                       c->beta2, c->beta3, c->vort.p, c->rhsv.p);
    hipLaunchKernelGGL(time_rhs_kernel, dim3(ceil_div(nn, 256)), dim3(256), 0, s, (size_t)nn, c->beta2, c->beta3,
                       c->temp.p, c->ttime.p);
    TT_LAUNCH_CHECK();
    for (int k = 0; k < 9; ++k) c->dlse[k] = c->cv.lsqs_eps;
    c->dense.since_anchor = 0;
    c->it = 0;
    c->finished = false;
  }
}
}  // namespace

namespace {
// TT_PREC_FP64: FP64 everywhere.  TT_PREC_FP32 / FP16 / BF16: FP32 least-squares
// blocks, and the dense kinematics products on copies in that format, anchored
// by an FP64 step every anchor_period Picard iterations.
void apply_precision(tt_case &c, int32_t prec) {
  if (prec < TT_PREC_FP64 || prec > TT_PREC_FP16) throw std::runtime_error("unknown precision");
  // FP16 mode: half-precision blocks; FP32 and BF16 modes: single-precision blocks
  const int bp = prec == TT_PREC_FP64 ? kBlocks64 : prec == TT_PREC_FP16 ? kBlocks16 : kBlocks32;
  for (SubSys *ss : {&c.vte[0], &c.vte[1], &c.vte[2], &c.km, &c.tsys}) {
    if (ss == &c.tsys && c.cv.iDT != 1) continue;
    if (bp == kBlocks32) block_make32(ss->bs, c.s);
    if (bp == kBlocks16) block_make16(ss->bs, c.s);
  }
  c.lsqr_prec = bp;
  if (c.cv.iBw == 1) {
    const SdPrec sp = prec == TT_PREC_FP64   ? SdPrec::F64
                      : prec == TT_PREC_FP32 ? SdPrec::F32
                      : prec == TT_PREC_BF16 ? SdPrec::BF16
                                             : SdPrec::F16;
    dense_set_precision(c.dense, sp, c.s);
  }
  TT_HIP(hipStreamSynchronize(c.s));
}
}  // namespace

extern "C" int tt_case_create(const tt_case_view *cv, const double *vel, const double *vort, const double *qvort,
                              const double *temp, const double *qtemp, int32_t sd_prec, double refine_tol,
                              tt_case **out) {
  (void)refine_tol;
  return guarded("tt_case_create", [&] {
    if (!cv || !out || !vel || !vort || !qvort || !temp || !qtemp) throw std::runtime_error("NULL argument");
    if (cv->iDT < 0 || cv->iDT > 1) throw std::runtime_error("iDT must be 0 or 1");
    auto c = std::make_unique<tt_case>();
    c->cv = *cv;
    const auto &m = cv->mesh;
    const int nn = m.nnodes, nb = m.nbnodes, nc = m.nicell, nq = m.nq;
    TT_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    TT_HIP(hipEventCreateWithFlags(&c->join, hipEventDisableTiming));
    hipStream_t s = c->s;
    c->kode_h.assign(cv->kode, cv->kode + (size_t)nb * 3);
    bool need_kobc = false;
    for (int k : c->kode_h) need_kobc |= (k == 2 || k == 3);
    if (need_kobc && !cv->kobc) throw std::runtime_error("kode 2/3 walls need kobc");
    upload_mesh(m, c->mesh, s);
    build_quad(cv->quad, c->qh);
    // ---- integrals
    double ont[11];
    {
      EvTimer t(s);
      macro_integrals_dev(c->mesh, c->qh, cv->copy, c->macro, ont, s);
      // Ax - Dx, Ay - Dy, Az - Dz in place of Ax, Ay, Az (see MacroP)
      const size_t na = (size_t)nc * 51 * 27;
      for (int d = 0; d < 3; ++d)
        hipLaunchKernelGGL(adv_diff_kernel, dim3(ceil_div(na, 256)), dim3(256), 0, s, na,
                           d == 0 ? c->macro.Ax.p : d == 1 ? c->macro.Ay.p : c->macro.Az.p,
                           d == 0 ? c->macro.Dx.p : d == 1 ? c->macro.Dy.p : c->macro.Dz.p);
      TT_LAUNCH_CHECK();
      c->t_phase[7] = t.stop();
    }
    if (cv->iBw != 0 && cv->iBw != 1) throw std::runtime_error("iBw must be 0 or 1");
    if (cv->iBw == 1) {  // single-domain kinematics (main.f90:410-513)
      {
        EvTimer t(s);
        double onts[10];
        sd_integrals_dev(c->mesh, m, c->qh, 0, nb, c->dense.mats, onts, s);
        c->t_phase[8] = t.stop();
      }
      dense_setup(c->dense, c->mesh, c->kode_h.data(), SdPrec::F64, s, &c->t_phase[9]);
    }
    // ---- time scheme (scheme 1) and the subdomain systems
    c->beta = 1.50 / cv->tstep, c->beta2 = 2.00 / cv->tstep, c->beta3 = 0.50 / cv->tstep;
    c->ren = cv->ReN;
    if (cv->iDT > 0) c->ren = 1.0 / cv->PrN;
    EvTimer tsys(s);
    const MacroDev &M = c->macro;
    size_t tot51 = (size_t)nc * 51 * 51;
    c->Af_vte.alloc(tot51);
    hipLaunchKernelGGL(afull_kernel, dim3(ceil_div(tot51, 256)), dim3(256), 0, s, nc, 51, 51, M.H.p, M.G.p, M.B.p,
                       c->beta, c->ren / 1.0, c->Af_vte.p);
    TT_LAUNCH_CHECK();
    for (int e = 0; e < 3; ++e)
      build_subsys(*c, c->vte[e], 51, 51, c->Af_vte.p, cv->wkode + (size_t)nn * e, cv->wqkode + (size_t)nq * e,
                   cv->nunk[e], c->lsqr_prec == kBlocks32);
    if (cv->iDT == 1) {
      c->Af_t.alloc(tot51);
      hipLaunchKernelGGL(afull_kernel, dim3(ceil_div(tot51, 256)), dim3(256), 0, s, nc, 51, 51, M.H.p, M.G.p, M.B.p,
                         c->beta, c->ren * cv->PrN / 1.0, c->Af_t.p);
      TT_LAUNCH_CHECK();
      build_subsys(*c, c->tsys, 51, 51, c->Af_t.p, cv->Tkode, cv->Tqkode, cv->Tnunk, c->lsqr_prec == kBlocks32);
    }
    c->Af_km.alloc((size_t)nc * 729);
    hipLaunchKernelGGL(km_block_kernel, dim3(ceil_div((size_t)nc * 729, 256)), dim3(256), 0, s, nc, M.H.p,
                       c->Af_km.p);
    TT_LAUNCH_CHECK();
    build_subsys(*c, c->km, 27, 27, c->Af_km.p, cv->kmu, nullptr, cv->nsmcol, c->lsqr_prec == kBlocks32);
    c->t_phase[10] = tsys.stop();
    c->b.alloc((size_t)nc * 51);
    int nmax = std::max({cv->nunk[0], cv->nunk[1], cv->nunk[2], cv->Tnunk, cv->nsmcol});
    c->x.alloc(nmax);
    c->cpart.alloc((size_t)ceil_div(nn, 256) * 14);
    c->kode.upload(c->kode_h, s);
    if (cv->kobc) c->kobc.upload(cv->kobc, nb, s);
    reset_fields(*c, vel, vort, qvort, temp, qtemp);
    apply_precision(*c, sd_prec);
    TT_HIP(hipStreamSynchronize(s));
    *out = c.release();
    return 0;
  });
}


namespace {
// outcome of one case's Picard iteration
struct StepOut {
  double err[9] = {0};
  int32_t nit[12] = {0};
  int status = 1;  // 1: another iteration follows, 0: finished, -1: failed (error)
  std::string error;
};

// One Picard iteration (main.f90:1200-1299) of every case in cs, all on the
// stream of cs[0].  The phases run case after case; each linear solve of a
// phase is one block_lsqr per case when cs holds one case (captured chunks
// and sampled timings as configured), and one block_lsqr_multi over the
// cases otherwise, so that a batch of cases shares every LSQR launch.  The
// kernels and their arithmetic are the same either way: a case's iterates do
// not depend on the batch it runs in.  A case whose vorticity or energy solve
// stops with istop > 3 (condition or iteration limit) is failed and leaves the
// remaining phases, as WarnErr level 4 stops the reference
// (lib_solvFDwTE.f:297, lib_solvFDdae.f:265); the kinematics solves only
// count it and keep the updated x (lib_km.f:274).  Each solve runs with the
// solve-history slot of its field component (component e for the kinematics
// system shared by the three components).
void picard_multi(const std::vector<tt_case *> &cs, std::vector<StepOut> &out) {
  const size_t K = cs.size();
  hipStream_t s = cs[0]->s;
  out.assign(K, StepOut());
  std::vector<char> ok(K, 1);
  for (tt_case *c : cs) c->it++;
  double tp[7] = {0};
  EvTimer ttot(s);
  // the same subsystem of every live case that has it: hot start, solve, result
  auto solve = [&](auto sys, auto fnode, auto fflux, auto eps, int slot, bool fatal, int its_k, auto has) {
    std::vector<size_t> who;
    for (size_t i = 0; i < K; ++i)
      if (ok[i] && has(*cs[i])) who.push_back(i);
    if (who.empty()) return;
    for (size_t i : who) {
      tt_case &c = *cs[i];
      SubSys &S = sys(c);
      hipLaunchKernelGGL(gather_kernel, dim3(ceil_div(S.n, 256)), dim3(256), 0, s, S.n, c.cv.mesh.nnodes,
                         S.unk_field.p, fnode(c), fflux(c), c.x.p);
      TT_LAUNCH_CHECK();
    }
    std::vector<LsqrResult> res(who.size());
    if (K == 1) {
      tt_case &c = *cs[0];
      SubSys &S = sys(c);
      res[0] = block_lsqr(S.bs, c.work, c.b.p, c.x.p, c.cv.lsqs_maxit, eps(c), c.lsqr_prec, s,
                          S.R == 51 ? &c.probe51 : &c.probe27, c.graphs, slot);
    } else {
      std::vector<LsqrJob> jobs(who.size());
      for (size_t q = 0; q < who.size(); ++q) {
        tt_case &c = *cs[who[q]];
        LsqrJob &j = jobs[q];
        j.s = &sys(c).bs, j.wk = &c.work, j.b = c.b.p, j.x = c.x.p;
        j.maxit = c.cv.lsqs_maxit, j.eps4 = eps(c), j.slot = slot;
      }
      block_lsqr_multi(jobs, cs[0]->lsqr_prec, s);
      for (size_t q = 0; q < who.size(); ++q) res[q] = jobs[q].res;
    }
    for (size_t q = 0; q < who.size(); ++q) {
      const size_t i = who[q];
      out[i].nit[its_k] = res[q].nit;
      if (res[q].istop > 3) {
        if (fatal) {
          ok[i] = 0;
          out[i].status = -1;
          out[i].error = "LSQR stopped with istop > 3 (ill-conditioned or iteration limit)";
        } else {
          ++cs[i]->km_solver_warnings;
        }
      }
    }
  };
  auto all = [](tt_case &) { return true; };
  // ---- boundary vorticities: the cases' steps are independent; in a batch
  // each runs on its own stream after the batch stream's earlier work, and
  // the batch stream waits for them all
  {
    EvTimer t(s);
    if (K > 1) {
      TT_HIP(hipEventRecord(cs[0]->join, s));
      for (size_t i = 1; i < K; ++i) TT_HIP(hipStreamWaitEvent(cs[i]->s, cs[0]->join, 0));
    }
    for (size_t i = 0; i < K; ++i) {
      tt_case &c = *cs[i];
      if (c.cv.iBw == 1) dense_bw_step(c.dense, c.mesh, c.vel.p, c.vort.p, c.cv.urBw, i == 0 ? s : c.s);
    }
    for (size_t i = 1; i < K; ++i) {
      TT_HIP(hipEventRecord(cs[i]->join, cs[i]->s));
      TT_HIP(hipStreamWaitEvent(s, cs[i]->join, 0));
    }
    tp[0] = t.stop();
  }
  // ---- domain velocities
  {
    EvTimer t(s);
    for (int e = 0; e < 3; ++e) {
      for (size_t i = 0; i < K; ++i) {
        tt_case &c = *cs[i];
        const int nc = c.cv.mesh.nicell, nn = c.cv.mesh.nnodes;
        hipLaunchKernelGGL(km_rhs_kernel, dim3(ceil_div((size_t)nc, 2)), dim3(64), 0, s, nc, nn, e,
                           macro_ptrs(c.macro), c.mesh.idc.p, c.Af_km.p, c.km.colknown.p, c.km.colfield.p, c.vel.p,
                           c.vort.p, c.b.p);
        TT_LAUNCH_CHECK();
      }
      solve([](tt_case &c) -> SubSys & { return c.km; },
            [e](tt_case &c) { return (const double *)(c.vel.p + (size_t)c.cv.mesh.nnodes * e); },
            [](tt_case &) { return (const double *)nullptr; }, [e](tt_case &c) { return c.dlse[e]; }, e, false,
            3 + e, all);
      for (size_t i = 0; i < K; ++i) {
        tt_case &c = *cs[i];
        const int nn = c.cv.mesh.nnodes, nb = c.cv.mesh.nbnodes;
        hipLaunchKernelGGL(scatter_kernel, dim3(ceil_div(c.km.n, 256)), dim3(256), 0, s, c.km.n, nn, c.cv.urDv[e],
                           c.km.unk_field.p, c.x.p, c.vel.p + (size_t)nn * e, nullptr);
        if (c.kobc.p)
          hipLaunchKernelGGL(dvdn_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, nb, nn, e, c.kode.p,
                             c.mesh.gbn.p, c.kobc.p, c.vel.p);
        TT_LAUNCH_CHECK();
      }
    }
    tp[1] = t.stop();
  }
  // ---- temperature
  auto has_t = [](tt_case &c) { return c.cv.iDT == 1; };
  bool any_t = false;
  for (size_t i = 0; i < K; ++i) any_t |= ok[i] && cs[i]->cv.iDT == 1;
  if (any_t) {
    EvTimer t(s);
    for (size_t i = 0; i < K; ++i) {
      tt_case &c = *cs[i];
      if (!ok[i] || c.cv.iDT != 1) continue;
      const int nc = c.cv.mesh.nicell, nn = c.cv.mesh.nnodes;
      const double enazalfa = c.ren * c.cv.PrN / 1.0;
      hipLaunchKernelGGL(t_rhs_kernel, dim3(nc), dim3(64), 0, s, nc, nn, macro_ptrs(c.macro), c.mesh.idc.p,
                         c.Af_t.p, c.tsys.colknown.p, c.tsys.colfield.p, c.vel.p, c.temp.p, c.qtemp.p, c.ttime.p,
                         enazalfa, c.b.p);
      TT_LAUNCH_CHECK();
    }
    solve([](tt_case &c) -> SubSys & { return c.tsys; }, [](tt_case &c) { return (const double *)c.temp.p; },
          [](tt_case &c) { return (const double *)c.qtemp.p; }, [](tt_case &c) { return c.dlse[6]; }, 0, true, 9,
          has_t);
    for (size_t i = 0; i < K; ++i) {
      tt_case &c = *cs[i];
      if (!ok[i] || c.cv.iDT != 1) continue;
      hipLaunchKernelGGL(scatter_kernel, dim3(ceil_div(c.tsys.n, 256)), dim3(256), 0, s, c.tsys.n, c.cv.mesh.nnodes,
                         c.cv.urDT, c.tsys.unk_field.p, c.x.p, c.temp.p, c.qtemp.p);
      TT_LAUNCH_CHECK();
    }
    tp[2] = t.stop();
  }
  // ---- domain vorticities
  {
    EvTimer t(s);
    for (int e = 0; e < 3; ++e) {
      for (size_t i = 0; i < K; ++i) {
        tt_case &c = *cs[i];
        if (!ok[i]) continue;
        const auto &cv = c.cv;
        const int nc = cv.mesh.nicell, nn = cv.mesh.nnodes, nq = cv.mesh.nq;
        const double fac = cv.RaN / cv.PrN / c.ren * 1.0 / 1.0;
        hipLaunchKernelGGL(vte_rhs_kernel, dim3(nc), dim3(64), 0, s, nc, nn, e, macro_ptrs(c.macro), c.mesh.idc.p,
                           c.Af_vte.p, c.vte[e].colknown.p, c.vte[e].colfield.p, c.vel.p, c.vort.p,
                           c.qvort.p + (size_t)nq * e, c.rhsv.p + (size_t)nn * e, c.temp.p, c.ren,
                           cv.iDT > 0 ? 1 : 0, fac, cv.gx, cv.gy, cv.gz, c.b.p);
        TT_LAUNCH_CHECK();
      }
      solve([e](tt_case &c) -> SubSys & { return c.vte[e]; },
            [e](tt_case &c) { return (const double *)(c.vort.p + (size_t)c.cv.mesh.nnodes * e); },
            [e](tt_case &c) { return (const double *)(c.qvort.p + (size_t)c.cv.mesh.nq * e); },
            [e](tt_case &c) { return c.dlse[3 + e]; }, 0, true, 6 + e, all);
      for (size_t i = 0; i < K; ++i) {
        tt_case &c = *cs[i];
        if (!ok[i]) continue;
        const int nn = c.cv.mesh.nnodes, nq = c.cv.mesh.nq;
        hipLaunchKernelGGL(scatter_kernel, dim3(ceil_div(c.vte[e].n, 256)), dim3(256), 0, s, c.vte[e].n, nn,
                           c.cv.urDw[e], c.vte[e].unk_field.p, c.x.p, c.vort.p + (size_t)nn * e,
                           c.qvort.p + (size_t)nq * e);
        TT_LAUNCH_CHECK();
      }
    }
    tp[3] = t.stop();
  }
  // ---- relative changes (CheckConvergencevw)
  EvTimer tconv(s);
  std::vector<std::vector<double>> parts(K);
  for (size_t i = 0; i < K; ++i) {
    tt_case &c = *cs[i];
    if (!ok[i]) continue;
    const int nn = c.cv.mesh.nnodes;
    const unsigned nbk = ceil_div(nn, 256);
    hipLaunchKernelGGL(conv_kernel, dim3(nbk), dim3(256), 0, s, nn, c.vel.p, c.vort.p, c.temp.p, c.pv.p, c.pw.p,
                       c.pt.p, c.cpart.p);
    TT_LAUNCH_CHECK();
    parts[i].resize((size_t)nbk * 14);
    TT_HIP(hipMemcpyAsync(parts[i].data(), c.cpart.p, sizeof(double) * parts[i].size(), hipMemcpyDeviceToHost, s));
  }
  TT_HIP(hipStreamSynchronize(s));
  tp[4] = tconv.stop();
  tp[6] = ttot.stop();
  for (size_t i = 0; i < K; ++i) {
    tt_case &c = *cs[i];
    if (!ok[i]) continue;
    const auto &cv = c.cv;
    const unsigned nbk = ceil_div(cv.mesh.nnodes, 256);
    const std::vector<double> &part = parts[i];
    double rnd[9] = {0}, rno[9] = {0};
    for (unsigned k = 0; k < nbk; ++k) {
      for (int j = 0; j < 7; ++j) rnd[j] += part[(size_t)k * 14 + j], rno[j] += part[(size_t)k * 14 + 7 + j];
    }
    double e9[9];
    for (int j = 0; j < 9; ++j) e9[j] = 9.999999999;
    double mx = std::max({rno[0], rno[1], rno[2]});
    if (mx > 1.0e-15)
      for (int j = 0; j < 3; ++j) e9[j] = std::sqrt(rnd[j] / mx);
    mx = std::max({rno[3], rno[4], rno[5]});
    if (mx > 1.0e-15)
      for (int j = 3; j < 6; ++j) e9[j] = std::sqrt(rnd[j] / mx);
    for (int j = 6; j < 9; ++j)
      if (rno[j] > 1.0e-15) e9[j] = std::sqrt(rnd[j] / rno[j]);
    for (int j = 0; j < 9; ++j) {
      if (rnd[j] == 0 && rno[j] == 0.0) e9[j] = 0.0;
      if (e9[j] > 9.999999999) e9[j] = 9.999999999;
    }
    // ---- DLSE
    if (cv.iDLSE == 1 && c.it > 1) {
      if (c.it == 2)
        for (int k = 0; k < 9; ++k) c.dlse[k] = cv.dlse_epsth;
      for (int k = 0; k < 8; ++k) {
        float old = c.dlse[k];
        if (e9[k] / (double)cv.dlse_r < (double)cv.dlse_epsth && e9[k] != 0.0)
          c.dlse[k] = (float)(e9[k] / (double)cv.dlse_r);
        else
          c.dlse[k] = cv.dlse_epsth;
        if (old < c.dlse[k]) c.dlse[k] = old;
        if (c.dlse[k] < cv.lsqs_eps) c.dlse[k] = cv.lsqs_eps;
      }
    }
    std::memcpy(out[i].err, e9, sizeof(e9));
    for (int k = 0; k < 7; ++k) c.t_phase[k] = tp[k];
    bool more = false, finite = true;
    for (int j = 0; j < 9; ++j) {
      finite &= std::isfinite(e9[j]);
      more |= e9[j] > cv.eps;
    }
    if (!finite) {
      c.finished = true;  // the reference would stop here as if converged (NaN > eps is false)
      out[i].status = -1;
      out[i].error = "Picard iteration diverged (non-finite relative change)";
      continue;
    }
    if (!(c.it < cv.nnlit && more)) {
      c.finished = true;
      out[i].status = 0;
    }
  }
}
}  // namespace

extern "C" int tt_case_picard_step(tt_case *c, double err[9], int32_t nit[12]) {
  return guarded("tt_case_picard_step", [&]() -> int {
    if (!c) throw std::runtime_error("NULL case");
    if (c->finished) return 0;
    std::vector<StepOut> out;
    picard_multi({c}, out);
    if (err) std::memcpy(err, out[0].err, sizeof(out[0].err));
    if (nit) std::memcpy(nit, out[0].nit, sizeof(out[0].nit));
    if (out[0].status < 0) throw std::runtime_error(out[0].error);
    return out[0].status;
  });
}

extern "C" int tt_case_picard_step_batch(tt_case **cases, int32_t ncase, double *err, int32_t *nit,
                                         int32_t *status) {
  return guarded("tt_case_picard_step_batch", [&]() -> int {
    if (!cases || ncase < 1 || ncase > 16) throw std::runtime_error("1..16 cases per batch");
    std::vector<tt_case *> live;
    std::vector<int> pos;
    for (int i = 0; i < ncase; ++i) {
      tt_case *c = cases[i];
      if (!c) throw std::runtime_error("NULL case");
      for (int j = 0; j < i; ++j)
        if (cases[j] == c) throw std::runtime_error("a case appears twice in the batch");
      if (status) status[i] = 0;
      if (err) std::memset(err + (size_t)9 * i, 0, sizeof(double) * 9);
      if (nit) std::memset(nit + (size_t)12 * i, 0, sizeof(int32_t) * 12);
      if (c->finished) continue;
      if (!live.empty() && c->lsqr_prec != live[0]->lsqr_prec)
        throw std::runtime_error("the cases of a batch must iterate with the same block precision");
      live.push_back(c), pos.push_back(i);
    }
    if (live.empty()) return 0;
    // the batch runs on the first live case's stream: earlier work of the
    // others on their own streams completes first
    for (size_t i = 1; i < live.size(); ++i) TT_HIP(hipStreamSynchronize(live[i]->s));
    std::vector<StepOut> out;
    picard_multi(live, out);
    int failed = 0;
    for (size_t q = 0; q < live.size(); ++q) {
      const int i = pos[q];
      if (err) std::memcpy(err + (size_t)9 * i, out[q].err, sizeof(out[q].err));
      if (nit) std::memcpy(nit + (size_t)12 * i, out[q].nit, sizeof(out[q].nit));
      if (status) status[i] = out[q].status;
      if (out[q].status < 0 && !failed++) tt_set_error("tt_case_picard_step_batch: case %d: %s", i, out[q].error.c_str());
    }
    return 0;
  });
}

extern "C" int tt_case_run(tt_case *c, int32_t max_iter, double *sta, int32_t *ite) {
  int n = 0;
  for (; n < max_iter;) {
    double e[9];
    int32_t it[12];
    int r = tt_case_picard_step(c, e, it);
    if (r < 0) return r;
    if (sta) std::memcpy(sta + (size_t)n * 9, e, sizeof(e));
    if (ite) std::memcpy(ite + (size_t)n * 12, it, sizeof(it));
    ++n;
    if (r == 0) break;
  }
  return n;
}

extern "C" int tt_case_fields(tt_case *c, double *vel, double *vort, double *qvort, double *temp, double *qtemp) {
  return guarded("tt_case_fields", [&] {
    const int nn = c->cv.mesh.nnodes, nq = c->cv.mesh.nq;
    if (vel) c->vel.download(vel, (size_t)nn * 3, c->s);
    if (vort) c->vort.download(vort, (size_t)nn * 3, c->s);
    if (qvort) c->qvort.download(qvort, (size_t)nq * 3, c->s);
    if (temp) c->temp.download(temp, nn, c->s);
    if (qtemp) c->qtemp.download(qtemp, nq, c->s);
    TT_HIP(hipStreamSynchronize(c->s));
    return 0;
  });
}

extern "C" int tt_case_write_restart(tt_case *c, const char *path, int32_t tstep, int32_t glnnlit, double rtime) {
  return guarded("tt_case_write_restart", [&] {
    if (!c || !path) throw std::runtime_error("NULL argument");
    const size_t nn = c->cv.mesh.nnodes, nq = c->cv.mesh.nq;
    std::vector<double> vel(nn * 3), pvel(nn * 3), vort(nn * 3), pvort(nn * 3), qvort(nq * 3), temp(nn), ptemp(nn),
        qtemp(nq), zn(nn, 0.0), zq(nq, 0.0);
    c->vel.download(vel.data(), nn * 3, c->s), c->pv.download(pvel.data(), nn * 3, c->s);
    c->vort.download(vort.data(), nn * 3, c->s), c->pw.download(pvort.data(), nn * 3, c->s);
    c->qvort.download(qvort.data(), nq * 3, c->s);
    c->temp.download(temp.data(), nn, c->s), c->pt.download(ptemp.data(), nn, c->s);
    c->qtemp.download(qtemp.data(), nq, c->s);
    TT_HIP(hipStreamSynchronize(c->s));
    const double *f[17] = {vel.data(),     pvel.data(),     vort.data(),     c->vort0.data(), c->vort0.data(),
                           pvort.data(),   qvort.data(),    temp.data(),     c->temp0.data(), c->temp0.data(),
                           ptemp.data(),   qtemp.data(),    zn.data(),       zn.data(),       zn.data(),
                           zn.data(),      zq.data()};
    if (tt_write_restart(path, (int32_t)nn, (int32_t)nq, tstep, glnnlit, rtime, f) < 0)
      throw std::runtime_error(tt_last_error());
    return 0;
  });
}

extern "C" int tt_case_timings(tt_case *c, double t[11]) {
  std::memcpy(t, c->t_phase, sizeof(c->t_phase));
  return 0;
}

extern "C" int64_t tt_case_device_bytes(tt_case *c) {
  int64_t b = 0;
  auto add = [&](auto &buf) { b += (int64_t)buf.bytes(); };
  const MacroDev &M = c->macro;
  for (auto *p : {&M.H, &M.G, &M.B, &M.Ax, &M.Ay, &M.Az, &M.Htx, &M.Hty, &M.Htz, &M.Dx, &M.Dy, &M.Dz}) add(*p);
  const SdDev &S = c->dense.mats;
  for (auto *p : {&S.H, &S.Htx, &S.Hty, &S.Htz, &S.Dx, &S.Dy, &S.Dz}) add(*p);
  for (auto &p : c->dense.Ainv) add(p);
  for (auto &p : c->dense.lo16) add(p);
  for (auto &p : c->dense.lo32) add(p);
  add(c->Af_vte), add(c->Af_t), add(c->Af_km);
  for (SubSys *ss : {&c->vte[0], &c->vte[1], &c->vte[2], &c->tsys, &c->km}) add(ss->bs.A), add(ss->bs.A32), add(ss->bs.A16);
  return b;
}

extern "C" void tt_case_free(tt_case *c) {
  if (!c) return;
  hipStream_t s = c->s;
  // the case's last launches (graph replays included) finish before their
  // buffers and executable graphs are released
  if (s) (void)hipStreamSynchronize(s);
  hipEvent_t j = c->join;
  delete c;
  if (j) (void)hipEventDestroy(j);
  if (s) (void)hipStreamDestroy(s);
}

extern "C" int tt_device_info(char *name, int32_t name_len, int32_t *num_cu, int64_t *hbm_bytes) {
  return guarded("tt_device_info", [&] {
    int dev = 0;
    TT_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    TT_HIP(hipGetDeviceProperties(&p, dev));
    if (name && name_len > 0) {
      std::snprintf(name, name_len, "%s%s%s", p.name, p.name[0] ? " " : "", p.gcnArchName);
    }
    if (num_cu) *num_cu = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (int64_t)p.totalGlobalMem;
    return 0;
  });
}

extern "C" int tt_case_reset(tt_case *c, const double *vel, const double *vort, const double *qvort,
                             const double *temp, const double *qtemp) {
  return guarded("tt_case_reset", [&] {
    if (!c) throw std::runtime_error("NULL case");
    reset_fields(*c, vel, vort, qvort, temp, qtemp);
    TT_HIP(hipStreamSynchronize(c->s));
    return 0;
  });
}


extern "C" int tt_case_set_precision(tt_case *c, int32_t prec) {
  return guarded("tt_case_set_precision", [&] {
    if (!c) throw std::runtime_error("NULL case");
    apply_precision(*c, prec);
    return 0;
  });
}

extern "C" int tt_case_set_anchor_period(tt_case *c, int32_t period) {
  return guarded("tt_case_set_anchor_period", [&] {
    if (!c) throw std::runtime_error("NULL case");
    if (period < 1) throw std::runtime_error("the anchor period must be at least 1");
    c->dense.anchor_period = period;
    c->dense.since_anchor = 0;
    return 0;
  });
}

extern "C" int tt_case_probe(tt_case *c, int32_t stride) {
  return guarded("tt_case_probe", [&] {
    if (!c) throw std::runtime_error("NULL case");
    for (LsqrProbe *p : {&c->probe51, &c->probe27}) {
      p->stride = stride > 0 ? stride : 0;
      p->seen = p->sampled = 0;
      p->total_ms = 0.0;
      p->used = 0;
    }
    return 0;
  });
}

extern "C" int tt_case_probe_stats(tt_case *c, int32_t rows, double *avg_ms, int64_t *sampled, int64_t *launches) {
  return guarded("tt_case_probe_stats", [&] {
    if (!c) throw std::runtime_error("NULL case");
    if (rows != 51 && rows != 27) throw std::runtime_error("rows must be 51 (VTE, T) or 27 (KM)");
    const LsqrProbe &p = rows == 51 ? c->probe51 : c->probe27;
    if (avg_ms) *avg_ms = p.sampled ? p.total_ms / (double)p.sampled : 0.0;
    if (sampled) *sampled = (int64_t)p.sampled;
    if (launches) *launches = (int64_t)p.seen;
    return 0;
  });
}

extern "C" int tt_case_set_grid(tt_case *c, int32_t grid) {
  return guarded("tt_case_set_grid", [&] {
    if (!c || grid < 0) throw std::runtime_error("grid must be >= 0");
    for (SubSys *ss : {&c->vte[0], &c->vte[1], &c->vte[2], &c->tsys, &c->km})
      for (int p = 0; p < 3; ++p) ss->bs.grid_a[p] = grid;  // 0: tuned again on next use
    return 0;
  });
}

extern "C" int tt_case_set_unknown_threads(tt_case *c, int32_t threads) {
  return guarded("tt_case_set_unknown_threads", [&] {
    if (!c || (threads != 0 && threads != 256 && threads != 512 && threads != 1024))
      throw std::runtime_error("threads must be 0 (tuned), 256, 512 or 1024");
    for (SubSys *ss : {&c->vte[0], &c->vte[1], &c->vte[2], &c->tsys, &c->km})
      for (int p = 0; p < 3; ++p) ss->bs.nb_b[p] = threads;
    return 0;
  });
}

extern "C" int tt_case_set_graphs(tt_case *c, int32_t on) {
  return guarded("tt_case_set_graphs", [&] {
    if (!c || on < 0 || on > 1) throw std::runtime_error("on must be 0 or 1");
    c->graphs = on == 1;
    return 0;
  });
}

extern "C" int tt_case_system_info(tt_case *c, int32_t sys, int64_t info[9]) {
  return guarded("tt_case_system_info", [&] {
    if (!c || !info) throw std::runtime_error("NULL argument");
    if (sys < 0 || sys > 4) throw std::runtime_error("sys must be 0..4 (VTE x, y, z, T, KM)");
    if (sys == 3 && c->cv.iDT != 1) throw std::runtime_error("the case has no energy equation");
    const SubSys &S = sys < 3 ? c->vte[sys] : sys == 3 ? c->tsys : c->km;
    info[0] = (int64_t)S.ncell * S.R, info[1] = S.n, info[2] = S.nnz, info[3] = S.ncell, info[4] = S.R, info[5] = S.L;
    for (int p = 0; p < 3; ++p) info[6 + p] = S.bs.grid_a[p];
    return 0;
  });
}

extern "C" int tt_case_system_tuning(tt_case *c, int32_t sys, int32_t grid[3], int32_t threads[3]) {
  return guarded("tt_case_system_tuning", [&] {
    if (!c || !grid || !threads) throw std::runtime_error("NULL argument");
    if (sys < 0 || sys > 4) throw std::runtime_error("sys must be 0..4 (VTE x, y, z, T, KM)");
    if (sys == 3 && c->cv.iDT != 1) throw std::runtime_error("the case has no energy equation");
    const SubSys &S = sys < 3 ? c->vte[sys] : sys == 3 ? c->tsys : c->km;
    for (int p = 0; p < 3; ++p) grid[p] = S.bs.grid_a[p], threads[p] = S.bs.nb_b[p];
    return 0;
  });
}
