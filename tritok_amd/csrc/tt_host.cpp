// tt_host.cpp -- host-side preprocessing: the part of the reference's Fortran
// host that turns tri.inp / .geo / .bic into the arrays the hot path consumes.
//
//   ReadMeshGid3D       lib_mesh.f:279-504   (.geo reader, gbn/lbn first-appearance order)
//   CheckLinearGeometry lib_mesh.f:135-202   (re-place non-corner nodes trilinearly)
//   SourcePointNormals  lib_mesh.f:658-761
//   GenMakroMesh        lib_mesh.f:865-982   (flux nodes xq, ibf, ibcf, flpm)
//   ReadBic             lib_bic.f:50-502     (walls, codes, unknown numbering)
//   DetKMunk3           lib_km.f:676-727
//   FillKobc            lib_mesh.f:230-276
//   ReadInputFile       lib_IO.f:2-359
//
// The reference's O(n^2) searches (first-appearance boundary list, flux node
// deduplication ImamVsez, FindAdjCell, unknown back-indices) are replaced by
// hash lookups that produce identical numbering.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "tritok_bem.h"
#include "tt_internal.h"

struct tt_host {
  int nn = 0, nb = 0, ne = 0, nc = 0, nq = 0;
  std::vector<double> x;            // (nn,3)
  std::vector<int> ibc, idc, bewn;  // (ne,9), (nc,27), (ne)
  std::vector<int> gbn, lbn;
  std::vector<double> nx, ny, nz;
  std::vector<int> ibf, flpm, ibcf;  // (nc,6,4), (nc,6), (ne,4)
  std::vector<double> xq;            // (nq,3)
  // boundary conditions
  bool have_bc = false;
  std::vector<int> iwn, kode, wkode, wqkode, Tkode, Tqkode, kmu, kobc;
  int nunk[3] = {0, 0, 0}, nbk[3] = {0, 0, 0}, Tnunk = 0, Tnb = 0, nsmcol = 0, nrhscol = 0;
  std::vector<double> vel, vort, qvort, temp, qtemp;
  // modified-Helmholtz codes of the single-domain form: nodes (nn) and the
  // boundary flux points element-major (4 ne); values prescribed by MHEL
  std::vector<int> mHkode, mHqfmkode, ibcf_em;
  std::vector<double> mHu, fmqmH;
  int mHnb = 0, mHfmnunk = 0;
  // tri.inp
  tt_quad quad{8, 8, 4, 4, 1};
  int copy = 0, iBw = 1, iDv = 1, iDw = 1, iDT = 0, iDC = 0, imH = 0;
  double ReN = 1.0, RaN = 1.0, PrN = 1.0, gx = 0.0, gy = 0.0, gz = -1.0, tstep = 1.0e20;
  double urBw[3] = {0.1, 0.1, 0.1}, urDv[3] = {0.1, 0.1, 0.1}, urDw[3] = {0.1, 0.1, 0.1}, urDT = 0.1;
  double eps = 1.0e-7;
  int nnlit = 1000, lsqs_maxit = 5000, iDLSE = 0;
  float lsqs_eps = 1.0e-6f, dlse_r = 100.0f, dlse_epsth = 1.0e-3f;
};

namespace {

// local node positions of the 27-node cell in [0,1]^3 (SetLC, lib_interpol.f:2-100)
const double kLc1[27] = {0, 0, 1, 1, 0, 0, 1, 1, 0, .5, 1, .5, 0, 0, 1, 1, 0, .5, 1, .5, .5, 0, .5, 1, .5, .5, .5};
const double kLc2[27] = {1, 1, 1, 1, 0, 0, 0, 0, 1, 1, 1, 1, .5, .5, .5, .5, 0, 0, 0, 0, 1, .5, .5, .5, .5, 0, .5};
const double kLc3[27] = {1, 0, 0, 1, 1, 0, 0, 1, .5, 0, .5, 1, 1, 0, 0, 1, .5, 0, .5, 1, .5, .5, 0, .5, 1, .5, .5};
// geometry corners in the reference's order idc(6,7,2,3,5,8,1,4)
const int kGeo8[8] = {6, 7, 2, 3, 5, 8, 1, 4};
// cell faces for flux nodes (AddSide calls, lib_mesh.f:889-895): corner local ids
const int kSide4[6][4] = {{6, 2, 3, 7}, {6, 7, 8, 5}, {7, 3, 4, 8}, {3, 2, 1, 4}, {2, 6, 5, 1}, {5, 8, 4, 1}};

bool data_line(std::istream &in, std::string &line) {
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    bool blank = line.find_first_not_of(" \t") == std::string::npos;
    if (blank || line.find('#') != std::string::npos) continue;
    return true;
  }
  return false;
}

// zv4nez4 (lib_interpol.f:704-754): flux point j near corner j of a face
void flux_points(const double c[4][3], double q[4][3]) {
  const double a = 49.0 / 64.0, b = 7.0 / 64.0, cc = 1.0 / 64.0;
  const double m[4][4] = {{a, b, cc, b}, {b, a, b, cc}, {cc, b, a, b}, {b, cc, b, a}};
  for (int r = 0; r < 4; ++r)
    for (int d = 0; d < 3; ++d) q[r][d] = m[r][0] * c[0][d] + m[r][1] * c[1][d] + m[r][2] * c[2][d] + m[r][3] * c[3][d];
}

// Point set with the reference's 1e-10 coordinate tolerance (ImamVsez,
// lib_mesh.f:1095-1117); returns the first-inserted matching index.
struct PointIndex {
  double cell = 1e-6;
  std::unordered_map<long long, std::vector<int>> buckets;
  std::vector<double> *pts;
  static long long key(long long i, long long j, long long k) { return (i * 73856093LL) ^ (j * 19349663LL) ^ (k * 83492791LL); }
  int find(const double p[3]) const {
    long long bi = (long long)std::floor(p[0] / cell), bj = (long long)std::floor(p[1] / cell),
              bk = (long long)std::floor(p[2] / cell);
    int best = -1;
    for (int di = -1; di <= 1; ++di)
      for (int dj = -1; dj <= 1; ++dj)
        for (int dk = -1; dk <= 1; ++dk) {
          auto it = buckets.find(key(bi + di, bj + dj, bk + dk));
          if (it == buckets.end()) continue;
          for (int idx : it->second) {
            const double *q = &(*pts)[3 * idx];
            if (std::fabs(q[0] - p[0]) < 1e-10 && std::fabs(q[1] - p[1]) < 1e-10 && std::fabs(q[2] - p[2]) < 1e-10)
              if (best < 0 || idx < best) best = idx;
          }
        }
    return best;
  }
  int insert(const double p[3]) {
    int idx = (int)(pts->size() / 3);
    pts->insert(pts->end(), p, p + 3);
    long long bi = (long long)std::floor(p[0] / cell), bj = (long long)std::floor(p[1] / cell),
              bk = (long long)std::floor(p[2] / cell);
    buckets[key(bi, bj, bk)].push_back(idx);
    return idx;
  }
};

void finish_mesh(tt_host &h) {
  const int nn = h.nn, ne = h.ne, nc = h.nc;
  auto X = [&](int node, int d) -> double & { return h.x[(node - 1) + (size_t)nn * d]; };
  auto IDC = [&](int ic, int k) { return h.idc[ic + (size_t)nc * (k - 1)]; };
  auto IBC = [&](int e, int k) { return h.ibc[e + (size_t)ne * (k - 1)]; };
  // boundary node list in first-appearance order over elements
  h.gbn.clear();
  h.lbn.assign(nn, 0);
  for (int e = 0; e < ne; ++e)
    for (int k = 1; k <= 9; ++k) {
      int dn = IBC(e, k);
      if (h.lbn[dn - 1] == 0) {
        h.gbn.push_back(dn);
        h.lbn[dn - 1] = (int)h.gbn.size();
      }
    }
  h.nb = (int)h.gbn.size();
  // CheckLinearGeometry: nodes 9..27 of every cell to trilinear positions
  for (int ic = 0; ic < nc; ++ic) {
    double cx[8][3];
    for (int i = 0; i < 8; ++i)
      for (int d = 0; d < 3; ++d) cx[i][d] = X(IDC(ic, kGeo8[i]), d);
    for (int i = 9; i <= 27; ++i) {
      double xi = kLc1[i - 1] * 2.0 - 1.0, eta = kLc2[i - 1] * 2.0 - 1.0, ze = kLc3[i - 1] * 2.0 - 1.0;
      double e1m = 1.0 - xi, e1p = 1.0 + xi, e2m = 1.0 - eta, e2p = 1.0 + eta, e3m = 1.0 - ze, e3p = 1.0 + ze;
      double fi[8] = {0.125 * e1m * e2m * e3m, 0.125 * e1p * e2m * e3m, 0.125 * e1m * e2p * e3m,
                      0.125 * e1p * e2p * e3m, 0.125 * e1m * e2m * e3p, 0.125 * e1p * e2m * e3p,
                      0.125 * e1m * e2p * e3p, 0.125 * e1p * e2p * e3p};
      double p[3] = {0, 0, 0};
      for (int j = 0; j < 8; ++j)
        for (int d = 0; d < 3; ++d) p[d] = p[d] + fi[j] * cx[j][d];
      for (int d = 0; d < 3; ++d) X(IDC(ic, i), d) = p[d];
    }
  }
  // SourcePointNormals: element normals at the face centre, summed per node
  // in element order, then normalised
  h.nx.assign(h.nb, 0.0);
  h.ny.assign(h.nb, 0.0);
  h.nz.assign(h.nb, 0.0);
  for (int e = 0; e < ne; ++e) {
    double c[4][3];
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) c[k][d] = X(IBC(e, 2 * k + 1), d);
    const double e1m = 1.0, e1p = 1.0, e2m = 1.0, e2p = 1.0;
    double xks = 0.25 * (-e2m * c[0][0] + e2m * c[1][0] + e2p * c[2][0] - e2p * c[3][0]);
    double yks = 0.25 * (-e2m * c[0][1] + e2m * c[1][1] + e2p * c[2][1] - e2p * c[3][1]);
    double zks = 0.25 * (-e2m * c[0][2] + e2m * c[1][2] + e2p * c[2][2] - e2p * c[3][2]);
    double xet = 0.25 * (-e1m * c[0][0] - e1p * c[1][0] + e1p * c[2][0] + e1m * c[3][0]);
    double yet = 0.25 * (-e1m * c[0][1] - e1p * c[1][1] + e1p * c[2][1] + e1m * c[3][1]);
    double zet = 0.25 * (-e1m * c[0][2] - e1p * c[1][2] + e1p * c[2][2] + e1m * c[3][2]);
    double anx = yks * zet - yet * zks, any = xet * zks - xks * zet, anz = xks * yet - xet * yks;
    double aj = 1.0 / std::sqrt(anx * anx + any * any + anz * anz);
    anx *= aj;
    any *= aj;
    anz *= aj;
    for (int k = 1; k <= 9; ++k) {
      int i = h.lbn[IBC(e, k) - 1] - 1;
      h.nx[i] += anx;
      h.ny[i] += any;
      h.nz[i] += anz;
    }
  }
  for (int i = 0; i < h.nb; ++i) {
    double aj = 1.0 / std::sqrt(h.nx[i] * h.nx[i] + h.ny[i] * h.ny[i] + h.nz[i] * h.nz[i]);
    h.nx[i] *= aj;
    h.ny[i] *= aj;
    h.nz[i] *= aj;
  }
  // GenMakroMesh: flux nodes of every cell face, deduplicated
  h.xq.clear();
  h.ibf.assign((size_t)nc * 24, 0);
  PointIndex pix;
  std::vector<double> qpts;
  pix.pts = &qpts;
  for (int ic = 0; ic < nc; ++ic)
    for (int s = 0; s < 6; ++s) {
      double c[4][3], q[4][3];
      for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) c[k][d] = X(IDC(ic, kSide4[s][k]), d);
      flux_points(c, q);
      for (int j = 0; j < 4; ++j) {
        int f = pix.find(q[j]);
        if (f < 0) f = pix.insert(q[j]);
        h.ibf[ic + (size_t)nc * (s + 6 * j)] = f + 1;
      }
    }
  h.nq = (int)(qpts.size() / 3);
  h.xq.assign((size_t)h.nq * 3, 0.0);
  for (int i = 0; i < h.nq; ++i)
    for (int d = 0; d < 3; ++d) h.xq[i + (size_t)h.nq * d] = qpts[3 * i + d];
  // ibcf (AddSideBound)
  h.ibcf.assign((size_t)ne * 4, 0);
  for (int e = 0; e < ne; ++e) {
    double c[4][3], q[4][3];
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) c[k][d] = X(IBC(e, 2 * k + 1), d);
    flux_points(c, q);
    for (int j = 0; j < 4; ++j) {
      int f = pix.find(q[j]);
      if (f < 0) throw std::runtime_error("boundary element flux point not found among cell flux points");
      h.ibcf[e + (size_t)ne * j] = f + 1;
    }
  }
  // flpm: +1 for the first of two cells sharing a face, -1 for the other
  h.flpm.assign((size_t)nc * 6, 0);
  std::unordered_map<std::string, std::vector<std::pair<int, int>>> faces;
  auto fkey = [&](int ic, int s) {
    std::array<int, 4> f;
    for (int j = 0; j < 4; ++j) f[j] = h.ibf[ic + (size_t)nc * (s + 6 * j)];
    std::sort(f.begin(), f.end());
    return std::string(reinterpret_cast<const char *>(f.data()), sizeof(int) * 4);
  };
  for (int ic = 0; ic < nc; ++ic)
    for (int s = 0; s < 6; ++s) faces[fkey(ic, s)].push_back({ic, s});
  for (int ic = 0; ic < nc; ++ic)
    for (int s = 0; s < 6; ++s) {
      if (h.flpm[ic + (size_t)nc * s] != 0) continue;
      h.flpm[ic + (size_t)nc * s] = 1;
      const auto &lst = faces[fkey(ic, s)];
      for (auto &pr : lst)  // FindAdjCell: lowest other cell, lowest side
        if (pr.first != ic) {
          h.flpm[pr.first + (size_t)nc * pr.second] = -1;
          break;
        }
    }
}

void set_error_text(const std::string &s) { tt_set_error("%s", s.c_str()); }

}  // namespace

// ---------------------------------------------------------------------------
extern "C" int tt_host_from_arrays(int32_t nn, int32_t nb, int32_t ne, int32_t nc, const double *x, const int32_t *ibc,
                                   const int32_t *idc, const int32_t *bewn, tt_host **out) {
  try {
    auto *h = new tt_host;
    h->nn = nn;
    h->ne = ne;
    h->nc = nc;
    h->x.assign(x, x + (size_t)nn * 3);
    h->ibc.assign(ibc, ibc + (size_t)ne * 9);
    h->idc.assign(idc, idc + (size_t)nc * 27);
    h->bewn.assign(bewn, bewn + ne);
    finish_mesh(*h);
    if (nb > 0 && nb != h->nb) {
      delete h;
      tt_set_error("tt_host_from_arrays: header says %d boundary nodes, elements give %d", nb, h->nb);
      return -2;
    }
    *out = h;
    return 0;
  } catch (std::exception &e) {
    set_error_text(e.what());
    return -1;
  }
}

extern "C" int tt_host_read_geo(const char *path, tt_host **out) {
  std::ifstream in(path);
  if (!in) {
    tt_set_error("could not open mesh: %s", path);
    return -1;
  }
  std::string line;
  int nbelem, nicell, nnodes, nbnodes, npob, npoc, npx;
  if (!data_line(in, line) || !data_line(in, line) || !data_line(in, line)) {
    tt_set_error("truncated .geo header: %s", path);
    return -1;
  }
  {
    std::istringstream s(line);
    s >> nbelem >> nicell >> nnodes >> nbnodes >> npob >> npoc >> npx;
    if (!s || npob != 9 || npoc != 27 || npx != 3) {
      tt_set_error("unsupported .geo sizes (need npob=9 npoc=27 npx=3): %s", path);
      return -1;
    }
  }
  std::vector<double> x((size_t)nnodes * 3);
  std::vector<int> ibc((size_t)nbelem * 9), idc((size_t)nicell * 27), bewn(nbelem);
  auto fail = [&](const char *what) {
    tt_set_error("bad .geo %s section: %s", what, path);
    return -1;
  };
  for (int i = 0; i < nnodes; ++i) {
    if (!data_line(in, line)) return fail("node");
    std::istringstream s(line);
    int k;
    double a, b, c;
    s >> k >> a >> b >> c;
    if (!s || k < 1 || k > nnodes) return fail("node");
    x[k - 1] = a;
    x[k - 1 + nnodes] = b;
    x[k - 1 + 2 * (size_t)nnodes] = c;
  }
  for (int i = 0; i < nbelem; ++i) {
    if (!data_line(in, line)) return fail("boundary element");
    std::istringstream s(line);
    int k;
    s >> k;
    if (k < 1 || k > nbelem) return fail("boundary element");
    for (int j = 0; j < 9; ++j) s >> ibc[(k - 1) + (size_t)nbelem * j];
    if (!s) return fail("boundary element");
  }
  for (int i = 0; i < nicell; ++i) {
    if (!data_line(in, line)) return fail("cell");
    std::istringstream s(line);
    int k;
    s >> k;
    if (k < 1 || k > nicell) return fail("cell");
    for (int j = 0; j < 27; ++j) s >> idc[(k - 1) + (size_t)nicell * j];
    if (!s) return fail("cell");
  }
  for (int i = 0; i < nbelem; ++i) {
    if (!data_line(in, line)) return fail("wall number");
    std::istringstream s(line);
    int k, w;
    s >> k >> w;
    if (!s || k < 1 || k > nbelem) return fail("wall number");
    bewn[k - 1] = w;
  }
  return tt_host_from_arrays(nnodes, nbnodes, nbelem, nicell, x.data(), ibc.data(), idc.data(), bewn.data(), out);
}

extern "C" int tt_host_cube(int32_t n, tt_host **out) {
  if (n < 1 || n > 256) {
    tt_set_error("tt_host_cube: n=%d out of range 1..256", n);
    return -1;
  }
  const int g = 2 * n + 1;
  const int nn = g * g * g, nc = n * n * n, ne = 6 * n * n;
  auto nid = [&](int i, int j, int k) { return 1 + i + g * (j + g * k); };
  std::vector<double> x((size_t)nn * 3);
  for (int k = 0; k < g; ++k)
    for (int j = 0; j < g; ++j)
      for (int i = 0; i < g; ++i) {
        int id = nid(i, j, k) - 1;
        x[id] = (double)i / (double)(2 * n);
        x[id + nn] = (double)j / (double)(2 * n);
        x[id + 2 * (size_t)nn] = (double)k / (double)(2 * n);
      }
  std::vector<int> idc((size_t)nc * 27);
  for (int c = 0, cz = 0; cz < n; ++cz)
    for (int cy = 0; cy < n; ++cy)
      for (int cx = 0; cx < n; ++cx, ++c)
        for (int l = 0; l < 27; ++l)
          idc[c + (size_t)nc * l] = nid(2 * cx + (int)(2 * kLc1[l]), 2 * cy + (int)(2 * kLc2[l]), 2 * cz + (int)(2 * kLc3[l]));
  // walls: (normal axis, plane index, in-plane axes u,v with u x v = outward normal)
  struct Wall { int ax, lvl, u, v; };
  const Wall walls[6] = {{2, 0, 1, 0}, {1, 0, 0, 2}, {0, g - 1, 1, 2}, {1, g - 1, 2, 0}, {0, 0, 2, 1}, {2, g - 1, 0, 1}};
  std::vector<int> ibc((size_t)ne * 9), bewn(ne);
  int e = 0;
  for (int w = 0; w < 6; ++w)
    for (int b = 0; b < n; ++b)
      for (int a = 0; a < n; ++a, ++e) {
        // corners 1,3,5,7 at (u0,v0),(u1,v0),(u1,v1),(u0,v1); mids 2,4,6,8; centre 9
        const int uu[9] = {0, 1, 2, 2, 2, 1, 0, 0, 1}, vv[9] = {0, 0, 0, 1, 2, 2, 2, 1, 1};
        for (int k = 0; k < 9; ++k) {
          int ijk[3];
          ijk[walls[w].ax] = walls[w].lvl;
          ijk[walls[w].u] = 2 * a + uu[k];
          ijk[walls[w].v] = 2 * b + vv[k];
          ibc[e + (size_t)ne * k] = nid(ijk[0], ijk[1], ijk[2]);
        }
        bewn[e] = w + 1;
      }
  return tt_host_from_arrays(nn, 0, ne, nc, x.data(), ibc.data(), idc.data(), bewn.data(), out);
}

extern "C" int tt_host_write_geo(tt_host *h, const char *path) {
  FILE *f = std::fopen(path, "w");
  if (!f) {
    tt_set_error("cannot write %s", path);
    return -1;
  }
  std::fprintf(f, "VERS GiD-3D-1.0\r\nstructured cube\r\n");
  std::fprintf(f, "%7d %7d %7d %7d %7d %7d %7d\r\n", h->ne, h->nc, h->nn, h->nb, 9, 27, 3);
  for (int i = 0; i < h->nn; ++i)
    std::fprintf(f, "%8d %23.16e %23.16e %23.16e\r\n", i + 1, h->x[i], h->x[i + h->nn], h->x[i + 2 * (size_t)h->nn]);
  for (int e = 0; e < h->ne; ++e) {
    std::fprintf(f, "%7d", e + 1);
    for (int k = 0; k < 9; ++k) std::fprintf(f, " %7d", h->ibc[e + (size_t)h->ne * k]);
    std::fprintf(f, "\r\n");
  }
  for (int c = 0; c < h->nc; ++c) {
    std::fprintf(f, "%7d", c + 1);
    for (int k = 0; k < 27; ++k) std::fprintf(f, " %7d", h->idc[c + (size_t)h->nc * k]);
    std::fprintf(f, "\r\n");
  }
  for (int e = 0; e < h->ne; ++e) std::fprintf(f, "%7d %d\r\n", e + 1, h->bewn[e]);
  std::fclose(f);
  return 0;
}

// ---------------------------------------------------------------------------
// ReadBic
namespace {
double fnum(const std::string &s) {
  std::string t = s;
  for (auto &ch : t)
    if (ch == 'd' || ch == 'D') ch = 'e';
  return std::strtod(t.c_str(), nullptr);
}
float fnumf(const std::string &s) {
  std::string t = s;
  for (auto &ch : t)
    if (ch == 'd' || ch == 'D') ch = 'e';
  return std::strtof(t.c_str(), nullptr);
}
std::vector<std::string> toks(const std::string &line) {
  std::istringstream s(line);
  std::vector<std::string> v;
  std::string t;
  while (s >> t) {
    // list-directed input: commas separate too
    size_t p;
    while ((p = t.find(',')) != std::string::npos) {
      if (p > 0) v.push_back(t.substr(0, p));
      t = t.substr(p + 1);
    }
    if (!t.empty()) v.push_back(t);
  }
  return v;
}
}  // namespace

extern "C" int tt_host_read_bic(tt_host *hp, const char *bic_text) {
  tt_host &h = *hp;
  const int nn = h.nn, nb = h.nb, ne = h.ne, nq = h.nq;
  auto IBC = [&](int e, int k) { return h.ibc[e + (size_t)ne * (k - 1)]; };
  h.iwn.assign(nb, 0);
  for (int e = 0; e < ne; ++e)
    for (int k = 1; k <= 9; ++k) h.iwn[h.lbn[IBC(e, k) - 1] - 1] = h.bewn[e];
  h.kode.assign((size_t)nb * 3, 0);
  h.wkode.assign((size_t)nn * 3, 0);
  h.wqkode.assign((size_t)nq * 3, 0);
  h.Tkode.assign(nn, 0);
  h.Tqkode.assign(nq, 0);
  for (int d = 0; d < 3; ++d) h.nunk[d] = h.nbk[d] = 0;
  h.Tnunk = h.Tnb = 0;
  h.vel.assign((size_t)nn * 3, 0.0);
  h.vort.assign((size_t)nn * 3, 0.0);
  h.qvort.assign((size_t)nq * 3, 0.0);
  h.temp.assign(nn, 0.0);
  h.qtemp.assign(nq, 0.0);
  h.mHkode.assign(nn, 0);
  h.mHqfmkode.assign((size_t)ne * 4, 0);
  h.mHu.assign(nn, 0.0);
  h.fmqmH.assign((size_t)ne * 4, 0.0);
  h.mHnb = h.mHfmnunk = 0;
  // walls of the elements containing each boundary node (bewniwnEdge/Corner, lib_bic.f:506-576)
  std::vector<std::vector<int>> node_walls(nb);
  for (int e = 0; e < ne; ++e)
    for (int k = 1; k <= 9; ++k) node_walls[h.lbn[IBC(e, k) - 1] - 1].push_back(h.bewn[e]);
  auto has_wall = [&](int i, int w) {
    for (int x : node_walls[i])
      if (x == w) return true;
    return false;
  };
  std::istringstream in(bic_text);
  std::string line;
  auto block = [&](auto &&fn) {
    while (data_line(in, line)) {
      if (line.compare(0, 3, "END") == 0 || line.compare(0, 3, "EOF") == 0) return;
      fn();
    }
  };
  bool iperi = false;
  while (data_line(in, line)) {
    if (line.compare(0, 3, "EOF") == 0) break;
    auto t = toks(line);
    if (t.empty()) continue;
    std::string kw = t[0].substr(0, 4);
    if (kw == "EDGE") {
      block([&] {
        auto v = toks(line);
        int w1 = std::atoi(v[0].c_str()), w2 = std::atoi(v[1].c_str()), w = std::atoi(v[2].c_str());
        for (int i = 0; i < nb; ++i)
          if (has_wall(i, w1) && has_wall(i, w2)) h.iwn[i] = w;
      });
    } else if (kw == "CORN") {
      block([&] {
        auto v = toks(line);
        int w1 = std::atoi(v[0].c_str()), w2 = std::atoi(v[1].c_str()), w3 = std::atoi(v[2].c_str()),
            w = std::atoi(v[3].c_str());
        for (int i = 0; i < nb; ++i)
          if (has_wall(i, w1) && has_wall(i, w2) && has_wall(i, w3)) h.iwn[i] = w;
      });
    } else if (kw == "PERI") {
      iperi = true;
      block([&] {});
    } else if (kw == "INIT") {
      block([&] {
        auto v = toks(line);
        int ff = std::atoi(v[0].c_str());
        double val = fnum(v[1]);
        if (ff == 10) std::fill(h.temp.begin(), h.temp.end(), val);
        else if (ff == 11 || ff == 12) {
        } else if (ff >= 7) std::fill(h.vort.begin() + (size_t)nn * (ff - 7), h.vort.begin() + (size_t)nn * (ff - 6), val);
        else if (ff >= 4) std::fill(h.vort.begin() + (size_t)nn * (ff - 4), h.vort.begin() + (size_t)nn * (ff - 3), val);
        else std::fill(h.vel.begin() + (size_t)nn * (ff - 1), h.vel.begin() + (size_t)nn * ff, val);
      });
    } else if (kw == "N-VO") {
      block([&] {});
    } else if (kw == "X-KI" || kw == "Y-KI" || kw == "Z-KI") {
      int ff = kw == "X-KI" ? 0 : kw == "Y-KI" ? 1 : 2;
      block([&] {
        auto v = toks(line);
        int wall = std::atoi(v[0].c_str()), kd = std::atoi(v[1].c_str()), bct = std::atoi(v[2].c_str());
        double a = 0, b = 0, c = 0, d = 0;
        if (bct == 0) {
          data_line(in, line);
          a = fnum(toks(line)[0]);
        } else if (bct >= 1 && bct <= 3) {
          data_line(in, line);
          auto w = toks(line);
          a = fnum(w[0]), b = fnum(w[1]), c = fnum(w[2]), d = fnum(w[3]);
        }
        if (bct == 3) throw std::runtime_error("X-KINEMATICS bctype 3 (channel profile) is not supported");
        for (int i = 0; i < nb; ++i) {
          if (h.iwn[i] != wall) continue;
          int dn = h.gbn[i] - 1;
          double *u = &h.vel[(size_t)nn * ff];
          if (bct == 0) u[dn] = a;
          else if (bct == 1) u[dn] = a * h.x[dn] + b * h.x[dn + nn] + c * h.x[dn + 2 * (size_t)nn] + d;
          else if (bct == 2) {
            double z = h.x[dn + 2 * (size_t)nn];
            u[dn] = a * (z * z) + b * z + c + d;
          }
          h.kode[i + (size_t)nb * ff] = kd;
        }
      });
    } else if (kw == "X-VO" || kw == "Y-VO" || kw == "Z-VO") {
      int sm = kw == "X-VO" ? 0 : kw == "Y-VO" ? 1 : 2;
      block([&] {
        auto v = toks(line);
        int ff = std::atoi(v[0].c_str()), wall = std::atoi(v[1].c_str());
        double val = fnum(v[2]);
        if (ff == 1) {
          for (int i = 0; i < nb; ++i)
            if (h.iwn[i] == wall) {
              int dn = h.gbn[i] - 1;
              h.vort[dn + (size_t)nn * sm] = val;
              h.wkode[dn + (size_t)nn * sm] = ++h.nbk[sm];
            }
        } else if (ff == 2) {
          for (int e = 0; e < ne; ++e)
            if (h.bewn[e] == wall)
              for (int j = 0; j < 4; ++j) {
                int f = h.ibcf[e + (size_t)ne * j] - 1;
                h.qvort[f + (size_t)nq * sm] = val;
                h.wqkode[f + (size_t)nq * sm] = ++h.nbk[sm];
              }
        }
      });
    } else if (kw == "TEMP") {
      block([&] {
        auto v = toks(line);
        int ff = std::atoi(v[0].c_str()), wall = std::atoi(v[1].c_str());
        double val = fnum(v[2]);
        if (ff == 1) {
          for (int i = 0; i < nb; ++i)
            if (h.iwn[i] == wall) {
              int dn = h.gbn[i] - 1;
              h.temp[dn] = val;
              h.Tkode[dn] = ++h.Tnb;
            }
        } else if (ff == 2) {
          for (int e = 0; e < ne; ++e)
            if (h.bewn[e] == wall)
              for (int j = 0; j < 4; ++j) {
                int f = h.ibcf[e + (size_t)ne * j] - 1;
                h.qtemp[f] = val;
                h.Tqkode[f] = ++h.Tnb;
              }
        }
      });
    } else if (kw == "MHEL") {
      // modified Helmholtz: 1 wall value (u on the wall's nodes), 2 wall value
      // (q on the wall's elements' flux points) (ReadBic, lib_bic.f:366-395)
      block([&] {
        auto v = toks(line);
        int ff = std::atoi(v[0].c_str()), wall = std::atoi(v[1].c_str());
        double val = fnum(v[2]);
        if (ff == 1) {
          for (int i = 0; i < nb; ++i)
            if (h.iwn[i] == wall) {
              int dn = h.gbn[i] - 1;
              h.mHu[dn] = val;
              h.mHkode[dn] = ++h.mHnb;
            }
        } else if (ff == 2) {
          for (int e = 0; e < ne; ++e)
            if (h.bewn[e] == wall)
              for (int j = 0; j < 4; ++j) {
                h.fmqmH[(size_t)e * 4 + j] = val;
                h.mHqfmkode[(size_t)e * 4 + j] = ++h.mHnb;
              }
        }
      });
    } else if (kw == "GEDC") {
      block([&] {});
    }
  }
  if (iperi) {
    tt_set_error("periodic boundary conditions (PERI) are not supported");
    return -3;
  }
  // unknown numbering: nodes first, then flux nodes (lib_bic.f:411-430)
  for (int sm = 0; sm < 3; ++sm) {
    for (int i = 0; i < nn; ++i)
      if (h.wkode[i + (size_t)nn * sm] == 0) h.wkode[i + (size_t)nn * sm] = -(++h.nunk[sm]);
    for (int i = 0; i < nq; ++i)
      if (h.wqkode[i + (size_t)nq * sm] == 0) h.wqkode[i + (size_t)nq * sm] = -(++h.nunk[sm]);
  }
  for (int i = 0; i < nn; ++i)
    if (h.Tkode[i] == 0) h.Tkode[i] = -(++h.Tnunk);
  // modified Helmholtz, single domain: nodes, then the boundary flux points (lib_bic.f:463-480)
  for (int i = 0; i < nn; ++i)
    if (h.mHkode[i] == 0) h.mHkode[i] = -(++h.mHfmnunk);
  for (size_t i = 0; i < h.mHqfmkode.size(); ++i)
    if (h.mHqfmkode[i] == 0) h.mHqfmkode[i] = -(++h.mHfmnunk);
  h.ibcf_em.resize((size_t)ne * 4);
  for (int e = 0; e < ne; ++e)
    for (int j = 0; j < 4; ++j) h.ibcf_em[(size_t)e * 4 + j] = h.ibcf[e + (size_t)ne * j];
  for (int i = 0; i < nq; ++i)
    if (h.Tqkode[i] == 0) h.Tqkode[i] = -(++h.Tnunk);
  // DetKMunk3
  h.kmu.assign(nn, 0);
  h.nsmcol = h.nrhscol = 0;
  for (int i = 0; i < nn; ++i) h.kmu[i] = h.lbn[i] != 0 ? ++h.nrhscol : -(++h.nsmcol);
  // FillKobc only matters for dv/dn = 0 walls (kode 2 or 3)
  h.kobc.assign(nb, -1);
  bool need = false;
  for (int k : h.kode)
    if (k == 2 || k == 3) need = true;
  if (need) {
    for (int i = 0; i < nb; ++i) {
      int dn = h.gbn[i] - 1;
      double mind = 1e20;
      for (int j = 0; j < nn; ++j) {
        bool ok = false;
        if (h.lbn[j] == 0) ok = true;
        else if (h.iwn[i] != h.iwn[h.lbn[j] - 1]) {
          double a = h.nx[i] * (h.x[dn] - h.x[j]), b = h.ny[i] * (h.x[dn + nn] - h.x[j + nn]),
                 c = h.nz[i] * (h.x[dn + 2 * (size_t)nn] - h.x[j + 2 * (size_t)nn]);
          if (std::sqrt(a * a + b * b + c * c) > 1.0e-5) ok = true;
        }
        if (ok) {
          double a = h.x[dn] - h.x[j], b = h.x[dn + nn] - h.x[j + nn], c = h.x[dn + 2 * (size_t)nn] - h.x[j + 2 * (size_t)nn];
          double dist = std::sqrt(a * a + b * b + c * c);
          if (dist < mind) {
            mind = dist;
            h.kobc[i] = j + 1;
          }
        }
      }
    }
  }
  h.have_bc = true;
  return 0;
}

extern "C" int tt_host_mh_codes(tt_host *h, const int32_t **kode, const int32_t **qkode, const double **u,
                                const double **q, const int32_t **ibcf, const int32_t **iwn,
                                const int32_t **bewn, int32_t *nunk, int32_t *nknown) {
  if (!h || !h->have_bc) {
    tt_set_error("tt_host_mh_codes: no boundary conditions read");
    return -1;
  }
  *kode = h->mHkode.data(), *qkode = h->mHqfmkode.data(), *u = h->mHu.data(), *q = h->fmqmH.data();
  *ibcf = h->ibcf_em.data(), *iwn = h->iwn.data(), *bewn = h->bewn.data();
  *nunk = h->mHfmnunk, *nknown = h->mHnb;
  return 0;
}

extern "C" int tt_host_read_inp(tt_host *h, const char *text) {
  std::istringstream in(text);
  std::string line;
  while (data_line(in, line)) {
    if (line.compare(0, 3, "EOF") == 0) break;
    auto t = toks(line);
    if (t.empty()) continue;
    const std::string kw = t[0].substr(0, 4);
    auto I = [&](size_t k) { return k < t.size() ? std::atoi(t[k].c_str()) : 0; };
    auto D = [&](size_t k) { return k < t.size() ? fnum(t[k]) : 0.0; };
    // REAL(4) fields: convert the decimal text straight to single precision
    auto F = [&](size_t k) { return k < t.size() ? fnumf(t[k]) : 0.0f; };
    if (kw == "TIME") h->tstep = D(1);
    else if (kw == "LSQS") { h->lsqs_maxit = I(1); h->lsqs_eps = F(2); }
    else if (kw == "DLSE") { h->iDLSE = I(1); h->dlse_r = F(2); h->dlse_epsth = F(3); }
    else if (kw == "EQNS") { h->iBw = I(1); h->iDv = I(2); h->iDw = I(3); h->iDT = I(4); h->iDC = I(5); h->imH = I(6); }
    else if (kw == "GRAV") { h->gx = D(1); h->gy = D(2); h->gz = D(3); }
    else if (kw == "REYN") h->ReN = D(1);
    else if (kw == "RAPR") { h->RaN = D(1); h->PrN = D(2); }
    else if (kw == "URBW") { for (int k = 0; k < 3; ++k) h->urBw[k] = D(1 + k); }
    else if (kw == "URDV") { for (int k = 0; k < 3; ++k) h->urDv[k] = D(1 + k); }
    else if (kw == "URDW") { for (int k = 0; k < 3; ++k) h->urDw[k] = D(1 + k); }
    else if (kw == "URDT") h->urDT = D(1);
    else if (kw == "RELA") { h->nnlit = I(1); h->eps = D(3); }
    else if (kw == "COPY") h->copy = I(1);
    else if (kw == "GAUS") { h->quad.kmBr = I(1); h->quad.kmBs = I(2); h->quad.kmDr = I(3); h->quad.kmDs = I(4); }
    else if (kw == "IIDI") h->quad.iiDiv = I(1);
  }
  if (h->iBw < 0 || h->iBw > 1 || h->iDv != 1 || h->iDw != 1 || h->iDT < 0 || h->iDT > 1 || h->iDC != 0 || h->imH != 0) {
    tt_set_error("EQNS %d %d %d %d %d %d: only the least-squares path {0,1} 1 1 {0,1} 0 0 is implemented", h->iBw,
                 h->iDv, h->iDw, h->iDT, h->iDC, h->imH);
    return -4;
  }
  return 0;
}

extern "C" int tt_host_case_view(tt_host *hp, tt_case_view *v) {
  tt_host &h = *hp;
  std::memset(v, 0, sizeof(*v));
  v->mesh.nnodes = h.nn;
  v->mesh.nbnodes = h.nb;
  v->mesh.nbelem = h.ne;
  v->mesh.nicell = h.nc;
  v->mesh.nq = h.nq;
  v->mesh.x = h.x.data();
  v->mesh.ibc = h.ibc.data();
  v->mesh.idc = h.idc.data();
  v->mesh.ibf = h.ibf.data();
  v->mesh.xq = h.xq.data();
  v->mesh.flpm = h.flpm.data();
  v->mesh.gbn = h.gbn.data();
  v->mesh.lbn = h.lbn.data();
  v->mesh.nx = h.nx.data();
  v->mesh.ny = h.ny.data();
  v->mesh.nz = h.nz.data();
  v->quad = h.quad;
  v->copy = h.copy;
  if (!h.have_bc) return 0;  // mesh part only
  v->kode = h.kode.data();
  v->kobc = h.kobc.data();
  v->wkode = h.wkode.data();
  v->wqkode = h.wqkode.data();
  v->Tkode = h.Tkode.data();
  v->Tqkode = h.Tqkode.data();
  for (int d = 0; d < 3; ++d) {
    v->nunk[d] = h.nunk[d];
    v->nbk[d] = h.nbk[d];
    v->urBw[d] = h.urBw[d];
    v->urDv[d] = h.urDv[d];
    v->urDw[d] = h.urDw[d];
  }
  v->Tnunk = h.Tnunk;
  v->Tnb = h.Tnb;
  v->kmu = h.kmu.data();
  v->nsmcol = h.nsmcol;
  v->nrhscol = h.nrhscol;
  v->iDT = h.iDT;
  v->iBw = h.iBw;
  v->ReN = h.ReN;
  v->RaN = h.RaN;
  v->PrN = h.PrN;
  v->gx = h.gx;
  v->gy = h.gy;
  v->gz = h.gz;
  v->tstep = h.tstep;
  v->urDT = h.urDT;
  v->eps = h.eps;
  v->nnlit = h.nnlit;
  v->lsqs_maxit = h.lsqs_maxit;
  v->iDLSE = h.iDLSE;
  v->lsqs_eps = h.lsqs_eps;
  v->dlse_r = h.dlse_r;
  v->dlse_epsth = h.dlse_epsth;
  return 0;
}

extern "C" int tt_host_fields(tt_host *h, const double **vel, const double **vort, const double **qvort,
                              const double **temp, const double **qtemp) {
  if (!h->have_bc) {
    tt_set_error("tt_host_fields: boundary conditions not read");
    return -1;
  }
  *vel = h->vel.data();
  *vort = h->vort.data();
  *qvort = h->qvort.data();
  *temp = h->temp.data();
  *qtemp = h->qtemp.data();
  return 0;
}

extern "C" void tt_host_free(tt_host *h) { delete h; }
