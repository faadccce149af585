// tt_files.cpp -- the reference's on-disk formats for integrals and results,
// written and read on the host, so that matrices and fields computed on the
// device drop into an unmodified Fortran run and back:
//
//   tri.macro.int     WriteMacroIntDisk / ReadMacroIntDisk (lib_int.f:1341-1441)
//   tri.sdkm.int.N    WriteSDIntDisk / ReadSDIntDisk (lib_sdkm.f:2641-2719)
//   tri.rst           WrstF / RrstF, restart version 4 (lib_IO.f:505-685)
//   tri.sta, tri.ite  the per-iteration lines of CheckConvergencevw
//                     (lib_util.f:1141-1146) under the headers of lib_IO.f:1466-1485
//
// The binary files are Fortran unformatted sequential files: every record is
// framed by its byte length as a 4-byte integer before and after.  Vectors
// are written as wrvec does (lib_int.f:3231-3252): full records of 8192
// doubles, then one record with the rest -- an empty record when the length is
// a multiple of 8192 -- and matrices column by column (WrSMat, :3174-3186).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "tritok_bem.h"
#include "tt_internal.h"

namespace {

constexpr int64_t kChunk = 8192;  // wrvec's MAXSIZE

struct File {
  FILE *f = nullptr;
  std::string path;
  File(const char *p, const char *mode) : path(p ? p : "") {
    if (!p) throw std::runtime_error("NULL path");
    f = std::fopen(p, mode);
    if (!f) throw std::runtime_error("cannot open " + path);
  }
  ~File() {
    if (f) std::fclose(f);
  }
  void put(const void *d, size_t n) {
    if (n && std::fwrite(d, 1, n, f) != n) throw std::runtime_error("write failed: " + path);
  }
  void get(void *d, size_t n) {
    if (n && std::fread(d, 1, n, f) != n) throw std::runtime_error("unexpected end of file: " + path);
  }
  // one record of n bytes
  void record(const void *d, size_t n) {
    if (n > 0x7fffffffu) throw std::runtime_error("record too long");
    const int32_t m = (int32_t)n;
    put(&m, 4), put(d, n), put(&m, 4);
  }
  // reads one record, which must hold exactly n bytes
  void read_record(void *d, size_t n) {
    int32_t m = 0, m2 = 0;
    get(&m, 4);
    if (m < 0 || (size_t)m != n)
      throw std::runtime_error(path + ": record of " + std::to_string(m) + " bytes where " + std::to_string(n) +
                               " were expected");
    get(d, n);
    get(&m2, 4);
    if (m2 != m) throw std::runtime_error(path + ": record markers differ");
  }
  void vec(const double *v, int64_t n) {  // wrvec
    const int64_t nblock = n / kChunk;
    for (int64_t j = 0; j < nblock; ++j) record(v + j * kChunk, sizeof(double) * kChunk);
    record(v + nblock * kChunk, sizeof(double) * (size_t)(n - nblock * kChunk));
  }
  void read_vec(double *v, int64_t n) {  // rdvec
    const int64_t nblock = n / kChunk;
    for (int64_t j = 0; j < nblock; ++j) read_record(v + j * kChunk, sizeof(double) * kChunk);
    read_record(v + nblock * kChunk, sizeof(double) * (size_t)(n - nblock * kChunk));
  }
  void mat(const double *a, int64_t nrow, int64_t ncol) {  // WrSMat, column-major
    for (int64_t j = 0; j < ncol; ++j) vec(a + j * nrow, nrow);
  }
  void read_mat(double *a, int64_t nrow, int64_t ncol) {
    for (int64_t j = 0; j < ncol; ++j) read_vec(a + j * nrow, nrow);
  }
};

template <typename F>
int guard(const char *fn, F &&f) {
  try {
    return f();
  } catch (std::exception &e) {
    tt_set_error("%s: %s", fn, e.what());
    return -1;
  }
}

// Fortran G10.5 (two leading blanks come from the 2X before each item): F
// editing with d - k decimals and four trailing blanks for 0.1 <= |x| < 1e5
// (after rounding to five significant digits), E10.5 otherwise; a leading zero
// that does not fit the field is left out, as the Fortran runtime does.
std::string g10_5(double x) {
  char b[64];
  if (!std::isfinite(x)) {
    std::snprintf(b, sizeof(b), "%10s", std::isnan(x) ? "NaN" : (x > 0 ? "Infinity" : "-Infinity"));
    return b;
  }
  if (x == 0.0) return "0.0000    ";
  std::snprintf(b, sizeof(b), "%.4E", x);  // five significant digits: [-]d.ddddE+ee
  const char *e = std::strchr(b, 'E');
  const int e10 = std::atoi(e + 1);
  auto fit = [](std::string s, size_t w) {
    if (s.size() > w) {  // drop the optional leading zero
      const size_t z = s[0] == '-' ? 1 : 0;
      if (s.size() > z + 1 && s[z] == '0' && s[z + 1] == '.') s.erase(z, 1);
    }
    if (s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
  };
  if (e10 >= -1 && e10 <= 4) {  // F(6).(5-k) with k = e10 + 1, then 4 blanks
    const int dec = 5 - (e10 + 1);
    char f[64];
    std::snprintf(f, sizeof(f), dec > 0 ? "%.*f" : "%.*f.", dec, x);  // F editing always shows the point
    return fit(f, 6) + "    ";
  }
  // E10.5: [-]0.dddddE+ee from d.dddd x 10^e10
  const bool neg = b[0] == '-';
  const char *m = b + (neg ? 1 : 0);
  std::string digits;
  digits += m[0];
  digits.append(m + 2, 4);
  const int ex = e10 + 1;
  char t[64];
  if (std::abs(ex) <= 99)
    std::snprintf(t, sizeof(t), "%s0.%sE%c%02d", neg ? "-" : "", digits.c_str(), ex < 0 ? '-' : '+', std::abs(ex));
  else
    std::snprintf(t, sizeof(t), "%s0.%s%c%03d", neg ? "-" : "", digits.c_str(), ex < 0 ? '-' : '+', std::abs(ex));
  return fit(t, 10);
}

}  // namespace

extern "C" int tt_write_macro_int(const char *path, int32_t nicell, int32_t nnodes, int32_t nbnodes, const double *H,
                                  const double *G, const double *B, const double *Abdx, const double *Abdy,
                                  const double *Abdz, const double *Htx, const double *Hty, const double *Htz,
                                  const double *Dx, const double *Dy, const double *Dz, const double ont[11]) {
  return guard("tt_write_macro_int", [&] {
    if (nicell < 1 || !H || !G || !B || !Abdx || !Abdy || !Abdz || !Htx || !Hty || !Htz || !Dx || !Dy || !Dz || !ont)
      throw std::runtime_error("bad argument");
    File f(path, "wb");
    const int32_t ver = 1;  // inp%INTversion (main.f90:249)
    const int32_t hdr[2] = {nnodes, nbnodes};
    f.record(&ver, 4);
    f.record(hdr, 8);
    const int64_t m1 = (int64_t)nicell * 51, m2 = (int64_t)nicell * 27;
    f.mat(H, m1, 27), f.mat(G, m1, 24), f.mat(B, m1, 27);
    f.mat(Abdx, m1, 27), f.mat(Abdy, m1, 27), f.mat(Abdz, m1, 27);
    f.mat(Htx, m2, 27), f.mat(Hty, m2, 27), f.mat(Htz, m2, 27);
    f.mat(Dx, m1, 27), f.mat(Dy, m1, 27), f.mat(Dz, m1, 27);
    f.record(ont, 11 * sizeof(double));
    return 0;
  });
}

extern "C" int tt_read_macro_int(const char *path, int32_t nicell, int32_t *nnodes, int32_t *nbnodes, double *H,
                                 double *G, double *B, double *Abdx, double *Abdy, double *Abdz, double *Htx,
                                 double *Hty, double *Htz, double *Dx, double *Dy, double *Dz, double ont[11]) {
  return guard("tt_read_macro_int", [&] {
    if (nicell < 1 || !H || !G || !B || !Abdx || !Abdy || !Abdz || !Htx || !Hty || !Htz || !Dx || !Dy || !Dz || !ont)
      throw std::runtime_error("bad argument");
    File f(path, "rb");
    int32_t ver = 0, hdr[2] = {0, 0};
    f.read_record(&ver, 4);
    f.read_record(hdr, 8);
    if (nnodes) *nnodes = hdr[0];
    if (nbnodes) *nbnodes = hdr[1];
    const int64_t m1 = (int64_t)nicell * 51, m2 = (int64_t)nicell * 27;
    f.read_mat(H, m1, 27), f.read_mat(G, m1, 24), f.read_mat(B, m1, 27);
    f.read_mat(Abdx, m1, 27), f.read_mat(Abdy, m1, 27), f.read_mat(Abdz, m1, 27);
    f.read_mat(Htx, m2, 27), f.read_mat(Hty, m2, 27), f.read_mat(Htz, m2, 27);
    f.read_mat(Dx, m1, 27), f.read_mat(Dy, m1, 27), f.read_mat(Dz, m1, 27);
    f.read_record(ont, 11 * sizeof(double));
    return 0;
  });
}

extern "C" int tt_write_sdkm_int(const char *path, int32_t nnodes, int32_t nbnodes, int32_t nproc, int32_t rows,
                                 const double *H, const double *Htx, const double *Hty, const double *Htz,
                                 const double *Dx, const double *Dy, const double *Dz, const double ont[10]) {
  return guard("tt_write_sdkm_int", [&] {
    if (rows < 1 || nbnodes < 1 || nnodes < 1 || !H || !Htx || !Hty || !Htz || !Dx || !Dy || !Dz || !ont)
      throw std::runtime_error("bad argument");
    File f(path, "wb");
    const int32_t ver = 1;
    const int32_t hdr[3] = {nnodes, nbnodes, nproc};
    f.record(&ver, 4);
    f.record(hdr, 12);
    f.mat(H, rows, nbnodes), f.mat(Htx, rows, nbnodes), f.mat(Hty, rows, nbnodes), f.mat(Htz, rows, nbnodes);
    f.mat(Dx, rows, nnodes), f.mat(Dy, rows, nnodes), f.mat(Dz, rows, nnodes);
    f.record(ont, 10 * sizeof(double));
    return 0;
  });
}

extern "C" int tt_read_sdkm_int(const char *path, int32_t rows, int32_t hdr_out[3], double *H, double *Htx,
                                double *Hty, double *Htz, double *Dx, double *Dy, double *Dz, double ont[10]) {
  return guard("tt_read_sdkm_int", [&] {
    if (rows < 1 || !H || !Htx || !Hty || !Htz || !Dx || !Dy || !Dz || !ont) throw std::runtime_error("bad argument");
    File f(path, "rb");
    int32_t ver = 0, hdr[3] = {0, 0, 0};
    f.read_record(&ver, 4);
    f.read_record(hdr, 12);
    if (hdr_out) std::memcpy(hdr_out, hdr, sizeof(hdr));
    const int32_t nnodes = hdr[0], nbnodes = hdr[1];
    if (nnodes < 1 || nbnodes < 1) throw std::runtime_error("bad header");
    f.read_mat(H, rows, nbnodes), f.read_mat(Htx, rows, nbnodes), f.read_mat(Hty, rows, nbnodes),
        f.read_mat(Htz, rows, nbnodes);
    f.read_mat(Dx, rows, nnodes), f.read_mat(Dy, rows, nnodes), f.read_mat(Dz, rows, nnodes);
    f.read_record(ont, 10 * sizeof(double));
    return 0;
  });
}

namespace {
// the seventeen fields of WrstF in file order: (nnodes or nq) x (3 or 1)
struct RstLayout {
  int64_t rows;
  int cols;
};
void rst_layout(int32_t nn, int32_t nq, RstLayout out[17]) {
  const RstLayout L[17] = {{nn, 3}, {nn, 3}, {nn, 3}, {nn, 3}, {nn, 3}, {nn, 3}, {nq, 3}, {nn, 1}, {nn, 1},
                           {nn, 1}, {nn, 1}, {nq, 1}, {nn, 1}, {nn, 1}, {nn, 1}, {nn, 1}, {nq, 1}};
  std::memcpy(out, L, sizeof(L));
}
}  // namespace

// WritemHfullIntDisk (lib_fm.f:1822-1895): version, (nnodes, nbnodes), beta,
// sum of diff, mumax, the ten U matrices (nnodes rows) and the ten Q
// matrices (nbnpof rows) in the order H, G, B, Bp, Abdx, Abdy, Abdz, Dx, Dy,
// Dz, then the five accuracy estimates.
extern "C" int tt_write_mhsd_int(const char *path, int32_t version, int32_t nnodes, int32_t nbnodes, int32_t nbnpof,
                                 double beta, double sumdiff, double mumax, const double *const U[10],
                                 const double *const Q[10], const double ont[5]) {
  return guard("tt_write_mhsd_int", [&] {
    if (nnodes < 1 || nbnodes < 1 || nbnpof < 1 || !U || !Q || !ont) throw std::runtime_error("bad argument");
    const int32_t cols[10] = {nbnodes, nbnpof, nnodes, nnodes, nbnodes, nbnodes, nbnodes, nnodes, nnodes, nnodes};
    for (int k = 0; k < 10; ++k)
      if (!U[k] || !Q[k]) throw std::runtime_error("NULL matrix");
    File f(path, "wb");
    const int32_t hdr[2] = {nnodes, nbnodes};
    f.record(&version, 4);
    f.record(hdr, 8);
    f.record(&beta, 8), f.record(&sumdiff, 8), f.record(&mumax, 8);
    for (int k = 0; k < 10; ++k) f.mat(U[k], nnodes, cols[k]);
    for (int k = 0; k < 10; ++k) f.mat(Q[k], nbnpof, cols[k]);
    f.record(ont, 5 * sizeof(double));
    return 0;
  });
}

// ReadmHfullIntDisk (lib_fm.f:1746-1820) after CheckmHfullIntFile (:1897-1916):
// hdr_out = version, nnodes, nbnodes; hd_out = beta, sum of diff, mumax.
extern "C" int tt_read_mhsd_int(const char *path, int32_t nbnpof, int32_t hdr_out[3], double hd_out[3],
                                double *const U[10], double *const Q[10], double ont[5]) {
  return guard("tt_read_mhsd_int", [&] {
    if (nbnpof < 1 || !hdr_out || !hd_out || !U || !Q || !ont) throw std::runtime_error("bad argument");
    File f(path, "rb");
    int32_t ver = 0, hdr[2] = {0, 0};
    f.read_record(&ver, 4);
    f.read_record(hdr, 8);
    hdr_out[0] = ver, hdr_out[1] = hdr[0], hdr_out[2] = hdr[1];
    f.read_record(&hd_out[0], 8), f.read_record(&hd_out[1], 8), f.read_record(&hd_out[2], 8);
    const int32_t nnodes = hdr[0], nbnodes = hdr[1];
    if (nnodes < 1 || nbnodes < 1) throw std::runtime_error("bad header");
    const int32_t cols[10] = {nbnodes, nbnpof, nnodes, nnodes, nbnodes, nbnodes, nbnodes, nnodes, nnodes, nnodes};
    for (int k = 0; k < 10; ++k) f.read_mat(U[k], nnodes, cols[k]);
    for (int k = 0; k < 10; ++k) f.read_mat(Q[k], nbnpof, cols[k]);
    f.read_record(ont, 5 * sizeof(double));
    return 0;
  });
}

extern "C" int tt_write_restart(const char *path, int32_t nnodes, int32_t nq, int32_t tstep, int32_t glnnlit,
                                double rtime, const double *const fields[17]) {
  return guard("tt_write_restart", [&] {
    if (nnodes < 1 || nq < 1 || !fields) throw std::runtime_error("bad argument");
    RstLayout L[17];
    rst_layout(nnodes, nq, L);
    for (int k = 0; k < 17; ++k)
      if (!fields[k]) throw std::runtime_error("NULL field");
    File f(path, "wb");
    const int32_t ver = 4;  // inp%RSTversion (main.f90:250)
    const int32_t cnt[2] = {tstep, glnnlit};
    f.record(&ver, 4);
    f.record(cnt, 8);
    f.record(&rtime, 8);
    for (int k = 0; k < 17; ++k)
      for (int j = 0; j < L[k].cols; ++j) f.vec(fields[k] + j * L[k].rows, L[k].rows);
    return 0;
  });
}

extern "C" int tt_read_restart(const char *path, int32_t nnodes, int32_t nq, int32_t *version, int32_t *tstep,
                               int32_t *glnnlit, double *rtime, double *const fields[17]) {
  return guard("tt_read_restart", [&] {
    if (nnodes < 1 || nq < 1 || !fields) throw std::runtime_error("bad argument");
    RstLayout L[17];
    rst_layout(nnodes, nq, L);
    File f(path, "rb");
    int32_t ver = 0, cnt[2] = {0, 0};
    double rt = 0.0;
    f.read_record(&ver, 4);
    if (ver < 3 || ver > 4) throw std::runtime_error("restart version " + std::to_string(ver) + " (3 or 4 read)");
    f.read_record(cnt, 8);
    f.read_record(&rt, 8);
    const int nf = ver >= 4 ? 17 : 12;  // version 3 has no DC fields
    for (int k = 0; k < nf; ++k)
      for (int j = 0; j < L[k].cols; ++j) {
        if (!fields[k]) throw std::runtime_error("NULL field");
        f.read_vec(fields[k] + j * L[k].rows, L[k].rows);
      }
    if (version) *version = ver;
    if (tstep) *tstep = cnt[0];
    if (glnnlit) *glnnlit = cnt[1];
    if (rtime) *rtime = rt;
    return 0;
  });
}

extern "C" int tt_write_sta_ite(const char *sta_path, const char *ite_path, const char *id_line, int32_t nrows,
                                const int32_t *tstep, const int32_t *tsnnlit, const double *err, const int32_t *nit) {
  return guard("tt_write_sta_ite", [&] {
    if (nrows < 0 || (nrows > 0 && (!tstep || !tsnnlit))) throw std::runtime_error("bad argument");
    const std::string id = id_line ? id_line : "TriTok (tritok_amd)";
    if (sta_path) {
      if (nrows > 0 && !err) throw std::runtime_error("NULL err");
      File f(sta_path, "w");
      // '(A,8(10X,A2))' with the first column's name inside the A item
      std::string s = id + "\n" + " tstep  tsite      vx";
      const char *cols[8] = {"vy", "vz", "wx", "wy", "wz", " T", "DC", "mH"};
      for (int k = 0; k < 8; ++k) s += std::string(10, ' ') + cols[k];
      s += "\n";
      for (int r = 0; r < nrows; ++r) {
        char b[32];
        std::snprintf(b, sizeof(b), "%6d %6d", tstep[r], tsnnlit[r]);
        s += b;
        for (int k = 0; k < 9; ++k) s += "  " + g10_5(err[(size_t)r * 9 + k]);
        s += "\n";
      }
      f.put(s.data(), s.size());
    }
    if (ite_path) {
      if (nrows > 0 && !nit) throw std::runtime_error("NULL nit");
      File f(ite_path, "w");
      std::string s = id + "\n" + " tstep  tsite";
      const char *cols[12] = {"Bwx", "Bwy", "Bwz", "Dvx", "Dvy", "Dvz", "Dwx", "Dwy", "Dwz", "  T", " DC", " mH"};
      for (int k = 0; k < 12; ++k) s += std::string("  ") + cols[k];
      s += "\n";
      for (int r = 0; r < nrows; ++r) {
        char b[128];
        std::snprintf(b, sizeof(b), "%6d %6d", tstep[r], tsnnlit[r]);
        s += b;
        for (int k = 0; k < 12; ++k) {
          std::snprintf(b, sizeof(b), "%5d", nit[(size_t)r * 12 + k]);
          s += std::strlen(b) > 5 ? "*****" : b;
        }
        s += "\n";
      }
      f.put(s.data(), s.size());
    }
    return 0;
  });
}
