// output.hip -- the reference's output formats, host side.
//
// * HIST energy files (text): fld_ene (report_energy_emf, emf/diagnostics/
//   os-emf-diagnostics.f03:44-85), par_ene (report_energy_particles, particles/
//   os-particles.f03:1710-1780) and parNN_ene (report_energy_species, spec/diagnostics/
//   os-spec-diagnostics.f03:1840-1885).  The rows are written with the reference's Fortran
//   edit descriptors (I6, G15.8, ES23.16, I14, A6/A15/A23), emulated here: ES from the
//   correctly rounded decimal expansion, G choosing F or E editing from the value rounded to
//   d significant digits (Fortran 2008 10.7.5.2.2 with the NEAREST rounding mode).
// * ZDF files (binary, little endian, 4-byte units): the record stream of zdf/zdf.c --
//   magic "ZDF1"; records of (uint32 id, string name, uint64 length) + payload; strings as
//   uint32 length + bytes padded to 4; datasets as a header (uint32 id, int32 type, uint32
//   ndims, uint64 count[ndims]) + raw values.  Grid files are laid out as OSIRIS's diagnostic
//   files write them (dutil/os-diagfile-zdf.f03:297-347, 481-595): TYPE, grid info,
//   iteration, then a chunked dataset with one chunk per writer; particle files as
//   zdf_open_part_file + zdf_add_quant_part_file.
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace osr_out {

// ---- Fortran edit descriptors ------------------------------------------------------
std::string stars(int w) { return std::string((size_t)w, '*'); }

std::string pad_left(const std::string& s, int w) {
  if ((int)s.size() > w) return stars(w);
  return std::string((size_t)(w - (int)s.size()), ' ') + s;
}

// Iw
std::string f_i(long long v, int w) {
  char b[64];
  snprintf(b, sizeof b, "%lld", v);
  return pad_left(b, w);
}

// Aw: right-justified when w exceeds the length, else the leftmost w characters
std::string f_a(const std::string& s, int w) {
  if ((int)s.size() >= w) return s.substr(0, (size_t)w);
  return std::string((size_t)(w - (int)s.size()), ' ') + s;
}

// decimal exponent suffix of E/ES editing without an Ee field: E+dd up to 99, +ddd above
bool exp_field(int e, std::string& out) {
  const int a = std::abs(e);
  char b[16];
  if (a <= 99)
    snprintf(b, sizeof b, "E%c%02d", e < 0 ? '-' : '+', a);
  else if (a <= 999)
    snprintf(b, sizeof b, "%c%03d", e < 0 ? '-' : '+', a);
  else
    return false;
  out = b;
  return true;
}

// d significant digits of |v| (correctly rounded) and the decimal exponent x of the
// leading one: |v| ~ D1.D2...Dd x 10^x
void sig_digits(double v, int d, std::string& digits, int& x) {
  char b[64];
  snprintf(b, sizeof b, "%.*E", d - 1, std::fabs(v));
  digits.clear();
  const char* p = b;
  for (; *p && *p != 'E'; p++)
    if (*p >= '0' && *p <= '9') digits.push_back(*p);
  x = *p ? atoi(p + 1) : 0;
}

// ESw.d
std::string f_es(double v, int w, int d) {
  if (std::isnan(v)) return pad_left("NaN", w);
  if (std::isinf(v)) return pad_left(v < 0 ? "-Infinity" : "Infinity", w);
  std::string dg, ex;
  int x = 0;
  if (v == 0.0) {
    dg.assign((size_t)(d + 1), '0');
  } else {
    sig_digits(v, d + 1, dg, x);
  }
  if (!exp_field(x, ex)) return stars(w);
  std::string s = (std::signbit(v) ? "-" : "") + dg.substr(0, 1) + "." + dg.substr(1) + ex;
  return pad_left(s, w);
}

// Gw.d (no exponent width): F(w-4).(d-k) plus four blanks when the value rounded to d
// significant digits lies in [10^(k-1), 10^k) with 0 <= k <= d, otherwise Ew.d
std::string f_g(double v, int w, int d) {
  if (std::isnan(v) || std::isinf(v)) return f_es(v, w, d);
  char b[64];
  if (v == 0.0) {
    snprintf(b, sizeof b, "%.*f", d - 1, 0.0);
    return pad_left(b, w - 4) + "    ";
  }
  std::string dg, ex;
  int x = 0;
  sig_digits(v, d, dg, x);
  const int k = x + 1;  // |v| ~ 0.D1D2...Dd x 10^k
  if (k >= 0 && k <= d) {
    snprintf(b, sizeof b, "%.*f", d - k, v);
    std::string f = b;
    if (d == k) f += ".";  // F editing always shows the decimal point
    return pad_left(f, w - 4) + "    ";
  }
  if (!exp_field(k, ex)) return stars(w);
  std::string s = (v < 0 ? "-0." : "0.") + dg + ex;
  return pad_left(s, w);
}

// ---- files ---------------------------------------------------------------------------
int make_dirs(const std::string& path) {
  std::string cur;
  size_t i = 0;
  while (i <= path.size()) {
    const size_t j = path.find('/', i);
    cur = path.substr(0, j == std::string::npos ? path.size() : j);
    if (!cur.empty() && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return errno;
    if (j == std::string::npos) break;
    i = j + 1;
  }
  return 0;
}

std::string join(const std::string& dir, const std::string& name) {
  if (dir.empty()) return name;
  return dir.back() == '/' ? dir + name : dir + "/" + name;
}

// ---- ZDF record stream ---------------------------------------------------------------
constexpr uint32_t kZdfString = 0x00030000, kZdfDataset = 0x00100002,
                   kZdfCdsetStart = 0x00110000, kZdfCdsetChunk = 0x00120000,
                   kZdfCdsetEnd = 0x00130000, kZdfIteration = 0x00200001,
                   kZdfGridInfo = 0x00210001, kZdfPartInfo = 0x00220002;

size_t zdf_type_size(int t) {
  switch (t) {
    case 1: case 2: return 1;      // int8, uint8
    case 3: case 4: return 2;      // int16, uint16
    case 5: case 6: case 9: return 4;  // int32, uint32, float32
    case 7: case 8: case 10: return 8; // int64, uint64, float64
    default: return 0;
  }
}

struct ZdfWriter {
  std::vector<unsigned char> buf;  // the whole file, written out at the end
  uint32_t ndatasets = 0;
  template <class T>
  void put(T v) {
    const unsigned char* p = (const unsigned char*)&v;  // little-endian host
    buf.insert(buf.end(), p, p + sizeof(T));
  }
  void bytes(const void* p, size_t n) {
    const unsigned char* c = (const unsigned char*)p;
    buf.insert(buf.end(), c, c + n);
  }
  static size_t pad4(size_t n) { return (n + 3) / 4 * 4; }
  static size_t str_size(const char* s) {
    const size_t n = s ? strlen(s) : 0;
    return 4 + (n ? pad4(n) : 0);
  }
  void str(const char* s) {
    const uint32_t n = s ? (uint32_t)strlen(s) : 0;
    put<uint32_t>(n);
    if (n) {
      bytes(s, n);
      for (size_t k = n; k < pad4(n); k++) buf.push_back(0);
    }
  }
  void record(uint32_t id, const char* name, uint64_t len) {
    put<uint32_t>(id);
    str(name);
    put<uint64_t>(len);
  }
  ZdfWriter() { bytes("ZDF1", 4); }
  void add_string(const char* name, const char* s) {
    record(kZdfString, name, str_size(s));
    str(s);
  }
  void add_iteration(const char* name, int32_t n, double t, const char* units) {
    record(kZdfIteration, name, 4 + 8 + str_size(units));
    put<int32_t>(n);
    put<double>(t);
    str(units);
  }
  void add_grid_info(const osr_zdf_grid_info* g) {
    size_t len = 4 + 8 * (size_t)g->ndims + str_size(g->label) + str_size(g->units) + 4;
    if (g->axis)
      for (int i = 0; i < g->ndims; i++)
        len += str_size(g->axis[i].name) + 4 + 16 + str_size(g->axis[i].label) +
               str_size(g->axis[i].units);
    record(kZdfGridInfo, g->name, len);
    put<uint32_t>((uint32_t)g->ndims);
    for (int i = 0; i < g->ndims; i++) put<uint64_t>(g->count[i]);
    str(g->label);
    str(g->units);
    put<int32_t>(g->axis ? 1 : 0);
    if (g->axis)
      for (int i = 0; i < g->ndims; i++) {
        str(g->axis[i].name);
        put<int32_t>(g->axis[i].type);
        put<double>(g->axis[i].min);
        put<double>(g->axis[i].max);
        str(g->axis[i].label);
        str(g->axis[i].units);
      }
  }
  void dataset_header(uint32_t id, int type, int ndims, const uint64_t* count) {
    put<uint32_t>(id);
    put<int32_t>(type);
    put<uint32_t>((uint32_t)ndims);
    for (int i = 0; i < ndims; i++) put<uint64_t>(count[i]);
  }
  void add_dataset(const char* name, int type, int ndims, const uint64_t* count,
                   const void* data) {
    size_t n = 1;
    for (int i = 0; i < ndims; i++) n *= (size_t)count[i];
    const size_t nb = n * zdf_type_size(type);
    record(kZdfDataset, name, nb + 12 + 8 * (size_t)ndims);
    dataset_header(++ndatasets, type, ndims, count);
    bytes(data, nb);
    if (zdf_type_size(type) == 1)
      for (size_t k = nb; k < pad4(nb); k++) buf.push_back(0);
  }
  // a chunked dataset written as one chunk covering all of it
  void add_cdset_one_chunk(const char* name, int type, int ndims, const uint64_t* count,
                           const void* data) {
    record(kZdfCdsetStart, name, 12 + 8 * (size_t)ndims);
    const uint32_t id = ++ndatasets;
    dataset_header(id, type, ndims, count);
    size_t n = 1;
    for (int i = 0; i < ndims; i++) n *= (size_t)count[i];
    const size_t nb = n * zdf_type_size(type);
    char cname[16];
    snprintf(cname, sizeof cname, "%08llx-chunk", (unsigned long long)id);
    record(kZdfCdsetChunk, cname, nb + 4 + 24 * (size_t)ndims);
    put<uint32_t>(id);
    for (int i = 0; i < ndims; i++) put<uint64_t>(count[i]);  // count
    for (int i = 0; i < ndims; i++) put<uint64_t>(0);         // start
    for (int i = 0; i < ndims; i++) put<uint64_t>(1);         // stride
    bytes(data, nb);
    if (zdf_type_size(type) == 1)
      for (size_t k = nb; k < pad4(nb); k++) buf.push_back(0);
    snprintf(cname, sizeof cname, "%08llx-end", (unsigned long long)id);
    record(kZdfCdsetEnd, cname, 0);
  }
  int save(const std::string& file) const {
    FILE* f = fopen(file.c_str(), "wb");
    if (!f) return errno ? errno : 1;
    const size_t w = fwrite(buf.data(), 1, buf.size(), f);
    const int c = fclose(f);
    return (w != buf.size() || c) ? 1 : 0;
  }
};

}  // namespace osr_out

extern "C" {

int osr_hist_append(const char* dir, int nspec, const char* const* names, int64_t nrows,
                    const int64_t* iter, const double* time, const double* fld,
                    const double* par, const int64_t* npar) {
  using namespace osr_out;
  if (!dir || nspec < 0 || nspec > 99 || nrows < 0 || (nrows && (!iter || !time || !fld)) ||
      (nspec && nrows && !par))
    return fail(OSR_ERR_ARG, "bad argument");
  if (make_dirs(dir)) return fail(OSR_ERR_ARG, "cannot create directory %s", dir);
  auto name_of = [&](int s) -> std::string {
    return names && names[s] ? std::string(names[s]) : "species " + std::to_string(s + 1);
  };
  for (int64_t r = 0; r < nrows; r++) {
    const bool fresh = time[r] == 0.0;
    // fld_ene
    {
      const std::string fn = join(dir, "fld_ene");
      FILE* f = fopen(fn.c_str(), fresh ? "w" : "a");
      if (!f) return fail(OSR_ERR_ARG, "cannot open %s", fn.c_str());
      std::string s;
      if (fresh) {
        s += "! EM field energy per field component\n";
        s += f_a("Iter", 6) + " " + f_a("Time    ", 15);
        const char* h[6] = {"B1", "B2", "B3", "E1", "E2", "E3"};
        for (int c = 0; c < 6; c++) s += " " + f_a(h[c], 23);
        s += "\n";
      }
      s += f_i(iter[r], 6) + " " + f_g(time[r], 15, 8);
      for (int c = 0; c < 6; c++) s += " " + f_es(fld[6 * r + c], 23, 16);
      s += "\n";
      fputs(s.c_str(), f);
      fclose(f);
    }
    if (nspec == 0) continue;
    // par_ene
    {
      const std::string fn = join(dir, "par_ene");
      FILE* f = fopen(fn.c_str(), fresh ? "w" : "a");
      if (!f) return fail(OSR_ERR_ARG, "cannot open %s", fn.c_str());
      std::string s;
      if (fresh) {
        s += "! Total and per species particle kinetic energy\n";
        s += f_a("Iter", 6) + " " + f_a("Time    ", 15) + " " + f_a("Total", 23);
        for (int k = 0; k < nspec; k++) s += " " + f_a(name_of(k), 23);
        s += "\n";
      }
      double tot = 0;
      for (int k = 0; k < nspec; k++) tot = tot + par[nspec * r + k];
      s += f_i(iter[r], 6) + " " + f_g(time[r], 15, 8) + " " + f_es(tot, 23, 16);
      for (int k = 0; k < nspec; k++) s += " " + f_es(par[nspec * r + k], 23, 16);
      s += "\n";
      fputs(s.c_str(), f);
      fclose(f);
    }
    // parNN_ene
    for (int k = 0; k < nspec; k++) {
      char nm[32];
      snprintf(nm, sizeof nm, "par%02d_ene", k + 1);
      const std::string fn = join(dir, nm);
      FILE* f = fopen(fn.c_str(), fresh ? "w" : "a");
      if (!f) return fail(OSR_ERR_ARG, "cannot open %s", fn.c_str());
      std::string s;
      if (fresh) {
        s += "! Kinetic energy for species \"" + name_of(k) + "\"\n";
        s += f_a("Iter", 6) + " " + f_a("Time    ", 15) + " " + f_a("Total Par.", 15) + " " +
             f_a("Kin. Energy", 23) + "\n";
      }
      s += f_i(iter[r], 6) + " " + f_g(time[r], 15, 8) + " " + f_i(npar ? npar[k] : 0, 14) +
           " " + f_es(par[nspec * r + k], 23, 16) + "\n";
      fputs(s.c_str(), f);
      fclose(f);
    }
  }
  return OSR_OK;
}

int osr_zdf_save_grid(const void* data, int data_type, const osr_zdf_grid_info* info,
                      const osr_zdf_iteration* iter, const char* path) {
  using namespace osr_out;
  if (!data || !info || !iter || !path || info->ndims < 1 || info->ndims > 3 ||
      !zdf_type_size(data_type) || !info->label)
    return fail(OSR_ERR_ARG, "bad argument");
  if (make_dirs(path)) return fail(OSR_ERR_ARG, "cannot create directory %s", path);
  char fn[1024];
  snprintf(fn, sizeof fn, "%s/%s-%06u.zdf", path, info->label, (unsigned)iter->n);
  ZdfWriter z;
  z.add_string("TYPE", "grid");
  z.add_grid_info(info);
  z.add_iteration(iter->name, iter->n, iter->t, iter->time_units);
  z.add_dataset(info->name, data_type, info->ndims, info->count, data);
  if (z.save(fn)) return fail(OSR_ERR_ARG, "cannot write %s", fn);
  return OSR_OK;
}

int osr_zdf_write_grid_file(const char* file, const void* data, int data_type,
                            const osr_zdf_grid_info* info, const osr_zdf_iteration* iter) {
  using namespace osr_out;
  if (!file || !data || !info || !iter || info->ndims < 1 || info->ndims > 3 ||
      !zdf_type_size(data_type))
    return fail(OSR_ERR_ARG, "bad argument");
  ZdfWriter z;
  z.add_string("TYPE", "grid");
  z.add_grid_info(info);
  z.add_iteration(iter->name, iter->n, iter->t, iter->time_units);
  z.add_cdset_one_chunk(info->name, data_type, info->ndims, info->count, data);
  if (z.save(file)) return fail(OSR_ERR_ARG, "cannot write %s", file);
  return OSR_OK;
}

int osr_zdf_save_particles(const osr_zdf_part_info* info, const osr_zdf_iteration* iter,
                           const char* path, const float* const* quants) {
  using namespace osr_out;
  if (!info || !iter || !path || !info->name || info->nquants < 0 ||
      (info->nquants && (!quants || !info->quants || !info->qlabels || !info->qunits)))
    return fail(OSR_ERR_ARG, "bad argument");
  if (make_dirs(path)) return fail(OSR_ERR_ARG, "cannot create directory %s", path);
  char fn[1024];
  snprintf(fn, sizeof fn, "%s/%s-%s-%06u.zdf", path, "particles", info->name, (unsigned)iter->n);
  ZdfWriter z;
  z.add_string("TYPE", "particles");
  size_t len = ZdfWriter::str_size(info->label) + 8 + 4;
  for (int i = 0; i < info->nquants; i++)
    len += ZdfWriter::str_size(info->quants[i]) + ZdfWriter::str_size(info->qlabels[i]) +
           ZdfWriter::str_size(info->qunits[i]);
  z.record(kZdfPartInfo, info->name, len);
  z.str(info->label);
  z.put<uint64_t>(info->np);
  z.put<uint32_t>((uint32_t)info->nquants);
  for (int i = 0; i < info->nquants; i++) z.str(info->quants[i]);
  for (int i = 0; i < info->nquants; i++) z.str(info->qlabels[i]);
  for (int i = 0; i < info->nquants; i++) z.str(info->qunits[i]);
  z.add_iteration(iter->name, iter->n, iter->t, iter->time_units);
  for (int i = 0; i < info->nquants; i++) {
    const uint64_t c = info->np;
    z.add_dataset(info->quants[i], OSR_ZDF_FLOAT32, 1, &c, quants[i]);
  }
  if (z.save(fn)) return fail(OSR_ERR_ARG, "cannot write %s", fn);
  return OSR_OK;
}

}  // extern "C"
