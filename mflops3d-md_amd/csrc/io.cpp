// Field files in the reference's format: one 3-D variable per call in a
// NetCDF classic file (write_var3d / read_var3d / get_var3d_info,
// src/class_io/class_io.f90:23-223, through write_field / read_field,
// src/class_field/class_field.f90:528-621).
//
// The reference links the NetCDF library; this is a self-contained reader and
// writer of the classic format (CDF-1, and CDF-2 "64-bit offset" when a file
// exceeds 2 GiB), the format nf90_create(..., nf90_clobber) produces:
//   header = 'C' 'D' 'F' version, numrecs, dim_list, gatt_list, var_list
//   big-endian 32-bit counts, names padded to 4 bytes, then each variable's
//   data at its `begin` offset, big-endian, padded to 4 bytes.
// A Fortran variable f(nx,ny,nz) defined on dimensions (x, y, z) is stored
// with the dimension list reversed (z, y, x), x varying fastest: the file's
// data order is the Fortran memory order.  Writing to an existing file keeps
// its other variables, checks that dimensions of the same name have the same
// length (the reference's "wrong dimensions" stop) and adds or replaces the
// variable.  Attributes are skipped on read and none are written (the
// reference writes none).  Record (unlimited) dimensions are not supported.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"
#include "mflops/mf.h"

namespace mf {
namespace nc {

enum { NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12 };
enum { NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };

struct Var {
  std::string name;
  std::vector<int> dimids;  // file order (slowest first)
  int type = NC_DOUBLE;
  std::vector<unsigned char> data;  // big-endian, unpadded
};

struct File {
  std::vector<std::string> dim_names;
  std::vector<uint64_t> dim_lens;
  std::vector<Var> vars;
};

inline size_t pad4(size_t n) { return (n + 3) & ~size_t(3); }

inline int type_size(int t) {
  switch (t) {
    case NC_BYTE: case NC_CHAR: return 1;
    case NC_SHORT: return 2;
    case NC_INT: case NC_FLOAT: return 4;
    case NC_DOUBLE: return 8;
  }
  return 0;
}

struct Reader {
  const std::vector<unsigned char> &b;
  size_t p = 0;
  bool ok = true;
  explicit Reader(const std::vector<unsigned char> &buf) : b(buf) {}
  uint32_t u32() {
    if (p + 4 > b.size()) { ok = false; return 0; }
    const uint32_t v = (uint32_t)b[p] << 24 | (uint32_t)b[p + 1] << 16 | (uint32_t)b[p + 2] << 8 | b[p + 3];
    p += 4;
    return v;
  }
  uint64_t u64() {
    const uint64_t hi = u32();
    return hi << 32 | u32();
  }
  std::string name() {
    const uint32_t n = u32();
    if (!ok || p + pad4(n) > b.size()) { ok = false; return ""; }
    std::string s(reinterpret_cast<const char *>(&b[p]), n);
    p += pad4(n);
    return s;
  }
  void skip_atts() {
    const uint32_t tag = u32(), n = u32();
    if (tag == 0 && n == 0) return;
    if (tag != NC_ATTRIBUTE) { ok = false; return; }
    for (uint32_t a = 0; a < n && ok; ++a) {
      name();
      const int t = (int)u32();
      const uint32_t cnt = u32();
      const size_t bytes = pad4((size_t)cnt * type_size(t));
      if (type_size(t) == 0 || p + bytes > b.size()) { ok = false; return; }
      p += bytes;
    }
  }
};

bool read_file(const std::string &path, File &f, std::string &err) {
  FILE *fp = std::fopen(path.c_str(), "rb");
  if (!fp) { err = "cannot open " + path; return false; }
  std::vector<unsigned char> buf;
  std::fseek(fp, 0, SEEK_END);
  const long sz = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  if (sz > 0) {
    buf.resize((size_t)sz);
    if (std::fread(buf.data(), 1, buf.size(), fp) != buf.size()) { std::fclose(fp); err = "short read of " + path; return false; }
  }
  std::fclose(fp);
  if (buf.size() < 8 || buf[0] != 'C' || buf[1] != 'D' || buf[2] != 'F' || (buf[3] != 1 && buf[3] != 2)) {
    err = path + " is not a NetCDF classic (CDF-1/CDF-2) file";
    return false;
  }
  const int version = buf[3];
  Reader r(buf);
  r.p = 4;
  r.u32();  // numrecs
  uint32_t tag = r.u32(), n = r.u32();
  if (tag == NC_DIMENSION) {
    for (uint32_t d = 0; d < n && r.ok; ++d) {
      f.dim_names.push_back(r.name());
      f.dim_lens.push_back(r.u32());
      if (f.dim_lens.back() == 0) { err = path + ": record dimensions are not supported"; return false; }
    }
  } else if (tag != 0 || n != 0) {
    err = path + ": bad dimension list";
    return false;
  }
  r.skip_atts();  // global attributes
  tag = r.u32();
  n = r.u32();
  if (tag != NC_VARIABLE && (tag != 0 || n != 0)) { err = path + ": bad variable list"; return false; }
  std::vector<uint64_t> begins;
  for (uint32_t v = 0; v < n && r.ok; ++v) {
    Var var;
    var.name = r.name();
    const uint32_t nd = r.u32();
    for (uint32_t d = 0; d < nd && r.ok; ++d) {
      const uint32_t id = r.u32();
      if (id >= f.dim_names.size()) { err = path + ": bad dimension id"; return false; }
      var.dimids.push_back((int)id);
    }
    r.skip_atts();
    var.type = (int)r.u32();
    r.u32();  // vsize
    begins.push_back(version == 1 ? r.u32() : r.u64());
    if (type_size(var.type) == 0) { err = path + ": unknown variable type"; return false; }
    f.vars.push_back(var);
  }
  if (!r.ok) { err = path + ": truncated header"; return false; }
  for (size_t v = 0; v < f.vars.size(); ++v) {
    uint64_t cnt = 1;
    for (int id : f.vars[v].dimids) cnt *= f.dim_lens[id];
    const uint64_t bytes = cnt * type_size(f.vars[v].type);
    if (begins[v] + bytes > buf.size()) { err = path + ": truncated data of " + f.vars[v].name; return false; }
    f.vars[v].data.assign(buf.begin() + begins[v], buf.begin() + begins[v] + bytes);
  }
  return true;
}

struct Writer {
  std::vector<unsigned char> b;
  void u32(uint32_t v) {
    b.push_back(v >> 24); b.push_back((v >> 16) & 255); b.push_back((v >> 8) & 255); b.push_back(v & 255);
  }
  void u64(uint64_t v) { u32((uint32_t)(v >> 32)); u32((uint32_t)v); }
  void name(const std::string &s) {
    u32((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    b.resize(pad4(b.size()), 0);
  }
};

bool write_file(const std::string &path, const File &f, std::string &err) {
  auto header = [&](int version, const std::vector<uint64_t> &begins) {
    Writer w;
    w.b = {'C', 'D', 'F', (unsigned char)version};
    w.u32(0);  // numrecs
    if (f.dim_names.empty()) { w.u32(0); w.u32(0); }
    else {
      w.u32(NC_DIMENSION);
      w.u32((uint32_t)f.dim_names.size());
      for (size_t d = 0; d < f.dim_names.size(); ++d) { w.name(f.dim_names[d]); w.u32((uint32_t)f.dim_lens[d]); }
    }
    w.u32(0); w.u32(0);  // no global attributes
    if (f.vars.empty()) { w.u32(0); w.u32(0); }
    else {
      w.u32(NC_VARIABLE);
      w.u32((uint32_t)f.vars.size());
      for (size_t v = 0; v < f.vars.size(); ++v) {
        const Var &var = f.vars[v];
        w.name(var.name);
        w.u32((uint32_t)var.dimids.size());
        for (int id : var.dimids) w.u32((uint32_t)id);
        w.u32(0); w.u32(0);  // no attributes
        w.u32((uint32_t)var.type);
        const uint64_t vs = pad4(var.data.size());
        w.u32(vs > 0xffffffffull ? 0xffffffffu : (uint32_t)vs);
        if (version == 1) w.u32((uint32_t)begins[v]); else w.u64(begins[v]);
      }
    }
    return w.b;
  };
  // offsets: header size does not depend on the offset values
  std::vector<uint64_t> begins(f.vars.size(), 0);
  int version = 1;
  size_t hsize = header(version, begins).size();
  uint64_t off = hsize;
  for (size_t v = 0; v < f.vars.size(); ++v) { begins[v] = off; off += pad4(f.vars[v].data.size()); }
  if (off > 0x7fffffffull) {
    version = 2;
    hsize = header(version, begins).size();
    off = hsize;
    for (size_t v = 0; v < f.vars.size(); ++v) { begins[v] = off; off += pad4(f.vars[v].data.size()); }
  }
  std::vector<unsigned char> h = header(version, begins);
  FILE *fp = std::fopen(path.c_str(), "wb");
  if (!fp) { err = "cannot create " + path; return false; }
  bool good = std::fwrite(h.data(), 1, h.size(), fp) == h.size();
  const unsigned char zeros[4] = {0, 0, 0, 0};
  for (const Var &var : f.vars) {
    if (!good) break;
    good = std::fwrite(var.data.data(), 1, var.data.size(), fp) == var.data.size();
    const size_t pad = pad4(var.data.size()) - var.data.size();
    if (good && pad) good = std::fwrite(zeros, 1, pad, fp) == pad;
  }
  good = (std::fclose(fp) == 0) && good;
  if (!good) err = "write error on " + path;
  return good;
}

void put_be(unsigned char *d, double x, int type) {
  if (type == NC_DOUBLE) {
    uint64_t u;
    std::memcpy(&u, &x, 8);
    for (int q = 7; q >= 0; --q) { d[q] = u & 255; u >>= 8; }
  } else {
    const float fx = (float)x;
    uint32_t u;
    std::memcpy(&u, &fx, 4);
    for (int q = 3; q >= 0; --q) { d[q] = u & 255; u >>= 8; }
  }
}

double get_be(const unsigned char *s, int type) {
  switch (type) {
    case NC_DOUBLE: {
      uint64_t u = 0;
      for (int q = 0; q < 8; ++q) u = u << 8 | s[q];
      double x;
      std::memcpy(&x, &u, 8);
      return x;
    }
    case NC_FLOAT: {
      uint32_t u = 0;
      for (int q = 0; q < 4; ++q) u = u << 8 | s[q];
      float x;
      std::memcpy(&x, &u, 4);
      return x;
    }
    case NC_INT: return (double)(int32_t)((uint32_t)s[0] << 24 | (uint32_t)s[1] << 16 | (uint32_t)s[2] << 8 | s[3]);
    case NC_SHORT: return (double)(int16_t)((uint16_t)s[0] << 8 | s[1]);
    default: return (double)(signed char)s[0];
  }
}

const Var *find(const File &f, const char *name) {
  for (const Var &v : f.vars)
    if (v.name == name) return &v;
  return nullptr;
}

}  // namespace nc

int nc_write_field(const char *path, const char *var_name, const char *const dim_names[3], int nx,
                   int ny, int nz, const double *f, bool dbl, std::string &err) {
  nc::File file;
  FILE *probe = std::fopen(path, "rb");
  if (probe) {
    std::fclose(probe);
    if (!nc::read_file(path, file, err)) return MF_ERR_ARG;
  }
  const int n3[3] = {nx, ny, nz};
  int ids[3];
  for (int a = 0; a < 3; ++a) {
    int id = -1;
    for (size_t d = 0; d < file.dim_names.size(); ++d)
      if (file.dim_names[d] == dim_names[a]) id = (int)d;
    if (id < 0) {
      file.dim_names.push_back(dim_names[a]);
      file.dim_lens.push_back((uint64_t)n3[a]);
      id = (int)file.dim_names.size() - 1;
    } else if (file.dim_lens[id] != (uint64_t)n3[a]) {
      err = std::string(path) + ": wrong dimensions (" + dim_names[a] + ")";
      return MF_ERR_ARG;
    }
    ids[a] = id;
  }
  nc::Var *var = nullptr;
  for (nc::Var &v : file.vars)
    if (v.name == var_name) var = &v;
  if (var) {  // an existing variable keeps its definition
    if (var->dimids != std::vector<int>{ids[2], ids[1], ids[0]}) {
      err = std::string(path) + ": variable " + var_name + " has other dimensions";
      return MF_ERR_ARG;
    }
  } else {
    file.vars.push_back(nc::Var());
    var = &file.vars.back();
    var->name = var_name;
    var->dimids = {ids[2], ids[1], ids[0]};
    var->type = dbl ? nc::NC_DOUBLE : nc::NC_FLOAT;
  }
  if (var->type != nc::NC_DOUBLE && var->type != nc::NC_FLOAT) {
    err = std::string(path) + ": variable " + var_name + " is not floating point";
    return MF_ERR_ARG;
  }
  const size_t N = (size_t)nx * ny * nz, ts = nc::type_size(var->type);
  var->data.resize(N * ts);
  for (size_t q = 0; q < N; ++q) nc::put_be(&var->data[q * ts], f[q], var->type);
  return nc::write_file(path, file, err) ? MF_OK : MF_ERR_ARG;
}

int nc_read_field(const char *path, const char *var_name, int dims[3], std::string dim_names[3],
                  bool *is_double, double *f, std::string &err) {
  nc::File file;
  if (!nc::read_file(path, file, err)) return MF_ERR_ARG;
  const nc::Var *v = nc::find(file, var_name);
  if (!v) { err = std::string(path) + ": no variable " + var_name; return MF_ERR_ARG; }
  if (v->dimids.size() != 3) { err = std::string(path) + ": " + var_name + " is not 3-D"; return MF_ERR_ARG; }
  int d[3];
  for (int a = 0; a < 3; ++a) {  // file order (z, y, x)
    d[a] = (int)file.dim_lens[v->dimids[2 - a]];
    if (dim_names) dim_names[a] = file.dim_names[v->dimids[2 - a]];
  }
  if (f && dims && (dims[0] != d[0] || dims[1] != d[1] || dims[2] != d[2])) {
    err = std::string(path) + ": " + var_name + " has other dimensions";
    return MF_ERR_ARG;
  }
  if (dims) { dims[0] = d[0]; dims[1] = d[1]; dims[2] = d[2]; }
  if (is_double) *is_double = v->type == nc::NC_DOUBLE;
  if (f) {
    const size_t N = (size_t)d[0] * d[1] * d[2], ts = nc::type_size(v->type);
    for (size_t q = 0; q < N; ++q) f[q] = nc::get_be(&v->data[q * ts], v->type);
  }
  return MF_OK;
}

}  // namespace mf

extern "C" int mf_io_write_field(const char *path, const char *var_name,
                                 const char *const dim_names[3], int nx, int ny, int nz,
                                 const double *f, int dbl) {
  static const char *const def[3] = {"resolution_x", "resolution_y", "resolution_z"};
  if (!path || !var_name || !f || nx < 1 || ny < 1 || nz < 1)
    return mf::fail(MF_ERR_ARG, "mf_io_write_field: invalid argument");
  std::string err;
  const int rc = mf::nc_write_field(path, var_name, dim_names ? dim_names : def, nx, ny, nz, f,
                                    dbl != 0, err);
  return rc == MF_OK ? MF_OK : mf::fail(rc, "mf_io_write_field: " + err);
}

extern "C" int mf_io_field_info(const char *path, const char *var_name, int dims[3],
                                char *dim_names, int name_len, int *is_double) {
  if (!path || !var_name || !dims) return mf::fail(MF_ERR_ARG, "mf_io_field_info: invalid argument");
  std::string err, names[3];
  bool dbl = false;
  const int rc = mf::nc_read_field(path, var_name, dims, names, &dbl, nullptr, err);
  if (rc != MF_OK) return mf::fail(rc, "mf_io_field_info: " + err);
  if (dim_names && name_len > 0)
    for (int a = 0; a < 3; ++a) {
      std::strncpy(dim_names + a * name_len, names[a].c_str(), name_len - 1);
      dim_names[a * name_len + name_len - 1] = 0;
    }
  if (is_double) *is_double = dbl;
  return MF_OK;
}

extern "C" int mf_io_read_field(const char *path, const char *var_name, int nx, int ny, int nz,
                                double *f) {
  if (!path || !var_name || !f) return mf::fail(MF_ERR_ARG, "mf_io_read_field: invalid argument");
  int dims[3] = {nx, ny, nz};
  std::string err;
  const int rc = mf::nc_read_field(path, var_name, dims, nullptr, nullptr, f, err);
  return rc == MF_OK ? MF_OK : mf::fail(rc, "mf_io_read_field: " + err);
}
