// VASP HDF5 files: the structure / charge / POTCAR inputs of vaspout.h5
// and the vaspwave.h5 export of the solved bands.
//
// The HDF5 C library is loaded when first needed (dlopen of
// libhdf5.so.103, the 1.10 ABI; HALF_HDF5_LIBRARY overrides the name), so
// the library and its hosts have no link-time dependency on it and contexts
// built from CHGCAR/POTCAR text never touch it.  The subset of the API used
// here is declared below with the 1.10 types (hid_t 64-bit); its predefined
// datatypes are the library's H5T_*_g globals, valid after H5open().
//
// The file layout is the reference's (src/half_hdf5_bridge.c:20-130): the
// same group and dataset names, element types and shapes, so files written
// here read in the reference and in VASP, and vaspout.h5 files it reads
// read here.
#include "h5io.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>

namespace half {

namespace {

using hid_t = int64_t;
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;

constexpr hid_t kDefault = 0;  // H5P_DEFAULT, H5S_ALL, H5E_DEFAULT
constexpr unsigned kAccRdonly = 0x0000u, kAccTrunc = 0x0002u;
constexpr int kScalar = 0;     // H5S_SCALAR
constexpr int kSelectSet = 0;  // H5S_SELECT_SET

struct H5 {
  herr_t (*open)();
  herr_t (*eset_auto)(hid_t, void *, void *);
  hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
  hid_t (*fopen)(const char *, unsigned, hid_t);
  herr_t (*fclose)(hid_t);
  hid_t (*gcreate)(hid_t, const char *, hid_t, hid_t, hid_t);
  hid_t (*gopen)(hid_t, const char *, hid_t);
  herr_t (*gclose)(hid_t);
  hid_t (*screate)(int);
  hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
  herr_t (*sclose)(hid_t);
  int (*sndims)(hid_t);
  int (*sdims)(hid_t, hsize_t *, hsize_t *);
  herr_t (*sselect)(hid_t, int, const hsize_t *, const hsize_t *, const hsize_t *, const hsize_t *);
  hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
  hid_t (*dopen)(hid_t, const char *, hid_t);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
  herr_t (*dclose)(hid_t);
  hid_t (*dspace)(hid_t);
  hid_t (*dtype)(hid_t);
  hid_t (*tcopy)(hid_t);
  herr_t (*tset_size)(hid_t, size_t);
  herr_t (*tclose)(hid_t);
  size_t (*tsize)(hid_t);
  htri_t (*tvarstr)(hid_t);
  htri_t (*lexists)(hid_t, const char *, hid_t);
  herr_t (*free_memory)(void *);
  const hid_t *native_int, *native_double, *native_float, *c_s1;
};

H5 *g_h5 = nullptr;
std::string g_h5_error;
std::once_flag g_h5_once;

template <typename F>
bool sym(void *lib, const char *name, F &f) {
  f = reinterpret_cast<F>(dlsym(lib, name));
  return f != nullptr;
}

void load_h5() {
  std::vector<std::string> names;
  if (const char *env = std::getenv("HALF_HDF5_LIBRARY")) names.push_back(env);
  names.insert(names.end(), {"libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103"});
  void *lib = nullptr;
  for (auto &n : names)
    if ((lib = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
  if (!lib) {
    g_h5_error = "the HDF5 1.10 C library (libhdf5.so.103) is not available; set HALF_HDF5_LIBRARY";
    return;
  }
  static H5 h;
  bool ok = sym(lib, "H5open", h.open) && sym(lib, "H5Eset_auto2", h.eset_auto) &&
            sym(lib, "H5Fcreate", h.fcreate) && sym(lib, "H5Fopen", h.fopen) && sym(lib, "H5Fclose", h.fclose) &&
            sym(lib, "H5Gcreate2", h.gcreate) && sym(lib, "H5Gopen2", h.gopen) && sym(lib, "H5Gclose", h.gclose) &&
            sym(lib, "H5Screate", h.screate) && sym(lib, "H5Screate_simple", h.screate_simple) &&
            sym(lib, "H5Sclose", h.sclose) && sym(lib, "H5Sget_simple_extent_ndims", h.sndims) &&
            sym(lib, "H5Sget_simple_extent_dims", h.sdims) && sym(lib, "H5Sselect_hyperslab", h.sselect) &&
            sym(lib, "H5Dcreate2", h.dcreate) && sym(lib, "H5Dopen2", h.dopen) && sym(lib, "H5Dwrite", h.dwrite) &&
            sym(lib, "H5Dread", h.dread) && sym(lib, "H5Dclose", h.dclose) && sym(lib, "H5Dget_space", h.dspace) &&
            sym(lib, "H5Dget_type", h.dtype) && sym(lib, "H5Tcopy", h.tcopy) && sym(lib, "H5Tset_size", h.tset_size) &&
            sym(lib, "H5Tclose", h.tclose) && sym(lib, "H5Tget_size", h.tsize) &&
            sym(lib, "H5Tis_variable_str", h.tvarstr) && sym(lib, "H5Lexists", h.lexists) &&
            sym(lib, "H5free_memory", h.free_memory) && sym(lib, "H5T_NATIVE_INT_g", h.native_int) &&
            sym(lib, "H5T_NATIVE_DOUBLE_g", h.native_double) && sym(lib, "H5T_NATIVE_FLOAT_g", h.native_float) &&
            sym(lib, "H5T_C_S1_g", h.c_s1);
  if (!ok || h.open() < 0) {
    g_h5_error = "the HDF5 library lacks part of the 1.10 C API";
    return;
  }
  h.eset_auto(kDefault, nullptr, nullptr);  // errors come back as status codes only
  g_h5 = &h;
}

H5 &h5() {
  std::call_once(g_h5_once, load_h5);
  if (!g_h5) throw Error(kUnavailable, g_h5_error);
  return *g_h5;
}

// Closes an identifier with its class's close call at scope exit.
struct Id {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  Id(const Id &) = delete;
  Id &operator=(const Id &) = delete;
  ~Id() {
    if (id >= 0 && close) close(id);
  }
  operator hid_t() const { return id; }
};

void check(bool ok, const std::string &what) {
  if (!ok) throw Error(kIo, what);
}

// one dataset: scalar (rank 0) or simple
void put(H5 &h, hid_t parent, const char *name, hid_t type, int rank, const hsize_t *dims, const void *data) {
  Id space(rank == 0 ? h.screate(kScalar) : h.screate_simple(rank, dims, nullptr), h.sclose);
  check(space >= 0, std::string("HDF5: cannot describe ") + name);
  Id set(h.dcreate(parent, name, type, space, kDefault, kDefault, kDefault), h.dclose);
  check(set >= 0 && h.dwrite(set, type, kDefault, kDefault, kDefault, data) >= 0,
        std::string("HDF5: cannot write ") + name);
}

void put_string(H5 &h, hid_t parent, const char *name, const std::string &text) {
  Id type(h.tcopy(*h.c_s1), h.tclose);
  const std::string v = text.empty() ? std::string(" ") : text;
  check(type >= 0 && h.tset_size(type, v.size()) >= 0, "HDF5: string type");
  put(h, parent, name, type, 0, nullptr, v.data());
}

hid_t group(H5 &h, hid_t parent, const char *name) {
  hid_t g = h.gcreate(parent, name, kDefault, kDefault, kDefault);
  check(g >= 0, std::string("HDF5: cannot create group ") + name);
  return g;
}

bool exists(H5 &h, hid_t loc, const char *path) { return h.lexists(loc, path, kDefault) > 0; }

std::vector<hsize_t> dims_of(H5 &h, hid_t set) {
  Id space(h.dspace(set), h.sclose);
  check(space >= 0, "HDF5: dataset without a dataspace");
  const int r = h.sndims(space);
  check(r >= 0 && r <= 8, "HDF5: unsupported dataset rank");
  std::vector<hsize_t> d(r);
  if (r) h.sdims(space, d.data(), nullptr);
  return d;
}

template <typename T>
std::vector<T> get(H5 &h, hid_t loc, const char *name, hid_t type, size_t expect) {
  Id set(h.dopen(loc, name, kDefault), h.dclose);
  check(set >= 0, std::string("HDF5 input lacks ") + name);
  size_t count = 1;
  for (auto d : dims_of(h, set)) count *= d;
  check(count == expect, std::string("HDF5 input: unexpected size of ") + name);
  std::vector<T> v(count);
  check(h.dread(set, type, kDefault, kDefault, kDefault, v.data()) >= 0, std::string("HDF5: cannot read ") + name);
  return v;
}

std::string trim(const std::string &s) {
  size_t b = s.find_first_not_of(" \t\r\n");
  if (b == std::string::npos) return "";
  size_t e = s.find_last_not_of(" \t\r\n");
  return s.substr(b, e - b + 1);
}

}  // namespace

bool is_hdf5(const std::string &path) {
  static const unsigned char sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
  std::ifstream in(path, std::ios::binary);
  if (!in) return false;
  in.seekg(0, std::ios::end);
  const long long size = in.tellg();
  for (long long off = 0; off + 8 <= size; off = off ? off * 2 : 512) {
    unsigned char buf[8];
    in.seekg(off);
    if (!in.read(reinterpret_cast<char *>(buf), 8)) return false;
    if (!std::memcmp(buf, sig, 8)) return true;
  }
  return false;
}

void read_h5_chgcar(const std::string &path, Crystal &crystal, ChargeGrid &charge) {
  H5 &h = h5();
  Id file(h.fopen(path.c_str(), kAccRdonly, kDefault), h.fclose);
  check(file >= 0, "cannot open HDF5 input: " + path);
  const char *groups[] = {"structure/positions", "input/poscar", "results/positions"};
  const char *gname = nullptr;
  for (const char *g : groups) {
    // H5Lexists needs every intermediate link of the path to exist
    const std::string top = std::string(g).substr(0, std::string(g).find('/'));
    if (exists(h, file, top.c_str()) && exists(h, file, g)) {
      gname = g;
      break;
    }
  }
  check(gname != nullptr, "invalid vaspwave.h5 charge/structure input: no structure group");
  Id grp(h.gopen(file, gname, kDefault), h.gclose);
  check(grp >= 0, "HDF5: cannot open the structure group");
  // counts first: they size the rest
  std::vector<int> counts;
  {
    Id set(h.dopen(grp, "number_ion_types", kDefault), h.dclose);
    check(set >= 0, "HDF5 input lacks number_ion_types");
    auto d = dims_of(h, set);
    check(d.size() == 1 && d[0] >= 1 && d[0] <= 64, "HDF5 input: invalid number_ion_types");
    counts.resize(d[0]);
    check(h.dread(set, *h.native_int, kDefault, kDefault, kDefault, counts.data()) >= 0,
          "HDF5: cannot read number_ion_types");
  }
  const int ntypes = static_cast<int>(counts.size());
  long long nions = 0;
  for (int c : counts) {
    check(c >= 0, "HDF5 input: negative ion count");
    nions += c;
  }
  check(nions >= 1 && nions <= 100000, "HDF5 input: invalid ion count");
  crystal.system = "VASP HDF5 structure";
  if (exists(h, grp, "system")) {
    Id set(h.dopen(grp, "system", kDefault), h.dclose);
    Id type(h.dtype(set), h.tclose);
    check(set >= 0 && type >= 0, "HDF5: cannot open system");
    check(dims_of(h, set).empty(), "HDF5 input: system is not a scalar string");
    if (h.tvarstr(type) > 0) {
      char *v = nullptr;
      check(h.dread(set, type, kDefault, kDefault, kDefault, &v) >= 0, "HDF5: cannot read system");
      crystal.system = v ? trim(v) : "";
      if (v) h.free_memory(v);
    } else {
      std::vector<char> buf(h.tsize(type) + 1, '\0');
      check(h.dread(set, type, kDefault, kDefault, kDefault, buf.data()) >= 0, "HDF5: cannot read system");
      crystal.system = trim(buf.data());
    }
  }
  double scale = 1.0;
  if (exists(h, grp, "scale")) scale = get<double>(h, grp, "scale", *h.native_double, 1)[0];
  auto lat = get<double>(h, grp, "lattice_vectors", *h.native_double, 9);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) crystal.lattice[i][j] = lat[3 * i + j] * scale;
  crystal.update_geometry();
  auto pos = get<double>(h, grp, "position_ions", *h.native_double, size_t(nions) * 3);
  std::vector<char> species(size_t(16) * ntypes, ' ');
  {
    Id set(h.dopen(grp, "ion_types", kDefault), h.dclose);
    Id mem(h.tcopy(*h.c_s1), h.tclose);
    check(set >= 0 && mem >= 0 && h.tset_size(mem, 16) >= 0, "HDF5 input lacks ion_types");
    // the read fills one 16-byte slot per stored entry: the extent must be
    // exactly the type count the buffer was sized for
    const auto d = dims_of(h, set);
    check(d.size() == 1 && d[0] == hsize_t(ntypes), "HDF5 input: ion_types does not hold number_ion_types entries");
    check(h.dread(set, mem, kDefault, kDefault, kDefault, species.data()) >= 0, "HDF5: cannot read ion_types");
  }
  const int direct = get<int>(h, grp, "direct_coordinates", *h.native_int, 1)[0];
  crystal.species.clear();
  for (int t = 0; t < ntypes; ++t) {
    std::string s(species.data() + 16 * t, 16);
    s = s.substr(0, s.find('\0'));
    crystal.species.push_back(trim(s));
  }
  crystal.counts = counts;
  crystal.positions = pos;
  if (direct == 0) {  // Cartesian: fractional f_i = r . b_i / (2 pi), b_i the reciprocal rows
    for (long long a = 0; a < nions; ++a) {
      const double *r = &pos[3 * a];
      for (int i = 0; i < 3; ++i)
        crystal.positions[3 * a + i] =
            (r[0] * crystal.reciprocal[i][0] + r[1] * crystal.reciprocal[i][1] + r[2] * crystal.reciprocal[i][2]) /
            (2.0 * kPi);
    }
  }
  // charge: [spin][n3][n2][n1] (first spin component) or [n3][n2][n1]
  auto grid = get<int>(h, file, "charge/grid", *h.native_int, 3);
  for (int d = 0; d < 3; ++d) check(grid[d] >= 1 && grid[d] <= 4096, "HDF5 input: invalid charge grid");
  Id set(h.dopen(file, "charge/charge", kDefault), h.dclose);
  check(set >= 0, "HDF5 input lacks charge/charge");
  auto d = dims_of(h, set);
  const hsize_t n1 = grid[0], n2 = grid[1], n3 = grid[2];
  charge.values.assign(size_t(n1) * n2 * n3, 0.0);
  for (int i = 0; i < 3; ++i) charge.shape[i] = grid[i];
  if (d.size() == 4) {
    check(d[0] >= 1 && d[1] == n3 && d[2] == n2 && d[3] == n1, "HDF5 input: charge shape differs from its grid");
    Id space(h.dspace(set), h.sclose);
    const hsize_t start[4] = {0, 0, 0, 0}, count[4] = {1, n3, n2, n1};
    Id mem(h.screate_simple(4, count, nullptr), h.sclose);
    check(space >= 0 && mem >= 0 && h.sselect(space, kSelectSet, start, nullptr, count, nullptr) >= 0,
          "HDF5: cannot select the first spin component");
    check(h.dread(set, *h.native_double, mem, space, kDefault, charge.values.data()) >= 0,
          "HDF5: cannot read charge/charge");
  } else {
    check(d.size() == 3 && d[0] == n3 && d[1] == n2 && d[2] == n1, "HDF5 input: charge shape differs from its grid");
    check(h.dread(set, *h.native_double, kDefault, kDefault, kDefault, charge.values.data()) >= 0,
          "HDF5: cannot read charge/charge");
  }
}

std::string read_h5_potcar(const std::string &path) {
  H5 &h = h5();
  Id file(h.fopen(path.c_str(), kAccRdonly, kDefault), h.fclose);
  check(file >= 0, "cannot open HDF5 input: " + path);
  check(exists(h, file, "input") && exists(h, file, "input/potcar") && exists(h, file, "input/potcar/content"),
        "HDF5 input has no /input/potcar/content");
  Id set(h.dopen(file, "input/potcar/content", kDefault), h.dclose);
  Id type(h.dtype(set), h.tclose);
  check(set >= 0 && type >= 0, "HDF5: cannot open /input/potcar/content");
  if (h.tvarstr(type) > 0) {
    char *v = nullptr;
    check(h.dread(set, type, kDefault, kDefault, kDefault, &v) >= 0 && v, "HDF5: cannot read the POTCAR");
    std::string out(v);
    h.free_memory(v);
    return out;
  }
  std::vector<char> buf(h.tsize(type) + 1, '\0');
  check(h.dread(set, type, kDefault, kDefault, kDefault, buf.data()) >= 0, "HDF5: cannot read the POTCAR");
  return std::string(buf.data());
}

std::vector<int64_t> vasp_order(const Basis &b) {
  const int64_t n = b.npw();
  const int n1 = b.shape[0], n2 = b.shape[1], n3 = b.shape[2];
  // C-order index i1 n2 n3 + i2 n3 + i3  ->  key i3 n2 n1 + i2 n1 + i1
  std::vector<std::pair<int64_t, int64_t>> key(n);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t c = b.fft_index[i];
    const int64_t i1 = c / (int64_t(n2) * n3), i2 = (c / n3) % n2, i3 = c % n3;
    key[i] = {(i3 * n2 + i2) * n1 + i1, i};
  }
  std::sort(key.begin(), key.end());
  std::vector<int64_t> order(n);
  for (int64_t i = 0; i < n; ++i) order[i] = key[i].second;
  return order;
}

void write_vaspwave(const std::string &path, const WaveExport &w) {
  H5 &h = h5();
  const Crystal &c = *w.crystal;
  Id file(h.fcreate(path.c_str(), kAccTrunc, kDefault, kDefault), h.fclose);
  check(file >= 0, "cannot create " + path);
  {
    Id v(group(h, file, "version"), h.gclose);
    const int major = 6, minor = 6, patch = 0;
    put(h, v, "major", *h.native_int, 0, nullptr, &major);
    put(h, v, "minor", *h.native_int, 0, nullptr, &minor);
    put(h, v, "patch", *h.native_int, 0, nullptr, &patch);
  }
  const int ntypes = static_cast<int>(c.species.size()), nions = c.nions();
  {
    Id s(group(h, file, "structure"), h.gclose);
    Id p(group(h, s, "positions"), h.gclose);
    put_string(h, p, "system", w.system);
    const double one = 1.0;
    put(h, p, "scale", *h.native_double, 0, nullptr, &one);
    double lat[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) lat[3 * i + j] = c.lattice[i][j];
    const hsize_t d33[2] = {3, 3};
    put(h, p, "lattice_vectors", *h.native_double, 2, d33, lat);
    std::vector<char> species(size_t(16) * ntypes, ' ');
    for (int t = 0; t < ntypes; ++t)
      std::memcpy(species.data() + 16 * t, c.species[t].data(), std::min<size_t>(16, c.species[t].size()));
    Id st(h.tcopy(*h.c_s1), h.tclose);
    check(st >= 0 && h.tset_size(st, 16) >= 0, "HDF5: string type");
    const hsize_t dt[1] = {hsize_t(ntypes)};
    put(h, p, "ion_types", st, 1, dt, species.data());
    put(h, p, "number_ion_types", *h.native_int, 1, dt, c.counts.data());
    const hsize_t dp[2] = {hsize_t(nions), 3};
    put(h, p, "position_ions", *h.native_double, 2, dp, c.positions.data());
    const int direct = 1;
    put(h, p, "direct_coordinates", *h.native_int, 0, nullptr, &direct);
  }
  {
    Id g(group(h, file, "charge"), h.gclose);
    const hsize_t d3[1] = {3};
    put(h, g, "grid", *h.native_int, 1, d3, w.charge->shape);
    const hsize_t d4[4] = {1, hsize_t(w.charge->shape[2]), hsize_t(w.charge->shape[1]), hsize_t(w.charge->shape[0])};
    put(h, g, "charge", *h.native_double, 4, d4, w.charge->values.data());
  }
  Id wave(group(h, file, "wave"), h.gclose);
  double lat[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) lat[3 * i + j] = c.lattice[i][j];
  const hsize_t d33[2] = {3, 3};
  put(h, wave, "amat", *h.native_double, 2, d33, lat);
  const double rdum = 8.0 * *std::max_element(w.npw.begin(), w.npw.end());
  const double rispin = 1.0, rnb = w.nb, rnk = w.nk;
  put(h, wave, "efermi", *h.native_double, 0, nullptr, &w.fermi);
  put(h, wave, "enmax", *h.native_double, 0, nullptr, &w.encut);
  put(h, wave, "rdum", *h.native_double, 0, nullptr, &rdum);
  put(h, wave, "rispin", *h.native_double, 0, nullptr, &rispin);
  put(h, wave, "rnb_tot", *h.native_double, 0, nullptr, &rnb);
  put(h, wave, "rnkpts", *h.native_double, 0, nullptr, &rnk);
  Id spin(group(h, wave, "spin_1"), h.gclose);
  std::vector<double> celtot(size_t(2) * w.nb), fertot(w.nb);
  for (int k = 0; k < w.nk; ++k) {
    const std::string name = "kpoint_" + std::to_string(k + 1);
    Id pt(group(h, spin, name.c_str()), h.gclose);
    for (int b = 0; b < w.nb; ++b) {
      celtot[2 * b] = w.eigenvalues[size_t(k) * w.nb + b];
      celtot[2 * b + 1] = 0.0;
      fertot[b] = 0.5 * w.occupations[size_t(k) * w.nb + b];
    }
    const hsize_t dc[2] = {hsize_t(w.nb), 2}, df[1] = {hsize_t(w.nb)}, dk[1] = {3};
    put(h, pt, "celtot", *h.native_double, 2, dc, celtot.data());
    put(h, pt, "fertot", *h.native_double, 1, df, fertot.data());
    put(h, pt, "num_planewaves", *h.native_int, 0, nullptr, &w.npw[k]);
    put(h, pt, "vkpt", *h.native_double, 1, dk, w.kpoints + 3 * k);
    const hsize_t dw[3] = {hsize_t(w.nb), hsize_t(w.npw[k]), 2};
    put(h, pt, "wave", *h.native_float, 3, dw, w.coefficients[k].data());
  }
}

}  // namespace half
