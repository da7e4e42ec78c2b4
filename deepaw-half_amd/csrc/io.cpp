// Host-side input handling: CHGCAR and POTCAR readers, crystal geometry,
// plane-wave basis enumeration, radial spline set-up and the lm expansion of
// DION/QPAW.  These run once per context (readers) or once per k point
// (basis) on the host; all per-band work is on the device.
#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "common.hpp"
#include "h5io.hpp"

namespace half {

namespace {

std::string lower(std::string s) {
  for (auto &c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return s;
}

bool blank(const std::string &s) {
  return s.find_first_not_of(" \t\r") == std::string::npos;
}

std::vector<std::string> words(const std::string &line) {
  std::vector<std::string> out;
  std::string cur;
  for (char c : line) {
    if (c == ' ' || c == '\t' || c == '\r' || c == ',') {
      if (!cur.empty()) out.push_back(cur), cur.clear();
    } else {
      cur.push_back(c);
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

bool parse_double(const std::string &tok, double &v) {
  // Fortran list-directed reals may use D exponents.
  std::string t = tok;
  for (auto &c : t)
    if (c == 'D' || c == 'd') c = 'E';
  char *end = nullptr;
  v = std::strtod(t.c_str(), &end);
  return end && end != t.c_str() && *end == '\0';
}

bool parse_int(const std::string &tok, long &v) {
  char *end = nullptr;
  v = std::strtol(tok.c_str(), &end, 10);
  return end && end != tok.c_str() && *end == '\0';
}

// Line cursor over a whole file with list-directed numeric reads that may
// span records, as Fortran's read(unit,*) array does.
struct Cursor {
  std::vector<std::string> lines;
  size_t pos = 0;
  std::string what;
  explicit Cursor(const std::string &path, const std::string &kind) : what(kind) {
    std::ifstream in(path);
    if (!in) throw Error(kIo, "cannot open " + kind + ": " + path);
    std::string l;
    while (std::getline(in, l)) lines.push_back(l);
  }
  // over text already in memory
  Cursor(const std::string &text, const std::string &kind, bool) : what(kind) {
    std::istringstream in(text);
    std::string l;
    while (std::getline(in, l)) lines.push_back(l);
  }
  bool eof() const { return pos >= lines.size(); }
  const std::string &line() {
    if (eof()) throw Error(kIo, "unexpected end of " + what);
    return lines[pos++];
  }
  const std::string &next_nonblank() {
    while (!eof() && blank(lines[pos])) ++pos;
    return line();
  }
  std::vector<double> reals(size_t n, const char *field) {
    std::vector<double> v;
    v.reserve(n);
    while (v.size() < n) {
      if (eof()) throw Error(kIo, std::string("failed to read ") + field + " in " + what);
      for (auto &w : words(lines[pos++])) {
        double x;
        if (!parse_double(w, x)) throw Error(kIo, std::string("invalid number in ") + field);
        v.push_back(x);
        if (v.size() == n) break;
      }
    }
    return v;
  }
};

// Value of "TAG = value" inside a POTCAR PSCTR line
// (reference: src/half_potcar.F90:220-237).
bool tag_text(const std::string &line, const std::string &tag, std::string &out) {
  size_t at = line.find(tag);
  if (at == std::string::npos) return false;
  size_t eq = line.find('=', at);
  if (eq == std::string::npos) return false;
  std::string part = line.substr(eq + 1);
  size_t b = part.find_first_not_of(" \t");
  part = b == std::string::npos ? "" : part.substr(b);
  out = part;
  return true;
}

void real_tag(const std::string &line, const std::string &tag, double &value) {
  std::string part;
  if (!tag_text(line, tag, part)) return;
  size_t semi = part.find(';');
  if (semi != std::string::npos) part = part.substr(0, semi);
  auto w = words(part);
  double v;
  if (!w.empty() && parse_double(w[0], v)) value = v;
}

void string_tag(const std::string &line, const std::string &tag, std::string &value) {
  std::string part;
  if (!tag_text(line, tag, part)) return;
  size_t last = part.find_first_of(" :;\t");
  value = part.substr(0, last);
}

double det3(const double a[3][3]) {
  return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
         a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
         a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}

// reference: src/half_math.F90:19-38
void inverse3(const double a[3][3], double inv[3][3]) {
  double d = det3(a);
  if (std::fabs(d) < 1e-18) throw Error(kInvalidArgument, "singular lattice");
  inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / d;
  inv[0][1] = -(a[0][1] * a[2][2] - a[0][2] * a[2][1]) / d;
  inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / d;
  inv[1][0] = -(a[1][0] * a[2][2] - a[1][2] * a[2][0]) / d;
  inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / d;
  inv[1][2] = -(a[0][0] * a[1][2] - a[0][2] * a[1][0]) / d;
  inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / d;
  inv[2][1] = -(a[0][0] * a[2][1] - a[0][1] * a[2][0]) / d;
  inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / d;
}

}  // namespace

void Crystal::update_geometry() {
  volume = std::fabs(det3(lattice));
  double inv[3][3];
  inverse3(lattice, inv);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) reciprocal[i][j] = 2.0 * kPi * inv[j][i];
}

void read_chgcar(const std::string &path, Crystal &crystal, ChargeGrid &charge) {
  // a VASP HDF5 file in place of the CHGCAR (reference: src/half_chgcar.F90:25)
  if (is_hdf5(path)) {
    read_h5_chgcar(path, crystal, charge);
    return;
  }
  Cursor c(path, "CHGCAR");
  {
    const std::string &name = c.line();
    const size_t b = name.find_first_not_of(" \t"), e = name.find_last_not_of(" \t\r");
    crystal.system = b == std::string::npos ? "" : name.substr(b, e - b + 1);
  }
  auto sw = words(c.line());
  double scale;
  if (sw.empty() || !parse_double(sw[0], scale)) throw Error(kIo, "failed to read lattice scale");
  if (scale <= 0) throw Error(kIo, "only positive POSCAR scale is supported");
  for (int i = 0; i < 3; ++i) {
    auto w = words(c.line());
    if (w.size() < 3) throw Error(kIo, "failed to read lattice vector");
    for (int j = 0; j < 3; ++j) {
      if (!parse_double(w[j], crystal.lattice[i][j])) throw Error(kIo, "invalid lattice vector");
      crystal.lattice[i][j] *= scale;
    }
  }
  crystal.update_geometry();
  auto w = words(c.line());
  if (w.empty()) throw Error(kIo, "empty species/count line");
  bool all_int = true;
  long iv;
  for (auto &x : w) all_int = all_int && parse_int(x, iv);
  crystal.species.clear();
  crystal.counts.clear();
  if (all_int) {
    for (size_t i = 0; i < w.size(); ++i) {
      crystal.species.push_back("Type" + std::to_string(i + 1));
      parse_int(w[i], iv);
      crystal.counts.push_back(static_cast<int>(iv));
    }
  } else {
    crystal.species = w;
    auto cw = words(c.line());
    if (cw.size() != w.size()) throw Error(kIo, "invalid ion counts");
    for (auto &x : cw) {
      if (!parse_int(x, iv)) throw Error(kIo, "invalid ion counts");
      crystal.counts.push_back(static_cast<int>(iv));
    }
  }
  std::string mode = c.line();
  size_t b = mode.find_first_not_of(" \t");
  char m0 = b == std::string::npos ? ' ' : static_cast<char>(std::tolower(mode[b]));
  if (mode.size() && std::tolower(static_cast<unsigned char>(mode[0])) == 's') {
    mode = c.line();
    b = mode.find_first_not_of(" \t");
    m0 = b == std::string::npos ? ' ' : static_cast<char>(std::tolower(mode[b]));
  }
  // the reference inspects the first column of the line
  char first = mode.empty() ? ' ' : static_cast<char>(std::tolower(static_cast<unsigned char>(mode[0])));
  (void)m0;
  bool cartesian = first == 'c' || first == 'k';
  int nions = crystal.nions();
  crystal.positions.assign(size_t(nions) * 3, 0.0);
  for (int i = 0; i < nions; ++i) {
    auto pw = words(c.line());
    if (pw.size() < 3) throw Error(kIo, "failed to read atomic position values");
    for (int j = 0; j < 3; ++j)
      if (!parse_double(pw[j], crystal.positions[3 * i + j])) throw Error(kIo, "invalid atomic position");
  }
  if (cartesian) {
    double inv[3][3];
    inverse3(crystal.lattice, inv);
    for (int i = 0; i < nions; ++i) {
      double r[3] = {crystal.positions[3 * i], crystal.positions[3 * i + 1], crystal.positions[3 * i + 2]};
      for (int j = 0; j < 3; ++j)
        crystal.positions[3 * i + j] = r[0] * inv[0][j] + r[1] * inv[1][j] + r[2] * inv[2][j];
    }
  }
  std::string gl = c.next_nonblank();
  auto gw = words(gl);
  if (gw.size() < 3) throw Error(kIo, "failed to read FFT grid dimensions");
  for (int i = 0; i < 3; ++i) {
    if (!parse_int(gw[i], iv) || iv <= 0) throw Error(kIo, "invalid FFT grid dimensions");
    charge.shape[i] = static_cast<int>(iv);
  }
  charge.values = c.reals(static_cast<size_t>(charge.size()), "smooth charge grid");
}

namespace {

void read_dataset(Cursor &c, const std::string &header, Dataset &p) {
  auto hw = words(header);
  p.element = hw.size() > 1 ? hw[1] : "";
  p.zval = c.reals(1, "POTCAR ZVAL")[0];
  // PSCTR block (reference: src/half_potcar.F90:65-78)
  while (true) {
    const std::string &l = c.line();
    if (lower(l).find("end of psctr") != std::string::npos) break;
    real_tag(l, "EATOM", p.eatom);
    real_tag(l, "ZVAL", p.zval);
    string_tag(l, "VRHFIN", p.element);
  }
  if (lower(c.next_nonblank()).find("local part") == std::string::npos)
    throw Error(kIo, "expected POTCAR local part");
  // local part (reference: src/half_potcar.F90:80-107)
  p.psp_gmax = c.reals(1, "POTCAR local grid")[0];
  p.psp_local = c.reals(1000, "POTCAR local potential");
  while (true) {
    std::string l = lower(c.next_nonblank());
    if (l.find("gradient") != std::string::npos) {
      c.line();
    } else if (l.find("core charge-density (partial)") != std::string::npos) {
      p.pspcor = c.reals(1000, "POTCAR partial core charge");
      p.has_core = true;
    } else if (l.find("kinetic energy density") != std::string::npos) {
      c.reals(1000, "POTCAR kinetic density");
    } else if (l.find("atomic pseudo charge-density") != std::string::npos) {
      c.reals(1000, "POTCAR atomic pseudo charge");
      break;
    } else {
      throw Error(kIo, "unknown POTCAR local-data selector");
    }
  }
  // non-local projectors (reference: src/half_potcar.F90:109-146)
  p.pspnl_gmax = c.reals(1, "POTCAR projector grid")[0];
  std::vector<std::vector<double>> blocks;
  std::vector<std::vector<double>> tables;
  while (true) {
    std::string l = lower(c.next_nonblank());
    if (l.find("paw radial sets") != std::string::npos) break;
    if (l.find("non local part") == std::string::npos)
      throw Error(kIo, "expected non-local projector block");
    auto hdr = c.reals(3, "POTCAR projector channels");
    int lval = static_cast<int>(hdr[0]), npro = static_cast<int>(hdr[1]);
    if (npro < 1 || lval < 0 || lval > 3) throw Error(kIo, "invalid POTCAR projector channel header");
    blocks.push_back(c.reals(size_t(npro) * npro, "POTCAR DION"));
    for (int i = 0; i < npro; ++i) {
      if (lower(c.next_nonblank()).find("reciprocal space part") == std::string::npos)
        throw Error(kIo, "expected reciprocal projector");
      tables.push_back(c.reals(100, "POTCAR reciprocal projector"));
      if (lower(c.next_nonblank()).find("real space part") == std::string::npos)
        throw Error(kIo, "expected real projector");
      c.reals(100, "POTCAR real projector");
      p.lps.push_back(lval);
    }
  }
  int n = p.channels();
  p.dion.assign(size_t(n) * n, 0.0);
  int o = 0;
  for (auto &blk : blocks) {
    int k = static_cast<int>(std::lround(std::sqrt(double(blk.size()))));
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) p.dion[size_t(o + i) * n + (o + j)] = blk[size_t(i) * k + j];
    o += k;
  }
  p.pspnl.clear();
  for (auto &t : tables) p.pspnl.insert(p.pspnl.end(), t.begin(), t.end());
  // PAW radial sets: augmentation charges, radial grid and partial waves;
  // the other radial tables are skipped (reference: src/half_potcar.F90:148-205)
  auto hdr = c.reals(2, "POTCAR PAW grid header");
  const int nmax = static_cast<int>(hdr[0]);
  if (nmax < 1 || nmax > 100000) throw Error(kIo, "invalid POTCAR PAW grid size");
  p.paw_rmax = hdr[1];
  c.line();  // format line
  if (lower(c.next_nonblank()).find("augmentation charges") == std::string::npos)
    throw Error(kIo, "expected augmentation charges");
  p.qpaw = c.reals(size_t(n) * n, "POTCAR QPAW");
  p.rgrid.assign(nmax, 0.0);
  p.wps.assign(size_t(nmax) * n, 0.0);
  p.wae.assign(size_t(nmax) * n, 0.0);
  while (true) {
    std::string key = lower(c.next_nonblank());
    const size_t b = key.find_first_not_of(" \t");
    key = b == std::string::npos ? std::string() : key.substr(b);
    while (!key.empty() && (key.back() == ' ' || key.back() == '\r' || key.back() == '\t')) key.pop_back();
    if (key.find("occupancies in atom") != std::string::npos || key.find("uccopancies in atom") != std::string::npos) {
      c.reals(size_t(n) * n, "POTCAR QATO");
    } else if (key == "grid") {
      p.rgrid = c.reals(nmax, "POTCAR radial grid");
    } else if (key == "aepotential" || key.find("core charge-density") != std::string::npos ||
               key.find("local pseudopotential core") != std::string::npos ||
               key.find("pspotential valence only") != std::string::npos ||
               key.find("kinetic energy-density") != std::string::npos) {
      c.reals(nmax, "POTCAR PAW radial table");
    } else if (key.find("pseudo wavefunction") != std::string::npos) {
      for (int ich = 0; ich < n; ++ich) {
        if (ich > 0 && lower(c.next_nonblank()).find("pseudo wavefunction") == std::string::npos)
          throw Error(kIo, "expected POTCAR pseudo wavefunction");
        auto ps = c.reals(nmax, "POTCAR pseudo wavefunction");
        c.next_nonblank();  // ae wavefunction header
        auto ae = c.reals(nmax, "POTCAR AE wavefunction");
        std::copy(ps.begin(), ps.end(), p.wps.begin() + size_t(ich) * nmax);
        std::copy(ae.begin(), ae.end(), p.wae.begin() + size_t(ich) * nmax);
      }
      break;
    } else if (!key.empty() && key[0] == 't') {
      c.reals(size_t(n) * n, "POTCAR total charge");
    } else {
      throw Error(kIo, "unknown POTCAR PAW selector");
    }
  }
  while (!c.eof()) {
    const std::string &l = c.line();
    if (l.find("End of Dataset") != std::string::npos) break;
  }
}

}  // namespace

std::vector<Dataset> read_potcar(const std::string &path) {
  // a VASP HDF5 file carries the POTCAR text at /input/potcar/content
  // (reference: src/half_potcar.F90:34)
  Cursor c = is_hdf5(path) ? Cursor(read_h5_potcar(path), "POTCAR", true) : Cursor(path, "POTCAR");
  std::vector<Dataset> out;
  while (true) {
    while (!c.eof() && blank(c.lines[c.pos])) ++c.pos;
    if (c.eof()) break;
    std::string header = c.line();
    if (header.find("End of Dataset") != std::string::npos) continue;
    Dataset d;
    read_dataset(c, header, d);
    out.push_back(std::move(d));
    if (out.size() > 32) throw Error(kIo, "too many POTCAR datasets");
  }
  return out;
}

// reference: src/half_potcar.F90:12-22
void validate_datasets(const std::vector<Dataset> &sets, const Crystal &crystal) {
  if (sets.size() != crystal.species.size())
    throw Error(kInvalidArgument, "POTCAR dataset count does not match structure species count");
  for (size_t i = 0; i < sets.size(); ++i) {
    std::string s = lower(crystal.species[i]);
    if (s.rfind("type", 0) == 0) continue;
    if (lower(sets[i].element) != s)
      throw Error(kInvalidArgument, "POTCAR dataset order does not match structure species order");
  }
}

Basis build_basis(const Crystal &crystal, const int shape[3], double encut, const double kpoint[3]) {
  if (!(encut > 0)) throw Error(kInvalidArgument, "ENCUT must be positive");
  for (int i = 0; i < 3; ++i)
    if (!std::isfinite(kpoint[i])) throw Error(kInvalidArgument, "k point must be finite");
  Basis b;
  for (int i = 0; i < 3; ++i) b.shape[i] = shape[i], b.kpoint[i] = kpoint[i];
  b.encut = encut;
  const auto &B = crystal.reciprocal;
  for (int n1 = -shape[0] / 2; n1 <= (shape[0] - 1) / 2; ++n1)
    for (int n2 = -shape[1] / 2; n2 <= (shape[1] - 1) / 2; ++n2)
      for (int n3 = -shape[2] / 2; n3 <= (shape[2] - 1) / 2; ++n3) {
        double f[3] = {n1 + kpoint[0], n2 + kpoint[1], n3 + kpoint[2]};
        double q[3];
        for (int j = 0; j < 3; ++j) q[j] = f[0] * B[0][j] + f[1] * B[1][j] + f[2] * B[2][j];
        double kin = kHsqdtm * (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        if (kin < encut) {
          b.gvec.insert(b.gvec.end(), {n1, n2, n3});
          b.q.insert(b.q.end(), {q[0], q[1], q[2]});
          b.kinetic.push_back(kin);
          int i1 = ((n1 % shape[0]) + shape[0]) % shape[0];
          int i2 = ((n2 % shape[1]) + shape[1]) % shape[1];
          int i3 = ((n3 % shape[2]) + shape[2]) % shape[2];
          b.fft_index.push_back((i1 * shape[1] + i2) * shape[2] + i3);
        }
      }
  return b;
}

// reference: src/half_paw.F90:78-99 (one-based indices kept in the comments)
std::vector<double> spline_second(const std::vector<double> &y, double xmax) {
  const int n = static_cast<int>(y.size());
  if (n < 5) throw Error(kInvalidArgument, "radial table too short");
  const double h = xmax / n;
  std::vector<double> diag(n + 1, 0.0), rhs(n + 1, 0.0), upper(n + 1, 0.0), m2(n + 1, 0.0);
  auto Y = [&](int i) { return y[i - 1]; };
  diag[2] = 1;
  rhs[2] = ((Y(3) - Y(2)) / h - (Y(2) - Y(1)) / h) / h;
  for (int i = 3; i <= n - 2; ++i) {
    diag[i] = 4 * h;
    upper[i] = h;
    rhs[i] = 6 * ((Y(i + 1) - Y(i)) / h - (Y(i) - Y(i - 1)) / h);
  }
  diag[n - 1] = 1;
  rhs[n - 1] = ((Y(n) - Y(n - 1)) / h - (Y(n - 1) - Y(n - 2)) / h) / h;
  for (int i = 3; i <= n - 1; ++i) {
    double w = (i <= n - 2 ? h : 0.0) / diag[i - 1];
    diag[i] = diag[i] - w * upper[i - 1];
    rhs[i] = rhs[i] - w * rhs[i - 1];
  }
  m2[n - 1] = rhs[n - 1] / diag[n - 1];
  for (int i = n - 2; i >= 2; --i) m2[i] = (rhs[i] - upper[i] * m2[i + 1]) / diag[i];
  m2[1] = 2 * m2[2] - m2[3];
  m2[n] = 2 * m2[n - 1] - m2[n - 2];
  return std::vector<double>(m2.begin() + 1, m2.end());
}

void lm_expand(const Dataset &d, std::vector<double> &dlm, std::vector<double> &qlm) {
  const int nlm = d.nlm(), nch = d.channels();
  dlm.assign(size_t(nlm) * nlm, 0.0);
  qlm.assign(size_t(nlm) * nlm, 0.0);
  std::vector<int> ch, ll, mm;
  for (int ic = 0; ic < nch; ++ic)
    for (int m = -d.lps[ic]; m <= d.lps[ic]; ++m) ch.push_back(ic), ll.push_back(d.lps[ic]), mm.push_back(m);
  for (int a = 0; a < nlm; ++a)
    for (int b = 0; b < nlm; ++b)
      if (ll[a] == ll[b] && mm[a] == mm[b]) {
        dlm[size_t(a) * nlm + b] = d.dion[size_t(ch[a]) * nch + ch[b]];
        qlm[size_t(a) * nlm + b] = d.qpaw[size_t(ch[a]) * nch + ch[b]];
      }
}

}  // namespace half
