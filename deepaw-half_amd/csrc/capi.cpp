// The versioned HALF C ABI (half.h) and its device-level extension
// (half_amd.h) on top of the MI355X backend.
//
// Handle registry, argument validation order, status codes and error-buffer
// semantics follow the reference's bind(C) layer (reference:
// src/half_c_api.F90:48-665, src/half_library.F90:51-388).  The context owns
// one HIP stream, one rocBLAS handle, the device-resident System (Veff built
// once at creation) and a cache of per-k operators, so repeated solves at one
// k point reuse the projector matrix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "../../include/half_amd.h"
#include "common.hpp"
#include "contract.hpp"
#include "forces.hpp"
#include "energy.hpp"
#include "fset.hpp"
#include "h5io.hpp"
#include "kmesh.hpp"
#include "noalloc.hpp"
#include "paw.hpp"
#include "qdep.hpp"
#include "solver.hpp"

namespace half {
namespace {

constexpr int kMaxContexts = 32;

struct KeyK {
  double k[3];
  bool operator<(const KeyK &o) const { return std::lexicographical_compare(k, k + 3, o.k, o.k + 3); }
};

struct Context {
  System sys;
  int solver = kSolverEvd;
  bool use_uspp = false;
  hipStream_t stream = nullptr;
  Blas blas;
  // operators cached per k point; device-level handles share ownership, so
  // evicting a cache entry never frees an operator a host still holds
  std::map<KeyK, std::shared_ptr<KOperator<double>>> ops64;
  std::map<KeyK, std::shared_ptr<KOperator<float>>> ops32;
  // request geometry (reference: src/half_library.F90:31-35)
  bool has_request = false;
  int32_t req_nions = 0, req_ntypes = 0, req_grid[3] = {0, 0, 0};
  double req_lattice[9] = {};
  std::vector<int32_t> req_species;
  std::vector<double> req_positions;
  std::vector<double> acc_history;  // residual history of the last all-band solve

  ~Context() {
    // Drain the stream, release device objects, destroy the rocBLAS handle
    // while its stream still exists, and only then the stream.
    if (stream) (void)hipStreamSynchronize(stream);
    ops64.clear();
    ops32.clear();
    blas.release();
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
  template <typename R>
  std::shared_ptr<KOperator<R>> cached(std::map<KeyK, std::shared_ptr<KOperator<R>>> &cache, const double k[3]) {
    KeyK key{{k[0], k[1], k[2]}};
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    if (cache.size() >= 8) cache.clear();
    auto op = std::make_shared<KOperator<R>>();
    op->build(sys, k, stream);
    cache[key] = op;
    return op;
  }
  KOperator<double> &op64(const double k[3]) { return *cached(ops64, k); }
  KOperator<float> &op32(const double k[3]) { return *cached(ops32, k); }
};

Context *g_contexts[kMaxContexts + 1] = {};
int64_t g_generation[kMaxContexts + 1] = {};

struct OperatorEntry {
  int64_t owner;
  int precision;
  std::shared_ptr<KOperator<double>> op64;
  std::shared_ptr<KOperator<float>> op32;
  Context *ctx;
};
std::unordered_map<int64_t, OperatorEntry> g_ops;
int64_t g_next_op = 1;

void clear_error(char *error, int capacity) {
  if (error && capacity > 0) error[0] = '\0';
}

int export_error(int code, const std::string &msg, char *error, int capacity) {
  if (error && capacity > 0) {
    const size_t n = std::min(msg.size(), size_t(capacity - 1));
    std::memcpy(error, msg.data(), n);
    error[n] = '\0';
  }
  return code;
}

// reference: src/half_c_api.F90:627-634
int context_slot(int64_t handle) {
  int slot = static_cast<int>(((handle % kMaxContexts) + kMaxContexts) % kMaxContexts);
  if (slot == 0) slot = kMaxContexts;
  const int64_t g = (handle - slot) / kMaxContexts;
  if (handle <= 0 || !g_contexts[slot] || g != g_generation[slot]) return 0;
  return slot;
}

int bad_handle(char *error, int capacity) {
  return export_error(kInvalidHandle, "invalid HALF context handle", error, capacity);
}

template <typename F>
int guarded(char *error, int capacity, F &&body) {
  try {
    return body();
  } catch (const Error &e) {
    return export_error(e.code, e.what(), error, capacity);
  } catch (const std::bad_alloc &) {
    return export_error(kInternal, "host memory allocation failed", error, capacity);
  } catch (const std::exception &e) {
    return export_error(kInternal, e.what(), error, capacity);
  }
}

bool file_exists(const char *p) {
  std::ifstream f(p);
  return static_cast<bool>(f);
}

void require_device() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
    throw Error(kUnavailable, "no HIP device is available for the GPU backend");
}

double det9(const double *a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}

int create(const char *charge, const char *potential, double encut, int xc, int use_uspp, int backend, int solver,
           half_handle *handle, char *error, int capacity) {
  if (handle) *handle = 0;
  clear_error(error, capacity);
  if (!charge || !potential || !*charge || !*potential || (xc != 1 && xc != 2) || backend < 0 || backend > 2 ||
      solver < 1 || solver > 4)
    return export_error(kInvalidArgument, "invalid path, XC, backend, or solver selector", error, capacity);
  if (!handle) return export_error(kInvalidArgument, "handle output pointer is required", error, capacity);
  int slot = 0;
  for (int i = 1; i <= kMaxContexts; ++i)
    if (!g_contexts[i]) {
      slot = i;
      break;
    }
  if (!slot) return export_error(kInternal, "HALF context registry is full", error, capacity);
  return guarded(error, capacity, [&]() {
    // reference: src/half_library.F90:61-109
    if (!(encut > 0) || !std::isfinite(encut))
      return export_error(kInvalidArgument, "ENCUT must be positive and finite", error, capacity);
    if (backend == HALF_BACKEND_CPU)
      return export_error(kUnavailable, "this library provides the GPU backend only; CPU backend is not compiled",
                          error, capacity);
    if (!file_exists(charge)) return export_error(kIo, "charge input does not exist", error, capacity);
    if (!file_exists(potential)) return export_error(kIo, "potential input does not exist", error, capacity);
    auto ctx = std::make_unique<Context>();
    read_chgcar(charge, ctx->sys.crystal, ctx->sys.charge);
    ctx->sys.sets = read_potcar(potential);
    validate_datasets(ctx->sys.sets, ctx->sys.crystal);
    for (int d = 0; d < 3; ++d)
      if (!FftPlan::supported(ctx->sys.charge.shape[d]))
        return export_error(kUnavailable, "FFT grid dimension has a prime factor above 7", error, capacity);
    ctx->sys.encut = encut;
    ctx->sys.use_pbe = xc == 2;
    ctx->solver = solver;
    ctx->use_uspp = use_uspp != 0;
    ctx->sys.use_qdep = ctx->use_uspp;
    require_device();
    HALF_HIP(hipStreamCreate(&ctx->stream));
    ctx->blas.init(ctx->stream);
    ctx->sys.build(ctx->stream);
    g_contexts[slot] = ctx.release();
    ++g_generation[slot];
    *handle = g_generation[slot] * kMaxContexts + slot;
    return int(kSuccess);
  });
}

Basis make_basis(Context &c, const double *k) {
  if (!k) throw Error(kInvalidArgument, "k point pointer is required");
  for (int i = 0; i < 3; ++i)
    if (!std::isfinite(k[i])) throw Error(kInvalidArgument, "k point must be finite");
  return build_basis(c.sys.crystal, c.sys.charge.shape, c.sys.encut, k);
}

void copy_vectors_out(const DeviceArray<double2> &v, int n, int nb, double *out, int64_t ld, hipStream_t s) {
  HALF_HIP(hipMemcpy2DAsync(out, size_t(ld) * sizeof(double2), v.get(), size_t(n) * sizeof(double2),
                            size_t(n) * sizeof(double2), nb, hipMemcpyDeviceToHost, s));
  HALF_HIP(hipStreamSynchronize(s));
}

// Makes a request geometry the context's crystal.  The density file fixes
// the grid and the composition, so the request must have the file's grid,
// one type per potential dataset and the file's ion count per type; its
// lattice and fractional positions (any ion order, regrouped by type, stable)
// replace the file's.  Veff, the projector tables, the atom positions and,
// with the potential-dependent correction, D are rebuilt on it; cached
// operators and the context's operator handles are dropped, since their
// basis and projectors belong to the old cell.  On a failed rebuild the
// file's crystal is restored.  This is the fix the reference records in
// docs/validation/HFO2_NWAVE_BOUNDARY_MISMATCH.md:43-50: the reference stores
// the geometry (src/half_library.F90:227-247) but keeps building from the
// density file (the same document, :33-36).
void apply_request_geometry(Context &c, int64_t handle, const half_request_geometry_v1 &g) {
  System &s = c.sys;
  const int ntypes = static_cast<int>(s.sets.size());
  if (g.ntypes != ntypes)
    throw Error(kInvalidArgument, "request geometry has " + std::to_string(g.ntypes) +
                                      " types but the potential file has " + std::to_string(ntypes) + " datasets");
  for (int d = 0; d < 3; ++d)
    if (g.grid[d] != s.charge.shape[d])
      throw Error(kInvalidArgument, "request grid " + std::to_string(g.grid[0]) + "x" + std::to_string(g.grid[1]) +
                                        "x" + std::to_string(g.grid[2]) + " differs from the density grid " +
                                        std::to_string(s.charge.shape[0]) + "x" + std::to_string(s.charge.shape[1]) +
                                        "x" + std::to_string(s.charge.shape[2]));
  std::vector<int> counts(ntypes, 0);
  for (int i = 0; i < g.nions; ++i) ++counts[g.species[i] - 1];
  if (counts != s.crystal.counts)
    throw Error(kInvalidArgument, "request geometry ion count per type differs from the density file");
  Crystal cell = s.crystal;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) cell.lattice[i][j] = g.lattice[3 * i + j];
  cell.positions.clear();
  for (int t = 1; t <= ntypes; ++t)
    for (int i = 0; i < g.nions; ++i)
      if (g.species[i] == t)
        cell.positions.insert(cell.positions.end(), g.positions_fractional + 3 * i, g.positions_fractional + 3 * i + 3);
  cell.update_geometry();
  HALF_HIP(hipStreamSynchronize(c.stream));
  for (auto it = g_ops.begin(); it != g_ops.end();)
    it = it->second.owner == handle ? g_ops.erase(it) : std::next(it);
  c.ops64.clear();
  c.ops32.clear();
  const Crystal file_cell = s.crystal;
  s.crystal = cell;
  try {
    s.build(c.stream);
    HALF_HIP(hipStreamSynchronize(c.stream));
  } catch (...) {
    s.crystal = file_cell;
    s.build(c.stream);
    throw;
  }
}

// First G vector that lies in only one of the two sets, as "(x, y, z) in the
// host basis only" / "... in the HALF basis only"; empty when the sets agree.
std::string first_differing_g(const Basis &b, int64_t host_npw, const int32_t *host_gvec) {
  auto key = [](int32_t x, int32_t y, int32_t z) {
    return (uint64_t(uint32_t(x) & 0x1fffff) << 42) | (uint64_t(uint32_t(y) & 0x1fffff) << 21) |
           uint64_t(uint32_t(z) & 0x1fffff);
  };
  auto text = [](const int32_t *g, const char *where) {
    return "(" + std::to_string(g[0]) + ", " + std::to_string(g[1]) + ", " + std::to_string(g[2]) + ") in the " +
           where + " basis only";
  };
  std::unordered_map<uint64_t, int64_t> host, own;
  host.reserve(size_t(host_npw) * 2);
  own.reserve(size_t(b.npw()) * 2);
  for (int64_t i = 0; i < host_npw; ++i) host.emplace(key(host_gvec[3 * i], host_gvec[3 * i + 1], host_gvec[3 * i + 2]), i);
  for (int64_t i = 0; i < b.npw(); ++i) {
    const int32_t *g = &b.gvec[3 * i];
    own.emplace(key(g[0], g[1], g[2]), i);
    if (!host.count(key(g[0], g[1], g[2]))) return text(g, "HALF");
  }
  for (int64_t i = 0; i < host_npw; ++i)
    if (!own.count(key(host_gvec[3 * i], host_gvec[3 * i + 1], host_gvec[3 * i + 2])))
      return text(host_gvec + 3 * i, "host");
  return std::string();
}

void run_solve(Context &c, const double *k, int nbands, bool want_vectors, DeviceArray<double2> &vec,
               SolveResult &res) {
  KOperator<double> &op = c.op64(k);
  if (nbands < 1 || nbands > op.n) throw Error(kInvalidArgument, "invalid requested band count");
  if (c.solver == kSolverAcc) {
    // reference controls: tol 1e-4 eV, 40 iterations, every band's correction
    // column in each Rayleigh-Ritz step, each step through the Cholesky
    // reduction (src/half_cuda_iterative_solver.cuf:113-119)
    AccControls ctl;
    ctl.lock = 0;
    ctl.precond = 0;
    ctl.identity = 0;
    solve_acc(c.blas, op, nullptr, nbands, ctl, want_vectors, vec, res);
  } else {
    solve_dense(c.blas, op, c.solver, nbands, want_vectors, vec, res);
  }
}

}  // namespace
}  // namespace half

using namespace half;

extern "C" {

int half_get_abi_version(void) { return HALF_ABI_VERSION; }

const char *half_get_version_string(void) { return "0.6.0"; }

int half_get_capabilities(void) { return HALF_CAP_CUDA; }

int half_escn_health(const char *, int, char *error, int capacity) {
  clear_error(error, capacity);
  return export_error(kUnavailable, "DeePAW-eSCN network client is not part of this library", error, capacity);
}

int half_escn_predict(const char *, const half_escn_request_v1 *, half_escn_result_v1 *, int, char *error,
                      int capacity) {
  clear_error(error, capacity);
  return export_error(kUnavailable, "DeePAW-eSCN network client is not part of this library", error, capacity);
}

int half_create_from_files(const char *charge, const char *potential, double encut, int xc, int use_uspp,
                           half_handle *handle, char *error, int capacity) {
  if (handle) *handle = 0;
  clear_error(error, capacity);
  if (!charge || !potential || !*charge || !*potential || (xc != 1 && xc != 2))
    return export_error(kInvalidArgument, "invalid path or XC selector", error, capacity);
  return create(charge, potential, encut, xc, use_uspp, HALF_BACKEND_AUTO, HALF_SOLVER_EVD, handle, error, capacity);
}

int half_create_from_files_backend(const char *charge, const char *potential, double encut, int xc, int use_uspp,
                                   int backend, int solver, half_handle *handle, char *error, int capacity) {
  return create(charge, potential, encut, xc, use_uspp, backend, solver, handle, error, capacity);
}

int half_create_from_escn(const char *base_url, const char *potential, const half_request_geometry_v1 *geometry,
                          double, int xc, int, int backend, int solver, int normalization, int,
                          half_handle *handle, char *error, int capacity) {
  if (handle) *handle = 0;
  clear_error(error, capacity);
  if (!base_url || !potential || !*base_url || !*potential || !geometry || (xc != 1 && xc != 2) || backend < 0 ||
      backend > 2 || solver < 1 || solver > 4 || (normalization != 0 && normalization != 1))
    return export_error(kInvalidArgument, "invalid eSCN URL, path, geometry, execution policy, or normalization",
                        error, capacity);
  return export_error(kUnavailable, "DeePAW-eSCN network density source is not part of this library", error,
                      capacity);
}

int half_destroy(half_handle handle, char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  return guarded(error, capacity, [&]() {
    for (auto it = g_ops.begin(); it != g_ops.end();)
      it = it->second.owner == handle ? g_ops.erase(it) : std::next(it);
    delete g_contexts[slot];
    g_contexts[slot] = nullptr;
    return int(kSuccess);
  });
}

int half_set_request_geometry(half_handle handle, const half_request_geometry_v1 *g, char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  if (!g) return export_error(kInvalidArgument, "request geometry pointer is required", error, capacity);
  if (g->struct_size < sizeof(half_request_geometry_v1) || g->flags != 0)
    return export_error(kInvalidArgument, "unsupported request geometry structure size or flags", error, capacity);
  if (g->nions < 1 || g->ntypes < 1 || !g->species || !g->positions_fractional)
    return export_error(kInvalidArgument, "request geometry arrays and positive dimensions are required", error,
                        capacity);
  // reference: src/half_library.F90:227-247
  if (g->grid[0] < 1 || g->grid[1] < 1 || g->grid[2] < 1)
    return export_error(kInvalidArgument, "invalid request geometry dimensions", error, capacity);
  for (int i = 0; i < g->nions; ++i)
    if (g->species[i] < 1 || g->species[i] > g->ntypes)
      return export_error(kInvalidArgument, "invalid request geometry values", error, capacity);
  // the reference rejects NaN; the geometry is used here, so infinities too
  for (int i = 0; i < 9; ++i)
    if (!std::isfinite(g->lattice[i]))
      return export_error(kInvalidArgument, "invalid request geometry values", error, capacity);
  for (int i = 0; i < 3 * g->nions; ++i)
    if (!std::isfinite(g->positions_fractional[i]))
      return export_error(kInvalidArgument, "invalid request geometry values", error, capacity);
  if (std::fabs(det9(g->lattice)) < 1e-12)
    return export_error(kInvalidArgument, "request lattice is singular", error, capacity);
  return guarded(error, capacity, [&]() {
    Context &c = *g_contexts[slot];
    apply_request_geometry(c, handle, *g);
    c.req_nions = g->nions;
    c.req_ntypes = g->ntypes;
    for (int i = 0; i < 3; ++i) c.req_grid[i] = g->grid[i];
    std::memcpy(c.req_lattice, g->lattice, sizeof(c.req_lattice));
    c.req_species.assign(g->species, g->species + g->nions);
    c.req_positions.assign(g->positions_fractional, g->positions_fractional + 3 * g->nions);
    c.has_request = true;
    return int(kSuccess);
  });
}

int half_get_request_geometry(half_handle handle, int32_t *nions, int32_t *ntypes, int32_t grid[3],
                              double lattice[9], int64_t species_capacity, int32_t *species,
                              int64_t positions_capacity, double *positions, char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  Context &c = *g_contexts[slot];
  if (!c.has_request) return export_error(kUnavailable, "request geometry has not been set", error, capacity);
  if (!nions || !ntypes)
    return export_error(kInvalidArgument, "nions and ntypes output pointers are required", error, capacity);
  *nions = c.req_nions;
  *ntypes = c.req_ntypes;
  if (grid)
    for (int i = 0; i < 3; ++i) grid[i] = c.req_grid[i];
  if (lattice) std::memcpy(lattice, c.req_lattice, sizeof(c.req_lattice));
  if (species) {
    if (species_capacity < c.req_nions)
      return export_error(kCapacity, "request species output capacity is too small", error, capacity);
    std::copy(c.req_species.begin(), c.req_species.end(), species);
  }
  if (positions) {
    if (positions_capacity < 3 * int64_t(c.req_nions))
      return export_error(kCapacity, "request positions output capacity is too small", error, capacity);
    std::copy(c.req_positions.begin(), c.req_positions.end(), positions);
  }
  return int(kSuccess);
}

int half_get_system_info(half_handle handle, int *nions, int *ntypes, int grid[3], double lattice[9],
                         double *volume, char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  if (!nions || !ntypes || !grid || !lattice || !volume)
    return export_error(kInvalidArgument, "system info output pointers are required", error, capacity);
  const System &s = g_contexts[slot]->sys;
  *nions = s.crystal.nions();
  *ntypes = static_cast<int>(s.crystal.species.size());
  for (int i = 0; i < 3; ++i) grid[i] = s.charge.shape[i];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) lattice[3 * i + j] = s.crystal.lattice[i][j];
  *volume = s.crystal.volume;
  return int(kSuccess);
}

int half_get_basis_size(half_handle handle, const double kpoint[3], int64_t *npw, char *error, int capacity) {
  clear_error(error, capacity);
  if (npw) *npw = 0;
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  if (!npw) return export_error(kInvalidArgument, "npw output pointer is required", error, capacity);
  return guarded(error, capacity, [&]() {
    *npw = make_basis(*g_contexts[slot], kpoint).npw();
    return int(kSuccess);
  });
}

int half_get_basis(half_handle handle, const double kpoint[3], int64_t cap, int32_t *gvec, double *kinetic,
                   char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  return guarded(error, capacity, [&]() {
    Basis b = make_basis(*g_contexts[slot], kpoint);
    if (cap < b.npw()) return export_error(kCapacity, "basis output capacity is too small", error, capacity);
    if (!gvec || !kinetic) return export_error(kInvalidArgument, "basis output pointers are required", error, capacity);
    std::copy(b.gvec.begin(), b.gvec.end(), gvec);
    std::copy(b.kinetic.begin(), b.kinetic.end(), kinetic);
    return int(kSuccess);
  });
}

int half_assemble_hs(half_handle handle, const double kpoint[3], int64_t ld, HALF_COMPLEX *h, HALF_COMPLEX *s,
                     char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  if (!h || !s) return export_error(kInvalidArgument, "H and S output pointers are required", error, capacity);
  return guarded(error, capacity, [&]() {
    Context &c = *g_contexts[slot];
    Basis b = make_basis(c, kpoint);
    const int64_t n = b.npw();
    if (ld < n) return export_error(kCapacity, "matrix leading dimension is too small", error, capacity);
    KOperator<double> &op = c.op64(kpoint);
    DeviceArray<double2> H(size_t(n) * n), S(size_t(n) * n);
    assemble_dense(op, H.get(), S.get(), n, c.stream);
    HALF_HIP(hipMemcpy2DAsync(h, size_t(ld) * sizeof(double2), H.get(), size_t(n) * sizeof(double2),
                              size_t(n) * sizeof(double2), n, hipMemcpyDeviceToHost, c.stream));
    HALF_HIP(hipMemcpy2DAsync(s, size_t(ld) * sizeof(double2), S.get(), size_t(n) * sizeof(double2),
                              size_t(n) * sizeof(double2), n, hipMemcpyDeviceToHost, c.stream));
    HALF_HIP(hipStreamSynchronize(c.stream));
    return int(kSuccess);
  });
}

int half_apply_hs(half_handle handle, const double kpoint[3], int nstates, int64_t ld, const HALF_COMPLEX *psi,
                  HALF_COMPLEX *hpsi, HALF_COMPLEX *spsi, char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  if (nstates < 1 || !psi || !hpsi || !spsi)
    return export_error(kInvalidArgument, "invalid state block arguments", error, capacity);
  return guarded(error, capacity, [&]() {
    Context &c = *g_contexts[slot];
    Basis b = make_basis(c, kpoint);
    const int64_t n = b.npw();
    if (ld < n) return export_error(kCapacity, "state leading dimension is too small", error, capacity);
    KOperator<double> &op = c.op64(kpoint);
    DeviceArray<double2> dpsi(size_t(n) * nstates), dh(size_t(n) * nstates), ds(size_t(n) * nstates);
    HALF_HIP(hipMemcpy2DAsync(dpsi.get(), size_t(n) * sizeof(double2), psi, size_t(ld) * sizeof(double2),
                              size_t(n) * sizeof(double2), nstates, hipMemcpyHostToDevice, c.stream));
    op.apply(nstates, dpsi.get(), n, dh.get(), ds.get(), n, c.stream);
    HALF_HIP(hipMemcpy2DAsync(hpsi, size_t(ld) * sizeof(double2), dh.get(), size_t(n) * sizeof(double2),
                              size_t(n) * sizeof(double2), nstates, hipMemcpyDeviceToHost, c.stream));
    HALF_HIP(hipMemcpy2DAsync(spsi, size_t(ld) * sizeof(double2), ds.get(), size_t(n) * sizeof(double2),
                              size_t(n) * sizeof(double2), nstates, hipMemcpyDeviceToHost, c.stream));
    HALF_HIP(hipStreamSynchronize(c.stream));
    return int(kSuccess);
  });
}

int half_solve_kpoint(half_handle handle, const double kpoint[3], int nbands, double *eigenvalues,
                      HALF_COMPLEX *eigenvectors, int64_t ld, double *overlap_min, double *overlap_max, char *error,
                      int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  if (nbands < 1) return export_error(kInvalidArgument, "band count must be positive", error, capacity);
  if (!eigenvalues || !overlap_min || !overlap_max)
    return export_error(kInvalidArgument, "eigenvalue and overlap output pointers are required", error, capacity);
  return guarded(error, capacity, [&]() {
    Context &c = *g_contexts[slot];
    Basis b = make_basis(c, kpoint);
    const int64_t n = b.npw();
    if (nbands > n) return export_error(kInvalidArgument, "invalid requested band count", error, capacity);
    if (c.solver == kSolverAcc && 2 * int64_t(nbands) > n)
      return export_error(kInvalidArgument, "ACC solver requires NBANDS <= NPL/2", error, capacity);
    if (eigenvectors && ld < n)
      return export_error(kCapacity, "eigenvector leading dimension is too small", error, capacity);
    DeviceArray<double2> vec;
    SolveResult res;
    run_solve(c, kpoint, nbands, eigenvectors != nullptr, vec, res);
    std::copy(res.eigenvalues.begin(), res.eigenvalues.end(), eigenvalues);
    *overlap_min = res.overlap_min;
    *overlap_max = res.overlap_max;
    if (eigenvectors) copy_vectors_out(vec, static_cast<int>(n), nbands, eigenvectors, ld, c.stream);
    return int(kSuccess);
  });
}

int half_solve_kpoint_mapped(half_handle handle, const double kpoint[3], int nbands, int64_t host_npw,
                             const int32_t *host_gvec, double *eigenvalues, HALF_COMPLEX *eigenvectors, int64_t ld,
                             double *overlap_min, double *overlap_max, char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  if (nbands < 1 || host_npw < 1 || ld < host_npw || !eigenvectors || !host_gvec)
    return export_error(kInvalidArgument,
                        "invalid mapped-solve band count, basis size, leading dimension, or output pointer", error,
                        capacity);
  if (!eigenvalues || !overlap_min || !overlap_max)
    return export_error(kInvalidArgument, "eigenvalue and overlap output pointers are required", error, capacity);
  return guarded(error, capacity, [&]() {
    Context &c = *g_contexts[slot];
    Basis b = make_basis(c, kpoint);
    const int64_t n = b.npw();
    if (nbands > n) return export_error(kInvalidArgument, "invalid requested band count", error, capacity);
    if (c.solver == kSolverAcc && 2 * int64_t(nbands) > n)
      return export_error(kInvalidArgument, "ACC solver requires NBANDS <= NPL/2", error, capacity);
    // The bases are compared before the solve (the reference solves first,
    // src/half_c_api.F90:543-551); a mismatch names both sizes and the first
    // G vector found in one basis only.
    if (host_npw != n)
      return export_error(kInvalidArgument,
                          "host and HALF plane-wave bases have different sizes (host " + std::to_string(host_npw) +
                              ", HALF " + std::to_string(n) + "); first differing G vector " +
                              first_differing_g(b, host_npw, host_gvec),
                          error, capacity);
    // G-set permutation (reference: src/half_c_api.F90:553-585)
    std::unordered_map<uint64_t, int64_t> table;
    table.reserve(size_t(n) * 2);
    auto key = [](int32_t x, int32_t y, int32_t z) {
      return (uint64_t(uint32_t(x) & 0x1fffff) << 42) | (uint64_t(uint32_t(y) & 0x1fffff) << 21) |
             uint64_t(uint32_t(z) & 0x1fffff);
    };
    for (int64_t i = 0; i < n; ++i) {
      if (!table.emplace(key(host_gvec[3 * i], host_gvec[3 * i + 1], host_gvec[3 * i + 2]), i).second)
        return export_error(kInvalidArgument, "host G-vector list contains a duplicate", error, capacity);
    }
    std::vector<int64_t> slot_of(n);
    for (int64_t i = 0; i < n; ++i) {
      auto it = table.find(key(b.gvec[3 * i], b.gvec[3 * i + 1], b.gvec[3 * i + 2]));
      if (it == table.end())
        return export_error(kInvalidArgument,
                            "host and HALF plane-wave bases contain different G vectors; first differing G vector " +
                                first_differing_g(b, host_npw, host_gvec),
                            error, capacity);
      slot_of[i] = it->second;
    }
    DeviceArray<double2> vec;
    SolveResult res;
    run_solve(c, kpoint, nbands, true, vec, res);
    std::vector<double2> hv(size_t(n) * nbands);
    HALF_HIP(hipMemcpyAsync(hv.data(), vec.get(), hv.size() * sizeof(double2), hipMemcpyDeviceToHost, c.stream));
    HALF_HIP(hipStreamSynchronize(c.stream));
    double2 *out = reinterpret_cast<double2 *>(eigenvectors);
    for (int bnd = 0; bnd < nbands; ++bnd)
      for (int64_t i = 0; i < ld; ++i) out[i + bnd * ld] = make_double2(0, 0);
    for (int64_t i = 0; i < n; ++i)
      for (int bnd = 0; bnd < nbands; ++bnd) out[slot_of[i] + bnd * ld] = hv[i + size_t(bnd) * n];
    std::copy(res.eigenvalues.begin(), res.eigenvalues.end(), eigenvalues);
    *overlap_min = res.overlap_min;
    *overlap_max = res.overlap_max;
    return int(kSuccess);
  });
}

// ---------------------------------------------------------------------------
// device-level extension (half_amd.h)
// ---------------------------------------------------------------------------

int half_amd_operator_create(half_handle handle, const double kpoint[3], int precision, half_amd_operator *op,
                             int64_t *npw, int32_t *nproj, char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  if (!op || (precision != 64 && precision != 32))
    return export_error(kInvalidArgument, "operator output pointer and precision 64 or 32 are required", error,
                        capacity);
  return guarded(error, capacity, [&]() {
    Context &c = *g_contexts[slot];
    make_basis(c, kpoint);
    OperatorEntry e{handle, precision, nullptr, nullptr, &c};
    int64_t n;
    int32_t np;
    if (precision == 64) {
      e.op64 = c.cached(c.ops64, kpoint);
      n = e.op64->n;
      np = e.op64->np;
    } else {
      e.op32 = c.cached(c.ops32, kpoint);
      n = e.op32->n;
      np = e.op32->np;
    }
    HALF_HIP(hipStreamSynchronize(c.stream));
    const int64_t id = g_next_op++;
    g_ops[id] = e;
    *op = id;
    if (npw) *npw = n;
    if (nproj) *nproj = np;
    return int(kSuccess);
  });
}

int half_amd_operator_apply(half_amd_operator op, int stage, int nstates, const void *d_psi, int64_t ldpsi,
                            void *d_hpsi, void *d_spsi, int64_t ldout, int box_batch, void *stream, char *error,
                            int capacity) {
  clear_error(error, capacity);
  auto it = g_ops.find(op);
  if (it == g_ops.end() || !context_slot(it->second.owner))
    return export_error(kInvalidHandle, "invalid HALF operator handle", error, capacity);
  if (nstates < 1 || !d_psi || !d_hpsi || !d_spsi || stage < 0 || stage > 2)
    return export_error(kInvalidArgument, "invalid state block arguments", error, capacity);
  return guarded(error, capacity, [&]() {
    OperatorEntry &e = it->second;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : e.ctx->stream;
    auto run = [&](auto *k, auto *psi, auto *hp, auto *sp) {
      if (ldpsi < k->n || ldout < k->n) throw Error(kCapacity, "state leading dimension is too small");
      if (box_batch > 0) k->box_batch = box_batch;
      if (stage == 0)
        k->apply(nstates, psi, ldpsi, hp, sp, ldout, s);
      else if (stage == 1)
        k->apply_local(nstates, psi, ldpsi, hp, sp, ldout, s);
      else
        k->apply_projectors(nstates, psi, ldpsi, hp, sp, ldout, s);
    };
    if (e.precision == 64)
      run(e.op64.get(), static_cast<const double2 *>(d_psi), static_cast<double2 *>(d_hpsi), static_cast<double2 *>(d_spsi));
    else
      run(e.op32.get(), static_cast<const float2 *>(d_psi), static_cast<float2 *>(d_hpsi), static_cast<float2 *>(d_spsi));
    return int(kSuccess);
  });
}

int half_amd_operator_apply_noalloc(half_amd_operator op, int nstates, const void *d_psi, int64_t ldpsi,
                                    void *d_hpsi, void *d_spsi, int64_t ldout, void *stream, char *error,
                                    int capacity) {
  clear_error(error, capacity);
  auto it = g_ops.find(op);
  if (it == g_ops.end() || !context_slot(it->second.owner))
    return export_error(kInvalidHandle, "invalid HALF operator handle", error, capacity);
  if (nstates < 1 || !d_psi || !d_hpsi || !d_spsi)
    return export_error(kInvalidArgument, "invalid state block arguments", error, capacity);
  return guarded(error, capacity, [&]() {
    OperatorEntry &e = it->second;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : e.ctx->stream;
    void *work = noalloc_scratch();
    auto run = [&](auto *k, auto *psi, auto *hp, auto *sp) {
      if (ldpsi < k->n || ldout < k->n) throw Error(kCapacity, "state leading dimension is too small");
      k->apply_external(nstates, psi, ldpsi, hp, sp, ldout, s, work, kNoAllocScratchBytes);
    };
    if (e.precision == 64)
      run(e.op64.get(), static_cast<const double2 *>(d_psi), static_cast<double2 *>(d_hpsi), static_cast<double2 *>(d_spsi));
    else
      run(e.op32.get(), static_cast<const float2 *>(d_psi), static_cast<float2 *>(d_hpsi), static_cast<float2 *>(d_spsi));
    return int(kSuccess);
  });
}

int half_amd_operator_destroy(half_amd_operator op, char *error, int capacity) {
  clear_error(error, capacity);
  auto it = g_ops.find(op);
  if (it == g_ops.end()) return export_error(kInvalidHandle, "invalid HALF operator handle", error, capacity);
  g_ops.erase(it);
  return int(kSuccess);
}

int half_amd_solve_acc(half_handle handle, const double kpoint[3], int nbands, int mixed, double tol,
                       double switch_tol, int maxiter, int block, int extra, double *eigenvalues,
                       HALF_COMPLEX *eigenvectors, int64_t ld, half_amd_solve_stats *stats, char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  if (nbands < 1 || !eigenvalues)
    return export_error(kInvalidArgument, "band count must be positive", error, capacity);
  return guarded(error, capacity, [&]() {
    Context &c = *g_contexts[slot];
    Basis b = make_basis(c, kpoint);
    const int64_t n = b.npw();
    if (2 * int64_t(nbands + std::max(extra, 0)) > n)
      return export_error(kInvalidArgument, "ACC solver requires NBANDS <= NPL/2", error, capacity);
    if (eigenvectors && ld < n)
      return export_error(kCapacity, "eigenvector leading dimension is too small", error, capacity);
    AccControls ctl;
    ctl.tol = tol;
    ctl.maxiter = maxiter;
    ctl.block = block;
    ctl.mixed = mixed != 0;
    if (switch_tol > 0) ctl.switch_tol = switch_tol;
    ctl.extra = extra;
    KOperator<double> &op = c.op64(kpoint);
    KOperator<float> *op32 = ctl.mixed ? &c.op32(kpoint) : nullptr;
    DeviceArray<double2> vec;
    SolveResult res;
    solve_acc(c.blas, op, op32, nbands, ctl, eigenvectors != nullptr, vec, res);
    c.acc_history = res.history;
    std::copy(res.eigenvalues.begin(), res.eigenvalues.end(), eigenvalues);
    if (eigenvectors) copy_vectors_out(vec, static_cast<int>(n), nbands, eigenvectors, ld, c.stream);
    if (stats) {
      stats->iterations = res.iterations;
      stats->fp64_iterations = res.fp64_iterations;
      stats->max_residual_eV = res.residual;
      stats->seconds = res.seconds_solver;
      stats->band_applications = res.applies;
      stats->metric_fallbacks = res.metric_fallbacks;
      stats->rr_order_mean = static_cast<int32_t>(std::lround(res.rr_order_mean));
    }
    return int(kSuccess);
  });
}

int half_amd_solve_kmesh(half_handle handle, int nk, const double *kpoints, int nbands, int mixed, double tol,
                         double switch_tol, int maxiter, int extra, int graph, double *eigenvalues,
                         half_amd_solve_stats *stats, char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  if (nbands < 1 || !eigenvalues || !kpoints || nk < 1 || nk > 65535)
    return export_error(kInvalidArgument, "k points (1..65535), band count and outputs are required", error,
                        capacity);
  return guarded(error, capacity, [&]() {
    Context &c = *g_contexts[slot];
    std::vector<std::array<double, 3>> ks(nk);
    for (int j = 0; j < nk; ++j) {
      ks[j] = {kpoints[3 * j], kpoints[3 * j + 1], kpoints[3 * j + 2]};
      if (2 * int64_t(nbands + std::max(extra, 0)) > make_basis(c, ks[j].data()).npw())
        return export_error(kInvalidArgument, "ACC solver requires NBANDS <= NPL/2 at every k point", error,
                            capacity);
    }
    AccControls ctl;
    ctl.tol = tol;
    ctl.maxiter = maxiter;
    ctl.mixed = mixed != 0;
    if (switch_tol > 0) ctl.switch_tol = switch_tol;
    ctl.extra = extra;
    KMeshControls kc;
    kc.graph = graph != 0;
    KMeshOperator<double> op64;
    op64.build(c.sys, ks, c.stream);
    std::unique_ptr<KMeshOperator<float>> op32;
    if (ctl.mixed) {
      op32.reset(new KMeshOperator<float>());
      op32->build(c.sys, ks, c.stream);
    }
    std::vector<double> w;
    SolveResult res;
    solve_acc_kmesh(c.blas, op64, op32.get(), nbands, ctl, kc, w, res);
    c.acc_history = res.history;
    std::copy(w.begin(), w.end(), eigenvalues);
    if (stats) {
      stats->iterations = res.iterations;
      stats->fp64_iterations = res.fp64_iterations;
      stats->max_residual_eV = res.residual;
      stats->seconds = res.seconds_solver;
      stats->band_applications = res.applies;
      stats->metric_fallbacks = res.metric_fallbacks;
      stats->rr_order_mean = static_cast<int32_t>(std::lround(res.rr_order_mean));
    }
    return int(kSuccess);
  });
}

int half_amd_force_set(half_handle handle, int nk, const double *kpoints, const double *kweights, int nbands,
                       const double *eigenvalues, const double *occupations, const HALF_COMPLEX *const *vectors,
                       const int64_t *ld, double *forces, double *components, double *density, double *charges,
                       char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  if (nk < 1 || nbands < 1 || !kpoints || !kweights || !eigenvalues || !occupations || !vectors || !ld || !forces)
    return export_error(kInvalidArgument, "k points, bands, eigenvalues, occupations, vectors and forces are required",
                        error, capacity);
  return guarded(error, capacity, [&]() {
    Context &c = *g_contexts[slot];
    for (int j = 0; j < nk; ++j) {
      if (!vectors[j]) return export_error(kInvalidArgument, "eigenvectors are required", error, capacity);
      if (ld[j] < make_basis(c, kpoints + 3 * j).npw())
        return export_error(kCapacity, "eigenvector leading dimension is too small", error, capacity);
    }
    ForceSetState st;
    st.reset(c.sys, c.stream);
    for (int j = 0; j < nk; ++j) {
      std::vector<double> eps(eigenvalues + size_t(j) * nbands, eigenvalues + size_t(j + 1) * nbands);
      std::vector<double> occ(occupations + size_t(j) * nbands, occupations + size_t(j + 1) * nbands);
      KOperator<double> &op = c.op64(kpoints + 3 * j);
      force_set_add_k(c.sys, op, nbands, reinterpret_cast<const double2 *>(vectors[j]), ld[j], eps, occ, kweights[j],
                      st, c.stream);
    }
    ForceSetResult r;
    force_set_finish(c.sys, st, density != nullptr, r, c.stream);
    std::copy(r.total.begin(), r.total.end(), forces);
    if (components) {
      const std::vector<double> *parts[7] = {&r.ewald, &r.local, &r.nonlocal, &r.projector,
                                             &r.augmentation, &r.nlcc, &r.harris};
      for (int i = 0; i < 7; ++i) std::copy(parts[i]->begin(), parts[i]->end(), components + i * r.total.size());
    }
    if (density) std::copy(r.rho_out.begin(), r.rho_out.end(), density);
    if (charges) {
      charges[0] = r.smooth_charge;
      charges[1] = r.aug_charge;
    }
    return int(kSuccess);
  });
}

int half_amd_write_vaspwave(half_handle handle, const char *path, int nk, const double *kpoints, int nbands,
                            const double *eigenvalues, const double *occupations, double fermi,
                            const HALF_COMPLEX *const *vectors, const int64_t *ld, char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  if (!path || !*path || nk < 1 || nbands < 1 || !kpoints || !eigenvalues || !occupations || !vectors || !ld)
    return export_error(kInvalidArgument, "path, k points, bands, eigenvalues, occupations and vectors are required",
                        error, capacity);
  // reference: src/half_vaspwave.F90:110-118
  for (int64_t i = 0; i < int64_t(nk) * nbands; ++i)
    if (!(occupations[i] >= -1e-10 && occupations[i] <= 2.0 + 1e-10))
      return export_error(kInvalidArgument, "invalid HDF5 occupations", error, capacity);
  if (!std::isfinite(fermi)) return export_error(kInvalidArgument, "Fermi level must be finite", error, capacity);
  return guarded(error, capacity, [&]() {
    Context &c = *g_contexts[slot];
    WaveExport w;
    w.system = c.sys.crystal.system;
    w.crystal = &c.sys.crystal;
    w.charge = &c.sys.charge;
    w.encut = c.sys.encut;
    w.fermi = fermi;
    w.nk = nk;
    w.nb = nbands;
    w.kpoints = kpoints;
    w.eigenvalues = eigenvalues;
    w.occupations = occupations;
    for (int j = 0; j < nk; ++j) {
      const Basis b = make_basis(c, kpoints + 3 * j);
      const int64_t n = b.npw();
      if (nbands > n) return export_error(kInvalidArgument, "invalid requested band count", error, capacity);
      if (!vectors[j]) return export_error(kInvalidArgument, "eigenvectors are required", error, capacity);
      if (ld[j] < n) return export_error(kCapacity, "eigenvector leading dimension is too small", error, capacity);
      // coefficients in VASP's plane-wave order, single precision
      // (reference: src/half_vaspwave.F90:121-130)
      const std::vector<int64_t> order = vasp_order(b);
      const double2 *v = reinterpret_cast<const double2 *>(vectors[j]);
      std::vector<float> packed(size_t(2) * n * nbands);
      for (int band = 0; band < nbands; ++band)
        for (int64_t g = 0; g < n; ++g) {
          const double2 x = v[order[g] + band * ld[j]];
          packed[2 * (band * n + g)] = static_cast<float>(x.x);
          packed[2 * (band * n + g) + 1] = static_cast<float>(x.y);
        }
      w.npw.push_back(static_cast<int>(n));
      w.coefficients.push_back(std::move(packed));
    }
    write_vaspwave(path, w);
    return int(kSuccess);
  });
}

int half_amd_solve_history(half_handle handle, int capacity, double *residuals, int32_t *count, char *error,
                           int ecap) {
  clear_error(error, ecap);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, ecap);
  if (!count || (capacity > 0 && !residuals))
    return export_error(kInvalidArgument, "output pointers are required", error, ecap);
  const std::vector<double> &h = g_contexts[slot]->acc_history;
  *count = static_cast<int32_t>(h.size());
  if (capacity < static_cast<int64_t>(h.size()))
    return export_error(kCapacity, "residual history capacity is too small", error, ecap);
  std::copy(h.begin(), h.end(), residuals);
  return int(kSuccess);
}

int half_amd_set_solver(half_handle handle, int solver, char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  if (solver < 1 || solver > 4) return export_error(kInvalidArgument, "solver must be 1..4", error, capacity);
  g_contexts[slot]->solver = solver;
  return int(kSuccess);
}

int half_amd_energy_terms(half_handle handle, double *eh, double *exc, double *exv, char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  if (!eh || !exc || !exv) return export_error(kInvalidArgument, "output pointers are required", error, capacity);
  const System &s = g_contexts[slot]->sys;
  *eh = s.e_hartree;
  *exc = s.e_xc;
  *exv = s.e_xc_potential;
  return int(kSuccess);
}

int half_amd_harris_energy(half_handle handle, int nk, const double *kweights, int nbands, const double *eigenvalues,
                           double sigma, int ismear, double *occupations, half_amd_harris_terms *terms, char *error,
                           int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  return guarded(error, capacity, [&]() {
    if (nk < 1 || nbands < 1 || !kweights || !eigenvalues || !terms)
      return export_error(kInvalidArgument, "invalid band arguments", error, capacity);
    Context &c = *g_contexts[slot];
    std::vector<double> occ;
    const HarrisTerms t = harris_energy(c.sys, nk, nbands, eigenvalues, kweights, sigma, ismear, occ, c.stream);
    terms->band_energy = t.band_energy;
    terms->fermi = t.fermi;
    terms->entropy = t.entropy;
    terms->e_hartree = t.e_hartree;
    terms->e_xc = t.e_xc;
    terms->e_xc_potential = t.e_xc_potential;
    terms->ewald = t.ewald;
    terms->local_g0 = t.local_g0;
    terms->atomic_reference = t.atomic_reference;
    terms->harris_energy = t.harris;
    terms->free_energy = t.free_energy;
    terms->nelect = t.nelect;
    if (occupations) std::copy(occ.begin(), occ.end(), occupations);
    return int(kSuccess);
  });
}

int half_amd_set_projector_positions(half_handle handle, int nions, const double *frac, char *error,
                                     int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  return guarded(error, capacity, [&]() {
    Context &c = *g_contexts[slot];
    if (!frac) return export_error(kInvalidArgument, "positions pointer is required", error, capacity);
    if (nions != c.sys.natoms) return export_error(kInvalidArgument, "ion count does not match the context", error,
                                                   capacity);
    std::vector<double> cart;
    for (int a = 0; a < nions; ++a) {
      const double *f = frac + 3 * a;
      if (!std::isfinite(f[0]) || !std::isfinite(f[1]) || !std::isfinite(f[2]))
        return export_error(kInvalidArgument, "invalid position values", error, capacity);
      for (int j = 0; j < 3; ++j)
        cart.push_back(f[0] * c.sys.crystal.lattice[0][j] + f[1] * c.sys.crystal.lattice[1][j] +
                       f[2] * c.sys.crystal.lattice[2][j]);
    }
    HALF_HIP(hipStreamSynchronize(c.stream));
    c.ops64.clear();
    c.ops32.clear();
    c.sys.positions_cart = cart;
    c.sys.d_positions.upload(cart, c.stream);
    if (c.sys.use_qdep) build_qdep(c.sys, c.stream);  // D depends on the atom positions
    HALF_HIP(hipStreamSynchronize(c.stream));
    return int(kSuccess);
  });
}

int half_amd_get_veff(half_handle handle, int64_t cap, double *veff, char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  return guarded(error, capacity, [&]() {
    Context &c = *g_contexts[slot];
    const int64_t nr = c.sys.charge.size();
    if (!veff) return export_error(kInvalidArgument, "output pointer is required", error, capacity);
    if (cap < nr) return export_error(kCapacity, "Veff output capacity is too small", error, capacity);
    HALF_HIP(hipMemcpyAsync(veff, c.sys.veff.get(), sizeof(double) * nr, hipMemcpyDeviceToHost, c.stream));
    HALF_HIP(hipStreamSynchronize(c.stream));
    return int(kSuccess);
  });
}

int half_amd_nonlocal_forces(half_handle handle, const double kpoint[3], int nbands, const double *eps,
                             const double *occ, double kweight, const HALF_COMPLEX *vectors, int64_t ld,
                             double *forces, char *error, int capacity) {
  clear_error(error, capacity);
  const int slot = context_slot(handle);
  if (!slot) return bad_handle(error, capacity);
  if (nbands < 1 || !eps || !occ || !vectors || !forces)
    return export_error(kInvalidArgument, "invalid force arguments", error, capacity);
  return guarded(error, capacity, [&]() {
    Context &c = *g_contexts[slot];
    Basis b = make_basis(c, kpoint);
    const int64_t n = b.npw();
    if (ld < n) return export_error(kCapacity, "eigenvector leading dimension is too small", error, capacity);
    KOperator<double> &op = c.op64(kpoint);
    DeviceArray<double2> d(size_t(n) * nbands);
    HALF_HIP(hipMemcpy2DAsync(d.get(), size_t(n) * sizeof(double2), vectors, size_t(ld) * sizeof(double2),
                              size_t(n) * sizeof(double2), nbands, hipMemcpyHostToDevice, c.stream));
    std::vector<double> f(size_t(c.sys.natoms) * 3, 0.0);
    nonlocal_forces(op, nbands, d.get(), n, std::vector<double>(eps, eps + nbands),
                    std::vector<double>(occ, occ + nbands), kweight, f, c.stream);
    for (size_t i = 0; i < f.size(); ++i) forces[i] += f[i];
    return int(kSuccess);
  });
}

int half_amd_set_tuning(const char *key, int value, char *error, int capacity) {
  clear_error(error, capacity);
  if (!key || !(set_local_tuning(key, value) || set_operator_tuning(key, value) || set_solver_tuning(key, value) ||
                set_contract_tuning(key, value)))
    return export_error(kInvalidArgument, "unknown tuning key or value out of range", error, capacity);
  return int(kSuccess);
}

int half_amd_get_tuning(const char *key, int *value, char *error, int capacity) {
  clear_error(error, capacity);
  if (!key || !value ||
      !(get_local_tuning(key, *value) || get_operator_tuning(key, *value) || get_solver_tuning(key, *value) ||
        get_contract_tuning(key, *value)))
    return export_error(kInvalidArgument, "unknown tuning key", error, capacity);
  return int(kSuccess);
}

int half_amd_device_info(char *name, int name_capacity, int32_t *cus, char *error, int capacity) {
  clear_error(error, capacity);
  return guarded(error, capacity, [&]() {
    require_device();
    int dev = 0;
    HALF_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HALF_HIP(hipGetDeviceProperties(&p, dev));
    if (name && name_capacity > 0) {
      std::string s = std::string(p.name) + " " + p.gcnArchName;
      export_error(0, s, name, name_capacity);
    }
    if (cus) *cus = p.multiProcessorCount;
    return int(kSuccess);
  });
}

}  // extern "C"
