// k-point mesh batching (reference behaviour per k point:
// src/half_cuda_iterative_solver.cuf:79-207 for the solve,
// src/half_cuda_assembly.cuf:468-519 for the operator).
//
// The reference solves k points one after another, each with its own
// operator preparation and per-call cuBLAS/cuSOLVER handles.  Here all k
// points of a mesh advance together:
//   - the local stage is one launch per pass for every k point (per-k index
//     tables in a device array, ragged line / plane counts retired per
//     workgroup);
//   - B^H psi and B (D c), B (Q c) are one batched contraction launch each,
//     with every k point's projector matrix padded to the mesh's largest
//     plane-wave count (zero rows), so ragged sizes need no special cases;
//   - the Rayleigh-Ritz algebra of all k points is batched: Grams, the
//     S-orthogonalisation update, the metric normalisation and the Ritz
//     rotations on the batched MFMA contractions (the tall right-multiplies
//     in place); Cholesky factors with their explicit inverses, the
//     reductions to tridiagonal form and the tridiagonal eigensolves are
//     one-workgroup-per-k kernels (smalleig.hip, tridiag.hip) where every
//     subspace problem has order <= kSmallMaxOrder, else
//     rocsolver_zpotrf/ztrtri/zheevd_strided_batched with
//     rocblas_zgemm_strided_batched;
//   - with the one-workgroup kernels an iteration needs no host decision
//     between its launches (the metric's eigen-decomposition form is chosen
//     per k point on the device), so the whole iteration -- correction
//     columns, their H.psi, orthogonalisation, normalisation, Rayleigh-Ritz,
//     rotations and the next residual norms -- is captured once per
//     precision and correction width into a hipGraph and replayed.
// One host round trip per iteration carries the nk x m residual norms (and
// the per-k status flags of the device-side factorisations).
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>

#include "contract.hpp"
#include "cplx.hpp"
#include "kmesh.hpp"
#include "projector.hpp"
#include "smalleig.hpp"
#include "trace.hpp"
#include "tridiag.hpp"

namespace half {

namespace {

constexpr int kThreads = 256;
inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

#define KM_BLAS(call)                                                                                    \
  do {                                                                                                   \
    rocblas_status st_ = (call);                                                                         \
    if (st_ != rocblas_status_success)                                                                   \
      throw Error(kInternal, std::string("rocBLAS/rocSOLVER status ") + std::to_string(int(st_)) + " in " #call); \
  } while (0)

inline rocblas_double_complex *rz(double2 *p) { return reinterpret_cast<rocblas_double_complex *>(p); }

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Per-k batched helpers: grid.y (or the column index) selects the k point.

// x_k[g, b] = (g == idx[k m + b]), N rows per column, k stride zk
__global__ void init_states_kernel(double2 *x, int64_t ld, int64_t zk, const int *idx, int N, int m) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(N) * m) return;
  const int k = blockIdx.y;
  const int b = static_cast<int>(t / N), g = static_cast<int>(t - int64_t(b) * N);
  x[k * zk + g + b * ld] = make_double2(g == idx[k * m + b] ? 1.0 : 0.0, 0.0);
}

// p_k(g, j) = exp(i theta) for g < npw_k, zero on the padding rows: dense
// probe columns of norm sqrt(npw_k), a pseudo-random phase per element
// (integer hash of g and j, the same at every k point)
__global__ void phase_block_kernel(double2 *p, int64_t zk, const int *npw, int N, int m, int64_t ld) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(N) * m) return;
  const int k = blockIdx.y;
  const int j = static_cast<int>(t / N), g = static_cast<int>(t - int64_t(j) * N);
  double2 v = make_double2(0.0, 0.0);
  if (g < npw[k]) {
    uint32_t h = static_cast<uint32_t>(g) * 2654435761u ^ (static_cast<uint32_t>(j) + 0x9e3779b9u) * 2246822519u;
    h ^= h >> 15;
    h *= 2654435761u;
    h ^= h >> 13;
    sincos(6.283185307179586 * (h * 2.3283064365386963e-10), &v.y, &v.x);
  }
  p[k * zk + g + j * ld] = v;
}

// R_k = HX_k - SX_k diag(w_k) (reference :29-40)
__global__ void residual_kernel(const double2 *hx, const double2 *sx, int64_t zk, const double *w, int64_t wk,
                                double2 *r, int64_t rk, int N, int m, int64_t ld) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(N) * m) return;
  const int k = blockIdx.y;
  const int b = static_cast<int>(t / N), g = static_cast<int>(t - int64_t(b) * N);
  const int64_t o = k * zk + g + b * ld;
  const double e = w[k * wk + b];
  r[k * rk + g + b * ld] = make_double2(hx[o].x - e * sx[o].x, hx[o].y - e * sx[o].y);
}

// squared norms of ncols columns of length N (fixed-order tree per column)
__global__ void column_norm2_kernel(const double2 *r, int64_t ld, int N, double *out) {
  __shared__ double red[kThreads];
  const double2 *c = r + blockIdx.x * ld;
  double acc = 0;
  for (int g = threadIdx.x; g < N; g += blockDim.x) acc += c[g].x * c[g].x + c[g].y * c[g].y;
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// P = -R/(|T - eps| + 1) or a unit seed vector (reference :42-64).  Column c
// (< np) of k point k's P corrects band cols[k m + c] with seed seeds[k m + c].
__global__ void precondition_kernel(const double2 *r, int64_t rk, const double *kin, int64_t kk, const double *w,
                                    int64_t wk, const int *cols, const int *seeds, double2 *p, int64_t zk, int N,
                                    int m, int np, int64_t ld) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(N) * np) return;
  const int k = blockIdx.y;
  const int c = static_cast<int>(t / N), g = static_cast<int>(t - int64_t(c) * N);
  const int b = cols[k * m + c];
  const int64_t o = k * zk + g + c * ld;
  const int seed = seeds[k * m + c];
  if (seed >= 0) {
    p[o] = make_double2(g == seed ? 1.0 : 0.0, 0.0);
  } else {
    const double2 v = r[k * rk + g + b * ld];
    const double d = fabs(kin[k * kk + g] - w[k * wk + b]) + 1.0;
    p[o] = make_double2(-v.x / d, -v.y / d);
  }
}

// metric_k <- D metric_k D, D = diag(metric_k)^-1/2 (zero where not positive)
__global__ void unit_diagonal_kernel(double2 *M, double *d, int m) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int k = blockIdx.y;
  const double v = M[int64_t(k) * m * m + j + int64_t(j) * m].x;
  d[k * m + j] = v > 0 ? 1.0 / sqrt(v) : 0.0;
}
__global__ void scale_both_kernel(double2 *M, const double *d, int m) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(m) * m) return;
  const int k = blockIdx.y;
  const int j = static_cast<int>(t / m), i = static_cast<int>(t - int64_t(j) * m);
  const double f = d[k * m + i] * d[k * m + j];
  double2 &v = M[int64_t(k) * m * m + t];
  v = make_double2(v.x * f, v.y * f);
}
// out[k (m + 1) + j] = U_jj^2, out[k (m + 1) + m] = potrf info of k
__global__ void pivots_kernel(const double2 *U, const rocblas_int *info, double *out, int m) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (j < m) {
    const double2 u = U[int64_t(k) * m * m + j + int64_t(j) * m];
    out[k * (m + 1) + j] = u.x * u.x;
  }
  if (j == 0) out[k * (m + 1) + m] = static_cast<double>(info[k]);
}
// U_k <- U_k D_k^-1
__global__ void unscale_columns_kernel(double2 *U, const double *d, int m) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(m) * m) return;
  const int k = blockIdx.y;
  const int j = static_cast<int>(t / m);
  const double dj = d[k * m + j];
  const double f = dj > 0 ? 1.0 / dj : 0.0;
  double2 &v = U[int64_t(k) * m * m + t];
  v = make_double2(v.x * f, v.y * f);
}
// zero the strictly lower triangle of k x k matrices at stride gstride
__global__ void upper_only_kernel(double2 *M, int k, int64_t gstride) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(k) * k) return;
  const int j = static_cast<int>(t / k), i = static_cast<int>(t - int64_t(j) * k);
  if (i > j) M[blockIdx.y * gstride + t] = make_double2(0.0, 0.0);
}
// V_k <- V_k diag(max(mu_k, floor))^-1/2 (reference :66-77)
__global__ void scale_columns_kernel(double2 *u, const double *mu, int m, double floor_value) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(m) * m) return;
  const int k = blockIdx.y;
  const int col = static_cast<int>(t / m);
  const double f = 1.0 / sqrt(fmax(mu[k * m + col], floor_value));
  double2 &v = u[int64_t(k) * m * m + t];
  v = make_double2(v.x * f, v.y * f);
}

// b_k <- a_k - b_k for m columns of N rows (a at stride za, b at stride zb)
__global__ void difference_kernel(const double2 *a, int64_t za, double2 *b, int64_t zb, int N, int m, int64_t ld) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(N) * m) return;
  const int k = blockIdx.y;
  const int c = static_cast<int>(t / N), g = static_cast<int>(t - int64_t(c) * N);
  const double2 x = a[k * za + g + c * ld];
  double2 &y = b[k * zb + g + c * ld];
  y = make_double2(x.x - y.x, x.y - y.y);
}

// strided precision conversion of m columns of N rows for every k point
template <typename VI, typename VO>
__global__ void convert_kernel(const VI *in, int64_t ldi, int64_t ski, VO *out, int64_t ldo, int64_t sko, int N,
                               int m) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(N) * m) return;
  const int k = blockIdx.y;
  const int b = static_cast<int>(t / N), g = static_cast<int>(t - int64_t(b) * N);
  const VI v = in[k * ski + g + b * ldi];
  VO o;
  o.x = v.x;
  o.y = v.y;
  out[k * sko + g + b * ldo] = o;
}

}  // namespace

// ---------------------------------------------------------------------------
// operator
// ---------------------------------------------------------------------------

template <typename R>
void KMeshOperator<R>::build(const System &s, const std::vector<std::array<double, 3>> &kpoints, hipStream_t st) {
  using V = cplx_t<R>;
  sys = &s;
  nk = static_cast<int>(kpoints.size());
  np = s.nproj;
  if (nk < 1 || nk > 65535) throw Error(kInvalidArgument, "k-point count must be 1..65535");
  npw.assign(nk, 0);
  nmax = 0;
  for (int j = 0; j < nk; ++j) {
    npw[j] = static_cast<int>(build_basis(s.crystal, s.charge.shape, s.encut, kpoints[j].data()).npw());
    nmax = std::max(nmax, npw[j]);
  }
  B.resize(size_t(nk) * std::max(np, 1) * nmax);
  kin.resize(size_t(nk) * nmax);
  HALF_HIP(hipMemsetAsync(B.get(), 0, sizeof(V) * B.size(), st));
  HALF_HIP(hipMemsetAsync(kin.get(), 0, sizeof(R) * kin.size(), st));
  const bool mixed = np > 0 && np <= kSkinnyMaxProjectors;
  if (mixed) {
    BD.resize(B.size());
    BQ.resize(B.size());
    HALF_HIP(hipMemsetAsync(BD.get(), 0, sizeof(V) * BD.size(), st));
    HALF_HIP(hipMemsetAsync(BQ.get(), 0, sizeof(V) * BQ.size(), st));
  }
  maps.clear();
  maps.reserve(nk);
  kinetic.assign(nk, {});
  std::vector<LocalTables> tabs(nk);
  lb = LocalBatch{};
  lb.nk = nk;
  lb.contiguous = true;
  for (int j = 0; j < nk; ++j) {
    KOperator<R> op;
    op.build(s, kpoints[j].data(), st);
    if (np > 0)
      HALF_HIP(hipMemcpy2DAsync(B.get() + size_t(j) * np * nmax, sizeof(V) * nmax, op.B.get(), sizeof(V) * op.ldb,
                                sizeof(V) * op.n, np, hipMemcpyDeviceToDevice, st));
    if (mixed) {
      op.mixed_blocks(st);
      HALF_HIP(hipMemcpy2DAsync(BD.get() + size_t(j) * np * nmax, sizeof(V) * nmax, op.BD.get(),
                                sizeof(V) * op.ldb, sizeof(V) * op.n, np, hipMemcpyDeviceToDevice, st));
      HALF_HIP(hipMemcpy2DAsync(BQ.get() + size_t(j) * np * nmax, sizeof(V) * nmax, op.BQ.get(),
                                sizeof(V) * op.ldb, sizeof(V) * op.n, np, hipMemcpyDeviceToDevice, st));
    }
    HALF_HIP(hipMemcpyAsync(kin.get() + size_t(j) * nmax, op.kinetic.get(), sizeof(R) * op.n,
                            hipMemcpyDeviceToDevice, st));
    kinetic[j] = op.basis.kinetic;
    lb.max_lines = std::max(lb.max_lines, op.lmap.nlines);
    lb.max_planes = std::max(lb.max_planes, op.lmap.nplanes);
    lb.contiguous = lb.contiguous && op.lmap.planes_contiguous;
    maps.push_back(std::move(op.lmap));
    tabs[j] = maps.back().tables();
    HALF_HIP(hipStreamSynchronize(st));  // the copies read op's arrays
  }
  dtabs.upload(tabs, st);
  lb.tabs = dtabs.get();
  lb.kin_k = nmax;
  HALF_HIP(hipStreamSynchronize(st));
}

template <typename R>
void KMeshOperator<R>::apply(int nb, const cplx_t<R> *psi, int64_t ld, int64_t sk_in, cplx_t<R> *hpsi,
                             cplx_t<R> *spsi, int64_t ldo, int64_t sk_out, hipStream_t st) {
  if (nb <= 0) return;
  if (ld < nmax || ldo < nmax) throw Error(kInvalidArgument, "k-mesh states need nmax rows");
  const int64_t nr = sys->fft.nr;
  boxes.ensure(size_t(nk) * nb * nr);
  LocalBatch kb = lb;
  kb.psi_k = sk_in;
  kb.out_k = sk_out;
  kb.box_k = int64_t(nb) * nr;
  const R *veff;
  const cplx_t<R> *vx;
  if constexpr (std::is_same<R, double>::value) {
    veff = sys->veff.get();
    vx = sys->veff_x.get();
  } else {
    veff = sys->veff32.get();
    vx = sys->veff_x32.get();
  }
  if (np > 0 && operator_overlap() && !trace_enabled()) {
    // projector stage on the second stream, added by the local stage's last
    // pass (as KOperator::apply does)
    const int64_t cnt = int64_t(np) * nb;
    coeff.ensure(size_t(cnt) * nk);
    dcoeff.ensure(size_t(cnt) * nk);
    qcoeff.ensure(size_t(cnt) * nk);
    nl_h.ensure(size_t(nk) * nb * nmax);
    nl_s.ensure(size_t(nk) * nb * nmax);
    aux.ensure();
    HALF_HIP(hipEventRecord(aux.fork, st));
    HALF_HIP(hipStreamWaitEvent(aux.stream, aux.fork, 0));
    if (projector_skinny(np, nb) && BD.size()) {
      projector_coefficients_batched<R>(nk, nmax, np, nb, B.get(), nmax, int64_t(nmax) * np, psi, ld, sk_in,
                                        coeff.get(), np, cnt, cwork, aux.stream);
      projector_expand_batched<R>(nk, nmax, np, nb, BD.get(), BQ.get(), nmax, int64_t(nmax) * np, coeff.get(), np,
                                  cnt, false, nl_h.get(), nl_s.get(), nmax, int64_t(nb) * nmax, aux.stream);
    } else {
      gram_batched<R>(nk, nmax, np, nb, B.get(), nmax, int64_t(nmax) * np, psi, ld, sk_in, coeff.get(), np, cnt,
                      nullptr, 0, 0, nullptr, 0, 0, work, aux.stream);
      mix_projector_coefficients<R>(*sys, nk * nb, coeff.get(), dcoeff.get(), qcoeff.get(), aux.stream);
      gemm_tall_batched<R>(nk, nmax, nb, np, R(1), B.get(), nmax, int64_t(nmax) * np, dcoeff.get(), np, cnt, R(0),
                           nl_h.get(), nmax, int64_t(nb) * nmax, qcoeff.get(), np, cnt, nl_s.get(), nmax,
                           int64_t(nb) * nmax, aux.stream);
    }
    HALF_HIP(hipEventRecord(aux.join, aux.stream));
    LocalAddend<R> add;
    add.h = nl_h.get();
    add.s = nl_s.get();
    add.ld = nmax;
    add.kstride = int64_t(nb) * nmax;
    add.wait = aux.join;
    local_apply_batch<R>(sys->fft, kb, veff, vx, kin.get(), nb, psi, ld, hpsi, spsi, ldo, boxes.get(), st, add);
    return;
  }
  local_apply_batch<R>(sys->fft, kb, veff, vx, kin.get(), nb, psi, ld, hpsi, spsi, ldo, boxes.get(), st);
  if (np > 0) {
    const int64_t cnt = int64_t(np) * nb;
    coeff.ensure(size_t(cnt) * nk);
    dcoeff.ensure(size_t(cnt) * nk);
    qcoeff.ensure(size_t(cnt) * nk);
    if (projector_skinny(np, nb) && BD.size()) {
      trace_mark("proj_coeff_mfma", st);
      projector_coefficients_batched<R>(nk, nmax, np, nb, B.get(), nmax, int64_t(nmax) * np, psi, ld, sk_in,
                                        coeff.get(), np, cnt, cwork, st);
      trace_mark("proj_expand_mfma", st);
      projector_expand_batched<R>(nk, nmax, np, nb, BD.get(), BQ.get(), nmax, int64_t(nmax) * np, coeff.get(), np,
                                  cnt, true, hpsi, spsi, ldo, sk_out, st);
      trace_close(st);
      return;
    }
    trace_mark("proj_coeff_mfma", st);
    gram_batched<R>(nk, nmax, np, nb, B.get(), nmax, int64_t(nmax) * np, psi, ld, sk_in, coeff.get(), np, cnt,
                    nullptr, 0, 0, nullptr, 0, 0, work, st);
    trace_mark("proj_mix", st);
    mix_projector_coefficients<R>(*sys, nk * nb, coeff.get(), dcoeff.get(), qcoeff.get(), st);
    trace_mark("proj_expand_mfma", st);
    gemm_tall_batched<R>(nk, nmax, nb, np, R(1), B.get(), nmax, int64_t(nmax) * np, dcoeff.get(), np, cnt, R(1), hpsi,
                         ldo, sk_out, qcoeff.get(), np, cnt, spsi, ldo, sk_out, st);
  }
  trace_close(st);
}

template struct KMeshOperator<double>;
template struct KMeshOperator<float>;

// ---------------------------------------------------------------------------
// solver
// ---------------------------------------------------------------------------

void solve_acc_kmesh(Blas &blas, KMeshOperator<double> &op64, KMeshOperator<float> *op32, int nbands,
                     const AccControls &ctl, const KMeshControls &kc, std::vector<double> &eigenvalues,
                     SolveResult &res) {
  hipStream_t s = blas.stream;
  const int nk = op64.nk, N = op64.nmax;
  const int mreq = nbands;
  const int m = nbands + std::max(0, ctl.extra);
  const int m2 = 2 * m;
  if (mreq < 1) throw Error(kInvalidArgument, "invalid requested band count");
  for (int j = 0; j < nk; ++j)
    if (2 * m > op64.npw[j]) throw Error(kInvalidArgument, "ACC solver requires NBANDS <= NPL/2 at every k point");
  if (!(ctl.tol > 0) || ctl.maxiter < 1) throw Error(kInvalidArgument, "invalid ACC solver controls");
  if (ctl.mixed && (!op32 || op32->nk != nk || op32->nmax != N))
    throw Error(kInternal, "mixed k-mesh solve needs the matching FP32 operator");
  const double t0 = now();
  const int64_t ld = N, zk = int64_t(N) * m2, rk = int64_t(N) * m;
  const int64_t gk = int64_t(m2) * m2;  // Gram stride per k point
  // every subspace problem (orders m .. 2m, metrics up to m) fits the
  // one-workgroup kernels: no host decision inside an iteration
  const bool small = m2 <= kSmallMaxOrder && solver_smalleig() != 0;
  const bool force_metric_eig = solver_smalleig() == 3;
  DeviceArray<double2> Z(size_t(nk) * zk), HZ(size_t(nk) * zk), SZ(size_t(nk) * zk), Rr(size_t(nk) * rk),
      tmp(size_t(nk) * rk), cross(size_t(nk) * m * m), metric(size_t(nk) * m * m), hsub(size_t(nk) * gk),
      ssub(size_t(nk) * gk), gvec(size_t(nk) * gk), tau(size_t(nk) * m2), gwork;
  DeviceArray<double> w(size_t(nk) * m2), E(size_t(nk) * m2), mu(size_t(nk) * m), res2(size_t(nk) * m),
      dscale(size_t(nk) * m), piv(size_t(nk) * (m + 1));
  DeviceArray<int> dseeds(size_t(nk) * m), dcols(size_t(nk) * m);
  // device-side status per k point: [0, nk) Rayleigh-Ritz Cholesky, [nk, 2nk)
  // its QL iterations, [2nk, 3nk) metric in the eigen-decomposition form,
  // [3nk, 4nk) that form's QL iterations (written only where it ran)
  DeviceArray<int> flags(size_t(4) * nk);
  DeviceArray<rocblas_int> info(size_t(3) * nk);
  DeviceArray<float2> p32, hp32, sp32;
  for (auto *a : {&Z, &HZ, &SZ, &Rr, &tmp})
    HALF_HIP(hipMemsetAsync(a->get(), 0, sizeof(double2) * a->size(), s));
  HALF_HIP(hipMemsetAsync(flags.get(), 0, sizeof(int) * flags.size(), s));
  if (ctl.mixed) {
    p32.resize(size_t(nk) * rk);
    hp32.resize(size_t(nk) * rk);
    sp32.resize(size_t(nk) * rk);
    for (auto *a : {&p32, &hp32, &sp32}) HALF_HIP(hipMemsetAsync(a->get(), 0, sizeof(float2) * a->size(), s));
  }
  double2 *X = Z.get(), *P = Z.get() + rk;
  double2 *HX = HZ.get(), *HP = HZ.get() + rk;
  double2 *SX = SZ.get(), *SP = SZ.get() + rk;
  const rocblas_double_complex one{1, 0}, zero{0, 0};
  const dim3 gNm(blocks_for(int64_t(N) * m), nk);
  bool fp64_phase = !ctl.mixed;
  double fp32_floor = 0;
  res.applies = 0;
  res.metric_fallbacks = 0;
  res.history.clear();

  // H and S applied to `cnt` columns of every k point
  auto apply = [&](int cnt, double2 *in, double2 *hout, double2 *sout) {
    if (fp64_phase) {
      op64.apply(cnt, in, ld, zk, hout, sout, ld, zk, s);
    } else {
      const dim3 gNc(blocks_for(int64_t(N) * cnt), nk);
      hipLaunchKernelGGL((convert_kernel<double2, float2>), gNc, dim3(kThreads), 0, s, in, ld, zk, p32.get(), ld, rk,
                         N, cnt);
      op32->apply(cnt, p32.get(), ld, rk, hp32.get(), sp32.get(), ld, rk, s);
      hipLaunchKernelGGL((convert_kernel<float2, double2>), gNc, dim3(kThreads), 0, s, hp32.get(), ld, rk, hout, ld,
                         zk, N, cnt);
      hipLaunchKernelGGL((convert_kernel<float2, double2>), gNc, dim3(kThreads), 0, s, sp32.get(), ld, rk, sout, ld,
                         zk, N, cnt);
      HALF_HIP(hipGetLastError());
    }
    res.applies += int64_t(cnt) * nk;
  };
  // k x k generalized problems (A_j = hsub_j, B_j = ssub_j at stride gstride):
  // B = U^H U, C = U^-H A U^-1, C v = e v, eigenvectors U^-1 V of the lowest m
  // into gvec (k x m), all k values ascending into w (stride m2)
  auto generalized_small = [&](int k, int64_t gstride) {
    ContractFp32Scope fp64_products(false);  // the subspace problem stays FP64 in either phase
    potrf_inverse_batched(k, nk, ssub.get(), k, gstride, ssub.get(), k, gstride, flags.get(), false, nullptr, 0, s);
    gemm_tall_batched<double>(nk, k, k, k, 1.0, hsub.get(), k, gstride, ssub.get(), k, gstride, 0.0, gvec.get(), k,
                              gstride, nullptr, 0, 0, nullptr, 0, 0, s);
    gram_batched<double>(nk, k, k, k, ssub.get(), k, gstride, gvec.get(), k, gstride, hsub.get(), k, gstride, nullptr,
                         0, 0, nullptr, 0, 0, gwork, s, true);
    hetrd_upper(k, nk, hsub.get(), k, gstride, w.get(), m2, E.get(), m2, tau.get(), m2, s);
    tridiag_eig_batched(k, nk, w.get(), m2, E.get(), m2, hsub.get(), k, gstride, tau.get(), m2, m, ssub.get(), k,
                        gstride, 0.0, gvec.get(), k, gstride, w.get(), m2, flags.get() + nk, nullptr, s);
  };
  // the same with rocSOLVER (orders above kSmallMaxOrder): U^-1 formed
  // explicitly (batched triangular solves are several times slower than
  // trtri plus products at these sizes); statuses checked on the host
  auto generalized_library = [&](int k, int64_t gstride) {
    KM_BLAS(rocsolver_zpotrf_strided_batched(blas.h, rocblas_fill_upper, k, rz(ssub.get()), k, gstride, info.get(),
                                             nk));
    KM_BLAS(rocsolver_ztrtri_strided_batched(blas.h, rocblas_fill_upper, rocblas_diagonal_non_unit, k,
                                             rz(ssub.get()), k, gstride, info.get() + nk, nk));
    hipLaunchKernelGGL(upper_only_kernel, dim3(blocks_for(int64_t(k) * k), nk), dim3(kThreads), 0, s, ssub.get(), k,
                       gstride);
    HALF_HIP(hipGetLastError());
    KM_BLAS(rocblas_zgemm_strided_batched(blas.h, rocblas_operation_none, rocblas_operation_none, k, k, k, &one,
                                          rz(hsub.get()), k, gstride, rz(ssub.get()), k, gstride, &zero,
                                          rz(gvec.get()), k, gstride, nk));
    KM_BLAS(rocblas_zgemm_strided_batched(blas.h, rocblas_operation_conjugate_transpose, rocblas_operation_none, k, k,
                                          k, &one, rz(ssub.get()), k, gstride, rz(gvec.get()), k, gstride, &zero,
                                          rz(hsub.get()), k, gstride, nk));
    KM_BLAS(rocsolver_zheevd_strided_batched(blas.h, rocblas_evect_original, rocblas_fill_upper, k, rz(hsub.get()), k,
                                             gstride, w.get(), m2, E.get(), m2, info.get() + 2 * nk, nk));
    KM_BLAS(rocblas_zgemm_strided_batched(blas.h, rocblas_operation_none, rocblas_operation_none, k, m, k, &one,
                                          rz(ssub.get()), k, gstride, rz(hsub.get()), k, gstride, &zero,
                                          rz(gvec.get()), k, gstride, nk));
    std::vector<rocblas_int> h(3 * nk);
    HALF_HIP(hipMemcpyAsync(h.data(), info.get(), sizeof(rocblas_int) * h.size(), hipMemcpyDeviceToHost, s));
    HALF_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < nk; ++j) {
      if (h[j] != 0 || h[nk + j] != 0)
        throw Error(kInternal, "k-mesh ACC subspace solve failed at k point " + std::to_string(j) +
                                   ": overlap not positive definite");
      if (h[2 * nk + j] != 0)
        throw Error(kInternal, "k-mesh ACC subspace eigensolve failed at k point " + std::to_string(j));
    }
  };
  auto generalized = [&](int k, int64_t gstride) {
    if (small)
      generalized_small(k, gstride);
    else
      generalized_library(k, gstride);
  };
  // the device-side statuses of the small kernels since the last check
  // (read with the residual norms); failures end the solve
  std::vector<int> hflags(size_t(4) * nk);
  bool metric_ran = false;  // an iteration (with its metric factorisation) since the last check
  auto check_flags = [&]() {
    bool fallback = false;
    for (int j = 0; j < nk; ++j) {
      if (hflags[j] != 0)
        throw Error(kInternal, "k-mesh ACC subspace solve failed at k point " + std::to_string(j) +
                                   ": overlap not positive definite");
      if (hflags[nk + j] != 0)
        throw Error(kInternal, "k-mesh ACC subspace eigensolve failed at k point " + std::to_string(j));
      if (metric_ran && hflags[2 * nk + j] != 0) {
        fallback = true;
        if (hflags[3 * nk + j] != 0) throw Error(kInternal, "k-mesh ACC metric eigensolve failed");
      }
    }
    if (fallback && metric_ran) ++res.metric_fallbacks;
    metric_ran = false;
  };
  // blk(:, :c) <- blk(:, :k) W(:k, :c) for every k point: in place on the
  // matrix cores when the c output columns fit one tile, else through tmp
  auto right_multiply = [&](double2 *blk, const double2 *W, int k, int64_t ldw, int64_t sw, int c) {
    if (c <= gemm_tall_inplace_max) {
      gemm_tall_batched<double>(nk, N, c, k, 1.0, blk, ld, zk, W, ldw, sw, 0.0, blk, ld, zk, nullptr, 0, 0, nullptr,
                                0, 0, s);
    } else {
      gemm_tall_batched<double>(nk, N, c, k, 1.0, blk, ld, zk, W, ldw, sw, 0.0, tmp.get(), ld, rk, nullptr, 0, 0,
                                nullptr, 0, 0, s);
      HALF_HIP(hipMemcpy2DAsync(blk, sizeof(double2) * zk, tmp.get(), sizeof(double2) * rk,
                                sizeof(double2) * int64_t(N) * c, nk, hipMemcpyDeviceToDevice, s));
    }
  };
  // X, HX, SX <- first k columns of (Z, HZ, SZ) times the lowest m Ritz vectors
  auto rotate = [&](int k, int64_t gstride) {
    for (double2 *blk : {Z.get(), HZ.get(), SZ.get()}) right_multiply(blk, gvec.get(), k, k, gstride, m);
  };
  auto rayleigh_ritz_x = [&]() {
    gram_batched<double>(nk, N, m, m, X, ld, zk, HX, ld, zk, hsub.get(), m, int64_t(m) * m, SX, ld, zk, ssub.get(), m,
                         int64_t(m) * m, gwork, s, true);
    generalized(m, int64_t(m) * m);
    rotate(m, int64_t(m) * m);
  };
  // R = HX - SX diag(w) and its squared column norms
  auto residual_norms = [&]() {
    trace_mark("acc_residual", s);
    hipLaunchKernelGGL(residual_kernel, gNm, dim3(kThreads), 0, s, HX, SX, zk, w.get(), int64_t(m2), Rr.get(), rk, N,
                       m, ld);
    hipLaunchKernelGGL(column_norm2_kernel, dim3(nk * m), dim3(kThreads), 0, s, Rr.get(), ld, N, res2.get());
    HALF_HIP(hipGetLastError());
  };

  // initial subspace: unit vectors at each k point's m lowest kinetic energies
  {
    std::vector<int> idx(size_t(nk) * m);
    for (int j = 0; j < nk; ++j) {
      auto order = kinetic_order(op64.kinetic[j]);
      std::copy(order.begin(), order.begin() + m, idx.begin() + size_t(j) * m);
    }
    DeviceArray<int> d_idx;
    d_idx.upload(idx, s);
    hipLaunchKernelGGL(init_states_kernel, gNm, dim3(kThreads), 0, s, X, ld, zk, d_idx.get(), N, m);
    HALF_HIP(hipGetLastError());
    DeviceArray<int> d_npw;
    apply(m, X, HX, SX);
    if (!fp64_phase) {
      // FP32 residual floor: the FP32 operator's error on a dense probe
      // block per k point, |dH p| + max|eps| |dS p| per unit column (P, HP,
      // SP, Rr and tmp are free until the loop; padding rows stay zero).
      // Unit vectors would understate it: their projector coefficients
      // involve no accumulation.
      d_npw.upload(op64.npw, s);
      hipLaunchKernelGGL(phase_block_kernel, gNm, dim3(kThreads), 0, s, P, zk, d_npw.get(), N, m, ld);
      HALF_HIP(hipGetLastError());
      apply(m, P, HP, SP);
      op64.apply(m, P, ld, zk, Rr.get(), tmp.get(), ld, rk, s);
      res.applies += int64_t(m) * nk;
      hipLaunchKernelGGL(difference_kernel, gNm, dim3(kThreads), 0, s, HP, zk, Rr.get(), rk, N, m, ld);
      hipLaunchKernelGGL(difference_kernel, gNm, dim3(kThreads), 0, s, SP, zk, tmp.get(), rk, N, m, ld);
      hipLaunchKernelGGL(column_norm2_kernel, dim3(nk * m), dim3(kThreads), 0, s, Rr.get(), ld, N, res2.get());
      hipLaunchKernelGGL(column_norm2_kernel, dim3(nk * m), dim3(kThreads), 0, s, tmp.get(), ld, N, mu.get());
      HALF_HIP(hipGetLastError());
    }
    rayleigh_ritz_x();
    if (!fp64_phase) {
      std::vector<double> eh(size_t(nk) * m), es(size_t(nk) * m), ev(size_t(nk) * m2);
      HALF_HIP(hipMemcpyAsync(eh.data(), res2.get(), sizeof(double) * eh.size(), hipMemcpyDeviceToHost, s));
      HALF_HIP(hipMemcpyAsync(es.data(), mu.get(), sizeof(double) * es.size(), hipMemcpyDeviceToHost, s));
      HALF_HIP(hipMemcpyAsync(ev.data(), w.get(), sizeof(double) * ev.size(), hipMemcpyDeviceToHost, s));
      HALF_HIP(hipStreamSynchronize(s));
      std::vector<double> e(size_t(nk) * m);
      for (int j = 0; j < nk; ++j) {
        std::copy(ev.begin() + size_t(j) * m2, ev.begin() + size_t(j) * m2 + m, e.begin() + size_t(j) * m);
        for (int b = 0; b < m; ++b) {  // per unit column: |p|^2 = npw_k
          eh[size_t(j) * m + b] /= op64.npw[j];
          es[size_t(j) * m + b] /= op64.npw[j];
        }
      }
      fp32_floor = fp32_floor_estimate(eh, es, e);
    }
    residual_norms();
    HALF_HIP(hipStreamSynchronize(s));
  }

  // one iteration after the host's decisions (correction columns cols and
  // seeds already on the device): preconditioned corrections, their H.psi,
  // S-orthogonalisation against X, normalisation in the S metric,
  // Rayleigh-Ritz in span[X P] for every k point, the rotations, and the
  // residual norms the next decision reads
  auto iteration = [&](int np) {
    const dim3 gNp(blocks_for(int64_t(N) * np), nk), gpp(blocks_for(int64_t(np) * np), nk), gp(blocks_for(np), nk);
    const int64_t pstride = int64_t(np) * np;  // metric stride at width np
    trace_mark("acc_precondition", s);
    hipLaunchKernelGGL(precondition_kernel, gNp, dim3(kThreads), 0, s, Rr.get(), rk, op64.kin.get(),
                       int64_t(op64.nmax), w.get(), int64_t(m2), dcols.get(), dseeds.get(), P, zk, N, m, np, ld);
    HALF_HIP(hipGetLastError());
    apply(np, P, HP, SP);
    // FP32 phase: the dense algebra below with FP32 products (as solve_acc)
    ContractFp32Scope fp32_products(!fp64_phase && solver_fp32_algebra());
    // S-orthogonalise against X (cross: m x np per k point)
    trace_mark("acc_orthogonalise", s);
    gram_batched<double>(nk, N, m, np, X, ld, zk, SP, ld, zk, cross.get(), m, int64_t(m) * m, nullptr, 0, 0, nullptr,
                         0, 0, gwork, s);
    for (auto pr : {std::make_pair(X, P), std::make_pair(HX, HP), std::make_pair(SX, SP)})
      gemm_tall_batched<double>(nk, N, np, m, -1.0, pr.first, ld, zk, cross.get(), m, int64_t(m) * m, 1.0,
                                pr.second, ld, zk, nullptr, 0, 0, nullptr, 0, 0, s);
    // normalise in the S metric: W = U^-1 of the unit-diagonal Cholesky
    // factor, or the reference's eigen-decomposition form V mu^-1/2 where a
    // k point's factor is ill-conditioned
    trace_mark("acc_metric_cholesky", s);
    gram_batched<double>(nk, N, np, np, P, ld, zk, SP, ld, zk, metric.get(), np, pstride, nullptr, 0, 0, nullptr, 0,
                         0, gwork, s, true);
    if (small) {
      // the copy for the eigen-decomposition form goes to `cross` (free now)
      potrf_inverse_batched(np, nk, metric.get(), np, pstride, metric.get(), np, pstride, flags.get() + 2 * nk, true,
                            cross.get(), pstride, s, force_metric_eig);
      trace_mark("acc_metric_zheevd", s);
      hetrd_upper(np, nk, cross.get(), np, pstride, mu.get(), m, E.get(), m2, tau.get(), m2, s, flags.get() + 2 * nk);
      tridiag_eig_batched(np, nk, mu.get(), m, E.get(), m2, cross.get(), np, pstride, tau.get(), m2, np, nullptr, 0, 0,
                          1e-14, metric.get(), np, pstride, nullptr, 0, flags.get() + 3 * nk, flags.get() + 2 * nk, s);
      trace_mark("acc_normalise", s);
      for (double2 *blk : {P, HP, SP}) right_multiply(blk, metric.get(), np, np, pstride, np);
    } else {
      hipLaunchKernelGGL(unit_diagonal_kernel, gp, dim3(kThreads), 0, s, metric.get(), dscale.get(), np);
      hipLaunchKernelGGL(scale_both_kernel, gpp, dim3(kThreads), 0, s, metric.get(), dscale.get(), np);
      HALF_HIP(hipGetLastError());
      KM_BLAS(rocsolver_zpotrf_strided_batched(blas.h, rocblas_fill_upper, np, rz(metric.get()), np, pstride,
                                               info.get(), nk));
      hipLaunchKernelGGL(pivots_kernel, gp, dim3(kThreads), 0, s, metric.get(), info.get(), piv.get(), np);
      HALF_HIP(hipGetLastError());
      std::vector<double> pivh(size_t(nk) * (np + 1));
      HALF_HIP(hipMemcpyAsync(pivh.data(), piv.get(), sizeof(double) * pivh.size(), hipMemcpyDeviceToHost, s));
      HALF_HIP(hipStreamSynchronize(s));
      bool chol_ok = true;
      for (int j = 0; j < nk && chol_ok; ++j) {
        const double *pj = pivh.data() + size_t(j) * (np + 1);
        chol_ok = pj[np] == 0.0;
        for (int b = 0; b < np && chol_ok; ++b) chol_ok = std::isfinite(pj[b]) && pj[b] > 1e-10;
      }
      if (chol_ok) {
        hipLaunchKernelGGL(unscale_columns_kernel, gpp, dim3(kThreads), 0, s, metric.get(), dscale.get(), np);
        HALF_HIP(hipGetLastError());
        trace_mark("acc_normalise", s);
        // P <- P U^-1 with U^-1 explicit (a triangular factor whose scaled
        // pivots all exceed 1e-10)
        KM_BLAS(rocsolver_ztrtri_strided_batched(blas.h, rocblas_fill_upper, rocblas_diagonal_non_unit, np,
                                                 rz(metric.get()), np, pstride, info.get(), nk));
        hipLaunchKernelGGL(upper_only_kernel, gpp, dim3(kThreads), 0, s, metric.get(), np, pstride);
        HALF_HIP(hipGetLastError());
        for (double2 *blk : {P, HP, SP}) right_multiply(blk, metric.get(), np, np, pstride, np);
      } else {
        ++res.metric_fallbacks;
        gram_batched<double>(nk, N, np, np, P, ld, zk, SP, ld, zk, metric.get(), np, pstride, nullptr, 0, 0, nullptr,
                             0, 0, gwork, s, true);
        trace_mark("acc_metric_zheevd", s);
        KM_BLAS(rocsolver_zheevd_strided_batched(blas.h, rocblas_evect_original, rocblas_fill_upper, np,
                                                 rz(metric.get()), np, pstride, mu.get(), np, E.get(), m2,
                                                 info.get(), nk));
        std::vector<rocblas_int> h(nk);
        HALF_HIP(hipMemcpyAsync(h.data(), info.get(), sizeof(rocblas_int) * nk, hipMemcpyDeviceToHost, s));
        HALF_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < nk; ++j)
          if (h[j] != 0) throw Error(kInternal, "k-mesh ACC metric eigensolve failed");
        hipLaunchKernelGGL(scale_columns_kernel, gpp, dim3(kThreads), 0, s, metric.get(), mu.get(), np, 1e-14);
        HALF_HIP(hipGetLastError());
        for (double2 *blk : {P, HP, SP}) right_multiply(blk, metric.get(), np, np, pstride, np);
      }
    }
    // Rayleigh-Ritz in span[X P] for every k point
    const int kz = m + np;
    trace_mark("acc_projection_gram", s);
    gram_batched<double>(nk, N, kz, kz, Z.get(), ld, zk, HZ.get(), ld, zk, hsub.get(), kz, gk, SZ.get(), ld, zk,
                         ssub.get(), kz, gk, gwork, s, true);
    trace_mark("acc_subspace_eigensolve", s);
    generalized(kz, gk);
    trace_mark("acc_rotate_gemm", s);
    rotate(kz, gk);
    residual_norms();
    trace_close(s);
  };

  // Captured iterations, one graph per (precision phase, correction width):
  // a width's first iteration runs eagerly (every workspace then has its
  // size for that shape), is recorded right after and replayed from the next
  // iteration of that shape on.  A graph also records the contraction
  // workspace it was captured with and is captured again if that moved.
  const bool use_graph = kc.graph && small && !trace_enabled();
  struct Captured {
    int phase, np;
    const void *work;
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<Captured> graphs;
  struct GraphGuard {
    std::vector<Captured> &g;
    ~GraphGuard() {
      for (auto &c : g) {
        if (c.exec) (void)hipGraphExecDestroy(c.exec);
        if (c.graph) (void)hipGraphDestroy(c.graph);
      }
    }
  } guard{graphs};
  auto run_iteration = [&](int np) {
    metric_ran = true;
    if (!use_graph) {
      iteration(np);
      return;
    }
    const int phase = fp64_phase ? 1 : 0;
    for (auto it = graphs.begin(); it != graphs.end(); ++it) {
      if (it->phase != phase || it->np != np) continue;
      if (it->work == static_cast<const void *>(gwork.get())) {
        HALF_HIP(hipGraphLaunch(it->exec, s));
        res.applies += int64_t(np) * nk;
        return;
      }
      HALF_HIP(hipGraphExecDestroy(it->exec));  // its last replay has ended at a residual sync
      HALF_HIP(hipGraphDestroy(it->graph));
      graphs.erase(it);
      break;
    }
    iteration(np);  // this iteration's work, eager
    HALF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int64_t applies0 = res.applies;
    const int fallbacks0 = res.metric_fallbacks;
    Captured c{phase, np, gwork.get(), nullptr, nullptr};
    try {
      iteration(np);
    } catch (...) {
      hipGraph_t broken = nullptr;  // leave the stream usable
      if (hipStreamEndCapture(s, &broken) == hipSuccess && broken) (void)hipGraphDestroy(broken);
      throw;
    }
    res.applies = applies0;
    res.metric_fallbacks = fallbacks0;
    HALF_HIP(hipStreamEndCapture(s, &c.graph));
    if (c.work != static_cast<const void *>(gwork.get())) {
      (void)hipGraphDestroy(c.graph);  // a workspace moved while recording: replay nothing of it
      throw Error(kInternal, "k-mesh ACC: contraction workspace resized during graph capture");
    }
    const hipError_t ie = hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      (void)hipGraphDestroy(c.graph);
      HALF_HIP(ie);
    }
    graphs.push_back(c);
  };

  std::vector<double> r2(size_t(nk) * m);
  std::vector<int> seeds(size_t(nk) * m), cols(size_t(nk) * m), pseeds(size_t(nk) * m);
  const bool lock = ctl.locking();
  int64_t order_sum = 0;
  int rr_steps = 0;
  StallWatch fp32_stalled;
  double max_res = HUGE_VAL;
  int iter = 0, fp64_iters = 0;
  for (iter = 1; iter <= ctl.maxiter; ++iter) {
    trace_mark("acc_norms_to_host", s);
    HALF_HIP(hipMemcpyAsync(r2.data(), res2.get(), sizeof(double) * r2.size(), hipMemcpyDeviceToHost, s));
    if (small)
      HALF_HIP(hipMemcpyAsync(hflags.data(), flags.get(), sizeof(int) * hflags.size(), hipMemcpyDeviceToHost, s));
    HALF_HIP(hipStreamSynchronize(s));
    if (small) check_flags();
    max_res = 0;
    for (int j = 0; j < nk; ++j)
      for (int b = 0; b < m; ++b) {
        const double r = std::sqrt(std::max(0.0, r2[size_t(j) * m + b]));
        if (b < mreq) max_res = std::max(max_res, r);
        seeds[size_t(j) * m + b] = r < std::max(ctl.tol * 1e-2, 1e-12) ? (m + b) % op64.npw[j] : -1;
      }
    res.history.push_back(max_res);
    if (fp64_phase) ++fp64_iters;
    const double fp32_target = std::max({ctl.switch_tol, ctl.tol, fp32_floor});
    const bool fp32_stall = !fp64_phase && ((max_res < kStallArm * fp32_target && fp32_stalled.update(max_res)) ||
                                            2 * iter > ctl.maxiter);
    if (!fp64_phase && (fp32_stall || max_res <= fp32_target)) {
      fp64_phase = true;
      apply(m, X, HX, SX);
      rayleigh_ritz_x();
      residual_norms();
      --iter;
      continue;
    }
    if (fp64_phase && max_res <= ctl.tol) break;
    // correction columns, one width np for all k points (the batched algebra
    // and the replayed graph need one shape): without locking every band;
    // with it, each k point's buffer bands and requested bands above the
    // current target first, np = the largest such count over the k points
    // rounded up to a multiple of 8, the rest of a k point's columns filled
    // with its other bands in order (extra columns only widen its subspace)
    const double lock_at = fp64_phase ? ctl.tol : fp32_target;
    auto active = [&](size_t o, int b) {
      return !lock || b >= mreq || std::sqrt(std::max(0.0, r2[o + b])) > lock_at;
    };
    int np = m;
    if (lock) {
      int most = 0;
      for (int j = 0; j < nk; ++j) {
        int c = 0;
        for (int b = 0; b < m; ++b) c += active(size_t(j) * m, b) ? 1 : 0;
        most = std::max(most, c);
      }
      np = std::min(m, std::max(8, (most + 7) / 8 * 8));
    }
    for (int j = 0; j < nk; ++j) {
      int c = 0;
      const size_t o = size_t(j) * m;
      for (int pass = 0; pass < 2; ++pass)
        for (int b = 0; b < m && c < np; ++b) {
          if (active(o, b) == (pass == 0)) {
            cols[o + c] = b;
            pseeds[o + c] = seeds[o + b];
            ++c;
          }
        }
    }
    order_sum += m + np;
    ++rr_steps;
    dseeds.upload(pseeds, s);
    dcols.upload(cols, s);
    run_iteration(np);
  }
  res.iterations = std::min(iter, ctl.maxiter);
  res.fp64_iterations = fp64_iters;
  res.residual = max_res;
  res.rr_order_mean = rr_steps > 0 ? double(order_sum) / rr_steps : 0.0;
  std::vector<double> wall(size_t(nk) * m2);
  HALF_HIP(hipMemcpyAsync(wall.data(), w.get(), sizeof(double) * wall.size(), hipMemcpyDeviceToHost, s));
  HALF_HIP(hipStreamSynchronize(s));
  eigenvalues.resize(size_t(nk) * mreq);
  for (int j = 0; j < nk; ++j)
    std::copy(wall.begin() + size_t(j) * m2, wall.begin() + size_t(j) * m2 + mreq,
              eigenvalues.begin() + size_t(j) * mreq);
  res.overlap_min = res.overlap_max = -1.0;
  res.seconds_solver = now() - t0;
}

}  // namespace half
