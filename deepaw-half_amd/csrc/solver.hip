// Dense and all-band eigensolvers on the device (see solver.hpp).
//
// The all-band minimiser keeps Z = [X | P], HZ = [HX | HP], SZ = [SX | SP]
// contiguous so the 2m x 2m Rayleigh-Ritz projection needs no block copies.
// Gram matrices (A^H B over the plane waves) run on the LDS-tiled MFMA
// contraction of contract.hip with a fixed-order split over plane waves;
// rotations are rocBLAS ZGEMMs.  The correction block is normalised by the
// Cholesky factor of its m x m metric (rocSOLVER zheevd of the metric when
// that factor is ill-conditioned); the 2m x 2m generalized problem H c = e S c is reduced
// with the Cholesky factor S = U^H U (zpotrf, two ztrsm) to a standard one
// solved by zheevd, and c = U^-1 y (ztrsm): the same reduction zhegvd makes,
// about two and a half times faster on the device at these sizes.  All FP64.
// Handles and workspaces live for the whole solve (the reference creates
// cuBLAS/cuSOLVER handles per call and reallocates per iteration).  The one
// host round trip per iteration is the m residual norms used for the
// convergence test and the seed selection.
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>

#include "contract.hpp"
#include "cplx.hpp"
#include "smalleig.hpp"
#include "solver.hpp"
#include "trace.hpp"
#include "tridiag.hpp"

namespace half {

namespace {

constexpr int kThreads = 256;
inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

#define HALF_BLAS(call)                                                                         \
  do {                                                                                          \
    rocblas_status st_ = (call);                                                                \
    if (st_ != rocblas_status_success)                                                          \
      throw Error(kInternal, std::string("rocBLAS/rocSOLVER status ") + std::to_string(int(st_)) + \
                                 " in " #call);                                                 \
  } while (0)

inline rocblas_double_complex *rz(double2 *p) { return reinterpret_cast<rocblas_double_complex *>(p); }
inline const rocblas_double_complex *rz(const double2 *p) {
  return reinterpret_cast<const rocblas_double_complex *>(p);
}

// C = op(A) * op(B) (+ beta C), all column-major FP64 complex
void zgemm(Blas &b, rocblas_operation ta, rocblas_operation tb, int m, int n, int k, double2 alpha,
           const double2 *A, int64_t lda, const double2 *B, int64_t ldb, double2 beta, double2 *C, int64_t ldc) {
  rocblas_double_complex al{alpha.x, alpha.y}, be{beta.x, beta.y};
  HALF_BLAS(rocblas_zgemm(b.h, ta, tb, m, n, k, &al, rz(A), static_cast<rocblas_int>(lda), rz(B),
                          static_cast<rocblas_int>(ldb), &be, rz(C), static_cast<rocblas_int>(ldc)));
}

__global__ void init_states_kernel(double2 *x, int64_t ld, const int *indices, int n, int m) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(n) * m) return;
  const int b = static_cast<int>(t / n);
  const int g = static_cast<int>(t - int64_t(b) * n);
  x[g + b * ld] = make_double2(g == indices[b] ? 1.0 : 0.0, 0.0);
}

// out column i = in column perm[i] (i < cnt), for reordering blocks of states
__global__ void gather_columns_kernel(const double2 *in, double2 *out, const int *perm, int n, int cnt, int64_t ld) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(n) * cnt) return;
  const int b = static_cast<int>(t / n);
  const int g = static_cast<int>(t - int64_t(b) * n);
  out[g + b * ld] = in[g + perm[b] * ld];
}

// R = HX - SX diag(w) (reference: src/half_cuda_iterative_solver.cuf:29-40)
__global__ void residual_kernel(const double2 *hx, const double2 *sx, const double *w, double2 *r, int n, int m,
                                int64_t ld) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(n) * m) return;
  const int b = static_cast<int>(t / n);
  const int g = static_cast<int>(t - int64_t(b) * n);
  const int64_t o = g + b * ld;
  r[o] = make_double2(hx[o].x - w[b] * sx[o].x, hx[o].y - w[b] * sx[o].y);
}

// p(g, j) = exp(i theta) with a pseudo-random phase per element (integer
// hash of g and j): dense columns of norm sqrt(n)
__global__ void phase_block_kernel(double2 *p, int n, int m, int64_t ld) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(n) * m) return;
  const int j = static_cast<int>(t / n);
  const int g = static_cast<int>(t - int64_t(j) * n);
  uint32_t h = static_cast<uint32_t>(g) * 2654435761u ^ (static_cast<uint32_t>(j) + 0x9e3779b9u) * 2246822519u;
  h ^= h >> 15;
  h *= 2654435761u;
  h ^= h >> 13;
  double sn, cs;
  sincos(6.283185307179586 * (h * 2.3283064365386963e-10), &sn, &cs);
  p[g + j * ld] = make_double2(cs, sn);
}

// zero the strictly lower triangle of a k x k matrix
__global__ void upper_only_kernel(double2 *M, int k) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(k) * k) return;
  const int j = static_cast<int>(t / k), i = static_cast<int>(t - int64_t(j) * k);
  if (i > j) M[t] = make_double2(0.0, 0.0);
}

// out[blockIdx.x] = max over this workgroup's share of the upper triangle of
// a k x k matrix of |M(i, j) - delta_ij| (|re| + |im|); a NaN entry makes the
// share NaN
__global__ void identity_deviation_kernel(const double2 *M, int k, double *out) {
  __shared__ double red[kThreads / 64];
  double mx = 0.0;
  for (int j = blockIdx.x; j < k; j += gridDim.x)
    for (int i = threadIdx.x; i <= j; i += kThreads) {
      double2 v = M[i + int64_t(j) * k];
      if (i == j) v.x -= 1.0;
      const double d = fabs(v.x) + fabs(v.y);
      if (!(d <= mx)) mx = d;
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_xor(mx, o);
    if (!(u <= mx)) mx = u;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < kThreads / 64; ++w)
      if (!(red[w] <= r)) r = red[w];
    out[blockIdx.x] = r;
  }
}

// Column j < nv of V (order k) and its value w[j], with b = Re(v^H T(:, j))
// for T = B V: w[j] /= b and v /= sqrt(b) (one workgroup per column, fixed
// summation order)
__global__ void metric_rayleigh_kernel(double2 *V, const double2 *T, int k, double *w) {
  __shared__ double red[kThreads / 64];
  __shared__ double s_b;
  const int j = blockIdx.x;
  double2 *v = V + int64_t(j) * k;
  const double2 *t = T + int64_t(j) * k;
  double acc = 0.0;
  for (int i = threadIdx.x; i < k; i += kThreads) acc += v[i].x * t[i].x + v[i].y * t[i].y;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.0;
    for (int q = 0; q < kThreads / 64; ++q) b += red[q];
    s_b = b;
    w[j] /= b;
  }
  __syncthreads();
  const double f = 1.0 / sqrt(s_b);
  for (int i = threadIdx.x; i < k; i += kThreads) v[i] = make_double2(v[i].x * f, v[i].y * f);
}

// b <- a - b for m columns of n rows
__global__ void difference_kernel(const double2 *a, double2 *b, int n, int m, int64_t ld) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(n) * m) return;
  const int j = static_cast<int>(t / n);
  const int64_t o = t - int64_t(j) * n + j * ld;
  b[o] = make_double2(a[o].x - b[o].x, a[o].y - b[o].y);
}

// squared column norms, fixed-order per column
__global__ void column_norm2_kernel(const double2 *r, int64_t ld, int n, double *out) {
  __shared__ double red[kThreads];
  const double2 *c = r + blockIdx.x * ld;
  double acc = 0;
  for (int g = threadIdx.x; g < n; g += blockDim.x) acc += c[g].x * c[g].x + c[g].y * c[g].y;
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// P = -R/(|T - eps|+1), or a unit seed vector for converged bands
// (reference: src/half_cuda_iterative_solver.cuf:42-64)
// Column c of P is the correction of band cols[c] (c < p): a unit vector at
// plane wave seeds[c] where that is set, else the preconditioned residual.
__global__ void precondition_kernel(const double2 *r, const double *kin, const double *w, const int *cols,
                                    const int *seeds, double2 *p, int n, int np, int64_t ld) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(n) * np) return;
  const int c = static_cast<int>(t / n);
  const int g = static_cast<int>(t - int64_t(c) * n);
  const int b = cols[c];
  const int64_t o = g + c * ld, ro = g + b * ld;
  if (seeds[c] >= 0) {
    p[o] = make_double2(g == seeds[c] ? 1.0 : 0.0, 0.0);
  } else {
    const double d = fabs(kin[g] - w[b]) + 1.0;
    p[o] = make_double2(-r[ro].x / d, -r[ro].y / d);
  }
}

// Kinetic energy expectation <x|T|x> / <x|x> of band cols[c] of X (one
// workgroup per correction column, fixed-order tree sums)
__global__ void band_kinetic_kernel(const double2 *x, const double *kin, const int *cols, int n, int64_t ld,
                                    double *ekin) {
  __shared__ double red[2][kThreads];
  const double2 *c = x + cols[blockIdx.x] * ld;
  double a = 0, b = 0;
  for (int g = threadIdx.x; g < n; g += blockDim.x) {
    const double v = c[g].x * c[g].x + c[g].y * c[g].y;
    a += kin[g] * v;
    b += v;
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) ekin[blockIdx.x] = red[1][0] > 0 ? red[0][0] / red[1][0] : 1.0;
}

// Teter-Payne-Allan preconditioner: P = -K(x) R with x = T_G / T_b (T_b the
// band's kinetic energy expectation) and
// K(x) = (27 + 18x + 12x^2 + 8x^3) / (27 + 18x + 12x^2 + 8x^3 + 16x^4)
// (Teter, Payne and Allan, Phys. Rev. B 40, 12255 (1989)); seeds as in
// precondition_kernel
__global__ void precondition_tpa_kernel(const double2 *r, const double *kin, const double *ekin, const int *cols,
                                        const int *seeds, double2 *p, int n, int np, int64_t ld) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(n) * np) return;
  const int c = static_cast<int>(t / n);
  const int g = static_cast<int>(t - int64_t(c) * n);
  const int b = cols[c];
  const int64_t o = g + c * ld, ro = g + b * ld;
  if (seeds[c] >= 0) {
    p[o] = make_double2(g == seeds[c] ? 1.0 : 0.0, 0.0);
  } else {
    const double x = kin[g] / fmax(ekin[c], 1e-12);
    const double num = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x));
    const double k = num / (num + 16.0 * x * x * x * x);
    p[o] = make_double2(-k * r[ro].x, -k * r[ro].y);
  }
}

// w[j] = Re(x_j^H (H x)_j) / Re(x_j^H (S x)_j) for column j = blockIdx.x + c0
// (fixed-order tree sums)
__global__ void rayleigh_quotient_kernel(const double2 *x, const double2 *hx, const double2 *sx, int n, int64_t ld,
                                         int c0, double *w) {
  __shared__ double red[2][kThreads];
  const int j = c0 + blockIdx.x;
  const double2 *a = x + j * ld, *h = hx + j * ld, *q = sx + j * ld;
  double num = 0, den = 0;
  for (int g = threadIdx.x; g < n; g += blockDim.x) {
    num += a[g].x * h[g].x + a[g].y * h[g].y;
    den += a[g].x * q[g].x + a[g].y * q[g].y;
  }
  red[0][threadIdx.x] = num;
  red[1][threadIdx.x] = den;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[1][0] > 0) w[j] = red[0][0] / red[1][0];
}

// M <- D M D with D = diag(M)^-1/2 (zero where the diagonal is not positive);
// d keeps D's diagonal
__global__ void unit_diagonal_kernel(double2 *M, double *d, int m) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const double v = M[j + int64_t(j) * m].x;
  d[j] = v > 0 ? 1.0 / sqrt(v) : 0.0;
}
__global__ void scale_both_kernel(double2 *M, const double *d, int m) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(m) * m) return;
  const int j = static_cast<int>(t / m), i = static_cast<int>(t - int64_t(j) * m);
  const double f = d[i] * d[j];
  M[t] = make_double2(M[t].x * f, M[t].y * f);
}
// out[j] = squared pivot U_jj^2 of the factor, out[m] = factorisation info
__global__ void pivots_kernel(const double2 *U, const rocblas_int *info, double *out, int m) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) out[j] = U[j + int64_t(j) * m].x * U[j + int64_t(j) * m].x;
  if (j == 0) out[m] = static_cast<double>(info[0]);
}
// U <- U D^-1 (column j divided by d_j)
__global__ void unscale_columns_kernel(double2 *U, const double *d, int m) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(m) * m) return;
  const int j = static_cast<int>(t / m);
  const double f = d[j] > 0 ? 1.0 / d[j] : 0.0;
  U[t] = make_double2(U[t].x * f, U[t].y * f);
}

// U <- U diag(max(mu, floor))^-1/2 (reference: :66-77)
__global__ void scale_columns_kernel(double2 *u, const double *mu, int m, double floor_value) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(m) * m) return;
  const int col = static_cast<int>(t / m);
  const double f = 1.0 / sqrt(fmax(mu[col], floor_value));
  u[t] = make_double2(u[t].x * f, u[t].y * f);
}

void check_info(const DeviceArray<rocblas_int> &info, hipStream_t s, const char *what) {
  rocblas_int h = 0;
  HALF_HIP(hipMemcpyAsync(&h, info.get(), sizeof(h), hipMemcpyDeviceToHost, s));
  HALF_HIP(hipStreamSynchronize(s));
  if (h != 0) throw Error(kInternal, std::string(what) + " failed (info " + std::to_string(h) + ")");
}

// Correction blocks of at least this many elements (n x np) are normalised
// by the explicit inverse of the Cholesky factor and one batched product
// (216-atom cell: n x np ~ 6e6); smaller ones by three rocBLAS triangular
// solves, which were faster there (HfO2-like cell, 3,407 x 72: 3.0 against
// 3.3 ms per solve with the product in place).
constexpr int64_t kNormaliseByInverse = int64_t(1) << 20;

bool g_hetrd = true;
int g_hetrd_mode = 1;  // solver.hetrd value: 0 rocSOLVER, 1 own reductions, 2 cooperative one with columns in A
// Locking in the all-band solve (key "solver.lock").  1 (soft): a band
// whose residual already meets the current target contributes no correction
// column to the next Rayleigh-Ritz subspace (it stays in X).  2 (soft plus
// freezing): in addition every requested band that meets the target is
// frozen -- moved to the front of X, its vector, H and S images and Ritz
// value no longer enter the Rayleigh-Ritz problem, which covers only the
// remaining bands and the correction columns (the corrections stay
// S-orthogonal to every band, frozen ones included); at the end the pairs
// are sorted by Ritz value.  0: every band keeps its column, the reference's
// 2m-wide subspace.
int g_lock = 2;
// Preconditioner of the correction columns (key "solver.precond"): 0 the
// reference's -R / (|T - eps| + 1 eV) (default; the drop-in half_solve_kpoint
// path always uses it), 1 Teter-Payne-Allan: 63 instead of 96 iterations on
// the 216-atom cell (512 bands), but slower where a requested band has no
// buffer above it (2-atom cell, 16 bands: 148 instead of 90 iterations to
// 1e-7 eV in the oracle's restatement).
int g_precond = 0;
// Mixed form: every `g_refresh` FP32 iterations (key "solver.refresh", 0 =
// never) H X and S X of the bands still in the Rayleigh-Ritz problem are
// recomputed with the FP32 operator and their Ritz values replaced by the
// Rayleigh quotients: the rotations otherwise accumulate one FP32 operator
// error per iteration into H X (about sqrt(iterations) times the error of
// one application), which sets the floor the FP32 phase stalls at.
int g_refresh = 0;
// Mixed form: the FP32 phase's dense algebra with FP32 products (1, default)
// or FP64 (0), key "solver.fp32algebra".
int g_fp32_algebra = 1;
// Subspace problems of order <= kSmallMaxOrder on the one-workgroup kernels
// of smalleig.hip (key "solver.smalleig"): 0 rocSOLVER everywhere; 1 (default)
// the k-point mesh solve (which then captures whole iterations); 2 also the
// single-k Rayleigh-Ritz steps and dense generalized solves; 3 as 1 with the
// mesh's correction-block metric always in the eigen-decomposition form.
int g_smalleig = 1;

// zheevd with vectors (the full Hermitian matrix in A; eigenvectors replace
// A, eigenvalues ascending in w; e is workspace of n values).  Orders up to
// kTridiagMaxOrder reduce in one workgroup (hetrd_upper), orders up to
// kTridiagCoopMaxOrder in one cooperative launch of several (hetrd_upper_coop),
// then rocSOLVER solves the tridiagonal problem by divide and conquer
// (zstedc) and applies the reflectors (zunmtr): the reduction is the part of
// rocSOLVER's zheevd that is a chain of launches per column at these orders.
void heevd_vectors(Blas &blas, int k, double2 *A, double *w, double *e, rocblas_int *info) {
  if (!g_hetrd || k > kTridiagCoopMaxOrder) {
    HALF_BLAS(rocsolver_zheevd(blas.h, rocblas_evect_original, rocblas_fill_upper, k, rz(A), k, w, e, info));
    return;
  }
  blas.eig_tau.ensure(size_t(k));
  blas.eig_z.ensure(size_t(k) * k);
  double2 *Z = blas.eig_z.get(), *tau = blas.eig_tau.get();
  if (k <= kTridiagMaxOrder) {
    hetrd_upper(k, 1, A, k, 0, w, 0, e, 0, tau, 0, blas.stream);
  } else {
    const size_t need = coop_work_elems(k, kTridiagCoopMaxGroups);
    blas.eig_coop.ensure(need);
    if (!hetrd_upper_coop(k, A, k, w, e, tau, blas.eig_coop.get(), need, blas.stream, g_hetrd_mode == 2)) {
      HALF_BLAS(rocsolver_zheevd(blas.h, rocblas_evect_original, rocblas_fill_upper, k, rz(A), k, w, e, info));
      return;
    }
  }
  HALF_BLAS(rocsolver_zstedc(blas.h, rocblas_evect_tridiagonal, k, w, e, rz(Z), k, info));
  HALF_BLAS(rocsolver_zunmtr(blas.h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none, k, k, rz(A), k,
                             rz(tau), rz(Z), k));
  HALF_HIP(hipMemcpyAsync(A, Z, sizeof(double2) * size_t(k) * k, hipMemcpyDeviceToDevice, blas.stream));
}

// The generalized problem below on the one-workgroup kernels (order <=
// kSmallMaxOrder): U^-1 from potrf_inverse_batched, C = U^-H (A U^-1) by two
// products, hetrd_upper, then the tridiagonal eigensolve with the
// back-transformation by the reflectors and U^-1 of the lowest nv vectors
// (written to A(:, :nv)); eigenvalues ascending in w.  One host round trip
// checks both statuses.
void generalized_eigh_small(Blas &blas, int k, int nv, double2 *A, double2 *B, double *w, double *e, const char *what) {
  hipStream_t s = blas.stream;
  blas.eig_tau.ensure(size_t(k));
  blas.eig_z.ensure(size_t(k) * k);
  blas.eig_info.ensure(2);
  double2 *T = blas.eig_z.get(), *tau = blas.eig_tau.get();
  int *info = blas.eig_info.get();
  potrf_inverse_batched(k, 1, B, k, 0, B, k, 0, info, false, nullptr, 0, s);
  const double2 one = make_double2(1, 0), zero = make_double2(0, 0);
  zgemm(blas, rocblas_operation_none, rocblas_operation_none, k, k, k, one, A, k, B, k, zero, T, k);  // A U^-1
  zgemm(blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, k, k, k, one, B, k, T, k, zero, A, k);
  hetrd_upper(k, 1, A, k, 0, w, 0, e, 0, tau, 0, s);
  tridiag_eig_batched(k, 1, w, 0, e, 0, A, k, 0, tau, 0, nv, B, k, 0, 0.0, T, k, 0, w, 0, info + 1, nullptr, s);
  HALF_HIP(hipMemcpyAsync(A, T, sizeof(double2) * size_t(k) * nv, hipMemcpyDeviceToDevice, s));
  int h[2] = {0, 0};
  HALF_HIP(hipMemcpyAsync(h, info, sizeof(h), hipMemcpyDeviceToHost, s));
  HALF_HIP(hipStreamSynchronize(s));
  if (h[0] != 0)
    throw Error(kInternal, std::string(what) + " failed: overlap not positive definite (info " +
                               std::to_string(h[0]) + ")");
  if (h[1] != 0) throw Error(kInternal, std::string(what) + " failed (QL iterations)");
}

// Type-1 generalized Hermitian eigenproblem A y = e B y of order k, upper
// triangles referenced (LAPACK zhegvd semantics): B = U^H U, C = U^-H A U^-1,
// C v = e v, y = U^-1 v.  Eigenvectors replace A, eigenvalues go to w
// (ascending), B holds U.  info[0] is the Cholesky status, info[1] the
// eigensolver status; both are checked with a single host round trip.
// Without vectors A is left holding the reduced matrix's workspace.
void generalized_eigh(Blas &blas, int k, double2 *A, double2 *B, double *w, double *e, rocblas_int *info,
                      const char *what, bool want_vectors = true) {
  if (g_smalleig == 2 && want_vectors && k <= kSmallMaxOrder) {
    generalized_eigh_small(blas, k, k, A, B, w, e, what);
    return;
  }
  const rocblas_double_complex one{1.0, 0.0};
  HALF_BLAS(rocsolver_zpotrf(blas.h, rocblas_fill_upper, k, rz(B), k, info));
  HALF_BLAS(rocblas_ztrsm(blas.h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_conjugate_transpose,
                          rocblas_diagonal_non_unit, k, k, &one, rz(B), k, rz(A), k));
  HALF_BLAS(rocblas_ztrsm(blas.h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                          rocblas_diagonal_non_unit, k, k, &one, rz(B), k, rz(A), k));
  if (want_vectors)
    heevd_vectors(blas, k, A, w, e, info + 1);
  else
    HALF_BLAS(rocsolver_zheevd(blas.h, rocblas_evect_none, rocblas_fill_upper, k, rz(A), k, w, e, info + 1));
  if (want_vectors)
    HALF_BLAS(rocblas_ztrsm(blas.h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                            rocblas_diagonal_non_unit, k, k, &one, rz(B), k, rz(A), k));
  rocblas_int h[2] = {0, 0};
  HALF_HIP(hipMemcpyAsync(h, info, sizeof(h), hipMemcpyDeviceToHost, blas.stream));
  HALF_HIP(hipStreamSynchronize(blas.stream));
  if (h[0] != 0)
    throw Error(kInternal, std::string(what) + " failed: overlap not positive definite (info " +
                               std::to_string(h[0]) + ")");
  if (h[1] != 0) throw Error(kInternal, std::string(what) + " failed (info " + std::to_string(h[1]) + ")");
}

// Rayleigh-Ritz steps whose metric is the identity to within 10^-g_identity
// entrywise (key "solver.identity", 0 = never) skip the Cholesky reduction.
// The subspace the all-band solve hands over is S-orthonormal by
// construction (Ritz vectors, then corrections S-orthogonalised against them
// and normalised), so in FP64 iterations B = I + E with E at rounding level.
// The standard problem A v = e' v is solved instead and each kept pair is
// corrected by its metric Rayleigh quotient, b = v^H B v: e = e' / b,
// y = v / sqrt(b).  The residual of (e, y) in A y = e B y is then
// e (b - 1) y - e E y / sqrt(b), orthogonal to y to first order in E, so an
// eigenvalue set apart from the others by more than |e| |E| is exact to
// second order (|e| |E|^2 / gap); within a (nearly) degenerate cluster the
// off-diagonal part of E still splits the values at first order, by at most
// |e| max|E| (1e-10 eV for |e| <= 10 eV at the default 1e-11).  The kept
// vectors are B-normal and B-orthogonal to within |E|.  Where the
// normalisation was ill conditioned (pivots near its 1e-10 floor), or in
// FP32-product iterations (E ~ 1e-7), the deviation exceeds the threshold
// and the generalized form runs.  216-atom cell, 512 + 128 bands, tol 1e-5
// eV: eigensolves 485 -> 415 ms per FP64 solve at 1e-11 (the same steps as at
// 1e-10; 1.18 -> 1.11 s; eigenvalues unchanged to 1e-13 eV); HfO2 0.050 ->
// 0.043 s.
int g_identity = 11;
// The same for the Rayleigh-Ritz steps of the mixed form's FP32 phase, whose
// Grams are formed with FP32 products (key "solver.identity32", default 5,
// 0 = never): their metrics deviate by ~1e-7, the pairs they give only steer
// the FP32 iterations (the switch to FP64 redoes the Rayleigh-Ritz step with
// FP64 products and the FP64 threshold), and a cluster split of |e| 1e-5 is
// far below the FP32 operator's own error there (~2e-4 eV in the residual).
int g_identity32 = 5;
constexpr int kDeviationGroups = 256;

// max |B - I| over B's upper triangle (order k), read back to the host
double identity_deviation(Blas &blas, int k, const double2 *B) {
  blas.eig_dev.ensure(kDeviationGroups);
  hipLaunchKernelGGL(identity_deviation_kernel, dim3(kDeviationGroups), dim3(kThreads), 0, blas.stream, B, k,
                     blas.eig_dev.get());
  HALF_HIP(hipGetLastError());
  double h[kDeviationGroups];
  HALF_HIP(hipMemcpyAsync(h, blas.eig_dev.get(), sizeof(h), hipMemcpyDeviceToHost, blas.stream));
  HALF_HIP(hipStreamSynchronize(blas.stream));
  double mx = 0.0;
  for (double v : h)
    if (!(v <= mx)) mx = v;
  return mx;
}

// The same problem for the Rayleigh-Ritz steps of the all-band solve, where
// only the lowest `nv` eigenvectors are used (written to A(:, :nv)).  At
// orders up to kTridiagMaxOrder the triangular factor is inverted once
// (ztrtri) and applied by matrix products: three triangular solves cost
// about 50 us each at these orders, the inverse and three products less.
// Larger orders take generalized_eigh.
void generalized_eigh_ritz(Blas &blas, int k, int nv, double2 *A, double2 *B, double *w, double *e,
                           rocblas_int *info, const char *what, bool unit_metric = true) {
  if (g_smalleig == 2 && k <= kSmallMaxOrder) {
    generalized_eigh_small(blas, k, nv, A, B, w, e, what);
    return;
  }
  // (k max|E| bounds |E| in the 2-norm: at most 0.1 keeps every b = v^H B v
  // within 10% of 1 whatever the threshold key allows)
  const int level = unit_metric ? (contract_fp32_products() ? g_identity32 : g_identity) : 0;
  double dev = HUGE_VAL;
  if (level > 0) dev = identity_deviation(blas, k, B);
  if (dev <= std::pow(10.0, -level) && dev * k <= 0.1) {
    trace_mark("acc_subspace_eigensolve_unit_metric", blas.stream);
    heevd_vectors(blas, k, A, w, e, info + 1);
    // metric Rayleigh quotients of the kept pairs: T = B V(:, :nv)
    blas.eig_z.ensure(size_t(k) * k);
    double2 *T = blas.eig_z.get();
    const double2 one = make_double2(1, 0), zero = make_double2(0, 0);
    if (nv > 0) {
      zgemm(blas, rocblas_operation_none, rocblas_operation_none, k, nv, k, one, B, k, A, k, zero, T, k);
      hipLaunchKernelGGL(metric_rayleigh_kernel, dim3(nv), dim3(kThreads), 0, blas.stream, A, T, k, w);
      HALF_HIP(hipGetLastError());
    }
    rocblas_int h1 = 0;
    HALF_HIP(hipMemcpyAsync(&h1, info + 1, sizeof(h1), hipMemcpyDeviceToHost, blas.stream));
    HALF_HIP(hipStreamSynchronize(blas.stream));
    if (h1 != 0) throw Error(kInternal, std::string(what) + " failed (info " + std::to_string(h1) + ")");
    return;
  }
  if (!g_hetrd || k > kTridiagMaxOrder) {
    generalized_eigh(blas, k, A, B, w, e, info, what);
    return;
  }
  const double2 one = make_double2(1, 0), zero = make_double2(0, 0);
  blas.eig_z.ensure(size_t(k) * k);
  double2 *T = blas.eig_z.get();
  HALF_BLAS(rocsolver_zpotrf(blas.h, rocblas_fill_upper, k, rz(B), k, info));
  HALF_BLAS(rocsolver_ztrtri(blas.h, rocblas_fill_upper, rocblas_diagonal_non_unit, k, rz(B), k, info + 2));
  hipLaunchKernelGGL(upper_only_kernel, dim3(blocks_for(int64_t(k) * k)), dim3(kThreads), 0, blas.stream, B, k);
  HALF_HIP(hipGetLastError());
  // A <- U^-H (A U^-1); A holds the full Hermitian matrix
  zgemm(blas, rocblas_operation_none, rocblas_operation_none, k, k, k, one, A, k, B, k, zero, T, k);
  zgemm(blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, k, k, k, one, B, k, T, k, zero, A, k);
  heevd_vectors(blas, k, A, w, e, info + 1);
  // A(:, :nv) <- U^-1 V(:, :nv)
  zgemm(blas, rocblas_operation_none, rocblas_operation_none, k, nv, k, one, B, k, A, k, zero, T, k);
  HALF_HIP(hipMemcpyAsync(A, T, sizeof(double2) * size_t(k) * nv, hipMemcpyDeviceToDevice, blas.stream));
  rocblas_int h[3] = {0, 0, 0};
  HALF_HIP(hipMemcpyAsync(h, info, sizeof(h), hipMemcpyDeviceToHost, blas.stream));
  HALF_HIP(hipStreamSynchronize(blas.stream));
  if (h[0] != 0 || h[2] != 0)
    throw Error(kInternal, std::string(what) + " failed: overlap not positive definite (info " +
                               std::to_string(h[0] != 0 ? h[0] : h[2]) + ")");
  if (h[1] != 0) throw Error(kInternal, std::string(what) + " failed (info " + std::to_string(h[1]) + ")");
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

int g_stall = 4;
// Buffer bands given correction columns under soft locking, percent (lowest
// first; key "solver.bufcols").  Without them the buffer bands still take
// part in every Rayleigh-Ritz step through X and absorb what the requested
// bands' corrections bring in, and the problems are m + (active requested
// bands) wide instead of m + (active requested) + (buffer).  216-atom cell,
// 512 + 128 bands, tol 1e-5 eV: 100 percent 25 iterations at mean order 729,
// FP64 1.385 s, mixed 1.267 s; 50: 1.276 / 1.150 s; 25: 1.219 / 1.102 s;
// 0: 39 iterations at order 442, 1.188 / 1.089 s.  HfO2 64 + 16 bands: FP64
// 0.051 -> 0.046 s, mixed 0.044 -> 0.038 s.  (The k-point mesh solve, with
// 8 buffer bands of 40, keeps them all: dropping their columns took it from
// 26 to 44 iterations.)
int g_bufcols = 0;

int solver_stall_window() { return g_stall; }

double fp32_floor_estimate(const std::vector<double> &dh2, const std::vector<double> &ds2,
                           const std::vector<double> &eps) {
  double h = 0, sn = 0, e = 0;
  for (double v : dh2) h = std::max(h, std::sqrt(std::max(0.0, v)));
  for (double v : ds2) sn = std::max(sn, std::sqrt(std::max(0.0, v)));
  for (double v : eps) e = std::max(e, std::fabs(v));
  return h + e * sn;
}

bool set_solver_tuning(const std::string &key, int value) {
  if (key == "solver.stall") {
    if (value < 0 || value > 1000) return false;
    g_stall = value;
    return true;
  }
  if (key == "solver.lock") {
    if (value < 0 || value > 2) return false;
    g_lock = value;
    return true;
  }
  if (key == "solver.precond") {
    if (value < 0 || value > 1) return false;
    g_precond = value;
    return true;
  }
  if (key == "solver.refresh") {
    if (value < 0 || value > 1000) return false;
    g_refresh = value;
    return true;
  }
  if (key == "solver.fp32algebra") {
    if (value < 0 || value > 1) return false;
    g_fp32_algebra = value;
    return true;
  }
  if (key == "solver.smalleig") {
    if (value < 0 || value > 3) return false;
    g_smalleig = value;
    return true;
  }
  if (key == "solver.bufcols") {
    if (value < 0 || value > 100) return false;
    g_bufcols = value;
    return true;
  }
  if (key == "solver.identity") {
    if (value != 0 && (value < 6 || value > 15)) return false;
    g_identity = value;
    return true;
  }
  if (key == "solver.identity32") {
    if (value != 0 && (value < 3 || value > 15)) return false;
    g_identity32 = value;
    return true;
  }
  if (key != "solver.hetrd" || value < 0 || value > 2) return false;
  g_hetrd = value != 0;
  g_hetrd_mode = value;
  return true;
}

bool solver_locking() { return g_lock != 0; }
bool solver_fp32_algebra() { return g_fp32_algebra != 0; }
int solver_smalleig() { return g_smalleig; }

int AccControls::lock_mode() const { return lock < 0 ? g_lock : lock; }
int AccControls::precond_mode() const { return precond < 0 ? g_precond : precond; }
bool AccControls::locking() const { return lock_mode() != 0; }

bool get_solver_tuning(const std::string &key, int &value) {
  if (key == "solver.stall") {
    value = g_stall;
  } else if (key == "solver.lock") {
    value = g_lock;
  } else if (key == "solver.precond") {
    value = g_precond;
  } else if (key == "solver.refresh") {
    value = g_refresh;
  } else if (key == "solver.fp32algebra") {
    value = g_fp32_algebra;
  } else if (key == "solver.smalleig") {
    value = g_smalleig;
  } else if (key == "solver.bufcols") {
    value = g_bufcols;
  } else if (key == "solver.identity") {
    value = g_identity;
  } else if (key == "solver.identity32") {
    value = g_identity32;
  } else if (key == "solver.hetrd") {
    value = g_hetrd_mode;
  } else {
    return false;
  }
  return true;
}

void Blas::init(hipStream_t s) {
  stream = s;
  if (!h) HALF_BLAS(rocblas_create_handle(&h));
  HALF_BLAS(rocblas_set_stream(h, s));
}

void Blas::release() {
  if (h) rocblas_destroy_handle(h);
  h = nullptr;
}

Blas::~Blas() { release(); }

std::vector<int> kinetic_order(const std::vector<double> &kin) {
  const int n = static_cast<int>(kin.size());
  std::vector<int> idx(n + 1);
  for (int i = 1; i <= n; ++i) idx[i] = i;
  auto K = [&](int i) { return kin[idx[i] - 1]; };
  for (int i = 2; i <= n; ++i) {
    int child = i;
    while (child > 1) {
      const int parent = child / 2;
      if (K(parent) >= K(child)) break;
      std::swap(idx[parent], idx[child]);
      child = parent;
    }
  }
  for (int last = n; last >= 2; --last) {
    std::swap(idx[1], idx[last]);
    int parent = 1;
    while (true) {
      int child = 2 * parent;
      if (child > last - 1) break;
      if (child + 1 <= last - 1 && K(child + 1) > K(child)) child = child + 1;
      if (K(parent) >= K(child)) break;
      std::swap(idx[parent], idx[child]);
      parent = child;
    }
  }
  std::vector<int> out(n);
  for (int i = 0; i < n; ++i) out[i] = idx[i + 1] - 1;
  return out;
}

void solve_dense(Blas &blas, KOperator<double> &op, int solver, int nbands, bool want_vectors,
                 DeviceArray<double2> &vectors, SolveResult &res) {
  const int n = op.n;
  hipStream_t s = blas.stream;
  if (nbands < 1 || nbands > n) throw Error(kInvalidArgument, "invalid requested band count");
  const double t0 = now();
  DeviceArray<double2> H(size_t(n) * n), S(size_t(n) * n);
  assemble_dense(op, H.get(), S.get(), n, s);
  const double t1 = now();
  DeviceArray<double> W(n), E(n);
  DeviceArray<rocblas_int> info(1);
  res.overlap_min = res.overlap_max = -1.0;
  if (solver != kSolverEvx) {
    // overlap spectrum (reference: src/half_cuda_solver.cuf:145-150)
    DeviceArray<double2> S2(size_t(n) * n);
    HALF_HIP(hipMemcpyAsync(S2.get(), S.get(), sizeof(double2) * size_t(n) * n, hipMemcpyDeviceToDevice, s));
    HALF_BLAS(rocsolver_zheevd(blas.h, rocblas_evect_none, rocblas_fill_upper, n, rz(S2.get()), n, W.get(),
                               E.get(), info.get()));
    check_info(info, s, "overlap eigensolve");
    double lo = 0, hi = 0;
    HALF_HIP(hipMemcpyAsync(&lo, W.get(), sizeof(double), hipMemcpyDeviceToHost, s));
    HALF_HIP(hipMemcpyAsync(&hi, W.get() + (n - 1), sizeof(double), hipMemcpyDeviceToHost, s));
    HALF_HIP(hipStreamSynchronize(s));
    res.overlap_min = lo;
    res.overlap_max = hi;
  }
  const rocblas_evect ev = want_vectors ? rocblas_evect_original : rocblas_evect_none;
  if (solver == kSolverEvx) {
    // lowest nbands only (reference: src/half_cuda_solver.cuf:119-143)
    rocblas_int nev = 0;
    DeviceArray<rocblas_int> dnev(1);
    // Z must hold n columns: rocSOLVER may use all of them as workspace,
    // whatever the index range asks for
    DeviceArray<double2> Z(size_t(n) * n);
    HALF_BLAS(rocsolver_zhegvdx(blas.h, rocblas_eform_ax, ev, rocblas_erange_index, rocblas_fill_upper, n,
                                rz(H.get()), n, rz(S.get()), n, 0.0, 0.0, 1, nbands, dnev.get(), W.get(),
                                rz(Z.get()), n, info.get()));
    check_info(info, s, "index-range generalized eigensolve");
    HALF_HIP(hipMemcpyAsync(&nev, dnev.get(), sizeof(nev), hipMemcpyDeviceToHost, s));
    HALF_HIP(hipStreamSynchronize(s));
    if (nev != nbands) throw Error(kInternal, "index-range eigensolve returned a different band count");
    if (want_vectors) {
      vectors.resize(size_t(n) * nbands);
      HALF_HIP(hipMemcpyAsync(vectors.get(), Z.get(), sizeof(double2) * size_t(n) * nbands, hipMemcpyDeviceToDevice,
                              s));
      HALF_HIP(hipStreamSynchronize(s));
    }
  } else if (solver == kSolverEvj) {
    double resid = 0;
    rocblas_int sweeps = 0;
    DeviceArray<double> dres(1);
    DeviceArray<rocblas_int> dsw(1);
    HALF_BLAS(rocsolver_zhegvj(blas.h, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, n,
                               rz(H.get()), n, rz(S.get()), n, 1e-12, dres.get(), 100, dsw.get(), W.get(),
                               info.get()));
    check_info(info, s, "Jacobi generalized eigensolve");
    (void)resid;
    (void)sweeps;
  } else {
    // zhegvd's reduction (Cholesky, two triangular solves, zheevd) written out
    DeviceArray<rocblas_int> info2(2);
    generalized_eigh(blas, n, H.get(), S.get(), W.get(), E.get(), info2.get(), "generalized eigensolve",
                     want_vectors);
  }
  res.eigenvalues.resize(nbands);
  HALF_HIP(hipMemcpyAsync(res.eigenvalues.data(), W.get(), sizeof(double) * nbands, hipMemcpyDeviceToHost, s));
  if (want_vectors && solver != kSolverEvx) {
    vectors.resize(size_t(n) * nbands);
    HALF_HIP(hipMemcpyAsync(vectors.get(), H.get(), sizeof(double2) * size_t(n) * nbands,
                            hipMemcpyDeviceToDevice, s));
  }
  HALF_HIP(hipStreamSynchronize(s));
  res.seconds_operator = t1 - t0;
  res.seconds_solver = now() - t1;
}

namespace {

struct AccWork {
  int n, m, m2;
  // ZZ holds Z, HZ, SZ (n x 2m each) at a stride of n 2m, so the three
  // blocks' updates are one batched launch; tmp3 the three n x m results of
  // an out-of-place update at a stride of n m
  DeviceArray<double2> ZZ, tmp3, R, tmp, cross, metric, hsub, ssub, gwork;
  double2 *Z = nullptr, *HZ = nullptr, *SZ = nullptr;
  DeviceArray<float2> p32, hp32, sp32;
  DeviceArray<double> w, mu, E, res2, kin, dscale, pivots, ekin;
  DeviceArray<int> seeds, cols;
  DeviceArray<rocblas_int> info;
};

}  // namespace

void solve_acc(Blas &blas, KOperator<double> &op64, KOperator<float> *op32, int nbands, const AccControls &ctl,
               bool want_vectors, DeviceArray<double2> &vectors, SolveResult &res) {
  hipStream_t s = blas.stream;
  const int n = op64.n;
  const int mreq = nbands;
  const int m = nbands + std::max(0, ctl.extra);
  if (mreq < 1 || mreq > n) throw Error(kInvalidArgument, "invalid requested band count");
  if (2 * m > n) throw Error(kInvalidArgument, "ACC solver requires NBANDS <= NPL/2");
  if (!(ctl.tol > 0) || ctl.maxiter < 1 || ctl.block < 0) throw Error(kInvalidArgument, "invalid ACC solver controls");
  if (ctl.mixed && !op32) throw Error(kInternal, "mixed ACC needs an FP32 operator");
  op64.box_batch = ctl.block;
  if (op32) op32->box_batch = ctl.block;
  const double t0 = now();
  const int m2 = 2 * m;
  const int64_t ld = n;
  AccWork wk;
  const int64_t zs = int64_t(n) * m2;  // block stride in ZZ
  wk.ZZ.resize(3 * size_t(zs));
  wk.Z = wk.ZZ.get();
  wk.HZ = wk.Z + zs;
  wk.SZ = wk.HZ + zs;
  wk.tmp3.resize(3 * size_t(n) * m);
  wk.R.resize(size_t(n) * m);
  wk.tmp.resize(size_t(n) * m);
  wk.cross.resize(size_t(m) * m);
  wk.metric.resize(size_t(m) * m);
  wk.hsub.resize(size_t(m2) * m2);
  wk.ssub.resize(size_t(m2) * m2);
  wk.w.resize(m2);
  wk.mu.resize(m);
  wk.E.resize(m2);
  wk.res2.resize(m);
  wk.seeds.resize(m);
  wk.cols.resize(m);
  wk.info.resize(3);
  wk.dscale.resize(m);
  wk.pivots.resize(m + 1);
  wk.ekin.resize(m);
  wk.kin.upload(op64.basis.kinetic, s);
  if (ctl.mixed) {
    wk.p32.resize(size_t(n) * m);
    wk.hp32.resize(size_t(n) * m);
    wk.sp32.resize(size_t(n) * m);
  }
  double2 *X = wk.Z, *P = wk.Z + size_t(n) * m;
  double2 *HX = wk.HZ, *HP = wk.HZ + size_t(n) * m;
  double2 *SX = wk.SZ, *SP = wk.SZ + size_t(n) * m;
  const double2 one = make_double2(1, 0), zero = make_double2(0, 0);
  bool fp64_phase = !ctl.mixed;
  double fp32_floor = 0;
  const bool locking = ctl.locking();
  const bool hard = ctl.lock_mode() == 2;
  const bool tpa = ctl.precond_mode() == 1;
  int L = 0;              // frozen bands: columns 0..L-1 of X, HX, SX
  bool hard_used = false;
  res.applies = 0;

  // H and S applied to `cnt` columns
  auto apply = [&](int cnt, double2 *in, double2 *hout, double2 *sout) {
    if (fp64_phase) {
      op64.apply(cnt, in, ld, hout, sout, ld, s);
    } else {
      convert_states(in, ld, wk.p32.get(), ld, n, cnt, s);
      op32->apply(cnt, wk.p32.get(), ld, wk.hp32.get(), wk.sp32.get(), ld, s);
      convert_states(wk.hp32.get(), ld, hout, ld, n, cnt, s);
      convert_states(wk.sp32.get(), ld, sout, ld, n, cnt, s);
    }
    res.applies += cnt;
  };
  auto copy = [&](double2 *dst, const double2 *src, int cnt) {
    HALF_HIP(hipMemcpyAsync(dst, src, sizeof(double2) * size_t(n) * cnt, hipMemcpyDeviceToDevice, s));
  };
  // the three results in tmp3 (cnt columns each) back to column c0.. of
  // Z, HZ, SZ: one strided copy
  auto copy_back3 = [&](int c0, int cnt) {
    HALF_HIP(hipMemcpy2DAsync(wk.Z + size_t(n) * c0, sizeof(double2) * size_t(zs), wk.tmp3.get(),
                              sizeof(double2) * size_t(n) * m, sizeof(double2) * size_t(n) * cnt, 3,
                              hipMemcpyDeviceToDevice, s));
  };
  // generalized k x k problem in (hsub, ssub): the lowest m eigenvectors in
  // hsub(:, :m), values in w
  // the lowest nv (default m) eigenvectors in hsub(:, :nv), values from w + w0
  auto small_generalized = [&](int k, int nv = -1, int w0 = 0) {
    generalized_eigh_ritz(blas, k, nv < 0 ? m : nv, wk.hsub.get(), wk.ssub.get(), wk.w.get() + w0, wk.E.get(),
                          wk.info.get(), "ACC small generalized eigensolve", ctl.identity != 0);
  };
  // (hsub, ssub) = (A^H HA, A^H SA) for k columns of Z-shaped blocks
  auto gram_hs = [&](int k, const double2 *A, const double2 *HA, const double2 *SA) {
    gram<double>(n, k, k, A, ld, HA, ld, wk.hsub.get(), k, SA, ld, wk.ssub.get(), k, wk.gwork, s, true);
  };
  // rotate k columns of (Z, HZ, SZ) from column c0 by hsub(:, :m - c0)
  // into columns c0.. of X, HX, SX (one batched product for the three
  // blocks into tmp3, then one strided copy back)
  auto ritz_rotate = [&](int k, int c0 = 0) {
    const int nv = m - c0;
    const size_t off = size_t(n) * c0;
    gemm_tall_batched<double>(3, n, nv, k, 1.0, wk.Z + off, ld, zs, wk.hsub.get(), k, 0, 0.0, wk.tmp3.get(), ld,
                              int64_t(n) * m, nullptr, 0, 0, nullptr, 0, 0, s);
    copy_back3(c0, nv);
  };
  auto rayleigh_ritz_x = [&]() {
    gram_hs(m, X, HX, SX);
    small_generalized(m);
    ritz_rotate(m);
  };

  // initial subspace: unit vectors at the m lowest kinetic energies
  {
    auto order = kinetic_order(op64.basis.kinetic);
    std::vector<int> first(order.begin(), order.begin() + m);
    DeviceArray<int> d_idx;
    d_idx.upload(first, s);
    hipLaunchKernelGGL(init_states_kernel, dim3(blocks_for(int64_t(n) * m)), dim3(kThreads), 0, s, X, ld,
                       d_idx.get(), n, m);
    HALF_HIP(hipGetLastError());
    apply(m, X, HX, SX);
    if (!fp64_phase) {
      // FP32 residual floor: the FP32 operator's error on a dense probe
      // block, |dH p| + max|eps| |dS p| per unit column (P, HP, SP, R and
      // tmp are free until the loop).  Unit vectors would understate it:
      // their projector coefficients involve no accumulation.
      hipLaunchKernelGGL(phase_block_kernel, dim3(blocks_for(int64_t(n) * m)), dim3(kThreads), 0, s, P, n, m, ld);
      HALF_HIP(hipGetLastError());
      apply(m, P, HP, SP);
      op64.apply(m, P, ld, wk.R.get(), wk.tmp.get(), ld, s);
      res.applies += m;
      hipLaunchKernelGGL(difference_kernel, dim3(blocks_for(int64_t(n) * m)), dim3(kThreads), 0, s, HP, wk.R.get(),
                         n, m, ld);
      hipLaunchKernelGGL(difference_kernel, dim3(blocks_for(int64_t(n) * m)), dim3(kThreads), 0, s, SP,
                         wk.tmp.get(), n, m, ld);
      hipLaunchKernelGGL(column_norm2_kernel, dim3(m), dim3(kThreads), 0, s, wk.R.get(), ld, n, wk.res2.get());
      hipLaunchKernelGGL(column_norm2_kernel, dim3(m), dim3(kThreads), 0, s, wk.tmp.get(), ld, n, wk.mu.get());
      HALF_HIP(hipGetLastError());
    }
    rayleigh_ritz_x();
    if (!fp64_phase) {
      std::vector<double> eh(m), es(m), ev(m);
      HALF_HIP(hipMemcpyAsync(eh.data(), wk.res2.get(), sizeof(double) * m, hipMemcpyDeviceToHost, s));
      HALF_HIP(hipMemcpyAsync(es.data(), wk.mu.get(), sizeof(double) * m, hipMemcpyDeviceToHost, s));
      HALF_HIP(hipMemcpyAsync(ev.data(), wk.w.get(), sizeof(double) * m, hipMemcpyDeviceToHost, s));
      HALF_HIP(hipStreamSynchronize(s));
      for (int j = 0; j < m; ++j) {  // per unit column: |p_j|^2 = n
        eh[j] /= n;
        es[j] /= n;
      }
      fp32_floor = fp32_floor_estimate(eh, es, ev);
    }
    HALF_HIP(hipStreamSynchronize(s));
  }
  std::vector<double> r2(m), wh(m);
  std::vector<int> seeds(m), cols(m), perm(m);
  DeviceArray<int> d_perm(m);
  int64_t order_sum = 0;
  int rr_steps = 0;
  StallWatch fp32_stalled;
  std::vector<double> pivots(m + 1);
  res.metric_fallbacks = 0;
  res.history.clear();
  double max_res = HUGE_VAL;
  int iter = 0;
  bool converged = false;
  int fp64_iters = 0;
  int fp32_iters = 0;
  for (iter = 1; iter <= ctl.maxiter; ++iter) {
    if (!fp64_phase && g_refresh > 0 && fp32_iters > 0 && fp32_iters % g_refresh == 0 && L < m) {
      trace_mark("acc_fp32_refresh", s);
      const size_t off = size_t(n) * L;
      apply(m - L, X + off, HX + off, SX + off);
      hipLaunchKernelGGL(rayleigh_quotient_kernel, dim3(m - L), dim3(kThreads), 0, s, X, HX, SX, n, ld, L,
                         wk.w.get());
      HALF_HIP(hipGetLastError());
    }
    if (!fp64_phase) ++fp32_iters;
    trace_mark("acc_residual", s);
    hipLaunchKernelGGL(residual_kernel, dim3(blocks_for(int64_t(n) * m)), dim3(kThreads), 0, s, HX, SX,
                       wk.w.get(), wk.R.get(), n, m, ld);
    hipLaunchKernelGGL(column_norm2_kernel, dim3(m), dim3(kThreads), 0, s, wk.R.get(), ld, n, wk.res2.get());
    HALF_HIP(hipGetLastError());
    trace_mark("acc_norms_to_host", s);
    HALF_HIP(hipMemcpyAsync(r2.data(), wk.res2.get(), sizeof(double) * m, hipMemcpyDeviceToHost, s));
    if (hard) HALF_HIP(hipMemcpyAsync(wh.data(), wk.w.get(), sizeof(double) * m, hipMemcpyDeviceToHost, s));
    HALF_HIP(hipStreamSynchronize(s));
    max_res = 0;
    for (int j = 0; j < m; ++j) {
      const double r = std::sqrt(std::max(0.0, r2[j]));
      if (j < mreq) max_res = std::max(max_res, r);
    }
    res.history.push_back(max_res);
    if (fp64_phase) ++fp64_iters;
    const double fp32_target = std::max({ctl.switch_tol, ctl.tol, fp32_floor});
    const bool fp32_stall = !fp64_phase && ((max_res < kStallArm * fp32_target && fp32_stalled.update(max_res)) ||
                                            2 * iter > ctl.maxiter);
    if (!fp64_phase && (fp32_stall || max_res <= fp32_target)) {
      // FP64 correction sweeps: exact HX, SX for the current subspace; the
      // bands frozen at FP32 accuracy take part again
      fp64_phase = true;
      L = 0;
      apply(m, X, HX, SX);
      rayleigh_ritz_x();
      --iter;
      continue;
    }
    if (fp64_phase && max_res <= ctl.tol) {
      converged = true;
      break;
    }
    // correction columns: every band (the reference), or with soft locking
    // only the buffer bands and the requested bands above the current
    // phase's target (tol; switch_tol in the FP32 phase of the mixed form).
    // The tests above still read every requested band, so a locked band
    // whose residual grows past the target gets its column back.
    const double lock_at = fp64_phase ? ctl.tol : fp32_target;
    if (hard) {
      // freeze the requested bands that meet the target: they move, in
      // order, to just after the frozen ones (columns 0..L-1); the other
      // bands keep their order behind them, so columns below mreq still
      // hold requested bands and the buffer bands stay last
      int nconv = 0, moved = 0;
      for (int j = L; j < mreq; ++j)
        if (std::sqrt(std::max(0.0, r2[j])) <= lock_at) perm[L + nconv++] = j;
      for (int j = L, c = L + nconv; j < m; ++j)
        if (!(j < mreq && std::sqrt(std::max(0.0, r2[j])) <= lock_at)) perm[c++] = j;
      for (int i = L; i < m; ++i) moved += perm[i] != i;
      if (moved) {
        trace_mark("acc_freeze_reorder", s);
        const int cnt = m - L;
        d_perm.upload(perm.data() + L, size_t(cnt), s);
        for (double2 *blk : {X, HX, SX, wk.R.get()}) {
          hipLaunchKernelGGL(gather_columns_kernel, dim3(blocks_for(int64_t(n) * cnt)), dim3(kThreads), 0, s,
                             blk, wk.tmp.get(), d_perm.get(), n, cnt, ld);
          HALF_HIP(hipGetLastError());
          copy(blk + size_t(n) * L, wk.tmp.get(), cnt);
        }
        std::vector<double> w2(wh), r22(r2);
        for (int i = L; i < m; ++i) {
          wh[i] = w2[perm[i]];
          r2[i] = r22[perm[i]];
        }
        HALF_HIP(hipMemcpyAsync(wk.w.get() + L, wh.data() + L, sizeof(double) * cnt, hipMemcpyHostToDevice, s));
        HALF_HIP(hipStreamSynchronize(s));  // wh is reused by the next iteration's download
      }
      L += nconv;
      if (L > 0) hard_used = true;
    }
    for (int j = 0; j < m; ++j)
      seeds[j] = std::sqrt(std::max(0.0, r2[j])) < std::max(ctl.tol * 1e-2, 1e-12) ? (m + j) % n : -1;
    // With soft locking only the lowest solver.bufcols percent of the buffer
    // bands get columns (all of them where no other column is left)
    int np = 0;
    const int buf_end = mreq + (m - mreq) * g_bufcols / 100;
    for (int pass = 0; pass < 2 && np == 0; ++pass)
      for (int j = L; j < m; ++j) {
        if (locking && j < mreq && std::sqrt(std::max(0.0, r2[j])) <= lock_at) continue;
        if (locking && pass == 0 && j >= buf_end) continue;
        cols[np] = j;
        seeds[np] = seeds[j];
        ++np;
      }
    if (np == 0) break;  // unreachable: the test above stops first
    order_sum += m - L + np;
    ++rr_steps;
    trace_mark("acc_precondition", s);
    wk.seeds.upload(seeds, s);
    wk.cols.upload(cols, s);
    if (tpa) {
      hipLaunchKernelGGL(band_kinetic_kernel, dim3(np), dim3(kThreads), 0, s, X, wk.kin.get(), wk.cols.get(), n, ld,
                         wk.ekin.get());
      hipLaunchKernelGGL(precondition_tpa_kernel, dim3(blocks_for(int64_t(n) * np)), dim3(kThreads), 0, s,
                         wk.R.get(), wk.kin.get(), wk.ekin.get(), wk.cols.get(), wk.seeds.get(), P, n, np, ld);
    } else {
      hipLaunchKernelGGL(precondition_kernel, dim3(blocks_for(int64_t(n) * np)), dim3(kThreads), 0, s, wk.R.get(),
                         wk.kin.get(), wk.w.get(), wk.cols.get(), wk.seeds.get(), P, n, np, ld);
    }
    HALF_HIP(hipGetLastError());
    apply(np, P, HP, SP);
    // the rest of an FP32-phase iteration forms its Grams, updates and
    // rotations with FP32 products on the matrix cores (FP64 blocks; the
    // subspace eigensolves stay FP64); the switch to FP64 recomputes H X,
    // S X and the Rayleigh-Ritz step in FP64 (key solver.fp32algebra)
    ContractFp32Scope fp32_products(!fp64_phase && g_fp32_algebra != 0);
    // S-orthogonalise the correction block against X (cross: m x np)
    trace_mark("acc_orthogonalise_gemm", s);
    gram<double>(n, m, np, X, ld, SP, ld, wk.cross.get(), m, nullptr, 0, nullptr, 0, wk.gwork, s);
    // (P, HP, SP) -= (X, HX, SX) cross: one batched launch over the three blocks
    gemm_tall_batched<double>(3, n, np, m, -1.0, X, ld, zs, wk.cross.get(), m, 0, 1.0, P, ld, zs, nullptr, 0, 0,
                              nullptr, 0, 0, s);
    // normalise it in the S metric.  With D = diag(P^H S P)^-1/2 and
    // D P^H S P D = U^H U, P <- P D U^-1 spans the same space as the
    // reference's P V mu^-1/2 (eigenpairs (mu, V) of P^H S P, mu floored at
    // 1e-14), so the Rayleigh-Ritz step that follows yields the same Ritz
    // pairs.  It is used when the unit-diagonal factor's squared pivots all
    // exceed 1e-10 (no column within 1e-5 of the span of the others);
    // otherwise the reference's eigen-decomposition form runs.
    gram<double>(n, np, np, P, ld, SP, ld, wk.metric.get(), np, nullptr, 0, nullptr, 0, wk.gwork, s, true);
    trace_mark("acc_metric_cholesky", s);
    hipLaunchKernelGGL(unit_diagonal_kernel, dim3(blocks_for(np)), dim3(kThreads), 0, s, wk.metric.get(),
                       wk.dscale.get(), np);
    hipLaunchKernelGGL(scale_both_kernel, dim3(blocks_for(int64_t(np) * np)), dim3(kThreads), 0, s,
                       wk.metric.get(), wk.dscale.get(), np);
    HALF_HIP(hipGetLastError());
    HALF_BLAS(rocsolver_zpotrf(blas.h, rocblas_fill_upper, np, rz(wk.metric.get()), np, wk.info.get()));
    hipLaunchKernelGGL(pivots_kernel, dim3(blocks_for(np)), dim3(kThreads), 0, s, wk.metric.get(), wk.info.get(),
                       wk.pivots.get(), np);
    HALF_HIP(hipGetLastError());
    HALF_HIP(hipMemcpyAsync(pivots.data(), wk.pivots.get(), sizeof(double) * (np + 1), hipMemcpyDeviceToHost, s));
    HALF_HIP(hipStreamSynchronize(s));
    bool chol_ok = pivots[np] == 0.0;
    for (int j = 0; j < np && chol_ok; ++j) chol_ok = std::isfinite(pivots[j]) && pivots[j] > 1e-10;
    if (chol_ok) {
      hipLaunchKernelGGL(unscale_columns_kernel, dim3(blocks_for(int64_t(np) * np)), dim3(kThreads), 0, s,
                         wk.metric.get(), wk.dscale.get(), np);
      HALF_HIP(hipGetLastError());
      trace_mark("acc_normalise_trsm", s);
      if (int64_t(n) * np >= kNormaliseByInverse) {
        // U^-1 formed once (ztrtri, its strictly lower part cleared) and
        // applied to the three blocks by one batched product
        HALF_BLAS(rocsolver_ztrtri(blas.h, rocblas_fill_upper, rocblas_diagonal_non_unit, np,
                                   rz(wk.metric.get()), np, wk.info.get() + 1));
        // (info not read: every pivot was checked above, so U is invertible)
        hipLaunchKernelGGL(upper_only_kernel, dim3(blocks_for(int64_t(np) * np)), dim3(kThreads), 0, s,
                           wk.metric.get(), np);
        HALF_HIP(hipGetLastError());
        gemm_tall_batched<double>(3, n, np, np, 1.0, P, ld, zs, wk.metric.get(), np, 0, 0.0, wk.tmp3.get(), ld,
                                  int64_t(n) * m, nullptr, 0, 0, nullptr, 0, 0, s);
        copy_back3(m, np);
      } else {
        const rocblas_double_complex z1{1.0, 0.0};
        for (double2 *blk : {P, HP, SP})
          HALF_BLAS(rocblas_ztrsm(blas.h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                                  rocblas_diagonal_non_unit, n, np, &z1, rz(wk.metric.get()), np, rz(blk),
                                  static_cast<rocblas_int>(ld)));
      }
    } else {
      ++res.metric_fallbacks;
      gram<double>(n, np, np, P, ld, SP, ld, wk.metric.get(), np, nullptr, 0, nullptr, 0, wk.gwork, s);
      trace_mark("acc_metric_zheevd", s);
      HALF_BLAS(rocsolver_zheevd(blas.h, rocblas_evect_original, rocblas_fill_upper, np, rz(wk.metric.get()), np,
                                 wk.mu.get(), wk.E.get(), wk.info.get()));
      check_info(wk.info, s, "ACC correction metric eigensolve");
      hipLaunchKernelGGL(scale_columns_kernel, dim3(blocks_for(int64_t(np) * np)), dim3(kThreads), 0, s,
                         wk.metric.get(), wk.mu.get(), np, 1e-14);
      HALF_HIP(hipGetLastError());
      trace_mark("acc_normalise_gemm", s);
      for (double2 *blk : {P, HP, SP}) {
        zgemm(blas, rocblas_operation_none, rocblas_operation_none, n, np, np, one, blk, ld, wk.metric.get(), np,
              zero, wk.tmp.get(), ld);
        copy(blk, wk.tmp.get(), np);
      }
    }
    // Rayleigh-Ritz in span[X P] without the frozen bands (columns L..)
    const int k = m - L + np;
    const size_t zoff = size_t(n) * L;
    trace_mark("acc_projection_gemm", s);
    gram_hs(k, wk.Z + zoff, wk.HZ + zoff, wk.SZ + zoff);
    trace_mark("acc_subspace_eigensolve", s);
    small_generalized(k, m - L, L);
    trace_mark("acc_rotate_gemm", s);
    ritz_rotate(k, L);
    trace_close(s);
  }
  if (hard_used && L > 0) {
    // frozen bands keep the Ritz values they had when frozen, so the bands
    // are out of order: sort the pairs by value (a permutation, so every
    // vector keeps its residual)
    trace_mark("acc_final_sort", s);
    HALF_HIP(hipMemcpyAsync(wh.data(), wk.w.get(), sizeof(double) * m, hipMemcpyDeviceToHost, s));
    HALF_HIP(hipStreamSynchronize(s));
    std::vector<int> order(m);
    for (int j = 0; j < m; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return wh[a] < wh[b]; });
    bool sorted = true;
    for (int j = 0; j < m; ++j) sorted = sorted && order[j] == j;
    if (!sorted) {
      d_perm.upload(order, s);
      for (double2 *blk : {X, HX, SX}) {
        hipLaunchKernelGGL(gather_columns_kernel, dim3(blocks_for(int64_t(n) * m)), dim3(kThreads), 0, s, blk,
                           wk.tmp.get(), d_perm.get(), n, m, ld);
        HALF_HIP(hipGetLastError());
        copy(blk, wk.tmp.get(), m);
      }
      std::vector<double> w2(m), r22(m);
      for (int j = 0; j < m; ++j) {
        w2[j] = wh[order[j]];
        r22[j] = r2[order[j]];
      }
      HALF_HIP(hipMemcpyAsync(wk.w.get(), w2.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
      HALF_HIP(hipStreamSynchronize(s));
      // the residual reported is that of the bands now first
      max_res = 0;
      for (int j = 0; j < mreq; ++j) max_res = std::max(max_res, std::sqrt(std::max(0.0, r22[j])));
    }
    trace_close(s);
  }
  res.iterations = std::min(iter, ctl.maxiter);
  res.fp64_iterations = fp64_iters;
  res.residual = max_res;
  res.rr_order_mean = rr_steps > 0 ? double(order_sum) / rr_steps : 0.0;
  res.eigenvalues.resize(mreq);
  HALF_HIP(hipMemcpyAsync(res.eigenvalues.data(), wk.w.get(), sizeof(double) * mreq, hipMemcpyDeviceToHost, s));
  if (want_vectors) {
    vectors.resize(size_t(n) * mreq);
    HALF_HIP(hipMemcpyAsync(vectors.get(), X, sizeof(double2) * size_t(n) * mreq, hipMemcpyDeviceToDevice, s));
  }
  HALF_HIP(hipStreamSynchronize(s));
  res.overlap_min = res.overlap_max = -1.0;
  res.seconds_solver = now() - t0;
  (void)converged;
}

}  // namespace half
