// Batched small dense kernels for the subspace problems of the k-point mesh
// solve (orders up to kSmallMaxOrder: the Rayleigh-Ritz problems of meshes
// with up to 48 bands per k point and their correction-block metrics).  One
// workgroup per matrix and no host synchronisation, so the mesh solver's
// whole iteration can be captured in a hipGraph: rocSOLVER's batched zheevd
// fails under stream capture, and its factorisations report through info
// arrays the host has to read before the next step.
//
// potrf_inverse_kernel: the upper triangle, packed column by column, in LDS
// (96 x 97 / 2 complex = 74.5 KB); right-looking Cholesky (thread 0 takes the
// pivot, all threads scale row k, then update the trailing triangle); then
// the inverse in place, column by column:
//   U^-1(0:j, j) = -U^-1(0:j, 0:j) U(0:j, j) / U_jj,   U^-1(j, j) = 1 / U_jj.
//
// tridiag_eig_kernel: implicit QL iterations with Wilkinson's shift on the
// real tridiagonal matrix (the EISPACK tql2 iteration).  Wave 0 runs the
// scalar recurrence and writes each sweep's rotations (c, s) to LDS while the
// threads that own the rows of Z (real, n x n, in LDS) apply the previous
// sweep's rotations to their row: two sweep buffers, one barrier per sweep.
// The eigenvalues are then ranked, and the lowest mc columns of Z (complex
// from here on) are multiplied by the reflectors H(0), H(1), ..., H(n-2) of
// hetrd_upper in chunks of 32 columns (a wave owns whole columns, so the
// reflectors need no barrier; each is read where hetrd_upper left it), then
// by the optional triangular factor or the metric scaling.
#include "smalleig.hpp"

#include <cfloat>
#include <cmath>
#include <mutex>

#include "common.hpp"
#include "cplx.hpp"

namespace half {

namespace {

constexpr int kSeThreads = 256;
constexpr int kSeWaves = kSeThreads / 64;
constexpr int NM = kSmallMaxOrder;  // row stride of Z and V in LDS
constexpr int kChunk = 32;          // eigenvector columns per back-transform chunk
constexpr int kRowBase = 64;        // thread kRowBase + r owns row r of Z; wave 0 runs the QL recurrence
constexpr int kMaxQlIter = 30;      // QL iterations allowed per eigenvalue
static_assert(kRowBase + NM <= kSeThreads, "one thread per row of Z outside wave 0");
static_assert(kSeThreads == 256, "16 x 16 grid of the trailing update");
static_assert(NM % 2 == 0, "Z occupies a whole number of double2 slots");

constexpr size_t kPotrfLds = sizeof(double2) * size_t(NM) * (NM + 1) / 2;
constexpr size_t kEigLds = sizeof(double) * size_t(NM) * NM + sizeof(double2) * size_t(NM) * kChunk;

__device__ inline int packed(int i, int j) { return j * (j + 1) / 2 + i; }  // i <= j

__device__ inline double wave_sum_d(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

__global__ __launch_bounds__(kSeThreads) void potrf_inverse_kernel(int n, const double2 *B, int64_t ldb, int64_t sB,
                                                                    double2 *W, int64_t ldw, int64_t sW, int *info,
                                                                    int metric, double min_pivot, double2 *copy,
                                                                    int64_t sC) {
  extern __shared__ double2 se_lds[];
  double2 *L = se_lds;  // packed upper triangle, column j at j (j + 1) / 2
  __shared__ double dsc[NM];
  __shared__ double2 y[NM];
  __shared__ int s_fail;
  __shared__ double s_inv;
  const int b = blockIdx.x, t = threadIdx.x;
  B += b * sB;
  const int nn = n * n;
  if (t == 0) s_fail = 0;
  for (int j = t; j < n; j += kSeThreads) {
    const double v = B[j + int64_t(j) * ldb].x;
    dsc[j] = metric ? (v > 0 ? 1.0 / sqrt(v) : 0.0) : 1.0;
  }
  __syncthreads();
  for (int idx = t; idx < nn; idx += kSeThreads) {
    const int j = idx / n, i = idx - j * n;
    if (i <= j) L[packed(i, j)] = cscale(B[i + int64_t(j) * ldb], dsc[i] * dsc[j]);
  }
  __syncthreads();

  // B = U^H U: row k of U, then A(i, j) -= conj(U(k, i)) U(k, j) for k < i <= j
  for (int k = 0; k < n; ++k) {
    if (t == 0) {
      const double p = L[packed(k, k)].x;
      const bool ok = metric ? (isfinite(p) && p > min_pivot) : (p > 0.0);
      if (!ok) {
        s_fail = k + 1;
      } else {
        const double u = sqrt(p);
        L[packed(k, k)] = mk<double2>(u, 0.0);
        s_inv = 1.0 / u;
      }
    }
    __syncthreads();
    if (s_fail) break;
    const double inv = s_inv;
    for (int j = k + 1 + t; j < n; j += kSeThreads) L[packed(k, j)] = cscale(L[packed(k, j)], inv);
    __syncthreads();
    // trailing triangle on a 16 x 16 thread grid (rows fastest)
    const int r = n - 1 - k, ti = t & 15, tj = t >> 4;
    for (int jj = tj; jj < r; jj += 16) {
      const int j = k + 1 + jj, col = packed(0, j);
      const double2 ukj = L[packed(k, j)];
      for (int ii = ti; ii <= jj; ii += 16) {
        const int i = k + 1 + ii;
        L[col + i] = csub(L[col + i], cmulc(ukj, L[packed(k, i)]));
      }
    }
    __syncthreads();
  }

  if (s_fail) {
    if (t == 0) info[b] = metric ? 1 : s_fail;
    if (metric && copy) {  // the matrix as a full Hermitian block for the eigen-decomposition form
      copy += b * sC;
      for (int idx = t; idx < nn; idx += kSeThreads) {
        const int j = idx / n, i = idx - j * n;
        const double2 v = i <= j ? B[i + int64_t(j) * ldb] : B[j + int64_t(i) * ldb];
        copy[i + int64_t(j) * n] = i <= j ? v : mk<double2>(v.x, -v.y);
      }
    }
    return;
  }
  if (t == 0) info[b] = 0;

  // U^-1 in place: columns 0..j-1 already hold the inverse of the leading block
  for (int j = 0; j < n; ++j) {
    const double ujj_inv = 1.0 / L[packed(j, j)].x;
    if (t < j) {
      double2 acc = mk<double2>(0.0, 0.0);
      for (int l = t; l < j; ++l) acc = cadd(acc, cmul(L[packed(t, l)], L[packed(l, j)]));
      y[t] = acc;
    }
    __syncthreads();
    if (t < j) L[packed(t, j)] = cscale(y[t], -ujj_inv);
    if (t == j) L[packed(j, j)] = mk<double2>(ujj_inv, 0.0);
    __syncthreads();
  }

  // W = D U'^-1 (D = I without the metric scaling)
  W += b * sW;
  for (int idx = t; idx < nn; idx += kSeThreads) {
    const int j = idx / n, i = idx - j * n;
    W[i + int64_t(j) * ldw] = i <= j ? cscale(L[packed(i, j)], dsc[i]) : mk<double2>(0.0, 0.0);
  }
}

struct Sweep {
  double c[NM], s[NM];  // rotation i acts on columns i, i + 1 of Z
  int lo, hi, done;     // rotations hi, hi - 1, ..., lo in this order
};

__global__ __launch_bounds__(kSeThreads) void tridiag_eig_kernel(int n, const double *D, int64_t sd, const double *E,
                                                                  int64_t se, const double2 *A, int64_t lda,
                                                                  int64_t sA, const double2 *Tau, int64_t stau,
                                                                  int mc, const double2 *post, int64_t ldp,
                                                                  int64_t sP, double floor_value, double2 *out,
                                                                  int64_t ldo, int64_t sO, double *w, int64_t sw,
                                                                  int *info, const int *only) {
  const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wid = t >> 6;
  if (only && only[b] == 0) return;  // the whole workgroup, before any barrier
  extern __shared__ double2 se_lds[];
  double *Z = reinterpret_cast<double *>(se_lds);  // Z(r, c) at c NM + r
  double2 *V = se_lds + NM * NM / 2;                // V(r, jj) at jj NM + r
  __shared__ Sweep sweeps[2];
  __shared__ double dd[NM], ee[NM];
  __shared__ int perm[NM];
  __shared__ int s_fail;
  D += b * sd;
  E += b * se;
  A += b * sA;
  Tau += b * stau;
  for (int i = t; i < n; i += kSeThreads) {
    dd[i] = D[i];
    ee[i] = i < n - 1 ? E[i] : 0.0;
    perm[i] = 0;
  }
  for (int idx = t; idx < n * n; idx += kSeThreads) {
    const int c = idx / n, r = idx - c * n;
    Z[c * NM + r] = r == c ? 1.0 : 0.0;
  }
  __syncthreads();

  // wave 0: QL recurrence, one sweep of rotations per call (or done).  Its
  // lanes run the scalar recurrence in lockstep (uniform values; they store
  // the same values) and search for the next negligible off-diagonal
  // together, 64 entries per step.
  int l = 0, iter = 0;
  auto produce = [&](Sweep &sw) {
    sw.done = 0;
    while (l < n) {
      int mm = n - 1;
      for (int base = l; base < n - 1; base += 64) {
        const int idx = base + lane;
        bool split = true;  // entries past n - 2 end the search at n - 1
        if (idx < n - 1) split = fabs(ee[idx]) <= DBL_EPSILON * (fabs(dd[idx]) + fabs(dd[idx + 1]));
        const unsigned long long mask = __ballot(split);
        if (mask) {  // uniform over the wave
          mm = min(n - 1, base + __ffsll(mask) - 1);
          break;
        }
      }
      if (mm == l) {  // dd[l] has converged
        ++l;
        iter = 0;
        continue;
      }
      if (iter++ == kMaxQlIter) {
        s_fail = 1;
        l = n;
        break;
      }
      double g = (dd[l + 1] - dd[l]) / (2.0 * ee[l]);
      double r = sqrt(g * g + 1.0);
      g = dd[mm] - dd[l] + ee[l] / (g + copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      int i = mm - 1;
      // step i reads e(i), d(i) and d(i + 1) and writes e(i + 1), d(i + 1):
      // d(i + 1) is still the value step i + 1 read as its d(i), and step
      // i - 1's e, d are loaded ahead, so no LDS latency is on the chain
      double e_i = ee[i], d_i = dd[i], d_i1 = dd[i + 1];
      for (; i >= l; --i) {
        const int ip = i > l ? i - 1 : i;  // (the last step's load is unused)
        const double e_n = ee[ip], d_n = dd[ip];
        // entries are of the matrix's scale (|g| < 1 / eps above): f^2 + g^2
        // cannot overflow, and an underflow to zero is the split below.
        // r = sqrt(x) and 1 / r from the hardware's reciprocal square root
        // refined by two Newton steps (y <- y + y (1 - x y^2) / 2): the
        // chain's latency is in these two operations, and IEEE sqrt and
        // division take about twice the dependent instructions
        const double f = s * e_i, bb = c * e_i;
        const double x = fma(f, f, g * g);
        double rinv = 0.0;
        r = 0.0;
        if (x > 1e-280) {
          double y = __builtin_amdgcn_rsq(x);
          double h = fma(-x * y, y, 1.0);
          y = fma(0.5 * y, h, y);
          h = fma(-x * y, y, 1.0);
          y = fma(0.5 * y, h, y);
          rinv = y;
          r = x * y;
        } else if (x > 0.0) {
          r = sqrt(x);
          rinv = 1.0 / r;
        }
        ee[i + 1] = r;
        if (r == 0.0) {  // split here and search again
          dd[i + 1] = d_i1 - p;
          ee[mm] = 0.0;
          break;
        }
        s = f * rinv;
        c = g * rinv;
        g = d_i1 - p;
        r = (d_i - g) * s + 2.0 * c * bb;
        p = s * r;
        dd[i + 1] = g + p;
        g = c * r - bb;
        sw.c[i] = c;
        sw.s[i] = s;
        e_i = e_n;
        d_i1 = d_i;
        d_i = d_n;
      }
      sw.hi = mm - 1;
      sw.lo = i + 1;
      if (r == 0.0 && i >= l) {
        if (sw.lo <= sw.hi) return;
        continue;
      }
      dd[l] -= p;
      ee[l] = g;
      ee[mm] = 0.0;
      return;
    }
    sw.done = 1;
  };
  if (wid == 0) {
    s_fail = 0;
    produce(sweeps[0]);
  }
  __syncthreads();
  for (int k = 0;; ++k) {
    const Sweep &cur = sweeps[k & 1];
    if (cur.done) break;  // uniform: read after the barrier
    if (wid == 0) {
      produce(sweeps[(k + 1) & 1]);
    } else if (t >= kRowBase && t < kRowBase + n) {
      const int r = t - kRowBase;
      double carry = Z[(cur.hi + 1) * NM + r];
      for (int i = cur.hi; i >= cur.lo; --i) {
        const double c = cur.c[i], s = cur.s[i], zi = Z[i * NM + r];
        Z[(i + 1) * NM + r] = s * zi + c * carry;
        carry = c * zi - s * carry;
      }
      Z[cur.lo * NM + r] = carry;
    }
    __syncthreads();
  }
  if (t == 0 && info) info[b] = s_fail;

  // ascending order (ties by index)
  if (t < n) {
    const double v = dd[t];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double u = dd[j];
      rank += (u < v || (u == v && j < t)) ? 1 : 0;
    }
    perm[rank] = t;
  }
  __syncthreads();
  if (w)
    for (int j = t; j < n; j += kSeThreads) w[b * sw + j] = dd[perm[j]];

  constexpr int RQ = (NM + 63) / 64;  // rows per lane
  for (int c0 = 0; c0 < mc; c0 += kChunk) {
    const int nc = min(kChunk, mc - c0);
    for (int idx = t; idx < nc * NM; idx += kSeThreads) {
      const int jj = idx / NM, r = idx - jj * NM;
      V[jj * NM + r] = mk<double2>(r < n ? Z[perm[c0 + jj] * NM + r] : 0.0, 0.0);
    }
    __syncthreads();
    // V <- H(n-2) ... H(0) V, H(i) = I - tau_i v v^H, v(0:i-1) = A(0:i-1, i+1), v(i) = 1
    for (int i = 0; i < n - 1; ++i) {
      const double2 tau = Tau[i];
      if (tau.x == 0.0 && tau.y == 0.0) continue;
      double2 v[RQ];
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        const int r = lane + 64 * q;
        v[q] = r < i ? A[r + int64_t(i + 1) * lda] : mk<double2>(r == i ? 1.0 : 0.0, 0.0);
      }
      for (int jj = wid; jj < nc; jj += kSeWaves) {  // uniform over the wave
        double2 *col = V + jj * NM;
        double2 acc = mk<double2>(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
          const int r = lane + 64 * q;
          if (r <= i) acc = cadd(acc, cmulc(col[r], v[q]));  // conj(v_r) V(r, jj)
        }
        acc.x = wave_sum_d(acc.x);
        acc.y = wave_sum_d(acc.y);
        const double2 f = cmul(tau, acc);
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
          const int r = lane + 64 * q;
          if (r <= i) col[r] = csub(col[r], cmul(v[q], f));
        }
      }
    }
    __syncthreads();
    double2 *o = out + b * sO;
    if (post) {
      const double2 *P = post + b * sP;
      for (int idx = t; idx < nc * n; idx += kSeThreads) {
        const int jj = idx / n, r = idx - jj * n;
        double2 acc = mk<double2>(0.0, 0.0);
        for (int l2 = r; l2 < n; ++l2) acc = cadd(acc, cmul(P[r + int64_t(l2) * ldp], V[jj * NM + l2]));
        o[r + int64_t(c0 + jj) * ldo] = acc;
      }
    } else {
      for (int idx = t; idx < nc * n; idx += kSeThreads) {
        const int jj = idx / n, r = idx - jj * n;
        const double f = floor_value > 0 ? 1.0 / sqrt(fmax(dd[perm[c0 + jj]], floor_value)) : 1.0;
        o[r + int64_t(c0 + jj) * ldo] = cscale(V[jj * NM + r], f);
      }
    }
    __syncthreads();
  }
}

}  // namespace

void smalleig_init() {
  static std::mutex mu;
  static bool done_devices[64] = {};
  int dev = 0;
  HALF_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if (dev >= 0 && dev < 64 && done_devices[dev]) return;
  HALF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&potrf_inverse_kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kPotrfLds)));
  HALF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&tridiag_eig_kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kEigLds)));
  if (dev >= 0 && dev < 64) done_devices[dev] = true;
}

void potrf_inverse_batched(int n, int batch, const double2 *B, int64_t ldb, int64_t sB, double2 *W, int64_t ldw,
                           int64_t sW, int *info, bool metric, double2 *copy, int64_t sC, hipStream_t s,
                           bool force_eig) {
  if (n < 1 || n > kSmallMaxOrder || batch < 1 || ldb < n || ldw < n || !info || (metric && copy && sC < int64_t(n) * n))
    throw Error(kInvalidArgument, "potrf_inverse_batched: order " + std::to_string(n) + " outside 1.." +
                                      std::to_string(kSmallMaxOrder) + " or bad leading dimension / batch");
  smalleig_init();
  hipLaunchKernelGGL(potrf_inverse_kernel, dim3(batch), dim3(kSeThreads), kPotrfLds, s, n, B, ldb, sB, W, ldw, sW,
                     info, metric ? 1 : 0, force_eig ? HUGE_VAL : 1e-10, copy, sC);
  HALF_HIP(hipGetLastError());
}

void tridiag_eig_batched(int n, int batch, const double *d, int64_t sd, const double *e, int64_t se,
                         const double2 *A, int64_t lda, int64_t sA, const double2 *tau, int64_t stau, int mc,
                         const double2 *post, int64_t ldp, int64_t sP, double floor_value, double2 *out, int64_t ldo,
                         int64_t sO, double *w, int64_t sw, int *info, const int *only, hipStream_t s) {
  if (n < 1 || n > kSmallMaxOrder || batch < 1 || mc < 0 || mc > n || lda < n || ldo < n || (post && ldp < n))
    throw Error(kInvalidArgument, "tridiag_eig_batched: order " + std::to_string(n) + " outside 1.." +
                                      std::to_string(kSmallMaxOrder) + " or bad column count / leading dimension");
  smalleig_init();
  hipLaunchKernelGGL(tridiag_eig_kernel, dim3(batch), dim3(kSeThreads), kEigLds, s, n, d, sd, e, se, A, lda, sA, tau,
                     stau, mc, post, ldp, sP, floor_value, out, ldo, sO, w, sw, info, only);
  HALF_HIP(hipGetLastError());
}

}  // namespace half
