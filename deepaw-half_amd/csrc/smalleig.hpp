// Batched small dense kernels for the subspace problems of the k-point mesh
// solve: one workgroup per matrix, no library calls, no host synchronisation
// and no memset, so a whole solver iteration can be captured in a hipGraph
// (see smalleig.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace half {

// Largest order the kernels take (the matrix or its factor in one CU's LDS).
constexpr int kSmallMaxOrder = 96;

// Cholesky factor B = U^H U of `batch` Hermitian matrices of order n (upper
// triangle read; entry b at B + b sB, column-major, leading dimension ldb) and
// its explicit inverse W = U^-1 (upper triangular, zeros below the diagonal,
// at W + b sW, leading dimension ldw).
//   metric == false: info[b] = 0, or j + 1 at the first pivot that is not
//     positive (rocSOLVER zpotrf's convention; W is then not written).
//   metric == true: the factor of the matrix scaled to unit diagonal,
//     D B D = U'^H U' with D = diag(B)^-1/2 (zero where not positive); when
//     every pivot U'_jj^2 is finite and above 1e-10, W = (U' D^-1)^-1 = D U'^-1
//     and info[b] = 0; otherwise info[b] = 1, W is not written and the
//     Hermitian matrix B (its upper triangle, mirrored) is copied to
//     copy + b sC (leading dimension n) for the eigen-decomposition form
//     (hetrd_upper, then tridiag_eig_batched with a floor); force_eig
//     takes that form for every matrix.
// W may be B itself (same leading dimension and stride): B is read whole
// before W is written.
void potrf_inverse_batched(int n, int batch, const double2 *B, int64_t ldb, int64_t sB, double2 *W, int64_t ldw,
                           int64_t sW, int *info, bool metric, double2 *copy, int64_t sC, hipStream_t s,
                           bool force_eig = false);

// Eigenpairs of `batch` Hermitian matrices A = Q T Q^H already reduced by
// hetrd_upper (d, e: the real tridiagonal T; A's strictly upper triangle and
// tau: the reflectors of Q), by implicit QL iterations on T, in ascending
// order.  The lowest mc eigenvectors are back-transformed, V = Q Z(:, 0:mc),
// and written to out + b sO (n x mc, leading dimension ldo) as
//   post != nullptr: post V, post an upper triangular n x n matrix at
//     post + b sP (the generalized problem's U^-1: eigenvectors of
//     A x = lambda U^H U x);
//   post == nullptr and floor_value > 0: V diag(max(lambda, floor))^-1/2.
// Eigenvalues (all n, ascending) go to w + b sw when w is given.  info[b] = 0,
// or 1 when an eigenvalue took more than 30 QL iterations.  With `only`
// given, matrix b is skipped (nothing read or written) where only[b] == 0.
void tridiag_eig_batched(int n, int batch, const double *d, int64_t sd, const double *e, int64_t se,
                         const double2 *A, int64_t lda, int64_t sA, const double2 *tau, int64_t stau, int mc,
                         const double2 *post, int64_t ldp, int64_t sP, double floor_value, double2 *out, int64_t ldo,
                         int64_t sO, double *w, int64_t sw, int *info, const int *only, hipStream_t s);

// Sets the kernels' LDS limits; called once per process before any capture
// (potrf_inverse_batched and tridiag_eig_batched also call it themselves).
void smalleig_init();

}  // namespace half
