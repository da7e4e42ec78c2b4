// Householder reduction of small Hermitian matrices to real symmetric
// tridiagonal form, one workgroup per matrix (see tridiag.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace half {

// Largest order the one-workgroup reduction takes.
constexpr int kTridiagMaxOrder = 256;

// LAPACK zhetrd with uplo = 'U' on `batch` matrices of order n <= kTridiagMaxOrder
// (column-major, leading dimension lda, matrix b at A + b * sA): on return
// d (n) and e (n - 1) hold the tridiagonal matrix T = Q^H A Q, tau (n - 1)
// and the strictly upper triangle of A the reflectors
// Q = H(n-2) ... H(0), H(i) = I - tau_i v v^H, v(i) = 1, v(0:i-1) stored in
// A(0:i-1, i+1), v(i+1:) = 0; the rest of A is overwritten.  The output
// feeds rocSOLVER's zstedc / dsterf (d, e) and zunmtr (A, tau) unchanged.
// With `only` given, matrix b is skipped (nothing read or written) where
// only[b] == 0.
void hetrd_upper(int n, int batch, double2 *A, int64_t lda, int64_t sA, double *d, int64_t sd, double *e,
                 int64_t se, double2 *tau, int64_t stau, hipStream_t s, const int *only = nullptr);

// The same reduction (uplo = 'U', one matrix, the full Hermitian matrix in
// A on entry) by a cooperative launch of several workgroups, for orders above
// kTridiagMaxOrder: the matrix's columns in their LDS where at most
// kCoopLdsMaxGroups workgroups hold it, else (or with force_global) in A,
// each workgroup updating its own columns in place.  Returns false, launching
// nothing, when the workgroups cannot all be resident; `work` needs
// coop_work_elems(n, kTridiagCoopMaxGroups) elements.
bool hetrd_upper_coop(int n, double2 *A, int64_t lda, double *d, double *e, double2 *tau, double2 *work,
                      size_t work_elems, hipStream_t s, bool force_global = false);
size_t coop_work_elems(int n, int groups);
constexpr int kCoopMinGroups = 8;
constexpr int kCoopLdsMaxGroups = 32;
// Largest order handed to hetrd_upper_coop: above it rocSOLVER's reduction
// is faster (dense solves of order 869 / 1067: 47.5 / 64.7 ms with the
// columns in A and 32 workgroups, 44.9 / 56.2 ms with rocSOLVER's; their
// columns would need more workgroups than a cheap grid barrier allows to
// stay in LDS).
constexpr int kTridiagCoopMaxOrder = 704;
constexpr int kTridiagCoopMaxGroups = 1024;

}  // namespace half
