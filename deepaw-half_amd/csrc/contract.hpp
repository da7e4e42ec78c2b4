// Dense complex contractions on the matrix cores with LDS-staged tiles
// (see contract.hip).  All blocks are column-major.
#pragma once

#include "common.hpp"
#include "fft.hpp"

namespace half {

// C = A^H B summed over the n rows of A (n x ka) and B (n x kb); C is ka x kb
// with leading dimension ldc.  When B2 is given, also C2 = A^H B2 (A is read
// once for both).  The row range is split into chunks whose partial tiles are
// added in a fixed order, so the result does not depend on scheduling.
// `work` grows to hold the partial tiles.
// hermitian (ka == kb): the results are known to be Hermitian (A^H H A,
// A^H S A); only the tiles on and above the diagonal are contracted and the
// strictly lower triangle is set to the conjugate of the upper one.
template <typename R>
void gram(int n, int ka, int kb, const cplx_t<R> *A, int64_t lda, const cplx_t<R> *B, int64_t ldb, cplx_t<R> *C,
          int64_t ldc, const cplx_t<R> *B2, int64_t ldb2, cplx_t<R> *C2, int64_t ldc2,
          DeviceArray<cplx_t<R>> &work, hipStream_t s, bool hermitian = false);

// The same for `batch` independent problems of equal shape, entry b's
// operands at A + b sA, B + b sB, C + b sC (one launch for all of them).
template <typename R>
void gram_batched(int batch, int n, int ka, int kb, const cplx_t<R> *A, int64_t lda, int64_t sA, const cplx_t<R> *B,
                  int64_t ldb, int64_t sB, cplx_t<R> *C, int64_t ldc, int64_t sC, const cplx_t<R> *B2, int64_t ldb2,
                  int64_t sB2, cplx_t<R> *C2, int64_t ldc2, int64_t sC2, DeviceArray<cplx_t<R>> &work, hipStream_t s,
                  bool hermitian = false);

// Elements of partial-tile workspace gram needs for C = A^H B (n rows, ka x
// kb result, one output, not Hermitian); 0 when the rows are not split.
template <typename R>
size_t gram_work_elems(int n, int ka, int kb);

// That gram with a caller-provided workspace of work_elems elements (no
// device allocation); kInternal if it is smaller than gram_work_elems.
template <typename R>
void gram_into(int n, int ka, int kb, const cplx_t<R> *A, int64_t lda, const cplx_t<R> *B, int64_t ldb, cplx_t<R> *C,
               int64_t ldc, cplx_t<R> *work, size_t work_elems, hipStream_t s);

// C = beta C + alpha A W for A (n x k) and W (k x m); C is n x m.  When W2 is
// given, also C2 = beta C2 + alpha A W2.  beta = 0 does not read C.  C must
// not overlap W, and either does not overlap A or is A itself (same pointer,
// leading dimension and batch stride, m <= gemm_tall_inplace_max): the
// in-place form gives each workgroup all output columns of its rows, which
// it writes only after reading every inner column of those rows.
template <typename R>
void gemm_tall(int n, int m, int k, R alpha, const cplx_t<R> *A, int64_t lda, const cplx_t<R> *W, int64_t ldw,
               R beta, cplx_t<R> *C, int64_t ldc, const cplx_t<R> *W2, int64_t ldw2, cplx_t<R> *C2, int64_t ldc2,
               hipStream_t s);

constexpr int gemm_tall_inplace_max = 64;

// While alive, gram / gemm_tall on FP64 blocks in the calling thread form
// their products in FP32 on the matrix cores (operands rounded to FP32 as
// they are staged, results and the beta C update in FP64): the FP32 phase
// of a mixed solve, whose iterates carry FP32 operator errors anyway.
// Scopes nest; the previous setting is restored on destruction.
class ContractFp32Scope {
 public:
  explicit ContractFp32Scope(bool on);
  ~ContractFp32Scope();
  ContractFp32Scope(const ContractFp32Scope &) = delete;
  ContractFp32Scope &operator=(const ContractFp32Scope &) = delete;

 private:
  bool prev_;
};

// Whether the calling thread's FP64 contractions currently form FP32
// products (inside a ContractFp32Scope(true)).
bool contract_fp32_products();

// Tuning keys contract.<fp64|fp32>.gram.wt / tall.wt (0 = shape rule, 1 =
// 32-wide tiles, 2 = 64-wide) and contract.<fp64|fp32>.gram.fill (10-100:
// percent of the device's workgroup slots every round of a row-chunked gram
// must fill).  False for an unknown key or a value out of range.
bool set_contract_tuning(const std::string &key, int value);
bool get_contract_tuning(const std::string &key, int &value);

// Batched form of gemm_tall (entry b's operands at base + b stride).
template <typename R>
void gemm_tall_batched(int batch, int n, int m, int k, R alpha, const cplx_t<R> *A, int64_t lda, int64_t sA,
                       const cplx_t<R> *W, int64_t ldw, int64_t sW, R beta, cplx_t<R> *C, int64_t ldc, int64_t sC,
                       const cplx_t<R> *W2, int64_t ldw2, int64_t sW2, cplx_t<R> *C2, int64_t ldc2, int64_t sC2,
                       hipStream_t s);

}  // namespace half
