// Projector stage of H.psi at tens to hundreds of bands (see projector.hip).
// All blocks are column-major.
#pragma once

#include "common.hpp"
#include "fft.hpp"

namespace half {

// Scratch of projector_coefficients: the workgroups' partial tiles (calls on
// one workspace must not overlap).
template <typename R>
struct CoeffWork {
  DeviceArray<cplx_t<R>> partial;
};

// c = B^H psi: np x nb (leading dimension ldc) from B (n x np) and psi
// (n x nb).  Each workgroup sums one slab of rows of a 16 x 16 tile, its
// waves' partial sums added in wave order; a second kernel adds the slabs in
// slab order.  The order of every sum is fixed, so the result does not
// depend on scheduling.
template <typename R>
void projector_coefficients(int n, int np, int nb, const cplx_t<R> *B, int64_t ldb, const cplx_t<R> *psi,
                            int64_t ldpsi, cplx_t<R> *c, int64_t ldc, CoeffWork<R> &work, hipStream_t s);

// The same for `batch` problems of one shape in one launch, entry b's
// operands at B + b sB, psi + b sPsi, c + b sC.
template <typename R>
void projector_coefficients_batched(int batch, int n, int np, int nb, const cplx_t<R> *B, int64_t ldb, int64_t sB,
                                    const cplx_t<R> *psi, int64_t ldpsi, int64_t sPsi, cplx_t<R> *c, int64_t ldc,
                                    int64_t sC, CoeffWork<R> &work, hipStream_t s);

// H = [H +] BD c and S = [S +] BQ c for BD, BQ (n x np, leading dimension
// ldb) and c (np x nb); add = false writes without reading H and S.
template <typename R>
void projector_expand(int n, int np, int nb, const cplx_t<R> *BD, const cplx_t<R> *BQ, int64_t ldb,
                      const cplx_t<R> *c, int64_t ldc, bool add, cplx_t<R> *H, cplx_t<R> *S, int64_t ldo,
                      hipStream_t s);

// The same for `batch` problems, entry b's operands at BD + b sB, BQ + b sB,
// c + b sC, H + b sO, S + b sO.
template <typename R>
void projector_expand_batched(int batch, int n, int np, int nb, const cplx_t<R> *BD, const cplx_t<R> *BQ,
                              int64_t ldb, int64_t sB, const cplx_t<R> *c, int64_t ldc, int64_t sC, bool add,
                              cplx_t<R> *H, cplx_t<R> *S, int64_t ldo, int64_t sO, hipStream_t s);

}  // namespace half
