// Projector stage of H.psi on the matrix cores, shaped for its skinny
// operands: B is n x np with n in the thousands to tens of thousands and np,
// nb (bands) in the tens to hundreds, so the 64-wide LDS-tiled contractions
// of contract.hip give a few hundred workgroups of long serial slab loops.
// Here every wave owns one 16 x 16 complex output tile and streams its
// operands straight from global memory into the MFMA operand layout (lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16]: a lane group reads four
// consecutive rows of 16 columns), with no LDS slabs and no barriers in the
// inner loop; the inner index is split over the waves of a workgroup (and,
// for the coefficients, over workgroups) to give thousands of waves, and the
// partial tiles are added in a fixed order.
//
//   coefficients  c = B^H psi           (reference: the Z^H contraction of
//                                        src/half_cuda_assembly.cuf:500-505)
//   expansion     H += (B D) c, S += (B Q) c with B D, B Q formed once per
//                 operator, so the per-atom D / Q mix of c needs no launch of
//                 its own (reference: :506-517)
#include <algorithm>

#include "cplx.hpp"
#include "mfma.hpp"
#include "projector.hpp"

namespace half {

namespace {

constexpr int kCoeffWaves = 8;    // row slabs per coefficient workgroup
constexpr int kExpandWaves = 4;   // projector quarters per expansion workgroup


// Accumulator entry r of lane l of a 16 x 16 tile holds row mfma_row(l, r),
// column l % 16; the lane and entry of (row, col):
template <typename R>
__device__ inline void tile_slot(int row, int col, int &lane, int &r) {
  if (std::is_same<R, double>::value) {  // row = 4 r + l / 16
    lane = (row & 3) * 16 + col;
    r = row >> 2;
  } else {  // row = 4 (l / 16) + r
    lane = (row >> 2) * 16 + col;
    r = row & 3;
  }
}

// The chip's XCDs (each with its own L2) take consecutive workgroup ids in
// turn: id % kXcds is the XCD of workgroup id.
constexpr int kXcds = 8;

// batch strides (elements) of the operands of one launch
struct Strides3 {
  int64_t a, b, c;
};

// Workgroup (tile, slab): wave w sums rows [r0, r1) of chunk
// slab * kCoeffWaves + w.  One-dimensional grid of tiles * slabs workgroups,
// slabs a multiple of kXcds: slab = id % kXcds + kXcds ((id / kXcds) / tiles),
// so every tile of a slab runs on one XCD and its B and psi rows are read
// into that XCD's L2 once.
template <typename R>
__global__ __launch_bounds__(64 * kCoeffWaves) void coeff_kernel(const cplx_t<R> *B, int64_t ldb,
                                                                 const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *c,
                                                                 int64_t ldc, cplx_t<R> *partial, int n, int np,
                                                                 int nb, int chunk, int ptiles, int tiles,
                                                                 Strides3 bs) {
  using V = cplx_t<R>;
  using A4 = typename Acc<R>::type;
  __shared__ V red[kCoeffWaves][4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t bz = blockIdx.y;  // batch entry
  B += bz * bs.a;
  psi += bz * bs.b;
  partial += bz * int64_t(gridDim.x) * 256;
  const int slot = blockIdx.x / kXcds;
  const int tile = slot % tiles, slab = blockIdx.x % kXcds + kXcds * (slot / tiles);
  const int p0 = (tile % ptiles) * 16, b0 = (tile / ptiles) * 16;
  const int i = lane & 15, kq = lane >> 4;
  const bool pv = p0 + i < np, bv = b0 + i < nb;
  // out-of-range columns read a valid column and are zeroed
  const V *pa = B + int64_t(min(p0 + i, np - 1)) * ldb;
  const V *pb = psi + int64_t(min(b0 + i, nb - 1)) * ldpsi;
  const int ch = slab * kCoeffWaves + wave;
  const int r0 = ch * chunk, r1 = min(n, r0 + chunk);
  A4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
  // (MFMAs are convergent: the compiler keeps this runtime-bounded loop rolled)
  for (int k = r0; k < r1; k += 4) {
    const int kk = k + kq;
    const int kc = min(kk, n - 1);
    V a = pa[kc], b = pb[kc];
    if (!(kk < r1 && pv)) a = mk<V>(R(0), R(0));
    if (!(kk < r1 && bv)) b = mk<V>(R(0), R(0));
    cmfma<R>(a.x, -a.y, b.x, b.y, cr, ci);  // A = conj(B)
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) red[wave][r][lane] = mk<V>(cr[r], ci[r]);
  __syncthreads();
  // 256 entries per tile: thread e < 256 takes (row e % 16, column e / 16)
  const int e = threadIdx.x;
  int el = 0, er = 0;
  if (e < 256) tile_slot<R>(e & 15, e >> 4, el, er);
  V sum = mk<V>(R(0), R(0));
  if (e < 256)
    for (int w = 0; w < kCoeffWaves; ++w) sum = cadd(sum, red[w][er][el]);
  if (e < 256) partial[(int64_t(slab) * tiles + tile) * 256 + e] = sum;
}

// c[p + b ldc] = sum over slabs g (in order) of the partial tiles; thread per
// entry, consecutive threads down a column of c
template <typename V>
__global__ __launch_bounds__(256) void coeff_reduce_kernel(const V *partial, int slabs, int tiles, int ptiles,
                                                           V *c, int64_t ldc, int np, int nb, int64_t sc) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(np) * nb) return;
  partial += blockIdx.y * int64_t(slabs) * tiles * 256;
  c += blockIdx.y * sc;
  const int b = static_cast<int>(t / np), p = static_cast<int>(t - int64_t(b) * np);
  const int tile = (b >> 4) * ptiles + (p >> 4);
  const V *src = partial + int64_t(tile) * 256 + (p & 15) + 16 * (b & 15);
  const int64_t step = int64_t(tiles) * 256;
  V s = src[0];
  for (int g = 1; g < slabs; ++g) s = cadd(s, src[g * step]);
  c[p + int64_t(b) * ldc] = s;
}

// Workgroup per 16 x 16 output tile (16 plane waves x 16 bands) of both H
// and S; wave w sums projectors [w pchunk, (w + 1) pchunk).  One-dimensional
// grid: the band tiles of one row tile run on one XCD, one after another
// (row tile r on XCD r % kXcds), so the rows of B D and B Q are read into
// that XCD's L2 once.
template <typename R>
__global__ __launch_bounds__(64 * kExpandWaves) void expand_kernel(const cplx_t<R> *BD, const cplx_t<R> *BQ,
                                                                   int64_t ldb, const cplx_t<R> *c, int64_t ldc,
                                                                   cplx_t<R> *H, cplx_t<R> *S, int64_t ldo, int n,
                                                                   int np, int nb, int add, int pchunk, int gtiles,
                                                                   int btiles, Strides3 bs) {
  using V = cplx_t<R>;
  using A4 = typename Acc<R>::type;
  __shared__ V red[2][kExpandWaves][4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t bz = blockIdx.y;  // batch entry
  BD += bz * bs.a;
  BQ += bz * bs.a;
  c += bz * bs.b;
  H += bz * bs.c;
  S += bz * bs.c;
  const int slot = blockIdx.x / kXcds;
  const int gt = (slot / btiles) * kXcds + blockIdx.x % kXcds, bt = slot % btiles;
  if (gt >= gtiles) return;  // uniform per workgroup: no barrier reached
  const int g0 = gt * 16, b0 = bt * 16;
  const int i = lane & 15, kq = lane >> 4;
  const int64_t grow = min(g0 + i, n - 1);                 // A rows: out-of-range rows are never stored
  const V *pc = c + int64_t(min(b0 + i, nb - 1)) * ldc;   // B columns: likewise
  const int pbeg = wave * pchunk, pend = min(np, pbeg + pchunk);
  A4 hr = {0, 0, 0, 0}, hi = {0, 0, 0, 0}, sr = {0, 0, 0, 0}, si = {0, 0, 0, 0};
  for (int p = pbeg; p < pend; p += 4) {
    const int pp = p + kq;
    const int64_t pcl = min(pp, np - 1);
    const V ad = BD[grow + pcl * ldb], aq = BQ[grow + pcl * ldb];
    V b = pc[pcl];
    if (pp >= pend) b = mk<V>(R(0), R(0));
    cmfma<R>(ad.x, ad.y, b.x, b.y, hr, hi);
    cmfma<R>(aq.x, aq.y, b.x, b.y, sr, si);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    red[0][wave][r][lane] = mk<V>(hr[r], hi[r]);
    red[1][wave][r][lane] = mk<V>(sr[r], si[r]);
  }
  __syncthreads();
  // thread t: output t / 256 (H, S), row t % 16, column (t / 16) % 16
  const int t = threadIdx.x;
  for (int u = t; u < 512; u += 64 * kExpandWaves) {
    const int j = u >> 8, row = u & 15, col = (u >> 4) & 15;
    const int g = g0 + row, b = b0 + col;
    if (g >= n || b >= nb) continue;
    int el, er;
    tile_slot<R>(row, col, el, er);
    V sum = red[j][0][er][el];
    for (int w = 1; w < kExpandWaves; ++w) sum = cadd(sum, red[j][w][er][el]);
    V *dst = (j ? S : H) + g + int64_t(b) * ldo;
    *dst = add ? cadd(*dst, sum) : sum;
  }
}

}  // namespace

template <typename R>
void projector_coefficients(int n, int np, int nb, const cplx_t<R> *B, int64_t ldb, const cplx_t<R> *psi,
                            int64_t ldpsi, cplx_t<R> *c, int64_t ldc, CoeffWork<R> &work, hipStream_t s) {
  projector_coefficients_batched<R>(1, n, np, nb, B, ldb, 0, psi, ldpsi, 0, c, ldc, 0, work, s);
}

template <typename R>
void projector_coefficients_batched(int batch, int n, int np, int nb, const cplx_t<R> *B, int64_t ldb, int64_t sB,
                                    const cplx_t<R> *psi, int64_t ldpsi, int64_t sPsi, cplx_t<R> *c, int64_t ldc,
                                    int64_t sC, CoeffWork<R> &work, hipStream_t s) {
  if (np <= 0 || nb <= 0 || batch <= 0) return;
  if (batch > 65535) throw Error(kInvalidArgument, "projector coefficients: batch too large");
  if (n <= 0) throw Error(kInvalidArgument, "projector coefficients: empty basis");
  if (ldb < n || ldpsi < n || ldc < np) throw Error(kInvalidArgument, "projector coefficients: leading dimension");
  const int ptiles = (np + 15) / 16, btiles = (nb + 15) / 16;
  const int64_t tiles = int64_t(ptiles) * btiles;
  if (tiles > (1ll << 31) - 1) throw Error(kInvalidArgument, "projector coefficients: too many tiles");
  // about two thousand waves (over the batch), each at least 32 rows (8 MFMA
  // steps), in slabs of kCoeffWaves chunks, a multiple of kXcds slabs
  const int64_t target = std::max<int64_t>(1, 2048 / (tiles * batch));
  int chunk = static_cast<int>((n + target - 1) / target);
  chunk = std::max(32, (chunk + 3) / 4 * 4);
  int slabs = (n + chunk * kCoeffWaves - 1) / (chunk * kCoeffWaves);
  slabs = (slabs + kXcds - 1) / kXcds * kXcds;
  chunk = (n + slabs * kCoeffWaves - 1) / (slabs * kCoeffWaves);
  chunk = (chunk + 3) / 4 * 4;
  if (int64_t(slabs) * tiles > (1ll << 31) - 1) throw Error(kInvalidArgument, "projector coefficients: grid");
  work.partial.ensure(size_t(tiles) * slabs * 256 * batch);
  hipLaunchKernelGGL((coeff_kernel<R>), dim3(static_cast<unsigned>(tiles * slabs), static_cast<unsigned>(batch)),
                     dim3(64 * kCoeffWaves), 0, s, B, ldb, psi, ldpsi, c, ldc, work.partial.get(), n, np, nb, chunk,
                     ptiles, static_cast<int>(tiles), Strides3{sB, sPsi, 0});
  HALF_HIP(hipGetLastError());
  const int64_t entries = int64_t(np) * nb;
  hipLaunchKernelGGL((coeff_reduce_kernel<cplx_t<R>>),
                     dim3(static_cast<unsigned>((entries + 255) / 256), static_cast<unsigned>(batch)), dim3(256), 0,
                     s, work.partial.get(), slabs, static_cast<int>(tiles), ptiles, c, ldc, np, nb, sC);
  HALF_HIP(hipGetLastError());
}

template <typename R>
void projector_expand(int n, int np, int nb, const cplx_t<R> *BD, const cplx_t<R> *BQ, int64_t ldb,
                      const cplx_t<R> *c, int64_t ldc, bool add, cplx_t<R> *H, cplx_t<R> *S, int64_t ldo,
                      hipStream_t s) {
  projector_expand_batched<R>(1, n, np, nb, BD, BQ, ldb, 0, c, ldc, 0, add, H, S, ldo, 0, s);
}

template <typename R>
void projector_expand_batched(int batch, int n, int np, int nb, const cplx_t<R> *BD, const cplx_t<R> *BQ,
                              int64_t ldb, int64_t sB, const cplx_t<R> *c, int64_t ldc, int64_t sC, bool add,
                              cplx_t<R> *H, cplx_t<R> *S, int64_t ldo, int64_t sO, hipStream_t s) {
  if (n <= 0 || nb <= 0 || batch <= 0) return;
  if (batch > 65535) throw Error(kInvalidArgument, "projector expansion: batch too large");
  if (np <= 0) throw Error(kInvalidArgument, "projector expansion: no projectors");
  if (ldb < n || ldc < np || ldo < n) throw Error(kInvalidArgument, "projector expansion: leading dimension");
  const int gt = (n + 15) / 16, bt = (nb + 15) / 16;
  const int64_t groups = int64_t((gt + kXcds - 1) / kXcds) * kXcds * bt;
  if (groups > (1ll << 31) - 1) throw Error(kInvalidArgument, "projector expansion: grid");
  int pchunk = (np + kExpandWaves - 1) / kExpandWaves;
  pchunk = (pchunk + 3) / 4 * 4;
  hipLaunchKernelGGL((expand_kernel<R>), dim3(static_cast<unsigned>(groups), static_cast<unsigned>(batch)),
                     dim3(64 * kExpandWaves), 0, s, BD, BQ, ldb, c, ldc, H, S, ldo, n, np, nb, add ? 1 : 0, pchunk,
                     gt, bt, Strides3{sB, sC, sO});
  HALF_HIP(hipGetLastError());
}

#define HALF_PROJECTOR(R)                                                                                          \
  template void projector_coefficients<R>(int, int, int, const cplx_t<R> *, int64_t, const cplx_t<R> *, int64_t,  \
                                          cplx_t<R> *, int64_t, CoeffWork<R> &, hipStream_t);                     \
  template void projector_coefficients_batched<R>(int, int, int, int, const cplx_t<R> *, int64_t, int64_t,        \
                                                  const cplx_t<R> *, int64_t, int64_t, cplx_t<R> *, int64_t,      \
                                                  int64_t, CoeffWork<R> &, hipStream_t);                          \
  template void projector_expand<R>(int, int, int, const cplx_t<R> *, const cplx_t<R> *, int64_t,                \
                                    const cplx_t<R> *, int64_t, bool, cplx_t<R> *, cplx_t<R> *, int64_t,          \
                                    hipStream_t);                                                                 \
  template void projector_expand_batched<R>(int, int, int, int, const cplx_t<R> *, const cplx_t<R> *, int64_t,   \
                                            int64_t, const cplx_t<R> *, int64_t, int64_t, bool, cplx_t<R> *,      \
                                            cplx_t<R> *, int64_t, int64_t, hipStream_t);
HALF_PROJECTOR(double)
HALF_PROJECTOR(float)

}  // namespace half
