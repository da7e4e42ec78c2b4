// Dense complex contractions on the gfx950 matrix cores.
//
// Two shapes occur on the operator and solver paths:
//   gram       C = A^H B    with A, B tall (n rows >> columns): the projector
//              coefficients B^H psi and the subspace Grams Z^H H Z, Z^H S Z,
//              X^H S P, P^H S P;
//   gemm_tall  C = A W      with A tall and W small: the projector expansion
//              B (D c), B (Q c).
// Both run 256-thread workgroups of 2 x 2 waves.  Each wave owns WT x WT
// 16x16 complex tiles (16 WT square), so a workgroup covers a 32 WT square
// of the output.  Operand slabs of 16 inner-index rows are read from HBM
// with each thread loading a short contiguous run of one column, stored to
// LDS split into real and imaginary planes, and consumed by
// v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32 (four real products per
// complex tile product).  The next slab is loaded into registers while the
// current one is multiplied.  gram splits the rows into chunks whose partial
// tiles are summed in a fixed order by a second kernel.
#include <algorithm>
#include <map>
#include <mutex>
#include <string>

#include "contract.hpp"
#include "cplx.hpp"
#include "mfma.hpp"

namespace half {

namespace {

constexpr int KS = 16;  // inner-index rows per LDS slab
// LDS row padding (elements): rows of 65 let the loaders' stores, four
// threads per column at rows 0, 4, 8, 12 (+i), fall in distinct banks
constexpr int PAD = 1;
constexpr int kXcds = 8;  // XCDs of an MI355X: workgroup id w is dispatched to XCD w % 8

template <typename R>
struct GramArgs {
  const cplx_t<R> *A;
  int64_t lda;
  const cplx_t<R> *B[2];
  int64_t ldb[2];
  cplx_t<R> *out[2];
  int64_t ldo[2];
  int64_t chunk_stride;  // distance between the partial tiles of consecutive chunks
  int64_t bA, bB[2], bOut[2];  // batch strides (blockIdx.z)
  int n, ka, kb, chunk, tiles_a;
  int upper;             // grid enumerates only tiles with ta <= tb (tiles_a per side)
};

// At least two waves per SIMD: with one (the widest FP64 forms otherwise
// need 258 registers) every barrier and slab store leaves the matrix core idle.
// S: the operands' storage type in memory, R: the type of the LDS slabs and
// the matrix-core products (S = double, R = float: FP32 products of FP64
// blocks, the FP32 phase of the mixed solve).
template <typename S, typename R, int WT, int NB, int PARTS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void gram_kernel(GramArgs<S> g) {
  using VS = cplx_t<S>;
  using A4 = typename Acc<R>::type;
  constexpr int TW = 32 * WT;        // workgroup tile width
  constexpr int LW = TW + PAD;       // LDS row length
  constexpr int L = TW * KS / 256;   // elements per thread and operand
  constexpr int TPC = KS / L;        // threads per column
  __shared__ R sa[2][KS][LW];
  __shared__ R sb[NB][2][KS][LW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int row = lane & 15, kq = lane >> 4;
  const int wa = (wave & 1) * 16 * WT, wb = (wave >> 1) * 16 * WT;
  int ta = blockIdx.x % g.tiles_a, tb = blockIdx.x / g.tiles_a;
  if (g.upper) {  // linear index over the upper-triangular tile pairs, column by column
    int t = blockIdx.x;
    tb = 0;
    while (t > tb) {
      t -= tb + 1;
      ++tb;
    }
    ta = t;
  }
  const int a0 = ta * TW, b0 = tb * TW;
  const int r_begin = blockIdx.y * g.chunk;
  const int r_end = min(g.n, r_begin + g.chunk);
  // loader: column lc of the tile, rows lr .. lr + L - 1 of the slab
  const int lc = tid / TPC, lr = (tid % TPC) * L;
  const bool ca = a0 + lc < g.ka, cb = b0 + lc < g.kb;
  const int64_t bz = blockIdx.z;
  const VS *pa = g.A + bz * g.bA + int64_t(min(a0 + lc, g.ka - 1)) * g.lda;
  const VS *pb[NB];
  for (int j = 0; j < NB; ++j) pb[j] = g.B[j] + bz * g.bB[j] + int64_t(min(b0 + lc, g.kb - 1)) * g.ldb[j];
  VS ra[L], rb[NB][L];
  auto load = [&](int r0) {
    for (int i = 0; i < L; ++i) {
      const int r = r0 + lr + i;
      const bool rv = r < r_end;
      ra[i] = (rv && ca) ? pa[r] : mk<VS>(0, 0);
      for (int j = 0; j < NB; ++j) rb[j][i] = (rv && cb) ? pb[j][r] : mk<VS>(0, 0);
    }
  };
  A4 acc[NB][WT][WT][PARTS];
  for (int j = 0; j < NB; ++j)
    for (int t = 0; t < WT; ++t)
      for (int u = 0; u < WT; ++u)
        for (int q = 0; q < PARTS; ++q) acc[j][t][u][q] = A4{0, 0, 0, 0};
  if (r_begin < r_end) load(r_begin);
  for (int r0 = r_begin; r0 < r_end; r0 += KS) {
    __syncthreads();
    for (int i = 0; i < L; ++i) {
      sa[0][lr + i][lc] = static_cast<R>(ra[i].x);
      sa[1][lr + i][lc] = -static_cast<R>(ra[i].y);  // conj(A)
      for (int j = 0; j < NB; ++j) {
        sb[j][0][lr + i][lc] = static_cast<R>(rb[j][i].x);
        sb[j][1][lr + i][lc] = static_cast<R>(rb[j][i].y);
      }
    }
    __syncthreads();
    if (r0 + KS < r_end) load(r0 + KS);
#pragma unroll
    for (int k4 = 0; k4 < KS / 4; ++k4) {
      const int kk = 4 * k4 + kq;
      R ar[WT], ai[WT];
      for (int t = 0; t < WT; ++t) {
        ar[t] = sa[0][kk][wa + 16 * t + row];
        ai[t] = sa[1][kk][wa + 16 * t + row];
      }
      for (int j = 0; j < NB; ++j)
        for (int u = 0; u < WT; ++u) {
          const R br = sb[j][0][kk][wb + 16 * u + row], bi = sb[j][1][kk][wb + 16 * u + row];
          for (int t = 0; t < WT; ++t) cmac<R>(ar[t], ai[t], br, bi, acc[j][t][u]);
        }
    }
  }
  const int64_t off = int64_t(blockIdx.y) * g.chunk_stride;
  for (int j = 0; j < NB; ++j)
    for (int t = 0; t < WT; ++t)
      for (int u = 0; u < WT; ++u) {
        const int cb2 = b0 + wb + 16 * u + row;
        if (cb2 >= g.kb) continue;
        for (int i = 0; i < 4; ++i) {
          const int ra2 = a0 + wa + 16 * t + mfma_row<R>(lane, i);
          if (ra2 < g.ka) {
            R re, im;
            cmac_result<R>(acc[j][t][u], i, re, im);
            g.out[j][bz * g.bOut[j] + off + ra2 + int64_t(cb2) * g.ldo[j]] =
                mk<VS>(static_cast<S>(re), static_cast<S>(im));
          }
        }
      }
}

// C[p + b ldc] = sum over chunks c of partial[c stride + p + b ka].  A
// workgroup takes 64 entries; its four waves sum the chunks c = q, q + 4,
// q + 8, ... (q = wave) with eight loads in flight per step, and the four
// partial sums are added in wave order: a fixed association, so the result
// does not depend on scheduling.
template <typename V>
__global__ __launch_bounds__(256) void gram_reduce_kernel(const V *partial, int64_t stride, int nchunks, int ka,
                                                          int kb, V *C, int64_t ldc, int64_t sC) {
  __shared__ V red[4][64];
  const int e = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int64_t t = blockIdx.x * int64_t(64) + e;
  const bool live = t < int64_t(ka) * kb;
  partial += blockIdx.y * stride * nchunks;  // batch entry blockIdx.y
  C += blockIdx.y * sC;
  using R = typename scalar_of<V>::type;
  V acc = mk<V>(R(0), R(0));
  if (live) {
    int c = q;
    for (; c + 28 < nchunks; c += 32) {
      V v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = partial[int64_t(c + 4 * i) * stride + t];
#pragma unroll
      for (int i = 0; i < 8; ++i) acc = cadd(acc, v[i]);
    }
    for (; c < nchunks; c += 4) acc = cadd(acc, partial[int64_t(c) * stride + t]);
  }
  red[q][e] = acc;
  __syncthreads();
  if (q == 0 && live) {
    const V s = cadd(cadd(red[0][e], red[1][e]), cadd(red[2][e], red[3][e]));
    const int b = static_cast<int>(t / ka), p = static_cast<int>(t - int64_t(b) * ka);
    C[p + int64_t(b) * ldc] = s;
  }
}

// C[j + i ldc] = conj(C[i + j ldc]) for i < j (lower triangle from the upper)
template <typename V>
__global__ void mirror_lower_kernel(V *C, int64_t ldc, int k, int64_t sC) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(k) * k) return;
  C += blockIdx.y * sC;
  const int j = static_cast<int>(t / k), i = static_cast<int>(t - int64_t(j) * k);
  if (i <= j) return;  // (i, j) strictly lower: row i > column j
  const V u = C[j + int64_t(i) * ldc];
  C[i + int64_t(j) * ldc] = mk<V>(u.x, -u.y);
}

template <typename R>
struct TallArgs {
  const cplx_t<R> *A;
  int64_t lda;
  const cplx_t<R> *W[2];
  int64_t ldw[2];
  cplx_t<R> *C[2];
  int64_t ldc[2];
  int64_t bA, bW[2], bC[2];  // batch strides (blockIdx.z)
  int n, m, k;
  R alpha, beta;
  int split;  // 1: the third grid index selects the output (W[z], C[z]) of one problem, not a batch entry
  // 0: grid (row tiles, column tiles, z), row tiles fastest; 1: grid
  // (column tiles x outputs when split, row tiles, batch), so consecutive
  // workgroups share one row slab of A; 2: as 1, with the workgroups of
  // each z renumbered so that every XCD (dispatch round-robins consecutive
  // workgroups over the 8) takes a contiguous range of row tiles
  int order, col_tiles;
};

// S: storage type, R: product type (as for gram_kernel); the epilogue
// (alpha, beta C) runs in S.
template <typename S, typename R, int WT, int NB, int PARTS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void tall_kernel(TallArgs<S> g) {
  using VS = cplx_t<S>;
  using A4 = typename Acc<R>::type;
  constexpr int TW = 32 * WT;
  constexpr int LW = TW + PAD;
  constexpr int L = TW * KS / 256;
  constexpr int TPA = TW / L;  // A loader: threads per inner-index column
  constexpr int TPW = KS / L;  // W loader: threads per output column
  __shared__ R sa[2][KS][LW];       // [inner][row]
  __shared__ R sw[NB][2][KS][LW];   // [inner][column]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int row = lane & 15, kq = lane >> 4;
  const int wr = (wave & 1) * 16 * WT, wc = (wave >> 1) * 16 * WT;
  int tr = blockIdx.x, tc = blockIdx.y, tz = blockIdx.z;
  if (g.order) {
    int bx = blockIdx.x, by = blockIdx.y;
    if (g.order == 2) {
      const int total = gridDim.x * gridDim.y, hw = blockIdx.x + gridDim.x * blockIdx.y;
      const int per = total / kXcds, extra = total % kXcds, xcd = hw % kXcds;
      const int lin = xcd * per + min(xcd, extra) + hw / kXcds;
      bx = lin % gridDim.x;
      by = lin / gridDim.x;
    }
    tc = bx % g.col_tiles;
    tr = by;
    tz = g.split ? bx / g.col_tiles : blockIdx.z;
  }
  const int g0 = tr * TW, c0 = tc * TW;
  const int akc = tid / TPA, ar0 = tid % TPA;  // A: inner column akc, rows ar0 + i TPA
  const int wcol = tid / TPW, wk0 = (tid % TPW) * L;  // W: column wcol, inner rows wk0..
  const bool wcv = c0 + wcol < g.m;
  const int64_t bz = g.split ? 0 : tz;
  const int zo = g.split ? tz : 0;  // first output of this workgroup
  VS ra[L], rw[NB][L];
  auto load = [&](int k0) {
    const int kc = k0 + akc;
    const bool kv = kc < g.k;
    const VS *pa = g.A + bz * g.bA + int64_t(min(kc, g.k - 1)) * g.lda;
    for (int i = 0; i < L; ++i) {
      const int r = g0 + ar0 + i * TPA;
      ra[i] = (kv && r < g.n) ? pa[r] : mk<VS>(0, 0);
    }
    for (int j = 0; j < NB; ++j) {
      const VS *pw = g.W[j + zo] + bz * g.bW[j + zo] + int64_t(min(c0 + wcol, g.m - 1)) * g.ldw[j + zo];
      for (int i = 0; i < L; ++i) {
        const int kr = k0 + wk0 + i;
        rw[j][i] = (wcv && kr < g.k) ? pw[kr] : mk<VS>(0, 0);
      }
    }
  };
  A4 acc[NB][WT][WT][PARTS];
  for (int j = 0; j < NB; ++j)
    for (int t = 0; t < WT; ++t)
      for (int u = 0; u < WT; ++u)
        for (int q = 0; q < PARTS; ++q) acc[j][t][u][q] = A4{0, 0, 0, 0};
  load(0);
  for (int k0 = 0; k0 < g.k; k0 += KS) {
    __syncthreads();
    for (int i = 0; i < L; ++i) {
      sa[0][akc][ar0 + i * TPA] = static_cast<R>(ra[i].x);
      sa[1][akc][ar0 + i * TPA] = static_cast<R>(ra[i].y);
      for (int j = 0; j < NB; ++j) {
        sw[j][0][wk0 + i][wcol] = static_cast<R>(rw[j][i].x);
        sw[j][1][wk0 + i][wcol] = static_cast<R>(rw[j][i].y);
      }
    }
    __syncthreads();
    if (k0 + KS < g.k) load(k0 + KS);
#pragma unroll
    for (int k4 = 0; k4 < KS / 4; ++k4) {
      const int kk = 4 * k4 + kq;
      R xr[WT], xi[WT];
      for (int t = 0; t < WT; ++t) {
        xr[t] = sa[0][kk][wr + 16 * t + row];
        xi[t] = sa[1][kk][wr + 16 * t + row];
      }
      for (int j = 0; j < NB; ++j)
        for (int u = 0; u < WT; ++u) {
          const R yr = sw[j][0][kk][wc + 16 * u + row], yi = sw[j][1][kk][wc + 16 * u + row];
          for (int t = 0; t < WT; ++t) cmac<R>(xr[t], xi[t], yr, yi, acc[j][t][u]);
        }
    }
  }
  // epilogue per output: with beta != 0 every old value of C this thread
  // updates is loaded first (independent loads in flight together) and
  // only then are the results stored
  for (int j = 0; j < NB; ++j) {
    VS *const Cj = g.C[j + zo] + bz * g.bC[j + zo];
    const int64_t ldcj = g.ldc[j + zo];
    VS old[WT][WT][4];
    if (g.beta != S(0)) {
      for (int u = 0; u < WT; ++u)
        for (int t = 0; t < WT; ++t)
          for (int i = 0; i < 4; ++i) {
            const int col = c0 + wc + 16 * u + row;
            const int r = g0 + wr + 16 * t + mfma_row<R>(lane, i);
            old[u][t][i] = (col < g.m && r < g.n) ? Cj[r + int64_t(col) * ldcj] : mk<VS>(S(0), S(0));
          }
    }
    for (int u = 0; u < WT; ++u) {
      const int col = c0 + wc + 16 * u + row;
      if (col >= g.m) continue;
      for (int t = 0; t < WT; ++t)
        for (int i = 0; i < 4; ++i) {
          const int r = g0 + wr + 16 * t + mfma_row<R>(lane, i);
          if (r >= g.n) continue;
          R pr, pi;
          cmac_result<R>(acc[j][t][u], i, pr, pi);
          S vr = static_cast<S>(pr) * g.alpha, vi = static_cast<S>(pi) * g.alpha;
          if (g.beta != S(0)) {
            vr += g.beta * old[u][t][i].x;
            vi += g.beta * old[u][t][i].y;
          }
          Cj[r + int64_t(col) * ldcj] = mk<VS>(vr, vi);
        }
    }
  }
}

template <typename S, typename R, int WT, int NB, int PARTS>
void launch_gram(dim3 grid, const GramArgs<S> &a, hipStream_t s) {
  hipLaunchKernelGGL((gram_kernel<S, R, WT, NB, PARTS>), grid, dim3(256), 0, s, a);
}

template <typename S, typename R, int WT, int NB, int PARTS>
void launch_tall(dim3 grid, const TallArgs<S> &a, hipStream_t s) {
  hipLaunchKernelGGL((tall_kernel<S, R, WT, NB, PARTS>), grid, dim3(256), 0, s, a);
}

}  // namespace

namespace {

// Tuning (half_amd_set_tuning keys contract.<fp64|fp32>.*; index 0 FP64,
// 1 FP32): tile width of gram and gemm_tall (0 = the shape rule below, 1 =
// 32-wide, 2 = 64-wide) and the fill a gram row-chunk count must reach in
// every round of workgroups (percent).
int g_gram_wt[2] = {0, 0};
int g_gram_fill[2] = {90, 90};
int g_tall_wt[2] = {0, 0};
// Workgroup order of gemm_tall (TallArgs::order).  Column tiles first (1):
// the workgroups sharing a row slab of A run together, so the tall operand
// comes from HBM about once instead of once per column tile and output
// (216-atom cell, 512 bands, B (D c), B (Q c): 8.5 -> 7.7 ms FP64,
// 4.1 -> 3.95 ms FP32; the XCD-contiguous form 2: 7.9 ms).
int g_tall_order[2] = {1, 1};
// FP64 complex products with three MFMAs (Gauss) instead of four where a
// workgroup has one output; the expansion with both outputs then runs as one
// launch with one output per workgroup (grid.z), which keeps a 64-wide
// tile's three accumulators within the register budget of three waves per
// SIMD.  216-atom cell, 512 bands: B^H psi 4.70 -> 3.96 ms, B (D c), B (Q c)
// 9.52 -> 8.05 ms.  FP32 always uses three.
int g_gram_gauss = 1;
int g_tall_gauss = 1;

template <typename R>
constexpr int prec_index() {
  return sizeof(R) == 8 ? 0 : 1;
}

// Tile width, grid and row chunking of a gram launch.
struct GramPlan {
  int wt, tiles_a, tiles, nchunks, chunk;
  size_t work_elems;  // partial tiles (all outputs and batch entries); 0 without row chunks
};

// FP32 products for FP64 blocks, set for the calling thread by
// ContractFp32Scope (the FP32 phase of a mixed solve).
thread_local bool t_fp32_products = false;

// Product form of a launch on blocks of type S: 0 S products with four
// MFMAs per complex product (FP64 only), 1 S products with three, 2 FP32
// products (three MFMAs) of FP64 blocks.
template <typename S>
int product_form(bool gauss_ok) {
  if (sizeof(S) == 4) return 1;
  if (t_fp32_products) return 2;
  return gauss_ok ? 1 : 0;
}

// Workgroups of gram_kernel<S, R, WT, NB, PARTS> the device holds at once
// (occupancy per CU x CUs of the current device), cached per variant and
// device.
template <typename S, typename R, int WT, int NB, int PARTS>
int gram_capacity() {
  static std::mutex mu;
  static std::map<int, int> cache;
  int dev = 0;
  HALF_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  int per_cu = 0, cus = 0;
  HALF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void *>(&gram_kernel<S, R, WT, NB, PARTS>), 256, 0));
  HALF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int cap = std::max(1, per_cu) * std::max(1, cus);
  cache[dev] = cap;
  return cap;
}

// (the two-output 64-wide FP64 tile would spill with three FP64 accumulators)
template <typename S>
int gram_form(int nout) {
  return product_form<S>(g_gram_gauss && nout == 1);
}

template <typename S, typename R, int PARTS>
int gram_capacity_wt(int wt, int nout) {
  if (wt == 1) return nout == 1 ? gram_capacity<S, R, 1, 1, PARTS>() : gram_capacity<S, R, 1, 2, PARTS>();
  return nout == 1 ? gram_capacity<S, R, 2, 1, PARTS>() : gram_capacity<S, R, 2, 2, PARTS>();
}

template <typename S>
int gram_capacity_of(int wt, int nout) {
  const int form = gram_form<S>(nout);
  if constexpr (std::is_same<S, double>::value) {
    if (form == 0) return gram_capacity_wt<double, double, 2>(wt, nout);
    if (form == 2) return gram_capacity_wt<double, float, 3>(wt, nout);
    return gram_capacity_wt<double, double, 3>(wt, nout);
  } else {
    return gram_capacity_wt<float, float, 3>(wt, nout);
  }
}

template <typename S, typename R, int PARTS>
void launch_gram_wt(int wt, int nout, dim3 grid, const GramArgs<S> &a, hipStream_t s) {
  if (wt == 1)
    nout == 1 ? launch_gram<S, R, 1, 1, PARTS>(grid, a, s) : launch_gram<S, R, 1, 2, PARTS>(grid, a, s);
  else
    nout == 1 ? launch_gram<S, R, 2, 1, PARTS>(grid, a, s) : launch_gram<S, R, 2, 2, PARTS>(grid, a, s);
}

// Row chunks for `units` output tiles (over the batch) on a device holding
// `cap` workgroups at once: the fewest chunks that fill the device in rounds
// at least `fill` percent full (default 90: a grid of 1.1 rounds leaves the
// second round nearly empty), at most 64 and at least 64 rows each; failing
// that, the count with the fullest rounds.
int choose_chunks(int units, int n, int cap, int fill) {
  const int max_chunks = std::max(1, std::min(64, (n + 63) / 64));
  int best = 1;
  double best_eff = -1;
  for (int c = 1; c <= max_chunks; ++c) {
    const int64_t total = int64_t(units) * c;
    const int64_t rounds = (total + cap - 1) / cap;
    const double eff = double(total) / double(rounds * cap);
    if (total >= cap && eff >= 0.01 * fill) return c;
    if (eff > best_eff + 1e-9) {
      best_eff = eff;
      best = c;
    }
  }
  return best;
}

GramPlan plan_gram(int batch, int n, int ka, int kb, int nout, bool hermitian, int cap_wt1, int cap_wt2, int wt_force,
                   int fill) {
  GramPlan p{};
  // 64-wide tiles; for the subspace Grams (Hermitian, matrix-core bound)
  // 32-wide ones where they cover the result with less padding (80 columns:
  // 96 against 128) below a few hundred columns, where the padding of the
  // wider tiles stops mattering and their halved operand re-reads win
  const int kmax = std::max(ka, kb);
  const bool pad_less_32 = (kmax + 31) / 32 * 32 < (kmax + 63) / 64 * 64;
  p.wt = wt_force ? wt_force : (kmax <= 32 || (hermitian && pad_less_32 && kmax < 256)) ? 1 : 2;
  const int tw = 32 * p.wt;
  p.tiles_a = (ka + tw - 1) / tw;
  p.tiles = hermitian ? p.tiles_a * (p.tiles_a + 1) / 2 : p.tiles_a * ((kb + tw - 1) / tw);
  int nchunks = choose_chunks(p.tiles * batch, n, p.wt == 1 ? cap_wt1 : cap_wt2, fill);
  int chunk = (n + nchunks - 1) / nchunks;
  chunk = (chunk + KS - 1) / KS * KS;
  p.nchunks = (n + chunk - 1) / chunk;
  p.chunk = chunk;
  p.work_elems = p.nchunks > 1 ? size_t(ka) * kb * p.nchunks * batch * nout : 0;
  return p;
}

// gram_batched with its partial tiles in `work` (at least plan.work_elems
// elements) or, when `grow` is given, in that array grown to fit
template <typename R>
void gram_batched_work(int batch, int n, int ka, int kb, const cplx_t<R> *A, int64_t lda, int64_t sA,
                       const cplx_t<R> *B, int64_t ldb, int64_t sB, cplx_t<R> *C, int64_t ldc, int64_t sC,
                       const cplx_t<R> *B2, int64_t ldb2, int64_t sB2, cplx_t<R> *C2, int64_t ldc2, int64_t sC2,
                       DeviceArray<cplx_t<R>> *grow, cplx_t<R> *work, size_t work_elems, hipStream_t s,
                       bool hermitian) {
  using V = cplx_t<R>;
  if (ka <= 0 || kb <= 0 || batch <= 0) return;
  if (batch > 65535) throw Error(kInvalidArgument, "gram: batch too large");
  if (hermitian && ka != kb) throw Error(kInvalidArgument, "gram: a Hermitian result must be square");
  const int nb = B2 ? 2 : 1;
  if (n <= 0) {
    for (int b = 0; b < batch; ++b) {
      HALF_HIP(hipMemset2DAsync(C + b * sC, sizeof(V) * ldc, 0, sizeof(V) * ka, kb, s));
      if (B2) HALF_HIP(hipMemset2DAsync(C2 + b * sC2, sizeof(V) * ldc2, 0, sizeof(V) * ka, kb, s));
    }
    return;
  }
  if (lda < n || ldb < n || ldc < ka || (B2 && (ldb2 < n || ldc2 < ka)))
    throw Error(kInvalidArgument, "gram: leading dimension too small");
  const GramPlan plan = plan_gram(batch, n, ka, kb, nb, hermitian, gram_capacity_of<R>(1, nb),
                                  gram_capacity_of<R>(2, nb), g_gram_wt[prec_index<R>()],
                                  g_gram_fill[prec_index<R>()]);
  const int nchunks = plan.nchunks;
  if (plan.work_elems > 0) {
    if (grow) {
      grow->ensure(plan.work_elems);
      work = grow->get();
    } else if (!work || work_elems < plan.work_elems) {
      throw Error(kInternal, "gram: workspace too small");
    }
  }
  GramArgs<R> a{};
  a.A = A;
  a.lda = lda;
  a.bA = sA;
  a.B[0] = B;
  a.ldb[0] = ldb;
  a.bB[0] = sB;
  a.B[1] = B2 ? B2 : B;
  a.ldb[1] = B2 ? ldb2 : ldb;
  a.bB[1] = B2 ? sB2 : sB;
  a.n = n;
  a.ka = ka;
  a.kb = kb;
  a.chunk = plan.chunk;
  a.tiles_a = plan.tiles_a;
  a.upper = hermitian ? 1 : 0;
  const int64_t tile_elems = int64_t(ka) * kb;
  if (nchunks == 1) {
    a.out[0] = C;
    a.ldo[0] = ldc;
    a.bOut[0] = sC;
    a.out[1] = C2;
    a.ldo[1] = ldc2;
    a.bOut[1] = sC2;
    a.chunk_stride = 0;
  } else {
    a.out[0] = work;
    a.out[1] = work + tile_elems * nchunks * batch;
    a.ldo[0] = a.ldo[1] = ka;
    a.bOut[0] = a.bOut[1] = tile_elems * nchunks;
    a.chunk_stride = tile_elems;
  }
  const dim3 grid(plan.tiles, nchunks, batch);
  if constexpr (std::is_same<R, double>::value) {
    const int form = gram_form<R>(nb);
    if (form == 0)
      launch_gram_wt<double, double, 2>(plan.wt, nb, grid, a, s);
    else if (form == 2)
      launch_gram_wt<double, float, 3>(plan.wt, nb, grid, a, s);
    else
      launch_gram_wt<double, double, 3>(plan.wt, nb, grid, a, s);
  } else {
    launch_gram_wt<float, float, 3>(plan.wt, nb, grid, a, s);
  }
  HALF_HIP(hipGetLastError());
  if (nchunks > 1) {
    const dim3 rg(static_cast<unsigned>((tile_elems + 63) / 64), batch);
    hipLaunchKernelGGL((gram_reduce_kernel<V>), rg, dim3(256), 0, s, a.out[0], tile_elems, nchunks, ka, kb, C, ldc,
                       sC);
    if (B2)
      hipLaunchKernelGGL((gram_reduce_kernel<V>), rg, dim3(256), 0, s, a.out[1], tile_elems, nchunks, ka, kb, C2,
                         ldc2, sC2);
    HALF_HIP(hipGetLastError());
  }
  if (hermitian) {
    const dim3 mg(static_cast<unsigned>((int64_t(ka) * ka + 255) / 256), batch);
    hipLaunchKernelGGL((mirror_lower_kernel<V>), mg, dim3(256), 0, s, C, ldc, ka, sC);
    if (B2) hipLaunchKernelGGL((mirror_lower_kernel<V>), mg, dim3(256), 0, s, C2, ldc2, ka, sC2);
    HALF_HIP(hipGetLastError());
  }
}

}  // namespace

template <typename R>
void gram_batched(int batch, int n, int ka, int kb, const cplx_t<R> *A, int64_t lda, int64_t sA, const cplx_t<R> *B,
                  int64_t ldb, int64_t sB, cplx_t<R> *C, int64_t ldc, int64_t sC, const cplx_t<R> *B2, int64_t ldb2,
                  int64_t sB2, cplx_t<R> *C2, int64_t ldc2, int64_t sC2, DeviceArray<cplx_t<R>> &work, hipStream_t s,
                  bool hermitian) {
  gram_batched_work<R>(batch, n, ka, kb, A, lda, sA, B, ldb, sB, C, ldc, sC, B2, ldb2, sB2, C2, ldc2, sC2, &work,
                       nullptr, 0, s, hermitian);
}

template <typename R>
size_t gram_work_elems(int n, int ka, int kb) {
  if (n <= 0 || ka <= 0 || kb <= 0) return 0;
  return plan_gram(1, n, ka, kb, 1, false, gram_capacity_of<R>(1, 1), gram_capacity_of<R>(2, 1),
                   g_gram_wt[prec_index<R>()], g_gram_fill[prec_index<R>()])
      .work_elems;
}

template <typename R>
void gram_into(int n, int ka, int kb, const cplx_t<R> *A, int64_t lda, const cplx_t<R> *B, int64_t ldb, cplx_t<R> *C,
               int64_t ldc, cplx_t<R> *work, size_t work_elems, hipStream_t s) {
  gram_batched_work<R>(1, n, ka, kb, A, lda, 0, B, ldb, 0, C, ldc, 0, nullptr, 0, 0, nullptr, 0, 0, nullptr, work,
                       work_elems, s, false);
}

template <typename R>
void gram(int n, int ka, int kb, const cplx_t<R> *A, int64_t lda, const cplx_t<R> *B, int64_t ldb, cplx_t<R> *C,
          int64_t ldc, const cplx_t<R> *B2, int64_t ldb2, cplx_t<R> *C2, int64_t ldc2,
          DeviceArray<cplx_t<R>> &work, hipStream_t s, bool hermitian) {
  gram_batched<R>(1, n, ka, kb, A, lda, 0, B, ldb, 0, C, ldc, 0, B2, ldb2, 0, C2, ldc2, 0, work, s, hermitian);
}

template <typename R>
void gemm_tall_batched(int batch, int n, int m, int k, R alpha, const cplx_t<R> *A, int64_t lda, int64_t sA,
                       const cplx_t<R> *W, int64_t ldw, int64_t sW, R beta, cplx_t<R> *C, int64_t ldc, int64_t sC,
                       const cplx_t<R> *W2, int64_t ldw2, int64_t sW2, cplx_t<R> *C2, int64_t ldc2, int64_t sC2,
                       hipStream_t s) {
  if (n <= 0 || m <= 0 || batch <= 0) return;
  if (k <= 0) throw Error(kInvalidArgument, "gemm_tall: empty inner dimension");
  if (batch > 65535) throw Error(kInvalidArgument, "gemm_tall: batch too large");
  if (lda < n || ldw < k || ldc < n || (W2 && (ldw2 < k || ldc2 < n)))
    throw Error(kInvalidArgument, "gemm_tall: leading dimension too small");
  const int nb = W2 ? 2 : 1;
  const bool inplace = C == A || (W2 && C2 == A);
  if (inplace) {
    if (m > gemm_tall_inplace_max || (C == A && (ldc != lda || sC != sA)) ||
        (W2 && C2 == A && (ldc2 != lda || sC2 != sA)))
      throw Error(kInvalidArgument, "gemm_tall: in-place form needs C = A with the same layout and m <= 64");
  }
  // 64-wide tiles when they still give a few hundred workgroups (always in
  // place: one column tile, so a workgroup owns whole rows)
  const int64_t t64 = int64_t((n + 63) / 64) * ((m + 63) / 64) * batch;
  const int forced = g_tall_wt[prec_index<R>()];
  const int wt = inplace ? 2 : forced ? forced : t64 >= 512 ? 2 : 1;
  const int tw = 32 * wt;
  TallArgs<R> a{};
  a.A = A;
  a.lda = lda;
  a.bA = sA;
  a.W[0] = W;
  a.ldw[0] = ldw;
  a.bW[0] = sW;
  a.C[0] = C;
  a.ldc[0] = ldc;
  a.bC[0] = sC;
  a.W[1] = W2 ? W2 : W;
  a.ldw[1] = W2 ? ldw2 : ldw;
  a.bW[1] = W2 ? sW2 : sW;
  a.C[1] = W2 ? C2 : C;
  a.ldc[1] = W2 ? ldc2 : ldc;
  a.bC[1] = W2 ? sC2 : sC;
  a.n = n;
  a.m = m;
  a.k = k;
  a.alpha = alpha;
  a.beta = beta;
  // FP64 Gauss form: both outputs of one problem as grid.z = 2, one each
  const bool fp32_products = sizeof(R) == 8 && t_fp32_products;
  const bool split = sizeof(R) == 8 && !fp32_products && g_tall_gauss && nb == 2 && batch == 1 && !inplace;
  a.split = split ? 1 : 0;
  const int row_tiles = (n + tw - 1) / tw, col_tiles = (m + tw - 1) / tw;
  a.col_tiles = col_tiles;
  a.order = row_tiles <= 65535 ? g_tall_order[prec_index<R>()] : 0;
  const dim3 grid = a.order ? dim3(col_tiles * (split ? 2 : 1), row_tiles, split ? 1 : batch)
                            : dim3(row_tiles, col_tiles, split ? 2 : batch);
  if (col_tiles > 65535) throw Error(kInvalidArgument, "gemm_tall: too many columns");
  const int nbl = split ? 1 : nb;  // outputs per workgroup
  if constexpr (std::is_same<R, double>::value) {
    // (in place too: a workgroup owns whole rows and has read all of them
    // before its epilogue stores, whichever product form it runs)
    const int form = product_form<double>(g_tall_gauss && nbl == 1);
    auto go = [&](auto s_tag, auto r_tag, auto parts) {
      using SS = decltype(s_tag);
      using RR = decltype(r_tag);
      constexpr int PP = decltype(parts)::value;
      if (wt == 1)
        nbl == 1 ? launch_tall<SS, RR, 1, 1, PP>(grid, a, s) : launch_tall<SS, RR, 1, 2, PP>(grid, a, s);
      else
        nbl == 1 ? launch_tall<SS, RR, 2, 1, PP>(grid, a, s) : launch_tall<SS, RR, 2, 2, PP>(grid, a, s);
    };
    if (form == 0)
      go(double(), double(), std::integral_constant<int, 2>());
    else if (form == 2)
      go(double(), float(), std::integral_constant<int, 3>());
    else
      go(double(), double(), std::integral_constant<int, 3>());
  } else {
    if (wt == 1)
      nbl == 1 ? launch_tall<float, float, 1, 1, 3>(grid, a, s) : launch_tall<float, float, 1, 2, 3>(grid, a, s);
    else
      nbl == 1 ? launch_tall<float, float, 2, 1, 3>(grid, a, s) : launch_tall<float, float, 2, 2, 3>(grid, a, s);
  }
  HALF_HIP(hipGetLastError());
}

template <typename R>
void gemm_tall(int n, int m, int k, R alpha, const cplx_t<R> *A, int64_t lda, const cplx_t<R> *W, int64_t ldw,
               R beta, cplx_t<R> *C, int64_t ldc, const cplx_t<R> *W2, int64_t ldw2, cplx_t<R> *C2, int64_t ldc2,
               hipStream_t s) {
  gemm_tall_batched<R>(1, n, m, k, alpha, A, lda, 0, W, ldw, 0, beta, C, ldc, 0, W2, ldw2, 0, C2, ldc2, 0, s);
}

ContractFp32Scope::ContractFp32Scope(bool on) : prev_(t_fp32_products) { t_fp32_products = on; }
ContractFp32Scope::~ContractFp32Scope() { t_fp32_products = prev_; }
bool contract_fp32_products() { return t_fp32_products; }

bool set_contract_tuning(const std::string &key, int value) {
  for (int p = 0; p < 2; ++p) {
    const std::string pre = p == 0 ? "contract.fp64." : "contract.fp32.";
    if (key == pre + "gram.wt" || key == pre + "tall.wt") {
      if (value < 0 || value > 2) return false;
      (key == pre + "gram.wt" ? g_gram_wt : g_tall_wt)[p] = value;
      return true;
    }
    if (key == pre + "gram.fill") {
      if (value < 10 || value > 100) return false;
      g_gram_fill[p] = value;
      return true;
    }
    if (key == pre + "tall.order") {
      if (value < 0 || value > 2) return false;
      g_tall_order[p] = value;
      return true;
    }
  }
  if (key == "contract.fp64.gram.gauss" || key == "contract.fp64.tall.gauss") {
    if (value < 0 || value > 1) return false;
    (key == "contract.fp64.gram.gauss" ? g_gram_gauss : g_tall_gauss) = value;
    return true;
  }
  return false;
}

bool get_contract_tuning(const std::string &key, int &value) {
  for (int p = 0; p < 2; ++p) {
    const std::string pre = p == 0 ? "contract.fp64." : "contract.fp32.";
    if (key == pre + "gram.wt") {
      value = g_gram_wt[p];
      return true;
    }
    if (key == pre + "tall.wt") {
      value = g_tall_wt[p];
      return true;
    }
    if (key == pre + "gram.fill") {
      value = g_gram_fill[p];
      return true;
    }
    if (key == pre + "tall.order") {
      value = g_tall_order[p];
      return true;
    }
  }
  if (key == "contract.fp64.gram.gauss" || key == "contract.fp64.tall.gauss") {
    value = key == "contract.fp64.gram.gauss" ? g_gram_gauss : g_tall_gauss;
    return true;
  }
  return false;
}

#define HALF_CONTRACT(R)                                                                                     \
  template void gram<R>(int, int, int, const cplx_t<R> *, int64_t, const cplx_t<R> *, int64_t, cplx_t<R> *, \
                        int64_t, const cplx_t<R> *, int64_t, cplx_t<R> *, int64_t, DeviceArray<cplx_t<R>> &, \
                        hipStream_t, bool);                                                                  \
  template size_t gram_work_elems<R>(int, int, int);                                                        \
  template void gram_into<R>(int, int, int, const cplx_t<R> *, int64_t, const cplx_t<R> *, int64_t, cplx_t<R> *,  \
                             int64_t, cplx_t<R> *, size_t, hipStream_t);                                     \
  template void gram_batched<R>(int, int, int, int, const cplx_t<R> *, int64_t, int64_t, const cplx_t<R> *,    \
                                int64_t, int64_t, cplx_t<R> *, int64_t, int64_t, const cplx_t<R> *, int64_t,   \
                                int64_t, cplx_t<R> *, int64_t, int64_t, DeviceArray<cplx_t<R>> &, hipStream_t, \
                                bool);                                                                       \
  template void gemm_tall<R>(int, int, int, R, const cplx_t<R> *, int64_t, const cplx_t<R> *, int64_t, R,       \
                             cplx_t<R> *, int64_t, const cplx_t<R> *, int64_t, cplx_t<R> *, int64_t, hipStream_t); \
  template void gemm_tall_batched<R>(int, int, int, int, R, const cplx_t<R> *, int64_t, int64_t,               \
                                     const cplx_t<R> *, int64_t, int64_t, R, cplx_t<R> *, int64_t, int64_t,    \
                                     const cplx_t<R> *, int64_t, int64_t, cplx_t<R> *, int64_t, int64_t,       \
                                     hipStream_t);
HALF_CONTRACT(double)
HALF_CONTRACT(float)
#undef HALF_CONTRACT

}  // namespace half
