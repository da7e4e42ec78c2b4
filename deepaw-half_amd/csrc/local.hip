// Pruned local stage of H.psi: T psi + F[Veff F^-1[psi]] / N_r.
//
// Only the plane-wave sphere is populated in G space, so most 1D transforms
// of a full 3D FFT act on zeros.  With the C-order box (i1 slowest, i3
// contiguous) and the basis enumerated n1, n2, n3 ascending (so the plane
// waves of one (i1, i2) line are contiguous), the transform is done in five
// kernels that touch only what is needed:
//
//   z_bwd   active (i1,i2) lines: plane waves scattered into LDS straight
//           from psi (no box clearing), backward transform along i3
//   y_bwd   active i1 planes: backward transform along i2, loading only the
//           active lines of the plane (the rest are zeros, never stored)
//   x_mfma  every (i2,i3) column: restricted to the active i1 planes the
//           backward transform along i1, the multiply by Veff/N_r and the
//           forward transform are one circular convolution,
//             out[p'] = sum_p Vx[(p' - p) mod n1][c] in[p],
//           with Vx[d][c] = sum_x Veff[x][c]/N_r e^{-2 pi i d x/n1} computed
//           once per potential, run as a Toeplitz product on the matrix
//           cores (x_conv: the same on the vector units); each active plane
//           is read and written once (x_veff, the transform-based form,
//           serves boxes whose active plane count exceeds both)
//   y_fwd   active planes: forward transform along i2, store active lines
//   z_fwd   active lines: forward transform along i3, gather the plane-wave
//           coefficients and add T psi (S psi = psi is written here too)
//
// Every kernel serves a batch of k points that share the box (one launch per
// stage for all of them): the k point of a workgroup is a grid coordinate,
// its index tables come from a device array, and its states and boxes sit
// at fixed strides.  Ragged sizes (lines and planes differ between k points)
// are handled by launching for the largest and retiring workgroups past a k
// point's own count before they touch anything.
//
// Box entries outside the active lines/planes are never read or written.
// Numerically this is the same circular convolution as the reference's
// scatter / inverse FFT / multiply / forward FFT / gather sequence
// (reference: src/half_cuda_assembly.cuf:180-218, 484-498).
#include <algorithm>

#include "local.hpp"
#include "mfma.hpp"
#include "stockham.hpp"
#include "trace.hpp"

namespace half {

namespace {

// Tunables (FFT points per workgroup tile per kernel class and precision,
// and bands sharing one Veff tile in the x pass); set through
// half_amd_set_tuning, chosen by the autotuner in half_amd/autotune.py.
struct Tuning {
  // defaults from the autotuner on the 48^3 HfO2-like box, 64 bands
  int tile[2][3] = {{1536, 1024, 768}, {3072, 2048, 1024}};  // [fp64|fp32][x|y|z]
  int xgroup[2] = {8, 8};
  // x pass where every k point's active planes are one run: 0 transforms,
  // 1 per-column convolution on the vector units, 2 the same as a Toeplitz
  // product on the matrix cores (four waves per workgroup), 3 the same with
  // eight waves and three-MFMA products (96^3 box, 41 planes, 512 bands:
  // FP64 1.64 ms (3) / 1.75 ms (2) against 3.22 ms (1) and 8.9 ms (0); FP32
  // 0.92 ms (2) / 0.94 ms (3) against 1.49 ms (1))
  int xconv[2] = {3, 2};
  int fused[2] = {1, 1};  // y, x, y passes in one kernel where the convolution applies
  int ygroup[2] = {4, 4}; // bands per workgroup of the fused kernel (fewer where LDS is short)
  int vxlds[2] = {1, 1};  // fused kernel with the slice's Vx rows in LDS where it applies
};
Tuning g_tune;

// acc += c v as four fused multiply-adds (the contraction of
// acc + (c.x v.x - c.y v.y) would leave a multiply and an add per part)
__device__ inline float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <typename V>
__device__ inline void cmac(V &acc, const V &c, const V &v) {
  acc.x = fma_r(c.x, v.x, acc.x);
  acc.x = fma_r(-c.y, v.y, acc.x);
  acc.y = fma_r(c.x, v.y, acc.y);
  acc.y = fma_r(c.y, v.x, acc.y);
}

// Elements per thread of a tile (M * n <= tile budget <= 256 * KMAX).
template <typename R>
struct Tile {
  static constexpr int KMAX = std::is_same<R, double>::value ? 8 : 16;
};

// k-point strides of the batch (elements)
struct KStride {
  long long psi, out, kin, box;
};

// Lines of one band per workgroup: blockIdx.x = block of a.M lines,
// blockIdx.y = band, blockIdx.z = k point.  NN > 0: the transform length
// n3 as a compile-time constant (runtime a.n otherwise).  a.M is a power of
// two, so the element -> (line, point) splits below are shifts and masks.
// TL: the transformed lines go to an i3-major line buffer, [band][i3][line]
// (per band nlines n3 values), for the fused kernel; else to the box.
template <typename R, bool TL, int NN>
__global__ __launch_bounds__(256) void z_bwd_kernel(const cplx_t<R> *psi, long long ldpsi, cplx_t<R> *box,
                                                    const LocalTables *tabs, KStride ks, int nb, FftArgs a) {
  using V = cplx_t<R>;
  constexpr int K = Tile<R>::KMAX;
  const int n = NN > 0 ? NN : a.n;
  const int kp = blockIdx.z, band = blockIdx.y;
  const LocalTables t = tabs[kp];
  const int line0 = blockIdx.x * a.M;
  if (line0 >= t.nlines) return;  // past this k point's lines: no barrier reached
  psi += kp * ks.psi + band * ldpsi;
  box += kp * ks.box;
  const int mshift = __ffs(a.M) - 1;
  extern __shared__ double2 smem_raw[];
  V *buf0 = reinterpret_cast<V *>(smem_raw);
  V *buf1 = buf0 + a.M * a.ld;
  V *tws = buf1 + a.M * a.ld;
  load_twiddles(tws, a);
  const int MN = a.M * n;
  for (int e = threadIdx.x; e < a.M * a.ld; e += blockDim.x) buf0[e] = mk<V>(R(0), R(0));
  __syncthreads();
  // plane waves of each line, placed at their i3 (the rest of the line is 0)
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = threadIdx.x + k * 256;
    if (e < MN) {  // consecutive threads walk one line's plane waves (contiguous in psi)
      const int l = e / n, j = e - l * n;
      const int line = line0 + l;
      if (line < t.nlines && j < t.line_count[line]) {
        const int g = t.line_start[line] + j;
        buf0[l * a.ld + t.pw_i3[g]] = psi[g];
      }
    }
  }
  __syncthreads();
  V *res = stages<NN>(buf0, buf1, tws, a, +1);
  const long long lstride = static_cast<long long>(t.nlines) * t.n3;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = threadIdx.x + k * 256;
    if (e >= MN) break;
    if (TL) {  // consecutive threads take consecutive lines at one i3
      const int l = e & (a.M - 1), p = e >> mshift;
      const int line = line0 + l;
      if (line < t.nlines) box[band * lstride + static_cast<long long>(p) * t.nlines + line] = res[l * a.ld + p];
    } else {  // consecutive threads walk one line
      const int l = e / n, p = e - l * n;
      const int line = line0 + l;
      if (line < t.nlines)
        box[band * t.nr + static_cast<long long>(t.line_i12[line]) * t.n3 + p] = res[l * a.ld + p];
    }
  }
}

// Along i2 on the active planes; BWD loads active lines only and stores all,
// the forward pass loads all and stores active lines only.  Each thread keeps
// one line (M divides the 256 threads) and walks its points.
template <typename R, bool BWD, int NN>
__global__ __launch_bounds__(256) void y_kernel(cplx_t<R> *box, const LocalTables *tabs, KStride ks, int nb,
                                                FftArgs a) {
  using V = cplx_t<R>;
  constexpr int K = Tile<R>::KMAX;
  const int kp = blockIdx.y;
  const LocalTables t = tabs[kp];
  const long long lines_total = static_cast<long long>(nb) * t.nplanes * t.n3;
  if (static_cast<long long>(blockIdx.x) * a.M >= lines_total) return;
  box += kp * ks.box;
  extern __shared__ double2 smem_raw[];
  V *buf0 = reinterpret_cast<V *>(smem_raw);
  V *buf1 = buf0 + a.M * a.ld;
  V *tws = buf1 + a.M * a.ld;
  load_twiddles(tws, a);
  const int l = threadIdx.x % a.M, p0 = threadIdx.x / a.M, pstep = blockDim.x / a.M;
  const int L = blockIdx.x * a.M + l;
  const bool live = L < lines_total;
  const int per_band = t.nplanes * t.n3;
  const int band = live ? L / per_band : 0;
  const int rem = L - band * per_band;
  const int ip = rem / t.n3, i3 = rem - ip * t.n3;
  const int i1 = live ? t.plane_i1[ip] : 0;
  const long long base = band * t.nr + static_cast<long long>(i1) * t.n2 * t.n3 + i3;
  const int *lof = t.line_of + i1 * t.n2;
  V v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int p = p0 + k * pstep;
    v[k] = mk<V>(R(0), R(0));
    if (live && p < a.n && (!BWD || lof[p] >= 0)) v[k] = box[base + static_cast<long long>(p) * t.n3];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int p = p0 + k * pstep;
    if (p < a.n) buf0[l * a.ld + p] = v[k];
  }
  __syncthreads();
  V *res = stages<NN>(buf0, buf1, tws, a, BWD ? +1 : -1);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int p = p0 + k * pstep;
    if (live && p < a.n && (BWD || lof[p] >= 0)) box[base + static_cast<long long>(p) * t.n3] = res[l * a.ld + p];
  }
}

// Along i1 for every (i2, i3): backward, * Veff/N_r, forward.  A workgroup
// takes Mc consecutive (i2,i3) columns for G bands (M = G * Mc lines); the
// Veff tile of the Mc columns is read once into LDS and shared by the bands.
template <typename R>
__global__ __launch_bounds__(256) void x_veff_kernel(cplx_t<R> *box, const R *veff, R scale,
                                                     const LocalTables *tabs, KStride ks, FftArgs a, int G, int nb) {
  using V = cplx_t<R>;
  constexpr int K = Tile<R>::KMAX;
  const int kp = blockIdx.y;
  const LocalTables t = tabs[kp];
  box += kp * ks.box;
  extern __shared__ double2 smem_raw[];
  V *buf0 = reinterpret_cast<V *>(smem_raw);
  V *buf1 = buf0 + a.M * a.ld;
  V *tws = buf1 + a.M * a.ld;
  R *wbuf = reinterpret_cast<R *>(tws + a.n);
  load_twiddles(tws, a);
  const int C = a.C;  // n2 * n3
  const int Mc = a.M / G;
  const int ncb = (C + Mc - 1) / Mc;
  const int cblock = blockIdx.x % ncb, bgroup = blockIdx.x / ncb;
  for (int e = threadIdx.x; e < Mc * a.n; e += blockDim.x) {
    const int p = e / Mc, ci = e - p * Mc;
    const int c = cblock * Mc + ci;
    wbuf[ci * a.n + p] = c < C ? veff[c + static_cast<long long>(p) * C] * scale : R(0);
  }
  const int l = threadIdx.x % a.M, p0 = threadIdx.x / a.M, pstep = blockDim.x / a.M;
  const int gi = l / Mc, ci = l - gi * Mc;
  const int c = cblock * Mc + ci, band = bgroup * G + gi;
  const bool live = c < C && band < nb;
  const long long base = static_cast<long long>(live ? band : 0) * t.nr + (live ? c : 0);
  V v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int p = p0 + k * pstep;
    v[k] = mk<V>(R(0), R(0));
    if (live && p < a.n && t.plane_on[p]) v[k] = box[base + static_cast<long long>(p) * C];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int p = p0 + k * pstep;
    if (p < a.n) buf0[l * a.ld + p] = v[k];
  }
  __syncthreads();
  V *res = run_stages(buf0, buf1, tws, a, +1);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int p = p0 + k * pstep;
    if (p < a.n) res[l * a.ld + p] = cscale(res[l * a.ld + p], wbuf[ci * a.n + p]);
  }
  __syncthreads();
  res = run_stages(res, res == buf0 ? buf1 : buf0, tws, a, -1);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int p = p0 + k * pstep;
    if (live && p < a.n && t.plane_on[p]) box[base + static_cast<long long>(p) * C] = res[l * a.ld + p];
  }
}

// x pass as a convolution over the active planes (see the header), for a
// contiguous run of P planes s0 .. s0 + P - 1 (signed, modulo n1): with the
// planes in that order, plane p' takes sum_p Vx[(p' - p) mod n1] in[p].  A
// workgroup takes kXcCols consecutive (i2,i3) columns and kXcBands bands;
// lanes 0-31 and 32-63 of a wave hold two bands of the same columns.  LDS
// holds the 2 pw - 1 rows of Vx around the difference 0, ordered by
// p' - p + pw - 1 (pw = planes of the widest k point of the batch; narrower
// k points use the middle rows), so every term of the sum is one LDS read at
// a fixed offset; each thread keeps its column's P values in registers
// (PMAX >= P) and overwrites them.
constexpr int kXcCols = 32;
constexpr int kXcBands = 8;

template <typename R, int PMAX>
__global__ __launch_bounds__(256) void x_conv_kernel(cplx_t<R> *box, const cplx_t<R> *vx, const LocalTables *tabs,
                                                     KStride ks, int nb, int pw) {
  using V = cplx_t<R>;
  const int kp = blockIdx.z;
  const LocalTables t = tabs[kp];
  box += kp * ks.box;
  extern __shared__ double2 smem_raw[];
  V *vs = reinterpret_cast<V *>(smem_raw);  // [2 pw - 1][kXcCols]
  const int C = t.n2 * t.n3;
  const int P = t.nplanes;
  const int c0 = blockIdx.x * kXcCols;
  for (int e = threadIdx.x; e < (2 * pw - 1) * kXcCols; e += blockDim.x) {
    const int k = e / kXcCols, c = c0 + (e - k * kXcCols);
    int d = (k - (pw - 1)) % t.n1;
    d += d < 0 ? t.n1 : 0;
    vs[e] = c < C ? vx[static_cast<long long>(d) * C + c] : mk<V>(R(0), R(0));
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ci = lane & 31;
  const int c = c0 + ci;
  const int band = blockIdx.y * kXcBands + wave * 2 + (lane >> 5);
  if (c >= C || band >= nb) return;  // no barrier follows
  V *col = box + static_cast<long long>(band) * t.nr + c;
  const int s0 = t.plane_s0 < 0 ? t.plane_s0 + t.n1 : t.plane_s0;
  V in[PMAX];
#pragma unroll
  for (int q = 0; q < PMAX; ++q) {
    int i1 = s0 + q;
    i1 -= i1 >= t.n1 ? t.n1 : 0;
    in[q] = q < P ? col[static_cast<long long>(i1) * C] : mk<V>(R(0), R(0));
  }
  for (int op = 0; op < P; ++op) {
    const V *w = vs + (op + pw - 1) * kXcCols + ci;  // w[-q kXcCols] = Vx[(op - q) mod n1]
    V acc = mk<V>(R(0), R(0));
#pragma unroll
    for (int q = 0; q < PMAX; ++q)
      if (q < P) cmac(acc, w[-q * kXcCols], in[q]);
    int i1 = s0 + op;
    i1 -= i1 >= t.n1 ? t.n1 : 0;
    col[static_cast<long long>(i1) * C] = acc;
  }
}

// The same convolution on the matrix cores.  For one (i2, i3) column c the x
// pass over the P active planes and a group of bands is the product
//   Out[p'][b] = sum_p T_c[p'][p] In[p][b],  T_c[p'][p] = Vx[(p' - p) mod n1][c],
// a P x P Toeplitz matrix times a P x 16 block of 16 bands.  A workgroup
// takes kXmCols consecutive columns and one 16-band tile; each wave owns two
// columns and accumulates their complex 16x16 output tiles (ceil(P / 16) per
// column) over ceil(P / 4) steps of 4 planes.  Per step the workgroup stages
// the In block [4 planes][kXmCols columns][16 bands] in LDS (loads along the
// columns: each plane row of the tile is one contiguous run), double-
// buffered so the next step's loads are in flight during this step's MFMAs;
// the B operand comes from there, the A operand (T_c) from the column's
// 2P - 1 Vx values, also in LDS.  The results go back through LDS, one
// 16-plane output tile at a time, and are stored along the columns over the
// input planes (every step has been read before the first store).  Column
// groups are numbered XCD-major: workgroup x runs on XCD x mod 8 (round-robin
// dispatch), so groups that share 128-byte lines (FP32: two per line) meet in
// one L2 (gridDim.x is a multiple of 8; groups past the end are empty).
constexpr int kXmCols = 8;

// WAVES = 4: two columns per wave, the default complex-product form of R,
// at least three waves per SIMD (the 48-plane form needs up to 168
// registers; left to itself the compiler took more and the kernel ran 10%
// slower); WAVES = 8: one column per wave, three-MFMA (Gauss) products, whose
// extra accumulators fit because each wave holds half the tiles.
template <typename R, int PMAX, int WAVES>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(WAVES == 4 ? 3 : 2))) void x_mfma_kernel(
    cplx_t<R> *box, const cplx_t<R> *vx, const LocalTables *tabs, KStride ks, int nb, int groups_per_xcd) {
  using V = cplx_t<R>;
  using A4 = typename Acc<R>::type;
  constexpr int NT = 64 * WAVES;        // threads
  constexpr int CPW = kXmCols / WAVES;  // columns per wave
  constexpr int PARTS = WAVES == 8 ? 3 : kCmacParts<R>;
  constexpr int JS = 4 * kXmCols * 16 / NT;  // staged elements per thread and step
  constexpr int MT = (PMAX + 15) / 16;  // 16-plane output tiles
  constexpr int SB = 17;                // band row length of a staged step (16 + 1 pad)
  constexpr int OC = kXmCols + 1;       // column row length of an output tile (8 + 1 pad)
  constexpr int STAGE = 4 * kXmCols * SB;
  constexpr int UNION = (2 * STAGE > 16 * 16 * OC) ? 2 * STAGE : 16 * 16 * OC;
  __shared__ V tw[kXmCols][2 * PMAX - 1];
  __shared__ V buf[UNION];  // two staged steps, then one output tile
  const int kp = blockIdx.z;
  const LocalTables t = tabs[kp];
  const int C = t.n2 * t.n3;
  const int P = t.nplanes;
  if (P > PMAX) return;  // the same for the whole workgroup; the host never launches it
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c0 = ((blockIdx.x & 7) * groups_per_xcd + (blockIdx.x >> 3)) * kXmCols;
  const int b0 = blockIdx.y * 16;
  box += kp * ks.box;
  const int s0 = t.plane_s0 < 0 ? t.plane_s0 + t.n1 : t.plane_s0;
  // tw[w][d + P - 1] = Vx[d mod n1][c0 + w] for d in [-(P - 1), P - 1]
  for (int e = tid; e < kXmCols * (2 * P - 1); e += blockDim.x) {
    const int w = e % kXmCols, k = e / kXmCols;
    const int c = c0 + w;
    int d = (k - (P - 1)) % t.n1;
    d += d < 0 ? t.n1 : 0;
    tw[w][k] = c < C ? vx[static_cast<long long>(d) * C + c] : mk<V>(R(0), R(0));
  }
  // staging: element e = tid + NT j of a step is column e & 7, band
  // (e >> 3) & 15, plane (e >> 7) of the step
  auto fetch = [&](int k4, V (&r)[JS]) {
#pragma unroll
    for (int j = 0; j < JS; ++j) {
      const int e = tid + NT * j;
      const int cc = e & 7, band = (e >> 3) & 15, q = 4 * k4 + (e >> 7);
      int i1 = s0 + q;
      i1 -= i1 >= t.n1 ? t.n1 : 0;
      const bool ok = q < P && b0 + band < nb && c0 + cc < C;
      r[j] = ok ? box[static_cast<long long>(b0 + band) * t.nr + static_cast<long long>(i1) * C + c0 + cc]
                : mk<V>(R(0), R(0));
    }
  };
  auto stash = [&](int st, const V (&r)[JS]) {
#pragma unroll
    for (int j = 0; j < JS; ++j) {
      const int e = tid + NT * j;
      buf[st * STAGE + ((e >> 7) * kXmCols + (e & 7)) * SB + ((e >> 3) & 15)] = r[j];
    }
  };
  const int nk4 = (P + 3) / 4;
  const int row = lane & 15, kq = lane >> 4;
  A4 acc[CPW][MT][PARTS];
#pragma unroll
  for (int j = 0; j < CPW; ++j)
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int u = 0; u < PARTS; ++u) acc[j][mt][u] = A4{0, 0, 0, 0};
  V r[JS];
  fetch(0, r);
  stash(0, r);
  for (int k4 = 0; k4 < nk4; ++k4) {
    __syncthreads();  // step k4 staged; every reader of the other buffer is done
    if (k4 + 1 < nk4) fetch(k4 + 1, r);
    const V *stg = buf + (k4 & 1) * STAGE;
    const int q = 4 * k4 + kq;
#pragma unroll
    for (int j = 0; j < CPW; ++j) {
      const int cc = CPW * wave + j;
      const V bv = stg[(kq * kXmCols + cc) * SB + row];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        if (16 * mt >= P) continue;  // uniform
        const int op = 16 * mt + row;
        const bool live = op < P && q < P;
        const V a = tw[cc][live ? op - q + P - 1 : 0];
        half::cmac<R>(live ? a.x : R(0), live ? a.y : R(0), bv.x, bv.y, acc[j][mt]);
      }
    }
    if (k4 + 1 < nk4) stash((k4 + 1) & 1, r);
  }
  // output tiles through LDS: buf[(op' * 16 + band) * OC + column]
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    if (16 * mt >= P) continue;  // uniform
    __syncthreads();  // the staged steps (or the previous tile) are no longer read
#pragma unroll
    for (int j = 0; j < CPW; ++j)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        R re, im;
        cmac_result<R>(acc[j][mt], rr, re, im);
        buf[(mfma_row<R>(lane, rr) * 16 + row) * OC + CPW * wave + j] = mk<V>(re, im);
      }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 16 * 16 * kXmCols / NT; ++it) {
      const int e = tid + NT * it;
      const int cc = e & 7, band = (e >> 3) & 15, opl = e >> 7;
      const int op = 16 * mt + opl;
      int i1 = s0 + op;
      i1 -= i1 >= t.n1 ? t.n1 : 0;
      if (op < P && b0 + band < nb && c0 + cc < C)
        box[static_cast<long long>(b0 + band) * t.nr + static_cast<long long>(i1) * C + c0 + cc] =
            buf[(opl * 16 + band) * OC + cc];
    }
  }
}

// Fused y, x and y passes on one i3 slice of G bands, reading and writing
// the i3-major line buffer of z_bwd_kernel<R, true> in place:
//   1. the slice's active lines go to LDS planes [g][q][i2] (q = signed
//      plane - first signed plane; zeros elsewhere);
//   2. backward transform along i2 for every (g, q);
//   3. the x pass per (g, i2) column as the circular convolution
//        out[q'] = sum_q Vx[(q' - q) mod n1][(i2, i3)] in[q],
//      an item taking kYxyOut consecutive outputs with its window of
//      PMAX + kYxyOut - 1 Vx values (read from the [i3][d][i2] copy, so the
//      loads are contiguous along i2) in registers and the inputs from LDS;
//   4. forward transform along i2;
//   5. the slice's active lines stored back.
// All band groups of one i3 fall on the same XCD (the grid is i3-fastest and
// n3 is a multiple of the XCD count on the boxes used), so each XCD's L2
// holds the Vx rows of its own i3 values.
constexpr int kYxyOut = 4;

template <typename R, int PMAX, int NN>
__global__ __launch_bounds__(256) void yxy_kernel(cplx_t<R> *lines, const cplx_t<R> *vxT, const LocalTables *tabs,
                                                  KStride ks, int nb, int G, FftArgs a) {
  using V = cplx_t<R>;
  const int kp = blockIdx.y;
  const LocalTables t = tabs[kp];
  const int P = t.nplanes;
  const int n2 = NN > 0 ? NN : t.n2;
  const int i3 = blockIdx.x % t.n3;
  const int band0 = (blockIdx.x / t.n3) * G;
  if (band0 >= nb || P == 0) return;  // uniform per workgroup: no barrier reached
  lines += kp * ks.box;
  FftArgs ay = a;
  ay.M = G * P;  // y lines of the slice
  const int ld = ay.ld;
  extern __shared__ double2 smem_raw[];
  V *A = reinterpret_cast<V *>(smem_raw);
  V *B = A + ay.M * ld;
  V *tws = B + ay.M * ld;
  load_twiddles(tws, ay);
  for (int e = threadIdx.x; e < ay.M * ld; e += blockDim.x) A[e] = mk<V>(R(0), R(0));
  __syncthreads();
  const long long lstride = static_cast<long long>(t.nlines) * t.n3;
  const int half1 = (t.n1 - 1) / 2;
  // line l of plane q at i2 sits at LDS offset (q ld + i2) of each band
  auto slot = [&](int l) {
    const int i12 = t.line_i12[l];
    const int i1 = i12 / n2, i2 = i12 - i1 * n2;
    return ((i1 <= half1 ? i1 : i1 - t.n1) - t.plane_s0) * ld + i2;
  };
  for (int g = 0; g < G && band0 + g < nb; ++g) {
    const V *src = lines + (band0 + g) * lstride + static_cast<long long>(i3) * t.nlines;
    for (int l = threadIdx.x; l < t.nlines; l += blockDim.x) A[g * P * ld + slot(l)] = src[l];
  }
  __syncthreads();
  V *r1 = stages<NN>(A, B, tws, ay, +1);
  V *o1 = r1 == A ? B : A;
  const V *vcol = vxT + static_cast<long long>(i3) * t.n1 * n2;  // [d][i2] of this i3
  const int nparts = (P + kYxyOut - 1) / kYxyOut;
  for (int it = threadIdx.x; it < nparts * G * n2; it += blockDim.x) {
    const int i2 = it % n2, rest = it / n2;
    const int g = rest % G, qa = (rest / G) * kYxyOut;
    // w[j] = Vx[(qa - (PMAX - 1) + j) mod n1]: output qa + o, input q uses
    // w[o + PMAX - 1 - q] (a compile-time index)
    // (-n1 <= qa - (PMAX - 1) + j < 2 n1: PMAX <= n1 + 1 is checked at launch)
    V w[PMAX + kYxyOut - 1];
#pragma unroll
    for (int j = 0; j < PMAX + kYxyOut - 1; ++j) {
      int d = qa - (PMAX - 1) + j;
      d += d < 0 ? t.n1 : 0;
      d -= d >= t.n1 ? t.n1 : 0;
      w[j] = vcol[d * n2 + i2];
    }
    V acc[kYxyOut];
#pragma unroll
    for (int o = 0; o < kYxyOut; ++o) acc[o] = mk<V>(R(0), R(0));
    const V *col = r1 + g * P * ld + i2;
#pragma unroll
    for (int q = 0; q < PMAX; ++q) {
      if (q < P) {
        const V v = *col;
#pragma unroll
        for (int o = 0; o < kYxyOut; ++o) cmac(acc[o], w[o + PMAX - 1 - q], v);
      }
      col += ld;
    }
#pragma unroll
    for (int o = 0; o < kYxyOut; ++o)
      if (qa + o < P) o1[(g * P + qa + o) * ld + i2] = acc[o];
  }
  __syncthreads();
  V *r2 = stages<NN>(o1, r1, tws, ay, -1);
  for (int g = 0; g < G && band0 + g < nb; ++g) {
    V *dst = lines + (band0 + g) * lstride + static_cast<long long>(i3) * t.nlines;
    for (int l = threadIdx.x; l < t.nlines; l += blockDim.x) dst[l] = r2[g * P * ld + slot(l)];
  }
}

// yxy_kernel with the slice's Vx rows in LDS, for eight waves per
// workgroup.  The convolution needs Vx[d][i2] for d = -(P-1) .. P-1 only,
// 2P - 1 rows: every thread loads its share into registers at the start (the
// latency runs alongside the line loads and the forward y transform) and
// writes it to the second plane buffer once that transform has freed it.  The convolution keeps a sliding window of kYxyOut values (one
// LDS load per input, instead of PMAX + kYxyOut - 1 values held per item),
// so the kernel fits 128 registers and two workgroups of eight waves share
// a CU (the plane buffers of two bands, 60 KB in FP64 at 48^3, allow two).
// It writes in place: all parts of one column are lanes of one wave in one
// round, part-major so a wave's reads stay on consecutive i2, and a wave's
// LDS reads of a round precede its writes.  Needs NN > 0 and G >= 2
// ((2P - 1) n2 values fit in the G P ld of a plane buffer).
constexpr int kYxyLdsThreads = 512;

template <typename R, int PMAX, int NN>
__global__ __attribute__((amdgpu_flat_work_group_size(kYxyLdsThreads, kYxyLdsThreads), amdgpu_waves_per_eu(4)))
void yxy_lds_kernel(cplx_t<R> *lines, const cplx_t<R> *vxT, const LocalTables *tabs, KStride ks, int nb, int G,
                    FftArgs a) {
  using V = cplx_t<R>;
  static_assert(NN > 0, "compile-time y length");
  constexpr int kVxPer = ((2 * PMAX - 1) * NN + kYxyLdsThreads - 1) / kYxyLdsThreads;
  constexpr int kWaves = kYxyLdsThreads / 64;
  const int kp = blockIdx.y;
  const LocalTables t = tabs[kp];
  const int P = t.nplanes;
  const int i3 = blockIdx.x % t.n3;
  const int band0 = (blockIdx.x / t.n3) * G;
  if (band0 >= nb || P == 0) return;  // uniform per workgroup: no barrier reached
  // row r of the staged block is Vx[(r - (P - 1)) mod n1][i2] of this i3; the
  // loads are issued here so their latency overlaps the line loads and the
  // forward transform, and the values go to LDS once that transform is done
  const V *vcol = vxT + static_cast<long long>(i3) * t.n1 * NN;
  V vreg[kVxPer];
#pragma unroll
  for (int k = 0; k < kVxPer; ++k) {
    const int e = threadIdx.x + k * kYxyLdsThreads;
    vreg[k] = mk<V>(R(0), R(0));
    if (e < (2 * P - 1) * NN) {
      const int r = e / NN, i2 = e - r * NN;
      const int d = r - (P - 1);
      vreg[k] = vcol[(d < 0 ? d + t.n1 : d) * NN + i2];
    }
  }
  lines += kp * ks.box;
  FftArgs ay = a;
  ay.M = G * P;
  constexpr int ld = NN + 1;  // the y axis' LDS line stride (checked at launch)
  extern __shared__ double2 smem_raw[];
  V *A = reinterpret_cast<V *>(smem_raw);
  V *B = A + ay.M * ld;
  V *tws = B + ay.M * ld;
  load_twiddles(tws, ay);
  for (int e = threadIdx.x; e < ay.M * ld; e += blockDim.x) A[e] = mk<V>(R(0), R(0));
  __syncthreads();
  const long long lstride = static_cast<long long>(t.nlines) * t.n3;
  const int half1 = (t.n1 - 1) / 2;
  auto slot = [&](int l) {
    const int i12 = t.line_i12[l];
    const int i1 = i12 / NN, i2 = i12 - i1 * NN;
    return ((i1 <= half1 ? i1 : i1 - t.n1) - t.plane_s0) * ld + i2;
  };
  for (int g = 0; g < G && band0 + g < nb; ++g) {
    const V *src = lines + (band0 + g) * lstride + static_cast<long long>(i3) * t.nlines;
    for (int l = threadIdx.x; l < t.nlines; l += blockDim.x) A[g * P * ld + slot(l)] = src[l];
  }
  __syncthreads();
  V *r1 = stages<NN>(A, B, tws, ay, +1);  // ends with a barrier
  V *o1 = r1 == A ? B : A;
#pragma unroll
  for (int k = 0; k < kVxPer; ++k) {
    const int e = threadIdx.x + k * kYxyLdsThreads;
    if (e < (2 * P - 1) * NN) o1[e] = vreg[k];
  }
  __syncthreads();
  const int nparts = (P + kYxyOut - 1) / kYxyOut;
  const int cpw = 64 / nparts;  // columns per wave and round
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ncol = G * NN;
  if (lane < cpw * nparts) {
    const int part = lane / cpw, cl = lane - part * cpw;
    const int qa = part * kYxyOut;
    // the round loop depends on the wave only
    for (int c0 = wave * cpw; c0 < ncol; c0 += kWaves * cpw) {
      const int col = c0 + cl;
      if (col >= ncol) continue;
      const int g = col / NN, i2 = col - g * NN;
      // Output qa + o takes input q with row qa + o - q + P - 1, in
      // [0, 2P - 2] for every kept pair (qa + o < P, q < P).  wv[o] holds
      // that row for the current q; the next q needs the previous entry
      // (wv[o - 1]) and one new row, qa + P - 2 - q.  Row offsets are
      // clamped into [0, 2P - 2] (a median of three); the clamp only
      // touches values no kept output uses.
      // The loads and the shift run for every q < PMAX (inputs past P
      // read row 0 of the column and are not used), so the window stays in
      // the same registers with no branch joins.
      const int lo = i2, hi = (2 * P - 2) * NN + i2;
      const int top = (qa + P - 2) * NN + i2;  // offset of row qa + P - 2
      auto clamp_off = [&](int off) { return off < lo ? lo : (off > hi ? hi : off); };
      V wv[kYxyOut];
#pragma unroll
      for (int o = 0; o < kYxyOut; ++o) wv[o] = o1[clamp_off(top + (o + 1) * NN)];
      V acc[kYxyOut];
#pragma unroll
      for (int o = 0; o < kYxyOut; ++o) acc[o] = mk<V>(R(0), R(0));
      V *col_p = r1 + g * P * ld + i2;
      const V *pin = col_p;
#pragma unroll
      for (int q = 0; q < PMAX; ++q) {
        const V v = *(q < P ? pin : col_p);
        if (q < P) {
#pragma unroll
          for (int o = 0; o < kYxyOut; ++o) cmac(acc[o], wv[o], v);
        }
#pragma unroll
        for (int o = kYxyOut - 1; o > 0; --o) wv[o] = wv[o - 1];
        wv[0] = o1[clamp_off(top - q * NN)];
        pin += ld;
      }
      V *pout = col_p + qa * ld;
#pragma unroll
      for (int o = 0; o < kYxyOut; ++o) {
        if (qa + o < P) *pout = acc[o];
        pout += ld;
      }
    }
  }
  __syncthreads();
  V *r2 = stages<NN>(r1, o1, tws, ay, -1);
  for (int g = 0; g < G && band0 + g < nb; ++g) {
    V *dst = lines + (band0 + g) * lstride + static_cast<long long>(i3) * t.nlines;
    for (int l = threadIdx.x; l < t.nlines; l += blockDim.x) dst[l] = r2[g * P * ld + slot(l)];
  }
}

// Vx[d][c] = scale sum_x veff[x][c] e^{-2 pi i d x / n1} for the C = n2 n3
// columns of the C-order box, in FP64 and FP32
__global__ void veff_x_dft_kernel(const double *veff, double2 *vx, float2 *vx32, int n1, int n2, int n3,
                                  double scale) {
  const int C = n2 * n3;
  const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<long long>(n1) * C) return;
  const int d = static_cast<int>(t / C);
  const int c = static_cast<int>(t - static_cast<long long>(d) * C);
  double sr = 0, si = 0;
  for (int x = 0; x < n1; ++x) {
    double sn, cs;
    sincospi(-2.0 * ((d * x) % n1) / n1, &sn, &cs);
    const double v = veff[static_cast<long long>(x) * C + c];
    sr += v * cs;
    si += v * sn;
  }
  vx[t] = make_double2(sr * scale, si * scale);
  vx32[t] = make_float2(static_cast<float>(sr * scale), static_cast<float>(si * scale));
  // the same value in the [i3][d][i2] layout after the first N_r entries
  const int i2 = c / n3, i3 = c - i2 * n3;
  const long long u = static_cast<long long>(n1) * C + (static_cast<long long>(i3) * n1 + d) * n2 + i2;
  vx[u] = make_double2(sr * scale, si * scale);
  vx32[u] = make_float2(static_cast<float>(sr * scale), static_cast<float>(si * scale));
}

// addh / adds (may be null): added to H psi / S psi (band b of k point j at
// + j addk + b addld).
template <typename R, bool TL, int NN>
__global__ __launch_bounds__(256) void z_fwd_kernel(const cplx_t<R> *psi, long long ldpsi, const cplx_t<R> *box,
                                                    const R *kin, cplx_t<R> *hpsi, cplx_t<R> *spsi,
                                                    long long ldout, const LocalTables *tabs, KStride ks, int nb,
                                                    FftArgs a, const cplx_t<R> *addh, const cplx_t<R> *adds,
                                                    long long addld, long long addk) {
  using V = cplx_t<R>;
  constexpr int K = Tile<R>::KMAX;
  const int n = NN > 0 ? NN : a.n;
  const int kp = blockIdx.z, band = blockIdx.y;
  const LocalTables t = tabs[kp];
  const int line0 = blockIdx.x * a.M;
  if (line0 >= t.nlines) return;
  psi += kp * ks.psi + band * ldpsi;
  box += kp * ks.box;
  kin += kp * ks.kin;
  hpsi += kp * ks.out + band * ldout;
  spsi += kp * ks.out + band * ldout;
  if (addh) {
    addh += kp * addk + band * addld;
    adds += kp * addk + band * addld;
  }
  const int mshift = __ffs(a.M) - 1;
  extern __shared__ double2 smem_raw[];
  V *buf0 = reinterpret_cast<V *>(smem_raw);
  V *buf1 = buf0 + a.M * a.ld;
  V *tws = buf1 + a.M * a.ld;
  load_twiddles(tws, a);
  const int MN = a.M * n;
  const long long lstride = static_cast<long long>(t.nlines) * t.n3;
  V v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = threadIdx.x + k * 256;
    v[k] = mk<V>(R(0), R(0));
    if (e < MN) {
      if (TL) {
        const int l = e & (a.M - 1), p = e >> mshift;
        const int line = line0 + l;
        if (line < t.nlines) v[k] = box[band * lstride + static_cast<long long>(p) * t.nlines + line];
      } else {
        const int l = e / n, p = e - l * n;
        const int line = line0 + l;
        if (line < t.nlines) v[k] = box[band * t.nr + static_cast<long long>(t.line_i12[line]) * t.n3 + p];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = threadIdx.x + k * 256;
    if (e < MN) {
      if (TL)
        buf0[(e & (a.M - 1)) * a.ld + (e >> mshift)] = v[k];
      else
        buf0[(e / n) * a.ld + (e - (e / n) * n)] = v[k];
    }
  }
  __syncthreads();
  V *res = stages<NN>(buf0, buf1, tws, a, -1);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = threadIdx.x + k * 256;
    if (e >= MN) break;
    const int l = e / n, j = e - l * n;  // consecutive threads walk one line's plane waves
    const int line = line0 + l;
    if (line < t.nlines && j < t.line_count[line]) {
      const int g = t.line_start[line] + j;
      const V p = psi[g];
      V r = res[l * a.ld + t.pw_i3[g]];
      V q = p;
      if (addh) {
        r = cadd(r, addh[g]);
        q = cadd(q, adds[g]);
      }
      hpsi[g] = mk<V>(r.x + kin[g] * p.x, r.y + kin[g] * p.y);
      spsi[g] = q;
    }
  }
}

template <typename R>
FftArgs axis_args(const FftPlan &plan, int axis, long long lines, int C, int group = 1) {
  const FftAxis &ax = plan.axis[axis];
  FftArgs a{};
  a.n = ax.n;
  a.nstages = ax.nstages;
  for (int i = 0; i < 8; ++i) a.radix[i] = ax.radix[i];
  a.tw = ax.twiddle.get();
  a.lines_total = lines;
  a.C = C;
  // lines per workgroup: the largest power of two with M * n <= budget
  // (points per tile; at most 256 threads x KMAX elements)
  const int prec = std::is_same<R, double>::value ? 0 : 1;
  const int budget = std::min(256 * Tile<R>::KMAX, g_tune.tile[prec][axis]);
  int M = 1;
  while (2 * M * ax.n <= budget && M < 64) M *= 2;
  M = std::max(M, group);  // at least one column per band of the group
  if (M * ax.n > 256 * Tile<R>::KMAX) throw Error(kUnavailable, "FFT line too long for the local-stage tile");
  a.M = M;
  a.ld = ax.n + 1;
  return a;
}

template <typename R>
size_t lds_bytes(const FftArgs &a) {
  return (2 * size_t(a.M) * a.ld + a.n) * sizeof(cplx_t<R>);
}

template <typename R>
void launch(const char *name, const void *fn, const FftArgs &a, hipStream_t s, void **args, dim3 grid,
            size_t extra_lds = 0) {
  trace_mark(name, s);
  HALF_HIP(hipLaunchKernel(fn, grid, dim3(256), args, lds_bytes<R>(a) + extra_lds, s));
}

// register budget of x_conv_kernel per precision: the largest PMAX instantiated
template <typename R>
const void *x_conv_for(int planes, int &pmax) {
  if constexpr (std::is_same<R, double>::value) {
    if (planes <= 16) return pmax = 16, reinterpret_cast<const void *>(&x_conv_kernel<R, 16>);
    if (planes <= 24) return pmax = 24, reinterpret_cast<const void *>(&x_conv_kernel<R, 24>);
    if (planes <= 32) return pmax = 32, reinterpret_cast<const void *>(&x_conv_kernel<R, 32>);
    if (planes <= 48) return pmax = 48, reinterpret_cast<const void *>(&x_conv_kernel<R, 48>);
  } else {
    if (planes <= 16) return pmax = 16, reinterpret_cast<const void *>(&x_conv_kernel<R, 16>);
    if (planes <= 32) return pmax = 32, reinterpret_cast<const void *>(&x_conv_kernel<R, 32>);
    if (planes <= 48) return pmax = 48, reinterpret_cast<const void *>(&x_conv_kernel<R, 48>);
    if (planes <= 64) return pmax = 64, reinterpret_cast<const void *>(&x_conv_kernel<R, 64>);
  }
  pmax = 0;
  return nullptr;
}

template <typename R, int WAVES>
const void *x_mfma_for_w(int planes) {
  if (planes <= 16) return reinterpret_cast<const void *>(&x_mfma_kernel<R, 16, WAVES>);
  if (planes <= 32) return reinterpret_cast<const void *>(&x_mfma_kernel<R, 32, WAVES>);
  if (planes <= 48) return reinterpret_cast<const void *>(&x_mfma_kernel<R, 48, WAVES>);
  return nullptr;
}

// xconv 2: four waves per workgroup; 3: eight, with three-MFMA products
template <typename R>
const void *x_mfma_for(int planes, int form, int &threads) {
  threads = form == 3 ? 512 : 256;
  return form == 3 ? x_mfma_for_w<R, 8>(planes) : x_mfma_for_w<R, 4>(planes);
}

// Transform lengths with compile-time kernels (the boxes of the workloads
// in use); any other length runs the runtime-plan kernels.
#define HALF_CT_LENGTHS(X) X(24) X(40) X(48) X(96)

template <typename R, bool TL>
const void *z_bwd_for(int n) {
#define HALF_CASE(N) \
  if (n == N) return reinterpret_cast<const void *>(&z_bwd_kernel<R, TL, N>);
  HALF_CT_LENGTHS(HALF_CASE)
#undef HALF_CASE
  return reinterpret_cast<const void *>(&z_bwd_kernel<R, TL, 0>);
}

template <typename R, bool TL>
const void *z_fwd_for(int n) {
#define HALF_CASE(N) \
  if (n == N) return reinterpret_cast<const void *>(&z_fwd_kernel<R, TL, N>);
  HALF_CT_LENGTHS(HALF_CASE)
#undef HALF_CASE
  return reinterpret_cast<const void *>(&z_fwd_kernel<R, TL, 0>);
}

template <typename R, bool BWD>
const void *y_for(int n) {
#define HALF_CASE(N) \
  if (n == N) return reinterpret_cast<const void *>(&y_kernel<R, BWD, N>);
  HALF_CT_LENGTHS(HALF_CASE)
#undef HALF_CASE
  return reinterpret_cast<const void *>(&y_kernel<R, BWD, 0>);
}

// A fused kernel and its workgroup size.
struct YxyKernel {
  const void *fn = nullptr;
  int threads = 256;
};

// Vx rows staged in LDS (yxy_lds_kernel): compile-time y length, a band
// group of two or more, and at most 16 staged values per thread
template <typename R, int PMAX>
YxyKernel yxy_len(int n2, bool vx_lds) {
#define HALF_CASE(N)                                                                             \
  if (n2 == N) {                                                                                 \
    if (vx_lds && (2 * PMAX - 1) * N <= 16 * kYxyLdsThreads)                                     \
      return {reinterpret_cast<const void *>(&yxy_lds_kernel<R, PMAX, N>), kYxyLdsThreads};      \
    return {reinterpret_cast<const void *>(&yxy_kernel<R, PMAX, N>), 256};                       \
  }
  HALF_CT_LENGTHS(HALF_CASE)
#undef HALF_CASE
  return {reinterpret_cast<const void *>(&yxy_kernel<R, PMAX, 0>), 256};
}

// register budget of yxy_kernel per precision (window of PMAX + 3 values)
template <typename R>
YxyKernel yxy_for(int planes, int n1, int n2, int G, int ld) {
  int pmax = planes <= 16 ? 16 : planes <= 24 ? 24 : planes <= 32 ? 32 : 48;
  // above 32 planes the separate passes are faster in either precision
  // (96^3 box, 41 planes, FP32: 2.34 ms fused against 1.92 ms per 303 bands)
  if (pmax > 32) return {};
  if (planes > 48 || pmax > n1 + 1) return {};  // window arithmetic assumes PMAX <= n1 + 1
  const bool vx_lds = g_tune.vxlds[std::is_same<R, double>::value ? 0 : 1] && G >= 2 && ld == n2 + 1;
  switch (pmax) {
    case 16: return yxy_len<R, 16>(n2, vx_lds);
    case 24: return yxy_len<R, 24>(n2, vx_lds);
    case 32: return yxy_len<R, 32>(n2, vx_lds);
    default: return yxy_len<R, 48>(n2, vx_lds);
  }
}

}  // namespace

void LocalMap::build(const Basis &b, hipStream_t s) {
  n1 = b.shape[0];
  n2 = b.shape[1];
  n3 = b.shape[2];
  const int n = static_cast<int>(b.npw());
  std::vector<int> i12(n), i3(n), line_of_h(size_t(n1) * n2, -1), plane_on_h(n1, 0);
  std::vector<int> starts, counts, lines;
  for (int g = 0; g < n; ++g) {
    i12[g] = b.fft_index[g] / n3;
    i3[g] = b.fft_index[g] % n3;
    if (g == 0 || i12[g] != i12[g - 1]) {
      if (line_of_h[i12[g]] >= 0)
        throw Error(kInternal, "plane waves of one FFT line are not contiguous in the basis");
      line_of_h[i12[g]] = static_cast<int>(lines.size());
      lines.push_back(i12[g]);
      starts.push_back(g);
      counts.push_back(0);
    }
    ++counts.back();
    plane_on_h[i12[g] / n2] = 1;
  }
  std::vector<int> planes;
  for (int i = 0; i < n1; ++i)
    if (plane_on_h[i]) planes.push_back(i);
  // signed plane indices in [-n1/2, (n1-1)/2], the basis's enumeration
  // range: one contiguous run for a sphere
  std::vector<int> signed_planes;
  for (int i : planes) signed_planes.push_back(i <= (n1 - 1) / 2 ? i : i - n1);
  std::sort(signed_planes.begin(), signed_planes.end());
  planes_contiguous = !signed_planes.empty();
  for (size_t q = 1; q < signed_planes.size(); ++q)
    if (signed_planes[q] != signed_planes[0] + static_cast<int>(q)) planes_contiguous = false;
  plane_s0 = signed_planes.empty() ? 0 : signed_planes[0];
  nlines = static_cast<int>(lines.size());
  nplanes = static_cast<int>(planes.size());
  line_i12.upload(lines, s);
  line_start.upload(starts, s);
  line_count.upload(counts, s);
  line_of.upload(line_of_h, s);
  plane_i1.upload(planes, s);
  plane_on.upload(plane_on_h, s);
  pw_i3.upload(i3, s);
  const LocalTables t = tables();
  dtab.upload(&t, 1, s);
  HALF_HIP(hipStreamSynchronize(s));
}

LocalTables LocalMap::tables() const {
  LocalTables t{};
  t.n1 = n1;
  t.n2 = n2;
  t.n3 = n3;
  t.nr = static_cast<long long>(n1) * n2 * n3;
  t.nlines = nlines;
  t.line_i12 = line_i12.get();
  t.line_start = line_start.get();
  t.line_count = line_count.get();
  t.line_of = line_of.get();
  t.nplanes = nplanes;
  t.plane_i1 = plane_i1.get();
  t.plane_on = plane_on.get();
  t.pw_i3 = pw_i3.get();
  t.plane_s0 = plane_s0;
  return t;
}

void build_veff_x_dft(const double *veff, int n1, int n2, int n3, double2 *vx, float2 *vx32, hipStream_t s) {
  const long long total = static_cast<long long>(n1) * n2 * n3;
  const double scale = 1.0 / static_cast<double>(total);
  hipLaunchKernelGGL(veff_x_dft_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s, veff, vx,
                     vx32, n1, n2, n3, scale);
  HALF_HIP(hipGetLastError());
}

template <typename R>
void local_apply_batch(const FftPlan &plan, const LocalBatch &kb, const R *veff, const cplx_t<R> *vx, const R *kin,
                       int nb, const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi, cplx_t<R> *spsi,
                       int64_t ldout, cplx_t<R> *boxes, hipStream_t s, const LocalAddend<R> &add) {
  using V = cplx_t<R>;
  const int n1 = plan.n[0], n2 = plan.n[1], n3 = plan.n[2];
  if (nb <= 0 || kb.nk <= 0 || kb.max_lines == 0) return;
  if (kb.nk > 65535) throw Error(kInvalidArgument, "too many k points in one batch");
  if (static_cast<long long>(nb) * (static_cast<long long>(n1) * n2 + static_cast<long long>(n2) * n3 +
                                   static_cast<long long>(n1) * n3) >= (1ll << 31))
    throw Error(kInvalidArgument, "band batch too large for the local stage");
  long long lp = ldpsi, lo = ldout;
  KStride ks{kb.psi_k, kb.out_k, kb.kin_k, kb.box_k};
  const LocalTables *tabs = kb.tabs;
  const int prec = std::is_same<R, double>::value ? 0 : 1;
  const int C = n2 * n3;
  const unsigned nk = static_cast<unsigned>(kb.nk);
  FftArgs az = axis_args<R>(plan, 2, 0, 1);
  FftArgs ay = axis_args<R>(plan, 1, 0, n3);
  auto blocks = [](long long lines, int M) { return static_cast<unsigned>((lines + M - 1) / M); };
  if (nb > 65535) throw Error(kInvalidArgument, "band batch too large for the local stage");
  const dim3 gz(blocks(kb.max_lines, az.M), static_cast<unsigned>(nb), nk);
  const dim3 gy(blocks(static_cast<long long>(nb) * kb.max_planes * n3, ay.M), nk);
  int nbi = nb;
  // fused y-x-y pass: every k point's active planes one run, the widest
  // within the register budget, the LDS of a band group within 64 KiB
  {
    const int pwf = kb.max_planes;
    int G = std::max(1, std::min(g_tune.ygroup[prec], nb));
    auto lds_of = [&](int g) { return (2 * size_t(g) * pwf * ay.ld + n2) * sizeof(V); };
    while (G > 1 && lds_of(G) > 65536) --G;
    const size_t lds = lds_of(G);
    const YxyKernel yxy =
        (vx && g_tune.fused[prec] && kb.contiguous && lds <= 65536) ? yxy_for<R>(pwf, n1, n2, G, ay.ld) : YxyKernel{};
    if (yxy.fn) {
      const cplx_t<R> *vxT = vx + static_cast<long long>(n1) * C;
      {
        void *args[] = {&psi, &lp, &boxes, &tabs, &ks, &nbi, &az};
        launch<R>("local_z_bwd", z_bwd_for<R, true>(n3), az, s, args, gz);
      }
      {
        void *args[] = {&boxes, &vxT, &tabs, &ks, &nbi, &G, &ay};
        trace_mark("local_yxy", s);
        HALF_HIP(hipLaunchKernel(yxy.fn, dim3(static_cast<unsigned>(n3 * ((nb + G - 1) / G)), nk),
                                 dim3(static_cast<unsigned>(yxy.threads)), args, lds, s));
      }
      {
        const cplx_t<R> *ah = add.h, *as = add.s;
        long long ald = add.ld, akk = add.kstride;
        void *args[] = {&psi, &lp, &boxes, &kin, &hpsi, &spsi, &lo, &tabs, &ks, &nbi, &az, &ah, &as, &ald, &akk};
        if (add.wait) HALF_HIP(hipStreamWaitEvent(s, add.wait, 0));
        launch<R>("local_z_fwd", z_fwd_for<R, true>(n3), az, s, args, gz);
      }
      return;
    }
  }
  {
    void *args[] = {&psi, &lp, &boxes, &tabs, &ks, &nbi, &az};
    launch<R>("local_z_bwd", z_bwd_for<R, false>(n3), az, s, args, gz);
  }
  {
    void *args[] = {&boxes, &tabs, &ks, &nbi, &ay};
    launch<R>("local_y_bwd", y_for<R, true>(n2), ay, s, args, gy);
  }
  // x pass: per-column convolution where every k point's active planes are
  // one run and the widest fits the register budget (on the matrix cores
  // with xconv 2), else transform-based
  int pmax = 0;
  int pw = kb.max_planes;
  const size_t xconv_lds = size_t(2 * pw - 1) * kXcCols * sizeof(V);
  const bool conv_ok = vx && kb.contiguous && pw <= n1;
  int xm_threads = 256;
  const void *xmfma = conv_ok && g_tune.xconv[prec] >= 2 ? x_mfma_for<R>(pw, g_tune.xconv[prec], xm_threads)
                                                          : nullptr;
  const bool xconv_ok = conv_ok && g_tune.xconv[prec] && xconv_lds <= 65536;
  const void *xconv = !xmfma && xconv_ok ? x_conv_for<R>(pw, pmax) : nullptr;
  if (xmfma) {
    const int groups = (C + kXmCols - 1) / kXmCols;
    int per_xcd = (groups + 7) / 8;
    const unsigned gx = static_cast<unsigned>(8 * per_xcd);
    const unsigned gyb = static_cast<unsigned>((nb + 15) / 16);
    void *args[] = {&boxes, &vx, &tabs, &ks, &nbi, &per_xcd};
    trace_mark("local_x_mfma", s);
    HALF_HIP(hipLaunchKernel(xmfma, dim3(gx, gyb, nk), dim3(static_cast<unsigned>(xm_threads)), args, 0, s));
  } else if (xconv) {
    void *args[] = {&boxes, &vx, &tabs, &ks, &nbi, &pw};
    trace_mark("local_x_conv", s);
    HALF_HIP(hipLaunchKernel(xconv, dim3((C + kXcCols - 1) / kXcCols, (nb + kXcBands - 1) / kXcBands, nk),
                             dim3(256), args, xconv_lds, s));
  } else {
    int G = std::max(1, std::min(g_tune.xgroup[prec], nb));
    while (G > 1 && (G & (G - 1))) --G;  // power of two, divides M
    FftArgs ax = axis_args<R>(plan, 0, 0, C, G);
    if (ax.M % G) G = 1;
    const int Mc = ax.M / G;
    const dim3 gx(static_cast<unsigned>(((C + Mc - 1) / Mc) * ((nb + G - 1) / G)), nk);
    R scale = static_cast<R>(1.0 / double(static_cast<long long>(n1) * C));
    void *args[] = {&boxes, &veff, &scale, &tabs, &ks, &ax, &G, &nbi};
    launch<R>("local_x_veff", reinterpret_cast<const void *>(&x_veff_kernel<R>), ax, s, args, gx,
              size_t(Mc) * ax.n * sizeof(R));
  }
  {
    void *args[] = {&boxes, &tabs, &ks, &nbi, &ay};
    launch<R>("local_y_fwd", y_for<R, false>(n2), ay, s, args, gy);
  }
  {
    const cplx_t<R> *ah = add.h, *as = add.s;
    long long ald = add.ld, akk = add.kstride;
    void *args[] = {&psi, &lp, &boxes, &kin, &hpsi, &spsi, &lo, &tabs, &ks, &nbi, &az, &ah, &as, &ald, &akk};
    if (add.wait) HALF_HIP(hipStreamWaitEvent(s, add.wait, 0));
    launch<R>("local_z_fwd", z_fwd_for<R, false>(n3), az, s, args, gz);
  }
}

template <typename R>
void local_apply(const FftPlan &plan, const LocalMap &m, const R *veff, const cplx_t<R> *vx, const R *kin, int nb,
                 const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi, cplx_t<R> *spsi, int64_t ldout,
                 cplx_t<R> *boxes, hipStream_t s, const LocalAddend<R> &add) {
  if (m.nlines == 0) return;
  LocalBatch kb;
  kb.nk = 1;
  kb.tabs = m.dtab.get();
  kb.max_lines = m.nlines;
  kb.max_planes = m.nplanes;
  kb.contiguous = m.planes_contiguous;
  local_apply_batch<R>(plan, kb, veff, vx, kin, nb, psi, ldpsi, hpsi, spsi, ldout, boxes, s, add);
}

bool set_local_tuning(const std::string &key, int value) {
  static const char *axes[3] = {"x", "y", "z"};
  for (int p = 0; p < 2; ++p) {
    const std::string pre = std::string("local.") + (p == 0 ? "fp64" : "fp32");
    for (int a = 0; a < 3; ++a)
      if (key == pre + ".tile." + axes[a]) {
        if (value < 16 || value > 256 * 16) return false;
        g_tune.tile[p][a] = value;
        return true;
      }
    if (key == pre + ".xgroup") {
      if (value < 1 || value > 64) return false;
      g_tune.xgroup[p] = value;
      return true;
    }
    if (key == pre + ".xconv") {
      if (value < 0 || value > 3) return false;
      g_tune.xconv[p] = value;
      return true;
    }
    if (key == pre + ".fused") {
      if (value < 0 || value > 1) return false;
      g_tune.fused[p] = value;
      return true;
    }
    if (key == pre + ".ygroup") {
      if (value < 1 || value > 8) return false;
      g_tune.ygroup[p] = value;
      return true;
    }
    if (key == pre + ".vxlds") {
      if (value < 0 || value > 1) return false;
      g_tune.vxlds[p] = value;
      return true;
    }
  }
  return false;
}

bool get_local_tuning(const std::string &key, int &value) {
  static const char *axes[3] = {"x", "y", "z"};
  for (int p = 0; p < 2; ++p) {
    const std::string pre = std::string("local.") + (p == 0 ? "fp64" : "fp32");
    for (int a = 0; a < 3; ++a)
      if (key == pre + ".tile." + axes[a]) {
        value = g_tune.tile[p][a];
        return true;
      }
    if (key == pre + ".xgroup") {
      value = g_tune.xgroup[p];
      return true;
    }
    if (key == pre + ".xconv") {
      value = g_tune.xconv[p];
      return true;
    }
    if (key == pre + ".fused") {
      value = g_tune.fused[p];
      return true;
    }
    if (key == pre + ".ygroup") {
      value = g_tune.ygroup[p];
      return true;
    }
    if (key == pre + ".vxlds") {
      value = g_tune.vxlds[p];
      return true;
    }
  }
  return false;
}

#define HALF_LOCAL(R)                                                                                              \
  template void local_apply_batch<R>(const FftPlan &, const LocalBatch &, const R *, const cplx_t<R> *, const R *, \
                                     int, const cplx_t<R> *, int64_t, cplx_t<R> *, cplx_t<R> *, int64_t,             \
                                     cplx_t<R> *, hipStream_t, const LocalAddend<R> &);                            \
  template void local_apply<R>(const FftPlan &, const LocalMap &, const R *, const cplx_t<R> *, const R *, int,     \
                               const cplx_t<R> *, int64_t, cplx_t<R> *, cplx_t<R> *, int64_t, cplx_t<R> *,         \
                               hipStream_t, const LocalAddend<R> &);
HALF_LOCAL(double)
HALF_LOCAL(float)
#undef HALF_LOCAL

}  // namespace half
