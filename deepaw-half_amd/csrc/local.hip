// Pruned local stage of H.psi: T psi + F[Veff F^-1[psi]] / N_r.
//
// Only the plane-wave sphere is populated in G space, so most 1D transforms
// of a full 3D FFT act on zeros.  With the C-order box (i1 slowest, i3
// contiguous) and the basis enumerated n1, n2, n3 ascending (so the plane
// waves of one (i1, i2) line are contiguous), the transform is done in four
// kernels that touch only what is needed:
//
//   z_bwd   active (i1,i2) lines: plane waves scattered into LDS straight
//           from psi (no box clearing), backward transform along i3
//   y_bwd   active i1 planes: backward transform along i2, loading only the
//           active lines of the plane (the rest are zeros, never stored)
//   x_conv  every (i2,i3) column: restricted to the active i1 planes the
//           backward transform along i1, the multiply by Veff/N_r and the
//           forward transform are one circular convolution,
//             out[p'] = sum_p Vx[(p' - p) mod n1][c] in[p],
//           with Vx[d][c] = sum_x Veff[x][c]/N_r e^{-2 pi i d x/n1} computed
//           once per potential; each active plane is read and written once
//           (x_veff, the transform-based form, serves boxes whose active
//           plane count exceeds the register budget)
//   y_fwd   active planes: forward transform along i2, store active lines
//   z_fwd   active lines: forward transform along i3, gather the plane-wave
//           coefficients and add T psi (S psi = psi is written here too)
//
// Where a whole n2 x n3 plane fits in LDS, z and y are fused per active
// plane instead (plane_bwd: scatter, z transforms of the plane's active
// lines, y transforms of all its columns; plane_fwd: the reverse, ending in
// the gather), and only the active planes are stored, compactly in signed
// order, so the local stage is three kernels and each active plane crosses
// HBM four times per band application.
//
// Box entries outside the active lines/planes are never read or written.
// Numerically this is the same circular convolution as the reference's
// scatter / inverse FFT / multiply / forward FFT / gather sequence
// (reference: src/half_cuda_assembly.cuf:180-218, 484-498).
#include <algorithm>

#include "local.hpp"
#include "stockham.hpp"
#include "trace.hpp"

namespace half {

namespace {

// Tunables (FFT points per workgroup tile per kernel class and precision,
// and bands sharing one Veff tile in the x pass); set through
// half_amd_set_tuning, chosen by the autotuner in half_amd/autotune.py.
struct Tuning {
  // defaults from the autotuner on the 48^3 HfO2-like box, 64 bands
  int tile[2][3] = {{1024, 1536, 768}, {1536, 1536, 768}};  // [fp64|fp32][x|y|z]
  int xgroup[2] = {4, 8};
  int xconv[2] = {1, 1};  // x pass as a per-column convolution when it fits
  // lines per LDS batch of the plane-fused z/y passes; off by default: with
  // one plane per workgroup its transform batches keep too few threads busy
  // (measured slower than the separate passes on the 48^3 box)
  int plane[2] = {0, 0};
};
Tuning g_tune;

struct Tables {
  int n1, n2, n3;
  long long nr;
  int nlines;
  const int *line_i12, *line_start, *line_count, *line_of;
  int nplanes;
  const int *plane_i1, *plane_on;
  const int *pw_i3;
  int plane_s0;  // first signed plane of the contiguous run (x_conv_kernel)
  int compact;   // planes stored compactly: plane q of a band at (band P + q) n2 n3
};

// Elements per thread of a tile (M * n <= tile budget <= 256 * KMAX).
template <typename R>
struct Tile {
  static constexpr int KMAX = std::is_same<R, double>::value ? 8 : 16;
};

template <typename R>
__global__ __launch_bounds__(256) void z_bwd_kernel(const cplx_t<R> *psi, long long ldpsi, cplx_t<R> *box,
                                                    Tables t, FftArgs a) {
  using V = cplx_t<R>;
  constexpr int K = Tile<R>::KMAX;
  extern __shared__ double2 smem_raw[];
  V *buf0 = reinterpret_cast<V *>(smem_raw);
  V *buf1 = buf0 + a.M * a.ld;
  V *tws = buf1 + a.M * a.ld;
  load_twiddles(tws, a);
  const int L0 = blockIdx.x * a.M;
  const int MN = a.M * a.n;
  for (int e = threadIdx.x; e < a.M * a.ld; e += blockDim.x) buf0[e] = mk<V>(R(0), R(0));
  __syncthreads();
  // plane waves of each line, placed at their i3 (the rest of the line is 0)
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = threadIdx.x + k * 256;
    if (e < MN) {
      const int l = e / a.n, j = e - l * a.n;
      const int L = L0 + l;
      if (L < a.lines_total) {
        const int band = L / t.nlines, line = L - band * t.nlines;
        if (j < t.line_count[line]) {
          const int g = t.line_start[line] + j;
          buf0[l * a.ld + t.pw_i3[g]] = psi[g + band * ldpsi];
        }
      }
    }
  }
  __syncthreads();
  V *res = run_stages(buf0, buf1, tws, a, +1);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = threadIdx.x + k * 256;
    if (e >= MN) break;
    const int l = e / a.n, p = e - l * a.n;
    const int L = L0 + l;
    if (L < a.lines_total) {
      const int band = L / t.nlines, line = L - band * t.nlines;
      box[band * t.nr + static_cast<long long>(t.line_i12[line]) * t.n3 + p] = res[l * a.ld + p];
    }
  }
}

// Along i2 on the active planes; BWD loads active lines only and stores all,
// the forward pass loads all and stores active lines only.  Each thread keeps
// one line (M divides the 256 threads) and walks its points.
template <typename R, bool BWD>
__global__ __launch_bounds__(256) void y_kernel(cplx_t<R> *box, Tables t, FftArgs a) {
  using V = cplx_t<R>;
  constexpr int K = Tile<R>::KMAX;
  extern __shared__ double2 smem_raw[];
  V *buf0 = reinterpret_cast<V *>(smem_raw);
  V *buf1 = buf0 + a.M * a.ld;
  V *tws = buf1 + a.M * a.ld;
  load_twiddles(tws, a);
  const int l = threadIdx.x % a.M, p0 = threadIdx.x / a.M, pstep = blockDim.x / a.M;
  const int L = blockIdx.x * a.M + l;
  const bool live = L < a.lines_total;
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
  V *res = run_stages(buf0, buf1, tws, a, BWD ? +1 : -1);
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
__global__ __launch_bounds__(256) void x_veff_kernel(cplx_t<R> *box, const R *veff, R scale, Tables t, FftArgs a,
                                                     int G, int nb) {
  using V = cplx_t<R>;
  constexpr int K = Tile<R>::KMAX;
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

// Plane-fused z and y passes.  Workgroup = one (band, active plane q).
// LDS: the plane pb[n2][n3 + 1], two transform batches of M lines of
// stride ldw = max(n2, n3) + 1, and both axes' twiddles.
struct PlaneArgs {
  const int *line0, *nline;  // lines of plane q (signed plane order)
  int P, M, ldp, ldw;
};

template <typename R>
__device__ inline void plane_lds(cplx_t<R> *smem, const PlaneArgs &pa, int n2, int n3, cplx_t<R> *&pb,
                                 cplx_t<R> *&w0, cplx_t<R> *&w1, cplx_t<R> *&tz, cplx_t<R> *&ty) {
  pb = smem;
  w0 = pb + n2 * pa.ldp;
  w1 = w0 + pa.M * pa.ldw;
  tz = w1 + pa.M * pa.ldw;
  ty = tz + n3;
}

// backward: psi -> plane q of `planes` (transform along i3 of the plane's
// active lines, then along i2 of every column)
template <typename R>
__global__ __launch_bounds__(256) void plane_bwd_kernel(const cplx_t<R> *psi, long long ldpsi, cplx_t<R> *planes,
                                                        Tables t, PlaneArgs pa, FftArgs az, FftArgs ay) {
  using V = cplx_t<R>;
  extern __shared__ double2 smem_raw[];
  V *pb, *w0, *w1, *tz, *ty;
  plane_lds<R>(reinterpret_cast<V *>(smem_raw), pa, t.n2, t.n3, pb, w0, w1, tz, ty);
  load_twiddles(tz, az);
  load_twiddles(ty, ay);
  const int q = blockIdx.x % pa.P, band = blockIdx.x / pa.P;
  const int n2 = t.n2, n3 = t.n3, M = pa.M;
  for (int e = threadIdx.x; e < n2 * pa.ldp; e += blockDim.x) pb[e] = mk<V>(R(0), R(0));
  const int l0 = pa.line0[q], nl = pa.nline[q];
  const V *pband = psi + band * ldpsi;
  for (int c0 = 0; c0 < nl; c0 += M) {
    __syncthreads();
    for (int e = threadIdx.x; e < M * pa.ldw; e += blockDim.x) w0[e] = mk<V>(R(0), R(0));
    __syncthreads();
    for (int e = threadIdx.x; e < M * n3; e += blockDim.x) {
      const int li = e / n3, j = e - li * n3;
      if (c0 + li < nl) {
        const int line = l0 + c0 + li;
        if (j < t.line_count[line]) {
          const int g = t.line_start[line] + j;
          w0[li * pa.ldw + t.pw_i3[g]] = pband[g];
        }
      }
    }
    __syncthreads();
    const V *res = run_stages(w0, w1, tz, az, +1);
    for (int e = threadIdx.x; e < M * n3; e += blockDim.x) {
      const int li = e / n3, j = e - li * n3;
      if (c0 + li < nl) pb[(t.line_i12[l0 + c0 + li] % n2) * pa.ldp + j] = res[li * pa.ldw + j];
    }
  }
  for (int j0 = 0; j0 < n3; j0 += M) {
    __syncthreads();
    for (int e = threadIdx.x; e < M * n2; e += blockDim.x) {
      const int ci = e / n2, i2 = e - ci * n2;
      w0[ci * pa.ldw + i2] = j0 + ci < n3 ? pb[i2 * pa.ldp + j0 + ci] : mk<V>(R(0), R(0));
    }
    __syncthreads();
    const V *res = run_stages(w0, w1, ty, ay, +1);
    for (int e = threadIdx.x; e < M * n2; e += blockDim.x) {
      const int ci = e / n2, i2 = e - ci * n2;
      if (j0 + ci < n3) pb[i2 * pa.ldp + j0 + ci] = res[ci * pa.ldw + i2];
    }
  }
  __syncthreads();
  V *out = planes + (static_cast<long long>(band) * pa.P + q) * n2 * n3;
  for (int e = threadIdx.x; e < n2 * n3; e += blockDim.x) {
    const int i2 = e / n3, i3 = e - i2 * n3;
    out[e] = pb[i2 * pa.ldp + i3];
  }
}

// forward: plane q -> hpsi = (transform) + T psi, spsi = psi for the plane's
// plane waves (transform along i2 of every column, then along i3 of the
// active lines, then gather)
template <typename R>
__global__ __launch_bounds__(256) void plane_fwd_kernel(const cplx_t<R> *psi, long long ldpsi, const cplx_t<R> *planes,
                                                        const R *kin, cplx_t<R> *hpsi, cplx_t<R> *spsi,
                                                        long long ldout, Tables t, PlaneArgs pa, FftArgs az,
                                                        FftArgs ay) {
  using V = cplx_t<R>;
  extern __shared__ double2 smem_raw[];
  V *pb, *w0, *w1, *tz, *ty;
  plane_lds<R>(reinterpret_cast<V *>(smem_raw), pa, t.n2, t.n3, pb, w0, w1, tz, ty);
  load_twiddles(tz, az);
  load_twiddles(ty, ay);
  const int q = blockIdx.x % pa.P, band = blockIdx.x / pa.P;
  const int n2 = t.n2, n3 = t.n3, M = pa.M;
  const V *in = planes + (static_cast<long long>(band) * pa.P + q) * n2 * n3;
  for (int e = threadIdx.x; e < n2 * n3; e += blockDim.x) {
    const int i2 = e / n3, i3 = e - i2 * n3;
    pb[i2 * pa.ldp + i3] = in[e];
  }
  for (int j0 = 0; j0 < n3; j0 += M) {
    __syncthreads();
    for (int e = threadIdx.x; e < M * n2; e += blockDim.x) {
      const int ci = e / n2, i2 = e - ci * n2;
      w0[ci * pa.ldw + i2] = j0 + ci < n3 ? pb[i2 * pa.ldp + j0 + ci] : mk<V>(R(0), R(0));
    }
    __syncthreads();
    const V *res = run_stages(w0, w1, ty, ay, -1);
    for (int e = threadIdx.x; e < M * n2; e += blockDim.x) {
      const int ci = e / n2, i2 = e - ci * n2;
      if (j0 + ci < n3) pb[i2 * pa.ldp + j0 + ci] = res[ci * pa.ldw + i2];
    }
  }
  const int l0 = pa.line0[q], nl = pa.nline[q];
  const V *pband = psi + band * ldpsi;
  for (int c0 = 0; c0 < nl; c0 += M) {
    __syncthreads();
    for (int e = threadIdx.x; e < M * n3; e += blockDim.x) {
      const int li = e / n3, j = e - li * n3;
      w0[li * pa.ldw + j] = c0 + li < nl ? pb[(t.line_i12[l0 + c0 + li] % n2) * pa.ldp + j] : mk<V>(R(0), R(0));
    }
    __syncthreads();
    const V *res = run_stages(w0, w1, tz, az, -1);
    for (int e = threadIdx.x; e < M * n3; e += blockDim.x) {
      const int li = e / n3, j = e - li * n3;
      if (c0 + li >= nl) continue;
      const int line = l0 + c0 + li;
      if (j < t.line_count[line]) {
        const int g = t.line_start[line] + j;
        const V p = pband[g];
        const V r = res[li * pa.ldw + t.pw_i3[g]];
        hpsi[g + band * ldout] = mk<V>(r.x + kin[g] * p.x, r.y + kin[g] * p.y);
        spsi[g + band * ldout] = p;
      }
    }
  }
}

// x pass as a convolution over the active planes (see the header), for a
// contiguous run of P planes s0 .. s0 + P - 1 (signed, modulo n1): with the
// planes in that order, plane p' takes sum_p Vx[(p' - p) mod n1] in[p].  A
// workgroup takes kXcCols consecutive (i2,i3) columns and kXcBands bands;
// lanes 0-31 and 32-63 of a wave hold two bands of the same columns.  LDS
// holds the 2P - 1 needed rows of Vx ordered by the difference p' - p + P - 1,
// so every term of the sum is one LDS read at a fixed offset; each thread
// keeps its column's P values in registers (PMAX >= P) and overwrites them.
constexpr int kXcCols = 32;
constexpr int kXcBands = 8;

template <typename R, int PMAX>
__global__ __launch_bounds__(256) void x_conv_kernel(cplx_t<R> *box, const cplx_t<R> *vx, Tables t, int nb) {
  using V = cplx_t<R>;
  extern __shared__ double2 smem_raw[];
  V *vs = reinterpret_cast<V *>(smem_raw);  // [2P - 1][kXcCols]
  const int C = t.n2 * t.n3;
  const int P = t.nplanes;
  const int c0 = blockIdx.x * kXcCols;
  for (int e = threadIdx.x; e < (2 * P - 1) * kXcCols; e += blockDim.x) {
    const int k = e / kXcCols, c = c0 + (e - k * kXcCols);
    int d = (k - (P - 1)) % t.n1;
    d += d < 0 ? t.n1 : 0;
    vs[e] = c < C ? vx[static_cast<long long>(d) * C + c] : mk<V>(R(0), R(0));
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ci = lane & 31;
  const int c = c0 + ci;
  const int band = blockIdx.y * kXcBands + wave * 2 + (lane >> 5);
  if (c >= C || band >= nb) return;  // no barrier follows
  V *col = box + (t.compact ? static_cast<long long>(band) * P * C : static_cast<long long>(band) * t.nr) + c;
  const int s0 = t.plane_s0 < 0 ? t.plane_s0 + t.n1 : t.plane_s0;
  V in[PMAX];
#pragma unroll
  for (int q = 0; q < PMAX; ++q) {
    int i1 = s0 + q;
    i1 -= i1 >= t.n1 ? t.n1 : 0;
    if (t.compact) i1 = q;
    in[q] = q < P ? col[static_cast<long long>(i1) * C] : mk<V>(R(0), R(0));
  }
  for (int op = 0; op < P; ++op) {
    const V *w = vs + (op + P - 1) * kXcCols + ci;  // w[-q kXcCols] = Vx[(op - q) mod n1]
    R ar = 0, ai = 0;
#pragma unroll
    for (int q = 0; q < PMAX; ++q) {
      if (q < P) {
        const V v = w[-q * kXcCols];
        ar += v.x * in[q].x - v.y * in[q].y;
        ai += v.x * in[q].y + v.y * in[q].x;
      }
    }
    int i1 = s0 + op;
    i1 -= i1 >= t.n1 ? t.n1 : 0;
    if (t.compact) i1 = op;
    col[static_cast<long long>(i1) * C] = mk<V>(ar, ai);
  }
}

// Vx[d][c] = scale sum_x veff[x][c] e^{-2 pi i d x / n1} for the C = n2 n3
// columns of the C-order box, in FP64 and FP32
__global__ void veff_x_dft_kernel(const double *veff, double2 *vx, float2 *vx32, int n1, int C, double scale) {
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
}

template <typename R>
__global__ __launch_bounds__(256) void z_fwd_kernel(const cplx_t<R> *psi, long long ldpsi, const cplx_t<R> *box,
                                                    const R *kin, cplx_t<R> *hpsi, cplx_t<R> *spsi,
                                                    long long ldout, Tables t, FftArgs a) {
  using V = cplx_t<R>;
  constexpr int K = Tile<R>::KMAX;
  extern __shared__ double2 smem_raw[];
  V *buf0 = reinterpret_cast<V *>(smem_raw);
  V *buf1 = buf0 + a.M * a.ld;
  V *tws = buf1 + a.M * a.ld;
  load_twiddles(tws, a);
  const int L0 = blockIdx.x * a.M;
  const int MN = a.M * a.n;
  V v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = threadIdx.x + k * 256;
    v[k] = mk<V>(R(0), R(0));
    if (e < MN) {
      const int l = e / a.n, p = e - l * a.n;
      const int L = L0 + l;
      if (L < a.lines_total) {
        const int band = L / t.nlines, line = L - band * t.nlines;
        v[k] = box[band * t.nr + static_cast<long long>(t.line_i12[line]) * t.n3 + p];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = threadIdx.x + k * 256;
    if (e < MN) {
      const int l = e / a.n, p = e - l * a.n;
      buf0[l * a.ld + p] = v[k];
    }
  }
  __syncthreads();
  V *res = run_stages(buf0, buf1, tws, a, -1);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = threadIdx.x + k * 256;
    if (e >= MN) break;
    const int l = e / a.n, j = e - l * a.n;
    const int L = L0 + l;
    if (L >= a.lines_total) continue;
    const int band = L / t.nlines, line = L - band * t.nlines;
    if (j < t.line_count[line]) {
      const int g = t.line_start[line] + j;
      const V p = psi[g + band * ldpsi];
      const V r = res[l * a.ld + t.pw_i3[g]];
      hpsi[g + band * ldout] = mk<V>(r.x + kin[g] * p.x, r.y + kin[g] * p.y);
      spsi[g + band * ldout] = p;
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
void launch(const char *name, const void *fn, const FftArgs &a, hipStream_t s, void **args, long long blocks = -1,
            size_t extra_lds = 0) {
  trace_mark(name, s);
  if (blocks < 0) blocks = (a.lines_total + a.M - 1) / a.M;
  HALF_HIP(hipLaunchKernel(fn, dim3(static_cast<unsigned>(blocks)), dim3(256), args, lds_bytes<R>(a) + extra_lds, s));
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
  auto signed_of = [&](int i) { return i <= (n1 - 1) / 2 ? i : i - n1; };
  std::vector<int> signed_planes;
  for (int i : planes) signed_planes.push_back(signed_of(i));
  std::sort(signed_planes.begin(), signed_planes.end());
  planes_contiguous = !signed_planes.empty();
  for (size_t q = 1; q < signed_planes.size(); ++q)
    if (signed_planes[q] != signed_planes[0] + static_cast<int>(q)) planes_contiguous = false;
  plane_s0 = signed_planes.empty() ? 0 : signed_planes[0];
  // lines of each plane in signed plane order; the basis enumerates the
  // signed i1 outermost, so each plane's lines form one run
  std::vector<int> q_of(n1, -1);
  for (size_t q = 0; q < signed_planes.size(); ++q) q_of[(signed_planes[q] % n1 + n1) % n1] = static_cast<int>(q);
  std::vector<int> qline0(signed_planes.size(), -1), qnline(signed_planes.size(), 0);
  lines_by_plane = true;
  int prev_q = -1;
  for (size_t l = 0; l < lines.size(); ++l) {
    const int q = q_of[lines[l] / n2];
    if (q != prev_q) {
      if (q < prev_q || qline0[q] >= 0) lines_by_plane = false;
      qline0[q] = static_cast<int>(l);
      prev_q = q;
    }
    ++qnline[q];
  }
  qplane_line0.upload(qline0, s);
  qplane_nline.upload(qnline, s);
  nlines = static_cast<int>(lines.size());
  nplanes = static_cast<int>(planes.size());
  line_i12.upload(lines, s);
  line_start.upload(starts, s);
  line_count.upload(counts, s);
  line_of.upload(line_of_h, s);
  plane_i1.upload(planes, s);
  plane_on.upload(plane_on_h, s);
  pw_i3.upload(i3, s);
  HALF_HIP(hipStreamSynchronize(s));
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

void build_veff_x_dft(const double *veff, int n1, int n2, int n3, double2 *vx, float2 *vx32, hipStream_t s) {
  const long long total = static_cast<long long>(n1) * n2 * n3;
  const double scale = 1.0 / static_cast<double>(total);
  hipLaunchKernelGGL(veff_x_dft_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s, veff, vx,
                     vx32, n1, n2 * n3, scale);
  HALF_HIP(hipGetLastError());
}

template <typename R>
void local_apply(const FftPlan &plan, const LocalMap &m, const R *veff, const cplx_t<R> *vx, const R *kin, int nb,
                 const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi, cplx_t<R> *spsi, int64_t ldout,
                 cplx_t<R> *boxes, hipStream_t s) {
  using V = cplx_t<R>;
  if (nb <= 0 || m.nlines == 0) return;
  if (static_cast<long long>(nb) * (static_cast<long long>(m.n1) * m.n2 + static_cast<long long>(m.n2) * m.n3 +
                                   static_cast<long long>(m.n1) * m.n3) >= (1ll << 31))
    throw Error(kInvalidArgument, "band batch too large for the local stage");
  Tables t{m.n1, m.n2, m.n3, static_cast<long long>(m.n1) * m.n2 * m.n3, m.nlines, m.line_i12.get(),
           m.line_start.get(), m.line_count.get(), m.line_of.get(), m.nplanes, m.plane_i1.get(), m.plane_on.get(),
           m.pw_i3.get(), m.plane_s0, 0};
  long long lp = ldpsi, lo = ldout;
  const int prec = std::is_same<R, double>::value ? 0 : 1;
  const int C = m.n2 * m.n3;

  // x pass form: per-column convolution where the active planes are one run
  // and fit the register budget, else transform / multiply / transform
  int pmax = 0;
  const size_t xconv_lds = size_t(2 * m.nplanes - 1) * kXcCols * sizeof(V);
  const bool xconv_ok = vx && g_tune.xconv[prec] && m.planes_contiguous && m.nplanes <= m.n1 && xconv_lds <= 65536;
  const void *xconv = xconv_ok ? x_conv_for<R>(m.nplanes, pmax) : nullptr;

  // z and y fused per plane when a plane and two transform batches fit LDS
  const int ldp = m.n3 + 1, ldw = std::max(m.n2, m.n3) + 1;
  int Mp = g_tune.plane[prec];
  auto plane_bytes = [&](int M) { return (size_t(m.n2) * ldp + 2 * size_t(M) * ldw + m.n2 + m.n3) * sizeof(V); };
  while (Mp > 1 && plane_bytes(Mp) > 65536) Mp /= 2;
  const bool fused = xconv && Mp >= 1 && m.lines_by_plane && plane_bytes(Mp) <= 65536;

  if (fused) {
    t.compact = 1;
    FftArgs az = axis_args<R>(plan, 2, 0, 1), ay = axis_args<R>(plan, 1, 0, m.n3);
    az.M = ay.M = Mp;
    az.ld = ay.ld = ldw;
    PlaneArgs pa{m.qplane_line0.get(), m.qplane_nline.get(), m.nplanes, Mp, ldp, ldw};
    const dim3 grid(static_cast<unsigned>(int64_t(m.nplanes) * nb));
    const size_t lds = plane_bytes(Mp);
    {
      void *args[] = {&psi, &lp, &boxes, &t, &pa, &az, &ay};
      trace_mark("local_plane_bwd", s);
      HALF_HIP(hipLaunchKernel(reinterpret_cast<const void *>(&plane_bwd_kernel<R>), grid, dim3(256), args, lds, s));
    }
    {
      int nbi = nb;
      void *args[] = {&boxes, &vx, &t, &nbi};
      trace_mark("local_x_conv", s);
      HALF_HIP(hipLaunchKernel(xconv, dim3((C + kXcCols - 1) / kXcCols, (nb + kXcBands - 1) / kXcBands), dim3(256),
                               args, xconv_lds, s));
    }
    {
      void *args[] = {&psi, &lp, &boxes, &kin, &hpsi, &spsi, &lo, &t, &pa, &az, &ay};
      trace_mark("local_plane_fwd", s);
      HALF_HIP(hipLaunchKernel(reinterpret_cast<const void *>(&plane_fwd_kernel<R>), grid, dim3(256), args, lds, s));
    }
    return;
  }

  FftArgs az = axis_args<R>(plan, 2, static_cast<long long>(nb) * m.nlines, 1);
  FftArgs ay = axis_args<R>(plan, 1, static_cast<long long>(nb) * m.nplanes * m.n3, m.n3);
  int G = std::max(1, std::min(g_tune.xgroup[prec], nb));
  while (G > 1 && (G & (G - 1))) --G;  // power of two, divides M
  FftArgs ax = axis_args<R>(plan, 0, static_cast<long long>(nb) * m.n2 * m.n3, m.n2 * m.n3, G);
  if (ax.M % G) G = 1;
  const int Mc = ax.M / G;
  const long long xblocks = static_cast<long long>((m.n2 * m.n3 + Mc - 1) / Mc) * ((nb + G - 1) / G);
  R scale = static_cast<R>(1.0 / double(t.nr));
  {
    void *args[] = {&psi, &lp, &boxes, &t, &az};
    launch<R>("local_z_bwd", reinterpret_cast<const void *>(&z_bwd_kernel<R>), az, s, args);
  }
  {
    void *args[] = {&boxes, &t, &ay};
    launch<R>("local_y_bwd", reinterpret_cast<const void *>(&y_kernel<R, true>), ay, s, args);
  }
  if (xconv) {
    int nbi = nb;
    void *args[] = {&boxes, &vx, &t, &nbi};
    trace_mark("local_x_conv", s);
    HALF_HIP(hipLaunchKernel(xconv, dim3((C + kXcCols - 1) / kXcCols, (nb + kXcBands - 1) / kXcBands), dim3(256),
                             args, xconv_lds, s));
  } else {
    int nbi = nb;
    void *args[] = {&boxes, &veff, &scale, &t, &ax, &G, &nbi};
    launch<R>("local_x_veff", reinterpret_cast<const void *>(&x_veff_kernel<R>), ax, s, args, xblocks,
              size_t(Mc) * ax.n * sizeof(R));
  }
  {
    void *args[] = {&boxes, &t, &ay};
    launch<R>("local_y_fwd", reinterpret_cast<const void *>(&y_kernel<R, false>), ay, s, args);
  }
  {
    void *args[] = {&psi, &lp, &boxes, &kin, &hpsi, &spsi, &lo, &t, &az};
    launch<R>("local_z_fwd", reinterpret_cast<const void *>(&z_fwd_kernel<R>), az, s, args);
  }
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
      if (value < 0 || value > 1) return false;
      g_tune.xconv[p] = value;
      return true;
    }
    if (key == pre + ".plane") {
      if (value < 0 || value > 64 || (value & (value - 1))) return false;
      g_tune.plane[p] = value;
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
    if (key == pre + ".plane") {
      value = g_tune.plane[p];
      return true;
    }
  }
  return false;
}

template void local_apply<double>(const FftPlan &, const LocalMap &, const double *, const double2 *,
                                  const double *, int, const double2 *, int64_t, double2 *, double2 *, int64_t,
                                  double2 *, hipStream_t);
template void local_apply<float>(const FftPlan &, const LocalMap &, const float *, const float2 *, const float *, int,
                                 const float2 *, int64_t, float2 *, float2 *, int64_t, float2 *, hipStream_t);

}  // namespace half
