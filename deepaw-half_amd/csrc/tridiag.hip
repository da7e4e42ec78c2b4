// Householder reduction of a Hermitian matrix to real symmetric tridiagonal
// form, LAPACK zhetrd 'U' conventions (the unblocked zhetd2 sequence:
// for i = n-2 .. 0, a reflector annihilates A(0:i-1, i+1), then
// A(0:i, 0:i) -= v w^H + w v^H with w = tau A v - tau/2 (tau A v)^H v v).
//
// The Rayleigh-Ritz problems of the all-band solve are small (2m = 80-300)
// and one matrix at a time, where a library reduction is a chain of a few
// launches per column and takes milliseconds.  Here one workgroup does the
// whole reduction: 512 threads, the matrix in global memory (it stays in
// the L2 of the XCD that runs the workgroup), the vectors in LDS.  Each
// column step is one pass over the active block that applies the previous
// step's rank-2 update and forms this step's product A v in the same read:
// the reflector only needs the one column the update changes first, so
// that column is updated ahead of the pass.  Lanes own rows (column-major
// loads are coalesced), waves own columns; the eight per-wave partial
// products are added in a fixed order, so results do not depend on timing.
// A step has three barriers: wave 0 forms the reflector from the column the
// previous pass left in LDS, all waves make the pass, every thread forms
// its row of w.  The pass streams the block through one CU (about 1 ms at
// order 144, against 2 ms for rocSOLVER's reduction); orders above 256 are
// left to rocSOLVER.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <map>
#include <mutex>

#include "cplx.hpp"
#include "tridiag.hpp"

#include "common.hpp"

namespace half {

namespace {

constexpr int kTrdThreads = 512;
constexpr int kTrdWaves = kTrdThreads / 64;
constexpr int kTrdGroup = 4;  // columns per load group in the pass

__device__ inline double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// sum of the per-wave values red[0..kTrdWaves) in a fixed order
__device__ inline double wave_total(const double *red) {
  double s = 0;
#pragma unroll
  for (int k = 0; k < kTrdWaves; ++k) s += red[k];
  return s;
}
__device__ inline double wave_total4(const double *red) { return (red[0] + red[1]) + (red[2] + red[3]); }

template <int NMAX>
__global__ __launch_bounds__(kTrdThreads) void hetrd_upper_kernel(int n, double2 *A, int64_t lda, int64_t sA,
                                                                   double *D, int64_t sD, double *E, int64_t sE,
                                                                   double2 *Tau, int64_t sT, const int *only) {
  constexpr int RC = NMAX / 64;  // row chunks per lane
  static_assert(NMAX <= kTrdThreads && NMAX % 64 == 0, "one thread per row in the vector phases");
  __shared__ double2 vb[2][NMAX], wb[2][NMAX];  // v, w of this step and of the previous one
  __shared__ double2 part[kTrdWaves][NMAX];     // per-wave partial rows of A v
  __shared__ double2 nextcol[NMAX];             // column i of the block after the pass of step i
  __shared__ double red[kTrdWaves];             // per-wave partial v^H A v
  __shared__ double2 s_alpha, s_tau;

  if (only && only[blockIdx.x] == 0) return;  // the whole workgroup, before any barrier
  A += blockIdx.x * sA;
  D += blockIdx.x * sD;
  E += blockIdx.x * sE;
  Tau += blockIdx.x * sT;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  int cur = 0;
  bool prev = false;  // whether the previous step left a rank-2 update to apply

  for (int i = n - 2; i >= 0; --i) {
    const int c = i + 1, m = i + 1;
    const double2 *vp = vb[cur ^ 1], *wp = wb[cur ^ 1];
    double2 *v = vb[cur], *w = wb[cur];

    // wave 0: column c (rows 0..c) with the previous step's update, then
    // the reflector (zlarfg) that annihilates rows 0..i-1
    if (wid == 0) {
      double2 x[RC];
      double nrm = 0;
      const double2 cwc = prev ? wp[c] : mk<double2>(0, 0), cvc = prev ? vp[c] : mk<double2>(0, 0);
#pragma unroll
      for (int k = 0; k < RC; ++k) {
        const int r = lane + 64 * k;
        x[k] = mk<double2>(0, 0);
        if (r <= c) {
          x[k] = (i == n - 2) ? A[r + c * lda] : nextcol[r];
          if (prev) x[k] = csub(x[k], cadd(cmulc(vp[r], cwc), cmulc(wp[r], cvc)));
          if (r < i) nrm += x[k].x * x[k].x + x[k].y * x[k].y;
          if (r == i) s_alpha = x[k];
          if (r == c) D[c] = x[k].x;
        }
      }
      const double xn2 = wave_sum(nrm);
      const double2 alpha = s_alpha;
      const bool refl = !(xn2 == 0.0 && alpha.y == 0.0);
      double beta = alpha.x;
      double2 taui = mk<double2>(0, 0), scal = mk<double2>(1, 0);
      if (refl) {
        beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2), alpha.x);
        taui = mk<double2>((beta - alpha.x) / beta, -alpha.y / beta);
        const double zr = alpha.x - beta, zi = alpha.y, q = 1.0 / (zr * zr + zi * zi);
        scal = mk<double2>(zr * q, -zi * q);  // 1 / (alpha - beta)
      }
#pragma unroll
      for (int k = 0; k < RC; ++k) {
        const int r = lane + 64 * k;
        if (r < i) {
          const double2 y = cmul(x[k], scal);
          A[r + c * lda] = y;
          v[r] = y;
        } else if (r == i) {
          A[i + c * lda] = mk<double2>(beta, 0);
          v[i] = mk<double2>(1, 0);
        }
      }
      if (lane == 0) {
        E[i] = beta;
        Tau[i] = taui;
        s_tau = taui;
      }
    }
    __syncthreads();

    // one pass over A(0:i, 0:i): the previous update, then the partial A v.
    // A wave's columns j = wid + 8 s go in groups of kTrdGroup; the loads of
    // the next group are issued before the current one is updated and
    // stored (distinct columns, so the order is safe, but the compiler
    // cannot see that and would wait out the L2 latency once per column).
    // Loads are clamped into the block instead of masked, so they need no
    // branches; what a clamped lane computes is never stored or summed.
    double2 acc[RC], vpr[RC], wpr[RC];
#pragma unroll
    for (int k = 0; k < RC; ++k) {
      const int r = min(lane + 64 * k, m - 1);
      acc[k] = mk<double2>(0, 0);
      vpr[k] = prev ? vp[r] : mk<double2>(0, 0);
      wpr[k] = prev ? wp[r] : mk<double2>(0, 0);
    }
    const int ng = (m > wid) ? ((m - wid + kTrdWaves - 1) / kTrdWaves + kTrdGroup - 1) / kTrdGroup : 0;
    auto load = [&](int g, double2 (&b)[kTrdGroup][RC]) {
#pragma unroll
      for (int u = 0; u < kTrdGroup; ++u) {
        const int64_t j = min(wid + kTrdWaves * (g * kTrdGroup + u), m - 1);
#pragma unroll
        for (int k = 0; k < RC; ++k) b[u][k] = A[min(lane + 64 * k, m - 1) + j * lda];
      }
    };
    auto process = [&](int g, const double2 (&b)[kTrdGroup][RC]) {
#pragma unroll
      for (int u = 0; u < kTrdGroup; ++u) {
        const int j = wid + kTrdWaves * (g * kTrdGroup + u);
        if (j < m) {  // uniform over the wave
          const double2 vj = v[j];
          const double2 cw = prev ? wp[j] : mk<double2>(0, 0), cv = prev ? vp[j] : mk<double2>(0, 0);
#pragma unroll
          for (int k = 0; k < RC; ++k) {
            const int r = lane + 64 * k;
            const double2 a = csub(b[u][k], cadd(cmulc(vpr[k], cw), cmulc(wpr[k], cv)));
            if (r < m) {
              A[r + int64_t(j) * lda] = a;
              if (j == i) nextcol[r] = a;  // the next step's Householder column
            }
            acc[k] = cadd(acc[k], cmul(a, vj));
          }
        }
      }
    };
    {
      double2 ba[kTrdGroup][RC], bb[kTrdGroup][RC];
      if (ng > 0) load(0, ba);
      for (int g = 0; g < ng; g += 2) {
        if (g + 1 < ng) load(g + 1, bb);
        process(g, ba);
        if (g + 1 < ng) {
          if (g + 2 < ng) load(g + 2, ba);
          process(g + 1, bb);
        }
      }
    }
    double s = 0;  // this wave's part of v^H A v
#pragma unroll
    for (int k = 0; k < RC; ++k) {
      const int r = lane + 64 * k;
      if (r < m) {
        part[wid][r] = acc[k];
        const double2 vr = v[r];
        s += vr.x * acc[k].x + vr.y * acc[k].y;
      }
    }
    s = wave_sum(s);
    if (lane == 0) red[wid] = s;
    __syncthreads();

    // w = tau A v - tau/2 ((tau A v)^H v) v = tau A v - |tau|^2/2 (v^H A v) v
    // (v^H A v is real for Hermitian A)
    if (t < m) {
      double2 y = part[0][t];
#pragma unroll
      for (int k = 1; k < kTrdWaves; ++k) y = cadd(y, part[k][t]);
      const double2 taui = s_tau;
      const double a2 = -0.5 * (taui.x * taui.x + taui.y * taui.y) * wave_total(red);
      w[t] = cadd(cmul(taui, y), cscale(v[t], a2));
    }
    prev = !(s_tau.x == 0.0 && s_tau.y == 0.0);
    cur ^= 1;
    __syncthreads();
  }

  if (t == 0) {
    double2 a = A[0];
    if (prev) {
      const double2 v0 = vb[cur ^ 1][0], w0 = wb[cur ^ 1][0];
      a = csub(a, cadd(cmulc(v0, w0), cmulc(w0, v0)));
    }
    D[0] = a.x;
  }
}

// Orders above the one-workgroup kernel: the same reduction by G
// workgroups of one cooperative launch, the matrix held in their LDS.
// Workgroup g owns columns g, g + G, g + 2G, ... (all n rows each, so the
// shrinking active block stays spread over every workgroup), and by
// Hermitian symmetry forms (A v)_j = sum_r conj(A(r, j)) v(r) for its own
// columns without any exchange.  One grid barrier per column step:
//   A  every workgroup reads column c = i + 1 (published by its owner in the
//      previous step), applies the pending rank-2 update and forms the
//      reflector itself (the same operations on the same data everywhere,
//      so every copy of v is identical);
//   B  each workgroup applies the pending update to its active columns,
//      forms their entries of A v and its share of v^H A v, and the owner of
//      column i publishes that column for the next step;
//   -- grid barrier --
//   C  every workgroup reads A v and the G shares (summed in workgroup order)
//      and forms w; workgroup 0 stores the reflector, e(i) and tau(i).
// Published columns, A v and the shares alternate between two buffers by
// step parity: a workgroup can only reach step i - 1's writes after every
// workgroup has passed step i's barrier, i.e. finished reading step i + 1's.
// The grid barrier (cooperative groups) costs a few microseconds and grows
// with the workgroup count, so a column step takes about 10 us at order 340
// (16 workgroups) against about 18 us for rocSOLVER's chain of launches;
// dense solves of order 339 / 531 / 645 took 14.7 / 24.7 / 31.4 ms against
// 17.3 / 27.3 / 32.3 ms with rocSOLVER's reduction (one MI355X).
// LDS_COLS false: the columns stay in A (global memory, L2 / Infinity Cache
// resident at these orders) and each workgroup updates its own in place,
// for orders whose columns would need more workgroups than a cheap barrier
// allows.
constexpr int kCoopThreads = 256;

template <bool LDS_COLS>
__global__ __launch_bounds__(kCoopThreads) void hetrd_upper_coop_kernel(int n, double2 *A, int64_t lda, double *D,
                                                                        double *E, double2 *Tau, double2 *colbuf,
                                                                        double2 *pbuf, double *share, int ncol) {
  namespace cg = cooperative_groups;
  cg::grid_group grid = cg::this_grid();
  const int G = static_cast<int>(gridDim.x), g = static_cast<int>(blockIdx.x);
  extern __shared__ double2 coop_lds[];
  double2 *cols = coop_lds;                   // LDS_COLS: [ncol][n], column g + lc G at lc n
  double2 *vc = cols + (LDS_COLS ? static_cast<size_t>(ncol) * n : 0);  // this step's v
  double2 *vp = vc + n, *wp = vp + n;         // the pending update's v, w
  __shared__ double red[kCoopThreads / 64];
  __shared__ double2 s_alpha, s_tau;
  __shared__ double s_beta;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  if (LDS_COLS)
    for (int lc = 0; lc < ncol && g + lc * G < n; ++lc)
      for (int r = t; r < n; r += kCoopThreads)
        cols[static_cast<size_t>(lc) * n + r] = A[r + int64_t(g + lc * G) * lda];
  bool prev = false;
  for (int i = n - 2; i >= 0; --i) {
    const int c = i + 1, par = i & 1;
    // A: column c with the pending update, then the reflector (zlarfg)
    const double2 *src = (i == n - 2) ? A + int64_t(c) * lda : colbuf + (par ^ 1) * n;
    const double2 cwc = prev ? wp[c] : mk<double2>(0, 0), cvc = prev ? vp[c] : mk<double2>(0, 0);
    double nrm = 0;
    for (int r = t; r <= c; r += kCoopThreads) {
      double2 x = src[r];
      if (prev) x = csub(x, cadd(cmulc(vp[r], cwc), cmulc(wp[r], cvc)));
      vc[r] = x;
      if (r < i) nrm += x.x * x.x + x.y * x.y;
      if (r == i) s_alpha = x;
      if (r == c && g == 0) D[c] = x.x;
    }
    nrm = wave_sum(nrm);
    if (lane == 0) red[wid] = nrm;
    __syncthreads();
    {
      double xn2 = 0;
#pragma unroll
      for (int k = 0; k < kCoopThreads / 64; ++k) xn2 += red[k];
      const double2 alpha = s_alpha;
      const bool refl = !(xn2 == 0.0 && alpha.y == 0.0);
      double beta = alpha.x;
      double2 taui = mk<double2>(0, 0), scal = mk<double2>(1, 0);
      if (refl) {
        beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2), alpha.x);
        taui = mk<double2>((beta - alpha.x) / beta, -alpha.y / beta);
        const double zr = alpha.x - beta, zi = alpha.y, q = 1.0 / (zr * zr + zi * zi);
        scal = mk<double2>(zr * q, -zi * q);  // 1 / (alpha - beta)
      }
      for (int r = t; r <= i; r += kCoopThreads) vc[r] = r < i ? cmul(vc[r], scal) : mk<double2>(1, 0);
      if (t == 0) {
        s_tau = taui;
        s_beta = beta;
      }
    }
    __syncthreads();
    // B: own active columns j <= i (ascending in lc), one wave per column
    double sv = 0;  // this wave's part of v^H A v
    for (int lc = wid; lc < ncol; lc += kCoopThreads / 64) {
      const int j = g + lc * G;
      if (j > i) break;  // uniform over the wave
      double2 *col = LDS_COLS ? cols + static_cast<size_t>(lc) * n : A + int64_t(j) * lda;
      const double2 cw = prev ? wp[j] : mk<double2>(0, 0), cv = prev ? vp[j] : mk<double2>(0, 0);
      double2 acc = mk<double2>(0, 0);
      for (int r = lane; r <= i; r += 64) {
        double2 a = col[r];
        if (prev) {
          a = csub(a, cadd(cmulc(vp[r], cw), cmulc(wp[r], cv)));
          col[r] = a;
        }
        acc = cadd(acc, cmulc(vc[r], a));  // conj(A(r, j)) v(r)
        if (j == i) colbuf[par * n + r] = a;
      }
      acc.x = wave_sum(acc.x);
      acc.y = wave_sum(acc.y);
      if (lane == 0) pbuf[par * n + j] = acc;
      const double2 vj = vc[j];
      sv += vj.x * acc.x + vj.y * acc.y;
    }
    if (lane == 0) red[wid] = sv;
    __syncthreads();
    if (t == 0) share[par * G + g] = wave_total4(red);
    grid.sync();
    // C: w = tau A v - |tau|^2 / 2 (v^H A v) v; v, w become the pending update
    double vav = 0;
    for (int h = 0; h < G; ++h) vav += share[par * G + h];
    const double2 taui = s_tau;
    const double a2 = -0.5 * (taui.x * taui.x + taui.y * taui.y) * vav;
    for (int r = t; r <= i; r += kCoopThreads) {
      const double2 v = vc[r];
      wp[r] = cadd(cmul(taui, pbuf[par * n + r]), cscale(v, a2));
      vp[r] = v;
      if (g == 0) A[r + int64_t(c) * lda] = r < i ? v : mk<double2>(s_beta, 0);
    }
    if (t == 0 && g == 0) {
      E[i] = s_beta;
      Tau[i] = taui;
    }
    prev = !(taui.x == 0.0 && taui.y == 0.0);
    __syncthreads();
  }
  if (g == 0 && t == 0) {  // column 0 is workgroup 0's first column
    double2 a = LDS_COLS ? cols[0] : A[0];
    if (prev) a = csub(a, cadd(cmulc(vp[0], wp[0]), cmulc(wp[0], vp[0])));
    D[0] = a.x;
  }
}

}  // namespace

bool hetrd_upper_coop(int n, double2 *A, int64_t lda, double *d, double *e, double2 *tau, double2 *work,
                      size_t work_elems, hipStream_t s, bool force_global) {
  if (n <= 1 || lda < n) throw Error(kInvalidArgument, "hetrd_upper_coop: order or leading dimension");
  static std::mutex mu;
  static std::map<int, std::pair<int, int>> dev_info;  // device -> (CUs, LDS bytes per workgroup)
  int dev = 0;
  HALF_HIP(hipGetDevice(&dev));
  std::pair<int, int> info;
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = dev_info.find(dev);
    if (it == dev_info.end()) {
      int cus = 0, lds = 0;
      HALF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      HALF_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
      it = dev_info.emplace(dev, std::make_pair(cus, lds)).first;
    }
    info = it->second;
  }
  const int cus = info.first;
  const size_t lds_max = static_cast<size_t>(info.second) - 1024;  // the kernel's static shared data
  const size_t vec_bytes = 3 * sizeof(double2) * size_t(n);
  if (vec_bytes >= lds_max) return false;
  // fewest workgroups whose columns fit in LDS (fewer workgroups: a cheaper
  // grid barrier); above kCoopLdsMaxGroups of them the columns stay in A
  const int ncol_fit = static_cast<int>(std::min<size_t>(n, (lds_max - vec_bytes) / (sizeof(double2) * size_t(n))));
  const int G_lds = ncol_fit >= 1 ? std::max(kCoopMinGroups, (n + ncol_fit - 1) / ncol_fit) : cus + 1;
  const bool lds_cols = !force_global && G_lds <= kCoopLdsMaxGroups;
  const int G = lds_cols ? G_lds : std::min(kCoopLdsMaxGroups, cus);
  if (G > cus || G < 1) return false;
  const int ncol_used = (n + G - 1) / G;
  const size_t lds = vec_bytes + (lds_cols ? sizeof(double2) * size_t(ncol_used) * n : 0);
  if (work_elems < coop_work_elems(n, G)) throw Error(kInternal, "hetrd_upper_coop: workspace too small");
  const void *fn = lds_cols ? reinterpret_cast<const void *>(&hetrd_upper_coop_kernel<true>)
                            : reinterpret_cast<const void *>(&hetrd_upper_coop_kernel<false>);
  int per_cu = 0;
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)) != hipSuccess) {
    (void)hipGetLastError();  // this runtime cannot raise the limit: leave the order to the library
    return false;
  }
  HALF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kCoopThreads, lds));
  if (per_cu < 1 || G > per_cu * cus) return false;  // every workgroup must be resident at once
  double2 *colbuf = work, *pbuf = work + 2 * size_t(n);
  double *share = reinterpret_cast<double *>(work + 4 * size_t(n));
  int nn = n, nc = ncol_used;
  void *args[] = {&nn, &A, &lda, &d, &e, &tau, &colbuf, &pbuf, &share, &nc};
  HALF_HIP(hipLaunchCooperativeKernel(fn, dim3(G), dim3(kCoopThreads), args, static_cast<unsigned>(lds), s));
  return true;
}

size_t coop_work_elems(int n, int groups) { return 4 * size_t(n) + size_t(groups); }

void hetrd_upper(int n, int batch, double2 *A, int64_t lda, int64_t sA, double *d, int64_t sd, double *e,
                 int64_t se, double2 *tau, int64_t stau, hipStream_t s, const int *only) {
  if (n < 1 || n > kTridiagMaxOrder || lda < n || batch < 1)
    throw Error(kInvalidArgument, "hetrd_upper: order " + std::to_string(n) + " outside 1.." +
                                      std::to_string(kTridiagMaxOrder) + " or bad leading dimension / batch");
  const dim3 grid(batch), block(kTrdThreads);
  if (n <= 64)
    hipLaunchKernelGGL(hetrd_upper_kernel<64>, grid, block, 0, s, n, A, lda, sA, d, sd, e, se, tau, stau, only);
  else if (n <= 128)
    hipLaunchKernelGGL(hetrd_upper_kernel<128>, grid, block, 0, s, n, A, lda, sA, d, sd, e, se, tau, stau, only);
  else if (n <= 192)
    hipLaunchKernelGGL(hetrd_upper_kernel<192>, grid, block, 0, s, n, A, lda, sA, d, sd, e, se, tau, stau, only);
  else
    hipLaunchKernelGGL(hetrd_upper_kernel<256>, grid, block, 0, s, n, A, lda, sA, d, sd, e, se, tau, stau, only);
  HALF_HIP(hipGetLastError());
}

}  // namespace half
