// LDS-staged mixed-radix Stockham FFT passes (see fft.hpp).
#include <cmath>

#include "cplx.hpp"
#include "fft.hpp"
#include "stockham.hpp"
#include "trace.hpp"

namespace half {

namespace {

template <typename R, bool CONTIG, bool VEFF>
__global__ __launch_bounds__(256) void fft_lines_kernel(cplx_t<R> *data, FftArgs a, int sign,
                                                          const R *veff, R scale) {
  using V = cplx_t<R>;
  extern __shared__ double2 smem_raw[];
  V *buf0 = reinterpret_cast<V *>(smem_raw);
  V *buf1 = buf0 + a.M * a.ld;
  V *tws = buf1 + a.M * a.ld;
  for (int i = threadIdx.x; i < a.n; i += blockDim.x)
    tws[i] = mk<V>(static_cast<R>(a.tw[i].x), static_cast<R>(a.tw[i].y));
  const long long L0 = static_cast<long long>(blockIdx.x) * a.M;
  const int MN = a.M * a.n;
  const long long nC = static_cast<long long>(a.n) * a.C;
  for (int e = threadIdx.x; e < MN; e += blockDim.x) {
    int l, p;
    if (CONTIG) {
      l = e / a.n;
      p = e - l * a.n;
    } else {
      p = e / a.M;
      l = e - p * a.M;
    }
    const long long L = L0 + l;
    V v = mk<V>(R(0), R(0));
    if (L < a.lines_total) {
      const long long blk = L / a.C;
      const long long c = L - blk * a.C;
      v = data[blk * nC + c + static_cast<long long>(p) * a.C];
    }
    buf0[l * a.ld + p] = v;
  }
  __syncthreads();
  V *res = run_stages(buf0, buf1, tws, a, VEFF ? +1 : sign);
  if (VEFF) {
    for (int e = threadIdx.x; e < MN; e += blockDim.x) {
      const int l = e / a.n;
      const int p = e - l * a.n;
      const long long L = L0 + l;
      if (L < a.lines_total) {
        const long long c = L % a.C;
        res[l * a.ld + p] = cscale(res[l * a.ld + p], veff[c + static_cast<long long>(p) * a.C] * scale);
      }
    }
    __syncthreads();
    res = run_stages(res, res == buf0 ? buf1 : buf0, tws, a, -1);
  }
  for (int e = threadIdx.x; e < MN; e += blockDim.x) {
    int l, p;
    if (CONTIG) {
      l = e / a.n;
      p = e - l * a.n;
    } else {
      p = e / a.M;
      l = e - p * a.M;
    }
    const long long L = L0 + l;
    if (L < a.lines_total) {
      const long long blk = L / a.C;
      const long long c = L - blk * a.C;
      data[blk * nC + c + static_cast<long long>(p) * a.C] = res[l * a.ld + p];
    }
  }
}

bool factor(int n, int *radix, int &count) {
  count = 0;
  const int order[] = {8, 6, 4, 5, 3, 2, 7};
  for (int r : order)
    while (n % r == 0 && n > 1) {
      if (count == 8) return false;
      radix[count++] = r;
      n /= r;
    }
  return n == 1;
}

template <typename R>
FftArgs make_args(const FftPlan &plan, int axis, int64_t nbatch, int &M, size_t &lds) {
  const FftAxis &ax = plan.axis[axis];
  FftArgs a{};
  a.n = ax.n;
  a.nstages = ax.nstages;
  for (int i = 0; i < 8; ++i) a.radix[i] = ax.radix[i];
  a.tw = ax.twiddle.get();
  int64_t C = 1;
  for (int d = axis + 1; d < 3; ++d) C *= plan.n[d];
  a.C = static_cast<int>(C);
  a.lines_total = static_cast<long long>(nbatch * (plan.nr / ax.n));
  const int budget = std::is_same<R, double>::value ? 1536 : 3072;
  M = 1;
  while (M * 2 * ax.n <= budget && M < 64) M *= 2;
  a.M = M;
  a.ld = ax.n + 1;
  lds = (2 * size_t(M) * a.ld + ax.n) * sizeof(cplx_t<R>);
  return a;
}

}  // namespace

bool FftPlan::supported(int n) {
  int r[8], c;
  return n >= 2 && n <= 512 && factor(n, r, c);
}

void FftPlan::init(const int shape[3], hipStream_t s) {
  nr = 1;
  for (int d = 0; d < 3; ++d) {
    n[d] = shape[d];
    nr *= shape[d];
    FftAxis &ax = axis[d];
    ax.n = shape[d];
    if (!factor(ax.n, ax.radix, ax.nstages) || ax.n > 512)
      throw Error(kUnavailable, "FFT box dimension " + std::to_string(ax.n) +
                                    " is not a product of 2, 3, 5, 7 up to 512");
    std::vector<double2> tw(ax.n);
    for (int k = 0; k < ax.n; ++k) {
      long double ang = -2.0L * 3.141592653589793238462643383279502884L * k / ax.n;
      tw[k].x = static_cast<double>(cosl(ang));
      tw[k].y = static_cast<double>(sinl(ang));
    }
    ax.twiddle.upload(tw, s);
  }
}

template <typename R>
void fft_axis(const FftPlan &plan, cplx_t<R> *data, int64_t nbatch, int axis, int sign,
              hipStream_t s) {
  int M;
  size_t lds;
  FftArgs a = make_args<R>(plan, axis, nbatch, M, lds);
  const long long blocks = (a.lines_total + M - 1) / M;
  static const char *names[3][2] = {{"fft_axis0_fwd", "fft_axis0_bwd"}, {"fft_axis1_fwd", "fft_axis1_bwd"},
                                    {"fft_axis2_fwd", "fft_axis2_bwd"}};
  trace_mark(names[axis][sign > 0], s);
  if (axis == 2)
    hipLaunchKernelGGL((fft_lines_kernel<R, true, false>), dim3(blocks), dim3(256), lds, s, data, a,
                       sign, static_cast<const R *>(nullptr), R(1));
  else
    hipLaunchKernelGGL((fft_lines_kernel<R, false, false>), dim3(blocks), dim3(256), lds, s, data, a,
                       sign, static_cast<const R *>(nullptr), R(1));
  HALF_HIP(hipGetLastError());
}

template <typename R>
void fft_axis0_veff(const FftPlan &plan, cplx_t<R> *data, int64_t nbatch, const R *veff, R scale,
                    hipStream_t s) {
  int M;
  size_t lds;
  FftArgs a = make_args<R>(plan, 0, nbatch, M, lds);
  const long long blocks = (a.lines_total + M - 1) / M;
  trace_mark("fft_axis0_veff", s);
  hipLaunchKernelGGL((fft_lines_kernel<R, false, true>), dim3(blocks), dim3(256), lds, s, data, a, 1,
                     veff, scale);
  HALF_HIP(hipGetLastError());
}

template <typename R>
void fft3(const FftPlan &plan, cplx_t<R> *data, int64_t nbatch, int sign, hipStream_t s) {
  for (int axis = 2; axis >= 0; --axis) fft_axis<R>(plan, data, nbatch, axis, sign, s);
}

template void fft_axis<double>(const FftPlan &, double2 *, int64_t, int, int, hipStream_t);
template void fft_axis<float>(const FftPlan &, float2 *, int64_t, int, int, hipStream_t);
template void fft_axis0_veff<double>(const FftPlan &, double2 *, int64_t, const double *, double,
                                     hipStream_t);
template void fft_axis0_veff<float>(const FftPlan &, float2 *, int64_t, const float *, float,
                                    hipStream_t);
template void fft3<double>(const FftPlan &, double2 *, int64_t, int, hipStream_t);
template void fft3<float>(const FftPlan &, float2 *, int64_t, int, hipStream_t);

}  // namespace half
