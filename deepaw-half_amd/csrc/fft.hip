// LDS-staged mixed-radix Stockham FFT passes (see fft.hpp).
#include <cmath>

#include "cplx.hpp"
#include "fft.hpp"

namespace half {

namespace {

struct FftArgs {
  int n;
  int nstages;
  int radix[8];
  const double2 *tw;
  long long lines_total;  // number of 1D lines in the batch
  int C;                  // element stride of the transform axis
  int M;                  // lines per workgroup
};

template <typename V>
__device__ inline V twiddle(const V *tws, int q, int sign) {
  V w = tws[q];
  if (sign > 0) w.y = -w.y;
  return w;
}

// In-register DFT of length R with exponent sign `sign`.
template <typename V>
__device__ inline void butterfly(V *v, int R, int sign, const V *tws, int n) {
  using S = typename scalar_of<V>::type;
  const S sg = static_cast<S>(sign);
  switch (R) {
    case 2: {
      V a = v[0], b = v[1];
      v[0] = cadd(a, b);
      v[1] = csub(a, b);
      break;
    }
    case 3: {
      const S c = S(-0.5), s = sg * S(0.86602540378443864676372317075294);
      V a1 = cadd(v[1], v[2]);
      V t = cadd(v[0], cscale(a1, c));
      V u = cmul_i(csub(v[1], v[2]), s);
      v[0] = cadd(v[0], a1);
      v[1] = cadd(t, u);
      v[2] = csub(t, u);
      break;
    }
    case 4: {
      V t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
      V t2 = cadd(v[1], v[3]), t3 = cmul_i(csub(v[1], v[3]), sg);
      v[0] = cadd(t0, t2);
      v[1] = cadd(t1, t3);
      v[2] = csub(t0, t2);
      v[3] = csub(t1, t3);
      break;
    }
    case 5: {
      const S c1 = S(0.30901699437494742410229341718282);
      const S c2 = S(-0.80901699437494742410229341718282);
      const S s1 = sg * S(0.95105651629515357211643933337938);
      const S s2 = sg * S(0.58778525229247312916870595463907);
      V a1 = cadd(v[1], v[4]), a2 = cadd(v[2], v[3]);
      V b1 = csub(v[1], v[4]), b2 = csub(v[2], v[3]);
      V r1 = cadd(v[0], cadd(cscale(a1, c1), cscale(a2, c2)));
      V r2 = cadd(v[0], cadd(cscale(a1, c2), cscale(a2, c1)));
      V i1 = cmul_i(cadd(cscale(b1, s1), cscale(b2, s2)), S(1));
      V i2 = cmul_i(csub(cscale(b1, s2), cscale(b2, s1)), S(1));
      v[0] = cadd(v[0], cadd(a1, a2));
      v[1] = cadd(r1, i1);
      v[4] = csub(r1, i1);
      v[2] = cadd(r2, i2);
      v[3] = csub(r2, i2);
      break;
    }
    case 8: {
      const S h = S(0.70710678118654752440084436210485);
      // radix-2 split into two radix-4 halves
      V e0 = cadd(v[0], v[4]), e1 = csub(v[0], v[4]);
      V e2 = cadd(v[2], v[6]), e3 = cmul_i(csub(v[2], v[6]), sg);
      V E0 = cadd(e0, e2), E2 = csub(e0, e2), E1 = cadd(e1, e3), E3 = csub(e1, e3);
      V o0 = cadd(v[1], v[5]), o1 = csub(v[1], v[5]);
      V o2 = cadd(v[3], v[7]), o3 = cmul_i(csub(v[3], v[7]), sg);
      V O0 = cadd(o0, o2), O2 = csub(o0, o2), O1 = cadd(o1, o3), O3 = csub(o1, o3);
      // twiddles w8^s = exp(sign i pi s / 4)
      V w1 = mk<V>(h, sg * h), w3 = mk<V>(-h, sg * h);
      O1 = cmul(O1, w1);
      O2 = cmul_i(O2, sg);
      O3 = cmul(O3, w3);
      v[0] = cadd(E0, O0);
      v[4] = csub(E0, O0);
      v[1] = cadd(E1, O1);
      v[5] = csub(E1, O1);
      v[2] = cadd(E2, O2);
      v[6] = csub(E2, O2);
      v[3] = cadd(E3, O3);
      v[7] = csub(E3, O3);
      break;
    }
    default: {  // generic small radix (7): direct DFT from the twiddle table
      V y[8];
      const int step = n / R;
      for (int s = 0; s < R; ++s) {
        V acc = v[0];
        for (int r = 1; r < R; ++r) acc = cadd(acc, cmul(v[r], twiddle(tws, (r * s % R) * step, sign)));
        y[s] = acc;
      }
      for (int s = 0; s < R; ++s) v[s] = y[s];
    }
  }
}

// Stockham stages on M lines held in LDS; returns the buffer with the result.
template <typename V>
__device__ V *run_stages(V *in, V *out, const V *tws, const FftArgs &a, int sign) {
  const int n = a.n;
  int Ns = 1;
  for (int s = 0; s < a.nstages; ++s) {
    const int R = a.radix[s];
    const int nr = n / R;
    const int tasks = a.M * nr;
    const int tstep = n / (Ns * R);
    for (int t = threadIdx.x; t < tasks; t += blockDim.x) {
      const int l = t / nr;
      const int j = t - l * nr;
      const int k = j % Ns;
      V v[8];
      const V *src = in + l * n + j;
#pragma unroll
      for (int r = 0; r < 8; ++r)
        if (r < R) v[r] = src[r * nr];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < 8; ++r)
          if (r < R) v[r] = cmul(v[r], twiddle(tws, k * r * tstep, sign));
      }
      butterfly(v, R, sign, tws, n);
      V *dst = out + l * n + (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < 8; ++r)
        if (r < R) dst[r * Ns] = v[r];
    }
    __syncthreads();
    V *tmp = in;
    in = out;
    out = tmp;
    Ns *= R;
  }
  return in;
}

template <typename R, bool CONTIG, bool VEFF>
__global__ __launch_bounds__(256) void fft_lines_kernel(cplx_t<R> *data, FftArgs a, int sign,
                                                          const R *veff, R scale) {
  using V = cplx_t<R>;
  extern __shared__ double2 smem_raw[];
  V *buf0 = reinterpret_cast<V *>(smem_raw);
  V *buf1 = buf0 + a.M * a.n;
  V *tws = buf1 + a.M * a.n;
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
    buf0[l * a.n + p] = v;
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
        res[e] = cscale(res[e], veff[c + static_cast<long long>(p) * a.C] * scale);
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
      data[blk * nC + c + static_cast<long long>(p) * a.C] = res[l * a.n + p];
    }
  }
}

bool factor(int n, int *radix, int &count) {
  count = 0;
  const int order[] = {8, 4, 5, 3, 2, 7};
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
  lds = (2 * size_t(M) * ax.n + ax.n) * sizeof(cplx_t<R>);
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
