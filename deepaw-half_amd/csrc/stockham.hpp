// Device core of the LDS-staged mixed-radix Stockham FFT: radix butterflies
// and the stage loop over M lines held in LDS (used by fft.hip for whole
// boxes and by local.hip for the pruned plane-wave transforms).
#pragma once

#include "cplx.hpp"
#include "fft.hpp"

namespace half {

struct FftArgs {
  int n;
  int nstages;
  int radix[8];
  const double2 *tw;
  long long lines_total;  // number of 1D lines in the batch
  int C;                  // element stride of the transform axis
  int M;                  // lines per workgroup
  int ld;                 // LDS line stride (n + 1: odd, so strided LDS accesses spread over banks)
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
    case 6: {
      // pairs at distance 3 (radix 2), then radix 3 on the sums (even
      // outputs) and on the twiddled differences (odd outputs):
      // X[2k] = DFT3(v_j + v_j+3)[k], X[2k+1] = DFT3((v_j - v_j+3) w6^j)[k]
      const S c = S(-0.5), s3 = sg * S(0.86602540378443864676372317075294);
      V a[3], b[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        a[j] = cadd(v[j], v[j + 3]);
        b[j] = csub(v[j], v[j + 3]);
      }
      b[1] = cmul(b[1], mk<V>(S(0.5), s3));  // w6 = exp(sign i pi / 3)
      b[2] = cmul(b[2], mk<V>(S(-0.5), s3));  // w6^2
      V x[3], y[3];
      {
        V a1 = cadd(a[1], a[2]), t = cadd(a[0], cscale(a1, c)), u = cmul_i(csub(a[1], a[2]), s3);
        x[0] = cadd(a[0], a1);
        x[1] = cadd(t, u);
        x[2] = csub(t, u);
      }
      {
        V a1 = cadd(b[1], b[2]), t = cadd(b[0], cscale(a1, c)), u = cmul_i(csub(b[1], b[2]), s3);
        y[0] = cadd(b[0], a1);
        y[1] = cadd(t, u);
        y[2] = csub(t, u);
      }
      v[0] = x[0];
      v[2] = x[1];
      v[4] = x[2];
      v[1] = y[0];
      v[3] = y[1];
      v[5] = y[2];
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
    case 7: {  // direct DFT with twiddles from the table (all indices compile-time)
      V y[7];
      const int step = n / 7;
#pragma unroll
      for (int q = 0; q < 7; ++q) {
        V acc = v[0];
#pragma unroll
        for (int r = 1; r < 7; ++r) acc = cadd(acc, cmul(v[r], twiddle(tws, ((r * q) % 7) * step, sign)));
        y[q] = acc;
      }
#pragma unroll
      for (int q = 0; q < 7; ++q) v[q] = y[q];
      break;
    }
    default:
      break;  // the plan only produces radices 8, 7, 6, 5, 4, 3, 2
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
      const V *src = in + l * a.ld + j;
#pragma unroll
      for (int r = 0; r < 8; ++r)
        if (r < R) v[r] = src[r * nr];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < 8; ++r)
          if (r < R) v[r] = cmul(v[r], twiddle(tws, k * r * tstep, sign));
      }
      butterfly(v, R, sign, tws, n);
      V *dst = out + l * a.ld + (j / Ns) * Ns * R + k;
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


// Compile-time form of run_stages for a transform length N known when the
// kernel is built: the same radix order as the runtime plan (fft.hip's
// factor: 8, 6, 4, 5, 3, 2, 7 greedily), so the arithmetic is identical, but
// every divisor and stride is a constant (the runtime form spends most of
// its instructions on integer division).  M (lines) must be below 2^16.
struct CtRadices {
  int count;
  int r[8];
};

constexpr CtRadices ct_factor(int n) {
  CtRadices p{0, {0, 0, 0, 0, 0, 0, 0, 0}};
  const int order[7] = {8, 6, 4, 5, 3, 2, 7};
  for (int i = 0; i < 7; ++i)
    while (n % order[i] == 0 && n > 1 && p.count < 8) {
      p.r[p.count++] = order[i];
      n /= order[i];
    }
  if (n != 1) p.count = -1;
  return p;
}

template <typename V, int N, int S, int NS>
__device__ inline V *ct_stages(V *in, V *out, const V *tws, int M, int ld, int sign) {
  constexpr CtRadices P = ct_factor(N);
  static_assert(P.count > 0, "transform length without a plan");
  if constexpr (S == P.count) {
    return in;
  } else {
    constexpr int R = P.r[S];
    constexpr int nr = N / R;
    constexpr int tstep = N / (NS * R);
    const int tasks = M * nr;
    // line-fastest task order: the lanes of a 16-lane LDS cycle take one
    // point of 16 consecutive lines, whose odd stride ld spreads them over
    // distinct banks, and share one twiddle (a broadcast).  j = t / M in
    // single precision is exact here: (t + 0.5) / M lies at least 0.5 / M
    // from an integer and t, M < 2^16.
    const float inv_m = 1.0f / static_cast<float>(M);
    for (int t = threadIdx.x; t < tasks; t += blockDim.x) {
      const int j = __float2int_rz((static_cast<float>(t) + 0.5f) * inv_m);
      const int l = t - j * M;
      const int k = j % NS;
      V v[8];
      const V *src = in + l * ld + j;
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = src[r * nr];
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twiddle(tws, k * r * tstep, sign));
      }
      butterfly(v, R, sign, tws, N);
      V *dst = out + l * ld + (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * NS] = v[r];
    }
    __syncthreads();
    return ct_stages<V, N, S + 1, NS * R>(out, in, tws, M, ld, sign);
  }
}

// Stages of length NN when NN > 0 (compile-time), else of a.n (runtime).
template <int NN, typename V>
__device__ inline V *stages(V *in, V *out, const V *tws, const FftArgs &a, int sign) {
  if constexpr (NN > 0)
    return ct_stages<V, NN, 0, 1>(in, out, tws, a.M, a.ld, sign);
  else
    return run_stages(in, out, tws, a, sign);
}

// Twiddle table of one axis into LDS (converted to the working precision).
template <typename V>
__device__ inline void load_twiddles(V *tws, const FftArgs &a) {
  using S = typename scalar_of<V>::type;
  for (int i = threadIdx.x; i < a.n; i += blockDim.x)
    tws[i] = mk<V>(static_cast<S>(a.tw[i].x), static_cast<S>(a.tw[i].y));
}

}  // namespace half
