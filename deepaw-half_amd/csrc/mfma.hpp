// Matrix-core helpers for gfx950: 16x16x4 MFMA in FP64 and FP32 and complex
// tile products built from four real MFMAs.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace half {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename R>
struct Acc;
template <>
struct Acc<double> {
  using type = v4d;
  __device__ static inline v4d mfma(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Acc<float> {
  using type = v4f;
  __device__ static inline v4f mfma(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// Operand layout of both instructions: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] of the 16x4 and 4x16 operands.  The result column of
// accumulator element r is l % 16; its row differs: the FP32 instruction keeps
// four consecutive rows per lane group (row = 4 (l / 16) + r), the FP64 one
// interleaves them (row = 4 r + l / 16).
template <typename R>
__device__ inline int mfma_row(int lane, int r) {
  return std::is_same<R, double>::value ? 4 * r + (lane >> 4) : 4 * (lane >> 4) + r;
}

// Complex 16x16 tile accumulate: (cr, ci) += A(16x4) * B(4x16), A and B
// given as this lane's real/imaginary parts.
template <typename R>
__device__ inline void cmfma(R ar, R ai, R br, R bi, typename Acc<R>::type &cr, typename Acc<R>::type &ci) {
  cr = Acc<R>::mfma(ar, br, cr);
  cr = Acc<R>::mfma(-ai, bi, cr);
  ci = Acc<R>::mfma(ar, bi, ci);
  ci = Acc<R>::mfma(ai, br, ci);
}

// Complex multiply-accumulate into the accumulators of one 16x16 tile:
// PARTS = 3 uses three real MFMAs (Gauss): p0 += ar br, p1 += ai bi,
// p2 += (ar + ai)(br + bi), with the product's real part p0 - p1 and its
// imaginary part p2 - p0 - p1; PARTS = 2 the four-MFMA form with the real
// and imaginary parts accumulated directly.  The default: three for FP32
// (measured faster for the FP32 contractions), four for FP64, whose smaller
// accumulator set holds more waves per SIMD in the widest tiles.
template <typename R>
constexpr int kCmacParts = sizeof(R) == 4 ? 3 : 2;

template <typename R, int PARTS>
__device__ inline void cmac(R ar, R ai, R br, R bi, typename Acc<R>::type (&p)[PARTS]) {
  static_assert(PARTS == 2 || PARTS == 3, "two or three accumulators per tile");
  if constexpr (PARTS == 3) {
    p[0] = Acc<R>::mfma(ar, br, p[0]);
    p[1] = Acc<R>::mfma(ai, bi, p[1]);
    p[2] = Acc<R>::mfma(ar + ai, br + bi, p[2]);
  } else {
    cmfma<R>(ar, ai, br, bi, p[0], p[1]);
  }
}

// element i of the tile accumulated by cmac
template <typename R, int PARTS>
__device__ inline void cmac_result(const typename Acc<R>::type (&p)[PARTS], int i, R &re, R &im) {
  if constexpr (PARTS == 3) {
    re = p[0][i] - p[1][i];
    im = p[2][i] - p[0][i] - p[1][i];
  } else {
    re = p[0][i];
    im = p[1][i];
  }
}

}  // namespace half
