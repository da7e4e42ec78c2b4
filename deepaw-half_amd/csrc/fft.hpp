// Batched 3D complex FFT on C-order boxes, written for gfx950.
//
// Each 1D pass stages a tile of M lines in LDS and runs mixed-radix Stockham
// autosort stages (radices 8, 6, 4, 5, 3, 2, generic 7) between two LDS
// buffers, so one pass is one HBM read and one HBM write of the tile.  Tiles
// of the strided axes take M lines that are adjacent in the contiguous
// direction, so every global access is a coalesced run.  The Veff multiply
// of the local operator is fused into the slowest-axis pass: inverse
// transform, multiply by Veff/N_r, forward transform, one round trip.
//
// Sign convention (reference: src/half_fft.F90:25-48): forward is
// exp(-i G.r), backward exp(+i G.r), both unnormalised.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace half {

template <typename R>
using cplx_t = typename std::conditional<std::is_same<R, double>::value, double2, float2>::type;

struct FftAxis {
  int n = 0;
  int nstages = 0;
  int radix[8] = {};
  DeviceArray<double2> twiddle;  // exp(-2 pi i k / n), k < n
};

struct FftPlan {
  int n[3] = {0, 0, 0};
  int64_t nr = 0;
  FftAxis axis[3];
  void init(const int shape[3], hipStream_t s);
  static bool supported(int n);
};

// Transform along one axis (0 = slowest, 2 = contiguous) of nbatch boxes
// stored back to back.  sign = -1 forward, +1 backward.
template <typename R>
void fft_axis(const FftPlan &plan, cplx_t<R> *data, int64_t nbatch, int axis, int sign,
              hipStream_t s);

// Along axis 0: backward transform, multiply by veff[r] * scale, forward
// transform.  This is the middle of the local-potential sandwich.
template <typename R>
void fft_axis0_veff(const FftPlan &plan, cplx_t<R> *data, int64_t nbatch, const R *veff,
                    R scale, hipStream_t s);

// Full unnormalised 3D transform of nbatch boxes.
template <typename R>
void fft3(const FftPlan &plan, cplx_t<R> *data, int64_t nbatch, int sign, hipStream_t s);

}  // namespace half
