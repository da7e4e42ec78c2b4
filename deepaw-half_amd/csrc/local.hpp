// Pruned local stage of the PAW operator (see local.hip).
#pragma once

#include <string>

#include "fft.hpp"

namespace half {

// Index tables of the plane-wave sphere on the FFT box at one k point.
struct LocalMap {
  int n1 = 0, n2 = 0, n3 = 0;
  int nlines = 0;   // (i1, i2) lines holding plane waves
  int nplanes = 0;  // i1 planes holding plane waves
  DeviceArray<int> line_i12, line_start, line_count, line_of, plane_i1, plane_on, pw_i3;
  void build(const Basis &b, hipStream_t s);
};

// hpsi = T psi + F[Veff F^-1[psi]]/N_r, spsi = psi for nb bands; boxes holds
// nb full boxes of workspace.
template <typename R>
void local_apply(const FftPlan &plan, const LocalMap &m, const R *veff, const R *kin, int nb, const cplx_t<R> *psi,
                 int64_t ldpsi, cplx_t<R> *hpsi, cplx_t<R> *spsi, int64_t ldout, cplx_t<R> *boxes, hipStream_t s);

// Tunables of the local stage, keys "local.{fp64|fp32}.tile.{x|y|z}" (FFT
// points per workgroup tile) and "local.{fp64|fp32}.xgroup" (bands sharing a
// Veff tile).  Return false for an unknown key or an out-of-range value.
bool set_local_tuning(const std::string &key, int value);
bool get_local_tuning(const std::string &key, int &value);

}  // namespace half
