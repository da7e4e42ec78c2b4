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
  // the active planes as signed indices form one run plane_s0 .. plane_s0 +
  // nplanes - 1 (modulo n1) when planes_contiguous (true for a sphere)
  int plane_s0 = 0;
  bool planes_contiguous = false;
  // lines of each active plane in signed order q (plane q holds lines
  // qplane_line0[q] .. + qplane_nline[q] - 1); valid when lines_by_plane
  bool lines_by_plane = false;
  DeviceArray<int> line_i12, line_start, line_count, line_of, plane_i1, plane_on, pw_i3;
  DeviceArray<int> qplane_line0, qplane_nline;
  void build(const Basis &b, hipStream_t s);
};

// hpsi = T psi + F[Veff F^-1[psi]]/N_r, spsi = psi for nb bands; boxes holds
// nb full boxes of workspace.  vx is the i1-axis DFT of Veff/N_r from
// build_veff_x_dft (null: transform-based x pass).
template <typename R>
void local_apply(const FftPlan &plan, const LocalMap &m, const R *veff, const cplx_t<R> *vx, const R *kin, int nb,
                 const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi, cplx_t<R> *spsi, int64_t ldout,
                 cplx_t<R> *boxes, hipStream_t s);

// vx[d][c] = sum_x veff[x][c] e^{-2 pi i d x / n1} / (n1 n2 n3) for the
// n2 n3 columns c of a C-order n1 x n2 x n3 box, FP64 and FP32 copies.
void build_veff_x_dft(const double *veff, int n1, int n2, int n3, double2 *vx, float2 *vx32, hipStream_t s);

// Tunables of the local stage, keys "local.{fp64|fp32}.tile.{x|y|z}" (FFT
// points per workgroup tile), "local.{fp64|fp32}.xgroup" (bands sharing a
// Veff tile in the transform-based x pass), "local.{fp64|fp32}.xconv"
// (1: x pass as a per-column convolution where the active planes fit, 0:
// transform-based) and "local.{fp64|fp32}.plane" (0: separate z and y
// passes over full boxes; M > 0: z and y fused per active plane with M
// lines per LDS transform batch, planes stored compactly).  Return false for
// an unknown key or an out-of-range value.
bool set_local_tuning(const std::string &key, int value);
bool get_local_tuning(const std::string &key, int &value);

}  // namespace half
