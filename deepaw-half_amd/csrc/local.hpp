// Pruned local stage of the PAW operator (see local.hip).
#pragma once

#include <string>

#include "fft.hpp"

namespace half {

// Index tables of one k point's plane-wave sphere on the FFT box, as the
// kernels read them (device pointers into a LocalMap).
struct LocalTables {
  int n1, n2, n3;
  long long nr;
  int nlines;  // (i1, i2) lines holding plane waves
  const int *line_i12, *line_start, *line_count, *line_of;
  int nplanes;  // i1 planes holding plane waves
  const int *plane_i1, *plane_on;
  const int *pw_i3;
  int plane_s0;  // first signed plane of the contiguous run
};

// Index tables of the plane-wave sphere on the FFT box at one k point.
struct LocalMap {
  int n1 = 0, n2 = 0, n3 = 0;
  int nlines = 0;   // (i1, i2) lines holding plane waves
  int nplanes = 0;  // i1 planes holding plane waves
  // the active planes as signed indices form one run plane_s0 .. plane_s0 +
  // nplanes - 1 (modulo n1) when planes_contiguous (true for a sphere)
  int plane_s0 = 0;
  bool planes_contiguous = false;
  DeviceArray<int> line_i12, line_start, line_count, line_of, plane_i1, plane_on, pw_i3;
  DeviceArray<LocalTables> dtab;  // one-entry device copy of tables()
  void build(const Basis &b, hipStream_t s);
  LocalTables tables() const;
};

// A batch of k points sharing the FFT box: k point j's tables are tabs[j]
// (device array), and its states, outputs, kinetic energies and workspace
// boxes sit at base + j * stride of the arrays passed to local_apply_batch.
struct LocalBatch {
  int nk = 1;
  const LocalTables *tabs = nullptr;
  int max_lines = 0, max_planes = 0;
  bool contiguous = false;  // every k point's active planes form one run
  long long psi_k = 0, out_k = 0, kin_k = 0, box_k = 0;
};

// Terms added to H psi and S psi as the local stage writes them: band b of
// k point j at h + j kstride + b ld (likewise s); the stream waits on
// `wait` (if set) before the final pass reads them.
template <typename R>
struct LocalAddend {
  const cplx_t<R> *h = nullptr, *s = nullptr;
  long long ld = 0, kstride = 0;
  hipEvent_t wait = nullptr;
};

// hpsi = T psi + F[Veff F^-1[psi]]/N_r, spsi = psi for nb bands of every k
// point of the batch; boxes holds nb full boxes per k point.  vx is the
// i1-axis DFT of Veff/N_r in both layouts from build_veff_x_dft (2 N_r
// values; null: transform-based x pass).
template <typename R>
void local_apply_batch(const FftPlan &plan, const LocalBatch &kb, const R *veff, const cplx_t<R> *vx, const R *kin,
                       int nb, const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi, cplx_t<R> *spsi,
                       int64_t ldout, cplx_t<R> *boxes, hipStream_t s, const LocalAddend<R> &add = {});

// One k point.
template <typename R>
void local_apply(const FftPlan &plan, const LocalMap &m, const R *veff, const cplx_t<R> *vx, const R *kin, int nb,
                 const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi, cplx_t<R> *spsi, int64_t ldout,
                 cplx_t<R> *boxes, hipStream_t s, const LocalAddend<R> &add = {});

// vx[d][c] = sum_x veff[x][c] e^{-2 pi i d x / n1} / (n1 n2 n3) for the
// n2 n3 columns c = i2 n3 + i3 of a C-order n1 x n2 x n3 box, FP64 and FP32
// copies, each 2 n1 n2 n3 long: [d][c] first, then [i3][d][i2].
void build_veff_x_dft(const double *veff, int n1, int n2, int n3, double2 *vx, float2 *vx32, hipStream_t s);

// Tunables of the local stage, keys "local.{fp64|fp32}.tile.{x|y|z}" (FFT
// points per workgroup tile), "local.{fp64|fp32}.xgroup" (bands sharing a
// Veff tile in the transform-based x pass), "local.{fp64|fp32}.xconv"
// (1: x pass as a per-column convolution where the active planes fit, 0:
// transform-based), "local.{fp64|fp32}.fused" (1: y, x and y passes in one
// kernel per (band, i3) slice where the convolution applies, 0: separate
// passes), "local.{fp64|fp32}.ygroup" (bands per workgroup of the fused
// kernel) and "local.{fp64|fp32}.vxlds" (1: the fused kernel stages the
// slice's Vx rows in LDS where its band group is two or more, 0: it reads
// them from L2).  Return false for an unknown key or an out-of-range value.
bool set_local_tuning(const std::string &key, int value);
bool get_local_tuning(const std::string &key, int &value);

}  // namespace half
