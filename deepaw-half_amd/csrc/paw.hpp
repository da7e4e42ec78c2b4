// Device-resident PAW system and per-k-point operator.
//
// System  : built once per context.  Veff on the C-order box (FP64 master
//           plus an FP32 copy), Veff(G)/N_r for dense assembly, the FFT plan,
//           radial projector tables with their spline second derivatives,
//           Cartesian atom positions and the lm-expanded D/Q blocks.
// KOperator<R>: built once per (context, k point, precision).  Basis index
//           map and kinetic energies, the projector matrix
//           B (npw x nproj, column-major, column = atom-major (atom, lm)),
//           per-atom D and per-species Q blocks, FFT workspace.
#pragma once

#include <memory>

#include "common.hpp"
#include "fft.hpp"
#include "local.hpp"
#include "projector.hpp"

namespace half {

struct System {
  Crystal crystal;
  ChargeGrid charge;
  std::vector<Dataset> sets;
  bool use_pbe = true;
  double encut = 0;
  hipStream_t stream = nullptr;

  FftPlan fft;
  DeviceArray<double> veff;    // C-order, eV
  DeviceArray<float> veff32;
  DeviceArray<double2> veff_g;  // forward FFT(veff) / N_r
  // i1-axis DFT of veff / N_r per (i2,i3) column: N_r values as [d][column],
  // then the same N_r values as [i3][d][i2] (build_veff_x_dft)
  DeviceArray<double2> veff_x;
  DeviceArray<float2> veff_x32;
  double e_hartree = 0, e_xc = 0, e_xc_potential = 0;
  bool ready = false;

  // projector description (global projector index p = atom-major (atom, lm))
  int nproj = 0;
  int natoms = 0;
  std::vector<int> proj_atom, proj_lm, proj_l, proj_m, proj_table;  // per projector
  std::vector<double> proj_gmax;
  std::vector<int> atom_first, atom_nlm, atom_species;              // per atom
  std::vector<double> tables, tables_m2;                           // 100 points per channel
  std::vector<double> positions_cart;                              // natoms x 3
  std::vector<double> dblk, qblk;   // per atom nlm x nlm row-major (D and Q)
  bool use_qdep = false;            // D of each atom corrected for Veff (QDEP)
  std::vector<double> ddblk;        // with QDEP: dD/dR per atom, [alpha][a][b] at 3 x block offset
  DeviceArray<double> d_ddblk;
  std::vector<int> atom_boff;       // offset of an atom's block in dblk/qblk
  DeviceArray<int> d_proj_atom, d_proj_lm, d_proj_l, d_proj_m, d_proj_table;
  DeviceArray<double> d_proj_gmax, d_tables, d_tables_m2, d_positions;
  DeviceArray<int> d_atom_first, d_atom_nlm, d_atom_boff;
  DeviceArray<double> d_dblk, d_qblk;
  DeviceArray<float> d_dblk32, d_qblk32;

  void build(hipStream_t s);  // Veff + tables; reads crystal/charge/sets
  int max_nlm() const;
};

// A second stream (default priority) and a fork/join event pair, created on
// first use: the projector stage runs there while the local stage runs on
// the caller's stream.  Destroying it drains the stream first.
struct AuxStream {
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  AuxStream() = default;
  AuxStream(const AuxStream &) = delete;
  AuxStream &operator=(const AuxStream &) = delete;
  AuxStream(AuxStream &&o) noexcept : stream(o.stream), fork(o.fork), join(o.join) {
    o.stream = nullptr;
    o.fork = o.join = nullptr;
  }
  ~AuxStream() { release(); }
  void ensure() {
    if (stream) return;
    HALF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HALF_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    HALF_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
  }
  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    fork = join = nullptr;
  }
};

// Whether H.psi overlaps its projector stage (second stream) with the local
// stage; off, everything runs in order on the caller's stream.  Key
// "operator.overlap" (0 / 1) through half_amd_set_tuning.  Key
// "operator.projector" (0 / 1): the projector stage on the skinny kernels of
// projector.hip (default) or on the generic contractions of contract.hip
// with a separate D / Q mix.
bool operator_overlap();
bool set_operator_tuning(const std::string &key, int value);
bool get_operator_tuning(const std::string &key, int &value);

template <typename R>
struct KOperator {
  const System *sys = nullptr;
  Basis basis;
  int n = 0;       // plane waves
  int np = 0;      // projectors
  int64_t ldb = 0;  // leading dimension of B
  DeviceArray<int32_t> fft_index;
  DeviceArray<R> kinetic;
  DeviceArray<double> qvec;  // npw x 3 (for the force adjoint)
  DeviceArray<cplx_t<R>> B;
  DeviceArray<cplx_t<R>> BD, BQ;  // B D and B Q (per-atom blocks), leading dimension ldb; formed on first use
  LocalMap lmap;  // plane-wave sphere on the box (pruned transforms)
  // workspace
  int box_batch = 0;  // bands per FFT launch; 0 = lines and planes within 256 MiB (>= 16), boxes within 2 GiB
  DeviceArray<cplx_t<R>> boxes;
  DeviceArray<cplx_t<R>> coeff, dcoeff, qcoeff, partial;
  CoeffWork<R> cwork;
  DeviceArray<cplx_t<R>> nl_h, nl_s;  // projector-stage results of the overlapped apply
  AuxStream aux;

  void build(const System &s, const double kpoint[3], hipStream_t st);
  // H psi and S psi for nb states (column-major, leading dimensions ld*).
  void apply(int nb, const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi, cplx_t<R> *spsi,
             int64_t ldout, hipStream_t st);
  // Stages, exposed for the local-only / projector-only checks and timing.
  void apply_local(int nb, const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi, cplx_t<R> *spsi,
                   int64_t ldout, hipStream_t st);
  // local stage that adds addh / adds (layout of hpsi / spsi; null: none)
  // after waiting on `wait` (may be null)
  void apply_local_added(int nb, const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi, cplx_t<R> *spsi,
                         int64_t ldout, hipStream_t st, const cplx_t<R> *addh, const cplx_t<R> *adds,
                         hipEvent_t wait);
  // B D B^H psi and B Q B^H psi written (not added) to hout / sout
  void projector_terms(int nb, const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hout, cplx_t<R> *sout,
                       int64_t ldout, hipStream_t st);
  void apply_projectors(int nb, const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi,
                        cplx_t<R> *spsi, int64_t ldout, hipStream_t st);
  // c = B^H psi  (np x nb, leading dimension np) into `coeff`
  void project(int nb, const cplx_t<R> *psi, int64_t ldpsi, hipStream_t st);
  // BD and BQ, once
  void mixed_blocks(hipStream_t st);
  // H psi and S psi in order on st with every temporary (FFT boxes,
  // projector coefficients, contraction partials) in the caller's `work`
  // of `capacity` bytes, bands in batches that fit it: no device
  // allocation.  kCapacity if one band does not fit.
  void apply_external(int nb, const cplx_t<R> *psi, int64_t ldpsi, cplx_t<R> *hpsi, cplx_t<R> *spsi,
                      int64_t ldout, hipStream_t st, void *work, size_t capacity);
  size_t external_bytes(int bands) const;              // workspace of one batch of `bands`
  int external_batch(int nb, size_t capacity) const;   // largest batch within capacity (0: none)
  // projector stage on the skinny kernels for nb bands
  bool skinny(int nb) const;
};

// Sizes up to which the projector stage uses the skinny kernels.
constexpr int kSkinnyMaxProjectors = 512;
constexpr int kSkinnyMaxBands = 256;
// Whether np projectors and nb bands take the skinny kernels (and the
// "operator.projector" key allows them).
bool projector_skinny(int np, int nb);

// dc = D c, qc = Q c per atom block for ncols coefficient columns of length
// np (column-major, leading dimension np), in the operator's precision.
template <typename R>
void mix_projector_coefficients(const System &sys, int ncols, const cplx_t<R> *c, cplx_t<R> *dc, cplx_t<R> *qc,
                                hipStream_t st);

// Dense H/S assembly on the device (FP64), column-major n x n with
// leading dimension ld, hermitised (reference: src/half_dense_solver.F90:53-88).
void assemble_dense(KOperator<double> &op, double2 *h, double2 *s, int64_t ld, hipStream_t st);

// Device Veff construction (reference: src/half_potential.F90:175-238).
void build_veff_device(System &sys, hipStream_t s);

// V_xc (eV, C order) of a C-order charge (cell volume x density, device)
// plus the datasets' partial core (reference: the vxc_out of
// src/half_potential.F90:31-94).
void xc_potential_device(const System &sys, const double *charge_c, double *vxc, hipStream_t s);

// f_G = FFT(f) / N_r of a real C-order field (device).
void field_transform(const System &sys, const double *field_c, double2 *field_g, hipStream_t s);

// F_I = -sum_G q t(|G|) Im[conj(f_G) e^{-i q.R_I}] for every atom (natoms x 3,
// written): t the local part over the cell volume (core = false; f = the
// output density) or the partial core table (core = true; f = V_xc or the
// XC response) (reference: src/half_local_forces.F90:11-98, 160-190).
void table_forces(const System &sys, const double2 *field_g, bool core, std::vector<double> &forces,
                  hipStream_t s);

// Conversions between precisions of state blocks.
void convert_states(const double2 *in, int64_t ldin, float2 *out, int64_t ldout, int n, int nb,
                    hipStream_t s);
void convert_states(const float2 *in, int64_t ldin, double2 *out, int64_t ldout, int n, int nb,
                    hipStream_t s);

}  // namespace half
