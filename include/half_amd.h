/* Device-level extension of the HALF C ABI for the MI355X backend.
 *
 * The reference ABI (half.h) exchanges host buffers only.  These entry
 * points expose the device-resident operator and solver controls that the
 * reference keeps internal, for hosts that already hold states in HBM and for
 * measurement.  All pointers named d_* are device pointers on the current
 * HIP device; complex data are interleaved (re, im) pairs of the operator's
 * precision (complex128 for precision 64, complex64 for precision 32);
 * matrices are column-major.  Every function returns a half_status code and
 * fills the caller-owned error buffer like the reference ABI.
 */
#ifndef DEEPAW_HALF_AMD_H
#define DEEPAW_HALF_AMD_H

#include <stdint.h>

#include "half.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t half_amd_operator;

/* Matrix-free operator at one k point, built on the device: basis index
 * map, kinetic energies, projector matrix B (npw x nproj) and workspace.
 * Counterpart of prepare_hs_operator_cuda (reference:
 * src/half_cuda_assembly.cuf:402-466).  precision is 64 or 32. */
int half_amd_operator_create(half_handle handle, const double kpoint[3], int precision,
                             half_amd_operator *op, int64_t *npw, int32_t *nproj, char *error,
                             int error_capacity);
/* H.psi and S.psi for nstates device columns (reference:
 * src/half_cuda_assembly.cuf:468-519).  stage: 0 full operator, 1 local
 * part only (T + Veff), 2 projector part only (B D B^H, B Q B^H added to
 * the outputs).  stream: a hipStream_t, or null for the context stream.
 * box_batch: bands per FFT batch (0 keeps the current value). */
int half_amd_operator_apply(half_amd_operator op, int stage, int nstates, const void *d_psi,
                            int64_t ldpsi, void *d_hpsi, void *d_spsi, int64_t ldout, int box_batch,
                            void *stream, char *error, int error_capacity);
int half_amd_operator_destroy(half_amd_operator op, char *error, int error_capacity);

/* H.psi and S.psi of half_amd_operator_apply (stage 0) with no device
 * allocation.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Every temporary (FFT boxes,
 * projector coefficients, contraction partials) lives in a 128 MiB array
 * the library's code object holds on each device until the process ends;
 * the bands are applied in batches that fit it, the two stages in order on
 * the stream.  HALF_ERROR_CAPACITY if one band of this operator needs more
 * than the array (an FFT box of 8 M points in FP64, for example).  All calls
 * on one device share that array: calls that overlap in time on one device
 * (other streams, other threads) race and must be serialised by the caller.
 * Results equal half_amd_operator_apply's to rounding (the projector stage
 * always takes the LDS-tiled contractions here). */
int half_amd_operator_apply_noalloc(half_amd_operator op, int nstates, const void *d_psi, int64_t ldpsi,
                                    void *d_hpsi, void *d_spsi, int64_t ldout, void *stream, char *error,
                                    int error_capacity);

/* Solver statistics of the last solve. */
typedef struct half_amd_solve_stats {
  int32_t iterations;
  int32_t fp64_iterations;
  double max_residual_eV;
  double seconds;
  int64_t band_applications;
  int32_t metric_fallbacks; /* iterations whose correction block was normalised
                               by the metric's eigen-decomposition instead of
                               its Cholesky factor */
  int32_t rr_order_mean;    /* mean order of the per-iteration Rayleigh-Ritz
                               problems, rounded (2 x bands without
                               locking, tuning key "solver.lock") */
} half_amd_solve_stats;

/* All-band solve with explicit controls (the reference ABI fixes tol 1e-4 eV,
 * 40 iterations, 16-band blocks: src/half_cuda_iterative_solver.cuf:113-119).
 * block: bands per FFT launch, 0 = as many as keep the pass-to-pass lines and
 * planes within 256 MiB (at least 16) and the boxes within 2 GiB.
 * mixed != 0 runs FP32 H.psi iterations, then FP64 H.psi sweeps to tol.  The
 * FP32 phase ends at the first of: max residual <= max(switch_tol (0: 2e-3),
 * tol, the FP32 operator's error measured on a probe block at the start);
 * no 10% gain in `solver.stall` iterations (default 4) once within 10x of
 * that target; half of maxiter.  Rayleigh-Ritz, Cholesky-type metric
 * normalisation and the subspace eigensolves are FP64 throughout.  Locking
 * (key "solver.lock"): 1 soft -- requested bands whose residual meets the
 * current phase's target add no correction column to the next Rayleigh-Ritz
 * step; 2 (default) soft plus a frozen prefix -- the lowest requested bands
 * that all meet it also leave the Rayleigh-Ritz problem (the corrections
 * stay S-orthogonal to them; the FP32 phase's frozen bands rejoin at the
 * switch to FP64), and one Rayleigh-Ritz step over all bands ends the solve;
 * 0 off.  The stopping rule still reads every requested band.
 * extra buffer bands are carried internally and not reported; with locking
 * on, only the lowest "solver.bufcols" percent of them (default 0) get
 * correction columns (all of them once no requested band needs one).
 * Rayleigh-Ritz steps whose subspace metric is the identity to within
 * 10^-"solver.identity" entrywise (default 11; 0 = never) solve the standard
 * problem and correct each kept pair by its metric Rayleigh quotient
 * instead of the Cholesky reduction (eigenvalues exact to second order in
 * the deviation, within degenerate clusters to |eigenvalue| x 10^-11);
 * "solver.identity32" (default 5) is the same threshold for the mixed
 * form's FP32-phase iterations.  eigenvectors
 * (host, ld >= npw) may be null.  ACC through half_solve_kpoint never locks
 * and reduces every Rayleigh-Ritz step by Cholesky (the reference's
 * iteration). */
int half_amd_solve_acc(half_handle handle, const double kpoint[3], int nbands, int mixed, double tol,
                       double switch_tol, int maxiter, int block, int extra, double *eigenvalues_eV,
                       HALF_COMPLEX *eigenvectors, int64_t ld, half_amd_solve_stats *stats,
                       char *error, int error_capacity);

/* Residual history of the last half_amd_solve_acc on this context: the
 * maximum residual norm (eV) over the requested bands at each iteration.
 * count receives the number of iterations; HALF_ERROR_CAPACITY if capacity
 * is smaller. */
int half_amd_solve_history(half_handle handle, int capacity, double *residuals_eV, int32_t *count,
                           char *error, int error_capacity);

/* All-band solves at nk k points (fractional, nk x 3) of one context in
 * lockstep: the H.psi of every k point is one launch per stage, the
 * Rayleigh-Ritz algebra batched over k, and, when graph != 0 and every
 * subspace problem has order <= 96 ("solver.smalleig"), each iteration
 * after the host's choice of correction columns (their H.psi,
 * orthogonalisation, normalisation, Rayleigh-Ritz, rotations, residual
 * norms) replayed from a hipGraph captured once per precision and
 * correction width; one host round trip per iteration.  Controls and
 * FP32-phase exit rules as in half_amd_solve_acc; "solver.lock" 1 and 2
 * both give soft locking here, with one correction width for all k points
 * (the largest active count rounded up to a multiple of 8); the solve stops
 * when every k point's
 * requested bands meet tol.  eigenvalues: nk x nbands, k-major.  stats: max_residual_eV is
 * the largest over k, band_applications counts all k points.  The
 * reference has no batched entry; it is the loop of half_solve_kpoint over
 * k points (src/half_library.F90:120-150) at ACC (solver 4). */
int half_amd_solve_kmesh(half_handle handle, int nk, const double *kpoints, int nbands, int mixed, double tol,
                         double switch_tol, int maxiter, int extra, int graph, double *eigenvalues_eV,
                         half_amd_solve_stats *stats, char *error, int error_capacity);

/* Full analytic force set (reference: app/half_cli.F90:921-962) for the
 * bands of nk k points (fractional kpoints nk x 3, weights kweights):
 * eigenvalues / occupations nk x nbands (k-major), vectors[j] the host
 * eigenvectors of k point j (npw_j x nbands, leading dimension ld[j]).
 * Output density = sum_k w sum_n f_n |psi|^2 + augmentation density of
 * rho_ab = sum w f Re[conj(c_a) c_b]; forces (natoms x 3, eV/A, written) =
 * Ewald + local (output density) + non-local (projector + dD/dR) + NLCC
 * (V_xc of the context's density) + Harris correction (XC response to the
 * output - input density on the partial core, central difference t = 1e-4).
 * components (may be null): 7 x natoms x 3, in the order Ewald, local,
 * non-local, its projector part, its dD/dR part, NLCC, Harris.  density (may
 * be null): the output density, CHGCAR order (x fastest), cell volume x
 * density.  charges (may be null): grid means of the smooth and the
 * augmentation parts. */
int half_amd_force_set(half_handle handle, int nk, const double *kpoints, const double *kweights, int nbands,
                       const double *eigenvalues, const double *occupations, const HALF_COMPLEX *const *vectors,
                       const int64_t *ld, double *forces, double *components, double *density, double *charges,
                       char *error, int error_capacity);

/* vaspwave.h5 export of solved bands (reference: src/half_vaspwave.F90:94-155
 * and src/half_hdf5_bridge.c:77-130, called from app/half_cli.F90:601,899):
 * the context's structure and input charge, ENCUT, the Fermi level, and per
 * k point its eigenvalues (celtot), occupations / 2 (fertot), plane-wave
 * count, k point and the nbands eigenvectors in VASP's plane-wave order as
 * single precision (wave, nbands x npw x 2).  Arguments as
 * half_amd_force_set; occupations must lie in [0, 2] (HALF_ERROR_INVALID_ARGUMENT
 * otherwise, where the reference stops).  The HDF5 1.10 C library
 * (libhdf5.so.103, or the path in HALF_HDF5_LIBRARY) is loaded on first use;
 * without it the call returns HALF_ERROR_UNAVAILABLE.  The same library
 * lets half_create_from_files read a VASP HDF5 file (vaspout.h5 or a
 * vaspwave.h5) given as the charge path and/or the potential path, as the
 * reference's readers do (src/half_chgcar.F90:25, src/half_potcar.F90:34). */
int half_amd_write_vaspwave(half_handle handle, const char *path, int nk, const double *kpoints, int nbands,
                            const double *eigenvalues, const double *occupations, double fermi,
                            const HALF_COMPLEX *const *vectors, const int64_t *ld, char *error,
                            int error_capacity);

/* Select the solver used by half_solve_kpoint for this context (1..4). */
int half_amd_set_solver(half_handle handle, int solver, char *error, int error_capacity);

/* Harris energy terms of the context's potential (reference:
 * src/half_cuda_potential.cuf:338-348): e_H, e_xc, e_xc,pot in eV. */
int half_amd_energy_terms(half_handle handle, double *e_hartree, double *e_xc,
                          double *e_xc_potential, char *error, int error_capacity);
/* Harris energy of solved bands, as the reference CLI assembles it
 * (app/half_cli.F90:896-920): occupations of the nk x nbands eigenvalues
 * (k-major; k weights summing to one) for the context's valence electron
 * count sum ZVAL x ions (src/half_energy.F90:41-120: sigma = 0 fills states
 * in ascending order with degenerate groups sharing electrons; sigma > 0
 * Gaussian (ismear 0) or Fermi-Dirac (ismear -1) smearing), the Ewald energy
 * of the ions (src/half_energy.F90:122-162), local_g0 = sum psp_local(0) x
 * count x nelect / volume and sum EATOM x count.  harris_energy = band -
 * e_H - e_xc,pot + e_xc + ewald + local_g0 + atomic_reference; free_energy
 * adds the entropy term.  The PAW atomic double-counting term the CLI adds
 * for use_uspp contexts is not included.  occupations (may be null): nk x
 * nbands, k-major.  Invalid weights, sigma, ismear or too few bands:
 * HALF_ERROR_INVALID_ARGUMENT with the reference's message, as is an
 * occupied highest band under smearing (the CLI stops there). */
typedef struct half_amd_harris_terms {
  double band_energy;
  double fermi;
  double entropy;
  double e_hartree;
  double e_xc;
  double e_xc_potential;
  double ewald;
  double local_g0;
  double atomic_reference;
  double harris_energy;
  double free_energy;
  double nelect;
} half_amd_harris_terms;

int half_amd_harris_energy(half_handle handle, int nk, const double *kweights, int nbands,
                           const double *eigenvalues_eV, double sigma, int ismear, double *occupations,
                           half_amd_harris_terms *terms, char *error, int error_capacity);

/* Veff on the C-order box (n1 slowest), eV; capacity in doubles. */
int half_amd_get_veff(half_handle handle, int64_t capacity, double *veff, char *error,
                      int error_capacity);

/* Hellmann-Feynman non-local projector force (reference:
 * src/half_forces.F90:9-73) through the adjoint projector contraction on
 * the device: F_I = -2 w sum_n f_n Re[(D_I - eps_n Q)c_n . conj(dc_n)],
 * c = B^H psi, dc_alpha = B^H (i q_alpha psi).  vectors: host, npw x nbands,
 * leading dimension ld; forces: natoms x 3 (atom-major, species order),
 * accumulated into (not overwritten). */
int half_amd_nonlocal_forces(half_handle handle, const double kpoint[3], int nbands,
                             const double *eigenvalues_eV, const double *occupations,
                             double kweight, const HALF_COMPLEX *vectors, int64_t ld,
                             double *forces, char *error, int error_capacity);

/* Frozen-potential displacement: move the atoms seen by the projectors
 * (and the force contraction) to `positions_fractional` (nions x 3, species
 * order, fractional) while Veff, the basis and the energy terms keep the
 * geometry the context was created with.  Operators built afterwards use
 * the new positions; operator handles already created keep theirs.  The
 * band energy sum_n f_n eps_n then varies with the positions exactly as the
 * non-local force of half_amd_nonlocal_forces predicts, which is what a
 * finite-difference check of that force needs.  No counterpart in the
 * reference ABI. */
int half_amd_set_projector_positions(half_handle handle, int nions, const double *positions_fractional,
                                     char *error, int error_capacity);

/* Launch-boundary timeline for measurement.  half_amd_trace(1) starts
 * recording an event before every operator/FFT kernel launch (clearing
 * earlier marks), half_amd_trace(0) stops.  half_amd_trace_read returns, per
 * kernel name (64-byte slots), the summed duration between a launch and the
 * next mark and the launch count; nkinds receives the number of names. */
int half_amd_trace(int enable, char *error, int error_capacity);
int half_amd_trace_read(int max_kernels, char *names, double *total_ms, int64_t *launches,
                        int32_t *nkinds, char *error, int error_capacity);

/* Kernel tunables, e.g. "local.fp64.tile.x" (FFT points per workgroup tile
 * of the x/y/z local-stage passes), "local.fp64.xgroup" (bands sharing a
 * Veff tile in the transform-based x pass), "local.fp64.xconv" (1: x pass
 * as a per-column convolution, 0: transform-based), "local.fp64.fused" and
 * "local.fp64.ygroup" (y, x, y passes in one kernel per slice of ygroup
 * bands), "local.fp64.vxlds" (1: that kernel stages the slice's Vx rows in
 * LDS); "fp32" keys for the FP32 operator; "operator.overlap" (projector
 * stage on a second stream) and "operator.projector" (1: projector stage on
 * the skinny 16x16-tile kernels, 0: on the 64-wide contractions with a
 * separate D/Q mix); "contract.fp64.tall.order" / "contract.fp32.tall.order"
 * (workgroup order of those contractions' expansion B (D c), B (Q c):
 * 0 row tiles first, 1 column tiles first, the default, 2 as 1 with each
 * XCD taking a contiguous range of row tiles); "solver.hetrd" (1: Hermitian eigensolves with vectors
 * of order <= 256 reduce to tridiagonal form in one workgroup before
 * rocSOLVER's zstedc / zunmtr, 0: rocSOLVER's zheevd throughout);
 * "solver.smalleig" (1, the default: the k-point mesh solve's subspace
 * problems of order <= 96 on one-workgroup-per-k kernels with whole
 * iterations captured in hipGraphs; 0: rocSOLVER's batched routines with
 * eager launches; 2: also single-k Rayleigh-Ritz problems and dense
 * generalized solves of order <= 96; 3: as 1 with the correction-block
 * metric always in the eigen-decomposition form).
 * Process-wide. */
int half_amd_set_tuning(const char *key, int value, char *error, int error_capacity);
int half_amd_get_tuning(const char *key, int *value, char *error, int error_capacity);

/* Name and compute-unit count of the current HIP device. */
int half_amd_device_info(char *name, int name_capacity, int32_t *compute_units, char *error,
                         int error_capacity);

#ifdef __cplusplus
}
#endif
#endif
