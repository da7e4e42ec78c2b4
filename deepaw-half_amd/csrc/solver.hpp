// Eigensolvers around the PAW operator: dense generalized solves on the
// device (rocSOLVER) and the all-band block Rayleigh-Ritz minimiser ("ACC")
// with FP64 or FP32-mixed H.psi.
#pragma once

#include <rocblas/rocblas.h>

#include "paw.hpp"

namespace half {

enum SolverKind { kSolverEvd = 1, kSolverEvj = 2, kSolverEvx = 3, kSolverAcc = 4 };

struct Blas {
  rocblas_handle h = nullptr;
  hipStream_t stream = nullptr;
  DeviceArray<double2> eig_z, eig_tau;  // workspace of the small eigensolver
  void init(hipStream_t s);
  void release();
  ~Blas();
};

struct SolveResult {
  std::vector<double> eigenvalues;  // nbands
  double overlap_min = -1, overlap_max = -1;
  int iterations = 0;
  double residual = 0;
  int fp64_iterations = 0;  // iterations done with the FP64 operator
  double seconds_operator = 0, seconds_solver = 0;
  int64_t applies = 0;      // band applications performed
  int metric_fallbacks = 0; // iterations normalised by the metric's eigen-decomposition
  double rr_order_mean = 0;  // mean order of the iterations' Rayleigh-Ritz problems (2m without locking)
  std::vector<double> history;  // max residual over the requested bands, per iteration
};

struct AccControls {
  double tol = 1e-4;      // max residual norm, eV (reference default)
  int maxiter = 40;
  int block = 0;          // bands per FFT launch (0: all within 2 GiB; the reference used 16)
  bool mixed = false;     // FP32 H.psi iterations before FP64 correction sweeps
  double switch_tol = 2e-3;  // mixed form: residual at which FP64 sweeps start
  int extra = 0;          // buffer bands carried internally, not reported
};

// Dense solve of the generalized problem on the device.  Eigenvectors (if
// requested) are left on the device in `vectors` (n x nbands, ld = n).
void solve_dense(Blas &blas, KOperator<double> &op, int solver, int nbands, bool want_vectors,
                 DeviceArray<double2> &vectors, SolveResult &res);

// All-band solve (reference: src/half_cuda_iterative_solver.cuf:79-207).
// op32 is required when ctl.mixed.  Eigenvectors (n x nbands, ld = n) are
// left in `vectors` when requested.
void solve_acc(Blas &blas, KOperator<double> &op64, KOperator<float> *op32, int nbands,
               const AccControls &ctl, bool want_vectors, DeviceArray<double2> &vectors, SolveResult &res);

// Key "solver.hetrd" (0 / 1): Hermitian eigensolves with vectors of order
// up to kTridiagMaxOrder reduce to tridiagonal form in one workgroup
// (tridiag.hip) before rocSOLVER's zstedc and zunmtr (default), or run
// rocSOLVER's zheevd throughout.
bool set_solver_tuning(const std::string &key, int value);
bool get_solver_tuning(const std::string &key, int &value);
// Whether the all-band solves drop the correction columns of converged bands
// (tuning key "solver.lock").
bool solver_locking();

// Heap-sort order of plane waves by kinetic energy, ties as in the reference
// (reference: src/half_cuda_iterative_solver.cuf:277-310).  Zero-based.
std::vector<int> kinetic_order(const std::vector<double> &kinetic);

}  // namespace half
