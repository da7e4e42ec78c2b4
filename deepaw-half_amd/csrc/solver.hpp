// Eigensolvers around the PAW operator: dense generalized solves on the
// device (rocSOLVER) and the all-band block Rayleigh-Ritz minimiser ("ACC")
// with FP64 or FP32-mixed H.psi.
#pragma once

#include <cmath>
#include <vector>

#include <rocblas/rocblas.h>

#include "paw.hpp"

namespace half {

enum SolverKind { kSolverEvd = 1, kSolverEvj = 2, kSolverEvx = 3, kSolverAcc = 4 };

struct Blas {
  rocblas_handle h = nullptr;
  hipStream_t stream = nullptr;
  DeviceArray<double2> eig_z, eig_tau, eig_coop;  // workspace of the small eigensolver
  DeviceArray<int> eig_info;                      // statuses of the one-workgroup eigensolver
  DeviceArray<double> eig_dev;                    // per-workgroup shares of a metric's deviation from I
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

// FP32 H.psi has a noise floor in the residual (about 1e-4 eV on the
// 216-atom cell); a switch_tol below it would never be reached.  A mixed
// solve therefore estimates the floor at the start (the FP32 operator's
// error on the initial subspace, fp32_floor_estimate) and ends the FP32
// phase at max(switch_tol, tol, floor), or when the residual stagnates
// within 10x of that target.  The max residual is far from monotone early
// in a solve (it can rise a hundredfold when higher bands reorder), so the
// stagnation rule is only armed near the target.  The measured floor can
// still sit above the estimate (the residual also carries the FP32 error
// accumulated in HX and SX through the rotations: on the 216-atom cell the
// estimate stays below 1e-5 eV while the residual stalls at 2e-4), so the
// FP32 phase also ends after half of maxiter, leaving the FP64 sweeps the
// rest.  update() takes the max
// residual of each armed FP32 iteration and returns true once `window`
// iterations in a row (tuning key "solver.stall"; 0 turns the rule off)
// have not brought it below 0.9 times the best value so far.
constexpr double kStallArm = 10.0;
int solver_stall_window();
// Floor estimate from the initial subspace: max_j |dH x_j| + max_j |eps_j|
// max_j |dS x_j| from squared column norms of the FP32 - FP64 differences.
double fp32_floor_estimate(const std::vector<double> &dh2, const std::vector<double> &ds2,
                           const std::vector<double> &eps);
struct StallWatch {
  int window = solver_stall_window();
  double best = HUGE_VAL;
  int since = 0;
  bool update(double r) {
    if (r < 0.9 * best) {
      best = r;
      since = 0;
    } else {
      ++since;
    }
    return window > 0 && since >= window;
  }
};

struct AccControls {
  double tol = 1e-4;      // max residual norm, eV (reference default)
  int maxiter = 40;
  int block = 0;          // bands per FFT launch (0: the operator's cache rule; the reference used 16)
  bool mixed = false;     // FP32 H.psi iterations before FP64 correction sweeps
  double switch_tol = 2e-3;  // mixed form: residual at which FP64 sweeps start
  int extra = 0;          // buffer bands carried internally, not reported
  int lock = -1;          // 0 off (the reference's 2m-wide subspace), 1 soft
                          // locking, 2 soft locking plus freezing of
                          // converged bands, -1 the process-wide key "solver.lock"
  int precond = -1;       // 0 the reference's preconditioner, 1 Teter-Payne-Allan,
                          // -1 the process-wide key "solver.precond"
  int identity = -1;      // 0: every Rayleigh-Ritz step through the Cholesky
                          // reduction (the reference's zhegv); -1 the
                          // process-wide keys "solver.identity" / "solver.identity32"
  int lock_mode() const;
  int precond_mode() const;
  bool locking() const;
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
// key solver.fp32algebra: FP32 products in the dense algebra of a mixed
// solve's FP32 phase
bool solver_fp32_algebra();
// Key "solver.smalleig" (0 rocSOLVER for every subspace problem, 1 the k-point
// mesh on the one-workgroup kernels, 2 also single-k problems of order <= 96,
// 3 as 1 with the mesh's metric always in the eigen-decomposition form).
int solver_smalleig();
bool get_solver_tuning(const std::string &key, int &value);
// Whether the all-band solves drop the correction columns of converged bands
// (tuning key "solver.lock").
bool solver_locking();

// Heap-sort order of plane waves by kinetic energy, ties as in the reference
// (reference: src/half_cuda_iterative_solver.cuf:277-310).  Zero-based.
std::vector<int> kinetic_order(const std::vector<double> &kinetic);

}  // namespace half
