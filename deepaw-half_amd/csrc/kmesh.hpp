// k-point mesh batching: the operator of every k point of a mesh in one
// launch per stage, and the all-band solves of all k points in lockstep
// (see kmesh.hip).
#pragma once

#include <array>
#include <vector>

#include "solver.hpp"

namespace half {

// The operator at nk k points sharing the FFT box.  States of k point j are
// stored in a uniform layout: nmax rows (the largest plane-wave count of the
// mesh), rows past that k point's own count kept zero.  Projector matrices
// are padded the same way, so the projector contractions of all k points are
// one batched launch each.
template <typename R>
struct KMeshOperator {
  const System *sys = nullptr;
  int nk = 0, nmax = 0, np = 0;
  std::vector<int> npw;                     // plane waves per k point
  std::vector<std::vector<double>> kinetic;  // host copies (initial subspace order)
  std::vector<LocalMap> maps;
  DeviceArray<LocalTables> dtabs;
  LocalBatch lb;
  DeviceArray<cplx_t<R>> B;       // [k][np][nmax]
  DeviceArray<cplx_t<R>> BD, BQ;  // B D and B Q, [k][np][nmax] (for the skinny projector kernels)
  CoeffWork<R> cwork;
  DeviceArray<R> kin;        // [k][nmax]
  DeviceArray<cplx_t<R>> boxes, coeff, dcoeff, qcoeff, work;
  DeviceArray<cplx_t<R>> nl_h, nl_s;  // projector-stage results of the overlapped apply, [k][band][nmax]
  AuxStream aux;

  void build(const System &s, const std::vector<std::array<double, 3>> &kpoints, hipStream_t st);
  // H psi and S psi for nb states of every k point: state b of k point j is
  // psi + j sk_in + b ld (nmax long, zero past npw[j]); outputs likewise with
  // (ldo, sk_out).  Output rows past npw[j] keep their (zero) contents.
  void apply(int nb, const cplx_t<R> *psi, int64_t ld, int64_t sk_in, cplx_t<R> *hpsi, cplx_t<R> *spsi,
             int64_t ldo, int64_t sk_out, hipStream_t st);
};

struct KMeshControls {
  bool graph = true;  // replay the correction-block H.psi from a captured hipGraph
};

// All-band solves (the ACC algorithm of solve_acc) at every k point of the
// mesh in lockstep: per-k Grams, Cholesky factors, eigensolves and rotations
// are strided-batched rocBLAS / rocSOLVER calls or batched contraction
// launches.  Iterates until every k point meets ctl.tol (or ctl.maxiter).
// eigenvalues: nk x nbands, k-major.  res.residual is the largest over k.
void solve_acc_kmesh(Blas &blas, KMeshOperator<double> &op64, KMeshOperator<float> *op32, int nbands,
                     const AccControls &ctl, const KMeshControls &kc, std::vector<double> &eigenvalues,
                     SolveResult &res);

}  // namespace half
