// Analytic non-local projector forces through the adjoint contraction.
#pragma once

#include "paw.hpp"

namespace half {

// forces: natoms x 3, accumulated.  psi: device, npw x nb, leading dimension ldpsi.
void nonlocal_forces(const KOperator<double> &op, int nb, const double2 *d_psi, int64_t ldpsi,
                     const std::vector<double> &eps, const std::vector<double> &occ, double kweight,
                     std::vector<double> &forces, hipStream_t st);

}  // namespace half
