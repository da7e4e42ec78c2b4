// Potential-dependent correction of the per-atom non-local strengths D
// ("QDEP", MIMIC_US) on the device (see qdep.hip).
#pragma once

#include "paw.hpp"

namespace half {

// Replaces the D block of every atom by DION + the QDEP correction of the
// current Veff at the current atom positions (sys.positions_cart), and
// fills dD/dR (sys.ddblk / sys.d_ddblk, [alpha][a][b] per atom).  Requires
// the datasets' PAW partial waves; all-zero partial waves give D = DION.
void build_qdep(System &sys, hipStream_t s);

// Augmentation part of the output density on the C-order box (cell volume x
// density, written to aug, device): inside each atom's PAW sphere,
// Omega sum_k q_k g_l(r) Y_k(r_hat) with q_k = sum_ab rho_ab M_ab,k (M: Gaunt
// coefficient x (AE - PS) partial-wave moment, g_l the two-Bessel
// compensation shape).  occupancy: per atom nlm x nlm (row-major a, b) at
// the atom's D-block offset.  Returns the grid mean (the augmentation
// charge) (reference: src/half_uspp.F90:47-111).
double augmentation_density(const System &sys, const std::vector<double> &occupancy, double *aug, hipStream_t s);

}  // namespace half
